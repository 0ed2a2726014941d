/*
 * physis.h -- umbrella header.  DSL sources are parsed with PHYSIS_USER
 * defined; code generated by physisc-hip is compiled with PHYSIS_HIP
 * defined (the reference selects its targets the same way,
 * include/physis/physis.h:12-32).
 */
#ifndef PHYSIS_PHYSIS_H_
#define PHYSIS_PHYSIS_H_

#if defined(PHYSIS_USER)
#include "physis/physis_user.h"
#endif

#if defined(PHYSIS_HIP)
#include "physis/physis_hip.h"
#endif

#include <math.h>

#endif /* PHYSIS_PHYSIS_H_ */
