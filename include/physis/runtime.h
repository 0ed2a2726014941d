/*
 * runtime.h -- trace hooks called by generated run functions.
 * `__ps_trace` is set to stderr by `--physis-trace` (reference
 * runtime/runtime.cc:81-91); the message texts match the reference's
 * include/physis/runtime.h:22-34 so existing log scrapers keep working.
 */
#ifndef PHYSIS_RUNTIME_H_
#define PHYSIS_RUNTIME_H_

#include <stdio.h>

#include "physis/stopwatch.h"

#ifdef __cplusplus
extern "C" {
#endif

extern FILE *__ps_trace;

static inline void __PSTraceStencilPre(const char *msg) {
  if (__ps_trace) fprintf(__ps_trace, "Physis: Stencil started (%s)\n", msg);
}

static inline void __PSTraceStencilPost(float time) {
  if (__ps_trace) fprintf(__ps_trace, "Physis: Stencil finished (time: %f)\n", time);
}

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_RUNTIME_H_ */
