/*
 * config.h -- build-time switches of the MI355X target.  The reference
 * generates this file from CMake (include/physis/config.h.cmake); this
 * target has exactly one configuration, so it is a plain header.
 */
#ifndef PHYSIS_CONFIG_H_
#define PHYSIS_CONFIG_H_

#define PHYSIS_HIP_ENABLED 1

#endif /* PHYSIS_CONFIG_H_ */
