/*
 * stopwatch.h -- wall-clock stopwatch used by generated tracing code.
 * Same names and millisecond unit as the reference
 * (include/physis/stopwatch.h:17-47); implemented on the POSIX monotonic
 * clock where <time.h> declares it, else on C11 timespec_get (strict ISO C
 * modes).
 */
#ifndef PHYSIS_STOPWATCH_H_
#define PHYSIS_STOPWATCH_H_

#include <time.h>

typedef struct {
  struct timespec ts;
} __PSStopwatch;

static inline void __PSStopwatchQuery(__PSStopwatch *w) {
#if defined(CLOCK_MONOTONIC)
  clock_gettime(CLOCK_MONOTONIC, &w->ts);
#else
  timespec_get(&w->ts, TIME_UTC);
#endif
}

/* Milliseconds from begin to end. */
static inline float __PSStopwatchDiff(const __PSStopwatch *begin,
                                      const __PSStopwatch *end) {
  double s = (double)(end->ts.tv_sec - begin->ts.tv_sec);
  double ns = (double)(end->ts.tv_nsec - begin->ts.tv_nsec);
  return (float)(s * 1.0e3 + ns * 1.0e-6);
}

static inline void __PSStopwatchStart(__PSStopwatch *w) {
  __PSStopwatchQuery(w);
}

static inline float __PSStopwatchStop(__PSStopwatch *w) {
  __PSStopwatch now;
  __PSStopwatchQuery(&now);
  return __PSStopwatchDiff(w, &now);
}

#endif /* PHYSIS_STOPWATCH_H_ */
