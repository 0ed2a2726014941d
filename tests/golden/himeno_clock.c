/* A stand-in wall clock for the reference's Himeno program
 * (examples/himeno/himenobmtxpa_physis.c), linked in place of gettimeofday
 * when tests/golden/build_corpus.py builds it (-Dgettimeofday=ps_himeno_clock).
 * The program sizes its measured loop to one minute from a timed rehearsal of
 * four sweeps (himenobmtxpa_physis.c:168-178: nn = 60 / (cpu / 3), made even);
 * this clock advances ten seconds per reading, so the rehearsal "takes" 10 s
 * and the measured loop runs 18 sweeps.  Its timings are meaningless; the
 * program is run for its results. */
#include <sys/time.h>

int ps_himeno_clock(struct timeval *tv, void *tz) {
  static long readings = 0;
  (void)tz;
  tv->tv_sec = 1000 + 10 * readings++;
  tv->tv_usec = 0;
  return 0;
}
