/*
 * oracle.h -- CPU restatement of the reference's sequential "ref" target.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libphysis_rt_hip, the
 * generated code, physisc-hip) links, loads or calls this code.  It is used
 * by tests/ as the checker, by __graft_entry__.smoke() as the checker, and
 * by bench.py's cpu_baseline leg as the timed CPU path.
 *
 * Contents and what each follows (reference paths, file:line):
 *   ps_ref_grid_*        runtime/libphysis_rt_ref.cc:59-131 (grid new/free/
 *                        copy/swap/mirror/set; zero-filled, p1 == p0)
 *   ps_ref_reduce_*      runtime/libphysis_rt_ref.cc:25-36 + runtime/reduce.h:
 *                        34-52 (left fold in T starting from d[0])
 *   ps_ref_diffusion7_*  the ref lowering (translator/reference_translator.cc:
 *                        479-548: z outer, x inner, inclusive domain bounds;
 *                        600-649: iter x {map0, map1}) of the kernel in
 *                        examples/diffusion-benchmark/diffusion3d_physis.c:29-58
 *                        (clamped) and of the same update read through
 *                        PSGridGetPeriodic (test_7-pt-periodic.c:12-24 form)
 *   ps_ref_himeno_*      ref lowering of examples/himeno/himenobmtxpa_physis.c:
 *                        331-361 over the interior domain (:372-374), plus
 *                        the FP64 sum of ss^2 per sweep
 *   ps_uniform_field     SplitMix64 counter-based field, uniform [0,1)
 *
 * Pinning (checks that fail on a dropped term, wrong sign/index or swapped
 * operand; see tests/test_oracle.py):
 *   - diffusion clamped  : bit-exact vs the reference's own
 *     test_7-pt-neumann-cond.manual.ref.c output (golden fixture); mass
 *     conservation; closed-form cosine eigenmode decay (FP64).
 *   - diffusion periodic : bit-exact vs test_7-pt-periodic.manual.ref.c
 *     (unit coefficients, golden fixture); Fourier-mode decay (FP64).
 *   - reductions         : test_11..test_14 closed forms.
 *   - himeno             : reduction to 7-point Jacobi with a closed-form
 *     eigen-decay when a=c=1, b=0, a3=1/6; one sweep vs the reference's
 *     original himenobmtxpa_original.c jacobi compiled from its source into
 *     oracle/_ref (uniform-coefficient input).
 */
#ifndef PHYSIS_ORACLE_H_
#define PHYSIS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PS_REF_MAX = 0, PS_REF_MIN = 1, PS_REF_SUM = 2, PS_REF_PROD = 3 };

typedef struct {
  int elm_size;
  int num_dims;
  int64_t num_elms;
  int dim[3];
  void *p0, *p1;
} ps_ref_grid;

ps_ref_grid *ps_ref_grid_new(int elm_size, int num_dims, const int *dim);
void ps_ref_grid_free(ps_ref_grid *g);
void ps_ref_grid_copyin(ps_ref_grid *g, const void *src);
void ps_ref_grid_copyout(const ps_ref_grid *g, void *dst);
void ps_ref_grid_swap(ps_ref_grid *g);
void ps_ref_grid_mirror(ps_ref_grid *g);
void ps_ref_grid_set(ps_ref_grid *g, const void *value, const int *idx);

/* Returns 0 on success, 1 for an invalid operator. */
int ps_ref_reduce_float(float *out, int op, const ps_ref_grid *g);
int ps_ref_reduce_double(double *out, int op, const ps_ref_grid *g);
int ps_ref_reduce_float_array(float *out, int op, const float *d, int64_t n);
int ps_ref_reduce_double_array(double *out, int op, const double *d, int64_t n);

/* Counter-based uniform [0,1) field: value i = (splitmix64(seed + (i+1)*G) >> 40) * 2^-24. */
void ps_uniform_field_f32(float *out, int64_t n, uint64_t seed);
void ps_uniform_field_f64(double *out, int64_t n, uint64_t seed);
void ps_uniform_range_f32(float *out, int64_t n, uint64_t seed, float lo, float hi);

/* 7-point diffusion, coef = {ce, cw, cn, cs, ct, cb, cc}; dom = {minx, maxx,
 * miny, maxy, minz, maxz} (half open).  `steps` map applications alternate
 * g1->g2, g2->g1 (PSStencilRun(map(g1,g2), map(g2,g1), steps/2), plus one
 * trailing g1->g2 map when steps is odd).  periodic != 0 selects the
 * PSGridGetPeriodic form. */
void ps_ref_diffusion7_f32(float *g1, float *g2, const int *dim, const int *dom,
                           const float *coef, int periodic, int steps);
void ps_ref_diffusion7_f64(double *g1, double *g2, const int *dim, const int *dom,
                           const double *coef, int periodic, int steps);
/* Multithreaded FP64 variant (same result: each point is independent). */
void ps_ref_diffusion7_f64_mt(double *g1, double *g2, const int *dim, const int *dom,
                              const double *coef, int periodic, int steps, int threads);

/* Himeno 19-point sweep.  coefs = 12 arrays {a0,a1,a2,a3,b0,b1,b2,c0,c1,c2,
 * bnd,wrk1}, each dim[0]*dim[1]*dim[2].  Runs `sweeps` sweeps alternating
 * p0->p1, p1->p0 over the interior; gosa[s] = FP64 sum of ss*ss of sweep s
 * (may be NULL). */
void ps_ref_himeno_f32(float *p0, float *p1, const float *const *coefs, const int *dim,
                       float omega, int sweeps, double *gosa);
void ps_ref_himeno_f64(double *p0, double *p1, const double *const *coefs, const int *dim,
                       double omega, int sweeps, double *gosa);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_ORACLE_H_ */
