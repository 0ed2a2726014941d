/*
 * oracle.c -- CPU restatement of the reference's sequential ref target.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for scope, sources and pinning).
 *
 * Built with -O2 -ffp-contract=off: every expression is evaluated as written
 * in the DSL source, in C's left-to-right order, with one rounding per
 * operation, which is what the ref target compiled by gcc on x86-64 does.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- runtime: runtime/libphysis_rt_ref.cc:59-131 -------------------------- */

ps_ref_grid *ps_ref_grid_new(int elm_size, int num_dims, const int *dim) {
  ps_ref_grid *g = (ps_ref_grid *)malloc(sizeof(ps_ref_grid));
  if (!g) return NULL;
  g->elm_size = elm_size;
  g->num_dims = num_dims;
  g->num_elms = 1;
  for (int i = 0; i < 3; ++i) g->dim[i] = i < num_dims ? dim[i] : 1;
  for (int i = 0; i < num_dims; ++i) g->num_elms *= dim[i];
  g->p0 = calloc((size_t)g->num_elms, (size_t)elm_size);
  if (!g->p0) {
    free(g);
    return NULL; /* INVALID_GRID on allocation failure */
  }
  g->p1 = g->p0; /* no implicit double buffering */
  return g;
}

void ps_ref_grid_free(ps_ref_grid *g) {
  if (!g) return;
  if (g->p0) free(g->p0);
  if (g->p1 && g->p1 != g->p0) free(g->p1);
  free(g);
}

void ps_ref_grid_copyin(ps_ref_grid *g, const void *src) {
  memcpy(g->p0, src, (size_t)g->elm_size * (size_t)g->num_elms);
}

void ps_ref_grid_copyout(const ps_ref_grid *g, void *dst) {
  memcpy(dst, g->p0, (size_t)g->elm_size * (size_t)g->num_elms);
}

void ps_ref_grid_swap(ps_ref_grid *g) {
  void *t = g->p1;
  g->p1 = g->p0;
  g->p0 = t;
}

void ps_ref_grid_mirror(ps_ref_grid *g) {
  if (g->p0 != g->p1) memcpy(g->p1, g->p0, (size_t)g->elm_size * (size_t)g->num_elms);
}

void ps_ref_grid_set(ps_ref_grid *g, const void *value, const int *idx) {
  int64_t off = 0, stride = 1;
  for (int i = 0; i < g->num_dims; ++i) {
    off += (int64_t)idx[i] * stride;
    stride *= g->dim[i];
  }
  memcpy((char *)g->p0 + off * g->elm_size, value, (size_t)g->elm_size);
}

/* ---- reductions: libphysis_rt_ref.cc:25-36, reduce.h:34-52 ---------------- */

#define DEFINE_REDUCE(NAME, T)                                            \
  int NAME(T *out, int op, const T *d, int64_t n) {                       \
    if (op < PS_REF_MAX || op > PS_REF_PROD) return 1;                    \
    T v = d[0];                                                           \
    for (int64_t i = 1; i < n; ++i) {                                     \
      T x = d[i];                                                         \
      switch (op) {                                                       \
        case PS_REF_MAX: v = (v > x) ? v : x; break;                      \
        case PS_REF_MIN: v = (v < x) ? v : x; break;                      \
        case PS_REF_SUM: v = v + x; break;                                \
        default: v = v * x; break;                                        \
      }                                                                   \
    }                                                                     \
    *out = v;                                                             \
    return 0;                                                             \
  }
DEFINE_REDUCE(ps_ref_reduce_float_array, float)
DEFINE_REDUCE(ps_ref_reduce_double_array, double)

int ps_ref_reduce_float(float *out, int op, const ps_ref_grid *g) {
  return ps_ref_reduce_float_array(out, op, (const float *)g->p0, g->num_elms);
}
int ps_ref_reduce_double(double *out, int op, const ps_ref_grid *g) {
  return ps_ref_reduce_double_array(out, op, (const double *)g->p0, g->num_elms);
}

/* ---- seeded fields --------------------------------------------------------- */

static inline uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline float unit_f32(uint64_t seed, int64_t i) {
  uint64_t z = splitmix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
  return (float)(z >> 40) * (1.0f / 16777216.0f);
}

void ps_uniform_field_f32(float *out, int64_t n, uint64_t seed) {
  for (int64_t i = 0; i < n; ++i) out[i] = unit_f32(seed, i);
}

void ps_uniform_field_f64(double *out, int64_t n, uint64_t seed) {
  for (int64_t i = 0; i < n; ++i) out[i] = (double)unit_f32(seed, i);
}

void ps_uniform_range_f32(float *out, int64_t n, uint64_t seed, float lo, float hi) {
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * unit_f32(seed, i);
}

/* ---- 7-point diffusion ----------------------------------------------------- */

/* One map application over the domain, ref loop order (z, y, x). */
#define DEFINE_DIFF_MAP(NAME, T)                                                   \
  static void NAME(const T *g1, T *g2, const int *dim, const int *dom,             \
                   const T *k, int periodic, int z_lo, int z_hi) {                  \
    const int nx = dim[0], ny = dim[1], nz = dim[2];                                \
    const T ce = k[0], cw = k[1], cn = k[2], cs = k[3], ct = k[4], cb = k[5],       \
            cc = k[6];                                                              \
    const int64_t pl = (int64_t)nx * ny;                                            \
    for (int z = z_lo; z <= z_hi - 1; ++z)                                          \
      for (int y = dom[2]; y <= dom[3] - 1; ++y)                                    \
        for (int x = dom[0]; x <= dom[1] - 1; ++x) {                                \
          T c, w, e, n, s, b, t;                                                    \
          if (periodic) {                                                           \
            c = g1[x + y * nx + z * pl];                                            \
            w = g1[(x - 1 + nx) % nx + y * nx + z * pl];                            \
            e = g1[(x + 1 + nx) % nx + y * nx + z * pl];                            \
            n = g1[x + (y - 1 + ny) % ny * nx + z * pl];                            \
            s = g1[x + (y + 1 + ny) % ny * nx + z * pl];                            \
            b = g1[x + y * nx + (int64_t)((z - 1 + nz) % nz) * pl];                 \
            t = g1[x + y * nx + (int64_t)((z + 1 + nz) % nz) * pl];                 \
          } else {                                                                  \
            c = g1[x + y * nx + z * pl];                                            \
            if (x == 0) w = g1[x + y * nx + z * pl];                                \
            else w = g1[x - 1 + y * nx + z * pl];                                   \
            if (x == nx - 1) e = g1[x + y * nx + z * pl];                           \
            else e = g1[x + 1 + y * nx + z * pl];                                   \
            if (y == 0) n = g1[x + y * nx + z * pl];                                \
            else n = g1[x + (y - 1) * nx + z * pl];                                 \
            if (y == ny - 1) s = g1[x + y * nx + z * pl];                           \
            else s = g1[x + (y + 1) * nx + z * pl];                                 \
            if (z == 0) b = g1[x + y * nx + z * pl];                                \
            else b = g1[x + y * nx + (z - 1) * pl];                                 \
            if (z == nz - 1) t = g1[x + y * nx + z * pl];                           \
            else t = g1[x + y * nx + (z + 1) * pl];                                 \
          }                                                                         \
          g2[x + y * nx + z * pl] = cc * c + cw * w + ce * e + cs * s + cn * n +    \
                                    cb * b + ct * t;                                \
        }                                                                           \
  }
DEFINE_DIFF_MAP(diff_map_f32, float)
DEFINE_DIFF_MAP(diff_map_f64, double)

void ps_ref_diffusion7_f32(float *g1, float *g2, const int *dim, const int *dom,
                           const float *coef, int periodic, int steps) {
  for (int i = 0; i < steps; ++i) {
    if (i & 1) diff_map_f32(g2, g1, dim, dom, coef, periodic, dom[4], dom[5]);
    else diff_map_f32(g1, g2, dim, dom, coef, periodic, dom[4], dom[5]);
  }
}

void ps_ref_diffusion7_f64(double *g1, double *g2, const int *dim, const int *dom,
                           const double *coef, int periodic, int steps) {
  for (int i = 0; i < steps; ++i) {
    if (i & 1) diff_map_f64(g2, g1, dim, dom, coef, periodic, dom[4], dom[5]);
    else diff_map_f64(g1, g2, dim, dom, coef, periodic, dom[4], dom[5]);
  }
}

typedef struct {
  const double *in;
  double *out;
  const int *dim, *dom;
  const double *coef;
  int periodic, z_lo, z_hi;
} diff_job;

static void *diff_worker(void *p) {
  diff_job *j = (diff_job *)p;
  diff_map_f64(j->in, j->out, j->dim, j->dom, j->coef, j->periodic, j->z_lo, j->z_hi);
  return NULL;
}

void ps_ref_diffusion7_f64_mt(double *g1, double *g2, const int *dim, const int *dom,
                              const double *coef, int periodic, int steps, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  diff_job jobs[64];
  const int zr = dom[5] - dom[4];
  for (int i = 0; i < steps; ++i) {
    const double *in = (i & 1) ? g2 : g1;
    double *out = (i & 1) ? g1 : g2;
    for (int t = 0; t < threads; ++t) {
      jobs[t].in = in;
      jobs[t].out = out;
      jobs[t].dim = dim;
      jobs[t].dom = dom;
      jobs[t].coef = coef;
      jobs[t].periodic = periodic;
      jobs[t].z_lo = dom[4] + (int)((int64_t)zr * t / threads);
      jobs[t].z_hi = dom[4] + (int)((int64_t)zr * (t + 1) / threads);
      pthread_create(&th[t], NULL, diff_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  }
}

/* ---- Himeno 19-point: himenobmtxpa_physis.c:331-361 ------------------------ */

#define DEFINE_HIMENO(NAME, T)                                                       \
  void NAME(T *p0, T *p1, const T *const *cf, const int *dim, T omega, int sweeps,    \
            double *gosa) {                                                          \
    const int nx = dim[0], ny = dim[1], nz = dim[2];                                 \
    const int64_t pl = (int64_t)nx * ny;                                             \
    const T *a0 = cf[0], *a1 = cf[1], *a2 = cf[2], *a3 = cf[3];                      \
    const T *b0 = cf[4], *b1 = cf[5], *b2 = cf[6];                                   \
    const T *c0 = cf[7], *c1 = cf[8], *c2 = cf[9], *bnd = cf[10], *wrk1 = cf[11];    \
    for (int sw = 0; sw < sweeps; ++sw) {                                            \
      const T *p = (sw & 1) ? p1 : p0;                                               \
      T *q = (sw & 1) ? p0 : p1;                                                     \
      double acc = 0.0;                                                              \
      for (int k = 1; k <= nz - 2; ++k)                                              \
        for (int j = 1; j <= ny - 2; ++j)                                            \
          for (int i = 1; i <= nx - 2; ++i) {                                        \
            const int64_t o = i + (int64_t)j * nx + k * pl;                          \
            T s0, ss;                                                                \
            s0 = a0[o] * p[o + pl] + a1[o] * p[o + nx] + a2[o] * p[o + 1] +          \
                 b0[o] * (p[o + nx + pl] - p[o - nx + pl] - p[o + nx - pl] +         \
                          p[o - nx - pl]) +                                          \
                 b1[o] * (p[o + 1 + nx] - p[o + 1 - nx] - p[o - 1 + nx] +            \
                          p[o - 1 - nx]) +                                           \
                 b2[o] * (p[o + 1 + pl] - p[o + 1 - pl] - p[o - 1 + pl] +            \
                          p[o - 1 - pl]) +                                           \
                 c0[o] * p[o - pl] + c1[o] * p[o - nx] + c2[o] * p[o - 1] + wrk1[o]; \
            ss = (s0 * a3[o] - p[o]) * bnd[o];                                       \
            acc += (double)ss * (double)ss;                                          \
            q[o] = p[o] + omega * ss;                                                \
          }                                                                          \
      if (gosa) gosa[sw] = acc;                                                      \
    }                                                                                \
  }
DEFINE_HIMENO(ps_ref_himeno_f32, float)
DEFINE_HIMENO(ps_ref_himeno_f64, double)
