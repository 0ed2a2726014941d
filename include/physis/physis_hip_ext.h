/*
 * physis_hip_ext.h -- additive C ABI of libphysis_rt_hip.
 *
 * Nothing here replaces a reference entry point; these calls extend the
 * runtime for callers that need more than the void/abort ABI of
 * physis_hip.h: status codes instead of process exit, 16-bit grid storage,
 * FP64 reduction results, device-side timing, and the built-in stencil
 * programs (the code physisc-hip emits for the diffusion and Himeno
 * benchmarks, compiled into the library).
 *
 * All calls are ordered on the runtime stream (__PSGetStream()).
 */
#ifndef PHYSIS_PHYSIS_HIP_EXT_H_
#define PHYSIS_PHYSIS_HIP_EXT_H_

#include "physis/physis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the non-aborting calls. */
enum PSHipStatus {
  PS_HIP_OK = 0,
  PS_HIP_ERR_INVALID = 1,   /* bad argument (shape, op, storage, domain) */
  PS_HIP_ERR_NOMEM = 2,     /* device allocation failed or budget exceeded */
  PS_HIP_ERR_DEVICE = 3,    /* a HIP call failed */
  PS_HIP_ERR_UNSUPPORTED = 4
};

/* How a grid's elements are held in HBM.  16-bit storage applies to FP32
 * grids only: host copies and all arithmetic stay FP32. */
enum PSHipStorage {
  PS_HIP_STORAGE_NATIVE = 0,
  PS_HIP_STORAGE_BF16 = 1,
  PS_HIP_STORAGE_FP16 = 2
};

/* __PSGridNew with the grid attribute of PSGrid3DFloatNew(nx, ny, nz, attr):
 * PS_GRID_STORAGE_BF16 / PS_GRID_STORAGE_FP16 select 16-bit storage. */
extern __PSGrid *__PSGridNewAttr(int elm_size, int num_dims, PSVectorInt dim,
                                 int attr);
/* Non-aborting grid creation; *out is NULL on failure. */
extern int PSHipGridNew(__PSGrid **out, int elm_size, int num_dims,
                        const int *dim, int storage);
extern int PSHipGridFree(__PSGrid *g);
extern int PSHipGridCopyin(__PSGrid *g, const void *src);
extern int PSHipGridCopyout(__PSGrid *g, void *dst);
/* Asynchronous device-to-device copy of a whole grid (same shape/storage). */
extern int PSHipGridCopyDevice(__PSGrid *dst, const __PSGrid *src);
extern int PSHipGridStorage(const __PSGrid *g);
/* Bytes per element as stored in HBM. */
extern int PSHipGridStoredElmSize(const __PSGrid *g);
/* Raw device pointer (same as g->p0). */
extern void *PSHipGridDevicePtr(const __PSGrid *g);

/* Fills a grid on the device with the seeded uniform [lo, hi) field
 * lo + (hi - lo) * u_i, u_i = (splitmix64(seed + (i+1)*0x9E3779B97F4A7C15) >> 40) * 2^-24,
 * i the linear (x fastest) index.  Asynchronous. */
extern int PSHipGridFillUniform(__PSGrid *g, uint64_t seed, float lo, float hi);
/* Fills a float grid with `value` (profile = 0) or with the Himeno initial
 * pressure (float)(z*z) / ((nz-1)*(nz-1)) (profile = 1).  Asynchronous. */
extern int PSHipGridFillPattern(__PSGrid *g, int profile, float value);

/* Streaming device copy of src into dst (same byte size, a multiple of 16)
 * with one plain kernel: the HBM bandwidth yardstick.  Asynchronous. */
extern int PSHipStreamCopy(__PSGrid *dst, const __PSGrid *src);

/* Reduction with an FP64 result.  SUM and PROD accumulate in FP64 in a
 * fixed order (bitwise reproducible); MAX and MIN are exact. */
extern int PSHipReduceGrid(double *out, enum PSReduceOp op, __PSGrid *g);

/* PSHipReduceGrid for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It calls no hipMalloc,
 * hipFree or other allocating routine: the partials and the result live in
 * a file-scope device array of 1 + 16384 doubles (128 KiB + 8 bytes) that
 * every device the library runs on holds until the process ends, and the
 * result is copied to *out.  Pass 1 never uses more than 16384 workgroups,
 * so every grid fits; a size that would not returns PS_HIP_ERR_UNSUPPORTED.
 * Same kernels and work split as PSHipReduceGrid, so the same result bits.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (from several host threads) race. */
extern int PSHipReduceGridNoAlloc(double *out, enum PSReduceOp op, __PSGrid *g);

/* A device buffer of at least `bytes`, owned by the runtime and reused by
 * every caller on its stream (fused time steps use it as a third grid
 * buffer).  NULL if it cannot be allocated. */
extern void *__PSHipAuxBuffer(size_t bytes);

/* Kernel reductions, PSReduce(&v, op, kernel, dom, ...) (documented in the
 * reference's docs/programming.md "Reduction" but not implemented there,
 * reference_translator.cc:1357-1360), as physisc-hip lowers them
 * (physis/hip/generated.h RunReduceMap): the reduction kernel's pass writes
 * one FP64 partial per workgroup into __PSHipReduceMapPartials(n) (n <= 16384;
 * NULL outside that range), then __PSHipReduceMapFinish folds them in index
 * order with op on the runtime stream, waits, and stores the FP64 result in
 * *out (a PSHipStatus). */
extern double *__PSHipReduceMapPartials(int n);
extern int __PSHipReduceMapFinish(double *out, enum PSReduceOp op, const double *partials, int n);

/* Blocks until the runtime stream is idle; returns a PSHipStatus. */
extern int PSHipSynchronize(void);
/* Device bytes held by live grids. */
extern int64_t PSHipGridBytesInUse(void);
extern int PSHipDeviceName(char *buf, int len);
/* Free and total device memory in bytes. */
extern int PSHipMemInfo(int64_t *free_bytes, int64_t *total_bytes);

/* Event timer on the runtime stream: Begin records a start event, End a stop
 * event; Elapsed waits and returns milliseconds between them. */
extern int PSHipTimerBegin(void);
extern int PSHipTimerEnd(void);
extern int PSHipTimerElapsed(float *ms);

/* ---- built-in stencil programs -------------------------------------------
 * 7-point diffusion (reference examples/diffusion-benchmark/diffusion3d_physis.c:29-76):
 *   out(x,y,z) = cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t
 * with the neighbour taken from the same point at a clamped face
 * (boundary = PS_HIP_CLAMPED) or from the wrapped point
 * (boundary = PS_HIP_PERIODIC).  coef = {ce, cw, cn, cs, ct, cb, cc}.
 * Runs `steps` steps alternating g1->g2 and g2->g1 (PSStencilRun with two
 * maps and count steps/2; an odd count ends with one g1->g2 map).
 * variant: 0 = two fused steps per pass with the launch shape chosen from
 * the grid size (waves per row from nx, z-chunk to fill the device; one step
 * per pass where rows are not whole 16-byte vectors), otherwise a kernel id from
 * PSHipDiffusion7Variants. */
enum PSHipBoundary { PS_HIP_CLAMPED = 0, PS_HIP_PERIODIC = 1 };
extern int PSHipDiffusion7Run(__PSGrid *g1, __PSGrid *g2,
                              const __PSDomain *dom, const float *coef,
                              int boundary, int steps, int variant);
/* Number of kernel variants and a short name for each (1-based ids). */
extern int PSHipDiffusion7Variants(void);
extern const char *PSHipDiffusion7VariantName(int variant);
/* Diagnostic (host only, no reference counterpart): the packed-row tiles'
 * lane and halo geometry for rows of nx points and 64*nw-lane workgroups.
 * lane_out: 8 ints per lane (row, vector, first point, shadow flag, and the
 * slot and position offset of the x neighbours across the vector's ends);
 * halo_out: 4 ints per point of a halo row (two slot/position places).
 * Returns the rows per workgroup. */
extern int PSHipPackGeometry(int nx, int nw, int periodic, int *lane_out, int *halo_out);

/* Himeno 19-point Jacobi sweep (reference examples/himeno/himenobmtxpa_physis.c:331-361).
 * grids[] = {a0,a1,a2,a3,b0,b1,b2,c0,c1,c2,bnd,wrk1}; p0/p1 ping-pong.
 * Runs `sweeps` sweeps over the interior [1,n-1) per axis, one sweep per
 * launch; gosa[i] receives the FP64 sum of ss^2 of sweep i when gosa is not
 * NULL.  variant: 0 = the default tile (256 points, 256/128/64 wide: the
 * widest that leaves at most 10% of a row's lanes idle; z-chunk of 4 or 2
 * planes for FP32, 16, 8, 4 or 2 for 16-bit storage: the largest that still
 * gives 1024 workgroups), 1-3 and 5-7 fixed generic-skeleton tiles, 4 the
 * one-sweep shape of the fused kernel (coefficients loaded a plane ahead). */
extern int PSHipHimenoRun(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids,
                          float omega, int sweeps, double *gosa, int variant);
/* The same sweeps on the generic skeleton with bx x by threads computing
 * `rows` tile rows each (64 x 2/4/8, 128 x 2/4, 256 x 1/2 with one row;
 * 64 x 4, 128 x 2, 256 x 1 with two) and zchunk planes per workgroup (0:
 * all); for tuning.  PS_HIP_ERR_UNSUPPORTED for other tiles. */
extern int PSHipHimenoRunTile(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids,
                              float omega, int sweeps, double *gosa, int bx, int by, int rows,
                              int zchunk);
/* The same sweeps with `degree` (1..4) of them fused per HBM pass (temporal
 * blocking; no reference counterpart): each coefficient element is loaded
 * once per pass instead of once per sweep.  Tiles of 64 x by points with
 * rows_per_wave tile rows per wave: (by, rows_per_wave) = (8, 1) or (8, 2)
 * for degree <= 3, (16, 1) or (16, 2) for any degree; zchunk = output
 * planes per workgroup (0: all); prefetch = march steps between a plane's
 * load and its first use (0: the default, 1; 2..4 for some shapes,
 * PS_HIP_ERR_UNSUPPORTED otherwise).  p0,
 * p1 and gosa end exactly as after PSHipHimenoRun; degree > 1 uses a third
 * runtime-owned buffer of the grid's size; grids must be below 4 GiB. */
extern int PSHipHimenoRunTB(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids,
                            float omega, int sweeps, double *gosa, int degree,
                            int by, int rows_per_wave, int zchunk, int prefetch);
/* The pass schedule of PSHipHimenoRunTB (host only, no device work): pass i
 * is passes[4i..4i+3] = {first sweep index, sweeps in the pass, source,
 * destination} with buffers 0 = p0, 1 = p1, 2 = the third buffer.  Returns
 * the number of passes, or -1 if max_passes is too small. */
extern int PSHipHimenoTBSchedule(int sweeps, int degree, int *passes, int max_passes);

/* ---- batched many-grid diffusion (no reference counterpart) -------------
 * Grid pairs (a[i], b[i]) of any shape (nx <= 512), FP32 or BF16 storage
 * (one kind per batch), coef = 7 floats per pair (ce, cw, cn, cs, ct, cb, cc).
 * Each step is one launch over all grids; runs are captured into hipGraphs. */
extern int PSHipBatchCreate(void **out, __PSGrid *const *a, __PSGrid *const *b, int n, const float *coef,
                            int boundary);
/* aux[i] (if aux and aux[i] are not NULL): caller-owned third buffer of the
 * grid's byte size used by fused steps; otherwise the batch allocates one. */
extern int PSHipBatchCreateEx(void **out, __PSGrid *const *a, __PSGrid *const *b, void *const *aux, int n,
                              const float *coef, int boundary);
extern int PSHipBatchDestroy(void *batch);
/* steps steps of every grid (A->B, B->A, ...); use_graph: capture and replay. */
extern int PSHipBatchRun(void *batch, int steps, int use_graph);
/* The same on a caller stream (a hipStream_t), without a graph.  Runs of one
 * batch must not overlap in time (they share its grids, third buffers and
 * the device queue its packed launches take work items from): queue them on
 * one stream, or order them with events. */
extern int PSHipBatchRunOn(void *batch, int steps, void *stream);
/* Launch shape: points per lane np (row width 64*np), rows per workgroup by,
 * planes per work item zchunk, temporal degree 1 or 2.  np = 0: mixed mode,
 * every grid in its class (the fewest points per lane that cover its rows;
 * grids of whole 16-byte rows longer than 128 points: the vectorised star
 * tile) and each class on its own stream (forked from and joined to the
 * caller's).  np = -1: the mixed mode with 16 rows per workgroup for the
 * classes of 1 and 2 points per lane.  np = -2: packed mode (FP32 grids
 * with rows of 32 to 512 points, degree 2 only): the fused pairs of all
 * grids in one launch on tiles that pack rows of any length across lanes
 * (four points per lane), `by` (8 or 16) waves per workgroup; single steps
 * as in the mixed mode. */
extern int PSHipBatchConfigure(void *batch, int np, int by, int zchunk, int degree);
extern int PSHipBatchConfigName(void *batch, char *buf, int len);
/* Times every candidate shape for `steps` steps on the batch (contents
 * restored) and keeps the fastest. */
extern int PSHipBatchTune(void *batch, int steps, float *best_ms);

/* ---- stencil-solve server (no reference counterpart) ---------------------- */
typedef struct {
  int kind;         /* 0: 7-point diffusion; 1: 19-point Himeno-form Jacobi */
  int n;            /* cubic edge, 3..512 */
  int storage;      /* PS_HIP_STORAGE_NATIVE or PS_HIP_STORAGE_BF16 */
  uint64_t seed;    /* initial field / coefficients */
  double arrival;   /* seconds after the start of the trace */
  int max_steps;    /* step cap */
  double tol;       /* residual tolerance */
} PSHipRequest;

typedef struct {
  int status;       /* PSHipStatus */
  int steps;
  double residual;
  double admitted;  /* seconds after the trace start */
  double finished;
  double checksum;  /* FP64 sum of the final field */
} PSHipResult;

/* Serves the requests in real time.  mode 0: continuous batching under an
 * HBM budget: diffusion requests of one storage kind advance together in
 * batched launches, Jacobi requests on side streams (those with grids of
 * 256 MiB or more one at a time in arrival order), each batch and request
 * starting its next tick as soon as its previous one completes; mode 1: one
 * request at a time through the single-grid paths.  tick_steps (even):
 * steps between residual checks.  Results are the same in both modes. */
extern int PSHipServeTrace(const PSHipRequest *reqs, int n, int64_t budget_bytes, int tick_steps, int mode,
                           PSHipResult *results);
/* Tunes, ahead of serving, the Jacobi launch shape of every (storage, size
 * class) among the requests (a trace tunes the classes left untuned when
 * it starts); the choices are kept for the process. */
extern int PSHipServeTune(const PSHipRequest *reqs, int n, int64_t budget_bytes, int tick_steps);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_HIP_EXT_H_ */
