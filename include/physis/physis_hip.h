/*
 * physis_hip.h -- C ABI of libphysis_rt_hip, the MI355X (gfx950) runtime
 * that code generated by physisc-hip links against.
 *
 * The entry points and struct layouts replace the CUDA target's ABI one for
 * one, so generated host code keeps its shape:
 *
 *   reference symbol / type                      replaced by (this header)
 *   -------------------------------------------  ---------------------------
 *   __PSGrid{1,2,3}D*_dev  physis_cuda.h:18-97    same layouts below
 *   __PSGrid               physis_cuda.h:99-106   same layout below
 *   __PSGridNew            physis_cuda.h:131-132  __PSGridNew
 *   __PSGridSwap/Mirror/GetID/Set physis_cuda.h:133-136  same names
 *   __PSGridFree           physis_cuda.h:137-138  __PSGridFree
 *   __PSGridCopyin/out     physis_cuda.h:139-145  __PSGridCopyin/out
 *   __PSCheckCudaError     physis_cuda.h:150      __PSCheckDeviceError
 *   __PSGridGetOffset*[Dev] physis_cuda.h:158-231 same names
 *   __PSReduceGridFloat/Double physis_cuda.h:234-238 same names
 *   PSInit/PSFinalize/PSDomain*New physis_common.h:76-99 same names
 *
 * Runtime errors print a message to stderr and abort the process, as the
 * reference runtimes do (runtime/runtime_common_cuda.h:16-23).  Status-code
 * variants for callers that must not exit live in physis_hip_ext.h.
 */
#ifndef PHYSIS_PHYSIS_HIP_H_
#define PHYSIS_PHYSIS_HIP_H_

#include "physis/physis_common.h"

#if defined(__HIPCC__)
#define PS_HD __host__ __device__
#else
#define PS_HD
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device views: passed by value to kernels ---------------------------- */
typedef struct { int dim[1]; void *p0; } __PSGrid1D_dev;
typedef struct { int dim[1]; float *p0; } __PSGrid1DFloat_dev;
typedef struct { int dim[1]; double *p0; } __PSGrid1DDouble_dev;
typedef struct { int dim[1]; int *p0; } __PSGrid1DInt_dev;
typedef struct { int dim[1]; long *p0; } __PSGrid1DLong_dev;

typedef struct { int dim[2]; void *p0; } __PSGrid2D_dev;
typedef struct { int dim[2]; float *p0; } __PSGrid2DFloat_dev;
typedef struct { int dim[2]; double *p0; } __PSGrid2DDouble_dev;
typedef struct { int dim[2]; int *p0; } __PSGrid2DInt_dev;
typedef struct { int dim[2]; long *p0; } __PSGrid2DLong_dev;

typedef struct { int dim[3]; void *p0; } __PSGrid3D_dev;
typedef struct { int dim[3]; float *p0; } __PSGrid3DFloat_dev;
typedef struct { int dim[3]; double *p0; } __PSGrid3DDouble_dev;
typedef struct { int dim[3]; int *p0; } __PSGrid3DInt_dev;
typedef struct { int dim[3]; long *p0; } __PSGrid3DLong_dev;

typedef struct { int dim[3]; void *p0; } __PSGrid_dev;

/* ---- host handle --------------------------------------------------------- */
/* p0 is the device buffer (primitive point types only); dev points to a
 * host-side copy of the device view that generated code passes by value. */
typedef struct {
  void *p0;
  PSVectorInt dim;
  int elm_size;
  int num_dims;
  int64_t num_elms;
  __PSGrid_dev *dev;
} __PSGrid;

#ifndef PHYSIS_USER
typedef __PSGrid *PSGrid1DFloat;
typedef __PSGrid *PSGrid2DFloat;
typedef __PSGrid *PSGrid3DFloat;
typedef __PSGrid *PSGrid1DDouble;
typedef __PSGrid *PSGrid2DDouble;
typedef __PSGrid *PSGrid3DDouble;
typedef __PSGrid *PSGrid1DInt;
typedef __PSGrid *PSGrid2DInt;
typedef __PSGrid *PSGrid3DInt;
typedef __PSGrid *PSGrid1DLong;
typedef __PSGrid *PSGrid2DLong;
typedef __PSGrid *PSGrid3DLong;
#define DeclareGrid1D(name, type) typedef __PSGrid *PSGrid1D##name;
#define DeclareGrid2D(name, type) typedef __PSGrid *PSGrid2D##name;
#define DeclareGrid3D(name, type) typedef __PSGrid *PSGrid3D##name;
#define PSGridDim(p, d) ((p)->dim[(d)])
#define __PSGridDimDev(p, d) ((p)->dim[d])
#endif

typedef void *(*__PSGrid_devNewFunc)(int num_dims, PSVectorInt dim);
typedef void (*__PSGrid_devFreeFunc)(void *);
typedef void (*__PSGrid_devCopyinFunc)(void *g, const void *src);
typedef void (*__PSGrid_devCopyoutFunc)(void *g, void *dst);

/* Allocates a zero-filled dense grid (x fastest, no padding) on the current
 * device.  With a non-NULL func the point type is user-defined and func
 * builds the device representation. */
extern __PSGrid *__PSGridNew(int elm_size, int num_dims, PSVectorInt dim,
                             __PSGrid_devNewFunc func);
extern void __PSGridSwap(__PSGrid *g);
extern void __PSGridMirror(__PSGrid *g);
extern int __PSGridGetID(__PSGrid *g);
/* Writes one element; the trailing varargs are num_dims PSIndex values. */
extern void __PSGridSet(__PSGrid *g, void *buf, ...);
extern void __PSGridFree(__PSGrid *g, __PSGrid_devFreeFunc func);
/* Blocking host->device / device->host copy of num_elms*elm_size bytes.
 * Copyout waits for every kernel queued on the runtime's stream. */
extern void __PSGridCopyin(void *g, const void *src_array,
                           __PSGrid_devCopyinFunc func);
extern void __PSGridCopyout(void *g, void *dst_array,
                            __PSGrid_devCopyoutFunc func);
/* Aborts with `message` if a device error is pending.  Emitted only when
 * kernel error checking is configured (reference CUDA_KERNEL_ERROR_CHECK). */
extern void __PSCheckDeviceError(const char *message);

/* Reduces every element of the grid (not only a domain) with op and stores
 * the result as the grid's element type in *buf.  Waits for queued kernels. */
extern void __PSReduceGridFloat(void *buf, enum PSReduceOp op, __PSGrid *g);
extern void __PSReduceGridDouble(void *buf, enum PSReduceOp op, __PSGrid *g);

/* The stream every runtime call and generated launch is ordered on
 * (a hipStream_t). */
extern void *__PSGetStream(void);

/* ---- user-defined point types ----------------------------------------------
 * A grid of a C struct point type (DeclareGrid3D(Name, struct T)) lives in
 * HBM as a structure of arrays: each scalar of T -- every element of an
 * array member counts as one -- is a plane of num_elms values, x fastest.
 * Planes are plane_stride elements apart (num_elms rounded up to 64, so each
 * plane starts 256-byte aligned) and ordered by member, then by the member's
 * row-major array index.  Kernels then load one member of 64 neighbouring
 * points with one coalesced access.
 *
 * physisc-hip describes T with a PSHipUtype and emits the devNewFunc /
 * devFreeFunc / devCopyinFunc / devCopyoutFunc that __PSGridNew,
 * __PSGridFree, __PSGridCopyin and __PSGridCopyout take, as thin wrappers
 * of the calls below.  They replace the per-type functions the reference's
 * CUDA builder generates (translator/cuda_runtime_builder.cc:319-368 device
 * type, :369-473 new, :474-718 free, :719-871 copyin/copyout), which
 * transpose on the host; here the AoS <-> SoA transpose runs on the device.
 * The plain PSGridCopyin/PSGridCopyout/PSGridFree also accept such grids. */
typedef struct {
  int elm_size;            /* sizeof(T) on the host */
  int num_members;
  const int *offset;       /* byte offset of member m in T */
  const int *scalar_size;  /* bytes of one scalar of member m: 1, 2, 4 or 8 */
  const int *count;        /* scalars in member m (product of its extents) */
} PSHipUtype;

/* The device representation a devNewFunc returns (g->dev of such a grid). */
typedef struct {
  int dim[3];
  void *p0;                /* first plane */
  int64_t num_elms;
  int64_t plane_stride;    /* elements from one plane to the next */
  size_t bytes;
  const PSHipUtype *type;
  unsigned magic;
} __PSGridUtype_dev;

/* Allocates the zero-filled planes of a grid of `type` (aborts on error). */
extern void *__PSHipUtypeDevNew(const PSHipUtype *type, int num_dims,
                                PSVectorInt dim);
extern void __PSHipUtypeDevFree(void *dev);
/* Blocking host AoS array -> device planes, and back. */
extern void __PSHipUtypeCopyin(void *dev, const void *src);
extern void __PSHipUtypeCopyout(void *dev, void *dst);
/* Device address of the first plane of member m of a user-type grid. */
extern void *__PSHipUtypeMember(const __PSGrid *g, int member);

/* ---- linear offsets (x fastest) ------------------------------------------ */
static inline PSIndex __PSGridGetOffset1D(__PSGrid *g, PSIndex i1) {
  (void)g;
  return i1;
}
static inline PSIndex __PSGridGetOffset2D(__PSGrid *g, PSIndex i1, PSIndex i2) {
  return i1 + i2 * g->dim[0];
}
static inline PSIndex __PSGridGetOffset3D(__PSGrid *g, PSIndex i1, PSIndex i2,
                                          PSIndex i3) {
  return i1 + i2 * g->dim[0] + i3 * g->dim[0] * g->dim[1];
}
/* Periodic variants wrap each index with (i + n) % n, which is exact for
 * i >= -n (reference physis_ref.h:69-81). */
static inline PSIndex __PSGridGetOffsetPeriodic1D(__PSGrid *g, PSIndex i1) {
  return (i1 + g->dim[0]) % g->dim[0];
}
static inline PSIndex __PSGridGetOffsetPeriodic2D(__PSGrid *g, PSIndex i1,
                                                  PSIndex i2) {
  return __PSGridGetOffsetPeriodic1D(g, i1) +
         (i2 + g->dim[1]) % g->dim[1] * g->dim[0];
}
static inline PSIndex __PSGridGetOffsetPeriodic3D(__PSGrid *g, PSIndex i1,
                                                  PSIndex i2, PSIndex i3) {
  return __PSGridGetOffsetPeriodic2D(g, i1, i2) +
         (i3 + g->dim[2]) % g->dim[2] * g->dim[0] * g->dim[1];
}

/* Device-view forms take the by-value view struct (any of the *_dev types,
 * whose leading member is the int dim[] array). */
PS_HD static inline PSIndex __PSGridGetOffset1DDev(const void *g, PSIndex i1) {
  (void)g;
  return i1;
}
PS_HD static inline PSIndex __PSGridGetOffset2DDev(const void *g, PSIndex i1,
                                                   PSIndex i2) {
  const int *dim = (const int *)g;
  return i1 + i2 * dim[0];
}
PS_HD static inline PSIndex __PSGridGetOffset3DDev(const void *g, PSIndex i1,
                                                   PSIndex i2, PSIndex i3) {
  const int *dim = (const int *)g;
  return i1 + i2 * dim[0] + i3 * dim[0] * dim[1];
}
PS_HD static inline PSIndex __PSGridGetOffsetPeriodic1DDev(const void *g,
                                                           PSIndex i1) {
  const int *dim = (const int *)g;
  return (i1 + dim[0]) % dim[0];
}
PS_HD static inline PSIndex __PSGridGetOffsetPeriodic2DDev(const void *g,
                                                           PSIndex i1,
                                                           PSIndex i2) {
  const int *dim = (const int *)g;
  return (i1 + dim[0]) % dim[0] + (i2 + dim[1]) % dim[1] * dim[0];
}
PS_HD static inline PSIndex __PSGridGetOffsetPeriodic3DDev(const void *g,
                                                           PSIndex i1,
                                                           PSIndex i2,
                                                           PSIndex i3) {
  const int *dim = (const int *)g;
  return (i1 + dim[0]) % dim[0] + (i2 + dim[1]) % dim[1] * dim[0] +
         (i3 + dim[2]) % dim[2] * dim[0] * dim[1];
}

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_HIP_H_ */
