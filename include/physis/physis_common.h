/*
 * physis_common.h -- declarations shared by DSL user code and the code that
 * physisc-hip generates for the MI355X target.
 *
 * ABI notes (each item is an interface the generated code and existing user
 * programs already depend on, so its shape is kept):
 *   - PSIndex is a 32-bit signed index (reference include/physis/physis_common.h:30-44).
 *   - PSVectorInt is int[3] (physis_common.h:56).
 *   - __PSDomain is four PSIndex[3] arrays in the order min, max,
 *     local_min, local_max; every box is half-open [min, max)
 *     (physis_common.h:86-99).
 *   - PSInit / PSFinalize / PSDomain{1,2,3}DNew / PSGridCopyin / PSGridCopyout /
 *     PSGridFree keep their C signatures (physis_common.h:76-84, 95-99).
 */
#ifndef PHYSIS_PHYSIS_COMMON_H_
#define PHYSIS_PHYSIS_COMMON_H_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>

#include "physis/config.h"
#include "physis/stopwatch.h"
#include "physis/types.h"
#include "physis/reduce.h"
#include "physis/runtime.h"

#define PS_MAX_DIM (3)

#define PSAssert(e) assert(e)

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t PSIndex;
#define PSINDEX_MAX INT32_MAX
#define PSINDEX_MIN INT32_MIN

enum physis_error_code {
  PHYSIS_OUTOFMEMORY = 1
};

typedef int PSVectorInt[PS_MAX_DIM];
typedef PSVectorInt PSPoint;

static inline void PSVectorIntInit(PSVectorInt v, int value) {
  for (int d = 0; d < PS_MAX_DIM; ++d) v[d] = value;
}
static inline void PSVectorIntCopy(PSVectorInt dst, const PSVectorInt src) {
  for (int d = 0; d < PS_MAX_DIM; ++d) dst[d] = src[d];
}
static inline void PSPointInit(PSPoint p, int value) {
  PSVectorIntInit(p, value);
}

extern void PSInit(int *argc, char ***argv, int grid_num_dims, ...);
extern void PSFinalize(void);
extern void PSPrintInternalInfo(FILE *out);

extern void PSGridCopyin(void *g, const void *src_array);
extern void PSGridCopyout(void *g, void *dst_array);
extern void PSGridFree(void *g);

typedef struct {
  PSIndex min[PS_MAX_DIM];
  PSIndex max[PS_MAX_DIM];
  PSIndex local_min[PS_MAX_DIM];
  PSIndex local_max[PS_MAX_DIM];
} __PSDomain;
typedef __PSDomain PSDomain1D;
typedef __PSDomain PSDomain2D;
typedef __PSDomain PSDomain3D;

extern PSDomain1D PSDomain1DNew(PSIndex minx, PSIndex maxx);
extern PSDomain2D PSDomain2DNew(PSIndex minx, PSIndex maxx,
                                PSIndex miny, PSIndex maxy);
extern PSDomain3D PSDomain3DNew(PSIndex minx, PSIndex maxx,
                                PSIndex miny, PSIndex maxy,
                                PSIndex minz, PSIndex maxz);

/* Terminates the program; every runtime error in the void C ABI ends here. */
static inline void PSAbort(int code) {
  exit(code);
}

/* Grid attributes passed as the optional trailing argument of
 * PSGrid{1,2,3}D<Type>New.  The reference defines only a dummy value
 * (physis_common.h:146-150); the storage attributes are additions of this
 * target: FP32 grids whose device copy is held as 16-bit values while all
 * arithmetic stays FP32. */
enum PS_GRID_ATTRIBUTE {
  PS_GRID_ATTRIBUTE_DUMMY = 1 << 0,
  PS_GRID_STORAGE_BF16 = 1 << 1,
  PS_GRID_STORAGE_FP16 = 1 << 2
};

#define INVALID_GRID (NULL)

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_COMMON_H_ */
