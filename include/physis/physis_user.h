/*
 * physis_user.h -- declarations that let a DSL source compile as plain C or
 * C++ (cc -DPHYSIS_USER -fsyntax-only prog.c): a type check of the user's
 * program before physisc-hip translates it.  Nothing declared here is ever
 * called: the translator replaces every intrinsic, so the program built
 * from this header alone is not meant to run.
 *
 * Interface (reference include/physis/physis_user.h:52-116): grid types
 * PSGrid{1,2,3}D{Float,Double,Int,Long} with their PSGrid<r>D<T>New
 * constructors, DeclareGrid{1,2,3}D(name, type) for grids of a user point
 * type, the element accessors PSGridGet / PSGridGetPeriodic / PSGridSet /
 * PSGridEmit / PSGridEmitUtype, PSGridDim, PSStencilMap (and its red-black
 * forms), PSStencilRun and PSReduce.
 *
 * Element types.  A grid handle points to a struct whose only member has the
 * grid's element type; the accessors are comma expressions that end in that
 * member, so PSGridGet(g, ...) has the element type (and .member access on a
 * point type checks), and PSGridEmit(g, v) converts v to it as an
 * assignment would.  Index arguments are evaluated and converted to PSIndex.
 */
#ifndef PHYSIS_PHYSIS_USER_H_
#define PHYSIS_PHYSIS_USER_H_

#include "physis/physis_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Consumes the index arguments of an accessor (type checks them). */
extern int __PSUserIndex(const void *grid, ...);
/* Stand-in for the location an emit writes. */
extern void *__PSUserEmitTarget(const char *grid_and_member, ...);

#define __PS_USER_GRID(R, name, type)                     \
  struct __PSUserGrid##R##D##name {                       \
    type __elem;                                          \
  };                                                      \
  typedef struct __PSUserGrid##R##D##name *PSGrid##R##D##name

#define DeclareGrid1D(name, type) \
  __PS_USER_GRID(1, name, type);  \
  extern PSGrid1D##name PSGrid1D##name##New(PSIndex, ...)
#define DeclareGrid2D(name, type) \
  __PS_USER_GRID(2, name, type);  \
  extern PSGrid2D##name PSGrid2D##name##New(PSIndex, PSIndex, ...)
#define DeclareGrid3D(name, type) \
  __PS_USER_GRID(3, name, type);  \
  extern PSGrid3D##name PSGrid3D##name##New(PSIndex, PSIndex, PSIndex, ...)

DeclareGrid1D(Float, float);
DeclareGrid1D(Double, double);
DeclareGrid1D(Int, int);
DeclareGrid1D(Long, long);
DeclareGrid2D(Float, float);
DeclareGrid2D(Double, double);
DeclareGrid2D(Int, int);
DeclareGrid2D(Long, long);
DeclareGrid3D(Float, float);
DeclareGrid3D(Double, double);
DeclareGrid3D(Int, int);
DeclareGrid3D(Long, long);

#define PSGridGet(g, ...) (__PSUserIndex((g), __VA_ARGS__), (g)->__elem)
#define PSGridGetPeriodic(g, ...) (__PSUserIndex((g), __VA_ARGS__), (g)->__elem)
/* PSGridSet(g, i[, j[, k]], v): the last argument is the value. */
#define PSGridSet(g, ...) ((void)__PSUserIndex((g), __VA_ARGS__))
#define PSGridEmit(g, v) ((void)((g)->__elem = (v)))
/* PSGridEmitUtype(g.member[...], v): one member of a point-type grid; the
 * grid and member are named, not evaluated. */
#define PSGridEmitUtype(gm, v) ((void)__PSUserEmitTarget(#gm, (v)))
#define PSGridEmitDirichlet(g, v) PSGridEmit(g, v)
#define PSGridEmitNeumann(g, v, grad) ((void)(grad), PSGridEmit(g, v))

extern PSIndex PSGridDim(void *g, int d);

typedef int PSStencil;
/* The kernel is passed as an untyped pointer (a C++ compiler would reject
 * the implicit conversion of a function pointer). */
extern PSStencil __PSStencilMap(void *kernel, ...);
extern PSStencil __PSStencilMapRedBlack(void *kernel, ...);
extern PSStencil __PSStencilMapRed(void *kernel, ...);
extern PSStencil __PSStencilMapBlack(void *kernel, ...);
#define PSStencilMap(k, ...) __PSStencilMap((void *)(k), __VA_ARGS__)
#define PSStencilMapRedBlack(k, ...) __PSStencilMapRedBlack((void *)(k), __VA_ARGS__)
#define PSStencilMapRed(k, ...) __PSStencilMapRed((void *)(k), __VA_ARGS__)
#define PSStencilMapBlack(k, ...) __PSStencilMapBlack((void *)(k), __VA_ARGS__)
extern void PSStencilRun(PSStencil s, ...);

/* PSReduce(&result, op, grid), or PSReduce(&result, op, kernel, domain,
 * args...) with a reduction kernel; PSGridReduce is accepted as another name
 * for it. */
extern void PSReduce(void *result, ...);
extern void PSGridReduce(void *result, ...);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_USER_H_ */
