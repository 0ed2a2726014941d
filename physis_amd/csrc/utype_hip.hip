// utype_hip.hip -- grids of user-defined point types (C structs).
//
// The device form is a structure of arrays: one plane of num_elms values per
// scalar of the struct, so a kernel reading member `p` of 64 neighbouring
// points issues one coalesced load instead of 64 strided ones.  Host arrays
// stay arrays of structs (the DSL's PSGridCopyin/Copyout contract), and the
// transpose between the two runs on the device: the host array is copied as
// is into a staging buffer and a kernel scatters its members into the planes
// (copyout gathers them back).  The reference's CUDA builder does the same
// transpose on the host, one member at a time through pinned buffers
// (translator/cuda_runtime_builder.cc:719-871).
#include "runtime_hip.h"

#include <cstdlib>
#include <cstring>

namespace physis_hip {
namespace {

constexpr unsigned kUtypeMagic = 0x55545950u;  // "UTYP"
constexpr int kMaxMembers = 64;
constexpr int64_t kPlaneAlign = 64;  // elements; 256 bytes for 4-byte scalars
constexpr int kThreads = 256;

struct MemberTable {
  int n;
  int elm_size;
  int offset[kMaxMembers];
  int size[kMaxMembers];
  int count[kMaxMembers];
  int64_t base[kMaxMembers];  // byte offset of the member's first plane
};

template <class T>
__device__ __forceinline__ void Move(char *dst, const char *src) {
  *reinterpret_cast<T *>(dst) = *reinterpret_cast<const T *>(src);
}

__device__ __forceinline__ void MoveScalar(char *dst, const char *src, int size) {
  switch (size) {
    case 8: Move<uint64_t>(dst, src); break;
    case 4: Move<uint32_t>(dst, src); break;
    case 2: Move<uint16_t>(dst, src); break;
    default: *dst = *src; break;
  }
}

// Elements [e0, e0 + n) of the grid: aos holds them as structs (element
// e0 + i at aos + i * elm_size); planes is the grid's first plane.  to_planes
// selects the direction.  Every (element, scalar) pair is moved by exactly
// one thread; consecutive threads take consecutive elements, so the plane
// side of each move is coalesced.
__global__ void __launch_bounds__(kThreads)
Transpose(char *aos, char *planes, int64_t e0, int64_t n, int64_t stride, MemberTable t, int to_planes) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    char *rec = aos + i * t.elm_size;
    for (int m = 0; m < t.n; ++m) {
      const int sz = t.size[m];
      for (int c = 0; c < t.count[m]; ++c) {
        char *s = rec + t.offset[m] + c * sz;
        char *p = planes + t.base[m] + ((int64_t)c * stride + e0 + i) * sz;
        if (to_planes) MoveScalar(p, s, sz);
        else MoveScalar(s, p, sz);
      }
    }
  }
}

bool ValidType(const PSHipUtype *t) {
  if (!t || t->elm_size <= 0 || t->num_members <= 0 || t->num_members > kMaxMembers || !t->offset ||
      !t->scalar_size || !t->count)
    return false;
  for (int m = 0; m < t->num_members; ++m) {
    const int s = t->scalar_size[m];
    if ((s != 1 && s != 2 && s != 4 && s != 8) || t->count[m] <= 0 || t->offset[m] < 0 ||
        t->offset[m] % s != 0 || (int64_t)t->offset[m] + (int64_t)s * t->count[m] > t->elm_size)
      return false;
  }
  return true;
}

MemberTable Table(const __PSGridUtype_dev *d) {
  MemberTable mt;
  memset(&mt, 0, sizeof(mt));
  const PSHipUtype *t = d->type;
  mt.n = t->num_members;
  mt.elm_size = t->elm_size;
  int64_t base = 0;
  for (int m = 0; m < t->num_members; ++m) {
    mt.offset[m] = t->offset[m];
    mt.size[m] = t->scalar_size[m];
    mt.count[m] = t->count[m];
    mt.base[m] = base;
    base += (int64_t)t->count[m] * t->scalar_size[m] * d->plane_stride;
  }
  return mt;
}

__PSGridUtype_dev *AsUtype(void *dev) {
  __PSGridUtype_dev *d = static_cast<__PSGridUtype_dev *>(dev);
  return d && d->magic == kUtypeMagic ? d : nullptr;
}

// Moves the whole grid between a host array of structs and the planes,
// through the runtime's staging buffer in chunks of about 32 MiB.
int CopyHost(__PSGridUtype_dev *d, void *host, bool in) {
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  const int64_t es = d->type->elm_size;
  const int64_t chunk = (int64_t)(32 << 20) / es > 0 ? (int64_t)(32 << 20) / es : 1;
  const MemberTable mt = Table(d);
  char *stage = static_cast<char *>(rt.Scratch((size_t)(chunk * es)));
  char *h = static_cast<char *>(host);
  for (int64_t e0 = 0; e0 < d->num_elms; e0 += chunk) {
    const int64_t n = d->num_elms - e0 < chunk ? d->num_elms - e0 : chunk;
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 4096) blocks = 4096;
    if (in) {
      if (hipMemcpyAsync(stage, h + e0 * es, (size_t)(n * es), hipMemcpyHostToDevice, s) != hipSuccess)
        return PS_HIP_ERR_DEVICE;
      hipLaunchKernelGGL(Transpose, dim3((unsigned)blocks), dim3(kThreads), 0, s, stage,
                         static_cast<char *>(d->p0), e0, n, d->plane_stride, mt, 1);
      if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
    } else {
      // Padding bytes of the host structs come back as zeros.
      if (hipMemsetAsync(stage, 0, (size_t)(n * es), s) != hipSuccess) return PS_HIP_ERR_DEVICE;
      hipLaunchKernelGGL(Transpose, dim3((unsigned)blocks), dim3(kThreads), 0, s, stage,
                         static_cast<char *>(d->p0), e0, n, d->plane_stride, mt, 0);
      if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
      if (hipMemcpyAsync(h + e0 * es, stage, (size_t)(n * es), hipMemcpyDeviceToHost, s) != hipSuccess)
        return PS_HIP_ERR_DEVICE;
    }
    // The staging buffer is reused by the next chunk.
    if (hipStreamSynchronize(s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

}  // namespace

bool IsUtypeDev(const void *dev) { return AsUtype(const_cast<void *>(dev)) != nullptr; }

int UtypeCopyin(void *dev, const void *src) {
  __PSGridUtype_dev *d = AsUtype(dev);
  if (!d || !src) return PS_HIP_ERR_INVALID;
  return CopyHost(d, const_cast<void *>(src), true);
}

int UtypeCopyout(void *dev, void *dst) {
  __PSGridUtype_dev *d = AsUtype(dev);
  if (!d || !dst) return PS_HIP_ERR_INVALID;
  return CopyHost(d, dst, false);
}

int UtypeFree(void *dev) {
  __PSGridUtype_dev *d = AsUtype(dev);
  if (!d) return PS_HIP_ERR_INVALID;
  Runtime &rt = Runtime::Get();
  // Kernels queued on the stream may still read the planes.
  if (hipStreamSynchronize(rt.stream()) != hipSuccess) return PS_HIP_ERR_DEVICE;
  if (d->p0 && hipFree(d->p0) != hipSuccess) return PS_HIP_ERR_DEVICE;
  rt.AddGridBytes(-(ptrdiff_t)d->bytes);
  d->magic = 0;
  free(d);
  return PS_HIP_OK;
}

}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

void *__PSHipUtypeDevNew(const PSHipUtype *type, int num_dims, PSVectorInt dim) {
  if (!ValidType(type)) Fatal("__PSHipUtypeDevNew", "invalid point type description");
  if (num_dims < 1 || num_dims > PS_MAX_DIM) Fatal("__PSHipUtypeDevNew", "invalid grid rank");
  Runtime &rt = Runtime::Get();
  rt.EnsureDevice();
  __PSGridUtype_dev *d = static_cast<__PSGridUtype_dev *>(calloc(1, sizeof(__PSGridUtype_dev)));
  if (!d) Fatal("__PSHipUtypeDevNew", "out of host memory");
  d->num_elms = 1;
  for (int i = 0; i < 3; ++i) {
    d->dim[i] = i < num_dims ? dim[i] : 1;
    if (d->dim[i] <= 0) Fatal("__PSHipUtypeDevNew", "invalid grid shape");
    d->num_elms *= d->dim[i];
  }
  if (d->num_elms > (int64_t)INT32_MAX) Fatal("__PSHipUtypeDevNew", "grid too large for PSIndex");
  d->plane_stride = (d->num_elms + kPlaneAlign - 1) / kPlaneAlign * kPlaneAlign;
  int64_t scalars_bytes = 0;
  for (int m = 0; m < type->num_members; ++m) scalars_bytes += (int64_t)type->count[m] * type->scalar_size[m];
  d->bytes = (size_t)(scalars_bytes * d->plane_stride);
  d->type = type;
  if (hipMalloc(&d->p0, d->bytes) != hipSuccess || !d->p0) {
    (void)hipGetLastError();
    free(d);
    Fatal("__PSHipUtypeDevNew", "out of device memory");
  }
  // Zero fill, like the primitive grids.
  CheckHip(hipMemsetAsync(d->p0, 0, d->bytes, rt.stream()), "utype zero fill");
  rt.AddGridBytes((ptrdiff_t)d->bytes);
  d->magic = kUtypeMagic;
  return d;
}

void __PSHipUtypeDevFree(void *dev) {
  if (UtypeFree(dev) != PS_HIP_OK) Fatal("__PSHipUtypeDevFree", "not a user-type grid, or a device error");
}

void __PSHipUtypeCopyin(void *dev, const void *src) {
  if (UtypeCopyin(dev, src) != PS_HIP_OK) Fatal("__PSHipUtypeCopyin", "copy failed");
}

void __PSHipUtypeCopyout(void *dev, void *dst) {
  // Kernels queued on the runtime stream are ordered before the copy.
  if (UtypeCopyout(dev, dst) != PS_HIP_OK) Fatal("__PSHipUtypeCopyout", "copy failed");
}

void *__PSHipUtypeMember(const __PSGrid *g, int member) {
  __PSGridUtype_dev *d = g ? AsUtype(g->dev) : nullptr;
  if (!d || member < 0 || member >= d->type->num_members) Fatal("__PSHipUtypeMember", "invalid grid or member");
  const MemberTable mt = Table(d);
  return static_cast<char *>(d->p0) + mt.base[member];
}

}  // extern "C"
