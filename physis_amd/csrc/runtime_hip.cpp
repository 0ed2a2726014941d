// runtime_hip.cpp -- host side of libphysis_rt_hip: initialisation, grid
// lifetime, host<->device copies, domains, tracing and error checks.
//
// Entry points follow the CUDA target's runtime one for one
// (reference runtime/libphysis_rt_cuda.cc:28-191, runtime/runtime.cc:16-91);
// the implementation is HIP on gfx950 with one runtime-owned stream.
#include "runtime_hip.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>

FILE *__ps_trace = nullptr;

namespace physis_hip {

void Fatal(const char *where, const char *what) {
  fprintf(stderr, "physis-hip: ERROR: %s: %s\n", where, what);
  fflush(stderr);
  PSAbort(1);
  abort();
}

void CheckHip(hipError_t e, const char *where) {
  if (e != hipSuccess) Fatal(where, hipGetErrorString(e));
}

Runtime &Runtime::Get() {
  static Runtime rt;
  return rt;
}

void Runtime::EnsureDevice() {
  if (inited_) return;
  std::lock_guard<std::mutex> lk(mu_);
  if (inited_) return;
  int n = 0;
  CheckHip(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (n < 1) Fatal("PSInit", "no HIP device");
  CheckHip(hipSetDevice(device_), "hipSetDevice");
  // Default priority: streams of different priorities are not used.
  CheckHip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
           "hipStreamCreate");
  // The no-allocation reduction's scratch address on this device; a failure
  // here is reported by that entry point when it is called.
  (void)NoAllocScratchInit();
  inited_ = true;
}

void *Runtime::Scratch(size_t bytes) {
  EnsureDevice();
  if (bytes > scratch_bytes_) {
    if (scratch_) {
      CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      CheckHip(hipFree(scratch_), "hipFree(scratch)");
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    CheckHip(hipMalloc(&scratch_, want), "hipMalloc(scratch)");
    scratch_bytes_ = want;
  }
  return scratch_;
}

void *Runtime::PinnedScratch(size_t bytes) {
  EnsureDevice();
  if (bytes > pinned_bytes_) {
    if (pinned_) {
      CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      CheckHip(hipHostFree(pinned_), "hipHostFree");
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    CheckHip(hipHostMalloc(&pinned_, want, hipHostMallocDefault),
             "hipHostMalloc");
    pinned_bytes_ = want;
  }
  return pinned_;
}

void *Runtime::Aux(size_t bytes) {
  EnsureDevice();
  if (bytes > aux_bytes_) {
    if (aux_) {
      CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      CheckHip(hipFree(aux_), "hipFree(aux)");
      aux_ = nullptr;
      aux_bytes_ = 0;
    }
    if (hipMalloc(&aux_, bytes) != hipSuccess) {
      (void)hipGetLastError();
      aux_ = nullptr;
      return nullptr;
    }
    aux_bytes_ = bytes;
  }
  return aux_;
}

void Runtime::Shutdown() {
  if (!inited_) return;
  (void)hipStreamSynchronize(stream_);
  ReleaseHostCopyPool();
  if (scratch_) (void)hipFree(scratch_);
  if (pinned_) (void)hipHostFree(pinned_);
  if (aux_) (void)hipFree(aux_);
  scratch_ = nullptr;
  pinned_ = nullptr;
  aux_ = nullptr;
  scratch_bytes_ = pinned_bytes_ = aux_bytes_ = 0;
  (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  inited_ = false;
}

size_t StoredElmSize(int elm_size, int storage) {
  return storage == PS_HIP_STORAGE_NATIVE ? (size_t)elm_size : 2u;
}

namespace {

// Removes "-name"/"--name" from argv in place (reference runtime/runtime.cc:16-44).
bool TakeOption(int *argc, char ***argv, const char *name) {
  if (!argc || !argv || !*argv) return false;
  std::string a1 = std::string("-") + name, a2 = std::string("--") + name;
  for (int i = 0; i < *argc; ++i) {
    const char *s = (*argv)[i];
    if (!s || (a1 != s && a2 != s)) continue;
    for (int j = i; j + 1 < *argc; ++j) (*argv)[j] = (*argv)[j + 1];
    --*argc;
    return true;
  }
  return false;
}

int GridNewImpl(__PSGrid **out, int elm_size, int num_dims, const int *dim,
                int storage, __PSGrid_devNewFunc func) {
  *out = nullptr;
  if (num_dims < 1 || num_dims > PS_MAX_DIM || elm_size <= 0)
    return PS_HIP_ERR_INVALID;
  if (storage != PS_HIP_STORAGE_NATIVE && elm_size != 4)
    return PS_HIP_ERR_INVALID;  // 16-bit storage is for float grids only
  Runtime &rt = Runtime::Get();
  rt.EnsureDevice();
  GridRec *r = static_cast<GridRec *>(calloc(1, sizeof(GridRec)));
  if (!r) return PS_HIP_ERR_NOMEM;
  __PSGrid *g = &r->pub;
  g->elm_size = elm_size;
  g->num_dims = num_dims;
  g->num_elms = 1;
  for (int d = 0; d < PS_MAX_DIM; ++d) {
    g->dim[d] = d < num_dims ? dim[d] : 1;
    if (d < num_dims) {
      if (dim[d] <= 0) {
        free(r);
        return PS_HIP_ERR_INVALID;
      }
      g->num_elms *= dim[d];
    }
  }
  if (g->num_elms > (int64_t)INT32_MAX) {
    // PSIndex is int32: offsets beyond 2^31-1 elements are not addressable.
    free(r);
    return PS_HIP_ERR_INVALID;
  }
  r->storage = storage;
  r->stored_elm_size = StoredElmSize(elm_size, storage);
  if (func) {
    // User-defined point type: the generated helper builds the device form.
    r->user_type = true;
    g->dev = static_cast<__PSGrid_dev *>(func(num_dims, g->dim));
    // p0 stays NULL: calls that address a grid as one array of primitive
    // elements (fills, reductions, device copies, built-in programs) reject
    // it; the planes are reached through __PSHipUtypeMember.
    g->p0 = nullptr;
    *out = g;
    return PS_HIP_OK;
  }
  r->bytes = r->stored_elm_size * (size_t)g->num_elms;
  void *buf = nullptr;
  if (hipMalloc(&buf, r->bytes) != hipSuccess || !buf) {
    (void)hipGetLastError();
    free(r);
    return PS_HIP_ERR_NOMEM;
  }
  // Zero fill, like calloc in the ref runtime (libphysis_rt_ref.cc:70).
  if (hipMemsetAsync(buf, 0, r->bytes, rt.stream()) != hipSuccess) {
    (void)hipFree(buf);
    free(r);
    return PS_HIP_ERR_DEVICE;
  }
  g->p0 = buf;
  r->view.p0 = buf;
  for (int d = 0; d < PS_MAX_DIM; ++d) r->view.dim[d] = g->dim[d];
  g->dev = &r->view;
  rt.AddGridBytes((ptrdiff_t)r->bytes);
  *out = g;
  return PS_HIP_OK;
}

// The staging buffers PSGridCopyin/Copyout of grid g will use; host arrays
// hold elm_size-byte elements (FP32 for 16-bit storage).
void PrepareCopies(const __PSGrid *g) {
  if (g && g != INVALID_GRID && g->p0) PrepareHostCopy((size_t)g->elm_size * (size_t)g->num_elms);
}

}  // namespace

int GridNewDevice(__PSGrid **out, int elm_size, int num_dims, const int *dim, int storage) {
  return GridNewImpl(out, elm_size, num_dims, dim, storage, nullptr);
}

}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

void PSInit(int *argc, char ***argv, int grid_num_dims, ...) {
  (void)grid_num_dims;  // the varargs maximum extents are unused on one device
  __ps_trace = nullptr;
  if (TakeOption(argc, argv, "physis-trace")) {
    __ps_trace = stderr;
    Runtime::Get().set_trace(true);
    fprintf(stderr, "physis-hip: tracing enabled\n");
  }
  Runtime::Get().EnsureDevice();
}

void PSFinalize(void) {
  Runtime::Get().Shutdown();
}

void PSPrintInternalInfo(FILE *out) {
  Runtime &rt = Runtime::Get();
  rt.EnsureDevice();
  hipDeviceProp_t p;
  PSH_CHECK(hipGetDeviceProperties(&p, rt.device()));
  fprintf(out, "physis-hip: device %d %s (%s), %d CUs, %.1f GiB, grids %.3f GiB\n",
          rt.device(), p.name, p.gcnArchName, p.multiProcessorCount,
          p.totalGlobalMem / 1073741824.0, rt.grid_bytes() / 1073741824.0);
}

int __PSGridGetID(__PSGrid *g) {
  (void)g;
  return 0;  // ids only matter for multi-process runtimes
}

__PSGrid *__PSGridNew(int elm_size, int num_dims, PSVectorInt dim,
                      __PSGrid_devNewFunc func) {
  __PSGrid *g = nullptr;
  int st = GridNewImpl(&g, elm_size, num_dims, dim, PS_HIP_STORAGE_NATIVE, func);
  if (st == PS_HIP_ERR_NOMEM) return INVALID_GRID;
  if (st != PS_HIP_OK) Fatal("__PSGridNew", "invalid grid shape");
  PrepareCopies(g);
  return g;
}

__PSGrid *__PSGridNewAttr(int elm_size, int num_dims, PSVectorInt dim,
                          int attr) {
  int storage = PS_HIP_STORAGE_NATIVE;
  if (attr & PS_GRID_STORAGE_BF16) storage = PS_HIP_STORAGE_BF16;
  else if (attr & PS_GRID_STORAGE_FP16) storage = PS_HIP_STORAGE_FP16;
  __PSGrid *g = nullptr;
  int st = GridNewImpl(&g, elm_size, num_dims, dim, storage, nullptr);
  if (st == PS_HIP_ERR_NOMEM) return INVALID_GRID;
  if (st != PS_HIP_OK) Fatal("__PSGridNewAttr", "invalid grid shape or storage");
  PrepareCopies(g);
  return g;
}

int PSHipGridNew(__PSGrid **out, int elm_size, int num_dims, const int *dim,
                 int storage) {
  if (!out || !dim) return PS_HIP_ERR_INVALID;
  if (storage < PS_HIP_STORAGE_NATIVE || storage > PS_HIP_STORAGE_FP16)
    return PS_HIP_ERR_INVALID;
  const int st = GridNewImpl(out, elm_size, num_dims, dim, storage, nullptr);
  if (st == PS_HIP_OK) PrepareCopies(*out);
  return st;
}

int PSHipGridFree(__PSGrid *g) {
  if (!g) return PS_HIP_ERR_INVALID;
  GridRec *r = Rec(g);
  Runtime &rt = Runtime::Get();
  if (r->user_type) {
    if (g->dev && !IsUtypeDev(g->dev)) return PS_HIP_ERR_INVALID;  // needs its own devFreeFunc
    if (g->dev && UtypeFree(g->dev) != PS_HIP_OK) return PS_HIP_ERR_DEVICE;
  } else if (g->dev && g->p0) {
    // Kernels queued on the stream may still read the buffer.
    if (hipStreamSynchronize(rt.stream()) != hipSuccess) return PS_HIP_ERR_DEVICE;
    if (hipFree(g->p0) != hipSuccess) return PS_HIP_ERR_DEVICE;
    rt.AddGridBytes(-(ptrdiff_t)r->bytes);
  }
  g->p0 = nullptr;
  g->dev = nullptr;
  free(r);
  return PS_HIP_OK;
}

void __PSGridFree(__PSGrid *g, __PSGrid_devFreeFunc func) {
  if (!g) return;
  if (func && g->dev) {
    func(g->dev);
    g->dev = nullptr;
    free(Rec(g));
    return;
  }
  if (PSHipGridFree(g) != PS_HIP_OK) Fatal("__PSGridFree", "device error");
}

void PSGridFree(void *g) { __PSGridFree(static_cast<__PSGrid *>(g), nullptr); }

int PSHipGridCopyin(__PSGrid *g, const void *src) {
  if (!g || !src) return PS_HIP_ERR_INVALID;
  GridRec *r = Rec(g);
  if (r->user_type) return UtypeCopyin(g->dev, src);
  if (!g->p0) return PS_HIP_ERR_INVALID;
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  if (r->storage == PS_HIP_STORAGE_NATIVE) return CopyHostToDevice(g->p0, src, r->bytes, s);
  // FP32 host data -> 16-bit device storage, converted on the device in
  // chunks through a staging buffer.
  const int64_t chunk = 8 << 20;  // elements (32 MiB of FP32)
  float *stage = static_cast<float *>(rt.Scratch(chunk * sizeof(float)));
  const float *h = static_cast<const float *>(src);
  char *d = static_cast<char *>(g->p0);
  for (int64_t off = 0; off < g->num_elms; off += chunk) {
    int64_t n = g->num_elms - off < chunk ? g->num_elms - off : chunk;
    // Blocking, and ordered after the previous chunk's conversion.
    if (CopyHostToDevice(stage, h + off, n * sizeof(float), s) != PS_HIP_OK) return PS_HIP_ERR_DEVICE;
    ConvertToStorage(d + off * r->stored_elm_size, stage, n, r->storage, s);
  }
  return hipStreamSynchronize(s) == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

int PSHipGridCopyout(__PSGrid *g, void *dst) {
  if (!g || !dst) return PS_HIP_ERR_INVALID;
  GridRec *r = Rec(g);
  if (r->user_type) return UtypeCopyout(g->dev, dst);
  if (!g->p0) return PS_HIP_ERR_INVALID;
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  if (r->storage == PS_HIP_STORAGE_NATIVE) return CopyDeviceToHost(dst, g->p0, r->bytes, s);
  const int64_t chunk = 8 << 20;
  float *stage = static_cast<float *>(rt.Scratch(chunk * sizeof(float)));
  float *h = static_cast<float *>(dst);
  const char *d = static_cast<const char *>(g->p0);
  for (int64_t off = 0; off < g->num_elms; off += chunk) {
    int64_t n = g->num_elms - off < chunk ? g->num_elms - off : chunk;
    ConvertFromStorage(stage, d + off * r->stored_elm_size, n, r->storage, s);
    // Blocking: the staging buffer is reused by the next chunk.
    if (CopyDeviceToHost(h + off, stage, n * sizeof(float), s) != PS_HIP_OK) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

int PSHipGridCopyDevice(__PSGrid *dst, const __PSGrid *src) {
  if (!dst || !src || !dst->p0 || !src->p0) return PS_HIP_ERR_INVALID;
  const GridRec *rs = Rec(src);
  GridRec *rd = Rec(dst);
  if (rs->bytes != rd->bytes || rs->storage != rd->storage) return PS_HIP_ERR_INVALID;
  return hipMemcpyAsync(dst->p0, src->p0, rs->bytes, hipMemcpyDeviceToDevice,
                        Runtime::Get().stream()) == hipSuccess
             ? PS_HIP_OK
             : PS_HIP_ERR_DEVICE;
}

void __PSGridCopyin(void *p, const void *src_array, __PSGrid_devCopyinFunc func) {
  __PSGrid *g = static_cast<__PSGrid *>(p);
  if (func) {
    func(g->dev, src_array);
    return;
  }
  if (PSHipGridCopyin(g, src_array) != PS_HIP_OK) Fatal("__PSGridCopyin", "copy failed");
}

void __PSGridCopyout(void *p, void *dst_array, __PSGrid_devCopyoutFunc func) {
  __PSGrid *g = static_cast<__PSGrid *>(p);
  if (func) {
    PSH_CHECK(hipStreamSynchronize(Runtime::Get().stream()));
    func(g->dev, dst_array);
    return;
  }
  if (PSHipGridCopyout(g, dst_array) != PS_HIP_OK) Fatal("__PSGridCopyout", "copy failed");
}

void PSGridCopyin(void *g, const void *src_array) { __PSGridCopyin(g, src_array, nullptr); }
void PSGridCopyout(void *g, void *dst_array) { __PSGridCopyout(g, dst_array, nullptr); }

// Double buffering is explicit (two grids, two maps) on this target, as on
// the CUDA target (libphysis_rt_cuda.cc:141-142): swap and mirror are no-ops.
void __PSGridSwap(__PSGrid *g) { (void)g; }
void __PSGridMirror(__PSGrid *g) { (void)g; }

void __PSGridSet(__PSGrid *g, void *buf, ...) {
  va_list vl;
  va_start(vl, buf);
  int64_t offset = 0, stride = 1;
  for (int d = 0; d < g->num_dims; ++d) {
    PSIndex idx = va_arg(vl, PSIndex);
    offset += (int64_t)idx * stride;
    stride *= g->dim[d];
  }
  va_end(vl);
  GridRec *r = Rec(g);
  if (r->user_type) Fatal("__PSGridSet", "not supported for user-defined point types");
  hipStream_t s = Runtime::Get().stream();
  if (r->storage == PS_HIP_STORAGE_NATIVE) {
    PSH_CHECK(hipMemcpyAsync(static_cast<char *>(g->p0) + offset * g->elm_size, buf,
                             g->elm_size, hipMemcpyHostToDevice, s));
  } else {
    float *stage = static_cast<float *>(Runtime::Get().Scratch(sizeof(float)));
    PSH_CHECK(hipMemcpyAsync(stage, buf, sizeof(float), hipMemcpyHostToDevice, s));
    ConvertToStorage(static_cast<char *>(g->p0) + offset * 2, stage, 1, r->storage, s);
  }
  PSH_CHECK(hipStreamSynchronize(s));
}

void __PSCheckDeviceError(const char *message) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(Runtime::Get().stream());
  if (e != hipSuccess) {
    fprintf(stderr, "ERROR: %s: %s\n", message, hipGetErrorString(e));
    PSAbort(1);
  }
}

void *__PSGetStream(void) { return Runtime::Get().stream(); }

PSDomain1D PSDomain1DNew(PSIndex minx, PSIndex maxx) {
  PSDomain1D d;
  memset(&d, 0, sizeof(d));
  d.min[0] = d.local_min[0] = minx;
  d.max[0] = d.local_max[0] = maxx;
  return d;
}

PSDomain2D PSDomain2DNew(PSIndex minx, PSIndex maxx, PSIndex miny, PSIndex maxy) {
  PSDomain2D d;
  memset(&d, 0, sizeof(d));
  d.min[0] = d.local_min[0] = minx;
  d.max[0] = d.local_max[0] = maxx;
  d.min[1] = d.local_min[1] = miny;
  d.max[1] = d.local_max[1] = maxy;
  return d;
}

PSDomain3D PSDomain3DNew(PSIndex minx, PSIndex maxx, PSIndex miny, PSIndex maxy,
                         PSIndex minz, PSIndex maxz) {
  PSDomain3D d;
  d.min[0] = d.local_min[0] = minx;
  d.max[0] = d.local_max[0] = maxx;
  d.min[1] = d.local_min[1] = miny;
  d.max[1] = d.local_max[1] = maxy;
  d.min[2] = d.local_min[2] = minz;
  d.max[2] = d.local_max[2] = maxz;
  return d;
}

int PSHipGridStorage(const __PSGrid *g) { return g ? Rec(g)->storage : -1; }
int PSHipGridStoredElmSize(const __PSGrid *g) {
  return g ? (int)Rec(g)->stored_elm_size : -1;
}
void *PSHipGridDevicePtr(const __PSGrid *g) { return g ? g->p0 : nullptr; }

void *__PSHipAuxBuffer(size_t bytes) { return Runtime::Get().Aux(bytes); }

int PSHipSynchronize(void) {
  return hipStreamSynchronize(Runtime::Get().stream()) == hipSuccess ? PS_HIP_OK
                                                                      : PS_HIP_ERR_DEVICE;
}

int64_t PSHipGridBytesInUse(void) { return (int64_t)Runtime::Get().grid_bytes(); }

int PSHipDeviceName(char *buf, int len) {
  if (!buf || len <= 0) return PS_HIP_ERR_INVALID;
  Runtime &rt = Runtime::Get();
  rt.EnsureDevice();
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, rt.device()) != hipSuccess) return PS_HIP_ERR_DEVICE;
  snprintf(buf, len, "%s %s %dCU", p.name, p.gcnArchName, p.multiProcessorCount);
  return PS_HIP_OK;
}

int PSHipMemInfo(int64_t *free_bytes, int64_t *total_bytes) {
  Runtime::Get().EnsureDevice();
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return PS_HIP_ERR_DEVICE;
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return PS_HIP_OK;
}

static hipEvent_t g_ev_begin = nullptr, g_ev_end = nullptr;

int PSHipTimerBegin(void) {
  Runtime &rt = Runtime::Get();
  if (!g_ev_begin) {
    if (hipEventCreate(&g_ev_begin) != hipSuccess) return PS_HIP_ERR_DEVICE;
    if (hipEventCreate(&g_ev_end) != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return hipEventRecord(g_ev_begin, rt.stream()) == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

int PSHipTimerEnd(void) {
  if (!g_ev_end) return PS_HIP_ERR_INVALID;
  return hipEventRecord(g_ev_end, Runtime::Get().stream()) == hipSuccess ? PS_HIP_OK
                                                                          : PS_HIP_ERR_DEVICE;
}

int PSHipTimerElapsed(float *ms) {
  if (!g_ev_end || !ms) return PS_HIP_ERR_INVALID;
  if (hipEventSynchronize(g_ev_end) != hipSuccess) return PS_HIP_ERR_DEVICE;
  return hipEventElapsedTime(ms, g_ev_begin, g_ev_end) == hipSuccess ? PS_HIP_OK
                                                                      : PS_HIP_ERR_DEVICE;
}

}  // extern "C"
