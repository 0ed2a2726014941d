// runtime_hip.h -- internal interface of libphysis_rt_hip (not installed).
//
// The public handle type __PSGrid is the first member of GridRec, so a
// __PSGrid* returned to generated code can be turned back into the record.
// Generated code only ever touches the public prefix (p0, dim, elm_size,
// num_dims, num_elms, dev), exactly the fields the CUDA target reads
// (reference translator/cuda_translator.cc:426-429, 476-489).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "physis/physis_hip.h"
#include "physis/physis_hip_ext.h"

namespace physis_hip {

[[noreturn]] void Fatal(const char *where, const char *what);
void CheckHip(hipError_t e, const char *where);

#define PSH_CHECK(call) ::physis_hip::CheckHip((call), #call)

// Per-grid record.  `pub` must stay the first member.
struct GridRec {
  __PSGrid pub;
  __PSGrid_dev view;       // pub.dev points here
  int storage;             // PSHipStorage
  size_t stored_elm_size;  // bytes per element in HBM
  size_t bytes;            // device allocation size
  bool user_type;          // allocated through a devNewFunc
};

inline GridRec *Rec(__PSGrid *g) { return reinterpret_cast<GridRec *>(g); }
inline const GridRec *Rec(const __PSGrid *g) {
  return reinterpret_cast<const GridRec *>(g);
}

class Runtime {
 public:
  static Runtime &Get();
  // Device selection and stream creation; idempotent.
  void EnsureDevice();
  hipStream_t stream() {
    EnsureDevice();
    return stream_;
  }
  int device() const { return device_; }
  // Device bytes currently held by grids created through this runtime.
  size_t grid_bytes() const { return grid_bytes_; }
  void AddGridBytes(ptrdiff_t d) { grid_bytes_ += d; }
  // A device scratch buffer of at least `bytes`, reused across calls that
  // are ordered on stream().  Grows, never shrinks.
  void *Scratch(size_t bytes);
  // Pinned host scratch of at least `bytes` (for small results).
  void *PinnedScratch(size_t bytes);
  // A device buffer of at least `bytes` that stencil programs use as a third
  // grid buffer (fused time steps); separate from Scratch().
  void *Aux(size_t bytes);
  bool trace() const { return trace_; }
  void set_trace(bool t) { trace_ = t; }
  void Shutdown();

 private:
  Runtime() = default;
  bool inited_ = false;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  size_t grid_bytes_ = 0;
  void *scratch_ = nullptr;
  size_t scratch_bytes_ = 0;
  void *pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
  void *aux_ = nullptr;
  size_t aux_bytes_ = 0;
  bool trace_ = false;
  std::mutex mu_;
};

// Blocking transfers between host memory (pageable or not) and device memory,
// ordered after the work already queued on s; large ones are pipelined
// through pinned staging buffers by several host threads (hostcopy.cpp).
int CopyHostToDevice(void *dst, const void *src, size_t bytes, hipStream_t s);
int CopyDeviceToHost(void *dst, const void *src, size_t bytes, hipStream_t s);
// Creates the staging buffers a transfer of `bytes` uses (pinning host
// memory takes milliseconds: grid creation pays it, not the first copy).
void PrepareHostCopy(size_t bytes);
// Frees the staging buffers, streams and events of those transfers.
void ReleaseHostCopyPool();

// Device-side storage conversions for 16-bit grids (buffer_hip.hip).
void ConvertToStorage(void *dst, const float *src, int64_t n, int storage,
                      hipStream_t s);
void ConvertFromStorage(float *dst, const void *src, int64_t n, int storage,
                        hipStream_t s);

size_t StoredElmSize(int elm_size, int storage);

// Stream-explicit building blocks used by the solve server (serve.hip).
// PSHipGridNew without preparing host-copy staging (grids that never leave
// the device).
int GridNewDevice(__PSGrid **out, int elm_size, int num_dims, const int *dim, int storage);
int FillUniformOn(__PSGrid *g, uint64_t seed, float lo, float hi, hipStream_t s);
int FillPatternOn(__PSGrid *g, int profile, float value, hipStream_t s);
// FP64 reduction of n elements; result at result_dev[0], partials in
// result_dev[1 .. ReduceScratchDoubles()).
int ReduceAsync(const void *p, int64_t n, int elm_size, int storage, enum PSReduceOp op, double *result_dev,
                double *partials_dev, hipStream_t s);
size_t ReduceScratchDoubles();
// Resolves the device address of the no-allocation reduction's scratch array
// on the current device (once per device); PS_HIP_OK or an error status.
int NoAllocScratchInit();
// Himeno sweeps on stream s with caller scratch: gosa_dev[i] (may be NULL)
// receives sweep i's FP64 sum of ss^2; partials needs HimenoPartials(p0).
int HimenoEnqueue(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                  double *gosa_dev, double *partials, hipStream_t s);
int HimenoPartials(const __PSGrid *p0);
// The same with `degree` sweeps fused per pass (tiles of 64 x by points,
// rpw rows per wave, zchunk planes per workgroup); aux is a third buffer of
// the grid's size (unused for degree 1); partials needs HimenoTBPartials(...)
// doubles.
int HimenoEnqueueTB(__PSGrid *p0, __PSGrid *p1, void *aux, __PSGrid *const *grids, float omega, int sweeps,
                    int degree, int by, int rpw, int zchunk, double *gosa_dev, double *partials, hipStream_t s);
int HimenoTBPartials(const __PSGrid *p0, int degree, int by, int rpw, int zchunk);

// User-defined point types (utype_hip.hip): `dev` is the __PSGridUtype_dev a
// generated devNewFunc returned through __PSHipUtypeDevNew.
bool IsUtypeDev(const void *dev);
int UtypeCopyin(void *dev, const void *src);
int UtypeCopyout(void *dev, void *dst);
int UtypeFree(void *dev);

}  // namespace physis_hip
