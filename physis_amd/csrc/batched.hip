// batched.hip -- many uneven grids advanced together (7-point diffusion).
//
// A batch is a set of independent grid pairs (A_i, B_i) of any shape with
// their own coefficients.  Each time step is ONE launch covering every grid:
// the launch's workgroups are persistent and stride over a work list of
// (grid, row tile, z chunk) items sorted largest grid first.  With fused
// steps on (temporal degree 2) each launch advances every grid two steps
// through a per-grid third buffer, as the single-grid path does, and the last
// fused pass also stores its intermediate step, so every B_i ends as the
// reference leaves it.  All the
// launches of a run are captured once into a hipGraph and replayed as one
// graph launch.
//
// The launch shape (points per lane NP, rows per workgroup BY, z chunk,
// temporal degree) is picked by an in-process autotuner that times the
// candidates on the batch itself (grids restored afterwards) and caches the
// winner per batch signature.
//
// No reference counterpart: the reference runs one grid per PSStencilRun.
#include "runtime_hip.h"
#include "diffusion7_body.h"
#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"
#include "physis/hip/skeleton_star_tb.h"
#include "physis/hip/skeleton_pack.h"
#include "physis/hip/skeleton_stride.h"
#include "physis/hip/storage.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace physis_hip {
namespace {

// The diffusion body (reference examples/diffusion-benchmark/diffusion3d_physis.c:29-58;
// periodic form test_7-pt-periodic.c:12-24), one value per point.
// Grid pointers are untyped so that one argument table serves every storage
// kind.
struct BatchArgs {
  const void *in;
  void *out;
  int dim[3];
  Coef7 k;
};

template <class S, bool kPeriodic>
struct BatchPlan {
  using Storage = S;
  using st = typename S::stored_t;
  static constexpr bool kWrap = kPeriodic;
  using Args = BatchArgs;
  __host__ __device__ static const st *stage_ptr(const Args &a) { return static_cast<const st *>(a.in); }
  __host__ __device__ static st *out_ptr(const Args &a) { return static_cast<st *>(a.out); }
  template <class C>
  __device__ static float compute(const C &ctx, const Args &a, const int x, const int y, const int z) {
    return Diffusion7Body<kPeriodic, false, false, false, kFma16<S>>(ctx, a.k, x, y, z, a.dim[0], a.dim[1], a.dim[2]);
  }
  // compute() where the stride skeleton supplies clamped-coordinate
  // neighbours along the flagged axes.
  static constexpr bool kExactNeighbours = true;
  template <bool kX, bool kY, bool kZ, class C>
  __device__ static float compute_exact(const C &ctx, const Args &a, const int x, const int y, const int z) {
    return Diffusion7Body<kPeriodic, kX, kY, kZ, kFma16<S>>(ctx, a.k, x, y, z, a.dim[0], a.dim[1], a.dim[2]);
  }
};

struct Item {
  int grid, y0, zb, ze;
};

// A work item and its grids' arguments are the same for the whole
// workgroup, but the compiler loads tables with vector loads; broadcast from
// the first lane, the values live in scalar registers instead of holding
// some twenty vector registers for the whole z-march.
__device__ __forceinline__ int Uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float Uni(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
template <class T>
__device__ __forceinline__ T *Uni(T *p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  return reinterpret_cast<T *>(((uint64_t)(uint32_t)Uni((int)(a >> 32)) << 32) | (uint32_t)Uni((int)a));
}
__device__ __forceinline__ Item Uniform(const Item &it) { return Item{Uni(it.grid), Uni(it.y0), Uni(it.zb), Uni(it.ze)}; }
__device__ __forceinline__ BatchArgs Uniform(const BatchArgs &a) {
  BatchArgs u;
  u.in = Uni(a.in);
  u.out = Uni(a.out);
  for (int d = 0; d < 3; ++d) u.dim[d] = Uni(a.dim[d]);
  u.k = Coef7{Uni(a.k.ce), Uni(a.k.cw), Uni(a.k.cn), Uni(a.k.cs), Uni(a.k.ct), Uni(a.k.cb), Uni(a.k.cc)};
  return u;
}

template <class P, int NP, int BY>
__global__ void __launch_bounds__(64 * BY)
BatchStep(const typename P::Args *args, const Item *items, int nitems) {
  __shared__ psk::StrideLds<NP, BY> L;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const Item it = Uniform(items[w]);
    const typename P::Args a = Uniform(args[it.grid]);
    psk::StrideStep<P, NP, BY>(a, a.dim[0], a.dim[1], a.dim[2], it.y0, it.zb, it.ze, L);
    __syncthreads();  // the next item reuses the LDS rows
  }
}

// kMid: the input table's `out` receives the intermediate step (the last
// fused pass of a run).
template <class P, int NP, int BY, bool kMid>
__global__ void __launch_bounds__(64 * BY)
BatchTB2(const typename P::Args *in_args, const typename P::Args *out_args, const Item *items, int nitems) {
  __shared__ psk::StrideTB2Lds<NP, BY> L;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const Item it = Uniform(items[w]);
    const typename P::Args a1 = Uniform(in_args[it.grid]);
    const typename P::Args a2 = Uniform(out_args[it.grid]);
    psk::StrideTB2<P, NP, BY, kMid>(a1, a2, a1.dim[0], a1.dim[1], a1.dim[2], it.y0, it.zb, it.ze, L);
    __syncthreads();
  }
}

// Classes of the mixed mode: the stride widths (ClassNp below), and two classes for
// grids of whole 16-byte rows longer than 128 points, whose fused steps run on the vectorised star
// tile (psk::StarTB2Tile, 8 rows per workgroup, one or two waves per row of
// up to 256 or 512 points) and whose unfused steps on the 4- or 8-point
// stride kernels.
constexpr int kStar1 = 101, kStar2 = 102;
constexpr int kMaxClasses = 8;
constexpr int kStarBY = 8;
inline bool IsStar(int cls) { return cls == kStar1 || cls == kStar2; }
inline int StarNwx(int cls) { return cls == kStar1 ? 1 : 2; }
// Stride width of a class's one-step kernel.
inline int StepNp(int cls) { return IsStar(cls) ? (cls == kStar1 ? 4 : 8) : cls; }

template <class P, int NWX, bool kMid>
__global__ void __launch_bounds__(64 * NWX * kStarBY, 8)
BatchStarTB2(const typename P::Args *in_args, const typename P::Args *out_args, const Item *items, int nitems) {
  __shared__ psk::StarTB2Lds<NWX, kStarBY> L;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const Item it = Uniform(items[w]);
    const typename P::Args a1 = Uniform(in_args[it.grid]);
    const typename P::Args a2 = Uniform(out_args[it.grid]);
    psk::StarTB2Tile<P, NWX, kStarBY, true, kMid>(a1, a2, a1.dim[0], a1.dim[1], a1.dim[2], it.y0, it.zb, it.ze,
                                                  L);
    __syncthreads();
  }
}

// Packed mode (FP32 and BF16 storage): every grid's fused pairs on the packed-row tile
// (psk::PackTB2Tile: rows of any length, several short rows per wave), all
// grids in one work list and one launch; rows of up to NXMAX points.
//
// The workgroups take items from the list through a queue: the first thread
// of a workgroup that is free claims the next item with an atomic add on
// queue[0], so a workgroup that drew short items takes more of them (the
// list is ordered longest first).  The last workgroup to find the queue
// empty (queue[1] counts them) resets both words for the next launch, so a
// captured graph replays without a memset.
template <class P, int NW, int NXMAX, bool kMid>
__global__ void __launch_bounds__(64 * NW, 8)
BatchPackTB2(const typename P::Args *in_args, const typename P::Args *out_args, const Item *items, int nitems,
             int *queue) {
  __shared__ psk::PackTB2Lds<NW, NXMAX> L;
  __shared__ int claim;
  for (;;) {
    if (threadIdx.x == 0) claim = atomicAdd(&queue[0], 1);
    __syncthreads();
    const int w = Uni(claim);
    if (w >= nitems) break;
    const Item it = Uniform(items[w]);
    const typename P::Args a1 = Uniform(in_args[it.grid]);
    const typename P::Args a2 = Uniform(out_args[it.grid]);
    psk::PackTB2Tile<P, NW, NXMAX, kMid>(a1, a2, a1.dim[0], a1.dim[1], a1.dim[2], it.y0, it.zb, it.ze, L);
    __syncthreads();  // LDS and `claim` are reused by the next item
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&queue[0], 0);
      atomicExch(&queue[1], 0);
    }
  }
}

// Longest row the packed kernels are built for, given the batch's longest.
inline int PackNxMax(int max_nx) { return max_nx <= 256 ? 256 : 512; }

// A launch configuration of the batch.
struct Config {
  int np;      // points per lane (row width 64*np); with `mixed`, of one class
  int by;      // rows per workgroup
  int zchunk;  // planes per work item
  int degree;  // 1: one step per launch; 2: two fused steps
  // 1: every grid runs in its class (the fewest points per lane that cover
  // its rows, or a star class); the grids of each class run on their own
  // stream (forked from and joined to the caller's), so the classes overlap.
  // 2: the same, with 16 rows per workgroup for the 1- and 2-point classes
  // (their kernels keep full occupancy at 16 rows, which halves the share
  // of halo rows).
  // 3: packed: the fused pairs of every grid on the packed-row tile in one
  // launch (`by` = waves per workgroup; rows of 4 to 512 points), single
  // steps as in mode 1.
  int mixed = 0;
  bool operator<(const Config &o) const {
    return std::tie(np, by, zchunk, degree, mixed) < std::tie(o.np, o.by, o.zchunk, o.degree, o.mixed);
  }
};

std::string ConfigName(const Config &c) {
  if (c.mixed == 3)
    return "packed_w" + std::to_string(c.by) + "_z" + std::to_string(c.zchunk) + "_t" + std::to_string(c.degree);
  return (c.mixed ? std::string(c.mixed == 2 ? "mixedn16" : "mixed") : "np" + std::to_string(c.np)) + "_by" +
         std::to_string(c.by) + "_z" +
         std::to_string(c.zchunk) + "_t" + std::to_string(c.degree);
}

// Points per lane of a grid's width class: the fewest that cover its rows
// (1, 2, 3, 4, 6 or 8).
int ClassNp(int nx) { return psk::StrideNp(nx); }



using KernelFn = const void *;

template <class S, bool kPer>
KernelFn StepKernelT(int np, int by) {
  using P = BatchPlan<S, kPer>;
#define PSH_K(NP, BY) \
  if (np == NP && by == BY) return reinterpret_cast<KernelFn>(&BatchStep<P, NP, BY>);
  PSH_K(1, 4) PSH_K(1, 8) PSH_K(1, 16) PSH_K(2, 4) PSH_K(2, 8) PSH_K(2, 16) PSH_K(3, 8) PSH_K(3, 16) PSH_K(4, 4)
  PSH_K(4, 8) PSH_K(4, 16) PSH_K(6, 8) PSH_K(8, 4) PSH_K(8, 8) PSH_K(8, 16)
#undef PSH_K
  return nullptr;
}

template <class S, bool kPer, bool kMid>
KernelFn TB2KernelT(int np, int by) {
  using P = BatchPlan<S, kPer>;
#define PSH_K(NP, BY) \
  if (np == NP && by == BY) return reinterpret_cast<KernelFn>(&BatchTB2<P, NP, BY, kMid>);
  PSH_K(1, 8) PSH_K(2, 8) PSH_K(3, 8) PSH_K(4, 8) PSH_K(6, 8) PSH_K(8, 8) PSH_K(1, 16) PSH_K(2, 16) PSH_K(3, 16)
  PSH_K(4, 16)
#undef PSH_K
  return nullptr;
}

template <class S, bool kPer, bool kMid>
KernelFn StarTB2KernelT(int cls) {
  using P = BatchPlan<S, kPer>;
  return cls == kStar1 ? reinterpret_cast<KernelFn>(&BatchStarTB2<P, 1, kMid>)
                       : reinterpret_cast<KernelFn>(&BatchStarTB2<P, 2, kMid>);
}

template <class S, bool kPer>
KernelFn TB2KernelFor(int np, int by, bool mid) {
  if (IsStar(np)) return mid ? StarTB2KernelT<S, kPer, true>(np) : StarTB2KernelT<S, kPer, false>(np);
  return mid ? TB2KernelT<S, kPer, true>(np, by) : TB2KernelT<S, kPer, false>(np, by);
}

// Kernel for a storage kind (PSHipStorage), boundary form and class (a
// stride width, or a star class: one-step kernels of star classes are the
// stride ones of StepNp, fused ones the star tile whatever `by` is).
KernelFn StepKernel(int storage, bool per, int np, int by) {
  np = StepNp(np);
  if (storage == PS_HIP_STORAGE_BF16)
    return per ? StepKernelT<psk::BF16, true>(np, by) : StepKernelT<psk::BF16, false>(np, by);
  return per ? StepKernelT<psk::Native<float>, true>(np, by) : StepKernelT<psk::Native<float>, false>(np, by);
}

// mid: the kernel that also stores the intermediate step (last fused pass).
KernelFn TB2Kernel(int storage, bool per, int np, int by, bool mid = false) {
  if (storage == PS_HIP_STORAGE_BF16)
    return per ? TB2KernelFor<psk::BF16, true>(np, by, mid) : TB2KernelFor<psk::BF16, false>(np, by, mid);
  return per ? TB2KernelFor<psk::Native<float>, true>(np, by, mid)
             : TB2KernelFor<psk::Native<float>, false>(np, by, mid);
}

// Packed fused kernel (FP32 and BF16 storage): nw waves, rows up to nxmax
// points.
template <class S, bool kPer>
KernelFn PackKernelT(int nw, int nxmax, bool mid) {
  using P = BatchPlan<S, kPer>;
#define PSH_K(NW, NXMAX)                                                                          \
  if (nw == NW && nxmax == NXMAX)                                                                 \
    return mid ? reinterpret_cast<KernelFn>(&BatchPackTB2<P, NW, NXMAX, true>)                    \
               : reinterpret_cast<KernelFn>(&BatchPackTB2<P, NW, NXMAX, false>);
  PSH_K(8, 256) PSH_K(16, 256) PSH_K(16, 512)
#undef PSH_K
  return nullptr;
}
KernelFn PackKernel(int storage, bool per, int nw, int nxmax, bool mid) {
  if (storage == PS_HIP_STORAGE_NATIVE)
    return per ? PackKernelT<psk::Native<float>, true>(nw, nxmax, mid)
               : PackKernelT<psk::Native<float>, false>(nw, nxmax, mid);
  if (storage == PS_HIP_STORAGE_BF16)
    return per ? PackKernelT<psk::BF16, true>(nw, nxmax, mid) : PackKernelT<psk::BF16, false>(nw, nxmax, mid);
  return nullptr;
}

// Rows per workgroup of class `cls` in mixed configuration c.
int ClassBy(const Config &c, int cls) { return c.mixed == 2 && !IsStar(cls) && cls <= 2 ? 16 : c.by; }

// Launch block of a configuration's fused (degree 2) or one-step kernel.
dim3 PairBlock(int np, int by) { return IsStar(np) ? dim3(64 * StarNwx(np), kStarBY) : dim3(64, by); }
// Output rows of a fused tile.
int PairRows(int np, int by) { return (IsStar(np) ? kStarBY : by) - 2; }

// A device array of T, reallocated on growth.
template <class T>
struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  void Assign(const std::vector<T> &v, hipStream_t s) {
    if (v.size() > n) {
      if (p) CheckHip(hipFree(p), "hipFree");
      CheckHip(hipMalloc(&p, v.size() * sizeof(T)), "hipMalloc(batch table)");
      n = v.size();
    }
    if (!v.empty()) {
      // Synchronous: the host vector may go away as soon as this returns.
      CheckHip(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
      CheckHip(hipStreamSynchronize(s), "upload sync");
    }
  }
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
};

// One batch of grid pairs.
struct Batch {
  std::vector<__PSGrid *> a, b;
  std::vector<Coef7> coef;
  std::vector<void *> aux;   // per-grid third buffer (fused steps)
  std::vector<bool> owns_aux;  // false: borrowed from the caller
  int storage = PS_HIP_STORAGE_NATIVE;
  bool periodic = false;
  int max_nx = 0;
  size_t bytes = 0;
  // Per-config work lists and arg tables, and cached graphs per (config, steps).
  struct Plan {
    DevArray<Item> items;
    DevArray<int> queue;  // packed mode: next item, workgroups done (zero between launches)
    int nitems = 0;
    int blocks = 0;
  };
  std::map<Config, Plan> plans;
  // Single steps: t_ab (A -> B), t_ba (B -> A).  Fused passes: input tables
  // t_in[b] and output tables t_out[b] for the buffers b = A, B, X of
  // psk::FusedPathBuffer, and t_in_mid (input X, intermediate into B).
  DevArray<BatchArgs> t_ab, t_ba, t_in[3], t_out[3], t_in_mid;
  bool tables_ready = false;
  struct Graph {
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::pair<Config, int>, Graph> graphs;
  Config tuned{0, 0, 0, 0};
  // Mixed mode: width classes present (ascending), and per-class streams and
  // fork/join events (created on first use; default priority).
  std::vector<int> classes;
  std::vector<int> cls;  // class of each grid
  hipStream_t class_stream[kMaxClasses] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kMaxClasses] = {};

  ~Batch() {
    for (auto &g : graphs)
      if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    for (auto &st : class_stream)
      if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
      }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (auto &e : join_ev)
      if (e) (void)hipEventDestroy(e);
    for (size_t i = 0; i < aux.size(); ++i)
      if (aux[i] && owns_aux[i]) (void)hipFree(aux[i]);
  }

  int Count() const { return (int)a.size(); }

  void BuildTables(hipStream_t s) {
    if (tables_ready) return;
    using Args = BatchArgs;
    std::vector<Args> ab, ba, in[3], out[3], in_mid;
    for (int i = 0; i < Count(); ++i) {
      Args x;
      for (int d = 0; d < 3; ++d) x.dim[d] = a[i]->dim[d];
      x.k = coef[i];
      void *A = a[i]->p0, *B = b[i]->p0, *X = aux[i];
      void *buf[3] = {A, B, X};
      x.in = A; x.out = B; ab.push_back(x);  // map 1: A -> B
      x.in = B; x.out = A; ba.push_back(x);  // map 2: B -> A
      for (int j = 0; j < 3; ++j) {
        x.in = buf[j]; x.out = nullptr; in[j].push_back(x);
        x.in = nullptr; x.out = buf[j]; out[j].push_back(x);
      }
      x.in = X; x.out = B; in_mid.push_back(x);
    }
    t_ab.Assign(ab, s);
    t_ba.Assign(ba, s);
    for (int j = 0; j < 3; ++j) {
      t_in[j].Assign(in[j], s);
      t_out[j].Assign(out[j], s);
    }
    t_in_mid.Assign(in_mid, s);
    tables_ready = true;
  }

  // A degree-2 run also needs the one-step kernel of the same shape for its
  // unfused steps.
  bool Supports(const Config &c) const {
    if (c.mixed == 3) {
      const int nxmax = PackNxMax(max_nx);
      if (c.degree != 2 || c.zchunk <= 0 || PackKernel(storage, periodic, c.by, nxmax, false) == nullptr)
        return false;
      for (int i = 0; i < Count(); ++i)
        if (!psk::PackSupported(a[i]->dim[0], c.by, nxmax) ||
            !psk::PackPlaneFits(a[i]->dim[0], a[i]->dim[1], Rec(a[i])->storage == PS_HIP_STORAGE_NATIVE ? 4 : 2))
          return false;
      return Supports(SingleSteps(c));
    }
    if (c.mixed) {
      if (c.zchunk <= 0) return false;
      for (int np : classes)
        if (StepKernel(storage, false, np, ClassBy(c, np)) == nullptr ||
            (c.degree == 2 && TB2Kernel(storage, false, np, ClassBy(c, np)) == nullptr))
          return false;
      return true;
    }
    if (64 * c.np < max_nx || c.zchunk <= 0) return false;
    if (c.degree == 2 && TB2Kernel(storage, false, c.np, c.by) == nullptr) return false;
    return StepKernel(storage, false, c.np, c.by) != nullptr;
  }

  Plan &PlanFor(const Config &c, hipStream_t s) {
    auto it = plans.find(c);
    if (it != plans.end()) return it->second;
    Plan &p = plans[c];
    // Largest grids first; items of one grid adjacent.
    std::vector<int> order(Count());
    for (int i = 0; i < Count(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return a[x]->num_elms > a[y]->num_elms; });
    std::vector<Item> items;
    const int rows_out = c.degree == 2 ? PairRows(c.np, c.by) : c.by;
    if (c.mixed == 3) {
      // Packed tiles of R - 2 output rows (R rows per workgroup), z chunks.
      for (int g : order) {
        const int ny = a[g]->dim[1], nz = a[g]->dim[2];
        const int rows_out = psk::PackRows(a[g]->dim[0], c.by) - 2;
        const int zc = c.zchunk > nz ? nz : c.zchunk;
        for (int ty = 0; ty * rows_out < ny; ++ty)
          for (int zb = 0; zb < nz; zb += zc) items.push_back(Item{g, ty * rows_out - 1, zb, std::min(zb + zc, nz)});
      }
      // Longest items (most planes) first, in list order among equals, so
      // the queue ends on short items.
      std::stable_sort(items.begin(), items.end(),
                       [](const Item &x, const Item &y) { return x.ze - x.zb > y.ze - y.zb; });
      p.queue.Assign(std::vector<int>{0, 0}, s);
    }
    for (int g : (c.mixed == 3 ? std::vector<int>() : order)) {
      if (c.mixed && cls[g] != c.np) continue;  // another class
      const int ny = a[g]->dim[1], nz = a[g]->dim[2];
      const int zc = c.zchunk <= 0 || c.zchunk > nz ? nz : c.zchunk;
      for (int ty = 0; ty * rows_out < ny; ++ty)
        for (int zb = 0; zb < nz; zb += zc) {
          Item it;
          it.grid = g;
          it.y0 = c.degree == 2 ? ty * rows_out - 1 : ty * rows_out;
          it.zb = zb;
          it.ze = std::min(zb + zc, nz);
          items.push_back(it);
        }
    }
    p.items.Assign(items, s);
    p.nitems = (int)items.size();
    // Persistent grid: as many workgroups as fit on the device at once.
    KernelFn k = c.mixed == 3 ? PackKernel(storage, periodic, c.by, PackNxMax(max_nx), false)
                 : c.degree == 2 ? TB2Kernel(storage, false, c.np, c.by)
                                 : StepKernel(storage, false, c.np, c.by);
    int per_cu = 1, cus = 1;
    const dim3 blk = c.mixed == 3 ? dim3(64 * c.by) : c.degree == 2 ? PairBlock(c.np, c.by) : dim3(64, c.by);
    CheckHip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, (int)(blk.x * blk.y), 0), "occupancy");
    hipDeviceProp_t prop;
    CheckHip(hipGetDeviceProperties(&prop, Runtime::Get().device()), "props");
    cus = prop.multiProcessorCount;
    p.blocks = std::max(1, std::min(p.nitems, std::max(1, per_cu) * cus));
    return p;
  }

  // Builds the tables, every work list configuration c uses and, in the
  // mixed mode, the class streams and events (uploads synchronise, so this
  // must happen before a graph capture).
  void Prepare(const Config &c, hipStream_t s) {
    BuildTables(s);
    if (c.mixed == 3) {
      PlanFor(c, s);
      Prepare(SingleSteps(c), s);
      return;
    }
    if (!c.mixed) {
      PlanFor(Config{c.np, c.by, c.zchunk, 1}, s);
      PlanFor(c, s);
      return;
    }
    for (int np : classes) {
      PlanFor(Config{np, ClassBy(c, np), c.zchunk, 1, 1}, s);
      PlanFor(Config{np, ClassBy(c, np), c.zchunk, c.degree, 1}, s);
    }
    // One stream and join event per class present (every stream a process
    // keeps takes a share of the device's few hardware queues).
    if (!fork_ev) {
      CheckHip(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "hipEventCreate");
      for (size_t i = 0; i < classes.size(); ++i) {
        CheckHip(hipStreamCreateWithFlags(&class_stream[i], hipStreamNonBlocking), "hipStreamCreate");
        CheckHip(hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming), "hipEventCreate");
      }
    }
  }

  // Queues `steps` steps of every grid with configuration c.
  hipError_t Enqueue(const Config &c, int steps, hipStream_t s) {
    Prepare(c, s);
    if (c.mixed == 3) return EnqueuePacked(c, steps, s);
    if (!c.mixed) return EnqueueOne(c, steps, s);
    // One stream per width class: fork from s, run each class's steps, join.
    hipError_t e = hipEventRecord(fork_ev, s);
    for (size_t i = 0; i < classes.size() && e == hipSuccess; ++i) {
      hipStream_t cs = class_stream[i];
      e = hipStreamWaitEvent(cs, fork_ev, 0);
      if (e == hipSuccess) e = EnqueueOne(Config{classes[i], ClassBy(c, classes[i]), c.zchunk, c.degree, 1}, steps, cs);
      if (e == hipSuccess) e = hipEventRecord(join_ev[i], cs);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, join_ev[i], 0);
    }
    return e;
  }

  // The configuration that runs the single steps of packed configuration c:
  // the width classes on their own streams.
  static Config SingleSteps(const Config &c) { return Config{0, 8, c.zchunk, 1, 1}; }

  // Packed mode: every pair fused along psk::FusedPathBuffer's path in one
  // launch over all grids; an even count's last pass also stores u_(steps-1)
  // into B, an odd count ends with one single step A -> B; runs of fewer than
  // two pairs are single steps.
  hipError_t EnqueuePacked(const Config &c, int steps, hipStream_t s) {
    const int pairs = steps / 2;
    if (pairs < 2) return Enqueue(SingleSteps(c), steps, s);
    Plan &p = PlanFor(c, s);
    const bool odd = steps & 1;
    const int nxmax = PackNxMax(max_nx);
    hipError_t e = hipSuccess;
    for (int j = 0; j < pairs && e == hipSuccess; ++j) {
      const int src = psk::FusedPathBuffer(j, pairs), dst = psk::FusedPathBuffer(j + 1, pairs);
      const bool mid = j == pairs - 1 && !odd;
      KernelFn k = PackKernel(storage, periodic, c.by, nxmax, mid);
      const void *ti = mid ? t_in_mid.p : t_in[src].p, *to = t_out[dst].p;
      const Item *iv = p.items.p;
      int n = p.nitems;
      int *q = p.queue.p;
      void *args[] = {&ti, &to, &iv, &n, &q};
      e = hipLaunchKernel(k, dim3(p.blocks), dim3(64 * c.by), args, 0, s);
    }
    if (odd && e == hipSuccess) e = Enqueue(SingleSteps(c), 1, s);
    return e;
  }

  // Queues `steps` steps of the grids configuration c covers (all grids, or
  // with c.mixed the width class c.np) on stream s.
  hipError_t EnqueueOne(const Config &c, int steps, hipStream_t s) {
    // Unfused steps always run on the one-step work list (row tiles start at
    // their first row); fused pairs on the two-step list (tiles start one row
    // above their first output row).
    const Config c1{c.np, c.by, c.zchunk, 1, c.mixed};
    Plan &p1 = PlanFor(c1, s);
    Plan &p2 = PlanFor(c, s);
    if (p1.nitems == 0 || steps <= 0) return hipSuccess;
    const dim3 block(64, c.by), pair_block = PairBlock(c.np, c.by);
    auto step = [&](const BatchArgs *t) {
      KernelFn k = StepKernel(storage, periodic, c.np, c.by);
      const void *tv = t;
      const Item *iv = p1.items.p;
      int n = p1.nitems;
      void *args[] = {&tv, &iv, &n};
      return hipLaunchKernel(k, dim3(p1.blocks), block, args, 0, s);
    };
    auto pair = [&](const BatchArgs *tin, const BatchArgs *tout, bool mid) {
      KernelFn k = TB2Kernel(storage, periodic, c.np, c.by, mid);
      const void *ti = tin, *to = tout;
      const Item *iv = p2.items.p;
      int n = p2.nitems;
      void *args[] = {&ti, &to, &iv, &n};
      return hipLaunchKernel(k, dim3(p2.blocks), pair_block, args, 0, s);
    };
    hipError_t e = hipSuccess;
    const int pairs = steps / 2;
    if (c.degree == 2 && pairs >= 2) {
      // Every pair fused along psk::FusedPathBuffer's path from A back to A.
      // An even count: the last pass also stores u_(steps-1) into B.  An odd
      // count: the pairs leave u_(steps-1) in A, and map 1 (A -> B) follows.
      const bool odd = steps & 1;
      for (int j = 0; j < pairs && e == hipSuccess; ++j) {
        const int src = psk::FusedPathBuffer(j, pairs), dst = psk::FusedPathBuffer(j + 1, pairs);
        const bool mid = j == pairs - 1 && !odd;
        e = pair(mid ? t_in_mid.p : t_in[src].p, t_out[dst].p, mid);
      }
      if (odd && e == hipSuccess) e = step(t_ab.p);
      return e;
    }
    for (int i = 0; i < steps && e == hipSuccess; ++i) e = step((i & 1) ? t_ba.p : t_ab.p);
    return e;
  }

  hipError_t Run(const Config &c, int steps, bool use_graph, hipStream_t s) {
    if (!use_graph) return Enqueue(c, steps, s);
    auto key = std::make_pair(c, steps);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      // Everything that allocates, uploads or creates streams and events
      // happens before the capture.
      Prepare(c, s);
      CheckHip(hipStreamSynchronize(s), "sync before capture");
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return e;
      hipError_t le = Enqueue(c, steps, s);
      e = hipStreamEndCapture(s, &g);
      if (le != hipSuccess) return le;
      if (e != hipSuccess) return e;
      hipGraphExec_t exec = nullptr;
      e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return e;
      it = graphs.emplace(key, Graph{exec}).first;
    }
    return hipGraphLaunch(it->second.exec, s);
  }
};

std::vector<Config> Candidates(const Batch &b) {
  std::vector<Config> out;
  if (b.classes.size() > 1 || (b.classes.size() == 1 && IsStar(b.classes[0]))) {
    for (int deg : {1, 2})
      for (int by : {8, 16})
        for (int zc : {16, 32, 64}) {
          Config c{0, by, zc, deg, 1};
          if (b.Supports(c)) out.push_back(c);
        }
    for (int zc : {32, 64}) {
      Config c{0, 8, zc, 2, 2};
      if (b.Supports(c)) out.push_back(c);
    }
  }
  // Packed tiles.
  for (int nw : {8, 16})
    for (int zc : {12, 16, 24}) {
      Config c{0, nw, zc, 2, 3};
      if (b.Supports(c)) out.push_back(c);
    }
  // One width for all grids: the narrowest that fits the widest row only
  // (wider rows idle lanes).
  const int np = ClassNp(b.max_nx);
  for (int deg : {1, 2})
    for (int by : {4, 8, 16})
      for (int zc : {16, 32, 64}) {
        Config c{np, by, zc, deg};
        if (b.Supports(c)) out.push_back(c);
      }
  return out;
}

// Before tuning: the packed mode where it serves the batch (rows of 4 to
// 512 points), else one width for all grids.
Config DefaultConfig(const Batch &b) {
  for (int nw : {8, 16}) {
    const Config c{0, nw, 16, 2, 3};
    if (b.Supports(c)) return c;
  }
  return Config{ClassNp(b.max_nx), 8, 32, 2};
}

}  // namespace
}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

// Like PSHipBatchCreate; aux (may be NULL) supplies a third buffer per grid
// (same byte size as the grid, owned by the caller) where aux[i] != NULL.
int PSHipBatchCreateEx(void **out, __PSGrid *const *a, __PSGrid *const *b, void *const *aux, int n,
                       const float *coef, int boundary) {
  if (!out || !a || !b || !coef || n <= 0) return PS_HIP_ERR_INVALID;
  if (boundary != PS_HIP_CLAMPED && boundary != PS_HIP_PERIODIC) return PS_HIP_ERR_INVALID;
  *out = nullptr;
  Batch *bt = new Batch();
  bt->periodic = boundary == PS_HIP_PERIODIC;
  bt->storage = a[0] ? Rec(a[0])->storage : PS_HIP_STORAGE_NATIVE;
  if (bt->storage == PS_HIP_STORAGE_FP16) {
    delete bt;
    return PS_HIP_ERR_UNSUPPORTED;
  }
  for (int i = 0; i < n; ++i) {
    const __PSGrid *x = a[i], *y = b[i];
    if (!x || !y || x->num_dims != 3 || y->num_dims != 3 || x->elm_size != 4 || y->elm_size != 4 ||
        Rec(x)->storage != bt->storage || Rec(y)->storage != bt->storage || x->dim[0] != y->dim[0] ||
        x->dim[1] != y->dim[1] || x->dim[2] != y->dim[2] || x->dim[0] > 512) {
      delete bt;
      return PS_HIP_ERR_INVALID;
    }
    bt->a.push_back(a[i]);
    bt->b.push_back(b[i]);
    bt->coef.push_back(Coef7{coef[7 * i], coef[7 * i + 1], coef[7 * i + 2], coef[7 * i + 3], coef[7 * i + 4],
                             coef[7 * i + 5], coef[7 * i + 6]});
    bt->max_nx = std::max(bt->max_nx, x->dim[0]);
    void *x_aux = aux ? aux[i] : nullptr;
    bt->owns_aux.push_back(x_aux == nullptr);
    if (!x_aux && hipMalloc(&x_aux, Rec(x)->bytes) != hipSuccess) {
      (void)hipGetLastError();
      bt->owns_aux.pop_back();
      delete bt;
      return PS_HIP_ERR_NOMEM;
    }
    bt->aux.push_back(x_aux);
    bt->bytes += Rec(x)->bytes;
  }
  for (int i = 0; i < bt->Count(); ++i) {
    // Star classes need whole 16-byte rows in 16-byte aligned buffers, and
    // pay only for rows of more than 128 points: a star wave spans 256
    // points, so narrower rows would leave most of its lanes idle.
    const int nx = bt->a[i]->dim[0];
    auto aligned = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    const bool star = nx % 4 == 0 && nx > 128 && aligned(bt->a[i]->p0) && aligned(bt->b[i]->p0) &&
                      aligned(bt->aux[i]);
    const int c = star ? (nx <= 256 ? kStar1 : kStar2) : ClassNp(nx);
    bt->cls.push_back(c);
    if (std::find(bt->classes.begin(), bt->classes.end(), c) == bt->classes.end()) bt->classes.push_back(c);
  }
  std::sort(bt->classes.begin(), bt->classes.end());
  bt->tuned = DefaultConfig(*bt);
  *out = bt;
  return PS_HIP_OK;
}

int PSHipBatchCreate(void **out, __PSGrid *const *a, __PSGrid *const *b, int n, const float *coef,
                     int boundary) {
  return PSHipBatchCreateEx(out, a, b, nullptr, n, coef, boundary);
}

// Queues `steps` steps on a caller stream (a hipStream_t), without a graph.
int PSHipBatchRunOn(void *h, int steps, void *stream) {
  if (!h || steps < 0) return PS_HIP_ERR_INVALID;
  Batch *b = static_cast<Batch *>(h);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : Runtime::Get().stream();
  return b->Enqueue(b->tuned, steps, s) == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

int PSHipBatchDestroy(void *h) {
  if (!h) return PS_HIP_ERR_INVALID;
  (void)hipStreamSynchronize(Runtime::Get().stream());
  delete static_cast<Batch *>(h);
  return PS_HIP_OK;
}

// Runs `steps` steps of every grid of the batch (map A->B, B->A, ...; after an
// even count A holds the field).  use_graph: capture once, replay afterwards.
int PSHipBatchRun(void *h, int steps, int use_graph) {
  if (!h || steps < 0) return PS_HIP_ERR_INVALID;
  Batch *b = static_cast<Batch *>(h);
  hipError_t e = b->Run(b->tuned, steps, use_graph != 0, Runtime::Get().stream());
  return e == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

// Sets the launch configuration explicitly (np, by, zchunk, degree); np = 0
// selects the mixed mode (each grid in its class, classes on their own
// streams), np = -1 the mixed mode with 16-row tiles for the 1- and 2-point
// classes, np = -2 the packed mode with `by` waves per workgroup (fused
// pairs only, degree 2).
int PSHipBatchConfigure(void *h, int np, int by, int zchunk, int degree) {
  if (!h) return PS_HIP_ERR_INVALID;
  Batch *b = static_cast<Batch *>(h);
  Config c{np, by, zchunk, degree, np == 0 ? 1 : np == -1 ? 2 : np == -2 ? 3 : 0};
  if (np < 0) c.np = 0;
  if ((degree != 1 && degree != 2) || !b->Supports(c)) return PS_HIP_ERR_UNSUPPORTED;
  b->tuned = c;
  return PS_HIP_OK;
}

int PSHipBatchConfigName(void *h, char *buf, int len) {
  if (!h || !buf || len <= 0) return PS_HIP_ERR_INVALID;
  snprintf(buf, len, "%s", ConfigName(static_cast<Batch *>(h)->tuned).c_str());
  return PS_HIP_OK;
}

// Times every candidate configuration for `steps` steps (graph launches) on
// the batch itself and keeps the fastest; the grids' contents are saved
// before and restored after.  best_ms receives the winner's time.
int PSHipBatchTune(void *h, int steps, float *best_ms) {
  if (!h || steps <= 0) return PS_HIP_ERR_INVALID;
  Batch *b = static_cast<Batch *>(h);
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  std::vector<void *> save_a(b->Count()), save_b(b->Count());
  for (int i = 0; i < b->Count(); ++i) {
    size_t by = Rec(b->a[i])->bytes;
    if (hipMalloc(&save_a[i], by) != hipSuccess || hipMalloc(&save_b[i], by) != hipSuccess) {
      (void)hipGetLastError();
      for (void *p : save_a) if (p) (void)hipFree(p);
      for (void *p : save_b) if (p) (void)hipFree(p);
      return PS_HIP_ERR_NOMEM;
    }
    CheckHip(hipMemcpyAsync(save_a[i], b->a[i]->p0, by, hipMemcpyDeviceToDevice, s), "save");
    CheckHip(hipMemcpyAsync(save_b[i], b->b[i]->p0, by, hipMemcpyDeviceToDevice, s), "save");
  }
  hipEvent_t e0, e1;
  CheckHip(hipEventCreate(&e0), "event");
  CheckHip(hipEventCreate(&e1), "event");
  float best = 1e30f;
  Config best_c = b->tuned;
  for (const Config &c : Candidates(*b)) {
    if (b->Run(c, steps, true, s) != hipSuccess) continue;  // warm: capture + first replay
    CheckHip(hipEventRecord(e0, s), "record");
    if (b->Run(c, steps, true, s) != hipSuccess) continue;
    CheckHip(hipEventRecord(e1, s), "record");
    CheckHip(hipEventSynchronize(e1), "sync");
    float ms = 0.f;
    CheckHip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (rt.trace()) fprintf(stderr, "physis-hip: batch tune %s: %.3f ms\n", ConfigName(c).c_str(), ms);
    if (ms < best) {
      best = ms;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int i = 0; i < b->Count(); ++i) {
    size_t by = Rec(b->a[i])->bytes;
    CheckHip(hipMemcpyAsync(b->a[i]->p0, save_a[i], by, hipMemcpyDeviceToDevice, s), "restore");
    CheckHip(hipMemcpyAsync(b->b[i]->p0, save_b[i], by, hipMemcpyDeviceToDevice, s), "restore");
  }
  CheckHip(hipStreamSynchronize(s), "sync");
  for (int i = 0; i < b->Count(); ++i) {
    (void)hipFree(save_a[i]);
    (void)hipFree(save_b[i]);
  }
  b->tuned = best_c;
  if (best_ms) *best_ms = best;
  return PS_HIP_OK;
}

}  // extern "C"
