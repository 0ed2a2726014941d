// serve.hip -- a continuous-batching server for stencil solves.
//
// Requests (7-point diffusion or 19-point Himeno-form Jacobi, cubic grids,
// FP32 or BF16 storage) arrive over time.  The server keeps every admitted
// request's grids resident on the device and advances all of them in
// "ticks" of a fixed number of steps:
//   - arrived requests wait in FIFO order and are admitted as soon as their
//     grids fit the HBM budget;
//   - all active diffusion requests of one storage kind advance together in
//     one batched launch per step (batched.hip) on the main stream, while the
//     Jacobi requests sweep on a pool of side streams (request i on stream
//     i mod 4), so the two kinds, and Jacobi requests among themselves,
//     overlap.  Jacobi requests run with a launch shape (one sweep per
//     launch, or several per HBM pass, himeno19.hip) an in-process tuner
//     picks per (storage, size class) by timing the candidates on the first
//     admitted request of that class;
//   - each diffusion batch and each Jacobi request starts its next tick as
//     soon as its previous one has completed (polled with events), so small
//     requests are not held to the pace of large ones; at the end of a tick
//     the request's residual has been computed on the device (diffusion:
//     ||u_n - u_{n-1}|| / ||u_n||, FP64; Jacobi: rms of ss over the interior
//     from the fused FP64 epilogue) and finished requests (residual <= tol,
//     or max_steps) retire: an FP64 checksum of the final field is taken.
//     Their memory stays counted against the budget until it is freed, which
//     happens when admission needs it or at the end of the trace (freeing
//     device memory waits for the whole device).
// Every request runs the same number of steps and the same arithmetic alone
// or in any mix (batching does not change per-grid results, reductions have
// a fixed order), so results are bitwise independent of the mix.
// Mode 1 ("sequential") serves the same trace one request at a time through
// the plain single-grid paths (one kernel per sweep or per fused diffusion
// step pair), as a loop of PSStencilRun calls would: the baseline the
// batched mode is compared with.  Fused sweeps are bitwise equal to single
// sweeps, so both modes compute the same fields.
//
// No reference counterpart (the reference has no serving layer).
#include "runtime_hip.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <list>
#include <map>
#include <thread>
#include <vector>

extern "C" int PSHipBatchCreateEx(void **out, __PSGrid *const *a, __PSGrid *const *b, void *const *aux, int n,
                                  const float *coef, int boundary);
extern "C" int PSHipBatchRunOn(void *h, int steps, void *stream);
extern "C" int PSHipBatchDestroy(void *h);

namespace physis_hip {
namespace {

constexpr int kNormThreads = 256;
constexpr int kNormBlocks = 512;

// Partial sums of (a-b)^2 and a^2 per workgroup (fixed assignment).
template <class S>
__global__ void __launch_bounds__(kNormThreads)
DiffNormPass1(const typename S::stored_t *a, const typename S::stored_t *b, int64_t n, double *partials) {
  __shared__ double red[kNormThreads / 64];
  double d2 = 0.0, a2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kNormThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNormThreads) {
    const double x = S::load(a + i), y = S::load(b + i);
    d2 += (x - y) * (x - y);
    a2 += x * x;
  }
  d2 = psk::block_reduce<kNormThreads>(d2, psk::OpSum(), red, threadIdx.x);
  a2 = psk::block_reduce<kNormThreads>(a2, psk::OpSum(), red, threadIdx.x);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = d2;
    partials[2 * blockIdx.x + 1] = a2;
  }
}

__global__ void __launch_bounds__(kNormThreads) DiffNormPass2(const double *partials, int nb, double *out) {
  __shared__ double red[kNormThreads / 64];
  double d2 = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kNormThreads) {
    d2 += partials[2 * i];
    a2 += partials[2 * i + 1];
  }
  d2 = psk::block_reduce<kNormThreads>(d2, psk::OpSum(), red, threadIdx.x);
  a2 = psk::block_reduce<kNormThreads>(a2, psk::OpSum(), red, threadIdx.x);
  if (threadIdx.x == 0) {
    out[0] = d2;
    out[1] = a2;
  }
}

void DiffNorm(const __PSGrid *a, const __PSGrid *b, double *partials, double *out, hipStream_t s) {
  const int64_t n = a->num_elms;
  int nb = (int)std::min<int64_t>(kNormBlocks, (n + kNormThreads * 4 - 1) / (kNormThreads * 4));
  nb = std::max(nb, 1);
  if (Rec(a)->storage == PS_HIP_STORAGE_BF16)
    hipLaunchKernelGGL(DiffNormPass1<psk::BF16>, dim3(nb), dim3(kNormThreads), 0, s,
                       static_cast<const uint16_t *>(a->p0), static_cast<const uint16_t *>(b->p0), n, partials);
  else
    hipLaunchKernelGGL(DiffNormPass1<psk::Native<float>>, dim3(nb), dim3(kNormThreads), 0, s,
                       static_cast<const float *>(a->p0), static_cast<const float *>(b->p0), n, partials);
  hipLaunchKernelGGL(DiffNormPass2, dim3(1), dim3(kNormThreads), 0, s, partials, nb, out);
}

const float kDiffCoef[7] = {0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.4f};
constexpr float kOmega = 0.8f;
constexpr int kSideStreams = 4;
constexpr size_t kLargeJacobiBytes = size_t(256) << 20;
constexpr size_t kMaxDeadBatches = 64;

// A launch shape of the Jacobi sweeps: degree 0 is one sweep per launch on
// the default tile (HimenoEnqueue); degree >= 1 the fused kernel with that
// many sweeps per pass, tile rows, rows per wave and z chunk.
struct JacobiShape {
  int degree, by, rpw, zchunk;
};
// Candidates the tuner times (all compiled shapes of himeno19.hip).
const JacobiShape kJacobiCandidates[] = {{0, 0, 0, 0},    {1, 8, 1, 16},  {2, 8, 1, 8},
                                         {2, 8, 1, 16},   {2, 8, 1, 32},  {2, 16, 1, 32}, {3, 16, 1, 32}};
constexpr int kTuneSweeps = 8;

int EnqueueJacobi(const JacobiShape &c, __PSGrid *A, __PSGrid *B, void *aux, __PSGrid *const *coef, int sweeps,
                  double *gosa, double *partials, hipStream_t s) {
  if (c.degree == 0) return HimenoEnqueue(A, B, coef, kOmega, sweeps, gosa, partials, s);
  return HimenoEnqueueTB(A, B, aux, coef, kOmega, sweeps, c.degree, c.by, c.rpw, c.zchunk, gosa, partials, s);
}

double Seconds(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace physis_hip

using namespace physis_hip;

// PSHipRequest / PSHipResult are declared in physis_hip_ext.h.  Request
// kinds: 0 = 7-point diffusion (clamped, cc=0.4, 0.1 neighbours, field
// uniform [0,1) from seed); 1 = 19-point Jacobi of the Himeno form
// (a0..a2, c0..c2 uniform [0.9,1.0] from seed, a3=1/6, b0..b2=0, bnd=1,
// wrk1=0, p=(z/(n-1))^2, omega=0.8: a convergent relaxation).

namespace physis_hip {
namespace {

struct Live {
  int idx = -1;
  bool inflight = false;      // a tick has been enqueued and not yet seen complete
  bool started = false;       // Jacobi: first tick enqueued (it waits for `init`)
  bool retired = false;
  hipStream_t js = nullptr;   // Jacobi: side stream
  hipEvent_t ev = nullptr;    // Jacobi: end of the in-flight tick
  hipEvent_t init = nullptr;  // Jacobi: end of the initial fills on the main stream
  const PSHipRequest *rq = nullptr;
  __PSGrid *A = nullptr, *B = nullptr;  // diffusion u_n, u_{n-1} / Jacobi p0, p1
  void *aux = nullptr;                  // third buffer (fused steps / sweeps)
  JacobiShape shape{1, 8, 1, 16};       // Jacobi launch shape (batched mode)
  __PSGrid *coef[12] = {};
  double *dev = nullptr;                // per-request scratch (device)
  int dev_doubles = 0;
  size_t bytes = 0;
  int steps = 0;
  double residual = INFINITY;
  double admitted = 0.0;
};

size_t RequestBytes(const PSHipRequest &r) {
  const size_t pts = (size_t)r.n * r.n * r.n;
  const size_t es = r.storage == PS_HIP_STORAGE_NATIVE ? 4 : 2;
  return pts * es * (r.kind == 0 ? 3 : 15);
}

class Server {
 public:
  Server(const PSHipRequest *reqs, int n, int64_t budget, int tick, PSHipResult *out)
      : reqs_(reqs), n_(n), budget_(budget), tick_(tick), out_(out) {
    Runtime &rt = Runtime::Get();
    main_ = rt.stream();
    for (auto &st : side_) CheckHip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    CheckHip(hipHostMalloc(&host_, sizeof(double) * 4 * (size_t)std::max(n, 1), hipHostMallocDefault),
             "hipHostMalloc");
  }
  ~Server() {
    for (auto &b : batches_) (void)PSHipBatchDestroy(b.second.handle);
    for (void *h : dead_batches_) (void)PSHipBatchDestroy(h);
    for (auto &g : groups_)
      if (g.ev) (void)hipEventDestroy(g.ev);
    for (auto &st : side_) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
    (void)hipHostFree(host_);
  }

  int RunSequential();
  int RunContinuous();
  void Tune() { PretuneJacobi(); }

 private:
  int Admit(int idx, double now, hipStream_t s, bool batched);
  int Allocate(Live &lv, const PSHipRequest &r);
  void PretuneJacobi();
  void Retire(Live &lv, double now, hipStream_t s);
  void FreeRetired();
  void Finish(Live &lv);  // residual and step count after a completed tick
  void LaunchDiffusion(int storage, std::vector<Live *> &members);
  void AdvanceTick(std::vector<Live *> &act);
  void Release(Live &lv);
  JacobiShape TuneJacobi(Live &lv, hipStream_t s);

  struct BatchEntry {
    void *handle = nullptr;
    std::vector<int> ids;
  };

  const PSHipRequest *reqs_;
  int n_;
  int64_t budget_;
  int tick_;
  PSHipResult *out_;
  hipStream_t main_, side_[kSideStreams];
  double *host_ = nullptr;
  int64_t used_ = 0;
  std::vector<Live> live_;               // sequential mode
  std::list<Live> clive_;                // continuous mode (stable addresses)
  std::vector<Live> retired_;            // allocations awaiting release
  std::vector<void *> dead_batches_;     // replaced batch handles awaiting release
  struct Group {
    bool inflight = false;
    hipEvent_t ev = nullptr;
    std::vector<Live *> members;
  };
  Group groups_[2];                      // continuous-mode diffusion batches per storage kind
  std::map<int, BatchEntry> batches_;  // key: storage kind
  std::chrono::steady_clock::time_point t0_;
};

// Tuned Jacobi launch shapes per (storage, size class of 64 points): one
// table for the process, so later servers reuse it.
std::map<std::pair<int, int>, JacobiShape> &JacobiTuned() {
  static std::map<std::pair<int, int>, JacobiShape> t;
  return t;
}

std::pair<int, int> JacobiKey(const PSHipRequest &r) { return {r.storage, (r.n + 63) / 64}; }

// Times every candidate shape for kTuneSweeps sweeps on a request's
// (zero-filled) grids and returns the fastest.  Each candidate runs once
// untimed (code loading, caches, clocks), then every candidate is timed in
// two rounds and keeps its shorter time.  Called before a trace starts, with
// nothing else on the device, so that the timings are not disturbed by other
// requests' work and the scheduler loop never waits on them.
JacobiShape Server::TuneJacobi(Live &lv, hipStream_t s) {
  CheckHip(hipDeviceSynchronize(), "sync before tuning");
  constexpr int kN = sizeof(kJacobiCandidates) / sizeof(kJacobiCandidates[0]);
  hipEvent_t e0, e1;
  CheckHip(hipEventCreate(&e0), "event");
  CheckHip(hipEventCreate(&e1), "event");
  double *gosa = lv.dev, *partials = lv.dev + std::max(tick_, kTuneSweeps);
  bool ok[kN];
  float best_ms[kN];
  for (int i = 0; i < kN; ++i) {
    ok[i] = EnqueueJacobi(kJacobiCandidates[i], lv.A, lv.B, lv.aux, lv.coef, kTuneSweeps, gosa, partials, s) ==
            PS_HIP_OK;
    best_ms[i] = 1e30f;
  }
  CheckHip(hipStreamSynchronize(s), "tuning warm-up");
  for (int round = 0; round < 2; ++round) {
    for (int i = 0; i < kN; ++i) {
      if (!ok[i]) continue;
      CheckHip(hipEventRecord(e0, s), "record");
      if (EnqueueJacobi(kJacobiCandidates[i], lv.A, lv.B, lv.aux, lv.coef, kTuneSweeps, gosa, partials, s) !=
          PS_HIP_OK) {
        ok[i] = false;
        continue;
      }
      CheckHip(hipEventRecord(e1, s), "record");
      CheckHip(hipEventSynchronize(e1), "sync");
      float ms = 0.f;
      CheckHip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      best_ms[i] = std::min(best_ms[i], ms);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int best = 0;
  for (int i = 1; i < kN; ++i)
    if (ok[i] && best_ms[i] < best_ms[best]) best = i;
  return kJacobiCandidates[best];
}

// Tunes the Jacobi launch shape of every (storage, size class) the trace
// holds that the process has not tuned yet, on temporary grids of the
// class's first request, before the trace's clock starts.  A class whose
// grids cannot be allocated keeps the default shape.
void Server::PretuneJacobi() {
  for (int i = 0; i < n_; ++i) {
    const PSHipRequest &r = reqs_[i];
    if (r.kind != 1 || JacobiTuned().count(JacobiKey(r))) continue;
    if ((int64_t)RequestBytes(r) > budget_) continue;  // never admitted
    Live lv;
    lv.rq = &r;
    if (Allocate(lv, r) == PS_HIP_OK) JacobiTuned()[JacobiKey(r)] = TuneJacobi(lv, main_);
    CheckHip(hipStreamSynchronize(main_), "tuning sync");
    Release(lv);
  }
}

// Device memory of a request: its grids, the fused-step buffer and its
// scratch; on failure everything allocated so far is released.
int Server::Allocate(Live &lv, const PSHipRequest &r) {
  const int dim[3] = {r.n, r.n, r.n};
  auto grid = [&](__PSGrid **g) { return GridNewDevice(g, 4, 3, dim, r.storage); };
  int st = grid(&lv.A);
  if (st == PS_HIP_OK) st = grid(&lv.B);
  if (st == PS_HIP_OK && hipMalloc(&lv.aux, (size_t)r.n * r.n * r.n * (r.storage ? 2 : 4)) != hipSuccess) {
    (void)hipGetLastError();
    st = PS_HIP_ERR_NOMEM;
  }
  if (r.kind == 0) {
    lv.dev_doubles = 2 + 2 * kNormBlocks + (int)ReduceScratchDoubles();
  } else {
    for (int i = 0; i < 12 && st == PS_HIP_OK; ++i) st = grid(&lv.coef[i]);
    if (st == PS_HIP_OK) {
      int parts = HimenoPartials(lv.A);
      for (const JacobiShape &c : kJacobiCandidates)
        if (c.degree > 0) parts = std::max(parts, HimenoTBPartials(lv.A, c.degree, c.by, c.rpw, c.zchunk));
      lv.dev_doubles = std::max(tick_, kTuneSweeps) + parts + (int)ReduceScratchDoubles() + 1;
    }
  }
  if (st == PS_HIP_OK && hipMalloc(&lv.dev, sizeof(double) * lv.dev_doubles) != hipSuccess) {
    (void)hipGetLastError();
    st = PS_HIP_ERR_NOMEM;
  }
  if (st != PS_HIP_OK) Release(lv);
  return st;
}

int Server::Admit(int idx, double now, hipStream_t s, bool batched) {
  const PSHipRequest &r = reqs_[idx];
  Live lv;
  lv.idx = idx;
  lv.rq = &r;
  lv.admitted = now;
  const int st = Allocate(lv, r);
  if (st != PS_HIP_OK) return st;
  lv.bytes = RequestBytes(r);
  used_ += (int64_t)lv.bytes;
  if (r.kind == 1 && batched) {
    auto it = JacobiTuned().find(JacobiKey(r));
    lv.shape = it != JacobiTuned().end() ? it->second : kJacobiCandidates[0];
  }
  // Initial state.
  if (r.kind == 0) {
    FillUniformOn(lv.A, r.seed, 0.f, 1.f, s);
    FillPatternOn(lv.B, 0, 0.f, s);
  } else {
    // A convergent Himeno-form problem: a0..a2, c0..c2 uniform in [0.9, 1.0]
    // (seeded), a3 = 1/6, b0..b2 = 0, bnd = 1, wrk1 = 0, so a3 * sum(a, c)
    // <= 1 and the Jacobi iteration contracts.
    FillPatternOn(lv.A, 1, 0.f, s);
    FillPatternOn(lv.B, 1, 0.f, s);
    for (int i = 0; i < 12; ++i) {
      if (i == 3) FillPatternOn(lv.coef[i], 0, 1.0f / 6.0f, s);
      else if (i >= 4 && i <= 6) FillPatternOn(lv.coef[i], 0, 0.f, s);
      else if (i == 10) FillPatternOn(lv.coef[i], 0, 1.f, s);
      else if (i == 11) FillPatternOn(lv.coef[i], 0, 0.f, s);
      else FillUniformOn(lv.coef[i], r.seed + i, 0.9f, 1.0f, s);
    }
  }
  live_.push_back(lv);
  return PS_HIP_OK;
}

void Server::Release(Live &lv) {
  for (__PSGrid **g : {&lv.A, &lv.B}) {
    if (*g) PSHipGridFree(*g);
    *g = nullptr;
  }
  for (auto &c : lv.coef) {
    if (c) PSHipGridFree(c);
    c = nullptr;
  }
  if (lv.aux) (void)hipFree(lv.aux);
  if (lv.dev) (void)hipFree(lv.dev);
  lv.aux = nullptr;
  lv.dev = nullptr;
  for (hipEvent_t *e : {&lv.ev, &lv.init}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
}

void Server::Retire(Live &lv, double now, hipStream_t s) {
  // Checksum of the final field (A holds u_n for diffusion; after an even
  // number of sweeps p0 holds the Jacobi field), on the stream that computed
  // it.
  double *res = lv.dev;
  CheckHip((hipError_t)(ReduceAsync(lv.A->p0, lv.A->num_elms, 4, Rec(lv.A)->storage, PS_SUM, res, nullptr, s) ==
                                PS_HIP_OK
                            ? hipSuccess
                            : hipErrorUnknown),
           "checksum");
  double *slot = host_ + 2 * (size_t)n_;  // past the per-request residual slots
  CheckHip(hipMemcpyAsync(slot, res, sizeof(double), hipMemcpyDeviceToHost, s), "checksum copy");
  CheckHip(hipStreamSynchronize(s), "checksum sync");
  PSHipResult &o = out_[lv.idx];
  o.status = PS_HIP_OK;
  o.steps = lv.steps;
  o.residual = lv.residual;
  o.admitted = lv.admitted;
  o.finished = now;
  o.checksum = *slot;
  lv.retired = true;
  // Freed later (FreeRetired): its bytes stay in used_ until then.
  retired_.push_back(lv);
  lv.A = lv.B = nullptr;
  for (auto &c : lv.coef) c = nullptr;
  lv.aux = nullptr;
  lv.dev = nullptr;
  lv.ev = lv.init = nullptr;
}

void Server::FreeRetired() {
  if (retired_.empty() && dead_batches_.empty()) return;
  CheckHip(hipDeviceSynchronize(), "device sync");
  for (Live &lv : retired_) {
    used_ -= (int64_t)lv.bytes;
    Release(lv);
  }
  retired_.clear();
  for (void *h : dead_batches_) (void)PSHipBatchDestroy(h);
  dead_batches_.clear();
}

void Server::Finish(Live &lv) {
  lv.steps += tick_;
  const double *h = host_ + 2 * lv.idx;
  if (lv.rq->kind == 0) {
    lv.residual = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
  } else {
    const double interior = std::pow((double)(lv.rq->n - 2), 3.0);
    lv.residual = std::sqrt(h[0] / interior);
  }
}

// One tick of the request being served in sequential mode (mode 1), through
// the plain single-grid paths on the main stream, then a sync.
void Server::AdvanceTick(std::vector<Live *> &act) {
  for (Live *lv : act) {
    if (lv->rq->kind == 0) {
      __PSDomain dom = PSDomain3DNew(0, lv->rq->n, 0, lv->rq->n, 0, lv->rq->n);
      if (PSHipDiffusion7Run(lv->A, lv->B, &dom, kDiffCoef, PS_HIP_CLAMPED, tick_, 0) != PS_HIP_OK)
        Fatal("serve", "diffusion run failed");
      DiffNorm(lv->A, lv->B, lv->dev + 2, lv->dev, main_);
      CheckHip(hipMemcpyAsync(host_ + 2 * lv->idx, lv->dev, 2 * sizeof(double), hipMemcpyDeviceToHost, main_),
               "norm copy");
    } else {
      double *gosa = lv->dev;
      double *partials = lv->dev + std::max(tick_, kTuneSweeps);
      if (HimenoEnqueue(lv->A, lv->B, lv->coef, kOmega, tick_, gosa, partials, main_) != PS_HIP_OK)
        Fatal("serve", "jacobi run failed");
      CheckHip(hipMemcpyAsync(host_ + 2 * lv->idx, gosa + tick_ - 1, sizeof(double), hipMemcpyDeviceToHost, main_),
               "gosa copy");
    }
  }
  CheckHip(hipStreamSynchronize(main_), "tick sync");
  for (Live *lv : act) Finish(*lv);
}

// One tick of the diffusion batch of one storage kind on the main stream,
// with each member's residual inputs copied to its host slots; the batch is
// rebuilt when its membership changed (the old handle is released later).
void Server::LaunchDiffusion(int storage, std::vector<Live *> &v) {
  std::vector<int> ids;
  for (Live *lv : v) ids.push_back(lv->idx);
  BatchEntry &be = batches_[storage];
  if (be.ids != ids) {
    if (be.handle) dead_batches_.push_back(be.handle);
    be.handle = nullptr;
    std::vector<__PSGrid *> a, b;
    std::vector<void *> aux;
    std::vector<float> coef;
    for (Live *lv : v) {
      a.push_back(lv->A);
      b.push_back(lv->B);
      aux.push_back(lv->aux);
      coef.insert(coef.end(), kDiffCoef, kDiffCoef + 7);
    }
    if (PSHipBatchCreateEx(&be.handle, a.data(), b.data(), aux.data(), (int)v.size(), coef.data(),
                           PS_HIP_CLAMPED) != PS_HIP_OK)
      Fatal("serve", "batch creation failed");
    be.ids = ids;
  }
  if (PSHipBatchRunOn(be.handle, tick_, main_) != PS_HIP_OK) Fatal("serve", "batched run failed");
  for (Live *lv : v) DiffNorm(lv->A, lv->B, lv->dev + 2, lv->dev, main_);
  for (Live *lv : v)
    CheckHip(hipMemcpyAsync(host_ + 2 * lv->idx, lv->dev, 2 * sizeof(double), hipMemcpyDeviceToHost, main_),
             "norm copy");
}

// Continuous batching (mode 0): admission as in Run, then every diffusion
// batch and every Jacobi request starts its next tick as soon as its last
// one has completed.
int Server::RunContinuous() {
  PretuneJacobi();  // classes PSHipServeTune has not tuned
  t0_ = std::chrono::steady_clock::now();
  for (int i = 0; i < n_; ++i) out_[i] = PSHipResult{PS_HIP_ERR_INVALID, 0, 0.0, 0.0, 0.0, 0.0};
  for (auto &g : groups_) CheckHip(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming), "hipEventCreate");
  std::vector<int> order(n_);
  for (int i = 0; i < n_; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs_[a].arrival < reqs_[b].arrival; });
  size_t next = 0;
  std::deque<int> waiting;
  int done = 0;
  auto finished = [](const Live &lv) { return lv.residual <= lv.rq->tol || lv.steps >= lv.rq->max_steps; };
  while (done < n_) {
    const double now = Seconds(t0_);
    while (next < order.size() && reqs_[order[next]].arrival <= now) waiting.push_back(order[next++]);
    // Admission by HBM budget, FIFO; retired requests' memory is released
    // first when the budget needs it.
    while (!waiting.empty()) {
      const int idx = waiting.front();
      const int64_t need = (int64_t)RequestBytes(reqs_[idx]);
      if (need > budget_) {
        out_[idx].status = PS_HIP_ERR_NOMEM;  // can never fit
        waiting.pop_front();
        ++done;
        continue;
      }
      if (used_ + need > budget_) {
        if (!retired_.empty()) {
          FreeRetired();
          continue;
        }
        break;
      }
      const int st = Admit(idx, Seconds(t0_), main_, true);
      if (st == PS_HIP_ERR_NOMEM && (!clive_.empty() || !retired_.empty())) {
        FreeRetired();
        if (!clive_.empty()) break;  // wait for memory to free up
        continue;
      }
      waiting.pop_front();
      if (st != PS_HIP_OK) {
        out_[idx].status = st;
        ++done;
        continue;
      }
      clive_.push_back(live_.back());
      live_.pop_back();
      Live &lv = clive_.back();
      if (lv.rq->kind == 1) {
        lv.js = side_[lv.idx % kSideStreams];
        CheckHip(hipEventCreateWithFlags(&lv.ev, hipEventDisableTiming), "hipEventCreate");
        CheckHip(hipEventCreateWithFlags(&lv.init, hipEventDisableTiming), "hipEventCreate");
        CheckHip(hipEventRecord(lv.init, main_), "record");
      }
    }
    bool progressed = false;
    // Diffusion batches, one per storage kind.
    for (int sk = 0; sk < 2; ++sk) {
      Group &g = groups_[sk];
      if (g.inflight) {
        const hipError_t q = hipEventQuery(g.ev);
        if (q == hipErrorNotReady) continue;
        CheckHip(q, "event query");
        g.inflight = false;
        progressed = true;
        const double t = Seconds(t0_);
        for (Live *lv : g.members) {
          Finish(*lv);
          if (finished(*lv)) {
            Retire(*lv, t, main_);
            ++done;
          }
        }
      }
      std::vector<Live *> members;
      for (Live &lv : clive_)
        if (!lv.retired && lv.rq->kind == 0 && lv.rq->storage == sk) members.push_back(&lv);
      if (members.empty()) continue;
      LaunchDiffusion(sk, members);
      CheckHip(hipEventRecord(g.ev, main_), "record");
      g.members = members;
      g.inflight = true;
    }
    // Jacobi requests, each on its side stream.  Large ones (grids of
    // kLargeJacobiBytes or more) sweep one at a time in arrival order: each
    // one fills the device alone, and sharing it among them only delays the
    // earliest (the trace's worst latency) without finishing the set sooner.
    // Small ones and the diffusion batches overlap with them.
    const Live *large_head = nullptr;
    for (const Live &lv : clive_)
      if (!lv.retired && lv.rq->kind == 1 && lv.bytes >= kLargeJacobiBytes) {
        large_head = &lv;
        break;
      }
    for (Live &lv : clive_) {
      if (lv.retired || lv.rq->kind != 1) continue;
      if (lv.inflight) {
        const hipError_t q = hipEventQuery(lv.ev);
        if (q == hipErrorNotReady) continue;
        CheckHip(q, "event query");
        lv.inflight = false;
        progressed = true;
        Finish(lv);
        if (finished(lv)) {
          Retire(lv, Seconds(t0_), lv.js);
          ++done;
          continue;
        }
      }
      if (lv.bytes >= kLargeJacobiBytes && &lv != large_head) continue;
      if (!lv.started) {
        CheckHip(hipStreamWaitEvent(lv.js, lv.init, 0), "wait");
        lv.started = true;
      }
      double *gosa = lv.dev;
      double *partials = lv.dev + std::max(tick_, kTuneSweeps);
      if (EnqueueJacobi(lv.shape, lv.A, lv.B, lv.aux, lv.coef, tick_, gosa, partials, lv.js) != PS_HIP_OK)
        Fatal("serve", "jacobi run failed");
      CheckHip(hipMemcpyAsync(host_ + 2 * lv.idx, gosa + tick_ - 1, sizeof(double), hipMemcpyDeviceToHost, lv.js),
               "gosa copy");
      CheckHip(hipEventRecord(lv.ev, lv.js), "record");
      lv.inflight = true;
    }
    clive_.remove_if([](const Live &lv) { return lv.retired; });
    // Replaced batch handles hold device work lists; release them in bulk
    // before they accumulate on a long trace.
    if (dead_batches_.size() >= kMaxDeadBatches) FreeRetired();
    if (!progressed) {
      if (clive_.empty() && next >= order.size() && waiting.empty()) break;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  FreeRetired();
  return PS_HIP_OK;
}

// Sequential mode (mode 1): one request at a time, admitted in arrival
// order, one tick then a sync.
int Server::RunSequential() {
  t0_ = std::chrono::steady_clock::now();
  for (int i = 0; i < n_; ++i) out_[i] = PSHipResult{PS_HIP_ERR_INVALID, 0, 0.0, 0.0, 0.0, 0.0};
  std::vector<int> order(n_);
  for (int i = 0; i < n_; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs_[a].arrival < reqs_[b].arrival; });
  size_t next = 0;
  std::deque<int> waiting;
  int done = 0;
  live_.reserve(n_);
  while (done < n_) {
    const double now = Seconds(t0_);
    while (next < order.size() && reqs_[order[next]].arrival <= now) waiting.push_back(order[next++]);
    // Admission in arrival order, one request at a time.
    while (!waiting.empty()) {
      const int idx = waiting.front();
      if (!live_.empty()) break;
      const int64_t need = (int64_t)RequestBytes(reqs_[idx]);
      if (need > budget_) {
        out_[idx].status = PS_HIP_ERR_NOMEM;  // can never fit
        waiting.pop_front();
        ++done;
        continue;
      }
      if (used_ + need > budget_) break;
      const int st = Admit(idx, Seconds(t0_), main_, false);
      if (st == PS_HIP_ERR_NOMEM && !live_.empty()) break;  // wait for memory to free up
      waiting.pop_front();
      if (st != PS_HIP_OK) {
        out_[idx].status = st;
        ++done;
      }
    }
    if (live_.empty()) {
      if (next >= order.size() && waiting.empty()) break;
      std::this_thread::sleep_for(std::chrono::microseconds(200));
      continue;
    }
    std::vector<Live *> act;
    for (Live &lv : live_) act.push_back(&lv);
    AdvanceTick(act);
    // Retire finished requests.
    const double t = Seconds(t0_);
    std::vector<Live> keep;
    for (Live &lv : live_) {
      if (lv.residual <= lv.rq->tol || lv.steps >= lv.rq->max_steps) {
        Retire(lv, t, main_);
        FreeRetired();
        ++done;
      } else {
        keep.push_back(lv);
      }
    }
    live_.swap(keep);
  }
  return PS_HIP_OK;
}

}  // namespace
}  // namespace physis_hip

extern "C" {

// Serves `n` requests in real time (arrival offsets from now).  mode 0:
// continuous batching; mode 1: one request at a time through the plain
// single-grid paths.  budget_bytes bounds the device memory of admitted
// requests; tick_steps (even) is the number of steps between residual checks.
int PSHipServeTrace(const PSHipRequest *reqs, int n, int64_t budget_bytes, int tick_steps, int mode,
                    PSHipResult *results) {
  if (!reqs || !results || n <= 0 || budget_bytes <= 0 || tick_steps <= 0 || (tick_steps & 1)) return PS_HIP_ERR_INVALID;
  for (int i = 0; i < n; ++i) {
    const PSHipRequest &r = reqs[i];
    if ((r.kind != 0 && r.kind != 1) || r.n < 3 || r.n > 512 ||
        (r.storage != PS_HIP_STORAGE_NATIVE && r.storage != PS_HIP_STORAGE_BF16) || r.max_steps <= 0)
      return PS_HIP_ERR_INVALID;
  }
  Runtime::Get().EnsureDevice();
  Server srv(reqs, n, budget_bytes, tick_steps, results);
  return mode == 1 ? srv.RunSequential() : srv.RunContinuous();
}

// Tunes the Jacobi launch shape of every (storage, size class) among the
// requests, as a continuous-batching trace would at its start, and keeps the
// results for the process (later traces reuse them).
int PSHipServeTune(const PSHipRequest *reqs, int n, int64_t budget_bytes, int tick_steps) {
  if (!reqs || n <= 0 || budget_bytes <= 0 || tick_steps <= 0 || (tick_steps & 1)) return PS_HIP_ERR_INVALID;
  for (int i = 0; i < n; ++i) {
    const PSHipRequest &r = reqs[i];
    if ((r.kind != 0 && r.kind != 1) || r.n < 3 || r.n > 512 ||
        (r.storage != PS_HIP_STORAGE_NATIVE && r.storage != PS_HIP_STORAGE_BF16) || r.max_steps <= 0)
      return PS_HIP_ERR_INVALID;
  }
  Runtime::Get().EnsureDevice();
  std::vector<PSHipResult> scratch_results((size_t)n);
  Server srv(reqs, n, budget_bytes, tick_steps, scratch_results.data());
  srv.Tune();
  return PS_HIP_OK;
}

}  // extern "C"
