// hostcopy.cpp -- PSGridCopyin/Copyout of large grids between pageable host
// memory and HBM.
//
// One hipMemcpy from or to pageable memory is staged by the HIP runtime
// through pinned memory with a single host thread: ~10 GB/s device to host,
// ~30 GB/s host to device on MI355X.  Here W host threads each own two pinned
// staging buffers and move chunks w, w + W, w + 2W, ...: the DMA of a
// thread's next chunk overlaps the memcpy of its current one, and the
// memcpys (and the first-touch page faults of a freshly allocated
// destination array) run on W cores while the link serves the other
// threads' chunks.  The DMAs share two streams, threads alternating (one
// stream reached ~28 GB/s device to host, two fill the link, and every
// stream a process keeps takes a share of the device's few hardware
// queues); each thread waits on its own events.  Every transfer is ordered
// after the work already queued on the caller's stream (an event the copy
// stream waits on), and the call returns when all bytes have arrived, as the
// blocking reference calls do (reference runtime/libphysis_rt_cuda.cc:116-139).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

#include "runtime_hip.h"

namespace physis_hip {
namespace {

constexpr size_t kChunk = 4u << 20;         // bytes per staged transfer
constexpr size_t kPipelineMin = 16u << 20;  // smaller copies: one hipMemcpy
constexpr int kMaxWorkers = 8;
constexpr int kDmaStreams = 2;

struct Worker {
  void *buf[2] = {};  // pinned, kChunk bytes each
  hipEvent_t ev[2] = {};
};

struct Pool {
  std::mutex mu;  // one pipelined copy at a time
  Worker w[kMaxWorkers];
  int n = 0;
  hipStream_t stream[kDmaStreams] = {};  // DMA streams (default priority)
  hipEvent_t start = nullptr;

  // Creates workers up to `want` (kept until ReleaseHostCopyPool) and
  // returns how many exist.  Pinned memory, streams or events that cannot be
  // made leave fewer workers (0: the copy falls back to one hipMemcpy)
  // instead of failing the copy.
  int Ensure(int want) {
    if (!start) {
      if (hipEventCreateWithFlags(&start, hipEventDisableTiming) != hipSuccess) {
        start = nullptr;
        return 0;
      }
      for (auto &st : stream) {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
          st = nullptr;
          Release();
          return 0;
        }
      }
    }
    for (; n < want; ++n) {
      Worker &k = w[n];
      bool made = true;
      for (int b = 0; b < 2 && made; ++b) {
        if (hipHostMalloc(&k.buf[b], kChunk, hipHostMallocDefault) != hipSuccess) {
          k.buf[b] = nullptr;
          made = false;
        } else if (hipEventCreateWithFlags(&k.ev[b], hipEventDisableTiming) != hipSuccess) {
          k.ev[b] = nullptr;
          made = false;
        }
      }
      if (!made) {
        FreeWorker(k);
        (void)hipGetLastError();  // clear the failed call's error
        break;
      }
    }
    return n;
  }

  static void FreeWorker(Worker &k) {
    for (int b = 0; b < 2; ++b) {
      if (k.buf[b]) (void)hipHostFree(k.buf[b]);
      if (k.ev[b]) (void)hipEventDestroy(k.ev[b]);
    }
    k = Worker{};
  }

  // Frees every worker, the streams and the start event (caller holds mu).
  void Release() {
    for (auto &st : stream)
      if (st) (void)hipStreamSynchronize(st);
    for (int i = 0; i < n; ++i) FreeWorker(w[i]);
    n = 0;
    for (auto &st : stream) {
      if (st) (void)hipStreamDestroy(st);
      st = nullptr;
    }
    if (start) (void)hipEventDestroy(start);
    start = nullptr;
  }
};

// One blocking copy through the HIP runtime's own staging.
int PlainCopy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (hipMemcpyAsync(dst, src, bytes, kind, s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  return hipStreamSynchronize(s) == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

Pool &ThePool() {
  static Pool p;
  return p;
}

int WorkersFor(size_t bytes) {
  const unsigned hc = std::thread::hardware_concurrency();
  int w = hc >= 2 ? (int)std::min<unsigned>(hc / 2, kMaxWorkers) : 1;
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  return (int)std::min<size_t>((size_t)w, chunks);
}

// Moves `bytes` between device `dev` and host `host` (to_device: host ->
// device) through the worker pool.
int Pipeline(void *dev, void *host, size_t bytes, bool to_device, hipStream_t s) {
  Pool &P = ThePool();
  std::lock_guard<std::mutex> lk(P.mu);
  const int W = std::min(WorkersFor(bytes), P.Ensure(WorkersFor(bytes)));
  if (W < 1)
    return to_device ? PlainCopy(dev, host, bytes, hipMemcpyHostToDevice, s)
                     : PlainCopy(host, dev, bytes, hipMemcpyDeviceToHost, s);
  if (hipEventRecord(P.start, s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  for (auto &st : P.stream)
    if (hipStreamWaitEvent(st, P.start, 0) != hipSuccess) return PS_HIP_ERR_DEVICE;
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  std::atomic<bool> ok{true};
  const int device = Runtime::Get().device();
  char *d = static_cast<char *>(dev);
  char *h = static_cast<char *>(host);

  auto run = [&](int wi) {
    Worker &k = P.w[wi];
    auto len = [&](size_t i) { return std::min(kChunk, bytes - i * kChunk); };
    // A new host thread starts on device 0: select the runtime's device.
    bool good = hipSetDevice(device) == hipSuccess;
    hipStream_t cs = P.stream[wi % kDmaStreams];
    if (to_device) {
      bool used[2] = {false, false};
      int b = 0;
      for (size_t i = wi; good && i < chunks && ok.load(std::memory_order_relaxed); i += W, b ^= 1) {
        // Buffer b still feeds the DMA of chunk i - 2W until its event fires.
        if (used[b] && hipEventSynchronize(k.ev[b]) != hipSuccess) {
          good = false;
          break;
        }
        std::memcpy(k.buf[b], h + i * kChunk, len(i));
        good = hipMemcpyAsync(d + i * kChunk, k.buf[b], len(i), hipMemcpyHostToDevice, cs) == hipSuccess &&
               hipEventRecord(k.ev[b], cs) == hipSuccess;
        used[b] = true;
      }
    } else {
      auto issue = [&](size_t i, int b) {
        return hipMemcpyAsync(k.buf[b], d + i * kChunk, len(i), hipMemcpyDeviceToHost, cs) == hipSuccess &&
               hipEventRecord(k.ev[b], cs) == hipSuccess;
      };
      int b = 0;
      if (good && (size_t)wi < chunks) good = issue(wi, 0);
      for (size_t i = wi; good && i < chunks && ok.load(std::memory_order_relaxed); i += W, b ^= 1) {
        // Buffer b ^ 1 was emptied by the previous iteration's memcpy.
        if (i + W < chunks && !issue(i + W, b ^ 1)) {
          good = false;
          break;
        }
        if (hipEventSynchronize(k.ev[b]) != hipSuccess) {
          good = false;
          break;
        }
        std::memcpy(h + i * kChunk, k.buf[b], len(i));
      }
    }
    // Every exit waits for this worker's transfers: no DMA into or out of a
    // staging buffer outlives the call.
    for (int b = 0; b < 2; ++b)
      if (hipEventSynchronize(k.ev[b]) != hipSuccess) good = false;
    if (!good) ok.store(false, std::memory_order_relaxed);
  };

  std::vector<std::thread> threads;
  std::vector<int> inline_workers;  // workers no thread could be started for
  threads.reserve(W - 1);
  for (int wi = 1; wi < W; ++wi) {
    try {
      threads.emplace_back(run, wi);
    } catch (const std::system_error &) {
      inline_workers.push_back(wi);
    }
  }
  run(0);
  for (int wi : inline_workers) run(wi);
  for (auto &t : threads) t.join();
  return ok.load() ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

}  // namespace

void PrepareHostCopy(size_t bytes) {
  if (bytes < kPipelineMin) return;
  Pool &P = ThePool();
  std::lock_guard<std::mutex> lk(P.mu);
  P.Ensure(WorkersFor(bytes));
}

int CopyHostToDevice(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes >= kPipelineMin) return Pipeline(dst, const_cast<void *>(src), bytes, true, s);
  return PlainCopy(dst, src, bytes, hipMemcpyHostToDevice, s);
}

int CopyDeviceToHost(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes >= kPipelineMin) return Pipeline(const_cast<void *>(src), dst, bytes, false, s);
  return PlainCopy(dst, src, bytes, hipMemcpyDeviceToHost, s);
}

void ReleaseHostCopyPool() {
  Pool &P = ThePool();
  std::lock_guard<std::mutex> lk(P.mu);
  P.Release();
}

}  // namespace physis_hip
