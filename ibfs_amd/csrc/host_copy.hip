// Device-to-host copies into pageable host memory.
//
// A copy into memory the runtime did not allocate goes through the runtime's
// own staging, which measured 4 GB/s into untouched pages and 17.5 GB/s into
// touched ones on an MI355X host (2-GiB copies), against 57 GB/s into pinned
// memory.  Depth arrays are large (2 GiB per 128-lane group at R-MAT scale
// 22), so a large copy here streams through two 64-MiB pinned chunks: the DMA
// engine fills one chunk while eight host threads copy the other out, which
// measured 55 GB/s into pageable memory (8 and 16 threads alike; 4 threads
// 46 GB/s).
#include "host_copy.h"

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace ibfs {
namespace {

constexpr size_t kChunk = size_t(64) << 20;
constexpr int kThreads = 8;
// copies below this size keep the runtime's path (no pipeline to fill)
constexpr size_t kDirect = size_t(4) << 20;

// Host threads that copy one buffer in kThreads slices; copy() returns when
// every slice is done.  The threads live as long as the process.
class CopyPool {
 public:
  CopyPool() {
    for (int k = 0; k < kThreads; ++k) th_.emplace_back([this, k] { run(k); });
  }

  void copy(void *dst, const void *src, size_t bytes) {
    std::unique_lock<std::mutex> lk(mu_);
    dst_ = static_cast<char *>(dst);
    src_ = static_cast<const char *>(src);
    bytes_ = bytes;
    left_ = kThreads;
    ++gen_;
    go_.notify_all();
    done_.wait(lk, [this] { return left_ == 0; });
  }

 private:
  void run(int k) {
    uint64_t seen = 0;
    for (;;) {
      char *d;
      const char *s;
      size_t b;
      {
        std::unique_lock<std::mutex> lk(mu_);
        go_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        d = dst_;
        s = src_;
        b = bytes_;
      }
      const size_t lo = b * (size_t)k / kThreads, hi = b * (size_t)(k + 1) / kThreads;
      if (hi > lo) std::memcpy(d + lo, s + lo, hi - lo);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_one();
    }
  }

  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable go_, done_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
  int left_ = 0;
  uint64_t gen_ = 0;
};

// The pinned chunks and the pool, made on first use and kept for the
// process; one pipelined copy at a time.
struct Stage {
  std::mutex mu;
  char *pin = nullptr;
  CopyPool *pool = nullptr;
};

Stage &stage() {
  static Stage s;
  return s;
}

int wait_event(hipEvent_t ev) {
  WaitClock clk;
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return IBFS_OK;
    if (q != hipErrorNotReady) return fail(IBFS_ERR_HIP, "copy to host: %s", hipGetErrorString(q));
    if (clk.expired())
      return fail(IBFS_ERR_STATE, "copy to host did not finish within the wait limit (%.1f s)",
                  wait_timeout_s());
  }
}

// The runtime's path: every piece on `st`, then one wait.
int copy_direct(const D2HPiece *p, size_t n, hipStream_t st) {
  for (size_t i = 0; i < n; ++i) {
    const hipError_t e = hipMemcpyAsync(p[i].dst, p[i].src, p[i].bytes, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return fail(IBFS_ERR_HIP, "copy to host: %s", hipGetErrorString(e));
  }
  hipEvent_t ev = nullptr;
  hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ev, st);
  int rc = e == hipSuccess ? wait_event(ev)
                           : fail(IBFS_ERR_HIP, "copy to host: %s", hipGetErrorString(e));
  if (ev) (void)hipEventDestroy(ev);
  return rc;
}

}  // namespace

int copy_to_host(const D2HPiece *pieces, size_t n, hipStream_t st) {
  size_t total = 0;
  for (size_t i = 0; i < n; ++i) total += pieces[i].bytes;
  if (total == 0) return IBFS_OK;
  if (total < kDirect) return copy_direct(pieces, n, st);
  Stage &s = stage();
  std::lock_guard<std::mutex> lk(s.mu);
  if (!s.pin) {
    void *p = nullptr;
    if (hipHostMalloc(&p, 2 * kChunk, hipHostMallocPortable) != hipSuccess || !p) {
      (void)hipGetLastError();
      return copy_direct(pieces, n, st);  // no pinned memory: the runtime's path
    }
    s.pin = static_cast<char *>(p);
  }
  if (!s.pool) {
    try {
      s.pool = new CopyPool();
    } catch (...) {
      return copy_direct(pieces, n, st);  // no threads: the runtime's path
    }
  }
  // the pieces cut into chunk-sized parts, copied in order
  std::vector<D2HPiece> q;
  for (size_t i = 0; i < n; ++i)
    for (size_t off = 0; off < pieces[i].bytes; off += kChunk)
      q.push_back({static_cast<char *>(pieces[i].dst) + off,
                   static_cast<const char *>(pieces[i].src) + off,
                   std::min(kChunk, pieces[i].bytes - off)});
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  int rc = e == hipSuccess ? IBFS_OK
                           : fail(IBFS_ERR_HIP, "copy to host: %s", hipGetErrorString(e));
  // part j goes to chunk j & 1; that chunk's previous part (j - 2) was copied
  // out before part j - 1 was waited for, so the DMA never overwrites data
  // still being read
  size_t issued = 0;
  for (size_t j = 0; rc == IBFS_OK && j < q.size(); ++j) {
    char *slot = s.pin + (j & 1) * kChunk;
    e = hipMemcpyAsync(slot, q[j].src, q[j].bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(ev[j & 1], st);
    if (e != hipSuccess) {
      rc = fail(IBFS_ERR_HIP, "copy to host: %s", hipGetErrorString(e));
      break;
    }
    issued = j + 1;
    if (j >= 1) {
      rc = wait_event(ev[(j - 1) & 1]);
      if (rc == IBFS_OK) s.pool->copy(q[j - 1].dst, s.pin + ((j - 1) & 1) * kChunk, q[j - 1].bytes);
    }
  }
  if (rc == IBFS_OK) {
    const size_t last = q.size() - 1;
    rc = wait_event(ev[last & 1]);
    if (rc == IBFS_OK) s.pool->copy(q[last].dst, s.pin + (last & 1) * kChunk, q[last].bytes);
  } else if (issued > 0) {
    // An error part-way (a failed copy or a wait past the limit): DMA already
    // issued may still write the chunks, and waiting for it could outlast the
    // limit.  The chunks are abandoned, never freed, and the next pipelined
    // copy allocates new ones.
    s.pin = nullptr;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

}  // namespace ibfs
