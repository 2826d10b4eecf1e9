#include "runtime.hpp"

#include <algorithm>
#include <array>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <thread>

namespace pk {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PK_ERR_OOM : PK_ERR_DEVICE;
  (void)hipGetLastError();  // clear the sticky error of this call
  throw Error(code, std::string(hipGetErrorString(e)) + " in " + what + " (" + file + ":" +
                        std::to_string(line) + ")");
}

static std::mutex g_dev_mtx;
static std::map<int, std::unique_ptr<Device>>& devices() {
  static std::map<int, std::unique_ptr<Device>> m;
  return m;
}

Device& device_get(int id) {
  std::lock_guard<std::mutex> lk(g_dev_mtx);
  auto& m = devices();
  auto it = m.find(id);
  if (it != m.end()) return *it->second;
  auto d = std::make_unique<Device>();
  d->id = id;
  // Scratch buffers come from the device's stream-ordered pool.  By default
  // the pool hands freed memory back to the driver at every synchronisation,
  // so the next call maps it again (a 128 MiB buffer costs about as much as
  // the kernel that uses it); keep it cached instead.
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, id) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  (void)hipGetLastError();
  Device& ref = *d;
  m.emplace(id, std::move(d));
  return ref;
}

Device& device_current() {
  int id = 0;
  PK_HIP(hipGetDevice(&id));
  return device_get(id);
}

// Requests are rounded up to one of eight size classes per power of two (at
// most 12.5% over), so that calls whose sizes differ slightly (three matrices
// with nearly equal entry counts) reuse the blocks the pool keeps instead of
// mapping new memory each time.
Scratch::Scratch(size_t bytes, hipStream_t st) : s(st) {
  if (bytes == 0) bytes = 16;
  if (bytes > 4096) {
    int lg = 63 - __builtin_clzll(bytes);
    const size_t step = size_t(1) << (lg - 3);
    bytes = (bytes + step - 1) / step * step;
  }
  PK_HIP(hipMallocAsync(&p, bytes, s));
}
Scratch::~Scratch() {
  if (p) (void)hipFreeAsync(p, s);
}

namespace {
struct ThreadStreams {
  std::map<int, hipStream_t> by_dev;
  ~ThreadStreams() {
    for (auto& kv : by_dev) (void)hipStreamDestroy(kv.second);
  }
};
}  // namespace

std::shared_mutex& capture_mutex() {
  static std::shared_mutex m;
  return m;
}

hipStream_t thread_stream() {
  thread_local ThreadStreams ts;
  int id = 0;
  PK_HIP(hipGetDevice(&id));
  auto it = ts.by_dev.find(id);
  if (it != ts.by_dev.end()) return it->second;
  hipStream_t s;
  PK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  ts.by_dev[id] = s;
  return s;
}

}  // namespace pk

// ------------------------------------------------------------ staged copies
namespace pk {
namespace {

// Pinned staging for large copies from pageable host memory.  A chunk is
// copied into one of three pinned buffers by a small pool of host threads
// at once, then sent by DMA while the next chunk is being staged; a buffer
// is refilled only after the DMA that read it has finished (an event).  The
// buffers, events and threads live until the process ends.
class Stager {
 public:
  static constexpr int kBufs = 3;
  static constexpr size_t kChunk = size_t(16) << 20;

  static Stager& get() {
    static Stager* st = new Stager();  // never destroyed: its threads wait forever
    return *st;
  }

  void copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(copy_mtx_);
    if (!buf_[0])
      for (int b = 0; b < kBufs; b++)
        PK_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf_[b]), kChunk, hipHostMallocDefault));
    int dev = 0;
    PK_HIP(hipGetDevice(&dev));
    auto& ev = events_[dev];
    if (!ev[0])
      for (int b = 0; b < kBufs; b++) PK_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    const uint8_t* from = static_cast<const uint8_t*>(src);
    uint8_t* to = static_cast<uint8_t*>(dst);
    for (size_t off = 0, i = 0; off < bytes; off += kChunk, i++) {
      const int b = (int)(i % kBufs);
      if (last_[b]) PK_HIP(hipEventSynchronize(last_[b]));  // the DMA that read it is done
      const size_t len = bytes - off < kChunk ? bytes - off : kChunk;
      parallel_memcpy(buf_[b], from + off, len);
      PK_HIP(hipMemcpyAsync(to + off, buf_[b], len, hipMemcpyHostToDevice, s));
      PK_HIP(hipEventRecord(ev[b], s));
      last_[b] = ev[b];
    }
  }

 private:
  Stager() {
    const unsigned hc = std::thread::hardware_concurrency();
    parts_ = (int)std::max(1u, std::min(16u, hc ? hc : 1u));
    for (int w = 0; w + 1 < parts_; w++) std::thread([this, w] { work(w); }).detach();
  }

  // dst[0, len) = src[0, len), split over the pool and the calling thread
  void parallel_memcpy(uint8_t* dst, const uint8_t* src, size_t len) {
    if (parts_ == 1) {
      std::memcpy(dst, src, len);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      job_dst_ = dst;
      job_src_ = src;
      job_len_ = len;
      pending_ = parts_ - 1;
      gen_++;
    }
    cv_.notify_all();
    part(parts_ - 1, dst, src, len);
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
  }
  void part(int k, uint8_t* dst, const uint8_t* src, size_t len) const {
    const size_t per = (len + parts_ - 1) / parts_;
    const size_t a = std::min(len, per * k), e = std::min(len, a + per);
    if (e > a) std::memcpy(dst + a, src + a, e - a);
  }
  void work(int k) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      cv_.wait(lk, [&] { return gen_ != seen; });
      seen = gen_;
      uint8_t* dst = job_dst_;
      const uint8_t* src = job_src_;
      const size_t len = job_len_;
      lk.unlock();
      part(k, dst, src, len);
      lk.lock();
      if (--pending_ == 0) done_cv_.notify_one();
    }
  }

  std::mutex copy_mtx_, m_;
  std::condition_variable cv_, done_cv_;
  uint8_t* buf_[kBufs] = {};
  hipEvent_t last_[kBufs] = {};
  std::map<int, std::array<hipEvent_t, kBufs>> events_;
  int parts_ = 1;
  uint64_t gen_ = 0;
  int pending_ = 0;
  uint8_t* job_dst_ = nullptr;
  const uint8_t* job_src_ = nullptr;
  size_t job_len_ = 0;
};

}  // namespace

void* pinned_small() {
  thread_local void* p = nullptr;
  if (!p) PK_HIP(hipHostMalloc(&p, 4096, hipHostMallocDefault));
  return p;
}

void h2d_from_pageable(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < (size_t(32) << 20)) {
    if (bytes) PK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  Stager::get().copy(dst, src, bytes, s);
}

}  // namespace pk
