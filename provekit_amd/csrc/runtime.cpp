#include "runtime.hpp"

#include <map>
#include <memory>

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
