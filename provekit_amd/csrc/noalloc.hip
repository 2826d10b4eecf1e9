// Scratch for the allocation-free entry point (pkgpu_sparse_mul_device_noalloc):
// a file-scope device array of a fixed size, one copy per GPU for the life of
// the process, so a call allocates nothing.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current, and only that address is used.
#include "kernels.hpp"
#include "sumcheck.hpp"

namespace pk {

namespace {

// flags (2 * chunks + 1 u32, rounded up to 256 B), then 2 * chunks partial sums
constexpr uint64_t kMaxChunks = (kNoAllocSpmvMaxNnz + 2047) / 2048;
constexpr uint64_t kFlagBytes = ((4 * (2 * kMaxChunks + 1)) + 255) / 256 * 256;
constexpr uint64_t kScratchBytes = kFlagBytes + 2 * sizeof(fr_t) * kMaxChunks;

__device__ __align__(256) static unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mtx;
void* g_ptr[kMaxDevices];

void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_mtx);
  PK_REQUIRE(device >= 0 && device < kMaxDevices, PK_ERR_ARG, "no-alloc scratch: device id");
  if (!g_ptr[device]) {
    void* p = nullptr;
    PK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PK_REQUIRE(p != nullptr, PK_ERR_DEVICE, "no-alloc scratch: no device address");
    g_ptr[device] = p;
  }
  return g_ptr[device];
}

}  // namespace

void noalloc_setup(int device) { (void)scratch_ptr(device); }

void spmv_dev_noalloc(const CsrDev& A, const fr_t* z, fr_t* out, hipStream_t s) {
  PK_REQUIRE(A.nnz <= kNoAllocSpmvMaxNnz, PK_ERR_SIZE,
             "sparse_mul_noalloc: more than 2^26 entries (the fixed scratch holds 2^26)");
  if (!A.rows) return;
  int dev = 0;
  PK_HIP(hipGetDevice(&dev));
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));
  static_assert(kFlagBytes % 256 == 0, "partials stay 256-byte aligned");
  spmv_launch(A, z, out, reinterpret_cast<uint32_t*>(base),
              reinterpret_cast<fr_t*>(base + kFlagBytes), s);
}

}  // namespace pk
