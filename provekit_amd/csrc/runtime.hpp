// Device runtime shared by every entry point: errors, per-device state
// (streams, cached twiddle and coset-power tables), stream-ordered scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fr.hpp"

namespace pk {

// Status codes of the C ABI (include/provekit_gpu.h mirrors them).
enum Status : int {
  PK_OK = 0,
  PK_ERR_ARG = 1,        // null pointer, length mismatch, bad enum
  PK_ERR_NOT_POW2 = 2,   // size must be a power of two
  PK_ERR_SIZE = 3,       // divisibility / range violations
  PK_ERR_OOM = 4,        // device allocation failed
  PK_ERR_DEVICE = 5,     // HIP runtime error
  PK_ERR_BUDGET = 6,     // serving admission refused (HBM budget)
  PK_ERR_STATE = 7,      // handle in the wrong state / unknown handle
  PK_ERR_INTERNAL = 8,   // a self-check failed
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);
#define PK_HIP(x)                                                   \
  do {                                                              \
    hipError_t _e = (x);                                            \
    if (_e != hipSuccess) ::pk::throw_hip(_e, #x, __FILE__, __LINE__); \
  } while (0)
#define PK_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) throw ::pk::Error((code), (msg)); \
  } while (0)

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline int log2_exact(uint64_t x) {
  int k = 0;
  while ((1ull << k) < x) k++;
  return k;
}

// A growable device table of prefix-consistent values (roots in bit-reversed
// order, powers of the coset generator).  Growth builds a larger table and
// publishes it; the smaller one is retired but never freed while the process
// runs, so kernels already enqueued against it stay valid without readers
// holding a lock (the sizes grow geometrically, so retired tables cost less
// than the live one).
struct GrowTable {
  fr_t* ptr = nullptr;
  int log = -1;  // table covers 2^log entries
  std::vector<fr_t*> retired;
};

struct Device {
  int id = 0;
  std::mutex tables_mtx;
  GrowTable roots;   // rr[j] = w^{bitrev(j)}, w of order 2^(log+1)
  GrowTable coset;   // g^j, g = 5 (multiplicative generator of Fr)
  bool p2_ready = false;  // Poseidon2 parameters uploaded to constant memory
};

Device& device_current();
Device& device_get(int id);

// Tables holding at least 2^log entries (built on first use, then cached).
const fr_t* roots_table(Device& d, int half_log);  // >= 2^half_log roots
const fr_t* coset_table(Device& d, int log);       // >= 2^log powers of 5
void poseidon2_consts(Device& d);

// Stream-ordered scratch allocation freed at scope exit (on the same stream).
struct Scratch {
  void* p = nullptr;
  hipStream_t s = nullptr;
  Scratch() = default;
  Scratch(size_t bytes, hipStream_t st);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  Scratch(Scratch&& o) noexcept : p(o.p), s(o.s) { o.p = nullptr; }
  Scratch& operator=(Scratch&& o) noexcept {
    std::swap(p, o.p);
    std::swap(s, o.s);
    return *this;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// The calling thread's work stream on the current device (default priority).
hipStream_t thread_stream();

// HIP invalidates a stream capture when another thread makes a synchronous
// call meanwhile (a hipMemcpy on the legacy stream fails with "would make
// the legacy stream depend on a capturing blocking stream", and launches on
// other streams then fail too), even for a thread-local capture on a
// non-blocking stream.  Every entry point holds this lock shared while it
// makes HIP calls; the entry points that capture a graph hold it exclusively
// for the capture, which only enqueues.
std::shared_mutex& capture_mutex();

// Host-to-device copy on stream s from pageable host memory.  Copies of 32 MiB
// and more go through pinned staging buffers filled by several host threads
// while the previous chunk is in flight; smaller ones are a plain async copy.
// src may be reused when the call returns.
void h2d_from_pageable(void* dst, const void* src, size_t bytes, hipStream_t s);

// A 4 KiB pinned host buffer of the calling thread (kept until the process
// ends), for small device-to-host reads that the thread waits on at once.
void* pinned_small();

}  // namespace pk
