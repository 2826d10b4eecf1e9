// C ABI (include/provekit_gpu.h) and the C++ host mirror (provekit.hpp).
#include <string.h>

#include <algorithm>

#include <string>

#include "../../include/provekit_gpu.h"
#include "host_fr.hpp"
#include "kernels.hpp"
#include "provekit.hpp"
#include "batch.hpp"
#include "serve.hpp"
#include "sumcheck.hpp"
#include "wire.hpp"

using pk::Error;
using pk::fr_t;

namespace {

thread_local std::string g_last_error;

// Entry points hold the capture lock shared (see capture_mutex); the ones
// that capture a graph take it exclusively instead, through guarded_capture.
// Entry points that only wait on the server's own condition variables take
// no lock (guarded_nolock), so a capture is not held up by a long wait.
enum class Lock { Shared, Exclusive, None };
template <Lock L, class F>
int guarded_impl(F&& f) {
  try {
    std::shared_lock<std::shared_mutex> shared(pk::capture_mutex(), std::defer_lock);
    std::unique_lock<std::shared_mutex> excl(pk::capture_mutex(), std::defer_lock);
    if (L == Lock::Exclusive) excl.lock();
    if (L == Lock::Shared) shared.lock();
    f();
    g_last_error.clear();
    return PKGPU_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return PKGPU_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return PKGPU_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "unknown failure";
    return PKGPU_ERR_INTERNAL;
  }
}
template <class F>
int guarded(F&& f) {
  return guarded_impl<Lock::Shared>(std::forward<F>(f));
}
template <class F>
int guarded_capture(F&& f) {
  return guarded_impl<Lock::Exclusive>(std::forward<F>(f));
}
template <class F>
int guarded_nolock(F&& f) {
  return guarded_impl<Lock::None>(std::forward<F>(f));
}

hipStream_t stream_or_thread(void* s) {
  return s ? static_cast<hipStream_t>(s) : pk::thread_stream();
}

// Host <-> device staging for the slice-taking entry points.
struct Staged {
  pk::Scratch buf;
  Staged(size_t bytes, hipStream_t s) : buf(bytes, s) {}
  template <class T>
  T* get() const { return buf.as<T>(); }
};

// host slices of 32 MiB and more go through the pinned staging pipeline
void h2d(void* d, const void* h, size_t bytes, hipStream_t s) { pk::h2d_from_pageable(d, h, bytes, s); }
void d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
  if (bytes) PK_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
}
void sync(hipStream_t s) { PK_HIP(hipStreamSynchronize(s)); }

#define REQ(c, code, msg) PK_REQUIRE(c, code, msg)

}  // namespace

namespace provekit_gpu {

// ntt/src/ntt.rs:105-108 with the asserts of interleaved_ntt_nr (:160-169)
void ntt_nr(Fr* values, size_t len, size_t codeword_size, size_t num_groups) {
  if (codeword_size <= 1) return;  // ntt.rs:160-162
  REQ(num_groups >= 1 && len % num_groups == 0, pk::PK_ERR_SIZE,
      "values.len() must be divisible by num_groups");
  REQ(pk::is_pow2(codeword_size), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  REQ(len % codeword_size == 0 && pk::is_pow2(num_groups) && num_groups <= codeword_size,
      pk::PK_ERR_SIZE,
      "ntt_nr: len must be a multiple of codeword_size and num_groups a power of two up to it");
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged b(sizeof(fr_t) * len, s);
  h2d(b.get<void>(), values, sizeof(fr_t) * len, s);
  pk::ntt_nr_dev(d, b.get<fr_t>(), len, codeword_size, num_groups, s);
  d2h(values, b.get<void>(), sizeof(fr_t) * len, s);
  sync(s);
}

void intt_rn(Fr* values, size_t len) {
  if (len <= 1) return;
  REQ(pk::is_pow2(len), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged b(sizeof(fr_t) * len, s);
  h2d(b.get<void>(), values, sizeof(fr_t) * len, s);
  pk::intt_rn_dev(d, b.get<fr_t>(), len, s);
  d2h(values, b.get<void>(), sizeof(fr_t) * len, s);
  sync(s);
}

// ntt's intt_nr (ntt/src/ntt.rs:333-348), a private helper of intt_rn there.
void intt_nr(Fr* values, size_t len) {
  if (len <= 1) return;
  REQ(pk::is_pow2(len), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged b(sizeof(fr_t) * len, s);
  h2d(b.get<void>(), values, sizeof(fr_t) * len, s);
  pk::intt_nr_dev(d, b.get<fr_t>(), len, s);
  d2h(values, b.get<void>(), sizeof(fr_t) * len, s);
  sync(s);
}

// ntt's reverse_order (ntt/src/ntt.rs:310-323) on the first codeword_size
// entries.
void reverse_order(Fr* values, size_t len, size_t codeword_size) {
  REQ(pk::is_pow2(codeword_size), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  REQ(codeword_size <= len, pk::PK_ERR_SIZE, "reverse_order: codeword_size above the slice length");
  if (codeword_size <= 1) return;
  hipStream_t s = pk::thread_stream();
  Staged b(sizeof(fr_t) * codeword_size, s);
  h2d(b.get<void>(), values, sizeof(fr_t) * codeword_size, s);
  pk::reverse_order_dev(b.get<fr_t>(), codeword_size, s);
  d2h(values, b.get<void>(), sizeof(fr_t) * codeword_size, s);
  sync(s);
}

std::optional<size_t> RSFr::next_order(size_t size) const {
  size_t o = 1;
  while (o < size) o <<= 1;
  if (o <= (size_t(1) << 28)) return o;
  return std::nullopt;
}

Fr RSFr::generator(size_t codeword_length) const {
  REQ(pk::is_pow2(codeword_length) && codeword_length <= (size_t(1) << 28), pk::PK_ERR_SIZE,
      "get_root_of_unity: codeword length must be a power of two up to 2^28");
  pk::hfr g = pk::hf::root_of_unity_log((unsigned)pk::log2_exact(codeword_length));
  Fr r;
  memcpy(r.data(), g.l, 32);
  return r;
}

std::vector<Fr> RSFr::evaluation_points(size_t, size_t codeword_length,
                                        const std::vector<size_t>& indices) const {
  REQ(pk::is_pow2(codeword_length) && codeword_length <= (size_t(1) << 28), pk::PK_ERR_SIZE,
      "get_root_of_unity: codeword length must be a power of two up to 2^28");
  std::vector<Fr> out(indices.size());
  if (indices.empty()) return out;
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged di(sizeof(uint64_t) * indices.size(), s);
  Staged dout(sizeof(fr_t) * indices.size(), s);
  std::vector<uint64_t> idx(indices.begin(), indices.end());
  h2d(di.get<void>(), idx.data(), sizeof(uint64_t) * idx.size(), s);
  pk::rs_points_dev(d, codeword_length, di.get<uint64_t>(), idx.size(), dout.get<fr_t>(), s);
  d2h(out.data(), dout.get<void>(), sizeof(fr_t) * out.size(), s);
  sync(s);
  return out;
}

std::vector<Fr> RSFr::interleaved_encode(const std::vector<const Fr*>& messages,
                                         size_t message_length, const std::vector<Fr>& masks,
                                         size_t codeword_length) const {
  if (messages.empty()) return {};
  const size_t M = messages.size();
  REQ(M * message_length + masks.size() <= M * codeword_length, pk::PK_ERR_SIZE,
      "interleaved_encode: messages and masks longer than the codeword");
  if (codeword_length == 0) return {};  // empty messages into an empty codeword
  REQ(pk::is_pow2(codeword_length), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged dm(sizeof(fr_t) * M * message_length, s);
  Staged dk(sizeof(fr_t) * masks.size(), s);
  Staged dout(sizeof(fr_t) * M * codeword_length, s);
  for (size_t r = 0; r < M; r++)
    h2d(dm.get<fr_t>() + r * message_length, messages[r], sizeof(fr_t) * message_length, s);
  h2d(dk.get<void>(), masks.data(), sizeof(fr_t) * masks.size(), s);
  pk::rs_encode_dev(d, dm.get<fr_t>(), M, message_length, dk.get<fr_t>(), masks.size(),
                    codeword_length, dout.get<fr_t>(), 0, s);
  std::vector<Fr> out(M * codeword_length);
  d2h(out.data(), dout.get<void>(), sizeof(fr_t) * out.size(), s);
  sync(s);
  return out;
}

namespace skyscraper {
void compress_many(const uint8_t* messages, size_t messages_len, uint8_t* hashes,
                   size_t hashes_len) {
  REQ(messages_len % 64 == 0, pk::PK_ERR_SIZE, "Message length not a multiple of 64");
  REQ(hashes_len % 32 == 0, pk::PK_ERR_SIZE, "Hashes length not a multiple of 32");
  REQ(messages_len / 64 == hashes_len / 32, pk::PK_ERR_SIZE,
      "Messages and hashes length mismatch");
  const size_t n = hashes_len / 32;
  if (!n) return;
  hipStream_t s = pk::thread_stream();
  Staged dm(messages_len, s), dh(hashes_len, s);
  h2d(dm.get<void>(), messages, messages_len, s);
  pk::sky_compress_many_dev(dm.get<fr_t>(), dh.get<fr_t>(), n, s);
  d2h(hashes, dh.get<void>(), hashes_len, s);
  sync(s);
}
Fr compress(const Fr& l, const Fr& r) {
  uint8_t m[64];
  memcpy(m, l.data(), 32);
  memcpy(m + 32, r.data(), 32);
  Fr o;
  compress_many(m, 64, reinterpret_cast<uint8_t*>(o.data()), 32);
  return o;
}

namespace pow {
Fr threshold(double difficulty) {
  Fr t;
  pk::pow_threshold(difficulty, t.data());
  return t;
}
bool verify(const Fr& challenge, double difficulty, uint64_t nonce) {
  if (difficulty == 0.0) return true;
  const Fr t = threshold(difficulty);
  const Fr h = compress(challenge, Fr{nonce, 0, 0, 0});
  for (int i = 3; i >= 0; i--)
    if (h[i] != t[i]) return h[i] < t[i];
  return false;
}
uint64_t solve(const Fr& challenge, double difficulty) {
  uint64_t nonce = 0;
  const bool found =
      pk::sky_pow_solve_dev(challenge.data(), difficulty, ~0ULL, &nonce, pk::thread_stream());
  REQ(found, pk::PK_ERR_INTERNAL, "proof of work: no nonce below 2^64 - 1");
  return nonce;
}
}  // namespace pow
}  // namespace skyscraper

SkyscraperPoW::SkyscraperPoW(const std::array<uint8_t, 32>& challenge, double bits)
    : challenge_(challenge), bits_(bits) {
  REQ(bits >= 0.0 && bits < 60.0, pk::PK_ERR_ARG, "bits must be smaller than 60");
}
bool SkyscraperPoW::check(uint64_t nonce) const {
  Fr c;
  memcpy(c.data(), challenge_.data(), 32);
  return skyscraper::pow::verify(c, bits_, nonce);
}
SkyscraperPoW::Solution SkyscraperPoW::solve() const {
  Fr c;
  memcpy(c.data(), challenge_.data(), 32);
  return Solution{challenge_, skyscraper::pow::solve(c, bits_)};
}

void SkyscraperHashEngine::hash_many(size_t size, const uint8_t* input, size_t input_len,
                                     Hash* output, size_t count) const {
  REQ(supports_size(size), pk::PK_ERR_SIZE,
      "skyscraper: unsupported message size " + std::to_string(size) +
          " (must be a positive multiple of 32)");
  REQ(input_len == size * count, pk::PK_ERR_SIZE,
      "skyscraper: input length " + std::to_string(input_len) + " != size " +
          std::to_string(size) + " * count " + std::to_string(count));
  if (!count) return;
  if (size == 32) {
    memcpy(output, input, input_len);
    return;
  }
  hipStream_t s = pk::thread_stream();
  Staged di(input_len, s), dout(32 * count, s);
  h2d(di.get<void>(), input, input_len, s);
  pk::sky_hash_many_dev(size / 32, di.get<fr_t>(), count, dout.get<fr_t>(), s);
  d2h(output, dout.get<void>(), 32 * count, s);
  sync(s);
}

Hash commit(const Fr* coeffs, size_t n, size_t blowup, bool coset, size_t leaf_elems,
            Fr* codeword_out, Hash* nodes_out) {
  REQ(pk::is_pow2(n) && pk::is_pow2(blowup), pk::PK_ERR_NOT_POW2,
      "commit: length and blowup must be powers of two");
  REQ(leaf_elems >= 1 && pk::is_pow2(leaf_elems) && (n * blowup) % leaf_elems == 0,
      pk::PK_ERR_SIZE, "commit: leaf size must be a power of two dividing the codeword");
  const size_t N = n * blowup, leaves = N / leaf_elems, nodes = 2 * leaves - 1;
  hipStream_t s = pk::thread_stream();
  pk::Device& d = pk::device_current();
  Staged dc(sizeof(fr_t) * n, s), dw(sizeof(fr_t) * N, s), dn(sizeof(fr_t) * nodes, s);
  h2d(dc.get<void>(), coeffs, sizeof(fr_t) * n, s);
  pk::commit_dev(d, dc.get<fr_t>(), n, blowup, coset ? 1 : 0, leaf_elems, dw.get<fr_t>(),
                 dn.get<fr_t>(), s);
  Hash root;
  d2h(root.data(), dn.get<fr_t>() + (nodes - 1), 32, s);
  if (codeword_out) d2h(codeword_out, dw.get<void>(), sizeof(fr_t) * N, s);
  if (nodes_out) d2h(nodes_out, dn.get<void>(), 32 * nodes, s);
  sync(s);
  return root;
}

}  // namespace provekit_gpu

// ================================================================== C ABI
using namespace provekit_gpu;

extern "C" {

int pkgpu_abi_version(void) { return PKGPU_ABI_VERSION; }

const char* pkgpu_last_error(void) { return g_last_error.c_str(); }

int pkgpu_init(int device) {
  return guarded([&] {
    int n = 0;
    PK_HIP(hipGetDeviceCount(&n));
    REQ(device >= 0 && device < n, pk::PK_ERR_ARG, "pkgpu_init: no such device");
    PK_HIP(hipSetDevice(device));
    pk::device_get(device);
    (void)pk::thread_stream();
    pk::noalloc_setup(device);
  });
}

int pkgpu_device_synchronize(void) {
  return guarded([&] { PK_HIP(hipDeviceSynchronize()); });
}

int pkgpu_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    REQ(ptr, pk::PK_ERR_ARG, "pkgpu_malloc: null out-pointer");
    PK_HIP(hipMalloc(ptr, bytes ? bytes : 16));
  });
}
int pkgpu_free(void* ptr) {
  return guarded([&] {
    if (ptr) PK_HIP(hipFree(ptr));
  });
}
int pkgpu_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    REQ(bytes == 0 || (dst && src), pk::PK_ERR_ARG, "memcpy: null pointer");
    hipStream_t s = stream_or_thread(stream);
    h2d(dst, src, bytes, s);
    sync(s);
  });
}
int pkgpu_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    REQ(bytes == 0 || (dst && src), pk::PK_ERR_ARG, "memcpy: null pointer");
    hipStream_t s = stream_or_thread(stream);
    d2h(dst, src, bytes, s);
    sync(s);
  });
}
int pkgpu_stream_synchronize(void* stream) {
  return guarded([&] { sync(stream_or_thread(stream)); });
}

int pkgpu_ntt_nr(uint64_t* values, size_t len, size_t codeword_size, size_t num_groups) {
  return guarded([&] {
    REQ(len == 0 || values, pk::PK_ERR_ARG, "ntt_nr: null values");
    ntt_nr(reinterpret_cast<Fr*>(values), len, codeword_size, num_groups);
  });
}
int pkgpu_ntt_nr_device(uint64_t* values, size_t len, size_t codeword_size, size_t num_groups,
                        void* stream) {
  return guarded([&] {
    if (codeword_size <= 1) return;
    REQ(values, pk::PK_ERR_ARG, "ntt_nr: null values");
    REQ(num_groups >= 1 && len % num_groups == 0, pk::PK_ERR_SIZE,
        "values.len() must be divisible by num_groups");
    REQ(pk::is_pow2(codeword_size), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
    REQ(len % codeword_size == 0 && pk::is_pow2(num_groups) && num_groups <= codeword_size,
        pk::PK_ERR_SIZE,
        "ntt_nr: len must be a multiple of codeword_size and num_groups a power of two up to it");
    pk::ntt_nr_dev(pk::device_current(), reinterpret_cast<fr_t*>(values), len, codeword_size,
                   num_groups, stream_or_thread(stream));
  });
}
int pkgpu_ntt_autotune(int log_min, int log_max, int max_skip, const char* save_path,
                       double* seconds) {
  return guarded([&] {
    REQ(log_min >= 1 && log_max <= 27 && log_min <= log_max, pk::PK_ERR_ARG,
        "ntt autotune: log range must lie in [1, 27]");
    REQ(max_skip >= 0 && max_skip <= 8, pk::PK_ERR_ARG, "ntt autotune: max_skip must be 0..8");
    hipStream_t s = pk::thread_stream();
    const double t = pk::ntt_autotune(pk::device_current(), log_min, log_max, max_skip, s);
    if (save_path) pk::ntt_tuning_save(save_path);
    if (seconds) *seconds = t;
  });
}
int pkgpu_ntt_tuning_load(const char* path) {
  return guarded([&] {
    REQ(path, pk::PK_ERR_ARG, "ntt tuning: null path");
    pk::ntt_tuning_load(path);
  });
}
int pkgpu_ntt_tuning_get(int log_codeword, int stages, int* tile_log, int* bpt, int* min_run) {
  return guarded([&] {
    REQ(tile_log && bpt && min_run, pk::PK_ERR_ARG, "ntt tuning: null output");
    REQ(log_codeword >= 0 && stages >= 0 && stages <= log_codeword, pk::PK_ERR_ARG,
        "ntt tuning: stages must lie in [0, log_codeword]");
    const pk::NttShape sh = pk::ntt_tuned_shape(log_codeword, stages);
    *tile_log = sh.tile_log;
    *bpt = sh.bpt ? sh.bpt : pk::ntt_default_bpt(sh.tile_log);
    *min_run = sh.min_run < 0 ? 2 : sh.min_run;
  });
}
int pkgpu_intt_rn(uint64_t* values, size_t len) {
  return guarded([&] {
    REQ(len == 0 || values, pk::PK_ERR_ARG, "intt_rn: null values");
    intt_rn(reinterpret_cast<Fr*>(values), len);
  });
}
int pkgpu_intt_rn_device(uint64_t* values, size_t len, void* stream) {
  return guarded([&] {
    if (len <= 1) return;
    REQ(values, pk::PK_ERR_ARG, "intt_rn: null values");
    pk::intt_rn_dev(pk::device_current(), reinterpret_cast<fr_t*>(values), len,
                    stream_or_thread(stream));
  });
}

int pkgpu_intt_nr(uint64_t* values, size_t len) {
  return guarded([&] {
    REQ(len == 0 || values, pk::PK_ERR_ARG, "intt_nr: null values");
    intt_nr(reinterpret_cast<Fr*>(values), len);
  });
}
int pkgpu_intt_nr_device(uint64_t* values, size_t len, void* stream) {
  return guarded([&] {
    if (len <= 1) return;
    REQ(values, pk::PK_ERR_ARG, "intt_nr: null values");
    pk::intt_nr_dev(pk::device_current(), reinterpret_cast<fr_t*>(values), len,
                    stream_or_thread(stream));
  });
}
int pkgpu_reverse_order(uint64_t* values, size_t len, size_t codeword_size) {
  return guarded([&] {
    REQ(codeword_size == 0 || values, pk::PK_ERR_ARG, "reverse_order: null values");
    reverse_order(reinterpret_cast<Fr*>(values), len, codeword_size);
  });
}
int pkgpu_reverse_order_device(uint64_t* values, size_t len, size_t codeword_size,
                               void* stream) {
  return guarded([&] {
    REQ(pk::is_pow2(codeword_size), pk::PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
    REQ(codeword_size <= len, pk::PK_ERR_SIZE, "reverse_order: codeword_size above the slice length");
    if (codeword_size <= 1) return;
    REQ(values, pk::PK_ERR_ARG, "reverse_order: null values");
    pk::reverse_order_dev(reinterpret_cast<fr_t*>(values), codeword_size,
                          stream_or_thread(stream));
  });
}

int pkgpu_rs_next_order(size_t size, size_t* order) {
  return guarded([&] {
    REQ(order, pk::PK_ERR_ARG, "next_order: null out-pointer");
    auto o = RSFr().next_order(size);
    REQ(o.has_value(), pk::PK_ERR_SIZE, "next_order: order above 2^28");
    *order = *o;
  });
}
int pkgpu_rs_generator(size_t codeword_length, uint64_t out[4]) {
  return guarded([&] {
    REQ(out, pk::PK_ERR_ARG, "generator: null out");
    Fr g = RSFr().generator(codeword_length);
    memcpy(out, g.data(), 32);
  });
}
int pkgpu_rs_evaluation_points(size_t masked_message_length, size_t codeword_length,
                               const size_t* indices, size_t n, uint64_t* out) {
  return guarded([&] {
    REQ(n == 0 || (indices && out), pk::PK_ERR_ARG, "evaluation_points: null pointer");
    std::vector<size_t> idx(indices, indices + n);
    auto pts = RSFr().evaluation_points(masked_message_length, codeword_length, idx);
    if (n) memcpy(out, pts.data(), 32 * n);
  });
}
int pkgpu_rs_interleaved_encode(const uint64_t* const* messages, size_t num_messages,
                                size_t message_length, const uint64_t* masks, size_t masks_len,
                                size_t codeword_length, uint64_t* out) {
  return guarded([&] {
    if (num_messages == 0) return;
    REQ(messages && out && (masks_len == 0 || masks), pk::PK_ERR_ARG,
        "interleaved_encode: null pointer");
    std::vector<const Fr*> msgs(num_messages);
    for (size_t r = 0; r < num_messages; r++) {
      REQ(message_length == 0 || messages[r], pk::PK_ERR_ARG, "interleaved_encode: null message");
      msgs[r] = reinterpret_cast<const Fr*>(messages[r]);
    }
    std::vector<Fr> mk(masks_len);
    if (masks_len) memcpy(mk.data(), masks, 32 * masks_len);
    auto res = RSFr().interleaved_encode(msgs, message_length, mk, codeword_length);
    memcpy(out, res.data(), 32 * res.size());
  });
}
int pkgpu_rs_interleaved_encode_device(const uint64_t* messages, size_t num_messages,
                                       size_t message_length, const uint64_t* masks,
                                       size_t masks_len, size_t codeword_length, int coset,
                                       uint64_t* out, void* stream) {
  return guarded([&] {
    if (num_messages == 0) return;
    REQ(messages && out && (masks_len == 0 || masks), pk::PK_ERR_ARG,
        "interleaved_encode: null pointer");
    REQ(num_messages * message_length + masks_len <= num_messages * codeword_length,
        pk::PK_ERR_SIZE, "interleaved_encode: messages and masks longer than the codeword");
    pk::rs_encode_dev(pk::device_current(), reinterpret_cast<const fr_t*>(messages), num_messages,
                      message_length, reinterpret_cast<const fr_t*>(masks), masks_len,
                      codeword_length, reinterpret_cast<fr_t*>(out), coset,
                      stream_or_thread(stream));
  });
}

int pkgpu_skyscraper_compress(const uint64_t l[4], const uint64_t r[4], uint64_t out[4]) {
  return guarded([&] {
    REQ(l && r && out, pk::PK_ERR_ARG, "compress: null pointer");
    Fr a, b;
    memcpy(a.data(), l, 32);
    memcpy(b.data(), r, 32);
    Fr o = skyscraper::compress(a, b);
    memcpy(out, o.data(), 32);
  });
}
int pkgpu_skyscraper_compress_many(const uint8_t* messages, size_t messages_len,
                                   uint8_t* hashes, size_t hashes_len) {
  return guarded([&] {
    REQ(hashes_len == 0 || (messages && hashes), pk::PK_ERR_ARG, "compress_many: null pointer");
    skyscraper::compress_many(messages, messages_len, hashes, hashes_len);
  });
}
int pkgpu_skyscraper_hash_many(size_t size, const uint8_t* input, size_t input_len,
                               uint8_t* output, size_t count) {
  return guarded([&] {
    REQ(count == 0 || (input && output), pk::PK_ERR_ARG, "hash_many: null pointer");
    SkyscraperHashEngine().hash_many(size, input, input_len, reinterpret_cast<Hash*>(output),
                                     count);
  });
}
int pkgpu_skyscraper_hash_many_device(size_t size, const void* input, size_t count, void* output,
                                      void* stream) {
  return guarded([&] {
    REQ(size > 0 && size % 32 == 0, pk::PK_ERR_SIZE,
        "skyscraper: unsupported message size " + std::to_string(size) +
            " (must be a positive multiple of 32)");
    if (!count) return;
    REQ(input && output, pk::PK_ERR_ARG, "hash_many: null pointer");
    pk::sky_hash_many_dev(size / 32, static_cast<const fr_t*>(input), count,
                          static_cast<fr_t*>(output), stream_or_thread(stream));
  });
}
int pkgpu_skyscraper_merkle_device(const void* leaves, size_t n, void* nodes, void* stream) {
  return guarded([&] {
    REQ(leaves && nodes, pk::PK_ERR_ARG, "merkle: null pointer");
    pk::sky_merkle_dev(static_cast<const fr_t*>(leaves), n, static_cast<fr_t*>(nodes),
                       stream_or_thread(stream));
  });
}

int pkgpu_skyscraper_pow_threshold(double difficulty, uint64_t out[4]) {
  return guarded([&] {
    REQ(out, pk::PK_ERR_ARG, "pow_threshold: null pointer");
    pk::pow_threshold(difficulty, out);
  });
}
int pkgpu_skyscraper_pow_verify(const uint64_t challenge[4], double difficulty, uint64_t nonce,
                                int* ok) {
  return guarded([&] {
    REQ(challenge && ok, pk::PK_ERR_ARG, "pow_verify: null pointer");
    Fr c;
    memcpy(c.data(), challenge, 32);
    *ok = skyscraper::pow::verify(c, difficulty, nonce) ? 1 : 0;
  });
}
int pkgpu_skyscraper_pow_solve(const uint64_t challenge[4], double difficulty,
                               uint64_t max_nonces, uint64_t* nonce, int* found, void* stream) {
  return guarded([&] {
    REQ(challenge && nonce && found, pk::PK_ERR_ARG, "pow_solve: null pointer");
    *found = pk::sky_pow_solve_dev(challenge, difficulty, max_nonces ? max_nonces : ~0ULL, nonce,
                                   stream_or_thread(stream))
                 ? 1
                 : 0;
  });
}

int pkgpu_merkle_open_device(const uint64_t* codeword, const void* nodes, size_t leaves,
                             size_t leaf_elems, const uint64_t* leaf_indices, size_t n,
                             uint64_t* values, uint8_t* paths, void* stream) {
  return guarded([&] {
    if (!n) return;
    REQ(codeword && nodes && leaf_indices && values && (leaves < 2 || paths), pk::PK_ERR_ARG,
        "merkle open: null pointer");
    REQ(pk::is_pow2(leaves), pk::PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
    REQ(leaf_elems > 0, pk::PK_ERR_SIZE, "merkle open: empty leaves");
    for (size_t q = 0; q < n; q++)
      REQ(leaf_indices[q] < leaves, pk::PK_ERR_ARG, "merkle open: leaf index out of range");
    hipStream_t s = stream_or_thread(stream);
    const int depth = pk::log2_exact(leaves);
    Staged di(8 * n, s), dv(sizeof(fr_t) * n * leaf_elems, s),
        dp(sizeof(fr_t) * n * (depth ? depth : 1), s);
    h2d(di.get<void>(), leaf_indices, 8 * n, s);
    pk::merkle_open_dev(reinterpret_cast<const fr_t*>(codeword), static_cast<const fr_t*>(nodes),
                        leaves, leaf_elems, di.get<uint64_t>(), n, dv.get<fr_t>(), dp.get<fr_t>(),
                        s);
    d2h(values, dv.get<void>(), sizeof(fr_t) * n * leaf_elems, s);
    if (depth) d2h(paths, dp.get<void>(), sizeof(fr_t) * n * depth, s);
    sync(s);
  });
}

int pkgpu_poseidon2_permutation_many(const uint64_t* states, size_t n, uint64_t* out) {
  return guarded([&] {
    if (!n) return;
    REQ(states && out, pk::PK_ERR_ARG, "poseidon2_permutation: null pointer");
    hipStream_t s = pk::thread_stream();
    Staged di(sizeof(fr_t) * 4 * n, s), dout(sizeof(fr_t) * 4 * n, s);
    h2d(di.get<void>(), states, sizeof(fr_t) * 4 * n, s);
    pk::p2_permutation_many_dev(pk::device_current(), di.get<fr_t>(), dout.get<fr_t>(), n, s);
    d2h(out, dout.get<void>(), sizeof(fr_t) * 4 * n, s);
    sync(s);
  });
}
int pkgpu_poseidon2_hash_many(const uint64_t* inputs, size_t len, size_t count, uint64_t* out) {
  return guarded([&] {
    if (!count) return;
    REQ(out && (len == 0 || inputs), pk::PK_ERR_ARG, "poseidon2_hash: null pointer");
    hipStream_t s = pk::thread_stream();
    Staged di(sizeof(fr_t) * len * count, s), dout(sizeof(fr_t) * count, s);
    h2d(di.get<void>(), inputs, sizeof(fr_t) * len * count, s);
    pk::p2_hash_many_dev(pk::device_current(), di.get<fr_t>(), len, count, dout.get<fr_t>(), s);
    d2h(out, dout.get<void>(), sizeof(fr_t) * count, s);
    sync(s);
  });
}
int pkgpu_poseidon2_hash(const uint64_t* inputs, size_t len, uint64_t out[4]) {
  return pkgpu_poseidon2_hash_many(inputs, len, 1, out);
}
int pkgpu_poseidon2_hash_many_device(const uint64_t* inputs, size_t len, size_t count,
                                     uint64_t* out, void* stream) {
  return guarded([&] {
    if (!count) return;
    REQ(out && (len == 0 || inputs), pk::PK_ERR_ARG, "poseidon2_hash: null pointer");
    pk::p2_hash_many_dev(pk::device_current(), reinterpret_cast<const fr_t*>(inputs), len, count,
                         reinterpret_cast<fr_t*>(out), stream_or_thread(stream));
  });
}
int pkgpu_poseidon2_merkle(const uint64_t* leaves, size_t n, uint64_t* nodes, uint64_t root[4]) {
  return guarded([&] {
    REQ(leaves && root, pk::PK_ERR_ARG, "poseidon2_merkle: null pointer");
    REQ(pk::is_pow2(n), pk::PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
    hipStream_t s = pk::thread_stream();
    const size_t total = 2 * n - 1;
    Staged dn(sizeof(fr_t) * total, s);
    h2d(dn.get<void>(), leaves, sizeof(fr_t) * n, s);
    pk::p2_merkle_dev(pk::device_current(), dn.get<fr_t>(), n, dn.get<fr_t>(), s);
    d2h(root, dn.get<fr_t>() + (total - 1), 32, s);
    if (nodes) d2h(nodes, dn.get<void>(), sizeof(fr_t) * total, s);
    sync(s);
  });
}
int pkgpu_poseidon2_merkle_device(const uint64_t* leaves, size_t n, uint64_t* nodes,
                                  void* stream) {
  return guarded([&] {
    REQ(leaves && nodes, pk::PK_ERR_ARG, "poseidon2_merkle: null pointer");
    pk::p2_merkle_dev(pk::device_current(), reinterpret_cast<const fr_t*>(leaves), n,
                      reinterpret_cast<fr_t*>(nodes), stream_or_thread(stream));
  });
}
int pkgpu_poseidon2_merkle_graph(const uint64_t* leaves, size_t n, uint64_t* nodes,
                                 void** graph) {
  return guarded_capture([&] {
    REQ(leaves && nodes && graph, pk::PK_ERR_ARG, "poseidon2_merkle_graph: null pointer");
    REQ(pk::is_pow2(n), pk::PK_ERR_NOT_POW2, "merkle: leaf count must be a power of two");
    *graph = pk::p2_merkle_graph(pk::device_current(), reinterpret_cast<const fr_t*>(leaves), n,
                                 reinterpret_cast<fr_t*>(nodes));
  });
}
int pkgpu_graph_launch(void* graph, void* stream) {
  return guarded([&] {
    REQ(graph, pk::PK_ERR_STATE, "graph_launch: null graph");
    PK_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), stream_or_thread(stream)));
  });
}
int pkgpu_graph_destroy(void* graph) {
  return guarded([&] {
    if (graph) PK_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
  });
}

namespace {
pk::CsrDev csr_view(size_t rows, size_t cols_n, const uint32_t* rs, size_t nnz, const uint32_t* cols,
                    const uint64_t* vals) {
  pk::CsrDev A;
  A.rows = rows;
  A.ncols = cols_n;
  A.nnz = nnz;
  A.row_starts = rs;
  A.cols = cols;
  A.vals = reinterpret_cast<const fr_t*>(vals);
  return A;
}
// Upload one host CSR matrix; the returned view points into `bufs`.
pk::CsrDev upload_csr(size_t rows, size_t cols_n, const uint32_t* rs, size_t nnz,
                      const uint32_t* cols, const uint64_t* vals, std::vector<pk::Scratch>& bufs,
                      hipStream_t s) {
  bufs.emplace_back(sizeof(uint32_t) * rows, s);
  h2d(bufs.back().p, rs, sizeof(uint32_t) * rows, s);
  const uint32_t* d_rs = bufs.back().as<uint32_t>();
  bufs.emplace_back(sizeof(uint32_t) * nnz, s);
  h2d(bufs.back().p, cols, sizeof(uint32_t) * nnz, s);
  const uint32_t* d_cols = bufs.back().as<uint32_t>();
  bufs.emplace_back(sizeof(fr_t) * nnz, s);
  h2d(bufs.back().p, vals, sizeof(fr_t) * nnz, s);
  return csr_view(rows, cols_n, d_rs, nnz, d_cols, bufs.back().as<uint64_t>());
}
void check_csr(size_t rows, const uint32_t* rs, size_t nnz, const uint32_t* cols,
               const uint64_t* vals) {
  REQ(rows == 0 || rs, pk::PK_ERR_ARG, "sparse matrix: null row starts");
  REQ(nnz == 0 || (cols && vals), pk::PK_ERR_ARG, "sparse matrix: null entries");
}
// Host-side CSR well-formedness: non-decreasing row starts within nnz, column
// indices below num_cols (the device kernel trusts both).
void validate_csr(size_t rows, size_t cols_n, const uint32_t* rs, size_t nnz,
                  const uint32_t* cols) {
  uint64_t prev = 0;
  for (size_t r = 0; r < rows; r++) {
    REQ(rs[r] >= prev && rs[r] <= nnz, pk::PK_ERR_ARG, "sparse matrix: bad row starts");
    prev = rs[r];
  }
  for (size_t i = 0; i < nnz; i++)
    REQ(cols[i] < cols_n, pk::PK_ERR_ARG, "sparse matrix: column index out of range");
}
}  // namespace

int pkgpu_sparse_mul(size_t num_rows, size_t num_cols, const uint32_t* row_starts, size_t nnz,
                     const uint32_t* cols, const uint64_t* vals, const uint64_t* z, size_t z_len,
                     uint64_t* out) {
  return guarded([&] {
    REQ(num_cols == z_len, pk::PK_ERR_SIZE, "Vector length does not match number of columns.");
    check_csr(num_rows, row_starts, nnz, cols, vals);
    REQ(num_rows == 0 || out, pk::PK_ERR_ARG, "sparse_mul: null output");
    REQ(z_len == 0 || z, pk::PK_ERR_ARG, "sparse_mul: null vector");
    validate_csr(num_rows, num_cols, row_starts, nnz, cols);
    if (!num_rows) return;
    hipStream_t s = pk::thread_stream();
    std::vector<pk::Scratch> bufs;
    pk::CsrDev A = upload_csr(num_rows, num_cols, row_starts, nnz, cols, vals, bufs, s);
    Staged dz(sizeof(fr_t) * z_len, s), dout(sizeof(fr_t) * num_rows, s);
    h2d(dz.get<void>(), z, sizeof(fr_t) * z_len, s);
    pk::spmv_dev(A, dz.get<fr_t>(), dout.get<fr_t>(), s);
    d2h(out, dout.get<void>(), sizeof(fr_t) * num_rows, s);
    sync(s);
  });
}
int pkgpu_sparse_mul_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                            size_t nnz, const uint32_t* cols, const uint64_t* vals,
                            const uint64_t* z, uint64_t* out, void* stream) {
  return guarded([&] {
    check_csr(num_rows, row_starts, nnz, cols, vals);
    REQ(num_rows == 0 || (out && (num_cols == 0 || z)), pk::PK_ERR_ARG, "sparse_mul: null pointer");
    pk::spmv_dev(csr_view(num_rows, num_cols, row_starts, nnz, cols, vals),
                 reinterpret_cast<const fr_t*>(z), reinterpret_cast<fr_t*>(out),
                 stream_or_thread(stream));
  });
}
int pkgpu_sparse_mul_interned(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                              size_t nnz, const uint32_t* cols, const uint32_t* ids,
                              const uint64_t* interner, size_t interner_len, const uint64_t* z,
                              size_t z_len, uint64_t* out) {
  return guarded([&] {
    REQ(num_cols == z_len, pk::PK_ERR_SIZE, "Vector length does not match number of columns.");
    REQ(num_rows == 0 || row_starts, pk::PK_ERR_ARG, "sparse matrix: null row starts");
    REQ(nnz == 0 || (cols && ids), pk::PK_ERR_ARG, "sparse matrix: null entries");
    REQ(interner_len == 0 || interner, pk::PK_ERR_ARG, "sparse_mul: null interner");
    REQ(num_rows == 0 || out, pk::PK_ERR_ARG, "sparse_mul: null output");
    REQ(z_len == 0 || z, pk::PK_ERR_ARG, "sparse_mul: null vector");
    validate_csr(num_rows, num_cols, row_starts, nnz, cols);
    for (size_t i = 0; i < nnz; i++)
      REQ(ids[i] < interner_len, pk::PK_ERR_ARG, "Value not in interner.");  // iter_row's expect
    if (!num_rows) return;
    hipStream_t s = pk::thread_stream();
    Staged drs(4 * num_rows, s), dc(4 * std::max<size_t>(nnz, 1), s),
        di(4 * std::max<size_t>(nnz, 1), s), dt(sizeof(fr_t) * std::max<size_t>(interner_len, 1), s),
        dz(sizeof(fr_t) * std::max<size_t>(z_len, 1), s), dout(sizeof(fr_t) * num_rows, s);
    h2d(drs.get<void>(), row_starts, 4 * num_rows, s);
    h2d(dc.get<void>(), cols, 4 * nnz, s);
    h2d(di.get<void>(), ids, 4 * nnz, s);
    h2d(dt.get<void>(), interner, sizeof(fr_t) * interner_len, s);
    h2d(dz.get<void>(), z, sizeof(fr_t) * z_len, s);
    pk::CsrDev A = csr_view(num_rows, num_cols, drs.get<uint32_t>(), nnz, dc.get<uint32_t>(), nullptr);
    A.ids = di.get<uint32_t>();
    A.interner = dt.get<fr_t>();
    A.interner_len = interner_len;
    pk::spmv_dev(A, dz.get<fr_t>(), dout.get<fr_t>(), s);
    d2h(out, dout.get<void>(), sizeof(fr_t) * num_rows, s);
    sync(s);
  });
}
int pkgpu_sparse_mul_interned_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                     size_t nnz, const uint32_t* cols, const uint32_t* ids,
                                     const uint64_t* interner, size_t interner_len,
                                     const uint64_t* z, uint64_t* out, void* stream) {
  return guarded([&] {
    REQ(num_rows == 0 || row_starts, pk::PK_ERR_ARG, "sparse matrix: null row starts");
    REQ(nnz == 0 || (cols && ids), pk::PK_ERR_ARG, "sparse matrix: null entries");
    REQ(interner_len == 0 || interner, pk::PK_ERR_ARG, "sparse_mul: null interner");
    REQ(num_rows == 0 || (out && (num_cols == 0 || z)), pk::PK_ERR_ARG, "sparse_mul: null pointer");
    pk::CsrDev A = csr_view(num_rows, num_cols, row_starts, nnz, cols, nullptr);
    A.ids = ids;
    A.interner = reinterpret_cast<const fr_t*>(interner);
    A.interner_len = interner_len;
    pk::spmv_dev(A, reinterpret_cast<const fr_t*>(z), reinterpret_cast<fr_t*>(out),
                 stream_or_thread(stream));
  });
}
int pkgpu_sparse_mul_device_noalloc(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                    size_t nnz, const uint32_t* cols, const uint64_t* vals,
                                    const uint64_t* z, uint64_t* out, void* stream) {
  return guarded([&] {
    check_csr(num_rows, row_starts, nnz, cols, vals);
    REQ(num_rows == 0 || (out && (num_cols == 0 || z)), pk::PK_ERR_ARG, "sparse_mul: null pointer");
    REQ(nnz <= PKGPU_SPMV_NOALLOC_MAX_NNZ, pk::PK_ERR_SIZE,
        "sparse_mul_noalloc: more than 2^26 entries (the fixed scratch holds 2^26)");
    // NULL: the device's null stream (no per-thread stream is created here)
    pk::spmv_dev_noalloc(csr_view(num_rows, num_cols, row_starts, nnz, cols, vals),
                         reinterpret_cast<const fr_t*>(z), reinterpret_cast<fr_t*>(out),
                         static_cast<hipStream_t>(stream));
  });
}
int pkgpu_sparse_transpose(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                           size_t nnz, const uint32_t* cols, const uint64_t* vals,
                           uint32_t* t_row_starts, uint32_t* t_cols, uint64_t* t_vals) {
  return guarded([&] {
    check_csr(num_rows, row_starts, nnz, cols, vals);
    REQ((num_cols == 0 || t_row_starts) && (nnz == 0 || (t_cols && t_vals)), pk::PK_ERR_ARG,
        "transpose: null output");
    validate_csr(num_rows, num_cols, row_starts, nnz, cols);
    if (!num_cols) return;
    hipStream_t s = pk::thread_stream();
    std::vector<pk::Scratch> bufs;
    pk::CsrDev A = upload_csr(num_rows, num_cols, row_starts, nnz, cols, vals, bufs, s);
    Staged trs(sizeof(uint32_t) * num_cols, s), tcs(sizeof(uint32_t) * (nnz ? nnz : 1), s),
        tvs(sizeof(fr_t) * (nnz ? nnz : 1), s);
    pk::sparse_transpose_dev(A, trs.get<uint32_t>(), tcs.get<uint32_t>(), tvs.get<fr_t>(), s);
    d2h(t_row_starts, trs.get<void>(), sizeof(uint32_t) * num_cols, s);
    d2h(t_cols, tcs.get<void>(), sizeof(uint32_t) * nnz, s);
    d2h(t_vals, tvs.get<void>(), sizeof(fr_t) * nnz, s);
    sync(s);
  });
}
int pkgpu_sparse_transpose_device(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                  size_t nnz, const uint32_t* cols, const uint64_t* vals,
                                  uint32_t* t_row_starts, uint32_t* t_cols, uint64_t* t_vals,
                                  void* stream) {
  return guarded([&] {
    check_csr(num_rows, row_starts, nnz, cols, vals);
    REQ((num_cols == 0 || t_row_starts) && (nnz == 0 || (t_cols && t_vals)), pk::PK_ERR_ARG,
        "transpose: null output");
    pk::sparse_transpose_dev(csr_view(num_rows, num_cols, row_starts, nnz, cols, vals),
                             t_row_starts, t_cols, reinterpret_cast<fr_t*>(t_vals),
                             stream_or_thread(stream));
  });
}
int pkgpu_sparse_transpose_interned(size_t num_rows, size_t num_cols, const uint32_t* row_starts,
                                    size_t nnz, const uint32_t* cols, const uint32_t* ids,
                                    uint32_t* t_row_starts, uint32_t* t_cols, uint32_t* t_ids) {
  return guarded([&] {
    REQ(num_rows == 0 || row_starts, pk::PK_ERR_ARG, "sparse matrix: null row starts");
    REQ(nnz == 0 || (cols && ids), pk::PK_ERR_ARG, "sparse matrix: null entries");
    REQ((num_cols == 0 || t_row_starts) && (nnz == 0 || (t_cols && t_ids)), pk::PK_ERR_ARG,
        "transpose: null output");
    validate_csr(num_rows, num_cols, row_starts, nnz, cols);
    if (!num_cols) return;
    hipStream_t s = pk::thread_stream();
    Staged drs(4 * std::max<size_t>(num_rows, 1), s), dc(4 * std::max<size_t>(nnz, 1), s),
        di(4 * std::max<size_t>(nnz, 1), s), trs(4 * num_cols, s),
        tcs(4 * std::max<size_t>(nnz, 1), s), tis(4 * std::max<size_t>(nnz, 1), s);
    h2d(drs.get<void>(), row_starts, 4 * num_rows, s);
    h2d(dc.get<void>(), cols, 4 * nnz, s);
    h2d(di.get<void>(), ids, 4 * nnz, s);
    pk::CsrDev A = csr_view(num_rows, num_cols, drs.get<uint32_t>(), nnz, dc.get<uint32_t>(), nullptr);
    A.ids = di.get<uint32_t>();
    pk::sparse_transpose_interned_dev(A, trs.get<uint32_t>(), tcs.get<uint32_t>(),
                                      tis.get<uint32_t>(), s);
    d2h(t_row_starts, trs.get<void>(), 4 * num_cols, s);
    d2h(t_cols, tcs.get<void>(), 4 * nnz, s);
    d2h(t_ids, tis.get<void>(), 4 * nnz, s);
    sync(s);
  });
}
int pkgpu_sparse_transpose_interned_device(size_t num_rows, size_t num_cols,
                                           const uint32_t* row_starts, size_t nnz,
                                           const uint32_t* cols, const uint32_t* ids,
                                           uint32_t* t_row_starts, uint32_t* t_cols,
                                           uint32_t* t_ids, void* stream) {
  return guarded([&] {
    REQ(num_rows == 0 || row_starts, pk::PK_ERR_ARG, "sparse matrix: null row starts");
    REQ(nnz == 0 || (cols && ids), pk::PK_ERR_ARG, "sparse matrix: null entries");
    REQ((num_cols == 0 || t_row_starts) && (nnz == 0 || (t_cols && t_ids)), pk::PK_ERR_ARG,
        "transpose: null output");
    pk::CsrDev A = csr_view(num_rows, num_cols, row_starts, nnz, cols, nullptr);
    A.ids = ids;
    pk::sparse_transpose_interned_dev(A, t_row_starts, t_cols, t_ids, stream_or_thread(stream));
  });
}
int pkgpu_multiply_transposed_by_eq_alpha_device(size_t t_num_rows, size_t t_num_cols,
                                                 const uint32_t* t_row_starts, size_t nnz,
                                                 const uint32_t* t_cols, const uint64_t* t_vals,
                                                 const uint64_t* alpha, size_t m, uint64_t* out,
                                                 void* stream) {
  return guarded([&] {
    check_csr(t_num_rows, t_row_starts, nnz, t_cols, t_vals);
    REQ(m < 63 && t_num_cols <= (size_t(1) << m), pk::PK_ERR_SIZE,
        "assertion failed: num_entries <= full_size");
    REQ(t_num_rows == 0 || (out && (m == 0 || alpha)), pk::PK_ERR_ARG,
        "transposed eq product: null pointer");
    pk::transposed_eq_mul_dev(csr_view(t_num_rows, t_num_cols, t_row_starts, nnz, t_cols, t_vals),
                              reinterpret_cast<const fr_t*>(alpha), (int)m,
                              reinterpret_cast<fr_t*>(out), stream_or_thread(stream));
  });
}
int pkgpu_multiply_transposed_by_eq_alpha_interned_device(
    size_t t_num_rows, size_t t_num_cols, const uint32_t* t_row_starts, size_t nnz,
    const uint32_t* t_cols, const uint32_t* t_ids, const uint64_t* interner, size_t interner_len,
    const uint64_t* alpha, size_t m, uint64_t* out, void* stream) {
  return guarded([&] {
    REQ(t_num_rows == 0 || t_row_starts, pk::PK_ERR_ARG, "sparse matrix: null row starts");
    REQ(nnz == 0 || (t_cols && t_ids), pk::PK_ERR_ARG, "sparse matrix: null entries");
    REQ(interner_len == 0 || interner, pk::PK_ERR_ARG, "transposed eq product: null interner");
    REQ(m < 63 && t_num_cols <= (size_t(1) << m), pk::PK_ERR_SIZE,
        "assertion failed: num_entries <= full_size");
    REQ(t_num_rows == 0 || (out && (m == 0 || alpha)), pk::PK_ERR_ARG,
        "transposed eq product: null pointer");
    pk::CsrDev At = csr_view(t_num_rows, t_num_cols, t_row_starts, nnz, t_cols, nullptr);
    At.ids = t_ids;
    At.interner = reinterpret_cast<const fr_t*>(interner);
    At.interner_len = interner_len;
    pk::transposed_eq_mul_dev(At, reinterpret_cast<const fr_t*>(alpha), (int)m,
                              reinterpret_cast<fr_t*>(out), stream_or_thread(stream));
  });
}
int pkgpu_interner_decode_device(const uint8_t* blob, size_t len, uint64_t* out, size_t* count,
                                 void* stream) {
  return guarded([&] {
    REQ(count && (blob || !len), pk::PK_ERR_ARG, "interner decode: null pointer");
    *count = pk::interner_decode_dev(blob, len, reinterpret_cast<fr_t*>(out),
                                     stream_or_thread(stream));
  });
}
int pkgpu_sparse_matrix_decode_device(const uint8_t* blob, size_t len, const uint64_t* interner,
                                      size_t interner_len, pkgpu_csr* m, void* stream) {
  return guarded([&] {
    REQ(m && (blob || !len), pk::PK_ERR_ARG, "sparse matrix decode: null pointer");
    const bool fill = m->row_starts || m->cols || m->vals;
    REQ(!fill || interner_len == 0 || interner, pk::PK_ERR_ARG,
        "sparse matrix decode: null interner");
    pk::WireCsrOut o{m->row_starts, m->cols, reinterpret_cast<fr_t*>(m->vals)};
    const pk::WireCsrSizes z = pk::sparse_matrix_decode_dev(
        blob, len, reinterpret_cast<const fr_t*>(interner), interner_len, fill ? &o : nullptr,
        stream_or_thread(stream));
    m->num_rows = z.num_rows;
    m->num_cols = z.num_cols;
    m->nnz = z.nnz;
  });
}
int pkgpu_r1cs_decode_device(const uint8_t* blob, size_t len, size_t* num_public_inputs,
                             uint64_t* interner, size_t* interner_len, pkgpu_csr mats[3],
                             void* stream) {
  return guarded([&] {
    REQ(num_public_inputs && interner_len && mats && (blob || !len), pk::PK_ERR_ARG,
        "r1cs decode: null pointer");
    pk::WireCsrOut outs[3];
    for (int i = 0; i < 3; i++)
      outs[i] = pk::WireCsrOut{mats[i].row_starts, mats[i].cols,
                               reinterpret_cast<fr_t*>(mats[i].vals)};
    const pk::R1csSizes z = pk::r1cs_decode_dev(blob, len, reinterpret_cast<fr_t*>(interner),
                                                interner ? outs : nullptr,
                                                stream_or_thread(stream));
    *num_public_inputs = z.num_public_inputs;
    *interner_len = z.interner_len;
    for (int i = 0; i < 3; i++) {
      mats[i].num_rows = z.mats[i].num_rows;
      mats[i].num_cols = z.mats[i].num_cols;
      mats[i].nnz = z.mats[i].nnz;
    }
  });
}
int pkgpu_calculate_witness_bounds(size_t num_rows, size_t num_cols, const uint32_t* a_rs,
                                   size_t a_nnz, const uint32_t* a_cols, const uint64_t* a_vals,
                                   const uint32_t* b_rs, size_t b_nnz, const uint32_t* b_cols,
                                   const uint64_t* b_vals, const uint64_t* w, size_t w_len,
                                   uint64_t* a_out, uint64_t* b_out, uint64_t* c_out,
                                   size_t* padded_len) {
  return guarded([&] {
    REQ(num_cols == w_len, pk::PK_ERR_SIZE, "Vector length does not match number of columns.");
    check_csr(num_rows, a_rs, a_nnz, a_cols, a_vals);
    check_csr(num_rows, b_rs, b_nnz, b_cols, b_vals);
    REQ(padded_len && a_out && b_out && c_out && (w_len == 0 || w), pk::PK_ERR_ARG,
        "calculate_witness_bounds: null pointer");
    validate_csr(num_rows, num_cols, a_rs, a_nnz, a_cols);
    validate_csr(num_rows, num_cols, b_rs, b_nnz, b_cols);
    size_t n = 1;
    while (n < num_rows) n <<= 1;
    if (num_rows == 0) n = 0;
    *padded_len = n;
    if (!n) return;
    hipStream_t s = pk::thread_stream();
    std::vector<pk::Scratch> bufs;
    pk::CsrDev A = upload_csr(num_rows, num_cols, a_rs, a_nnz, a_cols, a_vals, bufs, s);
    pk::CsrDev B = upload_csr(num_rows, num_cols, b_rs, b_nnz, b_cols, b_vals, bufs, s);
    Staged dw(sizeof(fr_t) * w_len, s), da(sizeof(fr_t) * n, s), db(sizeof(fr_t) * n, s),
        dc(sizeof(fr_t) * n, s);
    h2d(dw.get<void>(), w, sizeof(fr_t) * w_len, s);
    pk::zero_dev(da.get<fr_t>() + num_rows, n - num_rows, s);
    pk::zero_dev(db.get<fr_t>() + num_rows, n - num_rows, s);
    pk::spmv_dev(A, dw.get<fr_t>(), da.get<fr_t>(), s);
    pk::spmv_dev(B, dw.get<fr_t>(), db.get<fr_t>(), s);
    pk::hadamard_dev(da.get<fr_t>(), db.get<fr_t>(), dc.get<fr_t>(), num_rows, n, s);
    d2h(a_out, da.get<void>(), sizeof(fr_t) * n, s);
    d2h(b_out, db.get<void>(), sizeof(fr_t) * n, s);
    d2h(c_out, dc.get<void>(), sizeof(fr_t) * n, s);
    sync(s);
  });
}
int pkgpu_eq_evaluations(const uint64_t* r, size_t m, size_t num_entries, uint64_t* out) {
  return guarded([&] {
    if (!num_entries) return;
    REQ(m < 63 && num_entries <= (size_t(1) << m), pk::PK_ERR_SIZE,
        "assertion failed: num_entries <= full_size");
    REQ(out && (m == 0 || r), pk::PK_ERR_ARG, "eq: null pointer");
    hipStream_t s = pk::thread_stream();
    Staged dr(sizeof(fr_t) * (m ? m : 1), s), dout(sizeof(fr_t) * num_entries, s);
    h2d(dr.get<void>(), r, sizeof(fr_t) * m, s);
    pk::eq_table_dev(dr.get<fr_t>(), (int)m, num_entries, dout.get<fr_t>(), s);
    d2h(out, dout.get<void>(), sizeof(fr_t) * num_entries, s);
    sync(s);
  });
}
int pkgpu_eq_evaluations_device(const uint64_t* r, size_t m, size_t num_entries, uint64_t* out,
                                void* stream) {
  return guarded([&] {
    if (!num_entries) return;
    REQ(m < 63 && num_entries <= (size_t(1) << m), pk::PK_ERR_SIZE,
        "assertion failed: num_entries <= full_size");
    REQ(out && (m == 0 || r), pk::PK_ERR_ARG, "eq: null pointer");
    pk::eq_table_dev(reinterpret_cast<const fr_t*>(r), (int)m, num_entries,
                     reinterpret_cast<fr_t*>(out), stream_or_thread(stream));
  });
}
int pkgpu_sumcheck_r1cs_round_device(uint64_t* a, uint64_t* b, uint64_t* c, uint64_t* eq,
                                     size_t size, const uint64_t* fold, uint64_t* out,
                                     void* stream) {
  return guarded([&] {
    REQ(a && b && c && eq && out, pk::PK_ERR_ARG, "sumcheck round: null pointer");
    REQ(pk::is_pow2(size), pk::PK_ERR_NOT_POW2, "assertion failed: size.is_power_of_two()");
    REQ(size >= 2, pk::PK_ERR_SIZE, "assertion failed: size >= 2");
    REQ(!fold || size >= 4, pk::PK_ERR_SIZE, "assertion failed: size >= 4");
    hipStream_t s = stream_or_thread(stream);
    Staged work(sizeof(fr_t) * (pk::sc_partials_elems() + 4), s);
    fr_t* partials = work.get<fr_t>();
    fr_t* out3 = partials + pk::sc_partials_elems();
    fr_t* dfold = nullptr;
    if (fold) {
      dfold = out3 + 3;
      h2d(dfold, fold, 32, s);
    }
    pk::sc_round_dev(reinterpret_cast<fr_t*>(a), reinterpret_cast<fr_t*>(b),
                     reinterpret_cast<fr_t*>(c), reinterpret_cast<fr_t*>(eq), size, dfold,
                     partials, out3, s);
    d2h(out, out3, sizeof(fr_t) * 3, s);
    sync(s);
  });
}
namespace {
// The calling thread's pinned copy of the sumcheck state.  A copy from
// pageable memory is staged by the runtime and can hold the host until the
// stream reaches it, so the launches behind it start late; from pinned memory
// the copies stay asynchronous.  Each call syncs before returning, so the
// next call on the thread finds it free.  It is kept until the process ends
// (freeing it from a thread-exit destructor could run after the runtime).
pk::ScState* pinned_state() {
  thread_local pk::ScState* p = nullptr;
  if (!p) PK_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(pk::ScState), hipHostMallocDefault));
  return p;
}
}  // namespace

int pkgpu_sumcheck_prove_device(uint64_t* a, uint64_t* b, uint64_t* c, size_t m,
                                const uint64_t* blinding, uint64_t seed, uint64_t* coeffs,
                                uint64_t* alphas, uint64_t* r, uint64_t* final_claim,
                                uint64_t* rho, uint64_t* blinding_eval, void* stream) {
  return guarded([&] {
    REQ(a && b && c && final_claim && (m == 0 || (coeffs && alphas && r)), pk::PK_ERR_ARG,
        "sumcheck: null pointer");
    REQ(m <= (size_t)pk::kMaxSumcheckVars, pk::PK_ERR_SIZE, "sumcheck: more than 28 variables");
    // compute_blinding_coefficients_for_round's compute_for = n - 1 underflows
    // for an empty blinding polynomial (whir_r1cs.rs:437-439)
    REQ(!(blinding && m == 0), pk::PK_ERR_SIZE, "attempt to subtract with overflow");
    hipStream_t s = stream_or_thread(stream);
    const size_t size = m ? (size_t(1) << m) : 2;
    pk::ScState& hst = *pinned_state();
    pk::sc_state_init_host(hst, (int)m, blinding ? 1 : 0, reinterpret_cast<const fr_t*>(blinding));
    Staged dst(sizeof(pk::ScState), s), deq(sizeof(fr_t) * size, s),
        work(sizeof(fr_t) * (pk::sc_partials_elems() + 3), s),
        pad(m ? 0 : 6 * sizeof(fr_t), s);
    fr_t* da = reinterpret_cast<fr_t*>(a);
    fr_t* db = reinterpret_cast<fr_t*>(b);
    fr_t* dc = reinterpret_cast<fr_t*>(c);
    if (m == 0) {
      // pad_to_pow2_len_min2: the caller's MLEs hold one element each; the
      // rounds read two, so run on zero-padded copies
      fr_t* p = pad.get<fr_t>();
      pk::zero_dev(p, 6, s);
      fr_t* src[3] = {da, db, dc};
      for (int k = 0; k < 3; k++)
        PK_HIP(hipMemcpyAsync(p + 2 * k, src[k], sizeof(fr_t), hipMemcpyDeviceToDevice, s));
      da = p;
      db = p + 2;
      dc = p + 4;
    }
    h2d(dst.get<void>(), &hst, sizeof hst, s);
    pk::sc_prove_dev(dst.get<pk::ScState>(), da, db, dc, deq.get<fr_t>(), (int)m,
                     blinding ? 1 : 0, seed, work.get<fr_t>(), s);
    d2h(&hst, dst.get<void>(), sizeof hst, s);
    sync(s);
    if (m) {
      memcpy(coeffs, hst.coeffs, sizeof(fr_t) * 4 * m);
      memcpy(alphas, hst.alpha, sizeof(fr_t) * m);
      memcpy(r, hst.r, sizeof(fr_t) * m);
    }
    memcpy(final_claim, &hst.S, 32);
    if (rho) memcpy(rho, &hst.rho, 32);
    if (blinding_eval) memcpy(blinding_eval, &hst.blinding_eval, 32);
    REQ(!hst.error, pk::PK_ERR_INTERNAL,
        "assertion `left == right` failed: sumcheck round identity c0 + c0 + c1 + c2 + c3 = S");
  });
}

int pkgpu_commit(const uint64_t* coeffs, size_t n, size_t blowup, int coset, size_t leaf_elems,
                 uint64_t* codeword, uint8_t* nodes, uint8_t root[32]) {
  return guarded([&] {
    REQ(coeffs && root, pk::PK_ERR_ARG, "commit: null pointer");
    Hash h = commit(reinterpret_cast<const Fr*>(coeffs), n, blowup, coset != 0, leaf_elems,
                    reinterpret_cast<Fr*>(codeword), reinterpret_cast<Hash*>(nodes));
    memcpy(root, h.data(), 32);
  });
}
int pkgpu_commit_device(const uint64_t* coeffs, size_t n, size_t blowup, int coset,
                        size_t leaf_elems, uint64_t* codeword, void* nodes, void* stream) {
  return guarded([&] {
    REQ(coeffs && codeword && nodes, pk::PK_ERR_ARG, "commit: null pointer");
    pk::commit_dev(pk::device_current(), reinterpret_cast<const fr_t*>(coeffs), n, blowup, coset,
                   leaf_elems, reinterpret_cast<fr_t*>(codeword), static_cast<fr_t*>(nodes),
                   stream_or_thread(stream));
  });
}

int pkgpu_whir_r1cs_scheme_size(size_t num_witnesses, size_t num_constraints, size_t w1_size,
                                size_t* m, size_t* m0) {
  return guarded([&] {
    REQ(m && m0, pk::PK_ERR_ARG, "scheme size: null pointer");
    REQ(w1_size <= num_witnesses, pk::PK_ERR_SIZE, "w1_size exceeds total witnesses");
    auto clog = [](size_t n) {
      size_t k = 0;
      while ((size_t(1) << k) < n) k++;
      return k;
    };
    size_t mr = std::max<size_t>({clog(w1_size), clog(num_witnesses - w1_size), 13});
    const size_t z = std::max<size_t>(clog(num_constraints), 1);
    if ((size_t(1) << mr) - w1_size < 4 * z) mr++;
    *m = mr;
    *m0 = z;
  });
}

static void whir_commit_args(const uint64_t* witness, size_t witness_size, size_t num_vars,
                             const uint64_t* blinding, size_t m0) {
  REQ(witness || witness_size == 0, pk::PK_ERR_ARG, "commit: null witness");
  REQ(!blinding || m0 > 0, pk::PK_ERR_ARG, "commit: blinding without univariates");
  REQ(num_vars <= 28, pk::PK_ERR_SIZE, "commit: more than 2^28 coefficients");
  REQ(witness_size <= (size_t(1) << num_vars), pk::PK_ERR_SIZE,
      "R1CS witness length exceeds scheme capacity");
  REQ(!blinding || witness_size + 4 * m0 <= (size_t(1) << num_vars), pk::PK_ERR_SIZE,
      "index out of bounds: the blinding coefficients run past the padded witness");
}

int pkgpu_whir_r1cs_commit(const uint64_t* witness, size_t witness_size, size_t num_vars,
                           const uint64_t* blinding, size_t m0, size_t blowup, int coset,
                           size_t leaf_elems, uint64_t* polynomial, uint64_t* codeword,
                           uint8_t* nodes, uint8_t root[32]) {
  return guarded([&] {
    REQ(root, pk::PK_ERR_ARG, "commit: null pointer");
    whir_commit_args(witness, witness_size, num_vars, blinding, m0);
    const size_t n = size_t(1) << num_vars, N = n * blowup;
    REQ(pk::is_pow2(blowup) && leaf_elems >= 1 && pk::is_pow2(leaf_elems) && N % leaf_elems == 0,
        pk::PK_ERR_SIZE, "commit: leaf size must be a power of two dividing the codeword");
    const size_t leaves = N / leaf_elems, nn = 2 * leaves - 1;
    const size_t ng = blinding ? 4 * m0 : 0;
    hipStream_t s = pk::thread_stream();
    Staged dw(sizeof(fr_t) * std::max<size_t>(witness_size, 1), s), dg(sizeof(fr_t) * (ng + 1), s),
        dp(sizeof(fr_t) * n, s), dc(sizeof(fr_t) * N, s), dn(sizeof(fr_t) * nn, s);
    h2d(dw.get<void>(), witness, sizeof(fr_t) * witness_size, s);
    h2d(dg.get<void>(), blinding, sizeof(fr_t) * ng, s);
    pk::whir_r1cs_commit_dev(pk::device_current(), dw.get<fr_t>(), witness_size, (int)num_vars,
                             blinding ? dg.get<fr_t>() : nullptr, m0, blowup, coset,
                             leaf_elems, dp.get<fr_t>(), dc.get<fr_t>(), dn.get<fr_t>(), s);
    d2h(root, dn.get<fr_t>() + (nn - 1), 32, s);
    if (polynomial) d2h(polynomial, dp.get<void>(), sizeof(fr_t) * n, s);
    if (codeword) d2h(codeword, dc.get<void>(), sizeof(fr_t) * N, s);
    if (nodes) d2h(nodes, dn.get<void>(), 32 * nn, s);
    sync(s);
  });
}
int pkgpu_whir_r1cs_commit_device(const uint64_t* witness, size_t witness_size, size_t num_vars,
                                  const uint64_t* blinding, size_t m0, size_t blowup, int coset,
                                  size_t leaf_elems, uint64_t* polynomial, uint64_t* codeword,
                                  void* nodes, void* stream) {
  return guarded([&] {
    REQ(polynomial && codeword && nodes, pk::PK_ERR_ARG, "commit: null pointer");
    whir_commit_args(witness, witness_size, num_vars, blinding, m0);
    pk::whir_r1cs_commit_dev(pk::device_current(), reinterpret_cast<const fr_t*>(witness),
                             witness_size, (int)num_vars, reinterpret_cast<const fr_t*>(blinding),
                             m0, blowup, coset, leaf_elems, reinterpret_cast<fr_t*>(polynomial),
                             reinterpret_cast<fr_t*>(codeword), static_cast<fr_t*>(nodes),
                             stream_or_thread(stream));
  });
}

static void scalar_host(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
  REQ(n == 0 || (a && out), pk::PK_ERR_ARG, "scalar: null pointer");
  if (!n) return;
  hipStream_t s = pk::thread_stream();
  Staged da(sizeof(fr_t) * n, s), db(b ? sizeof(fr_t) * n : 0, s), dout(sizeof(fr_t) * n, s);
  h2d(da.get<void>(), a, sizeof(fr_t) * n, s);
  if (b) h2d(db.get<void>(), b, sizeof(fr_t) * n, s);
  pk::scalar_mul_many_dev(da.get<fr_t>(), b ? db.get<fr_t>() : nullptr, dout.get<fr_t>(), n, s);
  d2h(out, dout.get<void>(), sizeof(fr_t) * n, s);
  sync(s);
}
int pkgpu_scalar_mul_many(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
  return guarded([&] {
    REQ(n == 0 || b, pk::PK_ERR_ARG, "scalar: null pointer");
    scalar_host(a, b, out, n);
  });
}
int pkgpu_scalar_sqr_many(const uint64_t* a, uint64_t* out, size_t n) {
  return guarded([&] { scalar_host(a, nullptr, out, n); });
}
int pkgpu_scalar_mul_many_device(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n,
                                 void* stream) {
  return guarded([&] {
    REQ(n == 0 || (a && b && out), pk::PK_ERR_ARG, "scalar: null pointer");
    pk::scalar_mul_many_dev(reinterpret_cast<const fr_t*>(a), reinterpret_cast<const fr_t*>(b),
                            reinterpret_cast<fr_t*>(out), n, stream_or_thread(stream));
  });
}
int pkgpu_scalar_sqr_many_device(const uint64_t* a, uint64_t* out, size_t n, void* stream) {
  return guarded([&] {
    REQ(n == 0 || (a && out), pk::PK_ERR_ARG, "scalar: null pointer");
    pk::scalar_mul_many_dev(reinterpret_cast<const fr_t*>(a), nullptr, reinterpret_cast<fr_t*>(out),
                            n, stream_or_thread(stream));
  });
}

int pkgpu_random_fr_device(const uint8_t key[32], uint64_t first, size_t n, uint64_t* out,
                           void* stream) {
  return guarded([&] {
    REQ(key && (n == 0 || out), pk::PK_ERR_ARG, "random_fr: null pointer");
    pk::random_fr_dev(key, first, n, reinterpret_cast<fr_t*>(out), stream_or_thread(stream));
  });
}

int pkgpu_batch_plan_create(size_t count, const uint64_t* const* coeffs, const size_t* n,
                            const size_t* blowup, const int* coset, size_t leaf_elems,
                            uint64_t* const* codewords, void* const* nodes, int tile_log,
                            void** plan) {
  return guarded([&] {
    REQ(plan && count && coeffs && n && blowup && codewords && nodes, pk::PK_ERR_ARG,
        "batch: null pointer");
    std::vector<pk::BatchPolySpec> specs(count);
    for (size_t i = 0; i < count; i++) {
      specs[i].coeffs = reinterpret_cast<const fr_t*>(coeffs[i]);
      specs[i].n = n[i];
      specs[i].blowup = blowup[i];
      specs[i].coset = coset ? coset[i] : 0;
      specs[i].leaf_elems = leaf_elems;
      specs[i].codeword = reinterpret_cast<fr_t*>(codewords[i]);
      specs[i].nodes = static_cast<fr_t*>(nodes[i]);
    }
    *plan = pk::batch_plan_create(pk::device_current(), specs, tile_log);
  });
}
int pkgpu_batch_run(void* plan, void* stream) {
  return guarded([&] {
    REQ(plan, pk::PK_ERR_STATE, "batch: null plan");
    pk::batch_run(*static_cast<pk::BatchPlan*>(plan), stream_or_thread(stream));
  });
}
int pkgpu_batch_graph_launch(void* plan, void* stream) {
  return guarded_capture([&] {
    REQ(plan, pk::PK_ERR_STATE, "batch: null plan");
    hipGraphExec_t g = pk::batch_graph(*static_cast<pk::BatchPlan*>(plan));
    PK_HIP(hipGraphLaunch(g, stream_or_thread(stream)));
  });
}
int pkgpu_batch_roots(void* plan, uint8_t* roots, void* stream) {
  return guarded([&] {
    REQ(plan && roots, pk::PK_ERR_ARG, "batch: null pointer");
    auto* p = static_cast<pk::BatchPlan*>(plan);
    hipStream_t s = stream_or_thread(stream);
    uint32_t err = 0;
    d2h(&err, p->err, sizeof err, s);
    d2h(roots, p->roots, 32 * (size_t)p->count, s);
    sync(s);
    REQ(!err, pk::PK_ERR_INTERNAL, "batch: hashing queue wait timed out");
  });
}
int pkgpu_batch_plan_destroy(void* plan) {
  return guarded([&] { delete static_cast<pk::BatchPlan*>(plan); });
}

static bool is_host_pointer(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return true;  // unregistered pageable memory
  }
  return at.type != hipMemoryTypeDevice;
}

int pkgpu_server_create(int device, size_t hbm_budget_bytes, int slots, void** server) {
  return guarded([&] {
    REQ(server, pk::PK_ERR_ARG, "server: null out-pointer");
    *server = new pk::Server(device, hbm_budget_bytes, slots);
  });
}
int pkgpu_server_submit(void* server, const pkgpu_job* job, uint64_t* job_id) {
  return guarded([&] {
    REQ(server && job && job_id, pk::PK_ERR_ARG, "server: null pointer");
    pk::JobSpec sp;
    sp.m = (int)job->m;
    const size_t n = job->m < 63 ? (size_t(1) << job->m) : 0;
    for (int k = 0; k < 3; k++) {
      sp.rs[k] = job->row_starts[k];
      sp.cols[k] = job->cols[k];
      sp.vals[k] = reinterpret_cast<const fr_t*>(job->vals[k]);
      sp.nnz[k] = job->nnz[k];
      // host-side structure check for host-resident matrices (device ones are
      // trusted: the kernels index z with their column indices)
      if (sp.rs[k] && is_host_pointer(sp.rs[k]) && (!sp.nnz[k] || is_host_pointer(sp.cols[k])))
        validate_csr(n, n, sp.rs[k], sp.nnz[k], sp.cols[k]);
    }
    sp.z = reinterpret_cast<const fr_t*>(job->z);
    sp.seed = job->seed;
    sp.blowup = job->blowup;
    sp.coset = job->coset;
    sp.leaf_elems = job->leaf_elems;
    *job_id = static_cast<pk::Server*>(server)->submit(sp);
  });
}
int pkgpu_server_wait(void* server, uint64_t job_id, double timeout_ms, pkgpu_job_result* result,
                      uint64_t* coeffs, uint64_t* alphas, uint64_t* r) {
  return guarded_nolock([&] {
    REQ(server && result, pk::PK_ERR_ARG, "server: null pointer");
    pk::JobResult jr = static_cast<pk::Server*>(server)->wait(job_id, timeout_ms);
    memcpy(result->root, &jr.root, 32);
    memcpy(result->final_claim, &jr.final_claim, 32);
    result->latency_ms = jr.latency_ms;
    result->queue_ms = jr.queue_ms;
    result->run_ms = jr.run_ms;
    result->status = jr.status;
    if (coeffs && !jr.coeffs.empty()) memcpy(coeffs, jr.coeffs.data(), 32 * jr.coeffs.size());
    if (alphas && !jr.alphas.empty()) memcpy(alphas, jr.alphas.data(), 32 * jr.alphas.size());
    if (r && !jr.r.empty()) memcpy(r, jr.r.data(), 32 * jr.r.size());
    REQ(jr.status == 0, jr.status, "server: job failed: " + jr.error);
  });
}
int pkgpu_server_poll(void* server, uint64_t job_id, int* done) {
  return guarded_nolock([&] {
    REQ(server && done, pk::PK_ERR_ARG, "server: null pointer");
    *done = static_cast<pk::Server*>(server)->poll(job_id) ? 1 : 0;
  });
}
int pkgpu_server_open(void* server, uint64_t job_id, const uint64_t* leaf_indices, size_t n,
                      uint64_t* leaf_values, uint8_t* paths) {
  return guarded([&] {
    REQ(server && (n == 0 || (leaf_indices && leaf_values && paths)), pk::PK_ERR_ARG,
        "server: null pointer");
    static_cast<pk::Server*>(server)->open(job_id, leaf_indices, n,
                                           reinterpret_cast<fr_t*>(leaf_values),
                                           reinterpret_cast<fr_t*>(paths));
  });
}
int pkgpu_server_release(void* server, uint64_t job_id) {
  return guarded([&] {
    REQ(server, pk::PK_ERR_ARG, "server: null pointer");
    static_cast<pk::Server*>(server)->release(job_id);
  });
}
int pkgpu_server_stats(void* server, uint64_t* bytes_in_use, uint64_t* peak_in_flight) {
  return guarded([&] {
    REQ(server, pk::PK_ERR_ARG, "server: null pointer");
    auto* s = static_cast<pk::Server*>(server);
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use();
    if (peak_in_flight) *peak_in_flight = s->peak_in_flight();
  });
}
int pkgpu_server_destroy(void* server) {
  return guarded([&] { delete static_cast<pk::Server*>(server); });
}

}  // extern "C"
