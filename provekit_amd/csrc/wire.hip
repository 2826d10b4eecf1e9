// R1CS wire format decoded on gfx950: the postcard encodings of R1CS
// (provekit/common/src/r1cs.rs:59-66), Interner (interner.rs:6-10 through
// utils/serde_ark.rs:11-30) and SparseMatrix (sparse_matrix.rs:73-160, delta
// coded columns) straight into device CSR with hydrated Fr values.
//
// Postcard writes every integer wider than a byte as a LEB128 varint, and
// SparseMatrix, its three vectors and their lengths are nothing but varints,
// so the matrices of an R1CS are one flat varint stream.  The device decodes
// the whole stream at once: a byte ends a varint when its high bit is clear;
// an exclusive scan of those flags numbers the varints, and the thread of each
// ending byte reads back over its (at most ten) bytes and writes the value and
// its byte count.  The host then walks the few header fields (three small
// reads per matrix) and the section kernels build the CSR:
//   row starts   the row-start vector, checked (u32, starting at 0,
//                non-decreasing, within the value count);
//   columns      a segmented inclusive scan of the deltas, a new segment at
//                the first entry of every non-empty row (decode_col_deltas,
//                sparse_matrix.rs:118-147: the first column absolute, the
//                rest running sums, u32 wrapping), checked below num_cols;
//   values       interned ids gathered from the interner table (checked).
// The interner arrives as ark's compressed Vec<Fr> (u64 count, then 32-byte
// canonical integers) and is taken to Montgomery form by one product with
// R^2.  Malformed input raises a device flag, read once at the end; every
// kernel stays in bounds whatever the bytes say.
#include <cstring>
#include <string>

#include "host_fr.hpp"
#include "kernels.hpp"
#include "wire.hpp"

namespace pk {

namespace {

inline unsigned nblocks(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

constexpr int kWT = 256;            // threads per workgroup
constexpr int kVarintBytes = 16;    // bytes per thread
constexpr uint64_t kVarintBlock = (uint64_t)kWT * kVarintBytes;
constexpr int kSegPer = 4;          // scan elements per thread
constexpr uint64_t kSegBlock = (uint64_t)kWT * kSegPer;

// Exclusive scan over the workgroup (kWT threads) of one value per thread
// under an associative op with identity id; *total is the fold of all.
// Every thread must call it.
template <int NT, class T, class Op>
__device__ __forceinline__ T block_excl_scan(T v, Op op, T id, T* total) {
  __shared__ T ws[NT / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x = op(y, x);
  }
  T ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = id;
  if (lane == 63) ws[w] = x;
  __syncthreads();
  T pre = id, tot = id;
#pragma unroll
  for (int i = 0; i < NT / 64; i++) {
    if (i < w) pre = op(pre, ws[i]);
    tot = op(tot, ws[i]);
  }
  __syncthreads();
  *total = tot;
  return op(pre, ex);
}

struct AddU32 {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};

// varint ends (bytes with the high bit clear) in one 16-byte group
__device__ __forceinline__ uint32_t ends16(const uint4& q) {
  const uint32_t hi = __popc(q.x & 0x80808080u) + __popc(q.y & 0x80808080u) +
                      __popc(q.z & 0x80808080u) + __popc(q.w & 0x80808080u);
  return 16u - hi;
}

// The stream lives in a buffer padded to whole workgroups with 0x80 bytes
// (continuations, so the padding ends no varint).
__global__ __launch_bounds__(kWT) void k_varint_count(const uint4* b16, uint32_t* bsum) {
  const uint64_t c = ((uint64_t)blockIdx.x * kWT + threadIdx.x) * (kVarintBytes / 16);
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < kVarintBytes / 16; k++) n += ends16(b16[c + k]);
  uint32_t total;
  (void)block_excl_scan<kWT>(n, AddU32(), 0u, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// One workgroup: exclusive prefix of v[0, n) in place under op (identity
// id), the fold of all in *total when total is set.  Each thread takes a
// contiguous run of the values.
template <class T, class Op>
__global__ __launch_bounds__(1024) void k_scan_one(T* v, uint64_t n, Op op, T id, T* total) {
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t i0 = (uint64_t)threadIdx.x * per < n ? (uint64_t)threadIdx.x * per : n;
  const uint64_t i1 = i0 + per < n ? i0 + per : n;
  // loads go out eight at a time (the runs are ~40 values at 2^22 entries)
  T a = id;
  uint64_t i = i0;
  for (; i + 8 <= i1; i += 8) {
    T w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = v[i + k];
#pragma unroll
    for (int k = 0; k < 8; k++) a = op(a, w[k]);
  }
  for (; i < i1; i++) a = op(a, v[i]);
  T tot;
  T run = block_excl_scan<1024>(a, op, id, &tot);
  i = i0;
  for (; i + 8 <= i1; i += 8) {
    T w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = v[i + k];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      v[i + k] = run;
      run = op(run, w[k]);
    }
  }
  for (; i < i1; i++) {
    const T c = v[i];
    v[i] = run;
    run = op(run, c);
  }
  if (total && threadIdx.x == 0) *total = tot;
}

// flat[i], nb[i]: value and byte count of varint i (nb 11: longer than ten
// bytes, or past 64 bits), for the varints ending in this workgroup's bytes.
// Each thread decodes its 16 bytes forward in registers; a varint that began
// before them takes its low bytes from the 16 in front (the previous lane's,
// by a shuffle; a run of 16 continuation bytes there is already too long).
// The workgroup's varints are numbered consecutively, so they are staged in
// LDS and written out as one contiguous, coalesced range.
__global__ __launch_bounds__(kWT) void k_varint_decode(const uint8_t* b, uint64_t len,
                                                       const uint32_t* boff, uint64_t* flat,
                                                       uint8_t* nb) {
  __shared__ uint64_t s_flat[kVarintBlock];
  __shared__ uint8_t s_nb[kVarintBlock];
  const int lane = threadIdx.x & 63;
  const uint64_t j0 = ((uint64_t)blockIdx.x * kWT + threadIdx.x) * kVarintBytes;
  const uint4* b16 = reinterpret_cast<const uint4*>(b + j0);
  const uint4 q = b16[0];
  const uint32_t n = ends16(q);
  uint32_t total;
  uint32_t li = block_excl_scan<kWT>(n, AddU32(), 0u, &total);
  uint4 pv;
  pv.x = __shfl(q.x, (lane + 63) & 63, 64);
  pv.y = __shfl(q.y, (lane + 63) & 63, 64);
  pv.z = __shfl(q.z, (lane + 63) & 63, 64);
  pv.w = __shfl(q.w, (lane + 63) & 63, 64);
  if (lane == 0) pv = j0 ? b16[-1] : make_uint4(0, 0, 0, 0);
  if (n) {
    uint64_t acc = 0;
    int cnt = 0;
    const uint32_t pw[4] = {pv.x, pv.y, pv.z, pv.w};
    int back = 0;
    bool run = true;
#pragma unroll
    for (int i = 15; i >= 0; i--) {
      run = run && ((pw[i >> 2] >> (8 * (i & 3))) & 0x80u);
      back += run ? 1 : 0;
    }
    if (back == 16) {
      cnt = 11;
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int k = i - (16 - back);  // position in the varint
        if (k >= 0 && k < 10)
          acc |= (uint64_t)((pw[i >> 2] >> (8 * (i & 3))) & 0x7fu) << (7 * k);
      }
      cnt = back;
    }
    const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t c = (w4[i >> 2] >> (8 * (i & 3))) & 0xffu;
      if (j0 + i >= len) break;
      if (cnt < 10) acc |= (uint64_t)(c & 0x7fu) << (7 * cnt);
      cnt++;
      if (!(c & 0x80u)) {
        const bool bad = cnt > 10 || (cnt == 10 && c > 1);  // a 10th byte holds bit 63 only
        s_flat[li] = acc;
        s_nb[li] = (uint8_t)(bad ? 11 : cnt);
        li++;
        acc = 0;
        cnt = 0;
      }
    }
  }
  __syncthreads();
  const uint64_t base = boff[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < total; i += kWT) {
    flat[base + i] = s_flat[i];
    nb[base + i] = s_nb[i];
  }
}

// ---------------------------------------------------------------- headers
// Up to three matrix headers walked on the device from varint `first`, so
// the host reads them (and the varint count) in one copy instead of one round
// trip per field.  err: the first failed check, in the host walk's order.
enum WireHeadErr : uint32_t {
  kHeadOk = 0,
  kHeadEnd,      // unexpected end of input
  kHeadVarint,   // a header field is not a valid varint
  kHeadRows,     // num_rows differs from the row-start count
  kHeadDeltas,   // column delta count differs from value count
  kHeadNoRows,   // entries without rows
  kHeadSize,     // sizes beyond u32 indices
};
struct WireHeads {
  uint64_t count;
  uint64_t nr[3], nc[3], n1[3], n3[3], first[3];
  uint32_t err, pad;
};

__global__ void k_wire_heads(const uint64_t* flat, const uint8_t* nb, const uint32_t* count_p,
                             int nmats, WireHeads* out) {
  if (threadIdx.x || blockIdx.x) return;
  const uint64_t count = *count_p;
  out->count = count;
  uint32_t e = kHeadOk;
  uint64_t o = 0;
  for (int m = 0; m < nmats && !e; m++) {
    auto at = [&](uint64_t i) -> uint64_t {
      if (e) return 0;
      if (i >= count) {
        e = kHeadEnd;
        return 0;
      }
      if (nb[i] > 10) {
        e = kHeadVarint;
        return 0;
      }
      return flat[i];
    };
    const uint64_t nr = at(o), nc = at(o + 1), n1 = at(o + 2);
    if (!e && n1 > count) e = kHeadEnd;
    const uint64_t n2 = e ? 0 : at(o + 3 + n1);
    if (!e && n2 > count) e = kHeadEnd;
    const uint64_t n3 = e ? 0 : at(o + 4 + n1 + n2);
    if (!e && (n3 > count || o + 5 + n1 + n2 + n3 > count)) e = kHeadEnd;
    if (!e && nr != n1) e = kHeadRows;
    if (!e && n2 != n3) e = kHeadDeltas;
    if (!e && !n1 && n3) e = kHeadNoRows;
    if (!e && (n1 >= (1ull << 32) || n3 >= (1ull << 32))) e = kHeadSize;
    if (e) break;
    out->nr[m] = nr;
    out->nc[m] = nc;
    out->n1[m] = n1;
    out->n3[m] = n3;
    out->first[m] = o;
    o += 5 + n1 + n2 + n3;
  }
  out->err = e;
}

__device__ __forceinline__ bool u32_ok(uint64_t v, uint8_t n) { return n <= 5 && v <= 0xffffffffull; }

__global__ __launch_bounds__(256) void k_rows(const uint64_t* flat, const uint8_t* nb, uint64_t rows,
                                              uint64_t nnz, uint32_t* row_starts, uint32_t* err) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const uint64_t v = flat[r];
  bool ok = u32_ok(v, nb[r]);
  if (r == 0) ok = ok && v == 0;
  if (r > 0) ok = ok && v >= flat[r - 1];
  if (r + 1 == rows) ok = ok && v <= nnz;
  if (!ok) atomicOr(err, 1u);
  row_starts[r] = (uint32_t)v;
}

constexpr uint64_t kHead = 1ull << 63;

__global__ __launch_bounds__(256) void k_deltas(const uint64_t* flat, const uint8_t* nb, uint64_t n,
                                                uint64_t* packed, uint32_t* err) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t v = flat[k];
  if (!u32_ok(v, nb[k])) atomicOr(err, 1u);
  packed[k] = v & 0xffffffffull;
}

// the first entry of every non-empty row opens a segment
__global__ __launch_bounds__(256) void k_heads(const uint32_t* row_starts, uint64_t rows,
                                               uint64_t nnz, uint64_t* packed) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const uint64_t s = row_starts[r];
  const uint64_t e = r + 1 < rows ? row_starts[r + 1] : nnz;
  if (s < e && s < nnz) packed[s] |= kHead;
}

// segmented sum: (a, b) -> b if b opens a segment, else a + b (mod 2^32),
// keeping a's head bit; associative, identity 0
struct SegSum {
  __device__ uint64_t operator()(uint64_t a, uint64_t b) const {
    if (b & kHead) return b;
    return (a & kHead) | ((a + b) & 0xffffffffull);
  }
};

// this thread's kSegPer (= 4) entries, one 32-byte run; past n they are the
// identity
__device__ __forceinline__ void seg_load4(const uint64_t* x, uint64_t n, uint64_t i0, uint64_t (&e)[4]) {
  if (i0 + 4 <= n) {
    const ulonglong2* v = reinterpret_cast<const ulonglong2*>(x + i0);
    const ulonglong2 a = v[0], b = v[1];
    e[0] = a.x;
    e[1] = a.y;
    e[2] = b.x;
    e[3] = b.y;
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) e[k] = i0 + k < n ? x[i0 + k] : 0;
  }
}

// per workgroup of kSegBlock entries: the fold of its entries
__global__ __launch_bounds__(kWT) void k_seg_reduce(const uint64_t* x, uint64_t n, uint64_t* agg) {
  const uint64_t i0 = ((uint64_t)blockIdx.x * kWT + threadIdx.x) * kSegPer;
  uint64_t e[kSegPer];
  seg_load4(x, n, i0, e);
  uint64_t a = 0;
#pragma unroll
  for (int k = 0; k < kSegPer; k++) a = SegSum()(a, e[k]);
  uint64_t total;
  (void)block_excl_scan<kWT>(a, SegSum(), (uint64_t)0, &total);
  if (threadIdx.x == 0) agg[blockIdx.x] = total;
}

// the inclusive segmented scan from each workgroup's carry-in, written as
// columns (u32) checked below ncols
__global__ __launch_bounds__(kWT) void k_seg_cols(const uint64_t* x, uint64_t n, const uint64_t* carry,
                                                  uint64_t ncols, uint32_t* cols, uint32_t* err) {
  const uint64_t i0 = ((uint64_t)blockIdx.x * kWT + threadIdx.x) * kSegPer;
  uint64_t e[kSegPer];
  seg_load4(x, n, i0, e);
  uint64_t a = 0;
#pragma unroll
  for (int k = 0; k < kSegPer; k++) a = SegSum()(a, e[k]);
  uint64_t total;
  uint64_t run = SegSum()(carry[blockIdx.x], block_excl_scan<kWT>(a, SegSum(), (uint64_t)0, &total));
  uint32_t c[kSegPer];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < kSegPer; k++) {
    run = SegSum()(run, e[k]);
    c[k] = (uint32_t)run;
    bad |= i0 + k < n && c[k] >= ncols;
  }
  if (i0 + 4 <= n) {
    *reinterpret_cast<uint4*>(cols + i0) = make_uint4(c[0], c[1], c[2], c[3]);
  } else {
#pragma unroll
    for (int k = 0; k < kSegPer; k++)
      if (i0 + k < n) cols[i0 + k] = c[k];
  }
  if (bad) atomicOr(err, 1u);
}

__global__ __launch_bounds__(256) void k_hydrate(const uint64_t* flat, const uint8_t* nb, uint64_t nnz,
                                                 const fr_t* interner, uint64_t ni, fr_t* vals,
                                                 uint32_t* err) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  uint64_t id = flat[k];
  if (nb[k] > 10 || id >= ni) {
    atomicOr(err, 1u);
    id = 0;
  }
  fr_t v;
  if (ni) {
    v = fr_load(interner + id);
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) v.l[i] = 0;
  }
  fr_store(vals + k, v);
}

// canonical integer -> Montgomery form (x R^2 R^-1), rejecting x >= p
__global__ __launch_bounds__(256) void k_to_mont(const fr_t* raw, uint64_t n, fr_t r2, fr_t* out,
                                                 uint32_t* err) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fr_t x = fr_load(raw + i);
  fr_t d = x;
  uint32_t br = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) d.l[k] = subb(d.l[k], p_limb(k), br);
  if (!br) atomicOr(err, 1u);  // x >= p
  // a rejected x (its output is never used) enters the product as 0: fr_mul
  // needs operands within its bound on the two top limbs, and x >= p may not be
  fr_t xs;
#pragma unroll
  for (int k = 0; k < 8; k++) xs.l[k] = br ? x.l[k] : 0u;
  fr_t m = fr_mul(xs, r2);
  fr_canon(m);
  fr_store(out + i, m);
}

// ------------------------------------------------------------------ host side
struct HostReader {
  const uint8_t* b;
  size_t len, pos = 0;
  uint64_t varint(int max_bytes, int bits) {
    uint64_t v = 0;
    for (int i = 0; i < max_bytes; i++) {
      PK_REQUIRE(pos < len, PK_ERR_ARG, "wire format: unexpected end of input");
      const uint8_t c = b[pos++];
      const int sh = 7 * i;
      if (i == max_bytes - 1)
        PK_REQUIRE(!(c & 0x80) && (bits - sh >= 7 || (c >> (bits - sh)) == 0), PK_ERR_ARG,
                   "wire format: bad varint");
      v |= (uint64_t)(c & 0x7f) << sh;
      if (!(c & 0x80)) return v;
    }
    throw Error(PK_ERR_ARG, "wire format: bad varint");
  }
  uint64_t usize() { return varint(10, 64); }
};

// Located interner payload: count and the canonical integers.
struct InternerBytes {
  uint64_t count;
  const uint8_t* data;
};
InternerBytes read_interner(HostReader& r) {
  const uint64_t nbytes = r.usize();
  PK_REQUIRE(nbytes <= r.len - r.pos && nbytes >= 8, PK_ERR_ARG,
             "wire format: interner bytes out of range");
  uint64_t n = 0;
  for (int i = 0; i < 8; i++) n |= (uint64_t)r.b[r.pos + i] << (8 * i);
  PK_REQUIRE(n <= (nbytes - 8) / 32 && nbytes == 8 + 32 * n, PK_ERR_ARG,
             "wire format: interner length mismatch");
  InternerBytes ib{n, r.b + r.pos + 8};
  r.pos += nbytes;
  return ib;
}

void interner_to_device(const InternerBytes& ib, fr_t* out, uint32_t* err, hipStream_t s) {
  if (!ib.count) return;
  Scratch raw(32 * ib.count, s);
  PK_HIP(hipMemcpyAsync(raw.p, ib.data, 32 * ib.count, hipMemcpyHostToDevice, s));
  fr_t r2;
  memcpy(r2.l, hf::R2.l, 32);
  k_to_mont<<<nblocks(ib.count, 256), 256, 0, s>>>(raw.as<fr_t>(), ib.count, r2, out, err);
  PK_HIP(hipGetLastError());
  PK_HIP(hipStreamSynchronize(s));  // raw is released at scope exit
}

// The varint stream of bytes b[0, len) on the device.
struct FlatStream {
  Scratch bytes, flat, nb, bsum;
  uint64_t len, count = 0;
  hipStream_t s;
  static uint64_t padded(uint64_t len) {
    return len ? (len + kVarintBlock - 1) / kVarintBlock * kVarintBlock : kVarintBlock;
  }
  // the varints decoded on stream st (no synchronisation)
  FlatStream(const uint8_t* host, uint64_t n, hipStream_t st)
      : bytes(padded(n), st), flat(8 * (n ? n : 1), st), nb(n ? n : 1, st),
        bsum(4 * (padded(n) / kVarintBlock + 1), st), len(n), s(st) {
    PK_REQUIRE(len < (1ull << 32), PK_ERR_SIZE, "wire format: input above 4 GiB");
    if (!len) return;
    const uint64_t plen = padded(len), blocks = plen / kVarintBlock;
    h2d_from_pageable(bytes.p, host, len, s);
    if (plen > len) PK_HIP(hipMemsetAsync(bytes.as<uint8_t>() + len, 0x80, plen - len, s));
    k_varint_count<<<(unsigned)blocks, kWT, 0, s>>>(bytes.as<uint4>(), bsum.as<uint32_t>());
    k_scan_one<<<1, 1024, 0, s>>>(bsum.as<uint32_t>(), blocks, AddU32(), 0u,
                                  bsum.as<uint32_t>() + blocks);
    k_varint_decode<<<(unsigned)blocks, kWT, 0, s>>>(bytes.as<uint8_t>(), len, bsum.as<uint32_t>(),
                                                     flat.as<uint64_t>(), nb.as<uint8_t>());
    PK_HIP(hipGetLastError());
  }
  // The varint count and the headers of nmats matrices stored one after
  // another from varint 0 (checked in the order of SparseMatrix's fields).
  void heads(int nmats, WireCsrSizes* z) {
    PK_REQUIRE(len, PK_ERR_ARG, "wire format: unexpected end of input");
    const uint64_t blocks = padded(len) / kVarintBlock;
    Scratch dh(sizeof(WireHeads), s);
    k_wire_heads<<<1, 64, 0, s>>>(flat.as<uint64_t>(), nb.as<uint8_t>(),
                                  bsum.as<uint32_t>() + blocks, nmats, dh.as<WireHeads>());
    PK_HIP(hipGetLastError());
    WireHeads* h = static_cast<WireHeads*>(pinned_small());
    PK_HIP(hipMemcpyAsync(h, dh.p, sizeof(WireHeads), hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    count = h->count;
    switch (h->err) {
      case kHeadOk: break;
      case kHeadEnd: throw Error(PK_ERR_ARG, "wire format: unexpected end of input");
      case kHeadVarint: throw Error(PK_ERR_ARG, "wire format: bad varint");
      case kHeadRows: throw Error(PK_ERR_ARG, "wire format: num_rows differs from the row-start count");
      case kHeadDeltas:
        throw Error(PK_ERR_ARG, "wire format: column delta count differs from value count");
      case kHeadNoRows: throw Error(PK_ERR_ARG, "wire format: entries without rows");
      default: throw Error(PK_ERR_SIZE, "wire format: sizes must fit u32 indices");
    }
    for (int m = 0; m < nmats; m++) {
      z[m].num_rows = h->n1[m];
      z[m].num_cols = h->nc[m];
      z[m].nnz = h->n3[m];
      z[m].first = h->first[m];
    }
  }
};

void matrix_fill(const FlatStream& fs, const WireCsrSizes& z, const fr_t* interner, uint64_t ni,
                 const WireCsrOut& out, uint32_t* err, hipStream_t s) {
  const uint64_t o = z.first, n1 = z.num_rows, n3 = z.nnz;
  const uint64_t* flat = fs.flat.as<uint64_t>();
  const uint8_t* nb = fs.nb.as<uint8_t>();
  if (n1) {
    PK_REQUIRE(out.row_starts, PK_ERR_ARG, "wire format: null row starts");
    k_rows<<<nblocks(n1, 256), 256, 0, s>>>(flat + o + 3, nb + o + 3, n1, n3, out.row_starts, err);
    PK_HIP(hipGetLastError());
  }
  if (!n3) return;
  PK_REQUIRE(out.cols && out.vals, PK_ERR_ARG, "wire format: null entries");
  const uint64_t nblk = (n3 + kSegBlock - 1) / kSegBlock;
  Scratch packed(8 * n3, s), agg(8 * nblk, s);
  k_deltas<<<nblocks(n3, 256), 256, 0, s>>>(flat + o + 4 + n1, nb + o + 4 + n1, n3,
                                            packed.as<uint64_t>(), err);
  k_heads<<<nblocks(n1, 256), 256, 0, s>>>(out.row_starts, n1, n3, packed.as<uint64_t>());
  k_seg_reduce<<<(unsigned)nblk, kWT, 0, s>>>(packed.as<uint64_t>(), n3, agg.as<uint64_t>());
  k_scan_one<<<1, 1024, 0, s>>>(agg.as<uint64_t>(), nblk, SegSum(), (uint64_t)0,
                                static_cast<uint64_t*>(nullptr));
  k_seg_cols<<<(unsigned)nblk, kWT, 0, s>>>(packed.as<uint64_t>(), n3, agg.as<uint64_t>(),
                                             z.num_cols, out.cols, err);
  k_hydrate<<<nblocks(n3, 256), 256, 0, s>>>(flat + o + 5 + n1 + n3, nb + o + 5 + n1 + n3,
                                             n3, interner, ni, out.vals, err);
  PK_HIP(hipGetLastError());
}

void check_flag(uint32_t* err, hipStream_t s, const char* what) {
  uint32_t* h = static_cast<uint32_t*>(pinned_small());
  PK_HIP(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, s));
  PK_HIP(hipStreamSynchronize(s));
  PK_REQUIRE(!*h, PK_ERR_ARG, std::string("wire format: ") + what);
}

}  // namespace

uint64_t interner_decode_dev(const uint8_t* blob, size_t len, fr_t* out, hipStream_t s) {
  HostReader r{blob, len};
  const InternerBytes ib = read_interner(r);
  if (out && ib.count) {
    Scratch err(4, s);
    PK_HIP(hipMemsetAsync(err.p, 0, 4, s));
    interner_to_device(ib, out, err.as<uint32_t>(), s);
    check_flag(err.as<uint32_t>(), s, "interner value not below the modulus");
  }
  return ib.count;
}

WireCsrSizes sparse_matrix_decode_dev(const uint8_t* blob, size_t len, const fr_t* interner,
                                      uint64_t ni, const WireCsrOut* out, hipStream_t s) {
  FlatStream fs(blob, len, s);
  WireCsrSizes z;
  fs.heads(1, &z);
  if (out) {
    Scratch err(4, s);
    PK_HIP(hipMemsetAsync(err.p, 0, 4, s));
    matrix_fill(fs, z, interner, ni, *out, err.as<uint32_t>(), s);
    check_flag(err.as<uint32_t>(), s, "malformed sparse matrix");
  }
  return z;
}

R1csSizes r1cs_decode_dev(const uint8_t* blob, size_t len, fr_t* interner_out,
                          const WireCsrOut* outs, hipStream_t s) {
  HostReader r{blob, len};
  R1csSizes z;
  z.num_public_inputs = r.usize();
  const InternerBytes ib = read_interner(r);
  z.interner_len = ib.count;
  FlatStream fs(blob + r.pos, len - r.pos, s);
  fs.heads(3, z.mats);
  if (interner_out) {
    Scratch err(4, s);
    PK_HIP(hipMemsetAsync(err.p, 0, 4, s));
    interner_to_device(ib, interner_out, err.as<uint32_t>(), s);
    check_flag(err.as<uint32_t>(), s, "interner value not below the modulus");
    for (int m = 0; m < 3; m++)
      matrix_fill(fs, z.mats[m], interner_out, ib.count, outs[m], err.as<uint32_t>(), s);
    check_flag(err.as<uint32_t>(), s, "malformed sparse matrix");
  }
  return z;
}

}  // namespace pk
