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
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <cstring>
#include <string>

#include "host_fr.hpp"
#include "kernels.hpp"
#include "wire.hpp"

namespace pk {

namespace {

inline unsigned nblocks(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct EndsVarint {
  __device__ __host__ uint32_t operator()(uint8_t c) const { return (c & 0x80) ? 0u : 1u; }
};

// flat[i], nb[i]: value and byte count of varint i (nb 11: longer than ten
// bytes, or past 64 bits)
__global__ __launch_bounds__(256) void k_varints(const uint8_t* b, uint64_t len,
                                                 const uint32_t* idx, uint64_t* flat,
                                                 uint8_t* nb) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  const uint8_t c = b[j];
  if (c & 0x80) return;
  uint64_t v = c;
  int n = 1;
  bool bad = false;
  for (uint64_t k = j; k > 0 && (b[k - 1] & 0x80); k--) {
    if (n == 10) {
      bad = true;
      break;
    }
    v = (v << 7) | (b[k - 1] & 0x7f);
    n++;
  }
  if (n == 10 && c > 1) bad = true;  // a 10th byte holds bit 63 only
  flat[idx[j]] = v;
  nb[idx[j]] = (uint8_t)(bad ? 11 : n);
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
// keeping a's head bit; associative
struct SegSum {
  __device__ __host__ uint64_t operator()(uint64_t a, uint64_t b) const {
    if (b & kHead) return b;
    return (a & kHead) | ((a + b) & 0xffffffffull);
  }
};

__global__ __launch_bounds__(256) void k_cols(const uint64_t* scanned, uint64_t nnz, uint64_t ncols,
                                              uint32_t* cols, uint32_t* err) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t c = (uint32_t)scanned[k];
  if (c >= ncols) atomicOr(err, 1u);
  cols[k] = c;
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
  Scratch bytes, flat, nb;
  uint64_t count = 0;
  hipStream_t s;
  FlatStream(const uint8_t* host, uint64_t len, hipStream_t st)
      : bytes(len, st), flat(8 * (len ? len : 1), st), nb(len ? len : 1, st), s(st) {
    PK_REQUIRE(len < (1ull << 32), PK_ERR_SIZE, "wire format: input above 4 GiB");
    if (!len) return;
    PK_HIP(hipMemcpyAsync(bytes.p, host, len, hipMemcpyHostToDevice, s));
    Scratch idx(4 * len, s);
    rocprim::transform_iterator<const uint8_t*, EndsVarint, uint32_t> ends(bytes.as<uint8_t>(),
                                                                          EndsVarint());
    size_t tb = 0;
    PK_HIP(rocprim::exclusive_scan(nullptr, tb, ends, idx.as<uint32_t>(), 0u, (size_t)len,
                                   rocprim::plus<uint32_t>(), s));
    Scratch temp(tb ? tb : 16, s);
    PK_HIP(rocprim::exclusive_scan(temp.p, tb, ends, idx.as<uint32_t>(), 0u, (size_t)len,
                                   rocprim::plus<uint32_t>(), s));
    k_varints<<<nblocks(len, 256), 256, 0, s>>>(bytes.as<uint8_t>(), len, idx.as<uint32_t>(),
                                                flat.as<uint64_t>(), nb.as<uint8_t>());
    PK_HIP(hipGetLastError());
    uint32_t last_idx = 0;
    PK_HIP(hipMemcpyAsync(&last_idx, idx.as<uint32_t>() + (len - 1), 4, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    count = (uint64_t)last_idx + ((host[len - 1] & 0x80) ? 0 : 1);
  }
  // usize header field i (checked)
  uint64_t usize_at(uint64_t i) const {
    PK_REQUIRE(i < count, PK_ERR_ARG, "wire format: unexpected end of input");
    uint64_t v = 0;
    uint8_t n = 0;
    PK_HIP(hipMemcpyAsync(&v, flat.as<uint64_t>() + i, 8, hipMemcpyDeviceToHost, s));
    PK_HIP(hipMemcpyAsync(&n, nb.as<uint8_t>() + i, 1, hipMemcpyDeviceToHost, s));
    PK_HIP(hipStreamSynchronize(s));
    PK_REQUIRE(n <= 10, PK_ERR_ARG, "wire format: bad varint");
    return v;
  }
};

// Header of the matrix whose first varint is `o`; returns the varint after it.
uint64_t matrix_header(const FlatStream& fs, uint64_t o, WireCsrSizes& z) {
  const uint64_t nr = fs.usize_at(o), nc = fs.usize_at(o + 1), n1 = fs.usize_at(o + 2);
  PK_REQUIRE(n1 <= fs.count, PK_ERR_ARG, "wire format: unexpected end of input");
  const uint64_t n2 = fs.usize_at(o + 3 + n1);
  PK_REQUIRE(n2 <= fs.count, PK_ERR_ARG, "wire format: unexpected end of input");
  const uint64_t n3 = fs.usize_at(o + 4 + n1 + n2);
  PK_REQUIRE(n3 <= fs.count && o + 5 + n1 + n2 + n3 <= fs.count, PK_ERR_ARG,
             "wire format: unexpected end of input");
  PK_REQUIRE(nr == n1, PK_ERR_ARG, "wire format: num_rows differs from the row-start count");
  PK_REQUIRE(n2 == n3, PK_ERR_ARG, "wire format: column delta count differs from value count");
  PK_REQUIRE(n1 || !n3, PK_ERR_ARG, "wire format: entries without rows");
  PK_REQUIRE(n1 < (1ull << 32) && n3 < (1ull << 32), PK_ERR_SIZE,
             "wire format: sizes must fit u32 indices");
  z.num_rows = n1;
  z.num_cols = nc;
  z.nnz = n3;
  z.first = o;
  return o + 5 + n1 + n2 + n3;
}

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
  Scratch packed(8 * n3, s), scanned(8 * n3, s);
  k_deltas<<<nblocks(n3, 256), 256, 0, s>>>(flat + o + 4 + n1, nb + o + 4 + n1, n3,
                                            packed.as<uint64_t>(), err);
  k_heads<<<nblocks(n1, 256), 256, 0, s>>>(out.row_starts, n1, n3, packed.as<uint64_t>());
  PK_HIP(hipGetLastError());
  size_t tb = 0;
  PK_HIP(rocprim::inclusive_scan(nullptr, tb, packed.as<uint64_t>(), scanned.as<uint64_t>(),
                                 (size_t)n3, SegSum(), s));
  Scratch temp(tb ? tb : 16, s);
  PK_HIP(rocprim::inclusive_scan(temp.p, tb, packed.as<uint64_t>(), scanned.as<uint64_t>(),
                                 (size_t)n3, SegSum(), s));
  k_cols<<<nblocks(n3, 256), 256, 0, s>>>(scanned.as<uint64_t>(), n3, z.num_cols, out.cols, err);
  k_hydrate<<<nblocks(n3, 256), 256, 0, s>>>(flat + o + 5 + n1 + n3, nb + o + 5 + n1 + n3,
                                             n3, interner, ni, out.vals, err);
  PK_HIP(hipGetLastError());
}

void check_flag(uint32_t* err, hipStream_t s, const char* what) {
  uint32_t h = 0;
  PK_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, s));
  PK_HIP(hipStreamSynchronize(s));
  PK_REQUIRE(!h, PK_ERR_ARG, std::string("wire format: ") + what);
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
  matrix_header(fs, 0, z);
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
  uint64_t o = 0;
  for (int m = 0; m < 3; m++) o = matrix_header(fs, o, z.mats[m]);
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
