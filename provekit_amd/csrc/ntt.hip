// Reed-Solomon encoding by NTT over BN254 Fr on gfx950.
//
// Algorithm (same transform as ntt/src/ntt.rs:150-248): decimation in time,
// natural input order, bit-reversed output.  Stage s (s = 0 .. log2(codeword)-1)
// splits the array into 2^s groups; group k pairs (i, i + half) with twiddle
// rr[k], rr[j] = w^{bitrev(j)}.  Exact modular arithmetic makes the result
// independent of how the stages are scheduled.
//
// Schedule on the GPU: the stages are cut into passes of up to 11 stages.  A
// pass covering stages [s0, s0 + r) touches index bits [q, q + r) with
// q = L - s0 - r; a workgroup stages a tile of 2^r "window" positions times
// 2^jb consecutive low indices (coalesced runs of 2^jb * 32 bytes) in LDS,
// runs its r stages there with one barrier per stage, and writes the tile
// back.  Elements stay in the lazy range [0, 2p) between stages and are
// made canonical in the last pass.
#include "kernels.hpp"
#include "host_fr.hpp"

namespace pk {

namespace {

__device__ __forceinline__ uint64_t bitrev(uint64_t v, int bits) {
  return bits == 0 ? 0 : (__builtin_bitreverse64(v) >> (64 - bits));
}

// out[i] = base^i, i < n, by square-and-multiply per element (set-up only).
__global__ void k_pow_table(fr_t* out, fr_t base, uint64_t n, fr_t one) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t r = one, b = base;
  uint64_t e = i;
  while (e) {
    if (e & 1) r = fr_mul(r, b);
    b = fr_mul(b, b);
    e >>= 1;
  }
  fr_canon(r);
  out[i] = r;
}

// out[j] = lo[e & mask] * hi[e >> lo_bits] with e = bitrev_{bits}(j) if
// bitrev_bits > 0, else e = j.  lo/hi hold base^i and base^(i << lo_bits).
__global__ void k_two_level(fr_t* out, const fr_t* lo, const fr_t* hi, int lo_bits, uint64_t n,
                            int bitrev_bits) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t e = bitrev_bits > 0 ? bitrev(j, bitrev_bits) : j;
  fr_t v = fr_mul(lo[e & ((1ull << lo_bits) - 1)], hi[e >> lo_bits]);
  fr_canon(v);
  out[j] = v;
}

// Build a table of 2^log entries of base^{e(j)} (e = bitrev if requested).
void build_power_table(fr_t* out, const hfr& base_h, int log, bool bit_reversed, hipStream_t s) {
  uint64_t n = 1ull << log;
  int lo_bits = (log + 1) / 2;
  int hi_bits = log - lo_bits;
  fr_t *lo, *hi;
  PK_HIP(hipMallocAsync((void**)&lo, sizeof(fr_t) << lo_bits, s));
  PK_HIP(hipMallocAsync((void**)&hi, sizeof(fr_t) << hi_bits, s));
  fr_t base, one, step;
  memcpy(&base, &base_h, 32);
  memcpy(&one, &hf::ONE, 32);
  hfr step_h = base_h;
  for (int i = 0; i < lo_bits; i++) step_h = hf::mul(step_h, step_h);
  memcpy(&step, &step_h, 32);
  uint64_t nlo = 1ull << lo_bits, nhi = 1ull << hi_bits;
  k_pow_table<<<(unsigned)((nlo + 255) / 256), 256, 0, s>>>(lo, base, nlo, one);
  k_pow_table<<<(unsigned)((nhi + 255) / 256), 256, 0, s>>>(hi, step, nhi, one);
  k_two_level<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(out, lo, hi, lo_bits, n,
                                                          bit_reversed ? log : 0);
  PK_HIP(hipGetLastError());
  PK_HIP(hipFreeAsync(lo, s));
  PK_HIP(hipFreeAsync(hi, s));
}

using PassArgs = NttPassDesc;

// A pass kernel is instantiated per tile size 2^TL and butterflies per
// thread per stage BPT: 2^(TL-1) / BPT threads.  BPT trades occupancy (more
// waves per tile to hide the per-stage barrier and the tile load) against
// per-thread independent work; the autotuner picks it per length.
template <int TL, int BPT>
struct PassShape {
  static constexpr int BF = 1 << (TL - 1);  // butterflies per stage
  static constexpr int T = BF / BPT;
  static_assert(BF % BPT == 0 && T >= 1 && T <= 1024, "pass shape");
};

// Flat element index of codeword position i, column c: i * M + c.  A pass
// over stages [s0, s0 + r) splits i = (h, w, i_lo) with w the r window bits
// and i_lo the q bits below; the element sits at
//   h * (2^r * Q) + w * Q + l,   l = i_lo * M + c in [0, Q).
// A tile is one h, all 2^r windows, and 2^jb consecutive l (2^jb divides Q).
//
// Butterfly (Harvey's lazy form): values live in [0, 4p) between stages.
// x is first brought below 2p, the twiddle product m = w * y is below 2p for
// y < 4p (w canonical), and x + m, x - m + 2p stay below 4p < 2^256.
template <int TL, int BPT>
__device__ __forceinline__ void ntt_pass_tile(const PassArgs& a, uint64_t tile, fr_t* lds, fr_t* tws) {
  constexpr int T = PassShape<TL, BPT>::T;
  constexpr int N = 1 << TL;
  const int tid = threadIdx.x;
  const int jb = a.jb;
  const uint64_t Q = a.Q;
  const uint64_t tiles_per_h = Q >> jb;
  const uint64_t h = tile / tiles_per_h;
  const uint64_t j0 = (tile % tiles_per_h) << jb;
  const uint64_t base = h * (Q << a.r) + j0;
  const uint32_t jmask = (1u << jb) - 1;

  // Twiddle of butterfly k of this thread in stage t: rr[(h << t) | (w0 >> (r - t))],
  // one of the 2^t entries rr[(h << t) | j], j < 2^t, consecutive in the table.
  // The early stages (t < tl, tl = r - 2) share them across the tile: they
  // are staged in LDS once per tile (coalesced loads, tws[2^t - 1 + j],
  // 2^tl - 1 entries, a quarter of the tile).  In the last two stages every
  // butterfly has its own, so those come from the table, loaded one stage
  // ahead (the loads stay in flight across the barrier, which waits on LDS
  // only).
  const int tl = a.r > 2 ? a.r - 2 : 0;
  auto tw_index = [&](int t, int k) -> uint64_t {
    const int pos = a.r - 1 - t;
    const int wb = (tid + k * T) >> jb;
    const int w0 = ((wb >> pos) << (pos + 1)) | (wb & ((1 << pos) - 1));
    return (h << t) | ((uint64_t)w0 >> (a.r - t));
  };
  for (int i = tid; i < (1 << tl) - 1; i += T) {
    const int t = 31 - __builtin_clz((unsigned)(i + 1));
    const uint64_t j = (uint64_t)(i + 1 - (1 << t));
    tws[i] = fr_load(a.rr + ((h << t) | j));
  }
  fr_t tw_next[BPT];
  if (tl == 0 && a.r > 0) {
#pragma unroll
    for (int k = 0; k < BPT; k++) tw_next[k] = fr_load(a.rr + tw_index(0, k));
  }

  for (int e = tid; e < N; e += T) {
    const uint64_t g = base + (uint64_t)(e >> jb) * Q + (e & jmask);
    fr_t v;
    if (a.src) {
      // a 64-bit remainder is ~50 instructions; the chunk is a power of two
      // whenever the interleave width is
      const uint64_t gc = a.chunk_pow2 ? (g & (a.chunk - 1)) : g % a.chunk;
      v = fr_load(a.src + gc);
    } else {
      v = fr_load(a.data + g);
    }
    lds[e] = v;
  }
  __syncthreads();

  for (int t = 0; t < a.r; t++) {
    const int pos = a.r - 1 - t;
    const int span = 1 << (pos + jb);
    fr_t x[BPT], y[BPT], w[BPT];
    int e0[BPT];
    if (t >= tl) {
#pragma unroll
      for (int k = 0; k < BPT; k++) w[k] = tw_next[k];
    }
    if (t + 1 < a.r && t + 1 >= tl) {
#pragma unroll
      for (int k = 0; k < BPT; k++) tw_next[k] = fr_load(a.rr + tw_index(t + 1, k));
    }
#pragma unroll
    for (int k = 0; k < BPT; k++) {
      const int b = tid + k * T;
      const int jl = b & jmask;
      const int wb = b >> jb;
      const int w0 = ((wb >> pos) << (pos + 1)) | (wb & ((1 << pos) - 1));
      e0[k] = (w0 << jb) | jl;
      x[k] = lds[e0[k]];
      y[k] = lds[e0[k] + span];
      if (t < tl) w[k] = tws[(1 << t) - 1 + (w0 >> (a.r - t))];
    }
#pragma unroll
    for (int k = 0; k < BPT; k++) {
      const fr_t m = fr_mul(w[k], y[k]);
      fr_t u = x[k];
      cond_sub<2>(u);
      lds[e0[k]] = fr_add_raw(u, m);
      lds[e0[k] + span] = fr_sub_plus_2p(u, m);
    }
    __syncthreads();
  }

  for (int e = tid; e < N; e += T) {
    const uint64_t g = base + (uint64_t)(e >> jb) * Q + (e & jmask);
    fr_t v = lds[e];
    if (a.canon) {
      cond_sub<2>(v);
      fr_canon(v);
    }
    fr_store(a.data + g, v);
  }
}

__device__ __forceinline__ uint64_t pass_tile_of(const PassArgs& a, uint64_t b) {
  const uint64_t T = a.Q >> a.jb;  // tiles per row h
  return ntt_tile_order(a.src != nullptr, T, gridDim.x / T, b);
}

template <int TL, int BPT>
__global__ __launch_bounds__((PassShape<TL, BPT>::T)) void k_ntt_pass(PassArgs a) {
  __shared__ fr_t lds[1 << TL];
  __shared__ fr_t tws[TL > 2 ? 1 << (TL - 2) : 1];
  ntt_pass_tile<TL, BPT>(a, pass_tile_of(a, blockIdx.x), lds, tws);
}

// Batched pass: workgroup b runs tile map[b].y of descriptor map[b].x, so one
// launch advances many polynomials' NTTs (all with the same kernel shape).
template <int TL, int BPT>
__global__ __launch_bounds__((PassShape<TL, BPT>::T)) void k_ntt_pass_batch(const PassArgs* descs,
                                                                          const uint2* map) {
  __shared__ fr_t lds[1 << TL];
  __shared__ fr_t tws[TL > 2 ? 1 << (TL - 2) : 1];
  const uint2 m = map[blockIdx.x];
  ntt_pass_tile<TL, BPT>(descs[m.x], m.y, lds, tws);
}

struct PassVariant {
  int tl, bpt, threads;
  void (*plain)(PassArgs);
  void (*batch)(const PassArgs*, const uint2*);
};

template <int TL, int BPT>
constexpr PassVariant variant() {
  return {TL, BPT, PassShape<TL, BPT>::T, k_ntt_pass<TL, BPT>, k_ntt_pass_batch<TL, BPT>};
}

// Small tiles have one shape; tiles of 2^6..2^11 come with 1, 2 or 4
// butterflies per thread.
const PassVariant kVariants[] = {
    variant<1, 1>(),  variant<2, 2>(),  variant<3, 4>(),  variant<4, 4>(),  variant<5, 4>(),
    variant<6, 1>(),  variant<6, 2>(),  variant<6, 4>(),  variant<7, 1>(),  variant<7, 2>(),
    variant<7, 4>(),  variant<8, 1>(),  variant<8, 2>(),  variant<8, 4>(),  variant<9, 1>(),
    variant<9, 2>(),  variant<9, 4>(),  variant<10, 1>(), variant<10, 2>(), variant<10, 4>(),
    variant<11, 1>(), variant<11, 2>(), variant<11, 4>(),
};

const PassVariant& find_variant(int tl, int bpt) {
  for (const PassVariant& v : kVariants)
    if (v.tl == tl && v.bpt == bpt) return v;
  throw Error(PK_ERR_INTERNAL, "ntt: no pass kernel for this tile shape");
}

// Copy with replication: data[g] = src[g % chunk].
__global__ void k_replicate(fr_t* data, const fr_t* src, uint64_t chunk, uint64_t len) {
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < len) fr_store(data + g, fr_load(src + g % chunk));
}

}  // namespace

bool ntt_shape_supported(int tile_log, int bpt) {
  for (const PassVariant& v : kVariants)
    if (v.tl == tile_log && v.bpt == bpt) return true;
  return false;
}

int ntt_default_bpt(int tile_log) {
  return tile_log >= 3 ? 4 : (tile_log == 2 ? 2 : 1);
}

struct NttPass {
  int s0, r, jb;
};

// Pass plan for stages [s_begin, s_end) of codeword index bits (L = s_end)
// with interleave width M = 2^m_twos * odd.  The last pass takes as many
// stages as fit a full tile; earlier passes split the rest evenly, each
// keeping at least 2^min_run consecutive elements per run where the layout
// allows.  jb is capped by the power of two dividing the low stride 2^q * M.
static std::vector<NttPass> plan_passes(int L, int s_begin, int s_end, int m_twos, int tile_log,
                                        int min_run) {
  std::vector<NttPass> out;
  int rem = s_end - s_begin;
  if (rem <= 0) return out;
  int r_fin = rem < tile_log ? rem : tile_log;
  int rest = rem - r_fin;
  const int per = tile_log - min_run;
  int nfirst = (rest + per - 1) / per;
  int s0 = s_begin;
  for (int i = 0; i < nfirst; i++) {
    int r = rest / nfirst + (i < rest % nfirst ? 1 : 0);
    int q2 = (L - s0 - r) + m_twos;
    int jb = q2 < (tile_log - r) ? q2 : (tile_log - r);
    out.push_back({s0, r, jb});
    s0 += r;
  }
  int q2 = (L - s_end) + m_twos;
  int jb = q2 < (tile_log - r_fin) ? q2 : (tile_log - r_fin);
  out.push_back({s0, r_fin, jb});
  return out;
}

std::vector<NttPassDesc> ntt_plan(Device& d, fr_t* data, uint64_t len, uint64_t codeword,
                                  uint64_t num_groups, const fr_t* src, uint64_t chunk,
                                  NttShape shape) {
  PK_REQUIRE(codeword >= 1 && is_pow2(codeword) && len % codeword == 0, PK_ERR_SIZE,
             "ntt: length must be a multiple of the power-of-two codeword size");
  PK_REQUIRE(is_pow2(num_groups) && num_groups <= codeword, PK_ERR_SIZE,
             "ntt: num_groups must be a power of two up to the codeword size");
  const uint64_t M = len / codeword;
  int m_twos = 0;
  while (((M >> m_twos) & 1) == 0) m_twos++;
  const int L = log2_exact(codeword);
  const int s_begin = log2_exact(num_groups);
  if (shape.tile_log == 0) shape = ntt_tuned_shape(L, L - s_begin);
  if (shape.bpt == 0) shape.bpt = ntt_default_bpt(shape.tile_log);
  if (shape.min_run < 0) shape.min_run = 2;
  PK_REQUIRE(ntt_shape_supported(shape.tile_log, shape.bpt) && shape.min_run < shape.tile_log,
             PK_ERR_ARG, "ntt: unsupported tile shape");
  std::vector<NttPass> plan = plan_passes(L, s_begin, L, m_twos, shape.tile_log, shape.min_run);
  std::vector<NttPassDesc> out;
  if (plan.empty()) return out;
  const fr_t* rr = roots_table(d, L - 1);
  for (size_t i = 0; i < plan.size(); i++) {
    NttPassDesc a;
    a.data = data;
    a.src = (i == 0) ? src : nullptr;
    a.chunk = chunk;
    a.chunk_pow2 = chunk != 0 && is_pow2(chunk);
    a.rr = rr;
    a.Q = (1ull << (L - plan[i].s0 - plan[i].r)) * M;
    a.r = plan[i].r;
    a.jb = plan[i].jb;
    a.canon = (i + 1 == plan.size());
    a.tl = plan[i].r + plan[i].jb;
    // a pass narrower than the tuned tile (short transforms) takes its
    // tile's default shape
    a.bpt = a.tl == shape.tile_log ? shape.bpt : ntt_default_bpt(a.tl);
    if (!ntt_shape_supported(a.tl, a.bpt)) a.bpt = ntt_default_bpt(a.tl);
    out.push_back(a);
  }
  return out;
}

void ntt_run_plan(const std::vector<NttPassDesc>& plan, uint64_t len, hipStream_t s) {
  for (const NttPassDesc& a : plan) {
    const PassVariant& v = find_variant(a.tl, a.bpt);
    v.plain<<<(unsigned)(len >> a.tl), v.threads, 0, s>>>(a);
    PK_HIP(hipGetLastError());
  }
}

void ntt_nr_dev(Device& d, fr_t* data, uint64_t len, uint64_t codeword, uint64_t num_groups,
                hipStream_t s, const fr_t* src, uint64_t chunk) {
  std::vector<NttPassDesc> plan = ntt_plan(d, data, len, codeword, num_groups, src, chunk);
  if (plan.empty()) {
    if (src) {
      k_replicate<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(data, src, chunk, len);
      PK_HIP(hipGetLastError());
    }
    return;
  }
  ntt_run_plan(plan, len, s);
}

void ntt_batch_round_dev(const NttPassDesc* descs, const uint2* map, uint32_t blocks, int tl,
                         int bpt, hipStream_t s) {
  if (!blocks) return;
  const PassVariant& v = find_variant(tl, bpt);
  v.batch<<<blocks, v.threads, 0, s>>>(descs, map);
  PK_HIP(hipGetLastError());
}

namespace {
// y[i] = x[e((n - i) mod n)] with e = bitrev when `rev` (intt_rn: the
// reverse_order then values[1..].reverse() of ntt/src/ntt.rs:326-341 as one
// gather), else the identity (intt_nr: values[1..].reverse() alone).
__global__ void k_intt_in(const fr_t* x, fr_t* y, uint64_t n, int bits, int rev) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = (n - i) & (n - 1);
  fr_store(y + i, fr_load(x + (rev ? bitrev(k, bits) : k)));
}
// out[i] = f * w[e(i)]: the 1/n scaling, then reverse_order when `rev`.
__global__ void k_intt_out(const fr_t* w, fr_t* out, uint64_t n, int bits, fr_t f, int rev) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_t v = fr_mul(f, fr_load(w + (rev ? bitrev(i, bits) : i)));
  fr_canon(v);
  fr_store(out + i, v);
}
__global__ void k_bitrev(const fr_t* x, fr_t* y, uint64_t n, int bits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fr_store(y + i, fr_load(x + bitrev(i, bits)));
}
// In place: thread i swaps entries i and bitrev(i) when i < bitrev(i), so
// every pair has exactly one owner.
__global__ void k_bitrev_inplace(fr_t* x, uint64_t n, int bits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = bitrev(i, bits);
  if (i < r) {
    const fr_t a = fr_load(x + i), b = fr_load(x + r);
    fr_store(x + i, b);
    fr_store(x + r, a);
  }
}

// The inverse transforms share one schedule: gather (k_intt_in), forward
// NTT out of place in scratch, scale by 1/n and gather back (k_intt_out).
void intt_dev(Device& d, fr_t* data, uint64_t n, bool bit_reversed, hipStream_t s) {
  const int bits = log2_exact(n);
  Scratch tmp(sizeof(fr_t) * n, s);
  fr_t* w = tmp.as<fr_t>();
  const unsigned g = (unsigned)((n + 255) / 256);
  k_intt_in<<<g, 256, 0, s>>>(data, w, n, bits, bit_reversed ? 1 : 0);
  PK_HIP(hipGetLastError());
  ntt_nr_dev(d, w, n, n, 1, s);
  hfr inv_n = hf::inv(hf::from_u64(n));
  fr_t f;
  memcpy(&f, &inv_n, 32);
  k_intt_out<<<g, 256, 0, s>>>(w, data, n, bits, f, bit_reversed ? 1 : 0);
  PK_HIP(hipGetLastError());
}
}  // namespace

void bitrev_permute_dev(const fr_t* in, fr_t* out, uint64_t n, hipStream_t s) {
  PK_REQUIRE(is_pow2(n), PK_ERR_NOT_POW2, "reverse_order: size must be a power of two");
  k_bitrev<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, out, n, log2_exact(n));
  PK_HIP(hipGetLastError());
}

void reverse_order_dev(fr_t* data, uint64_t codeword, hipStream_t s) {
  PK_REQUIRE(is_pow2(codeword), PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  if (codeword <= 1) return;
  k_bitrev_inplace<<<(unsigned)((codeword + 255) / 256), 256, 0, s>>>(data, codeword,
                                                                       log2_exact(codeword));
  PK_HIP(hipGetLastError());
}

void intt_rn_dev(Device& d, fr_t* data, uint64_t n, hipStream_t s) {
  if (n <= 1) return;  // 1/1 scaling of a single element is the identity
  PK_REQUIRE(is_pow2(n), PK_ERR_NOT_POW2, "intt_rn: length must be a power of two");
  intt_dev(d, data, n, true, s);
}

void intt_nr_dev(Device& d, fr_t* data, uint64_t n, hipStream_t s) {
  if (n <= 1) return;  // 1/1 scaling of a single element is the identity
  PK_REQUIRE(is_pow2(n), PK_ERR_NOT_POW2, "codeword_size.is_power_of_two()");
  intt_dev(d, data, n, false, s);
}

namespace {
// The first coset chunk of RSFr::interleaved_encode
// (provekit/common/src/ntt.rs:70-84): message coefficients interleaved
// column-major, masks appended, zeros up to the chunk length.  With a coset
// table, coefficient index i / M is scaled by 5^(i / M).
__global__ void k_rs_chunk(fr_t* chunk, const fr_t* msgs, uint64_t M, uint64_t msg_len,
                           const fr_t* masks, uint64_t masks_len, uint64_t chunk_len,
                           const fr_t* coset) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= chunk_len) return;
  const uint64_t nm = msg_len * M;
  fr_t v;
  if (i < nm) {
    v = fr_load(msgs + (i % M) * msg_len + i / M);
  } else if (i < nm + masks_len) {
    v = fr_load(masks + (i - nm));
  } else {
    for (int k = 0; k < 8; k++) v.l[k] = 0;
  }
  if (coset && i < nm + masks_len) {
    v = fr_mul(v, fr_load(coset + i / M));
    fr_canon(v);
  }
  fr_store(chunk + i, v);
}
__global__ void k_rs_points(const fr_t* rr, const uint64_t* idx, uint64_t n, fr_t* out,
                            int bits, fr_t one) {
  // w^{bitrev_bits(i)} = rr-table entry... computed by square-and-multiply
  // from w = rr[1]^... is not prefix-friendly; use the table of the full
  // order: rr[j] = w^{bitrev_{bits-1}(j)} covers exponents below n/2, and
  // w^{k + n/2} = -w^k.
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t N = 1ull << bits;
  uint64_t i = idx[t];
  uint64_t k = bits == 0 ? i : bitrev(i & (N - 1), bits);
  // exponent k mod N (points repeat with period N)
  k &= N - 1;
  fr_t v;
  if (bits == 0) {
    v = one;
  } else {
    const uint64_t halfN = N >> 1;
    const uint64_t kk = k & (halfN - 1);
    v = rr[bits >= 2 ? bitrev(kk, bits - 1) : 0];
    if (k >= halfN) {
      fr_t z;
      for (int q = 0; q < 8; q++) z.l[q] = 0;
      v = fr_sub_canon(z, v);
    }
  }
  out[t] = v;
}
}  // namespace

void rs_points_dev(Device& d, uint64_t codeword, const uint64_t* idx, uint64_t n, fr_t* out,
                   hipStream_t s) {
  PK_REQUIRE(codeword >= 1 && is_pow2(codeword), PK_ERR_NOT_POW2,
             "evaluation_points: codeword length must be a power of two");
  if (n == 0) return;
  const int bits = log2_exact(codeword);
  fr_t one;
  memcpy(&one, &hf::ONE, 32);
  const fr_t* rr = roots_table(d, bits >= 1 ? bits - 1 : 0);
  k_rs_points<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(rr, idx, n, out, bits, one);
  PK_HIP(hipGetLastError());
}

void rs_encode_dev(Device& d, const fr_t* msgs, uint64_t M, uint64_t msg_len, const fr_t* masks,
                   uint64_t masks_len, uint64_t codeword, fr_t* out, int coset, hipStream_t s) {
  if (M == 0) return;
  // every message column and mask lands inside the M x codeword result, or
  // the reference's indexing panics (provekit/common/src/ntt.rs:70-80)
  PK_REQUIRE(msg_len <= codeword && msg_len * M + masks_len <= codeword * M, PK_ERR_SIZE,
             "interleaved_encode: messages and masks longer than the codeword");
  if (codeword == 0) return;  // empty messages into an empty codeword
  PK_REQUIRE(is_pow2(codeword), PK_ERR_NOT_POW2, "interleaved_encode: codeword length must be a power of two");
  const uint64_t mask_len = masks_len / M;
  uint64_t coset_size = 1;
  while (coset_size < msg_len + mask_len) coset_size <<= 1;
  PK_REQUIRE(coset_size <= (1ull << 28), PK_ERR_SIZE, "interleaved_encode: order above 2^28");
  while (codeword % coset_size) coset_size <<= 1;
  PK_REQUIRE(coset_size <= codeword, PK_ERR_SIZE,
             "interleaved_encode: message longer than the codeword");
  const uint64_t num_cosets = codeword / coset_size;
  const uint64_t chunk_len = coset_size * M;
  const uint64_t total = codeword * M;
  const fr_t* src = msgs;
  Scratch chunk;
  bool direct = (M == 1 && masks_len == 0 && !coset && msg_len == coset_size);
  if (!direct) {
    chunk = Scratch(sizeof(fr_t) * chunk_len, s);
    const fr_t* cpow = nullptr;
    if (coset) {
      uint64_t need = msg_len + mask_len;
      int lg = 0;
      while ((1ull << lg) < need) lg++;
      cpow = coset_table(d, lg);
    }
    k_rs_chunk<<<(unsigned)((chunk_len + 255) / 256), 256, 0, s>>>(
        chunk.as<fr_t>(), msgs, M, msg_len, masks, masks_len, chunk_len, cpow);
    PK_HIP(hipGetLastError());
    src = chunk.as<fr_t>();
  }
  ntt_nr_dev(d, out, total, codeword, num_cosets, s, src, chunk_len);
}

// ---- tables (declared in runtime.hpp) ----
static const fr_t* grow(Device& d, GrowTable& t, int log, int min_log, int max_log,
                        const hfr& base, bool bit_reversed, const char* what) {
  std::lock_guard<std::mutex> lk(d.tables_mtx);
  if (t.log < log) {
    int lg = log < min_log ? min_log : log;
    PK_REQUIRE(lg <= max_log, PK_ERR_SIZE, what);
    fr_t* p;
    PK_HIP(hipMalloc((void**)&p, sizeof(fr_t) << lg));
    hipStream_t bs;
    PK_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    build_power_table(p, bit_reversed ? hf::root_of_unity_log(lg + 1) : base, lg, bit_reversed, bs);
    PK_HIP(hipStreamSynchronize(bs));
    PK_HIP(hipStreamDestroy(bs));
    if (t.ptr) t.retired.push_back(t.ptr);
    t.ptr = p;
    t.log = lg;
  }
  return t.ptr;
}

const fr_t* roots_table(Device& d, int half_log) {
  if (half_log < 0) half_log = 0;
  // At least 2^20 roots (order 2^21, 32 MiB) on first use.
  return grow(d, d.roots, half_log, 20, 27, hfr{}, true, "ntt: order above 2^28");
}

const fr_t* coset_table(Device& d, int log) {
  return grow(d, d.coset, log, 20, 28, hf::from_u64(5), false, "coset: length above 2^28");
}

}  // namespace pk
