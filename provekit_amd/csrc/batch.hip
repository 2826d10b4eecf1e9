// Batched commitment of many polynomials of uneven lengths.
//
//  1. coset scaling of every polynomial's coefficients: one launch;
//  2. the NTT passes of all polynomials: one launch per pass round, the
//     workgroup -> (polynomial, tile) map in a device table (k_ntt_pass_batch);
//  3. all leaf hashing and all Merkle trees: ONE persistent launch fed by a
//     device work queue (k_hash_queue).
//
// The queue holds, in topological order, leaf tasks (64 leaves of one
// polynomial), node tasks (64 k nodes of one tree level, k per lane) and one
// tail task per tree (the last levels from <= 64 nodes to the root inside one
// wave).  A wave takes the next task index from a global counter, waits for
// the tasks that produce its inputs (a consecutive range), runs it and
// publishes a completion flag.  Publishing costs a release fence, which
// writes back the XCD's L2, and reading costs an acquire fence, which
// invalidates it: with one node per lane those fences, not the compressions,
// bound the tree levels (2^23 one-node tasks: 7 ms of the queue's 22.6 ms,
// against 1.1 ms of compressions).  Wide levels therefore give each lane up
// to 8 nodes per task.  Tasks are handed out in topological order, so every task a
// wave waits for was already taken by a running wave that waits only on
// earlier tasks: the earliest unfinished task always progresses, whatever the
// residency.  Each wait is bounded; on timeout the wave raises the error flag
// and leaves, so every wave exits.
//
// The upper tree levels of all polynomials overlap with other polynomials'
// leaf hashing instead of running as a latency-bound tail per polynomial.
#include <algorithm>

#include "batch.hpp"
#include "host_fr.hpp"
#include "kernels.hpp"
#include "sky.hpp"

namespace pk {

namespace {

constexpr int kQueueWave = 64;
constexpr uint64_t kSpinLimit = 1ull << 24;  // x ~64-cycle sleeps: well above any real wait

__global__ __launch_bounds__(256) void k_coset_scale_batch(const BatchPolyDev* polys,
                                                           const uint2* map, const fr_t* coset) {
  const uint2 m = map[blockIdx.x];
  const BatchPolyDev& p = polys[m.x];
  const uint64_t i = (uint64_t)m.y * 256 + threadIdx.x;
  if (i >= p.n) return;
  fr_t v = fr_mul(fr_load(p.coeffs + i), fr_load(coset + i));
  fr_canon(v);
  fr_store(p.chunk + i, v);
}

// A polynomial of one coefficient has no NTT stage: its codeword is that
// coefficient (coset-scaled by 5^0 = 1 when asked) in every position, as the
// per-call encoder's replication leaves it.
__global__ __launch_bounds__(256) void k_replicate_batch(const BatchPolyDev* polys, const uint2* map) {
  const uint2 m = map[blockIdx.x];
  const BatchPolyDev& p = polys[m.x];
  const uint64_t i = (uint64_t)m.y * 256 + threadIdx.x;
  if (i >= p.N) return;
  fr_store(p.codeword + i, fr_load(p.chunk ? p.chunk : p.coeffs));
}

__global__ void k_fill_zero_u32(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Offset of tree level l (level 0 = n0 leaves) in the level-by-level node
// array: sum_{i<l} n0 >> i = 2 n0 - 2 (n0 >> l).
__device__ __forceinline__ uint64_t level_off(uint64_t n0, uint32_t l) {
  return 2 * n0 - 2 * (n0 >> l);
}

// Workgroups of four waves, each wave an independent worker taking tasks:
// the dispatcher places one wave per SIMD per workgroup (one-wave
// workgroups can stack several waves on one SIMD while another idles).  No
// workgroup barrier: broadcasts and the tail's level exchange are wave-level.
// Every task kind runs as a sequence of steps with one compression each, so
// the kernel holds a single copy of the unrolled compression (three copies
// would overflow the instruction cache); at most 128 VGPRs keep the 16 waves
// per CU of the launch resident.
constexpr int kQueueBlock = 256;

__global__ __launch_bounds__(kQueueBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_hash_queue(
    const BatchPolyDev* polys, const HashTask* tasks, uint32_t ntasks, uint32_t* done,
    uint32_t* next, uint32_t* err) {
  // the compression's reduction table, filled before the waves go their own way
  __shared__ SkyTab tab;
  sky_tab_fill(tab);
  __syncthreads();
  const int t = threadIdx.x & (kQueueWave - 1);
  for (;;) {
    const uint32_t id = wave_take(next);
    if (id >= ntasks) return;
    const HashTask tk = tasks[id];
    const BatchPolyDev& p = polys[tk.poly];
    if (!wave_wait_range(done, tk.dep_first, tk.dep_count, err, kSpinLimit)) return;
    // the dependency flags were seen with relaxed loads; one acquire fence
    // (it invalidates this XCD's L2) makes the data they guard visible to the
    // wave.  Leaf tasks read only codewords written by earlier launches and
    // skip it.
    if (tk.dep_first >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // The task as steps of one compression each (wave-uniform count):
    //  leaf: acc = m0, then acc = C(acc, m_k), k = 1 .. leaf_elems - 1, the
    //        words read four at a time (one 128-byte line per lane; see
    //        sky_hash_message) through a shifting window;
    //  node: k steps, C(in[2i], in[2i + 1]) for the lane's nodes
    //        i = 64 k j + 64 step + lane;
    //  tail: level tk.level (<= 64 nodes, one per lane) up to the root, one
    //        step per level, operands from the level below by shuffles.
    const uint32_t kind = tk.kind;
    uint64_t lane_i = (uint64_t)tk.j * kQueueWave + t;  // leaf or node of this lane
    bool act = false;
    int nsteps = 0;
    fr_t acc = {}, w[4] = {};
    int have = 0;            // words left in the window (leaf)
    uint64_t next_word = 0;  // first word of the next window (leaf)
    const fr_t* src = nullptr;
    fr_t* dst = nullptr;
    uint64_t n = 0;  // tail: nodes at the current level
    if (kind == kTaskLeaf) {
      act = lane_i < p.leaves;
      src = p.codeword + (act ? lane_i : 0) * p.leaf_elems;
      const uint64_t E = p.leaf_elems;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if ((uint64_t)q < E) w[q] = fr_load(src + q);
      acc = w[0];
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
      have = E < 4 ? (int)E - 1 : 3;
      next_word = 4;
      nsteps = (int)E - 1;
      dst = p.nodes + lane_i;
    } else if (kind == kTaskNode) {
      n = p.leaves >> tk.level;  // nodes at this level
      lane_i = ((uint64_t)tk.j << (6 + tk.k_log)) + t;
      src = p.nodes + level_off(p.leaves, tk.level - 1);
      dst = p.nodes + level_off(p.leaves, tk.level);
      nsteps = 1 << tk.k_log;
    } else {
      n = p.leaves >> tk.level;
      const uint64_t off = level_off(p.leaves, tk.level);
      if ((uint64_t)t < n) acc = fr_load(p.nodes + off + t);
      dst = p.nodes + off + n;  // the next level starts here
      nsteps = 0;
      for (uint64_t m = n; m > 1; m >>= 1) nsteps++;
    }
#pragma unroll 1
    for (int step = 0; step < nsteps; step++) {
      fr_t x, y;
      bool on;
      if (kind == kTaskLeaf) {
        if (have == 0) {  // the next four words (wave-uniform)
          const uint64_t E = p.leaf_elems;
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (next_word + q < E) w[q] = fr_load(src + next_word + q);
          have = E - next_word < 4 ? (int)(E - next_word) : 4;
          next_word += 4;
        }
        x = acc;
        y = w[0];
        w[0] = w[1];
        w[1] = w[2];
        w[2] = w[3];
        have--;
        on = act;
      } else if (kind == kTaskNode) {
        const uint64_t i = lane_i + (uint64_t)step * kQueueWave;
        on = i < n;
        const fr_t* in = src + 2 * (on ? i : 0);
        x = fr_load(in);
        y = fr_load(in + 1);
      } else {
        x = fr_shfl(acc, (2 * t) & (kQueueWave - 1));
        y = fr_shfl(acc, (2 * t + 1) & (kQueueWave - 1));
        on = (uint64_t)t < (n >> 1);
      }
      fr_t v = {};
      if (on) v = sky_compress_t(x, y, tab);
      if (kind == kTaskLeaf) {
        acc = v;
      } else if (kind == kTaskNode) {
        if (on) fr_store(dst + lane_i + (uint64_t)step * kQueueWave, v);
      } else {
        if (on) fr_store(dst + t, v);
        acc = v;
        dst += n >> 1;
        n >>= 1;
      }
    }
    if (kind == kTaskLeaf && act) fr_store(dst, acc);
    // publish: the wave's stores reach device-coherent memory (a release
    // fence writes back this XCD's L2; no invalidate) before the flag, which
    // every lane writes (one store of one address)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(done + id, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_gather_roots(const BatchPolyDev* polys, uint32_t count, fr_t* roots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const BatchPolyDev& p = polys[i];
  fr_store(roots + i, fr_load(p.nodes + 2 * p.leaves - 2));
}

template <class T>
T* upload_vec(const std::vector<T>& v, std::vector<void*>& owned) {
  void* p = nullptr;
  const size_t bytes = sizeof(T) * (v.empty() ? 1 : v.size());
  PK_HIP(hipMalloc(&p, bytes));
  owned.push_back(p);
  if (!v.empty()) PK_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return static_cast<T*>(p);
}

}  // namespace

BatchPlan::~BatchPlan() {
  if (graph) (void)hipGraphExecDestroy(graph);
  for (void* p : owned) (void)hipFree(p);
}

BatchPlan* batch_plan_create(Device& d, const std::vector<BatchPolySpec>& specs, int tile_log) {
  PK_REQUIRE(!specs.empty(), PK_ERR_ARG, "batch: no polynomials");
  PK_REQUIRE(tile_log == 0 || (tile_log >= 6 && tile_log <= 11), PK_ERR_ARG,
             "batch: tile log must be 0 (tuned) or 6..11");
  auto plan = std::make_unique<BatchPlan>();
  plan->count = (uint32_t)specs.size();
  std::vector<BatchPolyDev> polys(specs.size());
  std::vector<uint2> coset_map, rep_map;
  uint64_t chunk_total = 0, max_n = 0;
  for (size_t i = 0; i < specs.size(); i++) {
    const BatchPolySpec& s = specs[i];
    PK_REQUIRE(is_pow2(s.n) && is_pow2(s.blowup) && s.coeffs && s.codeword && s.nodes,
               PK_ERR_NOT_POW2, "batch: lengths and blowups must be powers of two");
    PK_REQUIRE(s.leaf_elems >= 2 && is_pow2(s.leaf_elems) && s.n * s.blowup >= s.leaf_elems,
               PK_ERR_SIZE, "batch: leaf size must be a power of two >= 2 dividing the codeword");
    PK_REQUIRE(s.n * s.blowup <= (1ull << 28), PK_ERR_SIZE, "batch: codeword above 2^28");
    if (s.coset) chunk_total += s.n;
    max_n = std::max<uint64_t>(max_n, s.n);
  }
  fr_t* chunks = nullptr;
  if (chunk_total) {
    PK_HIP(hipMalloc((void**)&chunks, sizeof(fr_t) * chunk_total));
    plan->owned.push_back(chunks);
  }
  const fr_t* coset = nullptr;
  if (chunk_total) coset = coset_table(d, log2_exact(max_n));
  plan->coset = coset;
  uint64_t chunk_off = 0;
  // NTT plans per polynomial
  std::vector<std::vector<NttPassDesc>> nplans(specs.size());
  size_t max_rounds = 0;
  for (size_t i = 0; i < specs.size(); i++) {
    const BatchPolySpec& s = specs[i];
    BatchPolyDev& p = polys[i];
    p.coeffs = s.coeffs;
    p.codeword = s.codeword;
    p.nodes = s.nodes;
    p.n = s.n;
    p.N = s.n * s.blowup;
    p.leaf_elems = s.leaf_elems;
    p.leaves = p.N / s.leaf_elems;
    p.chunk = s.coset ? chunks + chunk_off : nullptr;
    if (s.coset) {
      for (uint64_t b = 0; b * 256 < s.n; b++) coset_map.push_back(make_uint2((uint32_t)i, (uint32_t)b));
      chunk_off += s.n;
    }
    const fr_t* src = s.coset ? p.chunk : s.coeffs;
    NttShape shape;
    shape.tile_log = tile_log;  // 0: the tuned shape of this length
    nplans[i] = ntt_plan(d, s.codeword, p.N, p.N, s.blowup, src, s.n, shape);
    if (nplans[i].empty())  // n = 1: no stage, the codeword is the coefficient replicated
      for (uint64_t b = 0; b * 256 < p.N; b++) rep_map.push_back(make_uint2((uint32_t)i, (uint32_t)b));
    max_rounds = std::max(max_rounds, nplans[i].size());
  }
  plan->polys = upload_vec(polys, plan->owned);
  plan->coset_map = upload_vec(coset_map, plan->owned);
  plan->coset_blocks = (uint32_t)coset_map.size();
  plan->rep_map = upload_vec(rep_map, plan->owned);
  plan->rep_blocks = (uint32_t)rep_map.size();
  // pass rounds, grouped by kernel shape (tile size, butterflies per thread)
  std::vector<NttPassDesc> descs;
  for (size_t r = 0; r < max_rounds; r++) {
    std::vector<std::pair<int, int>> shapes;
    for (size_t i = 0; i < specs.size(); i++) {
      if (r >= nplans[i].size()) continue;
      const std::pair<int, int> sh(nplans[i][r].tl, nplans[i][r].bpt);
      if (std::find(shapes.begin(), shapes.end(), sh) == shapes.end()) shapes.push_back(sh);
    }
    std::sort(shapes.begin(), shapes.end());
    for (const auto& sh : shapes) {
      std::vector<uint2> map;
      for (size_t i = 0; i < specs.size(); i++) {
        if (r >= nplans[i].size() || nplans[i][r].tl != sh.first || nplans[i][r].bpt != sh.second)
          continue;
        const uint32_t di = (uint32_t)descs.size();
        descs.push_back(nplans[i][r]);
        const uint64_t tiles = polys[i].N >> sh.first;
        const NttPassDesc& a = nplans[i][r];
        const uint64_t T = a.Q >> a.jb;
        for (uint64_t t = 0; t < tiles; t++)
          map.push_back(make_uint2(di, (uint32_t)ntt_tile_order(a.src != nullptr, T, tiles / T, t)));
      }
      BatchPlan::Round rd;
      rd.tl = sh.first;
      rd.bpt = sh.second;
      rd.blocks = (uint32_t)map.size();
      rd.map = upload_vec(map, plan->owned);
      plan->rounds.push_back(rd);
    }
  }
  plan->descs = upload_vec(descs, plan->owned);
  // hashing queue, topological: all leaf tasks, node tasks level by level, tails.
  // prev_first[i] / prev_width[i]: the first task id and the nodes per task
  // of polynomial i's previous level (its tasks are consecutive).
  std::vector<HashTask> tasks;
  std::vector<int32_t> prev_first(specs.size());
  std::vector<uint64_t> prev_width(specs.size(), kQueueWave), prev_count(specs.size());
  int max_levels = 0;
  std::vector<int> tail_level(specs.size());
  for (size_t i = 0; i < specs.size(); i++) {
    const uint64_t n0 = polys[i].leaves;
    const int lg = log2_exact(n0);
    tail_level[i] = lg > 6 ? lg - 6 : 0;  // first level with <= 64 nodes
    max_levels = std::max(max_levels, tail_level[i]);
    prev_first[i] = (int32_t)tasks.size();
    for (uint64_t j = 0; j * kQueueWave < n0; j++)
      tasks.push_back(HashTask{(uint32_t)i, 0, (uint32_t)j, kTaskLeaf, -1, 0, 0, 0});
    prev_count[i] = tasks.size() - (size_t)prev_first[i];
  }
  for (int l = 1; l <= max_levels; l++) {
    for (size_t i = 0; i < specs.size(); i++) {
      if (l > tail_level[i]) continue;
      const uint64_t nl = polys[i].leaves >> l;
      // nodes per lane: up to 8 while the level keeps >= 32 tasks
      uint32_t k_log = 0;
      while (k_log < 3 && (nl >> (6 + k_log + 1)) >= 32) k_log++;
      const uint64_t width = (uint64_t)kQueueWave << k_log;
      const int32_t first = (int32_t)tasks.size();
      for (uint64_t j = 0; j * width < nl; j++) {
        // inputs: nodes [2 width j, 2 width (j + 1)) of the level below
        const uint64_t t0 = (2 * width * j) / prev_width[i];
        const uint64_t t1 = std::min<uint64_t>((2 * width * (j + 1) + prev_width[i] - 1) / prev_width[i],
                                               prev_count[i]);
        tasks.push_back(HashTask{(uint32_t)i, (uint32_t)l, (uint32_t)j, kTaskNode,
                                 prev_first[i] + (int32_t)t0, (uint32_t)(t1 - t0), k_log, 0});
      }
      prev_first[i] = first;
      prev_width[i] = width;
      prev_count[i] = tasks.size() - (size_t)first;
    }
  }
  for (size_t i = 0; i < specs.size(); i++) {
    if (polys[i].leaves <= 1) continue;
    // level tail_level[i] (<= 64 nodes) was produced by one task
    tasks.push_back(HashTask{(uint32_t)i, (uint32_t)tail_level[i], 0, kTaskTail, prev_first[i],
                             (uint32_t)prev_count[i], 0, 0});
  }
  plan->ntasks = (uint32_t)tasks.size();
  plan->tasks = upload_vec(tasks, plan->owned);
  std::vector<uint32_t> zeros(plan->ntasks + 2, 0);
  plan->done = upload_vec(zeros, plan->owned);  // [0, ntasks) flags, then counter, error
  plan->next = plan->done + plan->ntasks;
  plan->err = plan->next + 1;
  fr_t* roots;
  PK_HIP(hipMalloc((void**)&roots, sizeof(fr_t) * specs.size()));
  plan->owned.push_back(roots);
  plan->roots = roots;
  int dev;
  PK_HIP(hipGetDevice(&dev));
  int cus = 0;
  PK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  plan->queue_blocks = (uint32_t)cus * 4;  // 16 waves per CU
  return plan.release();
}

void batch_run(BatchPlan& plan, hipStream_t s) {
  const uint64_t nflags = (uint64_t)plan.ntasks + 2;
  k_fill_zero_u32<<<(unsigned)((nflags + 255) / 256), 256, 0, s>>>(plan.done, nflags);
  if (plan.coset_blocks)
    k_coset_scale_batch<<<plan.coset_blocks, 256, 0, s>>>(plan.polys, plan.coset_map, plan.coset);
  PK_HIP(hipGetLastError());
  if (plan.rep_blocks) k_replicate_batch<<<plan.rep_blocks, 256, 0, s>>>(plan.polys, plan.rep_map);
  for (const BatchPlan::Round& rd : plan.rounds) ntt_batch_round_dev(plan.descs, rd.map, rd.blocks, rd.tl, rd.bpt, s);
  k_hash_queue<<<plan.queue_blocks, kQueueBlock, 0, s>>>(plan.polys, plan.tasks, plan.ntasks,
                                                         plan.done, plan.next, plan.err);
  k_gather_roots<<<(plan.count + 63) / 64, 64, 0, s>>>(plan.polys, plan.count, plan.roots);
  PK_HIP(hipGetLastError());
}

hipGraphExec_t batch_graph(BatchPlan& plan) {
  if (plan.graph) return plan.graph;
  hipStream_t cs;
  PK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  PK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  try {
    batch_run(plan, cs);
  } catch (...) {
    hipGraph_t junk = nullptr;
    (void)hipStreamEndCapture(cs, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    (void)hipStreamDestroy(cs);
    throw;
  }
  PK_HIP(hipStreamEndCapture(cs, &g));
  PK_HIP(hipGraphInstantiate(&plan.graph, g, nullptr, nullptr, 0));
  PK_HIP(hipGraphDestroy(g));
  PK_HIP(hipStreamDestroy(cs));
  return plan.graph;
}

}  // namespace pk
