// Batched commitment of many polynomials (batch.hip).
#pragma once
#include <memory>
#include <vector>

#include "kernels.hpp"

namespace pk {

// One polynomial of a batch (device pointers).
struct BatchPolySpec {
  const fr_t* coeffs;
  uint64_t n;
  uint64_t blowup;
  int coset;
  uint64_t leaf_elems;
  fr_t* codeword;  // n * blowup
  fr_t* nodes;     // 2 * leaves - 1
};

struct BatchPolyDev {
  const fr_t* coeffs;
  fr_t* chunk;     // coset-scaled coefficients, or null
  fr_t* codeword;
  fr_t* nodes;
  uint64_t n, N, leaves, leaf_elems;
};

enum : uint32_t { kTaskLeaf = 0, kTaskNode = 1, kTaskTail = 2 };
// A leaf task hashes leaves [64 j, 64 j + 64); a node task computes nodes
// [64 k j, 64 k (j + 1)) of one tree level, k = 1 << k_log per lane; a tail
// task runs the last levels (<= 64 nodes) to the root.  Its inputs come from
// tasks [dep_first, dep_first + dep_count), consecutive in the queue.
struct HashTask {
  uint32_t poly, level, j, kind;
  int32_t dep_first;   // -1: no inputs from other tasks
  uint32_t dep_count;
  uint32_t k_log;      // node tasks: log2 nodes per lane
  uint32_t pad;
};

struct BatchPlan {
  struct Round {
    int tl, bpt;
    uint32_t blocks;
    const uint2* map;
  };
  uint32_t count = 0;
  BatchPolyDev* polys = nullptr;
  const uint2* coset_map = nullptr;
  uint32_t coset_blocks = 0;
  const uint2* rep_map = nullptr;  // (polynomial, 256-element block) of single-coefficient codewords
  uint32_t rep_blocks = 0;
  const fr_t* coset = nullptr;
  NttPassDesc* descs = nullptr;
  std::vector<Round> rounds;
  HashTask* tasks = nullptr;
  uint32_t ntasks = 0;
  uint32_t* done = nullptr;  // ntasks completion flags, then the queue counter, then the error flag
  uint32_t* next = nullptr;
  uint32_t* err = nullptr;
  uint32_t queue_blocks = 0;
  fr_t* roots = nullptr;     // count roots after a run
  hipGraphExec_t graph = nullptr;
  std::vector<void*> owned;
  ~BatchPlan();
};

BatchPlan* batch_plan_create(Device& d, const std::vector<BatchPolySpec>& specs, int tile_log);
// Enqueue one batched commit of every polynomial of the plan.
void batch_run(BatchPlan& plan, hipStream_t s);
// The same launches recorded once into an executable graph (cached in the plan).
hipGraphExec_t batch_graph(BatchPlan& plan);

}  // namespace pk
