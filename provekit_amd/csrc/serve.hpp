// Serving loop: continuous batching of R1CS sumcheck + witness-commit jobs
// on one GPU (serve.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "kernels.hpp"
#include "sumcheck.hpp"

namespace pk {

struct JobSpec {
  int m;                       // 2^m constraints, witness of 2^m elements
  const uint32_t* rs[3];       // A, B, C row starts (2^m each)
  const uint32_t* cols[3];
  const fr_t* vals[3];
  uint64_t nnz[3];
  const fr_t* z;               // 2^m
  uint64_t seed;               // transcript seed
  uint64_t blowup;             // witness commit
  int coset;
  uint64_t leaf_elems;
};

struct JobResult {
  fr_t root;
  fr_t final_claim;
  std::vector<fr_t> coeffs, alphas, r;
  double latency_ms = 0, queue_ms = 0, run_ms = 0;
  int status = 0;  // 0 ok, else a Status code
  std::string error;
};

class Server {
 public:
  Server(int device, uint64_t hbm_budget, int slots);
  ~Server();
  // Returns the job id; throws PK_ERR_BUDGET if the job can never fit.
  uint64_t submit(const JobSpec& job);
  // Blocks until the job finished; the result stays available until release().
  // timeout_ms <= 0 waits without limit; else throws PK_ERR_STATE on timeout.
  JobResult wait(uint64_t id, double timeout_ms = 0);
  bool poll(uint64_t id);
  // Leaf openings from the job's resident codeword and tree.
  void open(uint64_t id, const uint64_t* leaves, size_t n, fr_t* leaf_values, fr_t* paths);
  // Frees the job's resident codeword and tree.
  void release(uint64_t id);
  uint64_t footprint(const JobSpec& job) const;
  uint64_t bytes_in_use();
  uint64_t peak_in_flight();

 private:
  struct Job;
  void scheduler();
  void launch(Job& j);
  void finish(Job& j, int slot);
  static void on_done(void* arg);

  int device_;
  uint64_t budget_;
  std::vector<hipStream_t> streams_;
  std::vector<bool> stream_busy_;
  std::vector<ScState*> host_state_;  // pinned, one per slot
  std::vector<fr_t*> host_root_;
  std::mutex mu_;
  std::condition_variable cv_;        // scheduler wake-ups
  // Bumped under mu_ by every change the scheduler acts on (a submit, a job
  // done, a release, shut-down, a failed launch); the scheduler sleeps only
  // while it is unchanged since its last scan, so an event that lands while
  // it is launching (mu_ released) is not lost.
  uint64_t events_ = 0;
  std::condition_variable done_cv_;   // waiters
  std::deque<uint64_t> pending_;
  std::map<uint64_t, std::unique_ptr<Job>> jobs_;
  uint64_t next_id_ = 1;
  uint64_t in_use_ = 0;
  uint64_t in_flight_ = 0, peak_in_flight_ = 0;
  bool stop_ = false;
  std::thread thread_;
};

}  // namespace pk
