// Serving loop for R1CS sumcheck + witness-commit jobs (continuous batching).
//
// A scheduler thread admits pending jobs in arrival order whenever a stream
// slot is free and the job's HBM footprint fits the remaining budget.  An
// admitted job runs on its own stream: Az, Bz, Cz (SpMV), the whole sumcheck
// over eq(tau, x)(Az Bz - Cz) with the on-device transcript, and the commit of
// the witness as WhirR1CSScheme::commit shapes it (zero-padded to at least
// 2^13 coefficients), RS-NTT + leaves + Skyscraper tree.  Jobs on different streams
// overlap on the device.  Completion is signalled by a host function enqueued
// behind the job's work (no HIP call inside it); the scheduler then frees the
// job's transient buffers, records its latency and admits more work.  The
// codeword and tree stay resident (and counted against the budget) until the
// client releases the job, so Merkle openings are served from HBM.
#include <algorithm>

#include "host_fr.hpp"
#include "serve.hpp"

namespace pk {

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

uint64_t align_up(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// The witness commitment has 2^max(m, 13) coefficients: WhirR1CSScheme pads
// the witness to at least MIN_WHIR_NUM_VARIABLES = 13 variables
// (provekit/r1cs-compiler/src/whir_r1cs.rs:7,69-73).
constexpr int kMinWhirVars = 13;
int commit_vars(const JobSpec& j) { return std::max(j.m, kMinWhirVars); }
uint64_t commit_len(const JobSpec& j) { return 1ull << commit_vars(j); }

}  // namespace

struct Server::Job {
  uint64_t id = 0;
  JobSpec spec;
  uint64_t transient_bytes = 0, resident_bytes = 0;
  enum State { Pending, Running, Done } state = Pending;
  int stream = -1;
  Clock::time_point t_submit, t_start, t_done;
  void* transient = nullptr;
  fr_t* codeword = nullptr;
  fr_t* nodes = nullptr;
  bool failed = false;
  int readers = 0;  // open() calls using the resident buffers right now
  JobResult result;
  Server* srv = nullptr;
};

Server::Server(int device, uint64_t hbm_budget, int slots) : device_(device), budget_(hbm_budget) {
  PK_REQUIRE(slots >= 1 && slots <= 64, PK_ERR_ARG, "server: slots must be 1..64");
  PK_HIP(hipSetDevice(device_));
  streams_.resize(slots);
  stream_busy_.assign(slots, false);
  host_state_.assign(slots, nullptr);
  host_root_.assign(slots, nullptr);
  for (int i = 0; i < slots; i++) {
    PK_HIP(hipStreamCreateWithFlags(&streams_[i], hipStreamNonBlocking));
    // per-slot pinned result buffers, allocated once (no host-memory calls
    // on the job path)
    PK_HIP(hipHostMalloc((void**)&host_state_[i], sizeof(ScState), 0));
    PK_HIP(hipHostMalloc((void**)&host_root_[i], sizeof(fr_t), 0));
  }
  thread_ = std::thread([this] { scheduler(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
    events_++;
    // jobs never admitted are dropped; running ones finish first
    for (uint64_t id : pending_) {
      Job& j = *jobs_.at(id);
      j.failed = true;
      j.result.status = PK_ERR_STATE;
      j.result.error = "server shut down";
      j.state = Job::Done;
    }
    pending_.clear();
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  done_cv_.notify_all();
  for (auto& kv : jobs_) {
    Job& j = *kv.second;
    if (j.codeword) (void)hipFreeAsync(j.codeword, streams_[0]);
    if (j.nodes) (void)hipFreeAsync(j.nodes, streams_[0]);
  }
  (void)hipStreamSynchronize(streams_[0]);
  for (size_t i = 0; i < streams_.size(); i++) {
    (void)hipStreamDestroy(streams_[i]);
    (void)hipHostFree(host_state_[i]);
    (void)hipHostFree(host_root_[i]);
  }
}

uint64_t Server::footprint(const JobSpec& job) const {
  const uint64_t n = 1ull << job.m;
  uint64_t t = 0;
  for (int k = 0; k < 3; k++) t += align_up(4 * n) + align_up(4 * job.nnz[k]) + align_up(32 * job.nnz[k]);
  t += align_up(32 * n);                  // z
  t += 4 * align_up(32 * n);              // a, b, c, eq
  t += align_up(sizeof(ScState)) + align_up(32 * (sc_partials_elems() + 3));
  const uint64_t cn = commit_len(job);
  if (cn > n) t += align_up(32 * cn);     // the zero-padded witness polynomial
  t += align_up(32 * cn);                 // coset chunk of the commit
  return t;
}

uint64_t Server::submit(const JobSpec& spec) {
  PK_REQUIRE(spec.m >= 1 && spec.m <= kMaxSumcheckVars, PK_ERR_SIZE, "server: m must be 1..28");
  PK_REQUIRE(is_pow2(spec.blowup) && is_pow2(spec.leaf_elems) && spec.leaf_elems >= 2 &&
                 commit_len(spec) * spec.blowup >= spec.leaf_elems &&
                 commit_len(spec) * spec.blowup <= (1ull << 28),
             PK_ERR_SIZE, "server: bad commit shape");
  for (int k = 0; k < 3; k++)
    PK_REQUIRE(spec.rs[k] && (spec.nnz[k] == 0 || (spec.cols[k] && spec.vals[k])), PK_ERR_ARG,
               "server: null matrix");
  PK_REQUIRE(spec.z, PK_ERR_ARG, "server: null witness");
  auto j = std::make_unique<Job>();
  j->spec = spec;
  j->srv = this;
  j->transient_bytes = footprint(spec);
  const uint64_t N = commit_len(spec) * spec.blowup;
  const uint64_t leaves = N / spec.leaf_elems;
  j->resident_bytes = align_up(32 * N) + align_up(32 * (2 * leaves - 1));
  PK_REQUIRE(j->transient_bytes + j->resident_bytes <= budget_, PK_ERR_BUDGET,
             "server: job footprint exceeds the HBM budget");
  j->t_submit = Clock::now();
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    id = next_id_++;
    j->id = id;
    pending_.push_back(id);
    jobs_.emplace(id, std::move(j));
    events_++;
  }
  cv_.notify_all();
  return id;
}

void Server::on_done(void* arg) {
  Job* j = static_cast<Job*>(arg);
  Server* s = j->srv;
  {
    std::lock_guard<std::mutex> lk(s->mu_);
    j->t_done = Clock::now();
    j->state = Job::Done;  // results are in pinned memory; scheduler finishes bookkeeping
    s->events_++;
  }
  s->cv_.notify_all();
}

void Server::launch(Job& j) {
  hipStream_t s = streams_[j.stream];
  const JobSpec& sp = j.spec;
  const uint64_t n = 1ull << sp.m;
  const uint64_t cn = commit_len(sp);
  const uint64_t N = cn * sp.blowup;
  const uint64_t leaves = N / sp.leaf_elems;
  ScState* hst = host_state_[j.stream];
  fr_t* hroot = host_root_[j.stream];
  try {
    PK_HIP(hipMallocAsync(&j.transient, j.transient_bytes, s));
    PK_HIP(hipMallocAsync((void**)&j.codeword, 32 * N, s));
    PK_HIP(hipMallocAsync((void**)&j.nodes, 32 * (2 * leaves - 1), s));
    char* p = static_cast<char*>(j.transient);
    auto take = [&](uint64_t bytes) {
      char* q = p;
      p += align_up(bytes);
      return q;
    };
    CsrDev M[3];
    for (int k = 0; k < 3; k++) {
      uint32_t* rs = (uint32_t*)take(4 * n);
      uint32_t* cols = (uint32_t*)take(4 * sp.nnz[k]);
      fr_t* vals = (fr_t*)take(32 * sp.nnz[k]);
      PK_HIP(hipMemcpyAsync(rs, sp.rs[k], 4 * n, hipMemcpyDefault, s));
      if (sp.nnz[k]) {
        PK_HIP(hipMemcpyAsync(cols, sp.cols[k], 4 * sp.nnz[k], hipMemcpyDefault, s));
        PK_HIP(hipMemcpyAsync(vals, sp.vals[k], 32 * sp.nnz[k], hipMemcpyDefault, s));
      }
      M[k] = CsrDev{n, n, sp.nnz[k], rs, cols, vals};
    }
    fr_t* z = (fr_t*)take(32 * n);
    PK_HIP(hipMemcpyAsync(z, sp.z, 32 * n, hipMemcpyDefault, s));
    fr_t* abc[4];
    for (int k = 0; k < 4; k++) abc[k] = (fr_t*)take(32 * n);
    ScState* st = (ScState*)take(sizeof(ScState));
    fr_t* work = (fr_t*)take(32 * (sc_partials_elems() + 3));
    fr_t* poly = cn > n ? (fr_t*)take(32 * cn) : z;  // shaped in place when no padding
    for (int k = 0; k < 3; k++) spmv_dev(M[k], z, abc[k], s);
    sc_state_init_host(*hst, sp.m, 0, nullptr);
    PK_HIP(hipMemcpyAsync(st, hst, sizeof(ScState), hipMemcpyHostToDevice, s));
    sc_prove_dev(st, abc[0], abc[1], abc[2], abc[3], sp.m, 0, sp.seed, work, s);
    // WhirR1CSScheme::commit of the witness (plain mode: no blinding
    // coefficients), after the SpMVs have read z
    whir_r1cs_commit_dev(device_get(device_), z, n, commit_vars(sp), nullptr, 0, sp.blowup,
                         sp.coset, sp.leaf_elems, poly, j.codeword, j.nodes, s);
    PK_HIP(hipMemcpyAsync(hst, st, sizeof(ScState), hipMemcpyDeviceToHost, s));
    PK_HIP(hipMemcpyAsync(hroot, j.nodes + (2 * leaves - 2), 32, hipMemcpyDeviceToHost, s));
    PK_HIP(hipFreeAsync(j.transient, s));
    j.transient = nullptr;
    PK_HIP(hipLaunchHostFunc(s, &Server::on_done, &j));
  } catch (...) {
    // undo this job's allocations in stream order, then let the caller mark it failed
    if (j.transient) (void)hipFreeAsync(j.transient, s);
    if (j.codeword) (void)hipFreeAsync(j.codeword, s);
    if (j.nodes) (void)hipFreeAsync(j.nodes, s);
    j.transient = nullptr;
    j.codeword = nullptr;
    j.nodes = nullptr;
    throw;
  }
}

void Server::finish(Job& j, int slot) {
  JobResult& r = j.result;
  r.latency_ms = ms_between(j.t_submit, j.t_done);
  r.queue_ms = ms_between(j.t_submit, j.t_start);
  r.run_ms = ms_between(j.t_start, j.t_done);
  if (j.failed) return;
  const ScState& st = *host_state_[slot];
  r.root = *host_root_[slot];
  r.final_claim = st.S;
  const int m = j.spec.m;
  r.coeffs.assign(&st.coeffs[0][0], &st.coeffs[0][0] + 4 * m);
  r.alphas.assign(st.alpha, st.alpha + m);
  r.r.assign(st.r, st.r + m);
  if (st.error) {
    r.status = PK_ERR_INTERNAL;
    r.error = "sumcheck round identity failed";
  }
}

void Server::scheduler() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    const uint64_t seen = events_;  // what this pass accounts for
    // finished jobs: free their slot and transient budget
    for (auto& kv : jobs_) {
      Job& j = *kv.second;
      if (j.state == Job::Done && j.stream >= 0) {
        const int slot = j.stream;
        finish(j, slot);  // copies out of the slot's pinned buffers, no HIP call
        stream_busy_[slot] = false;
        j.stream = -1;
        in_use_ -= j.transient_bytes;
        if (j.failed) in_use_ -= j.resident_bytes;
        in_flight_--;
        done_cv_.notify_all();
      }
    }
    // admit in arrival order while a slot and the budget allow
    while (!pending_.empty()) {
      Job& j = *jobs_.at(pending_.front());
      auto it = std::find(stream_busy_.begin(), stream_busy_.end(), false);
      if (it == stream_busy_.end() || in_use_ + j.transient_bytes + j.resident_bytes > budget_) break;
      pending_.pop_front();
      j.stream = (int)(it - stream_busy_.begin());
      stream_busy_[j.stream] = true;
      in_use_ += j.transient_bytes + j.resident_bytes;
      in_flight_++;
      peak_in_flight_ = std::max(peak_in_flight_, in_flight_);
      j.state = Job::Running;
      j.t_start = Clock::now();
      lk.unlock();
      try {
        std::shared_lock<std::shared_mutex> cap(capture_mutex());  // no capture meanwhile
        launch(j);
      } catch (const std::exception& e) {
        const Error* pe = dynamic_cast<const Error*>(&e);
        lk.lock();
        j.failed = true;
        j.result.status = pe ? pe->code : PK_ERR_INTERNAL;
        j.result.error = e.what();
        j.t_done = Clock::now();
        j.state = Job::Done;  // finished (as failed) on the next pass of the loop
        events_++;
        continue;
      }
      lk.lock();
    }
    if (stop_ && in_flight_ == 0 && pending_.empty()) break;
    // a job that completed while launch() ran unlocked has bumped events_:
    // scan again instead of sleeping past its completion
    cv_.wait(lk, [&] { return events_ != seen; });
  }
}

JobResult Server::wait(uint64_t id, double timeout_ms) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  PK_REQUIRE(it != jobs_.end(), PK_ERR_STATE, "server: unknown job");
  Job& j = *it->second;
  auto ready = [&] { return j.state == Job::Done && j.stream < 0; };
  if (timeout_ms > 0) {
    const bool ok = done_cv_.wait_for(
        lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
    PK_REQUIRE(ok, PK_ERR_STATE, "server: wait timed out");
  } else {
    done_cv_.wait(lk, ready);
  }
  return j.result;
}

bool Server::poll(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  PK_REQUIRE(it != jobs_.end(), PK_ERR_STATE, "server: unknown job");
  return it->second->state == Job::Done && it->second->stream < 0;
}

void Server::open(uint64_t id, const uint64_t* leaves_idx, size_t n, fr_t* values, fr_t* paths) {
  const fr_t* codeword;
  const fr_t* nodes;
  uint64_t N, leaf_elems;
  Job* j;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    PK_REQUIRE(it != jobs_.end(), PK_ERR_STATE, "server: unknown job");
    j = it->second.get();
    PK_REQUIRE(j->state == Job::Done && j->stream < 0 && j->codeword && !j->failed, PK_ERR_STATE,
               "server: job not finished, failed or released");
    codeword = j->codeword;
    nodes = j->nodes;
    N = commit_len(j->spec) * j->spec.blowup;
    leaf_elems = j->spec.leaf_elems;
    j->readers++;  // release() waits until the buffers are no longer read
  }
  // drop the reader mark on every exit path, after the stream work is done
  struct Reader {
    Server* s;
    Job* j;
    ~Reader() {
      {
        std::lock_guard<std::mutex> lk(s->mu_);
        j->readers--;
      }
      s->done_cv_.notify_all();
    }
  } reader{this, j};
  if (!n) return;
  const uint64_t leaves = N / leaf_elems;
  const int depth = log2_exact(leaves);
  for (size_t q = 0; q < n; q++)
    PK_REQUIRE(leaves_idx[q] < leaves, PK_ERR_ARG, "server: leaf index out of range");
  hipStream_t s = thread_stream();
  {
    Scratch di(8 * n, s), dv(32 * n * leaf_elems, s), dp(32 * n * (depth ? depth : 1), s);
    PK_HIP(hipMemcpyAsync(di.p, leaves_idx, 8 * n, hipMemcpyHostToDevice, s));
    merkle_open_dev(codeword, nodes, leaves, leaf_elems, di.as<uint64_t>(), n, dv.as<fr_t>(),
                    dp.as<fr_t>(), s);
    PK_HIP(hipMemcpyAsync(values, dv.p, 32 * n * leaf_elems, hipMemcpyDeviceToHost, s));
    if (depth) PK_HIP(hipMemcpyAsync(paths, dp.p, 32 * n * depth, hipMemcpyDeviceToHost, s));
  }
  // also on an error above: nothing may still read the buffers when the
  // reader mark goes
  (void)hipStreamSynchronize(s);
}

void Server::release(uint64_t id) {
  std::unique_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    PK_REQUIRE(it != jobs_.end(), PK_ERR_STATE, "server: unknown job");
    PK_REQUIRE(it->second->state == Job::Done && it->second->stream < 0, PK_ERR_STATE,
               "server: job still running");
    // openings in progress on other threads finish first (the job is looked
    // up again: a concurrent release may have taken it meanwhile)
    done_cv_.wait(lk, [&] {
      auto f = jobs_.find(id);
      return f == jobs_.end() || f->second->readers == 0;
    });
    it = jobs_.find(id);
    PK_REQUIRE(it != jobs_.end(), PK_ERR_STATE, "server: unknown job");
    j = std::move(it->second);
    jobs_.erase(it);
    if (!j->failed) in_use_ -= j->resident_bytes;
    events_++;  // the freed budget may admit a pending job
  }
  hipStream_t s = thread_stream();
  if (j->codeword) PK_HIP(hipFreeAsync(j->codeword, s));
  if (j->nodes) PK_HIP(hipFreeAsync(j->nodes, s));
  cv_.notify_all();
}

uint64_t Server::bytes_in_use() {
  std::lock_guard<std::mutex> lk(mu_);
  return in_use_;
}
uint64_t Server::peak_in_flight() {
  std::lock_guard<std::mutex> lk(mu_);
  return peak_in_flight_;
}

}  // namespace pk
