// NTT schedule autotuner and its cache.
//
// The reference's ntt crate runs one fixed recursive schedule on the CPU
// (ntt/src/ntt.rs:150-248).  On the GPU the same stages can be cut into
// passes in many ways; which one is fastest depends on the length (how many
// workgroups a pass has to fill 256 CUs, how many waves share a CU's LDS).
// The knobs per transform are
//   tile_log  tile of 2^tile_log elements in LDS (stages per pass, LDS split),
//   bpt       butterflies per thread per stage (waves per tile),
//   min_run   consecutive elements kept per tile row in early passes,
// and the table maps (log2 codeword, stages run) to the fastest measured
// shape.  It is saved as JSON and read back when the library starts.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "kernels.hpp"

namespace pk {

namespace {

struct Entry {
  NttShape shape;
  double us;
};

std::mutex g_mu;
std::map<std::pair<int, int>, Entry> g_table;  // (log codeword, stages)
std::once_flag g_loaded;

std::string default_path() {
  if (const char* e = std::getenv("PKGPU_NTT_TUNING")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&ntt_tuned_shape), &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/ntt_tuning.json";
  }
  return "ntt_tuning.json";
}

// Reads {"key": number, ...} objects from the "entries" array of the file;
// anything that is not such an object is skipped.
bool parse_entries(const std::string& text, std::map<std::pair<int, int>, Entry>& out) {
  size_t pos = text.find("\"entries\"");
  if (pos == std::string::npos) return false;
  while ((pos = text.find('{', pos)) != std::string::npos) {
    const size_t end = text.find('}', pos);
    if (end == std::string::npos) break;
    std::map<std::string, double> kv;
    size_t k = pos;
    while ((k = text.find('"', k)) != std::string::npos && k < end) {
      const size_t k2 = text.find('"', k + 1);
      if (k2 == std::string::npos || k2 > end) break;
      const std::string key = text.substr(k + 1, k2 - k - 1);
      size_t c = text.find(':', k2);
      if (c == std::string::npos || c > end) break;
      kv[key] = std::strtod(text.c_str() + c + 1, nullptr);
      k = c + 1;
    }
    if (kv.count("log") && kv.count("stages") && kv.count("tile_log") && kv.count("bpt") &&
        kv.count("min_run")) {
      NttShape s;
      s.tile_log = (int)kv["tile_log"];
      s.bpt = (int)kv["bpt"];
      s.min_run = (int)kv["min_run"];
      if (ntt_shape_supported(s.tile_log, s.bpt) && s.min_run >= 0 && s.min_run < s.tile_log)
        out[{(int)kv["log"], (int)kv["stages"]}] = {s, kv.count("us") ? kv["us"] : 0.0};
    }
    pos = end + 1;
  }
  return true;
}

void load_locked(const std::string& path, bool required) {
  std::ifstream f(path);
  if (!f) {
    PK_REQUIRE(!required, PK_ERR_ARG, "ntt tuning: cannot open " + path);
    return;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::map<std::pair<int, int>, Entry> t;
  const bool ok = parse_entries(ss.str(), t);
  PK_REQUIRE(ok || !required, PK_ERR_ARG, "ntt tuning: no \"entries\" array in " + path);
  if (ok) g_table.swap(t);
}

void ensure_loaded() {
  std::call_once(g_loaded, [] {
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked(default_path(), false);
  });
}

}  // namespace

NttShape ntt_tuned_shape(int log_codeword, int stages) {
  ensure_loaded();
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find({log_codeword, stages});
  if (it != g_table.end()) return it->second.shape;
  // nearest length with the same number of skipped stages, else any length
  const int skip = log_codeword - stages;
  const Entry* best = nullptr;
  int best_d = 1 << 30;
  for (int pass = 0; pass < 2 && !best; pass++) {
    for (const auto& kv : g_table) {
      if (pass == 0 && kv.first.first - kv.first.second != skip) continue;
      const int d = std::abs(kv.first.first - log_codeword);
      if (d < best_d) {
        best_d = d;
        best = &kv.second;
      }
    }
  }
  if (best) return best->shape;
  NttShape s;
  s.tile_log = 11;
  s.bpt = 4;
  s.min_run = 2;
  return s;
}

void ntt_tuning_put(int log_codeword, int stages, NttShape shape, double us) {
  ensure_loaded();
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[{log_codeword, stages}] = {shape, us};
}

void ntt_tuning_load(const std::string& path) {
  ensure_loaded();
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked(path, true);
}

void ntt_tuning_save(const std::string& path) {
  ensure_loaded();
  std::lock_guard<std::mutex> lk(g_mu);
  std::ofstream f(path);
  PK_REQUIRE(bool(f), PK_ERR_ARG, "ntt tuning: cannot write " + path);
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    arch = prop.gcnArchName;
  f << "{\n  \"device\": \"" << arch << "\",\n  \"entries\": [\n";
  size_t i = 0;
  for (const auto& kv : g_table) {
    char line[256];
    std::snprintf(line, sizeof line,
                  "    {\"log\": %d, \"stages\": %d, \"tile_log\": %d, \"bpt\": %d, \"min_run\": %d, "
                  "\"us\": %.2f}%s\n",
                  kv.first.first, kv.first.second, kv.second.shape.tile_log, kv.second.shape.bpt,
                  kv.second.shape.min_run, kv.second.us, ++i == g_table.size() ? "" : ",");
    f << line;
  }
  f << "  ]\n}\n";
  PK_REQUIRE(bool(f), PK_ERR_ARG, "ntt tuning: write failed for " + path);
}

double ntt_autotune(Device& d, int log_min, int log_max, int max_skip, hipStream_t s) {
  PK_REQUIRE(log_min >= 1 && log_max <= 27 && log_min <= log_max && max_skip >= 0, PK_ERR_ARG,
             "ntt autotune: log range must lie in [1, 27]");
  const auto t_begin = std::chrono::steady_clock::now();
  roots_table(d, log_max - 1);
  Scratch buf(sizeof(fr_t) << log_max, s);
  PK_HIP(hipMemsetAsync(buf.p, 0, sizeof(fr_t) << log_max, s));
  hipEvent_t e0, e1;
  PK_HIP(hipEventCreate(&e0));
  PK_HIP(hipEventCreate(&e1));
  struct Cand {
    NttShape shape;
    std::vector<NttPassDesc> plan;
  };
  try {
    for (int L = log_min; L <= log_max; L++) {
      const uint64_t len = 1ull << L;
      for (int skip = 0; skip <= max_skip && skip < L; skip++) {
        std::vector<Cand> cands;
        for (int tl = 6; tl <= 11; tl++) {
          for (int bpt : {1, 2, 4}) {
            if (!ntt_shape_supported(tl, bpt)) continue;
            for (int mr = 0; mr <= 3 && mr < tl; mr++) {
              NttShape sh;
              sh.tile_log = tl;
              sh.bpt = bpt;
              sh.min_run = mr;
              std::vector<NttPassDesc> plan =
                  ntt_plan(d, buf.as<fr_t>(), len, len, 1ull << skip, nullptr, 0, sh);
              bool dup = false;
              for (const Cand& c : cands) {
                if (c.plan.size() != plan.size()) continue;
                bool same = true;
                for (size_t i = 0; i < plan.size() && same; i++)
                  same = plan[i].r == c.plan[i].r && plan[i].jb == c.plan[i].jb &&
                         plan[i].tl == c.plan[i].tl && plan[i].bpt == c.plan[i].bpt;
                dup |= same;
              }
              if (!dup) cands.push_back({sh, plan});
            }
          }
        }
        // repetitions: about 2^25 element-stages per measurement, 2..64 runs
        const int reps = (int)std::min<uint64_t>(64, std::max<uint64_t>(2, (1ull << 25) / (len * L)));
        double best_us = 1e30;
        NttShape best;
        for (const Cand& c : cands) {
          ntt_run_plan(c.plan, len, s);  // warm-up
          PK_HIP(hipEventRecord(e0, s));
          for (int r = 0; r < reps; r++) ntt_run_plan(c.plan, len, s);
          PK_HIP(hipEventRecord(e1, s));
          PK_HIP(hipEventSynchronize(e1));
          float ms = 0;
          PK_HIP(hipEventElapsedTime(&ms, e0, e1));
          const double us = 1e3 * ms / reps;
          if (us < best_us) {
            best_us = us;
            best = c.shape;
          }
        }
        ntt_tuning_put(L, L - skip, best, best_us);
      }
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  PK_HIP(hipStreamSynchronize(s));
  PK_HIP(hipEventDestroy(e0));
  PK_HIP(hipEventDestroy(e1));
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
}

}  // namespace pk
