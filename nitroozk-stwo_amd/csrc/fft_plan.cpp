// Circle-FFT launch plans: the tile of the low-stage kernel for each transform shape.
//
// launch_fft runs the stages [0, S) of a transform of k stages in one kernel whose workgroup owns
// 2^S consecutive values of a column (11 <= S <= min(k, 15), 2^S words of LDS) and the stages
// [S, k) in strided passes of at most 8 stages. The largest tile reads and writes the data once,
// but one workgroup of 2^15 values runs for tens of microseconds, so a transform of few columns
// (the composition's 4 or 8, the interaction's 32, or 100 columns of 2^16) occupies fewer
// workgroups than the device has CUs, or ends in a partly filled wave of workgroups. A smaller
// tile spreads the same transform over more workgroups, at the cost of one more pass over HBM.
//
// The low stages of 2^15-value tiles can also run in a looping kernel (FFT_LOW_LOOP): one
// workgroup per CU, loading its next tile while it computes the current one, which hides the loads
// of a tile that no second workgroup on the CU could overlap. It writes 16 bytes at a time, so a plan
// that uses it is only taken for 16-byte aligned output columns (Ctx::fft_plan_for).
//
// The autotuner times every candidate (each S, and the looping kernel at S = 15) once per
// (device, direction, columns, sizes) on scratch columns and keeps the fastest; transforms of at
// most 15 stages whose largest-tile grid fills the device four times over take S = k without
// timing (looping when forward and S = 15), larger ones are timed on a subset of their columns.
// A transform beside critical-path work on another stream is planned, and timed, without the
// looping kernel: its workgroups would hold every CU until the transform ends. Choices persist across processes in the cache directory as
// fft_plans_<arch>_<CUs>.txt, one "direction n_cols log_in log_n plan" line per shape (direction: 1
// for inverse, plus 2 for a background transform; plan: S, plus 256 for the looping kernel).
// STWO_HIP_FFT_TUNE=0 turns the timing off (a fixed rule picks S instead).
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "runtime.h"

namespace sh {

std::string cache_directory() {
  const char* dir = std::getenv("STWO_HIP_KERNEL_CACHE");
  const char* home = std::getenv("HOME");
  std::string d = dir ? dir : (home ? std::string(home) + "/.cache/stwo_hip" : std::string());
  for (size_t k = 1; k <= d.size(); ++k)  // the directory and its parents
    if (k == d.size() || d[k] == '/') (void)mkdir(d.substr(0, k).c_str(), 0755);
  return d;
}

namespace {

// device, inverse (+2 for background transforms), n_cols, log_in, log_n
using PlanKey = std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t>;
std::mutex g_mu;
std::map<PlanKey, uint32_t> g_plans;
std::set<int> g_loaded;  // devices whose plan file has been read

std::string plan_file(int device, int cus) {
  const std::string d = cache_directory();
  if (d.empty()) return {};
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return {};
  std::string arch = prop.gcnArchName;
  if (auto c = arch.find(':'); c != std::string::npos) arch.resize(c);
  return d + "/fft_plans_" + arch + "_" + std::to_string(cus) + ".txt";
}

// With g_mu held.
void load_plans(int device, int cus) {
  if (!g_loaded.insert(device).second) return;
  const std::string path = plan_file(device, cus);
  if (path.empty()) return;
  std::FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) return;
  unsigned dir, n, li, ln, s;
  while (std::fscanf(fp, "%u %u %u %u %u", &dir, &n, &li, &ln, &s) == 5) {
    const uint32_t k = (dir & 1) ? ln : li, tile = s & ~FFT_LOW_LOOP, loop = s & FFT_LOW_LOOP;
    if (dir < 4 && k >= 11 && tile >= 11 && tile <= std::min<uint32_t>(k, 15) && li <= ln && ln <= 31 &&
        (s & ~(FFT_LOW_LOOP | 0xFFu)) == 0 && (!loop || (tile == 15 && dir < 2)))
      g_plans[PlanKey{device, dir, n, li, ln}] = s;
  }
  std::fclose(fp);
}

void persist_plan(int device, int cus, const PlanKey& key, uint32_t s) {
  const std::string path = plan_file(device, cus);
  if (path.empty()) return;
  char line[96];
  const int len = std::snprintf(line, sizeof(line), "%u %u %u %u %u\n", std::get<1>(key), std::get<2>(key),
                                std::get<3>(key), std::get<4>(key), s);
  // one short append per line: whole lines even when processes tune at the same time
  const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND, 0644);
  if (fd < 0) return;
  (void)!write(fd, line, (size_t)len);
  close(fd);
}

bool tuning_enabled() {
  const char* e = std::getenv("STWO_HIP_FFT_TUNE");
  return !(e && e[0] == '0');
}

}  // namespace

uint32_t Ctx::fft_plan_for(uint32_t plan, uint32_t* const* dst, size_t n_cols) {
  if (plan & FFT_LOW_LOOP)
    for (size_t c = 0; c < n_cols; ++c)
      if (reinterpret_cast<uintptr_t>(dst[c]) & 15) return plan & ~FFT_LOW_LOOP;
  return plan;
}

uint32_t Ctx::fft_low_log(bool inverse, uint32_t n_cols, uint32_t log_in, uint32_t log_n, const uint32_t* tree2,
                          uint32_t tree_len, bool background) {
  const uint32_t k = inverse ? log_n : log_in;
  if (k < 11 || n_cols == 0) return 0;
  const uint32_t smax = k < 15 ? k : 15;
  if (fft_low_override) {
    const uint32_t s = std::min(std::max(fft_low_override & ~FFT_LOW_LOOP, 11u), smax);
    const bool loop = (fft_low_override & FFT_LOW_LOOP) && s == 15 && !background;
    return s | (loop ? FFT_LOW_LOOP : 0u);
  }
  const uint64_t cus = (uint64_t)(cu_count > 0 ? cu_count : 1);
  if (smax == 11) return smax;
  const bool fills = ((uint64_t)n_cols << (log_n - smax)) >= 4 * cus;
  // transforms of at most 15 stages that fill the device four times over: one pass of the largest
  // tile, for forward transforms in the looping kernel, whose workgroups load their next tile while
  // they compute (the inverse gains nothing from it)
  if (fills && k <= 15) return smax | (!inverse && !background && smax == 15 ? FFT_LOW_LOOP : 0u);
  const PlanKey key{device, (inverse ? 1u : 0u) + (background ? 2u : 0u), n_cols, log_in, log_n};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    load_plans(device, cu_count);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) return it->second;
  }
  if (!tuning_enabled()) {
    // the largest tile whose grid still covers every CU twice
    uint32_t s = smax;
    while (s > 11 && ((uint64_t)n_cols << (log_n - s)) < 2 * cus) --s;
    return s;
  }
  // Time each candidate on zeroed scratch columns of the same shape (the values do not change the
  // work), after what is already queued on the stream; two runs each, the faster one counts. A
  // transform that fills the device (more than 15 stages: strided passes follow the low stages
  // whatever the tile) is timed on enough of its columns to fill it four times over.
  if (fills) n_cols = std::min<uint64_t>(n_cols, (4 * cus + (1u << (log_n - 15)) - 1) >> (log_n - 15));
  const size_t in_words = (size_t)1 << (inverse ? log_n : log_in), out_words = (size_t)1 << log_n;
  uint32_t* src = nullptr;
  uint32_t* dst = nullptr;
  HIP_CHECK(hipMallocFromPoolAsync((void**)&src, (size_t)n_cols * in_words * 4, pool, stream));
  HIP_CHECK(hipMallocFromPoolAsync((void**)&dst, (size_t)n_cols * out_words * 4, pool, stream));
  launch_fill_zero(src, (size_t)n_cols * in_words, stream);
  std::vector<const uint32_t*> s_tab(n_cols);
  std::vector<uint32_t*> d_tab(n_cols);  // 256-byte aligned blocks: every candidate applies
  for (uint32_t c = 0; c < n_cols; ++c) {
    s_tab[c] = src + (size_t)c * in_words;
    d_tab[c] = dst + (size_t)c * out_words;
  }
  const auto* ds = (const uint32_t* const*)upload_vec(s_tab);
  auto* dd = (uint32_t* const*)upload_vec(d_tab);
  // candidates: every tile, and the looping kernel with 2^15 values
  std::vector<uint32_t> cand;
  for (uint32_t s = 11; s <= smax; ++s) cand.push_back(s);
  if (smax == 15 && !background) cand.push_back(15 | FFT_LOW_LOOP);
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  for (uint32_t c : cand) {
    launch_fft(inverse, ds, dd, n_cols, log_in, log_n, tree2, tree_len, stream, false, c);  // warm-up
    for (int rep = 0; rep < 2; ++rep) {
      hipEvent_t a = get_event(), b = get_event();
      HIP_CHECK(hipEventRecord(a, stream));
      launch_fft(inverse, ds, dd, n_cols, log_in, log_n, tree2, tree_len, stream, false, c);
      HIP_CHECK(hipEventRecord(b, stream));
      ev.push_back({a, b});
    }
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  uint32_t best = smax;
  float best_ms = 0;
  for (size_t j = 0, i = 0; j < cand.size(); ++j) {
    for (int rep = 0; rep < 2; ++rep, ++i) {
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[i].first, ev[i].second));
      if (best_ms == 0 || ms < best_ms) {
        best_ms = ms;
        best = cand[j];
      }
    }
  }
  for (auto& [a, b] : ev) {
    event_pool.push_back(a);
    event_pool.push_back(b);
  }
  HIP_CHECK(hipFreeAsync(src, stream));
  HIP_CHECK(hipFreeAsync(dst, stream));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_plans.emplace(key, best).second) persist_plan(device, cu_count, key, best);
  }
  return best;
}

}  // namespace sh
