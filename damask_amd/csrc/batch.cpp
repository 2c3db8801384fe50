// Many independent RVEs on one GPU.  Each RVE is a solver handle with its own
// HIP stream; a pool of host threads drives the handles' increment loops
// (dmk_grid_run_loadcase) and adjoint sweeps concurrently, so that the
// kernels, copies and host synchronisations of different RVEs overlap.  Work
// is handed out largest RVE first from a shared counter.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "dmk_common.h"
#include "../../include/damask_hip.h"

namespace {

template <class F>
int run_pool(int n, const std::vector<int>& order, int n_threads, F&& job, int32_t* rc_out, double* wall_s) {
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      int k = next.fetch_add(1);
      if (k >= n) break;
      int i = order[k];
      auto t0 = std::chrono::steady_clock::now();
      int rc = job(i);
      auto t1 = std::chrono::steady_clock::now();
      if (rc_out) rc_out[i] = rc;
      if (wall_s) wall_s[i] = std::chrono::duration<double>(t1 - t0).count();
    }
  };
  int nt = std::max(1, std::min(n_threads, n));
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; t++) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  return 0;
}

std::vector<int> largest_first(dmk_handle** h, int n, const int64_t* cells) {
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  if (cells) std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a] > cells[b]; });
  return order;
}

}  // namespace

extern "C" int dmk_batch_run(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells,
                             int32_t n_threads, int32_t* rc_out, double* wall_s) {
  if (!h || !lc || n < 1) { dmk_set_error("dmk_batch_run: invalid argument"); return DMK_ERR_ARG; }
  std::vector<int> order = largest_first(h, n, cells);
  std::vector<int32_t> rc(n, 0);
  run_pool(n, order, n_threads,
           [&](int i) { return dmk_grid_run_loadcase(h[i], &lc[i], nullptr, nullptr, nullptr, nullptr, nullptr); },
           rc.data(), wall_s);
  int bad = 0;
  for (int i = 0; i < n; i++) {
    if (rc_out) rc_out[i] = rc[i];
    if (rc[i]) bad++;
  }
  if (bad) { dmk_set_error("dmk_batch_run: %d of %d RVEs failed", bad, n); return DMK_ERR_NOCONV; }
  return 0;
}

extern "C" int dmk_batch_adjoint(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                                 double* const* theta_bar, const int64_t* cells, int32_t n_threads, int32_t* rc_out,
                                 double* wall_s) {
  if (!h || !P_av_bar || !theta_bar || n < 1) {
    dmk_set_error("dmk_batch_adjoint: invalid argument");
    return DMK_ERR_ARG;
  }
  std::vector<int> order = largest_first(h, n, cells);
  std::vector<int32_t> rc(n, 0);
  run_pool(n, order, n_threads,
           [&](int i) { return dmk_adjoint_backward(h[i], P_av_bar[i], rtol, theta_bar[i], nullptr); }, rc.data(),
           wall_s);
  int bad = 0;
  for (int i = 0; i < n; i++) {
    if (rc_out) rc_out[i] = rc[i];
    if (rc[i]) bad++;
  }
  if (bad) { dmk_set_error("dmk_batch_adjoint: %d of %d RVEs failed", bad, n); return DMK_ERR_NOCONV; }
  return 0;
}
