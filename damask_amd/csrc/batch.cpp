// Many independent RVEs on one GPU.  Each RVE is a solver handle with its own
// HIP stream; a pool of host threads drives the handles' increment loops
// (dmk_grid_run_loadcase) and adjoint sweeps concurrently, so that the
// kernels, copies and host synchronisations of different RVEs overlap.  Work
// is handed out largest RVE first from a shared counter.
//
// Lockstep mode (dmk_batch_run_lockstep): the RVEs in flight also share their
// stress integrations.  Each residual evaluation submits its (RVE, phase)
// items to a DmkStressBatch; when every RVE still running has submitted, the
// last one enqueues one stress_kernel_batch launch per slip-system variant
// over all of them on the batch's stream, and every RVE's stream waits for it.
// The many small per-RVE launches, which the few hardware queues serialise,
// become one launch that fills the GPU.
//
// For the basic scheme the whole residual evaluation joins the round: each
// RVE submits, with its stress items, the arguments of its response
// reductions, transform with Gamma, dot products and report (DmkResidDev).
// The round then runs every kernel once over all its RVEs (the transform
// kernels once per FFT configuration), gathers the reports into one buffer
// and copies it to the host: one host synchronisation per round instead of
// one per RVE.  The kernels run the per-block code of the single-RVE
// kernels, so each RVE's results are those of its own run, bit for bit.
// DMK_LOCKSTEP=stress in the environment keeps the residual evaluations per
// RVE (the stress integrations alone are batched).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "dmk_common.h"
#include "dmk_batch.h"
#include "dmk_resid.h"
#include "dmk_linalg.h"
#include "../../include/damask_hip.h"

struct DmkStressBatch {
  struct Req {
    int slot, n;
    DmkStressItem items[DMK_MAX_PHASES];
    int key[DMK_MAX_PHASES];
    bool resid;                      // the rest of a residual evaluation follows the stress
    const DmkFFT* fft;
    DmkResidDev rq;
  };
  static constexpr int NSTAGE = 4;   // staging buffers of the item tables in flight
  std::mutex mu;
  std::condition_variable cv;
  int active = 0, waiting = 0;
  uint64_t gen = 0;                  // batches launched
  long wait_us = -1;                 // >= 0: a round starts after this wait without the RVEs not yet submitted
  hipStream_t st = nullptr;
  // per stage: the launch status and completion event of the stage's last
  // round, and how many of its submitters still use them (or its reports);
  // a new round on the stage waits until none does
  int rc[NSTAGE] = {};
  hipEvent_t done[NSTAGE] = {};
  int users[NSTAGE] = {};
  std::condition_variable cv_free;   // a stage's users reached zero
  bool launching = false;            // a launch is waiting for its stage: no other starts
  std::vector<hipEvent_t> ev_in;     // per slot: the submitter's input is ready
  std::vector<Req> reqs;
  DmkStressItem* host[NSTAGE] = {};
  DmkStressItem* dev[NSTAGE] = {};
  hipEvent_t staged[NSTAGE] = {};
  int cap = 0, stage = 0;
  int64_t launches = 0;
  // residual evaluations of the round: argument tables, reports (device and
  // pinned host), and per slot where its last report went
  bool resid = false;
  int slots = 0;
  DmkResidDev* rhost[NSTAGE] = {};
  DmkResidDev* rdev[NSTAGE] = {};
  double* rep_dev[NSTAGE] = {};
  double* rep_host[NSTAGE] = {};
  std::vector<int> slot_stage, rep_idx, order;
  // DMK_BATCH_TIMING: host time of the submitting threads spent waiting for
  // the round to start (the other RVEs' submissions and the launch), waiting
  // for the round on the device, and the launches themselves
  bool timing = false;
  std::atomic<int64_t> t_wait_ns{0}, t_gpu_ns{0}, t_launch_ns{0}, n_sub{0};
  // and on the device, per round: the batch stream waiting for the RVEs'
  // streams (and the table uploads), the stress launches, the rest
  hipEvent_t tev[NSTAGE][4] = {};
  double d_ms[3] = {0.0, 0.0, 0.0};
  int64_t d_rounds = 0;
  // and per round (by number, the first MAX_STAMPED rounds): steady-clock
  // times of the launch's end, of the round's end on the device (a host
  // function enqueued behind it) and of the first submitter's return with its
  // report
  static constexpr uint64_t MAX_STAMPED = 1 << 16;
  std::unique_ptr<std::atomic<int64_t>[]> stamp;
};

namespace {

int batch_init(DmkStressBatch* b, int slots, bool resid, long wait_us) {
  b->cap = slots * DMK_MAX_PHASES;
  b->slots = slots;
  b->resid = resid;
  b->reqs.reserve(slots);
  b->slot_stage.assign(slots, 0);
  const char* w = getenv("DMK_LOCKSTEP_WAIT_US");
  b->wait_us = w && w[0] ? std::atol(w) : wait_us;
  b->rep_idx.assign(slots, 0);
  b->order.reserve(slots);
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) return DMK_ERR_HIP;
  for (auto& e : b->done)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return DMK_ERR_HIP;
  for (auto& u : b->users) u = 0;
  b->ev_in.assign(slots, nullptr);
  for (auto& e : b->ev_in)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return DMK_ERR_HIP;
  for (int s = 0; s < DmkStressBatch::NSTAGE; s++) {
    if (b->timing)
      for (auto& e : b->tev[s])
        if (hipEventCreate(&e) != hipSuccess) return DMK_ERR_HIP;
    if (hipHostMalloc((void**)&b->host[s], sizeof(DmkStressItem) * b->cap, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&b->dev[s], sizeof(DmkStressItem) * b->cap) != hipSuccess ||
        hipEventCreateWithFlags(&b->staged[s], hipEventDisableTiming) != hipSuccess)
      return DMK_ERR_HIP;
    if (resid && (hipHostMalloc((void**)&b->rhost[s], sizeof(DmkResidDev) * slots, hipHostMallocDefault) != hipSuccess ||
                  hipMalloc((void**)&b->rdev[s], sizeof(DmkResidDev) * slots) != hipSuccess ||
                  hipMalloc((void**)&b->rep_dev[s], sizeof(double) * REP_SIZE * slots) != hipSuccess ||
                  hipHostMalloc((void**)&b->rep_host[s], sizeof(double) * REP_SIZE * slots, hipHostMallocDefault) !=
                      hipSuccess))
      return DMK_ERR_HIP;
  }
  return 0;
}

void batch_free(DmkStressBatch* b) {
  if (b->st) hipStreamSynchronize(b->st);
  for (int s = 0; s < DmkStressBatch::NSTAGE; s++) {
    if (b->host[s]) hipHostFree(b->host[s]);
    if (b->dev[s]) hipFree(b->dev[s]);
    if (b->staged[s]) hipEventDestroy(b->staged[s]);
    for (auto& e : b->tev[s])
      if (e) hipEventDestroy(e);
    if (b->rhost[s]) hipHostFree(b->rhost[s]);
    if (b->rdev[s]) hipFree(b->rdev[s]);
    if (b->rep_dev[s]) hipFree(b->rep_dev[s]);
    if (b->rep_host[s]) hipHostFree(b->rep_host[s]);
  }
  for (auto& e : b->ev_in) if (e) hipEventDestroy(e);
  for (auto& e : b->done) if (e) hipEventDestroy(e);
  if (b->st) hipStreamDestroy(b->st);
}

// FFT configurations that may share one launch of the transform kernels
bool same_fft(const DmkFFT* a, const DmkFFT* c) {
  for (int d = 0; d < 3; d++)
    if (a->n[d] != c->n[d]) return false;
  return a->xy_threads == c->xy_threads && a->xy_ip == c->xy_ip && a->zgz_g == c->zgz_g &&
         a->zgz_threads == c->zgz_threads && a->zgz_ip == c->zgz_ip;
}
bool fft_less(const DmkFFT* a, const DmkFFT* c) {
  const int ka[8] = {a->n[0], a->n[1], a->n[2], a->xy_threads, a->xy_ip, a->zgz_g, a->zgz_threads, a->zgz_ip};
  const int kc[8] = {c->n[0], c->n[1], c->n[2], c->xy_threads, c->xy_ip, c->zgz_g, c->zgz_threads, c->zgz_ip};
  return std::lexicographical_compare(ka, ka + 8, kc, kc + 8);
}

// with b->mu held, after the stress launches of stage s: the rest of the
// round's residual evaluations (their table uploaded to rdev[s] already)
int resid_launch_locked(DmkStressBatch* b, int s, int nr) {
  const DmkResidDev* R = b->rdev[s];
  int max_nc_resp = 0, max_nc_dots = 0, max_jobs = 0, max_sums = 0, max_nc_adj = 0, ntr = 0;
  long max_N_adj = 0, max_pre = 0, max_post = 0;
  bool any_res = false, any_adj = false;
  for (int k = 0; k < nr; k++) {
    const DmkResidDev& q = b->rhost[s][k];
    for (int o = 0; o < q.npre; o++) max_pre = std::max(max_pre, q.pre[o].n);
    if (q.npost) {
      if (q.npost != 1 || q.post.type != DMK_PRE_AXPBY) return DMK_ERR_ARG;   // the step after a transform is z = a x + b y
      max_post = std::max(max_post, q.post.n);
    }
    max_nc_dots = std::max(max_nc_dots, q.nc_dots);
    max_jobs = std::max(max_jobs, q.J.n);
    max_sums = std::max(max_sums, q.S.k);
    if (q.kind == DMK_RQ_DOTS) continue;     // sorted last: no transform
    ntr = k + 1;
    if (q.kind == DMK_RQ_RESIDUAL) {
      any_res = true;
      max_nc_resp = std::max(max_nc_resp, q.nc_resp);
    } else {
      any_adj = true;
      max_nc_adj = std::max(max_nc_adj, q.nc_resp);
      max_N_adj = std::max(max_N_adj, q.N);
    }
  }
  // the requests' vector steps first; then residual evaluations: response
  // reductions; adjoint operators: A^T v and the mean rule (each kernel
  // skips the other kind)
  if (max_pre > 0 && dmk_pre_batch(R, nr, max_pre, b->st)) return DMK_ERR_HIP;
  if (any_res && dmk_resid_reduce_batch(R, ntr, max_nc_resp, b->st)) return DMK_ERR_HIP;
  if (any_adj && (dmk_adj_atmul_batch(R, ntr, max_N_adj, b->st) || dmk_adj_mean_batch(R, ntr, max_nc_adj, b->st)))
    return DMK_ERR_HIP;
  for (int k0 = 0; k0 < ntr;) {
    const DmkFFT* f = b->reqs[b->order[k0]].fft;
    int k1 = k0 + 1;
    while (k1 < ntr && same_fft(f, b->reqs[b->order[k1]].fft)) k1++;
    const int rc = dmk_transform_batch(f, R + k0, k1 - k0, b->st);
    if (rc == 1) {
      // no fused kernels for this configuration: each RVE's transform on its own
      for (int k = k0; k < k1; k++) {
        const DmkResidDev& q = b->rhost[s][k];
        const DmkFFT* fk = b->reqs[b->order[k]].fft;
        int nb = 0;
        if (dmk_fft_r2c_xy(fk, 9, q.Pnew, q.Z, b->st) || dmk_gamma_conv(fk, q.Z, &q.g, q.divp, &nb, b->st) ||
            dmk_fft_c2r_xy(fk, 9, q.Z, q.r, b->st))
          return DMK_ERR_HIP;
      }
    } else if (rc) {
      return DMK_ERR_HIP;
    }
    k0 = k1;
  }
  // the vector steps that follow a request's transform, before its dot products
  if (max_post > 0 && dmk_post_batch(R, nr, max_post, b->st)) return DMK_ERR_HIP;
  if (dmk_resid_dots_batch(R, nr, max_nc_dots, max_jobs, max_sums, b->rep_dev[s], b->st)) return DMK_ERR_HIP;
  if (hipMemcpyAsync(b->rep_host[s], b->rep_dev[s], sizeof(double) * REP_SIZE * nr, hipMemcpyDeviceToHost, b->st) !=
      hipSuccess)
    return DMK_ERR_HIP;
  return 0;
}

// with b->mu held (through lk) and no launch in progress: one launch per
// slip-system variant over every waiting request (and the rest of the
// residual evaluations among them), then release the waiters
void batch_launch_locked(DmkStressBatch* b, std::unique_lock<std::mutex>& lk) {
  const auto t_start = std::chrono::steady_clock::now();
  const int s = b->stage;
  b->stage = (s + 1) % DmkStressBatch::NSTAGE;
  // the stage's previous round: every submitter done with its event and
  // report.  The wait releases the lock (those submitters may still have to
  // take it to leave their own wait); requests submitted meanwhile join this
  // round, and no other launch starts.
  b->launching = true;
  b->cv_free.wait(lk, [&] { return b->users[s] == 0; });
  b->launching = false;
  int rc = 0;
  if (hipEventSynchronize(b->staged[s]) != hipSuccess) rc = DMK_ERR_HIP;   // slot's previous upload done
  const bool tm = b->timing && !rc;
  if (tm) hipEventRecord(b->tev[s][0], b->st);
  for (auto& r : b->reqs)
    if (!rc && !(r.resid && r.rq.idle) && hipStreamWaitEvent(b->st, b->ev_in[r.slot], 0) != hipSuccess)
      rc = DMK_ERR_HIP;
  struct Group { int key, off, n, blocks; };
  Group g[4];
  int ng = 0, off = 0;
  for (auto& r : b->reqs)
    for (int i = 0; i < r.n; i++) {
      int k = 0;
      while (k < ng && g[k].key != r.key[i]) k++;
      if (k == ng) {
        if (ng == 4) { rc = DMK_ERR_CONFIG; break; }
        g[ng++] = Group{r.key[i], 0, 0, 0};
      }
    }
  for (int k = 0; k < ng && !rc; k++) {
    g[k].off = off;
    for (auto& r : b->reqs)
      for (int i = 0; i < r.n; i++) {
        if (r.key[i] != g[k].key) continue;
        if (off + g[k].n >= b->cap) { rc = DMK_ERR_ARG; break; }   // at most one request per slot
        DmkStressItem it = r.items[i];
        it.block0 = g[k].blocks;
        g[k].blocks += (it.mp.N + DMK_STRESS_WG - 1) / DMK_STRESS_WG;
        b->host[s][off + g[k].n++] = it;
      }
    off += g[k].n;
  }
  // residual evaluations, ordered by FFT configuration
  b->order.clear();
  for (int i = 0; i < (int)b->reqs.size(); i++)
    if (b->reqs[i].resid) b->order.push_back(i);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
    const bool dx = b->reqs[x].rq.kind == DMK_RQ_DOTS, dy = b->reqs[y].rq.kind == DMK_RQ_DOTS;
    if (dx || dy) return !dx && dy;                  // dot products alone last
    return fft_less(b->reqs[x].fft, b->reqs[y].fft);
  });
  const int nr = (int)b->order.size();
  if (nr > b->slots) rc = DMK_ERR_ARG;
  for (auto& r : b->reqs) b->slot_stage[r.slot] = s;
  for (int k = 0; k < nr && !rc; k++) {
    const DmkStressBatch::Req& r = b->reqs[b->order[k]];
    b->rhost[s][k] = r.rq;
    b->rep_idx[r.slot] = k;
  }
  if (!rc && (off > 0 || nr > 0)) {
    if ((off > 0 && hipMemcpyAsync(b->dev[s], b->host[s], sizeof(DmkStressItem) * off, hipMemcpyHostToDevice,
                                   b->st) != hipSuccess) ||
        (nr > 0 && hipMemcpyAsync(b->rdev[s], b->rhost[s], sizeof(DmkResidDev) * nr, hipMemcpyHostToDevice, b->st) !=
                       hipSuccess) ||
        hipEventRecord(b->staged[s], b->st) != hipSuccess)
      rc = DMK_ERR_HIP;
    if (tm) hipEventRecord(b->tev[s][1], b->st);
    for (int k = 0; k < ng && !rc; k++)
      if (dmk_launch_stress_batch(b->dev[s] + g[k].off, g[k].n, g[k].blocks, g[k].key / 2, g[k].key & 1, b->st))
        rc = DMK_ERR_HIP;
    if (tm) hipEventRecord(b->tev[s][2], b->st);
    if (!rc && nr > 0) rc = resid_launch_locked(b, s, nr);
    if (tm) hipEventRecord(b->tev[s][3], b->st);
  }
  if (!rc && hipEventRecord(b->done[s], b->st) != hipSuccess) rc = DMK_ERR_HIP;
  if (rc) dmk_set_error("lockstep batch: launch of the round failed");
  b->rc[s] = rc;
  b->users[s] = (int)b->reqs.size();
  if (b->timing) {
    const auto t_end = std::chrono::steady_clock::now();
    b->t_launch_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t_end - t_start).count();
    if (!rc && b->stamp && b->gen < DmkStressBatch::MAX_STAMPED) {
      std::atomic<int64_t>* st = &b->stamp[3 * b->gen];
      st[0].store(t_end.time_since_epoch().count());
      hipLaunchHostFunc(
          b->st, [](void* p) {
            static_cast<std::atomic<int64_t>*>(p)->store(std::chrono::steady_clock::now().time_since_epoch().count());
          },
          st + 1);
    }
  }
  b->gen++;
  b->launches++;
  b->reqs.clear();
  b->waiting = 0;
  b->cv.notify_all();
}

void batch_enter(DmkStressBatch* b) {
  std::lock_guard<std::mutex> lk(b->mu);
  b->active++;
}

void batch_leave(DmkStressBatch* b) {
  std::unique_lock<std::mutex> lk(b->mu);
  b->active--;
  if (!b->launching && b->waiting > 0 && b->waiting == b->active) batch_launch_locked(b, lk);
}

// DMK_BATCH_PROGRESS set (and not "0") in the environment: one line on stderr
// per finished RVE
bool progress_lines() {
  const char* v = getenv("DMK_BATCH_PROGRESS");
  return v && v[0] && !(v[0] == '0' && v[1] == 0);
}

template <class F>
int run_pool(int n, const std::vector<int>& order, int n_threads, F&& job, int32_t* rc_out, double* wall_s) {
  std::atomic<int> next{0}, done{0};
  const bool report = progress_lines();
  const auto start = std::chrono::steady_clock::now();
  auto worker = [&](int slot) {
    for (;;) {
      int k = next.fetch_add(1);
      if (k >= n) break;
      int i = order[k];
      auto t0 = std::chrono::steady_clock::now();
      int rc = job(i, slot);
      auto t1 = std::chrono::steady_clock::now();
      if (rc_out) rc_out[i] = rc;
      if (wall_s) wall_s[i] = std::chrono::duration<double>(t1 - t0).count();
      if (report) {
        int d = done.fetch_add(1) + 1;
        fprintf(stderr, "batch: RVE %d finished (rc %d), %d of %d after %.1f s\n", i, rc, d, n,
                std::chrono::duration<double>(t1 - start).count());
        fflush(stderr);
      }
    }
  };
  int nt = std::max(1, std::min(n_threads, n));
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; t++) pool.emplace_back(worker, t);
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

static int submit(DmkStressBatch* b, int slot, const DmkStressItem* items, const int* key, int n,
                  const DmkResidDev* rq, const DmkFFT* fft, hipStream_t st, double* report, hipEvent_t* done_out) {
  const bool needs_fft = rq && rq->kind != DMK_RQ_DOTS;
  if (n < (rq ? 0 : 1) || n > DMK_MAX_PHASES || slot < 0 || slot >= (int)b->ev_in.size() || (rq && !b->resid) ||
      (needs_fft && !fft) || (rq && rq->kind == DMK_RQ_RESIDUAL && n < 1)) {
    dmk_set_error("lockstep batch: bad submission");
    return DMK_ERR_ARG;
  }
  if (!(rq && rq->idle) && hipEventRecord(b->ev_in[slot], st) != hipSuccess) {
    dmk_set_error("lockstep batch: event record");
    return DMK_ERR_HIP;
  }
  const auto t0 = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(b->mu);
  DmkStressBatch::Req r;
  r.slot = slot;
  r.n = n;
  if (n > 0) std::memcpy(r.items, items, sizeof(DmkStressItem) * n);
  std::memcpy(r.key, key, sizeof(int) * n);
  r.resid = rq != nullptr;
  r.fft = fft;
  if (rq) r.rq = *rq;
  b->reqs.push_back(r);
  b->waiting++;
  // the first round launched after this submission holds it: launched here
  // when every RVE still running has submitted, or by a waiter whose wait
  // ran out (wait_us >= 0; the RVEs not yet submitted join the next round)
  const uint64_t my = b->gen;
  if (!b->launching && b->waiting == b->active) {
    batch_launch_locked(b, lk);
  } else if (b->wait_us < 0) {
    b->cv.wait(lk, [&] { return b->gen != my; });
  } else {
    // system_clock: its timed wait is pthread_cond_timedwait, which race
    // checkers see release the mutex (the steady clock's wait is not seen so)
    while (b->gen == my)
      if (!b->cv.wait_until(lk, std::chrono::system_clock::now() + std::chrono::microseconds(b->wait_us),
                            [&] { return b->gen != my; }) &&
          !b->launching)
        batch_launch_locked(b, lk);
  }
  const int s = b->slot_stage[slot], ri = b->rep_idx[slot];
  const hipEvent_t ev = b->done[s];
  const int rc = b->rc[s];
  lk.unlock();
  const auto t1 = std::chrono::steady_clock::now();
  int err = 0;
  float ms[3] = {-1.0f, 0.0f, 0.0f};
  if (rc) {
    dmk_set_error("lockstep batch: launch of the round failed");
    err = rc;
  } else if (!done_out && hipStreamWaitEvent(st, ev, 0) != hipSuccess) {
    dmk_set_error("lockstep batch: stream wait");
    err = DMK_ERR_HIP;
  } else if (rq && rq->rep_n > 0) {
    // the report reaches the host at the end of the round; a request without
    // one (an operator application whose result only later requests read,
    // vector steps) returns at the launch, and its stage may then be taken by
    // a later round while this one runs: the stage's uploads are ordered
    // behind it on the batch stream, and the reports of the stage are read
    // only by submitters that hold it
    if (hipEventSynchronize(ev) != hipSuccess) {
      dmk_set_error("lockstep batch: round failed");
      err = DMK_ERR_HIP;
    } else {
      std::memcpy(report, b->rep_host[s] + (size_t)ri * REP_SIZE, sizeof(double) * rq->rep_n);
      if (b->timing && ri == 0) {
        // one thread per round: the round's device intervals
        for (int k = 0; k < 3; k++) hipEventElapsedTime(&ms[k], b->tev[s][k], b->tev[s][k + 1]);
      }
      if (b->timing && b->stamp && my < DmkStressBatch::MAX_STAMPED) {
        const int64_t now = std::chrono::steady_clock::now().time_since_epoch().count();
        std::atomic<int64_t>& w = b->stamp[3 * my + 2];
        int64_t cur = w.load();
        while ((cur == 0 || now < cur) && !w.compare_exchange_weak(cur, now)) {
        }
      }
    }
  }
  if (done_out) *done_out = ev;
  {
    std::lock_guard<std::mutex> g(b->mu);
    if (--b->users[s] == 0) b->cv_free.notify_all();
  }
  if (b->timing) {
    const auto t2 = std::chrono::steady_clock::now();
    b->t_wait_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    b->t_gpu_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t2 - t1).count();
    b->n_sub++;
    if (ms[0] >= 0.0f) {
      std::lock_guard<std::mutex> g(b->mu);
      for (int k = 0; k < 3; k++) b->d_ms[k] += ms[k];
      b->d_rounds++;
    }
  }
  return err;
}

int dmk_stress_batch_submit(DmkStressBatch* b, int slot, const DmkStressItem* items, const int* key, int n,
                            hipStream_t st) {
  return submit(b, slot, items, key, n, nullptr, nullptr, st, nullptr, nullptr);
}

int dmk_resid_batch_submit(DmkStressBatch* b, int slot, const DmkStressItem* items, const int* key, int n,
                           const DmkResidDev* rq, const DmkFFT* fft, hipStream_t st, double* report,
                           hipEvent_t* done_out) {
  if (!rq || (rq->rep_n > 0 && !report)) { dmk_set_error("lockstep batch: bad submission"); return DMK_ERR_ARG; }
  return submit(b, slot, items, key, n, rq, fft, st, report, done_out);
}

bool dmk_stress_batch_resid(const DmkStressBatch* b) { return b && b->resid; }

// lockstep groups (DMK_LOCKSTEP_GROUPS, default 1): thread t belongs to
// group t % ng.  The groups' rounds are independent, so one group's host
// work between rounds overlaps the other groups' rounds on the device.  On
// the 256-RVE batch the rounds are bound by the device (the stress
// integrations), and 2 or 4 groups ran no faster than one.
// wait_us: how long a round waits for RVEs not yet submitted (< 0: for
// all), unless DMK_LOCKSTEP_WAIT_US sets it
static int make_groups(std::vector<std::unique_ptr<DmkStressBatch>>& grp, int nt, int& ng, long wait_us) {
  const char* gs = getenv("DMK_LOCKSTEP_GROUPS");
  ng = gs && gs[0] ? std::atoi(gs) : 1;
  ng = std::max(1, std::min(ng, nt));
  const char* mode = getenv("DMK_LOCKSTEP");
  const char* tm = getenv("DMK_BATCH_TIMING");
  const bool timing = tm && tm[0] && !(tm[0] == '0' && tm[1] == 0);
  int e = 0;
  for (int k = 0; k < ng && !e; k++) {
    grp.emplace_back(new DmkStressBatch);
    grp.back()->timing = timing;
    if (timing) {
      grp.back()->stamp.reset(new std::atomic<int64_t>[3 * DmkStressBatch::MAX_STAMPED]);
      for (uint64_t i = 0; i < 3 * DmkStressBatch::MAX_STAMPED; i++) grp.back()->stamp[i].store(0);
    }
    e = batch_init(grp.back().get(), (nt - k + ng - 1) / ng, !(mode && std::strcmp(mode, "stress") == 0), wait_us);
  }
  if (e) {
    for (auto& b : grp) batch_free(b.get());
    grp.clear();
  }
  return e;
}

static void print_timing(const std::vector<std::unique_ptr<DmkStressBatch>>& grp) {
  for (size_t k = 0; k < grp.size(); k++) {
    const DmkStressBatch& b = *grp[k];
    if (!b.timing || b.n_sub == 0) continue;
    const double ns = (double)b.n_sub, nr = (double)std::max<uint64_t>(b.gen, 1);
    fprintf(stderr,
            "lockstep group %zu timing: %llu rounds, %lld submissions; per submission %.3f ms waiting for the round "
            "to start, %.3f ms for the device; launches %.3f ms per round\n",
            k, (unsigned long long)b.gen, (long long)b.n_sub.load(), b.t_wait_ns * 1e-6 / ns, b.t_gpu_ns * 1e-6 / ns,
            b.t_launch_ns * 1e-6 / nr);
    if (b.stamp) {
      hipStreamSynchronize(b.st);     // the host functions of the last rounds
      double dev = 0.0, wake = 0.0;
      int64_t m = 0;
      for (uint64_t g = 0; g < std::min<uint64_t>(b.gen, DmkStressBatch::MAX_STAMPED); g++) {
        const int64_t t0 = b.stamp[3 * g].load(), t1 = b.stamp[3 * g + 1].load(), t2 = b.stamp[3 * g + 2].load();
        if (!t0 || !t1 || !t2) continue;
        dev += (double)(t1 - t0);
        wake += (double)(t2 - t1);
        m++;
      }
      if (m)
        fprintf(stderr,
                "lockstep group %zu timing per round (%lld rounds): %.3f ms from the launch's end to the round's end on "
                "the device, %.3f ms from there to the first submitter's return with its report\n",
                k, (long long)m, dev * 1e-6 / m, wake * 1e-6 / m);
    }
    if (b.d_rounds)
      fprintf(stderr,
              "lockstep group %zu timing on the device, per round: %.3f ms waiting for the RVE streams and uploads, "
              "%.3f ms stress, %.3f ms the rest of the requests\n",
              k, b.d_ms[0] / b.d_rounds, b.d_ms[1] / b.d_rounds, b.d_ms[2] / b.d_rounds);
    fflush(stderr);
  }
}

static int batch_run(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells, int32_t n_threads,
                     int32_t* rc_out, double* wall_s, bool lockstep, const char* name) {
  if (!h || !lc || n < 1) { dmk_set_error("%s: invalid argument", name); return DMK_ERR_ARG; }
  std::vector<int> order = largest_first(h, n, cells);
  std::vector<int32_t> rc(n, 0);
  const int nt = std::max(1, std::min<int>(n_threads, n));
  std::vector<std::unique_ptr<DmkStressBatch>> grp;
  int ng = 0;
  if (lockstep) {
    int e = make_groups(grp, nt, ng, -1);
    if (e) {
      dmk_set_error("%s: HIP resources of the lockstep batch", name);
      return e;
    }
  }
  run_pool(n, order, nt,
           [&](int i, int slot) {
             if (!lockstep)
               return dmk_grid_run_loadcase(h[i], &lc[i], nullptr, nullptr, nullptr, nullptr, nullptr);
             DmkStressBatch* b = grp[slot % ng].get();
             batch_enter(b);
             dmk_set_stress_batch(h[i], b, slot / ng);
             int r = dmk_grid_run_loadcase(h[i], &lc[i], nullptr, nullptr, nullptr, nullptr, nullptr);
             dmk_set_stress_batch(h[i], nullptr, -1);
             batch_leave(b);
             return r;
           },
           rc.data(), wall_s);
  print_timing(grp);
  for (auto& b : grp) batch_free(b.get());
  int bad = 0;
  for (int i = 0; i < n; i++) {
    if (rc_out) rc_out[i] = rc[i];
    if (rc[i]) bad++;
  }
  if (bad) { dmk_set_error("%s: %d of %d RVEs failed", name, bad, n); return DMK_ERR_NOCONV; }
  return 0;
}

extern "C" int dmk_batch_run(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells,
                             int32_t n_threads, int32_t* rc_out, double* wall_s) {
  return batch_run(h, n, lc, cells, n_threads, rc_out, wall_s, false, "dmk_batch_run");
}

extern "C" int dmk_batch_run_lockstep(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells,
                                      int32_t n_threads, int32_t* rc_out, double* wall_s) {
  return batch_run(h, n, lc, cells, n_threads, rc_out, wall_s, true, "dmk_batch_run_lockstep");
}

static int batch_adjoint(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                         double* const* theta_bar, const int64_t* cells, int32_t n_threads, int32_t* rc_out,
                         double* wall_s, bool lockstep, const char* name) {
  if (!h || !P_av_bar || !theta_bar || n < 1) {
    dmk_set_error("%s: invalid argument", name);
    return DMK_ERR_ARG;
  }
  std::vector<int> order = largest_first(h, n, cells);
  std::vector<int32_t> rc(n, 0);
  // each sweep in flight holds its work buffers (for a 24-system phase about
  // 33 kB per cell, most of it the factorised local Jacobians of two
  // increments); a sweep releases them when it ends, and no more sweeps run
  // at once than the free device memory holds, less a sixteenth of the device
  // for the runtime's kernel scratch and queues
  size_t need = 0;
  for (int i = 0; i < n; i++) need = std::max(need, dmk_adjoint_work_bytes(h[i]));
  int nt = std::max(1, std::min<int>(n_threads, n));
  size_t free_b = 0, total_b = 0;
  if (need && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const size_t margin = total_b / 16;
    const size_t fit = free_b > margin ? (free_b - margin) / need : 0;
    nt = (int)std::max<size_t>(1, std::min<size_t>(nt, fit));
  }
  if (progress_lines()) {
    fprintf(stderr, "%s: %d sweeps at a time, %.2f GB of work buffers each, %.1f GB free\n", name, nt,
            need * 1e-9, free_b * 1e-9);
    fflush(stderr);
  }
  std::vector<std::unique_ptr<DmkStressBatch>> grp;
  int ng = 0;
  if (lockstep) {
    // the reverse sweeps: an RVE at an increment boundary (its local
    // factorisation and cotangent reductions) would hold every round up; on
    // the 256-RVE batch a round that waits 1 ms for it ran 10 % faster
    int e = make_groups(grp, nt, ng, 1000);
    if (e) {
      dmk_set_error("%s: HIP resources of the lockstep batch", name);
      return e;
    }
  }
  run_pool(n, order, nt,
           [&](int i, int slot) {
             DmkStressBatch* b = lockstep ? grp[slot % ng].get() : nullptr;
             if (b) {
               batch_enter(b);
               dmk_set_stress_batch(h[i], b, slot / ng);
             }
             int r = dmk_adjoint_backward(h[i], P_av_bar[i], rtol, theta_bar[i], nullptr);
             if (b) {
               dmk_set_stress_batch(h[i], nullptr, -1);
               batch_leave(b);
             }
             dmk_adjoint_work_release(h[i]);
             return r;
           },
           rc.data(), wall_s);
  print_timing(grp);
  for (auto& b : grp) batch_free(b.get());
  int bad = 0;
  for (int i = 0; i < n; i++) {
    if (rc_out) rc_out[i] = rc[i];
    if (rc[i]) bad++;
  }
  if (bad) { dmk_set_error("%s: %d of %d RVEs failed", name, bad, n); return DMK_ERR_NOCONV; }
  return 0;
}

extern "C" int dmk_batch_adjoint(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                                 double* const* theta_bar, const int64_t* cells, int32_t n_threads, int32_t* rc_out,
                                 double* wall_s) {
  return batch_adjoint(h, n, P_av_bar, rtol, theta_bar, cells, n_threads, rc_out, wall_s, false, "dmk_batch_adjoint");
}

extern "C" int dmk_batch_adjoint_lockstep(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                                          double* const* theta_bar, const int64_t* cells, int32_t n_threads,
                                          int32_t* rc_out, double* wall_s) {
  return batch_adjoint(h, n, P_av_bar, rtol, theta_bar, cells, n_threads, rc_out, wall_s, true,
                       "dmk_batch_adjoint_lockstep");
}
