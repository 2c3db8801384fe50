// Host harness of the lockstep batch (damask_amd/csrc/batch.cpp).
//
// batch.cpp is compiled unchanged against the host stand-in of the HIP runtime
// in hip/hip_runtime.h: every stream is an in-order queue run by its own host
// thread, events carry generation counters, copies happen when their stream
// reaches them and the batched stress launch runs as an operation of the
// batch's stream.  The solver side (dmk_grid_run_loadcase, the handle) is
// replaced by RVEs that, per residual evaluation, write a round tag into their
// "input F" on their own stream, submit one item per phase to the batch, and
// check on their stream afterwards that the batched launch wrote the same tag
// into their "P".  The launch checks that it reads the tag of the round it was
// submitted for, i.e. that it ran after the submitter's input and before the
// submitter's next input.  Half of the RVEs submit whole residual
// evaluations instead: the round's reductions, transforms (batched, or per
// RVE for configurations without fused kernels) and report gather read their
// argument tables as uploaded for that round, and the thread checks that the
// report it is handed is its own, of that round.  A third of the RVEs submit
// dot-product requests as the adjoint GMRES does: in blocks whose requests
// after the first are idle (their stream had no work, so the round does not
// wait for it), each carrying a vector step that must see the block's input,
// with the stream's wait for the rounds deferred to the end of the block.
// Built with ThreadSanitizer,
// every host-side access of the batch (requests, staging tables, reports,
// generation counter, events) is also checked for data races.  Exit status 0:
// every check passed.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "dmk_batch.h"
#include "dmk_common.h"
#include "dmk_linalg.h"
#include "dmk_resid.h"
#include "../../include/damask_hip.h"

// ------------------------------------------------------------- runtime stand-in
struct MockStream {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  bool stop = false, busy = false;
  std::thread th;
  MockStream() : th([this] { run(); }) {}
  void run() {
    for (;;) {
      std::function<void()> f;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        f = std::move(q.front());
        q.pop_front();
        busy = true;
      }
      f();
      {
        std::lock_guard<std::mutex> lk(mu);
        busy = false;
      }
      cv.notify_all();
    }
  }
  void push(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(mu);
      q.push_back(std::move(f));
    }
    cv.notify_all();
  }
  void sync() {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return q.empty() && !busy; });
  }
};

struct MockEvent {
  std::mutex mu;
  std::condition_variable cv;
  uint64_t recorded = 0, completed = 0;
  void wait_for(uint64_t g) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return completed >= g; });
  }
};

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new MockStream; return hipSuccess; }
hipError_t hipLaunchHostFunc(hipStream_t s, hipHostFn_t fn, void* user) {
  s->push([=] { fn(user); });
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  s->sync();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->th.join();
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) { s->sync(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = new MockEvent; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  uint64_t g;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    g = ++e->recorded;
  }
  s->push([e, g] {
    {
      std::lock_guard<std::mutex> lk(e->mu);
      if (g > e->completed) e->completed = g;
    }
    e->cv.notify_all();
  });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  uint64_t g;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    g = e->recorded;
  }
  e->wait_for(g);
  return hipSuccess;
}
// set by the "drop-waits" run: stream waits are ignored, which the checks
// below must detect (shows that they can fail)
static bool g_drop_waits = false;

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  if (g_drop_waits) return hipSuccess;
  uint64_t g;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    g = e->recorded;
  }
  s->push([e, g] { e->wait_for(g); });
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  s->push([dst, src, bytes] { std::memcpy(dst, src, bytes); });
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { *p = std::malloc(bytes); return *p ? 0 : 1; }
hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipMalloc(void** p, size_t bytes) { *p = std::malloc(bytes); return *p ? 0 : 1; }
hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "mock"; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new MockEvent; return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.0f; return hipSuccess; }
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes) {
  *free_bytes = *total_bytes = (size_t)1 << 40;
  return hipSuccess;
}

// ------------------------------------------------------------- solver stand-in
struct dmk_handle {
  int id = 0, rounds = 0, n_phases = 1, nsl[DMK_MAX_PHASES] = {}, cells = 0;
  hipStream_t st = nullptr;
  DmkStressBatch* batch = nullptr;
  int slot = -1;
  int64_t Ftag = 0, Ptag = 0;            // "device" input and output of the stress integration
  int64_t Ttag = 0, Rtag = 0;            // output of the transform and of the report gather
  int64_t Vtag = 0;                      // output of a request's vector step
  bool resid = false;                    // submits whole residual evaluations
  bool dots = false;                     // submits dot-product requests as the adjoint GMRES does
  DmkFFT fft;
  std::atomic<int> errors{0};
  std::atomic<int64_t> launched{0}, transformed{0};
};

static thread_local char g_err[256];
void dmk_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
extern "C" const char* dmk_last_error(void) { return g_err; }

void dmk_set_stress_batch(dmk_handle* h, DmkStressBatch* b, int slot) {
  h->batch = b;
  h->slot = slot;
}

static std::atomic<int64_t> g_launches{0}, g_bad{0};

extern "C" int dmk_launch_stress_batch(const DmkStressItem* items_dev, int n_items, int n_blocks, int nsl,
                                       int quarter_pow, hipStream_t s) {
  g_launches++;
  s->push([=] {
    int blocks = 0;
    for (int i = 0; i < n_items; i++) {
      const DmkStressItem& it = items_dev[i];
      dmk_handle* h = reinterpret_cast<dmk_handle*>(const_cast<double*>(it.F));
      const int64_t round = (int64_t)it.dt;
      if (it.block0 != blocks || h->nsl[it.phase_id] != nsl || !quarter_pow) g_bad++;
      blocks += (it.mp.N + DMK_STRESS_WG - 1) / DMK_STRESS_WG;
      // the submitter's input of this round, not of an earlier or later one
      if (h->Ftag != round) g_bad++;
      if (it.phase_id == h->n_phases - 1) h->Ptag = round;
      h->launched++;
    }
    if (blocks != n_blocks) g_bad++;
  });
  return 0;
}

// the round's residual kernels: a request is identified by its handle (rep)
// and round (wgt); each kernel checks that it runs after the submitter's input
// of that round and after the round's earlier kernels
static dmk_handle* req_handle(const DmkResidDev& q) { return reinterpret_cast<dmk_handle*>(q.rep); }

int dmk_resid_reduce_batch(const DmkResidDev* R, int nr, int, hipStream_t s) {
  s->push([=] {
    for (int k = 0; k < nr; k++) {
      dmk_handle* h = req_handle(R[k]);
      const int64_t round = (int64_t)R[k].wgt;
      if (h->Ftag != round || h->Ptag != round) g_bad++;
    }
  });
  return 0;
}

static void transform_one(dmk_handle* h, int64_t round) {
  if (h->Ptag != round) g_bad++;
  h->Ttag = round;
  h->transformed++;
}

int dmk_transform_batch(const DmkFFT* fft, const DmkResidDev* R, int nr, hipStream_t s) {
  if (!fft->xy_threads) return 1;
  const int xy = fft->xy_threads;
  s->push([=] {
    for (int k = 0; k < nr; k++) {
      dmk_handle* h = req_handle(R[k]);
      if (h->fft.xy_threads != xy) g_bad++;      // one configuration per batched launch
      transform_one(h, (int64_t)R[k].wgt);
    }
  });
  return 0;
}

// per-RVE transforms (configuration without fused kernels): the request's
// handle rides in Pnew and its round in the Gamma arguments
int dmk_fft_r2c_xy(const DmkFFT* fft, int, const double* in, double2*, hipStream_t s) {
  if (fft->xy_threads) g_bad++;
  dmk_handle* h = reinterpret_cast<dmk_handle*>(const_cast<double*>(in));
  s->push([=] { if (h->Ptag != h->Ftag) g_bad++; });
  return 0;
}
int dmk_gamma_conv(const DmkFFT*, double2*, const GammaArgs* g, double* div_partial, int* nb, hipStream_t s) {
  dmk_handle* h = reinterpret_cast<dmk_handle*>(div_partial);
  const int64_t round = (int64_t)g->aim[0];
  *nb = 1;
  s->push([=] { transform_one(h, round); });
  return 0;
}
int dmk_fft_c2r_xy(const DmkFFT*, int, double2*, double*, hipStream_t) { return 0; }

int dmk_resid_dots_batch(const DmkResidDev* R, int nr, int, int, int, double* rep_out, hipStream_t s) {
  s->push([=] {
    for (int k = 0; k < nr; k++) {
      dmk_handle* h = req_handle(R[k]);
      const int64_t round = (int64_t)R[k].wgt;
      if (R[k].kind == DMK_RQ_DOTS ? h->Vtag != round : h->Ttag != round) g_bad++;
      rep_out[(size_t)k * REP_SIZE] = (double)round;
      rep_out[(size_t)k * REP_SIZE + 1] = (double)h->id;
      h->Rtag = round;
    }
  });
  return 0;
}

// adjoint-operator requests (the harness submits none): must never be reached with one
int dmk_adj_atmul_batch(const DmkResidDev*, int, long, hipStream_t) {
  g_bad++;
  return 0;
}
// the harness's requests carry no vector step after a transform
int dmk_post_batch(const DmkResidDev*, int, long, hipStream_t) {
  g_bad++;
  return 0;
}
int dmk_adj_mean_batch(const DmkResidDev*, int, int, hipStream_t) {
  g_bad++;
  return 0;
}
// vector steps: each dot-product request of the harness carries one, which
// must run after the submitter's input of its round
int dmk_pre_batch(const DmkResidDev* R, int nr, long, hipStream_t s) {
  s->push([=] {
    for (int k = 0; k < nr; k++) {
      if (R[k].npre == 0) continue;
      dmk_handle* h = req_handle(R[k]);
      const int64_t round = (int64_t)R[k].wgt;
      if (R[k].npre != 1 || h->Ftag != R[k].pre[0].a) g_bad++;
      h->Vtag = round;
    }
  });
  return 0;
}

size_t dmk_adjoint_work_bytes(const dmk_handle*) { return 0; }
void dmk_adjoint_work_release(dmk_handle*) {}

extern "C" int dmk_grid_run_loadcase(dmk_handle* h, const dmk_loadcase*, double*, double*, double*, double*,
                                     int64_t*) {
  std::mt19937 rng(1000u + (unsigned)h->id);
  for (int r = 1; r <= h->rounds; r++) {
    if (!h->dots || (r - 1) % 5 == 0)
      h->st->push([h, r] { h->Ftag = r; });              // input of this residual evaluation (or block)
    DmkStressItem items[DMK_MAX_PHASES];
    int key[DMK_MAX_PHASES];
    for (int p = 0; p < h->n_phases; p++) {
      std::memset(&items[p], 0, sizeof items[p]);
      items[p].mp.N = h->cells;
      items[p].F = reinterpret_cast<const double*>(h);
      items[p].dt = (double)r;
      items[p].phase_id = p;
      key[p] = 2 * h->nsl[p] + 1;
    }
    if (!h->batch) return DMK_ERR_ARG;
    if (h->dots && dmk_stress_batch_resid(h->batch)) {
      // as the adjoint GMRES: blocks of 5 rounds, the stream busy (input
      // written) before the first request of a block only, the stream's wait
      // for the rounds deferred to the end of the block
      static thread_local hipEvent_t last = nullptr;
      const bool first = (r - 1) % 5 == 0, lastr = r % 5 == 0 || r == h->rounds;
      DmkResidDev q{};
      q.kind = DMK_RQ_DOTS;
      q.idle = first ? 0 : 1;
      q.rep = reinterpret_cast<double*>(h);
      q.wgt = (double)r;
      // every third request carries no report (as the operator applications
      // whose results only later requests read): it returns at the launch
      const bool norep = r % 3 == 0;
      q.rep_n = norep ? 0 : 2;
      q.npre = 1;
      q.pre[0].type = DMK_PRE_AXPBY;
      q.pre[0].n = 1;
      q.pre[0].a = (double)(r - (r - 1) % 5);             // the block's input tag
      double rep[2] = {-1.0, -1.0};
      int rc = dmk_resid_batch_submit(h->batch, h->slot, nullptr, nullptr, 0, &q, nullptr, h->st, rep, &last);
      if (rc) return rc;
      if (!norep && (rep[0] != (double)r || rep[1] != (double)h->id)) h->errors++;
      if (norep && (rep[0] != -1.0 || rep[1] != -1.0)) h->errors++;
      if (lastr) {
        hipStreamWaitEvent(h->st, last, 0);
        h->st->push([h, r] {
          if (h->Rtag != r) h->errors++;                   // the deferred wait covers the block's last round
        });
      }
      continue;
    }
    if (h->resid && dmk_stress_batch_resid(h->batch)) {
      DmkResidDev q{};
      q.rep = reinterpret_cast<double*>(h);
      q.wgt = (double)r;
      q.Pnew = reinterpret_cast<double*>(h);
      q.divp = reinterpret_cast<double*>(h);
      q.g.aim[0] = (double)r;
      q.rep_n = 2;
      double rep[2] = {-1.0, -1.0};
      int rc = dmk_resid_batch_submit(h->batch, h->slot, items, key, h->n_phases, &q, &h->fft, h->st, rep, nullptr);
      if (rc) return rc;
      // the report of this RVE and round, read on the host after the round
      if (rep[0] != (double)r || rep[1] != (double)h->id) h->errors++;
      h->st->push([h, r] {
        if (h->Rtag != r) h->errors++;
      });
    } else {
      int rc = dmk_stress_batch_submit(h->batch, h->slot, items, key, h->n_phases, h->st);
      if (rc) return rc;
      h->st->push([h, r] {
        if (h->Ptag != r) h->errors++;                   // the stream waited for the batched launch
      });
    }
    if (rng() % 4 == 0) hipStreamSynchronize(h->st);     // a host round trip, as the solver makes
    if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
  }
  hipStreamSynchronize(h->st);
  return 0;
}

extern "C" int dmk_adjoint_backward(dmk_handle*, const double*, double, double*, double*) { return 0; }

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 200;
  const int threads = argc > 2 ? std::atoi(argv[2]) : 32;
  const int reps = argc > 3 ? std::atoi(argv[3]) : 3;
  g_drop_waits = argc > 4 && std::strcmp(argv[4], "drop-waits") == 0;
  std::mt19937 rng(7);
  int64_t expected_total = 0, bad_total = 0;
  for (int rep = 0; rep < reps; rep++) {
    std::vector<dmk_handle*> hs(n);
    std::vector<int64_t> cells(n);
    std::vector<dmk_loadcase> lcs(n);
    int64_t expected = 0, expected_tr = 0;
    for (int i = 0; i < n; i++) {
      dmk_handle* h = new dmk_handle;
      h->id = i;
      h->rounds = 1 + (int)(rng() % 40);
      h->n_phases = 1 + (int)(rng() % 3);
      for (int p = 0; p < h->n_phases; p++) h->nsl[p] = (rng() % 2) ? 12 : 24;
      h->cells = 512 * (1 + (int)(rng() % 64));
      const unsigned kind = rng() % 3;
      h->resid = kind == 1;
      h->dots = kind == 2;
      const int edge = 8 * (1 + (int)(rng() % 3));
      for (int d = 0; d < 3; d++) h->fft.n[d] = edge;
      h->fft.xy_threads = (rng() % 3) ? 256 : 0;
      h->fft.zgz_g = h->fft.xy_threads ? 2 : 0;
      if (h->resid) expected_tr += h->rounds;
      hipStreamCreateWithFlags(&h->st, 0);
      hs[i] = h;
      cells[i] = h->cells;
      if (!h->dots) expected += (int64_t)h->rounds * h->n_phases;
    }
    std::vector<int32_t> rc(n, -1);
    std::vector<double> wall(n, 0.0);
    int r = dmk_batch_run_lockstep(hs.data(), n, lcs.data(), cells.data(), threads, rc.data(), wall.data());
    int64_t launched = 0, errors = 0, transformed = 0;
    for (int i = 0; i < n; i++) {
      launched += hs[i]->launched;
      transformed += hs[i]->transformed;
      errors += hs[i]->errors;
      if (rc[i] != 0) errors++;
      hipStreamDestroy(hs[i]->st);
      delete hs[i];
    }
    std::printf("rep %d: rc %d, RVEs %d, threads %d, items launched %lld of %lld, residual transforms %lld of %lld, "
                "stream-order errors %lld, launch-check errors %lld\n", rep, r, n, threads, (long long)launched,
                (long long)expected, (long long)transformed, (long long)expected_tr, (long long)errors,
                (long long)g_bad.load());
    if (r != 0 || launched != expected || transformed != expected_tr || errors) bad_total++;
    expected_total += expected;
  }
  if (g_bad.load()) bad_total++;
  std::printf("batched launches %lld for %lld items: %s\n", (long long)g_launches.load(), (long long)expected_total,
              bad_total ? "FAILED" : "ok");
  return bad_total ? 1 : 0;
}
