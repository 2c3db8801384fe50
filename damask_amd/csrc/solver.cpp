// Host side of the MI355X spectral-basic solver and its C-ABI.
//
// Restates src/grid/grid_mech_spectral_basic.f90 (init :103-237, solution
// :243-277, forward :284-355, converged :422-458, form_residual :464-521),
// src/grid/grid_utilities.f90 (maskedCompliance :74-127, constitutiveResponse
// :133-207, calculateRate :245-264, forwardTensorField :271-297) and the
// increment/cutback loop of src/grid/DAMASK_grid.f90:288-460.  Every field
// lives on the GPU; the host keeps the 3x3 / 9x9 bookkeeping, the small NGMRES
// least-squares problem and the convergence logic.
//
// The nonlinear solver is the reference's choice, PETSc SNES "ngmres"
// (grid_mech_spectral_basic.f90:150), restated from the published method
// (Oosterlee & Washio, SIAM J. Sci. Comput. 21(5) 2000; Brune, Knepley, Smith,
// Tu, SIAM Review 57(4) 2015, sec. 4.2) with PETSc's default parameters:
// m = 30, x^M = x - r(x) (unit basic line search), difference-based selection
// (gamma_A = 2, epsilon_B = 0.1, delta_B = 0.9) and restart (gamma_C = 2,
// after 2 consecutive restart conditions), max 50 iterations.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dmk_common.h"
#include "dmk_adjoint_k.h"
#include "dmk_rot.h"
#include "dmk_linalg.h"
#include "dmk_spectral.h"
#include "dmk_polar.h"
#include "dmk_galerkin.h"
#include "dmk_batch.h"
#include "dmk_resid.h"
#include "../../include/damask_hip.h"

int dmk_build_phase(const dmk_phase_desc* d, DmkPhaseDev* p);
void dmk_qu2om(const double* qu, double* om);

static thread_local char g_err[512];

void dmk_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char* dmk_last_error(void) { return g_err; }

#define CK(x)                                                                            \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      dmk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));      \
      return DMK_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)
#define CKL(x)                                                                           \
  do {                                                                                   \
    int r_ = (x);                                                                        \
    if (r_ != 0) {                                                                       \
      dmk_set_error("%s:%d kernel launch %s failed (%d): %s", __FILE__, __LINE__, #x, r_, \
                    hipGetErrorString(hipGetLastError()));                                \
      return DMK_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)

// ------------------------------------------------------------------ host 3x3 helpers
// Fortran column-major 3x3 / 3x3x3x3 on the host side
static inline int c2(int i, int j) { return i + 3 * j; }
static inline int c4(int i, int j, int k, int l) { return i + 3 * j + 9 * k + 27 * l; }

static void rot_vec(const double* q, const double* v, double* o) { dmk_rot_vec(q, v, o); }

// rotations.f90:282-301 rotate3x3 (active conjugates q)
static void rotate33(const double* q0, const double* T, bool active, double* out) {
  dmk_rotate33(q0, T, active, out);
}

// rotations.f90:310-339 rotate3x3x3x3 (passive)
static void rotate3333(const double* q, const double* T, double* out) {
  double R[81], v[3], o[3];
  std::memcpy(R, T, sizeof R);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) {
    for (int x = 0; x < 3; x++) v[x] = R[c4(i, j, k, x)];
    rot_vec(q, v, o);
    for (int x = 0; x < 3; x++) R[c4(i, j, k, x)] = o[x];
  }
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[c4(i, j, x, l)];
    rot_vec(q, v, o);
    for (int x = 0; x < 3; x++) R[c4(i, j, x, l)] = o[x];
  }
  for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[c4(i, x, k, l)];
    rot_vec(q, v, o);
    for (int x = 0; x < 3; x++) R[c4(i, x, k, l)] = o[x];
  }
  for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[c4(x, j, k, l)];
    rot_vec(q, v, o);
    for (int x = 0; x < 3; x++) R[c4(x, j, k, l)] = o[x];
  }
  std::memcpy(out, R, sizeof R);
}

// dense inverse by LU with partial pivoting (column-major), returns true on failure
static bool invert_n(int n, const double* A, double* inv) {
  std::vector<double> a(A, A + n * n);
  std::vector<int> piv(n);
  for (int j = 0; j < n; j++) {
    int p = j;
    for (int i = j + 1; i < n; i++) if (std::fabs(a[i + n * j]) > std::fabs(a[p + n * j])) p = i;
    piv[j] = p;
    if (a[p + n * j] == 0.0) return true;
    if (p != j) for (int k = 0; k < n; k++) std::swap(a[j + n * k], a[p + n * k]);
    for (int i = j + 1; i < n; i++) a[i + n * j] /= a[j + n * j];
    for (int k = j + 1; k < n; k++)
      for (int i = j + 1; i < n; i++) a[i + n * k] -= a[i + n * j] * a[j + n * k];
  }
  for (int c = 0; c < n; c++) {
    std::vector<double> b(n, 0.0);
    b[c] = 1.0;
    for (int i = 0; i < n; i++) if (piv[i] != i) std::swap(b[i], b[piv[i]]);
    for (int j = 0; j < n; j++) for (int i = j + 1; i < n; i++) b[i] -= a[i + n * j] * b[j];
    for (int j = n - 1; j >= 0; j--) { b[j] /= a[j + n * j]; for (int i = 0; i < j; i++) b[i] -= a[i + n * j] * b[j]; }
    for (int i = 0; i < n; i++) inv[i + n * c] = b[i];
  }
  return false;
}

// Jacobi eigen-decomposition based least squares of a symmetric m x m system,
// singular values below m * eps * max treated as zero (rcond = -1 of gelss)
static void lsq_sym(int m, const double* H, const double* b, double* x) {
  std::vector<double> A(H, H + m * m), V(m * m, 0.0);
  for (int i = 0; i < m; i++) V[i + m * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < m; p++) for (int q = p + 1; q < m; q++) off += A[p + m * q] * A[p + m * q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        double apq = A[p + m * q];
        if (apq == 0.0) continue;
        double th = (A[q + m * q] - A[p + m * p]) / (2.0 * apq);
        double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; k++) {
          double akp = A[k + m * p], akq = A[k + m * q];
          A[k + m * p] = c * akp - s * akq;
          A[k + m * q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {
          double apk = A[p + m * k], aqk = A[q + m * k];
          A[p + m * k] = c * apk - s * aqk;
          A[q + m * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          double vkp = V[k + m * p], vkq = V[k + m * q];
          V[k + m * p] = c * vkp - s * vkq;
          V[k + m * q] = s * vkp + c * vkq;
        }
      }
  }
  double smax = 0.0;
  for (int i = 0; i < m; i++) smax = std::fmax(smax, std::fabs(A[i + m * i]));
  double cut = smax * 2.220446049250313e-16 * m;
  for (int i = 0; i < m; i++) x[i] = 0.0;
  for (int k = 0; k < m; k++) {
    double lam = A[k + m * k];
    if (std::fabs(lam) <= cut) continue;
    double c = 0.0;
    for (int i = 0; i < m; i++) c += V[i + m * k] * b[i];
    c /= lam;
    for (int i = 0; i < m; i++) x[i] += c * V[i + m * k];
  }
}

// ------------------------------------------------------------------ handle
// the driver's inputs and state at the start of one increment
// (DAMASK_grid.f90:318-330): enough to run the increment again
struct IncStart {
  dmk_loadcase lc;
  double t_0 = 0, t = 0, Delta_t = 0, Delta_t_prev = 0;
  int cutBack = 0, cutBackLevel = 0, guess = 0;
  bool valid = false;
};

struct dmk_handle {
  int cells[3];
  double size[3];
  long N, Nk, n9;
  int h;
  double wgt;
  hipStream_t st = nullptr;
  DmkFFT fft;
  int n_phases;
  std::vector<DmkPhaseDev> ph;
  DmkPhaseDev* ph_dev = nullptr;
  int nsl_max;
  DmkNumMech nm;
  dmk_numerics num;
  DmkMPDev mp;
  std::vector<void*> allocs;
  double *P_hom, *dPdF;
  double *F, *F_lastinc, *Fdot, *R, *XM, *FM, *XA, *FA, *Xh, *Fh, *Fext, *Pext;
  double2 *Z, *U;
  double *u3, *xn_dev, *xp_dev;
  double *partial, *small, *nrm, *pmax, *pmin, *ext;
  double* divp = nullptr;            // per-block divergence sums of the basic residual
  int div_nb = 0;                    // how many
  bool div_pending = false;          // divp holds sums the next residual_finish adds up
  // a basic residual evaluation left to residual_finish, which hands it to
  // the lockstep batch whole (input X, output r)
  bool lk_pending = false;
  const double* lk_X = nullptr;
  double* lk_r = nullptr;
  long *sel, *csel;
  int* flag;
  unsigned long long* cnt;
  double* hsmall = nullptr;          // pinned host mirror of small results
  // solver state
  double F_aimDot[9], F_aim[9], F_aim_lastinc[9], P_av[9], P_aim[9];
  double C_volAvg[81], C_volAvgLastInc[81], C_minMaxAvg[81], C_minMaxAvgLastInc[81], S[81];
  double C_minMaxAvgRestart[81];     // reference stiffness of Gamma, what restartWrite stores
  double err_div = 0, err_BC = 0, err_curl = 0;
  // scheme 0 spectral_basic, 1 spectral_polarization (the iterate is [F; F_tau],
  // 2 n9 doubles; C_scale / S_scale and alpha, beta travel in pol),
  // 2 spectral_Galerkin
  int scheme = 0;
  long nx = 0;
  PolArgs pol{};
  double* T = nullptr;
  int status = 0, totalIter = 0, nfuncs = 0;
  int stress_mask[9];
  double q_R[4];
  double dt = 0;
  GammaArgs ga;
  std::vector<double> Q;
  std::vector<double*> poolX, poolF;  // NGMRES iterate / residual buffers (M + 3 each)
  struct AdjState* adj = nullptr;    // adjoint tape and work buffers (dmk_adjoint_*)
  // driver state (DAMASK_grid.f90:78-100)
  double t = 0.0, Delta_t = 1.0, Delta_t_prev = 0.0;
  int cutBack = 0, cutBackLevel = 0, totalIncs = 0;
  int restartInc = -1;               // increments up to this one are skipped (restart)
  int64_t n_updates = 0, n_residuals = 0;
  // optional per-kernel timing with HIP events on the solver stream
  int profile = 0;
  hipEvent_t ev[8][2];
  double kt_ms[8] = {0};
  int64_t kt_n[8] = {0};
  bool ev_ready = false;
  // lockstep batch (batch.cpp): when set, the stress integration of each
  // residual evaluation joins one launch with the other RVEs of the batch
  DmkStressBatch* batch = nullptr;
  int batch_slot = -1;
  // driver state at the start of the increment in progress (run_increment),
  // kept per recorded increment by the adjoint tape
  IncStart inc_start;
};

void dmk_set_stress_batch(dmk_handle* h, DmkStressBatch* b, int slot) {
  h->batch = b;
  h->batch_slot = slot;
}

enum { KT_STRESS = 0, KT_TANGENT = 1, KT_REDUCE = 2, KT_FFT_FWD = 3, KT_GAMMA = 4, KT_FFT_INV = 5, KT_CAVG = 6,
       KT_N = 7 };

// profile level 1 times the stress kernel only, level 2 every group (each
// event record costs the stream a few microseconds between kernels)
static bool kt_on(const dmk_handle* h, int k) { return h->profile >= (k == KT_STRESS ? 1 : 2); }
static void kt_begin(dmk_handle* h, int k) {
  if (kt_on(h, k)) hipEventRecord(h->ev[k][0], h->st);
}
static void kt_end(dmk_handle* h, int k) {
  if (kt_on(h, k)) hipEventRecord(h->ev[k][1], h->st);
}
// after a stream synchronisation: accumulate the recorded intervals
static void kt_collect(dmk_handle* h, int k) {
  if (!kt_on(h, k)) return;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, h->ev[k][0], h->ev[k][1]) == hipSuccess) {
    h->kt_ms[k] += ms;
    h->kt_n[k] += 1;
  }
}

template <typename T>
static int dalloc(dmk_handle* h, T** p, size_t n) {
  void* v = nullptr;
  CK(hipMalloc(&v, sizeof(T) * (n ? n : 1)));
  CK(hipMemsetAsync(v, 0, sizeof(T) * (n ? n : 1), h->st));
  h->allocs.push_back(v);
  *p = (T*)v;
  return 0;
}

enum { SM_PAV = 0, SM_CVOL = 9, SM_CMM = 90, SM_DOTS = 180, SM_REP = 256, SM_FAIM = 352, SM_AIM = 368,
       // adjoint: aim bookkeeping, report, mean, dot products, maxima, tangent average
       SM_AFAIM = 384, SM_AAIM = 400, SM_AREP = 416, SM_APAV = 504, SM_ADOTS = 520, SM_AMAX = 584, SM_ACAV = 592,
       // polarization: mean of the compatible field F; Galerkin: DC of the residual
       SM_FAV = 680, SM_GDC = 690, SM_SIZE = 704 };
static_assert(SM_AREP + REP_SIZE <= SM_APAV, "adjoint report region");
static_assert(SM_ADOTS + DMK_MAX_DOTS <= SM_AMAX, "adjoint dot region");
static_assert(SM_REP + REP_SIZE <= SM_FAIM, "report region");

static void gamma_coef(double wgt, const double* C, GammaArgs* g) {
  // A_lm = sum_jq C_ref(l,j,m,q) a_j a_q with C_ref = C / wgt; products a_j a_q
  // ordered (00, 11, 22, 12, 02, 01)
  static const int pj[6] = {0, 1, 2, 1, 0, 0}, pq[6] = {0, 1, 2, 2, 2, 1};
  for (int l = 0; l < 3; l++)
    for (int m = 0; m < 3; m++)
      for (int p = 0; p < 6; p++) {
        int j = pj[p], q = pq[p];
        double v = C[c4(l, j, m, q)] / wgt;
        if (j != q) v += C[c4(l, q, m, j)] / wgt;
        g->Ccoef[3 * l + m][p] = v;
      }
}

static void set_gamma_coef(dmk_handle* h, const double* C) { gamma_coef(h->wgt, C, &h->ga); }

// Orientation, dPdF and the stiffness averages C_volAvg / C_minMaxAvg of the
// last response (homogenization_mechanical_response's second loop and the
// averages of utilities_constitutiveResponse, grid_utilities.f90:163-214).
// The spectral-basic solver reads the averages only between solves (solution
// start, forward, restart; grid_mech_spectral_basic.f90:261-262,306-311), so
// the solver runs this once after each solve instead of after every residual
// evaluation: the values it reads are those of the last response either way.
// Skipped when the last response did not converge (the reference then keeps
// the previous dPdF).
static int tangent_pass(dmk_handle* h, double dt) {
  if (h->status != ST_OK) return 0;
  kt_begin(h, KT_TANGENT);
  for (int p = 0; p < h->n_phases; p++)
    CKL(dmk_launch_tangent(&h->mp, h->ph_dev + p, &h->ph[p], p, dt, h->flag, h->dPdF, h->nrm, h->st));
  kt_end(h, KT_TANGENT);
  kt_begin(h, KT_CAVG);
  CKL(dmk_colsum(h->dPdF, h->N, h->N, 81, h->partial, h->small + SM_CVOL, h->wgt, h->st));
  CKL(dmk_cminmax(h->dPdF, h->N, h->num.cref_rule == 1 ? 1.0e-12 : 0.0, h->nrm, h->pmax, h->pmin, h->csel, h->sel,
                  h->ext, h->small + SM_CMM, h->st));
  kt_end(h, KT_CAVG);
  CK(hipMemcpyAsync(h->hsmall + SM_CVOL, h->small + SM_CVOL, sizeof(double) * 162, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  kt_collect(h, KT_TANGENT);
  kt_collect(h, KT_CAVG);
  std::memcpy(h->C_volAvg, h->hsmall + SM_CVOL, sizeof h->C_volAvg);
  std::memcpy(h->C_minMaxAvg, h->hsmall + SM_CMM, sizeof h->C_minMaxAvg);
  return 0;
}

// utilities_constitutiveResponse on the device field F (component-major),
// enqueued without synchronisation: stress integration of every cell,
// convergence flag, accepted P and its raw volume average, and with a the
// report and stress-BC bookkeeping of dmk_residual_aim.  The stress kernel
// writes into the spare P buffer; when a cell failed, the previous P (and its
// average) stays: it is copied back into that buffer (the reference's skipped
// second loop leaves P as it was, homogenization.f90:349), and the buffers
// swap roles.
static int response_enqueue(dmk_handle* h, const double* F, double dt, const AimArgs* a) {
  kt_begin(h, KT_STRESS);
  if (h->batch) {
    DmkStressItem items[DMK_MAX_PHASES];
    int key[DMK_MAX_PHASES];
    for (int p = 0; p < h->n_phases; p++) {
      items[p] = DmkStressItem{h->mp, F, h->ph_dev + p, h->nm, dt, p, 0};
      key[p] = 2 * h->ph[p].nsl + (h->ph[p].quarter_pow ? 1 : 0);
    }
    int rc = dmk_stress_batch_submit(h->batch, h->batch_slot, items, key, h->n_phases, h->st);
    if (rc) return rc;
  } else {
    for (int p = 0; p < h->n_phases; p++)
      CKL(dmk_launch_stress(&h->mp, F, h->ph_dev + p, &h->ph[p], p, &h->nm, dt, h->st));
  }
  kt_end(h, KT_STRESS);
  kt_begin(h, KT_REDUCE);
  CKL(dmk_response_reduce(h->mp.status, h->mp.n_newton, h->mp.n_fpi, h->mp.Pph, h->N, h->wgt, h->partial, h->flag,
                          h->cnt, h->small + SM_PAV, a, h->small + SM_FAIM, h->small + SM_REP, h->small + SM_AIM,
                          h->st));
  CKL(dmk_accept_P(h->flag, h->P_hom, h->mp.Pph, h->n9, h->st));
  std::swap(h->P_hom, h->mp.Pph);
  kt_end(h, KT_REDUCE);
  h->n_updates += h->N;
  return 0;
}

static void aim_args(const dmk_handle* h, bool rotate, bool update, AimArgs* a) {
  std::memcpy(a->S, h->S, sizeof a->S);
  std::memcpy(a->P_aim, h->P_aim, sizeof a->P_aim);
  std::memcpy(a->q_R, h->q_R, sizeof a->q_R);
  for (int k = 0; k < 9; k++) a->stress_mask[k] = h->stress_mask[k];
  a->rotate = rotate;
  a->update = update;
}

// constitutive response with the report read back (init, dmk_constitutive*)
static int constitutive(dmk_handle* h, const double* F, double dt, bool rotate, bool tangent) {
  AimArgs a;
  aim_args(h, rotate, false, &a);
  int rc = response_enqueue(h, F, dt, &a);
  if (rc) return rc;
  CK(hipMemcpyAsync(h->hsmall + SM_REP, h->small + SM_REP, sizeof(double) * REP_DIV2, hipMemcpyDeviceToHost,
                    h->st));
  CK(hipStreamSynchronize(h->st));
  kt_collect(h, KT_STRESS);
  kt_collect(h, KT_REDUCE);
  h->status = (int)h->hsmall[SM_REP + REP_FLAG];
  std::memcpy(h->P_av, h->hsmall + SM_REP + REP_PAV, sizeof h->P_av);
  if (tangent) return tangent_pass(h, dt);
  return 0;
}

// grid_utilities.f90:74-127 utilities_maskedCompliance of stiffness C for the
// load-case frame q_R and stress mask (1 = component not prescribed)
static int masked_compliance_of(const double* C, const int* stress_mask, const double* q_R, double* S) {
  int msv[9], sel[9], n = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) msv[3 * a + b] = !stress_mask[c2(a, b)];
  for (int i = 0; i < 9; i++) if (msv[i]) sel[n++] = i;
  double T99[81] = {0};
  if (n > 0) {
    double Cr[81], C99[81], c[81], s[81];
    rotate3333(q_R, C, Cr);
    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) for (int cc = 0; cc < 3; cc++) for (int d = 0; d < 3; d++)
      C99[(3 * a + b) + 9 * (3 * cc + d)] = Cr[c4(a, b, cc, d)];
    for (int q = 0; q < n; q++) for (int p = 0; p < n; p++) c[p + n * q] = C99[sel[p] + 9 * sel[q]];
    bool err = invert_n(n, c, s);
    for (int k = 0; k < n * n; k++) if (s[k] != s[k]) err = true;
    if (err) { dmk_set_error("matrix inversion error (masked compliance)"); return DMK_ERR_CONFIG; }
    for (int q = 0; q < n; q++) for (int p = 0; p < n; p++) T99[sel[p] + 9 * sel[q]] = s[p + n * q];
  }
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) for (int cc = 0; cc < 3; cc++) for (int d = 0; d < 3; d++)
    S[c4(a, b, cc, d)] = T99[(3 * a + b) + 9 * (3 * cc + d)];
  return 0;
}

static int masked_compliance(dmk_handle* h) {
  return masked_compliance_of(h->C_volAvg, h->stress_mask, h->q_R, h->S);
}

// C_scale = C_minMaxAvg, S_scale = invSym3333(C_minMaxAvg)
// (grid_mech_spectral_polarization.f90:266-268, :301-305)
static int set_scale(dmk_handle* h) {
  std::memcpy(h->pol.C, h->C_minMaxAvg, sizeof h->pol.C);
  if (dmk_inv_sym3333(h->C_minMaxAvg, h->pol.S)) { dmk_set_error("matrix inversion error (S_scale)"); return DMK_ERR_CONFIG; }
  return 0;
}

// formResidual (grid_mech_spectral_basic.f90:464-521), enqueued: response,
// stress-BC update of F_aim and the aim on the device, one forward FFT for
// the divergence RMS and the Gamma convolution, inverse FFT into r.
static int residual_enqueue_pol(dmk_handle* h, const double* X, double* r);
static int residual_enqueue_gk(dmk_handle* h, const double* X, double* r);

static int residual_enqueue(dmk_handle* h, const double* X, double* r, int iter) {
  if (h->nfuncs == 0 && iter == 0) h->totalIter = -1;
  if (h->totalIter <= iter) h->totalIter++;
  h->nfuncs++;
  h->n_residuals++;
  if (h->scheme == 1) return residual_enqueue_pol(h, X, r);
  if (h->scheme == 2) return residual_enqueue_gk(h, X, r);
  if (h->batch && dmk_stress_batch_resid(h->batch)) {
    h->lk_pending = true;
    h->lk_X = X;
    h->lk_r = r;
    return 0;
  }
  AimArgs a;
  aim_args(h, true, true, &a);
  int rc = response_enqueue(h, X, h->dt, &a);
  if (rc) return rc;
  // x-y passes, then the z passes with the Gamma convolution and the
  // divergence between them (fused when the FFT tuner found that faster)
  kt_begin(h, KT_FFT_FWD);
  CKL(dmk_fft_r2c_xy(&h->fft, 9, h->P_hom, h->Z, h->st));
  kt_end(h, KT_FFT_FWD);
  h->ga.do_div = 1;
  h->ga.do_gamma = 1;
  h->ga.aim_dev = h->small + SM_AIM;
  kt_begin(h, KT_GAMMA);
  CKL(dmk_gamma_conv(&h->fft, h->Z, &h->ga, h->divp, &h->div_nb, h->st));
  kt_end(h, KT_GAMMA);
  h->div_pending = true;            // summed with the dot products in residual_finish
  kt_begin(h, KT_FFT_INV);
  CKL(dmk_fft_c2r_xy(&h->fft, 9, h->Z, r, h->st));
  kt_end(h, KT_FFT_INV);
  return 0;
}

// formResidual of the polarization scheme (grid_mech_spectral_polarization.f90
// :530-634), enqueued.  X = [F; F_tau], r = [r_F; r_F_tau]:
//   r_F_tau = beta F - Gamma(beta C:(F - I) - alpha F (C:(F_tau - F - I))),
//             the convolution's mean set to rotate(beta F_aim);
//   response at F - r_F_tau / beta (written into r_F's slot), curl RMS of F,
//   divergence RMS of P, consistent tangent of every cell;
//   r_F = invSym3333(dPdF + C):(P - F (C:(F_tau - F - I))) + r_F_tau;
//   F_aim -= S:(P_av - P_aim) and err_BC on the device.
static int residual_enqueue_pol(dmk_handle* h, const double* X, double* r) {
  const long N = h->N, n9 = h->n9;
  const double* F = X;
  const double* Ft = X + n9;
  double* rF = r;
  double* rt = r + n9;
  CKL(dmk_colsum(F, N, N, 9, h->partial, h->small + SM_FAV, h->wgt, h->st));
  CKL(dmk_pol_aim_pre(h->q_R, h->pol.beta, h->small + SM_FAIM, h->small + SM_AIM, h->st));
  CKL(dmk_pol_tau(N, &h->pol, F, Ft, h->T, h->st));
  kt_begin(h, KT_FFT_FWD);
  CKL(dmk_fft_r2c(&h->fft, 9, h->T, h->Z, h->st));
  kt_end(h, KT_FFT_FWD);
  GammaArgs g = h->ga;
  g.do_div = 0;
  g.do_gamma = 1;
  g.aim_dev = h->small + SM_AIM;
  kt_begin(h, KT_GAMMA);
  CKL(dmk_gamma_div(h->Z, &g, h->partial, h->st));
  kt_end(h, KT_GAMMA);
  kt_begin(h, KT_FFT_INV);
  CKL(dmk_fft_c2r(&h->fft, 9, h->Z, h->T, h->st));
  kt_end(h, KT_FFT_INV);
  CKL(dmk_pol_stage2(N, h->pol.beta, F, h->T, rt, rF, h->st));
  int rc = response_enqueue(h, rF, h->dt, nullptr);
  if (rc) return rc;
  int nb = dmk_gamma_div_blocks(&h->ga);
  CKL(dmk_fft_r2c(&h->fft, 9, F, h->Z, h->st));
  CKL(dmk_curl_partial(h->Z, &h->ga, h->partial, h->st));
  CKL(dmk_colsum(h->partial, nb, nb, 1, h->partial + nb, h->small + SM_REP + REP_CURL2, 1.0, h->st));
  CKL(dmk_fft_r2c(&h->fft, 9, h->P_hom, h->Z, h->st));
  g.do_div = 1;
  g.do_gamma = 0;
  CKL(dmk_gamma_div(h->Z, &g, h->partial, h->st));
  CKL(dmk_colsum(h->partial, nb, nb, 1, h->partial + nb, h->small + SM_REP + REP_DIV2, 1.0, h->st));
  kt_begin(h, KT_TANGENT);
  for (int p = 0; p < h->n_phases; p++)
    CKL(dmk_launch_tangent(&h->mp, h->ph_dev + p, &h->ph[p], p, h->dt, h->flag, h->dPdF, h->nrm, h->st));
  kt_end(h, KT_TANGENT);
  CKL(dmk_pol_rF(N, &h->pol, F, Ft, h->P_hom, h->dPdF, rt, rF, h->st));
  AimArgs a;
  aim_args(h, true, true, &a);
  CKL(dmk_pol_residual_aim(&a, &h->pol, h->small + SM_PAV, h->small + SM_FAV, h->small + SM_FAIM, h->flag,
                           h->small + SM_REP, h->st));
  return 0;
}

// formResidual of the Galerkin scheme (grid_mech_spectral_Galerkin.f90:543-601),
// enqueued: response and its consistent tangent (read by the matrix-free
// Jacobian at this iterate), the stress-BC term dP, the divergence RMS and the
// G convolution of P from one forward FFT, DC term dP / wgt on the
// stress-controlled components and zero on the others, inverse FFT into r.
// F_aim is not touched here: it changes once per Newton step (set_F_aim).
static int residual_enqueue_gk(dmk_handle* h, const double* X, double* r) {
  int rc = response_enqueue(h, X, h->dt, nullptr);
  if (rc) return rc;
  kt_begin(h, KT_TANGENT);
  for (int p = 0; p < h->n_phases; p++)
    CKL(dmk_launch_tangent(&h->mp, h->ph_dev + p, &h->ph[p], p, h->dt, h->flag, h->dPdF, h->nrm, h->st));
  kt_end(h, KT_TANGENT);
  AimArgs a;
  aim_args(h, true, false, &a);
  CKL(dmk_gk_bc(&a, h->small + SM_PAV, h->flag, h->small + SM_REP, h->small + SM_GDC, h->st));
  kt_begin(h, KT_FFT_FWD);
  CKL(dmk_fft_r2c(&h->fft, 9, h->P_hom, h->Z, h->st));
  kt_end(h, KT_FFT_FWD);
  GalerkinArgs ga{};
  ga.do_div = 1;
  ga.residual = 1;
  for (int k = 0; k < 9; k++) ga.stress_mask[k] = h->stress_mask[k];
  ga.dc_dev = h->small + SM_GDC;
  ga.dc_scale = 1.0 / h->wgt;
  kt_begin(h, KT_GAMMA);
  CKL(dmk_galerkin(h->Z, &h->ga, &ga, h->partial, h->st));
  kt_end(h, KT_GAMMA);
  int nb = dmk_gamma_div_blocks(&h->ga);
  CKL(dmk_colsum(h->partial, nb, nb, 1, h->partial + nb, h->small + SM_REP + REP_DIV2, 1.0, h->st));
  kt_begin(h, KT_FFT_INV);
  CKL(dmk_fft_c2r(&h->fft, 9, h->Z, r, h->st));
  kt_end(h, KT_FFT_INV);
  return 0;
}

static int residual_report(dmk_handle* h, int nd, double* dots_out);
static int residual_lockstep(dmk_handle* h, const DotJobs* jobs, double* dots_out);

// The one synchronisation of a residual evaluation: optional dot products
// (written after the report), report to the host, solver scalars updated.
static int residual_finish(dmk_handle* h, const DotJobs* jobs, double* dots_out) {
  if (h->lk_pending) return residual_lockstep(h, jobs, dots_out);
  int nd = jobs ? jobs->n : 0;
  if (nd > DMK_MAX_DOTS) { dmk_set_error("too many dot products"); return DMK_ERR_ARG; }
  SumJobs extra;
  extra.k = 0;
  if (h->div_pending) {
    extra.p[0] = h->divp;
    extra.n[0] = h->div_nb;
    extra.scale[0] = 1.0;
    extra.out[0] = h->small + SM_REP + REP_DIV2;
    extra.k = 1;
    h->div_pending = false;
  }
  if (nd || extra.k) CKL(dmk_dots_sums(jobs, h->nx, h->partial, h->small + SM_REP + REP_DOTS, &extra, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_REP, h->small + SM_REP, sizeof(double) * (REP_DOTS + nd), hipMemcpyDeviceToHost,
                    h->st));
  CK(hipStreamSynchronize(h->st));
  kt_collect(h, KT_STRESS);
  kt_collect(h, KT_REDUCE);
  kt_collect(h, KT_FFT_FWD);
  kt_collect(h, KT_GAMMA);
  kt_collect(h, KT_FFT_INV);
  if (h->scheme != 0) kt_collect(h, KT_TANGENT);
  return residual_report(h, nd, dots_out);
}

// the report of a residual evaluation (in hsmall) into the solver scalars
static int residual_report(dmk_handle* h, int nd, double* dots_out) {
  const double* rep = h->hsmall + SM_REP;
  std::memcpy(h->P_av, rep + REP_PAV, sizeof h->P_av);
  if (h->scheme != 2) std::memcpy(h->F_aim, rep + REP_FAIM, sizeof h->F_aim);
  h->err_BC = rep[REP_ERRBC];
  h->status = (int)rep[REP_FLAG];
  h->err_div = std::sqrt(rep[REP_DIV2]) * h->wgt;
  if (h->cells[0] == 1) h->err_div *= 0.5;
  if (h->scheme == 1) {
    h->err_curl = std::sqrt(rep[REP_CURL2]) * h->wgt;
    if (h->cells[0] == 1) h->err_curl *= 0.5;
  }
  if (nd) std::memcpy(dots_out, rep + REP_DOTS, sizeof(double) * nd);
  return 0;
}

// The basic residual evaluation left by residual_enqueue, as one RVE of a
// lockstep round: the same launches as response_enqueue, residual_enqueue
// and residual_finish, with the same arguments, run by the round over all its
// RVEs; the round's report block of this RVE replaces the copy to hsmall.
static int residual_lockstep(dmk_handle* h, const DotJobs* jobs, double* dots_out) {
  h->lk_pending = false;
  const int nd = jobs ? jobs->n : 0;
  if (nd > DMK_MAX_DOTS) { dmk_set_error("too many dot products"); return DMK_ERR_ARG; }
  DmkStressItem items[DMK_MAX_PHASES];
  int key[DMK_MAX_PHASES];
  for (int p = 0; p < h->n_phases; p++) {
    items[p] = DmkStressItem{h->mp, h->lk_X, h->ph_dev + p, h->nm, h->dt, p, 0};
    key[p] = 2 * h->ph[p].nsl + (h->ph[p].quarter_pow ? 1 : 0);
  }
  DmkResidDev q{};
  q.st = h->mp.status;
  q.nn = h->mp.n_newton;
  q.nf = h->mp.n_fpi;
  q.Pnew = h->mp.Pph;
  q.N = h->N;
  q.wgt = h->wgt;
  q.partial = h->partial;
  q.nc_resp = dmk_nchunk(h->N);
  q.flag = h->flag;
  q.cnt = h->cnt;
  q.pav = h->small + SM_PAV;
  q.faim = h->small + SM_FAIM;
  q.rep = h->small + SM_REP;
  q.aim = h->small + SM_AIM;
  aim_args(h, true, true, &q.a);
  q.Pold = h->P_hom;
  q.n9 = h->n9;
  std::swap(h->P_hom, h->mp.Pph);
  h->n_updates += h->N;
  h->ga.do_div = 1;
  h->ga.do_gamma = 1;
  h->ga.aim_dev = h->small + SM_AIM;
  q.Z = h->Z;
  q.r = h->lk_r;
  q.divp = h->divp;
  q.g = h->ga;
  if (jobs) q.J = *jobs;
  q.J.n = nd;
  q.nx = h->nx;
  q.nc_dots = dmk_nchunk(h->nx);
  // the divergence sum first, then the dot products (as dmk_dots_sums orders them)
  q.S.p[0] = h->divp;
  q.S.n[0] = dmk_gamma_conv_nb(&h->fft);
  q.S.scale[0] = 1.0;
  q.S.out[0] = h->small + SM_REP + REP_DIV2;
  q.S.k = 1;
  for (int j = 0; j < nd; j++, q.S.k++) {
    q.S.p[q.S.k] = h->partial + (long)j * q.nc_dots;
    q.S.n[q.S.k] = q.nc_dots;
    q.S.scale[q.S.k] = 1.0;
    q.S.out[q.S.k] = h->small + SM_REP + REP_DOTS + j;
  }
  q.rep_n = REP_DOTS + nd;
  int rc = dmk_resid_batch_submit(h->batch, h->batch_slot, items, key, h->n_phases, &q, &h->fft, h->st,
                                  h->hsmall + SM_REP, nullptr);
  if (rc) return rc;
  return residual_report(h, nd, dots_out);
}

static int form_residual(dmk_handle* h, const double* X, double* r, int iter) {
  int rc = residual_enqueue(h, X, r, iter);
  return rc ? rc : residual_finish(h, nullptr, nullptr);
}

// grid_mech_spectral_basic.f90:422-458 converged; the polarization scheme
// adds the curl criterion (grid_mech_spectral_polarization.f90:482-506)
static int converged(const dmk_handle* h) {
  double pmax = 0.0;
  for (int k = 0; k < 9; k++) pmax = std::fmax(pmax, std::fabs(h->P_av[k]));
  double divTol = std::fmax(pmax * h->num.eps_rel_div, h->num.eps_abs_div);
  double BCTol = std::fmax(pmax * h->num.eps_rel_P, h->num.eps_abs_P);
  bool curl_ok = true;
  if (h->scheme == 1) {
    double fmax = 0.0;
    for (int k = 0; k < 9; k++) fmax = std::fmax(fmax, std::fabs(h->F_aim[k] - (k % 4 == 0 ? 1.0 : 0.0)));
    double curlTol = std::fmax(fmax * h->num.eps_rel_curl, h->num.eps_abs_curl);
    curl_ok = h->err_curl / curlTol < 1.0;
  }
  if (h->totalIter >= h->num.N_iter_min && h->err_div / divTol < 1.0 && curl_ok && h->err_BC / BCTol < 1.0 &&
      h->status == ST_OK)
    return 1;
  if (h->totalIter >= h->num.N_iter_max) return -1;
  return 0;
}

static void add_dot(DotJobs* J, const double* x, const double* y, int mode) {
  J->x[J->n] = x;
  J->y[J->n] = y;
  J->mode[J->n] = mode;
  J->n++;
}

// PETSc SNES NGMRES (snes/impls/ngmres, difference selection and restart,
// basic line search on x^M = x - r(x)), restated; the dot products each step
// needs are computed with its residual so that a residual evaluation costs
// one host synchronisation.  Iterates and residuals live in two fixed pools
// of M + 3 buffers: the history slots, the current pair and the two trial
// pairs are pointers into them, so a step moves no field (the current pair
// enters the history by reference; trial buffers are whichever pool buffers
// neither the current pair nor the history holds).
static double* pool_free(const std::vector<double*>& pool, const double* cur, double* const* hist, int l,
                         const double* also) {
  for (double* b : pool) {
    if (b == cur || b == also) continue;
    bool used = false;
    for (int i = 0; i < l && !used; i++) used = hist[i] == b;
    if (!used) return b;
  }
  return nullptr;
}

static int solve_ngmres(dmk_handle* h, int* reason_out) {
  const int M = h->num.ngmres_m;
  const double gammaA = 2.0, gammaC = 2.0, epsB = 0.1, deltaB = 0.9;
  const int restart_it = 2;
  int rc;
  DotJobs J;
  J.n = 0;
  add_dot(&J, h->R, h->R, 0);
  double d[DMK_MAX_DOTS];
  if ((rc = residual_enqueue(h, h->F, h->R, 0)) || (rc = residual_finish(h, &J, d))) return rc;
  double fnorm = std::sqrt(d[0]);
  int reason = converged(h);
  if (reason || fnorm != fnorm) { *reason_out = reason ? reason : -1; return 0; }
  double fminnorm = fnorm;
  std::vector<double*> Xs(M, nullptr), Fs(M, nullptr);
  // history slot 0 <- (X, F), by reference
  Xs[0] = h->F;
  Fs[0] = h->R;
  h->Q[0] = d[0];
  int k_restart = 1, l = 1, restart_count = 0;
  for (int k = 1; k <= h->num.snes_max_it; k++) {
    int ivec = k_restart % M;
    double* XM = pool_free(h->poolX, h->F, Xs.data(), l, nullptr);
    double* FM = pool_free(h->poolF, h->R, Fs.data(), l, nullptr);
    double* XA = pool_free(h->poolX, h->F, Xs.data(), l, XM);
    double* FA = pool_free(h->poolF, h->R, Fs.data(), l, FM);
    if (!XM || !FM || !XA || !FA) { dmk_set_error("NGMRES buffer pool exhausted"); return DMK_ERR_CONFIG; }
    CKL(dmk_axpby(h->nx, 1.0, h->F, -1.0, h->R, XM, h->st));
    J.n = 0;
    add_dot(&J, FM, FM, 0);
    for (int i = 0; i < l; i++) add_dot(&J, FM, Fs[i], 0);
    if ((rc = residual_enqueue(h, XM, FM, k - 1)) || (rc = residual_finish(h, &J, d))) return rc;
    double fMnorm = std::sqrt(d[0]);
    double xi[32];
    for (int i = 0; i < l; i++) xi[i] = d[1 + i];
    double nu = fMnorm * fMnorm;
    double beta[32], H[32 * 32], alpha[32];
    for (int i = 0; i < l; i++) beta[i] = nu - xi[i];
    for (int j = 0; j < l; j++)
      for (int i = 0; i < l; i++) H[i + l * j] = h->Q[i + M * j] - xi[i] - xi[j] + nu;
    lsq_sym(l, H, beta, alpha);
    CombineArgs ca;
    ca.l = l;
    ca.atot = 0.0;
    for (int i = 0; i < l; i++) { ca.alpha[i] = alpha[i]; ca.atot += alpha[i]; ca.X[i] = Xs[i]; }
    CKL(dmk_combine(h->nx, XM, &ca, XA, h->st));
    J.n = 0;
    add_dot(&J, FA, FA, 0);
    add_dot(&J, XA, XM, 1);
    for (int i = 0; i < l; i++) add_dot(&J, XA, Xs[i], 1);
    for (int i = 0; i < l; i++) add_dot(&J, FA, Fs[i], 0);
    if ((rc = residual_enqueue(h, XA, FA, k - 1)) || (rc = residual_finish(h, &J, d))) return rc;
    double fAnorm = std::sqrt(d[0]);
    double dnorm = std::sqrt(d[1]);
    double dminnorm = -1.0, eta[32];
    for (int i = 0; i < l; i++) {
      double dc = std::sqrt(d[2 + i]);
      if (dc < dminnorm || dminnorm < 0.0) dminnorm = dc;
      eta[i] = d[2 + l + i];
    }
    bool selectA = true;
    if (fAnorm >= gammaA * fminnorm) selectA = false;
    if (!(epsB * dnorm < dminnorm || std::sqrt(fnorm) < deltaB * std::sqrt(fminnorm))) selectA = false;
    const double* fsel_dots = selectA ? eta : xi;
    double fsel2 = selectA ? fAnorm * fAnorm : nu;
    h->F = selectA ? XA : XM;
    h->R = selectA ? FA : FM;
    fnorm = selectA ? fAnorm : fMnorm;
    bool selectRestart = false;
    if (epsB * dnorm > dminnorm && std::sqrt(fAnorm) > deltaB * std::sqrt(fminnorm) && l > 0) selectRestart = true;
    if (std::sqrt(fAnorm) > gammaC * std::sqrt(fminnorm)) selectRestart = true;
    restart_count = selectRestart ? restart_count + 1 : 0;
    if (restart_count >= restart_it) {
      // history restarts from the M pair
      restart_count = 0;
      k_restart = 1;
      l = 1;
      Xs[0] = XM;
      Fs[0] = FM;
      h->Q[0] = nu;
    } else {
      int l_old = l;
      if (l < M) l++;
      k_restart++;
      if (fminnorm > fnorm) fminnorm = fnorm;
      // history slot ivec <- (X, F), by reference; Gram entries against the other slots
      Xs[ivec] = h->F;
      Fs[ivec] = h->R;
      for (int i = 0; i < l; i++) {
        double v = i == ivec ? fsel2 : (i < l_old ? fsel_dots[i] : 0.0);
        h->Q[ivec + M * i] = v;
        h->Q[i + M * ivec] = v;
      }
    }
    reason = converged(h);
    if (reason || fnorm != fnorm) { *reason_out = reason ? reason : -1; return 0; }
  }
  *reason_out = -1;
  return 0;
}

static int solve_richardson(dmk_handle* h, int* reason_out) {
  for (int k = 0; k <= h->num.snes_max_it; k++) {
    int rc = form_residual(h, h->F, h->R, k);
    if (rc) return rc;
    int reason = converged(h);
    if (reason) { *reason_out = reason; return 0; }
    CKL(dmk_axpby(h->nx, 1.0, h->F, -1.0, h->R, h->F, h->st));
  }
  *reason_out = -1;
  return 0;
}

// ------------------------------------------------------------------ Galerkin
// PETSc SNES "newtonls" with "-ksp_type gmres -snes_linesearch_type bt"
// (grid_mech_spectral_Galerkin.f90:217-218) on the matrix-free Jacobian GK_op
// (:607-660), restated from the published methods with PETSc's defaults:
// Newton's method with the backtracking line search of Dennis & Schnabel,
// "Numerical Methods for Unconstrained Optimization and Nonlinear Equations"
// (1983), algorithm A6.3.1 (cubic order, alpha = 1e-4, minimum step 1e-12,
// maximum step 1e8, 40 backtracking steps, initial slope from the Jacobian
// applied to the step), one failed line search ending the solve; linear
// solves by restarted GMRES (Saad & Schultz, SIAM J. Sci. Stat. Comput. 7(3)
// 1986): restart 30, classical Gram-Schmidt (the new vector's norm from the
// same reduction as its projections), zero initial guess, Givens
// residual estimate, relative tolerance 1e-5, no preconditioner (a shell
// matrix gets none).  SNESSetUpdate's set_F_aim (:666-681) runs at the start
// of every Newton step.

static int sync_dots(dmk_handle* h, const DotJobs& J, double* out) {
  CKL(dmk_dots(&J, h->nx, h->partial, h->small + SM_ADOTS, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_ADOTS, h->small + SM_ADOTS, sizeof(double) * J.n, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  std::memcpy(out, h->hsmall + SM_ADOTS, sizeof(double) * J.n);
  return 0;
}

// GK_op: out = G * (K : v), DC term kept on the stress-controlled components
static int gk_jv(dmk_handle* h, const double* v, double* out) {
  CKL(dmk_gk_apply(h->N, h->dPdF, v, h->T, h->st));
  CKL(dmk_fft_r2c(&h->fft, 9, h->T, h->Z, h->st));
  GalerkinArgs ga{};
  for (int k = 0; k < 9; k++) ga.stress_mask[k] = h->stress_mask[k];
  CKL(dmk_galerkin(h->Z, &h->ga, &ga, h->partial, h->st));
  CKL(dmk_fft_c2r(&h->fft, 9, h->Z, out, h->st));
  return 0;
}

// Krylov basis vector i < 31 (the Galerkin scheme allocates Xh with at least 31 vectors)
static double* gk_basis(dmk_handle* h, int i) { return h->Xh + (size_t)i * h->n9; }

// J x = b, x overwritten; returns the iteration count in *its, *ok = converged
static int gk_gmres(dmk_handle* h, const double* b, double* x, int* its, bool* ok) {
  const int m = 30, maxit = 10000;
  const double rtol = 1.0e-5, atol = 1.0e-50;
  const long n = h->n9;
  DotJobs J;
  J.n = 0;
  add_dot(&J, b, b, 0);
  double d[DMK_MAX_DOTS];
  int rc;
  if ((rc = sync_dots(h, J, d))) return rc;
  const double bnorm = std::sqrt(d[0]);
  CK(hipMemsetAsync(x, 0, sizeof(double) * n, h->st));
  *its = 0;
  *ok = std::isfinite(bnorm);
  if (bnorm == 0.0 || !*ok) return 0;
  const double tol = std::fmax(rtol * bnorm, atol);
  std::vector<double> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m), y(m);
  int total = 0;
  bool first = true;
  for (;;) {
    double* v0 = gk_basis(h, 0);
    double beta;
    if (first) {
      beta = bnorm;
      CKL(dmk_axpby(n, 1.0 / beta, b, 0.0, b, v0, h->st));
    } else {
      if ((rc = gk_jv(h, x, v0))) return rc;
      CKL(dmk_axpby(n, 1.0, b, -1.0, v0, v0, h->st));
      J.n = 0;
      add_dot(&J, v0, v0, 0);
      if ((rc = sync_dots(h, J, d))) return rc;
      beta = std::sqrt(d[0]);
      if (beta <= tol) break;
      CKL(dmk_axpby(n, 1.0 / beta, v0, 0.0, v0, v0, h->st));
    }
    first = false;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    bool done = false;
    for (int j = 0; j < m; j++) {
      // classical Gram-Schmidt with the norm from the same reduction,
      // |w - V h|^2 = |w|^2 - |h|^2, so one synchronisation per step; the
      // norm is recomputed when that difference cancels by more than 100x
      double* w = gk_basis(h, j + 1);
      if ((rc = gk_jv(h, gk_basis(h, j), w))) return rc;
      J.n = 0;
      for (int i = 0; i <= j; i++) add_dot(&J, w, gk_basis(h, i), 0);
      add_dot(&J, w, w, 0);
      if ((rc = sync_dots(h, J, d))) return rc;
      double hh = 0.0;
      for (int i = 0; i <= j; i++) hh += d[i] * d[i];
      double hn2 = d[j + 1] - hh;
      CombineArgs ca;
      ca.l = j + 1;
      ca.atot = 0.0;
      for (int i = 0; i <= j; i++) {
        H[i + (m + 1) * j] = d[i];
        ca.X[i] = gk_basis(h, i);
      }
      double hn;
      if (hn2 > 0.01 * d[j + 1]) {
        hn = std::sqrt(hn2);
        for (int i = 0; i <= j; i++) ca.alpha[i] = -d[i] / hn;
        CKL(dmk_lincomb(n, 1.0 / hn, w, &ca, w, h->st));
      } else {
        for (int i = 0; i <= j; i++) ca.alpha[i] = -d[i];
        CKL(dmk_lincomb(n, 1.0, w, &ca, w, h->st));
        J.n = 0;
        add_dot(&J, w, w, 0);
        if ((rc = sync_dots(h, J, d))) return rc;
        hn = std::sqrt(d[0]);
        if (hn > 0.0) CKL(dmk_axpby(n, 1.0 / hn, w, 0.0, w, w, h->st));
      }
      if (!std::isfinite(hn)) { *ok = false; *its = total; return 0; }
      H[j + 1 + (m + 1) * j] = hn;
      for (int i = 0; i < j; i++) {
        double a = H[i + (m + 1) * j], c = H[i + 1 + (m + 1) * j];
        H[i + (m + 1) * j] = cs[i] * a + sn[i] * c;
        H[i + 1 + (m + 1) * j] = -sn[i] * a + cs[i] * c;
      }
      double a = H[j + (m + 1) * j], c = H[j + 1 + (m + 1) * j];
      double r = std::hypot(a, c);
      cs[j] = r > 0.0 ? a / r : 1.0;
      sn[j] = r > 0.0 ? c / r : 0.0;
      H[j + (m + 1) * j] = r;
      H[j + 1 + (m + 1) * j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      total++;
      k = j + 1;
      if (std::fabs(g[j + 1]) <= tol || hn == 0.0) { done = true; break; }
      if (total >= maxit) break;
    }
    for (int i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int q = i + 1; q < k; q++) s -= H[i + (m + 1) * q] * y[q];
      y[i] = s / H[i + (m + 1) * i];
    }
    CombineArgs ca;
    ca.l = k;
    ca.atot = 0.0;
    for (int i = 0; i < k; i++) { ca.alpha[i] = y[i]; ca.X[i] = gk_basis(h, i); }
    CKL(dmk_lincomb(n, 1.0, x, &ca, x, h->st));
    if (done) break;
    if (total >= maxit) { *ok = false; break; }
  }
  *its = total;
  return 0;
}

// trial point W = X - lambda Y and its residual G; g = |G|^2
static int gk_trial(dmk_handle* h, double lambda, const double* Y, double* W, double* G, int iter, double* g) {
  CKL(dmk_axpby(h->n9, 1.0, h->F, -lambda, Y, W, h->st));
  DotJobs J;
  J.n = 0;
  add_dot(&J, G, G, 0);
  int rc;
  if ((rc = residual_enqueue(h, W, G, iter)) || (rc = residual_finish(h, &J, g))) return rc;
  return 0;
}

// backtracking line search from X = h->F, F = h->R along -Y; on success the
// accepted pair becomes (h->F, h->R) and *fnorm its residual norm
static int gk_linesearch(dmk_handle* h, double* Y, int iter, double* fnorm, bool* ok) {
  const double alpha = 1.0e-4, minlambda = 1.0e-12, maxstep = 1.0e8;
  const int max_its = 40;
  double* W = pool_free(h->poolX, h->F, nullptr, 0, Y);
  double* G = pool_free(h->poolF, h->R, nullptr, 0, nullptr);
  double* JY = pool_free(h->poolF, h->R, nullptr, 0, G);
  if (!W || !G || !JY) { dmk_set_error("Galerkin buffer pool exhausted"); return DMK_ERR_CONFIG; }
  DotJobs J;
  J.n = 0;
  add_dot(&J, Y, Y, 0);
  double d[DMK_MAX_DOTS];
  int rc;
  if ((rc = sync_dots(h, J, d))) return rc;
  double ynorm = std::sqrt(d[0]);
  *ok = false;
  if (ynorm == 0.0) return 0;
  if (ynorm > maxstep) {
    CKL(dmk_axpby(h->n9, maxstep / ynorm, Y, 0.0, Y, Y, h->st));
    ynorm = maxstep;
  }
  const double f = (*fnorm) * (*fnorm);
  if ((rc = gk_jv(h, Y, JY))) return rc;
  J.n = 0;
  add_dot(&J, h->R, JY, 0);
  if ((rc = sync_dots(h, J, d))) return rc;
  double initslope = d[0];
  if (initslope > 0.0) initslope = -initslope;
  if (initslope == 0.0) initslope = -1.0;
  double lambda = 1.0, g;
  for (;;) {
    if ((rc = gk_trial(h, lambda, Y, W, G, iter, &g))) return rc;
    if (std::isfinite(g)) break;
    lambda *= 0.5;
    if (lambda <= minlambda) return 0;
  }
  bool accept = 0.5 * g <= 0.5 * f + lambda * alpha * initslope;
  if (!accept) {
    double lambdatemp = -initslope / (g - f - 2.0 * lambda * initslope);
    double lambdaprev = lambda, gprev = g;
    if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
    lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
    if ((rc = gk_trial(h, lambda, Y, W, G, iter, &g))) return rc;
    accept = 0.5 * g < 0.5 * f + lambda * alpha * initslope;
    for (int count = 0; !accept && count < max_its; count++) {
      if (lambda <= minlambda) return 0;
      double t1 = 0.5 * (g - f) - lambda * initslope;
      double t2 = 0.5 * (gprev - f) - lambdaprev * initslope;
      double a = (t1 / (lambda * lambda) - t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
      double b = (-lambdaprev * t1 / (lambda * lambda) + lambda * t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
      double dd = b * b - 3.0 * a * initslope;
      if (dd < 0.0) dd = 0.0;
      lambdatemp = a == 0.0 ? -initslope / (2.0 * b) : (-b + std::sqrt(dd)) / (3.0 * a);
      lambdaprev = lambda;
      gprev = g;
      if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
      lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
      if ((rc = gk_trial(h, lambda, Y, W, G, iter, &g))) return rc;
      accept = 0.5 * g < 0.5 * f + lambda * alpha * initslope;
    }
  }
  if (!accept) return 0;
  h->F = W;
  h->R = G;
  *fnorm = std::sqrt(g);
  *ok = true;
  return 0;
}

static int solve_newtonls(dmk_handle* h, int* reason_out) {
  int rc;
  DotJobs J;
  J.n = 0;
  add_dot(&J, h->R, h->R, 0);
  double d[DMK_MAX_DOTS];
  if ((rc = residual_enqueue(h, h->F, h->R, 0)) || (rc = residual_finish(h, &J, d))) return rc;
  double fnorm = std::sqrt(d[0]);
  int reason = converged(h);
  if (reason || fnorm != fnorm) { *reason_out = reason ? reason : -1; return 0; }
  for (int i = 0; i < h->num.snes_max_it; i++) {
    // set_F_aim (grid_mech_spectral_Galerkin.f90:666-681)
    for (int q = 0; q < 9; q++) {
      double s = 0.0;
      for (int p = 0; p < 9; p++) s += h->S[q + 9 * p] * (h->P_av[p] - h->P_aim[p]);
      h->F_aim[q] -= s;
    }
    double* Y = pool_free(h->poolX, h->F, nullptr, 0, nullptr);
    if (!Y) { dmk_set_error("Galerkin buffer pool exhausted"); return DMK_ERR_CONFIG; }
    int lits;
    bool kok, lok;
    if ((rc = gk_gmres(h, h->R, Y, &lits, &kok))) return rc;
    if (!kok) { *reason_out = -1; return 0; }
    if ((rc = gk_linesearch(h, Y, i, &fnorm, &lok))) return rc;
    if (!lok) { *reason_out = -1; return 0; }
    reason = converged(h);
    if (reason || fnorm != fnorm) { *reason_out = reason ? reason : -1; return 0; }
  }
  *reason_out = -1;
  return 0;
}

// ------------------------------------------------------------------ adjoint
// Reverse sweep over recorded increments.  This has no counterpart in the
// reference: it differentiates the converged spectral-basic increments with
// respect to the phenopowerlaw parameters.  Increment n's converged field F
// satisfies equilibrium (Gamma_0 * P = 0 off the mean, prescribed mean stress
// components, spectral_utilities.f90:436-501 and grid_mech_spectral_basic.f90
// :478-496) with P = P(F, state_{n-1}; theta).  Its adjoint is the same linear
// problem with the transposed tangent A^T of the increment map: find mu,
// compatible plus a mean on the stress-controlled components, with
// Phi(A^T mu) = Phi(g), where Phi is the Gamma_0 convolution whose mean is
// S : <.> (S the masked compliance of <A^T>, the basic scheme's own mean
// rule).  It is solved by restarted GMRES.  The local reverse steps
// (dmk_adjoint.h) give g, then carry the cotangents of F_p and xi to
// increment n-1 and into the parameters.
// Solver and material-point state at the end of an increment, enough to run
// the following increments again (device part in AdjState::ckpt)
struct HostState {
  double F_aimDot[9], F_aim[9], F_aim_lastinc[9], P_av[9], P_aim[9];
  double C_volAvg[81], C_volAvgLastInc[81], C_minMaxAvg[81], C_minMaxAvgLastInc[81], S[81];
  double C_minMaxAvgRestart[81];
  double err_div, err_BC, dt, q_R[4];
  int stress_mask[9];
  GammaArgs ga;
  double t, Delta_t, Delta_t_prev;
  int cutBack, cutBackLevel, totalIncs;
};

struct AdjState {
  int cap = 0, n = 0;
  size_t slot = 0;                 // doubles per tape entry
  double* tape = nullptr;          // entries F, F_p, L_p, xi (component-major): 0..cap, or one segment
  std::vector<double> dt, pav, qR;
  std::vector<int> mask, maxsub;
  // segmented tape (seg > 0): the tape holds the seg + 1 entries of one
  // segment of increments (entry k at k - base); a checkpoint of the full
  // state every seg increments lets the reverse sweep run each earlier
  // segment again before it is differentiated
  int seg = 0, base = 0;
  std::vector<double*> ckpt;       // device part of the checkpoints (allocated when first saved); the last
                                   // one (index cap / seg + 1) holds the end-of-run state during a sweep
  size_t ckpt_slot = 0;
  std::vector<HostState> ckpt_host;
  std::vector<IncStart> incs;      // driver state at the start of each recorded increment
  bool replaying = false;
  int replay_k = 0;                // the increment being run again
  int on_tape = 0;                 // segment whose entries the tape holds
  bool swept = false;              // a segmented reverse sweep ran: the tape no longer ends the run
  int M = 30;                      // GMRES restart length
  GammaArgs ga;                    // Gamma_0 of the adjoint solves (reference medium sym <A>)
  bool work_ready = false;
  double *A = nullptr, *g = nullptr, *res[2] = {nullptr, nullptr}, *scr[2] = {nullptr, nullptr},
         *Fpb[2] = {nullptr, nullptr}, *xib[2] = {nullptr, nullptr}, *thb = nullptr, *mu = nullptr, *y = nullptr,
         *b = nullptr, *r = nullptr, *V = nullptr, *red = nullptr, *mu_prev = nullptr;
  std::vector<void*> work;
  // the factorisation of increment k - 1 (tape only) runs on st2 while the
  // GMRES of increment k runs on the solver stream; scr / res alternate
  hipStream_t st2 = nullptr;
  hipEvent_t fac[2] = {nullptr, nullptr};
  // in a lockstep batch: the GMRES's vector steps waiting for its next round
  // request, whether the solver stream has had work since the last round, and
  // the last round's completion event the stream has yet to wait for
  DmkPreOp pre[DMK_PRE_OPS];
  int npre = 0;
  bool lk_busy = true;
  hipEvent_t lk_done = nullptr;
};

enum { TP_F = 0, TP_FP = 9, TP_LP = 18, TP_XI = 27 };   // tape entry offsets, in fields of N

static int run_increment(dmk_handle* h, const dmk_loadcase* lc, double t_0, int* guess, int* its_out, int* nf_out);

static double* tape_entry(const dmk_handle* h, int k) {
  return h->adj->tape + (size_t)(k - h->adj->base) * h->adj->slot;
}

// device part of a checkpoint: F, F_lastinc (nx each), the phase F, F_p,
// L_p, S and the accepted P (9 N each), xi and gamma (nsl_max N each) and the
// small reduction buffer
static size_t ckpt_doubles(const dmk_handle* h) {
  return 2 * (size_t)h->nx + 5 * (size_t)h->n9 + 2 * (size_t)h->nsl_max * h->N + SM_SIZE;
}

static int ckpt_copy(dmk_handle* h, double* c, bool save) {
  const size_t N = h->N, nx = h->nx, n9 = h->n9, ns = (size_t)h->nsl_max * N;
  double* dev[] = {h->F, h->F_lastinc, h->mp.Fphase, h->mp.Fp, h->mp.Lp, h->mp.S, h->P_hom, h->mp.xi, h->mp.gam,
                   h->small};
  const size_t n[] = {nx, nx, n9, n9, n9, n9, n9, ns, ns, (size_t)SM_SIZE};
  size_t off = 0;
  for (int q = 0; q < 10; q++) {
    if (save) CK(hipMemcpyAsync(c + off, dev[q], sizeof(double) * n[q], hipMemcpyDeviceToDevice, h->st));
    else CK(hipMemcpyAsync(dev[q], c + off, sizeof(double) * n[q], hipMemcpyDeviceToDevice, h->st));
    off += n[q];
  }
  return 0;
}

static int ckpt_save(dmk_handle* h, int i) {
  AdjState& a = *h->adj;
  if (!a.ckpt[i] && hipMalloc((void**)&a.ckpt[i], sizeof(double) * a.ckpt_slot) != hipSuccess) {
    hipGetLastError();
    a.ckpt[i] = nullptr;
    dmk_set_error("adjoint checkpoint: %zu bytes of device memory not available", sizeof(double) * a.ckpt_slot);
    return DMK_ERR_HIP;
  }
  int rc = ckpt_copy(h, a.ckpt[i], true);
  if (rc) return rc;
  HostState& q = a.ckpt_host[i];
  std::memcpy(q.F_aimDot, h->F_aimDot, sizeof q.F_aimDot);
  std::memcpy(q.F_aim, h->F_aim, sizeof q.F_aim);
  std::memcpy(q.F_aim_lastinc, h->F_aim_lastinc, sizeof q.F_aim_lastinc);
  std::memcpy(q.P_av, h->P_av, sizeof q.P_av);
  std::memcpy(q.P_aim, h->P_aim, sizeof q.P_aim);
  std::memcpy(q.C_volAvg, h->C_volAvg, sizeof q.C_volAvg);
  std::memcpy(q.C_volAvgLastInc, h->C_volAvgLastInc, sizeof q.C_volAvgLastInc);
  std::memcpy(q.C_minMaxAvg, h->C_minMaxAvg, sizeof q.C_minMaxAvg);
  std::memcpy(q.C_minMaxAvgLastInc, h->C_minMaxAvgLastInc, sizeof q.C_minMaxAvgLastInc);
  std::memcpy(q.S, h->S, sizeof q.S);
  std::memcpy(q.C_minMaxAvgRestart, h->C_minMaxAvgRestart, sizeof q.C_minMaxAvgRestart);
  q.err_div = h->err_div;
  q.err_BC = h->err_BC;
  q.dt = h->dt;
  std::memcpy(q.q_R, h->q_R, sizeof q.q_R);
  std::memcpy(q.stress_mask, h->stress_mask, sizeof q.stress_mask);
  q.ga = h->ga;
  q.t = h->t;
  q.Delta_t = h->Delta_t;
  q.Delta_t_prev = h->Delta_t_prev;
  q.cutBack = h->cutBack;
  q.cutBackLevel = h->cutBackLevel;
  q.totalIncs = h->totalIncs;
  return 0;
}

static int ckpt_restore(dmk_handle* h, int i) {
  AdjState& a = *h->adj;
  int rc = ckpt_copy(h, a.ckpt[i], false);
  if (rc) return rc;
  const HostState& q = a.ckpt_host[i];
  std::memcpy(h->F_aimDot, q.F_aimDot, sizeof q.F_aimDot);
  std::memcpy(h->F_aim, q.F_aim, sizeof q.F_aim);
  std::memcpy(h->F_aim_lastinc, q.F_aim_lastinc, sizeof q.F_aim_lastinc);
  std::memcpy(h->P_av, q.P_av, sizeof q.P_av);
  std::memcpy(h->P_aim, q.P_aim, sizeof q.P_aim);
  std::memcpy(h->C_volAvg, q.C_volAvg, sizeof q.C_volAvg);
  std::memcpy(h->C_volAvgLastInc, q.C_volAvgLastInc, sizeof q.C_volAvgLastInc);
  std::memcpy(h->C_minMaxAvg, q.C_minMaxAvg, sizeof q.C_minMaxAvg);
  std::memcpy(h->C_minMaxAvgLastInc, q.C_minMaxAvgLastInc, sizeof q.C_minMaxAvgLastInc);
  std::memcpy(h->S, q.S, sizeof q.S);
  std::memcpy(h->C_minMaxAvgRestart, q.C_minMaxAvgRestart, sizeof q.C_minMaxAvgRestart);
  h->err_div = q.err_div;
  h->err_BC = q.err_BC;
  h->dt = q.dt;
  std::memcpy(h->q_R, q.q_R, sizeof q.q_R);
  std::memcpy(h->stress_mask, q.stress_mask, sizeof q.stress_mask);
  h->ga = q.ga;
  h->t = q.t;
  h->Delta_t = q.Delta_t;
  h->Delta_t_prev = q.Delta_t_prev;
  h->cutBack = q.cutBack;
  h->cutBackLevel = q.cutBackLevel;
  h->totalIncs = q.totalIncs;
  return 0;
}

static void adj_free(dmk_handle* h) {
  if (!h->adj) return;
  if (h->st) hipStreamSynchronize(h->st);
  if (h->adj->st2) {
    hipStreamSynchronize(h->adj->st2);
    hipStreamDestroy(h->adj->st2);
  }
  for (hipEvent_t e : h->adj->fac)
    if (e) hipEventDestroy(e);
  if (h->adj->tape) hipFree(h->adj->tape);
  for (double* c : h->adj->ckpt)
    if (c) hipFree(c);
  for (void* p : h->adj->work) hipFree(p);
  delete h->adj;
  h->adj = nullptr;
}

// device bytes of the reverse-sweep work buffers (allocated by the first
// dmk_adjoint_backward of a tape, listed in adj_work_alloc)
size_t dmk_adjoint_work_bytes(const dmk_handle* h) {
  if (!h || !h->adj || h->adj->work_ready) return 0;
  const size_t N = h->N, n9 = h->n9, mmax = h->nsl_max, nslot = ADJ_TH_SYS * mmax + DMK_NH_MAX;
  const size_t d = 81 * N + 8 * n9 + 2 * N + 2 * dmk_adj_scratch_slots((int)mmax) * N + 2 * mmax * N + nslot * N +
                   (size_t)(h->adj->M + 1) * n9 + (size_t)h->n_phases * (nslot + mmax);
  return sizeof(double) * d;
}

// free the reverse-sweep work buffers and keep the tape: the next sweep
// allocates them again
void dmk_adjoint_work_release(dmk_handle* h) {
  if (!h || !h->adj) return;
  AdjState& a = *h->adj;
  if (h->st) hipStreamSynchronize(h->st);
  if (a.st2) {
    hipStreamSynchronize(a.st2);
    hipStreamDestroy(a.st2);
    a.st2 = nullptr;
  }
  for (hipEvent_t& e : a.fac)
    if (e) {
      hipEventDestroy(e);
      e = nullptr;
    }
  for (void* p : a.work) hipFree(p);
  a.work.clear();
  a.A = a.g = a.thb = a.mu = a.y = a.b = a.r = a.V = a.red = a.mu_prev = nullptr;
  for (int i = 0; i < 2; i++) a.res[i] = a.scr[i] = a.Fpb[i] = a.xib[i] = nullptr;
  a.work_ready = false;
}

static int adj_walloc(dmk_handle* h, double** p, size_t n) {
  void* v = nullptr;
  CK(hipMalloc(&v, sizeof(double) * (n ? n : 1)));
  h->adj->work.push_back(v);
  *p = (double*)v;
  return 0;
}

static int tape_store(dmk_handle* h, int k) {
  double* E = tape_entry(h, k);
  const size_t N = h->N, b9 = sizeof(double) * 9 * N;
  CK(hipMemcpyAsync(E + TP_F * N, h->mp.Fphase, b9, hipMemcpyDeviceToDevice, h->st));
  CK(hipMemcpyAsync(E + TP_FP * N, h->mp.Fp, b9, hipMemcpyDeviceToDevice, h->st));
  CK(hipMemcpyAsync(E + TP_LP * N, h->mp.Lp, b9, hipMemcpyDeviceToDevice, h->st));
  CK(hipMemcpyAsync(E + TP_XI * N, h->mp.xi, sizeof(double) * h->nsl_max * N, hipMemcpyDeviceToDevice, h->st));
  return 0;
}

extern "C" int dmk_adjoint_tape_segmented(dmk_handle* h, int32_t max_increments, int32_t segment) {
  if (!h || max_increments < 1 || segment < 0) { dmk_set_error("dmk_adjoint_tape: invalid argument"); return DMK_ERR_ARG; }
  if (h->scheme != 0) { dmk_set_error("dmk_adjoint_tape: the adjoint differentiates the spectral-basic scheme"); return DMK_ERR_CONFIG; }
  adj_free(h);
  h->adj = new AdjState;
  AdjState& a = *h->adj;
  a.cap = max_increments;
  a.seg = segment >= max_increments ? 0 : segment;
  a.slot = (size_t)(27 + h->nsl_max) * h->N;
  const size_t entries = a.seg ? (size_t)a.seg + 1 : (size_t)a.cap + 1;
  const int nck = a.seg ? a.cap / a.seg + 2 : 0;     // checkpoints 0..cap/seg and the final state
  a.ckpt_slot = a.seg ? ckpt_doubles(h) : 0;
  if (hipMalloc((void**)&a.tape, sizeof(double) * a.slot * entries) != hipSuccess) {
    hipGetLastError();
    a.tape = nullptr;
    adj_free(h);
    dmk_set_error("dmk_adjoint_tape: %zu bytes of device memory not available", sizeof(double) * a.slot * entries);
    return DMK_ERR_HIP;
  }
  a.ckpt.assign(nck, nullptr);
  a.ckpt_host.resize(nck);
  int rc = tape_store(h, 0);
  if (rc) return rc;
  return a.seg ? ckpt_save(h, 0) : 0;
}

extern "C" int dmk_adjoint_tape(dmk_handle* h, int32_t max_increments) {
  return dmk_adjoint_tape_segmented(h, max_increments, 0);
}

extern "C" int dmk_adjoint_record(dmk_handle* h) {
  if (!h || !h->adj) { dmk_set_error("dmk_adjoint_record: no tape"); return DMK_ERR_ARG; }
  AdjState& a = *h->adj;
  int rc;
  if (a.replaying) {
    // an increment of a segment run again from its checkpoint: the same state
    // as in the forward run, bit for bit
    const int k = ++a.replay_k;
    if ((rc = tape_store(h, k))) return rc;
    CK(hipStreamSynchronize(h->st));
    if (std::memcmp(h->P_av, &a.pav[9 * (size_t)(k - 1)], sizeof h->P_av) != 0) {
      dmk_set_error("dmk_adjoint_backward: increment %d run again from its checkpoint differs from the forward run", k);
      return DMK_ERR_NOCONV;
    }
    return 0;
  }
  if (a.n >= a.cap) { dmk_set_error("dmk_adjoint_record: tape full (%d increments)", a.cap); return DMK_ERR_ARG; }
  if (a.swept) {
    dmk_set_error("dmk_adjoint_record: the segmented tape was swept; start a new tape to record further increments");
    return DMK_ERR_ARG;
  }
  const int k = a.n + 1;
  if (a.seg) {
    if (!h->inc_start.valid) {
      dmk_set_error("dmk_adjoint_record: a segmented tape records increments of dmk_grid_run_loadcase only");
      return DMK_ERR_ARG;
    }
    if (k > 1 && (k - 1) % a.seg == 0) {
      // a new segment starts at k - 1: its entry moves to the segment's first slot
      const double* last = tape_entry(h, k - 1);
      a.base = k - 1;
      a.on_tape = (k - 1) / a.seg;
      CK(hipMemcpyAsync(tape_entry(h, k - 1), last, sizeof(double) * a.slot, hipMemcpyDeviceToDevice, h->st));
    }
    a.incs.push_back(h->inc_start);
  }
  if ((rc = tape_store(h, k))) return rc;
  if (a.seg && k % a.seg == 0 && (rc = ckpt_save(h, k / a.seg))) return rc;
  CKL(dmk_adj_imax(h->mp.n_sub, h->N, h->small + SM_AMAX, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_AMAX, h->small + SM_AMAX, sizeof(double), hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  a.n++;
  a.dt.push_back(h->dt);
  a.pav.insert(a.pav.end(), h->P_av, h->P_av + 9);
  a.qR.insert(a.qR.end(), h->q_R, h->q_R + 4);
  a.mask.insert(a.mask.end(), h->stress_mask, h->stress_mask + 9);
  a.maxsub.push_back((int)h->hsmall[SM_AMAX]);
  return 0;
}

extern "C" int dmk_adjoint_info(dmk_handle* h, int32_t* n, double* P_av, double* dt) {
  if (!h || !h->adj) { dmk_set_error("dmk_adjoint_info: no tape"); return DMK_ERR_ARG; }
  const AdjState& a = *h->adj;
  if (n) *n = a.n;
  if (P_av) std::memcpy(P_av, a.pav.data(), sizeof(double) * 9 * a.n);
  if (dt) std::memcpy(dt, a.dt.data(), sizeof(double) * a.n);
  return 0;
}

extern "C" int dmk_adjoint_release(dmk_handle* h) {
  if (!h) { dmk_set_error("null handle"); return DMK_ERR_ARG; }
  adj_free(h);
  return 0;
}

static bool in_lockstep(const dmk_handle* h) { return h->batch && dmk_stress_batch_resid(h->batch); }

// the dot products of J over the nx-long fields into the report region
// SM_ADOTS: as a request of the batch's round (kind DMK_RQ_ADJ_OP or
// DMK_RQ_DOTS, q filled by the caller) or enqueued on the handle's stream
static void adj_dot_request(dmk_handle* h, const DotJobs* J, DmkResidDev& q) {
  const int nd = J ? J->n : 0;
  if (J) q.J = *J;
  q.J.n = nd;
  q.partial = h->partial;
  q.nx = h->nx;
  q.nc_dots = dmk_nchunk(h->nx);
  q.S.k = 0;
  for (int j = 0; j < nd; j++, q.S.k++) {
    q.S.p[q.S.k] = h->partial + (long)j * q.nc_dots;
    q.S.n[q.S.k] = q.nc_dots;
    q.S.scale[q.S.k] = 1.0;
    q.S.out[q.S.k] = h->small + SM_ADOTS + j;
  }
  q.rep = h->small + SM_ADOTS;
  q.rep_n = nd;
}

// one request of the handle's lockstep round, carrying the pending vector
// steps; defer: the solver stream's wait for the round is left to
// adj_lk_sync (within the GMRES nothing else runs on that stream)
static int adj_submit(dmk_handle* h, DmkResidDev& q, double* d, bool defer) {
  AdjState& a = *h->adj;
  q.npre = a.npre;
  for (int k = 0; k < a.npre; k++) q.pre[k] = a.pre[k];
  q.idle = a.lk_busy ? 0 : 1;
  hipEvent_t ev = nullptr;
  int rc = dmk_resid_batch_submit(h->batch, h->batch_slot, nullptr, nullptr, 0, &q, &h->fft, h->st, d,
                                  defer ? &ev : nullptr);
  if (rc) return rc;
  a.npre = 0;
  if (defer) {
    a.lk_done = ev;
    a.lk_busy = false;
  } else {
    a.lk_done = nullptr;
    a.lk_busy = true;
  }
  return 0;
}

// the pending vector steps as a request of their own
static int adj_flush_pre(dmk_handle* h) {
  if (h->adj->npre == 0) return 0;
  DmkResidDev q{};
  q.kind = DMK_RQ_DOTS;
  adj_dot_request(h, nullptr, q);
  return adj_submit(h, q, nullptr, true);
}

// before other work on the solver stream: the vector steps still pending,
// and the stream's wait for the last round
static int adj_lk_sync(dmk_handle* h) {
  if (!in_lockstep(h)) return 0;
  AdjState& a = *h->adj;
  int rc = adj_flush_pre(h);
  if (rc) return rc;
  if (a.lk_done) CK(hipStreamWaitEvent(h->st, a.lk_done, 0));
  a.lk_done = nullptr;
  a.lk_busy = true;
  return 0;
}

// w -= sum_k c_k V_k (dmk_adj_gs), or the same step queued for the next round
static int adj_vgs(dmk_handle* h, long n, double* w, const double* V, long ldV, const GsArgs& c) {
  if (!in_lockstep(h)) {
    CKL(dmk_adj_gs(n, w, V, ldV, &c, h->st));
    return 0;
  }
  AdjState& a = *h->adj;
  if (c.l > DMK_PRE_GS_MAX) { dmk_set_error("adjoint GMRES: basis longer than a round's step takes"); return DMK_ERR_CONFIG; }
  int rc;
  if (a.npre == DMK_PRE_OPS && (rc = adj_flush_pre(h))) return rc;
  DmkPreOp& o = a.pre[a.npre++];
  o = DmkPreOp{};
  o.type = DMK_PRE_GS;
  o.n = n;
  o.w = w;
  o.V = V;
  o.ldV = ldV;
  o.l = c.l;
  for (int k = 0; k < c.l; k++) o.c[k] = c.c[k];
  return 0;
}

// z = a x + b y (dmk_axpby), or the same step queued for the next round
static int adj_vaxpby(dmk_handle* h, long n, double av, const double* x, double bv, const double* y, double* z) {
  if (!in_lockstep(h)) {
    CKL(dmk_axpby(n, av, x, bv, y, z, h->st));
    return 0;
  }
  AdjState& a = *h->adj;
  int rc;
  if (a.npre == DMK_PRE_OPS && (rc = adj_flush_pre(h))) return rc;
  DmkPreOp& o = a.pre[a.npre++];
  o = DmkPreOp{};
  o.type = DMK_PRE_AXPBY;
  o.n = n;
  o.a = av;
  o.x = x;
  o.b = bv;
  o.y = y;
  o.z = z;
  return 0;
}

static int adj_dots(dmk_handle* h, const DotJobs& J, double* out) {
  if (J.n > DMK_MAX_DOTS) { dmk_set_error("too many dot products"); return DMK_ERR_ARG; }
  if (in_lockstep(h)) {
    DmkResidDev q{};
    q.kind = DMK_RQ_DOTS;
    adj_dot_request(h, &J, q);
    return adj_submit(h, q, out, true);
  }
  CKL(dmk_dots_sums(&J, h->nx, h->partial, h->small + SM_ADOTS, nullptr, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_ADOTS, h->small + SM_ADOTS, sizeof(double) * J.n, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  std::memcpy(out, h->hsmall + SM_ADOTS, sizeof(double) * J.n);
  return 0;
}

// out = Phi(A^T v) (v given) or Phi(y): the Gamma_0 convolution with the
// mean rule of the increment; then the vector step post (optional) and the
// dot products J (optional) into d.  In
// a lockstep batch all of it is one request of the round; otherwise the same
// kernels on the handle's stream.
static int adj_apply(dmk_handle* h, const AimArgs& aa, const double* v, const double* y, double* out,
                     const DotJobs* J, double* d, bool defer = false, const DmkPreOp* post = nullptr) {
  if (v) y = h->adj->y;
  GammaArgs g = h->adj->ga;
  g.do_div = 0;
  g.do_gamma = 1;
  g.aim_dev = h->small + SM_AAIM;
  if (in_lockstep(h)) {
    DmkResidDev q{};
    q.kind = DMK_RQ_ADJ_OP;
    q.A = v ? h->adj->A : nullptr;
    q.v = v;
    q.Pnew = const_cast<double*>(y);
    q.N = h->N;
    q.wgt = h->wgt;
    q.nc_resp = dmk_nchunk(h->N);
    q.flag = h->flag;
    q.a = aa;
    q.apav = h->small + SM_APAV;
    q.afaim = h->small + SM_AFAIM;
    q.arep = h->small + SM_AREP;
    q.aaim = h->small + SM_AAIM;
    q.Z = h->Z;
    q.r = out;
    q.divp = h->divp;
    q.g = g;
    if (post) {
      q.npost = 1;
      q.post = *post;
    }
    adj_dot_request(h, J, q);
    return adj_submit(h, q, d, defer);
  }
  if (v) CKL(dmk_adj_atmul(h->N, h->adj->A, v, h->adj->y, h->st));
  CKL(dmk_colsum(y, h->N, h->N, 9, h->partial, h->small + SM_APAV, h->wgt, h->st));
  CKL(dmk_residual_aim(&aa, h->small + SM_APAV, h->small + SM_AFAIM, h->flag, h->small + SM_AREP,
                       h->small + SM_AAIM, h->st));
  CKL(dmk_fft_r2c_xy(&h->fft, 9, y, h->Z, h->st));
  int nb = 0;
  CKL(dmk_gamma_conv(&h->fft, h->Z, &g, h->divp, &nb, h->st));
  CKL(dmk_fft_c2r_xy(&h->fft, 9, h->Z, out, h->st));
  if (post) CKL(dmk_axpby(post->n, post->a, post->x, post->b, post->y, post->z, h->st));
  return J && J->n ? adj_dots(h, *J, d) : 0;
}

// z = a x + b y as the vector step that follows an operator application
static DmkPreOp axpby_step(long n, double av, const double* x, double bv, const double* y, double* z) {
  DmkPreOp o{};
  o.type = DMK_PRE_AXPBY;
  o.n = n;
  o.a = av;
  o.x = x;
  o.b = bv;
  o.y = y;
  o.z = z;
  return o;
}

// Restarted GMRES for Phi(A^T mu) = b, started from the mu in place (the
// caller's extrapolation of the previous increments' adjoint fields, which
// differ little from one increment to the next).  Arnoldi by classical
// Gram-Schmidt with the norm from the same reduction (|w - V h|^2 = |w|^2 -
// |h|^2), so one host synchronisation per step; a second pass when that
// cancels by more than a factor 100.
static int adj_gmres(dmk_handle* h, const AimArgs& aa, double rtol, int maxit, int* its_out, double* rel_out) {
  AdjState& a = *h->adj;
  const long n = h->n9;
  const int M = a.M;
  DotJobs J;
  double d[DMK_MAX_DOTS];
  int rc;
  // r = b - Phi(A^T mu), |b|, |r|.  In a lockstep batch every step below is
  // a round request or a vector step carried by the next one; the solver
  // stream waits for the rounds at the end (adj_lk_sync).
  const DmkPreOp resid = axpby_step(n, 1.0, a.b, -1.0, a.r, a.r);
  J.n = 0;
  add_dot(&J, a.b, a.b, 0);
  add_dot(&J, a.r, a.r, 0);
  if ((rc = adj_apply(h, aa, a.mu, nullptr, a.r, &J, d, true, &resid))) return rc;
  const double bnorm = std::sqrt(d[0]);
  double rnorm = std::sqrt(d[1]);
  *its_out = 0;
  *rel_out = 0.0;
  if (bnorm != bnorm || rnorm != rnorm) { dmk_set_error("adjoint: non-finite right-hand side"); return DMK_ERR_NOCONV; }
  if (bnorm == 0.0) {
    if ((rc = adj_lk_sync(h))) return rc;
    CK(hipMemsetAsync(a.mu, 0, sizeof(double) * n, h->st));
    return 0;
  }
  int its = 0;
  std::vector<double> H((size_t)(M + 1) * M), cs(M), sn(M), gv(M + 1), yv(M);
  while (rnorm > rtol * bnorm && its < maxit) {
    if ((rc = adj_vaxpby(h, n, 1.0 / rnorm, a.r, 0.0, a.r, a.V))) return rc;
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = rnorm;
    int j = 0;
    for (; j < M && its < maxit; j++) {
      its++;
      double* w = a.V + (size_t)(j + 1) * n;
      double hc[DMK_GS_MAX] = {0}, hn = 0.0;
      for (int pass = 0; pass < 2; pass++) {
        J.n = 0;
        for (int i = 0; i <= j; i++) add_dot(&J, a.V + (size_t)i * n, w, 0);
        add_dot(&J, w, w, 0);
        // w = Phi(A^T V_j) and its products with the basis (second pass: again after the update)
        if (pass == 0) rc = adj_apply(h, aa, a.V + (size_t)j * n, nullptr, w, &J, d, true);
        else rc = adj_dots(h, J, d);
        if (rc) return rc;
        GsArgs c;
        c.l = j + 1;
        double hh = 0.0;
        for (int i = 0; i <= j; i++) { c.c[i] = d[i]; hc[i] += d[i]; hh += d[i] * d[i]; }
        if ((rc = adj_vgs(h, n, w, a.V, n, c))) return rc;
        double w2 = d[j + 1], r2 = w2 - hh;
        hn = std::sqrt(std::fmax(r2, 0.0));
        if (r2 > 1e-4 * w2) break;
      }
      for (int i = 0; i <= j; i++) H[i + (size_t)(M + 1) * j] = hc[i];
      H[j + 1 + (size_t)(M + 1) * j] = hn;
      if (hn > 0.0 && (rc = adj_vaxpby(h, n, 1.0 / hn, w, 0.0, w, w))) return rc;
      for (int i = 0; i < j; i++) {
        double* c0 = &H[i + (size_t)(M + 1) * j];
        double t = cs[i] * c0[0] + sn[i] * c0[1];
        c0[1] = -sn[i] * c0[0] + cs[i] * c0[1];
        c0[0] = t;
      }
      double* cj = &H[j + (size_t)(M + 1) * j];
      double den = std::hypot(cj[0], cj[1]);
      cs[j] = den > 0.0 ? cj[0] / den : 1.0;
      sn[j] = den > 0.0 ? cj[1] / den : 0.0;
      cj[0] = den;
      cj[1] = 0.0;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      double est = std::fabs(gv[j + 1]);
      if (est != est) { dmk_set_error("adjoint GMRES: non-finite residual"); return DMK_ERR_NOCONV; }
      if (est <= 0.5 * rtol * bnorm || hn == 0.0) { j++; break; }
    }
    for (int i = j - 1; i >= 0; i--) {
      double s = gv[i];
      for (int k = i + 1; k < j; k++) s -= H[i + (size_t)(M + 1) * k] * yv[k];
      yv[i] = s / H[i + (size_t)(M + 1) * i];
    }
    GsArgs c;
    c.l = j;
    for (int i = 0; i < j; i++) c.c[i] = -yv[i];
    if ((rc = adj_vgs(h, n, a.mu, a.V, n, c))) return rc;
    // true residual (restart vector and the reported value)
    J.n = 0;
    add_dot(&J, a.r, a.r, 0);
    if ((rc = adj_apply(h, aa, a.mu, nullptr, a.r, &J, d, true, &resid))) return rc;
    rnorm = std::sqrt(d[0]);
  }
  if ((rc = adj_lk_sync(h))) return rc;
  *its_out = its;
  *rel_out = rnorm / bnorm;
  return 0;
}

extern "C" int dmk_adjoint_backward(dmk_handle* h, const double* P_av_bar, double rtol, double* theta_bar,
                                    double* stats) {
  if (!h || !h->adj || !P_av_bar || !theta_bar) { dmk_set_error("dmk_adjoint_backward: invalid argument"); return DMK_ERR_ARG; }
  AdjState& a = *h->adj;
  if (a.n < 1) { dmk_set_error("dmk_adjoint_backward: no recorded increment"); return DMK_ERR_ARG; }
  for (int k = 0; k < a.n; k++)
    if (a.maxsub[k] > 1) {
      dmk_set_error("dmk_adjoint_backward: increment %d was integrated in material-point sub-steps; "
                    "the recorded state does not define the single-step increment map", k + 1);
      return DMK_ERR_CONFIG;
    }
  const size_t N = h->N, n9 = h->n9;
  const int mmax = h->nsl_max, nslot = ADJ_TH_SYS * mmax + DMK_NH_MAX;
  int rc;
  if (!a.work_ready) {
    // sizes as counted by dmk_adjoint_work_bytes
    if ((rc = adj_walloc(h, &a.A, 81 * N)) || (rc = adj_walloc(h, &a.g, n9)) || (rc = adj_walloc(h, &a.res[0], N)) ||
        (rc = adj_walloc(h, &a.res[1], N)) || (rc = adj_walloc(h, &a.scr[0], dmk_adj_scratch_slots(mmax) * N)) ||
        (rc = adj_walloc(h, &a.scr[1], dmk_adj_scratch_slots(mmax) * N)) || (rc = adj_walloc(h, &a.Fpb[0], n9)) ||
        (rc = adj_walloc(h, &a.Fpb[1], n9)) || (rc = adj_walloc(h, &a.xib[0], (size_t)mmax * N)) ||
        (rc = adj_walloc(h, &a.xib[1], (size_t)mmax * N)) || (rc = adj_walloc(h, &a.thb, (size_t)nslot * N)) ||
        (rc = adj_walloc(h, &a.mu, n9)) || (rc = adj_walloc(h, &a.mu_prev, n9)) || (rc = adj_walloc(h, &a.y, n9)) ||
        (rc = adj_walloc(h, &a.b, n9)) ||
        (rc = adj_walloc(h, &a.r, n9)) || (rc = adj_walloc(h, &a.V, (size_t)(a.M + 1) * n9)) ||
        (rc = adj_walloc(h, &a.red, (size_t)h->n_phases * (nslot + mmax)))) {
      hipGetLastError();
      dmk_adjoint_work_release(h);
      dmk_set_error("dmk_adjoint_backward: %zu bytes of device memory for the reverse sweep not available",
                    dmk_adjoint_work_bytes(h));
      return rc;
    }
    CK(hipStreamCreate(&a.st2));
    CK(hipEventCreateWithFlags(&a.fac[0], hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&a.fac[1], hipEventDisableTiming));
    a.work_ready = true;
  }
  CK(hipMemsetAsync(a.Fpb[0], 0, sizeof(double) * n9, h->st));
  CK(hipMemsetAsync(a.xib[0], 0, sizeof(double) * mmax * N, h->st));
  CK(hipMemsetAsync(a.thb, 0, sizeof(double) * nslot * N, h->st));
  CK(hipMemsetAsync(a.red, 0, sizeof(double) * h->n_phases * (nslot + mmax), h->st));
  CK(hipMemsetAsync(a.mu, 0, sizeof(double) * n9, h->st));
  // the tape-only arguments of increment k, scratch buffer k % 2
  auto local_args = [&](int k) {
    const double* E = tape_entry(h, k);
    const double* Ep = tape_entry(h, k - 1);
    AdjKArgs ka{};
    ka.N = (int)N;
    ka.mmax = mmax;
    ka.dt = a.dt[k - 1];
    ka.eps_rel_Lp = h->nm.eps_rel_Lp;
    ka.eps_abs_Lp = h->nm.eps_abs_Lp;
    ka.eps_rel_state = h->nm.eps_rel_state;
    ka.phase = h->mp.phase;
    ka.F = E + TP_F * N;
    ka.Lp = E + TP_LP * N;
    ka.xi = E + TP_XI * N;
    ka.Fp0 = Ep + TP_FP * N;
    ka.xi0 = Ep + TP_XI * N;
    ka.scr = a.scr[k % 2];
    ka.res = a.res[k % 2];
    return ka;
  };
  // K1 + K2 of increment k on st2; its buffers were last read by increment
  // k + 2, whose work the solver stream finished before the enqueue
  auto factor = [&](int k) -> int {
    AdjKArgs ka = local_args(k);
    for (int p = 0; p < h->n_phases; p++) {
      ka.phase_id = p;
      ka.nsl = h->ph[p].nsl;
      CKL(dmk_adj_factor(&ka, h->ph_dev + p, a.st2));
    }
    CK(hipEventRecord(a.fac[k % 2], a.st2));
    return 0;
  };
  int cur = 0;
  int nsolved = 0;     // increments solved so far, across the segments
  // the reverse sweep over increments k1 .. k0 + 1 of the tape in memory
  auto sweep = [&](int k1, int k0) -> int {
    CK(hipStreamSynchronize(h->st));
    if ((rc = factor(k1))) return rc;
    for (int k = k1; k > k0; k--) {
      const double* qR = &a.qR[4 * (k - 1)];
      AdjKArgs ka = local_args(k);
      // P_av = R <P> R^T (passive rotation of the raw average): <P>_bar = R^T P_av_bar R
      double pr[9];
      rotate33(qR, P_av_bar + 9 * (k - 1), true, pr);
      for (int c = 0; c < 9; c++) ka.Pu[c] = h->wgt * pr[c];
      ka.Fpb_in = a.Fpb[cur];
      ka.xib_in = a.xib[cur];
      ka.A = a.A;
      ka.g = a.g;
      ka.mode = 0;
      CK(hipStreamWaitEvent(h->st, a.fac[k % 2], 0));
      for (int p = 0; p < h->n_phases; p++) {
        ka.phase_id = p;
        ka.nsl = h->ph[p].nsl;
        CKL(dmk_adj_vjp(&ka, h->ph_dev + p, h->st));
      }
      CKL(dmk_adj_max(ka.res, N, h->small + SM_AMAX, h->st));
      CKL(dmk_colsum(a.A, N, N, 81, h->partial, h->small + SM_ACAV, h->wgt, h->st));
      CK(hipMemcpyAsync(h->hsmall + SM_AMAX, h->small + SM_AMAX, sizeof(double) * (SM_ACAV + 81 - SM_AMAX),
                        hipMemcpyDeviceToHost, h->st));
      CK(hipStreamSynchronize(h->st));
      const double resmax = h->hsmall[SM_AMAX];
      if (!(resmax < HUGE_VAL)) {
        dmk_set_error("dmk_adjoint_backward: increment %d: singular increment Jacobian or non-finite state", k);
        return DMK_ERR_NOCONV;
      }
      if (k - 1 > k0 && (rc = factor(k - 1))) return rc;
      // mean rule: masked compliance of <A^T>
      double Ct[81], Sadj[81];
      const double* Abar = h->hsmall + SM_ACAV;
      for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) Ct[p + 9 * q] = Abar[q + 9 * p];
      if ((rc = masked_compliance_of(Ct, &a.mask[9 * (k - 1)], qR, Sadj))) return rc;
      // preconditioner: Gamma_0 of the reference medium sym <A> of this increment
      // (any positive-definite medium gives the same solution)
      double Cs[81];
      for (int p = 0; p < 81; p++) Cs[p] = 0.5 * (Abar[p] + Ct[p]);
      a.ga = h->ga;
      gamma_coef(h->wgt, Cs, &a.ga);
      AimArgs aa;
      std::memcpy(aa.S, Sadj, sizeof aa.S);
      for (int c = 0; c < 9; c++) { aa.P_aim[c] = 0.0; aa.stress_mask[c] = a.mask[9 * (k - 1) + c]; }
      std::memcpy(aa.q_R, qR, sizeof aa.q_R);
      aa.rotate = 1;
      aa.update = 1;
      // initial guess of the GMRES: the linear extrapolation 2 mu_(k+1) -
      // mu_(k+2) of the two increments solved before (after one, its mu; the
      // quadratic extrapolation measured more GMRES steps than the linear one)
      if (nsolved >= 2) {
        CKL(dmk_axpby((long)n9, 2.0, a.mu, -1.0, a.mu_prev, a.mu_prev, h->st));
        std::swap(a.mu, a.mu_prev);
      } else if (nsolved == 1) {
        CKL(dmk_axpby((long)n9, 1.0, a.mu, 0.0, a.mu, a.mu_prev, h->st));
      }
      CK(hipMemsetAsync(h->small + SM_AFAIM, 0, sizeof(double) * 9, h->st));
      if ((rc = adj_apply(h, aa, nullptr, a.g, a.b, nullptr, nullptr, true))) return rc;
      int its = 0;
      double rel = 0.0;
      if ((rc = adj_gmres(h, aa, rtol, 600, &its, &rel))) return rc;
      nsolved++;
      ka.mode = 1;
      ka.mu = a.mu;
      ka.Fpb_out = a.Fpb[1 - cur];
      ka.xib_out = a.xib[1 - cur];
      ka.thb = a.thb;
      for (int p = 0; p < h->n_phases; p++) {
        ka.phase_id = p;
        ka.nsl = h->ph[p].nsl;
        CKL(dmk_adj_vjp(&ka, h->ph_dev + p, h->st));
      }
      // parameter cotangents of this increment, summed per phase on the device
      CKL(dmk_adj_phase_sum(a.thb, N, h->mp.phase, h->n_phases, nslot, 1, a.red + (size_t)h->n_phases * mmax, h->st));
      cur = 1 - cur;
      if (stats) {
        stats[4 * (k - 1) + 0] = its;
        stats[4 * (k - 1) + 1] = rel;
        stats[4 * (k - 1) + 2] = resmax;
        stats[4 * (k - 1) + 3] = a.maxsub[k - 1];
      }
    }
    return 0;
  };
  if (!a.seg) {
    if ((rc = sweep(a.n, 0))) return rc;
  } else {
    // segments from the last: the last one is still on the tape; every earlier
    // one is run again from its checkpoint, recording its entries, before it is
    // differentiated; the end-of-run state is put back afterwards
    const int nseg = (a.n + a.seg - 1) / a.seg, fin = a.cap / a.seg + 1;
    if ((rc = ckpt_save(h, fin))) return rc;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
      const int k0 = sgi * a.seg, k1 = std::min(a.n, k0 + a.seg);
      if (sgi != a.on_tape) {
        if ((rc = ckpt_restore(h, sgi))) return rc;
        a.base = k0;
        if ((rc = tape_store(h, k0))) return rc;
        a.replaying = true;
        a.replay_k = k0;
        for (int k = k0 + 1; k <= k1 && !rc; k++) {
          const IncStart& is = a.incs[k - 1];
          h->t = is.t;
          h->Delta_t = is.Delta_t;
          h->Delta_t_prev = is.Delta_t_prev;
          h->cutBack = is.cutBack;
          h->cutBackLevel = is.cutBackLevel;
          int guess = is.guess, its = 0, nf = 0;
          const dmk_loadcase lc = is.lc;
          rc = run_increment(h, &lc, is.t_0, &guess, &its, &nf);
          if (!rc && a.replay_k != k) {
            dmk_set_error("dmk_adjoint_backward: increment %d did not converge when run again", k);
            rc = DMK_ERR_NOCONV;
          }
        }
        a.replaying = false;
        a.on_tape = sgi;
        if (rc) return rc;
      }
      if ((rc = sweep(k1, k0))) return rc;
    }
    if ((rc = ckpt_restore(h, fin))) return rc;
    a.swept = true;
  }
  // xi_0 enters as the initial state (phenopowerlaw.f90:259-262)
  CKL(dmk_adj_phase_sum(a.xib[cur], N, h->mp.phase, h->n_phases, mmax, 0, a.red, h->st));
  std::vector<double> hr((size_t)h->n_phases * (nslot + mmax));
  CK(hipMemcpyAsync(hr.data(), a.red, sizeof(double) * hr.size(), hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  const double* hx = hr.data();
  const double* ht = hr.data() + (size_t)h->n_phases * mmax;
  for (int p = 0; p < h->n_phases; p++) {
    double* o = theta_bar + (size_t)p * DMK_ADJ_NOUT;
    for (int k = 0; k < DMK_ADJ_NOUT; k++) o[k] = 0.0;
    const int m = h->ph[p].nsl;
    for (int s = 0; s < m; s++) o[s] = hx[(size_t)p * mmax + s];
    for (int slot = 0; slot < ADJ_TH_SYS; slot++)
      for (int s = 0; s < m; s++) o[(1 + slot) * DMK_NSL_MAX + s] = ht[(size_t)p * nslot + slot * mmax + s];
    for (int t = 0; t < DMK_NH_MAX; t++)
      o[(ADJ_TH_SYS + 1) * DMK_NSL_MAX + t] = ht[(size_t)p * nslot + ADJ_TH_SYS * mmax + t];
  }
  return 0;
}

// ------------------------------------------------------------------ C-ABI
extern "C" int dmk_device_synchronize(void) {
  CK(hipDeviceSynchronize());
  return 0;
}

extern "C" int dmk_nsl(const dmk_handle* h) { return h ? h->nsl_max : 0; }

extern "C" void dmk_mech_destroy(dmk_handle* h) {
  if (!h) return;
  if (h->st) hipStreamSynchronize(h->st);
  if (h->ev_ready)
    for (int k = 0; k < 8; k++) { hipEventDestroy(h->ev[k][0]); hipEventDestroy(h->ev[k][1]); }
  adj_free(h);
  for (void* p : h->allocs) hipFree(p);
  if (h->ph_dev) hipFree(h->ph_dev);
  if (h->hsmall) hipHostFree(h->hsmall);
  dmk_fft_free(&h->fft);
  if (h->st) hipStreamDestroy(h->st);
  delete h;
}

static int init_impl(dmk_handle* h, const int32_t cells[3], const double size[3], int32_t n_phases,
                     const dmk_phase_desc* phases, const int32_t* phase_of_cell, const double* q0,
                     const dmk_numerics* num) {
  for (int d = 0; d < 3; d++) {
    if (cells[d] < 1 || size[d] <= 0.0) { dmk_set_error("invalid grid"); return DMK_ERR_ARG; }
    h->cells[d] = cells[d];
    h->size[d] = size[d];
  }
  if (!num) { dmk_set_error("numerics missing"); return DMK_ERR_ARG; }
  h->num = *num;
  if (h->num.ngmres_m < 1 || h->num.ngmres_m > 30) { dmk_set_error("ngmres_m must be in [1,30]"); return DMK_ERR_CONFIG; }
  if (h->num.N_iter_max < 1 || h->num.N_iter_min < 1 || h->num.N_iter_min > h->num.N_iter_max) {
    dmk_set_error("IO_error 301: N_iter_min / N_iter_max");
    return DMK_ERR_CONFIG;
  }
  if (h->num.scheme < 0 || h->num.scheme > 2) { dmk_set_error("IO_error 601: not a valid grid solver type"); return DMK_ERR_CONFIG; }
  h->scheme = h->num.scheme;
  if (h->scheme == 1) {
    // grid_mech_spectral_polarization.f90:141-147; beta = 0, which the
    // reference accepts, divides by zero in its residual (:601), so it is refused
    const dmk_numerics& n = h->num;
    if (!(n.eps_abs_curl > 0.0) || !(n.eps_rel_curl > 0.0) || !(n.alpha > 0.0 && n.alpha <= 2.0) ||
        !(n.beta > 0.0 && n.beta <= 2.0)) {
      dmk_set_error("IO_error 301: eps_abs_curl(F) / eps_rel_curl(F) / alpha / beta");
      return DMK_ERR_CONFIG;
    }
    h->pol.alpha = n.alpha;
    h->pol.beta = n.beta;
  }
  h->N = (long)cells[0] * cells[1] * cells[2];
  // per-cell kernels address a field through a 32-bit byte offset of the cell
  if (h->N >= (1L << 29)) { dmk_set_error("grid of %ld cells exceeds the 2^29 - 1 cells of one device", h->N); return DMK_ERR_ARG; }
  h->h = cells[0] / 2 + 1;
  h->Nk = (long)h->h * cells[1] * cells[2];
  h->n9 = 9 * h->N;
  h->nx = h->scheme == 1 ? 2 * h->n9 : h->n9;
  h->wgt = 1.0 / (double)h->N;
  CK(hipStreamCreate(&h->st));
  if (dmk_fft_setup(&h->fft, cells)) { dmk_set_error("FFT setup failed"); return DMK_ERR_HIP; }
  // phases
  if (n_phases < 1 || n_phases > DMK_MAX_PHASES) { dmk_set_error("1..8 phases supported"); return DMK_ERR_CONFIG; }
  h->n_phases = n_phases;
  h->ph.resize(n_phases);
  h->nsl_max = 0;
  for (int p = 0; p < n_phases; p++) {
    int rc = dmk_build_phase(&phases[p], &h->ph[p]);
    if (rc) { dmk_set_error("phase %d: unsupported phenopowerlaw description (%d)", p, rc); return DMK_ERR_CONFIG; }
    if (h->ph[p].nsl > h->nsl_max) h->nsl_max = h->ph[p].nsl;
  }
  CK(hipMalloc(&h->ph_dev, sizeof(DmkPhaseDev) * n_phases));
  CK(hipMemcpy(h->ph_dev, h->ph.data(), sizeof(DmkPhaseDev) * n_phases, hipMemcpyHostToDevice));
  h->nm.r_cutback_min = num->r_cutback_min;
  h->nm.r_cutback = num->r_cutback;
  h->nm.r_increase = num->r_increase;
  h->nm.eps_rel_state = num->eps_rel_state;
  h->nm.N_iter_state_max = num->N_iter_state_max;
  h->nm.N_iter_Lp_max = num->N_iter_Lp_max;
  h->nm.r_linesearch_Lp = num->r_linesearch_Lp;
  h->nm.eps_rel_Lp = num->eps_rel_Lp;
  h->nm.eps_abs_Lp = num->eps_abs_Lp;
  h->nm.f_update_jacobi_Lp = num->f_update_jacobi_Lp;
  if (num->f_update_jacobi_Lp != 1) { dmk_set_error("f_update_jacobi_Lp other than 1 is not on this path"); return DMK_ERR_CONFIG; }
  const long N = h->N, n9 = h->n9;
  const int ns = h->nsl_max;
  // material point
  DmkMPDev& mp = h->mp;
  mp.N = (int)N;
  mp.stride_sl = ns;
  int32_t* phase_dev;
  int rc;
  if ((rc = dalloc(h, &phase_dev, N))) return rc;
  mp.phase = phase_dev;
  double** f9[] = {&mp.Fphase, &mp.F0, &mp.Fp, &mp.Fp0, &mp.Lp, &mp.Lp0, &mp.S, &mp.S0, &mp.Fe, &mp.Pph};
  for (auto pp : f9) if ((rc = dalloc(h, pp, n9))) return rc;
  double** fs[] = {&mp.xi, &mp.xi0, &mp.gam, &mp.gam0};
  for (auto pp : fs) if ((rc = dalloc(h, pp, (size_t)ns * N))) return rc;
  double** f2s[] = {&mp.st0s, &mp.last1, &mp.last2};
  for (auto pp : f2s) if ((rc = dalloc(h, pp, (size_t)2 * ns * N))) return rc;
  if ((rc = dalloc(h, &mp.sub, (size_t)DMK_SUB_ROWS * N))) return rc;
  if ((rc = dalloc(h, &mp.O, 4 * N))) return rc;
  if ((rc = dalloc(h, &mp.status, N))) return rc;
  if ((rc = dalloc(h, &mp.n_newton, N))) return rc;
  if ((rc = dalloc(h, &mp.n_fpi, N))) return rc;
  if ((rc = dalloc(h, &mp.n_sub, N))) return rc;
  if ((rc = dalloc(h, &h->P_hom, n9))) return rc;
  if ((rc = dalloc(h, &h->dPdF, 81 * N))) return rc;
  // solver iterates: [F] (basic) or [F; F_tau] (polarization), nx doubles each
  const long nx = h->nx;
  double** sol[] = {&h->F, &h->F_lastinc, &h->Fdot, &h->R, &h->XM, &h->FM, &h->XA, &h->FA};
  for (auto pp : sol) if ((rc = dalloc(h, pp, nx))) return rc;
  if ((rc = dalloc(h, &h->Fext, n9)) || (rc = dalloc(h, &h->Pext, n9))) return rc;
  if (h->scheme != 0 && (rc = dalloc(h, &h->T, n9))) return rc;
  // NGMRES history (M pairs), or for the Galerkin scheme the 31 Krylov vectors of GMRES(30)
  if (h->scheme == 2) {
    if ((rc = dalloc(h, &h->Xh, (size_t)31 * nx))) return rc;
    h->Fh = nullptr;
  } else {
    if ((rc = dalloc(h, &h->Xh, (size_t)h->num.ngmres_m * nx))) return rc;
    if ((rc = dalloc(h, &h->Fh, (size_t)h->num.ngmres_m * nx))) return rc;
  }
  h->poolX = {h->F, h->XM, h->XA};
  h->poolF = {h->R, h->FM, h->FA};
  for (int i = 0; h->scheme != 2 && i < h->num.ngmres_m; i++) {
    h->poolX.push_back(h->Xh + (size_t)i * nx);
    h->poolF.push_back(h->Fh + (size_t)i * nx);
  }
  if ((rc = dalloc(h, &h->Z, 9 * h->Nk))) return rc;
  // FFT tile shapes for this grid, timed on the device (no effect on results:
  // every line is transformed with the same operations whatever the tiling)
  if (dmk_fft_tune(&h->fft, h->Z, h->Fext, h->st)) { dmk_set_error("FFT tuning failed"); return DMK_ERR_HIP; }
  if ((rc = dalloc(h, &h->U, 3 * h->Nk))) return rc;
  if ((rc = dalloc(h, &h->u3, 3 * N))) return rc;
  long nn = (long)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1);
  if ((rc = dalloc(h, &h->xn_dev, 3 * nn))) return rc;
  if ((rc = dalloc(h, &h->xp_dev, 3 * N))) return rc;
  long np = 82L * dmk_nchunk(std::max(nx, 9 * h->Nk)) + 4096;
  if ((rc = dalloc(h, &h->partial, np))) return rc;
  if ((rc = dalloc(h, &h->small, SM_SIZE))) return rc;
  if ((rc = dalloc(h, &h->divp, (size_t)dmk_gamma_conv_partials(&h->fft)))) return rc;
  if ((rc = dalloc(h, &h->nrm, N))) return rc;
  if ((rc = dalloc(h, &h->pmax, dmk_nchunk(N)))) return rc;
  if ((rc = dalloc(h, &h->pmin, dmk_nchunk(N)))) return rc;
  if ((rc = dalloc(h, &h->ext, 2))) return rc;
  if ((rc = dalloc(h, &h->sel, 2))) return rc;
  if ((rc = dalloc(h, &h->csel, 2L * dmk_nchunk(N)))) return rc;
  if ((rc = dalloc(h, &h->flag, 4))) return rc;
  if ((rc = dalloc(h, &h->cnt, 4))) return rc;
  CK(hipHostMalloc((void**)&h->hsmall, sizeof(double) * SM_SIZE, hipHostMallocDefault));
  h->Q.assign((size_t)h->num.ngmres_m * h->num.ngmres_m, 0.0);
  // initial material-point state (phase_mechanical.f90:249-266, phenopowerlaw.f90:259-285)
  {
    std::vector<double> Fp(n9), Fe(n9), I(n9, 0.0), xi((size_t)ns * N, 0.0), O(4 * N);
    std::vector<int32_t> ph(phase_of_cell, phase_of_cell + N);
    for (long e = 0; e < N; e++) {
      if (ph[e] < 0 || ph[e] >= n_phases) { dmk_set_error("phase index out of range at cell %ld", e); return DMK_ERR_ARG; }
      double om[9];
      dmk_qu2om(q0 + 4 * e, om);   // row-major
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          Fp[(size_t)(i + 3 * j) * N + e] = om[3 * i + j];
          Fe[(size_t)(i + 3 * j) * N + e] = om[3 * j + i];
        }
      for (int i = 0; i < 3; i++) I[(size_t)(4 * i) * N + e] = 1.0;
      const DmkPhaseDev& P = h->ph[ph[e]];
      const dmk_phase_desc& D = phases[ph[e]];
      int a = 0;
      for (int f = 0; f < D.n_fam; f++)
        for (int s = 0; s < D.N_sl[f]; s++, a++) xi[(size_t)a * N + e] = D.xi_0[f];
      (void)P;
      for (int k = 0; k < 4; k++) O[k * N + e] = q0[4 * e + k];
    }
    CK(hipMemcpy((void*)mp.phase, ph.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.Fp, Fp.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.Fp0, Fp.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.Fe, Fe.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.Fphase, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.F0, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->F, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->F_lastinc, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    if (h->scheme == 1) {
      // F_tau = 2 F, F_tau_lastinc = 2 F_lastinc (grid_mech_spectral_polarization.f90:243-244)
      for (auto& v : I) v *= 2.0;
      CK(hipMemcpy(h->F + n9, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
      CK(hipMemcpy(h->F_lastinc + n9, I.data(), sizeof(double) * n9, hipMemcpyHostToDevice));
    }
    CK(hipMemcpy(mp.xi, xi.data(), sizeof(double) * ns * N, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.xi0, xi.data(), sizeof(double) * ns * N, hipMemcpyHostToDevice));
    CK(hipMemcpy(mp.O, O.data(), sizeof(double) * 4 * N, hipMemcpyHostToDevice));
  }
  // solver module defaults (grid_mech_spectral_basic.f90:69-82)
  std::memset(h->F_aimDot, 0, sizeof h->F_aimDot);
  std::memset(h->P_av, 0, sizeof h->P_av);
  std::memset(h->P_aim, 0, sizeof h->P_aim);
  std::memset(h->F_aim, 0, sizeof h->F_aim);
  h->F_aim[0] = h->F_aim[4] = h->F_aim[8] = 1.0;
  std::memcpy(h->F_aim_lastinc, h->F_aim, sizeof h->F_aim);
  std::memset(h->C_volAvgLastInc, 0, sizeof h->C_volAvgLastInc);
  std::memset(h->C_minMaxAvgLastInc, 0, sizeof h->C_minMaxAvgLastInc);
  std::memset(h->S, 0, sizeof h->S);
  for (int k = 0; k < 9; k++) h->stress_mask[k] = 1;
  h->q_R[0] = 1.0; h->q_R[1] = h->q_R[2] = h->q_R[3] = 0.0;
  // spectral set-up (spectral_utilities.f90:173-197)
  for (int d = 0; d < 3; d++) { h->ga.n[d] = cells[d]; h->ga.size[d] = size[d]; }
  h->ga.h = h->h;
  dmk_div_scale(cells, size, num->divergence_correction, h->ga.rg);
  // init: F = I, response at dt = 0, C_ref (grid_mech_spectral_basic.f90:213-235)
  if ((rc = constitutive(h, h->F, 0.0, false, true))) return rc;
  set_gamma_coef(h, h->C_minMaxAvg);
  std::memcpy(h->C_minMaxAvgRestart, h->C_minMaxAvg, sizeof h->C_minMaxAvg);
  if (h->scheme == 1 && (rc = set_scale(h))) return rc;
  return 0;
}

extern "C" int dmk_mech_init(const int32_t cells[3], const double size[3], int32_t n_phases,
                             const dmk_phase_desc* phases, const int32_t* phase_of_cell, const double* q0,
                             const dmk_numerics* num, dmk_handle** out) {
  if (!cells || !size || !phases || !phase_of_cell || !q0 || !out) {
    dmk_set_error("null argument");
    return DMK_ERR_ARG;
  }
  dmk_handle* h = new dmk_handle();
  int rc = init_impl(h, cells, size, n_phases, phases, phase_of_cell, q0, num);
  if (rc) { dmk_mech_destroy(h); *out = nullptr; return rc; }
  *out = h;
  return 0;
}

extern "C" int dmk_mech_forward(dmk_handle* h, int32_t cutBack, int32_t guess, double dt, double dt_old,
                                double t_remaining, const dmk_loadcase* bc) {
  if (!h || !bc) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  ForwardArgs fa{};
  fa.cutback = cutBack;
  if (cutBack) {
    std::memcpy(h->C_volAvg, h->C_volAvgLastInc, sizeof h->C_volAvg);
    std::memcpy(h->C_minMaxAvg, h->C_minMaxAvgLastInc, sizeof h->C_minMaxAvg);
  } else {
    std::memcpy(h->C_volAvgLastInc, h->C_volAvg, sizeof h->C_volAvg);
    std::memcpy(h->C_minMaxAvgLastInc, h->C_minMaxAvg, sizeof h->C_minMaxAvg);
    for (int k = 0; k < 9; k++) {
      int smask = bc->stress_type < 0 ? 1 : bc->stress_mask[k];
      h->F_aimDot[k] = guess ? (smask ? 0.0 : (h->F_aim[k] - h->F_aim_lastinc[k]) / dt_old) : 0.0;
    }
    std::memcpy(h->F_aim_lastinc, h->F_aim, sizeof h->F_aim);
    double V[9];
    for (int k = 0; k < 9; k++) V[k] = bc->def_mask[k] ? 0.0 : bc->def_values[k];
    if (bc->def_type == 0) {
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          double s = 0.0;
          for (int k = 0; k < 3; k++) s += V[c2(i, k)] * h->F_aim_lastinc[c2(k, j)];
          h->F_aimDot[c2(i, j)] += s;
        }
    } else if (bc->def_type == 1) {
      for (int k = 0; k < 9; k++) h->F_aimDot[k] += V[k];
    } else {
      for (int k = 0; k < 9; k++)
        h->F_aimDot[k] += bc->def_mask[k] ? 0.0 : (bc->def_values[k] - h->F_aim_lastinc[k]) / t_remaining;
    }
    rotate33(bc->q_R, h->F_aimDot, true, fa.avRate);
  }
  for (int k = 0; k < 9; k++) h->F_aim[k] = h->F_aim_lastinc[k] + h->F_aimDot[k] * dt;
  if (bc->stress_type == 0)
    for (int k = 0; k < 9; k++)
      h->P_aim[k] = h->P_aim[k] + (bc->stress_mask[k] ? 0.0 : (bc->stress_values[k] - h->P_aim[k]) / t_remaining) * dt;
  if (bc->stress_type == 1)
    for (int k = 0; k < 9; k++) h->P_aim[k] = h->P_aim[k] + (bc->stress_mask[k] ? 0.0 : bc->stress_values[k]) * dt;
  // F = F_lastinc + Fdot dt, mean corrected to the rotated aim
  CKL(dmk_forward_field(h->N, guess, dt_old, dt, &fa, h->F, h->F_lastinc, h->Fdot, h->F, h->st));
  CKL(dmk_colsum(h->F, h->N, h->N, 9, h->partial, h->small + SM_DOTS, h->wgt, h->st));
  rotate33(bc->q_R, h->F_aim, true, fa.aim);
  fa.mean = h->small + SM_DOTS;
  CKL(dmk_shift_field(h->N, &fa, h->F, h->st));
  if (h->scheme == 1) {
    // F_tau (grid_mech_spectral_polarization.f90:370-377,389-399): its rate and
    // last-increment copy as F's; with a guess forwarded by its own rate (no
    // mean correction), otherwise F_tau = I + S_scale:(F_l (C_scale:(F_l^T F_l - I)) / 2) + F
    // with F_l = F_tau - F at the new F
    const long n9 = h->n9;
    double* Ft = h->F + n9;
    if (guess) {
      CKL(dmk_forward_field(h->N, 1, dt_old, dt, &fa, Ft, h->F_lastinc + n9, h->Fdot + n9, Ft, h->st));
    } else {
      if (!cutBack) CKL(dmk_forward_field(h->N, 0, dt_old, dt, &fa, Ft, h->F_lastinc + n9, h->Fdot + n9, h->T, h->st));
      CKL(dmk_pol_lambda(h->N, &h->pol, h->F, Ft, h->st));
    }
  }
  for (int k = 0; k < 9; k++) h->stress_mask[k] = bc->stress_type < 0 ? 1 : bc->stress_mask[k];
  std::memcpy(h->q_R, bc->q_R, sizeof h->q_R);
  h->dt = dt;
  return 0;
}

// the six state copies in one launch
extern "C" int dmk_mech_materialpoint_forward(dmk_handle* h) {
  DmkMPDev& mp = h->mp;
  const long n9 = h->n9, ns = (long)h->nsl_max * h->N;
  CopyJobs j;
  const double* src[6] = {mp.Fp, mp.Fphase, mp.Lp, mp.S, mp.xi, mp.gam};
  double* dst[6] = {mp.Fp0, mp.F0, mp.Lp0, mp.S0, mp.xi0, mp.gam0};
  const long n[6] = {n9, n9, n9, n9, ns, ns};
  j.k = 6;
  for (int q = 0; q < 6; q++) { j.src[q] = src[q]; j.dst[q] = dst[q]; j.n[q] = n[q]; }
  CKL(dmk_copy_jobs(&j, h->st));
  return 0;
}

extern "C" int dmk_mech_solution(dmk_handle* h, int32_t* iterations, int32_t* conv, int32_t* n_residuals) {
  if (!h) { dmk_set_error("null handle"); return DMK_ERR_ARG; }
  int rc = masked_compliance(h);
  if (rc) return rc;
  if (h->num.update_gamma) {
    set_gamma_coef(h, h->C_minMaxAvg);
    if (h->scheme == 1 && (rc = set_scale(h))) return rc;
  }
  h->nfuncs = 0;
  int reason = 0;
  // F_aim is updated on the device during the solve (residual_aim_kernel)
  CK(hipMemcpyAsync(h->small + SM_FAIM, h->F_aim, sizeof h->F_aim, hipMemcpyHostToDevice, h->st));
  if (h->scheme == 2) rc = solve_newtonls(h, &reason);
  else rc = h->num.solver == 1 ? solve_richardson(h, &reason) : solve_ngmres(h, &reason);
  if (rc) return rc;
  if ((rc = tangent_pass(h, h->dt))) return rc;
  if (iterations) *iterations = h->totalIter;
  if (conv) *conv = reason > 0;
  if (n_residuals) *n_residuals = h->nfuncs;
  for (int k = 0; k < 9; k++) if (h->stress_mask[k]) h->P_aim[k] = h->P_av[k];
  return 0;
}

// enqueued; synchronises only when positions are copied out
extern "C" int dmk_mech_update_coords(dmk_handle* h, double* x_p, double* x_n) {
  CKL(dmk_fft_r2c(&h->fft, 9, h->F, h->Z, h->st));
  CKL(dmk_coords_freq(h->Z, h->U, &h->ga, h->st));
  CKL(dmk_fft_c2r(&h->fft, 3, h->U, h->u3, h->st));
  CoordsArgs ca;
  for (int d = 0; d < 3; d++) { ca.n[d] = h->cells[d]; ca.step[d] = h->size[d] / h->cells[d]; }
  ca.Fdc = h->Z;
  ca.Nk = h->Nk;
  ca.wgt = h->wgt;
  CKL(dmk_coords_nodes(h->u3, &ca, h->xn_dev, h->xp_dev, h->st));
  long nn = (long)(h->cells[0] + 1) * (h->cells[1] + 1) * (h->cells[2] + 1);
  if (x_n) CK(hipMemcpyAsync(x_n, h->xn_dev, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, h->st));
  if (x_p) CK(hipMemcpyAsync(x_p, h->xp_dev, sizeof(double) * 3 * h->N, hipMemcpyDeviceToHost, h->st));
  if (x_n || x_p) CK(hipStreamSynchronize(h->st));
  return 0;
}

extern "C" int dmk_constitutive(dmk_handle* h, double dt, const double* F, double* P, double* P_av,
                                double* C_volAvg, double* C_minMaxAvg, int32_t* status) {
  if (!h || !F) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  CK(hipMemcpyAsync(h->Pext, F, sizeof(double) * h->n9, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(h->Pext, h->N, 9, h->Fext, h->st));
  int rc = constitutive(h, h->Fext, dt, false, true);
  if (rc) return rc;
  if (P) {
    CKL(dmk_soa_to_aos(h->P_hom, h->N, 9, h->Pext, h->st));
    CK(hipMemcpyAsync(P, h->Pext, sizeof(double) * h->n9, hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
  }
  if (P_av) std::memcpy(P_av, h->P_av, sizeof h->P_av);
  if (C_volAvg) std::memcpy(C_volAvg, h->C_volAvg, sizeof h->C_volAvg);
  if (C_minMaxAvg) std::memcpy(C_minMaxAvg, h->C_minMaxAvg, sizeof h->C_minMaxAvg);
  if (status) *status = h->status;
  return 0;
}

extern "C" int dmk_constitutive_device(dmk_handle* h, double dt, const double* F_dev, double* P_dev, int32_t* status) {
  if (!h || !F_dev) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  CKL(dmk_aos_to_soa(F_dev, h->N, 9, h->Fext, h->st));
  int rc = constitutive(h, h->Fext, dt, false, true);
  if (rc) return rc;
  if (P_dev) {
    CKL(dmk_soa_to_aos(h->P_hom, h->N, 9, P_dev, h->st));
    CK(hipStreamSynchronize(h->st));
  }
  if (status) *status = h->status;
  return 0;
}

static int field_dev(dmk_handle* h, int what, const double** p, int* w) {
  const DmkMPDev& mp = h->mp;
  *w = 9;
  switch (what) {
    case 0: *p = h->F; return 0;
    case 1: *p = h->P_hom; return 0;
    case 2: *p = mp.Fe; return 0;
    case 3: *p = mp.Fp; return 0;
    case 4: *p = mp.Lp; return 0;
    case 5: *p = mp.S; return 0;
    case 7: *p = mp.O; *w = 4; return 0;
    case 8: *p = mp.xi; *w = h->nsl_max; return 0;
    case 9: *p = mp.gam; *w = h->nsl_max; return 0;
    case 10: *p = h->dPdF; *w = 81; return 0;
    case 11: *p = h->F_lastinc; return 0;
    case 12: if (h->scheme != 1) return 1; *p = h->F + h->n9; return 0;
    case 13: if (h->scheme != 1) return 1; *p = h->F_lastinc + h->n9; return 0;
    default: return 1;
  }
}

extern "C" int dmk_mech_get_field(dmk_handle* h, int32_t what, double* dst) {
  if (!h || !dst) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  if (what == 6) {   // F_i = I
    for (long e = 0; e < h->N; e++)
      for (int k = 0; k < 9; k++) dst[9 * e + k] = (k % 4 == 0) ? 1.0 : 0.0;
    return 0;
  }
  const double* p;
  int w;
  if (field_dev(h, what, &p, &w)) { dmk_set_error("unknown field %d", what); return DMK_ERR_ARG; }
  double* tmp;
  CK(hipMallocAsync((void**)&tmp, sizeof(double) * w * h->N, h->st));
  CKL(dmk_soa_to_aos(p, h->N, w, tmp, h->st));
  CK(hipMemcpyAsync(dst, tmp, sizeof(double) * w * h->N, hipMemcpyDeviceToHost, h->st));
  CK(hipFreeAsync(tmp, h->st));
  CK(hipStreamSynchronize(h->st));
  return 0;
}

extern "C" int dmk_mech_set_field(dmk_handle* h, int32_t what, const double* src) {
  if (!h || !src) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  const double* p;
  int w;
  if (field_dev(h, what, &p, &w) || what == 1 || what == 2 || what == 7 || what == 10) {
    dmk_set_error("field %d cannot be set", what);
    return DMK_ERR_ARG;
  }
  double* tmp;
  CK(hipMallocAsync((void**)&tmp, sizeof(double) * w * h->N, h->st));
  CK(hipMemcpyAsync(tmp, src, sizeof(double) * w * h->N, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(tmp, h->N, w, (double*)p, h->st));
  CK(hipFreeAsync(tmp, h->st));
  // the restart also sets the increment-start copies (phase_mechanical.f90:1265-1280)
  const DmkMPDev& mp = h->mp;
  size_t b = sizeof(double) * w * h->N;
  if (what == 3) CK(hipMemcpyAsync(mp.Fp0, mp.Fp, b, hipMemcpyDeviceToDevice, h->st));
  if (what == 4) CK(hipMemcpyAsync(mp.Lp0, mp.Lp, b, hipMemcpyDeviceToDevice, h->st));
  if (what == 5) CK(hipMemcpyAsync(mp.S0, mp.S, b, hipMemcpyDeviceToDevice, h->st));
  if (what == 8) CK(hipMemcpyAsync(mp.xi0, mp.xi, b, hipMemcpyDeviceToDevice, h->st));
  if (what == 9) CK(hipMemcpyAsync(mp.gam0, mp.gam, b, hipMemcpyDeviceToDevice, h->st));
  if (what == 0) {
    CK(hipMemcpyAsync(mp.Fphase, h->F, b, hipMemcpyDeviceToDevice, h->st));
    CK(hipMemcpyAsync(mp.F0, h->F, b, hipMemcpyDeviceToDevice, h->st));
  }
  CK(hipStreamSynchronize(h->st));
  return 0;
}

extern "C" int dmk_mech_get_scalars(dmk_handle* h, int32_t what, double* dst) {
  if (!h || !dst) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  switch (what) {
    case 0: std::memcpy(dst, h->F_aim, 9 * sizeof(double)); break;
    case 1: std::memcpy(dst, h->P_av, 9 * sizeof(double)); break;
    case 2: std::memcpy(dst, h->P_aim, 9 * sizeof(double)); break;
    case 3: std::memcpy(dst, h->C_volAvg, 81 * sizeof(double)); break;
    case 4: std::memcpy(dst, h->C_minMaxAvg, 81 * sizeof(double)); break;
    case 5: dst[0] = h->err_div; dst[1] = h->err_BC; break;
    case 10: dst[0] = h->err_curl; break;
    case 11:
      // grid_mech_spectral_basic.f90:388,408: the stiffness Gamma was built from
      // (the current C_minMaxAvg when update_gamma is set)
      // the Galerkin scheme never assigns it and stores zeros (grid_mech_spectral_Galerkin.f90:82,488)
      if (h->scheme == 2) std::memset(dst, 0, 81 * sizeof(double));
      else std::memcpy(dst, h->num.update_gamma ? h->C_minMaxAvg : h->C_minMaxAvgRestart, 81 * sizeof(double));
      break;
    case 6: std::memcpy(dst, h->F_aimDot, 9 * sizeof(double)); break;
    case 7: std::memcpy(dst, h->F_aim_lastinc, 9 * sizeof(double)); break;
    case 8: std::memcpy(dst, h->C_volAvgLastInc, 81 * sizeof(double)); break;
    case 9: std::memcpy(dst, h->C_minMaxAvgLastInc, 81 * sizeof(double)); break;
    case 12:
      dst[0] = h->fft.xy_threads; dst[1] = h->fft.xy_ip;
      dst[2] = h->fft.zgz_g; dst[3] = h->fft.zgz_threads; dst[4] = h->fft.zgz_ip;
      break;
    default: dmk_set_error("unknown scalar %d", what); return DMK_ERR_ARG;
  }
  return 0;
}

extern "C" int dmk_mech_set_scalars(dmk_handle* h, int32_t what, const double* src) {
  if (!h || !src) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  switch (what) {
    case 0: std::memcpy(h->F_aim, src, 9 * sizeof(double)); break;
    case 2: std::memcpy(h->P_aim, src, 9 * sizeof(double)); break;
    case 3: std::memcpy(h->C_volAvg, src, 81 * sizeof(double)); break;
    case 4: {
      // as the solver init after reading restart data: Gamma operator and,
      // for the polarization scheme, C_scale / S_scale from C_minMaxAvg
      std::memcpy(h->C_minMaxAvg, src, 81 * sizeof(double));
      std::memcpy(h->C_minMaxAvgRestart, src, 81 * sizeof(double));
      set_gamma_coef(h, src);
      if (h->scheme == 1) { int rc = set_scale(h); if (rc) return rc; }
      break;
    }
    case 6: std::memcpy(h->F_aimDot, src, 9 * sizeof(double)); break;
    case 7: std::memcpy(h->F_aim_lastinc, src, 9 * sizeof(double)); break;
    case 8: std::memcpy(h->C_volAvgLastInc, src, 81 * sizeof(double)); break;
    case 9: std::memcpy(h->C_minMaxAvgLastInc, src, 81 * sizeof(double)); break;
    default: dmk_set_error("scalar %d cannot be set", what); return DMK_ERR_ARG;
  }
  return 0;
}

extern "C" int dmk_mech_counters(dmk_handle* h, int64_t out[4]) {
  unsigned long long c[2];
  CK(hipMemcpyAsync(c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  out[0] = h->n_updates;
  out[1] = (int64_t)c[0];
  out[2] = (int64_t)c[1];
  out[3] = h->n_residuals;
  return 0;
}

extern "C" int dmk_mech_cell_counters(dmk_handle* h, int32_t* newton, int32_t* fpi, int32_t* substeps) {
  if (!h) return DMK_ERR_ARG;
  const size_t b = sizeof(int32_t) * (size_t)h->N;
  if (newton) CK(hipMemcpyAsync(newton, h->mp.n_newton, b, hipMemcpyDeviceToHost, h->st));
  if (fpi) CK(hipMemcpyAsync(fpi, h->mp.n_fpi, b, hipMemcpyDeviceToHost, h->st));
  if (substeps) CK(hipMemcpyAsync(substeps, h->mp.n_sub, b, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  return 0;
}

// math.f90:1222-1234
static double geometric_fraction(double r, int i, int N) {
  if (std::fabs(r - 1.0) <= 1.0e-9) return 1.0 / (double)N;
  return (1.0 - r) * std::pow(r, i) / (r - std::pow(r, N + 1));
}

// One increment of the DAMASK_grid loop (DAMASK_grid.f90:330-415): forward,
// materialpoint_forward, solution, and on non-convergence the cutback that
// halves the time step, until the increment's full time has converged.  The
// driver state at entry (t, Delta_t, Delta_t_prev, cutBack, cutBackLevel,
// guess) is recorded for the adjoint tape so that a segment of increments can
// be run again from a checkpoint with the same decisions.
static int run_increment(dmk_handle* h, const dmk_loadcase* lc, double t_0, int* guess, int* its_out,
                         int* nf_out) {
  if (h->adj) {
    IncStart& s = h->inc_start;
    s.lc = *lc;
    s.t_0 = t_0;
    s.t = h->t;
    s.Delta_t = h->Delta_t;
    s.Delta_t_prev = h->Delta_t_prev;
    s.cutBack = h->cutBack;
    s.cutBackLevel = h->cutBackLevel;
    s.guess = *guess;
    s.valid = true;
  }
  int stepFraction = 0, its = 0, nf_total = 0;
  while (stepFraction < (1 << h->cutBackLevel)) {
    double t_remaining = lc->t + t_0 - h->t;
    h->t = h->t + h->Delta_t;
    stepFraction++;
    int rc = dmk_mech_forward(h, h->cutBack, *guess, h->Delta_t, h->Delta_t_prev, t_remaining, lc);
    if (rc) return rc;
    if (!h->cutBack && (rc = dmk_mech_materialpoint_forward(h))) return rc;
    int32_t it, conv, nf;
    if ((rc = dmk_mech_solution(h, &it, &conv, &nf))) return rc;
    its = it;
    nf_total += nf;
    if (conv) {
      if ((rc = dmk_mech_update_coords(h, nullptr, nullptr))) return rc;
      if (h->adj && (rc = dmk_adjoint_record(h))) return rc;
      h->Delta_t_prev = h->Delta_t;
      h->cutBack = 0;
      *guess = 1;
    } else if (h->cutBackLevel < h->num.N_cutback_max) {
      h->cutBack = 1;
      stepFraction = (stepFraction - 1) * 2;
      h->cutBackLevel++;
      h->t = h->t - h->Delta_t;
      h->Delta_t = h->Delta_t / 2.0;
    } else {
      *its_out = its;
      *nf_out = nf_total;
      dmk_set_error("IO_error 950: no convergence after %d cutbacks", h->cutBackLevel);
      return 950;
    }
  }
  *its_out = its;
  *nf_out = nf_total;
  return 0;
}

extern "C" int dmk_grid_run_loadcase(dmk_handle* h, const dmk_loadcase* lc, double* t_io, double* stats,
                                     double* P_av_hist, double* F_aim_hist, int64_t* n_updates) {
  if (!h || !lc) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  double t_0 = h->t;
  int guess = lc->estimate_rate;
  int64_t u0 = h->n_updates;
  for (int inc = 1; inc <= lc->N; inc++) {
    h->totalIncs++;
    h->Delta_t_prev = h->Delta_t;
    h->Delta_t = lc->t * geometric_fraction(lc->r, inc, lc->N) * std::pow(2.0, -(double)h->cutBackLevel);
    if (h->totalIncs <= h->restartInc) {
      // already computed before the restart: advance the time only (DAMASK_grid.f90:300-303)
      h->t += h->Delta_t;
      guess = 1;
      if (stats) {
        stats[4 * (inc - 1) + 0] = h->t;
        stats[4 * (inc - 1) + 1] = h->cutBackLevel;
        stats[4 * (inc - 1) + 2] = 0;
        stats[4 * (inc - 1) + 3] = 0;
      }
      if (P_av_hist) std::memcpy(P_av_hist + 9 * (inc - 1), h->P_av, 9 * sizeof(double));
      if (F_aim_hist) std::memcpy(F_aim_hist + 9 * (inc - 1), h->F_aim, 9 * sizeof(double));
      continue;
    }
    int its = 0, nf_total = 0;
    int rc = run_increment(h, lc, t_0, &guess, &its, &nf_total);
    if (rc) {
      if (t_io) *t_io = h->t;
      if (n_updates) *n_updates = h->n_updates - u0;
      return rc;
    }
    if (stats) {
      stats[4 * (inc - 1) + 0] = h->t;
      stats[4 * (inc - 1) + 1] = h->cutBackLevel;
      stats[4 * (inc - 1) + 2] = its;
      stats[4 * (inc - 1) + 3] = nf_total;
    }
    h->cutBackLevel = h->cutBackLevel - 1 > 0 ? h->cutBackLevel - 1 : 0;
    if (P_av_hist) std::memcpy(P_av_hist + 9 * (inc - 1), h->P_av, 9 * sizeof(double));
    if (F_aim_hist) std::memcpy(F_aim_hist + 9 * (inc - 1), h->F_aim, 9 * sizeof(double));
  }
  if (t_io) *t_io = h->t;
  if (n_updates) *n_updates = h->n_updates - u0;
  return 0;
}

extern "C" int dmk_mech_restart_response(dmk_handle* h) {
  if (!h) { dmk_set_error("null handle"); return DMK_ERR_ARG; }
  int rc = constitutive(h, h->F, 0.0, false, true);
  if (rc) return rc;
  return dmk_mech_update_coords(h, nullptr, nullptr);
}

extern "C" int dmk_grid_set_restart(dmk_handle* h, int32_t restart_inc) {
  if (!h) { dmk_set_error("null handle"); return DMK_ERR_ARG; }
  h->restartInc = restart_inc;
  return 0;
}

// ------------------------------------------------------------------ spectral utilities on host data
// utilities_GammaConvolution (spectral_utilities.f90:436-501): field (3,3,N) -> out (3,3,N)
extern "C" int dmk_gamma_convolution(dmk_handle* h, const double* field, const double aim[9], double* out) {
  if (!h || !field || !aim || !out) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  CK(hipMemcpyAsync(h->Pext, field, sizeof(double) * h->n9, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(h->Pext, h->N, 9, h->Fext, h->st));
  // the transform path of the basic residual (x-y passes, z passes with Gamma)
  CKL(dmk_fft_r2c_xy(&h->fft, 9, h->Fext, h->Z, h->st));
  GammaArgs g = h->ga;
  g.do_div = 0;
  g.do_gamma = 1;
  g.aim_dev = nullptr;
  std::memcpy(g.aim, aim, sizeof g.aim);
  int nb;
  CKL(dmk_gamma_conv(&h->fft, h->Z, &g, h->divp, &nb, h->st));
  CKL(dmk_fft_c2r_xy(&h->fft, 9, h->Z, h->Fext, h->st));
  CKL(dmk_soa_to_aos(h->Fext, h->N, 9, h->Pext, h->st));
  CK(hipMemcpyAsync(out, h->Pext, sizeof(double) * h->n9, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  return 0;
}

// utilities_divergenceRMS (spectral_utilities.f90:650-691)
extern "C" int dmk_divergence_rms(dmk_handle* h, const double* field, double* out) {
  if (!h || !field || !out) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  CK(hipMemcpyAsync(h->Pext, field, sizeof(double) * h->n9, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(h->Pext, h->N, 9, h->Fext, h->st));
  CKL(dmk_fft_r2c(&h->fft, 9, h->Fext, h->Z, h->st));
  GammaArgs g = h->ga;
  g.do_div = 1;
  g.do_gamma = 0;
  CKL(dmk_gamma_div(h->Z, &g, h->partial, h->st));
  int nb = dmk_gamma_div_blocks(&g);
  CKL(dmk_colsum(h->partial, nb, nb, 1, h->partial + nb, h->small + SM_DOTS, 1.0, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_DOTS, h->small + SM_DOTS, sizeof(double), hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  double r = std::sqrt(h->hsmall[SM_DOTS]) * h->wgt;
  if (h->cells[0] == 1) r *= 0.5;
  *out = r;
  return 0;
}

// utilities_curlRMS (spectral_utilities.f90:697-757)
extern "C" int dmk_curl_rms(dmk_handle* h, const double* field, double* out) {
  if (!h || !field || !out) { dmk_set_error("null argument"); return DMK_ERR_ARG; }
  CK(hipMemcpyAsync(h->Pext, field, sizeof(double) * h->n9, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(h->Pext, h->N, 9, h->Fext, h->st));
  CKL(dmk_fft_r2c(&h->fft, 9, h->Fext, h->Z, h->st));
  CKL(dmk_curl_partial(h->Z, &h->ga, h->partial, h->st));
  int nb = dmk_gamma_div_blocks(&h->ga);
  CKL(dmk_colsum(h->partial, nb, nb, 1, h->partial + nb, h->small + SM_DOTS, 1.0, h->st));
  CK(hipMemcpyAsync(h->hsmall + SM_DOTS, h->small + SM_DOTS, sizeof(double), hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  double r = std::sqrt(h->hsmall[SM_DOTS]) * h->wgt;
  if (h->cells[0] == 1) r *= 0.5;
  *out = r;
  return 0;
}

// unnormalised r2c of `howmany` (<= 9) interleaved fields (howmany, N1, N2, N3) ->
// (howmany, N1/2+1, N2, N3) complex as (re, im) pairs
extern "C" int dmk_fft_forward(dmk_handle* h, int32_t howmany, const double* in, double* out) {
  if (!h || !in || !out || howmany < 1 || howmany > 9) { dmk_set_error("bad argument"); return DMK_ERR_ARG; }
  CK(hipMemcpyAsync(h->Pext, in, sizeof(double) * howmany * h->N, hipMemcpyHostToDevice, h->st));
  CKL(dmk_aos_to_soa(h->Pext, h->N, howmany, h->Fext, h->st));
  CKL(dmk_fft_r2c(&h->fft, howmany, h->Fext, h->Z, h->st));
  std::vector<double> tmp(2 * howmany * h->Nk);
  CK(hipMemcpyAsync(tmp.data(), h->Z, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  for (long f = 0; f < h->Nk; f++)
    for (int c = 0; c < howmany; c++) {
      out[2 * (c + howmany * f)] = tmp[2 * (c * h->Nk + f)];
      out[2 * (c + howmany * f) + 1] = tmp[2 * (c * h->Nk + f) + 1];
    }
  return 0;
}

// inverse of dmk_fft_forward (c2r, unnormalised)
extern "C" int dmk_fft_backward(dmk_handle* h, int32_t howmany, const double* in, double* out) {
  if (!h || !in || !out || howmany < 1 || howmany > 9) { dmk_set_error("bad argument"); return DMK_ERR_ARG; }
  std::vector<double> tmp(2 * howmany * h->Nk);
  for (long f = 0; f < h->Nk; f++)
    for (int c = 0; c < howmany; c++) {
      tmp[2 * (c * h->Nk + f)] = in[2 * (c + howmany * f)];
      tmp[2 * (c * h->Nk + f) + 1] = in[2 * (c + howmany * f) + 1];
    }
  CK(hipMemcpyAsync(h->Z, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice, h->st));
  CKL(dmk_fft_c2r(&h->fft, howmany, h->Z, h->Fext, h->st));
  CKL(dmk_soa_to_aos(h->Fext, h->N, howmany, h->Pext, h->st));
  CK(hipMemcpyAsync(out, h->Pext, sizeof(double) * howmany * h->N, hipMemcpyDeviceToHost, h->st));
  CK(hipStreamSynchronize(h->st));
  return 0;
}

// per-kernel timing: 0 off, 1 the stress kernel only, 2 every kernel group;
// totals reset when switched on
extern "C" int dmk_mech_profile(dmk_handle* h, int32_t enable) {
  if (!h) { dmk_set_error("null handle"); return DMK_ERR_ARG; }
  if (enable && !h->ev_ready) {
    for (int k = 0; k < 8; k++) {
      CK(hipEventCreate(&h->ev[k][0]));
      CK(hipEventCreate(&h->ev[k][1]));
    }
    h->ev_ready = true;
  }
  if (enable)
    for (int k = 0; k < 8; k++) { h->kt_ms[k] = 0.0; h->kt_n[k] = 0; }
  h->profile = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
  return 0;
}

// kernel groups: 0 stress integration, 1 orientation + tangent, 2 response
// reductions (status, P_av), 3 forward FFT, 4 Gamma + div, 5 inverse FFT,
// 6 stiffness averages: total ms and launch count
extern "C" int dmk_mech_kernel_time(dmk_handle* h, int32_t kernel, double* total_ms, int64_t* launches) {
  if (!h || kernel < 0 || kernel >= KT_N) { dmk_set_error("bad argument"); return DMK_ERR_ARG; }
  if (total_ms) *total_ms = h->kt_ms[kernel];
  if (launches) *launches = h->kt_n[kernel];
  return 0;
}
