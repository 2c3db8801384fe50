// Reduction / vector-kernel interfaces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// XA = XM (1 - sum alpha) + sum_k alpha_k X_k
struct CombineArgs {
  int l;
  double atot;
  double alpha[32];
  const double* X[32];
};

struct ForwardArgs {
  int cutback;
  double avRate[9];   // column-major rate for the homogeneous guess
  double aim[9];      // mean the forwarded field is corrected to (column-major)
  const double* mean; // device: the forwarded field's mean (the correction is mean - aim)
};

// up to 8 device-to-device copies in one launch (blockIdx.y = copy)
struct CopyJobs {
  const double* src[8];
  double* dst[8];
  long n[8];
  int k;
};

// several dot products in one launch: out[j] = x_j . y_j (mode 0) or
// |y_j - x_j|^2 (mode 1), each over n doubles
#define DMK_MAX_DOTS 64
struct DotJobs {
  const double* x[DMK_MAX_DOTS];
  const double* y[DMK_MAX_DOTS];
  int mode[DMK_MAX_DOTS];
  int n;
};

// per-residual stress-BC bookkeeping on the device (grid_mech_spectral_basic.f90:478-496):
// P_av (rotated to the load frame), F_aim -= S:(P_av - P_aim), err_BC, the
// aim of the Gamma convolution; report = {P_av 9, F_aim 9, err_BC, flag}
struct AimArgs {
  double S[81];
  double P_aim[9];
  double q_R[4];
  int32_t stress_mask[9];   // 1 where the component is not prescribed
  int32_t rotate;           // rotate P_av by q_R
  int32_t update;           // update F_aim and the aim (0: report P_av and flag only)
};
enum { REP_PAV = 0, REP_FAIM = 9, REP_ERRBC = 18, REP_FLAG = 19, REP_DIV2 = 20, REP_CURL2 = 21, REP_DOTS = 24, REP_SIZE = 24 + DMK_MAX_DOTS };

// sums of partial results: out[j][0] = scale[j] * sum of p[j][0..n[j])
#define DMK_MAX_SUMS (DMK_MAX_DOTS + 2)
struct SumJobs {
  const double* p[DMK_MAX_SUMS];
  double* out[DMK_MAX_SUMS];
  double scale[DMK_MAX_SUMS];
  int n[DMK_MAX_SUMS];
  int k;
};

int dmk_nchunk(long n);
int dmk_dots(const DotJobs* jobs, long n, double* partial, double* out, hipStream_t s);
// flag, iteration counters, raw P_av (kept when a cell failed) and, with a,
// the stress-BC bookkeeping of dmk_residual_aim, in two launches; partial
// needs 12 * dmk_nchunk(N) doubles
int dmk_response_reduce(const int32_t* st, const int32_t* nn, const int32_t* nf, const double* P, long N,
                        double wgt, double* partial, int* flag, unsigned long long* cnt, double* pav,
                        const AimArgs* a, double* faim, double* report, double* aim, hipStream_t s);
// the dot products of jobs (out[j]) and the sums of extra, with one finalize launch
int dmk_dots_sums(const DotJobs* jobs, long n, double* partial, double* out, const SumJobs* extra, hipStream_t s);
int dmk_residual_aim(const AimArgs* a, const double* pav_raw, double* faim, const int* flag, double* report,
                     double* aim, hipStream_t s);
// P_new <- P_old where the response failed (flag[0] != 0); nothing otherwise
int dmk_accept_P(const int* flag, const double* P_old, double* P_new, long n, hipStream_t s);
int dmk_colsum(const double* base, long ld, long n, int ncols, double* partial, double* out, double scale,
               hipStream_t s);
int dmk_mdot(const double* x, const double* Y, long ldY, int l, int mode, long n, double* partial, double* out,
             hipStream_t s);
int dmk_axpby(long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s);
int dmk_combine(long n, const double* XM, const CombineArgs* c, double* XA, hipStream_t s);
// C_minMaxAvg from the per-cell norms nrm (written by the tangent kernel):
// partial extrema, global extrema, per-chunk first qualifying cell, first overall, gather
int dmk_cminmax(const double* dPdF, long N, double band, const double* nrm, double* pmax, double* pmin,
                long* csel, long* sel, double* ext, double* out, hipStream_t s);
// convergence flag and iteration counters; partial needs 3 * dmk_nchunk(N) doubles
int dmk_status(const int32_t* st, const int32_t* nn, const int32_t* nf, long N, double* partial, int* flag,
               unsigned long long* cnt, hipStream_t s);
int dmk_forward_field(long N, int guess, double dt_old, double dt, const ForwardArgs* fa, const double* F,
                      double* F_lastinc, double* Fdot, double* Fnew, hipStream_t s);
int dmk_shift_field(long N, const ForwardArgs* fa, double* F, hipStream_t s);
int dmk_copy_jobs(const CopyJobs* j, hipStream_t s);
int dmk_aos_to_soa(const double* a, long N, int w, double* soa, hipStream_t s);
int dmk_soa_to_aos(const double* soa, long N, int w, double* a, hipStream_t s);

// constitutive launchers (constitutive.hip)
struct DmkMPDev;
struct DmkPhaseDev;
struct DmkNumMech;
extern "C" int dmk_launch_stress(const DmkMPDev* mp, const double* Ftarget, const DmkPhaseDev* ph,
                                 const DmkPhaseDev* ph_host, int phase_id, const DmkNumMech* num, double dt,
                                 hipStream_t s);
struct DmkStressItem;
// stress integration of several RVEs / phases with the same slip-system count
// and power evaluation in one launch (items on the device, block0 ascending)
extern "C" int dmk_launch_stress_batch(const DmkStressItem* items_dev, int n_items, int n_blocks, int nsl,
                                       int quarter_pow, hipStream_t s);
extern "C" int dmk_launch_tangent(const DmkMPDev* mp, const DmkPhaseDev* ph, const DmkPhaseDev* ph_host,
                                  int phase_id, double dt, const int* flag, double* dPdF, double* nrm,
                                  hipStream_t s);
