// One RVE's request in a lockstep round (batch.cpp).  Kind DMK_RQ_RESIDUAL:
// a residual evaluation of the basic scheme, i.e. what the batch stream needs
// after the stress integration: the arguments of dmk_response_reduce,
// dmk_accept_P, the x-y transform, dmk_gamma_conv, the inverse x-y transform
// and dmk_dots_sums of that RVE.  Kind DMK_RQ_ADJ_OP: one application of the
// adjoint's operator, y = A^T v per cell, the increment's mean rule on y's
// average (dmk_colsum, dmk_residual_aim), the same transform of y with Gamma,
// and the dot products.  Kind DMK_RQ_DOTS: the dot products alone.  The
// round's kernels read one of these per RVE (blockIdx.y or .z) from device
// memory and run the same per-block code as the single-RVE kernels.
#pragma once

#include "dmk_linalg.h"
#include "dmk_spectral.h"

enum { DMK_RQ_RESIDUAL = 0, DMK_RQ_ADJ_OP = 1, DMK_RQ_DOTS = 2 };

// vector steps of the adjoint GMRES carried by its next request and applied,
// in order, before the request's own work: w -= sum_k c_k V_k (DMK_PRE_GS,
// l <= DMK_PRE_GS_MAX terms, as dmk_adj_gs) or z = a x + b y (DMK_PRE_AXPBY,
// as dmk_axpby), over n doubles
enum { DMK_PRE_GS = 1, DMK_PRE_AXPBY = 2 };
#define DMK_PRE_OPS 3
#define DMK_PRE_GS_MAX 32
struct DmkPreOp {
  int type;
  int l;
  long n;
  double* w;
  const double* V;
  long ldV;
  double c[DMK_PRE_GS_MAX];
  double a, b;
  const double* x;
  const double* y;
  double* z;
};

struct DmkResidDev {
  int kind;
  int idle;                    // the RVE's stream has had no work since its last round: no input event
  int npre;
  DmkPreOp pre[DMK_PRE_OPS];
  // a vector step (DMK_PRE_AXPBY) after the request's transform and before
  // its dot products, e.g. r = b - Phi(A^T mu) ahead of |r|
  int npost;
  DmkPreOp post;
  // response reductions: status, iteration counts and P columns of the new
  // response (Pnew, N cells) -> flag, counters, raw P_av, aim bookkeeping
  const int32_t* st;
  const int32_t* nn;
  const int32_t* nf;
  double* Pnew;
  long N;
  double wgt;
  double* partial;
  int nc_resp;                 // dmk_nchunk(N)
  int* flag;
  unsigned long long* cnt;
  double* pav;
  double* faim;
  double* rep;                 // report block (REP_*)
  double* aim;
  AimArgs a;
  // the previous response's P, copied into Pnew where a cell failed
  const double* Pold;
  long n9;
  // forward x-y transform of Pnew into Z, z passes with Gamma and the
  // divergence partial sums (divp), inverse transform into r
  double2* Z;
  double* r;
  double* divp;
  GammaArgs g;
  // dot products over nx doubles (partial sums in partial) and the sums of S
  DotJobs J;
  long nx;
  int nc_dots;                 // dmk_nchunk(nx)
  SumJobs S;
  int rep_n;                   // doubles of the report block the host reads
  // adjoint operator: Pnew = A^T v (A: 81 per cell, A[p + 9 q]), its column
  // means times wgt into apav, the mean rule (a) into afaim, arep and aaim
  const double* A;
  const double* v;
  double* apav;
  double* afaim;
  double* arep;
  double* aaim;
};

// the transform part of the round for nr RVEs (R on the device) that share
// the FFT configuration fft; 1 when that configuration has no fused x-y or
// fused z kernels (the caller then runs each RVE's transform on its own)
int dmk_transform_batch(const DmkFFT* fft, const DmkResidDev* R, int nr, hipStream_t s);
// number of divergence partial sums dmk_gamma_conv writes for this configuration
int dmk_gamma_conv_nb(const DmkFFT* fft);
// reductions and report gather of the round (linalg.hip): max_* bound the
// per-RVE chunk and job counts; the reports go to rep_out[i * REP_SIZE]
int dmk_resid_reduce_batch(const DmkResidDev* R, int nr, int max_nc_resp, hipStream_t s);
// the adjoint-operator requests' A^T v (adjoint.hip) and mean rule (linalg.hip)
int dmk_adj_atmul_batch(const DmkResidDev* R, int nr, long max_N, hipStream_t s);
// the requests' vector steps (linalg.hip; max_n bounds their lengths): those
// before the request's work, and those after its transform
int dmk_pre_batch(const DmkResidDev* R, int nr, long max_n, hipStream_t s);
int dmk_post_batch(const DmkResidDev* R, int nr, long max_n, hipStream_t s);
int dmk_adj_mean_batch(const DmkResidDev* R, int nr, int max_nc, hipStream_t s);
int dmk_resid_dots_batch(const DmkResidDev* R, int nr, int max_nc_dots, int max_jobs, int max_sums, double* rep_out,
                         hipStream_t s);
