// One RVE's residual evaluation of the basic scheme in a lockstep round
// (batch.cpp): what the batch stream needs after the stress integration,
// i.e. the arguments of dmk_response_reduce, dmk_accept_P, the x-y transform,
// dmk_gamma_conv, the inverse x-y transform and dmk_dots_sums of that RVE.
// The round's kernels read one of these per RVE (blockIdx.y or .z) from
// device memory and run the same per-block code as the single-RVE kernels.
#pragma once

#include "dmk_linalg.h"
#include "dmk_spectral.h"

struct DmkResidDev {
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
int dmk_resid_dots_batch(const DmkResidDev* R, int nr, int max_nc_dots, int max_jobs, int max_sums, double* rep_out,
                         hipStream_t s);
