// Reduction / vector-kernel interfaces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct CombineArgs {
  int l;
  double atot;
  double alpha[32];
};

struct ForwardArgs {
  int cutback;
  double avRate[9];   // column-major rate for the homogeneous guess
  double shift[9];    // mean correction (column-major)
};

int dmk_nchunk(long n);
int dmk_colsum(const double* base, long ld, long n, int ncols, double* partial, double* out, double scale,
               hipStream_t s);
int dmk_mdot(const double* x, const double* Y, long ldY, int l, int mode, long n, double* partial, double* out,
             hipStream_t s);
int dmk_axpby(long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s);
int dmk_combine(long n, const double* XM, const double* Xh, long ldX, const CombineArgs* c, double* XA,
                hipStream_t s);
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
int dmk_aos_to_soa(const double* a, long N, int w, double* soa, hipStream_t s);
int dmk_soa_to_aos(const double* soa, long N, int w, double* a, hipStream_t s);

// constitutive launchers (constitutive.hip)
struct DmkMPDev;
struct DmkPhaseDev;
struct DmkNumMech;
extern "C" int dmk_launch_stress(const DmkMPDev* mp, const double* Ftarget, const DmkPhaseDev* ph,
                                 const DmkPhaseDev* ph_host, int phase_id, const DmkNumMech* num, double dt,
                                 hipStream_t s);
extern "C" int dmk_launch_tangent(const DmkMPDev* mp, const DmkPhaseDev* ph, const DmkPhaseDev* ph_host,
                                  int phase_id, double dt, const int* flag, double* dPdF, double* nrm,
                                  hipStream_t s);
