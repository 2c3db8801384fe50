// Per-cell kernels of the spectral-polarization scheme
// (src/grid/grid_mech_spectral_polarization.f90).  Fields are component-major
// (component c = i + 3 j of a Fortran (3,3) tensor, each over N cells);
// 3x3x3x3 tensors are Fortran column-major (i + 3 j + 9 k + 27 l).
#pragma once

#include <hip/hip_runtime.h>

struct AimArgs;

struct PolArgs {
  double C[81];        // C_scale
  double S[81];        // S_scale = invSym3333(C_scale)
  double alpha, beta;
};

// T = beta C:(F - I) - alpha F (C:(F_tau - F - I))   (:582-587)
int dmk_pol_tau(long N, const PolArgs* a, const double* F, const double* F_tau, double* T, hipStream_t s);
// r_F_tau = beta F - G (G the Gamma convolution of T), and the deformation
// gradient of the constitutive response, F - r_F_tau / beta   (:588-589,601)
int dmk_pol_stage2(long N, double beta, const double* F, const double* G, double* r_tau, double* F_resp,
                   hipStream_t s);
// r_F = invSym3333(dPdF + C_scale):(P - F (C_scale:(F_tau - F - I))) + r_F_tau   (:608-620)
int dmk_pol_rF(long N, const PolArgs* a, const double* F, const double* F_tau, const double* P,
               const double* dPdF, const double* r_tau, double* r_F, hipStream_t s);
// forward without a rate guess (:392-398):
// F_tau = I + S_scale:(1/2 F_l (C_scale:(F_l^T F_l - I))) + F,  F_l = F_tau - F
int dmk_pol_lambda(long N, const PolArgs* a, const double* F, double* F_tau, hipStream_t s);

// small bookkeeping kernels (one thread)
// aim of the Gamma convolution of the polarization residual: rotate(beta F_aim, active)  (:589)
int dmk_pol_aim_pre(const double q_R[4], double beta, const double* faim, double* aim, hipStream_t s);

// end of the residual (:627-632): P_av (load frame), F_aim update, err_BC with
// the mean F_av of the compatible field; report layout of dmk_residual_aim
int dmk_pol_residual_aim(const AimArgs* a, const PolArgs* pa, const double* pav_raw, const double* fav_raw,
                         double* faim, const int* flag, double* report, hipStream_t s);

// math_invSym3333 on the host (declared in the public header)
extern "C" int dmk_inv_sym3333(const double* A, double* out);
