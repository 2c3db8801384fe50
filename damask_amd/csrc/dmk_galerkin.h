// Per-cell and bookkeeping kernels of the spectral-Galerkin scheme
// (src/grid/grid_mech_spectral_Galerkin.f90) and the vector kernels of its
// Newton-Krylov solver.  Fields are component-major (c = i + 3 j, each over
// N cells); dPdF is 81 components, c4 = i + 3 j + 9 k + 27 l.
#pragma once

#include <hip/hip_runtime.h>

struct AimArgs;
struct CombineArgs;

// out = K:dF per cell, K = dPdF (the matrix-free Jacobian's first factor, :634-640)
int dmk_gk_apply(long N, const double* dPdF, const double* dF, double* out, hipStream_t s);
// out = a0 x0 + sum_k alpha_k X_k (x0 may be null)
int dmk_lincomb(long n, double a0, const double* x0, const CombineArgs* c, double* out, hipStream_t s);
// form_residual's stress bookkeeping (:592-596): P_av in the load frame into
// the report with the flag, dP = merge(0, P_av - P_aim, stress_mask) into dc,
// err_BC = max |dP|; the report's F_aim slots are left as they are
int dmk_gk_bc(const AimArgs* a, const double* pav_raw, const int* flag, double* report, double* dc, hipStream_t s);
