// Launch interfaces of the adjoint kernels (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmk_common.h"
#include "../../include/damask_hip.h"

#define DMK_NH_MAX 32
// parameter cotangents per cell: slot * nsl_max + system for the per-system
// parameters, then one slot per interaction type
enum { ADJ_TH_N = 0, ADJ_TH_XISAT = 1, ADJ_TH_H0 = 2, ADJ_TH_A = 3, ADJ_TH_G0 = 4, ADJ_TH_SYS = 5 };
// per-phase result of the backward sweep: xi_0 per system, the per-system
// slots above, the interaction types
#define DMK_ADJ_NOUT ((ADJ_TH_SYS + 1) * DMK_NSL_MAX + DMK_NH_MAX)

// per-system temporaries in the scratch: shear rate, its derivative by the
// active resolved stress, that stress, the sense (+-1), phi(xi), dphi/dxi,
// H |gdot|, then d tau_a / d L_p (9 each)
enum { AV_GD = 0, AV_KT, AV_TM, AV_SG, AV_PHI, AV_DPHI, AV_R, AV_T, AV_N = AV_T + 9 };

// scratch doubles per cell: per-system temporaries, the matrix block (G_y and
// its pivots) and the 9 tangent right-hand sides
inline size_t dmk_adj_scratch_slots(int mmax) {
  size_t n = 9 + (size_t)mmax;
  return (size_t)AV_N * mmax + n * n + n + 9 * n;
}


// local reverse step of one increment for the cells of one phase.
// mode 0: consistency residual (res, HUGE_VAL for a singular G_y), LU of G_y
//         (scr), tangent (A), F cotangent (g) for the output cotangent
//         (Pu, Fpb_in, xib_in);
// mode 1: with the LU of mode 0, the cotangents of F_p0 (Fpb_out), xi_0
//         (xib_out) and the parameters of this increment (thb, every slot of
//         the phase written) for the output cotangent (Pu - mu, Fpb_in, xib_in).
// Fields component-major ([c][N], c = i + 3 j); per-system arrays [a][N].
struct AdjKArgs {
  int N, phase_id, mode, mmax, nsl;    // nsl: slip systems of this phase
  double dt;
  double eps_rel_Lp, eps_abs_Lp, eps_rel_state;   // tolerances of the consistency residual
  const int32_t* phase;
  const double *F, *Fp0, *xi0, *Lp, *xi;
  double Pu[9];
  const double* mu;
  const double *Fpb_in, *xib_in;
  double *Fpb_out, *xib_out, *thb;
  double* scr;
  double *A, *g, *res;
};

#define DMK_GS_MAX 64
struct GsArgs {
  int l;
  double c[DMK_GS_MAX];
};

int dmk_adj_local(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s);
// the two halves of dmk_adj_local: K1 + K2 (read the tape only; write scr and
// res) and K3 (mode 0 or 1)
int dmk_adj_factor(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s);
int dmk_adj_vjp(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s);
int dmk_adj_atmul(long N, const double* A, const double* mu, double* y, hipStream_t s);
// out[p * nslot + s] (+)= sum over the cells of phase p of v[s * N + e]
int dmk_adj_phase_sum(const double* v, long N, const int32_t* phase, int n_phases, int nslot, int accumulate,
                      double* out, hipStream_t s);
int dmk_adj_max(const double* v, long N, double* out, hipStream_t s);
int dmk_adj_imax(const int32_t* v, long N, double* out, hipStream_t s);
int dmk_adj_gs(long n, double* w, const double* V, long ldV, const GsArgs* c, hipStream_t s);
