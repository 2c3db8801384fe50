// Shared definitions of the MI355X grid-mechanics library (device + host).
//
// Device layout: every per-cell tensor field is stored component-major
// ("SoA"): component c of cell e at [c * N + e], with c = i + 3 j for T_ij
// (the reference's Fortran column-major order inside one 3x3), cells x fastest.
// Fourier fields: component c, frequency f = kx + h (ky + N2 kz) at [c * Nk + f],
// h = N1/2 + 1, as interleaved double2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DMK_NSL_MAX 24
#define DMK_MAX_PHASES 8
#define DMK_STRESS_WG 64       // work-group (one wavefront) of the per-cell kernels
#define DMK_SUB_ROWS 54        // rows of the per-cell stress-integration scratch (DmkMPDev::sub)

#define DMK_CHECK(x)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      dmk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));    \
      return DMK_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

enum {
  DMK_OK = 0,
  DMK_ERR_ARG = 1,
  DMK_ERR_HIP = 2,
  DMK_ERR_CONFIG = 3,
  DMK_ERR_NOCONV = 950,
};

// status codes of src/constants.f90:28-45
enum {
  ST_OK = 0,
  ST_ITERATING = 1,
  ST_FAIL_MECH = 2,
  ST_FAIL_STATE = 3,
  ST_FAIL_DELTASTATE = 4,
  ST_FAIL_STRESS = 5,
  ST_FAIL_LP_MAXIT = 6,
  ST_FAIL_LP_NAN = 7,
  ST_FAIL_LP_DGESV = 8,
};

void dmk_set_error(const char* fmt, ...);

// Per-phase constant tables, device resident (read through uniform loads).
// Tensors T_ij at [3 i + j] (row-major) for the tables used only inside the kernels.
struct DmkPhaseDev {
  int nsl;
  int nonschmid;                    // P_nS_neg differs from -P_nS_pos
  int quarter_pow;                  // every n_sl and a_sl is a positive multiple of 1/4, <= 64
  double C11, C12, C44;             // cubic stiffness (Voigt constants), lattice frame
  double P[DMK_NSL_MAX][9];         // Schmid tensor d (x) n
  double Pp[DMK_NSL_MAX][9];        // non-Schmid projection, positive sense
  double Pm[DMK_NSL_MAX][9];        // non-Schmid projection, negative sense
  double Qp[DMK_NSL_MAX][9];        // C : Pp
  double Qm[DMK_NSL_MAX][9];        // C : Pm
  double H[DMK_NSL_MAX][DMK_NSL_MAX];  // H[reacting][acting]
  double g0[DMK_NSL_MAX], n[DMK_NSL_MAX], a[DMK_NSL_MAX], h0[DMK_NSL_MAX];
  double xi_inf[DMK_NSL_MAX], c1[DMK_NSL_MAX], c2[DMK_NSL_MAX], fsat[DMK_NSL_MAX];
  // without twins (sum F_tw = 0): xi_inf + f_sat sqrt(0) and h_0 (1 + c_1 0^c_2)
  double xi_sat[DMK_NSL_MAX], hf[DMK_NSL_MAX];
  double atol_xi, atol_gamma;
  // interaction type (index into h_sl_sl) of H[reacting][acting]: parameter
  // cotangents of the adjoint
  int n_h;
  int32_t Ht[DMK_NSL_MAX][DMK_NSL_MAX];
};

struct DmkNumMech {
  double r_cutback_min, r_cutback, r_increase, eps_rel_state;
  int N_iter_state_max, N_iter_Lp_max;
  double r_linesearch_Lp, eps_rel_Lp, eps_abs_Lp;
  int f_update_jacobi_Lp;
  double eps_rel_Li, eps_abs_Li;
};

// Material-point state on the device (all SoA, N cells)
struct DmkMPDev {
  int N;
  int stride_sl;                 // nsl_max: state arrays are [nsl_max][N]
  const int32_t* phase;          // phase index per cell
  double* Fphase;                // phase F (9N), set by the partition
  double* F0;                    // F at the last converged increment
  double* Fp;
  double* Fp0;
  double* Lp;
  double* Lp0;
  double* S;
  double* S0;
  double* Fe;
  double* Pph;                   // phase P from the stress integration
  double* xi;
  double* xi0;
  double* gam;
  double* gam0;
  double* O;                     // 4N quaternion
  int32_t* status;
  int32_t* n_newton;             // per cell Newton iterations of the last response
  int32_t* n_fpi;
  int32_t* n_sub;                // per cell state integrations (sub-steps) of the last response
  // scratch of the state integrator: [2*nsl_max][N] each
  double* st0s;                  // state at the start of the sub-step (after the first sub-step)
  double* last1;                 // state-rate history, ping-pong pair
  double* last2;
  double* sub;                   // [DMK_SUB_ROWS][N]: sub-step F_0, F, F_p0, L_p0, line-search state
};

// one (RVE, phase) of a batched stress launch (stress_kernel_batch)
struct DmkStressItem {
  DmkMPDev mp;
  const double* F;               // target F of the RVE's cells (component-major)
  const DmkPhaseDev* ph;         // device phase tables
  DmkNumMech num;
  double dt;
  int phase_id;
  int block0;                    // first block of the item in the launch
};
