/*
 * CPU oracle for the DAMASK grid spectral-basic mechanics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithm, used as the checker by tests/, by the smoke() entry
 * point and as the CPU baseline leg of bench.py.  No product code path may
 * load it.
 *
 * Layout conventions (identical to the reference's Fortran arrays):
 *   - a 3x3 tensor at one cell is 9 contiguous doubles, T_ij at [i + 3*j]
 *     (column-major), cells run x fastest, then y, then z;
 *   - a 3x3x3x3 tensor is 81 contiguous doubles, T_ijkl at
 *     [i + 3*j + 9*k + 27*l];
 *   - quaternions are (w, x, y, z) with w >= 0 (passive, P = -1 convention,
 *     src/rotations.f90:57).
 */
#ifndef DAMASK_ORACLE_H
#define DAMASK_ORACLE_H

#include <stdint.h>

#if defined(__cplusplus) && !defined(ORC_NO_EXTERN_C)
extern "C" {
#endif

#define ORC_NSL_MAX 48

enum { ORC_LATTICE_CF = 0, ORC_LATTICE_CI = 1 };

/* status codes: src/constants.f90:28-45 */
enum {
  ORC_STATUS_OK = 0,
  ORC_STATUS_ITERATING,
  ORC_STATUS_FAIL_PHASE_MECHANICAL,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_STATE,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_DELTASTATE,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_MAXIT,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_NAN,
  ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_DGESV
};

/* raw phenopowerlaw + Hooke description, as read from material.yaml */
typedef struct {
  int lattice;                 /* ORC_LATTICE_* */
  int n_fam;                   /* number of slip families given in N_sl */
  int N_sl[4];
  double C_11, C_12, C_44;
  double dot_gamma_0[4], n[4], a[4], h_0[4], xi_0[4], xi_inf[4];
  double c_1[4], c_2[4], f_sat[4];
  int n_h;                     /* number of h_sl-sl values */
  double h_sl_sl[32];
  int n_nS;                    /* number of non-Schmid coefficients (cI only) */
  double a_nS[6];              /* coefficients of family 1 */
  double atol_xi, atol_gamma;
} orc_phase_desc;

/* derived per-phase tables (phenopowerlaw.f90:76-295, elastic.f90:22-84) */
typedef struct {
  int nsl;
  double C66[6][6];
  double C3333[81];
  double P_sl[ORC_NSL_MAX][9];
  double P_nS_pos[ORC_NSL_MAX][9];
  double P_nS_neg[ORC_NSL_MAX][9];
  double H[ORC_NSL_MAX * ORC_NSL_MAX];   /* H[r*nsl + a]: reacting r, acting a */
  double dot_gamma_0[ORC_NSL_MAX], n[ORC_NSL_MAX], a[ORC_NSL_MAX], h_0[ORC_NSL_MAX];
  double xi_0[ORC_NSL_MAX], xi_inf[ORC_NSL_MAX], c_1[ORC_NSL_MAX], c_2[ORC_NSL_MAX];
  double f_sat[ORC_NSL_MAX];
  double atol[2 * ORC_NSL_MAX];
} orc_phase;

/* phase:mechanical numerics (phase_mechanical.f90:281-295) */
typedef struct {
  double r_cutback_min, r_cutback, r_increase, eps_rel_state;
  int N_iter_state_max;
  int N_iter_Lp_max;
  double r_linesearch_Lp, eps_rel_Lp, eps_abs_Lp;
  int f_update_jacobi_Lp;
  int N_iter_Li_max;
  double r_linesearch_Li, eps_rel_Li, eps_abs_Li;
  int f_update_jacobi_Li;
} orc_num_mech;

/* grid mechanical solver numerics (grid_mech_spectral_basic.f90:130-136) */
typedef struct {
  int N_iter_min, N_iter_max;
  double eps_abs_div, eps_rel_div, eps_abs_P, eps_rel_P;
  int update_gamma;
  int divergence_correction;   /* 0 none, 1 grid+size, 2 size */
  int solver;                  /* 0 NGMRES (reference default), 1 Richardson */
  int ngmres_m;                /* history size (PETSc default 30) */
  int snes_max_it;             /* PETSc default 50 */
  int N_cutback_max;
} orc_num_grid;

void orc_default_num_mech(orc_num_mech *n);
void orc_default_num_grid(orc_num_grid *n);

int orc_phase_build(const orc_phase_desc *d, orc_phase *p);

/* ---- math / rotation / crystal building blocks (exported for tests) ---- */
void orc_qu2om(const double q[4], double om[9]);
void orc_om2qu(const double om[9], double q[4]);
void orc_rotate33(const double q[4], const double T[9], int active, double out[9]);
void orc_rotate3333(const double q[4], const double T[81], int active, double out[81]);
void orc_rotational_part(const double F[9], double R[9]);
int  orc_invert(int n, const double *A, double *invA);            /* column-major */
int  orc_inv_sym3333(const double *A, double *invA);                /* math_invSym3333 */
void orc_schmid_slip(int lattice, int n_fam, const int *N_sl, int sense,
                     int n_nS, const double *a_nS, double *P /* [nsl][9] */);
void orc_interaction_slipslip(int lattice, int n_fam, const int *N_sl,
                              int n_val, const double *val, double *H);

/* ---- material point ---- */
typedef struct orc_mp orc_mp;

orc_mp *orc_mp_create(int n_cells, int n_phases, const orc_phase *phases,
                      const int32_t *phase_of_cell,        /* 0-based */
                      const double *q0_of_cell,             /* 4 per cell */
                      const orc_num_mech *num);
void orc_mp_destroy(orc_mp *m);

/* homogenization_mechanical_response + crystallite orientations + homogenize
 * for all cells (homogenization.f90:294-362).  F_target/P_out are (9,N).
 * dPdF_out (81,N) may be NULL.  Returns the max status over cells. */
int orc_mp_response(orc_mp *m, double dt, const double *F_target, double *P_out,
                    double *dPdF_out);
/* materialpoint_forward: state0 <- state (phase_mechanical.f90:987-1002) */
void orc_mp_forward(orc_mp *m);
/* fetch per-cell fields: 0 F, 1 P, 2 F_e, 3 F_p, 4 L_p, 5 S, 6 F_i, 7 O (4/cell),
 * 8 xi_sl (nsl/cell), 9 gamma_sl (nsl/cell) */
void orc_mp_get(const orc_mp *m, int what, double *dst);
void orc_mp_set(orc_mp *m, int what, const double *src);
int  orc_mp_nsl(const orc_mp *m);
/* per-cell statistics of the last response (Newton and FPI iteration counts) */
void orc_mp_counters(const orc_mp *m, int64_t *newton, int64_t *fpi, int64_t *cutbacks);
void orc_mp_cell_counters(const orc_mp *m, int64_t *newton, int64_t *fpi);
void orc_mp_cell_substeps(const orc_mp *m, int64_t *substeps);
/* single-cell kernels used by tests */
int  orc_integrate_stress(orc_mp *m, int ce, const double F[9], const double Fp0[9],
                          const double Fi0[9], double dt);
void orc_dPdF(const orc_mp *m, int ce, double dt, double out[81]);
void orc_kinetics_sl(const orc_phase *p, const double Mp[9], const double *xi,
                     double *dot_gamma, double *ddot_gamma_dtau);
void orc_hooke(const orc_phase *p, const double Fe[9], const double Fi[9],
               double S[9], double dS_dFe[81], double dS_dFi[81]);
void orc_set_threads(int n);

/* ---- spectral utilities ---- */
typedef struct orc_grid orc_grid;

orc_grid *orc_grid_create(const int cells[3], const double size[3],
                          const orc_num_grid *num);
void orc_grid_destroy(orc_grid *g);
/* unnormalised r2c / c2r of `howmany` interleaved fields: real (howmany, N1, N2, N3),
 * complex (howmany, N1/2+1, N2, N3) as re/im pairs */
void orc_fft_r2c(const orc_grid *g, int howmany, const double *in, double *out);
void orc_fft_c2r(const orc_grid *g, int howmany, const double *in, double *out);
void orc_update_gamma(orc_grid *g, const double C[81]);
void orc_gamma_convolution(orc_grid *g, const double *field, const double aim[9], double *out);
double orc_divergence_rms(orc_grid *g, const double *field);
double orc_curl_rms(orc_grid *g, const double *field);
void orc_update_coords(orc_grid *g, const double *F, double *x_p, double *x_n);

/* ---- grid mechanical spectral basic solver + DAMASK_grid increment loop ---- */
typedef struct orc_solver orc_solver;

typedef struct {
  int def_type;           /* 0 L, 1 dot_F, 2 F */
  double def_values[9];   /* column-major 3x3 */
  int32_t def_mask[9];    /* 1 = 'x' (not prescribed) */
  int stress_type;        /* -1 none, 0 P, 1 dot_P */
  double stress_values[9];
  int32_t stress_mask[9];
  double q_R[4];          /* load frame rotation */
  double t, r;
  int N, f_out, estimate_rate;
} orc_loadcase;

orc_solver *orc_solver_create(orc_grid *g, orc_mp *m, const orc_num_grid *num);
void orc_solver_destroy(orc_solver *s);
void orc_solver_forward(orc_solver *s, int cutBack, int guess, double dt, double dt_old,
                        double t_remaining, const orc_loadcase *lc);
/* returns converged flag; iterations in *its, residual evaluations in *nfuncs */
int orc_solver_solution(orc_solver *s, int *its, int *nfuncs);
void orc_solver_update_coords(orc_solver *s, double *x_p, double *x_n);
void orc_solver_get(const orc_solver *s, int what, double *dst); /* 0 F field, 1 P field,
                        2 F_aim, 3 P_av, 4 C_volAvg, 5 C_minMaxAvg, 6 P_aim, 7 err (div, BC),
                        8 S (masked compliance), 9 F_tau field, 10 err_curl, 11 C_scale (polarization) */
/* switch a fresh solver to spectral_polarization (grid_mech_spectral_polarization.f90);
 * nonzero for beta <= 0 */
/* spectral_basic only: dPdF, orientation and the stiffness averages once per
 * solve from the last response instead of after every response (the scheme
 * reads them only between solves); nonzero for other schemes */
int orc_solver_set_tangent_once(orc_solver *s, int once);
int orc_solver_set_polarization(orc_solver *s, double alpha, double beta, double eps_abs_curl, double eps_rel_curl);
/* one load case of DAMASK_grid (DAMASK_grid.f90:288-460).  stats (per increment,
 * 8 doubles): t, cutBackLevel, iterations, residual evaluations, P_av_11..  P_av
 * written separately.  Returns 0 on success, 950 when no convergence. */
int orc_run_loadcase(orc_solver *s, const orc_loadcase *lc, double *t_io,
                     double *stats /* N*4 */, double *P_av_hist /* N*9 */,
                     double *F_aim_hist /* N*9 */, int64_t *n_updates);

/* The NGMRES iteration of the solver (PETSc SNES NGMRES restated, see
 * oracle_solver.c) on an affine residual r(x) = A x - b (A n x n column-major),
 * from x (in/out), history M <= 30, at most max_it iterations, converged when
 * |r| <= atol.  trace (optional, trace_cap records) receives per iteration
 * l, selectA, restarted, |r| before, |r(x^M)|, |r(x^A)|, alpha[M], then x^M's
 * residual, x^M, x^A, x^A's residual, the accepted x, and the history of
 * residuals and iterates (M n each) the least squares used; *n_trace records.
 * Returns the convergence reason (> 0 converged). */
int orc_ngmres_affine(int n, const double *A, const double *b, double *x, int M, int max_it, double atol,
                      double *trace, int trace_cap, int *n_trace);

#if defined(__cplusplus) && !defined(ORC_NO_EXTERN_C)
}
#endif
#endif
