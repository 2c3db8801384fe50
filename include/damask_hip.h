/*
 * C-ABI of the MI355X grid spectral-basic mechanics library (libdamask_hip.so).
 *
 * Drop-in boundary for DAMASK_grid's mechanical solver plug-in: the five
 * procedure pointers bound in src/grid/DAMASK_grid.f90:110-119,171-178
 * (init / forward / solution / updateCoords / restartWrite of
 * src/grid/grid_mech_spectral_basic.f90:91-96) plus the material-point entry
 * homogenization_mechanical_response (src/homogenization.f90:294).
 *
 * Conventions (those of the Fortran arrays, so ISO_C_BINDING callers pass
 * their arrays unchanged):
 *   - a 3x3 tensor is 9 doubles, T_ij at [i + 3 j] (column-major);
 *   - per-cell fields are (3,3,N) column-major: cell e, T_ij at [i + 3 j + 9 e];
 *     cells run x fastest, then y, then z;
 *   - 3x3x3x3 tensors are 81 doubles, T_ijkl at [i + 3 j + 9 k + 27 l];
 *   - masks are int32, 1 where the load case has 'x' (component not prescribed);
 *   - quaternions (w, x, y, z), passive, P = -1 (src/rotations.f90:57).
 * Every function returns 0 on success, a DMK_ERR_* code otherwise; the text of
 * the last error is available from dmk_last_error().  Non-convergence of a
 * solution is not an error: it is reported through *converged = 0 so that the
 * caller applies its cutback (src/grid/DAMASK_grid.f90:398-408).
 * One handle = one RVE on one GPU; calls on one handle must not overlap.
 */
#ifndef DAMASK_HIP_H
#define DAMASK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMK_ERR_ARG 1
#define DMK_ERR_HIP 2
#define DMK_ERR_CONFIG 3

/* phenopowerlaw + Hooke phase as given in material.yaml
 * (src/phase_mechanical_plastic_phenopowerlaw.f90:146-285,
 *  src/phase_mechanical_elastic.f90:58-79).  Per-family arrays hold n_fam entries. */
typedef struct {
  int32_t lattice;          /* 0 cF, 1 cI */
  int32_t n_fam;
  int32_t N_sl[4];
  double C_11, C_12, C_44;
  double dot_gamma_0[4], n[4], a[4], h_0[4], xi_0[4], xi_inf[4];
  double c_1[4], c_2[4], f_sat[4];
  int32_t n_h;
  double h_sl_sl[32];
  int32_t n_nS;             /* number of non-Schmid coefficients (cI) */
  double a_nS[6];
  double atol_xi, atol_gamma;
} dmk_phase_desc;

/* numerics.yaml: solver:grid:mechanical, solver:grid, phase:mechanical
 * (src/grid/grid_mech_spectral_basic.f90:130-136, src/phase_mechanical.f90:281-295),
 * and the scheme of load.yaml solver:mechanical (src/grid/DAMASK_grid.f90:171-186)
 * with the spectral-polarization parameters
 * (src/grid/grid_mech_spectral_polarization.f90:136-147) */
typedef struct {
  int32_t N_iter_min, N_iter_max;
  double eps_abs_div, eps_rel_div, eps_abs_P, eps_rel_P;
  int32_t update_gamma;
  int32_t divergence_correction;   /* 0 none, 1 grid+size, 2 size */
  int32_t solver;                  /* 0 NGMRES (reference default), 1 Richardson */
  int32_t ngmres_m;
  int32_t snes_max_it;
  int32_t N_cutback_max;
  double r_cutback_min, r_cutback, r_increase, eps_rel_state;
  int32_t N_iter_state_max, N_iter_Lp_max;
  double r_linesearch_Lp, eps_rel_Lp, eps_abs_Lp;
  int32_t f_update_jacobi_Lp;
  int32_t cref_rule;               /* 0 strict first-extremum rule, 1 with 1e-12 dead band */
  int32_t scheme;                  /* 0 spectral_basic, 1 spectral_polarization, 2 spectral_Galerkin */
  double alpha, beta;              /* polarization: 0 < alpha <= 2, 0 < beta <= 2 */
  double eps_abs_curl, eps_rel_curl;
} dmk_numerics;

/* mechanical boundary conditions of one load step (src/grid/DAMASK_grid.f90:496-600) */
typedef struct {
  int32_t def_type;                /* 0 L, 1 dot_F, 2 F */
  double def_values[9];
  int32_t def_mask[9];
  int32_t stress_type;             /* -1 none, 0 P, 1 dot_P */
  double stress_values[9];
  int32_t stress_mask[9];
  double q_R[4];                   /* rotation of the load frame */
  double t, r;
  int32_t N, f_out, estimate_rate;
} dmk_loadcase;

typedef struct dmk_handle dmk_handle;

/* Allocate an RVE: cells/size (discretization_grid.f90:106-254), one phase
 * index (0-based) and initial orientation per cell (material.f90:95-218,
 * phase_mechanical.f90:249-266), then evaluate the constitutive response at
 * dt = 0 with F = I and set C_ref (grid_mech_spectral_basic.f90:103-237). */
int dmk_mech_init(const int32_t cells[3], const double size[3], int32_t n_phases,
                  const dmk_phase_desc* phases, const int32_t* phase_of_cell,
                  const double* q0_of_cell, const dmk_numerics* num, dmk_handle** h);

/* grid_mech_spectral_basic_forward (grid_mech_spectral_basic.f90:284-355) */
int dmk_mech_forward(dmk_handle* h, int32_t cutBack, int32_t guess, double dt, double dt_old,
                     double t_remaining, const dmk_loadcase* bc);

/* materialpoint_forward: state0 <- state (phase_mechanical.f90:987-1002) */
int dmk_mech_materialpoint_forward(dmk_handle* h);

/* grid_mech_spectral_basic_solution (grid_mech_spectral_basic.f90:243-277) */
int dmk_mech_solution(dmk_handle* h, int32_t* iterations, int32_t* converged, int32_t* n_residuals);

/* grid_mech_spectral_basic_updateCoords / utilities_updateCoords
 * (spectral_utilities.f90:857-968): x_p 3N cell centres, x_n
 * 3 (N1+1)(N2+1)(N3+1) nodes, either may be NULL */
int dmk_mech_update_coords(dmk_handle* h, double* x_p, double* x_n);

/* homogenization_mechanical_response + utilities_constitutiveResponse
 * (homogenization.f90:294-362, grid_utilities.f90:133-207) for a given F field
 * (host arrays, (3,3,N)).  P (3,3,N) out; P_av (9), C_volAvg (81),
 * C_minMaxAvg (81) out; status = max cell status (0 ok). */
int dmk_constitutive(dmk_handle* h, double dt, const double* F, double* P, double* P_av,
                     double* C_volAvg, double* C_minMaxAvg, int32_t* status);

/* same with device-resident F and P ((3,3,N) layout, for PyTorch callers) */
int dmk_constitutive_device(dmk_handle* h, double dt, const double* F_dev, double* P_dev,
                            int32_t* status);

/* per-cell fields to host buffers: 0 F, 1 P, 2 F_e, 3 F_p, 4 L_p, 5 S, 6 F_i,
 * 7 O (4/cell), 8 xi_sl (nsl/cell), 9 gamma_sl (nsl/cell), 10 dPdF (81/cell),
 * 11 F_lastinc; spectral-polarization handles also 12 F_tau, 13 F_tau_lastinc
 * (grid_mech_spectral_polarization.f90:452-455).
 * O and dPdF are those of the last response of the last completed
 * dmk_mech_solution or dmk_constitutive call: the solver evaluates them once
 * per solve, since grid_mech_spectral_basic reads the stiffness only between
 * solves (grid_mech_spectral_basic.f90:261-262,306-311). */
int dmk_mech_get_field(dmk_handle* h, int32_t what, double* dst);
/* restart data (grid_mech_spectral_basic.f90:198-232, phase_mechanical.f90:1248-1280):
 * 0 F, 3 F_p, 4 L_p, 5 S, 8 xi_sl, 9 gamma_sl, 11 F_lastinc, and for the
 * polarization scheme 12 F_tau, 13 F_tau_lastinc (grid_mech_spectral_polarization.f90:226-233) */
int dmk_mech_set_field(dmk_handle* h, int32_t what, const double* src);

/* solver scalars: 0 F_aim (9), 1 P_av (9), 2 P_aim (9), 3 C_volAvg (81),
 * 4 C_minMaxAvg (81), 5 {err_div, err_BC}, 6 F_aimDot (9), 7 F_aim_lastinc (9),
 * 8 C_volAvgLastInc (81), 9 C_minMaxAvgLastInc (81), 10 err_curl (1, polarization),
 * 11 C_minMaxAvgRestart (81, get only): the stiffness the Gamma operator was
 * built from, which restartWrite stores as C_minMaxAvg
 * (grid_mech_spectral_basic.f90:235,388,408), 12 transform path chosen at
 * setup (5, get only; no counterpart in the reference): threads of the fused
 * x-y kernels (0: separate passes), 1 if their single-buffer variant runs,
 * columns per block of the fused z+Gamma+z kernel (0: separate), its threads,
 * 1 if its single-buffer variant runs.  Setting 4 also sets the Gamma
 * operator, C_minMaxAvgRestart and, for the polarization scheme, C_scale and
 * S_scale, as the solver init does after reading restart data. */
int dmk_mech_get_scalars(dmk_handle* h, int32_t what, double* dst);
int dmk_mech_set_scalars(dmk_handle* h, int32_t what, const double* src);
/* restart read (grid_mech_spectral_basic.f90:198-232): after setting the
 * fields and F_aim, F_aim_lastinc, F_aimDot, P_aim, call this for the
 * material response at the restored F with dt = 0, then set C_volAvg,
 * C_volAvgLastInc and C_minMaxAvg */
int dmk_mech_restart_response(dmk_handle* h);
/* increments 1..restart_inc of the following dmk_grid_run_loadcase calls only
 * advance the time, with the rate guess on (DAMASK_grid.f90:300-303) */
int dmk_grid_set_restart(dmk_handle* h, int32_t restart_inc);

/* DAMASK_grid increment/cutback loop for one load case (DAMASK_grid.f90:288-460).
 * stats: 4 doubles per increment (t, cutback level, iterations, residual
 * evaluations); P_av_hist/F_aim_hist: 9 per increment (may be NULL);
 * n_updates: material-point updates performed.  Returns 0, or 950 when the
 * cutback limit is exhausted. */
int dmk_grid_run_loadcase(dmk_handle* h, const dmk_loadcase* lc, double* t_io, double* stats,
                          double* P_av_hist, double* F_aim_hist, int64_t* n_updates);

/* utilities_GammaConvolution (spectral_utilities.f90:436-501) on a (3,3,N)
 * host field with the handle's reference stiffness; aim = DC value */
int dmk_gamma_convolution(dmk_handle* h, const double* field, const double aim[9], double* out);
/* utilities_divergenceRMS (spectral_utilities.f90:650-691) of a (3,3,N) host field */
int dmk_divergence_rms(dmk_handle* h, const double* field, double* out);

/* utilities_divergenceRMS (spectral_utilities.f90:650-691) of a device-resident
 * (3,3,N) field, allocating no device memory.  No counterpart in the reference:
 * a variant of dmk_divergence_rms for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops).  It needs no handle: cells, size and
 * divergence_correction (0 none, 1 grid+size, 2 size) describe the grid, and
 * the RMS is written to rms_dev (one double, device memory) on stream (a
 * hipStream_t, NULL for the default stream); nothing is synchronised.  The
 * transform runs in one library-owned device array per GPU, sized for at most
 * DMK_NOALLOC_MAX_CELLS cells and 1024 cells per axis (larger grids are refused
 * with DMK_ERR_ARG); calls that overlap on one GPU share that array and race. */
#define DMK_NOALLOC_MAX_CELLS 262144
int dmk_divergence_rms_noalloc(const int32_t cells[3], const double size[3], int32_t divergence_correction,
                               const double* field_dev, double* rms_dev, void* stream);

/* utilities_curlRMS (src/grid/spectral_utilities.f90:697-757) of a (3,3,N)
 * field, the convergence measure of the spectral-polarization scheme; on the
 * x = 0 and x = N1/2 frequency planes it sums what the reference sums there
 * (nine times the squared (3,3) curl entry of the last row) */
int dmk_curl_rms(dmk_handle* h, const double* field, double* out);

/* math_invSym3333 (src/math.f90:511-532): inverse through the Mandel 6x6 form
 * (:826-904) of a column-major 3x3x3x3; nonzero on a zero pivot */
int dmk_inv_sym3333(const double* A, double* out);
/* x^y of the phenopowerlaw kinetics and hardening ((tau/xi)^n,
 * phase_mechanical_plastic_phenopowerlaw.f90:474-476; |1 - xi/xi_inf|^a,
 * :386-387) for n points
 * (host arrays), evaluated on the device as the stress kernel evaluates it:
 * quarter != 0 the quarter-power path of phases whose exponents are all
 * positive multiples of 1/4 up to 64 (other exponents are rejected), else the
 * library pow.  A check of the kernels' arithmetic; it allocates. */
int dmk_pow_law(const double* x, const double* y, double* out, int32_t n, int32_t quarter);
/* unnormalised 3-D r2c / c2r of howmany (<= 9) interleaved fields, FFTW's
 * conventions (spectral_utilities.f90:239-261): real (howmany, N1, N2, N3) <->
 * complex (howmany, N1/2+1, N2, N3) as (re, im) pairs; c2r ignores the
 * imaginary parts of the self-conjugate bins */
int dmk_fft_forward(dmk_handle* h, int32_t howmany, const double* in, double* out);
int dmk_fft_backward(dmk_handle* h, int32_t howmany, const double* in, double* out);

/* timing of the kernel groups with HIP events on the solver stream:
 * dmk_mech_profile(h, level) with level 1 times group 0 only, 2 every group
 * (both reset the totals), 0 disables.  Groups: 0 stress
 * integration, 1 orientation + tangent, 2 response reductions (status, P_av),
 * 3 forward FFT, 4 Gamma convolution + div RMS, 5 inverse FFT, 6 stiffness
 * averages C_volAvg / C_minMaxAvg */
int dmk_mech_profile(dmk_handle* h, int32_t enable);
int dmk_mech_kernel_time(dmk_handle* h, int32_t kernel, double* total_ms, int64_t* launches);

/* totals since init: material-point updates, Newton and state iterations,
 * residual evaluations */
int dmk_mech_counters(dmk_handle* h, int64_t out[4]);

/* per-cell counts of the last material-point response (N each, any may be
 * NULL): Lp Newton iterations, state fixed-point iterations, and state
 * integrations, i.e. sub-steps tried, counting the first full step and every
 * cut-back retry of phase_mechanical_constitutive (phase_mechanical.f90:1008-1090).
 * No counterpart in the reference (diagnostics of the material-point loop). */
int dmk_mech_cell_counters(dmk_handle* h, int32_t* newton, int32_t* fpi, int32_t* substeps);

/* Miller indices (direction d1 d2 d3, normal n1 n2 n3) of active slip system a
 * of a phase description (lattice, n_fam, N_sl): the system labels of the
 * xi_sl / gamma_sl results (crystal.f90:1642-1676, 2133-2181). */
int dmk_slip_system(int32_t lattice, int32_t n_fam, const int32_t* N_sl, int32_t a, double out[6]);

/* ---- Adjoint (parameter gradients of the converged increments).
 * None of these has a counterpart in the reference: they differentiate
 * grid_mech_spectral_basic's converged increments (implicit-function theorem
 * through the material-point increment, src/phase_mechanical.f90:379-676, and
 * through the spectral equilibrium, spectral_utilities.f90:436-501). */

/* Start a tape: the current state becomes entry 0, then up to max_increments
 * converged increments are recorded (dmk_grid_run_loadcase records each one
 * itself; a host driving forward/solution calls dmk_adjoint_record after each
 * converged increment).  Device memory: (max_increments + 1) (27 + nsl) N
 * doubles.  A new tape replaces the old one. */
int dmk_adjoint_tape(dmk_handle* h, int32_t max_increments);
/* The same with per-increment checkpointing and recomputation: the tape holds
 * the segment + 1 entries of one segment of increments, and every segment
 * increments a checkpoint of the full solver and material-point state
 * ((2 nx + 45 N + 2 nsl N) doubles) is kept.  The reverse sweep runs each
 * earlier segment again from its checkpoint, recording its entries, before it
 * differentiates it; the gradients are those of the full tape, bit for bit.
 * Device memory: (segment + 1) (27 + nsl) N doubles for the tape plus one
 * checkpoint per segment recorded (allocated as the run reaches it) and one
 * for the end-of-run state during a sweep.  Increments must be recorded by
 * dmk_grid_run_loadcase (the driver state of every increment is kept for the
 * recomputation); after a reverse sweep the handle holds the end-of-run state
 * again and the tape records no further increments.  segment = 0 or
 * >= max_increments: a full tape (dmk_adjoint_tape). */
int dmk_adjoint_tape_segmented(dmk_handle* h, int32_t max_increments, int32_t segment);
int dmk_adjoint_record(dmk_handle* h);
/* recorded increments n, their P_av (9 n, load frame) and time steps (n) */
int dmk_adjoint_info(dmk_handle* h, int32_t* n, double* P_av, double* dt);
/* Reverse sweep.  P_av_bar: cotangent of each recorded P_av (9 n).  rtol:
 * relative tolerance of the adjoint equilibrium solves.  theta_bar
 * (n_phases x 176): per phase the gradient with respect to xi_0 (24, per
 * slip system), n, xi_inf, h_0, a, dot_gamma_0 (24 each, per slip system;
 * sum over a family for the yaml value) and h_sl-sl (32, per interaction
 * type).  stats (4 n, optional): GMRES iterations, final relative residual,
 * largest local consistency residual (in units of the convergence
 * tolerances), largest sub-step count. */
int dmk_adjoint_backward(dmk_handle* h, const double* P_av_bar, double rtol, double* theta_bar, double* stats);
int dmk_adjoint_release(dmk_handle* h);
/* The local reverse step of one cell on the host (no GPU): tangent A (81,
 * dP_p/dF_q at p + 9 q) and the cotangents of F, F_p0, xi_0 and of the
 * parameters (accumulated into thbar, layout of theta_bar without the xi_0
 * block) for output cotangents (Pbar, Fpbar, xibar).  3x3 in field order. */
int dmk_adjoint_point_host(const dmk_phase_desc* d, double dt, const double F[9], const double Fp0[9],
                           const double* xi0, const double Lp[9], const double* xi, const double Pbar[9],
                           const double Fpbar[9], const double* xibar, double Fbar[9], double Fp0bar[9],
                           double* xi0bar, double* thbar, double A[81], double* residual);

/* ---- Many RVEs.  No counterpart in the reference (the batched
 * configuration of independent RVEs).  Each RVE is one handle with its own
 * stream; a pool of n_threads host threads runs dmk_grid_run_loadcase(h[i],
 * &lc[i]) for every i, largest cells[i] first (cells may be NULL).  rc_out
 * and wall_s (optional, n entries) receive each RVE's return code and wall
 * time.  Failures of single RVEs are reported through rc_out; the call then
 * returns 950. */
int dmk_batch_run(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells, int32_t n_threads,
                  int32_t* rc_out, double* wall_s);
/* as dmk_batch_run, with the RVEs in flight (one per host thread) in
 * lockstep: the stress integration of every residual evaluation waits until
 * each RVE still running has reached its own, and all of them run as one
 * launch per slip-system variant.  Results are those of dmk_batch_run. */
int dmk_batch_run_lockstep(dmk_handle** h, int32_t n, const dmk_loadcase* lc, const int64_t* cells,
                           int32_t n_threads, int32_t* rc_out, double* wall_s);
/* dmk_adjoint_backward(h[i], P_av_bar[i], rtol, theta_bar[i], NULL) for every i, concurrently:
 * at most n_threads sweeps at once, and no more than the free device memory
 * holds work buffers for.  Each handle's work buffers are released when its
 * sweep ends (the tape is kept; a later sweep allocates them again). */
int dmk_batch_adjoint(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                      double* const* theta_bar, const int64_t* cells, int32_t n_threads, int32_t* rc_out,
                      double* wall_s);
/* as dmk_batch_adjoint, with the sweeps in flight (one per host thread) in
 * lockstep: each application of the adjoint's operator (A^T v, the mean
 * rule, the transform with Gamma_0) with its dot products, and each set of
 * dot products alone, waits until every sweep still running has reached its
 * own, and all of them run as one round (one launch per kernel, the
 * transforms per FFT configuration) with one host synchronisation.
 * Gradients are those of dmk_batch_adjoint. */
int dmk_batch_adjoint_lockstep(dmk_handle** h, int32_t n, const double* const* P_av_bar, double rtol,
                               double* const* theta_bar, const int64_t* cells, int32_t n_threads,
                               int32_t* rc_out, double* wall_s);

void dmk_mech_destroy(dmk_handle* h);
const char* dmk_last_error(void);
int dmk_nsl(const dmk_handle* h);
int dmk_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
