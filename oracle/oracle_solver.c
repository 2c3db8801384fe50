/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of src/grid/grid_mech_spectral_basic.f90 (init/forward/solution/
 * form_residual/converged), src/grid/grid_utilities.f90 (maskedCompliance,
 * constitutiveResponse reductions, calculateRate, forwardTensorField) and the
 * increment/cutback loop of src/grid/DAMASK_grid.f90:288-460.
 *
 * The nonlinear solver the reference selects (PETSc SNES type "ngmres",
 * grid_mech_spectral_basic.f90:150) is not vendored.  It is restated from the
 * published algorithm (C. W. Oosterlee and T. Washio, SIAM J. Sci. Comput.
 * 21(5), 2000; P. R. Brune, M. G. Knepley, B. F. Smith, X. Tu, SIAM Review
 * 57(4), 2015, section 4.2) with PETSc's documented defaults: history m = 30,
 * difference-based candidate selection (gamma_A = 2, epsilon_B = 0.1,
 * delta_B = 0.9), difference-based restart (gamma_C = 2, restart after 2
 * consecutive restart conditions), preconditioning step x^M = x - r(x) with a
 * unit basic line search, and SNES max_it = 50.  The iterate sequence of PETSc
 * itself is therefore not pinned; a plain Richardson fixed point (x <- x - r(x))
 * is available as solver = 1.
 */
#include "oracle_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

const double *orc_mp_dPdF_ptr(const orc_mp *m);
void orc_mp_tangent(orc_mp *m, double dt);
void orc_mp_set_lazy_tangent(orc_mp *m, int lazy);
int orc_mp_ncells(const orc_mp *m);
double orc_grid_wgt(const orc_grid *g);
const int *orc_grid_cells(const orc_grid *g);

struct orc_solver {
  orc_grid *g;
  orc_mp *m;
  orc_num_grid num;
  int N;
  double *F, *F_lastinc, *Fdot, *P;
  double F_aimDot[9], F_aim[9], F_aim_lastinc[9], P_av[9], P_aim[9];
  double C_volAvg[81], C_volAvgLastInc[81], C_minMaxAvg[81], C_minMaxAvgLastInc[81];
  double S[81];
  double err_BC, err_div;
  int status;
  int totalIter;
  int32_t stress_mask[9];
  double q_R[4];
  double dt;
  int64_t n_updates;
  int nfuncs;
  /* NGMRES storage */
  double *Xh, *Fh, *XM, *FM, *XA, *FA, *R, *D;
  double *Q;
  /* driver state (DAMASK_grid.f90:78-100) */
  double t, Delta_t, Delta_t_prev;
  int cutBack, cutBackLevel, totalIncs;
  int cref_rule;
  /* spectral_polarization (grid_mech_spectral_polarization.f90): the iterate
   * is [F; F_tau] (18 N doubles, F_tau = F + 9 N) */
  int scheme;
  double alpha, beta, eps_abs_curl, eps_rel_curl;
  double *F_tau_lastinc, *F_tauDot;
  double C_scale[81], S_scale[81];
  double err_curl, F_av[9];
  /* tangent once per solve: the basic scheme reads C_volAvg and C_minMaxAvg
   * only between solves (grid_mech_spectral_basic.f90:261-262,306-311), so
   * dPdF, the orientation and the stiffness averages of the last response can
   * be formed after the solve instead of after every response */
  int tangent_once;
};

/* grid_utilities.f90:74-127 utilities_maskedCompliance */
static void masked_compliance(const double q[4], const int32_t *mask_stress, const double C[81], double *Sout) {
  int msv[9], sel[9], n = 0;
  double T99[81] = {0};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) msv[3 * a + b] = !mask_stress[I2(a, b)];
  for (int i = 0; i < 9; i++) if (msv[i]) sel[n++] = i;
  if (n > 0) {
    double Cr[81], C99[81], c[81], s[81];
    orc_rotate3333(q, C, 0, Cr);
    t3333_to99(Cr, C99);
    for (int qq = 0; qq < n; qq++)
      for (int p = 0; p < n; p++) c[p + n * qq] = C99[sel[p] + 9 * sel[qq]];
    orc_invert(n, c, s);
    for (int qq = 0; qq < n; qq++)
      for (int p = 0; p < n; p++) T99[sel[p] + 9 * sel[qq]] = s[p + n * qq];
  }
  t99_to3333(T99, Sout);
}

static void stiffness_averages(orc_solver *s);

/* grid_utilities.f90:133-207 utilities_constitutiveResponse */
static void constitutive_response(orc_solver *s, const double *F, double dt, int rotate) {
  int N = s->N;
  s->status = orc_mp_response(s->m, dt, F, s->P, NULL);
  s->n_updates += N;
  double wgt = orc_grid_wgt(s->g);
  double Pav[9] = {0};
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 9; k++) Pav[k] += s->P[9 * ce + k];
  for (int k = 0; k < 9; k++) Pav[k] *= wgt;
  if (rotate) orc_rotate33(s->q_R, Pav, 0, Pav);
  memcpy(s->P_av, Pav, sizeof Pav);
  if (!s->tangent_once) stiffness_averages(s);
}

/* grid_utilities.f90:172-206: C_minMaxAvg and C_volAvg from the tangents of
 * the last successful response */
static void stiffness_averages(orc_solver *s) {
  int N = s->N;
  double wgt = orc_grid_wgt(s->g);
  const double *dPdF = orc_mp_dPdF_ptr(s->m);
  /* argmax / argmin of |dPdF|^2 (grid_utilities.f90:172-199): first cell, in
   * Fortran order, whose norm reaches the extremum.  With cref_rule 1 a cell
   * within a relative 1e-12 of the extremum counts as reaching it, so that the
   * choice among norms that agree up to rounding does not depend on rounding. */
  double band = s->cref_rule == 1 ? 1.0e-12 : 0.0;
  double *nrm = (double *)malloc(sizeof(double) * N);
  double gmax = 0.0, gmin = DBL_MAX;
  for (int ce = 0; ce < N; ce++) {
    double v = 0.0;
    for (int k = 0; k < 81; k++) v += dPdF[(size_t)81 * ce + k] * dPdF[(size_t)81 * ce + k];
    nrm[ce] = v;
    if (v > gmax) gmax = v;
    if (v < gmin) gmin = v;
  }
  int imax = -1, imin = -1;
  for (int ce = 0; ce < N; ce++) {
    if (imax < 0 && gmax > 0.0 && nrm[ce] * (1.0 + band) >= gmax) imax = ce;
    if (imin < 0 && gmin < DBL_MAX && nrm[ce] * (1.0 - band) <= gmin) imin = ce;
  }
  free(nrm);
  for (int k = 0; k < 81; k++) {
    double mx = imax >= 0 ? dPdF[(size_t)81 * imax + k] : 0.0;
    double mn = imin >= 0 ? dPdF[(size_t)81 * imin + k] : DBL_MAX;
    s->C_minMaxAvg[k] = 0.5 * (mx + mn);
  }
  double Cv[81] = {0};
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 81; k++) Cv[k] += dPdF[(size_t)81 * ce + k];
  for (int k = 0; k < 81; k++) s->C_volAvg[k] = Cv[k] * wgt;
}

orc_solver *orc_solver_create(orc_grid *g, orc_mp *m, const orc_num_grid *num) {
  orc_solver *s = (orc_solver *)calloc(1, sizeof *s);
  s->g = g;
  s->m = m;
  if (num) s->num = *num; else orc_default_num_grid(&s->num);
  s->cref_rule = 1;
  s->N = orc_mp_ncells(m);
  size_t N9 = (size_t)9 * s->N;
  s->F = (double *)calloc(N9, sizeof(double));
  s->F_lastinc = (double *)calloc(N9, sizeof(double));
  s->Fdot = (double *)calloc(N9, sizeof(double));
  s->P = (double *)calloc(N9, sizeof(double));
  int M = s->num.ngmres_m > 0 ? s->num.ngmres_m : 1;
  s->Xh = (double *)calloc(N9 * M, sizeof(double));
  s->Fh = (double *)calloc(N9 * M, sizeof(double));
  s->XM = (double *)calloc(N9, sizeof(double));
  s->FM = (double *)calloc(N9, sizeof(double));
  s->XA = (double *)calloc(N9, sizeof(double));
  s->FA = (double *)calloc(N9, sizeof(double));
  s->R = (double *)calloc(N9, sizeof(double));
  s->D = (double *)calloc(N9, sizeof(double));
  s->Q = (double *)calloc((size_t)M * M, sizeof(double));
  /* grid_mech_spectral_basic.f90:69-82 module defaults */
  m33_eye(s->F_aim);
  m33_eye(s->F_aim_lastinc);
  s->q_R[0] = 1.0;
  for (int k = 0; k < 9; k++) s->stress_mask[k] = 1;
  /* init (grid_mech_spectral_basic.f90:213-235): F = I, response at dt = 0, C_ref */
  for (int ce = 0; ce < s->N; ce++) { m33_eye(s->F_lastinc + 9 * ce); m33_eye(s->F + 9 * ce); }
  constitutive_response(s, s->F, 0.0, 0);
  orc_update_gamma(g, s->C_minMaxAvg);
  /* DAMASK_grid.f90:78-100 */
  s->t = 0.0;
  s->Delta_t = 1.0;
  s->Delta_t_prev = 0.0;
  return s;
}

void orc_solver_destroy(orc_solver *s) {
  if (!s) return;
  free(s->F); free(s->F_lastinc); free(s->Fdot); free(s->P);
  free(s->Xh); free(s->Fh); free(s->XM); free(s->FM); free(s->XA); free(s->FA);
  free(s->R); free(s->D); free(s->Q);
  free(s->F_tau_lastinc); free(s->F_tauDot);
  free(s);
}

/* grid_mech_spectral_basic.f90:284-355 forward */
void orc_solver_forward(orc_solver *s, int cutBack, int guess, double dt, double dt_old,
                        double t_remaining, const orc_loadcase *lc) {
  int N = s->N;
  double wgt = orc_grid_wgt(s->g);
  if (cutBack) {
    memcpy(s->C_volAvg, s->C_volAvgLastInc, sizeof s->C_volAvg);
    memcpy(s->C_minMaxAvg, s->C_minMaxAvgLastInc, sizeof s->C_minMaxAvg);
  } else {
    memcpy(s->C_volAvgLastInc, s->C_volAvg, sizeof s->C_volAvg);
    memcpy(s->C_minMaxAvgLastInc, s->C_minMaxAvg, sizeof s->C_minMaxAvg);
    for (int k = 0; k < 9; k++) {
      int smask = lc->stress_type < 0 ? 1 : lc->stress_mask[k];
      s->F_aimDot[k] = guess ? (smask ? 0.0 : (s->F_aim[k] - s->F_aim_lastinc[k]) / dt_old) : 0.0;
    }
    memcpy(s->F_aim_lastinc, s->F_aim, sizeof s->F_aim);
    double V[9];
    for (int k = 0; k < 9; k++) V[k] = lc->def_mask[k] ? 0.0 : lc->def_values[k];
    if (lc->def_type == 0) {
      double LF[9];
      m33_mul(V, s->F_aim_lastinc, LF);
      for (int k = 0; k < 9; k++) s->F_aimDot[k] += LF[k];
    } else if (lc->def_type == 1) {
      for (int k = 0; k < 9; k++) s->F_aimDot[k] += V[k];
    } else {
      for (int k = 0; k < 9; k++)
        s->F_aimDot[k] += lc->def_mask[k] ? 0.0 : (lc->def_values[k] - s->F_aim_lastinc[k]) / t_remaining;
    }
    /* grid_utilities.f90:245-264 utilities_calculateRate */
    double avRate[9];
    orc_rotate33(lc->q_R, s->F_aimDot, 1, avRate);
    for (int ce = 0; ce < N; ce++)
      for (int k = 0; k < 9; k++)
        s->Fdot[9 * ce + k] = guess ? (s->F[9 * ce + k] - s->F_lastinc[9 * ce + k]) / dt_old : avRate[k];
    if (s->scheme == 1) {
      const double *F_tau = s->F + (size_t)9 * N;
      for (int ce = 0; ce < N; ce++)
        for (int k = 0; k < 9; k++)
          s->F_tauDot[9 * ce + k] = guess ? (F_tau[9 * ce + k] - s->F_tau_lastinc[9 * ce + k]) / dt_old : avRate[k];
      memcpy(s->F_tau_lastinc, F_tau, sizeof(double) * 9 * N);
    }
    memcpy(s->F_lastinc, s->F, sizeof(double) * 9 * N);
  }
  for (int k = 0; k < 9; k++) s->F_aim[k] = s->F_aim_lastinc[k] + s->F_aimDot[k] * dt;
  if (lc->stress_type == 0)
    for (int k = 0; k < 9; k++)
      s->P_aim[k] = s->P_aim[k] + (lc->stress_mask[k] ? 0.0 : (lc->stress_values[k] - s->P_aim[k]) / t_remaining) * dt;
  if (lc->stress_type == 1)
    for (int k = 0; k < 9; k++)
      s->P_aim[k] = s->P_aim[k] + (lc->stress_mask[k] ? 0.0 : lc->stress_values[k]) * dt;
  /* grid_utilities.f90:271-297 utilities_forwardTensorField */
  double aim[9], diff[9] = {0};
  orc_rotate33(lc->q_R, s->F_aim, 1, aim);
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 9; k++) s->F[9 * ce + k] = s->F_lastinc[9 * ce + k] + s->Fdot[9 * ce + k] * dt;
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 9; k++) diff[k] += s->F[9 * ce + k];
  for (int k = 0; k < 9; k++) diff[k] = diff[k] * wgt - aim[k];
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 9; k++) s->F[9 * ce + k] -= diff[k];
  /* grid_mech_spectral_polarization.f90:383-394: F_tau forwarded with its own
   * rate (no average to match), or, without a guess, rebuilt from the new F */
  if (s->scheme == 1) {
    double *F_tau = s->F + (size_t)9 * N;
    if (guess) {
      for (int ce = 0; ce < N; ce++)
        for (int k = 0; k < 9; k++) F_tau[9 * ce + k] = s->F_tau_lastinc[9 * ce + k] + s->F_tauDot[9 * ce + k] * dt;
    } else {
      for (int ce = 0; ce < N; ce++) {
        double Fl[9], Flt[9], E[9], T1[9], T2[9], T3[9];
        for (int k = 0; k < 9; k++) Fl[k] = F_tau[9 * ce + k] - s->F[9 * ce + k];
        m33_transpose(Fl, Flt);
        m33_mul(Flt, Fl, E);
        for (int i = 0; i < 3; i++) E[I2(i, i)] -= 1.0;
        t3333_mul33(s->C_scale, E, T1);
        m33_mul(Fl, T1, T2);
        for (int k = 0; k < 9; k++) T2[k] *= 0.5;
        t3333_mul33(s->S_scale, T2, T3);
        for (int i = 0; i < 3; i++) T3[I2(i, i)] += 1.0;
        for (int k = 0; k < 9; k++) F_tau[9 * ce + k] = T3[k] + s->F[9 * ce + k];
      }
    }
  }
  for (int k = 0; k < 9; k++) s->stress_mask[k] = lc->stress_type < 0 ? 1 : lc->stress_mask[k];
  memcpy(s->q_R, lc->q_R, sizeof s->q_R);
  s->dt = dt;
}

/* grid_mech_spectral_basic.f90:464-521 form_residual; F in, r out */
static void form_residual(orc_solver *s, const double *F, double *r, int iter) {
  if (s->nfuncs == 0 && iter == 0) s->totalIter = -1;
  if (s->totalIter <= iter) s->totalIter++;
  s->nfuncs++;
  constitutive_response(s, F, s->dt, 1);
  s->err_div = orc_divergence_rms(s->g, s->P);
  double dP[9], dFa[9], dFr[9];
  for (int k = 0; k < 9; k++) dP[k] = s->P_av[k] - s->P_aim[k];
  t3333_mul33(s->S, dP, dFa);
  for (int k = 0; k < 9; k++) s->F_aim[k] = s->F_aim[k] - dFa[k];
  double e = 0.0;
  for (int k = 0; k < 9; k++) {
    double v = s->stress_mask[k] ? 0.0 : fabs(dP[k]);
    if (v > e) e = v;
  }
  s->err_BC = e;
  orc_rotate33(s->q_R, dFa, 1, dFr);
  orc_gamma_convolution(s->g, s->P, dFr, r);
}

/* grid_mech_spectral_basic.f90:422-458 converged: 1 converged, -1 diverged, 0 iterating */
static int converged(const orc_solver *s) {
  double pmax = 0.0;
  for (int k = 0; k < 9; k++) if (fabs(s->P_av[k]) > pmax) pmax = fabs(s->P_av[k]);
  double divTol = pmax * s->num.eps_rel_div;
  if (divTol < s->num.eps_abs_div) divTol = s->num.eps_abs_div;
  double BCTol = pmax * s->num.eps_rel_P;
  if (BCTol < s->num.eps_abs_P) BCTol = s->num.eps_abs_P;
  if (s->totalIter >= s->num.N_iter_min && s->err_div / divTol < 1.0 && s->err_BC / BCTol < 1.0
      && s->status == ORC_STATUS_OK)
    return 1;
  if (s->totalIter >= s->num.N_iter_max) return -1;
  return 0;
}

/* grid_mech_spectral_polarization.f90:528-630 form_residual; X = [F; F_tau] in,
 * r = [r_F; r_F_tau] out */
static void form_residual_pol(orc_solver *s, const double *X, double *r, int iter) {
  int N = s->N;
  size_t n9 = (size_t)9 * N;
  const double *F = X, *F_tau = X + n9;
  double *r_F = r, *r_F_tau = r + n9;
  double wgt = orc_grid_wgt(s->g);
  for (int k = 0; k < 9; k++) s->F_av[k] = 0.0;
  for (int ce = 0; ce < N; ce++)
    for (int k = 0; k < 9; k++) s->F_av[k] += F[9 * ce + k];
  for (int k = 0; k < 9; k++) s->F_av[k] *= wgt;
  if (s->nfuncs == 0 && iter == 0) s->totalIter = -1;
  if (s->totalIter <= iter) s->totalIter++;
  s->nfuncs++;
  /* polarization field: beta C_scale:(F - I) - alpha F (C_scale:(F_tau - F - I)) */
  double *T = (double *)malloc(sizeof(double) * n9);
  for (int ce = 0; ce < N; ce++) {
    double A[9], B[9], CA[9], CB[9], FCB[9];
    for (int k = 0; k < 9; k++) {
      A[k] = F[9 * ce + k];
      B[k] = F_tau[9 * ce + k] - F[9 * ce + k];
    }
    for (int i = 0; i < 3; i++) { A[I2(i, i)] -= 1.0; B[I2(i, i)] -= 1.0; }
    t3333_mul33(s->C_scale, A, CA);
    t3333_mul33(s->C_scale, B, CB);
    m33_mul(F + 9 * ce, CB, FCB);
    for (int k = 0; k < 9; k++) T[9 * ce + k] = s->beta * CA[k] - s->alpha * FCB[k];
  }
  double aim[9], baim[9];
  for (int k = 0; k < 9; k++) baim[k] = s->beta * s->F_aim[k];
  orc_rotate33(s->q_R, baim, 1, aim);
  orc_gamma_convolution(s->g, T, aim, r_F_tau);
  for (size_t e = 0; e < n9; e++) r_F_tau[e] = s->beta * F[e] - r_F_tau[e];
  s->err_curl = orc_curl_rms(s->g, F);
  /* material response at F - r_F_tau / beta */
  for (size_t e = 0; e < n9; e++) T[e] = F[e] - r_F_tau[e] / s->beta;
  constitutive_response(s, T, s->dt, 1);
  s->err_div = orc_divergence_rms(s->g, s->P);
  const double *dPdF = orc_mp_dPdF_ptr(s->m);
  for (int ce = 0; ce < N; ce++) {
    double K[81], Ki[81], B[9], CB[9], FCB[9], V[9], W[9];
    for (int k = 0; k < 81; k++) K[k] = dPdF[(size_t)81 * ce + k] + s->C_scale[k];
    orc_inv_sym3333(K, Ki);
    for (int k = 0; k < 9; k++) B[k] = F_tau[9 * ce + k] - F[9 * ce + k];
    for (int i = 0; i < 3; i++) B[I2(i, i)] -= 1.0;
    t3333_mul33(s->C_scale, B, CB);
    m33_mul(F + 9 * ce, CB, FCB);
    for (int k = 0; k < 9; k++) V[k] = s->P[9 * ce + k] - FCB[k];
    t3333_mul33(Ki, V, W);
    for (int k = 0; k < 9; k++) r_F[9 * ce + k] = W[k] + r_F_tau[9 * ce + k];
  }
  free(T);
  /* stress BC: F_aim update, then err_BC on both kinds of components */
  double dP[9], dFa[9], Fr[9], D[9], CD[9];
  for (int k = 0; k < 9; k++) dP[k] = s->P_av[k] - s->P_aim[k];
  t3333_mul33(s->S, dP, dFa);
  for (int k = 0; k < 9; k++) s->F_aim[k] = s->F_aim[k] - dFa[k];
  orc_rotate33(s->q_R, s->F_av, 0, Fr);
  for (int k = 0; k < 9; k++) D[k] = s->F_aim[k] - Fr[k];
  t3333_mul33(s->C_scale, D, CD);
  double e = 0.0;
  for (int k = 0; k < 9; k++) {
    double v = fabs(s->stress_mask[k] ? CD[k] : dP[k]);
    if (v > e) e = v;
  }
  s->err_BC = e;
}

/* grid_mech_spectral_polarization.f90:482-522 converged */
static int converged_pol(const orc_solver *s) {
  double pmax = 0.0, fmax = 0.0;
  for (int k = 0; k < 9; k++) if (fabs(s->P_av[k]) > pmax) pmax = fabs(s->P_av[k]);
  for (int k = 0; k < 9; k++) {
    double v = fabs(s->F_aim[k] - ((k % 4) == 0 ? 1.0 : 0.0));
    if (v > fmax) fmax = v;
  }
  double curlTol = fmax * s->eps_rel_curl;
  if (curlTol < s->eps_abs_curl) curlTol = s->eps_abs_curl;
  double divTol = pmax * s->num.eps_rel_div;
  if (divTol < s->num.eps_abs_div) divTol = s->num.eps_abs_div;
  double BCTol = pmax * s->num.eps_rel_P;
  if (BCTol < s->num.eps_abs_P) BCTol = s->num.eps_abs_P;
  if (s->totalIter >= s->num.N_iter_min && s->err_div / divTol < 1.0 && s->err_curl / curlTol < 1.0
      && s->err_BC / BCTol < 1.0 && s->status == ORC_STATUS_OK)
    return 1;
  if (s->totalIter >= s->num.N_iter_max) return -1;
  return 0;
}

static double vdot(size_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s)
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* the tangent-linear build (oracle_fwd.cpp) waits for the derivative part of
 * the iteration to settle as well, and differentiates the least squares
 * solution in closed form instead of through the Jacobi sweeps */
#ifndef ORC_TANGENT_CONVERGED
#define ORC_TANGENT_CONVERGED(r, x, n) 1
#endif
#ifndef ORC_LSQ_TANGENT
#define ORC_LSQ_TANGENT(m, H, b, lam, V, cut, a)
#endif

/* symmetric least squares H a = b via Jacobi eigen-decomposition with
 * singular values below eps*max*m treated as zero (LAPACK gelss rcond = -1) */
static void lsq_sym(int m, const double *H, const double *b, double *a) {
  double A[30 * 30], V[30 * 30];
  memcpy(A, H, sizeof(double) * m * m);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) V[i + m * j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) off += A[p + m * q] * A[p + m * q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        double apq = A[p + m * q];
        if (apq == 0.0) continue;
        /* rotation t = tan(phi); for a negligible pivot t = apq / (a_qq - a_pp)
         * to first order, which also keeps theta^2 from overflowing */
        double hd = A[q + m * q] - A[p + m * p], t;
        if (fabs(apq) < 1e-150 * fabs(hd)) {
          t = apq / hd;
        } else {
          double theta = hd / (2.0 * apq);
          t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        }
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < m; k++) {
          double akp = A[k + m * p], akq = A[k + m * q];
          A[k + m * p] = c * akp - sn * akq;
          A[k + m * q] = sn * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {
          double apk = A[p + m * k], aqk = A[q + m * k];
          A[p + m * k] = c * apk - sn * aqk;
          A[q + m * k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          double vkp = V[k + m * p], vkq = V[k + m * q];
          V[k + m * p] = c * vkp - sn * vkq;
          V[k + m * q] = sn * vkp + c * vkq;
        }
      }
  }
  double smax = 0.0;
  for (int i = 0; i < m; i++) if (fabs(A[i + m * i]) > smax) smax = fabs(A[i + m * i]);
  double cut = smax * 2.220446049250313e-16 * m;
  for (int i = 0; i < m; i++) a[i] = 0.0;
  for (int k = 0; k < m; k++) {
    double lam = A[k + m * k];
    if (fabs(lam) <= cut) continue;
    double c = 0.0;
    for (int i = 0; i < m; i++) c += V[i + m * k] * b[i];
    c /= lam;
    for (int i = 0; i < m; i++) a[i] += c * V[i + m * k];
  }
  ORC_LSQ_TANGENT(m, H, b, A, V, cut, a);
}

/* Workspace of one NGMRES solve: the problem (residual and convergence test
 * through callbacks), the iterate and its residual, and the history pools.
 * The spectral solver and the affine test problem (orc_ngmres_affine) run the
 * same iteration below. */
typedef struct {
  size_t n;
  int M, max_it;
  void *ctx;
  void (*residual)(void *ctx, const double *X, double *F, int iter);
  int (*converged)(void *ctx);          /* 1 converged, -1 diverged, 0 iterating */
  double *X, *F;                        /* iterate and its residual, updated in place */
  double *Xh, *Fh, *XM, *FM, *XA, *FA, *D, *Q;
  double *trace;                        /* optional record per iteration (orc_ngmres_affine) */
  int trace_cap, n_trace;
} ngmres_ws;

static void update_subspace(ngmres_ws *w, int ivec, int l, const double *F, const double *X) {
  size_t n = w->n;
  int M = w->M;
  memcpy(w->Xh + n * ivec, X, sizeof(double) * n);
  memcpy(w->Fh + n * ivec, F, sizeof(double) * n);
  for (int i = 0; i < (l > 0 ? l : 1); i++) {
    double d = vdot(n, F, w->Fh + n * i);
    w->Q[ivec + M * i] = d;
    w->Q[i + M * ivec] = d;
  }
}

/* one trace record: l, selectA, restarted, fnorm (before), fMnorm, fAnorm,
 * alpha[M], then FM, XM, XA, FA, X (after) of n each, then Fh and Xh (M n each) */
static void trace_record(ngmres_ws *w, int l, int selectA, int restarted, double fnorm0, double fMnorm,
                         double fAnorm, const double *alpha, const double *Fh_before, const double *Xh_before) {
  if (!w->trace || w->n_trace >= w->trace_cap) return;
  size_t n = w->n, M = (size_t)w->M;
  double *r = w->trace + (size_t)w->n_trace * (6 + M + 5 * n + 2 * M * n);
  r[0] = l; r[1] = selectA; r[2] = restarted; r[3] = fnorm0; r[4] = fMnorm; r[5] = fAnorm;
  for (size_t i = 0; i < M; i++) r[6 + i] = i < (size_t)l ? alpha[i] : 0.0;
  double *v = r + 6 + M;
  memcpy(v, w->FM, sizeof(double) * n); v += n;
  memcpy(v, w->XM, sizeof(double) * n); v += n;
  memcpy(v, w->XA, sizeof(double) * n); v += n;
  memcpy(v, w->FA, sizeof(double) * n); v += n;
  memcpy(v, w->X, sizeof(double) * n); v += n;
  memcpy(v, Fh_before, sizeof(double) * M * n); v += M * n;
  memcpy(v, Xh_before, sizeof(double) * M * n);
  w->n_trace++;
}

/* PETSc SNES NGMRES (published algorithm, see the file header): returns the
 * convergence reason, > 0 converged, < 0 diverged or out of iterations */
static int ngmres_run(ngmres_ws *w) {
  size_t n = w->n;
  const int M = w->M;
  const double gammaA = 2.0, gammaC = 2.0, epsB = 0.1, deltaB = 0.9;
  const int restart_it = 2;
  double *X = w->X, *F = w->F;
  double *Fh0 = NULL, *Xh0 = NULL;
  if (w->trace) {
    Fh0 = (double *)malloc(sizeof(double) * n * M);
    Xh0 = (double *)malloc(sizeof(double) * n * M);
  }
  w->residual(w->ctx, X, F, 0);
  double fnorm = sqrt(vdot(n, F, F));
  int reason = w->converged(w->ctx);
  if (reason > 0 && !ORC_TANGENT_CONVERGED(F, X, n)) reason = 0;
  if (reason || fnorm != fnorm) {
    free(Fh0); free(Xh0);
    return reason ? reason : -1;
  }
  double fminnorm = fnorm;
  update_subspace(w, 0, 0, F, X);
  int k_restart = 1, l = 1, restart_count = 0;
  for (int k = 1; k <= w->max_it; k++) {
    int ivec = k_restart % M;
    double fnorm0 = fnorm;
    if (w->trace) {
      memcpy(Fh0, w->Fh, sizeof(double) * n * M);
      memcpy(Xh0, w->Xh, sizeof(double) * n * M);
    }
    /* x^M = x - r(x), basic line search of unit length */
    for (size_t i = 0; i < n; i++) w->XM[i] = X[i] - F[i];
    w->residual(w->ctx, w->XM, w->FM, k - 1);
    double fMnorm = sqrt(vdot(n, w->FM, w->FM));
    /* combined solution over the history */
    double xi[30], beta[30], H[30 * 30], alpha[30];
    double nu = fMnorm * fMnorm;
    int lh = l;
    for (int i = 0; i < l; i++) xi[i] = vdot(n, w->FM, w->Fh + n * i);
    for (int i = 0; i < l; i++) beta[i] = nu - xi[i];
    for (int j = 0; j < l; j++)
      for (int i = 0; i < l; i++) H[i + l * j] = w->Q[i + M * j] - xi[i] - xi[j] + nu;
    lsq_sym(l, H, beta, alpha);
    double atot = 0.0;
    for (int i = 0; i < l; i++) atot += alpha[i];
    for (size_t e = 0; e < n; e++) {
      double v = w->XM[e] * (1.0 - atot);
      for (int i = 0; i < l; i++) v += alpha[i] * w->Xh[n * i + e];
      w->XA[e] = v;
    }
    w->residual(w->ctx, w->XA, w->FA, k - 1);
    double fAnorm = sqrt(vdot(n, w->FA, w->FA));
    for (size_t e = 0; e < n; e++) w->D[e] = w->XA[e] - w->XM[e];
    double dnorm = sqrt(vdot(n, w->D, w->D));
    double dminnorm = -1.0;
    for (int i = 0; i < l; i++) {
      for (size_t e = 0; e < n; e++) w->D[e] = w->Xh[n * i + e] - w->XA[e];
      double dc = sqrt(vdot(n, w->D, w->D));
      if (dc < dminnorm || dminnorm < 0.0) dminnorm = dc;
    }
    /* difference-based selection */
    int selectA = 1;
    if (fAnorm >= gammaA * fminnorm) selectA = 0;
    if (!(epsB * dnorm < dminnorm || sqrt(fnorm) < deltaB * sqrt(fminnorm))) selectA = 0;
    if (selectA) {
      memcpy(X, w->XA, sizeof(double) * n);
      memcpy(F, w->FA, sizeof(double) * n);
      fnorm = fAnorm;
    } else {
      memcpy(X, w->XM, sizeof(double) * n);
      memcpy(F, w->FM, sizeof(double) * n);
      fnorm = fMnorm;
    }
    /* difference-based restart */
    int selectRestart = 0;
    if (epsB * dnorm > dminnorm && sqrt(fAnorm) > deltaB * sqrt(fminnorm) && l > 0) selectRestart = 1;
    if (sqrt(fAnorm) > gammaC * sqrt(fminnorm)) selectRestart = 1;
    restart_count = selectRestart ? restart_count + 1 : 0;
    int restarted = 0;
    if (restart_count >= restart_it) {
      restart_count = 0;
      k_restart = 1;
      l = 1;
      restarted = 1;
      update_subspace(w, 0, 0, w->FM, w->XM);
    } else {
      if (l < M) l++;
      k_restart++;
      if (fminnorm > fnorm) fminnorm = fnorm;
      update_subspace(w, ivec, l, F, X);
    }
    if (w->trace) trace_record(w, lh, selectA, restarted, fnorm0, fMnorm, fAnorm, alpha, Fh0, Xh0);
    reason = w->converged(w->ctx);
    if (reason > 0 && !ORC_TANGENT_CONVERGED(F, X, n)) reason = 0;
    if (reason || fnorm != fnorm) {
      free(Fh0); free(Xh0);
      return reason ? reason : -1;
    }
  }
  free(Fh0); free(Xh0);
  return -1;
}

static void spectral_residual(void *ctx, const double *X, double *F, int iter) {
  orc_solver *s = (orc_solver *)ctx;
  if (s->scheme == 1) form_residual_pol(s, X, F, iter); else form_residual(s, X, F, iter);
}

static int spectral_converged(void *ctx) {
  const orc_solver *s = (const orc_solver *)ctx;
  return s->scheme == 1 ? converged_pol(s) : converged(s);
}

static int solve_ngmres(orc_solver *s) {
  ngmres_ws w;
  memset(&w, 0, sizeof w);
  w.n = (size_t)(s->scheme == 1 ? 18 : 9) * s->N;
  w.M = s->num.ngmres_m;
  w.max_it = s->num.snes_max_it;
  w.ctx = s;
  w.residual = spectral_residual;
  w.converged = spectral_converged;
  w.X = s->F;
  w.F = s->R;
  w.Xh = s->Xh; w.Fh = s->Fh; w.XM = s->XM; w.FM = s->FM; w.XA = s->XA; w.FA = s->FA; w.D = s->D; w.Q = s->Q;
  return ngmres_run(&w);
}

/* ---- NGMRES on an affine residual r(x) = A x - b (test entry): the same
 * iteration as the spectral solver, converged when |r| <= atol, with a record
 * of every iteration in trace (see trace_record; trace_cap records of
 * 6 + M + 5 n + 2 M n doubles) */
typedef struct {
  size_t n;
  const double *A, *b;
  double atol, rnorm;
} affine_ctx;

static void affine_residual(void *ctx, const double *X, double *F, int iter) {
  (void)iter;
  affine_ctx *c = (affine_ctx *)ctx;
  double r2 = 0.0;
  for (size_t i = 0; i < c->n; i++) {
    double v = -c->b[i];
    for (size_t j = 0; j < c->n; j++) v += c->A[i + c->n * j] * X[j];
    F[i] = v;
    r2 += v * v;
  }
  c->rnorm = sqrt(r2);
}

static int affine_converged(void *ctx) {
  affine_ctx *c = (affine_ctx *)ctx;
  return c->rnorm <= c->atol ? 1 : 0;
}

int orc_ngmres_affine(int n, const double *A, const double *b, double *x, int M, int max_it, double atol,
                      double *trace, int trace_cap, int *n_trace) {
  if (n < 1 || M < 1 || M > 30) return -2;
  affine_ctx c = {(size_t)n, A, b, atol, 0.0};
  ngmres_ws w;
  memset(&w, 0, sizeof w);
  w.n = (size_t)n;
  w.M = M;
  w.max_it = max_it;
  w.ctx = &c;
  w.residual = affine_residual;
  w.converged = affine_converged;
  double *buf = (double *)calloc((size_t)n * (2 * M + 7) + (size_t)M * M, sizeof(double));
  w.X = x;
  w.F = buf;
  w.Xh = buf + n; w.Fh = w.Xh + (size_t)n * M;
  w.XM = w.Fh + (size_t)n * M; w.FM = w.XM + n; w.XA = w.FM + n; w.FA = w.XA + n; w.D = w.FA + n;
  w.Q = w.D + n;
  w.trace = trace;
  w.trace_cap = trace ? trace_cap : 0;
  int reason = ngmres_run(&w);
  if (n_trace) *n_trace = w.n_trace;
  free(buf);
  return reason;
}

static int solve_richardson(orc_solver *s) {
  size_t n = (size_t)9 * s->N;
  double *X = s->F, *F = s->R;
  for (int k = 0; k <= s->num.snes_max_it; k++) {
    form_residual(s, X, F, k);
    int reason = converged(s);
    if (reason > 0 && !ORC_TANGENT_CONVERGED(F, X, n)) reason = 0;
    if (reason) return reason;
    for (size_t i = 0; i < n; i++) X[i] = X[i] - F[i];
  }
  return -1;
}

/* grid_mech_spectral_basic.f90:243-277 solution */
int orc_solver_solution(orc_solver *s, int *its, int *nfuncs) {
  masked_compliance(s->q_R, s->stress_mask, s->C_volAvg, s->S);
  if (s->num.update_gamma) {
    orc_update_gamma(s->g, s->C_minMaxAvg);
    if (s->scheme == 1) {       /* grid_mech_spectral_polarization.f90:301-305 */
      memcpy(s->C_scale, s->C_minMaxAvg, sizeof s->C_scale);
      orc_inv_sym3333(s->C_minMaxAvg, s->S_scale);
    }
  }
  s->nfuncs = 0;
  int reason = s->num.solver == 1 ? solve_richardson(s) : solve_ngmres(s);
  if (s->tangent_once && s->status == ORC_STATUS_OK) {
    orc_mp_tangent(s->m, s->dt);
    stiffness_averages(s);
  }
  if (its) *its = s->totalIter;
  if (nfuncs) *nfuncs = s->nfuncs;
  for (int k = 0; k < 9; k++) if (s->stress_mask[k]) s->P_aim[k] = s->P_av[k];
  return reason > 0;
}

void orc_solver_update_coords(orc_solver *s, double *x_p, double *x_n) {
  orc_update_coords(s->g, s->F, x_p, x_n);
}

void orc_solver_get(const orc_solver *s, int what, double *dst) {
  size_t n = (size_t)9 * s->N;
  switch (what) {
    case 0: memcpy(dst, s->F, sizeof(double) * n); break;
    case 1: memcpy(dst, s->P, sizeof(double) * n); break;
    case 2: memcpy(dst, s->F_aim, sizeof(double) * 9); break;
    case 3: memcpy(dst, s->P_av, sizeof(double) * 9); break;
    case 4: memcpy(dst, s->C_volAvg, sizeof(double) * 81); break;
    case 5: memcpy(dst, s->C_minMaxAvg, sizeof(double) * 81); break;
    case 6: memcpy(dst, s->P_aim, sizeof(double) * 9); break;
    case 7: dst[0] = s->err_div; dst[1] = s->err_BC; break;
    case 8: memcpy(dst, s->S, sizeof(double) * 81); break;
    case 9: if (s->scheme == 1) memcpy(dst, s->F + n, sizeof(double) * n); break;
    case 10: dst[0] = s->err_curl; break;
    case 11: memcpy(dst, s->C_scale, sizeof(double) * 81); break;
    default: break;
  }
}

void orc_solver_set_cref_rule(orc_solver *s, int rule) { s->cref_rule = rule; }

/* dPdF, orientation and stiffness averages once per solve, from the last
 * response (spectral_basic only: the polarization residual reads dPdF) */
int orc_solver_set_tangent_once(orc_solver *s, int once) {
  if (once && s->scheme != 0) return 1;
  s->tangent_once = once;
  orc_mp_set_lazy_tangent(s->m, once);
  return 0;
}

/* switch a fresh solver to the spectral-polarization scheme
 * (grid_mech_spectral_polarization.f90:239-267 init): F_tau = 2 F,
 * C_scale = C_minMaxAvg, S_scale = invSym3333(C_minMaxAvg); the NGMRES
 * vectors become [F; F_tau] */
int orc_solver_set_polarization(orc_solver *s, double alpha, double beta, double eps_abs_curl, double eps_rel_curl) {
  if (s->scheme == 1) return 0;
  if (!(beta > 0.0)) return 1;     /* the residual divides by beta */
  size_t n9 = (size_t)9 * s->N, nx = 2 * n9;
  int M = s->num.ngmres_m > 0 ? s->num.ngmres_m : 1;
  double *F = (double *)calloc(nx, sizeof(double));
  memcpy(F, s->F, sizeof(double) * n9);
  for (size_t e = 0; e < n9; e++) F[n9 + e] = 2.0 * s->F[e];
  free(s->F);
  s->F = F;
  s->F_tau_lastinc = (double *)calloc(n9, sizeof(double));
  for (size_t e = 0; e < n9; e++) s->F_tau_lastinc[e] = 2.0 * s->F_lastinc[e];
  s->F_tauDot = (double *)calloc(n9, sizeof(double));
  double **v[] = {&s->XM, &s->FM, &s->XA, &s->FA, &s->R, &s->D};
  for (int i = 0; i < 6; i++) { free(*v[i]); *v[i] = (double *)calloc(nx, sizeof(double)); }
  free(s->Xh); free(s->Fh);
  s->Xh = (double *)calloc(nx * M, sizeof(double));
  s->Fh = (double *)calloc(nx * M, sizeof(double));
  s->alpha = alpha;
  s->beta = beta;
  s->eps_abs_curl = eps_abs_curl;
  s->eps_rel_curl = eps_rel_curl;
  memcpy(s->C_scale, s->C_minMaxAvg, sizeof s->C_scale);
  orc_inv_sym3333(s->C_minMaxAvg, s->S_scale);
  s->scheme = 1;
  return 0;
}
int64_t orc_solver_updates(const orc_solver *s) { return s->n_updates; }

/* math.f90:1222-1234 math_geometricFraction */
static double geometric_fraction(double r, int i, int N) {
  if (fabs(r - 1.0) <= 1.0e-9) return 1.0 / (double)N;
  return (1.0 - r) * pow(r, i) / (r - pow(r, N + 1));
}

/* DAMASK_grid.f90:288-415: one load case; stats per increment:
 * t, cutBackLevel, iterations of the accepted solve, residual evaluations */
int orc_run_loadcase(orc_solver *s, const orc_loadcase *lc, double *t_io, double *stats,
                     double *P_av_hist, double *F_aim_hist, int64_t *n_updates) {
  double t_0 = s->t;
  int guess = lc->estimate_rate;
  int64_t upd0 = s->n_updates;
  for (int inc = 1; inc <= lc->N; inc++) {
    s->totalIncs++;
    s->Delta_t_prev = s->Delta_t;
    s->Delta_t = lc->t * geometric_fraction(lc->r, inc, lc->N) * pow(2.0, -(double)s->cutBackLevel);
    int stepFraction = 0, its = 0, nf_total = 0;
    while (stepFraction < (1 << s->cutBackLevel)) {
      double t_remaining = lc->t + t_0 - s->t;
      s->t = s->t + s->Delta_t;
      stepFraction++;
      orc_solver_forward(s, s->cutBack, guess, s->Delta_t, s->Delta_t_prev, t_remaining, lc);
      if (!s->cutBack) orc_mp_forward(s->m);
      int nf;
      int conv = orc_solver_solution(s, &its, &nf);
      nf_total += nf;
      if (conv) {
        orc_update_coords(s->g, s->F, NULL, NULL);
        s->Delta_t_prev = s->Delta_t;
        s->cutBack = 0;
        guess = 1;
      } else if (s->cutBackLevel < s->num.N_cutback_max) {
        s->cutBack = 1;
        stepFraction = (stepFraction - 1) * 2;
        s->cutBackLevel++;
        s->t = s->t - s->Delta_t;
        s->Delta_t = s->Delta_t / 2.0;
      } else {
        if (t_io) *t_io = s->t;
        if (n_updates) *n_updates = s->n_updates - upd0;
        return 950;
      }
    }
    if (stats) {
      stats[4 * (inc - 1) + 0] = s->t;
      stats[4 * (inc - 1) + 1] = s->cutBackLevel;
      stats[4 * (inc - 1) + 2] = its;
      stats[4 * (inc - 1) + 3] = nf_total;
    }
    s->cutBackLevel = s->cutBackLevel - 1 > 0 ? s->cutBackLevel - 1 : 0;
    if (P_av_hist) memcpy(P_av_hist + 9 * (inc - 1), s->P_av, sizeof(double) * 9);
    if (F_aim_hist) memcpy(F_aim_hist + 9 * (inc - 1), s->F_aim, sizeof(double) * 9);
  }
  if (t_io) *t_io = s->t;
  if (n_updates) *n_updates = s->n_updates - upd0;
  return 0;
}
