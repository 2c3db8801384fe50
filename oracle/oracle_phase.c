/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of the per-material-point constitutive update: Hooke's law
 * (src/phase_mechanical_elastic.f90), phenopowerlaw kinetics and hardening
 * (src/phase_mechanical_plastic_phenopowerlaw.f90), the plastic dispatcher
 * (src/phase_mechanical_plastic.f90), stress integration, the fixed-point state
 * integrator, substepping and the consistent tangent (src/phase_mechanical.f90)
 * and the isostrain/pass homogenization loop (src/homogenization.f90:294-362).
 * Written with general F_i as the reference does; the eigen kinematics return
 * L_i = 0 with zero tangents (src/phase_mechanical_eigen.f90:134-165).
 */
#include "oracle_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_default_num_mech(orc_num_mech *n) {
  /* phase_mechanical.f90:281-295 */
  n->r_cutback_min = 1.0e-3;
  n->r_cutback = 0.25;
  n->r_increase = 1.5;
  n->eps_rel_state = 1.0e-6;
  n->N_iter_state_max = 20;
  n->N_iter_Lp_max = 40;
  n->r_linesearch_Lp = 0.5;
  n->eps_rel_Lp = 1.0e-6;
  n->eps_abs_Lp = 1.0e-8;
  n->f_update_jacobi_Lp = 1;
  n->N_iter_Li_max = 40;
  n->r_linesearch_Li = 0.5;
  n->eps_rel_Li = n->eps_rel_Lp;
  n->eps_abs_Li = n->eps_abs_Lp;
  n->f_update_jacobi_Li = 1;
}

/* phenopowerlaw.f90:146-285 + elastic.f90:58-79: derived tables of one phase */
int orc_phase_build(const orc_phase_desc *d, orc_phase *p) {
  memset(p, 0, sizeof *p);
  int nsl = 0;
  for (int f = 0; f < d->n_fam; f++) nsl += d->N_sl[f];
  if (nsl <= 0 || nsl > ORC_NSL_MAX) return 1;
  p->nsl = nsl;
  orc_cubic_stiffness(d->C_11, d->C_12, d->C_44, p->C66);
  m66_to_sym3333(p->C66, p->C3333);
  orc_schmid_slip(d->lattice, d->n_fam, d->N_sl, 1, 0, NULL, &p->P_sl[0][0]);
  orc_schmid_slip(d->lattice, d->n_fam, d->N_sl, +1, d->n_nS, d->a_nS, &p->P_nS_pos[0][0]);
  orc_schmid_slip(d->lattice, d->n_fam, d->N_sl, -1, d->n_nS, d->a_nS, &p->P_nS_neg[0][0]);
  orc_interaction_slipslip(d->lattice, d->n_fam, d->N_sl, d->n_h, d->h_sl_sl, p->H);
  int a = 0;
  for (int f = 0; f < d->n_fam; f++)
    for (int s = 0; s < d->N_sl[f]; s++, a++) {
      p->dot_gamma_0[a] = d->dot_gamma_0[f];
      p->n[a] = d->n[f];
      p->a[a] = d->a[f];
      p->h_0[a] = d->h_0[f];
      p->xi_0[a] = d->xi_0[f];
      p->xi_inf[a] = d->xi_inf[f];
      p->c_1[a] = d->c_1[f];
      p->c_2[a] = d->c_2[f];
      p->f_sat[a] = d->f_sat[f];
    }
  for (int i = 0; i < nsl; i++) {
    p->atol[i] = d->atol_xi;
    p->atol[nsl + i] = d->atol_gamma;
  }
  for (int i = 0; i < nsl * nsl; i++)
    if (isnan(p->H[i])) return 2;
  return 0;
}

/* ------------------------------------------------------------------ elastic */
/* elastic.f90:172-205 phase_hooke_SandItsTangents */
void orc_hooke(const orc_phase *p, const double Fe[9], const double Fi[9],
               double S[9], double dS_dFe[81], double dS_dFi[81]) {
  const double *C = p->C3333;
  double Fet[9], E[9], Fit[9], T1[9], T2[9], v6[6], s6[6];
  m33_transpose(Fe, Fet);
  m33_mul(Fet, Fe, E);
  E[0] -= 1.0; E[4] -= 1.0; E[8] -= 1.0;
  for (int k = 0; k < 9; k++) E[k] *= 0.5;
  m33_transpose(Fi, Fit);
  m33_mul(Fit, E, T1);
  m33_mul(T1, Fi, T2);
  m33_sym_to6(T2, v6);
  for (int r = 0; r < 6; r++) {
    double s = 0.0;
    for (int c = 0; c < 6; c++) s += p->C66[r][c] * v6[c];
    s6[r] = s;
  }
  m6_to_sym33(s6, S);
  double EFi[9];
  m33_mul(E, Fi, EFi);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double Cij[9], A[9], B[9], R[9];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) Cij[I2(k, l)] = C[I4(i, j, k, l)];
      m33_mul(Fi, Cij, A);
      m33_mul(A, Fit, B);
      m33_mul(Fe, B, R);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) dS_dFe[I4(i, j, k, l)] = R[I2(k, l)];
      if (dS_dFi) {
        m33_mul(EFi, Cij, R);
        for (int l = 0; l < 3; l++)
          for (int k = 0; k < 3; k++) dS_dFi[I4(i, j, k, l)] = 2.0 * R[I2(k, l)];
      }
    }
}

/* ------------------------------------------------------------------ plastic */
static double sum33(const double *A, const double *B) {   /* sum(A*B), column order */
  double s = 0.0;
  for (int k = 0; k < 9; k++) s += A[k] * B[k];
  return s;
}

/* phenopowerlaw.f90:449-488 kinetics_sl */
void orc_kinetics_sl(const orc_phase *p, const double Mp[9], const double *xi,
                     double *dot_gamma, double *ddot_gamma_dtau) {
  for (int i = 0; i < p->nsl; i++) {
    double tp = sum33(Mp, p->P_nS_pos[i]);
    double tn = sum33(Mp, p->P_nS_neg[i]);
    double sgn = tp > tn ? 1.0 : -1.0;
    double tmax = tp > tn ? tp : tn;
    if (tn != tn) tmax = tn;       /* Fortran max() with a NaN operand: NaN propagates here */
    if (tp != tp) tmax = tp;
    dot_gamma[i] = sgn * p->dot_gamma_0[i] * pow(tmax / xi[i], p->n[i]);
    if (ddot_gamma_dtau)
      ddot_gamma_dtau[i] = !orc_dEq0(dot_gamma[i]) ? dot_gamma[i] * p->n[i] / tmax : 0.0;
  }
}

/* phenopowerlaw.f90:303-350 phenopowerlaw_LpAndItsTangent */
static void pheno_Lp(const orc_phase *p, const double Mp[9], const double *xi, double Lp[9],
                     double dLp_dMp[81]) {
  double dg[ORC_NSL_MAX], ddg[ORC_NSL_MAX];
  orc_kinetics_sl(p, Mp, xi, dg, ddg);
  memset(Lp, 0, 9 * sizeof(double));
  memset(dLp_dMp, 0, 81 * sizeof(double));
  for (int i = 0; i < p->nsl; i++) {
    const double *Pn = dg[i] > 0.0 ? p->P_nS_pos[i] : p->P_nS_neg[i];
    for (int k = 0; k < 9; k++) Lp[k] += dg[i] * p->P_sl[i][k];
    for (int n = 0; n < 3; n++)
      for (int m = 0; m < 3; m++)
        for (int l = 0; l < 3; l++)
          for (int k = 0; k < 3; k++)
            dLp_dMp[I4(k, l, m, n)] += ddg[i] * p->P_sl[i][I2(k, l)] * Pn[I2(m, n)];
  }
}

/* phase_mechanical_plastic.f90:239-298 plastic_LpAndItsTangents */
static void plastic_Lp(const orc_phase *p, const double S[9], const double Fi[9], const double *xi,
                       double Lp[9], double dLp_dS[81], double dLp_dFi[81]) {
  double Fit[9], FitFi[9], Mp[9], dLp_dMp[81];
  m33_transpose(Fi, Fit);
  m33_mul(Fit, Fi, FitFi);
  m33_mul(FitFi, S, Mp);
  pheno_Lp(p, Mp, xi, Lp, dLp_dMp);
  double FiS[9];
  m33_mul(Fi, S, FiS);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double D[9], Dt[9], T1[9], T2[9], T3[9];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) D[I2(k, l)] = dLp_dMp[I4(i, j, k, l)];
      if (dLp_dFi) {
        m33_transpose(D, Dt);
        m33_mul(FiS, Dt, T1);
        m33_mul(Fi, D, T2);
        m33_mul(T2, S, T3);
        for (int l = 0; l < 3; l++)
          for (int k = 0; k < 3; k++) dLp_dFi[I4(i, j, k, l)] = T1[I2(k, l)] + T3[I2(k, l)];
      }
      m33_mul(FitFi, D, T1);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) dLp_dS[I4(i, j, k, l)] = T1[I2(k, l)];
    }
}

/* phenopowerlaw.f90:356-399 phenopowerlaw_dotState (no twins: sumF = 0) */
static void pheno_dotState(const orc_phase *p, const double Mp[9], const double *state,
                           double *dotState) {
  int nsl = p->nsl;
  double dg[ORC_NSL_MAX];
  orc_kinetics_sl(p, Mp, state, dg, NULL);
  for (int i = 0; i < nsl; i++) dg[i] = fabs(dg[i]);
  const double sumF = 0.0;
  for (int r = 0; r < nsl; r++) {
    double off = p->f_sat[r] * sqrt(sumF);
    double x = 1.0 - state[r] / (p->xi_inf[r] + off);
    double left = copysign(pow(fabs(x), p->a[r]), x);
    double hg = 0.0;
    for (int a = 0; a < nsl; a++) hg += p->H[r * nsl + a] * dg[a];
    dotState[r] = p->h_0[r] * (1.0 + p->c_1[r] * pow(sumF, p->c_2[r])) * left * hg + 0.0;
    dotState[nsl + r] = dg[r];
  }
}

/* ------------------------------------------------------------------ material point */
struct orc_mp {
  int N, n_phases, stride;   /* stride = 2 * max nsl */
  orc_phase *ph;
  int32_t *phase_of;
  orc_num_mech num;
  double *F, *F0, *Fp, *Fp0, *Fi, *Fi0, *Lp, *Lp0, *Li, *Li0, *S, *S0, *P, *Fe, *O;
  double *state, *state0;
  double *hom_P, *hom_dPdF;
  int32_t *status;
  int64_t *newton, *fpi, *cut, *nsub;
  int lazy_tangent;          /* 1: orientation and dPdF only on request (orc_mp_tangent) */
};

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

orc_mp *orc_mp_create(int n_cells, int n_phases, const orc_phase *phases,
                      const int32_t *phase_of_cell, const double *q0, const orc_num_mech *num) {
  orc_mp *m = (orc_mp *)calloc(1, sizeof *m);
  m->N = n_cells;
  m->n_phases = n_phases;
  m->ph = (orc_phase *)malloc(sizeof(orc_phase) * n_phases);
  memcpy(m->ph, phases, sizeof(orc_phase) * n_phases);
  int maxn = 0;
  for (int i = 0; i < n_phases; i++) if (phases[i].nsl > maxn) maxn = phases[i].nsl;
  m->stride = 2 * maxn;
  m->phase_of = (int32_t *)malloc(sizeof(int32_t) * n_cells);
  memcpy(m->phase_of, phase_of_cell, sizeof(int32_t) * n_cells);
  if (num) m->num = *num; else orc_default_num_mech(&m->num);
  size_t N9 = (size_t)9 * n_cells;
  m->F = dalloc(N9); m->F0 = dalloc(N9); m->Fp = dalloc(N9); m->Fp0 = dalloc(N9);
  m->Fi = dalloc(N9); m->Fi0 = dalloc(N9); m->Lp = dalloc(N9); m->Lp0 = dalloc(N9);
  m->Li = dalloc(N9); m->Li0 = dalloc(N9); m->S = dalloc(N9); m->S0 = dalloc(N9);
  m->P = dalloc(N9); m->Fe = dalloc(N9); m->O = dalloc((size_t)4 * n_cells);
  m->state = dalloc((size_t)m->stride * n_cells);
  m->state0 = dalloc((size_t)m->stride * n_cells);
  m->hom_P = dalloc(N9);
  m->hom_dPdF = dalloc((size_t)81 * n_cells);
  m->status = (int32_t *)calloc(n_cells, sizeof(int32_t));
  m->newton = (int64_t *)calloc(n_cells, sizeof(int64_t));
  m->fpi = (int64_t *)calloc(n_cells, sizeof(int64_t));
  m->cut = (int64_t *)calloc(n_cells, sizeof(int64_t));
  m->nsub = (int64_t *)calloc(n_cells, sizeof(int64_t));
  /* phase_mechanical.f90:249-266 with V_e = I */
  for (int ce = 0; ce < n_cells; ce++) {
    double *F = m->F + 9 * ce, *Fp = m->Fp + 9 * ce, I[9];
    m33_eye(F);
    orc_qu2om(q0 + 4 * ce, Fp);
    m33_transpose(Fp, m->Fe + 9 * ce);
    m33_eye(I);
    double Iinv[9];
    m33_inv(I, Iinv);
    orc_rotate33(q0 + 4 * ce, Iinv, 0, m->Fi + 9 * ce);
    memcpy(m->O + 4 * ce, q0 + 4 * ce, 4 * sizeof(double));
    const orc_phase *p = &m->ph[m->phase_of[ce]];
    double *st = m->state + (size_t)m->stride * ce;
    for (int i = 0; i < p->nsl; i++) { st[i] = p->xi_0[i]; st[p->nsl + i] = 0.0; }
  }
  memcpy(m->F0, m->F, N9 * sizeof(double));
  memcpy(m->Fp0, m->Fp, N9 * sizeof(double));
  memcpy(m->Fi0, m->Fi, N9 * sizeof(double));
  memcpy(m->state0, m->state, (size_t)m->stride * n_cells * sizeof(double));
  return m;
}

void orc_mp_destroy(orc_mp *m) {
  if (!m) return;
  free(m->ph); free(m->phase_of);
  free(m->F); free(m->F0); free(m->Fp); free(m->Fp0); free(m->Fi); free(m->Fi0);
  free(m->Lp); free(m->Lp0); free(m->Li); free(m->Li0); free(m->S); free(m->S0);
  free(m->P); free(m->Fe); free(m->O); free(m->state); free(m->state0);
  free(m->hom_P); free(m->hom_dPdF); free(m->status); free(m->newton); free(m->fpi); free(m->cut);
  free(m->nsub);
  free(m);
}

int orc_mp_nsl(const orc_mp *m) { return m->stride / 2; }

/* phase_mechanical.f90:379-584 integrateStress */
int orc_integrate_stress(orc_mp *m, int ce, const double F[9], const double Fp0[9],
                         const double Fi0[9], double dt) {
  const orc_phase *p = &m->ph[m->phase_of[ce]];
  const orc_num_mech *num = &m->num;
  const double *xi = m->state + (size_t)m->stride * ce;
  int status = ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS;
  double Lpguess[9], Liguess[9], invFp_current[9], invFi_current[9], A[9];
  memcpy(Lpguess, m->Lp + 9 * ce, sizeof Lpguess);
  memcpy(Liguess, m->Li + 9 * ce, sizeof Liguess);
  if (m33_invert(Fp0, invFp_current, NULL)) return status;
  if (m33_invert(Fi0, invFi_current, NULL)) return status;
  m33_mul(F, invFp_current, A);

  double steplengthLi = 1.0, Liguess_old[9], deltaLi[9] = {0};
  double resLi_old = DBL_MAX;
  int jacoLi = 0, nLi = 0;
  memcpy(Liguess_old, Liguess, sizeof Liguess);
  double B[9], S[9], Fi_new[9], invFi_new[9];
  for (;;) {
    nLi++;
    if (nLi > num->N_iter_Li_max) return status;
    double T[9];
    for (int k = 0; k < 9; k++) T[k] = (k % 4 == 0 ? 1.0 : 0.0) - dt * Liguess[k];
    m33_mul(invFi_current, T, invFi_new);
    m33_inv(invFi_new, Fi_new);

    int jacoLp = 0, nLp = 0;
    double steplengthLp = 1.0, Lpguess_old[9], deltaLp[9] = {0}, resLp_old = DBL_MAX;
    memcpy(Lpguess_old, Lpguess, sizeof Lpguess);
    for (;;) {
      nLp++;
      if (nLp > num->N_iter_Lp_max) return ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_MAXIT;
      m->newton[ce]++;
      for (int k = 0; k < 9; k++) B[k] = (k % 4 == 0 ? 1.0 : 0.0) - dt * Lpguess[k];
      double AB[9], Fe[9], dS_dFe[81], dS_dFi[81], Lp_c[9], dLp_dS[81], dLp_dFi[81];
      m33_mul(A, B, AB);
      m33_mul(AB, invFi_new, Fe);
      orc_hooke(p, Fe, Fi_new, S, dS_dFe, dS_dFi);
      plastic_Lp(p, S, Fi_new, xi, Lp_c, dLp_dS, dLp_dFi);
      double ng = m33_norm2(Lpguess), nc = m33_norm2(Lp_c);
      double atol = num->eps_rel_Lp * (ng > nc ? ng : nc);
      if (atol < num->eps_abs_Lp) atol = num->eps_abs_Lp;
      double res[9];
      int nan = 0;
      for (int k = 0; k < 9; k++) { res[k] = Lpguess[k] - Lp_c[k]; if (res[k] != res[k]) nan = 1; }
      double nres = m33_norm2(res);
      if (nan) return ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_NAN;
      else if (nres < atol) break;
      else if (nres < resLp_old) {
        resLp_old = nres;
        memcpy(Lpguess_old, Lpguess, sizeof Lpguess);
        steplengthLp = 1.0;
      } else {
        steplengthLp = num->r_linesearch_Lp * steplengthLp;
        for (int k = 0; k < 9; k++) Lpguess[k] = Lpguess_old[k] + deltaLp[k] * steplengthLp;
        continue;
      }
      if (jacoLp % num->f_update_jacobi_Lp == 0) {
        jacoLp++;
        double dFe_dLp[81], X[81], Y[81], J[81], t9[9];
        for (int o = 0; o < 3; o++)
          for (int pp = 0; pp < 3; pp++)
            for (int j = 0; j < 3; j++)
              for (int l = 0; l < 3; l++)
                dFe_dLp[I4(o, j, pp, l)] = -dt * A[I2(o, pp)] * invFi_new[I2(l, j)];
        t3333_mul(dLp_dS, dS_dFe, X);
        t3333_mul(X, dFe_dLp, Y);
        t3333_to99(Y, J);
        for (int r = 0; r < 9; r++)
          for (int c = 0; c < 9; c++) J[r + 9 * c] = (r == c ? 1.0 : 0.0) - J[r + 9 * c];
        m33_to9(res, t9);
        if (dgesv_n(9, J, t9)) return ORC_STATUS_FAIL_PHASE_MECHANICAL_STRESS_LP_DGESV;
        double d33[9];
        m9_to33(t9, d33);
        for (int k = 0; k < 9; k++) deltaLp[k] = -d33[k];
      }
      for (int k = 0; k < 9; k++) Lpguess[k] = Lpguess[k] + deltaLp[k] * steplengthLp;
    }

    /* eigen kinematics absent: Li_constitutive = 0, all tangents 0 */
    double ngi = m33_norm2(Liguess);
    double atolLi = num->eps_rel_Li * ngi;
    if (atolLi < num->eps_abs_Li) atolLi = num->eps_abs_Li;
    double resLi[9];
    int nan = 0;
    for (int k = 0; k < 9; k++) { resLi[k] = Liguess[k] - 0.0; if (resLi[k] != resLi[k]) nan = 1; }
    double nres = m33_norm2(resLi);
    if (nan) return status;
    else if (nres < atolLi) break;
    else if (nres < resLi_old) {
      resLi_old = nres;
      memcpy(Liguess_old, Liguess, sizeof Liguess);
      steplengthLi = 1.0;
    } else {
      steplengthLi = num->r_linesearch_Li * steplengthLi;
      for (int k = 0; k < 9; k++) Liguess[k] = Liguess_old[k] + deltaLi[k] * steplengthLi;
      continue;
    }
    if (jacoLi % num->f_update_jacobi_Li == 0) {
      jacoLi++;
      /* dRLi_dLi = I9 since every Li tangent vanishes */
      for (int k = 0; k < 9; k++) deltaLi[k] = -resLi[k];
    }
    for (int k = 0; k < 9; k++) Liguess[k] = Liguess[k] + deltaLi[k] * steplengthLi;
  }

  double invFp_new[9], Fp_new[9];
  m33_mul(invFp_current, B, invFp_new);
  if (m33_invert(invFp_new, Fp_new, NULL)) return status;
  double T1[9], T2[9], T3[9], invFpt[9];
  m33_mul(F, invFp_new, T1);
  m33_transpose(invFp_new, invFpt);
  m33_mul(S, invFpt, T2);
  m33_mul(T1, T2, T3);
  memcpy(m->P + 9 * ce, T3, sizeof T3);
  memcpy(m->S + 9 * ce, S, sizeof S);
  memcpy(m->Lp + 9 * ce, Lpguess, sizeof Lpguess);
  memcpy(m->Li + 9 * ce, Liguess, sizeof Liguess);
  double dfp = pow(m33_det(Fp_new), 1.0 / 3.0);
  for (int k = 0; k < 9; k++) m->Fp[9 * ce + k] = Fp_new[k] / dfp;
  memcpy(m->Fi + 9 * ce, Fi_new, sizeof Fi_new);
  m33_mul(T1, invFi_new, m->Fe + 9 * ce);
  return ORC_STATUS_OK;
}

/* phase_mechanical_plastic.f90:304-345 plastic_dotState */
static int dot_state(orc_mp *m, int ce, double *dotState) {
  const orc_phase *p = &m->ph[m->phase_of[ce]];
  double Fit[9], FitFi[9], Mp[9];
  m33_transpose(m->Fi + 9 * ce, Fit);
  m33_mul(Fit, m->Fi + 9 * ce, FitFi);
  m33_mul(FitFi, m->S + 9 * ce, Mp);
  pheno_dotState(p, Mp, m->state + (size_t)m->stride * ce, dotState);
  for (int i = 0; i < 2 * p->nsl; i++)
    if (dotState[i] != dotState[i]) return 1;
  return 0;
}

/* phase_mechanical.f90:653-674 damper */
static double damper(int n, const double *w0, const double *w1, const double *w2) {
  double d12 = 0.0, d22 = 0.0, d01 = 0.0;
  for (int i = 0; i < n; i++) d12 += (w0[i] - w1[i]) * (w1[i] - w2[i]);
  for (int i = 0; i < n; i++) d22 += (w1[i] - w2[i]) * (w1[i] - w2[i]);
  for (int i = 0; i < n; i++) d01 += w0[i] * w1[i];
  double mn = d01 < d12 ? d01 : d12;
  if (mn < 0.0 && d22 > 0.0) return 0.75 + 0.25 * tanh(2.0 + 4.0 * d12 / d22);
  return 1.0;
}

/* phase_mechanical.f90:591-676 integrateStateFPI */
static int integrate_state_fpi(orc_mp *m, int ce, const double F[9], const double Fp0[9],
                               const double Fi0[9], const double *state0, double dt) {
  const orc_phase *p = &m->ph[m->phase_of[ce]];
  int n = 2 * p->nsl;
  double *state = m->state + (size_t)m->stride * ce;
  double dotState[2 * ORC_NSL_MAX], last1[2 * ORC_NSL_MAX], last2[2 * ORC_NSL_MAX], r[2 * ORC_NSL_MAX];
  int status = ORC_STATUS_FAIL_PHASE_MECHANICAL_STATE;
  if (dot_state(m, ce, dotState)) return status;
  for (int i = 0; i < n; i++) state[i] = state0[i] + dotState[i] * dt;
  for (int it = 1; it <= m->num.N_iter_state_max; it++) {
    for (int i = 0; i < n; i++) last2[i] = it > 1 ? last1[i] : 0.0;
    for (int i = 0; i < n; i++) last1[i] = dotState[i];
    m->fpi[ce]++;
    status = orc_integrate_stress(m, ce, F, Fp0, Fi0, dt);
    if (status != ORC_STATUS_OK) break;
    if (dot_state(m, ce, dotState)) break;
    double zeta = damper(n, dotState, last1, last2);
    for (int i = 0; i < n; i++) dotState[i] = dotState[i] * zeta + last1[i] * (1.0 - zeta);
    for (int i = 0; i < n; i++) r[i] = state[i] - state0[i] - dotState[i] * dt;
    for (int i = 0; i < n; i++) state[i] = state[i] - r[i];
    int conv = 1;
    for (int i = 0; i < n; i++) {
      double tol = m->num.eps_rel_state * fabs(state[i]);
      if (p->atol[i] > tol) tol = p->atol[i];
      if (!(fabs(r[i]) <= tol)) { conv = 0; break; }
    }
    if (conv) { status = ORC_STATUS_OK; break; }
  }
  return status;
}

/* phase_mechanical.f90:1008-1090 phase_mechanical_constitutive */
static int constitutive(orc_mp *m, int ce, double dt) {
  const orc_num_mech *num = &m->num;
  int ns = m->stride;
  double state0[2 * ORC_NSL_MAX], Li0[9], Lp0[9], Fp0[9], Fi0[9], F0[9], F[9] = {0};
  memcpy(state0, m->state0 + (size_t)ns * ce, sizeof(double) * ns);
  memcpy(Li0, m->Li0 + 9 * ce, sizeof Li0);
  memcpy(Lp0, m->Lp0 + 9 * ce, sizeof Lp0);
  memcpy(Fp0, m->Fp0 + 9 * ce, sizeof Fp0);
  memcpy(Fi0, m->Fi0 + 9 * ce, sizeof Fi0);
  memcpy(F0, m->F0 + 9 * ce, sizeof F0);
  double stepFrac = 0.0, step = 1.0 / num->r_cutback;
  int status = ORC_STATUS_ITERATING, todo = 1;
  while (todo) {
    if (status == ORC_STATUS_OK) {
      stepFrac = stepFrac + step;
      double a = 1.0 - stepFrac, b = num->r_increase * step;
      step = a < b ? a : b;
      todo = step > 0.0;
      if (todo) {
        memcpy(F0, F, sizeof F);
        memcpy(Lp0, m->Lp + 9 * ce, sizeof Lp0);
        memcpy(Li0, m->Li + 9 * ce, sizeof Li0);
        memcpy(Fp0, m->Fp + 9 * ce, sizeof Fp0);
        memcpy(Fi0, m->Fi + 9 * ce, sizeof Fi0);
        memcpy(state0, m->state + (size_t)ns * ce, sizeof(double) * ns);
      }
    } else {
      step = num->r_cutback * step;
      memcpy(m->Fp + 9 * ce, Fp0, sizeof Fp0);
      memcpy(m->Fi + 9 * ce, Fi0, sizeof Fi0);
      memcpy(m->S + 9 * ce, m->S0 + 9 * ce, 9 * sizeof(double));
      if (step < 1.0) {
        memcpy(m->Lp + 9 * ce, Lp0, sizeof Lp0);
        memcpy(m->Li + 9 * ce, Li0, sizeof Li0);
        m->cut[ce]++;
      }
      memcpy(m->state + (size_t)ns * ce, state0, sizeof(double) * ns);
      todo = step > num->r_cutback_min;
    }
    if (todo) {
      for (int k = 0; k < 9; k++) F[k] = F0[k] + step * (m->F[9 * ce + k] - m->F0[9 * ce + k]);
      status = integrate_state_fpi(m, ce, F, Fp0, Fi0, state0, step * dt);
      m->nsub[ce]++;
    }
  }
  return status;
}

/* phase_mechanical.f90:1121-1242 phase_mechanical_dPdF (dLi_dS = 0 branch) */
void orc_dPdF(const orc_mp *m, int ce, double dt, double out[81]) {
  const orc_phase *p = &m->ph[m->phase_of[ce]];
  const double *Fe = m->Fe + 9 * ce, *Fi = m->Fi + 9 * ce, *S = m->S + 9 * ce;
  const double *F = m->F + 9 * ce, *xi = m->state + (size_t)m->stride * ce;
  double dSdFe[81], dSdFi[81], dev[9], dLpdS[81], dLpdFi[81];
  double dFidS[81] = {0}, dLidS[81] = {0};
  double S_[9];
  orc_hooke(p, Fe, Fi, S_, dSdFe, dSdFi);
  double invFp[9], invFi[9], invSubFp0[9], invSubFi0[9];
  m33_inv(m->Fp + 9 * ce, invFp);
  m33_inv(Fi, invFi);
  m33_inv(m->Fp0 + 9 * ce, invSubFp0);
  m33_inv(m->Fi0 + 9 * ce, invSubFi0);
  plastic_Lp(p, S, Fi, xi, dev, dLpdS, dLpdFi);
  double T[81];
  t3333_mul(dLpdFi, dFidS, T);
  for (int k = 0; k < 81; k++) dLpdS[k] = T[k] + dLpdS[k];

  double t1[9], t2[9], t3[9], M[9];
  m33_mul(invFp, invFi, M);
  m33_transpose(M, t1);
  m33_mul(F, invSubFp0, t2);
  m33_mul(F, invFp, M);
  m33_mul(M, invSubFi0, t3);
  double rhs[81], tmp[81], lhs[81];
  for (int o = 0; o < 3; o++)
    for (int pp = 0; pp < 3; pp++) {
      double D[9], R[9], L1[9], L2[9];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) D[I2(k, l)] = dSdFe[I4(pp, o, k, l)];
      m33_mul(D, t1, R);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) rhs[I4(pp, o, k, l)] = R[I2(k, l)];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) D[I2(k, l)] = dLpdS[I4(k, l, pp, o)];
      m33_mul(t2, D, L1);
      m33_mul(L1, invFi, L2);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) D[I2(k, l)] = dLidS[I4(k, l, pp, o)];
      m33_mul(t3, D, L1);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) tmp[I4(k, l, pp, o)] = L2[I2(k, l)] + L1[I2(k, l)];
    }
  double A1[81], A2[81];
  t3333_mul(dSdFe, tmp, A1);
  t3333_mul(dSdFi, dFidS, A2);
  for (int k = 0; k < 81; k++) lhs[k] = A1[k] * dt + A2[k];
  double K[81], Kinv[81], dSdF[81];
  t3333_to99(lhs, K);
  for (int r = 0; r < 9; r++) K[r + 9 * r] += 1.0;
  /* math_eye(9)+lhs: eye added first in the reference; addition is commutative */
  if (orc_invert(9, K, Kinv)) {
    memcpy(dSdF, rhs, sizeof rhs);
  } else {
    double K4[81];
    t99_to3333(Kinv, K4);
    t3333_mul(K4, rhs, dSdF);
  }
  double dFpinvdF[81];
  t3333_mul(dLpdS, dSdF, tmp);
  for (int o = 0; o < 3; o++)
    for (int pp = 0; pp < 3; pp++) {
      double D[9], R1[9], R2[9];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) D[I2(k, l)] = tmp[I4(k, l, pp, o)];
      m33_mul(D, invFi, R1);
      m33_mul(invSubFp0, R1, R2);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) dFpinvdF[I4(k, l, pp, o)] = -R2[I2(k, l)] * dt;
    }
  double invFpt[9], u1[9], u2[9], u3[9], W[9], Wt[9];
  m33_transpose(invFp, invFpt);
  m33_mul(S, invFpt, u1);
  m33_mul(F, invFp, u2);
  m33_mul(u2, S, u3);
  memset(out, 0, 81 * sizeof(double));
  m33_mul(invFp, u1, W);
  m33_transpose(W, Wt);
  for (int pp = 0; pp < 3; pp++)
    for (int j = 0; j < 3; j++)
      for (int l = 0; l < 3; l++) out[I4(pp, j, pp, l)] = Wt[I2(j, l)];
  for (int o = 0; o < 3; o++)
    for (int pp = 0; pp < 3; pp++) {
      double D[9], G[9], R1[9], R2[9], R3[9], R4[9], Dt[9];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) { D[I2(k, l)] = dFpinvdF[I4(k, l, pp, o)]; G[I2(k, l)] = dSdF[I4(k, l, pp, o)]; }
      m33_mul(F, D, R1);
      m33_mul(R1, u1, R2);
      m33_mul(u2, G, R3);
      m33_mul(R3, invFpt, R4);
      m33_transpose(D, Dt);
      m33_mul(u3, Dt, R1);
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++)
          out[I4(k, l, pp, o)] = out[I4(k, l, pp, o)] + R2[I2(k, l)] + R4[I2(k, l)] + R1[I2(k, l)];
    }
}

/* phase.f90:659-673 crystallite_orientations */
static void orientation(orc_mp *m, int ce) {
  double R[9], Rt[9];
  orc_rotational_part(m->Fe + 9 * ce, R);
  m33_transpose(R, Rt);
  orc_om2qu(Rt, m->O + 4 * ce);
}

/* homogenization.f90:294-362 homogenization_mechanical_response (isostrain/pass,
 * one constituent, v = 1) plus the dPdF/P copies of utilities_constitutiveResponse */
int orc_mp_response(orc_mp *m, double dt, const double *F_target, double *P_out, double *dPdF_out) {
  int N = m->N;
  int ns = m->stride;
#pragma omp parallel for schedule(dynamic, 64)
  for (int ce = 0; ce < N; ce++) {
    m->newton[ce] = m->fpi[ce] = m->cut[ce] = m->nsub[ce] = 0;
    /* phase_restore -> mechanical_restore (phase_mechanical.f90:1096-1115) */
    memcpy(m->Fp + 9 * ce, m->Fp0 + 9 * ce, 9 * sizeof(double));
    memcpy(m->Fi + 9 * ce, m->Fi0 + 9 * ce, 9 * sizeof(double));
    memcpy(m->S + 9 * ce, m->S0 + 9 * ce, 9 * sizeof(double));
    memcpy(m->state + (size_t)ns * ce, m->state0 + (size_t)ns * ce, ns * sizeof(double));
    /* mechanical_partition: isostrain, one constituent */
    memcpy(m->F + 9 * ce, F_target + 9 * ce, 9 * sizeof(double));
    m->status[ce] = constitutive(m, ce, dt);
  }
  int status = ORC_STATUS_OK;
  for (int ce = 0; ce < N; ce++)
    if (m->status[ce] != ORC_STATUS_OK) status = ORC_STATUS_FAIL_PHASE_MECHANICAL;
  if (status == ORC_STATUS_OK) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int ce = 0; ce < N; ce++) {
      for (int k = 0; k < 9; k++) m->hom_P[9 * ce + k] = m->P[9 * ce + k] * 1.0;
      if (m->lazy_tangent) continue;
      orientation(m, ce);
      double d[81];
      orc_dPdF(m, ce, dt, d);
      for (int k = 0; k < 81; k++) m->hom_dPdF[(size_t)81 * ce + k] = d[k] * 1.0;
    }
  }
  if (P_out) memcpy(P_out, m->hom_P, sizeof(double) * 9 * N);
  if (dPdF_out) memcpy(dPdF_out, m->hom_dPdF, sizeof(double) * 81 * N);
  return status;
}

/* the second loop of homogenization_mechanical_response (homogenization.f90:352-360)
 * for the state of the last (converged) response: orientation and dPdF of every cell */
void orc_mp_tangent(orc_mp *m, double dt) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int ce = 0; ce < m->N; ce++) {
    orientation(m, ce);
    double d[81];
    orc_dPdF(m, ce, dt, d);
    for (int k = 0; k < 81; k++) m->hom_dPdF[(size_t)81 * ce + k] = d[k] * 1.0;
  }
}

void orc_mp_set_lazy_tangent(orc_mp *m, int lazy) { m->lazy_tangent = lazy; }

const double *orc_mp_dPdF_ptr(const orc_mp *m) { return m->hom_dPdF; }
int orc_mp_ncells(const orc_mp *m) { return m->N; }

void orc_mp_forward(orc_mp *m) {
  size_t N9 = (size_t)9 * m->N;
  memcpy(m->Fi0, m->Fi, N9 * sizeof(double));
  memcpy(m->Fp0, m->Fp, N9 * sizeof(double));
  memcpy(m->F0, m->F, N9 * sizeof(double));
  memcpy(m->Li0, m->Li, N9 * sizeof(double));
  memcpy(m->Lp0, m->Lp, N9 * sizeof(double));
  memcpy(m->S0, m->S, N9 * sizeof(double));
  memcpy(m->state0, m->state, (size_t)m->stride * m->N * sizeof(double));
}

void orc_mp_counters(const orc_mp *m, int64_t *newton, int64_t *fpi, int64_t *cutbacks) {
  int64_t a = 0, b = 0, c = 0;
  for (int i = 0; i < m->N; i++) { a += m->newton[i]; b += m->fpi[i]; c += m->cut[i]; }
  *newton = a; *fpi = b; *cutbacks = c;
}

static double *field_ptr(const orc_mp *m, int what) {
  switch (what) {
    case 0: return m->F;
    case 1: return m->P;
    case 2: return m->Fe;
    case 3: return m->Fp;
    case 4: return m->Lp;
    case 5: return m->S;
    case 6: return m->Fi;
    case 10: return m->Fp0;
    case 11: return m->Lp0;
    case 12: return m->S0;
    case 13: return m->F0;
    case 14: return m->Fi0;
    default: return NULL;
  }
}

void orc_mp_get(const orc_mp *m, int what, double *dst) {
  int N = m->N, nsl = m->stride / 2;
  if (what == 7) { memcpy(dst, m->O, sizeof(double) * 4 * N); return; }
  if (what == 8 || what == 9 || what == 15) {
    const double *st = what == 15 ? m->state0 : m->state;
    int off = what == 9 ? nsl : 0;
    int w = what == 15 ? 2 * nsl : nsl;
    for (int ce = 0; ce < N; ce++)
      for (int i = 0; i < w; i++) dst[(size_t)w * ce + i] = st[(size_t)m->stride * ce + off + i];
    return;
  }
  double *src = field_ptr(m, what);
  if (src) memcpy(dst, src, sizeof(double) * 9 * N);
}

void orc_mp_set(orc_mp *m, int what, const double *src) {
  int N = m->N, nsl = m->stride / 2;
  if (what == 7) { memcpy(m->O, src, sizeof(double) * 4 * N); return; }
  if (what == 8 || what == 9 || what == 15) {
    double *st = what == 15 ? m->state0 : m->state;
    int off = what == 9 ? nsl : 0;
    int w = what == 15 ? 2 * nsl : nsl;
    for (int ce = 0; ce < N; ce++)
      for (int i = 0; i < w; i++) st[(size_t)m->stride * ce + off + i] = src[(size_t)w * ce + i];
    return;
  }
  double *d = field_ptr(m, what);
  if (d) memcpy(d, src, sizeof(double) * 9 * N);
}

/* per-cell Newton and state-iteration counts of the last response */
void orc_mp_cell_counters(const orc_mp *m, int64_t *newton, int64_t *fpi) {
  for (int i = 0; i < m->N; i++) { newton[i] = m->newton[i]; fpi[i] = m->fpi[i]; }
}

/* per-cell state integrations (sub-steps tried, the first full step and every
 * cut-back retry included) of the last response */
void orc_mp_cell_substeps(const orc_mp *m, int64_t *substeps) {
  for (int i = 0; i < m->N; i++) substeps[i] = m->nsub[i];
}
