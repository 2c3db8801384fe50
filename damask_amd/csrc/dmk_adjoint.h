// Adjoint of one converged material-point increment, shared by the device
// kernels (adjoint.hip) and a host entry point used by the CPU tests.
//
// A converged increment of one cell solves, for y = (L_p, xi),
//   G1 = L_p - sum_a gdot_a(tau_a, xi_a) P_a                       = 0   (9)
//   G2 = xi - xi_0 - dt h_0 phi(xi) (H |gdot|)                      = 0   (m)
// with F_e = F F_p0^-1 (I - dt L_p), S = C : E(F_e), tau_a = S : P_nS_a
// (src/phase_mechanical.f90:379-584 integrateStress, :591-676 the fixed point
// of integrateStateFPI, src/phase_mechanical_plastic_phenopowerlaw.f90:303-399
// and 449-488), and returns
//   P    = F M S M^T,  M = F_p0^-1 (I - dt L_p)                    (:576)
//   F_p  = M^-1 / det(M^-1)^(1/3)                                   (:579)
//   xi.
// The reference's Newton and fixed point only find y; their damping, line
// search and warm start do not change it.  The derivatives of the outputs
// follow from the implicit-function theorem at the converged y: with the
// Jacobian G_y (9 + m square, assembled analytically and LU-factorised with
// partial pivoting), dy = -G_y^-1 (G_x dx), and for a cotangent o_bar of the
// outputs the input cotangents are x_bar = (do/dx)^T o_bar - G_x^T lambda with
// G_y^T lambda = (do/dy)^T o_bar.
//
// Twins, f_sat and c_1 enter only through xi_sat and h_0 (1 + c_1 0^c_2)
// (sum F_tw = 0), as in the forward kernels.  F_i = I, L_i = 0.
//
// 3x3 tensors are row-major here (T[3 i + j] = T_ij); fields and the tangent
// use the component order c = i + 3 j.
#pragma once

#include "dmk_common.h"
#include "dmk_math.h"
#include "dmk_adjoint_k.h"

// Per-cell storage: per-system temporaries av ([slot][avN] at cell e) and the
// matrix block g (G_y row-major with row stride nmax, then the pivots), both
// entry-major over the cells (entry r of cell e at r * avN + e), so that a
// wavefront of consecutive cells reads and writes each entry coalesced; host
// check: avN = 1, e = 0.
struct AdjAcc {
  double* av;
  size_t avN, e;
  double* g;
  int nmax, mmax;       // nmax = 9 + mmax
  DHI double& G(int i, int j) const { return g[((size_t)i * nmax + j) * avN + e]; }
  DHI double& piv(int i) const { return g[((size_t)nmax * nmax + i) * avN + e]; }
  DHI double& v(int k, int a) const { return av[((size_t)k * mmax + a) * avN + e]; }
};

// strided per-system vector (xi, xi_0 and their cotangents)
struct AdjVec {
  double* p;
  size_t ld, e;
  DHI double& operator()(int a) const { return p[(size_t)a * ld + e]; }
};

struct AdjPoint {
  double F[9], Fp0[9], Lp[9];   // row-major
  double dt;
};

struct AdjFwd {
  double Fp0i[9], A0[9], B[9], Fe[9], S[9], W[9], M[9];
};

DHI double adj_dot9(const double* a, const double* b) {
  double s = 0.0;
  for (int k = 0; k < 9; k++) s = fma(a[k], b[k], s);
  return s;
}

// C : sym(X) (minor symmetries of C make C : X = C : sym X)
DHI void adj_hooke(const DmkPhaseDev& ph, const double* X, double* S) {
  double Y[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Y[3 * i + j] = 0.5 * (X[3 * i + j] + X[3 * j + i]);
  hooke_sym(ph.C11, ph.C12, ph.C44, Y, S);
}

// 3x3 forward intermediates at the converged y; returns 1 when F_p0 is singular
DHI int adj_fwd33(const DmkPhaseDev& ph, const AdjPoint& pt, AdjFwd& f) {
  if (m_invert(pt.Fp0, f.Fp0i)) return 1;
  m_mul(pt.F, f.Fp0i, f.A0);
  for (int k = 0; k < 9; k++) f.B[k] = -pt.dt * pt.Lp[k];
  f.B[0] += 1.0; f.B[4] += 1.0; f.B[8] += 1.0;
  m_mul(f.A0, f.B, f.Fe);
  double E[9];
  m_mul_tn(f.Fe, f.Fe, E);
  for (int k = 0; k < 9; k++) E[k] *= 0.5;
  E[0] -= 0.5; E[4] -= 0.5; E[8] -= 0.5;
  hooke_sym(ph.C11, ph.C12, ph.C44, E, f.S);
  m_mul_tn(f.A0, f.Fe, f.W);
  m_mul(f.Fp0i, f.B, f.M);
  return 0;
}

// per-system temporaries (rates, their derivatives, hardening factors)
DHI void adj_fwd_sys(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjVec& xi, const AdjFwd& f,
                     const AdjAcc& acc) {
  const int m = ph.nsl;
  for (int a = 0; a < m; a++) {
    double tp = adj_dot9(f.S, ph.Pp[a]), tn = adj_dot9(f.S, ph.Pm[a]);
    bool pos = tp > tn;
    double tau = pos ? tp : tn, x = xi(a);
    double gd = (pos ? 1.0 : -1.0) * ph.g0[a] * pow(tau / x, ph.n[a]);
    double kt = gd != 0.0 ? gd * ph.n[a] / tau : 0.0;
    acc.v(AV_GD, a) = gd;
    acc.v(AV_KT, a) = kt;
    acc.v(AV_TM, a) = tau;
    acc.v(AV_SG, a) = pos ? 1.0 : -1.0;
    double u = 1.0 - x / ph.xi_sat[a], au = fabs(u);
    double phi = pow(au, ph.a[a]);
    if (u < 0.0) phi = -phi;                   // sign(|u|^a, u)
    acc.v(AV_PHI, a) = phi;
    acc.v(AV_DPHI, a) = au > 0.0 ? -ph.a[a] * pow(au, ph.a[a] - 1.0) / ph.xi_sat[a] : 0.0;
    // d tau_a / d L_p = -dt (A0^T F_e) (C : P_act)
    const double* Q = pos ? ph.Qp[a] : ph.Qm[a];
    double T[9];
    m_mul(f.W, Q, T);
    for (int k = 0; k < 9; k++) acc.v(AV_T + k, a) = -pt.dt * T[k];
  }
  for (int a = 0; a < m; a++) {
    double r = 0.0;
    for (int b = 0; b < m; b++) r = fma(ph.H[a][b], fabs(acc.v(AV_GD, b)), r);
    acc.v(AV_R, a) = r;
  }
}

DHI int adj_forward(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjVec& xi, AdjFwd& f, const AdjAcc& acc) {
  if (adj_fwd33(ph, pt, f)) return 1;
  adj_fwd_sys(ph, pt, xi, f, acc);
  return 0;
}

// residual of the increment equations at the recorded y in units of the
// reference's convergence tolerances (consistency check): max of
// |G1| / max(eps_rel_Lp max(|L_p|, |L_p,c|), eps_abs_Lp)   (phase_mechanical.f90:440-445)
// |G2_a| / max(atol_xi, eps_rel_state |xi_a|)              (phase.f90:703-716)
// The state is one fixed-point update past the stress solve, so values of a
// few units are normal.
DHI double adj_residual(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjVec& xi, const AdjVec& xi0,
                        const AdjAcc& acc, double eps_rel_Lp, double eps_abs_Lp, double eps_rel_state) {
  const int m = ph.nsl;
  double Lc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < m; a++)
    for (int k = 0; k < 9; k++) Lc[k] = fma(acc.v(AV_GD, a), ph.P[a][k], Lc[k]);
  double r1 = 0.0, nl = 0.0, nc = 0.0;
  for (int k = 0; k < 9; k++) {
    r1 += (pt.Lp[k] - Lc[k]) * (pt.Lp[k] - Lc[k]);
    nl += pt.Lp[k] * pt.Lp[k];
    nc += Lc[k] * Lc[k];
  }
  double res = sqrt(r1) / fmax(eps_rel_Lp * sqrt(fmax(nl, nc)), eps_abs_Lp);
  for (int a = 0; a < m; a++) {
    double g2 = xi(a) - xi0(a) - pt.dt * ph.hf[a] * acc.v(AV_PHI, a) * acc.v(AV_R, a);
    res = fmax(res, fabs(g2) / fmax(ph.atol_xi, eps_rel_state * fabs(xi(a))));
  }
  return res;
}

// Jacobian G_y of the increment equations
DHI void adj_assemble(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjVec& xi, const AdjAcc& acc) {
  const int m = ph.nsl;
  const double dt = pt.dt;
  // d G1 / d L_p = I - sum_b P_b kt_b T_b^T ; d G1 / d xi_b = P_b n_b gdot_b / xi_b
  for (int k = 0; k < 9; k++) {
    double row[9];
    for (int l = 0; l < 9; l++) row[l] = k == l ? 1.0 : 0.0;
    for (int b = 0; b < m; b++) {
      double c = ph.P[b][k] * acc.v(AV_KT, b);
      if (c != 0.0)
        for (int l = 0; l < 9; l++) row[l] = fma(-c, acc.v(AV_T + l, b), row[l]);
    }
    for (int l = 0; l < 9; l++) acc.G(k, l) = row[l];
    for (int b = 0; b < m; b++) acc.G(k, 9 + b) = ph.P[b][k] * ph.n[b] * acc.v(AV_GD, b) / xi(b);
  }
  // d G2_a / d L_p = -dt hf_a phi_a sum_b H_ab s_b kt_b T_b ;
  // d G2_a / d xi_b = delta_ab (1 - dt hf_a phi'_a r_a) + dt hf_a phi_a H_ab n_b |gdot_b| / xi_b
  for (int a = 0; a < m; a++) {
    double ca = -dt * ph.hf[a] * acc.v(AV_PHI, a);
    double row[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < m; b++) {
      double c = ca * ph.H[a][b] * acc.v(AV_SG, b) * acc.v(AV_KT, b);
      if (c != 0.0)
        for (int l = 0; l < 9; l++) row[l] = fma(c, acc.v(AV_T + l, b), row[l]);
    }
    for (int l = 0; l < 9; l++) acc.G(9 + a, l) = row[l];
    for (int b = 0; b < m; b++)
      acc.G(9 + a, 9 + b) = (a == b ? 1.0 - dt * ph.hf[a] * acc.v(AV_DPHI, a) * acc.v(AV_R, a) : 0.0)
                          - ca * ph.H[a][b] * ph.n[b] * fabs(acc.v(AV_GD, b)) / xi(b);
  }
}

// LU factorisation with partial pivoting in place (row interchanges in piv;
// the first largest pivot, as dgetrf); returns 1 for an exactly zero pivot
DHI int adj_lu(const AdjAcc& acc, int n) {
  for (int j = 0; j < n; j++) {
    int p = j;
    double amax = fabs(acc.G(j, j));
    for (int i = j + 1; i < n; i++) {
      double v = fabs(acc.G(i, j));
      if (v > amax) { amax = v; p = i; }
    }
    acc.piv(j) = (double)p;
    if (amax == 0.0) return 1;
    if (p != j)
      for (int k = 0; k < n; k++) {
        double t = acc.G(j, k);
        acc.G(j, k) = acc.G(p, k);
        acc.G(p, k) = t;
      }
    double r = 1.0 / acc.G(j, j);
    for (int i = j + 1; i < n; i++) acc.G(i, j) *= r;
    for (int k = j + 1; k < n; k++) {
      double gjk = acc.G(j, k);
      if (gjk != 0.0)
        for (int i = j + 1; i < n; i++) acc.G(i, k) = fma(-acc.G(i, j), gjk, acc.G(i, k));
    }
  }
  return 0;
}

DHI int adj_factor(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjVec& xi, const AdjAcc& acc) {
  adj_assemble(ph, pt, xi, acc);
  return adj_lu(acc, 9 + ph.nsl);
}

// G_y x = b in place
DHI void adj_solve(const AdjAcc& acc, int n, double* x) {
  for (int j = 0; j < n; j++) {
    int p = (int)acc.piv(j);
    if (p != j) { double t = x[j]; x[j] = x[p]; x[p] = t; }
  }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) x[i] = fma(-acc.G(i, j), x[j], x[i]);
  for (int j = n - 1; j >= 0; j--) {
    x[j] /= acc.G(j, j);
    for (int i = 0; i < j; i++) x[i] = fma(-acc.G(i, j), x[j], x[i]);
  }
}

// G_y^T x = b in place
DHI void adj_solve_t(const AdjAcc& acc, int n, double* x) {
  for (int j = 0; j < n; j++) {
    double s = x[j];
    for (int i = 0; i < j; i++) s = fma(-acc.G(i, j), x[i], s);
    x[j] = s / acc.G(j, j);
  }
  for (int j = n - 1; j >= 0; j--) {
    double s = x[j];
    for (int i = j + 1; i < n; i++) s = fma(-acc.G(i, j), x[i], s);
    x[j] = s;
  }
  for (int j = n - 1; j >= 0; j--) {
    int p = (int)acc.piv(j);
    if (p != j) { double t = x[j]; x[j] = x[p]; x[p] = t; }
  }
}

// dP/dF of the increment map with the internal variables re-equilibrated,
// one direction dF = e_q0 (x) e_q1 (q = q0 + 3 q1) at a time:
// right-hand side -dG/dF dF of G_y dy = -dG/dF dF: L_p part in xl (9), xi
// part in xs (m); w (m) is workspace.  The xi rows are formed one at a time
// from the rate changes so that nothing accumulates through memory.
DHI void adj_tangent_rhs(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, const AdjAcc& acc, int q,
                         double* xl, double* xs, double* w) {
  const int m = ph.nsl, q0 = q % 3, q1 = q / 3;
  const double dt = pt.dt;
  double dA0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 3; j++) dA0[3 * q0 + j] = f.Fp0i[3 * q1 + j];
  double dFe[9], dE[9], dS[9];
  m_mul(dA0, f.B, dFe);
  m_mul_tn(f.Fe, dFe, dE);
  adj_hooke(ph, dE, dS);
  double x0 = 0, x1 = 0, x2 = 0, x3 = 0, x4 = 0, x5 = 0, x6 = 0, x7 = 0, x8 = 0;
  for (int b = 0; b < m; b++) {
    const double sg = acc.v(AV_SG, b);
    const double* Pa = sg > 0.0 ? ph.Pp[b] : ph.Pm[b];
    const double dg = adj_dot9(Pa, dS) * acc.v(AV_KT, b);     // d gdot_b
    const double* P = ph.P[b];
    x0 = fma(dg, P[0], x0); x1 = fma(dg, P[1], x1); x2 = fma(dg, P[2], x2);
    x3 = fma(dg, P[3], x3); x4 = fma(dg, P[4], x4); x5 = fma(dg, P[5], x5);
    x6 = fma(dg, P[6], x6); x7 = fma(dg, P[7], x7); x8 = fma(dg, P[8], x8);
    w[b] = sg * dg;                                            // d |gdot_b|
  }
  xl[0] = x0; xl[1] = x1; xl[2] = x2; xl[3] = x3; xl[4] = x4; xl[5] = x5; xl[6] = x6; xl[7] = x7; xl[8] = x8;
  for (int a = 0; a < m; a++) {
    double s = 0.0;
    for (int b = 0; b < m; b++) s = fma(ph.H[a][b], w[b], s);
    xs[a] = dt * ph.hf[a] * acc.v(AV_PHI, a) * s;
  }
}

// ... and column q of A[p + 9 q] = d P_p / d F_q (field order) from dL_p = x[0..8],
// written to c[p] (p = 0..8)
DHI void adj_tangent_dP(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, int q, const double* x,
                        double* c) {
  const int q0 = q % 3, q1 = q / 3;
  const double dt = pt.dt;
  double SMt[9], MSMt[9], X[9];
  m_mul_nt(f.S, f.M, SMt);
  m_mul(f.M, SMt, MSMt);
  m_mul(pt.F, f.M, X);
  double dA0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 3; j++) dA0[3 * q0 + j] = f.Fp0i[3 * q1 + j];
  double dB[9], dM[9], dFe2[9], dE2[9], dS2[9];
  for (int k = 0; k < 9; k++) dB[k] = -dt * x[k];
  m_mul(f.Fp0i, dB, dM);
  m_mul(f.A0, dB, dFe2);
  double T1[9];
  m_mul(dA0, f.B, T1);
  for (int k = 0; k < 9; k++) dFe2[k] += T1[k];
  m_mul_tn(f.Fe, dFe2, dE2);
  adj_hooke(ph, dE2, dS2);
  // dP = dF M S M^T + F dM S M^T + F M dS M^T + F M S dM^T
  double dP[9], T2[9];
  for (int k = 0; k < 9; k++) dP[k] = 0.0;
  for (int j = 0; j < 3; j++) dP[3 * q0 + j] = MSMt[3 * q1 + j];
  m_mul(pt.F, dM, T1);
  m_mul(T1, SMt, T2);
  for (int k = 0; k < 9; k++) dP[k] += T2[k];
  m_mul_nt(dS2, f.M, T1);
  m_mul(X, T1, T2);
  for (int k = 0; k < 9; k++) dP[k] += T2[k];
  m_mul(X, f.S, T1);
  m_mul_nt(T1, dM, T2);
  for (int k = 0; k < 9; k++) dP[k] += T2[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c[i + 3 * j] = dP[3 * i + j];
}

DHI void adj_tangent_col(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, int q, const double* x,
                         double* A) {
  adj_tangent_dP(ph, pt, f, q, x, A + 9 * q);
}

DHI void adj_tangent(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, const AdjAcc& acc, double* A) {
  for (int q = 0; q < 9; q++) {
    double x[9 + DMK_NSL_MAX], w[DMK_NSL_MAX];
    adj_tangent_rhs(ph, pt, f, acc, q, x, x + 9, w);
    adj_solve(acc, 9 + ph.nsl, x);
    adj_tangent_col(ph, pt, f, q, x, A);
  }
}

// chain S_bar -> (F_bar, Fp0i_bar, B_bar) through F_e = F Fp0^-1 B, E, S = C : E
DHI void adj_back_S(const DmkPhaseDev& ph, const double* F, const AdjFwd& f, const double* Sb, double* Fb,
                    double* Fp0ib, double* Bb, double sgn) {
  double Eb[9], Feb[9], A0b[9], T[9];
  adj_hooke(ph, Sb, Eb);
  m_mul(f.Fe, Eb, Feb);
  m_mul_nt(Feb, f.B, A0b);
  m_mul_nt(A0b, f.Fp0i, T);
  for (int k = 0; k < 9; k++) Fb[k] = fma(sgn, T[k], Fb[k]);
  m_mul_tn(F, A0b, T);
  for (int k = 0; k < 9; k++) Fp0ib[k] = fma(sgn, T[k], Fp0ib[k]);
  if (Bb) {
    m_mul_tn(f.A0, Feb, T);
    for (int k = 0; k < 9; k++) Bb[k] = fma(sgn, T[k], Bb[k]);
  }
}

// Vector-Jacobian product of the increment map, in two halves around the
// transposed solve.  Cotangents of the outputs: Pb (P), Fpb (F_p, may be
// null), xib (xi, ld 0 for none).  The first half returns the right-hand side
// (do/dy)^T o_bar in lam and the explicit parts of the F and F_p0^-1
// cotangents in Fb, Fp0ib.
DHI void adj_vjp_rhs(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, const double* Pb,
                     const double* Fpb, const AdjVec& xib, double* lam, double* Fb, double* Fp0ib) {
  const int m = ph.nsl;
  const double dt = pt.dt;
  double X[9], K[9], T[9], T2[9], Sb[9], Mb[9], Bb[9];
  m_mul(pt.F, f.M, X);
  // P = X S M^T
  m_mul_nt(f.S, f.M, T);
  m_mul(f.M, T, K);
  m_mul(Pb, K, Fb);                       // F_bar = P_bar M S M^T
  m_mul_tn(X, Pb, T);
  m_mul(T, f.M, Sb);                      // S_bar = X^T P_bar M
  m_mul_tn(pt.F, Pb, T);
  m_mul(T, f.M, T2);
  m_mul(T2, f.S, Mb);                     // F^T P_bar M S
  m_mul_tn(Pb, X, T);
  m_mul(T, f.S, T2);
  for (int k = 0; k < 9; k++) Mb[k] += T2[k];   // + P_bar^T X S
  if (Fpb) {
    // F_p = c Fpr, Fpr = M^-1, c = det(Fpr)^(-1/3)
    double Fpr[9];
    m_invert(f.M, Fpr);
    double c = pow(m_det(Fpr), -1.0 / 3.0), fp = adj_dot9(Fpb, Fpr), Fprb[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Fprb[3 * i + j] = c * (Fpb[3 * i + j] - fp / 3.0 * f.M[3 * j + i]);
    m_mul_tn(Fpr, Fprb, T);
    m_mul_nt(T, Fpr, T2);
    for (int k = 0; k < 9; k++) Mb[k] -= T2[k];
  }
  m_mul_nt(Mb, f.B, Fp0ib);               // M = Fp0i B
  m_mul_tn(f.Fp0i, Mb, Bb);
  adj_back_S(ph, pt.F, f, Sb, Fb, Fp0ib, Bb, 1.0);
  for (int k = 0; k < 9; k++) lam[k] = -dt * Bb[k];
  for (int a = 0; a < m; a++) lam[9 + a] = xib.ld ? xib(a) : 0.0;
}

// Second half, with lam = G_y^-T (do/dy)^T o_bar: F_bar -= G_F^T lam, and with
// params the cotangents of F_p0 (Fp0b), xi_0 (xi0b) and the parameters (thb:
// slot * ldth + system, interaction types at ADJ_TH_SYS * ldth + t); every
// slot of the phase is written once.
DHI void adj_vjp_finish(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, const AdjVec& xi,
                        const AdjAcc& acc, const double* lam, double* Fb, double* Fp0ib, bool params, double* Fp0b,
                        const AdjVec& xi0b, const AdjVec& thb, int ldth) {
  const int m = ph.nsl;
  const double dt = pt.dt;
  double T[9], T2[9];
  // lambda . G through tau: w_c = -kt_c (<Lambda, P_c> + dt s_c psi_c)
  double SbG[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < m; c++) {
    double psi = 0.0;
    for (int a = 0; a < m; a++) psi = fma(lam[9 + a] * ph.hf[a] * acc.v(AV_PHI, a), ph.H[a][c], psi);
    double Gc = adj_dot9(lam, ph.P[c]) + dt * acc.v(AV_SG, c) * psi;
    double w = -acc.v(AV_KT, c) * Gc;
    const double* Pa = acc.v(AV_SG, c) > 0.0 ? ph.Pp[c] : ph.Pm[c];
    for (int k = 0; k < 9; k++) SbG[k] = fma(w, Pa[k], SbG[k]);
    if (params) {
      double gd = acc.v(AV_GD, c);
      // d/dn (tau/xi)^n = (tau/xi)^n ln(tau/xi) for tau > 0.  With non-Schmid
      // terms both projections can be negative; (tau/xi)^n then exists only for
      // integer n and has no derivative in n, taken as 0 (the C library's
      // log domain, as the forward-mode reference does)
      const double tm = acc.v(AV_TM, c);
      thb(ADJ_TH_N * ldth + c) = gd != 0.0 && tm > 0.0 ? Gc * gd * log(tm / xi(c)) : 0.0;
      thb(ADJ_TH_G0 * ldth + c) = gd != 0.0 ? Gc * gd / ph.g0[c] : 0.0;
    }
  }
  adj_back_S(ph, pt.F, f, SbG, Fb, Fp0ib, nullptr, -1.0);
  if (!params) return;
  // F_p0 = (Fp0i)^-1: Fp0_bar = -Fp0i^T Fp0i_bar Fp0i^T
  m_mul_tn(f.Fp0i, Fp0ib, T);
  m_mul_nt(T, f.Fp0i, T2);
  for (int k = 0; k < 9; k++) Fp0b[k] = -T2[k];
  for (int a = 0; a < m; a++) {
    double l2 = lam[9 + a];
    xi0b(a) = l2;
    double phi = acc.v(AV_PHI, a), r = acc.v(AV_R, a), x = xi(a);
    double u = 1.0 - x / ph.xi_sat[a], au = fabs(u);
    double base = dt * l2 * r;
    thb(ADJ_TH_XISAT * ldth + a) = base * ph.hf[a] * ph.a[a] * pow(au, ph.a[a] - 1.0) * x / (ph.xi_sat[a] * ph.xi_sat[a]);
    thb(ADJ_TH_H0 * ldth + a) = base * phi * (ph.h0[a] != 0.0 ? ph.hf[a] / ph.h0[a] : 1.0);
    thb(ADJ_TH_A * ldth + a) = au > 0.0 ? base * ph.hf[a] * phi * log(au) : 0.0;
  }
  // interaction types: sum over the (reacting, acting) pairs of each type
  for (int t = 0; t < ph.n_h; t++) {
    double st = 0.0;
    for (int a = 0; a < m; a++) {
      double ca = 0.0;
      for (int b = 0; b < m; b++)
        if (ph.Ht[a][b] == t) ca += fabs(acc.v(AV_GD, b));
      st = fma(dt * lam[9 + a] * ph.hf[a] * acc.v(AV_PHI, a), ca, st);
    }
    thb(ADJ_TH_SYS * ldth + t) = st;
  }
}

DHI void adj_vjp(const DmkPhaseDev& ph, const AdjPoint& pt, const AdjFwd& f, const AdjVec& xi, const AdjAcc& acc,
                 const double* Pb, const double* Fpb, const AdjVec& xib, double* Fb, bool params, double* Fp0b,
                 const AdjVec& xi0b, const AdjVec& thb, int ldth) {
  double lam[9 + DMK_NSL_MAX], Fp0ib[9];
  adj_vjp_rhs(ph, pt, f, Pb, Fpb, xib, lam, Fb, Fp0ib);
  adj_solve_t(acc, 9 + ph.nsl, lam);
  adj_vjp_finish(ph, pt, f, xi, acc, lam, Fb, Fp0ib, params, Fp0b, xi0b, thb, ldth);
}
