// Per-material-point constitutive update on MI355X: one thread per cell, FP64.
//
// stress_kernel: homogenization_mechanical_response's first loop
//   (src/homogenization.f90:309-348): restore, isostrain partition,
//   phase_mechanical_constitutive substepping (src/phase_mechanical.f90:1008-1090),
//   integrateStateFPI (:591-676) and integrateStress (:379-584) with
//   phenopowerlaw kinetics (src/phase_mechanical_plastic_phenopowerlaw.f90:303-399,449-488)
//   and Hooke's law (src/phase_mechanical_elastic.f90:172-205).
// tangent_kernel: the second loop (:352-360): crystallite orientation
//   (src/phase.f90:659-673) and phase_mechanical_dPdF (src/phase_mechanical.f90:1121-1242).
//
// The material point carries no eigen kinematics and V_e = I, so F_i = I and
// L_i = 0 throughout; the 3x3x3x3 chains of the reference then collapse to
// rank-N_sl updates of the 9x9 identity:
//   dR_Lp/dL_p = I9 + dt * sum_a dgamma_a vec(P_a) vec(G^T Q_a)^T,  G = F_e^T A,
//   Q_a = C : P_nS,a (tabulated per phase), vec row-major (math_33to9),
// which is what is assembled in registers here.
#include "dmk_common.h"
#include "dmk_math.h"

namespace {

template <int NSL>
DI void kinetics(const DmkPhaseDev* __restrict__ ph, const double* Mp, const double* xi, double* g,
                 double* dg, bool* pos) {
#pragma unroll
  for (int a = 0; a < NSL; a++) {
    double tp = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) tp = fma(Mp[k], ph->Pp[a][k], tp);
    double tn;
    if (ph->nonschmid) {
      tn = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) tn = fma(Mp[k], ph->Pm[a][k], tn);
    } else {
      tn = -tp;
    }
    double sgn = tp > tn ? 1.0 : -1.0;
    double tmax = fmax(tp, tn);
    double gg = sgn * ph->g0[a] * pow(tmax / xi[a], ph->n[a]);
    g[a] = gg;
    if (dg) dg[a] = !dEq0(gg) ? gg * ph->n[a] / tmax : 0.0;
    if (pos) pos[a] = gg > 0.0;
  }
}

// E = (Fe^T Fe - I)/2 and S = C:E (lattice frame)
DI void hooke(const DmkPhaseDev* __restrict__ ph, const double* Fe, double* S) {
  double E[9];
  m_mul_tn(Fe, Fe, E);
  E[0] -= 1.0; E[4] -= 1.0; E[8] -= 1.0;
#pragma unroll
  for (int k = 0; k < 9; k++) E[k] *= 0.5;
  hooke_sym(ph->C11, ph->C12, ph->C44, E, S);
}

// LU factorisation with partial pivoting (first maximum wins, as dgetrf) of a
// row-major 9x9 in registers and solve for one right-hand side.  Rows are
// swapped physically with selects so every index stays compile-time.
DI bool lu9_factor(double* J, int* piv) {
  bool sing = false;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    int p = k;
    double amax = fabs(J[9 * k + k]);
#pragma unroll
    for (int i = k + 1; i < 9; i++) {
      double v = fabs(J[9 * i + k]);
      if (v > amax) { amax = v; p = i; }
    }
    piv[k] = p;
#pragma unroll
    for (int i = k + 1; i < 9; i++) {
      bool sw = (i == p);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        double a = J[9 * k + c], b = J[9 * i + c];
        J[9 * k + c] = sw ? b : a;
        J[9 * i + c] = sw ? a : b;
      }
    }
    double d = J[9 * k + k];
    if (d == 0.0) sing = true;
    double r = 1.0 / d;
#pragma unroll
    for (int i = k + 1; i < 9; i++) {
      double l = J[9 * i + k] * r;
      J[9 * i + k] = l;
#pragma unroll
      for (int c = k + 1; c < 9; c++) J[9 * i + c] = fma(-l, J[9 * k + c], J[9 * i + c]);
    }
  }
  return sing;
}

DI void lu9_solve(const double* J, const int* piv, double* b) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = k + 1; i < 9; i++) {
      bool sw = (i == piv[k]);
      double a = b[k], c = b[i];
      b[k] = sw ? c : a;
      b[i] = sw ? a : c;
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int i = k + 1; i < 9; i++) b[i] = fma(-J[9 * i + k], b[k], b[i]);
#pragma unroll
  for (int k = 8; k >= 0; k--) {
    b[k] = b[k] / J[9 * k + k];
#pragma unroll
    for (int i = 0; i < k; i++) b[i] = fma(-J[9 * i + k], b[k], b[i]);
  }
}

// integrateStress (src/phase_mechanical.f90:379-584) with F_i = I, L_i = 0.
// On success writes P, S, L_p, F_p (det-regularised) and F_e of the cell.
template <int NSL>
DI int integrate_stress(const DmkPhaseDev* __restrict__ ph, const DmkNumMech& num, DmkMPDev& mp,
                        int e, const double* F, const double* Fp0, double dt, const double* xi,
                        int& newton) {
  const int N = mp.N;
  double invFp_cur[9], A[9];
  if (m_invert(Fp0, invFp_cur)) return ST_FAIL_STRESS;
  m_mul(F, invFp_cur, A);

  double Lpg[9], Lpg_old[9], dLp[9], B[9], S[9];
  m_load(mp.Lp, N, e, Lpg);
#pragma unroll
  for (int k = 0; k < 9; k++) { Lpg_old[k] = Lpg[k]; dLp[k] = 0.0; }
  double step = 1.0, res_old = DBL_MAX;
  int it = 0;
  for (;;) {
    it++;
    if (it > num.N_iter_Lp_max) return ST_FAIL_LP_MAXIT;
    newton++;
#pragma unroll
    for (int k = 0; k < 9; k++) B[k] = ((k % 4) == 0 ? 1.0 : 0.0) - dt * Lpg[k];
    double Fe[9];
    m_mul(A, B, Fe);
    hooke(ph, Fe, S);
    double g[NSL], dg[NSL];
    bool pos[NSL];
    kinetics<NSL>(ph, S, xi, g, dg, pos);
    double Lpc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Lpc[k] = 0.0;
#pragma unroll
    for (int a = 0; a < NSL; a++)
#pragma unroll
      for (int k = 0; k < 9; k++) Lpc[k] = fma(g[a], ph->P[a][k], Lpc[k]);
    double ng = m_norm(Lpg), nc = m_norm(Lpc);
    double atol = fmax(num.eps_rel_Lp * fmax(ng, nc), num.eps_abs_Lp);
    double res[9];
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) { res[k] = Lpg[k] - Lpc[k]; nan |= (res[k] != res[k]); }
    double nres = m_norm(res);
    if (nan) return ST_FAIL_LP_NAN;
    if (nres < atol) break;
    if (nres < res_old) {
      res_old = nres;
#pragma unroll
      for (int k = 0; k < 9; k++) Lpg_old[k] = Lpg[k];
      step = 1.0;
    } else {
      step = num.r_linesearch_Lp * step;
#pragma unroll
      for (int k = 0; k < 9; k++) Lpg[k] = fma(dLp[k], step, Lpg_old[k]);
      continue;
    }
    // Jacobian, rebuilt every iteration (f_update_jacobi_Lp = 1)
    double G[9];
    m_mul_tn(Fe, A, G);
    double J[81];
#pragma unroll
    for (int r = 0; r < 81; r++) J[r] = (r % 10 == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int a = 0; a < NSL; a++) {
      const double* Q = pos[a] ? ph->Qp[a] : ph->Qm[a];
      double Y[9];
      m_mul_tn(G, Q, Y);
      double u = dt * dg[a];
#pragma unroll
      for (int r = 0; r < 9; r++) {
        double pr = u * ph->P[a][r];
#pragma unroll
        for (int c = 0; c < 9; c++) J[9 * r + c] = fma(pr, Y[c], J[9 * r + c]);
      }
    }
    int piv[9];
    if (lu9_factor(J, piv)) return ST_FAIL_LP_DGESV;
    lu9_solve(J, piv, res);
#pragma unroll
    for (int k = 0; k < 9; k++) dLp[k] = -res[k];
#pragma unroll
    for (int k = 0; k < 9; k++) Lpg[k] = fma(dLp[k], step, Lpg[k]);
  }

  double invFp_new[9], Fp_new[9];
  m_mul(invFp_cur, B, invFp_new);
  if (m_invert(invFp_new, Fp_new)) return ST_FAIL_STRESS;
  double T1[9], T2[9], P[9];
  m_mul(F, invFp_new, T1);
  m_mul_nt(S, invFp_new, T2);
  m_mul(T1, T2, P);
  m_store(mp.Pph, N, e, P);
  m_store(mp.S, N, e, S);
  m_store(mp.Lp, N, e, Lpg);
  double d = pow(m_det(Fp_new), 1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 9; k++) Fp_new[k] = Fp_new[k] / d;
  m_store(mp.Fp, N, e, Fp_new);
  m_store(mp.Fe, N, e, T1);
  return ST_OK;
}

// plastic_dotState -> phenopowerlaw_dotState (no twins): dot[0..NSL) xi, [NSL..2NSL) gamma
template <int NSL>
DI bool dot_state(const DmkPhaseDev* __restrict__ ph, const double* S, const double* xi, double* dot) {
  double g[NSL];
  kinetics<NSL>(ph, S, xi, g, nullptr, nullptr);
  bool nan = false;
#pragma unroll
  for (int a = 0; a < NSL; a++) g[a] = fabs(g[a]);
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    double x = 1.0 - xi[r] / (ph->xi_inf[r] + ph->fsat[r] * 0.0);
    double left = copysign(pow(fabs(x), ph->a[r]), x);
    double hg = 0.0;
#pragma unroll
    for (int a = 0; a < NSL; a++) hg = fma(ph->H[r][a], g[a], hg);
    double f = ph->h0[r] * (1.0 + ph->c1[r] * pow(0.0, ph->c2[r]));
    dot[r] = f * left * hg;
    dot[NSL + r] = g[r];
    nan |= (dot[r] != dot[r]) || (g[r] != g[r]);
  }
  return nan;
}

// integrateStateFPI (src/phase_mechanical.f90:591-676); xi in registers,
// gamma and the iteration history in per-cell SoA scratch.
template <int NSL>
DI int integrate_state_fpi(const DmkPhaseDev* __restrict__ ph, const DmkNumMech& num, DmkMPDev& mp,
                           int e, const double* F, const double* Fp0, double dt, double* xi,
                           const double* xi0s, int& newton, int& fpi) {
  const int N = mp.N;
  const size_t sN = (size_t)N;
  double S[9];
  m_load(mp.S, N, e, S);
  double dot[2 * NSL];
  if (dot_state<NSL>(ph, S, xi, dot)) return ST_FAIL_STATE;
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    xi[r] = fma(dot[r], dt, xi0s[r]);
    mp.gam[r * sN + e] = fma(dot[NSL + r], dt, mp.st0s[(NSL + r) * sN + e]);
  }
  int status = ST_FAIL_STATE;
  for (int it = 1; it <= num.N_iter_state_max; it++) {
#pragma unroll
    for (int r = 0; r < 2 * NSL; r++) {
      mp.last2[r * sN + e] = it > 1 ? mp.last1[r * sN + e] : 0.0;
      mp.last1[r * sN + e] = dot[r];
    }
    fpi++;
    status = integrate_stress<NSL>(ph, num, mp, e, F, Fp0, dt, xi, newton);
    if (status != ST_OK) break;
    m_load(mp.S, N, e, S);
    if (dot_state<NSL>(ph, S, xi, dot)) break;
    // damper (src/phase_mechanical.f90:653-674)
    double d12 = 0.0, d22 = 0.0, d01 = 0.0;
#pragma unroll
    for (int r = 0; r < 2 * NSL; r++) {
      double w1 = mp.last1[r * sN + e], w2 = mp.last2[r * sN + e];
      d12 = fma(dot[r] - w1, w1 - w2, d12);
      d22 = fma(w1 - w2, w1 - w2, d22);
      d01 = fma(dot[r], w1, d01);
    }
    double zeta = (fmin(d01, d12) < 0.0 && d22 > 0.0) ? 0.75 + 0.25 * tanh(2.0 + 4.0 * d12 / d22) : 1.0;
    bool conv = true;
#pragma unroll
    for (int r = 0; r < 2 * NSL; r++) {
      double w1 = mp.last1[r * sN + e];
      double d = dot[r] * zeta + w1 * (1.0 - zeta);
      dot[r] = d;
      double s0 = mp.st0s[r * sN + e];
      double s = r < NSL ? xi[r] : mp.gam[(r - NSL) * sN + e];
      double res = s - s0 - d * dt;
      s = s - res;
      double atol = r < NSL ? ph->atol_xi : ph->atol_gamma;
      conv = conv && (fabs(res) <= fmax(atol, num.eps_rel_state * fabs(s)));
      if (r < NSL) xi[r] = s; else mp.gam[(r - NSL) * sN + e] = s;
    }
    if (conv) { status = ST_OK; break; }
  }
  return status;
}

template <int NSL>
__global__ void __launch_bounds__(128) stress_kernel(DmkMPDev mp, const double* __restrict__ Ftarget,
                                                     const DmkPhaseDev* __restrict__ ph, int phase_id,
                                                     DmkNumMech num, double dt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = mp.N;
  if (e >= N) return;
  if (mp.phase[e] != phase_id) return;
  const size_t sN = (size_t)N;

  // phase_restore (src/phase_mechanical.f90:1096-1115) + isostrain partition
  double F[9], F0i[9];
  m_load(Ftarget, N, e, F);
  m_store(mp.Fphase, N, e, F);
  m_load(mp.F0, N, e, F0i);
  double Fp0s[9], Lp0s[9], F0s[9], Fsub[9];
  m_load(mp.Fp0, N, e, Fp0s);
  m_load(mp.Lp0, N, e, Lp0s);
#pragma unroll
  for (int k = 0; k < 9; k++) { F0s[k] = F0i[k]; Fsub[k] = 0.0; }
  double xi[NSL], xi0s[NSL];
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    xi0s[r] = mp.xi0[r * sN + e];
    xi[r] = xi0s[r];
    double g0 = mp.gam0[r * sN + e];
    mp.gam[r * sN + e] = g0;
    mp.st0s[r * sN + e] = xi0s[r];
    mp.st0s[(NSL + r) * sN + e] = g0;
  }
  for (int k = 0; k < 9; k++) mp.S[k * sN + e] = mp.S0[k * sN + e];
  m_store(mp.Fp, N, e, Fp0s);

  int newton = 0, fpi = 0;
  // phase_mechanical_constitutive substepping
  double stepFrac = 0.0, step = 1.0 / num.r_cutback;
  int status = ST_ITERATING;
  bool todo = true;
  while (todo) {
    if (status == ST_OK) {
      stepFrac = stepFrac + step;
      step = fmin(1.0 - stepFrac, num.r_increase * step);
      todo = step > 0.0;
      if (todo) {
#pragma unroll
        for (int k = 0; k < 9; k++) F0s[k] = Fsub[k];
        m_load(mp.Lp, N, e, Lp0s);
        m_load(mp.Fp, N, e, Fp0s);
#pragma unroll
        for (int r = 0; r < NSL; r++) {
          xi0s[r] = xi[r];
          mp.st0s[r * sN + e] = xi[r];
          mp.st0s[(NSL + r) * sN + e] = mp.gam[r * sN + e];
        }
      }
    } else {
      step = num.r_cutback * step;
      m_store(mp.Fp, N, e, Fp0s);
      for (int k = 0; k < 9; k++) mp.S[k * sN + e] = mp.S0[k * sN + e];
      if (step < 1.0) m_store(mp.Lp, N, e, Lp0s);
#pragma unroll
      for (int r = 0; r < NSL; r++) {
        xi[r] = xi0s[r];
        mp.gam[r * sN + e] = mp.st0s[(NSL + r) * sN + e];
      }
      todo = step > num.r_cutback_min;
    }
    if (todo) {
#pragma unroll
      for (int k = 0; k < 9; k++) Fsub[k] = fma(step, F[k] - F0i[k], F0s[k]);
      status = integrate_state_fpi<NSL>(ph, num, mp, e, Fsub, Fp0s, step * dt, xi, xi0s, newton, fpi);
    }
  }
#pragma unroll
  for (int r = 0; r < NSL; r++) mp.xi[r * sN + e] = xi[r];
  mp.status[e] = status;
  mp.n_newton[e] = newton;
  mp.n_fpi[e] = fpi;
}

// math_rotationalPart (src/math.f90:1004-1043)
DI void rotational_part(const double* F, double* R) {
  const double TAU = 6.283185307179586;
  double C[9], FF[9], FC[9];
  m_mul_tn(F, F, C);
  double I1 = C[0] + C[4] + C[8];
  double I2 = C[0] * C[4] + C[0] * C[8] + C[4] * C[8] - (C[1] * C[1] + C[2] * C[2] + C[5] * C[5]);
  double I3 = -(C[0] * C[5] * C[5] + C[4] * C[2] * C[2] + C[8] * C[1] * C[1]) + C[0] * C[4] * C[8]
            + 2.0 * C[1] * C[2] * C[5];
  m_mul(F, F, FF);
  double trF = F[0] + F[4] + F[8];
  double IF2 = 0.5 * (trF * trF - (FF[0] + FF[4] + FF[8]));
  double l[3];
  double q = fmax(I1 * I1 - 3.0 * I2, 0.0);
  double x = pow(q, 1.5);
  if (!dEq0(x)) {
    double arg = fmin(fmax((I1 * I1 * I1 - 4.5 * I1 * I2 + 13.5 * I3) / x, -1.0), 1.0);
    double Phi = acos(arg);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double lam = I1 + (2.0 * sqrt(q)) * cos((Phi - TAU * (double)(k + 1)) / 3.0);
      l[k] = sqrt(fmax(lam, 0.0) / 3.0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) l[k] = sqrt(I1 / 3.0);
  }
  double U1 = l[0] + l[1] + l[2], U2 = l[0] * l[1] + l[1] * l[2] + l[2] * l[0];
  m_mul(F, C, FC);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      R[3 * i + j] = U1 * IF2 * (i == j ? 1.0 : 0.0) + (U1 * U1 - U2) * F[3 * i + j] - U1 * trF * F[3 * j + i]
                   + U1 * FF[3 * j + i] - FC[3 * i + j];
  double d = pow(m_det(R), -1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] *= d;
}

// om2qu (src/rotations.f90:486-514), P = -1
DI void om2qu(const double* om, double* qu) {
  double tr = om[0] + om[4] + om[8];
  double s;
  if (tr > 0.0) {
    s = 0.5 / sqrt(tr + 1.0);
    qu[0] = 0.25 / s; qu[1] = (om[7] - om[5]) * s; qu[2] = (om[2] - om[6]) * s; qu[3] = (om[3] - om[1]) * s;
  } else if (om[0] > om[4] && om[0] > om[8]) {
    s = 2.0 * sqrt(1.0 + om[0] - om[4] - om[8]);
    qu[0] = (om[7] - om[5]) / s; qu[1] = 0.25 * s; qu[2] = (om[1] + om[3]) / s; qu[3] = (om[2] + om[6]) / s;
  } else if (om[4] > om[8]) {
    s = 2.0 * sqrt(1.0 + om[4] - om[0] - om[8]);
    qu[0] = (om[2] - om[6]) / s; qu[1] = (om[1] + om[3]) / s; qu[2] = 0.25 * s; qu[3] = (om[5] + om[7]) / s;
  } else {
    s = 2.0 * sqrt(1.0 + om[8] - om[0] - om[4]);
    qu[0] = (om[3] - om[1]) / s; qu[1] = (om[2] + om[6]) / s; qu[2] = (om[5] + om[7]) / s; qu[3] = 0.25 * s;
  }
  if (signbit(qu[0])) for (int k = 0; k < 4; k++) qu[k] = -qu[k];
  for (int k = 1; k < 4; k++) if (!dEq0(qu[k])) qu[k] = -qu[k];
  double n = sqrt(qu[0] * qu[0] + qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3]);
  for (int k = 0; k < 4; k++) qu[k] /= n;
}

// Orientation, consistent tangent and homogenization (one constituent, v = 1).
// Runs only when every cell of the response converged (flag[0] == ST_OK).
template <int NSL>
__global__ void __launch_bounds__(128) tangent_kernel(DmkMPDev mp, const DmkPhaseDev* __restrict__ ph,
                                                      int phase_id, double dt, const int* __restrict__ flag,
                                                      double* __restrict__ Phom, double* __restrict__ dPdF) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = mp.N;
  if (e >= N || flag[0] != ST_OK) return;
  if (mp.phase[e] != phase_id) return;
  const size_t sN = (size_t)N;
  double Fe[9], Fp[9], S[9], F[9], Fp0[9];
  m_load(mp.Fe, N, e, Fe);
  m_load(mp.Fp, N, e, Fp);
  m_load(mp.S, N, e, S);
  m_load(mp.Fphase, N, e, F);
  m_load(mp.Fp0, N, e, Fp0);
  {
    double R[9], Rt[9], q[4];
    rotational_part(Fe, R);
    m_transpose(R, Rt);
    om2qu(Rt, q);
#pragma unroll
    for (int k = 0; k < 4; k++) mp.O[k * sN + e] = q[k];
  }
  for (int k = 0; k < 9; k++) Phom[k * sN + e] = mp.Pph[k * sN + e] * 1.0;

  double xi[NSL], g[NSL], dg[NSL];
  bool pos[NSL];
#pragma unroll
  for (int r = 0; r < NSL; r++) xi[r] = mp.xi[r * sN + e];
  kinetics<NSL>(ph, S, xi, g, dg, pos);

  double invFp[9], invSubFp0[9], A2[9], G2[9];
  m_invert(Fp, invFp);
  m_invert(Fp0, invSubFp0);
  m_mul(F, invSubFp0, A2);
  m_mul_tn(Fe, A2, G2);
  // K = I9 + dt sum_a dg_a vec(C:(G2 P_a)) vec(Pn_a)^T
  double K[81];
#pragma unroll
  for (int r = 0; r < 81; r++) K[r] = (r % 10 == 0) ? 1.0 : 0.0;
#pragma unroll
  for (int a = 0; a < NSL; a++) {
    double X[9], Xs[9], W[9];
    m_mul(G2, ph->P[a], X);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) Xs[3 * i + j] = 0.5 * (X[3 * i + j] + X[3 * j + i]);
    hooke_sym(ph->C11, ph->C12, ph->C44, Xs, W);
    const double* Pn = pos[a] ? ph->Pp[a] : ph->Pm[a];
    double u = dt * dg[a];
#pragma unroll
    for (int r = 0; r < 9; r++) {
      double wr = u * W[r];
#pragma unroll
      for (int c = 0; c < 9; c++) K[9 * r + c] = fma(wr, Pn[c], K[9 * r + c]);
    }
  }
  int piv[9];
  bool sing = lu9_factor(K, piv);

  double M1[9], M2[9], M3[9], Mid[9];
  m_mul_nt(S, invFp, M1);        // S invFp^T
  m_mul(F, invFp, M2);           // F invFp
  m_mul(M2, S, M3);              // F invFp S
  m_mul(invFp, M1, Mid);         // invFp S invFp^T
#pragma unroll 1
  for (int k = 0; k < 3; k++) {
#pragma unroll 1
    for (int l = 0; l < 3; l++) {
      // rhs column: C : sym(Fe(k,:) (x) invFp(l,:))
      double X[9], y[9];
#pragma unroll
      for (int n = 0; n < 3; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) X[3 * n + m] = 0.5 * (Fe[3 * k + n] * invFp[3 * l + m] + Fe[3 * k + m] * invFp[3 * l + n]);
      hooke_sym(ph->C11, ph->C12, ph->C44, X, y);
      if (!sing) lu9_solve(K, piv, y);
      // t = sum_a dg_a P_a (Pn_a : dSdF)
      double t[9];
#pragma unroll
      for (int q = 0; q < 9; q++) t[q] = 0.0;
#pragma unroll
      for (int a = 0; a < NSL; a++) {
        const double* Pn = pos[a] ? ph->Pp[a] : ph->Pm[a];
        double c = 0.0;
#pragma unroll
        for (int q = 0; q < 9; q++) c = fma(Pn[q], y[q], c);
        c *= dg[a];
#pragma unroll
        for (int q = 0; q < 9; q++) t[q] = fma(c, ph->P[a][q], t[q]);
      }
      double D[9];
      m_mul(invSubFp0, t, D);
#pragma unroll
      for (int q = 0; q < 9; q++) D[q] = -D[q] * dt;
      double R1[9], R2[9], R3[9], R4[9], col[9];
      m_mul(F, D, R1);
      m_mul(R1, M1, R2);            // F dFpinv S invFp^T
      m_mul(M2, y, R3);
      m_mul_nt(R3, invFp, R4);      // F invFp dSdF invFp^T
      m_mul_nt(M3, D, R1);          // F invFp S dFpinv^T
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double base = (i == k) ? Mid[3 * l + j] : 0.0;
          col[3 * i + j] = base + R2[3 * i + j] + R4[3 * i + j] + R1[3 * i + j];
        }
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          dPdF[(size_t)(i + 3 * j + 9 * k + 27 * l) * sN + e] = col[3 * i + j];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------- launchers
extern "C" int dmk_launch_stress(const DmkMPDev* mp, const double* Ftarget, const DmkPhaseDev* ph, int nsl,
                                 int phase_id, const DmkNumMech* num, double dt, hipStream_t s) {
  dim3 blk(128), grd((mp->N + 127) / 128);
  if (nsl == 12)
    hipLaunchKernelGGL(stress_kernel<12>, grd, blk, 0, s, *mp, Ftarget, ph, phase_id, *num, dt);
  else if (nsl == 24)
    hipLaunchKernelGGL(stress_kernel<24>, grd, blk, 0, s, *mp, Ftarget, ph, phase_id, *num, dt);
  else
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int dmk_launch_tangent(const DmkMPDev* mp, const DmkPhaseDev* ph, int nsl, int phase_id, double dt,
                                  const int* flag, double* Phom, double* dPdF, hipStream_t s) {
  dim3 blk(128), grd((mp->N + 127) / 128);
  if (nsl == 12)
    hipLaunchKernelGGL(tangent_kernel<12>, grd, blk, 0, s, *mp, ph, phase_id, dt, flag, Phom, dPdF);
  else if (nsl == 24)
    hipLaunchKernelGGL(tangent_kernel<24>, grd, blk, 0, s, *mp, ph, phase_id, dt, flag, Phom, dPdF);
  else
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
