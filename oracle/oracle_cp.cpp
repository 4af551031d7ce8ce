// CPU oracle, crystal plasticity part -- TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// The reference's CM-path return map lives in the un-vendored GCM library
// (GcmCpIntegrator, src/plasticity_model.cc:1173-1199).  This restatement follows the
// in-tree legacy integrator, which solves the same kind of system:
//   unknowns  M = pF(n) pF(n+1)^-1 (9), the incompressibility multiplier lambda and the
//             slip resistance g;
//   R_M      = M - I + dt sum_a gdot_a P_a - dt lambda pJ M^-T          (int_alg_res,
//              src/integration.cc:54-100, multiplied by dt)
//   R_lambda = pJ - 1,  pJ = det(M^-1 pF(n))                            (:88-95)
//   R_g      = g (gs - g0 + dt h0 gdot) - (gs - g0) g_n - dt h0 gs gdot (int_HA :395-403)
//   gdot_a   = gdot0 (tau_a/g) |tau_a/g|^(1/m-1),  gdot = sum |gdot_a|  (TA_GA,
//              src/TA_GA.cc:8-41; compute_drdtau src/plasticity_model.cc:567-576)
//   tau_a    = (C_e S) : P_a, C_e = eF^T eF, eF = F pF(n+1)^-1, S = L : (C_e - I)/2
//              (src/TA_GA.cc:21-31, get_SPK_stress src/stress_strain.cc:4-40)
// with P_a = s_a (x) m_a of the 12 FCC {111}<110> systems rotated by the grain rotation R,
// and L the cubic (C11, C12, C44) stiffness rotated by R -- or, for a material with
// cp_elastic = 1, S and L = dS/dE of the material's hyperelastic potentials (devPotFlag,
// volPotFlag) evaluated at eF, as the CM model builds its elasticity
// (src/constitutive_model.cc:998-1014, HyperElasticity::update_elasticity called on eF at
// src/plasticity_model.cc:747-772; the potentials are orc_he_update's).  The three equations are solved
// together by Newton with backtracking; GCM staggers g (max_itr_stag) around the M solve,
// which has the same converged point.  Plain start: the reference's fully plastic
// predictor (initial_guess, src/initial_guess.cc:100-110: the elastic part of step n is
// kept, M_p = pF(n) F(n+1)^-1 F(n) pF(n)^-1 scaled to det 1), pulled back along
// M(a) = I + a (M_p - I) to the first a at which max_a |tau_a|/g_n = 1 (8 bisection steps
// on [0, 1], upper end kept); a = 0 when the elastic trial state is below yield.  The consistent tangent dM/dF is the implicit
// derivative of the converged system (the role of plasticity_compute_dMdu,
// src/plasticity_model.cc:578-699).
//
// Pinning: tests/test_oracle_cp.py -- dM/dF and the element K against central finite
// differences of the re-solved return map, det pF = 1, slip off (gdot0 -> 0) equals
// St Venant-Kirchhoff elasticity, the hardening relation satisfied at convergence, and a
// single-crystal uniaxial test against an independent scipy solve of the same equations.
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

inline int id4(int i, int j, int k, int l) { return ((i * 3 + j) * 3 + k) * 3 + l; }

double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) + m[1] * (m[5] * m[6] - m[3] * m[8]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}
bool inv3(const double *m, double *mi) {
  const double A = det3(m);
  if (!(std::fabs(A) >= 1.0e-15)) return false;
  mi[0] = (m[4] * m[8] - m[5] * m[7]) / A;
  mi[1] = -(m[1] * m[8] - m[2] * m[7]) / A;
  mi[2] = (m[1] * m[5] - m[2] * m[4]) / A;
  mi[3] = -(m[3] * m[8] - m[5] * m[6]) / A;
  mi[4] = (m[0] * m[8] - m[2] * m[6]) / A;
  mi[5] = -(m[0] * m[5] - m[2] * m[3]) / A;
  mi[6] = (m[3] * m[7] - m[4] * m[6]) / A;
  mi[7] = -(m[0] * m[7] - m[1] * m[6]) / A;
  mi[8] = (m[0] * m[4] - m[1] * m[3]) / A;
  return true;
}
void mm(const double *A, const double *B, double *C) {
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
  std::memcpy(C, T, sizeof(T));
}
void mtm(const double *A, const double *B, double *C) {  // A^T B
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
  std::memcpy(C, T, sizeof(T));
}
void mmt(const double *A, const double *B, double *C) {  // A B^T
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
  std::memcpy(C, T, sizeof(T));
}
void L_dot(const double *L, const double *E, double *S) {  // S_ij = L_ijkl E_kl
  for (int i = 0; i < 9; ++i) {
    double s = 0.0;
    for (int k = 0; k < 9; ++k) s += L[i * 9 + k] * E[k];
    S[i] = s;
  }
}

// Dense LU with partial pivoting, n <= 16; returns false if singular.
bool lu_factor(double *A, int *piv, int n) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > amax) { amax = std::fabs(A[i * n + k]); p = i; }
    if (!(amax > 0.0)) return false;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
    const double r = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = A[i * n + k] * r;
      A[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
    }
  }
  return true;
}
void lu_solve(const double *A, const int *piv, int n, double *b) {
  // lu_factor swaps whole rows (multipliers included), so P is applied to b first
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) b[i] -= A[i * n + k] * b[k];
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
}

struct Grain {
  double P[12][9];  // rotated Schmid tensors s (x) m
  double L[81];     // rotated cubic stiffness
};

// Elastic PK2 S and tangent dS/dE of the intermediate configuration at eF (returns the
// tangent: the grain's rotated cubic stiffness, or Lbuf filled from the potentials).
const double *elastic_SL(const orc_material &m, const Grain &gr, const double *eF, double *S, double *Lbuf,
                         bool *ok) {
  *ok = true;
  if (m.cp_elastic == 1) {
    *ok = orc_he_update(&m, eF, S, Lbuf) == 0;
    return Lbuf;
  }
  double Ce[9], E[9];
  mtm(eF, eF, Ce);
  for (int i = 0; i < 9; ++i) E[i] = 0.5 * (Ce[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  L_dot(gr.L, E, S);
  return gr.L;
}

struct CPEval {
  double eF[9], Ce[9], S[9], tau[12], gdot[12], dgdt[12], dgdg[12];
  double Dtau[12][9];  // d tau_a / d M
  double Mi[9], pJ, gsum, gs, dgs;
  double R[11];
};

const int kN = 11;

// Residual (and the pieces of the Jacobian) at x = (M, lambda, g).
bool cp_eval(const orc_material &m, const Grain &gr, const double *A /*F pFn^-1*/, double det_pFn,
             const double *g_n_p, double dt, const double *x, CPEval &ev) {
  const double *M = x;
  const double lam = x[9], g = x[10];
  if (!(g > 0.0)) return false;
  if (!inv3(M, ev.Mi)) return false;
  const double dM = det3(M);
  if (!(dM > 0.0)) return false;
  ev.pJ = det_pFn / dM;
  mm(A, M, ev.eF);
  mtm(ev.eF, ev.eF, ev.Ce);
  double Lbuf[81];
  bool sok = true;
  const double *Lp = elastic_SL(m, gr, ev.eF, ev.S, Lbuf, &sok);
  if (!sok) return false;
  double Sig[9];
  mm(ev.Ce, ev.S, Sig);  // Mandel stress C_e S
  const double expo = 1.0 / m.m_rate - 1.0;
  ev.gsum = 0.0;
  double AtA_eF[9];
  mtm(A, ev.eF, AtA_eF);  // A^T eF
  for (int a = 0; a < 12; ++a) {
    double t = 0.0;
    for (int i = 0; i < 9; ++i) t += Sig[i] * gr.P[a][i];
    ev.tau[a] = t;
    const double r = t / g;
    const double pw = std::pow(std::fabs(r), expo);
    ev.gdot[a] = m.gamma_dot_0 * r * pw;
    ev.dgdt[a] = m.gamma_dot_0 / (m.m_rate * g) * pw;
    ev.dgdg[a] = -ev.dgdt[a] * t / g;
    ev.gsum += std::fabs(ev.gdot[a]);
    // Q = P S + 1/2 L:(C_e P); Dtau = 2 A^T eF sym(Q)
    double PS[9], CP[9], LCP[9], Q[9];
    mm(gr.P[a], ev.S, PS);
    mm(ev.Ce, gr.P[a], CP);
    L_dot(Lp, CP, LCP);
    for (int i = 0; i < 9; ++i) Q[i] = PS[i] + 0.5 * LCP[i];
    double sQ[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) sQ[i * 3 + j] = 0.5 * (Q[i * 3 + j] + Q[j * 3 + i]);
    double D[9];
    mm(AtA_eF, sQ, D);
    for (int i = 0; i < 9; ++i) ev.Dtau[a][i] = 2.0 * D[i];
  }
  // saturation stress gs = gs0 (gdot/gdot_s)^w, guarded as src/plasticity_model.cc:617-622
  if (m.w_sat == 0.0) {
    ev.gs = m.gs_0;
    ev.dgs = 0.0;
  } else if (ev.gsum / m.gamma_dot_s > 1.0e-15) {
    ev.gs = m.gs_0 * std::pow(ev.gsum / m.gamma_dot_s, m.w_sat);
    ev.dgs = m.w_sat * ev.gs / ev.gsum;
  } else {
    ev.gs = 0.0;
    ev.dgs = 0.0;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 12; ++a) s += ev.gdot[a] * gr.P[a][i * 3 + j];
      ev.R[i * 3 + j] = M[i * 3 + j] - ((i == j) ? 1.0 : 0.0) + dt * s - dt * lam * ev.pJ * ev.Mi[j * 3 + i];
    }
  ev.R[9] = ev.pJ - 1.0;
  const double g_n = *g_n_p, h0 = m.G0, g0 = m.g0;
  ev.R[10] = g * (ev.gs - g0 + dt * h0 * ev.gsum) - (ev.gs - g0) * g_n - dt * h0 * ev.gs * ev.gsum;
  return true;
}

// Jacobian dR/dx at the evaluated point (row-major 11x11).
void cp_jacobian(const orc_material &m, const Grain &gr, double dt, const double *x, const CPEval &ev,
                 double g_n, double *J) {
  const double lam = x[9], g = x[10];
  std::memset(J, 0, sizeof(double) * kN * kN);
  const double *Mi = ev.Mi;
  // R_M rows
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int r = i * 3 + j;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const int c = k * 3 + l;
          double v = (i == k && j == l) ? 1.0 : 0.0;
          double s = 0.0;
          for (int a = 0; a < 12; ++a) s += gr.P[a][r] * ev.dgdt[a] * ev.Dtau[a][c];
          v += dt * s;
          v += dt * lam * ev.pJ * (Mi[j * 3 + i] * Mi[l * 3 + k] + Mi[l * 3 + i] * Mi[j * 3 + k]);
          J[r * kN + c] = v;
        }
      J[r * kN + 9] = -dt * ev.pJ * Mi[j * 3 + i];
      double s = 0.0;
      for (int a = 0; a < 12; ++a) s += gr.P[a][r] * ev.dgdg[a];
      J[r * kN + 10] = dt * s;
    }
  // R_lambda
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) J[9 * kN + k * 3 + l] = -ev.pJ * Mi[l * 3 + k];
  // R_g
  const double h0 = m.G0, g0 = m.g0;
  const double dRg_dgsum = g * (ev.dgs + dt * h0) - ev.dgs * g_n - dt * h0 * (ev.dgs * ev.gsum + ev.gs);
  for (int c = 0; c < 9; ++c) {
    double s = 0.0;
    for (int a = 0; a < 12; ++a) s += ((ev.gdot[a] < 0) ? -1.0 : 1.0) * ev.dgdt[a] * ev.Dtau[a][c];
    J[10 * kN + c] = dRg_dgsum * s;
  }
  double sg = 0.0;
  for (int a = 0; a < 12; ++a) sg += ((ev.gdot[a] < 0) ? -1.0 : 1.0) * ev.dgdg[a];
  J[10 * kN + 10] = (ev.gs - g0 + dt * h0 * ev.gsum) + dRg_dgsum * sg;
}

// max_a |tau_a| / g at M (elastic part eF = A M).
double tau_ratio_max(const orc_material &m, const Grain &gr, const double *A, const double *M, double g) {
  double eF[9], Ce[9], S[9], Sig[9], Lbuf[81];
  mm(A, M, eF);
  mtm(eF, eF, Ce);
  bool ok = true;
  elastic_SL(m, gr, eF, S, Lbuf, &ok);
  if (!ok) return 2.0;  // no elastic state: treat as beyond yield
  mm(Ce, S, Sig);
  double mx = 0.0;
  for (int a = 0; a < 12; ++a) {
    double t = 0.0;
    for (int i = 0; i < 9; ++i) t += Sig[i] * gr.P[a][i];
    mx = std::fmax(mx, std::fabs(t));
  }
  return mx / g;
}

// Plain start x = (M, lambda_n, g_n): the fully plastic predictor pulled back to yield.
bool cp_start(const orc_material &m, const Grain &gr, const double *A, const double *F, const double *Fn,
              const double *pFn, const double *pFni, double g_n, double lam_n, double *x) {
  double Fi[9], T[9], Mp[9];
  if (!inv3(F, Fi)) return false;
  mm(pFn, Fi, T);
  mm(T, Fn, Mp);
  mm(Mp, pFni, T);
  const double dp = det3(T);
  if (!(dp > 0.0)) return false;
  const double c = 1.0 / std::cbrt(dp);
  for (int i = 0; i < 9; ++i) Mp[i] = T[i] * c;
  const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double al = 0.0;
  if (tau_ratio_max(m, gr, A, I9, g_n) > 1.0) {
    al = 1.0;
    double lo = 0.0, hi = 1.0;
    if (!(tau_ratio_max(m, gr, A, Mp, g_n) > 1.0)) {
      for (int k = 0; k < 8; ++k) {
        const double mid = 0.5 * (lo + hi);
        double Mm[9];
        for (int i = 0; i < 9; ++i) Mm[i] = I9[i] + mid * (Mp[i] - I9[i]);
        if (tau_ratio_max(m, gr, A, Mm, g_n) > 1.0) lo = mid;
        else hi = mid;
      }
      al = hi;
    }
  }
  for (int i = 0; i < 9; ++i) x[i] = I9[i] + al * (Mp[i] - I9[i]);
  if (al > 0.0) {
    const double cm = 1.0 / std::cbrt(det3(x));
    for (int i = 0; i < 9; ++i) x[i] *= cm;
  }
  x[9] = lam_n;
  x[10] = g_n;
  return true;
}

// Line-search merit: half the squared 2-norm of the scaled residual (the Newton direction
// is a descent direction for it).
double merit(const CPEval &ev, double gscale) {
  double r = 0.0;
  for (int i = 0; i < 10; ++i) r += ev.R[i] * ev.R[i];
  const double rg = ev.R[10] * gscale;
  return 0.5 * (r + rg * rg);
}

}  // namespace

extern "C" {

// FCC {111}<110> slip systems in the crystal frame (unit vectors), P0[a*9 + i*3+j] = s_i m_j.
void orc_cp_slip_systems(double *P0) {
  const double n[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
  const double s[12][3] = {{0, 1, -1}, {-1, 0, 1}, {1, -1, 0}, {0, 1, -1}, {1, 0, 1}, {1, 1, 0},
                           {0, 1, 1},  {1, 0, -1}, {1, 1, 0},  {0, 1, 1},  {1, 0, 1}, {1, -1, 0}};
  const double rn = 1.0 / std::sqrt(3.0), rs = 1.0 / std::sqrt(2.0);
  for (int a = 0; a < 12; ++a)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) P0[a * 9 + i * 3 + j] = (s[a][i] * rs) * (n[a / 3][j] * rn);
}

// Grain tables: Schmid tensors and cubic stiffness rotated to the sample frame by R
// (crystal -> sample): P = R P0 R^T, L_ijkl = R_ia R_jb R_kc R_ld L0_abcd.
void orc_cp_grain(const orc_material *mat, const double *R, double *P_out, double *L_out) {
  double P0[108];
  orc_cp_slip_systems(P0);
  for (int a = 0; a < 12; ++a) {
    double T[9];
    mm(R, P0 + a * 9, T);
    mmt(T, R, P_out + a * 9);
  }
  double L0[81];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double v = 0.0;
          if (i == j && k == l) v += (i == k) ? mat->C11 : mat->C12;
          if (i != j && ((i == k && j == l) || (i == l && j == k))) v += mat->C44;
          L0[id4(i, j, k, l)] = v;
        }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double v = 0.0;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d)
                  v += R[i * 3 + a] * R[j * 3 + b] * R[k * 3 + c] * R[l * 3 + d] * L0[id4(a, b, c, d)];
          L_out[id4(i, j, k, l)] = v;
        }
}

// One QP: solve the return map and return the n+1 state, the elastic PK2 S (intermediate
// configuration), tau_a, gdot_a and dM/dF (row-major [9 M comps][9 F comps]).
// x0 = optional starting point (M 9, lambda, g) -- NULL = (I, lambda_n, g_n); with solve = 0
// x0 is taken as the converged point (tangent from a stored state).
// Returns 0 on success, 1 on failure (non-invertible or no convergence).
int orc_cp_update(const orc_material *mat, const double *P_grain, const double *L_grain, const double *F,
                  const double *Fn, const double *pFn, double g_n, double lam_n, double dt, const double *x0, int solve,
                  double *pFnp1, double *g_np1, double *lam_np1, double *S, double *tau, double *gdot,
                  double *dMdF, int *iters) {
  const orc_material &m = *mat;
  Grain gr;
  std::memcpy(gr.P, P_grain, sizeof(gr.P));
  std::memcpy(gr.L, L_grain, sizeof(gr.L));
  double pFni[9], A[9];
  if (!inv3(pFn, pFni)) return 1;
  mm(F, pFni, A);
  const double det_pFn = det3(pFn);
  double x[kN];
  if (x0) {
    std::memcpy(x, x0, sizeof(x));
  } else {
    const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (!cp_start(m, gr, A, F, Fn ? Fn : I9, pFn, pFni, g_n, lam_n, x)) return 1;
  }
  const double tol = (m.tol_M > 0.0) ? m.tol_M : 1.0e-10;
  const int maxit = (m.max_itr_M > 0) ? m.max_itr_M : 50;
  const double gscale = 1.0 / (std::fabs(m.gs_0 - m.g0) * g_n + 1e-300);
  CPEval ev;
  if (!cp_eval(m, gr, A, det_pFn, &g_n, dt, x, ev)) return 1;
  double J[kN * kN];
  int piv[kN];
  int it = 0;
  bool conv = !solve;  // solve == 0: x0 is the converged point, only the tangent is wanted
  while (solve && it < maxit) {
    ++it;
    cp_jacobian(m, gr, dt, x, ev, g_n, J);
    double rhs[kN];
    for (int i = 0; i < kN; ++i) rhs[i] = -ev.R[i];
    if (!lu_factor(J, piv, kN)) return 1;
    double dx[kN];
    for (int i = 0; i < kN; ++i) dx[i] = rhs[i];
    lu_solve(J, piv, kN, dx);
    // size of the full Newton step; below tol the iteration has converged (the line search
    // is skipped: at round-off level the merit need not decrease)
    double du_full = 0.0;
    for (int i = 0; i < 9; ++i) du_full = std::fmax(du_full, std::fabs(dx[i]));
    du_full = std::fmax(du_full, std::fabs(dx[9]) * dt);
    du_full = std::fmax(du_full, std::fabs(dx[10]) / g_n);
    if (du_full < tol) {
      for (int i = 0; i < kN; ++i) x[i] += dx[i];
      if (!cp_eval(m, gr, A, det_pFn, &g_n, dt, x, ev)) return 1;
      conv = true;
      break;
    }
    const double f0 = merit(ev, gscale);
    double alpha = 1.0;
    double xn[kN];
    CPEval evn;
    bool ok = false;
    for (int ls = 0; ls < 40; ++ls) {
      for (int i = 0; i < kN; ++i) xn[i] = x[i] + alpha * dx[i];
      if (cp_eval(m, gr, A, det_pFn, &g_n, dt, xn, evn) && merit(evn, gscale) <= (1.0 - 1e-4 * alpha) * f0) {
        ok = true;
        break;
      }
      alpha *= 0.5;
    }
    if (!ok) {  // no decrease found: take the full step if it evaluates
      alpha = 1.0;
      for (int i = 0; i < kN; ++i) xn[i] = x[i] + dx[i];
      if (!cp_eval(m, gr, A, det_pFn, &g_n, dt, xn, evn)) return 1;
    }
    std::memcpy(x, xn, sizeof(x));
    ev = evn;
  }
  if (iters) *iters = it;
  if (!conv) return 1;
  // state
  double Mi[9];
  inv3(x, Mi);
  mm(Mi, pFn, pFnp1);  // pF(n+1) = M^-1 pF(n)
  *g_np1 = x[10];
  *lam_np1 = x[9];
  std::memcpy(S, ev.S, sizeof(ev.S));
  std::memcpy(tau, ev.tau, sizeof(ev.tau));
  std::memcpy(gdot, ev.gdot, sizeof(ev.gdot));
  if (dMdF) {
    // dx/dF = -J^-1 dR/dF at the converged point
    cp_jacobian(m, gr, dt, x, ev, g_n, J);
    if (!lu_factor(J, piv, kN)) return 1;
    const double h0 = m.G0;
    const double dRg_dgsum = x[10] * (ev.dgs + dt * h0) - ev.dgs * g_n - dt * h0 * (ev.dgs * ev.gsum + ev.gs);
    // dtau_a/dF = 2 eF sym(Q_a) (pFn^-1 M)^T = Dtau_a-like: reuse Q through Dtau = 2 A^T eF sQ
    // -> 2 eF sQ = A^-T Dtau;  dtau/dF = A^-T Dtau (pFn^-1 M)^T
    double Ai[9], AiT[9], B[9];
    inv3(A, Ai);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) AiT[i * 3 + j] = Ai[j * 3 + i];
    mm(pFni, x, B);  // pFn^-1 M
    double dtdF[12][9];
    for (int a = 0; a < 12; ++a) {
      double T[9];
      mm(AiT, ev.Dtau[a], T);
      mmt(T, B, dtdF[a]);
    }
    for (int c = 0; c < 9; ++c) {
      double rhs[kN];
      for (int r = 0; r < 9; ++r) {
        double s = 0.0;
        for (int a = 0; a < 12; ++a) s += gr.P[a][r] * ev.dgdt[a] * dtdF[a][c];
        rhs[r] = -dt * s;
      }
      rhs[9] = 0.0;
      double s = 0.0;
      for (int a = 0; a < 12; ++a) s += ((ev.gdot[a] < 0) ? -1.0 : 1.0) * ev.dgdt[a] * dtdF[a][c];
      rhs[10] = -dRg_dgsum * s;
      lu_solve(J, piv, kN, rhs);
      for (int r = 0; r < 9; ++r) dMdF[r * 9 + c] = rhs[r];
    }
  }
  return 0;
}

}  // extern "C"

extern "C" {

// fd_residuals_MP / stiffmat_fd_MP for CP_PARAM elements (TL, src/constitutive_model.cc
// :2474-2692 and :3322-3539): M = pF(n+1)^-1, eFn = I, pJ = det pF(n+1), S the elastic PK2,
// L the elastic stiffness, dM/du from the return-map linearisation.  Grain tables are
// [n_grain][12*9] Schmid tensors and [n_grain][81] stiffness; the state arrays are
// [comp][n_qp] in mesh QP order.
int orc_assemble_cp(int64_t n_nodes, const double *X, int64_t n_elem, const int64_t *elem_ptr,
                    const int64_t *elem_conn, const int32_t *elem_mat, const orc_material *mats,
                    const int32_t *elem_grain, const double *grain_P, const double *grain_L,
                    const int64_t *id, int64_t ndof, const int32_t *Ap, const int32_t *Ai, const double *r,
                    const double *d_r, const double *defl, const double *defl_d, double dt, int do_residual,
                    int run_integration, int do_tangent, double *F_np1, const double *pF_n, double *pF_np1,
                    const double *g_n, double *g_np1, const double *lam_n, double *lam_np1, double *tau_out,
                    double *gdot_out, int32_t *iters_out, const double *F_n, const double *warm, double *f, double *vals);
}

namespace {
void matmul_plain(const double *A, const double *B, double *C) { mm(A, B, C); }
}

extern "C" int orc_assemble_cp(int64_t n_nodes, const double *X, int64_t n_elem, const int64_t *elem_ptr,
                               const int64_t *elem_conn, const int32_t *elem_mat, const orc_material *mats,
                               const int32_t *elem_grain, const double *grain_P, const double *grain_L,
                               const int64_t *id, int64_t ndof, const int32_t *Ap, const int32_t *Ai,
                               const double *r, const double *d_r, const double *defl, const double *defl_d,
                               double dt, int do_residual, int run_integration, int do_tangent, double *F_np1,
                               const double *pF_n, double *pF_np1, const double *g_n, double *g_np1,
                               const double *lam_n, double *lam_np1, double *tau_out, double *gdot_out,
                               int32_t *iters_out, const double *F_n, const double *warm, double *f, double *vals) {
  // elements colour by colour, in parallel within a colour (as orc_assemble)
  std::vector<int64_t> qoff(n_elem + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) qoff[e + 1] = qoff[e] + orc_nint(int(elem_ptr[e + 1] - elem_ptr[e]));
  const int64_t NQ = qoff[n_elem];
  std::vector<int64_t> order(n_elem), cptr(n_elem + 1);
  const int nc = orc_colour_elements(n_nodes, n_elem, elem_ptr, elem_conn, order.data(), cptr.data());
  if (do_residual && f) std::memset(f, 0, sizeof(double) * ndof);
  if (do_tangent && vals) std::memset(vals, 0, sizeof(double) * Ap[ndof]);
  int err = 0, missing = 0;
  auto element = [&](int64_t e) -> int {
    int err = 0;
    const int64_t qp0 = qoff[e];
    std::vector<double> lk;
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]);
    const int nint = orc_nint(nne);
    const int ndofe = nne * 3;
    const int64_t *nod = elem_conn + elem_ptr[e];
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    const double *gP = grain_P + int64_t(elem_grain[e]) * 108;
    const double *gL = grain_L + int64_t(elem_grain[e]) * 81;
    double x[30], u[30];
    int64_t cn[30];
    for (int a = 0; a < nne; ++a)
      for (int b = 0; b < 3; ++b) {
        x[a * 3 + b] = X[nod[a] * 3 + b];
        const int64_t dof = id[nod[a] * 3 + b];
        cn[a * 3 + b] = dof;
        const int64_t aid = (dof < 0 ? -dof : dof) - 1;
        if (dof == 0) u[a * 3 + b] = 0.0;
        else if (dof > 0) u[a * 3 + b] = r[aid] + (d_r ? d_r[aid] : 0.0);
        else u[a * 3 + b] = (defl ? defl[aid] : 0.0) + (defl_d ? defl_d[aid] : 0.0);
      }
    double fe[30] = {0};
    lk.assign(size_t(ndofe) * ndofe, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 0 && !do_residual) continue;
      if (pass == 1 && !do_tangent) continue;
      for (int ip = 0; ip < nint; ++ip) {
        const int64_t q = qp0 + ip;
        double N[10], dNdX[30], detJxW;
        if (orc_basis(nne, x, ip, N, dNdX, &detJxW)) { ++err; continue; }
        double ST[270];
        for (int a = 0; a < nne; ++a)
          for (int b = 0; b < 3; ++b)
            for (int i = 0; i < 3; ++i)
              for (int J = 0; J < 3; ++J) ST[a * 27 + b * 9 + i * 3 + J] = (b == i) ? dNdX[a * 3 + J] : 0.0;
        double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int a = 0; a < nne; a++)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) Fr[b * 3 + c] += u[a * 3 + b] * dNdX[a * 3 + c];
        double pFn[9], pF1[9], Fs[9];
        for (int i = 0; i < 9; ++i) pFn[i] = pF_n[i * NQ + q];
        if (pass == 0) {
          if (!(det3(Fr) > 0.0)) { ++err; continue; }
          if (run_integration) {
            for (int i = 0; i < 9; ++i) F_np1[i * NQ + q] = Fr[i];
            double S0[9], t12[12], gd12[12], g1, l1;
            int its = 0;
            double x0[11], Fn[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            const double *x0p = nullptr;
            if (warm) {  // caller-provided start (M, lambda, g)
              for (int i = 0; i < 11; ++i) x0[i] = warm[i * NQ + q];
              x0p = x0;
            }
            if (F_n)
              for (int i = 0; i < 9; ++i) Fn[i] = F_n[i * NQ + q];
            if (orc_cp_update(&mat, gP, gL, Fr, Fn, pFn, g_n[q], lam_n[q], dt, x0p, 1, pF1, &g1, &l1, S0, t12, gd12,
                              nullptr, &its)) {
              ++err;
              if (iters_out) iters_out[q] = its;
              continue;
            }
            for (int i = 0; i < 9; ++i) pF_np1[i * NQ + q] = pF1[i];
            g_np1[q] = g1;
            lam_np1[q] = l1;
            for (int a = 0; a < 12; ++a) {
              tau_out[a * NQ + q] = t12[a];
              gdot_out[a * NQ + q] = gd12[a];
            }
            if (iters_out) iters_out[q] = its;
          }
        }
        for (int i = 0; i < 9; ++i) {
          Fs[i] = F_np1[i * NQ + q];
          pF1[i] = pF_np1[i * NQ + q];
        }
        double Me[9];
        if (!inv3(pF1, Me)) { ++err; continue; }  // M = pF(n+1)^-1
        const double pJ = det3(pF1);
        double MeT[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) MeT[i * 3 + j] = Me[j * 3 + i];
        // elastic PK2 and tangent from the stored state
        double eF[9], S[9], Lbuf[81];
        matmul_plain(Fs, Me, eF);
        Grain gel;
        std::memcpy(gel.L, gL, sizeof(gel.L));
        bool sok = true;
        const double *Lel = elastic_SL(mat, gel, eF, S, Lbuf, &sok);
        if (!sok) { ++err; continue; }
        if (pass == 0) {
          for (int a = 0; a < nne; a++)  // compute_residual_vector
            for (int b = 0; b < 3; b++) {
              const double *STab = ST + a * 27 + b * 9;
              double AA[9], sAA[9], T[9], V[9];
              mtm(Fr, STab, AA);
              for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) sAA[i * 3 + j] = 0.5 * (AA[i * 3 + j] + AA[j * 3 + i]);
              mm(MeT, sAA, T);
              mm(T, Me, V);
              double v = 0.0;
              for (int i = 0; i < 9; ++i) v += V[i] * S[i];
              fe[a * 3 + b] += pJ * detJxW * v;
            }
        } else {
          // dM/dF at the stored converged state, then dMe/du = pFn^-1 dM/du
          double pFni[9], Mcp[9], x0[11], dMdF[81], S1[9], t12[12], gd12[12], g1, l1, pF2[9];
          inv3(pFn, pFni);
          matmul_plain(pFn, Me, Mcp);
          for (int i = 0; i < 9; ++i) x0[i] = Mcp[i];
          x0[9] = lam_np1[q];
          x0[10] = g_np1[q];
          if (orc_cp_update(&mat, gP, gL, Fs, nullptr, pFn, g_n[q], lam_n[q], dt, x0, 0, pF2, &g1, &l1, S1, t12, gd12, dMdF,
                            nullptr)) { ++err; continue; }
          double dMdu_all[270];
          for (int w = 0; w < nne; ++w)
            for (int g = 0; g < 3; ++g) {
              double dM[9], dMe[9];
              for (int rr = 0; rr < 9; ++rr) {
                double s = 0.0;
                for (int Lx = 0; Lx < 3; ++Lx) s += dMdF[rr * 9 + g * 3 + Lx] * dNdX[w * 3 + Lx];
                dM[rr] = s;
              }
              matmul_plain(pFni, dM, dMe);
              for (int rr = 0; rr < 9; ++rr) dMdu_all[w * 27 + g * 9 + rr] = dMe[rr];
            }
          const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
          double FrTFr[9];
          mtm(Fr, Fr, FrTFr);
          orc_stiffness_contract(lk.data(), nne, ST, detJxW, Fr, MeT, I9, Me, FrTFr, Me, S, Lel, dMdu_all, pJ);
        }
      }
    }
    if (do_tangent)
      for (int a = 0; a < ndofe; ++a)
        if (!std::isnormal(lk[a * ndofe + a])) ++err;
    if (do_residual && f)
      for (int k = 0; k < ndofe; ++k) {
        const int64_t II = cn[k] - 1;
        if (II >= 0) f[II] += fe[k];
      }
    if (do_tangent && vals)
      for (int i = 0; i < ndofe; ++i) {
        const int64_t row = cn[i] - 1;
        if (row < 0) continue;
        for (int j = 0; j < ndofe; ++j) {
          const int64_t col = cn[j] - 1;
          if (col < 0) continue;
          const int32_t *b = Ai + Ap[row], *en = Ai + Ap[row + 1];
          const int32_t *p = b;
          while (p < en && *p != col) ++p;
          if (p == en) return -1;
          vals[p - Ai] += lk[i * ndofe + j];
        }
      }
    return err;
  };
  for (int c = 0; c < nc; ++c) {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : err, missing)
    for (int64_t k = cptr[c]; k < cptr[c + 1]; ++k) {
      const int r = element(order[k]);
      if (r < 0) ++missing;
      else err += r;
    }
  }
  return missing ? -1 : err;
}

// Output stress of one QP from its stored state (constitutive_model_update_output_variables,
// src/constitutive_model.cc:1738-1875): the model's elastic stress on eF, the Cauchy stress
// sigma = eF eS eF^T / theta_e (theta_e = det eF) and the PK2 S = J F^-1 sigma F^-T.
//   model 0 (hyperelastic): eF = F, eS from the potentials (orc_he_update);
//   model 4 (J2, state Cp^-1): eS pulled back to the reference configuration directly,
//     S = (1 - w) [G J^-2/3 (Cp^-1 - tr(C Cp^-1)/3 C^-1) + kappa J U'(J) C^-1]
//     (the stress of the elastic left stretch b_e = F Cp^-1 F^T; w the damage, 0 without);
//   model 1 (crystal plasticity, state pF): eF = F pF^-1, eS the grain's elastic law.
// L_grain is the grain's rotated cubic stiffness [81] (law 0 of crystal plasticity).
extern "C" int orc_output_qp(const orc_material *mat, const double *L_grain, const double *F, const double *plastic,
                             double w, double *S_out, double *sig_out) {
  const orc_material &m = *mat;
  const double J = det3(F);
  double Fi[9];
  if (!(J > 0.0) || !inv3(F, Fi)) return 1;
  double sig[9];
  if (m.model == 1) {
    Grain gr{};
    std::memcpy(gr.L, L_grain, sizeof(gr.L));
    double pFi[9], eF[9], Se[9], Lb[81], T[9];
    if (!inv3(plastic, pFi)) return 1;
    mm(F, pFi, eF);
    bool ok = true;
    elastic_SL(m, gr, eF, Se, Lb, &ok);
    if (!ok) return 1;
    const double te = det3(eF);
    mm(eF, Se, T);
    mmt(T, eF, sig);
    for (int i = 0; i < 9; ++i) sig[i] /= te;
  } else {
    double S[9], T[9];
    if (m.model == 4) {
      double C[9], CI[9], Cp[9], CCp[9];
      mtm(F, F, C);
      if (!inv3(C, CI)) return 1;
      Cp[0] = plastic[0]; Cp[4] = plastic[1]; Cp[8] = plastic[2];
      Cp[1] = Cp[3] = plastic[3];
      Cp[5] = Cp[7] = plastic[4];
      Cp[2] = Cp[6] = plastic[5];
      mm(C, Cp, CCp);
      const double tr3 = (CCp[0] + CCp[4] + CCp[8]) / 3.0;
      const double kappa = m.E / (3.0 * (1.0 - 2.0 * m.nu));
      double dU;
      const int flag = m.volPotFlag;
      if (flag == 3) dU = (-1 + J) / (2. * J) + std::log(J) * .5;  // src/new_potentials.cc:575-616
      else if (flag == 99) dU = J - 1;
      else dU = (std::exp(-1 + J) - 1. / J) * .5;
      const double Jm23 = std::pow(J, -2.0 / 3.0);
      for (int i = 0; i < 9; ++i) S[i] = (1.0 - w) * (m.G * Jm23 * (Cp[i] - tr3 * CI[i]) + kappa * J * dU * CI[i]);
    } else {
      if (orc_he_update(&m, F, S, nullptr)) return 1;
    }
    mm(F, S, T);
    mmt(T, F, sig);
    for (int i = 0; i < 9; ++i) sig[i] /= J;
  }
  double T2[9], S2[9];
  mm(Fi, sig, T2);
  mmt(T2, Fi, S2);
  for (int i = 0; i < 9; ++i) {
    S_out[i] = J * S2[i];
    sig_out[i] = sig[i];
  }
  return 0;
}
