// Crystal-plasticity kernels for gfx950 (FP64).
//
// Constitutive update per QP (the role of CP_PARAM::integration_algorithm,
// src/plasticity_model.cc:1148-1224, whose return map lives in the un-vendored GCM library):
//   unknowns x = (M, lambda, g), M = pF(n) pF(n+1)^-1
//   R_M      = M - I + dt sum_a gdot_a P_a - dt lambda pJ M^-T        (src/integration.cc:54-100)
//   R_lambda = pJ - 1,  pJ = det(pF(n)) / det(M)                      (:88-95)
//   R_g      = g (gs - g0 + dt h0 gdot) - (gs - g0) g_n - dt h0 gs gdot (:395-403)
//   gdot_a   = gdot0 (tau_a/g)|tau_a/g|^(1/m-1), tau_a = (C_e S):P_a,
//   C_e = eF^T eF, eF = F pF(n)^-1 M, and S either the cubic S = L : (C_e - I)/2
//   (src/TA_GA.cc:8-41) or the hyperelastic law of the material line (he_law below); the
//   kernels are instantiated per law
// Plain start: the reference's fully plastic predictor (initial_guess,
// src/initial_guess.cc:100-110: elastic part of step n kept, M_p = pF(n) F^-1 F(n) pF(n)^-1
// scaled to det 1) pulled back along M(a) = I + a (M_p - I) to max |tau_a|/g_n = 1 (8
// bisection steps, upper end kept; a = 0 below yield).  Newton on all eleven unknowns with
// backtracking on half the squared residual norm; converged when the full Newton step is
// below tol.  The 11x11 system is
// solved through the Schur complement of its (M, M) block: with the bordered structure
//   [A b c; d^T 0 0; e^T 0 f]
// one 9x9 LU (register-resident, pivoting by predicated row swaps so nothing is indexed at
// run time) serves the two right-hand sides of a Newton step and the nine of the tangent.
//
// Kernels:
//   cp_kinematics  F(n+1) at every QP of a block (gather of the element nodes)
//   cp_start_kernel  the starting points (and the choice of a caller's start), one per lane
//   cp_newton      persistent: each wave owns a QP range; a lane that converges takes the
//                  next QP of its wave at once, so uneven iteration counts do not idle lanes
//   cp_tangent     at the converged state: S, tau, gdot and dMe/dF (Me = pF(n+1)^-1), for the
//                  element kernel (qpbuf)
//   cp_elem        coloured element rows: the reference contraction (compute_stiffness_matrix,
//                  src/constitutive_model.cc:1520-1616) regrouped per row (a,b) as
//                  K_ab,wg = pJ w [ z_ab : (e_g (x) dN_w) + (Fr W_ab dN_w)_g + d_bg dN_a.Sh.dN_w ]
//                  with Z_ab = Cr Me Y + 2 sAA Me S - (B:S) Me^-T, z_ab = Z_ab : dMe/dF,
//                  B = Me^T sAA Me, Y = L:B, W = Me Y Me^T, Sh = Me S Me^T, Cr = Fr^T Fr.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pgf_device.h"
#include "pgf_math.h"

namespace pgf {
namespace {

using namespace dev;

constexpr int kWave = 64;
typedef double dbl2 __attribute__((ext_vector_type(2)));

// Voigt helpers: index I -> (i,j) in xx,yy,zz,xy,yz,xz
__device__ __forceinline__ int vi(int I) { return I < 3 ? I : (I == 3 ? 0 : (I == 4 ? 1 : 0)); }
__device__ __forceinline__ int vj(int I) { return I < 3 ? I : (I == 3 ? 1 : 2); }

__device__ __forceinline__ void mm3(const double *A, const double *B, double *C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
__device__ __forceinline__ void mtm3(const double *A, const double *B, double *C) {  // A^T B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
// S (Voigt tensor comps) = Lv : E, E symmetric 3x3 given as full 3x3
__device__ __forceinline__ void lv_dot(const double *Lv, const double *E9, double *S6) {
  const double e[6] = {E9[0], E9[4], E9[8], 0.5 * (E9[1] + E9[3]), 0.5 * (E9[5] + E9[7]), 0.5 * (E9[2] + E9[6])};
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    double s = 0.0;
#pragma unroll
    for (int J = 0; J < 6; ++J) s += Lv[I * 6 + J] * e[J];
    S6[I] = s;
  }
}
// S (Voigt tensor comps) = L : sym(X) for the grain's cubic stiffness, through the crystal
// frame: E' = R^T sym(X) R, S'_ii = (C11 - C12) E'_ii + C12 tr E', S'_ij = 2 C44 E'_ij,
// S = R S' R^T.  12 numbers of the grain (R and the constants) instead of the 36 of its rotated
// stiffness; the same tensor up to rounding.  Every multiply-add is an explicit fma and nothing
// else is contracted, so the result does not depend on the code around a call (the residual
// and the fused residual-tangent passes give the same bits).
__device__ __forceinline__ double fma3(double a0, double b0, double a1, double b1, double a2, double b2) {
#pragma clang fp contract(off)
  return __builtin_fma(a0, b0, __builtin_fma(a1, b1, a2 * b2));
}
__device__ __forceinline__ void cubic_dot(const double *R, const double *cub, const double *X9, double *S6) {
#pragma clang fp contract(off)
  double E[9], T[9], Ep[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) E[i * 3 + j] = 0.5 * (X9[i * 3 + j] + X9[j * 3 + i]);
#pragma unroll
  for (int i = 0; i < 3; ++i)  // T = E R
#pragma unroll
    for (int j = 0; j < 3; ++j) T[i * 3 + j] = fma3(E[i * 3 + 0], R[0 * 3 + j], E[i * 3 + 1], R[1 * 3 + j], E[i * 3 + 2], R[2 * 3 + j]);
#pragma unroll
  for (int i = 0; i < 3; ++i)  // E' = R^T T (symmetric: the upper triangle)
#pragma unroll
    for (int j = i; j < 3; ++j) Ep[i * 3 + j] = fma3(R[0 * 3 + i], T[0 * 3 + j], R[1 * 3 + i], T[1 * 3 + j], R[2 * 3 + i], T[2 * 3 + j]);
  const double tr = (Ep[0] + Ep[4]) + Ep[8];
  const double ct = cub[1] * tr;
  double Sp[9];
  Sp[0] = __builtin_fma(cub[0], Ep[0], ct);
  Sp[4] = __builtin_fma(cub[0], Ep[4], ct);
  Sp[8] = __builtin_fma(cub[0], Ep[8], ct);
  Sp[1] = Sp[3] = cub[2] * Ep[1];
  Sp[2] = Sp[6] = cub[2] * Ep[2];
  Sp[5] = Sp[7] = cub[2] * Ep[5];
#pragma unroll
  for (int i = 0; i < 3; ++i)  // T = S' R^T
#pragma unroll
    for (int j = 0; j < 3; ++j) T[i * 3 + j] = fma3(Sp[i * 3 + 0], R[j * 3 + 0], Sp[i * 3 + 1], R[j * 3 + 1], Sp[i * 3 + 2], R[j * 3 + 2]);
  double S[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)  // S = R T (symmetric: the upper triangle)
#pragma unroll
    for (int j = i; j < 3; ++j) S[i * 3 + j] = fma3(R[i * 3 + 0], T[0 * 3 + j], R[i * 3 + 1], T[1 * 3 + j], R[i * 3 + 2], T[2 * 3 + j]);
  S6[0] = S[0];
  S6[1] = S[4];
  S6[2] = S[8];
  S6[3] = S[1];
  S6[4] = S[5];
  S6[5] = S[2];
}
__device__ __forceinline__ void v6_to_9(const double *s, double *S) {
  S[0] = s[0]; S[4] = s[1]; S[8] = s[2];
  S[1] = S[3] = s[3];
  S[5] = S[7] = s[4];
  S[2] = S[6] = s[5];
}

// Hyperelastic law of the intermediate configuration (law 1, as the CM model's elasticity on
// eF, src/constitutive_model.cc:998-1014): neo-Hookean deviatoric part (devStress_Mooney_Rivlin
// with m01 = 0, src/new_potentials.cc:329-358) and the volumetric potential of volPotFlag
// (:575-616), kappa = E / (3 (1 - 2 nu)):
//   S = 2 m10 d (I - tr(Ce)/3 Ci) + kappa J U'(J) Ci,   Ci = Ce^-1, d = det(Ce)^(-1/3), J = det eF
// Its tangent L = dS/dE (matStiff_Mooney_Rivlin, :376-432, with the volumetric terms) is used
// as a directional derivative on the symmetric part of X:
//   L : X = c1 (Ci:X) Ci + c2 ((Ci:X) I + tr(X) Ci) + c3 Ci X Ci
//   c1 = 4 m10 d tr(Ce)/9 + kappa (J U' + J^2 U''),  c2 = -4 m10 d / 3,  c3 = 4 m10 d tr(Ce)/3 - 2 kappa J U'
struct HeTan {
  double Ci[6], c1, c2, c3;
};

__device__ __forceinline__ void he_law(const CPMaterial &cm, const double *eF, const double *Ce, double *S6,
                                       HeTan &t) {
  const double J = det3(eF);
  double Ci9[9];
  inv3(Ce, Ci9);
  const double d = pow(det3(Ce), -1.0 / 3.0);
  const double trC = Ce[0] + Ce[4] + Ce[8];
  double dU, d2U;
  vol_pot(cm.vol_flag, J, dU, d2U);
  const double a = 2.0 * cm.m10 * d, pv = cm.kappa * J * dU;
  const double ci = pv - a * trC * (1.0 / 3.0);
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    t.Ci[I] = Ci9[vi(I) * 3 + vj(I)];
    S6[I] = (I < 3 ? a : 0.0) + ci * t.Ci[I];
  }
  t.c1 = 2.0 * a * trC * (1.0 / 9.0) + cm.kappa * (J * dU + J * J * d2U);
  t.c2 = -2.0 * a * (1.0 / 3.0);
  t.c3 = 2.0 * a * trC * (1.0 / 3.0) - 2.0 * pv;
}

__device__ __forceinline__ void he_dot(const HeTan &t, const double *X9, double *out6) {
  const double x[6] = {X9[0], X9[4], X9[8], 0.5 * (X9[1] + X9[3]), 0.5 * (X9[5] + X9[7]), 0.5 * (X9[2] + X9[6])};
  const double *C = t.Ci;
  const double cx = C[0] * x[0] + C[1] * x[1] + C[2] * x[2] + 2.0 * (C[3] * x[3] + C[4] * x[4] + C[5] * x[5]);
  const double tx = x[0] + x[1] + x[2];
  double Ci9[9], X[9], T[9];
  v6_to_9(C, Ci9);
  v6_to_9(x, X);
  mm3(Ci9, X, T);
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = vi(I), j = vj(I);
    const double r = T[i * 3] * Ci9[j] + T[i * 3 + 1] * Ci9[3 + j] + T[i * 3 + 2] * Ci9[6 + j];
    out6[I] = t.c1 * cx * C[I] + t.c2 * ((I < 3 ? cx : 0.0) + tx * C[I]) + t.c3 * r;
  }
}

// A pointer whose value the compiler cannot see through: loads from it are not hoisted out of
// loops or merged with earlier loads from the same address, so the data is read from cache
// where it is used instead of being held in registers across a long stretch of code.
__device__ __forceinline__ const double *opaque(const double *p) {
  asm volatile("" : "+v"(p));
  return p;
}

// Elastic PK2 from eF and Ce = eF^T eF: law 0 is S = Lv : (Ce - I)/2 with the grain's cubic
// stiffness; law 1 is he_law, which also fills the tangent coefficients.
template <bool HE>
__device__ __forceinline__ void el_stress(const CPMaterial &cm, const DevGrain &gr, const double *eF,
                                          const double *Ce, double *S6, HeTan &t) {
  if constexpr (HE) {
    he_law(cm, eF, Ce, S6, t);
  } else {
    double E[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) E[i] = 0.5 * (Ce[i] - ((i % 4 == 0) ? 1.0 : 0.0));
    cubic_dot(gr.R, gr.cub, E, S6);
  }
}
// L : X at the state el_stress was evaluated at
template <bool HE>
__device__ __forceinline__ void el_dot(const DevGrain &gr, const HeTan &t, const double *X9, double *out6) {
  if constexpr (HE) he_dot(t, X9, out6);
  else lv_dot(gr.Lv, X9, out6);
}

// |r|^e of the power-law flow rule, e = 1/m - 1.  When e is a whole number (ie >= 0; the rate
// sensitivity m = 0.05 gives 19) it is formed by binary powering, a handful of multiplies where
// pow takes about 130 FP64 operations; the two agree to a few ulps.  ie is the same for every
// lane, so the branch does not diverge.
__device__ __forceinline__ double flow_pow(double x, double e, int ie) {
  if (ie >= 0) {
    double r = 1.0, b = x;
    for (int k = ie; k; k >>= 1) {
      if (k & 1) r *= b;
      b *= b;
    }
    return r;
  }
  return pow(x, e);
}

// The resolved shear stress X : P_a with the Schmid tensor in its factored form P_a = s_a (x) m_a:
// s_a . (X m_a).  The kernels of the local solve keep the 6 numbers of each system's s and m
// rather than the 9 of P_a: 72 doubles of the grain instead of 108 in the registers.
__device__ __forceinline__ double resolved(const DevGrain &gr, int a, const double *X) {
  const double *s = gr.s[a], *m = gr.m[a / 3];
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) t += s[i] * (X[i * 3 + 0] * m[0] + X[i * 3 + 1] * m[1] + X[i * 3 + 2] * m[2]);
  return t;
}

// 9x9 LU with partial pivoting; row swaps are predicated so every index is static.
__device__ __forceinline__ bool lu9(double (&A)[81], int (&piv)[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int p = k;
    double amax = fabs(A[k * 9 + k]);
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const double v = fabs(A[i * 9 + k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    if (!(amax > 0.0)) return false;
    piv[k] = p;
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const bool sw = (i == p);
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double t = A[k * 9 + j];
        A[k * 9 + j] = sw ? A[i * 9 + j] : t;
        A[i * 9 + j] = sw ? t : A[i * 9 + j];
      }
    }
    const double r = 1.0 / A[k * 9 + k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const double l = A[i * 9 + k] * r;
      A[i * 9 + k] = l;
#pragma unroll
      for (int j = k + 1; j < 9; ++j) A[i * 9 + j] -= l * A[k * 9 + j];
    }
  }
  return true;
}
__device__ __forceinline__ void lu9_solve(const double (&A)[81], const int (&piv)[9], double (&b)[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k)
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const bool sw = (i == piv[k]);
      const double t = b[k];
      b[k] = sw ? b[i] : t;
      b[i] = sw ? t : b[i];
    }
#pragma unroll
  for (int k = 0; k < 9; ++k)
#pragma unroll
    for (int i = k + 1; i < 9; ++i) b[i] -= A[i * 9 + k] * b[k];
#pragma unroll
  for (int i = 8; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int j = i + 1; j < 9; ++j) s -= A[i * 9 + j] * b[j];
    b[i] = s / A[i * 9 + i];
  }
}

// Material-point evaluation at x = (M, lambda, g).  Returns false outside the domain.
struct CPPoint {
  double Mi[9], pJ, eF[9], Ce[9], S[6], Sig[9], tau[12], gdot[12], dgdt[12], dgdg[12];
  double gsum, gs, dgs, R[11];
  HeTan he;  // law 1: tangent coefficients at Ce
};

template <bool HE>
__device__ __forceinline__ bool cp_point(const CPMaterial &cm, const DevGrain &gr, const double *A,
                                         double det_pFn, double g_n, double dt, const double *x,
                                         CPPoint &p) {
  const double *M = x;
  const double lam = x[9], g = x[10];
  if (!(g > 0.0)) return false;
  const double dM = det3(M);
  if (!(dM > 0.0)) return false;
  if (!inv3(M, p.Mi)) return false;
  p.pJ = det_pFn / dM;
  mm3(A, M, p.eF);
  mtm3(p.eF, p.eF, p.Ce);
  el_stress<HE>(cm, gr, p.eF, p.Ce, p.S, p.he);
  double S9[9];
  v6_to_9(p.S, S9);
  mm3(p.Ce, S9, p.Sig);
  const double expo = 1.0 / cm.m - 1.0;
  // one reciprocal of g for the twelve systems (the divisions are software sequences)
  const double rg = 1.0 / g, cdg = cm.gd0 * rg / cm.m;
  p.gsum = 0.0;
#pragma unroll
  for (int a = 0; a < 12; ++a) {
    const double t = resolved(gr, a, p.Sig);
    p.tau[a] = t;
    const double r = t * rg;
    const double pw = flow_pow(fabs(r), expo, cm.ipow);
    p.gdot[a] = cm.gd0 * r * pw;
    p.dgdt[a] = cdg * pw;
    p.dgdg[a] = -p.dgdt[a] * r;
    p.gsum += fabs(p.gdot[a]);
  }
  if (cm.w == 0.0) {
    p.gs = cm.gs0;
    p.dgs = 0.0;
  } else if (p.gsum / cm.gds > 1.0e-15) {
    p.gs = cm.gs0 * pow(p.gsum / cm.gds, cm.w);
    p.dgs = cm.w * p.gs / p.gsum;
  } else {
    p.gs = 0.0;
    p.dgs = 0.0;
  }
  double sl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum_a gdot_a s_a (x) m_a, in system order
#pragma unroll
  for (int a = 0; a < 12; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double c = p.gdot[a] * gr.s[a][i];
#pragma unroll
      for (int j = 0; j < 3; ++j) sl[i * 3 + j] += c * gr.m[a / 3][j];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      p.R[i * 3 + j] = M[i * 3 + j] - ((i == j) ? 1.0 : 0.0) + dt * sl[i * 3 + j] - dt * lam * p.pJ * p.Mi[j * 3 + i];
  p.R[9] = p.pJ - 1.0;
  p.R[10] = g * (p.gs - cm.g0 + dt * cm.h0 * p.gsum) - (p.gs - cm.g0) * g_n - dt * cm.h0 * p.gs * p.gsum;
  return true;
}

// Q_a = sym(P_a S + 1/2 L:(C_e P_a)) in Voigt order, so that d tau_a / d M = 2 A^T eF Q_a
template <bool HE>
__device__ __forceinline__ void qsym(const DevGrain &gr, int a, const CPPoint &p, double *Q6) {
  // with P = s (x) m: P S = s (x) (S m) (S symmetric) and C_e P = (C_e s) (x) m
  const double *s = gr.s[a], *m = gr.m[a / 3];
  double S9[9], PS[9], CP[9], v[3], w[3];
  v6_to_9(p.S, S9);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v[i] = S9[i * 3 + 0] * m[0] + S9[i * 3 + 1] * m[1] + S9[i * 3 + 2] * m[2];
    w[i] = p.Ce[i * 3 + 0] * s[0] + p.Ce[i * 3 + 1] * s[1] + p.Ce[i * 3 + 2] * s[2];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      PS[i * 3 + j] = s[i] * v[j];
      CP[i * 3 + j] = w[i] * m[j];
    }
  double LCP[6];
  if constexpr (HE) he_dot(p.he, CP, LCP);
  else cubic_dot(gr.R, gr.cub, CP, LCP);
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = vi(I), j = vj(I);
    Q6[I] = 0.5 * (PS[i * 3 + j] + PS[j * 3 + i]) + 0.5 * LCP[I];
  }
}

// d tau_a / d M = 2 A^T eF sym(P_a S + 1/2 L:(C_e P_a))
template <bool HE>
__device__ __forceinline__ void dtau_dM(const DevGrain &gr, int a, const double *AtE, const CPPoint &p, double *D) {
  double Q6[6], sQ[9];
  qsym<HE>(gr, a, p, Q6);
  v6_to_9(Q6, sQ);
  double T[9];
  mm3(AtE, sQ, T);
#pragma unroll
  for (int i = 0; i < 9; ++i) D[i] = 2.0 * T[i];
}

// d tau_a / d M from a Q_a formed earlier: 2 A^T eF sym(Q_a)
__device__ __forceinline__ void dtau_dM_q(const double *AtE, const double *Q6, double *D) {
  double sQ[9], T[9];
  v6_to_9(Q6, sQ);
  mm3(AtE, sQ, T);
#pragma unroll
  for (int i = 0; i < 9; ++i) D[i] = 2.0 * T[i];
}

// Bordered Jacobian [Ah b c; d^T 0 0; e^T 0 f] at p.
struct Border {
  double b[9], c[9], d[9], e[9], f, dRg;
};

// Bordered Jacobian of the local system at p; the per-system slip quantities are recomputed
// from the Mandel stress with cp_point's arithmetic.  QPRE: the Q_a of qsym were formed
// earlier and sit in LDS at qpre[(a * 6 + I) * 64].
template <bool HE, bool QPRE = false>
__device__ __forceinline__ void cp_jacobian_re(const CPMaterial &cm, const DevGrain &gr, double dt, const double *x,
                                            const double *A, const CPPoint &p, double g_n,
                                            double (&Ah)[81], Border &B, const double *qpre = nullptr) {
  const double lam = x[9], g = x[10];
  double AtE[9];
  mtm3(A, p.eF, AtE);
#pragma unroll
  for (int i = 0; i < 81; ++i) Ah[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    B.c[i] = 0.0;
    B.e[i] = 0.0;
  }
  double sg = 0.0;
  const double expo = 1.0 / cm.m - 1.0;
  const double rg = 1.0 / g, cdg = cm.gd0 * rg / cm.m;
  for (int a = 0; a < 12; ++a) {
    // slip-system quantities recomputed from the Mandel stress (as cp_point evaluates them)
    const double t = resolved(gr, a, p.Sig);
    const double r = t * rg;
    const double pw = flow_pow(fabs(r), expo, cm.ipow);
    const double gd = cm.gd0 * r * pw;
    const double dgdt = cdg * pw;
    const double dgdg = -dgdt * r;
    double D[9];
    if constexpr (QPRE) {
      double Q6[6];
#pragma unroll
      for (int I = 0; I < 6; ++I) Q6[I] = qpre[(a * 6 + I) * 64];
      dtau_dM_q(AtE, Q6, D);
    } else {
      dtau_dM<HE>(gr, a, AtE, p, D);
    }
    const double co = dt * dgdt;
    const double sgn = (gd < 0) ? -1.0 : 1.0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const double Pr = gr.s[a][r / 3] * gr.m[a / 3][r % 3];
      const double pr = Pr * co;
#pragma unroll
      for (int c = 0; c < 9; ++c) Ah[r * 9 + c] += pr * D[c];
      B.c[r] += dt * Pr * dgdg;
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) B.e[c] += sgn * dgdt * D[c];
    sg += sgn * dgdg;
  }
  const double *Mi = p.Mi;
  const double lp = dt * lam * p.pJ;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          Ah[(i * 3 + j) * 9 + k * 3 + l] +=
              ((i == k && j == l) ? 1.0 : 0.0) + lp * (Mi[j * 3 + i] * Mi[l * 3 + k] + Mi[l * 3 + i] * Mi[j * 3 + k]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      B.b[i * 3 + j] = -dt * p.pJ * Mi[j * 3 + i];
      B.d[i * 3 + j] = -p.pJ * Mi[j * 3 + i];
    }
  B.dRg = g * (p.dgs + dt * cm.h0) - p.dgs * g_n - dt * cm.h0 * (p.dgs * p.gsum + p.gs);
#pragma unroll
  for (int c = 0; c < 9; ++c) B.e[c] *= B.dRg;
  B.f = (p.gs - cm.g0 + dt * cm.h0 * p.gsum) + B.dRg * sg;
  // Schur: Ah <- A - c e^T / f
  const double rf = 1.0 / B.f;
#pragma unroll
  for (int r = 0; r < 9; ++r)
#pragma unroll
    for (int c = 0; c < 9; ++c) Ah[r * 9 + c] -= B.c[r] * B.e[c] * rf;
}

// Solve [A b c; d^T 0 0; e^T 0 f] [u; y; z] = [r; s; t] given LU(Ah) and v = Ah^-1 b.
__device__ __forceinline__ void bordered_solve(const double (&LU)[81], const int (&piv)[9], const Border &B,
                                               const double (&v)[9], double dTv, const double *r, double s,
                                               double t, double *out) {
  double u[9];
  const double tf = t / B.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) u[i] = r[i] - B.c[i] * tf;
  lu9_solve(LU, piv, u);
  double dTu = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) dTu += B.d[i] * u[i];
  const double y = (dTu - s) / dTv;
  double eTx = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    out[i] = u[i] - y * v[i];
    eTx += B.e[i] * out[i];
  }
  out[9] = y;
  out[10] = (t - eTx) / B.f;
}

__device__ __forceinline__ double cp_merit(const CPPoint &p, double gscale) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < 10; ++i) r += p.R[i] * p.R[i];
  const double rg = p.R[10] * gscale;
  return 0.5 * (r + rg * rg);
}

// The merit of x without keeping the point: the same arithmetic as cp_point's residual (slip
// contributions accumulated in slip-system order), so acceptance decisions match it.
template <bool HE>
__device__ __forceinline__ bool cp_merit_at(const CPMaterial &cm, const DevGrain &gr, const double *A,
                                            double det_pFn, double g_n, double dt, const double *x, double gscale,
                                            double &merit) {
  const double *M = x;
  const double lam = x[9], g = x[10];
  if (!(g > 0.0)) return false;
  const double dM = det3(M);
  if (!(dM > 0.0)) return false;
  double Mi[9];
  if (!inv3(M, Mi)) return false;
  const double pJ = det_pFn / dM;
  double eF[9], Ce[9], S[6], S9[9], Sig[9];
  mm3(A, M, eF);
  mtm3(eF, eF, Ce);
  HeTan ht;
  el_stress<HE>(cm, gr, eF, Ce, S, ht);
  v6_to_9(S, S9);
  mm3(Ce, S9, Sig);
  const double expo = 1.0 / cm.m - 1.0;
  const double rg = 1.0 / g;
  double gsum = 0.0, sl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < 12; ++a) {
    const double t = resolved(gr, a, Sig);
    const double r = t * rg;
    const double gd = cm.gd0 * r * flow_pow(fabs(r), expo, cm.ipow);
    gsum += fabs(gd);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double c = gd * gr.s[a][i];
#pragma unroll
      for (int j = 0; j < 3; ++j) sl[i * 3 + j] += c * gr.m[a / 3][j];
    }
  }
  double gs;
  if (cm.w == 0.0) gs = cm.gs0;
  else if (gsum / cm.gds > 1.0e-15) gs = cm.gs0 * pow(gsum / cm.gds, cm.w);
  else gs = 0.0;
  double r2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double R = M[i * 3 + j] - ((i == j) ? 1.0 : 0.0) + dt * sl[i * 3 + j] - dt * lam * pJ * Mi[j * 3 + i];
      r2 += R * R;
    }
  const double R9 = pJ - 1.0;
  r2 += R9 * R9;
  const double R10 = (g * (gs - cm.g0 + dt * cm.h0 * gsum) - (gs - cm.g0) * g_n - dt * cm.h0 * gs * gsum) * gscale;
  merit = 0.5 * (r2 + R10 * R10);
  return true;
}

// max_a |tau_a| / g at M (elastic part eF = A M)
template <bool HE>
__device__ __forceinline__ double tau_ratio_max(const CPMaterial &cm, const DevGrain &gr, const double *A,
                                                const double *M, double rg) {
  double eF[9], Ce[9], S[6], S9[9], Sig[9];
  mm3(A, M, eF);
  mtm3(eF, eF, Ce);
  HeTan ht;
  el_stress<HE>(cm, gr, eF, Ce, S, ht);
  v6_to_9(S, S9);
  mm3(Ce, S9, Sig);
  double mx = 0.0;
#pragma unroll
  for (int a = 0; a < 12; ++a) mx = fmax(mx, fabs(resolved(gr, a, Sig)));
  return mx * rg;
}

// Plain start x = (M, lambda_n, g_n): the fully plastic predictor pulled back to yield.
template <bool HE>
__device__ __forceinline__ bool cp_start(const CPMaterial &cm, const DevGrain &gr, const double *A,
                                         const double *F, const double *Fn,
                                         const double *pFn, const double *pFi, double g_n, double lam_n,
                                         double *x) {
  double Fi[9], T[9], Mp[9];
  if (!inv3(F, Fi)) return false;
  mm3(pFn, Fi, T);
  mm3(T, Fn, Mp);
  mm3(Mp, pFi, T);
  const double dp = det3(T);
  if (!(dp > 0.0)) return false;
  const double c = 1.0 / cbrt(dp);
#pragma unroll
  for (int i = 0; i < 9; ++i) Mp[i] = T[i] * c;
  const double rg = 1.0 / g_n;
  const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double al = 0.0;
  if (tau_ratio_max<HE>(cm, gr, A, I9, rg) > 1.0) {
    al = 1.0;
    if (!(tau_ratio_max<HE>(cm, gr, A, Mp, rg) > 1.0)) {
      double lo = 0.0, hi = 1.0;
      for (int k = 0; k < 8; ++k) {
        const double mid = 0.5 * (lo + hi);
        double Mm[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Mm[i] = I9[i] + mid * (Mp[i] - I9[i]);
        if (tau_ratio_max<HE>(cm, gr, A, Mm, rg) > 1.0) lo = mid;
        else hi = mid;
      }
      al = hi;
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] = I9[i] + al * (Mp[i] - I9[i]);
  if (al > 0.0) {
    const double sc = 1.0 / cbrt(det3(x));
#pragma unroll
    for (int i = 0; i < 9; ++i) x[i] *= sc;
  }
  x[9] = lam_n;
  x[10] = g_n;
  return true;
}

// ------------------------------------------------------------------- kinematics kernel
template <int NNE>
__global__ __launch_bounds__(256) void cp_kinematics(CPArgs a) {
  constexpr int NQP = nqp_of(NNE);
  const int64_t tq = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tq >= a.ne * NQP) return;
  const int64_t e = tq / NQP;
  const int ip = int(tq % NQP);
  double x[NNE * 3], u[NNE * 3];
#pragma unroll
  for (int n = 0; n < NNE; ++n) {
    const int32_t node = a.conn[e * NNE + n];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      x[n * 3 + b] = a.X[int64_t(node) * 3 + b];
      const int32_t d = a.dofid[int64_t(node) * 3 + b];
      double uu = 0.0;  // def_elem_total, src/utils.cc:1391-1412
      if (d > 0) uu = a.r[d - 1] + (a.d_r ? a.d_r[d - 1] : 0.0);
      else if (d < 0) uu = (a.defl ? a.defl[-d - 1] : 0.0) + (a.defl_d ? a.defl_d[-d - 1] : 0.0);
      u[n * 3 + b] = uu;
    }
  }
  double k, ee, z, wt;
  int_point(NNE, ip, k, ee, z, wt);
  double dN[NNE * 3];
  shape_dN<NNE>(k, ee, z, dN);
  double dX[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < NNE; ++n)
#pragma unroll
    for (int ib = 0; ib < 3; ++ib)
#pragma unroll
      for (int ic = 0; ic < 3; ++ic) dX[ib * 3 + ic] += dN[n * 3 + ic] * x[n * 3 + ib];
  double Xi[9];
  const bool ok = det3(dX) > 0.0 && inv3(dX, Xi);
  double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
  for (int n = 0; n < NNE; ++n) {
    double g[3];
#pragma unroll
    for (int ib = 0; ib < 3; ++ib)
      g[ib] = dN[n * 3 + 0] * Xi[0 * 3 + ib] + dN[n * 3 + 1] * Xi[1 * 3 + ib] + dN[n * 3 + 2] * Xi[2 * 3 + ib];
#pragma unroll
    for (int ib = 0; ib < 3; ++ib)
#pragma unroll
      for (int ic = 0; ic < 3; ++ic) F[ib * 3 + ic] += u[n * 3 + ib] * g[ic];
  }
  const int64_t q = a.qp_base + tq;
  const double J = det3(F);
  bool okF = ok && J > 0.0;  // getJacobian, src/utils.cc:513-527
  if (a.tf_V) {
    // CM3F: the model integrates at (theta/J)^(1/3) F and keeps it as its F(n+1)
    // (run_integration_algorithm with tf_factor, src/constitutive_model.cc:810-818, :1109-1115)
    const double th = a.tf_V[e] + a.tf_dV[e];
    okF = okF && th > 0.0;
    const double fac = okF ? cbrt(th / J) : 1.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] *= fac;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) a.F_np1[i * a.n_qp + q] = F[i];
  if (!okF) atomicAdd(a.err, 1);
}

// ------------------------------------------------------------------- start kernel
// The local Newton's starting points, one thread per QP, ahead of the Newton kernel: the plain
// start (the plastic predictor pulled back to yield, about ten stress evaluations) and, with a
// caller's start attached, which of the two has the lower residual merit.  In the Newton
// kernel a lane takes a new QP whenever its last one converges, so work done at a refill runs
// with the other lanes of the wave masked off; here it runs on every lane.  The QP's record in
// qpbuf holds the plain start x(0..10) and the choice kStartFlag (bit 0: the plain start
// failed, bit 1: take the caller's start) until the tangent kernel overwrites it.
constexpr int kStartFlag = 11;
template <bool HE>
__global__ __launch_bounds__(64) void cp_start_kernel(CPArgs a) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.nq) return;
  const int64_t NQ = a.n_qp;
  const int64_t qg = a.qp_base + q;
  const CPMaterial cm = a.cpmats[0];
  const DevGrain &gr = a.grains[a.qgrain[q]];
  double F[9], pFn[9], pFi[9], Fn[9], A[9], x[11];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = a.F_np1[i * NQ + qg];
    pFn[i] = a.pF_n[i * NQ + qg];
    Fn[i] = a.F_n[i * NQ + qg];
  }
  inv3(pFn, pFi);
  mm3(F, pFi, A);
  const double g_n = a.g_n[qg];
  const bool bad = !cp_start<HE>(cm, gr, A, F, Fn, pFn, pFi, g_n, a.lam_n[qg], x);
  int flag = bad ? 1 : 0;
  if (a.warm) {
    // a caller's start is taken when its residual merit is below the plain start's
    const double det_pFn = det3(pFn);
    const double gscale = 1.0 / (fabs(cm.gs0 - cm.g0) * g_n + 1e-300);
    double xw[11];
#pragma unroll
    for (int i = 0; i < 11; ++i) xw[i] = a.warm[i * NQ + qg];
    double mw = 0.0, mp = 0.0;
    const bool okw = cp_merit_at<HE>(cm, gr, A, det_pFn, g_n, a.dt, xw, gscale, mw);
    const bool okp = !bad && cp_merit_at<HE>(cm, gr, A, det_pFn, g_n, a.dt, x, gscale, mp);
    if (okw && (!okp || mw < mp)) flag |= 2;
  }
  double *row = a.qpbuf + qrec(qg);
#pragma unroll
  for (int i = 0; i < 11; ++i) row[i * kRecStride] = x[i];
  row[kStartFlag * kRecStride] = double(flag);
}

// ------------------------------------------------------------------- Newton kernel
// One wave per workgroup.  With a.qnext set, every wave takes its QPs from one global queue:
// at each refill the wave's first lane that needs a QP claims as many consecutive QPs as lanes
// need one, so lanes stay busy until the queue runs dry.  Otherwise wave w owns QPs
// [w*chunk, (w+1)*chunk) of the block.  Which lane takes a QP does not change its result.  The
// starting points come from cp_start_kernel (qpbuf).
template <bool HE>
__global__ __launch_bounds__(64) void cp_newton(CPArgs a, int64_t chunk) {
  const int lane = threadIdx.x;
  const bool gq = a.qnext != nullptr;
  const int64_t q_lo = gq ? 0 : int64_t(blockIdx.x) * chunk;
  const int64_t q_hi = (gq || q_lo + chunk >= a.nq) ? a.nq : q_lo + chunk;
  int64_t cursor = q_lo;  // wave-uniform
  const int64_t NQ = a.n_qp;
  bool need = true, done = false;
  int64_t q = -1;
  // The lane's QP data that lives across turns sits in LDS (lane-fastest; each lane touches
  // only its own slots), read where used, rather than in registers held through the material
  // point and the Jacobian's LU, where it spilled: A (9), x (11), dx (11), det pF(n), g(n),
  // gscale.  17 KB per wave.
  __shared__ double sPk[34 * 64];
  double *pk = sPk + lane;  // field i at pk[i * 64]
  constexpr int kA = 0, kX = 9, kDX = 20, kDet = 31, kGn = 32, kGs = 33;
  const double dt = a.dt;
  int it = 0, it_att = 0, gid = 0;
  bool bad_start = false;
  int kind = 0;  // the QP's start: 0 plain, 1 the caller's (learned), 2 the caller's, then the plain fallback
  // the lane's phase: evaluate the start, a line-search trial x + alpha dx, or the full step
  // after 40 halvings without a decrease
  constexpr int kPhStart = 0, kPhTrial = 1, kPhFull = 2;
  int phase = kPhStart, ls = 0;
  double f0 = 0.0, alpha = 1.0;
  const CPMaterial cm = a.cpmats[0];  // one material for every CP point: wave-uniform, loaded once
  for (;;) {
    // refill the lanes that need a QP, in lane order, from the wave's range
    const uint64_t mask = __ballot(need && !done);
    const int nneed = __popcll(mask);
    if (gq && nneed > 0) {  // claim nneed QPs from the global queue (wave-uniform branch)
      const int leader = __ffsll((unsigned long long)mask) - 1;
      long long base = 0;
      if (lane == leader) base = (long long)atomicAdd(a.qnext, (unsigned long long)nneed);
      cursor = __shfl(base, leader);
    }
    if (need && !done) {
      const int rank = __popcll(mask & ((uint64_t(1) << lane) - 1));
      const int64_t cand = cursor + rank;
      if (cand >= q_hi) {
        done = true;
      } else {
        q = cand;
        need = false;
        const int64_t qg = a.qp_base + q;
        double F[9], pFn[9], pFi[9], A[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          F[i] = a.F_np1[i * NQ + qg];
          pFn[i] = a.pF_n[i * NQ + qg];
        }
        inv3(pFn, pFi);
        mm3(F, pFi, A);
#pragma unroll
        for (int i = 0; i < 9; ++i) pk[(kA + i) * 64] = A[i];
        pk[kDet * 64] = det3(pFn);
        const double g_n = a.g_n[qg];
        pk[kGn * 64] = g_n;
        gid = a.qgrain[q];
        pk[kGs * 64] = 1.0 / (fabs(cm.gs0 - cm.g0) * g_n + 1e-300);
        it = 0;
        it_att = 0;
        phase = kPhStart;
        // start: the caller's warm start where cp_start_kernel chose it, else the plain start
        const double *row = a.qpbuf + qrec(qg);
        const int flag = int(row[kStartFlag * kRecStride]);
        bad_start = (flag & 1) != 0;
        kind = 0;
        if (flag & 2) {
#pragma unroll
          for (int i = 0; i < 11; ++i) pk[(kX + i) * 64] = a.warm[i * NQ + qg];
          kind = 1;
          bad_start = false;
        } else {
#pragma unroll
          for (int i = 0; i < 11; ++i) pk[(kX + i) * 64] = row[i * kRecStride];
        }
      }
    }
    if (!gq) cursor += nneed;
    if (__all(done)) break;
    if (done || need) continue;

    // ---- one turn of this lane's QP: every live lane evaluates the material point once, at its
    // start or at its line-search trial x + alpha dx; a lane whose start or accepted trial this
    // is then takes a Newton step from it.  A trial's evaluation is thus also the point of the
    // next iteration (one evaluation per iteration, and none of it at a refill's masked width).
    const DevGrain &gr = a.grains[gid];
    const int64_t qg = a.qp_base + q;
    bool fail = false, conv = false, step = false;
    double xc[11], A[9];
#pragma unroll
    for (int i = 0; i < 11; ++i)
      xc[i] = (phase == kPhStart) ? pk[(kX + i) * 64]
                                  : pk[(kX + i) * 64] + (phase == kPhTrial ? alpha : 1.0) * pk[(kDX + i) * 64];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = pk[(kA + i) * 64];
    CPPoint p;
    const bool valid =
        !(phase == kPhStart && bad_start) && cp_point<HE>(cm, gr, A, pk[kDet * 64], pk[kGn * 64], dt, xc, p);
    if (phase == kPhStart) {
      ++it;
      ++it_att;
      if (valid) step = true;
      else fail = true;
    } else {
      bool accept = false;
      if (phase == kPhTrial) {
        accept = valid && cp_merit(p, pk[kGs * 64]) <= (1.0 - 1e-4 * alpha) * f0;
        if (!accept) {
          alpha *= 0.5;
          if (++ls == 40) phase = kPhFull;  // no decrease found: the full step, if it is valid
        }
      } else {
        accept = valid;
        if (!valid) fail = true;
      }
      if (accept) {
#pragma unroll
        for (int i = 0; i < 11; ++i) pk[(kX + i) * 64] = xc[i];
        if (it_att >= cm.maxit) {  // the end of the iteration that took this step
          fail = true;
        } else {
          ++it;
          ++it_att;
          step = true;
        }
      }
    }
    if (step) {
      double Ah[81];
      Border B;
      // the point of this step (the start, or the accepted trial); the Jacobian reads lambda and g
      double xj[11] = {};
      xj[9] = pk[(kX + 9) * 64];
      xj[10] = pk[(kX + 10) * 64];
      cp_jacobian_re<HE>(cm, gr, dt, xj, A, p, pk[kGn * 64], Ah, B);
      int piv[9];
      if (!lu9(Ah, piv)) {
        fail = true;
      } else {
        double v[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) v[i] = dt * B.d[i];  // b = dt d
        lu9_solve(Ah, piv, v);
        double dTv = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) dTv += B.d[i] * v[i];
        double r9[9], dx[11];
#pragma unroll
        for (int i = 0; i < 9; ++i) r9[i] = -p.R[i];
        bordered_solve(Ah, piv, B, v, dTv, r9, -p.R[9], -p.R[10], dx);
        double du = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) du = fmax(du, fabs(dx[i]));
        du = fmax(du, fabs(dx[9]) * dt);
        du = fmax(du, fabs(dx[10]) / pk[kGn * 64]);
        if (du < cm.tol) {
#pragma unroll
          for (int i = 0; i < 11; ++i) pk[(kX + i) * 64] = pk[(kX + i) * 64] + dx[i];
          conv = true;
        } else {  // backtracking on half the squared residual norm, from the full step
#pragma unroll
          for (int i = 0; i < 11; ++i) pk[(kDX + i) * 64] = dx[i];
          f0 = cp_merit(p, pk[kGs * 64]);
          alpha = 1.0;
          ls = 0;
          phase = kPhTrial;
        }
      }
    }
    if (fail && kind == 1) {
      // the warm start did not converge: fall back to the plain start with a fresh budget
      // (iterations keep counting)
      kind = 2;
      it_att = 0;
      const double *row = a.qpbuf + qrec(qg);
#pragma unroll
      for (int i = 0; i < 11; ++i) pk[(kX + i) * 64] = row[i * kRecStride];
      bad_start = (int(row[kStartFlag * kRecStride]) & 1) != 0;
      fail = false;
      phase = kPhStart;
    }
    if (conv || fail) {
      a.iters[qg] = it | (kind << 24);  // iterations, and the start in the high byte
      if (fail) {
        atomicAdd(a.err, 1);
      } else {
        // pF(n+1) = M^-1 pF(n)
        double x[11], Mi[9], pFn[9], pF1[9];
#pragma unroll
        for (int i = 0; i < 11; ++i) x[i] = pk[(kX + i) * 64];
        inv3(x, Mi);
#pragma unroll
        for (int i = 0; i < 9; ++i) pFn[i] = a.pF_n[i * NQ + qg];
        mm3(Mi, pFn, pF1);
#pragma unroll
        for (int i = 0; i < 9; ++i) a.pF_np1[i * NQ + qg] = pF1[i];
        a.Sc_np1[qg] = x[10];
        a.Sc_np1[NQ + qg] = x[9];
      }
      need = true;
    }
  }
}

// ------------------------------------------------------------------- tangent kernel
// At the stored converged state (F, pF(n), pF(n+1), g, lambda): S, tau, gdot, Me, pJ and,
// with WITH_D, dMe/dF = pF(n)^-1 dM/dF from the implicit derivative of the local system.
template <bool HE, bool WITH_D>
__global__ __launch_bounds__(64) void cp_tangent(CPArgs a, int64_t nqb) {
  // the Q_a of the twelve systems, formed once for the right-hand sides and read back by the
  // Jacobian (lane-fastest: conflict-free)
  __shared__ double sQ[WITH_D ? 72 * 64 : 1];
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nqb) return;
  const int64_t NQ = a.n_qp;
  const int64_t qg = a.qp_base + t;
  const CPMaterial cm = a.cpmats[0];
  const DevGrain &gr = a.grains[a.qgrain[t]];
  double F[9], pFn[9], pF1[9], pFi[9], A[9], Me[9], x[11];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = a.F_np1[i * NQ + qg];
    pFn[i] = a.pF_n[i * NQ + qg];
    pF1[i] = a.pF_np1[i * NQ + qg];
  }
  bool ok = inv3(pFn, pFi) && inv3(pF1, Me);
  mm3(F, pFi, A);
  mm3(pFn, Me, x);  // M = pF(n) pF(n+1)^-1
  x[9] = a.Sc_np1[NQ + qg];
  x[10] = a.Sc_np1[qg];
  const double g_n = a.g_n[qg];
  CPPoint p;
  ok = ok && cp_point<HE>(cm, gr, A, det3(pFn), g_n, a.dt, x, p);
  double *out = a.qpbuf + qrec(qg);  // field i at out[i * kRecStride]
  if (!ok) {
    atomicAdd(a.err, 1);
#pragma unroll
    for (int i = 0; i < kCPBuf; ++i) out[i * kRecStride] = 0.0;
    return;
  }
  const double expo = 1.0 / cm.m - 1.0;
  const double rg = 1.0 / x[10], cdg = cm.gd0 * rg / cm.m;
  for (int al = 0; al < 12; ++al) {  // tau and gamma_dot as cp_point evaluates them
    const double tt = resolved(gr, al, p.Sig);
    const double r = tt * rg;
    a.tau[al * NQ + qg] = tt;
    a.gdot[al * NQ + qg] = cm.gd0 * r * flow_pow(fabs(r), expo, cm.ipow);
  }
  // elastic PK2 from the stored state (eF = F pF(n+1)^-1)
  double eF[9], Ce[9], S[6];
  mm3(F, Me, eF);
  mtm3(eF, eF, Ce);
  {
    HeTan ht;
    el_stress<HE>(cm, gr, eF, Ce, S, ht);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i * kRecStride] = Me[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) out[(9 + i) * kRecStride] = S[i];
  out[15 * kRecStride] = det3(pF1);
  if constexpr (WITH_D) {
    // dtau_a/dF = A^-T D_a (pF(n)^-1 M)^T with D_a = 2 A^T eF Q_a, so dtau_a/dF = 2 eF Q_a BM^T
    // (BM = pF(n)^-1 M).  Column c = (i, j) of F has the right-hand side
    // rhs[r] = sum_a co_a P_a[r] q_a and the hardening row -sum_a sg_a q_a, with
    // q_a = 2 (eF Q_a BM^T)_ij = 2 (Q_a u).w (u = row i of eF, w = row j of BM).  They are formed
    // after the factorization from the Q_a in LDS, three columns (one row of eF) at a time, so
    // no accumulator over the slip systems is live across the Jacobian and its LU.
    double BM[9];
    mm3(pFi, x, BM);
    // d gamma_dot_a / d tau_a, with the sign of gamma_dot_a (co_a = -dt |dg_a|, sg_a = dg_a)
    double dg[12];
#pragma unroll
    for (int al = 0; al < 12; ++al) {
      double Q[6];
      qsym<HE>(gr, al, p, Q);
#pragma unroll
      for (int I = 0; I < 6; ++I) sQ[(al * 6 + I) * 64 + threadIdx.x] = Q[I];
      const double tt = resolved(gr, al, p.Sig);
      const double rr = tt * rg;
      const double pw = flow_pow(fabs(rr), expo, cm.ipow);
      const double dgdt = cdg * pw;
      dg[al] = (cm.gd0 * rr * pw < 0) ? -dgdt : dgdt;
    }
    // the local Jacobian at the converged point, factored once for the ten solves
    double Ah[81];
    Border B;
    cp_jacobian_re<HE, true>(cm, gr, a.dt, x, A, p, g_n, Ah, B, sQ + threadIdx.x);
    int piv[9];
    if (!lu9(Ah, piv)) {
      atomicAdd(a.err, 1);
      return;
    }
    double v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = B.b[i];
    lu9_solve(Ah, piv, v);
    double dTv = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) dTv += B.d[i] * v[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double *u = p.eF + i * 3;
      double rhs[3][9], rgc[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        rgc[j] = 0.0;
#pragma unroll
        for (int r = 0; r < 9; ++r) rhs[j][r] = 0.0;
      }
      // the slip directions and plane normals read where used: with the grain's address visible
      // the compiler keeps the Jacobian's loads of them live through the LU, and spills
      const DevGrain &go = *reinterpret_cast<const DevGrain *>(opaque(reinterpret_cast<const double *>(&gr)));
      // P_a = s_a (x) m_a: the three systems of plane k summed (vs = sum co_a q_a s_a) before
      // the outer product with m_k
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double vs[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) vs[j][0] = vs[j][1] = vs[j][2] = 0.0;
#pragma unroll
        for (int al = 3 * k; al < 3 * k + 3; ++al) {
          double Q[6];
#pragma unroll
          for (int I = 0; I < 6; ++I) Q[I] = sQ[(al * 6 + I) * 64 + threadIdx.x];
          const double Qu[3] = {Q[0] * u[0] + Q[3] * u[1] + Q[5] * u[2], Q[3] * u[0] + Q[1] * u[1] + Q[4] * u[2],
                                Q[5] * u[0] + Q[4] * u[1] + Q[2] * u[2]};
          const double cad = -a.dt * fabs(dg[al]);
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double *w = BM + j * 3;
            const double q = 2.0 * (Qu[0] * w[0] + Qu[1] * w[1] + Qu[2] * w[2]);
            const double ta = cad * q;
            rgc[j] -= dg[al] * q;
#pragma unroll
            for (int l = 0; l < 3; ++l) vs[j][l] += ta * go.s[al][l];
          }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int n = 0; n < 3; ++n) rhs[j][l * 3 + n] += vs[j][l] * go.m[k][n];
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int c = i * 3 + j;
        double sol[11];
        bordered_solve(Ah, piv, B, v, dTv, rhs[j], 0.0, B.dRg * rgc[j], sol);
        // dMe/dF[:, c] = pF(n)^-1 dM/dF[:, c]
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
            out[(16 + c * 9 + ii * 3 + jj) * kRecStride] =
                pFi[ii * 3 + 0] * sol[0 * 3 + jj] + pFi[ii * 3 + 1] * sol[1 * 3 + jj] + pFi[ii * 3 + 2] * sol[2 * 3 + jj];
      }
    }
  }
}

// ------------------------------------------------------------------- element kernel
// Elements per round of the LDS transpose of the staged rows (gather mode): the largest
// divisor of epb whose rows (nde x (nde+1) doubles each) fit in the dead dM/dF array.
constexpr int cp_stage_round(int epb, int nde, int room) {
  int re = epb;
  while (re > 1 && (epb % re != 0 || re * nde * (nde + 1) > room)) --re;
  return re;
}

template <bool HE, int NNE, int MODE, bool STG>
__global__ __launch_bounds__(256) void cp_elem(CPElemArgs a, int64_t e_begin, int64_t e_end) {
  constexpr int NQP = nqp_of(NNE);
  constexpr int NDE = NNE * 3;
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = (NNE == 8) ? 4 : 256 / TPE;
  constexpr bool RES = (MODE & MODE_RESIDUAL) != 0;
  constexpr bool TAN = (MODE & MODE_TANGENT) != 0;
  constexpr bool DFL = (MODE & MODE_DEFL_LOAD) != 0;  // K times the prescribed increments
  static_assert(!DFL || (TAN && !RES), "MODE_DEFL_LOAD goes with MODE_TANGENT alone");
  constexpr bool WK = TAN && !DFL;  // K rows are written
  constexpr bool WR = RES || DFL;   // the residual entry is written
  // dMe/dF to LDS by all of an element's lanes (tets), or by each QP's lane (hex8, where the
  // 648 values of eight QPs measured slower spread over 32 lanes)
  constexpr bool SPREAD_D = NNE != 8;
  __shared__ double s_x[EPB][NDE], s_u[EPB][NDE];
  __shared__ double s_dl[EPB][DFL ? NDE : 1];
  __shared__ int32_t s_dof[EPB][NDE], s_node[EPB][NNE];
  __shared__ double s_dN[EPB][NQP][NDE];
  __shared__ double s_Fr[EPB][NQP][9], s_Me[EPB][NQP][9], s_MeS[EPB][NQP][9], s_CrMe[EPB][NQP][9];
  __shared__ double s_MiT[EPB][NQP][9], s_Sh[EPB][NQP][6], s_S[EPB][NQP][6], s_w[EPB][NQP];
  __shared__ double s_D[EPB][NQP][TAN ? 81 : 1];
  __shared__ double s_he[EPB][NQP][(HE && TAN) ? 9 : 1];  // law 1: Ci, c1, c2, c3 at the QP

  const int slot = threadIdx.x / TPE, lane = threadIdx.x % TPE;
  const int64_t ie = e_begin + int64_t(blockIdx.x) * EPB + slot;
  const bool active = (threadIdx.x < EPB * TPE) && ie < e_end;
  const int64_t e = (a.elist && active) ? int64_t(a.elist[ie]) : ie;
  if (active && lane < NNE) {
    const int32_t node = a.conn[e * NNE + lane];
    s_node[slot][lane] = node;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      s_x[slot][lane * 3 + b] = a.X[int64_t(node) * 3 + b];
      const int32_t d = a.dofid[int64_t(node) * 3 + b];
      s_dof[slot][lane * 3 + b] = d;
      double u = 0.0, dl = 0.0;
      if (d > 0) {
        u = a.r[d - 1] + (a.d_r ? a.d_r[d - 1] : 0.0);
      } else if (d < 0) {
        dl = a.defl_d ? a.defl_d[-d - 1] : 0.0;
        u = (a.defl ? a.defl[-d - 1] : 0.0) + (DFL ? 0.0 : dl);  // the prescribed-load pass: K at defl
      }
      s_u[slot][lane * 3 + b] = u;
      if constexpr (DFL) s_dl[slot][lane * 3 + b] = dl;
    }
  }
  __syncthreads();
  if (active && lane < NQP) {
    const int ip = lane;
    double k, ee, z, wt;
    int_point(NNE, ip, k, ee, z, wt);
    double dN[NDE];
    shape_dN<NNE>(k, ee, z, dN);
    double dX[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NNE; ++n)
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) dX[ib * 3 + ic] += dN[n * 3 + ic] * s_x[slot][n * 3 + ib];
    const double J0 = det3(dX);
    double Xi[9];
    const bool ok = J0 > 0.0 && inv3(dX, Xi);
    if (!ok) atomicAdd(a.err, 1);
    double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int n = 0; n < NNE; ++n) {
      double g[3];
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
        g[ib] = dN[n * 3 + 0] * Xi[0 * 3 + ib] + dN[n * 3 + 1] * Xi[1 * 3 + ib] + dN[n * 3 + 2] * Xi[2 * 3 + ib];
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) {
        s_dN[slot][ip][n * 3 + ib] = g[ib];
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) Fr[ib * 3 + ic] += s_u[slot][n * 3 + ib] * g[ic];
      }
    }
    const double *qb = a.qpbuf + qrec(a.qp_base + e * NQP + ip);  // field i at qb[i * kRecStride]
    double Me[9], S9[9], Mi[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Me[i] = qb[i * kRecStride];
    double S6[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) S6[i] = qb[(9 + i) * kRecStride];
    v6_to_9(S6, S9);
    const double pJ = qb[15 * kRecStride];
    inv3(Me, Mi);
    double MeS[9], Sh9[9], Cr[9], CrMe[9];
    mm3(Me, S9, MeS);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Sh9[i * 3 + j] = MeS[i * 3 + 0] * Me[j * 3 + 0] + MeS[i * 3 + 1] * Me[j * 3 + 1] + MeS[i * 3 + 2] * Me[j * 3 + 2];
    mtm3(Fr, Fr, Cr);
    mm3(Cr, Me, CrMe);
    if constexpr (HE && TAN) {
      double eF[9], Ce[9], Sx[6];
      HeTan ht;
      mm3(Fr, Me, eF);
      mtm3(eF, eF, Ce);
      he_law(a.cpmats[0], eF, Ce, Sx, ht);
#pragma unroll
      for (int i = 0; i < 6; ++i) s_he[slot][ip][i] = ht.Ci[i];
      s_he[slot][ip][6] = ht.c1;
      s_he[slot][ip][7] = ht.c2;
      s_he[slot][ip][8] = ht.c3;
    }
    s_w[slot][ip] = ok ? J0 * wt * pJ : 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      s_Fr[slot][ip][i] = Fr[i];
      s_Me[slot][ip][i] = Me[i];
      s_MeS[slot][ip][i] = MeS[i];
      s_CrMe[slot][ip][i] = CrMe[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s_MiT[slot][ip][i * 3 + j] = Mi[j * 3 + i];
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      s_S[slot][ip][I] = S6[I];
      s_Sh[slot][ip][I] = Sh9[vi(I) * 3 + vj(I)];
    }
    if constexpr (TAN && !SPREAD_D) {
#pragma unroll
      for (int f = 0; f < 81; ++f) s_D[slot][ip][(f % 9) * 9 + f / 9] = qb[(16 + f) * kRecStride];
    }
  }
  if constexpr (TAN && SPREAD_D) {
    // the element's dMe/dF (NQP x 81) copied by all of its lanes, consecutive lanes on
    // consecutive record fields (field 16 + 9 c + r holds D[r][c]), rather than 81 loads on each
    // QP's lane
    if (active)
      for (int j = lane; j < NQP * 81; j += TPE) {
        const int ip = j / 81, f = j % 81;
        s_D[slot][ip][(f % 9) * 9 + f / 9] = a.qpbuf[qrec(a.qp_base + e * NQP + ip) + (16 + f) * kRecStride];
      }
  }
  __syncthreads();
  if constexpr (TAN) {
    // The row-dependent part of a K row is linear in sAA = sym(Fr^T (e_b (x) g_a)), so in the
    // row's g_a: per QP, the nine vectors z_bj of z = detJ w (Z : dMe/dF + Fr W) for
    // sAA = sym(Fr^T (e_b (x) e_j)) (b the row's component, j a component of g_a) replace the
    // per-row evaluation of B, L:B, Z, Z:dMe/dF and W; a row combines three of them.  They
    // are computed from dMe/dF and then overwrite it in LDS.
    constexpr int NT = EPB * TPE, NIT = EPB * NQP * 9, NPASS = (NIT + NT - 1) / NT;
    double zr[NPASS][9];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const int t = int(threadIdx.x) + p * NT;
      const int sl = t / (NQP * 9), ip = (t / 9) % NQP, bj = t % 9, b = bj / 3, j = bj - 3 * b;
      const int64_t iet = e_begin + int64_t(blockIdx.x) * EPB + sl;
#pragma unroll
      for (int c = 0; c < 9; ++c) zr[p][c] = 0.0;
      if (t < NIT && iet < e_end) {
        const double *Fr = s_Fr[sl][ip], *Me = s_Me[sl][ip];
        double sAA[9];  // sym(Fr^T (e_b (x) e_j))_ik = (Fr_bi d_jk + Fr_bk d_ij) / 2
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            sAA[i * 3 + k] = 0.5 * ((k == j ? Fr[b * 3 + i] : 0.0) + (i == j ? Fr[b * 3 + k] : 0.0));
        double T[9], Bm[9];
        mm3(sAA, Me, T);
        mtm3(Me, T, Bm);  // Me^T sAA Me
        double Y6[6];
        if constexpr (HE) {
          const double *hv = s_he[sl][ip];
          HeTan ht;
#pragma unroll
          for (int i = 0; i < 6; ++i) ht.Ci[i] = hv[i];
          ht.c1 = hv[6];
          ht.c2 = hv[7];
          ht.c3 = hv[8];
          he_dot(ht, Bm, Y6);
        } else {
          const int64_t et = a.elist ? int64_t(a.elist[iet]) : iet;
          lv_dot(a.grains[a.grain[et]].Lv, Bm, Y6);
        }
        double Y9[9];
        v6_to_9(Y6, Y9);
        const double *S6 = s_S[sl][ip];
        const double BS = Bm[0] * S6[0] + Bm[4] * S6[1] + Bm[8] * S6[2] +
                          2.0 * (0.5 * (Bm[1] + Bm[3]) * S6[3] + 0.5 * (Bm[5] + Bm[7]) * S6[4] +
                                 0.5 * (Bm[2] + Bm[6]) * S6[5]);
        // Z = CrMe Y + 2 sAA MeS - (B:S) Me^-T
        double Z[9], T2[9];
        mm3(s_CrMe[sl][ip], Y9, Z);
        mm3(sAA, s_MeS[sl][ip], T2);
        const double *MiT = s_MiT[sl][ip];
#pragma unroll
        for (int i = 0; i < 9; ++i) Z[i] += 2.0 * T2[i] - BS * MiT[i];
        // W = Me Y Me^T, FrW = Fr W
        double T3[9], W[9], FrW[9];
        mm3(Me, Y9, T3);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            W[i * 3 + k] = T3[i * 3 + 0] * Me[k * 3 + 0] + T3[i * 3 + 1] * Me[k * 3 + 1] + T3[i * 3 + 2] * Me[k * 3 + 2];
        mm3(Fr, W, FrW);
        // z[g L] = sum_ij Z_ij D[ij][g L] + (Fr W)[g L]
        const double *D = s_D[sl][ip];
        const double wq = s_w[sl][ip];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
          double sum = 0.0;
#pragma unroll
          for (int r = 0; r < 9; ++r) sum += Z[r] * D[r * 9 + c];
          zr[p][c] = wq * (sum + FrW[c]);
        }
      }
    }
    lds_barrier();  // every z_bj is computed from dMe/dF before any overwrites it
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const int t = int(threadIdx.x) + p * NT;
      if (t < NIT) {
        const int sl = t / (NQP * 9), ip = (t / 9) % NQP, bj = t % 9;
#pragma unroll
        for (int c = 0; c < 9; ++c) s_D[sl][ip][bj * 9 + c] = zr[p][c];
      }
    }
    lds_barrier();
  }
  const bool rowlane = active && lane < NDE;
  const int ra = lane / 3, rb = lane % 3;
  const double eb[3] = {rb == 0 ? 1.0 : 0.0, rb == 1 ? 1.0 : 0.0, rb == 2 ? 1.0 : 0.0};
  double K[TAN ? NDE : 1];
  if constexpr (TAN) {
#pragma unroll
    for (int i = 0; i < NDE; ++i) K[i] = 0.0;
  }
  double fe = 0.0;
  if (rowlane) {
#pragma unroll 1
  for (int ip = 0; ip < NQP; ++ip) {
    const double wq = s_w[slot][ip];
    const double *dN = s_dN[slot][ip];
    const double *Fr = s_Fr[slot][ip];
    const double ga[3] = {dN[ra * 3], dN[ra * 3 + 1], dN[ra * 3 + 2]};
    double Sh9[9];
    v6_to_9(s_Sh[slot][ip], Sh9);
    double ta[3];  // Sh dN_a
#pragma unroll
    for (int i = 0; i < 3; ++i) ta[i] = Sh9[i * 3] * ga[0] + Sh9[i * 3 + 1] * ga[1] + Sh9[i * 3 + 2] * ga[2];
    if constexpr (RES) fe += wq * (Fr[rb * 3] * ta[0] + Fr[rb * 3 + 1] * ta[1] + Fr[rb * 3 + 2] * ta[2]);
    if constexpr (TAN) {
      // z = sum_j g_a,j z_bj, plus the geometric term of row b: detJ w (Sh g_a)
      const double *zb = &s_D[slot][ip][rb * 27];
      double zz[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) zz[c] = ga[0] * zb[c] + ga[1] * zb[9 + c] + ga[2] * zb[18 + c];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double wt = wq * ta[k];
        zz[0 + k] += eb[0] * wt;
        zz[3 + k] += eb[1] * wt;
        zz[6 + k] += eb[2] * wt;
      }
#pragma unroll
      for (int w = 0; w < NNE; ++w) {
        const double h0 = dN[w * 3], h1 = dN[w * 3 + 1], h2 = dN[w * 3 + 2];
#pragma unroll
        for (int g = 0; g < 3; ++g) K[w * 3 + g] += zz[g * 3] * h0 + zz[g * 3 + 1] * h1 + zz[g * 3 + 2] * h2;
      }
    }
  }
  if constexpr (TAN) {
    double diag = 0.0;
#pragma unroll
    for (int i = 0; i < NDE; ++i) diag = (i == lane) ? K[i] : diag;
    if (!isnormal(diag)) atomicAdd(a.err, 1);
  }
  if constexpr (DFL) {  // f_defl row: K row . prescribed increments (src/load.cc:190-215)
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NDE; ++i) s += K[i] * s_dl[slot][i];
    fe = s;
  }
  }  // rowlane
  constexpr int LD = NDE + 1;
  if constexpr (WK && STG && NNE != 4) {
    // Gather mode: RE elements at a time, the rows go to LDS (in the dead dM/dF array) laid
    // out as the staging blocks, then the workgroup copies each node block, zero-padded to
    // whole lines, to its staging offset with consecutive threads on consecutive doubles.
    // (Linear tets store their 13-double rows directly: the transpose rounds cost more.)
    constexpr int PB = stage_block(NNE);
    constexpr int RE = cp_stage_round(EPB, NDE, EPB * NQP * 81);
    static_assert(RE * NDE * LD <= EPB * NQP * 81, "staging round does not fit");
    double *s_st = &s_D[0][0][0];
    lds_barrier();  // every lane is done with s_D
#pragma unroll 1
    for (int r0 = 0; r0 < EPB; r0 += RE) {
      if (rowlane && slot >= r0 && slot < r0 + RE) {
        double *row = s_st + ((slot - r0) * NDE + lane) * LD;
#pragma unroll
        for (int i = 0; i < NDE; ++i) row[i] = K[i];
        row[NDE] = fe;
      }
      lds_barrier();
      // two doubles per thread (blocks are whole 128-byte lines, so pairs are 16-byte aligned)
      constexpr int PB2 = PB / 2;
      for (int idx = threadIdx.x; idx < RE * NNE * PB2; idx += blockDim.x) {
        const int el = idx / (NNE * PB2), o = idx - el * (NNE * PB2);
        const int na = o / PB2, ob = 2 * (o - na * PB2);
        const int64_t ee = e_begin + int64_t(blockIdx.x) * EPB + r0 + el;
        if (ee >= e_end) continue;
        const int64_t so = a.soff[ee * NNE + na];
        if (so < 0) continue;
        const double *src = s_st + (el * NDE + na * 3) * LD + ob;
        dbl2 v;
        v.x = (ob < 3 * LD) ? src[0] : 0.0;
        v.y = (ob + 1 < 3 * LD) ? src[1] : 0.0;
        __builtin_nontemporal_store(v, reinterpret_cast<dbl2 *>(a.stage + so + ob));
      }
      lds_barrier();
    }
    return;
  }
  if (!rowlane) return;
  const int32_t dof = s_dof[slot][lane];
  if (dof <= 0) return;
  if constexpr (STG) {  // gather mode: the row straight into its staging block
    double *srow = a.stage + a.soff[e * NNE + ra] + rb * LD;
    if constexpr (WK) {
#pragma unroll
      for (int i = 0; i < NDE; ++i) srow[i] = K[i];
    }
    if constexpr (WR) srow[NDE] = fe;
    return;
  }
  const int16_t *off = a.off + (e * NNE + ra) * NNE;
  if constexpr (WR) {
    if (off[ra] <= -2) a.f[dof - 1] = fe;
    else a.f[dof - 1] += fe;
  }
  if constexpr (WK) {
    const int32_t node = s_node[slot][ra];
    int bp = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) bp += (j < rb && s_dof[slot][ra * 3 + j] > 0) ? 1 : 0;
    double *row = a.vals + (int64_t(a.row_start[node]) + int64_t(bp) * a.rowlen[node]);
    constexpr int H1 = (NNE + 1) / 2;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int w0 = half ? H1 : 0, w1 = half ? NNE : H1;
      int pos[3 * H1];
      double old[3 * H1];
#pragma unroll
      for (int w = w0; w < w1; ++w) {
        const int o = off[w];
        const bool first = o <= -2;
        const int base = first ? (-o - 2) : o;
        int gp = 0;
#pragma unroll
        for (int gg = 0; gg < 3; ++gg) {
          const bool fr = (o != -1) && s_dof[slot][w * 3 + gg] > 0;
          pos[(w - w0) * 3 + gg] = fr ? base + gp : -1;
          gp += fr ? 1 : 0;
          old[(w - w0) * 3 + gg] = (fr && !first) ? row[base + gp - 1] : 0.0;
        }
      }
#pragma unroll
      for (int w = w0; w < w1; ++w)
#pragma unroll
        for (int gg = 0; gg < 3; ++gg) {
          const int pp = pos[(w - w0) * 3 + gg];
          if (pp >= 0) row[pp] = old[(w - w0) * 3 + gg] + K[w * 3 + gg];
        }
    }
  }
}

// ------------------------------------------------------------------- CM3F element kernel
// The three-field (CM3F) element of crystal plasticity with the hyperelastic law, condensed to
// the displacement DOFs (ConstitutiveModelIntregrate::integrate_ss, src/constitutive_model.cc
// :720-880, total Lagrangian, one theta and one P per element; the blocks of CM_ThreeField,
// src/constitutive_model_3f.cc:189-470; condense_K_3F_to_1F / condense_F_3F_to_1F,
// src/condense.cc:304-385).  Per QP, from the tangent pass's qpbuf: M = pF(n+1)^-1, pJ and
// dM/dFb (Fb = (theta/J)^(1/3) F, the model's F); eF = F M with the unscaled F, Sd and Ld the
// deviatoric PK2 and tangent of the neo-Hookean law at eF, dU, ddU = kappa U', kappa U'' at
// theta_e = theta det M.  For a DOF k (node a, component b) with G_k = e_b (x) dN_a:
//   Phi_k = sym(G_k^T F), Psi_k = M^T Phi_k M, tr_k = tr(G_k F^-1),
//   dM_k = dM/dFb : dFb/du_k, dFb/du_k = f (G_k - tr_k/3 F), Lambda_k = tr(M^-1 dM_k),
//   X_k = Psi_k + sym(M^T F^T F dM_k) (the change of the elastic strain),
// and with cw = detJ w pJ (1/Jn), r_k = Psi_k : Sd + P J tr_k:
//   Ru_k  += cw r_k,                       Kup_k += cw J tr_k,
//   Kuu_kl += cw (Psi_k : Ld : X_l + P J (tr_k tr_l - tr(G_k F^-1 G_l F^-1))
//                 + sym(G_k^T G_l) : M Sd M^T + 2 (Phi_k M Sd) : dM_l - Lambda_l r_k),
//   Kut_k += cw (2 sym(M^T Phi_k dM_t) : Sd + Psi_k : Ld : X_t - Lambda_t r_k),
//   Ktu_k += cw (JM (ddU theta JM + dU) - (dU JM - P)) Lambda_k,
//   Rp += cw (J - theta), Rt += cw (dU JM - P), Ktt += cw (JM dU Lambda_t + JM^2 ddU
//   - Lambda_t (dU JM - P)), Ktp -= cw, with the theta direction dM_t = dM/dFb : Fb/(3 theta).
// The element's rows: one lane per DOF row, the QPs one at a time through LDS (the QP's
// tensors, then every column's LX = Ld : X, dM, tr, Lambda, then the row sums).
template <int NNE, int MODE, bool STG>
__global__ __launch_bounds__(256) void cp_tf_elem(CPElemArgs a, int64_t e_begin, int64_t e_end) {
  constexpr int NQP = nqp_of(NNE);
  constexpr int NDE = NNE * 3;
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = 256 / TPE;
  constexpr bool RES = (MODE & MODE_RESIDUAL) != 0;
  constexpr bool TAN = (MODE & MODE_TANGENT) != 0;
  constexpr bool DFL = (MODE & MODE_DEFL_LOAD) != 0;
  static_assert(!DFL || (TAN && !RES), "MODE_DEFL_LOAD goes with MODE_TANGENT alone");
  constexpr bool WK = TAN && !DFL;
  constexpr bool WR = RES || DFL;
  // per QP tensors in s_q: Fr 0, M 9, M^-1 18, M^T Cr 27, Sd 36, M Sd M^T 45, Ci (Voigt) 54,
  // c1 c2 c3 60, dM/dFb : Fr 63, dM_t 72, Ld : X_t (Voigt) 81, cw 87, J 88, JM 89, dU 90,
  // ddU 91, f 92, Lambda_t 93, dU JM - P 94, M Sd 95, M Sd dM_t^T 104, Fr^-1 113
  constexpr int QD = 122;
  constexpr int CD = 17;  // per DOF column: LX (Voigt) 0, dM 6, tr 15, Lambda 16
  __shared__ double s_x[EPB][NDE], s_u[EPB][NDE];
  __shared__ double s_dl[EPB][DFL ? NDE : 1];
  __shared__ int32_t s_dof[EPB][NDE], s_node[EPB][NNE];
  __shared__ double s_dN[EPB][NQP][NDE], s_w[EPB][NQP];
  __shared__ double s_q[EPB][QD], s_D[EPB][81], s_col[EPB][NDE][CD], s_gn[EPB][NDE];
  __shared__ double s_vec[EPB][2][NDE];

  const int slot = threadIdx.x / TPE, lane = threadIdx.x % TPE;
  const int64_t ie = e_begin + int64_t(blockIdx.x) * EPB + slot;
  const bool active = ie < e_end;
  const int64_t e = (a.elist && active) ? int64_t(a.elist[ie]) : ie;
  if (active && lane < NNE) {
    const int32_t node = a.conn[e * NNE + lane];
    s_node[slot][lane] = node;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      s_x[slot][lane * 3 + b] = a.X[int64_t(node) * 3 + b];
      const int32_t d = a.dofid[int64_t(node) * 3 + b];
      s_dof[slot][lane * 3 + b] = d;
      double u = 0.0, dl = 0.0;
      if (d > 0) {
        u = a.r[d - 1] + (a.d_r ? a.d_r[d - 1] : 0.0);
      } else if (d < 0) {
        dl = a.defl_d ? a.defl_d[-d - 1] : 0.0;
        u = (a.defl ? a.defl[-d - 1] : 0.0) + (DFL ? 0.0 : dl);
      }
      s_u[slot][lane * 3 + b] = u;
      if constexpr (DFL) s_dl[slot][lane * 3 + b] = dl;
    }
  }
  __syncthreads();
  if (active && lane < NQP) {  // physical shape-function gradients and detJ w of every QP
    const int ip = lane;
    double k, ee, z, wt;
    int_point(NNE, ip, k, ee, z, wt);
    double dN[NDE];
    shape_dN<NNE>(k, ee, z, dN);
    double dX[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NNE; ++n)
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) dX[ib * 3 + ic] += dN[n * 3 + ic] * s_x[slot][n * 3 + ib];
    const double J0 = det3(dX);
    double Xi[9];
    const bool ok = J0 > 0.0 && inv3(dX, Xi);
    if (!ok) atomicAdd(a.err, 1);
#pragma unroll
    for (int n = 0; n < NNE; ++n)
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
        s_dN[slot][ip][n * 3 + ib] =
            dN[n * 3 + 0] * Xi[0 * 3 + ib] + dN[n * 3 + 1] * Xi[1 * 3 + ib] + dN[n * 3 + 2] * Xi[2 * 3 + ib];
    s_w[slot][ip] = ok ? J0 * wt : 0.0;
  }
  const double theta = active ? a.tf_V[e] + a.tf_dV[e] : 1.0;
  const double Pp = active ? a.tf_P[e] + a.tf_dP[e] : 0.0;
  const CPMaterial cm = a.cpmats[0];
  const bool rowlane = active && lane < NDE;
  const int ra = lane / 3, rb = lane % 3;
  double K[NDE];
#pragma unroll
  for (int i = 0; i < NDE; ++i) K[i] = 0.0;
  double Ru = 0.0, Kup = 0.0, Kut = 0.0, Ktu = 0.0, Rp = 0.0, Rt = 0.0, Ktt = 0.0, Ktp = 0.0;
  __syncthreads();
#pragma unroll 1
  for (int ip = 0; ip < NQP; ++ip) {
    if (active) {
      const double *qb = a.qpbuf + qrec(a.qp_base + e * NQP + ip);
      for (int f = lane; f < 81; f += TPE) s_D[slot][(f % 9) * 9 + f / 9] = qb[(16 + f) * kRecStride];
    }
    lds_barrier();
    // the QP's tensors in three rounds, lanes of an element working side by side
    double *q = s_q[slot];
    if (active && lane == 0) {  // Fr, M, M^-1, Fr^-1, J, det M
      const double *dN = s_dN[slot][ip];
      double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
      for (int n = 0; n < NNE; ++n)
#pragma unroll
        for (int ib = 0; ib < 3; ++ib)
#pragma unroll
          for (int ic = 0; ic < 3; ++ic) Fr[ib * 3 + ic] += s_u[slot][n * 3 + ib] * dN[n * 3 + ic];
      const double *qb = a.qpbuf + qrec(a.qp_base + e * NQP + ip);
      double Me[9], Mi[9], Fi[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) Me[i] = qb[i * kRecStride];
      const double J = det3(Fr);
      const bool okq = J > 0.0 && theta > 0.0 && inv3(Fr, Fi) && inv3(Me, Mi);
      if (!okq) atomicAdd(a.err, 1);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        q[i] = Fr[i];
        q[9 + i] = Me[i];
        q[18 + i] = Mi[i];
        q[113 + i] = Fi[i];
      }
      q[87] = okq ? s_w[slot][ip] * qb[15 * kRecStride] : 0.0;  // detJ w / Jn, 1/Jn = pJ
      q[88] = J;
      q[89] = det3(Me);
      q[92] = okq ? cbrt(theta / J) : 1.0;
    }
    lds_barrier();
    if (active && lane == 0) {  // M^T Cr
      double Cr[9], MtC[9];
      mtm3(q, q, Cr);
      mtm3(q + 9, Cr, MtC);
#pragma unroll
      for (int i = 0; i < 9; ++i) q[27 + i] = MtC[i];
    } else if (active && lane == 1) {
      // neo-Hookean deviatoric PK2 and tangent coefficients at Ce = (F M)^T F M (he_law
      // without kappa), and the volumetric derivatives at theta_e = theta det M
      double eF[9], Ce[9], Ci9[9];
      mm3(q, q + 9, eF);
      mtm3(eF, eF, Ce);
      inv3(Ce, Ci9);
      const double d = pow(det3(Ce), -1.0 / 3.0);
      const double trC = Ce[0] + Ce[4] + Ce[8];
      const double aa = 2.0 * cm.m10 * d;
#pragma unroll
      for (int i = 0; i < 9; ++i) q[36 + i] = aa * ((i % 4 == 0) ? 1.0 : 0.0) - aa * trC * (1.0 / 3.0) * Ci9[i];
#pragma unroll
      for (int I = 0; I < 6; ++I) q[54 + I] = Ci9[vi(I) * 3 + vj(I)];
      q[60] = 2.0 * aa * trC * (1.0 / 9.0);
      q[61] = -2.0 * aa * (1.0 / 3.0);
      q[62] = 2.0 * aa * trC * (1.0 / 3.0);
      const double JM = q[89];
      double dU, ddU;
      vol_pot(cm.vol_flag, theta * JM, dU, ddU);
      q[90] = dU * cm.kappa;
      q[91] = ddU * cm.kappa;
      q[94] = dU * cm.kappa * JM - Pp;
    } else if (active && lane >= 2 && lane < 11) {  // row r of dM/dFb : Fr and of dM_t
      const int r = lane - 2;
      double sd = 0.0;
#pragma unroll
      for (int c = 0; c < 9; ++c) sd += s_D[slot][r * 9 + c] * q[c];
      q[63 + r] = sd;
      q[72 + r] = q[92] / (3.0 * theta) * sd;  // dM_t = f / (3 theta) dM/dFb : Fr
    }
    lds_barrier();
    if (active && lane == 0) {  // tr(M^-1 dM_t) and Ld : sym(M^T Cr dM_t)
      double Lt = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Lt += q[18 + i * 3 + j] * q[72 + j * 3 + i];
      q[93] = Lt;
      double T[9], Xt[9], LXt[6];
      mm3(q + 27, q + 72, T);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Xt[i * 3 + j] = 0.5 * (T[i * 3 + j] + T[j * 3 + i]);
      HeTan td;
#pragma unroll
      for (int I = 0; I < 6; ++I) td.Ci[I] = q[54 + I];
      td.c1 = q[60];
      td.c2 = q[61];
      td.c3 = q[62];
      he_dot(td, Xt, LXt);
#pragma unroll
      for (int I = 0; I < 6; ++I) q[81 + I] = LXt[I];
    } else if (active && lane == 1) {  // M Sd, M Sd M^T, M Sd dM_t^T
      double MeSd[9];
      mm3(q + 9, q + 36, MeSd);
#pragma unroll
      for (int i = 0; i < 9; ++i) q[95 + i] = MeSd[i];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          q[45 + i * 3 + j] = MeSd[i * 3 + 0] * q[9 + j * 3 + 0] + MeSd[i * 3 + 1] * q[9 + j * 3 + 1] +
                              MeSd[i * 3 + 2] * q[9 + j * 3 + 2];
          q[104 + i * 3 + j] = MeSd[i * 3 + 0] * q[72 + j * 3 + 0] + MeSd[i * 3 + 1] * q[72 + j * 3 + 1] +
                               MeSd[i * 3 + 2] * q[72 + j * 3 + 2];
        }
    }
    lds_barrier();
    const double *dN = s_dN[slot][ip];
    double Psi[9], Phi[9];
    if (rowlane) {  // the DOF's column data, and the row's own Psi
      const double *Fr = q, *Me = q + 9, *Fi = q + 113;
      const double ga[3] = {dN[ra * 3], dN[ra * 3 + 1], dN[ra * 3 + 2]};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Phi[i * 3 + j] = 0.5 * (ga[i] * Fr[rb * 3 + j] + ga[j] * Fr[rb * 3 + i]);
      double T[9];
      mm3(Phi, Me, T);
      mtm3(Me, T, Psi);
      // g_a = F^-T dN_a: tr_k = g_a,b
      double g[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i] = ga[0] * Fi[0 * 3 + i] + ga[1] * Fi[1 * 3 + i] + ga[2] * Fi[2 * 3 + i];
      s_gn[slot][lane] = g[rb];
      const double trk = g[rb], fac = q[92];
      double dM[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const double *Dr = s_D[slot] + r * 9 + rb * 3;
        dM[r] = fac * (Dr[0] * ga[0] + Dr[1] * ga[1] + Dr[2] * ga[2] - trk * (1.0 / 3.0) * q[63 + r]);
      }
      double Lk = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Lk += q[18 + i * 3 + j] * dM[j * 3 + i];
      double X[9], V[9];
      mm3(q + 27, dM, V);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[i * 3 + j] = Psi[i * 3 + j] + 0.5 * (V[i * 3 + j] + V[j * 3 + i]);
      HeTan td;
#pragma unroll
      for (int I = 0; I < 6; ++I) td.Ci[I] = q[54 + I];
      td.c1 = q[60];
      td.c2 = q[61];
      td.c3 = q[62];
      double LX[6];
      he_dot(td, X, LX);
      double *col = s_col[slot][lane];
#pragma unroll
      for (int I = 0; I < 6; ++I) col[I] = LX[I];
#pragma unroll
      for (int r = 0; r < 9; ++r) col[6 + r] = dM[r];
      col[15] = trk;
      col[16] = Lk;
    }
    lds_barrier();
    if (active) {
      const double cw = q[87], J = q[88], JM = q[89], dU = q[90], ddU = q[91], Lt = q[93], vt = q[94];
      Rp += cw * (J - theta);  // compute_Rp
      Rt += cw * vt;           // compute_Rt
      Ktt += cw * (JM * dU * Lt + JM * JM * ddU - Lt * vt);  // compute_Ktt
      Ktp -= cw;               // compute_Ktp
    }
    if (rowlane) {
      const double cw = q[87], J = q[88], JM = q[89], dU = q[90], ddU = q[91], Lt = q[93], vt = q[94];
      const double *Sd = q + 36;
      double PsiS = 0.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) PsiS += Psi[i] * Sd[i];
      const double trk = s_col[slot][lane][15], Lk = s_col[slot][lane][16];
      const double rk = PsiS + Pp * J * trk;
      // Psi : Ld : X_t and 2 sym(M^T Phi dM_t) : Sd = 2 Phi : (M Sd dM_t^T)
      const double *LXt = q + 81, *W = q + 104;
      const double PLt = Psi[0] * LXt[0] + Psi[4] * LXt[1] + Psi[8] * LXt[2] +
                         2.0 * (Psi[1] * LXt[3] + Psi[5] * LXt[4] + Psi[2] * LXt[5]);
      double DtS = 0.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) DtS += 2.0 * Phi[i] * W[i];
      Ru += cw * rk;                                                // compute_Ru
      Kup += cw * J * trk;                                          // compute_Kup
      Kut += cw * (DtS + PLt - Lt * rk);                            // compute_Kut
      Ktu += cw * (JM * (ddU * theta * JM + dU) * Lk - Lk * vt);    // compute_Ktu
      if constexpr (TAN) {
        // Z = 2 Phi_k M Sd, t = M Sd M^T dN_a
        double Z[9];
        mm3(Phi, q + 95, Z);
        const double *Sh = q + 45;
        const double ga[3] = {dN[ra * 3], dN[ra * 3 + 1], dN[ra * 3 + 2]};
        double ta[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) ta[i] = Sh[i * 3] * ga[0] + Sh[i * 3 + 1] * ga[1] + Sh[i * 3 + 2] * ga[2];
        double gk[3];  // g_a (F^-T dN_a), all three components
#pragma unroll
        for (int i = 0; i < 3; ++i) gk[i] = s_gn[slot][ra * 3 + i];
        // not unrolled: the column data are read from LDS one column at a time (unrolled, the
        // compiler hoists all of them and spills)
#pragma unroll 1
        for (int l = 0; l < NDE; ++l) {
          const int w = l / 3, bp = l % 3;
          const double *col = s_col[slot][l];
          const double PLX = Psi[0] * col[0] + Psi[4] * col[1] + Psi[8] * col[2] +
                             2.0 * (Psi[1] * col[3] + Psi[5] * col[4] + Psi[2] * col[5]);
          double dMS = 0.0;
#pragma unroll
          for (int i = 0; i < 9; ++i) dMS += 2.0 * Z[i] * col[6 + i];
          const double GGS = (bp == rb) ? ta[0] * dN[w * 3] + ta[1] * dN[w * 3 + 1] + ta[2] * dN[w * 3 + 2] : 0.0;
          const double tr2 = gk[bp] * s_gn[slot][w * 3 + rb];
          K[l] += cw * (PLX + Pp * J * (trk * col[15] - tr2) + GGS + dMS - col[16] * rk);  // compute_Kuu
        }
      }
    }
    lds_barrier();  // the next QP overwrites s_q, s_D, s_col, s_gn
  }
  // condensation: K = Kuu - Kup Ktu / Ktp + Kup Ktt Kpu / (Ktp Kpt) - Kut Kpu / Kpt, with
  // Kpu = Kup^T, Kpt = Ktp (IntegrateThreeFieldStiffness::out_loop, constitutive_model.cc:531-555)
  if (rowlane) {
    s_vec[slot][0][lane] = Kup;
    s_vec[slot][1][lane] = Ktu;
  }
  lds_barrier();
  double fe = 0.0;
  if (rowlane) {
    const double Kpt = Ktp;
    if constexpr (TAN) {
#pragma unroll
      for (int l = 0; l < NDE; ++l)
        K[l] = K[l] - Kup / Ktp * s_vec[slot][1][l] + Kup / Ktp * Ktt / Kpt * s_vec[slot][0][l] -
               Kut / Kpt * s_vec[slot][0][l];
      double diag = 0.0;
#pragma unroll
      for (int i = 0; i < NDE; ++i) diag = (i == lane) ? K[i] : diag;
      if (!isnormal(diag)) atomicAdd(a.err, 1);
    }
    if constexpr (RES) fe = Ru + Kup / Ktp * Ktt / Kpt * Rp - Kup / Ktp * Rt - Kut / Kpt * Rp;
    if constexpr (DFL) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < NDE; ++i) s += K[i] * s_dl[slot][i];
      fe = s;
    }
    if constexpr (RES) {  // the element update's data (compute_d_theta_dP)
      double *tb = a.tf_buf + e * kTfBuf;
      tb[lane] = Kup;
      tb[kTfKtu + lane] = Ktu;
      if (lane == 0) {
        tb[NDE] = Rp;
        tb[NDE + 1] = Rt;
        tb[NDE + 2] = Ktt;
        tb[NDE + 3] = -Ktp;
      }
    }
  }
  if (!rowlane) return;
  const int32_t dof = s_dof[slot][lane];
  if (dof <= 0) return;
  constexpr int LD = NDE + 1;
  if constexpr (STG) {  // gather mode: the row straight into its staging block
    double *srow = a.stage + a.soff[e * NNE + ra] + rb * LD;
    if constexpr (WK) {
#pragma unroll
      for (int i = 0; i < NDE; ++i) srow[i] = K[i];
    }
    if constexpr (WR) srow[NDE] = fe;
    return;
  }
  const int16_t *off = a.off + (e * NNE + ra) * NNE;
  if constexpr (WR) {
    if (off[ra] <= -2) a.f[dof - 1] = fe;
    else a.f[dof - 1] += fe;
  }
  if constexpr (WK) {
    const int32_t node = s_node[slot][ra];
    int bp = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) bp += (j < rb && s_dof[slot][ra * 3 + j] > 0) ? 1 : 0;
    double *row = a.vals + (int64_t(a.row_start[node]) + int64_t(bp) * a.rowlen[node]);
#pragma unroll
    for (int w = 0; w < NNE; ++w) {
      const int o = off[w];
      if (o == -1) continue;
      const bool first = o <= -2;
      const int base = first ? (-o - 2) : o;
      int gp = 0;
#pragma unroll
      for (int gg = 0; gg < 3; ++gg) {
        if (s_dof[slot][w * 3 + gg] <= 0) continue;
        double *p = row + base + gp;
        *p = first ? K[w * 3 + gg] : *p + K[w * 3 + gg];
        ++gp;
      }
    }
  }
}

template <int NNE, int MODE>
void launch_tfe(const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = 256 / TPE;
  const int64_t n = e1 - e0;
  if (n <= 0) return;
  const unsigned grid = unsigned((n + EPB - 1) / EPB);
  if (a.stage) cp_tf_elem<NNE, MODE, true><<<grid, 256, 0, s>>>(a, e0, e1);
  else cp_tf_elem<NNE, MODE, false><<<grid, 256, 0, s>>>(a, e0, e1);
}

template <int NNE>
void launch_tfe_mode(int mode, const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  switch (mode & (MODE_RESIDUAL | MODE_TANGENT | MODE_DEFL_LOAD)) {
    case MODE_RESIDUAL: launch_tfe<NNE, MODE_RESIDUAL>(a, e0, e1, s); break;
    case MODE_TANGENT: launch_tfe<NNE, MODE_TANGENT>(a, e0, e1, s); break;
    case MODE_RESIDUAL | MODE_TANGENT: launch_tfe<NNE, MODE_RESIDUAL | MODE_TANGENT>(a, e0, e1, s); break;
    case MODE_TANGENT | MODE_DEFL_LOAD: launch_tfe<NNE, MODE_TANGENT | MODE_DEFL_LOAD>(a, e0, e1, s); break;
    default: break;
  }
}

template <int NNE>
void launch_kin(const CPArgs &a, hipStream_t s) {
  const int64_t n = a.ne * nqp_of(NNE);
  if (n <= 0) return;
  cp_kinematics<NNE><<<unsigned((n + 255) / 256), 256, 0, s>>>(a);
}

template <int NNE, int MODE>
void launch_ce(const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = (NNE == 8) ? 4 : 256 / TPE;
  const int64_t n = e1 - e0;
  if (n <= 0) return;
  const unsigned grid = unsigned((n + EPB - 1) / EPB);
  if (a.law == 1) {
    if (a.stage) cp_elem<true, NNE, MODE, true><<<grid, EPB * TPE, 0, s>>>(a, e0, e1);
    else cp_elem<true, NNE, MODE, false><<<grid, EPB * TPE, 0, s>>>(a, e0, e1);
  } else {
    if (a.stage) cp_elem<false, NNE, MODE, true><<<grid, EPB * TPE, 0, s>>>(a, e0, e1);
    else cp_elem<false, NNE, MODE, false><<<grid, EPB * TPE, 0, s>>>(a, e0, e1);
  }
}

template <int NNE>
void launch_ce_mode(int mode, const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  switch (mode & (MODE_RESIDUAL | MODE_TANGENT | MODE_DEFL_LOAD)) {
    case MODE_RESIDUAL: launch_ce<NNE, MODE_RESIDUAL>(a, e0, e1, s); break;
    case MODE_TANGENT: launch_ce<NNE, MODE_TANGENT>(a, e0, e1, s); break;
    case MODE_RESIDUAL | MODE_TANGENT: launch_ce<NNE, MODE_RESIDUAL | MODE_TANGENT>(a, e0, e1, s); break;
    case MODE_TANGENT | MODE_DEFL_LOAD: launch_ce<NNE, MODE_TANGENT | MODE_DEFL_LOAD>(a, e0, e1, s); break;
    default: break;
  }
}

// ------------------------------------------------------------------- output variables
// One thread per element of a block: the stress from the stored state at the level, as
// constitutive_model_update_output_variables forms it (src/constitutive_model.cc:1738-1875):
// eF = F pF^-1 (J2 and the hyperelastic model: the elastic left stretch from Cp^-1), the elastic
// stress of the model on eF, sigma = eF eS eF^T / theta_e, S = theta F^-1 sigma F^-T,
// E = (F^T F - I)/2 per QP; sigma and e = (I - (F F^T)^-1)/2 (compute_Eulerian_Almansi_strain,
// src/utils.cc:46-56) averaged over the element with detJ w, shear strains doubled.
template <int NNE>
__global__ __launch_bounds__(256) void output_kernel(OutArgs o) {
  constexpr int NQP = nqp_of(NNE);
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= o.ne) return;
  double x[NNE * 3];
#pragma unroll
  for (int n = 0; n < NNE; ++n) {
    const int32_t node = o.conn[e * NNE + n];
#pragma unroll
    for (int b = 0; b < 3; ++b) x[n * 3 + b] = o.X[int64_t(node) * 3 + b];
  }
  // the reference's output order XX, YY, ZZ, YZ, XZ, XY as (i, j)
  const int oi[6] = {0, 1, 2, 1, 0, 0}, oj[6] = {0, 1, 2, 2, 2, 1};
  double vol = 0.0, sa[6] = {0, 0, 0, 0, 0, 0}, ea[6] = {0, 0, 0, 0, 0, 0};
  const DevMaterial m = o.mats[o.mat ? o.mat[e] : 0];
  for (int ip = 0; ip < NQP; ++ip) {
    double k, ee, z, wt;
    int_point(NNE, ip, k, ee, z, wt);
    double dN[NNE * 3];
    shape_dN<NNE>(k, ee, z, dN);
    double dX[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NNE; ++n)
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) dX[ib * 3 + ic] += dN[n * 3 + ic] * x[n * 3 + ib];
    const double w = det3(dX) * wt;
    const int64_t q = o.qp_base + e * NQP + ip;
    const int64_t NQ = o.n_qp;
    double F[9], Fi[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = o.F[i * NQ + q];
    const double J = det3(F);
    inv3(F, Fi);
    double sig[9];  // Cauchy
    if (o.model == PGF_MODEL_CRYSTAL_PLASTICITY) {
      const CPMaterial cm = o.cpmats[0];
      double pF[9], pFi[9], eF[9], Ce[9], S6[6], S9[9], T[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) pF[i] = o.P[i * NQ + q];
      inv3(pF, pFi);
      mm3(F, pFi, eF);
      mtm3(eF, eF, Ce);
      HeTan t;
      if (cm.law == 1) el_stress<true>(cm, o.grains[o.qgrain[q - o.qp_base]], eF, Ce, S6, t);
      else el_stress<false>(cm, o.grains[o.qgrain[q - o.qp_base]], eF, Ce, S6, t);
      v6_to_9(S6, S9);
      const double rte = 1.0 / det3(eF);
      mm3(eF, S9, T);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          sig[i * 3 + j] = rte * (T[i * 3] * eF[j * 3] + T[i * 3 + 1] * eF[j * 3 + 1] + T[i * 3 + 2] * eF[j * 3 + 2]);
    } else {
      // bbar_e = J^-2/3 F Cp^-1 F^T (Cp^-1 = I for the hyperelastic model); tau = G dev bbar_e +
      // kappa J U'(J) 1, scaled by (1 - w) with J2 damage; sigma = tau / J
      double Cn[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      if (o.model == PGF_MODEL_J2_PLASTICITY) {
        const double c6[6] = {o.P[0 * NQ + q], o.P[1 * NQ + q], o.P[2 * NQ + q],
                              o.P[3 * NQ + q], o.P[4 * NQ + q], o.P[5 * NQ + q]};
        Cn[0] = c6[0]; Cn[4] = c6[1]; Cn[8] = c6[2];
        Cn[1] = Cn[3] = c6[3];
        Cn[5] = Cn[7] = c6[4];
        Cn[2] = Cn[6] = c6[5];
      }
      double T[9], b[9];
      mm3(F, Cn, T);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          b[i * 3 + j] = T[i * 3] * F[j * 3] + T[i * 3 + 1] * F[j * 3 + 1] + T[i * 3 + 2] * F[j * 3 + 2];
      const double cj = cbrt(J);
      const double Jm23 = 1.0 / (cj * cj);
      const double tr3 = (b[0] + b[4] + b[8]) * (1.0 / 3.0);
      double dU, d2U;
      vol_pot(m.volPotFlag, J, dU, d2U);
      const double om = (o.model == PGF_MODEL_J2_PLASTICITY && m.mu > 0.0) ? 1.0 - o.Sc[2 * NQ + q] : 1.0;
      const double rJ = 1.0 / J;
      if (o.model == PGF_MODEL_HYPER_ELASTICITY && m.devPotFlag == 2) {
        // the linear law in C: tau = G (b b - b) + kappa J U'(J) 1
        double bb[9];
        mm3(b, b, bb);
#pragma unroll
        for (int i = 0; i < 9; ++i) sig[i] = rJ * (m.G * (bb[i] - b[i]) + ((i % 4 == 0) ? m.kappa * J * dU : 0.0));
      } else {
#pragma unroll
        for (int i = 0; i < 9; ++i)
          sig[i] = om * rJ * (m.G * Jm23 * (b[i] - ((i % 4 == 0) ? tr3 : 0.0)) + ((i % 4 == 0) ? m.kappa * J * dU : 0.0));
      }
    }
    // S = J F^-1 sigma F^-T, E = (F^T F - I)/2, e = (I - (F F^T)^-1)/2 = (I - Fi^T Fi)/2
    double T[9], S9[9], C[9], bi[9];
    mm3(Fi, sig, T);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        S9[i * 3 + j] = J * (T[i * 3] * Fi[j * 3] + T[i * 3 + 1] * Fi[j * 3 + 1] + T[i * 3 + 2] * Fi[j * 3 + 2]);
    mtm3(F, F, C);
    mtm3(Fi, Fi, bi);
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      const int ij = oi[I] * 3 + oj[I];
      o.S[I * NQ + q] = S9[ij];
      o.E[I * NQ + q] = 0.5 * (C[ij] - (I < 3 ? 1.0 : 0.0));
      sa[I] += w * sig[ij];
      ea[I] += w * (I < 3 ? 1.0 : 2.0) * 0.5 * ((I < 3 ? 1.0 : 0.0) - bi[ij]);
    }
    vol += w;
  }
  const double rv = 1.0 / vol;
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    o.sig[I * o.ne + e] = sa[I] * rv;
    o.eps[I * o.ne + e] = ea[I] * rv;
  }
}

// ------------------------------------------------------------------- learned warm start
// Features of one QP (device QP order rows of 34): F(n) pF(n)^-1, F(n+1) - F(n), tau_a(n)/g(n),
// grain quaternion; with shift/scale the row is normalised as (x - shift) * scale.
__global__ void cp_features(CPArgs a, const float *gquat, const float *shift, const float *scale, float *X) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const int64_t NQ = a.n_qp, qg = a.qp_base + t;
  double Fn[9], pFn[9], pFi[9], Fe[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    Fn[i] = a.F_n[i * NQ + qg];
    pFn[i] = a.pF_n[i * NQ + qg];
  }
  if (!inv3(pFn, pFi)) {
#pragma unroll
    for (int i = 0; i < 9; ++i) pFi[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  mm3(Fn, pFi, Fe);
  float f[34];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f[i] = float(Fe[i]);
    f[9 + i] = float(a.F_np1[i * NQ + qg] - Fn[i]);
  }
  // the resolved shear stresses at n over the slip resistance at n (the flow rule's ratio)
  const double rg = 1.0 / a.g_n[qg];
#pragma unroll
  for (int i = 0; i < 12; ++i) f[18 + i] = float(a.tau_n[i * NQ + qg] * rg);
  const int gid = a.qgrain[t];
#pragma unroll
  for (int i = 0; i < 4; ++i) f[30 + i] = gquat[gid * 4 + i];
  float *row = X + qg * 34;
#pragma unroll
  for (int i = 0; i < 34; ++i) row[i] = shift ? (f[i] - shift[i]) * scale[i] : f[i];
}

// Targets: dt * gamma_dot(n+1), rows of 12.
__global__ void cp_targets(CPArgs a, float *T) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const int64_t NQ = a.n_qp, qg = a.qp_base + t;
#pragma unroll
  for (int i = 0; i < 12; ++i) T[qg * 12 + i] = float(a.dt * a.gdot[i * NQ + qg]);
}

// Warm start from predicted slip increments (block-local rows of Y, ldy apart):
// M = I - sum_a dgamma_a P_a (first order in the increments), lambda = lambda(n) and g from
// the closed-form hardening update at the predicted total slip rate.
__global__ void cp_warm_from_slip(CPArgs a, const float *Y, int ldy, const float *out_scale, double *warm) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const int64_t NQ = a.n_qp, qg = a.qp_base + t;
  const DevGrain &gr = a.grains[a.qgrain[t]];
  const CPMaterial cm = a.cpmats[0];
  double dg[12], gsum = 0.0;
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    dg[i] = double(Y[t * ldy + i]) / double(out_scale[i]);
    gsum += fabs(dg[i]);
  }
  gsum /= a.dt;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double s = 0.0;
#pragma unroll
    for (int al = 0; al < 12; ++al) s += dg[al] * gr.P[al][i];
    warm[i * NQ + qg] = ((i % 4 == 0) ? 1.0 : 0.0) - s;
  }
  double gs = cm.gs0;
  if (cm.w != 0.0) gs = (gsum / cm.gds > 1.0e-15) ? cm.gs0 * pow(gsum / cm.gds, cm.w) : 0.0;
  const double g_n = a.g_n[qg];
  const double num = (gs - cm.g0) * g_n + a.dt * cm.h0 * gs * gsum;
  const double den = gs - cm.g0 + a.dt * cm.h0 * gsum;
  warm[9 * NQ + qg] = a.lam_n[qg];
  warm[10 * NQ + qg] = (den != 0.0 && num / den > 0.0) ? num / den : g_n;
}

}  // namespace

void launch_cp_features(const CPArgs &a, const float *gquat, const float *shift, const float *scale, float *X,
                        hipStream_t s) {
  if (a.nq > 0) cp_features<<<unsigned((a.nq + 255) / 256), 256, 0, s>>>(a, gquat, shift, scale, X);
}
void launch_cp_targets(const CPArgs &a, float *T, hipStream_t s) {
  if (a.nq > 0) cp_targets<<<unsigned((a.nq + 255) / 256), 256, 0, s>>>(a, T);
}
void launch_cp_warm_from_slip(const CPArgs &a, const float *Y, int ldy, const float *out_scale, double *warm,
                              hipStream_t s) {
  if (a.nq > 0) cp_warm_from_slip<<<unsigned((a.nq + 255) / 256), 256, 0, s>>>(a, Y, ldy, out_scale, warm);
}

void launch_output(int nne, const OutArgs &o, hipStream_t s) {
  if (o.ne <= 0) return;
  const unsigned grid = unsigned((o.ne + 255) / 256);
  if (nne == 4) output_kernel<4><<<grid, 256, 0, s>>>(o);
  else if (nne == 10) output_kernel<10><<<grid, 256, 0, s>>>(o);
  else if (nne == 8) output_kernel<8><<<grid, 256, 0, s>>>(o);
}

void launch_cp_kinematics(int nne, const CPArgs &a, hipStream_t s) {
  if (nne == 4) launch_kin<4>(a, s);
  else if (nne == 10) launch_kin<10>(a, s);
  else if (nne == 8) launch_kin<8>(a, s);
}

__global__ void cp_queue_reset(unsigned long long *q) { *q = 0; }
// resident waves of the persistent Newton kernel on an MI355X: 256 CUs x 4 SIMDs, one wave each
constexpr int64_t kQueueWaves = 1024;

void launch_cp_newton(const CPArgs &a, int64_t chunk, hipStream_t s) {
  const int64_t nq = a.nq;
  if (nq <= 0) return;
  if (a.qnext) {
    // one persistent wave per SIMD of the chip (the kernel holds all 512 registers of a lane),
    // draining the global queue; the counter is reset by a kernel, so a captured pass replays it
    cp_queue_reset<<<1, 1, 0, s>>>(a.qnext);
    chunk = 64;
  } else if (chunk < 64) {
    chunk = 64;
  }
  const unsigned grid = a.qnext ? unsigned(std::min<int64_t>((nq + 63) / 64, kQueueWaves))
                                : unsigned((nq + chunk - 1) / chunk);
  const unsigned sgrid = unsigned((nq + 63) / 64);
  if (a.law == 1) {
    cp_start_kernel<true><<<sgrid, 64, 0, s>>>(a);
    cp_newton<true><<<grid, 64, 0, s>>>(a, chunk);
  } else {
    cp_start_kernel<false><<<sgrid, 64, 0, s>>>(a);
    cp_newton<false><<<grid, 64, 0, s>>>(a, chunk);
  }
}

void launch_cp_tangent(const CPArgs &a, bool with_D, hipStream_t s) {
  const int64_t nq = a.nq;
  if (nq <= 0) return;
  const unsigned grid = unsigned((nq + 63) / 64);
  if (a.law == 1) {
    if (with_D) cp_tangent<true, true><<<grid, 64, 0, s>>>(a, nq);
    else cp_tangent<true, false><<<grid, 64, 0, s>>>(a, nq);
  } else {
    if (with_D) cp_tangent<false, true><<<grid, 64, 0, s>>>(a, nq);
    else cp_tangent<false, false><<<grid, 64, 0, s>>>(a, nq);
  }
}

void launch_cp_elem(int nne, int mode, const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  if (nne == 4) launch_ce_mode<4>(mode, a, e0, e1, s);
  else if (nne == 10) launch_ce_mode<10>(mode, a, e0, e1, s);
  else if (nne == 8) launch_ce_mode<8>(mode, a, e0, e1, s);
}
void launch_cp_tf_elem(int nne, int mode, const CPElemArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  if (nne == 4) launch_tfe_mode<4>(mode, a, e0, e1, s);
  else if (nne == 10) launch_tfe_mode<10>(mode, a, e0, e1, s);
  else if (nne == 8) launch_tfe_mode<8>(mode, a, e0, e1, s);
}

}  // namespace pgf
