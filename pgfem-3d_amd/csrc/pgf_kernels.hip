// Element hot-path kernels for gfx950 (FP64 throughout).
//
// One launch processes the elements of one (element type, colour) segment.  A 256-thread
// workgroup holds EPB elements; each element owns TPE consecutive lanes:
//   phase 0  lanes a < nne        gather node coordinates, DOF ids and displacements
//   phase 1  lanes q < nqp        shape gradients, F, constitutive update, tau, c  -> LDS
//   phase 2  lanes a*3+b < ndofe  one element row each: residual entry f_ab and the
//                                 tangent row K_ab,wg summed over the QPs, then written out.
// Output, gather mode: the row goes, whole and contiguous, to the element node's block of a
// node-major staging buffer; gather_kernel then sums each node's blocks and writes the node's
// CSR rows once.  Coloured mode: the launch covers one colour, whose elements share no node,
// so the row is added into the CSR values by node-pair blocks with a plain read-modify-write
// (no atomics); the first element (in launch order) to touch a node-pair block stores instead
// of adding, so neither f nor the CSR values need zeroing.  Either way every entry's sum runs
// in the same fixed order on every call.
//
// The contraction is the reference's compute_residual_vector / compute_stiffness_matrix
// (src/constitutive_model.cc:1520-1662) for TL with M = I, written in the current
// configuration: with Fs the stored F(n+1), tau = Fs S Fs^T, h_a = Fs^-T dN_a/dX and
// Gr = Fr Fs^-1 (Gr = I when the integration ran in the same pass):
//   f_ab      = w detJ (Gr tau h_a)_b
//   K_ab,wg   = w detJ [ (Gr R_ab h_w)_g + delta_bg h_a.tau.h_w ],  R_ab = mat(b_ab^T c)
// where c is the spatial (Kirchhoff) tangent in Voigt form and b_ab the engineering-strain
// vector of sym(Gr^T e_b (x) h_a).  This equals B_ab:L:B_wg + delta_bg dN_a.S.dN_w with
// L = Fs^-1 Fs^-1 Fs^-1 Fs^-1 : c.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pgf_device.h"
#include "pgf_math.h"

namespace pgf {
namespace {

using namespace dev;

typedef double dbl2 __attribute__((ext_vector_type(2)));

// Sum over aligned groups of W lanes (butterfly: every lane ends with the same bits, since
// each step adds the same two values in either order).
template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
  return v;
}

// symmetric 3x3 in Voigt (xx,yy,zz,xy,yz,xz) times vector
__device__ __forceinline__ void symv(const double *s, const double *x, double *y) {
  y[0] = s[0] * x[0] + s[3] * x[1] + s[5] * x[2];
  y[1] = s[3] * x[0] + s[1] * x[1] + s[4] * x[2];
  y[2] = s[5] * x[0] + s[4] * x[1] + s[2] * x[2];
}

// What the Voigt Kirchhoff tangent c of a material point is made of (j2_point):
//   c_IK = cv1 oo_IK + cv2 Is_IK + a1 cbar_IK - tm b3 n_I n_K - tm b4 n_I dn2_K,
//   cbar_IK = tm (Is_IK - oo_IK / 3) - 2/3 (s_tr_I 1_K + 1_I s_tr_K)
// (oo: 1 (x) 1 on the normal components, Is: the symmetric identity in Voigt form).
struct J2Tan {
  double cv1, cv2, a1, tm, b3, b4, s_tr[6], n[6], dn2[6];
};

// v[I] for a run-time I (selects, so v stays in registers)
__device__ __forceinline__ double pick6(double v0, double v1, double v2, double v3, double v4, double v5, int I) {
  return I == 0 ? v0 : I == 1 ? v1 : I == 2 ? v2 : I == 3 ? v3 : I == 4 ? v4 : v5;
}

// Continuum damage of the J2 point in the spatial form: tau = om taub and
// c = om cbar - hv (taub (x) tau0 + tau0 (x) taub), om = 1 - w, hv = evo / 2 (the push-forward
// of S = (1 - w) Sbar and L = (1 - w) Lbar - evo/2 (Sbar (x) S0 + S0 (x) Sbar),
// src/cm_j2_plasticity.cc:364-382); a dyad A (x) B is A_I B_K in this Voigt form.
struct J2Dam {
  double om, hv, tb[6], t0[6];
};

// Row I of c (row = stress component, column = engineering strain component).
template <bool DMG = false>
__device__ __forceinline__ void j2_c_row(const J2Tan &T, int I, double w, double *row, const J2Dam *D = nullptr) {
  const double nI = pick6(T.n[0], T.n[1], T.n[2], T.n[3], T.n[4], T.n[5], I);
  const double sI = pick6(T.s_tr[0], T.s_tr[1], T.s_tr[2], T.s_tr[3], T.s_tr[4], T.s_tr[5], I);
  const double nrm_I = I < 3 ? 1.0 : 0.0;
  double tbI = 0.0, t0I = 0.0;
  if constexpr (DMG) {
    tbI = pick6(D->tb[0], D->tb[1], D->tb[2], D->tb[3], D->tb[4], D->tb[5], I);
    t0I = pick6(D->t0[0], D->t0[1], D->t0[2], D->t0[3], D->t0[4], D->t0[5], I);
  }
#pragma unroll
  for (int K = 0; K < 6; ++K) {
    const double oo = (I < 3 && K < 3) ? 1.0 : 0.0;
    const double Is = (I == K) ? (I < 3 ? 1.0 : 0.5) : 0.0;
    const double cbar = T.tm * (Is - oo * (1.0 / 3.0)) - (2.0 / 3.0) * (sI * (K < 3 ? 1.0 : 0.0) + nrm_I * T.s_tr[K]);
    const double cv = T.cv1 * oo + T.cv2 * Is + T.a1 * cbar - T.tm * T.b3 * nI * T.n[K] - T.tm * T.b4 * nI * T.dn2[K];
    if constexpr (DMG)
      row[K] = w * (D->om * cv - D->hv * (tbI * D->t0[K] + t0I * D->tb[K]));
    else
      row[K] = w * cv;
  }
}

// The linear deviatoric law in C of the hyperelastic model (devPotFlag 2: devStress_Linear /
// matStiff_Linear, src/new_potentials.cc:360-373, :434-453, S = G (C - I), L = 2G I) with the
// volumetric potential, in the spatial form of the element kernel:
//   tau = G (b b - b) + kappa J U'(J) 1,  c = G (b_ik b_jl + b_il b_jk) + the volumetric terms
// (b = F F^T; the minor-symmetric form of the push-forward of 2G delta_ik delta_jl, which
// the element rows contract only with symmetric tensors).
struct LinTan {
  double b[6], G, cv1, cv2;
};

__device__ __forceinline__ void he_lin_point(const DevMaterial &m, const double *Fs, double J, double *tau,
                                             LinTan &t) {
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = voigt_i(I), j = voigt_j(I);
    t.b[I] = Fs[i * 3 + 0] * Fs[j * 3 + 0] + Fs[i * 3 + 1] * Fs[j * 3 + 1] + Fs[i * 3 + 2] * Fs[j * 3 + 2];
  }
  double B[9];
  B[0] = t.b[0]; B[4] = t.b[1]; B[8] = t.b[2];
  B[1] = B[3] = t.b[3];
  B[5] = B[7] = t.b[4];
  B[2] = B[6] = t.b[5];
  double dU, d2U;
  vol_pot(m.volPotFlag, J, dU, d2U);
  const double p = m.kappa * J * dU;
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = voigt_i(I), j = voigt_j(I);
    const double bb = B[i * 3 + 0] * B[0 * 3 + j] + B[i * 3 + 1] * B[1 * 3 + j] + B[i * 3 + 2] * B[2 * 3 + j];
    tau[I] = m.G * (bb - t.b[I]) + (I < 3 ? p : 0.0);
  }
  t.G = m.G;
  t.cv1 = m.kappa * (J * J * d2U + J * dU);
  t.cv2 = -2.0 * p;
}

__device__ __forceinline__ void lin_c_row(const LinTan &t, int I, double w, double *row) {
  const int i = voigt_i(I), j = voigt_j(I);
  auto bv = [&](int a, int c) {  // b_ac from the Voigt entries
    const int K = (a == c) ? a : (a + c == 1 ? 3 : (a + c == 3 ? 4 : 5));
    return pick6(t.b[0], t.b[1], t.b[2], t.b[3], t.b[4], t.b[5], K);
  };
#pragma unroll
  for (int K = 0; K < 6; ++K) {
    const int k = voigt_i(K), l = voigt_j(K);
    const double oo = (I < 3 && K < 3) ? 1.0 : 0.0;
    const double Is = (I == K) ? (I < 3 ? 1.0 : 0.5) : 0.0;
    row[K] = w * (t.G * (bv(i, k) * bv(j, l) + bv(i, l) * bv(j, k)) + t.cv1 * oo + t.cv2 * Is);
  }
}

// The damage update of a J2 point from the total F (Fs, J), restating the reference's
// CM_J2P_PARAM::integration_algorithm damage step (src/cm_j2_plasticity.cc:184-212):
// Y0 = G/2 (tr bbar - 3) + kappa U(J), bbar = J^-2/3 Fs Fs^T (j2d_compute_Y0, :162-182), and
// the viscous Weibull update of damage_int_alg (src/volumetric_damage.cc:196-262): with
// dmu = dt mu and g = G(Y0) - X_n, a damaging point (g > 0) takes w = w_n + dmu/(1+dmu) g,
// X = max(X_n, (X_n + dmu G)/(1+dmu)), H = G'(Y0); w is capped at wmax.  Sets D from the
// undamaged tau (taub) and tau0 = G dev(bbar) + kappa J U'(J) 1, the Kirchhoff stress of Y0.
__device__ __forceinline__ void j2_damage(const DevMaterial &m, const double *Fs, double J, double dt, double w_n,
                                          double X_n, const double *taub, double &w, double &X, double &H,
                                          double &damaged, J2Dam &D) {
  const double cj = cbrt(J);
  const double Jm23 = 1.0 / (cj * cj);
  double b[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = voigt_i(I), j = voigt_j(I);
    b[I] = Jm23 * (Fs[i * 3 + 0] * Fs[j * 3 + 0] + Fs[i * 3 + 1] * Fs[j * 3 + 1] + Fs[i * 3 + 2] * Fs[j * 3 + 2]);
  }
  const double tr = b[0] + b[1] + b[2];
  double dU, d2U;
  vol_pot(m.volPotFlag, J, dU, d2U);
  const double Y = 0.5 * m.G * (tr - 3.0) + m.kappa * vol_U(m.volPotFlag, J);
  const double p = m.kappa * J * dU;
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    D.t0[I] = m.G * (b[I] - (I < 3 ? tr * (1.0 / 3.0) : 0.0)) + (I < 3 ? p : 0.0);
    D.tb[I] = taub[I];
  }
  constexpr double kThresh = 0.9999;  // DAMAGE_THRESH, src/volumetric_damage.cc:16
  double G = 0.0, Hy = 0.0;
  if (Y > m.Yin) {  // weibull_function / weibull_evolution, src/volumetric_damage.cc:236-262
    const double z = (Y - m.Yin) / (m.p1 * m.Yin);
    const double e = exp(-pow(z, m.p2));
    G = kThresh - kThresh * e;
    Hy = kThresh * m.p2 / (m.p1 * m.Yin) * e * pow(z, m.p2 - 1.0);
  }
  const double dmu = dt * m.mu;
  const double g = G - X_n;
  const bool dmg = g > 0.0;
  w = dmg ? w_n + dmu / (1.0 + dmu) * g : w_n;
  X = dmg ? fmax(X_n, (X_n + dmu * G) / (1.0 + dmu)) : X_n;
  H = dmg ? Hy : 0.0;
  w = fmin(w, m.wmax >= 1.0 ? kThresh : m.wmax);
  damaged = dmg ? 1.0 : 0.0;
  D.om = 1.0 - w;
  D.hv = dmg ? 0.5 * dmu * H / (1.0 + dmu) : 0.0;
}

// Neo-Hookean / J2 material point in the current configuration.
// In: Fs, its inverse and J; J2 state (Cp^-1 at n, ep_n).  Out: tau (Voigt), the pieces of
// the tangent c (J2Tan, when WITH_C) and, when integ, the n+1 state (outputs by reference, so
// they stay in registers).  HE is the J2 elastic branch with Cp^-1 = I.
// TFV (three-field): the volumetric Kirchhoff stress is J P with the element pressure Ptf in
// place of J kappa U'(J), and its tangent J P (1 x 1 - 2 I) (CM_ThreeField::compute_Ru/Kuu,
// src/constitutive_model_3f.cc:190-286: P tJr Grad_du : tFr^-T and the pJrJn term).
template <bool WITH_C, bool TFV = false>
__device__ __forceinline__ void j2_point(const DevMaterial &m, bool plastic, const double *Fs,
                                         const double *Fi, double J, const double *Cn,
                                         double ep_n, bool integ, double (&Cnp1)[6], double &ep1,
                                         double &gam, double *tau, J2Tan &tg, double Ptf = 0.0) {
  // b_bar trial = J^-2/3 Fs Cn Fs^T (Voigt)
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double f0 = Fs[i * 3 + 0], f1 = Fs[i * 3 + 1], f2 = Fs[i * 3 + 2];
    T[i * 3 + 0] = f0 * Cn[0] + f1 * Cn[3] + f2 * Cn[5];
    T[i * 3 + 1] = f0 * Cn[3] + f1 * Cn[1] + f2 * Cn[4];
    T[i * 3 + 2] = f0 * Cn[5] + f1 * Cn[4] + f2 * Cn[2];
  }
  const double cj = cbrt(J);
  const double Jm23 = 1.0 / (cj * cj);
  double b[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    const int i = voigt_i(I), j = voigt_j(I);
    b[I] = Jm23 * (T[i * 3 + 0] * Fs[j * 3 + 0] + T[i * 3 + 1] * Fs[j * 3 + 1] + T[i * 3 + 2] * Fs[j * 3 + 2]);
  }
  const double G = m.G;
  const double Ibar = (b[0] + b[1] + b[2]) * (1.0 / 3.0);
  double s_tr[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) s_tr[I] = G * b[I] - (I < 3 ? G * Ibar : 0.0);
  const double nrm = sqrt(s_tr[0] * s_tr[0] + s_tr[1] * s_tr[1] + s_tr[2] * s_tr[2] +
                          2.0 * (s_tr[3] * s_tr[3] + s_tr[4] * s_tr[4] + s_tr[5] * s_tr[5]));
  const double mubar = G * Ibar;
  double f_tr = -1.0;
  if (plastic) f_tr = nrm - sqrt(2.0 / 3.0) * (m.k0 + m.hp * ep_n);
  const bool yield = f_tr > 0.0;
  double dg = 0.0, n[6] = {0, 0, 0, 0, 0, 0}, s[6];
  if (yield) {
    dg = f_tr / (2.0 * mubar + 2.0 / 3.0 * m.hp);
    const double rn = 1.0 / nrm;
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      n[I] = s_tr[I] * rn;
      s[I] = s_tr[I] - 2.0 * mubar * dg * n[I];
    }
  } else {
#pragma unroll
    for (int I = 0; I < 6; ++I) s[I] = s_tr[I];
  }
  if (integ) {
    if (yield) {
      // Cp^-1(n+1) = J^2/3 Fi (s/G + Ibar 1) Fi^T
      double be[6];
      const double rG = 1.0 / G;
#pragma unroll
      for (int I = 0; I < 6; ++I) be[I] = s[I] * rG + (I < 3 ? Ibar : 0.0);
      double U[9];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double f0 = Fi[i * 3 + 0], f1 = Fi[i * 3 + 1], f2 = Fi[i * 3 + 2];
        U[i * 3 + 0] = f0 * be[0] + f1 * be[3] + f2 * be[5];
        U[i * 3 + 1] = f0 * be[3] + f1 * be[1] + f2 * be[4];
        U[i * 3 + 2] = f0 * be[5] + f1 * be[4] + f2 * be[2];
      }
      const double Jp23 = cj * cj;
#pragma unroll
      for (int I = 0; I < 6; ++I) {
        const int i = voigt_i(I), j = voigt_j(I);
        Cnp1[I] = Jp23 * (U[i * 3 + 0] * Fi[j * 3 + 0] + U[i * 3 + 1] * Fi[j * 3 + 1] + U[i * 3 + 2] * Fi[j * 3 + 2]);
      }
    } else {
#pragma unroll
      for (int I = 0; I < 6; ++I) Cnp1[I] = Cn[I];
    }
    ep1 = ep_n + sqrt(2.0 / 3.0) * dg;
    gam = dg;
  }
  double dU = 0.0, d2U = 0.0;
  if constexpr (!TFV) vol_pot(m.volPotFlag, J, dU, d2U);
  const double Jp = TFV ? J * Ptf : J * m.kappa * dU;
#pragma unroll
  for (int I = 0; I < 6; ++I) tau[I] = s[I] + (I < 3 ? Jp : 0.0);
  if constexpr (WITH_C) {
    const double cv1 = TFV ? Jp : m.kappa * (J * J * d2U + J * dU);
    const double cv2 = TFV ? -2.0 * Jp : -2.0 * m.kappa * J * dU;
    double beta1 = 0.0, beta3 = 0.0, beta4 = 0.0, dn2[6] = {0, 0, 0, 0, 0, 0};
    if (yield) {
      // 1 / beta0 with beta0 = 1 + hp / (3 mubar)
      const double rb0 = (3.0 * mubar) / (3.0 * mubar + m.hp);
      const double nm = nrm / mubar;
      beta1 = 2.0 * dg / nm;
      const double beta2 = (1.0 - rb0) * (2.0 / 3.0) * nm * dg;
      beta3 = rb0 - beta1 + beta2;
      beta4 = (rb0 - beta1) * nm;
      // n^2 (symmetric), then its deviator
      double n2[6];
      n2[0] = n[0] * n[0] + n[3] * n[3] + n[5] * n[5];
      n2[1] = n[3] * n[3] + n[1] * n[1] + n[4] * n[4];
      n2[2] = n[5] * n[5] + n[4] * n[4] + n[2] * n[2];
      n2[3] = n[0] * n[3] + n[3] * n[1] + n[5] * n[4];
      n2[4] = n[3] * n[5] + n[1] * n[4] + n[4] * n[2];
      n2[5] = n[0] * n[5] + n[3] * n[4] + n[5] * n[2];
      const double t3 = (n2[0] + n2[1] + n2[2]) * (1.0 / 3.0);
#pragma unroll
      for (int I = 0; I < 6; ++I) dn2[I] = n2[I] - (I < 3 ? t3 : 0.0);
    }
    tg.cv1 = cv1;
    tg.cv2 = cv2;
    tg.a1 = 1.0 - beta1;
    tg.tm = 2.0 * mubar;
    tg.b3 = beta3;
    tg.b4 = beta4;
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      tg.s_tr[I] = s_tr[I];
      tg.n[I] = n[I];
      tg.dn2[I] = dn2[I];
    }
  }
}

// Elements per round of the LDS transpose that turns the lanes' element rows into
// contiguous staging-block stores: the largest divisor of epb whose rows (nde x (nde+1)
// doubles each) fit in the `room` doubles of the dead shape-gradient and tangent arrays.
constexpr int stage_round(int epb, int nde, int room) {
  int re = epb;
  while (re > 1 && (epb % re != 0 || re * nde * (nde + 1) > room)) --re;
  return re;
}

// TF: the CM3F three-field element of the hyperelastic model (one volume and one pressure
// variable per element, condensed): F(n+1) is always Fr (the model stores (theta/J)^1/3 Fr),
// tau and c take the element pressure, and the rows get the condensation terms.  Its staged
// variant runs at three waves per SIMD: at four (128 registers) the extra element terms spill.
// VAR: 0 the displacement element, 1 the three-field element (TF), 2 the displacement
// element with J2 continuum damage (DMG: the undamaged kernels keep their register budget),
// 3 the hyperelastic displacement element whose materials may take the linear deviatoric law
// (LIN, chosen per element from devPotFlag).
template <int NNE, int MODE, bool STG, int VAR>
__global__ __launch_bounds__(256, STG ? (VAR != 0 ? 3 : 4) : 1) void elem_kernel(ElemArgs a, int64_t e_begin,
                                                                                int64_t e_end) {
  constexpr bool TF = VAR == 1;
  constexpr bool DMG = VAR == 2;
  constexpr bool LIN = VAR == 3;
  constexpr int NQP = nqp_of(NNE);
  constexpr int NDE = NNE * 3;
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = 256 / TPE;
  constexpr bool INTEG = (MODE & MODE_INTEGRATE) != 0;
  constexpr bool RES = (MODE & MODE_RESIDUAL) != 0;
  constexpr bool TAN = (MODE & MODE_TANGENT) != 0;
  constexpr bool DFL = (MODE & MODE_DEFL_LOAD) != 0;  // K times the prescribed increments
  static_assert(!DFL || (TAN && !RES && !INTEG), "MODE_DEFL_LOAD goes with MODE_TANGENT alone");
  constexpr bool WK = TAN && !DFL;  // K rows are written
  constexpr bool WR = RES || DFL;   // the residual entry is written
  constexpr bool GEN = !INTEG && !TF;  // Fr may differ from the stored F(n+1)
  constexpr int LD = NDE + 1;   // staging row: K row, then the residual entry
  constexpr int kH = EPB * NQP * NDE, kC = EPB * NQP * (TAN ? 36 : 1);
  constexpr int RE = stage_round(EPB, NDE, kH + kC);
  static_assert(!TAN || RE * NDE * LD <= kH + kC, "staging round does not fit");

  __shared__ double s_u[EPB][NDE];
  __shared__ int32_t s_dof[EPB][NDE];
  __shared__ int32_t s_node[EPB][NNE];
  __shared__ double s_x[EPB][NDE];
  // shape gradients and tangents; reused as the staging transpose buffer after phase 2
  __shared__ double s_work[kH + kC];
  auto s_h = reinterpret_cast<double (*)[NQP][NDE]>(s_work);
  auto s_c = reinterpret_cast<double (*)[NQP][TAN ? 36 : 1]>(s_work + kH);
  __shared__ double s_tau[EPB][NQP][6];
  __shared__ double s_g[EPB][NQP][GEN ? 9 : 1];
  __shared__ double s_dl[EPB][DFL ? NDE : 1];  // prescribed increments of the element DOFs
  __shared__ double s_tf[EPB][NQP][TF ? 1 : 1];  // three-field: detJ w J per QP
  __shared__ double s_tfe[EPB][TF ? 6 : 1];      // three-field, per element: Ktt/V^2, the residual
                                                 // coefficient, fp, ft, Ktt, V

  const int slot = threadIdx.x / TPE;
  const int lane = threadIdx.x % TPE;
  // gather mode: all elements of the block; coloured mode: one colour through elist
  const int64_t blk = blockIdx.x;
  const int64_t ie = e_begin + blk * EPB + slot;
  const bool active = ie < e_end;
  const int64_t e = (a.elist && active) ? int64_t(a.elist[ie]) : ie;

  // The QP phase (phase 1 below) runs on one wave: the workgroup's QPs, LPQ lanes per QP,
  // thread p = slot * NQP * LPQ + ip * LPQ + sub of the wave (blk & 3).  Its lanes load their
  // QP's material and state now, so those loads are in flight during phase 0.
  constexpr int P1T = 64;
  constexpr int LPQ = P1T / (EPB * NQP);
  static_assert(LPQ * NQP * EPB == P1T && (LPQ & (LPQ - 1)) == 0 && LPQ <= 64, "lanes per QP");
  const int p1 = int(threadIdx.x) - int(blk & 3) * 64;
  const bool p1live = p1 >= 0 && p1 < P1T && e_begin + blk * EPB + p1 / (NQP * LPQ) < e_end;
  const int p1slot = p1live ? p1 / (NQP * LPQ) : 0, p1ip = p1live ? (p1 / LPQ) % NQP : 0;
  int64_t p1e = 0;
  DevMaterial pm{};
  double pCn[6] = {1, 1, 1, 0, 0, 0}, pep = 0.0, pth = 1.0, ppr = 0.0, pw = 0.0, pX = 0.0;
  if (p1live) {
    const int64_t ie1 = e_begin + blk * EPB + p1slot;
    p1e = a.elist ? int64_t(a.elist[ie1]) : ie1;
    pm = a.mats[a.mat[p1e]];
    if constexpr (TF) {
      pth = a.tf_V[p1e] + a.tf_dV[p1e];
      ppr = a.tf_P[p1e] + a.tf_dP[p1e];
    }
    if (pm.model == PGF_MODEL_J2_PLASTICITY) {
      const int64_t q = a.qp_base + p1e * NQP + p1ip;
#pragma unroll
      for (int I = 0; I < 6; ++I) pCn[I] = a.P_n[I * a.n_qp + q];
      pep = a.Sc_n[q];
      if constexpr (DMG) {
        pw = a.Sc_n[2 * a.n_qp + q];
        pX = a.Sc_n[3 * a.n_qp + q];
      }
    }
  }

  // ---- phase 0: gather, one lane per element DOF; the DOF ids come from the element's own
  // table (not through the node), so the displacements are two loads deep, not three
  if (active && lane < NDE) {
    const int nl = lane / 3, b = lane - 3 * nl;
    const int32_t node = a.conn[e * NNE + nl];
    const int32_t d = a.edof[e * NDE + lane];
    if (b == 0) s_node[slot][nl] = node;
    s_x[slot][lane] = a.X[int64_t(node) * 3 + b];
    s_dof[slot][lane] = d;
    double u = 0.0;  // def_elem_total, src/utils.cc:1391-1412
    double dl = 0.0;
    if (d > 0) {
      u = a.r[d - 1] + (a.d_r ? a.d_r[d - 1] : 0.0);
    } else if (d < 0) {
      const int k = -d - 1;
      dl = a.defl_d ? a.defl_d[k] : 0.0;
      // the prescribed-load pass takes K at defl (defl_d zeroed, src/stiffmat_fd.cc:722-745)
      u = (a.defl ? a.defl[k] : 0.0) + (DFL ? 0.0 : dl);
    }
    s_u[slot][lane] = u;
    if constexpr (DFL) s_dl[slot][lane] = dl;
  }
  lds_barrier();

  // ---- phase 1: quadrature points on one of the four waves (the others skip it, so a
  // workgroup issues the QP stream once, not four times); which wave works rotates with the
  // block, so the QP work spreads over the SIMDs.  Node loops are split over the QP's lanes
  // (node n on sub-lane n % LPQ, which alone reads and writes its entries of s_h) and summed
  // with butterfly shuffles, which leave every lane with the same bits; the return map runs on
  // every lane; row I of c goes to sub-lanes I, I+LPQ, ...
  if (p1live) {
    const int slot = p1slot;  // the QP's element (shadows the row lanes' slot)
    const int64_t e = p1e;
    const int ip = p1ip, sub = p1 % LPQ;
    double k, ee, z, wt;
    int_point(NNE, ip, k, ee, z, wt);
    double *hq = s_h[slot][ip];  // dN/dX first, h_a = Fs^-T dN_a/dX at the end
    {
      double dN[NDE];
      shape_dN<NNE>(k, ee, z, dN);
#pragma unroll
      for (int n = 0; n < NNE; ++n)
        if (n % LPQ == sub) {
          hq[n * 3 + 0] = dN[n * 3 + 0];
          hq[n * 3 + 1] = dN[n * 3 + 1];
          hq[n * 3 + 2] = dN[n * 3 + 2];
        }
    }
    double dX[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dX[i] = 0.0;
    for (int n = sub; n < NNE; n += LPQ) {
      const double d0 = hq[n * 3 + 0], d1 = hq[n * 3 + 1], d2 = hq[n * 3 + 2];
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) {
        const double xb = s_x[slot][n * 3 + ib];
        dX[ib * 3 + 0] += d0 * xb;
        dX[ib * 3 + 1] += d1 * xb;
        dX[ib * 3 + 2] += d2 * xb;
      }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) dX[i] = group_sum<LPQ>(dX[i]);
    const double J0 = det3(dX);
    bool ok = J0 > 0.0;
    double cof[9];
    cof[0] = dX[4] * dX[8] - dX[5] * dX[7];
    cof[1] = dX[2] * dX[7] - dX[1] * dX[8];
    cof[2] = dX[1] * dX[5] - dX[2] * dX[4];
    cof[3] = dX[5] * dX[6] - dX[3] * dX[8];
    cof[4] = dX[0] * dX[8] - dX[2] * dX[6];
    cof[5] = dX[2] * dX[3] - dX[0] * dX[5];
    cof[6] = dX[3] * dX[7] - dX[4] * dX[6];
    cof[7] = dX[1] * dX[6] - dX[0] * dX[7];
    cof[8] = dX[0] * dX[4] - dX[1] * dX[3];
    const double rJ0 = 1.0 / J0;
    double Fr[9];  // sum of u_n (x) dN_n/dX over the lane's nodes, then over the QP's lanes
#pragma unroll
    for (int i = 0; i < 9; ++i) Fr[i] = 0.0;
    for (int n = sub; n < NNE; n += LPQ) {
      const double d0 = hq[n * 3 + 0], d1 = hq[n * 3 + 1], d2 = hq[n * 3 + 2];
      double g[3];
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) g[ib] = rJ0 * (d0 * cof[0 * 3 + ib] + d1 * cof[1 * 3 + ib] + d2 * cof[2 * 3 + ib]);
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) {
        hq[n * 3 + ib] = g[ib];  // dN/dX
        const double uu = s_u[slot][n * 3 + ib];
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) Fr[ib * 3 + ic] += uu * g[ic];
      }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Fr[i] = group_sum<LPQ>(Fr[i]) + ((i % 4 == 0) ? 1.0 : 0.0);
    const int64_t q = a.qp_base + e * NQP + ip;  // e is the block-local element index
    const int64_t NQ = a.n_qp;
    double Fs[9];
    if (INTEG || TF) {
      // three-field: the model keeps (theta/J)^1/3 Fr (run_integration_algorithm's tf_factor,
      // src/constitutive_model.cc:815, :1115); the element works with Fr
      const double fac = TF ? cbrt(pth / det3(Fr)) : 1.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        Fs[i] = Fr[i];
        if (INTEG && i % LPQ == sub) a.F_np1[i * NQ + q] = TF ? fac * Fr[i] : Fr[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 9; ++i) Fs[i] = a.F_np1[i * NQ + q];
    }
    const double J = det3(Fs);
    double Fi[9];
    ok = ok && (J > 0.0) && inv3(Fs, Fi);
    if (!ok) {
      if (sub == 0) atomicAdd(a.err, 1);
#pragma unroll
      for (int i = 0; i < 9; ++i) Fi[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    const DevMaterial m = pm;
    const bool is_j2 = (m.model == PGF_MODEL_J2_PLASTICITY);
    const double ep_n = pep;
    double Cn[6];
#pragma unroll
    for (int I = 0; I < 6; ++I) Cn[I] = pCn[I];
    double Cnp1[6], ep1, gm, tau[6];
    J2Tan T;
    LinTan lt;
    const bool lin = LIN && m.devPotFlag == 2;
    if (lin)
      he_lin_point(m, Fs, ok ? J : 1.0, tau, lt);
    else
      j2_point<TAN, TF>(m, is_j2, Fs, Fi, ok ? J : 1.0, Cn, ep_n, INTEG && is_j2, Cnp1, ep1, gm, tau, T, ppr);
    J2Dam dam;
    double dw = pw, dXs = pX, dH = 0.0, dflag = 0.0;
    if constexpr (DMG) {
      // the damage step of every J2 point with mu > 0, from the state at n (an integrating and
      // a non-integrating pass at one F give the same values)
      if (is_j2 && m.mu > 0.0) {
        j2_damage(m, Fs, ok ? J : 1.0, a.dt, pw, pX, tau, dw, dXs, dH, dflag, dam);
      } else {
        dam.om = 1.0;
        dam.hv = 0.0;
#pragma unroll
        for (int I = 0; I < 6; ++I) dam.tb[I] = dam.t0[I] = 0.0;
      }
#pragma unroll
      for (int I = 0; I < 6; ++I) tau[I] *= dam.om;
    }
    // tau and c go to LDS weighted by the QP's detJ w (zero for a failed QP)
    const double wq = ok ? J0 * wt : 0.0;
    if constexpr (TF) {
      // element sums over the QP lanes of the element (NQP * LPQ consecutive lanes), then the
      // element's condensation coefficients (compute_Rp/Rt/Ktt/Ktp,
      // src/constitutive_model_3f.cc:219-246, :418-466): fp = sum detJ w (J - theta),
      // ft = V (kappa U'(theta) - P), Ktt = V kappa U''(theta), Ktp = Kpt = -V
      const double wJ = wq * (ok ? J : 0.0);
      if (sub == 0) s_tf[slot][ip][0] = wJ;
      const double Jw = group_sum<NQP * LPQ>(sub == 0 ? wJ : 0.0);
      const double V = group_sum<NQP * LPQ>(sub == 0 ? wq : 0.0);
      if (ip == 0 && sub == 0) {
        double dU, d2U;
        vol_pot(m.volPotFlag, pth, dU, d2U);
        const double rV = V > 0.0 ? 1.0 / V : 0.0;  // V = 0 only when every QP failed (counted)
        const double Ktt = V * m.kappa * d2U, ft = V * (m.kappa * dU - ppr), fp = Jw - pth * V;
        s_tfe[slot][0] = Ktt * rV * rV;
        s_tfe[slot][1] = Ktt * fp * rV * rV + ft * rV;
        s_tfe[slot][2] = fp;
        s_tfe[slot][3] = ft;
        s_tfe[slot][4] = Ktt;
        s_tfe[slot][5] = V;
        if (INTEG && RES) {  // the element data of the update (IntegrateThreeFieldUpdate)
          double *tb = a.tf_buf + e * kTfBuf + NDE;
          tb[0] = fp;
          tb[1] = ft;
          tb[2] = Ktt;
          tb[3] = V;
        }
      }
    }
#pragma unroll
    for (int I = 0; I < 6; ++I)
      if (I % LPQ == sub) {
        s_tau[slot][ip][I] = wq * tau[I];
        if (INTEG && is_j2) a.P_np1[I * NQ + q] = Cnp1[I];
      }
    if (sub == 0 && INTEG && is_j2) {
      a.Sc_np1[q] = ep1;
      a.Sc_np1[NQ + q] = gm;
      if constexpr (DMG) {
        a.Sc_np1[2 * NQ + q] = dw;
        a.Sc_np1[3 * NQ + q] = dXs;
        a.Sc_np1[4 * NQ + q] = dH;
        a.Sc_np1[5 * NQ + q] = dflag;
      }
    }
    if constexpr (TAN) {
#pragma unroll
      for (int t = 0; t < (6 + LPQ - 1) / LPQ; ++t) {
        const int I = sub + t * LPQ;
        if (I < 6) {
          if (lin)
            lin_c_row(lt, I, wq, &s_c[slot][ip][I * 6]);
          else
            j2_c_row<DMG>(T, I, wq, &s_c[slot][ip][I * 6], &dam);
        }
      }
    }
    if constexpr (GEN) {  // Gr = Fr Fs^-1
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          if ((i * 3 + j) % LPQ == sub)
            s_g[slot][ip][i * 3 + j] = Fr[i * 3 + 0] * Fi[0 * 3 + j] + Fr[i * 3 + 1] * Fi[1 * 3 + j] + Fr[i * 3 + 2] * Fi[2 * 3 + j];
    }
    for (int n = sub; n < NNE; n += LPQ) {
      const double d0 = hq[n * 3 + 0], d1 = hq[n * 3 + 1], d2 = hq[n * 3 + 2];
#pragma unroll
      for (int i = 0; i < 3; ++i) hq[n * 3 + i] = Fi[0 * 3 + i] * d0 + Fi[1 * 3 + i] * d1 + Fi[2 * 3 + i] * d2;
    }
  }
  lds_barrier();

  // ---- phase 2: element rows (lanes a*3+b < NDE of live elements)
  const bool rowlane = active && lane < NDE;
  const int ra = lane / 3, rb = lane % 3;
  const double eb0 = rb == 0 ? 1.0 : 0.0, eb1 = rb == 1 ? 1.0 : 0.0, eb2 = rb == 2 ? 1.0 : 0.0;
  double K[TAN ? NDE : 1];
  if constexpr (TAN) {
#pragma unroll
    for (int i = 0; i < NDE; ++i) K[i] = 0.0;
  }
  double fe = 0.0;
  if (rowlane) {
    // sym(e_b (x) h_a) in engineering Voigt order has three entries: h_a[b] at (b,b) and
    // h_a[j1], h_a[j2] at the shears (b,j1), (b,j2)
    const int j1 = (rb + 1) % 3, j2 = (rb + 2) % 3;
    const int s1 = 3 + rb, s2 = 3 + j2;
    // three-field: this row's Kup entry, sum detJ w J h_a[b]; the condensation's rank-one
    // term Ktt/V^2 Kup_row Kup_(w,g) is added per QP through the row block M (+ q 1 with
    // q = Ktt/V^2 Kup_row detJ w J), so it costs no extra pass over the columns
    double kup = 0.0;
    if constexpr (TF) {
#pragma unroll
      for (int ip = 0; ip < NQP; ++ip) kup += s_tf[slot][ip][0] * s_h[slot][ip][lane];
      if constexpr (INTEG && RES) a.tf_buf[e * kTfBuf + lane] = kup;  // Kup row of the update data
    }
    const double al = TF ? s_tfe[slot][0] * kup : 0.0;
#pragma unroll
    for (int ip = 0; ip < NQP; ++ip) {
      const double *h = s_h[slot][ip];
      const double *tau = s_tau[slot][ip];  // detJ w tau
      const double ha[3] = {h[ra * 3 + 0], h[ra * 3 + 1], h[ra * 3 + 2]};
      double t[3];
      symv(tau, ha, t);  // detJ w tau h_a
      double g[3] = {eb0, eb1, eb2};
      if constexpr (GEN) {
        g[0] = s_g[slot][ip][rb * 3 + 0];
        g[1] = s_g[slot][ip][rb * 3 + 1];
        g[2] = s_g[slot][ip][rb * 3 + 2];
      }
      if constexpr (RES) fe += g[0] * t[0] + g[1] * t[1] + g[2] * t[2];
      if constexpr (TAN) {
        const double *c = s_c[slot][ip];  // detJ w c
        double r[6];
        if constexpr (GEN) {
          // engineering strain vector of sym(g (x) h_a)
          const double bv[6] = {g[0] * ha[0], g[1] * ha[1], g[2] * ha[2], g[0] * ha[1] + g[1] * ha[0],
                                g[1] * ha[2] + g[2] * ha[1], g[0] * ha[2] + g[2] * ha[0]};
#pragma unroll
          for (int J = 0; J < 6; ++J) {
            double acc = 0.0;
#pragma unroll
            for (int I = 0; I < 6; ++I) acc += bv[I] * c[I * 6 + J];
            r[J] = acc;
          }
        } else {
          const double hb = h[ra * 3 + rb], hj1 = h[ra * 3 + j1], hj2 = h[ra * 3 + j2];
          const double *cb = c + rb * 6, *c1 = c + s1 * 6, *c2 = c + s2 * 6;
#pragma unroll
          for (int J = 0; J < 6; ++J) r[J] = hb * cb[J] + hj1 * c1[J] + hj2 * c2[J];
        }
        // M = Gr R with R = mat(r) (symmetric), plus the geometric term on row b:
        // K_(ab)(wg) = sum_k M_gk h_w,k with M_bk += (detJ w tau h_a)_k
        double M[9];
        if constexpr (GEN) {
          const double *Gm = s_g[slot][ip];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            M[i * 3 + 0] = Gm[i * 3 + 0] * r[0] + Gm[i * 3 + 1] * r[3] + Gm[i * 3 + 2] * r[5];
            M[i * 3 + 1] = Gm[i * 3 + 0] * r[3] + Gm[i * 3 + 1] * r[1] + Gm[i * 3 + 2] * r[4];
            M[i * 3 + 2] = Gm[i * 3 + 0] * r[5] + Gm[i * 3 + 1] * r[4] + Gm[i * 3 + 2] * r[2];
          }
        } else {
          M[0] = r[0]; M[1] = r[3]; M[2] = r[5];
          M[3] = r[3]; M[4] = r[1]; M[5] = r[4];
          M[6] = r[5]; M[7] = r[4]; M[8] = r[2];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          M[0 + k] += eb0 * t[k];
          M[3 + k] += eb1 * t[k];
          M[6 + k] += eb2 * t[k];
        }
        if constexpr (TF) {
          const double q = al * s_tf[slot][ip][0];
          M[0] += q;
          M[4] += q;
          M[8] += q;
        }
#pragma unroll
        for (int w = 0; w < NNE; ++w) {
          const double h0 = h[w * 3 + 0], h1 = h[w * 3 + 1], h2 = h[w * 3 + 2];
          K[w * 3 + 0] += M[0] * h0 + M[1] * h1 + M[2] * h2;
          K[w * 3 + 1] += M[3] * h0 + M[4] * h1 + M[5] * h2;
          K[w * 3 + 2] += M[6] * h0 + M[7] * h1 + M[8] * h2;
        }
      }
    }
    if constexpr (TF) {
      // Condensation of the element's volume and pressure (condense_K_3F_to_1F /
      // condense_F_3F_to_1F, src/condense.cc:304-385, with Pno = Vno = 1): Kpt = Ktp = -V,
      // Kup_(a,b) = sum detJ w J h_a[b] = Kpu, Kut = Ktu = 0 for this model, so
      //   K += Ktt / V^2 Kup Kup^T,   f += Kup (Ktt fp / V^2 + ft / V)
      // with fp = sum detJ w (J - theta), ft = V (kappa U'(theta) - P), Ktt = V kappa U''(theta)
      // (compute_Rp/Rt/Ktt/Ktp, src/constitutive_model_3f.cc:219-246, :418-466).
      if constexpr (RES) fe += kup * s_tfe[slot][1];
    }
    if constexpr (TAN) {  // non-normal diagonal, src/constitutive_model.cc:2682-2687
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
  }
  if constexpr (WK && STG) {
    // Gather mode: RE elements at a time, the rows go to LDS laid out as the staging blocks
    // ([node a][row b][K row, residual]); then the whole workgroup copies each node block
    // (3 (NDE+1) doubles, zero-padded to whole lines) to its staging offset, consecutive
    // threads on consecutive doubles.  Rows of prescribed DOFs are copied too (nothing reads
    // them; it keeps every line whole); nodes without free DOFs have no block.
    constexpr int PB = stage_block(NNE);
    lds_barrier();  // every lane is done with s_h / s_c
    const int64_t e_blk = e_begin + blk * EPB;
#pragma unroll 1
    for (int r0 = 0; r0 < EPB; r0 += RE) {
      if (rowlane && slot >= r0 && slot < r0 + RE) {
        double *row = s_work + ((slot - r0) * NDE + lane) * LD;
#pragma unroll
        for (int i = 0; i < NDE; ++i) row[i] = K[i];
        row[NDE] = fe;
      }
      lds_barrier();
      // two doubles per thread (blocks are whole 128-byte lines, so pairs are 16-byte aligned)
      constexpr int PB2 = PB / 2;
      for (int idx = threadIdx.x; idx < RE * NNE * PB2; idx += 256) {
        const int el = idx / (NNE * PB2), o = idx - el * (NNE * PB2);
        const int na = o / PB2, ob = 2 * (o - na * PB2);
        const int64_t ee = e_blk + r0 + el;
        if (ee >= e_end) continue;
        const int64_t so = a.soff[ee * NNE + na];
        if (so < 0) continue;
        const double *src = s_work + (el * NDE + na * 3) * LD + ob;
        dbl2 v;
        v.x = (ob < 3 * LD) ? src[0] : 0.0;
        v.y = (ob + 1 < 3 * LD) ? src[1] : 0.0;
        // streamed past L2: the gather reads it in a later launch
        __builtin_nontemporal_store(v, reinterpret_cast<dbl2 *>(a.stage + so + ob));
      }
      lds_barrier();
    }
    return;
  }
  if (!rowlane) return;
  const int32_t dof = s_dof[slot][lane];
  if (dof <= 0) return;
  if constexpr (STG) {  // gather mode, residual only: the row's entry into its staging block
    a.stage[a.soff[e * NNE + ra] + rb * LD + NDE] = fe;
    return;
  }
  const int16_t *off = a.off + (e * NNE + ra) * NNE;
  if constexpr (WR) {
    const bool first = off[ra] <= -2;
    if (first) a.f[dof - 1] = fe;
    else a.f[dof - 1] += fe;
  }
  if constexpr (WK) {
    const int32_t node = s_node[slot][ra];
    int bp = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) bp += (j < rb && s_dof[slot][ra * 3 + j] > 0) ? 1 : 0;
    double *row = a.vals + (int64_t(a.row_start[node]) + int64_t(bp) * a.rowlen[node]);
    // Read-modify-write in two halves: every load of a half is issued before its stores
    // (the addresses of one row are distinct), so the loads overlap in flight.
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
          const int p = pos[(w - w0) * 3 + gg];
          if (p >= 0) row[p] = old[(w - w0) * 3 + gg] + K[w * 3 + gg];
        }
    }
  }
}

// Gather assembly, one wavefront per node n with free DOFs: the node's rows (nr free DOFs x
// rowlen columns, contiguous in the CSR values) and residual entries are summed in LDS over
// the node's staging blocks, then stored once.  Each half-wave takes one block at a time
// (blocks k, k+1 side by side); lane c < 3nne of a half owns element column c -- it decodes
// the column's position once and loads the column's three rows -- and lane 3nne the three
// residual entries.  Blocks are loaded NB pairs ahead; the adds then follow in
// block order (the lower half's block before the upper half's, with a wave barrier between),
// and within one block every target is distinct: the sums run in the same order on every
// call.  Elements have at most 10 nodes (3nne + 1 <= 32 lanes).  NB and the nodes per
// workgroup are launch choices (autotuned, pgf_autotune).
template <bool RES, bool TAN, int NB>
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs g) {
  extern __shared__ double s_acc[];
  const int wpb = blockDim.x / 64;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  const int64_t node = int64_t(blockIdx.x) * wpb + wid;
  if (node >= g.n_nodes) return;
  int fm = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b) fm |= (g.dofid[node * 3 + b] > 0) ? (1 << b) : 0;
  if (fm == 0) return;
  const int nr = __popc(fm);
  const int len = g.rowlen[node];
  double *acc = s_acc + int64_t(wid) * (3 * g.maxlen + 3);
  double *facc = acc + 3 * g.maxlen;  // residual entries of the node's rows
  if constexpr (TAN)
    for (int i = lane; i < nr * len; i += 64) acc[i] = 0.0;
  if (lane < 3) facc[lane] = 0.0;
  __builtin_amdgcn_wave_barrier();
  int rowoff[3];  // offset of row b in the node's values (b free)
#pragma unroll
  for (int b = 0; b < 3; ++b) rowoff[b] = __popc(fm & ((1 << b) - 1)) * len;
  const int half = lane >> 5, c = lane & 31;
  const int64_t kb = g.inc_ptr[node], ke = g.inc_ptr[node + 1];
  for (int64_t k0 = kb; k0 < ke; k0 += 2 * NB) {
    double v[NB][3];
    int dst[NB];  // column in the rows, -2 = residual entries, -1 = nothing
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      dst[j] = -1;
      const int64_t k = k0 + 2 * j + half;
      if (k >= ke) continue;
      const uint64_t rec = g.inc_rec[k];
      const int nde = 3 * int(rec >> 60);
      const double *blk = g.stage + int64_t(rec & 0xffffffffu) * 16 + c;
      if (TAN && c < nde) {
        const int w = c / 3, gg = c - 3 * w;
        const uint32_t p = g.pos[int64_t((rec >> 32) & 0x0fffffffu) + w];
        const int mw = int(p >> 28);
        if ((mw >> gg) & 1) dst[j] = int(p & 0x0fffffffu) + __popc(mw & ((1 << gg) - 1));
      } else if (RES && c == nde) {
        dst[j] = -2;
      }
      if (dst[j] != -1) {
#pragma unroll
        for (int b = 0; b < 3; ++b)
          v[j][b] = ((fm >> b) & 1) ? __builtin_nontemporal_load(blk + b * (nde + 1)) : 0.0;
      }
    }
    // the three rows of one block are distinct targets: read the three sums, then write them
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (half == h && dst[j] != -1) {
          double *t0 = (dst[j] >= 0) ? acc + rowoff[0] + dst[j] : facc + 0;
          double *t1 = (dst[j] >= 0) ? acc + rowoff[1] + dst[j] : facc + 1;
          double *t2 = (dst[j] >= 0) ? acc + rowoff[2] + dst[j] : facc + 2;
          const double o0 = (fm & 1) ? *t0 : 0.0, o1 = (fm & 2) ? *t1 : 0.0, o2 = (fm & 4) ? *t2 : 0.0;
          if (fm & 1) *t0 = o0 + v[j][0];
          if (fm & 2) *t1 = o1 + v[j][1];
          if (fm & 4) *t2 = o2 + v[j][2];
        }
        __builtin_amdgcn_wave_barrier();
      }
  }
  if constexpr (TAN) {
    double *out = g.vals + g.row_start[node];
    for (int i = lane; i < nr * len; i += 64) __builtin_nontemporal_store(acc[i], out + i);
  }
  if (RES && lane < 3 && ((fm >> lane) & 1)) g.f[g.dofid[node * 3 + lane] - 1] = facc[lane];
}

__global__ void fill_kernel(double *p, int64_t n, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
// Largest per-QP subdivision measure over [0, n): scale * sum_c |a[c*stride + q]| over ncomp
// components (crystal plasticity: a = gamma_dot), plus scale_d * (d1[q] - d0[q]) when d1 is
// given (a damage increment).  The values are >= 0 (or NaN), so their bit patterns order as
// unsigned integers and one atomicMax per wave merges them; the result does not depend on
// the order.
__global__ __launch_bounds__(256) void rate_max_kernel(const double *a, int ncomp, int64_t stride, int64_t n,
                                                       double scale, const double *d1, const double *d0,
                                                       double scale_d, unsigned long long *out) {
  double m = 0.0;
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    double v = 0.0;
    if (a) {
      double s = 0.0;
      for (int c = 0; c < ncomp; ++c) s += fabs(a[c * stride + q]);
      v = scale * s;
    }
    if (d1) v = fmax(v, scale_d * (d1[q] - d0[q]));
    m = (v > m || v != v) ? v : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(m, o);
    m = (t > m || t != t) ? t : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

__global__ void copy_kernel(double *dst, const double *src, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Element update of the three-field formulation (constitutive_model_update_NR without inertia,
// src/constitutive_model.cc:3916-3999; compute_d_theta_dP, src/constitutive_model_3f.cc:494-549),
// one thread per element, from the element data of the last residual pass and the
// displacement increment du (prescribed DOFs zero: def_elem TYPE 2, src/utils.cc:1365-1389):
//   d_theta = -(Kpu du + fp) / Kpt = (Kup . du + fp) / V,
//   dP = -(ft + Ktu du + Ktt d_theta) / Ktp = (ft + Ktu . du + Ktt d_theta) / V
// (Ktu = 0 for the hyperelastic model, whose buffer keeps zeros there).
template <int NNE>
__global__ __launch_bounds__(256) void tf_update_kernel(TfArgs t) {
  constexpr int NDE = 3 * NNE;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= t.ne) return;
  const double *b = t.buf + e * kTfBuf;
  double s = 0.0, s2 = 0.0;
#pragma unroll
  for (int k = 0; k < NDE; ++k) {
    const int32_t d = t.edof[e * NDE + k];
    if (d > 0) {
      s += b[k] * t.du[d - 1];
      s2 += b[kTfKtu + k] * t.du[d - 1];
    }
  }
  const double V = b[NDE + 3];
  const double rV = V > 0.0 ? 1.0 / V : 0.0;
  const double dth = (s + b[NDE]) * rV;
  t.ddV[e] = dth;
  t.ddP[e] = (b[NDE + 1] + s2 + b[NDE + 2] * dth) * rV;
}

// Initial volume and pressure of each element (compute_3f_initial_conditions,
// src/three_field_element.cc:1328-1386): the volume averages of J and kappa U'(J) at the
// nodal input, one thread per element.
template <int NNE>
__global__ __launch_bounds__(256) void tf_initial_kernel(TfArgs t) {
  constexpr int NQP = nqp_of(NNE), NDE = 3 * NNE;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= t.ne) return;
  double x[NDE], u[NDE];
#pragma unroll
  for (int k = 0; k < NDE; ++k) {
    const int32_t node = t.conn[e * NNE + k / 3];
    x[k] = t.X[int64_t(node) * 3 + k % 3];
    const int32_t d = t.edof[e * NDE + k];
    u[k] = d > 0 ? t.r[d - 1] + (t.d_r ? t.d_r[d - 1] : 0.0)
                 : (d < 0 ? (t.defl ? t.defl[-d - 1] : 0.0) + (t.defl_d ? t.defl_d[-d - 1] : 0.0) : 0.0);
  }
  const DevMaterial m = t.mats[t.mat[e]];
  double vol = 0.0, Jv = 0.0, Up = 0.0;
  for (int ip = 0; ip < NQP; ++ip) {
    double k, ee, z, wt, dN[NDE], dX[9], cof[9], F[9];
    int_point(NNE, ip, k, ee, z, wt);
    shape_dN<NNE>(k, ee, z, dN);
#pragma unroll
    for (int i = 0; i < 9; ++i) dX[i] = 0.0;
#pragma unroll
    for (int n = 0; n < NNE; ++n)
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
#pragma unroll
        for (int j = 0; j < 3; ++j) dX[ib * 3 + j] += dN[n * 3 + j] * x[n * 3 + ib];
    const double J0 = det3(dX);
    cof[0] = dX[4] * dX[8] - dX[5] * dX[7];
    cof[1] = dX[2] * dX[7] - dX[1] * dX[8];
    cof[2] = dX[1] * dX[5] - dX[2] * dX[4];
    cof[3] = dX[5] * dX[6] - dX[3] * dX[8];
    cof[4] = dX[0] * dX[8] - dX[2] * dX[6];
    cof[5] = dX[2] * dX[3] - dX[0] * dX[5];
    cof[6] = dX[3] * dX[7] - dX[4] * dX[6];
    cof[7] = dX[1] * dX[6] - dX[0] * dX[7];
    cof[8] = dX[0] * dX[4] - dX[1] * dX[3];
    const double rJ0 = 1.0 / J0;
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = (i % 4 == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int n = 0; n < NNE; ++n) {
      double g[3];
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
        g[ib] = rJ0 * (dN[n * 3 + 0] * cof[0 * 3 + ib] + dN[n * 3 + 1] * cof[1 * 3 + ib] + dN[n * 3 + 2] * cof[2 * 3 + ib]);
#pragma unroll
      for (int ib = 0; ib < 3; ++ib)
#pragma unroll
        for (int ic = 0; ic < 3; ++ic) F[ib * 3 + ic] += u[n * 3 + ib] * g[ic];
    }
    double dU, d2U;
    vol_pot(m.volPotFlag, det3(F), dU, d2U);
    const double w = J0 * wt;
    vol += w;
    Jv += w * det3(F);
    Up += w * m.kappa * dU;
  }
  t.V[e] = Jv / vol;
  t.P[e] = Up / vol;
}

__global__ void axpy_kernel(double *y, const double *x, double g, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] += g * x[i];
}

template <int NNE, bool STG, int TF>
void launch_nne(int mode, const ElemArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s) {
  constexpr int TPE = (NNE == 4) ? 16 : 32;
  constexpr int EPB = 256 / TPE;
  const int64_t n = e_end - e_begin;
  if (n <= 0) return;
  const unsigned grid = unsigned((n + EPB - 1) / EPB);
  switch (mode) {
    case MODE_INTEGRATE | MODE_RESIDUAL:
      elem_kernel<NNE, MODE_INTEGRATE | MODE_RESIDUAL, STG, TF><<<grid, 256, 0, s>>>(a, e_begin, e_end);
      break;
    case MODE_RESIDUAL:
      elem_kernel<NNE, MODE_RESIDUAL, STG, TF><<<grid, 256, 0, s>>>(a, e_begin, e_end);
      break;
    case MODE_TANGENT:
      elem_kernel<NNE, MODE_TANGENT, STG, TF><<<grid, 256, 0, s>>>(a, e_begin, e_end);
      break;
    case MODE_INTEGRATE | MODE_RESIDUAL | MODE_TANGENT:
      elem_kernel<NNE, MODE_INTEGRATE | MODE_RESIDUAL | MODE_TANGENT, STG, TF><<<grid, 256, 0, s>>>(a, e_begin, e_end);
      break;
    case MODE_TANGENT | MODE_DEFL_LOAD:
      elem_kernel<NNE, MODE_TANGENT | MODE_DEFL_LOAD, STG, TF><<<grid, 256, 0, s>>>(a, e_begin, e_end);
      break;
    default:
      break;
  }
}

}  // namespace

void launch_elem(int nne, int mode, const ElemArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s,
                 int variant) {
  const bool stg = a.stage != nullptr;
  if (a.tf_V) {
    if (nne == 4) (stg ? launch_nne<4, true, 1> : launch_nne<4, false, 1>)(mode, a, e_begin, e_end, s);
    else if (nne == 10) (stg ? launch_nne<10, true, 1> : launch_nne<10, false, 1>)(mode, a, e_begin, e_end, s);
    else if (nne == 8) (stg ? launch_nne<8, true, 1> : launch_nne<8, false, 1>)(mode, a, e_begin, e_end, s);
    return;
  }
  if (variant == 2) {
    if (nne == 4) (stg ? launch_nne<4, true, 2> : launch_nne<4, false, 2>)(mode, a, e_begin, e_end, s);
    else if (nne == 10) (stg ? launch_nne<10, true, 2> : launch_nne<10, false, 2>)(mode, a, e_begin, e_end, s);
    else if (nne == 8) (stg ? launch_nne<8, true, 2> : launch_nne<8, false, 2>)(mode, a, e_begin, e_end, s);
    return;
  }
  if (variant == 3) {
    if (nne == 4) (stg ? launch_nne<4, true, 3> : launch_nne<4, false, 3>)(mode, a, e_begin, e_end, s);
    else if (nne == 10) (stg ? launch_nne<10, true, 3> : launch_nne<10, false, 3>)(mode, a, e_begin, e_end, s);
    else if (nne == 8) (stg ? launch_nne<8, true, 3> : launch_nne<8, false, 3>)(mode, a, e_begin, e_end, s);
    return;
  }
  if (nne == 4) (stg ? launch_nne<4, true, 0> : launch_nne<4, false, 0>)(mode, a, e_begin, e_end, s);
  else if (nne == 10) (stg ? launch_nne<10, true, 0> : launch_nne<10, false, 0>)(mode, a, e_begin, e_end, s);
  else if (nne == 8) (stg ? launch_nne<8, true, 0> : launch_nne<8, false, 0>)(mode, a, e_begin, e_end, s);
}

int gather_waves(int32_t maxlen) {
  const int64_t per_wave = (int64_t(3) * std::max<int32_t>(maxlen, 1) + 3) * int64_t(sizeof(double));
  const int64_t w = (int64_t(64) << 10) / per_wave;  // 64 KiB of LDS per workgroup
  return int(std::min<int64_t>(4, w));
}

template <int NB>
void launch_gather_nb(bool res, bool tan, const GatherArgs &g, unsigned grid, int threads, size_t lds,
                      hipStream_t s) {
  if (res && tan) gather_kernel<true, true, NB><<<grid, threads, lds, s>>>(g);
  else if (res) gather_kernel<true, false, NB><<<grid, threads, lds, s>>>(g);
  else if (tan) gather_kernel<false, true, NB><<<grid, threads, lds, s>>>(g);
}

void launch_gather(int mode, const GatherArgs &g, int waves, int batch, hipStream_t s) {
  const int wpb = std::min(gather_waves(g.maxlen), std::max(waves, 1));
  if (g.n_nodes <= 0 || wpb < 1) return;
  const unsigned grid = unsigned((g.n_nodes + wpb - 1) / wpb);
  const size_t lds = size_t(wpb) * (3 * size_t(std::max<int32_t>(g.maxlen, 1)) + 3) * sizeof(double);
  // the prescribed-load pass stages its vector in the residual entries and no K rows
  const bool dfl = (mode & MODE_DEFL_LOAD) != 0;
  const bool res = (mode & MODE_RESIDUAL) != 0 || dfl, tan = (mode & MODE_TANGENT) != 0 && !dfl;
  if (batch <= 2) launch_gather_nb<2>(res, tan, g, grid, 64 * wpb, lds, s);
  else if (batch <= 4) launch_gather_nb<4>(res, tan, g, grid, 64 * wpb, lds, s);
  else launch_gather_nb<8>(res, tan, g, grid, 64 * wpb, lds, s);
}

void launch_tf_update(int nne, const TfArgs &t, hipStream_t s) {
  if (t.ne <= 0) return;
  const unsigned grid = unsigned((t.ne + 255) / 256);
  if (nne == 4) tf_update_kernel<4><<<grid, 256, 0, s>>>(t);
  else if (nne == 10) tf_update_kernel<10><<<grid, 256, 0, s>>>(t);
  else if (nne == 8) tf_update_kernel<8><<<grid, 256, 0, s>>>(t);
}

void launch_tf_initial(int nne, const TfArgs &t, hipStream_t s) {
  if (t.ne <= 0) return;
  const unsigned grid = unsigned((t.ne + 255) / 256);
  if (nne == 4) tf_initial_kernel<4><<<grid, 256, 0, s>>>(t);
  else if (nne == 10) tf_initial_kernel<10><<<grid, 256, 0, s>>>(t);
  else if (nne == 8) tf_initial_kernel<8><<<grid, 256, 0, s>>>(t);
}

void launch_axpy(double *y, const double *x, double g, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  axpy_kernel<<<unsigned(blocks), 256, 0, s>>>(y, x, g, n);
}

void launch_fill(double *p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  fill_kernel<<<unsigned(blocks), 256, 0, s>>>(p, n, v);
}

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  fill_i32_kernel<<<unsigned(blocks), 256, 0, s>>>(p, n, v);
}

void launch_rate_max(const double *a, int ncomp, int64_t stride, int64_t n, double scale, const double *d1,
                     const double *d0, double scale_d, unsigned long long *out, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
  rate_max_kernel<<<unsigned(blocks), 256, 0, s>>>(a, ncomp, stride, n, scale, d1, d0, scale_d, out);
}

void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  copy_kernel<<<unsigned(blocks), 256, 0, s>>>(dst, src, n);
}

}  // namespace pgf
