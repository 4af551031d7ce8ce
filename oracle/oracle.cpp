// CPU oracle of the PGFem3D finite-strain element hot path -- TEST INFRASTRUCTURE ONLY.
//
// Plain FP64 restatement of the reference (C-SWARM/pgfem-3d) for one MPI rank, total
// Lagrangian CM path, no thermal coupling.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, as the checker; the product never does.
//
// Pinning (tests that would fail on a plausible mistake here):
//   orc_generate_local_dof_ids, orc_psparse_apai
//       brute force on hand-enumerated one-cell meshes (every DOF of a cell couples to every
//       other), pattern symmetry, diagonal present          -> tests/test_oracle_pattern.py
//   orc_basis   partition of unity, sum dN = 0, element volume = sum detJxW, exact
//               integration of linear fields, affine u gives F exactly -> test_oracle_kinematics.py
//   orc_he_update  S(I) = 0, closed-form uniaxial neo-Hookean S, L = central FD of S wrt
//               E (1e-7), frame invariance S(QF) = S(F)       -> test_oracle_materials.py
//   orc_j2_update  elastic range == orc_he_update, consistency ||s|| = sqrt(2/3)(k0+hp ep)
//               after return, small-strain uniaxial closed form (E_t = E hp/(E+hp)),
//               L = central FD of the return map (1e-6)        -> test_oracle_materials.py
//   orc_assemble   K = central FD of f (1e-6); rigid rotation gives f = 0; patch test
//               (affine u: f = 0 at interior DOFs)             -> test_oracle_element.py
// The J2 return map itself lives in the un-vendored GCM library in the reference
// (src/cm_j2_plasticity.cc:197-205); it is restated here from Simo & Hughes,
// Computational Inelasticity (1998), boxes 9.1-9.2, and pinned only by the checks above.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

inline int idx4(int i, int j, int k, int l) { return ((i * 3 + j) * 3 + k) * 3 + l; }

// det3x3, src/utils.cc:478-483
double det3x3(const double *m) {
  return ((m[0] * (m[4] * m[8] - m[5] * m[7])) + m[1] * (m[5] * m[6] - m[3] * m[8]) +
          m[2] * (m[3] * m[7] - m[4] * m[6]));
}

// inv3x3, src/utils.cc:542-562 (singular if |det| < 1e-15)
int inv3x3(const double *m, double *mi) {
  const double A = det3x3(m);
  if (!(std::fabs(A) >= 1.0e-15)) return 1;
  mi[0] = 1 / A * (m[4] * m[8] - m[5] * m[7]);
  mi[1] = -1 / A * (m[1] * m[8] - m[2] * m[7]);
  mi[2] = 1 / A * (m[1] * m[5] - m[2] * m[4]);
  mi[3] = -1 / A * (m[3] * m[8] - m[5] * m[6]);
  mi[4] = 1 / A * (m[0] * m[8] - m[2] * m[6]);
  mi[5] = -1 / A * (m[0] * m[5] - m[2] * m[3]);
  mi[6] = 1 / A * (m[3] * m[7] - m[4] * m[6]);
  mi[7] = -1 / A * (m[0] * m[7] - m[1] * m[6]);
  mi[8] = 1 / A * (m[0] * m[4] - m[1] * m[3]);
  return 0;
}

void matmul(const double *A, const double *B, double *C) {  // C = A B
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
      C[i * 3 + j] = s;
    }
}
void matmul_tn(const double *A, const double *B, double *C) {  // C = A^T B
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[k * 3 + i] * B[k * 3 + j];
      C[i * 3 + j] = s;
    }
}
void matmul_nt(const double *A, const double *B, double *C) {  // C = A B^T
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[j * 3 + k];
      C[i * 3 + j] = s;
    }
}

// ------------------------------------------------------------------ FEMLIB
// Quadrature (FEMLIB/src/femlib.cc:544-588): truncated constants kept verbatim.
const double kOneOverSqrt3 = 0.57735026918962584;
const double kTet4A = 0.58541020, kTet4B = 0.13819660;

int nint_of(int nne) {
  switch (nne) {
    case 4: return 1;   // TetIntegrationPoints1 (order 0)
    case 10: return 4;  // order upgraded to 1 -> TetIntegrationPoints4 (femlib.cc:1010-1016)
    case 8: return 8;   // gpno = order+1 = 2 -> 2x2x2 Gauss (femlib.cc:1067-1076)
  }
  return 0;
}

// integration point ip -> (ksi, eta, zet, weight); itg_ids as in femlib.cc:624-700
void int_point(int nne, int ip, double *k, double *e, double *z, double *w) {
  if (nne == 4) {
    *k = *e = *z = 1.0 / 4.0;
    *w = 1.0 / 6.0;
  } else if (nne == 10) {
    const double gk[4] = {kTet4A, kTet4B, kTet4B, kTet4B};
    const double ge[4] = {kTet4B, kTet4A, kTet4B, kTet4B};
    const double gz[4] = {kTet4B, kTet4B, kTet4A, kTet4B};
    *k = gk[ip];
    *e = ge[ip];
    *z = gz[ip];
    *w = (1.0 / 4.0) / 6.0;
  } else {  // hex: itg_ids(cnt) = (ia, ib, ic) in a triple loop, weights 1*1*1
    const double g[2] = {-kOneOverSqrt3, kOneOverSqrt3};
    *k = g[ip / 4];
    *e = g[(ip / 2) % 2];
    *z = g[ip % 2];
    *w = 1.0 * 1.0 * 1.0;
  }
}

// PGFem3D_N, femlib.cc:124-203
void shape_N(int nne, double ksi, double eta, double zet, double *N) {
  if (nne == 4) {
    N[0] = (1.0 - ksi - eta - zet);
    N[1] = ksi;
    N[2] = eta;
    N[3] = zet;
  } else if (nne == 10) {
    const double l = 1.0 - ksi - eta - zet;
    N[0] = (2.0 * l - 1.0) * l;
    N[1] = (2.0 * ksi - 1.0) * ksi;
    N[2] = (2.0 * eta - 1.0) * eta;
    N[3] = (2.0 * zet - 1.0) * zet;
    N[4] = 4.0 * l * ksi;
    N[5] = 4.0 * ksi * eta;
    N[6] = 4.0 * l * eta;
    N[7] = 4.0 * l * zet;
    N[8] = 4.0 * ksi * zet;
    N[9] = 4.0 * eta * zet;
  } else {
    const double o = 0.125;
    N[0] = o * (1.0 - ksi) * (1.0 - eta) * (1.0 - zet);
    N[1] = o * (1.0 + ksi) * (1.0 - eta) * (1.0 - zet);
    N[2] = o * (1.0 + ksi) * (1.0 + eta) * (1.0 - zet);
    N[3] = o * (1.0 - ksi) * (1.0 + eta) * (1.0 - zet);
    N[4] = o * (1.0 - ksi) * (1.0 - eta) * (1.0 + zet);
    N[5] = o * (1.0 + ksi) * (1.0 - eta) * (1.0 + zet);
    N[6] = o * (1.0 + ksi) * (1.0 + eta) * (1.0 + zet);
    N[7] = o * (1.0 - ksi) * (1.0 + eta) * (1.0 + zet);
  }
}

// compute_dN_dkez, femlib.cc:214-340; dN[a*3+c]
void shape_dN(int nne, double ksi, double eta, double zet, double *dN) {
  if (nne == 4) {
    const double t[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 12; ++i) dN[i] = t[i];
  } else if (nne == 10) {
    dN[0] = 4.0 * (ksi + eta + zet) - 3.0;
    dN[1] = 4.0 * (eta + ksi + zet) - 3.0;
    dN[2] = 4.0 * (zet + ksi + eta) - 3.0;
    dN[3] = 4.0 * ksi - 1.0; dN[4] = 0.0; dN[5] = 0.0;
    dN[6] = 0.0; dN[7] = 4.0 * eta - 1.0; dN[8] = 0.0;
    dN[9] = 0.0; dN[10] = 0.0; dN[11] = 4.0 * zet - 1.0;
    dN[12] = 4.0 * (-2.0 * ksi - eta - zet + 1.0); dN[13] = -4.0 * ksi; dN[14] = -4.0 * ksi;
    dN[15] = 4.0 * eta; dN[16] = 4.0 * ksi; dN[17] = 0.0;
    dN[18] = -4.0 * eta; dN[19] = 4.0 * (-ksi - 2.0 * eta - zet + 1.0); dN[20] = -4.0 * eta;
    dN[21] = -4.0 * zet; dN[22] = -4.0 * zet; dN[23] = 4.0 * (-ksi - eta - 2.0 * zet + 1.0);
    dN[24] = 4.0 * zet; dN[25] = 0.0; dN[26] = 4.0 * ksi;
    dN[27] = 0.0; dN[28] = 4.0 * zet; dN[29] = 4.0 * eta;
  } else {
    const double o = 0.125;
    const double s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int a = 0; a < 8; ++a) {
      dN[a * 3 + 0] = s[a][0] * o * (1.0 + s[a][1] * eta) * (1.0 + s[a][2] * zet);
      dN[a * 3 + 1] = s[a][1] * o * (1.0 + s[a][0] * ksi) * (1.0 + s[a][2] * zet);
      dN[a * 3 + 2] = s[a][2] * o * (1.0 + s[a][0] * ksi) * (1.0 + s[a][1] * eta);
    }
  }
}

// PGFem3D_dNdX, femlib.cc:403-467 (+ compute_dX_dkez :350-364 and det :373-392)
double dNdX_of(int nne, const double *dNk, const double *X, double *dNdX) {
  double dX[9] = {0};
  for (int ib = 0; ib < 3; ++ib)
    for (int ic = 0; ic < 3; ++ic)
      for (int ia = 0; ia < nne; ++ia) dX[ib * 3 + ic] += dNk[ia * 3 + ic] * X[ia * 3 + ib];
  const double J = dX[0] * (dX[4] * dX[8] - dX[5] * dX[7]) + dX[1] * (dX[5] * dX[6] - dX[3] * dX[8]) +
                   dX[2] * (dX[3] * dX[7] - dX[4] * dX[6]);
  if (J <= 0.0) return J;
  double I[9];
  I[0] = dX[4] * dX[8] - dX[5] * dX[7];
  I[1] = dX[2] * dX[7] - dX[1] * dX[8];
  I[2] = dX[1] * dX[5] - dX[2] * dX[4];
  I[3] = dX[5] * dX[6] - dX[3] * dX[8];
  I[4] = dX[0] * dX[8] - dX[2] * dX[6];
  I[5] = dX[2] * dX[3] - dX[0] * dX[5];
  I[6] = dX[3] * dX[7] - dX[4] * dX[6];
  I[7] = dX[1] * dX[6] - dX[0] * dX[7];
  I[8] = dX[0] * dX[4] - dX[1] * dX[3];
  for (int ia = 0; ia < nne; ++ia)
    for (int ib = 0; ib < 3; ++ib) {
      double s = 0.0;
      for (int ic = 0; ic < 3; ++ic) s += 1.0 / J * dNk[ia * 3 + ic] * I[ic * 3 + ib];
      dNdX[ia * 3 + ib] = s;
    }
  return J;
}

// ------------------------------------------------------------------ hyperelasticity
// kappa on the CM path: E/(3(1-2nu)) (src/constitutive_model.cc:1006)
double kappa_of(const orc_material &m) { return m.E / (3.0 * (1.0 - 2.0 * m.nu)); }

// dU/dJ and d2U/dJ2 (src/new_potentials.cc:575-616), flag selection :184-210
void vol_pot(int flag, double J, double *dU, double *d2U, double *U) {
  switch (flag) {
    case 3:  // Doll-Schweizerhof 8
      *U = 0.5 * (J - 1.0) * std::log(J);
      *dU = (-1 + J) / (2. * J) + std::log(J) * .5;
      *d2U = 1. / J - (-1 + J) / (2. * J * J);
      break;
    case 99:  // quadratic
      *U = 0.5 * (J - 1.0) * (J - 1.0);
      *dU = J - 1;
      *d2U = 1.0;
      break;
    default:  // 2: Doll-Schweizerhof 7
      *U = 0.5 * (std::exp(J - 1.0) - std::log(J) - 1.0);
      *dU = (std::exp(-1 + J) - 1. / J) * .5;
      *d2U = (std::exp(-1 + J) + 1. / (J * J)) * .5;
  }
}

// devStress_Mooney_Rivlin, src/new_potentials.cc:329-358
void dev_stress_mr(const double *C, double m10, double m01, double *S) {
  double invC[9];
  inv3x3(C, invC);
  const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double detC = det3x3(C);
  const double trC = C[0] + C[4] + C[8];
  double CC = 0.0;
  for (int i = 0; i < 9; ++i) CC += C[i] * C[i];
  for (int i = 0; i < 9; ++i)
    S[i] = ((2 * m10 * std::pow(detC, -1. / 3.) * (id[i] - 1. / 3. * trC * invC[i])) +
            (2 * m01 * std::pow(detC, -2. / 3.) *
             (trC * id[i] - C[i] + 1. / 3. * (CC - trC * trC) * invC[i])));
}

// matStiff_Mooney_Rivlin, src/new_potentials.cc:376-432
void mat_stiff_mr(const double *C, double m10, double m01, double *L) {
  double invC[9];
  inv3x3(C, invC);
  const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double d1 = std::pow(det3x3(C), -1. / 3.);
  const double d2 = d1 * d1;
  const double trC = C[0] + C[4] + C[8];
  double CC = 0.0;
  for (int i = 0; i < 9; ++i) CC += C[i] * C[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const int ij = i * 3 + j, kl = k * 3 + l, ik = i * 3 + k, lj = l * 3 + j;
          L[idx4(i, j, k, l)] =
              (4 * m10 * d1 *
               ((1. / 9. * trC * invC[ij] - 1. / 3. * id[ij]) * invC[kl] +
                1. / 3. * (trC * invC[ik] * invC[lj] - id[kl] * invC[ij]))) +
              (4 * m01 * d2 *
               (id[ij] * id[kl] - id[ik] * id[lj] + 2. / 3. * ((C[ij] - trC * id[ij]) * invC[kl]) -
                ((CC - trC * trC) * (2. / 9. * invC[ij] * invC[kl] + 1. / 3. * invC[ik] * invC[lj])) +
                2. / 3. * ((C[kl] - trC * id[kl]) * invC[ij])));
        }
}

// devStress_Linear / matStiff_Linear, src/new_potentials.cc:360-373, :434-453:
// S = G (C - I), L = 2 G delta_ik delta_jl
void dev_stress_linear(const double *C, double G, double *S) {
  for (int i = 0; i < 9; ++i) S[i] = G * (C[i] - ((i % 4 == 0) ? 1.0 : 0.0));
}
void mat_stiff_linear(double G, double *L) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) L[idx4(i, j, k, l)] = (i == k && j == l) ? 2.0 * G : 0.0;
}

// Total S and L of the hyperelastic model (HE_PARAM::update_elasticity,
// src/plasticity_model_none.cc:329-362, through the in-tree mirror get_material_stress /
// get_material_stiffness, src/displacement_based_element.cc:1368-1445, damage off).  The
// deviatoric law follows devPotFlag (getDevStressFunc, src/new_potentials.cc:104-129):
// 1 Mooney-Rivlin (m10, m01), 2 the linear law in C (G).
int he_update(const orc_material &m, const double *F, double *S, double *L) {
  double C[9], CI[9];
  matmul_tn(F, F, C);
  if (inv3x3(C, CI)) return 1;
  const double J = det3x3(F);
  if (!(J > 0.0)) return 1;
  const double kappa = kappa_of(m);
  double dU, d2U, U;
  vol_pot(m.volPotFlag, J, &dU, &d2U, &U);
  const bool lin = m.devPotFlag == 2;
  if (lin) dev_stress_linear(C, m.G, S);
  else dev_stress_mr(C, m.m10, m.m01, S);
  for (int i = 0; i < 9; ++i) S[i] += kappa * J * dU * CI[i];
  if (L) {
    if (lin) mat_stiff_linear(m.G, L);
    else mat_stiff_mr(C, m.m10, m.m01, L);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            const int q = idx4(i, j, k, l);
            L[q] = L[q] + kappa * (J * dU + J * J * d2U) * CI[i * 3 + j] * CI[k * 3 + l] -
                   2.0 * kappa * J * dU * CI[i * 3 + k] * CI[l * 3 + j];
          }
  }
  return 0;
}

// ------------------------------------------------------------------ J2 plasticity
// Simo & Hughes (1998) box 9.1 (return map) and box 9.2 (algorithmic tangent), total
// Lagrangian: internal variables Cp^-1 (reference configuration) and ep; S = F^-1 tau F^-T,
// L = F^-1 F^-1 F^-1 F^-1 : c.  The tangent's n (x) dev(n^2) term is kept unsymmetrised,
// which makes L the exact derivative of the return map.
void sym6_to_9(const double *a, double *A) {
  A[0] = a[0]; A[4] = a[1]; A[8] = a[2];
  A[1] = A[3] = a[3];
  A[5] = A[7] = a[4];
  A[2] = A[6] = a[5];
}
void sym9_to_6(const double *A, double *a) {
  a[0] = A[0]; a[1] = A[4]; a[2] = A[8];
  a[3] = 0.5 * (A[1] + A[3]);
  a[4] = 0.5 * (A[5] + A[7]);
  a[5] = 0.5 * (A[2] + A[6]);
}

int j2_update(const orc_material &m, const double *F, const double *Cpinv_n, double ep_n,
              double *Cpinv_np1, double *ep_np1, double *gam, double *S, double *L) {
  const double J = det3x3(F);
  if (!(J > 0.0)) return 1;
  double Fi[9];
  if (inv3x3(F, Fi)) return 1;
  const double G = m.G, H = m.hp, kappa = kappa_of(m);
  double Cn[9], T[9], b[9];
  sym6_to_9(Cpinv_n, Cn);
  matmul(F, Cn, T);
  matmul_nt(T, F, b);
  const double Jm23 = std::pow(J, -2.0 / 3.0);
  for (int i = 0; i < 9; ++i) b[i] *= Jm23;  // b_bar trial
  const double Ibar = (b[0] + b[4] + b[8]) / 3.0;
  double s_tr[9];
  for (int i = 0; i < 9; ++i) s_tr[i] = G * b[i];
  s_tr[0] -= G * Ibar;
  s_tr[4] -= G * Ibar;
  s_tr[8] -= G * Ibar;
  double nrm = 0.0;
  for (int i = 0; i < 9; ++i) nrm += s_tr[i] * s_tr[i];
  nrm = std::sqrt(nrm);
  const double f_tr = nrm - std::sqrt(2.0 / 3.0) * (m.k0 + H * ep_n);
  const double mubar = G * Ibar;
  double dg = 0.0, s[9], n[9] = {0};
  std::memcpy(s, s_tr, sizeof(s));
  if (f_tr > 0.0) {
    dg = f_tr / (2.0 * mubar + 2.0 / 3.0 * H);
    for (int i = 0; i < 9; ++i) {
      n[i] = s_tr[i] / nrm;
      s[i] = s_tr[i] - 2.0 * mubar * dg * n[i];
    }
  }
  if (Cpinv_np1) {
    if (f_tr > 0.0) {
      double be[9];  // b_bar_e(n+1) = s/G + Ibar 1
      for (int i = 0; i < 9; ++i) be[i] = s[i] / G;
      be[0] += Ibar; be[4] += Ibar; be[8] += Ibar;
      double T2[9], C2[9];
      matmul(Fi, be, T2);
      matmul_nt(T2, Fi, C2);
      const double Jp23 = std::pow(J, 2.0 / 3.0);
      for (int i = 0; i < 9; ++i) C2[i] *= Jp23;
      sym9_to_6(C2, Cpinv_np1);
    } else {
      for (int i = 0; i < 6; ++i) Cpinv_np1[i] = Cpinv_n[i];
    }
  }
  if (ep_np1) *ep_np1 = ep_n + std::sqrt(2.0 / 3.0) * dg;
  if (gam) *gam = dg;
  double dU, d2U, U;
  vol_pot(m.volPotFlag, J, &dU, &d2U, &U);
  double tau[9];
  for (int i = 0; i < 9; ++i) tau[i] = s[i];
  const double Jp = J * kappa * dU;
  tau[0] += Jp; tau[4] += Jp; tau[8] += Jp;
  if (S) {
    double T3[9];
    matmul(Fi, tau, T3);
    matmul_nt(T3, Fi, S);
  }
  if (L) {
    const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double cv1 = kappa * (J * J * d2U + J * dU), cv2 = -2.0 * kappa * J * dU;
    double beta1 = 0, beta3 = 0, beta4 = 0, dn2[9] = {0};
    if (f_tr > 0.0) {
      const double beta0 = 1.0 + H / (3.0 * mubar);
      beta1 = 2.0 * mubar * dg / nrm;
      const double beta2 = (1.0 - 1.0 / beta0) * (2.0 / 3.0) * (nrm / mubar) * dg;
      beta3 = 1.0 / beta0 - beta1 + beta2;
      beta4 = (1.0 / beta0 - beta1) * nrm / mubar;
      double n2[9];
      matmul(n, n, n2);
      const double t3 = (n2[0] + n2[4] + n2[8]) / 3.0;
      for (int i = 0; i < 9; ++i) dn2[i] = n2[i] - t3 * id[i];
    }
    double c[81];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            const double Isym = 0.5 * (id[i * 3 + k] * id[j * 3 + l] + id[i * 3 + l] * id[j * 3 + k]);
            const double oo = id[i * 3 + j] * id[k * 3 + l];
            const double cbar = 2.0 * mubar * (Isym - oo / 3.0) -
                                2.0 / 3.0 * (s_tr[i * 3 + j] * id[k * 3 + l] + id[i * 3 + j] * s_tr[k * 3 + l]);
            double v = cv1 * oo + cv2 * Isym + (1.0 - beta1) * cbar;
            if (f_tr > 0.0)
              v += -2.0 * mubar * beta3 * n[i * 3 + j] * n[k * 3 + l] -
                   2.0 * mubar * beta4 * n[i * 3 + j] * dn2[k * 3 + l];
            c[idx4(i, j, k, l)] = v;
          }
    // pull back: L_IJKL = Fi_Ii Fi_Jj Fi_Kk Fi_Ll c_ijkl, one index at a time
    double t1[81], t2[81];
    for (int I = 0; I < 3; ++I)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double a = 0.0;
            for (int i = 0; i < 3; ++i) a += Fi[I * 3 + i] * c[idx4(i, j, k, l)];
            t1[idx4(I, j, k, l)] = a;
          }
    for (int I = 0; I < 3; ++I)
      for (int Jx = 0; Jx < 3; ++Jx)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double a = 0.0;
            for (int j = 0; j < 3; ++j) a += Fi[Jx * 3 + j] * t1[idx4(I, j, k, l)];
            t2[idx4(I, Jx, k, l)] = a;
          }
    for (int I = 0; I < 3; ++I)
      for (int Jx = 0; Jx < 3; ++Jx)
        for (int K = 0; K < 3; ++K)
          for (int l = 0; l < 3; ++l) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a += Fi[K * 3 + k] * t2[idx4(I, Jx, k, l)];
            t1[idx4(I, Jx, K, l)] = a;
          }
    for (int I = 0; I < 3; ++I)
      for (int Jx = 0; Jx < 3; ++Jx)
        for (int K = 0; K < 3; ++K)
          for (int Lx = 0; Lx < 3; ++Lx) {
            double a = 0.0;
            for (int l = 0; l < 3; ++l) a += Fi[Lx * 3 + l] * t1[idx4(I, Jx, K, l)];
            L[idx4(I, Jx, K, Lx)] = a;
          }
  }
  return 0;
}

// ------------------------------------------------------------------ J2 + continuum damage
// CM_J2P_PARAM (src/cm_j2_plasticity.cc:184-212, 287-347): after the J2 return map the
// damage is integrated with the driving force Y0 = G/2 (tr bbar - 3) + kappa U(J), the
// undamaged elastic energy of the total F (j2d_compute_Y0, :162-182, bbar = J^-2/3 F F^T),
// then S = (1 - w) Sbar (apply_damage_on_stress) and, when the point is damaging,
// L = (1 - w) Lbar - evo/2 (Sbar (x) S0 + S0 (x) Sbar) with evo = dt mu H / (1 + dt mu)
// (:364-382).  S0 = 2 dY0/dC is the stress Y0 derives from.  The damage update,
// continuum_damage_integration_alg_public, lives in the absent library; it is restated from
// the in-tree viscous damage of the same parameters (damage_int_alg with the Weibull
// functions, src/volumetric_damage.cc:196-262; parameters mu, w_max, P1, P2, Yin read by
// CM_J2P_PARAM::read_param, :865-932, w_max >= 1 taken as DAMAGE_THRESH): with dmu = dt mu,
// g = G(Y0) - X_n; if g > 0: w = w_n + dmu/(1 + dmu) g, X = max(X_n, (X_n + dmu G)/(1 + dmu)),
// H = G'(Y0); else w = w_n, X = X_n, H = 0.  w is capped at w_max.
constexpr double kDamageThresh = 0.9999;  // DAMAGE_THRESH, src/volumetric_damage.cc:16

// weibull_function / weibull_evolution, src/volumetric_damage.cc:236-262
void weibull(const orc_material &m, double Y, double *G, double *H) {
  if (Y <= m.Yin) {
    *G = 0.0;
    *H = 0.0;
    return;
  }
  const double z = (Y - m.Yin) / (m.p1 * m.Yin);
  const double e = std::exp(-std::pow(z, m.p2));
  *G = kDamageThresh - kDamageThresh * e;
  *H = kDamageThresh * m.p2 / (m.p1 * m.Yin) * e * std::pow(z, m.p2 - 1.0);
}

double j2d_Y0(const orc_material &m, const double *F, double *S0) {
  const double J = det3x3(F);
  double C[9], CI[9];
  matmul_tn(F, F, C);
  inv3x3(C, CI);
  const double Jm23 = std::pow(J, -2.0 / 3.0);
  const double trC = C[0] + C[4] + C[8];
  double dU, d2U, U;
  vol_pot(m.volPotFlag, J, &dU, &d2U, &U);
  const double kappa = kappa_of(m);
  if (S0) {
    const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i)
      S0[i] = m.G * Jm23 * (id[i] - trC / 3.0 * CI[i]) + kappa * J * dU * CI[i];
  }
  return 0.5 * m.G * (Jm23 * trC - 3.0) + kappa * U;
}

int j2d_update(const orc_material &m, const double *F, const double *Cpinv_n, double ep_n, double w_n, double X_n,
               double dt, double *Cpinv_np1, double *ep_np1, double *gam, double *dam, double *S, double *L,
               double *S0_out) {
  double Sb[9], Lb[81], S0[9];
  const int err = j2_update(m, F, Cpinv_n, ep_n, Cpinv_np1, ep_np1, gam, Sb, L ? Lb : nullptr);
  if (err) return err;
  double w = w_n, X = X_n, H = 0.0;
  int damaged = 0;
  const double Y = j2d_Y0(m, F, S0);
  const double dmu = dt * m.mu;
  if (m.mu > 0.0) {
    double G, Gp;
    weibull(m, Y, &G, &Gp);
    const double g = G - X_n;
    if (g > 0.0) {
      damaged = 1;
      w = w_n + dmu / (1.0 + dmu) * g;
      X = std::max(X_n, (X_n + dmu * G) / (1.0 + dmu));
      H = Gp;
    }
    const double wmax = m.w_max >= 1.0 ? kDamageThresh : m.w_max;
    w = std::min(w, wmax);
  }
  if (dam) {
    dam[0] = w;
    dam[1] = X;
    dam[2] = H;
    dam[3] = damaged;
  }
  if (S0_out) std::memcpy(S0_out, S0, sizeof(S0));
  const double om = 1.0 - w;
  if (S)
    for (int i = 0; i < 9; ++i) S[i] = om * Sb[i];
  if (L) {
    const double evo = damaged ? dmu * H / (1.0 + dmu) : 0.0;
    for (int i = 0; i < 9; ++i)
      for (int k = 0; k < 9; ++k) L[i * 9 + k] = om * Lb[i * 9 + k] - 0.5 * evo * (Sb[i] * S0[k] + S0[i] * Sb[k]);
  }
  return 0;
}

// ------------------------------------------------------------------ element contractions
// compute_residual_vector, src/constitutive_model.cc:1628-1662
void compute_residual_vector(double *f, int nne, const double *ST, double detJxW,
                             const double *Fr, const double *eFnMT, const double *eFnM,
                             const double *S, double one_over_Jn) {
  for (int a = 0; a < nne; a++)
    for (int b = 0; b < 3; b++) {
      const double *STab = ST + a * 27 + b * 9;
      double AA[9], sAA[9], T[9], V[9];
      matmul_tn(Fr, STab, AA);  // AA(i,j) = Fr(k,i) ST_ab(k,j)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sAA[i * 3 + j] = 0.5 * (AA[i * 3 + j] + AA[j * 3 + i]);
      matmul(eFnMT, sAA, T);
      matmul(T, eFnM, V);
      double v = 0.0;
      for (int i = 0; i < 9; ++i) v += V[i] * S[i];
      f[a * 3 + b] += one_over_Jn * detJxW * v;
    }
}

// compute_stiffness_matrix, src/constitutive_model.cc:1520-1616
void compute_stiffness_matrix(double *lk, int nne, const double *ST, double detJxW,
                              const double *Fr, const double *eFnMT, const double *eFn,
                              const double *M, const double *FrTFr, const double *eFnM,
                              const double *S, const double *L, const double *dMdu_all,
                              double one_over_Jn) {
  const int ndofe = nne * 3;
  double MI[9];
  inv3x3(M, MI);
  for (int a = 0; a < nne; a++)
    for (int b = 0; b < 3; b++) {
      const double *STab = ST + a * 27 + b * 9;
      double AA[9], sAA[9], T[9], A_eFn[9], A_eFnM[9];
      matmul_tn(Fr, STab, AA);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sAA[i * 3 + j] = 0.5 * (AA[i * 3 + j] + AA[j * 3 + i]);
      matmul(eFnMT, sAA, T);
      matmul(T, eFn, A_eFn);      // MTeFnT_sAA_eFn
      matmul(A_eFn, M, A_eFnM);   // MTeFnT_sAA_eFnM
      for (int w = 0; w < nne; w++)
        for (int g = 0; g < 3; g++) {
          const double *STwg = ST + w * 27 + g * 9;
          const double *dMdu = dMdu_all + w * 27 + g * 9;
          double BB[9], sBB[9], CC[9], sCC[9], T1[9], T2[9], X1[9], dCdu[9];
          matmul_tn(Fr, STwg, BB);
          matmul_tn(STab, STwg, CC);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              sBB[i * 3 + j] = 0.5 * (BB[i * 3 + j] + BB[j * 3 + i]);
              sCC[i * 3 + j] = 0.5 * (CC[i * 3 + j] + CC[j * 3 + i]);
            }
          // dCdu = sym(M^T eFn^T Fr^T Fr eFn dMdu) + M^T eFn^T sBB eFn M
          matmul(eFnMT, FrTFr, T1);
          matmul(T1, eFn, T2);
          matmul(T2, dMdu, X1);
          double Y1[9], Y2[9];
          matmul(eFnMT, sBB, Y1);
          matmul(Y1, eFnM, Y2);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              dCdu[i * 3 + j] = 0.5 * (X1[i * 3 + j] + X1[j * 3 + i]) + Y2[i * 3 + j];
          double L_dCdu[9];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) s += L[idx4(i, j, k, l)] * dCdu[k * 3 + l];
              L_dCdu[i * 3 + j] = s;
            }
          double t_LdC = 0.0, t_CS = 0.0, t_dM = 0.0, t_S = 0.0, DpJ = 0.0;
          double Z1[9], Z2[9], W1[9];
          matmul(eFnMT, sCC, Z1);
          matmul(Z1, eFnM, Z2);
          matmul(A_eFn, dMdu, W1);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              t_LdC += A_eFnM[i * 3 + j] * L_dCdu[i * 3 + j];
              t_CS += Z2[i * 3 + j] * S[i * 3 + j];
              t_dM += 0.5 * (W1[i * 3 + j] + W1[j * 3 + i]) * S[i * 3 + j];
              t_S += A_eFnM[i * 3 + j] * S[i * 3 + j];
              DpJ += -MI[j * 3 + i] * dMdu[i * 3 + j];
            }
          lk[(a * 3 + b) * ndofe + (w * 3 + g)] +=
              one_over_Jn * detJxW * (t_LdC + 2.0 * t_dM + t_CS + DpJ * t_S);
        }
    }
}

}  // namespace

extern "C" {

void orc_stiffness_contract(double *lk, int nne, const double *ST, double detJxW, const double *Fr,
                            const double *eFnMT, const double *eFn, const double *M, const double *FrTFr,
                            const double *eFnM, const double *S, const double *L, const double *dMdu_all,
                            double one_over_Jn) {
  compute_stiffness_matrix(lk, nne, ST, detJxW, Fr, eFnMT, eFn, M, FrTFr, eFnM, S, L, dMdu_all, one_over_Jn);
}

int orc_nint(int nne) { return nint_of(nne); }

int orc_basis(int nne, const double *X, int ip, double *N, double *dNdX, double *detJxW) {
  double k, e, z, w, dNk[30];
  int_point(nne, ip, &k, &e, &z, &w);
  shape_N(nne, k, e, z, N);
  shape_dN(nne, k, e, z, dNk);
  const double J = dNdX_of(nne, dNk, X, dNdX);
  *detJxW = J * w;
  return J > 0.0 ? 0 : 1;
}

int orc_he_update(const orc_material *mat, const double *F, double *S, double *L) {
  return he_update(*mat, F, S, L);
}

int orc_j2_update(const orc_material *mat, const double *F, const double *Cpinv_n, double ep_n,
                  double *Cpinv_np1, double *ep_np1, double *gam, double *S, double *L) {
  return j2_update(*mat, F, Cpinv_n, ep_n, Cpinv_np1, ep_np1, gam, S, L);
}

int orc_j2d_update(const orc_material *mat, const double *F, const double *Cpinv_n, double ep_n, double w_n,
                   double X_n, double dt, double *Cpinv_np1, double *ep_np1, double *gam, double *dam_np1,
                   double *S, double *L, double *S0) {
  return j2d_update(*mat, F, Cpinv_n, ep_n, w_n, X_n, dt, Cpinv_np1, ep_np1, gam, dam_np1, S, L, S0);
}

void orc_weibull(const orc_material *mat, double Y, double *G, double *H) { weibull(*mat, Y, G, H); }

double orc_j2d_Y0(const orc_material *mat, const double *F) { return j2d_Y0(*mat, F, nullptr); }

// generate_local_dof_ids (+ _on_elem), src/generate_dof_ids.cc:69-171,306-412, without
// periodic nodes, element DOFs or bounding elements.
int orc_generate_local_dof_ids(int64_t n_nodes, int64_t n_elem, const int64_t *elem_ptr,
                               const int64_t *elem_conn, int64_t n_coel, const int64_t *coel_ptr,
                               const int64_t *coel_conn, int64_t *id, int64_t *ndof_out) {
  const int64_t ndofn = 3;
  std::vector<long> visited_node_dof(ndofn * n_nodes, 0);
  long ndof = 1;
  auto on_elem = [&](const int64_t *nod, int64_t nne) {
    long id_adder = 0;
    for (int64_t i = 0; i < nne; i++) {
      const int64_t node_id = nod[i];
      for (int64_t j = 0; j < ndofn; j++) {
        const int64_t dof_idx = node_id * ndofn + j;
        if (!visited_node_dof[dof_idx]) {
          if (id[dof_idx] < 0) {
            visited_node_dof[dof_idx] = 1;
          } else if (id[dof_idx] == 1) {
            id[dof_idx] = 0;
            visited_node_dof[dof_idx] = 1;
          } else if (id[dof_idx] >= 0) {
            id[dof_idx] = ndof + id_adder;
            id_adder++;
            visited_node_dof[dof_idx] = 1;
          }
        }
      }
    }
    return id_adder;
  };
  for (int64_t i = 0; i < n_elem; i++) ndof += on_elem(elem_conn + elem_ptr[i], elem_ptr[i + 1] - elem_ptr[i]);
  for (int64_t i = 0; i < n_coel; i++) ndof += on_elem(coel_conn + coel_ptr[i], coel_ptr[i + 1] - coel_ptr[i]);
  for (int64_t i = 0; i < ndofn * n_nodes; i++)
    if (!visited_node_dof[i]) return 1;  // hanging node
  *ndof_out = ndof - 1;
  return 0;
}

static int compare_long(const void *a, const void *b) {
  const long x = *(const long *)a, y = *(const long *)b;
  return (x > y) - (x < y);
}

// Psparse_ApAi for one rank, src/Psparse_ApAi.cc:238-338,600-658 with countDof/mapDof
// (:160-213); on one rank the local and global ids coincide (LG = identity).
int orc_psparse_apai(int64_t n_nodes, int64_t n_elem, const int64_t *elem_ptr,
                     const int64_t *elem_conn, int64_t n_coel, const int64_t *coel_ptr,
                     const int64_t *coel_conn, const int64_t *id, int64_t ndofd, int32_t *Ap,
                     int32_t *Ai, int64_t Ai_cap, int64_t *nnz) {
  (void)n_nodes;
  std::vector<long> ncols(ndofd, 0);
  auto for_each_dof = [&](const int64_t *nod, int64_t nne, std::vector<long> &lids) {
    lids.clear();
    for (int64_t a = 0; a < nne; ++a)
      for (int j = 0; j < 3; ++j) {
        const long lid = id[nod[a] * 3 + j];
        if (lid > 0) lids.push_back(lid - 1);
      }
  };
  std::vector<long> lids;
  // 1. countDof
  for (int64_t e = 0; e < n_elem; ++e) {
    for_each_dof(elem_conn + elem_ptr[e], elem_ptr[e + 1] - elem_ptr[e], lids);
    for (long l : lids) ncols[l] += lids.size();
  }
  for (int64_t c = 0; c < n_coel; ++c) {
    for_each_dof(coel_conn + coel_ptr[c], coel_ptr[c + 1] - coel_ptr[c], lids);
    for (long l : lids) ncols[l] += lids.size();
  }
  std::vector<std::vector<long>> AA(ndofd);
  for (int64_t i = 0; i < ndofd; i++) AA[i].resize(ncols[i]);
  std::vector<long> ap1(ndofd, 0);
  // 2. mapDof (gids = lids on one rank)
  auto map_dof = [&](const int64_t *nod, int64_t nne) {
    for_each_dof(nod, nne, lids);
    for (long l : lids)
      for (long g : lids) AA[l][ap1[l]++] = g;
  };
  for (int64_t e = 0; e < n_elem; ++e) map_dof(elem_conn + elem_ptr[e], elem_ptr[e + 1] - elem_ptr[e]);
  for (int64_t c = 0; c < n_coel; ++c) map_dof(coel_conn + coel_ptr[c], coel_ptr[c + 1] - coel_ptr[c]);
  // 3. sort each row, count unique
  for (int64_t k = 0; k < ndofd; k++) qsort(AA[k].data(), ap1[k], sizeof(long), compare_long);
  std::vector<long> ap(ndofd, 0);
  for (int64_t k = 0; k < ndofd; k++) {
    for (long j = 0; j < ap1[k] - 1; j++)
      if (AA[k][j] < AA[k][j + 1]) ap[k]++;
    ap[k]++;
  }
  // 4. Ap as running sum, Ai = unique sorted ids (:639-658)
  int64_t total = 0;
  for (int64_t k = 0; k < ndofd; ++k) total += ap[k];
  *nnz = total;
  Ap[0] = 0;
  for (int64_t k = 0; k < ndofd; ++k) Ap[k + 1] = Ap[k] + int32_t(ap[k]);
  if (!Ai) return 0;
  if (Ai_cap < total) return 1;
  for (int64_t k = 0; k < ndofd; k++) {
    int64_t p = Ap[k];
    for (long j = 0; j < ap1[k] - 1; j++)
      if (AA[k][j] < AA[k][j + 1]) Ai[p++] = int32_t(AA[k][j]);
    Ai[p++] = int32_t(AA[k][ap1[k] - 1]);
  }
  return 0;
}

// One rank of fd_residuals_MP (src/fd_residuals.cc:468-636) and stiffmat_fd_MP
// (src/stiffmat_fd.cc:831-993) on the CM path, total Lagrangian: elements in mesh order;
// per element def_elem_total (src/utils.cc:1391-1412), then
// residuals_el_constitutive_model_1f (src/constitutive_model.cc:3322-3539) assembled by
// fd_res_assemble (src/fd_residuals.cc:58-73) and stiffness_el_constitutive_model_1f
// (:2474-2692) assembled by PLoc_Sparse (src/PLoc_Sparse.cc:15-131) into the CSR values.
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_colour_elements(int64_t n_nodes, int64_t n_elem, const int64_t *elem_ptr, const int64_t *elem_conn,
                        int64_t *order, int64_t *cptr) {
  std::vector<std::vector<uint64_t>> used(1, std::vector<uint64_t>(n_nodes, 0));  // colour bits per node
  std::vector<int> colour(n_elem);
  int nc = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    int c = 0;
    for (;; ++c) {
      const size_t w = size_t(c / 64);
      const uint64_t bit = uint64_t(1) << (c % 64);
      if (w >= used.size()) used.emplace_back(n_nodes, 0);
      bool free = true;
      for (int64_t k = elem_ptr[e]; k < elem_ptr[e + 1] && free; ++k) free = !(used[w][elem_conn[k]] & bit);
      if (free) break;
    }
    colour[e] = c;
    for (int64_t k = elem_ptr[e]; k < elem_ptr[e + 1]; ++k) used[size_t(c / 64)][elem_conn[k]] |= uint64_t(1) << (c % 64);
    nc = std::max(nc, c + 1);
  }
  std::fill(cptr, cptr + nc + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) cptr[colour[e] + 1]++;
  for (int c = 0; c < nc; ++c) cptr[c + 1] += cptr[c];
  std::vector<int64_t> cur(cptr, cptr + nc);
  for (int64_t e = 0; e < n_elem; ++e) order[cur[colour[e]]++] = e;
  return nc;
}

int orc_assemble(int64_t n_nodes, const double *X, int64_t n_elem, const int64_t *elem_ptr,
                 const int64_t *elem_conn, const int32_t *elem_mat, const orc_material *mats,
                 const int64_t *id, int64_t ndof, const int32_t *Ap, const int32_t *Ai,
                 const double *r, const double *d_r, const double *defl, const double *defl_d,
                 double dt, int do_residual, int run_integration, int do_tangent,
                 double *F_np1, const double *plastic_n, double *plastic_np1,
                 const double *scalar_n, double *scalar_np1, double *f, double *vals) {
  // QP offsets of the elements, then the elements colour by colour: within a colour no two
  // elements share a node, so their scatters into f and vals touch distinct entries and run in
  // parallel; every entry's sum follows the colour order whatever the thread count.
  std::vector<int64_t> qoff(n_elem + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) qoff[e + 1] = qoff[e] + nint_of(int(elem_ptr[e + 1] - elem_ptr[e]));
  const int64_t NQ = qoff[n_elem];
  std::vector<int64_t> order(n_elem), cptr(n_elem + 1);
  const int nc = orc_colour_elements(n_nodes, n_elem, elem_ptr, elem_conn, order.data(), cptr.data());
  if (do_residual && f) std::memset(f, 0, sizeof(double) * ndof);
  if (do_tangent && vals) std::memset(vals, 0, sizeof(double) * Ap[ndof]);
  int err = 0, missing = 0;
  auto element = [&](int64_t e) -> int {
    int err = 0;
    const int64_t qp0 = qoff[e];
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]);
    const int nint = nint_of(nne);
    const int ndofe = nne * 3;
    const int64_t *nod = elem_conn + elem_ptr[e];
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    double x[30], u[30];
    int64_t cn[30];
    for (int a = 0; a < nne; ++a)
      for (int b = 0; b < 3; ++b) {
        x[a * 3 + b] = X[nod[a] * 3 + b];
        const int64_t dof = id[nod[a] * 3 + b];
        cn[a * 3 + b] = dof;
        const int64_t aid = std::llabs(dof) - 1;
        if (dof == 0) u[a * 3 + b] = 0.0;
        else if (dof > 0) u[a * 3 + b] = r[aid] + (d_r ? d_r[aid] : 0.0);
        else u[a * 3 + b] = (defl ? defl[aid] : 0.0) + (defl_d ? defl_d[aid] : 0.0);
      }
    double fe[30] = {0};
    std::vector<double> lk(size_t(ndofe) * ndofe, 0.0);
    const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double dMdu_all[270];
    std::memset(dMdu_all, 0, sizeof(dMdu_all));
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 0 && !do_residual) continue;
      if (pass == 1 && !do_tangent) continue;
      for (int ip = 0; ip < nint; ++ip) {
        const int64_t q = qp0 + ip;
        double N[10], dNdX[30], detJxW;
        if (orc_basis(nne, x, ip, N, dNdX, &detJxW)) { ++err; continue; }
        double ST[270];
        for (int a = 0; a < nne; ++a)  // compute_shape_tensor, femlib.cc:47-69
          for (int b = 0; b < 3; ++b)
            for (int i = 0; i < 3; ++i)
              for (int J = 0; J < 3; ++J) ST[a * 27 + b * 9 + i * 3 + J] = (b == i) ? dNdX[a * 3 + J] : 0.0;
        double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // compute_deformation_gradient :1289-1324
        for (int a = 0; a < nne; a++)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) Fr[b * 3 + c] += u[a * 3 + b] * dNdX[a * 3 + c];
        double Fs[9];
        if (pass == 0) {
          if (!(det3x3(Fr) > 0.0)) { ++err; continue; }  // getJacobian
          if (run_integration)
            for (int i = 0; i < 9; ++i) F_np1[i * NQ + q] = Fr[i];
        }
        for (int i = 0; i < 9; ++i) Fs[i] = F_np1[i * NQ + q];
        double S[9], L[81];
        double *Lp = (pass == 1) ? L : nullptr;
        int merr = 0;
        if (mat.model == 4) {  // J2: integration stores the n+1 state, S/L from the state
          double cn6[6], c6[6], ep_n = scalar_n[0 * NQ + q], ep1, gm, dam[4];
          for (int i = 0; i < 6; ++i) cn6[i] = plastic_n[i * NQ + q];
          // scalars (ep, gam, w, X, H, damaged); damage from (w_n, X_n) when mu > 0
          merr = j2d_update(mat, Fs, cn6, ep_n, scalar_n[2 * NQ + q], scalar_n[3 * NQ + q], dt, c6, &ep1, &gm, dam,
                            S, Lp, nullptr);
          if (pass == 0 && run_integration && !merr) {
            for (int i = 0; i < 6; ++i) plastic_np1[i * NQ + q] = c6[i];
            scalar_np1[0 * NQ + q] = ep1;
            scalar_np1[1 * NQ + q] = gm;
            for (int i = 0; i < 4; ++i) scalar_np1[(2 + i) * NQ + q] = dam[i];
          }
        } else {
          merr = he_update(mat, Fs, S, Lp);
        }
        if (merr) { ++err; continue; }
        // TL: eFn = I, pFnp1 = I -> M = I, pJ = 1 (HE_PARAM / CM_J2P_PARAM::get_pF)
        if (pass == 0) {
          compute_residual_vector(fe, nne, ST, detJxW, Fr, I9, I9, S, 1.0);
        } else {
          double FrTFr[9];
          matmul_tn(Fr, Fr, FrTFr);
          compute_stiffness_matrix(lk.data(), nne, ST, detJxW, Fr, I9, I9, I9, FrTFr, I9, S, L,
                                   dMdu_all, 1.0);
        }
      }
    }
    if (do_tangent)  // non-normal diagonal check, src/constitutive_model.cc:2682-2687
      for (int a = 0; a < ndofe; ++a)
        if (!std::isnormal(lk[a * ndofe + a])) ++err;
    if (do_residual && f)
      for (int k = 0; k < ndofe; ++k) {
        const int64_t II = cn[k] - 1;
        if (II < 0) continue;
        f[II] += fe[k];
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
          while (p < en && *p != col) ++p;  // SparseSystem::add by (row, col)
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

}  // extern "C"

// ------------------------------------------------------------------ three-field (CM3F)
// The CM3F mixed formulation of the hyperelastic model: per element one volume variable theta
// and one pressure P (get_3f_pressure_volume_number, src/three_field_element.cc:1399-1416:
// npres = nVol = 1, constant shape functions Np = Nt = 1, femlib.cc:1189-1193), total
// Lagrangian.  For this model pF = I (HE_PARAM::get_pF), so M = I, dM/du = dM/dtheta = 0
// (Model_parameters::compute_dMdu/dMdt, include/constitutive_model.h:662-698), eFn = I and
// Jn = tJn = theta_n = 1 (ConstitutiveModelIntregrate::integrate_ss,
// src/constitutive_model.cc:794-868).  The element quantities are those of CM_ThreeField,
// src/constitutive_model_3f.cc:189-470, written here in their material form:
//   S, L   deviatoric PK2 and tangent at eF = Fr (update_elasticity_dev, constitutive_model.h:413-434)
//   dU,ddU kappa U'(theta), kappa U''(theta) (compute_dudj/d2udj2, constitutive_model.h:319-384)
// Checks: Kuu, Kup, Ktt, Ktp against central differences of Ru, Rt, Rp; at theta = J and
// P = kappa U'(J) on linear tets Ru equals the displacement element's residual; the
// condensation equals a monolithic solve of the element's (u, theta, P) system, and the
// condensed global solve with orc_tf_update that of a whole mesh's system; a converged
// CM3F Newton on linear tets equals the displacement solution -> test_oracle_three_field.py.
namespace {

struct TfElem {
  double Kuu[900], Ru[30], Kup[30], Kut[30], Ktu[30];
  double Rp, Rt, Ktt, Ktp;
};

void tf_dof_grad(const double *dNdX, int k, double *G) {  // Grad_du = ST_(a,b): e_b (x) dN_a
  const int a = k / 3, b = k % 3;
  for (int i = 0; i < 9; ++i) G[i] = 0.0;
  for (int J = 0; J < 3; ++J) G[b * 3 + J] = dNdX[a * 3 + J];
}

// J2: the stored state enters through jst (per QP Cp^-1 at n [nint][6] and ep at n [nint]; the
// n+1 values are written to Cp1 / ep1 / gam1 when given).
struct TfJ2State {
  const double *Cpn, *epn;
  double *Cp1, *ep1, *gam1;
};

int tf_element(const orc_material &mat, int nne, const double *x, const double *u, double theta, double P,
               bool want_K, TfElem &E, double *Fbar_out /* [nint][9] or null */,
               const TfJ2State *jst = nullptr) {
  const int nd = nne * 3;
  std::memset(&E, 0, sizeof(E));
  const double kappa = kappa_of(mat);
  double dU, ddU, U;
  vol_pot(mat.volPotFlag, theta, &dU, &ddU, &U);
  dU *= kappa;
  ddU *= kappa;
  const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  int err = 0;
  for (int ip = 0; ip < nint_of(nne); ++ip) {
    double N[10], dNdX[30], w;
    if (orc_basis(nne, x, ip, N, dNdX, &w)) { ++err; continue; }
    double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int a = 0; a < nne; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) Fr[b * 3 + c] += u[a * 3 + b] * dNdX[a * 3 + c];
    const double tJr = det3x3(Fr);
    double Fi[9];
    if (!(tJr > 0.0) || inv3x3(Fr, Fi)) { ++err; continue; }
    if (Fbar_out) {  // the model's stored F: (theta/J)^(1/3) Fr (constitutive_model.cc:815, :1115)
      const double fac = std::cbrt(theta / tJr);
      for (int i = 0; i < 9; ++i) Fbar_out[ip * 9 + i] = fac * Fr[i];
    }
    double C[9], S[9], L[81];
    matmul_tn(Fr, Fr, C);
    if (mat.model == 4) {
      // J2 (CM_J2P_PARAM: pF = I, so M = I and dM/du = 0 as for the hyperelastic model): the
      // return map at Fr (its stored state does not change with the volumetric scaling of Fb,
      // so Fr and Fb integrate alike), and S, L less their volumetric parts kappa U(J) --
      // the deviatoric stress and tangent of update_elasticity_dev
      // (src/cm_j2_plasticity.cc:384-461)
      const double c6[6] = {1, 1, 1, 0, 0, 0};
      const double *Cn = (jst && jst->Cpn) ? jst->Cpn + ip * 6 : c6;
      const double epn = (jst && jst->epn) ? jst->epn[ip] : 0.0;
      double Cp1[6], ep1, gam, Lf[81], CI[9];
      if (orc_j2_update(&mat, Fr, Cn, epn, Cp1, &ep1, &gam, S, Lf)) { ++err; continue; }
      if (jst && jst->Cp1)
        for (int i = 0; i < 6; ++i) jst->Cp1[ip * 6 + i] = Cp1[i];
      if (jst && jst->ep1) jst->ep1[ip] = ep1;
      if (jst && jst->gam1) jst->gam1[ip] = gam;
      inv3x3(C, CI);
      double vdU, vd2U, vU;
      vol_pot(mat.volPotFlag, tJr, &vdU, &vd2U, &vU);
      const double kap = kappa_of(mat);
      for (int i = 0; i < 9; ++i) S[i] -= kap * tJr * vdU * CI[i];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
              const int q = idx4(i, j, k, l);
              L[q] = Lf[q] - (kap * (tJr * vdU + tJr * tJr * vd2U) * CI[i * 3 + j] * CI[k * 3 + l] -
                              2.0 * kap * tJr * vdU * CI[i * 3 + k] * CI[l * 3 + j]);
            }
    } else {
      dev_stress_mr(C, mat.m10, mat.m01, S);
      if (want_K) mat_stiff_mr(C, mat.m10, mat.m01, L);
    }
    const double *M = I9, *eFn = I9;
    double Psi[30][9], trGF[30];
    for (int k = 0; k < nd; ++k) {
      double G[9], A[9], As[9], Phi[9], T[9];
      tf_dof_grad(dNdX, k, G);
      matmul_tn(G, Fr, A);  // A = Grad_du^T tFr
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) As[i * 3 + j] = 0.5 * (A[i * 3 + j] + A[j * 3 + i]);
      for (int i = 0; i < 9; ++i) T[i] = 0.0;  // Phi = eFn^T As eFn, Psi = M^T Phi M
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int kk = 0; kk < 3; ++kk)
            for (int l = 0; l < 3; ++l) T[i * 3 + j] += eFn[kk * 3 + i] * As[kk * 3 + l] * eFn[l * 3 + j];
      for (int i = 0; i < 9; ++i) Phi[i] = T[i];
      for (int i = 0; i < 9; ++i) Psi[k][i] = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int kk = 0; kk < 3; ++kk)
            for (int l = 0; l < 3; ++l) Psi[k][i * 3 + j] += M[kk * 3 + i] * Phi[kk * 3 + l] * M[l * 3 + j];
      double tr = 0.0;  // Grad_du(j,i) tFrI(i,j)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tr += G[j * 3 + i] * Fi[i * 3 + j];
      trGF[k] = tr;
      double PsiS = 0.0;
      for (int i = 0; i < 9; ++i) PsiS += Psi[k][i] * S[i];
      E.Ru[k] += (PsiS + P * tJr * tr) * w;          // compute_Ru, :190-212
      E.Kup[k] += tJr * tr * w;                       // compute_Kup, :292-316 (dt factor -1)
    }
    E.Rp += (tJr - theta) * w;                        // compute_Rp, :219-226
    E.Rt += (dU - P) * w;                             // compute_Rt, :237-246
    E.Ktt += ddU * w;                                 // compute_Ktt, :438-466 (dM/dtheta = 0)
    E.Ktp -= w;                                       // compute_Ktp, :418-430
    if (!want_K) continue;
    for (int k = 0; k < nd; ++k) {
      double Gk[9];
      tf_dof_grad(dNdX, k, Gk);
      for (int l = 0; l < nd; ++l) {
        double Gl[9];
        tf_dof_grad(dNdX, l, Gl);
        double PLP = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk)
              for (int ll = 0; ll < 3; ++ll) PLP += Psi[k][i * 3 + j] * L[idx4(i, j, kk, ll)] * Psi[l][kk * 3 + ll];
        double tr2 = 0.0;  // Grad_du(j,i) tFrI(i,k) Grad_tu(k,l) tFrI(l,j)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk)
              for (int ll = 0; ll < 3; ++ll) tr2 += Gk[j * 3 + i] * Fi[i * 3 + kk] * Gl[kk * 3 + ll] * Fi[ll * 3 + j];
        const double pJ = P * tJr * (trGF[k] * trGF[l] - tr2);
        double GG[9], DS = 0.0;  // DPsi = eFnM^T sym(Grad_du^T Grad_tu) eFnM, eFnM = I
        matmul_tn(Gk, Gl, GG);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) DS += 0.5 * (GG[i * 3 + j] + GG[j * 3 + i]) * S[i * 3 + j];
        E.Kuu[k * nd + l] += w * (PLP + pJ + DS);  // compute_Kuu, :249-286 (dt factor -1)
      }
    }
  }
  return err;
}

// condense_K_3F_to_1F / condense_F_3F_to_1F, src/condense.cc:304-385, Pno = Vno = 1
// (Kpt = Ktp^T, Kpu = Kup^T: IntegrateThreeFieldStiffness::out_loop, constitutive_model.cc:531-555)
void tf_condense(int nd, const TfElem &E, double *K, double *f) {
  const double Kpt = E.Ktp;
  for (int i = 0; i < nd; ++i) {
    if (K)
      for (int j = 0; j < nd; ++j)
        K[i * nd + j] = E.Kuu[i * nd + j] - E.Kup[i] / E.Ktp * E.Ktu[j] +
                        E.Kup[i] / E.Ktp * E.Ktt / Kpt * E.Kup[j] - E.Kut[i] / Kpt * E.Kup[j];
    if (f) f[i] = E.Ru[i] + E.Kup[i] / E.Ktp * E.Ktt / Kpt * E.Rp - E.Kup[i] / E.Ktp * E.Rt - E.Kut[i] / Kpt * E.Rp;
  }
}

void tf_gather(const double *X, const int64_t *nod, int nne, const int64_t *id, const double *r, const double *d_r,
               const double *defl, const double *defl_d, double *x, double *u, int64_t *cn) {
  for (int a = 0; a < nne; ++a)
    for (int b = 0; b < 3; ++b) {  // def_elem_total, src/utils.cc:1391-1412
      x[a * 3 + b] = X[nod[a] * 3 + b];
      const int64_t dof = id[nod[a] * 3 + b];
      cn[a * 3 + b] = dof;
      const int64_t aid = std::llabs(dof) - 1;
      if (dof == 0) u[a * 3 + b] = 0.0;
      else if (dof > 0) u[a * 3 + b] = r[aid] + (d_r ? d_r[aid] : 0.0);
      else u[a * 3 + b] = (defl ? defl[aid] : 0.0) + (defl_d ? defl_d[aid] : 0.0);
    }
}

}  // namespace

extern "C" {

// One CM3F element (material form) for tests: the blocks of CM_ThreeField (Kuu row-major
// [3nne][3nne], Ru, Kup, Kut, Ktu; sc = Rp, Rt, Ktt, Ktp) and the condensed K and f.
int orc_tf_element(const orc_material *mat, int nne, const double *x, const double *u, double theta, double P,
                   double *Kuu, double *Ru, double *Kup, double *sc, double *Kc, double *fc, const double *Cpn,
                   const double *epn) {
  TfElem E;
  const TfJ2State jst{Cpn, epn, nullptr, nullptr, nullptr};
  const int err = tf_element(*mat, nne, x, u, theta, P, Kuu || Kc, E, nullptr, &jst);
  const int nd = 3 * nne;
  if (Kuu) std::memcpy(Kuu, E.Kuu, sizeof(double) * nd * nd);
  if (Ru) std::memcpy(Ru, E.Ru, sizeof(double) * nd);
  if (Kup) std::memcpy(Kup, E.Kup, sizeof(double) * nd);
  if (sc) { sc[0] = E.Rp; sc[1] = E.Rt; sc[2] = E.Ktt; sc[3] = E.Ktp; }
  if (Kc || fc) tf_condense(nd, E, Kc, fc);
  return err;
}

// fd_residuals_MP + stiffmat_fd_MP of the CM3F analysis (residuals_el_constitutive_model_3f /
// stiffness_el_constitutive_model_3f, src/constitutive_model.cc:2714-2735, :3565-3590):
// theta[e] = V_np1 + dV and P[e] = P_np1 + dP of each element; with run_integration the
// model's F(n+1) (per QP) is stored.
// The element loop of an assembly, colour by colour (orc_colour_elements): within a colour no two
// elements share a node, so their scatters touch distinct entries of f and vals and run on the
// host's threads; every entry's sum follows the colour order whatever the thread count.
// element(e) returns its error count, or -1 when a CSR entry is missing.
extern "C++" template <class Fn>
int assemble_by_colour(int64_t n_elem, const int64_t *elem_ptr, const int64_t *elem_conn, Fn element) {
  int64_t n_nodes = 0;
  for (int64_t k = 0; k < elem_ptr[n_elem]; ++k) n_nodes = std::max<int64_t>(n_nodes, elem_conn[k] + 1);
  std::vector<int64_t> order(n_elem), cptr(n_elem + 1);
  const int nc = orc_colour_elements(n_nodes, n_elem, elem_ptr, elem_conn, order.data(), cptr.data());
  int err = 0, missing = 0;
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

int orc_assemble_tf(int64_t n_elem, const double *X, const int64_t *elem_ptr, const int64_t *elem_conn,
                    const int32_t *elem_mat, const orc_material *mats, const int64_t *id, int64_t ndof,
                    const int32_t *Ap, const int32_t *Ai, const double *r, const double *d_r, const double *defl,
                    const double *defl_d, const double *theta, const double *P, int run_integration,
                    double *F_np1, int64_t n_qp, double *f, double *vals, const double *plastic_n,
                    double *plastic_np1, const double *scalar_n, double *scalar_np1) {
  if (f) std::memset(f, 0, sizeof(double) * ndof);
  if (vals) std::memset(vals, 0, sizeof(double) * Ap[ndof]);
  std::vector<int64_t> qoff(n_elem + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) qoff[e + 1] = qoff[e] + nint_of(int(elem_ptr[e + 1] - elem_ptr[e]));
  return assemble_by_colour(n_elem, elem_ptr, elem_conn, [&](int64_t e) -> int {
    int err = 0;
    const int64_t qp0 = qoff[e];
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]), nd = 3 * nne, nint = nint_of(nne);
    double x[30], u[30], Fb[72];
    int64_t cn[30];
    tf_gather(X, elem_conn + elem_ptr[e], nne, id, r, d_r, defl, defl_d, x, u, cn);
    TfElem E;
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    // J2: the element's QP state from the [comp][n_qp] arrays (Cp^-1 6 comps, scalars ep, gam, ...)
    double Cpn[48], epn[8], Cp1[48], ep1[8], gam1[8];
    for (int ip = 0; ip < nint; ++ip) {
      const int64_t q = qp0 + ip;
      for (int i = 0; i < 6; ++i) Cpn[ip * 6 + i] = plastic_n ? plastic_n[i * n_qp + q] : (i < 3 ? 1.0 : 0.0);
      epn[ip] = scalar_n ? scalar_n[q] : 0.0;
    }
    const TfJ2State jst{Cpn, epn, Cp1, ep1, gam1};
    err += tf_element(mat, nne, x, u, theta[e], P[e], vals != nullptr, E, Fb, &jst);
    if (run_integration && F_np1)
      for (int ip = 0; ip < nint; ++ip)
        for (int i = 0; i < 9; ++i) F_np1[i * n_qp + qp0 + ip] = Fb[ip * 9 + i];
    if (run_integration && mat.model == 4 && plastic_np1 && scalar_np1)
      for (int ip = 0; ip < nint; ++ip) {
        const int64_t q = qp0 + ip;
        for (int i = 0; i < 6; ++i) plastic_np1[i * n_qp + q] = Cp1[ip * 6 + i];
        scalar_np1[q] = ep1[ip];
        scalar_np1[n_qp + q] = gam1[ip];
      }
    std::vector<double> K(size_t(nd) * nd);
    double fe[30];
    tf_condense(nd, E, K.data(), fe);
    if (vals)
      for (int a = 0; a < nd; ++a)
        if (!std::isnormal(K[a * nd + a])) ++err;  // IntegrateThreeFieldStiffness::out_loop
    for (int i = 0; i < nd; ++i) {
      const int64_t row = cn[i] - 1;
      if (row < 0) continue;
      if (f) f[row] += fe[i];
      if (vals)
        for (int j = 0; j < nd; ++j) {
          const int64_t col = cn[j] - 1;
          if (col < 0) continue;
          const int32_t *p = Ai + Ap[row], *en = Ai + Ap[row + 1];
          while (p < en && *p != col) ++p;
          if (p == en) return -1;
          vals[p - Ai] += K[i * nd + j];
        }
    }
    return err;
  });
}

// constitutive_model_update_NR for CM3F without inertia (src/constitutive_model.cc:3916-3999,
// IntegrateThreeFieldUpdate :595-656, compute_d_theta_dP src/constitutive_model_3f.cc:494-549):
// at u = r + d_r (defl + defl_d), theta, P, the increments of theta and P that go with the
// displacement increment du (prescribed DOFs zero: def_elem TYPE 2, src/utils.cc:1365-1389).
int orc_tf_update(int64_t n_elem, const double *X, const int64_t *elem_ptr, const int64_t *elem_conn,
                  const int32_t *elem_mat, const orc_material *mats, const int64_t *id, const double *r,
                  const double *d_r, const double *defl, const double *defl_d, const double *theta, const double *P,
                  const double *du, double *ddV, double *ddP, int64_t n_qp, const double *plastic_n,
                  const double *scalar_n) {
  int err = 0;
  int64_t qp0 = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]), nd = 3 * nne, nint = nint_of(nne);
    double x[30], u[30];
    int64_t cn[30];
    tf_gather(X, elem_conn + elem_ptr[e], nne, id, r, d_r, defl, defl_d, x, u, cn);
    double Cpn[48], epn[8];
    for (int ip = 0; ip < nint; ++ip) {
      const int64_t q = qp0 + ip;
      for (int i = 0; i < 6; ++i) Cpn[ip * 6 + i] = plastic_n ? plastic_n[i * n_qp + q] : (i < 3 ? 1.0 : 0.0);
      epn[ip] = scalar_n ? scalar_n[q] : 0.0;
    }
    qp0 += nint;
    const TfJ2State jst{Cpn, epn, nullptr, nullptr, nullptr};
    TfElem E;
    err += tf_element(mats[elem_mat ? elem_mat[e] : 0], nne, x, u, theta[e], P[e], false, E, nullptr, &jst);
    const double Kpt = E.Ktp;
    double Kpu_du = 0.0, Ktu_du = 0.0;
    for (int k = 0; k < nd; ++k) {
      const double d = cn[k] > 0 ? du[cn[k] - 1] : 0.0;
      Kpu_du += E.Kup[k] * d;
      Ktu_du += E.Ktu[k] * d;
    }
    const double dth = -(Kpu_du + E.Rp) / Kpt;
    ddV[e] = dth;
    ddP[e] = -(E.Rt + Ktu_du + E.Ktt * dth) / E.Ktp;
  }
  return err;
}

// compute_3f_initial_conditions, src/three_field_element.cc:1328-1386: per element the
// volume averages of J and of kappa U'(J) at the nodal displacements u ([n_nodes][3]).
int orc_tf_initial(int64_t n_elem, const double *X, const int64_t *elem_ptr, const int64_t *elem_conn,
                   const int32_t *elem_mat, const orc_material *mats, const double *u_nodal, double *V,
                   double *Pr) {
  int err = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]);
    const int64_t *nod = elem_conn + elem_ptr[e];
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    double x[30], u[30];
    for (int a = 0; a < nne; ++a)
      for (int b = 0; b < 3; ++b) {
        x[a * 3 + b] = X[nod[a] * 3 + b];
        u[a * 3 + b] = u_nodal[nod[a] * 3 + b];
      }
    double vol = 0.0, Jv = 0.0, Up = 0.0;
    for (int ip = 0; ip < nint_of(nne); ++ip) {
      double N[10], dNdX[30], w;
      if (orc_basis(nne, x, ip, N, dNdX, &w)) { ++err; continue; }
      double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      for (int a = 0; a < nne; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) F[b * 3 + c] += u[a * 3 + b] * dNdX[a * 3 + c];
      const double J = det3x3(F);
      double dU, d2U, U;
      vol_pot(mat.volPotFlag, J, &dU, &d2U, &U);
      vol += w;
      Jv += w * J;
      Up += w * kappa_of(mat) * dU;
    }
    V[e] = Jv / vol;
    Pr[e] = Up / vol;
  }
  return err;
}

}  // extern "C"

// ------------------------------------------------------------------ CM3F, crystal plasticity
// The CM3F analysis of a CP_PARAM material (ConstitutiveModelIntregrate::integrate_ss,
// src/constitutive_model.cc:720-880, total Lagrangian, one theta and one P per element):
//   the model integrates at Fb = (theta/J)^(1/3) Fr (run_integration_algorithm with
//   tf_factor, :810-818, :1097-1115) and stores Fb as its F(n+1);
//   M = pF(n+1)^-1, eF = Fr M (compute_deformation_gradient_for_ss, :412-438), eFn = I,
//   theta_n = tJn = 1, Jn = 1/pJ (pJ = det pF(n+1)), JM = det M, theta_e = theta JM;
//   Sd, Ld: the deviatoric PK2 and tangent of the material's hyperelastic law at eF
//   (update_elasticity_dev, include/constitutive_model.h:413-434; Mooney-Rivlin, as the
//   hyperelastic CM3F above), dU, ddU = kappa U', kappa U'' at theta_e.
// The blocks are those of CM_ThreeField (src/constitutive_model_3f.cc:189-470) with M and its
// derivatives in place (Psi = M^T Phi M, DPsi, Lambda = tr(M^-1 dM)).  Three departures, all
// in the tangent (the residuals are the reference's):
//   - dM/du and dM/dtheta are the implicit derivatives of the converged return map at Fb,
//     chained through dFb/du = f (G - tr(Fr^-1 G)/3 Fr) and dFb/dtheta = Fb/(3 theta)
//     (the reference's CP_PARAM uses its linearised plasticity_compute_dMdu and dM/dtheta = 0,
//     src/plasticity_model.cc:578-699, include/constitutive_model.h:687-697);
//   - Kuu and Kut also carry Psi : Ld : sym(M^T Fr^T Fr dM), the change of the elastic strain
//     through M, which compute_Kuu/compute_Kut leave out;
//   - Kpu = Kup^T and Kpt = Ktp as the reference sets them (out_loop, :531-555): their terms
//     in Lambda vanish, since det pF(n+1) = 1 holds along the converged return map.
// Checks: every block against central differences of the re-integrated residuals; below
// yield (no slip) the element equals the hyperelastic CM3F element of the same law
// -> test_oracle_cp_three_field.py.
namespace {

int cp_tf_element(const orc_material &mat, const double *gP, const double *gL, int nne, const double *x,
                  const double *u, double theta, double P, double dt, bool integrate, const double *pFn_q,
                  const double *Fn_q, const double *gn_q, const double *lamn_q, double *Fb_q, double *pF1_q,
                  double *g1_q, double *lam1_q, double *tau_q, double *gdot_q, int32_t *it_q, bool want_K,
                  TfElem &E) {
  const int nd = nne * 3;
  std::memset(&E, 0, sizeof(E));
  const double kappa = kappa_of(mat);
  int err = 0;
  for (int ip = 0; ip < nint_of(nne); ++ip) {
    double N[10], dNdX[30], w;
    if (orc_basis(nne, x, ip, N, dNdX, &w)) { ++err; continue; }
    double Fr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int a = 0; a < nne; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) Fr[b * 3 + c] += u[a * 3 + b] * dNdX[a * 3 + c];
    const double J = det3x3(Fr);
    double Fi[9];
    if (!(J > 0.0) || !(theta > 0.0) || inv3x3(Fr, Fi)) { ++err; continue; }
    const double fac = std::cbrt(theta / J);
    double *Fb = Fb_q + ip * 9, *pF1 = pF1_q + ip * 9;
    const double *pFn = pFn_q + ip * 9;
    if (integrate) {
      for (int i = 0; i < 9; ++i) Fb[i] = fac * Fr[i];
      double S0[9], t12[12], gd12[12];
      int its = 0;
      const int bad = orc_cp_update(&mat, gP, gL, Fb, Fn_q + ip * 9, pFn, gn_q[ip], lamn_q[ip], dt, nullptr, 1, pF1,
                                    g1_q + ip, lam1_q + ip, S0, t12, gd12, nullptr, &its);
      if (it_q) it_q[ip] = its;
      if (bad) { ++err; continue; }
      for (int a = 0; a < 12; ++a) {
        if (tau_q) tau_q[ip * 12 + a] = t12[a];
        if (gdot_q) gdot_q[ip * 12 + a] = gd12[a];
      }
    }
    // the stored converged state: M = pF(n+1)^-1 and dM/dFb = pF(n)^-1 dMloc/dFb
    double Me[9], pFni[9], Mloc[9], x0[11], dMdF[81], S1[9], t1[12], gd1[12], pF2[9], g2, l2;
    if (inv3x3(pF1, Me) || inv3x3(pFn, pFni)) { ++err; continue; }
    const double pJ = det3x3(pF1), JM = det3x3(Me);
    matmul(pFn, Me, Mloc);
    for (int i = 0; i < 9; ++i) x0[i] = Mloc[i];
    x0[9] = lam1_q[ip];
    x0[10] = g1_q[ip];
    if (orc_cp_update(&mat, gP, gL, Fb, nullptr, pFn, gn_q[ip], lamn_q[ip], dt, x0, 0, pF2, &g2, &l2, S1, t1, gd1,
                      dMdF, nullptr)) { ++err; continue; }
    double dMedF[81];  // [Me comp][Fb comp]
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 9; ++c) {
          double s = 0.0;
          for (int m = 0; m < 3; ++m) s += pFni[i * 3 + m] * dMdF[(m * 3 + j) * 9 + c];
          dMedF[(i * 3 + j) * 9 + c] = s;
        }
    // deviatoric elasticity at eF = Fr M, volumetric terms at theta_e = theta JM
    double eF[9], eC[9], Sd[9], Ld[81], Cr[9];
    matmul(Fr, Me, eF);
    matmul_tn(eF, eF, eC);
    matmul_tn(Fr, Fr, Cr);
    dev_stress_mr(eC, mat.m10, mat.m01, Sd);
    mat_stiff_mr(eC, mat.m10, mat.m01, Ld);
    double dU, ddU, U;
    vol_pot(mat.volPotFlag, theta * JM, &dU, &ddU, &U);
    dU *= kappa;
    ddU *= kappa;
    // per DOF k: G_k, Phi_k = sym(G^T Fr), Psi_k = M^T Phi M, tr_k = tr(G Fr^-1), dM_k,
    // Lambda_k, X_k = Psi_k + sym(M^T Cr dM_k) (the change of the elastic strain), LX = Ld : X
    double G[30][9], Phi[30][9], Psi[30][9], trk[30], dM[30][9], Lam[30], LX[30][9];
    auto sym_MtCdM = [&](const double *dMx, double *out) {  // sym(M^T Cr dM)
      double T[9], V[9];
      matmul_tn(Me, Cr, T);
      matmul(T, dMx, V);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = 0.5 * (V[i * 3 + j] + V[j * 3 + i]);
    };
    auto L_on = [&](const double *X, double *out) {
      for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (int k = 0; k < 9; ++k) s += Ld[i * 9 + k] * X[k];
        out[i] = s;
      }
    };
    auto tr_MidM = [&](const double *dMx) {  // tr(M^-1 dM) = tr(pF(n+1) dM)
      double s = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += pF1[i * 3 + j] * dMx[j * 3 + i];
      return s;
    };
    for (int k = 0; k < nd; ++k) {
      tf_dof_grad(dNdX, k, G[k]);
      double A[9], T[9];
      matmul_tn(G[k], Fr, A);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Phi[k][i * 3 + j] = 0.5 * (A[i * 3 + j] + A[j * 3 + i]);
      matmul_tn(Me, Phi[k], T);
      matmul(T, Me, Psi[k]);
      double tr = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tr += G[k][j * 3 + i] * Fi[i * 3 + j];
      trk[k] = tr;
      double dFb[9];
      for (int i = 0; i < 9; ++i) dFb[i] = fac * (G[k][i] - tr / 3.0 * Fr[i]);
      for (int r = 0; r < 9; ++r) {
        double s = 0.0;
        for (int c = 0; c < 9; ++c) s += dMedF[r * 9 + c] * dFb[c];
        dM[k][r] = s;
      }
      Lam[k] = tr_MidM(dM[k]);
      double X[9], Y[9];
      sym_MtCdM(dM[k], Y);
      for (int i = 0; i < 9; ++i) X[i] = Psi[k][i] + Y[i];
      L_on(X, LX[k]);
    }
    // theta: dFb/dtheta = Fb / (3 theta) at the current u
    double dMt[9], Xt[9], LXt[9];
    for (int r = 0; r < 9; ++r) {
      double s = 0.0;
      for (int c = 0; c < 9; ++c) s += dMedF[r * 9 + c] * fac * Fr[c] / (3.0 * theta);
      dMt[r] = s;
    }
    const double Lamt = tr_MidM(dMt);
    sym_MtCdM(dMt, Xt);
    L_on(Xt, LXt);
    const double cw = pJ * w;  // detJxW / Jn
    const double vt = dU * JM - P;
    for (int k = 0; k < nd; ++k) {
      double PsiS = 0.0, PsiLXt = 0.0, DtS = 0.0;
      for (int i = 0; i < 9; ++i) {
        PsiS += Psi[k][i] * Sd[i];
        PsiLXt += Psi[k][i] * LXt[i];
      }
      {  // 2 sym(M^T Phi_k dM_t) : Sd
        double T[9], V[9];
        matmul_tn(Me, Phi[k], T);
        matmul(T, dMt, V);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) DtS += (V[i * 3 + j] + V[j * 3 + i]) * Sd[i * 3 + j];
      }
      const double rk = PsiS + P * J * trk[k];
      E.Ru[k] += cw * rk;                                          // compute_Ru
      E.Kup[k] += cw * J * trk[k];                                 // compute_Kup
      E.Kut[k] += cw * (DtS + PsiLXt - Lamt * rk);                 // compute_Kut (+ Ld term)
      E.Ktu[k] += cw * (JM * (ddU * theta * JM + dU) * Lam[k] - Lam[k] * vt);  // compute_Ktu
    }
    E.Rp += cw * (J - theta);                                      // compute_Rp
    E.Rt += cw * vt;                                               // compute_Rt
    E.Ktt += cw * (JM * dU * Lamt + JM * JM * ddU - Lamt * vt);    // compute_Ktt
    E.Ktp -= cw;                                                   // compute_Ktp
    if (!want_K) continue;
    double Sh[9];  // M Sd M^T: sym(G_k^T G_l) : Sh = M^T sym(G_k^T G_l) M : Sd
    {
      double T[9];
      matmul(Me, Sd, T);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int m = 0; m < 3; ++m) s += T[i * 3 + m] * Me[j * 3 + m];
          Sh[i * 3 + j] = s;
        }
    }
    for (int k = 0; k < nd; ++k) {
      double PhiMS[9];  // Phi_k M Sd: 2 sym(dM^T Phi_k M) : Sd = 2 (Phi_k M Sd) : dM
      {
        double T[9];
        matmul(Phi[k], Me, T);
        matmul(T, Sd, PhiMS);
      }
      double PsiS = 0.0;
      for (int i = 0; i < 9; ++i) PsiS += Psi[k][i] * Sd[i];
      const double rk = PsiS + P * J * trk[k];
      for (int l = 0; l < nd; ++l) {
        double PLX = 0.0, GGS = 0.0, dMS = 0.0, tr2 = 0.0;
        for (int i = 0; i < 9; ++i) {
          PLX += Psi[k][i] * LX[l][i];
          dMS += 2.0 * PhiMS[i] * dM[l][i];
        }
        double GG[9];
        matmul_tn(G[k], G[l], GG);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) GGS += 0.5 * (GG[i * 3 + j] + GG[j * 3 + i]) * Sh[i * 3 + j];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk)
              for (int ll = 0; ll < 3; ++ll) tr2 += G[k][j * 3 + i] * Fi[i * 3 + kk] * G[l][kk * 3 + ll] * Fi[ll * 3 + j];
        const double pJrJn = P * J * (trk[k] * trk[l] - tr2);
        E.Kuu[k * nd + l] += cw * (PLX + pJrJn + GGS + dMS - Lam[l] * rk);  // compute_Kuu (+ Ld term)
      }
    }
  }
  return err;
}

}  // namespace

extern "C" {

// One CM3F crystal-plasticity element for tests: state at n per QP ([nint][9] pF(n), F(n);
// [nint] g(n), lambda(n)), integration at (u, theta, P); out the n+1 state ([nint][9] Fb,
// pF(n+1); g, lambda; [nint][12] tau, gamma_dot), the blocks (Kuu [3nne][3nne], Ru, Kup, Kut,
// Ktu; sc = Rp, Rt, Ktt, Ktp) and the condensed K, f (any output may be null).
int orc_cp_tf_element(const orc_material *mat, const double *gP, const double *gL, int nne, const double *x,
                      const double *u, double theta, double P, double dt, const double *pFn, const double *Fn,
                      const double *g_n, const double *lam_n, double *Fb, double *pF1, double *g1, double *lam1,
                      double *tau, double *gdot, double *Kuu, double *Ru, double *Kup, double *Kut, double *Ktu,
                      double *sc, double *Kc, double *fc) {
  TfElem E;
  const int err = cp_tf_element(*mat, gP, gL, nne, x, u, theta, P, dt, true, pFn, Fn, g_n, lam_n, Fb, pF1, g1, lam1,
                                tau, gdot, nullptr, Kuu || Kc, E);
  const int nd = 3 * nne;
  if (Kuu) std::memcpy(Kuu, E.Kuu, sizeof(double) * nd * nd);
  if (Ru) std::memcpy(Ru, E.Ru, sizeof(double) * nd);
  if (Kup) std::memcpy(Kup, E.Kup, sizeof(double) * nd);
  if (Kut) std::memcpy(Kut, E.Kut, sizeof(double) * nd);
  if (Ktu) std::memcpy(Ktu, E.Ktu, sizeof(double) * nd);
  if (sc) { sc[0] = E.Rp; sc[1] = E.Rt; sc[2] = E.Ktt; sc[3] = E.Ktp; }
  if (Kc || fc) tf_condense(nd, E, Kc, fc);
  return err;
}

// fd_residuals_MP + stiffmat_fd_MP of the CM3F analysis for crystal plasticity: theta[e],
// P[e] per element; state arrays [comp][n_qp] in mesh QP order as orc_assemble_cp (with
// run_integration the return map runs at Fb and F_np1 = Fb is stored, otherwise the stored
// state is used); grain tables as orc_assemble_cp.
int orc_assemble_cp_tf(const double *X, int64_t n_elem, const int64_t *elem_ptr, const int64_t *elem_conn,
                       const int32_t *elem_mat, const orc_material *mats, const int32_t *elem_grain,
                       const double *grain_P, const double *grain_L, const int64_t *id, int64_t ndof,
                       const int32_t *Ap, const int32_t *Ai, const double *r, const double *d_r, const double *defl,
                       const double *defl_d, double dt, const double *theta, const double *P, int run_integration,
                       int64_t n_qp, double *F_np1, const double *F_n, const double *pF_n, double *pF_np1,
                       const double *g_n, double *g_np1, const double *lam_n, double *lam_np1, double *tau_out,
                       double *gdot_out, int32_t *iters_out, double *f, double *vals) {
  if (f) std::memset(f, 0, sizeof(double) * ndof);
  if (vals) std::memset(vals, 0, sizeof(double) * Ap[ndof]);
  std::vector<int64_t> qoff(n_elem + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) qoff[e + 1] = qoff[e] + nint_of(int(elem_ptr[e + 1] - elem_ptr[e]));
  return assemble_by_colour(n_elem, elem_ptr, elem_conn, [&](int64_t e) -> int {
    int err = 0;
    const int64_t qp0 = qoff[e];
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]), nd = 3 * nne, nint = nint_of(nne);
    double x[30], u[30];
    int64_t cn[30];
    tf_gather(X, elem_conn + elem_ptr[e], nne, id, r, d_r, defl, defl_d, x, u, cn);
    // the element's QP state, gathered from the [comp][n_qp] arrays
    double pFn[72], Fn[72], Fb[72], pF1[72], gn[8], ln[8], g1[8], l1[8], tq[96], gq[96];
    int32_t its[8] = {0};
    for (int ip = 0; ip < nint; ++ip) {
      const int64_t q = qp0 + ip;
      for (int i = 0; i < 9; ++i) {
        pFn[ip * 9 + i] = pF_n[i * n_qp + q];
        Fn[ip * 9 + i] = F_n ? F_n[i * n_qp + q] : ((i % 4 == 0) ? 1.0 : 0.0);
        Fb[ip * 9 + i] = F_np1[i * n_qp + q];
        pF1[ip * 9 + i] = pF_np1[i * n_qp + q];
      }
      gn[ip] = g_n[q];
      ln[ip] = lam_n[q];
      g1[ip] = g_np1[q];
      l1[ip] = lam_np1[q];
    }
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    const double *gP = grain_P + int64_t(elem_grain[e]) * 108, *gL = grain_L + int64_t(elem_grain[e]) * 81;
    TfElem E;
    err += cp_tf_element(mat, gP, gL, nne, x, u, theta[e], P[e], dt, run_integration != 0, pFn, Fn, gn, ln, Fb, pF1,
                         g1, l1, tq, gq, its, vals != nullptr, E);
    if (run_integration)
      for (int ip = 0; ip < nint; ++ip) {
        const int64_t q = qp0 + ip;
        for (int i = 0; i < 9; ++i) {
          F_np1[i * n_qp + q] = Fb[ip * 9 + i];
          pF_np1[i * n_qp + q] = pF1[ip * 9 + i];
        }
        g_np1[q] = g1[ip];
        lam_np1[q] = l1[ip];
        for (int a = 0; a < 12; ++a) {
          if (tau_out) tau_out[a * n_qp + q] = tq[ip * 12 + a];
          if (gdot_out) gdot_out[a * n_qp + q] = gq[ip * 12 + a];
        }
        if (iters_out) iters_out[q] = its[ip];
      }
    std::vector<double> K(size_t(nd) * nd);
    double fe[30];
    tf_condense(nd, E, K.data(), fe);
    if (vals)
      for (int a = 0; a < nd; ++a)
        if (!std::isnormal(K[a * nd + a])) ++err;
    for (int i = 0; i < nd; ++i) {
      const int64_t row = cn[i] - 1;
      if (row < 0) continue;
      if (f) f[row] += fe[i];
      if (vals)
        for (int j = 0; j < nd; ++j) {
          const int64_t col = cn[j] - 1;
          if (col < 0) continue;
          const int32_t *p = Ai + Ap[row], *en = Ai + Ap[row + 1];
          while (p < en && *p != col) ++p;
          if (p == en) return -1;
          vals[p - Ai] += K[i * nd + j];
        }
    }
    return err;
  });
}

// compute_d_theta_dP for the CM3F crystal-plasticity element at the stored state: the
// increments of theta and P that go with du (prescribed DOFs zero).
int orc_cp_tf_update(const double *X, int64_t n_elem, const int64_t *elem_ptr, const int64_t *elem_conn,
                     const int32_t *elem_mat, const orc_material *mats, const int32_t *elem_grain,
                     const double *grain_P, const double *grain_L, const int64_t *id, const double *r,
                     const double *d_r, const double *defl, const double *defl_d, double dt, const double *theta,
                     const double *P, int64_t n_qp, double *F_np1, const double *pF_n, double *pF_np1,
                     const double *g_n, double *g_np1, const double *lam_n, double *lam_np1, const double *du,
                     double *ddV, double *ddP) {
  int err = 0;
  int64_t qp0 = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    const int nne = int(elem_ptr[e + 1] - elem_ptr[e]), nd = 3 * nne, nint = nint_of(nne);
    double x[30], u[30];
    int64_t cn[30];
    tf_gather(X, elem_conn + elem_ptr[e], nne, id, r, d_r, defl, defl_d, x, u, cn);
    double pFn[72], Fn[72], Fb[72], pF1[72], gn[8], ln[8], g1[8], l1[8];
    for (int ip = 0; ip < nint; ++ip) {
      const int64_t q = qp0 + ip;
      for (int i = 0; i < 9; ++i) {
        pFn[ip * 9 + i] = pF_n[i * n_qp + q];
        Fn[ip * 9 + i] = (i % 4 == 0) ? 1.0 : 0.0;
        Fb[ip * 9 + i] = F_np1[i * n_qp + q];
        pF1[ip * 9 + i] = pF_np1[i * n_qp + q];
      }
      gn[ip] = g_n[q];
      ln[ip] = lam_n[q];
      g1[ip] = g_np1[q];
      l1[ip] = lam_np1[q];
    }
    const orc_material &mat = mats[elem_mat ? elem_mat[e] : 0];
    const double *gP = grain_P + int64_t(elem_grain[e]) * 108, *gL = grain_L + int64_t(elem_grain[e]) * 81;
    TfElem E;
    err += cp_tf_element(mat, gP, gL, nne, x, u, theta[e], P[e], dt, false, pFn, Fn, gn, ln, Fb, pF1, g1, l1, nullptr,
                         nullptr, nullptr, false, E);
    const double Kpt = E.Ktp;
    double Kpu_du = 0.0, Ktu_du = 0.0;
    for (int k = 0; k < nd; ++k) {
      const double d = cn[k] > 0 ? du[cn[k] - 1] : 0.0;
      Kpu_du += E.Kup[k] * d;
      Ktu_du += E.Ktu[k] * d;
    }
    const double dth = -(Kpu_du + E.Rp) / Kpt;
    ddV[e] = dth;
    ddP[e] = -(E.Rt + Ktu_du + E.Ktt * dth) / E.Ktp;
    qp0 += nint;
  }
  return err;
}

}  // extern "C"
