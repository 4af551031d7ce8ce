// Element kinematics shared by the device kernels: FEMLIB quadrature and shape-function
// derivatives, 3x3 determinant and inverse (FP64).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pgf {
namespace dev {

// volumetric potential derivatives (src/new_potentials.cc:575-616)
__device__ __forceinline__ void vol_pot(int flag, double J, double &dU, double &d2U) {
  if (flag == 3) {
    dU = (-1 + J) / (2. * J) + log(J) * .5;
    d2U = 1. / J - (-1 + J) / (2. * J * J);
  } else if (flag == 99) {
    dU = J - 1;
    d2U = 1.0;
  } else {
    const double ex = exp(J - 1.0), rJ = 1.0 / J;
    dU = (ex - rJ) * .5;
    d2U = (ex + rJ * rJ) * .5;
  }
}

// U(J) of the same potentials (src/new_potentials.cc:553-573)
__device__ __forceinline__ double vol_U(int flag, double J) {
  if (flag == 3) return 0.5 * (J - 1.0) * log(J);
  if (flag == 99) return 0.5 * (J - 1.0) * (J - 1.0);
  return 0.5 * (exp(J - 1.0) - log(J) - 1.0);
}

__host__ __device__ constexpr int nqp_of(int nne) { return nne == 4 ? 1 : (nne == 10 ? 4 : 8); }

// FEMLIB quadrature (FEMLIB/src/femlib.cc:544-588), truncated constants kept.
__device__ __forceinline__ void int_point(int nne, int ip, double &k, double &e, double &z,
                                          double &w) {
  if (nne == 4) {
    k = e = z = 1.0 / 4.0;
    w = 1.0 / 6.0;
  } else if (nne == 10) {
    const double A = 0.58541020, B = 0.13819660;
    k = (ip == 0) ? A : B;
    e = (ip == 1) ? A : B;
    z = (ip == 2) ? A : B;
    w = (1.0 / 4.0) / 6.0;
  } else {
    const double g = 0.57735026918962584;
    k = (ip / 4) ? g : -g;
    e = ((ip / 2) % 2) ? g : -g;
    z = (ip % 2) ? g : -g;
    w = 1.0;
  }
}

// dN/d(ksi,eta,zet), FEMLIB/src/femlib.cc:214-340; dN[a*3+c]
template <int NNE>
__device__ __forceinline__ void shape_dN(double k, double e, double z, double *dN) {
  if constexpr (NNE == 4) {
    dN[0] = -1.0; dN[1] = -1.0; dN[2] = -1.0;
    dN[3] = 1.0;  dN[4] = 0.0;  dN[5] = 0.0;
    dN[6] = 0.0;  dN[7] = 1.0;  dN[8] = 0.0;
    dN[9] = 0.0;  dN[10] = 0.0; dN[11] = 1.0;
  } else if constexpr (NNE == 10) {
    dN[0] = 4.0 * (k + e + z) - 3.0;
    dN[1] = 4.0 * (e + k + z) - 3.0;
    dN[2] = 4.0 * (z + k + e) - 3.0;
    dN[3] = 4.0 * k - 1.0; dN[4] = 0.0; dN[5] = 0.0;
    dN[6] = 0.0; dN[7] = 4.0 * e - 1.0; dN[8] = 0.0;
    dN[9] = 0.0; dN[10] = 0.0; dN[11] = 4.0 * z - 1.0;
    dN[12] = 4.0 * (-2.0 * k - e - z + 1.0); dN[13] = -4.0 * k; dN[14] = -4.0 * k;
    dN[15] = 4.0 * e; dN[16] = 4.0 * k; dN[17] = 0.0;
    dN[18] = -4.0 * e; dN[19] = 4.0 * (-k - 2.0 * e - z + 1.0); dN[20] = -4.0 * e;
    dN[21] = -4.0 * z; dN[22] = -4.0 * z; dN[23] = 4.0 * (-k - e - 2.0 * z + 1.0);
    dN[24] = 4.0 * z; dN[25] = 0.0; dN[26] = 4.0 * k;
    dN[27] = 0.0; dN[28] = 4.0 * z; dN[29] = 4.0 * e;
  } else {
    const double o = 0.125;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double sk = ((a + 1) & 2) ? 1.0 : -1.0;  // nodes 1,2,5,6: +ksi
      const double se = (a & 2) ? 1.0 : -1.0;        // nodes 2,3,6,7: +eta
      const double sz = (a & 4) ? 1.0 : -1.0;        // nodes 4..7: +zet
      dN[a * 3 + 0] = sk * o * (1.0 + se * e) * (1.0 + sz * z);
      dN[a * 3 + 1] = se * o * (1.0 + sk * k) * (1.0 + sz * z);
      dN[a * 3 + 2] = sz * o * (1.0 + sk * k) * (1.0 + se * e);
    }
  }
}

__device__ __forceinline__ double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) + m[1] * (m[5] * m[6] - m[3] * m[8]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// inv3x3 with the reference's singularity threshold (src/utils.cc:542-562)
__device__ __forceinline__ bool inv3(const double *m, double *mi) {
  const double A = det3(m);
  if (!(fabs(A) >= 1.0e-15)) return false;
  const double r = 1.0 / A;
  mi[0] = r * (m[4] * m[8] - m[5] * m[7]);
  mi[1] = -r * (m[1] * m[8] - m[2] * m[7]);
  mi[2] = r * (m[1] * m[5] - m[2] * m[4]);
  mi[3] = -r * (m[3] * m[8] - m[5] * m[6]);
  mi[4] = r * (m[0] * m[8] - m[2] * m[6]);
  mi[5] = -r * (m[0] * m[5] - m[2] * m[3]);
  mi[6] = r * (m[3] * m[7] - m[4] * m[6]);
  mi[7] = -r * (m[0] * m[7] - m[1] * m[6]);
  mi[8] = r * (m[0] * m[4] - m[1] * m[3]);
  return true;
}

}  // namespace dev
}  // namespace pgf
