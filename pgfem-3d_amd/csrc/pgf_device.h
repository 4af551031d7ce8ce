// Device-side data layout and QP math of the element hot path (gfx950, FP64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pgfem3d_amd.h"

namespace pgf {

// Voigt order of symmetric tensors: xx, yy, zz, xy, yz, xz.
__host__ __device__ constexpr int voigt_i(int I) { return I < 3 ? I : (I == 3 ? 0 : (I == 4 ? 1 : 0)); }
__host__ __device__ constexpr int voigt_j(int I) { return I < 3 ? I : (I == 3 ? 1 : 2); }

// Device copy of the per-material constants used by the kernels.
struct DevMaterial {
  int32_t model;
  int32_t volPotFlag;
  double G;       // shear modulus of the deviatoric part (neo-Hookean: 2*m10)
  double kappa;   // bulk modulus E/(3(1-2nu))
  double k0, hp;  // J2
};

// Everything one element launch reads (device pointers).
struct ElemArgs {
  const double *X;          // [n_nodes*3]
  const int32_t *dofid;     // [n_nodes*3] reference DOF ids
  const int32_t *row_start; // [n_nodes] Ap of the node's first free DOF, -1 if none
  const int32_t *rowlen;    // [n_nodes]
  const int32_t *conn;      // [ne*nne] block connectivity, device element order
  const int16_t *off;       // [ne*nne*nne] scatter offsets (>=0 add, <=-2 first touch, -1 none)
  const int32_t *mat;       // [ne] material index
  const DevMaterial *mats;
  int64_t qp_base;          // first QP of the block in device QP order
  int64_t n_qp;             // total QPs (stride of the SoA state)
  // nodal input (def_elem_total)
  const double *r, *d_r, *defl, *defl_d;
  // state, SoA [comp][n_qp]
  double *F_np1;            // 9 comps
  const double *P_n;        // plastic n   (J2: Cp^-1, 6 comps)
  double *P_np1;            // plastic n+1
  const double *Sc_n;       // scalars n   (J2: ep, gam)
  double *Sc_np1;           // scalars n+1
  double *cauchy;           // 6 comps (written by integration)
  // outputs
  double *f;                // [ndof]
  double *vals;             // [nnz]
  int32_t *err;             // failure counter
};

// Mode bits of the element kernel.
enum : int { MODE_INTEGRATE = 1, MODE_RESIDUAL = 2, MODE_TANGENT = 4 };

void launch_elem(int nne, int mode, const ElemArgs &a, int64_t e_begin, int64_t e_end,
                 hipStream_t s);
void launch_fill(double *p, int64_t n, double v, hipStream_t s);
void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s);

}  // namespace pgf
