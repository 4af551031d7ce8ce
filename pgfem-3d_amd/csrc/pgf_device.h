// Device-side data layout and QP math of the element hot path (gfx950, FP64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pgfem3d_amd.h"

namespace pgf {

// Voigt order of symmetric tensors: xx, yy, zz, xy, yz, xz.
__host__ __device__ constexpr int voigt_i(int I) { return I < 3 ? I : (I == 3 ? 0 : (I == 4 ? 1 : 0)); }
__host__ __device__ constexpr int voigt_j(int I) { return I < 3 ? I : (I == 3 ? 1 : 2); }

// Workgroup barrier that orders LDS only: global stores still in flight keep draining
// across it (__syncthreads would first wait for every outstanding global access).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Doubles of one staging block (gather assembly): the 3 rows of an element node, each the
// element's K row (3 nne) and residual entry, padded to whole 128-byte lines so every line
// of a block is written whole by the element that owns it.
__host__ __device__ constexpr int stage_block(int nne) { return (3 * (3 * nne + 1) + 15) & ~15; }

// Device copy of the per-material constants used by the kernels.
struct DevMaterial {
  int32_t model;
  int32_t volPotFlag;
  int32_t devPotFlag;  // hyperelastic deviatoric law: 1 neo-Hookean, 2 linear in C
  int32_t pad;
  double G;       // shear modulus of the deviatoric part (neo-Hookean: 2*m10; linear law: G)
  double kappa;   // bulk modulus E/(3(1-2nu))
  double k0, hp;  // J2
  // J2 continuum damage (viscosity mu, cap wmax, Weibull p1, p2, threshold Yin); mu <= 0: off
  double mu, wmax, p1, p2, Yin;
};

// Everything one element launch reads (device pointers).
struct ElemArgs {
  const double *X;          // [n_nodes*3]
  const int32_t *dofid;     // [n_nodes*3] reference DOF ids
  const int32_t *row_start; // [n_nodes] Ap of the node's first free DOF, -1 if none
  const int32_t *rowlen;    // [n_nodes]
  const int32_t *conn;      // [ne*nne] block connectivity, device element order
  const int32_t *edof;      // [ne*3nne] the element DOFs' ids (dofid of each node component)
  const int16_t *off;       // [ne*nne*nne] scatter offsets (>=0 add, <=-2 first touch, -1 none)
  const int32_t *elist;     // coloured mode: block elements in colour order (launch ranges index it)
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
  const double *Sc_n;       // scalars n   (J2: ep, gam; with damage also w, X, H, damaged)
  double *Sc_np1;           // scalars n+1
  double dt;                // time step (J2 damage viscosity)
  // outputs: coloured mode adds into f / vals through off; gather mode (stage != null)
  // writes row (a, b) of element e to stage[soff[e*nne + a] + b*(3nne+1) ...], residual last
  double *f;                // [ndof]
  double *vals;             // [nnz]
  const int64_t *soff;      // [ne*nne] staging offsets of the element's node blocks
  double *stage;
  int32_t *err;             // failure counter
  // three-field (CM3F) blocks, per block element (block-local index): theta = V(n+1) + dV and
  // P = P(n+1) + dP, and the element data of the last residual pass for the update (kTfBuf
  // doubles: kup[3nne], fp, ft, Ktt, V = -Ktp, then from kTfKtu on ktu[3nne], which only the
  // crystal-plasticity element fills).  Null on a displacement-only handle.
  const double *tf_V, *tf_dV, *tf_P, *tf_dP;
  double *tf_buf;
};
constexpr int kTfKtu = 34;
constexpr int kTfBuf = 64;
// per-element three-field slots of the handle: V(n+1), V(n), V(n-1), P(n+1), P(n), P(n-1), dV,
// ddV, dP, ddP, the full ddV and ddP of the last update, the dV and dP its trials start from
constexpr int kTfSlots = 14;

// Element update and initial conditions of the three-field formulation, one thread per
// element of a block (block-local arrays).
struct TfArgs {
  const double *X;
  const int32_t *conn, *edof, *mat;
  const DevMaterial *mats;
  int64_t ne;
  const double *buf;         // [ne][kTfBuf] from the last residual pass
  const double *du;          // free-DOF increments (the linear solve's solution)
  double *ddV, *ddP;         // out: the element's volume and pressure increments
  const double *r, *d_r, *defl, *defl_d;  // initial conditions: nodal input
  double *V, *P;             // out: volume-averaged J and kappa U'(J)
};

// Mode bits of the element kernel.
// MODE_DEFL_LOAD (with MODE_TANGENT alone): the element tangent at (r + d_r, defl) times the
// element's prescribed increments defl_d, summed into f in place of the residual; no K output.
enum : int { MODE_INTEGRATE = 1, MODE_RESIDUAL = 2, MODE_TANGENT = 4, MODE_DEFL_LOAD = 8 };

// ---------------------------------------------------------------- crystal plasticity
// Per-grain tables: Schmid tensors P_a = s_a (x) m_a rotated to the sample frame (12 x 9,
// row-major) and the rotated cubic stiffness in Voigt form Lv[I][J] (I, J in xx,yy,zz,xy,yz,
// xz) such that S_I = sum_J Lv[I][J] E_J with E_J the tensor components of a symmetric E.
struct DevGrain {
  double P[12][9];
  double Lv[36];
  double s[12][3];  // slip directions in the sample frame; P_a = s_a (x) m[a / 3]
  double m[4][3];   // the four slip-plane normals (systems 3k..3k+2 share plane k)
  double R[9];                // crystal -> sample rotation
  double cub[3];              // cubic constants C11 - C12, C12, 2 C44 (crystal frame)
};

// law 0: cubic St Venant-Kirchhoff through the grain's Lv; law 1: neo-Hookean deviatoric part
// (m10) and the volumetric potential vol_flag (kappa) on the elastic part eF.
struct CPMaterial {
  double gd0, m, h0, g0, gs0, gds, w, tol;
  double m10, kappa;
  int32_t maxit, law, vol_flag;
  int32_t ipow;  // 1/m - 1 when it is a whole number in [0, 64] (binary powering), else -1
};

// Arguments shared by the crystal-plasticity QP kernels (device pointers, SoA [comp][n_qp]).
struct CPArgs {
  // kinematics input
  const double *X;
  const int32_t *dofid;
  const int32_t *conn;      // block connectivity
  const int32_t *qgrain;    // [nq] grain id of the block's QPs
  const DevGrain *grains;
  const CPMaterial *cpmats;
  int64_t qp_base, n_qp, ne, nq;  // block QP offset, total QPs, block elements, block QPs
  const double *r, *d_r, *defl, *defl_d;
  double dt;
  // state
  double *F_np1;
  const double *pF_n;
  double *pF_np1;
  const double *g_n, *lam_n;  // Sc_n comps 0 and 1
  double *Sc_np1;             // g, lambda at n+1
  const double *F_n;          // 9 comps (plain start: plastic predictor)
  double *tau, *gdot;         // 12 comps at n+1
  const double *tau_n;        // 12 comps at n (learned-start features)
  int32_t *iters;             // [n_qp]
  const double *warm;         // optional start (M 9, lambda, g) [11][n_qp]
  // per-QP records (qrec): the plain start, then the tangent pass's output for the element
  // kernel, Me 9, S 6, pJ 1, D = dMe/dF 81 by columns (field 16 + 9 c + r holds D[r][c]);
  // indexed by device QP
  double *qpbuf;
  int32_t *err;
  unsigned long long *qnext;  // global work queue of the local Newton (cp_chunk = 1), else null
  int32_t law;                // cpmats[0].law, read on the host to pick the kernels
  // three-field (CM3F) blocks: per block element theta = tf_V + tf_dV; the kinematics then
  // stores the model's F(n+1) = (theta/J)^(1/3) F.  Null otherwise.
  const double *tf_V, *tf_dV;
};
constexpr int kCPBuf = 97;  // doubles per QP in qpbuf
// qpbuf holds the records of 8 QPs at a time with the field outermost: field f of device QP q
// is at qpbuf[qrec(q) + f * kRecStride].  A wave's store of one field over its 64 QPs then
// writes eight 64-byte runs, and the next field's store the other halves of the same 128-byte
// lines, where a 776-byte row per QP has each store touch 64 rows, and lines leave the L2
// partly written.  (Groups of 64 QPs make each store one 512-byte run, but the field offsets
// then exceed the stores' immediate offsets and the address registers spill.)  The buffer
// holds ceil(n_qp / 8) * 8 records.
constexpr int kRecStride = 8;
__host__ __device__ constexpr int64_t qrec(int64_t q) {  // q >= 0
  return int64_t((uint64_t(q) / kRecStride) * (uint64_t(kCPBuf) * kRecStride) + uint64_t(q) % kRecStride);
}

// Output variables of one block at one time level (constitutive_model_update_output_variables,
// src/constitutive_model.cc:1677-1881): per QP the PK2 stress S and the Green-Lagrange strain E,
// per element the volume averages of the Cauchy stress and of the Almansi strain (shears
// doubled), all in the reference's component order XX, YY, ZZ, YZ, XZ, XY.
struct OutArgs {
  const double *X;
  const int32_t *conn, *mat;        // block connectivity, material per block element
  const DevMaterial *mats;
  int model;                        // the block's model
  const int32_t *qgrain;            // crystal plasticity: grain per block QP
  const DevGrain *grains;
  const CPMaterial *cpmats;
  int64_t qp_base, n_qp, ne;
  const double *F, *P, *Sc;         // state at the level: F 9, plastic (Cp^-1 6 / pF 9), scalars
  double *S, *E;                    // [6][n_qp] device QP order
  double *sig, *eps;                // [6][ne] block elements
};
void launch_output(int nne, const OutArgs &o, hipStream_t s);

void launch_cp_kinematics(int nne, const CPArgs &a, hipStream_t s);
void launch_cp_newton(const CPArgs &a, int64_t chunk, hipStream_t s);
void launch_cp_tangent(const CPArgs &a, bool with_D, hipStream_t s);

// Element kernel of the crystal-plasticity blocks: residual and / or tangent from qpbuf.
struct CPElemArgs {
  const double *X;
  const int32_t *dofid, *row_start, *rowlen, *conn, *grain;
  const int16_t *off;
  const int32_t *elist;  // coloured mode: as ElemArgs
  const DevGrain *grains;
  const CPMaterial *cpmats;
  int32_t law;  // cpmats[0].law
  const double *qpbuf;
  int64_t qp_base;
  const double *r, *d_r, *defl, *defl_d;
  double *f, *vals;
  const int64_t *soff;  // gather mode: as ElemArgs
  double *stage;
  int32_t *err;
  // three-field (CM3F) blocks: as ElemArgs (block-local per element)
  const double *tf_V, *tf_dV, *tf_P, *tf_dP;
  double *tf_buf;
};
void launch_cp_elem(int nne, int mode, const CPElemArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s);
// The CM3F element of crystal plasticity (hyperelastic elastic law), condensed to the
// displacement DOFs; gather (a.stage) or coloured mode as launch_cp_elem.
void launch_cp_tf_elem(int nne, int mode, const CPElemArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s);

// Learned warm start: features / targets of a block's QPs (rows indexed by device QP) and the
// Newton start from predicted slip increments (block-local rows of Y).
void launch_cp_features(const CPArgs &a, const float *gquat, const float *shift, const float *scale, float *X,
                        hipStream_t s);
void launch_cp_targets(const CPArgs &a, float *T, hipStream_t s);
void launch_cp_warm_from_slip(const CPArgs &a, const float *Y, int ldy, const float *out_scale, double *warm,
                              hipStream_t s);
int mlp_forward_device(pgf_mlp m, const float *X, int64_t B, hipStream_t s, const float **Y, int *ldy);

// ---------------------------------------------------------------- interface elements
// Exponential traction-separation law parameters (Needleman_pkey order).
struct DevCohesive {
  double sig_c, chi_c, beta, alpha;
};

struct CohArgs {
  const double *X;
  const int32_t *dofid, *row_start, *rowlen;
  const int32_t *conn;     // [nc][2*nn] block connectivity (- side nodes, then + side)
  const int32_t *mat;      // [nc] property index
  const int16_t *off;      // [nc][2nn][2nn] scatter offsets (same encoding as ElemArgs::off)
  const int32_t *elist;    // coloured mode: as ElemArgs
  const DevCohesive *props;
  int64_t cqp_base, n_cqp; // first QP of the block, total interface QPs (SoA stride)
  const double *r, *d_r, *defl, *defl_d;
  const double *vars_n;    // [2][n_cqp] X_eff_max, t_max at n
  double *vars_np1;        // [2][n_cqp] at n+1 (written by integrating passes)
  double *f, *vals;
  const int64_t *soff;     // gather mode: as ElemArgs (node blocks of 2nn nodes)
  double *stage;
  int32_t *err;
};
void launch_coh_elem(int nn, int mode, const CohArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s);

// variant: 0 plain; 2 the block's J2 points carry continuum damage (state scalars w, X, H,
// damaged); 3 hyperelastic materials that may take the linear deviatoric law (devPotFlag 2).
// (The three-field variant is selected by a.tf_V.)
void launch_elem(int nne, int mode, const ElemArgs &a, int64_t e_begin, int64_t e_end, hipStream_t s,
                 int variant = 0);

// Gather assembly: node n owns the staging blocks [inc_ptr[n], inc_ptr[n+1]) (one per
// incident element, in a fixed order); block k starts at inc_soff[k], comes from an element
// of inc_nne[k] nodes, and pos[inc_poff[k] + w] packs, for the element's node w, the column
// of its first free DOF inside node n's rows (bits 0-27) and its free-DOF mask (bits 28-30;
// 0 = no column).
// inc_rec: per staging block (node-major), bits 0-31 its staging offset in units of 16 doubles,
// bits 32-59 the offset of its element's column positions in pos, bits 60-63 the element's
// node count.
struct GatherArgs {
  const int64_t *inc_ptr;
  const uint64_t *inc_rec;
  const uint32_t *pos;
  const int32_t *dofid, *row_start, *rowlen;
  const double *stage;
  double *f, *vals;
  int64_t n_nodes;
  int32_t maxlen;  // longest row of any node (LDS per wavefront: 3 maxlen doubles)
};
// Wavefronts per workgroup for a given maxlen (0: rows too long for LDS).
int gather_waves(int32_t maxlen);
// waves: nodes per workgroup (capped by the LDS each node needs); batch: staging blocks each
// half-wave keeps in flight (2, 4 or 8)
void launch_gather(int mode, const GatherArgs &g, int waves, int batch, hipStream_t s);
// Jacobi-preconditioned BiCGStab on a CSR matrix; work holds 9 n + 1025 doubles.
int bicgstab(hipStream_t s, const int32_t *Ap, const int32_t *Ai, const double *vals, const double *b, double *x,
             int64_t n, double tol, int maxit, int *iters, double *relres, double *work);

void launch_tf_update(int nne, const TfArgs &t, hipStream_t s);
void launch_tf_initial(int nne, const TfArgs &t, hipStream_t s);
void launch_axpy(double *y, const double *x, double g, int64_t n, hipStream_t s);  // y += g x
void launch_fill(double *p, int64_t n, double v, hipStream_t s);
void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s);
void launch_rate_max(const double *a, int ncomp, int64_t stride, int64_t n, double scale, const double *d1,
                     const double *d0, double scale_d, unsigned long long *out, hipStream_t s);
void launch_copy(double *dst, const double *src, int64_t n, hipStream_t s);

}  // namespace pgf
