// C ABI of the element hot path: device handle, the pass schedule (gather or coloured
// assembly), state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "pgf_device.h"
#include "pgf_internal.h"
#include "pgfem3d_amd.h"

namespace pgf {
namespace {

constexpr int nqp_of(int nne) { return nne == 4 ? 1 : (nne == 10 ? 4 : 8); }

// Elements of one (element type, material model) pair, in device order (mesh order, which
// keeps neighbouring elements together).  Their QPs occupy [qp_base, qp_base + ne*nqp) of
// the state arrays.  The coloured mode launches colour by colour through clist.
struct Block {
  int nne = 0, nqp = 0, model = 0;
  int64_t ne = 0, qp_base = 0;
  int64_t e_base = 0;              // first element of the block in per-element device arrays
  std::vector<int64_t> elems;      // mesh element ids in device order
  std::vector<int32_t> clist;      // block-local element indices sorted by colour
  std::vector<int64_t> color_ptr;  // [n_colors+1] colour segments of clist
  int32_t *d_clist = nullptr;
  int32_t *d_conn = nullptr;
  int32_t *d_edof = nullptr;  // [ne*3nne] DOF ids of the element DOFs
  int16_t *d_off = nullptr;
  int64_t *d_soff = nullptr;       // gather mode: [ne*nne] staging offsets
  int32_t *d_mat = nullptr;
  int32_t *d_egrain = nullptr;  // crystal plasticity: grain per element / per QP (device order)
  int32_t *d_qgrain = nullptr;
};

// Interface elements of one kind (nn nodes per side), in mesh order; colour lists as Block.
struct CohBlock {
  int nn = 0;
  int64_t nc = 0, cqp_base = 0;
  std::vector<int64_t> elems;      // mesh interface-element ids in device order
  std::vector<int32_t> clist;
  std::vector<int64_t> color_ptr;  // [n_ccolors+1]
  int32_t *d_clist = nullptr;
  int32_t *d_conn = nullptr, *d_mat = nullptr;
  int16_t *d_off = nullptr;
  int64_t *d_soff = nullptr;
};

template <class T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return PGF_OK;
  return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)) == hipSuccess ? PGF_OK : PGF_ENOMEM;
}
template <class T>
int upload(T **p, const T *h, size_t n) {
  int rc = dalloc(p, n);
  if (rc) return rc;
  if (n && hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return PGF_EHIP;
  return PGF_OK;
}

}  // namespace

// Unit quaternion (w, x, y, z), w >= 0, of a rotation matrix (Shepperd's method: the
// largest of the four squared components is taken from the trace or a diagonal entry).
void rot_to_quat(const double *R, float *q) {
  const double t = R[0] + R[4] + R[8];
  double w, x, y, z;
  if (t >= R[0] && t >= R[4] && t >= R[8]) {
    w = 0.5 * std::sqrt(1.0 + t);
    x = (R[7] - R[5]) / (4 * w);
    y = (R[2] - R[6]) / (4 * w);
    z = (R[3] - R[1]) / (4 * w);
  } else if (R[0] >= R[4] && R[0] >= R[8]) {
    x = 0.5 * std::sqrt(1.0 + 2 * R[0] - t);
    w = (R[7] - R[5]) / (4 * x);
    y = (R[1] + R[3]) / (4 * x);
    z = (R[2] + R[6]) / (4 * x);
  } else if (R[4] >= R[8]) {
    y = 0.5 * std::sqrt(1.0 + 2 * R[4] - t);
    w = (R[2] - R[6]) / (4 * y);
    x = (R[1] + R[3]) / (4 * y);
    z = (R[5] + R[7]) / (4 * y);
  } else {
    z = 0.5 * std::sqrt(1.0 + 2 * R[8] - t);
    w = (R[3] - R[1]) / (4 * z);
    x = (R[2] + R[6]) / (4 * z);
    y = (R[5] + R[7]) / (4 * z);
  }
  const double sg = (w < 0) ? -1.0 : 1.0;
  q[0] = float(sg * w);
  q[1] = float(sg * x);
  q[2] = float(sg * y);
  q[3] = float(sg * z);
}

// Host tables of one grain: the 12 FCC {111}<110> systems (unit slip direction s and plane
// normal m, Schmid tensor s (x) m) and the cubic stiffness, both rotated by R (crystal ->
// sample frame).
void make_grain(const double *R, double C11, double C12, double C44, DevGrain &g) {
  const double n[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
  const double sd[12][3] = {{0, 1, -1}, {-1, 0, 1}, {1, -1, 0}, {0, 1, -1}, {1, 0, 1}, {1, 1, 0},
                            {0, 1, 1},  {1, 0, -1}, {1, 1, 0},  {0, 1, 1},  {1, 0, 1}, {1, -1, 0}};
  const double rn = 1.0 / std::sqrt(3.0), rs = 1.0 / std::sqrt(2.0);
  for (int a = 0; a < 12; ++a) {
    double s[3], m[3];
    for (int i = 0; i < 3; ++i) {
      s[i] = R[i * 3] * sd[a][0] * rs + R[i * 3 + 1] * sd[a][1] * rs + R[i * 3 + 2] * sd[a][2] * rs;
      m[i] = R[i * 3] * n[a / 3][0] * rn + R[i * 3 + 1] * n[a / 3][1] * rn + R[i * 3 + 2] * n[a / 3][2] * rn;
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) g.P[a][i * 3 + j] = s[i] * m[j];
    for (int i = 0; i < 3; ++i) {
      g.s[a][i] = s[i];
      g.m[a / 3][i] = m[i];  // the same for the plane's three systems
    }
  }
  auto L0 = [&](int i, int j, int k, int l) {
    double v = 0.0;
    if (i == j && k == l) v += (i == k) ? C11 : C12;
    if (i != j && ((i == k && j == l) || (i == l && j == k))) v += C44;
    return v;
  };
  for (int i = 0; i < 9; ++i) g.R[i] = R[i];
  g.cub[0] = C11 - C12;
  g.cub[1] = C12;
  g.cub[2] = 2.0 * C44;
  static const int vi[6] = {0, 1, 2, 0, 1, 0}, vj[6] = {0, 1, 2, 1, 2, 2};
  for (int I = 0; I < 6; ++I)
    for (int J = 0; J < 6; ++J) {
      const int i = vi[I], j = vj[I], k = vi[J], l = vj[J];
      double v = 0.0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
              v += R[i * 3 + a] * R[j * 3 + b] * R[k * 3 + c] * R[l * 3 + d] * L0(a, b, c, d);
      g.Lv[I * 6 + J] = v * (J < 3 ? 1.0 : 2.0);  // tensor-component E: shear pairs count twice
    }
}

}  // namespace pgf

// Staging blocks of pgf_residual_tangent_static: one array per GPU, allocated with the code
// object, never by a call.  Host code reaches it only through the address
// hipGetSymbolAddress gives on each device (static_stage_ptr); the array's name in host
// code is not a device address.
static __device__ __align__(256) double scratch[PGF_STATIC_STAGE_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device of the calling handle),
// looked up once per device.
int static_stage_ptr(int device, double **ptr) {
  if (device < 0 || device >= kMaxDevices) return PGF_ENODEV;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return PGF_EHIP;
    if (cur != device && hipSetDevice(device) != hipSuccess) return PGF_EHIP;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) hipSetDevice(cur);
    if (e != hipSuccess || !p) return PGF_EHIP;
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[device];
  return PGF_OK;
}
}  // namespace

// What a captured pass depends on: a new key re-captures.
struct GraphKey {
  int mode;
  const void *r, *d_r, *defl, *defl_d, *f, *vals, *warm;
  double dt;
  int64_t chunk;
  int32_t gwaves, gbatch;
  bool operator==(const GraphKey &o) const {
    return mode == o.mode && r == o.r && d_r == o.d_r && defl == o.defl && defl_d == o.defl_d && f == o.f &&
           vals == o.vals && warm == o.warm && dt == o.dt && chunk == o.chunk && gwaves == o.gwaves &&
           gbatch == o.gbatch;
  }
};

struct pgf_handle_s {
  int device = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  int64_t n_nodes = 0, ndof = 0, nnz = 0, n_qp = 0, n_elem = 0;
  int32_t n_colors = 0;
  // three-field (CM3F) mode: per element in device order, d_tf holds [kTfSlots][n_elem] = V
  // at n+1, n, n-1; P at n+1, n, n-1; dV, ddV, dP, ddP (ThreeFieldVariables,
  // include/PGFem3D_data_structure.h:130-235); then the last update's full ddV, ddP and the
  // dV, dP its line-search trials start from; d_tfbuf the element data of the last residual
  // pass, [n_elem][kTfBuf]
  bool tf = false;
  double *d_tf = nullptr, *d_tfbuf = nullptr;
  bool has_j2 = false;
  double *d_X = nullptr;
  int32_t *d_dofid = nullptr, *d_row_start = nullptr, *d_rowlen = nullptr;
  pgf::DevMaterial *d_mats = nullptr;
  std::vector<pgf::Block> blocks;
  std::vector<int64_t> qp_mesh;  // device QP -> mesh QP
  // state, SoA [comp][n_qp]; index 0 = n+1, 1 = n, 2 = n-1
  double *d_F[3] = {nullptr, nullptr, nullptr};
  double *d_P[3] = {nullptr, nullptr, nullptr};   // J2: Cp^-1 (6)
  double *d_Sc[3] = {nullptr, nullptr, nullptr};  // J2: ep, gam (+ w, X, H, damaged); CP: g, lambda
  int nsc = 2;       // scalar components per QP of d_Sc
  bool j2_damage = false;  // a J2 material with continuum damage (mu > 0)
  bool he_linear = false;  // a hyperelastic material with the linear deviatoric law (devPotFlag 2)
  int32_t *d_err = nullptr;
  double *d_red = nullptr;  // one scalar of a device reduction (pgf_subdivision_param)
  // crystal plasticity
  bool has_cp = false;
  pgf::CPMaterial *d_cpmat = nullptr;
  pgf::DevGrain *d_grains = nullptr;
  double *d_tau[2] = {nullptr, nullptr}, *d_gdot[2] = {nullptr, nullptr};  // n+1, n
  int32_t *d_iters = nullptr;
  double *d_qpbuf = nullptr;
  const double *warm = nullptr;
  double cp_g0 = 0.0, dt = 1.0;
  int32_t cp_law = 0;  // elastic law of the crystal-plasticity material (pgf_material.cp_elastic)
  int64_t cp_nq = 0;         // crystal-plasticity QPs: device QPs [0, cp_nq)
  int64_t cp_chunk = 256;    // QPs per wave range of the local-Newton work queue (autotuned; 1:
                             // every wave draws from one global queue)
  unsigned long long *d_qnext = nullptr;  // the global queue's counter
  int32_t gather_waves = 4;  // nodes per gather workgroup (autotuned; capped by LDS)
  int32_t gather_batch = 4;  // staging blocks each gather half-wave keeps in flight (autotuned)
  int32_t *d_qgrain_all = nullptr;  // [n_qp] grain of every device QP
  float *d_gquat = nullptr;  // [n_grains][4] orientation quaternions (learned warm start features)
  // learned warm start
  pgf_mlp sur = nullptr;
  float *d_norm = nullptr;     // shift [34], scale [34], out_scale [12]
  float *d_feat = nullptr;     // [n_qp][34]
  double *d_sur_warm = nullptr;  // [11][n_qp]
  std::vector<double> cp_mat_host;  // C11, C12, C44
  // hipGraph of a whole pass
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  GraphKey gkey{};
  int32_t glaunches = 0;
  // linear solve
  int32_t *d_Ap = nullptr, *d_Ai = nullptr;
  double *d_solve_work = nullptr;
  // interface elements
  std::vector<pgf::CohBlock> cblocks;
  int32_t n_ccolors = 0;
  int64_t n_coel = 0, n_cqp = 0;
  std::vector<int64_t> cqp_mesh;  // device interface QP -> mesh interface QP
  double *d_cv[3] = {nullptr, nullptr, nullptr};  // [2][n_cqp] (X_eff_max, t_max): n+1, n, n-1
  pgf::DevCohesive *d_cprops = nullptr;
  // gather assembly (PGF_ASSEMBLY_GATHER): per-node staging blocks, see pgf::GatherArgs
  int32_t assembly = PGF_ASSEMBLY_COLOURED;
  int32_t maxlen = 0;
  int64_t stage_size = 0;  // doubles
  int64_t *d_inc_ptr = nullptr;
  uint64_t *d_inc_rec = nullptr;  // packed staging offset, pos offset and node count per block
  bool gather_fits = true;        // the packed fields hold the mesh's offsets
  uint32_t *d_pos = nullptr;
  double *d_stage = nullptr;
  bool profiling = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int32_t last_launches = 0;
};

namespace {

using pgf::Block;

int set_identity_state(pgf_handle h) {
  const int64_t NQ = h->n_qp;
  for (int t = 0; t < 3; ++t) {
    for (int i = 0; i < 9; ++i) pgf::launch_fill(h->d_F[t] + i * NQ, NQ, (i % 4 == 0) ? 1.0 : 0.0, h->stream);
    if (h->has_j2) {
      for (int I = 0; I < 6; ++I) pgf::launch_fill(h->d_P[t] + I * NQ, NQ, I < 3 ? 1.0 : 0.0, h->stream);
      pgf::launch_fill(h->d_Sc[t], h->nsc * NQ, 0.0, h->stream);
    }
    if (h->has_cp) {  // pF = I, g = g0, lambda = 0
      for (int i = 0; i < 9; ++i) pgf::launch_fill(h->d_P[t] + i * NQ, NQ, (i % 4 == 0) ? 1.0 : 0.0, h->stream);
      pgf::launch_fill(h->d_Sc[t], NQ, h->cp_g0, h->stream);
      pgf::launch_fill(h->d_Sc[t] + NQ, NQ, 0.0, h->stream);
    }
  }
  if (h->has_cp)
    for (int t = 0; t < 2; ++t) {
      pgf::launch_fill(h->d_tau[t], 12 * NQ, 0.0, h->stream);
      pgf::launch_fill(h->d_gdot[t], 12 * NQ, 0.0, h->stream);
    }
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

// Crystal-plasticity pre-pass of one block: F(n+1), the return map, and the per-QP data of
// the element kernel.
pgf::CPArgs cp_args(pgf_handle h, const Block &b, const pgf_nodal_input *in) {
  pgf::CPArgs c{};
  c.X = h->d_X;
  c.dofid = h->d_dofid;
  c.conn = b.d_conn;
  c.qgrain = b.d_qgrain;
  c.grains = h->d_grains;
  c.cpmats = h->d_cpmat;
  c.qp_base = b.qp_base;
  c.n_qp = h->n_qp;
  c.ne = b.ne;
  c.nq = b.ne * b.nqp;
  if (in) {
    c.r = in->r;
    c.d_r = in->d_r;
    c.defl = in->defl;
    c.defl_d = in->defl_d;
  }
  c.dt = h->dt;
  const int64_t NQ = h->n_qp;
  c.F_np1 = h->d_F[0];
  c.pF_n = h->d_P[1];
  c.pF_np1 = h->d_P[0];
  c.g_n = h->d_Sc[1];
  c.lam_n = h->d_Sc[1] + NQ;
  c.Sc_np1 = h->d_Sc[0];
  c.F_n = h->d_F[1];
  c.tau = h->d_tau[0];
  c.tau_n = h->d_tau[1];
  c.gdot = h->d_gdot[0];
  c.iters = h->d_iters;
  c.qnext = h->cp_chunk == 1 ? h->d_qnext : nullptr;
  c.warm = h->warm;
  c.qpbuf = h->d_qpbuf;  // indexed by device QP (pgf::qrec)
  c.err = h->d_err;
  c.law = h->cp_law;
  if (h->tf && b.ne > 0 && b.model == PGF_MODEL_CRYSTAL_PLASTICITY) {  // CM3F: the kinematics scales F
    c.tf_V = h->d_tf + b.e_base;
    c.tf_dV = h->d_tf + 6 * h->n_elem + b.e_base;
  }
  return c;
}

// All crystal-plasticity QPs as one range (device QPs [0, cp_nq)), for the kernels that do
// not depend on the element type.
pgf::CPArgs cp_args_all(pgf_handle h, const pgf_nodal_input *in) {
  Block all;
  all.qp_base = 0;
  all.ne = h->cp_nq;
  all.nqp = 1;
  all.d_qgrain = h->d_qgrain_all;
  pgf::CPArgs c = cp_args(h, all, in);
  c.conn = nullptr;
  c.tf_V = c.tf_dV = nullptr;
  return c;
}

// Crystal-plasticity pre-pass: F(n+1) per element type, then the local Newton (with the
// learned start when attached) and the per-QP tangent data over all CP QPs at once.
void cp_prepass(pgf_handle h, const pgf_nodal_input *in, int mode) {
  if (!h->cp_nq) return;
  pgf::CPArgs all = cp_args_all(h, in);
  if (mode & pgf::MODE_INTEGRATE) {
    for (auto &b : h->blocks)
      if (b.model == PGF_MODEL_CRYSTAL_PLASTICITY) pgf::launch_cp_kinematics(b.nne, cp_args(h, b, in), h->stream);
    if (h->sur && !h->warm) {  // learned start: features -> MLP -> (M, lambda, g)
      pgf::launch_cp_features(all, h->d_gquat, h->d_norm, h->d_norm + 34, h->d_feat, h->stream);
      const float *Y = nullptr;
      int ldy = 0;
      if (pgf::mlp_forward_device(h->sur, h->d_feat, all.nq, h->stream, &Y, &ldy) == PGF_OK) {
        pgf::launch_cp_warm_from_slip(all, Y, ldy, h->d_norm + 68, h->d_sur_warm, h->stream);
        all.warm = h->d_sur_warm;
      }
    }
    pgf::launch_cp_newton(all, h->cp_chunk, h->stream);
  }
  // the CM3F residual's condensation needs dM/dF too
  pgf::launch_cp_tangent(all, (mode & pgf::MODE_TANGENT) != 0 || h->tf, h->stream);
}

pgf::ElemArgs base_args(pgf_handle h, const pgf_nodal_input *in) {
  pgf::ElemArgs a{};
  a.X = h->d_X;
  a.dofid = h->d_dofid;
  a.row_start = h->d_row_start;
  a.rowlen = h->d_rowlen;
  a.mats = h->d_mats;
  a.n_qp = h->n_qp;
  if (in) {
    a.r = in->r;
    a.d_r = in->d_r;
    a.defl = in->defl;
    a.defl_d = in->defl_d;
  }
  a.F_np1 = h->d_F[0];
  a.P_n = h->d_P[1];
  a.P_np1 = h->d_P[0];
  a.Sc_n = h->d_Sc[1];
  a.Sc_np1 = h->d_Sc[0];
  a.dt = h->dt;
  a.err = h->d_err;
  return a;
}

// Gather tables (pgf::GatherArgs) of the element groups -- the volume blocks, then the
// interface blocks, each in device order.  Every (element, node) pair whose node has a free
// DOF gets a staging block; a node's blocks are contiguous and ordered by (group, element,
// local node), which fixes the order of every sum.  Blocks start on 128-byte lines and are
// padded to whole lines, so the element kernels' block stores never share a line with a
// block written at another time (a partly written line costs a read-modify-write).  off holds the groups' first-touch-marked
// node-pair offsets; soff receives each group's [n*nne] staging offsets (-1: no block).
struct GatherGroup {
  int nne;
  int64_t ne;
  const int32_t *conn;  // [ne*nne] device order
  const int16_t *off;   // [ne*nne*nne]
  std::vector<int64_t> soff;
};

void build_gather(int64_t n_nodes, const int32_t *dofid, std::vector<GatherGroup> &groups,
                  std::vector<int64_t> &inc_ptr, std::vector<int64_t> &inc_soff,
                  std::vector<int64_t> &inc_poff, std::vector<int32_t> &inc_nne, std::vector<uint32_t> &pos,
                  int64_t &stage_size) {
  auto fmask = [&](int64_t n) {
    return (dofid[n * 3] > 0 ? 1 : 0) | (dofid[n * 3 + 1] > 0 ? 2 : 0) | (dofid[n * 3 + 2] > 0 ? 4 : 0);
  };
  inc_ptr.assign(n_nodes + 1, 0);
  for (auto &g : groups)
    for (int64_t i = 0; i < g.ne * g.nne; ++i)
      if (fmask(g.conn[i])) inc_ptr[g.conn[i] + 1]++;
  for (int64_t n = 0; n < n_nodes; ++n) inc_ptr[n + 1] += inc_ptr[n];
  const int64_t n_inc = inc_ptr[n_nodes];
  std::vector<int32_t> s_group(n_inc);
  std::vector<int64_t> s_pair(n_inc);  // k*nne + a of the slot's element node
  std::vector<int64_t> cursor(inc_ptr.begin(), inc_ptr.end() - 1);
  inc_nne.resize(n_inc);
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    GatherGroup &g = groups[gi];
    g.soff.assign(g.ne * g.nne, -1);
    for (int64_t i = 0; i < g.ne * g.nne; ++i) {
      const int32_t node = g.conn[i];
      if (!fmask(node)) continue;
      const int64_t sl = cursor[node]++;
      s_group[sl] = int32_t(gi);
      s_pair[sl] = i;
      inc_nne[sl] = g.nne;
    }
  }
  inc_soff.resize(n_inc);
  inc_poff.resize(n_inc);
  int64_t so = 0, po = 0;
  for (int64_t sl = 0; sl < n_inc; ++sl) {
    inc_soff[sl] = so;
    inc_poff[sl] = po;
    so += pgf::stage_block(inc_nne[sl]);
    po += inc_nne[sl];
  }
  stage_size = so;
  pos.assign(po, 0u);
#pragma omp parallel for schedule(static)
  for (int64_t sl = 0; sl < n_inc; ++sl) {
    GatherGroup &g = groups[s_group[sl]];
    const int64_t i = s_pair[sl], k = i / g.nne;
    g.soff[i] = inc_soff[sl];
    const int16_t *orow = g.off + i * g.nne;
    for (int w = 0; w < g.nne; ++w) {
      const int o = orow[w];
      if (o == -1) continue;
      const uint32_t base = uint32_t(o >= 0 ? o : -o - 2);
      pos[inc_poff[sl] + w] = base | (uint32_t(fmask(g.conn[k * g.nne + w])) << 28);
    }
  }
}

// One pass over all elements in the pass schedule.
// Every launch of one pass, in order, on the handle's stream (also what a captured graph
// holds; the error counter is cleared by a kernel, not a memset node).
// stage: the gather mode's staging blocks (stage_size doubles); nullptr selects the coloured
// mode.
int32_t enqueue_pass(pgf_handle h, const pgf_nodal_input *in, int mode, double *f, double *vals, double *stage) {
  pgf::launch_fill_i32(h->d_err, 1, 0, h->stream);
  pgf::ElemArgs a = base_args(h, in);
  a.f = f;
  a.vals = vals;
  int32_t launches = 0;
  cp_prepass(h, in, mode);
  // gather mode: each block in one launch (all elements), then one gather over the nodes;
  // coloured mode: colour by colour, every block's elements of the colour
  const bool gather = stage != nullptr;
  double *fout = f, *vout = vals;
  if (gather) {
    a.f = nullptr;
    a.vals = nullptr;
    a.stage = stage;
    f = nullptr;
    vals = nullptr;
  }
  const int32_t nseg = gather ? 1 : h->n_colors;
  for (int32_t c = 0; c < nseg; ++c)
    for (auto &b : h->blocks) {
      const int64_t e0 = gather ? 0 : b.color_ptr[c], e1 = gather ? b.ne : b.color_ptr[c + 1];
      if (e1 <= e0) continue;
      if (b.model == PGF_MODEL_CRYSTAL_PLASTICITY) {
        pgf::CPElemArgs ce{};
        ce.X = h->d_X;
        ce.dofid = h->d_dofid;
        ce.row_start = h->d_row_start;
        ce.rowlen = h->d_rowlen;
        ce.conn = b.d_conn;
        ce.grain = b.d_egrain;
        ce.off = b.d_off;
        ce.grains = h->d_grains;
        ce.cpmats = h->d_cpmat;
        ce.law = h->cp_law;
        ce.qpbuf = h->d_qpbuf;  // indexed by device QP (pgf::qrec)
        ce.qp_base = b.qp_base;
        ce.r = in->r;
        ce.d_r = in->d_r;
        ce.defl = in->defl;
        ce.defl_d = in->defl_d;
        ce.f = f;
        ce.vals = vals;
        ce.soff = b.d_soff;
        ce.stage = gather ? stage : nullptr;
        ce.elist = gather ? nullptr : b.d_clist;
        ce.err = h->d_err;
        if (h->tf) {
          const int64_t NE = h->n_elem;
          ce.tf_V = h->d_tf + b.e_base;
          ce.tf_P = h->d_tf + 3 * NE + b.e_base;
          ce.tf_dV = h->d_tf + 6 * NE + b.e_base;
          ce.tf_dP = h->d_tf + 8 * NE + b.e_base;
          ce.tf_buf = h->d_tfbuf + b.e_base * pgf::kTfBuf;
          pgf::launch_cp_tf_elem(b.nne, mode, ce, e0, e1, h->stream);
        } else {
          pgf::launch_cp_elem(b.nne, mode, ce, e0, e1, h->stream);
        }
      } else {
        a.conn = b.d_conn;
        a.edof = b.d_edof;
        a.off = b.d_off;
        a.soff = b.d_soff;
        a.elist = gather ? nullptr : b.d_clist;
        a.mat = b.d_mat;
        a.qp_base = b.qp_base;
        if (h->tf) {
          const int64_t NE = h->n_elem;
          a.tf_V = h->d_tf + b.e_base;
          a.tf_P = h->d_tf + 3 * NE + b.e_base;
          a.tf_dV = h->d_tf + 6 * NE + b.e_base;
          a.tf_dP = h->d_tf + 8 * NE + b.e_base;
          a.tf_buf = h->d_tfbuf + b.e_base * pgf::kTfBuf;
        }
        const int var = (h->j2_damage && b.model == PGF_MODEL_J2_PLASTICITY)       ? 2
                        : (h->he_linear && b.model == PGF_MODEL_HYPER_ELASTICITY) ? 3
                                                                                  : 0;
        pgf::launch_elem(b.nne, mode, a, e0, e1, h->stream, var);
      }
      ++launches;
    }
  if (h->n_coel) {  // interface elements after every volume colour
    pgf::CohArgs ca{};
    ca.X = h->d_X;
    ca.dofid = h->d_dofid;
    ca.row_start = h->d_row_start;
    ca.rowlen = h->d_rowlen;
    ca.props = h->d_cprops;
    ca.n_cqp = h->n_cqp;
    ca.r = in->r;
    ca.d_r = in->d_r;
    ca.defl = in->defl;
    ca.defl_d = in->defl_d;
    ca.vars_n = h->d_cv[1];
    ca.vars_np1 = h->d_cv[0];
    ca.f = f;
    ca.vals = vals;
    ca.stage = gather ? stage : nullptr;
    const int32_t ncseg = gather ? 1 : h->n_ccolors;
    ca.err = h->d_err;
    for (int32_t c = 0; c < ncseg; ++c)
      for (auto &cb : h->cblocks) {
        const int64_t e0 = gather ? 0 : cb.color_ptr[c], e1 = gather ? cb.nc : cb.color_ptr[c + 1];
        if (e1 <= e0) continue;
        ca.conn = cb.d_conn;
        ca.mat = cb.d_mat;
        ca.off = cb.d_off;
        ca.soff = cb.d_soff;
        ca.elist = gather ? nullptr : cb.d_clist;
        ca.cqp_base = cb.cqp_base;
        pgf::launch_coh_elem(cb.nn, mode, ca, e0, e1, h->stream);
        ++launches;
      }
  }
  if (gather) {
    pgf::GatherArgs g{};
    g.inc_ptr = h->d_inc_ptr;
    g.inc_rec = h->d_inc_rec;
    g.pos = h->d_pos;
    g.dofid = h->d_dofid;
    g.row_start = h->d_row_start;
    g.rowlen = h->d_rowlen;
    g.stage = stage;
    g.f = fout;
    g.vals = vout;
    g.n_nodes = h->n_nodes;
    g.maxlen = h->maxlen;
    pgf::launch_gather(mode, g, h->gather_waves, h->gather_batch, h->stream);
    ++launches;
  }
  return launches;
}

// Finish the handle's work and forget its captured pass: for setters that free or replace a
// buffer whose address the graph's kernels hold.
int drop_graph(pgf_handle h) {
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (h->gexec) {
    hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  return PGF_OK;
}

double *pass_stage(pgf_handle h) { return h->assembly == PGF_ASSEMBLY_GATHER ? h->d_stage : nullptr; }

int run_pass(pgf_handle h, const pgf_nodal_input *in, int mode, double *f, double *vals) {
  if (!in || !in->r) return PGF_EINVAL;
  if (h->n_coel && !h->d_cprops) return PGF_EINVAL;  // interface properties not set
  if (h->assembly == PGF_ASSEMBLY_GATHER && !h->d_stage && h->stage_size) {
    const int rc = pgf::dalloc(&h->d_stage, size_t(h->stage_size));
    if (rc) return rc;
  }
  const bool graph = h->use_graph && !h->sur;  // the learned start may size buffers: eager
  GraphKey key{mode, in->r, in->d_r, in->defl, in->defl_d, f, vals, h->warm, h->dt, h->cp_chunk, h->gather_waves,
               h->gather_batch};
  if (graph && !(h->gexec && key == h->gkey)) {
    if (h->gexec) {
      hipGraphExecDestroy(h->gexec);
      h->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return PGF_EHIP;
    h->glaunches = enqueue_pass(h, in, mode, f, vals, pass_stage(h));
    if (hipStreamEndCapture(h->stream, &g) != hipSuccess) return PGF_EHIP;
    const hipError_t ie = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) {
      h->gexec = nullptr;
      return PGF_EHIP;
    }
    h->gkey = key;
  }
  if (h->profiling) hipEventRecord(h->ev0, h->stream);
  if (graph) {
    if (hipGraphLaunch(h->gexec, h->stream) != hipSuccess) return PGF_EHIP;
    h->last_launches = h->glaunches;
  } else {
    h->last_launches = enqueue_pass(h, in, mode, f, vals, pass_stage(h));
  }
  if (h->profiling) hipEventRecord(h->ev1, h->stream);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  int32_t err = 0;
  if (hipMemcpyAsync(&err, h->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return PGF_EHIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (h->profiling) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return PGF_EHIP;
    h->last_ms = ms;
  }
  return err;
}

int field_buffer(pgf_handle h, int32_t field, double **buf, int32_t *ncomp) {
  *buf = nullptr;
  *ncomp = 0;
  switch (field) {
    case PGF_STATE_F_NP1: case PGF_STATE_F_N: case PGF_STATE_F_NM1:
      *buf = h->d_F[field - PGF_STATE_F_NP1];
      *ncomp = 9;
      return PGF_OK;
    case PGF_STATE_PLASTIC_NP1: case PGF_STATE_PLASTIC_N: case PGF_STATE_PLASTIC_NM1:
      if (!h->has_j2 && !h->has_cp) return PGF_OK;
      *buf = h->d_P[field - PGF_STATE_PLASTIC_NP1];
      *ncomp = h->has_cp ? 9 : 6;
      return PGF_OK;
    case PGF_STATE_SCALAR_NP1: case PGF_STATE_SCALAR_N: case PGF_STATE_SCALAR_NM1:
      if (!h->has_j2 && !h->has_cp) return PGF_OK;
      *buf = h->d_Sc[field - PGF_STATE_SCALAR_NP1];
      *ncomp = h->nsc;
      return PGF_OK;
    case PGF_STATE_TAU: case PGF_STATE_TAU_N:
      if (!h->has_cp) return PGF_OK;
      *buf = h->d_tau[field - PGF_STATE_TAU];
      *ncomp = 12;
      return PGF_OK;
    case PGF_STATE_GDOT: case PGF_STATE_GDOT_N:
      if (!h->has_cp) return PGF_OK;
      *buf = h->d_gdot[field - PGF_STATE_GDOT];
      *ncomp = 12;
      return PGF_OK;
  }
  return PGF_EINVAL;
}

}  // namespace

extern "C" {

int pgf_create(const pgf_mesh *mesh, const int64_t *id, int64_t ndof, const int32_t *Ap,
               const int32_t *Ai, int32_t n_mat, const pgf_material *mats, pgf_handle *out) {
  if (!mesh || !id || !Ap || !Ai || n_mat < 1 || !mats || !out) return PGF_EINVAL;
  *out = nullptr;
  const pgf_mesh &m = *mesh;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PGF_ENODEV;
  // --- validate
  for (int64_t c = 0; c < m.n_coel; ++c) {
    const int64_t nt = m.coel_ptr[c + 1] - m.coel_ptr[c];
    if (!(nt == 6 || nt == 8)) return PGF_ENOTSUP;  // 3 + 3 or 4 + 4 node interfaces
    for (int64_t k = m.coel_ptr[c]; k < m.coel_ptr[c + 1]; ++k)
      if (m.coel_conn[k] < 0 || m.coel_conn[k] >= m.n_nodes) return PGF_EINVAL;
  }
  for (int64_t e = 0; e < m.n_elem; ++e) {
    const int64_t nne = m.elem_ptr[e + 1] - m.elem_ptr[e];
    if (!(nne == 4 || nne == 10 || nne == 8)) return PGF_ENOTSUP;
    const int32_t mi = m.elem_mat ? m.elem_mat[e] : 0;
    if (mi < 0 || mi >= n_mat) return PGF_EINVAL;
    for (int64_t k = m.elem_ptr[e]; k < m.elem_ptr[e + 1]; ++k)
      if (m.elem_conn[k] < 0 || m.elem_conn[k] >= m.n_nodes) return PGF_EINVAL;
  }
  int n_cp = 0, n_j2 = 0, n_dmg = 0, n_lin = 0, cp_index = -1;
  for (int32_t i = 0; i < n_mat; ++i) {
    const int md = mats[i].model;
    if (md == PGF_MODEL_HYPER_ELASTICITY) {
      if (mats[i].devPotFlag == 2) {  // the linear law in C (devStress_Linear)
        if (!(mats[i].G > 0.0)) return PGF_EINVAL;
        ++n_lin;
      } else if (mats[i].m01 != 0.0 || mats[i].devPotFlag != 1) {
        return PGF_ENOTSUP;  // Mooney-Rivlin with m01 = 0 (neo-Hookean) or the linear law
      }
    } else if (md == PGF_MODEL_J2_PLASTICITY) {
      ++n_j2;
      if (mats[i].mu > 0.0) {  // Weibull law: positive threshold and scales
        if (!(mats[i].Yin > 0.0) || !(mats[i].p1 > 0.0) || !(mats[i].p2 > 0.0) || !(mats[i].w_max > 0.0))
          return PGF_EINVAL;
        ++n_dmg;
      }
    } else if (md == PGF_MODEL_CRYSTAL_PLASTICITY) {
      ++n_cp;
      cp_index = i;
      if (!(mats[i].m_rate > 0.0) || !(mats[i].g0 > 0.0)) return PGF_EINVAL;
      if (mats[i].cp_elastic == 1) {
        // neo-Hookean deviatoric part only, as for the hyperelastic model
        if (mats[i].m01 != 0.0 || mats[i].devPotFlag != 1) return PGF_ENOTSUP;
        if (!(mats[i].m10 > 0.0) || !(mats[i].E > 0.0) || !(mats[i].nu < 0.5)) return PGF_EINVAL;
      } else if (mats[i].cp_elastic != 0) {
        return PGF_EINVAL;
      }
    } else {
      return PGF_ENOTSUP;
    }
  }
  // one crystal-plasticity material per handle, not mixed with J2 (they share state slots)
  if (n_cp > 1 || (n_cp && n_j2)) return PGF_ENOTSUP;
  // free DOFs of a node must be consecutive ids (no periodic nodes)
  std::vector<int32_t> dofid(size_t(m.n_nodes) * 3), row_start(m.n_nodes, -1), rowlen(m.n_nodes, 0);
  for (int64_t n = 0; n < m.n_nodes; ++n) {
    int64_t first = -1, prev = -1;
    for (int j = 0; j < 3; ++j) {
      const int64_t d = id[n * 3 + j];
      if (d > ndof || d < INT32_MIN) return PGF_EINVAL;
      dofid[n * 3 + j] = int32_t(d);
      if (d > 0) {
        if (prev > 0 && d != prev + 1) return PGF_ENOTSUP;
        if (first < 0) first = d;
        prev = d;
      }
    }
    if (first > 0) {
      row_start[n] = Ap[first - 1];
      rowlen[n] = Ap[first] - Ap[first - 1];
      for (int64_t d = first; d <= prev; ++d)
        if (Ap[d] - Ap[d - 1] != rowlen[n]) return PGF_EINVAL;  // rows of a node share a pattern
    }
  }
  // --- colouring and blocks
  std::vector<int32_t> color;
  const int ncol = pgf::color_elements(m, color);
  if (ncol < 0) return PGF_ENOTSUP;
  auto *h = new pgf_handle_s();
  h->n_nodes = m.n_nodes;
  h->ndof = ndof;
  h->nnz = Ap[ndof];
  h->n_colors = ncol;
  hipGetDevice(&h->device);
  if (hipStreamCreate(&h->own_stream) != hipSuccess) {
    delete h;
    return PGF_EHIP;
  }
  h->stream = h->own_stream;
  auto fail = [&](int rc) {
    pgf_destroy(h);
    return rc;
  };
  std::vector<std::pair<int, int>> keys;  // (nne, model)
  std::vector<int> elem_block(m.n_elem);
  for (int64_t e = 0; e < m.n_elem; ++e) {
    const int nne = int(m.elem_ptr[e + 1] - m.elem_ptr[e]);
    const int md = mats[m.elem_mat ? m.elem_mat[e] : 0].model;
    const auto key = std::make_pair(nne, md);
    auto it = std::find(keys.begin(), keys.end(), key);
    if (it == keys.end()) {
      keys.push_back(key);
      it = keys.end() - 1;
    }
    elem_block[e] = int(it - keys.begin());
    if (md == PGF_MODEL_J2_PLASTICITY) h->has_j2 = true;
    if (md == PGF_MODEL_CRYSTAL_PLASTICITY) h->has_cp = true;
  }
  h->j2_damage = h->has_j2 && n_dmg > 0;
  h->he_linear = n_lin > 0;
  h->nsc = h->j2_damage ? 6 : 2;
  // crystal-plasticity blocks first: their QPs then form one contiguous range that the
  // element-type-agnostic QP kernels (local Newton, tangent) cover in single launches
  std::vector<int> border(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) border[i] = int(i);
  std::stable_sort(border.begin(), border.end(), [&](int x, int y) {
    return (keys[x].second == PGF_MODEL_CRYSTAL_PLASTICITY) > (keys[y].second == PGF_MODEL_CRYSTAL_PLASTICITY);
  });
  std::vector<int> block_of_key(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) block_of_key[border[i]] = int(i);
  h->blocks.resize(keys.size());
  int64_t qp = 0;
  for (size_t bi = 0; bi < keys.size(); ++bi) {
    Block &b = h->blocks[bi];
    b.nne = keys[border[bi]].first;
    b.model = keys[border[bi]].second;
    b.nqp = pgf::nqp_of(b.nne);
    for (int64_t e = 0; e < m.n_elem; ++e)
      if (block_of_key[elem_block[e]] == int(bi)) b.elems.push_back(e);
    b.ne = int64_t(b.elems.size());
    b.clist.resize(b.ne);
    std::iota(b.clist.begin(), b.clist.end(), 0);
    std::stable_sort(b.clist.begin(), b.clist.end(),
                     [&](int32_t x, int32_t y) { return color[b.elems[x]] < color[b.elems[y]]; });
    b.color_ptr.assign(ncol + 1, 0);
    for (int64_t e : b.elems) b.color_ptr[color[e] + 1]++;
    for (int c = 0; c < ncol; ++c) b.color_ptr[c + 1] += b.color_ptr[c];
    b.qp_base = qp;
    qp += b.ne * b.nqp;
    b.e_base = h->n_elem;
    h->n_elem += b.ne;
  }
  h->n_qp = qp;
  for (auto &b : h->blocks)
    if (b.model == PGF_MODEL_CRYSTAL_PLASTICITY) h->cp_nq += b.ne * b.nqp;
  // mesh QP numbering: elements in mesh order, QPs in element order
  std::vector<int64_t> qp_of_elem(m.n_elem + 1, 0);
  for (int64_t e = 0; e < m.n_elem; ++e)
    qp_of_elem[e + 1] = qp_of_elem[e] + pgf::nqp_of(int(m.elem_ptr[e + 1] - m.elem_ptr[e]));
  h->qp_mesh.resize(h->n_qp);
  for (auto &b : h->blocks)
    for (int64_t k = 0; k < b.ne; ++k)
      for (int ip = 0; ip < b.nqp; ++ip) h->qp_mesh[b.qp_base + k * b.nqp + ip] = qp_of_elem[b.elems[k]] + ip;
  // --- scatter offsets in device order, with first-touch marks following the schedule
  std::vector<int64_t> pair_ptr;
  std::vector<int16_t> off;
  if (pgf::build_scatter_offsets(m, id, ndof, Ap, Ai, pair_ptr, off) != PGF_OK) return fail(PGF_EINVAL);
  std::vector<uint8_t> touched(size_t(h->nnz), 0);
  std::vector<std::vector<int16_t>> boff(h->blocks.size());
  for (size_t bi = 0; bi < h->blocks.size(); ++bi) boff[bi].resize(h->blocks[bi].ne * h->blocks[bi].nne * h->blocks[bi].nne);
  for (int32_t c = 0; c < ncol; ++c)
    for (size_t bi = 0; bi < h->blocks.size(); ++bi) {
      Block &b = h->blocks[bi];
      const int nne = b.nne;
      for (int64_t j = b.color_ptr[c]; j < b.color_ptr[c + 1]; ++j) {
        const int64_t k = b.clist[j], e = b.elems[k];
        const int64_t *conn = m.elem_conn + m.elem_ptr[e];
        for (int a = 0; a < nne; ++a)
          for (int w = 0; w < nne; ++w) {
            int16_t o = off[pair_ptr[e] + a * nne + w];
            if (o >= 0) {
              const int64_t pos = int64_t(row_start[conn[a]]) + o;
              if (!touched[pos]) {
                touched[pos] = 1;
                o = int16_t(-int(o) - 2);
              }
            }
            boff[bi][(k * nne + a) * nne + w] = o;
          }
      }
    }
  // --- interface elements: own colours after the volume ones, first-touch marks continued
  std::vector<std::vector<int16_t>> cboff;
  if (m.n_coel > 0) {
    std::vector<int32_t> ccolor;
    const int nccol = pgf::color_coels(m, ccolor);
    if (nccol < 0) return fail(PGF_ENOTSUP);
    h->n_coel = m.n_coel;
    h->n_ccolors = nccol;
    for (int nn : {3, 4}) {
      pgf::CohBlock cb;
      cb.nn = nn;
      for (int64_t c = 0; c < m.n_coel; ++c)
        if (m.coel_ptr[c + 1] - m.coel_ptr[c] == 2 * nn) cb.elems.push_back(c);
      if (cb.elems.empty()) continue;
      cb.nc = int64_t(cb.elems.size());
      cb.clist.resize(cb.nc);
      std::iota(cb.clist.begin(), cb.clist.end(), 0);
      std::stable_sort(cb.clist.begin(), cb.clist.end(),
                       [&](int32_t x, int32_t y) { return ccolor[cb.elems[x]] < ccolor[cb.elems[y]]; });
      cb.color_ptr.assign(nccol + 1, 0);
      for (int64_t c : cb.elems) cb.color_ptr[ccolor[c] + 1]++;
      for (int c = 0; c < nccol; ++c) cb.color_ptr[c + 1] += cb.color_ptr[c];
      cb.cqp_base = h->n_cqp;
      h->n_cqp += cb.nc * nn;
      h->cblocks.push_back(std::move(cb));
    }
    std::vector<int64_t> cq_of(m.n_coel + 1, 0);
    for (int64_t c = 0; c < m.n_coel; ++c) cq_of[c + 1] = cq_of[c] + (m.coel_ptr[c + 1] - m.coel_ptr[c]) / 2;
    h->cqp_mesh.resize(h->n_cqp);
    for (auto &cb : h->cblocks)
      for (int64_t k = 0; k < cb.nc; ++k)
        for (int q = 0; q < cb.nn; ++q) h->cqp_mesh[cb.cqp_base + k * cb.nn + q] = cq_of[cb.elems[k]] + q;
    std::vector<int64_t> cpair_ptr;
    std::vector<int16_t> coff;
    if (pgf::build_coel_offsets(m, id, Ap, Ai, cpair_ptr, coff) != PGF_OK) return fail(PGF_EINVAL);
    cboff.resize(h->cblocks.size());
    for (size_t bi = 0; bi < h->cblocks.size(); ++bi) {
      const int nt = 2 * h->cblocks[bi].nn;
      cboff[bi].resize(h->cblocks[bi].nc * nt * nt);
    }
    for (int32_t c = 0; c < nccol; ++c)
      for (size_t bi = 0; bi < h->cblocks.size(); ++bi) {
        pgf::CohBlock &cb = h->cblocks[bi];
        const int nt = 2 * cb.nn;
        for (int64_t j = cb.color_ptr[c]; j < cb.color_ptr[c + 1]; ++j) {
          const int64_t k = cb.clist[j], ce = cb.elems[k];
          const int64_t *conn = m.coel_conn + m.coel_ptr[ce];
          for (int a = 0; a < nt; ++a)
            for (int w = 0; w < nt; ++w) {
              int16_t o = coff[cpair_ptr[ce] + a * nt + w];
              if (o >= 0) {
                const int64_t pos = int64_t(row_start[conn[a]]) + o;
                if (!touched[pos]) {
                  touched[pos] = 1;
                  o = int16_t(-int(o) - 2);
                }
              }
              cboff[bi][(k * nt + a) * nt + w] = o;
            }
        }
      }
  }
  // --- device-order connectivity and the gather tables
  std::vector<std::vector<int32_t>> bconn(h->blocks.size()), cconn(h->cblocks.size());
  for (size_t bi = 0; bi < h->blocks.size(); ++bi) {
    Block &b = h->blocks[bi];
    bconn[bi].resize(b.ne * b.nne);
    for (int64_t k = 0; k < b.ne; ++k)
      for (int a = 0; a < b.nne; ++a) bconn[bi][k * b.nne + a] = int32_t(m.elem_conn[m.elem_ptr[b.elems[k]] + a]);
  }
  for (size_t bi = 0; bi < h->cblocks.size(); ++bi) {
    pgf::CohBlock &cb = h->cblocks[bi];
    const int nt = 2 * cb.nn;
    cconn[bi].resize(cb.nc * nt);
    for (int64_t k = 0; k < cb.nc; ++k)
      for (int a = 0; a < nt; ++a) cconn[bi][k * nt + a] = int32_t(m.coel_conn[m.coel_ptr[cb.elems[k]] + a]);
  }
  std::vector<GatherGroup> groups;
  for (size_t bi = 0; bi < h->blocks.size(); ++bi)
    groups.push_back({h->blocks[bi].nne, h->blocks[bi].ne, bconn[bi].data(), boff[bi].data(), {}});
  for (size_t bi = 0; bi < h->cblocks.size(); ++bi)
    groups.push_back({2 * h->cblocks[bi].nn, h->cblocks[bi].nc, cconn[bi].data(), cboff[bi].data(), {}});
  std::vector<int64_t> inc_ptr, inc_soff, inc_poff;
  std::vector<int32_t> inc_nne;
  std::vector<uint32_t> gpos;
  build_gather(m.n_nodes, dofid.data(), groups, inc_ptr, inc_soff, inc_poff, inc_nne, gpos, h->stage_size);
  for (int64_t n = 0; n < m.n_nodes; ++n) h->maxlen = std::max(h->maxlen, rowlen[n]);
  std::vector<uint64_t> inc_rec(inc_nne.size());
  for (size_t sl = 0; sl < inc_rec.size(); ++sl) {
    const uint64_t s16 = uint64_t(inc_soff[sl]) / 16, po = uint64_t(inc_poff[sl]);
    if (inc_soff[sl] % 16 != 0 || s16 >> 32 || po >> 28 || inc_nne[sl] > 15) h->gather_fits = false;
    inc_rec[sl] = (s16 & 0xffffffffull) | ((po & 0x0fffffffull) << 32) | (uint64_t(inc_nne[sl]) << 60);
  }
  h->assembly = (pgf::gather_waves(h->maxlen) >= 1 && h->gather_fits) ? PGF_ASSEMBLY_GATHER : PGF_ASSEMBLY_COLOURED;
  // --- upload
  int rc = PGF_OK;
  if ((rc = pgf::upload(&h->d_inc_ptr, inc_ptr.data(), inc_ptr.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_inc_rec, inc_rec.data(), inc_rec.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_pos, gpos.data(), gpos.size()))) return fail(rc);
  for (size_t bi = 0; bi < h->cblocks.size(); ++bi) {
    pgf::CohBlock &cb = h->cblocks[bi];
    const std::vector<int32_t> zero(cb.nc, 0);
    const std::vector<int64_t> &soff = groups[h->blocks.size() + bi].soff;
    if ((rc = pgf::upload(&cb.d_conn, cconn[bi].data(), cconn[bi].size()))) return fail(rc);
    if ((rc = pgf::upload(&cb.d_mat, zero.data(), zero.size()))) return fail(rc);
    if ((rc = pgf::upload(&cb.d_off, cboff[bi].data(), cboff[bi].size()))) return fail(rc);
    if ((rc = pgf::upload(&cb.d_soff, soff.data(), soff.size()))) return fail(rc);
    if ((rc = pgf::upload(&cb.d_clist, cb.clist.data(), cb.clist.size()))) return fail(rc);
  }
  if (h->n_cqp) {
    for (int t = 0; t < 3; ++t) {
      if ((rc = pgf::dalloc(&h->d_cv[t], size_t(h->n_cqp) * 2))) return fail(rc);
      pgf::launch_fill(h->d_cv[t], 2 * h->n_cqp, 0.0, h->stream);
    }
  }
  std::vector<pgf::DevMaterial> dm(n_mat);
  for (int32_t i = 0; i < n_mat; ++i) {
    dm[i].model = mats[i].model;
    dm[i].volPotFlag = mats[i].volPotFlag;
    dm[i].devPotFlag = mats[i].devPotFlag;
    dm[i].G = (mats[i].model == PGF_MODEL_HYPER_ELASTICITY && mats[i].devPotFlag != 2) ? 2.0 * mats[i].m10
                                                                                        : mats[i].G;
    dm[i].kappa = mats[i].E / (3.0 * (1.0 - 2.0 * mats[i].nu));
    dm[i].k0 = mats[i].k0;
    dm[i].hp = mats[i].hp;
    dm[i].mu = mats[i].mu;
    dm[i].wmax = mats[i].w_max;
    dm[i].p1 = mats[i].p1;
    dm[i].p2 = mats[i].p2;
    dm[i].Yin = mats[i].Yin;
  }
  if ((rc = pgf::upload(&h->d_mats, dm.data(), dm.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_X, m.X, size_t(m.n_nodes) * 3))) return fail(rc);
  if ((rc = pgf::upload(&h->d_dofid, dofid.data(), dofid.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_row_start, row_start.data(), row_start.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_rowlen, rowlen.data(), rowlen.size()))) return fail(rc);
  if ((rc = pgf::upload(&h->d_Ap, Ap, size_t(ndof) + 1))) return fail(rc);
  if ((rc = pgf::upload(&h->d_Ai, Ai, size_t(h->nnz)))) return fail(rc);
  for (size_t bi = 0; bi < h->blocks.size(); ++bi) {
    Block &b = h->blocks[bi];
    std::vector<int32_t> mat(b.ne);
    for (int64_t k = 0; k < b.ne; ++k) mat[k] = m.elem_mat ? m.elem_mat[b.elems[k]] : 0;
    if ((rc = pgf::upload(&b.d_conn, bconn[bi].data(), bconn[bi].size()))) return fail(rc);
    {  // the element DOF ids, element-major (element kernel; CM3F update and initial conditions)
      std::vector<int32_t> edof(bconn[bi].size() * 3);
      for (size_t i = 0; i < bconn[bi].size(); ++i)
        for (int c = 0; c < 3; ++c) edof[i * 3 + c] = dofid[size_t(bconn[bi][i]) * 3 + c];
      if ((rc = pgf::upload(&b.d_edof, edof.data(), edof.size()))) return fail(rc);
    }
    if ((rc = pgf::upload(&b.d_mat, mat.data(), mat.size()))) return fail(rc);
    if ((rc = pgf::upload(&b.d_off, boff[bi].data(), boff[bi].size()))) return fail(rc);
    if ((rc = pgf::upload(&b.d_soff, groups[bi].soff.data(), groups[bi].soff.size()))) return fail(rc);
    if ((rc = pgf::upload(&b.d_clist, b.clist.data(), b.clist.size()))) return fail(rc);
  }
  for (int t = 0; t < 3; ++t) {
    if ((rc = pgf::dalloc(&h->d_F[t], size_t(h->n_qp) * 9))) return fail(rc);
    if (h->has_j2 || h->has_cp) {
      if ((rc = pgf::dalloc(&h->d_P[t], size_t(h->n_qp) * (h->has_cp ? 9 : 6)))) return fail(rc);
      if ((rc = pgf::dalloc(&h->d_Sc[t], size_t(h->n_qp) * h->nsc))) return fail(rc);
    }
  }
  if (h->has_cp) {
    const pgf_material &cpm = mats[cp_index];
    pgf::CPMaterial c{};
    c.gd0 = cpm.gamma_dot_0;
    c.m = cpm.m_rate;
    c.h0 = cpm.G0;
    c.g0 = cpm.g0;
    c.gs0 = cpm.gs_0;
    c.gds = cpm.gamma_dot_s;
    c.w = cpm.w_sat;
    c.tol = cpm.tol_M > 0.0 ? cpm.tol_M : 1.0e-10;
    c.maxit = cpm.max_itr_M > 0 ? cpm.max_itr_M : 50;
    c.m10 = cpm.m10;
    c.kappa = cpm.E / (3.0 * (1.0 - 2.0 * cpm.nu));
    c.law = cpm.cp_elastic;
    c.vol_flag = cpm.volPotFlag;
    {
      // the flow rule's exponent, evaluated as the kernels do
      const double e = 1.0 / c.m - 1.0;
      c.ipow = (e >= 0.0 && e <= 64.0 && e == std::floor(e)) ? int32_t(e) : -1;
    }
    h->cp_g0 = cpm.g0;
    h->cp_law = cpm.cp_elastic;
    h->cp_mat_host = {cpm.C11, cpm.C12, cpm.C44};
    if ((rc = pgf::upload(&h->d_cpmat, &c, 1))) return fail(rc);
    for (int t = 0; t < 2; ++t) {
      if ((rc = pgf::dalloc(&h->d_tau[t], size_t(h->n_qp) * 12))) return fail(rc);
      if ((rc = pgf::dalloc(&h->d_gdot[t], size_t(h->n_qp) * 12))) return fail(rc);
    }
    if ((rc = pgf::dalloc(&h->d_iters, size_t(h->n_qp)))) return fail(rc);
    if ((rc = pgf::dalloc(&h->d_qnext, 1))) return fail(rc);
    if ((rc = pgf::dalloc(&h->d_qpbuf, size_t((h->n_qp + pgf::kRecStride - 1) / pgf::kRecStride) *
                                             pgf::kCPBuf * pgf::kRecStride)))
      return fail(rc);
    if (hipMemset(h->d_iters, 0, sizeof(int32_t) * h->n_qp) != hipSuccess) return fail(PGF_EHIP);
    if ((rc = pgf::dalloc(&h->d_qgrain_all, size_t(h->n_qp)))) return fail(rc);
    if (hipMemset(h->d_qgrain_all, 0, sizeof(int32_t) * h->n_qp) != hipSuccess) return fail(PGF_EHIP);
    for (auto &b : h->blocks) {
      if (b.model != PGF_MODEL_CRYSTAL_PLASTICITY) continue;
      if ((rc = pgf::dalloc(&b.d_egrain, size_t(b.ne)))) return fail(rc);
      if (hipMemset(b.d_egrain, 0, sizeof(int32_t) * b.ne) != hipSuccess) return fail(PGF_EHIP);
      b.d_qgrain = h->d_qgrain_all + b.qp_base;  // view into the handle's array
    }
    const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if ((rc = pgf_set_orientations(h, 1, I9, nullptr))) return fail(rc);
  }
  if ((rc = pgf::dalloc(&h->d_err, 1))) return fail(rc);
  if ((rc = pgf::dalloc(&h->d_red, 1))) return fail(rc);
  if ((rc = set_identity_state(h))) return fail(rc);
  *out = h;
  return PGF_OK;
}

int pgf_destroy(pgf_handle h) {
  if (!h) return PGF_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  hipFree(h->d_X);
  hipFree(h->d_dofid);
  hipFree(h->d_row_start);
  hipFree(h->d_rowlen);
  hipFree(h->d_mats);
  for (auto &b : h->blocks) {
    hipFree(b.d_conn);
    hipFree(b.d_edof);
    hipFree(b.d_off);
    hipFree(b.d_soff);
    hipFree(b.d_clist);
    hipFree(b.d_mat);
    hipFree(b.d_egrain);
  }
  for (auto &cb : h->cblocks) {
    hipFree(cb.d_conn);
    hipFree(cb.d_mat);
    hipFree(cb.d_off);
    hipFree(cb.d_soff);
    hipFree(cb.d_clist);
  }
  hipFree(h->d_inc_ptr);
  hipFree(h->d_inc_rec);
  hipFree(h->d_pos);
  hipFree(h->d_stage);
  for (int t = 0; t < 3; ++t) hipFree(h->d_cv[t]);
  hipFree(h->d_tf);
  hipFree(h->d_tfbuf);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  hipFree(h->d_Ap);
  hipFree(h->d_Ai);
  hipFree(h->d_solve_work);
  hipFree(h->d_cprops);
  hipFree(h->d_cpmat);
  hipFree(h->d_grains);
  hipFree(h->d_gquat);
  hipFree(h->d_qgrain_all);
  hipFree(h->d_norm);
  hipFree(h->d_feat);
  hipFree(h->d_sur_warm);
  for (int t = 0; t < 2; ++t) {
    hipFree(h->d_tau[t]);
    hipFree(h->d_gdot[t]);
  }
  hipFree(h->d_iters);
  hipFree(h->d_qnext);
  hipFree(h->d_qpbuf);
  for (int t = 0; t < 3; ++t) {
    hipFree(h->d_F[t]);
    hipFree(h->d_P[t]);
    hipFree(h->d_Sc[t]);
  }
  hipFree(h->d_err);
  hipFree(h->d_red);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->own_stream) hipStreamDestroy(h->own_stream);
  delete h;
  return PGF_OK;
}

int pgf_set_profiling(pgf_handle h, int on) {
  if (!h) return PGF_EINVAL;
  if (on && !h->ev0) {
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return PGF_EHIP;
  }
  h->profiling = on != 0;
  return PGF_OK;
}

int pgf_last_pass_ms(pgf_handle h, double *ms, int32_t *n_launches) {
  if (!h || !ms) return PGF_EINVAL;
  *ms = h->last_ms;
  if (n_launches) *n_launches = h->last_launches;
  return PGF_OK;
}

int pgf_set_stream(pgf_handle h, void *stream) {
  if (!h) return PGF_EINVAL;
  h->stream = stream ? reinterpret_cast<hipStream_t>(stream) : h->own_stream;
  return PGF_OK;
}

int pgf_synchronize(pgf_handle h) {
  if (!h) return PGF_EINVAL;
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_get_sizes(pgf_handle h, int64_t *n_qp, int64_t *nnz, int64_t *ndof, int32_t *n_colors) {
  if (!h) return PGF_EINVAL;
  if (n_qp) *n_qp = h->n_qp;
  if (nnz) *nnz = h->nnz;
  if (ndof) *ndof = h->ndof;
  if (n_colors) *n_colors = h->n_colors;
  return PGF_OK;
}

int pgf_residual(pgf_handle h, const pgf_nodal_input *in, double dt, int run_integration, double *f) {
  if (!h || !f) return PGF_EINVAL;
  if (dt > 0.0) h->dt = dt;
  return run_pass(h, in, run_integration ? (pgf::MODE_INTEGRATE | pgf::MODE_RESIDUAL) : pgf::MODE_RESIDUAL,
                  f, nullptr);
}

int pgf_tangent(pgf_handle h, const pgf_nodal_input *in, double dt, double *vals) {
  if (!h || !vals) return PGF_EINVAL;
  if (dt > 0.0) h->dt = dt;
  return run_pass(h, in, pgf::MODE_TANGENT, nullptr, vals);
}

int pgf_prescribed_load(pgf_handle h, const pgf_nodal_input *in, double dt, double *f_defl) {
  if (!h || !f_defl || !in || !in->defl_d) return PGF_EINVAL;
  if (dt > 0.0) h->dt = dt;
  return run_pass(h, in, pgf::MODE_TANGENT | pgf::MODE_DEFL_LOAD, f_defl, nullptr);
}

int pgf_residual_tangent(pgf_handle h, const pgf_nodal_input *in, double dt, double *f, double *vals) {
  if (!h || !f || !vals) return PGF_EINVAL;
  if (dt > 0.0) h->dt = dt;
  return run_pass(h, in, pgf::MODE_INTEGRATE | pgf::MODE_RESIDUAL | pgf::MODE_TANGENT, f, vals);
}

int pgf_residual_tangent_static(pgf_handle h, const pgf_nodal_input *in, double dt, double *f, double *vals) {
  if (!h || !f || !vals || !in || !in->r) return PGF_EINVAL;
  if (h->n_coel && !h->d_cprops) return PGF_EINVAL;  // interface properties not set
  // the gather tables exist and fit; the learned start sizes its MLP activations on first use
  if (pgf::gather_waves(h->maxlen) < 1 || !h->gather_fits || h->sur) return PGF_ENOTSUP;
  if (h->stage_size > PGF_STATIC_STAGE_DOUBLES) return PGF_ESCRATCH;
  double *stage = nullptr;
  int rc = static_stage_ptr(h->device, &stage);
  if (rc) return rc;
  if (dt > 0.0) h->dt = dt;
  // eager launches only: instantiating a graph may allocate.  Every staging block the pass
  // reads is written earlier in the same pass, so what an earlier call left is never read.
  if (h->profiling) hipEventRecord(h->ev0, h->stream);
  h->last_launches = enqueue_pass(h, in, pgf::MODE_INTEGRATE | pgf::MODE_RESIDUAL | pgf::MODE_TANGENT, f, vals,
                                  h->stage_size ? stage : nullptr);
  if (h->profiling) hipEventRecord(h->ev1, h->stream);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  int32_t err = 0;
  if (hipMemcpyAsync(&err, h->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return PGF_EHIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (h->profiling) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return PGF_EHIP;
    h->last_ms = ms;
  }
  return err;
}

int pgf_output_variables(pgf_handle h, int32_t level, double *S, double *E, double *sigma, double *eps) {
  if (!h || level < 0 || level > 2) return PGF_EINVAL;
  const int64_t NQ = h->n_qp, NE = h->n_elem;
  double *d_S = nullptr, *d_E = nullptr, *d_sig = nullptr, *d_eps = nullptr;
  int rc = pgf::dalloc(&d_S, size_t(6 * NQ));
  if (!rc) rc = pgf::dalloc(&d_E, size_t(6 * NQ));
  if (!rc) rc = pgf::dalloc(&d_sig, size_t(6 * NE));
  if (!rc) rc = pgf::dalloc(&d_eps, size_t(6 * NE));
  std::vector<double> tq(6 * NQ), te(6 * NE);
  if (!rc) {
    for (auto &b : h->blocks) {
      pgf::OutArgs o{};
      o.X = h->d_X;
      o.conn = b.d_conn;
      o.mat = b.d_mat;
      o.mats = h->d_mats;
      o.model = b.model;
      o.qgrain = b.d_qgrain;
      o.grains = h->d_grains;
      o.cpmats = h->d_cpmat;
      o.qp_base = b.qp_base;
      o.n_qp = NQ;
      o.ne = b.ne;
      o.F = h->d_F[level];
      o.P = (h->has_j2 || h->has_cp) ? h->d_P[level] : nullptr;
      o.Sc = (h->has_j2 || h->has_cp) ? h->d_Sc[level] : nullptr;
      o.S = d_S;
      o.E = d_E;
      o.sig = d_sig + 6 * b.e_base;  // the block's [6][ne] rows
      o.eps = d_eps + 6 * b.e_base;
      pgf::launch_output(b.nne, o, h->stream);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = PGF_EHIP;
  }
  for (int pass = 0; pass < 2 && !rc; ++pass) {
    double *out = pass == 0 ? S : E;
    if (!out) continue;
    if (hipMemcpy(tq.data(), pass == 0 ? d_S : d_E, sizeof(double) * tq.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = PGF_EHIP;
      break;
    }
    for (int c = 0; c < 6; ++c)
      for (int64_t q = 0; q < NQ; ++q) out[h->qp_mesh[q] * 6 + c] = tq[c * NQ + q];
  }
  for (int pass = 0; pass < 2 && !rc; ++pass) {
    double *out = pass == 0 ? sigma : eps;
    if (!out) continue;
    if (hipMemcpy(te.data(), pass == 0 ? d_sig : d_eps, sizeof(double) * te.size(), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      rc = PGF_EHIP;
      break;
    }
    for (auto &b : h->blocks)  // block b's rows are [6][b.ne] at e_base * 6
      for (int c = 0; c < 6; ++c)
        for (int64_t k = 0; k < b.ne; ++k) out[b.elems[k] * 6 + c] = te[b.e_base * 6 + c * b.ne + k];
  }
  hipFree(d_S);
  hipFree(d_E);
  hipFree(d_sig);
  hipFree(d_eps);
  return rc;
}

int pgf_cp_start_stats(pgf_handle h, int64_t *n_learned, int64_t *n_fallback) {
  if (!h || !n_learned || !n_fallback) return PGF_EINVAL;
  if (!h->has_cp) return PGF_ENOTSUP;
  std::vector<int32_t> it(h->cp_nq);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (h->cp_nq && hipMemcpy(it.data(), h->d_iters, sizeof(int32_t) * h->cp_nq, hipMemcpyDeviceToHost) != hipSuccess)
    return PGF_EHIP;
  int64_t nl = 0, nf = 0;
  for (const int32_t v : it) {
    const int kind = v >> 24;
    nl += kind == 1;
    nf += kind == 2;
  }
  *n_learned = nl;
  *n_fallback = nf;
  return PGF_OK;
}

int pgf_subdivision_param(pgf_handle h, double dt, double *alpha) {
  if (!h || !alpha) return PGF_EINVAL;
  if (!(dt > 0.0)) dt = h->dt;
  *alpha = 0.0;
  if (!h->has_cp && !h->j2_damage) return PGF_OK;  // the hyperelastic and undamaged models give 0
  pgf::launch_fill(h->d_red, 1, 0.0, h->stream);
  auto *out = reinterpret_cast<unsigned long long *>(h->d_red);
  if (h->has_cp)  // CP_PARAM::get_subdiv_param: dt sum_a |gamma_dot_a(n+1)| / MAX_D_GAMMA (0.005)
    pgf::launch_rate_max(h->d_gdot[0], 12, h->n_qp, h->cp_nq, dt / 0.005, nullptr, nullptr, 0.0, out, h->stream);
  if (h->j2_damage)  // CM_J2P_PARAM::get_subdiv_param: (w(n+1) - w(n)) / DELTA_W_MAX (0.05)
    pgf::launch_rate_max(nullptr, 0, 0, h->n_qp, 0.0, h->d_Sc[0] + 2 * h->n_qp, h->d_Sc[1] + 2 * h->n_qp,
                         1.0 / 0.05, out, h->stream);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  if (hipMemcpyAsync(alpha, h->d_red, sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return PGF_EHIP;
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_stage_doubles(pgf_handle h, int64_t *n) {
  if (!h || !n) return PGF_EINVAL;
  *n = h->stage_size;
  return PGF_OK;
}

int pgf_commit_step(pgf_handle h) {
  if (!h) return PGF_EINVAL;
  const int64_t NQ = h->n_qp;
  pgf::launch_copy(h->d_F[2], h->d_F[1], 9 * NQ, h->stream);
  pgf::launch_copy(h->d_F[1], h->d_F[0], 9 * NQ, h->stream);
  if (h->has_j2 || h->has_cp) {
    const int np = h->has_cp ? 9 : 6;
    pgf::launch_copy(h->d_P[2], h->d_P[1], np * NQ, h->stream);
    pgf::launch_copy(h->d_P[1], h->d_P[0], np * NQ, h->stream);
    pgf::launch_copy(h->d_Sc[2], h->d_Sc[1], h->nsc * NQ, h->stream);
    pgf::launch_copy(h->d_Sc[1], h->d_Sc[0], h->nsc * NQ, h->stream);
  }
  if (h->has_cp) {  // tau_n, gamma_dot_n (src/plasticity_model.cc:174-192)
    pgf::launch_copy(h->d_tau[1], h->d_tau[0], 12 * NQ, h->stream);
    pgf::launch_copy(h->d_gdot[1], h->d_gdot[0], 12 * NQ, h->stream);
  }
  if (h->n_cqp) {  // interface history (increment_cohesive_elements, src/cohesive_element.cc:645-700)
    pgf::launch_copy(h->d_cv[2], h->d_cv[1], 2 * h->n_cqp, h->stream);
    pgf::launch_copy(h->d_cv[1], h->d_cv[0], 2 * h->n_cqp, h->stream);
  }
  if (h->tf) {  // update_np1_from_increments, then update_for_next_time_step (TL)
    const int64_t NE = h->n_elem;
    double *t = h->d_tf;
    for (int v = 0; v < 2; ++v) {  // V, then P
      double *np1 = t + 3 * v * NE, *n = np1 + NE, *nm1 = n + NE, *inc = t + (6 + 2 * v) * NE;
      pgf::launch_axpy(np1, inc, 1.0, NE, h->stream);
      pgf::launch_copy(nm1, n, NE, h->stream);
      pgf::launch_copy(n, np1, NE, h->stream);
    }
    pgf::launch_fill(t + 6 * NE, (pgf::kTfSlots - 6) * NE, 0.0, h->stream);
  }
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_reset_step(pgf_handle h) {
  if (!h) return PGF_EINVAL;
  const int64_t NQ = h->n_qp;
  pgf::launch_copy(h->d_F[0], h->d_F[1], 9 * NQ, h->stream);
  if (h->has_j2 || h->has_cp) {
    pgf::launch_copy(h->d_P[0], h->d_P[1], (h->has_cp ? 9 : 6) * NQ, h->stream);
    pgf::launch_copy(h->d_Sc[0], h->d_Sc[1], h->nsc * NQ, h->stream);
  }
  if (h->has_cp) {
    pgf::launch_copy(h->d_tau[0], h->d_tau[1], 12 * NQ, h->stream);
    pgf::launch_copy(h->d_gdot[0], h->d_gdot[1], 12 * NQ, h->stream);
  }
  if (h->n_cqp) pgf::launch_copy(h->d_cv[0], h->d_cv[1], 2 * h->n_cqp, h->stream);
  if (h->tf) pgf::launch_fill(h->d_tf + 6 * h->n_elem, (pgf::kTfSlots - 6) * h->n_elem, 0.0, h->stream);  // tf.reset()
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

// ------------------------------------------------------------------ three-field (CM3F)
int pgf_set_three_field(pgf_handle h, int32_t on) {
  if (!h) return PGF_EINVAL;
  if (int rc = drop_graph(h)) return rc;  // a captured pass holds the old kernels
  if (!on) {
    hipFree(h->d_tf);
    hipFree(h->d_tfbuf);
    h->d_tf = h->d_tfbuf = nullptr;
    h->tf = false;
    return PGF_OK;
  }
  // the hyperelastic model, J2 plasticity (pF = I: its return map's deviatoric stress and
  // tangent with the element pressure, in the hyperelastic model's kernel), and crystal
  // plasticity with its hyperelastic law (the CM3F element splits the law into deviatoric and
  // volumetric parts; the cubic law has no split)
  for (auto &b : h->blocks)
    if (b.model != PGF_MODEL_HYPER_ELASTICITY && b.model != PGF_MODEL_J2_PLASTICITY &&
        !(b.model == PGF_MODEL_CRYSTAL_PLASTICITY && h->cp_law == 1))
      return PGF_ENOTSUP;
  if (h->he_linear) return PGF_ENOTSUP;  // the three-field element takes the neo-Hookean law
  if (h->j2_damage) return PGF_ENOTSUP;  // the three-field kernel carries no damage step
  if (h->tf) return PGF_OK;
  const int64_t NE = h->n_elem;
  int rc = pgf::dalloc(&h->d_tf, size_t(pgf::kTfSlots * NE));
  if (!rc) rc = pgf::dalloc(&h->d_tfbuf, size_t(NE * pgf::kTfBuf));
  if (rc) {
    hipFree(h->d_tf);
    h->d_tf = nullptr;
    return rc;
  }
  // ThreeFieldVariables::construct: V = 1, P = 0, increments 0
  pgf::launch_fill(h->d_tf, 3 * NE, 1.0, h->stream);
  pgf::launch_fill(h->d_tf + 3 * NE, (pgf::kTfSlots - 3) * NE, 0.0, h->stream);
  pgf::launch_fill(h->d_tfbuf, NE * pgf::kTfBuf, 0.0, h->stream);
  h->tf = true;
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

namespace {
pgf::TfArgs tf_args(pgf_handle h, const Block &b) {
  pgf::TfArgs t{};
  t.X = h->d_X;
  t.conn = b.d_conn;
  t.edof = b.d_edof;
  t.mat = b.d_mat;
  t.mats = h->d_mats;
  t.ne = b.ne;
  t.buf = h->d_tfbuf + b.e_base * pgf::kTfBuf;
  return t;
}
}  // namespace

int pgf_tf_initial(pgf_handle h, const pgf_nodal_input *in_n, const pgf_nodal_input *in_nm1) {
  if (!h || !h->tf || !in_n || !in_n->r) return PGF_EINVAL;
  const int64_t NE = h->n_elem;
  pgf::launch_fill_i32(h->d_err, 1, 0, h->stream);
  for (int s = 0; s < 2; ++s) {
    const pgf_nodal_input *in = (s == 1 && in_nm1) ? in_nm1 : in_n;
    if (!in->r) return PGF_EINVAL;
    for (auto &b : h->blocks) {
      pgf::TfArgs t = tf_args(h, b);
      t.r = in->r;
      t.d_r = in->d_r;
      t.defl = in->defl;
      t.defl_d = in->defl_d;
      t.V = h->d_tf + (s == 0 ? 1 : 2) * NE + b.e_base;      // V(n) / V(n-1)
      t.P = h->d_tf + (s == 0 ? 4 : 5) * NE + b.e_base;      // P(n) / P(n-1)
      pgf::launch_tf_initial(b.nne, t, h->stream);
      // the model's F(n) / F(n-1) from the same displacements (set_F in
      // compute_cm_initial_conditions, src/constitutive_model.cc:4097-4105)
      pgf::CPArgs c = cp_args(h, b, in);
      c.F_np1 = h->d_F[s == 0 ? 1 : 2];
      c.tf_V = c.tf_dV = nullptr;
      pgf::launch_cp_kinematics(b.nne, c, h->stream);
    }
  }
  pgf::launch_copy(h->d_tf, h->d_tf + NE, NE, h->stream);           // V(n+1) = V(n)
  pgf::launch_copy(h->d_tf + 3 * NE, h->d_tf + 4 * NE, NE, h->stream);  // P(n+1) = P(n)
  pgf::launch_fill(h->d_tf + 6 * NE, (pgf::kTfSlots - 6) * NE, 0.0, h->stream);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  int32_t err = 0;
  if (hipMemcpyAsync(&err, h->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return PGF_EHIP;
  return hipStreamSynchronize(h->stream) == hipSuccess ? err : PGF_EHIP;
}

int pgf_tf_update(pgf_handle h, const double *du) {
  if (!h || !h->tf || !du) return PGF_EINVAL;
  const int64_t NE = h->n_elem;
  for (auto &b : h->blocks) {
    pgf::TfArgs t = tf_args(h, b);
    t.du = du;
    t.ddV = h->d_tf + 10 * NE + b.e_base;  // the full increments, kept for every trial
    t.ddP = h->d_tf + 11 * NE + b.e_base;
    pgf::launch_tf_update(b.nne, t, h->stream);
  }
  double *tf = h->d_tf;
  pgf::launch_copy(tf + 7 * NE, tf + 10 * NE, NE, h->stream);   // ddV, ddP as computed
  pgf::launch_copy(tf + 9 * NE, tf + 11 * NE, NE, h->stream);
  pgf::launch_copy(tf + 12 * NE, tf + 6 * NE, NE, h->stream);   // dV, dP the trials start from
  pgf::launch_copy(tf + 13 * NE, tf + 8 * NE, NE, h->stream);
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_tf_accept(pgf_handle h, double gamma) {
  if (!h || !h->tf) return PGF_EINVAL;
  const int64_t NE = h->n_elem;
  double *tf = h->d_tf;
  for (int v = 0; v < 2; ++v) {  // V, then P: ddX = gamma ddX_full; dX = dX(start) + ddX
    double *dX = tf + (6 + 2 * v) * NE, *ddX = dX + NE, *full = tf + (10 + v) * NE, *start = tf + (12 + v) * NE;
    pgf::launch_fill(ddX, NE, 0.0, h->stream);
    pgf::launch_axpy(ddX, full, gamma, NE, h->stream);  // 0 + gamma x: the rounded product
    pgf::launch_copy(dX, start, NE, h->stream);
    pgf::launch_axpy(dX, ddX, 1.0, NE, h->stream);
  }
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_tf_get(pgf_handle h, int32_t field, double *host_out) {
  if (!h || !h->tf || !host_out || field < 0 || field > PGF_TF_DDP) return PGF_EINVAL;
  std::vector<double> tmp(h->n_elem);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (hipMemcpy(tmp.data(), h->d_tf + int64_t(field) * h->n_elem, sizeof(double) * h->n_elem,
                hipMemcpyDeviceToHost) != hipSuccess)
    return PGF_EHIP;
  for (auto &b : h->blocks)
    for (int64_t k = 0; k < b.ne; ++k) host_out[b.elems[k]] = tmp[b.e_base + k];
  return PGF_OK;
}

int pgf_tf_set(pgf_handle h, int32_t field, const double *host_in) {
  if (!h || !h->tf || !host_in || field < 0 || field > PGF_TF_DDP) return PGF_EINVAL;
  std::vector<double> tmp(h->n_elem);
  for (auto &b : h->blocks)
    for (int64_t k = 0; k < b.ne; ++k) tmp[b.e_base + k] = host_in[b.elems[k]];
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (hipMemcpy(h->d_tf + int64_t(field) * h->n_elem, tmp.data(), sizeof(double) * h->n_elem,
                hipMemcpyHostToDevice) != hipSuccess)
    return PGF_EHIP;
  return PGF_OK;
}

int pgf_state_ncomp(pgf_handle h, int32_t field, int32_t *ncomp) {
  if (!h || !ncomp) return PGF_EINVAL;
  if (field == PGF_STATE_COHESIVE_NP1 || field == PGF_STATE_COHESIVE_N) {
    *ncomp = h->n_cqp ? 2 : 0;
    return PGF_OK;
  }
  if (field == PGF_STATE_ITERS) {
    *ncomp = h->has_cp ? 1 : 0;
    return PGF_OK;
  }
  double *buf;
  return field_buffer(h, field, &buf, ncomp);
}

int pgf_get_state(pgf_handle h, int32_t field, double *host_out) {
  if (!h || !host_out) return PGF_EINVAL;
  if (field == PGF_STATE_COHESIVE_NP1 || field == PGF_STATE_COHESIVE_N) {
    if (!h->n_cqp) return PGF_EINVAL;
    const int64_t NC = h->n_cqp;
    std::vector<double> tmp(2 * NC);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
    if (hipMemcpy(tmp.data(), h->d_cv[field == PGF_STATE_COHESIVE_NP1 ? 0 : 1], sizeof(double) * 2 * NC,
                  hipMemcpyDeviceToHost) != hipSuccess)
      return PGF_EHIP;
    for (int c = 0; c < 2; ++c)
      for (int64_t q = 0; q < NC; ++q) host_out[c * NC + h->cqp_mesh[q]] = tmp[c * NC + q];
    return PGF_OK;
  }
  if (field == PGF_STATE_ITERS) {
    if (!h->has_cp) return PGF_EINVAL;
    std::vector<int32_t> it(h->n_qp);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
    if (hipMemcpy(it.data(), h->d_iters, sizeof(int32_t) * h->n_qp, hipMemcpyDeviceToHost) != hipSuccess)
      return PGF_EHIP;
    // the local Newton keeps the start it took in the high byte (pgf_cp_start_stats)
    for (int64_t q = 0; q < h->n_qp; ++q) host_out[h->qp_mesh[q]] = double(it[q] & 0xffffff);
    return PGF_OK;
  }
  double *buf;
  int32_t nc;
  int rc = field_buffer(h, field, &buf, &nc);
  if (rc || nc == 0) return rc ? rc : PGF_EINVAL;
  const int64_t NQ = h->n_qp;
  std::vector<double> tmp(size_t(nc) * NQ);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (hipMemcpy(tmp.data(), buf, tmp.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return PGF_EHIP;
  for (int32_t c = 0; c < nc; ++c)
    for (int64_t q = 0; q < NQ; ++q) host_out[c * NQ + h->qp_mesh[q]] = tmp[c * NQ + q];
  return PGF_OK;
}

int pgf_set_state(pgf_handle h, int32_t field, const double *host_in) {
  if (!h || !host_in) return PGF_EINVAL;
  if (field == PGF_STATE_COHESIVE_NP1 || field == PGF_STATE_COHESIVE_N) {
    if (!h->n_cqp) return PGF_EINVAL;
    const int64_t NC = h->n_cqp;
    std::vector<double> tmp(2 * NC);
    for (int c = 0; c < 2; ++c)
      for (int64_t q = 0; q < NC; ++q) tmp[c * NC + q] = host_in[c * NC + h->cqp_mesh[q]];
    if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
    if (hipMemcpy(h->d_cv[field == PGF_STATE_COHESIVE_NP1 ? 0 : 1], tmp.data(), sizeof(double) * 2 * NC,
                  hipMemcpyHostToDevice) != hipSuccess)
      return PGF_EHIP;
    return PGF_OK;
  }
  double *buf;
  int32_t nc;
  int rc = field_buffer(h, field, &buf, &nc);
  if (rc || nc == 0) return rc ? rc : PGF_EINVAL;
  const int64_t NQ = h->n_qp;
  std::vector<double> tmp(size_t(nc) * NQ);
  for (int32_t c = 0; c < nc; ++c)
    for (int64_t q = 0; q < NQ; ++q) tmp[c * NQ + q] = host_in[c * NQ + h->qp_mesh[q]];
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (hipMemcpy(buf, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return PGF_EHIP;
  return PGF_OK;
}

int pgf_set_time_step(pgf_handle h, double dt) {
  if (!h || !(dt > 0.0)) return PGF_EINVAL;
  h->dt = dt;
  return PGF_OK;
}

int pgf_set_warm_start(pgf_handle h, const double *warm) {
  if (!h) return PGF_EINVAL;
  if (warm && !h->has_cp) return PGF_ENOTSUP;
  h->warm = warm;
  return PGF_OK;
}

int pgf_get_qp_order(pgf_handle h, int64_t *mesh_qp) {
  if (!h || !mesh_qp) return PGF_EINVAL;
  std::copy(h->qp_mesh.begin(), h->qp_mesh.end(), mesh_qp);
  return PGF_OK;
}

int pgf_state_device_ptr(pgf_handle h, int32_t field, double **ptr) {
  if (!h || !ptr) return PGF_EINVAL;
  int32_t nc = 0;
  return field_buffer(h, field, ptr, &nc);
}

int pgf_set_orientations(pgf_handle h, int32_t n_grains, const double *R, const int32_t *elem_grain) {
  if (!h || n_grains < 1 || !R) return PGF_EINVAL;
  if (!h->has_cp) return PGF_ENOTSUP;
  // every host array is built and checked before device memory changes
  std::vector<pgf::DevGrain> g(n_grains);
  std::vector<float> q(size_t(n_grains) * 4);
  for (int32_t k = 0; k < n_grains; ++k) {
    pgf::make_grain(R + 9 * k, h->cp_mat_host[0], h->cp_mat_host[1], h->cp_mat_host[2], g[k]);
    pgf::rot_to_quat(R + 9 * k, q.data() + 4 * k);
  }
  std::vector<std::vector<int32_t>> egs(h->blocks.size()), qgs(h->blocks.size());
  if (elem_grain)
    for (size_t bi = 0; bi < h->blocks.size(); ++bi) {
      const Block &b = h->blocks[bi];
      if (b.model != PGF_MODEL_CRYSTAL_PLASTICITY) continue;
      egs[bi].resize(b.ne);
      qgs[bi].resize(b.ne * b.nqp);
      for (int64_t k = 0; k < b.ne; ++k) {
        const int32_t gk = elem_grain[b.elems[k]];
        if (gk < 0 || gk >= n_grains) return PGF_EINVAL;
        egs[bi][k] = gk;
        for (int ip = 0; ip < b.nqp; ++ip) qgs[bi][k * b.nqp + ip] = gk;
      }
    }
  int rc = drop_graph(h);  // a captured pass holds the grain table's address
  if (rc) return rc;
  hipFree(h->d_grains);
  hipFree(h->d_gquat);
  h->d_grains = nullptr;
  h->d_gquat = nullptr;
  if ((rc = pgf::upload(&h->d_grains, g.data(), g.size()))) return rc;
  if ((rc = pgf::upload(&h->d_gquat, q.data(), q.size()))) return rc;
  for (size_t bi = 0; bi < h->blocks.size(); ++bi) {
    if (egs[bi].empty()) continue;
    const Block &b = h->blocks[bi];
    if (hipMemcpy(b.d_egrain, egs[bi].data(), egs[bi].size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.d_qgrain, qgs[bi].data(), qgs[bi].size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
      return PGF_EHIP;
  }
  return PGF_OK;
}

int pgf_set_cohesive_props(pgf_handle h, int32_t n_props, const pgf_cohesive *props, const int32_t *coel_mat) {
  if (!h || n_props < 1 || !props) return PGF_EINVAL;
  if (!h->n_coel) return PGF_EINVAL;
  std::vector<pgf::DevCohesive> dp(n_props);
  for (int32_t i = 0; i < n_props; ++i) {
    if (!(props[i].chi_c > 0.0) || !(props[i].alpha != 0.0)) return PGF_EINVAL;
    dp[i] = {props[i].sig_c, props[i].chi_c, props[i].beta, props[i].alpha};
  }
  std::vector<std::vector<int32_t>> mats(h->cblocks.size());
  for (size_t bi = 0; bi < h->cblocks.size(); ++bi) {
    const pgf::CohBlock &cb = h->cblocks[bi];
    mats[bi].resize(cb.nc);
    for (int64_t k = 0; k < cb.nc; ++k) {
      const int32_t mi = coel_mat ? coel_mat[cb.elems[k]] : 0;
      if (mi < 0 || mi >= n_props) return PGF_EINVAL;
      mats[bi][k] = mi;
    }
  }
  int rc = drop_graph(h);  // a captured pass holds the property table's address
  if (rc) return rc;
  for (size_t bi = 0; bi < h->cblocks.size(); ++bi)
    if (hipMemcpy(h->cblocks[bi].d_mat, mats[bi].data(), mats[bi].size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
        hipSuccess)
      return PGF_EHIP;
  hipFree(h->d_cprops);
  h->d_cprops = nullptr;
  return pgf::upload(&h->d_cprops, dp.data(), dp.size());
}

int pgf_get_cohesive_sizes(pgf_handle h, int64_t *n_cqp, int32_t *n_colors) {
  if (!h) return PGF_EINVAL;
  if (n_cqp) *n_cqp = h->n_cqp;
  if (n_colors) *n_colors = h->n_ccolors;
  return PGF_OK;
}

int pgf_cp_log_samples(pgf_handle h, float *X, float *T) {
  if (!h || !X || !T) return PGF_EINVAL;
  if (!h->has_cp) return PGF_ENOTSUP;
  pgf::CPArgs c = cp_args_all(h, nullptr);
  pgf::launch_cp_features(c, h->d_gquat, nullptr, nullptr, X, h->stream);
  pgf::launch_cp_targets(c, T, h->stream);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  return hipStreamSynchronize(h->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_set_surrogate(pgf_handle h, pgf_mlp m, const float *shift, const float *scale, const float *out_scale) {
  if (!h) return PGF_EINVAL;
  if (!m) {
    h->sur = nullptr;
    return PGF_OK;
  }
  if (!h->has_cp) return PGF_ENOTSUP;
  if (!shift || !scale || !out_scale) return PGF_EINVAL;
  std::vector<float> nrm(80);
  std::copy(shift, shift + 34, nrm.begin());
  std::copy(scale, scale + 34, nrm.begin() + 34);
  std::copy(out_scale, out_scale + 12, nrm.begin() + 68);
  for (int i = 68; i < 80; ++i)
    if (!(nrm[i] != 0.0f)) return PGF_EINVAL;
  int rc;
  if (!h->d_norm && (rc = pgf::dalloc(&h->d_norm, 80))) return rc;
  if (!h->d_feat && (rc = pgf::dalloc(&h->d_feat, size_t(h->n_qp) * 34))) return rc;
  if (!h->d_sur_warm && (rc = pgf::dalloc(&h->d_sur_warm, size_t(h->n_qp) * 11))) return rc;
  if (hipMemcpy(h->d_norm, nrm.data(), 80 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return PGF_EHIP;
  h->sur = m;
  return PGF_OK;
}

int pgf_solve(pgf_handle h, const double *vals, const double *b, double *x, double tol, int32_t maxit,
              int32_t *iters, double *relres) {
  if (!h || !vals || !b || !x || !(tol > 0.0) || maxit < 1) return PGF_EINVAL;
  const int64_t n = h->ndof;
  if (!h->d_solve_work) {
    int rc = pgf::dalloc(&h->d_solve_work, size_t(9 * n + 1025));
    if (rc) return rc;
  }
  int it = 0;
  double rr = 0.0;
  if (pgf::bicgstab(h->stream, h->d_Ap, h->d_Ai, vals, b, x, n, tol, maxit, &it, &rr, h->d_solve_work))
    return PGF_EHIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return (rr <= tol) ? PGF_OK : 1;  // 1 = not converged within maxit
}

int pgf_set_graph(pgf_handle h, int on) {
  if (!h) return PGF_EINVAL;
  h->use_graph = on != 0;
  if (!h->use_graph && h->gexec) {
    hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  return PGF_OK;
}

int pgf_set_assembly(pgf_handle h, int32_t mode) {
  if (!h || (mode != PGF_ASSEMBLY_GATHER && mode != PGF_ASSEMBLY_COLOURED)) return PGF_EINVAL;
  if (mode == PGF_ASSEMBLY_GATHER && (pgf::gather_waves(h->maxlen) < 1 || !h->gather_fits)) return PGF_ENOTSUP;
  if (mode == h->assembly) return PGF_OK;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return PGF_EHIP;
  if (h->gexec) {  // a captured pass holds the old schedule
    hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  if (mode == PGF_ASSEMBLY_COLOURED && h->d_stage) {  // release the staging memory
    hipFree(h->d_stage);
    h->d_stage = nullptr;
  }
  h->assembly = mode;
  return PGF_OK;
}

int pgf_autotune(pgf_handle h, const pgf_nodal_input *in, double dt, int64_t *chunk_out) {
  if (!h || !in || !in->r) return PGF_EINVAL;
  if (dt > 0.0) h->dt = dt;
  const bool tune_cp = h->cp_nq > 0;
  const bool tune_gather = h->assembly == PGF_ASSEMBLY_GATHER && h->stage_size > 0;
  if (!tune_cp && !tune_gather) {
    if (chunk_out) *chunk_out = h->cp_chunk;
    return PGF_OK;
  }
  double *f = nullptr, *vals = nullptr;
  int rc = pgf::dalloc(&f, size_t(h->ndof));
  if (!rc && tune_gather) rc = pgf::dalloc(&vals, size_t(h->nnz));
  if (rc) {
    hipFree(f);
    return rc;
  }
  const bool prof = h->profiling, graph = h->use_graph;
  if ((rc = pgf_set_profiling(h, 1))) {  // creates the timing events on first use
    hipFree(f);
    hipFree(vals);
    return rc;
  }
  h->use_graph = false;
  // the best of three timed passes (pgf_last_pass_ms) of one configuration
  auto time_pass = [&](int mode, double *vv, double &ms) -> int {
    ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      const int r = run_pass(h, in, mode, f, vv);
      if (r < 0) return r;
      ms = std::min(ms, double(h->last_ms));
    }
    return PGF_OK;
  };
  // 1. points per wavefront of the crystal-plasticity work queue, on integrating passes
  if (tune_cp) {
    const int64_t cand[] = {1, 64, 128, 256, 512, 1024, 2048};  // 1: one global queue
    int64_t best = h->cp_chunk;
    double best_ms = 1e300;
    for (int64_t c : cand) {
      h->cp_chunk = c;
      double ms;
      if ((rc = time_pass(pgf::MODE_INTEGRATE | pgf::MODE_RESIDUAL, nullptr, ms)) < 0) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    h->cp_chunk = best;
  }
  // 2. the gather's tile: nodes per workgroup x staging blocks in flight, on tangent passes
  //    (every configuration gives the same sums in the same order)
  if (rc >= 0 && tune_gather) {
    const int32_t waves[] = {1, 2, 4}, batches[] = {2, 4, 8};
    int32_t bw = h->gather_waves, bb = h->gather_batch;
    double best_ms = 1e300;
    for (int32_t w : waves)
      for (int32_t b : batches) {
        if (w > pgf::gather_waves(h->maxlen)) continue;
        h->gather_waves = w;
        h->gather_batch = b;
        double ms;
        if ((rc = time_pass(pgf::MODE_RESIDUAL | pgf::MODE_TANGENT, vals, ms)) < 0) break;
        if (ms < best_ms) {
          best_ms = ms;
          bw = w;
          bb = b;
        }
      }
    h->gather_waves = bw;
    h->gather_batch = bb;
  }
  hipFree(f);
  hipFree(vals);
  h->profiling = prof;
  h->use_graph = graph;
  if (chunk_out) *chunk_out = h->cp_chunk;
  return rc < 0 ? rc : PGF_OK;
}

int pgf_get_tuning(pgf_handle h, int64_t *cp_chunk, int32_t *gather_waves, int32_t *gather_batch) {
  if (!h) return PGF_EINVAL;
  if (cp_chunk) *cp_chunk = h->cp_chunk;
  if (gather_waves) *gather_waves = std::min(h->gather_waves, std::max(pgf::gather_waves(h->maxlen), 1));
  if (gather_batch) *gather_batch = h->gather_batch;
  return PGF_OK;
}

int pgf_set_tuning(pgf_handle h, int64_t cp_chunk, int32_t gather_waves, int32_t gather_batch) {
  if (!h) return PGF_EINVAL;
  if (cp_chunk > 1 && cp_chunk < 64) return PGF_EINVAL;
  if (gather_waves > 0 && !(gather_waves == 1 || gather_waves == 2 || gather_waves == 4)) return PGF_EINVAL;
  if (gather_batch > 0 && !(gather_batch == 2 || gather_batch == 4 || gather_batch == 8)) return PGF_EINVAL;
  if (cp_chunk > 0) h->cp_chunk = cp_chunk;
  if (gather_waves > 0) h->gather_waves = gather_waves;
  if (gather_batch > 0) h->gather_batch = gather_batch;
  return PGF_OK;
}

}  // extern "C"
