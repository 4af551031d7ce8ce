// Interface (cohesive) element kernel for gfx950 (FP64).
//
// Zero-thickness elements of NN = 3 (triangle) or 4 (quad) nodes per side with the
// exponential traction-separation law (Needleman_traction / Needleman_tangents,
// src/cohesive_potentials.cc:200-416), residual and tangent as resid_co_elem /
// stiff_mat_coh (src/cohesive_element.cc:270-640):
//   fe(-) = -Rc, fe(+) = +Rc,            Rc_Ma = sum_q wJ t_M N_a
//   K(s,s') = s s' Kt + s Kn  (s = -1 for the - side, +1 for the + side)
//   Kt_MN,PR = sum_q wJ mat_tan_MN N_P N_R,  Kn_MN,PR = sum_q wJ/2 (geo_tan . dn/dxb_R)_MN N_P
// Positions are X + u (total displacement); the area Jacobian is that of the reference mean
// surface in its own frame; the normal, jump and dn/dxb are evaluated on the current mean
// surface.  State per QP: (X_eff_max, t_max) at n, updated into n+1 by integrating passes
// (update_state_variables_co, src/cohesive_element.cc:700-870; the update's effective
// opening is the normal opening because the element's beta member is never set).
//
// Work mapping: one 32-lane group per element, 8 elements per 256-thread workgroup.
// Phase 0: lanes < 2*NN gather nodes.  Phase 1: lanes < NN compute one QP each into LDS.
// Phase 2: lane r < 6*NN owns element row r: its residual entry and its K row, written to
// the node's staging block (gather mode) or scattered with the same first-touch offsets as
// the volume elements (coloured mode: interface colours are launched after the volume
// colours).
#include <hip/hip_runtime.h>

#include <cmath>

#include "pgf_device.h"

namespace pgf {
namespace {

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3d(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void unit3(double *a) {
  const double s = sqrt(dot3d(a, a));
  a[0] /= s;
  a[1] /= s;
  a[2] /= s;
}

template <int NN>
__device__ __forceinline__ void shape2d(double k, double e, double *N, double *Nk, double *Ne) {
  if constexpr (NN == 3) {
    N[0] = 1.0 - k - e;
    N[1] = k;
    N[2] = e;
    Nk[0] = -1.0; Nk[1] = 1.0; Nk[2] = 0.0;
    Ne[0] = -1.0; Ne[1] = 0.0; Ne[2] = 1.0;
  } else {
    N[0] = 0.25 * (1. - k) * (1. - e);
    N[1] = 0.25 * (1. + k) * (1. - e);
    N[2] = 0.25 * (1. + k) * (1. + e);
    N[3] = 0.25 * (1. - k) * (1. + e);
    Nk[0] = 0.25 * (e - 1.0); Ne[0] = 0.25 * (k - 1.0);
    Nk[1] = 0.25 * (1.0 - e); Ne[1] = 0.25 * (-1.0 - k);
    Nk[2] = 0.25 * (1.0 + e); Ne[2] = 0.25 * (1.0 + k);
    Nk[3] = 0.25 * (-1.0 - e); Ne[3] = 0.25 * (1.0 - k);
  }
}

// Interface quadrature: triangle edge midpoints with weight 0.333333333333333/2
// (int_tria_I type 3, src/cohesive_element_utils.cc:485-560); quad 2x2 Gauss
// (intpoints_2, src/elem3d.cc:92-102), ip = i*2 + j.
template <int NN>
__device__ __forceinline__ void coh_qp(int ip, double &k, double &e, double &w) {
  if constexpr (NN == 3) {
    k = (ip == 2) ? 0.0 : 0.5;
    e = (ip == 1) ? 0.0 : 0.5;
    w = 0.333333333333333 / 2.;
  } else {
    k = (ip / 2) ? 0.577350269189626 : -0.577350269189626;
    e = (ip % 2) ? 0.577350269189626 : -0.577350269189626;
    w = 1.0;
  }
}

// Traction and tangents of the exponential law at (jump, n) with state (Xmax, tmax).
// Returns true when the QP is killed (Xn > 0, X_eff > chi_c, t_hat < sig_c/100).
__device__ __forceinline__ bool needleman(const DevCohesive &p, const double *jump, const double *n,
                                          double Xmax, double tmax, double *t, double *mt, double *gt) {
  const double Xn = dot3d(jump, n);
  double Cs[3];
  double xs2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Cs[i] = jump[i] - Xn * n[i];
    xs2 += Cs[i] * Cs[i];
  }
  const double b2 = p.beta * p.beta;
  const double X_eff = sqrt(b2 * xs2 + Xn * Xn);  // sqrt(b2 Xs^2 + Xn^2), Xs = |jump - Xn n|
  const bool loading = (X_eff >= Xmax || Xmax < p.chi_c);
  const double th = loading ? p.sig_c / p.chi_c * exp(1.0 - X_eff / p.chi_c) : tmax / Xmax;
  const double dth = loading ? -th / p.chi_c : tmax / Xmax;
  const bool kill = (Xn > 0 && X_eff > p.chi_c && th < p.sig_c / 100.0);
  double hy = p.alpha * Xn / p.chi_c;
  if (fabs(hy) > 25.0) hy = hy / fabs(hy) * 25.0;
  const double ex = exp(1.0 - Xmax / p.chi_c);
  const double tn = (Xn >= 0.0) ? th * Xn : p.sig_c / p.alpha * ex * sinh(hy);
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = tn * n[i] + th * b2 * Cs[i];
  double dCdX[3] = {0, 0, 0}, dCdN[3] = {0, 0, 0}, Cz[3] = {0, 0, 0};
  if (X_eff > 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      Cz[i] = Cs[i];
      dCdX[i] = (b2 * jump[i] + (1 - b2) * Xn * n[i]) / X_eff;
      dCdN[i] = (1 - b2) * Xn / X_eff * jump[i];
    }
  }
  const double dtnh = p.sig_c / p.chi_c * ex * cosh(hy);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dij = (i == j) ? 1.0 : 0.0;
      double m = dth * b2 * Cz[i] * dCdX[j] + th * b2 * (dij - n[i] * n[j]);
      double g = dth * b2 * Cz[i] * dCdN[j] - th * b2 * (n[i] * jump[j] + Xn * dij);
      if (Xn >= 0.0) {
        m += dth * Xn * n[i] * dCdX[j] + th * n[i] * n[j];
        g += dth * Xn * n[i] * dCdN[j] + th * n[i] * jump[j] + th * Xn * dij;
      } else {
        m += dtnh * n[i] * n[j];
        g += dtnh * n[i] * jump[j] + tn * dij;
      }
      mt[i * 3 + j] = m;
      gt[i * 3 + j] = g;
    }
  return kill;
}

template <int NN, int MODE>
__global__ __launch_bounds__(256) void coh_elem(CohArgs a, int64_t e_begin, int64_t e_end) {
  constexpr int NT = 2 * NN;        // nodes
  constexpr int ND = 3 * NT;        // DOFs
  constexpr int QD = 13 + 10 * NN;  // per-QP data: wJ, N[NN], t[3], mt[9], AA[9][NN]
  constexpr int EPB = 8;
  constexpr bool INTEG = (MODE & MODE_INTEGRATE) != 0;
  constexpr bool RES = (MODE & MODE_RESIDUAL) != 0;
  constexpr bool TAN = (MODE & MODE_TANGENT) != 0;
  constexpr bool DFL = (MODE & MODE_DEFL_LOAD) != 0;  // K times the prescribed increments
  static_assert(!DFL || (TAN && !RES && !INTEG), "MODE_DEFL_LOAD goes with MODE_TANGENT alone");

  __shared__ double s_x[EPB][ND], s_u[EPB][ND];
  __shared__ double s_dl[EPB][DFL ? ND : 1];
  __shared__ int32_t s_dof[EPB][ND];
  __shared__ int32_t s_node[EPB][NT];
  __shared__ double s_q[EPB][NN][QD];
  __shared__ int32_t s_live[EPB][NN];

  const int slot = threadIdx.x / 32;
  const int lane = threadIdx.x % 32;
  const int64_t ie = e_begin + int64_t(blockIdx.x) * EPB + slot;
  const bool active = ie < e_end;
  const int64_t e = (a.elist && active) ? int64_t(a.elist[ie]) : ie;

  // ---- phase 0: gather
  if (active && lane < NT) {
    const int32_t node = a.conn[e * NT + lane];
    s_node[slot][lane] = node;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      s_x[slot][lane * 3 + b] = a.X[int64_t(node) * 3 + b];
      const int32_t d = a.dofid[int64_t(node) * 3 + b];
      s_dof[slot][lane * 3 + b] = d;
      double u = 0.0;  // def_elem_total, src/utils.cc:1391-1412
      double dl = 0.0;
      if (d > 0) {
        u = a.r[d - 1] + (a.d_r ? a.d_r[d - 1] : 0.0);
      } else if (d < 0) {
        const int k = -d - 1;
        dl = a.defl_d ? a.defl_d[k] : 0.0;
        u = (a.defl ? a.defl[k] : 0.0) + (DFL ? 0.0 : dl);  // the prescribed-load pass: K at defl
      }
      s_u[slot][lane * 3 + b] = u;
      if constexpr (DFL) s_dl[slot][lane * 3 + b] = dl;
    }
  }
  __syncthreads();

  // ---- phase 1: one lane per QP
  if (active && lane < NN) {
    const int ip = lane;
    const double *x = s_x[slot];
    const double *u = s_u[slot];
    double N[NN], Nk[NN], Ne[NN];
    // reference mean surface frame (at the element centre, as set up by read_cohe_elem)
    shape2d<NN>(0.0, 0.0, N, Nk, Ne);
    double g1[3] = {0, 0, 0}, g2[3] = {0, 0, 0}, xr[NN][3];
#pragma unroll
    for (int q = 0; q < NN; ++q)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        xr[q][i] = 0.5 * (x[(NN + q) * 3 + i] + x[q * 3 + i]);
        g1[i] += Nk[q] * xr[q][i];
        g2[i] += Ne[q] * xr[q][i];
      }
    double E1[3] = {g1[0], g1[1], g1[2]}, En[3], E2[3];
    cross3(g1, g2, En);
    unit3(E1);
    unit3(En);
    cross3(En, E1, E2);
    double qk, qe, w;
    coh_qp<NN>(ip, qk, qe, w);
    shape2d<NN>(qk, qe, N, Nk, Ne);
    double dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0;
    double e1[3] = {0, 0, 0}, e2h[3] = {0, 0, 0}, jump[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < NN; ++q) {
      const double xl = dot3d(E1, xr[q]), yl = dot3d(E2, xr[q]);
      dx0 += Nk[q] * xl;
      dx1 += Ne[q] * xl;
      dy0 += Nk[q] * yl;
      dy1 += Ne[q] * yl;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double xb = 0.5 * (x[(NN + q) * 3 + i] + x[q * 3 + i]) + 0.5 * (u[(NN + q) * 3 + i] + u[q * 3 + i]);
        e1[i] += Nk[q] * xb;
        e2h[i] += Ne[q] * xb;
        jump[i] += (x[(NN + q) * 3 + i] - x[q * 3 + i]) * N[q] + (u[(NN + q) * 3 + i] - u[q * 3 + i]) * N[q];
      }
    }
    const double J = dx0 * dy1 - dx1 * dy0;
    double n[3];
    cross3(e1, e2h, n);
    unit3(e1);
    unit3(e2h);
    unit3(n);
    const int64_t cq = a.cqp_base + e * NN + ip;
    const double Xmax = a.vars_n[cq], tmax = a.vars_n[a.n_cqp + cq];
    double t[3], mt[9], gt[9];
    const bool kill = needleman(a.props[a.mat[e]], jump, n, Xmax, tmax, t, mt, gt);
    if (INTEG) {  // committed-state candidate
      const double Xn = dot3d(jump, n);
      const bool up = (fabs(Xn) >= Xmax && Xn > 0.0);
      a.vars_np1[cq] = up ? fabs(Xn) : Xmax;
      a.vars_np1[a.n_cqp + cq] = up ? sqrt(dot3d(t, t)) : tmax;
    }
    const bool ok = J > 0.0;
    if (!ok) atomicAdd(a.err, 1);  // negative isoparametric mapping (dN3_xy)
    s_live[slot][ip] = (ok && !kill) ? 1 : 0;
    double *qd = s_q[slot][ip];
    qd[0] = w * J;
#pragma unroll
    for (int q = 0; q < NN; ++q) qd[1 + q] = N[q];
#pragma unroll
    for (int i = 0; i < 3; ++i) qd[1 + NN + i] = t[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) qd[4 + NN + i] = mt[i];
    if (TAN) {
      // dn/dxb (dN_dxb with the unit vectors e1, e2h); AA_MC,R = geo_tan_MU dn_U/dxb_C,R
      double a3[3];
      cross3(e1, e2h, a3);
      const double rB = 1.0 / sqrt(dot3d(a3, a3));
#pragma unroll
      for (int C = 0; C < 3; ++C)
#pragma unroll
        for (int R = 0; R < NN; ++R) {
          double cc[3];
#pragma unroll
          for (int Q = 0; Q < 3; ++Q) {
            // eps_QCS Nk_R e2h_S - eps_QCS e1_S Ne_R
            double v = 0.0;
#pragma unroll
            for (int S = 0; S < 3; ++S) {
              const double eps = (Q == C || C == S || Q == S) ? 0.0 : ((((C - Q + 3) % 3) == 1) ? 1.0 : -1.0);
              v += eps * (Nk[R] * e2h[S] - e1[S] * Ne[R]);
            }
            cc[Q] = v;
          }
          const double ncc = n[0] * cc[0] + n[1] * cc[1] + n[2] * cc[2];
          double dn[3];
#pragma unroll
          for (int Q = 0; Q < 3; ++Q) dn[Q] = (cc[Q] - n[Q] * ncc) * rB;
#pragma unroll
          for (int M = 0; M < 3; ++M)
            qd[13 + NN + (M * 3 + C) * NN + R] = gt[M * 3 + 0] * dn[0] + gt[M * 3 + 1] * dn[1] + gt[M * 3 + 2] * dn[2];
        }
    }
  }
  __syncthreads();
  if (!active || lane >= ND) return;

  // ---- phase 2: element row r = (side, node P, component M)
  const int r = lane;
  const int P = (r / 3) % NN, M = r % 3;
  const double sr = (r >= 3 * NN) ? 1.0 : -1.0;
  const int32_t dof = s_dof[slot][r];
  if (dof <= 0) return;
  const int ra = r / 3;  // row node within the element
  const int16_t *off = a.off + (e * NT + ra) * NT;
  if constexpr (RES) {
    double fr = 0.0;
#pragma unroll
    for (int q = 0; q < NN; ++q) {
      const double *qd = s_q[slot][q];
      fr += s_live[slot][q] ? qd[0] * qd[1 + NN + M] * qd[1 + P] : 0.0;
    }
    fr *= sr;
    if (a.stage) {
      a.stage[a.soff[e * NT + ra] + M * (ND + 1) + ND] = fr;
    } else {
      const bool first = off[ra] <= -2;
      if (first) a.f[dof - 1] = fr;
      else a.f[dof - 1] += fr;
    }
  }
  if constexpr (TAN) {
    double K[ND];
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      const int R = (c / 3) % NN, C = c % 3;
      const double sc = (c >= 3 * NN) ? 1.0 : -1.0;
      double kt = 0.0, kn = 0.0;
#pragma unroll
      for (int q = 0; q < NN; ++q) {
        const double *qd = s_q[slot][q];
        const double wl = s_live[slot][q] ? qd[0] * qd[1 + P] : 0.0;
        kt += wl * qd[4 + NN + M * 3 + C] * qd[1 + R];
        kn += 0.5 * wl * qd[13 + NN + (M * 3 + C) * NN + R];
      }
      K[c] = sr * sc * kt + sr * kn;
    }
    if constexpr (DFL) {  // f_defl row: K row . prescribed increments (src/load.cc:190-215)
      double fl = 0.0;
#pragma unroll
      for (int c = 0; c < ND; ++c) fl += K[c] * s_dl[slot][c];
      if (a.stage) {
        a.stage[a.soff[e * NT + ra] + M * (ND + 1) + ND] = fl;
      } else {
        const bool first = off[ra] <= -2;
        if (first) a.f[dof - 1] = fl;
        else a.f[dof - 1] += fl;
      }
      return;
    }
    if (a.stage) {  // gather mode: row (ra, M) into the node's staging block
      double *srow = a.stage + a.soff[e * NT + ra] + M * (ND + 1);
#pragma unroll
      for (int c = 0; c < ND; ++c) srow[c] = K[c];
      return;
    }
    const int32_t node = s_node[slot][ra];
    int bp = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) bp += (j < M && s_dof[slot][ra * 3 + j] > 0) ? 1 : 0;
    double *row = a.vals + (int64_t(a.row_start[node]) + int64_t(bp) * a.rowlen[node]);
#pragma unroll
    for (int w = 0; w < NT; ++w) {
      const int o = off[w];
      if (o == -1) continue;
      const bool first = o <= -2;
      const int base = first ? (-o - 2) : o;
      int gp = 0;
#pragma unroll
      for (int gg = 0; gg < 3; ++gg) {
        if (s_dof[slot][w * 3 + gg] > 0) {
          double *p = row + base + gp;
          *p = first ? K[w * 3 + gg] : *p + K[w * 3 + gg];
          ++gp;
        }
      }
    }
  }
}

template <int NN>
void launch_nn(int mode, const CohArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  const int64_t n = e1 - e0;
  if (n <= 0) return;
  const unsigned grid = unsigned((n + 7) / 8);
  switch (mode) {
    case MODE_INTEGRATE | MODE_RESIDUAL:
      coh_elem<NN, MODE_INTEGRATE | MODE_RESIDUAL><<<grid, 256, 0, s>>>(a, e0, e1);
      break;
    case MODE_RESIDUAL:
      coh_elem<NN, MODE_RESIDUAL><<<grid, 256, 0, s>>>(a, e0, e1);
      break;
    case MODE_TANGENT:
      coh_elem<NN, MODE_TANGENT><<<grid, 256, 0, s>>>(a, e0, e1);
      break;
    case MODE_INTEGRATE | MODE_RESIDUAL | MODE_TANGENT:
      coh_elem<NN, MODE_INTEGRATE | MODE_RESIDUAL | MODE_TANGENT><<<grid, 256, 0, s>>>(a, e0, e1);
      break;
    case MODE_TANGENT | MODE_DEFL_LOAD:
      coh_elem<NN, MODE_TANGENT | MODE_DEFL_LOAD><<<grid, 256, 0, s>>>(a, e0, e1);
      break;
    default:
      break;
  }
}

}  // namespace

void launch_coh_elem(int nn, int mode, const CohArgs &a, int64_t e0, int64_t e1, hipStream_t s) {
  if (nn == 3) launch_nn<3>(mode, a, e0, e1, s);
  else if (nn == 4) launch_nn<4>(mode, a, e0, e1, s);
}

}  // namespace pgf
