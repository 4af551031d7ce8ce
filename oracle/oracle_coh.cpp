// CPU oracle, cohesive interface elements -- TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// Restates, for 6-node (3 + 3) zero-thickness interface elements with the exponential
// (Needleman) traction-separation law:
//   Needleman_traction / Needleman_tangents     src/cohesive_potentials.cc:200-416
//   resid_co_elem / stiff_mat_coh               src/cohesive_element.cc:270-640
//   update_state_variables_co                   src/cohesive_element.cc:700-870
//   base_vec, get_jump, mean_map, dN_dxb, dN3_xy src/cohesive_element_utils.cc:139-460
//   int_tria_I (TYPE 3: edge midpoints, w = 0.333333333333333/2) :485-560
// Total Lagrangian: positions are X + u with u the total displacement, which equals the
// reference's updated coordinates x(n) plus the step increment; the area Jacobian is taken
// on the reference mean surface in its own (e1, e2, n) frame, as tran_coord + dN3_xy do.
// State per QP (vars): X_eff_max and t_max at step n; update at commit as the reference
// does it: if Xn > 0 and |Xn| >= X_eff_max then (X_eff_max, t_max) = (|Xn|, |t|) -- the
// effective opening of the update uses the element's beta member, which the reference
// never sets (0), so it is the normal opening.
//
// Pinning (tests/test_oracle_cohesive.py): pure opening peaks at sig_c at chi_c, unloading
// is secant to the origin, mat_tan = FD of the traction wrt the jump and geo_tan = FD wrt
// the normal, total force of a uniform opening = traction x area, K = FD of f on elements
// whose current edge vectors are unit (where the reference's dN_dxb is the exact
// derivative of the normal), f = 0 for a closed interface.
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

const double kMaxHy = 25.0;

void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void normalize(double *a) {
  const double s = std::sqrt(dot3(a, a));
  for (int i = 0; i < 3; ++i) a[i] /= s;
}

// t_hat = t / X_eff on the loading envelope, secant on unloading
double t_hat_of(const orc_cohesive &p, double X_eff, double Xmax, double tmax) {
  if (X_eff >= Xmax || Xmax < p.chi_c) return p.sig_c / p.chi_c * std::exp(1.0 - X_eff / p.chi_c);
  return tmax / Xmax;
}

void jump_nt(const double *jump, const double *n, double *Xn, double *Xs) {
  *Xn = dot3(jump, n);
  double s = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double d = jump[i] - *Xn * n[i];
    s += d * d;
  }
  *Xs = std::sqrt(s);
}

// Surface shape functions and their (xi, eta) derivatives (shape_2DC, dN_2D,
// src/cohesive_element_utils.cc:68-137): linear triangle (nn = 3) or bilinear quad (nn = 4).
void shape2d(int nn, double k, double e, double *N, double *Nk, double *Ne) {
  if (nn == 3) {
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

// Interface quadrature (integrateC): triangle = int_tria_I type 3 (edge midpoints, weight
// 0.333333333333333/2), quad = 2x2 Gauss (intpoints_2, +-0.577350269189626) with
// ip = i*2 + j, xi = g[i], eta = g[j].
int coh_qp(int nn, int ip, double *k, double *e, double *w) {
  if (nn == 3) {
    const double qk[3] = {0.5, 0.5, 0.0}, qe[3] = {0.5, 0.0, 0.5};
    *k = qk[ip];
    *e = qe[ip];
    *w = 0.333333333333333 / 2.;
    return 3;
  }
  const double g[2] = {-0.577350269189626, 0.577350269189626};
  *k = g[ip / 2];
  *e = g[ip % 2];
  *w = 1.0;
  return 4;
}

}  // namespace

extern "C" {

int orc_coh_traction(const orc_cohesive *pp, const double *jump, const double *n, const double *vars,
                     double *t) {
  const orc_cohesive &p = *pp;
  double Xn, Xs;
  jump_nt(jump, n, &Xn, &Xs);
  const double b2 = p.beta * p.beta;
  const double X_eff = std::sqrt(b2 * Xs * Xs + Xn * Xn);
  const double th = t_hat_of(p, X_eff, vars[0], vars[1]);
  int err = 0;
  if (Xn > 0 && X_eff > p.chi_c && th < p.sig_c / 100.0) err = 100;  // element killed
  double tn;
  if (Xn >= 0.0) {
    tn = th * Xn;
  } else {  // contact
    double hy = p.alpha * Xn / p.chi_c;
    if (std::fabs(hy) > kMaxHy) hy = hy / std::fabs(hy) * kMaxHy;
    tn = p.sig_c / p.alpha * std::exp(1.0 - vars[0] / p.chi_c) * std::sinh(hy);
  }
  for (int i = 0; i < 3; ++i) t[i] = tn * n[i] + th * b2 * (jump[i] - Xn * n[i]);
  return err;
}

int orc_coh_tangents(const orc_cohesive *pp, const double *jump, const double *n, const double *vars,
                     double *mat_tan, double *geo_tan) {
  const orc_cohesive &p = *pp;
  double Xn, Xs;
  jump_nt(jump, n, &Xn, &Xs);
  const double b2 = p.beta * p.beta;
  const double X_eff = std::sqrt(b2 * Xs * Xs + Xn * Xn);
  const double Xmax = vars[0], tmax = vars[1];
  const double th = t_hat_of(p, X_eff, Xmax, tmax);
  const double dth = (X_eff >= Xmax || Xmax < p.chi_c) ? -th / p.chi_c : tmax / Xmax;
  double Cs[3] = {0, 0, 0}, dCdX[3] = {0, 0, 0}, dCdN[3] = {0, 0, 0};
  if (X_eff > 0) {
    for (int i = 0; i < 3; ++i) {
      Cs[i] = jump[i] - Xn * n[i];
      dCdX[i] = (b2 * jump[i] + (1 - b2) * Xn * n[i]) / X_eff;
      dCdN[i] = (1 - b2) * Xn / X_eff * jump[i];
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double dij = (i == j) ? 1.0 : 0.0;
      mat_tan[i * 3 + j] = dth * b2 * Cs[i] * dCdX[j] + th * b2 * (dij - n[i] * n[j]);
      geo_tan[i * 3 + j] = dth * b2 * Cs[i] * dCdN[j] - th * b2 * (n[i] * jump[j] + Xn * dij);
    }
  if (Xn >= 0.0) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double dij = (i == j) ? 1.0 : 0.0;
        mat_tan[i * 3 + j] += dth * Xn * n[i] * dCdX[j] + th * n[i] * n[j];
        geo_tan[i * 3 + j] += dth * Xn * n[i] * dCdN[j] + th * n[i] * jump[j] + th * Xn * dij;
      }
  } else {
    // contact: the exponential uses X_eff_max, whose derivative wrt Xn is zero here
    const double ex = std::exp(1.0 - Xmax / p.chi_c);
    double hy = Xn * p.alpha / p.chi_c;
    if (std::fabs(hy) > kMaxHy) hy = hy / std::fabs(hy) * kMaxHy;
    const double tnh = p.sig_c / p.alpha * ex * std::sinh(hy);
    const double dtnh = p.sig_c / p.chi_c * ex * std::cosh(hy);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double dij = (i == j) ? 1.0 : 0.0;
        mat_tan[i * 3 + j] += dtnh * n[i] * n[j];
        geo_tan[i * 3 + j] += dtnh * n[i] * jump[j] + tnh * dij;
      }
  }
  return 0;
}

// One interface element of nn = 3 or 4 nodes per side.  x0[2nn][3] reference coordinates
// (nodes 0..nn-1: - side, nn..2nn-1: + side), u[6nn] total displacement, vars_n [nq][2]
// (nq = nn QPs).  fe[6nn], K[6nn*6nn] (row-major, node-major DOFs) are overwritten;
// vars_np1 [nq][2] gets the committed-state update.  Returns the number of failed QPs
// (non-positive area Jacobian); killed QPs contribute nothing, as in the reference.
int orc_coh_element(const orc_cohesive *pp, int nn, const double *x0, const double *u, const double *vars_n,
                    double *fe, double *K, double *vars_np1) {
  if (nn != 3 && nn != 4) return -1;
  double N[4], Nk[4], Ne[4];
  // reference mean surface and its frame at (0, 0) (read_cohe_elem: mean_map + base_vec)
  double xr[4][3], g1[3] = {0, 0, 0}, g2[3] = {0, 0, 0};
  for (int a = 0; a < nn; ++a)
    for (int i = 0; i < 3; ++i) xr[a][i] = 0.5 * (x0[(nn + a) * 3 + i] + x0[a * 3 + i]);
  shape2d(nn, 0.0, 0.0, N, Nk, Ne);
  for (int a = 0; a < nn; ++a)
    for (int i = 0; i < 3; ++i) {
      g1[i] += Nk[a] * xr[a][i];
      g2[i] += Ne[a] * xr[a][i];
    }
  double E1[3] = {g1[0], g1[1], g1[2]}, En[3], E2[3];
  cross(g1, g2, En);
  normalize(E1);
  normalize(En);
  cross(En, E1, E2);
  double xl[4], yl[4];
  for (int a = 0; a < nn; ++a) {
    xl[a] = dot3(E1, xr[a]);
    yl[a] = dot3(E2, xr[a]);
  }
  // current mean surface
  double xb[4][3];
  for (int a = 0; a < nn; ++a)
    for (int i = 0; i < 3; ++i)
      xb[a][i] = 0.5 * (x0[(nn + a) * 3 + i] + x0[a * 3 + i]) + 0.5 * (u[(nn + a) * 3 + i] + u[a * 3 + i]);
  double Kt[3][3][4][4], Kn[3][3][4][4], Rc[3][4];
  std::memset(Kt, 0, sizeof(Kt));
  std::memset(Kn, 0, sizeof(Kn));
  std::memset(Rc, 0, sizeof(Rc));
  int err = 0;
  for (int ip = 0; ip < nn; ++ip) {
    double qk, qe, w;
    coh_qp(nn, ip, &qk, &qe, &w);
    shape2d(nn, qk, qe, N, Nk, Ne);
    // base vectors of the current mean surface (base_vec)
    double e1[3] = {0, 0, 0}, e2h[3] = {0, 0, 0}, n[3];
    for (int a = 0; a < nn; ++a)
      for (int i = 0; i < 3; ++i) {
        e1[i] += Nk[a] * xb[a][i];
        e2h[i] += Ne[a] * xb[a][i];
      }
    cross(e1, e2h, n);
    normalize(e1);
    normalize(e2h);
    normalize(n);
    // area Jacobian on the reference surface in its local frame (dN3_xy)
    double dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0;
    for (int a = 0; a < nn; ++a) {
      dx0 += Nk[a] * xl[a];
      dx1 += Ne[a] * xl[a];
      dy0 += Nk[a] * yl[a];
      dy1 += Ne[a] * yl[a];
    }
    const double J = dx0 * dy1 - dx1 * dy0;
    if (!(J > 0.0)) {
      ++err;
      continue;
    }
    double jump[3] = {0, 0, 0};  // get_jump
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < nn; ++a)
        jump[i] += (x0[(nn + a) * 3 + i] - x0[a * 3 + i]) * N[a] + (u[(nn + a) * 3 + i] - u[a * 3 + i]) * N[a];
    double t[3], mt[9], gt[9];
    const double *vn = vars_n + 2 * ip;
    const int kill = orc_coh_traction(pp, jump, n, vn, t);
    orc_coh_tangents(pp, jump, n, vn, mt, gt);
    {  // committed-state update (update_state_variables_co)
      const double Xn = dot3(jump, n);
      const double Xxi = std::fabs(Xn);
      const double txi = std::sqrt(dot3(t, t));
      vars_np1[2 * ip] = vn[0];
      vars_np1[2 * ip + 1] = vn[1];
      if (Xxi >= vn[0] && Xn > 0.0) {
        vars_np1[2 * ip] = Xxi;
        vars_np1[2 * ip + 1] = txi;
      }
    }
    if (kill >= 100) continue;
    // d n / d xb (dN_dxb, evaluated with the unit vectors e1, e2h)
    double a3[3];
    cross(e1, e2h, a3);
    const double BB = std::sqrt(dot3(a3, a3));
    double CC[3][3][4], Nxb[3][3][4];
    for (int Q = 0; Q < 3; ++Q)
      for (int P = 0; P < 3; ++P)
        for (int R = 0; R < nn; ++R) {
          double v = 0.0;  // eps_QPS Nk_R e2h_S + eps_QSP e1_S Ne_R
          for (int S = 0; S < 3; ++S) {
            const double eQPS = (Q == P || P == S || Q == S) ? 0.0 : ((((P - Q + 3) % 3) == 1) ? 1.0 : -1.0);
            v += eQPS * Nk[R] * e2h[S] - eQPS * e1[S] * Ne[R];
          }
          CC[Q][P][R] = v;
        }
    for (int Q = 0; Q < 3; ++Q)
      for (int P = 0; P < 3; ++P)
        for (int R = 0; R < nn; ++R) {
          double v = CC[Q][P][R] / BB;
          for (int S = 0; S < 3; ++S) v -= 1.0 / BB * n[Q] * n[S] * CC[S][P][R];
          Nxb[Q][P][R] = v;
        }
    const double wJ = w * J;
    for (int M = 0; M < 3; ++M)
      for (int C = 0; C < 3; ++C)
        for (int R = 0; R < nn; ++R) {
          double AA = 0.0;  // geo_tan . dn/dxb of column node R
          for (int U = 0; U < 3; ++U) AA += gt[M * 3 + U] * Nxb[U][C][R];
          for (int P = 0; P < nn; ++P) {
            Kt[M][C][P][R] += wJ * mt[M * 3 + C] * N[P] * N[R];
            Kn[M][C][P][R] += wJ * 0.5 * AA * N[P];
          }
        }
    for (int M = 0; M < 3; ++M)
      for (int a = 0; a < nn; ++a) Rc[M][a] += wJ * t[M] * N[a];
  }
  const int h = 3 * nn, nd = 6 * nn;
  if (fe)
    for (int a = 0; a < nn; ++a)
      for (int M = 0; M < 3; ++M) {
        fe[a * 3 + M] = -Rc[M][a];
        fe[h + a * 3 + M] = Rc[M][a];
      }
  if (K)
    for (int M = 0; M < 3; ++M)
      for (int C = 0; C < 3; ++C)
        for (int P = 0; P < nn; ++P)
          for (int R = 0; R < nn; ++R) {
            const double kt = Kt[M][C][P][R], kn = Kn[M][C][P][R];
            K[(P * 3 + M) * nd + R * 3 + C] = kt - kn;
            K[(h + P * 3 + M) * nd + h + R * 3 + C] = kt + kn;
            K[(P * 3 + M) * nd + h + R * 3 + C] = -kt - kn;
            K[(h + P * 3 + M) * nd + R * 3 + C] = -kt + kn;
          }
  return err;
}

// All interface elements of a mesh, ADDED into f[ndof] and vals[nnz] (PLoc_Sparse /
// fd_res_assemble drop rows and columns of non-free DOFs).  vars_n / vars_np1 are
// [2][n_cqp] (X_eff_max, t_max), QPs numbered element by element (nn per element).
int orc_assemble_coh(const double *X, int64_t n_coel, const int64_t *coel_ptr, const int64_t *coel_conn,
                     const int32_t *coel_mat, const orc_cohesive *props, const int64_t *id, int64_t ndof,
                     const int32_t *Ap, const int32_t *Ai, const double *r, const double *d_r, const double *defl,
                     const double *defl_d, const double *vars_n, double *vars_np1, double *f, double *vals) {
  (void)ndof;
  int err = 0;
  int64_t NQ = 0;
  for (int64_t c = 0; c < n_coel; ++c) NQ += (coel_ptr[c + 1] - coel_ptr[c]) / 2;
  int64_t q0 = 0;
  for (int64_t c = 0; c < n_coel; ++c) {
    const int nnt = int(coel_ptr[c + 1] - coel_ptr[c]), nn = nnt / 2, nd = 3 * nnt;
    if (nnt != 6 && nnt != 8) return -1;
    const int64_t *nod = coel_conn + coel_ptr[c];
    double x0[24], u[24];
    int64_t cn[24];
    for (int a = 0; a < nnt; ++a)
      for (int b = 0; b < 3; ++b) {
        x0[a * 3 + b] = X[nod[a] * 3 + b];
        const int64_t dof = id[nod[a] * 3 + b];
        cn[a * 3 + b] = dof;
        const int64_t aid = (dof < 0 ? -dof : dof) - 1;
        if (dof == 0) u[a * 3 + b] = 0.0;
        else if (dof > 0) u[a * 3 + b] = r[aid] + (d_r ? d_r[aid] : 0.0);
        else u[a * 3 + b] = (defl ? defl[aid] : 0.0) + (defl_d ? defl_d[aid] : 0.0);
      }
    double vn[8], v1[8], fe[24], K[576];
    for (int q = 0; q < nn; ++q) {
      vn[2 * q] = vars_n[q0 + q];
      vn[2 * q + 1] = vars_n[NQ + q0 + q];
    }
    err += orc_coh_element(&props[coel_mat ? coel_mat[c] : 0], nn, x0, u, vn, fe, K, v1);
    if (vars_np1)
      for (int q = 0; q < nn; ++q) {
        vars_np1[q0 + q] = v1[2 * q];
        vars_np1[NQ + q0 + q] = v1[2 * q + 1];
      }
    q0 += nn;
    for (int i = 0; i < nd; ++i) {
      if (cn[i] <= 0) continue;
      const int64_t row = cn[i] - 1;
      if (f) f[row] += fe[i];
      if (!vals) continue;
      for (int j = 0; j < nd; ++j) {
        if (cn[j] <= 0) continue;
        const int32_t col = int32_t(cn[j] - 1);
        int32_t lo = Ap[row], hi = Ap[row + 1];
        while (lo < hi) {
          const int32_t mid = (lo + hi) / 2;
          if (Ai[mid] < col) lo = mid + 1;
          else hi = mid;
        }
        if (lo >= Ap[row + 1] || Ai[lo] != col) return -2;
        vals[lo] += K[i * nd + j];
      }
    }
  }
  return err;
}

}  // extern "C"
