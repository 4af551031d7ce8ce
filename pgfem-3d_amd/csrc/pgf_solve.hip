// Linear solve of the assembled tangent on the device: Jacobi-preconditioned BiCGStab over
// the handle's CSR pattern (the system the reference hands to hypre GMRES + Euclid,
// src/solvers/hypre/Hypre.cc, once per global Newton iteration).  Kernels: CSR SpMV with a
// group of lanes per row, element-wise vector updates and a deterministic two-pass dot
// product (per-block partials, then one block sums them in a fixed order).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "pgf_device.h"

namespace pgf {
namespace {

constexpr int kLanesPerRow = 16;

// y = A x; one 16-lane group per row
__global__ __launch_bounds__(256) void spmv(const int32_t *Ap, const int32_t *Ai, const double *vals,
                                            const double *x, double *y, int64_t n) {
  const int64_t row = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  double s = 0.0;
  if (row < n)
    for (int32_t k = Ap[row] + lane; k < Ap[row + 1]; k += kLanesPerRow) s += vals[k] * x[Ai[k]];
#pragma unroll
  for (int o = kLanesPerRow / 2; o > 0; o >>= 1) s += __shfl_down(s, o, kLanesPerRow);
  if (row < n && lane == 0) y[row] = s;
}

// inverse of the diagonal (Jacobi); zero / non-finite diagonals give 1
__global__ void inv_diag(const int32_t *Ap, const int32_t *Ai, const double *vals, double *dinv, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int32_t k = Ap[i]; k < Ap[i + 1]; ++k)
    if (Ai[k] == i) d = vals[k];
  dinv[i] = (d != 0.0 && isfinite(d)) ? 1.0 / d : 1.0;
}

__global__ void dot_partial(const double *a, const double *b, int64_t n, double *part) {
  __shared__ double sm[256];
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    s += a[i] * b[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void sum_partials(const double *part, int n, double *out) {
  __shared__ double sm[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// y = a x + b y  (elementwise helpers)
__global__ void axpby(double a, const double *x, double b, double *y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
// p = r + beta (p - omega v)
__global__ void bicg_p(const double *r, double beta, double omega, const double *v, double *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
// z = dinv * x
__global__ void jacobi(const double *dinv, const double *x, double *z, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) z[i] = dinv[i] * x[i];
}
// r = b - r
__global__ void residual_from(const double *b, double *r, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

inline unsigned nb(int64_t n) { return unsigned((n + 255) / 256); }

}  // namespace

int bicgstab(hipStream_t s, const int32_t *Ap, const int32_t *Ai, const double *vals, const double *b, double *x,
             int64_t n, double tol, int maxit, int *iters, double *relres, double *work) {
  // work: 9 vectors of n + 1024 partials + 1 scalar
  double *r = work, *rh = r + n, *p = rh + n, *v = p + n, *ph = v + n, *sv = ph + n, *sh = sv + n, *t = sh + n,
         *dinv = t + n, *part = dinv + n, *sc = part + 1024;
  const unsigned gs = unsigned(std::min<int64_t>(1024, (n + 255) / 256));
  // the first failed runtime call; a failed dot product reads NaN, which ends the iteration
  hipError_t err = hipSuccess;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && err == hipSuccess) err = e;
    return e == hipSuccess;
  };
  auto dot = [&](const double *a, const double *c) {
    dot_partial<<<gs, 256, 0, s>>>(a, c, n, part);
    sum_partials<<<1, 256, 0, s>>>(part, int(gs), sc);
    double h = 0.0;
    if (!chk(hipMemcpyAsync(&h, sc, sizeof(double), hipMemcpyDeviceToHost, s)) || !chk(hipStreamSynchronize(s)))
      return std::nan("");
    return h;
  };
  const unsigned sg = unsigned((n * kLanesPerRow + 255) / 256);
  inv_diag<<<nb(n), 256, 0, s>>>(Ap, Ai, vals, dinv, n);
  spmv<<<sg, 256, 0, s>>>(Ap, Ai, vals, x, r, n);
  residual_from<<<nb(n), 256, 0, s>>>(b, r, n);
  chk(hipMemcpyAsync(rh, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  chk(hipMemsetAsync(p, 0, n * sizeof(double), s));
  chk(hipMemsetAsync(v, 0, n * sizeof(double), s));
  const double bn = std::sqrt(dot(b, b));
  const double bnorm = bn > 0.0 ? bn : 1.0;
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  double res = std::sqrt(dot(r, r)) / bnorm;
  int it = 0;
  while (res > tol && it < maxit) {
    ++it;
    const double rho1 = dot(rh, r);
    if (rho1 == 0.0 || !std::isfinite(rho1)) break;
    const double beta = (rho1 / rho) * (alpha / omega);
    bicg_p<<<nb(n), 256, 0, s>>>(r, beta, omega, v, p, n);
    jacobi<<<nb(n), 256, 0, s>>>(dinv, p, ph, n);
    spmv<<<sg, 256, 0, s>>>(Ap, Ai, vals, ph, v, n);
    const double rv = dot(rh, v);
    if (rv == 0.0 || !std::isfinite(rv)) break;
    alpha = rho1 / rv;
    chk(hipMemcpyAsync(sv, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    axpby<<<nb(n), 256, 0, s>>>(-alpha, v, 1.0, sv, n);  // s = r - alpha v
    jacobi<<<nb(n), 256, 0, s>>>(dinv, sv, sh, n);
    spmv<<<sg, 256, 0, s>>>(Ap, Ai, vals, sh, t, n);
    const double tt = dot(t, t);
    omega = (tt > 0.0) ? dot(t, sv) / tt : 0.0;
    axpby<<<nb(n), 256, 0, s>>>(alpha, ph, 1.0, x, n);  // x += alpha ph + omega sh
    axpby<<<nb(n), 256, 0, s>>>(omega, sh, 1.0, x, n);
    chk(hipMemcpyAsync(r, sv, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    axpby<<<nb(n), 256, 0, s>>>(-omega, t, 1.0, r, n);  // r = s - omega t
    rho = rho1;
    res = std::sqrt(dot(r, r)) / bnorm;
    if (omega == 0.0) break;
  }
  *iters = it;
  *relres = res;
  return (err == hipSuccess && hipGetLastError() == hipSuccess) ? 0 : -1;
}

}  // namespace pgf
