// Small MLP on the gfx950 matrix cores: BF16 operands, FP32 accumulation (MFMA
// 16x16x16 bf16), hand-written forward, backward and Adam kernels.  It serves as the
// learned warm start of the crystal-plasticity local Newton (no reference counterpart: the
// reference has no learned component).
//
// Layout: layer l maps K_l inputs to N_l outputs, both padded to multiples of 16 (Kp, Np);
// weights are row-major [Np][Kp] with zero padding, biases [Np].  Master parameters and Adam
// moments are FP32; every step writes BF16 copies of the weights for the matrix cores.
// Activations are row-major [B][Kp] BF16; pre-activations Z are kept in FP32 for the
// backward pass.  GELU is the exact (erf) form.
//
// MFMA 16x16x16 bf16 fragment layout (lane l of a 64-wide wave):
//   A (16 x 16, rows m, cols k): a[i] = A[l % 16][4 (l / 16) + i]
//   B (16 x 16, rows k, cols n): b[i] = B[4 (l / 16) + i][l % 16]
//   D (16 x 16):                 d[i] = D[4 (l / 16) + i][l % 16],   i = 0..3
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "pgfem3d_amd.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi : int {
  EPI_HIDDEN = 0,  // Z = acc + b (fp32, kept), H = bf16(gelu(Z))
  EPI_OUT = 1,     // Y = acc + b (fp32), columns < n_real
  EPI_DGELU = 2,   // dZ = bf16(acc * gelu'(Zprev))
  EPI_ATOMIC = 3,  // C += acc (fp32 atomics; split-K weight gradient)
};

struct GemmArgs {
  const uint16_t *A;  // element (m, k) at A[m * a_m + k * a_k]
  int64_t a_m, a_k;
  const uint16_t *B;  // element (k, n) at B[k * b_k + n * b_n]
  int64_t b_k, b_n;
  int64_t M;          // rows of C (guarded)
  int N;              // columns of C, multiple of 16, <= 128
  int64_t K;          // reduction length, multiple of 16 (split into chunks of k_chunk)
  int64_t k_chunk;
  const float *bias;  // [N]
  float *Cf;          // fp32 output (Z, Y, or atomic target), leading dimension ldc
  uint16_t *Cb;       // bf16 output (H or dZ), leading dimension ldc
  const float *Zprev; // EPI_DGELU: pre-activation of the input layer, leading dimension ldc
  int64_t ldc;
  int n_real;         // EPI_OUT: valid columns
};

// One wave computes a 16-row strip of C over all N columns for one K chunk.
template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm(GemmArgs g) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t m0 = (int64_t(blockIdx.x) * 4 + wave) * 16;
  if (m0 >= g.M) return;
  const int64_t k_lo = int64_t(blockIdx.y) * g.k_chunk;
  const int64_t k_hi = (k_lo + g.k_chunk < g.K) ? k_lo + g.k_chunk : g.K;
  const int nt = g.N / 16;
  f32x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane % 16, q4 = (lane / 16) * 4;
  const int64_t am = m0 + r16;
  const bool arow = am < g.M;
  for (int64_t k0 = k_lo; k0 < k_hi; k0 += 16) {
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = arow ? short(g.A[am * g.a_m + (k0 + q4 + i) * g.a_k]) : short(0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (j < nt) {
        bf16x4 b;
        const int64_t n = int64_t(j) * 16 + r16;
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = short(g.B[(k0 + q4 + i) * g.b_k + n * g.b_n]);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (j >= nt) continue;
    const int n = j * 16 + r16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t m = m0 + q4 + i;
      if (m >= g.M) continue;
      const float v = acc[j][i];
      const int64_t o = m * g.ldc + n;
      if constexpr (EPI == EPI_HIDDEN) {
        const float z = v + g.bias[n];
        g.Cf[o] = z;
        g.Cb[o] = f2bf(gelu(z));
      } else if constexpr (EPI == EPI_OUT) {
        if (n < g.n_real) g.Cf[m * g.ldc + n] = v + g.bias[n];
      } else if constexpr (EPI == EPI_DGELU) {
        g.Cb[o] = f2bf(v * gelu_d(g.Zprev[o]));
      } else {
        atomicAdd(g.Cf + o, v);
      }
    }
  }
}

// features fp32 [B][n_in] -> bf16 [B][Kp] (zero padded)
__global__ void pack_input(const float *X, int64_t B, int n_in, int Kp, uint16_t *H) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= B * Kp) return;
  const int64_t m = t / Kp;
  const int k = int(t % Kp);
  H[t] = f2bf(k < n_in ? X[m * n_in + k] : 0.0f);
}

// dY = 2 (Y - T) / (B n_out) as bf16 [B][Np]; loss += sum (Y - T)^2 / (B n_out)
__global__ void loss_grad(const float *Y, int64_t ldy, const float *T, int64_t B, int n_out, int Np, uint16_t *dY,
                          float *loss) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (t < B * Np) {
    const int64_t m = t / Np;
    const int n = int(t % Np);
    float d = 0.f;
    if (n < n_out) {
      const float e = Y[m * ldy + n] - T[m * n_out + n];
      const float s = 1.0f / (float(B) * float(n_out));
      d = 2.0f * e * s;
      l = e * e * s;
    }
    dY[t] = f2bf(d);
  }
  // wave reduction of the loss, one atomic per wave
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  if ((threadIdx.x % 64) == 0 && l != 0.f) atomicAdd(loss, l);
}

// db[n] += sum_m dZ[m][n]
__global__ void bias_grad(const uint16_t *dZ, int64_t B, int Np, float *db) {
  const int n = threadIdx.x;
  if (n >= Np) return;
  const int64_t lo = int64_t(blockIdx.x) * 1024;
  const int64_t hi = (lo + 1024 < B) ? lo + 1024 : B;
  float s = 0.f;
  for (int64_t m = lo; m < hi; ++m) s += bf2f(dZ[m * Np + n]);
  atomicAdd(db + n, s);
}

__global__ void adam(float *p, float *g, float *mom, float *vel, int64_t n, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mom[i] + (1.f - b1) * gi;
  const float v = b2 * vel[i] + (1.f - b2) * gi * gi;
  mom[i] = m;
  vel[i] = v;
  p[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  g[i] = 0.f;
}

__global__ void to_bf16(const float *p, uint16_t *b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(p[i]);
}

__global__ void zero_f32(float *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Inference of the whole network in one launch (the learned start's forward pass).
//
// Each wave carries two 16-row strips of the batch through every layer with the
// activations in registers: the layer is computed transposed, D^T = W H^T, so the MFMA
// output fragment of a 16-feature tile (d[i] = D[4 (l / 16) + i][l % 16]: feature rows,
// batch columns) is exactly the B fragment of the next layer's 16-wide k block
// (b[i] = B[4 (l / 16) + i][l % 16]).  Only the weights go through LDS: the workgroup stages
// one layer's BF16 weights and FP32 biases at a time (<= 128 x 128 x 2 B + 512 B) and every
// wave reads its A fragments from there, so a layer's weights are read from L2 once per
// 128 rows instead of once per 16.  No pre-activations are written (inference keeps none).
constexpr int kInferWaves = 4, kInferStrips = 2, kMaxW = 128;

struct InferArgs {
  const float *X;      // [B][n_in] fp32 (normalised features)
  int64_t B;
  int n_in;
  const uint16_t *W;   // bf16 parameters (the d_wb copy)
  const float *P;      // fp32 parameters (biases)
  int L;
  int Kp[8], Np[8], n_out;
  int64_t woff[8], boff[8];
  float *Y;            // [B][ldy]
  int ldy;
};

__global__ __launch_bounds__(256) void mlp_infer(InferArgs g) {
  __shared__ uint16_t sW[kMaxW * kMaxW];
  __shared__ float sb[kMaxW];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r16 = lane % 16, q4 = (lane / 16) * 4;
  const int64_t rows_per_chunk = int64_t(kInferWaves) * kInferStrips * 16;
  for (int64_t c0 = int64_t(blockIdx.x) * rows_per_chunk; c0 < g.B; c0 += int64_t(gridDim.x) * rows_per_chunk) {
    // layer-0 B fragments straight from the features: b[i] = X[m][16 kb + q4 + i], m = row r16
    bf16x4 act[kInferStrips][kMaxW / 16];
#pragma unroll
    for (int s = 0; s < kInferStrips; ++s) {
      const int64_t m = c0 + (int64_t(wave) * kInferStrips + s) * 16 + r16;
#pragma unroll
      for (int kb = 0; kb < kMaxW / 16; ++kb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = kb * 16 + q4 + i;
          const float v = (m < g.B && k < g.n_in) ? g.X[m * g.n_in + k] : 0.0f;
          act[s][kb][i] = short(f2bf(v));
        }
      }
    }
    for (int l = 0; l < g.L; ++l) {
      const int Kp = g.Kp[l], Np = g.Np[l];
      __syncthreads();  // the previous layer's weights are no longer read
      {
        const uint64_t *src = reinterpret_cast<const uint64_t *>(g.W + g.woff[l]);
        uint64_t *dst = reinterpret_cast<uint64_t *>(sW);
        for (int t = threadIdx.x; t < Np * Kp / 4; t += blockDim.x) dst[t] = src[t];
        for (int t = threadIdx.x; t < Np; t += blockDim.x) sb[t] = g.P[g.boff[l] + t];
      }
      __syncthreads();
      const bool last = l + 1 == g.L;
#pragma unroll
      for (int s = 0; s < kInferStrips; ++s) {
        bf16x4 nxt[kMaxW / 16];
        const int64_t m = c0 + (int64_t(wave) * kInferStrips + s) * 16 + r16;
#pragma unroll
        for (int j = 0; j < kMaxW / 16; ++j) {
          if (j * 16 >= Np) break;
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kb = 0; kb < kMaxW / 16; ++kb) {
            if (kb * 16 >= Kp) break;
            // a[i] = W[16 j + r16][16 kb + q4 + i]: four consecutive bf16 of one weight row
            const bf16x4 a = *reinterpret_cast<const bf16x4 *>(sW + (j * 16 + r16) * Kp + kb * 16 + q4);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[s][kb], acc, 0, 0, 0);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int n = j * 16 + q4 + i;
            const float z = acc[i] + sb[n];
            if (last) {
              if (m < g.B && n < g.n_out) g.Y[m * g.ldy + n] = z;
            } else {
              nxt[j][i] = short(f2bf(gelu(z)));
            }
          }
        }
        if (!last) {
#pragma unroll
          for (int j = 0; j < kMaxW / 16; ++j) act[s][j] = nxt[j];
        }
      }
    }
  }
}

inline unsigned blocks_of(int64_t n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

struct pgf_mlp_s {
  int L = 0;
  std::vector<int> K, N, Kp, Np;     // per layer: real and padded widths
  std::vector<int64_t> woff, boff;   // offsets of W_l and b_l in the parameter vector
  int64_t n_param = 0;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;  // [n_param]
  uint16_t *d_wb = nullptr;          // bf16 copy of the parameter vector
  int64_t step = 0;
  hipStream_t stream = nullptr;
  // activation workspace for a batch capacity
  int64_t cap = 0;
  std::vector<uint16_t *> H;         // H[0] = packed input, H[l] = output of hidden layer l
  std::vector<float *> Z;            // pre-activations of hidden layers
  float *Y = nullptr;                // output [cap][Np_last]
  uint16_t *dA = nullptr, *dB = nullptr;  // bf16 gradient ping-pong [cap][128]
  float *d_loss = nullptr;
};

namespace {

int ensure_cap(pgf_mlp m, int64_t B) {
  if (B <= m->cap) return PGF_OK;
  for (auto p : m->H) hipFree(p);
  for (auto p : m->Z) hipFree(p);
  hipFree(m->Y);
  hipFree(m->dA);
  hipFree(m->dB);
  m->H.assign(m->L, nullptr);
  m->Z.assign(m->L, nullptr);
  const int64_t cap = ((B + 63) / 64) * 64;
  for (int l = 0; l < m->L; ++l) {
    if (hipMalloc(&m->H[l], size_t(cap) * m->Kp[l] * 2) != hipSuccess) return PGF_ENOMEM;
    if (l > 0 && hipMalloc(&m->Z[l], size_t(cap) * m->Kp[l] * 4) != hipSuccess) return PGF_ENOMEM;
  }
  if (hipMalloc(&m->Y, size_t(cap) * m->Np[m->L - 1] * 4) != hipSuccess) return PGF_ENOMEM;
  if (hipMalloc(&m->dA, size_t(cap) * 128 * 2) != hipSuccess) return PGF_ENOMEM;
  if (hipMalloc(&m->dB, size_t(cap) * 128 * 2) != hipSuccess) return PGF_ENOMEM;
  m->cap = cap;
  return PGF_OK;
}

// Whole-network inference (mlp_infer) of B rows of fp32 features X [B][K_0] into Y [B][ldy].
int infer(pgf_mlp m, const float *X, int64_t B, float *Y, int ldy, hipStream_t s) {
  if (m->L > 8) return PGF_ENOTSUP;
  InferArgs g{};
  g.X = X;
  g.B = B;
  g.n_in = m->K[0];
  g.W = m->d_wb;
  g.P = m->d_p;
  g.L = m->L;
  for (int l = 0; l < m->L; ++l) {
    g.Kp[l] = m->Kp[l];
    g.Np[l] = m->Np[l];
    g.woff[l] = m->woff[l];
    g.boff[l] = m->boff[l];
  }
  g.n_out = m->N[m->L - 1];
  g.Y = Y;
  g.ldy = ldy;
  const int64_t chunks = (B + kInferWaves * kInferStrips * 16 - 1) / (kInferWaves * kInferStrips * 16);
  const unsigned grid = unsigned(chunks < 2048 ? chunks : 2048);
  mlp_infer<<<grid, 64 * kInferWaves, 0, s>>>(g);
  return hipGetLastError() == hipSuccess ? PGF_OK : PGF_EHIP;
}

void refresh_bf16(pgf_mlp m) { to_bf16<<<blocks_of(m->n_param, 256), 256, 0, m->stream>>>(m->d_p, m->d_wb, m->n_param); }

// Forward over a batch already packed into H[0]; Y gets the outputs.
void forward_packed(pgf_mlp m, int64_t B) {
  for (int l = 0; l < m->L; ++l) {
    GemmArgs g{};
    g.A = m->H[l];
    g.a_m = m->Kp[l];
    g.a_k = 1;
    g.B = m->d_wb + m->woff[l];  // (k, n) = W[n][k]
    g.b_k = 1;
    g.b_n = m->Kp[l];
    g.M = B;
    g.N = m->Np[l];
    g.K = m->Kp[l];
    g.k_chunk = g.K;
    g.bias = m->d_p + m->boff[l];
    const dim3 grid(blocks_of(B, 64), 1);
    if (l + 1 < m->L) {
      g.Cf = m->Z[l + 1];
      g.Cb = m->H[l + 1];
      g.ldc = m->Np[l];
      mlp_gemm<EPI_HIDDEN><<<grid, 256, 0, m->stream>>>(g);
    } else {
      g.Cf = m->Y;
      g.ldc = m->Np[l];
      g.n_real = m->N[l];
      mlp_gemm<EPI_OUT><<<grid, 256, 0, m->stream>>>(g);
    }
  }
}

}  // namespace

extern "C" {

int pgf_mlp_create(int32_t n_layers, const int32_t *widths, uint64_t seed, pgf_mlp *out) {
  if (!widths || !out || n_layers < 1) return PGF_EINVAL;
  *out = nullptr;
  for (int l = 0; l <= n_layers; ++l)
    if (widths[l] < 1 || widths[l] > 128) return PGF_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PGF_ENODEV;
  auto *m = new pgf_mlp_s();
  m->L = n_layers;
  int64_t off = 0;
  for (int l = 0; l < n_layers; ++l) {
    m->K.push_back(widths[l]);
    m->N.push_back(widths[l + 1]);
    m->Kp.push_back(((widths[l] + 15) / 16) * 16);
    m->Np.push_back(((widths[l + 1] + 15) / 16) * 16);
    m->woff.push_back(off);
    off += int64_t(m->Np[l]) * m->Kp[l];
    m->boff.push_back(off);
    off += m->Np[l];
  }
  m->n_param = off;
  // He-style init for GELU layers, scaled uniform; deterministic from the seed
  std::vector<float> p(off, 0.f);
  std::mt19937_64 gen(seed);
  for (int l = 0; l < n_layers; ++l) {
    const float a = std::sqrt(6.0f / float(m->K[l])) * (l + 1 < n_layers ? 1.0f : 0.5f);
    for (int n = 0; n < m->N[l]; ++n)
      for (int k = 0; k < m->K[l]; ++k)
        p[m->woff[l] + int64_t(n) * m->Kp[l] + k] = a * (2.0f * float(double(gen() >> 11) * 0x1.0p-53) - 1.0f);
  }
  auto fail = [&](int rc) {
    pgf_mlp_destroy(m);
    return rc;
  };
  if (hipMalloc(&m->d_p, off * 4) != hipSuccess || hipMalloc(&m->d_g, off * 4) != hipSuccess ||
      hipMalloc(&m->d_m, off * 4) != hipSuccess || hipMalloc(&m->d_v, off * 4) != hipSuccess ||
      hipMalloc(&m->d_wb, off * 2) != hipSuccess || hipMalloc(&m->d_loss, 4) != hipSuccess)
    return fail(PGF_ENOMEM);
  if (hipMemcpy(m->d_p, p.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess) return fail(PGF_EHIP);
  zero_f32<<<blocks_of(off, 256), 256, 0, m->stream>>>(m->d_g, off);
  zero_f32<<<blocks_of(off, 256), 256, 0, m->stream>>>(m->d_m, off);
  zero_f32<<<blocks_of(off, 256), 256, 0, m->stream>>>(m->d_v, off);
  refresh_bf16(m);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return fail(PGF_EHIP);
  *out = m;
  return PGF_OK;
}

int pgf_mlp_destroy(pgf_mlp m) {
  if (!m) return PGF_OK;
  hipStreamSynchronize(m->stream);
  for (auto p : m->H) hipFree(p);
  for (auto p : m->Z) hipFree(p);
  hipFree(m->Y);
  hipFree(m->dA);
  hipFree(m->dB);
  hipFree(m->d_p);
  hipFree(m->d_g);
  hipFree(m->d_m);
  hipFree(m->d_v);
  hipFree(m->d_wb);
  hipFree(m->d_loss);
  delete m;
  return PGF_OK;
}

int pgf_mlp_param_count(pgf_mlp m, int64_t *n) {
  if (!m || !n) return PGF_EINVAL;
  *n = m->n_param;
  return PGF_OK;
}

int pgf_mlp_get_params(pgf_mlp m, float *host) {
  if (!m || !host) return PGF_EINVAL;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return PGF_EHIP;
  return hipMemcpy(host, m->d_p, m->n_param * 4, hipMemcpyDeviceToHost) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_mlp_set_params(pgf_mlp m, const float *host) {
  if (!m || !host) return PGF_EINVAL;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return PGF_EHIP;
  if (hipMemcpy(m->d_p, host, m->n_param * 4, hipMemcpyHostToDevice) != hipSuccess) return PGF_EHIP;
  refresh_bf16(m);
  return hipStreamSynchronize(m->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_mlp_get_grads(pgf_mlp m, float *host) {
  if (!m || !host) return PGF_EINVAL;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return PGF_EHIP;
  return hipMemcpy(host, m->d_g, m->n_param * 4, hipMemcpyDeviceToHost) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_mlp_forward(pgf_mlp m, const float *X, int64_t B, float *Y) {
  if (!m || !X || !Y || B < 1) return PGF_EINVAL;
  int rc = ensure_cap(m, B);
  if (rc) return rc;
  pack_input<<<blocks_of(B * m->Kp[0], 256), 256, 0, m->stream>>>(X, B, m->K[0], m->Kp[0], m->H[0]);
  forward_packed(m, B);
  const int no = m->N[m->L - 1], np = m->Np[m->L - 1];
  if (hipMemcpy2DAsync(Y, size_t(no) * 4, m->Y, size_t(np) * 4, size_t(no) * 4, size_t(B), hipMemcpyDeviceToDevice,
                       m->stream) != hipSuccess)
    return PGF_EHIP;
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  return hipStreamSynchronize(m->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_mlp_infer(pgf_mlp m, const float *X, int64_t B, float *Y) {
  if (!m || !X || !Y || B < 1) return PGF_EINVAL;
  int rc = infer(m, X, B, Y, m->N[m->L - 1], m->stream);
  if (rc) return rc;
  return hipStreamSynchronize(m->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

// Forward + backward of the mean squared error over a batch; gradients are ADDED to the
// gradient buffer (zeroed by each Adam step).  *loss = the batch loss.
int pgf_mlp_backward(pgf_mlp m, const float *X, const float *T, int64_t B, float *loss) {
  if (!m || !X || !T || B < 1) return PGF_EINVAL;
  int rc = ensure_cap(m, B);
  if (rc) return rc;
  pack_input<<<blocks_of(B * m->Kp[0], 256), 256, 0, m->stream>>>(X, B, m->K[0], m->Kp[0], m->H[0]);
  forward_packed(m, B);
  const int L = m->L;
  zero_f32<<<1, 64, 0, m->stream>>>(m->d_loss, 1);
  uint16_t *dZ = m->dA, *dN = m->dB;
  loss_grad<<<blocks_of(B * m->Np[L - 1], 256), 256, 0, m->stream>>>(m->Y, m->Np[L - 1], T, B, m->N[L - 1],
                                                                      m->Np[L - 1], dZ, m->d_loss);
  for (int l = L - 1; l >= 0; --l) {
    // weight gradient dW[n][k] += sum_m dZ[m][n] H[m][k]
    GemmArgs g{};
    g.A = dZ;  // (n, m) = dZ[m][n]
    g.a_m = 1;
    g.a_k = m->Np[l];
    g.B = m->H[l];  // (m, k) = H[m][k]
    g.b_k = m->Kp[l];
    g.b_n = 1;
    g.M = m->Np[l];
    g.N = m->Kp[l];
    g.K = ((B + 15) / 16) * 16;
    g.k_chunk = 1024;
    g.Cf = m->d_g + m->woff[l];
    g.ldc = m->Kp[l];
    // the batch is padded to 16 rows: rows >= B of dZ and H are never written, so they are
    // excluded by clamping the reduction to B rounded down plus a guarded tail chunk
    if (B % 16 != 0) return PGF_EINVAL;
    g.K = B;
    mlp_gemm<EPI_ATOMIC><<<dim3(blocks_of(g.M, 64), blocks_of(g.K, 1024)), 256, 0, m->stream>>>(g);
    bias_grad<<<blocks_of(B, 1024), 128, 0, m->stream>>>(dZ, B, m->Np[l], m->d_g + m->boff[l]);
    if (l == 0) break;
    // data gradient dH_{l} = dZ W_l, then dZ_{l} = dH * gelu'(Z_l)
    GemmArgs d{};
    d.A = dZ;
    d.a_m = m->Np[l];
    d.a_k = 1;
    d.B = m->d_wb + m->woff[l];  // (k = n_out, n = k_in) = W[n_out][k_in]
    d.b_k = m->Kp[l];
    d.b_n = 1;
    d.M = B;
    d.N = m->Kp[l];
    d.K = m->Np[l];
    d.k_chunk = d.K;
    d.Cb = dN;
    d.Zprev = m->Z[l];
    d.ldc = m->Kp[l];
    mlp_gemm<EPI_DGELU><<<dim3(blocks_of(B, 64), 1), 256, 0, m->stream>>>(d);
    std::swap(dZ, dN);
  }
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  if (loss && hipMemcpyAsync(loss, m->d_loss, 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) return PGF_EHIP;
  return hipStreamSynchronize(m->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

int pgf_mlp_adam_step(pgf_mlp m, float lr, float beta1, float beta2, float eps) {
  if (!m) return PGF_EINVAL;
  m->step += 1;
  const float c1 = 1.0f - std::pow(beta1, float(m->step));
  const float c2 = 1.0f - std::pow(beta2, float(m->step));
  adam<<<blocks_of(m->n_param, 256), 256, 0, m->stream>>>(m->d_p, m->d_g, m->d_m, m->d_v, m->n_param, lr, beta1,
                                                          beta2, eps, c1, c2);
  refresh_bf16(m);
  if (hipGetLastError() != hipSuccess) return PGF_EHIP;
  return hipStreamSynchronize(m->stream) == hipSuccess ? PGF_OK : PGF_EHIP;
}

}  // extern "C"

// Device-side forward for other translation units (the crystal-plasticity warm start):
// X fp32 [B][n_in] on stream s, outputs in the MLP's Y buffer [B][Np_last], by the fused
// inference kernel.
namespace pgf {
int mlp_forward_device(pgf_mlp m, const float *X, int64_t B, hipStream_t s, const float **Y, int *ldy) {
  int rc = ensure_cap(m, B);
  if (rc) return rc;
  if ((rc = infer(m, X, B, m->Y, m->Np[m->L - 1], s))) return rc;
  *Y = m->Y;
  *ldy = m->Np[m->L - 1];
  return PGF_OK;
}
}  // namespace pgf
