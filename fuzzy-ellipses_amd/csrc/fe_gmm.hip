// Gaussian-mixture log-likelihood of a point cloud, forward and backward.
//
// Reference: baseline/fm_render.py:231-259 (log_likelihood).  With the upper-triangular
// precision factor U_i (only triu read), A_i = U_i^T, mixture weights pi = softmax(w_log) and
// p = mu_i - x_j:
//   res_ij = -0.5 (|A_i p|^2 + 3 ln 2pi) + ln pi_i + sum_k ln|U_i,kk|
//   lse_j  = logsumexp_i res_ij,          nll = -mean_j lse_j
// The backward of nll (scaled by an upstream scalar g) uses the responsibilities
// r_ij = exp(res_ij - lse_j) and three per-Gaussian moments over the points,
//   R_i = sum_j r_ij,  S_i = sum_j r_ij p,  Q_i = sum_j r_ij p p^T,
// from which, with M points,
//   d nll / d mu_i     = U U^T S_i / M
//   d nll / d U_ab     = ((Q_i U)_ab - [a == b] R_i / U_aa) / M      (a <= b)
//   d nll / d w_log,i  = pi_i - R_i / M
//
// Layout.  A preparation pass turns each Gaussian into a 16-float record
//   [A' (6: a00 a10 a11 a20 a21 a22), mu (3), c', pad]
// with A' = A sqrt(0.5 log2 e) and c' = (ln pi + ln det U - 1.5 ln 2pi) log2 e, so that
// res_ij log2 e = c' - |A' p|^2: the log-sum-exp runs in base 2 with one exp2 per pair.
// The forward gives each thread one point and streams the records through LDS (broadcast
// reads); the backward gives each thread one Gaussian and streams the points (with their
// lse) through LDS.  Per-chunk moment partials are reduced in a fixed order, so both passes
// are deterministic.
#include <hip/hip_runtime.h>

#include "fe_launch.h"

namespace fe {

namespace {

constexpr int kThreads = 256;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;
constexpr float kHalfLn2Pi3 = 2.7568155996140185f;  // 1.5 ln(2 pi)

__device__ float block_sum(float x, float* sh) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__device__ float block_max(float x, float* sh) {
  for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
  __syncthreads();
  return fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// One block: log-normaliser of the weights, ln sum_i exp(w_i), in a fixed order.
__global__ __launch_bounds__(kThreads) void k_gmm_wnorm(int n, const float* __restrict__ wlog, float* out) {
  __shared__ float sh[4];
  __shared__ float sh2[4];
  float m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += kThreads) m = fmaxf(m, wlog[i]);
  m = block_max(m, sh);
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += kThreads) s += expf(wlog[i] - m);
  s = block_sum(s, sh2);
  if (threadIdx.x == 0) *out = m + logf(s);
}

__global__ __launch_bounds__(kThreads) void k_gmm_records(int n, const float* __restrict__ means,
                                                          const float* __restrict__ prec,
                                                          const float* __restrict__ wlog, const float* wnorm,
                                                          float* __restrict__ rec) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float* U = prec + 9 * (int64_t)i;
  const float s = sqrtf(0.5f * kLog2e);
  float* r = rec + 16 * (int64_t)i;
  // A = U^T (lower): A[r][c] = U[c][r]
  r[0] = U[0] * s;                   // a00 = u00
  r[1] = U[1] * s;                   // a10 = u01
  r[2] = U[4] * s;                   // a11 = u11
  r[3] = U[2] * s;                   // a20 = u02
  r[4] = U[5] * s;                   // a21 = u12
  r[5] = U[8] * s;                   // a22 = u22
  r[6] = means[3 * (int64_t)i + 0];
  r[7] = means[3 * (int64_t)i + 1];
  r[8] = means[3 * (int64_t)i + 2];
  const float logdet = logf(fabsf(U[0])) + logf(fabsf(U[4])) + logf(fabsf(U[8]));
  r[9] = ((wlog[i] - *wnorm) + logdet - kHalfLn2Pi3) * kLog2e;
  r[10] = r[11] = r[12] = r[13] = r[14] = r[15] = 0.0f;
}

// |A' p|^2 for the record at r (LDS) and p = mu - x.
__device__ __forceinline__ float quad(const float* r, float x, float y, float z, float& p0, float& p1, float& p2) {
  p0 = r[6] - x;
  p1 = r[7] - y;
  p2 = r[8] - z;
  const float z0 = r[0] * p0;
  const float z1 = fmaf(r[2], p1, r[1] * p0);
  const float z2 = fmaf(r[5], p2, fmaf(r[4], p1, r[3] * p0));
  return fmaf(z2, z2, fmaf(z1, z1, z0 * z0));
}

template <int GB>
__global__ __launch_bounds__(kThreads) void k_gmm_fwd(int n, int64_t m, const float* __restrict__ rec,
                                                      const float* __restrict__ pts, float* __restrict__ lse_out,
                                                      float* __restrict__ res_out, float* __restrict__ part) {
  __shared__ float sh_rec[GB * 16];
  __shared__ float sh[4];
  const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const bool live = j < m;
  const float x = live ? pts[3 * j + 0] : 0.0f;
  const float y = live ? pts[3 * j + 1] : 0.0f;
  const float z = live ? pts[3 * j + 2] : 0.0f;
  float mx = -INFINITY, sum = 0.0f;
  for (int base = 0; base < n; base += GB) {
    const int cnt = min(GB, n - base);
    __syncthreads();
    for (int k = threadIdx.x; k < cnt * 16; k += kThreads) sh_rec[k] = rec[16 * (int64_t)base + k];
    __syncthreads();
    for (int g0 = 0; g0 < cnt; g0 += 8) {
      float l[8];
      float bm = -INFINITY;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int g = g0 + u;
        if (g < cnt) {
          const float* r = sh_rec + 16 * g;
          float p0, p1, p2;
          l[u] = r[9] - quad(r, x, y, z, p0, p1, p2);
          if (res_out && live) res_out[(int64_t)(base + g) * m + j] = l[u] * kLn2;
        } else {
          l[u] = -INFINITY;
        }
        bm = fmaxf(bm, l[u]);
      }
      if (bm > mx) {
        sum *= exp2f(mx - bm);
        mx = bm;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) sum += exp2f(l[u] - mx);
    }
  }
  const float lse = (mx + log2f(sum)) * kLn2;
  if (live) lse_out[j] = lse;
  const float tot = block_sum(live ? lse : 0.0f, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void k_gmm_nll(const float* part, int nblocks, int64_t m, float* nll) {
  __shared__ float sh[4];
  float s = 0.0f;
  for (int i = threadIdx.x; i < nblocks; i += kThreads) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *nll = -s / (float)m;
}

// Backward moments: lane = Gaussian, points of one chunk streamed through LDS.
constexpr int kPtBatch = 256;

__global__ __launch_bounds__(kThreads) void k_gmm_moments(int n, int64_t m, int64_t chunk,
                                                          const float* __restrict__ rec,
                                                          const float* __restrict__ pts,
                                                          const float* __restrict__ lse,
                                                          float* __restrict__ part) {
  __shared__ float4 sh_pt[kPtBatch];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const bool live = i < n;
  float r[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) r[k] = live ? rec[16 * (int64_t)i + k] : 0.0f;
  float R = 0.f, S0 = 0.f, S1 = 0.f, S2 = 0.f, Q00 = 0.f, Q01 = 0.f, Q02 = 0.f, Q11 = 0.f, Q12 = 0.f, Q22 = 0.f;
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = j0 + chunk < m ? j0 + chunk : m;
  for (int64_t b = j0; b < j1; b += kPtBatch) {
    const int cnt = (int)((j1 - b) < kPtBatch ? (j1 - b) : kPtBatch);
    __syncthreads();
    if (threadIdx.x < cnt) {
      const int64_t j = b + threadIdx.x;
      sh_pt[threadIdx.x] = make_float4(pts[3 * j + 0], pts[3 * j + 1], pts[3 * j + 2], lse[j] * kLog2e);
    }
    __syncthreads();
    if (!live) continue;
    for (int t = 0; t < cnt; ++t) {
      const float4 q = sh_pt[t];
      float p0, p1, p2;
      const float l = r[9] - quad(r, q.x, q.y, q.z, p0, p1, p2);
      const float w = exp2f(l - q.w);
      R += w;
      const float w0 = w * p0, w1 = w * p1, w2 = w * p2;
      S0 += w0;
      S1 += w1;
      S2 += w2;
      Q00 = fmaf(w0, p0, Q00);
      Q01 = fmaf(w0, p1, Q01);
      Q02 = fmaf(w0, p2, Q02);
      Q11 = fmaf(w1, p1, Q11);
      Q12 = fmaf(w1, p2, Q12);
      Q22 = fmaf(w2, p2, Q22);
    }
  }
  if (live) {
    float* o = part + ((int64_t)blockIdx.y * n + i) * 10;
    o[0] = R; o[1] = S0; o[2] = S1; o[3] = S2;
    o[4] = Q00; o[5] = Q01; o[6] = Q02; o[7] = Q11; o[8] = Q12; o[9] = Q22;
  }
}

__global__ __launch_bounds__(kThreads) void k_gmm_grad(int n, int64_t m, int n_chunks, const float* __restrict__ part,
                                                       const float* __restrict__ prec, const float* __restrict__ wlog,
                                                       const float* wnorm, float g, float* __restrict__ d_means,
                                                       float* __restrict__ d_prec, float* __restrict__ d_wlog) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float s[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) s[k] = 0.0f;
  for (int c = 0; c < n_chunks; ++c) {
    const float* o = part + ((int64_t)c * n + i) * 10;
#pragma unroll
    for (int k = 0; k < 10; ++k) s[k] += o[k];
  }
  const float sc = g / (float)m;
  const float* U = prec + 9 * (int64_t)i;
  const float u00 = U[0], u01 = U[1], u02 = U[2], u11 = U[4], u12 = U[5], u22 = U[8];
  const float R = s[0], S0 = s[1], S1 = s[2], S2 = s[3];
  const float Q[3][3] = {{s[4], s[5], s[6]}, {s[5], s[7], s[8]}, {s[6], s[8], s[9]}};
  // y = U^T S, then U y
  const float y0 = u00 * S0;
  const float y1 = u01 * S0 + u11 * S1;
  const float y2 = u02 * S0 + u12 * S1 + u22 * S2;
  atomicAdd(&d_means[3 * (int64_t)i + 0], sc * (u00 * y0 + u01 * y1 + u02 * y2));
  atomicAdd(&d_means[3 * (int64_t)i + 1], sc * (u11 * y1 + u12 * y2));
  atomicAdd(&d_means[3 * (int64_t)i + 2], sc * (u22 * y2));
  // (Q U)_ab for a <= b, U upper: (Q U)_ab = sum_{c <= b} Q_ac U_cb
  const float Uc[3][3] = {{u00, u01, u02}, {0.f, u11, u12}, {0.f, 0.f, u22}};
  float* dU = d_prec + 9 * (int64_t)i;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      float qu = 0.0f;
      for (int c = 0; c <= b; ++c) qu += Q[a][c] * Uc[c][b];
      if (a == b) qu -= R / Uc[a][a];
      atomicAdd(&dU[3 * a + b], sc * qu);
    }
  }
  const float pi = expf(wlog[i] - *wnorm);
  atomicAdd(&d_wlog[i], g * (pi - R / (float)m));
}

}  // namespace

int64_t gmm_chunk(int64_t m) {
  // at most 256 point chunks, each at least 1024 points
  int64_t c = (m + 255) / 256;
  if (c < 1024) c = 1024;
  return c;
}

hipError_t launch_gmm_forward(hipStream_t s, int n, int64_t m, const float* means, const float* prec,
                              const float* wlog, const float* pts, float* rec, float* wnorm, float* part,
                              float* lse, float* res, float* nll) {
  k_gmm_wnorm<<<1, kThreads, 0, s>>>(n, wlog, wnorm);
  k_gmm_records<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(n, means, prec, wlog, wnorm, rec);
  const int64_t nb = (m + kThreads - 1) / kThreads;
  k_gmm_fwd<64><<<(unsigned)nb, kThreads, 0, s>>>(n, m, rec, pts, lse, res, part);
  k_gmm_nll<<<1, kThreads, 0, s>>>(part, (int)nb, m, nll);
  return hipGetLastError();
}

hipError_t launch_gmm_backward(hipStream_t s, int n, int64_t m, const float* prec, const float* wlog,
                               const float* pts, const float* rec, const float* wnorm, const float* lse,
                               float* part, float g, float* d_means, float* d_prec, float* d_wlog) {
  const int64_t chunk = gmm_chunk(m);
  const int n_chunks = (int)((m + chunk - 1) / chunk);
  const dim3 grid((n + kThreads - 1) / kThreads, n_chunks);
  k_gmm_moments<<<grid, kThreads, 0, s>>>(n, m, chunk, rec, pts, lse, part);
  k_gmm_grad<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(n, m, n_chunks, part, prec, wlog, wnorm, g,
                                                                d_means, d_prec, d_wlog);
  return hipGetLastError();
}

}  // namespace fe
