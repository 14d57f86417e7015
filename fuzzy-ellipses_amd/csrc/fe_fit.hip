// Kernels of the shape-from-silhouette caller of the rasterizer: the per-view objective
// (binary cross-entropy of the clipped alpha against a silhouette mask, plus a weighted L1
// of depth on the mask) together with its image cotangents, and the Adam update.
//
// Objective of one view v with P pixels and mask sum M (reference objective:
// shape_from_silhouette.ipynb "jax_objective", CLIP_ALPHA = 3e-8; the depth term is the
// workload's addition):
//   a_c  = clip(alpha, c, 1 - c)
//   L_v  = sum_p BCE(a_c, m) / P + w_depth * sum_p m |depth - target| / max(M, 1)
//   loss = sum_v L_v / V
// d loss / d alpha is zero where the clip is active (alpha outside (c, 1 - c)), as for the
// derivative of a clip; d |x| / dx is sign(x) with sign(0) = 0.
//
// All sums use a fixed order (per-block partials, then one block sums the partials in
// index order), so the loss is bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include "../../include/fuzzy_ellipses_amd.h"
#include "fe_launch.h"

namespace fe {

namespace {

constexpr int kThreads = 256;

// Sum over the 256 threads of a block; the result is valid on thread 0.
__device__ float block_sum(float x, float* sh) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wave] = x;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ __launch_bounds__(kThreads) void k_sil_count(const fe_sil_view* views, float* cnt_part) {
  __shared__ float sh[4];
  const fe_sil_view V = views[blockIdx.y];
  float s = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < V.n_pixels; i += (int64_t)gridDim.x * kThreads)
    s += V.mask[i];
  const float r = block_sum(s, sh);
  if (threadIdx.x == 0) cnt_part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
}

__global__ __launch_bounds__(kThreads) void k_sil_loss(const fe_sil_view* views, int n_views, const float* cnt_part,
                                                       float w_depth, float clip, float* loss_part) {
  __shared__ float sh[4];
  __shared__ float sh_inv_m;
  const int v = blockIdx.y;
  const fe_sil_view V = views[v];
  if (threadIdx.x == 0) {
    float c = 0.0f;
    for (int b = 0; b < (int)gridDim.x; ++b) c += cnt_part[(int64_t)v * gridDim.x + b];
    sh_inv_m = 1.0f / fmaxf(c, 1.0f);
  }
  __syncthreads();
  const float inv_m = sh_inv_m;
  const float inv_p = 1.0f / (float)V.n_pixels;
  const float inv_v = 1.0f / (float)n_views;
  const float hi = 1.0f - clip;
  float s = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < V.n_pixels; i += (int64_t)gridDim.x * kThreads) {
    const float a = V.alpha[i];
    const float m = V.mask[i];
    const float ac = fminf(fmaxf(a, clip), hi);
    const float bce = -(m * logf(ac) + (1.0f - m) * log1pf(-ac));
    const bool inside = (a > clip) && (a < hi);
    const float da = inside ? -(m / ac - (1.0f - m) / (1.0f - ac)) : 0.0f;
    const float diff = V.depth[i] - V.target_depth[i];
    const float sgn = (float)((diff > 0.0f) - (diff < 0.0f));
    V.d_alpha[i] = da * inv_p * inv_v;
    V.d_depth[i] = w_depth * sgn * m * inv_m * inv_v;
    s += bce * inv_p + w_depth * (m * fabsf(diff)) * inv_m;
  }
  const float r = block_sum(s, sh);
  if (threadIdx.x == 0) loss_part[(int64_t)v * gridDim.x + blockIdx.x] = r * inv_v;
}

__global__ __launch_bounds__(kThreads) void k_sil_final(const float* loss_part, int n, float* loss) {
  __shared__ float sh[4];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += kThreads) s += loss_part[i];
  const float r = block_sum(s, sh);
  if (threadIdx.x == 0) *loss = r;
}

__global__ __launch_bounds__(kThreads) void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                   float b2, float c1, float c2, float eps) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

}  // namespace

hipError_t launch_sil_loss(hipStream_t s, const fe_sil_view* views, int n_views, float* cnt_part,
                           float* loss_part, float w_depth, float clip, float* loss) {
  const dim3 grid(kSilBlocksPerView, n_views);
  k_sil_count<<<grid, kThreads, 0, s>>>(views, cnt_part);
  k_sil_loss<<<grid, kThreads, 0, s>>>(views, n_views, cnt_part, w_depth, clip, loss_part);
  k_sil_final<<<1, kThreads, 0, s>>>(loss_part, n_views * kSilBlocksPerView, loss);
  return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, int64_t n, float* p, const float* g, float* m, float* v, float lr,
                       float b1, float b2, float c1, float c2, float eps) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 4096) blocks = 4096;
  k_adam<<<(unsigned)blocks, kThreads, 0, s>>>(n, p, g, m, v, lr, b1, b2, c1, c2, eps);
  return hipGetLastError();
}

}  // namespace fe
