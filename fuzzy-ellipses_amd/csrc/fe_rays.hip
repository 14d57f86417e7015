// Per-ray optical flow between consecutive poses of a multi-pose ray batch.
//
// Reference: baseline/fm_render.py:170-224 (render_func_idx_quattrans_flow).  Ray k carries
// (x, y, z, pose) in pixel_posei; its camera-frame direction is (x invF, y invF, z) and its
// rendered depth d is the ray parameter of the surface point, so the point in the frame of
// pose i is (x invF, y invF, z) d.  It is moved to world coordinates with pose i
// (world = p R_i + t_i, row vectors) and projected into the next pose i+1 and the previous
// pose i-1 (indices clamped to [0, P-1]):  c = (world - t_j) R_j^T,  (u, v) = c.xy / (c.z invF).
// The flow is (x, y) - (u, v) for each of the two neighbours.
#include <hip/hip_runtime.h>

#include "fe_launch.h"

namespace fe {

namespace {

__device__ __forceinline__ void project(const float* w, const float* R, const float* t, float invF, float x, float y,
                                        float* out) {
  const float d0 = w[0] - t[0], d1 = w[1] - t[1], d2 = w[2] - t[2];
  // c = d R^T: c_k = sum_m d_m R[k][m]
  const float c0 = fmaf(d0, R[0], fmaf(d1, R[1], d2 * R[2]));
  const float c1 = fmaf(d0, R[3], fmaf(d1, R[4], d2 * R[5]));
  const float c2 = fmaf(d0, R[6], fmaf(d1, R[7], d2 * R[8]));
  const float den = c2 * invF;
  out[0] = x - c0 / den;
  out[1] = y - c1 / den;
}

__global__ __launch_bounds__(256) void k_ray_flow(int64_t n, const float* __restrict__ pp, float invF, int n_poses,
                                                  const float* __restrict__ rot, const float* __restrict__ trans,
                                                  const float* __restrict__ depth, float* __restrict__ fnext,
                                                  float* __restrict__ fprev) {
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
    const float x = pp[4 * k], y = pp[4 * k + 1], z = pp[4 * k + 2];
    int i = (int)pp[4 * k + 3];                      // truncation, as astype(int)
    i = min(max(i, 0), n_poses - 1);
    const int ip = min(i + 1, n_poses - 1), im = max(i - 1, 0);
    const float d = depth[k];
    const float p0 = x * invF * d, p1 = y * invF * d, p2 = z * d;
    const float* R = rot + 9 * i;
    const float* t = trans + 3 * i;
    float w[3];
    // world = p R + t: w_k = sum_m p_m R[m][k] + t_k
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c] = fmaf(p0, R[c], fmaf(p1, R[3 + c], fmaf(p2, R[6 + c], t[c])));
    float f[2];
    project(w, rot + 9 * ip, trans + 3 * ip, invF, x, y, f);
    fnext[2 * k] = f[0];
    fnext[2 * k + 1] = f[1];
    project(w, rot + 9 * im, trans + 3 * im, invF, x, y, f);
    fprev[2 * k] = f[0];
    fprev[2 * k + 1] = f[1];
  }
}

}  // namespace

hipError_t launch_ray_flow(hipStream_t s, int64_t n, const float* pp, float invF, int n_poses, const float* rot,
                           const float* trans, const float* depth, float* fnext, float* fprev) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  k_ray_flow<<<(unsigned)blocks, 256, 0, s>>>(n, pp, invF, n_poses, rot, trans, depth, fnext, fprev);
  return hipGetLastError();
}

}  // namespace fe
