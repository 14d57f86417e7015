// Backward kernels of the fuzzy-metaball rasterizer (gfx950).
//
// No reference counterpart: the reference obtained gradients from jax.grad of
// baseline/fm_render.py:11-109.  This pass recomputes every listed pair from the record and
// the saved per-pixel aggregates (M, S, Z, E, D) instead of storing per-pair values.
//
// Per pixel (upstream gD = dL/dD, gA = dL/dalpha), D = Z/S, alpha = 1 - exp(-E):
//   cE = gA exp(-E),  cD = gD / S,  s_j = [t_j>0] 2^(l2_j - M)
//   g_d2 = cE e_j + b3 cD s_j (t_j - D)
//   g_t  = cD s_j (1 - b2 (t_j - D))
// and through t* = (q.u)/(u.u), d2 = w - |t* u - q|^2 / 2 (envelope: dd0/dt = 0), v = t u - q:
//   g_q = g_d2 v + (g_t/vsv) u
//   g_u = -g_d2 t v + (g_t/vsv) (q - 2 t u)
//   g_w = g_d2
// Per (Gaussian, view) the kernel accumulates 13 sums: G_q (3), G_a = sum g_u x' (3),
// G_b = sum g_u y' (3), G_g = sum g_u (3), G_w; k_bwd_finalize maps them to d(mu), d(U), d(w).
#include "fe_common.h"

#include <string.h>

#include <algorithm>

namespace fe {

__global__ void k_zero(float* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Reduce 16 per-lane values over a wave64 with 17 cross-lane exchanges.  Afterwards lane
// 4*k holds the wave total of value perm(k), where perm(k) is the value index selected by
// lane bits 5..2; the function returns that index through *idx.
__device__ __forceinline__ float wave_reduce16(float (&v)[16], int lane, int* idx) {
  const bool b5 = (lane >> 5) & 1, b4 = (lane >> 4) & 1, b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1;
  float a8[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float mine = b5 ? v[i + 8] : v[i];
    const float other = b5 ? v[i] : v[i + 8];
    a8[i] = mine + __shfl_xor(other, 32, 64);
  }
  float a4[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float mine = b4 ? a8[i + 4] : a8[i];
    const float other = b4 ? a8[i] : a8[i + 4];
    a4[i] = mine + __shfl_xor(other, 16, 64);
  }
  float a2[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const float mine = b3 ? a4[i + 2] : a4[i];
    const float other = b3 ? a4[i] : a4[i + 2];
    a2[i] = mine + __shfl_xor(other, 8, 64);
  }
  float a1;
  {
    const float mine = b2 ? a2[1] : a2[0];
    const float other = b2 ? a2[0] : a2[1];
    a1 = mine + __shfl_xor(other, 4, 64);
  }
  a1 += __shfl_xor(a1, 2, 64);
  a1 += __shfl_xor(a1, 1, 64);
  *idx = (b5 ? 8 : 0) + (b4 ? 4 : 0) + (b3 ? 2 : 0) + (b2 ? 1 : 0);
  return a1;
}

struct BwdArgs {
  const Job* jobs;
  const BwdJob* bjobs;
  int n_jobs;
  const float* rec;
  const unsigned long long* off1;
  const unsigned long long* off2;
  const uint32_t* list;
  const float* st;
  int64_t total_pix;
  float* gacc;   // 16 floats per (Gaussian, view)
  int dense;
};

template <int TS, int GB>
__global__ __launch_bounds__(TS* TS) void k_bwd_blend(BwdArgs A) {
  constexpr int NT = TS * TS;
  constexpr int NW = NT / 64;
  __shared__ float4 lds[GB * 4];
  __shared__ float acc[GB * 16];
  __shared__ int gid_sh[GB];

  const int64_t gt = blockIdx.x;
  const int ji = find_job(A.n_jobs, gt, [&](int k) { return A.jobs[k].tile_base; });
  const Job& jb = A.jobs[ji];
  const BwdJob& bj = A.bjobs[ji];
  const int tile = (int)(gt - jb.tile_base);
  const int tx = tile % jb.tiles_x, ty = tile / jb.tiles_x;
  const int px = tx * TS + (threadIdx.x % TS);
  const int py = ty * TS + (threadIdx.x / TS);
  const bool inimg = (px < jb.W) && (py < jb.H);
  const int64_t lp = (int64_t)(inimg ? py : 0) * jb.W + (inimg ? px : 0);
  const int64_t pix = jb.pix_base + lp;
  const float xf = (float)px - jb.cx;
  const float yf = (float)py - jb.cy;
  const float b3 = jb.beta3, b2 = jb.beta2;
  const float nb2l = -jb.beta2 * kLog2e;
  const float kd = -0.5f * kLog2e;
  const int lane = threadIdx.x & 63;

  int64_t s1, n1, s2, n2;
  if (A.dense) { s1 = 0; n1 = jb.n; s2 = 0; n2 = 0; }
  else {
    s1 = (int64_t)A.off1[gt]; n1 = (int64_t)A.off1[gt + 1] - s1;
    s2 = (int64_t)A.off2[gt]; n2 = (int64_t)A.off2[gt + 1] - s2;
  }
  const float ncnt = (float)(n1 + n2);
  const float M = A.st[pix];
  const float S = A.st[A.total_pix + pix];
  const float E = A.st[3 * A.total_pix + pix];
  const float D = A.st[5 * A.total_pix + pix];
  const float gD = (inimg && bj.gD) ? bj.gD[lp] : 0.0f;
  const float gA = (inimg && bj.gA) ? bj.gA[lp] : 0.0f;
  const float cE = gA * expf(-E);
  const float den = S + 1e-20f * ncnt;
  const float cD = (den > 0.0f) ? gD / den : 0.0f;

  const int64_t total = n1 + n2;
  const float4* rec4 = reinterpret_cast<const float4*>(A.rec) + jb.rec_base * 4;
  for (int i = threadIdx.x; i < GB * 16; i += NT) acc[i] = 0.0f;
  for (int64_t base = 0; base < total; base += GB) {
    const int nb = (int)min<int64_t>(GB, total - base);
    for (int i = threadIdx.x; i < nb * 4; i += NT) {
      const int k = i >> 2, part = i & 3;
      const int64_t kk = base + k;
      int64_t gid;
      if (A.dense) gid = kk;
      else gid = (kk < n1) ? (int64_t)A.list[s1 + kk] : (int64_t)A.list[s2 + (kk - n1)];
      lds[i] = rec4[gid * 4 + part];
      if (part == 0) gid_sh[k] = (int)gid;
    }
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
      const float4 r0 = lds[k * 4 + 0];
      const float4 r1 = lds[k * 4 + 1];
      const float4 r2 = lds[k * 4 + 2];
      const float4 r3 = lds[k * 4 + 3];
      const float u0 = fmaf(r0.x, xf, fmaf(r0.w, yf, r1.z));
      const float u1 = fmaf(r0.y, xf, fmaf(r1.x, yf, r1.w));
      const float u2 = fmaf(r0.z, xf, fmaf(r1.y, yf, r2.x));
      const float q0 = r2.y, q1 = r2.z, q2 = r2.w;
      const float vsv = fmaf(u0, u0, fmaf(u1, u1, u2 * u2));
      const float psv = fmaf(q0, u0, fmaf(q1, u1, q2 * u2));
      const float iv = __builtin_amdgcn_rcpf(vsv);
      const float t = psv * iv;
      const float v0 = fmaf(t, u0, -q0);
      const float v1 = fmaf(t, u1, -q1);
      const float v2 = fmaf(t, u2, -q2);
      const float d0 = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
      const float d2l = fmaf(kd, d0, r3.y);
      const float e = __builtin_amdgcn_exp2f(d2l);
      const float l = fmaf(b3, d2l, nb2l * t);
      float s = __builtin_amdgcn_exp2f(l - M);
      s = (t > 0.0f) ? s : 0.0f;
      const float cs = cD * s;
      const float dt = t - D;
      float gd2 = fmaf(b3 * cs, dt, cE * e);
      float gt_ = cs * fmaf(-b2, dt, 1.0f);
      if (!inimg || !(gd2 == gd2) || !(gt_ == gt_)) { gd2 = 0.0f; gt_ = 0.0f; }
      const float gtv = gt_ * iv;
      const float m2t = -2.0f * t;
      const float gdt = gd2 * t;
      float val[16];
      val[0] = fmaf(gd2, v0, gtv * u0);
      val[1] = fmaf(gd2, v1, gtv * u1);
      val[2] = fmaf(gd2, v2, gtv * u2);
      const float gu0 = fmaf(gtv, fmaf(m2t, u0, q0), -(gdt * v0));
      const float gu1 = fmaf(gtv, fmaf(m2t, u1, q1), -(gdt * v1));
      const float gu2 = fmaf(gtv, fmaf(m2t, u2, q2), -(gdt * v2));
      val[3] = gu0 * xf; val[4] = gu1 * xf; val[5] = gu2 * xf;
      val[6] = gu0 * yf; val[7] = gu1 * yf; val[8] = gu2 * yf;
      val[9] = gu0; val[10] = gu1; val[11] = gu2;
      val[12] = gd2;
      val[13] = 0.0f; val[14] = 0.0f; val[15] = 0.0f;
      int idx;
      const float red = wave_reduce16(val, lane, &idx);
      if ((lane & 3) == 0 && idx < 13) atomicAdd(&acc[k * 16 + idx], red);
    }
    __syncthreads();
    // flush this batch's per-Gaussian sums
    for (int i = threadIdx.x; i < nb * 16; i += NT) {
      const int k = i >> 4, c = i & 15;
      if (c < 13) {
        const float vsum = acc[i];
        if (vsum != 0.0f) atomicAdd(A.gacc + (jb.rec_base + gid_sh[k]) * 16 + c, vsum);
      }
      acc[i] = 0.0f;
    }
    __syncthreads();
  }
  (void)NW;
}

// Per (Gaussian, view): chain rule from the 13 sums to d(mu), d(U upper), d(w).
//   M = A R^T, a = M[:,0]/fx, b = M[:,1]/fy, g = M[:,2], q = A (mu - o), A = U^T.
__global__ __launch_bounds__(256) void k_bwd_finalize(const Job* __restrict__ jobs,
                                                      const BwdJob* __restrict__ bjobs,
                                                      const float* __restrict__ gacc) {
  const Job& jb = jobs[blockIdx.y];
  const BwdJob& bj = bjobs[blockIdx.y];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= jb.n) return;
  const float* G = gacc + (jb.rec_base + j) * 16;
  const float* P = jb.prec + 9 * (int64_t)j;
  const float* mu = jb.means + 3 * (int64_t)j;
  const float A[3][3] = {{P[0], 0.f, 0.f}, {P[1], P[4], 0.f}, {P[2], P[5], P[8]}};
  const float p[3] = {mu[0] - jb.o[0], mu[1] - jb.o[1], mu[2] - jb.o[2]};
  float dM[3][3];
  for (int i = 0; i < 3; ++i) {
    dM[i][0] = G[3 + i] / jb.fx;
    dM[i][1] = G[6 + i] / jb.fy;
    dM[i][2] = G[9 + i];
  }
  float dA[3][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k)
      dA[i][k] = dM[i][0] * jb.R[0 * 3 + k] + dM[i][1] * jb.R[1 * 3 + k] + dM[i][2] * jb.R[2 * 3 + k] +
                 G[i] * p[k];
  float dmu[3];
  for (int k = 0; k < 3; ++k) dmu[k] = A[0][k] * G[0] + A[1][k] * G[1] + A[2][k] * G[2];
  if (bj.dmeans)
    for (int k = 0; k < 3; ++k) atomicAdd(bj.dmeans + 3 * (int64_t)j + k, dmu[k]);
  if (bj.dprec)
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k <= i; ++k) atomicAdd(bj.dprec + 9 * (int64_t)j + 3 * k + i, dA[i][k]);  // U[k][i] = A[i][k]
  if (bj.dw) atomicAdd(bj.dw + j, G[12]);
}

}  // namespace fe

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
namespace fe {

hipError_t launch_zero(hipStream_t s, float* p, int64_t n) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

template <int TS, int GB>
static hipError_t bwd_tg(hipStream_t s, int64_t total_tiles, const BwdArgs& a) {
  hipLaunchKernelGGL((k_bwd_blend<TS, GB>), dim3((unsigned)total_tiles), dim3(TS * TS), 0, s, a);
  return hipGetLastError();
}

template <int TS>
static hipError_t bwd_t(hipStream_t s, int gb, int64_t total_tiles, const BwdArgs& a) {
  switch (gb) {
    case 32: return bwd_tg<TS, 32>(s, total_tiles, a);
    case 64: return bwd_tg<TS, 64>(s, total_tiles, a);
    case 128: return bwd_tg<TS, 128>(s, total_tiles, a);
    case 256: return bwd_tg<TS, 256>(s, total_tiles, a);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_bwd_blend(hipStream_t s, int ts, int gb, int64_t total_tiles, const Job* jobs,
                            const void* bjobs, int n_jobs, const float* rec, const unsigned long long* off1,
                            const unsigned long long* off2, const uint32_t* list, const float* st,
                            int64_t total_pix, float* gacc, int dense) {
  BwdArgs a{jobs, reinterpret_cast<const BwdJob*>(bjobs), n_jobs, rec, off1, off2, list, st, total_pix, gacc, dense};
  switch (ts) {
    case 8: return bwd_t<8>(s, gb, total_tiles, a);
    case 16: return bwd_t<16>(s, gb, total_tiles, a);
    case 32: return bwd_t<32>(s, gb, total_tiles, a);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_bwd_finalize(hipStream_t s, const Job* jobs, const void* bjobs, int n_jobs, int max_n,
                               const float* gacc) {
  if (max_n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bwd_finalize, dim3((max_n + 255) / 256, n_jobs), dim3(256), 0, s, jobs,
                     reinterpret_cast<const BwdJob*>(bjobs), gacc);
  return hipGetLastError();
}

size_t bwd_job_bytes() { return sizeof(BwdJob); }

}  // namespace fe
