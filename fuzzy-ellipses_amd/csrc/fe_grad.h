// Per-pair gradient sums and the direct-form chain rule, shared by the backward blend and the
// feature backward (gfx950).
//
// Given, for one (pixel, Gaussian) pair, dL/dd2 (gd2, natural units; d2 = w - d0/2) and dL/dt
// (g1, t = t*):
//  * blend-record form (fe_common.h): with X, Y1, Y2 the linear forms of p = (x', y', 1),
//    rr = 1/(X^2 + W), t = X rr, Wr = W rr, the 11 sums per (Gaussian, view) are
//      sum g_X p (3), sum c Y1 p (3), sum c Y2 p (3), sum g_d2, sum g_d2 W r
//    with Gq = g_d2 |q|^2, h = 2 g1 + Gq X, g_X = rr (h Wr - g1), c = h t rr;
//  * direct form: with u the whitened ray, q the whitened centre offset, v = t u - q,
//    iv = 1/(u.u), the 13 sums G_q = sum (g_d2 v + g1 iv u), G_a/G_b/G_g = sum g_u (x', y', 1)
//    with g_u = -((g1 iv + g_d2 t) v + g1 iv t u), and G_w = sum g_d2.
#pragma once

#include "fe_common.h"

namespace fe {

__device__ __forceinline__ void rec2_pair_sums(float gd2, float g1, float X, float Y1, float Y2, float rr, float t,
                                               float Wr, float qq, float xf, float yf, float* G) {
  const float Gq = gd2 * qq;
  const float h = fmaf(Gq, X, g1 + g1);
  const float gX = rr * fmaf(h, Wr, -g1);   // = g1 rr + Gq t - t^2 h (1 - X^2 rr = W rr)
  const float c = h * (t * rr);
  const float c1 = c * Y1, c2 = c * Y2;
  G[0] = fmaf(gX, xf, G[0]); G[1] = fmaf(gX, yf, G[1]); G[2] += gX;
  G[3] = fmaf(c1, xf, G[3]); G[4] = fmaf(c1, yf, G[4]); G[5] += c1;
  G[6] = fmaf(c2, xf, G[6]); G[7] = fmaf(c2, yf, G[7]); G[8] += c2;
  G[9] += gd2;
  G[10] = fmaf(gd2, Wr, G[10]);
}

// G: [0..2] G_q, [3..5] G_a, [6..8] G_b, [9..11] G_g, [12] G_w
__device__ __forceinline__ void direct_pair_sums(float gd2, float g1, float t, float iv, const float* u,
                                                 const float* v, float xf, float yf, float* G) {
  const float gtv = g1 * iv;
  const float c1 = fmaf(gd2, t, gtv);     // g_t/vsv + g_d2 t
  const float c2 = gtv * t;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    G[i] = fmaf(gd2, v[i], fmaf(gtv, u[i], G[i]));
    const float gu = -fmaf(c1, v[i], c2 * u[i]);
    G[3 + i] = fmaf(gu, xf, G[3 + i]);
    G[6 + i] = fmaf(gu, yf, G[6 + i]);
    G[9 + i] += gu;
  }
  G[12] += gd2;
}

// Chain rule of one (tile, Gaussian) pair's 13 direct-form sums into the scene gradients of
// Gaussian sg (scene index):
//   dM = [G_a/fx, G_b/fy, G_g],  dA = dM R + G_q p^T (lower part),  dmu = A^T G_q,
//   d(prec)[k][i] += dA[i][k] for k <= i,  dw += G_w
__device__ __forceinline__ void direct_chain(const Job& jb, const BwdJob& bj, int64_t sg, float Gq0, float Gq1,
                                             float Gq2, float Ga0, float Ga1, float Ga2, float Gb0, float Gb1,
                                             float Gb2, float Gg0, float Gg1, float Gg2, float Gw) {
  const float ifx = 1.0f / jb.fx, ify = 1.0f / jb.fy;
  const float* Pp = jb.prec + 9 * sg;
  const float* mu = jb.means + 3 * sg;
  const float u00 = Pp[0], u01 = Pp[1], u02 = Pp[2], u11 = Pp[4], u12 = Pp[5], u22 = Pp[8];
  const float pv[3] = {mu[0] - jb.o[0], mu[1] - jb.o[1], mu[2] - jb.o[2]};
  const float dm[3][3] = {{Ga0 * ifx, Gb0 * ify, Gg0}, {Ga1 * ifx, Gb1 * ify, Gg1}, {Ga2 * ifx, Gb2 * ify, Gg2}};
  const float gq[3] = {Gq0, Gq1, Gq2};
  float dA[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k <= i; ++k)
      dA[i][k] = fmaf(dm[i][0], jb.R[k], fmaf(dm[i][1], jb.R[3 + k], fmaf(dm[i][2], jb.R[6 + k], gq[i] * pv[k])));
  if (bj.dmeans) {
    float* dmu = bj.dmeans + 3 * sg;
    atomicAdd(dmu + 0, fmaf(u00, Gq0, fmaf(u01, Gq1, u02 * Gq2)));
    atomicAdd(dmu + 1, fmaf(u11, Gq1, u12 * Gq2));
    atomicAdd(dmu + 2, u22 * Gq2);
  }
  if (bj.dprec) {
    float* dp = bj.dprec + 9 * sg;
    atomicAdd(dp + 0, dA[0][0]);
    atomicAdd(dp + 1, dA[1][0]);
    atomicAdd(dp + 2, dA[2][0]);
    atomicAdd(dp + 4, dA[1][1]);
    atomicAdd(dp + 5, dA[2][1]);
    atomicAdd(dp + 8, dA[2][2]);
  }
  if (bj.dw) atomicAdd(bj.dw + sg, Gw);
}

}  // namespace fe
