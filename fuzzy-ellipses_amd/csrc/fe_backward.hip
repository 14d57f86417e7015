// Backward kernels of the fuzzy-metaball rasterizer (gfx950).
//
// No reference counterpart: the reference obtained gradients from jax.grad of
// baseline/fm_render.py:11-109.  This pass recomputes every listed pair from the record and
// the saved per-pixel aggregates (M, S, Z, E, D) instead of storing per-pair values.
//
// Per pixel (upstream gD = dL/dD, gA = dL/dalpha), D = Z/S, alpha = 1 - exp(-E):
//   cE = gA exp(-E),  cD = gD / S,  s_j = [t_j>0] 2^(l2_j - M)
//   g_d2 = cE e_j + b3 cD s_j (t_j - D)
//   g_t  = cD s_j (1 - b2 (t_j - D)) + cD 1e-20
// with cD = gD/(S + 1e-20 n).  The last term and one more are the 1e-20 weight floor of
// fm_render.py:96 under reverse-mode differentiation: D = (Z + 1e-20 sum nan_to_num(t))/
// (S + 1e-20 n) gives every pair dD/dt = 1e-20/(S + 1e-20 n), and because the weights are
// exp(l - max l) (util_render.py:5-9) the pair holding the maximum logit also gets
//   dL/dl_max = cX = -cD (Z - D S),
// which is zero up to rounding unless the floor is visible (S below ~1e-11 n).  Tiles with
// such a pixel run the argmax variant; elsewhere |cX| is below 1e-9 of the other terms.
// and through t* = (q.u)/(u.u), d2 = w - |t* u - q|^2 / 2 (envelope: dd0/dt = 0), v = t u - q:
//   g_q = g_d2 v + (g_t/vsv) u
//   g_u = -g_d2 t v + (g_t/vsv) (q - 2 t u)
//   g_w = g_d2
// Per (tile, Gaussian) the kernel accumulates 13 sums: G_q (3), G_a = sum g_u x' (3),
// G_b = sum g_u y' (3), G_g = sum g_u (3), G_w, then maps them through the chain rule to the
// 10 gradient components of (mu, upper U, w) and adds those to the scene gradients.
//
// Views with blend records (fe_common.h) recompute each pair from the three linear forms
// X, Y1, Y2 of p = (x', y', 1) instead (t = X r, d0 = |q|^2 W r, r = 1/(X^2 + W)):
//   g_X  = g_t r + G t - t^2 h,   g_Yi = -c Y_i,
//   with G = g_d2 |q|^2, h = 2 g_t + G X, c = h t r,
// with 11 sums per (Gaussian, view): sum g_X p (3), sum c Y_i p (3 + 3), sum g_d2, and
// sum g_d2 W r (d2 depends on |q|^2 directly: dd2/d|q|^2 = -W r / 2).  The chain rule
// to (mu, U) goes through the blend record K = M'^T B, B = [q/|q|^2, e1/|q|, e2/|q|]
// (M' = [a b g]): dM' = B dK^T, dB = M' dK, and for the basis the derivative that keeps
// e1, e2 orthonormal with no rotation in their plane (W does not depend on that rotation):
//   dq = (dB_0 - 2 n (n.dB_0))/|q|^2 - sum_i (e_i (n.dB_i) + n (e_i.dB_i))/|q|^2 - q sum g_d2 W r,
// with n = q/|q|, once per (tile, Gaussian).
#include "fe_common.h"
#include "fe_grad.h"
#include "fe_launch.h"

#include <string.h>

#include <algorithm>

namespace fe {

// Loop over the tile's pixels for one Gaussian (this lane), accumulating its 13 sums.
// With v = t u - q, the t-derivative direction q - 2 t u equals -(v + t u), so
//   g_u = -(g_t/vsv + g_d2 t) v - (g_t/vsv) t u.
// SAFE zeroes non-finite per-pair terms (views with degenerate records only).
template <bool SAFE>
__device__ __forceinline__ void bwd_pixels(const float4* pxa, const float4* pxb, const float2* pxc, int npx, float a0,
                                           float a1,
                                           float a2, float bb0, float bb1, float bb2, float g0, float g1, float g2,
                                           float q0, float q1, float q2, float wl, float b3, float b2, float nb2l,
                                           float kd, float& Gq0, float& Gq1, float& Gq2, float& Ga0, float& Ga1,
                                           float& Ga2, float& Gb0, float& Gb1, float& Gb2, float& Gg0, float& Gg1,
                                           float& Gg2, float& Gw) {
  for (int p = 0; p < npx; ++p) {
    const float4 pa = pxa[p];
    const float4 pb = pxb[p];
    const float xf = pa.x, yf = pa.y, M = pa.z, D = pa.w, cE = pb.x, cD = pb.y, cF = pb.w;
    const float u0 = fmaf(a0, xf, fmaf(bb0, yf, g0));
    const float u1 = fmaf(a1, xf, fmaf(bb1, yf, g1));
    const float u2 = fmaf(a2, xf, fmaf(bb2, yf, g2));
    const float vsv = fmaf(u0, u0, fmaf(u1, u1, u2 * u2));
    const float psv = fmaf(q0, u0, fmaf(q1, u1, q2 * u2));
    const float iv = __builtin_amdgcn_rcpf(vsv);
    const float t = psv * iv;
    const float v0 = fmaf(t, u0, -q0);
    const float v1 = fmaf(t, u1, -q1);
    const float v2 = fmaf(t, u2, -q2);
    const float d0 = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
    const float d2l = fmaf(kd, d0, wl);
    const float e = __builtin_amdgcn_exp2f(d2l);
    const float l = fmaf(b3, d2l, nb2l * t);
    float s = __builtin_amdgcn_exp2f(l - M);
    s = (t > 0.0f) ? s : 0.0f;
    const float cs = cD * s;
    const float dt = t - D;
    const float ex = (l == M) ? pxc[p].y : 0.0f;   // the pair holding the maximum logit
    float gd2 = fmaf(b3, ex, fmaf(b3 * cs, dt, cE * e));
    float gtv = fmaf(-b2, ex, fmaf(cs, fmaf(-b2, dt, 1.0f), cF)) * iv;
    if (SAFE && (!(gd2 == gd2) || !(gtv == gtv))) { gd2 = 0.0f; gtv = 0.0f; }
    Gq0 = fmaf(gd2, v0, fmaf(gtv, u0, Gq0));
    Gq1 = fmaf(gd2, v1, fmaf(gtv, u1, Gq1));
    Gq2 = fmaf(gd2, v2, fmaf(gtv, u2, Gq2));
    const float c1 = fmaf(gd2, t, gtv);     // g_t/vsv + g_d2 t
    const float c2 = gtv * t;
    const float gu0 = -fmaf(c1, v0, c2 * u0);
    const float gu1 = -fmaf(c1, v1, c2 * u1);
    const float gu2 = -fmaf(c1, v2, c2 * u2);
    Ga0 = fmaf(gu0, xf, Ga0); Ga1 = fmaf(gu1, xf, Ga1); Ga2 = fmaf(gu2, xf, Ga2);
    Gb0 = fmaf(gu0, yf, Gb0); Gb1 = fmaf(gu1, yf, Gb1); Gb2 = fmaf(gu2, yf, Gb2);
    Gg0 += gu0; Gg1 += gu1; Gg2 += gu2;
    Gw += gd2;
  }
}

// Chain rule of the blend-record sums of Gaussian gid (scene index) of view jb into its scene
// gradients (10 atomics).  Float32 (float64 here doubles the kernel's registers and halves
// its occupancy); against a float64 chain it differs by ~1e-6 of the largest gradient.
__device__ __forceinline__ void rec2_chain(const Job& jb, const BwdJob& bj, int64_t gid, const float* GX,
                                           const float* GY1, const float* GY2, float Gw, float Gqq) {
  const float* P = jb.prec + 9 * gid;
  const float* mu = jb.means + 3 * gid;
  const float A[3][3] = {{P[0], 0.0f, 0.0f}, {P[1], P[4], 0.0f}, {P[2], P[5], P[8]}};
  const float pv[3] = {mu[0] - jb.o[0], mu[1] - jb.o[1], mu[2] - jb.o[2]};
  const float sc[3] = {1.0f / jb.fx, 1.0f / jb.fy, 1.0f};
  float Mp[3][3], q[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Mp[i][c] = fmaf(A[i][0], jb.R[3 * c], fmaf(A[i][1], jb.R[3 * c + 1], A[i][2] * jb.R[3 * c + 2])) * sc[c];
    q[i] = fmaf(A[i][0], pv[0], fmaf(A[i][1], pv[1], A[i][2] * pv[2]));
  }
  const float qq = dot3(q, q);
  const float iq = 1.0f / sqrtf(qq), iqq = 1.0f / qq;
  const float n[3] = {q[0] * iq, q[1] * iq, q[2] * iq};
  float e1[3], e2[3];
  ortho_basis(n, e1, e2);
  // dK[c][j] (c: coefficient of x', y', 1; j: X, Y1, Y2); the Y sums hold +c Y_i p
  float dK[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dK[c][0] = GX[c];
    dK[c][1] = -GY1[c];
    dK[c][2] = -GY2[c];
  }
  const float B[3][3] = {{q[0] * iqq, e1[0] * iq, e2[0] * iq},
                         {q[1] * iqq, e1[1] * iq, e2[1] * iq},
                         {q[2] * iqq, e1[2] * iq, e2[2] * iq}};
  float dM[3][3], dB[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      dM[i][c] = fmaf(B[i][0], dK[c][0], fmaf(B[i][1], dK[c][1], B[i][2] * dK[c][2])) * sc[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) dB[i][j] = fmaf(Mp[i][0], dK[0][j], fmaf(Mp[i][1], dK[1][j], Mp[i][2] * dK[2][j]));
  }
  const float nb0 = fmaf(n[0], dB[0][0], fmaf(n[1], dB[1][0], n[2] * dB[2][0]));
  const float nb1 = fmaf(n[0], dB[0][1], fmaf(n[1], dB[1][1], n[2] * dB[2][1]));
  const float nb2 = fmaf(n[0], dB[0][2], fmaf(n[1], dB[1][2], n[2] * dB[2][2]));
  const float eb = fmaf(e1[0], dB[0][1], fmaf(e1[1], dB[1][1], e1[2] * dB[2][1])) +
                   fmaf(e2[0], dB[0][2], fmaf(e2[1], dB[1][2], e2[2] * dB[2][2]));
  float dq[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dq[i] = fmaf(dB[i][0] - fmaf(n[i], fmaf(2.0f, nb0, eb), fmaf(e1[i], nb1, e2[i] * nb2)), iqq, -q[i] * Gqq);
  // dA = dM R + dq pv^T (lower part); dmu = A^T dq; d(prec)[k][i] = dA[i][k]
  if (bj.dmeans) {
    float* dmu = bj.dmeans + 3 * gid;
    atomicAdd(dmu + 0, fmaf(A[0][0], dq[0], fmaf(A[1][0], dq[1], A[2][0] * dq[2])));
    atomicAdd(dmu + 1, fmaf(A[1][1], dq[1], A[2][1] * dq[2]));
    atomicAdd(dmu + 2, A[2][2] * dq[2]);
  }
  if (bj.dprec) {
    float* dp = bj.dprec + 9 * gid;
    float dA[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k <= i; ++k)
        dA[i][k] = fmaf(dM[i][0], jb.R[k], fmaf(dM[i][1], jb.R[3 + k], fmaf(dM[i][2], jb.R[6 + k], dq[i] * pv[k])));
    atomicAdd(dp + 0, dA[0][0]);
    atomicAdd(dp + 1, dA[1][0]);
    atomicAdd(dp + 2, dA[2][0]);
    atomicAdd(dp + 4, dA[1][1]);
    atomicAdd(dp + 5, dA[2][1]);
    atomicAdd(dp + 8, dA[2][2]);
  }
  if (bj.dw) atomicAdd(bj.dw + gid, Gw);
}


struct BwdArgs {
  const Job* jobs;
  const BwdJob* bjobs;
  int n_jobs;
  const float* rec;
  const float* rec2;
  const uint32_t* tile_np;
  const unsigned long long* off1;
  const unsigned long long* off2;
  const uint32_t* list;
  const float* st;
  int64_t total_pix;
  const int32_t* job_flags;
  float* gsum;          // the 11 blend-record sums per (Gaussian, view) record, one plane each
  int64_t total_rec;    // plane stride
  int dense;
  const int32_t* tord;  // dispatch order (blocks of tile tord[b] are b kb .. b kb + kb - 1), or null
  int kb;               // blocks per tile (each takes every kb-th chunk of 64 Gaussians)
};

// Sums per record (kGsumPlanes = 11).  Planes, not records of 11 floats: device-scope float
// atomics to one cache line serialise, and a record's 11 sums are added together by every tile.

// Per-pixel terms of the backward (upstream gradients through the saved aggregates).
struct PixTerms {
  float xf, yf, M, D, cE, cD, cDb, cF, cX;
  float A1, A2, B1;   // g_t = s (A1 - A2 t) + cF, g_d2 = s (cDb t + B1) + cE e
  bool floor_px;
};

__device__ __forceinline__ PixTerms pix_terms(const BwdArgs& A, const Job& jb, const BwdJob& bj, int64_t lp, int px,
                                              int py, float ncnt, bool rec2) {
  PixTerms t;
  t.xf = jb.ray_xy ? jb.ray_xy[2 * lp] : (float)px - jb.cx;
  t.yf = jb.ray_xy ? jb.ray_xy[2 * lp + 1] : (float)py - jb.cy;
  const int64_t pix = jb.pix_base + lp;
  const float M = A.st[pix];
  const float S = A.st[A.total_pix + pix];
  const float Z = A.st[2 * A.total_pix + pix];
  const float E = A.st[3 * A.total_pix + pix];
  const float D = A.st[5 * A.total_pix + pix];
  const float gD = bj.gD ? bj.gD[lp] : 0.0f;
  const float gA = bj.gA ? bj.gA[lp] : 0.0f;
  const float den = S + 1e-20f * ncnt;
  float cE = gA * expf(-E);
  float cD = (den > 0.0f) ? gD / den : 0.0f;
  if (!(cE == cE)) cE = 0.0f;
  if (!(cD == cD) || !(D == D)) cD = 0.0f;
  float cX = -cD * fmaf(-D, S, Z);
  if (!(cX == cX) || fabsf(cX) > 3.0e38f) cX = 0.0f;
  t.M = M; t.D = D; t.cE = cE; t.cD = cD; t.cDb = jb.beta3 * cD; t.cF = 1e-20f * cD; t.cX = cX;
  // cD s (1 - b2 (t - D)) = s (A1 - A2 t) and b3 cD s (t - D) = s (cDb t + B1)
  t.A2 = jb.beta2 * cD;
  t.A1 = fmaf(t.A2, D, cD);
  t.B1 = -(t.cDb * D);
  if (cD == 0.0f || !(t.A1 == t.A1) || !(t.B1 == t.B1)) { t.A1 = 0.0f; t.A2 = 0.0f; t.B1 = 0.0f; }
  t.floor_px = !(S >= 1e-11f * ncnt);
  // A pixel whose weight sum is exactly that of its maximum-logit pair (every other weight
  // was rounded away in the forward's sums) has dD/dl = 0 for all its pairs in the forward's
  // arithmetic: its D is that pair's t, and the others' weights are below float resolution.
  // Computed as b3 cD s (t - D), the dominant pair's term would be a difference of two
  // separately rounded values of size |D| (D = Z/S from the forward's sums, t the pair's
  // own), an error of ~2^-23 |D| that the chain rule scales by dd2/dtheta, large at pixels
  // far (in whitened units) from the Gaussian.  Such pixels keep only dL/dt = cD s and the
  // alpha term.  The maximum pair's weight is recomputed with the forward's operations from
  // the exact maximum Mt it saved.
  if (rec2 && !t.floor_px && S > 0.0f) {
    const float Mt = A.st[6 * A.total_pix + pix];
    if (S == __builtin_amdgcn_exp2f(fmaf(jb.beta3, Mt, -M))) {
      t.A1 = cD;
      t.A2 = 0.0f;
      t.cDb = 0.0f;
      t.B1 = 0.0f;
      t.cX = 0.0f;
    }
  }
  return t;
}

// Blend-record loop over pixels [p0, p1) for one Gaussian (this lane); see the header
// comment.  Used for ray-list tiles: pxa holds (x, y, M, D) with (x, y) the ray's
// camera-plane coordinates, pxb (cE, cD, b3 cD, 1e-20 cD), pxc (1e-20 cD, the maximum-logit
// term cX).
// Sums: G[0..2] = sum g_X p, G[3..5] = sum c Y1 p, G[6..8] = sum c Y2 p, G[9] = sum g_d2,
// G[10] = sum g_d2 W r.
template <bool MASK, bool ARGMAX>
__device__ __forceinline__ void bwd_pixels_rec2(const float4* pxa, const float4* pxb, const float2* pxc, int p0,
                                                int p1, const float* k, float wL, float ncG, float qq, float b3,
                                                float b2, float kap, float* G) {
  const float kx0 = k[0], kx1 = k[1], kx2 = k[2], k10 = k[3], k11 = k[4], k12 = k[5], k20 = k[6], k21 = k[7],
              k22 = k[8];
  for (int p = p0; p < p1; ++p) {
    const float4 pa = pxa[p];
    const float4 pb = pxb[p];
    const float xf = pa.x, yf = pa.y, M = pa.z, D = pa.w, cE = pb.x, cD = pb.y, cDb = pb.z, cF = pb.w;
    const float X = fmaf(kx0, xf, fmaf(kx1, yf, kx2));
    const float Y1 = fmaf(k10, xf, fmaf(k11, yf, k12));
    const float Y2 = fmaf(k20, xf, fmaf(k21, yf, k22));
    const float W = fmaf(Y1, Y1, Y2 * Y2);
    const float rr = __builtin_amdgcn_rcpf(fmaf(X, X, W));
    const float t = X * rr;
    const float Wr = W * rr;
    const float d2l = fmaf(ncG, Wr, wL);
    const float e = __builtin_amdgcn_exp2f(d2l);
    const float lt = fmaf(kap, t, d2l);
    float s = __builtin_amdgcn_exp2f(fmaf(b3, lt, -M));
    if (MASK) s = (t > 0.0f) ? s : 0.0f;
    const float dt = t - D;
    float g1 = fmaf(cD * s, fmaf(-b2, dt, 1.0f), cF);     // dL/dt
    float gd2 = fmaf(cDb * s, dt, cE * e);                // dL/dd2
    if (ARGMAX) {
      const float ex = (b3 * lt == M) ? pxc[p].y : 0.0f;   // the pair holding the maximum logit
      g1 = fmaf(-b2, ex, g1);
      gd2 = fmaf(b3, ex, gd2);
    }
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
}

// Row form of the blend-record loop for pixel-grid tiles (row width wv), pixels [p0, p1).
// pxa holds (x', x'^2, M, cE) and pxb (A1, A2, b3 cD, B1) with (x', y') the pixel's
// coordinates relative to the centre of its 32x32-pixel block (form_coord); pxc (1e-20 cD,
// cX).  The forms X, Y1, Y2 take their y-part once per row, with the record's constant terms
// moved to the block's centre, by the forward blend's operations: every t*, d2 and weight s recomputed here is the
// forward's value bit for bit.  (Where logits are large, at pixels far in whitened units
// from every Gaussian, a logit rounded differently in the two passes would move its weight
// by a few percent, 2^-24 |logit| in the exponent, which the forward's S, made from its own
// weights, would no longer cancel; it also lets the argmax variant, on tiles where the
// 1e-20 floor is visible, find the pair holding the maximum logit by equality.)  The sums
// go per row and per tile-local moment:
//   sum g_X p' from sum g_X x' and per-row sums of g_X,
//   sum c Y_i p' = C k_i' with C = sum c p' p'^T (6 moments) and k_i' the form's local
//   coefficients, so the per-pixel work does not depend on the two forms Y1, Y2
//   (local coordinates keep the moments small: |x'|, |y'| <= 15.5),
// and are returned for the global p = p' + (ox, oy, 0) in the layout of bwd_pixels_rec2.
// HALVES: the two half-waves take consecutive row blocks of the range (lanes 32-63 are
// shifted down by roff rows) for the same 32 Gaussians; the caller adds the halves' sums.
template <bool MASK, bool ARGMAX, bool HALVES>
__device__ __forceinline__ void bwd_rows_rec2(const float4* pxa, const float4* pxb, const float2* pxc, int p0,
                                              int p1, int roff, int wv, int px0, int py0, float cx,
                                              float cy, const float* k, float wL, float ncG, float qq, float b3,
                                              float b2, float kap, float* G) {
  const float kx0 = k[0], kx1 = k[1], k10 = k[3], k11 = k[4], k20 = k[6], k21 = k[7];
  // global coordinates of the block's centre, and the forms' constants there (the operations
  // of localize_rec2, as the forward staged them)
  const float ox = form_origin(px0, cx), oy = form_origin(py0, cy);
  const float kxl = fmaf(kx0, ox, fmaf(kx1, oy, k[2]));
  const float k1l = fmaf(k10, ox, fmaf(k11, oy, k[5]));
  const float k2l = fmaf(k20, ox, fmaf(k21, oy, k[8]));
  float SgXx = 0.f, SgXy = 0.f, SgX = 0.f;
  float Cxx = 0.f, Cxy = 0.f, Cyy = 0.f, Cx = 0.f, Cy = 0.f, C1 = 0.f;
  float S9 = 0.f, S10 = 0.f;
  const int ro = HALVES ? roff : 0;   // this lane's row offset
  const int poff = ro * wv;
  int p = p0;
  while (p < p1) {
    const int iy = p / wv;
    const int pe = min(p1, (iy + 1) * wv);
    const float yl = form_coord(py0 + iy + ro);
    const float XR = fmaf(kx1, yl, kxl), Y1R = fmaf(k11, yl, k1l), Y2R = fmaf(k21, yl, k2l);
    float Rg = 0.f, Rcx = 0.f, Rc = 0.f;
    auto pixel = [&](int pp) {
      const float4 pa = pxa[pp + poff];
      const float4 pb = pxb[pp + poff];
      const float xl = pa.x, xl2 = pa.y, M = pa.z, cE = pa.w;
      const float A1 = pb.x, A2 = pb.y, cDb = pb.z, B1 = pb.w;
      const float X = fmaf(kx0, xl, XR);
      const float Y1 = fmaf(k10, xl, Y1R);
      const float Y2 = fmaf(k20, xl, Y2R);
      const float W = fmaf(Y1, Y1, Y2 * Y2);
      const float rr = __builtin_amdgcn_rcpf(fmaf(X, X, W));
      const float t = X * rr;
      const float Wr = W * rr;
      const float d2l = fmaf(ncG, Wr, wL);
      const float e = __builtin_amdgcn_exp2f(d2l);
      const float lt = fmaf(kap, t, d2l);
      float s = __builtin_amdgcn_exp2f(fmaf(b3, lt, -M));
      if (MASK) s = (t > 0.0f) ? s : 0.0f;
      float g1, gd2;
      if (ARGMAX) {
        const float2 pc = pxc[pp + poff];
        g1 = fmaf(s, fmaf(-A2, t, A1), pc.x);                  // dL/dt
        gd2 = fmaf(s, fmaf(cDb, t, B1), cE * e);                // dL/dd2
        const float ex = (b3 * lt == M) ? pc.y : 0.0f;          // the pair holding the maximum logit
        g1 = fmaf(-b2, ex, g1);
        gd2 = fmaf(b3, ex, gd2);
      } else {
        // the floor term 1e-20 cD of dL/dt is below 1e-9 of the rest on these tiles
        g1 = s * fmaf(-A2, t, A1);
        gd2 = fmaf(s, fmaf(cDb, t, B1), cE * e);
      }
      const float Gq = gd2 * qq;
      const float h = fmaf(Gq, X, g1 + g1);
      const float gX = rr * fmaf(h, Wr, -g1);   // = g1 rr + Gq t - t^2 h (1 - X^2 rr = W rr)
      const float c = h * (t * rr);
      SgXx = fmaf(gX, xl, SgXx); Rg += gX;
      Cxx = fmaf(c, xl2, Cxx); Rcx = fmaf(c, xl, Rcx); Rc += c;
      S9 += gd2;
      S10 = fmaf(gd2, Wr, S10);
    };
    // Whole runs of 16 pixels unrolled (LDS addresses as immediate offsets, no per-pixel
    // loop counter), then the rest of the row.  A scheduling barrier after each pixel keeps
    // the compiler from hoisting later pixels' LDS reads: hoisted, they take 94 VGPRs (5 waves
    // per SIMD); fenced, 79 (6).  Backward on the 16-view 480x640 shell: rolled loop
    // 1.31-1.32 ms, unrolled 1.29-1.30, unrolled and fenced 1.28-1.29 (interleaved runs).
    for (; pe - p >= 16; p += 16) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        pixel(p + k);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    for (; p < pe; ++p) pixel(p);
    SgX += Rg; SgXy = fmaf(yl, Rg, SgXy);
    Cx += Rcx; Cxy = fmaf(yl, Rcx, Cxy);
    C1 += Rc; Cy = fmaf(yl, Rc, Cy); Cyy = fmaf(yl * yl, Rc, Cyy);
  }
  G[0] = fmaf(ox, SgX, SgXx); G[1] = fmaf(oy, SgX, SgXy); G[2] = SgX;
  {
    const float sx = fmaf(k10, Cxx, fmaf(k11, Cxy, k1l * Cx));
    const float sy = fmaf(k10, Cxy, fmaf(k11, Cyy, k1l * Cy));
    const float s1 = fmaf(k10, Cx, fmaf(k11, Cy, k1l * C1));
    G[3] = fmaf(ox, s1, sx); G[4] = fmaf(oy, s1, sy); G[5] = s1;
  }
  {
    const float sx = fmaf(k20, Cxx, fmaf(k21, Cxy, k2l * Cx));
    const float sy = fmaf(k20, Cxy, fmaf(k21, Cyy, k2l * Cy));
    const float s1 = fmaf(k20, Cx, fmaf(k21, Cy, k2l * C1));
    G[6] = fmaf(ox, s1, sx); G[7] = fmaf(oy, s1, sy); G[8] = s1;
  }
  G[9] = S9; G[10] = S10;
}

// One block (4 waves) per tile, or kb on small grids (each taking every kb-th chunk of 64
// Gaussians); lane = one Gaussian of the tile's lists (phase 1 then phase 2, or all Gaussians
// in dense mode).  The tile's per-pixel terms (x', y', M, D and the cotangent factors) are
// staged once per block in LDS and read as wave-uniform broadcasts.
//  * Views with blend records: the Gaussians go in chunks of 64, one lane each, and every
//    wave takes the same chunk over its own quarter of the tile's pixels, so the 4 waves
//    carry equal work whatever the list length.  Each lane keeps 11 sums in registers; the 4
//    quarters meet in LDS and each (tile, Gaussian) adds its 11 sums to the per-(Gaussian,
//    view) sums in the workspace with one atomic each.  The chain rule is linear in those
//    sums, so k_bwd_chain applies it once per (Gaussian, view) afterwards.
//  * Other views (degenerate records, b3 <= 0): the waves take chunks of 64 Gaussians over
//    all pixels, with 13 sums per lane and the chain rule per (tile, Gaussian) straight into
//    the scene gradients.
template <int TS>
__global__ __launch_bounds__(256) void k_bwd_blend(BwdArgs A) {
  constexpr int NPIX = TS * TS;
  __shared__ float4 pxa[NPIX];   // x', y', M, D
  __shared__ float4 pxb[NPIX];   // cE, cD, b3 cD, 1e-20 cD
  __shared__ float2 pxc[NPIX];   // 1e-20 cD, cX (maximum-logit term)
  __shared__ float red[4][11][64];

  // kb blocks per tile (kb > 1 on small grids), next to each other in the dispatch order
  const int64_t kb = A.kb, cy = blockIdx.x % A.kb;
  const int64_t bt = blockIdx.x / A.kb, ntiles = gridDim.x / A.kb;
  const int64_t gt = A.tord ? (int64_t)A.tord[bt] : xcd_tile(bt, ntiles);
  const int ji = find_job(A.n_jobs, gt, [&](int k) { return A.jobs[k].tile_base; });
  const Job& jb = A.jobs[ji];
  const BwdJob& bj = A.bjobs[ji];
  const int tile = (int)(gt - jb.tile_base);
  const int tx = tile % jb.tiles_x, ty = tile / jb.tiles_x;
  const int wv = min(TS, jb.W - tx * TS);   // valid width (a ray-list tile is TS wide)
  const int npx = tile_valid(jb, tile, tx, ty, TS);

  int64_t s1, n1, s2, n2;
  if (A.dense) { s1 = 0; n1 = jb.n; s2 = 0; n2 = 0; }
  else {
    s1 = (int64_t)A.off1[gt]; n1 = (int64_t)A.off1[gt + 1] - s1;
    s2 = (int64_t)A.off2[gt]; n2 = (int64_t)A.off2[gt + 1] - s2;
  }
  const int64_t total = n1 + n2;
  // the tile's chunks of 64 Gaussians are shared out over its kb blocks (chunk c to block
  // c mod kb; the sums go out by atomics either way)
  if (total <= 64 * cy) return;   // no chunk for this block (block-uniform)
  const float ncnt = (float)total;

  const int32_t jflags = A.job_flags[ji];
  const bool rec2 = use_rec2(jb, jflags);
  const bool rows = rec2 && !jb.ray_xy;   // pixel-grid tile: row form, tile-local coordinates
  bool floor_px = false;
  for (int i = threadIdx.x; i < npx; i += 256) {
    const int ix = i % wv, iy = i / wv;
    const int px = tx * TS + ix, py = ty * TS + iy;
    const int64_t lp = (int64_t)py * jb.W + px;
    const PixTerms P = pix_terms(A, jb, bj, lp, px, py, ncnt, rec2);
    floor_px = floor_px || P.floor_px;
    if (rows) {
      const float xl = form_coord(tx * TS + ix);
      pxa[i] = make_float4(xl, xl * xl, P.M, P.cE);
      pxb[i] = make_float4(P.A1, P.A2, P.cDb, P.B1);
    } else {
      pxa[i] = make_float4(P.xf, P.yf, P.M, P.D);
      pxb[i] = make_float4(P.cE, P.cD, P.cDb, P.cF);
    }
    pxc[i] = make_float2(P.cF, P.cX);
  }
  const bool floor_tile = __syncthreads_or(floor_px ? 1 : 0) != 0;

  const float b3 = jb.beta3, b2 = jb.beta2;
  const float nb2l = -jb.beta2 * kLog2e;
  // wave index as a wave-uniform value: each wave's pixel range then runs as a scalar loop
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;

  if (rec2) {
    const bool masked = A.dense || !A.tile_np || (A.tile_np[2 * gt] | A.tile_np[2 * gt + 1]) != 0;
    const float kap = nb2l / b3;
    const float4* rec24 = reinterpret_cast<const float4*>(A.rec2) + jb.rec_base * 4;
    float* gsum = A.gsum + jb.rec_base;
    const int q = (npx + 3) >> 2;
    const int p0 = min(wave * q, npx), p1 = min(p0 + q, npx);
    // The next chunk's records are loaded before this chunk's atomics are issued: stores
    // and atomics count in the same wait counter as loads, and waiting for records loaded
    // after the atomics would wait for the atomics to complete at memory.
    // A last chunk of at most 32 Gaussians on a full pixel-grid tile runs on half-waves (lanes
    // l and l + 32 take the same Gaussian over the two halves of the wave's rows), one of at
    // most 16 on quarter-waves (lanes l, l + 16, l + 32, l + 48 over its four quarters) where a
    // quarter is whole rows.
    const bool can_halve = rows && npx == TS * TS;
    const bool can_quarter = can_halve && (q / 4) >= wv && (q / 4) % wv == 0;
    const int half_rows = (q / 2) / wv, quarter_rows = (q / 4) / wv;
    // lanes per Gaussian in the chunk starting at c0 (block-uniform)
    auto split_of = [&](int64_t c0) {
      const int64_t left = total - c0;
      return (can_quarter && left <= 16) ? 4 : ((can_halve && left <= 32) ? 2 : 1);
    };
    auto lane_of = [&](int sp) { return (sp == 4) ? (lane & 15) : ((sp == 2) ? (lane & 31) : lane); };
    int64_t gid_n = 0;
    float4 qa = make_float4(0.f, 0.f, 0.f, 0.f), qb = qa, qc = qa;
    auto fetch = [&](int64_t c0) {
      const int64_t kk = c0 + lane_of(split_of(c0));
      if (kk < total) {
        gid_n = A.dense ? kk : (int64_t)A.list[kk < n1 ? s1 + kk : s2 + (kk - n1)];
        qa = rec24[gid_n * 4 + 0];
        qb = rec24[gid_n * 4 + 1];
        qc = rec24[gid_n * 4 + 2];
      }
    };
    fetch(64 * cy);
    asm volatile("" ::"v"(qa.x), "v"(qa.y), "v"(qa.z), "v"(qa.w), "v"(qb.x), "v"(qb.y), "v"(qb.z), "v"(qb.w),
                 "v"(qc.x), "v"(qc.y), "v"(qc.z), "v"(qc.w));
    for (int64_t c0 = 64 * cy; c0 < total; c0 += 64 * kb) {
      const int sp = split_of(c0);   // block-uniform
      const bool act = c0 + lane_of(sp) < total;
      const int64_t gid = gid_n;
      const float4 q0 = qa, q1 = qb, q2 = qc;
      float G[11];
#pragma unroll
      for (int k = 0; k < 11; ++k) G[k] = 0.0f;
      if (act) {
        const float k9[9] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x};
        if (rows && sp > 1) {
          const int px0 = tx * TS, py0 = ty * TS;
          const int pe = p0 + q / sp, ro = (sp == 4) ? (lane >> 4) * quarter_rows : (lane >> 5) * half_rows;
          if (masked && floor_tile)
            bwd_rows_rec2<true, true, true>(pxa, pxb, pxc, p0, pe, ro, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                            q2.z, q2.w, b3, b2, kap, G);
          else if (masked)
            bwd_rows_rec2<true, false, true>(pxa, pxb, pxc, p0, pe, ro, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                             q2.z, q2.w, b3, b2, kap, G);
          else
            bwd_rows_rec2<false, false, true>(pxa, pxb, pxc, p0, pe, ro, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                              q2.z, q2.w, b3, b2, kap, G);
        } else if (rows) {
          const int px0 = tx * TS, py0 = ty * TS;
          if (masked && floor_tile)
            bwd_rows_rec2<true, true, false>(pxa, pxb, pxc, p0, p1, 0, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                             q2.z, q2.w, b3, b2, kap, G);
          else if (masked)
            bwd_rows_rec2<true, false, false>(pxa, pxb, pxc, p0, p1, 0, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                              q2.z, q2.w, b3, b2, kap, G);
          else
            bwd_rows_rec2<false, false, false>(pxa, pxb, pxc, p0, p1, 0, wv, px0, py0, jb.cx, jb.cy, k9, q2.y,
                                               q2.z, q2.w, b3, b2, kap, G);
        } else {
          if (masked && floor_tile)
            bwd_pixels_rec2<true, true>(pxa, pxb, pxc, p0, p1, k9, q2.y, q2.z, q2.w, b3, b2, kap, G);
          else if (masked) bwd_pixels_rec2<true, false>(pxa, pxb, pxc, p0, p1, k9, q2.y, q2.z, q2.w, b3, b2, kap, G);
          else bwd_pixels_rec2<false, false>(pxa, pxb, pxc, p0, p1, k9, q2.y, q2.z, q2.w, b3, b2, kap, G);
        }
      }
      if (sp > 1) {
        // add the other half- or quarter-waves' sums (the same Gaussian in lanes l ^ 16, l ^ 32)
        if (sp == 4) {
#pragma unroll
          for (int k = 0; k < 11; ++k) G[k] += __shfl_xor(G[k], 16, 64);
        }
#pragma unroll
        for (int k = 0; k < 11; ++k) G[k] += __shfl_xor(G[k], 32, 64);
      }
#pragma unroll
      for (int k = 0; k < 11; ++k) red[wave][k][lane] = G[k];
      __syncthreads();
      if (c0 + 64 * kb < total) fetch(c0 + 64 * kb);
      // wave w adds components k = w, w+4, w+8 of the chunk's 64 Gaussians
      float v[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int k = min(wave + 4 * i, 10);
        v[i] = red[0][k][lane] + red[1][k][lane] + red[2][k][lane] + red[3][k][lane];
      }
      // the next records must be in registers before the atomics go out (see above)
      asm volatile("" ::"v"(qa.x), "v"(qa.y), "v"(qa.z), "v"(qa.w), "v"(qb.x), "v"(qb.y), "v"(qb.z), "v"(qb.w),
                   "v"(qc.x), "v"(qc.y), "v"(qc.z), "v"(qc.w));
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int k = wave + 4 * i;
        if (act && lane < 64 / sp && k < 11) atomicAdd(gsum + (int64_t)k * A.total_rec + gid, v[i]);
      }
      __syncthreads();
    }
    return;
  }

  // ---- direct form (views without blend records) ----
  const float kd = -0.5f * kLog2e;
  const bool safe = (jflags & kFlagDegenerate) != 0;   // the view has degenerate records
  const float4* rec4 = reinterpret_cast<const float4*>(A.rec) + jb.rec_base * 4;

  for (int64_t c0 = (cy * 4 + wave) * 64; c0 < total; c0 += 256 * kb) {
    const int64_t kk = c0 + lane;
    if (kk >= total) break;
    int64_t gid;
    if (A.dense) gid = kk;
    else gid = (kk < n1) ? (int64_t)A.list[s1 + kk] : (int64_t)A.list[s2 + (kk - n1)];
    const float4 r0 = rec4[gid * 4 + 0];
    const float4 r1 = rec4[gid * 4 + 1];
    const float4 r2 = rec4[gid * 4 + 2];
    const float4 r3 = rec4[gid * 4 + 3];
    float Gq0 = 0.f, Gq1 = 0.f, Gq2 = 0.f, Ga0 = 0.f, Ga1 = 0.f, Ga2 = 0.f, Gb0 = 0.f, Gb1 = 0.f, Gb2 = 0.f;
    float Gg0 = 0.f, Gg1 = 0.f, Gg2 = 0.f, Gw = 0.f;
    if (safe) {
      bwd_pixels<true>(pxa, pxb, pxc, npx, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w,
                       r3.y, b3, b2, nb2l, kd, Gq0, Gq1, Gq2, Ga0, Ga1, Ga2, Gb0, Gb1, Gb2, Gg0, Gg1, Gg2, Gw);
    } else {
      bwd_pixels<false>(pxa, pxb, pxc, npx, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w,
                        r3.y, b3, b2, nb2l, kd, Gq0, Gq1, Gq2, Ga0, Ga1, Ga2, Gb0, Gb1, Gb2, Gg0, Gg1, Gg2, Gw);
    }
    // chain rule for this (tile, Gaussian) pair, straight into the scene gradients
    direct_chain(jb, bj, gauss_of(jb, gid), Gq0, Gq1, Gq2, Ga0, Ga1, Ga2, Gb0, Gb1, Gb2, Gg0, Gg1, Gg2, Gw);
  }
}

// Zero the per-(Gaussian, view) sums before the backward (a kernel, not hipMemset: graph-safe).
__global__ void k_fill_zero(float4* p, int64_t n4) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Chain rule once per (Gaussian, view) record of a view with blend records.
__global__ __launch_bounds__(256) void k_bwd_chain(const Job* __restrict__ jobs, const BwdJob* __restrict__ bjobs,
                                                   int n_jobs, const int32_t* __restrict__ job_flags,
                                                   const float* __restrict__ gsum, int64_t total_rec) {
  // one thread per (view, Gaussian) in scene order, so the gradient atomics of a wave fall on
  // consecutive Gaussians; the sums are read from the Gaussian's record
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= total_rec) return;
  const int ji = find_job(n_jobs, i, [&](int k) { return jobs[k].rec_base; });
  const Job& jb = jobs[ji];
  const int64_t gid = i - jb.rec_base;
  if (gid >= jb.n || !use_rec2(jb, job_flags[ji])) return;
  const int64_t r = jb.rec_base + (jb.rank ? jb.rank[gid] : gid);
  float g[kGsumPlanes];
  bool any = false;
#pragma unroll
  for (int k = 0; k < kGsumPlanes; ++k) {
    g[k] = gsum[(int64_t)k * total_rec + r];
    any = any || g[k] != 0.0f;
  }
  if (!any) return;   // no tile listed this Gaussian
  rec2_chain(jb, bjobs[ji], gid, g, g + 3, g + 6, g[9], g[10]);
}

}  // namespace fe

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
namespace fe {

template <int TS>
static hipError_t bwd_t(hipStream_t s, int64_t total_tiles, const BwdArgs& a) {
  hipLaunchKernelGGL((k_bwd_blend<TS>), dim3((unsigned)(total_tiles * a.kb)), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Blocks per tile of the backward blend.  A grid of few tiles leaves most of the GPU idle while
// each block walks its tile's list chunk by chunk, so up to 8 blocks share a tile until the grid
// has about 12 blocks per CU.  Lone views, N=4096: 128x128 (64 tiles, lists up to 3241
// records) 445 -> 79 us, 240x320 362 -> 78 us, 480x640 333 -> 197 us (2 blocks per tile); with
// about 4 blocks per CU the last two took 143 and 333 us.  Larger grids keep one block per tile:
// a second block repeats the tile's per-pixel set-up (the 16-view 480x640 shell: 1.38 -> 1.48 ms
// with two).
static int bwd_blocks_per_tile(int64_t total_tiles) {
  const int64_t target = (int64_t)12 * device_cus();
  if (total_tiles <= 0 || 2 * total_tiles > target) return 1;
  return (int)std::min<int64_t>(8, target / total_tiles);
}

hipError_t launch_bwd_blend(hipStream_t s, int ts, int64_t total_tiles, const Job* jobs, const void* bjobs,
                            int n_jobs, const float* rec, const float* rec2, const uint32_t* tile_np,
                            const unsigned long long* off1, const unsigned long long* off2, const uint32_t* list,
                            const float* st, int64_t total_pix, const int32_t* job_flags, float* gsum,
                            int64_t total_rec, int dense, const int32_t* tord, int feat_chunks) {
  const int64_t n4 = (total_rec * kGsumPlanes + 3) / 4;   // the region is padded to 256 B
  if (n4 > 0)
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<int64_t>((n4 + 255) / 256, 4096)), dim3(256), 0, s,
                       reinterpret_cast<float4*>(gsum), n4);
  BwdArgs a{jobs, reinterpret_cast<const BwdJob*>(bjobs), n_jobs, rec, rec2, tile_np, off1, off2, list, st,
            total_pix, job_flags, gsum, total_rec, dense, tord, bwd_blocks_per_tile(total_tiles)};
  hipError_t e = hipErrorInvalidValue;
  switch (ts) {
    case 8: e = bwd_t<8>(s, total_tiles, a); break;
    case 16: e = bwd_t<16>(s, total_tiles, a); break;
    case 32: e = bwd_t<32>(s, total_tiles, a); break;
  }
  if (e != hipSuccess) return e;
  if (feat_chunks > 0) {
    e = launch_feat(s, true, ts, total_tiles, feat_chunks, jobs, bjobs, n_jobs, rec, rec2, off1, off2, list, st,
                    total_pix, job_flags, gsum, total_rec, dense);
    if (e != hipSuccess) return e;
  }
  if (total_rec > 0)
    hipLaunchKernelGGL(k_bwd_chain, dim3((unsigned)((total_rec + 255) / 256)), dim3(256), 0, s, jobs,
                       reinterpret_cast<const BwdJob*>(bjobs), n_jobs, job_flags, gsum, total_rec);
  return hipGetLastError();
}

size_t bwd_job_bytes() { return sizeof(BwdJob); }
size_t bwd_gsum_floats() { return kGsumPlanes; }

}  // namespace fe
