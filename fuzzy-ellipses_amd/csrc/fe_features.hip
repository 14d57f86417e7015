// Depth-softmax-weighted per-Gaussian features, forward and backward (gfx950).
//
// render_func_rays returns, beside depth and alpha, the normalised weights w (N, P) of every
// (Gaussian, ray) pair "for color, DINO features, or any other per-Gaussian property"
// (baseline/fm_render.py:96-109).  Materialising (N, P) is impractical at scale; what callers
// do with it is contract it with a per-Gaussian array f (N, C).  Per pixel:
//   F = sum_j w_j f_j,   w_j = (s_j + 1e-20) / (S + 1e-20 n),   s_j = [t_j > 0] exp(l_j - max l),
// over the pixel's tile list (the same Gaussians, logits and 1e-20 floor as the depth; a tile
// where the floor may be visible lists every Gaussian).  k_feat_fwd recomputes each listed
// pair's logit from the forward's records with the forward's operations (online maximum,
// masking, NaN rules of views with degenerate records), one block per (tile, 4 channels), one
// thread per pixel, records and features staged in LDS.
//
// Backward, for a cotangent gF of the feature image: dF/df_j = w_j, and through the logits
//   dL/dl_j = s_j/(S + 1e-20 n) (gF.f_j - gF.F)  (+ for the pair holding max l:
//             1e-20 (gF.sum_j f_j - n gF.F)/(S + 1e-20 n), zero unless the floor is visible),
// with l = b3 d2 - b2 t, so dL/dd2 = b3 dL/dl and dL/dt = -b2 dL/dl.  k_feat_bwd (lane =
// Gaussian, loop over the tile's pixels) turns those into the same per-(Gaussian, view) sums as
// the backward blend (blend-record views, chained once per record by k_bwd_chain) or the
// direct-form chain rule, and adds w_j gF into d(features).  The forward's saved M and S and
// its feature image F are read back, so fe_backward must follow the matching fe_forward
// (save_for_backward) with the feature image untouched.
#include "fe_common.h"
#include "fe_grad.h"

#include <algorithm>

namespace fe {

constexpr int kFeatChunk = 4;    // channels per block
constexpr int kFeatBatch = 64;   // Gaussians per LDS batch

struct FeatArgs {
  const Job* jobs;
  const BwdJob* bjobs;
  int n_jobs;
  const float* rec;
  const float* rec2;
  const unsigned long long* off1;
  const unsigned long long* off2;
  const uint32_t* list;
  const float* st;
  int64_t total_pix;
  const int32_t* job_flags;
  float* gsum;
  int64_t total_rec;
  int dense;
};

// The tile's list: phase-1 entries [s1, s1 + n1), then phase-2 [s2, s2 + n2) (record ids of the
// view); in dense mode every record.
struct TileList {
  int64_t s1, n1, s2, n2;
  __device__ __forceinline__ int64_t total() const { return n1 + n2; }
};

__device__ __forceinline__ TileList tile_list(const FeatArgs& A, const Job& jb, int64_t gt) {
  TileList L;
  if (A.dense) {
    L.s1 = 0; L.n1 = jb.n; L.s2 = 0; L.n2 = 0;
  } else {
    L.s1 = (int64_t)A.off1[gt]; L.n1 = (int64_t)A.off1[gt + 1] - L.s1;
    L.s2 = (int64_t)A.off2[gt]; L.n2 = (int64_t)A.off2[gt + 1] - L.s2;
  }
  return L;
}

__device__ __forceinline__ int64_t list_record(const FeatArgs& A, const TileList& L, int64_t k) {
  if (A.dense) return k;
  return (k < L.n1) ? (int64_t)A.list[L.s1 + k] : (int64_t)A.list[L.s2 + (k - L.n1)];
}

__device__ __forceinline__ float4 feat_chunk(const Job& jb, int64_t g, int c0) {
  const float* f = jb.feat + g * (int64_t)jb.n_feat + c0;
  const int m = jb.n_feat - c0;
  return make_float4(f[0], m > 1 ? f[1] : 0.0f, m > 2 ? f[2] : 0.0f, m > 3 ? f[3] : 0.0f);
}

// Pixel coordinates of pixel i (tile-local, row-major over the valid width wv) of a tile.
__device__ __forceinline__ bool tile_pixel(const Job& jb, int tx, int ty, int ts, int i, int wv, int64_t* lp,
                                           float* xf, float* yf) {
  const int px = tx * ts + (i % wv), py = ty * ts + (i / wv);
  *lp = (int64_t)py * jb.W + px;
  const bool in = (px < jb.W) && (py < jb.H) && (*lp < jb.npix);
  if (jb.ray_xy) {
    *xf = in ? jb.ray_xy[2 * *lp] : 0.0f;
    *yf = in ? jb.ray_xy[2 * *lp + 1] : 0.0f;
  } else {
    *xf = (float)px - jb.cx;
    *yf = (float)py - jb.cy;
  }
  return in;
}

template <int TS>
__global__ __launch_bounds__(TS * TS) void k_feat_fwd(FeatArgs A) {
  constexpr int NT = TS * TS;
  __shared__ float4 lrec[kFeatBatch * 4];
  __shared__ float4 lf[kFeatBatch];
  const int64_t gt = blockIdx.x;
  const int ji = find_job(A.n_jobs, gt, [&](int k) { return A.jobs[k].tile_base; });
  const Job& jb = A.jobs[ji];
  const int C = jb.n_feat;
  const int c0 = (int)blockIdx.y * kFeatChunk;
  if (!jb.feat || !jb.feat_out || c0 >= C) return;   // block-uniform
  const int tile = (int)(gt - jb.tile_base);
  const int tx = tile % jb.tiles_x, ty = tile / jb.tiles_x;
  int64_t lp;
  float xf, yf;
  const bool inimg = tile_pixel(jb, tx, ty, TS, threadIdx.x, TS, &lp, &xf, &yf);
  const TileList L = tile_list(A, jb, gt);
  const int64_t ncnt = L.total();
  const int32_t jflags = A.job_flags[ji];
  const bool safe = (jflags & kFlagDegenerate) != 0;
  const bool r2 = use_rec2(jb, jflags);
  const float b3 = jb.beta3;
  const float nb2l = -jb.beta2 * kLog2e;
  const float kd = -0.5f * kLog2e;
  const float kap = r2 ? nb2l / b3 : 0.0f;
  const float4* rec4 = reinterpret_cast<const float4*>(r2 ? A.rec2 : A.rec) + jb.rec_base * 4;
  const int per = r2 ? 3 : 4;
  // pixel-grid tiles with blend records: the blend's forms, relative to the pixel's 32x32
  // block (fe_common.h, localize_rec2)
  const bool loc = r2 && !jb.ray_xy;
  const float ox = form_origin(tx * TS, jb.cx), oy = form_origin(ty * TS, jb.cy);
  if (loc) {
    xf = form_coord(tx * TS + (int)threadIdx.x % TS);
    yf = form_coord(ty * TS + (int)threadIdx.x / TS);
  }

  float M = -3.402823466e38f, Mt = -3.402823466e38f, S = 0.0f;
  float F[4] = {0.f, 0.f, 0.f, 0.f}, Fa[4] = {0.f, 0.f, 0.f, 0.f};
  bool nanf = false;
  for (int64_t base = 0; base < ncnt; base += kFeatBatch) {
    const int nb = (int)min<int64_t>(kFeatBatch, ncnt - base);
    if (loc) {
      for (int e = threadIdx.x; e < nb; e += NT) {
        const int64_t rid = list_record(A, L, base + e);
        float4 a0 = rec4[rid * 4], a1 = rec4[rid * 4 + 1], a2 = rec4[rid * 4 + 2];
        localize_rec2(a0, a1, a2, ox, oy);
        lrec[e * 4] = a0; lrec[e * 4 + 1] = a1; lrec[e * 4 + 2] = a2;
      }
    } else {
      for (int k = threadIdx.x; k < nb * per; k += NT) {
        const int e = k / per, part = k - e * per;
        lrec[e * 4 + part] = rec4[list_record(A, L, base + e) * 4 + part];
      }
    }
    for (int e = threadIdx.x; e < nb; e += NT) lf[e] = feat_chunk(jb, gauss_of(jb, list_record(A, L, base + e)), c0);
    __syncthreads();
    for (int e = 0; e < nb; ++e) {
      const float4 r0 = lrec[e * 4], r1 = lrec[e * 4 + 1], r2v = lrec[e * 4 + 2];
      float t, sv;
      if (r2) {
        // the forward blend's operations (fe_forward.hip, blend_batch_rec2)
        const float X = fmaf(r0.x, xf, fmaf(r0.y, yf, r0.z));   // (Y1, Y2 likewise)
        const float Y1 = fmaf(r0.w, xf, fmaf(r1.x, yf, r1.y));
        const float Y2 = fmaf(r1.z, xf, fmaf(r1.w, yf, r2v.x));
        const float W = fmaf(Y1, Y1, Y2 * Y2);
        const float rr = __builtin_amdgcn_rcpf(fmaf(X, X, W));
        t = X * rr;
        const float d2l = fmaf(r2v.z, W * rr, r2v.y);
        const float lt = fmaf(kap, t, d2l);
        if (lt > Mt) {
          const float Mn = b3 * lt;
          const float sc = __builtin_amdgcn_exp2f(M - Mn);
          S *= sc;
#pragma unroll
          for (int c = 0; c < 4; ++c) F[c] *= sc;
          M = Mn;
          Mt = lt;
        }
        sv = __builtin_amdgcn_exp2f(fmaf(b3, lt, -M));
      } else {
        // the direct form (fe_forward.hip, blend_batch)
        const float4 r3 = lrec[e * 4 + 3];
        const float u0 = fmaf(r0.x, xf, fmaf(r0.w, yf, r1.z));
        const float u1 = fmaf(r0.y, xf, fmaf(r1.x, yf, r1.w));
        const float u2 = fmaf(r0.z, xf, fmaf(r1.y, yf, r2v.x));
        const float vsv = fmaf(u0, u0, fmaf(u1, u1, u2 * u2));
        const float psv = fmaf(r2v.y, u0, fmaf(r2v.z, u1, r2v.w * u2));
        t = psv * __builtin_amdgcn_rcpf(vsv);
        const float v0 = fmaf(t, u0, -r2v.y), v1 = fmaf(t, u1, -r2v.z), v2 = fmaf(t, u2, -r2v.w);
        const float d0 = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
        const float d2l = fmaf(kd, d0, r3.y);
        const float l = fmaf(b3, d2l, nb2l * t);
        if (safe) nanf = nanf || (l != l);
        if (l > M) {
          const float sc = __builtin_amdgcn_exp2f(M - l);
          S *= sc;
#pragma unroll
          for (int c = 0; c < 4; ++c) F[c] *= sc;
          M = l;
        }
        sv = __builtin_amdgcn_exp2f(l - M);
        if (safe) sv = (sv == sv) ? sv : 0.0f;   // an infinite logit: exp(inf - inf) -> nan_to_num -> 0
      }
      sv = (t > 0.0f) ? sv : 0.0f;
      S += sv;
      const float4 fv = lf[e];
      F[0] = fmaf(sv, fv.x, F[0]); F[1] = fmaf(sv, fv.y, F[1]);
      F[2] = fmaf(sv, fv.z, F[2]); F[3] = fmaf(sv, fv.w, F[3]);
      Fa[0] += fv.x; Fa[1] += fv.y; Fa[2] += fv.z; Fa[3] += fv.w;
    }
    __syncthreads();
  }
  if (!inimg) return;
  const float nc = (float)ncnt;
  float* out = jb.feat_out + lp * (int64_t)C + c0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c0 + c >= C) break;
    float v;
    if (ncnt == 0) v = 0.0f;
    else if (nanf) v = Fa[c] / nc;   // a NaN logit leaves only the uniform 1e-20 floor
    else v = fmaf(1e-20f, Fa[c], F[c]) / fmaf(1e-20f, nc, S);
    out[c] = v;
  }
}

template <int TS>
__global__ __launch_bounds__(256) void k_feat_bwd(FeatArgs A) {
  constexpr int NPIX = TS * TS;
  __shared__ float4 pxa[NPIX];   // x', y', M, 1/(S + 1e-20 n)
  __shared__ float4 pxg[NPIX];   // gF (this chunk)
  __shared__ float2 pxb[NPIX];   // gF.F (this chunk), the maximum-logit term
  __shared__ float2 pxl[NPIX];   // x', y' relative to the pixel's 32x32 block (the blend's forms)
  __shared__ float4 red[4];
  const int64_t gt = blockIdx.x;
  const int ji = find_job(A.n_jobs, gt, [&](int k) { return A.jobs[k].tile_base; });
  const Job& jb = A.jobs[ji];
  const BwdJob& bj = A.bjobs[ji];
  const int C = jb.n_feat;
  const int c0 = (int)blockIdx.y * kFeatChunk;
  if (!jb.feat || !jb.feat_out || !bj.gF || c0 >= C) return;   // block-uniform
  const int tile = (int)(gt - jb.tile_base);
  const int tx = tile % jb.tiles_x, ty = tile / jb.tiles_x;
  const int wv = min(TS, jb.W - tx * TS);
  const int npx = tile_valid(jb, tile, tx, ty, TS);
  const TileList L = tile_list(A, jb, gt);
  const int64_t total = L.total();
  if (total == 0) return;
  const float nc = (float)total;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;

  // the tile's feature sum over its list (the 1e-20 floor's share of every pixel)
  float4 fs = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t k = threadIdx.x; k < total; k += 256) {
    const float4 f = feat_chunk(jb, gauss_of(jb, list_record(A, L, k)), c0);
    fs.x += f.x; fs.y += f.y; fs.z += f.z; fs.w += f.w;
  }
  for (int o = 32; o > 0; o >>= 1) {
    fs.x += __shfl_xor(fs.x, o, 64); fs.y += __shfl_xor(fs.y, o, 64);
    fs.z += __shfl_xor(fs.z, o, 64); fs.w += __shfl_xor(fs.w, o, 64);
  }
  if (lane == 0) red[wave] = fs;
  __syncthreads();
  const float FT[4] = {red[0].x + red[1].x + red[2].x + red[3].x, red[0].y + red[1].y + red[2].y + red[3].y,
                       red[0].z + red[1].z + red[2].z + red[3].z, red[0].w + red[1].w + red[2].w + red[3].w};

  for (int i = threadIdx.x; i < npx; i += 256) {
    int64_t lp;
    float xf, yf;
    tile_pixel(jb, tx, ty, TS, i, wv, &lp, &xf, &yf);
    const int64_t pix = jb.pix_base + lp;
    const float M = A.st[pix];
    const float S = A.st[A.total_pix + pix];
    const float den = fmaf(1e-20f, nc, S);
    float iS = (den > 0.0f) ? 1.0f / den : 0.0f;
    if (!(iS == iS) || !(fabsf(iS) <= 3.0e38f)) iS = 0.0f;
    const float* g = bj.gF + lp * (int64_t)C + c0;
    const float* Fo = jb.feat_out + lp * (int64_t)C + c0;
    float gv[4], GF = 0.0f, GT = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      gv[c] = (c0 + c < C) ? g[c] : 0.0f;
      const float Fv = (c0 + c < C) ? Fo[c] : 0.0f;
      GF = fmaf(gv[c], Fv, GF);
      GT = fmaf(gv[c], FT[c], GT);
    }
    float cX = 1e-20f * fmaf(-nc, GF, GT) * iS;   // gF . dF/d(max l) = 1e-20 (gF.sum f - n gF.F)/(S + 1e-20 n)
    if (!(cX == cX) || !(fabsf(cX) <= 3.0e38f)) cX = 0.0f;
    if (!(GF == GF)) GF = 0.0f;
    pxa[i] = make_float4(xf, yf, M, iS);
    pxg[i] = make_float4(gv[0], gv[1], gv[2], gv[3]);
    pxb[i] = make_float2(GF, cX);
    pxl[i] = make_float2(form_coord(tx * TS + i % wv), form_coord(ty * TS + i / wv));
  }
  __syncthreads();

  const int32_t jflags = A.job_flags[ji];
  const bool safe = (jflags & kFlagDegenerate) != 0;
  const bool r2 = use_rec2(jb, jflags);
  const float b3 = jb.beta3, b2 = jb.beta2;
  const float nb2l = -jb.beta2 * kLog2e;
  const float kd = -0.5f * kLog2e;
  const float kap = r2 ? nb2l / b3 : 0.0f;
  const float4* rec4 = reinterpret_cast<const float4*>(r2 ? A.rec2 : A.rec) + jb.rec_base * 4;
  for (int64_t cb = (int64_t)wave * 64; cb < total; cb += 256) {
    const int64_t kk = cb + lane;
    if (kk >= total) break;
    const int64_t rid = list_record(A, L, kk);
    const int64_t sg = gauss_of(jb, rid);
    float4 r0 = rec4[rid * 4], r1 = rec4[rid * 4 + 1], r2v = rec4[rid * 4 + 2];
    const float4 r3 = r2 ? make_float4(0.f, 0.f, 0.f, 0.f) : rec4[rid * 4 + 3];
    // pixel-grid tiles with blend records: the blend's forms (relative to the pixel's 32x32
    // block), so that s is the forward's weight bit for bit (the sums below stay in the
    // pixels' own coordinates)
    const bool loc = r2 && !jb.ray_xy;
    if (loc) localize_rec2(r0, r1, r2v, form_origin(tx * TS, jb.cx), form_origin(ty * TS, jb.cy));
    const float4 fj = feat_chunk(jb, sg, c0);
    float G[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) G[k] = 0.0f;
    float df[4] = {0.f, 0.f, 0.f, 0.f};
    for (int p = 0; p < npx; ++p) {
      const float4 pa = pxa[p];
      const float4 pg = pxg[p];
      const float2 pb = pxb[p];
      const float xf = pa.x, yf = pa.y, M = pa.z, iS = pa.w;
      float t, s, l2;   // l2: the logit as the forward compared it with M
      float X = 0.f, Y1 = 0.f, Y2 = 0.f, rr = 0.f, Wr = 0.f, iv = 0.f;
      float u[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f};
      if (r2) {
        const float2 pl = pxl[p];
        const float xe = loc ? pl.x : xf, ye = loc ? pl.y : yf;
        X = fmaf(r0.x, xe, fmaf(r0.y, ye, r0.z));
        Y1 = fmaf(r0.w, xe, fmaf(r1.x, ye, r1.y));
        Y2 = fmaf(r1.z, xe, fmaf(r1.w, ye, r2v.x));
        const float W = fmaf(Y1, Y1, Y2 * Y2);
        rr = __builtin_amdgcn_rcpf(fmaf(X, X, W));
        t = X * rr;
        Wr = W * rr;
        const float d2l = fmaf(r2v.z, Wr, r2v.y);
        const float lt = fmaf(kap, t, d2l);
        l2 = b3 * lt;
        s = __builtin_amdgcn_exp2f(fmaf(b3, lt, -M));
      } else {
        u[0] = fmaf(r0.x, xf, fmaf(r0.w, yf, r1.z));
        u[1] = fmaf(r0.y, xf, fmaf(r1.x, yf, r1.w));
        u[2] = fmaf(r0.z, xf, fmaf(r1.y, yf, r2v.x));
        const float vsv = fmaf(u[0], u[0], fmaf(u[1], u[1], u[2] * u[2]));
        const float psv = fmaf(r2v.y, u[0], fmaf(r2v.z, u[1], r2v.w * u[2]));
        iv = __builtin_amdgcn_rcpf(vsv);
        t = psv * iv;
        v[0] = fmaf(t, u[0], -r2v.y); v[1] = fmaf(t, u[1], -r2v.z); v[2] = fmaf(t, u[2], -r2v.w);
        const float d0 = fmaf(v[0], v[0], fmaf(v[1], v[1], v[2] * v[2]));
        const float d2l = fmaf(kd, d0, r3.y);
        l2 = fmaf(b3, d2l, nb2l * t);
        s = __builtin_amdgcn_exp2f(l2 - M);
      }
      s = (t > 0.0f) ? s : 0.0f;
      if (safe) s = (s == s) ? s : 0.0f;
      const float gfj = fmaf(pg.x, fj.x, fmaf(pg.y, fj.y, fmaf(pg.z, fj.z, pg.w * fj.w)));
      float gl = (s * iS) * (gfj - pb.x);
      if (l2 == M) gl += pb.y;   // the pair holding the maximum logit
      const float wf = (s + 1e-20f) * iS;
      df[0] = fmaf(wf, pg.x, df[0]); df[1] = fmaf(wf, pg.y, df[1]);
      df[2] = fmaf(wf, pg.z, df[2]); df[3] = fmaf(wf, pg.w, df[3]);
      if (safe && !(gl == gl)) gl = 0.0f;
      const float gd2 = b3 * gl, g1 = -b2 * gl;
      if (r2) rec2_pair_sums(gd2, g1, X, Y1, Y2, rr, t, Wr, r2v.w, xf, yf, G);
      else direct_pair_sums(gd2, g1, t, iv, u, v, xf, yf, G);
    }
    if (r2) {
      float* gs = A.gsum + jb.rec_base + rid;
#pragma unroll
      for (int k = 0; k < kGsumPlanes; ++k) atomicAdd(gs + (int64_t)k * A.total_rec, G[k]);
    } else {
      direct_chain(jb, bj, sg, G[0], G[1], G[2], G[3], G[4], G[5], G[6], G[7], G[8], G[9], G[10], G[11], G[12]);
    }
    if (bj.dfeat) {
      float* d = bj.dfeat + sg * (int64_t)C + c0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c0 + c < C) atomicAdd(d + c, df[c]);
    }
  }
}

template <int TS>
static hipError_t feat_fwd_t(hipStream_t s, int64_t tiles, int chunks, const FeatArgs& a) {
  hipLaunchKernelGGL((k_feat_fwd<TS>), dim3((unsigned)tiles, (unsigned)chunks), dim3(TS * TS), 0, s, a);
  return hipGetLastError();
}

template <int TS>
static hipError_t feat_bwd_t(hipStream_t s, int64_t tiles, int chunks, const FeatArgs& a) {
  hipLaunchKernelGGL((k_feat_bwd<TS>), dim3((unsigned)tiles, (unsigned)chunks), dim3(256), 0, s, a);
  return hipGetLastError();
}

int feat_chunks(int n_feat) { return (n_feat + kFeatChunk - 1) / kFeatChunk; }

hipError_t launch_feat(hipStream_t s, bool backward, int ts, int64_t total_tiles, int chunks, const Job* jobs,
                       const void* bjobs, int n_jobs, const float* rec, const float* rec2,
                       const unsigned long long* off1, const unsigned long long* off2, const uint32_t* list,
                       const float* st, int64_t total_pix, const int32_t* job_flags, float* gsum, int64_t total_rec,
                       int dense) {
  if (total_tiles <= 0 || chunks <= 0) return hipSuccess;
  const FeatArgs a{jobs, reinterpret_cast<const BwdJob*>(bjobs), n_jobs, rec, rec2, off1, off2, list, st, total_pix,
                   job_flags, gsum, total_rec, dense};
  switch (ts) {
    case 8: return backward ? feat_bwd_t<8>(s, total_tiles, chunks, a) : feat_fwd_t<8>(s, total_tiles, chunks, a);
    case 16: return backward ? feat_bwd_t<16>(s, total_tiles, chunks, a) : feat_fwd_t<16>(s, total_tiles, chunks, a);
    case 32: return backward ? feat_bwd_t<32>(s, total_tiles, chunks, a) : feat_fwd_t<32>(s, total_tiles, chunks, a);
  }
  return hipErrorInvalidValue;
}

}  // namespace fe
