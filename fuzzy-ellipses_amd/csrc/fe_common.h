// Shared device-side definitions of the fuzzy-metaball rasterizer (gfx950).
//
// Per (Gaussian, view) record, 16 floats (64 B, one cache-line quarter, four float4):
//   [0..2]  a = M[:,0] / fx      M = A R^T, A = triu(prec)^T (lower), R rows = camera axes
//   [3..5]  b = M[:,1] / fy
//   [6..8]  g = M[:,2]
//   [9..11] q = A (mu - o)
//   [12]    w        log-weight
//   [13]    w*log2(e)
//   [14]    q.q
//   [15]    flags (bit 0: degenerate record -- non-finite input or a zero diagonal of U)
// With these, the whitened ray of pixel (x, y) is u = a*(x-cx) + b*(y-cy) + g, which equals
// triu(prec)^T (r_cam @ R) of baseline/fm_render.py:33-52, and per pair
//   t* = (q.u)/(u.u),  d0 = |t* u - q|^2,  d2 = w - d0/2   (fm_render.py:53-70).
//
// Blend record, 16 floats per (Gaussian, view), computed in float64 and rounded once:
//   [0..2]  kX  = (a.q, b.q, g.q) / |q|^2
//   [3..5]  kY1 = (a.e1, b.e1, g.e1) / |q|
//   [6..8]  kY2 = (a.e2, b.e2, g.e2) / |q|
//   [9]     w*log2(e)
//   [10]    -|q|^2 log2(e) / 2   (stored negated: d2 log2(e) = [9] + [10] W/den is one fma)
//   [11]    |q|^2
//   [12..15] unused
// where (q/|q|, e1, e2) is an orthonormal basis.  For p = (x-cx, y-cy, 1) the linear forms
// X = kX.p and Y_i = kY_i.p are the components of the whitened ray u along q/|q|, e1, e2,
// divided by |q|.  With W = Y1^2 + Y2^2 and den = X^2 + W:
//   t* = (q.u)/(u.u) = X/den,   d0 = |q x u|^2/(u.u) = |q|^2 W/den,
// so every term of t* and d0 is a sum of squares (no cancellation near the centre ray),
// at 3 linear forms per pair instead of u, u.u, q.u and |t* u - q|^2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fe {

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;
constexpr int kRecFloats = 16;
constexpr int kRec2Floats = 16;
// A blend record (kX = a0.xyz, kY1 = (a0.w, a1.x, a1.y), kY2 = (a1.z, a1.w, a2.x), the
// last of each the constant term) with the constant terms moved to the point (ox, oy) of the
// camera plane: the forms of p = (ox + x', oy + y', 1) become k.(x', y', 1).  Pixel-grid views
// evaluate their forms this way in every pass (blend, backward, feature kernels), relative to
// the centre of the 32x32-pixel block holding the pixel, with these operations: the passes
// agree bit for bit, and so do the tile sizes (8, 16 and 32-pixel tiles lie inside one block).
constexpr int kFormBlock = 32;
constexpr float kFormHalf = 15.5f;   // (kFormBlock - 1) / 2
// first pixel of the block holding pixel column (or row) p >= 0
__device__ __forceinline__ int form_block(int p) { return p / kFormBlock * kFormBlock; }
// the pixel's coordinate relative to its block's centre (exact)
__device__ __forceinline__ float form_coord(int p) { return (float)(p - form_block(p)) - kFormHalf; }
// the block centre's camera-plane coordinate (principal point c)
__device__ __forceinline__ float form_origin(int p, float c) { return ((float)form_block(p) + kFormHalf) - c; }
__device__ __forceinline__ void localize_rec2(float4& a0, float4& a1, float4& a2, float ox, float oy) {
  a0.z = fmaf(a0.x, ox, fmaf(a0.y, oy, a0.z));
  a1.y = fmaf(a0.w, ox, fmaf(a1.x, oy, a1.y));
  a2.x = fmaf(a1.z, ox, fmaf(a1.w, oy, a2.x));
}
// Smallest |q|^2 the blend record form accepts (smaller: the view uses the direct form).
constexpr double kMinQQ = 1e-20;

// Job flag bits (per view, set by the preprocess)
constexpr int32_t kFlagDegenerate = 1;   // a non-finite input or a zero diagonal of U
constexpr int32_t kFlagNoBlendRec = 2;   // some Gaussian has no valid blend record

// Orthonormal vectors e1, e2 completing the unit vector n to a basis (Duff et al., "Building
// an Orthonormal Basis, Revisited", JCGT 6(1), 2017).  Only W = Y1^2 + Y2^2 is used, which
// does not depend on the choice of e1, e2 within the plane.
template <typename T>
__device__ __forceinline__ void ortho_basis(const T n[3], T e1[3], T e2[3]) {
  const T s = n[2] >= T(0) ? T(1) : T(-1);
  const T a = T(-1) / (s + n[2]);
  const T b = n[0] * n[1] * a;
  e1[0] = T(1) + s * n[0] * n[0] * a;
  e1[1] = s * b;
  e1[2] = -s * n[0];
  e2[0] = b;
  e2[1] = s + n[1] * n[1] * a;
  e2[2] = -n[1];
}

// One view of one scene, as the kernels see it.  Built on the host by the planner.
struct Job {
  const float* means;
  const float* prec;
  const float* wlog;
  float* depth;
  float* alpha;
  const float* pose;  // optional device pose R[9] t[3], copied into R / o by k_pose
  float R[9];
  float o[3];
  float fx, fy, cx, cy;
  float beta2, beta3;
  float c_alpha;   // alpha cull constant, natural-log units: drop if d2 < -c_alpha on a tile
  float delta2;    // depth cull margin, log2 units
  float floor_thr; // log2(1e13 / n): a dropped logit this far above the tile minimum makes the
                   // 1e-20 weight floor visible, so the tile is completed densely
  int32_t n, H, W;
  int32_t tiles_x, tiles_y, n_tiles;
  int32_t nw;      // 64-bit mask words per tile = ceil(n / 64)
  int32_t sf;      // super-tile factor: membership is pre-tested on sf x sf groups of tiles
  int32_t stiles_x, stiles_y;
  int64_t stile_base; // first global super-tile
  int32_t scene;
  int64_t rec_base;   // first record (Gaussian-view) of this job
  int64_t tile_base;  // first global tile
  int64_t mask_base;  // first mask word
  int64_t pix_base;   // first pixel of this view in the per-pixel state arrays
  // Ray-list views (ray_xy != nullptr): pixel k of the view is ray k, with camera-plane
  // coordinates (ray_xy[2k], ray_xy[2k+1]) in place of (x - cx, y - cy); the view is laid
  // out as an image TS pixels wide, so tile t holds rays [t*TS*TS, (t+1)*TS*TS), and each
  // tile's bounding rectangle (centre, half-extent) is rect[4*t .. 4*t+3].
  const float* ray_xy;
  float* rect;
  int64_t npix;       // valid pixels: H*W, or the number of rays
  // Record order: record j of this view (and every list id) is Gaussian order[j].  Written
  // by k_order each forward: the Gaussians sorted by a Z-order code of their projected
  // centres, in chunks of 1024, so a tile's list is a few runs of nearby ids.
  int32_t* order;
  int32_t* rank;      // inverse of order: Gaussian g is record rank[g]
  // Optional per-Gaussian features (scene order, (n, n_feat) row-major) and this view's
  // feature image (npix, n_feat): the pixel's depth-softmax-weighted feature sum.
  const float* feat;
  float* feat_out;
  int32_t n_feat;
};

// Tile of a block of a one-block-per-tile launch of n blocks.  The hardware deals blocks
// round-robin to the 8 XCDs (block b to XCD b % 8), each with its own L2; this gives XCD x
// one contiguous run of tiles instead of every 8th, so neighbouring tiles (which share the
// cache lines of their pixel rows and most of their records) run on the same L2.  A
// bijection of [0, n) whatever the dispatch does.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  const int64_t per = n >> 3, rem = n & 7;
  const int64_t x = b & 7, k = b >> 3;
  return x * per + (x < rem ? x : rem) + k;
}

// Dispatch order of the list-driven blends (the culled phase-1 blend and the backward).
// Blocks are dispatched in index order and a tile's time grows with its list length, so in
// spatial order the longest tiles can start in the last wave of blocks and leave the chip
// idle behind them.  The scan of the tile counts also sorts the tiles by a bucket of their
// count, longest first (longest-processing-time-first scheduling), and the blends map block
// b to tile tord[b].  The bucket is a float-like code of the count with 3 mantissa bits
// (monotone, within 12.5%), so the order is a counting sort over kTileBuckets buckets.
constexpr int kTileBuckets = 256;
__device__ __forceinline__ int tile_bucket(uint32_t c) {
  if (c < 8u) return (int)c;
  const int e = 31 - __clz(c);
  return 8 * (e - 2) + (int)((c >> (e - 3)) & 7u);   // at most 8 * 29 + 7 = 239
}

// Scene index of record j of a view.
__device__ __forceinline__ int64_t gauss_of(const Job& jb, int64_t j) { return jb.order ? jb.order[j] : j; }

// Per-view backward buffers: upstream gradients of the images and the scene's gradients.
struct BwdJob {
  const float* gD;
  const float* gA;
  float* dmeans;
  float* dprec;
  float* dw;
  const float* gF;     // cotangent of the feature image (npix, n_feat), or null
  float* dfeat;        // d(loss)/d(features) (n, n_feat) of the scene, or null
};

// Blend-record backward: sums per (Gaussian, view) record, one plane each (see fe_backward.hip).
constexpr int kGsumPlanes = 11;

// Device counters (zeroed by k_init each forward).  The per-block evaluation counts go to
// one of kCountSlots slots, each on its own 128-B line, picked by block index: a single
// counter would take one same-address atomic per blend block (19k per 16-view launch), and
// those serialise at memory (they bounded the phase-2 blend).  Readers add the slots.
constexpr int kCountSlots = 64;
// Membership of a small grid splits each group's Gaussians over several blocks (at most
// kMaskChunkBlocks blocks, for at most kMaskChunkGroups groups): each block leaves its per-tile
// partial count, flag, bound and threshold (4 words) in a workspace region, and a second kernel
// merges them and finishes the group's tiles; after them, one word per (block, tile) holds a
// phase-2 lower bound for tiles that phase 1 left without a threshold.
constexpr int kMaskChunkBlocks = 1024;
// The blend splits the tile lists of views of fewer 16-pixel tiles than this (8-pixel tiles:
// every view; 32-pixel tiles: none): up to 1280x720 and 960x1024.
constexpr int kSplitMinTiles = 4096;
__host__ __device__ inline bool view_splits_lists(int ts, int64_t n_tiles) {
  return ts == 8 || (ts == 16 && n_tiles < kSplitMinTiles);
}
constexpr int kMaskChunkGroups = kMaskChunkBlocks / 2;
constexpr int kMaskMaxSub = 16;   // tiles per group at most (4x4 8-pixel tiles)
struct Counters {
  unsigned long long floor_tiles;       // tiles completed densely because the floor could matter
  unsigned long long floor_unresolved;  // pixels' tiles where it still could (must stay 0)
  unsigned long long pad[14];
  unsigned long long slot[kCountSlots][16];   // [i][0]: evaluated pairs, [i][1]: of them phase 1
  uint32_t tile_hist[2][kTileBuckets];   // dispatch buckets: tiles per bucket, then bucket starts
};

// Whether the membership bounds apply to a view: they bound the logit b3 d2 - b2 t from above
// by the least distance and the least depth over a tile, which needs finite b2 >= 0 and
// b3 >= 0.  Other views list every pair (like views with degenerate records).
__device__ __forceinline__ bool bounds_apply(const Job& jb) {
  return jb.beta3 >= 0.0f && jb.beta3 <= 3.0e38f && jb.beta2 >= 0.0f && jb.beta2 <= 3.0e38f;
}

// Whether a view's blends use the blend records: no degenerate record, every record valid,
// and b3 > 0 (the running maximum is taken over l/b3).
__device__ __forceinline__ bool use_rec2(const Job& jb, int32_t flags) {
  return !(flags & (kFlagDegenerate | kFlagNoBlendRec)) && jb.beta3 > 0.0f && jb.beta3 < 3.0e38f &&
         (jb.beta2 == jb.beta2) && fabsf(jb.beta2) < 3.0e38f;
}

__device__ __forceinline__ float dot3(const float* x, const float* y) {
  return fmaf(x[0], y[0], fmaf(x[1], y[1], x[2] * y[2]));
}

// c = x cross y
__device__ __forceinline__ void cross3(const float* x, const float* y, float* c) {
  c[0] = fmaf(x[1], y[2], -(x[2] * y[1]));
  c[1] = fmaf(x[2], y[0], -(x[0] * y[2]));
  c[2] = fmaf(x[0], y[1], -(x[1] * y[0]));
}

// Binary search: the job whose [base, next base) range holds idx.  `bases` are read through
// the Job array (tile_base or rec_base selected by the accessor).
template <typename F>
__device__ __forceinline__ int find_job(int n_jobs, int64_t idx, F base_of) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (base_of(mid) <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Tile geometry in centred pixel coordinates: centre (Xc, Yc), half extents (hx, hy) over
// the tile's pixel centres, clipped to the image.
struct TileGeo {
  float Xc, Yc, hx, hy;
};

__device__ __forceinline__ TileGeo tile_geo_xy(const Job& jb, int tx, int ty, int ts) {
  if (jb.ray_xy) {   // ray lists use sf = 1, so (tx, ty) is always a tile
    const float4 r = reinterpret_cast<const float4*>(jb.rect)[(int64_t)ty * jb.tiles_x + tx];
    TileGeo g;
    g.Xc = r.x; g.Yc = r.y; g.hx = r.z; g.hy = r.w;
    return g;
  }
  int x0 = tx * ts, y0 = ty * ts;
  int xl = min(x0 + ts, jb.W) - 1, yl = min(y0 + ts, jb.H) - 1;
  TileGeo g;
  g.Xc = (float)(x0 + xl) * 0.5f - jb.cx;
  g.Yc = (float)(y0 + yl) * 0.5f - jb.cy;
  g.hx = (float)(xl - x0) * 0.5f;
  g.hy = (float)(yl - y0) * 0.5f;
  return g;
}

// Valid pixels (or rays) of tile (tx, ty).
__device__ __forceinline__ int tile_valid(const Job& jb, int tile, int tx, int ty, int TS) {
  if (jb.ray_xy) {
    const int64_t left = jb.npix - (int64_t)tile * TS * TS;
    return (int)(left < TS * TS ? left : TS * TS);
  }
  return min(TS, jb.W - tx * TS) * min(TS, jb.H - ty * TS);
}

// Bounding rectangle of a ray-list tile, in the units of the camera plane, widened by
// 2^-20 of the largest coordinate so that the rounded centre +- half-extent still covers
// every ray.
__device__ __forceinline__ float4 ray_rect(float xmin, float xmax, float ymin, float ymax) {
  const float mx = fmaxf(fabsf(xmin), fabsf(xmax)) * 9.5367431640625e-07f;
  const float my = fmaxf(fabsf(ymin), fabsf(ymax)) * 9.5367431640625e-07f;
  return make_float4(0.5f * (xmin + xmax), 0.5f * (ymin + ymax), 0.5f * (xmax - xmin) + mx,
                     0.5f * (ymax - ymin) + my);
}

// Conservative bounds of one Gaussian over one tile's pixel rays.
//   sd_lo, sd_hi : bounds of sqrt(d0) (distance of q to the ray line in whitened space)
//   t_lb         : lower bound of t* where the whole tile lies on the q side, else 0
//                  (pair_t_lb)
//   plus         : true when every ray of the tile has q.u > 0 (t* > 0 on the tile)
//   t_any_lb     : lower bound of t* on the tile regardless of side, -inf if unknown
//                  (inside masked_logit_ub)
//   t_ub         : upper bound of t* on the tile, +inf if unknown (pair_t_ub)
//   unbounded    : the tile cone is too wide for the bound (keep the pair)
// sqrt(d0) = |q| sin(angle(q, line)) is |q|-Lipschitz in the ray angle; the angle from the
// centre ray to any tile ray is at most the largest corner angle, bounded by its tangent.
//   front        : like plus, with a 2^-10 relative margin on sd_hi < |q| (used only to let
//                  the blend skip the t* > 0 test on a tile, never for membership)
//   can_mask, d0m_lo, negt_hi (inside masked_logit_ub) : where a tile ray can have t* <= 0,
//                  the angle theta between it
//                  and q is at least 90 deg; with theta_c the centre ray's angle and phi the
//                  tile's angular radius, theta <= theta_c + phi there, so
//                  d0 = |q|^2 sin^2(theta) >= d0m_lo and -t* <= negt_hi on the masked rays
//                  (can_mask false: no tile ray has t* <= 0).  This bounds the logit of a
//                  masked pair far below the one from (sd_lo, t_any_lb), whose extremes
//                  cannot occur at the same ray.
struct PairBound {
  float sd_lo, sd_hi;           // bounds of sqrt(d0) on the tile
  float qn, qu, sdc, tanm, dth; // |q|, q.u_c, sqrt(d0) at the centre ray, tan of the half-angle
  float uu, ap2, qq;            // |u_c|^2, the largest |a hx +- b hy|^2, |q|^2
  bool plus, unbounded, front;
};

// Core of the bound of a (tile, Gaussian) pair.  The parts that only some tests need (t_lb,
// t_ub, the masked-logit bound) are computed on demand from the stored intermediates, with
// the same operations as an eager evaluation, so every value is the same either way.
__device__ __forceinline__ PairBound pair_bound(const float* r, const TileGeo& tg) {
  const float* a = r;
  const float* b = r + 3;
  const float* g = r + 6;
  const float* q = r + 9;
  const float qq = r[14];
  float uc[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) uc[i] = fmaf(a[i], tg.Xc, fmaf(b[i], tg.Yc, g[i]));
  const float uu = dot3(uc, uc);
  const float qu = dot3(q, uc);
  const float tc = qu / uu;
  float v[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = fmaf(tc, uc[i], -q[i]);
  const float sdc = sqrtf(dot3(v, v));
  const float ua = dot3(uc, a);
  const float ub = dot3(uc, b);
  float ca[3], cb[3], pp[3], pm[3], ap[3], am[3];
  cross3(uc, a, ca);
  cross3(uc, b, cb);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pp[i] = fmaf(ca[i], tg.hx, cb[i] * tg.hy);
    pm[i] = fmaf(ca[i], tg.hx, -(cb[i] * tg.hy));
    ap[i] = fmaf(a[i], tg.hx, b[i] * tg.hy);
    am[i] = fmaf(a[i], tg.hx, -(b[i] * tg.hy));
  }
  // larger of |pp|, |pm| with one square root (sqrt is monotone and correctly rounded, so
  // this equals the larger of the two roots)
  const float nmax = sqrtf(fmaxf(dot3(pp, pp), dot3(pm, pm)));
  const float den = fmaf(-fabsf(ub), tg.hy, fmaf(-fabsf(ua), tg.hx, uu));
  PairBound pb;
  pb.unbounded = !(den > 0.0f);
  const float tanm = nmax / den;
  const float qn = sqrtf(qq);
  const float dth = qn * tanm;
  pb.sd_lo = fmaxf(sdc - dth, 0.0f);
  pb.sd_hi = sdc + dth;
  pb.qn = qn; pb.qu = qu; pb.sdc = sdc; pb.tanm = tanm; pb.dth = dth;
  pb.uu = uu; pb.ap2 = fmaxf(dot3(ap, ap), dot3(am, am)); pb.qq = qq;
  pb.plus = (qu > 0.0f) && (pb.sd_hi < qn);
  pb.front = (qu > 0.0f) && (pb.sd_hi * 1.0009765625f < qn);
  if (!(pb.sd_lo == pb.sd_lo)) pb.unbounded = true;  // NaN anywhere -> keep
  pb.front = pb.front && !pb.unbounded;
  return pb;
}

// Lower bound of t* on the tile for a pair on the q side everywhere (plus), else 0:
// |q.u|/|u| >= sqrt(qq - sd_hi^2), |u| <= |u_c| + rho.
__device__ __forceinline__ float pair_t_lb(const PairBound& pb) {
  return pb.plus ? sqrtf(fmaxf(fmaf(-pb.sd_hi, pb.sd_hi, pb.qq), 0.0f)) / (sqrtf(pb.uu) + sqrtf(pb.ap2)) : 0.0f;
}

// Upper bound of t* on the tile: q.u/|u| <= min(|q|, q.u_c/|u_c| + |q| tan(angle)),
// |u| >= |u_c| - rho.
__device__ __forceinline__ float pair_t_ub(const PairBound& pb) {
  const float un = sqrtf(pb.uu), rho = sqrtf(pb.ap2);
  return (un > rho) ? fminf(pb.qn, pb.qu / un + pb.dth) / (un - rho) : INFINITY;
}

// Upper bound (log2 units) of the depth logit of a pair on the tile rays where it is masked
// (t* <= 0), or -inf where it cannot be masked: the smaller of two bounds.  The first is
// the unmasked one with |q.u|/|u| <= sqrt(qq - d0_lo) on the tile; the second holds on
// the masked part of the tile cone (angle to q in [90 deg, theta_c + phi]), from the
// cos/sin of theta_c and of the half-angle phi: sin^2 >= sin^2(theta_c + phi) (0 past 180
// deg) and -cos <= min(-cos(theta_c + phi), 1).
__device__ __forceinline__ float masked_logit_ub(const PairBound& pb, float w, float b3l, float b2l) {
  const float un = sqrtf(pb.uu), rho = sqrtf(pb.ap2);
  const float t_any_lb = (un > rho) ? -(sqrtf(fmaxf(fmaf(-pb.sd_lo, pb.sd_lo, pb.qq), 0.0f)) / (un - rho)) : -INFINITY;
  const float cth = pb.qu / (pb.qn * un);
  const float sth = pb.sdc / pb.qn;
  const float cph = 1.0f / sqrtf(fmaf(pb.tanm, pb.tanm, 1.0f));
  const float sph = pb.tanm * cph;
  const float cmax = fmaf(cth, cph, -(sth * sph));
  const float smax = fmaf(sth, cph, cth * sph);
  const bool can_mask = !(cmax > 0.0f);
  const float d0m_lo = (smax > 0.0f) ? pb.qq * (smax * smax) : 0.0f;
  const float negt_hi = (un > rho) ? pb.qn * ((smax > 0.0f) ? fminf(-cmax, 1.0f) : 1.0f) / (un - rho) : INFINITY;
  const float dlo = fmaf(-0.5f * pb.sd_lo, pb.sd_lo, w);
  const float u1 = fmaf(b3l, dlo, -(b2l * t_any_lb));
  const float u2 = can_mask ? fmaf(b3l, fmaf(-0.5f, d0m_lo, w), b2l * negt_hi) : -INFINITY;
  return fminf(u1, u2);
}

// Phase-2 keep test of a bounded pair: the depth-logit bound b3l (w - sd_lo^2/2) - b2l t_lb
// reaches tau.  With b2l >= 0 and t_lb >= 0 the fma never exceeds the rounded b3l dlo, so
// a pair whose b3l dlo is below tau fails the full test too and t_lb is not needed.
__device__ __forceinline__ bool depth_keep(const PairBound& pb, float w, float b3l, float b2l, float tau) {
  const float dlo = fmaf(-0.5f * pb.sd_lo, pb.sd_lo, w);
  if (b2l >= 0.0f && !(b3l * dlo >= tau)) return false;
  return fmaf(b3l, dlo, -(b2l * pair_t_lb(pb))) >= tau;
}

}  // namespace fe
