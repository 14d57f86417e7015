// Forward kernels of the fuzzy-metaball rasterizer for MI355X (gfx950, wave64).
//
// Pipeline per batch of views (one launch of each kernel covers every view of the batch):
//   k_init        zero counters / per-job flags (a kernel, not hipMemset: graph-safe)
//   k_preprocess  (a) per (Gaussian, view): whitened camera-frame record (fe_common.h)
//   k_tile_mask   (b) per (tile, Gaussian): conservative bound test, wave ballot -> 64-bit
//                     membership words + per-tile histogram (count)
//   k_scan_*          exclusive prefix sum of the tile histogram -> tile ranges
//   k_tile_write      in-order compaction of the membership words -> per-tile Gaussian lists
//                     (tile-major emission makes the list sorted by (tile, gaussian id):
//                     this is the counting sort on the tile digit)
//   k_blend       (c) per tile: LDS-staged batches of records, per pixel ray the max-density
//                     point t*, density, alpha sum and the online-max depth softmax
// Culled mode runs (b)+(c) twice: phase 1 lists pairs whose density can reach the alpha
// threshold; the phase-1 blend yields, per tile, the exact minimum over its pixels of the
// log-sum-exp of the listed depth logits; phase 2 adds every pair whose logit upper bound is
// within the depth margin of it and finishes the pixels.  Dense mode evaluates all pairs.
#include <atomic>
#include <type_traits>

#include "fe_common.h"

namespace fe {

// ------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float nan_to_num(float t) {
  // jnp.nan_to_num: NaN -> 0, +-inf -> +-FLT_MAX
  return (t == t) ? fminf(fmaxf(t, -3.402823466e38f), 3.402823466e38f) : 0.0f;
}

// block-wide min (MAX = false) or max (MAX = true); every thread gets the result
template <int NT, bool MAX = false>
__device__ __forceinline__ float block_ext(float v, float* sh) {
  auto op = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (NT > 64) {
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < NT / 64; ++w) v = op(v, sh[w]);
      sh[0] = v;
    }
    __syncthreads();
    v = sh[0];
    __syncthreads();   // sh is reused by the next reduction
  }
  return v;
}

// ------------------------------------------------------------------------------------------
// k_init
// ------------------------------------------------------------------------------------------
__global__ void k_init(Counters* cnt, int32_t* job_flags, int n_jobs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < (int)(sizeof(Counters) / 8); k += gridDim.x * blockDim.x)
    reinterpret_cast<unsigned long long*>(cnt)[k] = 0ull;
  for (int k = i; k < n_jobs; k += gridDim.x * blockDim.x) job_flags[k] = 0;
}

// ------------------------------------------------------------------------------------------
// k_pose: jobs whose pose lives in device memory get R and o copied into their descriptor
// (one thread per job; runs before every kernel that reads the pose)
// ------------------------------------------------------------------------------------------
__global__ void k_pose(Job* jobs, int n_jobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs || jobs[j].pose == nullptr) return;
  const float* p = jobs[j].pose;
  for (int i = 0; i < 9; ++i) jobs[j].R[i] = p[i];
  for (int i = 0; i < 3; ++i) jobs[j].o[i] = p[9 + i];
}

// ------------------------------------------------------------------------------------------
// k_order: record order of each view.  The Gaussians of a view are sorted, in chunks of
// kOrderChunk, by a 20-bit Z-order code of their projected centre (10 bits per axis over
// twice the image extent around it; centres behind the camera last), so the Gaussians of a
// tile's list have nearby ids: the backward's per-record atomics of one wave then fall on
// few cache lines, and the blends' record loads on few pages.  The order changes nothing but
// the order of summation.  One 256-thread block per (chunk, view); bitonic sort of
// (code << 12 | index) keys, unique within a chunk, in registers, lanes and LDS.  (Lists are
// ascending in record id, so each chunk's share of a tile's list is a few runs: sorting
// within chunks of 1024 keeps the backward as fast as one sort of all 4096 and costs a third
// of the time.)
// ------------------------------------------------------------------------------------------
constexpr int kOrderChunk = 1024;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  // bits 0..9 of v to even bit positions 0..18
  v &= 0x3FFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// One compare-exchange of the bitonic network for element e (value v, partner value pv):
// the lower element of the pair keeps the minimum in ascending runs, the maximum otherwise.
__device__ __forceinline__ uint32_t bitonic_keep(uint32_t v, uint32_t pv, int e, int len, int stride) {
  const bool up = (e & len) == 0, lower = (e & stride) == 0;
  return (lower == up) ? min(v, pv) : max(v, pv);
}

constexpr int kOrderThreads = kOrderChunk / 4;

// It also clears the forward's counters and job flags (block (0, 0) the counters, block
// (0, j) job j's flag), which k_init does when no job has Gaussians: one launch fewer.
__global__ __launch_bounds__(kOrderThreads) void k_order(const Job* __restrict__ jobs, Counters* __restrict__ cnt,
                                                         int32_t* __restrict__ job_flags) {
  __shared__ uint32_t key[kOrderChunk];
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) job_flags[blockIdx.y] = 0;
    if (blockIdx.y == 0)
      for (int k = threadIdx.x; k < (int)(sizeof(Counters) / 8); k += kOrderThreads)
        reinterpret_cast<unsigned long long*>(cnt)[k] = 0ull;
  }
  const Job& jb = jobs[blockIdx.y];
  const int c0 = blockIdx.x * kOrderChunk;
  if (c0 >= jb.n || !jb.order) return;   // block-uniform
  const int m = min(kOrderChunk, jb.n - c0);
  // image-plane window of the code: pixels for grid views, camera-plane units for ray lists
  const bool grid = !jb.ray_xy;
  const float sx = grid ? jb.fx / (2.0f * (float)jb.W) : 0.25f, sy = grid ? jb.fy / (2.0f * (float)jb.H) : 0.25f;
  const float ox = grid ? (jb.cx + 0.5f * (float)jb.W) / (2.0f * (float)jb.W) : 0.5f;
  const float oy = grid ? (jb.cy + 0.5f * (float)jb.H) / (2.0f * (float)jb.H) : 0.5f;
  // thread t holds elements 4t .. 4t+3, so wave w holds elements [256 w, 256 w + 256)
  const int t = threadIdx.x;
  uint32_t v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * t + r;
    uint32_t k = 0xFFFFFFFFu;
    if (i < m) {
      const float* mu = jb.means + 3 * (int64_t)(c0 + i);
      const float p0 = mu[0] - jb.o[0], p1 = mu[1] - jb.o[1], p2 = mu[2] - jb.o[2];
      const float x = fmaf(jb.R[0], p0, fmaf(jb.R[1], p1, jb.R[2] * p2));
      const float y = fmaf(jb.R[3], p0, fmaf(jb.R[4], p1, jb.R[5] * p2));
      const float z = fmaf(jb.R[6], p0, fmaf(jb.R[7], p1, jb.R[8] * p2));
      uint32_t code = 0xFFFFFu;
      if (z > 0.0f) {
        // fraction of the window [-W/2, 3W/2] of pixel x (likewise y), clamped; NaN -> 0
        const float u = fminf(fmaxf(fmaf(sx, x / z, ox), 0.0f), 1.0f);
        const float w = fminf(fmaxf(fmaf(sy, y / z, oy), 0.0f), 1.0f);
        code = spread10((uint32_t)(u * 1023.0f)) | (spread10((uint32_t)(w * 1023.0f)) << 1);
      }
      k = (code << 12) | (uint32_t)i;
    }
    v[r] = k;
  }
  // bitonic sort: strides 1-2 within a thread, 4-128 across the lanes of a wave, 256 and up
  // through LDS
  for (int len = 2; len <= kOrderChunk; len <<= 1) {
    for (int stride = len >> 1; stride > 0; stride >>= 1) {
      if (kOrderChunk > 256 && stride >= 256) {
        __syncthreads();   // the previous LDS step's reads are done
#pragma unroll
        for (int r = 0; r < 4; ++r) key[4 * t + r] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = 4 * t + r;
          v[r] = bitonic_keep(v[r], key[e ^ stride], e, len, stride);
        }
      } else if (stride >= 4) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t pv = (uint32_t)__shfl_xor((int)v[r], stride >> 2, 64);
          v[r] = bitonic_keep(v[r], pv, 4 * t + r, len, stride);
        }
      } else {
        uint32_t nv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) nv[r] = bitonic_keep(v[r], v[r ^ stride], 4 * t + r, len, stride);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = nv[r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * t + r;
    if (i < m) {
      const int32_t g = c0 + (int32_t)(v[r] & 0xFFFu);
      jb.order[c0 + i] = g;
      if (jb.rank) jb.rank[g] = c0 + i;
    }
  }
}

// ------------------------------------------------------------------------------------------
// (a) preprocess: grid (ceil(max_n/256), n_jobs), one thread per Gaussian of the job.
//     Reads means (N,3), prec (N,3,3) (upper triangle), w (N,) -- a wave reads one
//     contiguous span of each array.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_preprocess(const Job* __restrict__ jobs,
                                                    float* __restrict__ rec,
                                                    float* __restrict__ rec2,
                                                    int32_t* __restrict__ job_flags) {
  const Job& jb = jobs[blockIdx.y];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= jb.n) return;
  const int64_t g = gauss_of(jb, j);   // record j holds Gaussian g
  const float* mu = jb.means + 3 * g;
  const float* P = jb.prec + 9 * g;
  const float u00 = P[0], u01 = P[1], u02 = P[2], u11 = P[4], u12 = P[5], u22 = P[8];
  const float w = jb.wlog[g];
  const float p0 = mu[0] - jb.o[0], p1 = mu[1] - jb.o[1], p2 = mu[2] - jb.o[2];
  // M = A R^T with A = U^T (lower): M[i][k] = A[i] . R[k]
  float M[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float r0 = jb.R[3 * k + 0], r1 = jb.R[3 * k + 1], r2 = jb.R[3 * k + 2];
    M[0][k] = u00 * r0;
    M[1][k] = fmaf(u01, r0, u11 * r1);
    M[2][k] = fmaf(u02, r0, fmaf(u12, r1, u22 * r2));
  }
  float out[16];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    out[0 + i] = M[i][0] / jb.fx;
    out[3 + i] = M[i][1] / jb.fy;
    out[6 + i] = M[i][2];
  }
  out[9] = u00 * p0;
  out[10] = fmaf(u01, p0, u11 * p1);
  out[11] = fmaf(u02, p0, fmaf(u12, p1, u22 * p2));
  out[12] = w;
  out[13] = w * kLog2e;
  out[14] = dot3(out + 9, out + 9);
  bool finite = true;
#pragma unroll
  for (int i = 0; i < 15; ++i) finite = finite && isfinite(out[i]);
  const bool degenerate = !finite || u00 == 0.0f || u11 == 0.0f || u22 == 0.0f;
  out[15] = __int_as_float(degenerate ? 1 : 0);
  if (degenerate) atomicOr(job_flags + blockIdx.y, kFlagDegenerate);
  float4* dst = reinterpret_cast<float4*>(rec + (jb.rec_base + j) * kRecFloats);
#pragma unroll
  for (int i = 0; i < 4; ++i) dst[i] = make_float4(out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);

  // blend record (fe_common.h), in float64 from the same float inputs
  const double A0[3] = {(double)u00, 0.0, 0.0};
  const double A1[3] = {(double)u01, (double)u11, 0.0};
  const double A2[3] = {(double)u02, (double)u12, (double)u22};
  const double* Ar[3] = {A0, A1, A2};
  const double pd[3] = {(double)mu[0] - (double)jb.o[0], (double)mu[1] - (double)jb.o[1],
                        (double)mu[2] - (double)jb.o[2]};
  const double sc[3] = {1.0 / (double)jb.fx, 1.0 / (double)jb.fy, 1.0};
  double Md[3][3], qd[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      Md[i][k] = (Ar[i][0] * (double)jb.R[3 * k] + Ar[i][1] * (double)jb.R[3 * k + 1] +
                  Ar[i][2] * (double)jb.R[3 * k + 2]) * sc[k];
    qd[i] = Ar[i][0] * pd[0] + Ar[i][1] * pd[1] + Ar[i][2] * pd[2];
  }
  const double qq = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
  float o2[16];
  bool ok = !degenerate && qq >= kMinQQ && qq < 1e30;
  if (ok) {
    const double iq = 1.0 / sqrt(qq);
    const double n[3] = {qd[0] * iq, qd[1] * iq, qd[2] * iq};
    double e1[3], e2[3];
    ortho_basis(n, e1, e2);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double mq = Md[0][c] * qd[0] + Md[1][c] * qd[1] + Md[2][c] * qd[2];
      const double m1 = Md[0][c] * e1[0] + Md[1][c] * e1[1] + Md[2][c] * e1[2];
      const double m2 = Md[0][c] * e2[0] + Md[1][c] * e2[1] + Md[2][c] * e2[2];
      o2[c] = (float)(mq / qq);
      o2[3 + c] = (float)(m1 * iq);
      o2[6 + c] = (float)(m2 * iq);
    }
    o2[9] = (float)((double)w * 1.4426950408889634);
    o2[10] = (float)(qq * -0.7213475204444817);
    o2[11] = (float)qq;
#pragma unroll
    for (int i = 0; i < 12; ++i) ok = ok && isfinite(o2[i]);
  }
  if (!ok) {
#pragma unroll
    for (int i = 0; i < 12; ++i) o2[i] = 0.0f;
    if (!degenerate) atomicOr(job_flags + blockIdx.y, kFlagNoBlendRec);
  }
  o2[12] = o2[13] = o2[14] = o2[15] = 0.0f;
  float4* dst2 = reinterpret_cast<float4*>(rec2 + (jb.rec_base + j) * kRec2Floats);
#pragma unroll
  for (int i = 0; i < 4; ++i) dst2[i] = make_float4(o2[4 * i], o2[4 * i + 1], o2[4 * i + 2], o2[4 * i + 3]);
}

// ------------------------------------------------------------------------------------------
// (b) tile membership.  One block (4 waves) per super-tile of sf x sf tiles; each wave tests
// 64 Gaussians at a time.  A lane loads its Gaussian's record once and bounds it over the
// whole super-tile (a pair the group bound rejects is rejected for every tile in it: the
// group bound is a bound over a superset of rays).  The survivors go to an LDS queue, and
// the per-tile bounds run on the queue with every lane busy (about 1 Gaussian in 10
// survives, so testing them in their original waves would leave most lanes idle).  Each
// tile's decisions are set as bits of 64-bit membership words held in LDS for a window of
// kWinWords words (4096 Gaussians), then written out and counted into the tile histogram.
// PHASE 1 keeps pairs whose density bound reaches exp(-c_alpha):  d0_lo/2 - w <= c_alpha,
//   writing mask1.
// PHASE 2 keeps pairs absent from mask1 whose depth-logit bound reaches the tile threshold
//   tau = min_pixels(log2-sum-exp of phase-1 logits) - delta2, writing mask2.  It also takes
//   the largest bound `ubm` of the logit, where masked, of every pair absent from mask1 that
//   might lie behind the camera, started from the phase-1 floor seed: if the max logit could
//   rise so far above the log2-sum-exp that the 1e-20 weight floor becomes visible
//   (ubm - lmin >= floor_thr, or a seed of +inf), the tile is completed densely (every
//   remaining pair), since the floor term sums over all n Gaussians.
// ------------------------------------------------------------------------------------------
constexpr int kWinWords = 64;   // membership words per LDS window

// Finish one tile of a group (one full wave): its pair count and not-known-in-front flag; in
// phase 2 also the largest masked-logit bound `ubm` (with the phase-1 floor seed), and the dense
// completion of the tile when the floor could become visible.
template <int PHASE>
__device__ void finish_tile(const Job& jb, int64_t tile, uint32_t total, uint32_t np, float ubm, float lmin,
                            int lane, const unsigned long long* __restrict__ mask1,
                            unsigned long long* __restrict__ mask2, uint32_t* __restrict__ count,
                            const float* __restrict__ tile_seed, float* __restrict__ tile_ubm,
                            uint32_t* __restrict__ tile_np, Counters* __restrict__ cnt_out) {
  const int64_t gt = jb.tile_base + tile;
  if (PHASE == 2) {
    const float ubm_t = fmaxf(ubm, tile_seed[gt]);
    const bool dense_tile = (ubm_t - lmin >= jb.floor_thr);   // false when nothing in doubt
    if (dense_tile) {
      // complete the tile: every pair not in phase 1
      const int64_t row = jb.mask_base + tile * jb.nw;
      uint32_t c2 = 0;
      for (int c = 0; c < jb.nw; ++c) {
        const int j = c * 64 + lane;
        const bool keep = (j < jb.n) && !((mask1[row + c] >> lane) & 1ull);
        const unsigned long long word = __ballot(keep);
        if (lane == 0) mask2[row + c] = word;
        c2 += (uint32_t)__popcll(word);
      }
      total = c2;
      np = 1;   // completed without bounds: none is known to be in front
      if (lane == 0) atomicAdd(&cnt_out->floor_tiles, 1ull);
    }
    if (lane == 0) tile_ubm[gt] = dense_tile ? -INFINITY : ubm_t;
  }
  if (lane == 0) {
    count[gt] = total;
    tile_np[2 * gt + (PHASE - 1)] = np;
  }
}

// Per lane, over the Gaussians of words [w0, w1) (lane = Gaussian, waves interleaved), for each
// tile k of the group with need(k) (block-uniform), the largest lower bound
// b3 (w - sd_hi^2/2) - b2 t_ub of the logit at every pixel of tile tg[k] of a Gaussian entirely
// in front of its rays (-inf if none), into lb[k]; then the block's maxima into sh[wave][k].
// One pass serves all such tiles: each record is loaded once and bounded against each of them
// (the same values, and maxima, as a pass per tile).
template <int NS, typename Need>
__device__ __forceinline__ void front_logit_lbs(const Job& jb, const float* __restrict__ rec, const TileGeo* tg,
                                                Need need, int w0, int w1, int wave, int lane, float b3l,
                                                float b2l, float (*sh)[NS]) {
  float lb[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) lb[k] = -INFINITY;
  for (int c = w0 + wave; c < w1; c += 4) {
    const int j = c * 64 + lane;
    if (j >= jb.n) continue;
    const float4* r4 = reinterpret_cast<const float4*>(rec + (jb.rec_base + j) * kRecFloats);
    float r[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 t4 = r4[i];
      r[4 * i] = t4.x; r[4 * i + 1] = t4.y; r[4 * i + 2] = t4.z; r[4 * i + 3] = t4.w;
    }
    if (__float_as_int(r[15]) & 1) continue;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      if (!need(k)) continue;   // block-uniform
      const PairBound pb = pair_bound(r, tg[k]);
      if (pb.plus && !pb.unbounded)
        lb[k] = fmaxf(lb[k], fmaf(b3l, fmaf(-0.5f * pb.sd_hi, pb.sd_hi, r[12]), -(b2l * pair_t_ub(pb))));
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (!need(k)) continue;   // block-uniform
    float v = lb[k];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if (lane == 0) sh[wave][k] = v;
  }
  __syncthreads();
}

// Phase 2 of a chunked membership (chunks > 1): for every tile of a group whose phase-1 lmin is
// -inf, the largest lower bound b3 (w - sd_hi^2/2) - b2 t_ub of the logit at every pixel of the
// tile over the Gaussians of this block's chunk that lie entirely in front of the tile's rays
// (-inf if none), to lbpart[(group * chunks + chunk) * kMaskMaxSub + tile of the group]
// (k_tile_mask takes the max over the chunks).
template <int SFMAX>
__global__ __launch_bounds__(256) void k_mask_lb(const Job* __restrict__ jobs, int n_jobs, int ts,
                                                 const float* __restrict__ rec, const float* __restrict__ tile_lmin,
                                                 const int32_t* __restrict__ job_flags, int chunks,
                                                 float* __restrict__ lbpart) {
  constexpr int NS = SFMAX * SFMAX;
  static_assert(NS <= kMaskMaxSub, "partials hold kMaskMaxSub tiles per group");
  const int64_t gs = blockIdx.x;
  const int ji = find_job(n_jobs, gs, [&](int k) { return jobs[k].stile_base; });
  const Job& jb = jobs[ji];
  if ((job_flags[ji] & 1) != 0 || !bounds_apply(jb)) return;   // dense job: no thresholds
  const int sf = jb.sf;
  const int st = (int)(gs - jb.stile_base);
  const int sx = st % jb.stiles_x, sy = st / jb.stiles_x;
  const int tx0 = sx * sf, ty0 = sy * sf;
  const int nsx = min(sf, jb.tiles_x - tx0), nsy = min(sf, jb.tiles_y - ty0);
  const int nsub = nsx * nsy;
  __shared__ TileGeo tg[NS];
  __shared__ int unc[NS];
  __shared__ int any_unc;
  __shared__ float sh_lb[4][NS];
  if (threadIdx.x == 0) any_unc = 0;
  __syncthreads();
  if (threadIdx.x < NS) {
    const int k = threadIdx.x;
    int u = 0;
    if (k < nsub) {
      const int txk = tx0 + k % nsx, tyk = ty0 + k / nsx;
      u = tile_lmin[jb.tile_base + (int64_t)tyk * jb.tiles_x + txk] == -INFINITY;
      tg[k] = tile_geo_xy(jb, txk, tyk, ts);
    }
    unc[k] = u;
    if (u) atomicOr(&any_unc, 1);
  }
  __syncthreads();
  if (!any_unc) return;   // block-uniform
  const int cw = (jb.nw + chunks - 1) / chunks;
  const int wlo = min(jb.nw, (int)blockIdx.y * cw), whi = min(jb.nw, wlo + cw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float b3l = jb.beta3 * kLog2e, b2l = jb.beta2 * kLog2e;
  front_logit_lbs<NS>(jb, rec, tg, [&](int k) { return unc[k] != 0; }, wlo, whi, wave, lane, b3l, b2l, sh_lb);
  if (threadIdx.x < nsub && unc[threadIdx.x]) {
    const int k = threadIdx.x;
    lbpart[((int64_t)gs * chunks + blockIdx.y) * kMaskMaxSub + k] =
        fmaxf(fmaxf(sh_lb[0][k], sh_lb[1][k]), fmaxf(sh_lb[2][k], sh_lb[3][k]));
  }
}

// Phase 2 with 3x3 groups is held to 128 VGPRs (4 waves per SIMD instead of 3; a few
// registers spill): 0.171 -> 0.152 ms.  Phase 1 at 5 waves (96 VGPRs) spilled into a slower
// kernel.
template <int PHASE, int SFMAX>
__global__ __launch_bounds__(256, (PHASE == 2 && SFMAX == 3) ? 4 : 1) void k_tile_mask(const Job* __restrict__ jobs, int n_jobs, int ts,
                                                   const float* __restrict__ rec,
                                                   unsigned long long* __restrict__ mask1,
                                                   unsigned long long* __restrict__ mask2,
                                                   uint32_t* __restrict__ count,
                                                   const float* __restrict__ tile_lmin,
                                                   const float* __restrict__ tile_seed,
                                                   float* __restrict__ tile_ubm,
                                                   uint32_t* __restrict__ tile_np,
                                                   const int32_t* __restrict__ job_flags,
                                                   Counters* __restrict__ cnt_out, int chunks,
                                                   uint32_t* __restrict__ mpart,
                                                   const float* __restrict__ lbpart) {
  constexpr int NS = SFMAX * SFMAX;   // sub-tiles per group at most (ray-list views use 1)
  static_assert(NS <= kMaskMaxSub, "partials hold kMaskMaxSub tiles per group");
  __shared__ float sh_ubm[4][NS];
  const int64_t gs = blockIdx.x;
  const int ji = find_job(n_jobs, gs, [&](int k) { return jobs[k].stile_base; });
  const Job& jb = jobs[ji];
  const int sf = jb.sf;
  const int st = (int)(gs - jb.stile_base);
  const int sx = st % jb.stiles_x, sy = st / jb.stiles_x;
  const TileGeo sg = tile_geo_xy(jb, sx, sy, ts * sf);
  // sub-tiles of this group that exist
  const int tx0 = sx * sf, ty0 = sy * sf;
  const int nsx = min(sf, jb.tiles_x - tx0), nsy = min(sf, jb.tiles_y - ty0);
  const int nsub = nsx * nsy;
  // per-sub-tile geometry and thresholds, block-uniform, kept in LDS
  __shared__ TileGeo tg[NS];
  __shared__ float tau[NS], lmin[NS];
  __shared__ float tau_grp;
  if (threadIdx.x < nsub) {
    const int k = threadIdx.x;
    const int txk = tx0 + k % nsx, tyk = ty0 + k / nsx;
    tg[k] = tile_geo_xy(jb, txk, tyk, ts);
    if (PHASE == 2) {
      lmin[k] = tile_lmin[jb.tile_base + (int64_t)tyk * jb.tiles_x + txk];
      tau[k] = lmin[k] - jb.delta2;
    }
  }
  __syncthreads();
  const bool dense_job = (job_flags[ji] & 1) != 0 || !bounds_apply(jb);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float b3l = jb.beta3 * kLog2e, b2l = jb.beta2 * kLog2e;
  if (PHASE == 2 && !dense_job) {
    // A sub-tile with a pixel that no phase-1 pair covers on the visible side has
    // lmin = -inf.  Its threshold then comes from bounds: any Gaussian entirely in front of
    // the tile's rays has logit >= lb = b3 (w - sd_hi^2/2) - b2 t_ub at every pixel, so the
    // log2-sum-exp is at least max_j lb_j there (and that Gaussian passes the test below).
    if (chunks > 1) {
      // the maximum of the chunks' maxima (k_mask_lb)
      if (threadIdx.x < nsub && lmin[threadIdx.x] == -INFINITY) {
        const int k = threadIdx.x;
        float m = -INFINITY;
        for (int q = 0; q < chunks; ++q) m = fmaxf(m, lbpart[((int64_t)gs * chunks + q) * kMaskMaxSub + k]);
        lmin[k] = m;
        tau[k] = m - jb.delta2;
      }
      __syncthreads();
    } else {
      __shared__ float sh_lb[4][NS];
      __shared__ int any_unc;
      if (threadIdx.x == 0) {
        int u = 0;
        for (int k = 0; k < nsub; ++k) u |= (lmin[k] == -INFINITY) ? 1 : 0;
        any_unc = u;
      }
      __syncthreads();
      if (any_unc) {   // block-uniform
        front_logit_lbs<NS>(jb, rec, tg, [&](int k) { return k < nsub && lmin[k] == -INFINITY; }, 0, jb.nw, wave,
                            lane, b3l, b2l, sh_lb);
        if (threadIdx.x < nsub && lmin[threadIdx.x] == -INFINITY) {
          const int k = threadIdx.x;
          const float m = fmaxf(fmaxf(sh_lb[0][k], sh_lb[1][k]), fmaxf(sh_lb[2][k], sh_lb[3][k]));
          lmin[k] = m;
          tau[k] = m - jb.delta2;
        }
        __syncthreads();
      }
    }
  }
  if (PHASE == 2 && threadIdx.x == 0) {
    float m = INFINITY;
    for (int k = 0; k < nsub; ++k) m = fminf(m, tau[k]);
    tau_grp = m;
  }
  __syncthreads();
  const float tau_s = (PHASE == 2) ? tau_grp : 0.0f;
  __shared__ unsigned long long wmask[NS][kWinWords];   // this window's membership words
  // phase 2: the phase-1 words of the window (phase 1 has none: its block then fits 7 times
  // per CU instead of 6)
  __shared__ unsigned long long wlist[PHASE == 2 ? NS : 1][kWinWords];
  auto listed_in_1 = [&](int k, int c, int bit) {
    if constexpr (PHASE == 2) return ((wlist[k][c] >> bit) & 1ull) != 0ull;
    else return false;
  };
  __shared__ uint32_t queue[kWinWords * 64];            // group survivors (Gaussian ids)
  __shared__ uint32_t qn;
  __shared__ uint32_t tcnt[NS], tnp[NS];
  if (threadIdx.x < NS) { tcnt[threadIdx.x] = 0; tnp[threadIdx.x] = 0; }
  float ubm[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) ubm[k] = -INFINITY;
  const bool single = (nsub == 1);   // the group is the tile: its bound decides directly
  // this block's share of the Gaussians (chunks > 1 on small grids): words [wlo, whi) of 64
  const int cw = (jb.nw + chunks - 1) / chunks;
  const int wlo = min(jb.nw, (int)blockIdx.y * cw), whi = min(jb.nw, wlo + cw);

  for (int w0 = wlo; w0 < whi; w0 += kWinWords) {
    const int nww = min(kWinWords, whi - w0);
    for (int i = threadIdx.x; i < NS * kWinWords; i += 256) {
      const int k = i / kWinWords, c = i - k * kWinWords;
      wmask[k][c] = 0ull;
      unsigned long long lw = 0ull;
      if (PHASE == 2 && k < nsub && c < nww) {
        const int64_t tile = (int64_t)(ty0 + k / nsx) * jb.tiles_x + (tx0 + k % nsx);
        lw = mask1[jb.mask_base + tile * jb.nw + w0 + c];
      }
      if constexpr (PHASE == 2) wlist[k][c] = lw;
    }
    if (threadIdx.x == 0) qn = 0;
    __syncthreads();

    // pass 1: group bound of every Gaussian of the window (lane = Gaussian)
    for (int c = wave; c < nww; c += 4) {
      const int j = (w0 + c) * 64 + lane;
      const bool valid = j < jb.n;
      float r[16];
      if (valid) {
        const float4* r4 = reinterpret_cast<const float4*>(rec + (jb.rec_base + j) * kRecFloats);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 t4 = r4[i];
          r[4 * i] = t4.x; r[4 * i + 1] = t4.y; r[4 * i + 2] = t4.z; r[4 * i + 3] = t4.w;
        }
      }
      const bool degenerate = valid && ((__float_as_int(r[15]) & 1) != 0);
      const bool forced = valid && (dense_job || degenerate);   // kept in phase 1, never in phase 2
      bool keep_s = false, plus_s = true, front_s = false;
      float ubm_s = -INFINITY;
      bool keep1 = false;   // single-tile groups: the decision itself
      if (valid && !forced) {
        const PairBound pb = pair_bound(r, sg);
        const float w = r[12];
        front_s = pb.front;
        if (PHASE == 1) {
          keep_s = pb.unbounded || (fmaf(0.5f * pb.sd_lo, pb.sd_lo, -w) <= jb.c_alpha);
        } else {
          keep_s = pb.unbounded || depth_keep(pb, w, b3l, b2l, tau_s);
          plus_s = pb.plus;
          if (!pb.plus) ubm_s = masked_logit_ub(pb, w, b3l, b2l);
        }
        keep1 = keep_s;
      }
      if (PHASE == 2 && valid && !forced) {
        // pairs listed in phase 1 are skipped; any other pair that may lie behind the camera
        // raises the tile's ubm: here from the group bound (which covers every sub-tile) when
        // the group drops it or the group is the tile, else from the tile bound in pass 2
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          if (k < nsub) {
            const bool listed = listed_in_1(k, c, lane);
            if (!listed && !plus_s && (single || !keep_s)) ubm[k] = fmaxf(ubm[k], ubm_s);
          }
        }
      }
      if (single) {
        // the group bound is the tile bound
        const bool keep = (PHASE == 1) ? (forced || keep1) : (keep1 && !listed_in_1(0, c, lane));
        const unsigned long long word = __ballot(keep);
        const unsigned long long npw = __ballot(keep && !front_s);   // forced pairs: front_s false
        if (lane == 0) {
          wmask[0][c] = word;
          if (npw != 0ull) tnp[0] = 1;
        }
        continue;
      }
      if (PHASE == 1) {
        const unsigned long long fw = __ballot(forced);
        if (fw != 0ull && lane == 0) {
          for (int k = 0; k < nsub; ++k) { wmask[k][c] = fw; tnp[k] = 1; }
        }
      }
      bool need = keep_s;
      if (PHASE == 2 && keep_s) {
        // a Gaussian listed in phase 1 on every sub-tile has nothing left to test
        bool all_listed = true;
#pragma unroll
        for (int k = 0; k < NS; ++k)
          if (k < nsub) all_listed = all_listed && listed_in_1(k, c, lane);
        need = !all_listed;
      }
      const unsigned long long sw = __ballot(need);
      if (sw != 0ull) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&qn, (uint32_t)__popcll(sw));
        base = __shfl(base, 0, 64);
        const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (need) queue[base + __popcll(sw & lt)] = (uint32_t)j;
      }
    }
    __syncthreads();

    // pass 2: per-tile bounds of the survivors, one survivor per lane
    if (!single) {
      const uint32_t nq = qn;
      for (uint32_t i = threadIdx.x; i < nq; i += 256) {
        const int j = (int)queue[i];
        const int c = (j >> 6) - w0;
        const int bit = j & 63;
        float r[16];
        const float4* r4 = reinterpret_cast<const float4*>(rec + (jb.rec_base + j) * kRecFloats);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 t4 = r4[q];
          r[4 * q] = t4.x; r[4 * q + 1] = t4.y; r[4 * q + 2] = t4.z; r[4 * q + 3] = t4.w;
        }
        const float w = r[12];
        // phase 1 keeps this loop rolled (it indexes only LDS by k): 71 VGPRs instead of 107,
        // more resident waves, 0.0997 -> 0.0964 ms; phase 2's per-lane maxima ubm[k] need it
        // unrolled to stay in registers
        constexpr int kSubUnroll = (PHASE == 1) ? 1 : NS;
#pragma unroll kSubUnroll
        for (int k = 0; k < NS; ++k) {
          if (k >= nsub) continue;
          if (listed_in_1(k, c, bit)) continue;   // listed in phase 1
          const PairBound pb = pair_bound(r, tg[k]);
          bool keep;
          if (PHASE == 1) {
            keep = pb.unbounded || (fmaf(0.5f * pb.sd_lo, pb.sd_lo, -w) <= jb.c_alpha);
          } else {
            keep = pb.unbounded || depth_keep(pb, w, b3l, b2l, tau[k]);
            if (!pb.plus) ubm[k] = fmaxf(ubm[k], masked_logit_ub(pb, w, b3l, b2l));
          }
          if (keep) {
            atomicOr(&wmask[k][c], 1ull << bit);
            if (!pb.front) tnp[k] = 1;
          }
        }
      }
    }
    __syncthreads();

    // write the window's words out and count them
    for (int i = threadIdx.x; i < nsub * kWinWords; i += 256) {
      const int k = i / kWinWords, c = i - k * kWinWords;
      if (c < nww) {
        const int64_t tile = (int64_t)(ty0 + k / nsx) * jb.tiles_x + (tx0 + k % nsx);
        const unsigned long long word = wmask[k][c];
        if (PHASE == 1) mask1[jb.mask_base + tile * jb.nw + w0 + c] = word;
        else mask2[jb.mask_base + tile * jb.nw + w0 + c] = word;
        if (word != 0ull) atomicAdd(&tcnt[k], (uint32_t)__popcll(word));
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (PHASE == 2) {
      float u = ubm[k];
      for (int o = 32; o > 0; o >>= 1) u = fmaxf(u, __shfl_xor(u, o, 64));
      ubm[k] = u;
    }
    if (lane == 0) sh_ubm[wave][k] = ubm[k];
  }
  __syncthreads();
  if (chunks > 1) {
    // leave this block's per-tile partials (count, flag, bound, threshold); k_mask_finish
    // merges them (sum, or, max: the same totals in any order) and finishes the tiles
    if (threadIdx.x < nsub) {
      const int k = threadIdx.x;
      uint32_t* mp = mpart + ((int64_t)gs * chunks + blockIdx.y) * (4 * kMaskMaxSub) + 4 * k;
      mp[0] = tcnt[k];
      mp[1] = tnp[k];
      mp[2] = __float_as_uint(fmaxf(fmaxf(sh_ubm[0][k], sh_ubm[1][k]), fmaxf(sh_ubm[2][k], sh_ubm[3][k])));
      mp[3] = __float_as_uint(PHASE == 2 ? lmin[k] : 0.0f);
    }
    return;
  }
  // one wave finishes the sub-tiles
  if (wave == 0) {
    for (int k = 0; k < nsub; ++k) {
      const int64_t tile = (int64_t)(ty0 + k / nsx) * jb.tiles_x + (tx0 + k % nsx);
      const float u = fmaxf(fmaxf(sh_ubm[0][k], sh_ubm[1][k]), fmaxf(sh_ubm[2][k], sh_ubm[3][k]));
      finish_tile<PHASE>(jb, tile, tcnt[k], tnp[k], u, lmin[k], lane, mask1, mask2, count, tile_seed, tile_ubm,
                         tile_np, cnt_out);
    }
  }
}

// Merges the partials of a group's chunks (k_tile_mask with chunks > 1) and finishes its
// tiles: one wave per group.  The lanes load every (tile, chunk) partial at once and combine
// them in LDS (count: sum, flag: or, bound: max through an order-preserving integer key; NaN
// bounds are dropped, as fmaxf drops them), then the wave finishes the tiles one by one.
__device__ __forceinline__ int float_order_key(float f) {
  const int b = __float_as_int(f);
  return b >= 0 ? b : (b ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float float_from_order_key(int k) {
  return __int_as_float(k >= 0 ? k : (k ^ 0x7FFFFFFF));
}

template <int PHASE>
__global__ __launch_bounds__(64) void k_mask_finish(const Job* __restrict__ jobs, int n_jobs, int chunks,
                                                    const uint32_t* __restrict__ mpart,
                                                    const unsigned long long* __restrict__ mask1,
                                                    unsigned long long* __restrict__ mask2,
                                                    uint32_t* __restrict__ count, const float* __restrict__ tile_seed,
                                                    float* __restrict__ tile_ubm, uint32_t* __restrict__ tile_np,
                                                    Counters* __restrict__ cnt_out) {
  __shared__ uint32_t s_cnt[kMaskMaxSub], s_np[kMaskMaxSub];
  __shared__ int s_ubm[kMaskMaxSub];
  const int64_t gs = blockIdx.x;
  const int ji = find_job(n_jobs, gs, [&](int k) { return jobs[k].stile_base; });
  const Job& jb = jobs[ji];
  const int sf = jb.sf;
  const int st = (int)(gs - jb.stile_base);
  const int sx = st % jb.stiles_x, sy = st / jb.stiles_x;
  const int tx0 = sx * sf, ty0 = sy * sf;
  const int nsx = min(sf, jb.tiles_x - tx0), nsy = min(sf, jb.tiles_y - ty0);
  const int nsub = nsx * nsy;
  const int lane = threadIdx.x;
  if (lane < kMaskMaxSub) {
    s_cnt[lane] = 0;
    s_np[lane] = 0;
    s_ubm[lane] = float_order_key(-INFINITY);
  }
  __syncthreads();
  const uint32_t* base = mpart + gs * chunks * (4 * kMaskMaxSub);
  for (int f = lane; f < nsub * chunks; f += 64) {
    const int q = f / nsub, k = f - q * nsub;
    const uint32_t* mp = base + (int64_t)q * (4 * kMaskMaxSub) + 4 * k;
    const uint32_t c = mp[0], np = mp[1];
    const float u = __uint_as_float(mp[2]);
    if (c) atomicAdd(&s_cnt[k], c);
    if (np) atomicOr(&s_np[k], np);
    if (u == u) atomicMax(&s_ubm[k], float_order_key(u));
  }
  __syncthreads();
  for (int k = 0; k < nsub; ++k) {
    const float lm = __uint_as_float(base[4 * k + 3]);   // every chunk holds the same threshold
    const int64_t tile = (int64_t)(ty0 + k / nsx) * jb.tiles_x + (tx0 + k % nsx);
    finish_tile<PHASE>(jb, tile, s_cnt[k], s_np[k], float_from_order_key(s_ubm[k]), lm, lane, mask1, mask2, count,
                       tile_seed, tile_ubm, tile_np, cnt_out);
  }
}

// ------------------------------------------------------------------------------------------
// ray-list views: bounding rectangle of each tile's rays (one block per tile)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ray_rects(const Job* __restrict__ jobs, int n_jobs, int ts,
                                                   int32_t* __restrict__ job_flags) {
  __shared__ float sh[4][4];
  const int64_t gt = blockIdx.x;
  const int ji = find_job(n_jobs, gt, [&](int k) { return jobs[k].tile_base; });
  const Job& jb = jobs[ji];
  if (!jb.ray_xy) return;   // block-uniform
  const int tile = (int)(gt - jb.tile_base);
  const int nv = tile_valid(jb, tile, 0, tile, ts);
  const int64_t k0 = (int64_t)tile * ts * ts;
  float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  bool bad = false;
  for (int i = threadIdx.x; i < nv; i += 256) {
    const float x = jb.ray_xy[2 * (k0 + i)], y = jb.ray_xy[2 * (k0 + i) + 1];
    bad |= !(fabsf(x) <= 3.0e38f) || !(fabsf(y) <= 3.0e38f);
    xmin = fminf(xmin, x); xmax = fmaxf(xmax, x);
    ymin = fminf(ymin, y); ymax = fmaxf(ymax, y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    xmin = fminf(xmin, __shfl_xor(xmin, o, 64)); xmax = fmaxf(xmax, __shfl_xor(xmax, o, 64));
    ymin = fminf(ymin, __shfl_xor(ymin, o, 64)); ymax = fmaxf(ymax, __shfl_xor(ymax, o, 64));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[wave][0] = xmin; sh[wave][1] = xmax; sh[wave][2] = ymin; sh[wave][3] = ymax; }
  const int any_bad = __syncthreads_or(bad ? 1 : 0);
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      xmin = fminf(xmin, sh[w][0]); xmax = fmaxf(xmax, sh[w][1]);
      ymin = fminf(ymin, sh[w][2]); ymax = fmaxf(ymax, sh[w][3]);
    }
    reinterpret_cast<float4*>(jb.rect)[tile] = ray_rect(xmin, xmax, ymin, ymax);
    // a non-finite ray makes the view dense with the NaN-safe blend, like degenerate records
    if (any_bad) atomicOr(&job_flags[ji], 1);
  }
}

// ------------------------------------------------------------------------------------------
// exclusive prefix sum: uint32 counts -> uint64 offsets (off[n] = base + total)
// ------------------------------------------------------------------------------------------
constexpr int kScanBlock = 1024;  // items per block (256 threads x 4)

__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned long long y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

// block-wide exclusive scan over 256 threads; returns exclusive prefix, total via *tot
__device__ __forceinline__ unsigned long long block_excl_scan256(unsigned long long v,
                                                                 unsigned long long* sh,
                                                                 unsigned long long* tot) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned long long inc = wave_incl_scan(v);
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  unsigned long long wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += sh[w];
  *tot = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return wbase + inc - v;
}

// Dispatch key of tile i: its list length, plus its phase-1 length when key_off (the
// phase-1 offsets) is given, so that the backward orders tiles by both lists.
__device__ __forceinline__ uint32_t tile_key(const uint32_t* cnt, const unsigned long long* key_off, int64_t i) {
  return cnt[i] + (key_off ? (uint32_t)(key_off[i + 1] - key_off[i]) : 0u);
}

// With hist, each block also adds its tiles' dispatch-bucket histogram into hist.
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ cnt, int64_t n,
                                                     unsigned long long* __restrict__ partial,
                                                     const unsigned long long* __restrict__ key_off,
                                                     uint32_t* __restrict__ hist) {
  static_assert(kTileBuckets == 256, "one bucket per thread");
  __shared__ unsigned long long sh[4];
  __shared__ uint32_t lh[kTileBuckets];
  if (hist) {   // grid-uniform
    lh[threadIdx.x] = 0u;
    __syncthreads();
  }
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock;
  unsigned long long s = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = b0 + threadIdx.x * 4 + k;
    if (i < n) {
      s += cnt[i];
      if (hist) atomicAdd(&lh[tile_bucket(tile_key(cnt, key_off, i))], 1u);
    }
  }
  unsigned long long tot;
  block_excl_scan256(s, sh, &tot);   // ends with a barrier
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
  if (hist && lh[threadIdx.x] != 0u) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

// single block: partial[i] <- base + exclusive prefix of partial; with hist, the bucket
// counts become the buckets' first positions, largest bucket first
__global__ __launch_bounds__(256) void k_scan_top(unsigned long long* __restrict__ partial, int64_t nb,
                                                  const unsigned long long* __restrict__ base_src,
                                                  uint32_t* __restrict__ hist) {
  __shared__ unsigned long long sh[4];
  unsigned long long carry = base_src ? *base_src : 0ull;
  for (int64_t c = 0; c < nb; c += 256) {
    const int64_t i = c + threadIdx.x;
    const unsigned long long v = (i < nb) ? partial[i] : 0ull;
    unsigned long long tot;
    const unsigned long long ex = block_excl_scan256(v, sh, &tot);
    if (i < nb) partial[i] = carry + ex;
    carry += tot;
  }
  if (hist) {
    const int b = kTileBuckets - 1 - (int)threadIdx.x;
    unsigned long long tot;
    const unsigned long long ex = block_excl_scan256(hist[b], sh, &tot);
    hist[b] = (uint32_t)ex;
  }
}

// With tord, tile i is also placed in the dispatch order: a block ranks its tiles within
// each bucket in LDS, reserves a run of the bucket's positions with one atomic, and writes
// tord[position] = i.  (Which run a block gets varies from call to call; every tile's
// result is independent of the order in which the tiles run.)
__global__ __launch_bounds__(256) void k_scan_final(const uint32_t* __restrict__ cnt, int64_t n,
                                                    const unsigned long long* __restrict__ partial,
                                                    unsigned long long* __restrict__ off,
                                                    const unsigned long long* __restrict__ key_off,
                                                    uint32_t* __restrict__ hist, int32_t* __restrict__ tord) {
  __shared__ unsigned long long sh[4];
  __shared__ uint32_t lh[kTileBuckets], lb[kTileBuckets];
  if (tord) {   // grid-uniform
    lh[threadIdx.x] = 0u;
    __syncthreads();
  }
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock;
  uint32_t v[4], rk[4];
  int bk[4];
  unsigned long long s = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = b0 + threadIdx.x * 4 + k;
    v[k] = (i < n) ? cnt[i] : 0u;
    s += v[k];
    bk[k] = 0;
    rk[k] = 0u;
    if (tord && i < n) {
      bk[k] = tile_bucket(tile_key(cnt, key_off, i));
      rk[k] = atomicAdd(&lh[bk[k]], 1u);
    }
  }
  unsigned long long tot;
  unsigned long long ex = block_excl_scan256(s, sh, &tot) + partial[blockIdx.x];   // ends with a barrier
  for (int k = 0; k < 4; ++k) {
    const int64_t i = b0 + threadIdx.x * 4 + k;
    if (i < n) off[i] = ex;
    ex += v[k];
    if (i == n - 1) off[n] = ex;
  }
  if (tord) {
    const uint32_t c = lh[threadIdx.x];
    if (c != 0u) lb[threadIdx.x] = atomicAdd(&hist[threadIdx.x], c);
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      const int64_t i = b0 + threadIdx.x * 4 + k;
      if (i < n) tord[lb[bk[k]] + rk[k]] = (int32_t)i;
    }
  }
}

// ------------------------------------------------------------------------------------------
// in-order compaction: one wave per tile, 4 tiles per block
// ------------------------------------------------------------------------------------------
// Position of the j-th set bit (j < popcount) of a 64-bit word, by halving.
__device__ __forceinline__ int select_bit(unsigned long long w, int j) {
  int pos = 0;
#pragma unroll
  for (int half = 32; half > 0; half >>= 1) {
    const unsigned long long lo = w & ((1ull << half) - 1ull);
    const int c = __popcll(lo);
    if (j >= c) { j -= c; w >>= half; pos += half; }
    else w = lo;
  }
  return pos;
}

__global__ __launch_bounds__(256) void k_tile_write(const Job* __restrict__ jobs, int n_jobs,
                                                    int64_t total_tiles,
                                                    const unsigned long long* __restrict__ mask,
                                                    const unsigned long long* __restrict__ off,
                                                    uint32_t* __restrict__ list) {
  // One wave per tile.  Per window of 64 membership words: lane c loads word c, the wave
  // scans the words' popcounts, and then lane k of each round writes list entry k of the
  // window: its word by binary search on the inclusive counts (in LDS), its bit by select.
  // The stores of a round are consecutive.
  __shared__ unsigned long long sw[4][64];
  __shared__ uint32_t sinc[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t gt = (int64_t)blockIdx.x * 4 + wave;
  if (gt >= total_tiles) return;   // wave-uniform; no block barrier below
  const int ji = find_job(n_jobs, gt, [&](int k) { return jobs[k].tile_base; });
  const Job& jb = jobs[ji];
  const int tile = (int)(gt - jb.tile_base);
  const unsigned long long* mrow = mask + jb.mask_base + (int64_t)tile * jb.nw;
  unsigned long long pos = off[gt];
  for (int c0 = 0; c0 < jb.nw; c0 += 64) {
    const int c = c0 + lane;
    const unsigned long long word = (c < jb.nw) ? mrow[c] : 0ull;
    uint32_t inc = (uint32_t)__popcll(word);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)inc, o, 64);
      if (lane >= o) inc += v;
    }
    sw[wave][lane] = word;
    sinc[wave][lane] = inc;
    const uint32_t tot = (uint32_t)__shfl((int)inc, 63, 64);
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = lane; k < tot; k += 64) {
      // first word whose inclusive count exceeds k
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (sinc[wave][lo + step - 1] <= k) lo += step;
      const uint32_t before = sinc[wave][lo] - (uint32_t)__popcll(sw[wave][lo]);
      list[pos + k] = (uint32_t)((c0 + lo) * 64 + select_bit(sw[wave][lo], (int)(k - before)));
    }
    pos += tot;
    __builtin_amdgcn_wave_barrier();
  }
}

// ------------------------------------------------------------------------------------------
// (c) blend.  One block per tile, one thread per pixel (TS*TS threads).
//   PHASE 0: dense (all N), PHASE 1: culled first pass, PHASE 2: culled second pass+finish.
// Per pixel state: M (running max of log2 logits), S = sum mask*2^(l-M), Z = sum
// mask*2^(l-M)*t, E = sum exp(d2), Zall = sum nan_to_num(t), nan flag.
// ------------------------------------------------------------------------------------------
// One LDS batch of `nbr` records (a multiple of SUB) for this thread's pixel.  The online
// maximum is updated once per SUB records.  SAFE keeps the exact NaN/inf semantics
// (nan_to_num of t for the weight floor, NaN-logit flag); it is only needed for views with
// degenerate records, whose t* or logits can be non-finite.
template <int SUB, bool SAFE>
__device__ __forceinline__ void blend_batch(const float4* lds, int nbr, float xf, float yf, float b3, float nb2l,
                                            float kd, float& M, float& S, float& Z, float& E, float& Zall,
                                            bool& nanf) {
  for (int k0 = 0; k0 < nbr; k0 += SUB) {
    float ls[SUB], tsv[SUB];
    float mx = M;
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      const float4 r0 = lds[(k0 + s) * 4 + 0];
      const float4 r1 = lds[(k0 + s) * 4 + 1];
      const float4 r2 = lds[(k0 + s) * 4 + 2];
      const float4 r3 = lds[(k0 + s) * 4 + 3];
      // a = r0.xyz, b = (r0.w, r1.x, r1.y), g = (r1.z, r1.w, r2.x), q = (r2.y, r2.z, r2.w)
      const float u0 = fmaf(r0.x, xf, fmaf(r0.w, yf, r1.z));
      const float u1 = fmaf(r0.y, xf, fmaf(r1.x, yf, r1.w));
      const float u2 = fmaf(r0.z, xf, fmaf(r1.y, yf, r2.x));
      const float vsv = fmaf(u0, u0, fmaf(u1, u1, u2 * u2));
      const float psv = fmaf(r2.y, u0, fmaf(r2.z, u1, r2.w * u2));
      const float t = psv * __builtin_amdgcn_rcpf(vsv);
      const float v0 = fmaf(t, u0, -r2.y);
      const float v1 = fmaf(t, u1, -r2.z);
      const float v2 = fmaf(t, u2, -r2.w);
      const float d0 = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
      const float d2l = fmaf(kd, d0, r3.y);           // log2(e) * d2
      E += __builtin_amdgcn_exp2f(d2l);
      const float l = fmaf(b3, d2l, nb2l * t);        // log2(e) * (b3 d2 - b2 t)
      if (SAFE) {
        nanf = nanf || (l != l);
        Zall += nan_to_num(t);
      } else {
        Zall += t;
      }
      ls[s] = l;
      tsv[s] = t;
      mx = fmaxf(mx, l);
    }
    if (mx > M) {
      const float sc = __builtin_amdgcn_exp2f(M - mx);
      S *= sc;
      Z *= sc;
      M = mx;
    }
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      float sv = __builtin_amdgcn_exp2f(ls[s] - M);
      sv = (tsv[s] > 0.0f) ? sv : 0.0f;
      if (SAFE) sv = (sv == sv) ? sv : 0.0f;   // an infinite logit: exp(inf - inf) -> nan_to_num -> 0
      S += sv;
      Z = fmaf(sv, tsv[s], Z);
    }
  }
}

// The same batch through blend records (fe_common.h), for the PPT pixels of this thread:
// each record is read from LDS once for all of them.  Per pair 3 linear forms, one
// reciprocal and two exponentials.  Logits are kept as lt = l/b3 (log2 units, b3 > 0), so
// the running maximum Mt is of lt and the reference M = b3*Mt is what S and Z are relative
// to (the value saved for the second pass and the backward).  MASK applies the t* > 0
// test; a tile whose listed pairs are all in front of the camera skips it.
template <int SUB, bool MASK, int PPT, bool ROW>
__device__ __forceinline__ void blend_batch_rec2(const float4* lds, int nbr, const float* xf, const float* yf, float b3,
                                                 float kap, float* M, float* Mt, float* S, float* Z, float* E,
                                                 float* Zall) {
  for (int k0 = 0; k0 < nbr; k0 += SUB) {
    float ls[PPT][SUB], tsv[PPT][SUB];
    float mx[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) mx[q] = Mt[q];
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      const float4 r0 = lds[(k0 + s) * 3 + 0];
      const float4 r1 = lds[(k0 + s) * 3 + 1];
      const float4 r2 = lds[(k0 + s) * 3 + 2];
      // kX = r0.xyz, kY1 = (r0.w, r1.x, r1.y), kY2 = (r1.z, r1.w, r2.x), wL = r2.y, -cG = r2.z
      // ROW: the thread's pixels share y, so the forms' y-parts are computed once (the same
      // operations as per pixel, so the values are identical)
      const float RX = ROW ? fmaf(r0.y, yf[0], r0.z) : 0.0f;
      const float RY1 = ROW ? fmaf(r1.x, yf[0], r1.y) : 0.0f;
      const float RY2 = ROW ? fmaf(r1.w, yf[0], r2.x) : 0.0f;
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const float X = fmaf(r0.x, xf[q], ROW ? RX : fmaf(r0.y, yf[q], r0.z));
        const float Y1 = fmaf(r0.w, xf[q], ROW ? RY1 : fmaf(r1.x, yf[q], r1.y));
        const float Y2 = fmaf(r1.z, xf[q], ROW ? RY2 : fmaf(r1.w, yf[q], r2.x));
        const float W = fmaf(Y1, Y1, Y2 * Y2);
        const float rr = __builtin_amdgcn_rcpf(fmaf(X, X, W));
        const float t = X * rr;
        // log2(e) * d2, rounded exactly as the backward recomputes it (it finds the pair
        // holding the maximum logit by equality)
        const float d2l = fmaf(r2.z, W * rr, r2.y);
        E[q] += __builtin_amdgcn_exp2f(d2l);
        const float lt = fmaf(kap, t, d2l);               // log2(e) * (b3 d2 - b2 t) / b3
        Zall[q] += t;
        ls[q][s] = lt;
        tsv[q][s] = t;
        if (SUB > 1) mx[q] = fmaxf(mx[q], lt);
      }
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      // With one record per update the maximum rises exactly when the record's logit exceeds
      // it (a NaN logit compares false, as fmaxf ignores it): one compare instead of a
      // canonicalising fmaxf and a compare per pair.
      const float mq = (SUB == 1) ? ls[q][0] : mx[q];
      if (mq > Mt[q]) {
        const float Mn = b3 * mq;
        const float sc = __builtin_amdgcn_exp2f(M[q] - Mn);
        S[q] *= sc;
        Z[q] *= sc;
        M[q] = Mn;
        Mt[q] = mq;
      }
    }
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        float sv = __builtin_amdgcn_exp2f(fmaf(b3, ls[q][s], -M[q]));
        if (MASK) sv = (tsv[q][s] > 0.0f) ? sv : 0.0f;
        S[q] += sv;
        Z[q] = fmaf(sv, tsv[q][s], Z[q]);
      }
    }
  }
}

struct BlendArgs {
  const Job* jobs;
  int n_jobs;
  const float* rec;
  const float* rec2;
  const uint32_t* tile_np;    // per tile and phase: 1 if a listed pair is not known to be in front
  const unsigned long long* off1;
  const unsigned long long* off2;
  const uint32_t* list;
  float* st;            // per-pixel state, 7 planes of total_pix: M S Z E Zall D Mt.  Zall, the
                        // NaN flag (direct form, in the D plane) and the exact maximum Mt (blend
                        // records) feed the second pass; the backward reads M S Z E D Mt
  int64_t total_pix;
  float* tile_lmin;
  float* tile_seed;     // phase 1 writes the floor seed of each tile (phase-2 membership input)
  const float* tile_ubm;
  Counters* cnt;
  const int32_t* job_flags;
  int save;
  const int32_t* tord;  // dispatch order (block b renders tile tord[b]), or null
};

// Pixels per thread of the blend: two for 16- and 32-pixel tiles (each record read from LDS
// serves two pixels; with one, three 16-byte LDS broadcasts per pair bound the kernel), one
// for 8-pixel tiles (whose 64 pixels fill one wave).
// The list of L records of a tile of a small view is split KS ways: split v blends the
// contiguous range [v h, min(L, (v + 1) h)) of it (h = L/KS rounded up to a multiple of SUB)
// from its own empty state, and the KS states are merged at the end (max, then S and Z
// rescaled to it; E and Zall added, NaN flags or-ed).  Views of kSplitMinTiles 16-pixel tiles
// or more are not split (view_splits_lists): their grids fill the GPU by themselves.  The split
// halves a view's latency where its longest tiles set it (one 128x128 view, N=4096: 0.74 ->
// 0.35 ms; one 480x640 view 0.63 -> 0.44 ms per forward call) and costs 1-2% in a batch that
// fills the GPU many times (the 16-view 480x640 shell's phase-1 blend: the second range
// restarts the online maximum, and the merge).  Whether a tile is split depends on its view
// alone, so a view's result does not depend on the batch around it.  How a launch runs (MODE): 0 when no view of the batch is
// split (the plain kernel), else 2, side by side on KS groups of NT threads, each round staging
// R = GB/KS records of every range (a tile runs on KS times the waves, for grids whose time is
// set by their longest tiles' latency), or 1, one after another on one group in rounds of GB
// records, for a grid that fills the GPU many times over (2-wave blocks, which wait less at the
// round barriers).  An unsplit view's tiles go through every mode with the same operations.
template <int TS, int PHASE>
struct BlendShape {
  static constexpr int PPT = (TS >= 16) ? 2 : 1;
  static constexpr int NT = TS * TS / PPT;   // threads per group (one or two pixels each)
  // splits: as many groups as make blocks of 4 waves (as many as 7 waves per SIMD hold exactly,
  // 28 per CU; 8-wave blocks fit only 3 times, and the phase-1 blend of the 16-view shell batch
  // took 1.10 ms with 4 splits of 2 waves against 0.89 unsplit).  The phase-2 lists are short
  // (its blend is bound by loading and storing the pixel state), so it is not split, nor are
  // 32-pixel tiles (8 waves already).
  static constexpr int KS = (PHASE == 2 || NT >= 256) ? 1 : 256 / NT;
  static constexpr int nb(int mode) { return NT * (mode == 2 ? KS : 1); }   // threads per block
};

// Waves per SIMD the blend of 16-pixel tiles is compiled for.  The phase-2 blend is held to 128
// VGPRs (4 waves instead of 3): its tiles are short and its time is mostly loading and storing
// the per-pixel state, which more resident waves hide better (0.112 -> 0.101 ms).  Phase 1 is
// held to 7, except where LDS caps it first: the 2-wave blocks of larger LDS batches fit 6 or 4
// waves per SIMD (staging, plus the merge buffer in mode 1: 8 and 16 KB for GB = 128 and 256
// unsplit, 14 and 22 KB in mode 1).
template <int TS, int GB, int PHASE, int MODE>
constexpr int blend_waves() {
  if (TS != 16) return 1;
  if (PHASE == 2) return 4;
  if (PHASE == 0) return 1;
  if (MODE == 2 || GB <= 64) return 7;
  if (MODE == 0) return GB == 128 ? 7 : 4;
  return GB == 128 ? 6 : 4;
}

template <int TS, int GB, int PHASE, int MODE>
__global__ __launch_bounds__((BlendShape<TS, PHASE>::nb(MODE)), (blend_waves<TS, GB, PHASE, MODE>()))
void k_blend(BlendArgs A) {
  constexpr int PPT = BlendShape<TS, PHASE>::PPT;
  constexpr int NT = BlendShape<TS, PHASE>::NT;
  constexpr int KS = (MODE == 0) ? 1 : BlendShape<TS, PHASE>::KS;   // splits of a split view
  constexpr int KP = (MODE == 2) ? KS : 1;     // groups side by side
  constexpr int NB = NT * KP;
  // records per online-max update: 2 pairs per thread (one record for threads of two pixels).
  // Fewer pairs in flight need fewer registers (70 and 88 VGPRs for 16-pixel tiles with 2 and
  // 4 pairs instead of 112 and 128 with 8), and the extra resident waves hide the LDS and
  // transcendental latencies (the first pass 6% faster with 2 pairs than with 8).  With one
  // record per update the update is a compare per pixel, with no maximum to take: phase-1
  // blend of the 16-view 480x640 shell 0.90-0.92 -> 0.83-0.84 ms (interleaved runs on one
  // box).  The second pass with one record per update instead of two: the same time on the
  // shell (0.085-0.087 ms), the 256-scene batch 2% faster, a lone 128x128 view's second pass
  // 15% slower (its call 3%).
  constexpr int SUB = 2 / PPT;
  constexpr int R = GB / KP;             // records staged per group and round
  static_assert(GB % (KS * SUB) == 0, "a round's slices fit the LDS batch");
  static_assert((R & (R - 1)) == 0, "slices of a power-of-two size");
  __shared__ float4 lds[GB * 4];
  __shared__ float sh_red[NB / 64 + 1];
  __shared__ float mrg[(KS > 1 ? KS - 1 : 1) * 6][KS > 1 ? NT * PPT : 1];   // splits 1..KS-1: M S Z E Zall nan per pixel
  const int split = (KP > 1) ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / NT) : 0;   // group (wave-uniform)
  const int ti = (int)threadIdx.x - split * NT;    // its pixel thread within the group

  // one block per tile
  const int64_t gt = A.tord ? (int64_t)A.tord[blockIdx.x] : xcd_tile(blockIdx.x, gridDim.x);
  const int ji = find_job(A.n_jobs, gt, [&](int k) { return A.jobs[k].tile_base; });
  const Job& jb = A.jobs[ji];
  const int tile = (int)(gt - jb.tile_base);
  const int tx = tile % jb.tiles_x, ty = tile / jb.tiles_x;
  const int32_t jflags = A.job_flags[ji];
  const bool safe = (jflags & kFlagDegenerate) != 0;   // the view has degenerate records
  const bool r2 = use_rec2(jb, jflags);
  // Pixel-grid tiles with blend records evaluate the forms in coordinates relative to the
  // centre (ox, oy) of their 32x32-pixel block (form_coord), with each record's constant
  // terms moved there when it is staged (localize_rec2): the backward evaluates the same
  // forms with the same operations, so every t*, d2 and weight it recomputes is this pass's
  // value bit for bit.  Ray lists and the direct form use the rays' or pixels' own
  // coordinates.
  const bool loc = r2 && !jb.ray_xy;
  const float ox = form_origin(tx * TS, jb.cx), oy = form_origin(ty * TS, jb.cy);
  // pixel q of this thread: tile pixel index PPT*ti + q, so a thread's pixels are neighbours
  // in one row (same y: the blend records' y-parts are shared between them)
  bool inimg[PPT];
  int32_t lpx[PPT];   // the pixel's index within the view (32 bits: fewer registers live in the loop)
  float xf[PPT], yf[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int i = PPT * ti + q;
    const int px = tx * TS + (i % TS);
    const int py = ty * TS + (i / TS);
    const int64_t lp = (int64_t)py * jb.W + px;
    inimg[q] = (px < jb.W) && (py < jb.H) && (lp < jb.npix);
    lpx[q] = inimg[q] ? (int32_t)lp : 0;
    if (jb.ray_xy) {
      xf[q] = inimg[q] ? jb.ray_xy[2 * lp] : 0.0f;
      yf[q] = inimg[q] ? jb.ray_xy[2 * lp + 1] : 0.0f;
    } else if (loc) {
      xf[q] = form_coord(px);
      yf[q] = form_coord(py);
    } else {
      xf[q] = (float)px - jb.cx;
      yf[q] = (float)py - jb.cy;
    }
  }
  const float b3 = jb.beta3;
  const float nb2l = -jb.beta2 * kLog2e;
  const float kd = -0.5f * kLog2e;
  const float kap = r2 ? nb2l / b3 : 0.0f;
  const bool masked = (PHASE == 0) || !A.tile_np || A.tile_np[2 * gt + (PHASE == 2 ? 1 : 0)] != 0;

  float M[PPT], S[PPT], Z[PPT], E[PPT], Zall[PPT], Mt[PPT];
  bool nanf[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    if (PHASE == 2) {   // (not split)
      M[q] = A.st[jb.pix_base + lpx[q]];
      S[q] = A.st[A.total_pix + jb.pix_base + lpx[q]];
      Z[q] = A.st[2 * A.total_pix + jb.pix_base + lpx[q]];
      E[q] = A.st[3 * A.total_pix + jb.pix_base + lpx[q]];
      Zall[q] = A.st[4 * A.total_pix + jb.pix_base + lpx[q]];
      // a NaN logit seen by the phase-1 pass (kept in the sixth state plane between the passes)
      // (only the NaN-safe blend of a view with degenerate records sets it)
      nanf[q] = safe && A.st[5 * A.total_pix + jb.pix_base + lpx[q]] != 0.0f;
    } else {
      M[q] = -3.402823466e38f; S[q] = 0.f; Z[q] = 0.f; E[q] = 0.f; Zall[q] = 0.f;
      nanf[q] = false;
    }
    Mt[q] = -3.402823466e38f;
  }

  int64_t start, end;
  if (PHASE == 0) { start = 0; end = jb.n; }
  else if (PHASE == 1) { start = (int64_t)A.off1[gt]; end = (int64_t)A.off1[gt + 1]; }
  else { start = (int64_t)A.off2[gt]; end = (int64_t)A.off2[gt + 1]; }
  // a second pass with pairs continues from the first pass's exact maximum (views with blend
  // records; without pairs it changes neither the maximum nor that plane)
  if (PHASE == 2 && r2 && end > start) {
#pragma unroll
    for (int q = 0; q < PPT; ++q) Mt[q] = A.st[6 * A.total_pix + jb.pix_base + lpx[q]];
  }
  const float4* rec4 = reinterpret_cast<const float4*>(A.rec) + jb.rec_base * 4;
  const float4* rec24 = reinterpret_cast<const float4*>(A.rec2) + jb.rec_base * 4;
  // (a tile lists each record of its view at most once: its list length fits 32 bits)
  const int L = (int)(end - start);
  const int ks = (KS > 1 && view_splits_lists(TS, jb.n_tiles)) ? KS : 1;   // this view's splits
  const int h = ((L + ks - 1) / ks + SUB - 1) / SUB * SUB;   // records per split (the last may hold fewer)
  auto range_len = [&](int v) { return min(L, (v + 1) * h) - min(L, v * h); };
  // stage this round's records (P float4 each) of the groups' ranges into their LDS slots,
  // padded with null records to the slot's last SUB group
  auto null_part = [](auto pc, int part) {
    constexpr int P = decltype(pc)::value;
    if (P == 3)   // null blend record: X = 0, Y1 = 1, wL = -inf -> contributes nothing
      return (part == 1) ? make_float4(0.f, 1.f, 0.f, 0.f)
           : (part == 2) ? make_float4(0.f, -INFINITY, 0.f, 0.f)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
    // null record: u = (1,0,0), q = 0, w = -inf -> contributes nothing
    return (part == 1) ? make_float4(0.f, 0.f, 1.f, 0.f)
         : (part == 3) ? make_float4(-INFINITY, -INFINITY, 0.f, 0.f)
                       : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  // a blend record staged with its constant terms at the block's centre (a null record stays
  // null: X = 0, Y1 = 1, wL = -inf)
  auto stage_local = [&](int k, bool real, int64_t gid) {
    float4 a0, a1, a2;
    if (real) {
      a0 = rec24[gid * 4];
      a1 = rec24[gid * 4 + 1];
      a2 = rec24[gid * 4 + 2];
    } else {
      a0 = make_float4(0.f, 0.f, 0.f, 0.f);
      a1 = make_float4(0.f, 1.f, 0.f, 0.f);
      a2 = make_float4(0.f, -INFINITY, 0.f, 0.f);
    }
    localize_rec2(a0, a1, a2, ox, oy);
    lds[k * 3] = a0;
    lds[k * 3 + 1] = a1;
    lds[k * 3 + 2] = a2;
  };
  auto stage = [&](auto pc, int v, int o, int len) {
    constexpr int P = decltype(pc)::value;
    if constexpr (P == 3) {
      if (loc) {   // one record per thread (its three parts together)
        if constexpr (KP == 1) {
          const int cnt = min(R, len - o), nbr = (cnt + SUB - 1) / SUB * SUB;
          const int li0 = v * h + o;
          for (int k = threadIdx.x; k < nbr; k += NB) {
            const bool real = k < cnt;
            const int64_t gid = !real ? 0 : (PHASE == 0) ? (int64_t)(li0 + k) : (int64_t)A.list[start + li0 + k];
            stage_local(k, real, gid);
          }
        } else {
          for (int k = threadIdx.x; k < KP * R; k += NB) {
            const int g = k / R, kk = k - g * R;
            const int cg = min(R, range_len(g) - o);
            if (kk >= (cg + SUB - 1) / SUB * SUB) continue;
            const bool real = kk < cg;
            const int li = g * h + o + kk;
            const int64_t gid = !real ? 0 : (PHASE == 0) ? (int64_t)li : (int64_t)A.list[start + li];
            stage_local(k, real, gid);
          }
        }
        return;
      }
    }
    if constexpr (KP == 1) {
      // one group: records [o, o + cnt) of range v, padded to the last SUB group
      const int cnt = min(R, len - o), nbr = (cnt + SUB - 1) / SUB * SUB;
      const int li0 = v * h + o;
      for (int i = threadIdx.x; i < nbr * P; i += NB) {
        const int k = i / P, part = i - P * k;
        float4 val;
        if (k < cnt) {
          const int64_t gid = (PHASE == 0) ? (int64_t)(li0 + k) : (int64_t)A.list[start + li0 + k];
          val = (P == 3) ? rec24[gid * 4 + part] : rec4[gid * 4 + part];
        } else {
          val = null_part(pc, part);
        }
        lds[i] = val;
      }
    } else {
      for (int i = threadIdx.x; i < KP * R * P; i += NB) {
        const int k = i / P, part = i - P * k;
        const int g = k / R, kk = k - g * R;   // group slot, record within it
        const int cg = min(R, range_len(g) - o);   // the slot's records this round (may be <= 0)
        if (kk >= (cg + SUB - 1) / SUB * SUB) continue;
        float4 val;
        if (kk < cg) {
          const int li = g * h + o + kk;
          const int64_t gid = (PHASE == 0) ? (int64_t)li : (int64_t)A.list[start + li];
          val = (P == 3) ? rec24[gid * 4 + part] : rec4[gid * 4 + part];
        } else {
          val = null_part(pc, part);
        }
        lds[i] = val;
      }
    }
  };
  const int parts = r2 ? 3 : 4;   // float4 per record

  for (int pass = 0; pass < KS / KP; ++pass) {
    // side by side: group `split` blends split `split`; one after another: splits 1, ..., KS-1,
    // then 0, whose state stays in registers for the merge
    const int v = (KP > 1) ? split : (pass + 1) % KS;
    const int len = range_len(v);
    const int span = (KP > 1) ? range_len(0) : len;   // rounds of the block (split 0 is the longest)
    for (int o = 0; o < span; o += R) {
      if (r2) stage(std::integral_constant<int, 3>{}, v, o, len);
      else stage(std::integral_constant<int, 4>{}, v, o, len);
      __syncthreads();
      const int cnt = max(0, min(R, len - o));
      // records this round, padded (block-uniform; made explicit in the split modes of 16- and
      // 32-pixel tiles, whose record loops otherwise count in vector registers: one more VALU
      // per record; 8-pixel tiles measured slower with it)
      int sq = (cnt + SUB - 1) / SUB * SUB;
      if (MODE != 0 && PPT > 1) sq = __builtin_amdgcn_readfirstlane(sq);
      const float4* sl = lds + split * R * parts;   // this group's slot
      if (r2) {
        if (PPT > 1 && !jb.ray_xy) {
          if (masked) blend_batch_rec2<SUB, true, PPT, true>(sl, sq, xf, yf, b3, kap, M, Mt, S, Z, E, Zall);
          else blend_batch_rec2<SUB, false, PPT, true>(sl, sq, xf, yf, b3, kap, M, Mt, S, Z, E, Zall);
        } else {
          if (masked) blend_batch_rec2<SUB, true, PPT, false>(sl, sq, xf, yf, b3, kap, M, Mt, S, Z, E, Zall);
          else blend_batch_rec2<SUB, false, PPT, false>(sl, sq, xf, yf, b3, kap, M, Mt, S, Z, E, Zall);
        }
      } else {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          if (safe) blend_batch<SUB, true>(sl, sq, xf[q], yf[q], b3, nb2l, kd, M[q], S[q], Z[q], E[q], Zall[q], nanf[q]);
          else blend_batch<SUB, false>(sl, sq, xf[q], yf[q], b3, nb2l, kd, M[q], S[q], Z[q], E[q], Zall[q], nanf[q]);
        }
      }
      __syncthreads();
    }
    // splits 1..KS-1 leave their states for the merge (one after another: in this thread's own
    // entries, which it reads back itself, then it starts the next split empty)
    if (KS > 1 && ks > 1 && v > 0) {
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int pi = q * NT + ti;   // consecutive lanes, consecutive words
        float* m = &mrg[(v - 1) * 6][0];
        m[pi] = M[q];
        m[NT * PPT + pi] = S[q];
        m[2 * NT * PPT + pi] = Z[q];
        m[3 * NT * PPT + pi] = E[q];
        m[4 * NT * PPT + pi] = Zall[q];
        m[5 * NT * PPT + pi] = r2 ? Mt[q] : (nanf[q] ? 1.0f : 0.0f);   // (the NaN flag: direct form only)
        M[q] = -3.402823466e38f; S[q] = 0.f; Z[q] = 0.f; E[q] = 0.f; Zall[q] = 0.f;
        nanf[q] = false;
        Mt[q] = -3.402823466e38f;
      }
    }
  }

  // merge the splits' states into split 0: the maximum, S and Z rescaled to it, E, Zall added
  if (KP > 1) __syncthreads();
  if (KS > 1 && ks > 1 && split == 0) {
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int pi = q * NT + ti;
      float Mm = M[q], Mtm = Mt[q];
#pragma unroll
      for (int k = 1; k < KS; ++k) Mm = fmaxf(Mm, mrg[(k - 1) * 6][pi]);
      // equal maxima (infinite ones included) keep their sums as they are
      const float s0 = (M[q] == Mm) ? 1.0f : __builtin_amdgcn_exp2f(M[q] - Mm);
      S[q] *= s0;
      Z[q] *= s0;
#pragma unroll
      for (int k = 1; k < KS; ++k) {
        const float* m = &mrg[(k - 1) * 6][0];
        const float Mk = m[pi];
        const float sk = (Mk == Mm) ? 1.0f : __builtin_amdgcn_exp2f(Mk - Mm);
        S[q] = fmaf(m[NT * PPT + pi], sk, S[q]);
        Z[q] = fmaf(m[2 * NT * PPT + pi], sk, Z[q]);
        E[q] += m[3 * NT * PPT + pi];
        Zall[q] += m[4 * NT * PPT + pi];
        if (r2) Mtm = fmaxf(Mtm, m[5 * NT * PPT + pi]);   // M = b3 Mt rounded is monotone in Mt
        else nanf[q] = nanf[q] || m[5 * NT * PPT + pi] != 0.0f;
      }
      M[q] = Mm;
      Mt[q] = Mtm;
    }
  }

  const unsigned long long ev = (unsigned long long)(end - start) *
                                (unsigned long long)tile_valid(jb, tile, tx, ty, TS);
  if (PHASE == 1) {
    // exact per-tile minimum of log2 sum exp of the unmasked listed logits; a pixel whose
    // phase-1 listed logits are NaN gets -inf (its tile then lists everything in phase 2)
    // The floor seed: the 1e-20 weight floor is already visible at a pixel whose listed max
    // logit M (masked pairs included) lies floor_thr above its log2-sum-exp M + log2 S, and
    // at a pixel with S = 0 any logit above the tile's threshold makes it so; the phase-2
    // membership completes such a tile densely (seed +inf), or starts its bound from M.
    float lm = INFINITY, gap = -INFINITY, m0 = -INFINITY;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (split == 0 && inimg[q]) {
        A.st[jb.pix_base + lpx[q]] = M[q];
        A.st[A.total_pix + jb.pix_base + lpx[q]] = S[q];
        A.st[2 * A.total_pix + jb.pix_base + lpx[q]] = Z[q];
        A.st[3 * A.total_pix + jb.pix_base + lpx[q]] = E[q];
        A.st[4 * A.total_pix + jb.pix_base + lpx[q]] = Zall[q];
        if (safe) A.st[5 * A.total_pix + jb.pix_base + lpx[q]] = nanf[q] ? 1.0f : 0.0f;
        if (r2) A.st[6 * A.total_pix + jb.pix_base + lpx[q]] = Mt[q];
        const bool pos = S[q] > 0.0f && !nanf[q];
        const float l2s = pos ? __log2f(S[q]) : 0.0f;
        lm = fminf(lm, pos ? M[q] + l2s : -INFINITY);
        if (nanf[q]) gap = INFINITY;
        else if (pos) gap = fmaxf(gap, -l2s);
        else m0 = fmaxf(m0, M[q]);
      }
    }
    lm = block_ext<NB>(lm, sh_red);
    gap = block_ext<NB, true>(gap, sh_red);
    m0 = block_ext<NB, true>(m0, sh_red);
    if (threadIdx.x == 0) {
      A.tile_lmin[gt] = lm;
      A.tile_seed[gt] = (gap >= jb.floor_thr) ? INFINITY : m0;
      unsigned long long* c = A.cnt->slot[blockIdx.x % kCountSlots];
      atomicAdd(c, ev);
      atomicAdd(c + 1, ev);
    }
    return;
  }

  // finish (PHASE 0 and 2)
  float ncnt;
  if (PHASE == 0) ncnt = (float)jb.n;
  else ncnt = (float)((A.off1[gt + 1] - A.off1[gt]) + (A.off2[gt + 1] - A.off2[gt]));
  bool fl = false;
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    float D;
    if (ncnt == 0.0f) D = 0.0f;
    else if (nanf[q]) D = Zall[q] / ncnt;
    else D = (Z[q] + 1e-20f * Zall[q]) / (S[q] + 1e-20f * ncnt);
    const float alpha = -expm1f(-E[q]);
    if (split == 0 && inimg[q]) {
      jb.depth[lpx[q]] = D;
      jb.alpha[lpx[q]] = alpha;
      if (A.save) {
        // what the backward reads: M, S, Z, E and, with blend records, the exact maximum Mt
        // (all unchanged by a second pass with no pairs), and D
        if (PHASE == 0 || end > start) {
          A.st[jb.pix_base + lpx[q]] = M[q];
          A.st[A.total_pix + jb.pix_base + lpx[q]] = S[q];
          A.st[2 * A.total_pix + jb.pix_base + lpx[q]] = Z[q];
          A.st[3 * A.total_pix + jb.pix_base + lpx[q]] = E[q];
          if (r2) A.st[6 * A.total_pix + jb.pix_base + lpx[q]] = Mt[q];
        }
        A.st[5 * A.total_pix + jb.pix_base + lpx[q]] = D;
      }
    }
    if (PHASE == 2) {
      // weight floor: 1e-20 * n relative to S can only matter if a dropped pair that may lie
      // behind the camera pushes the true max logit far above M
      const float ubm = A.tile_ubm[gt];
      const float lhs = -43.185f + __log2f((float)max(jb.n, 1)) + fmaxf(0.0f, ubm - M[q]);
      // (a view with degenerate records lists every pair, and so does a tile completed
      // densely: nothing was dropped there and the floor term is exact)
      fl = fl || (split == 0 && inimg[q] && !safe && ncnt < (float)jb.n &&
                  !(lhs < ((S[q] > 0.0f) ? __log2f(S[q]) : -INFINITY)));
    }
  }
  if (PHASE == 2) {
    const int any = __syncthreads_or(fl ? 1 : 0);
    if (threadIdx.x == 0) {
      if (any) atomicAdd(&A.cnt->floor_unresolved, 1ull);
      atomicAdd(&A.cnt->slot[blockIdx.x % kCountSlots][0], ev);
    }
  } else {
    if (threadIdx.x == 0) {
      unsigned long long* c = A.cnt->slot[blockIdx.x % kCountSlots];
      atomicAdd(c, ev);
      atomicAdd(c + 1, ev);
    }
  }
}

// ------------------------------------------------------------------------------------------
// launch helpers (called from the host code in fe_capi.cpp)
// ------------------------------------------------------------------------------------------
}  // namespace fe

#include "fe_launch.h"

namespace fe {

// CUs of the current device (cached per device; 256 if the query fails)
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cache[dev].load(std::memory_order_relaxed);
  if (c <= 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cache[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

hipError_t launch_init(hipStream_t s, Counters* cnt, int32_t* job_flags, int n_jobs) {
  const int blocks = max(1, min(64, (n_jobs + 255) / 256));
  hipLaunchKernelGGL(k_init, dim3(blocks), dim3(256), 0, s, cnt, job_flags, n_jobs);
  return hipGetLastError();
}

hipError_t launch_pose(hipStream_t s, Job* jobs, int n_jobs) {
  hipLaunchKernelGGL(k_pose, dim3((n_jobs + 63) / 64), dim3(64), 0, s, jobs, n_jobs);
  return hipGetLastError();
}

hipError_t launch_ray_rects(hipStream_t s, const Job* jobs, int n_jobs, int64_t total_tiles, int ts,
                            int32_t* job_flags) {
  if (total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ray_rects, dim3((unsigned)total_tiles), dim3(256), 0, s, jobs, n_jobs, ts, job_flags);
  return hipGetLastError();
}

hipError_t launch_order(hipStream_t s, const Job* jobs, int n_jobs, int max_n, Counters* cnt, int32_t* job_flags) {
  if (max_n <= 0) return launch_init(s, cnt, job_flags, n_jobs);
  hipLaunchKernelGGL(k_order, dim3((max_n + kOrderChunk - 1) / kOrderChunk, n_jobs), dim3(kOrderThreads), 0, s,
                     jobs, cnt, job_flags);
  return hipGetLastError();
}

hipError_t launch_preprocess(hipStream_t s, const Job* jobs, int n_jobs, int max_n, float* rec, float* rec2,
                             int32_t* job_flags) {
  if (max_n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_preprocess, dim3((max_n + 255) / 256, n_jobs), dim3(256), 0, s, jobs, rec, rec2,
                     job_flags);
  return hipGetLastError();
}

hipError_t launch_tile_mask(hipStream_t s, int phase, const Job* jobs, int n_jobs, int64_t total_stiles,
                            int ts, int max_n, const float* rec, unsigned long long* mask1, unsigned long long* mask2,
                            uint32_t* count, const float* tile_lmin, const float* tile_seed, float* tile_ubm,
                            uint32_t* tile_np, const int32_t* job_flags, Counters* cnt, uint32_t* mpart) {
  if (total_stiles <= 0) return hipSuccess;
  // A grid of few groups leaves most of the GPU idle while each block walks all n Gaussians:
  // up to 4 blocks per CU then share each group's Gaussians, in chunks of at least 4 words (256
  // Gaussians), and k_mask_finish merges their partials.  (One-tile groups for small views
  // instead: one 128x128 view, N=4096, 235 -> 49 us for both phases, but the 256-scene batch
  // 8% slower; chunks leave the groups, and the lists, as they are.)
  int chunks = 1;
  if (total_stiles <= kMaskChunkGroups) {
    const int64_t want = (int64_t)4 * device_cus() / total_stiles;
    const int64_t cap = std::min<int64_t>(kMaskChunkBlocks / total_stiles, std::max(1, ((max_n + 63) / 64) / 4));
    chunks = (int)std::max<int64_t>(1, std::min(want, cap));
  }
  // the group factor is 4 for 8-pixel tiles, 3 for 16, 1 for 32 and for ray lists
  const dim3 g((unsigned)total_stiles, (unsigned)chunks), b(256);
  // (the chunks' lower bounds follow the partials in the region)
  float* lbpart = reinterpret_cast<float*>(mpart + 4 * kMaskMaxSub * kMaskChunkBlocks);
#define FE_MASK_ARGS jobs, n_jobs, ts, rec, mask1, mask2, count, tile_lmin, tile_seed, tile_ubm, tile_np, job_flags, cnt, \
                     chunks, mpart, lbpart
#define FE_FINISH_ARGS jobs, n_jobs, chunks, mpart, mask1, mask2, count, tile_seed, tile_ubm, tile_np, cnt
  if (phase == 1) {
    if (ts == 8) hipLaunchKernelGGL((k_tile_mask<1, 4>), g, b, 0, s, FE_MASK_ARGS);
    else if (ts == 16) hipLaunchKernelGGL((k_tile_mask<1, 3>), g, b, 0, s, FE_MASK_ARGS);
    else hipLaunchKernelGGL((k_tile_mask<1, 1>), g, b, 0, s, FE_MASK_ARGS);
    if (chunks > 1) hipLaunchKernelGGL(k_mask_finish<1>, dim3((unsigned)total_stiles), dim3(64), 0, s, FE_FINISH_ARGS);
  } else {
    if (chunks > 1) {
      if (ts == 8) hipLaunchKernelGGL(k_mask_lb<4>, g, b, 0, s, jobs, n_jobs, ts, rec, tile_lmin, job_flags, chunks, lbpart);
      else if (ts == 16) hipLaunchKernelGGL(k_mask_lb<3>, g, b, 0, s, jobs, n_jobs, ts, rec, tile_lmin, job_flags, chunks, lbpart);
      else hipLaunchKernelGGL(k_mask_lb<1>, g, b, 0, s, jobs, n_jobs, ts, rec, tile_lmin, job_flags, chunks, lbpart);
    }
    if (ts == 8) hipLaunchKernelGGL((k_tile_mask<2, 4>), g, b, 0, s, FE_MASK_ARGS);
    else if (ts == 16) hipLaunchKernelGGL((k_tile_mask<2, 3>), g, b, 0, s, FE_MASK_ARGS);
    else hipLaunchKernelGGL((k_tile_mask<2, 1>), g, b, 0, s, FE_MASK_ARGS);
    if (chunks > 1) hipLaunchKernelGGL(k_mask_finish<2>, dim3((unsigned)total_stiles), dim3(64), 0, s, FE_FINISH_ARGS);
  }
#undef FE_MASK_ARGS
#undef FE_FINISH_ARGS
  return hipGetLastError();
}

int64_t scan_partial_count(int64_t n) { return (n + kScanBlock - 1) / kScanBlock; }

hipError_t launch_scan(hipStream_t s, const uint32_t* cnt, int64_t n, unsigned long long* partial,
                       unsigned long long* off, const unsigned long long* base_src,
                       const unsigned long long* key_off, uint32_t* hist, int32_t* tord) {
  if ((hist == nullptr) != (tord == nullptr)) return hipErrorInvalidValue;
  const int64_t nb = scan_partial_count(n);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(256), 0, s, cnt, n, partial, key_off, hist);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, partial, nb, base_src, hist);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(256), 0, s, cnt, n, partial, off, key_off, hist, tord);
  return hipGetLastError();
}

hipError_t launch_tile_write(hipStream_t s, const Job* jobs, int n_jobs, int64_t total_tiles,
                             const unsigned long long* mask, const unsigned long long* off,
                             uint32_t* list) {
  hipLaunchKernelGGL(k_tile_write, dim3((unsigned)((total_tiles + 3) / 4)), dim3(256), 0, s, jobs, n_jobs,
                     total_tiles, mask, off, list);
  return hipGetLastError();
}

// Where a split tile list runs (the same operations either way).  Side by side halves a block's
// latency, which pays wherever the grid's last blocks set the time: measured phase-1 blends of
// N=4096 views, side by side against one after another, 16-pixel tiles: one 128x128 view 0.35
// against 0.74 ms, 64 such views 0.41 against 0.74, 8 480x640 views 0.43 against 0.44, and the
// 16-view 480x640 shell (the grid 11 times what the GPU holds at once) 0.95 against 0.92 under
// sustained load.  8-pixel tiles run their 4 one-wave groups side by side always (the 16-view
// shell: 0.88 against 0.91 ms).
template <int TS, int PHASE>
static bool blend_side_by_side(int64_t total_tiles) {
  if (TS == 8) return true;
  constexpr int waves = BlendShape<TS, PHASE>::nb(2) / 64;
  return total_tiles * waves <= (int64_t)8 * device_cus() * 28;   // 28 waves per CU at once
}

template <int TS, int GB, int PHASE>
static void launch_blend_p(hipStream_t s, int64_t total_tiles, bool any_split, const BlendArgs& a) {
  const dim3 grid((unsigned)total_tiles);
  if constexpr (BlendShape<TS, PHASE>::KS > 1) {
    if (any_split) {
      if (blend_side_by_side<TS, PHASE>(total_tiles))
        hipLaunchKernelGGL((k_blend<TS, GB, PHASE, 2>), grid, dim3(BlendShape<TS, PHASE>::nb(2)), 0, s, a);
      else
        hipLaunchKernelGGL((k_blend<TS, GB, PHASE, 1>), grid, dim3(BlendShape<TS, PHASE>::nb(1)), 0, s, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_blend<TS, GB, PHASE, 0>), grid, dim3(BlendShape<TS, PHASE>::nb(0)), 0, s, a);
}

template <int TS, int GB>
static hipError_t launch_blend_tg(hipStream_t s, int phase, int64_t total_tiles, bool any_split,
                                  const BlendArgs& a) {
  if (phase == 0) launch_blend_p<TS, GB, 0>(s, total_tiles, any_split, a);
  else if (phase == 1) launch_blend_p<TS, GB, 1>(s, total_tiles, any_split, a);
  else launch_blend_p<TS, GB, 2>(s, total_tiles, any_split, a);
  return hipGetLastError();
}

template <int TS>
static hipError_t launch_blend_t(hipStream_t s, int gb, int phase, int64_t total_tiles, bool any_split,
                                 const BlendArgs& a) {
  switch (gb) {
    case 32: return launch_blend_tg<TS, 32>(s, phase, total_tiles, any_split, a);
    case 64: return launch_blend_tg<TS, 64>(s, phase, total_tiles, any_split, a);
    case 128: return launch_blend_tg<TS, 128>(s, phase, total_tiles, any_split, a);
    case 256: return launch_blend_tg<TS, 256>(s, phase, total_tiles, any_split, a);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_blend(hipStream_t s, int ts, int gb, int phase, int64_t total_tiles, const Job* jobs,
                        int n_jobs, const float* rec, const float* rec2, const uint32_t* tile_np,
                        const unsigned long long* off1,
                        const unsigned long long* off2, const uint32_t* list, float* st, int64_t total_pix,
                        float* tile_lmin, float* tile_seed, const float* tile_ubm, Counters* cnt,
                        const int32_t* job_flags, int save, const int32_t* tord, int split_views) {
  BlendArgs a{jobs, n_jobs, rec, rec2, tile_np, off1, off2, list, st, total_pix, tile_lmin, tile_seed, tile_ubm, cnt,
              job_flags, save, tord};
  switch (ts) {
    case 8: return launch_blend_t<8>(s, gb, phase, total_tiles, split_views > 0, a);
    case 16: return launch_blend_t<16>(s, gb, phase, total_tiles, split_views > 0, a);
    case 32: return launch_blend_t<32>(s, gb, phase, total_tiles, split_views > 0, a);
  }
  return hipErrorInvalidValue;
}

}  // namespace fe
