// C ABI of the MI355X fuzzy-metaball rasterizer: planning, forward and backward entry points.
//
// The host side only does O(views) arithmetic: it lays out one caller-provided workspace
// (sized for the worst case so nothing is ever read back), uploads the job descriptors and
// enqueues the kernels on the caller's stream.  No host synchronisation on the render path.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/fuzzy_ellipses_amd.h"
#include "fe_common.h"
#include "fe_launch.h"

namespace fe {

static thread_local char g_err[512] = "";

static void set_err(const char* what, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
}
static void set_msg(const char* what) { snprintf(g_err, sizeof(g_err), "%s", what); }
void set_error_message(const char* msg) { set_msg(msg); }

#define FE_HIP(call)                         \
  do {                                       \
    hipError_t e_ = (call);                  \
    if (e_ != hipSuccess) {                  \
      ::fe::set_err(#call, e_);              \
      return FE_ERR_HIP;                     \
    }                                        \
  } while (0)

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Workspace layout.  Every region is 256-byte aligned.
struct Layout {
  int n_jobs = 0;
  int max_n = 0;
  int has_rays = 0;
  int ts = 16, gb = 64, mode = FE_MODE_CULLED, save = 0, reuse = 0;
  int feat_chunks = 0;   // blocks per tile of the feature kernel (0: no view wants features)
  int split_views = 0;   // views whose tile lists the blend splits
  int64_t total_stiles = 0;
  int64_t total_rec = 0, total_tiles = 0, total_mask = 0, total_pix = 0, list_cap = 0;
  size_t o_header = 0, o_mask2 = 0, o_jobs = 0, o_counters = 0, o_flags = 0, o_rec = 0, o_rec2 = 0, o_tnp = 0,
         o_mask = 0, o_count = 0,
         o_off1 = 0, o_off2 = 0, o_partial = 0, o_list = 0, o_lmin = 0, o_ubm = 0, o_state = 0,
         o_bjobs = 0, o_gsum = 0, o_rect = 0, o_order = 0, o_rank = 0, o_tord = 0, o_mpart = 0, bytes = 0;
};

// A small header at the start of the workspace lets fe_read_stats find the counters.
struct WsHeader {
  uint64_t magic;
  uint64_t o_counters, o_off1, o_off2, total_tiles, mode;
  uint64_t o_rec, o_list, o_lmin, total_rec, o_state, total_pix;
  uint64_t o_order, o_tord;
};
constexpr uint64_t kMagic = 0x4645414d44313031ull;  // "FEAMD101"

static fe_status plan(const fe_scene* scenes, int32_t n_scenes, const fe_view* views, int32_t n_views,
                      const fe_options* opt, Layout* L, Job* jobs) {
  if (!scenes || !views || n_scenes <= 0 || n_views <= 0) {
    set_msg("fe_plan: need at least one scene and one view");
    return FE_ERR_INVALID_ARGUMENT;
  }
  fe_options o = {16, 64, FE_MODE_CULLED, 0, 0, 0.0f, 0.0f};
  if (opt) o = *opt;
  if (o.tile == 0) o.tile = 16;
  if (!(o.alpha_eps > 0.0f)) o.alpha_eps = 2e-5f;
  if (!(o.depth_eps > 0.0f)) o.depth_eps = 1e-6f;
  if (!(o.alpha_eps < 1.0f) || !(o.depth_eps < 1.0f)) {
    set_msg("fe_plan: alpha_eps and depth_eps must be in (0, 1)");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (o.batch == 0) o.batch = 64;
  if (o.tile != 8 && o.tile != 16 && o.tile != 32) {
    set_msg("fe_plan: tile must be 8, 16 or 32");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (o.batch != 32 && o.batch != 64 && o.batch != 128 && o.batch != 256) {
    set_msg("fe_plan: batch must be 32, 64, 128 or 256");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (o.mode != FE_MODE_CULLED && o.mode != FE_MODE_DENSE) {
    set_msg("fe_plan: unknown mode");
    return FE_ERR_INVALID_ARGUMENT;
  }
  L->ts = o.tile;
  L->gb = o.batch;
  L->mode = o.mode;
  L->save = o.save_for_backward ? 1 : 0;
  L->reuse = (o.flags & FE_FLAG_REUSE_DESCRIPTORS) ? 1 : 0;
  L->n_jobs = n_views;
  int64_t rec = 0, tiles = 0, stiles = 0, maskw = 0, pix = 0, cap = 0;
  // membership pre-test groups: 32x32 pixels for 8-pixel tiles, 48x48 for 16 (measured against
  // 32x32: membership 18% (phase 1) and 13% (phase 2) faster; 64x64 slows phase 2, whose 16
  // sub-tile accumulators cost occupancy), the tile itself for 32
  const int sf = (o.tile == 8) ? 4 : (o.tile == 16 ? 3 : 1);
  int max_n = 0;
  for (int v = 0; v < n_views; ++v) {
    const fe_view& vw = views[v];
    if (vw.scene < 0 || vw.scene >= n_scenes) {
      set_msg("fe_plan: view references a scene out of range");
      return FE_ERR_INVALID_ARGUMENT;
    }
    const fe_scene& sc = scenes[vw.scene];
    const fe_camera& c = vw.cam;
    const bool ray_view = vw.rays != nullptr;
    if (ray_view) {
      if (sc.n < 0 || vw.n_rays <= 0 || vw.n_rays > ((int64_t)1 << 30)) {
        set_msg("fe_plan: a ray-list view needs 1 <= n_rays <= 2^30");
        return FE_ERR_INVALID_ARGUMENT;
      }
    } else if (sc.n < 0 || c.height <= 0 || c.width <= 0 || !(c.fx != 0.0f) || !(c.fy != 0.0f)) {
      set_msg("fe_plan: invalid scene size or camera");
      return FE_ERR_INVALID_ARGUMENT;
    }
    if ((!ray_view && (int64_t)c.height * c.width > (int64_t)1 << 30) || sc.n > (1 << 26)) {
      set_msg("fe_plan: image or scene too large");
      return FE_ERR_INVALID_ARGUMENT;
    }
    if (sc.n_features < 0 || sc.n_features > (1 << 16) || (sc.n_features > 0 && sc.n > 0 && !sc.features)) {
      set_msg("fe_plan: a scene's features need 0 <= n_features <= 65536 and a feature array");
      return FE_ERR_INVALID_ARGUMENT;
    }
    const bool want_feat = sc.n_features > 0 && sc.n > 0 && vw.features_out != nullptr;
    if (want_feat) L->feat_chunks = std::max(L->feat_chunks, feat_chunks(sc.n_features));
    // a ray list is laid out as an image one tile wide: tile t holds rays [t*ts*ts, (t+1)*ts*ts)
    const int64_t tt = (int64_t)o.tile * o.tile;
    const int Wv = ray_view ? o.tile : c.width;
    const int Hv = ray_view ? (int)((vw.n_rays + o.tile - 1) / o.tile) : c.height;
    const int64_t npix = ray_view ? vw.n_rays : (int64_t)c.height * c.width;
    const int tx = ray_view ? 1 : (c.width + o.tile - 1) / o.tile;
    const int ty = ray_view ? (int)((vw.n_rays + tt - 1) / tt) : (c.height + o.tile - 1) / o.tile;
    const int sfv = ray_view ? 1 : sf;
    if (ray_view) L->has_rays = 1;
    const int nw = (sc.n + 63) / 64;
    if (jobs) {
      Job& jb = jobs[v];
      memset(&jb, 0, sizeof(Job));
      jb.means = sc.means;
      jb.prec = sc.prec;
      jb.wlog = sc.weights_log;
      jb.depth = vw.depth;
      jb.alpha = vw.alpha;
      jb.pose = vw.pose_device;
      memcpy(jb.R, c.R, sizeof(jb.R));
      memcpy(jb.o, c.t, sizeof(jb.o));
      if (ray_view) {
        jb.fx = 1.0f; jb.fy = 1.0f; jb.cx = 0.0f; jb.cy = 0.0f;
        jb.ray_xy = vw.rays;
      } else {
        jb.fx = c.fx; jb.fy = c.fy; jb.cx = c.cx; jb.cy = c.cy;
      }
      jb.npix = npix;
      if (want_feat) {
        jb.feat = sc.features;
        jb.feat_out = vw.features_out;
        jb.n_feat = sc.n_features;
      }
      jb.beta2 = vw.beta2; jb.beta3 = vw.beta3;
      const double ln_n = log((double)std::max(sc.n, 1));
      // alpha: total dropped density <= n * exp(-c_alpha) = alpha_eps
      jb.c_alpha = (float)(ln_n - log((double)o.alpha_eps));
      // depth: dropped softmax weight fraction <= n * exp(-delta) = depth_eps
      jb.delta2 = (float)((ln_n - log((double)o.depth_eps)) * 1.4426950408889634);
      jb.floor_thr = (float)(13.0 * 3.321928094887362 - ln_n * 1.4426950408889634);
      jb.n = sc.n; jb.H = Hv; jb.W = Wv;
      jb.tiles_x = tx; jb.tiles_y = ty; jb.n_tiles = tx * ty;
      jb.nw = nw;
      jb.sf = sfv;
      jb.stiles_x = (tx + sfv - 1) / sfv;
      jb.stiles_y = (ty + sfv - 1) / sfv;
      jb.stile_base = stiles;
      jb.scene = vw.scene;
      jb.rec_base = rec;
      jb.tile_base = tiles;
      jb.mask_base = maskw;
      jb.pix_base = pix;
    }
    max_n = std::max(max_n, sc.n);
    if (view_splits_lists(o.tile, (int64_t)tx * ty)) ++L->split_views;
    rec += sc.n;
    tiles += (int64_t)tx * ty;
    stiles += (int64_t)((tx + sfv - 1) / sfv) * ((ty + sfv - 1) / sfv);
    maskw += (int64_t)tx * ty * nw;
    pix += npix;
    cap += (int64_t)tx * ty * sc.n;
  }
  if (cap > ((int64_t)1 << 34)) {
    set_msg("fe_plan: batch too large (tile-pair capacity above 2^34)");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (tiles >= ((int64_t)1 << 31) || stiles >= ((int64_t)1 << 31)) {
    // one block per tile, and the tile dispatch order holds int32 tile indices
    set_msg("fe_plan: batch too large (2^31 tiles or more)");
    return FE_ERR_INVALID_ARGUMENT;
  }
  L->max_n = max_n;
  L->total_rec = rec;
  L->total_tiles = tiles;
  L->total_stiles = stiles;
  L->total_mask = maskw;
  L->total_pix = pix;
  L->list_cap = cap;
  size_t p = 0;
  L->o_header = p; p = align256(p + sizeof(WsHeader));
  L->o_jobs = p; p = align256(p + sizeof(Job) * n_views);
  L->o_counters = p; p = align256(p + sizeof(Counters));
  L->o_flags = p; p = align256(p + sizeof(int32_t) * n_views);
  L->o_rec = p; p = align256(p + sizeof(float) * kRecFloats * (size_t)std::max<int64_t>(rec, 1));
  L->o_rec2 = p; p = align256(p + sizeof(float) * kRec2Floats * (size_t)std::max<int64_t>(rec, 1));
  L->o_order = p; p = align256(p + sizeof(int32_t) * (size_t)std::max<int64_t>(rec, 1));
  L->o_rank = p; p = align256(p + sizeof(int32_t) * (size_t)std::max<int64_t>(rec, 1));
  L->o_lmin = p; p = align256(p + sizeof(float) * 2 * tiles);   // thresholds, then floor seeds
  L->o_ubm = p; p = align256(p + sizeof(float) * tiles);
  L->o_state = p; p = align256(p + sizeof(float) * 7 * pix);   // M S Z E Zall D Mt
  if (o.mode == FE_MODE_CULLED) {
    L->o_mask = p; p = align256(p + sizeof(unsigned long long) * std::max<int64_t>(maskw, 1));
    L->o_mask2 = p; p = align256(p + sizeof(unsigned long long) * std::max<int64_t>(maskw, 1));
    L->o_count = p; p = align256(p + sizeof(uint32_t) * tiles);
    L->o_tnp = p; p = align256(p + sizeof(uint32_t) * 2 * tiles);
    L->o_off1 = p; p = align256(p + sizeof(unsigned long long) * (tiles + 1));
    L->o_off2 = p; p = align256(p + sizeof(unsigned long long) * (tiles + 1));
    L->o_partial = p; p = align256(p + sizeof(unsigned long long) * (scan_partial_count(tiles) + 1));
    L->o_tord = p; p = align256(p + sizeof(int32_t) * 2 * tiles);   // phase-1 order, then union order
    L->o_mpart = p; p = align256(p + sizeof(uint32_t) * 5 * kMaskMaxSub * kMaskChunkBlocks);   // membership partials, bounds
    L->o_list = p; p = align256(p + sizeof(uint32_t) * (size_t)std::max<int64_t>(cap, 1));
  }
  if (L->save) {
    L->o_bjobs = p; p = align256(p + sizeof(BwdJob) * n_views);
    L->o_gsum = p; p = align256(p + sizeof(float) * bwd_gsum_floats() * (size_t)std::max<int64_t>(rec, 1));
  }
  if (L->has_rays) {
    L->o_rect = p; p = align256(p + sizeof(float) * 4 * tiles);
  }
  L->bytes = p;
  return FE_OK;
}

// Pinned staging for descriptor uploads: a ring of slots per device.  Before a slot is
// overwritten, the upload that last used it must have completed (normally long ago).
struct Slot {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};
constexpr int kSlots = 8;
struct Staging {
  std::mutex mu;
  Slot slot[kSlots];
  int next = 0;
};
static Staging g_staging[64];

static fe_status upload(hipStream_t s, void* dst, const void* src, size_t bytes) {
  int dev = 0;
  FE_HIP(hipGetDevice(&dev));
  Staging& stg = g_staging[dev & 63];
  std::lock_guard<std::mutex> lk(stg.mu);
  Slot& st = stg.slot[stg.next];
  stg.next = (stg.next + 1) % kSlots;
  if (st.pending) {
    FE_HIP(hipEventSynchronize(st.ev));
    st.pending = false;
  }
  if (!st.ev) FE_HIP(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
  if (st.cap < bytes) {
    if (st.host) FE_HIP(hipHostFree(st.host));
    st.host = nullptr;
    size_t c = std::max<size_t>(bytes, 1 << 16);
    FE_HIP(hipHostMalloc(&st.host, c, hipHostMallocDefault));
    st.cap = c;
  }
  memcpy(st.host, src, bytes);
  FE_HIP(hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, s));
  FE_HIP(hipEventRecord(st.ev, s));
  st.pending = true;
  return FE_OK;
}

// Optional per-kernel event timing.
struct Profiler {
  std::mutex mu;
  uint32_t mask = 0;   // kernels (bit 1 << FE_K_*) whose launches are timed
  std::vector<hipEvent_t> pool;
  struct Rec { int k; hipEvent_t a, b; };
  std::vector<Rec> pending;
  double ms[FE_K_COUNT] = {};
  int64_t launches[FE_K_COUNT] = {};
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};
static Profiler g_prof;

// Runs `launch` between two recorded events when kernel k is being timed.
template <typename F>
static hipError_t timed(hipStream_t s, int k, F launch, bool may_time = true) {
  if (!may_time || !((g_prof.mask >> k) & 1u)) return launch();
  hipEvent_t a, b;
  {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    a = g_prof.get();
    b = g_prof.get();
  }
  if (!a || !b) return launch();
  hipError_t e = hipEventRecord(a, s);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, s);
  std::lock_guard<std::mutex> lk(g_prof.mu);
  g_prof.pending.push_back({k, a, b});
  return e;
}

// Descriptors of a call that allocates nothing, passed by value as a kernel argument and
// copied into the workspace (one view's header and job fit in 1 KB).
struct ArgBlob {
  unsigned char b[1024];
};
static_assert(256 + ((sizeof(Job) + 255) / 256) * 256 <= sizeof(ArgBlob), "one view's header and job fit the blob");
__global__ void k_put_blob(ArgBlob blob, int bytes, unsigned char* dst) {
  for (int i = threadIdx.x; i < bytes; i += blockDim.x) dst[i] = blob.b[i];
}

template <typename T>
static T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(ws) + off);
}

// The forward of one call.  `jobs` has room for n_views descriptors.  With `noalloc` the call
// allocates nothing: the descriptors travel to the workspace as a kernel argument instead of
// through the pinned upload ring, and no launch is event-timed.
static fe_status forward_impl(hipStream_t s, const fe_scene* scenes, int32_t n_scenes, const fe_view* views,
                              int32_t n_views, const fe_options* opt, void* ws, size_t ws_bytes, Job* jobs,
                              bool noalloc) {
  Layout L;
  fe_status st = plan(scenes, n_scenes, views, n_views, opt, &L, jobs);
  if (st != FE_OK) return st;
  if (!ws || ws_bytes < L.bytes) {
    set_msg("fe_forward: workspace missing or too small (see fe_plan)");
    return FE_ERR_WORKSPACE;
  }
  for (int v = 0; v < n_views; ++v) {
    if (!views[v].depth || !views[v].alpha) {
      set_msg("fe_forward: null output buffer");
      return FE_ERR_INVALID_ARGUMENT;
    }
    const fe_scene& sc = scenes[views[v].scene];
    if (sc.n > 0 && (!sc.means || !sc.prec || !sc.weights_log)) {
      set_msg("fe_forward: null scene buffer");
      return FE_ERR_INVALID_ARGUMENT;
    }
  }
  for (int v = 0; v < n_views; ++v) {
    if (jobs[v].ray_xy) jobs[v].rect = at<float>(ws, L.o_rect) + 4 * jobs[v].tile_base;
    jobs[v].order = at<int32_t>(ws, L.o_order) + jobs[v].rec_base;
    jobs[v].rank = at<int32_t>(ws, L.o_rank) + jobs[v].rec_base;
  }
  // header + jobs in one upload
  const WsHeader h{kMagic, L.o_counters, L.o_off1, L.o_off2, (uint64_t)L.total_tiles, (uint64_t)L.mode,
                   L.o_rec, L.o_list, L.o_lmin, (uint64_t)L.total_rec, L.o_state, (uint64_t)L.total_pix,
                   L.o_order, L.o_tord};
  if (!L.reuse && noalloc) {
    if (L.o_counters > sizeof(ArgBlob)) {
      set_msg("forward: too many views for descriptors passed as a kernel argument");
      return FE_ERR_INVALID_ARGUMENT;
    }
    ArgBlob blob;
    memset(&blob, 0, sizeof(blob));
    memcpy(blob.b + L.o_header, &h, sizeof(h));
    memcpy(blob.b + L.o_jobs, jobs, sizeof(Job) * n_views);
    hipLaunchKernelGGL(k_put_blob, dim3(1), dim3(256), 0, s, blob, (int)L.o_counters,
                       static_cast<unsigned char*>(ws));
    FE_HIP(hipGetLastError());
  } else if (!L.reuse) {
    std::vector<char> blob(L.o_counters);
    memcpy(blob.data() + L.o_header, &h, sizeof(h));
    memcpy(blob.data() + L.o_jobs, jobs, sizeof(Job) * n_views);
    st = upload(s, ws, blob.data(), blob.size());
    if (st != FE_OK) return st;
  }
  const bool may_time = !noalloc;

  const Job* dj = at<Job>(ws, L.o_jobs);
  Counters* cnt = at<Counters>(ws, L.o_counters);
  int32_t* flags = at<int32_t>(ws, L.o_flags);
  float* rec = at<float>(ws, L.o_rec);
  float* rec2 = at<float>(ws, L.o_rec2);
  float* state = at<float>(ws, L.o_state);
  float* lmin = at<float>(ws, L.o_lmin);
  float* seed = lmin + L.total_tiles;   // per-tile floor seed, after the thresholds
  float* ubm = at<float>(ws, L.o_ubm);
  bool any_pose = false;
  for (int v = 0; v < n_views; ++v) any_pose = any_pose || views[v].pose_device != nullptr;
  if (any_pose) FE_HIP(launch_pose(s, at<Job>(ws, L.o_jobs), L.n_jobs));
  FE_HIP(timed(s, FE_K_PREPROCESS, [&] {
    const hipError_t e = launch_order(s, dj, L.n_jobs, L.max_n, cnt, flags);   // clears cnt, flags
    return e != hipSuccess ? e : launch_preprocess(s, dj, L.n_jobs, L.max_n, rec, rec2, flags);
  }, may_time));
  if (L.has_rays) FE_HIP(launch_ray_rects(s, dj, L.n_jobs, L.total_tiles, L.ts, flags));
  if (L.mode == FE_MODE_DENSE) {
    FE_HIP(timed(s, FE_K_BLEND_DENSE, [&] {
      return launch_blend(s, L.ts, L.gb, 0, L.total_tiles, dj, L.n_jobs, rec, rec2, nullptr, nullptr, nullptr,
                          nullptr, state, L.total_pix, lmin, seed, ubm, cnt, flags, L.save, nullptr, L.split_views);
    }, may_time));
    if (L.feat_chunks > 0)
      FE_HIP(launch_feat(s, false, L.ts, L.total_tiles, L.feat_chunks, dj, nullptr, L.n_jobs, rec, rec2, nullptr,
                         nullptr, nullptr, state, L.total_pix, flags, nullptr, L.total_rec, 1));
    return FE_OK;
  }
  unsigned long long* mask = at<unsigned long long>(ws, L.o_mask);
  unsigned long long* mask2 = at<unsigned long long>(ws, L.o_mask2);
  uint32_t* count = at<uint32_t>(ws, L.o_count);
  unsigned long long* off1 = at<unsigned long long>(ws, L.o_off1);
  unsigned long long* off2 = at<unsigned long long>(ws, L.o_off2);
  unsigned long long* partial = at<unsigned long long>(ws, L.o_partial);
  uint32_t* list = at<uint32_t>(ws, L.o_list);
  uint32_t* tnp = at<uint32_t>(ws, L.o_tnp);
  int32_t* tord = at<int32_t>(ws, L.o_tord);   // [0, T): phase-1 blend order; [T, 2T): backward order
  uint32_t* thist = at<uint32_t>(ws, L.o_counters + offsetof(Counters, tile_hist));
  uint32_t* mpart = at<uint32_t>(ws, L.o_mpart);
  // phase 1
  FE_HIP(timed(s, FE_K_MASK1, [&] {
    return launch_tile_mask(s, 1, dj, L.n_jobs, L.total_stiles, L.ts, L.max_n, rec, mask, mask2, count, lmin, seed,
                            ubm, tnp, flags, cnt, mpart);
  }, may_time));
  FE_HIP(timed(s, FE_K_SCAN1, [&] {
    return launch_scan(s, count, L.total_tiles, partial, off1, nullptr, nullptr, thist, tord);
  }, may_time));
  FE_HIP(timed(
      s, FE_K_WRITE1, [&] { return launch_tile_write(s, dj, L.n_jobs, L.total_tiles, mask, off1, list); }, may_time));
  FE_HIP(timed(s, FE_K_BLEND1, [&] {
    return launch_blend(s, L.ts, L.gb, 1, L.total_tiles, dj, L.n_jobs, rec, rec2, tnp, off1, off2, list, state,
                        L.total_pix, lmin, seed, ubm, cnt, flags, L.save, tord, L.split_views);
  }, may_time));
  // phase 2
  FE_HIP(timed(s, FE_K_MASK2, [&] {
    return launch_tile_mask(s, 2, dj, L.n_jobs, L.total_stiles, L.ts, L.max_n, rec, mask, mask2, count, lmin, seed,
                            ubm, tnp, flags, cnt, mpart);
  }, may_time));
  FE_HIP(timed(s, FE_K_SCAN2, [&] {
    // the backward's order: both lists of a tile (phase-1 length from off1)
    return launch_scan(s, count, L.total_tiles, partial, off2, off1 + L.total_tiles, off1,
                       thist + kTileBuckets, tord + L.total_tiles);
  }, may_time));
  FE_HIP(timed(
      s, FE_K_WRITE2, [&] { return launch_tile_write(s, dj, L.n_jobs, L.total_tiles, mask2, off2, list); }, may_time));
  FE_HIP(timed(s, FE_K_BLEND2, [&] {
    return launch_blend(s, L.ts, L.gb, 2, L.total_tiles, dj, L.n_jobs, rec, rec2, tnp, off1, off2, list, state,
                        L.total_pix, lmin, seed, ubm, cnt, flags, L.save, nullptr, L.split_views);
  }, may_time));
  if (L.feat_chunks > 0)
    FE_HIP(launch_feat(s, false, L.ts, L.total_tiles, L.feat_chunks, dj, nullptr, L.n_jobs, rec, rec2, off1, off2,
                       list, state, L.total_pix, flags, nullptr, L.total_rec, 0));
  return FE_OK;
}

fe_status forward(hipStream_t s, const fe_scene* scenes, int32_t n_scenes, const fe_view* views,
                  int32_t n_views, const fe_options* opt, void* ws, size_t ws_bytes) {
  std::vector<Job> jobs(n_views > 0 ? (size_t)n_views : 0);
  return forward_impl(s, scenes, n_scenes, views, n_views, opt, ws, ws_bytes, jobs.empty() ? nullptr : jobs.data(),
                      false);
}

fe_status backward(hipStream_t s, const fe_scene* scenes, int32_t n_scenes, const fe_view* views,
                   int32_t n_views, const fe_options* opt, const fe_view_cotangent* cot,
                   const fe_scene_grad* grads, void* ws, size_t ws_bytes) {
  fe_options o = {16, 64, FE_MODE_CULLED, 1, 0, 0.0f, 0.0f};
  if (opt) o = *opt;
  o.save_for_backward = 1;
  Layout L;
  fe_status st = plan(scenes, n_scenes, views, n_views, &o, &L, nullptr);
  if (st != FE_OK) return st;
  if (!cot || !grads) {
    set_msg("fe_backward: null cotangent or gradient array");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (!ws || ws_bytes < L.bytes) {
    set_msg("fe_backward: workspace missing or smaller than the forward's (save_for_backward=1)");
    return FE_ERR_WORKSPACE;
  }
  std::vector<BwdJob> bj(n_views);
  int feat_chunks_bwd = 0;
  for (int v = 0; v < n_views; ++v) {
    const fe_scene_grad& g = grads[views[v].scene];
    const fe_scene& sc = scenes[views[v].scene];
    bj[v].gD = cot[v].d_depth;
    bj[v].gA = cot[v].d_alpha;
    bj[v].dmeans = g.d_means;
    bj[v].dprec = g.d_prec;
    bj[v].dw = g.d_weights_log;
    const bool feat = sc.n_features > 0 && sc.n > 0 && views[v].features_out && cot[v].d_features;
    bj[v].gF = feat ? cot[v].d_features : nullptr;
    bj[v].dfeat = feat ? g.d_features : nullptr;
    if (feat) feat_chunks_bwd = std::max(feat_chunks_bwd, feat_chunks(sc.n_features));
  }
  if (!L.reuse) {
    st = upload(s, at<char>(ws, L.o_bjobs), bj.data(), sizeof(BwdJob) * n_views);
    if (st != FE_OK) return st;
  }
  const Job* dj = at<Job>(ws, L.o_jobs);
  const void* dbj = at<char>(ws, L.o_bjobs);
  const bool dense = (L.mode == FE_MODE_DENSE);
  FE_HIP(timed(s, FE_K_BWD_BLEND, [&] {
    return launch_bwd_blend(s, L.ts, L.total_tiles, dj, dbj, L.n_jobs, at<float>(ws, L.o_rec),
                            at<float>(ws, L.o_rec2), dense ? nullptr : at<uint32_t>(ws, L.o_tnp),
                            dense ? nullptr : at<unsigned long long>(ws, L.o_off1),
                            dense ? nullptr : at<unsigned long long>(ws, L.o_off2),
                            dense ? nullptr : at<uint32_t>(ws, L.o_list), at<float>(ws, L.o_state), L.total_pix,
                            at<int32_t>(ws, L.o_flags), at<float>(ws, L.o_gsum), L.total_rec, dense ? 1 : 0,
                            dense ? nullptr : at<int32_t>(ws, L.o_tord) + L.total_tiles, feat_chunks_bwd);
  }));
  return FE_OK;
}

// Workspace of fe_render_fwd_noalloc: a static device array, one copy per GPU for the life of
// the process.  Host code uses it only through the per-device address hipGetSymbolAddress
// returns (the array's own name is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[FE_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

static fe_status scratch_address(void** out) {
  int dev = 0;
  FE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_msg("fe_render_fwd_noalloc: device index out of range");
    return FE_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_msg("fe_render_fwd_noalloc: no device address for the scratch array");
      return FE_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FE_OK;
}

fe_status render_fwd_noalloc(hipStream_t s, int32_t n, const float* means, const float* prec,
                             const float* weights_log, const fe_camera* cam, float beta2, float beta3, float* depth,
                             float* alpha) {
  if (!cam) {
    set_msg("fe_render_fwd_noalloc: null camera");
    return FE_ERR_INVALID_ARGUMENT;
  }
  fe_scene sc{n, means, prec, weights_log};
  fe_view v;
  memset(&v, 0, sizeof(v));   // no device pose, no ray list
  v.cam = *cam;
  v.beta2 = beta2;
  v.beta3 = beta3;
  v.depth = depth;
  v.alpha = alpha;
  const fe_options o{16, 64, FE_MODE_CULLED, 0, 0, 0.0f, 0.0f};
  Layout L;
  fe_status st = plan(&sc, 1, &v, 1, &o, &L, nullptr);
  if (st != FE_OK) return st;
  if (L.bytes > (size_t)FE_NOALLOC_SCRATCH_BYTES) {
    snprintf(g_err, sizeof(g_err),
             "fe_render_fwd_noalloc: this view needs %zu bytes of scratch, more than the %zu the library holds "
             "per GPU (FE_NOALLOC_SCRATCH_BYTES); use fe_render_fwd with a workspace",
             L.bytes, (size_t)FE_NOALLOC_SCRATCH_BYTES);
    return FE_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  st = scratch_address(&ws);
  if (st != FE_OK) return st;
  Job job;
  return forward_impl(s, &sc, 1, &v, 1, &o, ws, (size_t)FE_NOALLOC_SCRATCH_BYTES, &job, true);
}

}  // namespace fe

extern "C" {

fe_status fe_render_fwd_noalloc(void* stream, int32_t n, const float* means, const float* prec,
                                const float* weights_log, const fe_camera* cam, float beta2, float beta3,
                                float* depth, float* alpha) {
  return fe::render_fwd_noalloc(reinterpret_cast<hipStream_t>(stream), n, means, prec, weights_log, cam, beta2, beta3,
                                depth, alpha);
}

fe_status fe_backward(void* stream, const fe_scene* scenes, int32_t n_scenes, const fe_view* views,
                      int32_t n_views, const fe_options* opt, const fe_view_cotangent* cot,
                      const fe_scene_grad* grads, void* workspace, size_t workspace_bytes) {
  return fe::backward(reinterpret_cast<hipStream_t>(stream), scenes, n_scenes, views, n_views, opt, cot, grads,
                      workspace, workspace_bytes);
}

int32_t fe_version(void) { return 102; }

const char* fe_status_string(fe_status s) {
  switch (s) {
    case FE_OK: return "ok";
    case FE_ERR_INVALID_ARGUMENT: return "invalid argument";
    case FE_ERR_WORKSPACE: return "workspace missing or too small";
    case FE_ERR_HIP: return "HIP runtime error";
    case FE_ERR_NO_DEVICE: return "no device";
  }
  return "unknown status";
}

const char* fe_last_error_string(void) { return fe::g_err; }

fe_status fe_plan(const fe_scene* scenes, int32_t n_scenes, const fe_view* views, int32_t n_views,
                  const fe_options* opt, size_t* workspace_bytes) {
  if (!workspace_bytes) return FE_ERR_INVALID_ARGUMENT;
  fe::Layout L;
  fe_status st = fe::plan(scenes, n_scenes, views, n_views, opt, &L, nullptr);
  if (st == FE_OK) *workspace_bytes = L.bytes;
  return st;
}

fe_status fe_forward(void* stream, const fe_scene* scenes, int32_t n_scenes, const fe_view* views,
                     int32_t n_views, const fe_options* opt, void* workspace, size_t workspace_bytes) {
  return fe::forward(reinterpret_cast<hipStream_t>(stream), scenes, n_scenes, views, n_views, opt,
                     workspace, workspace_bytes);
}

fe_status fe_read_stats(void* stream, const void* workspace, size_t workspace_bytes, fe_stats* out) {
  if (!workspace || !out || workspace_bytes < sizeof(fe::WsHeader)) return FE_ERR_INVALID_ARGUMENT;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  fe::WsHeader h;
  FE_HIP(hipMemcpyAsync(&h, workspace, sizeof(h), hipMemcpyDeviceToHost, s));
  FE_HIP(hipStreamSynchronize(s));
  if (h.magic != fe::kMagic) {
    fe::set_msg("fe_read_stats: workspace was not written by fe_forward");
    return FE_ERR_WORKSPACE;
  }
  fe::Counters c;
  const char* base = reinterpret_cast<const char*>(workspace);
  FE_HIP(hipMemcpyAsync(&c, base + h.o_counters, sizeof(c), hipMemcpyDeviceToHost, s));
  unsigned long long t1 = 0, t2 = 0;
  if (h.mode == FE_MODE_CULLED) {
    FE_HIP(hipMemcpyAsync(&t1, base + h.o_off1 + 8 * h.total_tiles, 8, hipMemcpyDeviceToHost, s));
    FE_HIP(hipMemcpyAsync(&t2, base + h.o_off2 + 8 * h.total_tiles, 8, hipMemcpyDeviceToHost, s));
  }
  FE_HIP(hipStreamSynchronize(s));
  out->pairs_phase1 = (int64_t)t1;
  out->pairs_phase2 = (int64_t)(t2 - t1);
  unsigned long long ev = 0, ev1 = 0;
  for (int k = 0; k < fe::kCountSlots; ++k) { ev += c.slot[k][0]; ev1 += c.slot[k][1]; }
  out->evals = (int64_t)ev;
  out->floor_tiles = (int64_t)c.floor_tiles;
  out->evals_phase1 = (int64_t)ev1;
  out->floor_unresolved = (int64_t)c.floor_unresolved;
  return FE_OK;
}

fe_status fe_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(fe::g_prof.mu);
  constexpr uint32_t all = (1u << FE_K_COUNT) - 1u;
  fe::g_prof.mask = (on & FE_PROFILE_SELECT) ? ((uint32_t)on & all) : (on != 0 ? all : 0u);
  return FE_OK;
}

fe_status fe_profile_read(fe_kernel_times* out, int32_t reset) {
  if (!out) return FE_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(fe::g_prof.mu);
  for (auto& r : fe::g_prof.pending) {
    FE_HIP(hipEventSynchronize(r.b));
    float ms = 0.0f;
    FE_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    fe::g_prof.ms[r.k] += ms;
    fe::g_prof.launches[r.k] += 1;
    fe::g_prof.pool.push_back(r.a);
    fe::g_prof.pool.push_back(r.b);
  }
  fe::g_prof.pending.clear();
  for (int k = 0; k < FE_K_COUNT; ++k) {
    out->ms[k] = fe::g_prof.ms[k];
    out->launches[k] = fe::g_prof.launches[k];
    if (reset) {
      fe::g_prof.ms[k] = 0.0;
      fe::g_prof.launches[k] = 0;
    }
  }
  return FE_OK;
}

fe_status fe_render_workspace_bytes(int32_t n, int32_t height, int32_t width, int32_t tile, size_t* bytes) {
  fe_scene sc{n, nullptr, nullptr, nullptr};
  fe_view v;
  memset(&v, 0, sizeof(v));
  v.cam.height = height;
  v.cam.width = width;
  v.cam.fx = v.cam.fy = 1.0f;
  fe_options o{tile, 64, FE_MODE_CULLED, 0, 0, 0.0f, 0.0f};
  return fe_plan(&sc, 1, &v, 1, &o, bytes);
}

fe_status fe_render_fwd(void* stream, int32_t n, const float* means, const float* prec,
                        const float* weights_log, const fe_camera* cam, float beta2, float beta3,
                        float* depth, float* alpha, void* workspace, size_t workspace_bytes) {
  if (!cam) return FE_ERR_INVALID_ARGUMENT;
  fe_scene sc{n, means, prec, weights_log};
  fe_view v;
  memset(&v, 0, sizeof(v));   // no device pose, no ray list
  v.scene = 0;
  v.cam = *cam;
  v.beta2 = beta2;
  v.beta3 = beta3;
  v.depth = depth;
  v.alpha = alpha;
  fe_options o{16, 64, FE_MODE_CULLED, 0, 0, 0.0f, 0.0f};
  return fe_forward(stream, &sc, 1, &v, 1, &o, workspace, workspace_bytes);
}

static size_t sil_align(size_t x) { return (x + 255) & ~(size_t)255; }

fe_status fe_silhouette_workspace_bytes(int32_t n_views, size_t* bytes) {
  if (!bytes || n_views <= 0 || n_views > (1 << 20)) return FE_ERR_INVALID_ARGUMENT;
  const size_t part = sizeof(float) * (size_t)n_views * fe::kSilBlocksPerView;
  *bytes = sil_align(sizeof(fe_sil_view) * (size_t)n_views) + 2 * sil_align(part);
  return FE_OK;
}

fe_status fe_silhouette_loss(void* stream, const fe_sil_view* views, int32_t n_views, float w_depth, float clip,
                             float* loss, int32_t flags, void* workspace, size_t workspace_bytes) {
  size_t need = 0;
  if (!views || !loss || fe_silhouette_workspace_bytes(n_views, &need) != FE_OK) {
    fe::set_msg("fe_silhouette_loss: need views, a loss pointer and 1 <= n_views <= 2^20");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (!(clip >= 0.0f && clip < 0.5f) || !(w_depth >= 0.0f)) {
    fe::set_msg("fe_silhouette_loss: clip must be in [0, 0.5) and w_depth >= 0");
    return FE_ERR_INVALID_ARGUMENT;
  }
  for (int32_t i = 0; i < n_views; ++i) {
    const fe_sil_view& v = views[i];
    if (!v.alpha || !v.depth || !v.mask || !v.target_depth || !v.d_alpha || !v.d_depth || v.n_pixels <= 0) {
      fe::set_msg("fe_silhouette_loss: null buffer or empty view");
      return FE_ERR_INVALID_ARGUMENT;
    }
  }
  if (!workspace || workspace_bytes < need) {
    fe::set_msg("fe_silhouette_loss: workspace missing or too small (see fe_silhouette_workspace_bytes)");
    return FE_ERR_WORKSPACE;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  char* base = reinterpret_cast<char*>(workspace);
  fe_sil_view* d_views = reinterpret_cast<fe_sil_view*>(base);
  const size_t part = sil_align(sizeof(float) * (size_t)n_views * fe::kSilBlocksPerView);
  float* cnt_part = reinterpret_cast<float*>(base + sil_align(sizeof(fe_sil_view) * (size_t)n_views));
  float* loss_part = reinterpret_cast<float*>(reinterpret_cast<char*>(cnt_part) + part);
  if (!(flags & FE_FLAG_REUSE_DESCRIPTORS)) {
    fe_status st = fe::upload(s, d_views, views, sizeof(fe_sil_view) * (size_t)n_views);
    if (st != FE_OK) return st;
  }
  FE_HIP(fe::launch_sil_loss(s, d_views, n_views, cnt_part, loss_part, w_depth, clip, loss));
  return FE_OK;
}

fe_status fe_adam_step(void* stream, int64_t n, float* param, const float* grad, float* m, float* v, float lr,
                       float beta1, float beta2, float eps, int32_t step) {
  if (n < 0 || (n > 0 && (!param || !grad || !m || !v)) || step < 1 || !(beta1 >= 0.0f && beta1 < 1.0f) ||
      !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f)) {
    fe::set_msg("fe_adam_step: need buffers, step >= 1, 0 <= beta < 1 and eps > 0");
    return FE_ERR_INVALID_ARGUMENT;
  }
  const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
  const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
  FE_HIP(fe::launch_adam(reinterpret_cast<hipStream_t>(stream), n, param, grad, m, v, lr, beta1, beta2, c1, c2,
                         eps));
  return FE_OK;
}

namespace {
struct GmmLayout {
  size_t rec, wnorm, fpart, bpart, bytes;
};
bool gmm_layout(int32_t n, int64_t m, GmmLayout* L) {
  if (n < 1 || n > (1 << 24) || m < 1 || m > ((int64_t)1 << 36)) return false;
  const int64_t chunk = fe::gmm_chunk(m);
  const int64_t n_chunks = (m + chunk - 1) / chunk;
  const int64_t nb = (m + 255) / 256;
  L->rec = 0;
  L->wnorm = sil_align(sizeof(float) * 16 * (size_t)n);
  L->fpart = L->wnorm + 256;
  L->bpart = L->fpart + sil_align(sizeof(float) * (size_t)nb);
  L->bytes = L->bpart + sil_align(sizeof(float) * 10 * (size_t)n * (size_t)n_chunks);
  return true;
}
}  // namespace

fe_status fe_gmm_workspace_bytes(int32_t n, int64_t m, size_t* bytes) {
  GmmLayout L;
  if (!bytes || !gmm_layout(n, m, &L)) return FE_ERR_INVALID_ARGUMENT;
  *bytes = L.bytes;
  return FE_OK;
}

fe_status fe_gmm_forward(void* stream, int32_t n, const float* means, const float* prec, const float* weights_log,
                         int64_t m, const float* points, float* point_lse, float* res, float* nll, void* workspace,
                         size_t workspace_bytes) {
  GmmLayout L;
  if (!gmm_layout(n, m, &L) || !means || !prec || !weights_log || !points || !point_lse || !nll) {
    fe::set_msg("fe_gmm_forward: need 1 <= n <= 2^24, 1 <= m <= 2^36 and non-null buffers (res may be NULL)");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (!workspace || workspace_bytes < L.bytes) {
    fe::set_msg("fe_gmm_forward: workspace missing or too small (see fe_gmm_workspace_bytes)");
    return FE_ERR_WORKSPACE;
  }
  char* b = reinterpret_cast<char*>(workspace);
  FE_HIP(fe::launch_gmm_forward(reinterpret_cast<hipStream_t>(stream), n, m, means, prec, weights_log, points,
                                reinterpret_cast<float*>(b + L.rec), reinterpret_cast<float*>(b + L.wnorm),
                                reinterpret_cast<float*>(b + L.fpart), point_lse, res, nll));
  return FE_OK;
}

fe_status fe_gmm_backward(void* stream, int32_t n, const float* prec, const float* weights_log, int64_t m,
                          const float* points, const float* point_lse, float d_nll, float* d_means, float* d_prec,
                          float* d_weights_log, void* workspace, size_t workspace_bytes) {
  GmmLayout L;
  if (!gmm_layout(n, m, &L) || !prec || !weights_log || !points || !point_lse || !d_means || !d_prec ||
      !d_weights_log) {
    fe::set_msg("fe_gmm_backward: need 1 <= n <= 2^24, 1 <= m <= 2^36 and non-null buffers");
    return FE_ERR_INVALID_ARGUMENT;
  }
  if (!workspace || workspace_bytes < L.bytes) {
    fe::set_msg("fe_gmm_backward: workspace missing or too small (see fe_gmm_workspace_bytes)");
    return FE_ERR_WORKSPACE;
  }
  char* b = reinterpret_cast<char*>(workspace);
  FE_HIP(fe::launch_gmm_backward(reinterpret_cast<hipStream_t>(stream), n, m, prec, weights_log, points,
                                 reinterpret_cast<const float*>(b + L.rec),
                                 reinterpret_cast<const float*>(b + L.wnorm), point_lse,
                                 reinterpret_cast<float*>(b + L.bpart), d_nll, d_means, d_prec, d_weights_log));
  return FE_OK;
}

fe_status fe_ray_flow(void* stream, int64_t n_rays, const float* pixel_posei, float inv_f, int32_t n_poses,
                      const float* rot, const float* trans, const float* depth, float* flow_next, float* flow_prev) {
  if (n_rays < 0 || n_poses < 1 || !(inv_f != 0.0f) ||
      (n_rays > 0 && (!pixel_posei || !rot || !trans || !depth || !flow_next || !flow_prev))) {
    fe::set_msg("fe_ray_flow: need n_rays >= 0, n_poses >= 1, inv_f != 0 and non-null buffers");
    return FE_ERR_INVALID_ARGUMENT;
  }
  FE_HIP(fe::launch_ray_flow(reinterpret_cast<hipStream_t>(stream), n_rays, pixel_posei, inv_f, n_poses, rot, trans,
                             depth, flow_next, flow_prev));
  return FE_OK;
}

}  // extern "C"
