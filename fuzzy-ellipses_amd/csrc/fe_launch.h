// Host-callable launchers of the device kernels (defined in fe_forward.hip / fe_backward.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "../../include/fuzzy_ellipses_amd.h"

namespace fe {

struct Job;
struct Counters;

// Blocks per view of the silhouette objective kernels (partial sums per view).
constexpr int kSilBlocksPerView = 64;

// Sets the text fe_last_error_string reports (this thread).
void set_error_message(const char* msg);

hipError_t launch_init(hipStream_t s, Counters* cnt, int32_t* job_flags, int n_jobs);
hipError_t launch_pose(hipStream_t s, Job* jobs, int n_jobs);
hipError_t launch_ray_rects(hipStream_t s, const Job* jobs, int n_jobs, int64_t total_tiles, int ts,
                            int32_t* job_flags);
// Also clears the forward's counters and job flags (as launch_init).
hipError_t launch_order(hipStream_t s, const Job* jobs, int n_jobs, int max_n, Counters* cnt, int32_t* job_flags);
hipError_t launch_preprocess(hipStream_t s, const Job* jobs, int n_jobs, int max_n, float* rec, float* rec2,
                             int32_t* job_flags);
// mpart: the membership partials region (5 * kMaskMaxSub * kMaskChunkBlocks words)
hipError_t launch_tile_mask(hipStream_t s, int phase, const Job* jobs, int n_jobs, int64_t total_stiles,
                            int ts, int max_n, const float* rec, unsigned long long* mask1, unsigned long long* mask2,
                            uint32_t* count, const float* tile_lmin, const float* tile_seed, float* tile_ubm,
                            uint32_t* tile_np, const int32_t* job_flags, Counters* cnt, uint32_t* mpart);
int64_t scan_partial_count(int64_t n);
int device_cus();   // CUs of the current device (cached per device; 256 if the query fails)
// With hist and tord (both or neither), also writes the tiles' dispatch order to tord (a
// counting sort over kTileBuckets buckets of the tile key; hist must start at zero).
hipError_t launch_scan(hipStream_t s, const uint32_t* cnt, int64_t n, unsigned long long* partial,
                       unsigned long long* off, const unsigned long long* base_src,
                       const unsigned long long* key_off, uint32_t* hist, int32_t* tord);
hipError_t launch_tile_write(hipStream_t s, const Job* jobs, int n_jobs, int64_t total_tiles,
                             const unsigned long long* mask, const unsigned long long* off,
                             uint32_t* list);
// split_views: views of the batch whose tile lists the blend splits (view_splits_lists)
hipError_t launch_blend(hipStream_t s, int ts, int gb, int phase, int64_t total_tiles, const Job* jobs,
                        int n_jobs, const float* rec, const float* rec2, const uint32_t* tile_np,
                        const unsigned long long* off1,
                        const unsigned long long* off2, const uint32_t* list, float* st, int64_t total_pix,
                        float* tile_lmin, float* tile_seed, const float* tile_ubm, Counters* cnt,
                        const int32_t* job_flags, int save, const int32_t* tord,
                        int split_views);

// With feat_chunks > 0 the feature backward (k_feat_bwd) runs between the blend and the
// per-record chain rule.
hipError_t launch_bwd_blend(hipStream_t s, int ts, int64_t total_tiles, const Job* jobs, const void* bjobs,
                            int n_jobs, const float* rec, const float* rec2, const uint32_t* tile_np,
                            const unsigned long long* off1, const unsigned long long* off2, const uint32_t* list,
                            const float* st, int64_t total_pix, const int32_t* job_flags, float* gsum,
                            int64_t total_rec, int dense, const int32_t* tord, int feat_chunks);
// Blocks per tile of the feature kernels for n_feat channels.
int feat_chunks(int n_feat);
// Feature image (forward) or its backward over every tile, `chunks` blocks per tile.
hipError_t launch_feat(hipStream_t s, bool backward, int ts, int64_t total_tiles, int chunks, const Job* jobs,
                       const void* bjobs, int n_jobs, const float* rec, const float* rec2,
                       const unsigned long long* off1, const unsigned long long* off2, const uint32_t* list,
                       const float* st, int64_t total_pix, const int32_t* job_flags, float* gsum, int64_t total_rec,
                       int dense);
size_t bwd_gsum_floats();

hipError_t launch_sil_loss(hipStream_t s, const fe_sil_view* views, int n_views, float* cnt_part,
                           float* loss_part, float w_depth, float clip, float* loss);
hipError_t launch_adam(hipStream_t s, int64_t n, float* p, const float* g, float* m, float* v, float lr,
                       float b1, float b2, float c1, float c2, float eps);

int64_t gmm_chunk(int64_t m);
hipError_t launch_gmm_forward(hipStream_t s, int n, int64_t m, const float* means, const float* prec,
                              const float* wlog, const float* pts, float* rec, float* wnorm, float* part,
                              float* lse, float* res, float* nll);
hipError_t launch_gmm_backward(hipStream_t s, int n, int64_t m, const float* prec, const float* wlog,
                               const float* pts, const float* rec, const float* wnorm, const float* lse,
                               float* part, float g, float* d_means, float* d_prec, float* d_wlog);

hipError_t launch_ray_flow(hipStream_t s, int64_t n, const float* pp, float invF, int n_poses, const float* rot,
                           const float* trans, const float* depth, float* fnext, float* fprev);

}  // namespace fe
