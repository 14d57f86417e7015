// The reference's XLA custom-call target, on the MI355X pipeline.
//
// rasterize_gaussians_fwd(stream, buffers, opaque, opaque_len) has the signature, buffer
// order and 20-byte descriptor of the reference's legacy custom call (rasterize_main.h:13-30,
// rasterize_main.cu:29-114, registered from ext.cpp:22-32, lowered by
// fuzzy_ellipses/__init__.py:111-166):
//   buffers[0] means (N,3)  [1] log-weights (N,)  [2] scales (N,3)  [3] quaternions (N,4, w first)
//   [4] camera rays (H*W,3)  [5] R (3,3)  [6] t (3,)  [7] view (4,4)  [8] projection (4,4)
//   [9] depth (H,W) out  [10] alpha (H,W) out
//   opaque = {int32 H, W, N; float tan_fovx, tan_fovy}
// Everything stays on the device and on the caller's stream:
//   k_xla_prec    precision factor U (upper triangular, U U^T = (R diag(s^2) R^T)^-1) from
//                 scales and quaternions, in float64 (quat_to_rot of baseline/util_render.py:51-76)
//   k_xla_rays    the (H*W,3) rays gathered into 16x16-pixel tile order as a ray list (tile t
//                 holds rays [256 t, 256 t + 256)), edge tiles padded with their last in-image
//                 pixel; the pose R|t copied into one device array
//   fe::forward   one ray-list view with the device pose (beta2 = 21.4, beta3 = 2.66, the
//                 reference's fixed values, rasterizer_impl.cu:196-197)
//   k_xla_scatter the tile-ordered results written back to row-major (H,W)
// View and projection matrices and tan-FOV only fed the reference's culling heuristic; the
// math does not use them.  Rays are used as given: the world ray is rays[k] @ R from t
// (rasterizer_impl.cu:200, baseline/fm_render.py:116).  A ray's depth depends on its length
// (t* is in units of the ray), and the ray-list form carries camera-plane rays (x, y, 1):
// rows whose third component is not 1 get alpha from the ray (x/z, y/z, 1) (alpha does not
// depend on the ray's length) and a NaN depth.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "../../include/fuzzy_ellipses_amd.h"
#include "fe_common.h"
#include "fe_launch.h"

namespace fe {

constexpr int kXlaTile = 16;
constexpr float kXlaBeta2 = 21.4f;
constexpr float kXlaBeta3 = 2.66f;

// Upper-triangular U with U U^T = Sigma^-1, Sigma = R diag(s^2) R^T: C = chol(Sigma) (lower),
// U = (C^-1)^T.  quat_to_rot normalises by 2/|q|^2 and gives the identity for |q|^2 <= 1e-12.
__global__ void k_xla_prec(int n, const float* __restrict__ scales, const float* __restrict__ quats,
                           float* __restrict__ prec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double w = quats[4 * i], x = quats[4 * i + 1], y = quats[4 * i + 2], z = quats[4 * i + 3];
  const double nq = w * w + x * x + y * y + z * z;
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (nq > 1e-12) {
    const double s = 2.0 / nq;
    const double X = x * s, Y = y * s, Z = z * s;
    const double wX = w * X, wY = w * Y, wZ = w * Z, xX = x * X, xY = x * Y, xZ = x * Z;
    const double yY = y * Y, yZ = y * Z, zZ = z * Z;
    R[0] = 1.0 - (yY + zZ); R[1] = xY - wZ;         R[2] = xZ + wY;
    R[3] = xY + wZ;         R[4] = 1.0 - (xX + zZ); R[5] = yZ - wX;
    R[6] = xZ - wY;         R[7] = yZ + wX;         R[8] = 1.0 - (xX + yY);
  }
  double s2[3];
  for (int k = 0; k < 3; ++k) {
    const double s = scales[3 * i + k];
    s2[k] = s * s;
  }
  double S[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      S[a][b] = R[3 * a] * s2[0] * R[3 * b] + R[3 * a + 1] * s2[1] * R[3 * b + 1] + R[3 * a + 2] * s2[2] * R[3 * b + 2];
  // Cholesky S = C C^T, C = [[c00,0,0],[c10,c11,0],[c20,c21,c22]]
  const double c00 = sqrt(S[0][0]);
  const double c10 = S[1][0] / c00, c20 = S[2][0] / c00;
  const double c11 = sqrt(S[1][1] - c10 * c10);
  const double c21 = (S[2][1] - c20 * c10) / c11;
  const double c22 = sqrt(S[2][2] - c20 * c20 - c21 * c21);
  // C^-1 = [[1/c00,0,0],[-c10/(c00 c11),1/c11,0],[(c10 c21 - c11 c20)/(c00 c11 c22),-c21/(c11 c22),1/c22]]
  float* U = prec + 9 * (int64_t)i;
  U[0] = (float)(1.0 / c00);
  U[1] = (float)(-c10 / (c00 * c11));
  U[2] = (float)((c10 * c21 - c11 * c20) / (c00 * c11 * c22));
  U[3] = 0.0f;
  U[4] = (float)(1.0 / c11);
  U[5] = (float)(-c21 / (c11 * c22));
  U[6] = 0.0f;
  U[7] = 0.0f;
  U[8] = (float)(1.0 / c22);
}

// Ray list in tile order: list entry j is pixel (x, y) of tile j / 256 (tiles row-major,
// ceil(W/16) per row), x and y clamped to the image (padding of edge tiles).
__device__ __forceinline__ int64_t xla_src_pixel(int64_t j, int H, int W, bool* inside) {
  const int tiles_x = (W + kXlaTile - 1) / kXlaTile;
  const int64_t tile = j / (kXlaTile * kXlaTile);
  const int l = (int)(j - tile * (kXlaTile * kXlaTile));
  const int x = (int)(tile % tiles_x) * kXlaTile + (l % kXlaTile);
  const int y = (int)(tile / tiles_x) * kXlaTile + (l / kXlaTile);
  *inside = (x < W) && (y < H);
  return (int64_t)min(y, H - 1) * W + min(x, W - 1);
}

__global__ void k_xla_rays(int64_t n_list, int H, int W, const float* __restrict__ rays, const float* __restrict__ R,
                           const float* __restrict__ t, float* __restrict__ xy, float* __restrict__ pose) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < 12) pose[j] = j < 9 ? R[j] : t[j - 9];
  if (j >= n_list) return;
  bool inside;
  const int64_t k = xla_src_pixel(j, H, W, &inside);
  const float rz = rays[3 * k + 2];
  float2 v = make_float2(rays[3 * k], rays[3 * k + 1]);
  if (rz != 1.0f) v = make_float2(v.x / rz, v.y / rz);
  reinterpret_cast<float2*>(xy)[j] = v;
}

__global__ void k_xla_scatter(int64_t n_list, int H, int W, const float* __restrict__ rays,
                              const float* __restrict__ dl, const float* __restrict__ al, float* __restrict__ depth,
                              float* __restrict__ alpha) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_list) return;
  bool inside;
  const int64_t k = xla_src_pixel(j, H, W, &inside);
  if (!inside) return;
  depth[k] = rays[3 * k + 2] == 1.0f ? dl[j] : __builtin_nanf("");
  alpha[k] = al[j];
}

__global__ void k_xla_fill(int64_t n, float v, float* __restrict__ a, float* __restrict__ b) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    a[j] = v;
    b[j] = v;
  }
}

static inline size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// Library-owned workspace of the custom call, one per (device, stream).  XLA's legacy custom
// call passes no scratch buffer, so the library keeps one and grows it when a larger call
// comes; growing waits for the stream (the old block may still be in use) and happens only
// at a new largest size.
struct XlaArena {
  void* ptr = nullptr;
  size_t cap = 0;
};
static std::mutex g_xla_mu;
static std::map<std::pair<int, hipStream_t>, XlaArena> g_xla_arenas;

static fe_status xla_fail(fe_status st, const char* msg) {
  char buf[400];
  snprintf(buf, sizeof(buf), "rasterize_gaussians_fwd: %s", msg);
  set_error_message(buf);
  return st;
}

static fe_status rasterize_xla(hipStream_t s, void** buffers, const char* opaque, size_t opaque_len) {
  if (!buffers || !opaque || opaque_len != sizeof(fe_rasterize_descriptor))
    return xla_fail(FE_ERR_INVALID_ARGUMENT, "opaque must be the 20-byte descriptor");
  fe_rasterize_descriptor d;
  memcpy(&d, opaque, sizeof(d));
  const int H = d.image_height, W = d.image_width, n = d.P;
  if (H <= 0 || W <= 0 || n < 0 || (int64_t)H * W > ((int64_t)1 << 30))
    return xla_fail(FE_ERR_INVALID_ARGUMENT, "invalid image size or Gaussian count");
  for (int b : {0, 1, 2, 3, 4, 5, 6, 9, 10})
    if (!buffers[b] && (n > 0 || b >= 9)) return xla_fail(FE_ERR_INVALID_ARGUMENT, "null buffer");
  float* depth = static_cast<float*>(buffers[9]);
  float* alpha = static_cast<float*>(buffers[10]);
  const int64_t npix = (int64_t)H * W;
  if (n == 0) {   // the reference leaves its zeroed outputs when there is no Gaussian
    hipLaunchKernelGGL(k_xla_fill, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, npix, 0.0f, depth, alpha);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FE_OK : xla_fail(FE_ERR_HIP, hipGetErrorString(e));
  }
  const int tiles = ((W + kXlaTile - 1) / kXlaTile) * ((H + kXlaTile - 1) / kXlaTile);
  const int64_t n_list = (int64_t)tiles * kXlaTile * kXlaTile;

  fe_scene sc{n, static_cast<const float*>(buffers[0]), nullptr, static_cast<const float*>(buffers[1])};
  fe_view v;
  memset(&v, 0, sizeof(v));
  v.scene = 0;
  v.cam.fx = v.cam.fy = 1.0f;
  v.cam.height = 1;
  v.cam.width = (int32_t)std::min<int64_t>(n_list, INT32_MAX);
  v.beta2 = kXlaBeta2;
  v.beta3 = kXlaBeta3;
  v.n_rays = n_list;
  fe_options o{kXlaTile, 64, FE_MODE_CULLED, 0, 0, 0.0f, 0.0f};
  // layout: pose | prec | ray list | list depth | list alpha | renderer workspace
  const size_t o_pose = 0;
  const size_t o_prec = al256(o_pose + 12 * sizeof(float));
  const size_t o_xy = al256(o_prec + 9 * sizeof(float) * (size_t)n);
  const size_t o_dl = al256(o_xy + 2 * sizeof(float) * (size_t)n_list);
  const size_t o_al = al256(o_dl + sizeof(float) * (size_t)n_list);
  const size_t o_ws = al256(o_al + sizeof(float) * (size_t)n_list);
  v.rays = reinterpret_cast<const float*>(16);   // planning only needs it non-null
  size_t ws_bytes = 0;
  fe_status st = fe_plan(&sc, 1, &v, 1, &o, &ws_bytes);
  if (st != FE_OK) return st;
  const size_t total = o_ws + ws_bytes;

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return xla_fail(FE_ERR_HIP, hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_xla_mu);   // one call at a time enqueues on an arena
  XlaArena& ar = g_xla_arenas[std::make_pair(dev, s)];
  if (ar.cap < total) {
    if (ar.ptr) {
      e = hipStreamSynchronize(s);
      if (e == hipSuccess) e = hipFree(ar.ptr);
      ar.ptr = nullptr;
      ar.cap = 0;
      if (e != hipSuccess) return xla_fail(FE_ERR_HIP, hipGetErrorString(e));
    }
    const size_t cap = total + total / 4;
    e = hipMalloc(&ar.ptr, cap);
    if (e != hipSuccess) {
      ar.ptr = nullptr;
      return xla_fail(FE_ERR_HIP, hipGetErrorString(e));
    }
    ar.cap = cap;
  }
  char* base = static_cast<char*>(ar.ptr);
  float* pose = reinterpret_cast<float*>(base + o_pose);
  float* prec = reinterpret_cast<float*>(base + o_prec);
  float* xy = reinterpret_cast<float*>(base + o_xy);
  float* dl = reinterpret_cast<float*>(base + o_dl);
  float* al = reinterpret_cast<float*>(base + o_al);
  const float* rays = static_cast<const float*>(buffers[4]);

  hipLaunchKernelGGL(k_xla_prec, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, n,
                     static_cast<const float*>(buffers[2]), static_cast<const float*>(buffers[3]), prec);
  hipLaunchKernelGGL(k_xla_rays, dim3((unsigned)((n_list + 255) / 256)), dim3(256), 0, s, n_list, H, W, rays,
                     static_cast<const float*>(buffers[5]), static_cast<const float*>(buffers[6]), xy, pose);
  e = hipGetLastError();
  if (e != hipSuccess) return xla_fail(FE_ERR_HIP, hipGetErrorString(e));
  sc.prec = prec;
  v.rays = xy;
  v.pose_device = pose;
  v.depth = dl;
  v.alpha = al;
  st = fe_forward(s, &sc, 1, &v, 1, &o, base + o_ws, ar.cap - o_ws);
  if (st != FE_OK) return st;
  hipLaunchKernelGGL(k_xla_scatter, dim3((unsigned)((n_list + 255) / 256)), dim3(256), 0, s, n_list, H, W, rays, dl,
                     al, depth, alpha);
  e = hipGetLastError();
  return e == hipSuccess ? FE_OK : xla_fail(FE_ERR_HIP, hipGetErrorString(e));
}

}  // namespace fe

extern "C" {

fe_status fe_rasterize_gaussians_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  return fe::rasterize_xla(reinterpret_cast<hipStream_t>(stream), buffers, opaque, opaque_len);
}

// The legacy target returns nothing: a failed call fills both outputs with NaN when the
// descriptor is readable (the error text stays in fe_last_error_string).
void rasterize_gaussians_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  const fe_status st = fe_rasterize_gaussians_fwd(stream, buffers, opaque, opaque_len);
  if (st == FE_OK || !buffers || !opaque || opaque_len != sizeof(fe_rasterize_descriptor)) return;
  fe_rasterize_descriptor d;
  memcpy(&d, opaque, sizeof(d));
  if (d.image_height <= 0 || d.image_width <= 0 || !buffers[9] || !buffers[10]) return;
  const int64_t npix = (int64_t)d.image_height * d.image_width;
  if (npix > ((int64_t)1 << 30)) return;
  hipLaunchKernelGGL(fe::k_xla_fill, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), npix, __builtin_nanf(""),
                     static_cast<float*>(buffers[9]), static_cast<float*>(buffers[10]));
}

}  // extern "C"
