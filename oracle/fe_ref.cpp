// CPU restatements for the fuzzy-metaball rasterizer -- TEST INFRASTRUCTURE ONLY.
//
// Only tests, the smoke check and the benchmark's CPU-baseline leg load this library, and
// only as the checker / reported baseline.  The product path never links it.
//
// 1. fe_ref_render_f64: dense float64 restatement of render_func_rays
//    (baseline/fm_render.py:11-109) for a pinhole camera whose rays follow
//    baseline/ycb_fmb_rotscale.py:175-188 and fm_render.py:113-120 (rays @ R, origin t).
//    Per pair: prc = triu(P)^T, p = mu - o, vsv = |prc r|^2, psv = (prc p).(prc r),
//    t* = psv/vsv, d0 = |prc (r t* - p)|^2, d2 = w - d0/2; per ray alpha = 1-exp(-sum e^d2),
//    weights = (t*>0) nan_to_num(exp(l - max l)) + 1e-20 with l = -b2 t* + b3 d2, depth =
//    sum(weights * nan_to_num(t*)) / sum(weights).  A NaN logit makes the max NaN, as jnp.max.
// 2. fe_ref_render_f32: the same dense math in float32 with the per-Gaussian hoisting and
//    one-pass online maximum of the GPU kernel -- the timed CPU baseline (OpenMP).
// 3. fe_ref_tiles: bit-exact restatement of the tile membership (records, bound test,
//    per-tile sorted lists) written independently of the HIP source, same float op order.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ------------------------------------------------------------------------------------------
// 1. float64 oracle
// ------------------------------------------------------------------------------------------
void fe_ref_render_f64(int n, const double* means, const double* prec9, const double* wlog, const double* R9,
                       const double* o3, double fx, double fy, double cx, double cy, int H, int W, double beta2,
                       double beta3, const int64_t* pixels, int64_t npix, double* depth, double* alpha,
                       int nthreads) {
  // hoist per Gaussian: prc (lower = triu(P)^T) and p = mu - o
  std::vector<double> L(9 * (size_t)n), P(3 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const double* U = prec9 + 9 * (size_t)j;
    double* A = &L[9 * (size_t)j];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[3 * r + c] = (c <= r) ? U[3 * c + r] : 0.0;  // (triu U)^T
    for (int k = 0; k < 3; ++k) P[3 * (size_t)j + k] = means[3 * (size_t)j + k] - o3[k];
  }
  if (npix <= 0) npix = (int64_t)H * W;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    std::vector<double> zs(n > 0 ? n : 1), d2s(n > 0 ? n : 1);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t ip = 0; ip < npix; ++ip) {
      const int64_t pid = pixels ? pixels[ip] : ip;
      const double x = (double)(pid % W), y = (double)(pid / W);
      const double rc[3] = {(x - cx) / fx, (y - cy) / fy, 1.0};
      double r[3];
      for (int k = 0; k < 3; ++k) r[k] = rc[0] * R9[0 * 3 + k] + rc[1] * R9[1 * 3 + k] + rc[2] * R9[2 * 3 + k];
      double esum = 0.0;
      for (int j = 0; j < n; ++j) {
        const double* A = &L[9 * (size_t)j];
        const double* p = &P[3 * (size_t)j];
        double pp[3], pr[3];
        for (int i = 0; i < 3; ++i) {
          pp[i] = A[3 * i] * p[0] + A[3 * i + 1] * p[1] + A[3 * i + 2] * p[2];
          pr[i] = A[3 * i] * r[0] + A[3 * i + 1] * r[1] + A[3 * i + 2] * r[2];
        }
        const double vsv = pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2];
        const double psv = pp[0] * pr[0] + pp[1] * pr[1] + pp[2] * pr[2];
        const double res = psv / vsv;
        double v[3], pv[3];
        for (int i = 0; i < 3; ++i) v[i] = r[i] * res - p[i];
        for (int i = 0; i < 3; ++i) pv[i] = A[3 * i] * v[0] + A[3 * i + 1] * v[1] + A[3 * i + 2] * v[2];
        const double d0 = pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2];
        zs[j] = res;
        d2s[j] = -0.5 * d0 + wlog[j];
        esum += exp(d2s[j]);
      }
      alpha[ip] = 1.0 - exp(-esum);
      // max over all logits, NaN-propagating
      double Lmax = -INFINITY;
      bool anynan = false;
      for (int j = 0; j < n; ++j) {
        const double l = -zs[j] * beta2 + beta3 * d2s[j];
        if (l != l) anynan = true;
        else if (l > Lmax) Lmax = l;
      }
      double sw = 0.0, sz = 0.0;
      for (int j = 0; j < n; ++j) {
        const double l = -zs[j] * beta2 + beta3 * d2s[j];
        double e = anynan ? NAN : exp(l - Lmax);
        if (e != e) e = 0.0;                       // nan_to_num
        else if (isinf(e)) e = e > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
        const double wj = (zs[j] > 0 ? e : 0.0) + 1e-20;
        double z = zs[j];
        if (z != z) z = 0.0;
        else if (isinf(z)) z = z > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
        sw += wj;
        sz += wj * z;
      }
      depth[ip] = (n > 0) ? sz / sw : 0.0;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 2. float32 dense baseline (same hoisted math as the GPU blend, one pass, online max)
// ------------------------------------------------------------------------------------------
void fe_ref_render_f32(int n, const float* means, const float* prec9, const float* wlog, const float* R9,
                       const float* o3, float fx, float fy, float cx, float cy, int H, int W, float beta2,
                       float beta3, const int64_t* pixels, int64_t npix, float* depth, float* alpha,
                       int nthreads) {
  // per Gaussian: M = A R^T, a = M[:,0]/fx, b = M[:,1]/fy, g = M[:,2], q = A (mu - o)
  std::vector<float> rec(13 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const float* U = prec9 + 9 * (size_t)j;
    const float A[3][3] = {{U[0], 0, 0}, {U[1], U[4], 0}, {U[2], U[5], U[8]}};
    float M[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) M[i][k] = A[i][0] * R9[3 * k] + A[i][1] * R9[3 * k + 1] + A[i][2] * R9[3 * k + 2];
    const float p[3] = {means[3 * j] - o3[0], means[3 * j + 1] - o3[1], means[3 * j + 2] - o3[2]};
    float* r = &rec[13 * (size_t)j];
    for (int i = 0; i < 3; ++i) {
      r[i] = M[i][0] / fx;
      r[3 + i] = M[i][1] / fy;
      r[6 + i] = M[i][2];
      r[9 + i] = A[i][0] * p[0] + A[i][1] * p[1] + A[i][2] * p[2];
    }
    r[12] = wlog[j];
  }
  if (npix <= 0) npix = (int64_t)H * W;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t ip = 0; ip < npix; ++ip) {
    const int64_t pid = pixels ? pixels[ip] : ip;
    const float xf = (float)(pid % W) - cx, yf = (float)(pid / W) - cy;
    float M = -3.402823466e38f, S = 0.f, Z = 0.f, E = 0.f, Zall = 0.f;
    for (int j = 0; j < n; ++j) {
      const float* r = &rec[13 * (size_t)j];
      const float u0 = r[0] * xf + r[3] * yf + r[6];
      const float u1 = r[1] * xf + r[4] * yf + r[7];
      const float u2 = r[2] * xf + r[5] * yf + r[8];
      const float vsv = u0 * u0 + u1 * u1 + u2 * u2;
      const float psv = r[9] * u0 + r[10] * u1 + r[11] * u2;
      const float t = psv / vsv;
      const float v0 = t * u0 - r[9], v1 = t * u1 - r[10], v2 = t * u2 - r[11];
      const float d2 = r[12] - 0.5f * (v0 * v0 + v1 * v1 + v2 * v2);
      E += expf(d2);
      const float l = beta3 * d2 - beta2 * t;
      Zall += (t == t) ? t : 0.f;
      if (l > M) {
        const float sc = expf(M - l);
        S *= sc;
        Z *= sc;
        M = l;
      }
      if (t > 0.f) {
        const float s = expf(l - M);
        S += s;
        Z += s * t;
      }
    }
    alpha[ip] = -expm1f(-E);
    depth[ip] = (n > 0) ? (Z + 1e-20f * Zall) / (S + 1e-20f * (float)n) : 0.f;
  }
}

// ------------------------------------------------------------------------------------------
// 3. tile membership, bit-exact restatement (compile with -ffp-contract=off)
// ------------------------------------------------------------------------------------------
static const float kLog2e = 1.4426950408889634f;

static inline float dot3f(const float* x, const float* y) { return fmaf(x[0], y[0], fmaf(x[1], y[1], x[2] * y[2])); }
static inline void crossf(const float* x, const float* y, float* c) {
  c[0] = fmaf(x[1], y[2], -(x[2] * y[1]));
  c[1] = fmaf(x[2], y[0], -(x[0] * y[2]));
  c[2] = fmaf(x[0], y[1], -(x[1] * y[0]));
}

// Per-Gaussian record of one view (16 floats), identical arithmetic to the GPU preprocess.
void fe_ref_records(int n, const float* means, const float* prec9, const float* wlog, const float* R9,
                    const float* o3, float fx, float fy, float* rec, int* any_degenerate) {
  int deg_any = 0;
  for (int j = 0; j < n; ++j) {
    const float* mu = means + 3 * (size_t)j;
    const float* P = prec9 + 9 * (size_t)j;
    const float u00 = P[0], u01 = P[1], u02 = P[2], u11 = P[4], u12 = P[5], u22 = P[8];
    const float w = wlog[j];
    const float p0 = mu[0] - o3[0], p1 = mu[1] - o3[1], p2 = mu[2] - o3[2];
    float M[3][3];
    for (int k = 0; k < 3; ++k) {
      const float r0 = R9[3 * k], r1 = R9[3 * k + 1], r2 = R9[3 * k + 2];
      M[0][k] = u00 * r0;
      M[1][k] = fmaf(u01, r0, u11 * r1);
      M[2][k] = fmaf(u02, r0, fmaf(u12, r1, u22 * r2));
    }
    float* out = rec + 16 * (size_t)j;
    for (int i = 0; i < 3; ++i) {
      out[i] = M[i][0] / fx;
      out[3 + i] = M[i][1] / fy;
      out[6 + i] = M[i][2];
    }
    out[9] = u00 * p0;
    out[10] = fmaf(u01, p0, u11 * p1);
    out[11] = fmaf(u02, p0, fmaf(u12, p1, u22 * p2));
    out[12] = w;
    out[13] = w * kLog2e;
    out[14] = dot3f(out + 9, out + 9);
    bool finite = true;
    for (int i = 0; i < 15; ++i) finite = finite && isfinite(out[i]);
    const int deg = (!finite || u00 == 0.0f || u11 == 0.0f || u22 == 0.0f) ? 1 : 0;
    int32_t bits = deg;
    memcpy(&out[15], &bits, 4);
    deg_any |= deg;
  }
  if (any_degenerate) *any_degenerate = deg_any;
}

struct Bound {
  float sd_lo, sd_hi, t_lb, t_any_lb, t_ub;
  float d0m_lo, negt_hi;
  bool plus, unbounded, can_mask;
};

static Bound bound_of(const float* r, float Xc, float Yc, float hx, float hy) {
  const float* a = r;
  const float* b = r + 3;
  const float* g = r + 6;
  const float* q = r + 9;
  const float qq = r[14];
  float uc[3];
  for (int i = 0; i < 3; ++i) uc[i] = fmaf(a[i], Xc, fmaf(b[i], Yc, g[i]));
  const float uu = dot3f(uc, uc);
  const float qu = dot3f(q, uc);
  const float tc = qu / uu;
  float v[3];
  for (int i = 0; i < 3; ++i) v[i] = fmaf(tc, uc[i], -q[i]);
  const float sdc = sqrtf(dot3f(v, v));
  const float ua = dot3f(uc, a);
  const float ub = dot3f(uc, b);
  float ca[3], cb[3], pp[3], pm[3], ap[3], am[3];
  crossf(uc, a, ca);
  crossf(uc, b, cb);
  for (int i = 0; i < 3; ++i) {
    pp[i] = fmaf(ca[i], hx, cb[i] * hy);
    pm[i] = fmaf(ca[i], hx, -(cb[i] * hy));
    ap[i] = fmaf(a[i], hx, b[i] * hy);
    am[i] = fmaf(a[i], hx, -(b[i] * hy));
  }
  const float nmax = sqrtf(fmaxf(dot3f(pp, pp), dot3f(pm, pm)));   // = max of the two roots
  const float den = fmaf(-fabsf(ub), hy, fmaf(-fabsf(ua), hx, uu));
  Bound o;
  o.unbounded = !(den > 0.0f);
  const float tanm = nmax / den;
  const float qn = sqrtf(qq);
  const float dth = qn * tanm;
  o.sd_lo = fmaxf(sdc - dth, 0.0f);
  o.sd_hi = sdc + dth;
  const float rho = sqrtf(fmaxf(dot3f(ap, ap), dot3f(am, am)));
  const float un = sqrtf(uu);
  o.plus = (qu > 0.0f) && (o.sd_hi < qn);
  o.t_lb = o.plus ? sqrtf(fmaxf(fmaf(-o.sd_hi, o.sd_hi, qq), 0.0f)) / (un + rho) : 0.0f;
  o.t_any_lb = (un > rho) ? -(sqrtf(fmaxf(fmaf(-o.sd_lo, o.sd_lo, qq), 0.0f)) / (un - rho)) : -INFINITY;
  o.t_ub = (un > rho) ? fminf(qn, qu / un + dth) / (un - rho) : INFINITY;
  // masked part of the tile cone (theta >= 90 deg, theta <= theta_c + phi)
  const float cth = qu / (qn * un);
  const float sth = sdc / qn;
  const float cph = 1.0f / sqrtf(fmaf(tanm, tanm, 1.0f));
  const float sph = tanm * cph;
  const float cmax = fmaf(cth, cph, -(sth * sph));
  const float smax = fmaf(sth, cph, cth * sph);
  o.can_mask = !(cmax > 0.0f);
  o.d0m_lo = (smax > 0.0f) ? qq * (smax * smax) : 0.0f;
  o.negt_hi = (un > rho) ? qn * ((smax > 0.0f) ? fminf(-cmax, 1.0f) : 1.0f) / (un - rho) : INFINITY;
  if (!(o.sd_lo == o.sd_lo)) o.unbounded = true;
  return o;
}

// Upper bound (log2 units) of the depth logit of a pair on the tile rays where it is masked.
static float masked_logit_ub(const Bound& o, float w, float b3l, float b2l) {
  const float dlo = fmaf(-0.5f * o.sd_lo, o.sd_lo, w);
  const float u1 = fmaf(b3l, dlo, -(b2l * o.t_any_lb));
  const float u2 = o.can_mask ? fmaf(b3l, fmaf(-0.5f, o.d0m_lo, w), b2l * o.negt_hi) : -INFINITY;
  return fminf(u1, u2);
}

// Tile lists of one view.  phase 1: alpha bound.  phase 2 (tile_lmin != NULL): pairs not in
// phase 1 whose depth-logit bound reaches tile_lmin[t] - delta2; a tile is completed with every
// pair not in phase 1 when the largest masked-logit bound of those pairs, or the phase-1 floor
// seed tile_seed[t] (NULL: none), reaches tile_lmin[t] + floor_thr.  Membership is pre-tested
// on groups of sf x sf tiles (sf = 4, 3, 1 for ts = 8, 16, 32): a pair the group bound
// rejects is rejected for every tile of the group.  Writes counts[T] and the concatenated,
// per-tile ascending Gaussian ids to `list` (capacity n*T); returns the total.
struct Geo {
  float Xc, Yc, hx, hy;
};

static Geo geo_of(int tx, int ty, int ts, int W, int H, float cx, float cy) {
  const int x0 = tx * ts, y0 = ty * ts;
  const int xl = std::min(x0 + ts, W) - 1, yl = std::min(y0 + ts, H) - 1;
  Geo g;
  g.Xc = (float)(x0 + xl) * 0.5f - cx;
  g.Yc = (float)(y0 + yl) * 0.5f - cy;
  g.hx = (float)(xl - x0) * 0.5f;
  g.hy = (float)(yl - y0) * 0.5f;
  return g;
}

}  // extern "C"

// Shared body of the grid and ray-list membership: tx x ty tiles, groups of sf x sf tiles,
// geo(txi, tyi) the tile's rectangle and ggeo(sx, sy) the group's.
template <typename GeoT, typename GroupGeoT>
static int64_t tiles_impl(int n, const float* rec, int any_degenerate, int tx, int ty, int sf, GeoT geo,
                          GroupGeoT ggeo, float beta2, float beta3, float c_alpha, float delta2, float floor_thr,
                          const float* tile_lmin, const float* tile_seed, uint32_t* counts, uint32_t* list) {
  const float b3l = beta3 * kLog2e, b2l = beta2 * kLog2e;
  // the bounds hold for finite beta2 >= 0 and beta3 >= 0; other views list every pair
  if (!(beta3 >= 0.0f && beta3 <= 3.0e38f && beta2 >= 0.0f && beta2 <= 3.0e38f)) any_degenerate = 1;
  int64_t pos = 0;
  std::vector<char> in1(n > 0 ? n : 1, 0);
  // phase-1 thresholds, with the bound-based replacement for tiles whose value is -inf
  std::vector<float> lm;
  if (tile_lmin) {
    lm.assign(tile_lmin, tile_lmin + tx * ty);
    if (!any_degenerate) {
      for (int tile = 0; tile < tx * ty; ++tile) {
        if (!(lm[tile] == -INFINITY)) continue;
        const Geo g = geo(tile % tx, tile / tx);
        float lb = -INFINITY;
        for (int j = 0; j < n; ++j) {
          const float* r = rec + 16 * (size_t)j;
          int32_t fl;
          memcpy(&fl, &r[15], 4);
          if (fl & 1) continue;
          const Bound o = bound_of(r, g.Xc, g.Yc, g.hx, g.hy);
          if (o.plus && !o.unbounded)
            lb = fmaxf(lb, fmaf(b3l, fmaf(-0.5f * o.sd_hi, o.sd_hi, r[12]), -(b2l * o.t_ub)));
        }
        lm[tile] = lb;
      }
    }
    tile_lmin = lm.data();
  }
  for (int tile = 0; tile < tx * ty; ++tile) {
    const int txi = tile % tx, tyi = tile / tx;
    const Geo g = geo(txi, tyi);
    // group of this tile
    const int sx = txi / sf, sy = tyi / sf;
    const Geo gg = ggeo(sx, sy);
    float tau_grp = INFINITY;
    if (tile_lmin) {
      const int nsx = std::min(sf, tx - sx * sf), nsy = std::min(sf, ty - sy * sf);
      for (int k = 0; k < nsx * nsy; ++k) {
        const int tk = (sy * sf + k / nsx) * tx + (sx * sf + k % nsx);
        tau_grp = fminf(tau_grp, tile_lmin[tk] - delta2);
      }
    }
    uint32_t c = 0;
    const int64_t start = pos;
    // largest logit bound, where masked, of the pairs absent from phase 1, started from the
    // phase-1 floor seed of the tile (+inf when the floor is already visible)
    float ubm = (tile_lmin && tile_seed) ? tile_seed[tile] : -INFINITY;
    for (int j = 0; j < n; ++j) {
      const float* r = rec + 16 * (size_t)j;
      int32_t fl;
      memcpy(&fl, &r[15], 4);
      const bool forced = any_degenerate || (fl & 1) != 0;
      const float w = r[12];
      bool k1 = forced, keep2 = false;
      if (!forced) {
        const Bound og = bound_of(r, gg.Xc, gg.Yc, gg.hx, gg.hy);
        const bool g1 = og.unbounded || (fmaf(0.5f * og.sd_lo, og.sd_lo, -w) <= c_alpha);
        bool sub_done = false;
        Bound o;
        if (g1) {
          o = bound_of(r, g.Xc, g.Yc, g.hx, g.hy);
          sub_done = true;
          k1 = o.unbounded || (fmaf(0.5f * o.sd_lo, o.sd_lo, -w) <= c_alpha);
        }
        if (tile_lmin && !k1) {
          const float dlo_g = fmaf(-0.5f * og.sd_lo, og.sd_lo, w);
          const float ub_g = fmaf(b3l, dlo_g, -(b2l * og.t_lb));
          const bool g2 = og.unbounded || (ub_g >= tau_grp);
          if (g2) {
            if (!sub_done) o = bound_of(r, g.Xc, g.Yc, g.hx, g.hy);
            const float tau = tile_lmin[tile] - delta2;
            const float dlo = fmaf(-0.5f * o.sd_lo, o.sd_lo, w);
            const float ub = fmaf(b3l, dlo, -(b2l * o.t_lb));
            keep2 = o.unbounded || (ub >= tau);
            if (!o.plus) ubm = fmaxf(ubm, masked_logit_ub(o, w, b3l, b2l));
          } else if (!og.plus) {
            ubm = fmaxf(ubm, masked_logit_ub(og, w, b3l, b2l));
          }
        }
      }
      in1[j] = k1;
      const bool keep = tile_lmin ? keep2 : k1;
      if (keep) {
        list[pos++] = (uint32_t)j;
        ++c;
      }
    }
    if (tile_lmin && (ubm - tile_lmin[tile] >= floor_thr)) {
      // the weight floor could matter: complete the tile with every pair not in phase 1
      pos = start;
      c = 0;
      for (int j = 0; j < n; ++j)
        if (!in1[j]) {
          list[pos++] = (uint32_t)j;
          ++c;
        }
    }
    counts[tile] = c;
  }
  return pos;
}

extern "C" {

int64_t fe_ref_tiles(int n, const float* rec, int any_degenerate, int H, int W, int ts, float cx, float cy,
                     float beta2, float beta3, float c_alpha, float delta2, float floor_thr,
                     const float* tile_lmin, const float* tile_seed, uint32_t* counts, uint32_t* list) {
  const int tx = (W + ts - 1) / ts, ty = (H + ts - 1) / ts;
  const int sf = (ts == 8) ? 4 : (ts == 16 ? 3 : 1);
  return tiles_impl(
      n, rec, any_degenerate, tx, ty, sf, [&](int a, int b) { return geo_of(a, b, ts, W, H, cx, cy); },
      [&](int a, int b) { return geo_of(a, b, ts * sf, W, H, cx, cy); }, beta2, beta3, c_alpha, delta2, floor_thr,
      tile_lmin, tile_seed, counts, list);
}

// Ray-list views: tile t holds rays [t*ts*ts, (t+1)*ts*ts) of (x', y') camera-plane
// coordinates; its rectangle is the rays' bounding box widened by 2^-20 of the largest
// coordinate (the GPU's ray_rect, same float operations).
// Check of the masked-logit bound: for every (tile, Gaussian) of a pixel-grid view, the exact
// (float64) logit b3l*d2 - b2l*t* at every pixel of the tile where t* <= 0 must not exceed
// masked_logit_ub of the tile's bound (with a relative slack of 1e-5 for the bound's own
// float32 rounding).  Returns the number of violating (tile, Gaussian) pairs; *checked gets
// the number of (pixel, Gaussian) masked evaluations compared.
int64_t fe_ref_masked_bound_check(int n, const float* rec, int H, int W, int ts, float cx, float cy, float beta2,
                                  float beta3, int64_t* checked) {
  const float b3l = beta3 * kLog2e, b2l = beta2 * kLog2e;
  const int tx = (W + ts - 1) / ts, ty = (H + ts - 1) / ts;
  int64_t bad = 0, cnt = 0;
#pragma omp parallel for reduction(+ : bad, cnt) schedule(dynamic)
  for (int tile = 0; tile < tx * ty; ++tile) {
    const int txi = tile % tx, tyi = tile / tx;
    const Geo g = geo_of(txi, tyi, ts, W, H, cx, cy);
    for (int j = 0; j < n; ++j) {
      const float* r = rec + 16 * (size_t)j;
      const Bound o = bound_of(r, g.Xc, g.Yc, g.hx, g.hy);
      if (o.unbounded) continue;
      const float ub = masked_logit_ub(o, r[12], b3l, b2l);
      double worst = -INFINITY;
      for (int y = tyi * ts; y < std::min(H, (tyi + 1) * ts); ++y)
        for (int x = txi * ts; x < std::min(W, (txi + 1) * ts); ++x) {
          const double xf = (double)x - cx, yf = (double)y - cy;
          double u[3], qu = 0.0, uu = 0.0;
          for (int i = 0; i < 3; ++i) {
            u[i] = (double)r[i] * xf + (double)r[3 + i] * yf + (double)r[6 + i];
            qu += (double)r[9 + i] * u[i];
            uu += u[i] * u[i];
          }
          const double t = qu / uu;
          if (!(t <= 0.0)) continue;
          double d0 = 0.0;
          for (int i = 0; i < 3; ++i) {
            const double v = t * u[i] - (double)r[9 + i];
            d0 += v * v;
          }
          const double l = (double)b3l * ((double)r[12] - 0.5 * d0) - (double)b2l * t;
          worst = std::max(worst, l);
          ++cnt;
        }
      if (worst > (double)ub + 1e-5 * std::fabs((double)ub) + 1e-4) ++bad;
    }
  }
  if (checked) *checked = cnt;
  return bad;
}

void fe_ref_ray_rects(int64_t n_rays, int ts, const float* xy, float* rects) {
  const int64_t tt = (int64_t)ts * ts;
  const int64_t T = (n_rays + tt - 1) / tt;
  for (int64_t t = 0; t < T; ++t) {
    float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (int64_t k = t * tt; k < std::min(n_rays, (t + 1) * tt); ++k) {
      xmin = fminf(xmin, xy[2 * k]); xmax = fmaxf(xmax, xy[2 * k]);
      ymin = fminf(ymin, xy[2 * k + 1]); ymax = fmaxf(ymax, xy[2 * k + 1]);
    }
    const float mx = fmaxf(fabsf(xmin), fabsf(xmax)) * 9.5367431640625e-07f;
    const float my = fmaxf(fabsf(ymin), fabsf(ymax)) * 9.5367431640625e-07f;
    rects[4 * t] = 0.5f * (xmin + xmax);
    rects[4 * t + 1] = 0.5f * (ymin + ymax);
    rects[4 * t + 2] = 0.5f * (xmax - xmin) + mx;
    rects[4 * t + 3] = 0.5f * (ymax - ymin) + my;
  }
}

int64_t fe_ref_tiles_rects(int n, const float* rec, int any_degenerate, int64_t n_tiles, const float* rects,
                           float beta2, float beta3, float c_alpha, float delta2, float floor_thr,
                           const float* tile_lmin, const float* tile_seed, uint32_t* counts, uint32_t* list) {
  auto g = [&](int, int b) {
    Geo o;
    o.Xc = rects[4 * (int64_t)b]; o.Yc = rects[4 * (int64_t)b + 1];
    o.hx = rects[4 * (int64_t)b + 2]; o.hy = rects[4 * (int64_t)b + 3];
    return o;
  };
  return tiles_impl(n, rec, any_degenerate, 1, (int)n_tiles, 1, g, g, beta2, beta3, c_alpha, delta2, floor_thr,
                    tile_lmin, tile_seed, counts, list);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// 4. Record order of one view (the renderer's k_order), restated from the inputs.  Gaussians
//    are ordered within consecutive chunks of 1024 by (code, index), where code is a 20-bit
//    Z-order (Morton) code of the projected centre: x = R p, the camera-frame position of
//    p = mu - o (fma order x = R0 p0 + (R1 p1 + (R2 p2))), then for z > 0 the fractions
//    u = clamp(sx (x/z) + ox, 0, 1) and v likewise over the window [-W/2, 3W/2] of a pixel grid
//    (sx = fx / 2W, ox = (cx + W/2) / 2W), or [-2, 2] of the camera plane for a ray list
//    (sx = 1/4, ox = 1/2); 10 bits each, trunc(u * 1023), x bits at even positions.  Centres
//    not in front of the camera get the largest code.  A NaN fraction clamps to 0.
// ------------------------------------------------------------------------------------------
static uint32_t morton2(uint32_t ix, uint32_t iy) {
  uint32_t c = 0;
  for (int b = 0; b < 10; ++b) c |= (((ix >> b) & 1u) << (2 * b)) | (((iy >> b) & 1u) << (2 * b + 1));
  return c;
}

extern "C" void fe_ref_order(int n, const float* means, const float* R9, const float* o3, float fx, float fy,
                             float cx, float cy, int H, int W, int ray_view, int32_t* order) {
  const float sx = ray_view ? 0.25f : fx / (2.0f * (float)W);
  const float sy = ray_view ? 0.25f : fy / (2.0f * (float)H);
  const float ox = ray_view ? 0.5f : (cx + 0.5f * (float)W) / (2.0f * (float)W);
  const float oy = ray_view ? 0.5f : (cy + 0.5f * (float)H) / (2.0f * (float)H);
  std::vector<std::pair<uint32_t, int32_t>> keyed((size_t)std::max(n, 0));
  for (int g = 0; g < n; ++g) {
    const float* mu = means + 3 * (size_t)g;
    const float p0 = mu[0] - o3[0], p1 = mu[1] - o3[1], p2 = mu[2] - o3[2];
    const float x = fmaf(R9[0], p0, fmaf(R9[1], p1, R9[2] * p2));
    const float y = fmaf(R9[3], p0, fmaf(R9[4], p1, R9[5] * p2));
    const float z = fmaf(R9[6], p0, fmaf(R9[7], p1, R9[8] * p2));
    uint32_t code = 0xFFFFFu;
    if (z > 0.0f) {
      const float u = fminf(fmaxf(fmaf(sx, x / z, ox), 0.0f), 1.0f);
      const float v = fminf(fmaxf(fmaf(sy, y / z, oy), 0.0f), 1.0f);
      code = morton2((uint32_t)(u * 1023.0f), (uint32_t)(v * 1023.0f));
    }
    keyed[g] = {code, g};
  }
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int c1 = std::min(n, c0 + 1024);
    std::sort(keyed.begin() + c0, keyed.begin() + c1);   // (code, index): unique keys
    for (int j = c0; j < c1; ++j) order[j] = keyed[j].second;
  }
}
