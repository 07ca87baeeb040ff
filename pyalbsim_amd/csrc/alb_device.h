// alb_device.h — device-side building blocks shared by the trace and gather
// kernels: Philox4x32-10 counter RNG, the per-photon uniform stream, wave
// helpers and the surface / phase-function formulas of the reference
// (src/alb_sim/physics/reflection/*.py, physics/scatter/energy.py).
//
// Arithmetic is compiled with -ffp-contract=off so every product and sum is
// rounded separately, in the order NumPy evaluates the reference expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/alb_abi.h"

#define ALB_PI 3.141592653589793
#define ALB_FLT_EPS 1.1920928955078125e-07  // np.finfo(np.float32).eps

namespace alb {

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)a << 21) | (uint64_t)(b >> 11);
  return (double)m * 0x1.0p-53;
}

// Per-photon stream of FP64 uniforms in [0,1): draw j comes from Philox block
// j/2 (counter = {j/2, photon lo, photon hi, stream}), words (x,y) for even j
// and (z,w) for odd j; or from the caller's replay buffer.
struct Stream {
  uint64_t photon;
  int64_t slot;
  uint32_t draw;
  bool exhausted;
  double cached;
};

__device__ __forceinline__ double uniform(const alb_rng_t& rng, Stream& s) {
  if (rng.mode == ALB_RNG_REPLAY) {
    if ((int64_t)s.draw >= rng.replay_len) {
      s.exhausted = true;
      return 0.5;
    }
    return rng.replay[s.slot * rng.replay_len + (int64_t)(s.draw++)];
  }
  const uint32_t j = s.draw++;
  if (j & 1u) return s.cached;
  const u32x4 r = philox4x32_10(
      u32x4{j >> 1, (uint32_t)s.photon, (uint32_t)(s.photon >> 32), rng.stream}, rng.key[0], rng.key[1]);
  s.cached = u53(r.z, r.w);
  return u53(r.x, r.y);
}

// ------------------------------------------------------------- wave helpers
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// number of set bits of mask below this lane
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t bcast_u64(uint64_t v, int src) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double bcast_f64(double v, int src) {
  return __longlong_as_double((long long)bcast_u64((uint64_t)__double_as_longlong(v), src));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------- vector math
__device__ __forceinline__ double norm3(double x, double y, double z) {
  double s = x * x + y * y;
  s = s + z * z;
  return sqrt(s);
}

// np.interp(x, xp, fp) for increasing xp: clamps, exact nodes, nan fallback
__device__ __forceinline__ double interp(double x, const double* __restrict__ xp,
                                         const double* __restrict__ fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  double r = slope * (x - xp[lo]) + fp[lo];
  if (isnan(r)) {
    r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    if (isnan(r) && fp[lo] == fp[lo + 1]) r = fp[lo];
  }
  return r;
}

// interp() that also reports where x fell: seg = 0 below xp[0] (the clamped
// atom at fp[0]), seg = i + 1 inside [xp[i], xp[i+1]), n - 1 at or above xp[n-1].
// Same arithmetic as interp(), so the returned value is identical.
__device__ __forceinline__ double interp_seg(double x, const double* __restrict__ xp,
                                             const double* __restrict__ fp, int n, int* seg) {
  if (x < xp[0]) {
    *seg = 0;
    return fp[0];
  }
  if (x >= xp[n - 1]) {
    *seg = n - 1;
    return fp[n - 1];
  }
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  *seg = lo + 1;
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  double r = slope * (x - xp[lo]) + fp[lo];
  if (isnan(r)) {
    r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    if (isnan(r) && fp[lo] == fp[lo + 1]) r = fp[lo];
  }
  return r;
}

// np.interp with a float32 grid (converted to FP64 as NumPy does)
__device__ __forceinline__ double interp_f32grid(double x, const float* __restrict__ xp,
                                                 const double* __restrict__ fp, int n) {
  if (x < (double)xp[0]) return fp[0];
  if (x >= (double)xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((double)xp[mid] <= x) lo = mid; else hi = mid;
  }
  if ((double)xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / ((double)xp[lo + 1] - (double)xp[lo]);
  double r = slope * (x - (double)xp[lo]) + fp[lo];
  if (isnan(r)) {
    r = slope * (x - (double)xp[lo + 1]) + fp[lo + 1];
    if (isnan(r) && fp[lo] == fp[lo + 1]) r = fp[lo];
  }
  return r;
}

// ------------------------------------------------------------ layered water
// (src/alb_sim/physics/models/water.py:253-416, physics/models/simulation.py:214-314)

// layer of depth d: the number of layer bottoms <= d (np.searchsorted
// side="right"), clamped to the last layer
__device__ __forceinline__ int layer_of(const alb_layer_t* __restrict__ L, int n, double d) {
  int i = 0;
  while (i < n - 1 && L[i].bottom <= d) ++i;
  return i;
}

// a layer property at local depth z (TurbidityLayerModel._eval with NumPy's
// dtype rules): a linear profile evaluates in float32 with t = float32(z / h),
// an exponential one in FP64 (parameter_profile.py:47-99); any other profile
// (kind 3) interpolates its samples tab[0, tab_n) at t = i / (tab_n - 1)
// linearly in t, in FP64
__device__ __forceinline__ double layer_profile(int kind, double s, double e, double z, double h,
                                                const double* __restrict__ tab, int tab_n) {
  if (kind == 0) return s;
  if (kind == 1) {
    const float t = (float)(z / h);
    const float r = (float)s + t * (float)(e - s);
    return (double)r;
  }
  if (kind == 3) {
    const double t = z / h;
    if (!(t > 0.0)) return tab[0];
    if (t >= 1.0) return tab[tab_n - 1];
    const double u = t * (double)(tab_n - 1);
    int i = (int)u;
    if (i > tab_n - 2) i = tab_n - 2;
    const double f = u - (double)i;
    return tab[i] + f * (tab[i + 1] - tab[i]);
  }
  const double ls = log(s), le = log(e);
  return exp(ls + (z / h) * (le - ls));
}

// speed (float32 value) at local depth z: c / n, in float32 arithmetic when n
// is a float32 profile value, else in FP64 then rounded to float32
__device__ __forceinline__ double layer_velocity(const alb_layer_t& L, double z) {
  if (!L.profiled) return L.v;
  if (L.n_kind == 1) {
    const float n32 = (float)layer_profile(1, L.n0, L.n1, z, L.height, L.n_tab, L.tab_n);
    return (double)(2.998e8f / n32);
  }
  return (double)(float)(2.998e8 / layer_profile(L.n_kind, L.n0, L.n1, z, L.height, L.n_tab, L.tab_n));
}

// lidar attenuation at local depth z: FP64 table value (np.interp at float32(z))
__device__ __forceinline__ double layer_alpha64(const alb_layer_t& L, double z) {
  return interp_f32grid((double)(float)z, L.z_lut, L.alpha_lut, L.lut_n);
}

// float32 attenuation used for the optical depth
__device__ __forceinline__ float layer_alpha32(const alb_layer_t& L, double z) {
  return L.profiled ? (float)layer_alpha64(L, z) : L.alpha;
}

// float32 single-scattering albedo (alpha - a) / alpha at local depth z
__device__ __forceinline__ float layer_omega0(const alb_layer_t& L, double z) {
  if (!L.profiled) return L.omega0;
  const double al = layer_alpha64(L, z);
  const double a = layer_profile(L.a_kind, L.a0, L.a1, z, L.height, L.a_tab, L.tab_n);
  return (float)((al - a) / al);
}

// Phase-bin mass table of the gather's scatter multiplier: the layer of the
// query vertex depth (y + flying height) * -1 in a layered column
// (physics/models/water.py:418-447), else the scene's own table.
__device__ __forceinline__ const double* gather_phase_mass(const alb_scene_t& sc, double qy, int* n) {
  if (!sc.layers) {
    *n = sc.phase_n;
    return sc.phase_mass;
  }
  const alb_layer_t& L = sc.layers[layer_of(sc.layers, sc.n_layers, (qy - sc.y_ws) * -1.0)];
  *n = L.phase_n;
  return L.phase_mass;
}

// np.interp on the theta grid itself (uniform spacing up to rounding): locate
// the node by division and correct by one in either direction.
__device__ __forceinline__ double interp_uniform_grid(double x, const double* __restrict__ xp,
                                                      const double* __restrict__ fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  const double step = (xp[n - 1] - xp[0]) / (double)(n - 1);
  int lo = (int)((x - xp[0]) / step);
  lo = lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
  while (lo > 0 && xp[lo] > x) --lo;
  while (lo < n - 2 && xp[lo + 1] <= x) ++lo;
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  double r = slope * (x - xp[lo]) + fp[lo];
  if (isnan(r)) {
    r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    if (isnan(r) && fp[lo] == fp[lo + 1]) r = fp[lo];
  }
  return r;
}

__device__ __forceinline__ double clipd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// Schlick: F0 + (1 - F0)(1 - cos)^5 (physics/reflection/fresnel.py:14-29)
__device__ __forceinline__ double schlick(double c, double F0) {
  return F0 + (1.0 - F0) * pow(1.0 - c, 5.0);
}

// GGX D with alpha2 = alpha^2 (microfacet_brdf.py:18-39)
__device__ __forceinline__ double ggx_d(double ndh, double alpha) {
  const double a2 = alpha * alpha;
  double den = (a2 - 1.0) * ndh * ndh + 1.0;
  den = ALB_PI * den * den;
  return a2 / (den + ALB_FLT_EPS);
}

// Smith G1 (microfacet_brdf.py:42-67)
__device__ __forceinline__ double smith_g1(double x, double alpha) {
  const double a2 = alpha * alpha;
  const double den = x + sqrt(a2 + (1.0 - a2) * (x * x));
  return 2.0 * x / den;
}

// float32 normalisation of a stored map direction (normalize_batch on float32)
__device__ __forceinline__ void normalize_f32(float& x, float& y, float& z) {
  float s = x * x + y * y;
  s = s + z * z;
  float n = sqrtf(s);
  if (n == 0.0f) n = 1.0f;
  x = x / n;
  y = y / n;
  z = z / n;
}

// x / y in float32 for float32 x and y, given ry = 1 / (double)y (IEEE).  The
// FP64 product is within 2^-52 of the exact quotient, relative: 2^-28 float32
// ulps.  An exact quotient of two float32 values lies at least 2^-25 ulps
// from every rounding midpoint (x = y * m for a midpoint m would make an odd
// 25-bit significand times y's significand fit in 24 bits).  So the product
// rounds to the IEEE float32 quotient.
__device__ __forceinline__ float div_f32_recip(float x, double ry) { return (float)((double)x * ry); }

// normalize_f32 with its three float32 divisions taken through one FP64
// reciprocal (div_f32_recip): the same results
__device__ __forceinline__ void normalize_f32_recip(float& x, float& y, float& z) {
  float s = x * x + y * y;
  s = s + z * z;
  float n = sqrtf(s);
  if (n == 0.0f) n = 1.0f;
  const double rn = 1.0 / (double)n;
  x = div_f32_recip(x, rn);
  y = div_f32_recip(y, rn);
  z = div_f32_recip(z, rn);
}

__device__ __forceinline__ void normalize_f64_keep_zero(double& x, double& y, double& z) {
  const double n = norm3(x, y, z);
  if (n != 0.0) {
    x = x / n;
    y = y / n;
    z = z / n;
  }
}

// Cook-Torrance reflection towards the sensor, float32 result
// (microfacet_brdf.py:96-147 called as sea_surface.py:35-62: wi = -d_p, wo = -d_s),
// with the outgoing direction o = normalised -s given (the gather stages it
// once per sensor vertex)
__device__ __forceinline__ float brdf_reflect_o(float px, float py, float pz, double ox, double oy, double oz,
                                                double alpha, double F0) {
  float wx = -px, wy = -py, wz = -pz;
  normalize_f32_recip(wx, wy, wz);
  const double ndwi = clipd((double)wy, 0.0, 1.0);
  const double ndwo = clipd(oy, 0.0, 1.0);
  if (!(ndwi > 0.0 && ndwo > 0.0)) return 0.0f;
  double hx = (double)wx + ox, hy = (double)wy + oy, hz = (double)wz + oz;
  {
    double n = norm3(hx, hy, hz);
    if (n == 0.0) n = 1.0;
    hx = hx / n;
    hy = hy / n;
    hz = hz / n;
  }
  double widh = (double)wx * hx;
  widh = widh + (double)wy * hy;
  widh = widh + (double)wz * hz;
  widh = clipd(widh, 0.0, 1.0);
  const double ndh = clipd(hy, 0.0, 1.0);
  const double D = ggx_d(ndh, alpha);
  const double G = smith_g1(ndwi, alpha) * smith_g1(ndwo, alpha);
  const double F = schlick(widh, F0);
  const double den = 4.0 * ndwi * ndwo + ALB_FLT_EPS;
  return (float)((D * G * F) / den);
}

__device__ __forceinline__ float brdf_reflect(float px, float py, float pz, double sx, double sy,
                                              double sz, double alpha, double F0) {
  double ox = -sx, oy = -sy, oz = -sz;
  normalize_f64_keep_zero(ox, oy, oz);
  return brdf_reflect_o(px, py, pz, ox, oy, oz, alpha, F0);
}

// brdf_reflect together with d log(BRDF) / d(alpha^2) (0 where the BRDF is 0):
// log D = log a2 - log(pi u^2 + eps) with u = (a2 - 1) c^2 + 1, and
// log G1(x) = log 2x - log(x + q) with q = sqrt(a2 + (1 - a2) x^2).
__device__ __forceinline__ float brdf_reflect_dlog(float px, float py, float pz, double sx, double sy,
                                                   double sz, double alpha, double F0, double* dlog_da2) {
  float wx = -px, wy = -py, wz = -pz;
  normalize_f32(wx, wy, wz);
  double ox = -sx, oy = -sy, oz = -sz;
  normalize_f64_keep_zero(ox, oy, oz);
  double hx = (double)wx + ox, hy = (double)wy + oy, hz = (double)wz + oz;
  {
    double n = norm3(hx, hy, hz);
    if (n == 0.0) n = 1.0;
    hx = hx / n;
    hy = hy / n;
    hz = hz / n;
  }
  const double ndwi = clipd((double)wy, 0.0, 1.0);
  const double ndwo = clipd(oy, 0.0, 1.0);
  double widh = (double)wx * hx;
  widh = widh + (double)wy * hy;
  widh = widh + (double)wz * hz;
  widh = clipd(widh, 0.0, 1.0);
  const double ndh = clipd(hy, 0.0, 1.0);
  *dlog_da2 = 0.0;
  if (!(ndwi > 0.0 && ndwo > 0.0)) return 0.0f;
  const double D = ggx_d(ndh, alpha);
  const double G = smith_g1(ndwi, alpha) * smith_g1(ndwo, alpha);
  const double F = schlick(widh, F0);
  const double den = 4.0 * ndwi * ndwo + ALB_FLT_EPS;
  const double a2 = alpha * alpha;
  const double c2 = ndh * ndh;
  const double u = (a2 - 1.0) * c2 + 1.0;
  double dl = 1.0 / a2 - (2.0 * ALB_PI * u * c2) / (ALB_PI * u * u + ALB_FLT_EPS);
  const double qi = sqrt(a2 + (1.0 - a2) * (ndwi * ndwi));
  const double qo = sqrt(a2 + (1.0 - a2) * (ndwo * ndwo));
  dl -= ((1.0 - ndwi * ndwi) / (2.0 * qi)) / (ndwi + qi);
  dl -= ((1.0 - ndwo * ndwo) / (2.0 * qo)) / (ndwo + qo);
  *dlog_da2 = dl;
  return (float)((D * G * F) / den);
}

// Walter BTDF + surface albedo / pi (microfacet_brdf.py:150-226 called as
// sea_surface.py:64-100 with eta_i = n_water, eta_o = 1)
// Walter BTDF towards the sensor plus the surface albedo / pi
// (microfacet_brdf.py:150-226 called as sea_surface.py:64-100), with the
// outgoing direction o = normalised -s given
__device__ __forceinline__ double btdf_transmit_o(float px, float py, float pz, double ox, double oy, double oz,
                                                  double alpha, double n_w, double F0, double albedo) {
  float wx = px, wy = py, wz = pz;
  normalize_f32_recip(wx, wy, wz);
  const double eta_i = n_w, eta_o = 1.0, eta = eta_i / eta_o;
  const double ndwi = clipd((double)wy, -1.0, 1.0);
  const double ndwo = clipd(oy, -1.0, 1.0);
  double btdf = 0.0;
  if (ndwi * ndwo < 0.0) {  // the half vector only matters here
    double hx = (double)wx + eta * ox, hy = (double)wy + eta * oy, hz = (double)wz + eta * oz;
    {
      double n = norm3(hx, hy, hz);
      if (n == 0.0) n = 1.0;
      hx = hx / n;
      hy = hy / n;
      hz = hz / n;
    }
    const double ndh = clipd(hy, 0.0, 1.0);
    double widh = (double)wx * hx;
    widh = widh + (double)wy * hy;
    widh = widh + (double)wz * hz;
    widh = clipd(widh, 0.0, 1.0);
    double wodh = hx * ox;
    wodh = wodh + hy * oy;
    wodh = wodh + hz * oz;
    wodh = clipd(wodh, 0.0, 1.0);
    const double D = ggx_d(ndh, alpha);
    const double G = smith_g1(fabs(ndwi), alpha) * smith_g1(fabs(ndwo), alpha);
    const double F = schlick(widh, F0);
    const double den = eta_i * widh + eta_o * wodh;
    const double factor = (eta_o * eta_o * wodh) / (den * den + ALB_FLT_EPS);
    btdf = (1.0 - F) * D * G * factor / (fabs(ndwi) * fabs(ndwo) + ALB_FLT_EPS);
  }
  return btdf + albedo / ALB_PI;
}

__device__ __forceinline__ double btdf_transmit(float px, float py, float pz, double sx, double sy,
                                                double sz, double alpha, double n_w, double F0,
                                                double albedo) {
  double ox = -sx, oy = -sy, oz = -sz;
  normalize_f64_keep_zero(ox, oy, oz);
  return btdf_transmit_o(px, py, pz, ox, oy, oz, alpha, n_w, F0, albedo);
}

// Phase-function mass of the scattering-angle bin towards the sensor
// (physics/scatter/energy.py:8-52).  delta is float32: pi - float32 eps
// stays float32 under NumPy 2 promotion.
__device__ __forceinline__ double scatter_mult(float px, float py, float pz, double sx, double sy,
                                               double sz, const double* __restrict__ theta,
                                               const double* __restrict__ cdf, int n) {
  normalize_f32(px, py, pz);
  double ox = -sx, oy = -sy, oz = -sz;
  normalize_f64_keep_zero(ox, oy, oz);
  double c = (double)px * ox;
  c = c + (double)py * oy;
  c = c + (double)pz * oz;
  c = clipd(c, -1.0, 1.0);
  const double ang = acos(c);
  const float delta_f = (3.14159265358979323846f - 1.1920928955078125e-07f) / (float)n;
  const double delta = (double)delta_f;
  long long i = (long long)((ang - ALB_FLT_EPS) / delta);
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  const double lo = (double)i * delta;
  const double hi = (double)(i + 1) * delta;
  return interp_uniform_grid(hi, theta, cdf, n) - interp_uniform_grid(lo, theta, cdf, n);
}

// BOTTOM multiplier (core/simulation.py:155-163): (albedo / pi) max(0, n . (-d_s))
// with n = +y, or the tilted floor's normal
__device__ __forceinline__ double bottom_mult(const alb_scene_t& sc, double sdx, double sdy, double sdz) {
  if (!sc.tilted) return (sc.floor_albedo / ALB_PI) * fmax(0.0, -sdy);
  const double c = (sc.floor_normal[0] * -sdx + sc.floor_normal[1] * -sdy) + sc.floor_normal[2] * -sdz;
  return (sc.floor_albedo / ALB_PI) * fmax(0.0, c);
}

// ------------------------------------------------------------- heat maps
// accumulate_to_heatmap (src/alb_sim/math/accumulate_heatmap.py:6-61) for one
// point: bins of (x - centre + extent) / (2 extent / bins) along x and z,
// truncated toward zero as astype(int64) does; NaN points and bins outside
// [0, bins) are skipped
__device__ __forceinline__ void heat_add(double* __restrict__ h, float px, float pz, const double* c, double extent,
                                         int bins, double e) {
  if (isnan(px)) return;
  const double bs = 2.0 * extent / (double)bins;
  const double vx = (((double)px - c[0]) + extent) / bs;
  const double vz = (((double)pz - c[2]) + extent) / bs;
  if (!(vx > -1.0 && vx < (double)bins && vz > -1.0 && vz < (double)bins)) return;
  const int64_t ix = (int64_t)vx, iz = (int64_t)vz;
  unsafeAtomicAdd(&h[ix * bins + iz], e);
}

// ------------------------------------------------------------- tree nodes
// 64-byte internal node of the photon-map radix tree (layout in alb_abi.h)
struct __attribute__((aligned(16))) Node {
  int32_t first, last, split, leaf_mask;
  float lb[6];  // left child box: min x,y,z, max x,y,z
  float rb[6];  // right child box
};
static_assert(sizeof(Node) == 64, "tree node must be one 64-byte line");

// The build writes child boxes only into the root and into nodes covering
// more than kBoxedRecords records: the gathers scan any subtree of at most
// that many records as one record range, so they never read its node's boxes.
// first/last/split/leaf_mask are written for every node.
constexpr int kBoxedRecords = 64;

// squared distance in FP64 from a query to a float32 record, terms summed x,
// y, z in this order (the order the oracle and the neighbour-set proofs use)
__device__ __forceinline__ double dist2(double qx, double qy, double qz, float x, float y, float z) {
  const double dx = (double)x - qx, dy = (double)y - qy, dz = (double)z - qz;
  double s = dx * dx + dy * dy;
  return s + dz * dz;
}

// distance from q to [lo, hi] along one axis; the same values as the branchy
// form (lo - q below, q - hi above, 0 inside) without divergent branches
__device__ __forceinline__ double axis_gap(double q, float lo, float hi) {
  return fmax(fmax((double)lo - q, q - (double)hi), 0.0);
}

// squared distance from q to a float32 box (min x,y,z, max x,y,z); never larger
// than dist2 of a record inside the box (monotone rounding)
__device__ __forceinline__ double box_dist2(double qx, double qy, double qz, const float* __restrict__ b) {
  const double gx = axis_gap(qx, b[0], b[3]);
  const double gy = axis_gap(qy, b[1], b[4]);
  const double gz = axis_gap(qz, b[2], b[5]);
  double s = gx * gx + gy * gy;
  return s + gz * gz;
}

// squared distance from q to the farthest corner of a box; never smaller than
// dist2 of a record inside the box (monotone rounding)
__device__ __forceinline__ double box_far2(double qx, double qy, double qz, const float* __restrict__ b) {
  const double gx = fmax(fabs((double)b[0] - qx), fabs((double)b[3] - qx));
  const double gy = fmax(fabs((double)b[1] - qy), fabs((double)b[4] - qy));
  const double gz = fmax(fabs((double)b[2] - qz), fabs((double)b[5] - qz));
  double s = gx * gx + gy * gy;
  return s + gz * gz;
}

// ----------------------------------------------------------- Morton keys
__device__ __forceinline__ uint64_t spread2(uint64_t v) {  // 32 bits -> even bit positions
  v &= 0xFFFFFFFFull;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

__device__ __forceinline__ uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// float32 <-> unsigned keys that order like the floats (bounding boxes by
// integer atomicMin / atomicMax)
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ uint64_t quant(double p, double lo, double inv, double maxc) {
  double c = floor((p - lo) * inv);
  c = c < 0.0 ? 0.0 : (c > maxc ? maxc : c);
  return (uint64_t)c;
}

// Morton key of a point on the map grid described by meta (origin, inverse
// cell width, dims, bits per axis); used by the map build and the gather
// bits that hold scene numbers 0 .. n_scenes-1 in scene-major keys (0 for one scene)
__host__ __device__ inline int scene_bits(int n_scenes) {
  int b = 0;
  while ((1 << b) < n_scenes) ++b;
  return b;
}

__device__ __forceinline__ uint64_t morton_key(double px, double py, double pz, const double* meta) {
  const int dims = (int)meta[6];
  const double maxc = ldexp(1.0, (int)meta[7]) - 1.0;
  const uint64_t qx = quant(px, meta[0], meta[3], maxc);
  const uint64_t qz = quant(pz, meta[2], meta[5], maxc);
  if (dims == 2) return (spread2(qx) << 1) | spread2(qz);
  const uint64_t qy = quant(py, meta[1], meta[4], maxc);
  return (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
}

// width of the gather's scattering-angle bins, (pi - eps) / n in float32
// (physics/scatter/energy.py:41-47)
__device__ __forceinline__ double scatter_bin_width(int n) {
  return (double)((3.14159265358979323846f - 1.1920928955078125e-07f) / (float)n);
}

// Scattering-angle bin towards the sensor, int((acos(c) - eps) / delta), with
// the outgoing direction o = normalised -s given, delta = scatter_bin_width(n)
// and rdelta = 1 / delta.  The quotient comes from the reciprocal: it is
// within 2^-37 of the FP64 quotient (q < 2^15), so the two truncate alike
// unless an integer lies within 2^-32 of it, and then the quotient is taken.
__device__ __forceinline__ int scatter_bin_o(float px, float py, float pz, double ox, double oy, double oz, int n,
                                             double delta, double rdelta) {
  normalize_f32_recip(px, py, pz);
  double c = (double)px * ox;
  c = c + (double)py * oy;
  c = c + (double)pz * oz;
  c = clipd(c, -1.0, 1.0);
  const double x = acos(c) - ALB_FLT_EPS;
  double q = x * rdelta;
  const double fr = q - floor(q);
  if (fr < 0x1p-32 || fr > 1.0 - 0x1p-32) q = x / delta;
  long long i = (long long)q;
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  return (int)i;
}

__device__ __forceinline__ int scatter_bin(float px, float py, float pz, double sx, double sy, double sz, int n) {
  normalize_f32(px, py, pz);
  double ox = -sx, oy = -sy, oz = -sz;
  normalize_f64_keep_zero(ox, oy, oz);
  double c = (double)px * ox;
  c = c + (double)py * oy;
  c = c + (double)pz * oz;
  c = clipd(c, -1.0, 1.0);
  const double ang = acos(c);
  const float delta_f = (3.14159265358979323846f - 1.1920928955078125e-07f) / (float)n;
  long long i = (long long)((ang - ALB_FLT_EPS) / (double)delta_f);
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  return (int)i;
}

__device__ __forceinline__ double scatter_mass(float px, float py, float pz, double sx, double sy, double sz,
                                               const double* __restrict__ mass, int n) {
  return mass[scatter_bin(px, py, pz, sx, sy, sz, n)];
}

}  // namespace alb
