/*
 * alb_oracle.c — CPU oracle for the photon-tracing / photon-mapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pyalbsim_amd, the HIP
 * library) links, loads or calls this file; only tests/, the smoke check and
 * the cpu_baseline leg of bench.py do, as the checker.
 *
 * It restates PyALBSim's algorithm literally: every photon is advanced by one
 * fixed time step per iteration, exactly as Simulation.simulate_batch does for
 * the whole batch (src/alb_sim/core/simulation.py:71-107), with the reference's
 * FP64 positions / FP32 energies and optical depths, its state classification
 * and its handler formulas.  The HIP trace solves straight legs in closed form
 * instead; agreement of the two is what the parity tests check.
 *
 * Random numbers: a per-photon stream of FP64 uniforms, either replayed from a
 * host buffer or drawn from Philox4x32-10 (Salmon et al., SC'11), consumed in
 * per-photon event order:
 *   forward emission  : theta (Rayleigh, rejected above the divergence), phi
 *   backward emission : disk radius, disk angle, cone z, cone phi
 *   enter: tau_target; scatter: theta, phi, tau_target; exit: roulette;
 *   floor: two uniforms of the cosine-hemisphere map.
 *
 * Each function cites the reference lines it restates.  Pinning status of each
 * function (what independent check covers it) is listed with the tests in
 * tests/test_oracle_*.py; every function here is pinned by at least one of:
 * Philox known-answer vectors, closed forms, distribution tests, a ballistic
 * known-answer scene, brute-force / scipy cross-checks, an FP64 re-sum.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off (see oracle/Makefile); no
 * fused multiply-add, so every product and sum rounds as NumPy's do.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/alb_abi.h"

#define ORC_PI 3.141592653589793
#define ORC_FLT_EPS 1.1920928955078125e-07 /* np.finfo(np.float32).eps */

/* ------------------------------------------------------------------ Philox */

static inline void orc_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

/* Philox4x32 with 10 rounds, Salmon, Moraes, Dror & Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57,
 * Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    orc_mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    orc_mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct {
  const alb_rng_t* rng;
  uint64_t photon;   /* global photon index (Philox counter words 1,2) */
  int64_t slot;      /* local photon index (replay slot)               */
  uint64_t draw;     /* next draw index                                */
  int exhausted;
} orc_stream;

/* Uniform in [0,1) with 53 random bits: draw j uses Philox block j/2, words
 * (0,1) for even j and (2,3) for odd j: u = ((a << 21) | (b >> 11)) * 2^-53. */
static double orc_uniform(orc_stream* s) {
  if (s->rng->mode == ALB_RNG_REPLAY) {
    if ((int64_t)s->draw >= s->rng->replay_len) {
      s->exhausted = 1;
      return 0.5;
    }
    return s->rng->replay[s->slot * s->rng->replay_len + (int64_t)(s->draw++)];
  }
  uint64_t j = s->draw++;
  uint32_t ctr[4] = {(uint32_t)(j >> 1), (uint32_t)s->photon, (uint32_t)(s->photon >> 32), s->rng->stream};
  uint32_t out[4];
  orc_philox4x32_10(ctr, s->rng->key, out);
  uint32_t a = (j & 1) ? out[2] : out[0];
  uint32_t b = (j & 1) ? out[3] : out[1];
  uint64_t m = ((uint64_t)a << 21) | (uint64_t)(b >> 11);
  return (double)m * 0x1.0p-53;
}

double orc_uniform_at(const alb_rng_t* rng, uint64_t photon, uint64_t draw) {
  orc_stream s = {rng, photon, 0, draw, 0};
  return orc_uniform(&s);
}

/* -------------------------------------------------------------- vector math */

/* np.linalg.norm(v, axis=1) on float64: sqrt((x*x + y*y) + z*z)
 * (src/alb_sim/math/vector_math.py:22-24) */
static inline double orc_norm3(const double v[3]) {
  double s = v[0] * v[0] + v[1] * v[1];
  s = s + v[2] * v[2];
  return sqrt(s);
}

/* normalize_batch (src/alb_sim/math/vector_math.py:14-19): zero norm -> 1 */
static inline void orc_normalize3(double v[3]) {
  double n = orc_norm3(v);
  if (n == 0.0) n = 1.0;
  v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n;
}

/* normalize_batch on a float32 array: the whole computation stays float32 */
static inline void orc_normalize3f(float v[3]) {
  float s = v[0] * v[0] + v[1] * v[1];
  s = s + v[2] * v[2];
  float n = sqrtf(s);
  if (n == 0.0f) n = 1.0f;
  v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n;
}

/* local @ R.T with R float32, result float64: d_i = sum_j local_j * R[i][j]
 * (src/alb_sim/math/direction_cone.py:46-50, 112-117) */
static inline void orc_rotate_f32(const float R[9], const double l[3], double out[3]) {
  for (int i = 0; i < 3; ++i) {
    double s = l[0] * (double)R[3 * i + 0];
    s = s + l[1] * (double)R[3 * i + 1];
    s = s + l[2] * (double)R[3 * i + 2];
    out[i] = s;
  }
}

/* |a - b| <= atol + rtol*|b| with numpy's defaults (np.isclose) */
static inline int orc_isclose(double a, double b) {
  return fabs(a - b) <= (1e-08 + 1e-05 * fabs(b));
}

/* rotation_from_z_batch for one vector, cast to float32
 * (src/alb_sim/math/rotation_matrix.py:39-102) */
void orc_rotation_from_z(const double vin[3], float R[9]) {
  double v[3] = {vin[0], vin[1], vin[2]};
  orc_normalize3(v);
  double c = v[2];
  double vx = -v[1], vy = v[0], vz = 0.0;
  double s2 = vx * vx + vy * vy;
  double M[9];
  if (orc_isclose(c, 1.0)) {
    double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(M, I, sizeof M);
  } else if (orc_isclose(c, -1.0)) {
    double F[9] = {1, 0, 0, 0, -1, 0, 0, 0, -1};
    memcpy(M, F, sizeof M);
  } else {
    double K[9] = {0.0, -vz, vy, vz, 0.0, -vx, -vy, vx, 0.0};
    double f = (1.0 - c) / s2;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double kk = K[3 * i + 0] * K[0 * 3 + j];
        kk = kk + K[3 * i + 1] * K[1 * 3 + j];
        kk = kk + K[3 * i + 2] * K[2 * 3 + j];
        double id = (i == j) ? 1.0 : 0.0;
        M[3 * i + j] = (id + K[3 * i + j]) + kk * f;
      }
  }
  for (int i = 0; i < 9; ++i) R[i] = (float)M[i];
}

/* np.interp(x, xp, fp) with numpy's clamping and exact-node rules */
double orc_interp(double x, const double* xp, const double* fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x > xp[n - 1]) return fp[n - 1];
  if (x == xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1; /* xp[lo] <= x < xp[hi] */
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  double r = slope * (x - xp[lo]) + fp[lo];
  if (isnan(r)) {
    r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    if (isnan(r) && fp[lo] == fp[lo + 1]) r = fp[lo];
  }
  return r;
}

/* ------------------------------------------------------- layered water
 * Per-depth lookups of WaterModel (src/alb_sim/physics/models/water.py:253-416)
 * and TurbidityLayerModel (water.py:39-250), with NumPy's dtype rules. */

/* layer_index (water.py:283-286): searchsorted(cumsum, d, side="right"),
 * clamped to the last layer */
static int orc_layer_index(const alb_scene_t* sc, double d) {
  int idx = 0;
  for (int i = 0; i < sc->n_layers; ++i)
    if (sc->layers[i].bottom <= d) idx = i + 1;
  return idx < sc->n_layers - 1 ? idx : sc->n_layers - 1;
}

/* _eval of a profile at t = z / h (utils/parameter_profile.py:47-99):
 * LinearParameter casts t to float32 and computes
 * float32(start) + t * float32(end - start) in float32; ExponentialParameter
 * exp(log s + t (log e - log s)) in FP64 */
static double orc_profile(int kind, double s, double e, double z, double h, const double* tab, int tab_n) {
  if (kind == 0) return s;
  double t = z / h;
  if (kind == 1) {
    float tf = (float)t;
    float span = (float)(e - s);
    float prod = tf * span;
    float r = (float)s + prod;
    return (double)r;
  }
  if (kind == 3) {
    /* any other profile: linear interpolation of its samples at
     * t = i / (tab_n - 1), FP64 */
    if (!(t > 0.0)) return tab[0];
    if (t >= 1.0) return tab[tab_n - 1];
    double u = t * (double)(tab_n - 1);
    int i = (int)u;
    if (i > tab_n - 2) i = tab_n - 2;
    double f = u - (double)i;
    return tab[i] + f * (tab[i + 1] - tab[i]);
  }
  double ls = log(s), le = log(e);
  return exp(ls + t * (le - ls));
}

/* np.interp at float32(z) over the float32 z grid (water.py:210-229) */
static double orc_layer_alpha64(const alb_layer_t* L, double z) {
  const int n = L->lut_n;
  const double x = (double)(float)z;
  if (x < (double)L->z_lut[0]) return L->alpha_lut[0];
  if (x >= (double)L->z_lut[n - 1]) return L->alpha_lut[n - 1];
  int lo = 0;
  while (lo + 1 < n - 1 && (double)L->z_lut[lo + 1] <= x) ++lo;
  const double x0 = (double)L->z_lut[lo], x1 = (double)L->z_lut[lo + 1];
  if (x == x0) return L->alpha_lut[lo];
  const double slope = (L->alpha_lut[lo + 1] - L->alpha_lut[lo]) / (x1 - x0);
  return slope * (x - x0) + L->alpha_lut[lo];
}

/* Lookups at depth d below the surface: out = {layer, velocity (float32
 * value), float32 attenuation, float32 single-scattering albedo}.
 *   velocity: velocities_by_position stores c / n into a float32 array
 *     (physics/models/simulation.py:214-242); a float32 n (linear profile)
 *     divides in float32 (water.py:198-201), an FP64 n in FP64;
 *   attenuation: a float32 array of the LUT value (water.py:330-349), or of
 *     the FP64 constant for a constant layer;
 *   albedo: (alpha - a) / alpha in FP64 from the FP64 LUT value, stored
 *     float32 (water.py:235-245, 351-370). */
void orc_layer_lookup(const alb_scene_t* sc, double d, double out[4]) {
  const int li = orc_layer_index(sc, d);
  const alb_layer_t* L = &sc->layers[li];
  const double z = d - L->top;  /* layer_local_depth: float32 tops (water.py:294-299) */
  double v, a32, w32;
  if (!L->profiled) {
    v = L->v;
    a32 = (double)L->alpha;
    w32 = (double)L->omega0;
  } else {
    if (L->n_kind == 1) {
      float n32 = (float)orc_profile(1, L->n0, L->n1, z, L->height, L->n_tab, L->tab_n);
      float c32 = (float)2.998e8;
      v = (double)(c32 / n32);
    } else {
      v = (double)(float)(2.998e8 / orc_profile(L->n_kind, L->n0, L->n1, z, L->height, L->n_tab, L->tab_n));
    }
    const double al = orc_layer_alpha64(L, z);
    a32 = (double)(float)al;
    const double a = orc_profile(L->a_kind, L->a0, L->a1, z, L->height, L->a_tab, L->tab_n);
    w32 = (double)(float)((al - a) / al);
  }
  out[0] = (double)li;
  out[1] = v;
  out[2] = a32;
  out[3] = w32;
}

/* Schlick Fresnel (src/alb_sim/physics/reflection/fresnel.py:14-29) */
static inline double orc_schlick(double cos_t, double F0) {
  return F0 + (1.0 - F0) * pow(1.0 - cos_t, 5.0);
}

/* calculate_refraction_direction (src/alb_sim/physics/reflection/
 * snell_refraction.py:7-45): returns the TIR flag */
int orc_refract(const double d[3], const double n[3], double eta, double out[3]) {
  double ci = (-d[0]) * n[0];
  ci = ci + (-d[1]) * n[1];
  ci = ci + (-d[2]) * n[2];
  if (ci < -1.0) ci = -1.0;
  if (ci > 1.0) ci = 1.0;
  double sin2t = (eta * eta) * (1.0 - ci * ci);
  int tir = sin2t > 1.0;
  double ct = sqrt(fmax(0.0, 1.0 - sin2t));
  double g = eta * ci - ct;
  for (int i = 0; i < 3; ++i) out[i] = eta * d[i] + g * n[i];
  orc_normalize3(out);
  return tir;
}

/* ---------------------------------------------------------- record sinks */

/* Path statistics of the adjoint (test of the replay trace): water length
 * whose optical depth was accumulated, phase-sampling score sum, scatters
 * weighted by omega0 and floor reflections so far. */
typedef struct orc_path {
  double lam, gs;
  int32_t nsc, nfl, dsc;  /* dsc: completed free paths ended by a scatter */
} orc_path;

/* optional per-record statistics, parallel to one sink */
typedef struct orc_stats_sink {
  double* lam;
  double* gs;
  int32_t* nsc;
  int32_t* nfl;
  int32_t* dsc;
  int32_t* rdl;  /* 1: a scatter record whose own free path ended in a counted layer */
} orc_stats_sink;

/* Per-call state of one trace: the optional statistics sinks and the state
 * of the photon being traced (its path statistics and its first water and
 * latest seafloor interaction, float32, NaN before them;
 * core/simulation.py:85-86, 393, 521).  Each call keeps its own, so calls
 * from several threads do not share anything mutable. */
typedef struct orc_ctx {
  const orc_stats_sink* stats;  /* [4] or NULL */
  int stat_layer;               /* -1: every layer counts; l: only layer l (water length, scatters) */
  orc_path path;
  float fwi[3], sfi[3];
} orc_ctx;

static int orc_stat_in(const orc_ctx* cx, int layer) { return cx->stat_layer < 0 || cx->stat_layer == layer; }

static void orc_emit_stats(const orc_ctx* cx, int type, unsigned long long i, int64_t capacity, int rdl) {
  const orc_stats_sink* st = cx->stats;
  if (!st || !st[type].lam || (int64_t)i >= capacity) return;
  st[type].lam[i] = cx->path.lam;
  st[type].gs[i] = cx->path.gs;
  st[type].nsc[i] = cx->path.nsc;
  st[type].nfl[i] = cx->path.nfl;
  if (st[type].dsc) st[type].dsc[i] = cx->path.dsc;
  if (st[type].rdl) st[type].rdl[i] = rdl;
}

/* stored with every forward record whose sink holds the interaction points */
static void orc_emit_record(const orc_ctx* cx, const alb_records_t* r, const double p[3], const double d[3],
                            float E, float t, int64_t tag) {
  unsigned long long i = (*r->count)++;
  if ((int64_t)i >= r->capacity) return;
  r->x[i] = (float)p[0]; r->y[i] = (float)p[1]; r->z[i] = (float)p[2];
  r->dx[i] = (float)d[0]; r->dy[i] = (float)d[1]; r->dz[i] = (float)d[2];
  r->energy[i] = E;
  r->time[i] = t;
  if (r->tag) r->tag[i] = tag;
  if (r->fwi_x) {
    r->fwi_x[i] = cx->fwi[0]; r->fwi_y[i] = cx->fwi[1]; r->fwi_z[i] = cx->fwi[2];
    r->sfi_x[i] = cx->sfi[0]; r->sfi_y[i] = cx->sfi[1]; r->sfi_z[i] = cx->sfi[2];
  }
}

/* accumulate_scatter_radius_heatmap for one forward seafloor hit
 * (math/accumulate_heatmap.py:64-124): surface radius from the float32 first
 * water interaction, floor radius from the FP64 hit point, each over its
 * extent times bins, truncated; photons without water entry are skipped */
static void orc_count_radius(const orc_ctx* cx, const alb_scene_t* sc, const double P[3]) {
  if (sc->heat_bins <= 0 || !sc->heat_radius || isnan(cx->fwi[0])) return;
  const double sx = (double)cx->fwi[0] - sc->heat_water_center[0];
  const double sz = (double)cx->fwi[2] - sc->heat_water_center[2];
  const double fx = P[0] - sc->heat_floor_center[0], fz = P[2] - sc->heat_floor_center[2];
  const double rs = sqrt(sx * sx + sz * sz), rf = sqrt(fx * fx + fz * fz);
  const int64_t bins = sc->heat_bins;
  const int64_t ix = (int64_t)((rs / sc->heat_water_extent) * (double)bins);
  const int64_t iy = (int64_t)((rf / sc->heat_floor_extent) * (double)bins);
  if (ix >= 0 && ix < bins && iy >= 0 && iy < bins) sc->heat_radius[ix * bins + iy] += 1u;
}

static void orc_emit_query(const alb_queries_t* q, const double p[3], const double d[3],
                           float E, float t, int64_t tag) {
  unsigned long long i = (*q->count)++;
  if ((int64_t)i >= q->capacity) return;
  q->x[i] = p[0]; q->y[i] = p[1]; q->z[i] = p[2];
  q->dx[i] = d[0]; q->dy[i] = d[1]; q->dz[i] = d[2];
  q->energy[i] = E;
  q->time[i] = t;
  if (q->tag) q->tag[i] = tag;
}

/* ------------------------------------------------------------ tilted floor
 * Extension of the reference's flat floor y = y_sf: the plane
 * floor_normal . p = floor_offset (water on the positive side); the floor
 * test, the intersection fraction and the cosine hemisphere follow the flat
 * forms with y replaced by the signed distance and +y by the normal. */
static inline double orc_floor_dot(const alb_scene_t* sc, const double v[3]) {
  return (sc->floor_normal[0] * v[0] + sc->floor_normal[1] * v[1]) + sc->floor_normal[2] * v[2];
}

static inline double orc_floor_dist(const alb_scene_t* sc, const double v[3]) {
  return orc_floor_dot(sc, v) - sc->floor_offset;
}

/* BOTTOM multiplier (core/simulation.py:155-163): (albedo / pi) max(0, n . (-d_s)) */
static inline double orc_bottom_mult(const alb_scene_t* sc, const double ds[3]) {
  if (!sc->tilted) return (sc->floor_albedo / ORC_PI) * fmax(0.0, -ds[1]);
  const double c = (sc->floor_normal[0] * -ds[0] + sc->floor_normal[1] * -ds[1]) + sc->floor_normal[2] * -ds[2];
  return (sc->floor_albedo / ORC_PI) * fmax(0.0, c);
}

/* ------------------------------------------------------------ the stepper */

enum { ST_AIR = 0, ST_ENTER = 1, ST_WATER = 2, ST_FLOOR = 3, ST_EXIT = 4, ST_DEAD = 5 };

/* -log of the uniform rounded to float32, kept in float32
 * (src/alb_sim/core/simulation.py:370-371, 428-429) */
static inline float orc_tau_target(double u) {
  float uf = (float)u;
  return (float)(-log((double)uf));
}

/* Starting direction of one photon (the position is the origin for both
 * passes, core/simulation.py:75): forward = Gaussian-cone laser sample,
 * backward = aperture disk point added to a cone-uniform direction (Q1). */
static void orc_emission(const alb_scene_t* sc, orc_stream* stp, int fwd, double d[3]) {
  orc_stream st_local = *stp;
  orc_stream* st = &st_local;
  const double two_pi = 2.0 * ORC_PI;
  if (fwd) {
    /* sample_directions_in_cone_gaussian (math/direction_cone.py:53-117) */
    double theta;
    do {
      double u = orc_uniform(st);
      theta = sc->rayleigh_sigma * sqrt(-2.0 * log(1.0 - u));
    } while (theta > sc->divergence && !st->exhausted);
    double phi = two_pi * orc_uniform(st);
    double s = sin(theta), c = cos(theta);
    double l[3] = {s * cos(phi), s * sin(phi), c};
    double w[3];
    orc_rotate_f32(sc->rot_laser, l, w);
    for (int i = 0; i < 3; ++i) d[i] = (double)(float)w[i];
  } else {
    /* sample_disk_points (math/disk_sampling.py:7-53) + cone-uniform
     * (math/direction_cone.py:7-50); disk point ADDED to the direction (Q1) */
    double xi1 = orc_uniform(st);
    double xi2 = orc_uniform(st);
    double r = sc->aperture_radius * sqrt(xi1);
    double th = two_pi * xi2;
    double ld[3] = {r * cos(th), r * sin(th), 0.0};
    double disk[3];
    orc_rotate_f32(sc->rot_laser, ld, disk);
    double z = sc->cos_fov + (1.0 - sc->cos_fov) * orc_uniform(st);
    double phi = two_pi * orc_uniform(st);
    double sn = sqrt(1.0 - z * z);
    double l[3] = {sn * cos(phi), sn * sin(phi), z};
    double w[3];
    orc_rotate_f32(sc->rot_laser, l, w);
    for (int i = 0; i < 3; ++i) d[i] = disk[i] + (double)(float)w[i];
  }
  *stp = st_local;
}

/* exported for the sampler distribution tests */
void orc_emission_dir(const alb_scene_t* sc, const alb_rng_t* rng, int32_t pass, uint64_t photon, double out[3]) {
  orc_stream st = {rng, photon, 0, 0, 0};
  orc_emission(sc, &st, pass == ALB_PASS_FORWARD, out);
}

/* cosine-weighted hemisphere about +y from two uniforms (the floor reflection;
 * same distribution as normalize(y + isotropic unit vector) of
 * physics/reflection/lambertian.py:7-27) */
void orc_lambert(double u1, double u2, double out[3]) {
  const double sr = sqrt(u1), ph = 2.0 * ORC_PI * u2;
  out[0] = sr * cos(ph);
  out[1] = sqrt(1.0 - u1);
  out[2] = sr * sin(ph);
}

/* new direction after a scatter: theta by inverse CDF of the phase table,
 * phi uniform, rotated from +z onto the incoming direction with a float32
 * matrix, stored float32 (physics/scatter/direction.py:7-51) */
/* inverse-CDF segment of u: 0 below cdf[0], i + 1 for cdf[i] <= u < cdf[i+1] */
static int orc_phase_segment(double u, const alb_scene_t* sc) {
  const double* c = sc->phase_cdf;
  const int n = sc->phase_n;
  if (u < c[0]) return 0;
  if (u >= c[n - 1]) return n - 1;
  int i = 0;
  while (i + 1 < n && c[i + 1] <= u) ++i;
  return i + 1;
}

static void orc_scatter_direction_tab(const double* theta_tab, const double* cdf, int n, const double din[3],
                                      double u1, double u2, double out[3]) {
  double theta = orc_interp(u1, cdf, theta_tab, n);
  double phi = 2.0 * ORC_PI * u2;
  double l[3] = {sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)};
  float R[9];
  orc_rotation_from_z(din, R);
  for (int i = 0; i < 3; ++i) {
    double acc = (double)R[3 * i + 0] * l[0];
    acc = acc + (double)R[3 * i + 1] * l[1];
    acc = acc + (double)R[3 * i + 2] * l[2];
    out[i] = (double)(float)acc;
  }
}

void orc_scatter_direction(const alb_scene_t* sc, const double din[3], double u1, double u2, double out[3]) {
  orc_scatter_direction_tab(sc->phase_theta, sc->phase_cdf, sc->phase_n, din, u1, u2, out);
}

/* Trace one photon through all S sub-steps.  Returns 0, or 1 if its replay
 * stream ran out (the photon is then abandoned). */
static int orc_trace_photon(orc_ctx* cx, const alb_scene_t* sc, const alb_rng_t* rng, int pass,
                            uint64_t photon, int64_t slot, const alb_records_t* out,
                            const alb_queries_t* q, unsigned long long* ctr) {
  orc_stream st = {rng, photon, slot, 0, 0};
  double p[3] = {0.0, 0.0, 0.0};
  double d[3];
  const int fwd = (pass == ALB_PASS_FORWARD);

  /* emission: physics/models/simulation.py:130-177 */
  orc_emission(sc, &st, fwd, d);
  if (ctr) ctr[ALB_CTR_PATHS]++;
  if (st.exhausted) return 1;

  float E = 1.0f, tau = 0.0f, tau_t = INFINITY;
  int64_t event_no = 0;
  for (int i = 0; i < 3; ++i) cx->fwi[i] = cx->sfi[i] = NAN;
  cx->path.lam = 0.0;
  cx->path.gs = 0.0;
  cx->path.nsc = 0;
  cx->path.nfl = 0;
  cx->path.dsc = 0;
  const double y_ws = sc->y_ws, y_sf = sc->y_sf;
  const double up[3] = {0.0, 1.0, 0.0}, down[3] = {0.0, -1.0, 0.0};
  const double alpha = (double)sc->alpha_l;
  const int layered = sc->n_layers > 0 && sc->layers != NULL;

  for (int64_t s = 0; s < sc->n_substeps; ++s) {
    double y = p[1];
    /* above the surface and not descending: no event can ever happen again */
    if (y > y_ws && d[1] >= 0.0) break;
    /* velocities_by_position (physics/models/simulation.py:214-242) */
    double v = (y > y_ws) ? sc->v_air : sc->v_water;
    double look[4];
    if (layered && y <= y_ws) {
      orc_layer_lookup(sc, (y - y_ws) * -1.0, look);
      v = look[1];
    }
    double nx[3];
    for (int i = 0; i < 3; ++i) nx[i] = p[i] + (d[i] * v) * sc->dt;
    double yn = nx[1];
    /* evaluate_state (core/simulation.py:220-244), later masks win */
    int state = ST_AIR;
    if (y > y_ws && yn <= y_ws) state = ST_ENTER;
    if (!sc->tilted) {
      if (y <= y_ws && yn > y_sf) state = ST_WATER;
      if (y > y_sf && yn <= y_sf) state = ST_FLOOR;
    } else {  /* tilted floor: the plane's signed distance, positive on the water side */
      /* The plane bounds the water column only below the surface: a step
       * in the air passes through it (where it rises above the surface), and
       * a photon in the water on its far side that does not cross back is
       * absorbed (it entered the water over dry ground). */
      const double g0 = orc_floor_dist(sc, p), g1 = orc_floor_dist(sc, nx);
      if (y <= y_ws) state = (g1 > 0.0) ? ST_WATER : (g0 > 0.0 ? ST_FLOOR : ST_DEAD);
    }
    if (y <= y_ws && yn > y_ws) state = ST_EXIT;
    if (state == ST_DEAD) break;
    const float t_evt = (float)s;
    const int64_t tag = (int64_t)(photon << 16) | (event_no & 0xFFFF);

    if (state == ST_AIR) {
      p[0] = nx[0]; p[1] = nx[1]; p[2] = nx[2];
    } else if (state == ST_WATER) {
      /* handle_water (core/simulation.py:328-373) */
      double sv[3] = {nx[0] - p[0], nx[1] - p[1], nx[2] - p[2]};
      double len = orc_norm3(sv);
      /* attenuation at the pre-step position (the lookup above) */
      tau = (float)((double)tau + (layered ? look[2] : alpha) * len);
      int scatter = tau >= tau_t;
      const int li = layered ? (int)look[0] : 0;  /* the sub-step's layer (pre-step position) */
      if (orc_stat_in(cx, li)) cx->path.lam += len;
      p[0] = nx[0]; p[1] = nx[1]; p[2] = nx[2];
      if (ctr) ctr[ALB_CTR_WATER_STEPS]++;
      if (scatter) {
        if (fwd) orc_emit_record(cx, &out[ALB_SCATTER], p, d, E, t_evt, tag);
        else orc_emit_query(&q[ALB_SCATTER], p, d, E, t_evt, tag);
        orc_emit_stats(cx, ALB_SCATTER, (fwd ? *out[ALB_SCATTER].count : *q[ALB_SCATTER].count) - 1,
                       fwd ? out[ALB_SCATTER].capacity : q[ALB_SCATTER].capacity, orc_stat_in(cx, li));
        event_no++;
        if (ctr) ctr[ALB_CTR_SCATTER]++;
        /* sample_scattering_directions_batch (physics/scatter/direction.py:7-51) */
        double u1 = orc_uniform(&st);
        double u2 = orc_uniform(&st);
        double nd[3];
        float w0 = sc->omega0;
        int li2 = 0;  /* the layer whose omega0 applies (new position) */
        if (layered) {
          /* phase function and albedo of the layer at the new position
           * (core/simulation.py:362-368, water.py:372-403) */
          double lk[4];
          orc_layer_lookup(sc, (p[1] - y_ws) * -1.0, lk);
          li2 = (int)lk[0];
          const alb_layer_t* L = &sc->layers[li2];
          orc_scatter_direction_tab(L->phase_theta, L->phase_cdf, L->phase_n, d, u1, u2, nd);
          w0 = (float)lk[3];
          /* phase-sampling score (one constant layer: its table is the scene's) */
          if (sc->phase_score) cx->path.gs += sc->phase_score[orc_phase_segment(u1, sc)];
        } else {
          orc_scatter_direction(sc, d, u1, u2, nd);
          if (sc->phase_score) cx->path.gs += sc->phase_score[orc_phase_segment(u1, sc)];
        }
        if (orc_stat_in(cx, li2)) cx->path.nsc++;
        if (orc_stat_in(cx, li)) cx->path.dsc++;
        d[0] = nd[0]; d[1] = nd[1]; d[2] = nd[2];
        E = E * w0;
        tau = tau - tau_t;
        tau_t = orc_tau_target(orc_uniform(&st));
      }
    } else if (state == ST_ENTER) {
      /* handle_enter (core/simulation.py:375-431) */
      double sv[3] = {nx[0] - p[0], nx[1] - p[1], nx[2] - p[2]};
      double f = (y_ws - p[1]) / sv[1];
      if (f < 0.0) f = 0.0;
      if (f > 1.0) f = 1.0;
      double P[3];
      for (int i = 0; i < 3; ++i) P[i] = p[i] + f * sv[i];
      double ci = -d[1];
      double Rf = orc_schlick(ci, sc->F0);
      double T = 1.0 - Rf;
      for (int i = 0; i < 3; ++i) cx->fwi[i] = (float)P[i];  /* core/simulation.py:393 */
      if (fwd) {
        orc_emit_record(cx, &out[ALB_SURFACE_REFLECTION], P, d, E, t_evt, tag);
        orc_emit_stats(cx, ALB_SURFACE_REFLECTION, *out[ALB_SURFACE_REFLECTION].count - 1,
                       out[ALB_SURFACE_REFLECTION].capacity, 0);
      } else {
        orc_emit_query(&q[ALB_SURFACE_REFLECTION], P, d, E, t_evt, tag);
        orc_emit_query(&q[ALB_SURFACE_TRANSMISSION_UP], P, d, E, t_evt, tag);
        orc_emit_stats(cx, ALB_SURFACE_REFLECTION, *q[ALB_SURFACE_REFLECTION].count - 1,
                       q[ALB_SURFACE_REFLECTION].capacity, 0);
        orc_emit_stats(cx, ALB_SURFACE_TRANSMISSION_UP, *q[ALB_SURFACE_TRANSMISSION_UP].count - 1,
                       q[ALB_SURFACE_TRANSMISSION_UP].capacity, 0);
      }
      event_no++;
      if (ctr) ctr[ALB_CTR_ENTER]++;
      double nd[3];
      orc_refract(d, up, sc->eta, nd);
      d[0] = nd[0]; d[1] = nd[1]; d[2] = nd[2];
      E = (float)((double)E * T);
      double len = orc_norm3(sv);
      double rem = (1.0 - f) * len;
      for (int i = 0; i < 3; ++i) p[i] = P[i] + d[i] * rem;
      tau_t = orc_tau_target(orc_uniform(&st));
    } else if (state == ST_EXIT) {
      /* handle_exit (core/simulation.py:433-502) */
      double sv[3] = {nx[0] - p[0], nx[1] - p[1], nx[2] - p[2]};
      double f = (y_ws - p[1]) / sv[1];
      if (f < 0.0) f = 0.0;
      if (f > 1.0) f = 1.0;
      double P[3];
      for (int i = 0; i < 3; ++i) P[i] = p[i] + f * sv[i];
      double ci = d[1];
      double Rf = orc_schlick(ci, sc->F0);
      int reflect = orc_uniform(&st) < Rf;
      if (!reflect && fwd) {
        orc_emit_record(cx, &out[ALB_SURFACE_TRANSMISSION_UP], P, d, E, t_evt, tag);
        orc_emit_stats(cx, ALB_SURFACE_TRANSMISSION_UP, *out[ALB_SURFACE_TRANSMISSION_UP].count - 1,
                       out[ALB_SURFACE_TRANSMISSION_UP].capacity, 0);
      }
      event_no++;
      if (ctr) ctr[ALB_CTR_EXIT]++;
      double rd[3];
      int tir = orc_refract(d, down, sc->eta, rd);
      reflect = reflect || tir;
      if (reflect) {
        d[1] = -d[1];
      } else {
        d[0] = rd[0]; d[1] = rd[1]; d[2] = rd[2];
      }
      double len = orc_norm3(sv);
      double rem = (1.0 - f) * len;
      for (int i = 0; i < 3; ++i) p[i] = P[i] + d[i] * rem;
      if (reflect) {
        /* attenuation at the position after the reflection (core/simulation.py:495-499) */
        double a_post = alpha;
        int lp = 0;
        if (layered) {
          double lk[4];
          orc_layer_lookup(sc, (p[1] - y_ws) * -1.0, lk);
          a_post = lk[2];
          lp = (int)lk[0];
        }
        tau = (float)((double)tau + a_post * len);
        if (orc_stat_in(cx, lp)) cx->path.lam += len;
      }
    } else { /* ST_FLOOR */
      /* handle_seafloor (core/simulation.py:504-570) */
      double sv[3] = {nx[0] - p[0], nx[1] - p[1], nx[2] - p[2]};
      double f = sc->tilted ? (sc->floor_offset - orc_floor_dot(sc, p)) / orc_floor_dot(sc, sv)
                            : (y_sf - p[1]) / sv[1];
      if (f < 0.0) f = 0.0;
      if (f > 1.0) f = 1.0;
      double P[3];
      for (int i = 0; i < 3; ++i) P[i] = p[i] + f * sv[i];
      for (int i = 0; i < 3; ++i) cx->sfi[i] = (float)P[i];  /* core/simulation.py:521 */
      if (fwd) orc_count_radius(cx, sc, P);                   /* core/simulation.py:523-537 */
      if (fwd) orc_emit_record(cx, &out[ALB_BOTTOM_REFLECTION], P, d, E, t_evt, tag);
      else orc_emit_query(&q[ALB_BOTTOM_REFLECTION], P, d, E, t_evt, tag);
      orc_emit_stats(cx, ALB_BOTTOM_REFLECTION,
                     (fwd ? *out[ALB_BOTTOM_REFLECTION].count : *q[ALB_BOTTOM_REFLECTION].count) - 1,
                     fwd ? out[ALB_BOTTOM_REFLECTION].capacity : q[ALB_BOTTOM_REFLECTION].capacity, 0);
      event_no++;
      if (ctr) ctr[ALB_CTR_FLOOR]++;
      /* cosine-weighted hemisphere about +y (physics/reflection/lambertian.py:7-27) */
      double u1 = orc_uniform(&st);
      double u2 = orc_uniform(&st);
      orc_lambert(u1, u2, d);
      if (sc->tilted) { /* the same hemisphere sample about the plane's normal */
        const double l[3] = {d[0], d[1], d[2]};
        for (int i = 0; i < 3; ++i)
          d[i] = (sc->floor_t1[i] * l[0] + sc->floor_normal[i] * l[1]) + sc->floor_t2[i] * l[2];
      }
      double len = orc_norm3(sv);
      double rem = (1.0 - f) * len;
      for (int i = 0; i < 3; ++i) p[i] = P[i] + d[i] * rem;
      /* attenuation at the position after the reflection (core/simulation.py:563-566) */
      double a_post = alpha;
      int lp = 0;
      if (layered) {
        double lk[4];
        orc_layer_lookup(sc, (p[1] - y_ws) * -1.0, lk);
        a_post = lk[2];
        lp = (int)lk[0];
      }
      tau = (float)((double)tau + a_post * len);
      if (orc_stat_in(cx, lp)) cx->path.lam += len;
      cx->path.nfl++;
      E = E * (float)sc->floor_albedo;
    }
    if (st.exhausted) return 1;
  }
  return 0;
}

/* Trace photons [photon_begin, photon_begin+n) of one pass.  Same sink
 * semantics as alb_trace (host pointers).  Returns the number of photons whose
 * replay stream ran out. */
static int64_t orc_trace_ctx(const alb_scene_t* sc, const alb_rng_t* rng, int32_t pass, int64_t photon_begin,
                             int64_t n, const alb_records_t* out, const alb_queries_t* q,
                             unsigned long long* counters, const orc_stats_sink* stats, int stat_layer) {
  orc_ctx cx;
  memset(&cx, 0, sizeof cx);
  cx.stats = stats;
  cx.stat_layer = stat_layer;
  int64_t exhausted = 0;
  for (int64_t i = 0; i < n; ++i) {
    int ex = orc_trace_photon(&cx, sc, rng, pass, (uint64_t)(photon_begin + i), i, out, q, counters);
    if (ex) {
      exhausted++;
      if (counters) counters[ALB_CTR_REPLAY_EXHAUSTED]++;
    }
  }
  return exhausted;
}

int64_t orc_trace(const alb_scene_t* sc, const alb_rng_t* rng, int32_t pass,
                  int64_t photon_begin, int64_t n, const alb_records_t* out,
                  const alb_queries_t* q, unsigned long long* counters) {
  return orc_trace_ctx(sc, rng, pass, photon_begin, n, out, q, counters, NULL, -1);
}

/* orc_trace with the adjoint path statistics of every record / query
 * written into stats[type] at the record's slot (test of the replay trace);
 * stat_layer >= 0 counts the water length and scatters of that layer only. */
int64_t orc_trace_stats(const alb_scene_t* sc, const alb_rng_t* rng, int32_t pass,
                        int64_t photon_begin, int64_t n, const alb_records_t* out,
                        const alb_queries_t* q, unsigned long long* counters,
                        const orc_stats_sink* stats, int32_t stat_layer) {
  return orc_trace_ctx(sc, rng, pass, photon_begin, n, out, q, counters, stats, stat_layer);
}

/* --------------------------------------------------------------- k-NN */

/* squared distance in FP64, terms summed x, y, z in order */
static inline double orc_d2(double qx, double qy, double qz, float x, float y, float z) {
  double dx = (double)x - qx, dy = (double)y - qy, dz = (double)z - qz;
  double s = dx * dx + dy * dy;
  return s + dz * dz;
}

/* (d2, idx) lexicographic "a before b" */
static inline int orc_before(double da, int32_t ia, double db, int32_t ib) {
  return da < db || (da == db && ia < ib);
}

/* Exact k nearest neighbours of one query by brute force over all n points,
 * ordered by (squared distance, index).  Replaces KDTree.query(pos, k)
 * (core/simulation.py:139-142).  Writes idx[k], d2[k]; returns 0, or -1 when
 * k > n (the reference raises IndexError there, core/simulation.py:143-153). */
int orc_knn_one(const float* x, const float* y, const float* z, int64_t n, double qx,
                double qy, double qz, int32_t k, int32_t* idx, double* d2) {
  if (k > n || k <= 0) return -1;
  int32_t m = 0; /* max-heap of the best k on (d2, idx) */
  for (int64_t i = 0; i < n; ++i) {
    double di = orc_d2(qx, qy, qz, x[i], y[i], z[i]);
    if (m < k) {
      int32_t c = m++;
      while (c > 0) {
        int32_t par = (c - 1) >> 1;
        if (orc_before(d2[par], idx[par], di, (int32_t)i)) {
          d2[c] = d2[par]; idx[c] = idx[par]; c = par;
        } else break;
      }
      d2[c] = di; idx[c] = (int32_t)i;
    } else if (orc_before(di, (int32_t)i, d2[0], idx[0])) {
      int32_t c = 0;
      for (;;) {
        int32_t l = 2 * c + 1, r = l + 1, big = c;
        double bd = di; int32_t bi = (int32_t)i;
        if (l < k && orc_before(bd, bi, d2[l], idx[l])) { big = l; bd = d2[l]; bi = idx[l]; }
        if (r < k && orc_before(bd, bi, d2[r], idx[r])) { big = r; }
        if (big == c) break;
        d2[c] = d2[big]; idx[c] = idx[big]; c = big;
      }
      d2[c] = di; idx[c] = (int32_t)i;
    }
  }
  /* heap sort ascending */
  for (int32_t end = k - 1; end > 0; --end) {
    double td = d2[0]; int32_t ti = idx[0];
    d2[0] = d2[end]; idx[0] = idx[end];
    d2[end] = td; idx[end] = ti;
    int32_t c = 0;
    for (;;) {
      int32_t l = 2 * c + 1, r = l + 1, big = c;
      if (l < end && orc_before(d2[big], idx[big], d2[l], idx[l])) big = l;
      if (r < end && orc_before(d2[big], idx[big], d2[r], idx[r])) big = r;
      if (big == c) break;
      double sd = d2[c]; int32_t si = idx[c];
      d2[c] = d2[big]; idx[c] = idx[big];
      d2[big] = sd; idx[big] = si;
      c = big;
    }
  }
  return 0;
}

/* ------------------------------------------------------------ multipliers */

static inline double orc_ggx_d(double ndh, double alpha) {
  double a2 = alpha * alpha;
  double den = (a2 - 1.0) * ndh * ndh + 1.0;
  den = ORC_PI * den * den;
  return a2 / (den + ORC_FLT_EPS);
}

static inline double orc_g1(double x, double alpha) {
  double a2 = alpha * alpha;
  double den = x + sqrt(a2 + (1.0 - a2) * (x * x));
  return 2.0 * x / den;
}

static inline double orc_clip(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

/* microfacet_reflected_energy(incident=-d_p, outgoing=-d_s, n=+y), float32 out
 * (physics/reflection/microfacet_brdf.py:96-147; physics/models/sea_surface.py:35-62) */
float orc_brdf_reflect(const float dp[3], const double ds[3], double alpha, double F0) {
  float wi[3] = {-dp[0], -dp[1], -dp[2]};
  orc_normalize3f(wi);
  double wo[3] = {-ds[0], -ds[1], -ds[2]};
  {
    double s = wo[0] * wo[0] + wo[1] * wo[1];
    s = s + wo[2] * wo[2];
    double nn = sqrt(s);
    if (nn != 0.0) { wo[0] /= nn; wo[1] /= nn; wo[2] /= nn; }
  }
  double h[3] = {(double)wi[0] + wo[0], (double)wi[1] + wo[1], (double)wi[2] + wo[2]};
  orc_normalize3(h);
  double ndwi = orc_clip((double)wi[1], 0.0, 1.0);
  double ndwo = orc_clip(wo[1], 0.0, 1.0);
  double widh = (double)wi[0] * h[0];
  widh = widh + (double)wi[1] * h[1];
  widh = widh + (double)wi[2] * h[2];
  widh = orc_clip(widh, 0.0, 1.0);
  double ndh = orc_clip(h[1], 0.0, 1.0);
  if (!(ndwi > 0.0 && ndwo > 0.0)) return 0.0f;
  double D = orc_ggx_d(ndh, alpha);
  double G = orc_g1(ndwi, alpha) * orc_g1(ndwo, alpha);
  double F = orc_schlick(widh, F0);
  double den = 4.0 * ndwi * ndwo + ORC_FLT_EPS;
  return (float)((D * G * F) / den);
}

/* microfacet_transmitted_energy(-d_p, -d_s, +y, r, n, 1, F0) + albedo/pi
 * (physics/reflection/microfacet_brdf.py:150-226; physics/models/sea_surface.py:64-100) */
double orc_btdf_transmit(const float dp[3], const double ds[3], double alpha, double n_w,
                         double F0, double surface_albedo) {
  float wi[3] = {dp[0], dp[1], dp[2]};
  orc_normalize3f(wi);
  double wo[3] = {-ds[0], -ds[1], -ds[2]};
  {
    double s = wo[0] * wo[0] + wo[1] * wo[1];
    s = s + wo[2] * wo[2];
    double nn = sqrt(s);
    if (nn != 0.0) { wo[0] /= nn; wo[1] /= nn; wo[2] /= nn; }
  }
  double eta_i = n_w, eta_o = 1.0, eta = eta_i / eta_o;
  double h[3] = {(double)wi[0] + eta * wo[0], (double)wi[1] + eta * wo[1], (double)wi[2] + eta * wo[2]};
  orc_normalize3(h);
  double ndwi = orc_clip((double)wi[1], -1.0, 1.0);
  double ndwo = orc_clip(wo[1], -1.0, 1.0);
  double ndh = orc_clip(h[1], 0.0, 1.0);
  double widh = (double)wi[0] * h[0];
  widh = widh + (double)wi[1] * h[1];
  widh = widh + (double)wi[2] * h[2];
  widh = orc_clip(widh, 0.0, 1.0);
  double wodh = h[0] * wo[0];
  wodh = wodh + h[1] * wo[1];
  wodh = wodh + h[2] * wo[2];
  wodh = orc_clip(wodh, 0.0, 1.0);
  double btdf = 0.0;
  if (ndwi * ndwo < 0.0) {
    double D = orc_ggx_d(ndh, alpha);
    double G = orc_g1(fabs(ndwi), alpha) * orc_g1(fabs(ndwo), alpha);
    double F = orc_schlick(widh, F0);
    double den = eta_i * widh + eta_o * wodh;
    double factor = (eta_o * eta_o * wodh) / (den * den + ORC_FLT_EPS);
    btdf = (1.0 - F) * D * G * factor / (fabs(ndwi) * fabs(ndwo) + ORC_FLT_EPS);
  }
  return btdf + surface_albedo / ORC_PI;
}

/* calculate_energy_batch: phase-function mass of the 1/18000-pi angle bin
 * (physics/scatter/energy.py:8-52).  delta is float32 arithmetic because
 * np.pi - np.float32 eps stays float32 under NumPy 2 promotion. */
double orc_scatter_mult(const float dp[3], const double ds[3], const double* theta,
                        const double* cdf, int n) {
  float pd[3] = {dp[0], dp[1], dp[2]};
  orc_normalize3f(pd);
  double sd[3] = {-ds[0], -ds[1], -ds[2]};
  {
    double s = sd[0] * sd[0] + sd[1] * sd[1];
    s = s + sd[2] * sd[2];
    double nn = sqrt(s);
    if (nn != 0.0) { sd[0] /= nn; sd[1] /= nn; sd[2] /= nn; }
  }
  double c = (double)pd[0] * sd[0];
  c = c + (double)pd[1] * sd[1];
  c = c + (double)pd[2] * sd[2];
  c = orc_clip(c, -1.0, 1.0);
  double ang = acos(c);
  float delta_f = ((float)ORC_PI - (float)ORC_FLT_EPS) / (float)n;
  double delta = (double)delta_f;
  double q = (ang - ORC_FLT_EPS) / delta;
  int64_t i = (int64_t)q; /* astype(int) truncates toward zero */
  if (i < 0) i = 0;
  if (i > n - 1) i = n - 1;
  double lo = (double)i * delta;
  double hi = (double)(i + 1) * delta;
  return orc_interp(hi, theta, cdf, n) - orc_interp(lo, theta, cdf, n);
}

double orc_ggx_d_x(double ndh, double alpha) { return orc_ggx_d(ndh, alpha); }
double orc_g1_x(double x, double alpha) { return orc_g1(x, alpha); }
double orc_schlick_x(double c, double F0) { return orc_schlick(c, F0); }

/* ---------------------------------------------------------------- gather */

/* sample_photons for queries [0, nq) of one type against one map
 * (core/simulation.py:132-195), brute-force k-NN, FP64 accumulation in query
 * order then neighbour order.  nbr_out (optional) [nq*k]; radius_out
 * (optional) [nq].  Returns 0, or -1 (k > n or empty map). */
/* scatter_energy's phase table: that of the layer at the query vertex depth
 * (physics/models/water.py:418-447), else the scene's */
typedef struct {
  const double* theta;
  const double* cdf;
  int n;
} orc_phase_tab;

static orc_phase_tab orc_gather_phase(const alb_scene_t* sc, double qy) {
  orc_phase_tab t = {sc->phase_theta, sc->phase_cdf, sc->phase_n};
  if (sc->n_layers > 0 && sc->layers) {
    const alb_layer_t* L = &sc->layers[orc_layer_index(sc, (qy + (-sc->y_ws)) * -1.0)];
    t.theta = L->phase_theta;
    t.cdf = L->phase_cdf;
    t.n = L->phase_n;
  }
  return t;
}

/* interaction points of the map records being gathered (heat maps); all
 * NULL: no heat maps */
typedef struct orc_map_pts {
  const float* fwi_x;
  const float* fwi_z;
  const float* sfi_x;
  const float* sfi_z;
} orc_map_pts;

/* accumulate_to_heatmap (math/accumulate_heatmap.py:6-61) for one point:
 * bins of (x - centre + extent) / (2 extent / bins) along x and z, truncated
 * toward zero (astype(int64)); NaN points and bins outside [0, bins) skipped */
static void orc_heat_add(double* h, float px, float pz, const double c[3], double extent, int bins, double e) {
  if (isnan(px)) return;
  const double bs = 2.0 * extent / (double)bins;
  const double xc = (double)px - c[0], zc = (double)pz - c[2];
  const double vx = (xc + extent) / bs, vz = (zc + extent) / bs;
  if (!(vx > -1.0 && vx < (double)bins && vz > -1.0 && vz < (double)bins)) return;
  const int64_t ix = (int64_t)vx, iz = (int64_t)vz;
  h[ix * bins + iz] += e;
}

/* one (query, record) contribution of sample_photons (core/simulation.py:155-195) */
static void orc_contribute(const alb_scene_t* sc, int32_t type, orc_phase_tab tab, double m_bottom,
                           const double ds[3], float q_energy, float q_time, int32_t i, const float* dx,
                           const float* dy, const float* dz, const float* energy, const float* time,
                           double knorm, double* waveform, int64_t n_bins, const orc_map_pts* pts) {
  float dp[3] = {dx[i], dy[i], dz[i]};
  double m;
  if (type == ALB_BOTTOM_REFLECTION) m = m_bottom;
  else if (type == ALB_SCATTER) m = orc_scatter_mult(dp, ds, tab.theta, tab.cdf, tab.n);
  else if (type == ALB_SURFACE_REFLECTION) m = (double)orc_brdf_reflect(dp, ds, sc->alpha_ggx, sc->F0);
  else m = orc_btdf_transmit(dp, ds, sc->alpha_ggx, sc->n_water, sc->F0, sc->surface_albedo);
  float ee = q_energy * energy[i];
  double e = (double)ee * knorm * m;
  float tt = q_time + time[i];
  int64_t bin = (int64_t)(tt / (float)sc->mult);
  if (bin >= 0 && bin < n_bins) waveform[bin] += e;
  if (sc->heat_bins > 0 && sc->heat_water && pts && pts->fwi_x) {
    orc_heat_add(sc->heat_water, pts->fwi_x[i], pts->fwi_z[i], sc->heat_water_center,
                 sc->heat_water_extent, sc->heat_bins, e);
    orc_heat_add(sc->heat_floor, pts->sfi_x[i], pts->sfi_z[i], sc->heat_floor_center,
                 sc->heat_floor_extent, sc->heat_bins, e);
  }
}

/* sample_photons as above; nbr_in (optional, [nq*k] indices into the map):
 * the neighbour sets are given (from any exact k-NN) instead of found by
 * brute force, and the kernel radius is the largest distance within each
 * set.  fwi_x / fwi_z / sfi_x / sfi_z (optional): the records' interaction
 * points for the heat maps. */
int orc_gather_ex(const alb_scene_t* sc, int32_t type, const float* x, const float* y,
                  const float* z, const float* dx, const float* dy, const float* dz,
                  const float* energy, const float* time, int64_t n, const alb_queries_t* q,
                  int64_t nq, double* waveform, int64_t n_bins, int32_t* nbr_out,
                  double* radius_out, const int32_t* nbr_in, const float* fwi_x, const float* fwi_z,
                  const float* sfi_x, const float* sfi_z) {
  const int32_t k = sc->k;
  if (nq == 0) return 0;
  if (n <= 0 || k > n) return -1;
  const orc_map_pts pts = {fwi_x, fwi_z, sfi_x, sfi_z};
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)k);
  double* d2 = (double*)malloc(sizeof(double) * (size_t)k);
  for (int64_t j = 0; j < nq; ++j) {
    double qx = q->x[j], qy = q->y[j], qz = q->z[j];
    double ds[3] = {q->dx[j], q->dy[j], q->dz[j]};
    double r;
    if (nbr_in) {
      double mx = 0.0;
      for (int32_t t = 0; t < k; ++t) {
        const int32_t i = nbr_in[j * k + t];
        if (i < 0 || i >= n) {
          free(idx);
          free(d2);
          return -2;
        }
        idx[t] = i;
        const double dd = orc_d2(qx, qy, qz, x[i], y[i], z[i]);
        if (dd > mx) mx = dd;
      }
      r = sqrt(mx);
    } else {
      orc_knn_one(x, y, z, n, qx, qy, qz, k, idx, d2);
      r = sqrt(d2[k - 1]);
    }
    double ksize;
    if (type == ALB_SCATTER) ksize = (4.0 / 3.0) * ORC_PI * pow(r, 3.0);
    else ksize = ORC_PI * (r * r);
    double knorm = 1.0 / ksize;
    double m_bottom = orc_bottom_mult(sc, ds);
    if (nbr_out) memcpy(nbr_out + j * k, idx, sizeof(int32_t) * (size_t)k);
    if (radius_out) radius_out[j] = r;
    const orc_phase_tab tab = orc_gather_phase(sc, qy);
    for (int32_t t = 0; t < k; ++t)
      orc_contribute(sc, type, tab, m_bottom, ds, q->energy[j], q->time[j], idx[t], dx, dy, dz, energy, time,
                     knorm, waveform, n_bins, &pts);
  }
  free(idx);
  free(d2);
  return 0;
}

int orc_gather(const alb_scene_t* sc, int32_t type, const float* x, const float* y,
               const float* z, const float* dx, const float* dy, const float* dz,
               const float* energy, const float* time, int64_t n, const alb_queries_t* q,
               int64_t nq, double* waveform, int64_t n_bins, int32_t* nbr_out,
               double* radius_out) {
  return orc_gather_ex(sc, type, x, y, z, dx, dy, dz, energy, time, n, q, nq, waveform, n_bins, nbr_out,
                       radius_out, NULL, NULL, NULL, NULL, NULL);
}

/* Fixed-radius gather: every record within distance r of a query (squared
 * distance as orc_d2, d2 <= r*r), brute force over the map in index order;
 * kernel 1/(pi r^2), or 1/((4/3) pi r^3) for the scatter map, from the given r;
 * multiplier, energy product and time bin as orc_gather
 * (core/simulation.py:155-195).  count_out / isum_out (optional): per query
 * the number of records taken and the sum of their indices. */
int orc_radius_gather(const alb_scene_t* sc, int32_t type, const float* x, const float* y,
                      const float* z, const float* dx, const float* dy, const float* dz,
                      const float* energy, const float* time, int64_t n, const alb_queries_t* q,
                      int64_t nq, double r, double* waveform, int64_t n_bins, int64_t* count_out,
                      int64_t* isum_out) {
  const double r2 = r * r;
  const double ksize = (type == ALB_SCATTER) ? (4.0 / 3.0) * ORC_PI * pow(r, 3.0) : ORC_PI * (r * r);
  const double knorm = 1.0 / ksize;
  for (int64_t j = 0; j < nq; ++j) {
    const double qx = q->x[j], qy = q->y[j], qz = q->z[j];
    const double ds[3] = {q->dx[j], q->dy[j], q->dz[j]};
    const double m_bottom = orc_bottom_mult(sc, ds);
    const orc_phase_tab tab = orc_gather_phase(sc, qy);
    int64_t cnt = 0, isum = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (orc_d2(qx, qy, qz, x[i], y[i], z[i]) > r2) continue;
      cnt++;
      isum += i;
      orc_contribute(sc, type, tab, m_bottom, ds, q->energy[j], q->time[j], (int32_t)i, dx, dy, dz, energy,
                     time, knorm, waveform, n_bins, NULL);
    }
    if (count_out) count_out[j] = cnt;
    if (isum_out) isum_out[j] = isum;
  }
  return 0;
}

/* layout of alb_layer_t as this compiler sees the header (checked against
 * the ctypes mirror) */
void orc_layer_layout(int64_t out[4]) {
  out[0] = (int64_t)sizeof(alb_layer_t);
  out[1] = (int64_t)offsetof(alb_layer_t, a_tab);
  out[2] = (int64_t)offsetof(alb_layer_t, tab_n);
  out[3] = (int64_t)offsetof(alb_layer_t, phase_n);
}
