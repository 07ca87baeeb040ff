// alb_trace.hip — photon tracing, one photon per lane.
//
// Replaces Simulation.simulate_batch / simulate_photon_step and the four
// interaction handlers (src/alb_sim/core/simulation.py:71-107, 220-570).  The
// reference advances every photon by one fixed time step per iteration; here a
// lane jumps from event to event:
//  * air legs are straight and draw no random numbers, so the sub-step at
//    which the photon crosses the surface is solved in closed form (FP64);
//  * in water the sub-step structure matters because the optical depth is a
//    float32 running sum (core/simulation.py:339-344): the lane reproduces that
//    float32 accumulation exactly, jumping a whole float32 binade at a time
//    (advance_optical_depth), until the target is reached or the closed-form
//    boundary sub-step (floor / surface) comes;
//  * events happen at the sub-step index the reference would give them and
//    use the reference's handler formulas (intersection fraction, Schlick,
//    Snell with eta = 1/n both ways, FF inverse-CDF sampling with a float32
//    rotation, cosine-hemisphere floor reflection).
// Records (forward) and gather queries (backward) leave through per-type
// sinks; each wave reserves its slots with one atomic per type.
#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

struct Sinks {
  alb_records_t rec[ALB_N_TYPES];
  alb_queries_t q[ALB_N_TYPES];
};

// first j >= 0 with (y0 + j*dy) + dy <= bound (dy < 0); >= cap means "none below cap"
__device__ __forceinline__ int64_t first_step_below(double y0, double dy, double bound, int64_t cap) {
  const double x = (bound - y0) / dy;
  if (!(x < (double)cap + 2.0)) return cap;
  int64_t j = x > 1.0 ? (int64_t)ceil(x) - 1 : 0;
  while (j > 0 && ((y0 + (double)(j - 1) * dy) + dy <= bound)) --j;
  while (j < cap && !((y0 + (double)j * dy) + dy <= bound)) ++j;
  return j;
}

// first j >= 0 with (y0 + j*dy) + dy > bound (dy > 0)
__device__ __forceinline__ int64_t first_step_above(double y0, double dy, double bound, int64_t cap) {
  const double x = (bound - y0) / dy;
  if (!(x < (double)cap + 2.0)) return cap;
  int64_t j = x > 1.0 ? (int64_t)ceil(x) - 1 : 0;
  while (j > 0 && ((y0 + (double)(j - 1) * dy) + dy > bound)) --j;
  while (j < cap && !((y0 + (double)j * dy) + dy > bound)) ++j;
  return j;
}

// Snell refraction about a horizontal interface with normal (0, ny, 0), ny = +-1
// (physics/reflection/snell_refraction.py:7-45 with refractive_index_1 = 1,
//  refractive_index_2 = n, i.e. eta = 1/n for entry and exit alike)
__device__ __forceinline__ void refract_y(double& dx, double& dy, double& dz, double ny, double eta) {
  double ci = -dy * ny;
  ci = clipd(ci, -1.0, 1.0);
  const double sin2t = (eta * eta) * (1.0 - ci * ci);
  const double ct = sqrt(fmax(0.0, 1.0 - sin2t));
  const double g = eta * ci - ct;
  double tx = eta * dx, ty = eta * dy + g * ny, tz = eta * dz;
  double n = norm3(tx, ty, tz);
  if (n == 0.0) n = 1.0;
  dx = tx / n;
  dy = ty / n;
  dz = tz / n;
}

__device__ __forceinline__ bool isclose1(double a, double b) {
  return fabs(a - b) <= (1e-08 + 1e-05 * fabs(b));
}

// float32 rotation taking +z to v (math/rotation_matrix.py:39-102)
__device__ __forceinline__ void rotation_from_z(double vx0, double vy0, double vz0, float R[9]) {
  double n = norm3(vx0, vy0, vz0);
  if (n == 0.0) n = 1.0;
  const double v0 = vx0 / n, v1 = vy0 / n, v2 = vz0 / n;
  const double c = v2;
  if (isclose1(c, 1.0)) {
    R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 0; R[4] = 1; R[5] = 0; R[6] = 0; R[7] = 0; R[8] = 1;
    return;
  }
  if (isclose1(c, -1.0)) {
    R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 0; R[4] = -1; R[5] = 0; R[6] = 0; R[7] = 0; R[8] = -1;
    return;
  }
  const double ax = -v1, ay = v0, az = 0.0;
  const double s2 = ax * ax + ay * ay;
  const double K[9] = {0.0, -az, ay, az, 0.0, -ax, -ay, ax, 0.0};
  const double f = (1.0 - c) / s2;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double kk = K[3 * i + 0] * K[j];
      kk = kk + K[3 * i + 1] * K[3 + j];
      kk = kk + K[3 * i + 2] * K[6 + j];
      const double id = (i == j) ? 1.0 : 0.0;
      R[3 * i + j] = (float)((id + K[3 * i + j]) + kk * f);
    }
}

__device__ __forceinline__ void rotate(const float R[9], double l0, double l1, double l2, double& o0,
                                       double& o1, double& o2) {
  double s;
  s = l0 * (double)R[0]; s = s + l1 * (double)R[1]; s = s + l2 * (double)R[2]; o0 = s;
  s = l0 * (double)R[3]; s = s + l1 * (double)R[4]; s = s + l2 * (double)R[5]; o1 = s;
  s = l0 * (double)R[6]; s = s + l1 * (double)R[7]; s = s + l2 * (double)R[8]; o2 = s;
}


// Exact emulation of up to jmax water sub-steps of the float32 optical-depth
// update tau = float32(float64(tau) + inc) (core/simulation.py:339-344, inc =
// alpha * step length) with a stop at the first sub-step where tau >= tau_t.
// Within one float32 binade [top/2, top) with ulp u, float64(tau) + inc rounds
// to tau + I (I = the same float64 value for every tau of the binade, because
// tau is a multiple of u and u = 2^29 float64 ulps), and rounding that to
// float32 adds round(I/u) * u, again the same for every step unless I/u ends in
// exactly .5 (a tie, whose rounding depends on the parity of tau).  So a
// binade is crossed in one jump; the sub-step that leaves the binade and any
// tie are taken one at a time with the literal formula.  Returns the number
// of sub-steps taken; *hit tells whether the last one reached tau_t.
__device__ __forceinline__ int64_t advance_optical_depth(float& tau, float tau_t, double inc, int64_t jmax,
                                                        bool* hit, uint32_t& solves) {
  int64_t j = 0;
  *hit = false;
  const double tt = (double)tau_t;
  while (j < jmax) {
    ++solves;
    const double t0 = (double)tau;
    bool single = !(tau >= 1.17549435e-38f);  // zero (or subnormal): literal step
    double top = 0.0, I = 0.0, u = 0.0, nu = 0.0;
    if (!single) {
      int e;
      frexpf(tau, &e);  // tau in [2^(e-1), 2^e)
      top = ldexp(1.0, e);
      u = ldexp(1.0, e - 24);
      const double sum = t0 + inc;
      if (!(sum < top)) {
        single = true;  // this sub-step leaves the binade
      } else {
        I = sum - t0;  // exact
        const double q = ldexp(I, 24 - e);  // I / u, exact (u is a power of two)
        const double fl = floor(q);
        const double fr = q - fl;
        if (fr == 0.5) single = true;
        else nu = fr > 0.5 ? fl + 1.0 : fl;
      }
    }
    if (single) {
      tau = (float)(t0 + inc);
      ++j;
      if (tau >= tau_t) {
        *hit = true;
        return j;
      }
      continue;
    }
    const int64_t left = jmax - j;
    if (nu == 0.0) {  // the add no longer changes tau
      if (t0 >= tt) {
        ++j;
        *hit = true;
        return j;
      }
      return jmax;
    }
    const double d = nu * u;
    // the counts below start from estimates through the hardware FP64
    // reciprocal and are then corrected one sub-step at a time on the exact
    // sums (the estimates land within a count or two; capping the target count
    // at the binade's, which would bound the corrections whatever the
    // estimate, cost 17 more spilled VGPRs and an 18 % slower trace)
    const double rd = __builtin_amdgcn_rcp(d);
    // sub-steps whose float64 sum stays inside the binade: t0 + c*d + I < top
    int64_t c = (int64_t)ceil((top - I - t0) * rd);
    if (c < 1) c = 1;
    while (c > 1 && !(t0 + (double)(c - 1) * d + I < top)) --c;
    while (t0 + (double)c * d + I < top) ++c;
    // first sub-step reaching the target: t0 + nt*d >= tau_t, nt >= 1
    int64_t nt = 1;
    if (t0 + d < tt) {
      nt = (int64_t)ceil((tt - t0) * rd);
      if (nt < 1) nt = 1;
      while (nt > 1 && t0 + (double)(nt - 1) * d >= tt) --nt;
      while (t0 + (double)nt * d < tt) ++nt;
    }
    if (nt <= c && nt <= left) {
      tau = (float)(t0 + (double)nt * d);
      j += nt;
      *hit = true;
      return j;
    }
    const int64_t take = c < left ? c : left;
    tau = (float)(t0 + (double)take * d);
    j += take;
  }
  return j;
}

__device__ __forceinline__ float tau_target(double u) {
  const float uf = (float)u;
  return (float)(-log((double)uf));
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ------------------------------------------------------------ per-photon engine

struct Photon {
  double px, py, pz, dx, dy, dz;
  float E, tau, tau_t;
  int64_t s;          // current sub-step
  uint32_t event_no;  // events so far (record tag)
  bool alive;
  // path statistics of the replay (adjoint) trace only
  // (restricted to one layer of a layered column when the replay asks for
  // it: lam, nsc and dsc then count only that layer's part)
  double lam;         // water path length whose optical depth was accumulated
  double gs;          // sum of phase-sampling scores d log p / dg
  int nsc, nfl;       // scatters (omega0 applied) and floor reflections so far
  int dsc;            // completed free paths ended by a scatter (their density)
  // heat maps only: first water interaction and latest seafloor interaction
  // (float32, NaN before the interaction; core/simulation.py:85-86, 393, 521)
  float fx, fy, fz, bx, by, bz;
};

// record / query produced by one event
struct Emit {
  int et;  // photon type, -1 for none
  double x, y, z, dx, dy, dz;
  float E, t;
  uint32_t eno;
  // replay trace: the path statistics of the record's prefix, and rdl = 1 for
  // a scatter record whose own free-path density counts (its sub-step's layer)
  double lam, gs;
  int nsc, nfl, dsc, rdl;
  // heat maps: the photon's interaction points at the event
  float fx, fy, fz, bx, by, bz;
};

template <bool STATS, bool HEAT = false>
__device__ __forceinline__ void snapshot(Emit& e, const Photon& ph) {
  if constexpr (STATS) {
    e.lam = ph.lam;
    e.gs = ph.gs;
    e.nsc = ph.nsc;
    e.nfl = ph.nfl;
    e.dsc = ph.dsc;
  }
  if constexpr (HEAT) {
    e.fx = ph.fx; e.fy = ph.fy; e.fz = ph.fz;
    e.bx = ph.bx; e.by = ph.by; e.bz = ph.bz;
  }
}

// forward scatter-radius histogram at a seafloor hit: surface radius from the
// float32 first water interaction, floor radius from the FP64 hit point
// (math/accumulate_heatmap.py:64-124, hooked at core/simulation.py:523-537);
// photons without a water entry (NaN) and radii outside the extents are not
// counted
__device__ __forceinline__ void count_radius(const alb_scene_t& sc, const Photon& ph, double Px, double Pz) {
  if (sc.heat_bins <= 0 || isnan(ph.fx)) return;
  const double sx = (double)ph.fx - sc.heat_water_center[0], sz = (double)ph.fz - sc.heat_water_center[2];
  const double fxd = Px - sc.heat_floor_center[0], fzd = Pz - sc.heat_floor_center[2];
  const double rs = sqrt(sx * sx + sz * sz);
  const double rf = sqrt(fxd * fxd + fzd * fzd);
  const int64_t bins = sc.heat_bins;
  const int64_t ix = (int64_t)((rs / sc.heat_water_extent) * (double)bins);
  const int64_t iy = (int64_t)((rf / sc.heat_floor_extent) * (double)bins);
  if (ix >= 0 && ix < bins && iy >= 0 && iy < bins) atomicAdd(&sc.heat_radius[ix * bins + iy], 1u);
}

// per-lane event counts of one launch: 32 bits hold a lane's share of every
// count but the water sub-steps (up to the time budget per photon)
struct Tally {
  uint32_t paths, air, sc, en, ex, fl, exh, solve;
  unsigned long long ws;
};

// Start a photon at the origin (core/simulation.py:75) with its emission
// direction (physics/models/simulation.py:130-177).
__device__ __forceinline__ void emit_photon(const alb_scene_t& sc, bool fwd, const alb_rng_t& rng, Stream& st,
                                            Photon& ph, Tally& c) {
  const double two_pi = 2.0 * ALB_PI;
  ph.px = ph.py = ph.pz = 0.0;
  if (fwd) {
    // sample_directions_in_cone_gaussian (math/direction_cone.py:53-117)
    double theta;
    do {
      const double u = uniform(rng, st);
      theta = sc.rayleigh_sigma * sqrt(-2.0 * log(1.0 - u));
    } while (theta > sc.divergence && !st.exhausted);
    const double phi = two_pi * uniform(rng, st);
    double s, cth, sp, cp;
    sincos(theta, &s, &cth);
    sincos(phi, &sp, &cp);
    double w0, w1, w2;
    rotate(sc.rot_laser, s * cp, s * sp, cth, w0, w1, w2);
    ph.dx = (double)(float)w0;
    ph.dy = (double)(float)w1;
    ph.dz = (double)(float)w2;
  } else {
    // aperture disk point added to a cone-uniform direction
    // (physics/models/simulation.py:157-177; math/disk_sampling.py:7-53)
    const double xi1 = uniform(rng, st);
    const double xi2 = uniform(rng, st);
    const double r = sc.aperture_radius * sqrt(xi1);
    const double th = two_pi * xi2;
    double st_, ct_;
    sincos(th, &st_, &ct_);
    double k0, k1, k2;
    rotate(sc.rot_laser, r * ct_, r * st_, 0.0, k0, k1, k2);
    const double z = sc.cos_fov + (1.0 - sc.cos_fov) * uniform(rng, st);
    const double phi = two_pi * uniform(rng, st);
    const double sn = sqrt(1.0 - z * z);
    double sp, cp;
    sincos(phi, &sp, &cp);
    double w0, w1, w2;
    rotate(sc.rot_laser, sn * cp, sn * sp, z, w0, w1, w2);
    ph.dx = k0 + (double)(float)w0;
    ph.dy = k1 + (double)(float)w1;
    ph.dz = k2 + (double)(float)w2;
  }
  ph.E = 1.0f;
  ph.tau = 0.0f;
  ph.tau_t = __builtin_inff();
  ph.s = 0;
  ph.event_no = 0;
  ph.alive = true;
  ph.fx = ph.fy = ph.fz = ph.bx = ph.by = ph.bz = __builtin_nanf("");
  c.paths++;
  if (st.exhausted) {
    ph.alive = false;
    c.exh++;
  }
}

// tilted floor: n . p (terms summed x, y, z) and the signed distance n . p - c
__device__ __forceinline__ double floor_dot(const alb_scene_t& sc, double x, double y, double z) {
  return (sc.floor_normal[0] * x + sc.floor_normal[1] * y) + sc.floor_normal[2] * z;
}

__device__ __forceinline__ double floor_dist(const alb_scene_t& sc, double x, double y, double z) {
  return floor_dot(sc, x, y, z) - sc.floor_offset;
}

// first j >= 0 whose sub-step ends on the floor side of a tilted plane,
// floor_dist((p + j*D) + D) <= 0, with g0 = floor_dist(p) and gD = n . D < 0
// (cap if none).  The signed distance is affine in j, so ceil(-g0/gD) - 1 is
// the index up to rounding; one correction on the positions the leg itself
// computes settles it.  (A correction loop, as the flat floor's search has,
// costs the batched kernels 35 VGPRs and half their occupancy; only a leg
// grazing the plane for ~1e5 sub-steps could need a second step.)
__device__ __forceinline__ int64_t first_step_through_plane(const alb_scene_t& sc, double px, double py, double pz,
                                                            double Dx, double Dy, double Dz, double g0, double gD,
                                                            int64_t cap) {
  const double x = -g0 / gD;
  if (!(x < (double)cap + 2.0)) return cap;
  auto ends_below = [&](int64_t j) {
    const double a = (double)j;
    return floor_dist(sc, (px + a * Dx) + Dx, (py + a * Dy) + Dy, (pz + a * Dz) + Dz) <= 0.0;
  };
  int64_t j = x > 1.0 ? (int64_t)ceil(x) - 1 : 0;
  if (j > 0 && ends_below(j - 1)) --j;
  else if (!ends_below(j)) ++j;
  return j < cap ? j : cap;
}

// Layered or depth-profiled water column, or a tilted floor: sub-steps are
// taken one at a time,
// as Simulation.simulate_photon_step does (core/simulation.py:246-322): the
// speed and the attenuation come from the layer of the pre-step position
// (physics/models/simulation.py:214-265), the phase table and the
// single-scattering albedo of a scatter from the layer of the new position
// (core/simulation.py:362-368), and the optical depth added after a floor or
// surface reflection from the layer of the position after the event
// (core/simulation.py:495-499, 563-566).  Positions are updated by the same
// repeated additions, so they match the reference's bit for bit.
// does layer l count in the replay's statistics (stat_layer < 0: every layer)
__device__ __forceinline__ bool stat_in(int stat_layer, int l) { return stat_layer < 0 || stat_layer == l; }

template <bool STATS, bool HEAT = false>
__device__ __forceinline__ void water_stepped(const alb_scene_t& sc, bool fwd, const alb_rng_t& rng, Stream& st,
                                              Photon& ph, Emit& e, Tally& c, int stat_layer) {
  const double two_pi = 2.0 * ALB_PI;
  const int64_t S = sc.n_substeps;
  const double y_ws = sc.y_ws, y_sf = sc.y_sf;
  const alb_layer_t* __restrict__ Ls = sc.layers;
  const int nl = sc.n_layers;
  for (;;) {
    if (ph.s >= S) {
      ph.alive = false;  // time budget used up in water
      return;
    }
    const double y = ph.py;
    const double dep = (y - y_ws) * -1.0;
    const int li = layer_of(Ls, nl, dep);
    const alb_layer_t& L = Ls[li];
    const double zl = dep - L.top;
    const double v = layer_velocity(L, zl);
    const double nx = ph.px + (ph.dx * v) * sc.dt, ny = ph.py + (ph.dy * v) * sc.dt, nz = ph.pz + (ph.dz * v) * sc.dt;
    // evaluate_state (core/simulation.py:220-244) for a position at or below
    // the surface; a tilted floor compares the plane's signed distance
    // (positive on the water side) where the flat floor compares y with y_sf
    int state = 0;
    if (!sc.tilted) {
      if (ny > y_sf) state = 1;                // in water
      if (y > y_sf && ny <= y_sf) state = 2;   // hitting the floor
    } else {
      const double g0 = floor_dist(sc, ph.px, y, ph.pz), g1 = floor_dist(sc, nx, ny, nz);
      if (g1 > 0.0) state = 1;
      if (g0 > 0.0 && g1 <= 0.0) state = 2;
    }
    if (ny > y_ws) state = 3;                // leaving the water
    if (state == 0) {
      ph.alive = false;  // below the floor and not rising: no event can follow
      return;
    }
    const double svx = nx - ph.px, svy = ny - ph.py, svz = nz - ph.pz;
    const double slen = norm3(svx, svy, svz);
    if (state == 1) {
      // handle_water (core/simulation.py:328-373)
      const float a32 = layer_alpha32(L, zl);
      ph.tau = (float)((double)ph.tau + (double)a32 * slen);
      c.ws++;
      if constexpr (STATS) {
        if (stat_in(stat_layer, li)) ph.lam += slen;
      }
      const bool scat = ph.tau >= ph.tau_t;
      ph.px = nx;
      ph.py = ny;
      ph.pz = nz;
      if (!scat) {
        ph.s++;
        continue;
      }
      c.sc++;
      e.et = ALB_SCATTER;
      e.x = ph.px; e.y = ph.py; e.z = ph.pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
      e.E = ph.E;
      e.t = (float)ph.s;
      e.eno = ph.event_no++;
      snapshot<STATS, HEAT>(e, ph);
      const double dep2 = (ph.py - y_ws) * -1.0;
      const int li2 = layer_of(Ls, nl, dep2);
      const alb_layer_t& L2 = Ls[li2];
      const double zl2 = dep2 - L2.top;
      if constexpr (STATS) e.rdl = stat_in(stat_layer, li) ? 1 : 0;  // the free path ended in the sub-step's layer
      const double u = uniform(rng, st);
      double theta;
      if constexpr (STATS) {  // phase-sampling score of the adjoint (one constant layer: the scene's table)
        int seg;
        theta = interp_seg(u, L2.phase_cdf, L2.phase_theta, L2.phase_n, &seg);
        if (sc.phase_score) ph.gs += sc.phase_score[seg];
      } else {
        theta = interp(u, L2.phase_cdf, L2.phase_theta, L2.phase_n);
      }
      const double phi = two_pi * uniform(rng, st);
      double sth, cth, sph, cph;
      sincos(theta, &sth, &cth);
      sincos(phi, &sph, &cph);
      float Rm[9];
      rotation_from_z(ph.dx, ph.dy, ph.dz, Rm);
      const double l0 = sth * cph, l1 = sth * sph, l2 = cth;
      double w0, w1, w2;
      w0 = (double)Rm[0] * l0; w0 = w0 + (double)Rm[1] * l1; w0 = w0 + (double)Rm[2] * l2;
      w1 = (double)Rm[3] * l0; w1 = w1 + (double)Rm[4] * l1; w1 = w1 + (double)Rm[5] * l2;
      w2 = (double)Rm[6] * l0; w2 = w2 + (double)Rm[7] * l1; w2 = w2 + (double)Rm[8] * l2;
      ph.dx = (double)(float)w0;
      ph.dy = (double)(float)w1;
      ph.dz = (double)(float)w2;
      if constexpr (STATS) {  // omega0 of the new position's layer; the density of the sub-step's
        if (stat_in(stat_layer, li2)) ph.nsc++;
        if (stat_in(stat_layer, li)) ph.dsc++;
      }
      ph.E = ph.E * layer_omega0(L2, zl2);
      ph.tau = ph.tau - ph.tau_t;
      ph.tau_t = tau_target(uniform(rng, st));
      ph.s++;
      break;
    }
    if (state == 2) {
      // handle_seafloor (core/simulation.py:504-570)
      c.fl++;
      double f = sc.tilted ? (sc.floor_offset - floor_dot(sc, ph.px, ph.py, ph.pz)) / floor_dot(sc, svx, svy, svz)
                           : (y_sf - ph.py) / svy;
      f = clipd(f, 0.0, 1.0);
      const double Px = ph.px + f * svx, Py = ph.py + f * svy, Pz = ph.pz + f * svz;
      if constexpr (HEAT) {  // seafloor interaction, set before the record (core/simulation.py:521)
        ph.bx = (float)Px;
        ph.by = (float)Py;
        ph.bz = (float)Pz;
        if (fwd && sc.heat_radius) count_radius(sc, ph, Px, Pz);
      }
      e.et = ALB_BOTTOM_REFLECTION;
      e.x = Px; e.y = Py; e.z = Pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
      e.E = ph.E;
      e.t = (float)ph.s;
      e.eno = ph.event_no++;
      snapshot<STATS, HEAT>(e, ph);
      const double u1 = uniform(rng, st);
      const double u2 = uniform(rng, st);
      const double sr = sqrt(u1);
      double sp, cp;
      sincos(two_pi * u2, &sp, &cp);
      if (!sc.tilted) {
        ph.dx = sr * cp;
        ph.dy = sqrt(1.0 - u1);
        ph.dz = sr * sp;
      } else {  // the same hemisphere sample about the plane's normal
        const double lx = sr * cp, ly = sqrt(1.0 - u1), lz = sr * sp;
        ph.dx = (sc.floor_t1[0] * lx + sc.floor_normal[0] * ly) + sc.floor_t2[0] * lz;
        ph.dy = (sc.floor_t1[1] * lx + sc.floor_normal[1] * ly) + sc.floor_t2[1] * lz;
        ph.dz = (sc.floor_t1[2] * lx + sc.floor_normal[2] * ly) + sc.floor_t2[2] * lz;
      }
      const double rem = (1.0 - f) * slen;
      ph.px = Px + ph.dx * rem;
      ph.py = Py + ph.dy * rem;
      ph.pz = Pz + ph.dz * rem;
      const double dep2 = (ph.py - y_ws) * -1.0;
      const int li2 = layer_of(Ls, nl, dep2);
      const alb_layer_t& L2 = Ls[li2];
      ph.tau = (float)((double)ph.tau + (double)layer_alpha32(L2, dep2 - L2.top) * slen);
      if constexpr (STATS) {
        if (stat_in(stat_layer, li2)) ph.lam += slen;
        ph.nfl++;
      }
      ph.E = ph.E * (float)sc.floor_albedo;
      ph.s++;
      break;
    }
    // handle_exit (core/simulation.py:433-502)
    c.ex++;
    double f = (y_ws - ph.py) / svy;
    f = clipd(f, 0.0, 1.0);
    const double Px = ph.px + f * svx, Py = ph.py + f * svy, Pz = ph.pz + f * svz;
    const double R = schlick(ph.dy, sc.F0);
    const bool reflect = uniform(rng, st) < R;
    const uint32_t no = ph.event_no++;
    if (!reflect && fwd) {
      e.et = ALB_SURFACE_TRANSMISSION_UP;
      e.x = Px; e.y = Py; e.z = Pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
      e.E = ph.E;
      e.t = (float)ph.s;
      e.eno = no;
      snapshot<STATS, HEAT>(e, ph);
    }
    if (reflect) {
      ph.dy = -ph.dy;
    } else {
      refract_y(ph.dx, ph.dy, ph.dz, -1.0, sc.eta);
    }
    const double rem = (1.0 - f) * slen;
    ph.px = Px + ph.dx * rem;
    ph.py = Py + ph.dy * rem;
    ph.pz = Pz + ph.dz * rem;
    if (reflect) {
      const double dep2 = (ph.py - y_ws) * -1.0;
      const int li2 = layer_of(Ls, nl, dep2);
      const alb_layer_t& L2 = Ls[li2];
      ph.tau = (float)((double)ph.tau + (double)layer_alpha32(L2, dep2 - L2.top) * slen);
      if constexpr (STATS) {
        if (stat_in(stat_layer, li2)) ph.lam += slen;
      }
    }
    ph.s++;
    break;
  }
  if (st.exhausted) {
    ph.alive = false;
    c.exh++;
  }
}

// Advance a live photon to its next event and apply it; fills e when the
// event leaves a record (forward) or a query (backward).  STATS: also keep the
// path statistics the adjoint needs (water length, scatter and floor counts,
// phase-sampling score) and snapshot them into e; the physics is unchanged.
// LAYERED: the scene may be a layered column or have a tilted floor (checked
// at run time); without it the kernel carries only the flat floor's
// closed-form water legs.  stat_layer >= 0 restricts the water length and
// scatter statistics to that layer of a layered column (a single constant
// layer is layer 0).
template <bool STATS, bool LAYERED = true, bool HEAT = false>
__device__ __forceinline__ void next_event(const alb_scene_t& sc, bool fwd, const alb_rng_t& rng, Stream& st,
                                           Photon& ph, Emit& e, Tally& c, int stat_layer = -1) {
  const double two_pi = 2.0 * ALB_PI;
  const int64_t S = sc.n_substeps;
  const double y_ws = sc.y_ws, y_sf = sc.y_sf;
  const double alpha = (double)sc.alpha_l;
  e.et = -1;
  if (ph.py > y_ws) {
    // ------------------------------------------------------------ air leg
    if (ph.dy >= 0.0) {
      ph.alive = false;  // above the surface, not descending: no more events
      return;
    }
    const double Dx = (ph.dx * sc.v_air) * sc.dt, Dy = (ph.dy * sc.v_air) * sc.dt, Dz = (ph.dz * sc.v_air) * sc.dt;
    const int64_t avail = S - ph.s;
    const int64_t j = first_step_below(ph.py, Dy, y_ws, avail);
    if (j >= avail) {
      ph.alive = false;
      return;
    }
    c.air++;
    c.en++;
    // handle_enter (core/simulation.py:375-431)
    const double qx = ph.px + (double)j * Dx, qy = ph.py + (double)j * Dy, qz = ph.pz + (double)j * Dz;
    const double nx = qx + Dx, ny = qy + Dy, nz = qz + Dz;
    const double svx = nx - qx, svy = ny - qy, svz = nz - qz;
    double f = (y_ws - qy) / svy;
    f = clipd(f, 0.0, 1.0);
    const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
    if constexpr (HEAT) {  // first water interaction, set before the record (core/simulation.py:393)
      ph.fx = (float)Px;
      ph.fy = (float)Py;
      ph.fz = (float)Pz;
    }
    const double R = schlick(-ph.dy, sc.F0);
    const double T = 1.0 - R;
    e.et = ALB_SURFACE_REFLECTION;
    e.x = Px; e.y = Py; e.z = Pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
    e.E = ph.E;
    e.t = (float)(ph.s + j);
    e.eno = ph.event_no++;
    snapshot<STATS, HEAT>(e, ph);
    refract_y(ph.dx, ph.dy, ph.dz, 1.0, sc.eta);
    ph.E = (float)((double)ph.E * T);
    const double len = norm3(svx, svy, svz);
    const double rem = (1.0 - f) * len;
    ph.px = Px + ph.dx * rem;
    ph.py = Py + ph.dy * rem;
    ph.pz = Pz + ph.dz * rem;
    ph.tau_t = tau_target(uniform(rng, st));
    ph.s = ph.s + j + 1;
  } else if (!(LAYERED && sc.tilted) && ph.py <= y_sf && ph.dy <= 0.0) {
    ph.alive = false;  // below the floor and not rising: the reference finds no event
    return;
  } else if (LAYERED && sc.layers) {
    water_stepped<STATS, HEAT>(sc, fwd, rng, st, ph, e, c, stat_layer);
    return;
  } else {
    // ---------------------------------------------------------- water leg
    const double Dx = (ph.dx * sc.v_water) * sc.dt, Dy = (ph.dy * sc.v_water) * sc.dt,
                 Dz = (ph.dz * sc.v_water) * sc.dt;
    const int64_t avail = S - ph.s;
    int64_t jb;
    bool to_floor;
    if (LAYERED && sc.tilted) {
      // A single constant layer over a tilted plane: the plane's signed
      // distance is linear in the sub-step index, so the leg keeps its closed
      // form, with the per-sub-step classification of the stepped legs (in
      // water while the step ends on the water side, the floor when it
      // crosses to the far side, the surface when it rises above it, which
      // wins a tie).
      const double g0 = floor_dist(sc, ph.px, ph.py, ph.pz);
      if (!(g0 > 0.0) && !(ph.py + Dy > y_ws) && !(floor_dist(sc, ph.px + Dx, ph.py + Dy, ph.pz + Dz) > 0.0)) {
        ph.alive = false;  // on the far side of the plane (entered over dry ground) and not crossing back
        return;
      }
      const double gD = floor_dot(sc, Dx, Dy, Dz);
      const int64_t jf = gD < 0.0 ? first_step_through_plane(sc, ph.px, ph.py, ph.pz, Dx, Dy, Dz, g0, gD, avail) : avail;
      const int64_t je = Dy > 0.0 ? first_step_above(ph.py, Dy, y_ws, avail) : avail;
      to_floor = jf < je;
      jb = to_floor ? jf : je;
    } else {
      if (Dy < 0.0) jb = first_step_below(ph.py, Dy, y_sf, avail);
      else if (Dy > 0.0) jb = first_step_above(ph.py, Dy, y_ws, avail);
      else jb = avail;
      to_floor = Dy < 0.0;
    }
    const int64_t jmax = jb < avail ? jb : avail;
    int64_t j = 0;
    bool scattered = false;
    const double len = norm3(Dx, Dy, Dz);
    const double inc = alpha * len;
    if (!__builtin_isinf(ph.tau_t)) {
      bool hit = false;
      const int64_t taken = advance_optical_depth(ph.tau, ph.tau_t, inc, jmax, &hit, c.solve);
      scattered = hit;
      j = hit ? taken - 1 : taken;  // index of the scattering sub-step
    } else {
      j = jmax;
    }
    c.ws += scattered ? (unsigned long long)(j + 1) : (unsigned long long)j;
    if (scattered) {
      // handle_water scatter branch (core/simulation.py:349-371)
      c.sc++;
      const double a1 = (double)(j + 1);
      ph.px = ph.px + a1 * Dx;
      ph.py = ph.py + a1 * Dy;
      ph.pz = ph.pz + a1 * Dz;
      e.et = ALB_SCATTER;
      e.x = ph.px; e.y = ph.py; e.z = ph.pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
      e.E = ph.E;
      e.t = (float)(ph.s + j);
      e.eno = ph.event_no++;
      const bool in0 = stat_in(stat_layer, 0);  // the closed-form legs: one constant layer
      if constexpr (STATS) {
        if (in0) ph.lam += a1 * len;
      }
      snapshot<STATS, HEAT>(e, ph);
      if constexpr (STATS) e.rdl = in0 ? 1 : 0;
      const double u = uniform(rng, st);
      double theta;
      if constexpr (STATS) {
        int seg;
        theta = interp_seg(u, sc.phase_cdf, sc.phase_theta, sc.phase_n, &seg);
        if (sc.phase_score) ph.gs += sc.phase_score[seg];
        if (in0) {
          ph.nsc++;
          ph.dsc++;
        }
      } else {
        theta = interp(u, sc.phase_cdf, sc.phase_theta, sc.phase_n);
      }
      const double phi = two_pi * uniform(rng, st);
      double sth, cth, sph, cph;
      sincos(theta, &sth, &cth);
      sincos(phi, &sph, &cph);
      float Rm[9];
      rotation_from_z(ph.dx, ph.dy, ph.dz, Rm);
      const double l0 = sth * cph, l1 = sth * sph, l2 = cth;
      double w0, w1, w2;
      w0 = (double)Rm[0] * l0; w0 = w0 + (double)Rm[1] * l1; w0 = w0 + (double)Rm[2] * l2;
      w1 = (double)Rm[3] * l0; w1 = w1 + (double)Rm[4] * l1; w1 = w1 + (double)Rm[5] * l2;
      w2 = (double)Rm[6] * l0; w2 = w2 + (double)Rm[7] * l1; w2 = w2 + (double)Rm[8] * l2;
      ph.dx = (double)(float)w0;
      ph.dy = (double)(float)w1;
      ph.dz = (double)(float)w2;
      ph.E = ph.E * sc.omega0;
      ph.tau = ph.tau - ph.tau_t;
      ph.tau_t = tau_target(uniform(rng, st));
      ph.s = ph.s + j + 1;
    } else if (jb < avail) {
      const double qx = ph.px + (double)jb * Dx, qy = ph.py + (double)jb * Dy, qz = ph.pz + (double)jb * Dz;
      const double nx = qx + Dx, ny = qy + Dy, nz = qz + Dz;
      const double svx = nx - qx, svy = ny - qy, svz = nz - qz;
      const double slen = norm3(svx, svy, svz);
      if constexpr (STATS) {
        if (stat_in(stat_layer, 0)) ph.lam += (double)jb * len;
      }
      if (to_floor) {
        // handle_seafloor (core/simulation.py:504-570)
        c.fl++;
        double f = (LAYERED && sc.tilted) ? (sc.floor_offset - floor_dot(sc, qx, qy, qz)) / floor_dot(sc, svx, svy, svz)
                                          : (y_sf - qy) / svy;
        f = clipd(f, 0.0, 1.0);
        const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
        if constexpr (HEAT) {  // seafloor interaction, set before the record (core/simulation.py:521)
          ph.bx = (float)Px;
          ph.by = (float)Py;
          ph.bz = (float)Pz;
          if (fwd && sc.heat_radius) count_radius(sc, ph, Px, Pz);
        }
        e.et = ALB_BOTTOM_REFLECTION;
        e.x = Px; e.y = Py; e.z = Pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
        e.E = ph.E;
        e.t = (float)(ph.s + jb);
        e.eno = ph.event_no++;
        snapshot<STATS, HEAT>(e, ph);
        if constexpr (STATS) {
          if (stat_in(stat_layer, 0)) ph.lam += slen;
          ph.nfl++;
        }
        const double u1 = uniform(rng, st);
        const double u2 = uniform(rng, st);
        const double sr = sqrt(u1);
        double sp, cp;
        sincos(two_pi * u2, &sp, &cp);
        if (!(LAYERED && sc.tilted)) {
          ph.dx = sr * cp;
          ph.dy = sqrt(1.0 - u1);
          ph.dz = sr * sp;
        } else {  // the same hemisphere sample about the plane's normal
          const double lx = sr * cp, ly = sqrt(1.0 - u1), lz = sr * sp;
          ph.dx = (sc.floor_t1[0] * lx + sc.floor_normal[0] * ly) + sc.floor_t2[0] * lz;
          ph.dy = (sc.floor_t1[1] * lx + sc.floor_normal[1] * ly) + sc.floor_t2[1] * lz;
          ph.dz = (sc.floor_t1[2] * lx + sc.floor_normal[2] * ly) + sc.floor_t2[2] * lz;
        }
        const double rem = (1.0 - f) * slen;
        ph.px = Px + ph.dx * rem;
        ph.py = Py + ph.dy * rem;
        ph.pz = Pz + ph.dz * rem;
        ph.tau = (float)((double)ph.tau + alpha * slen);
        ph.E = ph.E * (float)sc.floor_albedo;
      } else {
        // handle_exit (core/simulation.py:433-502)
        c.ex++;
        double f = (y_ws - qy) / svy;
        f = clipd(f, 0.0, 1.0);
        const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
        const double R = schlick(ph.dy, sc.F0);
        const bool reflect = uniform(rng, st) < R;
        const uint32_t no = ph.event_no++;
        if (!reflect && fwd) {
          e.et = ALB_SURFACE_TRANSMISSION_UP;
          e.x = Px; e.y = Py; e.z = Pz; e.dx = ph.dx; e.dy = ph.dy; e.dz = ph.dz;
          e.E = ph.E;
          e.t = (float)(ph.s + jb);
          e.eno = no;
          snapshot<STATS, HEAT>(e, ph);
        }
        if constexpr (STATS) {
          if (reflect && stat_in(stat_layer, 0)) ph.lam += slen;
        }
        if (reflect) {
          ph.dy = -ph.dy;
        } else {
          refract_y(ph.dx, ph.dy, ph.dz, -1.0, sc.eta);
        }
        const double rem = (1.0 - f) * slen;
        ph.px = Px + ph.dx * rem;
        ph.py = Py + ph.dy * rem;
        ph.pz = Pz + ph.dz * rem;
        if (reflect) ph.tau = (float)((double)ph.tau + alpha * slen);
      }
      ph.s = ph.s + jb + 1;
    } else {
      ph.alive = false;  // time budget used up in water
    }
  }
  if (st.exhausted) {
    ph.alive = false;  // the record of this event still leaves, as in the oracle
    c.exh++;
  }
}

// Wave-aggregated emission: one atomic per photon type per wave.  A backward
// surface entry goes to both surface types.
template <bool FWD, bool HEAT = false>
__device__ __forceinline__ void emit_record(const Sinks& sk, const Emit& e, int64_t tag, int32_t scene,
                                            uint32_t lane) {
#pragma unroll
  for (int t = 0; t < ALB_N_TYPES; ++t) {
    const bool mine = (e.et == t) || (!FWD && t == ALB_SURFACE_TRANSMISSION_UP && e.et == ALB_SURFACE_REFLECTION);
    const uint64_t m = ballot(mine);
    if (m == 0) continue;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    unsigned long long* cnt = FWD ? sk.rec[t].count : sk.q[t].count;
    if ((int)lane == leader) base = atomicAdd(cnt, (unsigned long long)__popcll(m));
    base = bcast_u64(base, leader);
    if (mine) {
      const int64_t slot = (int64_t)(base + mask_rank(m));
      if (FWD) {
        const alb_records_t& r = sk.rec[t];
        if (slot < r.capacity) {
          r.x[slot] = (float)e.x; r.y[slot] = (float)e.y; r.z[slot] = (float)e.z;
          r.dx[slot] = (float)e.dx; r.dy[slot] = (float)e.dy; r.dz[slot] = (float)e.dz;
          r.energy[slot] = e.E;
          r.time[slot] = e.t;
          if (r.tag) r.tag[slot] = tag;
          if (r.scene) r.scene[slot] = scene;
          if constexpr (HEAT) {
            if (r.fwi_x) {
              r.fwi_x[slot] = e.fx; r.fwi_y[slot] = e.fy; r.fwi_z[slot] = e.fz;
              r.sfi_x[slot] = e.bx; r.sfi_y[slot] = e.by; r.sfi_z[slot] = e.bz;
            }
          }
        }
      } else {
        const alb_queries_t& r = sk.q[t];
        if (slot < r.capacity) {
          r.x[slot] = e.x; r.y[slot] = e.y; r.z[slot] = e.z;
          r.dx[slot] = e.dx; r.dy[slot] = e.dy; r.dz[slot] = e.dz;
          r.energy[slot] = e.E;
          r.time[slot] = e.t;
          if (r.tag) r.tag[slot] = tag;
          if (r.scene) r.scene[slot] = scene;
        }
      }
    }
  }
}

__device__ __forceinline__ void flush_tally(const Tally& c, unsigned long long* ctr, uint32_t lane) {
  if (!ctr) return;
  const unsigned long long v[8] = {c.paths, c.air, c.ws, c.sc, c.en, c.ex, c.fl, c.exh};  // widened
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const unsigned long long tot = wave_sum(v[i]);
    if (lane == 0 && tot) atomicAdd(&ctr[i], tot);
  }
  const unsigned long long ts = wave_sum((unsigned long long)c.solve);
  if (lane == 0 && ts) atomicAdd(&ctr[ALB_CTR_DEPTH_SOLVES], ts);
}

// One scene in a persistent launch; LAYERED selects the stepped water legs of
// a layered column, HEAT the interaction points and radius histogram of the
// heat maps.  Each wavefront owns a contiguous range of photons; a lane whose
// photon has ended takes the range's next index in the following iteration
// (a wave-uniform counter held in a register), so the wave's lanes stay busy
// while path lengths differ and the wave ends with its range.  (A lane per
// photon, with the wave waiting for its longest path, took twice as long;
// fixed per-lane lists of photons left each wave waiting for its slowest
// lane.)  The photon index alone selects the random stream, so the records
// do not depend on the grid.
// Occupancy: the single-layer trace is latency-bound (FP64 chains, Philox,
// the optical-depth jumps); at four waves per SIMD it spills some VGPRs and
// still runs 21 % faster than at two (1e8-photon forward pass 297 -> 235 ms
// together with the reciprocal counts of advance_optical_depth).  The layered
// and heat-map variants keep two waves (fewer spills).
template <bool FWD, bool LAYERED, bool HEAT>
__global__ __launch_bounds__(256, (LAYERED || HEAT) ? 2 : 4) void trace_kernel(alb_scene_t sc, alb_rng_t rng, int64_t begin,
                                                    int64_t n, Sinks sk, unsigned long long* ctr) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (int64_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t per = (n + waves - 1) / waves;
  const int64_t w_lo = wid * per < n ? wid * per : n;
  const int64_t w_hi = w_lo + per < n ? w_lo + per : n;
  int64_t w_next = w_lo;  // wave-uniform
  int64_t i = 0;
  const uint32_t lane = lane_id();
  Stream st{0, 0, 0u, false, 0.0};
  Photon ph;
  ph.alive = false;
  ph.event_no = 0;
  bool done = false;
  Tally c{};
  for (;;) {
    const bool need = !ph.alive && !done;
    const uint64_t m = ballot(need);
    if (m) {
      const int64_t idx = w_next + (int64_t)mask_rank(m);
      w_next += (int64_t)__popcll(m);
      if (need) {
        if (idx < w_hi) {
          i = idx;
          st = Stream{(uint64_t)(begin + i), i, 0u, false, 0.0};
          emit_photon(sc, FWD, rng, st, ph, c);
        } else {
          done = true;
        }
      }
    }
    if (!ballot(ph.alive || !done)) break;
    Emit e{-1, 0, 0, 0, 0, 0, 0, 0.0f, 0.0f, 0u};
    if (ph.alive) next_event<false, LAYERED, HEAT>(sc, FWD, rng, st, ph, e, c);
    const int64_t tag = (int64_t)(((uint64_t)(begin + i) << 16) | (uint64_t)(e.eno & 0xFFFFu));
    emit_record<FWD, HEAT>(sk, e, tag, 0, lane);
  }
  flush_tally(c, ctr, lane);
}

// Many scenes in one persistent launch.  Photons [offsets[s], offsets[s+1])
// of the flattened index belong to scene s; scene s's photon p uses Philox
// counter words {draw/2, p, p>>32, rng.stream + 2 s}, so with rng.stream =
// pass scene 0 reproduces the single-scene trace.  Lanes whose photon has finished take the next index
// from a global counter every refill_interval iterations, which keeps the
// wave's lanes busy when path lengths differ.
// Three waves per SIMD (168 VGPRs, a few spilled): the 256-scene sweep's
// forward pass 77 -> 68 ms against two waves; four were slower (more spills).
template <bool FWD>
__global__ __launch_bounds__(256, 3) void trace_batched_kernel(const alb_scene_t* __restrict__ scenes, int32_t n_scenes,
                                                            const int64_t* __restrict__ offsets, alb_rng_t rng,
                                                            Sinks sk, unsigned long long* work,
                                                            int32_t refill_interval, unsigned long long* ctr) {
  const uint32_t lane = lane_id();
  const int64_t total = offsets[n_scenes];
  Photon ph;
  ph.alive = false;
  ph.event_no = 0;
  bool done = false;
  int32_t sid = 0;
  int64_t pidx = 0;
  Stream st{0, 0, 0u, false, 0.0};
  alb_rng_t lrng = rng;
  Tally c{};
  for (uint32_t it = 0;; ++it) {
    if (it % (uint32_t)refill_interval == 0) {
      const bool need = !ph.alive && !done;
      const uint64_t m = ballot(need);
      if (m) {
        const int leader = __ffsll((long long)m) - 1;
        unsigned long long base = 0;
        if ((int)lane == leader) base = atomicAdd(work, (unsigned long long)__popcll(m));
        base = bcast_u64(base, leader);
        if (need) {
          const int64_t g = (int64_t)(base + mask_rank(m));
          if (g >= total) {
            done = true;
          } else {
            int lo = 0, hi = n_scenes;  // offsets[lo] <= g < offsets[hi]
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (offsets[mid] <= g) lo = mid; else hi = mid;
            }
            sid = lo;
            pidx = g - offsets[lo];
            lrng.stream = rng.stream + ((uint32_t)sid << 1);
            st = Stream{(uint64_t)pidx, 0, 0u, false, 0.0};
            emit_photon(scenes[sid], FWD, lrng, st, ph, c);
          }
        }
      }
    }
    if (!ballot(ph.alive || !done)) break;
    Emit e{-1, 0, 0, 0, 0, 0, 0, 0.0f, 0.0f, 0u};
    if (ph.alive) next_event<false>(scenes[sid], FWD, lrng, st, ph, e, c);
    const int64_t tag = (int64_t)(((uint64_t)pidx << 16) | (uint64_t)(e.eno & 0xFFFFu));
    emit_record<FWD>(sk, e, tag, sid, lane);
  }
  flush_tally(c, ctr, lane);
}

// first position in keys[0, n) whose key is >= key
__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* __restrict__ keys, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

struct ReplayIdx {
  alb_replay_index_t t[ALB_N_TYPES];
};

// Per-scene sums of weight * (n, [scatter], L, m, G) of the lane's records.
constexpr int kReplaySums = 6;  // per scene: nsc, [scatter record], lam, nfl, gs, dsc (alb_replay_batched)

struct ReplayAcc {
  double v[kReplaySums];
  int32_t sid;
};

__device__ __forceinline__ void flush_acc(ReplayAcc& a, double* acc) {
  if (a.sid >= 0) {
#pragma unroll
    for (int c = 0; c < kReplaySums; ++c)
      if (a.v[c] != 0.0) unsafeAtomicAdd(&acc[kReplaySums * (int64_t)a.sid + c], a.v[c]);
  }
#pragma unroll
  for (int c = 0; c < kReplaySums; ++c) a.v[c] = 0.0;
}

// Path replay for the adjoint: the photons of trace_batched_kernel, traced
// again with the same streams (so the same events), now keeping the path
// statistics; every record / query is found through the sorted key index
// (the photon's records of one type are consecutive there, in event order,
// so one binary search per photon and type places a cursor) and its adjoint
// weight multiplies the prefix statistics.
template <bool FWD>
__global__ __launch_bounds__(256, 3) void replay_batched_kernel(const alb_scene_t* __restrict__ scenes, int32_t n_scenes,
                                                             const int64_t* __restrict__ offsets, alb_rng_t rng,
                                                             ReplayIdx ix, unsigned long long* work,
                                                             int32_t refill_interval, double* acc,
                                                             unsigned long long* ctr, int32_t stat_layer) {
  const uint32_t lane = lane_id();
  const int64_t total = offsets[n_scenes];
  Photon ph;
  ph.alive = false;
  ph.event_no = 0;
  bool done = false;
  int32_t sid = 0;
  int64_t g = 0;
  int64_t cur[ALB_N_TYPES] = {0, 0, 0, 0};
  Stream st{0, 0, 0u, false, 0.0};
  alb_rng_t lrng = rng;
  Tally c{};
  ReplayAcc a{{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, -1};
  unsigned long long misses = 0;
  for (uint32_t it = 0;; ++it) {
    if (it % (uint32_t)refill_interval == 0) {
      const bool need = !ph.alive && !done;
      const uint64_t m = ballot(need);
      if (m) {
        const int leader = __ffsll((long long)m) - 1;
        unsigned long long base = 0;
        if ((int)lane == leader) base = atomicAdd(work, (unsigned long long)__popcll(m));
        base = bcast_u64(base, leader);
        if (need) {
          g = (int64_t)(base + mask_rank(m));
          if (g >= total) {
            done = true;
          } else {
            int lo = 0, hi = n_scenes;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (offsets[mid] <= g) lo = mid; else hi = mid;
            }
            sid = lo;
            if (a.sid != sid) {
              flush_acc(a, acc);
              a.sid = sid;
            }
            const int64_t pidx = g - offsets[lo];
            lrng.stream = rng.stream + ((uint32_t)sid << 1);
            st = Stream{(uint64_t)pidx, 0, 0u, false, 0.0};
#pragma unroll
            for (int t = 0; t < ALB_N_TYPES; ++t)
              cur[t] = ix.t[t].n > 0 ? lower_bound_u64(ix.t[t].keys, ix.t[t].n, (uint64_t)g << 16) : 0;
            emit_photon(scenes[sid], FWD, lrng, st, ph, c);
            ph.lam = 0.0;
            ph.gs = 0.0;
            ph.nsc = 0;
            ph.nfl = 0;
            ph.dsc = 0;
          }
        }
      }
    }
    if (!ballot(ph.alive || !done)) break;
    if (ph.alive) {
      Emit e{-1, 0, 0, 0, 0, 0, 0, 0.0f, 0.0f, 0u, 0.0, 0.0, 0, 0, 0, 0};
      next_event<true, true>(scenes[sid], FWD, lrng, st, ph, e, c, stat_layer);  // tilted floors too
      if (e.et >= 0) {
        const uint64_t key = ((uint64_t)g << 16) | (uint64_t)(e.eno & 0xFFFFu);
#pragma unroll
        for (int t = 0; t < ALB_N_TYPES; ++t) {
          const bool mine = (e.et == t) || (!FWD && t == ALB_SURFACE_TRANSMISSION_UP && e.et == ALB_SURFACE_REFLECTION);
          if (!mine || ix.t[t].n == 0) continue;
          const int64_t p = cur[t];
          if (p < ix.t[t].n && ix.t[t].keys[p] == key) {
            cur[t] = p + 1;
            const double A = ix.t[t].weight[ix.t[t].slots[p]];
            if (A != 0.0) {
              a.v[0] += A * (double)e.nsc;
              if (t == ALB_SCATTER && e.rdl) a.v[1] += A;
              a.v[2] += A * e.lam;
              a.v[3] += A * (double)e.nfl;
              a.v[4] += A * e.gs;
              a.v[5] += A * (double)e.dsc;
            }
          } else {
            misses++;
          }
        }
      }
    }
  }
  flush_acc(a, acc);
  if (ctr) {
    const unsigned long long tm = wave_sum(misses);
    if (lane == 0 && tm) atomicAdd(&ctr[ALB_CTR_REPLAY_MISSES], tm);
  }
}

static int fill_sinks(int32_t pass, const alb_records_t* out, const alb_queries_t* q, Sinks* sk, const char* who) {
  for (int t = 0; t < ALB_N_TYPES; ++t) {
    if (pass == ALB_PASS_FORWARD) {
      if (!out || !out[t].count) return set_error(ALB_ERR_ARG, "%s: forward pass needs 4 record sinks", who);
      sk->rec[t] = out[t];
      sk->q[t] = alb_queries_t{};
    } else {
      if (!q || !q[t].count) return set_error(ALB_ERR_ARG, "%s: backward pass needs 4 query sinks", who);
      sk->q[t] = q[t];
      sk->rec[t] = alb_records_t{};
    }
  }
  return ALB_OK;
}

}  // namespace alb

extern "C" int alb_trace(const alb_scene_t* scene, const alb_rng_t* rng, int32_t pass,
                         int64_t photon_begin, int64_t n, const alb_records_t* out,
                         const alb_queries_t* q, unsigned long long* counters, void* stream) {
  using namespace alb;
  if (!scene || !rng) return set_error(ALB_ERR_ARG, "alb_trace: null scene or rng");
  if (n < 0 || photon_begin < 0) return set_error(ALB_ERR_ARG, "alb_trace: negative photon range");
  if (pass != ALB_PASS_FORWARD && pass != ALB_PASS_BACKWARD)
    return set_error(ALB_ERR_ARG, "alb_trace: pass must be 0 (forward) or 1 (backward)");
  if (rng->mode == ALB_RNG_REPLAY && (!rng->replay || rng->replay_len <= 0))
    return set_error(ALB_ERR_ARG, "alb_trace: replay mode needs a replay buffer");
  if (!scene->phase_cdf || !scene->phase_theta || scene->phase_n < 2)
    return set_error(ALB_ERR_ARG, "alb_trace: missing phase tables");
  if (scene->n_layers < 0 || (scene->n_layers > 0) != (scene->layers != nullptr))
    return set_error(ALB_ERR_ARG, "alb_trace: n_layers and layers disagree");
  if (scene->heat_bins < 0 || (scene->heat_bins > 0 && !(scene->heat_water_extent > 0.0 && scene->heat_floor_extent > 0.0)))
    return set_error(ALB_ERR_ARG, "alb_trace: heat maps need bins >= 0 and positive extents");
  if (photon_begin + n > ((int64_t)1 << 47))
    return set_error(ALB_ERR_ARG, "alb_trace: photon index exceeds 2^47");
  Sinks sk;
  int rc = fill_sinks(pass, out, q, &sk, "alb_trace");
  if (rc != ALB_OK) return rc;
  if (n == 0) return ALB_OK;
  const int block = 256;
  hipStream_t s = (hipStream_t)stream;
  const bool layered = scene->layers != nullptr || scene->tilted;
  // interaction points: written into forward records whose sinks hold them;
  // floor hits fill the radius histogram when the scene has one
  const bool heat = pass == ALB_PASS_FORWARD && ((scene->heat_bins > 0 && scene->heat_radius) || out[0].fwi_x);
  auto kern = pass == ALB_PASS_FORWARD
                  ? (heat ? (layered ? trace_kernel<true, true, true> : trace_kernel<true, false, true>)
                          : (layered ? trace_kernel<true, true, false> : trace_kernel<true, false, false>))
                  : (layered ? trace_kernel<false, true, false> : trace_kernel<false, false, false>);
  // persistent grid: the resident blocks of every CU, fewer when n is small
  int dev = 0, cus = 256, per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_trace: no HIP device");
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, block, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  int64_t grid = (int64_t)cus * per_cu;
  const int64_t need = (n + block - 1) / block;
  if (grid > need) grid = need;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), 0, s, *scene, *rng, photon_begin, n, sk, counters);
  return check_launch("alb_trace");
}

extern "C" int alb_trace_batched(const alb_scene_t* scenes, int32_t n_scenes, const int64_t* photon_offsets,
                                 const alb_rng_t* rng, int32_t pass, const alb_records_t* out,
                                 const alb_queries_t* q, unsigned long long* work_counter, int32_t block,
                                 int32_t refill_interval, unsigned long long* counters, void* stream) {
  using namespace alb;
  if (!scenes || !photon_offsets || !rng || !work_counter)
    return set_error(ALB_ERR_ARG, "alb_trace_batched: null argument");
  if (n_scenes <= 0 || n_scenes > (1 << 20)) return set_error(ALB_ERR_ARG, "alb_trace_batched: bad scene count");
  if (pass != ALB_PASS_FORWARD && pass != ALB_PASS_BACKWARD)
    return set_error(ALB_ERR_ARG, "alb_trace_batched: pass must be 0 (forward) or 1 (backward)");
  if (rng->mode != ALB_RNG_PHILOX) return set_error(ALB_ERR_UNSUPPORTED, "alb_trace_batched: Philox streams only");
  if ((int32_t)(rng->stream & 1u) != pass)
    return set_error(ALB_ERR_ARG, "alb_trace_batched: rng stream word must have the pass in its low bit");
  if (block != 64 && block != 128 && block != 256)
    return set_error(ALB_ERR_ARG, "alb_trace_batched: block must be 64, 128 or 256");
  if (refill_interval < 1) return set_error(ALB_ERR_ARG, "alb_trace_batched: refill_interval must be >= 1");
  Sinks sk;
  int rc = fill_sinks(pass, out, q, &sk, "alb_trace_batched");
  if (rc != ALB_OK) return rc;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_trace_batched: no HIP device");
  int per_cu = 1;
  const void* fn = pass == ALB_PASS_FORWARD ? (const void*)trace_batched_kernel<true> : (const void*)trace_batched_kernel<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  const unsigned grid = (unsigned)(cus * per_cu);
  hipStream_t s = (hipStream_t)stream;
  if (pass == ALB_PASS_FORWARD)
    hipLaunchKernelGGL(trace_batched_kernel<true>, dim3(grid), dim3(block), 0, s, scenes, n_scenes, photon_offsets,
                       *rng, sk, work_counter, refill_interval, counters);
  else
    hipLaunchKernelGGL(trace_batched_kernel<false>, dim3(grid), dim3(block), 0, s, scenes, n_scenes, photon_offsets,
                       *rng, sk, work_counter, refill_interval, counters);
  return check_launch("alb_trace_batched");
}

extern "C" int alb_replay_batched(const alb_scene_t* scenes, int32_t n_scenes, const int64_t* photon_offsets,
                                  const alb_rng_t* rng, int32_t pass, const alb_replay_index_t* index,
                                  unsigned long long* work_counter, int32_t block, int32_t refill_interval,
                                  double* acc, unsigned long long* counters, int32_t stat_layer, void* stream) {
  using namespace alb;
  if (!scenes || !photon_offsets || !rng || !index || !work_counter || !acc)
    return set_error(ALB_ERR_ARG, "alb_replay_batched: null argument");
  if (stat_layer < -1) return set_error(ALB_ERR_ARG, "alb_replay_batched: stat_layer must be -1 or a layer index");
  if (n_scenes <= 0 || n_scenes > (1 << 20)) return set_error(ALB_ERR_ARG, "alb_replay_batched: bad scene count");
  if (pass != ALB_PASS_FORWARD && pass != ALB_PASS_BACKWARD)
    return set_error(ALB_ERR_ARG, "alb_replay_batched: pass must be 0 (forward) or 1 (backward)");
  if (rng->mode != ALB_RNG_PHILOX) return set_error(ALB_ERR_UNSUPPORTED, "alb_replay_batched: Philox streams only");
  if ((int32_t)(rng->stream & 1u) != pass)
    return set_error(ALB_ERR_ARG, "alb_replay_batched: rng stream word must have the pass in its low bit");
  if (block != 64 && block != 128 && block != 256)
    return set_error(ALB_ERR_ARG, "alb_replay_batched: block must be 64, 128 or 256");
  if (refill_interval < 1) return set_error(ALB_ERR_ARG, "alb_replay_batched: refill_interval must be >= 1");
  ReplayIdx ix;
  for (int t = 0; t < ALB_N_TYPES; ++t) {
    ix.t[t] = index[t];
    if (ix.t[t].n < 0 || (ix.t[t].n > 0 && (!ix.t[t].keys || !ix.t[t].slots || !ix.t[t].weight)))
      return set_error(ALB_ERR_ARG, "alb_replay_batched: index %d incomplete", t);
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_replay_batched: no HIP device");
  int per_cu = 1;
  const void* fn = pass == ALB_PASS_FORWARD ? (const void*)replay_batched_kernel<true>
                                            : (const void*)replay_batched_kernel<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  const unsigned grid = (unsigned)(cus * per_cu);
  hipStream_t s = (hipStream_t)stream;
  if (pass == ALB_PASS_FORWARD)
    hipLaunchKernelGGL(replay_batched_kernel<true>, dim3(grid), dim3(block), 0, s, scenes, n_scenes, photon_offsets,
                       *rng, ix, work_counter, refill_interval, acc, counters, stat_layer);
  else
    hipLaunchKernelGGL(replay_batched_kernel<false>, dim3(grid), dim3(block), 0, s, scenes, n_scenes, photon_offsets,
                       *rng, ix, work_counter, refill_interval, acc, counters, stat_layer);
  return check_launch("alb_replay_batched");
}
