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
                                                        bool* hit, unsigned long long& solves) {
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
        const double q = I / u;
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
    // sub-steps whose float64 sum stays inside the binade: t0 + c*d + I < top
    int64_t c = (int64_t)ceil((top - I - t0) / d);
    if (c < 1) c = 1;
    while (c > 1 && !(t0 + (double)(c - 1) * d + I < top)) --c;
    while (t0 + (double)c * d + I < top) ++c;
    // first sub-step reaching the target: t0 + nt*d >= tau_t, nt >= 1
    int64_t nt = 1;
    if (t0 + d < tt) {
      nt = (int64_t)ceil((tt - t0) / d);
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

template <bool FWD>
__global__ __launch_bounds__(256) void trace_kernel(alb_scene_t sc, alb_rng_t rng, int64_t begin,
                                                    int64_t n, Sinks sk, unsigned long long* ctr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool alive = i < n;
  Stream st{(uint64_t)(begin + i), i, 0u, false, 0.0};
  const double two_pi = 2.0 * ALB_PI;
  double px = 0.0, py = 0.0, pz = 0.0, dx = 0.0, dy = -1.0, dz = 0.0;
  unsigned long long c_paths = alive ? 1 : 0, c_air = 0, c_ws = 0, c_sc = 0, c_en = 0, c_ex = 0,
                     c_fl = 0, c_exh = 0, c_solve = 0;

  if (alive) {
    if (FWD) {
      // sample_directions_in_cone_gaussian (math/direction_cone.py:53-117)
      double theta;
      do {
        const double u = uniform(rng, st);
        theta = sc.rayleigh_sigma * sqrt(-2.0 * log(1.0 - u));
      } while (theta > sc.divergence && !st.exhausted);
      const double phi = two_pi * uniform(rng, st);
      double s, c, sp, cp;
      sincos(theta, &s, &c);
      sincos(phi, &sp, &cp);
      double w0, w1, w2;
      rotate(sc.rot_laser, s * cp, s * sp, c, w0, w1, w2);
      dx = (double)(float)w0;
      dy = (double)(float)w1;
      dz = (double)(float)w2;
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
      dx = k0 + (double)(float)w0;
      dy = k1 + (double)(float)w1;
      dz = k2 + (double)(float)w2;
    }
    if (st.exhausted) {
      alive = false;
      c_exh++;
    }
  }

  float E = 1.0f, tau = 0.0f, tau_t = __builtin_inff();
  int64_t s = 0;
  uint32_t event_no = 0;
  const int64_t S = sc.n_substeps;
  const double y_ws = sc.y_ws, y_sf = sc.y_sf;
  const double alpha = (double)sc.alpha_l;
  const uint32_t lane = lane_id();

  while (ballot(alive)) {
    int et = -1;  // photon type of this iteration's record / query
    double ex = 0, ey = 0, ez = 0, edx = 0, edy = 0, edz = 0;
    float eE = 0.0f, etime = 0.0f;
    uint32_t eno = 0;
    if (alive) {
      if (py > y_ws) {
        // ---------------------------------------------------------- air leg
        if (dy >= 0.0) {
          alive = false;  // above the surface, not descending: no more events
        } else {
          const double Dx = (dx * sc.v_air) * sc.dt, Dy = (dy * sc.v_air) * sc.dt, Dz = (dz * sc.v_air) * sc.dt;
          const int64_t avail = S - s;
          const int64_t j = first_step_below(py, Dy, y_ws, avail);
          if (j >= avail) {
            alive = false;
          } else {
            c_air++;
            c_en++;
            // handle_enter (core/simulation.py:375-431)
            const double qx = px + (double)j * Dx, qy = py + (double)j * Dy, qz = pz + (double)j * Dz;
            const double nx = qx + Dx, ny = qy + Dy, nz = qz + Dz;
            const double svx = nx - qx, svy = ny - qy, svz = nz - qz;
            double f = (y_ws - qy) / svy;
            f = clipd(f, 0.0, 1.0);
            const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
            const double R = schlick(-dy, sc.F0);
            const double T = 1.0 - R;
            et = ALB_SURFACE_REFLECTION;
            ex = Px; ey = Py; ez = Pz; edx = dx; edy = dy; edz = dz;
            eE = E;
            etime = (float)(s + j);
            eno = event_no++;
            refract_y(dx, dy, dz, 1.0, sc.eta);
            E = (float)((double)E * T);
            const double len = norm3(svx, svy, svz);
            const double rem = (1.0 - f) * len;
            px = Px + dx * rem;
            py = Py + dy * rem;
            pz = Pz + dz * rem;
            tau_t = tau_target(uniform(rng, st));
            s = s + j + 1;
          }
        }
      } else if (py <= y_sf && dy <= 0.0) {
        alive = false;  // below the floor and not rising: the reference finds no event
      } else {
        // -------------------------------------------------------- water leg
        const double Dx = (dx * sc.v_water) * sc.dt, Dy = (dy * sc.v_water) * sc.dt,
                     Dz = (dz * sc.v_water) * sc.dt;
        const int64_t avail = S - s;
        int64_t jb;
        if (Dy < 0.0) jb = first_step_below(py, Dy, y_sf, avail);
        else if (Dy > 0.0) jb = first_step_above(py, Dy, y_ws, avail);
        else jb = avail;
        const int64_t jmax = jb < avail ? jb : avail;
        int64_t j = 0;
        bool scattered = false;
        const double len = norm3(Dx, Dy, Dz);
        const double inc = alpha * len;
        if (!__builtin_isinf(tau_t)) {
          bool hit = false;
          const int64_t taken = advance_optical_depth(tau, tau_t, inc, jmax, &hit, c_solve);
          scattered = hit;
          j = hit ? taken - 1 : taken;  // index of the scattering sub-step
        } else {
          j = jmax;
        }
        c_ws += scattered ? (unsigned long long)(j + 1) : (unsigned long long)j;
        if (scattered) {
          // handle_water scatter branch (core/simulation.py:349-371)
          c_sc++;
          const double a1 = (double)(j + 1);
          px = px + a1 * Dx;
          py = py + a1 * Dy;
          pz = pz + a1 * Dz;
          et = ALB_SCATTER;
          ex = px; ey = py; ez = pz; edx = dx; edy = dy; edz = dz;
          eE = E;
          etime = (float)(s + j);
          eno = event_no++;
          const double u = uniform(rng, st);
          const double theta = interp(u, sc.phase_cdf, sc.phase_theta, sc.phase_n);
          const double phi = two_pi * uniform(rng, st);
          double sth, cth, sph, cph;
          sincos(theta, &sth, &cth);
          sincos(phi, &sph, &cph);
          float Rm[9];
          rotation_from_z(dx, dy, dz, Rm);
          const double l0 = sth * cph, l1 = sth * sph, l2 = cth;
          double w0, w1, w2;
          w0 = (double)Rm[0] * l0; w0 = w0 + (double)Rm[1] * l1; w0 = w0 + (double)Rm[2] * l2;
          w1 = (double)Rm[3] * l0; w1 = w1 + (double)Rm[4] * l1; w1 = w1 + (double)Rm[5] * l2;
          w2 = (double)Rm[6] * l0; w2 = w2 + (double)Rm[7] * l1; w2 = w2 + (double)Rm[8] * l2;
          dx = (double)(float)w0;
          dy = (double)(float)w1;
          dz = (double)(float)w2;
          E = E * sc.omega0;
          tau = tau - tau_t;
          tau_t = tau_target(uniform(rng, st));
          s = s + j + 1;
        } else if (jb < avail) {
          const double qx = px + (double)jb * Dx, qy = py + (double)jb * Dy, qz = pz + (double)jb * Dz;
          const double nx = qx + Dx, ny = qy + Dy, nz = qz + Dz;
          const double svx = nx - qx, svy = ny - qy, svz = nz - qz;
          const double slen = norm3(svx, svy, svz);
          if (Dy < 0.0) {
            // handle_seafloor (core/simulation.py:504-570)
            c_fl++;
            double f = (y_sf - qy) / svy;
            f = clipd(f, 0.0, 1.0);
            const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
            et = ALB_BOTTOM_REFLECTION;
            ex = Px; ey = Py; ez = Pz; edx = dx; edy = dy; edz = dz;
            eE = E;
            etime = (float)(s + jb);
            eno = event_no++;
            const double u1 = uniform(rng, st);
            const double u2 = uniform(rng, st);
            const double sr = sqrt(u1);
            double sp, cp;
            sincos(two_pi * u2, &sp, &cp);
            dx = sr * cp;
            dy = sqrt(1.0 - u1);
            dz = sr * sp;
            const double rem = (1.0 - f) * slen;
            px = Px + dx * rem;
            py = Py + dy * rem;
            pz = Pz + dz * rem;
            tau = (float)((double)tau + alpha * slen);
            E = E * (float)sc.floor_albedo;
          } else {
            // handle_exit (core/simulation.py:433-502)
            c_ex++;
            double f = (y_ws - qy) / svy;
            f = clipd(f, 0.0, 1.0);
            const double Px = qx + f * svx, Py = qy + f * svy, Pz = qz + f * svz;
            const double R = schlick(dy, sc.F0);
            const bool reflect = uniform(rng, st) < R;
            const uint32_t no = event_no++;
            if (!reflect && FWD) {
              et = ALB_SURFACE_TRANSMISSION_UP;
              ex = Px; ey = Py; ez = Pz; edx = dx; edy = dy; edz = dz;
              eE = E;
              etime = (float)(s + jb);
              eno = no;
            }
            if (reflect) {
              dy = -dy;
            } else {
              refract_y(dx, dy, dz, -1.0, sc.eta);
            }
            const double rem = (1.0 - f) * slen;
            px = Px + dx * rem;
            py = Py + dy * rem;
            pz = Pz + dz * rem;
            if (reflect) tau = (float)((double)tau + alpha * slen);
          }
          s = s + jb + 1;
        } else {
          alive = false;  // time budget used up in water
        }
      }
      if (st.exhausted) {
        alive = false;  // the record of this event still leaves, as in the oracle
        c_exh++;
      }
    }

    // ------------------------------------------------ wave-aggregated emission
    const int64_t tag = (int64_t)(((uint64_t)(begin + i) << 16) | (uint64_t)(eno & 0xFFFFu));
#pragma unroll
    for (int t = 0; t < ALB_N_TYPES; ++t) {
      const bool mine = (et == t) || (!FWD && t == ALB_SURFACE_TRANSMISSION_UP && et == ALB_SURFACE_REFLECTION);
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
            r.x[slot] = (float)ex; r.y[slot] = (float)ey; r.z[slot] = (float)ez;
            r.dx[slot] = (float)edx; r.dy[slot] = (float)edy; r.dz[slot] = (float)edz;
            r.energy[slot] = eE;
            r.time[slot] = etime;
            if (r.tag) r.tag[slot] = tag;
          }
        } else {
          const alb_queries_t& r = sk.q[t];
          if (slot < r.capacity) {
            r.x[slot] = ex; r.y[slot] = ey; r.z[slot] = ez;
            r.dx[slot] = edx; r.dy[slot] = edy; r.dz[slot] = edz;
            r.energy[slot] = eE;
            r.time[slot] = etime;
            if (r.tag) r.tag[slot] = tag;
          }
        }
      }
    }
  }

  if (ctr) {
    const unsigned long long v[8] = {c_paths, c_air, c_ws, c_sc, c_en, c_ex, c_fl, c_exh};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const unsigned long long tot = wave_sum(v[c]);
      if (lane == 0 && tot) atomicAdd(&ctr[c], tot);
    }
    const unsigned long long ts = wave_sum(c_solve);
    if (lane == 0 && ts) atomicAdd(&ctr[ALB_CTR_DEPTH_SOLVES], ts);
  }
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
  if (photon_begin + n > ((int64_t)1 << 47))
    return set_error(ALB_ERR_ARG, "alb_trace: photon index exceeds 2^47");
  Sinks sk;
  for (int t = 0; t < ALB_N_TYPES; ++t) {
    if (pass == ALB_PASS_FORWARD) {
      if (!out || !out[t].count) return set_error(ALB_ERR_ARG, "alb_trace: forward pass needs 4 record sinks");
      sk.rec[t] = out[t];
      sk.q[t] = alb_queries_t{};
    } else {
      if (!q || !q[t].count) return set_error(ALB_ERR_ARG, "alb_trace: backward pass needs 4 query sinks");
      sk.q[t] = q[t];
      sk.rec[t] = alb_records_t{};
    }
  }
  if (n == 0) return ALB_OK;
  const int block = 256;
  const int64_t grid = (n + block - 1) / block;
  hipStream_t s = (hipStream_t)stream;
  if (pass == ALB_PASS_FORWARD)
    hipLaunchKernelGGL(trace_kernel<true>, dim3((unsigned)grid), dim3(block), 0, s, *scene, *rng,
                       photon_begin, n, sk, counters);
  else
    hipLaunchKernelGGL(trace_kernel<false>, dim3((unsigned)grid), dim3(block), 0, s, *scene, *rng,
                       photon_begin, n, sk, counters);
  return check_launch("alb_trace");
}
