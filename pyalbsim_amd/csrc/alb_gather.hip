// alb_gather.hip — exact k-nearest-neighbour photon gather into waveforms.
//
// Replaces Simulation.sample_photons (src/alb_sim/core/simulation.py:132-195):
// for every sensor-path vertex, the k nearest map records (KDTree.query), the
// type-specific energy multiplier, the area / volume kernel 1/(pi r_k^2) or
// 1/(4/3 pi r_k^3) from the k-th distance, and np.add.at into the waveform.
//
// One wavefront handles one query:
//  1. seed: the query's Morton key is located in the sorted keys with a
//     64-ary wave search; the W records around it are the first candidates
//     and their k-th smallest squared distance is the pruning bound theta;
//  2. traversal of the radix tree, up to 64 nodes per step (one per lane),
//     pruning nodes whose exact float32 box is farther than theta; nodes of at
//     most 64 records are scanned one record per lane, and records with
//     d2 <= theta are appended to an LDS candidate buffer;
//  3. when the buffer fills it is bitonic-sorted by (d2, index) and cut back
//     to k entries, which tightens theta.  The final sort gives the neighbour
//     set in (distance, index) order — bit-identical to brute force, because
//     squared distances are evaluated in FP64 in one fixed order and box
//     distances are, by monotone rounding, never larger than those of the
//     records inside;
//  4. each neighbour's FP64 energy goes into an LDS-privatised window of the
//     waveform (ds_add_f64); bins outside the window use global FP64 atomics.
#include <limits.h>

#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

constexpr int kWaves = 4;         // wavefronts per block
constexpr int kStack = 512;       // per-wave node stack (LDS)
constexpr int kBuf = 512;         // per-wave candidate buffer (LDS)
constexpr int kBucket = 64;       // nodes with <= kBucket records are scanned
constexpr int kMaxWindow = 2048;  // LDS-privatised waveform bins
constexpr int kMaxK = kBuf - 128;

struct GatherArgs {
  alb_scene_t sc;
  alb_map_t map;
  alb_queries_t q;
  int64_t q_begin, q_n;
  int32_t type;
  double* wf;
  int64_t n_bins;
  int32_t win0, win_len;
  int32_t* nbr_out;
  double* radius_out;
  unsigned long long* ctr;
};

__device__ __forceinline__ double dist2(double qx, double qy, double qz, float x, float y, float z) {
  const double dx = (double)x - qx, dy = (double)y - qy, dz = (double)z - qz;
  double s = dx * dx + dy * dy;
  return s + dz * dz;
}

__device__ __forceinline__ double axis_gap(double q, float lo, float hi) {
  if (q < (double)lo) return (double)lo - q;
  if (q > (double)hi) return q - (double)hi;
  return 0.0;
}

__device__ __forceinline__ double box_dist2(double qx, double qy, double qz, const float* __restrict__ b) {
  const double gx = axis_gap(qx, b[0], b[3]);
  const double gy = axis_gap(qy, b[1], b[4]);
  const double gz = axis_gap(qz, b[2], b[5]);
  double s = gx * gx + gy * gy;
  return s + gz * gz;
}

__device__ __forceinline__ bool before(double da, int32_t ia, double db, int32_t ib) {
  return da < db || (da == db && ia < ib);
}

// Sort the first n entries (padded to a power of two) of (d, idx) ascending.
__device__ void wave_sort(double* d, int32_t* idx, int n, uint32_t lane) {
  int N = 64;
  while (N < n) N <<= 1;
  for (int t = n + (int)lane; t < N; t += 64) {
    d[t] = __builtin_inf();
    idx[t] = INT_MAX;
  }
  wave_sync();
  for (int size = 2; size <= N; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = (int)lane; t < (N >> 1); t += 64) {
        const int a = (t / stride) * 2 * stride + (t & (stride - 1));
        const int b = a + stride;
        const bool up = (a & size) == 0;
        const double da = d[a], db = d[b];
        const int32_t ia = idx[a], ib = idx[b];
        const bool sw = up ? before(db, ib, da, ia) : before(da, ia, db, ib);
        if (sw) {
          d[a] = db; d[b] = da;
          idx[a] = ib; idx[b] = ia;
        }
      }
      wave_sync();
    }
  }
}

// first index in [0, n) whose key is >= key (n if none); wave-cooperative
__device__ int64_t wave_lower_bound(const uint64_t* __restrict__ keys, int64_t n, uint64_t key, uint32_t lane) {
  int64_t lo = 0, hi = n;
  while (hi - lo > 64) {
    const int64_t step = (hi - lo + 63) / 64;
    const int64_t p = lo + (int64_t)lane * step;
    const bool less = p < hi && keys[p] < key;
    const int c = __popcll(ballot(less));
    if (c == 0) {
      hi = lo;
      break;
    }
    const int64_t nlo = lo + (int64_t)(c - 1) * step + 1;
    const int64_t pc = lo + (int64_t)c * step;
    hi = (c < 64 && pc < hi) ? pc : hi;
    lo = nlo;
  }
  const int64_t p = lo + (int64_t)lane;
  const bool less = p < hi && keys[p] < key;
  return lo + __popcll(ballot(less));
}

__global__ __launch_bounds__(256) void gather_kernel(GatherArgs a) {
  __shared__ double hist[kMaxWindow];
  __shared__ int32_t stack_s[kWaves][kStack];
  __shared__ double bd_s[kWaves][kBuf];
  __shared__ int32_t bi_s[kWaves][kBuf];
  __shared__ int32_t rlo_s[kWaves][64];
  __shared__ int32_t rhi_s[kWaves][64];

  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x) hist[t] = 0.0;
  __syncthreads();

  const uint32_t lane = lane_id();
  const int w = threadIdx.x >> 6;
  int32_t* stack = stack_s[w];
  double* bd = bd_s[w];
  int32_t* bi = bi_s[w];
  int32_t* rlo = rlo_s[w];
  int32_t* rhi = rhi_s[w];
  const alb_map_t& m = a.map;
  const int64_t n = m.n;
  const int k = a.sc.k;
  const int W = (int)(n < (int64_t)(((k + 63) / 64) * 64 + 64) ? n : (int64_t)(((k + 63) / 64) * 64 + 64));
  unsigned long long c_q = 0, c_cand = 0, c_nodes = 0, c_ovf = 0;

  for (int64_t qi = (int64_t)blockIdx.x * kWaves + w; qi < a.q_n; qi += (int64_t)gridDim.x * kWaves) {
    const int64_t qg = a.q_begin + qi;
    const double qx = a.q.x[qg], qy = a.q.y[qg], qz = a.q.z[qg];
    c_q++;

    // ---- 1. Morton-window seed
    const uint64_t key = morton_key(qx, qy, qz, m.meta);
    const int64_t pos = wave_lower_bound(m.keys, n, key, lane);
    int64_t w0 = pos - W / 2;
    if (w0 > n - W) w0 = n - W;
    if (w0 < 0) w0 = 0;
    const int64_t w1 = w0 + W - 1;
    int buf_n = 0;
    for (int t = 0; t < W; t += 64) {
      const int64_t r = w0 + t + (int64_t)lane;
      const bool ok = t + (int)lane < W;
      if (ok) {
        bd[t + lane] = dist2(qx, qy, qz, m.x[r], m.y[r], m.z[r]);
        bi[t + lane] = (int32_t)r;
      }
    }
    buf_n = W;
    c_cand += W;
    wave_sort(bd, bi, buf_n, lane);
    buf_n = k;
    double theta = bd[k - 1];

    // ---- 2. pruned traversal of the rest of the map
    if (W < n) {
      int sp = 0;
      if (lane == 0) stack[0] = 0;
      sp = 1;
      wave_sync();
      while (sp > 0) {
        const int p = (sp >= kStack - 128) ? 1 : (sp < 64 ? sp : 64);
        const bool have = (int)lane < p;
        const int32_t node = have ? stack[sp - 1 - (int)lane] : 0;
        sp -= p;
        wave_sync();
        int kind = 0;  // 1: expand children, 2: scan record range
        int32_t lo = 0, hi = -1, c0 = 0, c1 = 0;
        if (have) {
          if (node < 0) {
            const int32_t r = ~node;
            if (!(r >= w0 && r <= w1)) {
              kind = 2;
              lo = hi = r;
            }
          } else {
            lo = m.node_range[2 * node];
            hi = m.node_range[2 * node + 1];
            if (!(lo >= w0 && hi <= w1)) {
              const double md2 = box_dist2(qx, qy, qz, m.node_box + 6 * (int64_t)node);
              if (md2 <= theta) {
                if (hi - lo + 1 <= kBucket) {
                  kind = 2;
                } else {
                  kind = 1;
                  c0 = m.node_child[2 * node];
                  c1 = m.node_child[2 * node + 1];
                }
              }
            }
          }
        }
        c_nodes += (unsigned long long)p;
        const uint64_t m1 = ballot(kind == 1);
        const int np = 2 * __popcll(m1);
        if (sp + np > kStack) {  // cannot happen with the pop policy above; refuse rather than corrupt
          c_ovf++;
          break;
        }
        if (kind == 1) {
          const int off = sp + 2 * (int)mask_rank(m1);
          stack[off] = c1;
          stack[off + 1] = c0;
        }
        sp += np;
        const uint64_t m2 = ballot(kind == 2);
        const int nr = __popcll(m2);
        if (kind == 2) {
          const int r = (int)mask_rank(m2);
          rlo[r] = lo;
          rhi[r] = hi;
        }
        wave_sync();
        for (int r = 0; r < nr; ++r) {
          const int32_t rl = rlo[r], rh = rhi[r];
          const int64_t idx = (int64_t)rl + lane;
          bool take = false;
          double d2 = 0.0;
          if (idx <= rh && !(idx >= w0 && idx <= w1)) {
            d2 = dist2(qx, qy, qz, m.x[idx], m.y[idx], m.z[idx]);
            take = d2 <= theta;
          }
          c_cand += (unsigned long long)(rh - rl + 1);
          const uint64_t mt = ballot(take);
          if (mt == 0) continue;
          if (buf_n + __popcll(mt) > kBuf) {
            wave_sort(bd, bi, buf_n, lane);
            buf_n = k;
            theta = bd[k - 1];
            take = take && d2 <= theta;
          }
          const uint64_t mt2 = ballot(take);
          if (take) {
            const int slot = buf_n + (int)mask_rank(mt2);
            bd[slot] = d2;
            bi[slot] = (int32_t)idx;
          }
          buf_n += __popcll(mt2);
          wave_sync();
        }
      }
    }

    // ---- 3. final neighbour set in (distance, index) order
    wave_sort(bd, bi, buf_n, lane);
    const double r2 = bd[k - 1];
    const double rk = sqrt(r2);

    // ---- 4. energies into the waveform
    const double sdx = a.q.dx[qg], sdy = a.q.dy[qg], sdz = a.q.dz[qg];
    const float qE = a.q.energy[qg], qt = a.q.time[qg];
    double ksize;
    if (a.type == ALB_SCATTER) ksize = (4.0 / 3.0) * ALB_PI * pow(rk, 3.0);
    else ksize = ALB_PI * (rk * rk);
    const double knorm = 1.0 / ksize;
    const double m_bottom = (a.sc.floor_albedo / ALB_PI) * fmax(0.0, -sdy);
    const float fmult = (float)a.sc.mult;
    for (int t = (int)lane; t < k; t += 64) {
      const int32_t i = bi[t];
      double mul;
      if (a.type == ALB_BOTTOM_REFLECTION) {
        mul = m_bottom;
      } else if (a.type == ALB_SCATTER) {
        mul = scatter_mult(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, a.sc.phase_theta, a.sc.phase_cdf,
                           a.sc.phase_n);
      } else if (a.type == ALB_SURFACE_REFLECTION) {
        mul = (double)brdf_reflect(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, a.sc.alpha_ggx, a.sc.F0);
      } else {
        mul = btdf_transmit(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, a.sc.alpha_ggx, a.sc.n_water, a.sc.F0,
                            a.sc.surface_albedo);
      }
      const float ee = qE * m.energy[i];
      const double e = (double)ee * knorm * mul;
      const float tt = qt + m.time[i];
      const int64_t bin = (int64_t)(tt / fmult);
      if (bin >= a.win0 && bin < (int64_t)a.win0 + a.win_len) atomicAdd(&hist[bin - a.win0], e);
      else if (bin >= 0 && bin < a.n_bins) unsafeAtomicAdd(&a.wf[bin], e);
      if (a.nbr_out) a.nbr_out[qi * k + t] = i;
    }
    if (a.radius_out && lane == 0) a.radius_out[qi] = rk;
    wave_sync();
  }

  __syncthreads();
  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x)
    if (hist[t] != 0.0) unsafeAtomicAdd(&a.wf[a.win0 + t], hist[t]);
  if (a.ctr && lane == 0) {
    if (c_q) atomicAdd(&a.ctr[ALB_CTR_QUERIES], c_q);
    if (c_cand) atomicAdd(&a.ctr[ALB_CTR_CANDIDATES], c_cand);
    if (c_nodes) atomicAdd(&a.ctr[ALB_CTR_NODES], c_nodes);
    if (c_ovf) atomicAdd(&a.ctr[ALB_CTR_OVERFLOW], c_ovf);
  }
}

}  // namespace alb

extern "C" int alb_gather(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                          const alb_queries_t* q, int64_t q_begin, int64_t q_n, double* waveform,
                          int64_t n_bins, int32_t win0, int32_t win_len, int32_t* nbr_out, double* radius_out,
                          unsigned long long* counters, void* stream) {
  using namespace alb;
  if (!scene || !map || !q || !waveform) return set_error(ALB_ERR_ARG, "alb_gather: null argument");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES) return set_error(ALB_ERR_ARG, "alb_gather: bad photon type");
  if (q_n < 0 || q_begin < 0) return set_error(ALB_ERR_ARG, "alb_gather: negative query range");
  if (q_n == 0) return ALB_OK;
  if (map->n <= 0) return set_error(ALB_ERR_ARG, "alb_gather: empty photon map for type %d", photon_type);
  if (scene->k <= 0 || scene->k > map->n)
    return set_error(ALB_ERR_ARG, "alb_gather: k = %d exceeds the map size %lld", scene->k, (long long)map->n);
  if (scene->k > kMaxK) return set_error(ALB_ERR_UNSUPPORTED, "alb_gather: k = %d above %d", scene->k, kMaxK);
  if (win_len < 0 || win_len > kMaxWindow || win0 < 0 || (int64_t)win0 + win_len > n_bins)
    return set_error(ALB_ERR_ARG, "alb_gather: window [%d, %d) invalid (max length %d, %lld bins)", win0,
                     win0 + win_len, kMaxWindow, (long long)n_bins);
  if (!map->keys || !map->meta || !map->x || (map->n > 1 && (!map->node_box || !map->node_child)))
    return set_error(ALB_ERR_ARG, "alb_gather: map not built");
  GatherArgs a;
  a.sc = *scene;
  a.map = *map;
  a.q = *q;
  a.q_begin = q_begin;
  a.q_n = q_n;
  a.type = photon_type;
  a.wf = waveform;
  a.n_bins = n_bins;
  a.win0 = win0;
  a.win_len = win_len;
  a.nbr_out = nbr_out;
  a.radius_out = radius_out;
  a.ctr = counters;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_gather: no HIP device");
  int64_t blocks = (q_n + kWaves - 1) / kWaves;
  const int64_t cap = (int64_t)cus * 8;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)blocks), dim3(64 * kWaves), 0, (hipStream_t)stream, a);
  return check_launch("alb_gather");
}
