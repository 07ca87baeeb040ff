// alb_radius.hip — fixed-radius photon gather into waveforms.
//
// The reference gathers the k nearest records of every sensor-path vertex
// (Simulation.sample_photons, src/alb_sim/core/simulation.py:132-195).  This
// is the fixed-radius form of the same estimate: every map record within
// distance r of a query contributes, with the kernel 1/(pi r^2) (planar maps)
// or 1/((4/3) pi r^3) (scatter map) from the given r and the reference's
// type multipliers, energy product and time bins.  It serves footprint-grid
// queries such as a 256 x 256 grid on the seafloor plane.
//
// radius_kernel — one wavefront per query over the map's radix tree:
//   * up to 64 stacked nodes are expanded per step (one per lane); a child
//     whose box is farther than r is dropped, a child whose box lies inside
//     the ball (farthest corner within r) becomes a record range taken whole,
//     a leaf or a subtree of <= kBucket records becomes a range to test, and
//     any other child is pushed;
//   * the step's ranges are concatenated and dealt to the lanes, four records
//     per lane at a time, so every lane works whatever the range sizes; a
//     tested record is taken when its FP64 squared distance (dist2, the
//     k-NN's own) is <= r^2, which is exactly the brute-force set: a box's
//     nearest and farthest corner distances bound those of its records under
//     monotone rounding;
//   * taken records contribute at once: each lane keeps a running FP64 sum
//     of its current time bin (a query's records fall in a few bins) and adds
//     it into an LDS window of the waveform (ds_add_f64), or to global
//     atomics outside the window, when the bin changes.
// Queries are visited in the caller's order (the Morton order of
// alb_gather_order keeps neighbouring waves on the same nodes and records).
#include <limits.h>

#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

constexpr int kRWaves = 4;          // wavefronts per block
constexpr int kRStack = 1024;       // per-wave node stack (LDS)
constexpr int kRRanges = 128;       // ranges per step: two children per lane
constexpr int kRBucket = 64;        // subtrees of <= kRBucket records are scanned
static_assert(kRBucket >= kBoxedRecords, "the radius gather reads the boxes of nodes of more than kRBucket records");
constexpr int kRMaxWindow = 2048;   // LDS-privatised waveform bins
constexpr int kRXcds = 8;           // MI355X XCDs; workgroups are dealt round-robin

struct RadiusArgs {
  alb_scene_t sc;
  alb_map_t map;
  alb_queries_t q;
  int64_t q_begin, q_n;
  const int32_t* order;  // slot -> query offset, or null for identity
  int32_t type;
  double r2;
  double knorm;
  double* wf;
  int64_t n_bins;
  int32_t win0, win_len;
  int64_t* count;        // [q_n] records taken per query, or null
  int64_t* isum;         // [q_n] sum of their map indices, or null
  unsigned long long* ctr;
};

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)((uint64_t)v >> 32), off, 64);
    v += (int64_t)(((uint64_t)hi << 32) | lo);
  }
  return v;
}

// add v into bin b: the LDS window when b is inside it, else the global
// waveform (bins outside [0, n_bins) are dropped, as np.add.at's index range)
__device__ __forceinline__ void add_bin(const RadiusArgs& a, double* hist, int64_t b, double v) {
  if (b >= a.win0 && b < (int64_t)a.win0 + a.win_len) atomicAdd(&hist[b - a.win0], v);
  else if (b >= 0 && b < a.n_bins) unsafeAtomicAdd(&a.wf[b], v);
}

// TYPE: the photon type (a template parameter, so each map's kernel carries
// only its own multiplier code and registers)
template <int TYPE>
__global__ __launch_bounds__(256) void radius_kernel(RadiusArgs a) {
  extern __shared__ double hist[];  // [win_len]
  __shared__ int32_t sn_s[kRWaves][kRStack];
  __shared__ int32_t rlo_s[kRWaves][kRRanges];
  __shared__ int32_t rpre_s[kRWaves][kRRanges + 1];
  __shared__ uint8_t rin_s[kRWaves][kRRanges];

  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x) hist[t] = 0.0;
  __syncthreads();

  const uint32_t lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar slot loop
  int32_t* sn = sn_s[w];
  int32_t* rlo = rlo_s[w];
  int32_t* rpre = rpre_s[w];
  uint8_t* rin = rin_s[w];
  const alb_map_t& m = a.map;
  const Node* __restrict__ nodes = (const Node*)m.nodes;
  const float* __restrict__ mx = m.x;
  const float* __restrict__ my = m.y;
  const float* __restrict__ mz = m.z;
  const int32_t n = (int32_t)m.n;
  const double r2 = a.r2;
  const float fmult = (float)a.sc.mult;
  uint32_t c_q = 0, c_nodes = 0, c_ovf = 0;
  uint64_t c_cand = 0, c_take = 0;

  // XCD-aware slot ranges, as in the k-NN: XCD x walks its own contiguous
  // 1/kRXcds of the slots, so the records its queries visit stay in its L2
  const int xcd = (int)(blockIdx.x % kRXcds);
  const int64_t xblocks = gridDim.x / kRXcds;
  const int64_t s_lo = a.q_n * xcd / kRXcds, s_hi = a.q_n * (xcd + 1) / kRXcds;
  for (int64_t si = s_lo + (int64_t)(blockIdx.x / kRXcds) * kRWaves + w; si < s_hi; si += xblocks * kRWaves) {
    const int64_t qrel = a.order ? (int64_t)a.order[si] : si;
    const int64_t qg = a.q_begin + qrel;
    const double qx = a.q.x[qg], qy = a.q.y[qg], qz = a.q.z[qg];
    const double sdx = a.q.dx[qg], sdy = a.q.dy[qg], sdz = a.q.dz[qg];
    const float qe = a.q.energy[qg], qt = a.q.time[qg];
    // multiplier parts fixed by the query
    const double m_bottom = bottom_mult(a.sc, sdx, sdy, sdz);
    int pn = 0;
    const double* pm = TYPE == ALB_SCATTER ? gather_phase_mass(a.sc, qy, &pn) : nullptr;
    c_q++;
    int64_t cnt = 0, isum = 0;
    int64_t bcur = -1;  // lane's current bin (-1: none) and its running sum
    double acc = 0.0;

    // the root: node 0, or the single record of a one-record map
    int sp = 0;
    int pending = 0;  // ranges waiting from the root set-up
    if (n == 1) {
      if (lane == 0) {
        rlo[0] = 0;
        rpre[0] = 0;
        rpre[1] = 1;
        rin[0] = 0;
      }
      pending = 1;
    } else {
      if (lane == 0) sn[0] = 0;
      sp = 1;
    }
    wave_sync();
    while (sp > 0 || pending > 0) {
      int nr = pending;
      pending = 0;
      if (nr == 0) {
        // pop policy: at most 64 nodes while the stack is below kRStack - 256
        // (then a step pushes at most 128), else one (then the stack grows by
        // at most one per level of the tree's depth, < 128)
        const int p = (sp > kRStack - 256) ? 1 : (sp < 64 ? sp : 64);
        const bool have = (int)lane < p;
        const int32_t node = have ? sn[sp - 1 - (int)lane] : 0;
        sp -= p;
        wave_sync();
        // per child (0 = left, 1 = right): 1 = push, 2 = range to test, 3 = range inside the ball
        int kind0 = 0, kind1 = 0;
        int32_t lo0 = 0, hi0 = -1, lo1 = 0, hi1 = -1, split = 0;
        if (have) {
          const Node nd = nodes[node];
          split = nd.split;
          const double dl = box_dist2(qx, qy, qz, nd.lb);
          const double dr = box_dist2(qx, qy, qz, nd.rb);
          lo0 = nd.first;
          hi0 = nd.split;
          lo1 = nd.split + 1;
          hi1 = nd.last;
          if (dl <= r2) {
            if ((nd.leaf_mask & 1) || hi0 - lo0 + 1 <= kRBucket) kind0 = 2;
            else kind0 = box_far2(qx, qy, qz, nd.lb) <= r2 ? 3 : 1;
          }
          if (dr <= r2) {
            if ((nd.leaf_mask & 2) || hi1 - lo1 + 1 <= kRBucket) kind1 = 2;
            else kind1 = box_far2(qx, qy, qz, nd.rb) <= r2 ? 3 : 1;
          }
          if (kind0 == 2 && box_far2(qx, qy, qz, nd.lb) <= r2) kind0 = 3;
          if (kind1 == 2 && box_far2(qx, qy, qz, nd.rb) <= r2) kind1 = 3;
        }
        c_nodes += (uint32_t)__popcll(ballot(have));
        const uint64_t pa = ballot(kind0 == 1), pb = ballot(kind1 == 1);
        const int npush = __popcll(pa) + __popcll(pb);
        if (sp + npush > kRStack) {  // excluded by the pop policy; refuse rather than corrupt
          c_ovf++;
          break;
        }
        {
          int off = sp + (int)mask_rank(pa) + (int)mask_rank(pb);
          if (kind0 == 1) sn[off++] = split;      // internal left child
          if (kind1 == 1) sn[off] = split + 1;    // internal right child
        }
        sp += npush;
        // ranges of this step, in lane order, with their sizes' prefix sums
        const uint64_t ra = ballot(kind0 >= 2), rb = ballot(kind1 >= 2);
        nr = __popcll(ra) + __popcll(rb);
        const int32_t sz = (kind0 >= 2 ? hi0 - lo0 + 1 : 0) + (kind1 >= 2 ? hi1 - lo1 + 1 : 0);
        int32_t incl = sz;
        for (int off = 1; off < 64; off <<= 1) {
          const int32_t v = __shfl_up(incl, off, 64);
          if ((int)lane >= off) incl += v;
        }
        const int32_t excl = incl - sz;
        {
          int off = (int)mask_rank(ra) + (int)mask_rank(rb);
          int32_t pre = excl;
          if (kind0 >= 2) {
            rlo[off] = lo0;
            rpre[off] = pre;
            rin[off] = kind0 == 3;
            pre += hi0 - lo0 + 1;
            ++off;
          }
          if (kind1 >= 2) {
            rlo[off] = lo1;
            rpre[off] = pre;
            rin[off] = kind1 == 3;
          }
          if (lane == 63) rpre[nr] = incl;
        }
        wave_sync();
      }
      if (nr == 0) continue;
      const int32_t total = rpre[nr];
      // deal the concatenated ranges to the lanes, four records per lane per
      // pass.  A lane's items increase, so it follows its current range with
      // a cursor (cur_*) and reads the range tables only when it crosses into
      // the next range; the coordinate loads of all four items issue before
      // one wait.
      int rc = 0;
      int32_t cur_lo = rlo[0], cur_pre = rpre[0];
      bool cur_in = rin[0] != 0;
      int32_t nxt = nr > 1 ? rpre[1] : INT_MAX;
      for (int32_t base = 0; base < total; base += 4 * 64) {
        int32_t ix[4];
        bool tk[4], in[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t t = base + 64 * u + (int32_t)lane;
          while (t >= nxt) {  // nxt is INT_MAX in the last range
            ++rc;
            cur_lo = rlo[rc];
            cur_pre = rpre[rc];
            cur_in = rin[rc] != 0;
            nxt = rc + 1 < nr ? rpre[rc + 1] : INT_MAX;
          }
          tk[u] = t < total;
          ix[u] = tk[u] ? cur_lo + (t - cur_pre) : cur_lo;
          in[u] = cur_in;
        }
        float xs[4], ys[4], zs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t li = in[u] ? cur_lo : ix[u];  // ranges inside the ball need no coordinates
          xs[u] = in[u] ? 0.0f : mx[li];
          ys[u] = in[u] ? 0.0f : my[li];
          zs[u] = in[u] ? 0.0f : mz[li];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (tk[u] && !in[u]) {
            c_cand++;
            tk[u] = dist2(qx, qy, qz, xs[u], ys[u], zs[u]) <= r2;
          }
        }
        // payload loads of the four items, issued together (every index is a
        // valid record; untaken items are not used)
        float es[4], ts[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          es[u] = m.energy[ix[u]];
          ts[u] = m.time[ix[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (!tk[u]) continue;
          const int32_t i = ix[u];
          double mul;
          if constexpr (TYPE == ALB_BOTTOM_REFLECTION) {
            mul = m_bottom;
          } else if constexpr (TYPE == ALB_SCATTER) {
            mul = scatter_mass(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, pm, pn);
          } else if constexpr (TYPE == ALB_SURFACE_REFLECTION) {
            mul = (double)brdf_reflect(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, a.sc.alpha_ggx, a.sc.F0);
          } else {
            mul = btdf_transmit(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, a.sc.alpha_ggx, a.sc.n_water, a.sc.F0,
                                a.sc.surface_albedo);
          }
          const float ee = qe * es[u];
          const double e = (double)ee * a.knorm * mul;
          const float tt = qt + ts[u];
          const int64_t bin = (int64_t)(tt / fmult);
          // a query's records fall in a few bins: keep a per-lane running sum
          // of the current bin and add it out when the bin changes
          if (bin != bcur) {
            add_bin(a, hist, bcur, acc);
            bcur = bin;
            acc = e;
          } else {
            acc += e;
          }
          cnt++;
          isum += i;
        }
      }
      wave_sync();
    }
    add_bin(a, hist, bcur, acc);
    c_take += (uint64_t)cnt;
    cnt = wave_sum_i64(cnt);
    isum = wave_sum_i64(isum);
    if (lane == 0) {
      if (a.count) a.count[qrel] = cnt;
      if (a.isum) a.isum[qrel] = isum;
    }
    wave_sync();
  }
  __syncthreads();
  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x)
    if (hist[t] != 0.0) unsafeAtomicAdd(&a.wf[a.win0 + t], hist[t]);
  c_cand = (uint64_t)wave_sum_i64((int64_t)c_cand);
  c_take = (uint64_t)wave_sum_i64((int64_t)c_take);
  if (a.ctr && lane == 0) {
    if (c_q) atomicAdd(&a.ctr[ALB_CTR_QUERIES], (unsigned long long)c_q);
    if (c_cand) atomicAdd(&a.ctr[ALB_CTR_CANDIDATES], (unsigned long long)c_cand);
    if (c_nodes) atomicAdd(&a.ctr[ALB_CTR_NODES], (unsigned long long)c_nodes);
    if (c_ovf) atomicAdd(&a.ctr[ALB_CTR_OVERFLOW], (unsigned long long)c_ovf);
    if (c_take) atomicAdd(&a.ctr[ALB_CTR_NEIGHBOURS], (unsigned long long)c_take);
  }
}

}  // namespace alb

extern "C" int alb_radius_gather(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                                 const alb_queries_t* q, int64_t q_begin, int64_t q_n, const int32_t* order,
                                 double radius, double* waveform, int64_t n_bins, int32_t win0, int32_t win_len,
                                 int64_t* count_out, int64_t* index_sum_out, unsigned long long* counters,
                                 void* stream) {
  using namespace alb;
  if (!scene || !map || !q || !waveform) return set_error(ALB_ERR_ARG, "alb_radius_gather: null argument");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES)
    return set_error(ALB_ERR_ARG, "alb_radius_gather: bad photon type");
  if (q_n < 0 || q_begin < 0 || q_n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_radius_gather: bad query range");
  if (!(radius > 0.0) || !(radius < 1e30)) return set_error(ALB_ERR_ARG, "alb_radius_gather: radius must be > 0");
  if (win_len < 0 || win_len > kRMaxWindow || win0 < 0 || (int64_t)win0 + win_len > n_bins)
    return set_error(ALB_ERR_ARG, "alb_radius_gather: window [%d, %d) invalid (max length %d, %lld bins)", win0,
                     win0 + win_len, kRMaxWindow, (long long)n_bins);
  if (photon_type == ALB_SCATTER && !scene->phase_mass)
    return set_error(ALB_ERR_ARG, "alb_radius_gather: scene without phase_mass table");
  if (map->n < 0 || map->n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_radius_gather: bad map size");
  if (q_n == 0) return ALB_OK;
  if (map->n == 0) return set_error(ALB_ERR_ARG, "alb_radius_gather: empty photon map");
  if (!map->x || (map->n > 1 && !map->nodes)) return set_error(ALB_ERR_ARG, "alb_radius_gather: map not built");
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_radius_gather: no HIP device");
  RadiusArgs ra;
  ra.sc = *scene;
  ra.map = *map;
  ra.q = *q;
  ra.q_begin = q_begin;
  ra.q_n = q_n;
  ra.order = order;
  ra.type = photon_type;
  ra.r2 = radius * radius;
  // kernel normalisation from the given radius (core/simulation.py:163-187 with r in place of r_k)
  const double ksize = photon_type == ALB_SCATTER ? (4.0 / 3.0) * ALB_PI * pow(radius, 3.0) : ALB_PI * (radius * radius);
  ra.knorm = 1.0 / ksize;
  ra.wf = waveform;
  ra.n_bins = n_bins;
  ra.win0 = win0;
  ra.win_len = win_len;
  ra.count = count_out;
  ra.isum = index_sum_out;
  ra.ctr = counters;
  int64_t blocks = (q_n + kRWaves - 1) / kRWaves;
  if (blocks > (int64_t)cus * 8) blocks = (int64_t)cus * 8;
  blocks = (blocks + kRXcds - 1) / kRXcds * kRXcds;
  auto kern = photon_type == ALB_BOTTOM_REFLECTION ? radius_kernel<ALB_BOTTOM_REFLECTION>
              : photon_type == ALB_SCATTER         ? radius_kernel<ALB_SCATTER>
              : photon_type == ALB_SURFACE_REFLECTION ? radius_kernel<ALB_SURFACE_REFLECTION>
                                                      : radius_kernel<ALB_SURFACE_TRANSMISSION_UP>;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * kRWaves), sizeof(double) * (size_t)win_len,
                     (hipStream_t)stream, ra);
  return check_launch("alb_radius_gather");
}
