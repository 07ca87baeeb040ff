// alb_adjoint.hip — adjoint of the batched gather and the replay index.
//
// The forward estimator (core/simulation.py:132-195 restated by the gather
// kernels) adds c = E_q E_p K mult into bin(t_q + t_p).  With the paths held
// fixed (detached sampling), d log c = d log mult + S_p + S_q, S being the
// log-weight derivative of a path prefix (see the replay kernel in
// alb_trace.hip).  This file computes, for a loss gradient G over the
// waveform bins, the adjoint weight w = G[bin] c of every contribution and
// routes it three ways:
//   * to the map record p and the query q (input slots), where the replay
//     multiplies it with S_p / S_q;
//   * into per-scene sums of w d log mult for the parameters mult depends on
//     (phase mass: g; GGX BRDF: roughness; bottom factor: albedo).
// The kNN selection and the kernel radius are piecewise constant in the
// parameters and are treated as constants.
#include <hipcub/hipcub.hpp>

#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

struct AdjointArgs {
  const alb_scene_t* scenes;
  alb_map_t map;
  alb_queries_t q;
  int64_t c0, q_n;
  const int32_t* order;
  int32_t type, k;
  const int32_t* nbr;
  const double* knorm;
  const double* dloss;
  const int64_t* wf_off;
  double* a_rec;
  double* a_q;
  double* direct;  // [n_scenes][3]
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// one wave-reduced atomic per channel for the lanes' per-scene sums
__device__ __forceinline__ void flush_direct(double* d, int32_t sid, double* dst) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double t = wave_sum_f64(d[c]);
    if (lane_id() == 0 && sid >= 0 && t != 0.0) unsafeAtomicAdd(&dst[3 * (int64_t)sid + c], t);
    d[c] = 0.0;
  }
}

// Tiles of 256 (query, neighbour) items in slot order; each block takes a
// contiguous run of tiles (slots are sorted by scene, so a block meets few
// scenes).  Direct sums stay in registers while the tile scene is unchanged.
__global__ __launch_bounds__(256) void adjoint_contrib_kernel(AdjointArgs a) {
  const int k = a.k;
  const int64_t total = a.q_n * (int64_t)k;
  const int64_t tiles = (total + 255) / 256;
  const int64_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * per;
  const int64_t t1 = t0 + per < tiles ? t0 + per : tiles;
  const alb_map_t& m = a.map;
  int32_t cur = -1;
  double d[3] = {0.0, 0.0, 0.0};
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t it0 = tile * 256;
    const int32_t ts = a.q.scene[a.order[a.c0 + it0 / k]];
    if (ts != cur) {  // uniform across the block
      flush_direct(d, cur, a.direct);
      cur = ts;
    }
    const int64_t it = it0 + threadIdx.x;
    if (it >= total) continue;
    const int64_t qi = it / k;
    const int64_t qg = a.order[a.c0 + qi];
    const int32_t sc = a.q.scene[qg];
    const alb_scene_t* S = a.scenes + sc;
    const int32_t i = a.nbr[it];
    if (i < 0) continue;
    const float tt = a.q.time[qg] + m.time[i];
    const int64_t bin = (int64_t)(tt / (float)S->mult);
    const int64_t nb = 2 * (int64_t)S->steps;
    if (bin < 0 || bin >= nb) continue;
    const double G = a.dloss[a.wf_off[sc] + (int64_t)a.type * nb + bin];
    if (G == 0.0) continue;
    const double sdx = a.q.dx[qg], sdy = a.q.dy[qg], sdz = a.q.dz[qg];
    double mul, dl = 0.0;
    int ch = -1;
    if (a.type == ALB_BOTTOM_REFLECTION) {
      mul = bottom_mult(*S, sdx, sdy, sdz);
      ch = 2;
      dl = 1.0;  // the host divides by the albedo
    } else if (a.type == ALB_SCATTER) {
      // the phase masses of the layer at the query vertex, as the gather takes them
      int pn;
      const double* pm = gather_phase_mass(*S, a.q.y[qg], &pn);
      const int b = scatter_bin(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, pn);
      mul = pm[b];
      if (S->phase_dlog_mass) {  // HG scenes: one table for the column
        ch = 0;
        dl = S->phase_dlog_mass[b];
      }
    } else if (a.type == ALB_SURFACE_REFLECTION) {
      double dla2;
      mul = (double)brdf_reflect_dlog(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, S->alpha_ggx, S->F0, &dla2);
      ch = 1;
      dl = dla2;
    } else {
      mul = btdf_transmit(m.dx[i], m.dy[i], m.dz[i], sdx, sdy, sdz, S->alpha_ggx, S->n_water, S->F0,
                          S->surface_albedo);
    }
    const float ee = a.q.energy[qg] * m.energy[i];
    const double c = (double)ee * a.knorm[qi] * mul;
    const double w = G * c;
    if (w == 0.0) continue;
    unsafeAtomicAdd(&a.a_rec[m.perm[i]], w);
    unsafeAtomicAdd(&a.a_q[qg], w);
    if (ch >= 0) {
      if (sc == cur) {
        d[ch] += w * dl;
      } else {
        unsafeAtomicAdd(&a.direct[3 * (int64_t)sc + ch], w * dl);
      }
    }
  }
  flush_direct(d, cur, a.direct);
}

__global__ __launch_bounds__(256) void replay_key_kernel(const int64_t* __restrict__ tags,
                                                         const int32_t* __restrict__ scenes, int64_t n,
                                                         const int64_t* __restrict__ offsets, uint64_t* keys,
                                                         int32_t* slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t tag = (uint64_t)tags[i];
  const int64_t base = scenes ? offsets[scenes[i]] : 0;
  keys[i] = ((uint64_t)(base + (int64_t)(tag >> 16)) << 16) | (tag & 0xFFFFu);
  slots[i] = (int32_t)i;
}

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct IdxWs {
  size_t keys, slots, temp, temp_bytes, total;
};

static int idx_layout(int64_t n, IdxWs* L) {
  size_t tb = 0;
  const int nn = (int)(n > 1 ? n : 1);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, nn, 0, 64,
                                                    (hipStream_t)0);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_replay_index: sort size query failed: %s", hipGetErrorString(e));
  size_t off = 0;
  L->keys = off; off = al256(off + 8 * (size_t)nn);
  L->slots = off; off = al256(off + 4 * (size_t)nn);
  L->temp = off; off = al256(off + tb);
  L->temp_bytes = tb;
  L->total = off;
  return ALB_OK;
}

}  // namespace alb

extern "C" int alb_adjoint_contrib(const alb_scene_t* scenes, int32_t n_scenes, int32_t k, const alb_map_t* map,
                                   int32_t photon_type, const alb_queries_t* q, const int32_t* order, int64_t c0,
                                   int64_t cn, const int32_t* nbr, const double* knorm, const double* dloss,
                                   const int64_t* wf_offset, double* a_record, double* a_query, double* direct,
                                   void* stream) {
  using namespace alb;
  if (!scenes || !map || !q || !order || !nbr || !knorm || !dloss || !wf_offset || !a_record || !a_query || !direct)
    return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: null argument");
  if (n_scenes < 1) return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: bad scene count");
  if (!q->scene) return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: queries carry no scene array");
  if (!map->perm) return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: map not built");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES)
    return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: bad photon type");
  if (k <= 0) return set_error(ALB_ERR_ARG, "alb_adjoint_contrib: bad k");
  if (cn <= 0) return ALB_OK;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_adjoint_contrib: no HIP device");
  AdjointArgs aa{scenes, *map, *q, c0, cn, order, photon_type, k, nbr, knorm, dloss, wf_offset, a_record, a_query,
                 direct};
  const int64_t tiles = (cn * k + 255) / 256;
  int64_t grid = (int64_t)cus * 8;
  if (grid > tiles) grid = tiles;
  hipLaunchKernelGGL(adjoint_contrib_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, aa);
  return check_launch("alb_adjoint_contrib");
}

extern "C" int alb_replay_index_workspace_size(int64_t n, size_t* bytes) {
  using namespace alb;
  if (!bytes || n < 0 || n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_replay_index_workspace_size: bad arguments");
  IdxWs L;
  const int rc = idx_layout(n, &L);
  if (rc != ALB_OK) return rc;
  *bytes = L.total;
  return ALB_OK;
}

extern "C" int alb_replay_index_build(const int64_t* tags, const int32_t* scenes, int64_t n,
                                      const int64_t* photon_offsets, uint64_t* keys_out, int32_t* slots_out,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  using namespace alb;
  if (n < 0 || n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_replay_index_build: bad record count");
  if (n == 0) return ALB_OK;
  if (!tags || !keys_out || !slots_out || (scenes && !photon_offsets))
    return set_error(ALB_ERR_ARG, "alb_replay_index_build: null argument");
  IdxWs L;
  int rc = idx_layout(n, &L);
  if (rc != ALB_OK) return rc;
  if (!workspace || workspace_bytes < L.total)
    return set_error(ALB_ERR_WORKSPACE, "alb_replay_index_build: workspace too small (%zu < %zu)", workspace_bytes,
                     L.total);
  char* ws = (char*)workspace;
  uint64_t* keys = (uint64_t*)(ws + L.keys);
  int32_t* slots = (int32_t*)(ws + L.slots);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(replay_key_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tags, scenes, n,
                     photon_offsets, keys, slots);
  size_t tb = L.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs((void*)(ws + L.temp), tb, keys, keys_out, slots, slots_out, (int)n,
                                                    0, 64, s);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_replay_index_build: sort failed: %s", hipGetErrorString(e));
  return check_launch("alb_replay_index_build");
}
