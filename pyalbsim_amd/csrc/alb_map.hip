// alb_map.hip — photon-map build: Morton-ordered records + binary radix tree.
//
// Replaces build_photon_map_data (src/alb_sim/photon_mapping/
// build_photon_map_data.py:8-62: concatenation into float32 SoA) and the
// scipy KDTree build of PhotonMapIndex (photon_map_index.py:17).
//
//  1. bounding box of the records (block reduction + ordered-int atomics)
//  2. quantise each record to a uniform grid of cubic cells (2^32 cells
//     along the largest extent of (x,z) for the planar surface / floor maps,
//     2^21 along the largest of (x,y,z) for scatter events) and interleave
//     the cell coordinates into a 64-bit Morton key
//  3. sort (key, record index) pairs (rocPRIM radix sort via hipCUB)
//  4. gather the float32 SoA into Morton order
//  5. build the binary radix tree over the sorted keys (Karras, "Maximizing
//     parallelism in the construction of BVHs, octrees, and k-d trees", HPG
//     2012): internal node i covers a contiguous record range; equal keys are
//     split by index
//  6. exact float32 bounding boxes of the children of the root and of every
//     node covering more than kBoxedRecords records: first every child of at
//     most kBoxedRecords records (or a single record) takes the box of its
//     records, all in one launch; then the other children, bottom-up: a
//     child's prefix length is always longer than its parent's, so one launch
//     per prefix length (longest first) finds every child box already written.
//     The gathers scan such small subtrees as record ranges and never read
//     their nodes' boxes.
#include <hipcub/hipcub.hpp>

#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

constexpr int kLevels = 128;  // prefix lengths 0..63 (+64..95 for equal keys)
constexpr int kNoBox = 255;   // level of a node whose boxes are not written


__global__ void bbox_init_kernel(uint32_t* bb, uint32_t* hist) {
  if (threadIdx.x < 3) bb[threadIdx.x] = 0xFFFFFFFFu;
  else if (threadIdx.x < 6) bb[threadIdx.x] = 0u;
  for (int i = threadIdx.x; i < kLevels; i += blockDim.x) hist[i] = 0u;
}

__global__ __launch_bounds__(256) void bbox_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                   const float* __restrict__ z, int64_t n, uint32_t* bb) {
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t a = f2ord(x[i]), b = f2ord(y[i]), c = f2ord(z[i]);
    mn[0] = min(mn[0], a); mx[0] = max(mx[0], a);
    mn[1] = min(mn[1], b); mx[1] = max(mx[1], b);
    mn[2] = min(mn[2], c); mx[2] = max(mx[2], c);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicMin(&bb[k], mn[k]);
      atomicMax(&bb[3 + k], mx[k]);
    }
  }
}

// meta = {origin x,y,z, inverse cell width x,y,z, dims, bits per axis}.
// The cells are cubes: one width on every keyed axis, set by the largest
// extent.  Per-axis widths would stretch the tree's boxes by the ratio of the
// extents (the scatter map of a 400 m shot spans ~240 m across and 5 m in
// depth, so its boxes were ~48 times flatter than wide), and a k-NN ball
// then meets many more of them.
__global__ void meta_kernel(const uint32_t* bb, double* meta, int dims) {
  const int bits = dims == 2 ? 32 : 21;
  const double cells = ldexp(1.0, bits);
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double lo = (double)ord2f(bb[k]);
    const double hi = (double)ord2f(bb[3 + k]);
    meta[k] = lo;
    if (dims == 3 || k != 1) ext = fmax(ext, hi - lo);  // planar maps key (x, z)
  }
  if (!(ext > 0.0)) ext = 1.0;
  for (int k = 0; k < 3; ++k) meta[3 + k] = cells / ext;
  meta[6] = (double)dims;
  meta[7] = (double)bits;
}

__global__ __launch_bounds__(256) void key_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, int64_t n,
                                                  const double* __restrict__ meta, uint64_t* __restrict__ keys,
                                                  int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = morton_key((double)x[i], (double)y[i], (double)z[i], meta);
  vals[i] = (int32_t)i;
}

// The sort permutation scatters records across the input: gathering the
// eight SoA arrays through it directly costs eight random cache lines per
// record.  The records are first packed into 32-byte rows (sequential reads
// and writes), so the gather touches one row per record.
__global__ __launch_bounds__(256) void pack_records_kernel(alb_records_t in, int64_t n, float4* __restrict__ rows) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  rows[2 * j] = make_float4(in.x[j], in.y[j], in.z[j], in.dx[j]);
  rows[2 * j + 1] = make_float4(in.dy[j], in.dz[j], in.energy[j], in.time[j]);
}

__global__ __launch_bounds__(256) void permute_kernel(const float4* __restrict__ rows, alb_map_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  const int64_t j = m.perm[i];
  const float4 a = rows[2 * j], b = rows[2 * j + 1];
  m.x[i] = a.x; m.y[i] = a.y; m.z[i] = a.z;
  m.dx[i] = a.w; m.dy[i] = b.x; m.dz[i] = b.y;
  m.energy[i] = b.z;
  m.time[i] = b.w;
}

// interaction points (heat maps) follow the records' permutation
__global__ __launch_bounds__(256) void permute_points_kernel(alb_records_t in, alb_map_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  const int32_t j = m.perm[i];
  m.fwi_x[i] = in.fwi_x[j]; m.fwi_y[i] = in.fwi_y[j]; m.fwi_z[i] = in.fwi_z[j];
  m.sfi_x[i] = in.sfi_x[j]; m.sfi_y[i] = in.sfi_y[j]; m.sfi_z[i] = in.sfi_z[j];
}

// common-prefix length of keys i and j (ki = k[i]); equal keys continue with
// the index bits
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, uint64_t ki, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t b = k[j];
  if (ki == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clzll(ki ^ b);
}

__device__ __forceinline__ void build_node(const uint64_t* __restrict__ k, int64_t n, int64_t i,
                                           Node* __restrict__ nodes, uint8_t* __restrict__ level, uint32_t* lh);

__global__ __launch_bounds__(256) void radix_tree_kernel(const uint64_t* __restrict__ k, int64_t n,
                                                         Node* __restrict__ nodes, uint8_t* __restrict__ level,
                                                         uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[kLevels];
  for (int t = threadIdx.x; t < kLevels; t += blockDim.x) lh[t] = 0u;
  __syncthreads();
  // grid-stride: a few thousand blocks, so the level histogram takes few
  // global atomics (a block per 256 nodes made millions on a few addresses)
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += (int64_t)gridDim.x * blockDim.x)
    build_node(k, n, i, nodes, level, lh);
  __syncthreads();
  for (int t = threadIdx.x; t < kLevels; t += blockDim.x)
    if (lh[t]) atomicAdd(&hist[t], lh[t]);
}

__device__ __forceinline__ void build_node(const uint64_t* __restrict__ k, int64_t n, int64_t i,
                                           Node* __restrict__ nodes, uint8_t* __restrict__ level, uint32_t* lh) {
  const uint64_t ki = k[i];
  const int d = (delta(k, n, i, ki, i + 1) - delta(k, n, i, ki, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(k, n, i, ki, i - d);
  int64_t lmax = 2;
  while (delta(k, n, i, ki, i + lmax * d) > dmin) lmax <<= 1;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(k, n, i, ki, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(k, n, i, ki, j);
  // split: binary search with steps ceil(l / 2^e), e = 1, 2, ... (a shift, not
  // a 64-bit division)
  int64_t s = 0;
  int e = 1;
  int64_t t = (l + 1) >> 1;
  while (true) {
    if (delta(k, n, i, ki, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
    ++e;
    t = (l + ((int64_t)1 << e) - 1) >> e;
  }
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  Node* nd = nodes + i;
  nd->first = (int32_t)lo;
  nd->last = (int32_t)hi;
  nd->split = (int32_t)g;
  nd->leaf_mask = (lo == g ? 1 : 0) | (hi == g + 1 ? 2 : 0);
  // only the root and nodes of more than kBoxedRecords records get boxes
  // (level kNoBox: skipped by the box passes)
  if (hi - lo + 1 > kBoxedRecords || i == 0) {
    const int lv = dnode < 0 ? 0 : (dnode >= kLevels ? kLevels - 1 : dnode);
    level[i] = (uint8_t)lv;
    atomicAdd(&lh[lv], 1u);
  } else {
    level[i] = (uint8_t)kNoBox;
  }
}

__global__ void level_scan_kernel(const uint32_t* hist, uint32_t* offset, uint32_t* cursor) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t acc = 0;
  for (int i = 0; i < kLevels; ++i) {
    offset[i] = acc;
    cursor[i] = acc;
    acc += hist[i];
  }
  offset[kLevels] = acc;
}

// Boxed nodes into level order.  Block b takes a contiguous chunk of nodes:
// it counts the chunk's nodes per level in LDS, reserves each level's run with
// one global atomic, then writes the chunk's node indices into the runs (in
// any order within a level: boxes do not depend on it).
__global__ __launch_bounds__(256) void level_scatter_kernel(const uint8_t* __restrict__ level, int64_t m,
                                                            uint32_t* cursor, int32_t* __restrict__ order) {
  __shared__ uint32_t cnt[kLevels], base[kLevels];
  for (int t = threadIdx.x; t < kLevels; t += blockDim.x) cnt[t] = 0u;
  __syncthreads();
  const int64_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < m ? lo + chunk : m;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int lv = (int)level[i];
    if (lv < kLevels) atomicAdd(&cnt[lv], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kLevels; t += blockDim.x) {
    if (cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
    cnt[t] = 0u;
  }
  __syncthreads();
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int lv = (int)level[i];
    if (lv < kLevels) order[base[lv] + atomicAdd(&cnt[lv], 1u)] = (int32_t)i;
  }
}

// box of a terminal child covering records [lo, hi]: the record itself for a
// leaf, else the records' own box (a subtree of at most kBoxedRecords
// records, whose node has no boxes)
__device__ __forceinline__ void terminal_box(const alb_map_t& m, int32_t lo, int32_t hi, float b[6]) {
  b[0] = b[3] = m.x[lo];
  b[1] = b[4] = m.y[lo];
  b[2] = b[5] = m.z[lo];
  // eight records per round, their loads issued together (a record past hi
  // repeats record lo, which changes no extreme)
  for (int32_t r0 = lo + 1; r0 <= hi; r0 += 8) {
    float x[8], y[8], z[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t r = r0 + u <= hi ? r0 + u : lo;
      x[u] = m.x[r];
      y[u] = m.y[r];
      z[u] = m.z[r];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      b[0] = fminf(b[0], x[u]);
      b[1] = fminf(b[1], y[u]);
      b[2] = fminf(b[2], z[u]);
      b[3] = fmaxf(b[3], x[u]);
      b[4] = fmaxf(b[4], y[u]);
      b[5] = fmaxf(b[5], z[u]);
    }
  }
}

// child side (0 left, 1 right) of a boxed node: its record range, and whether
// it is terminal (a single record or a subtree of at most kBoxedRecords)
__device__ __forceinline__ bool child_range(const Node& nd, int side, int32_t* lo, int32_t* hi) {
  *lo = side ? nd.split + 1 : nd.first;
  *hi = side ? nd.last : nd.split;
  return ((nd.leaf_mask >> side) & 1) != 0 || *hi - *lo + 1 <= kBoxedRecords;
}

// Terminal child boxes of every boxed node, in one launch: they depend on no
// other box, so no level order is needed, and the record scans (the bulk of
// the box work) run with the whole map's parallelism instead of one level's.
// Thread t takes side t & 1 of boxed node order[t >> 1].
__global__ __launch_bounds__(256) void box_terminal_kernel(alb_map_t m, const int32_t* __restrict__ order,
                                                           const uint32_t* __restrict__ offset) {
  Node* nodes = (Node*)m.nodes;
  const uint32_t items = 2u * offset[kLevels];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < items; t += gridDim.x * blockDim.x) {
    Node& nd = nodes[order[t >> 1]];
    const int side = (int)(t & 1u);
    int32_t lo, hi;
    if (!child_range(nd, side, &lo, &hi)) continue;
    float b[6];
    terminal_box(m, lo, hi, b);
    float* dst = side ? nd.rb : nd.lb;
#pragma unroll
    for (int c = 0; c < 6; ++c) dst[c] = b[c];
  }
}

// Boxed children of the boxed nodes of one prefix length: the union of the
// child node's two boxes, written by an earlier launch (a child's prefix is
// longer than its parent's, and the launches run longest first).
__global__ __launch_bounds__(256) void box_level_kernel(alb_map_t m, const int32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ offset, int lv) {
  Node* nodes = (Node*)m.nodes;
  const uint32_t beg = offset[lv], end = offset[lv + 1];
  for (uint32_t t = beg + blockIdx.x * blockDim.x + threadIdx.x; t < end; t += gridDim.x * blockDim.x) {
    Node& nd = nodes[order[t]];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      int32_t lo, hi;
      if (child_range(nd, side, &lo, &hi)) continue;
      const Node& cn = nodes[side ? nd.split + 1 : nd.split];
      float* dst = side ? nd.rb : nd.lb;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        dst[a] = fminf(cn.lb[a], cn.rb[a]);
        dst[3 + a] = fmaxf(cn.lb[3 + a], cn.rb[3 + a]);
      }
    }
  }
}

// Scene-major keys for maps of many scenes: the scene number in the top sb
// bits, the Morton key (truncated by sb bits) below, so one sort groups the
// records by scene and orders each scene along the Morton curve.
__global__ __launch_bounds__(256) void key_batched_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                          const float* __restrict__ z,
                                                          const int32_t* __restrict__ scene, int64_t n,
                                                          const double* __restrict__ meta, int key_bits, int sb,
                                                          uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mk = morton_key((double)x[i], (double)y[i], (double)z[i], meta);
  keys[i] = ((uint64_t)(uint32_t)scene[i] << (key_bits - sb)) | (mk >> sb);
  vals[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void scene_range_init_kernel(int32_t* first, int32_t* last, int32_t* root, int ns) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  first[s] = 0;
  last[s] = -1;
  root[s] = 0;
}

__global__ __launch_bounds__(256) void scene_range_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                          int32_t* first, int32_t* last) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = keys[i] >> shift;
  if (i == 0 || (keys[i - 1] >> shift) != s) first[s] = (int32_t)i;
  if (i == n - 1 || (keys[i + 1] >> shift) != s) last[s] = (int32_t)i;
}

// The records of one scene share the key prefix of the scene bits, so the
// radix tree holds a node covering exactly [first, last]; Karras nodes cover
// ranges that start or end at their own index, so it is node first or node last.
__global__ void scene_range_one_kernel(int32_t* first, int32_t* last, int64_t n) {
  first[0] = 0;
  last[0] = (int32_t)(n - 1);
}

__global__ __launch_bounds__(256) void scene_root_kernel(const Node* __restrict__ nodes, const int32_t* first,
                                                         const int32_t* last, int32_t* root, int ns) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int32_t f = first[s], l = last[s];
  if (l - f + 1 < 2) {
    root[s] = l == f ? ~f : 0;
    return;
  }
  const Node& a = nodes[f];
  root[s] = (a.first == f && a.last == l) ? f : l;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct WsLayout {
  size_t keys_in, vals_in, level, order, small, temp, rows, total, temp_bytes;
};

static int layout(int64_t n, WsLayout* L) {
  size_t temp_bytes = 0;
  const int nn = (int)(n > 1 ? n : 1);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (const uint64_t*)nullptr,
                                                    (uint64_t*)nullptr, (const int32_t*)nullptr,
                                                    (int32_t*)nullptr, nn, 0, 64, (hipStream_t)0);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_map: sort size query failed: %s", hipGetErrorString(e));
  size_t off = 0;
  L->keys_in = off; off = align256(off + 8 * (size_t)nn);
  L->vals_in = off; off = align256(off + 4 * (size_t)nn);
  L->level = off; off = align256(off + (size_t)nn);
  L->order = off; off = align256(off + 4 * (size_t)nn);
  L->small = off; off = align256(off + 4 * (3 * kLevels + 8) + 64);
  L->temp = off; off = align256(off + temp_bytes);
  L->rows = off; off = align256(off + 32 * (size_t)nn);  // packed records for the permutation
  L->total = off;
  L->temp_bytes = temp_bytes;
  return ALB_OK;
}

}  // namespace alb

extern "C" int alb_map_workspace_size(int64_t n, size_t* bytes) {
  using namespace alb;
  if (!bytes || n < 0) return set_error(ALB_ERR_ARG, "alb_map_workspace_size: bad arguments");
  if (n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_map_workspace_size: more than 2^31 records");
  WsLayout L;
  const int rc = layout(n, &L);
  if (rc != ALB_OK) return rc;
  *bytes = L.total;
  return ALB_OK;
}

static int map_build(const alb_records_t* in, int64_t n, int32_t dims, alb_map_t* map, alb_scene_index_t* idx,
                     void* workspace, size_t workspace_bytes, void* stream, const char* who) {
  using namespace alb;
  if (!in || !map) return set_error(ALB_ERR_ARG, "%s: null argument", who);
  if (dims != 2 && dims != 3) return set_error(ALB_ERR_ARG, "%s: dims must be 2 or 3", who);
  if (n < 0 || n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "%s: record count out of range", who);
  int sb = 0;
  if (idx) {
    if (idx->n_scenes < 1 || idx->n_scenes > (1 << 20)) return set_error(ALB_ERR_ARG, "%s: bad scene count", who);
    if (!idx->first || !idx->last || !idx->root) return set_error(ALB_ERR_ARG, "%s: scene index arrays missing", who);
    if (n > 0 && !in->scene) return set_error(ALB_ERR_ARG, "%s: records carry no scene array", who);
    sb = scene_bits(idx->n_scenes);
    idx->scene_bits = sb;
  }
  map->n = n;
  map->dims = dims;
  hipStream_t s = (hipStream_t)stream;
  const int B = 256;
  if (idx)
    hipLaunchKernelGGL(scene_range_init_kernel, dim3((idx->n_scenes + B - 1) / B), dim3(B), 0, s, idx->first,
                       idx->last, idx->root, idx->n_scenes);
  if (n == 0) return check_launch(who);
  if (!map->x || !map->y || !map->z || !map->dx || !map->dy || !map->dz || !map->energy || !map->time ||
      !map->perm || !map->keys || !map->meta)
    return set_error(ALB_ERR_ARG, "%s: map arrays missing", who);
  if (n > 1 && !map->nodes)
    return set_error(ALB_ERR_ARG, "%s: tree arrays missing", who);
  WsLayout L;
  int rc = layout(n, &L);
  if (rc != ALB_OK) return rc;
  if (!workspace || workspace_bytes < L.total)
    return set_error(ALB_ERR_WORKSPACE, "%s: workspace too small (%zu < %zu)", who, workspace_bytes, L.total);
  char* ws = (char*)workspace;
  uint64_t* keys_in = (uint64_t*)(ws + L.keys_in);
  int32_t* vals_in = (int32_t*)(ws + L.vals_in);
  uint8_t* level = (uint8_t*)(ws + L.level);
  int32_t* order = (int32_t*)(ws + L.order);
  uint32_t* small = (uint32_t*)(ws + L.small);
  uint32_t* bb = small;                // 6
  uint32_t* hist = small + 8;          // kLevels
  uint32_t* offset = hist + kLevels;   // kLevels + 1
  uint32_t* cursor = offset + kLevels + 1;
  const unsigned grid = (unsigned)((n + B - 1) / B);
  const int key_bits = dims == 2 ? 64 : 63;

  hipLaunchKernelGGL(bbox_init_kernel, dim3(1), dim3(128), 0, s, bb, hist);
  hipLaunchKernelGGL(bbox_kernel, dim3(grid < 2048 ? grid : 2048), dim3(B), 0, s, in->x, in->y, in->z, n, bb);
  hipLaunchKernelGGL(meta_kernel, dim3(1), dim3(1), 0, s, bb, map->meta, dims);
  if (sb > 0)
    hipLaunchKernelGGL(key_batched_kernel, dim3(grid), dim3(B), 0, s, in->x, in->y, in->z, in->scene, n, map->meta,
                       key_bits, sb, keys_in, vals_in);
  else
    hipLaunchKernelGGL(key_kernel, dim3(grid), dim3(B), 0, s, in->x, in->y, in->z, n, map->meta, keys_in, vals_in);
  if ((rc = check_launch(who)) != ALB_OK) return rc;
  size_t tb = L.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs((void*)(ws + L.temp), tb, keys_in, map->keys, vals_in,
                                                    map->perm, (int)n, 0, key_bits, s);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "%s: sort failed: %s", who, hipGetErrorString(e));
  alb_map_t m = *map;
  float4* rows = (float4*)(ws + L.rows);
  hipLaunchKernelGGL(pack_records_kernel, dim3(grid), dim3(B), 0, s, *in, n, rows);
  hipLaunchKernelGGL(permute_kernel, dim3(grid), dim3(B), 0, s, (const float4*)rows, m);
  if (in->fwi_x && m.fwi_x) hipLaunchKernelGGL(permute_points_kernel, dim3(grid), dim3(B), 0, s, *in, m);
  if (n > 1) {
    const unsigned g2 = (unsigned)((n - 1 + B - 1) / B);
    const unsigned gt = g2 < 2048 ? g2 : 2048;  // grid-stride launches
    hipLaunchKernelGGL(radix_tree_kernel, dim3(gt), dim3(B), 0, s, map->keys, n, (Node*)map->nodes, level, hist);
    hipLaunchKernelGGL(level_scan_kernel, dim3(1), dim3(1), 0, s, hist, offset, cursor);
    hipLaunchKernelGGL(level_scatter_kernel, dim3(gt), dim3(B), 0, s, level, n - 1, cursor, order);
    // boxed nodes cover more than kBoxedRecords records: fewer than n / 32
    const unsigned g3b = (unsigned)((n / 32 + B - 1) / B);
    const unsigned g3 = g3b < 1 ? 1 : (g3b < 1024 ? g3b : 1024);
    // terminal child boxes: two per boxed node
    const unsigned g4b = (unsigned)((n / 16 + B - 1) / B);
    const unsigned g4 = g4b < 1 ? 1 : (g4b < 4096 ? g4b : 4096);
    hipLaunchKernelGGL(box_terminal_kernel, dim3(g4), dim3(B), 0, s, m, order, offset);
    for (int lv = kLevels - 1; lv >= 0; --lv)
      hipLaunchKernelGGL(box_level_kernel, dim3(g3), dim3(B), 0, s, m, order, offset, lv);
  }
  if (idx) {
    if (sb > 0) {
      hipLaunchKernelGGL(scene_range_kernel, dim3(grid), dim3(B), 0, s, map->keys, n, key_bits - sb, idx->first,
                         idx->last);
    } else {
      hipLaunchKernelGGL(scene_range_one_kernel, dim3(1), dim3(1), 0, s, idx->first, idx->last, n);
    }
    hipLaunchKernelGGL(scene_root_kernel, dim3((idx->n_scenes + B - 1) / B), dim3(B), 0, s,
                       (const Node*)map->nodes, idx->first, idx->last, idx->root, idx->n_scenes);
  }
  return check_launch(who);
}

extern "C" int alb_map_build(const alb_records_t* in, int64_t n, int32_t dims, alb_map_t* map,
                             void* workspace, size_t workspace_bytes, void* stream) {
  return map_build(in, n, dims, map, nullptr, workspace, workspace_bytes, stream, "alb_map_build");
}

extern "C" int alb_map_build_batched(const alb_records_t* in, int64_t n, int32_t dims, alb_map_t* map,
                                     alb_scene_index_t* index, void* workspace, size_t workspace_bytes,
                                     void* stream) {
  if (!index) return alb::set_error(ALB_ERR_ARG, "alb_map_build_batched: null scene index");
  return map_build(in, n, dims, map, index, workspace, workspace_bytes, stream, "alb_map_build_batched");
}
