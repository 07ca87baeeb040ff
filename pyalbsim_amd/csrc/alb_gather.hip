// alb_gather.hip — exact k-nearest-neighbour photon gather into waveforms.
//
// Replaces Simulation.sample_photons (src/alb_sim/core/simulation.py:132-195):
// for every sensor-path vertex, the k nearest map records (KDTree.query), the
// type-specific energy multiplier, the area / volume kernel 1/(pi r_k^2) or
// 1/(4/3 pi r_k^3) from the k-th distance, and np.add.at into the waveform.
//
// Two kernels per chunk of queries:
//  knn_kernel — one wavefront per query:
//   1. seed: descend the radix tree into the child whose box is nearer while
//      it still holds >= k records, stacking every child not taken; the (at
//      most kSeed) records reached give a first bound theta = k-th smallest
//      squared distance;
//   2. traversal of the stacked subtrees, up to 64 nodes per step (one per
//      lane): each node carries its children's exact float32 boxes, so a child
//      is dropped, pushed (with its box distance, re-checked when popped) or,
//      at <= 64 records, scanned; scanned records with d2 <= theta go to an
//      LDS candidate buffer, four record ranges at a time so that their loads
//      are in flight together;
//   3. a full buffer (and the seed) is cut with a 64-bucket histogram of
//      [0, theta]: entries up to the bucket holding the k-th smallest stay,
//      and theta drops to the largest of them (an upper bound of the k-th);
//      when that would not free 64 slots, the cut is exact: a wave-wide
//      bitwise radix select of the FP64 bit patterns, ties by index.  The
//      final k are taken by the exact select.  The result is the brute-force
//      neighbour set bit for bit, because squared distances are evaluated in
//      FP64 in one fixed order, box distances are, by monotone rounding,
//      never larger than those of the records inside, and every cut keeps all
//      entries at or below its bound.
//  Queries are first sorted by the Morton key of their position on the map's
//  grid (on a grid over their own box where they spread more than four times
//  as wide as the records), so neighbouring wavefronts walk the same nodes and
//  buckets (L2 hits), also where the queries lie far outside the records.
//  contrib_kernel — one thread per (query, neighbour): energy multiplier,
//   kernel normalisation and time bin; FP64 sums go to an LDS-privatised
//   window of the waveform (ds_add_f64), bins outside it to global atomics.
//   Keeping the FP64 surface / phase-function code out of knn_kernel keeps the
//   traversal at low register pressure (higher occupancy).
#include <limits.h>

#include <mutex>

#include <hipcub/hipcub.hpp>

#include "alb_device.h"
#include "alb_internal.h"

namespace alb {

constexpr int kWaves = 4;         // wavefronts per k-NN block
constexpr int kStack = 256;       // per-wave node stack (LDS)
constexpr int kSeed = 192;        // records of the descent seed (<= kBuf)
constexpr int kBuf = 256;         // per-wave candidate buffer (LDS), power of two
// k above kMaxK runs a variant with a buffer of kBufLarge entries (and a seed
// of kBufLarge - 64), k above kMaxKLarge one of kBufHuge entries (two waves
// per block for its LDS); a buffer cut must free 64 slots, so k <= buffer - 64
constexpr int kBufLarge = 512;
constexpr int kMaxKLarge = kBufLarge - 64;
constexpr int kBufHuge = 1024;
constexpr int kMaxKHuge = kBufHuge - 64;  // the largest k of the gathers
constexpr int kBucket = 64;       // nodes with <= kBucket records are scanned
static_assert(kBucket >= kBoxedRecords, "the k-NN reads the boxes of nodes of more than kBucket records");
constexpr int kPath = 128;        // descent levels kept for the next query (LDS)
constexpr int kMaxWindow = 7680;  // LDS-privatised waveform bins (60 KB of FP64: with the row tile, 64 KB per block)
constexpr int kMaxWindowLarge = 16384;  // batched contributions with one 1024-thread block per CU (132 KB)
constexpr int kMaxK = kSeed;      // the seed must hold k records
constexpr int64_t kChunk = 1 << 22;  // queries per k-NN / contribution round
constexpr int kWorkStride = 32;      // words between the per-XCD query counters (256 B)
constexpr int kGrab = 4;             // queries a wavefront takes per counter increment
// slots of one k-NN launch: the kernel's slot arithmetic is 32-bit, and the
// per-XCD counters run past the end by at most kGrab per wavefront
constexpr int64_t kMaxSlots = 1ll << 30;
static_assert(8 * kWorkStride <= ALB_KNN_WORK_WORDS, "work array too small");
constexpr int kXcds = 8;          // MI355X XCDs; workgroups are dealt to them round-robin

// Query slots are processed in sorted order: slot si (c0 <= si < c0 + q_n)
// is query q_begin + order[si].  Output rows are slot - c0, or the query's own
// index when direct (caller-provided neighbour arrays).
struct KnnArgs {
  // map (sorted positions, tree) and queries: only the arrays the k-NN reads,
  // so the kernel's scalar registers hold few pointers
  const float* x;
  const float* y;
  const float* z;
  const Node* nodes;
  int64_t n;
  const double* qx;
  const double* qy;
  const double* qz;
  const int32_t* qscene;
  int64_t q_begin, c0, q_n;
  const int32_t* order;
  int32_t direct;
  int32_t k;
  int32_t* nbr;      // [rows * k]
  double* radius;    // [rows]
  unsigned long long* ctr;
  unsigned long long* work;  // query-dealing counters, one per XCD on its own 256-byte line
  // many scenes: each query searches its own scene's subtree (null for one scene)
  const int32_t* sfirst;
  const int32_t* slast;
  const int32_t* sroot;
};

static KnnArgs knn_args(const alb_map_t& m, const alb_queries_t& q, int64_t q_begin, int64_t c0, int64_t q_n,
                        const int32_t* order, int direct, int k, int32_t* nbr, double* radius,
                        unsigned long long* ctr, const alb_scene_index_t* idx, unsigned long long* work) {
  KnnArgs a;
  a.x = m.x; a.y = m.y; a.z = m.z;
  a.nodes = (const Node*)m.nodes;
  a.n = m.n;
  a.qx = q.x; a.qy = q.y; a.qz = q.z;
  a.qscene = q.scene;
  a.q_begin = q_begin; a.c0 = c0; a.q_n = q_n;
  a.order = order;
  a.direct = direct;
  a.k = k;
  a.nbr = nbr;
  a.radius = radius;
  a.ctr = ctr;
  a.work = work;
  a.sfirst = idx ? idx->first : nullptr;
  a.slast = idx ? idx->last : nullptr;
  a.sroot = idx ? idx->root : nullptr;
  return a;
}

// map record coordinate i: through a 32-bit byte offset from the array base
// when the map holds fewer than 2^30 records (one shift per index, shared by
// the three arrays), else through 64-bit addresses
template <bool OFF32>
__device__ __forceinline__ float rec_ld(const float* __restrict__ p, int32_t i) {
  if constexpr (OFF32) {
    return *(const float*)((const char*)p + ((uint32_t)i << 2));
  } else {
    return p[(uint32_t)i];
  }
}

struct ContribArgs {
  alb_scene_t sc;
  alb_map_t map;
  alb_queries_t q;
  int64_t q_begin, c0, q_n;
  const int32_t* order;
  int32_t direct;
  int32_t type;
  const int32_t* nbr;
  const double* knorm;
  double* wf;
  int64_t n_bins;
  int32_t win0, win_len;
  double* rep;     // [n_rep][n_bins] copies for the bins outside the window, or null
  int32_t n_rep;
};

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
  for (int lsize = 1; (1 << lsize) <= N; ++lsize) {
    const int size = 1 << lsize;
    for (int lstr = lsize - 1; lstr >= 0; --lstr) {
      const int stride = 1 << lstr;
      for (int t = (int)lane; t < (N >> 1); t += 64) {
        const int a = ((t >> lstr) << (lstr + 1)) | (t & (stride - 1));
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


__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(v, off);
    if ((int)lane >= off) v += o;
  }
  return v;
}

// Radix selection over at most kBuf buffer entries: the key of entry t is
// the bit pattern of d[t] (non-negative doubles order like their bits) when
// by_index is false, else idx[t] restricted to entries with d[t] == eq.
// Returns the key of rank kk (1-based) among the eligible entries.
__device__ uint64_t wave_select(const double* d, const int32_t* idx, int n, int kk, bool by_index, double eq,
                                uint32_t* hist, uint32_t lane) {
  uint64_t prefix = 0, pmask = 0;
  const int top = by_index ? 24 : 56;
  for (int shift = top; shift >= 0; shift -= 8) {
    for (int t = (int)lane; t < 256; t += 64) hist[t] = 0u;
    wave_sync();
    for (int t = (int)lane; t < n; t += 64) {
      bool ok = true;
      uint64_t key;
      if (by_index) {
        ok = d[t] == eq;
        key = (uint64_t)(uint32_t)idx[t];
      } else {
        key = (uint64_t)__double_as_longlong(d[t]);
      }
      if (ok && (key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    wave_sync();
    const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
    const uint32_t loc = h0 + h1 + h2 + h3;
    const uint32_t incl = wave_incl_scan(loc, lane);
    const uint32_t excl = incl - loc;
    const uint64_t bal = ballot(excl < (uint32_t)kk && (uint32_t)kk <= incl);
    const int L = __ffsll((long long)bal) - 1;
    const uint32_t e = __shfl(excl, L), g0 = __shfl(h0, L), g1 = __shfl(h1, L), g2 = __shfl(h2, L);
    uint32_t digit = 4u * (uint32_t)L, before_cnt = e;
    if ((uint32_t)kk > e + g0) {
      digit += 1; before_cnt += g0;
      if ((uint32_t)kk > e + g0 + g1) {
        digit += 1; before_cnt += g1;
        if ((uint32_t)kk > e + g0 + g1 + g2) {
          digit += 1; before_cnt += g2;
        }
      }
    }
    kk -= (int)before_cnt;
    prefix |= (uint64_t)digit << shift;
    pmask |= (uint64_t)0xFF << shift;
    wave_sync();
  }
  return prefix;
}

// k-th smallest (1-based) of the first n buffer distances.  The distances
// stay in registers, kBuf / 64 per lane.  Up to three histogram levels
// narrow the active set to the bucket holding rank kk; then bit by bit over
// the keys (FP64 bit patterns, ordered like the non-negative values): the
// bits on which all active keys agree are found with one wave OR / AND
// reduction and skipped; for every other bit, from the top, one ballot per
// slot counts the still-active keys with a 0 there, and the side holding rank
// kk stays active.  It stops when one key (or a run of equal keys) is left.
// *lt_out / *eq_out receive the counts of entries below and equal to the
// result.
template <int BUF>
__device__ uint64_t wave_kth_key(const double* d, int n, int kk, uint32_t lane, uint32_t* hist, int* lt_out,
                                 int* eq_out) {
  constexpr int U = BUF / 64;
  double dv[U];
  bool act[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = (int)lane + 64 * u;
    act[u] = t < n;
    dv[u] = act[u] ? d[t] : 0.0;
  }
  int m = n, below = 0;
  // Histogram levels first: the active entries are counted into 64 buckets
  // of [their smallest, their largest] (monotone in d, so equal values share
  // a bucket and the entries of lower buckets are strictly smaller), and the
  // bucket holding rank kk becomes the active set.  Each level takes at
  // least the largest entry out, and on spread values about 63/64 of them;
  // after a few levels (or once the active values are all equal) the bit
  // loop below finishes on the few left.
  for (int level = 0; level < 3 && m > 1; ++level) {
    double lo = __builtin_inf(), hi = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (act[u]) {
        lo = fmin(lo, dv[u]);
        hi = fmax(hi, dv[u]);
      }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, off));
      hi = fmax(hi, __shfl_xor(hi, off));
    }
    if (!(hi > lo)) break;  // all active values equal: the bit loop ends at once
    hist[lane] = 0u;
    wave_sync();
    const double scale = 64.0 / (hi - lo);
    int bk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bk[u] = 64;
      if (act[u]) {
        const double q = (dv[u] - lo) * scale;
        bk[u] = q < 63.0 ? (int)q : 63;
        atomicAdd(&hist[bk[u]], 1u);
      }
    }
    wave_sync();
    const uint32_t c = hist[lane];
    const uint32_t cum = wave_incl_scan(c, lane);
    const int sel = __ffsll((long long)ballot(cum >= (uint32_t)kk)) - 1;  // m >= kk: some bucket reaches it
    const int cnt = (int)__builtin_amdgcn_readlane((int)c, sel);
    const int lower = (int)__builtin_amdgcn_readlane((int)cum, sel) - cnt;
    below += lower;
    kk -= lower;
    m = cnt;
#pragma unroll
    for (int u = 0; u < U; ++u) act[u] = act[u] && bk[u] == sel;
    wave_sync();
  }
  // Bitwise radix select over the FP64 bit patterns of the active entries
  // (non-negative doubles order like their bits): bits shared by all of
  // them are skipped, each other bit costs one ballot per slot.
  uint64_t key[U];
  uint32_t olo = 0u, ohi = 0u, alo = ~0u, ahi = ~0u;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    key[u] = act[u] ? (uint64_t)__double_as_longlong(dv[u]) : 0ull;
    if (act[u]) {
      olo |= (uint32_t)key[u];
      ohi |= (uint32_t)(key[u] >> 32);
      alo &= (uint32_t)key[u];
      ahi &= (uint32_t)(key[u] >> 32);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    olo |= (uint32_t)__shfl_xor((int)olo, off);
    ohi |= (uint32_t)__shfl_xor((int)ohi, off);
    alo &= (uint32_t)__shfl_xor((int)alo, off);
    ahi &= (uint32_t)__shfl_xor((int)ahi, off);
  }
  const uint64_t diff = ((((uint64_t)ohi) << 32) | olo) ^ ((((uint64_t)ahi) << 32) | alo);
  for (int b = diff ? 63 - __clzll((long long)diff) : -1; b >= 0 && m > 1; --b) {
    if (!((diff >> b) & 1ull)) continue;  // all active keys share this bit
    int c0 = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) c0 += __popcll(ballot(act[u] && !((key[u] >> b) & 1ull)));
    const bool one = kk > c0;  // rank kk lies among the keys with a 1 here
    if (one) {
      kk -= c0;
      below += c0;
      m -= c0;
    } else {
      m = c0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) act[u] = act[u] && (((key[u] >> b) & 1ull) != 0ull) == one;
  }
  // every key still active equals the result
  uint64_t pk = 0;
  bool has = false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (act[u] && !has) {
      pk = key[u];
      has = true;
    }
  }
  *lt_out = below;
  *eq_out = m;
  return bcast_u64(pk, __ffsll((long long)ballot(has)) - 1);
}

// Keep the k smallest entries by (d, idx) at the front of the buffer (any
// order) and return the k-th smallest d.
template <int BUF>
__device__ double wave_keep_k(double* d, int32_t* idx, int n, int k, uint32_t* hist, uint32_t lane) {
  int lt = 0, eq = 0;
  const double th = __longlong_as_double((long long)wave_kth_key<BUF>(d, n, k, lane, hist, &lt, &eq));
  int32_t icut = INT_MAX;  // entries equal to th are kept up to this index
  if (lt + eq > k) icut = (int32_t)wave_select(d, idx, n, k - lt, true, th, hist, lane);
  double vd[BUF / 64];
  int32_t vi[BUF / 64];
  bool keep[BUF / 64];
#pragma unroll
  for (int u = 0; u < BUF / 64; ++u) {
    const int t = (int)lane + 64 * u;
    keep[u] = false;
    if (t < n) {
      vd[u] = d[t];
      vi[u] = idx[t];
      keep[u] = vd[u] < th || (vd[u] == th && vi[u] <= icut);
    }
  }
  wave_sync();
  int base = 0;
#pragma unroll
  for (int u = 0; u < BUF / 64; ++u) {
    const uint64_t m = ballot(keep[u]);
    if (keep[u]) {
      const int slot = base + (int)mask_rank(m);
      d[slot] = vd[u];
      idx[slot] = vi[u];
    }
    base += __popcll(m);
  }
  wave_sync();
  return th;
}

// Cheap buffer cut: an upper bound of the k-th smallest d, not the k-th
// itself.  The n <= BUF entries (all <= hi) are counted into 64 buckets of
// [lo, hi], lo the smallest entry (LDS histogram); the bucket holding the
// k-th smallest bounds it.
// Entries of the buckets up to that one are kept (the buckets are monotone
// in d, so the kept set is exactly {d <= theta'} with theta' the largest kept
// d) and theta' is returned.  Returns a negative value, changing nothing,
// when the cut would not free 64 slots (the caller then cuts exactly).
template <int BUF>
__device__ double wave_cut_hist(double* d, int32_t* idx, int n, int k, double hi, uint32_t* hist, uint32_t lane,
                                int* n_out) {
  constexpr int U = BUF / 64;
  double dv[U];
  double lo = __builtin_inf();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = (int)lane + 64 * u;
    dv[u] = t < n ? d[t] : 0.0;
    if (t < n) lo = fmin(lo, dv[u]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lo = fmin(lo, __shfl_xor(lo, off));
  // buckets span [smallest entry, hi]: entries far from the query crowd the
  // top of [0, hi], where a histogram of [0, hi] cannot separate them
  if (!(hi > lo)) return -1.0;
  hist[lane] = 0u;
  wave_sync();
  const double scale = 64.0 / (hi - lo);
  int bk[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = (int)lane + 64 * u;
    bk[u] = 64;
    if (t < n) {
      const double q = (dv[u] - lo) * scale;
      bk[u] = q < 63.0 ? (int)q : 63;
      atomicAdd(&hist[bk[u]], 1u);
    }
  }
  wave_sync();
  const uint32_t c = hist[lane];
  const uint32_t cum = wave_incl_scan(c, lane);
  const uint64_t reach = ballot(cum >= (uint32_t)k);
  const int sel = __ffsll((long long)reach) - 1;  // n >= k, so some bucket reaches k
  const int kept = (int)__builtin_amdgcn_readlane((int)cum, sel);
  if (kept > BUF - 64) return -1.0;
  double mx = 0.0;
  double vd[U];
  int32_t vi[U];
  bool keep[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = (int)lane + 64 * u;
    keep[u] = bk[u] <= sel;
    vd[u] = 0.0;
    vi[u] = 0;
    if (keep[u]) {
      vd[u] = dv[u];
      vi[u] = idx[t];
      mx = fmax(mx, vd[u]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  wave_sync();
  int base = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t m = ballot(keep[u]);
    if (keep[u]) {
      const int slot = base + (int)mask_rank(m);
      d[slot] = vd[u];
      idx[slot] = vi[u];
    }
    base += __popcll(m);
  }
  wave_sync();
  *n_out = kept;
  return mx;
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

__global__ __launch_bounds__(256) void knn_work_clear_kernel(unsigned long long* w) {
  for (int t = threadIdx.x; t < ALB_KNN_WORK_WORDS; t += blockDim.x) w[t] = 0ull;
}

template <bool OFF32, int BUF>
__global__ __launch_bounds__(256) void knn_kernel(KnnArgs a) {
  constexpr int kBufK = BUF;          // candidate buffer entries per wave
  constexpr int kSeedK = BUF - 64;    // seed records (at least k)
  constexpr int kMaxKK = BUF - 64;    // largest k of this variant
  // one wavefront's LDS in one struct: its arrays are one base address plus
  // constant offsets (one scalar register instead of one per array)
  struct WaveLds {
    double sd[kStack];
    double bd[kBufK];
    int32_t sn[kStack];
    int32_t bi[kBufK];
    int32_t rlo[64];  // one range per lane and step
    int32_t rhi[64];
    uint32_t hist[256];
    int32_t path[kPath];
  };
  constexpr int kW = BUF > kBufLarge ? 2 : kWaves;  // wavefronts per block (LDS)
  __shared__ WaveLds lds_s[kW];

  const uint32_t lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar slot loop
  WaveLds& L = lds_s[w];
  int32_t* sn = L.sn;
  double* sd = L.sd;
  double* bd = L.bd;
  int32_t* bi = L.bi;
  int32_t* rlo = L.rlo;
  int32_t* rhi = L.rhi;
  uint32_t* hist = L.hist;
  int32_t* path = L.path;
  int plen = 0;            // levels of the previous query's descent held in path
  int32_t prev_root = -1;  // the subtree root it started from
  int32_t nn_root = -1;    // subtree root of the previous query whose neighbours bi[0, k) holds
  const Node* __restrict__ nodes = a.nodes;
  const float* __restrict__ mx = a.x;
  const float* __restrict__ my = a.y;
  const float* __restrict__ mz = a.z;
  const int64_t n = a.n;
  const int k = a.k;
  // per-wave event counts of one launch (diagnostics only; thousands of waves
  // share a launch's queries, so 32 bits hold them)
  uint32_t c_q = 0, c_cand = 0, c_nodes = 0, c_ovf = 0, c_steps = 0, c_sorts = 0, c_desc = 0, c_small = 0;

  // XCD-aware slot ranges: the grid is a multiple of kXcds blocks and block b
  // runs on XCD b % kXcds, so XCD x walks its own contiguous 1/kXcds of the
  // Morton-sorted slots and the records its queries visit stay in its L2.
  // Within the range, wavefronts take kGrab slots at a time from the XCD's
  // counter (the next chunk is fetched when one starts, so the atomic's
  // latency overlaps the work): query costs differ several-fold, and fixed
  // shares left a launch waiting for its slowest wavefronts (SCATTER
  // queries 12 % faster dealt this way).
  const int xcd = (int)(blockIdx.x % kXcds);
  // slot arithmetic in 32 bits (a launch takes fewer than 2^31 slots; the
  // host checks), which keeps scalar registers free for the traversal
  const int32_t s_lo = (int32_t)(a.q_n * xcd / kXcds), s_hi = (int32_t)(a.q_n * (xcd + 1) / kXcds);
  unsigned long long* const wk = a.work + xcd * kWorkStride;
  auto grab = [&]() -> int32_t {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(wk, (unsigned long long)kGrab);
    return s_lo + (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
  };
  int32_t next = 0;
  for (int32_t chunk = grab(); chunk < s_hi; chunk = next) {
  next = grab();
  const int32_t chunk_end = chunk + kGrab < s_hi ? chunk + kGrab : s_hi;
  for (int32_t si = chunk; si < chunk_end; ++si) {
    const int32_t qrel = a.order[a.c0 + si];
    const int64_t qg = a.q_begin + qrel;
    const int64_t qi = a.direct ? qrel : si;  // output row
    const double qx = a.qx[qg], qy = a.qy[qg], qz = a.qz[qg];
    c_q++;

    // ---- 1. seed descent; every child not taken goes on the stack (with its
    // box distance), so that stack + seed cover the whole map.  Stack entries
    // are node indices, or ~record for a single record.
    int sp = 0;
    int32_t s0 = 0, s1 = (int32_t)n - 1, root = 0;
    if (a.sroot) {
      const int32_t sc = a.qscene[qg];
      s0 = a.sfirst[sc];
      s1 = a.slast[sc];
      root = a.sroot[sc];
      if (s1 - s0 + 1 < k) {  // fewer records than neighbours: no estimate (counted, rows marked)
        c_small++;
        for (int t = (int)lane; t < k; t += 64) a.nbr[qi * k + t] = -1;
        if (lane == 0) a.radius[qi] = __builtin_nan("");
        continue;
      }
    }
    if (s1 - s0 + 1 > kSeedK) {
      // Descent toward the nearer child box.  Queries are Morton-sorted, so
      // a wavefront's consecutive queries mostly take the same path: the
      // previous query's path (path[0..plen), root first) is tested for this
      // query at all levels at once, one level per lane.  Along the longest
      // prefix where this query takes the same child it is this query's own
      // descent, so the far children of that prefix are stacked with the
      // distances the lanes computed, and the descent continues one level at
      // a time only from the first level where it turns elsewhere.  Seed and
      // stack are exactly those of a descent from the root.
      int32_t node = root;
      int lvl = 0;
      if (root == prev_root && plen > 0) {
        const int lp = plen < 64 ? plen : 64;
        bool agree = false;
        int32_t far_code = 0;
        double far_d = 0.0;
        if ((int)lane < lp) {
          const Node nd = nodes[path[lane]];
          const double dl = box_dist2(qx, qy, qz, nd.lb);
          const double dr = box_dist2(qx, qy, qz, nd.rb);
          const bool gl = dl <= dr;
          const int32_t child = gl ? nd.split : nd.split + 1;
          agree = (int)lane + 1 < plen && child == path[lane + 1];
          const int32_t far_id = gl ? nd.split + 1 : nd.split;
          const bool far_leaf = gl ? (nd.leaf_mask & 2) : (nd.leaf_mask & 1);
          far_code = far_leaf ? ~far_id : far_id;
          far_d = gl ? dr : dl;
        }
        const uint64_t dis = ballot((int)lane < lp && !agree);
        const int m = dis ? __ffsll((long long)dis) - 1 : lp;  // levels shared with the previous path
        wave_sync();
        if ((int)lane < m) {
          sn[sp + (int)lane] = far_code;
          sd[sp + (int)lane] = far_d;
        }
        sp += m;
        c_nodes += (uint32_t)m;
        c_desc += (uint32_t)m;
        lvl = m;
        node = path[m];
        wave_sync();
      }
      for (;;) {
        if (lvl < kPath && lane == 0) path[lvl] = node;
        ++lvl;
        const Node nd = nodes[node];
        c_nodes++;
        c_desc++;
        const double dl = box_dist2(qx, qy, qz, nd.lb);
        const double dr = box_dist2(qx, qy, qz, nd.rb);
        const bool go_left = dl <= dr;
        const int32_t cnt_near = go_left ? nd.split - nd.first + 1 : nd.last - nd.split;
        const bool near_leaf = go_left ? (nd.leaf_mask & 1) : (nd.leaf_mask & 2);
        const bool far_leaf = go_left ? (nd.leaf_mask & 2) : (nd.leaf_mask & 1);
        const int32_t far_id = go_left ? nd.split + 1 : nd.split;
        bool stop = true;
        if (cnt_near >= k && !near_leaf && cnt_near > kSeedK) {
          stop = false;
        } else if (cnt_near >= k && !near_leaf) {  // the near child fits the seed buffer
          s0 = go_left ? nd.first : nd.split + 1;
          s1 = go_left ? nd.split : nd.last;
        } else if (nd.last - nd.first + 1 <= kSeedK) {
          s0 = nd.first;
          s1 = nd.last;
        } else if (go_left) {  // near child (< k records) plus the first records of the far one
          s0 = nd.first;
          s1 = nd.first + kSeedK - 1;
        } else {
          s0 = nd.last - kSeedK + 1;
          s1 = nd.last;
        }
        const bool far_in_seed = stop && (nd.last - nd.first + 1 <= kSeedK);
        if (!far_in_seed) {
          if (lane == 0) {
            sn[sp] = far_leaf ? ~far_id : far_id;
            sd[sp] = go_left ? dr : dl;
          }
          ++sp;
        }
        if (stop) break;
        node = go_left ? nd.split : nd.split + 1;
      }
      plen = lvl < kPath ? lvl : kPath;
      prev_root = root;
      wave_sync();
    }
    // A bound from the previous query's neighbours (still in bi[0, k)): k
    // distinct records of this query's map lie within their largest distance
    // from this query, so its k-th distance is no larger.
    double theta_prev = __builtin_inf();
    if (nn_root == root) {
      constexpr int KU = (kMaxKK + 63) / 64;
      float xs[KU], ys[KU], zs[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int t = (int)lane + 64 * u;
        const int32_t r = bi[t < k ? t : 0];
        xs[u] = rec_ld<OFF32>(mx, r);
        ys[u] = rec_ld<OFF32>(my, r);
        zs[u] = rec_ld<OFF32>(mz, r);
      }
      double m = 0.0;
#pragma unroll
      for (int u = 0; u < KU; ++u)
        if ((int)lane + 64 * u < k) m = fmax(m, dist2(qx, qy, qz, xs[u], ys[u], zs[u]));
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
      theta_prev = m;
      wave_sync();
    }
    const int W = s1 - s0 + 1;
    {
      // W <= kSeedK: the loads of all slots issue before one wait
      constexpr int SU = (kSeedK + 63) / 64;
      float xs[SU], ys[SU], zs[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = (int)lane + 64 * u;
        const int32_t r = s0 + (t < W ? t : 0);
        xs[u] = rec_ld<OFF32>(mx, r);
        ys[u] = rec_ld<OFF32>(my, r);
        zs[u] = rec_ld<OFF32>(mz, r);
      }
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = (int)lane + 64 * u;
        if (t < W) {
          bd[t] = dist2(qx, qy, qz, xs[u], ys[u], zs[u]);
          bi[t] = s0 + t;
        }
      }
    }
    c_cand += (uint32_t)W;
    wave_sync();
    // first bound: a histogram cut of the seed over [0, its largest distance]
    double smax = 0.0;
    for (int t = (int)lane; t < W; t += 64) smax = fmax(smax, bd[t]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
    int buf_n = 0;
    double theta;
    bool theta_exact = false;  // theta is the k-th of the buffer (true after an exact cut)
    if (theta_prev < smax) {
      // the previous query's bound is tighter: keep the seed entries within
      // it (possibly fewer than k; the traversal adds the rest)
      constexpr int SU = (kSeedK + 63) / 64;
      double vd[SU];
      int32_t vi[SU];
      bool keep[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = (int)lane + 64 * u;
        keep[u] = false;
        vd[u] = 0.0;
        vi[u] = 0;
        if (t < W) {
          vd[u] = bd[t];
          vi[u] = bi[t];
          keep[u] = vd[u] <= theta_prev;
        }
      }
      wave_sync();
      int base = 0;
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const uint64_t mk = ballot(keep[u]);
        if (keep[u]) {
          const int slot = base + (int)mask_rank(mk);
          bd[slot] = vd[u];
          bi[slot] = vi[u];
        }
        base += __popcll(mk);
      }
      wave_sync();
      buf_n = base;
      theta = theta_prev;
      if (buf_n >= k) {
        int kept = 0;
        const double th2 = wave_cut_hist<BUF>(bd, bi, buf_n, k, theta, hist, lane, &kept);
        if (th2 >= 0.0) {
          theta = th2;
          buf_n = kept;
        }
      }
    } else {
      theta = wave_cut_hist<BUF>(bd, bi, W, k, smax, hist, lane, &buf_n);
      if (theta < 0.0) {
        theta = wave_keep_k<BUF>(bd, bi, W, k, hist, lane);
        buf_n = k;
        theta_exact = true;
      }
    }

    // ---- 2. pruned traversal of the stacked subtrees
    while (sp > 0) {
      c_steps++;
      // one lane per child: lanes 2e and 2e + 1 take the left and the right
      // child of stack entry e, so a step pops at most 32 entries and each
      // lane computes one box distance.  Pop policy: at most 32 entries while
      // the stack is below kStack - 64 (a step then pushes at most 64), else
      // one (the stack then grows by at most one per tree level).
      const int p = (sp >= kStack - 64) ? 1 : (sp < 32 ? sp : 32);
      const int e = (int)(lane >> 1);
      const int side = (int)(lane & 1);
      bool have = e < p;
      int32_t node = 0;
      if (have) {
        node = sn[sp - 1 - e];
        have = sd[sp - 1 - e] <= theta;
      }
      sp -= p;
      wave_sync();
      // Branch free: every lane loads its entry's node header and its own
      // child's box (lanes without an entry load node 0, which exists
      // whenever the stack is used), and selects keep what applies.
      const bool rec = have && node < 0 && side == 0;  // a single record
      const bool inner = have && node >= 0;
      const Node* np = nodes + (inner ? node : 0);
      const int4 hd = *reinterpret_cast<const int4*>(np);  // first, last, split, leaf_mask
      const float2* bp = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(np) + 16 + 24 * side);
      const float2 b01 = bp[0], b23 = bp[1], b45 = bp[2];
      const float box[6] = {b01.x, b01.y, b23.x, b23.y, b45.x, b45.y};
      // a subtree of <= kBucket records stacked by the descent (far sides are
      // stacked whatever their size) has no boxes of its own: one range,
      // taken by the entry's left lane
      const bool whole = inner && hd.y - hd.x + 1 <= kBucket;
      const int32_t lo = side ? hd.z + 1 : hd.x, hi = side ? hd.y : hd.z;
      const bool leaf = ((hd.w >> side) & 1) != 0;
      const double d = box_dist2(qx, qy, qz, box);
      const bool ok = inner && !whole && d <= theta && !(lo >= s0 && hi <= s1);
      const int kind = (rec || (whole && side == 0)) ? 2 : (ok ? ((leaf || hi - lo + 1 <= kBucket) ? 2 : 1) : 0);
      const int32_t range_lo = rec ? ~node : (whole ? hd.x : lo), range_hi = rec ? ~node : (whole ? hd.y : hi);
      c_nodes += (uint32_t)__popcll(ballot(have && side == 0));
      const uint64_t pm = ballot(kind == 1);
      const int npush = __popcll(pm);
      if (sp + npush > kStack) {  // excluded by the pop policy; refuse rather than corrupt
        c_ovf++;
        break;
      }
      if (kind == 1) {
        const int off = sp + (int)mask_rank(pm);
        sn[off] = side ? hd.z + 1 : hd.z;  // an internal child's node index
        sd[off] = d;
      }
      sp += npush;
      const uint64_t rm = ballot(kind == 2);
      const int nr = __popcll(rm);
      if (kind == 2) {
        const int off = (int)mask_rank(rm);
        rlo[off] = range_lo;
        rhi[off] = range_hi;
      }
      wave_sync();
      // scan the ranges four at a time.  The twelve coordinate loads are
      // unconditional (lanes past a range's end read its first record) so that
      // they issue back to back before one wait, instead of one round trip per
      // range behind a branch.
      for (int r0 = 0; r0 < nr; r0 += 4) {
        double d2v[4];
        bool tk[4];
        int32_t ix[4];
        float xs[4], ys[4], zs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool vr = r0 + u < nr;
          const int32_t rl = vr ? rlo[r0 + u] : 0, rh = vr ? rhi[r0 + u] : -1;
          const int32_t idx = rl + (int32_t)lane;
          ix[u] = idx;
          c_cand += (uint32_t)(rh - rl + 1);
          tk[u] = idx <= rh && !(idx >= s0 && idx <= s1);
          const int32_t li = idx <= rh ? idx : rl;
          xs[u] = rec_ld<OFF32>(mx, li);
          ys[u] = rec_ld<OFF32>(my, li);
          zs[u] = rec_ld<OFF32>(mz, li);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          d2v[u] = dist2(qx, qy, qz, xs[u], ys[u], zs[u]);
          tk[u] = tk[u] && d2v[u] <= theta;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint64_t mt = ballot(tk[u]);
          if (mt == 0) continue;
          bool take = tk[u];
          uint64_t mt2 = mt;  // the takers, unless a cut lowers theta
          if (buf_n + __popcll(mt) > kBufK) {
            c_sorts++;
            int kept = 0;
            const double th2 = wave_cut_hist<BUF>(bd, bi, buf_n, k, theta, hist, lane, &kept);
            if (th2 >= 0.0) {
              theta = th2;
              buf_n = kept;
              theta_exact = false;
            } else {
              theta = wave_keep_k<BUF>(bd, bi, buf_n, k, hist, lane);
              buf_n = k;
              theta_exact = true;
            }
            take = take && d2v[u] <= theta;
            mt2 = ballot(take);
          }
          if (take) {
            const int slot = buf_n + (int)mask_rank(mt2);
            bd[slot] = d2v[u];
            bi[slot] = ix[u];
          }
          buf_n += __popcll(mt2);
          wave_sync();
        }
      }
    }

    // ---- 3. the k nearest: cut the buffer back to k; in (distance, index)
    // order when the caller asked for the neighbour sets
    // exact k-th distance and set (histogram cuts only bound them)
    const double r2 = (buf_n > k || !theta_exact) ? wave_keep_k<BUF>(bd, bi, buf_n, k, hist, lane) : theta;
    if (a.direct) wave_sort(bd, bi, k, lane);
    for (int t = (int)lane; t < k; t += 64) a.nbr[qi * k + t] = bi[t];
    if (lane == 0) a.radius[qi] = sqrt(r2);
    nn_root = root;
    wave_sync();
  }
  }
  if (a.ctr && lane == 0) {
    if (c_q) atomicAdd(&a.ctr[ALB_CTR_QUERIES], (unsigned long long)c_q);
    if (c_cand) atomicAdd(&a.ctr[ALB_CTR_CANDIDATES], (unsigned long long)c_cand);
    if (c_nodes) atomicAdd(&a.ctr[ALB_CTR_NODES], (unsigned long long)c_nodes);
    if (c_ovf) atomicAdd(&a.ctr[ALB_CTR_OVERFLOW], (unsigned long long)c_ovf);
    if (c_steps) atomicAdd(&a.ctr[ALB_CTR_KNN_STEPS], (unsigned long long)c_steps);
    if (c_sorts) atomicAdd(&a.ctr[ALB_CTR_KNN_SORTS], (unsigned long long)c_sorts);
    if (c_desc) atomicAdd(&a.ctr[ALB_CTR_KNN_DESCENT], (unsigned long long)c_desc);
    if (c_small) atomicAdd(&a.ctr[ALB_CTR_SMALL_SCENE], (unsigned long long)c_small);
  }
}

// kernel normalisation per row: 1 / (pi r_k^2), or 1 / ((4/3) pi r_k^3) for
// the scatter map (core/simulation.py:163-187); a separate pass keeps the FP64
// pow out of the traversal kernel's register budget
__global__ __launch_bounds__(256) void knorm_kernel(const double* __restrict__ radius, double* __restrict__ knorm,
                                                    int64_t rows, int32_t volume) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const double rk = radius[r];
  const double ksize = volume ? (4.0 / 3.0) * ALB_PI * pow(rk, 3.0) : ALB_PI * (rk * rk);
  knorm[r] = 1.0 / ksize;
}

__global__ __launch_bounds__(256) void knorm_perm_kernel(const double* __restrict__ radius, double* __restrict__ knorm,
                                                         const int32_t* __restrict__ rows, int64_t n, int32_t volume) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t r = rows[t];
  const double rk = radius[r];
  const double ksize = volume ? (4.0 / 3.0) * ALB_PI * pow(rk, 3.0) : ALB_PI * (rk * rk);
  knorm[r] = 1.0 / ksize;
}

// Rows staged per contribution block: what depends only on the query (its
// direction, the outgoing direction -s normalised for the scatter map, the
// BOTTOM multiplier, energy, time, kernel normalisation, phase table) is
// computed once per row into LDS, and the block's threads then take the
// tile's (row, neighbour) pairs.
constexpr int kRowTile = 64;

struct RowTile {
  double s0[kRowTile], s1[kRowTile], s2[kRowTile];  // -s normalised (BOTTOM: the query direction s)
  double kn[kRowTile];                              // kernel normalisation
  double mul[kRowTile];                             // BOTTOM multiplier
  const double* pm[kRowTile];                       // SCATTER phase masses of the query's layer
  int32_t pn[kRowTile];
  double dl[kRowTile], rdl[kRowTile];               // SCATTER angle-bin width and its reciprocal
  double rm[kRowTile];                              // 1 / float32(sample multiplier)
  float e[kRowTile], t[kRowTile];                   // query energy and time
  int32_t scene[kRowTile];
};

template <int TYPE>
__device__ __forceinline__ void stage_row(RowTile& R, int r, const alb_scene_t& sc, const alb_queries_t& q,
                                          int64_t qg, double knorm) {
  double sx = q.dx[qg], sy = q.dy[qg], sz = q.dz[qg];
  if constexpr (TYPE != ALB_BOTTOM_REFLECTION) {  // the outgoing direction -s normalised
    double ox = -sx, oy = -sy, oz = -sz;
    normalize_f64_keep_zero(ox, oy, oz);
    sx = ox;
    sy = oy;
    sz = oz;
  }
  if constexpr (TYPE == ALB_SCATTER) {
    int pn;
    R.pm[r] = gather_phase_mass(sc, q.y[qg], &pn);
    R.pn[r] = pn;
    R.dl[r] = scatter_bin_width(pn);
    R.rdl[r] = 1.0 / R.dl[r];
  } else if constexpr (TYPE == ALB_BOTTOM_REFLECTION) {
    R.mul[r] = bottom_mult(sc, sx, sy, sz);
  }
  R.s0[r] = sx;
  R.s1[r] = sy;
  R.s2[r] = sz;
  R.kn[r] = knorm;
  R.rm[r] = 1.0 / (double)(float)sc.mult;
  R.e[r] = q.energy[qg];
  R.t[r] = q.time[qg];
}

// it / d for it < kRowTile * kMaxKHuge < 2^24 (the float quotient is within one
// of it and corrected)
__device__ __forceinline__ int small_div(int it, int d, float inv_d) {
  int q = (int)(((float)it + 0.5f) * inv_d);
  if (q * d > it) --q;
  else if ((q + 1) * d <= it) ++q;
  return q;
}

// the kernels are instantiated per photon type: each carries only its own
// multiplier's code and registers (the BRDF and BTDF paths would otherwise set
// the register budget of the scatter kernel)
// a neighbour's direction and energy (the BOTTOM multiplier needs the energy only)
struct NbrRec {
  float dx, dy, dz, e;
};

template <int TYPE>
__device__ __forceinline__ NbrRec nbr_rec(const alb_map_t& m, int32_t i) {
  if constexpr (TYPE == ALB_BOTTOM_REFLECTION) {
    return {0.0f, 0.0f, 0.0f, m.energy[i]};
  } else {
    return {m.dx[i], m.dy[i], m.dz[i], m.energy[i]};
  }
}

template <int TYPE>
__device__ __forceinline__ double pair_mult(const RowTile& R, int r, const alb_scene_t& sc, const NbrRec& p) {
  if constexpr (TYPE == ALB_BOTTOM_REFLECTION) {
    return R.mul[r];
  } else if constexpr (TYPE == ALB_SCATTER) {
    return R.pm[r][scatter_bin_o(p.dx, p.dy, p.dz, R.s0[r], R.s1[r], R.s2[r], R.pn[r], R.dl[r], R.rdl[r])];
  } else if constexpr (TYPE == ALB_SURFACE_REFLECTION) {
    return (double)brdf_reflect_o(p.dx, p.dy, p.dz, R.s0[r], R.s1[r], R.s2[r], sc.alpha_ggx, sc.F0);
  } else {
    return btdf_transmit_o(p.dx, p.dy, p.dz, R.s0[r], R.s1[r], R.s2[r], sc.alpha_ggx, sc.n_water, sc.F0,
                           sc.surface_albedo);
  }
}

// HEAT: also add every contribution into the scene's water and seafloor heat
// maps at the record's interaction points (core/simulation.py:197-214)
template <int TYPE, bool HEAT>
__global__ __launch_bounds__(256) void contrib_kernel(ContribArgs a) {
  extern __shared__ double hist[];  // [win_len]
  __shared__ RowTile R;
  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x) hist[t] = 0.0;
  const int k = a.sc.k;
  const alb_map_t& m = a.map;
  // XCD-aware tile ranges (as in knn_kernel)
  const int64_t tiles = (a.q_n + kRowTile - 1) / kRowTile;
  const int xcd = (int)(blockIdx.x % kXcds);
  const int64_t xblocks = gridDim.x / kXcds;
  const int64_t t_lo = tiles * xcd / kXcds, t_hi = tiles * (xcd + 1) / kXcds;
  for (int64_t tile = t_lo + (int64_t)(blockIdx.x / kXcds); tile < t_hi; tile += xblocks) {
    const int64_t row0 = tile * kRowTile;
    const int nrow = (int)(a.q_n - row0 < kRowTile ? a.q_n - row0 : kRowTile);
    __syncthreads();  // the previous tile's rows are no longer read
    if ((int)threadIdx.x < nrow) {
      const int64_t qi = row0 + threadIdx.x;
      const int64_t qg = a.q_begin + (a.direct ? qi : (int64_t)a.order[a.c0 + qi]);
      stage_row<TYPE>(R, threadIdx.x, a.sc, a.q, qg, a.knorm[qi]);
    }
    __syncthreads();
    // item it -> (row it / k, neighbour it % k): consecutive lanes take a
    // row's consecutive neighbours (coalesced neighbour loads, records close
    // in the map); lanes taking different rows were slower
    const int items = nrow * k;
    const float inv_k = 1.0f / (float)k;
    const int32_t* __restrict__ nb = a.nbr + row0 * k;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
      const int r = small_div(it, k, inv_k);
      const int32_t i = nb[it];
      const NbrRec p = nbr_rec<TYPE>(m, i);
      const double mul = pair_mult<TYPE>(R, r, a.sc, p);
      const float ee = R.e[r] * p.e;
      const double e = (double)ee * R.kn[r] * mul;
      const float tt = R.t[r] + m.time[i];
      const int64_t bin = (int64_t)div_f32_recip(tt, R.rm[r]);  // tt / fmult
      if (bin >= a.win0 && bin < (int64_t)a.win0 + a.win_len) atomicAdd(&hist[bin - a.win0], e);
      else if (bin >= 0 && bin < a.n_bins)
        unsafeAtomicAdd(a.rep ? &a.rep[(int64_t)(blockIdx.x % a.n_rep) * a.n_bins + bin] : &a.wf[bin], e);
      if constexpr (HEAT) {
        heat_add(a.sc.heat_water, m.fwi_x[i], m.fwi_z[i], a.sc.heat_water_center, a.sc.heat_water_extent,
                 a.sc.heat_bins, e);
        heat_add(a.sc.heat_floor, m.sfi_x[i], m.sfi_z[i], a.sc.heat_floor_center, a.sc.heat_floor_extent,
                 a.sc.heat_bins, e);
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x)
    if (hist[t] != 0.0) unsafeAtomicAdd(&a.wf[a.win0 + t], hist[t]);
}

// Many scenes: slots are sorted by scene, and each block takes a contiguous
// run of 256-item tiles, so a block meets few scene changes.  The LDS window
// belongs to the scene of the block's current tile; it is flushed when the
// tile scene changes.  Items of another scene than the tile's (a scene border
// inside a tile) and bins outside the window go straight to global atomics.
// Scene s's waveform of photon type t is wf[wf_off[s] + t * 2 * steps_s ...].
struct ContribBatchedArgs {
  const alb_scene_t* scenes;
  alb_map_t map;
  alb_queries_t q;
  int64_t c0, q_n;
  const int32_t* order;
  int32_t type, k;
  const int32_t* nbr;
  const double* knorm;
  double* wf;
  const int64_t* wf_off;
  const int32_t* win0;
  int32_t win_len;
  // replicas of this type's waveforms, [n_rep][n_scenes][nb_max], or null
  double* rep;
  int32_t n_rep, n_scenes;
  int64_t nb_max;
};

__device__ __forceinline__ void flush_window(double* hist, int win_len, double* dst) {
  for (int t = threadIdx.x; t < win_len; t += blockDim.x) {
    const double v = hist[t];
    if (v != 0.0) {
      unsafeAtomicAdd(&dst[t], v);
      hist[t] = 0.0;
    }
  }
}

template <int TYPE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void contrib_batched_kernel(ContribBatchedArgs a) {
  extern __shared__ double hist[];  // [win_len]
  __shared__ RowTile R;
  for (int t = threadIdx.x; t < a.win_len; t += blockDim.x) hist[t] = 0.0;
  const int k = a.k;
  const int64_t tiles = (a.q_n + kRowTile - 1) / kRowTile;
  const int64_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * per;
  const int64_t t1 = t0 + per < tiles ? t0 + per : tiles;
  const alb_map_t& m = a.map;
  int32_t cur = -1, cw0 = 0;
  double* cdst = nullptr;  // global bins of the LDS window
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t row0 = tile * kRowTile;
    const int nrow = (int)(a.q_n - row0 < kRowTile ? a.q_n - row0 : kRowTile);
    const int32_t ts = a.q.scene[a.order[a.c0 + row0]];
    __syncthreads();  // the previous tile's rows and window adds are done
    if (ts != cur) {  // uniform: every thread read the same slot
      if (cur >= 0) flush_window(hist, a.win_len, cdst);
      __syncthreads();
      cur = ts;
      cw0 = a.win0[ts];
      const int64_t nb = 2 * (int64_t)a.scenes[ts].steps;
      cdst = a.wf + a.wf_off[ts] + (int64_t)a.type * nb + cw0;
    }
    if ((int)threadIdx.x < nrow) {
      const int64_t qi = row0 + threadIdx.x;
      const int64_t qg = a.order[a.c0 + qi];
      const int32_t sc = a.q.scene[qg];
      R.scene[threadIdx.x] = sc;
      stage_row<TYPE>(R, threadIdx.x, a.scenes[sc], a.q, qg, a.knorm[qi]);
    }
    __syncthreads();
    const int items = nrow * k;  // (row, neighbour) as in contrib_kernel
    const float inv_k = 1.0f / (float)k;
    const int32_t* __restrict__ nb = a.nbr + row0 * k;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
      const int r = small_div(it, k, inv_k);
      const int32_t i = nb[it];
      if (i < 0) continue;  // scene with fewer records than neighbours (counted by the k-NN pass)
      const int32_t sc = R.scene[r];
      const alb_scene_t& S = a.scenes[sc];
      const NbrRec p = nbr_rec<TYPE>(m, i);
      const double mul = pair_mult<TYPE>(R, r, S, p);
      const float ee = R.e[r] * p.e;
      const double e = (double)ee * R.kn[r] * mul;
      const float tt = R.t[r] + m.time[i];
      const int64_t bin = (int64_t)div_f32_recip(tt, R.rm[r]);  // tt / float32(mult)
      if (sc == cur && bin >= cw0 && bin < (int64_t)cw0 + a.win_len) {
        atomicAdd(&hist[bin - cw0], e);
      } else {
        const int64_t nbins = 2 * (int64_t)S.steps;
        if (bin >= 0 && bin < nbins)
          unsafeAtomicAdd(a.rep ? &a.rep[((int64_t)(blockIdx.x % a.n_rep) * a.n_scenes + sc) * a.nb_max + bin]
                                : &a.wf[a.wf_off[sc] + (int64_t)a.type * nbins + bin],
                          e);
      }
    }
  }
  __syncthreads();
  if (cur >= 0) flush_window(hist, a.win_len, cdst);
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// waveform[b] += sum of the replicas' bin b, in replica order
__global__ __launch_bounds__(256) void replica_reduce_kernel(const double* __restrict__ rep, int32_t n_rep,
                                                             int64_t n_bins, double* __restrict__ wf) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_bins; b += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < n_rep; ++r) s += rep[(int64_t)r * n_bins + b];
    if (s != 0.0) wf[b] += s;
  }
}

// Grid of the query sort: the map's, or the queries' own bounding box (cubic
// cells, the map's dims and bits per axis) where the queries spread far wider
// than the records.  On the map's grid, queries outside the
// records' box are clamped onto its boundary cells: the surface-reflection
// map's records lie in the laser spot while its queries spread tens of
// metres around it, so nearly all of them shared a few keys and reached the
// k-NN in trace order.  The order changes no result: each query's neighbour
// set is exact in any order.
__global__ void query_bbox_init_kernel(uint32_t* bb) {
  if (threadIdx.x < 3) bb[threadIdx.x] = 0xFFFFFFFFu;
  else if (threadIdx.x < 6) bb[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void query_bbox_kernel(alb_queries_t q, int64_t q_begin, int64_t q_n, uint32_t* bb) {
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q_n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t a = f2ord((float)q.x[q_begin + i]), b = f2ord((float)q.y[q_begin + i]),
                   c = f2ord((float)q.z[q_begin + i]);
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

// qmeta: the map meta's layout (origin, inverse cell width, dims, bits) over
// the queries' box; a query rounded outside it is clamped by quant, which only
// moves it to a boundary cell of the sort
__global__ void query_meta_kernel(const uint32_t* bb, const double* __restrict__ map_meta, double* qmeta) {
  const int dims = (int)map_meta[6];
  const double cells = ldexp(1.0, (int)map_meta[7]);
  double ext = 0.0;
  for (int k = 0; k < 3; ++k)
    if (dims == 3 || k != 1) ext = fmax(ext, (double)ord2f(bb[3 + k]) - (double)ord2f(bb[k]));
  if (!(ext > 0.0)) ext = 1.0;
  // the map's grid unless the queries spread more than four times as wide as
  // the records (a sort over all maps on the queries' grid was slower on the
  // sweep scenes, whose queries and records mostly share their extent)
  const double map_ext = cells / map_meta[3];
  const bool own = ext > 4.0 * map_ext;
  for (int k = 0; k < 8; ++k) qmeta[k] = map_meta[k];
  if (own) {
    for (int k = 0; k < 3; ++k) qmeta[k] = (double)ord2f(bb[k]);
    for (int k = 0; k < 3; ++k) qmeta[3 + k] = cells / ext;
  }
}

__global__ __launch_bounds__(256) void query_key_kernel(alb_queries_t q, int64_t q_begin, int64_t q_n,
                                                        const double* __restrict__ meta, uint64_t* __restrict__ keys,
                                                        int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= q_n) return;
  keys[i] = morton_key(q.x[q_begin + i], q.y[q_begin + i], q.z[q_begin + i], meta);
  vals[i] = (int32_t)i;
}

// waveform of scene s, type t, bin b += sum of the replicas' (s, b), in replica order
__global__ __launch_bounds__(256) void replica_reduce_batched_kernel(const double* __restrict__ rep, int32_t n_rep,
                                                                     int32_t n_scenes, int64_t nb_max,
                                                                     const alb_scene_t* __restrict__ scenes,
                                                                     const int64_t* __restrict__ wf_off, int32_t type,
                                                                     double* __restrict__ wf) {
  const int64_t total = (int64_t)n_scenes * nb_max;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sc = i / nb_max, b = i - sc * nb_max;
    const int64_t nb = 2 * (int64_t)scenes[sc].steps;
    if (b >= nb) continue;
    double s = 0.0;
    for (int r = 0; r < n_rep; ++r) s += rep[(int64_t)r * total + i];
    if (s != 0.0) wf[wf_off[sc] + (int64_t)type * nb + b] += s;
  }
}

__global__ __launch_bounds__(256) void query_key_batched_kernel(alb_queries_t q, int64_t q_n,
                                                                const double* __restrict__ meta, int key_bits, int sb,
                                                                uint64_t* __restrict__ keys,
                                                                int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= q_n) return;
  const uint64_t mk = morton_key(q.x[i], q.y[i], q.z[i], meta);
  keys[i] = ((uint64_t)(uint32_t)q.scene[i] << (key_bits - sb)) | (mk >> sb);
  vals[i] = (int32_t)i;
}

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct GatherWs {
  size_t keys_a, keys_b, vals_a, order, temp, temp_bytes, nbr, rad, knorm, work, qgrid, total;
  int64_t chunk, knorm_rows;
};

// workspace: query sort buffers, then one chunk of neighbour rows
static int gather_layout(int64_t q_n, int k, int64_t chunk, GatherWs* L) {
  size_t tb = 0;
  const int nn = (int)(q_n > 1 ? q_n : 1);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, nn, 0, 64,
                                                    (hipStream_t)0);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_gather: sort size query failed: %s", hipGetErrorString(e));
  size_t off = 0;
  L->keys_a = off; off = al256(off + 8 * (size_t)nn);
  L->keys_b = off; off = al256(off + 8 * (size_t)nn);
  L->vals_a = off; off = al256(off + 4 * (size_t)nn);
  L->order = off; off = al256(off + 4 * (size_t)nn);
  L->temp = off; off = al256(off + tb);
  L->temp_bytes = tb;
  L->chunk = chunk;
  L->nbr = off; off = al256(off + (size_t)(chunk > 0 ? chunk : 1) * (size_t)k * sizeof(int32_t));
  L->rad = off; off = al256(off + (size_t)(chunk > 0 ? chunk : 1) * sizeof(double));
  // per-query kernel normalisation; whole query range so that rows may follow query order
  L->knorm_rows = q_n > chunk ? q_n : chunk;
  L->knorm = off; off = al256(off + (size_t)(L->knorm_rows > 0 ? L->knorm_rows : 1) * sizeof(double));
  L->work = off; off = al256(off + ALB_KNN_WORK_WORDS * sizeof(unsigned long long));
  L->qgrid = off; off = al256(off + 8 * sizeof(double) + 8 * sizeof(uint32_t));  // query-sort grid
  L->total = off;
  return ALB_OK;
}

}  // namespace alb

extern "C" int alb_gather_workspace_size(int64_t q_n, int32_t k, size_t* bytes) {
  using namespace alb;
  if (!bytes || q_n < 0 || k <= 0) return set_error(ALB_ERR_ARG, "alb_gather_workspace_size: bad arguments");
  if (q_n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_gather_workspace_size: more than 2^31 queries");
  GatherWs L;
  const int rc = gather_layout(q_n, k, q_n < kChunk ? q_n : kChunk, &L);
  if (rc != ALB_OK) return rc;
  *bytes = L.total;
  return ALB_OK;
}

static int device_cus(int* cus) {
  int dev = 0;
  *cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return alb::set_error(ALB_ERR_HIP, "alb_gather: no HIP device");
  return ALB_OK;
}

// the k-NN variant for the map size (32-bit record offsets below 2^30
// records) and k (the larger candidate buffer above 192)
static void launch_knn(const alb_map_t& map, int k, int64_t blocks, const alb::KnnArgs& ka, hipStream_t s) {
  using namespace alb;
  const dim3 g((unsigned)blocks), b(64 * kWaves);
  const bool off32 = map.n < (1ll << 30);
  if (k <= kMaxK) {
    if (off32) hipLaunchKernelGGL((knn_kernel<true, kBuf>), g, b, 0, s, ka);
    else hipLaunchKernelGGL((knn_kernel<false, kBuf>), g, b, 0, s, ka);
  } else if (k <= kMaxKLarge) {
    if (off32) hipLaunchKernelGGL((knn_kernel<true, kBufLarge>), g, b, 0, s, ka);
    else hipLaunchKernelGGL((knn_kernel<false, kBufLarge>), g, b, 0, s, ka);
  } else {  // two wavefronts per block; the per-XCD dealing takes any block count
    const dim3 b2(64 * 2);
    if (off32) hipLaunchKernelGGL((knn_kernel<true, kBufHuge>), g, b2, 0, s, ka);
    else hipLaunchKernelGGL((knn_kernel<false, kBufHuge>), g, b2, 0, s, ka);
  }
}

static int check_map(const alb_map_t* map, int k, const char* who) {
  using namespace alb;
  if (!map) return set_error(ALB_ERR_ARG, "%s: null map", who);
  if (map->n <= 0) return set_error(ALB_ERR_ARG, "%s: empty photon map", who);
  if (k <= 0 || k > map->n)
    return set_error(ALB_ERR_ARG, "%s: k = %d exceeds the map size %lld", who, k, (long long)map->n);
  if (k > kMaxKHuge) return set_error(ALB_ERR_UNSUPPORTED, "%s: k = %d above %d", who, k, kMaxKHuge);
  if (!map->keys || !map->meta || !map->x || (map->n > 1 && !map->nodes))
    return set_error(ALB_ERR_ARG, "%s: map not built", who);
  return ALB_OK;
}

extern "C" int alb_gather_order(const alb_map_t* map, const alb_queries_t* q, int64_t q_begin, int64_t q_n,
                                int32_t* order, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace alb;
  if (!map || !q || !order) return set_error(ALB_ERR_ARG, "alb_gather_order: null argument");
  if (q_n < 0 || q_begin < 0 || q_n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_gather_order: bad query range");
  if (q_n == 0) return ALB_OK;
  if (!map->meta) return set_error(ALB_ERR_ARG, "alb_gather_order: map not built");
  GatherWs L;
  int rc = gather_layout(q_n, 1, 0, &L);
  if (rc != ALB_OK) return rc;
  if (!workspace || workspace_bytes < L.total)
    return set_error(ALB_ERR_WORKSPACE, "alb_gather_order: workspace too small (%zu < %zu)", workspace_bytes, L.total);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  uint64_t* keys_a = (uint64_t*)(ws + L.keys_a);
  uint64_t* keys_b = (uint64_t*)(ws + L.keys_b);
  int32_t* vals_a = (int32_t*)(ws + L.vals_a);
  double* qmeta = (double*)(ws + L.qgrid);
  uint32_t* qbb = (uint32_t*)(qmeta + 8);
  const unsigned gq = (unsigned)((q_n + 255) / 256);
  hipLaunchKernelGGL(query_bbox_init_kernel, dim3(1), dim3(64), 0, s, qbb);
  hipLaunchKernelGGL(query_bbox_kernel, dim3(gq < 2048 ? gq : 2048), dim3(256), 0, s, *q, q_begin, q_n, qbb);
  hipLaunchKernelGGL(query_meta_kernel, dim3(1), dim3(1), 0, s, qbb, map->meta, qmeta);
  hipLaunchKernelGGL(query_key_kernel, dim3(gq), dim3(256), 0, s, *q, q_begin, q_n, qmeta, keys_a, vals_a);
  size_t tb = L.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs((void*)(ws + L.temp), tb, keys_a, keys_b, vals_a, order,
                                                    (int)q_n, 0, map->dims == 2 ? 64 : 63, s);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_gather_order: query sort failed: %s", hipGetErrorString(e));
  return check_launch("alb_gather_order");
}

extern "C" int alb_knn(const alb_map_t* map, int32_t photon_type, int32_t k, const alb_queries_t* q,
                       int64_t q_begin, const int32_t* order, int64_t c0, int64_t cn, int32_t rows_by_query,
                       int32_t* nbr, double* radius, double* knorm, unsigned long long* counters,
                       unsigned long long* work, void* stream) {
  using namespace alb;
  int rc = check_map(map, k, "alb_knn");
  if (rc != ALB_OK) return rc;
  if (!q || !order || !nbr || !radius || !knorm || !work) return set_error(ALB_ERR_ARG, "alb_knn: null argument");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES) return set_error(ALB_ERR_ARG, "alb_knn: bad photon type");
  if (cn < 0 || c0 < 0) return set_error(ALB_ERR_ARG, "alb_knn: bad slot range");
  if (cn > kMaxSlots) return set_error(ALB_ERR_ARG, "alb_knn: %lld slots in one call (at most 2^30)", (long long)cn);
  if (cn == 0) return ALB_OK;
  int cus;
  if ((rc = device_cus(&cus)) != ALB_OK) return rc;
  KnnArgs ka = knn_args(*map, *q, q_begin, c0, cn, order, rows_by_query ? 1 : 0, k, nbr, radius, counters, nullptr,
                        work);
  int64_t kb = (cn + kWaves - 1) / kWaves;
  if (kb > (int64_t)cus * 16) kb = (int64_t)cus * 16;
  kb = (kb + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL(knn_work_clear_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, work);
  launch_knn(*map, k, kb, ka, (hipStream_t)stream);
  // rows written: [0, cn) in slot order, or the query offsets of the slots when rows follow queries
  if (rows_by_query) {
    // rows are query offsets: normalise every row of this chunk's queries (all rows when one chunk)
    hipLaunchKernelGGL(knorm_perm_kernel, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       radius, knorm, order + c0, cn, photon_type == ALB_SCATTER ? 1 : 0);
  } else {
    hipLaunchKernelGGL(knorm_kernel, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, (hipStream_t)stream, radius,
                       knorm, cn, photon_type == ALB_SCATTER ? 1 : 0);
  }
  return check_launch("alb_knn");
}

extern "C" int alb_contrib(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                           const alb_queries_t* q, int64_t q_begin, const int32_t* order, int64_t c0, int64_t cn,
                           int32_t rows_by_query, const int32_t* nbr, const double* knorm, double* waveform,
                           int64_t n_bins, int32_t win0, int32_t win_len, double* replicas, int32_t n_replicas,
                           void* stream) {
  using namespace alb;
  if (!scene || !q || !nbr || !knorm || !waveform || !order) return set_error(ALB_ERR_ARG, "alb_contrib: null argument");
  int rc = check_map(map, scene->k, "alb_contrib");
  if (rc != ALB_OK) return rc;
  if (photon_type < 0 || photon_type >= ALB_N_TYPES) return set_error(ALB_ERR_ARG, "alb_contrib: bad photon type");
  if (photon_type == ALB_SCATTER && !scene->phase_mass)
    return set_error(ALB_ERR_ARG, "alb_contrib: scene without phase_mass table");
  if (win_len < 0 || win_len > kMaxWindow || win0 < 0 || (int64_t)win0 + win_len > n_bins)
    return set_error(ALB_ERR_ARG, "alb_contrib: window [%d, %d) invalid (max length %d, %lld bins)", win0,
                     win0 + win_len, kMaxWindow, (long long)n_bins);
  if (n_replicas < 0 || (n_replicas > 0 && !replicas))
    return set_error(ALB_ERR_ARG, "alb_contrib: replicas missing");
  if (cn <= 0) return ALB_OK;
  int cus;
  if ((rc = device_cus(&cus)) != ALB_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t rn = (int64_t)n_replicas * n_bins;
  if (n_replicas > 0)
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((rn + 255) / 256 < 4096 ? (rn + 255) / 256 : 4096)),
                       dim3(256), 0, s, replicas, rn);
  ContribArgs ca{*scene, *map, *q, q_begin, c0, cn, order, rows_by_query ? 1 : 0, photon_type, nbr, knorm,
                 waveform, n_bins, win0, win_len, n_replicas > 0 ? replicas : nullptr, n_replicas > 0 ? n_replicas : 1};
  int64_t cb = (cn + kRowTile - 1) / kRowTile;
  if (cb > (int64_t)cus * 8) cb = (int64_t)cus * 8;
  cb = (cb + kXcds - 1) / kXcds * kXcds;
  const bool heat = scene->heat_bins > 0 && scene->heat_water && scene->heat_floor && map->fwi_x && map->sfi_x;
  using CK = void (*)(ContribArgs);
  static const CK kern[2][ALB_N_TYPES] = {
      {contrib_kernel<ALB_BOTTOM_REFLECTION, false>, contrib_kernel<ALB_SCATTER, false>,
       contrib_kernel<ALB_SURFACE_REFLECTION, false>, contrib_kernel<ALB_SURFACE_TRANSMISSION_UP, false>},
      {contrib_kernel<ALB_BOTTOM_REFLECTION, true>, contrib_kernel<ALB_SCATTER, true>,
       contrib_kernel<ALB_SURFACE_REFLECTION, true>, contrib_kernel<ALB_SURFACE_TRANSMISSION_UP, true>}};
  hipLaunchKernelGGL(kern[heat ? 1 : 0][photon_type], dim3((unsigned)cb), dim3(256),
                     sizeof(double) * (size_t)win_len, s, ca);
  if (n_replicas > 0)
    hipLaunchKernelGGL(replica_reduce_kernel, dim3((unsigned)((n_bins + 255) / 256)), dim3(256), 0, s, replicas,
                       n_replicas, n_bins, waveform);
  return check_launch("alb_contrib");
}

// The gather sequence of alb_gather over a laid-out workspace: query order,
// then k-NN + contributions per chunk of L.chunk rows (or all rows at once
// into the caller's neighbour arrays).
static int gather_run(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type, const alb_queries_t* q,
                      int64_t q_begin, int64_t q_n, double* waveform, int64_t n_bins, int32_t win0, int32_t win_len,
                      int32_t* nbr_out, double* radius_out, char* ws, size_t ws_bytes, const alb::GatherWs& L,
                      unsigned long long* counters, void* stream, const char* who) {
  using namespace alb;
  const int k = scene->k;
  const bool direct = nbr_out != nullptr;
  int32_t* order = (int32_t*)(ws + L.order);
  int rc = alb_gather_order(map, q, q_begin, q_n, order, ws, ws_bytes, stream);
  if (rc != ALB_OK) return rc;
  if (direct && L.knorm_rows < q_n) return set_error(ALB_ERR_WORKSPACE, "%s: workspace too small", who);
  const int64_t chunk = direct ? q_n : L.chunk;
  for (int64_t c0 = 0; c0 < q_n; c0 += chunk) {
    const int64_t cn = (q_n - c0) < chunk ? (q_n - c0) : chunk;
    int32_t* nbr = direct ? nbr_out : (int32_t*)(ws + L.nbr);
    double* rad = direct ? radius_out : (double*)(ws + L.rad);
    double* kn = (double*)(ws + L.knorm);
    rc = alb_knn(map, photon_type, k, q, q_begin, order, c0, cn, direct ? 1 : 0, nbr, rad, kn, counters,
                 (unsigned long long*)(ws + L.work), stream);
    if (rc != ALB_OK) return rc;
    rc = alb_contrib(scene, map, photon_type, q, q_begin, order, c0, cn, direct ? 1 : 0, nbr, kn, waveform, n_bins,
                     win0, win_len, nullptr, 0, stream);
    if (rc != ALB_OK) return rc;
  }
  return ALB_OK;
}

static int gather_args(const alb_scene_t* scene, const alb_map_t* map, const alb_queries_t* q, int64_t q_begin,
                       int64_t q_n, double* waveform, int32_t* nbr_out, double* radius_out, const char* who) {
  using namespace alb;
  if (!scene || !map || !q || !waveform) return set_error(ALB_ERR_ARG, "%s: null argument", who);
  if (q_n < 0 || q_begin < 0 || q_n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "%s: bad query range", who);
  if (q_n == 0) return ALB_OK;
  int rc = check_map(map, scene->k, who);
  if (rc != ALB_OK) return rc;
  if ((nbr_out == nullptr) != (radius_out == nullptr))
    return set_error(ALB_ERR_ARG, "%s: nbr_out and radius_out go together", who);
  return ALB_OK;
}

extern "C" int alb_gather(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                          const alb_queries_t* q, int64_t q_begin, int64_t q_n, double* waveform,
                          int64_t n_bins, int32_t win0, int32_t win_len, int32_t* nbr_out, double* radius_out,
                          void* workspace, size_t workspace_bytes, unsigned long long* counters, void* stream) {
  using namespace alb;
  int rc = gather_args(scene, map, q, q_begin, q_n, waveform, nbr_out, radius_out, "alb_gather");
  if (rc != ALB_OK || q_n == 0) return rc;
  GatherWs L;
  rc = gather_layout(q_n, scene->k, q_n < kChunk ? q_n : kChunk, &L);
  if (rc != ALB_OK) return rc;
  if (!workspace || workspace_bytes < L.total)
    return set_error(ALB_ERR_WORKSPACE, "alb_gather: workspace too small (%zu < %zu)", workspace_bytes, L.total);
  return gather_run(scene, map, photon_type, q, q_begin, q_n, waveform, n_bins, win0, win_len, nbr_out, radius_out,
                    (char*)workspace, workspace_bytes, L, counters, stream, "alb_gather");
}

// ------------------------------------------------ gather without allocation
// alb_gather_static keeps its workspace in `scratch`, a file-scope device
// array each GPU holds for the life of the process.  Host code never passes
// the array's name to a kernel or a copy: on the host it is only a stand-in
// for the symbol, and its device address on the current GPU comes from
// hipGetSymbolAddress, taken once per device (static_scratch) and kept.
namespace alb {
constexpr size_t kStaticBytes = (size_t)ALB_GATHER_STATIC_SCRATCH_BYTES;
constexpr int64_t kStaticChunk = 1 << 14;  // neighbour rows per k-NN round (12.6 MB at k = 192)
__device__ __align__(256) unsigned char scratch[kStaticBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];
}  // namespace alb

static int static_scratch(void** out) {
  using namespace alb;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "alb_gather_static: no HIP device: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(ALB_ERR_UNSUPPORTED, "alb_gather_static: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(ALB_ERR_HIP, "alb_gather_static: scratch address on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ALB_OK;
}

extern "C" int alb_gather_static(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                                 const alb_queries_t* q, int64_t q_begin, int64_t q_n, double* waveform,
                                 int64_t n_bins, int32_t win0, int32_t win_len, int32_t* nbr_out, double* radius_out,
                                 unsigned long long* counters, void* stream) {
  using namespace alb;
  int rc = gather_args(scene, map, q, q_begin, q_n, waveform, nbr_out, radius_out, "alb_gather_static");
  if (rc != ALB_OK || q_n == 0) return rc;
  if (q_n > ALB_GATHER_STATIC_MAX_QUERIES)
    return set_error(ALB_ERR_ARG, "alb_gather_static: %lld queries exceed the maximum of %d per call",
                     (long long)q_n, ALB_GATHER_STATIC_MAX_QUERIES);
  GatherWs L;
  if (scene->k > kMaxKLarge)
    return set_error(ALB_ERR_UNSUPPORTED, "alb_gather_static: k = %d above %d", scene->k, kMaxKLarge);
  rc = gather_layout(q_n, scene->k, q_n < kStaticChunk ? q_n : kStaticChunk, &L);
  if (rc != ALB_OK) return rc;
  if (L.total > kStaticBytes)
    return set_error(ALB_ERR_WORKSPACE, "alb_gather_static: %lld queries at k = %d need %zu B of scratch, %zu held",
                     (long long)q_n, scene->k, L.total, kStaticBytes);
  void* ws = nullptr;
  if ((rc = static_scratch(&ws)) != ALB_OK) return rc;
  return gather_run(scene, map, photon_type, q, q_begin, q_n, waveform, n_bins, win0, win_len, nbr_out, radius_out,
                    (char*)ws, kStaticBytes, L, counters, stream, "alb_gather_static");
}

extern "C" int alb_gather_order_batched(const alb_map_t* map, const alb_scene_index_t* index,
                                        const alb_queries_t* q, int64_t q_n, int32_t* order, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  using namespace alb;
  if (!map || !index || !q || !order) return set_error(ALB_ERR_ARG, "alb_gather_order_batched: null argument");
  if (q_n < 0 || q_n > 0x7FFFFFF0) return set_error(ALB_ERR_ARG, "alb_gather_order_batched: bad query count");
  if (q_n == 0) return ALB_OK;
  if (!map->meta) return set_error(ALB_ERR_ARG, "alb_gather_order_batched: map not built");
  if (!q->scene) return set_error(ALB_ERR_ARG, "alb_gather_order_batched: queries carry no scene array");
  if (index->scene_bits != scene_bits(index->n_scenes))
    return set_error(ALB_ERR_ARG, "alb_gather_order_batched: scene index not built by alb_map_build_batched");
  GatherWs L;
  int rc = gather_layout(q_n, 1, 0, &L);
  if (rc != ALB_OK) return rc;
  if (!workspace || workspace_bytes < L.total)
    return set_error(ALB_ERR_WORKSPACE, "alb_gather_order_batched: workspace too small (%zu < %zu)", workspace_bytes,
                     L.total);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  uint64_t* keys_a = (uint64_t*)(ws + L.keys_a);
  uint64_t* keys_b = (uint64_t*)(ws + L.keys_b);
  int32_t* vals_a = (int32_t*)(ws + L.vals_a);
  const int key_bits = map->dims == 2 ? 64 : 63;
  const int sb = index->scene_bits;
  const unsigned g = (unsigned)((q_n + 255) / 256);
  // the sort's grid, as in alb_gather_order
  double* qmeta = (double*)(ws + L.qgrid);
  uint32_t* qbb = (uint32_t*)(qmeta + 8);
  hipLaunchKernelGGL(query_bbox_init_kernel, dim3(1), dim3(64), 0, s, qbb);
  hipLaunchKernelGGL(query_bbox_kernel, dim3(g < 2048 ? g : 2048), dim3(256), 0, s, *q, (int64_t)0, q_n, qbb);
  hipLaunchKernelGGL(query_meta_kernel, dim3(1), dim3(1), 0, s, qbb, map->meta, qmeta);
  if (sb > 0)
    hipLaunchKernelGGL(query_key_batched_kernel, dim3(g), dim3(256), 0, s, *q, q_n, qmeta, key_bits, sb, keys_a,
                       vals_a);
  else
    hipLaunchKernelGGL(query_key_kernel, dim3(g), dim3(256), 0, s, *q, (int64_t)0, q_n, qmeta, keys_a, vals_a);
  size_t tb = L.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs((void*)(ws + L.temp), tb, keys_a, keys_b, vals_a, order,
                                                    (int)q_n, 0, key_bits, s);
  if (e != hipSuccess)
    return set_error(ALB_ERR_HIP, "alb_gather_order_batched: query sort failed: %s", hipGetErrorString(e));
  return check_launch("alb_gather_order_batched");
}

extern "C" int alb_knn_batched(const alb_map_t* map, const alb_scene_index_t* index, int32_t photon_type, int32_t k,
                               const alb_queries_t* q, const int32_t* order, int64_t c0, int64_t cn, int32_t* nbr,
                               double* radius, double* knorm, unsigned long long* counters,
                               unsigned long long* work, void* stream) {
  using namespace alb;
  if (!map || !index || !q || !order || !nbr || !radius || !knorm || !work)
    return set_error(ALB_ERR_ARG, "alb_knn_batched: null argument");
  if (map->n <= 0) return set_error(ALB_ERR_ARG, "alb_knn_batched: empty photon map");
  if (k <= 0 || k > kMaxKHuge)
    return set_error(ALB_ERR_UNSUPPORTED, "alb_knn_batched: k = %d outside [1, %d]", k, kMaxKHuge);
  if (!map->keys || !map->meta || !map->x || (map->n > 1 && !map->nodes))
    return set_error(ALB_ERR_ARG, "alb_knn_batched: map not built");
  if (!index->first || !index->last || !index->root || !q->scene)
    return set_error(ALB_ERR_ARG, "alb_knn_batched: scene index or query scenes missing");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES) return set_error(ALB_ERR_ARG, "alb_knn_batched: bad photon type");
  if (cn < 0 || c0 < 0) return set_error(ALB_ERR_ARG, "alb_knn_batched: bad slot range");
  if (cn > kMaxSlots)
    return set_error(ALB_ERR_ARG, "alb_knn_batched: %lld slots in one call (at most 2^30)", (long long)cn);
  if (cn == 0) return ALB_OK;
  int cus, rc;
  if ((rc = device_cus(&cus)) != ALB_OK) return rc;
  KnnArgs ka = knn_args(*map, *q, 0, c0, cn, order, 0, k, nbr, radius, counters, index, work);
  int64_t kb = (cn + kWaves - 1) / kWaves;
  if (kb > (int64_t)cus * 16) kb = (int64_t)cus * 16;
  kb = (kb + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL(knn_work_clear_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, work);
  launch_knn(*map, k, kb, ka, (hipStream_t)stream);
  hipLaunchKernelGGL(knorm_kernel, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, (hipStream_t)stream, radius,
                     knorm, cn, photon_type == ALB_SCATTER ? 1 : 0);
  return check_launch("alb_knn_batched");
}

extern "C" int alb_contrib_batched(const alb_scene_t* scenes, int32_t n_scenes, int32_t k, const alb_map_t* map,
                                   int32_t photon_type, const alb_queries_t* q, const int32_t* order, int64_t c0,
                                   int64_t cn, const int32_t* nbr, const double* knorm, double* waveforms,
                                   const int64_t* wf_offset, const int32_t* win0, int32_t win_len, double* replicas,
                                   int32_t n_replicas, int64_t nb_max, void* stream) {
  using namespace alb;
  if (!scenes || !map || !q || !order || !nbr || !knorm || !waveforms || !wf_offset || !win0)
    return set_error(ALB_ERR_ARG, "alb_contrib_batched: null argument");
  if (n_scenes < 1) return set_error(ALB_ERR_ARG, "alb_contrib_batched: bad scene count");
  if (!q->scene) return set_error(ALB_ERR_ARG, "alb_contrib_batched: queries carry no scene array");
  if (photon_type < 0 || photon_type >= ALB_N_TYPES) return set_error(ALB_ERR_ARG, "alb_contrib_batched: bad photon type");
  if (k <= 0 || k > kMaxKHuge)
    return set_error(ALB_ERR_UNSUPPORTED, "alb_contrib_batched: k = %d outside [1, %d]", k, kMaxKHuge);
  if (win_len < 0 || win_len > kMaxWindowLarge)
    return set_error(ALB_ERR_ARG, "alb_contrib_batched: window length %d outside [0, %d]", win_len, kMaxWindowLarge);
  if (n_replicas < 0 || (n_replicas > 0 && (!replicas || nb_max <= 0)))
    return set_error(ALB_ERR_ARG, "alb_contrib_batched: replicas missing");
  if (cn <= 0) return ALB_OK;
  int cus, rc;
  if ((rc = device_cus(&cus)) != ALB_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t rn = (int64_t)n_replicas * n_scenes * nb_max;
  if (n_replicas > 0)
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((rn + 255) / 256 < 8192 ? (rn + 255) / 256 : 8192)),
                       dim3(256), 0, s, replicas, rn);
  ContribBatchedArgs ca{scenes, *map, *q, c0, cn, order, photon_type, k, nbr, knorm, waveforms, wf_offset, win0,
                        win_len, n_replicas > 0 ? replicas : nullptr, n_replicas > 0 ? n_replicas : 1, n_scenes,
                        nb_max};
  const int64_t tiles = (cn + kRowTile - 1) / kRowTile;
  using CBK = void (*)(ContribBatchedArgs);
  static const CBK bkern[2][ALB_N_TYPES] = {
      {contrib_batched_kernel<ALB_BOTTOM_REFLECTION, 256>, contrib_batched_kernel<ALB_SCATTER, 256>,
       contrib_batched_kernel<ALB_SURFACE_REFLECTION, 256>, contrib_batched_kernel<ALB_SURFACE_TRANSMISSION_UP, 256>},
      {contrib_batched_kernel<ALB_BOTTOM_REFLECTION, 1024>, contrib_batched_kernel<ALB_SCATTER, 1024>,
       contrib_batched_kernel<ALB_SURFACE_REFLECTION, 1024>, contrib_batched_kernel<ALB_SURFACE_TRANSMISSION_UP, 1024>}};
  const size_t lds = sizeof(double) * (size_t)win_len;
  if (win_len <= kMaxWindow) {  // 256-thread blocks, several per CU
    int64_t cb = (int64_t)cus * 8;
    if (cb > tiles) cb = tiles;
    hipLaunchKernelGGL(bkern[0][photon_type], dim3((unsigned)cb), dim3(256), lds, s, ca);
  } else {  // a window beyond one block's default 64 KB of LDS: one 1024-thread block per CU
    int dev = 0, maxsh = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return set_error(ALB_ERR_HIP, "alb_contrib_batched: no HIP device");
    if (lds + sizeof(RowTile) > (size_t)maxsh)
      return set_error(ALB_ERR_UNSUPPORTED, "alb_contrib_batched: window of %d bins needs %zu B of LDS per block, %d available",
                       win_len, lds + sizeof(RowTile), maxsh);
    if (hipFuncSetAttribute((const void*)bkern[1][photon_type], hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return set_error(ALB_ERR_HIP, "alb_contrib_batched: LDS size %zu refused", lds);
    int64_t cb = (int64_t)cus;
    if (cb > tiles) cb = tiles;
    hipLaunchKernelGGL(bkern[1][photon_type], dim3((unsigned)cb), dim3(1024), lds, s, ca);
  }
  if (n_replicas > 0) {
    const int64_t tot = (int64_t)n_scenes * nb_max;
    hipLaunchKernelGGL(replica_reduce_batched_kernel, dim3((unsigned)((tot + 255) / 256 < 8192 ? (tot + 255) / 256 : 8192)),
                       dim3(256), 0, s, replicas, n_replicas, n_scenes, nb_max, scenes, wf_offset, photon_type,
                       waveforms);
  }
  return check_launch("alb_contrib_batched");
}
