/*
 * alb_abi.h — C ABI of the MI355X photon-tracing / photon-mapping library
 * (libalbsim_hip.so).  Plain structs of scalars and pointers; no C++ or
 * torch types cross this boundary.  Every entry point returns ALB_OK (0) or a
 * negative alb_status and never throws; alb_last_error() gives the message.
 *
 * The reference (PyALBSim) has no native code and no FFI: its boundary for this
 * path is the Python call run_linear(simulation_config, run_config, rng)
 * (src/alb_sim/execution/linear.py:77-117).  The entry points below replace the
 * NumPy/SciPy work that call performs:
 *
 *   alb_trace          <- Simulation.simulate_batch + simulate_photon_step +
 *                         handle_water/enter/exit/seafloor + store_photons
 *                         (src/alb_sim/core/simulation.py:71-130, 246-570)
 *   alb_map_build      <- build_photon_map_data + PhotonMapIndex (scipy KDTree)
 *                         (src/alb_sim/photon_mapping/build_photon_map_data.py:8-62,
 *                          src/alb_sim/photon_mapping/photon_map_index.py:6-17)
 *   alb_gather         <- Simulation.sample_photons (KDTree.query k-NN, energy
 *                         multipliers, np.add.at into the waveform)
 *                         (src/alb_sim/core/simulation.py:132-214)
 *
 * Conventions
 *  - Every device buffer is owned by the caller.  Scratch memory is caller-
 *    provided ("workspace"); its size comes from the *_workspace_size calls.
 *  - Work is enqueued asynchronously on the given HIP stream (a hipStream_t
 *    passed as void*; NULL = the default stream).
 *  - Record and query sinks count every item they are offered, also past
 *    capacity.  When a count exceeds capacity the caller reallocates and reruns;
 *    Philox counters make the rerun produce the same items.
 *  - Photon-type indices follow the reference enum order
 *    (src/alb_sim/photon_mapping/photon_type.py:4-10).
 */
#ifndef ALB_ABI_H
#define ALB_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALB_ABI_VERSION 11

/* photon types, reference enum order (BOTTOM_REFLECTION = auto() = 1, ...) */
enum alb_photon_type {
  ALB_BOTTOM_REFLECTION = 0,
  ALB_SCATTER = 1,
  ALB_SURFACE_REFLECTION = 2,
  ALB_SURFACE_TRANSMISSION_UP = 3,
  ALB_N_TYPES = 4
};

enum alb_status {
  ALB_OK = 0,
  ALB_ERR_ARG = -1,
  ALB_ERR_CAPACITY = -2,
  ALB_ERR_HIP = -3,
  ALB_ERR_UNSUPPORTED = -4,
  ALB_ERR_WORKSPACE = -5
};

enum alb_rng_mode { ALB_RNG_PHILOX = 0, ALB_RNG_REPLAY = 1 };

enum alb_pass { ALB_PASS_FORWARD = 0, ALB_PASS_BACKWARD = 1 };

/* device event counters (uint64 each), written with atomics by alb_trace and
 * alb_gather; used for the algorithmic-work (flop / byte) accounting */
enum alb_counter {
  ALB_CTR_PATHS = 0,         /* photons emitted                            */
  ALB_CTR_AIR_LEGS = 1,      /* straight air legs solved in closed form     */
  ALB_CTR_WATER_STEPS = 2,   /* in-water sub-steps covered (float32 optical
                                depth updates reproduced)                   */
  ALB_CTR_SCATTER = 3,       /* scatter events                              */
  ALB_CTR_ENTER = 4,         /* air->water crossings                        */
  ALB_CTR_EXIT = 5,          /* water->air crossings                        */
  ALB_CTR_FLOOR = 6,         /* seafloor hits                               */
  ALB_CTR_REPLAY_EXHAUSTED = 7, /* replay streams that ran out of uniforms  */
  ALB_CTR_QUERIES = 8,       /* gather queries processed                    */
  ALB_CTR_CANDIDATES = 9,    /* map records whose distance was evaluated    */
  ALB_CTR_NODES = 10,        /* tree nodes visited                          */
  ALB_CTR_OVERFLOW = 11,     /* gather queries that exceeded a kNN buffer   */
  ALB_CTR_DEPTH_SOLVES = 12, /* optical-depth binade jumps and single steps */
  ALB_CTR_KNN_STEPS = 13,    /* k-NN traversal steps (node batches)          */
  ALB_CTR_KNN_SORTS = 14,    /* candidate-buffer sorts                       */
  ALB_CTR_KNN_DESCENT = 15,  /* seed-descent levels                          */
  ALB_CTR_SMALL_SCENE = 16,  /* batched queries whose scene map holds fewer
                                than k records (rows left at -1 / NaN)      */
  ALB_CTR_REPLAY_MISSES = 17,/* replayed events whose record was not found   */
  ALB_CTR_NEIGHBOURS = 18,   /* records taken by the fixed-radius gather    */
  ALB_N_COUNTERS = 19
};

/* One turbidity layer of a layered or depth-profiled water column
 * (src/alb_sim/physics/models/water.py:39-416).  A photon at depth d (= -(y -
 * y_ws)) is in the first layer whose bottom exceeds d (searchsorted right,
 * clamped to the last layer) and at local depth d - top. */
typedef struct alb_layer {
  double top;               /* float32 cumulative height of the layers above       */
  double bottom;            /* FP64 cumulative height including this layer         */
  double height;
  int32_t profiled;         /* 0: constant properties, 1: depth profiles           */
  int32_t lut_n;            /* profiled: entries of the attenuation table          */
  float alpha;              /* constant: float32 lidar attenuation (1/m)           */
  float omega0;             /* constant: float32 single-scattering albedo          */
  double v;                 /* constant: float32(c / n)                            */
  int32_t a_kind;           /* absorption: 0 constant, 1 linear, 2 exponential,
                               3 tabulated (a_tab)                                  */
  int32_t n_kind;           /* refractive index: same codes (3: n_tab)             */
  double a0, a1;            /* absorption at the top / bottom (profile start, end) */
  double n0, n1;            /* refractive index at the top / bottom                */
  const float* z_lut;       /* profiled: [lut_n] float32 local depths              */
  const double* alpha_lut;  /* profiled: [lut_n] attenuation at z_lut              */
  const double* phase_theta;/* [phase_n] phase-function tables of this layer       */
  const double* phase_cdf;
  const double* phase_mass;
  int32_t phase_n;
  int32_t pad0;
  /* a_kind / n_kind 3: any other depth profile, sampled at the fractional
   * depths t = i / (tab_n - 1), i = 0 .. tab_n - 1, and interpolated linearly
   * in t in FP64 (NULL / 0 when neither property uses one) */
  const double* a_tab;
  const double* n_tab;
  int32_t tab_n;
  int32_t pad1;
} alb_layer_t;

/* Scene constants, computed on the host in FP64 from SimulationConfig
 * (src/alb_sim/physics/models/simulation.py:21-74, water.py:117-245,
 *  sea_surface.py:30-33, laser.py:22-33).  Pointer members are device pointers
 * for the HIP library and host pointers for the CPU oracle. */
typedef struct alb_scene {
  double y_ws;              /* water-surface plane y = -flying_height (m)          */
  double y_sf;              /* seafloor plane y = -(flying_height + depth)         */
  double laser_dir[3];      /* unit laser / sensor boresight                       */
  float rot_laser[9];       /* float32 rotation +z -> laser_dir, row major         */
  float pad0;
  double divergence;        /* forward cone half-angle (rad)                       */
  double rayleigh_sigma;    /* divergence / 3                                      */
  double cos_fov;           /* cos(receiver half-angle)                            */
  double aperture_radius;   /* receiver disk radius (m)                            */
  double v_air;             /* float32(2.998e8)                                    */
  double v_water;           /* float32(2.998e8 / n)                                */
  double dt;                /* 1 / (sample_multiplier * sample_rate)               */
  double n_water;           /* refractive index of layer 0                          */
  double eta;               /* 1 / n_water (entry and exit, as the reference)      */
  double F0;                /* ((1-n)/(1+n))^2                                     */
  double alpha_ggx;         /* roughness^2                                         */
  double surface_albedo;
  double floor_albedo;
  double absorption;        /* layer absorption coefficient a (1/m)               */
  float alpha_l;            /* float32 lidar attenuation coefficient (1/m)         */
  float omega0;             /* float32 single-scattering albedo                    */
  int64_t n_substeps;       /* S = steps * sample_multiplier                       */
  int32_t steps;            /* waveform has 2*steps bins                           */
  int32_t mult;             /* sample_multiplier                                   */
  int32_t k;                /* photon_mapping_k                                    */
  int32_t phase_n;          /* entries of the phase tables (18000)                 */
  const double* phase_theta;/* [phase_n] scattering-angle grid                     */
  const double* phase_cdf;  /* [phase_n] normalised CDF                            */
  const double* phase_mass; /* [phase_n] CDF mass of each scattering-angle bin used
                               by the gather (bin i spans [i, i+1) * float32(pi -
                               eps)/phase_n), evaluated with np.interp semantics */
  /* adjoint only (NULL otherwise): derivatives with respect to the phase
   * function's parameter (Henyey-Greenstein g) */
  const double* phase_score;     /* [phase_n] d log(sampling probability)/dg of the
                                    inverse-CDF segment: [0] the atom below cdf[0],
                                    [i] the interval [cdf[i-1], cdf[i])            */
  const double* phase_dlog_mass; /* [phase_n] d log(phase_mass[i])/dg               */
  /* layered column (0 / NULL: one constant layer described by the fields
   * above, traced with closed-form water legs); otherwise water sub-steps are
   * taken one at a time with per-depth properties */
  int32_t n_layers;
  int32_t pad1;
  const alb_layer_t* layers;     /* [n_layers] (device pointer for the HIP library)  */
  /* sampled heat maps (src/alb_sim/math/accumulate_heatmap.py:6-124, hooked at
   * core/simulation.py:197-214, 523-537); heat_bins = 0: none.  alb_trace
   * (forward) counts every seafloor hit into heat_radius[bins x bins] at
   * (surface radius / water extent, floor radius / floor extent) * bins;
   * alb_gather / alb_contrib add each contribution into heat_water and
   * heat_floor [bins x bins] at the record's first water / seafloor
   * interaction, (x - centre + extent) / (2 extent / bins) along x and z.
   * One-scene entry points only. */
  int32_t heat_bins;
  int32_t pad2;
  double heat_water_extent;      /* half-width around the surface centre (m)        */
  double heat_floor_extent;      /* half-width around the floor centre (m)          */
  double heat_water_center[3];   /* laser / water-surface intersection              */
  double heat_floor_center[3];   /* refracted beam / seafloor intersection          */
  unsigned int* heat_radius;     /* [bins * bins] counts                            */
  double* heat_water;            /* [bins * bins] energy                            */
  double* heat_floor;            /* [bins * bins] energy                            */
  /* tilted floor (extension of the reference's flat floor y = y_sf): the
   * plane floor_normal . p = floor_offset, water on the side where it is
   * positive, with the frame (floor_t1, floor_normal, floor_t2) for the
   * cosine-hemisphere reflection.  1: a tilted floor (single-layer scenes keep
   * closed-form water legs, layered ones the stepped legs); 0: the flat
   * floor, these fields unused. */
  int32_t tilted;
  int32_t pad3;
  double floor_normal[3];
  double floor_offset;
  double floor_t1[3];
  double floor_t2[3];
} alb_scene_t;

typedef struct alb_rng {
  int32_t mode;             /* ALB_RNG_PHILOX or ALB_RNG_REPLAY                     */
  uint32_t key[2];          /* Philox4x32-10 key                                    */
  uint32_t stream;          /* counter word 3 (pass / scene)                        */
  const double* replay;     /* [n * replay_len] uniforms in [0,1), slot = local idx */
  int64_t replay_len;       /* uniforms per photon slot                             */
} alb_rng_t;

/* Forward-pass record sink, one per photon type: float32 SoA like
 * PhotonMapData (src/alb_sim/photon_mapping/photon_map_data.py:6-15) without
 * the heat-map fields. */
typedef struct alb_records {
  float* x;
  float* y;
  float* z;
  float* dx;
  float* dy;
  float* dz;
  float* energy;
  float* time;              /* sub-step index of the event (float like reference) */
  int64_t* tag;             /* optional: (photon index << 16) | event ordinal     */
  int64_t capacity;
  unsigned long long* count;/* device counter                                     */
  int32_t* scene;           /* optional: scene of each record (batched traces)     */
  /* optional (heat maps, alb_trace only): the photon's first water
   * interaction (surface entry point) and latest seafloor interaction at the
   * event, float32, NaN before the interaction
   * (core/simulation.py:85-86, 393, 521) */
  float* fwi_x;
  float* fwi_y;
  float* fwi_z;
  float* sfi_x;
  float* sfi_y;
  float* sfi_z;
} alb_records_t;

/* Backward-pass query sink (sensor-path vertices), one per photon type. */
typedef struct alb_queries {
  double* x;
  double* y;
  double* z;
  double* dx;
  double* dy;
  double* dz;
  float* energy;
  float* time;
  int64_t* tag;             /* optional                                            */
  int64_t capacity;
  unsigned long long* count;
  int32_t* scene;           /* optional: scene of each query (batched traces)      */
  float* fwi_x;             /* unused for queries; same layout as alb_records_t    */
  float* fwi_y;
  float* fwi_z;
  float* sfi_x;
  float* sfi_y;
  float* sfi_z;
} alb_queries_t;

/* A built photon map: records sorted along a Morton curve plus a binary radix
 * tree over them.  Internal node i (0 <= i < n-1; node 0 is the root) is 64
 * bytes:
 *   int32 first, last    sorted records it covers
 *   int32 split          left child covers [first, split], right [split+1, last];
 *                        an internal child's node index is split (left) or
 *                        split+1 (right)
 *   int32 leaf_mask      bit 0: left child is the single record split,
 *                        bit 1: right child is the single record split+1
 *   float left_box[6]    min x,y,z, max x,y,z of the left child's records
 *   float right_box[6]   same for the right child
 * The boxes are written for the root and for every node covering more than
 * 64 records; a subtree of at most 64 records is searched as one record range,
 * and its nodes' boxes are left unwritten. */
typedef struct alb_map {
  int64_t n;                /* records                                             */
  int32_t dims;             /* 2: key over (x,z) (planar maps), 3: over (x,y,z)    */
  int32_t pad0;
  float* x;                 /* sorted SoA [n]                                      */
  float* y;
  float* z;
  float* dx;
  float* dy;
  float* dz;
  float* energy;
  float* time;
  int32_t* perm;            /* [n] sorted position -> input record index          */
  uint64_t* keys;           /* [n] sorted Morton keys                              */
  double* meta;             /* [8] grid origin x,y,z, inverse cell width x,y,z     */
  void* nodes;              /* [n-1] 64-byte internal nodes (layout above)         */
  /* optional sorted interaction points (heat maps): permuted from the records'
   * arrays by alb_map_build when both are set */
  float* fwi_x;
  float* fwi_y;
  float* fwi_z;
  float* sfi_x;
  float* sfi_y;
  float* sfi_z;
} alb_map_t;

/* Scene index of a map (or query set) holding records of many scenes: after
 * alb_map_build_batched, scene s owns sorted records [first[s], last[s]]
 * (first = 0, last = -1 when it has none) and root[s] is its subtree's root
 * node (~record when the scene has a single record).  scene_bits: bits the
 * scene number takes at the top of the sort keys (written by the build). */
typedef struct alb_scene_index {
  int32_t n_scenes;
  int32_t scene_bits;
  int32_t* first;           /* [n_scenes] device                                   */
  int32_t* last;            /* [n_scenes] device                                   */
  int32_t* root;            /* [n_scenes] device                                   */
} alb_scene_index_t;

/* ---------------------------------------------------------------- entry points */

int alb_abi_version(void);
size_t alb_scene_size(void);           /* sizeof(alb_scene_t), layout check   */
const char* alb_last_error(void);      /* thread-local message of last error  */

/* Trace photons [photon_begin, photon_begin + n) of one pass.  Forward pass
 * writes map records into out[4]; backward pass writes gather queries into
 * q[4] (enter vertices go to both surface types).  counters: device
 * uint64[ALB_N_COUNTERS] or NULL.
 * Replaces core/simulation.py:71-107 (simulate_batch) with a persistent grid
 * (each wavefront refills its lanes from its own range of photons; a photon's
 * random stream depends on its index only), closed-form air legs and exact
 * sub-step quantisation in water. */
int alb_trace(const alb_scene_t* scene, const alb_rng_t* rng, int32_t pass,
              int64_t photon_begin, int64_t n, const alb_records_t* out,
              const alb_queries_t* q, unsigned long long* counters,
              void* stream);

/* Many scenes in one persistent launch: scenes (device array [n_scenes]),
 * photon_offsets (device int64 [n_scenes + 1], prefix of photons per scene).
 * Photon p of scene s uses Philox counter words {draw/2, p, p>>32,
 * rng->stream + 2 s}; rng->stream must carry the pass in its low bit, and
 * with rng->stream = pass scene 0 reproduces alb_trace.  Records / queries
 * carry their scene in the sinks' scene arrays (optional) and the photon index
 * within the scene in the tags.  work_counter: device uint64 set to 0 before
 * the call; lanes whose photon has finished take new photons from it every
 * refill_interval events.  block: 64, 128 or 256 threads.  Philox streams
 * only.  No counterpart in the reference, which traces one scene per
 * run_linear call (core/simulation.py:71-107). */
int alb_trace_batched(const alb_scene_t* scenes, int32_t n_scenes, const int64_t* photon_offsets,
                      const alb_rng_t* rng, int32_t pass, const alb_records_t* out,
                      const alb_queries_t* q, unsigned long long* work_counter, int32_t block,
                      int32_t refill_interval, unsigned long long* counters, void* stream);

/* Workspace bytes needed to build a map of n records. */
int alb_map_workspace_size(int64_t n, size_t* bytes);

/* Sort n input records along a Morton curve (dims 2 or 3) into map's SoA and
 * build the radix tree with bounding boxes.  Replaces
 * build_photon_map_data.py:8-62 and photon_map_index.py:17 (KDTree build). */
int alb_map_build(const alb_records_t* in, int64_t n, int32_t dims,
                  alb_map_t* map, void* workspace, size_t workspace_bytes,
                  void* stream);

/* alb_map_build for the records of many scenes (in->scene required): one
 * sort by (scene, Morton key) and one tree, in which every scene is a subtree
 * (index receives the per-scene record ranges and roots).  Same workspace
 * size as alb_map_build.  No counterpart in the reference, which builds one
 * KDTree per scene and type (photon_mapping/photon_map_index.py:17). */
int alb_map_build_batched(const alb_records_t* in, int64_t n, int32_t dims, alb_map_t* map,
                          alb_scene_index_t* index, void* workspace, size_t workspace_bytes,
                          void* stream);

/* Workspace bytes alb_gather needs for q_n queries (query sort buffers plus
 * the neighbour rows of at most 2^22 queries, reused round by round). */
int alb_gather_workspace_size(int64_t q_n, int32_t k, size_t* bytes);

/* k-nearest-neighbour gather of queries [q_begin, q_begin + q_n) of one photon
 * type against its map, accumulating FP64 energy into waveform[n_bins]
 * (n_bins = 2*steps).  Bins [win0, win0 + win_len) are privatised in LDS.
 * nbr_out (optional, [q_n * k] int32) receives each query's neighbour set as
 * sorted-map indices in (distance, index) order and radius_out ([q_n]
 * double) the k-th neighbour distance.  Queries are processed in Morton order
 * of their positions.  Same result as alb_gather_order + alb_knn +
 * alb_contrib.  Replaces core/simulation.py:132-195 (sample_photons). */
int alb_gather(const alb_scene_t* scene, const alb_map_t* map,
               int32_t photon_type, const alb_queries_t* q, int64_t q_begin,
               int64_t q_n, double* waveform, int64_t n_bins, int32_t win0,
               int32_t win_len, int32_t* nbr_out, double* radius_out,
               void* workspace, size_t workspace_bytes,
               unsigned long long* counters, void* stream);

/* alb_gather without a workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as alb_gather; the query sort buffers, neighbour rows (in rounds of
 * 2^14 queries when nbr_out is NULL), kernel normalisations and k-NN
 * counters live in a static device array of ALB_GATHER_STATIC_SCRATCH_BYTES
 * that each GPU the library runs on holds until the process ends.
 * q_n <= ALB_GATHER_STATIC_MAX_QUERIES and k <= 448, within the array
 * (at k = 448 about 2^19 queries); larger calls fail with ALB_ERR_ARG,
 * ALB_ERR_WORKSPACE or ALB_ERR_UNSUPPORTED.  Every call on one GPU shares its array: calls that can
 * overlap on one GPU (other streams, other host threads) race and must be
 * serialised by the caller; calls in sequence on one stream are safe. */
#define ALB_GATHER_STATIC_MAX_QUERIES (1 << 20)
#define ALB_GATHER_STATIC_SCRATCH_BYTES (64u << 20)
int alb_gather_static(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                      const alb_queries_t* q, int64_t q_begin, int64_t q_n, double* waveform,
                      int64_t n_bins, int32_t win0, int32_t win_len, int32_t* nbr_out,
                      double* radius_out, unsigned long long* counters, void* stream);

/* The three stages of alb_gather, for callers that time or overlap them.
 * alb_gather_order: order[q_n] (int32) = query offsets sorted by the Morton
 *   key of their positions on the map's grid; workspace from
 *   alb_gather_workspace_size.
 * alb_knn: the k nearest records of the queries in sort slots [c0, c0 + cn),
 *   cn <= 2^30 (larger calls fail with ALB_ERR_ARG; alb_knn_batched alike):
 *   row r (= slot - c0, or the query offset when rows_by_query != 0) of
 *   nbr[rows * k] gets the neighbour indices, radius[r] the k-th distance and
 *   knorm[r] the kernel normalisation 1/(pi r^2), or 1/(4/3 pi r^3) for
 *   photon_type ALB_SCATTER.  Neighbours are in (distance, index) order when
 *   rows_by_query != 0, else in no particular order.
 *   work: device unsigned long long[ALB_KNN_WORK_WORDS] of scratch (the
 *   counters from which the wavefronts take queries); the call clears it on
 *   the stream, so calls in sequence on one stream may share it and calls
 *   that can overlap need their own.
 * alb_contrib: adds the energies of those rows to the waveform.  Bins outside
 *   the LDS window go through n_replicas copies of the waveform (replicas:
 *   device double[n_replicas * n_bins], contents need not be zero), so that
 *   concurrent FP64 atomics on one bin are spread over that many addresses;
 *   the copies are summed into waveform before the call returns (on the
 *   stream).  n_replicas = 0 adds straight into waveform. */
int alb_gather_order(const alb_map_t* map, const alb_queries_t* q, int64_t q_begin, int64_t q_n,
                     int32_t* order, void* workspace, size_t workspace_bytes, void* stream);
#define ALB_KNN_WORK_WORDS 256
/* Largest k that alb_knn, alb_knn_batched and alb_gather take (the
 * reference takes any photon_mapping_k, config/simulation.py:23-29): the
 * candidates live in a per-wavefront LDS buffer of 256, 512 or 1024 entries
 * (k <= 192, <= 448, <= 960), and a buffer cut must free 64 of them.  Larger
 * k fails with ALB_ERR_UNSUPPORTED; alb_gather_static takes k <= 448. */
#define ALB_MAX_K 960
int alb_knn(const alb_map_t* map, int32_t photon_type, int32_t k, const alb_queries_t* q,
            int64_t q_begin, const int32_t* order, int64_t c0, int64_t cn, int32_t rows_by_query,
            int32_t* nbr, double* radius, double* knorm, unsigned long long* counters,
            unsigned long long* work, void* stream);
int alb_contrib(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                const alb_queries_t* q, int64_t q_begin, const int32_t* order, int64_t c0, int64_t cn,
                int32_t rows_by_query, const int32_t* nbr, const double* knorm, double* waveform,
                int64_t n_bins, int32_t win0, int32_t win_len, double* replicas,
                int32_t n_replicas, void* stream);

/* Fixed-radius gather: for queries q_begin + [0, q_n) (in the slot order of
 * order[], e.g. from alb_gather_order, or in index order when order is NULL),
 * every record of the built map within distance radius (FP64 squared distance
 * <= radius^2, the brute-force set exactly) adds
 * E_q E_p mult(type) / (pi r^2)  [1 / ((4/3) pi r^3) for ALB_SCATTER]
 * into waveform[(int)((t_q + t_p) / sample_multiplier)], with the multipliers
 * of alb_gather.  The reference's gather is the k-NN form
 * (src/alb_sim/core/simulation.py:132-195); this is its fixed-radius variant
 * for footprint-grid queries.  count_out / index_sum_out (optional, [q_n],
 * indexed by query offset): records taken and the sum of their map indices.
 * counters[ALB_CTR_OVERFLOW] > 0 means a query's traversal stack overflowed
 * and its sum is incomplete.  Needs no workspace. */
int alb_radius_gather(const alb_scene_t* scene, const alb_map_t* map, int32_t photon_type,
                      const alb_queries_t* q, int64_t q_begin, int64_t q_n, const int32_t* order,
                      double radius, double* waveform, int64_t n_bins, int32_t win0, int32_t win_len,
                      int64_t* count_out, int64_t* index_sum_out, unsigned long long* counters,
                      void* stream);

/* Batched gather over a map built by alb_map_build_batched.  Each query
 * (q->scene required) is answered from its own scene's records only, so scene
 * s gets exactly the neighbour sets and waveform a one-scene gather of its
 * records would give.  alb_gather_order_batched sorts queries by (scene,
 * Morton key); alb_knn_batched writes rows in slot order (rows of scenes with
 * fewer than k records are -1 / NaN and counted in ALB_CTR_SMALL_SCENE; work
 * as for alb_knn);
 * alb_contrib_batched reads the scene table scenes[n_scenes] (device) and adds
 * scene s's type-t waveform into waveforms[wf_offset[s] + t * 2 * steps_s ...],
 * LDS-privatising bins [win0[s], win0[s] + win_len) (device arrays); other
 * bins go through n_replicas copies (replicas: device double[n_replicas *
 * n_scenes * nb_max], nb_max >= every scene's 2 * steps), summed into the
 * waveforms before the call returns (n_replicas = 0: straight atomics).  No
 * counterpart in the reference (core/simulation.py:132-195 gathers one scene
 * per Simulation). */
int alb_gather_order_batched(const alb_map_t* map, const alb_scene_index_t* index,
                             const alb_queries_t* q, int64_t q_n, int32_t* order,
                             void* workspace, size_t workspace_bytes, void* stream);
int alb_knn_batched(const alb_map_t* map, const alb_scene_index_t* index, int32_t photon_type,
                    int32_t k, const alb_queries_t* q, const int32_t* order, int64_t c0,
                    int64_t cn, int32_t* nbr, double* radius, double* knorm,
                    unsigned long long* counters, unsigned long long* work, void* stream);
int alb_contrib_batched(const alb_scene_t* scenes, int32_t n_scenes, int32_t k,
                        const alb_map_t* map, int32_t photon_type, const alb_queries_t* q,
                        const int32_t* order, int64_t c0, int64_t cn, const int32_t* nbr,
                        const double* knorm, double* waveforms, const int64_t* wf_offset,
                        const int32_t* win0, int32_t win_len, double* replicas,
                        int32_t n_replicas, int64_t nb_max, void* stream);


/* ------------------------------------------------------------------ adjoint
 * Waveform-loss gradients by path replay (Vicini, Speierer and Jakob, "Path
 * replay backpropagation", SIGGRAPH 2021) with detached sampling: the paths
 * of the forward and backward traces are held fixed and every gather
 * contribution c = E_q E_p K mult is differentiated through its log-weight,
 * d log c = d log mult + S(path of p) + S(path of q), where for a path prefix
 * with n scatters already weighted by omega0, a water length L whose optical
 * depth was accumulated, m floor reflections and phase-sampling score sum G,
 *   S = (n + [record is a scatter]) dlog(alpha_L) - L d(alpha_L)
 *       + n dlog(omega0) + m dlog(albedo) + G dg.
 * No counterpart in the reference (NumPy forward only). */

/* Adjoint of alb_contrib_batched for one photon type: with G = dloss (same
 * layout as the waveforms) and w = G[bin] * c for every (query, neighbour)
 * contribution c, adds w to a_record[map.perm[neighbour]] and to
 * a_query[query] (input slots of the forward / backward sinks), and per scene
 * to direct[3 s + 0] the sum of w d log(phase_mass)/dg (scatter map),
 * direct[3 s + 1] the sum of w d log(BRDF)/d(alpha^2) (surface reflection) and
 * direct[3 s + 2] the sum of w over bottom contributions (d log / d albedo =
 * 1 / albedo).  Same slot range, rows and scene table as alb_contrib_batched. */
int alb_adjoint_contrib(const alb_scene_t* scenes, int32_t n_scenes, int32_t k,
                        const alb_map_t* map, int32_t photon_type, const alb_queries_t* q,
                        const int32_t* order, int64_t c0, int64_t cn, const int32_t* nbr,
                        const double* knorm, const double* dloss, const int64_t* wf_offset,
                        double* a_record, double* a_query, double* direct, void* stream);

/* Index of one sink for the replay: keys ((photon_offsets[scene] + photon)
 * << 16 | event ordinal), from the sink's tag and scene arrays, sorted with
 * their slots. */
typedef struct alb_replay_index {
  const uint64_t* keys;     /* [n] sorted                                          */
  const int32_t* slots;     /* [n] sink slot of each key                           */
  int64_t n;
  const double* weight;     /* [sink slots] adjoint weight of each record / query  */
} alb_replay_index_t;

int alb_replay_index_workspace_size(int64_t n, size_t* bytes);
int alb_replay_index_build(const int64_t* tags, const int32_t* scenes, int64_t n,
                           const int64_t* photon_offsets, uint64_t* keys_out, int32_t* slots_out,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Replay the photons of alb_trace_batched (same scenes, offsets, rng, pass)
 * with the path statistics above, look every record / query up in index[4]
 * and add, per scene s, acc[6 s + ...] += weight * (n, [scatter record], L,
 * m, G, d) of its prefix, where d counts the completed free paths ended by a
 * scatter (their density; n counts the omega0 factors).  Nothing is stored
 * per path.  stat_layer = -1 counts the whole water column; stat_layer = l
 * counts only layer l of a layered column of constant layers (L: the water
 * length in layer l; n: omega0 factors of layer l; d and [scatter record]:
 * free paths ended in a sub-step of layer l; m and G: whole path), which
 * gives the gradients of each layer's absorption and scattering (a single
 * constant layer is layer 0).  work_counter as in alb_trace_batched;
 * counters[ALB_CTR_REPLAY_MISSES] counts events whose record was not in the
 * index (0 when the sinks did not overflow). */
int alb_replay_batched(const alb_scene_t* scenes, int32_t n_scenes, const int64_t* photon_offsets,
                       const alb_rng_t* rng, int32_t pass, const alb_replay_index_t* index,
                       unsigned long long* work_counter, int32_t block, int32_t refill_interval,
                       double* acc, unsigned long long* counters, int32_t stat_layer, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ALB_ABI_H */
