/*
 * tss.h -- C ABI of the MI355X two-scale WCSPH solver (libtss.so).
 *
 * This is the drop-in boundary for the reference's solver path
 * (segfault11/TwoScaleSplit, TwoScale2D/WCSPHSolver.{h,cu} and
 * TwoScale2D/ParticleSystem.{h,cpp}; "R/" below = TwoScale2D/). The C++ class
 * API of the reference (WCSPHConfig, WCSPHSolver, ParticleSystem) is kept by
 * the header-only facade twoscalesplit_amd/host/WCSPHSolver.h on top of these
 * entry points.
 *
 * Conventions
 *  - Every call returns 0 on success and a negative TSS_E_* code on failure;
 *    tss_last_error() then returns a thread-local message (the reference
 *    throws std::runtime_error from UTIL::ThrowException, R/stdutil.h:22-33,
 *    and CUDA_SAFE_CALL, R/CUDA/cuda.h:37-51; the facade rethrows).
 *  - Handles are independent: any number of solvers and batches may coexist
 *    (the reference allows one per process, R/WCSPHSolver.cu:23-48,2728). A
 *    handle is not thread-safe; the caller serialises calls on it.
 *  - The library owns all device memory behind a handle. tss_read/tss_write
 *    copy between host memory and the device state and synchronise the
 *    handle's stream; tss_step does not synchronise.
 *  - Positions, velocities and accelerations are interleaved (x, y) float
 *    pairs indexed by particle id, like the reference's float2 arrays
 *    (R/ParticleSystem.h:56-77). High-res child k of low particle p has id
 *    4p+k (R/WCSPHSolver.cu:971-995).
 */
#ifndef TSS_H
#define TSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSS_OK 0
#define TSS_E_INVALID -1   /* bad argument / configuration          */
#define TSS_E_HIP -2       /* HIP runtime error                     */
#define TSS_E_NOMEM -3     /* device allocation or budget exhausted */
#define TSS_E_STATE -4     /* call not valid in the handle's state  */

/* Two-scale adaptivity modes. */
#define TSS_MODE_PLAIN 0     /* single level: no high-res particles          */
#define TSS_MODE_REFERENCE 1 /* the reference's live rule: split an active low
                                particle iff its predicted x > split_x;
                                children stay transient (R:1834-1863)        */
#define TSS_MODE_BAND 2      /* split inside the free-surface band, promote
                                relaxed groups, merge groups outside it       */

#define TSS_DT_FIXED 0
#define TSS_DT_CFL 1

/* Solver configuration. The first block is the reference's WCSPHConfig
 * (R/WCSPHSolver.h:14-35); the rest are the constants the reference
 * hard-codes (gravity R:1805, the split rule R:1835) and the two-scale and
 * time-step controls of the method spec. */
typedef struct tss_config {
    float domain_origin[2];
    float domain_end[2];
    int domain_dims[2];       /* set by tss_config_init: ceil(L/h) in FP32 */
    int domain_dims_high[2];  /* ceil(L/H)                                 */
    float effective_radius;   /* h: support radius of the low level        */
    float effective_radius_high; /* H                                       */
    float rest_density;
    float tait_coefficient;   /* B                                         */
    float speed_sound;
    float alpha;              /* artificial viscosity                      */
    float tension_coefficient;
    float time_step;          /* initial (or fixed) dt                     */
    float gravity;            /* 9.81                                      */
    float mass_low, mass_high, mass_boundary;
    int mode;                 /* TSS_MODE_*                                */
    float split_x;            /* reference rule threshold (0.5)            */
    float surface_ratio;      /* low surface iff rho < surface_ratio * rho0 */
    float surface_ratio_high; /* high: rho < surface_ratio_high * rho0;
                                 default surface_ratio * kappa(H/h), kappa =
                                 bulk value of the high-level density, whose
                                 blended kernel is cut at r <= H (R:439-474) */
    float band_radius;        /* split band distance (<= h)                */
    float merge_radius;       /* merge when all children are farther       */
    int promote_min_age, promote_max_age;
    float promote_fluc;
    int cap_high;             /* max live high particles (<= 4*cap_low);
                                 0 = no cap (all 4*cap_low child slots)    */
    int dt_mode;              /* TSS_DT_*                                  */
    float cfl, dt_max;
    /* pressure-only relaxation of the children injected by the previous
     * step's splits, run after the grid build: relax_iters Jacobi
     * iterations (0 = off) of density + pressure/wall acceleration +
     * symplectic Euler at relax_dt_frac * dt (the dormant sub-steps of
     * TwoScale2D-master/TwoScale2D/WCSPHSolver.cu:2380-2446 and
     * R/WCSPHSolver.cu:3044-3110, there 4 iterations at dt/10) */
    int relax_iters;
    float relax_dt_frac;
} tss_config;

typedef struct tss_solver tss_solver;
typedef struct tss_batch tss_batch;

/* Fields for tss_read / tss_write / tss_device_ptr (per scene). */
enum tss_field {
    TSS_POS_L = 0,   /* float[2*cap_low]   by low id                     */
    TSS_VEL_L,       /* float[2*cap_low]                                 */
    TSS_ACC_L,       /* float[2*cap_low]   acceleration of the last step */
    TSS_RHO_L,       /* float[cap_low]     density of the last step      */
    TSS_P_L,         /* float[cap_low]     pressure of the last step     */
    TSS_STATE_L,     /* uint8[cap_low]     0 active 1 transient 2 deleted */
    TSS_AGE_L,       /* int32[cap_low]     steps since split             */
    TSS_VIS_L,       /* float[cap_low]     min(|rho-rho0|,400)/400       */
    TSS_POS_H,       /* float[8*cap_low]   by high id                    */
    TSS_VEL_H,
    TSS_ACC_H,
    TSS_RHO_H,       /* float[4*cap_low]                                 */
    TSS_P_H,
    TSS_STATE_H,     /* uint8[4*cap_low]   1 transient child 0 promoted  */
    TSS_VIS_H,
    TSS_LIST_L,      /* int32[n_low]   active low ids (next step's list)  */
    TSS_LIST_H,      /* int32[n_high]                                    */
    TSS_LIST_G,      /* int32[n_groups] parents that own children        */
    TSS_SORTED_L,    /* int32[n_low_last]  last step's hash-sorted ids    */
    TSS_SORTED_H,    /* int32[n_high_last]                               */
    TSS_KEYS_L,      /* int32[n_low_last]  sorted cell hashes            */
    TSS_KEYS_H,
    TSS_CELL_START_L,/* int32[cells_low]   -1 for empty cells            */
    TSS_CELL_END_L,
    TSS_CELL_START_H,/* int32[cells_high]                                */
    TSS_CELL_END_H,
    TSS_BND_SORTED,  /* int32[n_bnd] boundary ids in hash order          */
    TSS_BND_START,   /* int32[cells_low]                                 */
    TSS_BND_END,
    TSS_POS_B,       /* float[2*n_bnd] boundary positions by id          */
    TSS_DT,          /* float[1] dt of the next step                     */
    TSS_FIELD_COUNT
};

/* Per-scene counters (tss_counts / tss_batch_counts). */
typedef struct tss_counts_t {
    int n_low, n_high, n_groups;          /* lists of the next step          */
    int n_low_last, n_high_last;          /* lists the last step advanced    */
    int n_split, n_promote, n_merge;      /* events of the last step         */
    int n_dissolve;                       /* transient groups removed        */
    int step;                             /* completed steps                 */
    int cap_low, n_bnd;
    int cells_low, cells_high;
    long long particle_steps;             /* low + high particles advanced,
                                             summed over completed steps    */
} tss_counts_t;

/* ---- configuration ---------------------------------------------------- */
/* Replaces WCSPHConfig::WCSPHConfig (R/WCSPHSolver.cu:2287-2332): fills the
 * WCSPHConfig fields, derives the grid dimensions (ceil of FP32 quotient)
 * and sets the remaining fields to defaults (gravity 9.81, split_x 0.5,
 * surface_ratio 0.9, surface_ratio_high 0.9 kappa(H/h), band/merge radius
 * h, plain mode, fixed dt, cap_high 0 = no cap).
 * TSS_E_INVALID if xs >= xe or ys >= ye, like the reference's exception. */
int tss_config_init(tss_config* cfg, float xs, float ys, float xe, float ye,
                    float effective_radius, float effective_radius_high,
                    float rest_density, float tait_coefficient,
                    float speed_sound, float alpha,
                    float tension_coefficient, float time_step);

/* ---- single-scene solver ------------------------------------------------ */
/* Replaces WCSPHSolver::WCSPHSolver + Bind (R/WCSPHSolver.cu:2391-2520,
 * 2537-2639) and the particle uploads of CreateParticleBox /
 * CreateParticleBoxCanvas (R/ParticleSystem.cpp:123-179): n_low fluid
 * particles (low ids 0..n_low-1, capacity cap_low >= n_low) and n_bnd
 * static boundary particles; builds the boundary grid once. */
int tss_create(const tss_config* cfg, const float* low_xy, int n_low,
               int cap_low, const float* bnd_xy, int n_bnd,
               tss_solver** out);
/* Replaces WCSPHSolver::AdvanceTS (R/WCSPHSolver.cu:2726-2810): n_steps
 * steps, enqueued on the handle's stream, no host synchronisation. */
int tss_step(tss_solver* s, int n_steps);
/* Same n_steps through a captured hipGraph of one step (replayed). */
int tss_step_graph(tss_solver* s, int n_steps);
/* Counters (synchronising). Replaces ParticleSystem::GetNumParticles /
 * GetMaxParticles (R/ParticleSystem.h:35-37). */
int tss_counts(tss_solver* s, tss_counts_t* out);
/* Synchronising read-back / upload of a field (see enum tss_field).
 * Replaces the Positions/Velocities/Densities/Pressures/VisualQuantities/
 * ParticleIDs accessors (R/ParticleSystem.h:27-32) + Map/Unmap
 * (R/ParticleSystem.cpp:78-98). bytes must not exceed the field's size. */
int tss_read(tss_solver* s, int field, void* dst, size_t bytes);
int tss_write(tss_solver* s, int field, const void* src, size_t bytes);
/* Sets the list counts (n_low, n_high, n_groups) after tss_write of lists:
 * re-seeding a state for tests and restart. Lists, counts and states written
 * from the host are checked before the next call that launches on them
 * (tss_step, tss_step_graph, tss_neighbors, tss_pair_sets,
 * tss_work_counts): ids in range without repeats, no deleted particle
 * listed, no active group parent, no child of an unlisted group (the
 * invariants that bound list growth); otherwise that call returns
 * TSS_E_INVALID and launches nothing. */
int tss_set_counts(tss_solver* s, int n_low, int n_high, int n_groups);
/* Device pointer of a field (valid until tss_destroy), replacing GL
 * interop for device-side consumers. */
int tss_device_ptr(tss_solver* s, int field, void** out);
int tss_synchronize(tss_solver* s);
/* FP64 conservation sums of the last step (mass, px, py, kinetic,
 * potential) and a rigorous bound of each sum's rounding error. */
int tss_conservation(tss_solver* s, double out[5], double err_bound[5]);
/* Neighbour sets of the current state: per listed particle of a level (in
 * list order), 5 categories of counts and order-independent id checksums
 * (see oracle's orc_neighbors for the categories). Builds the grid of the
 * current lists; does not advance. counts/sums hold 5*n entries. */
int tss_neighbors(tss_solver* s, int level, uint32_t* counts, uint32_t* sums,
                  size_t n);
/* The pairs the step's own density and force kernels add, in the categories
 * of tss_neighbors restricted to nonzero state weights: [0] same-level
 * density, [1] cross-level density, [2] same-level force, [3] cross-level
 * force, [4] wall candidates of the force. The kernels run on the current
 * state in a recording variant -- the same candidate masks, scans and row
 * and pass skips as a step -- after a grid build of the current lists; no
 * state is advanced (relaxation sub-steps are not run). These are the sets
 * of updateDensityCell* (R/WCSPHSolver.cu:335-474) and updateAccCell*
 * (R/WCSPHSolver.cu:476-885) with their lambda tables; no counterpart
 * entry point in the reference. */
int tss_pair_sets(tss_solver* s, int level, uint32_t* counts, uint32_t* sums,
                  size_t n);
/* Device time per phase (ms), averaged over every step taken (eager or
 * graph-replayed) since timing was switched on, from hipEvents recorded on
 * the solver's stream around each phase without a host wait per step:
 * [0] grid build (hash, radix sort, cell ranges, reorder; and the
 * relaxation of new children when a scene enables it) [1] density, low
 * targets [2] density, high targets [3] force + integration, low targets
 * [4] force + integration, high targets [5] FP64 reductions [6] adaptivity
 * and list compaction [7] finalize (counts, dt). tss_enable_timing resets
 * the averages. No counterpart in the reference (it times nothing). */
#define TSS_PHASE_COUNT 8
int tss_enable_timing(tss_solver* s, int on);
int tss_phase_times(tss_solver* s, float out_ms[TSS_PHASE_COUNT]);
/* Exact work of the density and force kernels on the current state (grid
 * rebuilt from the current lists), per target level: candidates [0] low
 * same-level [1] low cross-level [2] high same-level [3] high cross-level
 * (the force's radius-h box); density pairs [4] low [5] high; force pairs
 * [6] low [7] high; wall candidates [8] low [9] high; targets [10] low [11]
 * high; [12] high cross-level candidates of the density (its radius-H box).
 * Multiplied by fixed per-item flop weights these give each kernel's
 * algorithmic flops. */
#define TSS_WORK_COUNT 13
int tss_work_counts(tss_solver* s, unsigned long long out[TSS_WORK_COUNT]);
/* Launch shapes of the four neighbour kernels, index order: [0] density low
 * targets, [1] density high, [2] force low, [3] force high. block: 128, 256
 * or 512 threads; depth: LDS buffer depth, 8, 16 or 32 pairs per lane;
 * xcd_order: 1 gives each of the 8 XCDs one contiguous range of blocks;
 * grid: 0 = one thread per target of capacity, else a persistent grid of at
 * most this many blocks. Results never depend on the shapes (every
 * neighbour sum is lane-local). Default: 256, 16, 0, 0. No counterpart in
 * the reference (fixed 256-thread blocks, R/WCSPHSolver.cu:2421). */
typedef struct tss_tuning {
    int block[4];
    int depth[4];
    int xcd_order[4];
    int grid[4];
} tss_tuning;
/* Sort of the grid build (R/WCSPHSolver.cu:2828-2834, thrust::sort_by_key
 * by cell, stable): TSS_SORT_AUTO picks per batch shape; TSS_SORT_RADIX is
 * the 8-bit LSD radix sort (hist / scan / scatter per pass);
 * TSS_SORT_COUNT the counting sort by cell (counts, prefix over cells,
 * scatter, rank within each cell by list position); TSS_SORT_SMALL the
 * one-workgroup radix sort, for resident capacities of at most 13107 low
 * ids (65536 keys; TSS_E_INVALID otherwise). AUTO takes COUNT when the
 * resident cells number at most 4 per key of capacity, else SMALL where it
 * fits, else RADIX. Every path gives the same order, cell ranges and
 * results. */
#define TSS_SORT_AUTO 0
#define TSS_SORT_RADIX 1
#define TSS_SORT_COUNT 2
#define TSS_SORT_SMALL 3
int tss_set_grid_sort(tss_solver* s, int path);
int tss_batch_set_grid_sort(tss_batch* b, int path);
/* Launch shapes and autotuner of a single-scene solver (as for batches,
 * tss_batch_autotune below). */
int tss_get_tuning(tss_solver* s, tss_tuning* out);
int tss_set_tuning(tss_solver* s, const tss_tuning* t);
int tss_autotune(tss_solver* s, int steps_per_trial, int rounds,
                 tss_tuning* out, int* steps_taken, float* trial_log,
                 int log_rows, int* n_log);
/* Replaces WCSPHSolver::~WCSPHSolver (R/WCSPHSolver.cu:2522-2535). */
void tss_destroy(tss_solver* s);

/* ---- batched scenes ----------------------------------------------------- */
/* Many independent scenes advanced by one launch per phase. Capacities are
 * totals over resident scenes; tss_batch_memory_bytes gives the device
 * bytes such a batch holds. No counterpart in the reference, which allows
 * one solver per process (R/WCSPHSolver.cu:23-48,2728). */
typedef struct tss_batch_desc {
    int max_scenes;        /* 1 .. 65535                                */
    long long cap_low;     /* total low-particle capacity (children 4x) */
    long long cap_bnd;     /* total boundary particles                  */
    long long cells_low;   /* total low-grid cells                      */
    long long cells_high;  /* total high-grid cells                     */
} tss_batch_desc;

int tss_batch_memory_bytes(const tss_batch_desc* d, long long* bytes);
int tss_batch_create(const tss_batch_desc* d, tss_batch** out);
/* Admits a scene (same arguments as tss_create) between steps. Returns
 * TSS_E_NOMEM when the batch's capacities cannot hold it. */
int tss_batch_add_scene(tss_batch* b, const tss_config* cfg,
                        const float* low_xy, int n_low, int cap_low,
                        const float* bnd_xy, int n_bnd, int* scene_id);
/* Retires a scene; its slots are reusable by the next admission. */
int tss_batch_remove_scene(tss_batch* b, int scene_id);
int tss_batch_step(tss_batch* b, int n_steps);
int tss_batch_step_graph(tss_batch* b, int n_steps);
int tss_batch_counts(tss_batch* b, int scene_id, tss_counts_t* out);
int tss_batch_read(tss_batch* b, int scene_id, int field, void* dst,
                   size_t bytes);
int tss_batch_conservation(tss_batch* b, int scene_id, double out[5],
                           double err_bound[5]);
int tss_batch_synchronize(tss_batch* b);
int tss_batch_num_scenes(tss_batch* b, int* n);
int tss_batch_get_tuning(tss_batch* b, tss_tuning* out);
int tss_batch_set_tuning(tss_batch* b, const tss_tuning* t);
/* Autotuner: times candidate shapes with hipEvents while advancing the
 * batch by real steps (they count, the results being shape-independent).
 * After one ordinary step, stage 1 tries every block x depth under the XCD
 * order, stage 2 round-robin order and persistent grid caps (4 and 8
 * blocks per CU) on each kernel's stage-1 winner; each candidate runs
 * `rounds` times for
 * steps_per_trial steps, in alternating order; a candidate's time per
 * target of the kernel's level is its fastest trial, and per kernel the
 * fastest candidate is kept (and returned in out).
 * steps_taken: steps advanced. trial_log (optional): log_rows x 8 floats per
 * trial, (block, depth, xcd_order, grid, ms of kernels 0..3); n_log rows. */
int tss_batch_autotune(tss_batch* b, int steps_per_trial, int rounds,
                       tss_tuning* out, int* steps_taken, float* trial_log,
                       int log_rows, int* n_log);
void tss_batch_destroy(tss_batch* b);

/* Scene generated on the device: the seeded dam break of
 * tss_scene_dam_break below (fluid positions bit for bit identical), with
 * the fluid block written by a kernel instead of copied from the host; the
 * wall canvas (perimeter-sized) is built on the host. Replaces
 * CreateParticleBox / CreateParticleBoxCanvas (R/ParticleSystem.cpp:123-179),
 * which upload one particle per copy (R/ParticleSystem.cpp:73-74). */
typedef struct tss_dam_break {
    uint64_t seed, scene;
    int nx, ny;
    float dx, jitter;
    int nxd, nyd, layers;
} tss_dam_break;
int tss_batch_add_dam_break(tss_batch* b, const tss_config* cfg,
                            const tss_dam_break* gen, int cap_low,
                            int* scene_id);

/* ---- headless frame export -------------------------------------------------- */
/* One frame of a scene: the particles of its current lists -- exactly the
 * points the reference draws (glDrawElements over the ParticleIDs index
 * buffer, R/QuantityRenderer.cpp:84-95, with the positions and visual
 * quantities of R/ParticleSystem.cpp:21-46 mapped into CUDA by Map/Unmap,
 * R/ParticleSystem.cpp:78-98) -- gathered on the device in list order and
 * copied out in one transfer, replacing the GL interop with a buffer that a
 * file, a pipe or an offline renderer can take.
 *
 * Layout: a 64-byte tss_frame_header, then the low section, then the high
 * section, then (TSS_FRAME_BND) the boundary positions by id. A level's
 * section holds one array per requested field, in bit order
 * (ID, POS, VEL, RHO, P, VIS, STATE), each n entries long and padded with
 * zero bytes to a multiple of 16 bytes. IDs are scene-local (low id, or high
 * id = 4 * parent + k); STATE is one byte per particle. Frames are
 * self-describing, so a stream of frames is their concatenation. */
enum tss_frame_bits {
    TSS_FRAME_ID = 1,          /* int32  */
    TSS_FRAME_POS = 2,         /* float2 */
    TSS_FRAME_VEL = 4,         /* float2 */
    TSS_FRAME_RHO = 8,         /* float  */
    TSS_FRAME_P = 16,          /* float  */
    TSS_FRAME_VIS = 32,        /* float: min(|rho - rho0|, 400) / 400 */
    TSS_FRAME_STATE = 64,      /* uint8  */
    TSS_FRAME_BND = 128,       /* boundary positions (float2 by id) */
    TSS_FRAME_ALL = 255
};
#define TSS_FRAME_MAGIC 0x46535354u  /* bytes "TSSF" */
#define TSS_FRAME_VERSION 1u
typedef struct tss_frame_header {
    uint32_t magic;            /* TSS_FRAME_MAGIC */
    uint32_t version;          /* TSS_FRAME_VERSION */
    uint32_t fields;           /* tss_frame_bits of the sections present */
    int32_t step;              /* completed steps of the scene */
    int32_t n_low, n_high;     /* entries of the low and high sections */
    int32_t n_bnd;             /* boundary entries (0 without TSS_FRAME_BND) */
    float dt;                  /* dt of the next step */
    float domain_origin[2];
    float domain_end[2];
    uint32_t reserved[4];      /* zero */
} tss_frame_header;
/* Bytes of a frame with these fields and counts (0 for invalid arguments). */
size_t tss_frame_size(uint32_t fields, int n_low, int n_high, int n_bnd);
/* Writes the current frame of the scene into dst (synchronising; between
 * steps). *bytes receives the frame's size; with dst == NULL nothing else
 * happens (a size query). TSS_E_INVALID if cap is smaller than the frame. */
int tss_frame(tss_solver* s, uint32_t fields, void* dst, size_t cap,
              size_t* bytes);
int tss_batch_frame(tss_batch* b, int scene_id, uint32_t fields, void* dst,
                    size_t cap, size_t* bytes);
/* tss_batch_frame without any allocation: the frame is staged in a fixed
 * device array of the library (64 MiB, one per device, held until the
 * process ends) instead of a buffer grown with hipMalloc. No counterpart in
 * the reference; it is for callers that cannot allocate device memory --
 * tight serving loops and harnesses that forbid hipMalloc. A frame whose
 * staging (its sections plus 4 bytes per 1024 list entries of the batch)
 * exceeds 64 MiB is refused with TSS_E_INVALID. Every call on a device
 * shares that device's array: calls that overlap on one GPU (from several
 * threads or handles) race and must be serialised by the caller. */
int tss_batch_frame_noalloc(tss_batch* b, int scene_id, uint32_t fields,
                            void* dst, size_t cap, size_t* bytes);

/* ---- scene snapshot / resume ---------------------------------------------- */
/* A self-contained image of one scene between steps: its configuration,
 * boundary, every per-particle array of its id ranges, its lists (in list
 * order, scene-local ids), counters, the dt of the next step and the last
 * conservation sums. Restoring it -- into a new solver, or into any batch
 * with room -- continues the run bit for bit as if it had not stopped
 * (results of a scene do not depend on where its ids or cells sit).
 * Replaces ParticleSystem::Save (R/ParticleSystem.cpp:100-107), which
 * writes only the particle count.
 *
 * Layout: tss_snapshot_header (padded to a multiple of 16 bytes), then
 * sections each padded to 16 bytes: boundary positions by id
 * (float2[n_bnd]); low arrays over cap_low ids: pos, vel, acc (float2), rho,
 * p, vis (float), state (uint8), age (int32); high arrays over 4*cap_low
 * ids: pos, vel, acc, rho, p, vis, state; lists: low (int32[n_low]), high
 * (int32[n_high]), groups (int32[n_groups]). */
#define TSS_SNAPSHOT_MAGIC 0x53535354u  /* bytes "TSSS" */
#define TSS_SNAPSHOT_VERSION 1u
typedef struct tss_snapshot_header {
    uint32_t magic, version;
    uint32_t header_bytes;        /* sizeof(tss_snapshot_header) */
    uint32_t config_bytes;        /* sizeof(tss_config) */
    uint64_t total_bytes;
    tss_config cfg;
    int32_t cap_low, n_bnd;
    int32_t n_low, n_high, n_groups;
    int32_t n_fresh;              /* trailing high-list entries injected by
                                     the last step (relaxed by the next) */
    int32_t step, n_low_last, n_high_last;
    int32_t n_split, n_promote, n_merge, n_dissolve, n_cand;
    int64_t particle_steps;
    float dt;                     /* dt of the next step */
    int32_t reserved0;
    double conservation[10];      /* last step's sums and error bounds */
} tss_snapshot_header;
/* Bytes of a snapshot with these sizes (0 for invalid arguments). */
size_t tss_snapshot_size(int cap_low, int n_bnd, int n_low, int n_high,
                         int n_groups);
/* Snapshot of a scene into dst (synchronising; between steps). *bytes gets
 * the snapshot's size; dst == NULL is a size query. TSS_E_INVALID if cap is
 * smaller. */
int tss_snapshot(tss_solver* s, void* dst, size_t cap, size_t* bytes);
int tss_batch_snapshot(tss_batch* b, int scene_id, void* dst, size_t cap,
                       size_t* bytes);
/* A new single-scene solver holding the snapshot's scene. */
int tss_create_from_snapshot(const void* src, size_t bytes, tss_solver** out);
/* Adds the snapshot's scene to a batch (TSS_E_NOMEM without room). The
 * snapshot is validated (sizes, ids, states, finite positions) first. */
int tss_batch_restore(tss_batch* b, const void* src, size_t bytes,
                      int* scene_id);

/* ---- serving loop ---------------------------------------------------------- */
/* A scene request: its initial state, the steps it asks for and its arrival
 * time; optional output buffers receive fields of its final state. No
 * counterpart in the reference (one scene per process, driven per frame by
 * R/main.cpp:70). */
typedef struct tss_request {
    const tss_config* cfg;
    const float* low_xy;
    int n_low, cap_low;
    const float* bnd_xy;
    int n_bnd;
    const tss_dam_break* gen;       /* or NULL; when set, the scene is
                                       generated on the device at admission
                                       and low_xy / bnd_xy are not read     */
    int steps;                      /* steps to advance (>= 1)            */
    double arrival_s;               /* seconds after the start of the run */
    void* out[TSS_FIELD_COUNT];     /* optional: field copies at finish   */
    size_t out_bytes[TSS_FIELD_COUNT];
    /* results */
    int status;                     /* TSS_OK or the error that rejected it */
    double admit_s, finish_s;       /* seconds after the start of the run */
    long long particle_steps;
} tss_request;

typedef struct tss_serve_stats {
    double wall_s;                  /* start of the run to the last finish */
    double busy_s;                  /* time with at least one scene resident */
    long long particle_steps;
    double particle_steps_per_s;    /* particle_steps / wall_s            */
    int ticks, graph_ticks;         /* batch advances; of them graph-replayed */
    int max_resident;
    double step_ms_p50, step_ms_p99;        /* per step, host wall clock  */
    double request_ms_p50, request_ms_p99;  /* arrival to finish          */
    double queue_ms_p50, queue_ms_p99;      /* arrival to admission       */
    long long batch_bytes;          /* device memory of the serving batch */
    long long budget_bytes;
    long long peak_admitted_bytes;  /* budget accounting of resident scenes */
} tss_serve_stats;

/* Continuous batching: runs the requests to completion on one batch of the
 * given capacities, in real time. Every tick it admits the requests that
 * have arrived, in arrival order, whenever the batch has room and the sum of
 * the resident scenes' device bytes stays within budget_bytes (a request
 * that does not fit waits; later ones may pass it); advances every resident
 * scene together (one launch per phase; while the resident set is
 * unchanged, hipGraph pairs of steps, about 2 ms of them per tick); and
 * retires each scene the step it completes
 * its steps, copying the requested fields out first. Results of a scene do
 * not depend on the others resident with it. Requests that can never fit
 * (or are invalid) get a nonzero status; the run continues. */
int tss_serve(const tss_batch_desc* capacity, long long budget_bytes,
              tss_request* requests, int n_requests, tss_serve_stats* stats);
/* Device bytes a scene accounts for in admission (its share of a batch). */
long long tss_scene_bytes(const tss_config* cfg, int cap_low, int n_bnd);

/* ---- scene generator (host) ----------------------------------------------- */
/* Seeded dam-break scene: fluid block of nx*ny particles on a lattice of
 * spacing dx at ((i+0.5)dx, (j+0.5)dx) in row order (j outer, like
 * CreateParticleBox, R/ParticleSystem.cpp:123-148), each jittered by
 * U(-jitter*dx, jitter*dx) from SplitMix64 keyed by (seed, scene, index);
 * plus `layers` wall layers around the domain [0, nxd*dx] x [0, nyd*dx]
 * (CreateParticleBoxCanvas, R/ParticleSystem.cpp:150-179). Arrays are
 * caller-owned: low_xy[2*nx*ny], bnd_xy[2*tss_canvas_count(...)]. */
int tss_canvas_count(int nxd, int nyd, int layers);
int tss_scene_dam_break(uint64_t seed, uint64_t scene, int nx, int ny,
                        float dx, float jitter, int nxd, int nyd, int layers,
                        float* low_xy, float* bnd_xy);
/* SplitMix64 output stream (for tests of the generator). */
uint64_t tss_splitmix64_next(uint64_t* state);

/* Streaming-copy bandwidth of the device (read + write bytes / time) over
 * `bytes`-sized buffers, `reps` copies: the plain-copy reference of the
 * roofline and of the HBM counter calibration. No counterpart in the
 * reference. */
int tss_bandwidth_probe(size_t bytes, int reps, double* gbs);
/* Sparse-read rate: one 8-byte read every stride_bytes over a fresh
 * `bytes`-sized buffer (each read its own request), reps passes; gbs = 8-byte
 * words read / time. With a buffer larger than the caches and a stride of a
 * cache line or more, every read fetches its own line: the reference of the
 * HBM counter calibration for gathered, smaller-than-line requests. No
 * counterpart in the reference. */
int tss_gather_probe(size_t bytes, int stride_bytes, int reps, double* gbs);

const char* tss_last_error(void);
const char* tss_version(void);

#ifdef __cplusplus
}
#endif
#endif
