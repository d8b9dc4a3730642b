// tss_device.h -- device-side data layout of a batch of two-scale WCSPH
// scenes and the launch interface of the HIP kernels (tss_kernels.hip).
//
// Layout (one "batch" = any number of independent scenes; a single-scene
// solver is a batch of one):
//  * Unified particle index space u: low ids [0, CL), high ids [CL, 5 CL);
//    child k of low id p is u = CL + 4p + k, so a scene that owns low ids
//    [lbase, lbase + lcap) owns the children [CL + 4 lbase, CL + 4(lbase+lcap)).
//  * Per-particle state is SoA by u: pos/vel/acc (float2), rho, p, vis
//    (float), state (u8); age and scene index by low id.
//  * Unified cell index space: low cells of all scenes first, then high
//    cells; a scene owns [cbase_l, cbase_l + dlx*dly) and
//    [cbase_h, cbase_h + dhx*dhy). The sort key of a particle is its unified
//    cell index, so one stable sort groups particles by level, scene, cell.
//  * After the sort, the step works on cell-sorted SoA copies (spos, svel,
//    srho, spr = p/rho^2, sstate, sscene) so every neighbour sum streams
//    contiguous rows of cells.
#pragma once

#include <cstdint>

namespace tss {

constexpr int kClearChunk = 4096;  // cells per block of the cell-table clear
constexpr int kRadixBits = 8;     // radix digit width
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = (32 + kRadixBits - 1) / kRadixBits;  // 32-bit keys
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per sort tile
// Batches whose resident scenes hold at most kSmallSortKeys keys (5 keys per
// low-id capacity) sort in one workgroup of kSmallSortThreads threads, every
// radix pass in one launch.
constexpr int kSmallSortThreads = 1024;
constexpr long long kSmallSortKeys = 65536;
// Counting sort by cell (the grid build when the resident cells number at
// most kCountCellsPerKey per key): cell counts are scanned in tiles of
// kCellTile cells (kCellThreads threads, kCellTile / kCellThreads cells
// each).
constexpr int kCellThreads = 256;
constexpr int kCellTile = 1024;
constexpr int kCountCellsPerKey = 4;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // entries per compaction tile
// FP64 reductions: a scene's sorted particles are cut into chunks of
// kRedChunk; a scene owns ceil(5 cap_low / kRedChunk) chunk slots (from its
// own capacity, so the partition -- and every rounding -- is the same alone
// and in a batch).
constexpr int kRedChunk = 4096;
// Spare entries past the capacity of the cell-sorted position and state
// arrays: the candidate fill reads up to 31 entries beyond a run's end.
constexpr int kSortedPad = 32;
// Bytes of padding past the last cell of the csurf and cband flag tables
// (read as aligned 4-byte words, up to 11 bytes past a row chunk's start).
constexpr int kFlagPad = 16;
// mass, px, py, ke, pe, max|v|^2, sum|px|, sum|py|, sum|pe|
constexpr int kRedFields = 9;

struct SceneDev {
    // geometry
    float ox, oy;             // domain origin
    float h, H;               // support radii (low / high)
    int dlx, dly, dhx, dhy;   // grid dimensions
    int cbase_l, cbase_h;     // unified cell bases
    int lbase, lcap;          // low ids owned
    int bbase, nb;            // boundary sorted range
    // physics
    float rho0, B, c, alpha, sigma, g;
    float m_l, m_h, m_b;
    float sig_h2, sig_H2;     // M4 coefficients / 2  (W_bar = sum of both)
    float gc_h2, gc_H2;       // gradient coefficients / 2
    float q_h, q_H;           // 2/h, 2/H
    float visc_h, visc_H, visc_x;  // 2*l*alpha*c for l = h, H, (h+H)/2
    float eps_h, eps_H, eps_x;     // 0.01*l^2
    float cb_l, cb_h;         // m_b/(m_l+m_b), m_b/(m_h+m_b)
    float cwall;              // 0.02*c^2
    // squared-distance thresholds equivalent to the FP32 sqrt predicates
    float t_le_h, t_lt_h, t_le_H, t_lt_H, t_band, t_merge;
    float r_band, r_merge;
    // adaptivity
    int mode;
    float split_x;
    float surf_t;             // surface_ratio * rho0 (FP32 product)
    float surf_t_h;           // surface_ratio_high * rho0
    int pmin, pmax;
    float pfluc;
    int cap_h;
    // time step
    int dt_mode;
    float cfl, dt_max;
    float dt;                 // dt of the next step (device-updated)
    int alive;
    int relax_iters;          // relaxation iterations of new children
    float relax_frac;         // relaxation step = relax_frac * dt
    int hh_inner;             // 1: 2H <= h, so r <= H gives q_h = 2r/h <= 1
    int hh_half;              // 1: H = h/2 exactly: sig_H2 = 4 sig_h2,
                              // gc_H2 = 8 gc_h2, q_H = 2 q_h in FP32
};

struct SceneCount {
    int n_l, n_h, n_g;        // next-step lists (per scene)
    int n_l_last, n_h_last;   // lists advanced by the last step
    int n_split, n_promote, n_merge, n_dissolve;
    int cand;                 // split candidates of the last step
    int step;
    int seg_l0, seg_l1, seg_h0, seg_h1;  // sorted segments of the last step
    // compaction prefixes at the scene's segment ends (kept low, kept high,
    // admitted splits): counts = end - begin, no hot atomics
    int kl0, kl1, kh0, kh1, sp0, sp1;
    int prom;                 // 1: a listed high particle is promoted (state 0)
    int act;                  // 1: a listed low particle is active (state 0)
    int hsurf;                // 1: a listed high particle is a free-surface particle
    unsigned long long psteps;  // particle-steps advanced since admission
};

struct DevCounts {
    int n_l, n_h, n_g;        // lists the next step reads
    int n_l_next, n_h_next, n_g_next;
    int n_l_last, n_h_last;   // lists the last step advanced (sorted sizes)
    // children injected by the last step's splits: the last n_fresh entries
    // of the high list (new children are appended after the kept ones)
    int n_fresh, n_fresh_next;
    // 1: the step has a split candidate, promotion, merge or dissolve, or
    // the host changed the state since the last step; 0: every listed
    // entry is kept as it is (the compaction needs no flags). Reset to
    // events_floor at the end of every step (1 once raw device pointers
    // were handed out: the state may change behind the library's back).
    int events, events_floor;
};

// Device buffers of a batch (all pointers are device memory).
// 32-bit words of a target's candidate-mask record (up to 256 candidates of
// its same-level scan rows; targets with more take the scan path).
constexpr int kMaskWords = 8;

struct DevState {
    int CL;                   // low capacity (unified index space size 5 CL)
    int NB;                   // boundary capacity
    int CCL, CCH;             // low / high cell capacity
    int S;                    // scene capacity
    SceneDev* scenes;
    SceneCount* scnt;
    DevCounts* counts;
    int* scene_of;            // [CL] scene index of each low id (-1 free)
    float2* pos;
    float2* vel;
    float2* acc;              // [5 CL]
    float* rho;
    float* p;
    float* vis;               // [5 CL]
    uint8_t* state;           // [5 CL]
    int* age;                 // [CL]
    int* list_l[2];           // [CL]   unified ids
    int* list_h[2];           // [4 CL]
    int* list_g[2];           // [CL]   low ids of parents
    uint32_t* keys[2];        // [5 CL] sort ping-pong
    uint32_t* vals[2];
    uint32_t* hist;           // [256 * max_tiles]
    uint32_t* digit_tot;      // [kRadix] digit totals of the current pass
    int* cstart;
    int* cend;                // [CCL + CCH]
    int* cscene;              // [CCL + CCH] scene owning each cell (-1: none)
    int* ccount;              // [CCL + CCH] keys per cell (counting sort;
                              // zero between grid builds)
    int* ctile;               // [CCL + CCH / kCellTile + 1] cell tile sums
    uint2* cslot;             // [5 CL] counting sort: (list position, id)
                              // per slot of a cell's range
    // [cells / kClearChunk + 2 S + 1] cell ranges [x, y) of the resident
    // scenes, at most kClearChunk cells each: one block of the step's
    // cell-table clear per entry
    int2* clear_blk;
    int* bstart;
    int* bend;                // [CCL]
    // [CCL] 1: a boundary particle in the low cells x..x+3, y..y+3 of this
    // cell (x, y); a box of at most 4x4 low cells starting at a cell
    // without the flag has no wall particle
    uint8_t* bnear;
    float2* bpos;             // [NB] boundary positions in cell order
    int* bids;                // [NB] boundary ids (scene-local) in cell order
    float2* spos;             // [5 CL + kSortedPad]
    float4* sxv;              // [5 CL] cell-sorted (x, y, vx, vy)
    float4* snew;             // [5 CL] cell-sorted state after the step
    float2* srp;              // [5 CL] cell-sorted (rho, p / rho^2)
    uint8_t* sstate;          // [5 CL + kSortedPad]
    uint8_t* ssurf;           // surface flag (sorted)
    uint16_t* sscene;         // [5 CL] scene of each sorted entry
    int* sidx;                // [5 CL] sorted index of unified id (written
                              // where read: band rule, relaxation, inspection)
    // [5 CL * kMaskWords] per cell-sorted target: the same-level
    // candidates its density pass accepted, one bit per candidate of its
    // scan rows (each row starting on a word), read by the force pass
    uint32_t* nmask;
    uint8_t* cand;            // [CL] split candidate flag per sorted low index
    uint8_t* merged;          // [CL] merged flag per low id
    int* tile_sums;           // [2 * max_scan_tiles]
    int* tile_base;           // [2 * max_scan_tiles]
    int* seg_base;            // [8]
    double* red;              // [red_slots * kRedFields]
    int* red_scene;           // [red_slots] scene of each chunk slot (-1 none)
    int* red_base;            // [S] first chunk slot of each scene
    int* red_n;               // [S] chunk slots of each scene
    double* cons;             // [S * 10] conservation sums + bounds
    uint8_t* cflag;           // [2 (CCL + CCH)] cflag[2c + s] = 1: state s in cell c
    uint8_t* csurf;           // [CCL + CCH] 1: a free-surface particle in cell c
    // [CCL] 1: a free-surface particle (either level) in this low cell or
    // one of the eight around it
    uint8_t* cband;
    float4* rtmp;             // [4 CL] relaxation scratch (x', v') per target;
                              // allocated when a scene enables relaxation
    // [5 CL * 10] pair-recording output of the density and force kernels'
    // recording variant (null in the step): per unified id, the counts of
    // the pairs each sum added (same-level density, cross-level density,
    // same-level force, cross-level force, wall candidates) and their id
    // checksums
    uint32_t* pairrec;
};

// Launch shape of one neighbour kernel: block size (128, 256, 512), LDS
// buffer depth in pairs per lane (8, 16, 32), XCD-aware block order (0/1),
// grid cap in blocks (0: capacity-sized grid; else a persistent grid).
struct PairTune {
    int block, depth, swz, grid;
};
enum { kTuneDensityLow = 0, kTuneDensityHigh, kTuneForceLow, kTuneForceHigh,
       kPairKernels };

// Host-known launch parameters of one step.
struct StepShape {
    int passes;               // radix passes
    int key_bits;
    // Grids follow the resident scenes' capacity (cap_alive low ids, 4x
    // as many high): a pool sized for many scenes launches for those it
    // holds. Every kernel loops over or tests the device-side counts.
    long long cap_alive;      // low-id capacity of the resident scenes
    int grid_elem;            // blocks of element-wise kernels
    int sort_tiles;           // sort tiles allocated (row stride of hist)
    int sort_grid;            // sort tiles launched (<= sort_tiles)
    int scan_tiles_l, scan_tiles_h, scan_tiles_g;  // compaction tiles launched
    PairTune tune[kPairKernels];
    int S;
    int red_slots;            // chunk slots in use (grid of the reduction)
    int list_parity;          // which list buffer holds the current lists
    int relax_iters;          // max relaxation iterations over the scenes
    int clear_blocks;         // blocks of the cell-table clear (clear_blk)
    int cap_binds;            // 1: a resident scene has cap_high < 4 cap_low
    int band_any;             // 1: a resident scene uses the surface band rule
    int small_sort;           // 1: the resident capacity (5 cap_alive keys)
                              // fits kSmallSortKeys: one-workgroup sort
    int count_sort;           // 1: counting sort by cell (not small_sort,
                              // key_cells <= kCountCellsPerKey * 5 cap_alive)
    int key_cells;            // cells [0, key_cells) hold every resident cell
    int cell_tiles;           // ceil(key_cells / kCellTile)
    int sorted_buf;           // keys/vals buffer holding the sorted order
    int solo;                 // the only resident scene's slot, else -1
};

// Timing points of one step (events recorded before / after each phase):
// 0 start | grid build | 1 | density low | 2 | density high | 3 | force low
// | 4 | force high | 5 | reductions | 6 | adaptivity + compaction | 7 |
// finalize | 8
constexpr int kTimingPoints = 9;
// Work counters of k_work: candidates (low same, low cross, high same, high
// cross), density pairs (low, high), force pairs (low, high), wall
// candidates (low, high), targets (low, high)
constexpr int kWorkCounts = 13;

// Kernel launchers (tss_kernels.hip). All enqueue on `stream`.
// timing_events: hipEvent_t[kTimingPoints] or nullptr; event_flags is passed
// to hipEventRecordWithFlags (hipEventRecordExternal under graph capture).
void launch_step(const DevState& d, const StepShape& sh, void* stream,
                 void* timing_events, unsigned int event_flags);
void launch_neighbors_n(const DevState& d, const StepShape& sh, int scene,
                        int level, int n, uint32_t* counts, uint32_t* sums,
                        void* stream);
void launch_grid_only(const DevState& d, const StepShape& sh, void* stream);
void launch_work_counts(const DevState& d, const StepShape& sh,
                        unsigned long long* out, void* stream);
// Grid build of the current lists, then the density and force kernels in
// their recording variant: no state is advanced; d.pairrec receives the
// pairs each sum added (kPairRecWords words per unified id).
constexpr int kPairRecWords = 10;
void launch_pair_record(const DevState& d, const StepShape& sh, void* stream);
void launch_fill_int(int* p, int value, long long n, void* stream);
void launch_iota(int* p, int first, int n, void* stream);
// Frame export sections (any pointer may be null: field not exported).
struct FrameOut {
    int* id;
    float2* pos;
    float2* vel;
    float* rho;
    float* p;
    float* vis;
    uint8_t* state;
};
// Tiles of the selection over a list of n entries; tile_cnt holds
// frame_tiles(n) + 1 ints (the last one receives the selected total).
int frame_tiles(int n);
// The entries of list[0..n) in [lo, hi), in list order, gathered into out
// (at most n_out entries written); enqueued on stream.
void launch_frame_pack(const DevState& d, const int* list, int n, int lo, int hi,
                       int* tile_cnt, int n_out, const FrameOut& out, void* stream);
// A scene's boundary positions by id: out[bids[k]] = bpos[k], k < nb (its
// cell-ordered copy and scene-local ids).
void launch_frame_bnd(const float2* bpos, const int* bids, int nb, float2* out,
                      void* stream);
// Fixed device scratch of the frame export that allocates nothing: a
// file-scope array of kNoAllocScratchBytes in the code object, one copy per
// device, held until the process ends. noalloc_scratch gives its address on
// `device` (looked up once per device with that device current; host code
// uses this address, never the array's name).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t noalloc_scratch(int device, void** out);
void launch_dam_break(float2* pos, int nx, int ny, float dx, float jitter,
                      unsigned long long seed, unsigned long long scene,
                      void* stream);
void launch_copy16(const void* src, void* dst, long long n16, void* stream);
// n 8-byte reads, one every `stride` 8-byte words of src (src holds at least
// n * stride words); sink: one 8-byte word.
void launch_gather_probe(const void* src, long long n, int stride, void* sink, void* stream);
// Stable removal of one scene's entries from the current lists (a scene
// retired between steps), into the other list buffers: the caller flips
// the list parity afterwards. n_l, n_h, n_g: the list sizes before.
void launch_filter_scene(const DevState& d, const StepShape& sh, int scene,
                         int n_l, int n_h, int n_g, void* stream);

}  // namespace tss
