// tss_host.cpp -- host runtime of libtss: batches of scenes in one device
// state, scene admission / retirement, the step driver (eager and hipGraph),
// field I/O, the seeded scene generator and the C ABI of include/tss.h.
//
// Reference: segfault11/TwoScaleSplit TwoScale2D/ ("R/"): WCSPHSolver.{h,cu},
// ParticleSystem.{h,cpp}.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tss.h"
#include "tss_device.h"

namespace {

thread_local std::string g_last_error;

// particle states (uint8): low 0 active, 1 transient parent, 2 deleted
constexpr uint8_t kStateActive = 0, kStateDeleted = 2;

// frame export layout: sections padded to 16 bytes, in field-bit order
size_t frame_pad(size_t x) { return (x + 15) & ~(size_t)15; }
size_t frame_level_bytes(uint32_t f, size_t n) {
    size_t b = 0;
    if (f & TSS_FRAME_ID) b += frame_pad(4 * n);
    if (f & TSS_FRAME_POS) b += frame_pad(8 * n);
    if (f & TSS_FRAME_VEL) b += frame_pad(8 * n);
    if (f & TSS_FRAME_RHO) b += frame_pad(4 * n);
    if (f & TSS_FRAME_P) b += frame_pad(4 * n);
    if (f & TSS_FRAME_VIS) b += frame_pad(4 * n);
    if (f & TSS_FRAME_STATE) b += frame_pad(n);
    return b;
}

struct TssError : std::runtime_error {
    int code;
    TssError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            throw TssError(TSS_E_HIP, std::string(#expr) + ": " +              \
                                          hipGetErrorString(e_));              \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return TSS_OK;
    } catch (const TssError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return TSS_E_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return TSS_E_INVALID;
    }
}

void require(bool ok, const std::string& msg, int code = TSS_E_INVALID) {
    if (!ok) throw TssError(code, msg);
}
// A literal message builds no string unless the check fails.
void require(bool ok, const char* msg, int code = TSS_E_INVALID) {
    if (!ok) throw TssError(code, msg);
}

// Largest non-negative float f with sqrtf(f) <= R (inclusive) or < R: the
// squared-distance threshold equivalent to the FP32 predicate on
// sqrt(dx*dx + dy*dy) (sqrtf is correctly rounded, hence monotone).
float sqrt_threshold(float R, bool inclusive) {
    auto ok = [&](uint32_t bits) {
        float f;
        std::memcpy(&f, &bits, 4);
        float r = std::sqrt(f);
        return inclusive ? (r <= R) : (r < R);
    };
    uint32_t lo = 0, hi = 0x7f800000u;  // ok(lo) holds for R > 0
    if (!ok(lo)) return -1.0f;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (ok(mid)) lo = mid; else hi = mid;
    }
    float f;
    std::memcpy(&f, &lo, 4);
    return f;
}

constexpr double kPi = 3.14159265358979323846;  // R/WCSPHSolver.cu:19
static_assert(TSS_PHASE_COUNT == tss::kTimingPoints - 1, "phase count");
static_assert(TSS_WORK_COUNT == tss::kWorkCounts, "work counter count");
static_assert(sizeof(tss_frame_header) == 64, "frame header is 64 bytes");
static_assert(offsetof(tss_snapshot_header, cfg) == 24, "snapshot configuration offset");

// Bulk value (in units of rho0) of the high-level density estimator: the
// blended kernel (W_h + W_H)/2 summed only over r <= H (R:402-474), i.e.
// (1 + F)/2 with F the fraction of the 2-D M4 kernel W_h inside radius H.
// F = M(2H/h)/M(2), M(Q) = int_0^Q f(q) q dq, f the M4 shape. 1 when H = h.
double high_density_ratio(double s) {
    auto f = [](double q) {
        double a = std::max(2.0 - q, 0.0), b = std::max(1.0 - q, 0.0);
        return a * a * a - 4.0 * b * b * b;
    };
    auto M = [&](double Q) {
        const int n = 20000;  // Simpson; f(q) q is piecewise cubic*linear
        double hq = Q / n, acc = 0.0;
        for (int i = 0; i <= n; ++i) {
            double q = i * hq, w = (i == 0 || i == n) ? 1.0 : (i % 2 ? 4.0 : 2.0);
            acc += w * f(q) * q;
        }
        return acc * hq / 3.0;
    };
    double Q = std::min(2.0 * s, 2.0);
    return 0.5 * (1.0 + M(Q) / M(2.0));
}

int axis_cell_host(float x, float o, float cell, int dim) {
    float t = ((x + 0.0f) - o) / cell;
    if (!(t >= 1.0f)) return 0;
    if (t >= (float)dim) return dim - 1;
    return (int)t;
}

// First-fit allocator over an index range.
class RangeAlloc {
public:
    explicit RangeAlloc(long long cap = 0) { reset(cap); }
    void reset(long long cap) {
        free_.clear();
        if (cap > 0) free_[0] = cap;
    }
    bool alloc(long long n, long long& start) {
        if (n == 0) {
            start = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                start = it->first;
                long long rest = it->second - n;
                free_.erase(it);
                if (rest > 0) free_[start + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(long long start, long long n) {
        if (n == 0) return;
        auto it = free_.emplace(start, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<long long, long long> free_;
};

struct SceneHost {
    bool alive = false;
    tss_config cfg{};
    int lbase = 0, lcap = 0;
    int bbase = 0, nb = 0;
    int cbl = 0, ncl = 0;     // low cells: unified [cbl, cbl + ncl)
    int cbh = 0, nch = 0;     // high cells: unified [cbh, cbh + nch)
};

template <class T>
T* dalloc(size_t n, std::vector<void*>& owned) {
    void* p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw TssError(TSS_E_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    owned.push_back(p);
    return static_cast<T*>(p);
}

}  // namespace

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
struct tss_batch {
    static constexpr long long kElemGridCap = 32768;  // blocks of 256
    static constexpr long long kPairGridFull = 16384;
    static constexpr int kPairGridCap = 8192;
    tss::DevState d{};
    tss::StepShape sh{};
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    std::vector<SceneHost> scenes;
    RangeAlloc a_low, a_bnd, a_cl, a_ch;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    tss::PairTune tune[tss::kPairKernels] = {
        {256, 16, 0, 0}, {256, 16, 0, 0}, {256, 16, 0, 0}, {256, 16, 0, 0}};
    long long cap_l = 0, cap_b = 0, ccl = 0, cch = 0;
    int max_scenes = 0;
    int sort_path = TSS_SORT_AUTO;  // grid-build sort (tss_set_grid_sort)
    int device = 0;                 // the device the batch lives on
    uint32_t* nb_counts = nullptr;  // neighbour inspection scratch
    uint32_t* nb_sums = nullptr;
    size_t nb_cap = 0;
    void* frame_buf = nullptr;      // frame export staging (grown on demand)
    size_t frame_cap = 0;

    ~tss_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        invalidate_graphs();
        for (auto& row : tev)
            for (auto& e : row)
                if (e) (void)hipEventDestroy(e);
        for (auto& row : placeholder)
            for (auto& e : row)
                if (e) (void)hipEventDestroy(e);
        for (void* p : owned) (void)hipFree(p);
        if (nb_counts) (void)hipFree(nb_counts);
        if (nb_sums) (void)hipFree(nb_sums);
        if (frame_buf) (void)hipFree(frame_buf);
        if (stream) (void)hipStreamDestroy(stream);
    }

    static long long bytes_for(const tss_batch_desc& ds) {
        const long long CL = ds.cap_low, U = 5 * CL;
        const long long tiles = (U + tss::kSortTile - 1) / tss::kSortTile + 1;
        long long b = 0;
        b += U * (8 * 3 + 4 * 3 + 1);             // pos vel acc rho p vis state
        b += CL * (4 + 4);                         // age scene_of
        b += 2 * (CL + 4 * CL + CL) * 4;           // lists
        b += 4 * U * 4;                            // keys/vals x2
        b += (long long)tss::kRadix * tiles * 4 + tss::kRadix * 4;
        b += (ds.cells_low + ds.cells_high) * 8 + ds.cells_low * 9;
        b += ds.cap_bnd * 12;
        b += U * (8 + 16 + 8 + 1 + 1 + 2 + 4);  // sorted copies + sidx
        b += tss::kSortedPad * (8 + 1);
        b += U * 16;                               // sorted post-step state
        b += U * 4 * tss::kMaskWords;              // candidate masks
        b += CL * 2;
        b += (long long)ds.max_scenes * (sizeof(tss::SceneDev) + sizeof(tss::SceneCount) + 10 * 8 + 12);
        b += (U / tss::kRedChunk + ds.max_scenes + 2) * (tss::kRedFields * 8 + 4);
        b += 3 * (ds.cells_low + ds.cells_high) + ds.cells_low;  // cell state and surface flags
        b += 2 * tss::kFlagPad;
        b += 4 * (ds.cells_low + ds.cells_high);  // cell -> scene
        // counting sort: counts per cell, tile sums, slots
        b += 4 * (ds.cells_low + ds.cells_high) + 4 * ((ds.cells_low + ds.cells_high) / tss::kCellTile + 1);
        b += 8 * U;
        b += 8 * ((ds.cells_low + ds.cells_high) / tss::kClearChunk + 2LL * ds.max_scenes + 1);
        b += 16 * 4 * CL;                         // relaxation scratch (when used)
        return b;
    }

    void create(const tss_batch_desc& ds) {
        // the cell-sorted scene index is 16-bit
        require(ds.max_scenes <= 65535, "a batch holds at most 65535 scenes");
        require(ds.max_scenes >= 1 && ds.cap_low >= 1 && ds.cap_bnd >= 0 &&
                    ds.cells_low >= 1 && ds.cells_high >= 1,
                "invalid batch capacities");
        require(5 * ds.cap_low < (1LL << 31) && ds.cells_low + ds.cells_high < (1LL << 31),
                "batch capacities exceed 32-bit indexing");
        // neighbour gathers use 32-bit byte offsets into the sorted copies
        require(16 * 5 * ds.cap_low < (1LL << 32), "batch capacities exceed 32-bit byte offsets");
        cap_l = ds.cap_low;
        cap_b = ds.cap_bnd;
        ccl = ds.cells_low;
        cch = ds.cells_high;
        max_scenes = ds.max_scenes;
        HIP_CHECK(hipGetDevice(&device));
        HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const long long CL = cap_l, U = 5 * CL;
        d.CL = (int)CL;
        d.NB = (int)cap_b;
        d.CCL = (int)ccl;
        d.CCH = (int)cch;
        d.S = max_scenes;
        d.scenes = dalloc<tss::SceneDev>(max_scenes, owned);
        d.scnt = dalloc<tss::SceneCount>(max_scenes, owned);
        d.counts = dalloc<tss::DevCounts>(1, owned);
        d.scene_of = dalloc<int>(CL, owned);
        d.pos = dalloc<float2>(U, owned);
        d.vel = dalloc<float2>(U, owned);
        d.acc = dalloc<float2>(U, owned);
        d.rho = dalloc<float>(U, owned);
        d.p = dalloc<float>(U, owned);
        d.vis = dalloc<float>(U, owned);
        d.state = dalloc<uint8_t>(U, owned);
        d.age = dalloc<int>(CL, owned);
        for (int i = 0; i < 2; ++i) {
            d.list_l[i] = dalloc<int>(CL, owned);
            d.list_h[i] = dalloc<int>(4 * CL, owned);
            d.list_g[i] = dalloc<int>(CL, owned);
            d.keys[i] = dalloc<uint32_t>(U, owned);
            d.vals[i] = dalloc<uint32_t>(U, owned);
        }
        sh.sort_tiles = (int)((U + tss::kSortTile - 1) / tss::kSortTile);
        d.hist = dalloc<uint32_t>((size_t)tss::kRadix * sh.sort_tiles, owned);
        d.digit_tot = dalloc<uint32_t>(tss::kRadix, owned);
        d.cstart = dalloc<int>(ccl + cch, owned);
        d.cscene = dalloc<int>(ccl + cch, owned);
        d.ccount = dalloc<int>(ccl + cch, owned);
        d.ctile = dalloc<int>((ccl + cch) / tss::kCellTile + 1, owned);
        d.cslot = dalloc<uint2>(U, owned);
        d.cend = dalloc<int>(ccl + cch, owned);
        d.bstart = dalloc<int>(ccl, owned);
        d.bend = dalloc<int>(ccl, owned);
        d.bnear = dalloc<uint8_t>(ccl, owned);
        d.bpos = dalloc<float2>(cap_b, owned);
        d.bids = dalloc<int>(cap_b, owned);
        d.spos = dalloc<float2>(U + tss::kSortedPad, owned);
        d.sxv = dalloc<float4>(U, owned);
        d.snew = dalloc<float4>(U, owned);
        d.srp = dalloc<float2>(U, owned);
        d.sstate = dalloc<uint8_t>(U + tss::kSortedPad, owned);
        d.ssurf = dalloc<uint8_t>(U, owned);
        d.sscene = dalloc<uint16_t>(U, owned);
        d.sidx = dalloc<int>(U, owned);
        d.nmask = dalloc<uint32_t>((size_t)U * tss::kMaskWords, owned);
        d.cand = dalloc<uint8_t>(CL, owned);
        d.merged = dalloc<uint8_t>(CL, owned);
        const int nst = (int)((CL + tss::kScanTile - 1) / tss::kScanTile) * 2 +
                        (int)((4 * CL + tss::kScanTile - 1) / tss::kScanTile);
        d.tile_sums = dalloc<int>(2 * (size_t)nst, owned);
        d.tile_base = dalloc<int>(2 * (size_t)nst, owned);
        d.seg_base = dalloc<int>(8, owned);
        red_cap = (int)((U + tss::kRedChunk - 1) / tss::kRedChunk) + max_scenes + 1;
        d.red = dalloc<double>((size_t)red_cap * tss::kRedFields, owned);
        d.red_scene = dalloc<int>(red_cap, owned);
        clear_cap = (int)((ccl + cch) / tss::kClearChunk) + 2 * max_scenes + 1;
        d.clear_blk = dalloc<int2>(clear_cap, owned);
        d.red_base = dalloc<int>(max_scenes, owned);
        d.red_n = dalloc<int>(max_scenes, owned);
        d.cons = dalloc<double>((size_t)max_scenes * 10, owned);
        d.cflag = dalloc<uint8_t>(2 * (ccl + cch), owned);
        // the band tests read these flags as aligned words: padded past
        // the last cell
        d.csurf = dalloc<uint8_t>(ccl + cch + tss::kFlagPad, owned);
        d.cband = dalloc<uint8_t>(ccl + tss::kFlagPad, owned);
        // initial contents
        HIP_CHECK(hipMemsetAsync(d.scenes, 0, sizeof(tss::SceneDev) * max_scenes, stream));
        HIP_CHECK(hipMemsetAsync(d.scnt, 0, sizeof(tss::SceneCount) * max_scenes, stream));
        HIP_CHECK(hipMemsetAsync(d.counts, 0, sizeof(tss::DevCounts), stream));
        HIP_CHECK(hipMemsetAsync(d.cons, 0, sizeof(double) * 10 * max_scenes, stream));
        HIP_CHECK(hipMemsetAsync(d.red, 0, sizeof(double) * tss::kRedFields * red_cap, stream));
        HIP_CHECK(hipMemsetAsync(d.cflag, 0, 2 * (size_t)(ccl + cch), stream));
        HIP_CHECK(hipMemsetAsync(d.csurf, 0, (size_t)(ccl + cch + tss::kFlagPad), stream));
        HIP_CHECK(hipMemsetAsync(d.cband, 0, (size_t)(ccl + tss::kFlagPad), stream));
        tss::launch_fill_int(d.scene_of, 0, CL, stream);
        HIP_CHECK(hipMemsetAsync(d.state, 0, (size_t)U, stream));
        HIP_CHECK(hipMemsetAsync(d.pos, 0, sizeof(float2) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.vel, 0, sizeof(float2) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.acc, 0, sizeof(float2) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.rho, 0, sizeof(float) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.p, 0, sizeof(float) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.vis, 0, sizeof(float) * U, stream));
        HIP_CHECK(hipMemsetAsync(d.age, 0, sizeof(int) * CL, stream));
        HIP_CHECK(hipMemsetAsync(d.merged, 0, (size_t)CL, stream));
        HIP_CHECK(hipMemsetAsync(d.cand, 0, (size_t)CL, stream));
        tss::launch_fill_int(d.cscene, -1, ccl + cch, stream);
        HIP_CHECK(hipMemsetAsync(d.ccount, 0, sizeof(int) * (size_t)(ccl + cch), stream));
        tss::launch_fill_int(d.cstart, -1, ccl + cch, stream);
        tss::launch_fill_int(d.cend, -1, ccl + cch, stream);
        tss::launch_fill_int(d.bstart, -1, ccl, stream);
        tss::launch_fill_int(d.bend, -1, ccl, stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        a_low.reset(cap_l);
        a_bnd.reset(cap_b);
        a_cl.reset(ccl);
        a_ch.reset(cch);
        scenes.assign(max_scenes, SceneHost{});
        // the neighbour kernels default to a persistent grid when a
        // capacity-sized grid would exceed kPairGridFull blocks
        for (int k = 0; k < tss::kPairKernels; ++k) {
            const long long n = (k == tss::kTuneDensityHigh || k == tss::kTuneForceHigh) ? 4 * CL : CL;
            if ((n + tune[k].block - 1) / tune[k].block > kPairGridFull) tune[k].grid = kPairGridCap;
            sh.tune[k] = tune[k];
        }
        sh.S = max_scenes;
        sh.list_parity = 0;
        update_shape();
    }

    // Reduction chunk slots: each live scene owns ceil(5 cap_low / chunk)
    // consecutive slots (a function of its own capacity only).
    void upload_reduction_tables() {
        std::vector<int> scene_of_slot, base(max_scenes, 0), n(max_scenes, 0);
        for (int s = 0; s < max_scenes; ++s) {
            if (!scenes[s].alive) continue;
            const long long u = 5LL * scenes[s].lcap;
            const int k = (int)((u + tss::kRedChunk - 1) / tss::kRedChunk);
            base[s] = (int)scene_of_slot.size();
            n[s] = k;
            for (int i = 0; i < k; ++i) scene_of_slot.push_back(s);
        }
        if (scene_of_slot.empty()) scene_of_slot.push_back(-1);
        require((int)scene_of_slot.size() <= red_cap, "reduction slots exhausted");
        sh.red_slots = (int)scene_of_slot.size();
        HIP_CHECK(hipMemcpyAsync(d.red_scene, scene_of_slot.data(), 4 * scene_of_slot.size(),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d.red_base, base.data(), 4 * (size_t)max_scenes,
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d.red_n, n.data(), 4 * (size_t)max_scenes,
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    int red_cap = 0;
    int clear_cap = 0;

    // The cell tables are cleared at the start of every step only over the
    // resident scenes' cell ranges (a pool sized for many scenes does not
    // pay for its empty capacity): one block per kClearChunk cells.
    void upload_clear_table() {
        std::vector<int2> blk;
        auto add = [&](long long a, long long n) {
            for (long long x = a; x < a + n; x += tss::kClearChunk)
                blk.push_back(make_int2((int)x, (int)std::min(a + n, x + tss::kClearChunk)));
        };
        for (const SceneHost& s : scenes)
            if (s.alive) {
                add(s.cbl, s.ncl);
                add(s.cbh, s.nch);
            }
        if (blk.empty()) blk.push_back(make_int2(0, 0));
        require((int)blk.size() <= clear_cap, "cell-clear table exhausted");
        sh.clear_blocks = (int)blk.size();
        HIP_CHECK(hipMemcpyAsync(d.clear_blk, blk.data(), sizeof(int2) * blk.size(),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    void update_shape() {
        upload_reduction_tables();
        upload_clear_table();
        // launch sizes from the resident scenes' capacity
        long long cap = 0;
        for (const SceneHost& s : scenes)
            if (s.alive) cap += s.lcap;
        cap = std::max<long long>(cap, 1);
        const long long U = 5 * cap;
        sh.cap_alive = cap;
        sh.grid_elem = (int)std::min<long long>(kElemGridCap, (U + 255) / 256);
        sh.sort_grid = (int)std::min<long long>(sh.sort_tiles, (U + tss::kSortTile - 1) / tss::kSortTile);
        sh.scan_tiles_l = (int)((cap + tss::kScanTile - 1) / tss::kScanTile);
        sh.scan_tiles_h = (int)((4 * cap + tss::kScanTile - 1) / tss::kScanTile);
        sh.scan_tiles_g = sh.scan_tiles_l;
        sh.cap_binds = 0;
        sh.band_any = 0;
        for (const SceneHost& s : scenes) {
            if (s.alive && effective_cap_high(s.cfg, s.lcap) < 4LL * s.lcap) sh.cap_binds = 1;
            if (s.alive && s.cfg.mode == 2) sh.band_any = 1;
        }
        long long maxkey = 1;
        for (auto& s : scenes)
            if (s.alive) maxkey = std::max<long long>(maxkey, (long long)s.cbh + s.nch);
        int bits = 1;
        while ((1LL << bits) < maxkey) ++bits;
        sh.key_bits = bits;
        sh.passes = (bits + tss::kRadixBits - 1) / tss::kRadixBits;
        sh.key_cells = (int)maxkey;
        sh.cell_tiles = (int)((maxkey + tss::kCellTile - 1) / tss::kCellTile);
        // the grid build's sort: the counting sort where the resident cells
        // number at most kCountCellsPerKey per key (five short launches,
        // faster than the one-workgroup radix sort even on small batches:
        // 0.104 vs 0.128 ms per C1/C2 step), else the one-workgroup radix
        // sort where the keys fit, else the multi-tile radix sort; a forced
        // one-workgroup sort that no longer fits (scenes added since) falls
        // back to this choice
        const bool small_fits = U <= tss::kSmallSortKeys;
        const bool count_ok = maxkey <= tss::kCountCellsPerKey * U;
        int path = sort_path;
        if (path == TSS_SORT_SMALL && !small_fits) path = TSS_SORT_AUTO;
        if (path == TSS_SORT_AUTO)
            path = count_ok ? TSS_SORT_COUNT : small_fits ? TSS_SORT_SMALL : TSS_SORT_RADIX;
        sh.count_sort = path == TSS_SORT_COUNT;
        sh.small_sort = path == TSS_SORT_SMALL;
        sh.sorted_buf = sh.count_sort ? 1 : (sh.passes & 1);
        sh.solo = -1;
        int resident = 0;
        for (int k = 0; k < max_scenes; ++k)
            if (scenes[k].alive) {
                ++resident;
                sh.solo = k;
            }
        if (resident != 1) sh.solo = -1;
        for (int k = 0; k < tss::kPairKernels; ++k) sh.tune[k] = tune[k];
        sh.relax_iters = 0;
        for (auto& s : scenes)
            if (s.alive) sh.relax_iters = std::max(sh.relax_iters, s.cfg.relax_iters);
        if (sh.relax_iters > 0 && !d.rtmp)  // between steps: never under capture
            d.rtmp = dalloc<float4>((size_t)4 * cap_l, owned);
        invalidate_graphs();
    }

    // Executable graphs are destroyed only once the stream has drained: a
    // replay still in flight keeps reading its exec's kernel arguments, and
    // HIP does not document a deferred destruction (update_shape is reached
    // from set_tuning and write paths right after enqueued replays).
    void invalidate_graphs() {
        bool any = false;
        for (auto& g : graph) any |= g != nullptr;
        for (auto& t : tgraph) any |= t.exec != nullptr;
        if (any && stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graph) {
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        }
        for (auto& t : tgraph) {
            if (t.exec) (void)hipGraphExecDestroy(t.exec);
            if (t.g) (void)hipGraphDestroy(t.g);
            t = TimedGraph{};
        }
    }

    tss::DevCounts read_counts() {
        tss::DevCounts c;
        HIP_CHECK(hipMemcpyAsync(&c, d.counts, sizeof(c), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        return c;
    }

    // Host changes to lists or counts: the next step's compaction takes the
    // full path (events = 1).
    void write_counts(const tss::DevCounts& c) {
        tss::DevCounts w = c;
        w.events = 1;
        HIP_CHECK(hipMemcpyAsync(d.counts, &w, sizeof(w), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    void mark_host_change() {
        const int one = 1;
        HIP_CHECK(hipMemcpyAsync(&d.counts->events, &one, sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    tss::SceneCount read_scnt(int s) {
        tss::SceneCount c;
        HIP_CHECK(hipMemcpyAsync(&c, d.scnt + s, sizeof(c), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        return c;
    }

    static tss::SceneDev make_scene_dev(const tss_config& c) {
        tss::SceneDev s{};
        s.ox = c.domain_origin[0];
        s.oy = c.domain_origin[1];
        s.h = c.effective_radius;
        s.H = c.effective_radius_high;
        s.dlx = c.domain_dims[0];
        s.dly = c.domain_dims[1];
        s.dhx = c.domain_dims_high[0];
        s.dhy = c.domain_dims_high[1];
        s.rho0 = c.rest_density;
        s.B = c.tait_coefficient;
        s.c = c.speed_sound;
        s.alpha = c.alpha;
        s.sigma = c.tension_coefficient;
        s.g = c.gravity;
        s.m_l = c.mass_low;
        s.m_h = c.mass_high;
        s.m_b = c.mass_boundary;
        // M4 coefficients with double pi, stored as float (R:2559-2590)
        const double h = s.h, H = s.H;
        s.sig_h2 = (float)(10.0 / (7.0 * kPi * h * h) / 2.0);
        s.sig_H2 = (float)(10.0 / (7.0 * kPi * H * H) / 2.0);
        s.gc_h2 = (float)(-60.0 / (7.0 * kPi * h * h * h) / 2.0);
        s.gc_H2 = (float)(-60.0 / (7.0 * kPi * H * H * H) / 2.0);
        s.q_h = (float)(2.0 / h);
        s.q_H = (float)(2.0 / H);
        const float hb = (s.h + s.H) / 2.0f;  // R:623
        s.visc_h = 2.0f * s.h * s.alpha * s.c;
        s.visc_H = 2.0f * s.H * s.alpha * s.c;
        s.visc_x = 2.0f * hb * s.alpha * s.c;
        s.eps_h = 0.01f * s.h * s.h;
        s.eps_H = 0.01f * s.H * s.H;
        s.eps_x = 0.01f * hb * hb;
        s.cb_l = s.m_b / (s.m_l + s.m_b);
        s.cb_h = s.m_b / (s.m_h + s.m_b);
        s.cwall = 0.02f * s.c * s.c;
        s.t_le_h = sqrt_threshold(s.h, true);
        s.t_lt_h = sqrt_threshold(s.h, false);
        s.t_le_H = sqrt_threshold(s.H, true);
        s.t_lt_H = sqrt_threshold(s.H, false);
        s.r_band = c.band_radius;
        s.r_merge = c.merge_radius;
        s.t_band = sqrt_threshold(c.band_radius, true);
        s.t_merge = sqrt_threshold(c.merge_radius, true);
        s.mode = c.mode;
        s.split_x = c.split_x;
        s.surf_t = c.surface_ratio * c.rest_density;
        s.surf_t_h = c.surface_ratio_high * c.rest_density;
        s.pmin = c.promote_min_age;
        s.pmax = c.promote_max_age;
        s.pfluc = c.promote_fluc;
        s.cap_h = c.cap_high;  // 0 (no cap) is resolved by add_scene
        s.dt_mode = c.dt_mode;
        s.cfl = c.cfl;
        s.dt_max = c.dt_max;
        s.dt = c.time_step;
        s.alive = 1;
        s.relax_iters = c.relax_iters;
        s.relax_frac = c.relax_dt_frac;
        s.hh_inner = 2.0 * (double)s.H <= (double)s.h ? 1 : 0;
        s.hh_half = s.hh_inner && 2.0f * s.H == s.h && s.sig_H2 == 4.0f * s.sig_h2 &&
                            s.gc_H2 == 8.0f * s.gc_h2 && s.q_H == 2.0f * s.q_h
                        ? 1
                        : 0;
        return s;
    }

    // cap_high 0 means no cap: every child slot (4 cap_low) is available
    static long long effective_cap_high(const tss_config& c, long long cap_low) {
        return c.cap_high > 0 ? (long long)c.cap_high : 4LL * cap_low;
    }

    static void validate(const tss_config& c, int n_low, int cap_low, int n_bnd) {
        require(c.domain_origin[0] < c.domain_end[0] && c.domain_origin[1] < c.domain_end[1],
                "Invalid configuration parameters: empty domain");
        require(c.effective_radius > 0 && c.effective_radius_high > 0 &&
                    c.effective_radius_high <= c.effective_radius,
                "Invalid configuration parameters: radii");
        require(c.domain_dims[0] >= 1 && c.domain_dims[1] >= 1 && c.domain_dims_high[0] >= 1 &&
                    c.domain_dims_high[1] >= 1,
                "Invalid configuration parameters: grid dimensions (call tss_config_init)");
        require((long long)c.domain_dims_high[0] * c.domain_dims_high[1] < (1LL << 30),
                "grid too large");
        require(n_low >= 0 && cap_low >= n_low && cap_low >= 1, "invalid particle counts");
        require(n_bnd >= 0, "invalid boundary count");
        require(c.mode >= 0 && c.mode <= 2, "invalid mode");
        require(c.cap_high >= 0 && (long long)c.cap_high <= 4LL * cap_low,
                "cap_high must be within 0..4*cap_low (child slots 4p+k)");
        require(c.band_radius > 0 && c.merge_radius > 0, "band/merge radius must be > 0");
        require(c.rest_density > 0 && c.mass_low > 0 && c.mass_high > 0, "masses/density must be > 0");
        require(c.dt_mode == TSS_DT_FIXED || (c.dt_mode == TSS_DT_CFL && c.cfl > 0 && c.dt_max > 0),
                "invalid time-step control");
        require(c.time_step > 0, "time step must be > 0");
        require(c.relax_iters >= 0 && c.relax_iters <= 64, "relax_iters must be within 0..64");
        require(c.relax_iters == 0 || (c.relax_dt_frac > 0 && c.relax_dt_frac <= 1),
                "relax_dt_frac must be within (0, 1]");
    }

    // low_xy == nullptr with gen: the fluid block is generated on the device
    // (tss_dam_break), bit for bit the positions of tss_scene_dam_break
    int add_scene(const tss_config& c, const float* low_xy, int n_low, int cap_low,
                  const float* bnd_xy, int n_bnd, const tss_dam_break* gen = nullptr) {
        validate(c, n_low, cap_low, n_bnd);
        require((low_xy != nullptr) != (gen != nullptr) || n_low == 0,
                "a scene takes positions or a generator");
        if (gen) {
            require(gen->nx >= 0 && gen->ny >= 0 && (long long)gen->nx * gen->ny == n_low &&
                        gen->dx > 0 && gen->jitter >= 0 && std::isfinite(gen->dx),
                    "invalid scene generator");
        } else {
            for (int i = 0; i < 2 * n_low; ++i)
                require(std::isfinite(low_xy[i]), "non-finite particle position");
        }
        for (int i = 0; i < 2 * n_bnd; ++i)
            require(std::isfinite(bnd_xy[i]), "non-finite boundary position");
        int sid = -1;
        for (int i = 0; i < max_scenes; ++i)
            if (!scenes[i].alive) {
                sid = i;
                break;
            }
        require(sid >= 0, "batch holds max_scenes scenes", TSS_E_NOMEM);
        const long long ncl = (long long)c.domain_dims[0] * c.domain_dims[1];
        const long long nch = (long long)c.domain_dims_high[0] * c.domain_dims_high[1];
        long long lb, bb, cl, ch;
        if (!a_low.alloc(cap_low, lb)) throw TssError(TSS_E_NOMEM, "low-particle capacity exhausted");
        if (!a_bnd.alloc(n_bnd, bb)) {
            a_low.release(lb, cap_low);
            throw TssError(TSS_E_NOMEM, "boundary capacity exhausted");
        }
        if (!a_cl.alloc(ncl, cl)) {
            a_low.release(lb, cap_low);
            a_bnd.release(bb, n_bnd);
            throw TssError(TSS_E_NOMEM, "low-cell capacity exhausted");
        }
        if (!a_ch.alloc(nch, ch)) {
            a_low.release(lb, cap_low);
            a_bnd.release(bb, n_bnd);
            a_cl.release(cl, ncl);
            throw TssError(TSS_E_NOMEM, "high-cell capacity exhausted");
        }
        SceneHost& sh_ = scenes[sid];
        sh_.cfg = c;
        sh_.lbase = (int)lb;
        sh_.lcap = cap_low;
        sh_.bbase = (int)bb;
        sh_.nb = n_bnd;
        sh_.cbl = (int)cl;
        sh_.ncl = (int)ncl;
        sh_.cbh = (int)(ccl + ch);
        sh_.nch = (int)nch;
        tss::SceneDev sd = make_scene_dev(c);
        sd.cbase_l = sh_.cbl;
        sd.cbase_h = sh_.cbh;
        sd.lbase = sh_.lbase;
        sd.lcap = cap_low;
        sd.cap_h = effective_cap_high(c, cap_low);
        sd.bbase = sh_.bbase;
        sd.nb = n_bnd;

        tss::DevCounts cnt = read_counts();
        const int CL = d.CL;
        const int lbase = sh_.lbase;
        // particle state
        if (gen)
            tss::launch_dam_break(d.pos + lbase, gen->nx, gen->ny, gen->dx, gen->jitter,
                                  gen->seed, gen->scene, stream);
        else
            HIP_CHECK(hipMemcpyAsync(d.pos + lbase, low_xy, sizeof(float) * 2 * (size_t)n_low,
                                     hipMemcpyHostToDevice, stream));
        if (cap_low > n_low)
            HIP_CHECK(hipMemsetAsync(d.pos + lbase + n_low, 0, sizeof(float2) * (size_t)(cap_low - n_low), stream));
        HIP_CHECK(hipMemsetAsync(d.vel + lbase, 0, sizeof(float2) * (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.acc + lbase, 0, sizeof(float2) * (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.rho + lbase, 0, sizeof(float) * (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.p + lbase, 0, sizeof(float) * (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.vis + lbase, 0, sizeof(float) * (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.state + lbase, 0, (size_t)cap_low, stream));
        HIP_CHECK(hipMemsetAsync(d.age + lbase, 0, sizeof(int) * (size_t)cap_low, stream));
        const size_t hb = (size_t)CL + 4 * (size_t)lbase, hn = 4 * (size_t)cap_low;
        HIP_CHECK(hipMemsetAsync(d.pos + hb, 0, sizeof(float2) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.vel + hb, 0, sizeof(float2) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.acc + hb, 0, sizeof(float2) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.rho + hb, 0, sizeof(float) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.p + hb, 0, sizeof(float) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.vis + hb, 0, sizeof(float) * hn, stream));
        HIP_CHECK(hipMemsetAsync(d.state + hb, 1, hn, stream));
        tss::launch_fill_int(d.scene_of + lbase, sid, cap_low, stream);
        tss::launch_fill_int(d.cscene + sh_.cbl, sid, ncl, stream);
        tss::launch_fill_int(d.cscene + sh_.cbh, sid, nch, stream);
        // boundary grid on the low grid (initBoundaries, R:3205-3238): iota
        // ids, hash, stable sort, cell ranges
        std::vector<int> bh(n_bnd), order(n_bnd);
        for (int i = 0; i < n_bnd; ++i) {
            int cx = axis_cell_host(bnd_xy[2 * i], c.domain_origin[0], c.effective_radius, c.domain_dims[0]);
            int cy = axis_cell_host(bnd_xy[2 * i + 1], c.domain_origin[1], c.effective_radius, c.domain_dims[1]);
            bh[i] = cy * c.domain_dims[0] + cx;
        }
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bh[a] < bh[b]; });
        std::vector<float2> bpos(n_bnd);
        std::vector<int> bids(n_bnd);
        std::vector<int> bst(ncl, -1), ben(ncl, -1);
        for (int k = 0; k < n_bnd; ++k) {
            int i = order[k];
            bpos[k] = make_float2(bnd_xy[2 * i], bnd_xy[2 * i + 1]);
            bids[k] = i;
            int key = bh[i];
            if (k == 0 || bh[order[k - 1]] != key) bst[key] = (int)bb + k;
            if (k == n_bnd - 1 || bh[order[k + 1]] != key) ben[key] = (int)bb + k + 1;
        }
        if (n_bnd) {
            HIP_CHECK(hipMemcpyAsync(d.bpos + bb, bpos.data(), sizeof(float2) * n_bnd, hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(d.bids + bb, bids.data(), sizeof(int) * n_bnd, hipMemcpyHostToDevice, stream));
        }
        HIP_CHECK(hipMemcpyAsync(d.bstart + cl, bst.data(), sizeof(int) * ncl, hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d.bend + cl, ben.data(), sizeof(int) * ncl, hipMemcpyHostToDevice, stream));
        // wall-near flags: any boundary particle in the 4x4 cells from (x, y)
        // (an OR over x..x+3, then over y..y+3)
        std::vector<uint8_t> near(ncl, 0);
        {
            const int nx = c.domain_dims[0], ny = c.domain_dims[1];
            std::vector<uint8_t> rowor(ncl, 0);
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    uint8_t f = 0;
                    for (int i = x; i < std::min(x + 4, nx); ++i) f |= bst[(size_t)y * nx + i] >= 0;
                    rowor[(size_t)y * nx + x] = f;
                }
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    uint8_t f = 0;
                    for (int j = y; j < std::min(y + 4, ny); ++j) f |= rowor[(size_t)j * nx + x];
                    near[(size_t)y * nx + x] = f;
                }
        }
        HIP_CHECK(hipMemcpyAsync(d.bnear + cl, near.data(), ncl, hipMemcpyHostToDevice, stream));
        // initial low list: ids in lattice order (the reference's iota list,
        // R/WCSPHSolver.cu:2358-2371), appended to the current list
        const int par = sh.list_parity;
        tss::launch_iota(d.list_l[par] + cnt.n_l, lbase, n_low, stream);
        cnt.n_l += n_low;
        tss::SceneCount sc{};
        sc.n_l = n_low;
        HIP_CHECK(hipMemcpyAsync(d.scnt + sid, &sc, sizeof(sc), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d.scenes + sid, &sd, sizeof(sd), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemsetAsync(d.cons + (size_t)sid * 10, 0, sizeof(double) * 10, stream));
        write_counts(cnt);
        sh_.alive = true;
        update_shape();
        return sid;
    }

    void remove_scene(int sid) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        SceneHost& s = scenes[sid];
        {
            const tss::DevCounts c = read_counts();
            tss::launch_filter_scene(d, sh, sid, c.n_l, c.n_h, c.n_g, stream);
            sh.list_parity ^= 1;  // the filtered lists are in the other buffers
        }
        tss::SceneDev sd{};
        HIP_CHECK(hipMemcpyAsync(d.scenes + sid, &sd, sizeof(sd), hipMemcpyHostToDevice, stream));
        tss::launch_fill_int(d.bstart + s.cbl, -1, s.ncl, stream);
        tss::launch_fill_int(d.bend + s.cbl, -1, s.ncl, stream);
        tss::launch_fill_int(d.cscene + s.cbl, -1, s.ncl, stream);
        tss::launch_fill_int(d.cscene + s.cbh, -1, s.nch, stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        a_low.release(s.lbase, s.lcap);
        a_bnd.release(s.bbase, s.nb);
        a_cl.release(s.cbl, s.ncl);
        a_ch.release(s.cbh - ccl, s.nch);
        s = SceneHost{};
        update_shape();
    }

    // ---- timing ------------------------------------------------------------
    // With timing on, every step records hipEvents at tss::kTimingPoints
    // points on the batch's stream into one slot of a ring of kTimingRing
    // step slots. Nothing waits per step: a slot's times are harvested (its
    // last event waited for) before the slot is reused, and all pending
    // slots when the phase times are read.
    static constexpr int kTimingRing = 64;
    static constexpr int kPhases = tss::kTimingPoints - 1;
    bool timing = false;
    bool tev_made = false;
    hipEvent_t tev[kTimingRing][tss::kTimingPoints] = {};
    bool slot_busy[kTimingRing] = {};
    int slot_next = 0;
    double phase_sum[kPhases] = {};
    long long phase_n = 0;

    void harvest(int slot) {
        if (!slot_busy[slot]) return;
        slot_busy[slot] = false;
        HIP_CHECK(hipEventSynchronize(tev[slot][tss::kTimingPoints - 1]));
        for (int p = 0; p < kPhases; ++p) {
            float t = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&t, tev[slot][p], tev[slot][p + 1]));
            phase_sum[p] += t;
        }
        ++phase_n;
    }
    void harvest_all() {
        for (int k = 0; k < kTimingRing; ++k) harvest((slot_next + k) % kTimingRing);
    }
    int acquire_slot() {
        const int s = slot_next;
        slot_next = (slot_next + 1) % kTimingRing;
        harvest(s);
        slot_busy[s] = true;
        return s;
    }
    void set_timing(bool on) {
        harvest_all();
        if (on && !tev_made) {
            for (auto& row : tev)
                for (auto& e : row) HIP_CHECK(hipEventCreate(&e));
            tev_made = true;
        }
        timing = on;
        for (double& x : phase_sum) x = 0.0;
        phase_n = 0;
    }
    void phase_times(float* out) {
        harvest_all();
        for (int p = 0; p < kPhases; ++p)
            out[p] = phase_n ? (float)(phase_sum[p] / (double)phase_n) : 0.0f;
    }

    void step_eager(int n) {
        check_host_writes();
        for (int i = 0; i < n; ++i) {
            void* ev = timing ? (void*)tev[acquire_slot()] : nullptr;
            tss::launch_step(d, sh, stream, ev, 0u);
            sh.list_parity ^= 1;
        }
        HIP_CHECK(hipGetLastError());
    }

    // Two consecutive steps captured once per starting list parity and
    // replayed: no host synchronisation, counts stay on the device.
    void capture_pair(hipGraph_t* g, hipEvent_t* ev0, hipEvent_t* ev1) {
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        tss::StepShape s2 = sh;
        const unsigned int fl = ev0 ? hipEventRecordExternal : 0u;
        tss::launch_step(d, s2, stream, ev0, fl);
        s2.list_parity ^= 1;
        tss::launch_step(d, s2, stream, ev1, fl);
        HIP_CHECK(hipStreamEndCapture(stream, g));
    }

    // The timed variant holds event-record nodes at every timing point of
    // both steps; before each replay they are pointed at the events of two
    // fresh ring slots (hipGraphExecEventRecordNodeSetEvent), so kernel times
    // are taken inside the replayed graphs themselves.
    struct TimedGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        hipGraphNode_t node[2][tss::kTimingPoints] = {};
    };
    TimedGraph tgraph[2];
    hipEvent_t placeholder[2][tss::kTimingPoints] = {};

    void build_timed_graph(int par) {
        TimedGraph& t = tgraph[par];
        if (!placeholder[0][0])
            for (auto& row : placeholder)
                for (auto& e : row) HIP_CHECK(hipEventCreate(&e));
        capture_pair(&t.g, placeholder[0], placeholder[1]);
        size_t n = 0;
        HIP_CHECK(hipGraphGetNodes(t.g, nullptr, &n));
        std::vector<hipGraphNode_t> nodes(n);
        HIP_CHECK(hipGraphGetNodes(t.g, nodes.data(), &n));
        int found = 0;
        for (hipGraphNode_t nd : nodes) {
            hipGraphNodeType ty;
            HIP_CHECK(hipGraphNodeGetType(nd, &ty));
            if (ty != hipGraphNodeTypeEventRecord) continue;
            hipEvent_t e = nullptr;
            HIP_CHECK(hipGraphEventRecordNodeGetEvent(nd, &e));
            for (int q = 0; q < 2; ++q)
                for (int k = 0; k < tss::kTimingPoints; ++k)
                    if (placeholder[q][k] == e && !t.node[q][k]) {
                        t.node[q][k] = nd;
                        ++found;
                    }
        }
        require(found == 2 * tss::kTimingPoints,
                "graph capture did not keep the timing event nodes");
        HIP_CHECK(hipGraphInstantiate(&t.exec, t.g, nullptr, nullptr, 0));
    }

    void step_graph(int n) {
        check_host_writes();
        const int pairs = n / 2;
        if (pairs > 0) {
            const int par = sh.list_parity;
            if (timing) {
                if (!tgraph[par].exec) build_timed_graph(par);
                TimedGraph& t = tgraph[par];
                for (int i = 0; i < pairs; ++i) {
                    for (int q = 0; q < 2; ++q) {
                        const int slot = acquire_slot();
                        for (int k = 0; k < tss::kTimingPoints; ++k)
                            HIP_CHECK(hipGraphExecEventRecordNodeSetEvent(t.exec, t.node[q][k],
                                                                          tev[slot][k]));
                    }
                    HIP_CHECK(hipGraphLaunch(t.exec, stream));
                }
            } else {
                if (!graph[par]) {
                    hipGraph_t g;
                    capture_pair(&g, nullptr, nullptr);
                    hipError_t e = hipGraphInstantiate(&graph[par], g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    HIP_CHECK(e);
                }
                for (int i = 0; i < pairs; ++i) HIP_CHECK(hipGraphLaunch(graph[par], stream));
            }
        }
        if (n & 1) step_eager(1);
        HIP_CHECK(hipGetLastError());
    }

    // ---- launch-shape tuning ------------------------------------------------
    static void validate_tune(const tss::PairTune& t) {
        require(t.block == 128 || t.block == 256 || t.block == 512,
                "tuning: block must be 128, 256 or 512");
        require(t.depth == 8 || t.depth == 16 || t.depth == 32,
                "tuning: depth must be 8, 16 or 32");
        require(t.swz == 0 || t.swz == 1, "tuning: xcd_order must be 0 or 1");
        require(t.grid >= 0, "tuning: grid must be >= 0");
    }
    void set_tuning(const tss::PairTune* t) {
        for (int k = 0; k < tss::kPairKernels; ++k) validate_tune(t[k]);
        for (int k = 0; k < tss::kPairKernels; ++k) tune[k] = t[k];
        update_shape();
    }

    // Project-written autotuner. Each trial applies one candidate shape to
    // all four neighbour kernels, advances the batch `steps` real steps
    // (results do not depend on the shape, so the steps count) and reads
    // each kernel's mean time from the phase events. After one ordinary
    // step, candidates are visited `rounds` times in alternating order; the
    // lowest mean time per target wins, per kernel. Stage 1: block x depth
    // with the plain grid; stage 2: XCD block order x persistent grid caps on
    // top of each kernel's stage-1 winner.
    int autotune(int steps, int rounds, float* trial_log, int log_cap, int* n_log) {
        require(steps >= 1 && rounds >= 1, "autotune: steps and rounds must be >= 1");
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive > 0, "autotune: no scene");
        const bool was_timing = timing;
        int dev = 0, cus = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        int taken = 0, logged = 0;
        // per-target time of each kernel: the population changes from step to
        // step (the first step of a scene splits before any child exists),
        // so a trial's kernel time is divided by the targets of its level in
        // that step; trials whose level had no target do not count. Each
        // candidate keeps its fastest trial (the other trials carry the
        // interference of whatever else the device did meanwhile)
        auto run_trials = [&](const std::vector<std::array<tss::PairTune, tss::kPairKernels>>& cand,
                              std::vector<std::array<double, tss::kPairKernels>>& tot) {
            tot.assign(cand.size(), {1e300, 1e300, 1e300, 1e300});
            for (int r = 0; r < rounds; ++r) {
                for (size_t i = 0; i < cand.size(); ++i) {
                    // alternate the visiting order so drift cancels
                    const size_t c = (r & 1) ? cand.size() - 1 - i : i;
                    set_tuning(cand[c].data());
                    set_timing(true);
                    step_eager(steps);
                    taken += steps;
                    float ph[kPhases];
                    phase_times(ph);
                    const tss::DevCounts dc = read_counts();
                    const int ntg[tss::kPairKernels] = {dc.n_l_last, dc.n_h_last, dc.n_l_last,
                                                        dc.n_h_last};
                    for (int k = 0; k < tss::kPairKernels; ++k) {
                        if (ntg[k] <= 0) continue;
                        tot[c][k] = std::min(tot[c][k], (double)ph[1 + k] / (double)ntg[k]);
                    }
                    if (trial_log && logged < log_cap) {
                        float* row = trial_log + (size_t)logged * 8;
                        const auto& t = cand[c][0];
                        row[0] = (float)t.block;
                        row[1] = (float)t.depth;
                        row[2] = (float)t.swz;
                        row[3] = (float)t.grid;
                        for (int k = 0; k < tss::kPairKernels; ++k) row[4 + k] = ph[1 + k];
                        ++logged;
                    }
                }
            }
        };
        // one ordinary step first: trials then start from a populated state
        set_timing(false);
        step_eager(1);
        taken += 1;
        // stage 1: block and depth, with the XCD order (the usual winner on
        // dense populations, where the block size that suits it can differ
        // from the one that suits round-robin blocks)
        std::vector<std::array<tss::PairTune, tss::kPairKernels>> cand;
        for (int b : {128, 256, 512})
            for (int dp : {8, 16, 32}) {
                std::array<tss::PairTune, tss::kPairKernels> a;
                for (auto& t : a) t = {b, dp, 1, 0};
                cand.push_back(a);
            }
        std::vector<std::array<double, tss::kPairKernels>> tot;
        run_trials(cand, tot);
        tss::PairTune best[tss::kPairKernels];
        double best_t[tss::kPairKernels];
        for (int k = 0; k < tss::kPairKernels; ++k) {
            size_t bi = 0;
            for (size_t c = 1; c < cand.size(); ++c)
                if (tot[c][k] < tot[bi][k]) bi = c;
            best[k] = cand[bi][k];
            best_t[k] = tot[bi][k];
        }
        // stage 2 (the plain-grid, XCD-order variant is stage 1's winner)
        cand.clear();
        for (int swz : {0, 1})
            for (int gcap : {0, 4 * cus, 8 * cus}) {
                if (swz == 1 && gcap == 0) continue;
                std::array<tss::PairTune, tss::kPairKernels> a;
                for (int k = 0; k < tss::kPairKernels; ++k) a[k] = {best[k].block, best[k].depth, swz, gcap};
                cand.push_back(a);
            }
        run_trials(cand, tot);
        for (int k = 0; k < tss::kPairKernels; ++k)
            for (size_t c = 0; c < cand.size(); ++c)
                if (tot[c][k] < best_t[k]) {
                    best_t[k] = tot[c][k];
                    best[k] = cand[c][k];
                }
        set_tuning(best);
        set_timing(was_timing);
        if (n_log) *n_log = logged;
        return taken;
    }

    // ---- field I/O -------------------------------------------------------
    struct FieldView {
        void* dev;          // device base of the scene's slice
        size_t elem;        // bytes per element
        size_t count;       // elements
        int kind;           // 0 plain, 1 low-id list, 2 high-id list, 3 sorted-low, 4 sorted-high,
                            // 5 keys-low, 6 keys-high, 7 cells-low (rel seg_l0), 8 cells-high (rel seg_h0),
                            // 9 bnd ranges (rel bbase), 10 dt
    };

    FieldView view(int sid, int field) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        const SceneHost& s = scenes[sid];
        const size_t lb = s.lbase, lc = s.lcap;
        const size_t hb = (size_t)d.CL + 4 * lb, hc = 4 * lc;
        switch (field) {
            case TSS_POS_L: return {d.pos + lb, 8, lc, 0};
            case TSS_VEL_L: return {d.vel + lb, 8, lc, 0};
            case TSS_ACC_L: return {d.acc + lb, 8, lc, 0};
            case TSS_RHO_L: return {d.rho + lb, 4, lc, 0};
            case TSS_P_L: return {d.p + lb, 4, lc, 0};
            case TSS_STATE_L: return {d.state + lb, 1, lc, 0};
            case TSS_AGE_L: return {d.age + lb, 4, lc, 0};
            case TSS_VIS_L: return {d.vis + lb, 4, lc, 0};
            case TSS_POS_H: return {d.pos + hb, 8, hc, 0};
            case TSS_VEL_H: return {d.vel + hb, 8, hc, 0};
            case TSS_ACC_H: return {d.acc + hb, 8, hc, 0};
            case TSS_RHO_H: return {d.rho + hb, 4, hc, 0};
            case TSS_P_H: return {d.p + hb, 4, hc, 0};
            case TSS_STATE_H: return {d.state + hb, 1, hc, 0};
            case TSS_VIS_H: return {d.vis + hb, 4, hc, 0};
            case TSS_LIST_L: return {d.list_l[sh.list_parity], 4, 0, 1};
            case TSS_LIST_H: return {d.list_h[sh.list_parity], 4, 0, 2};
            case TSS_LIST_G: return {d.list_g[sh.list_parity], 4, 0, 11};
            case TSS_SORTED_L: return {d.vals[sh.sorted_buf], 4, 0, 3};
            case TSS_SORTED_H: return {d.vals[sh.sorted_buf], 4, 0, 4};
            case TSS_KEYS_L: return {d.keys[sh.sorted_buf], 4, 0, 5};
            case TSS_KEYS_H: return {d.keys[sh.sorted_buf], 4, 0, 6};
            case TSS_CELL_START_L: return {d.cstart + s.cbl, 4, (size_t)s.ncl, 7};
            case TSS_CELL_END_L: return {d.cend + s.cbl, 4, (size_t)s.ncl, 7};
            case TSS_CELL_START_H: return {d.cstart + s.cbh, 4, (size_t)s.nch, 8};
            case TSS_CELL_END_H: return {d.cend + s.cbh, 4, (size_t)s.nch, 8};
            case TSS_BND_SORTED: return {d.bids + s.bbase, 4, (size_t)s.nb, 0};
            case TSS_BND_START: return {d.bstart + s.cbl, 4, (size_t)s.ncl, 9};
            case TSS_BND_END: return {d.bend + s.cbl, 4, (size_t)s.ncl, 9};
            case TSS_POS_B: return {nullptr, 8, (size_t)s.nb, 12};
            case TSS_DT: return {&d.scenes[sid].dt, 4, 1, 0};
            default: break;
        }
        throw TssError(TSS_E_INVALID, "unknown field " + std::to_string(field));
    }

    void read_field(int sid, int field, void* dst, size_t bytes) {
        FieldView v = view(sid, field);
        const SceneHost& s = scenes[sid];
        const int CL = d.CL;
        if (v.kind == 0) {
            require(bytes <= v.elem * v.count, "read exceeds field size");
            HIP_CHECK(hipMemcpyAsync(dst, v.dev, bytes, hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            return;
        }
        if (v.kind == 12) {
            // boundary positions by id, from the cell-ordered copy
            std::vector<float2> bp(s.nb);
            std::vector<int> bi(s.nb);
            if (s.nb) {
                HIP_CHECK(hipMemcpyAsync(bp.data(), d.bpos + s.bbase, 8 * (size_t)s.nb, hipMemcpyDeviceToHost, stream));
                HIP_CHECK(hipMemcpyAsync(bi.data(), d.bids + s.bbase, 4 * (size_t)s.nb, hipMemcpyDeviceToHost, stream));
            }
            HIP_CHECK(hipStreamSynchronize(stream));
            std::vector<float2> out(s.nb);
            for (int k = 0; k < s.nb; ++k) out[bi[k]] = bp[k];
            require(bytes <= 8 * (size_t)s.nb, "read exceeds field size");
            std::memcpy(dst, out.data(), bytes);
            return;
        }
        tss::DevCounts c = read_counts();
        tss::SceneCount sc = read_scnt(sid);
        std::vector<int> buf;
        auto fetch = [&](const void* src, size_t off, size_t n) {
            buf.resize(n);
            if (n) {
                HIP_CHECK(hipMemcpyAsync(buf.data(), (const int*)src + off, 4 * n, hipMemcpyDeviceToHost, stream));
                HIP_CHECK(hipStreamSynchronize(stream));
            }
        };
        std::vector<int> out;
        switch (v.kind) {
            case 1:  // current low list, this scene's ids (local)
                fetch(v.dev, 0, c.n_l);
                for (int u : buf)
                    if (u >= s.lbase && u < s.lbase + s.lcap) out.push_back(u - s.lbase);
                break;
            case 2:
                fetch(v.dev, 0, c.n_h);
                for (int u : buf) {
                    int a = u - CL;
                    if (a >= 4 * s.lbase && a < 4 * (s.lbase + s.lcap)) out.push_back(a - 4 * s.lbase);
                }
                break;
            case 11:
                fetch(v.dev, 0, c.n_g);
                for (int p : buf)
                    if (p >= s.lbase && p < s.lbase + s.lcap) out.push_back(p - s.lbase);
                break;
            case 3:
                fetch(v.dev, sc.seg_l0, sc.seg_l1 - sc.seg_l0);
                for (int u : buf) out.push_back(u - s.lbase);
                break;
            case 4:
                fetch(v.dev, sc.seg_h0, sc.seg_h1 - sc.seg_h0);
                for (int u : buf) out.push_back(u - CL - 4 * s.lbase);
                break;
            case 5:
                fetch(v.dev, sc.seg_l0, sc.seg_l1 - sc.seg_l0);
                for (int k : buf) out.push_back(k - s.cbl);
                break;
            case 6:
                fetch(v.dev, sc.seg_h0, sc.seg_h1 - sc.seg_h0);
                for (int k : buf) out.push_back(k - s.cbh);
                break;
            case 7:
                fetch(v.dev, 0, v.count);
                for (int x : buf) out.push_back(x < 0 ? -1 : x - sc.seg_l0);
                break;
            case 8:
                fetch(v.dev, 0, v.count);
                for (int x : buf) out.push_back(x < 0 ? -1 : x - sc.seg_h0);
                break;
            case 9:
                fetch(v.dev, 0, v.count);
                for (int x : buf) out.push_back(x < 0 ? -1 : x - s.bbase);
                break;
            default:
                throw TssError(TSS_E_INVALID, "field not readable");
        }
        require(bytes <= out.size() * 4, "read exceeds field size (" + std::to_string(out.size()) + " entries)");
        std::memcpy(dst, out.data(), bytes);
    }

    void write_field(int sid, int field, const void* src, size_t bytes) {
        FieldView v = view(sid, field);
        const SceneHost& s = scenes[sid];
        if (v.kind == 0 && field != TSS_DT) {
            require(bytes <= v.elem * v.count, "write exceeds field size");
            HIP_CHECK(hipMemcpyAsync(v.dev, src, bytes, hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            mark_host_change();
            invalidate_graphs();
            if (field == TSS_STATE_L || field == TSS_STATE_H) lists_dirty = true;
            return;
        }
        if (field == TSS_DT) {
            require(bytes == 4, "dt is one float");
            HIP_CHECK(hipMemcpyAsync(v.dev, src, 4, hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            return;
        }
        // lists: only for single-scene batches (re-seeding a state)
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive == 1, "list writes need a batch of one scene");
        const int n = (int)(bytes / 4);
        std::vector<int> buf(n);
        std::memcpy(buf.data(), src, bytes);
        if (v.kind == 1 || v.kind == 11) {
            require(n <= s.lcap, "list too long");
            for (int& x : buf) {
                require(x >= 0 && x < s.lcap, "list id out of range");
                x += s.lbase;
            }
        } else if (v.kind == 2) {
            require(n <= 4 * s.lcap, "list too long");
            for (int& x : buf) {
                require(x >= 0 && x < 4 * s.lcap, "list id out of range");
                x += d.CL + 4 * s.lbase;
            }
        } else {
            throw TssError(TSS_E_INVALID, "field not writable");
        }
        if (n) {
            HIP_CHECK(hipMemcpyAsync(v.dev, buf.data(), bytes, hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
        }
        mark_host_change();
        lists_dirty = true;
    }

    // ---- frame export ------------------------------------------------------
    void frame(int sid, uint32_t fields, void* dst, size_t cap, size_t* bytes) {
        frame_impl(sid, fields, dst, cap, bytes, nullptr, 0);
    }

    // The same frame staged in the fixed device scratch: no allocation.
    void frame_noalloc(int sid, uint32_t fields, void* dst, size_t cap, size_t* bytes) {
        void* p = nullptr;
        HIP_CHECK(tss::noalloc_scratch(device, &p));
        frame_impl(sid, fields, dst, cap, bytes, (char*)p, tss::kNoAllocScratchBytes);
    }

    // A frame of scene sid into the host buffer dst. Staging: the batch's
    // own buffer, grown on demand (fixed == nullptr), or the `fixed` device
    // buffer of fixed_cap bytes, where the frame must fit. Staging layout:
    // low sections | high sections | boundary section | tile counts.
    void frame_impl(int sid, uint32_t fields, void* dst, size_t cap, size_t* bytes,
                    char* fixed, size_t fixed_cap) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        require((fields & ~(uint32_t)TSS_FRAME_ALL) == 0, "unknown frame field bits");
        const SceneHost& s = scenes[sid];
        const tss::SceneCount sc = read_scnt(sid);
        const int nb = (fields & TSS_FRAME_BND) ? s.nb : 0;
        const size_t size = tss_frame_size(fields, sc.n_l, sc.n_h, nb);
        *bytes = size;
        if (!dst) return;
        if (cap < size)
            throw TssError(TSS_E_INVALID, "frame buffer too small (" + std::to_string(size) +
                                              " bytes needed)");
        const tss::DevCounts c = read_counts();
        const size_t lb = frame_level_bytes(fields, sc.n_l);
        const size_t hb = frame_level_bytes(fields, sc.n_h);
        const size_t bb = frame_pad(8 * (size_t)nb);
        const int tl_n = tss::frame_tiles(c.n_l) + 1, th_n = tss::frame_tiles(c.n_h) + 1;
        const size_t need = lb + hb + bb + 4 * (size_t)(tl_n + th_n);
        char* stage;
        if (fixed) {
            if (need > fixed_cap)
                throw TssError(TSS_E_INVALID,
                               "frame needs " + std::to_string(need) +
                                   " bytes of device staging; the no-allocation export holds at most " +
                                   std::to_string(fixed_cap) + " (use tss_batch_frame)");
            stage = fixed;
        } else {
            if (frame_cap < need) {
                if (frame_buf) (void)hipFree(frame_buf);
                frame_buf = nullptr;
                frame_cap = 0;
                HIP_CHECK(hipMalloc(&frame_buf, need));
                frame_cap = need;
            }
            stage = (char*)frame_buf;
        }
        int* tl = (int*)(stage + lb + hb + bb);
        int* th = tl + tl_n;
        auto sections = [&](char* base, size_t n) {
            tss::FrameOut o{};
            size_t off = 0;
            auto take = [&](uint32_t bit, size_t elem) -> void* {
                if (!(fields & bit)) return nullptr;
                void* p = base + off;
                off += frame_pad(elem * n);
                return p;
            };
            o.id = (int*)take(TSS_FRAME_ID, 4);
            o.pos = (float2*)take(TSS_FRAME_POS, 8);
            o.vel = (float2*)take(TSS_FRAME_VEL, 8);
            o.rho = (float*)take(TSS_FRAME_RHO, 4);
            o.p = (float*)take(TSS_FRAME_P, 4);
            o.vis = (float*)take(TSS_FRAME_VIS, 4);
            o.state = (uint8_t*)take(TSS_FRAME_STATE, 1);
            return o;
        };
        const int par = sh.list_parity;
        const int CL = d.CL;
        // padding bytes of the sections are zero (a fill kernel: sections
        // are multiples of 16 bytes)
        if (lb + hb + bb) tss::launch_fill_int((int*)stage, 0, (long long)((lb + hb + bb) / 4), stream);
        tss::launch_frame_pack(d, d.list_l[par], c.n_l, s.lbase, s.lbase + s.lcap, tl, sc.n_l,
                               sections(stage, sc.n_l), stream);
        tss::launch_frame_pack(d, d.list_h[par], c.n_h, CL + 4 * s.lbase,
                               CL + 4 * (s.lbase + s.lcap), th, sc.n_h,
                               sections(stage + lb, sc.n_h), stream);
        if (nb) tss::launch_frame_bnd(d.bpos + s.bbase, d.bids + s.bbase, nb,
                                      (float2*)(stage + lb + hb), stream);
        char* out = (char*)dst;
        int tot[2] = {0, 0};
        if (lb + hb + bb)
            HIP_CHECK(hipMemcpyAsync(out + sizeof(tss_frame_header), stage, lb + hb + bb,
                                     hipMemcpyDeviceToHost, stream));
        if (c.n_l) HIP_CHECK(hipMemcpyAsync(&tot[0], tl + tl_n - 1, 4, hipMemcpyDeviceToHost, stream));
        if (c.n_h) HIP_CHECK(hipMemcpyAsync(&tot[1], th + th_n - 1, 4, hipMemcpyDeviceToHost, stream));
        float dt = 0;
        HIP_CHECK(hipMemcpyAsync(&dt, &d.scenes[sid].dt, 4, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        require(tot[0] == sc.n_l && tot[1] == sc.n_h, "frame selection disagrees with the scene counts",
                TSS_E_STATE);
        tss_frame_header hd{};
        hd.magic = TSS_FRAME_MAGIC;
        hd.version = TSS_FRAME_VERSION;
        hd.fields = fields;
        hd.step = sc.step;
        hd.n_low = sc.n_l;
        hd.n_high = sc.n_h;
        hd.n_bnd = nb;
        hd.dt = dt;
        for (int k = 0; k < 2; ++k) {
            hd.domain_origin[k] = s.cfg.domain_origin[k];
            hd.domain_end[k] = s.cfg.domain_end[k];
        }
        std::memcpy(out, &hd, sizeof(hd));
    }

    // ---- snapshot / resume -------------------------------------------------
    struct Slice {
        void* dev;
        size_t bytes;
    };
    // the by-id arrays of a scene, in snapshot order
    std::vector<Slice> scene_slices(const SceneHost& s) const {
        const size_t lb = s.lbase, lc = s.lcap;
        const size_t hb = (size_t)d.CL + 4 * lb, hc = 4 * lc;
        return {{d.pos + lb, 8 * lc},  {d.vel + lb, 8 * lc}, {d.acc + lb, 8 * lc},
                {d.rho + lb, 4 * lc},  {d.p + lb, 4 * lc},   {d.vis + lb, 4 * lc},
                {d.state + lb, lc},    {d.age + lb, 4 * lc}, {d.pos + hb, 8 * hc},
                {d.vel + hb, 8 * hc},  {d.acc + hb, 8 * hc}, {d.rho + hb, 4 * hc},
                {d.p + hb, 4 * hc},    {d.vis + hb, 4 * hc}, {d.state + hb, hc}};
    }
    static size_t snapshot_size(long long cap_low, long long nb, long long nl, long long nh,
                                long long ng) {
        const size_t lc = (size_t)cap_low, hc = 4 * lc;
        size_t b = frame_pad(sizeof(tss_snapshot_header)) + frame_pad(8 * (size_t)nb);
        for (size_t e : {8, 8, 8, 4, 4, 4, 1, 4}) b += frame_pad(e * lc);
        for (size_t e : {8, 8, 8, 4, 4, 4, 1}) b += frame_pad(e * hc);
        return b + frame_pad(4 * (size_t)nl) + frame_pad(4 * (size_t)nh) + frame_pad(4 * (size_t)ng);
    }

    void snapshot(int sid, void* dst, size_t cap, size_t* bytes) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        const SceneHost& s = scenes[sid];
        const tss::SceneCount sc = read_scnt(sid);
        const size_t size = snapshot_size(s.lcap, s.nb, sc.n_l, sc.n_h, sc.n_g);
        *bytes = size;
        if (!dst) return;
        require(cap >= size, "snapshot buffer too small (" + std::to_string(size) + " bytes needed)");
        const tss::DevCounts c = read_counts();
        const int par = sh.list_parity;
        // the scene's high list and how many of its entries sit in the
        // global fresh tail (the children the next step relaxes)
        std::vector<int> gh(c.n_h), lh;
        if (c.n_h)
            HIP_CHECK(hipMemcpyAsync(gh.data(), d.list_h[par], 4 * (size_t)c.n_h, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        const long long h0 = (long long)d.CL + 4LL * s.lbase, h1 = h0 + 4LL * s.lcap;
        int fresh = 0;
        for (int i = 0; i < c.n_h; ++i)
            if (gh[i] >= h0 && gh[i] < h1) {
                lh.push_back((int)(gh[i] - h0));
                fresh += i >= c.n_h - c.n_fresh;
            }
        require((int)lh.size() == sc.n_h, "high list disagrees with the scene counts", TSS_E_STATE);
        char* out = (char*)dst;
        std::memset(out, 0, size);
        tss_snapshot_header hd{};
        hd.magic = TSS_SNAPSHOT_MAGIC;
        hd.version = TSS_SNAPSHOT_VERSION;
        hd.header_bytes = sizeof(tss_snapshot_header);
        hd.config_bytes = sizeof(tss_config);
        hd.total_bytes = size;
        hd.cfg = s.cfg;
        hd.cap_low = s.lcap;
        hd.n_bnd = s.nb;
        hd.n_low = sc.n_l;
        hd.n_high = sc.n_h;
        hd.n_groups = sc.n_g;
        hd.n_fresh = fresh;
        hd.step = sc.step;
        hd.n_low_last = sc.n_l_last;
        hd.n_high_last = sc.n_h_last;
        hd.n_split = sc.n_split;
        hd.n_promote = sc.n_promote;
        hd.n_merge = sc.n_merge;
        hd.n_dissolve = sc.n_dissolve;
        hd.n_cand = sc.cand;
        hd.particle_steps = (int64_t)sc.psteps;
        HIP_CHECK(hipMemcpyAsync(&hd.dt, &d.scenes[sid].dt, 4, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(hd.conservation, d.cons + (size_t)sid * 10, sizeof(hd.conservation),
                                 hipMemcpyDeviceToHost, stream));
        size_t off = frame_pad(sizeof(hd));
        const size_t off_bnd = off;
        off += frame_pad(8 * (size_t)s.nb);
        for (const Slice& sl : scene_slices(s)) {
            HIP_CHECK(hipMemcpyAsync(out + off, sl.dev, sl.bytes, hipMemcpyDeviceToHost, stream));
            off += frame_pad(sl.bytes);
        }
        HIP_CHECK(hipStreamSynchronize(stream));
        if (s.nb) read_field(sid, TSS_POS_B, out + off_bnd, 8 * (size_t)s.nb);
        if (sc.n_l) read_field(sid, TSS_LIST_L, out + off, 4 * (size_t)sc.n_l);
        off += frame_pad(4 * (size_t)sc.n_l);
        if (sc.n_h) std::memcpy(out + off, lh.data(), 4 * (size_t)sc.n_h);
        off += frame_pad(4 * (size_t)sc.n_h);
        if (sc.n_g) read_field(sid, TSS_LIST_G, out + off, 4 * (size_t)sc.n_g);
        off += frame_pad(4 * (size_t)sc.n_g);
        require(off == size, "snapshot layout", TSS_E_STATE);
        std::memcpy(out, &hd, sizeof(hd));
    }

    // header checks shared by restore and create_from_snapshot
    static const tss_snapshot_header& snapshot_header(const void* src, size_t bytes) {
        require(src != nullptr && bytes >= sizeof(tss_snapshot_header), "snapshot too short");
        const tss_snapshot_header& hd = *(const tss_snapshot_header*)src;
        require(hd.magic == TSS_SNAPSHOT_MAGIC, "not a snapshot");
        require(hd.version == TSS_SNAPSHOT_VERSION && hd.header_bytes == sizeof(tss_snapshot_header) &&
                    hd.config_bytes == sizeof(tss_config),
                "snapshot version or layout not supported");
        require(hd.cap_low > 0 && hd.n_bnd >= 0 && hd.n_low >= 0 && hd.n_low <= hd.cap_low &&
                    hd.n_high >= 0 && hd.n_high <= 4LL * hd.cap_low && hd.n_groups >= 0 &&
                    hd.n_groups <= hd.cap_low && hd.n_fresh >= 0 && hd.n_fresh <= hd.n_high,
                "snapshot counts out of range");
        require(hd.total_bytes == snapshot_size(hd.cap_low, hd.n_bnd, hd.n_low, hd.n_high, hd.n_groups) &&
                    hd.total_bytes <= bytes,
                "snapshot size does not match its header");
        return hd;
    }

    // section offsets of a snapshot whose header passed snapshot_header
    struct SnapLayout {
        size_t bnd;
        size_t arr[15], size[15];
        size_t ll, lh, lg;
    };
    static SnapLayout snapshot_layout(const tss_snapshot_header& hd) {
        SnapLayout L{};
        const size_t lc = (size_t)hd.cap_low, hc = 4 * lc;
        size_t off = frame_pad(sizeof(hd));
        L.bnd = off;
        off += frame_pad(8 * (size_t)hd.n_bnd);
        const size_t sizes[15] = {8 * lc, 8 * lc, 8 * lc, 4 * lc, 4 * lc, 4 * lc, lc, 4 * lc,
                                  8 * hc, 8 * hc, 8 * hc, 4 * hc, 4 * hc, 4 * hc, hc};
        for (int k = 0; k < 15; ++k) {
            L.arr[k] = off;
            L.size[k] = sizes[k];
            off += frame_pad(sizes[k]);
        }
        L.ll = off;
        off += frame_pad(4 * (size_t)hd.n_low);
        L.lh = off;
        off += frame_pad(4 * (size_t)hd.n_high);
        L.lg = off;
        return L;
    }

    // every check a snapshot must pass before anything is allocated for it
    static const tss_snapshot_header& check_snapshot(const void* src, size_t bytes) {
        const tss_snapshot_header& hd = snapshot_header(src, bytes);
        validate(hd.cfg, 0, hd.cap_low, hd.n_bnd);
        const char* in = (const char*)src;
        const SnapLayout L = snapshot_layout(hd);
        const size_t lc = (size_t)hd.cap_low, hc = 4 * lc;
        const size_t* offs = L.arr;
        const size_t* sizes = L.size;
        const float* bnd = (const float*)(in + L.bnd);
        for (size_t i = 0; i < 2 * (size_t)hd.n_bnd; ++i)
            require(std::isfinite(bnd[i]), "non-finite boundary position in snapshot");
        // per-particle arrays: finite positions and velocities, known states
        for (int k : {0, 1, 8, 9}) {
            const float* f = (const float*)(in + offs[k]);
            for (size_t i = 0; i < sizes[k] / 4; ++i)
                require(std::isfinite(f[i]), "non-finite position or velocity in snapshot");
        }
        for (int k : {6, 14}) {
            const uint8_t* st = (const uint8_t*)(in + offs[k]);
            for (size_t i = 0; i < sizes[k]; ++i) require(st[i] <= 2, "unknown particle state in snapshot");
        }
        (void)hc;
        check_lists((const int*)(in + L.ll), hd.n_low, (const int*)(in + L.lh), hd.n_high,
                    (const int*)(in + L.lg), hd.n_groups, (const uint8_t*)(in + offs[6]), lc,
                    "in snapshot");
        return hd;
    }

    // The list invariants that bound every later step's list growth (and
    // so keep the device-side compaction within the list buffers): each
    // list holds ids in range without repeats; listed low particles are
    // active or transient; group parents are transient or deleted (so they
    // cannot split again); every listed child belongs to a listed group.
    // Ids are scene-local; st_l holds the low states (lc of them).
    static void check_lists(const int* ll, int nl, const int* lh, int nh, const int* lg,
                            int ng, const uint8_t* st_l, size_t lc, const char* where) {
        const std::string at = std::string(" ") + where;
        auto check_list = [&](const int* l, int n, size_t range, const char* what) {
            std::vector<uint8_t> seen(range, 0);
            for (int i = 0; i < n; ++i) {
                require(l[i] >= 0 && (size_t)l[i] < range && !seen[l[i]],
                        std::string("invalid ") + what + " list" + at);
                seen[l[i]] = 1;
            }
        };
        check_list(ll, nl, lc, "low");
        check_list(lh, nh, 4 * lc, "high");
        check_list(lg, ng, lc, "group");
        std::vector<uint8_t> group(lc, 0);
        for (int i = 0; i < ng; ++i) {
            require(st_l[lg[i]] != kStateActive, "group parent is active" + at);
            group[lg[i]] = 1;
        }
        for (int i = 0; i < nl; ++i)
            require(st_l[ll[i]] != kStateDeleted, "deleted particle listed" + at);
        for (int i = 0; i < nh; ++i)
            require(group[lh[i] >> 2], "child of an unlisted group" + at);
    }

    // Host writes of lists, counts or states (tss_write, tss_set_counts) may
    // leave lists that break the invariants above; they are checked before
    // the next launch that reads the lists, on the device's current lists,
    // counts and states. A failing state is refused (TSS_E_INVALID) and
    // nothing is launched until the caller writes a valid one.
    bool lists_dirty = false;
    void check_host_writes() {
        if (!lists_dirty) return;
        const tss::DevCounts c = read_counts();
        const int par = sh.list_parity;
        for (int sid = 0; sid < max_scenes; ++sid) {
            const SceneHost& s = scenes[sid];
            if (!s.alive) continue;
            const tss::SceneCount sc = read_scnt(sid);
            require(sc.n_l == c.n_l && sc.n_h == c.n_h && sc.n_g == c.n_g,
                    "host writes need a batch of one scene");
            std::vector<int> ll(c.n_l), lh(c.n_h), lg(c.n_g);
            std::vector<uint8_t> st((size_t)s.lcap);
            if (c.n_l)
                HIP_CHECK(hipMemcpyAsync(ll.data(), d.list_l[par], 4 * (size_t)c.n_l, hipMemcpyDeviceToHost, stream));
            if (c.n_h)
                HIP_CHECK(hipMemcpyAsync(lh.data(), d.list_h[par], 4 * (size_t)c.n_h, hipMemcpyDeviceToHost, stream));
            if (c.n_g)
                HIP_CHECK(hipMemcpyAsync(lg.data(), d.list_g[par], 4 * (size_t)c.n_g, hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipMemcpyAsync(st.data(), d.state + s.lbase, (size_t)s.lcap, hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            for (int& x : ll) x -= s.lbase;
            for (int& x : lh) x -= d.CL + 4 * s.lbase;
            for (int& x : lg) x -= s.lbase;
            check_lists(ll.data(), c.n_l, lh.data(), c.n_h, lg.data(), c.n_g, st.data(),
                        (size_t)s.lcap, "after host writes");
        }
        lists_dirty = false;
    }

    int restore(const void* src, size_t bytes) {
        const tss_snapshot_header& hd = check_snapshot(src, bytes);
        const char* in = (const char*)src;
        const SnapLayout L = snapshot_layout(hd);
        const size_t* offs = L.arr;
        const size_t* sizes = L.size;
        const float* bnd = (const float*)(in + L.bnd);
        const int* ll = (const int*)(in + L.ll);
        const int* lh = (const int*)(in + L.lh);
        const int* lg = (const int*)(in + L.lg);
        const int sid = add_scene(hd.cfg, nullptr, 0, hd.cap_low, bnd, hd.n_bnd);
        const SceneHost& s = scenes[sid];
        const std::vector<Slice> sl = scene_slices(s);
        for (int k = 0; k < 15; ++k)
            HIP_CHECK(hipMemcpyAsync(sl[k].dev, in + offs[k], sizes[k], hipMemcpyHostToDevice, stream));
        // lists in unified ids; the scene's kept high entries go before the
        // batch's fresh tail and its own fresh entries after it, so every
        // scene's fresh children stay at the end of the high list
        tss::DevCounts c = read_counts();
        const int par = sh.list_parity;
        const int CL = d.CL;
        std::vector<int> ul(ll, ll + hd.n_low), ug(lg, lg + hd.n_groups);
        for (int& x : ul) x += s.lbase;
        for (int& x : ug) x += s.lbase;
        const int kept = hd.n_high - hd.n_fresh;
        std::vector<int> uh((size_t)hd.n_high + c.n_fresh);
        for (int i = 0; i < kept; ++i) uh[i] = lh[i] + CL + 4 * s.lbase;
        const int t0 = c.n_h - c.n_fresh;
        if (c.n_fresh)
            HIP_CHECK(hipMemcpyAsync(uh.data() + kept, d.list_h[par] + t0, 4 * (size_t)c.n_fresh,
                                     hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int i = kept; i < hd.n_high; ++i) uh[c.n_fresh + i] = lh[i] + CL + 4 * s.lbase;
        if (hd.n_low)
            HIP_CHECK(hipMemcpyAsync(d.list_l[par] + c.n_l, ul.data(), 4 * ul.size(), hipMemcpyHostToDevice, stream));
        if (hd.n_groups)
            HIP_CHECK(hipMemcpyAsync(d.list_g[par] + c.n_g, ug.data(), 4 * ug.size(), hipMemcpyHostToDevice, stream));
        if (!uh.empty())
            HIP_CHECK(hipMemcpyAsync(d.list_h[par] + t0, uh.data(), 4 * uh.size(), hipMemcpyHostToDevice, stream));
        c.n_l += hd.n_low;
        c.n_h += hd.n_high;
        c.n_g += hd.n_groups;
        c.n_fresh += hd.n_fresh;
        tss::SceneCount sc{};
        sc.n_l = hd.n_low;
        sc.n_h = hd.n_high;
        sc.n_g = hd.n_groups;
        sc.n_l_last = hd.n_low_last;
        sc.n_h_last = hd.n_high_last;
        sc.n_split = hd.n_split;
        sc.n_promote = hd.n_promote;
        sc.n_merge = hd.n_merge;
        sc.n_dissolve = hd.n_dissolve;
        sc.cand = hd.n_cand;
        sc.step = hd.step;
        sc.psteps = (unsigned long long)hd.particle_steps;
        HIP_CHECK(hipMemcpyAsync(d.scnt + sid, &sc, sizeof(sc), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(&d.scenes[sid].dt, &hd.dt, 4, hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d.cons + (size_t)sid * 10, hd.conservation, sizeof(hd.conservation),
                                 hipMemcpyHostToDevice, stream));
        write_counts(c);
        invalidate_graphs();
        return sid;
    }

    void set_counts(int sid, int nl, int nh, int ng) {
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive == 1 && scenes[sid].alive, "set_counts needs a batch of one scene");
        const SceneHost& s = scenes[sid];
        require(nl >= 0 && nl <= s.lcap && nh >= 0 && nh <= 4 * s.lcap && ng >= 0 && ng <= s.lcap,
                "counts out of range");
        tss::DevCounts c = read_counts();
        c.n_l = nl;
        c.n_h = nh;
        c.n_g = ng;
        c.n_fresh = 0;  // a re-seeded state has no children awaiting relaxation
        write_counts(c);
        tss::SceneCount sc = read_scnt(sid);
        sc.n_l = nl;
        sc.n_h = nh;
        sc.n_g = ng;
        HIP_CHECK(hipMemcpyAsync(d.scnt + sid, &sc, sizeof(sc), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        lists_dirty = true;
    }

    void counts(int sid, tss_counts_t* o) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        tss::SceneCount sc = read_scnt(sid);
        const SceneHost& s = scenes[sid];
        o->n_low = sc.n_l;
        o->n_high = sc.n_h;
        o->n_groups = sc.n_g;
        o->n_low_last = sc.n_l_last;
        o->n_high_last = sc.n_h_last;
        o->n_split = sc.n_split;
        o->n_promote = sc.n_promote;
        o->n_merge = sc.n_merge;
        o->n_dissolve = sc.n_dissolve;
        o->step = sc.step;
        o->cap_low = s.lcap;
        o->n_bnd = s.nb;
        o->cells_low = s.ncl;
        o->cells_high = s.nch;
        o->particle_steps = (long long)sc.psteps;
    }

    void conservation(int sid, double* out, double* err) {
        require(sid >= 0 && sid < max_scenes && scenes[sid].alive, "no such scene");
        double buf[10];
        HIP_CHECK(hipMemcpyAsync(buf, d.cons + (size_t)sid * 10, sizeof(buf), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        for (int i = 0; i < 5; ++i) {
            out[i] = buf[i];
            err[i] = buf[5 + i];
        }
    }

    void neighbors(int sid, int level, uint32_t* counts_out, uint32_t* sums_out, size_t n) {
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive == 1 && scenes[sid].alive, "neighbour inspection needs a batch of one scene");
        check_host_writes();
        tss::DevCounts c = read_counts();
        const int nl = level ? c.n_h : c.n_l;
        require(n == (size_t)nl, "n must equal the level's list length (" + std::to_string(nl) + ")");
        if (nb_cap < 5 * n + 5) {
            if (nb_counts) (void)hipFree(nb_counts);
            if (nb_sums) (void)hipFree(nb_sums);
            nb_counts = nb_sums = nullptr;
            HIP_CHECK(hipMalloc(&nb_counts, 4 * (5 * n + 5)));
            HIP_CHECK(hipMalloc(&nb_sums, 4 * (5 * n + 5)));
            nb_cap = 5 * n + 5;
        }
        tss::launch_neighbors_n(d, sh, sid, level, nl, nb_counts, nb_sums, stream);
        if (n) {
            HIP_CHECK(hipMemcpyAsync(counts_out, nb_counts, 20 * n, hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipMemcpyAsync(sums_out, nb_sums, 20 * n, hipMemcpyDeviceToHost, stream));
        }
        HIP_CHECK(hipStreamSynchronize(stream));
        HIP_CHECK(hipGetLastError());
    }

    // The step's density and force kernels in their recording variant on
    // the current state: per listed particle of a level (list order), the
    // pairs each sum adds, as counts and id checksums in the neighbour
    // categories.
    void pair_sets(int sid, int level, uint32_t* counts_out, uint32_t* sums_out, size_t n) {
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive == 1 && scenes[sid].alive, "pair recording needs a batch of one scene");
        check_host_writes();
        const tss::DevCounts c = read_counts();
        const int nl = level ? c.n_h : c.n_l;
        require(n == (size_t)nl, "n must equal the level's list length (" + std::to_string(nl) + ")");
        const size_t U = 5 * (size_t)d.CL, words = (size_t)tss::kPairRecWords * U;
        uint32_t* rec = nullptr;
        HIP_CHECK(hipMalloc(&rec, 4 * words));
        std::vector<uint32_t> host(words);
        std::vector<int> list(n);
        hipError_t e = hipMemsetAsync(rec, 0, 4 * words, stream);
        if (e == hipSuccess) {
            tss::DevState d2 = d;
            d2.pairrec = rec;
            tss::launch_pair_record(d2, sh, stream);
            e = hipMemcpyAsync(host.data(), rec, 4 * words, hipMemcpyDeviceToHost, stream);
        }
        const int par = sh.list_parity;
        if (e == hipSuccess && n)
            e = hipMemcpyAsync(list.data(), level ? d.list_h[par] : d.list_l[par], 4 * n,
                               hipMemcpyDeviceToHost, stream);
        const hipError_t e2 = hipStreamSynchronize(stream);
        (void)hipFree(rec);
        HIP_CHECK(e);
        HIP_CHECK(e2);
        HIP_CHECK(hipGetLastError());
        for (size_t t = 0; t < n; ++t) {
            const uint32_t* r = host.data() + (size_t)list[t] * tss::kPairRecWords;
            for (int q = 0; q < 5; ++q) {
                counts_out[5 * t + q] = r[q];
                sums_out[5 * t + q] = r[5 + q];
            }
        }
    }

    void work_counts(int sid, unsigned long long* out) {
        int alive = 0;
        for (auto& x : scenes) alive += x.alive;
        require(alive == 1 && scenes[sid].alive, "work counts need a batch of one scene");
        check_host_writes();
        unsigned long long* dev = nullptr;
        HIP_CHECK(hipMalloc(&dev, tss::kWorkCounts * sizeof(unsigned long long)));
        tss::launch_work_counts(d, sh, dev, stream);
        hipError_t e = hipMemcpyAsync(out, dev, tss::kWorkCounts * sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, stream);
        hipError_t e2 = hipStreamSynchronize(stream);
        (void)hipFree(dev);
        HIP_CHECK(e);
        HIP_CHECK(e2);
        HIP_CHECK(hipGetLastError());
    }

    // A raw device pointer lets the caller change the state between steps
    // unseen: from now on every step's compaction takes the full path.
    void* device_ptr(int sid, int field) {
        FieldView v = view(sid, field);
        require(v.kind == 0, "field has no plain device view");
        const int one[2] = {1, 1};
        HIP_CHECK(hipMemcpyAsync(&d.counts->events, one, sizeof(one), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        return v.dev;
    }
};

struct tss_solver {
    tss_batch b;
    int sid = 0;
};

// ---------------------------------------------------------------------------
// Scene generator
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float unit24(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }
}  // namespace

extern "C" {

uint64_t tss_splitmix64_next(uint64_t* state) {
    *state += kGolden;
    return mix64(*state);
}

int tss_canvas_count(int nxd, int nyd, int layers) {
    if (nxd < 0 || nyd < 0 || layers < 0) return -1;
    return (nxd + 2 * layers) * (nyd + 2 * layers) - nxd * nyd;
}

int tss_scene_dam_break(uint64_t seed, uint64_t scene, int nx, int ny, float dx, float jitter,
                        int nxd, int nyd, int layers, float* low_xy, float* bnd_xy) {
    return guarded([&] {
        require(nx >= 0 && ny >= 0 && dx > 0 && jitter >= 0 && nxd >= 0 && nyd >= 0 && layers >= 0,
                "invalid scene generator arguments");
        const float amp = jitter * dx;
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                const uint64_t idx = (uint64_t)j * (uint64_t)nx + (uint64_t)i;
                uint64_t st = mix64(mix64(seed + kGolden * (scene + 1)) + kGolden * (idx + 1));
                float ux = unit24(tss_splitmix64_next(&st));
                float uy = unit24(tss_splitmix64_next(&st));
                float jx = (2.0f * ux - 1.0f) * amp;
                float jy = (2.0f * uy - 1.0f) * amp;
                low_xy[2 * idx] = (0.5f + (float)i) * dx + jx;
                low_xy[2 * idx + 1] = (0.5f + (float)j) * dx + jy;
            }
        }
        if (bnd_xy) {
            size_t k = 0;
            for (int j = -layers; j < nyd + layers; ++j) {
                for (int i = -layers; i < nxd + layers; ++i) {
                    if (i < 0 || i >= nxd || j < 0 || j >= nyd) {
                        bnd_xy[2 * k] = (0.5f + (float)i) * dx;
                        bnd_xy[2 * k + 1] = (0.5f + (float)j) * dx;
                        ++k;
                    }
                }
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Configuration (WCSPHConfig, R/WCSPHSolver.cu:2287-2332)
// ---------------------------------------------------------------------------
int tss_config_init(tss_config* cfg, float xs, float ys, float xe, float ye, float h, float H,
                    float rho0, float B, float c, float alpha, float sigma, float dt) {
    return guarded([&] {
        require(cfg != nullptr, "null config");
        std::memset(cfg, 0, sizeof(*cfg));
        require(!(xs >= xe || ys >= ye), "Invalid configuration parameters");
        require(h > 0 && H > 0, "Invalid configuration parameters: radii must be > 0");
        cfg->domain_origin[0] = xs;
        cfg->domain_origin[1] = ys;
        cfg->domain_end[0] = xe;
        cfg->domain_end[1] = ye;
        cfg->domain_dims[0] = (int)std::ceil((xe - xs) / h);
        cfg->domain_dims[1] = (int)std::ceil((ye - ys) / h);
        cfg->domain_dims_high[0] = (int)std::ceil((xe - xs) / H);
        cfg->domain_dims_high[1] = (int)std::ceil((ye - ys) / H);
        cfg->effective_radius = h;
        cfg->effective_radius_high = H;
        cfg->rest_density = rho0;
        cfg->tait_coefficient = B;
        cfg->speed_sound = c;
        cfg->alpha = alpha;
        cfg->tension_coefficient = sigma;
        cfg->time_step = dt;
        cfg->gravity = 9.81f;
        cfg->mode = TSS_MODE_PLAIN;
        cfg->split_x = 0.5f;
        cfg->surface_ratio = 0.9f;
        cfg->surface_ratio_high = 0.9f * (float)high_density_ratio((double)H / (double)h);
        cfg->band_radius = h;
        cfg->merge_radius = h;
        cfg->promote_min_age = 5;
        cfg->promote_max_age = 0;  // promotion off (opt-in)
        cfg->promote_fluc = 0.05f;
        cfg->cap_high = 0;
        cfg->dt_mode = TSS_DT_FIXED;
        cfg->cfl = 0.4f;
        cfg->dt_max = dt;
        cfg->relax_iters = 0;     // relaxation off
        cfg->relax_dt_frac = 0.1f;
    });
}

// ---------------------------------------------------------------------------
// Single-scene solver
// ---------------------------------------------------------------------------
int tss_create(const tss_config* cfg, const float* low_xy, int n_low, int cap_low,
               const float* bnd_xy, int n_bnd, tss_solver** out) {
    return guarded([&] {
        require(cfg && out && (low_xy || n_low == 0) && (bnd_xy || n_bnd == 0), "null argument");
        *out = nullptr;
        tss_batch::validate(*cfg, n_low, cap_low, n_bnd);
        auto s = std::make_unique<tss_solver>();
        tss_batch_desc ds{};
        ds.max_scenes = 1;
        ds.cap_low = cap_low;
        ds.cap_bnd = n_bnd;
        ds.cells_low = (long long)cfg->domain_dims[0] * cfg->domain_dims[1];
        ds.cells_high = (long long)cfg->domain_dims_high[0] * cfg->domain_dims_high[1];
        s->b.create(ds);
        s->sid = s->b.add_scene(*cfg, low_xy, n_low, cap_low, bnd_xy, n_bnd);
        *out = s.release();
    });
}

int tss_step(tss_solver* s, int n) {
    return guarded([&] {
        require(s && n >= 0, "invalid argument");
        s->b.step_eager(n);
    });
}

int tss_step_graph(tss_solver* s, int n) {
    return guarded([&] {
        require(s && n >= 0, "invalid argument");
        s->b.step_graph(n);
    });
}

int tss_counts(tss_solver* s, tss_counts_t* o) {
    return guarded([&] {
        require(s && o, "null argument");
        s->b.counts(s->sid, o);
    });
}

int tss_read(tss_solver* s, int field, void* dst, size_t bytes) {
    return guarded([&] {
        require(s && (dst || bytes == 0), "null argument");
        s->b.read_field(s->sid, field, dst, bytes);
    });
}

int tss_write(tss_solver* s, int field, const void* src, size_t bytes) {
    return guarded([&] {
        require(s && (src || bytes == 0), "null argument");
        s->b.write_field(s->sid, field, src, bytes);
    });
}

int tss_set_counts(tss_solver* s, int nl, int nh, int ng) {
    return guarded([&] {
        require(s, "null argument");
        s->b.set_counts(s->sid, nl, nh, ng);
    });
}

int tss_device_ptr(tss_solver* s, int field, void** out) {
    return guarded([&] {
        require(s && out, "null argument");
        *out = s->b.device_ptr(s->sid, field);
    });
}

int tss_synchronize(tss_solver* s) {
    return guarded([&] {
        require(s, "null argument");
        HIP_CHECK(hipStreamSynchronize(s->b.stream));
    });
}

int tss_conservation(tss_solver* s, double out[5], double err[5]) {
    return guarded([&] {
        require(s && out && err, "null argument");
        s->b.conservation(s->sid, out, err);
    });
}

int tss_neighbors(tss_solver* s, int level, uint32_t* counts, uint32_t* sums, size_t n) {
    return guarded([&] {
        require(s && (level == 0 || level == 1) && ((counts && sums) || n == 0), "invalid argument");
        s->b.neighbors(s->sid, level, counts, sums, n);
    });
}

int tss_pair_sets(tss_solver* s, int level, uint32_t* counts, uint32_t* sums, size_t n) {
    return guarded([&] {
        require(s && (level == 0 || level == 1) && (n == 0 || (counts && sums)), "invalid argument");
        s->b.pair_sets(s->sid, level, counts, sums, n);
    });
}

int tss_enable_timing(tss_solver* s, int on) {
    return guarded([&] {
        require(s, "null argument");
        s->b.set_timing(on != 0);
    });
}

int tss_phase_times(tss_solver* s, float out[TSS_PHASE_COUNT]) {
    return guarded([&] {
        require(s && out, "null argument");
        s->b.phase_times(out);
    });
}

int tss_work_counts(tss_solver* s, unsigned long long out[TSS_WORK_COUNT]) {
    return guarded([&] {
        require(s && out, "null argument");
        s->b.work_counts(s->sid, out);
    });
}

void tss_destroy(tss_solver* s) { delete s; }

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
int tss_batch_memory_bytes(const tss_batch_desc* d, long long* bytes) {
    return guarded([&] {
        require(d && bytes, "null argument");
        *bytes = tss_batch::bytes_for(*d);
    });
}

int tss_batch_create(const tss_batch_desc* d, tss_batch** out) {
    return guarded([&] {
        require(d && out, "null argument");
        *out = nullptr;
        auto b = std::make_unique<tss_batch>();
        b->create(*d);
        *out = b.release();
    });
}

int tss_batch_add_scene(tss_batch* b, const tss_config* cfg, const float* low_xy, int n_low,
                        int cap_low, const float* bnd_xy, int n_bnd, int* scene_id) {
    return guarded([&] {
        require(b && cfg && scene_id && (low_xy || n_low == 0) && (bnd_xy || n_bnd == 0),
                "null argument");
        *scene_id = b->add_scene(*cfg, low_xy, n_low, cap_low, bnd_xy, n_bnd);
    });
}

// wall canvas of a generated scene (host; perimeter-sized)
static std::vector<float> generated_walls(const tss_dam_break& g) {
    const int nb = tss_canvas_count(g.nxd, g.nyd, g.layers);
    require(nb >= 0, "invalid scene generator canvas");
    std::vector<float> walls(2 * (size_t)nb + 2);
    const int rc = tss_scene_dam_break(g.seed, g.scene, 0, 0, g.dx, 0.0f, g.nxd, g.nyd,
                                       g.layers, nullptr, walls.data());
    if (rc != TSS_OK) throw TssError(rc, tss_last_error());
    walls.resize(2 * (size_t)nb);
    return walls;
}

int tss_batch_add_dam_break(tss_batch* b, const tss_config* cfg, const tss_dam_break* g,
                            int cap_low, int* scene_id) {
    return guarded([&] {
        require(b && cfg && g && scene_id, "null argument");
        require(g->nx >= 0 && g->ny >= 0 && (long long)g->nx * g->ny <= cap_low,
                "scene generator larger than cap_low");
        std::vector<float> walls = generated_walls(*g);
        *scene_id = b->add_scene(*cfg, nullptr, g->nx * g->ny, cap_low, walls.data(),
                                 (int)(walls.size() / 2), g);
    });
}

int tss_batch_remove_scene(tss_batch* b, int sid) {
    return guarded([&] {
        require(b, "null argument");
        b->remove_scene(sid);
    });
}

int tss_batch_step(tss_batch* b, int n) {
    return guarded([&] {
        require(b && n >= 0, "invalid argument");
        b->step_eager(n);
    });
}

int tss_batch_step_graph(tss_batch* b, int n) {
    return guarded([&] {
        require(b && n >= 0, "invalid argument");
        b->step_graph(n);
    });
}

int tss_batch_counts(tss_batch* b, int sid, tss_counts_t* o) {
    return guarded([&] {
        require(b && o, "null argument");
        b->counts(sid, o);
    });
}

size_t tss_frame_size(uint32_t fields, int n_low, int n_high, int n_bnd) {
    if ((fields & ~(uint32_t)TSS_FRAME_ALL) || n_low < 0 || n_high < 0 || n_bnd < 0) return 0;
    return sizeof(tss_frame_header) + frame_level_bytes(fields, (size_t)n_low) +
           frame_level_bytes(fields, (size_t)n_high) +
           ((fields & TSS_FRAME_BND) ? frame_pad(8 * (size_t)n_bnd) : 0);
}

int tss_frame(tss_solver* s, uint32_t fields, void* dst, size_t cap, size_t* bytes) {
    return guarded([&] {
        require(s && bytes, "null argument");
        s->b.frame(s->sid, fields, dst, cap, bytes);
    });
}

int tss_batch_frame_noalloc(tss_batch* b, int sid, uint32_t fields, void* dst, size_t cap,
                            size_t* bytes) {
    return guarded([&] {
        require(b && bytes, "null argument");
        b->frame_noalloc(sid, fields, dst, cap, bytes);
    });
}

int tss_batch_frame(tss_batch* b, int sid, uint32_t fields, void* dst, size_t cap,
                    size_t* bytes) {
    return guarded([&] {
        require(b && bytes, "null argument");
        b->frame(sid, fields, dst, cap, bytes);
    });
}

size_t tss_snapshot_size(int cap_low, int n_bnd, int n_low, int n_high, int n_groups) {
    if (cap_low <= 0 || n_bnd < 0 || n_low < 0 || n_low > cap_low || n_high < 0 ||
        (long long)n_high > 4LL * cap_low || n_groups < 0 || n_groups > cap_low)
        return 0;
    return tss_batch::snapshot_size(cap_low, n_bnd, n_low, n_high, n_groups);
}

int tss_snapshot(tss_solver* s, void* dst, size_t cap, size_t* bytes) {
    return guarded([&] {
        require(s && bytes, "null argument");
        s->b.snapshot(s->sid, dst, cap, bytes);
    });
}

int tss_batch_snapshot(tss_batch* b, int sid, void* dst, size_t cap, size_t* bytes) {
    return guarded([&] {
        require(b && bytes, "null argument");
        b->snapshot(sid, dst, cap, bytes);
    });
}

int tss_batch_restore(tss_batch* b, const void* src, size_t bytes, int* scene_id) {
    return guarded([&] {
        require(b && scene_id, "null argument");
        *scene_id = b->restore(src, bytes);
    });
}

int tss_create_from_snapshot(const void* src, size_t bytes, tss_solver** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = nullptr;
        const tss_snapshot_header& hd = tss_batch::check_snapshot(src, bytes);
        auto s = std::make_unique<tss_solver>();
        tss_batch_desc ds{};
        ds.max_scenes = 1;
        ds.cap_low = hd.cap_low;
        ds.cap_bnd = hd.n_bnd;
        ds.cells_low = (long long)hd.cfg.domain_dims[0] * hd.cfg.domain_dims[1];
        ds.cells_high = (long long)hd.cfg.domain_dims_high[0] * hd.cfg.domain_dims_high[1];
        s->b.create(ds);
        s->sid = s->b.restore(src, bytes);
        *out = s.release();
    });
}

int tss_batch_read(tss_batch* b, int sid, int field, void* dst, size_t bytes) {
    return guarded([&] {
        require(b && (dst || bytes == 0), "null argument");
        b->read_field(sid, field, dst, bytes);
    });
}

int tss_batch_conservation(tss_batch* b, int sid, double out[5], double err[5]) {
    return guarded([&] {
        require(b && out && err, "null argument");
        b->conservation(sid, out, err);
    });
}

int tss_batch_synchronize(tss_batch* b) {
    return guarded([&] {
        require(b, "null argument");
        HIP_CHECK(hipStreamSynchronize(b->stream));
    });
}

int tss_batch_num_scenes(tss_batch* b, int* n) {
    return guarded([&] {
        require(b && n, "null argument");
        int k = 0;
        for (auto& s : b->scenes) k += s.alive;
        *n = k;
    });
}

static void tune_to_abi(const tss::PairTune* t, tss_tuning* o) {
    for (int k = 0; k < tss::kPairKernels; ++k) {
        o->block[k] = t[k].block;
        o->depth[k] = t[k].depth;
        o->xcd_order[k] = t[k].swz;
        o->grid[k] = t[k].grid;
    }
}

static void tune_from_abi(const tss_tuning* o, tss::PairTune* t) {
    for (int k = 0; k < tss::kPairKernels; ++k)
        t[k] = {o->block[k], o->depth[k], o->xcd_order[k], o->grid[k]};
}

int tss_batch_get_tuning(tss_batch* b, tss_tuning* out) {
    return guarded([&] {
        require(b && out, "null argument");
        tune_to_abi(b->tune, out);
    });
}

int tss_batch_set_tuning(tss_batch* b, const tss_tuning* t) {
    return guarded([&] {
        require(b && t, "null argument");
        tss::PairTune pt[tss::kPairKernels];
        tune_from_abi(t, pt);
        b->set_tuning(pt);
    });
}

int tss_batch_autotune(tss_batch* b, int steps_per_trial, int rounds, tss_tuning* out,
                       int* steps_taken, float* trial_log, int log_rows, int* n_log) {
    return guarded([&] {
        require(b && (trial_log || log_rows == 0), "null argument");
        int taken = b->autotune(steps_per_trial, rounds, trial_log, log_rows, n_log);
        if (out) tune_to_abi(b->tune, out);
        if (steps_taken) *steps_taken = taken;
    });
}

int tss_batch_set_grid_sort(tss_batch* b, int path) {
    return guarded([&] {
        require(b, "null argument");
        require(path >= TSS_SORT_AUTO && path <= TSS_SORT_SMALL, "unknown sort path");
        long long cap = 0;
        for (const auto& sc : b->scenes)
            if (sc.alive) cap += sc.lcap;
        require(path != TSS_SORT_SMALL || 5 * std::max<long long>(cap, 1) <= tss::kSmallSortKeys,
                "the one-workgroup sort holds at most 65536 keys (13107 low ids)");
        b->sort_path = path;
        b->update_shape();
    });
}

int tss_set_grid_sort(tss_solver* s, int path) {
    return s ? tss_batch_set_grid_sort(&s->b, path) : tss_batch_set_grid_sort(nullptr, path);
}

int tss_get_tuning(tss_solver* s, tss_tuning* out) {
    return guarded([&] {
        require(s && out, "null argument");
        tune_to_abi(s->b.tune, out);
    });
}

int tss_set_tuning(tss_solver* s, const tss_tuning* t) {
    return guarded([&] {
        require(s && t, "null argument");
        tss::PairTune pt[tss::kPairKernels];
        tune_from_abi(t, pt);
        s->b.set_tuning(pt);
    });
}

int tss_autotune(tss_solver* s, int steps_per_trial, int rounds, tss_tuning* out,
                 int* steps_taken, float* trial_log, int log_rows, int* n_log) {
    return guarded([&] {
        require(s && (trial_log || log_rows == 0), "null argument");
        int taken = s->b.autotune(steps_per_trial, rounds, trial_log, log_rows, n_log);
        if (out) tune_to_abi(s->b.tune, out);
        if (steps_taken) *steps_taken = taken;
    });
}

void tss_batch_destroy(tss_batch* b) { delete b; }

// ---------------------------------------------------------------------------
// Serving loop: continuous batching of scene requests on one batch
// ---------------------------------------------------------------------------
namespace {

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

long long scene_bytes(const tss_config& c, int cap_low, int n_bnd) {
    tss_batch_desc ds{};
    ds.max_scenes = 1;
    ds.cap_low = cap_low;
    ds.cap_bnd = n_bnd;
    ds.cells_low = (long long)c.domain_dims[0] * c.domain_dims[1];
    ds.cells_high = (long long)c.domain_dims_high[0] * c.domain_dims_high[1];
    return tss_batch::bytes_for(ds);
}

struct Resident {
    int req;         // request index
    int sid;         // scene slot in the batch
    int left;        // steps still to take
    long long bytes; // budget accounting
};

}  // namespace

long long tss_scene_bytes(const tss_config* cfg, int cap_low, int n_bnd) {
    if (!cfg || cap_low < 1 || n_bnd < 0) return -1;
    return scene_bytes(*cfg, cap_low, n_bnd);
}

int tss_serve(const tss_batch_desc* cap, long long budget, tss_request* req, int n,
              tss_serve_stats* st) {
    return guarded([&] {
        require(cap && st && (req || n == 0) && n >= 0 && budget > 0, "invalid argument");
        std::memset(st, 0, sizeof(*st));
        auto b = std::make_unique<tss_batch>();
        b->create(*cap);
        st->batch_bytes = tss_batch::bytes_for(*cap);
        st->budget_bytes = budget;
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return req[x].arrival_s < req[y].arrival_s; });
        int done = 0;
        for (int i = 0; i < n; ++i) {
            tss_request& q = req[i];
            q.status = TSS_E_STATE;  // pending
            q.admit_s = q.finish_s = -1.0;
            q.particle_steps = 0;
            if (q.steps < 1 || !q.cfg || (!q.low_xy && !q.gen && q.n_low > 0)) {
                q.status = TSS_E_INVALID;
                ++done;
            }
        }
        std::vector<Resident> res;
        std::deque<int> queue;  // arrived, waiting for room
        std::vector<double> step_ms;
        size_t next = 0;
        long long admitted = 0;
        int stable = 0;  // ticks since the resident set last changed
        // a stable resident set advances several steps per tick (graph
        // pairs, one host wait), about kTickMs of device time, so that
        // arrivals wait at most about that long for admission
        constexpr double kTickMs = 2.0;
        double step_est_ms = 0.0;  // the last tick's time per step
        double busy = 0.0, last_finish = 0.0;
        const auto t0 = std::chrono::steady_clock::now();
        auto now = [&] {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        while (done < n) {
            const double t = now();
            while (next < order.size() && req[order[next]].arrival_s <= t) {
                const int r = order[next++];
                if (req[r].status == TSS_E_STATE) queue.push_back(r);
            }
            // admission, in arrival order; a request that does not fit waits
            bool changed = false;
            for (auto it = queue.begin(); it != queue.end();) {
                tss_request& q = req[*it];
                const int nbq = q.gen ? tss_canvas_count(q.gen->nxd, q.gen->nyd, q.gen->layers)
                                      : q.n_bnd;
                const long long sb = scene_bytes(*q.cfg, q.cap_low, nbq < 0 ? 0 : nbq);
                if (sb > budget || sb > st->batch_bytes) {  // can never fit
                    q.status = TSS_E_NOMEM;
                    q.finish_s = now();
                    ++done;
                    it = queue.erase(it);
                    continue;
                }
                if (admitted + sb > budget) {
                    ++it;
                    continue;
                }
                int sid = -1;
                try {
                    if (q.gen) {
                        std::vector<float> walls = generated_walls(*q.gen);
                        sid = b->add_scene(*q.cfg, nullptr, q.gen->nx * q.gen->ny, q.cap_low,
                                           walls.data(), (int)(walls.size() / 2), q.gen);
                    } else {
                        sid = b->add_scene(*q.cfg, q.low_xy, q.n_low, q.cap_low, q.bnd_xy,
                                           q.n_bnd);
                    }
                } catch (const TssError& e) {
                    if (e.code == TSS_E_NOMEM && !res.empty()) {  // wait for room
                        ++it;
                        continue;
                    }
                    q.status = e.code;
                    q.finish_s = now();
                    ++done;
                    it = queue.erase(it);
                    continue;
                }
                q.admit_s = now();
                admitted += sb;
                res.push_back({*it, sid, q.steps, sb});
                it = queue.erase(it);
                changed = true;
            }
            if (res.empty()) {
                if (next >= order.size()) {
                    // nothing resident, nothing to come: the queue cannot
                    // be admitted (it was rejected above) -- done
                    if (queue.empty()) break;
                    for (int r : queue) {
                        req[r].status = TSS_E_NOMEM;
                        req[r].finish_s = now();
                        ++done;
                    }
                    queue.clear();
                    break;
                }
                // idle until the next arrival
                const double wait = req[order[next]].arrival_s - now();
                if (wait > 0)
                    std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
                continue;
            }
            st->max_resident = std::max(st->max_resident, (int)res.size());
            st->peak_admitted_bytes = std::max(st->peak_admitted_bytes, admitted);
            if (changed) stable = 0;
            int minleft = res[0].left;
            for (auto& r : res) minleft = std::min(minleft, r.left);
            // graph pairs once the resident set has held for two ticks; no
            // scene is advanced past its last step
            int k = 1;
            if (stable >= 2 && minleft >= 2) {
                const int want = step_est_ms > 0.0 ? (int)(kTickMs / step_est_ms) : 2;
                k = std::min(std::max(want, 2), minleft);
            }
            const double a = now();
            if (k >= 2) {
                b->step_graph(k);
                ++st->graph_ticks;
            } else {
                b->step_eager(1);
            }
            HIP_CHECK(hipStreamSynchronize(b->stream));
            const double e = now();
            busy += e - a;
            step_est_ms = (e - a) * 1e3 / k;
            for (int i = 0; i < k; ++i) step_ms.push_back(step_est_ms);
            ++st->ticks;
            ++stable;
            bool removed = false;
            for (auto it = res.begin(); it != res.end();) {
                it->left -= k;
                if (it->left > 0) {
                    ++it;
                    continue;
                }
                tss_request& q = req[it->req];
                for (int f = 0; f < TSS_FIELD_COUNT; ++f)
                    if (q.out[f]) b->read_field(it->sid, f, q.out[f], q.out_bytes[f]);
                tss_counts_t c;
                b->counts(it->sid, &c);
                q.particle_steps = c.particle_steps;
                q.finish_s = e;
                q.status = TSS_OK;
                st->particle_steps += c.particle_steps;
                last_finish = e;
                b->remove_scene(it->sid);
                admitted -= it->bytes;
                ++done;
                removed = true;
                it = res.erase(it);
            }
            if (removed) stable = 0;
        }
        st->wall_s = last_finish;
        st->busy_s = busy;
        st->particle_steps_per_s = last_finish > 0 ? (double)st->particle_steps / last_finish : 0.0;
        std::vector<double> rl, ql;
        for (int i = 0; i < n; ++i) {
            if (req[i].status != TSS_OK) continue;
            rl.push_back((req[i].finish_s - req[i].arrival_s) * 1e3);
            ql.push_back((req[i].admit_s - req[i].arrival_s) * 1e3);
        }
        st->step_ms_p50 = percentile(step_ms, 0.50);
        st->step_ms_p99 = percentile(step_ms, 0.99);
        st->request_ms_p50 = percentile(rl, 0.50);
        st->request_ms_p99 = percentile(rl, 0.99);
        st->queue_ms_p50 = percentile(ql, 0.50);
        st->queue_ms_p99 = percentile(ql, 0.99);
    });
}

int tss_bandwidth_probe(size_t bytes, int reps, double* gbs) {
    return guarded([&] {
        require(gbs && bytes >= 16 && reps >= 1, "invalid argument");
        const size_t n16 = bytes / 16;
        void *a = nullptr, *b = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        try {
            HIP_CHECK(hipMalloc(&a, n16 * 16));
            HIP_CHECK(hipMalloc(&b, n16 * 16));
            HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            HIP_CHECK(hipMemsetAsync(a, 1, n16 * 16, st));
            tss::launch_copy16(a, b, (long long)n16, st);  // warm
            HIP_CHECK(hipEventRecord(e0, st));
            for (int i = 0; i < reps; ++i) tss::launch_copy16(a, b, (long long)n16, st);
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            *gbs = 2.0 * (double)(n16 * 16) * reps / (ms * 1e-3) / 1e9;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int tss_gather_probe(size_t bytes, int stride_bytes, int reps, double* gbs) {
    return guarded([&] {
        require(gbs && reps >= 1 && stride_bytes >= 8 && stride_bytes % 8 == 0 &&
                    bytes >= (size_t)stride_bytes,
                "invalid argument");
        const int stride = stride_bytes / 8;
        const long long n = (long long)(bytes / (size_t)stride_bytes);
        void *a = nullptr, *sink = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (a) (void)hipFree(a);
            if (sink) (void)hipFree(sink);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        try {
            HIP_CHECK(hipMalloc(&a, (size_t)n * (size_t)stride_bytes));
            HIP_CHECK(hipMalloc(&sink, 8));
            HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            // every byte 0x5A: the words XOR to 0 or 0x5A...5A, never the
            // probe's sink trigger
            HIP_CHECK(hipMemsetAsync(a, 0x5A, (size_t)n * (size_t)stride_bytes, st));
            tss::launch_gather_probe(a, n, stride, sink, st);  // warm
            HIP_CHECK(hipEventRecord(e0, st));
            for (int i = 0; i < reps; ++i) tss::launch_gather_probe(a, n, stride, sink, st);
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            *gbs = 8.0 * (double)n * reps / (ms * 1e-3) / 1e9;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

const char* tss_last_error(void) { return g_last_error.c_str(); }

const char* tss_version(void) { return "tss 0.1.0 gfx950"; }

}  // extern "C"
