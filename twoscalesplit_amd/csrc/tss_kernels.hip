// tss_kernels.hip -- HIP kernels (gfx950) of the two-scale WCSPH step.
//
// Reference: segfault11/TwoScaleSplit, TwoScale2D/WCSPHSolver.cu ("R:" lines).
// Compiled with -ffp-contract=off: every expression that decides an index or
// a set (cell coordinates, pair distances, cutoffs, integration, child
// placement) rounds each operation exactly like the FP32 method definition;
// the physics sums use explicit fmaf where a fused operation is wanted.
//
// Step (one launch per phase for all scenes of a batch):
//   fill -> hash -> radix sort (P passes x {hist, scan, scatter})
//   -> ranges+reorder -> density -> force+integrate -> reduce
//   -> split candidates -> groups -> cap -> compaction (3 kernels)
//   -> finalize
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "tss_device.h"

namespace tss {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
#define GRID_STRIDE(k, n)                                             \
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < (n);     \
         k += gridDim.x * blockDim.x)

constexpr uint8_t kActive = 0, kTransient = 1, kDeleted = 2;

// Grid coordinate along one axis: clamp((int)(((x+off)-o)/cell), 0, dim-1)
// (R:241-289); the clamp is taken in float so every input is defined.
__device__ __forceinline__ int axis_cell(float x, float off, float o,
                                         float cell, int dim) {
    float t = ((x + off) - o) / cell;
    if (!(t >= 1.0f)) return 0;
    if (t >= (float)dim) return dim - 1;
    return (int)t;
}

// Squared pair distance, each operation rounded (R:325-333 without sqrt:
// sqrt(r2) <= R  <=>  r2 <= T(R) for the precomputed threshold T).
__device__ __forceinline__ float dist2(float2 a, float2 b) {
    float dx = a.x - b.x;
    float dy = a.y - b.y;
    float p = dx * dx;
    float q = dy * dy;
    return p + q;
}

// Gather of element i of a device array through a 32-bit byte offset: with
// the base in scalar registers this is one 32-bit shift per load instead of
// 64-bit address arithmetic (batches keep every such array below 4 GiB).
template <class T>
__device__ __forceinline__ T gather(const T* __restrict__ base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       (size_t)(i * (uint32_t)sizeof(T)));
}

// Element j + i through the 32-bit byte offset of j, i a small constant: the
// loads of j + 0, 1, ... share one address and differ in their immediate
// offsets.
template <class T>
__device__ __forceinline__ T gather_at(const T* __restrict__ base, uint32_t j, int i) {
    const T* p = reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                            (size_t)(j * (uint32_t)sizeof(T)));
    return p[i];
}

// 0 < r2 <= t as one unsigned comparison of the bit patterns (r2 is a sum
// of squares, so never negative; NaN compares above every finite t).
__device__ __forceinline__ bool in_open_ball(float r2, float t) {
    return __float_as_uint(r2) - 1u < __float_as_uint(t);
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Order-independent id checksum of neighbour sets: (id + 1) * 2654435761
// mod 2^32, summed over a set (ids scene-local).
__device__ __forceinline__ uint32_t csum(int id) {
    return (uint32_t)(id + 1) * 2654435761u;
}

// Scene-local id of a unified particle id u (low ids from lbase, children
// from 4 lbase).
__device__ __forceinline__ int local_id(const DevState& d, const SceneDev& sc, int u) {
    return u < d.CL ? u - sc.lbase : (u - d.CL) - 4 * sc.lbase;
}

struct Box {
    int x0, x1, y0, y1;
};

__device__ __forceinline__ Box scan_box(const SceneDev& sc, bool high, float2 x,
                                        float r) {
    const float cell = high ? sc.H : sc.h;
    const int dx = high ? sc.dhx : sc.dlx;
    const int dy = high ? sc.dhy : sc.dly;
    Box b;
    b.x0 = axis_cell(x.x, -r, sc.ox, cell, dx);
    b.x1 = axis_cell(x.x, r, sc.ox, cell, dx);
    b.y0 = axis_cell(x.y, -r, sc.oy, cell, dy);
    b.y1 = axis_cell(x.y, r, sc.oy, cell, dy);
    return b;
}

// Contiguous sorted range covering cells x0..x1 of one grid row (cells are
// row-major, so neighbouring cells of a row are adjacent in sorted order).
// The entries of four cells are loaded together, then combined in order.
__device__ __forceinline__ void row_run(const int* __restrict__ cs,
                                        const int* __restrict__ ce, int row,
                                        int x0, int x1, int& s, int& e) {
    s = -1;
    e = -1;
    for (int c0 = x0; c0 <= x1; c0 += 4) {
        int st[4], en[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int cx = row + min(c0 + k, x1);
            st[k] = cs[cx];
            en[k] = ce[cx];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (c0 + k <= x1 && st[k] >= 0) {
                if (s < 0) s = st[k];
                e = en[k];
            }
        }
    }
}

// Blended M4 kernel (R:116-239, 358-361): W_bar = (W_h + W_H)/2 and the
// factor G with grad W_bar = G/r * x_ij. Branch-free: with a=max(2-q,0),
// b=max(1-q,0), f = a^3 - 4b^3 and f' = a^2 - 4b^2 on all of q >= 0.
__device__ __forceinline__ void m4_blend(const SceneDev& sc, float r, float& w,
                                         float& g) {
    float qh = r * sc.q_h, qH = r * sc.q_H;
    float ah = fmaxf(2.0f - qh, 0.0f), bh = fmaxf(1.0f - qh, 0.0f);
    float aH = fmaxf(2.0f - qH, 0.0f), bH = fmaxf(1.0f - qH, 0.0f);
    float ah2 = ah * ah, bh2 = bh * bh, aH2 = aH * aH, bH2 = bH * bH;
    float fh = fmaf(-4.0f * bh, bh2, ah2 * ah);
    float fH = fmaf(-4.0f * bH, bH2, aH2 * aH);
    float dh = fmaf(-4.0f, bh2, ah2);
    float dH = fmaf(-4.0f, bH2, aH2);
    w = fmaf(sc.sig_h2, fh, sc.sig_H2 * fH);
    g = fmaf(sc.gc_h2, dh, sc.gc_H2 * dH);
}

// m4_blend for pairs with q_h = 2r/h <= 1 (a high target's pairs at r <= H
// when 2H <= h): W_h takes its inner branch, (2-q)^3 - 4(1-q)^3 =
// 4 - 6q^2 + 3q^3, with a^2 - 4b^2 = q(4 - 3q); W_H as in m4_blend. Rounding
// can put q an ulp or two above 1, where both branches agree to first
// order.
__device__ __forceinline__ void m4_blend_inner(const SceneDev& sc, float r, float& w,
                                               float& g) {
    float qh = r * sc.q_h, qH = r * sc.q_H;
    float aH = fmaxf(2.0f - qH, 0.0f), bH = fmaxf(1.0f - qH, 0.0f);
    float aH2 = aH * aH, bH2 = bH * bH;
    float fh = fmaf(qh * qh, fmaf(3.0f, qh, -6.0f), 4.0f);
    float fH = fmaf(-4.0f * bH, bH2, aH2 * aH);
    float dh = qh * fmaf(-3.0f, qh, 4.0f);
    float dH = fmaf(-4.0f, bH2, aH2);
    w = fmaf(sc.sig_h2, fh, sc.sig_H2 * fH);
    g = fmaf(sc.gc_h2, dh, sc.gc_H2 * dH);
}

// W_bar and G of a high target's own-level pairs (r <= H) when H = h/2
// exactly, so that sigma_H = 4 sigma_h, g_H = 8 g_h and q_H = 2 q_h hold in
// FP32 (hh_half): the two M4 kernels of R:116-239 summed and expanded in
// q = 2r/h <= 1 and u = 1 - q,
//   W_bar = sigma_h/2 [(1 + 3u + 3u^2 + 29u^3) - 16 max(2u - 1, 0)^3],
//   G = g_h/2 q (68 - 99 q) for q < 1/2, g_h/2 (1 + u (2 + 29 u)) above,
// each form where it has no cancellation (ten and seven operations against
// sixteen and thirteen for the two kernels separately; as accurate in FP32).
__device__ __forceinline__ float wbar_half(const SceneDev& sc, float r) {
    const float u = fmaf(-r, sc.q_h, 1.0f);
    const float p = fmaf(fmaf(fmaf(29.0f, u, 3.0f), u, 3.0f), u, 1.0f);
    const float v = fmaxf(fmaf(2.0f, u, -1.0f), 0.0f);
    return sc.sig_h2 * fmaf(-16.0f, v * v * v, p);
}

__device__ __forceinline__ float gbar_half(const SceneDev& sc, float r) {
    const float q = r * sc.q_h;
    const float u = 1.0f - q;
    const float ga = q * fmaf(-99.0f, q, 68.0f);
    const float gb = fmaf(u, fmaf(29.0f, u, 2.0f), 1.0f);
    return sc.gc_h2 * (q < 0.5f ? ga : gb);
}

// gbar_half without its coefficient gc_h2 (the force passes fold it into
// the pair's mass factor).
__device__ __forceinline__ float gbar_half_unit(const SceneDev& sc, float r) {
    const float q = r * sc.q_h;
    const float u = 1.0f - q;
    const float ga = q * fmaf(-99.0f, q, 68.0f);
    const float gb = fmaf(u, fmaf(29.0f, u, 2.0f), 1.0f);
    return q < 0.5f ? ga : gb;
}

// Kernel forms, chosen per scene: KM 0 m4_blend, 1 m4_blend_inner (hh_inner),
// 2 wbar_half / gbar_half (hh_half).
template <int KM>
__device__ __forceinline__ void m4_pick(const SceneDev& sc, float r, float& w, float& g) {
    if (KM == 2) {
        w = wbar_half(sc, r);
        g = gbar_half(sc, r);
    } else if (KM == 1) {
        m4_blend_inner(sc, r, w, g);
    } else {
        m4_blend(sc, r, w, g);
    }
}

// Wall weight Gamma(r)/(mass factor) (R:66-114), qs = 2/ell.
__device__ __forceinline__ float wall_gamma(const SceneDev& sc, float r2,
                                            float r, float qs) {
    float q = r * qs;
    float c = sc.cwall / r2;
    float v;
    if (q < 2.0f / 3.0f) {
        v = 2.0f / 3.0f;
    } else if (q < 1.0f) {
        v = q * fmaf(-1.5f, q, 2.0f);
    } else if (q < 2.0f) {
        float a = 2.0f - q;
        v = 0.5f * a * a;
    } else {
        v = 0.0f;
    }
    return c * v;
}

// Tait pressure from a compensated density hi+lo (R:1421-1423 restated as
// B((1+d)^7 - 1) with d = ((hi - rho0) + lo)/rho0, binomially expanded so
// small |d| keeps its relative accuracy).
__device__ __forceinline__ float tait(const SceneDev& sc, float hi, float lo) {
    float d = ((hi - sc.rho0) + lo) / sc.rho0;
    float poly = fmaf(d, 1.0f, 7.0f);
    poly = fmaf(d, poly, 21.0f);
    poly = fmaf(d, poly, 35.0f);
    poly = fmaf(d, poly, 35.0f);
    poly = fmaf(d, poly, 21.0f);
    poly = fmaf(d, poly, 7.0f);
    return sc.B * d * poly;
}

// hi+lo of m1*(s1-c1) + m2*(s2-c2) where (s, c) are Kahan sums.
__device__ __forceinline__ void mass_sum(float m1, float s1, float c1,
                                         float m2, float s2, float c2,
                                         float& hi, float& lo) {
    float h1 = m1 * s1;
    float l1 = fmaf(m1, s1, -h1) - m1 * c1;
    float h2 = m2 * s2;
    float l2 = fmaf(m2, s2, -h2) - m2 * c2;
    float s = h1 + h2;
    float bb = s - h1;
    float e = (h1 - (s - bb)) + (h2 - bb);
    hi = s;
    lo = (l1 + l2) + e;
}

__device__ __forceinline__ float fast_rsq(float x) {
    return __builtin_amdgcn_rsqf(x);
}

__device__ __forceinline__ float fast_rcp(float x) {
    return __builtin_amdgcn_rcpf(x);
}

// |x_ij| from r2 >= 0: the hardware square root (one instruction, exact
// at r2 = 0, the self pair of the density sums).
__device__ __forceinline__ float dist_from2(float r2) {
    return __builtin_amdgcn_sqrtf(r2);
}

// Streaming stores (not kept in L2 for reuse): a kernel's results, written
// once and read by later kernels.
__device__ __forceinline__ void nt_store(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void nt_store(float2* p, float x, float y) {
    float* q = reinterpret_cast<float*>(p);
    __builtin_nontemporal_store(x, q);
    __builtin_nontemporal_store(y, q + 1);
}
__device__ __forceinline__ void nt_store(float4* p, float x, float y, float z, float w) {
    float* q = reinterpret_cast<float*>(p);
    __builtin_nontemporal_store(x, q);
    __builtin_nontemporal_store(y, q + 1);
    __builtin_nontemporal_store(z, q + 2);
    __builtin_nontemporal_store(w, q + 3);
}

__device__ __forceinline__ int low_of(const DevState& d, int u) {
    return u < d.CL ? u : (u - d.CL) >> 2;
}

// ---------------------------------------------------------------------------
// Fill kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_int(int* __restrict__ p, int v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Start of a step, one launch: blocks [0, nclear) clear the cell tables of
// the resident scenes (block b the cells of range clear_blk[b]) and reset
// the per-scene step accumulators; the other blocks hash (R:1246-1298):
// key = unified cell index of the particle. The two parts touch disjoint
// data; the cell tables are next written by the grid build's last kernels.
// count = 1 (counting sort by cell): the cell ranges are rewritten whole by
// k_cell_ranges, so they are not cleared here; each key is counted in its
// cell (one atomic per run of equal keys along a wave's lanes) and the ids
// are not copied (the scatter reads them from the lists). solo: the slot of
// the only resident scene, or -1 (the scene then comes from scene_of).
__global__ void k_begin_hash(DevState d, int nclear, int count, int solo,
                             const int* __restrict__ list_l,
                             const int* __restrict__ list_h,
                             uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
    if ((int)blockIdx.x < nclear) {
        const int2 r = d.clear_blk[blockIdx.x];
        for (int i = r.x + (int)threadIdx.x; i < r.y; i += (int)blockDim.x) {
            if (!count) {
                d.cstart[i] = -1;
                d.cend[i] = -1;
            }
            d.cflag[2 * i] = 0;
            d.cflag[2 * i + 1] = 0;
            d.csurf[i] = 0;
            if (i < d.CCL) d.cband[i] = 0;
        }
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < d.S;
             s += nclear * blockDim.x) {
            SceneCount& c = d.scnt[s];
            c.n_split = 0;
            c.n_promote = 0;
            c.n_merge = 0;
            c.n_dissolve = 0;
            c.cand = 0;
            c.seg_l0 = 0;
            c.seg_l1 = 0;
            c.seg_h0 = 0;
            c.seg_h1 = 0;
            c.kl0 = c.kl1 = c.kh0 = c.kh1 = c.sp0 = c.sp1 = 0;
            c.prom = 0;
            c.act = 0;
            c.hsurf = 0;
        }
        return;
    }
    const int nl = d.counts->n_l;
    const int n = nl + d.counts->n_h;
    const int stride = (gridDim.x - nclear) * blockDim.x;
    const int lane = threadIdx.x & 63;
    // block-uniform trip count: the run votes below see whole waves, whose
    // valid lanes are a prefix
    for (int b0 = (blockIdx.x - nclear) * blockDim.x; b0 < n; b0 += stride) {
        const int k = b0 + (int)threadIdx.x;
        const bool valid = k < n;
        uint32_t key = 0u;
        if (valid) {
            const int u = k < nl ? list_l[k] : list_h[k - nl];
            const bool high = u >= d.CL;
            // with one resident scene, no lookup by id
            const int s = solo >= 0 ? solo : d.scene_of[low_of(d, u)];
            const SceneDev& sc = d.scenes[s];
            const float2 x = d.pos[u];
            if (!high) {
                int cx = axis_cell(x.x, 0.0f, sc.ox, sc.h, sc.dlx);
                int cy = axis_cell(x.y, 0.0f, sc.oy, sc.h, sc.dly);
                key = (uint32_t)(sc.cbase_l + cy * sc.dlx + cx);
            } else {
                int cx = axis_cell(x.x, 0.0f, sc.ox, sc.H, sc.dhx);
                int cy = axis_cell(x.y, 0.0f, sc.oy, sc.H, sc.dhy);
                key = (uint32_t)(sc.cbase_h + cy * sc.dhx + cx);
            }
            keys[k] = key;
            if (!count) vals[k] = (uint32_t)u;
        }
        if (count) {
            const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 64);
            const bool start = valid && (lane == 0 || prev != key);
            const uint64_t starts = __ballot(start);
            const int nvalid = __popcll(__ballot(valid));
            if (start) {
                const uint64_t later = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
                const int next = later ? __ffsll((long long)later) - 1 : nvalid;
                atomicAdd(&d.ccount[key], next - lane);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Counting sort by cell (replaces thrust::sort_by_key, R:2828-2834, with the
// same stable order: by cell, ties by list position):
//   k_begin_hash (count = 1) counts the keys per cell ->
//   k_cell_sums (per tile of kCellTile cells) ->
//   k_cell_ranges: each block sums the earlier tiles' sums (L2-resident,
//   one read per tile) for its base; cstart = exclusive prefix of the
//   counts (-1 for an empty cell), cend = cstart (the scatter's fill
//   pointer), scene segments ->
//   k_cell_scatter: list positions into their cell's range (atomic fill,
//   any order within a cell; cend ends at cstart + count) ->
//   k_cell_reorder: each entry's rank among its cell's list positions gives
//   its sorted position; then the cell-sorted copies.
// The order within a cell is fixed by the rank, so the result is
// deterministic and equal to the radix sort's. The rank costs one read per
// entry of the cell, as a target's neighbour scan does per candidate.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kCellThreads) void k_cell_sums(DevState d, int ncell) {
    __shared__ int wsum[kCellThreads / 64];
    const int c0 = blockIdx.x * kCellTile + (int)threadIdx.x;
    int v = 0;
#pragma unroll
    for (int i = 0; i < kCellTile / kCellThreads; ++i) {
        const int c = c0 + i * kCellThreads;
        v += c < ncell ? d.ccount[c] : 0;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kCellThreads / 64; ++w) t += wsum[w];
        d.ctile[blockIdx.x] = t;
    }
}

// Per tile: the counts to LDS (coalesced), each thread's kCellTile /
// kCellThreads consecutive cells scanned in registers, the thread totals
// scanned over the block, the prefixes written back through LDS. Resets the
// counts for the next grid build. A scene's low and high cells are each one
// range of the unified cells (low ranges below CCL, high ranges above), so
// a cell whose owner differs from its neighbour's, or that sits at either
// side of CCL, bounds a range: it writes that scene's sorted segment bound.
__global__ __launch_bounds__(kCellThreads) void k_cell_ranges(DevState d, int ncell) {
    constexpr int kPer = kCellTile / kCellThreads;  // 4 cells per thread
    // one pad word per kPer: thread t's run starts at bank (t * 5) mod 32
    __shared__ int cells[kCellTile + kCellTile / kPer];
    __shared__ int owner[kCellTile];  // cscene of the tile's cells
    __shared__ int wsum[kCellThreads / 64];
    __shared__ int bsum[kCellThreads / 64];
    __shared__ int tbase;
    const int t = (int)threadIdx.x;
    const int base = blockIdx.x * kCellTile;
    auto slot = [](int x) { return x + x / kPer; };
    // the tile's base: the sum of the earlier tiles' sums
    {
        int v = 0;
        for (int q = t; q < (int)blockIdx.x; q += kCellThreads) v += d.ctile[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((t & 63) == 0) bsum[t >> 6] = v;
    }
    int cnt[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int x = i * kCellThreads + t;
        const int c = base + x;
        cnt[i] = c < ncell ? d.ccount[c] : 0;
        cells[slot(x)] = cnt[i];
        owner[x] = c < ncell ? d.cscene[c] : -1;
    }
    __syncthreads();
    if (t == 0) {
        int b = 0;
        for (int w = 0; w < kCellThreads / 64; ++w) b += bsum[w];
        tbase = b;
    }
    int run[kPer];
    int tot = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        run[j] = tot;
        tot += cells[slot(t * kPer + j)];
    }
    // exclusive scan of the thread totals: within the wave, then the waves
    const int lane = t & 63;
    int inc = tot;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[t >> 6] = inc;
    __syncthreads();
    int woff = tbase;
    for (int w = 0; w < (t >> 6); ++w) woff += wsum[w];
    const int toff = woff + inc - tot;
#pragma unroll
    for (int j = 0; j < kPer; ++j) cells[slot(t * kPer + j)] = toff + run[j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int x = i * kCellThreads + t;
        const int c = base + x;
        if (c >= ncell) continue;
        const int ex = cells[slot(x)];
        const int v = cnt[i] > 0 ? ex : -1;
        d.cstart[c] = v;
        d.cend[c] = v;
        if (cnt[i] != 0) d.ccount[c] = 0;
        const int s = owner[x];
        if (s >= 0) {
            const bool high = c >= d.CCL;
            const int sp = x > 0 ? owner[x - 1] : (c > 0 ? d.cscene[c - 1] : -1);
            const int sn = x + 1 < kCellTile ? owner[x + 1]
                                             : (c + 1 < ncell ? d.cscene[c + 1] : -1);
            const bool first = c == 0 || c == d.CCL || sp != s;
            const bool last = c == ncell - 1 || c == d.CCL - 1 || sn != s;
            SceneCount& sct = d.scnt[s];
            if (first) {
                if (high) sct.seg_h0 = ex; else sct.seg_l0 = ex;
            }
            if (last) {
                if (high) sct.seg_h1 = ex + cnt[i]; else sct.seg_l1 = ex + cnt[i];
            }
        }
    }
}

// List positions into their cells' ranges: one atomic per run of equal keys
// along a wave's lanes (lists are kept in cell order, so a cell's entries
// mostly come in one run). A slot holds the entry's list position and id
// (read here in list order); its key goes to the same slot of the sorted
// keys: a cell's slots all hold its key, whatever the order within the cell.
__global__ void k_cell_scatter(DevState d, const uint32_t* __restrict__ keys,
                               const int* __restrict__ list_l,
                               const int* __restrict__ list_h,
                               uint32_t* __restrict__ skeys) {
    const int nl = d.counts->n_l;
    const int n = nl + d.counts->n_h;
    const int stride = gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
        const int k = b0 + (int)threadIdx.x;
        const bool valid = k < n;
        uint32_t key = 0u;
        int u = 0;
        if (valid) {
            key = keys[k];
            u = k < nl ? list_l[k] : list_h[k - nl];
        }
        const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 64);
        const bool start = valid && (lane == 0 || prev != key);
        const uint64_t starts = __ballot(start);
        const int nvalid = __popcll(__ballot(valid));
        int pos = 0;
        if (start) {
            const uint64_t later = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
            const int next = later ? __ffsll((long long)later) - 1 : nvalid;
            pos = atomicAdd(&d.cend[key], next - lane);
        }
        // the run's leader: the highest start lane at or below this lane
        const uint64_t upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int leader = upto ? 63 - __clzll((long long)upto) : lane;
        pos = __shfl(pos, leader, 64) + (lane - leader);
        if (valid) {
            d.cslot[pos] = make_uint2((uint32_t)k, (uint32_t)u);
            skeys[pos] = key;
        }
    }
}

// Sorted position of the entry in slot p: its cell's start plus the number
// of the cell's list positions below its own. A cell whose slots all lie in
// the wave's 64 is ranked through lane shuffles (the cell's lanes are one
// run of equal sorted keys); a cell reaching past the wave's slots reads its
// range from the cell table. Then the cell-sorted copies (as k_reorder):
// ids, positions, velocities, states, the id -> sorted index map, scenes,
// cell state flags, scene flags. The id's data is loaded before the rank is
// formed, so its latency overlaps the ranking.
__global__ void k_cell_reorder(DevState d, const uint32_t* __restrict__ skeys,
                               uint32_t* __restrict__ svals, int with_sidx) {
    const int nl = d.counts->n_l;
    const int n = nl + d.counts->n_h;
    const int stride = gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
        const int wb = b0 + (int)(threadIdx.x & ~63u);  // the wave's first slot
        const int p = wb + lane;
        const bool valid = p < n;
        const uint2 eu = valid ? d.cslot[p] : make_uint2(0u, 0u);
        const int e = (int)eu.x;
        const int u = (int)eu.y;
        const uint32_t key = valid ? skeys[p] : 0xFFFFFFFFu;
        uint8_t sst = 0;
        float2 x = make_float2(0.f, 0.f), v = x;
        if (valid) {
            sst = d.state[u];
            x = d.pos[u];
            v = d.vel[u];
        }
        const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 64);
        const bool start = valid && (lane == 0 || prev != key);
        const uint64_t starts = __ballot(start);
        const int nvalid = __popcll(__ballot(valid));
        const uint64_t upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int lead = upto ? 63 - __clzll((long long)upto) : lane;
        const uint64_t later = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
        const int next = later ? __ffsll((long long)later) - 1 : nvalid;
        // the run reaches past the wave's slots when its first lane is 0 and
        // the slot before the wave holds the same key, or its end is the
        // wave's end and the slot after holds the same key
        bool outside = false;
        if (valid && lead == 0 && wb > 0) outside = skeys[wb - 1] == key;
        if (valid && next == 64 && wb + 64 < n) outside = outside || skeys[wb + 64] == key;
        // rank among the run's lanes: a pass per lane offset, up to the
        // longest run in the wave
        const int len = valid ? next - lead : 0;
        int maxlen = 0;
#pragma unroll
        for (int b = 6; b >= 0; --b) {
            const int c = maxlen | (1 << b);
            if (__ballot(len >= c) != 0ull) maxlen = c;
        }
        int rank = 0;
        for (int j = 0; j < maxlen; ++j) {
            const int ej = __shfl(e, min(lead + j, 63), 64);
            rank += (j < len && ej < e) ? 1 : 0;
        }
        int k = wb + lead + rank;
        if (valid && outside) {
            // eight of the cell's list positions in flight at a time
            const int st = d.cstart[key], en = d.cend[key];
            rank = 0;
            for (int q0 = st; q0 < en; q0 += 8) {
                uint32_t t[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) t[i] = d.cslot[min(q0 + i, en - 1)].x;
#pragma unroll
                for (int i = 0; i < 8; ++i) rank += (q0 + i < en && (int)t[i] < e) ? 1 : 0;
            }
            k = st + rank;
        }
        if (!valid) continue;
        svals[k] = (uint32_t)u;
        d.cflag[2 * (size_t)key + (sst & 1u)] = 1;
        nt_store(&d.spos[k], x.x, x.y);
        nt_store(&d.sxv[k], x.x, x.y, v.x, v.y);
        d.sstate[k] = sst;
        if (with_sidx) d.sidx[u] = k;
        const int s = d.cscene[key];
        d.sscene[k] = (uint16_t)s;
        if (!(sst & 1u)) {
            if (k >= nl) d.scnt[s].prom = 1;
            else d.scnt[s].act = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, kRadixBits-bit digits (replaces thrust::sort_by_key,
// R:2828-2834, 2962-2968). Per pass: tile histograms -> per-digit scan over
// tiles -> stable scatter ranked with wavefront ballots.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(
    const uint32_t* __restrict__ keys, const DevCounts* __restrict__ cnt,
    int shift, uint32_t* __restrict__ hist, int max_tiles) {
    __shared__ uint32_t h[kRadix];
    const int n = cnt->n_l + cnt->n_h;
    const int tile = blockIdx.x;
    // tiles past the live keys (the grid is sized by the resident scenes'
    // capacity) are never read by the scan: the whole block exits
    if (tile * kSortTile >= n) return;
    const int base = tile * kSortTile;
    // all of the thread's keys first: their loads are in flight together
    uint32_t kk[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int e = base + i * kSortThreads + threadIdx.x;
        kk[i] = e < n ? keys[e] : 0u;
    }
    for (int i = threadIdx.x; i < kRadix; i += kSortThreads) h[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    {
        // keys arrive nearly sorted (lists are kept in cell order), so
        // equal digits come in runs along the lanes: one LDS atomic per
        // run, from the run's first lane (valid lanes are a prefix)
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const int e = base + i * kSortThreads + threadIdx.x;
            const bool valid = e < n;
            const uint32_t dg = valid ? (kk[i] >> shift) & (kRadix - 1u) : (uint32_t)kRadix;
            const uint32_t prev = (uint32_t)__shfl_up((int)dg, 1, 64);
            const bool start = valid && (lane == 0 || prev != dg);
            const uint64_t starts = __ballot(start);
            const int nvalid = __popcll(__ballot(valid));
            if (start) {
                const uint64_t later = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
                const int next = later ? __ffsll((long long)later) - 1 : nvalid;
                atomicAdd(&h[dg], (uint32_t)(next - lane));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kSortThreads)
        hist[(size_t)i * max_tiles + tile] = h[i];
}

// One block per digit: exclusive scan over the tiles of that digit.
__global__ __launch_bounds__(256) void k_sort_scan(
    uint32_t* __restrict__ hist, const DevCounts* __restrict__ cnt,
    int max_tiles, uint32_t* __restrict__ digit_tot) {
    __shared__ uint32_t part[256];
    const int n = cnt->n_l + cnt->n_h;
    const int nt = (n + kSortTile - 1) / kSortTile;
    uint32_t* row = hist + (size_t)blockIdx.x * max_tiles;
    const int per = (nt + 255) / 256;
    const int t0 = threadIdx.x * per;
    uint32_t local = 0;
    for (int i = 0; i < per; ++i) {
        int t = t0 + i;
        if (t < nt) local += row[t];
    }
    part[threadIdx.x] = local;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 partials
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - local;
    for (int i = 0; i < per; ++i) {
        int t = t0 + i;
        if (t < nt) {
            uint32_t v = row[t];
            row[t] = run;
            run += v;
        }
    }
    if (threadIdx.x == 255) digit_tot[blockIdx.x] = part[255];
}

// Stable scatter of one tile. Per item (256 consecutive keys): each wave
// ranks its lanes among equal digits with kRadixBits ballots; the first
// lane of each digit group (its leader) publishes the group's size, reads
// the sizes of the other waves' groups of that digit and the digit's
// running count in the tile, and passes the group's offset to its lanes.
// The work per item does not depend on the number of digits.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
    const DevCounts* __restrict__ cnt, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ digit_tot,
    int max_tiles) {
    constexpr int kWaves = kSortThreads / 64;
    constexpr int kPer = kRadix / kSortThreads;  // digits per thread
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t part[kSortThreads];
    const int n = cnt->n_l + cnt->n_h;
    const int tile = blockIdx.x;
    const int tbase = tile * kSortTile;
    if (tbase >= n) return;  // whole block exits together
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    // all of the thread's keys and values first: their loads are in flight
    // together (and behind the digit-base scan below)
    uint32_t kk[kSortItems], vv[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int e = tbase + it * kSortThreads + t;
        kk[it] = e < n ? kin[e] : 0u;
        vv[it] = e < n ? vin[e] : 0u;
    }
    // digit bases: exclusive scan of the digit totals + this tile's prefix
    {
        uint32_t loc[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            loc[q] = digit_tot[t * kPer + q];
            sum += loc[q];
        }
        part[t] = sum;
        __syncthreads();
        for (int off = 1; off < kSortThreads; off <<= 1) {
            uint32_t a = t >= off ? part[t - off] : 0u;
            __syncthreads();
            part[t] += a;
            __syncthreads();
        }
        uint32_t run = part[t] - sum;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int dgt = t * kPer + q;
            base[dgt] = run + hist[(size_t)dgt * max_tiles + tile];
            run += loc[q];
            running[dgt] = 0;
#pragma unroll
            for (int v = 0; v < kWaves; ++v) wcnt[v][dgt] = 0;
        }
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int e = tbase + it * kSortThreads + t;
        const bool valid = e < n;
        const uint32_t key = kk[it];
        const uint32_t val = vv[it];
        const uint32_t dg = (key >> shift) & (kRadix - 1u);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t bb = __ballot((dg >> b) & 1u);
            m &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rk = (uint32_t)__popcll(m & lt);
        const bool leader = valid && rk == 0;
        if (leader) wcnt[w][dg] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = 0, tot = 0;
        bool last = true;
        if (leader) {
#pragma unroll
            for (int v = 0; v < kWaves; ++v) {
                const uint32_t c = wcnt[v][dg];
                off += v < w ? c : 0u;
                tot += c;
                last = last && (v <= w || c == 0u);
            }
            off += running[dg];
        }
        __syncthreads();
        if (leader) {
            if (last) running[dg] += tot;  // the digit's last wave this item
            wcnt[w][dg] = 0;
        }
        // the group's offset, from its leader (the lowest lane of m)
        const int src = valid ? __ffsll((long long)m) - 1 : lane;
        off = (uint32_t)__shfl((int)off, src, 64);
        if (valid) {
            const uint32_t dst = base[dg] + off + rk;
            kout[dst] = key;
            vout[dst] = val;
        }
    }
}

// The same stable LSD sort for small batches, in one workgroup and one
// launch: per pass a histogram of the live keys (LDS atomics), the digit
// bases (one wave), then the keys in order, in chunks of one key per thread,
// each ranked like an item of k_sort_scatter (kRadixBits ballots within the
// wave, the earlier waves' counts of its digit through LDS). Passes read
// keys/vals[pass & 1] and write the other buffer, as the multi-tile sort
// does, so the result is in buffer (passes & 1) and identical to it.
__global__ __launch_bounds__(kSmallSortThreads) void k_sort_small(DevState d, int passes) {
    constexpr int kWaves = kSmallSortThreads / 64;
    __shared__ uint32_t hcount[kRadix];
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    const int n = d.counts->n_l + d.counts->n_h;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = t; i < kWaves * kRadix; i += kSmallSortThreads) (&wcnt[0][0])[i] = 0;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = kRadixBits * pass;
        const uint32_t* __restrict__ kin = d.keys[pass & 1];
        const uint32_t* __restrict__ vin = d.vals[pass & 1];
        uint32_t* __restrict__ kout = d.keys[(pass & 1) ^ 1];
        uint32_t* __restrict__ vout = d.vals[(pass & 1) ^ 1];
        for (int i = t; i < kRadix; i += kSmallSortThreads) {
            hcount[i] = 0;
            running[i] = 0;
        }
        __syncthreads();
        for (int e = t; e < n; e += kSmallSortThreads)
            atomicAdd(&hcount[(kin[e] >> shift) & (kRadix - 1u)], 1u);
        __syncthreads();
        // digit bases: exclusive scan of the histogram by the first wave,
        // kRadix / 64 digits per lane
        if (w == 0) {
            constexpr int kPer = kRadix / 64;
            uint32_t loc[kPer], sum = 0;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                loc[q] = hcount[lane * kPer + q];
                sum += loc[q];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t a = (uint32_t)__shfl_up((int)incl, off, 64);
                incl += lane >= off ? a : 0u;
            }
            uint32_t run = incl - sum;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                base[lane * kPer + q] = run;
                run += loc[q];
            }
        }
        __syncthreads();
        for (int c0 = 0; c0 < n; c0 += kSmallSortThreads) {
            const int e = c0 + t;
            const bool valid = e < n;
            const uint32_t key = valid ? kin[e] : 0u;
            const uint32_t val = valid ? vin[e] : 0u;
            const uint32_t dg = (key >> shift) & (kRadix - 1u);
            uint64_t m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < kRadixBits; ++b) {
                const uint64_t bb = __ballot((dg >> b) & 1u);
                m &= ((dg >> b) & 1u) ? bb : ~bb;
            }
            const uint32_t rk = (uint32_t)__popcll(m & lt);
            const bool leader = valid && rk == 0;
            if (leader) wcnt[w][dg] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t off = 0, tot = 0;
            bool last = true;
            if (leader) {
                for (int v = 0; v < kWaves; ++v) {
                    const uint32_t c = wcnt[v][dg];
                    off += v < w ? c : 0u;
                    tot += c;
                    last = last && (v <= w || c == 0u);
                }
                off += running[dg];
            }
            __syncthreads();
            if (leader) {
                if (last) running[dg] += tot;  // the digit's last wave this chunk
                wcnt[w][dg] = 0;
            }
            const int src = valid ? __ffsll((long long)m) - 1 : lane;
            off = (uint32_t)__shfl((int)off, src, 64);
            if (valid) {
                const uint32_t dst = base[dg] + off + rk;
                kout[dst] = key;
                vout[dst] = val;
            }
        }
        // the pass's writes are read by the next pass (same workgroup)
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Cell ranges (R:1300-1340) + reorder into cell-sorted SoA + scene segments.
// ---------------------------------------------------------------------------
__global__ void k_reorder(DevState d, const uint32_t* __restrict__ keys,
                          const uint32_t* __restrict__ vals, int with_sidx) {
    const int nl = d.counts->n_l;
    const int n = nl + d.counts->n_h;
    GRID_STRIDE(k, n) {
        const uint32_t key = keys[k];
        const int u = (int)vals[k];
        if (k == 0 || keys[k - 1] != key) d.cstart[key] = k;
        if (k == n - 1 || keys[k + 1] != key) d.cend[key] = k + 1;
        // states present per cell (every particle stores the same 1 for its
        // state): lets the scans skip rows whose weights are all zero
        const uint8_t st = d.state[u];
        d.cflag[2 * (size_t)key + (st & 1u)] = 1;
        const float2 x = d.pos[u];
        const float2 v = d.vel[u];
        d.spos[k] = x;
        d.sxv[k] = make_float4(x.x, x.y, v.x, v.y);
        d.sstate[k] = st;
        if (with_sidx) d.sidx[u] = k;
        // the scene owns the cell (cells of a scene are one range)
        const int s = d.cscene[key];
        d.sscene[k] = (uint16_t)s;
        // scene segments per level: low part [0, nl), high part [nl, n)
        const bool high = k >= nl;
        // state-0 particles present: promoted children make the low
        // level's cross passes run, active low particles the high level's
        if (!(st & 1u)) {
            if (high) d.scnt[s].prom = 1;
            else d.scnt[s].act = 1;
        }
        const int first = high ? nl : 0;
        const int last = high ? n - 1 : nl - 1;
        const bool starts = k == first || d.cscene[keys[k - 1]] != s;
        const bool ends = k == last || d.cscene[keys[k + 1]] != s;
        if (starts) {
            if (high) d.scnt[s].seg_h0 = k; else d.scnt[s].seg_l0 = k;
        }
        if (ends) {
            if (high) d.scnt[s].seg_h1 = k + 1; else d.scnt[s].seg_l1 = k + 1;
        }
    }
}

#include "tss_pairs.inc"

// ---------------------------------------------------------------------------
// Relaxation of the children injected by the last step's splits (the last
// n_fresh entries of the high list): relax_iters Jacobi iterations, each
//   1. density + Tait of every target, with the high level's own sums
//      (lambda_HH over high r <= H, lambda_HL over low r <= H, blended M4);
//   2. pressure + wall acceleration of every target (no viscosity, tension
//      or gravity): -m_H sum_HH (p_i/rho_i^2 + p_j/rho_j^2) grad W_bar over
//      0 < r < H, -m_L sum_HL (...) over 0 < r < h, + walls; then
//      v' = v + dt_r a, x' = x + dt_r v' with dt_r = relax_frac * dt, into
//      scratch (every target reads the iteration's start state);
//   3. commit of x', v' to the by-id and cell-sorted arrays.
// Neighbours' rho and p are read by id: the targets' from step 1, everyone
// else's from the last step. Cells stay those of the grid just built.
// (TwoScale2D-master/TwoScale2D/WCSPHSolver.cu:2380-2446, 1133-1230,
// 1499-1625, dormant at R:3044-3110; here with the method's weights and
// without the in-place race.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool relax_target(const DevState& d,
                                             const int* __restrict__ lh, int t,
                                             int it, int& u, int& k) {
    const int nh = d.counts->n_h;
    const int nf = min(d.counts->n_fresh, nh);
    u = lh[nh - nf + t];
    k = d.sidx[u];
    return it < d.scenes[d.sscene[k]].relax_iters;
}

__global__ void k_relax_density(DevState d, const int* __restrict__ lh,
                                int it) {
    GRID_STRIDE(t, min(d.counts->n_fresh, d.counts->n_h)) {
        int u, k;
        if (!relax_target(d, lh, t, it, u, k)) continue;
        const SceneDev& sc = d.scenes[d.sscene[k]];
        const float2 xi = d.spos[k];
        const uint8_t si = d.sstate[k];
        const bool i_prom = !(si & 1u);
        float s1 = 0.f, c1 = 0.f, s2 = 0.f, c2 = 0.f;
        Box b = scan_box(sc, true, xi, sc.H);
        for (int cy = b.y0; cy <= b.y1; ++cy) {
            int st, en;
            row_run(d.cstart, d.cend, sc.cbase_h + cy * sc.dhx, b.x0, b.x1, st, en);
            for (int j = st; j < en; ++j) {
                const float r2 = dist2(xi, d.spos[j]);
                if (r2 <= sc.t_le_H && !(i_prom && (d.sstate[j] & 1u))) {
                    float w, g;
                    m4_blend(sc, dist_from2(r2), w, g);
                    kahan_add(s1, c1, w);
                }
            }
        }
        b = scan_box(sc, false, xi, sc.h);
        for (int cy = b.y0; cy <= b.y1; ++cy) {
            int st, en;
            row_run(d.cstart, d.cend, sc.cbase_l + cy * sc.dlx, b.x0, b.x1, st, en);
            for (int j = st; j < en; ++j) {
                const float r2 = dist2(xi, d.spos[j]);
                if (r2 <= sc.t_le_H && !((si & 1u) && (d.sstate[j] & 1u))) {
                    float w, g;
                    m4_blend(sc, dist_from2(r2), w, g);
                    kahan_add(s2, c2, w);
                }
            }
        }
        float hi, lo;
        mass_sum(sc.m_h, s1, c1, sc.m_l, s2, c2, hi, lo);
        d.rho[u] = hi + lo;
        d.p[u] = tait(sc, hi, lo);
    }
}

__global__ void k_relax_force(DevState d, const uint32_t* __restrict__ vals,
                              const int* __restrict__ lh, int it) {
    GRID_STRIDE(t, min(d.counts->n_fresh, d.counts->n_h)) {
        int u, k;
        if (!relax_target(d, lh, t, it, u, k)) continue;
        const SceneDev& sc = d.scenes[d.sscene[k]];
        const float4 xv = d.sxv[k];
        const float2 xi = make_float2(xv.x, xv.y);
        const uint8_t si = d.sstate[k];
        const bool i_prom = !(si & 1u);
        const float rhoi = d.rho[u];
        const float pri = d.p[u] / (rhoi * rhoi);
        float sx = 0.f, cx = 0.f, sy = 0.f, cy_ = 0.f;
        auto pair = [&](int j, float mneg) {
            const float2 xj = d.spos[j];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y;
            const float r2 = fmaf(dx, dx, dy * dy);
            const float rinv = fast_rsq(r2);
            float w, g;
            m4_blend(sc, r2 * rinv, w, g);
            const int uj = (int)vals[j];
            const float rhoj = d.rho[uj];
            const float coeff = pri + d.p[uj] / (rhoj * rhoj);
            const float f = mneg * coeff * g * rinv;
            kahan_add(sx, cx, f * dx);
            kahan_add(sy, cy_, f * dy);
        };
        Box b = scan_box(sc, true, xi, sc.H);
        for (int ry = b.y0; ry <= b.y1; ++ry) {
            int st, en;
            row_run(d.cstart, d.cend, sc.cbase_h + ry * sc.dhx, b.x0, b.x1, st, en);
            for (int j = st; j < en; ++j) {
                const float r2 = dist2(xi, d.spos[j]);
                if (in_open_ball(r2, sc.t_lt_H) && !(i_prom && (d.sstate[j] & 1u)))
                    pair(j, -sc.m_h);
            }
        }
        b = scan_box(sc, false, xi, sc.h);
        float bx = 0.f, by = 0.f;
        for (int ry = b.y0; ry <= b.y1; ++ry) {
            int st, en;
            row_run(d.cstart, d.cend, sc.cbase_l + ry * sc.dlx, b.x0, b.x1, st, en);
            for (int j = st; j < en; ++j) {
                const float r2 = dist2(xi, d.spos[j]);
                const uint8_t sj = d.sstate[j];
                if (in_open_ball(r2, sc.t_lt_h) && !((si & 1u) && (sj & 1u)) && sj != kDeleted)
                    pair(j, -sc.m_l);
            }
            int bs, be;
            row_run(d.bstart, d.bend, sc.cbase_l + ry * sc.dlx, b.x0, b.x1, bs, be);
            for (int j = bs; j < be; ++j) {
                const float2 xb = d.bpos[j];
                const float r2 = dist2(xi, xb);
                const float gm = sc.cb_h * wall_gamma(sc, r2, dist_from2(r2), sc.q_H);
                bx = fmaf(gm, xi.x - xb.x, bx);
                by = fmaf(gm, xi.y - xb.y, by);
            }
        }
        const float ax = (sx - cx) + bx;
        const float ay = (sy - cy_) + by;
        const float dt = sc.relax_frac * sc.dt;
        const float vx = xv.z + dt * ax;
        const float vy = xv.w + dt * ay;
        d.rtmp[t] = make_float4(xi.x + dt * vx, xi.y + dt * vy, vx, vy);
    }
}

__global__ void k_relax_commit(DevState d, const int* __restrict__ lh, int it) {
    GRID_STRIDE(t, min(d.counts->n_fresh, d.counts->n_h)) {
        int u, k;
        if (!relax_target(d, lh, t, it, u, k)) continue;
        const float4 v = d.rtmp[t];
        d.pos[u] = make_float2(v.x, v.y);
        d.vel[u] = make_float2(v.z, v.w);
        d.spos[k] = make_float2(v.x, v.y);
        d.sxv[k] = v;
    }
}


// ---------------------------------------------------------------------------
// FP64 reductions per scene: conservation sums over real particles (listed
// low, promoted high) and the max |v|^2 for the CFL time step. Fixed chunk
// and tree shape -> bitwise reproducible, independent of other scenes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce(DevState d) {
    __shared__ double sh[kRedFields][256];
    const int slot = blockIdx.x;
    const int s = d.red_scene[slot];
    if (s < 0) return;  // whole block: slot not owned by any scene
    const int chunk = slot - d.red_base[s];
    const SceneDev& sc = d.scenes[s];
    const SceneCount& c = d.scnt[s];
    const int nl_s = c.seg_l1 - c.seg_l0;
    const int m = nl_s + (c.seg_h1 - c.seg_h0);
    double acc[kRedFields] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (sc.alive && m > 0) {
        const long long i0 = (long long)chunk * kRedChunk;
        const long long i1 = i0 + kRedChunk < m ? i0 + kRedChunk : (long long)m;
        const double g = (double)sc.g;
        // four entries per thread and trip, their loads in flight together
        // (one load per trip left the loop waiting on each in turn); added
        // in the same order as one entry per trip
        const long long bd = blockDim.x;
        for (long long i = i0 + threadIdx.x; i < i1; i += 4 * bd) {
            float4 xv[4];
            uint8_t st[4];
            bool hi[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long long ii = i + q * bd < i1 ? i + q * bd : i;
                hi[q] = ii >= nl_s;
                const int k = !hi[q] ? c.seg_l0 + (int)ii : c.seg_h0 + (int)(ii - nl_s);
                // post-step position and velocity (the force pass's sorted copy)
                xv[q] = d.snew[k];
                st[q] = d.sstate[k];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (i + q * bd >= i1) break;
                const bool high = hi[q];
                double vx = xv[q].z, vy = xv[q].w;
                double v2 = vx * vx + vy * vy;
                acc[5] = fmax(acc[5], v2);
                if (!high || st[q] == 0) {
                    double mm = high ? (double)sc.m_h : (double)sc.m_l;
                    double y = (double)xv[q].y;
                    double ke = 0.5 * (mm * v2);
                    double pe = (mm * g) * y;
                    acc[0] += mm;
                    acc[1] += mm * vx;
                    acc[2] += mm * vy;
                    acc[3] += ke;
                    acc[4] += pe;
                    acc[6] += fabs(mm * vx);
                    acc[7] += fabs(mm * vy);
                    acc[8] += fabs(pe);
                }
            }
        }
    }
    for (int f = 0; f < kRedFields; ++f) sh[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            for (int f = 0; f < kRedFields; ++f) {
                if (f == 5)
                    sh[f][threadIdx.x] = fmax(sh[f][threadIdx.x],
                                              sh[f][threadIdx.x + off]);
                else
                    sh[f][threadIdx.x] += sh[f][threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* out = d.red + (size_t)slot * kRedFields;
        for (int f = 0; f < kRedFields; ++f) out[f] = sh[f][0];
    }
}

// ---------------------------------------------------------------------------
// Adaptivity.
// ---------------------------------------------------------------------------
// Cell flags (0 or 1 bytes) of cells c0 .. min(c0 + 7, x1) of a row as a
// mask, bit k for cell c0 + k: three aligned word loads cover any eight
// bytes (the tables are padded past their last cell), and bits past x1 are
// cleared.
__device__ __forceinline__ uint32_t flag_bits8(const uint8_t* __restrict__ f,
                                               int row, int c0, int x1) {
    const int a = row + c0;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(f) + (a >> 2);
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
    const int sh = (a & 3) * 8;
    uint64_t bytes = (((uint64_t)w1 << 32) | w0) >> sh;
    if (sh) bytes |= (uint64_t)w2 << (64 - sh);
    // byte k (0 or 1) -> bit k: the product gathers the eight low bits in
    // its top byte
    const uint32_t m = (uint32_t)(((bytes & 0x0101010101010101ull) *
                                   0x0102040810204080ull) >> 56);
    const int n = min(8, x1 - c0 + 1);
    return m & ((1u << n) - 1u);
}

// Any surface particle within sqrt-distance r of x (r2 <= t)? Scans the
// low box and the high box of radius r on the cell-sorted pre-step state.
// Only cells flagged as holding a surface particle (csurf, set by the
// density pass) are read, and the high box only where the scene has a high
// surface particle this step (hsurf).
__device__ bool in_band(const DevState& d, const SceneDev& sc, float2 x,
                        float r, float t, bool hsurf) {
    // coarse test first: a surface particle within r of x lies in a low cell
    // at most one cell outside x's low box of radius r (rounding included)
    // and marks its own low cell and the eight around it (cband), so a box
    // without a mark holds no such particle. Flags are read eight cells of
    // a row at a time, the loads issued together (indices clamped into the
    // box: a repeated cell changes no OR).
    {
        const Box b = scan_box(sc, false, x, r);
        bool any = false;
        for (int cy = b.y0; cy <= b.y1 && !any; ++cy) {
            const int row = sc.cbase_l + cy * sc.dlx;
            for (int c0 = b.x0; c0 <= b.x1 && !any; c0 += 8)
                any = flag_bits8(d.cband, row, c0, b.x1) != 0;
        }
        if (!any) return false;
    }
    // the high cells only where the scene has a high surface particle
    for (int lv = 0; lv < (hsurf ? 2 : 1); ++lv) {
        const bool high = lv == 1;
        const Box b = scan_box(sc, high, x, r);
        const int base = high ? sc.cbase_h : sc.cbase_l;
        const int dimx = high ? sc.dhx : sc.dlx;
        for (int cy = b.y0; cy <= b.y1; ++cy) {
            const int row = base + cy * dimx;
            for (int c0 = b.x0; c0 <= b.x1; c0 += 8) {
                // the row's surface-cell flags, eight at a time
                uint32_t m = flag_bits8(d.csurf, row, c0, b.x1);
                while (m) {
                    const int c = row + c0 + __builtin_ctz(m);
                    m &= m - 1u;
                    // the cell's particles four at a time, their loads in
                    // flight together (indices clamped into the cell: a
                    // repeated particle changes no OR)
                    const int en = d.cend[c];
                    for (int j0 = d.cstart[c]; j0 < en; j0 += 4) {
                        bool hit = false;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int j = min(j0 + k, en - 1);
                            hit |= (d.ssurf[j] != 0) & (dist2(x, d.spos[j]) <= t);
                        }
                        if (hit) return true;
                    }
                }
            }
        }
    }
    return false;
}

// Split candidates over the sorted low particles (R:1834-1836 rule or the
// surface band rule).
__global__ void k_split_cand(DevState d) {
    const int nl = d.counts->n_l;
    GRID_STRIDE(k, nl) {
        const int s = d.sscene[k];
        const SceneDev& sc = d.scenes[s];
        uint8_t cand = 0;
        if (sc.mode != 0 && d.sstate[k] == kActive) {
            if (sc.mode == 1) {
                // predicted x' (R:1835): the force pass's post-step
                // position, read in sorted order (the same value it
                // stored by id)
                cand = d.snew[k].x > sc.split_x;
            } else {
                cand = in_band(d, sc, d.spos[k], sc.r_band, sc.t_band, d.scnt[s].hsurf != 0);
            }
        }
        d.cand[k] = cand;
        // per-scene candidate count; one atomic per wave when the wave's
        // active lanes share a scene (the usual case)
        const int s0 = __builtin_amdgcn_readfirstlane(s);
        if (__ballot(s != s0) == 0ull) {
            const uint64_t cm = __ballot(cand != 0);
            const uint64_t act = __ballot(true);
            const int lane = threadIdx.x & 63;
            if (cm && lane == __ffsll((long long)act) - 1) {
                atomicAdd(&d.scnt[s0].cand, (int)__popcll(cm));
                d.counts->events = 1;
            }
        } else if (cand) {
            atomicAdd(&d.scnt[s].cand, 1);
            d.counts->events = 1;
        }
    }
}

// R's adjustPosition (R:1007-1022): the bound rmax = 0.75 sqrt(m_l /
// (rho pi)) is a double (double pi), the same for a group's four children
// (clamp_radius); min(float dist, double rmax) is stored to a float r, and
// r / dist and the offsets are FP32, each operation rounded (the build does
// not contract). A coincident child is left in place.
__device__ __forceinline__ double clamp_radius(float m_l, float rho) {
    return 0.75 * sqrt((double)m_l / ((double)rho * 3.14159265358979323846));
}

__device__ __forceinline__ float2 clamp_child(float2 ph, float2 pl, double rmax) {
    float dx = ph.x - pl.x;
    float dy = ph.y - pl.y;
    float a = dx * dx;
    float b = dy * dy;
    float dist = sqrtf(a + b);
    if (!(dist > 0.0f)) return ph;
    const float r = (double)dist < rmax ? dist : (float)rmax;
    const float t = r / dist;
    const float ox = t * dx;
    const float oy = t * dy;
    return make_float2(pl.x + ox, pl.y + oy);
}

// Existing groups: age, promotion (dormant test of R:1053-1095, enabled by
// the method), child clamp (R:1097-1172), merge decision and application.
// LANES lanes per group: under the band rule the merge test scans the
// neighbourhoods of up to four children, one child per lane; the lanes of a
// group sit in one wave and OR their results in a ballot, and lane 0
// applies the decision. LANES = 1 tests the children in turn.
template <int LANES>
__global__ void k_groups(DevState d, const int* __restrict__ list_g, int solo) {
    static_assert(LANES == 1 || LANES == 4, "one lane or one per child");
    const int ng = d.counts->n_g;
    GRID_STRIDE(i, LANES * ng) {
        const int q = i / LANES, kl = i % LANES;
        const int p = list_g[q];
        // with one resident scene (solo >= 0), no lookup by id
        const int s = solo >= 0 ? solo : d.scene_of[p];
        const SceneDev& sc = d.scenes[s];
        const int c0 = d.CL + 4 * p;
        const uint8_t st = d.state[p];
        int age = 0;
        bool promote = false;
        if (st == kTransient) {
            age = d.age[p] + 1;
            if (sc.mode == 2 && sc.pmax > 0) {
                float fluc = 0.0f;
                for (int k = 0; k < 4; ++k) fluc += fabsf(sc.rho0 - d.rho[c0 + k]);
                fluc = fluc / (4.0f * sc.rho0);
                promote = (age >= sc.pmin && fluc <= sc.pfluc) || age >= sc.pmax;
            }
        }
        // merge test (band rule): a child within the merge radius of a
        // surface particle keeps the group
        const bool test = sc.mode == 2 && ((st == kTransient && !promote) || st == kDeleted);
        bool keep = false;
        if (LANES == 1) {
            for (int k = 0; k < 4 && test && !keep; ++k)
                keep = in_band(d, sc, d.spos[d.sidx[c0 + k]], sc.r_merge, sc.t_merge,
                               d.scnt[s].hsurf != 0);
        } else {
            const bool mine = test && in_band(d, sc, d.spos[d.sidx[c0 + kl]],
                                              sc.r_merge, sc.t_merge,
                                              d.scnt[s].hsurf != 0);
            const uint64_t bal = __ballot(mine);
            const int g0 = (threadIdx.x & 63) & ~(LANES - 1);
            keep = ((bal >> g0) & ((1ull << LANES) - 1ull)) != 0ull;
            if (kl != 0) continue;
        }
        // merged: 1 = promoted group merged back (parent re-listed),
        //         2 = transient group dissolved (parent already listed)
        uint8_t merged = 0;
        if (st == kTransient) {
            d.age[p] = age;
            if (promote) {
                d.state[p] = kDeleted;
                for (int k = 0; k < 4; ++k) d.state[c0 + k] = 0;
                atomicAdd(&d.scnt[s].n_promote, 1);
                d.counts->events = 1;
            } else if (sc.mode == 2 && !keep) {
                // dissolve: children leave the lists, the parent stays
                merged = 2;
                d.state[p] = kActive;
                d.age[p] = 0;
                atomicAdd(&d.scnt[s].n_dissolve, 1);
                d.counts->events = 1;
            } else {
                const float2 pl = d.pos[p];
                const double rmax = clamp_radius(sc.m_l, d.rho[p]);
                float2 ph[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) ph[k] = d.pos[c0 + k];
                // stored only where the clamp changes a bit (within the
                // radius it returns the position itself, as a rule)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float2 c = clamp_child(ph[k], pl, rmax);
                    if (__float_as_uint(c.x) != __float_as_uint(ph[k].x) ||
                        __float_as_uint(c.y) != __float_as_uint(ph[k].y))
                        d.pos[c0 + k] = c;
                }
            }
        } else if (st == kDeleted && sc.mode == 2 && !keep) {
            merged = 1;
            float2 c[4], v[4];
            for (int k = 0; k < 4; ++k) {
                c[k] = d.pos[c0 + k];
                v[k] = d.vel[c0 + k];
            }
            d.pos[p] = make_float2(((c[0].x + c[1].x) + (c[2].x + c[3].x)) * 0.25f,
                                   ((c[0].y + c[1].y) + (c[2].y + c[3].y)) * 0.25f);
            d.vel[p] = make_float2(((v[0].x + v[1].x) + (v[2].x + v[3].x)) * 0.25f,
                                   ((v[0].y + v[1].y) + (v[2].y + v[3].y)) * 0.25f);
            d.state[p] = kActive;
            d.age[p] = 0;
            for (int k = 0; k < 4; ++k) d.state[c0 + k] = 1;
            atomicAdd(&d.scnt[s].n_merge, 1);
            d.counts->events = 1;
        }
        d.merged[p] = merged;
    }
}

// High-particle cap (spec): where a scene has more split candidates than
// its cap allows, keep the first ones in sorted order. One block per scene;
// exits at once when the cap does not bind.
__global__ __launch_bounds__(256) void k_cap(DevState d) {
    __shared__ int part[256];
    const int s = blockIdx.x;
    const SceneDev& sc = d.scenes[s];
    const SceneCount& c = d.scnt[s];
    if (!sc.alive) return;
    const long long allowed_ll = ((long long)sc.cap_h - 4LL * c.n_g) / 4;
    const int allowed = allowed_ll < 0 ? 0 : (int)allowed_ll;
    if (c.cand <= allowed) return;
    int run = 0;
    for (int base = c.seg_l0; base < c.seg_l1; base += 256) {
        int k = base + threadIdx.x;
        int f = (k < c.seg_l1) ? d.cand[k] : 0;
        part[threadIdx.x] = f;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += v;
            __syncthreads();
        }
        int excl = run + part[threadIdx.x] - f;
        if (f && excl >= allowed) d.cand[k] = 0;
        int tot = part[255];
        __syncthreads();
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Compaction of the next lists (stable; replaces the atomicAdd appends of
// R:1826, 1845, 2030, 1135-1139). Entries E = [low tiles][high tiles]
// [group tiles]; two counters per entry:
//   P0 = keep (low: not deleted; high: group not merged; group: not merged)
//   P1 = low: split admitted; group: merged.
// ---------------------------------------------------------------------------
struct Flags {
    int p0, p1;
};

__device__ __forceinline__ Flags entry_flags(const DevState& d,
                                             const uint32_t* __restrict__ vals,
                                             const int* __restrict__ list_g,
                                             int seg, int i, int nl, int nh,
                                             int ng) {
    Flags f = {0, 0};
    if (seg == 0) {
        if (i < nl) {
            int u = (int)vals[i];
            f.p0 = d.state[u] != kDeleted;
            f.p1 = d.cand[i];
        }
    } else if (seg == 1) {
        if (i < nh) {
            int u = (int)vals[nl + i];
            f.p0 = !d.merged[(u - d.CL) >> 2];
        }
    } else {
        if (i < ng) {
            int p = list_g[i];
            int m = d.merged[p];
            f.p0 = m == 0;
            f.p1 = m == 1;  // only merged-back parents are re-listed
        }
    }
    return f;
}

__device__ __forceinline__ void tile_seg(const StepShape& sh, int tile,
                                         int& seg, int& t) {
    if (tile < sh.scan_tiles_l) {
        seg = 0;
        t = tile;
    } else if (tile < sh.scan_tiles_l + sh.scan_tiles_h) {
        seg = 1;
        t = tile - sh.scan_tiles_l;
    } else {
        seg = 2;
        t = tile - sh.scan_tiles_l - sh.scan_tiles_h;
    }
}

// A compaction tile is kScanItems chunks of kScanThreads consecutive
// entries; chunk q of tile t holds entries t * kScanTile + q * kScanThreads
// + thread (coalesced), and entries are ranked in index order with wave
// ballots.
constexpr int kScanWaves = kScanThreads / 64;

__global__ __launch_bounds__(kScanThreads) void k_compact_reduce(
    DevState d, StepShape sh, const uint32_t* __restrict__ vals,
    const int* __restrict__ list_g) {
    __shared__ int wsum[2][kScanWaves];
    const int nl = d.counts->n_l, nh = d.counts->n_h, ng = d.counts->n_g;
    int seg, t;
    tile_seg(sh, blockIdx.x, seg, t);
    // a tile past its segment's live entries counts nothing (the grid is
    // sized by capacity): one write, and the whole block exits. A step
    // without events keeps every live entry and splits or re-lists none:
    // the counts follow from the segment sizes alone
    const int nseg = seg == 0 ? nl : seg == 1 ? nh : ng;
    if (t * kScanTile >= nseg || !d.counts->events) {
        if (threadIdx.x == 0) {
            d.tile_sums[2 * blockIdx.x] = max(0, min(kScanTile, nseg - t * kScanTile));
            d.tile_sums[2 * blockIdx.x + 1] = 0;
        }
        return;
    }
    int a = 0, b = 0;  // wave totals (wave-uniform)
    const int i0 = t * kScanTile + (int)threadIdx.x;
    for (int q = 0; q < kScanItems; ++q) {
        const Flags f = entry_flags(d, vals, list_g, seg, i0 + q * kScanThreads, nl, nh, ng);
        a += __popcll(__ballot(f.p0));
        b += __popcll(__ballot(f.p1));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wsum[0][w] = a;
        wsum[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0;
        for (int v = 0; v < kScanWaves; ++v) {
            ta += wsum[0][v];
            tb += wsum[1][v];
        }
        d.tile_sums[2 * blockIdx.x] = ta;
        d.tile_sums[2 * blockIdx.x + 1] = tb;
    }
}

// Single block: exclusive scan of the tile sums, segment bases, new counts.
// Only tiles holding live entries are scanned (the others count zero and
// the scatter never reads their bases): virtual tile v runs over the live
// tiles of the low, high and group segments in that order.
__global__ __launch_bounds__(1024) void k_compact_top(DevState d, StepShape sh) {
    __shared__ int2 part[1024];
    __shared__ int2 seg_pre[2];  // prefixes where the high / group tiles start
    const int th = sh.scan_tiles_l, tg = sh.scan_tiles_l + sh.scan_tiles_h;
    const DevCounts* c = d.counts;
    const int ll = min((c->n_l + kScanTile - 1) / kScanTile, sh.scan_tiles_l);
    const int lh = min((c->n_h + kScanTile - 1) / kScanTile, sh.scan_tiles_h);
    const int lg = min((c->n_g + kScanTile - 1) / kScanTile, sh.scan_tiles_g);
    const int nv = ll + lh + lg;
    auto real = [&](int v) {
        return v < ll ? v : (v < ll + lh ? th + (v - ll) : tg + (v - ll - lh));
    };
    const int per = (nv + 1023) / 1024;
    const int t0 = threadIdx.x * per;
    int a = 0, b = 0;
    for (int i = 0; i < per; ++i) {
        const int v = t0 + i;
        if (v < nv) {
            const int t = real(v);
            a += d.tile_sums[2 * t];
            b += d.tile_sums[2 * t + 1];
        }
    }
    part[threadIdx.x] = make_int2(a, b);
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int2 w = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off]
                                              : make_int2(0, 0);
        __syncthreads();
        part[threadIdx.x].x += w.x;
        part[threadIdx.x].y += w.y;
        __syncthreads();
    }
    const int tot0 = part[1023].x, tot1 = part[1023].y;
    if (threadIdx.x == 0) {
        // a segment without live tiles starts at the total
        seg_pre[0] = make_int2(tot0, tot1);
        seg_pre[1] = make_int2(tot0, tot1);
    }
    __syncthreads();
    int ra = part[threadIdx.x].x - a, rb = part[threadIdx.x].y - b;
    for (int i = 0; i < per; ++i) {
        const int v = t0 + i;
        if (v < nv) {
            const int t = real(v);
            const int va = d.tile_sums[2 * t], vb = d.tile_sums[2 * t + 1];
            if (v == ll) seg_pre[0] = make_int2(ra, rb);
            if (v == ll + lh) seg_pre[1] = make_int2(ra, rb);
            d.tile_base[2 * t] = ra;
            d.tile_base[2 * t + 1] = rb;
            ra += va;
            rb += vb;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int b0h = seg_pre[0].x, b1h = seg_pre[0].y;
        const int b0g = seg_pre[1].x, b1g = seg_pre[1].y;
        const int keepL = b0h, nsplit = b1h;
        const int keepH = b0g - b0h;
        const int keepG = tot0 - b0g, nmerge = tot1 - b1g;
        d.seg_base[0] = b0h;    // P0 base of high segment
        d.seg_base[1] = b0g;    // P0 base of group segment
        d.seg_base[2] = b1g;    // P1 base of group segment
        d.seg_base[3] = keepL;
        d.seg_base[4] = keepH;
        d.seg_base[5] = keepG;
        d.seg_base[6] = nsplit;
        d.seg_base[7] = nmerge;
        d.counts->n_l_next = keepL + nmerge;
        d.counts->n_h_next = keepH + 4 * nsplit;
        d.counts->n_fresh_next = 4 * nsplit;
        d.counts->n_g_next = keepG + nsplit;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_compact_scatter(
    DevState d, StepShape sh, const uint32_t* __restrict__ vals,
    const int* __restrict__ list_g, int* __restrict__ nlist_l,
    int* __restrict__ nlist_h, int* __restrict__ nlist_g) {
    __shared__ int wsum[2][kScanWaves];
    const int nl = d.counts->n_l, nh = d.counts->n_h, ng = d.counts->n_g;
    const bool events = d.counts->events != 0;
    int seg, t;
    tile_seg(sh, blockIdx.x, seg, t);
    if (!events) {
        // every live entry is kept, in order, and none splits or is
        // re-listed: the next lists are copies of the sorted ids and the
        // groups, and each scene's list bounds are its sorted segments
        if (blockIdx.x == 0) {
            for (int s = threadIdx.x; s < d.S; s += kScanThreads) {
                SceneCount& c = d.scnt[s];
                c.kl0 = c.seg_l0;
                c.kl1 = c.seg_l1;
                c.kh0 = c.seg_h0 - nl;
                c.kh1 = c.seg_h1 - nl;
                c.sp0 = 0;
                c.sp1 = 0;
            }
        }
        const int nseg = seg == 0 ? nl : seg == 1 ? nh : ng;
        const uint32_t* src = seg == 2 ? (const uint32_t*)list_g : vals + (seg == 1 ? nl : 0);
        int* dst = seg == 0 ? nlist_l : seg == 1 ? nlist_h : nlist_g;
        // all of the thread's entries in flight, then stored
        uint32_t v[kScanItems];
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            const int i = t * kScanTile + q * kScanThreads + (int)threadIdx.x;
            v[q] = i < nseg ? src[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            const int i = t * kScanTile + q * kScanThreads + (int)threadIdx.x;
            if (i < nseg) dst[i] = (int)v[q];
        }
        return;
    }
    // tiles past their segment's live entries write nothing: whole block
    if (t * kScanTile >= (seg == 0 ? nl : seg == 1 ? nh : ng)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // output positions of the tile's first entry
    int run0 = d.tile_base[2 * blockIdx.x];
    int run1 = d.tile_base[2 * blockIdx.x + 1];
    const int keepL = d.seg_base[3], keepH = d.seg_base[4],
              keepG = d.seg_base[5];
    // places entry i at output positions g0 / g1 (flags fq)
    auto emit = [&](int i, int g0, int g1, const Flags& fq) {
        if (seg == 0 && i < nl) {
            const int u = (int)vals[i];
            const int s = d.sscene[i];
            SceneCount& sct = d.scnt[s];
            // the scene's first / last sorted low entry
            if (i == 0 || d.sscene[i - 1] != s) {
                sct.kl0 = g0;
                sct.sp0 = g1;
            }
            if (i == nl - 1 || d.sscene[i + 1] != s) {
                sct.kl1 = g0 + fq.p0;
                sct.sp1 = g1 + fq.p1;
            }
            if (fq.p0) nlist_l[g0] = u;
            if (fq.p1) {
                // split (R:956-1005, 1834-1863): children at x' +-(rc, rc)
                const SceneDev& sc = d.scenes[s];
                nlist_g[keepG + g1] = u;
                const float2 x = d.pos[u];
                const float2 v = d.vel[u];
                const float rho = d.rho[u];
                const float two_rho = 2.0f * rho;
                const float rc = (float)(0.5 * sqrt((double)sc.m_l /
                                                    ((double)two_rho * 3.14159265358979323846)));
                const int c0 = d.CL + 4 * u;
                for (int k = 0; k < 4; ++k) {
                    const int a2 = c0 + k;
                    d.pos[a2] = make_float2((k & 1) ? x.x - rc : x.x + rc,
                                            (k & 2) ? x.y - rc : x.y + rc);
                    d.vel[a2] = v;
                    d.acc[a2] = make_float2(0.f, 0.f);
                    d.rho[a2] = 0.f;
                    d.p[a2] = 0.f;
                    d.state[a2] = 1;
                    nlist_h[keepH + 4 * g1 + k] = a2;
                }
                d.state[u] = kTransient;
                d.age[u] = 0;
            }
        } else if (seg == 1 && i < nh) {
            const int k = nl + i;
            const int s = d.sscene[k];
            SceneCount& sct = d.scnt[s];
            const int gh = g0 - d.seg_base[0];
            // the scene's first / last sorted high entry
            if (k == nl || d.sscene[k - 1] != s) sct.kh0 = gh;
            if (k == nl + nh - 1 || d.sscene[k + 1] != s) sct.kh1 = gh + fq.p0;
            if (fq.p0) nlist_h[gh] = (int)vals[k];
        } else if (seg == 2 && i < ng) {
            const int p = list_g[i];
            if (fq.p0) nlist_g[g0 - d.seg_base[1]] = p;
            if (fq.p1) nlist_l[keepL + (g1 - d.seg_base[2])] = p;
        }
    };
    for (int q = 0; q < kScanItems; ++q) {
        const int i = t * kScanTile + q * kScanThreads + (int)threadIdx.x;
        const Flags fq = entry_flags(d, vals, list_g, seg, i, nl, nh, ng);
        const uint64_t m0 = __ballot(fq.p0), m1 = __ballot(fq.p1);
        if (lane == 0) {
            wsum[0][wv] = __popcll(m0);
            wsum[1][wv] = __popcll(m1);
        }
        __syncthreads();
        int pre0 = 0, pre1 = 0, tot0 = 0, tot1 = 0;
        for (int v = 0; v < kScanWaves; ++v) {
            const int c0 = wsum[0][v], c1 = wsum[1][v];
            pre0 += v < wv ? c0 : 0;
            pre1 += v < wv ? c1 : 0;
            tot0 += c0;
            tot1 += c1;
        }
        __syncthreads();  // wsum is rewritten by the next chunk
        // exclusive output positions of this entry
        const int g0 = run0 + pre0 + __popcll(m0 & lt);
        const int g1 = run1 + pre1 + __popcll(m1 & lt);
        run0 += tot0;
        run1 += tot1;
        emit(i, g0, g1, fq);
    }
}

// ---------------------------------------------------------------------------
// End of step: per-scene dt, conservation sums and counters; global counts.
// ---------------------------------------------------------------------------
// One block per scene: the scene's chunk partials are summed in a fixed
// order (thread t takes slots t, t+256, ...; then a fixed tree), so the
// sums are bitwise reproducible.
__global__ __launch_bounds__(256) void k_finalize(DevState d) {
    __shared__ double sh[kRedFields][256];
    const int s = blockIdx.x;
    SceneDev& sc = d.scenes[s];
    SceneCount& c = d.scnt[s];
    const int nslots = d.red_n[s];
    if (sc.alive) {
        double part[kRedFields] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        // two slots per trip, their loads in flight together; added in
        // slot order as with one slot per trip
        for (int ch = threadIdx.x; ch < nslots; ch += 512) {
            const bool two = ch + 256 < nslots;
            const double* r0 = d.red + ((size_t)d.red_base[s] + ch) * kRedFields;
            const double* r1 = two ? r0 + 256 * kRedFields : r0;
            double v0[kRedFields], v1[kRedFields];
#pragma unroll
            for (int f = 0; f < kRedFields; ++f) {
                v0[f] = r0[f];
                v1[f] = r1[f];
            }
#pragma unroll
            for (int f = 0; f < kRedFields; ++f) {
                if (f == 5) part[f] = fmax(part[f], v0[f]);
                else part[f] += v0[f];
            }
            if (two) {
#pragma unroll
                for (int f = 0; f < kRedFields; ++f) {
                    if (f == 5) part[f] = fmax(part[f], v1[f]);
                    else part[f] += v1[f];
                }
            }
        }
        for (int f = 0; f < kRedFields; ++f) sh[f][threadIdx.x] = part[f];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (sc.alive && threadIdx.x < (unsigned)off) {
            for (int f = 0; f < kRedFields; ++f) {
                if (f == 5)
                    sh[f][threadIdx.x] = fmax(sh[f][threadIdx.x], sh[f][threadIdx.x + off]);
                else
                    sh[f][threadIdx.x] += sh[f][threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (sc.alive && threadIdx.x == 0) {
        double tot[kRedFields];
        for (int f = 0; f < kRedFields; ++f) tot[f] = sh[f][0];
        c.n_l_last = c.seg_l1 - c.seg_l0;
        c.n_h_last = c.seg_h1 - c.seg_h0;
        c.psteps += (unsigned long long)(c.n_l_last + c.n_h_last);
        // next lists: low = kept + merged; high = kept + 4 per split
        c.n_split = c.sp1 - c.sp0;
        c.n_l = (c.kl1 - c.kl0) + c.n_merge;
        c.n_h = (c.kh1 - c.kh0) + 4 * c.n_split;
        c.n_g = c.n_g + c.n_split - c.n_merge - c.n_dissolve;
        c.step += 1;
        // rounding bound of the sums: per-term rounding (<= 2u) + the
        // serial per-thread loop + 8-level tree of k_reduce, then the
        // per-thread slot loop + 8-level tree here
        const double serial = (double)(kRedChunk / 256);
        const double kk = serial + 8.0 + (double)((nslots + 255) / 256) + 8.0 + 2.0;
        const double u = 1.1102230246251565e-16;
        const double gam = kk * u / (1.0 - kk * u);
        double* o = d.cons + (size_t)s * 10;
        for (int f = 0; f < 5; ++f) o[f] = tot[f];
        o[5] = gam * tot[0];  // mass (positive terms)
        o[6] = gam * tot[6];  // sum |m vx|
        o[7] = gam * tot[7];  // sum |m vy|
        o[8] = gam * tot[3];  // kinetic (positive terms)
        o[9] = gam * tot[8];  // sum |m g y|
        if (sc.dt_mode == 1) {
            double ell = 0.5 * (double)(c.n_h > 0 ? sc.H : sc.h);
            double dt = (double)sc.cfl * ell / ((double)sc.c + sqrt(tot[5]));
            if (dt > (double)sc.dt_max) dt = (double)sc.dt_max;
            sc.dt = (float)dt;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        DevCounts& g = *d.counts;
        g.n_l_last = g.n_l;
        g.n_h_last = g.n_h;
        g.n_l = g.n_l_next;
        g.n_h = g.n_h_next;
        g.n_g = g.n_g_next;
        g.n_fresh = g.n_fresh_next;
        g.events = g.events_floor;
    }
}

// ---------------------------------------------------------------------------
// Neighbour-set inspection (tests): same boxes, runs and predicates as the
// density and force kernels.
// ---------------------------------------------------------------------------

__global__ void k_neighbors(DevState d, const uint32_t* __restrict__ vals,
                            const int* __restrict__ list, int n, int scene,
                            uint32_t* __restrict__ counts,
                            uint32_t* __restrict__ sums) {
    const SceneDev& sc = d.scenes[scene];
    GRID_STRIDE(t, n) {
        const int u = list[t];
        const int k = d.sidx[u];
        const bool high = u >= d.CL;
        const float2 xi = d.spos[k];
        uint32_t c[5] = {0, 0, 0, 0, 0}, su[5] = {0, 0, 0, 0, 0};
        auto local_id = [&](int v) {
            // scene-local id of a unified id
            return v < d.CL ? v - sc.lbase : (v - d.CL) - 4 * sc.lbase;
        };
        // same level
        {
            const float rs = high ? sc.H : sc.h;
            Box b = scan_box(sc, high, xi, rs);
            const float tle = high ? sc.t_le_H : sc.t_le_h;
            const float tlt = high ? sc.t_lt_H : sc.t_lt_h;
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                int row = high ? sc.cbase_h + cy * sc.dhx : sc.cbase_l + cy * sc.dlx;
                row_run(d.cstart, d.cend, row, b.x0, b.x1, st, en);
                for (int j = st; j < en; ++j) {
                    float r2 = dist2(xi, d.spos[j]);
                    int id = local_id((int)vals[j]);
                    if (r2 <= tle) { c[0]++; su[0] += csum(id); }
                    if (r2 != 0.0f && r2 <= tlt) { c[2]++; su[2] += csum(id); }
                }
            }
        }
        // cross level (box +-h on the other grid)
        {
            const bool oh = !high;
            Box b = scan_box(sc, oh, xi, sc.h);
            const float tle = high ? sc.t_le_H : sc.t_le_h;
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                int row = oh ? sc.cbase_h + cy * sc.dhx : sc.cbase_l + cy * sc.dlx;
                row_run(d.cstart, d.cend, row, b.x0, b.x1, st, en);
                for (int j = st; j < en; ++j) {
                    float r2 = dist2(xi, d.spos[j]);
                    int id = local_id((int)vals[j]);
                    if (r2 <= tle) { c[1]++; su[1] += csum(id); }
                    if (r2 != 0.0f && r2 <= sc.t_lt_h) { c[3]++; su[3] += csum(id); }
                }
            }
        }
        // boundary candidates of the low box
        {
            Box b = scan_box(sc, false, xi, sc.h);
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                row_run(d.bstart, d.bend, sc.cbase_l + cy * sc.dlx, b.x0, b.x1,
                        st, en);
                for (int j = st; j < en; ++j) {
                    c[4]++;
                    su[4] += csum(d.bids[j]);
                }
            }
        }
        for (int q = 0; q < 5; ++q) {
            counts[5 * t + q] = c[q];
            sums[5 * t + q] = su[q];
        }
    }
}

// Exact work of the density and force kernels on the current grid (the
// algorithmic flop counts of the roofline): candidates scanned per box kind,
// pairs that pass each kernel's cutoff and weight, wall candidates, targets.
__global__ void k_work(DevState d, unsigned long long* __restrict__ out) {
    const int nl = d.counts->n_l;
    const int n = nl + d.counts->n_h;
    unsigned long long w[kWorkCounts] = {};
    GRID_STRIDE(k, n) {
        const int s = d.sscene[k];
        const SceneDev& sc = d.scenes[s];
        const float2 xi = d.spos[k];
        const uint8_t si = d.sstate[k];
        const bool high = k >= nl;
        const int lv = high ? 1 : 0;
        const bool i_prom = !(si & 1u);
        w[10 + lv] += 1;
        // same level (rows skipped exactly as the neighbour kernels do)
        {
            Box b = scan_box(sc, high, xi, high ? sc.H : sc.h);
            const uint8_t need = (high && i_prom) ? 1u : 0u;
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                int row = high ? sc.cbase_h + cy * sc.dhx : sc.cbase_l + cy * sc.dlx;
                row_run_need(d.cstart, d.cend, d.cflag, row, b.x0, b.x1, need, st, en);
                for (int j = st; j < en; ++j) {
                    float r2 = dist2(xi, d.spos[j]);
                    const uint8_t sj = d.sstate[j];
                    w[high ? 2 : 0] += 1;
                    bool lam = high ? !(i_prom && (sj & 1u)) : true;
                    if (lam && r2 <= (high ? sc.t_le_H : sc.t_le_h)) w[4 + lv] += 1;
                    if (lam && r2 != 0.0f && r2 <= (high ? sc.t_lt_H : sc.t_lt_h)) w[6 + lv] += 1;
                }
            }
        }
        // cross level
        if (high || d.scnt[s].seg_h1 > d.scnt[s].seg_h0) {
            // the density's high <- low scan: the low cells of the radius-H
            // box (its cutoff), the same rows' state filter
            if (high) {
                const Box bd = scan_box(sc, false, xi, sc.H);
                const uint8_t need = i_prom ? 0u : 1u;
                for (int cy = bd.y0; cy <= bd.y1; ++cy) {
                    int st, en;
                    row_run_need(d.cstart, d.cend, d.cflag, sc.cbase_l + cy * sc.dlx, bd.x0,
                                 bd.x1, need, st, en);
                    if (en > st) w[12] += (unsigned long long)(en - st);
                }
            }
            Box b = scan_box(sc, !high, xi, sc.h);
            const uint8_t need = high ? (i_prom ? 0u : 1u) : 1u;
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                int row = !high ? sc.cbase_h + cy * sc.dhx : sc.cbase_l + cy * sc.dlx;
                row_run_need(d.cstart, d.cend, d.cflag, row, b.x0, b.x1, need, st, en);
                for (int j = st; j < en; ++j) {
                    float r2 = dist2(xi, d.spos[j]);
                    const uint8_t sj = d.sstate[j];
                    w[high ? 3 : 1] += 1;
                    bool lam = high ? !((si & 1u) && (sj & 1u)) : !(sj & 1u);
                    if (lam && r2 <= (high ? sc.t_le_H : sc.t_le_h)) w[4 + lv] += 1;
                    if (lam && r2 != 0.0f && r2 <= sc.t_lt_h && (!high || sj != kDeleted))
                        w[6 + lv] += 1;
                }
            }
        }
        // wall candidates of the force kernel
        {
            Box b = scan_box(sc, false, xi, sc.h);
            for (int cy = b.y0; cy <= b.y1; ++cy) {
                int st, en;
                row_run(d.bstart, d.bend, sc.cbase_l + cy * sc.dlx, b.x0, b.x1, st, en);
                if (en > st) w[8 + lv] += (unsigned long long)(en - st);
            }
        }
    }
    for (int q = 0; q < kWorkCounts; ++q) {
        unsigned long long v = w[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&out[q], v);
    }
}

// Removal of a scene's entries from the lists: a stable filter from the
// current list buffers into the other ones, in three launches over the
// compaction's tiles (low, high, group segments): kept entries per tile
// (and, in the high segment, kept entries of its fresh tail), one block's
// prefix over the tiles with the new counts, then the scatter (each entry
// ranked by wave ballots within its tile, in index order).
__device__ __forceinline__ bool filter_keep(const DevState& d, int seg, int v, int scene) {
    const int lid = seg == 1 ? (v - d.CL) >> 2 : v;
    return d.scene_of[lid] != scene;
}

__global__ __launch_bounds__(kScanThreads) void k_filter_count(
    DevState d, StepShape sh, int scene, const int* __restrict__ ll,
    const int* __restrict__ lh, const int* __restrict__ lg) {
    __shared__ int wsum[2][kScanWaves];
    int seg, t;
    tile_seg(sh, blockIdx.x, seg, t);
    const int n = seg == 0 ? d.counts->n_l : seg == 1 ? d.counts->n_h : d.counts->n_g;
    const int tail0 = seg == 1 ? n - d.counts->n_fresh : n;
    const int* L = seg == 0 ? ll : seg == 1 ? lh : lg;
    int a = 0, f = 0;  // wave totals: kept entries, kept in the fresh tail
    for (int q = 0; q < kScanItems; ++q) {
        const int i = t * kScanTile + q * kScanThreads + (int)threadIdx.x;
        const bool keep = i < n && filter_keep(d, seg, L[i], scene);
        a += __popcll(__ballot(keep));
        f += __popcll(__ballot(keep && i >= tail0));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wsum[0][w] = a;
        wsum[1][w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tf = 0;
        for (int v = 0; v < kScanWaves; ++v) {
            ta += wsum[0][v];
            tf += wsum[1][v];
        }
        d.tile_sums[2 * blockIdx.x] = ta;
        d.tile_sums[2 * blockIdx.x + 1] = tf;
    }
}

// One block: per segment, exclusive prefix of the tiles' kept counts (the
// tiles' output bases) and the segment's new size; the high segment's kept
// fresh entries give the new n_fresh.
__global__ __launch_bounds__(1024) void k_filter_top(DevState d, StepShape sh) {
    __shared__ int part[1024];
    const int start[3] = {0, sh.scan_tiles_l, sh.scan_tiles_l + sh.scan_tiles_h};
    const int len[3] = {sh.scan_tiles_l, sh.scan_tiles_h, sh.scan_tiles_g};
    int fresh = 0;
    for (int sg = 0; sg < 3; ++sg) {
        const int nt = len[sg];
        const int per = (nt + 1023) / 1024;
        const int t0 = (int)threadIdx.x * per;
        int a = 0, f = 0;
        for (int i = 0; i < per; ++i) {
            if (t0 + i < nt) {
                a += d.tile_sums[2 * (start[sg] + t0 + i)];
                f += d.tile_sums[2 * (start[sg] + t0 + i) + 1];
            }
        }
        part[threadIdx.x] = a;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += v;
            __syncthreads();
        }
        int run = part[threadIdx.x] - a;
        const int tot = part[1023];
        for (int i = 0; i < per; ++i) {
            if (t0 + i < nt) {
                const int tt = start[sg] + t0 + i;
                const int v = d.tile_sums[2 * tt];
                d.tile_base[2 * tt] = run;
                run += v;
            }
        }
        // the fresh tail's kept entries (high segment only)
        if (sg == 1) {
            __syncthreads();
            part[threadIdx.x] = f;
            __syncthreads();
            for (int off = 512; off > 0; off >>= 1) {
                if (threadIdx.x < (unsigned)off) part[threadIdx.x] += part[threadIdx.x + off];
                __syncthreads();
            }
            fresh = part[0];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (sg == 0) d.counts->n_l = tot;
            if (sg == 1) {
                d.counts->n_h = tot;
                d.counts->n_fresh = fresh;
            }
            if (sg == 2) d.counts->n_g = tot;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void k_filter_scatter(
    DevState d, StepShape sh, int scene, const int* __restrict__ ll,
    const int* __restrict__ lh, const int* __restrict__ lg, int* __restrict__ nll,
    int* __restrict__ nlh, int* __restrict__ nlg, int n_l, int n_h, int n_g) {
    __shared__ int wsum[kScanWaves];
    int seg, t;
    tile_seg(sh, blockIdx.x, seg, t);
    // the segment sizes before the filter (k_filter_top has rewritten the
    // counts): passed by the host, which read them before the launch
    const int n = seg == 0 ? n_l : seg == 1 ? n_h : n_g;
    if (t * kScanTile >= n) return;  // whole block
    const int* L = seg == 0 ? ll : seg == 1 ? lh : lg;
    int* O = seg == 0 ? nll : seg == 1 ? nlh : nlg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int run = d.tile_base[2 * blockIdx.x];
    for (int q = 0; q < kScanItems; ++q) {
        const int i = t * kScanTile + q * kScanThreads + (int)threadIdx.x;
        const int v = i < n ? L[i] : 0;
        const bool keep = i < n && filter_keep(d, seg, v, scene);
        const uint64_t m = __ballot(keep);
        if (lane == 0) wsum[wv] = __popcll(m);
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w = 0; w < kScanWaves; ++w) {
            pre += w < wv ? wsum[w] : 0;
            tot += wsum[w];
        }
        __syncthreads();  // wsum is rewritten by the next chunk
        if (keep) O[run + pre + __popcll(m & lt)] = v;
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static inline int blocks_for(long long n, int bs, int cap) {
    long long b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// Plain streaming copy (16-byte vectors): the reference workload of the
// bandwidth probe and of the HBM counter calibration.
__global__ void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst,
                         long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy16(const void* src, void* dst, long long n16, void* stream) {
    k_copy16<<<blocks_for(n16, 256, 8192), 256, 0, (hipStream_t)stream>>>(
        (const float4*)src, (float4*)dst, n16);
}

// Sparse reads: one 8-byte load every `stride` 8-byte words of src (n
// loads), each lane its own address -- the reference workload of the HBM
// counter calibration for requests smaller than a cache line. The sink is
// written only if a value equals `never`, which the caller's fill excludes.
__global__ void k_gather_probe(const uint64_t* __restrict__ src, long long n, int stride,
                               uint64_t never, uint64_t* __restrict__ sink) {
    uint64_t acc = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        acc ^= src[i * stride];
    if (acc == never) sink[0] = acc;
}

void launch_gather_probe(const void* src, long long n, int stride, void* sink, void* stream) {
    k_gather_probe<<<blocks_for(n, 256, 8192), 256, 0, (hipStream_t)stream>>>(
        (const uint64_t*)src, n, stride, 0x5A5A5A5A5A5A5A5Bull, (uint64_t*)sink);
}

// Seeded dam-break block on the device: particle idx = j * nx + i at
// ((0.5 + i) dx, (0.5 + j) dx) + (2u - 1) * amp per axis, u = the top 24 bits
// of SplitMix64 outputs keyed by (seed, scene, idx) -- the same arithmetic,
// operation by operation, as the host generator tss_scene_dam_break.
__device__ __forceinline__ uint64_t mix64_dev(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_dam_break(float2* __restrict__ pos, int nx, int n,
                            float dx, float amp, uint64_t seed,
                            uint64_t scene) {
    constexpr uint64_t kG = 0x9E3779B97F4A7C15ull;
    GRID_STRIDE(idx, n) {
        const int j = idx / nx, i = idx - j * nx;
        uint64_t st = mix64_dev(mix64_dev(seed + kG * (scene + 1)) +
                                kG * ((uint64_t)idx + 1));
        st += kG;
        const float ux = (float)(mix64_dev(st) >> 40) * (1.0f / 16777216.0f);
        st += kG;
        const float uy = (float)(mix64_dev(st) >> 40) * (1.0f / 16777216.0f);
        const float jx = (2.0f * ux - 1.0f) * amp;
        const float jy = (2.0f * uy - 1.0f) * amp;
        pos[idx] = make_float2((0.5f + (float)i) * dx + jx,
                               (0.5f + (float)j) * dx + jy);
    }
}

__global__ void k_iota(int* __restrict__ p, int first, int n) {
    GRID_STRIDE(i, n) p[i] = first + i;
}

// ---------------------------------------------------------------------------
// Frame export: the entries of a level's current list that a scene owns
// (unified ids in [lo, hi)), in list order, gathered field by field into
// packed sections. Stable three-kernel selection: per-tile counts, one
// exclusive scan over the tiles, then a ballot prefix inside each tile.
// ---------------------------------------------------------------------------
constexpr int kFrameThreads = 256;
constexpr int kFrameChunks = 4;
constexpr int kFrameTile = kFrameThreads * kFrameChunks;

__global__ __launch_bounds__(kFrameThreads) void k_frame_count(
    const int* __restrict__ list, int n, int lo, int hi, int* __restrict__ cnt) {
    __shared__ int wsum[kFrameThreads / 64];
    int c = 0;
    for (int k = 0; k < kFrameChunks; ++k) {
        const int i = blockIdx.x * kFrameTile + k * kFrameThreads + threadIdx.x;
        if (i < n) {
            const int v = list[i];
            c += v >= lo && v < hi;
        }
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kFrameThreads / 64; ++w) t += wsum[w];
        cnt[blockIdx.x] = t;
    }
}

// in-place exclusive scan of cnt[0..nt); cnt[nt] = total. One block.
__global__ __launch_bounds__(1024) void k_frame_scan(int* __restrict__ cnt, int nt) {
    __shared__ int part[1024];
    __shared__ int base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nt; b0 += 1024) {
        const int i = b0 + threadIdx.x;
        const int v = i < nt ? cnt[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int a = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nt) cnt[i] = base + part[threadIdx.x] - v;
        const int tot = part[1023];
        __syncthreads();
        if (threadIdx.x == 0) base += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[nt] = base;
}

__global__ __launch_bounds__(kFrameThreads) void k_frame_pack(
    DevState d, const int* __restrict__ list, int n, int lo, int hi,
    const int* __restrict__ offs, int n_out, FrameOut out) {
    __shared__ int wsum[kFrameThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int base = offs[blockIdx.x];
    for (int k = 0; k < kFrameChunks; ++k) {
        const int i = blockIdx.x * kFrameTile + k * kFrameThreads + threadIdx.x;
        const int v = i < n ? list[i] : -1;
        const bool keep = i < n && v >= lo && v < hi;
        const unsigned long long m = __ballot(keep);
        const int wpre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int bpre = 0, tot = 0;
        for (int q = 0; q < kFrameThreads / 64; ++q) {
            bpre += q < w ? wsum[q] : 0;
            tot += wsum[q];
        }
        const int o = base + bpre + wpre;
        if (keep && o < n_out) {
            if (out.id) out.id[o] = v - lo;
            if (out.pos) out.pos[o] = d.pos[v];
            if (out.vel) out.vel[o] = d.vel[v];
            if (out.rho) out.rho[o] = d.rho[v];
            if (out.p) out.p[o] = d.p[v];
            if (out.vis) out.vis[o] = d.vis[v];
            if (out.state) out.state[o] = d.state[v];
        }
        base += tot;
        __syncthreads();
    }
}

__global__ void k_frame_bnd(const float2* __restrict__ bpos, const int* __restrict__ bids,
                            int nb, float2* __restrict__ out) {
    GRID_STRIDE(k, nb) {
        const int i = bids[k];
        if (i >= 0 && i < nb) out[i] = bpos[k];
    }
}

void launch_frame_bnd(const float2* bpos, const int* bids, int nb, float2* out,
                      void* stream) {
    if (nb <= 0) return;
    k_frame_bnd<<<blocks_for(nb, 256, 1024), 256, 0, (hipStream_t)stream>>>(bpos, bids, nb, out);
}

// Staging of the frame export that allocates nothing. Its address differs
// per device and is known to host code only through hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        ptr[device] = p;
    }
    *out = ptr[device];
    return hipSuccess;
}

void launch_dam_break(float2* pos, int nx, int ny, float dx, float jitter,
                      unsigned long long seed, unsigned long long scene,
                      void* stream) {
    const int n = nx * ny;
    if (n <= 0) return;
    k_dam_break<<<blocks_for(n, 256, 4096), 256, 0, (hipStream_t)stream>>>(
        pos, nx, n, dx, jitter * dx, seed, scene);
}

void launch_iota(int* p, int first, int n, void* stream) {
    if (n <= 0) return;
    k_iota<<<blocks_for(n, 256, 4096), 256, 0, (hipStream_t)stream>>>(p, first, n);
}

void launch_fill_int(int* p, int value, long long n, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    k_fill_int<<<blocks_for(n, 256, 4096), 256, 0, st>>>(p, value, n);
}

// Grid build: clears, hash, sort, ranges + reorder. Returns the buffer index
// holding the sorted keys/vals.
// with_sidx: also the sorted index of each id (sidx), which only the band
// rule's merge test, the relaxation of new children and the neighbour
// inspection read
static int enqueue_grid(const DevState& d, const StepShape& sh,
                        hipStream_t st, int with_sidx) {
    const int par = sh.list_parity;
    k_begin_hash<<<sh.clear_blocks + sh.grid_elem, 256, 0, st>>>(
        d, sh.clear_blocks, sh.count_sort, sh.solo, d.list_l[par], d.list_h[par], d.keys[0],
        d.vals[0]);
    if (sh.count_sort) {
        k_cell_sums<<<sh.cell_tiles, kCellThreads, 0, st>>>(d, sh.key_cells);
        k_cell_ranges<<<sh.cell_tiles, kCellThreads, 0, st>>>(d, sh.key_cells);
        k_cell_scatter<<<sh.grid_elem, 256, 0, st>>>(d, d.keys[0], d.list_l[par],
                                                     d.list_h[par], d.keys[1]);
        k_cell_reorder<<<sh.grid_elem, 256, 0, st>>>(d, d.keys[1], d.vals[1], with_sidx);
        return 1;
    }
    int cur = 0;
    if (sh.small_sort) {
        k_sort_small<<<1, kSmallSortThreads, 0, st>>>(d, sh.passes);
        cur = sh.passes & 1;
    }
    for (int pass = 0; !sh.small_sort && pass < sh.passes; ++pass) {
        const int shift = kRadixBits * pass;
        k_sort_hist<<<sh.sort_grid, kSortThreads, 0, st>>>(
            d.keys[cur], d.counts, shift, d.hist, sh.sort_tiles);
        k_sort_scan<<<kRadix, 256, 0, st>>>(d.hist, d.counts, sh.sort_tiles,
                                             d.digit_tot);
        k_sort_scatter<<<sh.sort_grid, kSortThreads, 0, st>>>(
            d.keys[cur], d.vals[cur], d.keys[cur ^ 1], d.vals[cur ^ 1],
            d.counts, shift, d.hist, d.digit_tot, sh.sort_tiles);
        cur ^= 1;
    }
    k_reorder<<<sh.grid_elem, 256, 0, st>>>(d, d.keys[cur], d.vals[cur], with_sidx);
    return cur;
}

void launch_grid_only(const DevState& d, const StepShape& sh, void* stream) {
    enqueue_grid(d, sh, (hipStream_t)stream, 0);
}

void launch_step(const DevState& d, const StepShape& sh, void* stream,
                 void* timing_events, unsigned int event_flags) {
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t* ev = (hipEvent_t*)timing_events;
    // event points: see kTimingPoints (tss_device.h)
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecordWithFlags(ev[k], st, event_flags);
    };
    const int par = sh.list_parity;
    mark(0);
    const int cur = enqueue_grid(d, sh, st, sh.band_any || sh.relax_iters > 0);
    const uint32_t* vals = d.vals[cur];
    const uint32_t* keys = d.keys[cur];
    for (int it = 0; it < sh.relax_iters; ++it) {
        const int gr = blocks_for(4 * sh.cap_alive, 256, 4096);
        k_relax_density<<<gr, 256, 0, st>>>(d, d.list_h[par], it);
        k_relax_force<<<gr, 256, 0, st>>>(d, vals, d.list_h[par], it);
        k_relax_commit<<<gr, 256, 0, st>>>(d, d.list_h[par], it);
    }
    mark(1);
    launch_pair<false, false>(d, vals, keys, sh.tune[kTuneDensityLow], sh.cap_alive, st);
    mark(2);
    launch_pair<false, true>(d, vals, keys, sh.tune[kTuneDensityHigh], 4 * sh.cap_alive, st);
    mark(3);
    launch_pair<true, false>(d, vals, keys, sh.tune[kTuneForceLow], sh.cap_alive, st);
    mark(4);
    launch_pair<true, true>(d, vals, keys, sh.tune[kTuneForceHigh], 4 * sh.cap_alive, st);
    mark(5);
    k_reduce<<<sh.red_slots, 256, 0, st>>>(d);
    mark(6);
    k_split_cand<<<sh.grid_elem, 256, 0, st>>>(d);
    // few groups and long band tests: a lane per child, small blocks to
    // spread them over the CUs
    if (sh.band_any)
        k_groups<4><<<blocks_for(4 * sh.cap_alive, 64, 16384), 64, 0, st>>>(d, d.list_g[par],
                                                                           sh.solo);
    else
        k_groups<1><<<blocks_for(sh.cap_alive, 256, 4096), 256, 0, st>>>(d, d.list_g[par],
                                                                         sh.solo);
    // the cap can bind only where cap_high < 4 cap_low: a scene's groups
    // and candidates never exceed its low ids
    if (sh.cap_binds) k_cap<<<sh.S, 256, 0, st>>>(d);
    const int ntiles = sh.scan_tiles_l + sh.scan_tiles_h + sh.scan_tiles_g;
    k_compact_reduce<<<ntiles, kScanThreads, 0, st>>>(d, sh, vals,
                                                      d.list_g[par]);
    k_compact_top<<<1, 1024, 0, st>>>(d, sh);
    k_compact_scatter<<<ntiles, kScanThreads, 0, st>>>(
        d, sh, vals, d.list_g[par], d.list_l[par ^ 1], d.list_h[par ^ 1],
        d.list_g[par ^ 1]);
    mark(7);
    k_finalize<<<sh.S, 256, 0, st>>>(d);
    mark(8);
}

void launch_frame_pack(const DevState& d, const int* list, int n, int lo, int hi,
                       int* tile_cnt, int n_out, const FrameOut& out, void* stream) {
    if (n <= 0) return;
    hipStream_t st = (hipStream_t)stream;
    const int nt = (n + kFrameTile - 1) / kFrameTile;
    k_frame_count<<<nt, kFrameThreads, 0, st>>>(list, n, lo, hi, tile_cnt);
    k_frame_scan<<<1, 1024, 0, st>>>(tile_cnt, nt);
    k_frame_pack<<<nt, kFrameThreads, 0, st>>>(d, list, n, lo, hi, tile_cnt, n_out, out);
}

int frame_tiles(int n) { return (n + kFrameTile - 1) / kFrameTile; }

void launch_filter_scene(const DevState& d, const StepShape& sh, int scene,
                         int n_l, int n_h, int n_g, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int par = sh.list_parity;
    const int ntiles = sh.scan_tiles_l + sh.scan_tiles_h + sh.scan_tiles_g;
    k_filter_count<<<ntiles, kScanThreads, 0, st>>>(d, sh, scene, d.list_l[par],
                                                     d.list_h[par], d.list_g[par]);
    k_filter_top<<<1, 1024, 0, st>>>(d, sh);
    k_filter_scatter<<<ntiles, kScanThreads, 0, st>>>(
        d, sh, scene, d.list_l[par], d.list_h[par], d.list_g[par], d.list_l[par ^ 1],
        d.list_h[par ^ 1], d.list_g[par ^ 1], n_l, n_h, n_g);
}

void launch_work_counts(const DevState& d, const StepShape& sh,
                        unsigned long long* out, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    enqueue_grid(d, sh, st, 0);
    (void)hipMemsetAsync(out, 0, kWorkCounts * sizeof(unsigned long long), st);
    k_work<<<sh.grid_elem, 256, 0, st>>>(d, out);
}

// Pair recording (tests): the grid of the current lists, then both levels'
// density and force kernels in their recording variant.
void launch_pair_record(const DevState& d, const StepShape& sh, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int cur = enqueue_grid(d, sh, st, 0);
    const uint32_t* vals = d.vals[cur];
    const uint32_t* keys = d.keys[cur];
    const int gl = blocks_for(sh.cap_alive, 256, 16384);
    const int gh = blocks_for(4 * sh.cap_alive, 256, 16384);
    k_density<false, 256, 16, true><<<gl, 256, 0, st>>>(d, vals, keys, 0);
    k_density<true, 256, 16, true><<<gh, 256, 0, st>>>(d, vals, keys, 0);
    k_force<false, 256, 16, true><<<gl, 256, 0, st>>>(d, vals, 0);
    k_force<true, 256, 16, true><<<gh, 256, 0, st>>>(d, vals, 0);
}

// Neighbour inspection of the first n entries of a level's current list.
void launch_neighbors_n(const DevState& d, const StepShape& sh, int scene,
                        int level, int n, uint32_t* counts, uint32_t* sums,
                        void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int cur = enqueue_grid(d, sh, st, 1);
    const int par = sh.list_parity;
    const int* list = level ? d.list_h[par] : d.list_l[par];
    k_neighbors<<<blocks_for(n, 256, 4096), 256, 0, st>>>(d, d.vals[cur], list,
                                                          n, scene, counts, sums);
}

}  // namespace tss
