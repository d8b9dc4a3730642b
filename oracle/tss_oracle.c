/*
 * tss_oracle.c -- index, decision and update parts of the CPU restatement.
 * TEST INFRASTRUCTURE ONLY (see tss_oracle.h). Compiled with
 * -ffp-contract=off: every float expression below rounds each operation,
 * which is the method's definition of the bit-exact index arithmetic.
 *
 * R = TwoScale2D/WCSPHSolver.cu of the reference.
 */
#include "tss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846 /* R:19 (double literal) */

/* ------------------------------------------------------------------------ */
/* Parameters: WCSPHConfig's constructor, R:2287-2332.                       */
/* ------------------------------------------------------------------------ */
int orc_params_init(orc_params* p, float xs, float ys, float xe, float ye,
                    float h, float H, float rho0, float B, float c,
                    float alpha, float sigma)
{
    memset(p, 0, sizeof(*p));
    if (xs >= xe || ys >= ye || !(h > 0.0f) || !(H > 0.0f)) {
        return -1; /* R:2312-2316 throws "Invalid configuration parameters" */
    }
    p->origin[0] = xs;
    p->origin[1] = ys;
    p->end[0] = xe;
    p->end[1] = ye;
    /* dims = ceil(FP32 quotient), R:2323-2331 */
    p->dims_l[0] = (int)ceilf((xe - xs) / h);
    p->dims_l[1] = (int)ceilf((ye - ys) / h);
    p->dims_h[0] = (int)ceilf((xe - xs) / H);
    p->dims_h[1] = (int)ceilf((ye - ys) / H);
    p->h = h;
    p->H = H;
    p->rho0 = rho0;
    p->B = B;
    p->c = c;
    p->alpha = alpha;
    p->sigma = sigma;
    p->g = 9.81f;              /* R:1805, 2012 */
    p->mode = ORC_MODE_PLAIN;
    p->split_x = 0.5f;         /* R:1835 */
    p->surf_ratio = 0.9f;
    p->surf_ratio_h = 0.9f;
    p->r_band = h;
    p->r_merge = h;
    p->promote_min_age = 5;
    p->promote_max_age = 0;    /* promotion off */
    p->promote_fluc = 0.05f;
    p->cap_h = 0;
    p->dt_mode = ORC_DT_FIXED;
    p->cfl = 0.4f;
    p->dt_max = 1e-3f;
    p->relax_iters = 0;
    p->relax_frac = 0.1f;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Grid coordinates and hashes, R:241-323.                                   */
/* coord = clamp((int)(((x + off) - origin)/cell), 0, dim-1): truncation     */
/* toward zero, then clamp. The clamp is applied in float first so that     */
/* non-finite or huge quotients are defined (NaN -> 0).                      */
/* ------------------------------------------------------------------------ */
int orc_axis_cell(float x, float off, float origin, float cell, int dim)
{
    float t = ((x + off) - origin) / cell;
    if (!(t >= 1.0f)) {
        return 0;
    }
    if (t >= (float)dim) {
        return dim - 1;
    }
    return (int)t;
}

static void level_grid(const orc_params* p, int level, float* cell,
                       const int** dims)
{
    *cell = level ? p->H : p->h;
    *dims = level ? p->dims_h : p->dims_l;
}

int orc_hash(const orc_params* p, int level, float x, float y)
{
    float cell;
    const int* d;
    level_grid(p, level, &cell, &d);
    int cx = orc_axis_cell(x, 0.0f, p->origin[0], cell, d[0]);
    int cy = orc_axis_cell(y, 0.0f, p->origin[1], cell, d[1]);
    return cy * d[0] + cx;
}

void orc_hash_list(const orc_params* p, int level, const float* pos,
                   const int32_t* ids, int n, int32_t* keys)
{
    for (int k = 0; k < n; k++) {
        int id = ids[k];
        keys[k] = orc_hash(p, level, pos[2 * id], pos[2 * id + 1]);
    }
}

/* Stable sort of (key, value) by key: a counting sort, which keeps input
 * order among equal keys -- the property thrust::sort_by_key provides at
 * R:2828-2834 (it forwards to stable_sort_by_key). */
void orc_stable_sort(const int32_t* keys, const int32_t* vals, int n,
                     int nkeys, int32_t* keys_out, int32_t* vals_out)
{
    int32_t* cnt = (int32_t*)calloc((size_t)nkeys + 1, sizeof(int32_t));
    for (int k = 0; k < n; k++) {
        cnt[keys[k] + 1]++;
    }
    for (int c = 0; c < nkeys; c++) {
        cnt[c + 1] += cnt[c];
    }
    for (int k = 0; k < n; k++) {
        int dst = cnt[keys[k]]++;
        keys_out[dst] = keys[k];
        vals_out[dst] = vals[k];
    }
    free(cnt);
}

/* Cell start/end over sorted keys; empty cells = -1 (R:18, 1300-1340,
 * 2840-2843). */
void orc_cell_ranges(const int32_t* sk, int n, int ncells, int32_t* start,
                     int32_t* end)
{
    for (int c = 0; c < ncells; c++) {
        start[c] = -1;
        end[c] = -1;
    }
    for (int k = 0; k < n; k++) {
        if (k == 0 || sk[k] != sk[k - 1]) {
            start[sk[k]] = k;
            if (k > 0) {
                end[sk[k - 1]] = k;
            }
        }
        if (k == n - 1) {
            end[sk[k]] = n;
        }
    }
}

void orc_build_grid(const orc_params* p, const orc_state* s, orc_grid* g)
{
    int ncl = p->dims_l[0] * p->dims_l[1];
    int nch = p->dims_h[0] * p->dims_h[1];
    int nmax = s->n_l > s->n_h ? s->n_l : s->n_h;
    int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nmax + 1));
    orc_hash_list(p, 0, s->pos_l, s->list_l, s->n_l, keys);
    orc_stable_sort(keys, s->list_l, s->n_l, ncl, g->keys_l, g->sorted_l);
    orc_cell_ranges(g->keys_l, s->n_l, ncl, g->start_l, g->end_l);
    orc_hash_list(p, 1, s->pos_h, s->list_h, s->n_h, keys);
    orc_stable_sort(keys, s->list_h, s->n_h, nch, g->keys_h, g->sorted_h);
    orc_cell_ranges(g->keys_h, s->n_h, nch, g->start_h, g->end_h);
    free(keys);
}

/* Boundary grid: iota ids, low-grid hash, stable sort, ranges
 * (initBoundaries, R:3205-3238). */
void orc_build_boundary_grid(const orc_params* p, const orc_state* s,
                             orc_grid* g)
{
    int ncl = p->dims_l[0] * p->dims_l[1];
    int n = s->n_b;
    int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* sk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int k = 0; k < n; k++) {
        ids[k] = k;
    }
    orc_hash_list(p, 0, s->pos_b, ids, n, keys);
    orc_stable_sort(keys, ids, n, ncl, sk, g->bsorted);
    orc_cell_ranges(sk, n, ncl, g->bstart, g->bend);
    free(ids);
    free(keys);
    free(sk);
}

/* ------------------------------------------------------------------------ */
/* Neighbour sets.                                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int x0, x1, y0, y1, dx;
} box2;

/* Scan box: cells cell(x - r) .. cell(x + r) per axis on the given level's
 * grid (R:1374-1375, 1392-1393, 1457-1458, 1483-1484). */
static box2 scan_box(const orc_params* p, int level, float x, float y,
                     float r)
{
    float cell;
    const int* d;
    level_grid(p, level, &cell, &d);
    box2 b;
    b.x0 = orc_axis_cell(x, -r, p->origin[0], cell, d[0]);
    b.x1 = orc_axis_cell(x, r, p->origin[0], cell, d[0]);
    b.y0 = orc_axis_cell(y, -r, p->origin[1], cell, d[1]);
    b.y1 = orc_axis_cell(y, r, p->origin[1], cell, d[1]);
    b.dx = d[0];
    return b;
}

/* FP32 pair distance of the method: sqrt(dx*dx + dy*dy), each op rounded
 * (R:325-333). */
static inline float dist_f(float xi, float yi, float xj, float yj)
{
    float dx = xi - xj;
    float dy = yi - yj;
    float a = dx * dx;
    float b = dy * dy;
    return sqrtf(a + b);
}

static inline uint32_t csum(int id)
{
    return (uint32_t)(id + 1) * 2654435761u;
}

/* State weights of the pair sums, index 2(s_i & 1) + (s_j & 1) for the
 * lambda tables: lambda_LH (R:377, 602), lambda_HH (R:414, 681), lambda_HL
 * (R:451, 808); the deleted weight d(s_j) = {1, 1, 0} on forces from low
 * particles (R:500, 809). */
static const int kLamLH[4] = {1, 0, 1, 0};
static const int kLamHH[4] = {1, 0, 1, 1};
static const int kLamHL[4] = {1, 1, 1, 0};
static const int kDel[4] = {1, 1, 0, 0};
#define LAM(tab, si, sj) (tab[(((si) & 1u) << 1) | ((sj) & 1u)])

/* weighted = 0: geometric sets (every candidate within the cutoff);
 * weighted = 1: only pairs whose weight in the density or force sum is
 * nonzero, i.e. the terms those sums add. */
static void neighbour_sets(const orc_params* p, const orc_state* s,
                           const orc_grid* g, int level, int weighted,
                           uint32_t* counts, uint32_t* sums)
{
    int n = level ? s->n_h : s->n_l;
    const int32_t* list = level ? s->list_h : s->list_l;
    const float* pos = level ? s->pos_h : s->pos_l;
    const float h = p->h, H = p->H;
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < n; t++) {
        int i = list[t];
        float xi = pos[2 * i], yi = pos[2 * i + 1];
        const uint8_t si = level ? s->state_h[i] : s->state_l[i];
        uint32_t c[5] = {0, 0, 0, 0, 0}, su[5] = {0, 0, 0, 0, 0};
        /* same-level box, cross-level box, boundary (low) box */
        float rs = level ? H : h;
        box2 b = scan_box(p, level, xi, yi, rs);
        const int32_t* st = level ? g->start_h : g->start_l;
        const int32_t* en = level ? g->end_h : g->end_l;
        const int32_t* so = level ? g->sorted_h : g->sorted_l;
        float rd = level ? H : h;      /* density cutoff (<=)  */
        float rf = level ? H : h;      /* force cutoff (<)     */
        for (int cy = b.y0; cy <= b.y1; cy++) {
            for (int cx = b.x0; cx <= b.x1; cx++) {
                int cc = cy * b.dx + cx;
                for (int k = st[cc]; k < en[cc]; k++) {
                    int j = so[k];
                    float r = dist_f(xi, yi, pos[2 * j], pos[2 * j + 1]);
                    int wd = 1, wf = 1;
                    if (weighted) {
                        uint8_t sj = level ? s->state_h[j] : s->state_l[j];
                        /* LL: density every state (R:335-363), force d(s_j);
                         * HH: lambda_HH both */
                        wd = level ? LAM(kLamHH, si, sj) : 1;
                        wf = level ? LAM(kLamHH, si, sj) : kDel[sj & 3u];
                    }
                    if (wd && r <= rd) { c[0]++; su[0] += csum(j); }
                    if (wf && r != 0.0f && r < rf) { c[2]++; su[2] += csum(j); }
                }
            }
        }
        /* cross level: low target scans high grid with +-h; high target
         * scans the low grid with +-h */
        int ol = 1 - level;
        box2 bo = scan_box(p, ol, xi, yi, h);
        const int32_t* st2 = ol ? g->start_h : g->start_l;
        const int32_t* en2 = ol ? g->end_h : g->end_l;
        const int32_t* so2 = ol ? g->sorted_h : g->sorted_l;
        const float* pos2 = ol ? s->pos_h : s->pos_l;
        float rd2 = level ? H : h;     /* HL dens r<=H ; LH dens r<=h */
        float rf2 = h;                 /* both cross forces r<h       */
        for (int cy = bo.y0; cy <= bo.y1; cy++) {
            for (int cx = bo.x0; cx <= bo.x1; cx++) {
                int cc = cy * bo.dx + cx;
                for (int k = st2[cc]; k < en2[cc]; k++) {
                    int j = so2[k];
                    float r = dist_f(xi, yi, pos2[2 * j], pos2[2 * j + 1]);
                    int wd = 1, wf = 1;
                    if (weighted) {
                        uint8_t sj = ol ? s->state_h[j] : s->state_l[j];
                        /* LH: lambda_LH both; HL: lambda_HL, times d(s_j)
                         * on the force */
                        wd = level ? LAM(kLamHL, si, sj) : LAM(kLamLH, si, sj);
                        wf = level ? LAM(kLamHL, si, sj) * kDel[sj & 3u]
                                   : LAM(kLamLH, si, sj);
                    }
                    if (wd && r <= rd2) { c[1]++; su[1] += csum(j); }
                    if (wf && r != 0.0f && r < rf2) { c[3]++; su[3] += csum(j); }
                }
            }
        }
        /* boundary: low cells of the +-h box */
        box2 bb = scan_box(p, 0, xi, yi, h);
        for (int cy = bb.y0; cy <= bb.y1; cy++) {
            for (int cx = bb.x0; cx <= bb.x1; cx++) {
                int cc = cy * bb.dx + cx;
                for (int k = g->bstart[cc]; k < g->bend[cc]; k++) {
                    c[4]++;
                    su[4] += csum(g->bsorted[k]);
                }
            }
        }
        for (int q = 0; q < 5; q++) {
            counts[5 * t + q] = c[q];
            sums[5 * t + q] = su[q];
        }
    }
}

void orc_neighbors(const orc_params* p, const orc_state* s,
                   const orc_grid* g, int level, uint32_t* counts,
                   uint32_t* sums)
{
    neighbour_sets(p, s, g, level, 0, counts, sums);
}

void orc_pair_sets(const orc_params* p, const orc_state* s,
                   const orc_grid* g, int level, uint32_t* counts,
                   uint32_t* sums)
{
    neighbour_sets(p, s, g, level, 1, counts, sums);
}

/* ------------------------------------------------------------------------ */
/* Symplectic Euler: v += dt a; x += dt v (R:1635-1639, 2014-2018).          */
/* ------------------------------------------------------------------------ */
void orc_integrate(const orc_params* p, orc_state* s, const float* acc_l,
                   const float* acc_h)
{
    (void)p;
    const float dt = s->dt;
    for (int lv = 0; lv < 2; lv++) {
        int n = lv ? s->n_h : s->n_l;
        const int32_t* list = lv ? s->list_h : s->list_l;
        float* pos = lv ? s->pos_h : s->pos_l;
        float* vel = lv ? s->vel_h : s->vel_l;
        float* acc = lv ? s->acc_h : s->acc_l;
        const float* a = lv ? acc_h : acc_l;
        for (int k = 0; k < n; k++) {
            int i = list[k];
            float ax = a[2 * i], ay = a[2 * i + 1];
            float vx = vel[2 * i] + dt * ax;
            float vy = vel[2 * i + 1] + dt * ay;
            pos[2 * i] = pos[2 * i] + dt * vx;
            pos[2 * i + 1] = pos[2 * i + 1] + dt * vy;
            vel[2 * i] = vx;
            vel[2 * i + 1] = vy;
            acc[2 * i] = ax;
            acc[2 * i + 1] = ay;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Adaptivity.                                                               */
/* ------------------------------------------------------------------------ */

/* Is any surface particle within distance r of (x, y)? Surface set: listed
 * low particles and promoted (state 0) high particles whose FP32 density is
 * below surf_ratio*rho0. Positions are the pre-step ones the grid holds. */
static int in_band(const orc_params* p, const orc_grid* g, const orc_state* s,
                   const float* ppl, const float* pph, const uint8_t* surf_l,
                   const uint8_t* surf_h, float x, float y, float r)
{
    box2 b = scan_box(p, 0, x, y, r);
    for (int cy = b.y0; cy <= b.y1; cy++) {
        for (int cx = b.x0; cx <= b.x1; cx++) {
            int cc = cy * b.dx + cx;
            for (int k = g->start_l[cc]; k < g->end_l[cc]; k++) {
                int j = g->sorted_l[k];
                if (surf_l[j] &&
                    dist_f(x, y, ppl[2 * j], ppl[2 * j + 1]) <= r) {
                    return 1;
                }
            }
        }
    }
    if (s->n_h > 0) {
        box2 bh = scan_box(p, 1, x, y, r);
        for (int cy = bh.y0; cy <= bh.y1; cy++) {
            for (int cx = bh.x0; cx <= bh.x1; cx++) {
                int cc = cy * bh.dx + cx;
                for (int k = g->start_h[cc]; k < g->end_h[cc]; k++) {
                    int j = g->sorted_h[k];
                    if (surf_h[j] &&
                        dist_f(x, y, pph[2 * j], pph[2 * j + 1]) <= r) {
                        return 1;
                    }
                }
            }
        }
    }
    return 0;
}

/* Does a wall particle lie within sqrt-distance h of (x, y)? (low box +-h).
 * Walls carry no density in the method, so a density deficit next to a
 * wall is not a free surface. */
static int near_wall(const orc_params* p, const orc_grid* g,
                     const orc_state* s, float x, float y)
{
    box2 b = scan_box(p, 0, x, y, p->h);
    for (int cy = b.y0; cy <= b.y1; cy++) {
        for (int cx = b.x0; cx <= b.x1; cx++) {
            int cc = cy * b.dx + cx;
            for (int k = g->bstart[cc]; k < g->bend[cc]; k++) {
                int j = g->bsorted[k];
                if (dist_f(x, y, s->pos_b[2 * j], s->pos_b[2 * j + 1]) <=
                    p->h) {
                    return 1;
                }
            }
        }
    }
    return 0;
}

/* Child offset radius r_c = 0.5*sqrt(m/(2 rho pi)), evaluated like R:1848
 * (float 2*rho, then double arithmetic with double pi), rounded to float. */
static float child_radius(float m_l, float rho)
{
    float two_rho = 2.0f * rho;
    return (float)(0.5 * sqrt((double)m_l / ((double)two_rho * ORC_PI)));
}

/* R's adjustPosition (R:1007-1022): clamp the child's distance to its
 * parent at 0.75*sqrt(m/(rho pi)). The bound is a double (m/(rho*PI) with
 * the double PI, R:1018), the minimum min(float, double) is a double, and
 * it is stored to `float r`; then r/dist and the product with posji are
 * FP32 (R:1020-1021), each operation rounded. A coincident child (dist 0,
 * where the reference divides 0/0) is left in place. */
static void clamp_child(float* ph, const float* pl, float m_l, float rho)
{
    float dx = ph[0] - pl[0];
    float dy = ph[1] - pl[1];
    float a = dx * dx;
    float b = dy * dy;
    float dist = sqrtf(a + b);
    if (!(dist > 0.0f)) {
        return;
    }
    double rmax = 0.75 * sqrt((double)m_l / ((double)rho * ORC_PI));
    float r = (float)((double)dist < rmax ? (double)dist : rmax);
    float t = r / dist;
    float ox = t * dx;
    float oy = t * dy;
    ph[0] = pl[0] + ox;
    ph[1] = pl[1] + oy;
}

void orc_adapt(const orc_params* p, orc_state* s, const orc_grid* g,
               const float* ppl, const float* pph, orc_adapt_out* out)
{
    const int cap_l = s->cap_l;
    out->n_split = out->n_promote = out->n_merge = out->n_dissolve = 0;
    if (p->mode == ORC_MODE_PLAIN) {
        memcpy(s->list_l, g->sorted_l, sizeof(int32_t) * (size_t)s->n_l);
        memcpy(s->list_h, g->sorted_h, sizeof(int32_t) * (size_t)s->n_h);
        return;
    }
    const float ts = p->surf_ratio * p->rho0;
    const float tsh = p->surf_ratio_h * p->rho0;
    uint8_t* surf_l = out->surf_l;
    uint8_t* surf_h = out->surf_h;
    memset(surf_l, 0, (size_t)cap_l);
    memset(surf_h, 0, (size_t)cap_l * 4);
    for (int k = 0; k < s->n_l; k++) {
        int i = g->sorted_l[k];
        surf_l[i] = s->rho_l[i] < ts &&
                    !near_wall(p, g, s, ppl[2 * i], ppl[2 * i + 1]);
    }
    for (int k = 0; k < s->n_h; k++) {
        int a = g->sorted_h[k];
        surf_h[a] = s->state_h[a] == 0 && s->rho_h[a] < tsh &&
                    !near_wall(p, g, s, pph[2 * a], pph[2 * a + 1]);
    }

    /* 1. split candidates in sorted order, admitted against the cap
     * (cap_h 0: no cap, every child slot 4 cap_l) */
    const long long cap_h = p->cap_h > 0 ? (long long)p->cap_h : 4LL * s->cap_l;
    for (int k = 0; k < s->n_l; k++) {
        int i = g->sorted_l[k];
        if (s->state_l[i] != ORC_ACTIVE) {
            continue;
        }
        int cand = 0;
        if (p->mode == ORC_MODE_REFERENCE) {
            cand = s->pos_l[2 * i] > p->split_x; /* predicted x' (R:1835) */
        } else {
            cand = in_band(p, g, s, ppl, pph, surf_l, surf_h, ppl[2 * i],
                           ppl[2 * i + 1], p->r_band);
        }
        if (cand && 4LL * (s->n_g + out->n_split + 1) <= cap_h) {
            out->split_ids[out->n_split++] = i;
        }
    }

    /* 2. existing groups, in group-list order, with step-start states.
     * merged[p]: 1 = promoted group merged back (parent re-listed),
     *            2 = transient group dissolved (parent already listed). */
    uint8_t* merged = (uint8_t*)calloc((size_t)cap_l, 1);
    for (int q = 0; q < s->n_g; q++) {
        int pid = s->list_g[q];
        if (s->state_l[pid] == ORC_TRANSIENT) {
            s->age_l[pid] += 1;
            int promote = 0;
            if (p->mode == ORC_MODE_BAND && p->promote_max_age > 0) {
                /* children's density fluctuation, as the reference's dormant
                 * promotion test measures it (R:1073-1082) */
                float fluc = 0.0f;
                for (int k = 0; k < 4; k++) {
                    fluc += fabsf(p->rho0 - s->rho_h[4 * pid + k]);
                }
                fluc = fluc / (4.0f * p->rho0);
                int age = s->age_l[pid];
                promote = (age >= p->promote_min_age &&
                           fluc <= p->promote_fluc) ||
                          age >= p->promote_max_age;
            }
            if (promote) {
                s->state_l[pid] = ORC_DELETED;
                for (int k = 0; k < 4; k++) {
                    s->state_h[4 * pid + k] = 0;
                }
                out->promote_ids[out->n_promote++] = pid;
                continue;
            }
            int keep = 1;
            if (p->mode == ORC_MODE_BAND) {
                keep = 0;
                for (int k = 0; k < 4 && !keep; k++) {
                    int a = 4 * pid + k;
                    keep = in_band(p, g, s, ppl, pph, surf_l, surf_h,
                                   pph[2 * a], pph[2 * a + 1], p->r_merge);
                }
            }
            if (!keep) {
                /* dissolve: children removed, the parent stays listed */
                merged[pid] = 2;
                s->state_l[pid] = ORC_ACTIVE;
                s->age_l[pid] = 0;
                out->dissolve_ids[out->n_dissolve++] = pid;
            } else {
                /* child clamp (R:1097-1172) */
                for (int k = 0; k < 4; k++) {
                    clamp_child(&s->pos_h[2 * (4 * pid + k)],
                                &s->pos_l[2 * pid], p->m_l, s->rho_l[pid]);
                }
            }
        } else if (s->state_l[pid] == ORC_DELETED &&
                   p->mode == ORC_MODE_BAND) {
            int keep = 0;
            for (int k = 0; k < 4 && !keep; k++) {
                int a = 4 * pid + k;
                keep = in_band(p, g, s, ppl, pph, surf_l, surf_h, pph[2 * a],
                               pph[2 * a + 1], p->r_merge);
            }
            if (!keep) {
                merged[pid] = 1;
                out->merge_ids[out->n_merge++] = pid;
            }
        }
    }

    /* 3. apply merges: parent at the children's centroid, mean velocity */
    for (int q = 0; q < out->n_merge; q++) {
        int pid = out->merge_ids[q];
        const float* c = &s->pos_h[8 * pid];
        const float* v = &s->vel_h[8 * pid];
        s->pos_l[2 * pid] = ((c[0] + c[2]) + (c[4] + c[6])) * 0.25f;
        s->pos_l[2 * pid + 1] = ((c[1] + c[3]) + (c[5] + c[7])) * 0.25f;
        s->vel_l[2 * pid] = ((v[0] + v[2]) + (v[4] + v[6])) * 0.25f;
        s->vel_l[2 * pid + 1] = ((v[1] + v[3]) + (v[5] + v[7])) * 0.25f;
        s->state_l[pid] = ORC_ACTIVE;
        s->age_l[pid] = 0;
        for (int k = 0; k < 4; k++) {
            s->state_h[4 * pid + k] = 1;
        }
    }

    /* 4. apply splits: children at x' +-(r_c, r_c), order (+,+),(-,+),(+,-),
     * (-,-), velocity v', transient (R:956-1005, 1848) */
    for (int q = 0; q < out->n_split; q++) {
        int pid = out->split_ids[q];
        float rc = child_radius(p->m_l, s->rho_l[pid]);
        float x = s->pos_l[2 * pid], y = s->pos_l[2 * pid + 1];
        float vx = s->vel_l[2 * pid], vy = s->vel_l[2 * pid + 1];
        for (int k = 0; k < 4; k++) {
            int a = 4 * pid + k;
            s->pos_h[2 * a] = (k & 1) ? x - rc : x + rc;
            s->pos_h[2 * a + 1] = (k & 2) ? y - rc : y + rc;
            s->vel_h[2 * a] = vx;
            s->vel_h[2 * a + 1] = vy;
            s->acc_h[2 * a] = 0.0f;
            s->acc_h[2 * a + 1] = 0.0f;
            s->rho_h[a] = 0.0f;
            s->p_h[a] = 0.0f;
            s->state_h[a] = 1;
        }
        s->state_l[pid] = ORC_TRANSIENT;
        s->age_l[pid] = 0;
    }

    /* 5. next lists (stable, order-preserving compaction) */
    int nl = 0;
    for (int k = 0; k < s->n_l; k++) {
        int i = g->sorted_l[k];
        if (s->state_l[i] != ORC_DELETED) {
            s->list_l[nl++] = i;
        }
    }
    for (int q = 0; q < s->n_g; q++) {
        int pid = s->list_g[q];
        if (merged[pid] == 1) {
            s->list_l[nl++] = pid;
        }
    }
    int nh = 0;
    for (int k = 0; k < s->n_h; k++) {
        int a = g->sorted_h[k];
        if (!merged[a / 4]) {
            s->list_h[nh++] = a;
        }
    }
    for (int q = 0; q < out->n_split; q++) {
        for (int k = 0; k < 4; k++) {
            s->list_h[nh++] = 4 * out->split_ids[q] + k;
        }
    }
    int ng = 0;
    for (int q = 0; q < s->n_g; q++) {
        int pid = s->list_g[q];
        if (!merged[pid]) {
            s->list_g[ng++] = pid;
        }
    }
    for (int q = 0; q < out->n_split; q++) {
        s->list_g[ng++] = out->split_ids[q];
    }
    s->n_l = nl;
    s->n_h = nh;
    s->n_g = ng;
    s->n_fresh = 4 * out->n_split;  /* appended last: the tail of list_h */
    free(merged);
}

/* ------------------------------------------------------------------------ */
/* Adaptive dt: dt = min(dt_max, cfl * l / (c + max|v|)), l = smoothing      */
/* length (support/2) of the finest level present; FP64, rounded to float.   */
/* ------------------------------------------------------------------------ */
double orc_v2max(const orc_state* s)
{
    double v2max = 0.0;
    for (int lv = 0; lv < 2; lv++) {
        int n = lv ? s->n_h : s->n_l;
        const int32_t* list = lv ? s->list_h : s->list_l;
        const float* vel = lv ? s->vel_h : s->vel_l;
        for (int k = 0; k < n; k++) {
            int i = list[k];
            double vx = vel[2 * i], vy = vel[2 * i + 1];
            double v2 = vx * vx + vy * vy;
            if (v2 > v2max) {
                v2max = v2;
            }
        }
    }
    return v2max;
}

float orc_next_dt(const orc_params* p, const orc_state* s, double v2max)
{
    if (p->dt_mode != ORC_DT_CFL) {
        return s->dt;
    }
    double ell = 0.5 * (double)(s->n_h > 0 ? p->H : p->h);
    double dt = (double)p->cfl * ell / ((double)p->c + sqrt(v2max));
    if (dt > (double)p->dt_max) {
        dt = (double)p->dt_max;
    }
    return (float)dt;
}

/* ------------------------------------------------------------------------ */
/* FP64 conservation sums (long double Kahan: an accurate reference).        */
/* ------------------------------------------------------------------------ */
void orc_conservation(const orc_params* p, const orc_state* s_, double* out,
                      double* abs_sum)
{
    const orc_state* s = s_;
    long double sum[5] = {0, 0, 0, 0, 0}, comp[5] = {0, 0, 0, 0, 0};
    long double asum[5] = {0, 0, 0, 0, 0};
    for (int lv = 0; lv < 2; lv++) {
        int n = lv ? s->n_h : s->n_l;
        const int32_t* list = lv ? s->list_h : s->list_l;
        const float* pos = lv ? s->pos_h : s->pos_l;
        const float* vel = lv ? s->vel_h : s->vel_l;
        long double m = lv ? p->m_h : p->m_l;
        for (int k = 0; k < n; k++) {
            int i = list[k];
            if (lv && s->state_h[i] != 0) {
                continue; /* transient children are not yet real */
            }
            long double vx = vel[2 * i], vy = vel[2 * i + 1];
            long double t[5];
            t[0] = m;
            t[1] = m * vx;
            t[2] = m * vy;
            t[3] = 0.5L * m * (vx * vx + vy * vy);
            t[4] = m * (long double)p->g * (long double)pos[2 * i + 1];
            for (int q = 0; q < 5; q++) {
                long double y = t[q] - comp[q];
                long double u = sum[q] + y;
                comp[q] = (u - sum[q]) - y;
                sum[q] = u;
                asum[q] += t[q] < 0 ? -t[q] : t[q];
            }
        }
    }
    for (int q = 0; q < 5; q++) {
        out[q] = (double)sum[q];
        abs_sum[q] = (double)asum[q];
    }
}

/* ------------------------------------------------------------------------ */
/* Full step (R's AdvanceTS, R:2726-2810, with the spec's rulings).          */
/* ------------------------------------------------------------------------ */
int orc_step(const orc_params* p, orc_state* s, int precision, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) {
        omp_set_num_threads(nthreads);
    }
#else
    (void)nthreads;
#endif
    const int cap_l = s->cap_l;
    const int cap_h = 4 * cap_l;
    const int ncl = p->dims_l[0] * p->dims_l[1];
    const int nch = p->dims_h[0] * p->dims_h[1];
    orc_grid g;
    g.sorted_l = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap_l + 1));
    g.keys_l = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap_l + 1));
    g.sorted_h = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap_h + 1));
    g.keys_h = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap_h + 1));
    g.start_l = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncl);
    g.end_l = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncl);
    g.start_h = (int32_t*)malloc(sizeof(int32_t) * (size_t)nch);
    g.end_h = (int32_t*)malloc(sizeof(int32_t) * (size_t)nch);
    g.bsorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->n_b + 1));
    g.bstart = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncl);
    g.bend = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncl);
    float* ppl = (float*)malloc(sizeof(float) * 2 * (size_t)cap_l);
    float* pph = (float*)malloc(sizeof(float) * 2 * (size_t)cap_h);
    float* acc_l = (float*)calloc(2 * (size_t)cap_l, sizeof(float));
    float* acc_h = (float*)calloc(2 * (size_t)cap_h, sizeof(float));
    orc_adapt_out ao;
    ao.split_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap_l);
    ao.promote_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap_l);
    ao.merge_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap_l);
    ao.dissolve_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap_l);
    ao.surf_l = (uint8_t*)malloc((size_t)cap_l);
    ao.surf_h = (uint8_t*)malloc((size_t)cap_h);

    orc_build_boundary_grid(p, s, &g);
    orc_build_grid(p, s, &g);
    if (p->relax_iters > 0 && s->n_fresh > 0) {
        if (precision == 64) {
            orc_relax_f64(p, s, &g);
        } else {
            orc_relax_f32(p, s, &g);
        }
    }
    if (precision == 64) {
        double* rl = (double*)malloc(sizeof(double) * (size_t)cap_l);
        double* pl = (double*)malloc(sizeof(double) * (size_t)cap_l);
        double* rh = (double*)malloc(sizeof(double) * (size_t)cap_h);
        double* ph = (double*)malloc(sizeof(double) * (size_t)cap_h);
        orc_density_f64(p, s, &g, rl, pl, rh, ph);
        for (int k = 0; k < s->n_l; k++) {
            int i = s->list_l[k];
            s->rho_l[i] = (float)rl[i];
            s->p_l[i] = (float)pl[i];
        }
        for (int k = 0; k < s->n_h; k++) {
            int a = s->list_h[k];
            s->rho_h[a] = (float)rh[a];
            s->p_h[a] = (float)ph[a];
        }
        free(rl);
        free(pl);
        free(rh);
        free(ph);
        double* al = (double*)calloc(2 * (size_t)cap_l, sizeof(double));
        double* ah = (double*)calloc(2 * (size_t)cap_h, sizeof(double));
        orc_force_f64(p, s, &g, al, ah, NULL, NULL);
        for (int k = 0; k < s->n_l; k++) {
            int i = s->list_l[k];
            acc_l[2 * i] = (float)al[2 * i];
            acc_l[2 * i + 1] = (float)al[2 * i + 1];
        }
        for (int k = 0; k < s->n_h; k++) {
            int a = s->list_h[k];
            acc_h[2 * a] = (float)ah[2 * a];
            acc_h[2 * a + 1] = (float)ah[2 * a + 1];
        }
        free(al);
        free(ah);
    } else {
        orc_density_f32(p, s, &g, s->rho_l, s->p_l, s->rho_h, s->p_h);
        orc_force_f32(p, s, &g, acc_l, acc_h, NULL, NULL);
    }
    memcpy(ppl, s->pos_l, sizeof(float) * 2 * (size_t)cap_l);
    memcpy(pph, s->pos_h, sizeof(float) * 2 * (size_t)cap_h);
    orc_integrate(p, s, acc_l, acc_h);
    double v2max = orc_v2max(s);
    orc_adapt(p, s, &g, ppl, pph, &ao);
    s->dt = orc_next_dt(p, s, v2max);
    s->step += 1;

    free(g.sorted_l); free(g.keys_l); free(g.sorted_h); free(g.keys_h);
    free(g.start_l); free(g.end_l); free(g.start_h); free(g.end_h);
    free(g.bsorted); free(g.bstart); free(g.bend);
    free(ppl); free(pph); free(acc_l); free(acc_h);
    free(ao.split_ids); free(ao.promote_ids); free(ao.merge_ids);
    free(ao.dissolve_ids);
    free(ao.surf_l); free(ao.surf_h);
    return 0;
}
