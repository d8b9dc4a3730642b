/*
 * tss_oracle.h -- CPU restatement of the two-scale WCSPH step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (twoscalesplit_amd) never links or calls it.
 *
 * It restates, from the reference's source read as text:
 *   R = TwoScale2D/WCSPHSolver.cu (segfault11/TwoScaleSplit)
 *   - grid coordinates / hash           R:241-323
 *   - hash, stable key sort, ranges     R:1246-1340, 2814-2854, 2934-2979
 *   - M4 kernel, gradient, wall weight  R:66-239
 *   - density + Tait                    R:335-474, 1342-1516
 *   - pressure/viscosity/tension/wall   R:476-885, 1648-1865, 1867-2033
 *   - split 1->4, child clamp           R:956-1022, 1834-1863, 1097-1172
 *   - step order                        R:2726-2810
 * with the defect rulings of the method spec (high<-low density reads the
 * low arrays; accL zero-initialised; stable order-preserving compaction
 * instead of atomicAdd order; all forces read the pre-step state).
 *
 * Two precisions of the physics are built from one source
 * (tss_oracle_phys.inc): *_f64 (the FP64 method oracle, accuracy reference)
 * and *_f32 (a transcription of the reference's FP32 kernels with its naive
 * "+=" sums; used as the timed CPU baseline). Everything that decides an
 * index or a set -- cell coordinates, hashes, sort order, cell ranges,
 * neighbour predicates, surface/band/split/promote/merge decisions, the
 * symplectic-Euler update and the child placement -- is evaluated in FP32
 * with IEEE division/sqrt and no contraction (compiled -ffp-contract=off),
 * identically in both builds: that is the bit-exact reference.
 */
#ifndef TSS_ORACLE_H
#define TSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Low-particle states (R: mdStates, R/WCSPHSolver.h:105-123). */
#define ORC_ACTIVE    0u   /* ordinary low particle                       */
#define ORC_TRANSIENT 1u   /* split; children relaxing; parent still live */
#define ORC_DELETED   2u   /* replaced by its promoted children           */
/* High (child) states: 1 = transient child, 0 = promoted (active) child. */

/* Adaptivity modes. */
#define ORC_MODE_PLAIN     0  /* single level, no high particles            */
#define ORC_MODE_REFERENCE 1  /* reference rule: split iff x'.x > split_x,
                                 children stay transient (R:1834-1863)      */
#define ORC_MODE_BAND      2  /* split inside the free-surface band, promote
                                 relaxed groups, merge groups outside it    */

#define ORC_DT_FIXED 0
#define ORC_DT_CFL   1

typedef struct orc_params {
    float origin[2], end[2];
    int   dims_l[2], dims_h[2];   /* ceil((end-origin)/h) in FP32 (R:2323-2331) */
    float h, H;                   /* support radii: EffectiveRadius(High)       */
    float rho0, B, c, alpha, sigma, g;
    float m_l, m_h, m_b;
    int   mode;
    float split_x;                /* reference rule threshold (R:1835: 0.5)    */
    float surf_ratio;             /* low surface iff rho < surf_ratio*rho0     */
    float surf_ratio_h;           /* high surface iff rho < surf_ratio_h*rho0  */
    float r_band, r_merge;        /* split band / merge distance               */
    int   promote_min_age, promote_max_age;  /* max_age <= 0: no promotion */
    float promote_fluc;
    int   cap_h;                  /* live high-particle cap (<= 4*cap_l)        */
    int   dt_mode;
    float cfl, dt_max;
    int   relax_iters;            /* relaxation of new children (0: off)        */
    float relax_frac;             /* relaxation step = relax_frac * dt          */
} orc_params;

/* Scene state, by particle id. High id of child k of low parent p is 4p+k.
 * All arrays are owned by the caller. */
typedef struct orc_state {
    int cap_l, n_b;
    float *pos_l, *vel_l, *acc_l;      /* 2*cap_l */
    float *rho_l, *p_l;                /* cap_l   */
    uint8_t *state_l;                  /* cap_l   */
    int32_t *age_l;                    /* cap_l   */
    float *pos_h, *vel_h, *acc_h;      /* 8*cap_l */
    float *rho_h, *p_h;                /* 4*cap_l */
    uint8_t *state_h;                  /* 4*cap_l */
    float *pos_b;                      /* 2*n_b   */
    int n_l, n_h, n_g;
    int32_t *list_l, *list_h, *list_g; /* cap_l, 4*cap_l, cap_l */
    float dt;
    int step;
    int n_fresh;                       /* children injected by the last step:
                                          the last n_fresh of list_h       */
} orc_state;

/* Neighbour grid of one step (outputs of the grid phase). */
typedef struct orc_grid {
    int32_t *sorted_l, *sorted_h;      /* n_l, n_h: ids in stable hash order */
    int32_t *keys_l, *keys_h;          /* n_l, n_h: sorted hashes            */
    int32_t *start_l, *end_l;          /* dims_l[0]*dims_l[1]                */
    int32_t *start_h, *end_h;          /* dims_h[0]*dims_h[1]                */
    int32_t *bsorted, *bstart, *bend;  /* boundary: n_b ids; low-grid cells  */
} orc_grid;

/* Per-step outputs of the adaptivity phase. */
typedef struct orc_adapt_out {
    int n_split, n_promote, n_merge, n_dissolve;
    int32_t *split_ids, *promote_ids, *merge_ids, *dissolve_ids; /* each cap_l */
    uint8_t *surf_l, *surf_h;                       /* by id (cap_l, 4*cap_l)    */
} orc_adapt_out;

/* ---- parameters -------------------------------------------------------- */
/* Mirrors WCSPHConfig's constructor (R/WCSPHSolver.cu:2287-2332); returns
 * -1 on an invalid domain like the reference's exception. */
int  orc_params_init(orc_params* p, float xs, float ys, float xe, float ye,
                     float h, float H, float rho0, float B, float c,
                     float alpha, float sigma);

/* ---- grid / indices (FP32, bit-exact reference) ------------------------ */
int  orc_axis_cell(float x, float off, float origin, float cell, int dim);
int  orc_hash(const orc_params* p, int level, float x, float y);
void orc_hash_list(const orc_params* p, int level, const float* pos,
                   const int32_t* ids, int n, int32_t* keys);
void orc_stable_sort(const int32_t* keys, const int32_t* vals, int n,
                     int nkeys, int32_t* keys_out, int32_t* vals_out);
void orc_cell_ranges(const int32_t* sorted_keys, int n, int ncells,
                     int32_t* start, int32_t* end);
void orc_build_grid(const orc_params* p, const orc_state* s, orc_grid* g);
void orc_build_boundary_grid(const orc_params* p, const orc_state* s,
                             orc_grid* g);

/* Neighbour sets: per target (in list order), for 5 categories, the count
 * and an order-independent checksum sum((id+1)*2654435761) mod 2^32.
 * low targets : 0 LL dens (r<=h), 1 LH dens (r<=h), 2 LL force (0<r<h),
 *               3 LH force (0<r<h), 4 boundary candidates in the box
 * high targets: 0 HH dens (r<=H), 1 HL dens (r<=H), 2 HH force (0<r<H),
 *               3 HL force (0<r<h), 4 boundary candidates in the box     */
void orc_neighbors(const orc_params* p, const orc_state* s,
                   const orc_grid* g, int level, uint32_t* counts,
                   uint32_t* sums);
/* The same five categories restricted to the pairs whose state weight is
 * nonzero -- the terms the density and force sums add: lambda_LH on low <->
 * high pairs (R:377, 602), lambda_HH (R:414, 681), lambda_HL (R:451, 808)
 * and the deleted weight d(s_j) on forces from low particles (R:500, 809). */
void orc_pair_sets(const orc_params* p, const orc_state* s,
                   const orc_grid* g, int level, uint32_t* counts,
                   uint32_t* sums);

/* ---- physics ----------------------------------------------------------- */
/* Density + Tait pressure for every listed particle, by id. */
void orc_density_f64(const orc_params* p, const orc_state* s,
                     const orc_grid* g, double* rho_l, double* p_l,
                     double* rho_h, double* p_h);
void orc_density_f32(const orc_params* p, const orc_state* s,
                     const orc_grid* g, float* rho_l, float* p_l,
                     float* rho_h, float* p_h);
/* Accelerations from the state's FP32 rho/p (the force kernel's inputs),
 * by id (2 per particle). scale_* (may be NULL) receives, per particle,
 * sum over pair/wall terms of their magnitudes plus g: the conditioning
 * scale of the cancelling sum. */
void orc_force_f64(const orc_params* p, const orc_state* s,
                   const orc_grid* g, double* acc_l, double* acc_h,
                   double* scale_l, double* scale_h);
void orc_force_f32(const orc_params* p, const orc_state* s,
                   const orc_grid* g, float* acc_l, float* acc_h,
                   double* scale_l, double* scale_h); /* scales unused */

/* ---- update ------------------------------------------------------------ */
/* Symplectic Euler (R:1635-1639, 2014-2018) on the listed particles with
 * the given FP32 accelerations; writes s->acc_* too. */
void orc_integrate(const orc_params* p, orc_state* s, const float* acc_l,
                   const float* acc_h);
/* Surface band split / promotion / merge / child clamp and the next lists.
 * pre_pos_* are the pre-step positions the grid was built from; s holds the
 * post-integration positions/velocities and this step's FP32 rho. */
void orc_adapt(const orc_params* p, orc_state* s, const orc_grid* g,
               const float* pre_pos_l, const float* pre_pos_h,
               orc_adapt_out* out);
/* max |v|^2 (FP64) over the listed particles: evaluated after integration,
 * before the adaptivity phase changes the lists. */
double orc_v2max(const orc_state* s);
/* Next dt (CFL mode): min(dt_max, cfl*l/(c + sqrt(v2max))) with l the
 * smoothing length (support/2) of the finest level in the next lists. */
float orc_next_dt(const orc_params* p, const orc_state* s, double v2max);

/* FP64 kernel functions the oracle sums use (for closed-form pins). */
double orc_kernel_w(double r, double ell);      /* M4 W, ell = support */
double orc_kernel_gradf(double r, double ell);  /* dW/dr               */
double orc_wall_gamma(double r, double ell, double c);
double orc_tait(double rho, double rho0, double B);

/* FP64 conservation sums over real particles (listed low, promoted high):
 * out[0] mass, [1] px, [2] py, [3] kinetic, [4] potential (m g y);
 * err[i] = sum of |terms| (multiply by a summation bound to get an error
 * bound). Summed with long double + Kahan. */
void orc_conservation(const orc_params* p, const orc_state* s, double* out,
                      double* abs_sum);

/* Relaxation of the s->n_fresh children at the end of list_h on the grid g
 * (p->relax_iters Jacobi iterations of density, pressure + wall
 * acceleration and symplectic Euler at relax_frac * dt; see orc_step). */
void orc_relax_f64(const orc_params* p, orc_state* s, const orc_grid* g);
void orc_relax_f32(const orc_params* p, orc_state* s, const orc_grid* g);

/* One full step: grid, [relaxation], density, force, integrate, adapt, dt.
 * precision 64 or 32 selects the physics build. nthreads <= 0 keeps
 * OpenMP's default. */
int  orc_step(const orc_params* p, orc_state* s, int precision, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
