/* C ABI of the MI355X successive-orders-of-scattering engine (libsosfuse2_hip.so).
 *
 * Plain pointers and sizes only. Every pointer named *_dev is device memory owned by the caller (the Python host
 * allocates it through PyTorch); the library allocates nothing except where a function says so. Every call is
 * asynchronous on the HIP stream it is given (NULL = the default stream) and returns 0 or a negative SOS_E_* code;
 * sos_last_error() gives the message of the last failure on the calling thread.
 *
 * Reference interfaces these entry points replace (paths relative to the reference repository):
 *   sos_transport      Kernel.launch(bundle, s, field, first, phase, boundary) -> (moments, flux, query)
 *                      src/sosfuse2/solve/source_solver.py:2428-2456 (device transport of sos_fused,
 *                      src/sosfuse2/backends/kernel.cu:4138-4266), and the NumPy mirror
 *                      src/sosfuse2/backends/numpy_reference.py:461-643
 *   sos_project        the projection of kernel.cu fl_project :4323 / sh_project :4900 / sos_slim_proj :5702,
 *                      NumPy numpy_reference.py:683-711
 *   sos_compose        Solver.apply composition, src/sosfuse2/solve/source_iteration.py:176-212
 *   sos_make_plane     combined SH plane, src/sosfuse2/backends/sh_layout.py:135-152
 *   sos_source_metrics _SOURCE_METRICS kernel, src/sosfuse2/solve/source_iteration.py:52-97
 *   sos_anderson_*     Anderson residual, Gram products and mixing, source_iteration.py:432-443,618-635,25-33
 */
#ifndef SOSFUSE2_HIP_H
#define SOSFUSE2_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOS_OK 0
#define SOS_E_ARGUMENT (-1)
#define SOS_E_LAUNCH (-2)
#define SOS_E_UNSUPPORTED (-3)
#define SOS_E_DEVICE (-4)

/* Per-lane optical constants: float [SOS_LANE_CONSTANTS][lanes], row order below. */
#define SOS_LANE_EXT0 0 /* air (Rayleigh) extinction per km at unit profile */
#define SOS_LANE_EXT1 1 /* tropospheric aerosol */
#define SOS_LANE_EXT2 2 /* ozone */
#define SOS_LANE_EXT3 3 /* NO2 */
#define SOS_LANE_EXT4 4 /* O4 */
#define SOS_LANE_ASCA 5 /* aerosol scattering */
#define SOS_LANE_G 6    /* aerosol HG asymmetry */
#define SOS_LANE_MIX 7  /* Rayleigh depolarisation factor (1 - delta) / (1 + delta / 2) */
#define SOS_LANE_ALBEDO 8
#define SOS_LANE_SEXT 9  /* sulfate extinction */
#define SOS_LANE_SSCA 10 /* sulfate scattering */
#define SOS_LANE_SG 11   /* sulfate HG asymmetry */
#define SOS_LANE_CONSTANTS 12

/* Transport flags */
#define SOS_FIRST 1     /* first light (single scattering of the direct and ground-reflected sun) */
#define SOS_BOUNDARY 2  /* add the attenuated ground boundary term at ray ends that hit the ground */
#define SOS_RECORDS 8   /* read per-(node, corner) SH values from records_dev instead of reconstructing them */
#define SOS_UNIFORM_PHASE 16 /* first light: every lane of a ray has the same aerosol and sulfate asymmetry */

typedef struct sos_bundle {
    int64_t groups;               /* destination cells or query groups */
    int64_t rays;
    const int64_t *group_off_dev; /* [groups + 1] rays of each group */
    const double *starts_dev;     /* [groups][3] km */
    const float *targets_dev;     /* [groups][9] target frame, column i = basis vector i (ex, ey, up) */
    const int32_t *ray_group_dev; /* [rays] */
    const double *dirs_dev;       /* [rays][3] unit directions (global frame) */
    const float *dloc_dev;        /* [rays][3] directions in the target frame */
    const float *weights_dev;     /* [rays] solid-angle weights */
    const int64_t *edge_off_dev;  /* [rays + 1] */
    const double *edges_dev;      /* [edges] path coordinates (km) */
    const uint16_t *lower_dev;    /* [rays] bit i: segment i grows in radius */
    const int32_t *end_j_dev;     /* [rays] chi interval of the ray end */
    const float *end_f_dev;       /* [rays] fraction in that interval */
    const float *end_col_dev;     /* [rays][6] slant columns to the ground (5 components + sulfate) */
    const uint8_t *hits_dev;      /* [rays] 1 when the ray ends on the ground */
    const int64_t *node_off_dev;  /* [rays + 1] first quadrature node of every ray (records mode, from the plan) */
    const float *node_geom_dev;   /* [nodes][32] static node geometry (SOS_NODE_GEOMETRY, from sos_plan_geometry) */
    const float *node_sun_dev;    /* [nodes][12] sun geometry of first light (SOS_NODE_GEOMETRY with SOS_FIRST) */
} sos_bundle;

enum { SOS_PROFILE_SHAPES = 0, SOS_PROFILE_EXP = 1, SOS_PROFILE_GAUSS = 2, SOS_PROFILE_TRIANGLE = 3,
       SOS_PROFILE_TABLE = 4, SOS_PROFILE_AIR_SQUARED = 5, SOS_PROFILE_ZERO = 6 };

typedef struct sos_grid {
    int32_t nh, nc;            /* heights, solar zenith nodes; cells = nh * nc */
    int32_t L;                 /* SH order: 8, 16 or 32 */
    int32_t lanes;             /* lanes per ray: 1, 16 or 32 (>= active wavelengths) */
    int32_t nq;                /* Gauss-Legendre nodes per segment, 1..256 (the sun table takes 32 x its scale) */
    int32_t strato;            /* sulfate layer present */
    float top_km;
    float air_H, aer_H, o3_c, o3_s, no2_c, no2_s, st_c, st_s; /* profile shapes (km) */
    const double *gl_u_dev;    /* [nq] nodes on [0, 1] */
    const double *gl_w_dev;    /* [nq] weights */
    const float *h_dev;        /* [nh] */
    const float *chi_dev;      /* [nc] rad */
    const float *hr_dev;       /* [nh] air profile at the grid heights */
    const float *lane_dev;     /* [SOS_LANE_CONSTANTS][lanes] */
    int32_t sun_nh, sun_nx;
    const float *sun_h_dev;    /* [sun_nh] */
    const float *sun_col_dev;  /* [sun_nh][sun_nx][6] */
    const int64_t *goff_dev;   /* [cells + 1] ground pools per cell */
    const float *gdir_dev;     /* [pools][3] cell-frame directions */
    const float *gval_dev;     /* [pools][lanes] */
    const float *gm_dev;       /* [cells][4][lanes] Rayleigh tensor of the ground-reflected sun */
    int32_t blocks_per_cu;     /* transport launch: persistent workgroups per CU (0: 8); a tuning knob, results do
                                  not depend on it */
    int32_t sh_tile;           /* sos_sh_records FP32 workgroup form: 0 flow (double-buffered basis tiles, one
                                  barrier per round of 16 nodes), 1 stepwise (seeds, basis and products separated by
                                  barriers), 2 quad (the flow's rounds on twice the wavefronts, each holding half of
                                  the K range), 3 staggered quad (the two K halves form the next tile before and
                                  after their products); a tuning knob, the records are bitwise the same. The split-FP16
                                  records (mixed precision) at 32 lanes and L = 16, 24, 32 take the staggered form
                                  (8 whole-K wavefronts, tile formation of one under the products of the other on
                                  every SIMD) unless sh_tile is 1, which takes the quad form; other lane counts the
                                  quad form (the flow form at L = 8) */
    int32_t pool_max;          /* largest pool count of a cell (goff_dev differences); sos_ground_records holds a
                                  cell's pools in registers and takes at most 640 */
    /* Profile kinds of air, aerosol, O3, NO2, O4 and the sulfate layer (numpy_reference.py:187-223 evaluates the
     * reference's: exp(-h / 8), exp(-h / H), a triangle ozone, a tabulated NO2, air^2). SOS_PROFILE_SHAPES (0) takes
     * the analytic shape of the fields above for that component (exponential air and aerosol, Gaussian O3, NO2 and
     * sulfate, O4 as air^2); otherwise profile_a / profile_b are: SOS_PROFILE_EXP scale height (km) / -;
     * SOS_PROFILE_GAUSS sigma / centre; SOS_PROFILE_TRIANGLE half width / centre; SOS_PROFILE_TABLE step (km) / -,
     * values in row k of profile_table_dev [6][profile_table_n] from height 0; SOS_PROFILE_AIR_SQUARED the air
     * profile's surface value / - (O4 = (air / surface)^2); SOS_PROFILE_ZERO. */
    int32_t profile_kind[6];
    float profile_a[6], profile_b[6];
    const float *profile_table_dev;
    int32_t profile_table_n;
} sos_grid;

/* Radiance tensor of every ray: out_dev [rays][lanes][pairs] (pairs 4: xx yy zz xz; 6: + xy yz), FP32.
 * field_dev: source [cells][rows][lanes] (rows 0..9 read); plane_dev: the FP32 combined SH plane [cells][K][lanes]
 * of sos_make_plane (read without SOS_RECORDS); boundary_dev: [nc][lanes] ground radiance (with SOS_BOUNDARY).
 * counters_dev (may be NULL): uint64 [2] += quadrature nodes evaluated and SH corner refills, summed over rays.
 * records_dev (SOS_RECORDS): [nodes][4][lanes] SH values of every (node, corner) from sos_sh_records; the plane is
 * then not read. */
int sos_transport(const sos_grid *grid, const sos_bundle *bundle, int32_t flags, int32_t pairs, int32_t rows,
                  const float *field_dev, const void *plane_dev, const float *boundary_dev, const float *records_dev,
                  float *out_dev, uint64_t *counters_dev, void *stream);
/* first light with the ground term read from ground records [nodes][4][lanes] (sos_ground_records) */
#define SOS_GROUND_RECORDS 32
/* transport on the bundle's static node geometry (node_geom_dev; pairs 4): the warm records consumer (with
 * SOS_RECORDS), or first light (with SOS_FIRST and SOS_GROUND_RECORDS; also node_sun_dev) */
#define SOS_NODE_GEOMETRY 64
/* with SOS_NODE_GEOMETRY and SOS_RECORDS (warm): field_dev holds only the Rayleigh tensor rows, lane-major as
 * [cells][lanes][4] (rows 0..3 of each lane together, one 16-byte read per corner); rows is not read */
#define SOS_TENSOR4 128

/* Visit plan of a bundle (fixed by its geometry): counts_dev [rays] = quadrature nodes of every ray; then with
 * node_off_dev = exclusive prefix sum of the counts ([rays + 1]), every node's cell (ir * nc + j, its corner 0),
 * the ray's mu and the node's cos(phi). The caller sorts the visits (node * 4 + corner) by cell. */
int sos_plan_count(const sos_grid *grid, const sos_bundle *bundle, int64_t *counts_dev, void *stream);
int sos_plan_nodes(const sos_grid *grid, const sos_bundle *bundle, const int64_t *node_off_dev, int32_t *node_cell_dev,
                   float *node_mu_dev, float *node_cp_dev, float *node_dir_dev, void *stream);
/* Static geometry of every node for the warm consumer, geom_dev [nodes][32] floats: the Simpson panel ending at the
 * node and (last node of a segment) the segment's closing panel as weights of the six profiles (five components and
 * the sulfate layer), ds, the bilinear weights, the base cell with a last-of-segment bit, and the tensor rotation
 * (the per-node part of numpy_reference.py:294-351 and :236-291, computed once per bundle). */
int sos_plan_geometry(const sos_grid *grid, const sos_bundle *bundle, const int64_t *node_off_dev, float *geom_dev,
                      void *stream);
/* Sun geometry of every node for first light, sun_dev [nodes][12] floats: the six bilinear sun-table column
 * integrals at the node (FP64 lit test near the terminator, numpy_reference.py:431-458), the lit flag and the air,
 * aerosol and sulfate profiles there. */
int sos_plan_sun(const sos_grid *grid, const sos_bundle *bundle, const int64_t *node_off_dev, float *sun_dev,
                 void *stream);

/* First-light ground records: rec[id][w] = asca[w] p_aer(h) sum_p PHt gval[p][w] + ssca[w] p_sul(h) sum_p PHs gval[p][w]
 * over the ground pools of the visited cell (the pool sums of numpy_reference.py:526-558). The lanes fall into
 * `groups` phase groups (1 to 4), group i holding the lanes whose aerosol and sulfate asymmetries equal
 * group_g[i] and group_sg[i] (host arrays); the phase tiles are evaluated once per group. node_dir_dev:
 * [nodes][4] (ray direction in the node frame: mu, ex, |ey|; node height) from sos_plan_nodes. The records feed
 * sos_transport's first light through records_dev with SOS_GROUND_RECORDS. */
int sos_ground_records(const sos_grid *grid, int32_t groups, const float *group_g, const float *group_sg, int64_t items,
                       const int32_t *item_cell_dev, const int64_t *item_v0_dev, const int64_t *item_v1_dev,
                       const int32_t *vis_dev, const float *node_dir_dev, float *records_dev, void *stream);

/* Ground pools of the grid's lanes (the state-dependent part of the reference's pool preparation,
 * engine_core.py:103-147): for the pools of every cell (CSR off_dev [cells + 1] over the pools, cell = ir * nc + j),
 * gval_dev [pools][lanes] = albedo factor exp(-(column[0..4] . ext[0..4] + column[5] sulfate_ext)) from the lane
 * table, and gm_dev [cells][4][lanes] = 3/(8 pi) ext_air sum over the cell's pools of ((1 - x^2)/2, (1 - y^2)/2,
 * (1 - z^2)/2, -x z/2) gval with (x, y, z) = loc_dev [pools][3]; column_dev [pools][6], factor_dev [pools], loc_dev
 * in FP64, sums in FP64. The two outputs are the grid's gval_dev and gm_dev inputs of first light. */
int sos_ground_pools(const sos_grid *grid, const int64_t *off_dev, const double *column_dev, const double *factor_dev,
                     const double *loc_dev, float *gval_dev, float *gm_dev, void *stream);

/* Basis table of a bundle (fixed by its geometry): basis_dev [rays][KZ] (KZ = K rounded up to a multiple of 16)
 * holds fp32(x_m P_lm(mu)) in the m-major K order k' = m (L + 1) - m (m - 1) / 2 + l - m, zero for K <= k' < KZ,
 * with mu the ray direction's component along the sun axis (FP64) and x_m = sqrt 2 for m > 0: the normalised
 * associated-Legendre part of the SH basis (the FP64 recurrence of arithmetic.py:86-109), which is the same at every
 * node of a straight ray. sos_sh_records multiplies it by each node's cos(m phi). */
int sos_plan_basis(const sos_grid *grid, const sos_bundle *bundle, float *basis_dev, void *stream);

/* SH values of every (node, corner) as an MFMA contraction: records[4 n + c][w] = sum_k Z[n][k] plane[cell_c][k][w]
 * with Z[n][k] = basis[ray(n)][k] cos(m phi_n) the basis row of node n and cell_c its corners base, base + 1,
 * base + nc, base + nc + 1 (the replaced per-corner sum of numpy_reference.py:579-600). Work item i covers the nodes
 * node_dev[v0[i] .. v1[i]) whose base cell is item_cell[i]; node_dev holds node ids sorted by base cell and
 * node_ray_dev / node_cp_dev the ray and the node's cos(phi) in the same order; basis_dev is the bundle's table from
 * sos_plan_basis. half = 1: the plane is the split FP16 plane of sos_make_plane with its per-(cell, lane) scale_dev;
 * Z rows are split the same way and the products run as three v_mfma_f32_16x16x32_f16 (hi hi, hi lo, lo hi; the
 * K = 16 form v_mfma_f32_16x16x16f16 at L = 8) with FP32 accumulation (mixed precision, FP32-level accuracy). */
int sos_sh_records(const sos_grid *grid, const void *plane_dev, const float *scale_dev, int32_t half, int64_t items,
                   const int32_t *item_cell_dev, const int64_t *item_v0_dev, const int64_t *item_v1_dev,
                   const int32_t *node_dev, const int32_t *node_ray_dev, const float *node_cp_dev,
                   const float *basis_dev, float *records_dev, void *stream);

/* Moments [groups][4 + K][lanes] and ground flux [groups][lanes] of the per-ray radiances rad_dev
 * [rays][lanes][4] (the bundle's groups are the destination cells); the SH moments use the ray's basis row from
 * basis_dev (the bundle's table of sos_plan_basis) times cos(m phi) of the ray's azimuth. ground_only writes the flux
 * only (basis_dev may then be NULL). */
int sos_project(const sos_grid *grid, const sos_bundle *bundle, const float *rad_dev, int32_t ground_only,
                const float *basis_dev, float *moments_dev, float *flux_dev, void *stream);

/* source[cells][rows][lanes] = factor[rows][lanes] * (moments + first)[cells][row'][lanes] with row' the moment of
 * each row; values below FLT_MIN in magnitude become 0. first_dev may be NULL. */
int sos_compose(const sos_grid *grid, int32_t rows, const float *moments_dev, const float *first_dev,
                const float *factor_dev, float *source_dev, void *stream);

/* plane[cells][K][lanes] = source[:, 10:10+K] * ha[ih] + source[:, 10+K:10+2K] * sd[ih], K in m-major order.
 * half = 0: FP32 [cells][K][lanes], scale_dev unused. half = 1: split FP16, plane_dev [2][cells][KZ/4][lanes][4]
 * halves (KZ = K rounded up to 16, zeros beyond K; four consecutive K positions of a lane together) holding
 * hi = fp16(x) and lo = fp16((x - hi) 2^11) of x = c / scale, with scale_dev [cells][lanes] float the power of two
 * just above the column's largest magnitude (1 for an all-zero column). */
int sos_make_plane(const sos_grid *grid, int32_t rows, const float *source_dev, const float *ha_dev,
                   const float *sd_dev, int32_t half, void *plane_dev, float *scale_dev, void *stream);

/* Per lane: count of non-finite g, max |g|, max |g - x|; frozen lanes read x for g. out_dev [3][lanes] float
 * (count as float). */
int sos_source_metrics(int64_t rows, int32_t lanes, const float *g_dev, const float *x_dev, uint32_t frozen,
                       float *out_dev, void *stream);

/* Anderson residual of every lane (source_iteration.py:432-443):
 * f[lane][n], n = cells*rows + nobs, f = (g - x) / max(||g_cell||, 1e-8 max_cells ||g_cell||) over the source rows,
 * then obs[lane][nobs] (the scaled observer change). work_dev: float [cells*lanes + lanes] scratch. */
int sos_anderson_residual(int64_t cells, int32_t rows, int32_t lanes, const float *g_dev, const float *x_dev,
                          const float *obs_dev, int32_t nobs, float *work_dev, float *f_dev, void *stream);

/* Row and column `slot` of the per-lane Gram matrices gram_dev [lanes][depth][depth] (FP64):
 * gram[lane][slot][j] = gram[lane][j][slot] = sum_i f_ring[slot][lane][i] f_ring[j][lane][i] for every j with bit j
 * of filled set. f_ring: [depth][lanes][n]. work_dev: double [128 * lanes * depth] scratch (or more). Fixed-order sums. */
int sos_anderson_gram(int64_t n, int32_t lanes, int32_t depth, int32_t slot, uint32_t filled, const float *ring_dev,
                      double *work_dev, double *gram_dev, void *stream);

/* sos_anderson_gram without the work buffer; no counterpart in the reference. For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): it calls no allocator. The partial sums go to
 * a file-scope device array of 128 x 32 x 8 doubles (256 KiB per device, held until the process ends), enough for
 * every argument the limits allow (lanes <= 32, depth <= 8); a larger need is rejected with an error. Every call on a
 * device shares that device's array, so calls that overlap on one device race. Results equal sos_anderson_gram's
 * bit for bit (the same kernels and summation order). */
int sos_anderson_gram_static(int64_t n, int32_t lanes, int32_t depth, int32_t slot, uint32_t filled,
                             const float *ring_dev, double *gram_dev, void *stream);

/* Per lane, FP64 on the device: the regularised Gram system of the m slots order_dev[0..m-1] (oldest first),
 * G += max(tr G 1e-10, 1e-30) I, w = G^-1 1 / sum, accepted when finite and max |w| < 8 (source_iteration.py:25-33).
 * w_dev [lanes][depth] float; w[lane][0] = NaN marks a lane that keeps g. Lanes with bit set in frozen are skipped. */
int sos_anderson_solve(int32_t lanes, int32_t depth, int32_t m, const int32_t *order_dev, const double *gram_dev,
                       uint32_t frozen, float *w_dev, void *stream);

/* x[cells][rows][lanes] = sum_j w[lane][j] g_ring[order_j][cells][rows][lanes] for lanes with accepted weights;
 * other lanes are left as they are. g_ring: [depth][cells][rows][lanes]. */
int sos_anderson_mix(int64_t cells, int32_t rows, int32_t lanes, int32_t depth, int32_t m, const float *ring_dev,
                     const int32_t *order_dev, const float *w_dev, float *x_dev, void *stream);

/* Learned warm start: no counterpart in the reference (the network only seeds Solver.solve's iteration, which still
 * runs to its tolerance). All arrays FP32 in device memory; products round their operands to BF16 and accumulate
 * in FP32.
 * sos_mlp_gemm: Z[i][j] (row stride ldz) = sum_k A(i, k) B(k, j) (+ bias[j]) with A(i, k) = a[i sa_i + k sa_k],
 * B(k, j) = b[k sb_k + j sb_j]; act 1 also writes H = gelu(Z) (exact erf form). splits > 1 divides K over
 * workgroups (work_dev: splits * M * N floats) and sums the parts in a fixed order. */
int sos_mlp_gemm(int64_t M, int64_t N, int64_t K, const float *a_dev, int64_t sa_i, int64_t sa_k, const float *b_dev,
                 int64_t sb_k, int64_t sb_j, const float *bias_dev, int32_t act, float *z_dev, int64_t ldz,
                 float *h_dev, int64_t ldh, int32_t splits, float *work_dev, void *stream);
/* dz = dh * gelu'(z), elementwise over rows x cols with row strides ldd, ldz, ldo (dz may be dh). */
int sos_mlp_gelu_backward(int64_t rows, int64_t cols, const float *dh_dev, int64_t ldd, const float *z_dev,
                          int64_t ldz, float *dz_dev, int64_t ldo, void *stream);
/* loss_dev[0] = mean over rows x cols of w[j] (y - t)^2 (w_dev may be NULL = 1); dy_dev (may be NULL) = its
 * gradient 2 w (y - t) / (rows cols). work_dev: 1024 doubles. */
int sos_mlp_mse(int64_t rows, int64_t cols, const float *y_dev, int64_t ldy, const float *t_dev, int64_t ldt,
                const float *w_dev, float *dy_dev, int64_t lddy, float *loss_dev, double *work_dev, void *stream);
/* db[j] = sum_i dz[i][j] (fixed order). */
int sos_mlp_bias_grad(int64_t rows, int64_t cols, const float *dz_dev, int64_t ld, float *db_dev, void *stream);
/* One Adam step (bias-corrected, step counts from 1) on n parameters in place. */
int sos_mlp_adam(int64_t n, float *p_dev, const float *g_dev, float *m_dev, float *v_dev, int32_t step, float lr,
                 float beta1, float beta2, float eps, void *stream);

/* Column integrals on the device (FP64): the slant columns of the five component profiles and the sulfate profile
 * along rays, Gauss-Legendre nodes on every segment between the ray's boundaries (shell exit or ground hit, tangent
 * point, shadow-cylinder crossings, interior shells at 10, 25, 40 km and 80 km when the top is higher), graded
 * quadratically toward each segment's lower end (rays/shell_segments.py:43-66,115-147). Profile kinds as in sos_grid
 * (SOS_PROFILE_SHAPES is not accepted: give every kind), parameters in FP64, the table in FP64. */
typedef struct sos_columns {
    double top_km;
    int32_t nq;
    const double *gl_u_dev;         /* [nq] nodes on [0, 1] */
    const double *gl_w_dev;         /* [nq] weights */
    int32_t profile_kind[6];
    double profile_a[6], profile_b[6];
    const double *profile_table_dev; /* [6][profile_table_n] */
    int32_t profile_table_n;
} sos_columns;

/* Sun-column table (replaces rays/shell_segments.py:267-307, SunGeometry's table): out_dev [nh][nx][6] = columns from
 * (0, 0, R + h[i]) toward mu = mg + (1 - mg) x[j]^2 (mg the horizon's cosine there; mg + 1e-9 at j = 0). */
int sos_sun_table(const sos_columns *c, int32_t nh, int32_t nx, const double *h_dev, const double *x_dev,
                  double *out_dev, void *stream);

/* Columns [n][6] from each ray's start (starts_dev[group_dev[i]], or starts_dev[i] without groups) along dirs_dev[i]
 * to its end, for the rays with hit_dev[i] set (all rays without hit_dev; the others 0): the end columns of the
 * ground-hitting rays (replaces solve/ray_geometry.py:53-65 and rays/ray_columns.py) and the path columns of the
 * ground pools (solve/engine_core.py:103-147). */
int sos_ray_columns(const sos_columns *c, int64_t n, const double *starts_dev, const int32_t *group_dev,
                    const double *dirs_dev, const uint8_t *hit_dev, double *out_dev, void *stream);

/* Incident-direction quadrature of every cell of a grid (replaces rays/incoming_rings.py:63-126): per azimuth j of
 * nphi + 1 over [0, pi] (trapezoid, the y-mirror folded into the weights), nmu Gauss-Legendre nodes on every
 * sinh-graded mu interval between the horizon, shadow, tangent-pair and ring edges; weights sum to 4 pi per cell. */
typedef struct sos_incident {
    int32_t nh, nc, nmu, nphi, nlevels;
    const double *h_dev;            /* [nh] heights (km) */
    const double *chi_dev;          /* [nc] solar zenith nodes (rad) */
    double levels[8];               /* ring levels -sqrt((R + h)^2 - R^2) of the ring heights */
    const double *gl_u_dev;         /* [nmu] nodes on [0, 1] */
    const double *gl_w_dev;         /* [nmu] weights */
} sos_incident;

/* counts_dev [nh nc][nphi + 1]: the node count of every (cell, azimuth). */
int sos_incident_count(const sos_incident *a, int64_t *counts_dev, void *stream);
/* From the exclusive prefix sums off_dev of those counts: every node's direction (global frame, FP64 [rays][3]),
 * solid-angle weight (FP64) and cell, in cell, azimuth, mu order. */
int sos_incident_fill(const sos_incident *a, const int64_t *off_dev, double *dirs_dev, double *weights_dev,
                      int32_t *ray_cell_dev, void *stream);

/* Ray packs of a bundle (replaces solve/ray_geometry.py:28-82 pack and :148-180 _csr_vector): per ray from its cell's
 * start along dirs_dev, the sorted boundaries (sos_columns' top), kept intervals of strictly positive width.
 * sos_pack_count writes the compacted edge count of every ray (kept intervals + 1, or 1 with a single 0) and its
 * intervals wider than 1e-8 km (the sampled ones); sos_pack_fill writes, from the prefix sums edge_off_dev, the
 * compacted edges, the lower-end flag bits, the end's chi index and fraction, the ground hit and the direction in the
 * cell's target frame (targets_dev [cells][3][3], frame vectors as columns). */
int sos_pack_count(const sos_columns *c, int32_t nc, const double *chi_dev, const double *starts_dev, int64_t rays,
                   const int32_t *ray_cell_dev, const double *dirs_dev, int64_t *edge_count_dev, int64_t *sampled_dev,
                   void *stream);
int sos_pack_fill(const sos_columns *c, int32_t nc, const double *chi_dev, const double *starts_dev,
                  const double *targets_dev, int64_t rays, const int32_t *ray_cell_dev, const double *dirs_dev,
                  const int64_t *edge_off_dev, double *edges_dev, int16_t *lower_dev, int32_t *end_j_dev,
                  float *end_f_dev, uint8_t *hits_dev, float *dloc_dev, void *stream);

const char *sos_last_error(void);
/* Compute units, XCDs (hipDeviceAttributeNumberOfXccs) and the gfx architecture name of the current device. */
int sos_device_info(int32_t *cus, int32_t *xcds, char *arch, int32_t arch_len);

#ifdef __cplusplus
}
#endif
#endif
