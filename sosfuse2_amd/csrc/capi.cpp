// C ABI of libsosfuse2_hip.so: argument checks, error strings, and dispatch to the kernel launchers.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace sos {

static thread_local char g_error[512] = "";

int set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return SOS_E_LAUNCH;
}

static int arg_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return SOS_E_ARGUMENT;
}

int transport_dispatch(const TransportArgs &a, int G, int L, hipStream_t s);
int project_dispatch(const ProjectArgs &a, int G, int L, hipStream_t s);
int plan_count_launch(const TransportArgs &a, int64_t *counts, hipStream_t s);
int plan_geometry_launch(const TransportArgs &a, const int64_t *node_off, float *geom, hipStream_t s);
int plan_sun_launch(const TransportArgs &a, const int64_t *node_off, float *sun, hipStream_t s);
int plan_nodes_launch(const TransportArgs &a, const int64_t *node_off, int32_t *cell, float *mu, float *cp,
                      float *dir, hipStream_t s);
int ground_records_dispatch(const GroundArgs &a, hipStream_t s);
int pool_values_launch(int cells, int G, const int64_t *off, const double *column, const double *factor,
                       const double *loc, const float *lane, float *gval, float *gm, hipStream_t s);

int sh_records_dispatch(const ShArgs &a, bool half, hipStream_t s);
int plan_basis_launch(int L, int64_t rays, const double *dirs, float *basis, hipStream_t s);
int compose_launch(int64_t cells, int rows, int K, int G, const float *mom, const float *first, const float *factor,
                   float *out, hipStream_t s);
int plane_launch(int nh, int nc, int L, int rows, int G, int strato, const float *src, const float *ha,
                 const float *sd, int half, void *plane, float *scale, hipStream_t s);
int metrics_launch(int64_t rows, int G, const float *g, const float *x, unsigned frozen, float *out, hipStream_t s);
int residual_launch(int64_t cells, int rows, int G, const float *g, const float *x, const float *obs, int nobs,
                    float *work, float *f, hipStream_t s);
int gram_launch(int64_t n, int G, int depth, int slot, unsigned filled, const float *ring, double *work, double *gram,
                hipStream_t s);
int gram_static_launch(int64_t n, int G, int depth, int slot, unsigned filled, const float *ring, double *gram,
                       hipStream_t s);
int solve_launch(int G, int depth, int m, const int *order, const double *gram, unsigned frozen, float *w,
                 hipStream_t s);
int mix_launch(int64_t cells, int rows, int G, int depth, int m, const float *ring, const int *order, const float *w,
               float *x, hipStream_t s);

static bool lanes_ok(int G) { return G == 1 || G == 16 || G == 32; }
static bool order_ok(int L) { return L == 8 || L == 16 || L == 24 || L == 32; }

int mlp_gemm_launch(const GemmArgs &in, float *work, hipStream_t s);
int mlp_gelu_backward_launch(int64_t rows, int64_t cols, const float *dh, int64_t ldd, const float *z, int64_t ldz,
                             float *dz, int64_t ldo, hipStream_t s);
int mlp_mse_launch(int64_t rows, int64_t cols, const float *y, int64_t ldy, const float *t, int64_t ldt,
                   const float *w, float *dy, int64_t lddy, float *loss, double *work, hipStream_t s);
int mlp_bias_grad_launch(int64_t rows, int64_t cols, const float *dz, int64_t ld, float *db, hipStream_t s);
int mlp_adam_launch(int64_t n, float *p, const float *g, float *m, float *v, int step, float lr, float b1, float b2,
                    float eps, hipStream_t s);
int sun_table_launch(int nh, int nx, const double *h, const double *x, const ColumnArgs &c, double *out, hipStream_t s);
int incident_launch(const IncidentArgs &a, int64_t *counts, const int64_t *off, double *dirs, double *weights,
                    int32_t *ray_cell, hipStream_t s);
int pack_launch(const PackArgs &a, int64_t rays, const int32_t *ray_cell, const double *dirs, int64_t *edge_count,
                int64_t *sampled, const int64_t *edge_off, double *edges, int16_t *lower, int32_t *end_j, float *end_f,
                uint8_t *hits, float *dloc, hipStream_t s);
int ray_columns_launch(int64_t n, const double *starts, const int32_t *group, const double *dirs, const uint8_t *hit,
                       const ColumnArgs &c, double *out, hipStream_t s);

}  // namespace sos

using namespace sos;

// Column arguments from the C struct: interior shells by the top, profile kinds checked (every kind explicit).
static int column_args(const sos_columns *c, ColumnArgs &a, const char *who) {
    if (!c || !c->gl_u_dev || !c->gl_w_dev || c->nq < 1 || c->nq > 256 || !(c->top_km > 0))
        return arg_error("%s: bad column arguments", who);
    a.top = c->top_km;
    a.shells = c->top_km > 80.0 ? 4 : 3;
    const double sh[4] = {10.0, 25.0, 40.0, 80.0};
    for (int i = 0; i < 4; ++i) a.shell_h[i] = sh[i];
    a.nq = c->nq; a.gl_u = c->gl_u_dev; a.gl_w = c->gl_w_dev;
    for (int k = 0; k < 6; ++k) {
        const int kind = c->profile_kind[k];
        if (kind < SOS_PROFILE_EXP || kind > SOS_PROFILE_ZERO) return arg_error("%s: profile kind %d of %d", who, kind, k);
        if (kind == SOS_PROFILE_TABLE && (!c->profile_table_dev || c->profile_table_n < 2))
            return arg_error("%s: profile table missing", who);
        if (kind != SOS_PROFILE_ZERO && kind != SOS_PROFILE_AIR_SQUARED && !(c->profile_a[k] > 0))
            return arg_error("%s: profile %d scale %g", who, k, c->profile_a[k]);
        a.shape.kind[k] = kind;  // the device codes of geometry.hip follow the SOS_PROFILE_* values
        a.shape.a[k] = c->profile_a[k];
        a.shape.b[k] = c->profile_b[k];
    }
    a.shape.table = c->profile_table_dev;
    a.shape.table_n = c->profile_table_n;
    return SOS_OK;
}


extern "C" {

const char *sos_last_error(void) { return g_error; }

int sos_sun_table(const sos_columns *c, int32_t nh, int32_t nx, const double *h_dev, const double *x_dev,
                  double *out_dev, void *stream) {
    ColumnArgs a;
    if (int rc = column_args(c, a, "sos_sun_table")) return rc;
    if (nh < 2 || nx < 2 || !h_dev || !x_dev || !out_dev) return arg_error("sos_sun_table: bad table arguments");
    return sun_table_launch(nh, nx, h_dev, x_dev, a, out_dev, (hipStream_t)stream);
}

static int incident_args(const sos_incident *a, IncidentArgs &o, const char *who) {
    if (!a || a->nh < 1 || a->nc < 2 || a->nmu < 1 || a->nphi < 1 || a->nlevels < 0 || a->nlevels > 8 || !a->h_dev ||
        !a->chi_dev || !a->gl_u_dev || !a->gl_w_dev)
        return arg_error("%s: bad quadrature arguments", who);
    o.nh = a->nh; o.nc = a->nc; o.nmu = a->nmu; o.nphi = a->nphi; o.nlevels = a->nlevels;
    o.h = a->h_dev; o.chi = a->chi_dev; o.gl_u = a->gl_u_dev; o.gl_w = a->gl_w_dev;
    for (int i = 0; i < 8; ++i) o.levels[i] = a->levels[i];
    return SOS_OK;
}

int sos_incident_count(const sos_incident *a, int64_t *counts_dev, void *stream) {
    IncidentArgs o;
    if (int rc = incident_args(a, o, "sos_incident_count")) return rc;
    if (!counts_dev) return arg_error("sos_incident_count: null argument");
    return incident_launch(o, counts_dev, nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int sos_incident_fill(const sos_incident *a, const int64_t *off_dev, double *dirs_dev, double *weights_dev,
                      int32_t *ray_cell_dev, void *stream) {
    IncidentArgs o;
    if (int rc = incident_args(a, o, "sos_incident_fill")) return rc;
    if (!off_dev || !dirs_dev || !weights_dev || !ray_cell_dev) return arg_error("sos_incident_fill: null argument");
    return incident_launch(o, nullptr, off_dev, dirs_dev, weights_dev, ray_cell_dev, (hipStream_t)stream);
}

static int pack_args(const sos_columns *c, int32_t nc, const double *chi_dev, const double *starts_dev,
                     const double *targets_dev, PackArgs &o, const char *who) {
    if (int rc = column_args(c, o.c, who)) return rc;
    if (nc < 2 || !chi_dev || !starts_dev) return arg_error("%s: bad grid arguments", who);
    o.nc = nc; o.chi = chi_dev; o.starts = starts_dev; o.targets = targets_dev;
    return SOS_OK;
}

int sos_pack_count(const sos_columns *c, int32_t nc, const double *chi_dev, const double *starts_dev, int64_t rays,
                   const int32_t *ray_cell_dev, const double *dirs_dev, int64_t *edge_count_dev, int64_t *sampled_dev,
                   void *stream) {
    PackArgs o;
    if (int rc = pack_args(c, nc, chi_dev, starts_dev, nullptr, o, "sos_pack_count")) return rc;
    if (rays < 0 || (rays > 0 && (!ray_cell_dev || !dirs_dev || !edge_count_dev || !sampled_dev)))
        return arg_error("sos_pack_count: null argument");
    return pack_launch(o, rays, ray_cell_dev, dirs_dev, edge_count_dev, sampled_dev, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int sos_pack_fill(const sos_columns *c, int32_t nc, const double *chi_dev, const double *starts_dev,
                  const double *targets_dev, int64_t rays, const int32_t *ray_cell_dev, const double *dirs_dev,
                  const int64_t *edge_off_dev, double *edges_dev, int16_t *lower_dev, int32_t *end_j_dev,
                  float *end_f_dev, uint8_t *hits_dev, float *dloc_dev, void *stream) {
    PackArgs o;
    if (int rc = pack_args(c, nc, chi_dev, starts_dev, targets_dev, o, "sos_pack_fill")) return rc;
    if (rays < 0 || (rays > 0 && (!targets_dev || !ray_cell_dev || !dirs_dev || !edge_off_dev || !edges_dev ||
                                  !lower_dev || !end_j_dev || !end_f_dev || !hits_dev || !dloc_dev)))
        return arg_error("sos_pack_fill: null argument");
    return pack_launch(o, rays, ray_cell_dev, dirs_dev, nullptr, nullptr, edge_off_dev, edges_dev, lower_dev, end_j_dev,
                       end_f_dev, hits_dev, dloc_dev, (hipStream_t)stream);
}

int sos_ray_columns(const sos_columns *c, int64_t n, const double *starts_dev, const int32_t *group_dev,
                    const double *dirs_dev, const uint8_t *hit_dev, double *out_dev, void *stream) {
    ColumnArgs a;
    if (int rc = column_args(c, a, "sos_ray_columns")) return rc;
    if (n < 0 || (n > 0 && (!starts_dev || !dirs_dev || !out_dev))) return arg_error("sos_ray_columns: null argument");
    return ray_columns_launch(n, starts_dev, group_dev, dirs_dev, hit_dev, a, out_dev, (hipStream_t)stream);
}

int sos_device_info(int32_t *cus, int32_t *xcds, char *arch, int32_t arch_len) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error("no HIP device");
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return set_error("hipGetDeviceProperties failed");
    if (cus) *cus = p.multiProcessorCount;
    if (xcds) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || n < 1) n = 1;
        *xcds = n;
    }
    if (arch && arch_len > 0) {
        strncpy(arch, p.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return SOS_OK;
}

// The device form of a grid's profile kinds: SOS_PROFILE_SHAPES resolved to the analytic shape of its component,
// scales turned into the inverses the device multiplies by.
static ProfileShape profile_shape(const sos_grid *g) {
    ProfileShape s;
    const int def_kind[6] = {PROFILE_EXP, PROFILE_EXP, PROFILE_GAUSS, PROFILE_GAUSS, PROFILE_AIR_SQUARED, PROFILE_GAUSS};
    const float def_a[6] = {1.0f / g->air_H, 1.0f / g->aer_H, 1.0f / g->o3_s, 1.0f / g->no2_s, 1.0f, 1.0f / g->st_s};
    const float def_b[6] = {0.0f, 0.0f, g->o3_c, g->no2_c, 0.0f, g->st_c};
    for (int k = 0; k < 6; ++k) {
        const float a = g->profile_a[k], b = g->profile_b[k];
        switch (g->profile_kind[k]) {
        case SOS_PROFILE_EXP: s.kind[k] = PROFILE_EXP; s.a[k] = 1.0f / a; s.b[k] = 0.0f; break;
        case SOS_PROFILE_GAUSS: s.kind[k] = PROFILE_GAUSS; s.a[k] = 1.0f / a; s.b[k] = b; break;
        case SOS_PROFILE_TRIANGLE: s.kind[k] = PROFILE_TRIANGLE; s.a[k] = 1.0f / a; s.b[k] = b; break;
        case SOS_PROFILE_TABLE: s.kind[k] = PROFILE_TABLE; s.a[k] = 1.0f / a; s.b[k] = 0.0f; break;
        case SOS_PROFILE_AIR_SQUARED: s.kind[k] = PROFILE_AIR_SQUARED; s.a[k] = 1.0f / (a * a); s.b[k] = 0.0f; break;
        case SOS_PROFILE_ZERO: s.kind[k] = PROFILE_ZERO; s.a[k] = 0.0f; s.b[k] = 0.0f; break;
        default: s.kind[k] = def_kind[k]; s.a[k] = def_a[k]; s.b[k] = def_b[k]; break;
        }
    }
    s.table = g->profile_table_dev;
    s.table_n = g->profile_table_n;
    return s;
}

// A grid whose profile kinds ask for a table must carry one of at least two nodes.
static bool profile_table_ok(const sos_grid *g) {
    for (int k = 0; k < 6; ++k)
        if (g->profile_kind[k] == SOS_PROFILE_TABLE && (!g->profile_table_dev || g->profile_table_n < 2)) return false;
    return true;
}

static void fill_args(TransportArgs &a, const sos_grid *g, const sos_bundle *b) {
    memset(&a, 0, sizeof(a));
    a.nh = g->nh; a.nc = g->nc; a.nq = g->nq; a.strato = g->strato; a.top = g->top_km;
    a.gl_u = g->gl_u_dev; a.gl_w = g->gl_w_dev; a.h = g->h_dev; a.chi = g->chi_dev;
    a.rays = b->rays;
    a.starts = b->starts_dev; a.ray_group = b->ray_group_dev; a.dirs = b->dirs_dev;
    a.edge_off = b->edge_off_dev; a.edges = b->edges_dev; a.lower = b->lower_dev;
}

int sos_plan_count(const sos_grid *g, const sos_bundle *b, int64_t *counts_dev, void *stream) {
    if (!g || !b || !counts_dev) return arg_error("sos_plan_count: null argument");
    TransportArgs a;
    fill_args(a, g, b);
    return plan_count_launch(a, counts_dev, (hipStream_t)stream);
}

int sos_plan_nodes(const sos_grid *g, const sos_bundle *b, const int64_t *node_off_dev, int32_t *node_cell_dev,
                   float *node_mu_dev, float *node_cp_dev, float *node_dir_dev, void *stream) {
    if (!g || !b || !node_off_dev || !node_cell_dev || !node_mu_dev || !node_cp_dev)
        return arg_error("sos_plan_nodes: null argument");
    if (g->nh < 2 || g->nc < 2) return arg_error("sos_plan_nodes: grid too small");
    TransportArgs a;
    fill_args(a, g, b);
    return plan_nodes_launch(a, node_off_dev, node_cell_dev, node_mu_dev, node_cp_dev, node_dir_dev,
                             (hipStream_t)stream);
}

int sos_plan_geometry(const sos_grid *g, const sos_bundle *b, const int64_t *node_off_dev, float *geom_dev,
                      void *stream) {
    if (!g || !b || !node_off_dev || !geom_dev) return arg_error("sos_plan_geometry: null argument");
    TransportArgs a;
    fill_args(a, g, b);
    if (!profile_table_ok(g)) return arg_error("profile table missing for a table profile kind");
    a.shape = profile_shape(g);
    a.targets = b->targets_dev; a.dloc = b->dloc_dev; a.nc = g->nc; a.nh = g->nh;
    return plan_geometry_launch(a, node_off_dev, geom_dev, (hipStream_t)stream);
}

int sos_plan_sun(const sos_grid *g, const sos_bundle *b, const int64_t *node_off_dev, float *sun_dev, void *stream) {
    if (!g || !b || !node_off_dev || !sun_dev) return arg_error("sos_plan_sun: null argument");
    if (!g->sun_col_dev || g->sun_nh < 2 || g->sun_nx < 2) return arg_error("sos_plan_sun: no sun table");
    TransportArgs a;
    fill_args(a, g, b);
    if (!profile_table_ok(g)) return arg_error("profile table missing for a table profile kind");
    a.shape = profile_shape(g);
    a.sun_nh = g->sun_nh; a.sun_nx = g->sun_nx; a.sun_h = g->sun_h_dev; a.sun_col = g->sun_col_dev;
    return plan_sun_launch(a, node_off_dev, sun_dev, (hipStream_t)stream);
}

int sos_ground_records(const sos_grid *g, int32_t groups, const float *group_g, const float *group_sg, int64_t items,
                       const int32_t *item_cell_dev, const int64_t *item_v0_dev, const int64_t *item_v1_dev,
                       const int32_t *vis_dev, const float *node_dir_dev, float *records_dev, void *stream) {
    if (groups < 1 || groups > 4 || !group_g || !group_sg)
        return arg_error("sos_ground_records: %d phase groups (1 to 4 with their asymmetries)", groups);
    if (!g || !item_cell_dev || !item_v0_dev || !item_v1_dev || !vis_dev || !node_dir_dev || !records_dev ||
        !g->goff_dev || !g->gdir_dev || !g->gval_dev || !g->lane_dev)
        return arg_error("sos_ground_records: null argument");
    if (g->lanes != 16 && g->lanes != 32) return arg_error("sos_ground_records: lanes %d (16 or 32)", g->lanes);
    GroundArgs a;
    a.G = g->lanes; a.strato = g->strato;
    if (!profile_table_ok(g)) return arg_error("profile table missing for a table profile kind");
    a.shape = profile_shape(g);
    a.items = items; a.item_cell = item_cell_dev; a.item_v0 = item_v0_dev; a.item_v1 = item_v1_dev; a.vis = vis_dev;
    a.node_dir = reinterpret_cast<const float4 *>(node_dir_dev); a.goff = g->goff_dev; a.gdir = g->gdir_dev;
    a.gval = g->gval_dev; a.lane = g->lane_dev; a.records = records_dev;
    a.groups = groups;
    a.pool_max = g->pool_max;
    if (g->pool_max < 0 || g->pool_max > 640)
        return arg_error("sos_ground_records: pool_max %d (0 to 640 pools per cell)", g->pool_max);
    for (int i = 0; i < 4; ++i) {
        a.gg[i] = i < groups ? group_g[i] : 0.0f;
        a.sgg[i] = i < groups ? group_sg[i] : 0.0f;
    }
    return ground_records_dispatch(a, (hipStream_t)stream);
}

int sos_ground_pools(const sos_grid *g, const int64_t *off_dev, const double *column_dev, const double *factor_dev,
                     const double *loc_dev, float *gval_dev, float *gm_dev, void *stream) {
    if (!g || !off_dev || !column_dev || !factor_dev || !loc_dev || !gval_dev || !gm_dev || !g->lane_dev)
        return arg_error("sos_ground_pools: null argument");
    if (!lanes_ok(g->lanes)) return arg_error("sos_ground_pools: lanes %d not in {1, 16, 32}", g->lanes);
    if (g->nh < 2 || g->nc < 2) return arg_error("sos_ground_pools: grid too small");
    return pool_values_launch(g->nh * g->nc, g->lanes, off_dev, column_dev, factor_dev, loc_dev, g->lane_dev, gval_dev,
                              gm_dev, (hipStream_t)stream);
}

int sos_plan_basis(const sos_grid *g, const sos_bundle *b, float *basis_dev, void *stream) {
    if (!g || !b || !basis_dev) return arg_error("sos_plan_basis: null argument");
    if (!order_ok(g->L)) return arg_error("sos_plan_basis: L %d", g->L);
    if (b->rays <= 0) return SOS_OK;
    if (!b->dirs_dev) return arg_error("sos_plan_basis: bundle without directions");
    return plan_basis_launch(g->L, b->rays, b->dirs_dev, basis_dev, (hipStream_t)stream);
}

int sos_sh_records(const sos_grid *g, const void *plane_dev, const float *scale_dev, int32_t half, int64_t items,
                   const int32_t *item_cell_dev, const int64_t *item_v0_dev, const int64_t *item_v1_dev,
                   const int32_t *node_dev, const int32_t *node_ray_dev, const float *node_cp_dev,
                   const float *basis_dev, float *records_dev, void *stream) {
    if (!g || !plane_dev || !item_cell_dev || !item_v0_dev || !item_v1_dev || !node_dev || !node_ray_dev ||
        !node_cp_dev || !basis_dev || !records_dev)
        return arg_error("sos_sh_records: null argument");
    if (half && !scale_dev) return arg_error("sos_sh_records: FP16 plane without its scale");
    if (!lanes_ok(g->lanes) || !order_ok(g->L)) return arg_error("sos_sh_records: lanes %d / L %d", g->lanes, g->L);
    if (g->nh < 2 || g->nc < 2) return arg_error("sos_sh_records: grid too small");
    ShArgs a;
    a.L = g->L; a.G = g->lanes; a.K = (g->L + 1) * (g->L + 2) / 2; a.nc = g->nc; a.plane = plane_dev; a.items = items;
    a.item_cell = item_cell_dev; a.item_v0 = item_v0_dev; a.item_v1 = item_v1_dev; a.node = node_dev;
    a.node_ray = node_ray_dev; a.node_cp = node_cp_dev; a.basis = basis_dev; a.records = records_dev;
    a.scale = scale_dev;
    a.plane_lo = (int64_t)g->nh * g->nc * ((a.K + 15) / 16 * 16) * a.G;
    a.tile = g->sh_tile;
    return sh_records_dispatch(a, half != 0, (hipStream_t)stream);
}

int sos_transport(const sos_grid *g, const sos_bundle *b, int32_t flags, int32_t pairs, int32_t rows,
                  const float *field_dev, const void *plane_dev, const float *boundary_dev, const float *records_dev,
                  float *out_dev, uint64_t *counters_dev, void *stream) {
    if (!g || !b || !out_dev) return arg_error("sos_transport: null grid, bundle or output");
    if (!lanes_ok(g->lanes)) return arg_error("sos_transport: lanes %d not in {1, 16, 32}", g->lanes);
    if (!order_ok(g->L)) return arg_error("sos_transport: L %d not in {8, 16, 32}", g->L);
    if (pairs != 4 && pairs != 6) return arg_error("sos_transport: pairs %d not 4 or 6", pairs);
    if (g->nh < 2 || g->nc < 2 || g->nq < 1) return arg_error("sos_transport: grid too small");
    const bool first = flags & SOS_FIRST;
    const bool records = flags & SOS_RECORDS;
    if (!first && (!field_dev || (!plane_dev && !records)))
        return arg_error("sos_transport: warm transport needs field and plane");
    if (!first && records && (!records_dev || !b->node_off_dev))
        return arg_error("sos_transport: records mode needs the records and the bundle's node offsets");
    const bool tensor4 = flags & SOS_TENSOR4;
    if (tensor4 && (first || !records || !(flags & SOS_NODE_GEOMETRY)))
        return arg_error("sos_transport: SOS_TENSOR4 is a form of the warm records consumer on node geometry");
    if (tensor4 && (int64_t)g->nh * g->nc * g->lanes >= ((int64_t)1 << 28))
        return arg_error("sos_transport: SOS_TENSOR4 tensor copy of %d x %d cells too large", g->nh, g->nc);
    if (!first && !tensor4 && rows < 10) return arg_error("sos_transport: rows %d < 10", rows);
    if ((flags & SOS_BOUNDARY) && !boundary_dev) return arg_error("sos_transport: boundary flag without boundary");
    if (first && (flags & SOS_BOUNDARY)) return arg_error("sos_transport: first light takes no boundary");
    if (first && (!g->sun_col_dev || g->sun_nh < 2 || g->sun_nx < 2 || !g->gm_dev || !g->goff_dev))
        return arg_error("sos_transport: first light needs the sun table, ground pools and gm");
    if (b->rays <= 0) return SOS_OK;
    TransportArgs a;
    memset(&a, 0, sizeof(a));
    a.nh = g->nh; a.nc = g->nc; a.nq = g->nq; a.strato = g->strato; a.flags = flags; a.pairs = pairs; a.rows = rows;
    a.blocks_per_cu = g->blocks_per_cu;
    a.sun_nh = g->sun_nh; a.sun_nx = g->sun_nx; a.top = g->top_km;
    if (!profile_table_ok(g)) return arg_error("profile table missing for a table profile kind");
    a.shape = profile_shape(g);
    a.gl_u = g->gl_u_dev; a.gl_w = g->gl_w_dev; a.h = g->h_dev; a.chi = g->chi_dev; a.hr = g->hr_dev;
    a.lane = g->lane_dev; a.sun_h = g->sun_h_dev; a.sun_col = g->sun_col_dev; a.goff = g->goff_dev;
    a.gdir = g->gdir_dev; a.gval = g->gval_dev; a.gm = g->gm_dev;
    a.rays = b->rays;
    const int rpw = 64 / g->lanes;
    a.tiles = (b->rays + rpw - 1) / rpw;
    a.starts = b->starts_dev; a.targets = b->targets_dev; a.ray_group = b->ray_group_dev; a.dirs = b->dirs_dev;
    a.dloc = b->dloc_dev; a.edge_off = b->edge_off_dev; a.edges = b->edges_dev; a.lower = b->lower_dev;
    a.end_j = b->end_j_dev; a.end_f = b->end_f_dev; a.end_col = b->end_col_dev; a.hits = b->hits_dev;
    a.field = field_dev; a.plane = plane_dev; a.boundary = boundary_dev; a.out = out_dev;
    a.counters = (unsigned long long *)counters_dev;
    a.records = (flags & SOS_FIRST) ? nullptr : records_dev;
    a.ground_records = ((flags & SOS_FIRST) && (flags & SOS_GROUND_RECORDS)) ? records_dev : nullptr;
    a.node_off = b->node_off_dev;
    if (a.ground_records && !a.node_off) return arg_error("sos_transport: ground records need the node offsets");
    if (flags & SOS_NODE_GEOMETRY) {
        if (pairs != 4 || !b->node_geom_dev || !b->node_off_dev)
            return arg_error("sos_transport: node geometry needs pairs 4, the node offsets and the geometry");
        if (first && (!a.ground_records || !b->node_sun_dev))
            return arg_error("sos_transport: first light on node geometry needs ground records and the sun geometry");
        if (!first && !records) return arg_error("sos_transport: warm node geometry needs the records path");
        a.node_geom = reinterpret_cast<const float4 *>(b->node_geom_dev);
        a.node_sun = reinterpret_cast<const float4 *>(b->node_sun_dev);
    }
    return transport_dispatch(a, g->lanes, g->L, (hipStream_t)stream);
}

int sos_project(const sos_grid *g, const sos_bundle *b, const float *rad_dev, int32_t ground_only,
                const float *basis_dev, float *moments_dev, float *flux_dev, void *stream) {
    if (!g || !b || !rad_dev || !flux_dev) return arg_error("sos_project: null argument");
    if (!ground_only && (!moments_dev || !basis_dev)) return arg_error("sos_project: moments output or basis missing");
    if (!lanes_ok(g->lanes) || !order_ok(g->L)) return arg_error("sos_project: lanes %d / L %d", g->lanes, g->L);
    ProjectArgs a;
    a.groups = b->groups; a.group_off = b->group_off_dev; a.dirs = b->dirs_dev; a.dloc = b->dloc_dev;
    a.weights = b->weights_dev; a.rad = rad_dev; a.lane = g->lane_dev; a.ground_only = ground_only;
    a.basis = basis_dev;
    a.moments = moments_dev; a.flux = flux_dev;
    return project_dispatch(a, g->lanes, g->L, (hipStream_t)stream);
}

int sos_compose(const sos_grid *g, int32_t rows, const float *moments_dev, const float *first_dev,
                const float *factor_dev, float *source_dev, void *stream) {
    if (!g || !moments_dev || !factor_dev || !source_dev) return arg_error("sos_compose: null argument");
    const int K = (g->L + 1) * (g->L + 2) / 2;
    if (rows != 10 + K && rows != 10 + 2 * K) return arg_error("sos_compose: rows %d for K %d", rows, K);
    return compose_launch((int64_t)g->nh * g->nc, rows, K, g->lanes, moments_dev, first_dev, factor_dev, source_dev,
                          (hipStream_t)stream);
}

int sos_make_plane(const sos_grid *g, int32_t rows, const float *source_dev, const float *ha_dev, const float *sd_dev,
                   int32_t half, void *plane_dev, float *scale_dev, void *stream) {
    if (!g || !source_dev || !ha_dev || !plane_dev) return arg_error("sos_make_plane: null argument");
    if (half && !scale_dev) return arg_error("sos_make_plane: FP16 plane without its scale");
    const int K = (g->L + 1) * (g->L + 2) / 2;
    const int strato = rows == 10 + 2 * K;
    if (rows != 10 + K && !strato) return arg_error("sos_make_plane: rows %d for K %d", rows, K);
    if (strato && !sd_dev) return arg_error("sos_make_plane: sulfate plane without its density");
    return plane_launch(g->nh, g->nc, g->L, rows, g->lanes, strato, source_dev, ha_dev, sd_dev, half, plane_dev,
                        scale_dev, (hipStream_t)stream);
}

int sos_source_metrics(int64_t rows, int32_t lanes, const float *g_dev, const float *x_dev, uint32_t frozen,
                       float *out_dev, void *stream) {
    if (!g_dev || !x_dev || !out_dev || lanes < 1 || lanes > 32) return arg_error("sos_source_metrics: bad argument");
    if (!lanes_ok(lanes)) return arg_error("sos_source_metrics: lanes %d", lanes);
    return metrics_launch(rows, lanes, g_dev, x_dev, frozen, out_dev, (hipStream_t)stream);
}

int sos_anderson_residual(int64_t cells, int32_t rows, int32_t lanes, const float *g_dev, const float *x_dev,
                          const float *obs_dev, int32_t nobs, float *work_dev, float *f_dev, void *stream) {
    if (!g_dev || !x_dev || !work_dev || !f_dev || (nobs > 0 && !obs_dev) || lanes < 1 || lanes > 32)
        return arg_error("sos_anderson_residual: bad argument");
    return residual_launch(cells, rows, lanes, g_dev, x_dev, obs_dev, nobs, work_dev, f_dev, (hipStream_t)stream);
}

int sos_anderson_gram(int64_t n, int32_t lanes, int32_t depth, int32_t slot, uint32_t filled, const float *ring_dev,
                      double *work_dev, double *gram_dev, void *stream) {
    if (!ring_dev || !work_dev || !gram_dev || depth < 1 || depth > 8 || slot < 0 || slot >= depth || lanes < 1 ||
        lanes > 32)
        return arg_error("sos_anderson_gram: bad argument");
    return gram_launch(n, lanes, depth, slot, filled, ring_dev, work_dev, gram_dev, (hipStream_t)stream);
}

int sos_anderson_gram_static(int64_t n, int32_t lanes, int32_t depth, int32_t slot, uint32_t filled,
                             const float *ring_dev, double *gram_dev, void *stream) {
    if (!ring_dev || !gram_dev || depth < 1 || depth > 8 || slot < 0 || slot >= depth || lanes < 1 || lanes > 32)
        return arg_error("sos_anderson_gram_static: bad argument");
    return gram_static_launch(n, lanes, depth, slot, filled, ring_dev, gram_dev, (hipStream_t)stream);
}

int sos_anderson_solve(int32_t lanes, int32_t depth, int32_t m, const int32_t *order_dev, const double *gram_dev,
                       uint32_t frozen, float *w_dev, void *stream) {
    if (!order_dev || !gram_dev || !w_dev || m < 1 || m > depth || depth > 8 || lanes < 1 || lanes > 32)
        return arg_error("sos_anderson_solve: bad argument");
    return solve_launch(lanes, depth, m, order_dev, gram_dev, frozen, w_dev, (hipStream_t)stream);
}

int sos_anderson_mix(int64_t cells, int32_t rows, int32_t lanes, int32_t depth, int32_t m, const float *ring_dev,
                     const int32_t *order_dev, const float *w_dev, float *x_dev, void *stream) {
    if (!ring_dev || !order_dev || !w_dev || !x_dev || m < 1 || m > depth || lanes < 1 || lanes > 32)
        return arg_error("sos_anderson_mix: bad argument");
    return mix_launch(cells, rows, lanes, depth, m, ring_dev, order_dev, w_dev, x_dev, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------- learned warm start
int sos_mlp_gemm(int64_t M, int64_t N, int64_t K, const float *a_dev, int64_t sa_i, int64_t sa_k, const float *b_dev,
                 int64_t sb_k, int64_t sb_j, const float *bias_dev, int32_t act, float *z_dev, int64_t ldz,
                 float *h_dev, int64_t ldh, int32_t splits, float *work_dev, void *stream) {
    if (M < 0 || N < 0 || K < 1 || !a_dev || !b_dev || !z_dev || ldz < N || (act == 1 && (!h_dev || ldh < N)) ||
        act < 0 || act > 1 || splits < 1 || splits > 256 || (splits > 1 && !work_dev))
        return arg_error("sos_mlp_gemm: bad argument");
    GemmArgs a;
    a.M = M; a.N = N; a.K = K; a.A = a_dev; a.sa_i = sa_i; a.sa_k = sa_k; a.B = b_dev; a.sb_k = sb_k; a.sb_j = sb_j;
    a.bias = bias_dev; a.act = act; a.Z = z_dev; a.ldz = ldz; a.H = h_dev; a.ldh = ldh; a.splits = splits; a.kchunk = K;
    return mlp_gemm_launch(a, work_dev, (hipStream_t)stream);
}

int sos_mlp_gelu_backward(int64_t rows, int64_t cols, const float *dh_dev, int64_t ldd, const float *z_dev,
                          int64_t ldz, float *dz_dev, int64_t ldo, void *stream) {
    if (rows < 0 || cols < 0 || !dh_dev || !z_dev || !dz_dev) return arg_error("sos_mlp_gelu_backward: bad argument");
    return mlp_gelu_backward_launch(rows, cols, dh_dev, ldd, z_dev, ldz, dz_dev, ldo, (hipStream_t)stream);
}

int sos_mlp_mse(int64_t rows, int64_t cols, const float *y_dev, int64_t ldy, const float *t_dev, int64_t ldt,
                const float *w_dev, float *dy_dev, int64_t lddy, float *loss_dev, double *work_dev, void *stream) {
    if (rows < 1 || cols < 1 || !y_dev || !t_dev || !loss_dev || !work_dev) return arg_error("sos_mlp_mse: bad argument");
    return mlp_mse_launch(rows, cols, y_dev, ldy, t_dev, ldt, w_dev, dy_dev, lddy, loss_dev, work_dev,
                          (hipStream_t)stream);
}

int sos_mlp_bias_grad(int64_t rows, int64_t cols, const float *dz_dev, int64_t ld, float *db_dev, void *stream) {
    if (rows < 0 || cols < 0 || !dz_dev || !db_dev) return arg_error("sos_mlp_bias_grad: bad argument");
    return mlp_bias_grad_launch(rows, cols, dz_dev, ld, db_dev, (hipStream_t)stream);
}

int sos_mlp_adam(int64_t n, float *p_dev, const float *g_dev, float *m_dev, float *v_dev, int32_t step, float lr,
                 float beta1, float beta2, float eps, void *stream) {
    if (n < 0 || !p_dev || !g_dev || !m_dev || !v_dev || step < 1) return arg_error("sos_mlp_adam: bad argument");
    return mlp_adam_launch(n, p_dev, g_dev, m_dev, v_dev, step, lr, beta1, beta2, eps, (hipStream_t)stream);
}
