// Shared device helpers of the MI355X SOS kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/sosfuse2_hip.h"

#define SOS_R_EARTH 6371.0
#define SOS_FLT_MIN 1.17549435e-38f
#define SOS_PI 3.14159265358979323846

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace sos {

constexpr int kidx(int l, int m) { return l * (l + 1) / 2 + m; }

// Offset of order m in the m-major SH plane: orders 0..m-1 hold L+1-m' degrees each.
template <int L>
constexpr int moff(int m) { return m * (L + 1) - m * (m - 1) / 2; }

// Largest i with a[i] <= x, clipped to [0, n-2] (numpy searchsorted(side='right') - 1, then clip).
__device__ inline int interval(const float *__restrict__ a, int n, float x) {
    int lo = 0, hi = n;  // first index with a[i] > x lies in [lo, hi]
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    int i = lo - 1;
    i = i < 0 ? 0 : i;
    return i > n - 2 ? n - 2 : i;
}

__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// Base-2 logarithm and exponential on the hardware's instructions (v_log_f32, v_exp_f32, about 1 ulp; inputs below
// FLT_MIN and results below it are flushed as everywhere here). The transport's log-space trace blend runs in base 2
// (its scale (G / A)^W does not depend on the base) and its attenuations are exp2(-tau log2 e).
constexpr float SOS_LOG2E = 1.44269504088896341f;
__device__ inline float log2_hw(float x) { return __builtin_amdgcn_logf(x); }
__device__ inline float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * SOS_LOG2E); }

// Henyey-Greenstein per steradian (numpy_reference.py:390-392).
__device__ inline float hg(float u, float g) {
    float den = fmaxf(1e-20f, (1.0f + g * g) - (2.0f * g) * u);
    return (1.0f - g * g) / (12.566370614359172f * (den * sqrtf(den)));
}

// The paired Henyey-Greenstein phase (hg(up) + hg(dn)) / 2 of numpy_reference.py:526-558 with the constants of one
// asymmetry hoisted, c = (1 + g^2, 2 g, (1 - g^2) / (4 pi)), and den^-3/2 as rsq^3 (v_rsq_f32, about 1 ulp: the value
// agrees with hg to a few 1e-7 relative) instead of a correctly rounded square root and division; for |up|, |dn| <= 1
// and |g| < 1 the denominator 1 + g^2 - 2 g u >= (1 - |g|)^2 > 0 needs no floor, and the half norm is applied once.
// For the inner loops that evaluate many phases of one asymmetry.
struct HgConst {
    float a, b, norm;
};
__device__ inline HgConst hg_const(float g) { return HgConst{1.0f + g * g, 2.0f * g, (1.0f - g * g) / 12.566370614359172f}; }
__device__ inline float hg_pair(float up, float dn, const HgConst &c) {
    const float ru = rsqrtf(c.a - c.b * up), rd = rsqrtf(c.a - c.b * dn);
    return (0.5f * c.norm) * ((ru * ru) * ru + (rd * rd) * rd);
}

// Height profiles of the five components (air, tropospheric aerosol, O3, NO2, O4) and the sulfate layer (index 5).
// Each has a kind: exponential exp(-h a) (a = 1 / scale height), Gaussian exp(-((h - b) a)^2 / 2) (a = 1 / sigma,
// b = centre), triangle max(0, 1 - |h - b| a) (a = 1 / half width), table (linear interpolation of table[k][i] at
// x = h a, a = 1 / step, as numpy_reference.profile_at's interpolate32, density.py:677-690), square of the air
// profile times a (O4 = (air / air_surface)^2, numpy_reference.py:219-220), or zero.
enum ProfileKind { PROFILE_EXP = 0, PROFILE_GAUSS = 1, PROFILE_TRIANGLE = 2, PROFILE_TABLE = 3, PROFILE_AIR_SQUARED = 4,
                   PROFILE_ZERO = 5 };

struct ProfileShape {
    int kind[6];
    float a[6], b[6];
    const float *table;  // [6][table_n] rows of the table kinds
    int table_n;
};

__device__ inline float profile_value(const ProfileShape &s, int k, float h, float air) {
    const float a = s.a[k], b = s.b[k];
    switch (s.kind[k]) {
    case PROFILE_EXP:
        return expf(-h * a);
    case PROFILE_GAUSS: {
        const float t = (h - b) * a;
        return expf(-0.5f * (t * t));
    }
    case PROFILE_TRIANGLE:
        return fmaxf(0.0f, 1.0f - fabsf(h - b) * a);
    case PROFILE_TABLE: {
        const float x = fmaxf(h, 0.0f) * a;
        const int i = min((int)x, s.table_n - 2);
        const float f = x - (float)i;
        const float *t = s.table + k * s.table_n;
        return t[i] + f * (t[i + 1] - t[i]);
    }
    case PROFILE_AIR_SQUARED:
        return (air * air) * a;
    default:
        return 0.0f;
    }
}

// Five component profiles and the sulfate profile at height h.
__device__ inline void profiles(const ProfileShape &s, float h, float *p, float &st) {
    const float air = profile_value(s, 0, h, 0.0f);
    p[0] = air;
    p[1] = profile_value(s, 1, h, air);
    p[2] = profile_value(s, 2, h, air);
    p[3] = profile_value(s, 3, h, air);
    p[4] = profile_value(s, 4, h, air);
    st = profile_value(s, 5, h, air);
}

// FP64 profile kinds of the column integrals (geometry.hip): codes as SOS_PROFILE_* (1 exponential, 2 Gaussian,
// 3 triangle, 4 table, 5 air squared, 6 zero), scales in km.
struct ProfileShape64 {
    int kind[6];
    double a[6], b[6];      // scale (km) and centre, as the host's ProfileSpec
    const double *table;    // [6][table_n] rows of the table kinds (FP64)
    int table_n;
};

struct ColumnArgs {
    double top;
    int shells;              // interior shells: 3, or 4 with 80 km when top > 80
    double shell_h[4];
    int nq;
    const double *gl_u, *gl_w;
    ProfileShape64 shape;
};

// Incident quadrature of a grid
struct IncidentArgs {
    int nh, nc, nmu, nphi, nlevels;
    const double *h, *chi;      // grid heights (km) and solar zenith nodes (rad)
    double levels[8];           // ring levels: -sqrt((R + h_ring)^2 - R^2)
    const double *gl_u, *gl_w;  // [nmu] Gauss-Legendre nodes and weights on [0, 1]
};

// Ray packs of a bundle
struct PackArgs {
    int nc;
    const double *chi;       // [nc]
    const double *starts;    // [cells][3]
    const double *targets;   // [cells][3][3] frame vectors as columns
    ColumnArgs c;            // top and shells of the boundaries
};

struct TransportArgs {
    // grid
    int nh, nc, nq, strato, flags, pairs, rows, sun_nh, sun_nx;
    int blocks_per_cu;  // persistent workgroups per CU (0: 8)
    float top;
    ProfileShape shape;
    const double *gl_u, *gl_w;
    const float *h, *chi, *hr, *lane;
    const float *sun_h, *sun_col;
    const int64_t *goff;
    const float *gdir, *gval, *gm;
    // bundle
    int64_t rays, tiles;
    const double *starts;
    const float *targets;
    const int32_t *ray_group;
    const double *dirs;
    const float *dloc;
    const int64_t *edge_off;
    const double *edges;
    const uint16_t *lower;
    const int32_t *end_j;
    const float *end_f, *end_col;
    const uint8_t *hits;
    // fields
    const float *field;
    const void *plane;
    const float *boundary;
    float *out;
    unsigned long long *counters;
    const float *records;    // [nodes][4][G] SH values per (node, corner) (records mode)
    const int64_t *node_off;  // [rays + 1] first node of every ray (records mode)
    const float *ground_records;  // [nodes][4][G] scattered ground light per (node, corner) (first light)
    const float4 *node_geom;  // [nodes][8] static node geometry (SOS_NODE_GEOMETRY)
    const float4 *node_sun;   // [nodes][3] sun geometry of first light (SOS_NODE_GEOMETRY with SOS_FIRST)
};

struct ProjectArgs {
    int64_t groups;
    const int64_t *group_off;
    const double *dirs;
    const float *dloc, *weights, *rad, *lane;
    const float *basis;       // [rays][KZ] the bundle's basis table (sos_plan_basis)
    int ground_only;
    float *moments, *flux;
};

struct ShArgs {
    int L, G, K, nc;
    const void *plane;         // [cells][K][G] float, or [2][cells][KZ/4][G][4] half (split hi, lo)
    int64_t items;
    const int32_t *item_cell;  // [items] base cell (ir * nc + j) of the item's nodes
    const int64_t *item_v0, *item_v1;  // node range of the item in the sorted order
    const int32_t *node;       // [nodes] node ids sorted by base cell
    const int32_t *node_ray;   // [nodes] the ray of each node, in the same order
    const float *node_cp;      // [nodes] the node's cos(phi), in the same order
    const float *basis;        // [rays][KZ] fp32(x_m P_lm(mu_ray)), m-major K order, zero padded (sos_plan_basis)
    float *records;            // [nodes][4][G]: corner c of node n at (4 n + c) G
    const float *scale;        // [cells][G] scale of a split FP16 plane (mixed precision)
    int64_t plane_lo;          // offset (elements) of the lo plane after the hi plane: cells * KZ * G
    int tile;                  // FP32 workgroup form: stepwise (1) or flow (0)
};

struct GroundArgs {
    int G, strato;
    ProfileShape shape;
    int64_t items;
    const int32_t *item_cell;
    const int64_t *item_v0, *item_v1;
    const int32_t *vis;
    const float4 *node_dir;    // [nodes] (mu_s, exd, |eyd|, h)
    const int64_t *goff;       // [cells + 1]
    const float *gdir, *gval;  // [pools][3], [pools][G]
    const float *lane;         // [SOS_LANE_CONSTANTS][G]
    float *records;            // [nodes * 4][G]
    int groups;                // phase groups: distinct (aerosol g, sulfate g) among the lanes, at most 4
    float gg[4], sgg[4];
    int pool_max;              // largest pool count of a cell
};

struct GemmArgs {
    int64_t M, N, K;
    const float *A;
    int64_t sa_i, sa_k;
    const float *B;
    int64_t sb_k, sb_j;
    const float *bias;  // [N] or null
    int act;            // 0: none, 1: GELU
    float *Z;           // [M][ldz] pre-activation (or the product), or the split partials
    int64_t ldz;
    float *H;           // [M][ldh] activation (act 1) or null
    int64_t ldh;
    int splits;
    int64_t kchunk;     // K range of one split (multiple of 32)
};

}  // namespace sos
