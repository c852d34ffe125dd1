// Line-of-sight transport through the spherical shell (gfx950).
//
// One lane per (ray, wavelength lane): G lanes per ray, 64/G rays per wavefront. Each lane walks its ray's segments
// and Gauss-Legendre nodes, keeps its own optical depth (Simpson prefix with Kahan at segment ends) and its own Kahan
// radiance sum, and reconstructs the source at the four bilinear corners of every node. Per ray and corner cell the
// 33 azimuthal amplitudes sum_l Y_lm(mu) c_lm are cached in registers and refilled only when the corner cell
// changes, after the reference's adjacent-stencil reuse (src/sosfuse2/backends/numpy_reference.py:106-129,573-600).
// Node positions, radii and the shadow test are FP64 (the chord near the tangent point and the solar terminator are
// where FP32 cancels); everything per wavelength is FP32 with denormals flushed.
//
// Operation order follows numpy_reference.py:467-643 (transport), :294-387 (sample), :236-291 (Simpson prefix),
// :431-458 (sun columns), :89-103 (log-space trace blend).

#include "common.h"
#include "sh_constants.h"

namespace sos {

// Azimuthal amplitudes of one corner cell for a ray with direction cosine mu = d_z (global frame):
// out[m] = sqrt2^(m>0) * sum_{l>=m} P_lm(mu) c[cell][(l,m)], the recurrence of arithmetic.py:86-109 in FP32.
template <int L, int G>
__device__ inline void refill(const float *plane, int cell, int w, float mu, float st, float (&out)[L + 1]) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    const int64_t base = (int64_t)cell * K * G + w;
    float diag = 0.28209479177387814f;
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        if (m) diag = (diag * ShConst<L>::dd()[m]) * st;
        float z0 = diag;
        const int64_t o = base + (int64_t)moff<L>(m) * G;
        float acc = z0 * plane[o];
        if (m < L) {
            float z1 = (ShConst<L>::sq()[m] * mu) * z0;
            acc = acc + z1 * plane[o + G];
            const float *aa = ShConst<L>::aa(), *bb = ShConst<L>::bb();
#pragma unroll 4
            for (int l = m + 2; l <= L; ++l) {
                const int k = kidx(l, m);
                float z2 = (aa[k] * mu) * z1 - bb[k] * z0;
                acc = acc + z2 * plane[o + (int64_t)(l - m) * G];
                z0 = z1;
                z1 = z2;
            }
        }
        out[m] = m ? acc * 1.41421356237309515f : acc;
    }
}

template <int P>
struct PairIdx;
template <>
struct PairIdx<4> { static constexpr int i[4] = {0, 1, 2, 0}, k[4] = {0, 1, 2, 2}; };
template <>
struct PairIdx<6> { static constexpr int i[6] = {0, 1, 2, 0, 0, 1}, k[6] = {0, 1, 2, 2, 1, 2}; };

// Node geometry shared by the lanes of a ray.
struct Node {
    float h, ds, ah, ac, cx, sx, sn, cc;
    int ir, j;
    double r, qz;
};

__device__ inline void node_position(const double *p, const double *d, double t, double &r, double &qz, double &qx,
                                     double &qy) {
    qx = fma(t, d[0], p[0]);
    qy = fma(t, d[1], p[1]);
    qz = fma(t, d[2], p[2]);
    r = sqrt(fma(qx, qx, fma(qy, qy, qz * qz)));
}

__device__ inline float node_height(double r, float top) {
    return clampf((float)(r - SOS_R_EARTH), 0.0f, top);
}

// Extinction of this lane at a height.
__device__ inline float extinction(const ProfileShape &s, const float *lc, float sext, float h) {
    float p[5], st;
    profiles(s, h, p, st);
    float e = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) e = e + p[k] * lc[k];
    return e + st * sext;
}

// Geometry of one quadrature node: height, azimuth frame, grid interval, bilinear weights (numpy_reference.py:294-326).
struct NodeGeom {
    float h, cx, sx, ah, ac, sn, cc;
    int ir, j;
};

__device__ inline NodeGeom node_geometry(double qx, double qy, double qz, double r, const float *__restrict__ hg,
                                         int nh, const float *__restrict__ chi, int nc, float top) {
    NodeGeom g;
    g.h = node_height(r, top);
    const double rho = sqrt(fma(qx, qx, qy * qy));
    g.cx = rho > 1e-12 ? (float)(qx / rho) : 1.0f;
    g.sx = rho > 1e-12 ? (float)(qy / rho) : 0.0f;
    const float angle = atan2f((float)rho, (float)qz);
    g.ir = interval(hg, nh, g.h);
    g.j = interval(chi, nc, angle);
    const float h0 = hg[g.ir], h1 = hg[g.ir + 1], c0 = chi[g.j], c1 = chi[g.j + 1];
    g.ah = clampf((g.h - h0) / (h1 - h0), 0.0f, 1.0f);
    g.ac = clampf((angle - c0) / (c1 - c0), 0.0f, 1.0f);
    const float ang = fmaf(g.ac, c1 - c0, c0);
    sincosf(ang, &g.sn, &g.cc);
    return g;
}

// cos(phi) of the ray direction about the sun axis relative to the node's meridian (numpy_reference.py:354).
__device__ inline float node_cp(float st, float df0, float df1, float cx, float sx) {
    return clampf(st > 1e-12f ? fmaf(df0, cx, df1 * sx) / fmaxf(st, 1e-30f) : 1.0f, -1.0f, 1.0f);
}

// Path coordinate and weight of node u of a segment [sa, sb], graded quadratically toward the lower end
// (numpy_reference.py:299-309).
__device__ inline void node_t(double sa, double sb, bool lower, double u, double wu, double &t, float &ds) {
    const double len = sb - sa;
    const double v = lower ? u : 1.0 - u;
    t = lower ? fma(len, u * u, sa) : fma(-len, v * v, sb);
    ds = (float)(len * v * wu) * 2.0f;
}

// Visit plan of a bundle: nodes per ray, then per node its cell (ir nc + j), the ray's mu and the node's cos(phi).
__global__ void plan_count_kernel(TransportArgs a, int64_t *counts) {
    const int64_t ray = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ray >= a.rays) return;
    const int64_t e0 = a.edge_off[ray];
    const int nseg = (int)(a.edge_off[ray + 1] - e0) - 1;
    int64_t n = 0;
    for (int seg = 0; seg < nseg; ++seg)
        if (a.edges[e0 + seg + 1] - a.edges[e0 + seg] > 1e-8) n += a.nq;
    counts[ray] = n;
}

__global__ void plan_nodes_kernel(TransportArgs a, const int64_t *node_off, int32_t *node_cell, float *node_mu,
                                  float *node_cp_out, float4 *node_dir) {
    const int64_t ray = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ray >= a.rays) return;
    const int grp = a.ray_group[ray];
    const double p[3] = {a.starts[3 * grp], a.starts[3 * grp + 1], a.starts[3 * grp + 2]};
    const double d[3] = {a.dirs[3 * ray], a.dirs[3 * ray + 1], a.dirs[3 * ray + 2]};
    const float df0 = (float)d[0], df1 = (float)d[1], mu = (float)d[2];
    const float st = (float)sqrt(fmax(0.0, 1.0 - d[2] * d[2]));
    const int64_t e0 = a.edge_off[ray];
    const int nseg = (int)(a.edge_off[ray + 1] - e0) - 1;
    const unsigned bits = a.lower[ray];
    int64_t n = node_off[ray];
    for (int seg = 0; seg < nseg; ++seg) {
        const double sa = a.edges[e0 + seg], sb = a.edges[e0 + seg + 1];
        if (!(sb - sa > 1e-8)) continue;
        const bool lower = (bits >> seg) & 1u;
        for (int step = 0; step < a.nq; ++step, ++n) {
            double t;
            float ds;
            node_t(sa, sb, lower, a.gl_u[step], a.gl_w[step], t, ds);
            double r, qz, qx, qy;
            node_position(p, d, t, r, qz, qx, qy);
            const NodeGeom ng = node_geometry(qx, qy, qz, r, a.h, a.nh, a.chi, a.nc, a.top);
            node_cell[n] = ng.ir * a.nc + ng.j;
            node_mu[n] = mu;
            node_cp_out[n] = node_cp(st, df0, df1, ng.cx, ng.sx);
            if (node_dir) {
                // the ray direction in the node's frame (numpy_reference.py:527-533) and the node height
                const float df2 = mu;
                const float Q[9] = {-ng.cc * ng.cx, ng.sx, ng.sn * ng.cx, -ng.cc * ng.sx, -ng.cx, ng.sn * ng.sx,
                                    ng.sn, 0.0f, ng.cc};
                float4 o;
                o.x = (df0 * Q[2] + df1 * Q[5]) + df2 * Q[8];
                o.y = (df0 * Q[0] + df1 * Q[3]) + df2 * Q[6];
                o.z = fabsf((df0 * Q[1] + df1 * Q[4]) + df2 * Q[7]);
                o.w = ng.h;
                node_dir[n] = o;
            }
        }
    }
}

// Static per-node geometry of a bundle for the warm consumer (one thread per ray), 8 float4 per node:
//   [0] piece[0..3]  [1] piece[4], piece[5], post[4], post[5]  [2] post[0..3]  [3] ds, ah, ac, cell | last << 31
//   [4..7] the tensor rotation T[q][u] (pairs xx, yy, zz, xz)
// piece[k] = w6 ((p_prev + p_node) + 4 p_mid)[k] is the Simpson panel ending at the node with the five component
// profiles and the sulfate profile in place of the lane's extinction (the lane takes piece . (ext, sext)); post is
// the segment's closing panel (last node of a segment only). Everything is computed exactly as the transport kernel
// computes it per lane (FP64 positions, the same node geometry), once per bundle instead of once per evaluation and
// lane.
__global__ void plan_geometry_kernel(TransportArgs a, const int64_t *node_off, float4 *geom) {
    const int64_t ray = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ray >= a.rays) return;
    const int grp = a.ray_group[ray];
    const double p[3] = {a.starts[3 * grp], a.starts[3 * grp + 1], a.starts[3 * grp + 2]};
    const double d[3] = {a.dirs[3 * ray], a.dirs[3 * ray + 1], a.dirs[3 * ray + 2]};
    float B[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] = a.targets[9 * grp + k];
    const float df0 = (float)d[0], df1 = (float)d[1], df2 = (float)d[2];
    const float dl[3] = {a.dloc[3 * ray], a.dloc[3 * ray + 1], a.dloc[3 * ray + 2]};
    const int64_t e0 = a.edge_off[ray];
    const int nseg = (int)(a.edge_off[ray + 1] - e0) - 1;
    const unsigned bits = a.lower[ray];
    int64_t n = node_off[ray];
    auto prof6 = [&](double t, float *o) {
        double r, qz, qx, qy;
        node_position(p, d, t, r, qz, qx, qy);
        float st;
        profiles(a.shape, node_height(r, a.top), o, st);
        o[5] = st;
    };
    for (int seg = 0; seg < nseg; ++seg) {
        const double sa = a.edges[e0 + seg], sb = a.edges[e0 + seg + 1];
        if (!(sb - sa > 1e-8)) continue;
        const bool lower = (bits >> seg) & 1u;
        double tprev = sa;
        float pprev[6];
        prof6(sa, pprev);
        for (int step = 0; step < a.nq; ++step, ++n) {
            double t;
            float ds;
            node_t(sa, sb, lower, a.gl_u[step], a.gl_w[step], t, ds);
            double r, qz, qx, qy;
            node_position(p, d, t, r, qz, qx, qy);
            const NodeGeom ng = node_geometry(qx, qy, qz, r, a.h, a.nh, a.chi, a.nc, a.top);
            float pn[6], pm[6], piece[6], post[6];
            {
                float st;
                profiles(a.shape, ng.h, pn, st);
                pn[5] = st;
            }
            prof6(0.5 * (tprev + t), pm);
            const float w6 = (float)(t - tprev) / 6.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) piece[k] = w6 * ((pprev[k] + pn[k]) + 4.0f * pm[k]);
            const bool last = step == a.nq - 1;
#pragma unroll
            for (int k = 0; k < 6; ++k) post[k] = 0.0f;
            if (last) {
                float pe[6], pme[6];
                prof6(0.5 * (t + sb), pme);
                prof6(sb, pe);
                const float w6e = (float)(sb - t) / 6.0f;
#pragma unroll
                for (int k = 0; k < 6; ++k) post[k] = w6e * ((pn[k] + pe[k]) + 4.0f * pme[k]);
            }
            const float cx = ng.cx, sx = ng.sx, sn = ng.sn, cc = ng.cc;
            const float Q[9] = {-cc * cx, sx, sn * cx, -cc * sx, -cx, sn * sx, sn, 0.0f, cc};
            float T[4][4];
            {
                float A[3][3], dt[3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) A[i][jj] = (B[i] * Q[jj] + B[3 + i] * Q[3 + jj]) + B[6 + i] * Q[6 + jj];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) dt[jj] = (df0 * Q[jj] + df1 * Q[3 + jj]) + df2 * Q[6 + jj];
                float Kx[3][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) Kx[i][jj] = A[i][jj] - dl[i] * dt[jj];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = PairIdx<4>::i[q], k = PairIdx<4>::k[q];
                    T[q][0] = Kx[i][0] * Kx[k][0];
                    T[q][1] = Kx[i][1] * Kx[k][1];
                    T[q][2] = Kx[i][2] * Kx[k][2];
                    T[q][3] = Kx[i][0] * Kx[k][2] + Kx[i][2] * Kx[k][0];
                }
            }
            float4 *o = geom + n * 8;
            // base cell as (ring << 16) | column, so the consumers need no division by the column count
            const unsigned cellbits = ((unsigned)ng.ir << 16) | (unsigned)ng.j | (last ? 0x80000000u : 0u);
            o[0] = make_float4(piece[0], piece[1], piece[2], piece[3]);
            o[1] = make_float4(piece[4], piece[5], post[4], post[5]);
            o[2] = make_float4(post[0], post[1], post[2], post[3]);
            o[3] = make_float4(ds, ng.ah, ng.ac, __uint_as_float(cellbits));
#pragma unroll
            for (int q = 0; q < 4; ++q) o[4 + q] = make_float4(T[q][0], T[q][1], T[q][2], T[q][3]);
            pprev[0] = pn[0]; pprev[1] = pn[1]; pprev[2] = pn[2]; pprev[3] = pn[3]; pprev[4] = pn[4]; pprev[5] = pn[5];
            tprev = t;
        }
    }
}

// Warm consumer on the static node geometry (records mode, pairs xx yy zz xz): per node and lane only the lane's
// optical depth, attenuation, the four corners (Rayleigh tensor from the field, SH value from the records) with the
// log-space trace blend, and the Kahan radiance sum remain. Same operations as transport_kernel's warm path.
// TENSOR4: the field is the lane-major tensor copy [cells][G][4] (one 16-byte read per corner, 32-bit offsets).
template <int G, bool BOUNDARY, bool TENSOR4>
__global__ __launch_bounds__(256) void transport_geom_kernel(TransportArgs a) {
    constexpr int RPW = 64 / G;
    constexpr int P = 4;
    const int lane = threadIdx.x & 63;
    const int w = lane % G;
    const int sub = lane / G;
    float lc[SOS_LANE_CONSTANTS];
#pragma unroll
    for (int k = 0; k < SOS_LANE_CONSTANTS; ++k) lc[k] = a.lane[k * G + w];
    const float mix = lc[SOS_LANE_MIX], kmix = (1.0f - mix) / 3.0f;
    const float sext = a.strato ? lc[SOS_LANE_SEXT] : 0.0f;
    const float ext6[6] = {lc[0], lc[1], lc[2], lc[3], lc[4], sext};
    const int64_t nwg = gridDim.x;
    const int xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
    const int64_t per_xcd_wg = (nwg + 7 - xcd) >> 3;
    const int64_t waves_xcd = per_xcd_wg * 4;
    const int64_t chunk = (a.tiles + 7) >> 3;
    const int64_t t0 = xcd * chunk, t1 = min(a.tiles, t0 + chunk);
    const int nc = a.nc;
    for (int64_t tile = t0 + idx * 4 + (threadIdx.x >> 6); tile < t1; tile += waves_xcd) {
        const int64_t ray = tile * RPW + sub;
        if (ray >= a.rays) continue;
        const int grp = a.ray_group[ray];
        const float dl[3] = {a.dloc[3 * ray], a.dloc[3 * ray + 1], a.dloc[3 * ray + 2]};
        float pu[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int i = PairIdx<P>::i[q], k = PairIdx<P>::k[q];
            pu[q] = (i == k ? 1.0f : 0.0f) - dl[i] * dl[k];
        }
        (void)grp;
        const float p3 = (pu[0] + pu[1]) + pu[2];
        float total[P], corr[P];
#pragma unroll
        for (int q = 0; q < P; ++q) total[q] = corr[q] = 0.0f;
        float tau = 0.0f, tc = 0.0f, acc = 0.0f;
        const int64_t n0 = a.node_off[ray], n1 = a.node_off[ray + 1];
        // the panels and cell of node n + 1 are loaded while node n is evaluated (its records and tensor reads
        // depend on the cell)
        float4 p0 = float4{0.f, 0.f, 0.f, 0.f}, p1 = p0, p3n = p0;
        if (n0 < n1) {
            const float4 *gp = a.node_geom + n0 * 8;
            p0 = gp[0]; p1 = gp[1]; p3n = gp[3];
        }
        for (int64_t n = n0; n < n1; ++n) {
            const float4 *g = a.node_geom + n * 8;
            const float4 g0 = p0, g1 = p1, g3 = p3n;
            if (n + 1 < n1) {
                const float4 *gn = g + 8;
                p0 = gn[0]; p1 = gn[1]; p3n = gn[3];
            }
            acc = acc + (((((g0.x * ext6[0] + g0.y * ext6[1]) + g0.z * ext6[2]) + g0.w * ext6[3]) + g1.x * ext6[4]) +
                         g1.y * ext6[5]);
            const float mid = (tau - tc) + acc;
            const float att = g3.x * exp_hw(-mid);
            const float ah = g3.y, ac = g3.z;
            const unsigned cb = __float_as_uint(g3.w);
            const int ir = (int)((cb >> 16) & 0x7fffu);
            const int cell0 = ir * nc + (int)(cb & 0xffffu);
            const int cells[4] = {cell0, cell0 + 1, cell0 + nc, cell0 + nc + 1};
            const float sw[4] = {(1.0f - ah) * (1.0f - ac), (1.0f - ah) * ac, ah * (1.0f - ac), ah * ac};
            float T[P][4];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float4 t4 = g[4 + q];
                T[q][0] = t4.x; T[q][1] = t4.y; T[q][2] = t4.z; T[q][3] = t4.w;
            }
            // The node's rotation T is the same for its four corners, so the corners are blended before it is
            // applied: with d the corner's air density, a its SH value and k = (1 - mix) / 3 trace(rm), a corner's
            // tensor is tens[q] = d mix (T rm)[q] + pu[q] (d k + a / 2), so val[q] = sum_c sw tens[q] = mix (T R)[q]
            // + pu[q] S with R = sum_c sw d rm and S = sum_c sw (d k + a / 2); each corner's trace for the blend is
            // d (mix (t3 . rm) + k p3) + a p3 / 2 with t3 = T[0] + T[1] + T[2] and p3 = pu[0] + pu[1] + pu[2].
            float t3[4];
#pragma unroll
            for (int uu = 0; uu < 4; ++uu) t3[uu] = (T[0][uu] + T[1][uu]) + T[2][uu];
            const float *rec = a.records + n * 4 * G + w;
            float R[4] = {0.0f, 0.0f, 0.0f, 0.0f}, S = 0.0f;
            float logs = 0.0f, wg = 0.0f, am = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float aero = rec[c * G];
                float rm[4];
                if constexpr (TENSOR4) {
                    const float4 t = reinterpret_cast<const float4 *>(a.field)[(uint32_t)(cells[c] * G + w)];
                    rm[0] = t.x; rm[1] = t.y; rm[2] = t.z; rm[3] = t.w;
                } else {
                    const float *f = a.field + (int64_t)cells[c] * a.rows * G + w;
                    rm[0] = f[0]; rm[1] = f[G]; rm[2] = f[2 * G]; rm[3] = f[3 * G];
                }
                const float k = kmix * ((rm[0] + rm[1]) + rm[2]);
                const float density = a.hr[ir + (c >> 1)];
                const float rot3 = ((t3[0] * rm[0] + t3[1] * rm[1]) + t3[2] * rm[2]) + t3[3] * rm[3];
                const float half_a = aero * 0.5f;
                const float tr = density * (mix * rot3 + k * p3) + half_a * p3;
                const float sd = sw[c] * density;
#pragma unroll
                for (int uu = 0; uu < 4; ++uu) R[uu] = R[uu] + sd * rm[uu];
                S = S + sw[c] * (density * k + half_a);
                if (tr >= SOS_FLT_MIN) {
                    logs = logs + sw[c] * log2_hw(tr);
                    wg = wg + sw[c];
                    am = am + sw[c] * tr;
                }
            }
            float val[P];
#pragma unroll
            for (int q = 0; q < P; ++q)
                val[q] = mix * (((T[q][0] * R[0] + T[q][1] * R[1]) + T[q][2] * R[2]) + T[q][3] * R[3]) + pu[q] * S;
            const bool ok = (wg > 0.0f) && (am >= SOS_FLT_MIN);
            const float mean = ok ? am * __builtin_amdgcn_rcpf(wg) : 1.0f;
            const float scale = exp2_hw(ok ? logs - wg * log2_hw(mean) : 0.0f);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float y = (val[q] * scale) * att - corr[q];
                const float tn = total[q] + y;
                corr[q] = (tn - total[q]) - y;
                total[q] = tn;
            }
            if (cb & 0x80000000u) {
                const float4 g2 = g[2];
                acc = acc + (((((g2.x * ext6[0] + g2.y * ext6[1]) + g2.z * ext6[2]) + g2.w * ext6[3]) + g1.z * ext6[4]) +
                             g1.w * ext6[5]);
                const float y = acc - tc;
                const float tn = tau + y;
                tc = (tn - tau) - y;
                tau = tn;
                acc = 0.0f;
            }
        }
        if constexpr (BOUNDARY) {
            const float ef = a.end_f[ray];
            const int ej = a.end_j[ray];
            float E = (1.0f - ef) * a.boundary[ej * G + w] + ef * a.boundary[(ej + 1) * G + w];
            float optical = 0.0f;
#pragma unroll
            for (int k = 0; k < 5; ++k) optical = optical + a.end_col[6 * ray + k] * lc[k];
            optical = optical + a.end_col[6 * ray + 5] * sext;
            E = E * (exp_hw(-optical) * (a.hits[ray] ? 1.0f : 0.0f));
#pragma unroll
            for (int q = 0; q < P; ++q) total[q] = total[q] + (E * pu[q]) / 2.0f;
        }
        float *o = a.out + (ray * G + w) * P;
#pragma unroll
        for (int q = 0; q < P; ++q) o[q] = total[q];
    }
}

// Sun geometry of every node for first light (one thread per ray), 3 float4 per node:
//   [0] col[0..3]  [1] col[4], col[5], lit, air profile  [2] aerosol profile, sulfate profile, 0, 0
// col[k] are the bilinear sun-table column integrals of the six components at the node (FP64 lit test near the
// terminator, numpy_reference.py:431-458), computed exactly as the per-lane first-light path computes them.
__global__ void plan_sun_kernel(TransportArgs a, const int64_t *node_off, float4 *sun) {
    const int64_t ray = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ray >= a.rays) return;
    const int grp = a.ray_group[ray];
    const double p[3] = {a.starts[3 * grp], a.starts[3 * grp + 1], a.starts[3 * grp + 2]};
    const double d[3] = {a.dirs[3 * ray], a.dirs[3 * ray + 1], a.dirs[3 * ray + 2]};
    const int64_t e0 = a.edge_off[ray];
    const int nseg = (int)(a.edge_off[ray + 1] - e0) - 1;
    const unsigned bits = a.lower[ray];
    int64_t n = node_off[ray];
    for (int seg = 0; seg < nseg; ++seg) {
        const double sa = a.edges[e0 + seg], sb = a.edges[e0 + seg + 1];
        if (!(sb - sa > 1e-8)) continue;
        const bool lower = (bits >> seg) & 1u;
        for (int step = 0; step < a.nq; ++step, ++n) {
            double t;
            float ds;
            node_t(sa, sb, lower, a.gl_u[step], a.gl_w[step], t, ds);
            double r, qz, qx, qy;
            node_position(p, d, t, r, qz, qx, qy);
            const float h = node_height(r, a.top);
            double height = r - SOS_R_EARTH;
            if (!(height > 0.0)) height = 0.0;
            const double rad = SOS_R_EARTH + height;
            const double mg = -sqrt(height * (2.0 * SOS_R_EARTH + height)) / rad;
            double delta = qz / r - mg;
            if (fabs(delta) < (double)SOS_FLT_MIN) delta = 0.0;
            const bool lit = delta >= 0.0;
            const float x = clampf((float)sqrt((lit ? delta : 0.0) / (1.0 - mg)), 0.0f, 1.0f);
            const float hh = clampf(h, 0.0f, a.top);
            const int si = interval(a.sun_h, a.sun_nh, hh);
            const float al = (hh - a.sun_h[si]) / (a.sun_h[si + 1] - a.sun_h[si]);
            const float fx = x * (float)(a.sun_nx - 1);
            int sj = (int)fx;
            sj = sj < 0 ? 0 : (sj > a.sun_nx - 2 ? a.sun_nx - 2 : sj);
            const float bx = fx - (float)sj;
            const float w00 = (1.0f - al) * (1.0f - bx), w10 = al * (1.0f - bx), w01 = (1.0f - al) * bx, w11 = al * bx;
            const float *c00 = a.sun_col + ((int64_t)si * a.sun_nx + sj) * 6;
            const float *c10 = c00 + (int64_t)a.sun_nx * 6;
            float col[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                float v = w00 * c00[k];
                v = v + w10 * c10[k];
                v = v + w01 * c00[6 + k];
                v = v + w11 * c10[6 + k];
                col[k] = v;
            }
            float prof[5], stp;
            profiles(a.shape, h, prof, stp);
            float4 *o = sun + n * 3;
            o[0] = make_float4(col[0], col[1], col[2], col[3]);
            o[1] = make_float4(col[4], col[5], lit ? 1.0f : 0.0f, prof[0]);
            o[2] = make_float4(prof[1], stp, 0.0f, 0.0f);
        }
    }
}

// First light on the static node geometry and sun geometry, with the ground term from the ground records (pairs
// xx yy zz xz). Same operations as transport_kernel's first-light path with ground records.
template <int G>
__global__ __launch_bounds__(256) void transport_first_geom_kernel(TransportArgs a) {
    constexpr int RPW = 64 / G;
    constexpr int P = 4;
    const int lane = threadIdx.x & 63;
    const int w = lane % G;
    const int sub = lane / G;
    float lc[SOS_LANE_CONSTANTS];
#pragma unroll
    for (int k = 0; k < SOS_LANE_CONSTANTS; ++k) lc[k] = a.lane[k * G + w];
    const float mix = lc[SOS_LANE_MIX], gasy = lc[SOS_LANE_G], asca = lc[SOS_LANE_ASCA];
    const float sext = a.strato ? lc[SOS_LANE_SEXT] : 0.0f;
    const float ext6[6] = {lc[0], lc[1], lc[2], lc[3], lc[4], sext};
    const int64_t nwg = gridDim.x;
    const int xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
    const int64_t per_xcd_wg = (nwg + 7 - xcd) >> 3;
    const int64_t waves_xcd = per_xcd_wg * 4;
    const int64_t chunk = (a.tiles + 7) >> 3;
    const int64_t t0 = xcd * chunk, t1 = min(a.tiles, t0 + chunk);
    const int nc = a.nc;
    for (int64_t tile = t0 + idx * 4 + (threadIdx.x >> 6); tile < t1; tile += waves_xcd) {
        const int64_t ray = tile * RPW + sub;
        if (ray >= a.rays) continue;
        const int grp = a.ray_group[ray];
        const float df2 = (float)a.dirs[3 * ray + 2];
        const float dl[3] = {a.dloc[3 * ray], a.dloc[3 * ray + 1], a.dloc[3 * ray + 2]};
        float pu[P], direct[P];
        {
            const float u0 = a.targets[9 * grp + 6], u1 = a.targets[9 * grp + 7], u2 = a.targets[9 * grp + 8];
            const float dot = (dl[0] * u0 + dl[1] * u1) + dl[2] * u2;
            const float e[3] = {u0 - dl[0] * dot, u1 - dl[1] * dot, u2 - dl[2] * dot};
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int i = PairIdx<P>::i[q], k = PairIdx<P>::k[q];
                pu[q] = (i == k ? 1.0f : 0.0f) - dl[i] * dl[k];
                direct[q] = (pu[q] - e[i] * e[k]) / 2.0f;
            }
        }
        const float phase = hg(df2, gasy);
        const float sphase = a.strato ? hg(df2, lc[SOS_LANE_SG]) : 0.0f;
        float total[P], corr[P];
#pragma unroll
        for (int q = 0; q < P; ++q) total[q] = corr[q] = 0.0f;
        float tau = 0.0f, tc = 0.0f, acc = 0.0f;
        const int64_t n0 = a.node_off[ray], n1 = a.node_off[ray + 1];
        for (int64_t n = n0; n < n1; ++n) {
            const float4 *g = a.node_geom + n * 8;
            const float4 g0 = g[0], g1 = g[1], g3 = g[3];
            acc = acc + (((((g0.x * ext6[0] + g0.y * ext6[1]) + g0.z * ext6[2]) + g0.w * ext6[3]) + g1.x * ext6[4]) +
                         g1.y * ext6[5]);
            const float mid = (tau - tc) + acc;
            const float att = g3.x * exp_hw(-mid);
            const float ah = g3.y, ac = g3.z;
            const unsigned cb = __float_as_uint(g3.w);
            const int cell0 = (int)((cb >> 16) & 0x7fffu) * nc + (int)(cb & 0xffffu);
            const int cells[4] = {cell0, cell0 + 1, cell0 + nc, cell0 + nc + 1};
            const float sw[4] = {(1.0f - ah) * (1.0f - ac), (1.0f - ah) * ac, ah * (1.0f - ac), ah * ac};
            const float4 *sg = a.node_sun + n * 3;
            const float4 s0 = sg[0], s1 = sg[1], s2 = sg[2];
            const float col[6] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y};
            float optical = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) optical = optical + col[k] * (k < 5 ? lc[k] : sext);
            const float sun = s1.z != 0.0f ? exp_hw(-optical) : 0.0f;
            const float prof0 = s1.w, prof1 = s2.x, stp = s2.y;
            const float *rec = a.ground_records + n * 4 * G + w;
            const float grec = ((sw[0] * rec[0] + sw[1] * rec[G]) + sw[2] * rec[2 * G]) + sw[3] * rec[3 * G];
            const float beam = (prof1 * asca) * (sun * phase + 0.0f);
            const float ray_ = (sun * prof0) * lc[0];
            const float sbeam = (stp * lc[SOS_LANE_SSCA]) * (sun * sphase + 0.0f);
            float val[P];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                val[q] = ray_ * (mix * 0.119366207318921496f * direct[q] +
                                 (1.0f - mix) * 0.0397887357729738339f * pu[q]) +
                         beam * pu[q] / 2.0f;
                if (a.strato) val[q] = val[q] + (sbeam * pu[q]) / 2.0f;
                val[q] = val[q] + (grec * pu[q]) / 2.0f;
            }
            float T[P][4];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float4 t4 = g[4 + q];
                T[q][0] = t4.x; T[q][1] = t4.y; T[q][2] = t4.z; T[q][3] = t4.w;
            }
            // the corners' ground tensors are blended before the node's rotation (linear in them): R = sum_c sw rm,
            // S = sum_c sw trace(rm), one 4 x 4 rotation per node instead of four
            float R[4] = {0.0f, 0.0f, 0.0f, 0.0f}, S = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float *gm = a.gm + (int64_t)cells[c] * 4 * G + w;
                const float rm[4] = {gm[0], gm[G], gm[2 * G], gm[3 * G]};
#pragma unroll
                for (int uu = 0; uu < 4; ++uu) R[uu] = R[uu] + sw[c] * rm[uu];
                S = S + sw[c] * ((rm[0] + rm[1]) + rm[2]);
            }
            const float kS = (1.0f - mix) / 3.0f * S;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float rot = ((T[q][0] * R[0] + T[q][1] * R[1]) + T[q][2] * R[2]) + T[q][3] * R[3];
                val[q] = val[q] + prof0 * (mix * rot + kS * pu[q]);
            }
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float y = val[q] * att - corr[q];
                const float tn = total[q] + y;
                corr[q] = (tn - total[q]) - y;
                total[q] = tn;
            }
            if (cb & 0x80000000u) {
                const float4 g2 = g[2];
                acc = acc + (((((g2.x * ext6[0] + g2.y * ext6[1]) + g2.z * ext6[2]) + g2.w * ext6[3]) + g1.z * ext6[4]) +
                             g1.w * ext6[5]);
                const float y = acc - tc;
                const float tn = tau + y;
                tc = (tn - tau) - y;
                tau = tn;
                acc = 0.0f;
            }
        }
        float *o = a.out + (ray * G + w) * P;
#pragma unroll
        for (int q = 0; q < P; ++q) o[q] = total[q];
    }
}

enum Aero { DIRECT = 0, RECORDS = 1 };

template <int G, int L, int P, bool FIRST, bool BOUNDARY, int AERO, bool UNI = false>
__global__ __launch_bounds__(256) void transport_kernel(TransportArgs a) {
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63;
    const int w = lane % G;
    const int sub = lane / G;
    // lane constants
    float lc[SOS_LANE_CONSTANTS];
#pragma unroll
    for (int k = 0; k < SOS_LANE_CONSTANTS; ++k) lc[k] = a.lane[k * G + w];
    const float mix = lc[SOS_LANE_MIX], gasy = lc[SOS_LANE_G], asca = lc[SOS_LANE_ASCA];
    const float sext = a.strato ? lc[SOS_LANE_SEXT] : 0.0f;
    // XCD-contiguous tile ranges: workgroup b runs on XCD b % 8.
    const int64_t nwg = gridDim.x;
    const int xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
    const int64_t per_xcd_wg = (nwg + 7 - xcd) >> 3;  // workgroups on this XCD
    const int64_t waves_xcd = per_xcd_wg * 4;
    const int64_t chunk = (a.tiles + 7) >> 3;
    const int64_t t0 = xcd * chunk, t1 = min(a.tiles, t0 + chunk);
    for (int64_t tile = t0 + idx * 4 + (threadIdx.x >> 6); tile < t1; tile += waves_xcd) {
        const int64_t ray = tile * RPW + sub;
        if (ray >= a.rays) continue;
        const int grp = a.ray_group[ray];
        double p[3], d[3];
        p[0] = a.starts[3 * grp]; p[1] = a.starts[3 * grp + 1]; p[2] = a.starts[3 * grp + 2];
        d[0] = a.dirs[3 * ray]; d[1] = a.dirs[3 * ray + 1]; d[2] = a.dirs[3 * ray + 2];
        float B[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) B[k] = a.targets[9 * grp + k];  // B[3*k + i]: component k of basis vector i
        const float df0 = (float)d[0], df1 = (float)d[1], df2 = (float)d[2];
        const float dl[3] = {a.dloc[3 * ray], a.dloc[3 * ray + 1], a.dloc[3 * ray + 2]};
        const float st = (float)sqrt(fmax(0.0, 1.0 - d[2] * d[2]));
        const float mu = df2;
        float pu[P], direct[P];
        {
            // sun direction in the target frame: row 2 of B
            const float u0 = B[6], u1 = B[7], u2 = B[8];
            const float dot = (dl[0] * u0 + dl[1] * u1) + dl[2] * u2;
            const float e[3] = {u0 - dl[0] * dot, u1 - dl[1] * dot, u2 - dl[2] * dot};
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int i = PairIdx<P>::i[q], k = PairIdx<P>::k[q];
                pu[q] = (i == k ? 1.0f : 0.0f) - dl[i] * dl[k];
                direct[q] = (pu[q] - e[i] * e[k]) / 2.0f;
            }
        }
        float phase = 0.0f, sphase = 0.0f;
        if constexpr (FIRST) {
            phase = hg(df2, gasy);
            if (a.strato) sphase = hg(df2, lc[SOS_LANE_SG]);
        }
        float total[P], corr[P];
#pragma unroll
        for (int q = 0; q < P; ++q) total[q] = corr[q] = 0.0f;
        float tau = 0.0f, tc = 0.0f;
        float cache[AERO == RECORDS ? 1 : 4][AERO == RECORDS ? 1 : L + 1];
        int tags[4] = {-1, -1, -1, -1};
        unsigned nodes = 0, refills = 0;
        // node offsets exist only when records are read: warm records, or first light with ground records
        const bool uses_nodes = FIRST ? (a.ground_records != nullptr) : (AERO == RECORDS);
        const int64_t node_base = uses_nodes ? a.node_off[ray] : 0;
        const int64_t e0 = a.edge_off[ray];
        const int nseg = (int)(a.edge_off[ray + 1] - e0) - 1;
        const unsigned bits = a.lower[ray];
        for (int seg = 0; seg < nseg; ++seg) {
            const double sa = a.edges[e0 + seg], sb = a.edges[e0 + seg + 1];
            if (!(sb - sa > 1e-8)) continue;
            const bool lower = (bits >> seg) & 1u;
            const double len = sb - sa;
            // Simpson prefix: panel start at the segment start
            double tprev = sa;
            double r_, qz_, qx_, qy_;
            node_position(p, d, sa, r_, qz_, qx_, qy_);
            float bprev = extinction(a.shape, lc, sext, node_height(r_, a.top));
            float acc = 0.0f;
            for (int step = 0; step < a.nq; ++step) {
                ++nodes;
                double t;
                float ds;
                node_t(sa, sb, lower, a.gl_u[step], a.gl_w[step], t, ds);
                double r, qz, qx, qy;
                node_position(p, d, t, r, qz, qx, qy);
                const NodeGeom ng = node_geometry(qx, qy, qz, r, a.h, a.nh, a.chi, a.nc, a.top);
                const float h = ng.h, cx = ng.cx, sx = ng.sx, ah = ng.ah, ac = ng.ac, sn = ng.sn, cc = ng.cc;
                const int ir = ng.ir, j = ng.j;
                // rotation rows Q[k][j] (numpy_reference.py:327)
                const float Q[9] = {-cc * cx, sx, sn * cx, -cc * sx, -cx, sn * sx, sn, 0.0f, cc};
                float prof[5], stp;
                profiles(a.shape, h, prof, stp);
                float bnode = 0.0f;
#pragma unroll
                for (int k = 0; k < 5; ++k) bnode = bnode + prof[k] * lc[k];
                bnode = bnode + stp * sext;
                {
                    const double tm = 0.5 * (tprev + t);
                    double rm, a1, a2, a3;
                    node_position(p, d, tm, rm, a1, a2, a3);
                    const float bm = extinction(a.shape, lc, sext, node_height(rm, a.top));
                    const float w6 = (float)(t - tprev) / 6.0f;
                    acc = acc + w6 * ((bprev + bnode) + 4.0f * bm);
                    tprev = t;
                    bprev = bnode;
                }
                const float mid = (tau - tc) + acc;
                const float att = ds * exp_hw(-mid);
                const int nc = a.nc;
                const int cells[4] = {ir * nc + j, ir * nc + j + 1, (ir + 1) * nc + j, (ir + 1) * nc + j + 1};
                const float sw[4] = {(1.0f - ah) * (1.0f - ac), (1.0f - ah) * ac, ah * (1.0f - ac), ah * ac};
                // tensor rotation T[q][u] (numpy_reference.py:328-351)
                float T[P][4];
                {
                    float A[3][3], dt[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int jj = 0; jj < 3; ++jj)
                            A[i][jj] = (B[i] * Q[jj] + B[3 + i] * Q[3 + jj]) + B[6 + i] * Q[6 + jj];
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) dt[jj] = (df0 * Q[jj] + df1 * Q[3 + jj]) + df2 * Q[6 + jj];
                    float Kx[3][3];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int jj = 0; jj < 3; ++jj) Kx[i][jj] = A[i][jj] - dl[i] * dt[jj];
#pragma unroll
                    for (int q = 0; q < P; ++q) {
                        const int i = PairIdx<P>::i[q], k = PairIdx<P>::k[q];
                        T[q][0] = Kx[i][0] * Kx[k][0];
                        T[q][1] = Kx[i][1] * Kx[k][1];
                        T[q][2] = Kx[i][2] * Kx[k][2];
                        T[q][3] = Kx[i][0] * Kx[k][2] + Kx[i][2] * Kx[k][0];
                    }
                }
                float val[P];
#pragma unroll
                for (int q = 0; q < P; ++q) val[q] = 0.0f;
                if constexpr (FIRST) {
                    // sun columns (numpy_reference.py:431-458), FP64 near the terminator
                    double height = r - SOS_R_EARTH;
                    if (!(height > 0.0)) height = 0.0;
                    const double rad = SOS_R_EARTH + height;
                    const double mg = -sqrt(height * (2.0 * SOS_R_EARTH + height)) / rad;
                    double delta = qz / r - mg;
                    if (fabs(delta) < (double)SOS_FLT_MIN) delta = 0.0;
                    const bool lit = delta >= 0.0;
                    const float x = clampf((float)sqrt((lit ? delta : 0.0) / (1.0 - mg)), 0.0f, 1.0f);
                    const float hh = clampf(h, 0.0f, a.top);
                    const int si = interval(a.sun_h, a.sun_nh, hh);
                    const float al = (hh - a.sun_h[si]) / (a.sun_h[si + 1] - a.sun_h[si]);
                    const float fx = x * (float)(a.sun_nx - 1);
                    int sj = (int)fx;
                    sj = sj < 0 ? 0 : (sj > a.sun_nx - 2 ? a.sun_nx - 2 : sj);
                    const float bx = fx - (float)sj;
                    const float w00 = (1.0f - al) * (1.0f - bx), w10 = al * (1.0f - bx), w01 = (1.0f - al) * bx,
                                w11 = al * bx;
                    const float *c00 = a.sun_col + ((int64_t)si * a.sun_nx + sj) * 6;
                    const float *c10 = c00 + (int64_t)a.sun_nx * 6;
                    float optical = 0.0f;
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        float col = w00 * c00[k];
                        col = col + w10 * c10[k];
                        col = col + w01 * c00[6 + k];
                        col = col + w11 * c10[6 + k];
                        optical = optical + col * (k < 5 ? lc[k] : sext);
                    }
                    const float sun = lit ? exp_hw(-optical) : 0.0f;
                    // ground pools: exact paired HG phases (numpy_reference.py:526-557)
                    float mus = 0.0f, exd = 0.0f, eyd = 0.0f;
                    mus = (df0 * Q[2] + df1 * Q[5]) + df2 * Q[8];
                    exd = (df0 * Q[0] + df1 * Q[3]) + df2 * Q[6];
                    eyd = fabsf((df0 * Q[1] + df1 * Q[4]) + df2 * Q[7]);
                    float ground = 0.0f, sground = 0.0f;
                    const float sg = lc[SOS_LANE_SG];
                    float grec = 0.0f;  // scattered ground light of the four corners from the ground records
                    if (a.ground_records) {
                        const float *rec = a.ground_records + (node_base + nodes - 1) * 4 * G + w;
                        grec = ((sw[0] * rec[0] + sw[1] * rec[G]) + sw[2] * rec[2 * G]) + sw[3] * rec[3 * G];
                    } else {
    #pragma unroll 1
                        for (int c = 0; c < 4; ++c) {
                            const int64_t g0 = a.goff[cells[c]], g1 = a.goff[cells[c] + 1];
                            const float swc = sw[c];
                            if constexpr (UNI && G > 1) {
                                // the ray's lanes share g: lane j evaluates the phase of pool base + j once and the
                                // values reach every lane of the ray by a shuffle; sums keep the pool order
                                for (int64_t base = g0; base < g1; base += G) {
                                    const int64_t gi = base + w;
                                    float pt = 0.0f, ps = 0.0f;
                                    if (gi < g1) {
                                        const float gx = a.gdir[3 * gi], gy = a.gdir[3 * gi + 1], gz = a.gdir[3 * gi + 2];
                                        const float m0 = mus * gz + exd * gx;
                                        const float m1 = eyd * gy;
                                        const float up = clampf(m0 + m1, -1.0f, 1.0f), dn = clampf(m0 - m1, -1.0f, 1.0f);
                                        pt = (hg(up, gasy) + hg(dn, gasy)) / 2.0f;
                                        if (a.strato) ps = (hg(up, sg) + hg(dn, sg)) / 2.0f;
                                    }
                                    const int n = (int)min((int64_t)G, g1 - base);
                                    const float *gv = a.gval + base * G + w;
    #pragma unroll 4
                                    for (int jj = 0; jj < n; ++jj) {
                                        const float gvv = gv[jj * G];
                                        ground = ground + (swc * __shfl(pt, jj, G)) * gvv;
                                        if (a.strato) sground = sground + (swc * __shfl(ps, jj, G)) * gvv;
                                    }
                                }
                            } else {
                                for (int64_t gi = g0; gi < g1; ++gi) {
                                    const float gx = a.gdir[3 * gi], gy = a.gdir[3 * gi + 1], gz = a.gdir[3 * gi + 2];
                                    const float m0 = mus * gz + exd * gx;
                                    const float m1 = eyd * gy;
                                    const float up = clampf(m0 + m1, -1.0f, 1.0f), dn = clampf(m0 - m1, -1.0f, 1.0f);
                                    const float gvv = a.gval[gi * G + w];
                                    ground = ground + (swc * ((hg(up, gasy) + hg(dn, gasy)) / 2.0f)) * gvv;
                                    if (a.strato) sground = sground + (swc * ((hg(up, sg) + hg(dn, sg)) / 2.0f)) * gvv;
                                }
                            }
                        }

                    }
                    const float beam = (prof[1] * asca) * (sun * phase + ground);
                    const float ray_ = (sun * prof[0]) * lc[0];
                    const float sbeam = (stp * lc[SOS_LANE_SSCA]) * (sun * sphase + sground);
#pragma unroll
                    for (int q = 0; q < P; ++q) {
                        val[q] = ray_ * (mix * 0.119366207318921496f * direct[q] +
                                         (1.0f - mix) * 0.0397887357729738339f * pu[q]) +
                                 beam * pu[q] / 2.0f;
                        if (a.strato) val[q] = val[q] + (sbeam * pu[q]) / 2.0f;
                        if (a.ground_records) val[q] = val[q] + (grec * pu[q]) / 2.0f;
                    }
                    // Rayleigh scattering of the ground-reflected sun (gm), density at the node
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const float *g = a.gm + (int64_t)cells[c] * 4 * G + w;
                        const float rm[4] = {g[0], g[G], g[2 * G], g[3 * G]};
                        const float trace = (rm[0] + rm[1]) + rm[2];
#pragma unroll
                        for (int q = 0; q < P; ++q) {
                            float rot = 0.0f;
#pragma unroll
                            for (int uu = 0; uu < 4; ++uu) rot = rot + T[q][uu] * rm[uu];
                            const float tensor = prof[0] * (mix * rot + ((1.0f - mix) / 3.0f * trace) * pu[q]);
                            val[q] = val[q] + sw[c] * tensor;
                        }
                    }
                } else {
                    float aero[4];
                    if constexpr (AERO == RECORDS) {
                        // per (node, corner) SH values produced by the records kernel: [node][4][G]
                        const float *rec = a.records + (node_base + nodes - 1) * 4 * G + w;
#pragma unroll
                        for (int c = 0; c < 4; ++c) aero[c] = rec[c * G];
                    } else {
                        // adjacent-stencil reuse of cached corners (numpy_reference.py:106-129)
                        const int old = tags[0];
                        if (old >= 0) {
                            const int di = cells[0] / nc - old / nc, dj = cells[0] % nc - old % nc;
                            auto copy = [&](int dst, int src) {
    #pragma unroll
                                for (int m = 0; m <= L; ++m) cache[dst][m] = cache[src][m];
                                tags[dst] = cells[dst];
                            };
                            if (di == 0 && dj == 1) { copy(0, 1); copy(2, 3); }
                            else if (di == 0 && dj == -1) { copy(1, 0); copy(3, 2); }
                            else if (di == 1 && dj == 0) { copy(0, 2); copy(1, 3); }
                            else if (di == -1 && dj == 0) { copy(2, 0); copy(3, 1); }
                            else if (di == 1 && dj == 1) { copy(0, 3); }
                            else if (di == 1 && dj == -1) { copy(1, 2); }
                            else if (di == -1 && dj == 1) { copy(2, 1); }
                            else if (di == -1 && dj == -1) { copy(3, 0); }
                        }
    #pragma unroll 1
                        for (int c = 0; c < 4; ++c) {
                            const int cell = c == 0 ? cells[0] : c == 1 ? cells[1] : c == 2 ? cells[2] : cells[3];
                            const int tag = c == 0 ? tags[0] : c == 1 ? tags[1] : c == 2 ? tags[2] : tags[3];
                            if (tag != cell) {
                                ++refills;
                                float fresh[L + 1];
                                refill<L, G>(reinterpret_cast<const float *>(a.plane), cell, w, mu, st, fresh);
    #pragma unroll
                                for (int s = 0; s < 4; ++s)
                                    if (s == c) {
    #pragma unroll
                                        for (int m = 0; m <= L; ++m) cache[s][m] = fresh[m];
                                        tags[s] = cell;
                                    }
                            }
                        }
                        // cos(m phi) of this node (arithmetic.py:112-119) and the four azimuthal sums
                        const float cp = node_cp(st, df0, df1, cx, sx);
    #pragma unroll
                        for (int c = 0; c < 4; ++c) aero[c] = cache[c][0];
                        {
                            float cm0 = 1.0f, cm1 = cp;
    #pragma unroll
                            for (int c = 0; c < 4; ++c) aero[c] = aero[c] + cm1 * cache[c][1];
    #pragma unroll
                            for (int m = 2; m <= L; ++m) {
                                const float cm2 = (2.0f * cp) * cm1 - cm0;
    #pragma unroll
                                for (int c = 0; c < 4; ++c) aero[c] = aero[c] + cm2 * cache[c][m];
                                cm0 = cm1;
                                cm1 = cm2;
                            }
                        }

                    }
                    float logs = 0.0f, wg = 0.0f, am = 0.0f;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const float *f = a.field + (int64_t)cells[c] * a.rows * G + w;
                        const float rm[4] = {f[0], f[G], f[2 * G], f[3 * G]};
                        const float trace = (rm[0] + rm[1]) + rm[2];
                        const float density = a.hr[ir + (c >> 1)];
                        float tens[P];
#pragma unroll
                        for (int q = 0; q < P; ++q) {
                            float rot = 0.0f;
#pragma unroll
                            for (int uu = 0; uu < 4; ++uu) rot = rot + T[q][uu] * rm[uu];
                            tens[q] = density * (mix * rot + ((1.0f - mix) / 3.0f * trace) * pu[q]);
                            tens[q] = tens[q] + (aero[c] * pu[q]) / 2.0f;
                            val[q] = val[q] + sw[c] * tens[q];
                        }
                        const float tr = (tens[0] + tens[1]) + tens[2];
                        if (tr >= SOS_FLT_MIN) {
                            logs = logs + sw[c] * log2_hw(tr);
                            wg = wg + sw[c];
                            am = am + sw[c] * tr;
                        }
                    }
                    const bool ok = (wg > 0.0f) && (am >= SOS_FLT_MIN);
                    const float mean = ok ? am * __builtin_amdgcn_rcpf(wg) : 1.0f;
                    const float scale = exp2_hw(ok ? logs - wg * log2_hw(mean) : 0.0f);
#pragma unroll
                    for (int q = 0; q < P; ++q) val[q] = val[q] * scale;
                }
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float y = val[q] * att - corr[q];
                    const float tn = total[q] + y;
                    corr[q] = (tn - total[q]) - y;
                    total[q] = tn;
                }
            }
            // end piece of the segment, then the segment's optical depth into the ray total (Kahan)
            {
                const double tm = 0.5 * (tprev + sb);
                double rm, a1, a2, a3;
                node_position(p, d, tm, rm, a1, a2, a3);
                const float bm = extinction(a.shape, lc, sext, node_height(rm, a.top));
                node_position(p, d, sb, rm, a1, a2, a3);
                const float be = extinction(a.shape, lc, sext, node_height(rm, a.top));
                const float w6 = (float)(sb - tprev) / 6.0f;
                acc = acc + w6 * ((bprev + be) + 4.0f * bm);
                const float y = acc - tc;
                const float tn = tau + y;
                tc = (tn - tau) - y;
                tau = tn;
            }
        }
        if constexpr (BOUNDARY) {
            const float ef = a.end_f[ray];
            const int ej = a.end_j[ray];
            float E = (1.0f - ef) * a.boundary[ej * G + w] + ef * a.boundary[(ej + 1) * G + w];
            float optical = 0.0f;
#pragma unroll
            for (int k = 0; k < 5; ++k) optical = optical + a.end_col[6 * ray + k] * lc[k];
            optical = optical + a.end_col[6 * ray + 5] * sext;
            E = E * (exp_hw(-optical) * (a.hits[ray] ? 1.0f : 0.0f));
#pragma unroll
            for (int q = 0; q < P; ++q) total[q] = total[q] + (E * pu[q]) / 2.0f;
        }
        if (a.counters && w == 0) {
            atomicAdd(&a.counters[0], (unsigned long long)nodes);
            atomicAdd(&a.counters[1], (unsigned long long)refills);
        }
        float *o = a.out + (ray * G + w) * P;
#pragma unroll
        for (int q = 0; q < P; ++q) o[q] = total[q];
    }
}

}  // namespace sos

// ------------------------------------------------------------------------------------------------ host dispatch
namespace sos {
int set_error(const char *fmt, ...);

template <int G, int L, int P, bool FIRST, bool BOUNDARY, int AERO, bool UNI = false>
static int launch_one(const TransportArgs &a, hipStream_t s) {
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t wgs_needed = (a.tiles + 3) / 4;
    int64_t grid = (int64_t)cus * (a.blocks_per_cu > 0 ? a.blocks_per_cu : 8);
    if (grid > wgs_needed) grid = wgs_needed;
    grid = ((grid + 7) / 8) * 8;
    if (grid < 8) grid = 8;
    hipLaunchKernelGGL((transport_kernel<G, L, P, FIRST, BOUNDARY, AERO, UNI>), dim3((unsigned)grid), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("transport launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

template <int G, bool BOUNDARY, bool FIRST, bool TENSOR4 = false>
static int launch_geom(const TransportArgs &a, hipStream_t s) {
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t wgs_needed = (a.tiles + 3) / 4;
    int64_t grid = (int64_t)cus * (a.blocks_per_cu > 0 ? a.blocks_per_cu : 8);
    if (grid > wgs_needed) grid = wgs_needed;
    grid = ((grid + 7) / 8) * 8;
    if (grid < 8) grid = 8;
    if constexpr (FIRST) hipLaunchKernelGGL((transport_first_geom_kernel<G>), dim3((unsigned)grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((transport_geom_kernel<G, BOUNDARY, TENSOR4>), dim3((unsigned)grid), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("transport launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

// First light and the records consumer do not depend on L; the direct (register cache) path does.
template <int G, int P>
static int launch_flags(const TransportArgs &a, int L, hipStream_t s) {
    const bool first = a.flags & SOS_FIRST, bnd = a.flags & SOS_BOUNDARY;
    const bool records = a.flags & SOS_RECORDS;
    if (first && (a.flags & SOS_NODE_GEOMETRY)) {
        if constexpr (P == 4) return launch_geom<G, false, true>(a, s);
        else return set_error("node geometry: pairs must be 4");
    }
    if (first)
        return (a.flags & SOS_UNIFORM_PHASE) ? launch_one<G, 8, P, true, false, RECORDS, true>(a, s)
                                             : launch_one<G, 8, P, true, false, RECORDS, false>(a, s);
    if (records && (a.flags & SOS_NODE_GEOMETRY)) {
        if constexpr (P == 4) {
            if (a.flags & SOS_TENSOR4)
                return bnd ? launch_geom<G, true, false, true>(a, s) : launch_geom<G, false, false, true>(a, s);
            return bnd ? launch_geom<G, true, false>(a, s) : launch_geom<G, false, false>(a, s);
        } else {
            return set_error("node geometry: pairs must be 4");
        }
    }
    if (records)
        return bnd ? launch_one<G, 8, P, false, true, RECORDS>(a, s) : launch_one<G, 8, P, false, false, RECORDS>(a, s);
#define SOS_DIRECT(LL)                                                                                   \
    return bnd ? launch_one<G, LL, P, false, true, DIRECT>(a, s) : launch_one<G, LL, P, false, false, DIRECT>(a, s);
    switch (L) {
        case 8: SOS_DIRECT(8)
        case 16: SOS_DIRECT(16)
        case 24: SOS_DIRECT(24)
        case 32: SOS_DIRECT(32)
        default: return set_error("unsupported L %d (8, 16, 24, 32)", L);
    }
#undef SOS_DIRECT
}

template <int G>
static int launch_pairs(const TransportArgs &a, int L, hipStream_t s) {
    if (a.pairs == 4) return launch_flags<G, 4>(a, L, s);
    if (a.pairs == 6) return launch_flags<G, 6>(a, L, s);
    return set_error("pairs must be 4 or 6");
}

int transport_dispatch(const TransportArgs &a, int G, int L, hipStream_t s) {
    switch (G) {
        case 1: return launch_pairs<1>(a, L, s);
        case 16: return launch_pairs<16>(a, L, s);
        case 32: return launch_pairs<32>(a, L, s);
        default: return set_error("unsupported lane count %d (1, 16, 32)", G);
    }
}

int plan_count_launch(const TransportArgs &a, int64_t *counts, hipStream_t s) {
    const unsigned b = (unsigned)((a.rays + 255) / 256);
    if (b) hipLaunchKernelGGL(plan_count_kernel, dim3(b), dim3(256), 0, s, a, counts);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("plan_count: %s", hipGetErrorString(e));
}

int plan_nodes_launch(const TransportArgs &a, const int64_t *node_off, int32_t *cell, float *mu, float *cp,
                      float *dir, hipStream_t s) {
    const unsigned b = (unsigned)((a.rays + 255) / 256);
    if (b)
        hipLaunchKernelGGL(plan_nodes_kernel, dim3(b), dim3(256), 0, s, a, node_off, cell, mu, cp,
                           reinterpret_cast<float4 *>(dir));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("plan_nodes: %s", hipGetErrorString(e));
}

int plan_sun_launch(const TransportArgs &a, const int64_t *node_off, float *sun, hipStream_t s) {
    const unsigned b = (unsigned)((a.rays + 255) / 256);
    if (b) hipLaunchKernelGGL(plan_sun_kernel, dim3(b), dim3(256), 0, s, a, node_off, reinterpret_cast<float4 *>(sun));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("plan_sun: %s", hipGetErrorString(e));
}

int plan_geometry_launch(const TransportArgs &a, const int64_t *node_off, float *geom, hipStream_t s) {
    // the node geometry packs a base cell as 15 bits of ring and 16 bits of column
    if (a.nh > 32768 || a.nc > 65536) return set_error("plan_geometry: %d x %d cells exceed 32768 x 65536", a.nh, a.nc);
    const unsigned b = (unsigned)((a.rays + 255) / 256);
    if (b)
        hipLaunchKernelGGL(plan_geometry_kernel, dim3(b), dim3(256), 0, s, a, node_off, reinterpret_cast<float4 *>(geom));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("plan_geometry: %s", hipGetErrorString(e));
}

}  // namespace sos
