// Ray geometry and column integrals on the device (gfx950), FP64: the inputs of the transport that depend on the grid
// and the profile shapes only.
//
// boundaries   src/sosfuse2/rays/shell_segments.py:43-66 and solve/ray_geometry.py:28-45: path coordinates of the
//              shell exit or ground hit, the tangent point, the two crossings of the back half of the Earth's shadow
//              cylinder and the interior shells, sorted
// columns      shell_segments.py:115-147,223-227: slant columns of the five component profiles and the sulfate profile
//              along a ray, Gauss-Legendre nodes on every segment graded quadratically toward its lower end
// sun table    shell_segments.py:109-112,267-307: columns toward the sun on (height, x) with mu = mg + (1 - mg) x^2
// end columns  ray_geometry.py:53-65: columns from a ray's start to the ground for the rays that hit it
//
// One thread per ray; every sum runs in a fixed order, so results are bitwise repeatable.

#include "common.h"

// The FP64 arithmetic follows the host restatement (NumPy) operation by operation: no fused multiply-adds here.
#pragma clang fp contract(off)

namespace sos {
int set_error(const char *fmt, ...);

constexpr int MAX_EDGES = 13;  // 0, tangent, 2 shadow crossings, end, 2 x 4 interior shells

__device__ inline double profile64(const ProfileShape64 &s, int k, double h, double air) {
    const double a = s.a[k], b = s.b[k];
    switch (s.kind[k]) {
    case 1: return exp(-h / a);                                        // exponential
    case 2: { const double t = (h - b) / a; return exp(-0.5 * (t * t)); }  // Gaussian
    case 3: return fmax(0.0, 1.0 - fabs(h - b) / a);                   // triangle
    case 4: {                                                          // table of step a from height 0
        const double x = fmax(h, 0.0) * (1.0 / a);
        const int i = min((int)x, s.table_n - 2);
        const double f = x - (double)i;
        const double *t = s.table + k * s.table_n;
        return t[i] + f * (t[i + 1] - t[i]);
    }
    case 5: return (air * air) / (a * a);                              // (air / surface)^2
    default: return 0.0;
    }
}

__device__ inline void profiles64(const ProfileShape64 &s, double h, double *p) {
    h = fmax(h, 0.0);
    const double air = profile64(s, 0, h, 0.0);
    p[0] = air;
    for (int k = 1; k < 6; ++k) p[k] = profile64(s, k, h, air);
}

// Sorted boundaries of the ray p + t d (d a unit vector); returns their count (always 5 + 2 shells).
__device__ inline int ray_boundaries(const double *p, const double *d, const ColumnArgs &c, double *e) {
    const double R = SOS_R_EARTH;
    const double b = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
    const double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    double end = -b + sqrt(fmax(b * b - pp + (R + c.top) * (R + c.top), 0.0));
    const double disc = b * b - pp + R * R;
    const double hit = -b - sqrt(fmax(disc, 0.0));
    const bool ground = disc >= 0.0 && b < 0.0 && hit >= -1e-7;
    if (ground) end = fmax(hit, 0.0);
    end = fmax(end, 0.0);
    int n = 0;
    e[n++] = 0.0;
    e[n++] = (-b > 0.0 && -b < end) ? -b : end;
    const double aa = d[0] * d[0] + d[1] * d[1];
    const double bb = p[0] * d[0] + p[1] * d[1];
    const double cc = p[0] * p[0] + p[1] * p[1] - R * R;
    const double ds = bb * bb - aa * cc;
    for (int sign = -1; sign <= 1; sign += 2) {
        const double root = (-bb + sign * sqrt(fmax(ds, 0.0))) / fmax(aa, 1e-30);
        const bool valid = aa > 1e-20 && ds >= 0.0 && root > 0.0 && root < end && p[2] + root * d[2] < 0.0;
        e[n++] = valid ? root : end;
    }
    e[n++] = end;
    for (int s = 0; s < c.shells; ++s) {
        const double rs = R + c.shell_h[s];
        const double dd = b * b - pp + rs * rs;
        for (int sign = -1; sign <= 1; sign += 2) {
            const double rr = -b + sign * sqrt(fmax(dd, 0.0));
            e[n++] = (dd > 0.0 && rr > 0.0 && rr < end) ? rr : end;
        }
    }
    for (int i = 1; i < n; ++i) {  // insertion sort (at most 13 values)
        const double v = e[i];
        int j = i - 1;
        while (j >= 0 && e[j] > v) {
            e[j + 1] = e[j];
            --j;
        }
        e[j + 1] = v;
    }
    return n;
}

// Columns of the six profiles along p + t d over the sorted edges e[0 .. ne-1] (segments shorter than 1e-8 km skipped).
__device__ inline void ray_columns(const double *p, const double *d, const double *e, int ne, const ColumnArgs &c,
                                   double *out) {
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    for (int s = 0; s + 1 < ne; ++s) {
        const double a = e[s], b = e[s + 1];
        if (!(b - a > 1e-8)) continue;
        double ra = 0.0, rb = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double qa = p[i] + a * d[i], qb = p[i] + b * d[i];
            ra += qa * qa;
            rb += qb * qb;
        }
        const bool lower = ra <= rb;
        double seg[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < c.nq; ++j) {
            const double u = c.gl_u[j], wu = c.gl_w[j];
            const double t = lower ? a + (b - a) * u * u : b - (b - a) * (1.0 - u) * (1.0 - u);
            const double w = lower ? 2.0 * (b - a) * u * wu : 2.0 * (b - a) * (1.0 - u) * wu;
            double q2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                const double q = p[i] + t * d[i];
                q2 += q * q;
            }
            double prof[6];
            profiles64(c.shape, sqrt(q2) - SOS_R_EARTH, prof);
            for (int k = 0; k < 6; ++k) seg[k] += prof[k] * w;
        }
        for (int k = 0; k < 6; ++k) out[k] += seg[k];
    }
}

// Sun table: thread (i, j) takes height h[i] and x[j]; the ray starts at (0, 0, R + h) with mu = mg + (1 - mg) x^2
// (mg + 1e-9 at x = 0), direction (sqrt(1 - mu^2), 0, mu).
__global__ void sun_table_kernel(int nh, int nx, const double *h, const double *x, ColumnArgs c, double *out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)nh * nx) return;
    const int i = (int)(t / nx), j = (int)(t % nx);
    const double r = SOS_R_EARTH + h[i];
    const double q = SOS_R_EARTH / r;
    const double mg = -sqrt(fmax(0.0, 1.0 - q * q));
    double mu = mg + (1.0 - mg) * x[j] * x[j];
    if (j == 0) mu = mg + 1e-9;
    const double p[3] = {0.0, 0.0, r};
    const double d[3] = {sqrt(fmax(1.0 - mu * mu, 0.0)), 0.0, mu};
    double e[MAX_EDGES];
    const int ne = ray_boundaries(p, d, c, e);
    ray_columns(p, d, e, ne, c, out + t * 6);
}

// Columns from each ray's start along its direction to its end, for the rays with hit[ray] set (others 0). start of
// ray i: starts[group[i]] (group = nullptr: starts[i]).
__global__ void ray_columns_kernel(int64_t n, const double *starts, const int32_t *group, const double *dirs,
                                   const uint8_t *hit, ColumnArgs c, double *out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *o = out + i * 6;
    if (hit && !hit[i]) {
        for (int k = 0; k < 6; ++k) o[k] = 0.0;
        return;
    }
    const double *p = starts + 3 * (group ? (int64_t)group[i] : i);
    const double *d = dirs + 3 * i;
    double e[MAX_EDGES];
    const int ne = ray_boundaries(p, d, c, e);
    ray_columns(p, d, e, ne, c, o);
}

// ------------------------------------------------------------------------------- incident quadrature and packs
// (src/sosfuse2/rays/incoming_rings.py:11-126, solve/ray_geometry.py:28-82,148-180; the host restatement is
// geometry.incoming, geometry.pack and geometry.compact_edges)

constexpr int MAX_MU_EDGES = 32;  // 8 fixed + 2 shadow + 3 tangent-pair + 2 per ring level (at most 7 levels)


__device__ inline double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

// mu edges of azimuth phi (_mu_edges and ring_edges of the host restatement), appended to e; returns the count.
__device__ int mu_edges(const IncidentArgs &a, double r, double chi, double phi, double mg, double *e) {
    const double R = SOS_R_EARTH;
    const double sinchi = sin(chi), coschi = cos(chi);
    const double cphi = cos(phi), sphi = sin(phi);
    int n = 0;
    e[n++] = -1.0; e[n++] = 1.0; e[n++] = mg; e[n++] = 0.0;
    e[n++] = mg - 0.05234; e[n++] = mg + 0.05234; e[n++] = mg - 0.0061; e[n++] = mg + 0.0061;
    const double rp = r * sinchi;
    if (rp > R + 1e-8 && fabs(sinchi) > 1e-8) {
        const double kk = R / sqrt(rp * rp - R * R);
        for (int sign = -1; sign <= 1; sign += 2) {
            const double t = cphi * coschi / sinchi + sign * sphi / (kk * sinchi);
            const double mu = t / sqrt(1.0 + t * t);
            const double st = sqrt(1.0 - mu * mu);
            const double dx = mu * sinchi - st * cphi * coschi;
            const double dy = -st * sphi;
            const double dz = mu * coschi + st * cphi * sinchi;
            const double tmin = -rp * dx / fmax(dx * dx + dy * dy, 1e-25);
            if (fabs(dx) + fabs(dy) < 1e-10 || (tmin > 0.0 && r * coschi + tmin * dz < 0.0)) e[n++] = mu;
        }
    }
    if (phi < SOS_PI / 6.0) {
        const double mu = coschi / sqrt(coschi * coschi + cphi * cphi * sinchi * sinchi + 0.0);
        e[n++] = mu; e[n++] = mu - 0.012; e[n++] = mu + 0.012;
    }
    // ring edges: intersections of the view plane at azimuth phi with each ring level of the shadow
    const double up[3] = {sinchi, 0.0, coschi};
    const double ex[3] = {-coschi, 0.0, sinchi};
    const double ey[3] = {up[1] * ex[2] - up[2] * ex[1], up[2] * ex[0] - up[0] * ex[2], up[0] * ex[1] - up[1] * ex[0]};
    const double p[3] = {r * up[0], r * up[1], r * up[2]};
    double hz[3];
    for (int i = 0; i < 3; ++i) hz[i] = cphi * ex[i] + sphi * ey[i];
    const double nrm[3] = {up[1] * hz[2] - up[2] * hz[1], up[2] * hz[0] - up[0] * hz[2], up[0] * hz[1] - up[1] * hz[0]};
    const double amp = R * hypot(nrm[0], nrm[1]);
    if (amp < 1e-10) return n;
    const double angle = atan2(nrm[1], nrm[0]);
    for (int lv = 0; lv < a.nlevels; ++lv) {
        const double z = a.levels[lv];
        const double target = -nrm[2] * z / amp;
        if (fabs(target) > 1.0) continue;
        const double delta = acos(target);
        for (int side = 0; side < 2; ++side) {
            const double az = side ? angle + delta : angle - delta;
            double dv[3] = {R * cos(az) - p[0], R * sin(az) - p[1], z - p[2]};
            const double length = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (length < 1e-8) continue;
            for (int i = 0; i < 3; ++i) dv[i] /= length;
            if (dv[0] * hz[0] + dv[1] * hz[1] + dv[2] * hz[2] <= 1e-10) continue;
            const double b = p[0] * dv[0] + p[1] * dv[1] + p[2] * dv[2];
            const double disc = b * b - r * r + R * R;
            const double hit = -b - sqrt(fmax(0.0, disc));
            if (disc >= 0.0 && b < 0.0 && hit >= -1e-7 && hit < length - 1e-5) continue;
            e[n++] = dv[0] * up[0] + dv[1] * up[1] + dv[2] * up[2];
        }
    }
    return n;
}

// The sinh-graded intervals of the edges (_sinh_nodes): clip to [-1, 1], sort, drop exact duplicates, map to
// t = sign(d) asinh(|d| / den sa) / 4, keep intervals with dt >= 1e-10. Returns the kept count; their t bounds in
// tl, tr.
__device__ int sinh_intervals(double *e, int n, double mg, double *tl, double *tr) {
    const double A = 4.0, sa = sinh(A);
    for (int i = 0; i < n; ++i) e[i] = fmin(fmax(e[i], -1.0), 1.0);
    for (int i = 1; i < n; ++i) {
        const double v = e[i];
        int j = i - 1;
        while (j >= 0 && e[j] > v) {
            e[j + 1] = e[j];
            --j;
        }
        e[j + 1] = v;
    }
    int u = 0;
    for (int i = 0; i < n; ++i)
        if (i == 0 || e[i] != e[u - 1]) e[u++] = e[i];
    int kept = 0;
    double prev = 0.0;
    for (int i = 0; i < u; ++i) {
        const double d = e[i] - mg;
        const double den = d < 0.0 ? 1.0 + mg : 1.0 - mg;
        const double t = sgn(d) * asinh(fabs(d) / den * sa) / A;
        if (i > 0 && t - prev >= 1e-10) {
            tl[kept] = prev;
            tr[kept] = t;
            ++kept;
        }
        prev = t;
    }
    return kept;
}

// Task (cell, j): the nodes of azimuth j of the cell's incident quadrature. count_only: write the node count;
// otherwise write directions (global frame), weights and the cell of every node from offset off[task].
__global__ void incident_kernel(IncidentArgs a, int64_t *counts, const int64_t *off, double *dirs, double *weights,
                                int32_t *ray_cell) {
    const int64_t task = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int per = a.nphi + 1;
    if (task >= (int64_t)a.nh * a.nc * per) return;
    const int cell = (int)(task / per), j = (int)(task % per);
    const int ih = cell / a.nc, ic = cell % a.nc;
    const double R = SOS_R_EARTH;
    const double r = R + a.h[ih], chi = a.chi[ic];
    const double q = R / r;
    const double mg = -sqrt(fmax(0.0, 1.0 - q * q));
    const double phi = j == a.nphi ? SOS_PI : (double)j * (SOS_PI / (double)a.nphi);
    double e[MAX_MU_EDGES], tl[MAX_MU_EDGES], tr[MAX_MU_EDGES];
    const int n = mu_edges(a, r, chi, phi, mg, e);
    const int kept = sinh_intervals(e, n, mg, tl, tr);
    if (counts) {
        counts[task] = (int64_t)kept * a.nmu;
        return;
    }
    const double A = 4.0, sa = sinh(A);
    const double dphi = 2.0 * SOS_PI / (double)a.nphi, half = (j == 0 || j == a.nphi) ? 0.5 : 1.0;
    const double sinchi = sin(chi), coschi = cos(chi);
    const double up[3] = {sinchi, 0.0, coschi};
    const double ex[3] = {-coschi, 0.0, sinchi};
    const double ey[3] = {up[1] * ex[2] - up[2] * ex[1], up[2] * ex[0] - up[0] * ex[2], up[0] * ex[1] - up[1] * ex[0]};
    const double cphi = cos(phi), sphi = sin(phi);
    int64_t o = off[task];
    for (int k = 0; k < kept; ++k) {
        for (int g = 0; g < a.nmu; ++g, ++o) {
            const double t = tl[k] + (tr[k] - tl[k]) * a.gl_u[g];
            const double den = t < 0.0 ? 1.0 + mg : 1.0 - mg;
            const double mu = mg + sgn(t) * den * sinh(A * fabs(t)) / sa;
            const double w = (tr[k] - tl[k]) * a.gl_w[g] * den * A * cosh(A * fabs(t)) / sa;
            const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
            const double loc[3] = {st * cphi, st * sphi, mu};
            for (int i = 0; i < 3; ++i) dirs[3 * o + i] = loc[0] * ex[i] + loc[1] * ey[i] + loc[2] * up[i];
            weights[o] = w * dphi * half;
            ray_cell[o] = cell;
        }
    }
}


// Per ray (pack + compact_edges): boundaries, lower flags, the kept intervals (strictly positive widths), end chi
// index and fraction, ground hit, direction in the target frame. count_only: the compacted edge count and the
// sampled intervals (width > 1e-8); otherwise the compacted edges from edge_off[ray] and the lower bits.
__global__ void pack_kernel(PackArgs a, int64_t rays, const int32_t *ray_cell, const double *dirs, int64_t *edge_count,
                            int64_t *sampled, const int64_t *edge_off, double *edges, int16_t *lower, int32_t *end_j,
                            float *end_f, uint8_t *hits, float *dloc) {
    const int64_t ray = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ray >= rays) return;
    const int cell = ray_cell[ray];
    const double *p = a.starts + 3 * (int64_t)cell;
    const double *d = dirs + 3 * ray;
    double e[MAX_EDGES];
    const int ne = ray_boundaries(p, d, a.c, e);
    int count = 0, samp = 0, last = -1;
    for (int s = 0; s + 1 < ne; ++s) {
        if (e[s + 1] - e[s] > 0.0) last = s;
        if (e[s + 1] - e[s] > 1e-8) ++samp;
    }
    for (int s = 0; s + 1 < ne; ++s) count += e[s + 1] - e[s] > 0.0;
    const int total = last < 0 ? 1 : count + 1;
    if (edge_count) {
        edge_count[ray] = total;
        sampled[ray] = samp;
        return;
    }
    int64_t o = edge_off[ray];
    unsigned bits = 0;
    if (last < 0) {
        edges[o] = 0.0;
    } else {
        int rank = 0;
        for (int s = 0; s + 1 < ne; ++s) {
            if (!(e[s + 1] - e[s] > 0.0)) continue;
            edges[o++] = e[s];
            double ra = 0.0, rb = 0.0;
            for (int i = 0; i < 3; ++i) {
                const double qa = p[i] + e[s] * d[i], qb = p[i] + e[s + 1] * d[i];
                ra += qa * qa;
                rb += qb * qb;
            }
            if (ra <= rb) bits |= 1u << rank;
            ++rank;
        }
        edges[o] = e[last + 1];
    }
    lower[ray] = (int16_t)(uint16_t)bits;
    const double end = e[ne - 1];
    double qe[3], re2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        qe[i] = p[i] + end * d[i];
        re2 += qe[i] * qe[i];
    }
    const double re = sqrt(re2);
    const double ch = acos(fmin(fmax(qe[2] / re, -1.0), 1.0));
    int lo = 0, hi = a.nc;  // searchsorted(chi, ch, 'right') - 1, clipped to [0, nc - 2]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a.chi[mid] <= ch) lo = mid + 1; else hi = mid;
    }
    const int j = min(max(lo - 1, 0), a.nc - 2);
    end_j[ray] = j;
    end_f[ray] = (float)fmin(fmax((ch - a.chi[j]) / (a.chi[j + 1] - a.chi[j]), 0.0), 1.0);
    hits[ray] = re < SOS_R_EARTH + 1e-5;
    const double *B = a.targets + 9 * (int64_t)cell;
    for (int k = 0; k < 3; ++k) dloc[3 * ray + k] = (float)(B[0 * 3 + k] * d[0] + B[1 * 3 + k] * d[1] + B[2 * 3 + k] * d[2]);
}

int incident_launch(const IncidentArgs &a, int64_t *counts, const int64_t *off, double *dirs, double *weights,
                    int32_t *ray_cell, hipStream_t s) {
    const int64_t n = (int64_t)a.nh * a.nc * (a.nphi + 1);
    if (n > 0) hipLaunchKernelGGL(incident_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, a, counts, off, dirs,
                                  weights, ray_cell);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("incident: %s", hipGetErrorString(e));
}

int pack_launch(const PackArgs &a, int64_t rays, const int32_t *ray_cell, const double *dirs, int64_t *edge_count,
                int64_t *sampled, const int64_t *edge_off, double *edges, int16_t *lower, int32_t *end_j, float *end_f,
                uint8_t *hits, float *dloc, hipStream_t s) {
    if (rays > 0) hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((rays + 127) / 128)), dim3(128), 0, s, a, rays,
                                     ray_cell, dirs, edge_count, sampled, edge_off, edges, lower, end_j, end_f, hits,
                                     dloc);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("pack: %s", hipGetErrorString(e));
}

int sun_table_launch(int nh, int nx, const double *h, const double *x, const ColumnArgs &c, double *out,
                     hipStream_t s) {
    const int64_t n = (int64_t)nh * nx;
    if (n > 0) hipLaunchKernelGGL(sun_table_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, nh, nx, h, x, c,
                                  out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("sun_table: %s", hipGetErrorString(e));
}

int ray_columns_launch(int64_t n, const double *starts, const int32_t *group, const double *dirs, const uint8_t *hit,
                       const ColumnArgs &c, double *out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(ray_columns_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, n, starts,
                                  group, dirs, hit, c, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("ray_columns: %s", hipGetErrorString(e));
}

}  // namespace sos
