// Source composition, the combined SH plane, convergence metrics and the Anderson update on the device (gfx950).
//
// compose      src/sosfuse2/solve/source_iteration.py:176-212 (eigenvalue scaling, FP32 flush of |s| < FLT_MIN)
// make_plane   src/sosfuse2/backends/sh_layout.py:135-152 (both aerosol planes weighted by their densities),
//              stored m-major per cell so the transport's recurrence reads it in order
// metrics      source_iteration.py:52-97 (non-finite count, max |g|, max |g - x| per lane; integer atomics only)
// anderson_*   source_iteration.py:432-443 (scaled residual), :618-635 (Gram products and mixing), :25-33 (FP64
//              regularised solve and guard). All floating sums have a fixed order.

#include <mutex>

#include "common.h"

namespace sos {
int set_error(const char *fmt, ...);

static int check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("%s: %s", what, hipGetErrorString(e));
    return SOS_OK;
}

static unsigned blocks_for(int64_t n, int threads = 256, int64_t cap = 65535 * 8) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// log2 of a power-of-two lane count, -1 otherwise
static int lane_shift(int G) {
    int gs = 0;
    while ((1 << gs) < G) ++gs;
    return (1 << gs) == G ? gs : -1;
}

// ----------------------------------------------------------------------------------------------------- compose
// One workgroup per cell; G is a power of two (1, 16 or 32), so a row and lane come from a shift and a mask.
__global__ __launch_bounds__(256) void compose_kernel(int rows, int K, int G, int gs, const float *__restrict__ mom,
                                                      const float *__restrict__ first, const float *__restrict__ factor,
                                                      float *__restrict__ out) {
    const int64_t cell = blockIdx.x;
    const float *mc = mom + cell * (4 + K) * G;
    const float *fc = first ? first + cell * (4 + K) * G : nullptr;
    float *oc = out + cell * (int64_t)rows * G;
    const int n = rows * G;
    for (int idx = threadIdx.x; idx < n; idx += 256) {
        const int row = idx >> gs, w = idx & (G - 1);
        int mr = row;
        if (row >= 10) {
            mr = row - 10;
            while (mr >= K) mr -= K;
            mr += 4;
        }
        float m = mc[mr * G + w];
        if (fc) m = m + fc[mr * G + w];
        float v = factor[row * G + w] * m;
        if (fabsf(v) < SOS_FLT_MIN) v = 0.0f;
        oc[idx] = v;
    }
}

// ------------------------------------------------------------------------------------------------- make_plane
// Combined plane c[cell][j][w] = src[10 + k] ha(h) (+ src[10 + K + k] sd(h)), j the m-major position of k = (l, m).
__device__ inline int plane_k(int L, int j) {
    int m = 0, off = 0;
    while (off + (L + 1 - m) <= j) { off += L + 1 - m; ++m; }
    const int l = m + (j - off);
    return l * (l + 1) / 2 + m;
}

// One workgroup per cell: the l-major index of every m-major position in LDS, the cell's profile weights once.
__global__ __launch_bounds__(256) void plane_kernel(int nc, int L, int rows, int G, int gs, int strato,
                                                    const float *__restrict__ src, const float *__restrict__ ha,
                                                    const float *__restrict__ sd, float *__restrict__ plane) {
    const int K = (L + 1) * (L + 2) / 2;
    __shared__ short kl[(32 + 1) * (32 + 2) / 2];
    for (int j = threadIdx.x; j < K; j += 256) kl[j] = (short)plane_k(L, j);
    __syncthreads();
    const int64_t cell = blockIdx.x;
    const int ih = (int)(cell / nc);
    const float hav = ha[ih], sdv = strato ? sd[ih] : 0.0f;
    const float *sc = src + cell * rows * G;
    float *pc = plane + cell * (int64_t)K * G;
    const int n = K * G;
    for (int idx = threadIdx.x; idx < n; idx += 256) {
        const int j = idx >> gs, w = idx & (G - 1);
        const int k = kl[j];
        float v = sc[(10 + k) * G + w] * hav;
        if (strato) v = v + sc[(10 + K + k) * G + w] * sdv;
        pc[idx] = v;
    }
}

// Split FP16 plane for the mixed-precision contraction: one workgroup per cell. scale[cell][w] is a power of two
// with max_j |c[cell][j][w]| / scale in [2^14, 2^15) (1 for an all-zero column); x = c / scale is stored as
// hi = fp16(x) in the first plane and lo = fp16((x - hi) 2^11) in the second (cells x KZ x G further on), so
// hi + lo 2^-11 carries x to 2^-23 relative, neither part leaves FP16's range, and every entry down to 2^-28 of the
// column's largest is an FP16 normal number (the coefficients of a column span many decades; MFMA products of
// subnormal FP16 operands are not exact). Layout of each plane: quads of four
// consecutive K positions per lane, [cell][KZ / 4][G][4] with KZ = K rounded up to 16 and zeros beyond K, so that
// an MFMA B operand (four K positions of one lane) is one 8-byte load.
__global__ __launch_bounds__(256) void plane_half_kernel(int nc, int L, int rows, int G, int strato,
                                                         const float *__restrict__ src, const float *__restrict__ ha,
                                                         const float *__restrict__ sd, _Float16 *__restrict__ plane,
                                                         float *__restrict__ scale) {
    constexpr int KMAX = 33 * 34 / 2;
    __shared__ short kmap[KMAX];
    __shared__ float colmax[256];
    __shared__ float inv[32];
    const int K = (L + 1) * (L + 2) / 2;
    const int KZ = (K + 15) / 16 * 16;
    const int64_t cell = blockIdx.x;
    const int ih = (int)(cell / nc);
    const int tid = threadIdx.x, w = tid % G, q = tid / G, nq = 256 / G;
    for (int j = tid; j < K; j += 256) kmap[j] = (short)plane_k(L, j);
    __syncthreads();
    const float *s = src + cell * rows * G + w;
    const float h = ha[ih], d = strato ? sd[ih] : 0.0f;
    auto value = [&](int j) {
        const int k = kmap[j];
        float v = s[(10 + k) * G] * h;
        if (strato) v = v + s[(10 + K + k) * G] * d;
        return v;
    };
    float mx = 0.0f;
    for (int j = q; j < K; j += nq) mx = fmaxf(mx, fabsf(value(j)));
    colmax[tid] = mx;
    __syncthreads();
    if (tid < G) {
        for (int r = 1; r < nq; ++r) mx = fmaxf(mx, colmax[r * G + tid]);
        int ex = 0;
        frexpf(mx, &ex);  // mx = f 2^ex with f in [0.5, 1)
        // largest entry in [2^14, 2^15): FP16 keeps entries down to 2^-28 of it normal (MFMA FP16 products of
        // subnormal operands are not exact); the scale itself stays an FP32 normal power of two
        const float sc = mx > 0.0f ? ldexpf(1.0f, max(ex - 15, -126)) : 1.0f;
        scale[cell * G + tid] = sc;
        inv[tid] = 1.0f / sc;  // exact: a power of two
    }
    __syncthreads();
    const float iv = inv[w];
    _Float16 *hi = plane + cell * KZ * G;
    _Float16 *lo = hi + (int64_t)gridDim.x * KZ * G;
    for (int j = q; j < KZ; j += nq) {
        const float x = j < K ? value(j) * iv : 0.0f;
        const _Float16 hv = (_Float16)x;
        const int64_t o = ((int64_t)(j >> 2) * G + w) * 4 + (j & 3);
        hi[o] = hv;
        lo[o] = (_Float16)((x - (float)hv) * 2048.0f);
    }
}

// ---------------------------------------------------------------------------------------------------- metrics
__global__ void zero_u32(unsigned *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void metrics_kernel(int64_t rows, int G, const float *__restrict__ g, const float *__restrict__ x,
                               unsigned frozen, unsigned *__restrict__ acc) {
    // acc: [3][G] = count, max |value| bits, max |value - x| bits. The grid stride is a multiple of G, so every
    // thread sees one lane; it reduces in registers and publishes once (max and integer sums are order-free).
    __shared__ unsigned s[3][32];
    if (threadIdx.x < 3 * 32) s[threadIdx.x / 32][threadIdx.x % 32] = 0;
    __syncthreads();
    const int64_t total = rows * G;
    const int64_t start = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int w = (int)(start % G);
    unsigned bad = 0, mag = 0, dif = 0;
    for (int64_t e = start; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const float cur = g[e], old = x[e];
        const float val = (frozen >> w) & 1u ? old : cur;
        bad += !isfinite(cur);
        mag = max(mag, __float_as_uint(fabsf(val)));
        dif = max(dif, __float_as_uint(fabsf(val - old)));
    }
    atomicAdd(&s[0][w], bad);
    atomicMax(&s[1][w], mag);
    atomicMax(&s[2][w], dif);
    __syncthreads();
    if (threadIdx.x < G) {
        atomicAdd(&acc[threadIdx.x], s[0][threadIdx.x]);
        atomicMax(&acc[G + threadIdx.x], s[1][threadIdx.x]);
        atomicMax(&acc[2 * G + threadIdx.x], s[2][threadIdx.x]);
    }
}

__global__ void metrics_finish(int G, const unsigned *acc, float *out) {
    int w = threadIdx.x;
    if (w < G) {
        out[w] = (float)acc[w];
        out[G + w] = __uint_as_float(acc[G + w]);
        out[2 * G + w] = __uint_as_float(acc[2 * G + w]);
    }
}

// --------------------------------------------------------------------------------------------------- anderson
// One workgroup per cell: thread t sums the squares of rows t / G, t / G + 256 / G, ... of lane t mod G in FP64, and
// the row groups are added in a fixed order.
__global__ __launch_bounds__(256) void norms_kernel(int rows, int G, int gs, const float *__restrict__ g,
                                                    float *__restrict__ norms, unsigned *__restrict__ maxbits) {
    __shared__ double part[256];
    const int64_t cell = blockIdx.x;
    const int t = threadIdx.x, w = t & (G - 1), rg = t >> gs, groups = 256 >> gs;
    const float *p = g + cell * rows * G + w;
    double s = 0.0;
    for (int r = rg; r < rows; r += groups) {
        const double v = p[r * G];
        s += v * v;
    }
    part[t] = s;
    __syncthreads();
    if (t < G) {
        double tot = 0.0;
        for (int q = 0; q < groups; ++q) tot += part[(q << gs) + t];
        const float nrm = (float)sqrt(tot);
        norms[cell * G + t] = nrm;
        atomicMax(&maxbits[t], __float_as_uint(nrm));
    }
}

// f[w][i] = (g - x)[i][w] / max(norm[cell(i)][w], 1e-8 max_cell norm[w]) for the source rows i < cells * rows and
// f[w][nsrc + o] = obs[w][o]: the field is row-major over lanes and the residual lane-major, so each block moves a
// tile of 64 rows through LDS (coalesced reads of g and x, coalesced writes of every lane's row).
__global__ __launch_bounds__(256) void residual_kernel(int64_t cells, int rows, int G, const float *__restrict__ g,
                                                       const float *__restrict__ x, const float *__restrict__ norms,
                                                       const unsigned *__restrict__ maxbits,
                                                       const float *__restrict__ obs, int nobs,
                                                       float *__restrict__ f) {
    constexpr int R = 64;
    __shared__ float tile[32][R + 1];
    __shared__ float floor_[32];
    const int64_t nsrc = cells * rows;
    const int64_t n = nsrc + nobs;
    if (threadIdx.x < G) floor_[threadIdx.x] = __uint_as_float(maxbits[threadIdx.x]) * 1e-8f;
    __syncthreads();
    const int64_t tiles = (nsrc + R - 1) / R;
    __shared__ int64_t tcell[R];
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t r0 = t * R;
        if (threadIdx.x < R) tcell[threadIdx.x] = (r0 + threadIdx.x) / rows;  // the cell of each row of the tile
        __syncthreads();
        for (int e = threadIdx.x; e < R * G; e += blockDim.x) {
            const int r = e / G, w = e % G;
            const int64_t i = r0 + r;
            float v = 0.0f;
            if (i < nsrc) {
                const float scale = fmaxf(norms[tcell[r] * G + w], floor_[w]);
                v = (g[i * G + w] - x[i * G + w]) / scale;
            }
            tile[w][r] = v;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < R * G; e += blockDim.x) {
            const int w = e / R, r = e % R;
            const int64_t i = r0 + r;
            if (i < nsrc) f[(int64_t)w * n + i] = tile[w][r];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0)
        for (int e = threadIdx.x; e < G * nobs; e += blockDim.x) {
            const int w = e / nobs, o = e % nobs;
            f[(int64_t)w * n + nsrc + o] = obs[(int64_t)w * nobs + o];
        }
}

// Block partial dot products: part[b][w][j] = sum over chunk b of f[slot][w][i] f[j][w][i]. Block (b, w) streams
// chunk b of lane w's residuals (contiguous in the lane-major ring): each thread reads f[slot] once per element and
// every filled slot beside it, keeps the depth sums in registers (FP64), and the block reduces them once.
__global__ __launch_bounds__(256) void gram_partial(int64_t n, int G, int depth, int slot, unsigned filled,
                                                    const float *__restrict__ ring, double *__restrict__ part) {
    constexpr int DMAX = 8;
    __shared__ double red[4][DMAX];
    const int w = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = blockIdx.x * per, i1 = min(n, i0 + per);
    const float *fs = ring + ((int64_t)slot * G + w) * n;
    double s[DMAX];
#pragma unroll
    for (int j = 0; j < DMAX; ++j) s[j] = 0.0;
    // four strided elements per step, so every thread keeps up to 4 (depth + 1) loads in flight
    int64_t i = i0 + tid;
    for (; i + 768 < i1; i += 1024) {
        const double a0 = fs[i], a1 = fs[i + 256], a2 = fs[i + 512], a3 = fs[i + 768];
#pragma unroll
        for (int j = 0; j < DMAX; ++j)
            if (j < depth && ((filled >> j) & 1u)) {
                const float *fj = ring + ((int64_t)j * G + w) * n + i;
                const float b0 = fj[0], b1 = fj[256], b2 = fj[512], b3 = fj[768];
                s[j] += a0 * (double)b0;
                s[j] += a1 * (double)b1;
                s[j] += a2 * (double)b2;
                s[j] += a3 * (double)b3;
            }
    }
    for (; i < i1; i += 256) {
        const double a = fs[i];
#pragma unroll
        for (int j = 0; j < DMAX; ++j)
            if (j < depth && ((filled >> j) & 1u)) s[j] += a * (double)ring[((int64_t)j * G + w) * n + i];
    }
#pragma unroll
    for (int j = 0; j < DMAX; ++j) {
        double v = s[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (tid < depth) part[((int64_t)blockIdx.x * G + w) * depth + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

__global__ void gram_finish(int blocks, int G, int depth, int slot, unsigned filled, const double *__restrict__ part,
                            double *__restrict__ gram) {
    const int t = threadIdx.x;
    if (t >= G * depth) return;
    const int w = t / depth, j = t % depth;
    if (!((filled >> j) & 1u)) return;
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += part[((int64_t)b * G + w) * depth + j];
    double *gm = gram + (int64_t)w * depth * depth;
    gm[slot * depth + j] = s;
    gm[j * depth + slot] = s;
}

__global__ void solve_kernel(int G, int depth, int m, const int *__restrict__ order, const double *__restrict__ gram,
                             unsigned frozen, float *__restrict__ wout) {
    const int w = threadIdx.x;
    if (w >= G) return;
    float *wo = wout + w * depth;
    if ((frozen >> w) & 1u) { wo[0] = __int_as_float(0x7fc00000); return; }
    double A[8][9];
    const double *gm = gram + (int64_t)w * depth * depth;
    double tr = 0.0;
    for (int i = 0; i < m; ++i) tr += gm[order[i] * depth + order[i]];
    const double reg = fmax(tr * 1e-10, 1e-30);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) A[i][j] = gm[order[i] * depth + order[j]] + (i == j ? reg : 0.0);
        A[i][m] = 1.0;
    }
    bool ok = true;
    for (int c = 0; c < m; ++c) {  // Gaussian elimination, partial pivoting
        int piv = c;
        for (int r = c + 1; r < m; ++r) if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (A[piv][c] == 0.0) { ok = false; break; }
        if (piv != c) for (int k = 0; k <= m; ++k) { double t = A[c][k]; A[c][k] = A[piv][k]; A[piv][k] = t; }
        for (int r = c + 1; r < m; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k <= m; ++k) A[r][k] -= f * A[c][k];
        }
    }
    double x[8];
    if (ok) {
        for (int r = m - 1; r >= 0; --r) {
            double s = A[r][m];
            for (int k = r + 1; k < m; ++k) s -= A[r][k] * x[k];
            x[r] = s / A[r][r];
        }
        double sum = 0.0;
        for (int i = 0; i < m; ++i) sum += x[i];
        double mx = 0.0;
        for (int i = 0; i < m; ++i) {
            x[i] /= sum;
            if (!isfinite(x[i])) ok = false;
            mx = fmax(mx, fabs(x[i]));
        }
        if (!(mx < 8.0)) ok = false;
    }
    if (!ok) { wo[0] = __int_as_float(0x7fc00000); return; }
    for (int i = 0; i < m; ++i) wo[i] = (float)x[i];
}

// x = sum_j ring[order[j]] wts[lane][j] in float4 groups of four lanes (G is a multiple of 4, so a group never
// straddles a row); a lane whose weights are NaN (frozen) keeps x. The grid stride is a multiple of G / 4, so every
// thread keeps the same four lanes and reads their weights and the slot order once.
__global__ __launch_bounds__(256) void mix_kernel(int64_t elems, int G, int depth, int m, const float *__restrict__ ring,
                                                  const int *__restrict__ order, const float *__restrict__ wts,
                                                  float *__restrict__ x) {
    constexpr int DMAX = 8;
    const int64_t n4 = elems * G / 4;
    const float4 *r4 = reinterpret_cast<const float4 *>(ring);
    float4 *x4 = reinterpret_cast<float4 *>(x);
    const int g4 = G / 4;
    const int64_t e0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int w0 = 4 * (int)(e0 % g4);
    float wa[DMAX], wb[DMAX], wc[DMAX], wd[DMAX];
    int64_t off[DMAX];
#pragma unroll
    for (int j = 0; j < DMAX; ++j) {
        const bool in = j < m;
        wa[j] = in ? wts[w0 * depth + j] : 0.0f;
        wb[j] = in ? wts[(w0 + 1) * depth + j] : 0.0f;
        wc[j] = in ? wts[(w0 + 2) * depth + j] : 0.0f;
        wd[j] = in ? wts[(w0 + 3) * depth + j] : 0.0f;
        off[j] = in ? (int64_t)order[j] * n4 : 0;
    }
    const bool fa = isnan(wa[0]), fb = isnan(wb[0]), fc = isnan(wc[0]), fd = isnan(wd[0]);
    if (fa && fb && fc && fd) return;
    for (int64_t e = e0; e < n4; e += (int64_t)gridDim.x * blockDim.x) {
        float4 r = r4[off[0] + e];
        float4 v = float4{r.x * wa[0], r.y * wb[0], r.z * wc[0], r.w * wd[0]};
#pragma unroll
        for (int j = 1; j < DMAX; ++j)
            if (j < m) {
                r = r4[off[j] + e];
                v = float4{v.x + r.x * wa[j], v.y + r.y * wb[j], v.z + r.z * wc[j], v.w + r.w * wd[j]};
            }
        if (fa || fb || fc || fd) {
            const float4 o = x4[e];
            v = float4{fa ? o.x : v.x, fb ? o.y : v.y, fc ? o.z : v.z, fd ? o.w : v.w};
        }
        x4[e] = v;
    }
}

// The same for a lane count that is not a multiple of 4 (or unaligned buffers), one element per thread.
__global__ void mix_scalar_kernel(int64_t elems, int G, int depth, int m, const float *__restrict__ ring,
                                  const int *__restrict__ order, const float *__restrict__ wts, float *__restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < elems * G;
         e += (int64_t)gridDim.x * blockDim.x) {
        const float *wl = wts + (int)(e % G) * depth;
        if (isnan(wl[0])) continue;
        float v = ring[(int64_t)order[0] * elems * G + e] * wl[0];
        for (int j = 1; j < m; ++j) v = v + ring[(int64_t)order[j] * elems * G + e] * wl[j];
        x[e] = v;
    }
}

// ------------------------------------------------------------------------------------------------ host entries
int compose_launch(int64_t cells, int rows, int K, int G, const float *mom, const float *first, const float *factor,
                   float *out, hipStream_t s) {
    const int gs = lane_shift(G);
    if (gs < 0) return set_error("compose: lanes %d not a power of two", G);
    if (cells > 0)
        hipLaunchKernelGGL(compose_kernel, dim3((unsigned)cells), dim3(256), 0, s, rows, K, G, gs, mom, first, factor,
                           out);
    return check("compose");
}

int plane_launch(int nh, int nc, int L, int rows, int G, int strato, const float *src, const float *ha,
                 const float *sd, int half, void *plane, float *scale, hipStream_t s) {
    if (half) {
        if (nh * nc > 0)
            hipLaunchKernelGGL(plane_half_kernel, dim3((unsigned)(nh * nc)), dim3(256), 0, s, nc, L, rows, G, strato,
                               src, ha, sd, reinterpret_cast<_Float16 *>(plane), scale);
    } else {
        const int gs = lane_shift(G);
        if (gs < 0 || L > 32) return set_error("make_plane: lanes %d or L %d", G, L);
        if (nh * nc > 0)
            hipLaunchKernelGGL(plane_kernel, dim3((unsigned)(nh * nc)), dim3(256), 0, s, nc, L, rows, G, gs, strato,
                               src, ha, sd, reinterpret_cast<float *>(plane));
    }
    return check("make_plane");
}

int metrics_launch(int64_t rows, int G, const float *g, const float *x, unsigned frozen, float *out, hipStream_t s) {
    // out doubles as the uint accumulator [3][G] before the final conversion
    unsigned *acc = reinterpret_cast<unsigned *>(out);
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(128), 0, s, acc, 3 * G);
    hipLaunchKernelGGL(metrics_kernel, dim3(blocks_for(rows * G, 256, 4096)), dim3(256), 0, s, rows, G, g, x, frozen,
                       acc);
    hipLaunchKernelGGL(metrics_finish, dim3(1), dim3(64), 0, s, G, acc, out);
    return check("source_metrics");
}

int residual_launch(int64_t cells, int rows, int G, const float *g, const float *x, const float *obs, int nobs,
                    float *work, float *f, hipStream_t s) {
    float *norms = work;
    unsigned *maxbits = reinterpret_cast<unsigned *>(work + cells * G);
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s, maxbits, G);
    const int gs = lane_shift(G);
    if (gs < 0) return set_error("anderson_residual: lanes %d not a power of two", G);
    if (cells > 0) hipLaunchKernelGGL(norms_kernel, dim3((unsigned)cells), dim3(256), 0, s, rows, G, gs, g, norms, maxbits);
    const int64_t tiles = (cells * rows + 63) / 64;
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)max((int64_t)1, min(tiles, (int64_t)8192))), dim3(256), 0, s,
                       cells, rows, G, g, x, norms, maxbits, obs, nobs, f);
    return check("anderson_residual");
}

constexpr int GRAM_MAX_BLOCKS = 128;

// chunks per lane of gram_partial: about 64 Ki residuals each, at most GRAM_MAX_BLOCKS (the grid is chunks x lanes)
static int gram_blocks(int64_t n) {
    int blocks = (int)min((int64_t)GRAM_MAX_BLOCKS, (n + 65535) / 65536);
    return blocks < 1 ? 1 : blocks;
}

int gram_launch(int64_t n, int G, int depth, int slot, unsigned filled, const float *ring, double *work, double *gram,
                hipStream_t s) {
    const int blocks = gram_blocks(n);
    hipLaunchKernelGGL(gram_partial, dim3(blocks, G), dim3(256), 0, s, n, G, depth, slot, filled, ring, work);
    hipLaunchKernelGGL(gram_finish, dim3(1), dim3(256), 0, s, blocks, G, depth, slot, filled, work, gram);
    return check("anderson_gram");
}

// Allocation-free Gram products: the per-block partial sums live in a file-scope device array instead of a caller
// buffer. Its host-side name is only a stand-in for the symbol, so the launcher passes the device address that
// hipGetSymbolAddress reports for the current device, looked up once per device and kept.
constexpr int64_t GRAM_SCRATCH = (int64_t)GRAM_MAX_BLOCKS * 32 * 8;  // blocks x lanes x depth at their limits (256 KiB)
static __device__ __align__(256) double scratch[GRAM_SCRATCH];

static double *scratch_address() {
    static std::mutex mu;
    static double *addr[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        addr[dev] = static_cast<double *>(p);
    }
    return addr[dev];
}

int gram_static_launch(int64_t n, int G, int depth, int slot, unsigned filled, const float *ring, double *gram,
                       hipStream_t s) {
    const int blocks = gram_blocks(n);
    if ((int64_t)blocks * G * depth > GRAM_SCRATCH)
        return set_error("anderson_gram_static: %d blocks x %d lanes x depth %d exceed the %lld-double scratch", blocks,
                         G, depth, (long long)GRAM_SCRATCH);
    double *work = scratch_address();
    if (!work) return set_error("anderson_gram_static: hipGetSymbolAddress of the scratch array failed");
    hipLaunchKernelGGL(gram_partial, dim3(blocks, G), dim3(256), 0, s, n, G, depth, slot, filled, ring, work);
    hipLaunchKernelGGL(gram_finish, dim3(1), dim3(256), 0, s, blocks, G, depth, slot, filled, work, gram);
    return check("anderson_gram_static");
}

int solve_launch(int G, int depth, int m, const int *order, const double *gram, unsigned frozen, float *w,
                 hipStream_t s) {
    hipLaunchKernelGGL(solve_kernel, dim3(1), dim3(64), 0, s, G, depth, m, order, gram, frozen, w);
    return check("anderson_solve");
}

int mix_launch(int64_t cells, int rows, int G, int depth, int m, const float *ring, const int *order, const float *w,
               float *x, hipStream_t s) {
    const int64_t elems = cells * rows;
    if (G % 4 == 0 && ((uintptr_t)ring % 16) == 0 && ((uintptr_t)x % 16) == 0)
        // a grid of a few workgroups per CU: each thread streams many float4 groups after reading its weights once
        hipLaunchKernelGGL(mix_kernel, dim3(blocks_for(elems * G / 4, 256, 4096)), dim3(256), 0, s, elems, G, depth, m,
                           ring, order, w, x);
    else
        hipLaunchKernelGGL(mix_scalar_kernel, dim3(blocks_for(elems * G)), dim3(256), 0, s, elems, G, depth, m, ring,
                           order, w, x);
    return check("anderson_mix");
}

}  // namespace sos
