// Spherical-harmonic reconstruction of the source at every (node, corner) visit as an MFMA contraction (gfx950).
//
// aero[visit][w] = sum_k Z[visit][k] c[cell][k][w],  Z[visit][k] = sqrt2^(m>0) P_lm(mu) cos(m phi)
// where the visit belongs to a quadrature node of a ray (mu = the ray's direction cosine to the sun axis, cos(m phi)
// the node's azimuth) and c is the combined SH plane of the visited cell. This is the per-corner SH sum of
// src/sosfuse2/backends/numpy_reference.py:579-600 (cache[m] = sum_l Y c, aero = sum_m cm cache[m]), regrouped by
// source cell: visits are sorted by cell once per bundle (the geometry is fixed), so each cell's K x lanes
// coefficients are loaded once into registers as the B operand and every visiting row streams through
// v_mfma_f32_16x16x4f32. The plan kernels walk the same node geometry as the transport.
//
// K order (the plane layout): m-major, k' = moff(m) + (l - m). MFMA k-steps are permuted so that one ds_read_b128
// feeds four consecutive MFMAs: step s covers k' = 16 (s/4) + 4 (lane/16) + (s%4).

#include "common.h"

namespace sos {
int set_error(const char *fmt, ...);

template <int L, bool HALF>
__device__ inline float plane_at(const void *p, int64_t i) {
    if constexpr (HALF) return __half2float(reinterpret_cast<const __half *>(p)[i]);
    else return reinterpret_cast<const float *>(p)[i];
}

// Recurrence constants in FP64 (arithmetic.py:69-83), staged once per workgroup.
template <int L>
__device__ inline void stage_constants(double *aa, double *bb, int nt) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    for (int k = threadIdx.x; k < K; k += nt) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) ++l;
        const int m = k - l * (l + 1) / 2;
        const double fl = l, fm = m;
        aa[k] = l > m ? sqrt((4.0 * fl * fl - 1.0) / (fl * fl - fm * fm)) : 0.0;
        bb[k] = l > m + 1 ? sqrt((2.0 * fl + 1.0) * ((fl - 1.0) * (fl - 1.0) - fm * fm) / ((2.0 * fl - 3.0) * (fl * fl - fm * fm)))
                          : 0.0;
    }
}

// Per visit: the diagonal seeds 1/sqrt(4 pi) prod dd_j st (FP64, arithmetic.py:95-97) and cos(m phi) by the FP32
// Chebyshev recurrence (arithmetic.py:112-119).
template <int L, int V>
__device__ inline void visit_seeds(int v, double mu, float cp, const double *dd, double (*seed)[V], float (*cmv)[V]) {
    const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
    double diag = 0.28209479177387814;
    float c0 = 1.0f, c1 = cp;
    for (int m = 0; m <= L; ++m) {
        if (m) diag = (diag * dd[m]) * st;
        seed[m][v] = diag;
        float cm;
        if (m == 0) cm = 1.0f;
        else if (m == 1) cm = cp;
        else { cm = (2.0f * cp) * c1 - c0; c0 = c1; c1 = cm; }
        cmv[m][v] = cm;
    }
}

// Z rows of V visits into LDS. Task (visit, p) runs the Legendre recurrence over l in FP64 for the two orders
// m = p and m = L - p side by side (their lengths add up to L + 2, so tasks are balanced and each thread has two
// independent dependency chains); each basis value is rounded to FP32 before the FP32 product with cos(m phi)
// (the reference's FP64 basis, rounded).
template <int L, int V, int KZ, typename T = float>
__device__ inline void build_z(T (*Z)[KZ], const double *mus, const double (*seed)[V], const float (*cmv)[V],
                               const double *aa, const double *bb, int nv, int nt) {
    constexpr int PAIRS = L / 2 + 1;
    for (int task = threadIdx.x; task < V * PAIRS; task += nt) {
        const int v = task % V, p = task / V;
        const int ma = p, mb = L - p;
        const bool two = mb != ma;
        const int oa = moff<L>(ma), ob = moff<L>(mb);
        if (v >= nv) {
            for (int l = ma; l <= L; ++l) Z[v][oa + l - ma] = (T)0.0f;
            if (two)
                for (int l = mb; l <= L; ++l) Z[v][ob + l - mb] = (T)0.0f;
            continue;
        }
        const double mu = mus[v];
        const float ca = cmv[ma][v], cb = cmv[mb][v];
        const double xa = ma ? 1.41421356237309515 : 1.0, xb = 1.41421356237309515;
        // chain a: m = p (long); chain b: m = L - p (short, length p + 1)
        double a0 = seed[ma][v], b0 = seed[mb][v];
        Z[v][oa] = (T)((float)(xa * a0) * ca);
        if (two) Z[v][ob] = (T)((float)(xb * b0) * cb);
        double a1 = (sqrt(2.0 * ma + 3.0) * mu) * a0;  // ma < L always
        Z[v][oa + 1] = (T)((float)(xa * a1) * ca);
        double b1 = 0.0;
        if (two && mb < L) {
            b1 = (sqrt(2.0 * mb + 3.0) * mu) * b0;
            Z[v][ob + 1] = (T)((float)(xb * b1) * cb);
        }
        // both chains advance together while chain b lasts
        int la = ma + 2, lb = mb + 2;
#pragma unroll 2
        for (; two && lb <= L; ++la, ++lb) {
            const int ka = kidx(la, ma), kb = kidx(lb, mb);
            const double a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            const double b2 = (aa[kb] * mu) * b1 - bb[kb] * b0;
            Z[v][oa + la - ma] = (T)((float)(xa * a2) * ca);
            Z[v][ob + lb - mb] = (T)((float)(xb * b2) * cb);
            a0 = a1; a1 = a2;
            b0 = b1; b1 = b2;
        }
#pragma unroll 4
        for (; la <= L; ++la) {
            const int ka = kidx(la, ma);
            const double a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            Z[v][oa + la - ma] = (T)((float)(xa * a2) * ca);
            a0 = a1; a1 = a2;
        }
    }
}

template <int G, int L>
__global__ __launch_bounds__(G * 4) __attribute__((amdgpu_waves_per_eu(2, 2))) void sh_records_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;   // padded row
    constexpr int KS = KZ / 4;               // MFMA k-steps
    constexpr int NT = G * 4;                // one wavefront per 16-lane N-tile
    constexpr int V = 16;                    // visits per round (one M-tile)
    constexpr int ZS = KZ + 4;               // row stride = 4 (mod 32 banks): 16 rows spread over the LDS banks
    __shared__ float Z[V][ZS];
    __shared__ double aa[K], bb[K];
    __shared__ double seed[L + 1][V];
    __shared__ float cmv[L + 1][V];
    __shared__ double smu[V];
    __shared__ double dd[L + 1];
    __shared__ int sid[V];
    const int tid = threadIdx.x, lane = tid & 63;
    const int nt = tid >> 6;
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_constants<L>(aa, bb, NT);
    for (int m = tid; m <= L; m += NT) dd[m] = m ? -sqrt((2.0 * m + 1) / (2.0 * m)) : 0.0;
    for (int e = tid; e < V * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = 0.0f;
    // B operand in registers: b[s] = c[cell][k'(s, lane)][nt*16 + lane%16]
    float b[KS];
    {
        const int col = nt * 16 + (lane & 15);
        const int64_t base = (int64_t)cell * K * G + col;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 16 * (s >> 2) + 4 * (lane >> 4) + (s & 3);
            b[s] = k < K ? reinterpret_cast<const float *>(a.plane)[base + (int64_t)k * G] : 0.0f;
        }
    }
    // visit data of the next round is loaded one round ahead (contiguous: the plan stores mu and cos(phi) in sorted
    // visit order), so its memory latency overlaps the current round's work
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (tid < V && v0 + tid < v1) {
        nid = a.vis[v0 + tid];
        nmu = a.vis_mu[v0 + tid];
        ncp = a.vis_cp[v0 + tid];
    }
    for (int64_t r0 = v0; r0 < v1; r0 += V) {
        const int nv = (int)min((int64_t)V, v1 - r0);
        if (tid < V) {
            const int id = nid;
            const double mu = (double)nmu;
            const float cp = ncp;
            const int64_t nx = r0 + V + tid;
            nid = -1;
            if (nx < v1) {
                nid = a.vis[nx];
                nmu = a.vis_mu[nx];
                ncp = a.vis_cp[nx];
            }
            sid[tid] = tid < nv ? id : -1;
            smu[tid] = tid < nv ? mu : 0.0;
            visit_seeds<L, V>(tid, tid < nv ? mu : 0.0, tid < nv ? cp : 1.0f, dd, seed, cmv);
        }
        __syncthreads();
        build_z<L, V, ZS>(Z, smu, seed, cmv, aa, bb, nv, NT);
        __syncthreads();
        floatx4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
        const float *zrow = &Z[lane & 15][4 * (lane >> 4)];
#pragma unroll
        for (int q = 0; q < KS / 4; ++q) {
            const floatx4 av = *reinterpret_cast<const floatx4 *>(zrow + 16 * q);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], b[4 * q + u], acc[u], 0, 0, 0);
        }
        const floatx4 d = (acc[0] + acc[1]) + (acc[2] + acc[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[(int64_t)sid[row] * G + nt * 16 + (lane & 15)] = d[v];
        }
        __syncthreads();
    }
}

// Mixed precision: the plane is FP16 scaled per (cell, lane) (plane = c / scale[cell][lane]), Z rows are FP16, and
// v_mfma_f32_16x16x16f16 accumulates in FP32; the epilogue multiplies by the scale. 32 visits per round
// (two M-tiles); wavefront w takes M-tile w / (G/16) and N-tile w % (G/16).
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

template <int G, int L>
__global__ __launch_bounds__(G * 8) __attribute__((amdgpu_waves_per_eu(2, 2))) void sh_records_half_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KS = KZ / 16;               // MFMA k-steps of 16
    constexpr int NT = G * 8;
    constexpr int V = 32;
    constexpr int ZS = KZ + 8;                // halves: row stride = 4 (mod 32 banks) in 32-bit words
    __shared__ _Float16 Z[V][ZS];
    __shared__ double aa[K], bb[K];
    __shared__ double seed[L + 1][V];
    __shared__ float cmv[L + 1][V];
    __shared__ double smu[V];
    __shared__ double dd[L + 1];
    __shared__ int sid[V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mt = wave / (G / 16), nt = wave % (G / 16);
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_constants<L>(aa, bb, NT);
    for (int m = tid; m <= L; m += NT) dd[m] = m ? -sqrt((2.0 * m + 1) / (2.0 * m)) : 0.0;
    for (int e = tid; e < V * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = (_Float16)0.0f;
    const int col = nt * 16 + (lane & 15);
    const float scale = a.scale[(int64_t)cell * G + col];
    // B operand: b[s] holds plane[cell][16 s + 4 (lane / 16) + 0..3][col]
    half4_t b[KS];
    {
        const _Float16 *pl = reinterpret_cast<const _Float16 *>(a.plane) + (int64_t)cell * K * G + col;
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = 16 * s + 4 * (lane >> 4) + u;
                b[s][u] = k < K ? pl[(int64_t)k * G] : (_Float16)0.0f;
            }
    }
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (tid < V && v0 + tid < v1) {
        nid = a.vis[v0 + tid];
        nmu = a.vis_mu[v0 + tid];
        ncp = a.vis_cp[v0 + tid];
    }
    for (int64_t r0 = v0; r0 < v1; r0 += V) {
        const int nv = (int)min((int64_t)V, v1 - r0);
        if (tid < V) {
            const int id = nid;
            const double mu = (double)nmu;
            const float cp = ncp;
            const int64_t nx = r0 + V + tid;
            nid = -1;
            if (nx < v1) {
                nid = a.vis[nx];
                nmu = a.vis_mu[nx];
                ncp = a.vis_cp[nx];
            }
            sid[tid] = tid < nv ? id : -1;
            smu[tid] = tid < nv ? mu : 0.0;
            visit_seeds<L, V>(tid, tid < nv ? mu : 0.0, tid < nv ? cp : 1.0f, dd, seed, cmv);
        }
        __syncthreads();
        build_z<L, V, ZS, _Float16>(Z, smu, seed, cmv, aa, bb, nv, NT);
        __syncthreads();
        floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
        const _Float16 *zrow = &Z[mt * 16 + (lane & 15)][4 * (lane >> 4)];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const half4_t av = *reinterpret_cast<const half4_t *>(zrow + 16 * s);
            acc[s & 1] = __builtin_amdgcn_mfma_f32_16x16x16f16(av, b[s], acc[s & 1], 0, 0, 0);
        }
        const floatx4 d = acc[0] + acc[1];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = mt * 16 + 4 * (lane >> 4) + v;
            if (row < nv) a.records[(int64_t)sid[row] * G + col] = d[v] * scale;
        }
        __syncthreads();
    }
}

// W = 1 lanes: one thread per visit, dot product against the cell's coefficients staged in LDS (an FP16 plane is
// scaled back to FP32 there; the basis stays FP32).
template <int L, bool HALF>
__global__ __launch_bounds__(256) void sh_records_scalar(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    __shared__ float c[K];
    __shared__ double aa[K], bb[K];
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    const float sc = HALF ? a.scale[cell] : 1.0f;
    for (int k = threadIdx.x; k < K; k += 256) c[k] = plane_at<L, HALF>(a.plane, (int64_t)cell * K + k) * sc;
    stage_constants<L>(aa, bb, 256);
    __syncthreads();
    for (int64_t r = a.item_v0[item] + threadIdx.x; r < a.item_v1[item]; r += 256) {
        const int id = a.vis[r];
        const double mu = a.vis_mu[r];
        const float cp = a.vis_cp[r];
        const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
        float sum = 0.0f;
        float cm0 = 1.0f, cm1 = cp;
        double diag = 0.28209479177387814;
        for (int m = 0; m <= L; ++m) {
            float cm;
            if (m == 0) cm = 1.0f;
            else if (m == 1) cm = cp;
            else { cm = (2.0f * cp) * cm1 - cm0; cm0 = cm1; cm1 = cm; }
            if (m) diag = (diag * -sqrt((2.0 * m + 1) / (2.0 * m))) * st;
            const int o = moff<L>(m);
            const double mult = m ? 1.41421356237309515 : 1.0;
            double z0 = diag;
            float acc = ((float)(mult * z0) * cm) * c[o];
            if (m < L) {
                double z1 = (sqrt(2.0 * m + 3.0) * mu) * z0;
                acc = acc + ((float)(mult * z1) * cm) * c[o + 1];
                for (int l = m + 2; l <= L; ++l) {
                    const int k = kidx(l, m);
                    const double z2 = (aa[k] * mu) * z1 - bb[k] * z0;
                    acc = acc + ((float)(mult * z2) * cm) * c[o + l - m];
                    z0 = z1;
                    z1 = z2;
                }
            }
            sum = sum + acc;
        }
        a.records[id] = sum;
    }
}

template <int G, int L>
static int sh_launch(const ShArgs &a, bool half, hipStream_t s) {
    if (a.items <= 0) return SOS_OK;
    if constexpr (G == 1) {
        if (half) hipLaunchKernelGGL((sh_records_scalar<L, true>), dim3((unsigned)a.items), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((sh_records_scalar<L, false>), dim3((unsigned)a.items), dim3(256), 0, s, a);
    } else {
        if (half) hipLaunchKernelGGL((sh_records_half_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 8), 0, s, a);
        else hipLaunchKernelGGL((sh_records_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 4), 0, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("sh_records launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

template <int G>
static int sh_L(const ShArgs &a, bool half, hipStream_t s) {
    switch (a.L) {
        case 8: return sh_launch<G, 8>(a, half, s);
        case 16: return sh_launch<G, 16>(a, half, s);
        case 32: return sh_launch<G, 32>(a, half, s);
        default: return set_error("sh_records: L %d", a.L);
    }
}

int sh_records_dispatch(const ShArgs &a, bool half, hipStream_t s) {
    switch (a.G) {
        case 1: return sh_L<1>(a, half, s);
        case 16: return sh_L<16>(a, half, s);
        case 32: return sh_L<32>(a, half, s);
        default: return set_error("sh_records: lanes %d", a.G);
    }
}

}  // namespace sos
