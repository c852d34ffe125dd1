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

// 1/sqrt(4 pi) prod_{j<=m} dd_j: the diagonal seed without the st^m factor (arithmetic.py:95-97)
__device__ inline float seed_coef(int m) {
    double c = 0.28209479177387814;
    for (int j = 1; j <= m; ++j) c *= -sqrt((2.0 * j + 1) / (2.0 * j));
    return (float)c;
}



template <int L, bool HALF>
__device__ inline float plane_at(const void *p, int64_t i) {
    if constexpr (HALF) return __half2float(reinterpret_cast<const __half *>(p)[i]);
    else return reinterpret_cast<const float *>(p)[i];
}

// Z rows of up to 32 visits into LDS: chain (visit, m) runs the Legendre recurrence over l.
template <int L, int NT>
__device__ inline void build_z(float (*Z)[((L + 1) * (L + 2) / 2 + 15) / 16 * 16], const float *smu, const float *scp,
                               const float *sseed, int nv) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    for (int c = threadIdx.x; c < 32 * (L + 1); c += NT) {
        const int v = c & 31, m = c >> 5;
        const int o = moff<L>(m);
        if (v >= nv) {
            for (int l = m; l <= L; ++l) Z[v][o + l - m] = 0.0f;
            continue;
        }
        const float mu = smu[v], cp = scp[v];
        const float st = sqrtf(fmaxf(0.0f, 1.0f - mu * mu));
        const float cm = m == 0 ? 1.0f : cosf((float)m * acosf(cp));
        const float mult = m ? 1.41421356237309515f * cm : cm;
        float z0 = sseed[m] * (m ? powf(st, (float)m) : 1.0f);
        if (m > 0 && st < 1e-30f) z0 = 0.0f;
        Z[v][o] = mult * z0;
        if (m < L) {
            float z1 = (sqrtf(2.0f * m + 3.0f) * mu) * z0;
            Z[v][o + 1] = mult * z1;
            for (int l = m + 2; l <= L; ++l) {
                const float fl = (float)l, fm = (float)m;
                const float aa = sqrtf((4.0f * fl * fl - 1.0f) / (fl * fl - fm * fm));
                const float bb = sqrtf((2.0f * fl + 1.0f) * ((fl - 1.0f) * (fl - 1.0f) - fm * fm) /
                                       ((2.0f * fl - 3.0f) * (fl * fl - fm * fm)));
                const float z2 = (aa * mu) * z1 - bb * z0;
                Z[v][o + l - m] = mult * z2;
                z0 = z1;
                z1 = z2;
            }
        }
    }
}

template <int G, int L, bool HALF>
__global__ __launch_bounds__(G * 8) void sh_records_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;   // padded row
    constexpr int KS = KZ / 4;               // MFMA k-steps
    constexpr int NT = G * 8;                // threads: 2 M-tiles x (G/16) N-tiles wavefronts
    __shared__ float Z[32][KZ];
    __shared__ float smu[32], scp[32], sseed[L + 1];
    __shared__ int sid[32];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int mt = wave / (G / 16), nt = wave % (G / 16);
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    for (int m = tid; m <= L; m += NT) sseed[m] = seed_coef(m);
    // B operand in registers: b[s] = c[cell][k'(s, lane)][nt*16 + lane%16]
    float b[KS];
    {
        const int col = nt * 16 + (lane & 15);
        const int64_t base = (int64_t)cell * K * G + col;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 16 * (s >> 2) + 4 * (lane >> 4) + (s & 3);
            b[s] = k < K ? plane_at<L, HALF>(a.plane, base + (int64_t)k * G) : 0.0f;
        }
    }
    for (int64_t r0 = v0; r0 < v1; r0 += 32) {
        const int nv = (int)min((int64_t)32, v1 - r0);
        if (tid < 32) {
            int id = -1;
            float mu = 0.0f, cp = 1.0f;
            if (tid < nv) {
                id = a.vis[r0 + tid];
                const int64_t node = id >> 2;
                mu = a.node_mu[node];
                cp = a.node_cp[node];
            }
            sid[tid] = id;
            smu[tid] = mu;
            scp[tid] = cp;
        }
        __syncthreads();
        build_z<L, NT>(Z, smu, scp, sseed, nv);
        for (int e = tid; e < 32 * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = 0.0f;
        __syncthreads();
        floatx4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
        const float *zrow = &Z[mt * 16 + (lane & 15)][4 * (lane >> 4)];
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
            const int row = mt * 16 + 4 * (lane >> 4) + v;
            if (row < nv) a.records[(int64_t)sid[row] * G + nt * 16 + (lane & 15)] = d[v];
        }
        __syncthreads();
    }
}

// W = 1 lanes: one thread per visit, dot product against the cell's coefficients staged in LDS.
template <int L, bool HALF>
__global__ __launch_bounds__(256) void sh_records_scalar(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    __shared__ float c[K];
    __shared__ float sseed[L + 1];
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    for (int k = threadIdx.x; k < K; k += 256) c[k] = plane_at<L, HALF>(a.plane, (int64_t)cell * K + k);
    for (int m = threadIdx.x; m <= L; m += 256) sseed[m] = seed_coef(m);
    __syncthreads();
    for (int64_t r = a.item_v0[item] + threadIdx.x; r < a.item_v1[item]; r += 256) {
        const int id = a.vis[r];
        const int64_t node = id >> 2;
        const float mu = a.node_mu[node], cp = a.node_cp[node];
        const float st = sqrtf(fmaxf(0.0f, 1.0f - mu * mu));
        float sum = 0.0f;
        float cm0 = 1.0f, cm1 = cp;
        for (int m = 0; m <= L; ++m) {
            float cm;
            if (m == 0) cm = 1.0f;
            else if (m == 1) cm = cp;
            else { cm = (2.0f * cp) * cm1 - cm0; cm0 = cm1; cm1 = cm; }
            const int o = moff<L>(m);
            float z0 = sseed[m] * (m ? powf(st, (float)m) : 1.0f);
            if (m > 0 && st < 1e-30f) z0 = 0.0f;
            float acc = z0 * c[o];
            if (m < L) {
                float z1 = (sqrtf(2.0f * m + 3.0f) * mu) * z0;
                acc = acc + z1 * c[o + 1];
                for (int l = m + 2; l <= L; ++l) {
                    const float fl = (float)l, fm = (float)m;
                    const float aa = sqrtf((4.0f * fl * fl - 1.0f) / (fl * fl - fm * fm));
                    const float bb = sqrtf((2.0f * fl + 1.0f) * ((fl - 1.0f) * (fl - 1.0f) - fm * fm) /
                                           ((2.0f * fl - 3.0f) * (fl * fl - fm * fm)));
                    const float z2 = (aa * mu) * z1 - bb * z0;
                    acc = acc + z2 * c[o + l - m];
                    z0 = z1;
                    z1 = z2;
                }
            }
            sum = sum + (m ? 1.41421356237309515f * cm : 1.0f) * acc;
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
        if (half) hipLaunchKernelGGL((sh_records_kernel<G, L, true>), dim3((unsigned)a.items), dim3(G * 8), 0, s, a);
        else hipLaunchKernelGGL((sh_records_kernel<G, L, false>), dim3((unsigned)a.items), dim3(G * 8), 0, s, a);
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
