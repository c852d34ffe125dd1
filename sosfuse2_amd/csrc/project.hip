// Projection of per-ray radiances onto the incident-radiance moments of each destination cell (gfx950).
//
// moments[cell][k<4][w]  = sum_rays weight * rad[ray][w][k]                     (tensor moments)
// moments[cell][4+k][w]  = sum_rays (Y_k(mu) cos(m phi)) * (trace * weight)    (SH moments, k = l(l+1)/2 + m)
// flux[cell][w]          = albedo/pi * sum_rays trace * weight * max(0, dloc_z) (ground flux)
// (src/sosfuse2/backends/numpy_reference.py:683-711.) The SH moments are a dense contraction
// Z[rays x K]^T . V[rays x lanes]: one workgroup per cell stages 16 rays of Z (FP64 recurrence rounded to FP32, times
// the FP32 Chebyshev cos(m phi)) in LDS, and each wavefront accumulates 16-harmonic tiles with
// v_mfma_f32_16x16x4f32 (4 rays per instruction, lanes as the N dimension). Tensor moments and flux are FP64 sums in
// ray order. Sums are in a fixed order: results do not depend on timing.

#include "common.h"

namespace sos {



// Ground pass (flux only): FP64 in ray order, one thread per lane.
template <int G>
__device__ inline void project_flux_only(const ProjectArgs &a, int cell, int64_t r0, int64_t r1) {
    const int w = threadIdx.x;
    if (w >= G) return;
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const float *x = a.rad + (r * G + w) * 4;
        const float tr = (x[0] + x[1]) + x[2];
        s += (double)(tr * (a.weights[r] * fmaxf(0.0f, a.dloc[3 * r + 2])));
    }
    a.flux[(int64_t)cell * G + w] = (float)(s * (double)a.lane[SOS_LANE_ALBEDO * G + w] / SOS_PI);
}

// One workgroup per cell, its rays in rounds of 16, staged in LDS: Z (basis row x cos(m phi), into the l-major
// moment order), V (trace x weight), and the radiance tensors, from which threads 0 .. 5 G - 1 add the tensor moments
// and the flux term of each ray in ray order (FP64, the same order as a sequential pass over the cell's rays).
template <int G, int L>
__global__ __launch_bounds__(256) void project_kernel(ProjectArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KT = (K + 15) / 16;
    constexpr int KP = KT * 16;
    constexpr int NT = G >= 16 ? G / 16 : 1;
    constexpr int TPW = (KT + 3) / 4;        // harmonic tiles per wavefront
    constexpr int ZQ = 16 * (KP / 4);        // float4 chunks of a round's basis rows
    constexpr int RQ = 16 * G;               // float4 radiance tensors of a round
    __shared__ __attribute__((aligned(16))) float Z[16][KP];
    __shared__ float V[16][G];
    __shared__ float4 Rd[16][G];
    __shared__ float cmv[16][L + 1];
    __shared__ float wr[16], wf[16];        // weight; weight x max(0, dloc_z)
    __shared__ short kl[KP];                // l-major position kidx(l, m) of m-major position k' (K for the padding)
    __shared__ unsigned char km[KP];        // m of k'
    const int cell = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t r0 = a.group_off[cell], r1 = a.group_off[cell + 1];
    if (a.ground_only) {
        project_flux_only<G>(a, cell, r0, r1);
        return;
    }
    for (int k = tid; k < KP; k += 256) {
        int m = 0;
        if (k < K)
            while (moff<L>(m + 1) <= k) ++m;
        km[k] = (unsigned char)m;
        kl[k] = (short)(k < K ? kidx(k - moff<L>(m) + m, m) : K);
    }
    const int wave = tid >> 6, lane = tid & 63;
    floatx4 acc[TPW][NT];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = floatx4{0.f, 0.f, 0.f, 0.f};
    float accv[(K + 255) / 256];
#pragma unroll
    for (int i = 0; i < (K + 255) / 256; ++i) accv[i] = 0.f;
    double tsum = 0.0;  // thread tid < 5 G: tensor moment tid / G (0 .. 3) or the flux term (4) of lane tid % G
    for (int64_t c0 = r0; c0 < r1; c0 += 16) {
        const int n = (int)min((int64_t)16, r1 - c0);
        if (tid < 16) {
            // cos(m phi) of the ray's azimuth about the sun axis (FP32 Chebyshev, arithmetic.py:112-119)
            const int i = tid;
            float cp = 1.0f, wt = 0.0f, dl = 0.0f;
            if (i < n) {
                const int64_t r = c0 + i;
                const double dx = a.dirs[3 * r], dz = a.dirs[3 * r + 2];
                wt = a.weights[r];
                dl = a.dloc[3 * r + 2];
                const double st = sqrt(fmax(0.0, 1.0 - dz * dz));
                cp = (float)(st > 1e-12 ? dx / fmax(st, 1e-30) : 1.0);
                cp = clampf(cp, -1.0f, 1.0f);
            }
            float c0m = 1.0f, c1m = cp;
            for (int m = 0; m <= L; ++m) {
                float cm;
                if (m == 0) cm = 1.0f;
                else if (m == 1) cm = cp;
                else { cm = (2.0f * cp) * c1m - c0m; c0m = c1m; c1m = cm; }
                cmv[i][m] = cm;
            }
            wr[i] = i < n ? wt : 0.0f;
            wf[i] = i < n ? wt * fmaxf(0.0f, dl) : 0.0f;
        }
        for (int e = tid; e < RQ; e += 256) {
            const int i = e / G, w = e % G;
            Rd[i][w] = i < n ? reinterpret_cast<const float4 *>(a.rad)[(c0 + i) * G + w] : float4{0.f, 0.f, 0.f, 0.f};
        }
        __syncthreads();
        // basis rows: the ray's row of the bundle's basis table (fp32(x_m P_lm(mu))) times cos(m phi), in the table's
        // m-major order (one 16-byte store per chunk; the moments are put in l-major order as they are written)
        for (int e = tid; e < ZQ; e += 256) {
            const int i = e / (KP / 4), q = e % (KP / 4);
            float4 z = float4{0.0f, 0.0f, 0.0f, 0.0f};
            if (i < n) z = reinterpret_cast<const float4 *>(a.basis + (c0 + i) * KP)[q];
            const float *cm = cmv[i];
            const int k = 4 * q;
            reinterpret_cast<float4 *>(&Z[i][0])[q] =
                float4{z.x * cm[km[k]], z.y * cm[km[k + 1]], z.z * cm[km[k + 2]], z.w * cm[km[k + 3]]};
        }
        for (int e = tid; e < 16 * G; e += 256) {
            const int i = e / G, w = e % G;
            const float4 x = Rd[i][w];
            V[i][w] = ((x.x + x.y) + x.z) * wr[i];
        }
        if (tid < 5 * G) {
            const int comp = tid / G, w = tid % G;
            for (int i = 0; i < n; ++i) {
                const float4 x = Rd[i][w];
                if (comp == 4) tsum += (double)(((x.x + x.y) + x.z) * wf[i]);
                else tsum += (double)((comp == 0 ? x.x : comp == 1 ? x.y : comp == 2 ? x.z : x.w) * wr[i]);
            }
        }
        __syncthreads();
        if constexpr (G >= 16) {
            const int col = lane & 15, kk = lane >> 4;
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int kt = wave + 4 * t;
                if (kt < KT) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int ray = 4 * s + kk;
                        const float av = Z[ray][kt * 16 + col];
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[t][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, V[ray][nt * 16 + col], acc[t][nt], 0, 0, 0);
                    }
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < (K + 255) / 256; ++q) {
                const int k = tid + 256 * q;
                if (k < K)
                    for (int i = 0; i < n; ++i) accv[q] = accv[q] + Z[i][k] * V[i][0];
            }
        }
    }
    if (tid < 5 * G) {
        const int comp = tid / G, w = tid % G;
        if (comp < 4) a.moments[((int64_t)cell * (4 + K) + comp) * G + w] = (float)tsum;
        else a.flux[(int64_t)cell * G + w] = (float)(tsum * (double)a.lane[SOS_LANE_ALBEDO * G + w] / SOS_PI);
    }
    float *out = a.moments + (int64_t)cell * (4 + K) * G;
    if constexpr (G >= 16) {
        const int col = lane & 15, rb4 = 4 * (lane >> 4);
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int kt = wave + 4 * t;
            if (kt < KT) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k = kt * 16 + rb4 + v;  // m-major position: moment kl[k]
                        if (k < K) out[(4 + kl[k]) * G + nt * 16 + col] = acc[t][nt][v];
                    }
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < (K + 255) / 256; ++q) {
            const int k = tid + 256 * q;
            if (k < K) out[(4 + kl[k]) * G] = accv[q];
        }
    }
}

int set_error(const char *fmt, ...);

template <int G, int L>
static int project_launch(const ProjectArgs &a, hipStream_t s) {
    if (a.groups <= 0) return SOS_OK;
    hipLaunchKernelGGL((project_kernel<G, L>), dim3((unsigned)a.groups), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("project launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

template <int G>
static int project_L(const ProjectArgs &a, int L, hipStream_t s) {
    switch (L) {
        case 8: return project_launch<G, 8>(a, s);
        case 16: return project_launch<G, 16>(a, s);
        case 24: return project_launch<G, 24>(a, s);
        case 32: return project_launch<G, 32>(a, s);
        default: return set_error("unsupported L %d", L);
    }
}

int project_dispatch(const ProjectArgs &a, int G, int L, hipStream_t s) {
    switch (G) {
        case 1: return project_L<1>(a, L, s);
        case 16: return project_L<16>(a, L, s);
        case 32: return project_L<32>(a, L, s);
        default: return set_error("unsupported lane count %d", G);
    }
}

}  // namespace sos
