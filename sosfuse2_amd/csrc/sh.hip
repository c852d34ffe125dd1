// Spherical-harmonic reconstruction of the source at every (node, corner) visit as an MFMA contraction (gfx950).
//
// rec[node][c][w] = sum_k Z[node][k] c[cell_c][k][w],  Z[node][k] = sqrt2^(m>0) P_lm(mu) cos(m phi)
// with mu the ray's direction cosine to the sun axis and cos(m phi) the node's azimuth, cell_c the four bilinear
// corner cells of the node. This is the per-corner SH sum of src/sosfuse2/backends/numpy_reference.py:579-600
// (cache[m] = sum_l Y c, aero = sum_m cm cache[m]), regrouped: the basis row Z of a node is the same for its four
// corners, so nodes are sorted by their base cell once per bundle (the geometry is fixed) and one workgroup takes a
// run of nodes of one base cell. Every wavefront holds the K x 16 coefficients of one corner cell and one 16-lane
// N-tile in registers as the MFMA B operand; tiles of 16 basis rows are generated once in LDS and stream past all
// four corners. The plan kernels walk the same node geometry as the transport.
//
// K order (the plane layout): m-major, k' = moff(m) + (l - m). FP32 MFMA k-steps are permuted so that one
// ds_read_b128 feeds four consecutive MFMAs: step 4 i + u covers k' = 16 i + 4 (lane/16) + u.

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

// Per node: the diagonal seeds 1/sqrt(4 pi) prod dd_j st (FP64, arithmetic.py:95-97) and cos(m phi) by the FP32
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

// Z rows of V nodes into LDS. Task (node, p) runs the Legendre recurrence over l in FP64 for the two orders
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

// Mixed precision: split-FP16 operands on v_mfma_f32_16x16x16f16 with FP32 accumulation, the three-product scheme
// of Ootomo and Yokota (2022, "Recovering single precision accuracy from Tensor Cores while surpassing the FP32
// theoretical peak performance"): x = hi + lo 2^-11 with hi = fp16(x) and lo = fp16((x - hi) 2^11), and
// a b = ah bh + (ah bl + al bh) 2^-11 (the lo lo term, 2^-22 relative, is dropped). The plane is stored scaled per
// (cell, lane) to [-1, 1] (plane = c / scale, scale a power of two) so hi and lo stay in FP16's normal range; Z rows
// are generated by the FP32 recurrence and split on the way into LDS. Each k-step of 16 costs three 8-pass MFMAs
// instead of four 32-pass FP32 MFMAs; the error per record is about 2^-21 of sum_k |Z_k c_k|, the FP32 level.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

// Per node, FP32: the diagonal seeds 1/sqrt(4 pi) prod dd_j st and cos(m phi) by the Chebyshev recurrence. Threads
// v < V run the seed chain, threads V <= t < 2V the cosine chain.
template <int L, int V>
__device__ inline void visit_seeds32(int t, float mu, float cp, const float *dd, float (*seed)[V], float (*cmv)[V]) {
    if (t < V) {
        const float st = sqrtf(fmaxf(0.0f, 1.0f - mu * mu));
        float diag = 0.28209479177387814f;
        for (int m = 0; m <= L; ++m) {
            if (m) diag = (diag * dd[m]) * st;
            seed[m][t] = diag;
        }
    } else {
        const int v = t - V;
        float c0 = 1.0f, c1 = cp;
        for (int m = 0; m <= L; ++m) {
            float cm;
            if (m == 0) cm = 1.0f;
            else if (m == 1) cm = cp;
            else { cm = (2.0f * cp) * c1 - c0; c0 = c1; c1 = cm; }
            cmv[m][v] = cm;
        }
    }
}

__device__ inline void split_store(_Float16 *hi, _Float16 *lo, float z) {
    const _Float16 h = (_Float16)z;
    *hi = h;
    *lo = (_Float16)((z - (float)h) * 2048.0f);
}

// Z rows of V nodes, split into hi and lo planes: task (node, p) runs the FP32 recurrence of the orders m = p and
// m = L - p side by side (as build_z).
template <int L, int V, int ZS>
__device__ inline void build_z_split(_Float16 (*Zh)[ZS], _Float16 (*Zl)[ZS], const float *mus, const float (*seed)[V],
                                     const float (*cmv)[V], const float *aa, const float *bb, int nv, int nt) {
    constexpr int PAIRS = L / 2 + 1;
    for (int task = threadIdx.x; task < V * PAIRS; task += nt) {
        const int v = task % V, p = task / V;
        const int ma = p, mb = L - p;
        const bool two = mb != ma;
        const int oa = moff<L>(ma), ob = moff<L>(mb);
        if (v >= nv) {
            for (int l = ma; l <= L; ++l) Zh[v][oa + l - ma] = Zl[v][oa + l - ma] = (_Float16)0.0f;
            if (two)
                for (int l = mb; l <= L; ++l) Zh[v][ob + l - mb] = Zl[v][ob + l - mb] = (_Float16)0.0f;
            continue;
        }
        const float mu = mus[v];
        const float ca = ma ? 1.41421356237309515f * cmv[ma][v] : cmv[ma][v];
        const float cb = 1.41421356237309515f * cmv[mb][v];
        float a0 = seed[ma][v], b0 = seed[mb][v];
        split_store(&Zh[v][oa], &Zl[v][oa], a0 * ca);
        if (two) split_store(&Zh[v][ob], &Zl[v][ob], b0 * cb);
        float a1 = (sqrtf(2.0f * ma + 3.0f) * mu) * a0;
        split_store(&Zh[v][oa + 1], &Zl[v][oa + 1], a1 * ca);
        float b1 = 0.0f;
        if (two && mb < L) {
            b1 = (sqrtf(2.0f * mb + 3.0f) * mu) * b0;
            split_store(&Zh[v][ob + 1], &Zl[v][ob + 1], b1 * cb);
        }
        int la = ma + 2, lb = mb + 2;
#pragma unroll 1
        for (; two && lb <= L; ++la, ++lb) {
            const int ka = kidx(la, ma), kb = kidx(lb, mb);
            const float a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            const float b2 = (aa[kb] * mu) * b1 - bb[kb] * b0;
            split_store(&Zh[v][oa + la - ma], &Zl[v][oa + la - ma], a2 * ca);
            split_store(&Zh[v][ob + lb - mb], &Zl[v][ob + lb - mb], b2 * cb);
            a0 = a1; a1 = a2;
            b0 = b1; b1 = b2;
        }
#pragma unroll 1
        for (; la <= L; ++la) {
            const int ka = kidx(la, ma);
            const float a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            split_store(&Zh[v][oa + la - ma], &Zl[v][oa + la - ma], a2 * ca);
            a0 = a1; a1 = a2;
        }
    }
}

// Corner c of a node with base cell b (ir * nc + j): b, b + 1, b + nc, b + nc + 1, the transport's corner order.
__device__ inline int corner_cell(int b, int c, int nc) { return b + (c & 1) + (c >> 1) * nc; }

// FP32, two workgroup shapes (a tuning choice; the records do not depend on it):
//   TILE = false: one workgroup per work item, 4 x G/16 wavefronts, wavefront (c, nt) holds corner c's
//                 coefficients of lanes 16 nt .. 16 nt + 15 (whole K, about 250 VGPRs, one workgroup per CU);
//   TILE = true:  one workgroup per (work item, 16-lane tile), the K range split over KP wavefronts per corner
//                 (KP = 2 at L = 32, at most 128 VGPRs) so that two workgroups share a CU and one's basis generation
//                 overlaps the other's MFMA phase; the basis rows are then generated once per tile.
template <int G, int L, bool TILE>
struct ShShape {
    static constexpr int KP = (TILE && L >= 32) ? 2 : 1;
    static constexpr int THREADS = TILE ? 256 * KP : G * 16;
    static constexpr int WAVES_PER_EU = TILE ? 4 : 2;
};

template <int G, int L, bool TILE>
__global__ __launch_bounds__((ShShape<G, L, TILE>::THREADS))
__attribute__((amdgpu_waves_per_eu((ShShape<G, L, TILE>::WAVES_PER_EU), (ShShape<G, L, TILE>::WAVES_PER_EU))))
void sh_records_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;   // padded row
    constexpr int KQ = KZ / 16;              // groups of four MFMA k-steps (one ds_read_b128 each)
    constexpr int KP = ShShape<G, L, TILE>::KP;
    constexpr int KQH = (KQ + KP - 1) / KP;  // groups of one wavefront
    constexpr int NWB = TILE ? 1 : G / 16;   // 16-lane tiles per workgroup
    constexpr int NC = 4 * NWB;              // (corner, tile) wavefronts per K part
    constexpr int NT = 64 * NC * KP;
    constexpr int V = 16;                    // nodes per round (one M-tile)
    constexpr int ZS = KZ + 4;               // row stride = 4 (mod 32 banks): 16 rows spread over the LDS banks
    __shared__ float Z[V][ZS];
    __shared__ double aa[K], bb[K];
    __shared__ double seed[L + 1][V];
    __shared__ float cmv[L + 1][V];
    __shared__ double smu[V];
    __shared__ double dd[L + 1];
    __shared__ int sid[V];
    __shared__ floatx4 red[KP > 1 ? NC * 64 : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wc = KP > 1 ? wave % NC : wave, kh = KP > 1 ? wave / NC : 0;  // kh a constant 0 without a K split
    const int corner = wc / NWB, nt = TILE ? (int)blockIdx.y : wc % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_constants<L>(aa, bb, NT);
    for (int m = tid; m <= L; m += NT) dd[m] = m ? -sqrt((2.0 * m + 1) / (2.0 * m)) : 0.0;
    for (int e = tid; e < V * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = 0.0f;
    // B operand in registers: b[4 i + u] = c[cell][k][col] at k = 16 (kh KQH + i) + 4 (lane / 16) + u
    float b[4 * KQH];
    {
        const float *pl = reinterpret_cast<const float *>(a.plane) + (int64_t)cell * K * G + col;
#pragma unroll
        for (int i = 0; i < KQH; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = 16 * (kh * KQH + i) + 4 * (lane >> 4) + u;
                b[4 * i + u] = k < K ? pl[(int64_t)k * G] : 0.0f;
            }
    }
    // node data of the next round is loaded one round ahead (contiguous in the sorted order)
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (tid < V && v0 + tid < v1) {
        nid = a.node[v0 + tid];
        nmu = a.node_mu[v0 + tid];
        ncp = a.node_cp[v0 + tid];
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
                nid = a.node[nx];
                nmu = a.node_mu[nx];
                ncp = a.node_cp[nx];
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
        const float *zrow = &Z[lane & 15][4 * (lane >> 4) + 16 * kh * KQH];
#pragma unroll
        for (int i = 0; i < KQH; ++i) {
            if (kh * KQH + i < KQ) {
                const floatx4 av = *reinterpret_cast<const floatx4 *>(zrow + 16 * i);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], b[4 * i + u], acc[u], 0, 0, 0);
                if constexpr (TILE) __builtin_amdgcn_sched_barrier(0);
            }
        }
        floatx4 d = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        if constexpr (KP > 1) {
            if (kh == 1) red[wc * 64 + lane] = d;
            __syncthreads();
            if (kh == 0) d = d + red[wc * 64 + lane];
        }
        if (kh == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v;
                if (row < nv) a.records[((int64_t)sid[row] * 4 + corner) * G + col] = d[v];
            }
        }
        __syncthreads();
    }
}

// The split kernel runs 4 x G/16 wavefronts, wavefront (corner, N-tile) holding the whole K range of its hi / lo B
// operands (144 VGPRs at L = 32 in the quad layout, about 240 in all: two wavefronts per SIMD); the basis rows are
// generated once per node. (KP > 1 would split K over wavefronts and sum the parts through LDS.)
template <int L>
struct SplitK { static constexpr int KP = 1; };

template <int G, int L>
__global__ __launch_bounds__(G * 16 * SplitK<L>::KP) __attribute__((amdgpu_waves_per_eu(2, 2)))
void sh_records_split_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KS = KZ / 16;               // MFMA k-steps of 16
    constexpr int KP = SplitK<L>::KP;
    constexpr int KSH = (KS + KP - 1) / KP;   // k-steps of one wavefront
    constexpr int NW = G / 16;                // N-tiles
    constexpr int NC = 4 * NW;                // (corner, N-tile) wavefronts per K part
    constexpr int NT = 64 * NC * KP;
    constexpr int V = 16;                     // nodes per round (one M-tile)
    constexpr int ZS = KZ + 8;                // halves: row stride = 4 (mod 32 banks) in 32-bit words
    __shared__ _Float16 Zh[V][ZS], Zl[V][ZS];
    __shared__ float aa[K], bb[K];
    __shared__ float seed[L + 1][V];
    __shared__ float cmv[L + 1][V];
    __shared__ float smu[V], scp[V];
    __shared__ float dd[L + 1];
    __shared__ int sid[V];
    __shared__ floatx4 red[KP > 1 ? NC * 64 : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wc = KP > 1 ? wave % NC : wave, kh = KP > 1 ? wave / NC : 0;  // kh a constant 0 without a K split
    const int corner = wc / NW, nt = wc % NW;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    for (int k = tid; k < K; k += NT) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) ++l;
        const int m = k - l * (l + 1) / 2;
        const double fl = l, fm = m;
        aa[k] = l > m ? (float)sqrt((4.0 * fl * fl - 1.0) / (fl * fl - fm * fm)) : 0.0f;
        bb[k] = l > m + 1 ? (float)sqrt((2.0 * fl + 1.0) * ((fl - 1.0) * (fl - 1.0) - fm * fm) /
                                        ((2.0 * fl - 3.0) * (fl * fl - fm * fm)))
                          : 0.0f;
    }
    for (int m = tid; m <= L; m += NT) dd[m] = m ? -(float)sqrt((2.0 * m + 1) / (2.0 * m)) : 0.0f;
    for (int e = tid; e < V * (KZ - K); e += NT) {
        Zh[e / (KZ - K)][K + e % (KZ - K)] = (_Float16)0.0f;
        Zl[e / (KZ - K)][K + e % (KZ - K)] = (_Float16)0.0f;
    }
    const float scale = a.scale[(int64_t)cell * G + col];
    // B operands from the quad layout [cell][KZ / 4][G] (half4): bh[i], bl[i] hold K positions
    // 16 (kh KSH + i) + 4 (lane / 16) + 0..3 of column col
    half4_t bh[KSH], bl[KSH];
    {
        const half4_t *ph = reinterpret_cast<const half4_t *>(a.plane) + (int64_t)cell * (KZ / 4) * G + col;
        const half4_t *pl = ph + a.plane_lo / 4;
#pragma unroll
        for (int i = 0; i < KSH; ++i) {
            const int q = 4 * (kh * KSH + i) + (lane >> 4);
            bh[i] = q < KZ / 4 ? ph[(int64_t)q * G] : half4_t{0, 0, 0, 0};
            bl[i] = q < KZ / 4 ? pl[(int64_t)q * G] : half4_t{0, 0, 0, 0};
        }
    }
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (tid < V && v0 + tid < v1) {
        nid = a.node[v0 + tid];
        nmu = a.node_mu[v0 + tid];
        ncp = a.node_cp[v0 + tid];
    }
    for (int64_t r0 = v0; r0 < v1; r0 += V) {
        const int nv = (int)min((int64_t)V, v1 - r0);
        if (tid < V) {
            const int64_t nx = r0 + V + tid;
            sid[tid] = tid < nv ? nid : -1;
            smu[tid] = tid < nv ? nmu : 0.0f;
            scp[tid] = tid < nv ? ncp : 1.0f;
            nid = -1;
            if (nx < v1) {
                nid = a.node[nx];
                nmu = a.node_mu[nx];
                ncp = a.node_cp[nx];
            }
        }
        __syncthreads();
        if (tid < 2 * V) {
            const int v = tid % V;
            visit_seeds32<L, V>(tid, smu[v], scp[v], dd, seed, cmv);
        }
        __syncthreads();
        build_z_split<L, V, ZS>(Zh, Zl, smu, seed, cmv, aa, bb, nv, NT);
        __syncthreads();
        floatx4 am = floatx4{0.f, 0.f, 0.f, 0.f}, ac = am;
        const _Float16 *hrow = &Zh[lane & 15][4 * (lane >> 4) + 16 * kh * KSH];
        const _Float16 *lrow = &Zl[lane & 15][4 * (lane >> 4) + 16 * kh * KSH];
#pragma unroll
        for (int i = 0; i < KSH; ++i) {
            if (kh * KSH + i < KS) {
                const half4_t ah = *reinterpret_cast<const half4_t *>(hrow + 16 * i);
                const half4_t al = *reinterpret_cast<const half4_t *>(lrow + 16 * i);
                am = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh[i], am, 0, 0, 0);
                ac = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl[i], ac, 0, 0, 0);
                ac = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh[i], ac, 0, 0, 0);
            }
        }
        floatx4 d = am + ac * (1.0f / 2048.0f);
        if constexpr (KP > 1) {
            if (kh == 1) red[wc * 64 + lane] = d;
            __syncthreads();
            if (kh == 0) d = d + red[wc * 64 + lane];
        }
        if (kh == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v;
                if (row < nv) a.records[((int64_t)sid[row] * 4 + corner) * G + col] = d[v] * scale;
            }
        }
        __syncthreads();
    }
}

// One lane (W = 1): one thread per node, FP64 recurrence, dot products against the four corner cells' coefficients
// staged in LDS (a split FP16 plane is recombined and scaled back to FP32 there).
template <int L, bool HALF>
__global__ __launch_bounds__(256) void sh_records_scalar(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    __shared__ float c[4][K];
    __shared__ double aa[K], bb[K];
    const int64_t item = blockIdx.x;
    const int base = a.item_cell[item];
    for (int e = threadIdx.x; e < 4 * K; e += 256) {
        const int cell = corner_cell(base, e / K, a.nc), k = e % K;
        const int64_t i = (int64_t)cell * K + k;
        if constexpr (HALF) {
            const int64_t j = (int64_t)cell * ((K + 15) / 16 * 16) + k;  // quad layout with one lane
            c[e / K][k] = (plane_at<L, true>(a.plane, j) + plane_at<L, true>(a.plane, j + a.plane_lo) * (1.0f / 2048.0f)) *
                          a.scale[cell];
        }
        else
            c[e / K][k] = plane_at<L, false>(a.plane, i);
    }
    stage_constants<L>(aa, bb, 256);
    __syncthreads();
    for (int64_t r = a.item_v0[item] + threadIdx.x; r < a.item_v1[item]; r += 256) {
        const int id = a.node[r];
        const double mu = a.node_mu[r];
        const float cp = a.node_cp[r];
        const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
        float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
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
            float acc[4];
            float z = (float)(mult * z0) * cm;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = z * c[q][o];
            if (m < L) {
                double z1 = (sqrt(2.0 * m + 3.0) * mu) * z0;
                z = (float)(mult * z1) * cm;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = acc[q] + z * c[q][o + 1];
                for (int l = m + 2; l <= L; ++l) {
                    const int k = kidx(l, m);
                    const double z2 = (aa[k] * mu) * z1 - bb[k] * z0;
                    z = (float)(mult * z2) * cm;
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = acc[q] + z * c[q][o + l - m];
                    z0 = z1;
                    z1 = z2;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) sum[q] = sum[q] + acc[q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) a.records[(int64_t)id * 4 + q] = sum[q];
    }
}

template <int G, int L>
static int sh_launch(const ShArgs &a, bool half, hipStream_t s) {
    if (a.items <= 0) return SOS_OK;
    if constexpr (G == 1) {
        if (half) hipLaunchKernelGGL((sh_records_scalar<L, true>), dim3((unsigned)a.items), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((sh_records_scalar<L, false>), dim3((unsigned)a.items), dim3(256), 0, s, a);
    } else {
        if (half)
            hipLaunchKernelGGL((sh_records_split_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16 * SplitK<L>::KP),
                               0, s, a);
        else if (a.tile)
            hipLaunchKernelGGL((sh_records_kernel<G, L, true>), dim3((unsigned)a.items, G / 16),
                               dim3(ShShape<G, L, true>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((sh_records_kernel<G, L, false>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
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
