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

// Recurrence constants (arithmetic.py:69-83) in FP64, rounded to T, staged once per workgroup.
template <int L, typename T>
__device__ inline void stage_constants(T *aa, T *bb, int nt) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    for (int k = threadIdx.x; k < K; k += nt) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) ++l;
        const int m = k - l * (l + 1) / 2;
        const double fl = l, fm = m;
        aa[k] = l > m ? (T)sqrt((4.0 * fl * fl - 1.0) / (fl * fl - fm * fm)) : (T)0;
        bb[k] = l > m + 1 ? (T)sqrt((2.0 * fl + 1.0) * ((fl - 1.0) * (fl - 1.0) - fm * fm) /
                                    ((2.0 * fl - 3.0) * (fl * fl - fm * fm)))
                          : (T)0;
    }
}

// dd[m] = -sqrt((2m + 1) / 2m), the factor of the diagonal seed chain (arithmetic.py:95-97), rounded to T.
template <int L, typename T>
__device__ inline void stage_dd(T *dd, int nt) {
    for (int m = threadIdx.x; m <= L; m += nt) dd[m] = m ? (T)-sqrt((2.0 * m + 1) / (2.0 * m)) : (T)0;
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

// The same in FP32 for the split path: threads t < V run the seed chain of node t, threads V <= t < 2V the cosine
// chain of node t - V.
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

// Basis rows of V nodes into LDS, tasks first, first + stride, ... below last. Task (node, p) runs the Legendre
// recurrence over l for the two orders m = p and m = L - p side by side (their lengths add up to L + 2, so tasks
// are balanced and each thread has two independent dependency chains): a loop over both chains while the short one
// lasts (unrolled UB times), then the long one alone (unrolled UT times). Each value goes to sink(v, k, z). With
// T = double (the reference's FP64 basis) z = fp32(x P) cos(m phi), x = sqrt 2 for m > 0, each basis value rounded
// to FP32 before the product; with T = float (the split path) z = P (x cos(m phi)) in FP32.
template <int L, int V, int UB, int UT, typename T, typename Sink>
__device__ inline void build_rows(const T *mus, const T (*seed)[V], const float (*cmv)[V], const T *aa, const T *bb,
                                  int nv, int first, int stride, int last, Sink sink) {
    for (int task = first; task < last; task += stride) {
        const int v = task % V, p = task / V;
        const int ma = p, mb = L - p;
        const bool two = mb != ma;
        const int oa = moff<L>(ma), ob = moff<L>(mb);
        if (v >= nv) {
            for (int l = ma; l <= L; ++l) sink(v, oa + l - ma, 0.0f);
            if (two)
                for (int l = mb; l <= L; ++l) sink(v, ob + l - mb, 0.0f);
            continue;
        }
        const T mu = mus[v];
        const float ca = cmv[ma][v], cb = cmv[mb][v];
        const float cxa = ma ? 1.41421356237309515f * ca : ca, cxb = 1.41421356237309515f * cb;
        const double xa = ma ? 1.41421356237309515 : 1.0, xb = 1.41421356237309515;
        auto za = [&](T pv) -> float {
            if constexpr (sizeof(T) == 8) return (float)(xa * pv) * ca;
            else return pv * cxa;
        };
        auto zb = [&](T pv) -> float {
            if constexpr (sizeof(T) == 8) return (float)(xb * pv) * cb;
            else return pv * cxb;
        };
        // chain a: m = p (long); chain b: m = L - p (short, length p + 1)
        T a0 = seed[ma][v], b0 = seed[mb][v];
        sink(v, oa, za(a0));
        if (two) sink(v, ob, zb(b0));
        T a1 = (sqrt((T)(2 * ma + 3)) * mu) * a0;  // ma < L always
        sink(v, oa + 1, za(a1));
        T b1 = 0;
        if (two && mb < L) {
            b1 = (sqrt((T)(2 * mb + 3)) * mu) * b0;
            sink(v, ob + 1, zb(b1));
        }
        int la = ma + 2, lb = mb + 2;
#pragma unroll UB
        for (; two && lb <= L; ++la, ++lb) {
            const int ka = kidx(la, ma), kb = kidx(lb, mb);
            const T a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            const T b2 = (aa[kb] * mu) * b1 - bb[kb] * b0;
            sink(v, oa + la - ma, za(a2));
            sink(v, ob + lb - mb, zb(b2));
            a0 = a1; a1 = a2;
            b0 = b1; b1 = b2;
        }
#pragma unroll UT
        for (; la <= L; ++la) {
            const int ka = kidx(la, ma);
            const T a2 = (aa[ka] * mu) * a1 - bb[ka] * a0;
            sink(v, oa + la - ma, za(a2));
            a0 = a1; a1 = a2;
        }
    }
}

// Corner c of a node with base cell b (ir * nc + j): b, b + 1, b + nc, b + nc + 1, the transport's corner order.
__device__ inline int corner_cell(int b, int c, int nc) { return b + (c & 1) + (c >> 1) * nc; }

// FP32 MFMA products of one round: d = Z rows (16 nodes) x this wavefront's B operand (b[4 i + u] = plane column at
// k = 16 i + 4 (lane / 16) + u), four independent accumulators. The A operands (one ds_read_b128 per group of four
// products) are read D groups ahead, and the schedule is pinned to (four products, one read) per group so that the
// LDS latency stays behind the products in flight.
template <int KQ, int ZS, int D>
__device__ inline floatx4 sh_products(const float *Zt, const float *b, int lane) {
    floatx4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float *zrow = Zt + (lane & 15) * ZS + 4 * (lane >> 4);
    floatx4 av[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        av[j] = *reinterpret_cast<const floatx4 *>(zrow + 16 * j);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // reads of groups 0 .. D - 1 first
    }
#pragma unroll
    for (int i = 0; i < KQ; ++i) {
        const floatx4 cur = av[i % D];
        if (i + D < KQ) av[i % D] = *reinterpret_cast<const floatx4 *>(zrow + 16 * (i + D));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[u], b[4 * i + u], acc[u], 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);                  // the four products of group i
        if (i + D < KQ) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // then the read of group i + D
    }
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// B operand of wavefront (corner cell, 16-lane tile): b[4 i + u] = c[cell][k][col], k = 16 i + 4 (lane / 16) + u.
template <int G, int L, int KQ>
__device__ inline void load_b(float *b, const void *plane, int cell, int col, int lane) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    const float *pl = reinterpret_cast<const float *>(plane) + (int64_t)cell * K * G + col;
#pragma unroll
    for (int i = 0; i < KQ; ++i)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = 16 * i + 4 * (lane >> 4) + u;
            b[4 * i + u] = k < K ? pl[(int64_t)k * G] : 0.0f;
        }
}

// FP32 SH records, one workgroup per work item (a run of nodes of one base cell), 4 x G/16 wavefronts: wavefront
// (c, nt) holds corner c's coefficients of lanes 16 nt .. 16 nt + 15 in registers (whole K, about 250 VGPRs, two
// wavefronts per SIMD, one workgroup per CU). Two forms, a tuning choice (the records are bitwise the same):
//   sh_records_kernel (flow): per round of 16 nodes, the threads t < 16 (L / 2 + 1) build the basis tile while the
//     last wavefront computes the seeds of the next round, one barrier, then the products of all wavefronts. Tiles,
//     seeds and node ids are multi-buffered, so a wavefront that finishes its products starts building the next tile
//     at once and the others' products overlap it;
//   sh_records_step_kernel: seeds, barrier, build, barrier, products, barrier.
template <int G, int L>
__global__ __launch_bounds__(G * 16) __attribute__((amdgpu_waves_per_eu(2, 2)))
void sh_records_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;   // padded row
    constexpr int KQ = KZ / 16;              // groups of four MFMA k-steps (one ds_read_b128 each)
    constexpr int NWB = G / 16;              // 16-lane tiles
    constexpr int NT = 64 * 4 * NWB;
    constexpr int V = 16;                    // nodes per round (one M-tile)
    constexpr int ZS = KZ + 4;               // row stride = 4 (mod 32 banks): 16 rows spread over the LDS banks
    constexpr int SW = NT - 64;              // first thread of the seed wavefront
    constexpr int TASKS = V * (L / 2 + 1);
    __shared__ float Z[2][V][ZS];
    __shared__ double aa[K], bb[K];
    __shared__ double seed[2][L + 1][V];
    __shared__ float cmv[2][L + 1][V];
    __shared__ double smu[2][V];
    __shared__ double dd[L + 1];
    __shared__ int sid[3][V];  // node ids of round r in sid[r % 3]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NWB, nt = wave % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    const bool seeder = tid >= SW && tid < SW + V;
    stage_constants<L>(aa, bb, NT);
    stage_dd<L>(dd, NT);
    for (int e = tid; e < 2 * V * (KZ - K); e += NT) {
        const int q = e / (KZ - K);
        Z[q / V][q % V][K + e % (KZ - K)] = 0.0f;
    }
    float b[4 * KQ];
    load_b<G, L, KQ>(b, a.plane, cell, col, lane);
    __syncthreads();  // dd
    // seeds of round 0; node (id, mu, cos phi) of round 1 held in registers by the seed threads
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (seeder) {
        const int v = tid - SW;
        const bool in = v0 + v < v1;
        const double mu = in ? (double)a.node_mu[v0 + v] : 0.0;
        smu[0][v] = mu;
        sid[0][v] = in ? a.node[v0 + v] : -1;
        visit_seeds<L, V>(v, mu, in ? a.node_cp[v0 + v] : 1.0f, dd, seed[0], cmv[0]);
        if (v0 + V + v < v1) {
            nid = a.node[v0 + V + v];
            nmu = a.node_mu[v0 + V + v];
            ncp = a.node_cp[v0 + V + v];
        }
    }
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
        const int cur = r & 1, nxt = cur ^ 1;
        const int64_t r0 = v0 + (int64_t)r * V;
        const int nv = (int)min((int64_t)V, v1 - r0);
        // Z[cur] was last read by the products of round r - 2, seed[cur] by the build of round r - 2: both before
        // the barrier of round r - 1
        float (*Zc)[ZS] = Z[cur];
        build_rows<L, V, 2, 4, double>(smu[cur], seed[cur], cmv[cur], aa, bb, nv, tid, NT, TASKS,
                                       [&](int v, int k, float z) { Zc[v][k] = z; });
        if (seeder && r + 1 < rounds) {
            const int v = tid - SW;
            const bool in = r0 + V + v < v1;
            const double mu = in ? (double)nmu : 0.0;
            smu[nxt][v] = mu;
            sid[(r + 1) % 3][v] = in ? nid : -1;
            visit_seeds<L, V>(v, mu, in ? ncp : 1.0f, dd, seed[nxt], cmv[nxt]);
            const int64_t nx = r0 + 2 * V + v;
            if (nx < v1) {
                nid = a.node[nx];
                nmu = a.node_mu[nx];
                ncp = a.node_cp[nx];
            }
        }
        __syncthreads();
        const floatx4 d = sh_products<KQ, ZS, 3>(&Z[cur][0][0], b, lane);
        const int *ids = sid[r % 3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = d[v];
        }
    }
}

template <int G, int L>
__global__ __launch_bounds__(G * 16) __attribute__((amdgpu_waves_per_eu(2, 2)))
void sh_records_step_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KQ = KZ / 16;
    constexpr int NWB = G / 16;
    constexpr int NT = 64 * 4 * NWB;
    constexpr int V = 16;
    constexpr int ZS = KZ + 4;
    constexpr int TASKS = V * (L / 2 + 1);
    __shared__ float Z[V][ZS];
    __shared__ double aa[K], bb[K];
    __shared__ double seed[L + 1][V];
    __shared__ float cmv[L + 1][V];
    __shared__ double smu[V];
    __shared__ double dd[L + 1];
    __shared__ int sid[V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NWB, nt = wave % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_constants<L>(aa, bb, NT);
    stage_dd<L>(dd, NT);
    for (int e = tid; e < V * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = 0.0f;
    float b[4 * KQ];
    load_b<G, L, KQ>(b, a.plane, cell, col, lane);
    // node data of the next round is loaded one round ahead (contiguous in the sorted order)
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (tid < V && v0 + tid < v1) {
        nid = a.node[v0 + tid];
        nmu = a.node_mu[v0 + tid];
        ncp = a.node_cp[v0 + tid];
    }
    __syncthreads();  // dd
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
        build_rows<L, V, 2, 4, double>(smu, seed, cmv, aa, bb, nv, tid, NT, TASKS,
                                       [&](int v, int k, float z) { Z[v][k] = z; });
        __syncthreads();
        const floatx4 d = sh_products<KQ, ZS, 3>(&Z[0][0], b, lane);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)sid[row] * 4 + corner) * G + col] = d[v];
        }
        __syncthreads();
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

__device__ inline void split_store(_Float16 *hi, _Float16 *lo, float z) {
    const _Float16 h = (_Float16)z;
    *hi = h;
    *lo = (_Float16)((z - (float)h) * 2048.0f);
}

// Split-FP16 products of one round: three independent accumulators (hi hi, hi lo, lo hi) so that no product waits
// on the one before it, the hi / lo A operands read D k-steps ahead with the schedule pinned to (three products,
// two reads) per k-step.
template <int KS, int D>
__device__ inline floatx4 split_products(const _Float16 *hrow, const _Float16 *lrow, const half4_t *bh,
                                         const half4_t *bl) {
    floatx4 am = floatx4{0.f, 0.f, 0.f, 0.f}, a1 = am, a2 = am;
    half4_t ph[D], pl[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        ph[j] = *reinterpret_cast<const half4_t *>(hrow + 16 * j);
        pl[j] = *reinterpret_cast<const half4_t *>(lrow + 16 * j);
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
#pragma unroll
    for (int i = 0; i < KS; ++i) {
        const half4_t ah = ph[i % D], al = pl[i % D];
        if (i + D < KS) {
            ph[i % D] = *reinterpret_cast<const half4_t *>(hrow + 16 * (i + D));
            pl[i % D] = *reinterpret_cast<const half4_t *>(lrow + 16 * (i + D));
        }
        am = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh[i], am, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl[i], a1, 0, 0, 0);
        a2 = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh[i], a2, 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
        if (i + D < KS) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
    return am + (a1 + a2) * (1.0f / 2048.0f);
}

// Split-FP16 SH records in the flow form of sh_records_kernel: 4 x G/16 wavefronts, wavefront (corner, N-tile)
// holding the whole K range of its hi / lo B operands (144 VGPRs at L = 32 in the quad layout [cell][KZ/4][G][4],
// about 240 in all: two wavefronts per SIMD); the seed wavefront's first 2 x 16 threads run the seed and cosine
// chains of the next round.
template <int G, int L>
__global__ __launch_bounds__(G * 16) __attribute__((amdgpu_waves_per_eu(2, 2)))
void sh_records_split_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KS = KZ / 16;               // MFMA k-steps of 16
    constexpr int NW = G / 16;                // N-tiles
    constexpr int NT = 64 * 4 * NW;
    constexpr int V = 16;                     // nodes per round (one M-tile)
    constexpr int ZS = KZ + 8;                // halves: row stride = 4 (mod 32 banks) in 32-bit words
    constexpr int SW = NT - 64;
    constexpr int TASKS = V * (L / 2 + 1);
    __shared__ _Float16 Zh[2][V][ZS], Zl[2][V][ZS];
    __shared__ float aa[K], bb[K];
    __shared__ float seed[2][L + 1][V];
    __shared__ float cmv[2][L + 1][V];
    __shared__ float smu[2][V];
    __shared__ float dd[L + 1];
    __shared__ int sid[3][V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NW, nt = wave % NW;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    const bool seeder = tid >= SW && tid < SW + 2 * V;
    stage_constants<L>(aa, bb, NT);
    stage_dd<L>(dd, NT);
    for (int e = tid; e < 2 * V * (KZ - K); e += NT) {
        const int q = e / (KZ - K);
        Zh[q / V][q % V][K + e % (KZ - K)] = (_Float16)0.0f;
        Zl[q / V][q % V][K + e % (KZ - K)] = (_Float16)0.0f;
    }
    const float scale = a.scale[(int64_t)cell * G + col];
    // B operands from the quad layout: bh[i], bl[i] hold K positions 16 i + 4 (lane / 16) + 0..3 of column col
    half4_t bh[KS], bl[KS];
    {
        const half4_t *ph = reinterpret_cast<const half4_t *>(a.plane) + (int64_t)cell * (KZ / 4) * G + col;
        const half4_t *pl = ph + a.plane_lo / 4;
#pragma unroll
        for (int i = 0; i < KS; ++i) {
            const int q = 4 * i + (lane >> 4);
            bh[i] = ph[(int64_t)q * G];
            bl[i] = pl[(int64_t)q * G];
        }
    }
    __syncthreads();  // dd
    int nid = -1;
    float nmu = 0.0f, ncp = 1.0f;
    if (seeder) {
        const int t = tid - SW, v = t % V;
        const bool in = v0 + v < v1;
        const float mu = in ? a.node_mu[v0 + v] : 0.0f;
        if (t < V) {
            smu[0][v] = mu;
            sid[0][v] = in ? a.node[v0 + v] : -1;
        }
        visit_seeds32<L, V>(t, mu, in ? a.node_cp[v0 + v] : 1.0f, dd, seed[0], cmv[0]);
        if (v0 + V + v < v1) {
            nid = a.node[v0 + V + v];
            nmu = a.node_mu[v0 + V + v];
            ncp = a.node_cp[v0 + V + v];
        }
    }
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
        const int cur = r & 1, nxt = cur ^ 1;
        const int64_t r0 = v0 + (int64_t)r * V;
        const int nv = (int)min((int64_t)V, v1 - r0);
        _Float16 (*Hc)[ZS] = Zh[cur];
        _Float16 (*Lc)[ZS] = Zl[cur];
        build_rows<L, V, 1, 1, float>(smu[cur], seed[cur], cmv[cur], aa, bb, nv, tid, NT, TASKS,
                                      [&](int v, int k, float z) { split_store(&Hc[v][k], &Lc[v][k], z); });
        if (seeder && r + 1 < rounds) {
            const int t = tid - SW, v = t % V;
            const bool in = r0 + V + v < v1;
            const float mu = in ? nmu : 0.0f;
            if (t < V) {
                smu[nxt][v] = mu;
                sid[(r + 1) % 3][v] = in ? nid : -1;
            }
            visit_seeds32<L, V>(t, mu, in ? ncp : 1.0f, dd, seed[nxt], cmv[nxt]);
            const int64_t nx = r0 + 2 * V + v;
            if (nx < v1) {
                nid = a.node[nx];
                nmu = a.node_mu[nx];
                ncp = a.node_cp[nx];
            }
        }
        __syncthreads();
        const floatx4 d = split_products<KS, 3>(&Zh[cur][lane & 15][4 * (lane >> 4)],
                                                &Zl[cur][lane & 15][4 * (lane >> 4)], bh, bl);
        const int *ids = sid[r % 3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = d[v] * scale;
        }
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
            hipLaunchKernelGGL((sh_records_split_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
        else if (a.tile == 1)
            hipLaunchKernelGGL((sh_records_step_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
        else hipLaunchKernelGGL((sh_records_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
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
