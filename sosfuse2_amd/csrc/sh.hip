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

// dd[m] = -sqrt((2m + 1) / 2m), the factor of the diagonal seed chain (arithmetic.py:95-97), rounded to T.
template <int L, typename T>
__device__ inline void stage_dd(T *dd, int nt) {
    for (int m = threadIdx.x; m <= L; m += nt) dd[m] = m ? (T)-sqrt((2.0 * m + 1) / (2.0 * m)) : (T)0;
}

// Recurrence constants at the m-major positions: ab[moff(m) + l - m] = (aa, bb) of (l, m) (arithmetic.py:69-83, FP64;
// zero for l < m + 2), so that a chain of one order walks ab with its output position.
template <int L>
__device__ inline void stage_ab(double2 *ab, int nt) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    for (int k = threadIdx.x; k < K; k += nt) {
        int m = 0;
        while (moff<L>(m + 1) <= k) ++m;
        const int l = m + (k - moff<L>(m));
        const double fl = l, fm = m;
        const double x = l > m ? sqrt((4.0 * fl * fl - 1.0) / (fl * fl - fm * fm)) : 0.0;
        const double y = l > m + 1 ? sqrt((2.0 * fl + 1.0) * ((fl - 1.0) * (fl - 1.0) - fm * fm) /
                                          ((2.0 * fl - 3.0) * (fl * fl - fm * fm)))
                                   : 0.0;
        ab[k] = double2{x, y};
    }
}

// The seeds x_m P_mm(mu) of every order (x_m = sqrt 2 for m > 0): the FP64 diagonal chain of arithmetic.py:95-97
// (1 / sqrt(4 pi) times the products of dd[j] sin theta), times x_m; seed(m, value) for m = 0 .. L.
template <int L, typename Out>
__device__ inline void seeds_x(double mu, const double *dd, Out seed) {
    const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
    double diag = 0.28209479177387814;
    seed(0, diag);
    for (int m = 1; m <= L; ++m) {
        diag = (diag * dd[m]) * st;
        seed(m, 1.41421356237309515 * diag);
    }
}

// The basis values x_m P_lm(mu) of the orders ma = p and mb = L - p (p < L / 2; p = L / 2: the one order) from
// their seeds sa, sb (seeds_x), by the FP64 recurrence P_l = (aa mu) P_(l-1) - bb P_(l-2) (arithmetic.py:98-109),
// side by side while both last, then the longer alone; every value rounded to FP32 goes to sink_a(k, z) (order ma)
// or sink_b(k, z) (order mb) at its m-major position k. The two lengths add up to L + 2 for every p, so tasks over
// p are balanced, and positions advance by one per step (no index arithmetic per value).
template <int L, typename SinkA, typename SinkB>
__device__ inline void basis_pair(int p, double mu, double sa, double sb, const double2 *__restrict__ ab, SinkA sink_a,
                                  SinkB sink_b) {
    const int ma = p, mb = L - p;
    const bool two = mb != ma;
    int ka = moff<L>(ma), kb = moff<L>(mb);
    double a0 = sa, b0 = sb;
    sink_a(ka, (float)a0);
    if (two) sink_b(kb, (float)b0);
    double a1 = (sqrt((double)(2 * ma + 3)) * mu) * a0;  // ma < L always
    sink_a(ka + 1, (float)a1);
    double b1 = 0.0;
    if (two && mb < L) {
        b1 = (sqrt((double)(2 * mb + 3)) * mu) * b0;
        sink_b(kb + 1, (float)b1);
    }
    ka += 2;
    kb += 2;
    const int na = L - ma - 1, nb = two ? L - mb - 1 : 0;  // values still to come in each chain (l = m + 2 .. L)
    int i = 0;
    for (; i < nb; ++i, ++ka, ++kb) {
        const double2 ca = ab[ka], cb = ab[kb];
        const double a2 = (ca.x * mu) * a1 - ca.y * a0;
        const double b2 = (cb.x * mu) * b1 - cb.y * b0;
        sink_a(ka, (float)a2);
        sink_b(kb, (float)b2);
        a0 = a1; a1 = a2;
        b0 = b1; b1 = b2;
    }
    for (; i < na; ++i, ++ka) {
        const double2 ca = ab[ka];
        const double a2 = (ca.x * mu) * a1 - ca.y * a0;
        sink_a(ka, (float)a2);
        a0 = a1; a1 = a2;
    }
}

// Corner c of a node with base cell b (ir * nc + j): b, b + 1, b + nc, b + nc + 1, the transport's corner order.
__device__ inline int corner_cell(int b, int c, int nc) { return b + (c & 1) + (c >> 1) * nc; }

// Basis table of a bundle: one workgroup per 16 rays builds their rows fp32(x_m P_lm(mu)) with the FP64 recurrence
// (basis_pair over the tasks (ray, p)) in LDS and writes them out row by row, coalesced.
template <int L>
__global__ __launch_bounds__(256) void plan_basis_kernel(int64_t rays, const double *dirs, float *basis) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int V = 16;
    constexpr int NT = 256;
    __shared__ __attribute__((aligned(16))) float Z[V][KZ];
    __shared__ double2 ab[K];
    __shared__ double seed[L + 1][V];
    __shared__ double smu[V];
    __shared__ double dd[L + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * V;
    const int nv = (int)min((int64_t)V, rays - r0);
    stage_ab<L>(ab, NT);
    stage_dd<L>(dd, NT);
    for (int e = tid; e < V * (KZ - K); e += NT) Z[e / (KZ - K)][K + e % (KZ - K)] = 0.0f;
    __syncthreads();
    if (tid < V) {
        const double mu = tid < nv ? fmin(1.0, fmax(-1.0, dirs[3 * (r0 + tid) + 2])) : 0.0;
        smu[tid] = mu;
        seeds_x<L>(mu, dd, [&](int m, double x) { seed[m][tid] = x; });
    }
    __syncthreads();
    for (int task = tid; task < V * (L / 2 + 1); task += NT) {
        const int v = task % V, p = task / V;
        auto put = [&](int k, float z) { Z[v][k] = z; };
        basis_pair<L>(p, smu[v], seed[p][v], seed[L - p][v], ab, put, put);
    }
    __syncthreads();
    for (int e = tid; e < nv * (KZ / 4); e += NT) {
        const int v = e / (KZ / 4), q = e % (KZ / 4);
        reinterpret_cast<float4 *>(basis + (r0 + v) * KZ)[q] = reinterpret_cast<const float4 *>(&Z[v][0])[q];
    }
}

int plan_basis_launch(int L, int64_t rays, const double *dirs, float *basis, hipStream_t s) {
    const dim3 grid((unsigned)((rays + 15) / 16));
    switch (L) {
        case 8: hipLaunchKernelGGL(plan_basis_kernel<8>, grid, dim3(256), 0, s, rays, dirs, basis); break;
        case 16: hipLaunchKernelGGL(plan_basis_kernel<16>, grid, dim3(256), 0, s, rays, dirs, basis); break;
        case 24: hipLaunchKernelGGL(plan_basis_kernel<24>, grid, dim3(256), 0, s, rays, dirs, basis); break;
        case 32: hipLaunchKernelGGL(plan_basis_kernel<32>, grid, dim3(256), 0, s, rays, dirs, basis); break;
        default: return set_error("plan_basis: L %d", L);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("plan_basis launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

// Order m of every position k' of the m-major K order (0 in the padding K <= k' < KZ).
template <int L, int KZ>
__device__ inline void stage_km(unsigned char *km, int nt) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    for (int k = threadIdx.x; k < KZ; k += nt) {
        int m = 0;
        if (k < K)
            while (moff<L>(m + 1) <= k) ++m;
        km[k] = (unsigned char)m;
    }
}

// cos(m phi), m = 0 .. L, by the FP32 Chebyshev recurrence (arithmetic.py:112-119).
// With scale a power of two the values are cos(m phi) scale: the recurrence is linear, so running it from scale and
// cp scale gives the scaled values of the unscaled recurrence exactly.
template <int L>
__device__ inline void cos_chain(float cp, float *c, float scale = 1.0f) {
    float c0 = scale, c1 = cp * scale;
    c[0] = c0;
    c[1] = c1;
    for (int m = 2; m <= L; ++m) {
        const float cm = (2.0f * cp) * c1 - c0;
        c0 = c1;
        c1 = cm;
        c[m] = cm;
    }
}

// Per-round node data, prepared two rounds ahead by 16 threads: the node ids (record rows), the rays (basis rows)
// and cos(m phi) of the round's nodes, in slot r mod S.
template <int L, int V, int S = 4>
struct RoundMeta {
    int id[S][V];
    int ray[S][V];
    float cm[S][V][L + 1];
};

// Thread v < V: node v of the round starting at sorted position r0 into slot q (id -1 and ray -1 past v1).
template <int L, int V, int S>
__device__ inline void meta_fill(RoundMeta<L, V, S> &mt, int q, int v, bool in, int id, int ray, float cp,
                                 float scale = 1.0f) {
    mt.id[q][v] = in ? id : -1;
    mt.ray[q][v] = in ? ray : -1;
    cos_chain<L>(in ? cp : 1.0f, mt.cm[q][v], scale);
}

// Basis rows of the round in slot q, PT float4 chunks per thread (chunk c = tid + j NT: row c / (KZ/4)), read into
// registers; rows of absent nodes read as 0.
template <int V, int KZ, int NT, int PT, typename Meta>
__device__ inline void fetch_rows(float4 *pr, const float *basis, const Meta &mt, int q, int tid) {
#pragma unroll
    for (int j = 0; j < PT; ++j) {
        const int c = tid + j * NT;
        pr[j] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        if (c < V * (KZ / 4)) {
            const int ray = mt.ray[q][c / (KZ / 4)];
            if (ray >= 0) pr[j] = reinterpret_cast<const float4 *>(basis + (int64_t)ray * KZ)[c % (KZ / 4)];
        }
    }
}

// Z[v][k] = basis row [k] cos(m(k) phi_v), four positions per chunk, through sink(v, q, z4) (q = k / 4).
template <int V, int KZ, int NT, int PT, typename Meta, typename Sink>
__device__ inline void rows_to_z(const float4 *pr, const Meta &mt, int q, const unsigned char *km, int tid, Sink sink) {
#pragma unroll
    for (int j = 0; j < PT; ++j) {
        const int c = tid + j * NT;
        if (c < V * (KZ / 4)) {
            const int v = c / (KZ / 4), k4 = c % (KZ / 4);
            const uchar4 m = reinterpret_cast<const uchar4 *>(km)[k4];
            const float *cm = mt.cm[q][v];
            sink(v, k4, float4{pr[j].x * cm[m.x], pr[j].y * cm[m.y], pr[j].z * cm[m.z], pr[j].w * cm[m.w]});
        }
    }
}

// FP32 MFMA products of one round: d = Z rows (16 nodes) x this wavefront's B operand (b[4 i + u] = plane column at
// k = 16 i + 4 (lane / 16) + u). The K range is taken as two halves (groups 0 .. KQH - 1 and KQH .. KQ - 1, a group
// being four MFMA k-steps) processed side by side, each into two accumulators (even and odd u), and the result is
// (half 0) + (half 1): the same order as the quad form, whose wavefronts each hold one half, so every form gives
// the same records bit for bit. The A operands (one ds_read_b128 per group) are read D group pairs ahead, with the
// schedule pinned to (eight products, two reads) per pair so that the LDS latency stays behind the products.
template <int KQ, int ZS, int D>
__device__ inline floatx4 sh_products(const float *Zt, const float *b, int lane) {
    constexpr int KQH = (KQ + 1) / 2;
    floatx4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float *zrow = Zt + (lane & 15) * ZS + 4 * (lane >> 4);
    floatx4 av[D][2];
#pragma unroll
    for (int j = 0; j < D && j < KQH; ++j) {
        av[j][0] = *reinterpret_cast<const floatx4 *>(zrow + 16 * j);
        if (KQH + j < KQ) av[j][1] = *reinterpret_cast<const floatx4 *>(zrow + 16 * (KQH + j));
        if (KQH + j < KQ) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // reads of pairs 0 .. D - 1 first
        else __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
#pragma unroll
    for (int i = 0; i < KQH; ++i) {
        const floatx4 c0 = av[i % D][0], c1 = av[i % D][1];
        if (i + D < KQH) {
            av[i % D][0] = *reinterpret_cast<const floatx4 *>(zrow + 16 * (i + D));
            if (KQH + i + D < KQ) av[i % D][1] = *reinterpret_cast<const floatx4 *>(zrow + 16 * (KQH + i + D));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[u & 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0[u], b[4 * i + u], acc[u & 1], 0, 0, 0);
            if (KQH + i < KQ)
                acc[2 + (u & 1)] =
                    __builtin_amdgcn_mfma_f32_16x16x4f32(c1[u], b[4 * (KQH + i) + u], acc[2 + (u & 1)], 0, 0, 0);
        }
        if (KQH + i < KQ) __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);  // the products of pair i
        else __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        if (i + D < KQH) {  // then the reads of pair i + D
            if (KQH + i + D < KQ) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
            else __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
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

// Node data of the round starting at sorted position r0, node v (for meta_fill).
struct NodeRef {
    int id, ray;
    float cp;
};

__device__ inline NodeRef node_ref(const ShArgs &a, int64_t r, int64_t v1) {
    NodeRef n{-1, -1, 1.0f};
    if (r < v1) {
        n.id = a.node[r];
        n.ray = a.node_ray[r];
        n.cp = a.node_cp[r];
    }
    return n;
}

// FP32 SH records, one workgroup per work item (a run of nodes of one base cell), 4 x G/16 wavefronts: wavefront
// (c, nt) holds corner c's coefficients of lanes 16 nt .. 16 nt + 15 in registers (whole K, about 250 VGPRs, two
// wavefronts per SIMD, one workgroup per CU). Per round of 16 nodes the basis tile is the nodes' basis rows (read
// from the bundle's table a round ahead) times their cos(m phi). Three forms, a tuning choice (the records are
// bitwise the same):
//   sh_records_kernel (flow): tile r is formed from the rows fetched during round r - 1, the rows of round r + 1 are
//     fetched, 16 threads of the last wavefront prepare the node data of round r + 2, one barrier, then the
//     products of all wavefronts; tiles are double-buffered and node data four-buffered, so the fetches run under
//     the products and a wavefront that finishes its products goes on to the next tile at once;
//   sh_records_step_kernel: node data, barrier, fetch and tile, barrier, products, barrier;
//   sh_records_quad_kernel: the flow's rounds on twice the wavefronts, each holding half of the K range.
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
    constexpr int SW = NT - 64;              // first thread of the node-data wavefront
    constexpr int PT = (V * KZ / 4 + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) float Z[2][V][ZS];
    __shared__ RoundMeta<L, V> mt;
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NWB, nt = wave % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    const bool meta = tid >= SW && tid < SW + V;
    stage_km<L, KZ>(km, NT);
    float b[4 * KQ];
    load_b<G, L, KQ>(b, a.plane, cell, col, lane);
    // node data of rounds 0 and 1 (threads t < 2V); that of round 2 held in registers by the node-data threads
    if (tid < 2 * V) {
        const NodeRef n = node_ref(a, v0 + tid, v1);
        meta_fill<L, V>(mt, tid / V, tid % V, n.id >= 0, n.id, n.ray, n.cp);
    }
    NodeRef nn{-1, -1, 1.0f};
    if (meta) nn = node_ref(a, v0 + 2 * V + (tid - SW), v1);
    __syncthreads();
    float4 pr[PT];
    fetch_rows<V, KZ, NT, PT>(pr, a.basis, mt, 0, tid);
    for (int r = 0; r < rounds; ++r) {
        const int cur = r & 1;
        const int64_t r0 = v0 + (int64_t)r * V;
        const int nv = (int)min((int64_t)V, v1 - r0);
        // Z[cur] was last read by the products of round r - 2 (before the barrier of round r - 1)
        float (*Zc)[ZS] = Z[cur];
        rows_to_z<V, KZ, NT, PT>(pr, mt, r & 3, km, tid, [&](int v, int k4, float4 z) {
            reinterpret_cast<float4 *>(&Zc[v][0])[k4] = z;
        });
        if (r + 1 < rounds) fetch_rows<V, KZ, NT, PT>(pr, a.basis, mt, (r + 1) & 3, tid);
        if (meta && r + 2 < rounds) {
            // slot (r + 2) & 3 last held round r - 2, read by the products of round r - 2 (before this round)
            const int v = tid - SW;
            meta_fill<L, V>(mt, (r + 2) & 3, v, nn.id >= 0, nn.id, nn.ray, nn.cp);
            nn = node_ref(a, r0 + 3 * V + v, v1);
        }
        __syncthreads();
        const floatx4 d = sh_products<KQ, ZS, 3>(&Z[cur][0][0], b, lane);
        const int *ids = mt.id[r & 3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = d[v];
        }
    }
}

// Products over the groups g0 .. g0 + ng - 1 only (one wavefront's half of the K range) into two accumulators (even
// and odd u), in the order of one half of sh_products. The steps past ng (an odd group count leaves the second half
// one short) read the half's last group again, finite values against a zero B: they add exactly nothing. The reads
// are unconditional (no select on ng), so each keeps its alias information.
template <int NG, int ZS, int D>
__device__ inline floatx4 sh_products_part(const float *Z, int tile, const float *b, int lane, int g0, int ng) {
    floatx4 acc[2];
    acc[0] = acc[1] = floatx4{0.f, 0.f, 0.f, 0.f};
    // every read indexes Z itself (one integer offset), which keeps the reads' LDS alias information
    const int zrow = tile + (lane & 15) * ZS + 4 * (lane >> 4) + 16 * g0;
    const int glast = ng - 1;
    floatx4 av[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        av[j] = *reinterpret_cast<const floatx4 *>(Z + (zrow + 16 * min(j, glast)));
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const floatx4 cur = av[i % D];
        if (i + D < NG) av[i % D] = *reinterpret_cast<const floatx4 *>(Z + (zrow + 16 * min(i + D, glast)));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            acc[u & 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[u], b[4 * i + u], acc[u & 1], 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        if (i + D < NG) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    return acc[0] + acc[1];
}

// Node data of the round whose first sorted position is first into slot (3 V ints): lanes 0 .. 3 V - 1 of one
// wavefront (loader), one LDS-direct 4-byte load each, lane l writing slot[l]; src is the lane's array (node ids,
// rays or cos phi bits, by lane / V). Positions past the item read its last node (absent nodes are flagged by
// position where the data is used).
template <int V>
__device__ inline void quad_fetch_nodes(int *slot, const int32_t *src, int64_t first, int64_t v1, bool loader, int lane) {
    if (loader) {
        const int64_t pos = min(first + lane % V, v1 - 1);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + pos),
                                         (__attribute__((address_space(3))) void *)slot, 4, 0, 0);
    }
}

// Basis rows of one round (rays[v] of its 16 nodes) straight into raw with LDS-direct 16-byte loads, in a row-block
// layout: wavefront w takes rows 4 (w mod 4) .. + 3 and, in pass j, the group of 16 chunk positions KG j + w / 4
// (KG = waves / 4); lane l the row 4 (w mod 4) + l / 16 and the chunk 16 group + l mod 16. The hardware writes lane
// l's 16 bytes at the wavefront's base + 16 l, so a wavefront's 64 chunks land at raw[(group 4 + w mod 4) 64 ..] and
// raw holds raw_groups(KZ) x 256 chunks; each thread later reads back only its own slot, so its own wait for the
// loads is all the ordering needed. Groups past the row are skipped (uniform per wavefront) and chunk positions past
// it read the row's last chunk into their own slot, which the tile never reads. Absent nodes (ray -1) read row 0:
// their products are never stored. Row, chunk and LDS offsets are constants per thread and pass (no division).
template <int KZ>
constexpr int raw_groups() { return (KZ / 4 + 15) / 16; }

template <int V, int KZ, int NT>
__device__ inline void quad_fetch_rows_rb(float4 *raw, const float *basis, const int *rays, int tid) {
    constexpr int KG = NT / 256, NGR = raw_groups<KZ>(), PASSES = (NGR + KG - 1) / KG;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6, rb = wave & 3;
    const int ray = max(rays[4 * rb + (lane >> 4)], 0);
    const float *row = basis + (int64_t)ray * KZ;
#pragma unroll
    for (int j = 0; j < PASSES; ++j) {
        const int grp = KG * j + (wave >> 2);
        if (grp < NGR) {
            const int k4 = min(16 * grp + (lane & 15), KZ / 4 - 1);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(row + 4 * k4),
                                             (__attribute__((address_space(3))) void *)&raw[(grp * 4 + rb) * 64], 16,
                                             0, 0);
        }
    }
}

// The rounds of the quad form (below). Every LDS array comes in as a restrict pointer, so that each LDS access
// carries alias-scope information: the compiler's wait insertion can then tell the reads of Z, the node data and the
// partials from the LDS-direct loads into raw and ndata still in flight, instead of waiting for those loads before
// every LDS read.
// STAGGER = false: every round forms tile r from the rows fetched in round r - 1, prepares node data two rounds
// ahead, fetches the rows of round r + 1, then (after the barrier) runs the products of round r.
// STAGGER = true: the two K halves run their rounds in opposite orders, one wavefront of each half per SIMD (a
// workgroup's wavefronts are placed on the SIMDs round robin): the second half (kh = 1) forms its chunks of tile r + 1
// at the start of round r and then runs the products of round r, the first half runs the products first and then
// forms its chunks, so on every SIMD tile formation (VALU) runs under products (MFMA). A thread refetches its own
// chunks of the rows as soon as it has formed them (one raw buffer, about a round to land), and node data is
// prepared three rounds ahead. The products, their order and the records are those of the unstaggered rounds.
template <int G, int L, bool STAGGER = false>
__device__ __attribute__((always_inline)) inline void quad_rounds(
    const ShArgs &a, float *__restrict__ Z, float4 *__restrict__ raw, int *__restrict__ ndata,
    RoundMeta<L, 16, 8> *__restrict__ mtp, const unsigned char *__restrict__ km, floatx4 *__restrict__ red,
    const float *b, int64_t v0, int64_t v1, int corner, int col, int kh, int wc, int g0, int ng) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KQ = KZ / 16;
    constexpr int KQH = (KQ + 1) / 2;
    constexpr int NWB = G / 16;
    constexpr int NC = 4 * NWB;
    constexpr int NT = 64 * NC * 2;
    constexpr int V = 16;
    constexpr int ZS = KZ + 4;
    constexpr int SW = NT - 64;
    constexpr int NG = KQH;
    constexpr int AHEAD = STAGGER ? 3 : 2;  // rounds between a round's node data and its tile
    RoundMeta<L, V, 8> &mt = *mtp;
    const int32_t *node = a.node, *node_ray = a.node_ray, *node_cp = reinterpret_cast<const int32_t *>(a.node_cp);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    // the node-data wavefront: the last one (first K half in the staggered rounds, which forms late)
    const int mwave = STAGGER ? NC - 1 : NT / 64 - 1;
    const bool meta = wave == mwave && lane < V;
    if (tid < AHEAD * V) {
        const NodeRef n = node_ref(a, v0 + tid, v1);
        meta_fill<L, V, 8>(mt, tid / V, tid % V, n.id >= 0, n.id, n.ray, n.cp);
    }
    const bool nlane = wave == mwave && lane < 3 * V;  // the node-data loaders
    const int64_t ndel = lane < V ? 0 : lane < 2 * V ? (int64_t)(node_ray - node) : (int64_t)(node_cp - node);
    if (rounds > AHEAD) quad_fetch_nodes<V>(ndata + (AHEAD & 3) * 3 * V, node + ndel, v0 + AHEAD * V, v1, nlane, lane);
    __syncthreads();
    const float *basis = a.basis;
    // per thread and pass: its chunk of four positions (row v, chunk k4 of the row-block layout), the cosine offsets
    // of the chunk's first order mA and other order mB (a chunk holds at most two orders; mB = mA for one) and the
    // positions that take mB (bit 4 j + e; positions past K read mA, their basis values are zero)
    constexpr int KG = NT / 256, NGR = raw_groups<KZ>(), PASSES = (NGR + KG - 1) / KG;
    const int rb = wave & 3, vrow = 4 * rb + (lane >> 4);
    int offa[PASSES], offb[PASSES];
    unsigned sel = 0;
#pragma unroll
    for (int j = 0; j < PASSES; ++j) {
        const int k4 = 16 * (KG * j + (tid >> 8)) + (lane & 15);
        int ma = 0, mb = 0;
        if (k4 < KZ / 4) {
            ma = km[4 * k4];
            mb = ma;
#pragma unroll
            for (int e = 1; e < 4; ++e)
                if (4 * k4 + e < K && km[4 * k4 + e] != ma) {
                    mb = km[4 * k4 + e];
                    sel |= 1u << (4 * j + e);
                }
        }
        offa[j] = vrow * (L + 1) + ma;
        offb[j] = vrow * (L + 1) + mb;
    }
    // this thread's chunks of tile q (into Z[q & 1]) from its chunks of the rows of round q in raw
    auto form = [&](int q) {
        const float *cmr = &mt.cm[q & 7][0][0];
        float *zr = Z + (q & 1) * V * ZS + vrow * ZS;
#pragma unroll
        for (int j = 0; j < PASSES; ++j) {
            const int grp = KG * j + (tid >> 8);
            const int k4 = 16 * grp + (lane & 15);
            if (grp < NGR && k4 < KZ / 4) {
                const float4 p = raw[(grp * 4 + rb) * 64 + lane];
                const float ca = cmr[offa[j]], cb = cmr[offb[j]];
                const unsigned sb = sel >> (4 * j);
                reinterpret_cast<float4 *>(zr)[k4] = float4{p.x * ca, p.y * ((sb & 2u) ? cb : ca),
                                                            p.z * ((sb & 4u) ? cb : ca), p.w * ((sb & 8u) ? cb : ca)};
            }
        }
    };
    // node data of round q from ndata (fetched a round earlier), then the node data of round q + 1 fetched
    auto prepare = [&](int q) {
        if (meta && q < rounds) {
            const int v = lane;
            const int *nd = ndata + (q & 3) * 3 * V;
            meta_fill<L, V, 8>(mt, q & 7, v, v0 + (int64_t)q * V + v < v1, nd[v], nd[V + v], __int_as_float(nd[2 * V + v]));
        }
        if (q + 1 < rounds) quad_fetch_nodes<V>(ndata + ((q + 1) & 3) * 3 * V, node + ndel, v0 + (int64_t)(q + 1) * V, v1, nlane, lane);
    };
    floatx4 dprev = floatx4{0.f, 0.f, 0.f, 0.f};
    if constexpr (STAGGER) {
        if (rounds > 0) {
            quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[0], tid);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round 0 (and the node data)
            form(0);
            if (rounds > 1) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[1], tid);
        }
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else {
        if (rounds > 0) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[0], tid);
    }
    for (int r = 0; r <= rounds; ++r) {
        const int cur = r & 1;
        if constexpr (STAGGER) {
            if (kh == 1 && r + 1 < rounds) {
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round r + 1 are in raw
                form(r + 1);
                // the reads of raw above have returned (their values are in Z): the rows of round r + 2 may land
                if (r + 2 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[(r + 2) & 7], tid);
            }
        } else if (r < rounds) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round r are in raw
            form(r);
            if (r + 2 < rounds) prepare(r + 2);  // from ndata, fetched during round r - 1 (or before the loop)
            // the reads of raw above have returned (their values are in Z): the rows of round r + 1 may land
            if (r + 1 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[(r + 1) & 7], tid);
            // Workgroup barrier over the LDS only: the wavefronts share Z, the node data and the partials, all in
            // LDS, and nothing in global memory, so the barrier waits for this wavefront's LDS operations and
            // leaves its loads into raw in flight (a __syncthreads would wait for them too, as LDS writes).
        }
        if constexpr (!STAGGER) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        // records of round r - 1: the first half adds the second half's partials
        if (r > 0 && kh == 0) {
            const int64_t p0 = v0 + (int64_t)(r - 1) * V;
            const int nvp = (int)min((int64_t)V, v1 - p0);
            const floatx4 d = dprev + red[(((r - 1) & 1) * NC + wc) * 64 + lane];
            const int *ids = mt.id[(r - 1) & 7];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v;
                if (row < nvp) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = d[v];
            }
        }
        if (r < rounds) {
            const floatx4 d = sh_products_part<NG, ZS, 2>(Z, cur * V * ZS, b, lane, g0, ng);
            if (kh == 1) red[(cur * NC + wc) * 64 + lane] = d;
            else dprev = d;
        }
        if constexpr (STAGGER) {
            if (kh == 0 && r + 1 < rounds) {
                // this thread's rows of round r + 1 and (node-data wavefront) the node data of round r + 3, both
                // fetched in round r - 1, have landed
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                form(r + 1);
                if (r + 3 < rounds) prepare(r + 3);
                if (r + 2 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[(r + 2) & 7], tid);
            }
            // the partials, tile r + 1 and the node data are complete; the loads into raw stay in flight
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // no load into LDS outlives the workgroup
}

// Quad form: 8 G/16 wavefronts (4 corners x G/16 lane tiles x 2 halves of the K range), 72 B registers each at
// L = 32, four wavefronts per SIMD at 32 lanes. The basis rows of round r + 1 and the node data of round r + 3 are
// fetched before the barrier of round r with LDS-direct loads (no registers held), so they land during the products
// of round r. Per round every thread forms its chunks of the tile from the rows it fetched itself (its own wait is
// the only one needed), 16 threads prepare the node data of round r + 2, one LDS-only barrier, then the products.
// The second half's partial products of round r go through LDS and are added by the first half after the barrier
// of round r + 1, so the sum is (half 0) + (half 1) as in sh_products. Tiles and partials are double-buffered; the
// node ids of round r are read after the barrier of round r + 1, so node data takes eight slots.
template <int G, int L, bool STAGGER>
__global__ __launch_bounds__(G * 32) __attribute__((amdgpu_waves_per_eu(4, 4)))
void sh_records_quad_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KQ = KZ / 16;
    constexpr int KQH = (KQ + 1) / 2;
    constexpr int NWB = G / 16;
    constexpr int NC = 4 * NWB;               // (corner, tile) pairs
    constexpr int NT = 64 * NC * 2;
    constexpr int V = 16;
    constexpr int ZS = KZ + 4;
    __shared__ __attribute__((aligned(16))) float Z[2][V][ZS];
    __shared__ float4 raw[raw_groups<KZ>() * 256];  // the round's basis rows, row-block layout (quad_fetch_rows_rb)
    __shared__ RoundMeta<L, V, 8> mt;  // ids of round r are read after the barrier of round r + 1: eight slots
    __shared__ int ndata[4][3 * V];    // node id, ray and cos phi (bits) of round r in slot r & 3, as fetched
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    __shared__ floatx4 red[2][NC][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = wave / NC, wc = wave % NC;
    const int corner = wc / NWB, nt = wc % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_km<L, KZ>(km, NT);
    constexpr int NG = KQH;
    const int g0 = kh * KQH;
    const int ng = min(KQH, KQ - g0);  // groups of this half (KQH or fewer for the second half)
    float b[4 * NG];
    {
        const float *pl = reinterpret_cast<const float *>(a.plane) + (int64_t)cell * K * G + col;
#pragma unroll
        for (int i = 0; i < NG; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = 16 * (g0 + i) + 4 * (lane >> 4) + u;
                b[4 * i + u] = (i < ng && k < K) ? pl[(int64_t)k * G] : 0.0f;
            }
    }
    quad_rounds<G, L, STAGGER>(a, &Z[0][0][0], raw, &ndata[0][0], &mt, km, &red[0][0][0], b, v0, v1, corner, col, kh, wc,
                               g0, ng);
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
    constexpr int PT = (V * KZ / 4 + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) float Z[V][ZS];
    __shared__ RoundMeta<L, V> mt;
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NWB, nt = wave % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_km<L, KZ>(km, NT);
    float b[4 * KQ];
    load_b<G, L, KQ>(b, a.plane, cell, col, lane);
    for (int64_t r0 = v0; r0 < v1; r0 += V) {
        const int nv = (int)min((int64_t)V, v1 - r0);
        if (tid < V) {
            const NodeRef n = node_ref(a, r0 + tid, v1);
            meta_fill<L, V>(mt, 0, tid, n.id >= 0, n.id, n.ray, n.cp);
        }
        __syncthreads();
        float4 pr[PT];
        fetch_rows<V, KZ, NT, PT>(pr, a.basis, mt, 0, tid);
        rows_to_z<V, KZ, NT, PT>(pr, mt, 0, km, tid, [&](int v, int k4, float4 z) {
            reinterpret_cast<float4 *>(&Z[v][0])[k4] = z;
        });
        __syncthreads();
        const floatx4 d = sh_products<KQ, ZS, 3>(&Z[0][0], b, lane);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)mt.id[0][row] * 4 + corner) * G + col] = d[v];
        }
        __syncthreads();
    }
}

// Mixed precision: split-FP16 operands on v_mfma_f32_16x16x16f16 with FP32 accumulation, the three-product scheme
// of Ootomo and Yokota (2022, "Recovering single precision accuracy from Tensor Cores while surpassing the FP32
// theoretical peak performance"): x = hi + lo 2^-11 with hi = fp16(x) and lo = fp16((x - hi) 2^11), and
// a b = ah bh + (ah bl + al bh) 2^-11 (the lo lo term, 2^-22 relative, is dropped). Both operands are scaled by
// powers of two into the top of FP16's range so that their small entries stay normal numbers: the plane per
// (cell, lane) to a largest magnitude in [2^14, 2^15) (plane = c / scale), the FP32 basis tile (table row times
// cos(m phi)) by 2^13 on its way into LDS; the records are scaled back. Each k-step of 16 costs three 8-pass
// MFMAs instead of four 32-pass FP32 MFMAs; the error per record is about 2^-21 of sum_k |Z_k c_k|, the FP32 level.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

__device__ inline half4_t split_hi(float4 z) {
    return half4_t{(_Float16)z.x, (_Float16)z.y, (_Float16)z.z, (_Float16)z.w};
}

// Split by truncation: hi keeps the leading 11 significant bits of z (exact in FP16 for |z| in FP16's normal
// range), lo = (z - hi) 2^11 holds the next 13 bits rounded to 11; no conversion back from FP16 is needed.
__device__ inline void split_trunc(float4 z, half4_t &h, half4_t &l) {
    const float hx = __uint_as_float(__float_as_uint(z.x) & 0xffffe000u);
    const float hy = __uint_as_float(__float_as_uint(z.y) & 0xffffe000u);
    const float hz = __uint_as_float(__float_as_uint(z.z) & 0xffffe000u);
    const float hw = __uint_as_float(__float_as_uint(z.w) & 0xffffe000u);
    h = half4_t{(_Float16)hx, (_Float16)hy, (_Float16)hz, (_Float16)hw};
    l = half4_t{(_Float16)((z.x - hx) * 2048.0f), (_Float16)((z.y - hy) * 2048.0f), (_Float16)((z.z - hz) * 2048.0f),
                (_Float16)((z.w - hw) * 2048.0f)};
}

__device__ inline half4_t split_lo(float4 z, half4_t h) {
    return half4_t{(_Float16)((z.x - (float)h.x) * 2048.0f), (_Float16)((z.y - (float)h.y) * 2048.0f),
                   (_Float16)((z.z - (float)h.z) * 2048.0f), (_Float16)((z.w - (float)h.w) * 2048.0f)};
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

// Split-FP16 products (K = 32 MFMA) over the steps g0 .. g0 + ng - 1 only (one wavefront's half of the K range),
// tile at offset tile of Zh / Zl (halves); steps past ng read the half's last step again against a zero B.
template <int NS, int ZS, int D>
__device__ inline floatx4 split_products_part(const _Float16 *Zh, const _Float16 *Zl, int tile, const half8_t *bh,
                                              const half8_t *bl, int lane, int g0, int ng) {
    floatx4 am = floatx4{0.f, 0.f, 0.f, 0.f}, a1 = am, a2 = am;
    int zrow = tile + (lane & 15) * ZS + 8 * (lane >> 4) + 32 * g0;
    asm volatile("" : "+v"(zrow));  // one opaque offset: each read one constant offset away from Zh / Zl
    const int glast = ng - 1;
    half8_t ph[D], pl[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        ph[j] = *reinterpret_cast<const half8_t *>(Zh + (zrow + 32 * min(j, glast)));
        pl[j] = *reinterpret_cast<const half8_t *>(Zl + (zrow + 32 * min(j, glast)));
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const half8_t ah = ph[i % D], al = pl[i % D];
        if (i + D < NS) {
            ph[i % D] = *reinterpret_cast<const half8_t *>(Zh + (zrow + 32 * min(i + D, glast)));
            pl[i % D] = *reinterpret_cast<const half8_t *>(Zl + (zrow + 32 * min(i + D, glast)));
        }
        am = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[i], am, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[i], a1, 0, 0, 0);
        a2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[i], a2, 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
        if (i + D < NS) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
    return am + (a1 + a2) * (1.0f / 2048.0f);
}

// The rounds of the split-FP16 quad form: the FP32 quad form's rounds (quad_rounds) with the tile stored as split
// FP16 (hi, lo) times 2^13 and the K = 32 FP16 products.
template <int G, int L>
__device__ __attribute__((always_inline)) inline void quad_rounds_split(
    const ShArgs &a, _Float16 *__restrict__ Zh, _Float16 *__restrict__ Zl, float4 *__restrict__ raw, int *__restrict__ ndata,
    RoundMeta<L, 16, 8> *__restrict__ mtp, const unsigned char *__restrict__ km, floatx4 *__restrict__ red,
    const half8_t *bh, const half8_t *bl, float scale, int64_t v0, int64_t v1, int corner, int col, int kh, int wc,
    int g0, int ng) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KS2 = KZ / 32;
    constexpr int NS = (KS2 + 1) / 2;
    constexpr int NWB = G / 16;
    constexpr int NC = 4 * NWB;
    constexpr int NT = 64 * NC * 2;
    constexpr int V = 16;
    constexpr int ZS = KZ + 8;  // halves
    constexpr float ZSCALE = 8192.0f;
    constexpr int SW = NT - 64;
    RoundMeta<L, V, 8> &mt = *mtp;
    const int32_t *node = a.node, *node_ray = a.node_ray, *node_cp = reinterpret_cast<const int32_t *>(a.node_cp);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    const bool meta = tid >= SW && tid < SW + V;
    if (tid < 2 * V) {
        const NodeRef n = node_ref(a, v0 + tid, v1);
        meta_fill<L, V, 8>(mt, tid / V, tid % V, n.id >= 0, n.id, n.ray, n.cp, ZSCALE);
    }
    const bool nlane = tid >= SW && lane < 3 * V;  // the node-data loaders
    const int64_t ndel = lane < V ? 0 : lane < 2 * V ? (int64_t)(node_ray - node) : (int64_t)(node_cp - node);
    if (rounds > 2) quad_fetch_nodes<V>(ndata + 2 * 3 * V, node + ndel, v0 + 2 * V, v1, nlane, lane);
    __syncthreads();
    const float *basis = a.basis;
    if (rounds > 0) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[0], tid);
    floatx4 dprev = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int r = 0; r <= rounds; ++r) {
        const int cur = r & 1;
        const int64_t r0 = v0 + (int64_t)r * V;
        if (r < rounds) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round r are in raw
            const int zt = cur * V * ZS;
            constexpr int KG = NT / 256, NGR = raw_groups<KZ>(), PASSES = (NGR + KG - 1) / KG;
            int t = tid;
            asm volatile("" : "+v"(t));  // positions recomputed per round (registers)
            const int rb = (t >> 6) & 3, v = 4 * rb + ((t & 63) >> 4);
            const float *cm = mt.cm[r & 7][v];  // cos(m phi) ZSCALE
            _Float16 *zh = Zh + zt + v * ZS, *zl = Zl + zt + v * ZS;
#pragma unroll
            for (int j = 0; j < PASSES; ++j) {
                const int grp = KG * j + (t >> 8);
                const int k4 = 16 * grp + (t & 15);
                if (grp < NGR && k4 < KZ / 4) {
                    const float4 p = raw[(grp * 4 + rb) * 64 + (t & 63)];
                    const uchar4 m = reinterpret_cast<const uchar4 *>(km)[k4];
                    const float4 z = float4{p.x * cm[m.x], p.y * cm[m.y], p.z * cm[m.z], p.w * cm[m.w]};
                    half4_t h, l;
                    split_trunc(z, h, l);
                    reinterpret_cast<half4_t *>(zh)[k4] = h;
                    reinterpret_cast<half4_t *>(zl)[k4] = l;
                }
            }
            if (meta && r + 2 < rounds) {  // from ndata, fetched during round r - 1 (or before the loop)
                const int v = lane;
                const int *nd = ndata + ((r + 2) & 3) * 3 * V;
                meta_fill<L, V, 8>(mt, (r + 2) & 7, v, r0 + 2 * V + v < v1, nd[v], nd[V + v], __int_as_float(nd[2 * V + v]), ZSCALE);
            }
            if (r + 3 < rounds)  // slot (r + 3) & 3 was last read in round r - 3
                quad_fetch_nodes<V>(ndata + ((r + 3) & 3) * 3 * V, node + ndel, r0 + 3 * V, v1, nlane, lane);
            // the reads of raw above have returned (their values are in Z): the rows of round r + 1 may land
            if (r + 1 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw, basis, mt.ray[(r + 1) & 7], tid);
        }
        // Workgroup barrier over the LDS only: the wavefronts share Z, the node data and the partials, all in LDS,
        // and nothing in global memory, so the barrier waits for this wavefront's LDS operations and leaves its
        // loads into raw in flight (a __syncthreads would wait for them too, as LDS writes).
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        // records of round r - 1: the first half adds the second half's partials
        if (r > 0 && kh == 0) {
            const int64_t p0 = v0 + (int64_t)(r - 1) * V;
            const int nvp = (int)min((int64_t)V, v1 - p0);
            const floatx4 d = dprev + red[(((r - 1) & 1) * NC + wc) * 64 + lane];
            const int *ids = mt.id[(r - 1) & 7];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v;
                if (row < nvp) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = (d[v] * scale) * (1.0f / ZSCALE);
            }
        }
        if (r < rounds) {
            const floatx4 d = split_products_part<NS, ZS, 2>(Zh, Zl, cur * V * ZS, bh, bl, lane, g0, ng);
            if (kh == 1) red[(cur * NC + wc) * 64 + lane] = d;
            else dprev = d;
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // no load into LDS outlives the workgroup
}

// Split-FP16 SH records in the flow form of sh_records_kernel, for a K range that is not a multiple of 32 (L = 8;
// the other degrees take the quad form below): 4 x G/16 wavefronts, wavefront (corner, N-tile) holding the whole K
// range of its hi / lo B operands in the quad layout [cell][KZ/4][G][4], K = 16 FP16 products.
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
    constexpr int PT = (V * KZ / 4 + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) _Float16 Zh[2][V][ZS];
    __shared__ __attribute__((aligned(16))) _Float16 Zl[2][V][ZS];
    __shared__ RoundMeta<L, V> mt;
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int corner = wave / NW, nt = wave % NW;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    const bool meta = tid >= SW && tid < SW + V;
    stage_km<L, KZ>(km, NT);
    const float scale = a.scale[(int64_t)cell * G + col];
    // the basis tile is stored times ZSCALE = 2^13 (|Z| <= sqrt(2 (2L + 1) / 4 pi) < 3.3, so |Z| 2^13 < 2^15): its
    // entries down to 2^-27 are FP16 normal numbers
    constexpr float ZSCALE = 8192.0f;
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
    if (tid < 2 * V) {
        const NodeRef n = node_ref(a, v0 + tid, v1);
        meta_fill<L, V>(mt, tid / V, tid % V, n.id >= 0, n.id, n.ray, n.cp);
    }
    NodeRef nn{-1, -1, 1.0f};
    if (meta) nn = node_ref(a, v0 + 2 * V + (tid - SW), v1);
    __syncthreads();
    float4 pr[PT];
    fetch_rows<V, KZ, NT, PT>(pr, a.basis, mt, 0, tid);
    for (int r = 0; r < rounds; ++r) {
        const int cur = r & 1;
        const int64_t r0 = v0 + (int64_t)r * V;
        const int nv = (int)min((int64_t)V, v1 - r0);
        _Float16 (*Hc)[ZS] = Zh[cur];
        _Float16 (*Lc)[ZS] = Zl[cur];
        rows_to_z<V, KZ, NT, PT>(pr, mt, r & 3, km, tid, [&](int v, int k4, float4 z) {
            z = float4{z.x * ZSCALE, z.y * ZSCALE, z.z * ZSCALE, z.w * ZSCALE};
            const half4_t h = split_hi(z);
            reinterpret_cast<half4_t *>(&Hc[v][0])[k4] = h;
            reinterpret_cast<half4_t *>(&Lc[v][0])[k4] = split_lo(z, h);
        });
        if (r + 1 < rounds) fetch_rows<V, KZ, NT, PT>(pr, a.basis, mt, (r + 1) & 3, tid);
        if (meta && r + 2 < rounds) {
            const int v = tid - SW;
            meta_fill<L, V>(mt, (r + 2) & 3, v, nn.id >= 0, nn.id, nn.ray, nn.cp);
            nn = node_ref(a, r0 + 3 * V + v, v1);
        }
        __syncthreads();
        const floatx4 d = split_products<KS, 3>(&Zh[cur][lane & 15][4 * (lane >> 4)],
                                                &Zl[cur][lane & 15][4 * (lane >> 4)], bh, bl);
        const int *ids = mt.id[r & 3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nv) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = (d[v] * scale) * (1.0f / ZSCALE);
        }
    }
}

// Split-FP16 SH records in the quad form (KZ a multiple of 32, L = 16 and 32): 8 G/16 wavefronts (4 corners x G/16
// lane tiles x 2 halves of the K range), each holding its half of the hi / lo B operands (72 VGPRs at L = 32), the
// rounds of quad_rounds_split.
template <int G, int L>
__global__ __launch_bounds__(G * 32) __attribute__((amdgpu_waves_per_eu(4, 4)))
void sh_records_split_quad_kernel(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    static_assert(KZ % 32 == 0, "K = 32 products need KZ a multiple of 32");
    constexpr int KS2 = KZ / 32;
    constexpr int NS = (KS2 + 1) / 2;
    constexpr int NWB = G / 16;
    constexpr int NC = 4 * NWB;
    constexpr int NT = 64 * NC * 2;
    constexpr int V = 16;
    constexpr int ZS = KZ + 8;
    __shared__ __attribute__((aligned(16))) _Float16 Zh[2][V][ZS];
    __shared__ __attribute__((aligned(16))) _Float16 Zl[2][V][ZS];
    __shared__ float4 raw[raw_groups<KZ>() * 256];  // row-block layout (quad_fetch_rows_rb)
    __shared__ RoundMeta<L, V, 8> mt;
    __shared__ int ndata[4][3 * V];
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    __shared__ floatx4 red[2][NC][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = wave / NC, wc = wave % NC;
    const int corner = wc / NWB, nt = wc % NWB;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    stage_km<L, KZ>(km, NT);
    const float scale = a.scale[(int64_t)cell * G + col];
    const int g0 = kh * NS;
    const int ng = min(NS, KS2 - g0);
    // bh[i], bl[i]: K positions 32 (g0 + i) + 8 (lane / 16) + 0..7 of column col (quads 8 (g0 + i) + 2 (lane / 16)
    // and the next of the layout [cell][KZ/4][G][4]); zero past the half's steps
    half8_t bh[NS], bl[NS];
    {
        const half4_t *ph = reinterpret_cast<const half4_t *>(a.plane) + (int64_t)cell * (KZ / 4) * G + col;
        const half4_t *pl = ph + a.plane_lo / 4;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            half4_t h0{}, h1{}, l0{}, l1{};
            if (i < ng) {
                const int q = 8 * (g0 + i) + 2 * (lane >> 4);
                h0 = ph[(int64_t)q * G];
                h1 = ph[(int64_t)(q + 1) * G];
                l0 = pl[(int64_t)q * G];
                l1 = pl[(int64_t)(q + 1) * G];
            }
            bh[i] = half8_t{h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
            bl[i] = half8_t{l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
        }
    }
    quad_rounds_split<G, L>(a, &Zh[0][0][0], &Zl[0][0][0], raw, &ndata[0][0], &mt, km, &red[0][0][0], bh, bl, scale,
                            v0, v1, corner, col, kh, wc, g0, ng);
}

// Split FP16 of a tile chunk with the low part unscaled: hi = z truncated to 11 significant bits (exact in FP16 for
// |z| in FP16's normal range), lo = z - hi (exact in FP32, rounded or flushed in FP16: at most 2^-14 absolute, below
// 2^-28 of a tile scaled to a largest entry near 2^15).
__device__ inline void split_trunc_unscaled(float4 z, half4_t &h, half4_t &l) {
    const float hx = __uint_as_float(__float_as_uint(z.x) & 0xffffe000u);
    const float hy = __uint_as_float(__float_as_uint(z.y) & 0xffffe000u);
    const float hz = __uint_as_float(__float_as_uint(z.z) & 0xffffe000u);
    const float hw = __uint_as_float(__float_as_uint(z.w) & 0xffffe000u);
    h = half4_t{(_Float16)hx, (_Float16)hy, (_Float16)hz, (_Float16)hw};
    l = half4_t{(_Float16)(z.x - hx), (_Float16)(z.y - hy), (_Float16)(z.z - hz), (_Float16)(z.w - hw)};
}

// Split-FP16 products over the whole K range (KS steps of 32) into three accumulators: hi hi, hi lo (the plane's lo
// is scaled by 2^11) and lo hi (the tile's lo is unscaled); the A operands read D steps ahead, the schedule pinned
// to (three products, two reads) per step.
template <int KS, int ZS, int D>
__device__ inline floatx4 split_products_full(const _Float16 *Zh, const _Float16 *Zl, int tile, const half8_t *bh,
                                              const half8_t *bl, int lane) {
    floatx4 am = floatx4{0.f, 0.f, 0.f, 0.f}, a1 = am, a2 = am;
    int zrow = tile + (lane & 15) * ZS + 8 * (lane >> 4);
    asm volatile("" : "+v"(zrow));  // one opaque offset: each read one constant offset away from Zh / Zl
    half8_t ph[D], pl[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        ph[j] = *reinterpret_cast<const half8_t *>(Zh + (zrow + 32 * j));
        pl[j] = *reinterpret_cast<const half8_t *>(Zl + (zrow + 32 * j));
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
#pragma unroll
    for (int i = 0; i < KS; ++i) {
        const half8_t ah = ph[i % D], al = pl[i % D];
        if (i + D < KS) {
            ph[i % D] = *reinterpret_cast<const half8_t *>(Zh + (zrow + 32 * (i + D)));
            pl[i % D] = *reinterpret_cast<const half8_t *>(Zl + (zrow + 32 * (i + D)));
        }
        am = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[i], am, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[i], a1, 0, 0, 0);
        a2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[i], a2, 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
        if (i + D < KS) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
    return (am + a2) + a1 * (1.0f / 2048.0f);
}

// Split-FP16 SH records, staggered form (32 lanes; L = 16, 24, 32): 8 wavefronts, wavefront (corner, 16-lane tile)
// holding the hi / lo B operands of its whole K range (KS steps of 32, 8 KS registers), so no partial products
// cross wavefronts. The wavefronts form two groups of four with one wavefront of each per SIMD (a workgroup's
// wavefronts are placed on the SIMDs round robin): group A (corners 0 and 1) runs the products and records of round
// r and then forms its chunks of tile r + 1; group B (corners 2 and 3) forms its chunks of tile r + 1 first and then
// runs the products and records of round r. On every SIMD one wavefront's tile formation (VALU) thus runs under the
// other's products (MFMA). Tiles and basis rows are double-buffered: a thread fetches its chunks of the rows of round
// r + 2 (LDS-direct loads, the row-block layout of quad_fetch_rows_rb) right after forming its chunks of tile r + 1
// from those of round r + 1, so every fetch has about a round to land and each thread waits only for its own loads.
// Node data (record rows, rays, cos(m phi) 2^13) is prepared three rounds ahead in four slots by one wavefront of
// group A, from node ids, rays and cosines it fetched a round earlier. Tile formation per chunk of four positions:
// z = row x cos(m phi) 2^13 with at most two orders m per chunk (two cosine reads and a per-position select), hi and
// lo by split_trunc_unscaled. Records = (hi hi + lo hi + hi lo 2^-11) x plane scale / 2^13.
template <int L>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
void sh_records_split_stagger_kernel(ShArgs a) {
    constexpr int G = 32;
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    constexpr int KS = (KZ + 31) / 32;      // K = 32 steps
    constexpr int ZK = 32 * KS;             // tile positions; KZ .. ZK - 1 stay zero
    constexpr int ZS = ZK + 8;              // halves per tile row: rows 4 banks apart
    constexpr int V = 16;                   // nodes per round
    constexpr int NT = 512;
    constexpr int KG = NT / 256, NGR = raw_groups<KZ>(), PASSES = (NGR + KG - 1) / KG;
    constexpr int MW = 3;                   // the node-data wavefront (group A)
    constexpr float ZSCALE = 8192.0f;
    __shared__ __attribute__((aligned(16))) _Float16 Zt[2][2][V][ZS];  // tile buffer, hi / lo, node, position
    __shared__ float4 raw[2][NGR * 256];
    __shared__ RoundMeta<L, V, 4> mt;
    __shared__ int ndata[4][3 * V];
    __shared__ __attribute__((aligned(4))) unsigned char km[KZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool group_b = wave >= 4;
    const int corner = 2 * (wave >> 2) + ((wave >> 1) & 1), nt = wave & 1;
    const int col = nt * 16 + (lane & 15);
    const int64_t item = blockIdx.x;
    const int cell = corner_cell(a.item_cell[item], corner, a.nc);
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int rounds = (int)((v1 - v0 + V - 1) / V);
    stage_km<L, KZ>(km, NT);
    if constexpr (ZK > KZ) {
        for (int e = tid; e < 2 * V * (ZK - KZ); e += NT) {
            const int b = e / (V * (ZK - KZ)), rem = e % (V * (ZK - KZ));
            Zt[b][0][rem / (ZK - KZ)][KZ + rem % (ZK - KZ)] = (_Float16)0.0f;
            Zt[b][1][rem / (ZK - KZ)][KZ + rem % (ZK - KZ)] = (_Float16)0.0f;
        }
    }
    const float scale = a.scale[(int64_t)cell * G + col];
    // bh[i], bl[i]: K positions 32 i + 8 (lane / 16) + 0..7 of column col (quads 8 i + 2 (lane / 16) and the next of
    // the layout [cell][KZ/4][G][4]); zero past KZ
    half8_t bh[KS], bl[KS];
    {
        const half4_t *ph = reinterpret_cast<const half4_t *>(a.plane) + (int64_t)cell * (KZ / 4) * G + col;
        const half4_t *pl = ph + a.plane_lo / 4;
#pragma unroll
        for (int i = 0; i < KS; ++i) {
            const int q = 8 * i + 2 * (lane >> 4);
            half4_t h0{}, h1{}, l0{}, l1{};
            if (q < KZ / 4) {
                h0 = ph[(int64_t)q * G];
                l0 = pl[(int64_t)q * G];
            }
            if (q + 1 < KZ / 4) {
                h1 = ph[(int64_t)(q + 1) * G];
                l1 = pl[(int64_t)(q + 1) * G];
            }
            bh[i] = half8_t{h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
            bl[i] = half8_t{l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
        }
    }
    const int32_t *node = a.node, *node_ray = a.node_ray, *node_cp = reinterpret_cast<const int32_t *>(a.node_cp);
    const bool meta = wave == MW && lane < V;
    const bool nlane = wave == MW && lane < 3 * V;  // the node-data loaders
    const int64_t ndel = lane < V ? 0 : lane < 2 * V ? (int64_t)(node_ray - node) : (int64_t)(node_cp - node);
    if (tid < 3 * V) {  // node data of rounds 0, 1, 2
        const NodeRef n = node_ref(a, v0 + tid, v1);
        meta_fill<L, V, 4>(mt, tid / V, tid % V, n.id >= 0, n.id, n.ray, n.cp, ZSCALE);
    }
    if (rounds > 3) quad_fetch_nodes<V>(&ndata[3][0], node + ndel, v0 + 3 * V, v1, nlane, lane);
    __syncthreads();
    // per thread and pass: its chunk's row, the cosine offsets of the chunk's first order mA and other order mB
    // (mB = mA for a chunk of one order) and the positions that take mB (bit 4 j + e); positions past K read mA
    const int rb = wave & 3, vrow = 4 * rb + (lane >> 4);
    int offa[PASSES], offb[PASSES];
    unsigned sel = 0;
#pragma unroll
    for (int j = 0; j < PASSES; ++j) {
        const int k4 = 16 * (KG * j + (wave >> 2)) + (lane & 15);
        int ma = 0, mb = 0;
        if (k4 < KZ / 4) {
            ma = km[4 * k4];
            mb = ma;
#pragma unroll
            for (int e = 1; e < 4; ++e)
                if (4 * k4 + e < K && km[4 * k4 + e] != ma) {
                    mb = km[4 * k4 + e];
                    sel |= 1u << (4 * j + e);
                }
        }
        offa[j] = vrow * (L + 1) + ma;
        offb[j] = vrow * (L + 1) + mb;
    }
    // this thread's chunks of tile q from its chunks of the rows of round q
    auto form = [&](int q) {
        const float *cmr = &mt.cm[q & 3][0][0];
        const float4 *rw = raw[q & 1];
        _Float16 *zh = &Zt[q & 1][0][vrow][0], *zl = &Zt[q & 1][1][vrow][0];
#pragma unroll
        for (int j = 0; j < PASSES; ++j) {
            const int grp = KG * j + (wave >> 2);
            const int k4 = 16 * grp + (lane & 15);
            if (grp < NGR && k4 < KZ / 4) {
                const float4 p = rw[(grp * 4 + rb) * 64 + lane];
                const float ca = cmr[offa[j]], cb = cmr[offb[j]];
                const unsigned s = sel >> (4 * j);
                const float4 z = float4{p.x * ca, p.y * ((s & 2u) ? cb : ca), p.z * ((s & 4u) ? cb : ca),
                                        p.w * ((s & 8u) ? cb : ca)};
                half4_t h, l;
                split_trunc_unscaled(z, h, l);
                reinterpret_cast<half4_t *>(zh)[k4] = h;
                reinterpret_cast<half4_t *>(zl)[k4] = l;
            }
        }
    };
    auto store = [&](int q, floatx4 d) {
        const int64_t p0 = v0 + (int64_t)q * V;
        const int nvp = (int)min((int64_t)V, v1 - p0);
        const int *ids = mt.id[q & 3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v;
            if (row < nvp) a.records[((int64_t)ids[row] * 4 + corner) * G + col] = (d[v] * scale) * (1.0f / ZSCALE);
        }
    };
    const float *basis = a.basis;
    if (rounds > 0) quad_fetch_rows_rb<V, KZ, NT>(raw[0], basis, mt.ray[0], tid);
    if (rounds > 1) quad_fetch_rows_rb<V, KZ, NT>(raw[1], basis, mt.ray[1], tid);
    if (rounds > 0) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round 0 (and 1) are in raw
        form(0);
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    for (int r = 0; r < rounds; ++r) {
        const bool next = r + 1 < rounds;
        if (group_b) {
            if (next) {
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's rows of round r + 1 are in raw
                form(r + 1);
                // the reads of raw[(r + 1) & 1] have returned (their values are in the tile): raw[r & 1] was read
                // in round r - 1, so the rows of round r + 2 may land there
                if (r + 2 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw[r & 1], basis, mt.ray[(r + 2) & 3], tid);
            }
            store(r, split_products_full<KS, ZS, 2>(&Zt[r & 1][0][0][0], &Zt[r & 1][1][0][0], 0, bh, bl, lane));
        } else {
            store(r, split_products_full<KS, ZS, 2>(&Zt[r & 1][0][0][0], &Zt[r & 1][1][0][0], 0, bh, bl, lane));
            if (next) {
                // this thread's rows of round r + 1 and (node-data wavefront) the node data of round r + 3, both
                // fetched in round r - 1, have landed
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                form(r + 1);
                if (wave == MW && r + 3 < rounds) {
                    // node data of round r + 3 into slot (r + 3) & 3, whose round r - 1 was last read by the records
                    // of round r - 1; then the node data of round r + 4 into slot r & 3 (read in round r - 3)
                    if (meta) {
                        const int *nd = &ndata[(r + 3) & 3][0];
                        meta_fill<L, V, 4>(mt, (r + 3) & 3, lane, v0 + (int64_t)(r + 3) * V + lane < v1, nd[lane],
                                           nd[V + lane], __int_as_float(nd[2 * V + lane]), ZSCALE);
                    }
                    if (r + 4 < rounds)
                        quad_fetch_nodes<V>(&ndata[r & 3][0], node + ndel, v0 + (int64_t)(r + 4) * V, v1, nlane, lane);
                }
                if (r + 2 < rounds) quad_fetch_rows_rb<V, KZ, NT>(raw[r & 1], basis, mt.ray[(r + 2) & 3], tid);
            }
        }
        // LDS-only barrier: tile r + 1 and the node data are complete; the loads into raw stay in flight
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // no load into LDS outlives the workgroup
}

// One lane (W = 1): one thread per node, dot products of the node's basis row (table row times cos(m phi)) against
// the four corner cells' coefficients staged in LDS (a split FP16 plane is recombined and scaled back to FP32 there).
template <int L, bool HALF>
__global__ __launch_bounds__(256) void sh_records_scalar(ShArgs a) {
    constexpr int K = (L + 1) * (L + 2) / 2;
    constexpr int KZ = (K + 15) / 16 * 16;
    __shared__ float c[4][K];
    const int64_t item = blockIdx.x;
    const int base = a.item_cell[item];
    for (int e = threadIdx.x; e < 4 * K; e += 256) {
        const int cell = corner_cell(base, e / K, a.nc), k = e % K;
        const int64_t i = (int64_t)cell * K + k;
        if constexpr (HALF) {
            const int64_t j = (int64_t)cell * KZ + k;  // quad layout with one lane
            c[e / K][k] = (plane_at<L, true>(a.plane, j) + plane_at<L, true>(a.plane, j + a.plane_lo) * (1.0f / 2048.0f)) *
                          a.scale[cell];
        }
        else
            c[e / K][k] = plane_at<L, false>(a.plane, i);
    }
    __syncthreads();
    for (int64_t r = a.item_v0[item] + threadIdx.x; r < a.item_v1[item]; r += 256) {
        const int id = a.node[r];
        const float *row = a.basis + (int64_t)a.node_ray[r] * KZ;
        const float cp = a.node_cp[r];
        float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float cm0 = 1.0f, cm1 = cp;
        for (int m = 0; m <= L; ++m) {
            float cm;
            if (m == 0) cm = 1.0f;
            else if (m == 1) cm = cp;
            else { cm = (2.0f * cp) * cm1 - cm0; cm0 = cm1; cm1 = cm; }
            const int o = moff<L>(m);
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int l = m; l <= L; ++l) {
                const float z = row[o + l - m] * cm;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = acc[q] + z * c[q][o + l - m];
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
        if (half) {
            if constexpr (G == 32 && L >= 16) {
                if (a.tile != 1) {  // the staggered form unless the quad form (tile 1) is asked for
                    hipLaunchKernelGGL((sh_records_split_stagger_kernel<L>), dim3((unsigned)a.items), dim3(512), 0, s,
                                       a);
                    hipError_t e = hipGetLastError();
                    if (e != hipSuccess) return set_error("sh_records launch: %s", hipGetErrorString(e));
                    return SOS_OK;
                }
            }
            if constexpr (((L + 1) * (L + 2) / 2 + 15) / 16 * 16 % 32 == 0)
                hipLaunchKernelGGL((sh_records_split_quad_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 32), 0, s, a);
            else
                hipLaunchKernelGGL((sh_records_split_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
        }
        else if (a.tile == 1)
            hipLaunchKernelGGL((sh_records_step_kernel<G, L>), dim3((unsigned)a.items), dim3(G * 16), 0, s, a);
        else if (a.tile == 2)
            hipLaunchKernelGGL((sh_records_quad_kernel<G, L, false>), dim3((unsigned)a.items), dim3(G * 32), 0, s, a);
        else if (a.tile == 3)
            hipLaunchKernelGGL((sh_records_quad_kernel<G, L, true>), dim3((unsigned)a.items), dim3(G * 32), 0, s, a);
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
        case 24: return sh_launch<G, 24>(a, half, s);
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
