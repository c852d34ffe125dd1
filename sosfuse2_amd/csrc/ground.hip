// First-light ground term as an MFMA contraction per destination-grid cell (gfx950).
//
// For every (node, corner) visit of a first-light launch, the sunlit-ground light scattered by the aerosol layers:
//   rec[visit][w] = asca[w] p_aer(h) sum_p PHt[visit][p] gv[cell][p][w] + ssca[w] p_sul(h) sum_p PHs[visit][p] gv[p][w]
// with PH the paired Henyey-Greenstein phases (hg(m0 + m1) + hg(m0 - m1)) / 2 between the ray direction in the node
// frame and the cell's pool direction (src/sosfuse2/backends/numpy_reference.py:526-558). The phases depend on the
// lane only through its asymmetries (g, sulfate g), so for each of the launch's phase groups (lanes sharing both; one
// per state, at most four) the pool sum is a dense [visits x pools] x [pools x lanes] product: per cell the pool
// radiances are staged once in LDS, phase tiles of 16 visits are generated in LDS per group, and
// v_mfma_f32_16x16x4f32 accumulates them; each lane keeps its own group's sums. The transport then adds
// sum_corners sw * rec instead of walking the pools of four cells at every node.

#include "common.h"

namespace sos {
int set_error(const char *fmt, ...);

// Layout: a workgroup takes an item (at most 2048 visits of one cell) in rounds of 32 visits. Each wavefront
// (16-lane N-tile nt, 16-visit M-tile mt) holds the cell's pool radiances of its lanes as the MFMA B operand in
// registers (NCH chunks of PC pools, loaded once per item), so the pools are staged once however many there are;
// per round and chunk the 32 x PC phase tile of each phase group is generated in LDS by all threads and multiplied.
// Step 4 i + u of a chunk's products covers pool 16 i + 4 (lane / 16) + u, so one ds_read_b128 of the tile feeds four
// MFMAs. The tile is generated only up to the chunk's pools rounded up to 16 (the products stop there); a thread
// takes one visit (two at 16 lanes) and the pools p = 8 j + (thread mod 8), so its direction stays in registers and
// no index is divided. Slots past the cell's pools and absent visits evaluate the phase of a zero direction (finite),
// against a zero B or into a record never stored.
template <int G, int NCH>
__global__ __launch_bounds__(64 * (G / 16) * 2) void ground_records_kernel(GroundArgs a) {
    constexpr int NW = G / 16;               // N-tiles
    constexpr int NT = 64 * NW * 2;          // two M-tiles per N-tile
    constexpr int V = 32, PC = 160, PS = PC + 4, NQ = PC / 16;
    constexpr int VS = NT / 8, VT = V / VS;  // visits per pass of the tile generation, passes
    __shared__ __attribute__((aligned(16))) float PT[V][PS];
    __shared__ __attribute__((aligned(16))) float PU[V][PS];
    __shared__ float4 gd[NCH * PC];
    __shared__ float4 sdir[V];
    __shared__ int sid[V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nt = wave % NW, mt = wave / NW;
    const int col = nt * 16 + (lane & 15);
    const int kk = lane >> 4, row = mt * 16 + (lane & 15);
    const int pl = tid & 7, vt = tid >> 3;
    const int64_t item = blockIdx.x;
    const int cell = a.item_cell[item];
    const int64_t v0 = a.item_v0[item], v1 = a.item_v1[item];
    const int64_t p0 = a.goff[cell];
    const int np = (int)min((int64_t)(NCH * PC), a.goff[cell + 1] - p0);
    const int chunks = (np + PC - 1) / PC;
    const float asca = a.lane[SOS_LANE_ASCA * G + col], ssca = a.lane[SOS_LANE_SSCA * G + col];
    // this lane's phase group
    int mygroup = 0;
    {
        const float lg = a.lane[SOS_LANE_G * G + col], lsg = a.lane[SOS_LANE_SG * G + col];
        for (int i = 0; i < a.groups; ++i)
            if (a.gg[i] == lg && (!a.strato || a.sgg[i] == lsg)) { mygroup = i; break; }
    }
    // B operand: bq[ch][4 i + u] = gval[pool ch PC + 16 i + 4 kk + u][col] (0 past the cell's pools)
    float bq[NCH][PC / 4];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
        for (int i = 0; i < NQ; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int pp = ch * PC + 16 * i + 4 * kk + u;
                bq[ch][4 * i + u] = pp < np ? a.gval[(p0 + pp) * G + col] : 0.0f;
            }
    for (int p = tid; p < NCH * PC; p += NT)
        gd[p] = p < np ? make_float4(a.gdir[(p0 + p) * 3], a.gdir[(p0 + p) * 3 + 1], a.gdir[(p0 + p) * 3 + 2], 0.0f)
                       : make_float4(0.f, 0.f, 0.f, 0.f);
    // the visit ids and directions of round r + 1 are loaded (threads 0 .. V - 1) while round r is evaluated
    int nid = -1;
    float4 ndir = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < V && v0 + tid < v1) {
        nid = a.vis[v0 + tid];
        ndir = a.node_dir[nid >> 2];
    }
    for (int64_t r0 = v0; r0 < v1; r0 += V) {
        const int nv = (int)min((int64_t)V, v1 - r0);
        if (tid < V) {
            sid[tid] = nid;
            sdir[tid] = ndir;
            nid = -1;
            ndir = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r0 + V + tid < v1) {
                nid = a.vis[r0 + V + tid];
                ndir = a.node_dir[nid >> 2];
            }
        }
        __syncthreads();
        float4 dv[VT];
#pragma unroll
        for (int i = 0; i < VT; ++i) dv[i] = sdir[vt + VS * i];
        floatx4 acct = floatx4{0.f, 0.f, 0.f, 0.f}, accs = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (ch >= chunks) break;
            const int n = min(PC, np - ch * PC);
            const int n16 = (n + 15) & ~15;
            for (int grp = 0; grp < a.groups; ++grp) {
                const HgConst ht = hg_const(a.gg[grp]), hs = hg_const(a.sgg[grp]);
#pragma unroll
                for (int j = 0; j < PC / 8; ++j) {
                    if (8 * j >= n16) break;  // uniform: the chunk's pools rounded up to 16
                    const int p = 8 * j + pl;
                    const float4 g3 = gd[ch * PC + p];
#pragma unroll
                    for (int i = 0; i < VT; ++i) {
                        const float m0 = dv[i].x * g3.z + dv[i].y * g3.x;
                        const float m1 = dv[i].z * g3.y;
                        const float up = clampf(m0 + m1, -1.0f, 1.0f), dn = clampf(m0 - m1, -1.0f, 1.0f);
                        PT[vt + VS * i][p] = hg_pair(up, dn, ht);
                        if (a.strato) PU[vt + VS * i][p] = hg_pair(up, dn, hs);
                    }
                }
                __syncthreads();
                // every wavefront runs the products (the MFMA needs the whole wavefront); lanes of other groups
                // discard them
                floatx4 t = floatx4{0.f, 0.f, 0.f, 0.f}, u = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    if (16 * i >= n) break;  // past the chunk's pools both operands are zero
                    const float4 at = *reinterpret_cast<const float4 *>(&PT[row][16 * i + 4 * kk]);
                    t = __builtin_amdgcn_mfma_f32_16x16x4f32(at.x, bq[ch][4 * i], t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x4f32(at.y, bq[ch][4 * i + 1], t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x4f32(at.z, bq[ch][4 * i + 2], t, 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x4f32(at.w, bq[ch][4 * i + 3], t, 0, 0, 0);
                    if (a.strato) {
                        const float4 as = *reinterpret_cast<const float4 *>(&PU[row][16 * i + 4 * kk]);
                        u = __builtin_amdgcn_mfma_f32_16x16x4f32(as.x, bq[ch][4 * i], u, 0, 0, 0);
                        u = __builtin_amdgcn_mfma_f32_16x16x4f32(as.y, bq[ch][4 * i + 1], u, 0, 0, 0);
                        u = __builtin_amdgcn_mfma_f32_16x16x4f32(as.z, bq[ch][4 * i + 2], u, 0, 0, 0);
                        u = __builtin_amdgcn_mfma_f32_16x16x4f32(as.w, bq[ch][4 * i + 3], u, 0, 0, 0);
                    }
                }
                if (grp == mygroup) {
                    acct = acct + t;
                    accs = accs + u;
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = mt * 16 + 4 * (lane >> 4) + v;
            if (r < nv) {
                float prof[5], st;
                profiles(a.shape, sdir[r].w, prof, st);
                float rec = (prof[1] * asca) * acct[v];
                if (a.strato) rec = rec + (st * ssca) * accs[v];
                a.records[(int64_t)sid[r] * G + col] = rec;
            }
        }
        __syncthreads();
    }
}

// Ground pools of one set of lanes (engine_core.py:103-147), one workgroup per cell and one thread per lane:
// gval[p][w] = albedo_w factor_p exp(-tau), tau = column_p[0..4] . ext[0..4][w] + column_p[5] sulfate_ext[w], and
// the cell's Rayleigh tensor moments of the ground-reflected light gm[cell][j][w] = 3/(8 pi) ext_air[w]
// sum_p f_j(loc_p) gval[p][w], f = ((1 - x^2)/2, (1 - y^2)/2, (1 - z^2)/2, -x z/2); FP64 throughout, the outputs
// rounded to FP32 once.
__global__ void pool_values_kernel(int G, const int64_t *off, const double *column, const double *factor,
                                   const double *loc, const float *lane, float *gval, float *gm) {
    const int64_t cell = blockIdx.x;
    const int w = threadIdx.x;
    if (w >= G) return;
    double e[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) e[i] = lane[(SOS_LANE_EXT0 + i) * G + w];
    const double es = lane[SOS_LANE_SEXT * G + w], alb = lane[SOS_LANE_ALBEDO * G + w];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t p = off[cell]; p < off[cell + 1]; ++p) {
        const double *c = column + 6 * p;
        const double tau = c[0] * e[0] + c[1] * e[1] + c[2] * e[2] + c[3] * e[3] + c[4] * e[4] + c[5] * es;
        const double v = alb * factor[p] * exp(-tau);
        gval[p * G + w] = (float)v;
        const double x = loc[3 * p], y = loc[3 * p + 1], z = loc[3 * p + 2];
        acc[0] += (1.0 - x * x) * 0.5 * v;
        acc[1] += (1.0 - y * y) * 0.5 * v;
        acc[2] += (1.0 - z * z) * 0.5 * v;
        acc[3] += -x * z * 0.5 * v;
    }
    const double k = 3.0 / (8.0 * 3.14159265358979324) * e[0];
#pragma unroll
    for (int j = 0; j < 4; ++j) gm[(cell * 4 + j) * G + w] = (float)(acc[j] * k);
}

int pool_values_launch(int cells, int G, const int64_t *off, const double *column, const double *factor,
                       const double *loc, const float *lane, float *gval, float *gm, hipStream_t s) {
    if (cells <= 0) return SOS_OK;
    hipLaunchKernelGGL(pool_values_kernel, dim3((unsigned)cells), dim3(64), 0, s, G, off, column, factor, loc, lane,
                       gval, gm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("pool_values launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

template <int G, int NCH>
static int ground_launch(const GroundArgs &a, hipStream_t s) {
    if (a.items <= 0) return SOS_OK;
    hipLaunchKernelGGL((ground_records_kernel<G, NCH>), dim3((unsigned)a.items), dim3(64 * (G / 16) * 2), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error("ground_records launch: %s", hipGetErrorString(e));
    return SOS_OK;
}

template <int G>
static int ground_chunks(const GroundArgs &a, hipStream_t s) {
    // pools of a cell held in registers in chunks of 160
    switch ((a.pool_max + 159) / 160) {
        case 0:
        case 1: return ground_launch<G, 1>(a, s);
        case 2: return ground_launch<G, 2>(a, s);
        case 3: return ground_launch<G, 3>(a, s);
        case 4: return ground_launch<G, 4>(a, s);
        default: return set_error("ground_records: %d pools in a cell (at most 640)", a.pool_max);
    }
}

int ground_records_dispatch(const GroundArgs &a, hipStream_t s) {
    switch (a.G) {
        case 16: return ground_chunks<16>(a, s);
        case 32: return ground_chunks<32>(a, s);
        default: return set_error("ground_records: lanes %d", a.G);
    }
}

}  // namespace sos
