// Training and inference kernels of the learned warm start's network (gfx950): a tiled BF16 MFMA product with FP32
// accumulation and a fused bias + GELU epilogue, the GELU derivative, the weighted squared-error loss and its
// gradient, bias gradients and Adam. The network has no counterpart in the reference.
//
// Products: C[M][N] = op(A)[M][K] op(B)[K][N] with A(i, k) = A[i sa_i + k sa_k] and B(k, j) = B[k sb_k + j sb_j] in FP32
// memory; operands are rounded to BF16 (nearest, ties to even) on their way into LDS and multiplied by
// v_mfma_f32_16x16x16bf16_1k. Strides make every product of the forward and backward passes one kernel:
// forward X W^T, input gradient dZ W, weight gradient dZ^T X (the last one over the sample axis, split over
// gridDim.z into partial products summed in a fixed order, so every result is deterministic).

#include "common.h"

namespace sos {
int set_error(const char *fmt, ...);

typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad_f(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// 64 x 64 output tile per workgroup of 4 wavefronts (2 x 2 of 32 x 32), K staged 32 at a time; both tiles are kept
// K-contiguous in LDS (As[row][k], Bt[col][k]) so that every MFMA operand is one 8-byte LDS read.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs a) {
    constexpr int TM = 64, TN = 64, TK = 32, PAD = 8;
    __shared__ unsigned short As[TM][TK + PAD], Bt[TN][TK + PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t i0 = (int64_t)blockIdx.x * TM, j0 = (int64_t)blockIdx.y * TN;
    const int64_t kb = (int64_t)blockIdx.z * a.kchunk, ke = min(a.K, kb + a.kchunk);
    floatx4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bool a_krow = a.sa_k == 1, b_kcol = a.sb_k == 1;
    for (int64_t k0 = kb; k0 < ke; k0 += TK) {
        for (int e = tid; e < TM * TK; e += 256) {
            const int r = a_krow ? e / TK : e % TM, kk = a_krow ? e % TK : e / TM;
            const int64_t i = i0 + r, k = k0 + kk;
            As[r][kk] = (i < a.M && k < ke) ? to_bf16(a.A[i * a.sa_i + k * a.sa_k]) : (unsigned short)0;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            const int c = b_kcol ? e / TK : e % TN, kk = b_kcol ? e % TK : e / TN;
            const int64_t j = j0 + c, k = k0 + kk;
            Bt[c][kk] = (j < a.N && k < ke) ? to_bf16(a.B[k * a.sb_k + j * a.sb_j]) : (unsigned short)0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK / 16; ++ks) {
            bf16x4 av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x)
                av[x] = *reinterpret_cast<const bf16x4 *>(&As[wm * 32 + x * 16 + (lane & 15)][ks * 16 + 4 * (lane >> 4)]);
#pragma unroll
            for (int y = 0; y < 2; ++y)
                bv[y] = *reinterpret_cast<const bf16x4 *>(&Bt[wn * 32 + y * 16 + (lane & 15)][ks * 16 + 4 * (lane >> 4)]);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[x], bv[y], acc[x][y], 0, 0, 0);
        }
        __syncthreads();
    }
    float *Z = a.Z + (a.splits > 1 ? (int64_t)blockIdx.z * a.M * a.ldz : 0);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = i0 + wm * 32 + x * 16 + 4 * (lane >> 4) + v;
                const int64_t j = j0 + wn * 32 + y * 16 + (lane & 15);
                if (i < a.M && j < a.N) {
                    float z = acc[x][y][v];
                    if (a.splits == 1) {
                        if (a.bias) z = z + a.bias[j];
                        if (a.act == 1) a.H[i * a.ldh + j] = gelu_f(z);
                    }
                    Z[i * a.ldz + j] = z;
                }
            }
}

// Sum of the split partials in split order, then bias and activation.
__global__ void split_epilogue_kernel(GemmArgs a, const float *__restrict__ part, float *__restrict__ Z) {
    const int64_t total = a.M * a.N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / a.N, j = e % a.N;
        float z = 0.0f;
        for (int s = 0; s < a.splits; ++s) z = z + part[((int64_t)s * a.M + i) * a.N + j];
        if (a.bias) z = z + a.bias[j];
        if (a.act == 1) a.H[i * a.ldh + j] = gelu_f(z);
        Z[i * a.ldz + j] = z;
    }
}

__global__ void gelu_backward_kernel(int64_t rows, int64_t cols, const float *__restrict__ dh, int64_t ldd,
                                     const float *__restrict__ z, int64_t ldz, float *__restrict__ dz, int64_t ldo) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        dz[i * ldo + j] = dh[i * ldd + j] * gelu_grad_f(z[i * ldz + j]);
    }
}

// dy = 2 w (y - t) / (rows cols); per-block FP64 partial sums of w (y - t)^2 in part[block], summed by loss_finish.
__global__ __launch_bounds__(256) void mse_kernel(int64_t rows, int64_t cols, const float *__restrict__ y, int64_t ldy,
                                                  const float *__restrict__ t, int64_t ldt,
                                                  const float *__restrict__ w, float *__restrict__ dy, int64_t lddy,
                                                  double *__restrict__ part) {
    __shared__ double red[256];
    const int64_t total = rows * cols;
    const float scale = 2.0f / (float)total;
    double acc = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        const float d = y[i * ldy + j] - t[i * ldt + j];
        const float wj = w ? w[j] : 1.0f;
        if (dy) dy[i * lddy + j] = scale * wj * d;
        acc += (double)wj * (double)d * (double)d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void loss_finish_kernel(int blocks, int64_t total, const double *__restrict__ part, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < blocks; ++b) s += part[b];
        out[0] = (float)(s / (double)total);
    }
}

// db[j] = sum_i dz[i][j]: one workgroup per column, fixed-order tree.
__global__ __launch_bounds__(256) void bias_grad_kernel(int64_t rows, const float *__restrict__ dz, int64_t ld,
                                                        float *__restrict__ db) {
    __shared__ float red[256];
    const int64_t j = blockIdx.x;
    float acc = 0.0f;
    for (int64_t i = threadIdx.x; i < rows; i += 256) acc = acc + dz[i * ld + j];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = red[0];
}

__global__ void adam_kernel(int64_t n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        p[e] = p[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

static unsigned grid_for(int64_t n) { return (unsigned)max((int64_t)1, min((n + 255) / 256, (int64_t)65536)); }

static int check(const char *what) {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SOS_OK : set_error("%s: %s", what, hipGetErrorString(e));
}

int mlp_gemm_launch(const GemmArgs &in, float *work, hipStream_t s) {
    GemmArgs a = in;
    if (a.M <= 0 || a.N <= 0) return SOS_OK;
    const int splits = a.splits < 1 ? 1 : a.splits;
    a.splits = splits;
    a.kchunk = ((a.K + splits - 1) / splits + 31) / 32 * 32;
    dim3 grid((unsigned)((a.M + 63) / 64), (unsigned)((a.N + 63) / 64), (unsigned)splits);
    if (splits == 1) {
        hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, a);
        return check("mlp_gemm");
    }
    float *Z = a.Z;
    GemmArgs p = a;
    p.Z = work;
    p.ldz = a.N;
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, p);
    hipLaunchKernelGGL(split_epilogue_kernel, dim3(grid_for(a.M * a.N)), dim3(256), 0, s, a, work, Z);
    return check("mlp_gemm split");
}

int mlp_gelu_backward_launch(int64_t rows, int64_t cols, const float *dh, int64_t ldd, const float *z, int64_t ldz,
                             float *dz, int64_t ldo, hipStream_t s) {
    if (rows * cols > 0)
        hipLaunchKernelGGL(gelu_backward_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, dh, ldd, z,
                           ldz, dz, ldo);
    return check("mlp_gelu_backward");
}

int mlp_mse_launch(int64_t rows, int64_t cols, const float *y, int64_t ldy, const float *t, int64_t ldt,
                   const float *w, float *dy, int64_t lddy, float *loss, double *work, hipStream_t s) {
    const int blocks = 1024;
    hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, s, rows, cols, y, ldy, t, ldt, w, dy, lddy, work);
    hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, s, blocks, rows * cols, work, loss);
    return check("mlp_mse");
}

int mlp_bias_grad_launch(int64_t rows, int64_t cols, const float *dz, int64_t ld, float *db, hipStream_t s) {
    if (cols > 0) hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)cols), dim3(256), 0, s, rows, dz, ld, db);
    return check("mlp_bias_grad");
}

int mlp_adam_launch(int64_t n, float *p, const float *g, float *m, float *v, int step, float lr, float b1, float b2,
                    float eps, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    if (n > 0) hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
    return check("mlp_adam");
}

}  // namespace sos
