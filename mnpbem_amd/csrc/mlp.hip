// Per-face MLP of the GMRES warm start: GEMMs on the BF16 matrix cores with FP32 accumulation,
// GELU, mean-squared loss and Adam. No counterpart in the reference (BEMRetIter starts from
// zero, bem_ret_iter.py:1695).
//
// All matrices are row-major FP32 in memory; operands are rounded to BF16 (nearest even) when
// they are staged into LDS. C[M][N] = op(A)[M][K] op(B)[K][N], op = identity or transpose:
//   forward     Z = X W + b, H = gelu(Z)        (A = X, B = W)
//   backward    dX = dZ W^T                      (B transposed)
//               dW = X^T dZ                      (A transposed, K = number of faces)
// 64 x 64 tile per workgroup of 4 waves, one v_mfma_f32_32x32x8bf16_1k tile per wave and K step
// of 8; operand lane l holds P[l%32][4(l/32) .. 4(l/32)+3] and Q[4(l/32) .. +3][l%32], result
// register r holds D[(r%4) + 8(r/4) + 4(l/32)][l%32] (checked by mnp_probe_mfma_bf16).
#include "common.hpp"

namespace mnp {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);   // inf / nan kept
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

constexpr double INV_SQRT2 = 0.70710678118654752440;
constexpr double INV_SQRT2PI = 0.39894228040143267794;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * (float)INV_SQRT2)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * (float)INV_SQRT2)) + z * (float)INV_SQRT2PI * expf(-0.5f * z * z);
}

constexpr int MT = 64, MK = 32, MPAD = 4;   // tile, K slice (4 MFMA steps), LDS row padding (shorts)

// EPI: 0 = C = AB, 1 = C = AB accumulated into C, 2 = Z = AB + bias, H = gelu(Z), 3 = Z = AB + bias
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const float* __restrict__ A, long lda,
                                                  const float* __restrict__ B, long ldb, const float* __restrict__ bias,
                                                  float* __restrict__ Z, float* __restrict__ H, long ldc) {
    __shared__ short As[MT][MK + MPAD];   // As[i][k]
    __shared__ short Bs[MT][MK + MPAD];   // Bs[j][k]
    const int row0 = blockIdx.x * MT, col0 = blockIdx.y * MT;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0;
    for (int k0 = 0; k0 < K; k0 += MK) {
        // stage 64 x 32 of op(A) and 32 x 64 of op(B), 8 elements per lane each; consecutive lanes
        // follow the contiguous memory direction
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = t + 256 * q;
            int i, kk;
            if (TA) { i = e % MT; kk = e / MT; }        // A^T: memory A[k][i], i contiguous
            else { kk = e % MK; i = e / MK; }           // A: memory A[i][k], k contiguous
            const int gi = row0 + i, gk = k0 + kk;
            float v = 0.f;
            if (gi < M && gk < K) v = TA ? A[(long)gk * lda + gi] : A[(long)gi * lda + gk];
            As[i][kk] = to_bf16(v);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = t + 256 * q;
            int j, kk;
            if (TB) { kk = e % MK; j = e / MK; }        // B^T: memory B[j][k], k contiguous
            else { j = e % MT; kk = e / MT; }           // B: memory B[k][j], j contiguous
            const int gj = col0 + j, gk = k0 + kk;
            float v = 0.f;
            if (gj < N && gk < K) v = TB ? B[(long)gj * ldb + gk] : B[(long)gk * ldb + gj];
            Bs[j][kk] = to_bf16(v);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < MK; ks += 8) {
            const int kq = ks + 4 * (lane >> 5);
            bf4 a, b;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // MFMA fed (B^T, A^T): result rows run over C's columns, lanes over C's rows
                a[e] = Bs[wn + (lane & 31)][kq + e];
                b[e] = As[wm + (lane & 31)][kq + e];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int gi = row0 + wm + (lane & 31);
        const int gj = col0 + wn + (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
        if (gi < M && gj < N) {
            const long o = (long)gi * ldc + gj;
            float v = acc[r];
            if (EPI == 1) v += Z[o];
            if (EPI >= 2) v += bias[gj];
            Z[o] = v;
            if (EPI == 2) H[o] = gelu(v);
        }
    }
}

__global__ void k_gelu_bwd(long count, const float* dH, const float* Z, float* dZ) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// out[j] = sum_i X[i][j] (one thread per column, rows strided over the grid's y dimension)
__global__ void k_colsum(int M, int N, const float* X, long ld, float* out) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int part = threadIdx.x >> 6;
    __shared__ float red[4][64];
    float s = 0.f;
    if (j < N)
        for (int i = part; i < M; i += 4) s += X[(long)i * ld + j];
    red[part][threadIdx.x & 63] = s;
    __syncthreads();
    if (part == 0 && j < N) out[j] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// loss += sum (Y - T)^2 / count, dY = 2 (Y - T) / count
__global__ void k_mse(long count, const float* Y, const float* T, float* dY, float* loss) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < count) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d / (float)count;
        l = d * d / (float)count;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// Adam (Kingma & Ba 2015) with bias correction; p, g, m, v: count floats
__global__ void k_adam(long count, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero_f32(long count, float* p) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0.f;
}

// layout probe of v_mfma_f32_32x32x8bf16_1k: P[x][k] = x + 16 k, Q[k][y] = (k == 1) (y + 1)
__global__ void k_probe_bf16(float* out) {
    const int l = threadIdx.x;
    bf4 a, b;
    for (int e = 0; e < 4; ++e) {
        const int kk = 4 * (l / 32) + e;
        a[e] = to_bf16((float)((l % 32) + 16 * kk));
        b[e] = to_bf16(kk == 1 ? (float)((l % 32) + 1) : 0.f);
    }
    f16v c;
    for (int r = 0; r < 16; ++r) c[r] = 0;
    c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
    for (int r = 0; r < 16; ++r) out[l * 16 + r] = c[r];
}

// ------------------------------------------------------------------------------ launchers
template <bool TA, bool TB, int EPI>
static void launch_gemm(int M, int N, int K, const float* A, long lda, const float* B, long ldb, const float* bias,
                        float* Z, float* H, long ldc, hipStream_t s) {
    dim3 grid((M + MT - 1) / MT, (N + MT - 1) / MT);
    hipLaunchKernelGGL((k_mlp_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, Z, H, ldc);
}

int mlp_gemm(int ta, int tb, int epi, int M, int N, int K, const float* A, long lda, const float* B, long ldb,
             const float* bias, float* Z, float* H, long ldc, hipStream_t s) {
#define MLP_CASE(a, b, e)                                                                   \
    if (ta == a && tb == b && epi == e) {                                                   \
        launch_gemm<a, b, e>(M, N, K, A, lda, B, ldb, bias, Z, H, ldc, s);                  \
        return 0;                                                                           \
    }
    MLP_CASE(0, 0, 0) MLP_CASE(0, 0, 1) MLP_CASE(0, 0, 2) MLP_CASE(0, 0, 3)
    MLP_CASE(0, 1, 0) MLP_CASE(0, 1, 1)
    MLP_CASE(1, 0, 0) MLP_CASE(1, 0, 1)
#undef MLP_CASE
    return -1;
}

void mlp_gelu_bwd(long count, const float* dH, const float* Z, float* dZ, hipStream_t s) {
    hipLaunchKernelGGL(k_gelu_bwd, dim3((count + 255) / 256), dim3(256), 0, s, count, dH, Z, dZ);
}
void mlp_colsum(int M, int N, const float* X, long ld, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(256), 0, s, M, N, X, ld, out);
}
void mlp_mse(long count, const float* Y, const float* T, float* dY, float* loss, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, s, 1L, loss);
    hipLaunchKernelGGL(k_mse, dim3((count + 255) / 256), dim3(256), 0, s, count, Y, T, dY, loss);
}
void mlp_adam(long count, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
              int step, hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((count + 255) / 256), dim3(256), 0, s, count, p, g, m, v, lr, b1, b2, eps, c1, c2);
}
void mlp_probe(float* out, hipStream_t s) { hipLaunchKernelGGL(k_probe_bf16, dim3(1), dim3(64), 0, s, out); }

}  // namespace mnp
