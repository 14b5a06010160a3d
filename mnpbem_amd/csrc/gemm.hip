// Batched complex GEMM update C -= A * B on the matrix cores.
//
// Column-major, interleaved complex operands. A workgroup of 256 lanes (4 waves) owns a
// 64x64 tile of C; each wave a 32x32 quarter built from 2x2 MFMA 16x16 tiles. The K loop
// stages 16-deep slices of A and B through LDS split into real/imaginary planes (A stored
// negated, so the MFMA accumulation directly yields -A*B). The complex product uses four
// real MFMAs per 16x16x4 step (re*re, -im*im, re*im, im*re).
//
// The MFMA is fed (B^T, A^T) rather than (A, B): its result row index then runs over C's
// columns and its lane index over C's rows, so each group of 16 lanes writes 16 consecutive
// rows of one column (a contiguous 256 B / 128 B segment in column-major C).
//
// Lane layout of v_mfma_*_16x16x4 (64 lanes, l): operand-A lane l holds A[l%16][l/16],
// operand-B lane l holds B[l/16][l%16]; result register r of lane l holds D[row][l%16] with
// row = l/16 + 4 r for the f64 instruction and row = 4 (l/16) + r for the f32 one (measured
// on gfx950 by k_probe_layout; Mfma<T>::drow encodes it).
#include "common.hpp"

#include <vector>

namespace mnp {

// Optional per-launch timing of the GEMM (HIP events on the launch stream), for roofline
// reporting. Off by default; see mnp_prof_enable / mnp_prof_collect in capi.hip.
struct ProfRec {
    hipEvent_t a, b;
    double flops;
};
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_pool;

static hipEvent_t pool_get() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

void prof_enable(bool on) { g_prof_on = on; }

// Sum of GEMM kernel time (ms) and algorithmic flops since the last call; waits for the events.
void prof_collect(double* ms, double* flops, double* launches) {
    double t = 0, f = 0;
    for (auto& r : g_prof) {
        hipEventSynchronize(r.b);
        float e = 0;
        hipEventElapsedTime(&e, r.a, r.b);
        t += e;
        f += r.flops;
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    *ms = t;
    *flops = f;
    *launches = (double)g_prof.size();
    g_prof.clear();
}

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int GBM = 64, GBN = 64, GBK = 16, GPAD = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(int m, int n, int k, const cx<T>* __restrict__ A, long lda,
                                                  long sA, const cx<T>* __restrict__ B, long ldb, long sB,
                                                  cx<T>* __restrict__ C, long ldc, long sC) {
    using acc_t = typename Mfma<T>::acc_t;
    __shared__ T As_re[GBK][GBM + GPAD], As_im[GBK][GBM + GPAD];
    __shared__ T Bs_re[GBK][GBN + GPAD], Bs_im[GBK][GBN + GPAD];
    const int bz = blockIdx.z;
    A += (long)bz * sA;
    B += (long)bz * sB;
    C += (long)bz * sC;
    const int row0 = blockIdx.x * GBM, col0 = blockIdx.y * GBN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    acc_t cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }

    for (int k0 = 0; k0 < k; k0 += GBK) {
        // A slice: 64 rows x 16 k, thread -> row t%64, k = t/64 + 4q
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int ii = t & 63, kk = (t >> 6) + 4 * q;
            int gi = row0 + ii, gk = k0 + kk;
            cx<T> v = mk<T>(0, 0);
            if (gi < m && gk < k) v = A[gi + (long)gk * lda];
            As_re[kk][ii] = -v.re;
            As_im[kk][ii] = -v.im;
        }
        // B slice: 16 k x 64 cols, thread -> k = t%16, col = t/16 + 16q
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int kk = t & 15, jj = (t >> 4) + 16 * q;
            int gk = k0 + kk, gj = col0 + jj;
            cx<T> v = mk<T>(0, 0);
            if (gk < k && gj < n) v = B[gk + (long)gj * ldb];
            Bs_re[kk][jj] = v.re;
            Bs_im[kk][jj] = v.im;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 4) {
            int kk = ks + (lane >> 4);
            T are[2], aim[2], bre[2], bim[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                are[a] = As_re[kk][wm + a * 16 + (lane & 15)];
                aim[a] = As_im[kk][wm + a * 16 + (lane & 15)];
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                bre[b] = Bs_re[kk][wn + b * 16 + (lane & 15)];
                bim[b] = Bs_im[kk][wn + b * 16 + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // D[j][i] = sum_k B[k][j] * A[i][k]; A already negated
                    cre[a][b] = Mfma<T>::op(bre[b], are[a], cre[a][b]);
                    cre[a][b] = Mfma<T>::op(-bim[b], aim[a], cre[a][b]);
                    cim[a][b] = Mfma<T>::op(bim[b], are[a], cim[a][b]);
                    cim[a][b] = Mfma<T>::op(bre[b], aim[a], cim[a][b]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gi = row0 + wm + a * 16 + (lane & 15);
                int gj = col0 + wn + b * 16 + Mfma<T>::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<T>* c = C + gi + (long)gj * ldc;
                    cx<T> v = *c;
                    v.re += cre[a][b][r];
                    v.im += cim[a][b][r];
                    *c = v;
                }
            }
}

// C -= A * B for `batch` independent problems.
template <typename T>
void gemm_sub(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB, void* C,
              long ldc, long sC, int batch, hipStream_t s) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return;
    dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN, batch);
    ProfRec rec{};
    if (g_prof_on) {
        rec.a = pool_get();
        rec.b = pool_get();
        rec.flops = 8.0 * m * (double)n * k * batch;
        hipEventRecord(rec.a, s);
    }
    hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, s, m, n, k, (const cx<T>*)A, lda, sA, (const cx<T>*)B,
                       ldb, sB, (cx<T>*)C, ldc, sC);
    if (g_prof_on) {
        hipEventRecord(rec.b, s);
        g_prof.push_back(rec);
    }
}

template void gemm_sub<double>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                               int, hipStream_t);
template void gemm_sub<float>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                              int, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Lane-layout probe: D = A(16x4) * B(4x16) with A[i][k] = i + 16 k and B[k][j] = j + 1 for
// k = 1, 0 otherwise, so D[i][j] = (i + 16)(j + 1). One wave per instruction; out[0..255] =
// f64 result registers (lane * 4 + r), out[256..511] = f32 result registers.
__global__ void k_probe_layout(double* out) {
    int l = threadIdx.x;
    int i = l & 15, kk = l >> 4;
    {
        typedef double acc_t __attribute__((ext_vector_type(4)));
        double a = (double)(i + 16 * kk);
        double b = (kk == 1) ? (double)(i + 1) : 0.0;
        acc_t c = {0, 0, 0, 0};
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
    }
    {
        typedef float acc_t __attribute__((ext_vector_type(4)));
        float a = (float)(i + 16 * kk);
        float b = (kk == 1) ? (float)(i + 1) : 0.0f;
        acc_t c = {0, 0, 0, 0};
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[256 + l * 4 + r] = (double)c[r];
    }
}

void probe_layout(double* out, hipStream_t s) { hipLaunchKernelGGL(k_probe_layout, dim3(1), dim3(64), 0, s, out); }

}  // namespace mnp
