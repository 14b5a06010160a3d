// Batched complex GEMM update C -= A * B on the matrix cores.
//
// Column-major, interleaved complex operands. A workgroup of 256 lanes (4 waves) owns a
// 64x64 tile of C, each wave a 32x32 quarter: one v_mfma_f32_32x32x2f32 tile (complex float)
// or 2x2 v_mfma_f64_16x16x4f64 tiles (complex double). K slices of 16 are staged through LDS
// split into real/imaginary planes (A stored negated, so the accumulation yields -A*B), and
// the next slice is prefetched into registers while the current one is multiplied. A complex
// product takes four real MFMAs (re*re, -im*im, re*im, im*re).
//
// The MFMA is fed (B^T, A^T) rather than (A, B): its result row index then runs over C's
// columns and its lane index over C's rows, so each lane group writes consecutive rows of one
// column (a contiguous segment in column-major C). Lane layouts are in Mfma<T> below
// (measured on gfx950 by k_probe_layout).
#include "common.hpp"

#include <stdint.h>
#include <vector>

namespace mnp {

// Optional per-launch timing of the GEMM (HIP events on the launch stream), for roofline
// reporting. Off by default; see mnp_prof_enable / mnp_prof_collect in capi.hip.
struct ProfRec {
    hipEvent_t a, b;
    double flops, bytes;
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

// Sum of GEMM kernel time (ms), algorithmic flops and bytes since the last call; waits for the
// events.
void prof_collect(double* ms, double* flops, double* launches, double* bytes) {
    double t = 0, f = 0, by = 0;
    for (auto& r : g_prof) {
        hipEventSynchronize(r.b);
        float e = 0;
        hipEventElapsedTime(&e, r.a, r.b);
        t += e;
        f += r.flops;
        by += r.bytes;
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    *ms = t;
    *flops = f;
    *launches = (double)g_prof.size();
    *bytes = by;
    g_prof.clear();
}

// Matrix-core tile of one wave: 32 x 32 of C.
//   float : one v_mfma_f32_32x32x2f32 per real product, K step 2; operand a lane l holds
//           P[l%32][l/32], operand b lane l holds Q[l/32][l%32]; result register r of lane l
//           holds D[(r%4) + 8(r/4) + 4(l/32)][l%32] (measured on gfx950).
//   double: 2 x 2 v_mfma_f64_16x16x4f64, K step 4; operand a lane l holds P[l%16][l/16],
//           operand b lane l holds Q[l/16][l%16]; result register r of lane l holds
//           D[l/16 + 4r][l%16] (measured on gfx950 by k_probe_layout).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static constexpr int KS = 4, TILE = 16, NT = 2, NR = 4;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(16)));
    static constexpr int KS = 2, TILE = 32, NT = 1, NR = 16;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5); }
};

constexpr int GBM = 64, GBN = 64, GBK = 16, GPAD = 4;
constexpr int GNA = GBM * GBK / 256, GNB = GBK * GBN / 256;   // complex loads per lane per K slice

// C -= A B: 256 lanes (4 waves, 2 x 2) per 64 x 64 tile of C; K slices of 16 staged through
// LDS as real/imaginary planes (A negated, so accumulation yields -A B directly); the next
// slice is loaded into registers while the current one is multiplied. Each complex product
// takes four real MFMAs. The MFMA is fed (B^T, A^T) so that its lane index runs over C's rows
// and the stores of a lane group are contiguous in column-major C.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(int m, int n, int k, const cx<T>* __restrict__ A, long lda,
                                                  long sA, const cx<T>* __restrict__ B, long ldb, long sB,
                                                  cx<T>* __restrict__ C, long ldc, long sC) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int TL = M::TILE, NT = M::NT, KS = M::KS;
    __shared__ T As_re[GBK][GBM + GPAD], As_im[GBK][GBM + GPAD];
    __shared__ T Bs_re[GBK][GBN + GPAD], Bs_im[GBK][GBN + GPAD];
    const int bz = blockIdx.z;
    A += (long)bz * sA;
    B += (long)bz * sB;
    C += (long)bz * sC;
    const int row0 = blockIdx.x * GBM, col0 = blockIdx.y * GBN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    acc_t cre[NT][NT], cim[NT][NT];
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < M::NR; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }
    cx<T> pa[GNA], pb[GNB];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < GNA; ++q) {
            const int idx = t + 256 * q, ii = idx % GBM, kk = idx / GBM;
            const int gi = row0 + ii, gk = k0 + kk;
            pa[q] = (gi < m && gk < k) ? A[gi + (long)gk * lda] : mk<T>(0, 0);
        }
#pragma unroll
        for (int q = 0; q < GNB; ++q) {
            const int idx = t + 256 * q, kk = idx % GBK, jj = idx / GBK;
            const int gk = k0 + kk, gj = col0 + jj;
            pb[q] = (gk < k && gj < n) ? B[gk + (long)gj * ldb] : mk<T>(0, 0);
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < GNA; ++q) {
            const int idx = t + 256 * q, ii = idx % GBM, kk = idx / GBM;
            As_re[kk][ii] = -pa[q].re;
            As_im[kk][ii] = -pa[q].im;
        }
#pragma unroll
        for (int q = 0; q < GNB; ++q) {
            const int idx = t + 256 * q, kk = idx % GBK, jj = idx / GBK;
            Bs_re[kk][jj] = pb[q].re;
            Bs_im[kk][jj] = pb[q].im;
        }
        __syncthreads();
        if (k0 + GBK < k) load(k0 + GBK);
#pragma unroll
        for (int ks = 0; ks < GBK; ks += KS) {
            const int kk = ks + lane / TL;
            T are[NT], aim[NT], bre[NT], bim[NT];
#pragma unroll
            for (int a = 0; a < NT; ++a) {
                are[a] = As_re[kk][wm + a * TL + lane % TL];
                aim[a] = As_im[kk][wm + a * TL + lane % TL];
            }
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                bre[b] = Bs_re[kk][wn + b * TL + lane % TL];
                bim[b] = Bs_im[kk][wn + b * TL + lane % TL];
            }
#pragma unroll
            for (int a = 0; a < NT; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b) {
                    // D[j][i] = sum_k B[k][j] * A[i][k]; A already negated
                    cre[a][b] = M::op(bre[b], are[a], cre[a][b]);
                    cim[a][b] = M::op(bim[b], are[a], cim[a][b]);
                    cre[a][b] = M::op(-bim[b], aim[a], cre[a][b]);
                    cim[a][b] = M::op(bre[b], aim[a], cim[a][b]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < M::NR; ++r) {
                const int gi = row0 + wm + a * TL + lane % TL;
                const int gj = col0 + wn + b * TL + M::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<T>* c = C + gi + (long)gj * ldc;
                    cx<T> v = *c;
                    v.re += cre[a][b][r];
                    v.im += cim[a][b][r];
                    *c = v;
                }
            }
}

// Complex-float variant for 16-byte aligned operands (even leading dimensions): K slices of
// 8, each lane loads two consecutive complex values of A (rows i, i+1) or B (k, k+1) with one
// 16-byte load, LDS holds interleaved (re, im) pairs read with one 8-byte load per operand.
constexpr int VBK = 8;
__global__ __launch_bounds__(256) void k_gemm_sub_c64v(int m, int n, int k, const cx<float>* __restrict__ A,
                                                       long lda, long sA, const cx<float>* __restrict__ B, long ldb,
                                                       long sB, cx<float>* __restrict__ C, long ldc, long sC) {
    typedef float acc16 __attribute__((ext_vector_type(16)));
    __shared__ float2 As[VBK][GBM + 2];
    __shared__ float2 Bs[VBK][GBN + 2];
    const int bz = blockIdx.z;
    A += (long)bz * sA;
    B += (long)bz * sB;
    C += (long)bz * sC;
    const int row0 = blockIdx.x * GBM, col0 = blockIdx.y * GBN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    acc16 cre, cim;
#pragma unroll
    for (int r = 0; r < 16; ++r) { cre[r] = 0; cim[r] = 0; }
    // lane t moves rows 2 (t % 32), 2 (t % 32) + 1 of A at k = t / 32, and k = 2 (t % 4), 2 (t % 4) + 1
    // of B at column t / 4
    const int ai = 2 * (t % 32), ak = t / 32, bk = 2 * (t % 4), bj = t / 4;
    float4 pa, pb;
    auto load = [&](int k0) {
        const int gi = row0 + ai, gk = k0 + ak;
        pa = make_float4(0, 0, 0, 0);
        if (gk < k) {
            const cx<float>* p = A + gi + (long)gk * lda;
            if (gi + 1 < m) pa = *(const float4*)p;
            else if (gi < m) { pa.x = p->re; pa.y = p->im; }
        }
        const int gk2 = k0 + bk, gj = col0 + bj;
        pb = make_float4(0, 0, 0, 0);
        if (gj < n) {
            const cx<float>* p = B + gk2 + (long)gj * ldb;
            if (gk2 + 1 < k) pb = *(const float4*)p;
            else if (gk2 < k) { pb.x = p->re; pb.y = p->im; }
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += VBK) {
        As[ak][ai] = make_float2(-pa.x, -pa.y);
        As[ak][ai + 1] = make_float2(-pa.z, -pa.w);
        Bs[bk][bj] = make_float2(pb.x, pb.y);
        Bs[bk + 1][bj] = make_float2(pb.z, pb.w);
        __syncthreads();
        if (k0 + VBK < k) load(k0 + VBK);
#pragma unroll
        for (int ks = 0; ks < VBK; ks += 2) {
            const int kk = ks + (lane >> 5);
            const float2 a = As[kk][wm + (lane & 31)];
            const float2 b = Bs[kk][wn + (lane & 31)];
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(b.x, a.x, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x2f32(b.y, a.x, cim, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(-b.y, a.y, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x2f32(b.x, a.y, cim, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int gi = row0 + wm + (lane & 31);
        const int gj = col0 + wn + Mfma<float>::drow(lane, r);
        if (gi < m && gj < n) {
            cx<float>* c = C + gi + (long)gj * ldc;
            cx<float> v = *c;
            v.re += cre[r];
            v.im += cim[r];
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
        // A and B read once, C read and written once
        rec.bytes = 2.0 * sizeof(T) * ((double)m * k + (double)k * n + 2.0 * m * n) * batch;
        hipEventRecord(rec.a, s);
    }
    // 16-byte loads need even leading dimensions and batch strides and 16-byte aligned bases
    const bool vec = sizeof(T) == 4 && ((lda | ldb | sA | sB) & 1) == 0 && ((uintptr_t)A & 15) == 0 &&
                     ((uintptr_t)B & 15) == 0;
    if (vec)
        hipLaunchKernelGGL(k_gemm_sub_c64v, grid, dim3(256), 0, s, m, n, k, (const cx<float>*)A, lda, sA,
                           (const cx<float>*)B, ldb, sB, (cx<float>*)C, ldc, sC);
    else
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
// Lane-layout probe: D = P * Q with P[x][k] = x + 16 k (16x16x4) or x + 16 k (32x32x2) and
// Q[k][y] = y + 1 for k = 1, 0 otherwise, so D[x][y] = (x + 16)(y + 1). One wave per
// instruction; out[0..255] = f64 16x16x4 result registers (lane * 4 + r), out[256..511] =
// f32 16x16x4, out[512..1535] = f32 32x32x2 (lane * 16 + r).
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
    {
        typedef float acc_t __attribute__((ext_vector_type(16)));
        const int x = l & 31, k2 = l >> 5;
        float a = (float)(x + 16 * k2);
        float b = (k2 == 1) ? (float)(x + 1) : 0.0f;
        acc_t c;
        for (int r = 0; r < 16; ++r) c[r] = 0;
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
        for (int r = 0; r < 16; ++r) out[512 + l * 16 + r] = (double)c[r];
    }
}

void probe_layout(double* out, hipStream_t s) { hipLaunchKernelGGL(k_probe_layout, dim3(1), dim3(64), 0, s, out); }

}  // namespace mnp
