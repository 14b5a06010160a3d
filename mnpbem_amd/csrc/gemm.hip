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
    int plane;                              // 1: the split-BF16 plane kernel
};
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;
static ProfRec g_prof_pending;              // record being filled by the plane GEMM path
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

// Sum of GEMM kernel time (ms), algorithmic flops and bytes since the last call, over the
// launches of the split-BF16 plane kernel when there were any, else over all; waits for the
// events.
void prof_collect(double* ms, double* flops, double* launches, double* bytes) {
    double t = 0, f = 0, by = 0, nl = 0;
    bool any_plane = false;
    for (auto& r : g_prof) any_plane |= r.plane != 0;
    for (auto& r : g_prof) {
        if (!any_plane || r.plane) {
            hipEventSynchronize(r.b);
            float e = 0;
            hipEventElapsedTime(&e, r.a, r.b);
            t += e;
            f += r.flops;
            by += r.bytes;
            nl += 1;
        }
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    hipDeviceSynchronize();
    *ms = t;
    *flops = f;
    *launches = nl;
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

// ---------------------------------------------------------------------------------------------
// Complex-float GEMM on the BF16 matrix cores with split operands. Every float x is written as
// a sum of NS bfloat16 terms (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)), and a
// real product keeps the terms of order 0 and 1 (NS = 2: x0y0 + x0y1 + x1y0, about 16
// significant bits) or up to order 2 (NS = 3: six products, about 24 bits, the float's own
// precision), accumulated in FP32 by v_mfma_f32_32x32x16_bf16. The v_mfma_f32_32x32x16_bf16
// rate is 16x the FP32 MFMA rate, so even the six-product form leaves 2.7x headroom.
//
// Workgroup of 4 waves owns a 128x128 tile of C, each wave 64x64 (2x2 MFMA tiles of 32x32,
// real and imaginary accumulators). K slices of 16: each thread loads 8 complex values of A
// (one row, 8 consecutive k) or B (one column, 8 consecutive k), splits them and writes one
// 16-byte row segment per term and component into LDS rows of 16 k (+8 padding: 48-byte row
// stride, so 16 lanes of a 16-byte read hit distinct banks). A is stored negated (C -= A B);
// the imaginary term needs +Ai once, which is one sign flip per fragment in registers.
// Operand fragments: lane l reads 8 consecutive k at row l % 32, k offset 8 (l / 32); the MFMA
// is fed (B^T, A^T) as the FP32 kernels, so results store as contiguous column segments.
// Used for even leading dimensions and 16-byte aligned A, B (the other cases keep the FP32
// kernels).
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f16acc __attribute__((ext_vector_type(16)));
constexpr int ST = 128, SK = 16, SKP = 24;    // tile, K slice, LDS row length (bf16)

template <int NS>
__global__ __launch_bounds__(256) void k_gemm_sub_c64s(int m, int n, int k, const cx<float>* __restrict__ A, long lda,
                                                       long sA, const cx<float>* __restrict__ B, long ldb, long sB,
                                                       cx<float>* __restrict__ C, long ldc, long sC) {
    // [term][re/im][row][k]
    __shared__ __attribute__((aligned(16))) __bf16 As[NS][2][ST][SKP];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NS][2][ST][SKP];
    const int bz = blockIdx.z;
    A += (long)bz * sA;
    B += (long)bz * sB;
    C += (long)bz * sC;
    const int row0 = blockIdx.x * ST, col0 = blockIdx.y * ST;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    f16acc cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }
    // loads: A row ai, k = 8 akg .. +7; B column bj, k = 8 bkg .. +7
    const int ai = t & 127, akg = t >> 7;
    const int bj = t >> 1, bkg = t & 1;
    float2 va[8];
    float4 vb[4];
    auto load = [&](int k0) {
        const int gi = row0 + ai;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int gk = k0 + 8 * akg + e;
            va[e] = make_float2(0.f, 0.f);
            if (gi < m && gk < k) {
                const cx<float> v = A[gi + (long)gk * lda];
                va[e] = make_float2(v.re, v.im);
            }
        }
        const int gj = col0 + bj, gk = k0 + 8 * bkg;
#pragma unroll
        for (int e = 0; e < 4; ++e) vb[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gj < n) {
            const cx<float>* p = B + gk + (long)gj * ldb;
            if (gk + 8 <= k) {
#pragma unroll
                for (int e = 0; e < 4; ++e) vb[e] = ((const float4*)p)[e];
            } else {
                for (int e = 0; e < 8; ++e)
                    if (gk + e < k) {
                        const float2 w = make_float2(p[e].re, p[e].im);
                        if (e & 1) { vb[e >> 1].z = w.x; vb[e >> 1].w = w.y; }
                        else { vb[e >> 1].x = w.x; vb[e >> 1].y = w.y; }
                    }
            }
        }
    };
    // split 8 values into NS bf16 terms and store each term as one 16-byte row segment
    auto split_store = [&](const float (&x)[8], __bf16* dst0, long term_stride) {
        float r[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) r[e] = x[e];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            bf8v h;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                h[e] = (__bf16)r[e];
                r[e] -= (float)h[e];
            }
            *(bf8v*)(dst0 + q * term_stride) = h;
        }
    };
    const long tstride = 2L * ST * SKP;      // elements between terms
    load(0);
    for (int k0 = 0; k0 < k; k0 += SK) {
        {
            float xr[8], xi[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) { xr[e] = -va[e].x; xi[e] = -va[e].y; }
            split_store(xr, &As[0][0][ai][8 * akg], tstride);
            split_store(xi, &As[0][1][ai][8 * akg], tstride);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                xr[2 * e] = vb[e].x; xi[2 * e] = vb[e].y;
                xr[2 * e + 1] = vb[e].z; xi[2 * e + 1] = vb[e].w;
            }
            split_store(xr, &Bs[0][0][bj][8 * bkg], tstride);
            split_store(xi, &Bs[0][1][bj][8 * bkg], tstride);
        }
        __syncthreads();
        if (k0 + SK < k) load(k0 + SK);
        const int fr = lane & 31, fk = 8 * (lane >> 5);
        bf8v pr[2][NS], pi[2][NS];
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                pr[tj][q] = *(const bf8v*)&Bs[q][0][wn + 32 * tj + fr][fk];
                pi[tj][q] = *(const bf8v*)&Bs[q][1][wn + 32 * tj + fr][fk];
            }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
            bf8v qr[NS], qi[NS], qp[NS];      // -Ar, -Ai, +Ai terms
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                qr[q] = *(const bf8v*)&As[q][0][wm + 32 * ti + fr][fk];
                qi[q] = *(const bf8v*)&As[q][1][wm + 32 * ti + fr][fk];
                typedef unsigned u4v __attribute__((ext_vector_type(4)));
                u4v w = __builtin_bit_cast(u4v, qi[q]);
                w ^= 0x80008000u;
                qp[q] = __builtin_bit_cast(bf8v, w);
            }
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                // term pairs (p, q) with p + q < NS: order 0 .. NS - 1
#pragma unroll
                for (int p = 0; p < NS; ++p)
#pragma unroll
                    for (int q = 0; q < NS - p; ++q) {
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[tj][p], qr[q], cre[ti][tj], 0, 0, 0);
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[tj][p], qp[q], cre[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[tj][p], qr[q], cim[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[tj][p], qi[q], cim[ti][tj], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = row0 + wm + 32 * ti + (lane & 31);
                const int gj = col0 + wn + 32 * tj + Mfma<float>::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<float>* c = C + gi + (long)gj * ldc;
                    cx<float> v = *c;
                    v.re += cre[ti][tj][r];
                    v.im += cim[ti][tj][r];
                    *c = v;
                }
            }
}

// ---------------------------------------------------------------------------------------------
// Pre-split operands. The split of the kernel above runs once per loaded tile (every column
// tile of C re-splits the same rows of A); here A and B are split once into bfloat16 planes,
// and the GEMM only copies 16-byte chunks. Plane layout (per batch entry): element (qc, r, k)
// of plane qc = 2 term + component (0 re, 1 im), row r (A: row of A, B: column of B), k at
// (((qc * KS + k / 16) * 2 + (k / 8) % 2) * R + r) * 8 + k % 8, KS = ceil(K / 16), R = rows
// padded to 128: one k slice of 128 rows is 4 KB contiguous, as two halves of 8 k in which
// consecutive rows are 16 bytes apart (consecutive lanes read consecutive rows coalesced and
// store them to LDS rows 48 bytes apart, conflict-free). Padding rows and k are zero. A is
// stored negated (C -= A B).
// One thread per plane row and k slice of 16: reads 16 complex values (TRANS: 128 contiguous
// bytes of a column of B; else one element per k, coalesced across the threads' rows of A) and
// writes 16 bytes per half slice, term and component (consecutive threads: consecutive 16 bytes).
template <int NS, bool TRANS>
__global__ __launch_bounds__(256) void k_split_planes(int rows, int K, const cx<float>* __restrict__ X, long ldx,
                                                      long sX, __bf16* __restrict__ P, int R, long sP, float sgn,
                                                      int vec) {
    const int r = blockIdx.x * 256 + threadIdx.x;      // row of the plane (< R)
    const int s = blockIdx.y;                            // k slice
    const int KS = gridDim.y;
    X += (long)blockIdx.z * sX;
    P += (long)blockIdx.z * sP;
    if (r >= R) return;
    float re[16], im[16];
    const int k0 = 16 * s;
    if (TRANS && vec && r < rows && k0 + 16 <= K) {
        const float4* p = (const float4*)(X + k0 + (long)r * ldx);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float4 v = p[e];
            re[2 * e] = sgn * v.x;
            im[2 * e] = sgn * v.y;
            re[2 * e + 1] = sgn * v.z;
            im[2 * e + 1] = sgn * v.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int k = k0 + e;
            re[e] = 0.f;
            im[e] = 0.f;
            if (r < rows && k < K) {
                const cx<float> v = TRANS ? X[k + (long)r * ldx] : X[r + (long)k * ldx];
                re[e] = sgn * v.re;
                im[e] = sgn * v.im;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float* x = c ? im : re;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            bf8v h0, h1;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                h0[e] = (__bf16)x[e];
                x[e] -= (float)h0[e];
                h1[e] = (__bf16)x[8 + e];
                x[8 + e] -= (float)h1[e];
            }
            const long qc = 2 * q + c;
            *(bf8v*)(P + (((qc * KS + s) * 2) * R + r) * 8) = h0;
            *(bf8v*)(P + (((qc * KS + s) * 2 + 1) * R + r) * 8) = h1;
        }
    }
}

// XCD-aware tile order. Workgroups go to the 8 XCDs round-robin in dispatch order (linear id
// x fastest), and each XCD has its own L2. With gm > 0 the tile of a workgroup is chosen so that
// each XCD walks one contiguous range of the tile sequence, the sequence running through groups
// of gm row tiles column by column: the tiles in flight on one XCD then share few A row tiles
// and B column tiles. gm = 0 keeps the plain grid order.
__device__ inline void xcd_tile(int gm, int& bx, int& by, int& bz) {
    bx = blockIdx.x;
    by = blockIdx.y;
    bz = blockIdx.z;
    if (gm <= 0) return;
    const long X = gridDim.x, Y = gridDim.y, total = X * Y * (long)gridDim.z;
    const long L = blockIdx.x + X * (blockIdx.y + Y * (long)blockIdx.z);
    const long q = total / 8, r = total % 8, xcd = L % 8, j = L / 8;
    const long T = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
    bz = (int)(T / (X * Y));
    const long u = T % (X * Y), g = u / (gm * Y), x0 = g * gm;
    const long rows = (X - x0 < gm) ? X - x0 : gm;
    const long w = u - g * gm * Y;
    bx = (int)(x0 + w % rows);
    by = (int)(w / rows);
}

// Workgroup of WR x WC waves (64 WR WC lanes), each wave MI x NJ MFMA tiles of 32 x 32: a
// (32 MI WR) x (32 NJ WC) tile of C.
template <int NS, int MI, int NJ, int WR = 2, int WC = 2, int NB = 1>
__global__ __launch_bounds__(64 * WR * WC) void k_gemm_sub_c64p(int m, int n, int K, const __bf16* __restrict__ PA,
                                                                int RA, long sPA, const __bf16* __restrict__ PB,
                                                                int RB, long sPB, cx<float>* __restrict__ C, long ldc,
                                                                long sC, int gm) {
    constexpr int NT = 64 * WR * WC;
    constexpr int TM = 32 * MI * WR, TN = 32 * NJ * WC;
    constexpr int CA = 4 * NS * TM, CB = 4 * NS * TN;          // 16-byte chunks per k slice
    constexpr int UA = (CA + NT - 1) / NT, UB = (CB + NT - 1) / NT;
    // NB = 2: two LDS buffers, the next k slice is stored while the current one is read (one
    // barrier per slice)
    __shared__ __attribute__((aligned(16))) __bf16 As[NB][NS][2][TM][SKP];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NB][NS][2][TN][SKP];
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    int tx, ty, bz;
    xcd_tile(gm, tx, ty, bz);
    PA += (long)bz * sPA;
    PB += (long)bz * sPB;
    C += (long)bz * sC;
    const int KS = (K + 15) / 16;
    const int row0 = tx * TM, col0 = ty * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave % WR) * 32 * MI, wn = (wave / WR) * 32 * NJ;
    f16acc cre[MI][NJ], cim[MI][NJ];
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int b = 0; b < NJ; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }
    // chunk e = t + NT u of the A (B) tile: plane e / (2 TM), half (e % (2 TM)) / TM, row e % TM
    u4v ra[UA], rb[UB];
    auto load = [&](int s) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int e = t + NT * u, qc = e / (2 * TM), w = e % (2 * TM);
            if (CA % NT == 0 || e < CA)
                ra[u] = *(const u4v*)(PA + ((((long)qc * KS + s) * 2 + w / TM) * RA + row0 + w % TM) * 8);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = t + NT * u, qc = e / (2 * TN), w = e % (2 * TN);
            if (CB % NT == 0 || e < CB)
                rb[u] = *(const u4v*)(PB + ((((long)qc * KS + s) * 2 + w / TN) * RB + col0 + w % TN) * 8);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int e = t + NT * u, qc = e / (2 * TM), w = e % (2 * TM);
            if (CA % NT == 0 || e < CA) *(u4v*)&As[buf][qc >> 1][qc & 1][w % TM][8 * (w / TM)] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = t + NT * u, qc = e / (2 * TN), w = e % (2 * TN);
            if (CB % NT == 0 || e < CB) *(u4v*)&Bs[buf][qc >> 1][qc & 1][w % TN][8 * (w / TN)] = rb[u];
        }
    };
    auto compute = [&](int buf) {
        const int fr = lane & 31, fk = 8 * (lane >> 5);
        bf8v pr[NJ][NS], pi[NJ][NS], qr[MI][NS], qi[MI][NS], qp[MI][NS];
#pragma unroll
        for (int tj = 0; tj < NJ; ++tj)
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                pr[tj][q] = *(const bf8v*)&Bs[buf][q][0][wn + 32 * tj + fr][fk];
                pi[tj][q] = *(const bf8v*)&Bs[buf][q][1][wn + 32 * tj + fr][fk];
            }
#pragma unroll
        for (int ti = 0; ti < MI; ++ti)
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                qr[ti][q] = *(const bf8v*)&As[buf][q][0][wm + 32 * ti + fr][fk];
                qi[ti][q] = *(const bf8v*)&As[buf][q][1][wm + 32 * ti + fr][fk];
                u4v w = __builtin_bit_cast(u4v, qi[ti][q]);
                w ^= 0x80008000u;
                qp[ti][q] = __builtin_bit_cast(bf8v, w);
            }
        // term pairs (p, q), p + q < NS; the four real products of each pair go to 2 MI NJ
        // independent accumulators in turn
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
            for (int q = 0; q < NS - p; ++q) {
#pragma unroll
                for (int ti = 0; ti < MI; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NJ; ++tj) {
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[tj][p], qr[ti][q], cre[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[tj][p], qr[ti][q], cim[ti][tj], 0, 0, 0);
                    }
#pragma unroll
                for (int ti = 0; ti < MI; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NJ; ++tj) {
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[tj][p], qp[ti][q], cre[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[tj][p], qi[ti][q], cim[ti][tj], 0, 0, 0);
                    }
            }
    };
    load(0);
    if (NB == 1) {
        for (int s = 0; s < KS; ++s) {
            store(0);
            __syncthreads();
            if (s + 1 < KS) load(s + 1);
            compute(0);
            __syncthreads();
        }
    } else {
        store(0);
        __syncthreads();
        for (int s = 0; s < KS; ++s) {
            if (s + 1 < KS) load(s + 1);
            compute(s & 1);
            if (s + 1 < KS) store((s + 1) & 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int ti = 0; ti < MI; ++ti)
#pragma unroll
        for (int tj = 0; tj < NJ; ++tj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = row0 + wm + 32 * ti + (lane & 31);
                const int gj = col0 + wn + 32 * tj + Mfma<float>::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<float>* c = C + gi + (long)gj * ldc;
                    cx<float> v = *c;
                    v.re += cre[ti][tj][r];
                    v.im += cim[ti][tj][r];
                    *c = v;
                }
            }
}

// Variant without LDS: in the plane layout lane l's 16 bytes of an MFMA operand (row l % 32,
// k 8 (l / 32) .. +7) lie in two 512-byte runs (one per half slice), so each wave loads its
// operands straight from memory with one coalesced 16-byte load per operand fragment (the two waves sharing a row or column tile meet in the L1),
// and the next k slice is loaded while the current one is multiplied. No barriers.
template <int NS, int MI, int NJ>
__global__ __launch_bounds__(256) void k_gemm_sub_c64d(int m, int n, int K, const __bf16* __restrict__ PA, int RA,
                                                       long sPA, const __bf16* __restrict__ PB, int RB, long sPB,
                                                       cx<float>* __restrict__ C, long ldc, long sC, int gm) {
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    int tx, ty, bz;
    xcd_tile(gm, tx, ty, bz);
    PA += (long)bz * sPA;
    PB += (long)bz * sPB;
    C += (long)bz * sC;
    const int KS = (K + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = tx * 64 * MI + (wave & 1) * 32 * MI, wn = ty * 64 * NJ + (wave >> 1) * 32 * NJ;
    const int lr = lane & 31, lh = lane >> 5;
    f16acc cre[MI][NJ], cim[MI][NJ];
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int b = 0; b < NJ; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }
    // fragments of one k slice: A[ti][qc], B[tj][qc]
    u4v fa[2][MI][2 * NS], fb[2][NJ][2 * NS];
    auto load = [&](int s, int buf) {
#pragma unroll
        for (int qc = 0; qc < 2 * NS; ++qc) {
#pragma unroll
            for (int ti = 0; ti < MI; ++ti)
                fa[buf][ti][qc] = *(const u4v*)(PA + ((((long)qc * KS + s) * 2 + lh) * RA + wm + 32 * ti + lr) * 8);
#pragma unroll
            for (int tj = 0; tj < NJ; ++tj)
                fb[buf][tj][qc] = *(const u4v*)(PB + ((((long)qc * KS + s) * 2 + lh) * RB + wn + 32 * tj + lr) * 8);
        }
    };
    auto step = [&](int buf) {
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
            for (int q = 0; q < NS - p; ++q) {
#pragma unroll
                for (int ti = 0; ti < MI; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NJ; ++tj) {
                        const bf8v pr = __builtin_bit_cast(bf8v, fb[buf][tj][2 * p]);
                        const bf8v pi = __builtin_bit_cast(bf8v, fb[buf][tj][2 * p + 1]);
                        const bf8v qr = __builtin_bit_cast(bf8v, fa[buf][ti][2 * q]);
                        const bf8v qi = __builtin_bit_cast(bf8v, fa[buf][ti][2 * q + 1]);
                        const bf8v qp = __builtin_bit_cast(bf8v, fa[buf][ti][2 * q + 1] ^ 0x80008000u);
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr, qr, cre[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi, qr, cim[ti][tj], 0, 0, 0);
                        cre[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi, qp, cre[ti][tj], 0, 0, 0);
                        cim[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr, qi, cim[ti][tj], 0, 0, 0);
                    }
            }
    };
    load(0, 0);
    int s = 0;
    for (; s + 2 <= KS; s += 2) {
        load(s + 1, 1);
        step(0);
        if (s + 2 < KS) load(s + 2, 0);
        step(1);
    }
    if (s < KS) step(0);
#pragma unroll
    for (int ti = 0; ti < MI; ++ti)
#pragma unroll
        for (int tj = 0; tj < NJ; ++tj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = wm + 32 * ti + (lane & 31);
                const int gj = wn + 32 * tj + Mfma<float>::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<float>* c = C + gi + (long)gj * ldc;
                    cx<float> v = *c;
                    v.re += cre[ti][tj][r];
                    v.im += cim[ti][tj][r];
                    *c = v;
                }
            }
}

// Workgroup tile of the plane GEMM: 11 (64 x 64), 21 (128 x 64), 12 (64 x 128), 22 (128 x 128).
static int g_plane_tile = getenv("MNP_GEMM_TILE") ? atoi(getenv("MNP_GEMM_TILE")) : 11;
// Row tiles per group of the XCD-aware order (0: plain grid order).
static int g_plane_group = getenv("MNP_GEMM_GROUP") ? atoi(getenv("MNP_GEMM_GROUP")) : 8;

// Plane workspace per stream (the LU runs GEMMs on two streams at once), grown on demand.
struct PlaneWs {
    hipStream_t s;
    void* p;
    size_t bytes;
};
static std::vector<PlaneWs> g_plane_ws;
static void* plane_ws(hipStream_t s, size_t bytes) {
    for (auto& w : g_plane_ws)
        if (w.s == s) {
            if (w.bytes < bytes) {
                hipStreamSynchronize(s);
                hipFree(w.p);
                w.p = nullptr;
                if (hipMalloc(&w.p, bytes) != hipSuccess) {
                    w.p = nullptr;
                    w.bytes = 0;
                    return nullptr;
                }
                w.bytes = bytes;
            }
            return w.p;
        }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    g_plane_ws.push_back(PlaneWs{s, p, bytes});
    return p;
}

template <int NS>
static bool gemm_planes(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB,
                        void* C, long ldc, long sC, int batch, hipStream_t s) {
    const int KS = (k + 15) / 16;
    const int RA = (m + ST - 1) / ST * ST, RB = (n + ST - 1) / ST * ST;
    const long sPA = 2L * NS * KS * RA * 16, sPB = 2L * NS * KS * RB * 16;
    __bf16* P = (__bf16*)plane_ws(s, (size_t)(sPA + sPB) * batch * sizeof(__bf16));
    if (!P) return false;
    __bf16* PA = P;
    __bf16* PB = P + sPA * batch;
    ProfRec* rec = g_prof_on ? &g_prof_pending : nullptr;
    const int vb = ((ldb | sB) & 1) == 0 && ((uintptr_t)B & 15) == 0;   // 16-byte loads along a column of B
    hipLaunchKernelGGL((k_split_planes<NS, false>), dim3(RA / 256 + (RA % 256 != 0), KS, batch), dim3(256), 0, s, m,
                       k, (const cx<float>*)A, lda, sA, PA, RA, sPA, -1.0f, 0);
    hipLaunchKernelGGL((k_split_planes<NS, true>), dim3(RB / 256 + (RB % 256 != 0), KS, batch), dim3(256), 0, s, n,
                       k, (const cx<float>*)B, ldb, sB, PB, RB, sPB, 1.0f, vb);
    if (rec) hipEventRecord(rec->a, s);    // the timed region is the GEMM kernel alone
    switch (g_plane_tile) {
        case 11:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, s,
                               m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        case 21:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 2, 1>), dim3((m + 127) / 128, (n + 63) / 64, batch), dim3(256), 0,
                               s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        case 12:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 2>), dim3((m + 63) / 64, (n + 127) / 128, batch), dim3(256), 0,
                               s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        case 3011:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1, 2, 2, 2>), dim3((m + 63) / 64, (n + 63) / 64, batch),
                               dim3(256), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 3142:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1, 4, 2, 2>), dim3((m + 127) / 128, (n + 63) / 64, batch),
                               dim3(512), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 1142:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1, 4, 2>), dim3((m + 127) / 128, (n + 63) / 64, batch),
                               dim3(512), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 1124:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1, 2, 4>), dim3((m + 63) / 64, (n + 127) / 128, batch),
                               dim3(512), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 1144:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 1, 4, 4>), dim3((m + 127) / 128, (n + 127) / 128, batch),
                               dim3(1024), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 1242:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 1, 2, 4, 2>), dim3((m + 127) / 128, (n + 127) / 128, batch),
                               dim3(512), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 2124:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 2, 1, 2, 4>), dim3((m + 127) / 128, (n + 127) / 128, batch),
                               dim3(512), 0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC,
                               g_plane_group);
            break;
        case 111:
            hipLaunchKernelGGL((k_gemm_sub_c64d<NS, 1, 1>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, s,
                               m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        case 121:
            hipLaunchKernelGGL((k_gemm_sub_c64d<NS, 2, 1>), dim3((m + 127) / 128, (n + 63) / 64, batch), dim3(256), 0,
                               s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        case 112:
            hipLaunchKernelGGL((k_gemm_sub_c64d<NS, 1, 2>), dim3((m + 63) / 64, (n + 127) / 128, batch), dim3(256), 0,
                               s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
            break;
        default:
            hipLaunchKernelGGL((k_gemm_sub_c64p<NS, 2, 2>), dim3((m + 127) / 128, (n + 127) / 128, batch), dim3(256),
                               0, s, m, n, k, PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, g_plane_group);
    }
    return true;
}

// 0: FP32 matrix cores; 2 or 3: BF16 matrix cores with that many operand terms, split per
// tile; 12 or 13: the same with operands split once into planes.
static int g_gemm_split = getenv("MNP_GEMM_SPLIT") ? atoi(getenv("MNP_GEMM_SPLIT")) : 0;
int set_gemm_split(int v) {
    if (v != 0 && v != 2 && v != 3 && v != 12 && v != 13) return -2;
    g_gemm_split = v;
    return 0;
}
int get_gemm_split() { return g_gemm_split; }
int set_plane_tile(int v) {
    if (v != 11 && v != 21 && v != 12 && v != 22 && v != 111 && v != 121 && v != 112 && v != 1142 && v != 1124 &&
        v != 1144 && v != 1242 && v != 2124 && v != 3011 && v != 3142)
        return -2;
    g_plane_tile = v;
    return 0;
}
int get_plane_tile() { return g_plane_tile; }
int set_plane_group(int v) {
    if (v < 0 || v > 64) return -2;
    g_plane_group = v;
    return 0;
}
int get_plane_group() { return g_plane_group; }

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
    // The plane GEMM splits A (m x k) and B (k x n) per launch: with a narrow C the split passes
    // outweigh the faster products (measured at m = 2496, k = 256..2048: 3 terms lose to the
    // FP32 matrix cores below ~160 columns, 2 terms below ~96), so narrow launches keep FP32.
    const int narrow = (m < n ? m : n);
    if (sizeof(T) == 4 && g_gemm_split >= 12 && narrow >= (g_gemm_split == 13 ? 160 : 96)) {
        if (g_prof_on) {
            // the kernel reads A and B as bfloat16 planes (2 bytes x 2 components x terms)
            const double terms = g_gemm_split - 10;
            rec.bytes = (4.0 * terms * ((double)m * k + (double)k * n) + 16.0 * m * n) * batch;
            rec.plane = 1;
            g_prof_pending = rec;
        }
        const bool ok = g_gemm_split == 12
                            ? gemm_planes<2>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s)
                            : gemm_planes<3>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s);
        if (ok) {
            if (g_prof_on) {
                hipEventRecord(rec.b, s);
                g_prof.push_back(rec);
            }
            return;
        }
        if (g_prof_on) hipEventRecord(rec.a, s);    // fell through: time the kernel below
    }
    if (vec && g_gemm_split >= 2 && g_gemm_split < 12) {
        dim3 g2((m + ST - 1) / ST, (n + ST - 1) / ST, batch);
        if (g_gemm_split == 2)
            hipLaunchKernelGGL(k_gemm_sub_c64s<2>, g2, dim3(256), 0, s, m, n, k, (const cx<float>*)A, lda, sA,
                               (const cx<float>*)B, ldb, sB, (cx<float>*)C, ldc, sC);
        else
            hipLaunchKernelGGL(k_gemm_sub_c64s<3>, g2, dim3(256), 0, s, m, n, k, (const cx<float>*)A, lda, sA,
                               (const cx<float>*)B, ldb, sB, (cx<float>*)C, ldc, sC);
    } else if (vec)
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
