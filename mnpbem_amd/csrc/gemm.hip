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
#include "ctx.hpp"

#include <stdint.h>

#include <mutex>
#include <vector>

namespace mnp {

// Optional per-launch timing of the GEMM (HIP events on the launch stream), for roofline
// reporting. Off by default; see mnp_prof_enable / mnp_prof_collect in capi.hip. Records live in
// the calling thread's handle.
void prof_enable(bool on) { ctx().prof_on = on; }

// Sum of GEMM kernel time (ms), algorithmic flops and bytes since the last call, over the
// launches of the split-BF16 plane kernel when there were any, else over all; waits for the
// events.
void prof_collect(double* ms, double* flops, double* launches, double* bytes) {
    Ctx& c = ctx();
    double t = 0, f = 0, by = 0, nl = 0;
    bool any_plane = false;
    for (auto& r : c.prof) any_plane |= r.plane != 0;
    for (auto& r : c.prof) {
        if (!any_plane || r.plane) {
            hipEventSynchronize(r.b);
            float e = 0;
            hipEventElapsedTime(&e, r.a, r.b);
            t += e;
            f += r.flops;
            by += r.bytes;
            nl += 1;
        }
        c.pool.push_back(r.a);
        c.pool.push_back(r.b);
    }
    hipDeviceSynchronize();
    *ms = t;
    *flops = f;
    *launches = nl;
    *bytes = by;
    c.prof.clear();
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
// real product keeps the term pairs of order < NS (NS = 2: x0y0 + x0y1 + x1y0, about 16
// significant bits; NS = 3: six term products, about 24 bits, the float's own precision),
// accumulated in FP32 by v_mfma_f32_32x32x16_bf16.
//
// The complex product takes four real products per term pair, C += (-A) B with (re re, -im im)
// and (im re, re im). (The three-product Gauss form, P1 = ar br, P2 = ai bi, P3 = (ar + ai)
// (br + bi), needs a third operand plane per term; measured on this kernel it ran 7-15 %
// slower than four products, because the kernel is bound by moving operands through LDS, not
// by the matrix cores.)
//
// Plane layout (per batch entry): component c (0 re, 1 im) of term q is plane qc = 2 q + c.
// Element (qc, row r, k) (r: row of A or column of B) lies at
//   (((qc KS + k / 16) 2 + (k / 8) % 2) R + r) 8 + k % 8,   KS = ceil(K / 16), R = rows padded
// to 128: one k slice of a plane is two halves of 8 k, and in a half consecutive rows are 16
// bytes apart. A 16-byte chunk (8 k of one row) is one MFMA operand fragment of one lane: lane
// l of a wave takes row l % 32 and half l / 32, so a wave's fragment is two contiguous 512-byte
// runs. Padding rows and k are zero. A is stored negated (C -= A B).
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f16acc __attribute__((ext_vector_type(16)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
constexpr int ST = 128, SK = 16, SKP = 24;    // row padding, K slice, LDS row length (4M kernel, bf16)

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
    float v[2][16];
    const int k0 = 16 * s;
    if (TRANS && vec && r < rows && k0 + 16 <= K) {
        const float4* p = (const float4*)(X + k0 + (long)r * ldx);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float4 w = p[e];
            v[0][2 * e] = sgn * w.x;
            v[1][2 * e] = sgn * w.y;
            v[0][2 * e + 1] = sgn * w.z;
            v[1][2 * e + 1] = sgn * w.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int k = k0 + e;
            v[0][e] = 0.f;
            v[1][e] = 0.f;
            if (r < rows && k < K) {
                const cx<float> w = TRANS ? X[k + (long)r * ldx] : X[r + (long)k * ldx];
                v[0][e] = sgn * w.re;
                v[1][e] = sgn * w.im;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float* x = v[c];
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

// Workgroup of 4 waves (2 x 2), a 64 x 64 tile of C, each wave one 32 x 32 MFMA tile
// with real and imaginary accumulators. Each k slice is staged through LDS in rows of 16 k (+8
// padding: 48-byte rows, conflict-free 16-byte reads); the next slice is loaded into registers
// while the current one is multiplied. +Ai (for the -im im product) is one sign flip per
// fragment in registers.
template <int NS, bool F16 = false>
__global__ __launch_bounds__(256) void k_gemm_sub_c64p(int m, int n, int K, const __bf16* __restrict__ PA, int RA,
                                                       long sPA, const __bf16* __restrict__ PB, int RB, long sPB,
                                                       cx<float>* __restrict__ C, long ldc, long sC, int gm,
                                                       const float* __restrict__ scA = nullptr,
                                                       const float* __restrict__ scB = nullptr) {
    constexpr int TM = 64, TN = 64;
    constexpr int CA = 4 * NS * TM, CB = 4 * NS * TN;          // 16-byte chunks per k slice
    constexpr int UA = CA / 256, UB = CB / 256;
    __shared__ __attribute__((aligned(16))) __bf16 As[NS][2][TM][SKP];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NS][2][TN][SKP];
    int tx, ty, bz;
    xcd_tile(gm, tx, ty, bz);
    PA += (long)bz * sPA;
    PB += (long)bz * sPB;
    C += (long)bz * sC;
    const int KS = (K + 15) / 16;
    const int row0 = tx * TM, col0 = ty * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    f16acc cre, cim;
#pragma unroll
    for (int r = 0; r < 16; ++r) { cre[r] = 0; cim[r] = 0; }
    // chunk e = t + 256 u of the A (B) tile: plane e / (2 TM), half (e % (2 TM)) / TM, row e % TM
    u4v ra[UA], rb[UB];
    auto load = [&](int s) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int e = t + 256 * u, qc = e / (2 * TM), w = e % (2 * TM);
            ra[u] = *(const u4v*)(PA + ((((long)qc * KS + s) * 2 + w / TM) * RA + row0 + w % TM) * 8);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = t + 256 * u, qc = e / (2 * TN), w = e % (2 * TN);
            rb[u] = *(const u4v*)(PB + ((((long)qc * KS + s) * 2 + w / TN) * RB + col0 + w % TN) * 8);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int e = t + 256 * u, qc = e / (2 * TM), w = e % (2 * TM);
            *(u4v*)&As[qc >> 1][qc & 1][w % TM][8 * (w / TM)] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = t + 256 * u, qc = e / (2 * TN), w = e % (2 * TN);
            *(u4v*)&Bs[qc >> 1][qc & 1][w % TN][8 * (w / TN)] = rb[u];
        }
    };
    load(0);
    for (int s = 0; s < KS; ++s) {
        store();
        __syncthreads();
        if (s + 1 < KS) load(s + 1);
        const int fr = lane & 31, fk = 8 * (lane >> 5);
        bf8v pr[NS], pi[NS], qr[NS], qi[NS], qp[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            pr[q] = *(const bf8v*)&Bs[q][0][wn + fr][fk];
            pi[q] = *(const bf8v*)&Bs[q][1][wn + fr][fk];
            qr[q] = *(const bf8v*)&As[q][0][wm + fr][fk];
            qi[q] = *(const bf8v*)&As[q][1][wm + fr][fk];
            u4v w = __builtin_bit_cast(u4v, qi[q]);
            w ^= 0x80008000u;
            qp[q] = __builtin_bit_cast(bf8v, w);
        }
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
            for (int q = 0; q < NS - p; ++q) {
                if constexpr (F16) {
                    // the same bits as fp16 pairs (k_split_f16)
                    typedef _Float16 h8 __attribute__((ext_vector_type(8)));
                    const h8 br = __builtin_bit_cast(h8, pr[p]), bi = __builtin_bit_cast(h8, pi[p]);
                    const h8 ar = __builtin_bit_cast(h8, qr[q]), ai = __builtin_bit_cast(h8, qi[q]);
                    const h8 ap = __builtin_bit_cast(h8, qp[q]);
                    cre = __builtin_amdgcn_mfma_f32_32x32x16_f16(br, ar, cre, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi, ar, cim, 0, 0, 0);
                    cre = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi, ap, cre, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_32x32x16_f16(br, ai, cim, 0, 0, 0);
                } else {
                    cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[p], qr[q], cre, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[p], qr[q], cim, 0, 0, 0);
                    cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pi[p], qp[q], cre, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr[p], qi[q], cim, 0, 0, 0);
                }
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
            if constexpr (F16) {
                // undo the row and column scales (powers of two: exact)
                const float u = 1.f / (scA[(long)bz * RA + gi] * scB[(long)bz * RB + gj]);
                v.re += cre[r] * u;
                v.im += cim[r] * u;
            } else {
                v.re += cre[r];
                v.im += cim[r];
            }
            *c = v;
        }
    }
}

// Workgroup of 4 waves (2 x 2) on a (64 MT) x (64 NT) tile of C, each wave MT x NT MFMA tiles of
// 32 x 32 (real and imaginary accumulators each): every A fragment read from LDS feeds NT MFMA
// tiles and every B fragment MT, and a k slice of operands moved per workgroup feeds
// MT NT times the products of the 64 x 64 tile of k_gemm_sub_c64p. Operands go from the planes
// to LDS without registers (global_load_lds_dwordx4: a wave instruction copies 64 rows x 16
// bytes of one half slice of one plane, exactly the plane's own layout, so LDS holds chunk
// (plane, half, 64-row block) at 1 KB each); two LDS stages, the next slice's copies in flight
// while the current one is multiplied, one barrier per slice. Fragment reads: lane l takes row
// l % 32 of its MFMA tile from half l / 32, two contiguous 512-byte runs per wave
// (conflict-free). Each accumulator receives the term products of k_gemm_sub_c64p in the same
// order: results are identical. The MFMA tiles take turns, so that an accumulator is written
// again only MT NT MFMAs later.
template <int NS, int MT, int NT, bool F16 = false>
__global__ __launch_bounds__(256) void k_gemm_sub_c64q(int m, int n, int K, const __bf16* __restrict__ PA, int RA,
                                                       long sPA, const __bf16* __restrict__ PB, int RB, long sPB,
                                                       cx<float>* __restrict__ C, long ldc, long sC, int gm,
                                                       const float* __restrict__ scA = nullptr,
                                                       const float* __restrict__ scB = nullptr) {
    constexpr int TM = 64 * MT, TN = 64 * NT;
    constexpr int CA = 2 * NS * 2 * MT, CB = 2 * NS * 2 * NT;    // 1 KB chunks per slice: plane x half x block
    constexpr int CH = CA + CB;
    static_assert(CH % 4 == 0, "chunks split evenly over the four waves");
    __shared__ __attribute__((aligned(1024))) unsigned char lds[2][CH * 1024];   // [stage][A chunks, B chunks]
    int tx, ty, bz;
    xcd_tile(gm, tx, ty, bz);
    PA += (long)bz * sPA;
    PB += (long)bz * sPB;
    C += (long)bz * sC;
    const int KS = (K + 15) / 16;
    const int row0 = tx * TM, col0 = ty * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 32 * MT, wn = (wave >> 1) * 32 * NT;
    f16acc cre[MT][NT], cim[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) { cre[a][b][r] = 0; cim[a][b][r] = 0; }
    // chunk c < CA of A: plane c / (2 MT), half (c / MT) % 2, row block c % MT; then B likewise;
    // wave w copies chunks w, w + 4, ...
    auto issue = [&](int s, int st) {
#pragma unroll
        for (int i = 0; i < CH / 4; ++i) {
            const int ch = wave + 4 * i;
            const bool isb = ch >= CA;
            const int c = isb ? ch - CA : ch;
            const int nbk = isb ? NT : MT;
            const int qc = c / (2 * nbk), h = (c / nbk) & 1, rb = c % nbk;
            const __bf16* src = isb ? PB + ((((long)qc * KS + s) * 2 + h) * RB + col0 + 64 * rb) * 8
                                    : PA + ((((long)qc * KS + s) * 2 + h) * RA + row0 + 64 * rb) * 8;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + 8 * lane),
                                             (__attribute__((address_space(3))) void*)&lds[st][ch * 1024], 16, 0, 0);
        }
    };
    // fragment of MFMA tile row `tr` (a multiple of 32 below 64 MT or 64 NT), plane qc
    auto fragA = [&](int st, int qc, int tr) {
        const int r = tr + (lane & 31), h = lane >> 5;
        return *(const bf8v*)&lds[st][(((qc * 2 + h) * MT + (r >> 6)) * 64 + (r & 63)) * 16];
    };
    auto fragB = [&](int st, int qc, int tr) {
        const int r = tr + (lane & 31), h = lane >> 5;
        return *(const bf8v*)&lds[st][((CA + (qc * 2 + h) * NT + (r >> 6)) * 64 + (r & 63)) * 16];
    };
    issue(0, 0);
    for (int s = 0; s < KS; ++s) {
        const int st = s & 1;
        __builtin_amdgcn_s_waitcnt(0);     // this wave's copies of slice s have landed in LDS
        __syncthreads();                   // everyone's; and stage st ^ 1 is no longer read
        if (s + 1 < KS) issue(s + 1, st ^ 1);
        bf8v pr[NT][NS], pi[NT][NS], qr[MT][NS], qi[MT][NS], qp[MT][NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                pr[u][q] = fragB(st, 2 * q, wn + 32 * u);
                pi[u][q] = fragB(st, 2 * q + 1, wn + 32 * u);
            }
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                qr[u][q] = fragA(st, 2 * q, wm + 32 * u);
                qi[u][q] = fragA(st, 2 * q + 1, wm + 32 * u);
                u4v w = __builtin_bit_cast(u4v, qi[u][q]);
                w ^= 0x80008000u;
                qp[u][q] = __builtin_bit_cast(bf8v, w);
            }
        }
        // one real product of two operand planes into an accumulator: bfloat16 terms, or the
        // same bits read as fp16 pairs (k_split_f16)
        auto mfma = [](bf8v x, bf8v y, f16acc acc) {
            if constexpr (F16) {
                typedef _Float16 h8 __attribute__((ext_vector_type(8)));
                return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, x), __builtin_bit_cast(h8, y),
                                                              acc, 0, 0, 0);
            } else {
                return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc, 0, 0, 0);
            }
        };
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
            for (int q = 0; q < NS - p; ++q) {
#pragma unroll
                for (int a = 0; a < MT; ++a)
#pragma unroll
                    for (int b = 0; b < NT; ++b)
                        cre[a][b] = mfma(pr[b][p], qr[a][q], cre[a][b]);
#pragma unroll
                for (int a = 0; a < MT; ++a)
#pragma unroll
                    for (int b = 0; b < NT; ++b)
                        cim[a][b] = mfma(pi[b][p], qr[a][q], cim[a][b]);
#pragma unroll
                for (int a = 0; a < MT; ++a)
#pragma unroll
                    for (int b = 0; b < NT; ++b)
                        cre[a][b] = mfma(pi[b][p], qp[a][q], cre[a][b]);
#pragma unroll
                for (int a = 0; a < MT; ++a)
#pragma unroll
                    for (int b = 0; b < NT; ++b)
                        cim[a][b] = mfma(pr[b][p], qi[a][q], cim[a][b]);
            }
    }
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = row0 + wm + 32 * a + (lane & 31);
                const int gj = col0 + wn + 32 * b + Mfma<float>::drow(lane, r);
                if (gi < m && gj < n) {
                    cx<float>* c = C + gi + (long)gj * ldc;
                    cx<float> v = *c;
                    if constexpr (F16) {
                        // undo the row and column scales (powers of two: exact)
                        const float u = 1.f / (scA[(long)bz * RA + gi] * scB[(long)bz * RB + gj]);
                        v.re += cre[a][b][r] * u;
                        v.im += cim[a][b][r] * u;
                    } else {
                        v.re += cre[a][b][r];
                        v.im += cim[a][b][r];
                    }
                    *c = v;
                }
            }
}

// ---------------------------------------------------------------------------------------------
// Complex-float GEMM on the FP16 matrix cores with operands as scaled pairs of half floats
// (mode 16; a two-term splitting in the manner of Ootomo and Yokota, "Recovering single
// precision accuracy from Tensor Cores while surpassing the FP32 theoretical peak performance",
// IJHPCA 2022, with fp16 terms). Each row i of A and column j of B gets a power-of-two scale s
// (its largest |re|, |im| brought into [2^14, 2^15)), and every scaled value v = s x is held as
// hi = fp16(v) and lo = fp16(v - hi): hi carries 11 significant bits and lo the next 11, so the
// pair keeps about 22 bits; with the scale, lo is a normal fp16 number for every entry down to
// 2^-17 of its row's (column's) largest, and below that its absolute error stays under 2^-39 of
// that largest entry. A real product keeps hi hi + hi lo + lo hi (the dropped lo lo is 2^-22 of
// it): 3 MFMAs, as two bfloat16 terms and half the 6 of three, with FP32 accumulation; the
// epilogue divides by the two scales (powers of two, exact). The planes have the layout of the
// two-term bfloat16 planes (term 0 = hi, term 1 = lo) and are read by k_gemm_sub_c64p<2, true>.

// Scales of the rows of A (m x K, column-major): one thread per row. sc[r] = 2^(15 - e) with
// max_k max(|re|, |im|) in [2^(e-1), 2^e); 1 for a zero row and for the padding rows.
__global__ __launch_bounds__(256) void k_scale_rows(int rows, int R, int K, const cx<float>* __restrict__ X, long ldx,
                                                    long sX, float* __restrict__ sc) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    X += (long)blockIdx.y * sX;
    sc += (long)blockIdx.y * R;
    if (r >= R) return;
    float m = 0.f;
    if (r < rows)
        for (int k = 0; k < K; ++k) {
            const cx<float> v = X[r + (long)k * ldx];
            m = fmaxf(m, fmaxf(fabsf(v.re), fabsf(v.im)));
        }
    int e = 0;
    frexpf(m, &e);
    e = e < -100 ? -100 : e;
    sc[r] = m > 0.f ? ldexpf(1.f, 15 - e) : 1.f;
}

// Scales of the columns of B (K x n, column-major, a column contiguous): one wave per column,
// lanes along K, wave maximum by butterfly shuffles.
__global__ __launch_bounds__(256) void k_scale_cols(int cols, int R, int K, const cx<float>* __restrict__ X, long ldx,
                                                    long sX, float* __restrict__ sc) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    X += (long)blockIdx.y * sX;
    sc += (long)blockIdx.y * R;
    if (c >= R) return;
    float m = 0.f;
    if (c < cols)
        for (int k = lane; k < K; k += 64) {
            const cx<float> v = X[k + (long)c * ldx];
            m = fmaxf(m, fmaxf(fabsf(v.re), fabsf(v.im)));
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) {
        int e = 0;
        frexpf(m, &e);
        e = e < -100 ? -100 : e;
        sc[c] = m > 0.f ? ldexpf(1.f, 15 - e) : 1.f;
    }
}

typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// Scaled fp16 pair planes: as k_split_planes<2, TRANS> with v = sgn sc[r] x, term 0 = fp16(v),
// term 1 = fp16(v - term 0).
template <bool TRANS>
__global__ __launch_bounds__(256) void k_split_f16(int rows, int K, const cx<float>* __restrict__ X, long ldx, long sX,
                                                   _Float16* __restrict__ P, int R, long sP, float sgn,
                                                   const float* __restrict__ sc, int vec) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    const int KS = gridDim.y;
    X += (long)blockIdx.z * sX;
    P += (long)blockIdx.z * sP;
    sc += (long)blockIdx.z * R;
    if (r >= R) return;
    float v[2][16];
    const int k0 = 16 * s;
    if (TRANS && vec && r < rows && k0 + 16 <= K) {
        const float4* p = (const float4*)(X + k0 + (long)r * ldx);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float4 w = p[e];
            v[0][2 * e] = w.x;
            v[1][2 * e] = w.y;
            v[0][2 * e + 1] = w.z;
            v[1][2 * e + 1] = w.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int k = k0 + e;
            v[0][e] = 0.f;
            v[1][e] = 0.f;
            if (r < rows && k < K) {
                const cx<float> w = TRANS ? X[k + (long)r * ldx] : X[r + (long)k * ldx];
                v[0][e] = w.re;
                v[1][e] = w.im;
            }
        }
    }
    const float f = sgn * sc[r];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            h8v hi, lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float x = f * v[c][8 * hf + e];
                hi[e] = (_Float16)x;
                lo[e] = (_Float16)(x - (float)hi[e]);
            }
            *(h8v*)(P + (((long)(0 * 2 + c) * KS + s) * 2 + hf) * R * 8 + (long)r * 8) = hi;
            *(h8v*)(P + (((long)(1 * 2 + c) * KS + s) * 2 + hf) * R * 8 + (long)r * 8) = lo;
        }
    }
}

// Scales and fp16-pair planes in one pass over the operand, for K <= 64 HPT: a workgroup of 256
// lanes takes 32 plane rows (rows of A; TRANS: columns of B, K x n column-major) and all of K.
// Lane t holds row t % 32 and the half slices (8 consecutive k) h = t / 32 + 8 j, j < HPT, in
// registers; the row's largest |re|, |im| is reduced over its 8 lanes in LDS; then the scale is
// written as k_scale_rows / k_scale_cols write it and the planes as k_split_f16 writes them
// (bit-identical), so the operand is read once instead of twice.
template <bool TRANS, int HPT>
__global__ __launch_bounds__(256) void k_scale_split_f16(int rows, int K, const cx<float>* __restrict__ X, long ldx,
                                                         long sX, _Float16* __restrict__ P, int R, long sP, float sgn,
                                                         float* __restrict__ sc, int vec) {
    __shared__ float red[8][32];
    const int t = threadIdx.x, rl = t & 31, g = t >> 5;
    const int r = blockIdx.x * 32 + rl;          // < R: the grid covers R / 32 blocks, R a multiple of 128
    X += (long)blockIdx.y * sX;
    P += (long)blockIdx.y * sP;
    sc += (long)blockIdx.y * R;
    const int KS = (K + 15) / 16, NH = 2 * KS;
    float v[HPT][2][8];
    float mx = 0.f;
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
        const int h = g + 8 * j, k0 = 8 * h;
        if (TRANS && vec && r < rows && k0 + 8 <= K) {
            const float4* p = (const float4*)(X + k0 + (long)r * ldx);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float4 w = p[e];
                v[j][0][2 * e] = w.x;
                v[j][1][2 * e] = w.y;
                v[j][0][2 * e + 1] = w.z;
                v[j][1][2 * e + 1] = w.w;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + e;
                v[j][0][e] = 0.f;
                v[j][1][e] = 0.f;
                if (h < NH && r < rows && k < K) {
                    const cx<float> w = TRANS ? X[k + (long)r * ldx] : X[r + (long)k * ldx];
                    v[j][0][e] = w.re;
                    v[j][1][e] = w.im;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) mx = fmaxf(mx, fmaxf(fabsf(v[j][0][e]), fabsf(v[j][1][e])));
    }
    red[g][rl] = mx;
    __syncthreads();
    float m = red[0][rl];
#pragma unroll
    for (int i = 1; i < 8; ++i) m = fmaxf(m, red[i][rl]);
    int ex = 0;
    frexpf(m, &ex);
    ex = ex < -100 ? -100 : ex;
    const float scale = m > 0.f ? ldexpf(1.f, 15 - ex) : 1.f;
    if (g == 0) sc[r] = scale;
    const float f = sgn * scale;
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
        const int h = g + 8 * j;
        if (h >= NH) break;
        const int sl = h >> 1, hf = h & 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            h8v hi, lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float x = f * v[j][c][e];
                hi[e] = (_Float16)x;
                lo[e] = (_Float16)(x - (float)hi[e]);
            }
            *(h8v*)(P + (((long)(0 * 2 + c) * KS + sl) * 2 + hf) * R * 8 + (long)r * 8) = hi;
            *(h8v*)(P + (((long)(1 * 2 + c) * KS + sl) * 2 + hf) * R * 8 + (long)r * 8) = lo;
        }
    }
}

template <bool TRANS>
static bool scale_split_f16(int rows, int R, int K, const cx<float>* X, long ldx, long sX, _Float16* P, long sP,
                            float sgn, float* sc, int vec, int batch, hipStream_t s) {
    const int NH = 2 * ((K + 15) / 16);
    const dim3 g(R / 32, batch);
    if (NH <= 8)
        hipLaunchKernelGGL((k_scale_split_f16<TRANS, 1>), g, dim3(256), 0, s, rows, K, X, ldx, sX, P, R, sP, sgn, sc, vec);
    else if (NH <= 16)
        hipLaunchKernelGGL((k_scale_split_f16<TRANS, 2>), g, dim3(256), 0, s, rows, K, X, ldx, sX, P, R, sP, sgn, sc, vec);
    else if (NH <= 32)
        hipLaunchKernelGGL((k_scale_split_f16<TRANS, 4>), g, dim3(256), 0, s, rows, K, X, ldx, sX, P, R, sP, sgn, sc, vec);
    else if (NH <= 64)
        hipLaunchKernelGGL((k_scale_split_f16<TRANS, 8>), g, dim3(256), 0, s, rows, K, X, ldx, sX, P, R, sP, sgn, sc, vec);
    else
        return false;
    return true;
}

// Plane workspace of the calling thread's handle for stream s, grown on demand.
static void* plane_ws(hipStream_t s, size_t bytes) {
    Ctx& c = ctx();
    for (auto& w : c.plane_ws)
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
    c.plane_ws.push_back(PlaneWs{s, p, bytes});
    return p;
}

// Bytes of the plane workspace of one launch.
static size_t plane_bytes(int NS, int m, int n, int k, int batch) {
    const long KS = (k + 15) / 16;
    const long RA = (m + ST - 1) / ST * ST, RB = (n + ST - 1) / ST * ST;
    return (size_t)2 * NS * KS * 16 * (RA + RB) * batch * sizeof(__bf16);
}

// Bytes of the workspace of one fp16-pair launch: the two-term planes and the row / column scales.
static size_t f16_bytes(int m, int n, int k, int batch) {
    const long RA = (m + ST - 1) / ST * ST, RB = (n + ST - 1) / ST * ST;
    return plane_bytes(2, m, n, k, batch) + (size_t)(RA + RB) * batch * sizeof(float);
}

// Scales, fp16-pair planes of A (negated) and B at P0 (workspace of f16_bytes; null: the
// handle's own), and the product. False when the workspace cannot be allocated.
static bool gemm_f16(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB, void* C,
                     long ldc, long sC, int batch, hipStream_t s, void* P0, ProfRec* rec) {
    const int KS = (k + 15) / 16;
    const int RA = (m + ST - 1) / ST * ST, RB = (n + ST - 1) / ST * ST;
    const long sPA = 4L * KS * RA * 16, sPB = 4L * KS * RB * 16;
    _Float16* P = (_Float16*)(P0 ? P0 : plane_ws(s, f16_bytes(m, n, k, batch)));
    if (!P) return false;
    _Float16* PA = P;
    _Float16* PB = P + sPA * batch;
    float* scA = (float*)(PB + sPB * batch);
    float* scB = scA + (long)RA * batch;
    const int vb = ((ldb | sB) & 1) == 0 && ((uintptr_t)B & 15) == 0;
    // one pass per operand while its K fits the registers (K <= 512), else scales then planes
    if (!(ctx().f16_fused && scale_split_f16<false>(m, RA, k, (const cx<float>*)A, lda, sA, PA, sPA, -1.0f, scA, 0,
                                                     batch, s))) {
        hipLaunchKernelGGL(k_scale_rows, dim3((RA + 255) / 256, batch), dim3(256), 0, s, m, RA, k,
                           (const cx<float>*)A, lda, sA, scA);
        hipLaunchKernelGGL((k_split_f16<false>), dim3((RA + 255) / 256, KS, batch), dim3(256), 0, s, m, k,
                           (const cx<float>*)A, lda, sA, PA, RA, sPA, -1.0f, scA, 0);
    }
    if (!(ctx().f16_fused && scale_split_f16<true>(n, RB, k, (const cx<float>*)B, ldb, sB, PB, sPB, 1.0f, scB, vb,
                                                    batch, s))) {
        hipLaunchKernelGGL(k_scale_cols, dim3((RB + 3) / 4, batch), dim3(256), 0, s, n, RB, k,
                           (const cx<float>*)B, ldb, sB, scB);
        hipLaunchKernelGGL((k_split_f16<true>), dim3((RB + 255) / 256, KS, batch), dim3(256), 0, s, n, k,
                           (const cx<float>*)B, ldb, sB, PB, RB, sPB, 1.0f, scB, vb);
    }
    if (rec) hipEventRecord(rec->a, s);
    const __bf16 *pa = (const __bf16*)PA, *pb = (const __bf16*)PB;
    const int gm = ctx().plane_group;
    // the plane tile knob (11: 64 x 64 through registers; 111, 21, 12, 22: direct LDS copies)
    switch (ctx().plane_tile) {
        case 111:
            hipLaunchKernelGGL((k_gemm_sub_c64q<2, 1, 1, true>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256),
                               0, s, m, n, k, pa, RA, sPA, pb, RB, sPB, (cx<float>*)C, ldc, sC, gm, scA, scB);
            break;
        case 21:
            hipLaunchKernelGGL((k_gemm_sub_c64q<2, 2, 1, true>), dim3((m + 127) / 128, (n + 63) / 64, batch),
                               dim3(256), 0, s, m, n, k, pa, RA, sPA, pb, RB, sPB, (cx<float>*)C, ldc, sC, gm, scA, scB);
            break;
        case 12:
            hipLaunchKernelGGL((k_gemm_sub_c64q<2, 1, 2, true>), dim3((m + 63) / 64, (n + 127) / 128, batch),
                               dim3(256), 0, s, m, n, k, pa, RA, sPA, pb, RB, sPB, (cx<float>*)C, ldc, sC, gm, scA, scB);
            break;
        case 22:
            hipLaunchKernelGGL((k_gemm_sub_c64q<2, 2, 2, true>), dim3((m + 127) / 128, (n + 127) / 128, batch),
                               dim3(256), 0, s, m, n, k, pa, RA, sPA, pb, RB, sPB, (cx<float>*)C, ldc, sC, gm, scA, scB);
            break;
        default:
            hipLaunchKernelGGL((k_gemm_sub_c64p<2, true>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, s,
                               m, n, k, pa, RA, sPA, pb, RB, sPB, (cx<float>*)C, ldc, sC, gm, scA, scB);
    }
    return true;
}

// Split A and B into planes at P0 (workspace of plane_bytes; null: the handle's own) and run
// the product. False when the workspace cannot be allocated.
template <int NS>
static bool gemm_planes(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB,
                        void* C, long ldc, long sC, int batch, hipStream_t s, void* P0, ProfRec* rec) {
    const int KS = (k + 15) / 16;
    const int RA = (m + ST - 1) / ST * ST, RB = (n + ST - 1) / ST * ST;
    const long sPA = 2L * NS * KS * RA * 16, sPB = 2L * NS * KS * RB * 16;
    __bf16* P = (__bf16*)(P0 ? P0 : plane_ws(s, plane_bytes(NS, m, n, k, batch)));
    if (!P) return false;
    __bf16* PA = P;
    __bf16* PB = P + sPA * batch;
    const int vb = ((ldb | sB) & 1) == 0 && ((uintptr_t)B & 15) == 0;   // 16-byte loads along a column of B
    hipLaunchKernelGGL((k_split_planes<NS, false>), dim3(RA / 256 + (RA % 256 != 0), KS, batch), dim3(256), 0, s, m,
                       k, (const cx<float>*)A, lda, sA, PA, RA, sPA, -1.0f, 0);
    hipLaunchKernelGGL((k_split_planes<NS, true>), dim3(RB / 256 + (RB % 256 != 0), KS, batch), dim3(256), 0, s, n,
                       k, (const cx<float>*)B, ldb, sB, PB, RB, sPB, 1.0f, vb);
    if (rec) hipEventRecord(rec->a, s);    // the timed region is the GEMM kernel alone
    const int tile = ctx().plane_tile;
    const int gx = tile == 22 || tile == 21 ? (m + 127) / 128 : (m + 63) / 64;
    const int gy = tile == 22 || tile == 12 ? (n + 127) / 128 : (n + 63) / 64;
    if (tile == 22)
        hipLaunchKernelGGL((k_gemm_sub_c64q<NS, 2, 2>), dim3(gx, gy, batch), dim3(256), 0, s, m, n, k, PA, RA, sPA, PB,
                           RB, sPB, (cx<float>*)C, ldc, sC, ctx().plane_group);
    else if (tile == 21)
        hipLaunchKernelGGL((k_gemm_sub_c64q<NS, 2, 1>), dim3(gx, gy, batch), dim3(256), 0, s, m, n, k, PA, RA, sPA, PB,
                           RB, sPB, (cx<float>*)C, ldc, sC, ctx().plane_group);
    else if (tile == 12)
        hipLaunchKernelGGL((k_gemm_sub_c64q<NS, 1, 2>), dim3(gx, gy, batch), dim3(256), 0, s, m, n, k, PA, RA, sPA, PB,
                           RB, sPB, (cx<float>*)C, ldc, sC, ctx().plane_group);
    else if (tile == 111)
        hipLaunchKernelGGL((k_gemm_sub_c64q<NS, 1, 1>), dim3(gx, gy, batch), dim3(256), 0, s, m, n, k, PA, RA, sPA, PB,
                           RB, sPB, (cx<float>*)C, ldc, sC, ctx().plane_group);
    else
        hipLaunchKernelGGL((k_gemm_sub_c64p<NS>), dim3((m + 63) / 64, (n + 63) / 64, batch), dim3(256), 0, s, m, n, k,
                           PA, RA, sPA, PB, RB, sPB, (cx<float>*)C, ldc, sC, ctx().plane_group);
    return true;
}

// 0: FP32 matrix cores; 12 or 13: BF16 matrix cores, operands split once into planes of 2 or 3
// terms; 16: FP16 matrix cores, operands as scaled fp16 pairs.
int set_gemm_split(int v) {
    if (v != 0 && v != 12 && v != 13 && v != 16) return -2;
    ctx().gemm_split = v;
    return 0;
}
int get_gemm_split() { return ctx().gemm_split; }
// Plane kernel tile: 11 = 64 x 64 staged through registers (k_gemm_sub_c64p); with direct LDS
// copies (k_gemm_sub_c64q): 111 = 64 x 64, 21 = 128 x 64, 12 = 64 x 128, 22 = 128 x 128.
int set_plane_tile(int v) {
    if (v != 11 && v != 111 && v != 21 && v != 12 && v != 22) return -2;
    ctx().plane_tile = v;
    return 0;
}
int get_plane_tile() { return ctx().plane_tile; }
int set_plane_group(int v) {
    if (v < 0 || v > 64) return -2;
    ctx().plane_group = v;
    return 0;
}
int get_plane_group() { return ctx().plane_group; }

// Whether a complex-float launch of this shape runs on the split-BF16 plane kernels under the
// handle's mode: the split passes cost more than they save when C is narrow (measured at
// m = 2496, k = 256..2048: 3 terms lose to the FP32 matrix cores below ~160 columns, 2 terms
// below ~96).
static bool plane_launch(int mode, int m, int n) {
    const int narrow = (m < n ? m : n);
    if (mode == 16) return narrow >= 96;
    return (mode == 12 || mode == 13) && narrow >= (mode == 13 ? 160 : 96);
}

// C -= A * B for `batch` independent problems with complex-float GEMM mode `mode` (0, 12, 13).
// P: plane workspace of at least plane_bytes for this shape, or null for the handle's own;
// prof: record the launch when the handle's timing is on.
template <typename T>
static void gemm_sub_impl(int mode, bool prof, int m, int n, int k, const void* A, long lda, long sA, const void* B,
                          long ldb, long sB, void* C, long ldc, long sC, int batch, hipStream_t s, void* P) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return;
    Ctx& c = ctx();
    prof = prof && c.prof_on;
    dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN, batch);
    ProfRec rec{};
    if (prof) {
        rec.a = c.event();
        rec.b = c.event();
        rec.flops = 8.0 * m * (double)n * k * batch;
        // A and B read once, C read and written once
        rec.bytes = 2.0 * sizeof(T) * ((double)m * k + (double)k * n + 2.0 * m * n) * batch;
        hipEventRecord(rec.a, s);
    }
    // 16-byte loads need even leading dimensions and batch strides and 16-byte aligned bases
    const bool vec = sizeof(T) == 4 && ((lda | ldb | sA | sB) & 1) == 0 && ((uintptr_t)A & 15) == 0 &&
                     ((uintptr_t)B & 15) == 0;
    if (sizeof(T) == 4 && plane_launch(mode, m, n)) {
        const int terms = mode == 16 ? 2 : mode % 10;
        if (prof) {
            // the kernel reads A and B as 16-bit planes (2 bytes x 2 components x terms)
            rec.bytes = (4.0 * terms * ((double)m * k + (double)k * n) + 16.0 * m * n) * batch;
            rec.plane = 1;
        }
        ProfRec* rp = prof ? &rec : nullptr;
        const bool ok =
            mode == 16 ? gemm_f16(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, P, rp)
            : terms == 2 ? gemm_planes<2>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, P, rp)
                         : gemm_planes<3>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, P, rp);
        if (ok) {
            if (prof) {
                hipEventRecord(rec.b, s);
                c.prof.push_back(rec);
            }
            return;
        }
        if (prof) hipEventRecord(rec.a, s);    // fell through: time the kernel below
    }
    if (vec)
        hipLaunchKernelGGL(k_gemm_sub_c64v, grid, dim3(256), 0, s, m, n, k, (const cx<float>*)A, lda, sA,
                           (const cx<float>*)B, ldb, sB, (cx<float>*)C, ldc, sC);
    else
        hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, s, m, n, k, (const cx<T>*)A, lda, sA, (const cx<T>*)B,
                           ldb, sB, (cx<T>*)C, ldc, sC);
    if (prof) {
        hipEventRecord(rec.b, s);
        c.prof.push_back(rec);
    }
}

template <typename T>
void gemm_sub(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB, void* C,
              long ldc, long sC, int batch, hipStream_t s) {
    gemm_sub_impl<T>(ctx().gemm_split, true, m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, nullptr);
}

template void gemm_sub<double>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                               int, hipStream_t);
template void gemm_sub<float>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                              int, hipStream_t);

// Grows the handle's plane workspace of stream s (and of its lookahead side stream) to what a
// complex-float launch of m x n x k for `batch` matrices needs under the handle's mode, so that
// later launches of that size or smaller allocate nothing (work captured into a graph). False
// when the allocation fails.
bool gemm_reserve(int m, int n, int k, int batch, hipStream_t s, hipStream_t side) {
    const int mode = ctx().gemm_split;
    if (!plane_launch(mode, m, n)) return true;
    const size_t bytes = mode == 16 ? f16_bytes(m, n, k, batch) : plane_bytes(mode % 10, m, n, k, batch);
    if (!plane_ws(s, bytes)) return false;
    return side ? plane_ws(side, bytes) != nullptr : true;
}

// ---------------------------------------------------------------------------------------------
// Entry without allocation: the operand planes of the split-BF16 GEMM live in a static device
// array. In host code `scratch` names a host-side stand-in; its device address is taken once per
// device with hipGetSymbolAddress and only that address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = 32u << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEV = 64;
static void* g_scratch_ptr[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (null on failure).
static void* scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

// C -= A B in complex float with operands split into `terms` bfloat16 terms (as gemm_sub under
// mode 10 + terms), planes in the static scratch. Returns 0, 2 when the planes of this shape
// exceed the scratch (nothing enqueued), 3 when the scratch address is unavailable. Calls that
// overlap on one device race on the scratch; every plane element a launch reads is written by
// its own split pass, so nothing needs clearing between calls.
int gemm_sub_noalloc(int terms, int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb,
                     long sB, void* C, long ldc, long sC, int batch, hipStream_t s) {
    const int mode = 10 + terms;
    const bool planes = m > 0 && n > 0 && k > 0 && batch > 0 && plane_launch(mode, m, n);
    void* P = nullptr;
    if (planes) {
        if (plane_bytes(terms, m, n, k, batch) > NOALLOC_SCRATCH_BYTES) return 2;
        P = scratch_ptr();
        if (!P) return 3;
    }
    gemm_sub_impl<float>(mode, false, m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, P);
    return 0;
}

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
