// Batched blocked LU with partial pivoting, permutations and triangular solves.
//
// Right-looking blocked getrf (column-major, interleaved complex):
//   for each panel of nb columns starting at j0:
//     k_panel      : one workgroup per matrix factors the (n-j0) x nb panel (pivot = max |re|+|im|,
//                    first index on ties, as LAPACK's izamax), swapping rows inside the panel;
//     k_laswp      : applies the panel's row interchanges to the columns left and right of it;
//     k_trsm_lln   : U12 = L11^{-1} A12 (unit lower nb x nb block held in LDS);
//     gemm_sub     : A22 -= L21 * U12 on the matrix cores.
// Pivots are 0-based global row indices. `info` gets 1 + the first column whose pivot is zero.
//
// After the factorisation k_perm turns the interchange list into a row permutation pf
// (P b = b[pf]) and its inverse, so the solves gather instead of swapping sequentially.
#include "common.hpp"
#include "ctx.hpp"

#include <stdint.h>
#include <stdlib.h>

namespace mnp {

template <typename T>
void gemm_sub(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB, void* C,
              long ldc, long sC, int batch, hipStream_t s);

constexpr int PANEL_THREADS = 512;

template <typename T>
__global__ __launch_bounds__(PANEL_THREADS) void k_panel(int n, int j0, int nb, cx<T>* A, long lda, long sA,
                                                         int* ipiv, long sP, int* info) {
    __shared__ T red_v[PANEL_THREADS / 64];
    __shared__ int red_i[PANEL_THREADS / 64];
    __shared__ cx<T> urow[64];
    __shared__ int piv_s;
    const int b = blockIdx.x, t = threadIdx.x;
    A += (long)b * sA;
    ipiv += (long)b * sP;
    for (int c = 0; c < nb; ++c) {
        const int col = j0 + c;
        // pivot search in column col, rows col..n-1
        T best = -1;
        int bi = n;
        for (int r = col + t; r < n; r += PANEL_THREADS) {
            T v = cabs1(A[r + (long)col * lda]);
            if (v > best) { best = v; bi = r; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            T ov = __shfl_down(best, off, 64);
            int oi = __shfl_down(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if ((t & 63) == 0) { red_v[t >> 6] = best; red_i[t >> 6] = bi; }
        __syncthreads();
        if (t == 0) {
            T bv = red_v[0];
            int bx = red_i[0];
            for (int w = 1; w < PANEL_THREADS / 64; ++w)
                if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bx)) { bv = red_v[w]; bx = red_i[w]; }
            if (bx >= n) bx = col;
            piv_s = bx;
            ipiv[col] = bx;
        }
        __syncthreads();
        const int p = piv_s;
        // swap rows col and p within the panel, and stage the pivot row
        if (t < nb) {
            cx<T>* x = A + col + (long)(j0 + t) * lda;
            cx<T>* y = A + p + (long)(j0 + t) * lda;
            cx<T> u = *y;
            if (p != col) { *y = *x; *x = u; }
            urow[t] = u;
        }
        __syncthreads();
        cx<T> piv = urow[c];
        if (piv.re == 0 && piv.im == 0) {
            if (t == 0 && info[b] == 0) info[b] = col + 1;
            __syncthreads();
            continue;
        }
        cx<T> inv = crecip(piv);
        for (int r = col + 1 + t; r < n; r += PANEL_THREADS) {
            cx<T>* Ar = A + r + (long)j0 * lda;
            cx<T> l = Ar[(long)c * lda] * inv;
            Ar[(long)c * lda] = l;
            // trailing columns in chunks of 8: the chunk's loads are issued before its updates
            for (int c0 = c + 1; c0 < nb; c0 += 8) {
                cx<T> v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (c0 + u < nb) v[u] = Ar[(long)(c0 + u) * lda];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (c0 + u < nb) Ar[(long)(c0 + u) * lda] = fms(v[u], l, urow[c0 + u]);
            }
        }
        __syncthreads();
    }
}

// Register-resident panel: one workgroup of 1024 lanes per matrix holds the whole (n - j0) x w
// panel (n - j0 <= 1024 RPT rows, lane t owns rows j0 + t + 1024 q) in registers, CW columns at
// a time (left-looking over column chunks):
//   1. load the chunk's columns with the interchanges of the earlier chunks applied;
//   2. U part: forward substitution of the top rows with the unit-lower triangle of the earlier
//      chunks (in LDS), rows below: subtract L(rows, earlier columns) * U;
//   3. factor the chunk column by column: pivot = largest |re|+|im| on or below the diagonal,
//      smallest row on ties (as k_panel / izamax), the two rows trade places through LDS, rows
//      below are scaled and updated;
//   4. write the chunk back and apply its interchanges to the earlier chunks' columns.
// Interchanges outside the panel are applied by k_laswp afterwards, as for k_panel.
constexpr int RP_T = 1024, RP_RPT = 3, RP_MAXW = 32;
template <typename T>
struct RegChunk;
template <>
struct RegChunk<float> {
    static constexpr int CW = 8;
};
template <>
struct RegChunk<double> {
    static constexpr int CW = 4;
};

template <typename T, int CW>
__global__ __launch_bounds__(RP_T) void k_panel_reg(int n, int j0, int w, cx<T>* A, long lda, long sA, int* ipiv,
                                                    long sP, int* info) {
    constexpr int NW = RP_T / 64;
    __shared__ T red_v[NW];
    __shared__ int red_i[NW];
    __shared__ cx<T> urow[CW], crow[CW];
    __shared__ cx<T> ublk[RP_MAXW][CW];
    __shared__ int sw_p[RP_MAXW];          // pivot row chosen for panel column c (global index)
    __shared__ int aff[2 * CW], src_s[2 * CW], naff_s;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    A += (long)b * sA;
    ipiv += (long)b * sP;
    const int nc = (w + CW - 1) / CW;
    cx<T> x[RP_RPT][CW];
    for (int k = 0; k < nc; ++k) {
        const int c0 = k * CW;
        const int cw = min(CW, w - c0);
        // 1. load with the earlier interchanges applied (track each row back through the swaps)
#pragma unroll
        for (int q = 0; q < RP_RPT; ++q) {
            const int r = j0 + t + RP_T * q;
            int src = r;
            for (int c = c0 - 1; c >= 0; --c) {
                const int rc = j0 + c, pc = sw_p[c];
                if (src == rc) src = pc;
                else if (src == pc) src = rc;
            }
#pragma unroll
            for (int c = 0; c < CW; ++c)
                x[q][c] = (r < n && c < cw) ? A[src + (long)(j0 + c0 + c) * lda] : mk<T>(0, 0);
        }
        if (c0 > 0) {
            // 2a. U rows j0 .. j0+c0-1 of the chunk: forward substitution (lane c per column)
            if (t < c0)
#pragma unroll
                for (int c = 0; c < CW; ++c) ublk[t][c] = x[0][c];
            __syncthreads();
            if (t < cw) {
                for (int i = 1; i < c0; ++i) {
                    cx<T> v = ublk[i][t];
                    for (int qq = 0; qq < i; ++qq) v = fms(v, A[(j0 + i) + (long)(j0 + qq) * lda], ublk[qq][t]);
                    ublk[i][t] = v;
                }
            }
            __syncthreads();
            if (t < c0)
#pragma unroll
                for (int c = 0; c < CW; ++c) x[0][c] = ublk[t][c];
            // 2b. rows below: x -= L(row, 0..c0-1) * U
#pragma unroll
            for (int q = 0; q < RP_RPT; ++q) {
                const int r = j0 + t + RP_T * q;
                if (r < n && r >= j0 + c0) {
                    for (int s2 = 0; s2 < c0; ++s2) {
                        const cx<T> l = A[r + (long)(j0 + s2) * lda];
#pragma unroll
                        for (int c = 0; c < CW; ++c) x[q][c] = fms(x[q][c], l, ublk[s2][c]);
                    }
                }
            }
        }
        // 3. factor the chunk
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
            if (cc < cw) {
                const int col = j0 + c0 + cc;
                T best = (T)-1;
                int bi = n;
#pragma unroll
                for (int q = 0; q < RP_RPT; ++q) {
                    const int r = j0 + t + RP_T * q;
                    if (r < n && r >= col) {
                        const T v = cabs1(x[q][cc]);
                        if (v > best) { best = v; bi = r; }
                    }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    T ov = __shfl_xor(best, off, 64);
                    int oi = __shfl_xor(bi, off, 64);
                    if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
                }
                if (lane == 0) { red_v[wv] = best; red_i[wv] = bi; }
                __syncthreads();
                // every wave reduces the NW wave maxima itself (lanes 0 .. NW-1, same comparison),
                // so the pivot needs no second barrier to reach the other waves
                {
                    T bv = lane < NW ? red_v[lane] : (T)-1;
                    int bx = lane < NW ? red_i[lane] : n;
                    for (int off = NW / 2; off > 0; off >>= 1) {
                        T ov = __shfl_xor(bv, off, 64);
                        int oi = __shfl_xor(bx, off, 64);
                        if (ov > bv || (ov == bv && oi < bx)) { bv = ov; bx = oi; }
                    }
                    bi = __builtin_amdgcn_readfirstlane(bx);
                }
                if (bi >= n) bi = col;
                if (t == 0) {
                    sw_p[c0 + cc] = bi;
                    ipiv[col] = bi;
                }
                const int p = bi;
                // rows col and p trade places (chunk columns)
#pragma unroll
                for (int q = 0; q < RP_RPT; ++q) {
                    const int r = j0 + t + RP_T * q;
                    if (r == p)
#pragma unroll
                        for (int c = 0; c < CW; ++c) urow[c] = x[q][c];
                    if (r == col)
#pragma unroll
                        for (int c = 0; c < CW; ++c) crow[c] = x[q][c];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < RP_RPT; ++q) {
                    const int r = j0 + t + RP_T * q;
                    if (r == p && p != col)
#pragma unroll
                        for (int c = 0; c < CW; ++c) x[q][c] = crow[c];
                    if (r == col)
#pragma unroll
                        for (int c = 0; c < CW; ++c) x[q][c] = urow[c];
                }
                const cx<T> piv = urow[cc];
                const bool zero = piv.re == 0 && piv.im == 0;
                if (zero && t == 0 && info[b] == 0) info[b] = col + 1;
                if (!zero) {
                    const cx<T> inv = crecip(piv);
#pragma unroll
                    for (int q = 0; q < RP_RPT; ++q) {
                        const int r = j0 + t + RP_T * q;
                        if (r < n && r > col) {
                            const cx<T> l = x[q][cc] * inv;
                            x[q][cc] = l;
#pragma unroll
                            for (int c = cc + 1; c < CW; ++c) x[q][c] = fms(x[q][c], l, urow[c]);
                        }
                    }
                }
                // no barrier here: the next column writes red_v / red_i only after the barrier
                // above (every read of them is before it) and urow / crow only after its own
                // first barrier (every read of this column's is before that)
            }
        }
        // 4. write back; then this chunk's interchanges on the earlier chunks' columns
#pragma unroll
        for (int q = 0; q < RP_RPT; ++q) {
            const int r = j0 + t + RP_T * q;
            if (r < n)
#pragma unroll
                for (int c = 0; c < CW; ++c)
                    if (c < cw) A[r + (long)(j0 + c0 + c) * lda] = x[q][c];
        }
        if (c0 > 0) {
            if (t == 0) {
                // rows touched by this chunk's swaps and, for each, the row whose values it receives
                int na = 0;
                for (int cc = 0; cc < cw; ++cc) {
                    const int rr[2] = {j0 + c0 + cc, sw_p[c0 + cc]};
                    for (int e = 0; e < 2; ++e) {
                        bool seen = false;
                        for (int a2 = 0; a2 < na; ++a2) seen |= aff[a2] == rr[e];
                        if (!seen) aff[na++] = rr[e];
                    }
                }
                for (int a2 = 0; a2 < na; ++a2) {
                    int src = aff[a2];
                    for (int cc = cw - 1; cc >= 0; --cc) {
                        const int rc = j0 + c0 + cc, pc = sw_p[c0 + cc];
                        if (src == rc) src = pc;
                        else if (src == pc) src = rc;
                    }
                    src_s[a2] = src;
                }
                naff_s = na;
            }
            __syncthreads();
            const int na = naff_s;
            cx<T> v = mk<T>(0, 0);
            const bool act = t < na * c0;
            const int a2 = act ? t / c0 : 0, cc = act ? t % c0 : 0;
            if (act) v = A[src_s[a2] + (long)(j0 + cc) * lda];
            __syncthreads();
            if (act) A[aff[a2] + (long)(j0 + cc) * lda] = v;
        }
        __syncthreads();
    }
}

// Panel factorisation spread over the whole device (cooperative launch, grid G x batch).
// Each thread holds one row of the (n - j0) x w panel (w <= 32) in registers, 1024 rows per
// workgroup for complex float, 256 for complex double (few workgroups per matrix: the
// per-column barrier's atomics on one counter serialise, so its cost grows with their number). Per column: every workgroup publishes its best pivot candidate (value, row index
// and the candidate row itself) and the workgroup holding the diagonal row publishes that row;
// after one barrier among the matrix's workgroups every workgroup picks the same pivot
// (largest |re|+|im|, smallest row on ties), the two rows trade places and all rows below the
// diagonal are eliminated. Published data alternate between two buffers by column parity, so
// one barrier per column suffices. Same pivots and arithmetic order per element as k_panel.
//
// The barrier is a counter per (matrix, panel) in global memory (zeroed when the factorisation
// starts): each workgroup adds 1 with release semantics and waits (acquire) until all G have
// arrived for the column. The cooperative launch guarantees that all workgroups are resident.
// The wait is bounded: after CP_SPIN_LIMIT polls it gives up and flags info[b] = -1.
constexpr int CP_W = 32;
// rows (threads) per workgroup: 1024 for complex float; 256 for complex double, whose 32-column
// row (128 registers) leaves no room for more threads
template <typename T>
struct CoopRows {
    static constexpr int value = 1024;
};
template <>
struct CoopRows<double> {
    static constexpr int value = 256;
};
// candidate rows per workgroup of a tournament round: complex float, one wave holding two rows
// per lane (k_calu_round_wave); complex double, one row per lane (k_calu_round)
template <typename T>
struct CaluRows {
    static constexpr int value = 128;
};
template <>
struct CaluRows<double> {
    static constexpr int value = 256;
};
constexpr long CP_SPIN_LIMIT = 1L << 22;

__device__ __forceinline__ void matrix_barrier(unsigned* cnt, unsigned target, int* flag) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        long spins = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > CP_SPIN_LIMIT) {
                *flag = -1;
                break;
            }
        }
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(CoopRows<T>::value) void k_panel_coop(int n, int j0, int w, cx<T>* A, long lda, long sA,
                                                                   int* ipiv, long sP, int* info, T* cv, int* ci,
                                                                   cx<T>* crow, cx<T>* drow, unsigned* cnt) {
    constexpr int CP_T = CoopRows<T>::value;
    __shared__ T red_v[CP_T / 64];
    __shared__ int red_i[CP_T / 64];
    __shared__ T gv[CP_T];
    __shared__ int gi[CP_T];
    __shared__ cx<T> urow[CP_W];
    __shared__ int piv_s, gp_s;
    const int G = gridDim.x, g = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
    const int batch = gridDim.y;
    A += (long)b * sA;
    ipiv += (long)b * sP;
    const int row = j0 + g * CP_T + t;
    const bool have = row < n;
    cx<T> x[CP_W];
#pragma unroll
    for (int c = 0; c < CP_W; ++c) x[c] = (have && c < w) ? A[row + (long)(j0 + c) * lda] : mk<T>(0, 0);
#pragma clang loop unroll(full)
    for (int c = 0; c < CP_W; ++c) {
        if (c < w) {
            const int col = j0 + c, par = c & 1;
            const long slot = ((long)par * batch + b) * G + g;
            // local candidate
            T best = (have && row >= col) ? cabs1(x[c]) : (T)-1;
            int bi = (have && row >= col) ? row : n;
            for (int off = 32; off > 0; off >>= 1) {
                T ov = __shfl_xor(best, off, 64);
                int oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if ((t & 63) == 0) { red_v[t >> 6] = best; red_i[t >> 6] = bi; }
            __syncthreads();
            T bv = red_v[0];
            int bx = red_i[0];
#pragma unroll
            for (int q = 1; q < CP_T / 64; ++q)
                if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bx)) { bv = red_v[q]; bx = red_i[q]; }
            if (t == 0) { cv[slot] = bv; ci[slot] = bx; }
            if (have && row == bx)
#pragma unroll
                for (int u = 0; u < CP_W; ++u)
                    if (u < w) crow[slot * CP_W + u] = x[u];
            if (have && row == col)
#pragma unroll
                for (int u = 0; u < CP_W; ++u)
                    if (u < w) drow[((long)par * batch + b) * CP_W + u] = x[u];
            matrix_barrier(cnt + (long)b * n + j0, (unsigned)((c + 1) * G), info + b);
            // global pivot: same choice in every workgroup
            const long base = ((long)par * batch + b) * G;
            T v2 = (T)-1;
            int i2 = n, g2 = 0;
            for (int q = t; q < G; q += CP_T) {
                T qv = cv[base + q];
                int qi = ci[base + q];
                if (qv > v2 || (qv == v2 && qi < i2)) { v2 = qv; i2 = qi; g2 = q; }
            }
            gv[t] = v2;
            gi[t] = i2;
            __syncthreads();
            if (t == 0) {
                T bv2 = gv[0];
                int bi2 = gi[0], bg = 0;
                const int lim = G < CP_T ? G : CP_T;
                for (int q = 1; q < lim; ++q)
                    if (gv[q] > bv2 || (gv[q] == bv2 && gi[q] < bi2)) { bv2 = gv[q]; bi2 = gi[q]; bg = q; }
                // the winning entry's workgroup: entries q, q + CP_T, ... were reduced into slot q
                int wg = bg;
                for (int q = bg; q < G; q += CP_T)
                    if (cv[base + q] == bv2 && ci[base + q] == bi2) { wg = q; break; }
                piv_s = bi2;
                gp_s = wg;
            }
            (void)g2;
            __syncthreads();
            const int p = piv_s;
            if (t < w) urow[t] = crow[(base + gp_s) * CP_W + t];
            __syncthreads();
            if (g == 0 && t == 0) ipiv[col] = p;
            // trade rows col and p
            if (have && row == p && p != col)
#pragma unroll
                for (int u = 0; u < CP_W; ++u)
                    if (u < w) x[u] = drow[((long)par * batch + b) * CP_W + u];
            if (have && row == col)
#pragma unroll
                for (int u = 0; u < CP_W; ++u)
                    if (u < w) x[u] = urow[u];
            const cx<T> piv = urow[c];
            const bool zero = piv.re == 0 && piv.im == 0;
            if (zero && g == 0 && t == 0 && info[b] == 0) info[b] = col + 1;
            if (!zero && have && row > col) {
                const cx<T> l = x[c] * crecip(piv);
                x[c] = l;
#pragma unroll
                for (int u = c + 1; u < CP_W; ++u)
                    if (u < w) x[u] = fms(x[u], l, urow[u]);
            }
        }
    }
    if (have)
#pragma unroll
        for (int c = 0; c < CP_W; ++c)
            if (c < w) A[row + (long)(j0 + c) * lda] = x[c];
}

// Scratch of the cooperative panel (in the calling thread's handle, grown on demand): candidates,
// diagonal rows and the barrier counters (batch x n).
struct CoopWS {
    void* crow;
    void* drow;
    void* cv;
    int* ci;
    unsigned* cnt;
    void* top;    // factored top blocks of tournament panels (batch x CP_W x CP_W)
    int gmax;
};

__global__ void k_zero_u32(long count, unsigned* p) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Sizes the scratch for panels of up to n rows and clears the counters; false if unusable.
template <typename T>
static bool coop_prepare(int n, int batch, hipStream_t s, CoopWS& ws) {
    Ctx& h = ctx();
    if (n < h.coop_min_rows) return false;
    // workgroups per matrix of the cooperative panel or of a tournament round, at most
    constexpr int CP_T = CoopRows<T>::value < CaluRows<T>::value ? CoopRows<T>::value : CaluRows<T>::value;
    const int gmax = (n + CP_T - 1) / CP_T;
    const size_t ncand = 2 * (size_t)batch * gmax;
    const size_t tbytes = (size_t)batch * CP_W * CP_W * sizeof(cx<T>);
    const size_t bytes = ncand * CP_W * sizeof(cx<T>) + 2 * (size_t)batch * CP_W * sizeof(cx<T>) +
                         ncand * (sizeof(T) + sizeof(int)) + (size_t)batch * n * sizeof(unsigned) + tbytes + 1280;
    if (bytes > h.coop_bytes) {
        if (h.coop_ws) {
            hipStreamSynchronize(s);
            hipFree(h.coop_ws);
        }
        if (hipMalloc(&h.coop_ws, bytes) != hipSuccess) {
            h.coop_ws = nullptr;
            h.coop_bytes = 0;
            return false;
        }
        h.coop_bytes = bytes;
    }
    char* p = (char*)h.coop_ws;
    ws.crow = p;
    p += ncand * CP_W * sizeof(cx<T>);
    ws.drow = p;
    p += 2 * (size_t)batch * CP_W * sizeof(cx<T>);
    ws.cv = p;
    p += ncand * sizeof(T);
    ws.ci = (int*)p;
    p += ncand * sizeof(int);
    ws.cnt = (unsigned*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    ws.top = (void*)((((uintptr_t)(ws.cnt + (size_t)batch * n)) + 255) & ~(uintptr_t)255);
    ws.gmax = gmax;
    const long nc = (long)batch * n;
    hipLaunchKernelGGL(k_zero_u32, dim3((nc + 255) / 256), dim3(256), 0, s, nc, ws.cnt);
    return true;
}

template <typename T>
static bool panel_coop(int n, int j0, int w, cx<T>* a, long lda, long sA, int* ipiv, long sP, int* info, int batch,
                       const CoopWS& ws, hipStream_t s) {
    constexpr int CP_T = CoopRows<T>::value;
    static int capacity = -1;
    if (capacity < 0) {
        int dev = 0, per_cu = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_panel_coop<T>, CP_T, 0) != hipSuccess) per_cu = 0;
        capacity = per_cu * cus;
    }
    const int G = (n - j0 + CP_T - 1) / CP_T;
    // a barrier per column costs more than a single workgroup walking a short panel
    if (n - j0 < ctx().coop_min_rows || w > CP_W || G > ws.gmax || (long)G * batch > capacity) return false;
    // candidate slots are laid out for G = gmax
    T* cv = (T*)ws.cv;
    int* ci = ws.ci;
    cx<T>* crow = (cx<T>*)ws.crow;
    cx<T>* drow = (cx<T>*)ws.drow;
    unsigned* cnt = ws.cnt;
    void* args[] = {&n, &j0, &w, &a, &lda, &sA, &ipiv, &sP, &info, &cv, &ci, &crow, &drow, &cnt};
    return hipLaunchCooperativeKernel((const void*)k_panel_coop<T>, dim3(G, batch), dim3(CP_T), args, 0, s) ==
           hipSuccess;
}

// ---------------------------------------------------------------------------------------------
// Tall panels by tournament pivoting (CALU / TSLU): the w pivot rows of a panel of m rows are
// chosen by a reduction tree instead of a search across the whole column for every column.
//   k_calu_round (complex double), k_calu_round_wave (complex float): each workgroup takes up
//          to CaluRows candidate rows (a contiguous block of the panel, or the winners of the
//          previous round, from their original entries), runs LU with partial pivoting on them
//          in registers (largest |re|+|im|, the earlier candidate on a tie, as izamax) and keeps
//          its w pivot rows in pivot order; rounds repeat until one workgroup remains, which
//          also writes the pivot magnitudes;
//   k_calu_ipiv: LAPACK-style interchanges ipiv[j0 + c] that bring the winners to the top;
//   k_laswp on the panel's own columns;
//   k_calu_top: the w x w top block (the winners' rows) eliminated in pivot order;
//   k_calu_l21: L21 = A21 U11^{-1}, one row per lane.
// The panel's interchanges and factors keep the layout of the other panel kernels, so the rest
// of the blocked LU is unchanged. Pivots differ from partial pivoting's; no workgroup waits for
// another. In k_calu_round every elimination step costs one workgroup barrier (each wave
// publishes its best candidate row in LDS, double-buffered by step parity, and every wave then
// picks the same winner from those); k_calu_round_wave is a single wave and needs none.

// wave-wide maxima of the device library (DPP reductions; every lane gets the result)
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_max_u32(unsigned int);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(unsigned long long);

template <typename T>
struct PivKey {
    using type = unsigned;
    static __device__ __forceinline__ type of(T v) { return __float_as_uint(v) + 1u; }
    static __device__ __forceinline__ type wave_max(type k) { return __ockl_wfred_max_u32(k); }
    static __device__ __forceinline__ T value(type k) { return __uint_as_float(k - 1u); }
};
template <>
struct PivKey<double> {
    using type = unsigned long long;
    static __device__ __forceinline__ type of(double v) { return (type)__double_as_longlong(v) + 1ull; }
    static __device__ __forceinline__ type wave_max(type k) { return __ockl_wfred_max_u64(k); }
    static __device__ __forceinline__ double value(type k) { return __longlong_as_double((long long)(k - 1ull)); }
};

// One round. in == nullptr: candidates are rows j0 + c CP_T + t (< n); else in[b nin + c CP_T + t]
// (< nin; -1 = none). out[(b G + c) w + k]: k-th winner (global row, -1 = none); best (the last
// round, a single workgroup; else null): best[b w + k] = |pivot| of step k.
template <typename T>
__global__ __launch_bounds__(CaluRows<T>::value) void k_calu_round(int n, int j0, int w, const cx<T>* A, long lda,
                                                                   long sA, const int* in, int nin, int* out,
                                                                   T* best) {
    constexpr int CP_T = CaluRows<T>::value;
    constexpr int NW = CP_T / 64;
    using K = typename PivKey<T>::type;
    __shared__ K key_s[2][NW];
    __shared__ int row_s[2][NW];
    __shared__ cx<T> prow[2][NW][CP_W];
    const int G = gridDim.x, c = blockIdx.x, b = blockIdx.y, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    A += (long)b * sA;
    int row = -1;
    if (in == nullptr) {
        const int r = j0 + c * CP_T + t;
        row = r < n ? r : -1;
    } else {
        const int q = c * CP_T + t;
        row = q < nin ? in[(long)b * nin + q] : -1;
    }
    bool live = row >= 0;
    // columns from w on stay zero, so that every step updates all CP_W columns without a branch
    cx<T> x[CP_W];
#pragma unroll
    for (int u = 0; u < CP_W; ++u) x[u] = (live && u < w) ? A[row + (long)(j0 + u) * lda] : mk<T>(0, 0);
    int* o = out + ((long)b * G + c) * w;
#pragma clang loop unroll(full)
    for (int k = 0; k < CP_W; ++k) {
        if (k < w) {
            const int buf = k & 1;
            const K mine = live ? PivKey<T>::of(cabs1(x[k])) : (K)0;
            const K wmax = PivKey<T>::wave_max(mine);
            const unsigned long long hit = __ballot(live && mine == wmax);
            const int wl = hit ? __ffsll(hit) - 1 : -1;
            if (lane == wl) {
                key_s[buf][wv] = wmax;
                row_s[buf][wv] = row;
#pragma unroll
                for (int u = 0; u < CP_W; ++u) prow[buf][wv][u] = x[u];
            } else if (wl < 0 && lane == 0) {
                key_s[buf][wv] = 0;
            }
            __syncthreads();
            K bk = key_s[buf][0];
            int bw = 0;
#pragma unroll
            for (int q = 1; q < NW; ++q) {
                const K kq = key_s[buf][q];
                if (kq > bk) { bk = kq; bw = q; }
            }
            if (bk != 0) {
                const int win = row_s[buf][bw];
                live = live && row != win;
                const cx<T> piv = prow[buf][bw][k];
                const bool zero = piv.re == 0 && piv.im == 0;
                const cx<T> r = zero ? mk<T>(0, 0) : crecip(piv);
                // a zero pivot leaves every live x[k] at zero: no update, as partial pivoting
                const cx<T> l = live ? x[k] * r : mk<T>(0, 0);
                if (live) x[k] = l;
#pragma unroll
                for (int u = k + 1; u < CP_W; ++u) x[u] = fms(x[u], l, prow[buf][bw][u]);
                if (t == 0) {
                    o[k] = win;
                    if (best != nullptr) best[(long)b * w + k] = PivKey<T>::value(bk);
                }
            } else {
                if (t == 0) o[k] = -1;
                if (best != nullptr && t == 0) best[(long)b * w + k] = (T)0;
            }
        }
    }
}

extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_min_u32(unsigned int);

// One round for complex float in a single wave: candidate q of the block (q < 64 R) sits in
// lane q % 64, slot q / 64, R rows per lane in registers. The loop over the w elimination steps
// is not unrolled (the fully unrolled form is larger than the instruction cache and ran bound by
// instruction fetch): the rows are kept as shift registers, whose element 0 is the current pivot
// column; after each step the remaining columns move one place down, so every register index is
// a constant. A step needs no barrier: the wave maximum of the lanes' best |re|+|im| (the
// earlier slot on a tie inside a lane), the wave minimum of the candidate index among the lanes
// that hold it (the earlier candidate on a tie, as izamax), the winner's row through LDS to
// every lane, and the rank-1 update of the live rows on the columns still in the panel. Only the
// winners are kept (and, for the last round, the pivot magnitudes): the factored top block is
// formed afterwards from the winners' rows (k_calu_top). Pivot rule as k_calu_round, blocks of
// 64 R rows.
template <int R>
__global__ __launch_bounds__(64) void k_calu_round_wave(int n, int j0, int w, const cx<float>* A, long lda, long sA,
                                                        const int* in, int nin, int* out, float* best) {
    constexpr int CP = 64 * R;
    using K = PivKey<float>;
    __shared__ cx<float> prow[CP_W];
    __shared__ int prow_id;
    const int G = gridDim.x, c = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    A += (long)b * sA;
    int row[R];
    bool live[R];
    cx<float> x[R][CP_W];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int q = c * CP + r * 64 + lane;
        if (in == nullptr) {
            const int g = j0 + q;
            row[r] = g < n ? g : -1;
        } else {
            row[r] = q < nin ? in[(long)b * nin + q] : -1;
        }
        live[r] = row[r] >= 0;
#pragma unroll
        for (int u = 0; u < CP_W; ++u)
            x[r][u] = (live[r] && u < w) ? A[row[r] + (long)(j0 + u) * lda] : mk<float>(0, 0);
    }
    int* o = out + ((long)b * G + c) * w;
#pragma nounroll
    for (int k = 0; k < w; ++k) {
        unsigned key = 0, idx = 0xffffffffu;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const unsigned kr = live[r] ? K::of(cabs1(x[r][0])) : 0u;
            if (kr > key) {
                key = kr;
                idx = (unsigned)(r * 64 + lane);
            }
        }
        const unsigned wmax = K::wave_max(key);
        if (wmax == 0u) {
            // no live candidate left
            if (lane == 0) {
                o[k] = -1;
                if (best != nullptr) best[(long)b * w + k] = 0.0f;
            }
        } else {
            const unsigned win = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_u32(key == wmax ? idx : 0xffffffffu));
            const int wl = (int)(win & 63u), ws = (int)(win >> 6);
            if (lane == wl) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (ws == r) {
#pragma unroll
                        for (int u = 0; u < CP_W; ++u) prow[u] = x[r][u];
                        prow_id = row[r];
                        live[r] = false;
                    }
            }
            __syncthreads();
            cx<float> pr[CP_W];
#pragma unroll
            for (int u = 0; u < CP_W; ++u) pr[u] = prow[u];
            const int wrow = prow_id;
            const bool zero = pr[0].re == 0 && pr[0].im == 0;
            const cx<float> rp = zero ? mk<float>(0, 0) : crecip(pr[0]);
            // columns k + 1 .. w - 1 of the panel sit at 1 .. w - 1 - k; groups of 8 past them
            // are skipped (a branch the whole wave takes)
            const int left = w - k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                // a zero pivot leaves every live x[0] at zero: no update, as partial pivoting
                const cx<float> l = live[r] ? x[r][0] * rp : mk<float>(0, 0);
#pragma unroll
                for (int g = 0; g < CP_W / 8; ++g)
                    if (8 * g < left)
#pragma unroll
                        for (int u = 8 * g; u < 8 * g + 8; ++u)
                            if (u > 0) x[r][u] = fms(x[r][u], l, pr[u]);
#pragma unroll
                for (int u = 0; u + 1 < CP_W; ++u) x[r][u] = x[r][u + 1];
                x[r][CP_W - 1] = mk<float>(0, 0);
            }
            if (lane == 0) {
                o[k] = wrow;
                if (best != nullptr) best[(long)b * w + k] = K::value(wmax);
            }
            __syncthreads();      // prow is rewritten at the next step
        }
    }
}

// The factored w x w top block of a tournament panel: the winners' rows, moved to rows
// j0 .. j0 + w by the interchanges, eliminated in pivot order without further pivoting (the
// same arithmetic as the rounds: l = x_k / pivot through the reciprocal, x_u -= l p_u, and a
// zero pivot leaves its column unscaled), in place, with a row-major copy in top[] for
// k_calu_l21. One wave per matrix, the block in LDS, lane r owns row r.
template <typename T>
__global__ __launch_bounds__(64) void k_calu_top(int j0, int w, cx<T>* A, long lda, long sA, cx<T>* top) {
    __shared__ cx<T> t_[CP_W][CP_W + 1];
    const int b = blockIdx.x, r = threadIdx.x;
    A += (long)b * sA;
    top += (long)b * CP_W * CP_W;
    for (int e = r; e < w * w; e += 64) t_[e % w][e / w] = A[j0 + e % w + (long)(j0 + e / w) * lda];
    __syncthreads();
    for (int k = 0; k < w; ++k) {
        if (r > k && r < w) {
            const cx<T> piv = t_[k][k];
            const bool zero = piv.re == 0 && piv.im == 0;
            const cx<T> rp = zero ? mk<T>(0, 0) : crecip(piv);
            const cx<T> l = t_[r][k] * rp;
            t_[r][k] = l;
            for (int u = k + 1; u < w; ++u) t_[r][u] = fms(t_[r][u], l, t_[k][u]);
        }
        __syncthreads();
    }
    for (int e = r; e < w * w; e += 64) {
        const int i = e % w, jj = e / w;
        A[j0 + i + (long)(j0 + jj) * lda] = t_[i][jj];
        top[i * CP_W + jj] = t_[i][jj];
    }
}

// Interchanges that bring the winners sel[0 .. w) (global rows, in pivot order) to rows j0 + c,
// applied one after the other (LAPACK convention), and info from a zero pivot. One wave per
// matrix: lane q < w holds the original row now at position j0 + q, lane q < nd one position
// below the top block that a displaced row now occupies.
template <typename T>
__global__ __launch_bounds__(64) void k_calu_ipiv(int j0, int w, const int* sel, const T* best, int* ipiv, long sP,
                                                  int* info) {
    const int b = blockIdx.x, q = threadIdx.x;
    sel += (long)b * w;
    best += (long)b * w;
    int cur = q < w ? j0 + q : -1;
    int dpos = -1, drow = -1, nd = 0, first_zero = 0;
    for (int c = 0; c < w; ++c) {
        const int col = j0 + c;
        const int want = sel[c];
        int p = col;
        if (want >= 0) {
            const unsigned long long in_top = __ballot(q >= c && q < w && cur == want);
            const unsigned long long in_down = __ballot(q < nd && drow == want);
            int slot = -1;
            if (in_top) {
                p = j0 + __ffsll(in_top) - 1;
            } else if (in_down) {
                slot = __ffsll(in_down) - 1;
                p = __shfl(dpos, slot, 64);
            } else {
                p = want;
            }
            const int rc = __shfl(cur, c, 64);
            if (p != col) {
                if (p < j0 + w) {
                    if (q == p - j0) cur = rc;
                } else if (slot >= 0) {
                    if (q == slot) drow = rc;
                } else {
                    if (q == nd) { dpos = p; drow = rc; }
                    ++nd;
                }
                if (q == c) cur = want;
            }
        }
        if (q == 0) ipiv[(long)b * sP + col] = p;
        if (best[c] == (T)0 && first_zero == 0) first_zero = col + 1;
    }
    if (q == 0 && first_zero != 0 && info[b] == 0) info[b] = first_zero;
}

// L21 = A21 U11^{-1} for the rows j0 + w .. n, one lane per row, NT rows per workgroup. U11's
// entries are the same for every lane: they are read from top[] with uniform addresses (scalar
// loads), so that only the row itself lives in vector registers; the substitution runs
// right-looking (after x[k] is final, the later columns' updates are independent).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_calu_l21(int n, int j0, int w, cx<T>* A, long lda, long sA,
                                                 const cx<T>* __restrict__ top) {
    const int t = threadIdx.x, b = blockIdx.y;
    A += (long)b * sA;
    top += (long)b * CP_W * CP_W;
    const int row = j0 + w + blockIdx.x * NT + t;
    if (row >= n) return;
    cx<T> x[CP_W];
#pragma unroll
    for (int c = 0; c < CP_W; ++c) x[c] = c < w ? A[row + (long)(j0 + c) * lda] : mk<T>(0, 0);
#pragma unroll
    for (int k = 0; k < CP_W; ++k) {
        if (k < w) {
            const cx<T> d = top[k * CP_W + k];
            // a zero pivot leaves its column unscaled
            if (d.re != 0 || d.im != 0) x[k] = x[k] * crecip(d);
#pragma unroll
            for (int c = k + 1; c < CP_W; ++c)
                if (c < w) x[c] = fms(x[c], x[k], top[k * CP_W + c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CP_W; ++c)
        if (c < w) A[row + (long)(j0 + c) * lda] = x[c];
}

// apply interchanges ipiv[r0 .. r1) to the columns [a0, a1) and [b0, b1)
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(int r0, int r1, int a0, int a1, int b0, int b1, cx<T>* A, long lda,
                                               long sA, const int* ipiv, long sP) {
    const int b = blockIdx.y;
    int c = blockIdx.x * 256 + threadIdx.x;
    const int na = a1 - a0;
    if (c >= na + (b1 - b0)) return;
    c = (c < na) ? a0 + c : b0 + (c - na);
    A += (long)b * sA + (long)c * lda;
    ipiv += (long)b * sP;
    for (int r = r0; r < r1; ++r) {
        int p = ipiv[r];
        if (p != r) { cx<T> x = A[r]; A[r] = A[p]; A[p] = x; }
    }
}

// X = L^{-1} X for the unit-lower nb x nb block at (j0, j0), X = rows j0 .. j0+nb of the
// columns [c0, c1); one lane per column.
template <typename T, int NB>
__global__ __launch_bounds__(128) void k_trsm_lln(int c0, int c1, int j0, int nb, cx<T>* A, long lda, long sA) {
    __shared__ cx<T> L[NB][NB + 1];
    const int b = blockIdx.y;
    A += (long)b * sA;
    for (int e = threadIdx.x; e < nb * nb; e += 128) {
        int r = e % nb, c = e / nb;
        L[r][c] = A[(j0 + r) + (long)(j0 + c) * lda];
    }
    __syncthreads();
    int col = c0 + blockIdx.x * 128 + threadIdx.x;
    if (col >= c1) return;
    cx<T>* x = A + j0 + (long)col * lda;
    cx<T> v[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r)
        if (r < nb) v[r] = x[r];
#pragma unroll
    for (int r = 1; r < NB; ++r) {
        if (r < nb) {
            cx<T> s = v[r];
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (q < r) s = fms(s, L[r][q], v[q]);
            v[r] = s;
        }
    }
#pragma unroll
    for (int r = 0; r < NB; ++r)
        if (r < nb) x[r] = v[r];
}

// pf: P b = b[pf] for the interchange list; qinv = pf^{-1}. One workgroup per matrix; the
// sequential interchange pass runs in LDS on one lane.
__global__ __launch_bounds__(64) void k_perm(int n, const int* ipiv, long sP, int* pf, int* qinv, long sQ) {
    extern __shared__ int sp[];
    const int b = blockIdx.x, t = threadIdx.x;
    ipiv += (long)b * sP;
    pf += (long)b * sQ;
    qinv += (long)b * sQ;
    for (int i = t; i < n; i += 64) sp[i] = i;
    __syncthreads();
    if (t == 0) {
        for (int i = 0; i < n; ++i) {
            int p = ipiv[i];
            int x = sp[i];
            sp[i] = sp[p];
            sp[p] = x;
        }
    }
    __syncthreads();
    for (int i = t; i < n; i += 64) {
        int v = sp[i];
        pf[i] = v;
        qinv[v] = i;
    }
}

template <typename T>
static void laswp(int r0, int r1, int a0, int a1, int b0, int b1, cx<T>* a, long lda, long sA, int* ipiv, long sP,
                  int batch, hipStream_t s) {
    int cols = (a1 - a0) + (b1 - b0);
    if (cols <= 0 || r1 <= r0) return;
    hipLaunchKernelGGL(k_laswp<T>, dim3((cols + 255) / 256, batch), dim3(256), 0, s, r0, r1, a0, a1, b0, b1, a, lda,
                       sA, ipiv, sP);
}

static void round_launch(int G, int batch, int n, int j0, int w, const cx<float>* a, long lda, long sA, const int* in,
                         int nin, int* out, float* best, hipStream_t s) {
    static_assert(CaluRows<float>::value == 128, "two rows per lane of one wave");
    hipLaunchKernelGGL(k_calu_round_wave<2>, dim3(G, batch), dim3(64), 0, s, n, j0, w, a, lda, sA, in, nin, out, best);
}
static void round_launch(int G, int batch, int n, int j0, int w, const cx<double>* a, long lda, long sA,
                         const int* in, int nin, int* out, double* best, hipStream_t s) {
    hipLaunchKernelGGL(k_calu_round<double>, dim3(G, batch), dim3(CaluRows<double>::value), 0, s, n, j0, w, a, lda, sA,
                       in, nin, out, best);
}

// Tall panel by tournament pivoting (see k_calu_round); false when the scratch does not fit.
template <typename T>
static bool panel_calu(int n, int j0, int w, cx<T>* a, long lda, long sA, int* ipiv, long sP, int* info, int batch,
                       const CoopWS& ws, hipStream_t s) {
    constexpr int CP_T = CaluRows<T>::value;
    const int m = n - j0;
    int G = (m + CP_T - 1) / CP_T;
    if (m < ctx().coop_min_rows || w > CP_W || G > ws.gmax) return false;
    // two index lists of batch x gmax x CP_W ints in the candidate-row scratch; winners and their
    // magnitudes in the diagonal-row scratch; the factored top blocks in ws.top
    int* lst[2] = {(int*)ws.crow, (int*)ws.crow + (long)batch * ws.gmax * CP_W};
    int* sel = (int*)ws.drow;
    T* best = (T*)((int*)ws.drow + (long)batch * CP_W);
    cx<T>* top = (cx<T>*)ws.top;
    int cur = 0;
    const int* in = nullptr;
    int nin = 0;
    for (;;) {
        const bool last = G == 1;
        round_launch(G, batch, n, j0, w, (const cx<T>*)a, lda, sA, in, nin, last ? sel : lst[cur],
                     last ? best : (T*)nullptr, s);
        if (last) break;
        in = lst[cur];
        nin = G * w;
        cur = 1 - cur;
        G = (nin + CP_T - 1) / CP_T;
    }
    hipLaunchKernelGGL(k_calu_ipiv<T>, dim3(batch), dim3(64), 0, s, j0, w, (const int*)sel, (const T*)best, ipiv, sP,
                       info);
    laswp<T>(j0, j0 + w, j0, j0 + w, 0, 0, a, lda, sA, ipiv, sP, batch, s);
    hipLaunchKernelGGL(k_calu_top<T>, dim3(batch), dim3(64), 0, s, j0, w, a, lda, sA, top);
    // 64 rows per workgroup: a 25k-row panel of one matrix spreads over every CU
    const int gx = m > w ? (m - w + 63) / 64 : 1;
    hipLaunchKernelGGL((k_calu_l21<T, 64>), dim3(gx, batch), dim3(64), 0, s, n, j0, w, a, lda, sA, (const cx<T>*)top);
    return true;
}

template <typename T>
static void trsm_lln(int c0, int c1, int j0, int w, int nb, cx<T>* a, long lda, long sA, int batch, hipStream_t s) {
    if (c1 <= c0) return;
    dim3 g((c1 - c0 + 127) / 128, batch);
    if (nb <= 32)
        hipLaunchKernelGGL((k_trsm_lln<T, 32>), g, dim3(128), 0, s, c0, c1, j0, w, a, lda, sA);
    else
        hipLaunchKernelGGL((k_trsm_lln<T, 64>), g, dim3(128), 0, s, c0, c1, j0, w, a, lda, sA);
}

// Two-level blocking: outer blocks of NBO columns carry the large trailing update (a GEMM with
// k = NBO, so C is read and written once per NBO columns), inner panels of nb columns are
// factored inside the outer block, by recursive halving (default) or one after the other.
// Tuning knobs of the calling thread's handle: 0 = outer block width (columns, >= nb), 1 =
// lookahead on/off, 2 = register-resident panel on/off, 3 = rows from which panels use the
// cooperative kernel, 4 = complex-float GEMM mode (0: FP32 matrix cores; 12, 13: split-BF16
// planes of 2 or 3 terms; 16: scaled fp16 pairs), 5 = workgroup tile of the plane GEMM (11,
// 111, 21, 12, 22), 6 = row tiles per group of its XCD-aware tile order (0: grid order), 7 =
// rows per workgroup of the Green-function assembly (64, 128, 256), 8 = wavelengths per assembly
// thread (0: all), 9 = tall panels by tournament pivoting, 10 = recursive outer blocks, 11 =
// one-pass scale and split of fp16-pair operands.
int set_gemm_split(int v);
int get_gemm_split();
int set_plane_tile(int v);
int get_plane_tile();
int set_plane_group(int v);
int get_plane_group();
int set_asm_rows(int v);
int set_asm_lchunk(int v);
int set_tuning(int key, int value) {
    switch (key) {
        case 7: return set_asm_rows(value);
        case 8: return set_asm_lchunk(value);
        case 9: if (value != 0 && value != 1) return -2; ctx().tall_panel = value; return 0;
        case 10: if (value != 0 && value != 1) return -2; ctx().lu_recursive = value; return 0;
        case 11: if (value != 0 && value != 1) return -2; ctx().f16_fused = value; return 0;
        case 4: return set_gemm_split(value);
        case 5: return set_plane_tile(value);
        case 6: return set_plane_group(value);
        case 0: if (value < 16 || value > 1024) return -2; ctx().nbo = value; return 0;
        case 1: ctx().lookahead = value != 0; return 0;
        case 2: ctx().panel_reg = value != 0; return 0;
        case 3: if (value < 256) return -2; ctx().coop_min_rows = value; return 0;
        default: return -1;
    }
}
int get_tuning(int key) {
    switch (key) {
        case 0: return ctx().nbo;
        case 1: return ctx().lookahead ? 1 : 0;
        case 2: return ctx().panel_reg ? 1 : 0;
        case 3: return ctx().coop_min_rows;
        case 4: return get_gemm_split();
        case 5: return get_plane_tile();
        case 6: return get_plane_group();
        case 7: return ctx().asm_rows;
        case 8: return ctx().asm_lchunk;
        case 9: return ctx().tall_panel;
        case 10: return ctx().lu_recursive;
        case 11: return ctx().f16_fused;
        default: return -1;
    }
}

// Lookahead: the trailing update of an outer block is split into the next block's columns
// (on the caller's stream) and the rest (on a second stream, default priority). The rest runs
// while the next block's panels are factored: those touch only the next block's columns, and
// the interchanges outside the block wait for the rest to finish (event).
struct SideStream {
    hipStream_t st;
    hipEvent_t ready, done;
};
static SideStream side_stream() {
    Ctx& h = ctx();
    if (!h.side) {
        hipStreamCreateWithFlags(&h.side, hipStreamNonBlocking);
        hipEventCreateWithFlags(&h.side_ready, hipEventDisableTiming);
        hipEventCreateWithFlags(&h.side_done, hipEventDisableTiming);
    }
    return SideStream{h.side, h.side_ready, h.side_done};
}

bool gemm_reserve(int m, int n, int k, int batch, hipStream_t s, hipStream_t side);

// Workspaces the blocked LU and right-side solves of n x n matrices need on stream s: the plane
// workspaces of the widest trailing update (for s and the lookahead side stream). False when an
// allocation fails or the panels would need the cooperative scratch (n at or above its
// threshold, which is sized per call).
bool lu_reserve(int n, int batch, int nb, hipStream_t s) {
    Ctx& h = ctx();
    if (n >= h.coop_min_rows) return false;
    const int NB = (h.nbo > nb) ? (h.nbo / nb) * nb : nb;
    const int k = NB < n ? NB : n;
    return gemm_reserve(n, n, k, batch, s, side_stream().st);
}

template <typename T>
void getrf_batched(int n, int batch, void* A, long lda, long sA, int* ipiv, long sP, int* info, int nb,
                   hipStream_t s) {
    cx<T>* a = (cx<T>*)A;
    Ctx& h = ctx();
    const int NB = (h.nbo > nb) ? (h.nbo / nb) * nb : nb;
    CoopWS ws{};
    const bool coop = coop_prepare<T>(n, batch, s, ws);
    SideStream side = side_stream();
    const bool lookahead = h.lookahead;
    bool pending = false;   // a rest update is running on the side stream
    for (int J0 = 0; J0 < n; J0 += NB) {
        const int W = (n - J0 < NB) ? (n - J0) : NB;
        const int JE = J0 + W;
        // factor the tall outer panel A[J0:n, J0:JE]
        // one panel of w <= nb columns at j0, its interchanges applied to the other columns of the
        // outer panel
        auto panel = [&](int j0, int w) {
            const bool tall = coop && (h.tall_panel == 1 ? panel_calu<T>(n, j0, w, a, lda, sA, ipiv, sP, info,
                                                                          batch, ws, s)
                                                         : panel_coop<T>(n, j0, w, a, lda, sA, ipiv, sP, info,
                                                                         batch, ws, s));
            if (!tall) {
                // complex float only: complex double needs 4-column chunks and runs slower than k_panel
                if (sizeof(T) == 4 && n - j0 <= RP_T * RP_RPT && w <= RP_MAXW && h.panel_reg)
                    hipLaunchKernelGGL((k_panel_reg<T, RegChunk<T>::CW>), dim3(batch), dim3(RP_T), 0, s, n, j0, w, a, lda,
                                       sA, ipiv, sP, info);
                else
                    hipLaunchKernelGGL(k_panel<T>, dim3(batch), dim3(PANEL_THREADS), 0, s, n, j0, w, a, lda, sA, ipiv,
                                       sP, info);
            }
            laswp<T>(j0, j0 + w, J0, j0, j0 + w, JE, a, lda, sA, ipiv, sP, batch, s);
        };
        // U12 = L11^{-1} A12 for the rows [r0, r1) and columns [c0, c1), nb rows at a time
        auto solve_u12 = [&](int r0, int r1, int c0, int c1) {
            for (int j0 = r0; j0 < r1; j0 += nb) {
                int w = (r1 - j0 < nb) ? (r1 - j0) : nb;
                trsm_lln<T>(c0, c1, j0, w, nb, a, lda, sA, batch, s);
                if (j0 + w < r1)
                    gemm_sub<T>(r1 - j0 - w, c1 - c0, w, a + (j0 + w) + (long)j0 * lda, lda, sA,
                                a + j0 + (long)c0 * lda, lda, sA, a + (j0 + w) + (long)c0 * lda, lda, sA, batch, s);
            }
        };
        if (h.lu_recursive) {
            // columns [c0, c1): the left half recursively, then the right half's rows c0 .. mid by
            // L11 and its rows below by one GEMM of depth mid - c0, then the right half. Each
            // trailing column of the outer panel is read and written log2(W / nb) times by
            // GEMMs of depth W / 2, W / 4, ... instead of W / nb times by GEMMs of depth nb.
            auto rec = [&](auto&& self, int c0, int c1) -> void {
                const int np = (c1 - c0 + nb - 1) / nb;
                if (np <= 1) {
                    panel(c0, c1 - c0);
                    return;
                }
                const int mid = c0 + ((np + 1) / 2) * nb;
                self(self, c0, mid);
                solve_u12(c0, mid, mid, c1);
                gemm_sub<T>(n - mid, c1 - mid, mid - c0, a + mid + (long)c0 * lda, lda, sA, a + c0 + (long)mid * lda,
                            lda, sA, a + mid + (long)mid * lda, lda, sA, batch, s);
                self(self, mid, c1);
            };
            rec(rec, J0, JE);
        } else {
            for (int j0 = J0; j0 < JE; j0 += nb) {
                int w = (JE - j0 < nb) ? (JE - j0) : nb;
                panel(j0, w);
                int rest = JE - j0 - w;
                if (rest > 0) {
                    trsm_lln<T>(j0 + w, JE, j0, w, nb, a, lda, sA, batch, s);
                    gemm_sub<T>(n - j0 - w, rest, w, a + (j0 + w) + (long)j0 * lda, lda, sA,
                                a + j0 + (long)(j0 + w) * lda, lda, sA, a + (j0 + w) + (long)(j0 + w) * lda, lda, sA,
                                batch, s);
                }
            }
        }
        // interchanges of the outer panel on the columns outside it (after the previous block's
        // rest update, which reads the left columns and writes the right ones)
        if (pending) {
            hipStreamWaitEvent(s, side.done, 0);
            pending = false;
        }
        laswp<T>(J0, JE, 0, J0, JE, n, a, lda, sA, ipiv, sP, batch, s);
        if (JE < n) {
            // U12 = L11^{-1} A12 for the W x (n - JE) block row
            solve_u12(J0, JE, JE, n);
            // trailing update with k = W: the next block's columns here, the rest on the side stream
            const int first = lookahead ? ((n - JE < NB) ? (n - JE) : NB) : n - JE;
            gemm_sub<T>(n - JE, first, W, a + JE + (long)J0 * lda, lda, sA, a + J0 + (long)JE * lda, lda, sA,
                        a + JE + (long)JE * lda, lda, sA, batch, s);
            const int rest = n - JE - first;
            if (rest > 0) {
                hipEventRecord(side.ready, s);
                hipStreamWaitEvent(side.st, side.ready, 0);
                gemm_sub<T>(n - JE, rest, W, a + JE + (long)J0 * lda, lda, sA, a + J0 + (long)(JE + first) * lda, lda,
                            sA, a + JE + (long)(JE + first) * lda, lda, sA, batch, side.st);
                hipEventRecord(side.done, side.st);
                pending = true;
            }
        }
    }
    if (pending) hipStreamWaitEvent(s, side.done, 0);
}

// Same for n beyond what LDS holds: the interchanges run in pf itself (global memory).
__global__ __launch_bounds__(64) void k_perm_global(int n, const int* ipiv, long sP, int* pf, int* qinv, long sQ) {
    const int b = blockIdx.x, t = threadIdx.x;
    ipiv += (long)b * sP;
    pf += (long)b * sQ;
    qinv += (long)b * sQ;
    for (int i = t; i < n; i += 64) pf[i] = i;
    __syncthreads();
    if (t == 0) {
        for (int i = 0; i < n; ++i) {
            int p = ipiv[i];
            int x = pf[i];
            pf[i] = pf[p];
            pf[p] = x;
        }
    }
    __syncthreads();
    for (int i = t; i < n; i += 64) qinv[pf[i]] = i;
}

// the permutation is built in LDS up to 16384 rows (64 KiB), in global memory beyond
void perm_batched(int n, int batch, const int* ipiv, long sP, int* pf, int* qinv, long sQ, hipStream_t s) {
    if (n <= 16384)
        hipLaunchKernelGGL(k_perm, dim3(batch), dim3(64), (size_t)n * sizeof(int), s, n, ipiv, sP, pf, qinv, sQ);
    else
        hipLaunchKernelGGL(k_perm_global, dim3(batch), dim3(64), 0, s, n, ipiv, sP, pf, qinv, sQ);
}

template void getrf_batched<double>(int, int, void*, long, long, int*, long, int*, int, hipStream_t);
template void getrf_batched<float>(int, int, void*, long, long, int*, long, int*, int, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Right-side solve X = B A^{-1} with A = P^T L U, i.e. X = B U^{-1} L^{-1} P.
//   phase 1 (left to right):  B_j <- B_j U_jj^{-1};  B_{>j} -= B_j U_{j,>j}
//   phase 2 (right to left):  B_j <- B_j L_jj^{-1};  B_{<j} -= B_j L_{j,<j}
//   phase 3: X[:, c] = B[:, qinv[c]]  (out of place)
// One lane per row of B for the diagonal blocks.
template <typename T, int NB>
__global__ __launch_bounds__(128) void k_rtrsm_upper(int m, int j0, int w, const cx<T>* A, long lda, long sA,
                                                     cx<T>* B, long ldb, long sB) {
    __shared__ cx<T> U[NB][NB + 1];
    const int b = blockIdx.y;
    A += (long)b * sA;
    B += (long)b * sB;
    for (int e = threadIdx.x; e < w * w; e += 128) {
        int r = e % w, c = e / w;
        U[r][c] = A[(j0 + r) + (long)(j0 + c) * lda];
    }
    __syncthreads();
    int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= m) return;
    cx<T> v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) v[c] = B[i + (long)(j0 + c) * ldb];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < w) {
            cx<T> s = v[c];
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (q < c) s = fms(s, v[q], U[q][c]);
            v[c] = cdiv(s, U[c][c]);
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) B[i + (long)(j0 + c) * ldb] = v[c];
}

template <typename T, int NB>
__global__ __launch_bounds__(128) void k_rtrsm_lower_unit(int m, int j0, int w, const cx<T>* A, long lda, long sA,
                                                          cx<T>* B, long ldb, long sB) {
    __shared__ cx<T> L[NB][NB + 1];
    const int b = blockIdx.y;
    A += (long)b * sA;
    B += (long)b * sB;
    for (int e = threadIdx.x; e < w * w; e += 128) {
        int r = e % w, c = e / w;
        L[r][c] = A[(j0 + r) + (long)(j0 + c) * lda];
    }
    __syncthreads();
    int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= m) return;
    cx<T> v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) v[c] = B[i + (long)(j0 + c) * ldb];
#pragma unroll
    for (int c = NB - 1; c >= 0; --c) {
        if (c < w) {
            cx<T> s = v[c];
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (q > c && q < w) s = fms(s, v[q], L[q][c]);
            v[c] = s;
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) B[i + (long)(j0 + c) * ldb] = v[c];
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_cols(int m, int n, const cx<T>* B, long ldb, long sB, cx<T>* X,
                                                     long ldx, long sX, const int* q, long sQ) {
    const int b = blockIdx.z;
    int i = blockIdx.x * 256 + threadIdx.x;
    int c = blockIdx.y;
    if (i >= m) return;
    int src = q[(long)b * sQ + c];
    X[(long)b * sX + i + (long)c * ldx] = B[(long)b * sB + i + (long)src * ldb];
}

template <typename T>
void rsolve_batched(int m, int n, int batch, const void* LU, long lda, long sA, const int* qinv, long sQ, void* B,
                    long ldb, long sB, void* X, long ldx, long sX, int nb, hipStream_t s, bool identity) {
    const cx<T>* a = (const cx<T>*)LU;
    cx<T>* bm = (cx<T>*)B;
    dim3 g((m + 127) / 128, batch);
    const int nbo = ctx().nbo;
    const int NB = (nbo > nb) ? (nbo / nb) * nb : nb;
    // identity: B is the n x n identity (the explicit inverse), so in phase 1 the column block
    // [c0, c1) is zero in every row from c1 on (Y = U^{-1} is upper triangular): its solve and
    // the updates it feeds touch rows < c1 only, a third of phase 1's work
    auto rows = [&](int c1) { return identity && c1 < m ? c1 : m; };
    auto upper = [&](int j0, int w) {
        const int mr = rows(j0 + w);
        const dim3 gu((mr + 127) / 128, batch);
        if (nb <= 32)
            hipLaunchKernelGGL((k_rtrsm_upper<T, 32>), gu, dim3(128), 0, s, mr, j0, w, a, lda, sA, bm, ldb, sB);
        else
            hipLaunchKernelGGL((k_rtrsm_upper<T, 64>), gu, dim3(128), 0, s, mr, j0, w, a, lda, sA, bm, ldb, sB);
    };
    auto lower = [&](int j0, int w) {
        if (nb <= 32)
            hipLaunchKernelGGL((k_rtrsm_lower_unit<T, 32>), g, dim3(128), 0, s, m, j0, w, a, lda, sA, bm, ldb, sB);
        else
            hipLaunchKernelGGL((k_rtrsm_lower_unit<T, 64>), g, dim3(128), 0, s, m, j0, w, a, lda, sA, bm, ldb, sB);
    };
    const bool rec = ctx().lu_recursive;
    // phase 1 inside an outer block, columns [c0, c1): left half, B[:, right] -= B[:, left]
    // U[left, right] (one GEMM of depth mid - c0), right half; the diagonal blocks by k_rtrsm
    auto upper_rec = [&](auto&& self, int c0, int c1) -> void {
        const int np = (c1 - c0 + nb - 1) / nb;
        if (np <= 1) {
            upper(c0, c1 - c0);
            return;
        }
        const int mid = c0 + ((np + 1) / 2) * nb;
        self(self, c0, mid);
        gemm_sub<T>(rows(mid), c1 - mid, mid - c0, bm + (long)c0 * ldb, ldb, sB, a + c0 + (long)mid * lda, lda, sA,
                    bm + (long)mid * ldb, ldb, sB, batch, s);
        self(self, mid, c1);
    };
    // phase 2 inside an outer block: right half, B[:, left] -= B[:, right] L[right, left], left half
    auto lower_rec = [&](auto&& self, int c0, int c1) -> void {
        const int np = (c1 - c0 + nb - 1) / nb;
        if (np <= 1) {
            lower(c0, c1 - c0);
            return;
        }
        const int mid = c0 + ((np + 1) / 2) * nb;
        self(self, mid, c1);
        gemm_sub<T>(m, mid - c0, c1 - mid, bm + (long)mid * ldb, ldb, sB, a + mid + (long)c0 * lda, lda, sA,
                    bm + (long)c0 * ldb, ldb, sB, batch, s);
        self(self, c0, mid);
    };
    // phase 1: Y U = B, outer blocks left to right
    for (int J0 = 0; J0 < n; J0 += NB) {
        const int JE = (n - J0 < NB) ? n : J0 + NB;
        if (rec)
            upper_rec(upper_rec, J0, JE);
        else
            for (int j0 = J0; j0 < JE; j0 += nb) {
                int w = (JE - j0 < nb) ? (JE - j0) : nb;
                upper(j0, w);
                if (j0 + w < JE)
                    gemm_sub<T>(rows(j0 + w), JE - j0 - w, w, bm + (long)j0 * ldb, ldb, sB,
                                a + j0 + (long)(j0 + w) * lda, lda, sA, bm + (long)(j0 + w) * ldb, ldb, sB, batch, s);
            }
        if (JE < n)
            gemm_sub<T>(rows(JE), n - JE, JE - J0, bm + (long)J0 * ldb, ldb, sB, a + J0 + (long)JE * lda, lda, sA,
                        bm + (long)JE * ldb, ldb, sB, batch, s);
    }
    // phase 2: Z L = Y, outer blocks right to left
    const int lastJ = ((n - 1) / NB) * NB;
    for (int J0 = lastJ; J0 >= 0; J0 -= NB) {
        const int JE = (n - J0 < NB) ? n : J0 + NB;
        const int lastj = J0 + ((JE - J0 - 1) / nb) * nb;
        if (rec)
            lower_rec(lower_rec, J0, JE);
        else
            for (int j0 = lastj; j0 >= J0; j0 -= nb) {
                int w = (JE - j0 < nb) ? (JE - j0) : nb;
                lower(j0, w);
                if (j0 > J0)
                    gemm_sub<T>(m, j0 - J0, w, bm + (long)j0 * ldb, ldb, sB, a + j0 + (long)J0 * lda, lda, sA,
                                bm + (long)J0 * ldb, ldb, sB, batch, s);
            }
        if (J0 > 0)
            gemm_sub<T>(m, J0, JE - J0, bm + (long)J0 * ldb, ldb, sB, a + J0, lda, sA, bm, ldb, sB, batch, s);
    }
    hipLaunchKernelGGL(k_gather_cols<T>, dim3((m + 255) / 256, n, batch), dim3(256), 0, s, m, n, bm, ldb, sB,
                       (cx<T>*)X, ldx, sX, qinv, sQ);
}

template void rsolve_batched<double>(int, int, int, const void*, long, long, const int*, long, void*, long, long,
                                     void*, long, long, int, hipStream_t, bool);
template void rsolve_batched<float>(int, int, int, const void*, long, long, const int*, long, void*, long, long,
                                    void*, long, long, int, hipStream_t, bool);

// ---------------------------------------------------------------------------------------------
// Left solve with a few right-hand sides: X = A^{-1} B = U^{-1} L^{-1} (B[pf]).
constexpr int TS_LOADS = 32;

// value of v in lane c (c uniform across the wave) through scalar register reads
__device__ __forceinline__ float lane_read(float v, int c) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), c));
}
__device__ __forceinline__ double lane_read(double v, int c) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), c);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), c);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ cx<T> lane_bcast(cx<T> v, int c) {
    return mk<T>(lane_read(v.re, c), lane_read(v.im, c));
}

// k_trsv: one workgroup (512 lanes) per matrix solves an n <= 256 triangle in 64-row blocks:
// the diagonal block is staged in LDS by all lanes, wave q solves right-hand side q with lane
// shuffles (lane = row), then every lane updates the rows below (above) with the block's
// solution held in LDS, issuing the block's column loads independently of the accumulation.
// getrs_batched walks the full triangle in 256-row blocks: k_trsv on the diagonal block, then
// the rows below (above) are updated by the multi-workgroup GEMV (gemv_sub_batched).
constexpr int TS_MAXR = 8;
constexpr int TS_THREADS = 512;

template <typename T, bool LOWER>
__global__ __launch_bounds__(TS_THREADS) void k_trsv(int n, int r, const cx<T>* A, long lda, long sA, cx<T>* X,
                                                     long ldx, long sX) {
    __shared__ cx<T> D[64][65];
    __shared__ cx<T> xs[64][TS_MAXR];
    __shared__ cx<T> dinv[64];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63;
    A += (long)b * sA;
    X += (long)b * sX;
    const int nblk = (n + 63) / 64;
    for (int s = 0; s < nblk; ++s) {
        const int blk = LOWER ? s : nblk - 1 - s;
        const int i0 = blk * 64;
        const int bs = (n - i0 < 64) ? (n - i0) : 64;
        for (int e = t; e < 64 * 64; e += TS_THREADS) {
            int rr = e & 63, cc = e >> 6;
            D[rr][cc] = (rr < bs && cc < bs) ? A[(i0 + rr) + (long)(i0 + cc) * lda] : mk<T>(0, 0);
        }
        __syncthreads();
        if (!LOWER && t < bs) dinv[t] = crecip(D[t][t]);
        __syncthreads();
        // one wave per right-hand side; the solved entry is broadcast with a register read of
        // lane c (no LDS round trip on the dependency chain), U's diagonal by its reciprocal
        const int wave = t >> 6;
        for (int q = wave; q < r; q += TS_THREADS / 64) {
            {
                cx<T> v = lane < bs ? X[i0 + lane + (long)q * ldx] : mk<T>(0, 0);
                if (LOWER) {
                    for (int c = 0; c < bs; ++c) {
                        const cx<T> xc = lane_bcast(v, c);
                        if (lane > c) v = fms(v, D[lane][c], xc);
                    }
                } else {
                    for (int c = bs - 1; c >= 0; --c) {
                        if (lane == c) v = v * dinv[c];
                        const cx<T> xc = lane_bcast(v, c);
                        if (lane < c) v = fms(v, D[lane][c], xc);
                    }
                }
                if (lane < bs) {
                    xs[lane][q] = v;
                    X[i0 + lane + (long)q * ldx] = v;
                }
            }
        }
        __syncthreads();
        const int lo = LOWER ? i0 + bs : 0;
        const int hi = LOWER ? n : i0;
        for (int i = lo + t; i < hi; i += TS_THREADS) {
            cx<T> acc[TS_MAXR];
#pragma unroll
            for (int q = 0; q < TS_MAXR; ++q)
                if (q < r) acc[q] = X[i + (long)q * ldx];
            const cx<T>* Ai = A + i + (long)i0 * lda;
            int c = 0;
            // 32 independent loads in flight per lane (the update is latency bound)
            for (; c + TS_LOADS <= bs; c += TS_LOADS) {
                cx<T> av[TS_LOADS];
#pragma unroll
                for (int u = 0; u < TS_LOADS; ++u) av[u] = Ai[(long)(c + u) * lda];
#pragma unroll
                for (int u = 0; u < TS_LOADS; ++u)
#pragma unroll
                    for (int q = 0; q < TS_MAXR; ++q)
                        if (q < r) acc[q] = fms(acc[q], av[u], xs[c + u][q]);
            }
            for (; c < bs; ++c) {
                cx<T> a = Ai[(long)c * lda];
#pragma unroll
                for (int q = 0; q < TS_MAXR; ++q)
                    if (q < r) acc[q] = fms(acc[q], a, xs[c][q]);
            }
#pragma unroll
            for (int q = 0; q < TS_MAXR; ++q)
                if (q < r) X[i + (long)q * ldx] = acc[q];
        }
        __syncthreads();
    }
}

template <typename T>
void gemv_sub_batched(int m, int k, int r, int batch, const void* A, long lda, long sA, const void* X, long ldx,
                      long sX, void* Y, long ldy, long sY, hipStream_t s);

constexpr int TS_BLOCK = 256;

template <typename T>
__global__ __launch_bounds__(256) void k_gather_rows(int n, int r, const cx<T>* B, long ldb, long sB, cx<T>* X,
                                                     long ldx, long sX, const int* pf, long sQ) {
    const int b = blockIdx.y;
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int src = pf[(long)b * sQ + i];
    for (int q = 0; q < r; ++q) X[(long)b * sX + i + (long)q * ldx] = B[(long)b * sB + src + (long)q * ldb];
}

// X = A^{-1} B for r <= 8 columns; B and X must not alias.
template <typename T>
void getrs_batched(int n, int r, int batch, const void* LU, long lda, long sA, const int* pf, long sQ, const void* B,
                   long ldb, long sB, void* X, long ldx, long sX, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows<T>, dim3((n + 255) / 256, batch), dim3(256), 0, s, n, r, (const cx<T>*)B, ldb,
                       sB, (cx<T>*)X, ldx, sX, pf, sQ);
    const cx<T>* a = (const cx<T>*)LU;
    cx<T>* x = (cx<T>*)X;
    for (int i0 = 0; i0 < n; i0 += TS_BLOCK) {
        int bs = (n - i0 < TS_BLOCK) ? (n - i0) : TS_BLOCK;
        hipLaunchKernelGGL((k_trsv<T, true>), dim3(batch), dim3(TS_THREADS), 0, s, bs, r, a + i0 + (long)i0 * lda, lda,
                           sA, x + i0, ldx, sX);
        int rest = n - i0 - bs;
        if (rest > 0)
            gemv_sub_batched<T>(rest, bs, r, batch, a + (i0 + bs) + (long)i0 * lda, lda, sA, x + i0, ldx, sX,
                                x + i0 + bs, ldx, sX, s);
    }
    int last = ((n - 1) / TS_BLOCK) * TS_BLOCK;
    for (int i0 = last; i0 >= 0; i0 -= TS_BLOCK) {
        int bs = (n - i0 < TS_BLOCK) ? (n - i0) : TS_BLOCK;
        hipLaunchKernelGGL((k_trsv<T, false>), dim3(batch), dim3(TS_THREADS), 0, s, bs, r, a + i0 + (long)i0 * lda,
                           lda, sA, x + i0, ldx, sX);
        if (i0 > 0)
            gemv_sub_batched<T>(i0, bs, r, batch, a + (long)i0 * lda, lda, sA, x + i0, ldx, sX, x, ldx, sX, s);
    }
}

template void getrs_batched<double>(int, int, int, const void*, long, long, const int*, long, const void*, long,
                                    long, void*, long, long, hipStream_t);
template void getrs_batched<float>(int, int, int, const void*, long, long, const int*, long, const void*, long,
                                   long, void*, long, long, hipStream_t);

}  // namespace mnp
