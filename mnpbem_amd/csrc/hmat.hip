// H-matrix operator of BEMRetIter on the device: on-the-fly matrix entries, dense blocks,
// partially pivoted ACA for admissible blocks, blocked matrix-vector products.
//
// Matrix kinds: 0 = G1, 1 = G2, 2 = H1, 3 = H2 (inside / outside Green function and surface
// derivative, H = F +- 2 pi on self terms). Entry (i, j) of kind q at one wavelength:
//   far:     G = a_j/d e^{ikd},  F = n_i.r (ik - 1/d)/d^2 a_j e^{ikd}
//   refined: G = sum_n g_n (ik)^n e^{ikd},  F = n_i . sum_n f_n (ik)^n e^{ikd}
// with k of the material connecting the two faces' regions (greenret_refined.py:654-810,
// compgreen_ret.py:739-858), the same values the dense assembly writes. Refined pairs are
// found by binary search in a CSR whose columns are sorted within each row.
//
// Blocks are index ranges [r0, r0+m) x [c0, c0+n) in cluster order; c2p maps a cluster
// position to its face. ACA follows greenfun/hmatrix.py:195-336: first pivot row 0; pivot
// column = largest residual modulus of the row among unused columns; rank-1 term u v^T with
// v = residual row, u = residual column / pivot; |S|_F^2 grows by |u|^2 |v|^2 plus
// 2 Re sum_q (U_q^H u)(V_q^H v); stop when |u||v| < htol |S|_F, at rank min(m, n, kmax) or when
// the pivot modulus is below 1e-15; next pivot row = largest |u| among unused rows.
#include "common.hpp"
#include "args.hpp"
#include "hmat.hpp"

namespace mnp {

constexpr double HM_TWO_PI = 6.283185307179586476925286766559;

__device__ cx<double> hm_entry(const GreenGeom& a, int kind, int i, int j) {
    int pa = a.fpart[i], pb = a.fpart[j];
    // signed material: -m is the subtracted cross-region term of nested particles
    const int c = (kind & 1) ? a.con_out[pa * a.npart + pb] : a.con_in[pa * a.npart + pb];
    const int m = c < 0 ? -c : c;
    if (m == 0) return mk(0.0, 0.0);
    cx<double> k = reinterpret_cast<const cx<double>*>(a.kmat)[m - 1];
    double rx = a.pos[3 * i] - a.pos[3 * j];
    double ry = a.pos[3 * i + 1] - a.pos[3 * j + 1];
    double rz = a.pos[3 * i + 2] - a.pos[3 * j + 2];
    double d = sqrt(rx * rx + ry * ry + rz * rz);
    const bool isG = kind < 2;
    int lo = a.rowptr[i], hi = a.rowptr[i + 1] - 1, p = -1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int c = a.rcol[mid];
        if (c == j) {
            p = mid;
            break;
        }
        if (c < j) lo = mid + 1;
        else hi = mid - 1;
    }
    cx<double> v;
    if (p >= 0) {
        const int no = a.order + 1;
        cx<double> ik = mk(-k.im, k.re);
        cx<double> s = mk(0.0, 0.0);
        for (int q = no - 1; q >= 0; --q) {
            double c;
            if (isG) c = a.rg[(long)p * no + q];
            else
                c = a.nvec[3 * i] * a.rf[((long)p * 3) * no + q] + a.nvec[3 * i + 1] * a.rf[((long)p * 3 + 1) * no + q] +
                    a.nvec[3 * i + 2] * a.rf[((long)p * 3 + 2) * no + q];
            s = ik * s + mk(c, 0.0);
        }
        v = s * expikd(k, d);
    } else if (i == j) {
        v = mk(0.0, 0.0);
    } else {
        cx<double> ph = expikd(k, d);
        double inv_d = 1.0 / d;
        double aj = a.area[j];
        if (isG) v = (aj * inv_d) * ph;
        else {
            double nr = a.nvec[3 * i] * rx + a.nvec[3 * i + 1] * ry + a.nvec[3 * i + 2] * rz;
            v = (nr * aj * inv_d * inv_d) * (mk(-k.im - inv_d, k.re) * ph);
        }
    }
    if (c < 0) v = -1.0 * v;
    if (!isG && i == j) v.re += (kind == 2) ? HM_TWO_PI : -HM_TWO_PI;
    return v;
}

// ---------------------------------------------------------------------------- dense blocks
// blk: nd x 4 ints (r0, c0, m, n); off: complex offset of each block (column-major m x n);
// kind = blockIdx.z, stored at D + kind * kstride.
__global__ __launch_bounds__(256) void k_hm_dense(GreenGeom a, const int* c2p, const int* blk, const long* off,
                                                  cx<double>* D, long kstride) {
    const int b = blockIdx.x, kind = blockIdx.z;
    const int r0 = blk[4 * b], c0 = blk[4 * b + 1], m = blk[4 * b + 2], n = blk[4 * b + 3];
    cx<double>* out = D + (long)kind * kstride + off[b];
    const long tot = (long)m * n;
    for (long e = (long)blockIdx.y * 256 + threadIdx.x; e < tot; e += (long)gridDim.y * 256) {
        int ii = (int)(e % m), jj = (int)(e / m);
        out[e] = hm_entry(a, kind, c2p[r0 + ii], c2p[c0 + jj]);
    }
}

// ---------------------------------------------------------------------------- ACA
constexpr int ACA_T = 256;
constexpr int ACA_W = ACA_T / 64;
constexpr int ACA_CH = 8;    // cross-term dot products per reduction

// Sum K values over the workgroup; every thread gets the totals. sh holds ACA_W * K doubles.
template <int K>
__device__ void block_sum(double (&v)[K], double* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[w * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int q = 0; q < ACA_W; ++q) s += sh[q * K + k];
        v[k] = s;
    }
}

// Largest value over the workgroup, smallest index on ties (first maximum, as numpy.argmax).
__device__ void block_argmax(double& v, int& idx, double* shv, int* shi) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    __syncthreads();
    if (lane == 0) {
        shv[w] = v;
        shi[w] = idx;
    }
    __syncthreads();
    v = shv[0];
    idx = shi[0];
    for (int q = 1; q < ACA_W; ++q)
        if (shv[q] > v || (shv[q] == v && shi[q] < idx)) {
            v = shv[q];
            idx = shi[q];
        }
}

// One workgroup per admissible block. U: m x kr, V: n x kr complex double (kr = min(m, n, kmax))
// at uoff[b], voff[b]; rank[4 b + kind] = number of terms kept.
// Confirmation row of the ACA stopping test: an unused row chosen from a fixed pseudo-random
// sequence of the block and step (the same sequence as oracle/aca_oracle.py), m if none is left.
__device__ inline int aca_fresh_row(const unsigned char* used_r, int m, int b, int r) {
    const unsigned long long h = (6364136223846793005ULL * (unsigned long long)(b * 131 + r + 1) + 1442695040888963407ULL);
    const int start = (int)((h >> 33) % (unsigned long long)m);
    for (int q = 0; q < m; ++q) {
        const int i = (start + q) % m;
        if (!used_r[i]) return i;
    }
    return m;
}

__global__ __launch_bounds__(ACA_T) void k_hm_aca(GreenGeom a, const int* c2p, int kind, const int* lr,
                                                  const long* uoff, const long* voff, cx<double>* U, cx<double>* V,
                                                  int* rank_out, double htol, int kmax) {
    __shared__ unsigned char used_r[HM_MAXDIM], used_c[HM_MAXDIM];
    __shared__ double shs[ACA_W * 4 * ACA_CH];
    __shared__ double shv[ACA_W];
    __shared__ int shi[ACA_W];
    const int b = blockIdx.x, t = threadIdx.x;
    const int r0 = lr[4 * b], c0 = lr[4 * b + 1], m = lr[4 * b + 2], n = lr[4 * b + 3];
    cx<double>* Ub = U + uoff[b];
    cx<double>* Vb = V + voff[b];
    for (int i = t; i < m; i += ACA_T) used_r[i] = 0;
    for (int j = t; j < n; j += ACA_T) used_c[j] = 0;
    __syncthreads();
    const int maxr = min(min(m, n), kmax);
    int pr = 0, rank = 0;
    bool confirming = false;
    double frob = 0;
    for (int r = 0; r < maxr; ++r) {
        // residual row pr -> V[:, r]
        const int fi = c2p[r0 + pr];
        double best = -1;
        int bj = n;
        for (int j = t; j < n; j += ACA_T) {
            cx<double> v = hm_entry(a, kind, fi, c2p[c0 + j]);
            for (int q = 0; q < r; ++q) v = fms(v, Ub[pr + (long)q * m], Vb[j + (long)q * n]);
            Vb[j + (long)r * n] = v;
            double av = used_c[j] ? -1.0 : sqrt(v.re * v.re + v.im * v.im);
            if (av > best) {   // j increases within a thread: strict > keeps the first maximum
                best = av;
                bj = j;
            }
        }
        block_argmax(best, bj, shv, shi);
        if (bj >= n || best < 1e-15) break;
        const int pj = bj;
        const cx<double> piv = Vb[pj + (long)r * n];
        const int fj = c2p[c0 + pj];
        // residual column pj / pivot -> U[:, r]; next pivot row candidates
        double acc[2] = {0, 0};   // |u|^2, |v|^2
        double bu = -1;
        int bi = m;
        for (int i = t; i < m; i += ACA_T) {
            cx<double> v = hm_entry(a, kind, c2p[r0 + i], fj);
            for (int q = 0; q < r; ++q) v = fms(v, Ub[i + (long)q * m], Vb[pj + (long)q * n]);
            v = cdiv(v, piv);
            Ub[i + (long)r * m] = v;
            double a2 = v.re * v.re + v.im * v.im;
            acc[0] += a2;
            if (!used_r[i] && i != pr) {
                double av = sqrt(a2);
                if (av > bu) {
                    bu = av;
                    bi = i;
                }
            }
        }
        for (int j = t; j < n; j += ACA_T) {
            cx<double> v = Vb[j + (long)r * n];
            acc[1] += v.re * v.re + v.im * v.im;
        }
        block_sum<2>(acc, shs);
        // cross terms, ACA_CH previous terms per reduction
        double cross = 0;
        for (int q0 = 0; q0 < r; q0 += ACA_CH) {
            double p[4 * ACA_CH];
#pragma unroll
            for (int s = 0; s < 4 * ACA_CH; ++s) p[s] = 0;
            for (int i = t; i < m; i += ACA_T) {
                cx<double> u = Ub[i + (long)r * m];
#pragma unroll
                for (int s = 0; s < ACA_CH; ++s)
                    if (q0 + s < r) {
                        cx<double> w = Ub[i + (long)(q0 + s) * m];
                        p[4 * s] += w.re * u.re + w.im * u.im;
                        p[4 * s + 1] += w.re * u.im - w.im * u.re;
                    }
            }
            for (int j = t; j < n; j += ACA_T) {
                cx<double> v = Vb[j + (long)r * n];
#pragma unroll
                for (int s = 0; s < ACA_CH; ++s)
                    if (q0 + s < r) {
                        cx<double> w = Vb[j + (long)(q0 + s) * n];
                        p[4 * s + 2] += w.re * v.re + w.im * v.im;
                        p[4 * s + 3] += w.re * v.im - w.im * v.re;
                    }
            }
            block_sum<4 * ACA_CH>(p, shs);
#pragma unroll
            for (int s = 0; s < ACA_CH; ++s)
                if (q0 + s < r) cross += 2.0 * (p[4 * s] * p[4 * s + 2] - p[4 * s + 1] * p[4 * s + 3]);
        }
        block_argmax(bu, bi, shv, shi);
        if (t == 0) {
            used_r[pr] = 1;
            used_c[pj] = 1;
        }
        __syncthreads();
        rank = r + 1;
        const double term = acc[0] * acc[1];
        frob += term + cross;
        if (frob > 0 && sqrt(term) < htol * sqrt(fabs(frob))) {
            // converged by the cross just taken: accept only if a fresh row, taken away from the
            // pivot sequence, confirms it (its cross is small too)
            if (confirming) break;
            confirming = true;
            if (t == 0) shi[0] = aca_fresh_row(used_r, m, b, r);
            __syncthreads();
            bi = shi[0];
            __syncthreads();
        } else {
            confirming = false;
        }
        if (bi >= m) break;
        pr = bi;
    }
    if (t == 0) rank_out[4 * b + kind] = rank;
}

// compact complex-float copy of the first `rank` columns of U and V (leading dims m and n)
__global__ __launch_bounds__(256) void k_hm_compact(const int* lr, int kind, const int* rank, const long* uws,
                                                    const long* vws, const cx<double>* Uw, const cx<double>* Vw,
                                                    const long* uoff, const long* voff, cx<float>* U, cx<float>* V) {
    const int b = blockIdx.x;
    const int rk = rank[4 * b + kind];
    for (int which = 0; which < 2; ++which) {
        const int rows = lr[4 * b + 2 + which];
        const cx<double>* s = which ? Vw + vws[b] : Uw + uws[b];
        cx<float>* d = which ? V + voff[b] : U + uoff[b];
        const long tot = (long)rows * rk;
        for (long e = (long)blockIdx.y * 256 + threadIdx.x; e < tot; e += (long)gridDim.y * 256) {
            cx<double> v = s[e];
            d[e] = mk((float)v.re, (float)v.im);
        }
    }
}

// ---------------------------------------------------------------------------- matvec
// All vectors in cluster order: x (n x r, leading dim ldx), y (n x r, ldy) accumulated with
// atomics (blocks sharing rows run concurrently). r <= HM_MAXR.

// dense: 64 rows per pass, the 4 waves split the columns, partial sums reduced in LDS
__global__ __launch_bounds__(256) void k_hm_mv_dense(const int* blk, const long* off, const cx<double>* D, int r,
                                                     const cx<double>* x, long ldx, double* y, long ldy) {
    __shared__ cx<double> red[3][64][HM_MAXR];
    const int b = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blk[4 * b], c0 = blk[4 * b + 1], m = blk[4 * b + 2], n = blk[4 * b + 3];
    const cx<double>* Db = D + off[b];
    for (int i0 = 0; i0 < m; i0 += 64) {
        const int i = i0 + lane;
        cx<double> acc[HM_MAXR];
#pragma unroll
        for (int c = 0; c < HM_MAXR; ++c) acc[c] = mk(0.0, 0.0);
        if (i < m)
            for (int j = w; j < n; j += 4) {
                cx<double> d = Db[i + (long)j * m];
#pragma unroll
                for (int c = 0; c < HM_MAXR; ++c)
                    if (c < r) {
                        cx<double> xv = x[c0 + j + (long)c * ldx];
                        acc[c].re += d.re * xv.re - d.im * xv.im;
                        acc[c].im += d.re * xv.im + d.im * xv.re;
                    }
            }
        if (w > 0)
#pragma unroll
            for (int c = 0; c < HM_MAXR; ++c) red[w - 1][lane][c] = acc[c];
        __syncthreads();
        if (w == 0 && i < m) {
#pragma unroll
            for (int c = 0; c < HM_MAXR; ++c)
                if (c < r) {
                    cx<double> s = acc[c] + red[0][lane][c] + red[1][lane][c] + red[2][lane][c];
                    long o = 2 * ((long)r0 + i + (long)c * ldy);
                    atomicAdd(&y[o], s.re);
                    atomicAdd(&y[o + 1], s.im);
                }
        }
        __syncthreads();
    }
}

// tmp[toff[b] + q HM_MAXR + c] = sum_j V[j, q] x[c0 + j, c]. The block's slice of x is staged
// through LDS in chunks of 256 rows; each wave accumulates 4 terms q at once (lanes over j) and
// reduces across lanes once per group of 16 terms.
constexpr int VT_Q = 4, VT_CH = 256;
__global__ __launch_bounds__(256) void k_hm_mv_vt(const int* lr, const long* voff, const int* rank, int kind,
                                                  const cx<float>* V, int r, const cx<double>* x, long ldx,
                                                  cx<double>* tmp, const long* toff) {
    __shared__ cx<double> xs[VT_CH][HM_MAXR];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int c0 = lr[4 * b + 1], n = lr[4 * b + 3];
    const int rk = rank[4 * b + kind];
    const cx<float>* Vb = V + voff[b];
    for (int g0 = 0; g0 < rk; g0 += 4 * VT_Q) {
        const int q0 = g0 + w * VT_Q;
        cx<double> acc[VT_Q][HM_MAXR];
#pragma unroll
        for (int a = 0; a < VT_Q; ++a)
#pragma unroll
            for (int c = 0; c < HM_MAXR; ++c) acc[a][c] = mk(0.0, 0.0);
        for (int j0 = 0; j0 < n; j0 += VT_CH) {
            const int cnt = min(VT_CH, n - j0);
            __syncthreads();
            for (int e = t; e < VT_CH * HM_MAXR; e += 256) {
                const int jj = e % VT_CH, c = e / VT_CH;
                if (jj < cnt) xs[jj][c] = (c < r) ? x[c0 + j0 + jj + (long)c * ldx] : mk(0.0, 0.0);
            }
            __syncthreads();
            for (int jj = lane; jj < cnt; jj += 64) {
                cx<float> v[VT_Q];
#pragma unroll
                for (int a = 0; a < VT_Q; ++a)
                    v[a] = (q0 + a < rk) ? Vb[j0 + jj + (long)(q0 + a) * n] : mk(0.0f, 0.0f);
#pragma unroll
                for (int c = 0; c < HM_MAXR; ++c) {
                    if (c < r) {
                        const cx<double> xv = xs[jj][c];
#pragma unroll
                        for (int a = 0; a < VT_Q; ++a) {
                            acc[a][c].re += v[a].re * xv.re - v[a].im * xv.im;
                            acc[a][c].im += v[a].re * xv.im + v[a].im * xv.re;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int a = 0; a < VT_Q; ++a)
#pragma unroll
            for (int c = 0; c < HM_MAXR; ++c) {
                double sr = acc[a][c].re, si = acc[a][c].im;
                for (int o = 32; o > 0; o >>= 1) {
                    sr += __shfl_down(sr, o, 64);
                    si += __shfl_down(si, o, 64);
                }
                if (lane == 0 && c < r && q0 + a < rk) tmp[toff[b] + (long)(q0 + a) * HM_MAXR + c] = mk(sr, si);
            }
    }
}

// y[r0 + i, c] += sum_q U[i, q] tmp[q, c]
__global__ __launch_bounds__(256) void k_hm_mv_u(const int* lr, const long* uoff, const int* rank, int kind,
                                                 const cx<float>* U, int r, const cx<double>* tmp, const long* toff,
                                                 double* y, long ldy) {
    const int b = blockIdx.x;
    const int r0 = lr[4 * b], m = lr[4 * b + 2];
    const int rk = rank[4 * b + kind];
    if (rk == 0) return;
    const cx<float>* Ub = U + uoff[b];
    const cx<double>* tb = tmp + toff[b];
    for (int i = blockIdx.y * 256 + threadIdx.x; i < m; i += gridDim.y * 256) {
        cx<double> acc[HM_MAXR];
#pragma unroll
        for (int c = 0; c < HM_MAXR; ++c) acc[c] = mk(0.0, 0.0);
        for (int q = 0; q < rk; ++q) {
            cx<float> u = Ub[i + (long)q * m];
#pragma unroll
            for (int c = 0; c < HM_MAXR; ++c)
                if (c < r) {
                    cx<double> tv = tb[(long)q * HM_MAXR + c];
                    acc[c].re += u.re * tv.re - u.im * tv.im;
                    acc[c].im += u.re * tv.im + u.im * tv.re;
                }
        }
#pragma unroll
        for (int c = 0; c < HM_MAXR; ++c)
            if (c < r) {
                long o = 2 * ((long)r0 + i + (long)c * ldy);
                atomicAdd(&y[o], acc[c].re);
                atomicAdd(&y[o + 1], acc[c].im);
            }
    }
}

// face order <-> cluster order for r columns
__global__ void k_hm_gather(int n, int r, const int* c2p, const cx<double>* src, long lds, cx<double>* dst,
                            long ldd) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int f = c2p[i];
    for (int c = 0; c < r; ++c) dst[i + (long)c * ldd] = src[f + (long)c * lds];
}
__global__ void k_hm_scatter(int n, int r, const int* c2p, const cx<double>* src, long lds, cx<double>* dst,
                             long ldd) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int f = c2p[i];
    for (int c = 0; c < r; ++c) dst[f + (long)c * ldd] = src[i + (long)c * lds];
}
__global__ void k_hm_zero(long count, cx<double>* z) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) z[i] = mk(0.0, 0.0);
}

// ---------------------------------------------------------------------------- surface fields
// G and the Cartesian derivative Gp of pair (i, j) in region `outside` (0 inside, 1 outside):
//   far:     G = a_j/d e^{ikd},  Gp = (x_i - x_j) (ik - 1/d)/d^2 a_j e^{ikd}
//   refined: G = sum g_n (ik)^n e^{ikd},  Gp = sum f_n (ik)^n e^{ikd}   (greenret_refined.py:812-844)
__device__ void gp_entry(const GreenGeom& a, int outside, int i, int j, cx<double>& G, cx<double> (&Gp)[3]) {
    G = mk(0.0, 0.0);
    Gp[0] = Gp[1] = Gp[2] = mk(0.0, 0.0);
    const int pa = a.fpart[i], pb = a.fpart[j];
    const int m = outside ? a.con_out[pa * a.npart + pb] : a.con_in[pa * a.npart + pb];
    if (m <= 0) return;
    const cx<double> k = reinterpret_cast<const cx<double>*>(a.kmat)[m - 1];
    const double rx = a.pos[3 * i] - a.pos[3 * j];
    const double ry = a.pos[3 * i + 1] - a.pos[3 * j + 1];
    const double rz = a.pos[3 * i + 2] - a.pos[3 * j + 2];
    const double d = sqrt(rx * rx + ry * ry + rz * rz);
    int lo = a.rowptr[i], hi = a.rowptr[i + 1] - 1, p = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = a.rcol[mid];
        if (c == j) {
            p = mid;
            break;
        }
        if (c < j) lo = mid + 1;
        else hi = mid - 1;
    }
    const cx<double> ph = expikd(k, d);
    if (p >= 0) {
        const int no = a.order + 1;
        const cx<double> ik = mk(-k.im, k.re);
        cx<double> sg = mk(0.0, 0.0), sf[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
        for (int q = no - 1; q >= 0; --q) {
            sg = ik * sg + mk(a.rg[(long)p * no + q], 0.0);
            for (int c = 0; c < 3; ++c) sf[c] = ik * sf[c] + mk(a.rf[((long)p * 3 + c) * no + q], 0.0);
        }
        G = sg * ph;
        for (int c = 0; c < 3; ++c) Gp[c] = sf[c] * ph;
    } else if (i != j) {
        const double inv_d = 1.0 / d, aj = a.area[j];
        G = (aj * inv_d) * ph;
        const cx<double> f = (aj * inv_d * inv_d) * (mk(-k.im - inv_d, k.re) * ph);
        Gp[0] = rx * f;
        Gp[1] = ry * f;
        Gp[2] = rz * f;
    }
}

// One wave per observer face i: potentials phi = G s, phip = H s, a = G h, ap = H h and fields
// e = i k0 G h - Hp s, h = sum_j Hp_ij x h_j of region `outside` (H = n_i.Gp +- 2 pi on the
// self term, Hp = Gp +- 2 pi n_i; + inside, - outside), CompGreenRet.potential / .field
// (compgreen_ret.py:912-1040). s: n complex, hv: n x 3 complex (row-major); out: n x 14 complex
// (phi, phip, a[3], ap[3], e[3], h[3]).
__global__ __launch_bounds__(256) void k_surface_eval(GreenGeom a, int outside, double k0, const cx<double>* sg,
                                                      const cx<double>* hv, cx<double>* out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= a.n) return;
    const double nx = a.nvec[3 * i], ny = a.nvec[3 * i + 1], nz = a.nvec[3 * i + 2];
    const double s2pi = outside ? -HM_TWO_PI : HM_TWO_PI;
    cx<double> acc[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) acc[q] = mk(0.0, 0.0);
    for (int j = lane; j < a.n; j += 64) {
        cx<double> G, Gp[3];
        gp_entry(a, outside, i, j, G, Gp);
        const int pa = a.fpart[i], pb = a.fpart[j];
        const int m = outside ? a.con_out[pa * a.npart + pb] : a.con_in[pa * a.npart + pb];
        cx<double> H = nx * Gp[0] + ny * Gp[1] + nz * Gp[2];
        cx<double> P[3] = {Gp[0], Gp[1], Gp[2]};
        if (i == j && m > 0) {
            H.re += s2pi;
            P[0].re += s2pi * nx;
            P[1].re += s2pi * ny;
            P[2].re += s2pi * nz;
        }
        const cx<double> s = sg[j];
        const cx<double> h[3] = {hv[3 * j], hv[3 * j + 1], hv[3 * j + 2]};
        acc[0] = acc[0] + G * s;
        acc[1] = acc[1] + H * s;
        for (int c = 0; c < 3; ++c) {
            acc[2 + c] = acc[2 + c] + G * h[c];
            acc[5 + c] = acc[5 + c] + H * h[c];
            acc[8 + c] = acc[8 + c] + P[c] * s;
        }
        acc[11] = acc[11] + (P[1] * h[2] - P[2] * h[1]);
        acc[12] = acc[12] + (P[2] * h[0] - P[0] * h[2]);
        acc[13] = acc[13] + (P[0] * h[1] - P[1] * h[0]);
    }
#pragma unroll
    for (int q = 0; q < 14; ++q)
        for (int o = 32; o > 0; o >>= 1) {
            acc[q].re += __shfl_down(acc[q].re, o, 64);
            acc[q].im += __shfl_down(acc[q].im, o, 64);
        }
    if (lane == 0) {
        cx<double>* o = out + 14L * i;
        for (int q = 0; q < 8; ++q) o[q] = acc[q];
        const cx<double> ik0 = mk(0.0, k0);
        for (int c = 0; c < 3; ++c) o[8 + c] = ik0 * acc[2 + c] - acc[8 + c];
        for (int c = 0; c < 3; ++c) o[11 + c] = acc[11 + c];
    }
}

void surface_eval(const GreenGeom& a, int outside, double k0, const void* s, const void* hv, void* out,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_surface_eval, dim3((a.n + 3) / 4), dim3(256), 0, st, a, outside, k0, (const cx<double>*)s,
                       (const cx<double>*)hv, (cx<double>*)out);
}

// ---------------------------------------------------------------------------- fields at points
// One wave per observer point i in medium m: lanes over the faces j that touch m (inside
// medium m: sources sig1, h1; outside medium m: sig2, h2), G and Gp as for face pairs (refined
// Taylor pairs from the CSR over the points, far pairs in closed form). out: 14 complex per
// point as k_surface_eval (phi, 0, a, 0, e = i k0 a - Gp sigma, h = Gp x h).
__global__ __launch_bounds__(256) void k_point_eval(PointGeom a, double k0, const cx<double>* s1, const cx<double>* h1,
                                                    const cx<double>* s2, const cx<double>* h2, cx<double>* out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= a.npt) return;
    const int m = a.pmed[i];
    cx<double> acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = mk(0.0, 0.0);
    if (m > 0) {
        const cx<double> k = reinterpret_cast<const cx<double>*>(a.kmat)[m - 1];
        const cx<double> ik = mk(-k.im, k.re);
        const double px = a.ppos[3 * i], py = a.ppos[3 * i + 1], pz = a.ppos[3 * i + 2];
        const int r0 = a.rowptr[i], r1 = a.rowptr[i + 1];
        const int no = a.order + 1;
        for (int j = lane; j < a.nface; j += 64) {
            const int pj = a.fpart[j];
            const bool in = a.inout[2 * pj] == m, out_ = a.inout[2 * pj + 1] == m;
            if (!in && !out_) continue;
            const cx<double> s = in ? s1[j] : s2[j];
            const cx<double>* hv = in ? h1 + 3 * j : h2 + 3 * j;
            const double rx = px - a.fpos[3 * j], ry = py - a.fpos[3 * j + 1], rz = pz - a.fpos[3 * j + 2];
            const double d = sqrt(rx * rx + ry * ry + rz * rz);
            int lo = r0, hi = r1 - 1, p = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int c = a.rcol[mid];
                if (c == j) {
                    p = mid;
                    break;
                }
                if (c < j) lo = mid + 1;
                else hi = mid - 1;
            }
            const cx<double> ph = expikd(k, d);
            cx<double> G, Gp[3];
            if (p >= 0) {
                cx<double> sg = mk(0.0, 0.0), sf[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
                for (int q = no - 1; q >= 0; --q) {
                    sg = ik * sg + mk(a.rg[(long)p * no + q], 0.0);
                    for (int c = 0; c < 3; ++c) sf[c] = ik * sf[c] + mk(a.rf[((long)p * 3 + c) * no + q], 0.0);
                }
                G = sg * ph;
                for (int c = 0; c < 3; ++c) Gp[c] = sf[c] * ph;
            } else if (d > 0) {
                const double inv_d = 1.0 / d, aj = a.farea[j];
                G = (aj * inv_d) * ph;
                const cx<double> f = (aj * inv_d * inv_d) * (mk(-k.im - inv_d, k.re) * ph);
                Gp[0] = rx * f;
                Gp[1] = ry * f;
                Gp[2] = rz * f;
            } else {
                continue;      // a point on a centroid: excluded (ComPoint's mindist)
            }
            const cx<double> h[3] = {hv[0], hv[1], hv[2]};
            acc[0] = acc[0] + G * s;
            for (int c = 0; c < 3; ++c) {
                acc[1 + c] = acc[1 + c] + G * h[c];
                acc[4 + c] = acc[4 + c] + Gp[c] * s;
            }
            acc[7] = acc[7] + (Gp[1] * h[2] - Gp[2] * h[1]);
            acc[8] = acc[8] + (Gp[2] * h[0] - Gp[0] * h[2]);
            acc[9] = acc[9] + (Gp[0] * h[1] - Gp[1] * h[0]);
        }
    }
#pragma unroll
    for (int q = 0; q < 10; ++q)
        for (int o = 32; o > 0; o >>= 1) {
            acc[q].re += __shfl_down(acc[q].re, o, 64);
            acc[q].im += __shfl_down(acc[q].im, o, 64);
        }
    if (lane == 0) {
        cx<double>* o = out + 14L * i;
        const cx<double> zero = mk(0.0, 0.0), ik0 = mk(0.0, k0);
        o[0] = acc[0];
        o[1] = zero;
        for (int c = 0; c < 3; ++c) {
            o[2 + c] = acc[1 + c];
            o[5 + c] = zero;
            o[8 + c] = ik0 * acc[1 + c] - acc[4 + c];
            o[11 + c] = acc[7 + c];
        }
    }
}

void point_eval(const PointGeom& a, double k0, const void* s1, const void* h1, const void* s2, const void* h2,
                void* out, hipStream_t st) {
    if (a.npt <= 0) return;
    hipLaunchKernelGGL(k_point_eval, dim3((a.npt + 3) / 4), dim3(256), 0, st, a, k0, (const cx<double>*)s1,
                       (const cx<double>*)h1, (const cx<double>*)s2, (const cx<double>*)h2, (cx<double>*)out);
}

// ---------------------------------------------------------------------------- entries by key
// Key of one region-pair entry: 0 G, 1 F, 2 H1 = F + 2 pi, 3 H2 = F - 2 pi (self terms), 4 Gp,
// 5 H1p = Gp + 2 pi n_i, 6 H2p = Gp - 2 pi n_i (greenret_refined.py:654-866). Writes 1 value
// (keys 0-3) or 3 values at stride `cstride` (keys 4-6).
__device__ void key_entry(const GreenGeom& a, int outside, int key, int i, int j, cx<double>* o, long cstride) {
    cx<double> G, Gp[3];
    gp_entry(a, outside, i, j, G, Gp);
    double self = 0.0;
    if (i == j && (key == 2 || key == 3 || key == 5 || key == 6)) {
        const int pa = a.fpart[i];
        const int m = outside ? a.con_out[pa * a.npart + pa] : a.con_in[pa * a.npart + pa];
        if (m > 0) self = (key == 2 || key == 5) ? HM_TWO_PI : -HM_TWO_PI;
    }
    const double nx = a.nvec[3 * i], ny = a.nvec[3 * i + 1], nz = a.nvec[3 * i + 2];
    if (key == 0) {
        o[0] = G;
    } else if (key < 4) {
        cx<double> F = nx * Gp[0] + ny * Gp[1] + nz * Gp[2];
        F.re += self;
        o[0] = F;
    } else {
        Gp[0].re += self * nx;
        Gp[1].re += self * ny;
        Gp[2].re += self * nz;
        o[0] = Gp[0];
        o[cstride] = Gp[1];
        o[2 * cstride] = Gp[2];
    }
}

// Columns [c0, c0 + ncol) of the n-row matrix of `key`, row-major: (n, ncol) for keys 0-3,
// (n, 3, ncol) for keys 4-6 (CompGreenRet.eval / eval_block, compgreen_ret.py:507-737).
__global__ __launch_bounds__(256) void k_green_block(GreenGeom a, int outside, int key, int c0, int ncol,
                                                     cx<double>* out) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)a.n * ncol) return;
    const int i = (int)(e / ncol), jj = (int)(e % ncol);
    if (key < 4) key_entry(a, outside, key, i, c0 + jj, out + e, 0);
    else key_entry(a, outside, key, i, c0 + jj, out + (long)i * 3 * ncol + jj, ncol);
}

// Entries at row-major linear indices ind[q] = i * n + j: out (count) or (count, 3).
__global__ __launch_bounds__(256) void k_green_pick(GreenGeom a, int outside, int key, const long* ind, long count,
                                                    cx<double>* out) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= count) return;
    const long e = ind[q];
    if (e < 0 || e >= (long)a.n * a.n) {   // out of range: NaN, never an access outside the geometry
        const cx<double> nan = mk(__builtin_nan(""), __builtin_nan(""));
        for (int c = 0; c < (key < 4 ? 1 : 3); ++c) out[(key < 4 ? 1 : 3) * q + c] = nan;
        return;
    }
    const int i = (int)(e / a.n), j = (int)(e % a.n);
    if (key < 4) key_entry(a, outside, key, i, j, out + q, 0);
    else key_entry(a, outside, key, i, j, out + 3 * q, 1);
}

void green_block(const GreenGeom& a, int outside, int key, int c0, int ncol, void* out, hipStream_t st) {
    const long tot = (long)a.n * ncol;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_green_block, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a, outside, key, c0, ncol,
                       (cx<double>*)out);
}

void green_pick(const GreenGeom& a, int outside, int key, const long* ind, long count, void* out, hipStream_t st) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_green_pick, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, a, outside, key, ind,
                       count, (cx<double>*)out);
}

// ---------------------------------------------------------------------------- launchers
void hm_fill_dense(const GreenGeom& a, const int* c2p, int nd, const int* blk, const long* off, void* D, long kstride,
                   hipStream_t s) {
    if (nd <= 0) return;
    hipLaunchKernelGGL(k_hm_dense, dim3(nd, 4, 4), dim3(256), 0, s, a, c2p, blk, off, (cx<double>*)D, kstride);
}

void hm_aca(const GreenGeom& a, const int* c2p, int kind, int nlr, const int* lr, const long* uoff,
            const long* voff, void* U, void* V, int* rank, double htol, int kmax, hipStream_t s) {
    if (nlr <= 0) return;
    hipLaunchKernelGGL(k_hm_aca, dim3(nlr), dim3(ACA_T), 0, s, a, c2p, kind, lr, uoff, voff, (cx<double>*)U,
                       (cx<double>*)V, rank, htol, kmax);
}

void hm_compact(int nlr, const int* lr, int kind, const int* rank, const long* uws, const long* vws, const void* Uw,
                const void* Vw, const long* uoff, const long* voff, void* U, void* V, hipStream_t s) {
    if (nlr <= 0) return;
    hipLaunchKernelGGL(k_hm_compact, dim3(nlr, 8), dim3(256), 0, s, lr, kind, rank, uws, vws, (const cx<double>*)Uw,
                       (const cx<double>*)Vw, uoff, voff, (cx<float>*)U, (cx<float>*)V);
}

void hm_matvec(const HMat& h, int kind, int r, const void* x, long ldx, void* y, long ldy, hipStream_t s) {
    if (h.nd > 0)
        hipLaunchKernelGGL(k_hm_mv_dense, dim3(h.nd), dim3(256), 0, s, h.dblk, h.doff,
                           (const cx<double>*)h.D + (long)kind * h.dstride, r, (const cx<double>*)x, ldx, (double*)y,
                           ldy);
    if (h.nlr > 0) {
        const long* uo = h.uoff + (long)kind * h.nlr;
        const long* vo = h.voff + (long)kind * h.nlr;
        hipLaunchKernelGGL(k_hm_mv_vt, dim3(h.nlr), dim3(256), 0, s, h.lblk, vo, h.rank, kind, (const cx<float>*)h.V,
                           r, (const cx<double>*)x, ldx, (cx<double>*)h.tmp, h.toff);
        hipLaunchKernelGGL(k_hm_mv_u, dim3(h.nlr, 4), dim3(256), 0, s, h.lblk, uo, h.rank, kind, (const cx<float>*)h.U,
                           r, (const cx<double>*)h.tmp, h.toff, (double*)y, ldy);
    }
}

// The 8N operator of BEMRetIter with H-matrices (batch 1). work: 64 N complex double.
void bem_operator_hmat(const HMat& h, const Face& f, const void* x, const void* rhs, void* y, void* work,
                       double* norms, hipStream_t s) {
    const int n = f.n;
    const long N = n;
    cx<double>* w = (cx<double>*)work;
    cx<double>*X1 = w, *X2 = w + 8 * N, *Xc1 = w + 16 * N, *Xc2 = w + 24 * N, *Yc = w + 32 * N;
    cx<double>*GX1 = w + 40 * N, *GX2 = w + 48 * N, *HX1 = w + 56 * N, *HX2 = X1;
    const dim3 g1((n + 255) / 256);
    op_prep(f, 1, x, 8 * N, X1, X2, 8 * N, s);
    hipLaunchKernelGGL(k_hm_gather, g1, dim3(256), 0, s, n, 8, h.c2p, X1, N, Xc1, N);
    hipLaunchKernelGGL(k_hm_gather, g1, dim3(256), 0, s, n, 8, h.c2p, X2, N, Xc2, N);
    cx<double>* outs[4] = {GX1, GX2, HX1, HX2};
    for (int kind = 0; kind < 4; ++kind) {
        hipLaunchKernelGGL(k_hm_zero, dim3((8 * N + 255) / 256), dim3(256), 0, s, 8 * N, Yc);
        hm_matvec(h, kind, 8, (kind & 1) ? Xc2 : Xc1, N, Yc, N, s);
        hipLaunchKernelGGL(k_hm_scatter, g1, dim3(256), 0, s, n, 8, h.c2p, Yc, N, outs[kind], N);
    }
    op_finish(f, 1, GX1, GX2, HX1, 8 * N, HX2, 8 * N, rhs, 8 * N, y, 8 * N, norms, s);
}

}  // namespace mnp
