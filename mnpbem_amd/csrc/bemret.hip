// BEMRet chain on the device (Garcia de Abajo & Howie, PRB 65, 115418 (2002)).
//
// Factor (reference bem/bem_ret.py:438-523, GPU ordering :719-896):
//   LU(G1), LU(G2);  Sigma1 = H1 G1^{-1}, Sigma2 = H2 G2^{-1}  (right-side solves, no explicit G^{-1})
//   Delta = Sigma1 - Sigma2;  LU(Delta);  Dinv = Delta^{-1} (explicit: it enters a Hadamard product)
//   Sigma = Sigma1 L1 - Sigma2 L2 + k^2 L (Dinv o n n^T) L,  L = L1 - L2;  LU(Sigma)
// with L1, L2 the per-face inside/outside dielectric functions (diagonal matrices; a scalar
// when uniform), k = 2 pi / lambda.
//
// Solve (bem_ret.py:1708-1752), for phi, a, alpha, De packed as N x 8 columns [phi a De alpha]:
//   alpha' = alpha - Sigma1 a + ik n (L1 phi);  De' = De - Sigma1 L1 phi + ik n.(L1 a)
//   sig2 = Sigma^{-1}(De' + ik n.(L Dinv alpha'));  h2 = Dinv(ik n (L sig2) + alpha')
//   sig1 = G1^{-1}(sig2 + phi), h1 = G1^{-1}(h2 + a), sig2 = G2^{-1} sig2, h2 = G2^{-1} h2
// returned as N x 8 columns [sig1 h1 sig2 h2] (the packing of bem/bem_ret_iter.py:1196-1232).
//
// Exact operator (bem_ret_iter.py:1362-1501), used for FP64 iterative refinement:
//   phi = G1 s1 - G2 s2;  a = G1 h1 - G2 h2
//   De = H1(L1 s1) - H2(L2 s2) - ik n.(G1(L1 h1) - G2(L2 h2))
//   alpha = H1 h1 - H2 h2 - ik n (G1(L1 s1) - G2(L2 s2))
#include "common.hpp"
#include "args.hpp"

namespace mnp {

template <typename T>
void gemm_sub(int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb, long sB, void* C,
              long ldc, long sC, int batch, hipStream_t s);

// ----------------------------------------------------------------------------- elementwise
template <typename T>
__global__ __launch_bounds__(256) void k_sub(int n, const cx<T>* A, const cx<T>* B, cx<T>* C, long ld, long stride) {
    int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, b = blockIdx.z;
    if (i >= n) return;
    long o = (long)b * stride + i + (long)j * ld;
    C[o] = A[o] - B[o];
}

// Sigma (into S2) from S1, S2, Dinv
template <typename T>
__global__ __launch_bounds__(256) void k_sigma(int n, const cx<T>* S1, cx<T>* S2, const cx<T>* Dinv, long ld,
                                               long stride, const double* nvec, long sface, const double* L1,
                                               const double* L2, const double* kvac) {
    int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, b = blockIdx.z;
    if (i >= n) return;
    long o = (long)b * stride + i + (long)j * ld;
    nvec += (long)b * 3 * sface;
    const cx<double>* l1 = reinterpret_cast<const cx<double>*>(L1) + (long)b * n;
    const cx<double>* l2 = reinterpret_cast<const cx<double>*>(L2) + (long)b * n;
    double k = kvac[b];
    double nn = nvec[3 * i] * nvec[3 * j] + nvec[3 * i + 1] * nvec[3 * j + 1] + nvec[3 * i + 2] * nvec[3 * j + 2];
    cx<double> Li = l1[i] - l2[i], Lj = l1[j] - l2[j];
    cx<double> s1 = mk((double)S1[o].re, (double)S1[o].im);
    cx<double> s2 = mk((double)S2[o].re, (double)S2[o].im);
    cx<double> di = mk((double)Dinv[o].re, (double)Dinv[o].im);
    cx<double> v = s1 * l1[j] - s2 * l2[j] + (k * k * nn) * (Li * di * Lj);
    S2[o] = mk((T)v.re, (T)v.im);
}

// M_b = s A + diag(d_b) for every wavelength b (A one matrix, d: batch x n complex double).
// Quasistatic BEM matrix -(F + diag Lambda) (bem/bem_stat.py:330-342) from H1 = F + 2 pi.
template <typename T>
__global__ __launch_bounds__(256) void k_scale_diag(int n, double s, const cx<T>* A, long lda, const double* d,
                                                    cx<T>* M, long ld, long stride) {
    int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, b = blockIdx.z;
    if (i >= n) return;
    cx<T> a = A[i + (long)j * lda];
    cx<double> v = mk(s * (double)a.re, s * (double)a.im);
    if (i == j) v = v + reinterpret_cast<const cx<double>*>(d)[(long)b * n + i];
    M[(long)b * stride + i + (long)j * ld] = mk((T)v.re, (T)v.im);
}

template <typename T>
void scale_diag_batched(int n, int batch, double s, const void* A, long lda, const double* d, void* M, long ld,
                        long stride, hipStream_t st) {
    hipLaunchKernelGGL(k_scale_diag<T>, dim3((n + 255) / 256, n, batch), dim3(256), 0, st, n, s, (const cx<T>*)A, lda,
                       d, (cx<T>*)M, ld, stride);
}
template void scale_diag_batched<double>(int, int, double, const void*, long, const double*, void*, long, long,
                                         hipStream_t);
template void scale_diag_batched<float>(int, int, double, const void*, long, const double*, void*, long, long,
                                        hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_identity(int n, cx<T>* A, long ld, long stride) {
    int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, b = blockIdx.z;
    if (i >= n) return;
    A[(long)b * stride + i + (long)j * ld] = mk<T>(i == j ? 1 : 0, 0);
}

// ----------------------------------------------------------------------------- GEMV (r columns)
// Y[:, 0:r] = A X[:, 0:r]. A workgroup owns 64 rows; its four waves take every fourth column
// (lane = row, so each column read is a contiguous 64-element segment) with loads issued
// four at a time, X staged through LDS in 256-column chunks; partial sums of the waves are
// combined in LDS at the end.
constexpr int GV_MAXR = 8;
// SUB = false: Y = A X (m x k times k x r); SUB = true: Y -= A X.
template <typename T, bool SUB>
__global__ __launch_bounds__(256) void k_gemv(int m, int n, int r, const cx<T>* A, long lda, long sA, const cx<T>* X,
                                              long ldx, long sX, cx<T>* Y, long ldy, long sY) {
    __shared__ cx<T> xs[256][GV_MAXR];
    const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int i = blockIdx.x * 64 + lane;
    const bool row_ok = i < m;
    A += (long)b * sA;
    X += (long)b * sX;
    Y += (long)b * sY;
    cx<T> acc[GV_MAXR];
#pragma unroll
    for (int q = 0; q < GV_MAXR; ++q) acc[q] = mk<T>(0, 0);
    const cx<T>* Ai = A + (row_ok ? i : 0);
    for (int j0 = 0; j0 < n; j0 += 256) {
        int w = (n - j0 < 256) ? (n - j0) : 256;
        __syncthreads();
        if (t < w) {
#pragma unroll
            for (int q = 0; q < GV_MAXR; ++q)
                if (q < r) xs[t][q] = X[j0 + t + (long)q * ldx];
        }
        __syncthreads();
        if (row_ok) {
            int c = wave;
            for (; c + 12 < w; c += 16) {
                cx<T> a4[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) a4[u] = Ai[(long)(j0 + c + 4 * u) * lda];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < GV_MAXR; ++q)
                        if (q < r) {
                            cx<T> x = xs[c + 4 * u][q];
                            acc[q].re += a4[u].re * x.re - a4[u].im * x.im;
                            acc[q].im += a4[u].re * x.im + a4[u].im * x.re;
                        }
            }
            for (; c < w; c += 4) {
                cx<T> a = Ai[(long)(j0 + c) * lda];
#pragma unroll
                for (int q = 0; q < GV_MAXR; ++q)
                    if (q < r) {
                        cx<T> x = xs[c][q];
                        acc[q].re += a.re * x.re - a.im * x.im;
                        acc[q].im += a.re * x.im + a.im * x.re;
                    }
            }
        }
    }
    __syncthreads();
    // reduce the four waves' partial sums: xs reused as [wave][lane] x [q]
    cx<T>(*red)[GV_MAXR] = xs;
#pragma unroll
    for (int q = 0; q < GV_MAXR; ++q)
        if (q < r) red[wave * 64 + lane][q] = acc[q];
    __syncthreads();
    if (wave == 0 && row_ok) {
#pragma unroll
        for (int q = 0; q < GV_MAXR; ++q)
            if (q < r) {
                cx<T> v = red[lane][q];
                v = v + red[64 + lane][q];
                v = v + red[128 + lane][q];
                v = v + red[192 + lane][q];
                if (SUB) v = Y[i + (long)q * ldy] - v;
                Y[i + (long)q * ldy] = v;
            }
    }
}

template <typename T>
void gemv_batched(int n, int r, int batch, const void* A, long lda, long sA, const void* X, long ldx, long sX,
                  void* Y, long ldy, long sY, hipStream_t s) {
    hipLaunchKernelGGL((k_gemv<T, false>), dim3((n + 63) / 64, batch), dim3(256), 0, s, n, n, r, (const cx<T>*)A, lda,
                       sA, (const cx<T>*)X, ldx, sX, (cx<T>*)Y, ldy, sY);
}

// Y -= A X with A m x k (the off-diagonal updates of the blocked triangular solves).
template <typename T>
void gemv_sub_batched(int m, int k, int r, int batch, const void* A, long lda, long sA, const void* X, long ldx,
                      long sX, void* Y, long ldy, long sY, hipStream_t s) {
    if (m <= 0 || k <= 0) return;
    hipLaunchKernelGGL((k_gemv<T, true>), dim3((m + 63) / 64, batch), dim3(256), 0, s, m, k, r, (const cx<T>*)A, lda,
                       sA, (const cx<T>*)X, ldx, sX, (cx<T>*)Y, ldy, sY);
}
template void gemv_sub_batched<double>(int, int, int, int, const void*, long, long, const void*, long, long, void*,
                                       long, long, hipStream_t);
template void gemv_sub_batched<float>(int, int, int, int, const void*, long, long, const void*, long, long, void*,
                                      long, long, hipStream_t);

// ----------------------------------------------------------------------------- solve glue
// Every "vector block" is N x c complex, column-major with leading dimension n, stride per
// wavelength given in elements. Per-face L1/L2 are complex double (n per wavelength).

template <typename T>
__device__ __forceinline__ cx<T> Lval(const double* L, int b, int n, int i) {
    const cx<double>* l = reinterpret_cast<const cx<double>*>(L) + (long)b * n + i;
    return mk<T>((T)l->re, (T)l->im);
}

// V = [a, L1 phi]  (N x 4)  from IN = [phi a De alpha]
template <typename T>
__global__ void k_glue1(Face f, const cx<T>* IN, long sIn, cx<T>* V, long sV) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<T>* in = IN + (long)b * sIn;
    cx<T>* v = V + (long)b * sV;
    cx<T> l1 = Lval<T>(f.L1, b, n, i);
    for (int c = 0; c < 3; ++c) v[i + (long)c * n] = in[i + (long)(1 + c) * n];
    v[i + 3L * n] = l1 * in[i];
}

// alpha' (N x 3) and De' from IN and W = Sigma1 V; writes AP = [alpha'(3), De'(1)]
template <typename T>
__global__ void k_glue2(Face f, const cx<T>* IN, long sIn, const cx<T>* W, long sW, cx<T>* AP, long sAP) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<T>* in = IN + (long)b * sIn;
    const cx<T>* w = W + (long)b * sW;
    cx<T>* ap = AP + (long)b * sAP;
    T k = (T)f.kvac[b];
    cx<T> l1 = Lval<T>(f.L1, b, n, i);
    cx<T> l1phi = l1 * in[i];
    cx<T> ik = mk<T>(0, k);
    cx<T> dot = mk<T>(0, 0);
    for (int c = 0; c < 3; ++c) {
        T nc = (T)f.nv(b)[3 * i + c];
        cx<T> al = in[i + (long)(5 + c) * n] - w[i + (long)c * n] + (nc * ik) * l1phi;
        ap[i + (long)c * n] = al;
        dot = dot + nc * (l1 * in[i + (long)(1 + c) * n]);
    }
    ap[i + 3L * n] = in[i + 4L * n] - w[i + 3L * n] + ik * dot;
}

// RHS for Sigma: De' + ik n.(L U), U = Dinv alpha' (N x 3)
template <typename T>
__global__ void k_glue3(Face f, const cx<T>* AP, long sAP, const cx<T>* U, long sU, cx<T>* R, long sR) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<T>* ap = AP + (long)b * sAP;
    const cx<T>* u = U + (long)b * sU;
    cx<T> L = Lval<T>(f.L1, b, n, i) - Lval<T>(f.L2, b, n, i);
    T k = (T)f.kvac[b];
    cx<T> dot = mk<T>(0, 0);
    for (int c = 0; c < 3; ++c) dot = dot + (T)f.nv(b)[3 * i + c] * u[i + (long)c * n];
    R[(long)b * sR + i] = ap[i + 3L * n] + mk<T>(0, k) * (L * dot);
}

// t = ik n (L sig2) + alpha'  (N x 3)
template <typename T>
__global__ void k_glue4(Face f, const cx<T>* AP, long sAP, const cx<T>* S, long sS, cx<T>* Tt, long sT) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<T>* ap = AP + (long)b * sAP;
    cx<T> L = Lval<T>(f.L1, b, n, i) - Lval<T>(f.L2, b, n, i);
    T k = (T)f.kvac[b];
    cx<T> ls = mk<T>(0, k) * (L * S[(long)b * sS + i]);
    for (int c = 0; c < 3; ++c) Tt[(long)b * sT + i + (long)c * n] = ((T)f.nv(b)[3 * i + c]) * ls + ap[i + (long)c * n];
}

// P1 = [sig2 + phi, h2 + a], P2 = [sig2, h2]  (N x 4 each)
template <typename T>
__global__ void k_glue5(Face f, const cx<T>* IN, long sIn, const cx<T>* S, long sS, const cx<T>* H, long sH,
                        cx<T>* P1, cx<T>* P2, long sP) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<T>* in = IN + (long)b * sIn;
    cx<T> s = S[(long)b * sS + i];
    P1[(long)b * sP + i] = s + in[i];
    P2[(long)b * sP + i] = s;
    for (int c = 0; c < 3; ++c) {
        cx<T> h = H[(long)b * sH + i + (long)c * n];
        P1[(long)b * sP + i + (long)(1 + c) * n] = h + in[i + (long)(1 + c) * n];
        P2[(long)b * sP + i + (long)(1 + c) * n] = h;
    }
}

// ----------------------------------------------------------------------------- operator glue
// X1 = [s1, h1, L1 s1, L1 h1], X2 likewise from x = [s1 h1 s2 h2] (N x 8 each)
__global__ void k_op_prep(Face f, const cx<double>* x, long sx, cx<double>* X1, cx<double>* X2, long sX) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    if (i >= n) return;
    const cx<double>* xv = x + (long)b * sx;
    cx<double> l[2] = {Lval<double>(f.L1, b, n, i), Lval<double>(f.L2, b, n, i)};
    cx<double>* out[2] = {X1 + (long)b * sX, X2 + (long)b * sX};
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 4; ++c) {
            cx<double> v = xv[i + (long)(4 * s + c) * n];
            out[s][i + (long)c * n] = v;
            out[s][i + (long)(4 + c) * n] = l[s] * v;
        }
}

// y = [phi a De alpha] from GX1 = G1 X1, GX2 = G2 X2, HX1 = H1 X1, HX2 = H2 X2; if b != null,
// y = b - y (residual) and per-wavelength partial sums of |r|^2 and |b|^2 are accumulated.
__global__ void k_op_finish(Face f, const cx<double>* GX1, const cx<double>* GX2, const cx<double>* HX1,
                            long sH1, const cx<double>* HX2, long sX, const cx<double>* rhs, long sr, cx<double>* y,
                            long sy, double* norms) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = f.n;
    double rr = 0, bb = 0;
    if (i < n) {
        long o = (long)b * sX + i;
        long oh = (long)b * sH1 + i;
        const long N = n;
        double k = f.kvac[b];
        cx<double> ik = mk(0.0, k);
        cx<double> out[8];
        out[0] = GX1[o] - GX2[o];
        cx<double> dotn = mk(0.0, 0.0);
        for (int c = 0; c < 3; ++c) {
            out[1 + c] = GX1[o + (1 + c) * N] - GX2[o + (1 + c) * N];
            dotn = dotn + f.nv(b)[3 * i + c] * (GX1[o + (5 + c) * N] - GX2[o + (5 + c) * N]);
        }
        out[4] = HX1[oh + 4 * N] - HX2[o + 4 * N] - ik * dotn;
        cx<double> es = GX1[o + 4 * N] - GX2[o + 4 * N];
        for (int c = 0; c < 3; ++c)
            out[5 + c] = HX1[oh + (1 + c) * N] - HX2[o + (1 + c) * N] - (f.nv(b)[3 * i + c] * ik) * es;
        for (int c = 0; c < 8; ++c) {
            cx<double> v = out[c];
            if (rhs) {
                cx<double> bv = rhs[(long)b * sr + i + c * N];
                v = bv - v;
                bb += bv.re * bv.re + bv.im * bv.im;
            }
            rr += v.re * v.re + v.im * v.im;
            y[(long)b * sy + i + c * N] = v;
        }
    }
    if (norms) {
        for (int off = 32; off > 0; off >>= 1) {
            rr += __shfl_down(rr, off, 64);
            bb += __shfl_down(bb, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&norms[2 * b], rr);
            atomicAdd(&norms[2 * b + 1], bb);
        }
    }
}

// dst (T) = src (double) or dst (double) += src (T)
template <typename T>
__global__ void k_cast(long count, const cx<double>* src, long ss, cx<T>* dst, long sd) {
    long e = (long)blockIdx.x * 256 + threadIdx.x;
    int b = blockIdx.y;
    if (e >= count) return;
    cx<double> v = src[(long)b * ss + e];
    dst[(long)b * sd + e] = mk<T>((T)v.re, (T)v.im);
}
template <typename T>
__global__ void k_accum(long count, const cx<T>* src, long ss, cx<double>* dst, long sd, const int* active) {
    long e = (long)blockIdx.x * 256 + threadIdx.x;
    int b = blockIdx.y;
    if (e >= count) return;
    if (active && !active[b]) return;
    cx<T> v = src[(long)b * ss + e];
    cx<double>* d = dst + (long)b * sd + e;
    d->re += (double)v.re;
    d->im += (double)v.im;
}

// ----------------------------------------------------------------------------- plane wave
// Packed right-hand side [phi a De alpha] for a plane wave of polarisation pol, direction dir
// in medium `med` (1-based) (simulation/planewave_ret.py:257-361; bem_ret.py:1360-1429):
//   on faces whose outside (inside) medium is med: a2 (a1) = e^{i km r.dir}/(i k0) pol,
//   a' = i km (n.dir) a;  a = a2 - a1, alpha = a2' - a1', De = -i k0 (eps2 n.a2 - eps1 n.a1).

__global__ void k_planewave(PW p, cx<double>* out, long so) {
    int i = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, n = p.n;
    if (i >= n) return;
    cx<double> km = reinterpret_cast<const cx<double>*>(p.kmed)[b];
    double k0 = p.kvac[b];
    const double* pos = p.pos + (long)b * 3 * p.sface;
    const double* nvec = p.nvec + (long)b * 3 * p.sface;
    double rd = pos[3 * i] * p.dir[0] + pos[3 * i + 1] * p.dir[1] + pos[3 * i + 2] * p.dir[2];
    double nd = nvec[3 * i] * p.dir[0] + nvec[3 * i + 1] * p.dir[1] + nvec[3 * i + 2] * p.dir[2];
    // e^{i km rd} / (i k0) = -i/k0 * e^{i km rd}
    cx<double> e = expikd(km, rd);
    cx<double> ph = mk(e.im / k0, -e.re / k0);
    cx<double> dph = mk(-km.im * nd, km.re * nd) * ph;  // i km (n.dir) ph
    int sd = p.side[(long)b * p.sface + i];
    double w1 = (sd & 1) ? 1.0 : 0.0, w2 = (sd & 2) ? 1.0 : 0.0;
    cx<double> e1 = Lval<double>(p.L1, b, n, i), e2 = Lval<double>(p.L2, b, n, i);
    cx<double>* o = out + (long)b * so;
    const long N = n;
    double npol = nvec[3 * i] * p.pol[0] + nvec[3 * i + 1] * p.pol[1] + nvec[3 * i + 2] * p.pol[2];
    o[i] = mk(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
        o[i + (1 + c) * N] = ((w2 - w1) * p.pol[c]) * ph;
        o[i + (5 + c) * N] = ((w2 - w1) * p.pol[c]) * dph;
    }
    // De = -i k0 (eps2 w2 - eps1 w1) (n.pol) ph
    cx<double> ediff = w2 * e2 - w1 * e1;
    o[i + 4 * N] = mk(0.0, -k0) * ((npol) * (ediff * ph));
}

// ----------------------------------------------------------------------------- far field
// e(d) = i k0 sum_j ph_j h_j - i km d sum_j ph_j sig_j, h(d) = i km d x sum_j ph_j h_j,
// ph_j = e^{-i km d.r_j} a_j, over faces whose inside (sig1, h1) / outside (sig2, h2) medium is
// med (spectrum/spectrum_ret.py:231-392). One workgroup per (direction, wavelength);
// out = [e(3), h(3)] complex per direction and wavelength.
__global__ __launch_bounds__(256) void k_farfield(PW p, const double* area, int ndir, const double* dirs,
                                                  const cx<double>* x, long sx, cx<double>* out) {
    __shared__ double red[8][8];
    int d = blockIdx.x, b = blockIdx.y, t = threadIdx.x, n = p.n;
    cx<double> km = reinterpret_cast<const cx<double>*>(p.kmed)[b];
    double k0 = p.kvac[b];
    double dx = dirs[3 * d], dy = dirs[3 * d + 1], dz = dirs[3 * d + 2];
    const cx<double>* xv = x + (long)b * sx;
    const long N = n;
    const double* pos = p.pos + (long)b * 3 * p.sface;
    const int* side = p.side + (long)b * p.sface;
    area += (long)b * p.sface;
    // accumulators: S = sum ph sig, Hs = sum ph h (3)
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = t; i < n; i += 256) {
        double rd = dx * pos[3 * i] + dy * pos[3 * i + 1] + dz * pos[3 * i + 2];
        cx<double> e = expikd(km, -rd);
        cx<double> ph = area[i] * e;
        int sd = side[i];
        for (int s = 0; s < 2; ++s) {
            if (!(sd & (1 << s))) continue;
            int base = 4 * s;
            cx<double> sg = ph * xv[i + base * N];
            acc[0] += sg.re;
            acc[1] += sg.im;
            for (int c = 0; c < 3; ++c) {
                cx<double> hv = ph * xv[i + (base + 1 + c) * N];
                acc[2 + 2 * c] += hv.re;
                acc[3 + 2 * c] += hv.im;
            }
        }
    }
    for (int q = 0; q < 8; ++q) {
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((t & 63) == 0) red[t >> 6][q] = v;
    }
    __syncthreads();
    if (t == 0) {
        double s[8];
        for (int q = 0; q < 8; ++q) s[q] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
        cx<double> S = mk(s[0], s[1]);
        cx<double> Hs[3] = {mk(s[2], s[3]), mk(s[4], s[5]), mk(s[6], s[7])};
        double dv[3] = {dx, dy, dz};
        cx<double> ik0 = mk(0.0, k0), ikm = mk(-km.im, km.re);
        cx<double>* o = out + ((long)b * ndir + d) * 6;
        for (int c = 0; c < 3; ++c) o[c] = ik0 * Hs[c] - (dv[c] * ikm) * S;
        o[3] = ikm * (dv[1] * Hs[2] - dv[2] * Hs[1]);
        o[4] = ikm * (dv[2] * Hs[0] - dv[0] * Hs[2]);
        o[5] = ikm * (dv[0] * Hs[1] - dv[1] * Hs[0]);
    }
}

// ----------------------------------------------------------------------------- launchers
template <typename T>
void sub_batched(int n, int batch, const void* A, const void* B, void* C, long ld, long stride, hipStream_t s) {
    hipLaunchKernelGGL(k_sub<T>, dim3((n + 255) / 256, n, batch), dim3(256), 0, s, n, (const cx<T>*)A,
                       (const cx<T>*)B, (cx<T>*)C, ld, stride);
}
template <typename T>
void sigma_batched(int n, int batch, const void* S1, void* S2, const void* Dinv, long ld, long stride,
                   const double* nvec, long sface, const double* L1, const double* L2, const double* kvac,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_sigma<T>, dim3((n + 255) / 256, n, batch), dim3(256), 0, s, n, (const cx<T>*)S1, (cx<T>*)S2,
                       (const cx<T>*)Dinv, ld, stride, nvec, sface, L1, L2, kvac);
}

__global__ void k_zero_int(long count, int* v) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) v[i] = 0;
}

// Clears count ints with a kernel (no memset node when the caller captures the stream).
void zero_int(long count, int* v, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_int, dim3((count + 255) / 256), dim3(256), 0, s, count, v);
}

// Padding of particles smaller than the batch's matrix size: entries (i, j) with i or j >= nact[b]
// of matrix b become d (i == j) or 0, so the padded system is block diagonal with a constant
// diagonal block that the LU chain carries through unchanged.
template <typename T>
__global__ void k_pad(int n, const int* nact, cx<T>* A, long ld, long stride, double dre, double dim) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, b = blockIdx.z;
    const int na = nact[b];
    if (i >= n || (i < na && j < na)) return;
    A[(long)b * stride + i + (long)j * ld] = (i == j) ? mk<T>((T)dre, (T)dim) : mk<T>(0, 0);
}

void pad_batched(int prec, int n, int batch, const int* nact, void* A, long ld, long stride, double dre, double dim,
                 hipStream_t s) {
    dim3 g((n + 255) / 256, n, batch);
    if (prec == 0)
        hipLaunchKernelGGL(k_pad<double>, g, dim3(256), 0, s, n, nact, (cx<double>*)A, ld, stride, dre, dim);
    else
        hipLaunchKernelGGL(k_pad<float>, g, dim3(256), 0, s, n, nact, (cx<float>*)A, ld, stride, dre, dim);
}
template <typename T>
void identity_batched(int n, int batch, void* A, long ld, long stride, hipStream_t s) {
    hipLaunchKernelGGL(k_identity<T>, dim3((n + 255) / 256, n, batch), dim3(256), 0, s, n, (cx<T>*)A, ld, stride);
}

template <typename T>
void getrs_batched(int n, int r, int batch, const void* LU, long lda, long sA, const int* pf, long sQ, const void* B,
                   long ldb, long sB, void* X, long ldx, long sX, hipStream_t s);

// Workspace for the solve chain: 24 N complex values of T per wavelength (see offsets below).
template <typename T>
void solve_chain(const Face& f, int batch, const void* G1, const void* G2, const void* S1, const void* Dinv,
                 const void* Sg, long ld, long stride, const int* pf1, const int* pf2, const int* pfS, long sQ,
                 const void* IN, long sIn, void* OUT, long sOut, void* work, long sW, hipStream_t s) {
    const int n = f.n;
    const long N = n;
    cx<T>* w = (cx<T>*)work;
    cx<T>* V = w;             // N x 4
    cx<T>* W = w + 4 * N;     // N x 4
    cx<T>* AP = w + 8 * N;    // N x 4  [alpha' De']
    cx<T>* U = w + 12 * N;    // N x 3
    cx<T>* R = w + 15 * N;    // N x 1
    cx<T>* SG = w + 16 * N;   // N x 1 (sig2)
    cx<T>* TT = w + 17 * N;   // N x 3
    cx<T>* H = w + 20 * N;    // N x 3 (h2)   -> ends at 23 N
    cx<T>* P1 = W;            // reuse N x 4
    cx<T>* P2 = V;            // reuse N x 4
    cx<T>* out = (cx<T>*)OUT;
    const cx<T>* in = (const cx<T>*)IN;
    dim3 g((n + 255) / 256, batch);
    hipLaunchKernelGGL(k_glue1<T>, g, dim3(256), 0, s, f, in, sIn, V, sW);
    gemv_batched<T>(n, 4, batch, S1, ld, stride, V, N, sW, W, N, sW, s);
    hipLaunchKernelGGL(k_glue2<T>, g, dim3(256), 0, s, f, in, sIn, W, sW, AP, sW);
    gemv_batched<T>(n, 3, batch, Dinv, ld, stride, AP, N, sW, U, N, sW, s);
    hipLaunchKernelGGL(k_glue3<T>, g, dim3(256), 0, s, f, AP, sW, U, sW, R, sW);
    // pivots null: G1, G2 and Sg hold explicit inverses (products instead of substitutions)
    if (pfS)
        getrs_batched<T>(n, 1, batch, Sg, ld, stride, pfS, sQ, R, N, sW, SG, N, sW, s);
    else
        gemv_batched<T>(n, 1, batch, Sg, ld, stride, R, N, sW, SG, N, sW, s);
    hipLaunchKernelGGL(k_glue4<T>, g, dim3(256), 0, s, f, AP, sW, SG, sW, TT, sW);
    gemv_batched<T>(n, 3, batch, Dinv, ld, stride, TT, N, sW, H, N, sW, s);
    hipLaunchKernelGGL(k_glue5<T>, g, dim3(256), 0, s, f, in, sIn, SG, sW, H, sW, P1, P2, sW);
    if (pf1) {
        getrs_batched<T>(n, 4, batch, G1, ld, stride, pf1, sQ, P1, N, sW, out, N, sOut, s);
        getrs_batched<T>(n, 4, batch, G2, ld, stride, pf2, sQ, P2, N, sW, out + 4 * N, N, sOut, s);
    } else {
        gemv_batched<T>(n, 4, batch, G1, ld, stride, P1, N, sW, out, N, sOut, s);
        gemv_batched<T>(n, 4, batch, G2, ld, stride, P2, N, sW, out + 4 * N, N, sOut, s);
    }
}

template void solve_chain<double>(const Face&, int, const void*, const void*, const void*, const void*, const void*,
                                  long, long, const int*, const int*, const int*, long, const void*, long, void*, long,
                                  void*, long, hipStream_t);
template void solve_chain<float>(const Face&, int, const void*, const void*, const void*, const void*, const void*,
                                 long, long, const int*, const int*, const int*, long, const void*, long, void*, long,
                                 void*, long, hipStream_t);
template void sub_batched<double>(int, int, const void*, const void*, void*, long, long, hipStream_t);
template void sub_batched<float>(int, int, const void*, const void*, void*, long, long, hipStream_t);
template void sigma_batched<double>(int, int, const void*, void*, const void*, long, long, const double*, long,
                                    const double*, const double*, const double*, hipStream_t);
template void sigma_batched<float>(int, int, const void*, void*, const void*, long, long, const double*, long,
                                   const double*, const double*, const double*, hipStream_t);
template void identity_batched<double>(int, int, void*, long, long, hipStream_t);
template void identity_batched<float>(int, int, void*, long, long, hipStream_t);
template void gemv_batched<double>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                                   hipStream_t);
template void gemv_batched<float>(int, int, int, const void*, long, long, const void*, long, long, void*, long, long,
                                  hipStream_t);

// Exact FP64 operator / residual. work: 32 N complex double per wavelength.
void op_prep(const Face& f, int batch, const void* x, long sx, void* X1, void* X2, long sX, hipStream_t s);
void op_finish(const Face& f, int batch, const void* GX1, const void* GX2, const void* HX1, long sH1, const void* HX2,
               long sX, const void* rhs, long sr, void* y, long sy, double* norms, hipStream_t s);

void bem_operator(const Face& f, int batch, const void* G1, const void* G2, const void* H1, const void* H2, long ld,
                  long stride, const void* x, long sx, const void* rhs, long sr, void* y, long sy, void* work,
                  long sW, double* norms, hipStream_t s) {
    const int n = f.n;
    const long N = n;
    cx<double>* w = (cx<double>*)work;
    cx<double>*X1 = w, *X2 = w + 8 * N, *GX1 = w + 16 * N, *GX2 = w + 24 * N;
    op_prep(f, batch, x, sx, X1, X2, sW, s);
    gemv_batched<double>(n, 8, batch, G1, ld, stride, X1, N, sW, GX1, N, sW, s);
    gemv_batched<double>(n, 8, batch, G2, ld, stride, X2, N, sW, GX2, N, sW, s);
    // H1 X1 is staged in y (rewritten row by row by k_op_finish), H2 X2 in X1 (consumed by then);
    // y must not alias x.
    cx<double>* HX1 = (cx<double>*)y;
    gemv_batched<double>(n, 8, batch, H1, ld, stride, X1, N, sW, HX1, N, sy, s);
    gemv_batched<double>(n, 8, batch, H2, ld, stride, X2, N, sW, X1, N, sW, s);  // HX2 into X1 (X2 consumed)
    op_finish(f, batch, GX1, GX2, HX1, sy, X1, sW, rhs, sr, y, sy, norms, s);
}

__global__ void k_zero_norms(int count, double* v) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) v[i] = 0.0;
}

void op_prep(const Face& f, int batch, const void* x, long sx, void* X1, void* X2, long sX, hipStream_t s) {
    hipLaunchKernelGGL(k_op_prep, dim3((f.n + 255) / 256, batch), dim3(256), 0, s, f, (const cx<double>*)x, sx,
                       (cx<double>*)X1, (cx<double>*)X2, sX);
}

void op_finish(const Face& f, int batch, const void* GX1, const void* GX2, const void* HX1, long sH1, const void* HX2,
               long sX, const void* rhs, long sr, void* y, long sy, double* norms, hipStream_t s) {
    // norms are cleared by a kernel (no memset node if the caller captures the stream)
    if (norms) hipLaunchKernelGGL(k_zero_norms, dim3((2 * batch + 255) / 256), dim3(256), 0, s, 2 * batch, norms);
    hipLaunchKernelGGL(k_op_finish, dim3((f.n + 255) / 256, batch), dim3(256), 0, s, f, (const cx<double>*)GX1,
                       (const cx<double>*)GX2, (const cx<double>*)HX1, sH1, (const cx<double>*)HX2, sX,
                       (const cx<double>*)rhs, sr, (cx<double>*)y, sy, norms);
}

void planewave_rhs(const PW& p, int batch, void* out, long so, hipStream_t s) {
    hipLaunchKernelGGL(k_planewave, dim3((p.n + 255) / 256, batch), dim3(256), 0, s, p, (cx<double>*)out, so);
}

void farfield(const PW& p, const double* area, int ndir, const double* dirs, int batch, const void* x, long sx,
              void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_farfield, dim3(ndir, batch), dim3(256), 0, s, p, area, ndir, dirs, (const cx<double>*)x, sx,
                       (cx<double>*)out);
}

template <typename T>
void cast_batched(long count, int batch, const void* src, long ss, void* dst, long sd, hipStream_t s) {
    hipLaunchKernelGGL(k_cast<T>, dim3((count + 255) / 256, batch), dim3(256), 0, s, count, (const cx<double>*)src, ss,
                       (cx<T>*)dst, sd);
}
template <typename T>
void accum_batched(long count, int batch, const void* src, long ss, void* dst, long sd, const int* active,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_accum<T>, dim3((count + 255) / 256, batch), dim3(256), 0, s, count, (const cx<T>*)src, ss,
                       (cx<double>*)dst, sd, active);
}
template void cast_batched<float>(long, int, const void*, long, void*, long, hipStream_t);
template void cast_batched<double>(long, int, const void*, long, void*, long, hipStream_t);
template void accum_batched<float>(long, int, const void*, long, void*, long, const int*, hipStream_t);
template void accum_batched<double>(long, int, const void*, long, void*, long, const int*, hipStream_t);

}  // namespace mnp
