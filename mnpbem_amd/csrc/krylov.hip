// Krylov-space kernels for the device GMRES of BEMRetIter (bem/bem_iter.py:45-101 drives
// scipy's GMRES on the host in the reference; here the Arnoldi vectors stay in HBM).
//
// Basis vectors are complex double, stored one after another with leading dimension ldv.
//   k_mdot    : h[c] += sum_e conj(V_c[e]) w[e]   for c < nv   (h zeroed by the caller)
//   k_mupdate : w[e] += sum_c V_c[e] s[c]          for c < nv   (s complex, host-supplied sign)
// Both are HBM-bound streams over nv * len complex values; classical Gram-Schmidt with one
// re-orthogonalisation (CGS2) needs two of each per Arnoldi step.
#include "common.hpp"

namespace mnp {

constexpr int KMAX = 128;

__global__ __launch_bounds__(256) void k_mdot(long len, int nv, const cx<double>* V, long ldv, const cx<double>* w,
                                              double* h) {
    __shared__ double red[4][2];
    const int c = blockIdx.y, t = threadIdx.x;
    const cx<double>* v = V + (long)c * ldv;
    double re = 0, im = 0;
    for (long e = (long)blockIdx.x * 256 + t; e < len; e += (long)gridDim.x * 256) {
        cx<double> a = v[e], b = w[e];
        re += a.re * b.re + a.im * b.im;
        im += a.re * b.im - a.im * b.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_down(re, off, 64);
        im += __shfl_down(im, off, 64);
    }
    if ((t & 63) == 0) { red[t >> 6][0] = re; red[t >> 6][1] = im; }
    __syncthreads();
    if (t == 0) {
        double sr = red[0][0] + red[1][0] + red[2][0] + red[3][0];
        double si = red[0][1] + red[1][1] + red[2][1] + red[3][1];
        atomicAdd(&h[2 * c], sr);
        atomicAdd(&h[2 * c + 1], si);
    }
}

__global__ __launch_bounds__(256) void k_mupdate(long len, int nv, const cx<double>* V, long ldv, const double* s,
                                                 cx<double>* w) {
    __shared__ double ss[2 * KMAX];
    const int t = threadIdx.x;
    for (int q = t; q < 2 * nv; q += 256) ss[q] = s[q];
    __syncthreads();
    long e = (long)blockIdx.x * 256 + t;
    if (e >= len) return;
    cx<double> acc = w[e];
    for (int c = 0; c < nv; ++c) {
        cx<double> a = V[(long)c * ldv + e];
        acc.re += a.re * ss[2 * c] - a.im * ss[2 * c + 1];
        acc.im += a.re * ss[2 * c + 1] + a.im * ss[2 * c];
    }
    w[e] = acc;
}

void mdot(long len, int nv, const void* V, long ldv, const void* w, double* h, hipStream_t s) {
    long blocks = (len + 256 * 8 - 1) / (256 * 8);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipMemsetAsync(h, 0, sizeof(double) * 2 * nv, s);
    hipLaunchKernelGGL(k_mdot, dim3((unsigned)blocks, nv), dim3(256), 0, s, len, nv, (const cx<double>*)V, ldv,
                       (const cx<double>*)w, h);
}

void mupdate(long len, int nv, const void* V, long ldv, const double* sv, void* w, hipStream_t s) {
    hipLaunchKernelGGL(k_mupdate, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, nv,
                       (const cx<double>*)V, ldv, sv, (cx<double>*)w);
}

}  // namespace mnp
