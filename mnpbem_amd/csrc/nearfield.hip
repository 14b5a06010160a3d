// Near-field set-up on the device (wavelength independent, once per geometry).
//
// k_near_terms: Taylor coefficients of a near pair (observer i, source j) from the source
// face's quadrature points p_q with weights w_q (greenret_refined.py:287-516): with
// r_q = |x_i - p_q|, r0 = |x_j - x_i|, d_q = r_q - r0,
//   g_n = sum_q w_q d_q^n / (r_q n!)
//   f_0 = -sum_q w_q (x_i - p_q) / r_q^3
//   f_n = sum_q w_q (x_i - p_q) [ -d_q^n / (r_q^3 n!) + d_q^(n-1) / (r_q^2 (n-1)!) ],  n >= 1
// one thread per pair; the points of face j are the CSR range qptr[j] .. qptr[j+1].
//
// k_colsum_far: collocated part of the column sums of the quasistatic surface derivative
// over a closed surface (compgreen_stat.py:219-464, compgreen_ret.py:192-272):
//   s_j = sum_{i in loc} -a_i n_i.(x_i - x_j) / |x_i - x_j|^3   for j in ind,
// one thread per column, rows staged through LDS in tiles of 256 faces.
#include "common.hpp"

namespace mnp {

constexpr int NF_MAXO = 8;
constexpr double NF_TINY = 2.220446049250313e-16;   // double epsilon (distance floor)

__global__ __launch_bounds__(256) void k_near_terms(int npair, const int* rows, const int* cols, const long* slot,
                                                    const double* pos, const int* qptr, const double* qpts,
                                                    const double* qw, int order, double* g, double* f) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npair) return;
    const int i = rows[p], j = cols[p];
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double dx0 = pos[3 * j] - xi, dy0 = pos[3 * j + 1] - yi, dz0 = pos[3 * j + 2] - zi;
    const double r0 = sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    const int no = order + 1;
    double gs[NF_MAXO], fx[NF_MAXO], fy[NF_MAXO], fz[NF_MAXO];
#pragma unroll
    for (int n = 0; n < NF_MAXO; ++n) gs[n] = fx[n] = fy[n] = fz[n] = 0;
    double fact[NF_MAXO];
    fact[0] = 1;
#pragma unroll
    for (int n = 1; n < NF_MAXO; ++n) fact[n] = fact[n - 1] * n;
    for (int q = qptr[j]; q < qptr[j + 1]; ++q) {
        const double rx = xi - qpts[3 * q], ry = yi - qpts[3 * q + 1], rz = zi - qpts[3 * q + 2];
        const double r = fmax(sqrt(rx * rx + ry * ry + rz * rz), NF_TINY);
        const double w = qw[q];
        const double d = r - r0;
        const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
        double dp[NF_MAXO];
        dp[0] = 1;
#pragma unroll
        for (int n = 1; n < NF_MAXO; ++n) dp[n] = dp[n - 1] * d;
#pragma unroll
        for (int n = 0; n < NF_MAXO; ++n) {
            if (n < no) {
                gs[n] += w * dp[n] * ir;
                const double s = (n == 0) ? -ir3 : (-dp[n] * ir3 / fact[n] + dp[n - 1] * ir2 / fact[n - 1]);
                fx[n] += w * s * rx;
                fy[n] += w * s * ry;
                fz[n] += w * s * rz;
            }
        }
    }
    const long o = slot[p];
#pragma unroll
    for (int n = 0; n < NF_MAXO; ++n) {
        if (n < no) {
            g[o * no + n] = gs[n] / fact[n];
            f[(o * 3 + 0) * no + n] = fx[n];
            f[(o * 3 + 1) * no + n] = fy[n];
            f[(o * 3 + 2) * no + n] = fz[n];
        }
    }
}

__global__ __launch_bounds__(256) void k_colsum_far(int nloc, const int* loc, int nind, const int* ind,
                                                    const double* pos, const double* nvec, const double* area,
                                                    double* out) {
    __shared__ double sp[256][7];
    const int t = threadIdx.x;
    const int jj = blockIdx.x * 256 + t;
    const bool have = jj < nind;
    const int j = have ? ind[jj] : 0;
    const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2];
    double acc = 0;
    for (int t0 = 0; t0 < nloc; t0 += 256) {
        const int li = t0 + t;
        if (li < nloc) {
            const int i = loc[li];
            sp[t][0] = pos[3 * i];
            sp[t][1] = pos[3 * i + 1];
            sp[t][2] = pos[3 * i + 2];
            sp[t][3] = nvec[3 * i];
            sp[t][4] = nvec[3 * i + 1];
            sp[t][5] = nvec[3 * i + 2];
            sp[t][6] = area[i];
        }
        __syncthreads();
        const int cnt = min(256, nloc - t0);
        for (int s = 0; s < cnt; ++s) {
            const double rx = sp[s][0] - xj, ry = sp[s][1] - yj, rz = sp[s][2] - zj;
            const double d = fmax(sqrt(rx * rx + ry * ry + rz * rz), NF_TINY);
            const double nr = sp[s][3] * rx + sp[s][4] * ry + sp[s][5] * rz;
            acc += -sp[s][6] * nr / (d * d * d);
        }
        __syncthreads();
    }
    if (have) out[jj] = acc;
}

void near_terms(int npair, const int* rows, const int* cols, const long* slot, const double* pos, const int* qptr,
                const double* qpts, const double* qw, int order, double* g, double* f, hipStream_t s) {
    if (npair <= 0) return;
    hipLaunchKernelGGL(k_near_terms, dim3((npair + 255) / 256), dim3(256), 0, s, npair, rows, cols, slot, pos, qptr,
                       qpts, qw, order, g, f);
}

void colsum_far(int nloc, const int* loc, int nind, const int* ind, const double* pos, const double* nvec,
                const double* area, double* out, hipStream_t s) {
    if (nind <= 0) return;
    hipLaunchKernelGGL(k_colsum_far, dim3((nind + 255) / 256), dim3(256), 0, s, nloc, loc, nind, ind, pos, nvec, area,
                       out);
}

}  // namespace mnp
