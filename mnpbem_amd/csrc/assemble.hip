// Retarded Green function G and surface derivative H = F +- 2 pi assembly.
//
// For observer face i (row) and source face j (column), at wavenumber k:
//   G_ij = a_j / d_ij * e^{i k d_ij}
//   F_ij = (n_i . r_ij) (i k - 1/d_ij) / d_ij^2 * a_j * e^{i k d_ij},  r_ij = p_i - p_j
// (reference greenfun/greenret_refined.py:654-810). H1 = F + 2 pi, H2 = F - 2 pi on the
// diagonal of same-particle blocks. The wavenumber of block (particle a, particle b) is that
// of the material connecting the two regions, con_in[a][b] for the inside matrices
// (G1, H1) and con_out[a][b] for the outside ones (G2, H2) (greenfun/compgreen_ret.py:435-505,
// :739-858); 0 = not connected (entry 0).
//
// Refined pairs (self terms and near pairs) are overwritten afterwards with the Taylor series
//   G = sum_n g_n (ik)^n e^{ikd},  F = n_i . sum_n f_n (ik)^n e^{ikd}
// whose coefficients are wavelength independent (greenret_refined.py:738-777).
//
// Layout: column-major matrices, interleaved complex, leading dimension ld, one matrix per
// wavelength at stride `stride` elements. One launch assembles the four matrices for
// every wavelength of the batch; geometry terms are computed once per entry and reused.
#include "common.hpp"
#include "args.hpp"

namespace mnp {

constexpr double TWO_PI = 6.283185307179586476925286766559;


template <typename OT>
__global__ __launch_bounds__(256) void k_assemble_far(AsmArgs a, cx<OT>* G1, cx<OT>* G2, cx<OT>* H1,
                                                     cx<OT>* H2) {
    int i = blockIdx.x * 64 + (threadIdx.x & 63);
    int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= a.n || j >= a.n) return;
    double rx = a.pos[3 * i] - a.pos[3 * j];
    double ry = a.pos[3 * i + 1] - a.pos[3 * j + 1];
    double rz = a.pos[3 * i + 2] - a.pos[3 * j + 2];
    double d = sqrt(rx * rx + ry * ry + rz * rz);
    double nr = a.nvec[3 * i] * rx + a.nvec[3 * i + 1] * ry + a.nvec[3 * i + 2] * rz;
    double aj = a.area[j];
    int pa = a.fpart[i], pb = a.fpart[j];
    int m1 = a.con_in[pa * a.npart + pb];
    int m2 = a.con_out[pa * a.npart + pb];
    bool diag = (i == j);
    double inv_d = diag ? 0.0 : 1.0 / d;
    long off = (long)i + (long)j * a.ld;
    const cx<double>* kk = reinterpret_cast<const cx<double>*>(a.kmat);
    for (int l = 0; l < a.nlam; ++l) {
        long o = off + (long)l * a.stride;
        for (int side = 0; side < 2; ++side) {
            int m = side ? m2 : m1;
            cx<OT> g = mk<OT>(0, 0), f = mk<OT>(0, 0);
            if (m > 0 && !diag) {
                cx<double> k = kk[l * a.nmat + m - 1];
                cx<double> ph = expikd(k, d);
                cx<double> gg = (aj * inv_d) * ph;
                // (i k - 1/d) / d^2 * nr * a_j
                cx<double> fac = mk(-k.im - inv_d, k.re);
                cx<double> ff = (nr * aj * inv_d * inv_d) * (fac * ph);
                g = conv<OT>(gg);
                f = conv<OT>(ff);
            }
            if (side == 0) { G1[o] = g; H1[o] = f; }
            else { G2[o] = g; H2[o] = f; }
        }
    }
}

template <typename OT>
__global__ __launch_bounds__(256) void k_assemble_refined(AsmArgs a, int nref, const int* rrow, const int* rcol,
                                                         const double* rg, const double* rf, cx<OT>* G1,
                                                         cx<OT>* G2, cx<OT>* H1, cx<OT>* H2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nref) return;
    int i = rrow[p], j = rcol[p];
    int no = a.order + 1;
    double rx = a.pos[3 * i] - a.pos[3 * j];
    double ry = a.pos[3 * i + 1] - a.pos[3 * j + 1];
    double rz = a.pos[3 * i + 2] - a.pos[3 * j + 2];
    double d = sqrt(rx * rx + ry * ry + rz * rz);
    double nx = a.nvec[3 * i], ny = a.nvec[3 * i + 1], nz = a.nvec[3 * i + 2];
    // n_i . f_n for every order (f stored as n_ref x 3 x (order+1))
    double fn[8], gn[8];
    for (int q = 0; q < no; ++q) {
        gn[q] = rg[(long)p * no + q];
        fn[q] = nx * rf[((long)p * 3 + 0) * no + q] + ny * rf[((long)p * 3 + 1) * no + q] +
                nz * rf[((long)p * 3 + 2) * no + q];
    }
    int pa = a.fpart[i], pb = a.fpart[j];
    int m1 = a.con_in[pa * a.npart + pb];
    int m2 = a.con_out[pa * a.npart + pb];
    bool diag = (i == j);
    long off = (long)i + (long)j * a.ld;
    const cx<double>* kk = reinterpret_cast<const cx<double>*>(a.kmat);
    for (int l = 0; l < a.nlam; ++l) {
        long o = off + (long)l * a.stride;
        for (int side = 0; side < 2; ++side) {
            int m = side ? m2 : m1;
            if (m <= 0) continue;
            cx<double> k = kk[l * a.nmat + m - 1];
            cx<double> ik = mk(-k.im, k.re);
            // Horner in (ik)
            cx<double> sg = mk(gn[no - 1], 0.0), sf = mk(fn[no - 1], 0.0);
            for (int q = no - 2; q >= 0; --q) {
                sg = ik * sg + mk(gn[q], 0.0);
                sf = ik * sf + mk(fn[q], 0.0);
            }
            cx<double> ph = expikd(k, d);
            cx<double> g = sg * ph, f = sf * ph;
            if (diag) f.re += side ? -TWO_PI : TWO_PI;
            if (side == 0) { G1[o] = conv<OT>(g); H1[o] = conv<OT>(f); }
            else { G2[o] = conv<OT>(g); H2[o] = conv<OT>(f); }
        }
    }
}

template <typename OT>
void launch_assemble(const AsmArgs& a, int nref, const int* rrow, const int* rcol, const double* rg,
                     const double* rf, void* G1, void* G2, void* H1, void* H2, hipStream_t s) {
    dim3 grid((a.n + 63) / 64, (a.n + 3) / 4);
    hipLaunchKernelGGL(k_assemble_far<OT>, grid, dim3(256), 0, s, a, (cx<OT>*)G1, (cx<OT>*)G2, (cx<OT>*)H1,
                       (cx<OT>*)H2);
    if (nref > 0) {
        hipLaunchKernelGGL(k_assemble_refined<OT>, dim3((nref + 255) / 256), dim3(256), 0, s, a, nref, rrow, rcol,
                           rg, rf, (cx<OT>*)G1, (cx<OT>*)G2, (cx<OT>*)H1, (cx<OT>*)H2);
    }
}

template void launch_assemble<double>(const AsmArgs&, int, const int*, const int*, const double*, const double*,
                                      void*, void*, void*, void*, hipStream_t);
template void launch_assemble<float>(const AsmArgs&, int, const int*, const int*, const double*, const double*,
                                     void*, void*, void*, void*, hipStream_t);

}  // namespace mnp
