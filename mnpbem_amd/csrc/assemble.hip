// Retarded Green function G and surface derivative H = F +- 2 pi assembly.
//
// For observer face i (row) and source face j (column), at wavenumber k:
//   G_ij = a_j / d_ij * e^{i k d_ij}
//   F_ij = (n_i . r_ij) (i k - 1/d_ij) / d_ij^2 * a_j * e^{i k d_ij},  r_ij = p_i - p_j
// (reference greenfun/greenret_refined.py:654-810). H1 = F + 2 pi, H2 = F - 2 pi on the
// diagonal of same-particle blocks. The wavenumber of block (particle a, particle b) is that
// of the material connecting the two regions, con_in[a][b] for the inside matrices
// (G1, H1) and con_out[a][b] for the outside ones (G2, H2) (greenfun/compgreen_ret.py:435-505,
// :739-858); 0 = not connected (entry 0). A negative table entry -m is the cross-region term of
// nested particles, subtracted (G1 = G(in, in) - G(out, in), bem_ret.py:407-433): the entry is
// formed with material m and negated.
//
// Refined pairs (self terms and near pairs) are overwritten afterwards with the Taylor series
//   G = sum_n g_n (ik)^n e^{ikd},  F = n_i . sum_n f_n (ik)^n e^{ikd}
// whose coefficients are wavelength independent (greenret_refined.py:738-777).
//
// Layout: column-major matrices, interleaved complex, leading dimension ld, one matrix per
// wavelength at stride `stride` elements. One launch assembles the four matrices for every
// wavelength of the batch: a workgroup of 256 lanes owns R rows x (256 / R) columns of entries
// (lanes along the rows: each store instruction writes R consecutive elements of one column)
// and a chunk of LB wavelengths (grid z); the tile's observer and source face data are staged in
// LDS, and the geometry terms of an entry are computed once and reused for every wavelength of
// its chunk. R and LB are tuning knobs of the handle (small
// matrices fill the GPU only with short wavelength chunks); the autotuner times them per shape.
#include "common.hpp"
#include "args.hpp"
#include "ctx.hpp"

namespace mnp {

constexpr double TWO_PI = 6.283185307179586476925286766559;


// Second output set of the pair assembly: complex-float copies of the same entries (the factor
// matrices of the mixed chain next to the complex128 ones of its exact operator), at their own
// leading dimension and stride; one evaluation of each e^{ikd} serves both.
struct AsmOut2 {
    cx<float>* G1;
    cx<float>* G2;
    cx<float>* H1;
    cx<float>* H2;
    long ld, stride;
};

template <typename OT, int R, bool DUAL = false>
__global__ __launch_bounds__(256) void k_assemble_far(AsmArgs a, int lb, cx<OT>* G1, cx<OT>* G2, cx<OT>* H1,
                                                     cx<OT>* H2, AsmOut2 o2 = AsmOut2{}) {
    constexpr int C = 256 / R;
    // the workgroup's R observer faces (centroid, normal, particle) and C source faces
    // (centroid, area, particle), staged in LDS once and read by every entry of the tile
    __shared__ double ro[R][6];
    __shared__ double cs[C][4];
    __shared__ int rpart[R], cpart[C];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * R, j0 = blockIdx.y * C;
    if (t < R && i0 + t < a.n) {
        const int i = i0 + t;
        for (int c = 0; c < 3; ++c) {
            ro[t][c] = a.pos[3 * i + c];
            ro[t][3 + c] = a.nvec[3 * i + c];
        }
        rpart[t] = a.fpart[i];
    }
    if (t < C && j0 + t < a.n) {
        const int j = j0 + t;
        for (int c = 0; c < 3; ++c) cs[t][c] = a.pos[3 * j + c];
        cs[t][3] = a.area[j];
        cpart[t] = a.fpart[j];
    }
    __syncthreads();
    const int li = t % R, lj = t / R;
    const int i = i0 + li, j = j0 + lj;
    if (i >= a.n || j >= a.n) return;
    const int l0 = blockIdx.z * lb, l1 = min(a.nlam, l0 + lb);
    double rx = ro[li][0] - cs[lj][0];
    double ry = ro[li][1] - cs[lj][1];
    double rz = ro[li][2] - cs[lj][2];
    double d = sqrt(rx * rx + ry * ry + rz * rz);
    double nr = ro[li][3] * rx + ro[li][4] * ry + ro[li][5] * rz;
    double aj = cs[lj][3];
    int pa = rpart[li], pb = cpart[lj];
    int c1 = a.con_in[pa * a.npart + pb];
    int c2 = a.con_out[pa * a.npart + pb];
    bool diag = (i == j);
    double inv_d = diag ? 0.0 : 1.0 / d;
    long off = (long)i + (long)j * a.ld;
    const long off2 = (long)i + (long)j * o2.ld;
    const cx<double>* kk = reinterpret_cast<const cx<double>*>(a.kmat);
    for (int l = l0; l < l1; ++l) {
        long o = off + (long)l * a.stride;
        const long q = off2 + (long)l * o2.stride;
        for (int side = 0; side < 2; ++side) {
            const int c = side ? c2 : c1;
            const int m = c < 0 ? -c : c;
            const double sgn = c < 0 ? -1.0 : 1.0;
            cx<double> gg = mk(0.0, 0.0), ff = mk(0.0, 0.0);
            if (m > 0 && !diag) {
                cx<double> k = kk[l * a.nmat + m - 1];
                cx<double> ph = expikd(k, d);
                gg = (sgn * aj * inv_d) * ph;
                // (i k - 1/d) / d^2 * nr * a_j
                cx<double> fac = mk(-k.im - inv_d, k.re);
                ff = (sgn * nr * aj * inv_d * inv_d) * (fac * ph);
            }
            const cx<OT> g = conv<OT>(gg), f = conv<OT>(ff);
            if (side == 0) { G1[o] = g; H1[o] = f; }
            else { G2[o] = g; H2[o] = f; }
            if constexpr (DUAL) {
                const cx<float> gs = conv<float>(gg), fs = conv<float>(ff);
                if (side == 0) { o2.G1[q] = gs; o2.H1[q] = fs; }
                else { o2.G2[q] = gs; o2.H2[q] = fs; }
            }
        }
    }
}

template <typename OT, bool DUAL = false>
__global__ __launch_bounds__(256) void k_assemble_refined(AsmArgs a, int lb, int nref, const int* rrow,
                                                         const int* rcol, const double* rg, const double* rf,
                                                         cx<OT>* G1, cx<OT>* G2, cx<OT>* H1, cx<OT>* H2,
                                                         AsmOut2 o2 = AsmOut2{}) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nref) return;
    const int l0 = blockIdx.y * lb, l1 = min(a.nlam, l0 + lb);
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
    int c1 = a.con_in[pa * a.npart + pb];
    int c2 = a.con_out[pa * a.npart + pb];
    bool diag = (i == j);
    long off = (long)i + (long)j * a.ld;
    const cx<double>* kk = reinterpret_cast<const cx<double>*>(a.kmat);
    for (int l = l0; l < l1; ++l) {
        long o = off + (long)l * a.stride;
        for (int side = 0; side < 2; ++side) {
            const int c = side ? c2 : c1;
            const int m = c < 0 ? -c : c;
            const double sgn = c < 0 ? -1.0 : 1.0;
            if (m == 0) continue;
            cx<double> k = kk[l * a.nmat + m - 1];
            cx<double> ik = mk(-k.im, k.re);
            // Horner in (ik)
            cx<double> sg = mk(gn[no - 1], 0.0), sf = mk(fn[no - 1], 0.0);
            for (int q = no - 2; q >= 0; --q) {
                sg = ik * sg + mk(gn[q], 0.0);
                sf = ik * sf + mk(fn[q], 0.0);
            }
            cx<double> ph = expikd(k, d);
            cx<double> g = sgn * (sg * ph), f = sgn * (sf * ph);
            if (diag) f.re += side ? -TWO_PI : TWO_PI;
            if (side == 0) { G1[o] = conv<OT>(g); H1[o] = conv<OT>(f); }
            else { G2[o] = conv<OT>(g); H2[o] = conv<OT>(f); }
            if constexpr (DUAL) {
                const long q = (long)i + (long)j * o2.ld + (long)l * o2.stride;
                if (side == 0) { o2.G1[q] = conv<float>(g); o2.H1[q] = conv<float>(f); }
                else { o2.G2[q] = conv<float>(g); o2.H2[q] = conv<float>(f); }
            }
        }
    }
}

template <typename OT, bool DUAL>
static void launch_assemble_impl(const AsmArgs& a, int nref, const int* rrow, const int* rcol, const double* rg,
                                 const double* rf, void* G1, void* G2, void* H1, void* H2, const AsmOut2& o2,
                                 hipStream_t s) {
    const Ctx& h = ctx();
    const int R = h.asm_rows;
    const int lb = (h.asm_lchunk > 0 && h.asm_lchunk < a.nlam) ? h.asm_lchunk : a.nlam;
    const int nz = (a.nlam + lb - 1) / lb;
    dim3 grid((a.n + R - 1) / R, (a.n + 256 / R - 1) / (256 / R), nz);
    cx<OT>* g1 = (cx<OT>*)G1;
    cx<OT>* g2 = (cx<OT>*)G2;
    cx<OT>* h1 = (cx<OT>*)H1;
    cx<OT>* h2 = (cx<OT>*)H2;
    switch (R) {
        case 128:
            hipLaunchKernelGGL((k_assemble_far<OT, 128, DUAL>), grid, dim3(256), 0, s, a, lb, g1, g2, h1, h2, o2);
            break;
        case 256:
            hipLaunchKernelGGL((k_assemble_far<OT, 256, DUAL>), grid, dim3(256), 0, s, a, lb, g1, g2, h1, h2, o2);
            break;
        default:
            hipLaunchKernelGGL((k_assemble_far<OT, 64, DUAL>), grid, dim3(256), 0, s, a, lb, g1, g2, h1, h2, o2);
    }
    if (nref > 0) {
        hipLaunchKernelGGL((k_assemble_refined<OT, DUAL>), dim3((nref + 255) / 256, nz), dim3(256), 0, s, a, lb,
                           nref, rrow, rcol, rg, rf, g1, g2, h1, h2, o2);
    }
}

template <typename OT>
void launch_assemble(const AsmArgs& a, int nref, const int* rrow, const int* rcol, const double* rg,
                     const double* rf, void* G1, void* G2, void* H1, void* H2, hipStream_t s) {
    launch_assemble_impl<OT, false>(a, nref, rrow, rcol, rg, rf, G1, G2, H1, H2, AsmOut2{}, s);
}

// complex128 matrices and their complex-float copies from one pass
void launch_assemble_pair(const AsmArgs& a, int nref, const int* rrow, const int* rcol, const double* rg,
                          const double* rf, void* G1, void* G2, void* H1, void* H2, void* G1s, void* G2s, void* H1s,
                          void* H2s, long lds, long strides, hipStream_t s) {
    const AsmOut2 o2{(cx<float>*)G1s, (cx<float>*)G2s, (cx<float>*)H1s, (cx<float>*)H2s, lds, strides};
    launch_assemble_impl<double, true>(a, nref, rrow, rcol, rg, rf, G1, G2, H1, H2, o2, s);
}

// Assembly knobs of the calling thread's handle: rows per workgroup (64, 128, 256) and
// wavelengths per thread (0: all of the launch).
int set_asm_rows(int v) {
    if (v != 64 && v != 128 && v != 256) return -2;
    ctx().asm_rows = v;
    return 0;
}
int set_asm_lchunk(int v) {
    if (v < 0) return -2;
    ctx().asm_lchunk = v;
    return 0;
}

template void launch_assemble<double>(const AsmArgs&, int, const int*, const int*, const double*, const double*,
                                      void*, void*, void*, void*, hipStream_t);
template void launch_assemble<float>(const AsmArgs&, int, const int*, const int*, const double*, const double*,
                                     void*, void*, void*, void*, hipStream_t);

}  // namespace mnp
