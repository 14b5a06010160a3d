// C-ABI of libmnpbem_hip.so: plain pointers and sizes, no framework types.
//
// Conventions (declared in include/mnpbem_hip.h):
//   * every array argument is a device pointer unless stated otherwise;
//   * complex values are interleaved (re, im); matrices are column-major with leading dimension
//     `ld`, one matrix per wavelength at `stride` elements;
//   * `prec` selects the arithmetic of matrices: 0 = complex double, 1 = complex float;
//   * work is enqueued on `stream` (a hipStream_t, 0 = default) and returns immediately;
//   * return value 0 = ok, -k = argument k invalid, 1000 + hipError_t = launch failure.
#include "common.hpp"
#include "args.hpp"
#include "hmat.hpp"
#include "ctx.hpp"
#include "../../include/mnpbem_hip.h"

#include <rocprofiler-sdk-roctx/roctx.h>

namespace mnp {
template <typename OT>
void launch_assemble(const AsmArgs&, int, const int*, const int*, const double*, const double*, void*, void*, void*,
                     void*, hipStream_t);
void launch_assemble_pair(const AsmArgs&, int, const int*, const int*, const double*, const double*, void*, void*,
                          void*, void*, void*, void*, void*, void*, long, long, hipStream_t);
template <typename T>
void gemm_sub(int, int, int, const void*, long, long, const void*, long, long, void*, long, long, int, hipStream_t);
template <typename T>
void getrf_batched(int, int, void*, long, long, int*, long, int*, int, hipStream_t);
int gemm_sub_noalloc(int, int, int, int, const void*, long, long, const void*, long, long, void*, long, long, int,
                     hipStream_t);
size_t noalloc_scratch_bytes();
bool lu_reserve(int, int, int, hipStream_t);
void perm_batched(int, int, const int*, long, int*, int*, long, hipStream_t);
template <typename T>
void rsolve_batched(int, int, int, const void*, long, long, const int*, long, void*, long, long, void*, long, long, int,
                    hipStream_t, bool identity = false);
template <typename T>
void getrs_batched(int, int, int, const void*, long, long, const int*, long, const void*, long, long, void*, long,
                   long, hipStream_t);
template <typename T>
void gemv_batched(int, int, int, const void*, long, long, const void*, long, long, void*, long, long, hipStream_t);
template <typename T>
void sub_batched(int, int, const void*, const void*, void*, long, long, hipStream_t);
template <typename T>
void sigma_batched(int, int, const void*, void*, const void*, long, long, const double*, long, const double*,
                   const double*, const double*, hipStream_t);
void zero_int(long, int*, hipStream_t);
void pad_batched(int, int, int, const int*, void*, long, long, double, double, hipStream_t);
template <typename T>
void identity_batched(int, int, void*, long, long, hipStream_t);
template <typename T>
void scale_diag_batched(int, int, double, const void*, long, const double*, void*, long, long, hipStream_t);
template <typename T>
void solve_chain(const Face&, int, const void*, const void*, const void*, const void*, const void*, long, long,
                 const int*, const int*, const int*, long, const void*, long, void*, long, void*, long, hipStream_t);
void bem_operator(const Face&, int, const void*, const void*, const void*, const void*, long, long, const void*, long,
                  const void*, long, void*, long, void*, long, double*, hipStream_t);
void planewave_rhs(const PW&, int, void*, long, hipStream_t);
void farfield(const PW&, const double*, int, const double*, int, const void*, long, void*, hipStream_t);
template <typename T>
void cast_batched(long, int, const void*, long, void*, long, hipStream_t);
template <typename T>
void accum_batched(long, int, const void*, long, void*, long, const int*, hipStream_t);
void probe_layout(double*, hipStream_t);
void prof_enable(bool);
void mdot(long, int, const void*, long, const void*, double*, hipStream_t);
void mupdate(long, int, const void*, long, const double*, void*, hipStream_t);
void prof_collect(double*, double*, double*, double*);
void near_terms(int, const int*, const int*, const long*, const double*, const int*, const double*, const double*, int,
                double*, double*, hipStream_t);
void colsum_far(int, const int*, int, const int*, const double*, const double*, const double*, double*, hipStream_t);
int mlp_gemm(int, int, int, int, int, int, const float*, long, const float*, long, const float*, float*, float*, long,
             hipStream_t);
void mlp_gelu_bwd(long, const float*, const float*, float*, hipStream_t);
void mlp_colsum(int, int, const float*, long, float*, hipStream_t);
void mlp_mse(long, const float*, const float*, float*, float*, hipStream_t);
void mlp_adam(long, float*, const float*, float*, float*, float, float, float, float, int, hipStream_t);
void mlp_probe(float*, hipStream_t);
int set_tuning(int, int);
int get_tuning(int);
}  // namespace mnp

using namespace mnp;

// roctx range over one entry point (host enqueue time; the kernels it launches carry the same
// correlation in a marker trace). Names: mnp:<stage>.
struct StageRange {
    explicit StageRange(const char* name) { roctxRangePushA(name); }
    ~StageRange() { roctxRangePop(); }
};

static int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 1000 + (int)e;
}
#define REQUIRE(cond, k) \
    do {                 \
        if (!(cond)) return -(k); \
    } while (0)

extern "C" {

int mnp_abi_version(void) { return 3; }

// Handles: knobs, workspaces and the lookahead stream of the library, per host thread.
int mnp_handle_create(mnp_handle* h) {
    REQUIRE(h, 1);
    *h = (mnp_handle)ctx_create();
    return 0;
}

int mnp_handle_destroy(mnp_handle h) {
    REQUIRE(h && ctx_valid((Ctx*)h), 1);
    const bool touched = ((Ctx*)h)->release();
    ctx_destroy((Ctx*)h);
    return touched ? launch_status() : 0;
}

int mnp_handle_bind(mnp_handle h) {
    REQUIRE(!h || ctx_valid((Ctx*)h), 1);
    ctx_bind((Ctx*)h);
    return 0;
}

int mnp_release(void) {
    return ctx_release_all() ? launch_status() : 0;
}

int mnp_probe_mfma_layout(double* out, void* stream) {
    REQUIRE(out, 1);
    probe_layout(out, (hipStream_t)stream);
    return launch_status();
}

// GEMM timing for roofline reports: while enabled, every matrix-core GEMM launch of the calling
// thread's handle is bracketed by HIP events on its stream; meant for benchmarks.
int mnp_prof_enable(int on) {
    prof_enable(on != 0);
    return 0;
}

// out[0] = summed GEMM kernel time (ms), out[1] = algorithmic flops (8 m n k per complex
// GEMM), out[2] = launches, since the previous collect. Host array; waits for the events.
int mnp_range_push(const char* name) {
    REQUIRE(name, 1);
    roctxRangePushA(name);
    return 0;
}

int mnp_range_pop(void) {
    roctxRangePop();
    return 0;
}

int mnp_prof_collect(double* out) {
    REQUIRE(out, 1);
    prof_collect(out, out + 1, out + 2, out + 3);
    return 0;
}

int mnp_green_assemble(int n, int npart, int nmat, int nlam, int order, const double* pos, const double* nvec,
                       const double* area, const int* fpart, const int* con_in, const int* con_out,
                       const double* kmat, int nref, const int* ref_row, const int* ref_col, const double* ref_g,
                       const double* ref_f, int prec, void* G1, void* G2, void* H1, void* H2, long ld, long stride,
                       void* stream) {
    StageRange range_("mnp:assemble");
    REQUIRE(n > 0, 1);
    REQUIRE(npart > 0, 2);
    REQUIRE(nmat > 0, 3);
    REQUIRE(nlam > 0, 4);
    REQUIRE(order >= 0 && order <= 7, 5);
    REQUIRE(pos && nvec && area && fpart && con_in && con_out && kmat, 6);
    REQUIRE(nref >= 0, 13);
    REQUIRE(nref == 0 || (ref_row && ref_col && ref_g && ref_f), 14);
    REQUIRE(prec == 0 || prec == 1, 18);
    REQUIRE(G1 && G2 && H1 && H2, 19);
    REQUIRE(ld >= n, 23);
    REQUIRE(stride >= ld * (long)n || nlam == 1, 24);
    AsmArgs a{n, npart, nmat, nlam, order, pos, nvec, area, fpart, con_in, con_out, kmat, ld, stride};
    hipStream_t s = (hipStream_t)stream;
    if (prec == 0)
        launch_assemble<double>(a, nref, ref_row, ref_col, ref_g, ref_f, G1, G2, H1, H2, s);
    else
        launch_assemble<float>(a, nref, ref_row, ref_col, ref_g, ref_f, G1, G2, H1, H2, s);
    return launch_status();
}

int mnp_green_assemble_pair(int n, int npart, int nmat, int nlam, int order, const double* pos, const double* nvec,
                            const double* area, const int* fpart, const int* con_in, const int* con_out,
                            const double* kmat, int nref, const int* ref_row, const int* ref_col, const double* ref_g,
                            const double* ref_f, void* G1, void* G2, void* H1, void* H2, long ld, long stride,
                            void* G1s, void* G2s, void* H1s, void* H2s, long lds, long strides, void* stream) {
    StageRange range_("mnp:assemble");
    REQUIRE(n > 0, 1);
    REQUIRE(npart > 0, 2);
    REQUIRE(nmat > 0, 3);
    REQUIRE(nlam > 0, 4);
    REQUIRE(order >= 0 && order <= 7, 5);
    REQUIRE(pos && nvec && area && fpart && con_in && con_out && kmat, 6);
    REQUIRE(nref >= 0, 13);
    REQUIRE(nref == 0 || (ref_row && ref_col && ref_g && ref_f), 14);
    REQUIRE(G1 && G2 && H1 && H2, 18);
    REQUIRE(ld >= n, 22);
    REQUIRE(stride >= ld * (long)n || nlam == 1, 23);
    REQUIRE(G1s && G2s && H1s && H2s, 24);
    REQUIRE(lds >= n, 28);
    REQUIRE(strides >= lds * (long)n || nlam == 1, 29);
    AsmArgs a{n, npart, nmat, nlam, order, pos, nvec, area, fpart, con_in, con_out, kmat, ld, stride};
    launch_assemble_pair(a, nref, ref_row, ref_col, ref_g, ref_f, G1, G2, H1, H2, G1s, G2s, H1s, H2s, lds, strides,
                         (hipStream_t)stream);
    return launch_status();
}

int mnp_gemm_sub_batched(int prec, int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb,
                         long sB, void* C, long ldc, long sC, int batch, void* stream) {
    StageRange range_("mnp:gemm");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(m >= 0 && n >= 0 && k >= 0, 2);
    REQUIRE(A && B && C, 5);
    REQUIRE(lda >= (m > 0 ? m : 1) && ldb >= (k > 0 ? k : 1) && ldc >= (m > 0 ? m : 1), 6);
    REQUIRE(batch >= 1, 14);
    if (prec == 0)
        gemm_sub<double>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, (hipStream_t)stream);
    else
        gemm_sub<float>(m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, (hipStream_t)stream);
    return launch_status();
}

int mnp_gemm_sub_batched_noalloc(int terms, int m, int n, int k, const void* A, long lda, long sA, const void* B,
                                 long ldb, long sB, void* C, long ldc, long sC, int batch, void* stream) {
    REQUIRE(terms == 2 || terms == 3, 1);
    REQUIRE(m >= 0 && n >= 0 && k >= 0, 2);
    REQUIRE(A && B && C, 5);
    REQUIRE(lda >= (m > 0 ? m : 1) && ldb >= (k > 0 ? k : 1) && ldc >= (m > 0 ? m : 1), 6);
    REQUIRE(batch >= 1, 14);
    const int st = gemm_sub_noalloc(terms, m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, (hipStream_t)stream);
    return st ? st : launch_status();
}

int mnp_lu_reserve(int n, int batch, int nb, void* stream) {
    REQUIRE(n > 0, 1);
    REQUIRE(batch >= 1, 2);
    REQUIRE(nb >= 1 && nb <= 64, 3);
    return lu_reserve(n, batch, nb, (hipStream_t)stream) ? launch_status() : 4;
}

int mnp_noalloc_scratch_bytes(long* bytes) {
    REQUIRE(bytes, 1);
    *bytes = (long)noalloc_scratch_bytes();
    return 0;
}

int mnp_getrf_batched(int prec, int n, int batch, void* A, long lda, long sA, int* ipiv, long sP, int* info, int nb,
                      void* stream) {
    StageRange range_("mnp:getrf");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(A, 4);
    REQUIRE(lda >= n, 5);
    REQUIRE(ipiv && info, 7);
    REQUIRE(sP >= n, 8);
    REQUIRE(nb >= 1 && nb <= 64, 10);
    hipStream_t s = (hipStream_t)stream;
    zero_int(batch, info, s);
    if (prec == 0)
        getrf_batched<double>(n, batch, A, lda, sA, ipiv, sP, info, nb, s);
    else
        getrf_batched<float>(n, batch, A, lda, sA, ipiv, sP, info, nb, s);
    return launch_status();
}

int mnp_perm_batched(int n, int batch, const int* ipiv, long sP, int* pf, int* qinv, long sQ, void* stream) {
    REQUIRE(n > 0, 1);
    REQUIRE(batch >= 1, 2);
    REQUIRE(ipiv && pf && qinv, 3);
    perm_batched(n, batch, ipiv, sP, pf, qinv, sQ, (hipStream_t)stream);
    return launch_status();
}

int mnp_getrs_batched(int prec, int n, int r, int batch, const void* LU, long lda, long sA, const int* pf, long sQ,
                      const void* B, long ldb, long sB, void* X, long ldx, long sX, void* stream) {
    StageRange range_("mnp:getrs");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(r >= 1 && r <= 8, 3);
    REQUIRE(batch >= 1, 4);
    REQUIRE(LU && pf && B && X, 5);
    REQUIRE(B != X, 11);
    if (prec == 0)
        getrs_batched<double>(n, r, batch, LU, lda, sA, pf, sQ, B, ldb, sB, X, ldx, sX, (hipStream_t)stream);
    else
        getrs_batched<float>(n, r, batch, LU, lda, sA, pf, sQ, B, ldb, sB, X, ldx, sX, (hipStream_t)stream);
    return launch_status();
}

int mnp_rsolve_batched(int prec, int m, int n, int batch, const void* LU, long lda, long sA, const int* qinv, long sQ,
                       void* B, long ldb, long sB, void* X, long ldx, long sX, int nb, void* stream) {
    StageRange range_("mnp:trsm_right");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(m > 0, 2);
    REQUIRE(n > 0, 3);
    REQUIRE(batch >= 1, 4);
    REQUIRE(LU && qinv && B && X, 5);
    REQUIRE(B != X, 13);
    REQUIRE(nb >= 1 && nb <= 64, 16);
    if (prec == 0)
        rsolve_batched<double>(m, n, batch, LU, lda, sA, qinv, sQ, B, ldb, sB, X, ldx, sX, nb, (hipStream_t)stream);
    else
        rsolve_batched<float>(m, n, batch, LU, lda, sA, qinv, sQ, B, ldb, sB, X, ldx, sX, nb, (hipStream_t)stream);
    return launch_status();
}

int mnp_gemv_batched(int prec, int n, int r, int batch, const void* A, long lda, long sA, const void* X, long ldx,
                     long sX, void* Y, long ldy, long sY, void* stream) {
    StageRange range_("mnp:gemv");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(r >= 1 && r <= 8, 3);
    REQUIRE(batch >= 1, 4);
    REQUIRE(A && X && Y, 5);
    if (prec == 0)
        gemv_batched<double>(n, r, batch, A, lda, sA, X, ldx, sX, Y, ldy, sY, (hipStream_t)stream);
    else
        gemv_batched<float>(n, r, batch, A, lda, sA, X, ldx, sX, Y, ldy, sY, (hipStream_t)stream);
    return launch_status();
}

// M_b = s A + diag(d_b), b < batch: A one n x n matrix (lda), d batch x n complex double,
// M batch matrices (ld, stride).
int mnp_scale_diag_batched(int prec, int n, int batch, double s, const void* A, long lda, const double* d, void* M,
                           long ld, long stride, void* stream) {
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(A && d && M, 5);
    REQUIRE(lda >= n && ld >= n, 6);
    if (prec == 0)
        scale_diag_batched<double>(n, batch, s, A, lda, d, M, ld, stride, (hipStream_t)stream);
    else
        scale_diag_batched<float>(n, batch, s, A, lda, d, M, ld, stride, (hipStream_t)stream);
    return launch_status();
}

// Padding of matrices whose particle has fewer faces than the batch's n (see the header).
int mnp_pad_batched(int prec, int n, int batch, const int* nact, void* A, long ld, long stride, double dre, double dim,
                    void* stream) {
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(batch >= 1 && batch <= 65535, 3);
    REQUIRE(nact, 4);
    REQUIRE(A, 5);
    REQUIRE(ld >= n, 6);
    REQUIRE(stride >= ld * (long)n || batch == 1, 7);
    pad_batched(prec, n, batch, nact, A, ld, stride, dre, dim, (hipStream_t)stream);
    return launch_status();
}

// Factor the BEMRet system for a batch of wavelengths. Buffers (each `batch` matrices of n x n):
//   G1, G2 (in: Green matrices; out: LU factors), H1, H2 (in: H matrices; destroyed),
//   S1 (out: Sigma1), S2 (out: LU of Sigma), D (out: Delta^{-1}).
// ipiv: 4 x batch x n (G1, G2, Delta, Sigma); perm: 4 x 2 x batch x n (pf, qinv per factor);
// info: 4 x batch. nvec: n x 3 per geometry, sface faces apart (0: shared); L1, L2: batch x n
// complex double; kvac: batch.
int mnp_bemret_factor(int prec, int n, int batch, void* G1, void* G2, void* H1, void* H2, void* S1, void* S2, void* D,
                      long ld, long stride, int* ipiv, int* perm, int* info, const double* nvec, long sface,
                      const double* L1, const double* L2, const double* kvac, int nb, void* stream) {
    StageRange range_("mnp:bemret_factor");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(G1 && G2 && H1 && H2 && S1 && S2 && D, 4);
    REQUIRE(ld >= n, 11);
    REQUIRE(stride >= ld * (long)n, 12);
    REQUIRE(ipiv && perm && info, 13);
    REQUIRE(nvec && L1 && L2 && kvac, 16);
    REQUIRE(sface == 0 || sface >= n, 17);
    REQUIRE(nb >= 1 && nb <= 64, 21);
    hipStream_t s = (hipStream_t)stream;
    const long sP = n, sQ = n;
    const long PB = (long)batch * n;  // ints per factor
    int *ip1 = ipiv, *ip2 = ipiv + PB, *ipD = ipiv + 2 * PB, *ipS = ipiv + 3 * PB;
    int *pf1 = perm, *q1 = perm + PB, *pf2 = perm + 2 * PB, *q2 = perm + 3 * PB;
    int *pfD = perm + 4 * PB, *qD = perm + 5 * PB, *pfS = perm + 6 * PB, *qS = perm + 7 * PB;
    zero_int(4L * batch, info, s);
    auto run = [&](auto tag) {
        using T = decltype(tag);
        getrf_batched<T>(n, batch, G1, ld, stride, ip1, sP, info, nb, s);
        getrf_batched<T>(n, batch, G2, ld, stride, ip2, sP, info + batch, nb, s);
        perm_batched(n, batch, ip1, sP, pf1, q1, sQ, s);
        perm_batched(n, batch, ip2, sP, pf2, q2, sQ, s);
        rsolve_batched<T>(n, n, batch, G1, ld, stride, q1, sQ, H1, ld, stride, S1, ld, stride, nb, s);
        rsolve_batched<T>(n, n, batch, G2, ld, stride, q2, sQ, H2, ld, stride, S2, ld, stride, nb, s);
        sub_batched<T>(n, batch, S1, S2, H1, ld, stride, s);  // Delta into H1
        getrf_batched<T>(n, batch, H1, ld, stride, ipD, sP, info + 2 * batch, nb, s);
        perm_batched(n, batch, ipD, sP, pfD, qD, sQ, s);
        identity_batched<T>(n, batch, H2, ld, stride, s);
        rsolve_batched<T>(n, n, batch, H1, ld, stride, qD, sQ, H2, ld, stride, D, ld, stride, nb, s, true);
        sigma_batched<T>(n, batch, S1, S2, D, ld, stride, nvec, sface, L1, L2, kvac, s);
        getrf_batched<T>(n, batch, S2, ld, stride, ipS, sP, info + 3 * batch, nb, s);
        perm_batched(n, batch, ipS, sP, pfS, qS, sQ, s);
    };
    if (prec == 0)
        run(double(0));
    else
        run(float(0));
    return launch_status();
}

// Solve for packed right-hand sides IN (n x 8 per wavelength: phi a De alpha) -> OUT (n x 8:
// sig1 h1 sig2 h2), with the factors of mnp_bemret_factor. IN/OUT/work are of the factor's
// precision; work holds 24 n complex values per wavelength.
int mnp_bemret_solve(int prec, int n, int batch, const void* G1, const void* G2, const void* S1, const void* D,
                     const void* S2, long ld, long stride, const int* perm, const double* nvec, long sface,
                     const double* L1, const double* L2, const double* kvac, const void* in, long sIn, void* out,
                     long sOut, void* work, long sW, void* stream) {
    StageRange range_("mnp:bemret_solve");
    REQUIRE(prec == 0 || prec == 1, 1);
    REQUIRE(n > 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(G1 && G2 && S1 && D && S2, 4);
    REQUIRE(nvec && L1 && L2 && kvac, 12);
    REQUIRE(sface == 0 || sface >= n, 13);
    REQUIRE(in && out && work, 17);
    REQUIRE(sIn >= 8L * n && sOut >= 8L * n, 18);
    REQUIRE(sW >= 24L * n, 22);
    const long PB = (long)batch * n;
    Face f{n, nvec, L1, L2, kvac, sface};
    hipStream_t s = (hipStream_t)stream;
    // perm null: G1, G2 and S2 are the explicit inverses of G1, G2 and Sigma
    const int* pf1 = perm ? perm : nullptr;
    const int* pf2 = perm ? perm + 2 * PB : nullptr;
    const int* pfS = perm ? perm + 6 * PB : nullptr;
    if (prec == 0)
        solve_chain<double>(f, batch, G1, G2, S1, D, S2, ld, stride, pf1, pf2, pfS, n, in, sIn, out, sOut, work, sW,
                            s);
    else
        solve_chain<float>(f, batch, G1, G2, S1, D, S2, ld, stride, pf1, pf2, pfS, n, in, sIn, out, sOut, work, sW, s);
    return launch_status();
}

// y = A x (rhs == NULL) or y = rhs - A x, with A the exact complex-double BEM operator built
// from G1, G2, H1, H2. x, rhs, y: n x 8 per wavelength; work: 32 n per wavelength. When
// `norms` is given, norms[2b] = |y_b|^2 and norms[2b+1] = |rhs_b|^2 (device, double).
int mnp_bem_operator(int n, int batch, const void* G1, const void* G2, const void* H1, const void* H2, long ld,
                     long stride, const double* nvec, long sface, const double* L1, const double* L2,
                     const double* kvac, const void* x, long sx, const void* rhs, long sr, void* y, long sy, void* work,
                     long sW, double* norms, void* stream) {
    StageRange range_("mnp:refine_operator");
    REQUIRE(n > 0, 1);
    REQUIRE(batch >= 1, 2);
    REQUIRE(G1 && G2 && H1 && H2, 3);
    REQUIRE(nvec && L1 && L2 && kvac, 9);
    REQUIRE(sface == 0 || sface >= n, 10);
    REQUIRE(x && y && work, 14);
    REQUIRE(x != y, 18);
    REQUIRE(sW >= 32L * n, 21);
    Face f{n, nvec, L1, L2, kvac, sface};
    bem_operator(f, batch, G1, G2, H1, H2, ld, stride, x, sx, rhs, sr, y, sy, work, sW, norms, (hipStream_t)stream);
    return launch_status();
}

// Packed plane-wave right-hand side (n x 8 complex double per wavelength). side[i]: bit 0 =
// inside medium of face i is the excitation medium, bit 1 = outside medium is. kmed: complex
// wavenumber in that medium per wavelength; pol, dir: host arrays of 3 doubles.
int mnp_planewave_rhs(int n, int batch, const double* pos, const double* nvec, const int* side, long sface,
                      const double* kmed, const double* kvac, const double* L1, const double* L2, const double* pol,
                      const double* dir, void* out, long so, void* stream) {
    StageRange range_("mnp:planewave");
    REQUIRE(n > 0, 1);
    REQUIRE(batch >= 1, 2);
    REQUIRE(pos && nvec && side && kmed && kvac && L1 && L2, 3);
    REQUIRE(sface == 0 || sface >= n, 6);
    REQUIRE(pol && dir, 11);
    REQUIRE(out, 13);
    REQUIRE(so >= 8L * n, 14);
    PW p{n, pos, nvec, side, kmed, kvac, L1, L2, {pol[0], pol[1], pol[2]}, {dir[0], dir[1], dir[2]}, sface};
    planewave_rhs(p, batch, out, so, (hipStream_t)stream);
    return launch_status();
}

// Far fields for ndir directions: out[(b * ndir + d) * 6 + c] = e_c (c < 3) and h_{c-3}.
int mnp_farfield(int n, int batch, const double* pos, const double* area, const int* side, long sface,
                 const double* kmed, const double* kvac, int ndir, const double* dirs, const void* x, long sx, void* out,
                 void* stream) {
    StageRange range_("mnp:farfield");
    REQUIRE(n > 0, 1);
    REQUIRE(batch >= 1 && batch <= 65535, 2);
    REQUIRE(pos && area && side && kmed && kvac, 3);
    REQUIRE(sface == 0 || sface >= n, 6);
    REQUIRE(ndir >= 1, 9);
    REQUIRE(dirs && x && out, 10);
    PW p{n, pos, nullptr, side, kmed, kvac, nullptr, nullptr, {0, 0, 0}, {0, 0, 0}, sface};
    farfield(p, area, ndir, dirs, batch, x, sx, out, (hipStream_t)stream);
    return launch_status();
}

// h[c] = sum_e conj(V_c[e]) w[e] for c < nv (complex double vectors of length len, basis
// vectors ldv apart); h: device, 2 nv doubles (re, im), overwritten. nv <= 128.
int mnp_mdot(long len, int nv, const void* V, long ldv, const void* w, double* h, void* stream) {
    REQUIRE(len > 0, 1);
    REQUIRE(nv >= 1 && nv <= 128, 2);
    REQUIRE(V && w && h, 3);
    REQUIRE(ldv >= len || nv == 1, 4);
    mdot(len, nv, V, ldv, w, h, (hipStream_t)stream);
    return launch_status();
}

// w[e] += sum_c V_c[e] s[c] for c < nv; s: device, 2 nv doubles (complex). nv <= 128.
int mnp_mupdate(long len, int nv, const void* V, long ldv, const double* s, void* w, void* stream) {
    REQUIRE(len > 0, 1);
    REQUIRE(nv >= 1 && nv <= 128, 2);
    REQUIRE(V && s && w, 3);
    REQUIRE(ldv >= len || nv == 1, 4);
    mupdate(len, nv, V, ldv, s, w, (hipStream_t)stream);
    return launch_status();
}

// dst = (float or double) src for complex double src (count complex values per wavelength).
int mnp_cast_batched(int prec_dst, long count, int batch, const void* src, long ss, void* dst, long sd, void* stream) {
    REQUIRE(prec_dst == 0 || prec_dst == 1, 1);
    REQUIRE(count >= 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(src && dst, 4);
    if (count == 0) return 0;
    if (prec_dst == 1)
        cast_batched<float>(count, batch, src, ss, dst, sd, (hipStream_t)stream);
    else
        cast_batched<double>(count, batch, src, ss, dst, sd, (hipStream_t)stream);
    return launch_status();
}

// dst (complex double) += src (prec_src) for wavelengths with active[b] != 0 (active may be NULL).
int mnp_accum_batched(int prec_src, long count, int batch, const void* src, long ss, void* dst, long sd,
                      const int* active, void* stream) {
    REQUIRE(prec_src == 0 || prec_src == 1, 1);
    REQUIRE(count >= 0, 2);
    REQUIRE(batch >= 1, 3);
    REQUIRE(src && dst, 4);
    if (count == 0) return 0;
    if (prec_src == 1)
        accum_batched<float>(count, batch, src, ss, dst, sd, active, (hipStream_t)stream);
    else
        accum_batched<double>(count, batch, src, ss, dst, sd, active, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- H-matrices
static GreenGeom to_geom(const mnp_green_geom* g) {
    return GreenGeom{g->n,     g->npart,   g->order, g->pos,  g->nvec, g->area, g->fpart,
                     g->con_in, g->con_out, g->kmat,  g->rowptr, g->rcol, g->rg,   g->rf};
}
static HMat to_hmat(const mnp_hmat* h) {
    return HMat{h->n,    h->nd,   h->nlr,  h->c2p,  h->dblk, h->doff, h->D,    h->dstride,
                h->lblk, h->uoff, h->voff, h->U,    h->V,    h->rank, h->tmp,  h->toff};
}
static bool geom_ok(const mnp_green_geom* g) {
    return g && g->n > 0 && g->npart > 0 && g->order >= 0 && g->order < 8 && g->pos && g->nvec && g->area &&
           g->fpart && g->con_in && g->con_out && g->kmat && g->rowptr && g->rcol && g->rg && g->rf;
}

int mnp_hmat_fill_dense(const mnp_green_geom* g, const int* c2p, int nd, const int* dblk, const long* doff, void* D,
                        long dstride, void* stream) {
    StageRange range_("mnp:hmat_dense");
    REQUIRE(geom_ok(g), 1);
    REQUIRE(c2p, 2);
    REQUIRE(nd >= 0, 3);
    REQUIRE(nd == 0 || (dblk && doff && D), 4);
    REQUIRE(dstride >= 0, 7);
    hm_fill_dense(to_geom(g), c2p, nd, dblk, doff, D, dstride, (hipStream_t)stream);
    return launch_status();
}

int mnp_hmat_aca(const mnp_green_geom* g, const int* c2p, int kind, int nlr, const int* lblk, const long* uws,
                 const long* vws, void* U, void* V, int* rank, double htol, int kmax, void* stream) {
    StageRange range_("mnp:aca");
    REQUIRE(geom_ok(g), 1);
    REQUIRE(c2p, 2);
    REQUIRE(kind >= 0 && kind < 4, 3);
    REQUIRE(nlr >= 0, 4);
    REQUIRE(nlr == 0 || (lblk && uws && vws && U && V && rank), 5);
    REQUIRE(htol > 0, 11);
    REQUIRE(kmax >= 1, 12);
    hm_aca(to_geom(g), c2p, kind, nlr, lblk, uws, vws, U, V, rank, htol, kmax, (hipStream_t)stream);
    return launch_status();
}

int mnp_hmat_compact(int nlr, const int* lblk, int kind, const int* rank, const long* uws, const long* vws,
                     const void* Uws, const void* Vws, const long* uoff, const long* voff, void* U, void* V,
                     void* stream) {
    REQUIRE(nlr >= 0, 1);
    REQUIRE(kind >= 0 && kind < 4, 3);
    REQUIRE(nlr == 0 || (lblk && rank && uws && vws && Uws && Vws && uoff && voff && U && V), 2);
    hm_compact(nlr, lblk, kind, rank, uws, vws, Uws, Vws, uoff, voff, U, V, (hipStream_t)stream);
    return launch_status();
}

static bool hmat_ok(const mnp_hmat* h) {
    return h && h->n > 0 && h->nd >= 0 && h->nlr >= 0 && h->c2p && (h->nd == 0 || (h->dblk && h->doff && h->D)) &&
           (h->nlr == 0 || (h->lblk && h->uoff && h->voff && h->U && h->V && h->rank && h->tmp && h->toff));
}

int mnp_hmat_matvec(const mnp_hmat* h, int kind, int r, const void* x, long ldx, void* y, long ldy, void* stream) {
    StageRange range_("mnp:hmat_matvec");
    REQUIRE(hmat_ok(h), 1);
    REQUIRE(kind >= 0 && kind < 4, 2);
    REQUIRE(r >= 1 && r <= HM_MAXR, 3);
    REQUIRE(x, 4);
    REQUIRE(ldx >= h->n, 5);
    REQUIRE(y && y != x, 6);
    REQUIRE(ldy >= h->n, 7);
    hm_matvec(to_hmat(h), kind, r, x, ldx, y, ldy, (hipStream_t)stream);
    return launch_status();
}

int mnp_bem_operator_hmat(const mnp_hmat* h, const double* nvec, const double* L1, const double* L2,
                          const double* kvac, const void* x, const void* rhs, void* y, void* work, long sW,
                          double* norms, void* stream) {
    StageRange range_("mnp:gmres_operator");
    REQUIRE(hmat_ok(h), 1);
    REQUIRE(nvec && L1 && L2 && kvac, 2);
    REQUIRE(x, 6);
    REQUIRE(y && y != x, 8);
    REQUIRE(work, 9);
    REQUIRE(sW >= 64L * h->n, 10);
    Face f{h->n, nvec, L1, L2, kvac};
    bem_operator_hmat(to_hmat(h), f, x, rhs, y, work, norms, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- near-field set-up
int mnp_near_terms(int npair, const int* rows, const int* cols, const long* slot, const double* pos, const int* qptr,
                   const double* qpts, const double* qw, int order, double* g, double* f, void* stream) {
    StageRange range_("mnp:near_terms");
    REQUIRE(npair >= 0, 1);
    REQUIRE(npair == 0 || (rows && cols && slot), 2);
    REQUIRE(pos && qptr && qpts && qw, 5);
    REQUIRE(order >= 0 && order < 8, 9);
    REQUIRE(g && f, 10);
    near_terms(npair, rows, cols, slot, pos, qptr, qpts, qw, order, g, f, (hipStream_t)stream);
    return launch_status();
}

int mnp_colsum_far(int nloc, const int* loc, int nind, const int* ind, const double* pos, const double* nvec,
                   const double* area, double* out, void* stream) {
    StageRange range_("mnp:colsum_far");
    REQUIRE(nloc >= 0, 1);
    REQUIRE(nloc == 0 || loc, 2);
    REQUIRE(nind >= 0, 3);
    REQUIRE(nind == 0 || (ind && out), 4);
    REQUIRE(pos && nvec && area, 5);
    colsum_far(nloc, loc, nind, ind, pos, nvec, area, out, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- warm-start MLP
int mnp_mlp_gemm(int ta, int tb, int epi, int M, int N, int K, const float* A, long lda, const float* B, long ldb,
                 const float* bias, float* Z, float* H, long ldc, void* stream) {
    StageRange range_("mnp:mlp_gemm");
    REQUIRE(ta == 0 || ta == 1, 1);
    REQUIRE(tb == 0 || tb == 1, 2);
    REQUIRE(epi >= 0 && epi <= 3 && !(epi >= 2 && (ta || tb)), 3);
    REQUIRE(M > 0 && N > 0 && K > 0, 4);
    REQUIRE(A && lda >= (ta ? M : K), 7);
    REQUIRE(B && ldb >= (tb ? K : N), 9);
    REQUIRE(epi < 2 || bias, 11);
    REQUIRE(Z && ldc >= N, 12);
    REQUIRE(epi != 2 || H, 13);
    if (mlp_gemm(ta, tb, epi, M, N, K, A, lda, B, ldb, bias, Z, H, ldc, (hipStream_t)stream) != 0) return -3;
    return launch_status();
}

int mnp_mlp_gelu_bwd(long count, const float* dH, const float* Z, float* dZ, void* stream) {
    REQUIRE(count > 0, 1);
    REQUIRE(dH && Z && dZ, 2);
    mlp_gelu_bwd(count, dH, Z, dZ, (hipStream_t)stream);
    return launch_status();
}

int mnp_mlp_colsum(int M, int N, const float* X, long ld, float* out, void* stream) {
    REQUIRE(M > 0 && N > 0, 1);
    REQUIRE(X && ld >= N, 3);
    REQUIRE(out, 5);
    mlp_colsum(M, N, X, ld, out, (hipStream_t)stream);
    return launch_status();
}

int mnp_mlp_mse(long count, const float* Y, const float* T, float* dY, float* loss, void* stream) {
    REQUIRE(count > 0, 1);
    REQUIRE(Y && T && dY && loss, 2);
    mlp_mse(count, Y, T, dY, loss, (hipStream_t)stream);
    return launch_status();
}

int mnp_mlp_adam(long count, float* p, const float* g, float* m, float* v, double lr, double b1, double b2, double eps,
                 int step, void* stream) {
    REQUIRE(count > 0, 1);
    REQUIRE(p && g && m && v, 2);
    REQUIRE(lr > 0 && b1 >= 0 && b1 < 1 && b2 >= 0 && b2 < 1 && eps > 0, 6);
    REQUIRE(step >= 1, 10);
    mlp_adam(count, p, g, m, v, (float)lr, (float)b1, (float)b2, (float)eps, step, (hipStream_t)stream);
    return launch_status();
}

int mnp_probe_mfma_bf16(float* out, void* stream) {
    REQUIRE(out, 1);
    mlp_probe(out, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- tuning knobs
int mnp_set_tuning(int key, int value) {
    REQUIRE(key >= 0 && key <= 11, 1);
    return set_tuning(key, value) == 0 ? 0 : -2;
}

int mnp_get_tuning(int key, int* value) {
    REQUIRE(key >= 0 && key <= 11, 1);
    REQUIRE(value, 2);
    *value = get_tuning(key);
    return 0;
}


// ------------------------------------------------------------------------- surface potentials / fields
int mnp_surface_eval(const mnp_green_geom* g, int outside, double k0, const void* sig, const void* h, void* out,
                     void* stream) {
    StageRange range_("mnp:surface_eval");
    REQUIRE(geom_ok(g), 1);
    REQUIRE(outside == 0 || outside == 1, 2);
    REQUIRE(k0 > 0, 3);
    REQUIRE(sig && h && out, 4);
    surface_eval(to_geom(g), outside, k0, sig, h, out, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- entries by key
int mnp_green_block(const mnp_green_geom* g, int outside, int key, int c0, int ncol, void* out, void* stream) {
    REQUIRE(geom_ok(g), 1);
    REQUIRE(outside == 0 || outside == 1, 2);
    REQUIRE(key >= 0 && key <= 6, 3);
    REQUIRE(c0 >= 0 && ncol >= 0 && (long)c0 + ncol <= g->n, 4);
    REQUIRE(out || ncol == 0, 6);
    green_block(to_geom(g), outside, key, c0, ncol, out, (hipStream_t)stream);
    return launch_status();
}

int mnp_green_pick(const mnp_green_geom* g, int outside, int key, const long* ind, long count, void* out,
                   void* stream) {
    REQUIRE(geom_ok(g), 1);
    REQUIRE(outside == 0 || outside == 1, 2);
    REQUIRE(key >= 0 && key <= 6, 3);
    REQUIRE(count >= 0, 5);
    REQUIRE((ind && out) || count == 0, 4);
    green_pick(to_geom(g), outside, key, ind, count, out, (hipStream_t)stream);
    return launch_status();
}


// ------------------------------------------------------------------------- fields at points
int mnp_point_eval(const mnp_point_geom* g, double k0, const void* sig1, const void* h1, const void* sig2,
                   const void* h2, void* out, void* stream) {
    StageRange range_("mnp:point_eval");
    REQUIRE(g && g->npt >= 0 && g->nface > 0 && g->npart > 0 && g->order >= 0 && g->order < 8, 1);
    REQUIRE(g->ppos && g->pmed && g->fpos && g->farea && g->fpart && g->inout && g->kmat && g->rowptr, 1);
    REQUIRE(k0 > 0, 2);
    REQUIRE(sig1 && h1 && sig2 && h2, 3);
    REQUIRE(out || g->npt == 0, 7);
    PointGeom a{g->npt, g->nface, g->npart, g->order, g->ppos, g->pmed, g->fpos, g->farea, g->fpart, g->inout,
                g->kmat, g->rowptr, g->rcol, g->rg, g->rf};
    point_eval(a, k0, sig1, h1, sig2, h2, out, (hipStream_t)stream);
    return launch_status();
}

}  // extern "C"
