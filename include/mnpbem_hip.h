/*
 * libmnpbem_hip.so — C-ABI of the MI355X (gfx950) retarded BEM solve path.
 *
 * The reference (Yoo-JK/MNPBEM) is pure Python; its drop-in boundary for this path is the
 * Python class API (ComParticle, CompGreenRet, BEMRet.solve, PlaneWaveRet.extinction) plus
 * the dense linear-algebra dispatch of mnpbem/utils/gpu.py. Each entry point below replaces
 * a named piece of that path; the Python package mnpbem_amd binds them with ctypes.
 *
 * Conventions
 *   - Array arguments are device pointers (hipMalloc / torch tensors) unless noted "host".
 *   - Complex numbers are interleaved (re, im). Matrices are column-major with leading
 *     dimension ld; a batch holds one matrix per wavelength at `stride` elements.
 *   - prec: 0 = complex double, 1 = complex float.
 *   - Work is enqueued on `stream` (hipStream_t; NULL = default stream) and is asynchronous.
 *   - Return: 0 ok; -k argument k (1-based) rejected before any work was enqueued;
 *     1000 + hipError_t when a launch failed.
 *   - Pivots are 0-based row indices. LU info: 0 ok, j > 0 = zero pivot in column j-1
 *     (LAPACK zgetrf convention; the reference does not detect it, utils/gpu.py:189-191).
 *   - Geometry of a batch: entry points that take `sface` read the per-face arrays (nvec, pos,
 *     area, side) of batch entry b at sface * b faces (times 3 doubles for vectors): 0 = one
 *     geometry shared by every entry (a wavelength sweep), n = one geometry per entry (particles
 *     of different shape padded to one size, mnp_pad_batched).
 *   - State: tuning knobs, grown workspaces (split-BF16 operand planes, cooperative-panel
 *     scratch), the lookahead side stream and GEMM timing records belong to a handle. Every call
 *     works through the handle bound to the calling host thread (mnp_handle_bind); a thread that
 *     binds none has its own default handle. Threads that run at once must not share a handle.
 */
#ifndef MNPBEM_HIP_H
#define MNPBEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ABI revision of this header (3). */
int mnp_abi_version(void);

/* Library handle (opaque). The reference keeps its solver state in Python objects and runs one
 * solve at a time per process (bem/bem_ret.py:248-259); a handle gives each host thread its own
 * knobs, workspaces and side stream, as a per-thread handle replaces the reference's
 * process-global cupy/cuSolver context (utils/gpu.py:145-301). */
typedef struct mnp_handle_s* mnp_handle;
/* New handle with the default knobs (environment overrides MNP_LU_NBO, MNP_GEMM_SPLIT, ...). */
int mnp_handle_create(mnp_handle* h);
/* Waits for the handle's work, frees its workspaces, stream and events, and destroys it; a
 * thread bound to it falls back to its default handle. */
int mnp_handle_destroy(mnp_handle h);
/* Bind h to the calling thread (NULL: the thread's default handle). */
int mnp_handle_bind(mnp_handle h);
/* Waits for the device and frees the workspaces, side streams and events of every handle (they
 * are made again on demand). Meant for process exit: the Python binding registers it with
 * atexit so that nothing of the library is alive when the HIP runtime and tools shut down. */
int mnp_release(void);

/* Test hook: one wave runs v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32 and
 * v_mfma_f32_32x32x2f32 on known operands and writes their result registers to out[0..255],
 * out[256..511] (64 x 4 each) and out[512..1535] (64 x 16) (device doubles). No counterpart in
 * the reference. */
int mnp_probe_mfma_layout(double* out, void* stream);

/* GEMM timing hooks for roofline reports (no counterpart in the reference): while enabled,
 * each matrix-core GEMM launch is bracketed by HIP events on its stream; collect returns
 * out[0] = GEMM kernel ms, out[1] = algorithmic flops (8 m n k per complex GEMM),
 * out[2] = launches since the previous collect, out[3] = algorithmic bytes (A and B read once,
 * C read and written once) (host array of 4; waits for the events). When some launches ran the
 * split-BF16 plane kernel, the sums cover those launches only (the dominant kernel); the rest
 * are dropped. */
int mnp_prof_enable(int on);
int mnp_prof_collect(double* out);

/* Tracing: every compute entry point below opens a roctx range "mnp:<stage>" (assemble, getrf,
 * getrs, trsm_right, gemm, gemv, bemret_factor, bemret_solve, refine_operator, planewave,
 * farfield, near_terms, colsum_far, hmat_dense, aca, hmat_matvec, gmres_operator, surface_eval,
 * point_eval, mlp_gemm) for the time it takes to enqueue its work, visible to rocprofv3
 * --marker-trace. These two let the host wrap its own stages (a GMRES iteration, a refinement
 * step) the same way. The reference has no tracer; its stage timings are perf_counter prints
 * (bem/bem_ret.py:597-601, :931-937) and BEMIter.tocout (bem/bem_iter.py:134-190), which these
 * ranges replace. name: host string. */
int mnp_range_push(const char* name);
int mnp_range_pop(void);

/* G1, G2, H1, H2 of a composite particle for nlam wavelengths.
 * Replaces CompGreenRet.eval(i, i, 'G'|'H1'|'H2', enei) (greenfun/compgreen_ret.py:507-553,
 * :739-858) with GreenRetRefined.eval (greenfun/greenret_refined.py:654-810) per block.
 * pos, nvec: n x 3; area: n; fpart: n particle index per face; con_in, con_out: npart x npart
 * material index (1-based, 0 = not connected) for the inside / outside region pairs;
 * kmat: nlam x nmat complex wavenumbers; refined pairs: ref_row/ref_col (nref), ref_g
 * (nref x (order+1)), ref_f (nref x 3 x (order+1)) Taylor coefficients in (ik).
 * Outputs: column-major n x n matrices at G1 + l*stride etc. for wavelength l. */
int mnp_green_assemble(int n, int npart, int nmat, int nlam, int order, const double* pos, const double* nvec,
                       const double* area, const int* fpart, const int* con_in, const int* con_out,
                       const double* kmat, int nref, const int* ref_row, const int* ref_col, const double* ref_g,
                       const double* ref_f, int prec, void* G1, void* G2, void* H1, void* H2, long ld, long stride,
                       void* stream);

/* The same entries as mnp_green_assemble once in complex double (G1, G2, H1, H2 at ld, stride)
 * and once in complex float (G1s, G2s, H1s, H2s at lds, strides) from one evaluation of each
 * exponential: the exact operator and the factor matrices of a mixed-precision BEMRet sweep.
 * Replaces the same CompGreenRet.eval calls (compgreen_ret.py:739-858, bem_ret.py:407-433). */
int mnp_green_assemble_pair(int n, int npart, int nmat, int nlam, int order, const double* pos, const double* nvec,
                            const double* area, const int* fpart, const int* con_in, const int* con_out,
                            const double* kmat, int nref, const int* ref_row, const int* ref_col, const double* ref_g,
                            const double* ref_f, void* G1, void* G2, void* H1, void* H2, long ld, long stride,
                            void* G1s, void* G2s, void* H1s, void* H2s, long lds, long strides, void* stream);

/* C -= A * B (m x k times k x n), batched. Matrix-core kernel behind the LU and solves.
 * Replaces the cupy/numpy GEMMs of bem_ret.py:472-495 / :844-886. */
int mnp_gemm_sub_batched(int prec, int m, int n, int k, const void* A, long lda, long sA, const void* B, long ldb,
                         long sB, void* C, long ldc, long sC, int batch, void* stream);

/* C -= A B in complex float on the BF16 matrix cores with operands split into `terms` (2 or 3)
 * bfloat16 terms: the kernels and results of mnp_gemm_sub_batched(prec = 1) under tuning key 4
 * = 10 + terms, but allocating nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). The operand planes live in a static device array of
 * mnp_noalloc_scratch_bytes() (32 MiB) per GPU; a launch needs
 * 2 terms ceil(k/16) 16 (ceil(m/128) + ceil(n/128)) 128 batch 2 bytes of it when its C has at
 * least 160 (3 terms) or 96 (2 terms) rows and columns (narrower launches use the FP32 matrix
 * cores and no scratch). Returns 2 without enqueuing work when that exceeds the scratch, 3 when
 * the scratch's device address is unavailable. Calls that overlap on one GPU race on the
 * scratch: serialise them (one stream, or events). */
int mnp_gemm_sub_batched_noalloc(int terms, int m, int n, int k, const void* A, long lda, long sA, const void* B,
                                 long ldb, long sB, void* C, long ldc, long sC, int batch, void* stream);
/* Size in bytes of the static scratch behind mnp_gemm_sub_batched_noalloc (host long). */
int mnp_noalloc_scratch_bytes(long* bytes);

/* In-place LU with partial pivoting of `batch` n x n matrices (panel width nb, 1..64); panels
 * of at least knob 3's height choose their pivots by tournament (knob 9), so for n at or above
 * it the pivots may differ from LAPACK's (P A = L U holds all the same).
 * info[b]: 0 ok, j > 0 zero pivot in column j-1, -1 a multi-workgroup panel's barrier timed out
 * (factors invalid). Replaces lu_factor_dispatch (utils/gpu.py:145-191: scipy zgetrf / cupyx getrf). */
int mnp_getrf_batched(int prec, int n, int batch, void* A, long lda, long sA, int* ipiv, long sP, int* info, int nb,
                      void* stream);

/* Grow the calling thread's workspaces for LUs and right-side solves of `batch` n x n matrices
 * on `stream` (panel width nb, current knobs) so that mnp_getrf_batched, mnp_rsolve_batched and
 * mnp_bemret_factor of that size allocate nothing afterwards, e.g. while the stream is captured
 * into a HIP graph. Returns 4 when that is not possible (n at or above the cooperative-panel
 * threshold, or an allocation failed). No counterpart in the reference. */
int mnp_lu_reserve(int n, int batch, int nb, void* stream);

/* Row permutation pf (P b = b[pf]) and its inverse from the pivots of mnp_getrf_batched
 * (any n). Part of lu_solve_dispatch (utils/gpu.py:194-251). */
int mnp_perm_batched(int n, int batch, const int* ipiv, long sP, int* pf, int* qinv, long sQ, void* stream);

/* X = A^{-1} B for r <= 8 columns from an LU factor (B != X).
 * Replaces lu_solve_dispatch / lu_solve_native (utils/gpu.py:194-301). */
int mnp_getrs_batched(int prec, int n, int r, int batch, const void* LU, long lda, long sA, const int* pf, long sQ,
                      const void* B, long ldb, long sB, void* X, long ldx, long sX, void* stream);

/* X = B A^{-1} (m x n, B overwritten, X separate) from an LU factor of A.
 * Replaces Sigma_i = H_i G_i^{-1} (bem_ret.py:472-473; GPU form :818-819 getrs trans) and the
 * explicit Delta^{-1} (bem_ret.py:476-479, :838). */
int mnp_rsolve_batched(int prec, int m, int n, int batch, const void* LU, long lda, long sA, const int* qinv, long sQ,
                       void* B, long ldb, long sB, void* X, long ldx, long sX, int nb, void* stream);

/* Y = A X for r <= 8 columns. Matrix-vector products of the solve (bem_ret.py:1716-1752). */
int mnp_gemv_batched(int prec, int n, int r, int batch, const void* A, long lda, long sA, const void* X, long ldx,
                     long sX, void* Y, long ldy, long sY, void* stream);

/* M_b = s A + diag(d_b) for b < batch (A: one n x n matrix; d: batch x n complex double).
 * Builds the quasistatic BEM matrix -(F + diag Lambda) of BEMStat._init_matrices
 * (bem/bem_stat.py:241-404) from the k = 0 assembly. */
int mnp_scale_diag_batched(int prec, int n, int batch, double s, const void* A, long lda, const double* d, void* M,
                           long ld, long stride, void* stream);

/* BEMRet.init for a batch of wavelengths (bem_ret.py:293-525): LU(G1), LU(G2),
 * Sigma1 = H1 G1^{-1}, Sigma2 = H2 G2^{-1}, Delta^{-1}, Sigma = Sigma1 L1 - Sigma2 L2 +
 * k^2 L (Delta^{-1} o n n^T) L, LU(Sigma). In: G1, G2, H1, H2; out: G1, G2 (LU), S1 (Sigma1),
 * S2 (LU of Sigma), D (Delta^{-1}); H1, H2 destroyed. ipiv: 4 x batch x n; perm: 8 x batch x n;
 * info: 4 x batch; nvec: n x 3 per geometry (sface); L1, L2: batch x n complex (inside /
 * outside eps per face); kvac: batch (2 pi / lambda). */
int mnp_bemret_factor(int prec, int n, int batch, void* G1, void* G2, void* H1, void* H2, void* S1, void* S2, void* D,
                      long ld, long stride, int* ipiv, int* perm, int* info, const double* nvec, long sface,
                      const double* L1, const double* L2, const double* kvac, int nb, void* stream);

/* Particles of different face counts in one batch (no counterpart in the reference, which solves
 * one particle per process): matrix b holds a particle of nact[b] <= n faces in its leading
 * block; every entry (i, j) with i or j >= nact[b] is set to d = dre + i dim on the diagonal and
 * 0 elsewhere. With G1, G2 padded by 1, H1 by 1, H2 by -1, and L1 = L2 = 1, zero normals and
 * zero excitation on the padding faces, the LU chain of mnp_bemret_factor stays block diagonal
 * and the padding rows of every solution are 0. */
int mnp_pad_batched(int prec, int n, int batch, const int* nact, void* A, long ld, long stride, double dre, double dim,
                    void* stream);

/* BEMRet.solve (bem_ret.py:1569-1831, Eqs. 19-20): in = [phi a De alpha] (n x 8 per
 * wavelength, precision prec) -> out = [sig1 h1 sig2 h2] (n x 8); work: 24 n complex per
 * wavelength. Factors from mnp_bemret_factor. perm NULL: G1, G2 and S2 hold the explicit
 * inverses G1^{-1}, G2^{-1}, Sigma^{-1} (from mnp_rsolve_batched of the identity), applied as
 * products; for a factorisation reused by many solves (a fixed GMRES preconditioner). */
int mnp_bemret_solve(int prec, int n, int batch, const void* G1, const void* G2, const void* S1, const void* D,
                     const void* S2, long ld, long stride, const int* perm, const double* nvec, long sface,
                     const double* L1, const double* L2, const double* kvac, const void* in, long sIn, void* out,
                     long sOut, void* work, long sW, void* stream);

/* Exact complex-double BEM operator y = A x, or residual y = rhs - A x when rhs != NULL
 * (BEMRetIter._afun, bem/bem_ret_iter.py:1362-1501); norms (device, 2 x batch, may be NULL):
 * |y|^2 and |rhs|^2 per wavelength. x != y; work: 32 n complex double per wavelength. */
int mnp_bem_operator(int n, int batch, const void* G1, const void* G2, const void* H1, const void* H2, long ld,
                     long stride, const double* nvec, long sface, const double* L1, const double* L2,
                     const double* kvac, const void* x, long sx, const void* rhs, long sr, void* y, long sy, void* work,
                     long sW, double* norms, void* stream);

/* Plane-wave right-hand side [phi a De alpha] (PlaneWaveRet.potential,
 * simulation/planewave_ret.py:257-361, with BEMRet._excitation, bem_ret.py:1360-1429).
 * side: n flags (bit 0 inside medium = excitation medium, bit 1 outside); kmed: batch complex;
 * pol, dir: host double[3]. */
int mnp_planewave_rhs(int n, int batch, const double* pos, const double* nvec, const int* side, long sface,
                      const double* kmed, const double* kvac, const double* L1, const double* L2, const double* pol,
                      const double* dir, void* out, long so, void* stream);

/* Far fields e, h in ndir directions from x = [sig1 h1 sig2 h2] (SpectrumRet._farfield,
 * spectrum/spectrum_ret.py:231-392); out: batch x ndir x 6 complex. */
int mnp_farfield(int n, int batch, const double* pos, const double* area, const int* side, long sface,
                 const double* kmed, const double* kvac, int ndir, const double* dirs, const void* x, long sx, void* out,
                 void* stream);

/* Arnoldi kernels of the device GMRES (BEMRetIter; the reference calls scipy.sparse.linalg.gmres,
 * bem/bem_iter.py:45-101): h[c] = <V_c, w> and w += sum_c V_c s[c] over nv <= 128 complex double
 * vectors of length len stored ldv apart; h, s: device arrays of 2 nv doubles. */
int mnp_mdot(long len, int nv, const void* V, long ldv, const void* w, double* h, void* stream);
int mnp_mupdate(long len, int nv, const void* V, long ldv, const double* s, void* w, void* stream);

/* dst = cast(src) from complex double (refinement plumbing; no counterpart in the reference). */
int mnp_cast_batched(int prec_dst, long count, int batch, const void* src, long ss, void* dst, long sd, void* stream);

/* dst (complex double) += src where active[b] (refinement update; no counterpart). */
int mnp_accum_batched(int prec_src, long count, int batch, const void* src, long ss, void* dst, long sd,
                      const int* active, void* stream);


/* ---------------------------------------------------------------------------------------------
 * H-matrices (BEMRetIter with hmatrix=True: greenfun/aca_compgreen_ret.py:20-208,
 * greenfun/hmatrix.py:94-541, greenfun/clustertree.py:13-380). The cluster tree and the block
 * partition are built on the host; entries, ACA and products run here.
 *
 * Kinds: 0 = G1, 1 = G2, 2 = H1, 3 = H2. Blocks are (r0, c0, m, n) in cluster order. */

/* Geometry and refined-pair table for on-the-fly Green-function entries at one wavelength
 * (same values as mnp_green_assemble). rowptr/rcol: CSR of the refined pairs with ascending
 * columns per row; rg/rf: their Taylor coefficients in that order; kmat: nmat complex. */
typedef struct mnp_green_geom {
    int n, npart, order;
    const double* pos;
    const double* nvec;
    const double* area;
    const int* fpart;
    const int* con_in;
    const int* con_out;
    const double* kmat;
    const int* rowptr;
    const int* rcol;
    const double* rg;
    const double* rf;
} mnp_green_geom;

/* Observer points and the faces of a particle for fields at points (all device pointers):
 * ppos npt x 3, pmed npt media of the points (1-based material index); fpos nface x 3 face
 * centroids, farea, fpart (particle of each face), inout npart x 2 (inside, outside medium,
 * 1-based); kmat nmat complex wavenumbers; refined point-face pairs as a CSR over the points
 * (rowptr npt + 1, rcol ascending face indices, rg nref x (order + 1), rf nref x 3 x
 * (order + 1) Taylor coefficients). */
typedef struct mnp_point_geom {
    int npt, nface, npart, order;
    const double* ppos;
    const int* pmed;
    const double* fpos;
    const double* farea;
    const int* fpart;
    const int* inout;
    const double* kmat;
    const int* rowptr;
    const int* rcol;
    const double* rg;
    const double* rf;
} mnp_point_geom;

/* A built set of four H-matrices on one tree (all device pointers). D: complex double dense
 * blocks, kind q at D + q dstride, block b at doff[b]; U, V: complex float low-rank factors
 * (m x rank, n x rank) of kind q and block b at uoff[q nlr + b], voff[q nlr + b];
 * rank: nlr x 4; tmp/toff: complex double scratch of 8 x (largest rank of block b) per block. */
typedef struct mnp_hmat {
    int n, nd, nlr;
    const int* c2p;
    const int* dblk;
    const long* doff;
    const void* D;
    long dstride;
    const int* lblk;
    const long* uoff;
    const long* voff;
    const void* U;
    const void* V;
    const int* rank;
    void* tmp;
    const long* toff;
} mnp_hmat;

/* Dense (inadmissible) blocks of all four kinds, entries evaluated on the fly
 * (HMatrix.aca dense part, hmatrix.py:172-180; entries of CompGreenRet.eval). */
int mnp_hmat_fill_dense(const mnp_green_geom* g, const int* c2p, int nd, const int* dblk, const long* doff, void* D,
                        long dstride, void* stream);

/* Partially pivoted ACA of the admissible blocks of one kind (HMatrix._aca_block,
 * hmatrix.py:195-336), one workgroup per block, complex double factors U (m x kr) and
 * V (n x kr), kr = min(m, n, kmax), at uws[b], vws[b]; rank[4 b + kind] receives the rank.
 * Blocks must satisfy m, n <= 8192. */
int mnp_hmat_aca(const mnp_green_geom* g, const int* c2p, int kind, int nlr, const int* lblk, const long* uws,
                 const long* vws, void* U, void* V, int* rank, double htol, int kmax, void* stream);

/* Copy the first rank[4 b + kind] columns of the ACA factors to complex float storage at
 * uoff[b], voff[b] (leading dimensions m and n). No counterpart (the reference keeps complex128). */
int mnp_hmat_compact(int nlr, const int* lblk, int kind, const int* rank, const long* uws, const long* vws,
                     const void* Uws, const void* Vws, const long* uoff, const long* voff, void* U, void* V,
                     void* stream);

/* y += H_kind x for r <= 8 complex double columns in cluster order (HMatrix.mtimes_vec,
 * hmatrix.py:450-541, without its permutations). */
int mnp_hmat_matvec(const mnp_hmat* h, int kind, int r, const void* x, long ldx, void* y, long ldy, void* stream);

/* The packed 8N operator of BEMRetIter._afun (bem_ret_iter.py:1362-1501) on the four
 * H-matrices, one wavelength: y = A x, or y = rhs - A x with norms = (|y|^2, |rhs|^2) when rhs
 * is given. x, rhs, y: 8 n complex double (y != x); work: 64 n complex double. */
int mnp_bem_operator_hmat(const mnp_hmat* h, const double* nvec, const double* L1, const double* L2,
                          const double* kvac, const void* x, const void* rhs, void* y, void* work, long sW,
                          double* norms, void* stream);


/* ---------------------------------------------------------------------------------------------
 * Near-field set-up (wavelength independent). */

/* Taylor coefficients of near pairs (GreenRetRefined._refine_offdiagonal,
 * greenfun/greenret_refined.py:287-516): pair p = (observer rows[p], source cols[p]) is
 * integrated over the source's quadrature points qpts[qptr[j] .. qptr[j+1]) (3 doubles each)
 * with weights qw; results go to g[slot[p] (order+1) + n] and f[(slot[p] 3 + c)(order+1) + n]
 * (order < 8). */
int mnp_near_terms(int npair, const int* rows, const int* cols, const long* slot, const double* pos, const int* qptr,
                   const double* qpts, const double* qw, int order, double* g, double* f, void* stream);

/* Collocated column sums of the quasistatic surface derivative over one closed surface
 * (closed-surface correction, compgreen_stat.py:219-464 / compgreen_ret.py:192-272):
 * out[jj] = sum_{i in loc} -area_i nvec_i.(pos_i - pos_j) / |pos_i - pos_j|^3 for j = ind[jj]. */
int mnp_colsum_far(int nloc, const int* loc, int nind, const int* ind, const double* pos, const double* nvec,
                   const double* area, double* out, void* stream);


/* ---------------------------------------------------------------------------------------------
 * Learned GMRES warm start (no counterpart in the reference: BEMRetIter starts GMRES from zero,
 * bem_ret_iter.py:1695). A per-face MLP; all arrays row-major float, device pointers. */

/* C[M][N] = op(A)[M][K] op(B)[K][N] on the BF16 matrix cores (operands rounded to BF16, FP32
 * accumulation); op = transpose when ta / tb is 1. epi: 0 Z = AB, 1 Z += AB, 2 Z = AB + bias and
 * H = gelu(Z) (erf form), 3 Z = AB + bias (2 and 3 without transposes). */
int mnp_mlp_gemm(int ta, int tb, int epi, int M, int N, int K, const float* A, long lda, const float* B, long ldb,
                 const float* bias, float* Z, float* H, long ldc, void* stream);

/* dZ = dH * gelu'(Z), elementwise over count floats. */
int mnp_mlp_gelu_bwd(long count, const float* dH, const float* Z, float* dZ, void* stream);

/* out[j] = sum_i X[i][j] (bias gradients). */
int mnp_mlp_colsum(int M, int N, const float* X, long ld, float* out, void* stream);

/* loss[0] = mean (Y - T)^2 over count floats, dY = 2 (Y - T) / count. */
int mnp_mlp_mse(long count, const float* Y, const float* T, float* dY, float* loss, void* stream);

/* One Adam step (bias-corrected, `step` >= 1) on count parameters p with gradients g and moments m, v. */
int mnp_mlp_adam(long count, float* p, const float* g, float* m, float* v, double lr, double b1, double b2, double eps,
                 int step, void* stream);

/* Test hook: result registers of one v_mfma_f32_32x32x8bf16_1k on known operands (64 x 16 floats). */
int mnp_probe_mfma_bf16(float* out, void* stream);


/* ---------------------------------------------------------------------------------------------
 * LU tuning knobs of the calling thread's handle (host values; no counterpart in the
 * reference). key 0: outer block width of the blocked LU and right-side solves (16..1024
 * columns, default 128); 1: lookahead (trailing update split over two streams) on/off; 2:
 * register-resident panel for complex float on/off; 3: panel height from which the
 * multi-workgroup panel is used (>= 256, default 4096); 4: complex-float GEMM on the FP32 matrix
 * cores (0, default of the handle) or on the 16-bit matrix cores with operands split once into
 * planes: 2 or 3 bfloat16 terms (12, 13; 2 terms keep about 16 significant bits, 3 about 24, as
 * FP32) or scaled fp16 pairs (16: rows of A and columns of B scaled by powers of two, about 22
 * bits, 3 products per real product); 5: workgroup tile of the bfloat16 plane GEMM (11: 64 x 64
 * staged through registers; 111, 21, 12, 22: 64 x 64, 128 x 64, 64 x 128, 128 x 128 with direct
 * LDS copies; all give bit-identical results); 6: row tiles per group of the plane GEMM's
 * XCD-aware tile order (0: plain grid order); 7: rows per workgroup of the Green-function
 * assembly (64, 128, 256; the workgroup covers 256 / rows columns); 8: wavelengths per assembly
 * thread (0: all of the launch; shorter chunks put more workgroups on small matrices); 9:
 * multi-workgroup panels by tournament pivoting (1, default: each panel's pivot rows chosen by
 * a reduction tree of partial-pivoting eliminations over blocks of 128 (complex float) or 256
 * (complex double) rows, no barrier between workgroups) or by partial pivoting with a
 * cooperative launch (0: LAPACK's pivots, one grid barrier per column); 10: outer blocks
 * factored by recursive halving (1, default: the trailing columns of the block are updated by
 * GEMMs of depth W/2, W/4, ..., nb) or panel by panel (0: W/nb GEMMs of depth nb); 11: fp16-pair
 * operands with K <= 512 scaled and split in one pass over the operand (1, default) or in two
 * (0); both write the same bits.
 * mnpbem_amd.autotune picks them per problem shape by timing. */
int mnp_set_tuning(int key, int value);
int mnp_get_tuning(int key, int* value);


/* Potentials and fields on the faces from a solution, for one region (outside = 0 inside with
 * sigma1/h1, 1 outside with sigma2/h2): CompGreenRet.potential and .field
 * (greenfun/compgreen_ret.py:912-1040) for particles whose regions do not nest. sig: n complex,
 * h: n x 3 complex (row-major), k0: vacuum wavenumber; out: n x 14 complex per face
 * (phi, phip, a[3], ap[3], e[3], h[3]). Entries come from g (mnp_green_geom) on the fly. */
int mnp_surface_eval(const mnp_green_geom* g, int outside, double k0, const void* sig, const void* h, void* out,
                     void* stream);


/* Entries of one region-pair Green function by key (0 G, 1 F, 2 H1, 3 H2, 4 Gp, 5 H1p,
 * 6 H2p) from g on the fly, for the diagonal region pairs of non-nested particles
 * (outside = 0: inside-inside, 1: outside-outside).
 * mnp_green_block: columns [c0, c0 + ncol), row-major (n, ncol) complex, or (n, 3, ncol) for
 *   keys 4-6 -- CompGreenRet.eval_block (greenfun/compgreen_ret.py:555-737) and eval of the
 *   derivative keys (:507-553).
 * mnp_green_pick: entries at row-major linear indices ind[q] = i * n + j (device int64),
 *   out (count) or (count, 3) complex -- CompGreenRet.eval(..., ind) (:860-910). */
int mnp_green_block(const mnp_green_geom* g, int outside, int key, int c0, int ncol, void* out, void* stream);
int mnp_green_pick(const mnp_green_geom* g, int outside, int key, const long* ind, long count, void* out,
                   void* stream);


/* Potentials and fields at points from a solution: CompGreenRet(p1 = ComPoint, p2 =
 * ComParticle).potential / .field (greenfun/compgreen_ret.py:911-1034). A point in medium m
 * sees the faces whose inside medium is m through sig1, h1 and those whose outside medium is m
 * through sig2, h2 (n complex, n x 3 complex row-major). out: npt x 14 complex per point
 * (phi, phip = 0, a[3], ap[3] = 0, e[3], h[3]; bare points carry no normal). */
int mnp_point_eval(const mnp_point_geom* g, double k0, const void* sig1, const void* h1, const void* sig2,
                   const void* h2, void* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MNPBEM_HIP_H */
