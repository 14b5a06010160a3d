// H-matrix argument bundles (device pointers) shared by hmat.hip and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "args.hpp"

namespace mnp {

constexpr int HM_MAXDIM = 8192;   // rows / columns of one admissible block (ACA pivot masks in LDS)
constexpr int HM_MAXR = 8;        // right-hand sides of one H-matrix product

// Geometry and refined-pair table for Green-function entries at one wavelength.
struct GreenGeom {
    int n, npart, order;
    const double* pos;
    const double* nvec;
    const double* area;
    const int* fpart;
    const int* con_in;
    const int* con_out;
    const double* kmat;     // nmat complex wavenumbers
    const int* rowptr;      // n + 1, CSR of refined pairs
    const int* rcol;        // columns, ascending within a row
    const double* rg;       // nref x (order + 1)
    const double* rf;       // nref x 3 x (order + 1)
};

// Observer points against the faces of a particle (fields at points).
struct PointGeom {
    int npt, nface, npart, order;
    const double* ppos;
    const int* pmed;        // medium of each point (1-based)
    const double* fpos;
    const double* farea;
    const int* fpart;
    const int* inout;       // npart x 2 (inside, outside), 1-based
    const double* kmat;
    const int* rowptr;      // npt + 1, CSR of refined point-face pairs
    const int* rcol;
    const double* rg;
    const double* rf;
};

// Built H-matrices of the four kinds (G1, G2, H1, H2) on one cluster tree.
struct HMat {
    int n, nd, nlr;
    const int* c2p;         // cluster position -> face
    const int* dblk;        // nd x 4 (r0, c0, m, n)
    const long* doff;       // nd complex offsets of dense blocks (same for every kind)
    const void* D;          // complex double, kind q at D + q * dstride
    long dstride;
    const int* lblk;        // nlr x 4
    const long* uoff;       // 4 x nlr complex offsets into U (per kind)
    const long* voff;       // 4 x nlr complex offsets into V
    const void* U;          // complex float, m x rank per block
    const void* V;          // complex float, n x rank per block
    const int* rank;        // nlr x 4
    void* tmp;              // complex double scratch, HM_MAXR x (largest rank of the block) per block
    const long* toff;       // nlr offsets into tmp
};

void op_prep(const Face& f, int batch, const void* x, long sx, void* X1, void* X2, long sX, hipStream_t s);
void op_finish(const Face& f, int batch, const void* GX1, const void* GX2, const void* HX1, long sH1, const void* HX2,
               long sX, const void* rhs, long sr, void* y, long sy, double* norms, hipStream_t s);

void hm_fill_dense(const GreenGeom& a, const int* c2p, int nd, const int* blk, const long* off, void* D, long kstride,
                   hipStream_t s);
void hm_aca(const GreenGeom& a, const int* c2p, int kind, int nlr, const int* lr, const long* uoff,
            const long* voff, void* U, void* V, int* rank, double htol, int kmax, hipStream_t s);
void hm_compact(int nlr, const int* lr, int kind, const int* rank, const long* uws, const long* vws, const void* Uw,
                const void* Vw, const long* uoff, const long* voff, void* U, void* V, hipStream_t s);
void hm_matvec(const HMat& h, int kind, int r, const void* x, long ldx, void* y, long ldy, hipStream_t s);
void bem_operator_hmat(const HMat& h, const Face& f, const void* x, const void* rhs, void* y, void* work,
                       double* norms, hipStream_t s);
void surface_eval(const GreenGeom& a, int outside, double k0, const void* s, const void* hv, void* out,
                  hipStream_t st);
void point_eval(const PointGeom& a, double k0, const void* s1, const void* h1, const void* s2, const void* h2,
                void* out, hipStream_t st);
void green_block(const GreenGeom& a, int outside, int key, int c0, int ncol, void* out, hipStream_t st);
void green_pick(const GreenGeom& a, int outside, int key, const long* ind, long count, void* out, hipStream_t st);

}  // namespace mnp
