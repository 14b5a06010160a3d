// Argument bundles shared by the kernels and the C-ABI layer.
#pragma once

namespace mnp {

// Green-function assembly inputs (all device pointers; kmat is nlam x nmat complex).
struct AsmArgs {
    int n, npart, nmat, nlam, order;
    const double* pos;
    const double* nvec;
    const double* area;
    const int* fpart;
    const int* con_in;
    const int* con_out;
    const double* kmat;
    long ld, stride;
};

// Per-face data of the BEMRet chain: normals, inside/outside dielectric functions (batch x n
// complex double) and vacuum wavenumber per wavelength. sface: faces between the geometry of
// consecutive batch entries (0: one geometry shared by the batch, n: one per entry, as for
// particles of different shape padded to one size).
struct Face {
    int n;
    const double* nvec;
    const double* L1;
    const double* L2;
    const double* kvac;
    long sface;
    __host__ __device__ const double* nv(int b) const { return nvec + (long)b * 3 * sface; }
};

// Plane-wave excitation / far-field inputs; sface as in Face (pos, nvec, side and area advance
// by sface faces per batch entry).
struct PW {
    int n;
    const double* pos;
    const double* nvec;
    const int* side;
    const double* kmed;
    const double* kvac;
    const double* L1;
    const double* L2;
    double pol[3], dir[3];
    long sface;
};

}  // namespace mnp
