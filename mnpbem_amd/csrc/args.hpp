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
// complex double) and vacuum wavenumber per wavelength.
struct Face {
    int n;
    const double* nvec;
    const double* L1;
    const double* L2;
    const double* kvac;
};

// Plane-wave excitation / far-field inputs.
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
};

}  // namespace mnp
