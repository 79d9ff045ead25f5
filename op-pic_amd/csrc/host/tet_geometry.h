// Tet geometry shared by the host mesh builders (box meshes, mesh files): the FEM quantities the
// mover consumes, computed as the reference mesh loader computes them (apps/app_fempic/minifempic_funcs.h:
// det3/det4 :69-103, volume and the ABCD -> ADBC flip :284-313, alpha/beta/gamma/delta :316-380), and
// the mover's inside test. Built with -ffp-contract=off: every operation is one IEEE operation unless
// it is an explicit std::fma.
#pragma once

#include <cmath>
#include <cstdint>
#include <utility>

namespace oppx_host {

// Determinants as written in minifempic_funcs.h:69-103 (cofactor expansion).
inline double det3(const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

inline double det4(const double M[4][4]) {
    double M0[3][3], M1[3][3], M2[3][3], M3[3][3];
    for (int i = 0; i < 3; i++) {
        M0[i][0] = M[i + 1][1]; M0[i][1] = M[i + 1][2]; M0[i][2] = M[i + 1][3];
        M1[i][0] = M[i + 1][0]; M1[i][1] = M[i + 1][2]; M1[i][2] = M[i + 1][3];
        M2[i][0] = M[i + 1][0]; M2[i][1] = M[i + 1][1]; M2[i][2] = M[i + 1][3];
        M3[i][0] = M[i + 1][0]; M3[i][1] = M[i + 1][1]; M3[i][2] = M[i + 1][2];
    }
    return M[0][0] * det3(M0) - M[0][1] * det3(M1) + M[0][2] * det3(M2) - M[0][3] * det3(M3);
}

// The same determinants with the FMA contractions the reference's GCC -O3 build of the mesh loader
// forms (x86-64 with FMA): within a - b, the first product is fused with the second subtracted,
// fma(a1, a2, -(b1 * b2)); a following + p or - p fuses p into the running sum. Used by the mesh-file
// loader, whose arrays must equal the reference loader's bit for bit.
inline double det3_fma(const double M[3][3]) {
    const double x = std::fma(M[1][1], M[2][2], -(M[1][2] * M[2][1]));
    const double y = std::fma(M[1][0], M[2][2], -(M[1][2] * M[2][0]));
    const double z = std::fma(M[1][0], M[2][1], -(M[1][1] * M[2][0]));
    return std::fma(M[0][2], z, std::fma(M[0][0], x, -(M[0][1] * y)));
}

// det3_fma of a matrix whose first column is all ones, as the loader evaluates the beta, gamma and
// delta determinants (the ones are literal constants there): the build folds the products by 1 away,
// so two differences become plain subtractions and the second product fuses into the first term.
inline double det3_fma_ones(const double M[3][3]) {
    const double x = std::fma(M[1][1], M[2][2], -(M[1][2] * M[2][1]));
    const double y = M[2][2] - M[1][2];
    const double z = M[2][1] - M[1][1];
    return std::fma(M[0][2], z, std::fma(-M[0][1], y, x));
}

// det4 of the loader's volume matrix (rows (1, x, y, z)) as the reference build evaluates it: the
// three sub-determinants keep their products by 1 (det3_fma with the ones as values: the difference
// x - M01 y rounds the product first), while the outer sum folds the leading 1 and fuses the first
// product into it (checked bit for bit against the reference build's loader).
inline double det4_fma_ones(const double M[4][4]) {
    double M0[3][3], M1[3][3], M2[3][3], M3[3][3];
    for (int i = 0; i < 3; i++) {
        M0[i][0] = M[i + 1][1]; M0[i][1] = M[i + 1][2]; M0[i][2] = M[i + 1][3];
        M1[i][0] = M[i + 1][0]; M1[i][1] = M[i + 1][2]; M1[i][2] = M[i + 1][3];
        M2[i][0] = M[i + 1][0]; M2[i][1] = M[i + 1][1]; M2[i][2] = M[i + 1][3];
        M3[i][0] = M[i + 1][0]; M3[i][1] = M[i + 1][1]; M3[i][2] = M[i + 1][2];
    }
    const double d0 = det3_fma(M0), d1 = det3_fma(M1), d2 = det3_fma(M2), d3 = det3_fma(M3);
    const double s = std::fma(-M[0][1], d1, d0);
    return std::fma(-M[0][3], d3, std::fma(M[0][2], d2, s));
}

// Volume, vertex order and plane coefficients of the tet with nodes con[0..3]: volume = det4/6 with
// the rows (1, x, y, z) of the nodes; a negative volume swaps con[1] and con[3] and is negated.
// det16 receives (alpha, beta, gamma, delta) of vertex v at 4v..4v+3, each the 3x3 determinant of
// the other three nodes in the reference's vertex order, as packed in the app's c_det.
// FUSED selects the reference build's contractions (det3_fma / det4_fma).
template <bool FUSED = false>
inline void tet_geometry(const double* node_pos, int con[4], double* det16, double* vol) {
    auto d3 = [](const double A[3][3]) { return FUSED ? det3_fma(A) : det3(A); };
    auto d3_ones = [](const double A[3][3]) { return FUSED ? det3_fma_ones(A) : det3(A); };
    double M[4][4];
    for (int v = 0; v < 4; v++) {
        M[v][0] = 1.0;
        for (int d = 0; d < 3; d++) M[v][d + 1] = node_pos[3 * (int64_t)con[v] + d];
    }
    double V = (1.0 / 6.0) * (FUSED ? det4_fma_ones(M) : det4(M));
    if (V < 0) {
        std::swap(con[1], con[3]);
        V = -V;
    }
    *vol = V;
    for (int v = 0; v < 4; v++) {
        int v2, v3, v4;
        switch (v) {
            case 0: v2 = 1; v3 = 2; v4 = 3; break;
            case 1: v2 = 3; v3 = 2; v4 = 0; break;
            case 2: v2 = 3; v3 = 0; v4 = 1; break;
            default: v2 = 1; v3 = 0; v4 = 2; break;
        }
        const double* p2 = node_pos + 3 * (int64_t)con[v2];
        const double* p3 = node_pos + 3 * (int64_t)con[v3];
        const double* p4 = node_pos + 3 * (int64_t)con[v4];
        const double A[3][3] = {{p2[0], p2[1], p2[2]}, {p3[0], p3[1], p3[2]}, {p4[0], p4[1], p4[2]}};
        const double B[3][3] = {{1, p2[1], p2[2]}, {1, p3[1], p3[2]}, {1, p4[1], p4[2]}};
        const double G[3][3] = {{1, p2[0], p2[2]}, {1, p3[0], p3[2]}, {1, p4[0], p4[2]}};
        const double D[3][3] = {{1, p2[0], p2[1]}, {1, p3[0], p3[1]}, {1, p4[0], p4[1]}};
        det16[4 * v + 0] = d3(A);
        det16[4 * v + 1] = d3_ones(B);
        det16[4 * v + 2] = d3_ones(G);
        det16[4 * v + 3] = d3_ones(D);
    }
}

// Inside test of the mover's elemental kernel (apps/app_fempic/kernels.h:106-127), with the
// FMA contraction the reference's GCC -O3 build applies: lc = c2*fnma(d3,z,fma(d2,y,fnma(d1,x,d0))).
inline bool tet_contains(const double* det, double vol, double x, double y, double z) {
    const double c2 = (1.0 / 6.0) / vol;
    for (int i = 0; i < 4; i++) {
        const double* d = det + 4 * i;
        const double lc = c2 * std::fma(-d[3], z, std::fma(d[2], y, std::fma(-d[1], x, d[0])));
        if (!(lc > -1e-10 && (lc - 1.0) < 1e-10)) return false;
    }
    return true;
}

// c2c of a conforming tet mesh from its c2n (defined in box_mesh.cpp).
void face_neighbours(int64_t nn, int64_t nc, const int* c2n, int* c2c);

}  // namespace oppx_host
