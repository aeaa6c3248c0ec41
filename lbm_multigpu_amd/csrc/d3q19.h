// d3q19.h -- D3Q19 lattice tables shared by host and device code.
//
// Velocity order, weights, opposites and the moment basis follow the
// reference: velocities MRTtest/DynInlet3.cu:603-605 and :1625-1671,
// bounce-back permutation :1688-1705, weights CPU_3D/CPU_3D.cpp:135-153,
// moment rows (d'Humieres 2002 order) MRTtest/DynInlet3.cu:255-276.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

constexpr int Q = 19;

constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 0, 1, 0, -1, 0, 0, 1, 0, -1, 0};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6, 14, 17, 18, 15, 16, 9, 12, 13, 10, 11};

// weights as rationals w = WNUM/36
constexpr int WNUM[Q] = {12, 2, 2, 2, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1};

// The 9 opposite pairs (PP[j], OPP[PP[j]]) with PP[j] the first of each pair.
constexpr int NP = 9;
constexpr int PP[NP] = {1, 2, 5, 6, 9, 10, 11, 12, 13};

// Moment matrix entry M[k][i] from the defining polynomials of the basis
// (rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz,
// mx, my, mz).
__host__ __device__ constexpr int mrow(int k, int i)
{
    const int cx = CX[i], cy = CY[i], cz = CZ[i];
    const int c2 = cx * cx + cy * cy + cz * cz;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return cx;
    case 4: return (5 * c2 - 9) * cx;
    case 5: return cy;
    case 6: return (5 * c2 - 9) * cy;
    case 7: return cz;
    case 8: return (5 * c2 - 9) * cz;
    case 9: return 3 * cx * cx - c2;
    case 10: return (3 * c2 - 5) * (3 * cx * cx - c2);
    case 11: return cy * cy - cz * cz;
    case 12: return (3 * c2 - 5) * (cy * cy - cz * cz);
    case 13: return cx * cy;
    case 14: return cy * cz;
    case 15: return cx * cz;
    case 16: return cx * (cy * cy - cz * cz);
    case 17: return cy * (cz * cz - cx * cx);
    case 18: return cz * (cx * cx - cy * cy);
    }
    return 0;
}

// |row_k|^2: the rows are orthogonal, so M^-1 = M^T diag(1/NORM).
constexpr int NORM[Q] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8};

// Moments even under c -> -c (built from pair sums) and odd ones (pair differences).
constexpr bool even_moment(int k)
{
    return !(k == 3 || k == 4 || k == 5 || k == 6 || k == 7 || k == 8 || k == 16 || k == 17 ||
             k == 18);
}

// Node codes (low 7 bits of the device mask byte); bit 7 marks a fluid node
// with at least one link that bounces (halfway mode).
constexpr unsigned char NODE_FLUID = 0;
constexpr unsigned char NODE_SOLID = 1;
constexpr unsigned char NODE_SOLID_FORCE = 2;
constexpr unsigned char NODE_INLET_WEST = 3;
constexpr unsigned char NODE_OUTLET_EAST = 4;
constexpr unsigned char NODE_LID_WEST_REG = 5;
constexpr unsigned char NODE_SYM_XMIN = 6;
constexpr unsigned char NODE_SYM_XMAX = 7;
// y-streamwise inflow family and ghost copies (codes at and above
// NODE_OUTLET_NORTH take one branch off the common path)
constexpr unsigned char NODE_OUTLET_NORTH = 8;   // im 100
constexpr unsigned char NODE_INLET_SOUTH = 9;    // im 200
constexpr unsigned char NODE_RECYCLE = 10;       // im 80
constexpr unsigned char NODE_VEL_FROM_BELOW = 11; // im 70
constexpr unsigned char NODE_GHOST_W = 12;       // im 50 .. 54
constexpr unsigned char NODE_GHOST_E = 13;
constexpr unsigned char NODE_GHOST_S = 14;
constexpr unsigned char NODE_GHOST_N = 15;
constexpr unsigned char NODE_GHOST_DYN = 16;
constexpr unsigned char NODE_MAX = 16;
// Flags the host adds to a node's code in the device mask: NODE_NEAR, the
// node has a link that bounces (out through a wall face, or from a solid
// under link bounce-back); NODE_NEAR_SOLID, one of those links comes from a
// solid node, so the kernel must read the neighbours' codes (without it the
// bounced links follow from the domain faces alone).
constexpr unsigned char NODE_NEAR = 0x80;
constexpr unsigned char NODE_NEAR_SOLID = 0x40;
constexpr unsigned char NODE_CODE_MASK = 0x3f;

__host__ __device__ inline bool is_solid(unsigned char code)
{
    code &= NODE_CODE_MASK;
    return code == NODE_SOLID || code == NODE_SOLID_FORCE;
}

} // namespace lbm
