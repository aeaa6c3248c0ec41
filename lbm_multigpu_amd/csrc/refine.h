// refine.h -- 2:1 local refinement: the coarse-fine interface kernels.
//
// Geometry: fine node (i,j,k) sits at coarse coordinates
// (a_x + 0.25 + 0.5 i, a_y + 0.25 + 0.5 j, a_z + 0.25 + 0.5 k) (the
// reference's LRX0 = 128.25 layout, MRTtest/DynInlet3.cu:41-47).  The coarse
// footprint is nodes a .. a + n_f/2 per axis; its post-collision physical
// moments (PhysicalMoments, :396-407) are kept for the step's start (Mold)
// and end (Mnew).  Per coarse step:
//   coarse stream-collide -> extract Mnew -> [CF(Mold) -> fine step] ->
//   [CF(Mold/2 + Mnew/2) if temporal] -> fine step -> FC -> Mold = Mnew
// CF (InterpCF, :2202-2382) fills a frame two fine nodes thick on all six
// faces: trilinear interpolation of the 8 surrounding coarse moments, the
// ORDER-2 momentum correction (:2312-2357), and the LES-aware
// non-equilibrium rescale (InvertPhysicalMoments_LES_cf, :540-595).  FC
// (InterpFC, :2383-2560) replaces the coarse nodes a+2 .. a+n_f/2-2 by the
// mean of their 8 fine neighbours, the same correction with unit weights, and
// InvertPhysicalMoments_LES_fc (:483-538).
#pragma once

#include "collide.h"
#include "kernels.h"
#include "storage.h"

namespace lbm {

// Higher-moment equilibrium of the populations CF/FC rebuild.
// InvertPhysicalMoments (MRTtest/DynInlet3.cu:437-481) takes them from
// mrt_meq (eps = +475/63 j^2, 3pixx = piww = 0), which is not what the
// LES-family collide relaxes to (bgk_meq: eps = 3 rho - 11/2 j^2,
// 3pixx = -meq9/6, piww = -meq11/2): that mismatch writes an eps jump of
// ~3 rho into the frame, which streaming turns into density errors.
enum EqKind : int { EQ_BGK_MEQ = 0, EQ_BGK_PROJ = 1, EQ_MRT_MEQ = 2 };

struct RefineGeom {
    int a[3];     // footprint origin (coarse)
    int nf[3];    // fine nodes per axis
    int nc[3];    // footprint nodes per axis (nf/2 + 1)
    float omega_c, omega_f, cs;
    int eq_kind;  // EqKind
};

// PhysicalMoments: rho, jx, jy, jz, 3pxx row, pww row, pxy, pyz, pxz.
template <class R>
__device__ __forceinline__ void phys_moments(const R f[Q], R mom[9])
{
    R s[NP], d[NP];
    pair_split(f, s, d);
    mom[0] = moment(0, f[0], s, d);
    mom[1] = moment(3, f[0], s, d);
    mom[2] = moment(5, f[0], s, d);
    mom[3] = moment(7, f[0], s, d);
    mom[4] = moment(9, f[0], s, d);
    mom[5] = moment(11, f[0], s, d);
    mom[6] = moment(13, f[0], s, d);
    mom[7] = moment(14, f[0], s, d);
    mom[8] = moment(15, f[0], s, d);
}

// StrainRate (:367-393) as coded: omega = 1, m1 = 0.
template <class R>
__device__ __forceinline__ void strain_rate(const R m[9], R S[6])
{
    const R u = m[1], v = m[2], w = m[3];
    const R m9 = m[4] - (R(2.) * u * u - (v * v + w * w));
    const R m11 = m[5] - (v * v - w * w);
    S[0] = -R(0.026315789) * (R(19.) * m9);
    S[1] = -R(0.013157895) * (-R(19.) * (m9 - R(3.) * m11));
    S[2] = -R(0.013157895) * (-R(19.) * (m9 + R(3.) * m11));
    S[3] = -R(1.5) * (m[6] - u * v);
    S[4] = -R(1.5) * (m[7] - v * w);
    S[5] = -R(1.5) * (m[8] - u * w);
}

template <class R>
__device__ __forceinline__ R smag_of(const R S[6])
{
    return sqrt(R(2.) * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + R(2.) * S[3] * S[3] + R(2.) * S[4] * S[4] +
                        R(2.) * S[5] * S[5]));
}

// Populations from moments: conserved and non-equilibrium-scaled second
// moments; every other moment at the equilibrium selected by eq_kind.
template <class R>
__device__ __forceinline__ void invert_phys(R f[Q], const R mom[9], R SF, int eq_kind)
{
    R m[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) m[k] = R(0.);
    const R u = mom[1], v = mom[2], w = mom[3];
    const R usq = u * u + v * v + w * w;
    m[0] = mom[0];
    m[1] = -R(11.) * mom[0] + R(19.) * usq;
    if (eq_kind == EQ_MRT_MEQ) {
        m[2] = R(7.53968254) * usq;
    } else {
        const R m9eq = R(2.) * u * u - (v * v + w * w), m11eq = v * v - w * w;
        m[2] = R(3.) * mom[0] - R(5.5) * usq;
        m[10] = eq_kind == EQ_BGK_MEQ ? -R(0.5) * m9eq * (R(1.0) / R(3.0)) : -R(0.5) * m9eq;
        m[12] = -R(0.5) * m11eq;
    }
    m[3] = u; m[4] = -R(0.666666667) * u;
    m[5] = v; m[6] = -R(0.666666667) * v;
    m[7] = w; m[8] = -R(0.666666667) * w;
    m[9] = mom[4] * SF + (R(1.) - SF) * (R(2.) * u * u - (v * v + w * w));
    m[11] = mom[5] * SF + (R(1.) - SF) * (v * v - w * w);
    m[13] = mom[6] * SF + (R(1.) - SF) * u * v;
    m[14] = mom[7] * SF + (R(1.) - SF) * v * w;
    m[15] = mom[8] * SF + (R(1.) - SF) * u * w;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        R acc = R(0.);
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (mrow(k, i) != 0) acc += (R(mrow(k, i)) / R(NORM[k])) * m[k];
        f[i] = acc;
    }
}

// ORDER-2 correction (:2312-2357).  Corners c = 0..7 are (x,y,z) bits of c;
// fc_form = 1 gives the InterpFC weights (xpr = ypr = zpr = 0).
template <class R>
__device__ __forceinline__ void order2(const R (*Sc)[6], const R (*uc)[3], R mom[9], R xf, R yf, R zf,
                                       bool fc_form)
{
#define SXX(k) Sc[(k)-1][0]
#define SYY(k) Sc[(k)-1][1]
#define SZZ(k) Sc[(k)-1][2]
#define SXY(k) Sc[(k)-1][3]
#define SYZ(k) Sc[(k)-1][4]
#define SXZ(k) Sc[(k)-1][5]
#define U(k) uc[(k)-1][0]
#define V(k) uc[(k)-1][1]
#define W(k) uc[(k)-1][2]
    const R cx = -((SXX(8) - SXX(7) + SXX(6) - SXX(5) + SXX(4) - SXX(3) + SXX(2) - SXX(1))) * R(0.03125);
    const R cy = -((SXY(8) + SXY(7) - SXY(6) - SXY(5) + SXY(4) + SXY(3) - SXY(2) - SXY(1)) - V(8) + V(7) +
                       V(6) - V(5) - V(4) + V(3) + V(2) - V(1)) * R(0.0625);
    const R cz = -((SXZ(8) + SXZ(7) + SXZ(6) + SXZ(5) - SXZ(4) - SXZ(3) - SXZ(2) - SXZ(1)) - W(8) + W(7) -
                       W(6) + W(5) + W(4) - W(3) + W(2) - W(1)) * R(0.0625);
    const R dx = -((SXY(8) - SXY(7) + SXY(6) - SXY(5) + SXY(4) - SXY(3) + SXY(2) - SXY(1)) - U(8) + U(7) +
                       U(6) - U(5) - U(4) + U(3) + U(2) - U(1)) * R(0.0625);
    const R dy = -((SYY(8) + SYY(7) - SYY(6) - SYY(5) + SYY(4) + SYY(3) - SYY(2) - SYY(1))) * R(0.03125);
    const R dz = -((SYZ(8) + SYZ(7) + SYZ(6) + SYZ(5) - SYZ(4) - SYZ(3) - SYZ(2) - SYZ(1)) - W(8) - W(7) +
                       W(6) + W(5) + W(4) + W(3) - W(2) - W(1)) * R(0.0625);
    const R ex = -((SXZ(8) - SXZ(7) + SXZ(6) - SXZ(5) + SXZ(4) - SXZ(3) + SXZ(2) - SXZ(1)) - U(8) + U(7) -
                       U(6) + U(5) + U(4) - U(3) + U(2) - U(1)) * R(0.0625);
    const R ey = -((SYZ(8) + SYZ(7) - SYZ(6) - SYZ(5) + SYZ(4) + SYZ(3) - SYZ(2) - SYZ(1)) - V(8) - V(7) +
                       V(6) + V(5) + V(4) + V(3) - V(2) - V(1)) * R(0.0625);
    const R ez = -((SZZ(8) + SZZ(7) + SZZ(6) + SZZ(5) - SZZ(4) - SZZ(3) - SZZ(2) - SZZ(1))) * R(0.03125);
#undef SXX
#undef SYY
#undef SZZ
#undef SXY
#undef SYZ
#undef SXZ
#undef U
#undef V
#undef W
    R wx = R(1.), wy = R(1.), wz = R(1.);
    if (!fc_form) {
        wx = R(1.) - (R(4.) * xf * xf - R(4.) * xf + R(1.));
        wy = R(1.) - (R(4.) * yf * yf - R(4.) * yf + R(1.));
        wz = R(1.) - (R(4.) * zf * zf - R(4.) * zf + R(1.));
    }
    mom[1] += cx * wx + cy * wy + cz * wz;
    mom[2] += dx * wx + dy * wy + dz * wz;
    mom[3] += ex * wx + ey * wy + ez * wz;
}

// Footprint node data for one CF pass: 9 moments and their 6 strain rates,
// 16 values per node (AoS, so a CF thread fetches a corner with 16-byte
// loads: four for FP32 interface data, eight for FP64), in the interface's
// arithmetic type R (double for FP64 storage, float otherwise).
template <class R> struct CfVec;
template <> struct CfVec<float> { typedef float V __attribute__((ext_vector_type(4))); static constexpr int N = 4; };
template <> struct CfVec<double> { typedef double V __attribute__((ext_vector_type(2))); static constexpr int N = 2; };

template <class R>
__device__ __forceinline__ void cf_pack(const R m[9], R* P)
{
    R Sv[6];
    strain_rate(m, Sv);
    const R v[16] = {m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], Sv[0], Sv[1], Sv[2], Sv[3], Sv[4], Sv[5], R(0)};
    typedef typename CfVec<R>::V V;
    constexpr int N = CfVec<R>::N;
    V* pv = reinterpret_cast<V*>(P);
#pragma unroll
    for (int k = 0; k < 16 / N; ++k) {
        V x;
#pragma unroll
        for (int e = 0; e < N; ++e) x[e] = v[k * N + e];
        pv[k] = x;
    }
}

// The 16 values of one footprint node.
template <class R>
__device__ __forceinline__ void cf_load(const R* P, R out[16])
{
    typedef typename CfVec<R>::V V;
    constexpr int N = CfVec<R>::N;
    const V* pv = reinterpret_cast<const V*>(P);
#pragma unroll
    for (int k = 0; k < 16 / N; ++k) {
        const V x = pv[k];
#pragma unroll
        for (int e = 0; e < N; ++e) out[k * N + e] = x[e];
    }
}

// Footprint moments of the coarse lattice at the step's end (one thread per
// footprint node), and the CF data of the two fine sub-steps that follow:
// Pnew packs these moments (the first CF pass of the next coarse step, whose
// start they are), Phalf the midpoint (Mold + Mnew)/2 (the second CF pass of
// this step, temporal interpolation; skipped when Phalf is null).
template <class S>
__global__ void __launch_bounds__(256) k_extract(const S* fc, long long qs, int nxp, int ny, RefineGeom g,
                                                 const typename Arith<S>::R* Mold, typename Arith<S>::R* M,
                                                 typename Arith<S>::R* Pnew, typename Arith<S>::R* Phalf,
                                                 PopShift sh)
{
    using R = typename Arith<S>::R;
    const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
    const long long NC = (long long)ncx * ncy * ncz;
    const long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (o >= NC) return;
    const int i = int(o % ncx), j = int((o / ncx) % ncy), k = int(o / ((long long)ncx * ncy));
    const int cell = (g.a[0] + i) + nxp * ((g.a[1] + j) + ny * (g.a[2] + k));
    R f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = R(Store<S>::to_real(fc[q * qs + cell], q, sh));
    R mom[9];
    phys_moments(f, mom);
#pragma unroll
    for (int m = 0; m < 9; ++m) M[m * NC + o] = mom[m];
    cf_pack(mom, Pnew + 16 * o);
    if (Phalf) {
        R mid[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) mid[m] = (R(1) - R(0.5)) * Mold[m * NC + o] + R(0.5) * mom[m];
        cf_pack(mid, Phalf + 16 * o);
    }
}

// The same from inside the coarse stream-collide (the reference extracts in
// update_inn, MRTtest/DynInlet3.cu:1801-1815): a footprint node's populations
// as the launch leaves them in memory -- the stored post-collision values, or
// for a solid skipped under link bounce-back the values it keeps -- give the
// moments k_extract would read back, without the second pass over them.
template <class S, class R>
__device__ __forceinline__ void footprint_extract(const StepParams& p, int x, int y, int z, int cell, int act, R f[Q])
{
    const int i = x - p.fp_a[0], j = y - p.fp_a[1], k = z - p.fp_a[2];
    if ((unsigned)i >= (unsigned)p.fp_n[0] || (unsigned)j >= (unsigned)p.fp_n[1] || (unsigned)k >= (unsigned)p.fp_n[2])
        return;
    // k_interp_cf reads only the box's two outer layers on each face (the
    // corners of fine frame nodes 0, 1, nf - 2, nf - 1 along an axis are
    // coarse 0, 1, nc - 2, nc - 1); the inner nodes, which the restriction
    // overwrites, write nothing (on the 128x96x96 patch 79 % of the box)
    if (i >= 2 && i < p.fp_n[0] - 2 && j >= 2 && j < p.fp_n[1] - 2 && k >= 2 && k < p.fp_n[2] - 2) return;
    const long long NC = (long long)p.fp_n[0] * p.fp_n[1] * p.fp_n[2];
    const long long o = i + (long long)p.fp_n[0] * (j + (long long)p.fp_n[1] * k);
    R g[Q];
    if (act == ACT_SKIP) {
        const S* fo = static_cast<const S*>(p.fout);
#pragma unroll
        for (int q = 0; q < Q; ++q) g[q] = R(Store<S>::to_real(fo[q * p.qs + cell], q, p.sh));
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) g[q] = R(Store<S>::stored(f[q], q, p.sh));
    }
    R mom[9];
    phys_moments(g, mom);
    R* M = static_cast<R*>(p.fp_m);
#pragma unroll
    for (int m = 0; m < 9; ++m) M[m * NC + o] = mom[m];
    cf_pack(mom, static_cast<R*>(p.fp_pnew) + 16 * o);
    if (p.fp_phalf) {
        const R* Mold = static_cast<const R*>(p.fp_mold);
        R mid[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) mid[m] = (R(1) - R(0.5)) * Mold[m * NC + o] + R(0.5) * mom[m];
        cf_pack(mid, static_cast<R*>(p.fp_phalf) + 16 * o);
    }
}

// Frame node number o of the CF frame (two fine nodes thick on all six faces)
// -> (i, j, k).  Enumerated as the z slabs (k < 2 or k >= nf_z - 2), then the
// y slabs of the remaining planes, then the x slabs of the remaining rows; x
// runs fastest in the first two, so consecutive threads write consecutive x.
__device__ __forceinline__ bool frame_node(long long o, const RefineGeom& g, int& i, int& j, int& k)
{
    const int nx = g.nf[0], ny = g.nf[1], nz = g.nf[2];
    const long long nZ = 4LL * nx * ny, nY = 4LL * nx * (nz - 4), nX = 4LL * (ny - 4) * (nz - 4);
    if (o < nZ) {
        i = int(o % nx);
        const long long r = o / nx;
        j = int(r % ny);
        const int s = int(r / ny);
        k = s < 2 ? s : nz - 4 + s;
        return true;
    }
    o -= nZ;
    if (o < nY) {
        i = int(o % nx);
        const long long r = o / nx;
        k = 2 + int(r % (nz - 4));
        const int s = int(r / (nz - 4));
        j = s < 2 ? s : ny - 4 + s;
        return true;
    }
    o -= nY;
    if (o < nX) {
        const int s = int(o % 4);
        const long long r = o / 4;
        i = s < 2 ? s : nx - 4 + s;
        j = 2 + int(r % (ny - 4));
        k = 2 + int(r / (ny - 4));
        return true;
    }
    return false;
}

__host__ __device__ inline long long frame_nodes(const int nf[3])
{
    return 4LL * nf[0] * nf[1] + 4LL * nf[0] * (nf[2] - 4) + 4LL * (nf[1] - 4) * (nf[2] - 4);
}

// Coarse -> fine frame, one thread per frame node: the 8 surrounding
// footprint nodes' blended moments and strain rates (k_extract, L2-resident
// and shared by neighbouring threads), then trilinear interpolation, ORDER-2
// correction and the LES-aware rescale.
template <class S>
__device__ __forceinline__ void interp_cf_node(long long o, S* ff, long long fqs, int fnxp, const unsigned char* fmask,
                                               const typename Arith<S>::R* P, const RefineGeom& g, const PopShift& sh)
{
    using R = typename Arith<S>::R;
    int i, j, k;
    if (!frame_node(o, g, i, j, k)) return;
    const int nfy = g.nf[1];
    const int fcell = i + fnxp * (j + nfy * k);
    if (fmask && is_solid(fmask[fcell])) return;
    const int ncx = g.nc[0], ncy = g.nc[1];
    const int lx = i >> 1, ly = j >> 1, lz = k >> 1;
    const R xf = (i & 1) ? R(0.75) : R(0.25), yf = (j & 1) ? R(0.75) : R(0.25), zf = (k & 1) ? R(0.75) : R(0.25);
    R Sc[8][6], uc[8][3], mom[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) mom[m] = R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const long long o = (lx + dx) + (long long)ncx * ((ly + dy) + (long long)ncy * (lz + dz));
        R v[16]; // 9 moments, then the 6 strain rates
        cf_load(P + 16 * o, v);
        const R w = (dx ? xf : R(1) - xf) * (dy ? yf : R(1) - yf) * (dz ? zf : R(1) - zf);
#pragma unroll
        for (int q = 0; q < 9; ++q) mom[q] += v[q] * w;
#pragma unroll
        for (int q = 0; q < 6; ++q) Sc[c][q] = v[9 + q];
        uc[c][0] = v[1]; uc[c][1] = v[2]; uc[c][2] = v[3];
    }
    order2(Sc, uc, mom, xf, yf, zf, false);
    // InvertPhysicalMoments_LES_cf (:540-595)
    R Sv[6];
    strain_rate(mom, Sv);
    const R Smag_c = smag_of(Sv);
    R tau_c = R(1) / R(g.omega_c);
    const R tau_f = R(2) * tau_c - R(0.5) + R(1.5) * Smag_c * R(g.cs);
    const R omega_f = R(1) / tau_f;
    tau_c = tau_c + Smag_c * R(g.cs);
    const R omega_c = R(1) / tau_c;
    const R SF = (R(0.5) * omega_c * (R(1) - omega_f)) / ((R(1) - omega_c) * omega_f);
    R f[Q];
    invert_phys(f, mom, SF, g.eq_kind);
#pragma unroll
    for (int q = 0; q < Q; ++q) ff[q * fqs + fcell] = Store<S>::from_real(f[q], q, sh);
}

template <class S>
__global__ void __launch_bounds__(256) k_interp_cf(S* ff, long long fqs, int fnxp, const unsigned char* fmask,
                                                   const typename Arith<S>::R* P, RefineGeom g, PopShift sh)
{
    interp_cf_node<S>(blockIdx.x * 256LL + threadIdx.x, ff, fqs, fnxp, fmask, P, g, sh);
}

// Fine -> coarse.  One workgroup per tile of FC_TX x FC_TY x FC_TZ coarse
// nodes: its 256 threads load the (2 FC_TX) x (2 FC_TY) x (2 FC_TZ) fine nodes
// below them as contiguous 64-node rows (one wavefront per row and plane),
// reduce them to the 9 physical moments in LDS, and one thread per coarse
// node then finishes it.
#ifndef LBM_FC_TY
#define LBM_FC_TY 1
#endif
#ifndef LBM_FC_TZ
#define LBM_FC_TZ 1
#endif
constexpr int FC_TX = 32, FC_TY = LBM_FC_TY, FC_TZ = LBM_FC_TZ;
constexpr int FC_ROWS = 4 * FC_TY * FC_TZ; // fine rows of a tile (a multiple of the 4 wavefronts)

template <class S>
__device__ __forceinline__ void interp_fc_block(int blk, const S* ff, long long fqs, int fnxp, S* fc, long long cqs,
                                                int cnxp, int cny, const unsigned char* cmask, const RefineGeom& g,
                                                int nbx, int nby, const PopShift& sh)
{
    using R = typename Arith<S>::R;
    __shared__ R fm[2 * FC_TZ][2 * FC_TY][2 * FC_TX][9]; // fine moments [fz][fy][fx][m]
    const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
    // coarse nodes a + 2 + (FC_T* b) + l; fine nodes 2(c - a) - 1 .. 2(c - a)
    const int fx0 = 2 * (2 + FC_TX * bx) - 1, fy0 = 2 * (2 + FC_TY * by) - 1, fz0 = 2 * (2 + FC_TZ * bz) - 1;
    const int lxf = threadIdx.x & 63, row0 = threadIdx.x >> 6; // 4 wavefronts
#pragma unroll
    for (int rr = 0; rr < FC_ROWS / 4; ++rr) {
        const int row = row0 + 4 * rr;          // (fy, fz) of the tile
        const int ly = row % (2 * FC_TY), lz = row / (2 * FC_TY);
        const int i = fx0 + lxf, j = fy0 + ly, k = fz0 + lz;
        R mom[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) mom[m] = R(0);
        // only fine nodes 3 .. nf - 4 feed a restricted coarse node; the
        // frame beyond them (which k_interp_fc_cf's CF workgroups may be
        // writing in the same launch) is not read
        if (i <= g.nf[0] - 4 && j <= g.nf[1] - 4 && k <= g.nf[2] - 4) {
            const int fcell = i + fnxp * (j + g.nf[1] * k);
            R f[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = R(Store<S>::to_real(ff[q * fqs + fcell], q, sh));
            phys_moments(f, mom);
        }
#pragma unroll
        for (int m = 0; m < 9; ++m) fm[lz][ly][lxf][m] = mom[m];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= FC_TX * FC_TY * FC_TZ) return;
    const int lx = t % FC_TX, ly = (t / FC_TX) % FC_TY, lz = t / (FC_TX * FC_TY);
    const int cx = g.a[0] + 2 + FC_TX * bx + lx, cy = g.a[1] + 2 + FC_TY * by + ly, cz = g.a[2] + 2 + FC_TZ * bz + lz;
    if (cx > g.a[0] + g.nf[0] / 2 - 2 || cy > g.a[1] + g.nf[1] / 2 - 2 || cz > g.a[2] + g.nf[2] / 2 - 2) return;
    const int ccell = cx + cnxp * (cy + cny * cz);
    if (cmask && is_solid(cmask[ccell])) return;
    R Sc[8][6], uc[8][3], avg[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) avg[m] = R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        R m[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = fm[2 * lz + dz][2 * ly + dy][2 * lx + dx][q];
#pragma unroll
        for (int q = 0; q < 9; ++q) avg[q] += m[q];
        strain_rate(m, Sc[c]);
        uc[c][0] = m[1]; uc[c][1] = m[2]; uc[c][2] = m[3];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) avg[q] *= R(0.125);
    order2(Sc, uc, avg, R(0), R(0), R(0), true);
    // InvertPhysicalMoments_LES_fc (:483-538)
    R Sv[6];
    strain_rate(avg, Sv);
    const R Smag_f = smag_of(Sv);
    R tau_f = R(1) / R(g.omega_f);
    R tau_c = tau_f + R(0.5) + R(12) * Smag_f * R(g.cs);
    tau_c *= R(0.5);
    const R omega_c = R(1) / tau_c;
    tau_f = tau_f + Smag_f * R(g.cs);
    const R omega_f = R(1) / tau_f;
    const R SF = (R(1) - omega_c) * omega_f / (R(0.5) * omega_c * (R(1) - omega_f));
    R f[Q];
    invert_phys(f, avg, SF, g.eq_kind);
#pragma unroll
    for (int q = 0; q < Q; ++q) fc[q * cqs + ccell] = Store<S>::from_real(f[q], q, sh);
}

template <class S>
__global__ void __launch_bounds__(256) k_interp_fc(const S* ff, long long fqs, int fnxp, S* fc, long long cqs,
                                                   int cnxp, int cny, const unsigned char* cmask, RefineGeom g,
                                                   int nbx, int nby, PopShift sh)
{
    interp_fc_block<S>(blockIdx.x, ff, fqs, fnxp, fc, cqs, cnxp, cny, cmask, g, nbx, nby, sh);
}

// The restriction and the next step's first CF pass in one launch: the
// first nb_cf workgroups fill the frame of the fine lattice ff (from the
// coarse data P), the rest restrict ff's inner nodes onto the coarse
// lattice.  CF writes only frame nodes (fine 0, 1, nf - 2, nf - 1 along an
// axis) and reads only P; the restriction reads fine nodes 3 .. nf - 4 and
// writes coarse ones, so the two sets of workgroups touch nothing in common
// and the latency-bound CF pass runs beside the restriction instead of as a
// launch of its own.
template <class S>
__global__ void __launch_bounds__(256) k_interp_fc_cf(const S* ff_in, S* ff, long long fqs, int fnxp,
                                                      const unsigned char* fmask, const typename Arith<S>::R* P,
                                                      S* fc, long long cqs, int cnxp, int cny,
                                                      const unsigned char* cmask, RefineGeom g, int nbx, int nby,
                                                      int nb_cf, PopShift sh)
{
    if ((int)blockIdx.x < nb_cf)
        interp_cf_node<S>(blockIdx.x * 256LL + threadIdx.x, ff, fqs, fnxp, fmask, P, g, sh);
    else
        interp_fc_block<S>((int)blockIdx.x - nb_cf, ff_in, fqs, fnxp, fc, cqs, cnxp, cny, cmask, g, nbx, nby, sh);
}

} // namespace lbm
