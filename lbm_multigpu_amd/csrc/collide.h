// collide.h -- device-side D3Q19 collision operators and node boundary
// conditions, written for 64-wide CDNA wavefronts: every loop over the 19
// directions or 19 moments is unrolled at compile time so populations stay in
// VGPRs, and the moment transforms use the pair sums/differences of opposite
// directions (even moments from sums, odd moments from differences), which
// halves both the forward and the inverse transform.
#pragma once

#include "d3q19.h"

namespace lbm {

enum Model : int { MODEL_BGK = 0, MODEL_MRT_CPU3D = 1, MODEL_MRT_LES = 2, MODEL_MRT_DHUM = 3 };

template <class R>
struct CollideParams {
    R omega;       // base rate 1/tau0
    R cs;          // reference CS (Cs^2); 0 = no LES
    R dpdy;        // AddForce pressure gradient
    R s1, s2, s4, s10, s16; // d'Humieres rates (e, eps, q, pi, m)
};

template <class R>
__device__ __forceinline__ R wgt(int i)
{
    return R(WNUM[i]) / R(36);
}

// BGK equilibrium w_i (rho + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), u the momentum
// (CPU_3D/CPU_3D.cpp:135-153).
template <class R>
__device__ __forceinline__ void feq_bgk(R rho, R u, R v, R w, R feq[Q])
{
    const R base = rho - R(1.5) * (u * u + v * v + w * w);
    feq[0] = wgt<R>(0) * base;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int i = PP[j];
        const R cu = R(CX[i]) * u + R(CY[i]) * v + R(CZ[i]) * w;
        const R a = wgt<R>(i) * (base + R(4.5) * cu * cu);
        const R b = wgt<R>(i) * R(3) * cu;
        feq[i] = a + b;
        feq[OPP[i]] = a - b;
    }
}

// Pair sums s[j] = f_i + f_opp and differences d[j] = f_i - f_opp.
template <class R>
__device__ __forceinline__ void pair_split(const R f[Q], R s[NP], R d[NP])
{
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        s[j] = f[PP[j]] + f[OPP[PP[j]]];
        d[j] = f[PP[j]] - f[OPP[PP[j]]];
    }
}

// m_k = sum_i M[k][i] f_i from the pair sums/differences.
template <class R>
__device__ __forceinline__ R moment(int k, R f0, const R s[NP], const R d[NP])
{
    R acc = R(mrow(k, 0)) * f0;
    if (even_moment(k)) {
#pragma unroll
        for (int j = 0; j < NP; ++j)
            if (mrow(k, PP[j]) != 0) acc += R(mrow(k, PP[j])) * s[j];
    } else {
#pragma unroll
        for (int j = 0; j < NP; ++j)
            if (mrow(k, PP[j]) != 0) acc += R(mrow(k, PP[j])) * d[j];
    }
    return acc;
}

// All 19 moments m_k = sum_i M[k][i] f_i, factored along the lattice's
// structure: the three face pairs and the three planes of edge pairs share
// partial sums across moments (about 45 operations where the row-by-row sums
// take about 115).  Pair order (PP): x face, y face, xy edges (+,+) (-,+),
// z face, xz edge (+,+), yz edge (+,+), xz edge (-,+), yz edge (-,+).
template <class R>
__host__ __device__ __forceinline__ void moments_d3q19(R f0, const R s[NP], const R d[NP], R m[Q])
{
    // even moments from the pair sums
    const R exy = s[2] + s[3], exz = s[5] + s[7], eyz = s[6] + s[8];
    const R faces = s[0] + s[1] + s[4], edges = exy + exz + eyz;
    m[0] = f0 + faces + edges;
    m[1] = R(-30) * f0 + R(-11) * faces + R(8) * edges;
    m[2] = R(12) * f0 + R(-4) * faces + edges;
    const R t = exy + exz - R(2) * eyz, u = R(2) * s[0] - (s[1] + s[4]);
    m[9] = u + t;
    m[10] = t - R(2) * u;
    const R v = s[1] - s[4], w = exy - exz;
    m[11] = v + w;
    m[12] = w - R(2) * v;
    m[13] = s[2] - s[3];
    m[14] = s[6] - s[8];
    m[15] = s[5] - s[7];
    // odd moments from the pair differences
    const R xxy = d[2] - d[3], xxz = d[5] - d[7];
    const R yxy = d[2] + d[3], yyz = d[6] - d[8];
    const R zxz = d[5] + d[7], zyz = d[6] + d[8];
    const R ex = xxy + xxz, ey = yxy + yyz, ez = zxz + zyz;
    m[3] = d[0] + ex;
    m[4] = R(-4) * d[0] + ex;
    m[5] = d[1] + ey;
    m[6] = R(-4) * d[1] + ey;
    m[7] = d[4] + ez;
    m[8] = R(-4) * d[4] + ez;
    m[16] = xxy - xxz;
    m[17] = yyz - yxy;
    m[18] = zxz - zyz;
}

// f_i -= sum_k M[k][i] dm_k (the inverse transform of the scaled
// non-equilibrium moments; dm of the conserved moments is zero), factored
// the same way: per pair j, f_i -= E + O and f_opp -= E - O with E the even
// and O the odd part.
template <class R>
__host__ __device__ __forceinline__ void relax_d3q19(const R dm[Q], R f[Q])
{
    f[0] -= R(-30) * dm[1] + R(12) * dm[2];
    // faces: c^2 = 1
    const R a = R(-11) * dm[1] + R(-4) * dm[2];
    const R b = R(2) * dm[10] - dm[9];
    const R c = dm[11] - R(2) * dm[12];
    const R ex = a - R(2) * b, eyz = a + b;
    const R ey = eyz + c, ez = eyz - c;
    const R ox = R(-4) * dm[4], oy = R(-4) * dm[6], oz = R(-4) * dm[8];
    // edges: c^2 = 2
    const R g = R(8) * dm[1] + dm[2];
    const R k = dm[9] + dm[10], l = dm[11] + dm[12];
    const R gk = g + k;
    const R hxy = gk + l, hxz = gk - l, hyz = g - R(2) * k;
    const R pxy = dm[4] + dm[16], rxy = dm[6] - dm[17];
    const R pxz = dm[4] - dm[16], rxz = dm[8] + dm[18];
    const R pyz = dm[6] + dm[17], ryz = dm[8] - dm[18];
    const R E[NP] = {ex, ey, hxy + dm[13], hxy - dm[13], ez, hxz + dm[15], hyz + dm[14], hxz - dm[15], hyz - dm[14]};
    const R O[NP] = {ox, oy, pxy + rxy, rxy - pxy, oz, pxz + rxz, pyz + ryz, rxz - pxz, ryz - pyz};
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        f[PP[j]] -= E[j] + O[j];
        f[OPP[PP[j]]] -= E[j] - O[j];
    }
}

// Second moments P_ab = sum c_a c_b f and conserved moments.
template <class R>
struct Macro {
    R rho, jx, jy, jz, pxx, pyy, pzz, pxy, pyz, pxz;
};

template <class R>
__device__ __forceinline__ Macro<R> macro_from_pairs(R f0, const R s[NP], const R d[NP])
{
    Macro<R> m;
    R r = f0;
    R jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pyz = 0, pxz = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int i = PP[j];
        r += s[j];
        if (CX[i]) jx += R(CX[i]) * d[j];
        if (CY[i]) jy += R(CY[i]) * d[j];
        if (CZ[i]) jz += R(CZ[i]) * d[j];
        if (CX[i]) pxx += s[j];
        if (CY[i]) pyy += s[j];
        if (CZ[i]) pzz += s[j];
        if (CX[i] * CY[i]) pxy += R(CX[i] * CY[i]) * s[j];
        if (CY[i] * CZ[i]) pyz += R(CY[i] * CZ[i]) * s[j];
        if (CX[i] * CZ[i]) pxz += R(CX[i] * CZ[i]) * s[j];
    }
    m.rho = r; m.jx = jx; m.jy = jy; m.jz = jz;
    m.pxx = pxx; m.pyy = pyy; m.pzz = pzz; m.pxy = pxy; m.pyz = pyz; m.pxz = pxz;
    return m;
}

// Smagorinsky effective rate (MRTtest/DynInlet3.cu:628-735):
// Pi_neq = P - rho/3 I - j j (the second moment of f - feq_BGK),
// Q = sqrt(Pi:Pi), tau = tau0/2 + sqrt(tau0^2 + 18 sqrt(2) CS Q)/2.
template <class R>
__device__ __forceinline__ R les_omega(const Macro<R>& m, R omega0, R cs)
{
    const R third = m.rho * R(1.0 / 3.0);
    const R nxx = m.pxx - third - m.jx * m.jx;
    const R nyy = m.pyy - third - m.jy * m.jy;
    const R nzz = m.pzz - third - m.jz * m.jz;
    const R nxy = m.pxy - m.jx * m.jy;
    const R nyz = m.pyz - m.jy * m.jz;
    const R nxz = m.pxz - m.jx * m.jz;
    const R q2 = nxx * nxx + nyy * nyy + nzz * nzz + R(2) * (nxy * nxy + nyz * nyz + nxz * nxz);
    const R q = sqrt(q2);
    const R tau0 = R(1) / omega0;
    const R tau = R(0.5) * tau0 + R(0.5) * sqrt(tau0 * tau0 + R(18.0 * 1.4142135623730951) * cs * q);
    return R(1) / tau;
}

// The same Smagorinsky rate from the moments the MRT collide already has:
// tr P = (m1 + 30 rho)/19, P_xx = (tr P + m9)/3, P_yy - P_zz = m11,
// P_xy = m13, P_yz = m14, P_xz = m15 (rows of the d'Humieres basis).  FAST
// (the 16-bit storage kernels, which are bound by the vector ALUs) uses the
// hardware reciprocal and square root instead of the IEEE sequences.
template <bool FAST, class R>
__device__ __forceinline__ R les_omega_from_moments(const R m[Q], R omega0, R cs)
{
    const R rho = m[0], jx = m[3], jy = m[5], jz = m[7];
    const R tr = (m[1] + R(30) * rho) * R(1.0 / 19.0);
    const R pxx = (tr + m[9]) * R(1.0 / 3.0);
    const R pyz_sum = tr - pxx;
    const R pyy = R(0.5) * (pyz_sum + m[11]);
    const R pzz = R(0.5) * (pyz_sum - m[11]);
    const R third = rho * R(1.0 / 3.0);
    const R nxx = pxx - third - jx * jx;
    const R nyy = pyy - third - jy * jy;
    const R nzz = pzz - third - jz * jz;
    const R nxy = m[13] - jx * jy;
    const R nyz = m[14] - jy * jz;
    const R nxz = m[15] - jx * jz;
    const R q2 = nxx * nxx + nyy * nyy + nzz * nzz + R(2) * (nxy * nxy + nyz * nyz + nxz * nxz);
    if constexpr (FAST && sizeof(R) == 4) {
        const float q = __builtin_amdgcn_sqrtf(q2);
        const float tau0 = __builtin_amdgcn_rcpf(omega0);
        const float tau = 0.5f * tau0 + 0.5f * __builtin_amdgcn_sqrtf(tau0 * tau0 + float(18.0 * 1.4142135623730951) * cs * q);
        return __builtin_amdgcn_rcpf(tau);
    } else {
        const R q = sqrt(q2);
        const R tau0 = R(1) / omega0;
        const R tau = R(0.5) * tau0 + R(0.5) * sqrt(tau0 * tau0 + R(18.0 * 1.4142135623730951) * cs * q);
        return R(1) / tau;
    }
}

// AddForce (MRTtest/DynInlet3.cu:229-253): f_i -= 3 w_i c_iy dpdy.
template <class R>
__device__ __forceinline__ void add_force(R f[Q], R dpdy)
{
#pragma unroll
    for (int i = 0; i < Q; ++i)
        if (CY[i]) f[i] -= R(3 * CY[i]) * wgt<R>(i) * dpdy;
}

// BGK collide (CPU_3D/CPU_3D.cpp:116-160; DynInlet3.cu:802-830 adds the LES
// rate and AddForce).  Evaluated as the same IEEE operation sequence as the
// CPU restatement (index-order sums, feq = w (((rho + 3cu) + 4.5cu cu) - 1.5 u.u),
// f - omega (f - feq), no FMA contraction), so both produce the same bits;
// BGK is far from the VALU limit, so the unshared arithmetic costs nothing.
template <class R>
__device__ __forceinline__ void collide_bgk(R f[Q], const CollideParams<R>& cp)
{
#pragma clang fp contract(off)
    R rho = 0, u = 0, v = 0, w = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        rho += f[i];
        u += R(CX[i]) * f[i];
        v += R(CY[i]) * f[i];
        w += R(CZ[i]) * f[i];
    }
    R omega = cp.omega;
    if (cp.cs > R(0)) {
        R s[NP], d[NP];
        pair_split(f, s, d);
        const Macro<R> m = macro_from_pairs(f[0], s, d);
        omega = les_omega(m, cp.omega, cp.cs);
    }
    const R usqr = u * u + v * v + w * w;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const R cu = R(CX[i]) * u + R(CY[i]) * v + R(CZ[i]) * w;
        const R feq = wgt<R>(i) * (rho + R(3) * cu + R(4.5) * cu * cu - R(1.5) * usqr);
        f[i] -= omega * (f[i] - feq);
    }
    if (cp.dpdy != R(0)) add_force(f, cp.dpdy);
}

// MRT collide, f <- f - M^-1 S (m - meq), for the three presets:
//  MODEL_MRT_CPU3D: CPU_3D/CPU_3D.cpp:229-279.  meq has eps = -475/63 j^2,
//    3pixx = piww = m16..18 = 0; S = 1 except s9 = s11 = s15 = omega
//    (m13, m14 at rate 1 as coded there).  No LES, no body force.
//  MODEL_MRT_LES: MRTtest/DynInlet3.cu:759-789 with bgk_meq (:347-365,
//    whose 3pixx entry is -meq9/6 as coded); s9,11,13,14,15 = omega_LES,
//    every other moment to equilibrium; then AddForce.
//  MODEL_MRT_DHUM: as MRT_LES with the d'Humieres rates of :774-783.
template <int MODEL, class R, bool FAST = false>
__device__ __forceinline__ void collide_mrt(R f[Q], const CollideParams<R>& cp)
{
    R s[NP], d[NP];
    pair_split(f, s, d);
    R m[Q];
    moments_d3q19(f[0], s, d, m);
    const R rho = m[0], u = m[3], v = m[5], w = m[7];
    const R usqr = u * u + v * v + w * w;

    R omega = cp.omega;
    if (MODEL != MODEL_MRT_CPU3D && cp.cs > R(0)) omega = les_omega_from_moments<FAST>(m, cp.omega, cp.cs);

    R meq[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) meq[k] = R(0);
    meq[1] = R(-11) * rho + R(19) * usqr;
    meq[4] = R(-2.0 / 3.0) * u;
    meq[6] = R(-2.0 / 3.0) * v;
    meq[8] = R(-2.0 / 3.0) * w;
    meq[9] = R(2) * u * u - (v * v + w * w);
    meq[11] = v * v - w * w;
    meq[13] = u * v;
    meq[14] = v * w;
    meq[15] = u * w;
    if (MODEL == MODEL_MRT_CPU3D) {
        meq[2] = R(-475.0 / 63.0) * usqr;
    } else {
        meq[2] = R(3) * rho - R(5.5) * usqr;
        meq[10] = R(-0.5) * meq[9] * R(1.0 / 3.0);
        meq[12] = R(-0.5) * meq[11];
    }

    // rates
    R rate[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) rate[k] = R(1);
    rate[9] = omega; rate[11] = omega; rate[15] = omega;
    if (MODEL != MODEL_MRT_CPU3D) { rate[13] = omega; rate[14] = omega; }
    if (MODEL == MODEL_MRT_DHUM) {
        rate[1] = cp.s1; rate[2] = cp.s2;
        rate[4] = cp.s4; rate[6] = cp.s4; rate[8] = cp.s4;
        rate[10] = cp.s10; rate[12] = cp.s10;
        rate[16] = cp.s16; rate[17] = cp.s16; rate[18] = cp.s16;
    }

    // scaled non-equilibrium parts dm_k = (s_k / |row_k|^2) (m_k - meq_k);
    // moments sharing a rate and norm share the product
    R dm[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const bool conserved = (k == 0 || k == 3 || k == 5 || k == 7);
        dm[k] = conserved ? R(0) : (rate[k] * (R(1) / R(NORM[k]))) * (m[k] - meq[k]);
    }

    // f_i -= sum_k M[k][i] dm_k
    relax_d3q19(dm, f);
    if (MODEL != MODEL_MRT_CPU3D && cp.dpdy != R(0)) add_force(f, cp.dpdy);
}

template <int MODEL, class R, bool FAST = false>
__device__ __forceinline__ void collide(R f[Q], const CollideParams<R>& cp)
{
    if (MODEL == MODEL_BGK) collide_bgk(f, cp);
    else collide_mrt<MODEL, R, FAST>(f, cp);
}

// Velocity-prescribing Extrap conditions applied to a neighbour's unstreamed
// populations: f = M^-1 m with j = (ux, uy, uz) and every other moment (rho
// included) from f; as a correction f_i += c_i . (u - j)/10.  West_Extrap
// (MRTtest/DynInlet3.cu:973-1017) prescribes (u_in, 0, 0), South_Extrap
// (:881-924) (0, v_in, 0), and the im 70 plane of bgk_channel1/Re3000_Cyl2/
// DynInlet3.cu:1220-1230 the momentum of the node below.
template <class R>
__device__ __forceinline__ void vel_extrap(R f[Q], R ux, R uy, R uz)
{
    R jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (CX[i]) jx += R(CX[i]) * f[i];
        if (CY[i]) jy += R(CY[i]) * f[i];
        if (CZ[i]) jz += R(CZ[i]) * f[i];
    }
    const R ax = (ux - jx) * R(0.1), ay = (uy - jy) * R(0.1), az = (uz - jz) * R(0.1);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] += R(CX[i]) * ax + R(CY[i]) * ay + R(CZ[i]) * az;
}

template <class R>
__device__ __forceinline__ void west_extrap(R f[Q], R u_in)
{
    vel_extrap(f, u_in, R(0), R(0));
}

// Momentum sums of a population set.
template <class R>
__device__ __forceinline__ void momentum(const R f[Q], R& jx, R& jy, R& jz)
{
    jx = 0; jy = 0; jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (CX[i]) jx += R(CX[i]) * f[i];
        if (CY[i]) jy += R(CY[i]) * f[i];
        if (CZ[i]) jz += R(CZ[i]) * f[i];
    }
}

// East_Extrap (MRTtest/DynInlet3.cu:926-971) and North_Extrap (:834-879):
// f = M^-1 m with m0 = rho_out, i.e. f_i += (rho_out - sum f)/19.
template <class R>
__device__ __forceinline__ void east_extrap(R f[Q], R rho_out)
{
    R rho = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) rho += f[i];
    const R a = (rho_out - rho) * R(1.0 / 19.0);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] += a;
}

// DirichletWest_Regularized (CPU_3D/CPU_3D.cpp:974-1054) as coded there:
// velocity (0, u_lid, 0); feq11/feq13 use (v+w)*(u+w) and (-v+w)*(u+w); the
// unknown f8 becomes feq6 + f6 - feq6; then f_i = feq_i + 4.5 w_i Q_i:Pi_neq.
template <class R>
__device__ __forceinline__ void lid_west_reg(R f[Q], R u_lid)
{
    const R u = 0, v = u_lid, w = 0;
    const R usqr = v * v;
    const R rho = u + (f[0] + f[2] + f[4] + f[9] + f[11] + f[13] + f[14] + f[16] + f[18] +
                       R(2) * (f[3] + f[6] + f[7] + f[12] + f[17]));
    R feq[Q];
    feq_bgk(rho, u, v, w, feq);
    feq[11] = wgt<R>(11) * (rho + R(3) * (v + w) + R(4.5) * (v + w) * (u + w) - R(1.5) * usqr);
    feq[13] = wgt<R>(13) * (rho + R(3) * (-v + w) + R(4.5) * (-v + w) * (u + w) - R(1.5) * usqr);
    f[1] = feq[1] + f[3] - feq[3];
    f[5] = feq[5] + f[7] - feq[7];
    f[8] = feq[6] + f[6] - feq[6];
    f[10] = feq[10] + f[17] - feq[17];
    f[15] = feq[15] + f[12] - feq[12];
    R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pyz = 0, pxz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const R dd = f[i] - feq[i];
        if (CX[i]) pxx += dd;
        if (CY[i]) pyy += dd;
        if (CZ[i]) pzz += dd;
        if (CX[i] * CY[i]) pxy += R(CX[i] * CY[i]) * dd;
        if (CY[i] * CZ[i]) pyz += R(CY[i] * CZ[i]) * dd;
        if (CX[i] * CZ[i]) pxz += R(CX[i] * CZ[i]) * dd;
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const R qxx = R(CX[i] * CX[i]) - R(1.0 / 3.0);
        const R qyy = R(CY[i] * CY[i]) - R(1.0 / 3.0);
        const R qzz = R(CZ[i] * CZ[i]) - R(1.0 / 3.0);
        const R qs = qxx * pxx + qyy * pyy + qzz * pzz +
                     R(2) * (R(CX[i] * CY[i]) * pxy + R(CY[i] * CZ[i]) * pyz + R(CX[i] * CZ[i]) * pxz);
        f[i] = feq[i] + R(4.5) * wgt<R>(i) * qs;
    }
}

// x = 0 symmetry plane, xsymmetry_bot (MRTtest/DynInlet3.cu:1020-1075):
// the populations entering through the plane copy their mirror images, with
// the reference's own treatment of the y/z edges and corners.  The copies
// are sequential (later ones read earlier results), as coded.  Each case is
// applied under its own predicate (one at most holds), so every index stays
// a compile-time constant and the populations stay in registers.
#define LBM_SYM_COPY(cond, d, s_) f[d] = (cond) ? f[s_] : f[d]
template <class R>
__device__ __forceinline__ void xsym_bot(R f[Q], int y, int z, int ny, int nz)
{
    const bool y0 = y == 0, y1 = y == ny - 1, z0 = z == 0, z1 = z == nz - 1;
    const bool a = y0 && z0, b = y0 && z1 && !z0, c = y1 && z0 && !y0, d = y1 && z1 && !y0 && !z0;
    const bool e = y0 && !a && !b, h = y1 && !y0 && !c && !d;
    LBM_SYM_COPY(a, 2, 4); LBM_SYM_COPY(a, 13, 18); LBM_SYM_COPY(a, 11, 18); LBM_SYM_COPY(a, 16, 18);
    LBM_SYM_COPY(a, 6, 7); LBM_SYM_COPY(a, 9, 14); LBM_SYM_COPY(a, 12, 17);
    LBM_SYM_COPY(b, 4, 2); LBM_SYM_COPY(b, 11, 13); LBM_SYM_COPY(b, 18, 13); LBM_SYM_COPY(b, 16, 13);
    LBM_SYM_COPY(b, 6, 7); LBM_SYM_COPY(b, 14, 9); LBM_SYM_COPY(b, 17, 12);
    LBM_SYM_COPY(c, 4, 2); LBM_SYM_COPY(c, 11, 16); LBM_SYM_COPY(c, 18, 16); LBM_SYM_COPY(c, 13, 16);
    LBM_SYM_COPY(c, 7, 6); LBM_SYM_COPY(c, 9, 14); LBM_SYM_COPY(c, 12, 17);
    LBM_SYM_COPY(d, 4, 2); LBM_SYM_COPY(d, 16, 11); LBM_SYM_COPY(d, 18, 11); LBM_SYM_COPY(d, 13, 11);
    LBM_SYM_COPY(d, 7, 6); LBM_SYM_COPY(d, 14, 9); LBM_SYM_COPY(d, 17, 12);
    LBM_SYM_COPY(e, 2, 4); LBM_SYM_COPY(e, 11, 13); LBM_SYM_COPY(e, 16, 18); LBM_SYM_COPY(e, 8, 5);
    LBM_SYM_COPY(h, 4, 2); LBM_SYM_COPY(h, 13, 11); LBM_SYM_COPY(h, 18, 16); LBM_SYM_COPY(h, 5, 8);
    f[1] = f[3]; f[5] = f[6]; f[8] = f[7]; f[10] = f[12]; f[15] = f[17];
}

// x = nx-1 symmetry plane, xsymmetry_top (MRTtest/DynInlet3.cu:1077-1137).
template <class R>
__device__ __forceinline__ void xsym_top(R f[Q], int y, int z, int ny, int nz)
{
    const bool y0 = y == 0, y1 = y == ny - 1, z0 = z == 0, z1 = z == nz - 1;
    const bool a = y0 && z0, b = y0 && z1 && !z0, c = y1 && z0 && !y0, d = y1 && z1 && !y0 && !z0;
    const bool e = y0 && !a && !b, h = y1 && !y0 && !c && !d;
    LBM_SYM_COPY(a, 2, 4); LBM_SYM_COPY(a, 13, 18); LBM_SYM_COPY(a, 11, 18); LBM_SYM_COPY(a, 16, 18);
    LBM_SYM_COPY(a, 5, 8); LBM_SYM_COPY(a, 9, 14); LBM_SYM_COPY(a, 10, 15);
    LBM_SYM_COPY(b, 2, 4); LBM_SYM_COPY(b, 11, 13); LBM_SYM_COPY(b, 18, 13); LBM_SYM_COPY(b, 16, 13);
    LBM_SYM_COPY(b, 5, 8); LBM_SYM_COPY(b, 14, 9); LBM_SYM_COPY(b, 15, 10);
    LBM_SYM_COPY(c, 4, 2); LBM_SYM_COPY(c, 18, 16); LBM_SYM_COPY(c, 11, 16); LBM_SYM_COPY(c, 13, 16);
    LBM_SYM_COPY(c, 8, 5); LBM_SYM_COPY(c, 9, 14); LBM_SYM_COPY(c, 10, 15);
    LBM_SYM_COPY(d, 4, 2); LBM_SYM_COPY(d, 13, 11); LBM_SYM_COPY(d, 16, 11); LBM_SYM_COPY(d, 18, 11);
    LBM_SYM_COPY(d, 8, 5); LBM_SYM_COPY(d, 14, 9); LBM_SYM_COPY(d, 15, 10);
    LBM_SYM_COPY(e, 2, 4); LBM_SYM_COPY(e, 11, 13); LBM_SYM_COPY(e, 16, 18); LBM_SYM_COPY(e, 5, 8);
    LBM_SYM_COPY(h, 4, 2); LBM_SYM_COPY(h, 13, 11); LBM_SYM_COPY(h, 18, 16); LBM_SYM_COPY(h, 8, 5);
    f[3] = f[1]; f[6] = f[5]; f[7] = f[8]; f[12] = f[10]; f[17] = f[15];
}
#undef LBM_SYM_COPY

} // namespace lbm
