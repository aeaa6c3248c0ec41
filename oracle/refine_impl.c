/*
 * refine_impl.c -- CPU restatement of the reference's coarse-fine interface
 * (float).  TEST INFRASTRUCTURE ONLY: the checker of the HIP refinement path.
 *
 * Follows MRTtest/DynInlet3.cu: PhysicalMoments :396-407, StrainRate :367-393,
 * InvertPhysicalMoments_LES_fc :483-538, _LES_cf :540-595, InterpCF
 * :2202-2382 (trilinear :165-175, ORDER-2 correction :2312-2357), InterpFC
 * :2383-2560, and the sub-step sequence of main() :3180-3245.  The frame is
 * two fine nodes thick on all six faces (the reference's patch spans z
 * periodically, so it has x/y faces only); the coarse footprint moments are
 * optionally interpolated to t+1/2 for the second sub-step (the reference
 * reuses the state of time t).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define LRFACTOR 0.5f

void orc_phys_moments_f32(const float f[19], float mom[9])
{
    mom[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] +
             f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    mom[1] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + f[10] - f[12] + f[15] - f[17];
    mom[2] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + f[11] - f[13] + f[16] - f[18];
    mom[3] = f[9] + f[10] + f[11] + f[12] + f[13] - f[14] - f[15] - f[16] - f[17] - f[18];
    mom[4] = 2.f * f[1] - f[2] + 2.f * f[3] - f[4] + f[5] + f[6] + f[7] + f[8] - f[9] + f[10] - 2.f * f[11] +
             f[12] - 2.f * f[13] - f[14] + f[15] - 2.f * f[16] + f[17] - 2.f * f[18];
    mom[5] = f[2] + f[4] + f[5] + f[6] + f[7] + f[8] - f[9] - f[10] - f[12] - f[14] - f[15] - f[17];
    mom[6] = f[5] - f[6] + f[7] - f[8];
    mom[7] = f[11] - f[13] - f[16] + f[18];
    mom[8] = f[10] - f[12] - f[15] + f[17];
}

/* StrainRate with omega = 1 and m1 = 0 as coded (:367-393) */
void orc_strain_rate_f32(const float m[9], float S[6])
{
    const float omega = 1.f, m1 = 0.f;
    const float u = m[1], v = m[2], w = m[3];
    const float m9 = m[4] - (2.f * u * u - (v * v + w * w));
    const float m11 = m[5] - (v * v - w * w);
    const float m13 = m[6] - (u * v);
    const float m14 = m[7] - (v * w);
    const float m15 = m[8] - (u * w);
    S[0] = -0.026315789f * (m1 + 19.f * omega * m9);
    S[1] = -0.013157895f * (2.f * m1 - 19.f * omega * (m9 - 3.f * m11));
    S[2] = -0.013157895f * (2.f * m1 - 19.f * omega * (m9 + 3.f * m11));
    S[3] = -1.5f * omega * m13;
    S[4] = -1.5f * omega * m14;
    S[5] = -1.5f * omega * m15;
}

/* moments with the non-equilibrium part of the second moments scaled by SF,
 * everything else at equilibrium, then InvertMoments.  InvertPhysicalMoments
 * (:437-481) takes the higher moments from mrt_meq (eps = +475/63 j^2,
 * 3pixx = piww = 0), which is not the equilibrium the LES-family collide
 * relaxes to (bgk_meq: eps = 3 rho - 11/2 j^2, 3pixx = -meq9/6, piww =
 * -meq11/2); eq_kind selects which one is used. */
static void invert_phys(float f[19], const float mom[9], float SF, int eq_kind)
{
    float m[19] = {0};
    const float u = mom[1], v = mom[2], w = mom[3];
    const float usq = u * u + v * v + w * w;
    m[0] = mom[0];
    m[1] = -11.f * mom[0] + 19.f * usq;
    if (eq_kind == ORC_EQ_MRT_MEQ) {
        m[2] = 7.53968254f * usq;
    } else {
        const float m9eq = 2.f * u * u - (v * v + w * w), m11eq = v * v - w * w;
        m[2] = 3.f * mom[0] - 5.5f * usq;
        m[10] = eq_kind == ORC_EQ_BGK_MEQ ? -0.5f * m9eq * (1.0f / 3.0f) : -0.5f * m9eq;
        m[12] = -0.5f * m11eq;
    }
    m[3] = u; m[4] = -0.666666667f * u;
    m[5] = v; m[6] = -0.666666667f * v;
    m[7] = w; m[8] = -0.666666667f * w;
    m[9] = mom[4] * SF + (1.f - SF) * (2.f * u * u - (v * v + w * w));
    m[11] = mom[5] * SF + (1.f - SF) * (v * v - w * w);
    m[13] = mom[6] * SF + (1.f - SF) * u * v;
    m[14] = mom[7] * SF + (1.f - SF) * v * w;
    m[15] = mom[8] * SF + (1.f - SF) * u * w;
    double M[19][19];
    orc_moment_matrix(M);
    static const int norm[19] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8};
    for (int i = 0; i < 19; ++i) {
        float s = 0;
        for (int k = 0; k < 19; ++k)
            if (M[k][i] != 0) s += ((float)M[k][i] / (float)norm[k]) * m[k];
        f[i] = s;
    }
}

static float smag(const float mom[9])
{
    float S[6];
    orc_strain_rate_f32(mom, S);
    return sqrtf(2.f * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.f * S[3] * S[3] + 2.f * S[4] * S[4] +
                        2.f * S[5] * S[5]));
}

/* InvertPhysicalMoments_LES_cf (:540-595) */
void orc_invert_phys_les_cf_f32(float f[19], const float mom[9], float omega_c, float cs, int eq_kind)
{
    float tau_c = 1.f / omega_c;
    const float Smag_c = smag(mom);
    const float tau_f = 2.f * tau_c - 0.5f + 1.5f * Smag_c * cs;
    const float omega_f = 1.f / tau_f;
    tau_c = tau_c + Smag_c * cs;
    omega_c = 1.f / tau_c;
    const float SF = (LRFACTOR * omega_c * (1.f - omega_f)) / ((1.f - omega_c) * omega_f);
    invert_phys(f, mom, SF, eq_kind);
}

/* InvertPhysicalMoments_LES_fc (:483-538) */
void orc_invert_phys_les_fc_f32(float f[19], const float mom[9], float omega_f, float cs, int eq_kind)
{
    float tau_f = 1.f / omega_f;
    const float Smag_f = smag(mom);
    float tau_c = tau_f + 0.5f + 12.f * Smag_f * cs;
    tau_c *= 0.5f;
    const float omega_c = 1.f / tau_c;
    tau_f = tau_f + Smag_f * cs;
    omega_f = 1.f / tau_f;
    const float SF = (1.f - omega_c) * omega_f / (LRFACTOR * omega_c * (1.f - omega_f));
    invert_phys(f, mom, SF, eq_kind);
}

/* ORDER-2 correction of the interpolated momentum (:2312-2357, FC form
 * :2477-2532 with xpr = ypr = zpr = 0).  Corners 1..8 = (x,y,z) bits
 * 000,100,010,110,001,101,011,111 (x fastest). */
void orc_order2_correction_f32(const float c[8][9], float mom[9], float xf, float yf, float zf, int fc_form)
{
    float Sc[8][6];
    for (int k = 0; k < 8; ++k) orc_strain_rate_f32(c[k], Sc[k]);
#define SXX(k) Sc[(k)-1][0]
#define SYY(k) Sc[(k)-1][1]
#define SZZ(k) Sc[(k)-1][2]
#define SXY(k) Sc[(k)-1][3]
#define SYZ(k) Sc[(k)-1][4]
#define SXZ(k) Sc[(k)-1][5]
#define U(k) c[(k)-1][1]
#define V(k) c[(k)-1][2]
#define W(k) c[(k)-1][3]
    const float cx = -((SXX(8) - SXX(7) + SXX(6) - SXX(5) + SXX(4) - SXX(3) + SXX(2) - SXX(1))) * 0.03125f;
    const float cy = -((SXY(8) + SXY(7) - SXY(6) - SXY(5) + SXY(4) + SXY(3) - SXY(2) - SXY(1)) - V(8) + V(7) +
                       V(6) - V(5) - V(4) + V(3) + V(2) - V(1)) * 0.0625f;
    const float cz = -((SXZ(8) + SXZ(7) + SXZ(6) + SXZ(5) - SXZ(4) - SXZ(3) - SXZ(2) - SXZ(1)) - W(8) + W(7) -
                       W(6) + W(5) + W(4) - W(3) + W(2) - W(1)) * 0.0625f;
    const float dx = -((SXY(8) - SXY(7) + SXY(6) - SXY(5) + SXY(4) - SXY(3) + SXY(2) - SXY(1)) - U(8) + U(7) +
                       U(6) - U(5) - U(4) + U(3) + U(2) - U(1)) * 0.0625f;
    const float dy = -((SYY(8) + SYY(7) - SYY(6) - SYY(5) + SYY(4) + SYY(3) - SYY(2) - SYY(1))) * 0.03125f;
    const float dz = -((SYZ(8) + SYZ(7) + SYZ(6) + SYZ(5) - SYZ(4) - SYZ(3) - SYZ(2) - SYZ(1)) - W(8) - W(7) +
                       W(6) + W(5) + W(4) + W(3) - W(2) - W(1)) * 0.0625f;
    const float ex = -((SXZ(8) - SXZ(7) + SXZ(6) - SXZ(5) + SXZ(4) - SXZ(3) + SXZ(2) - SXZ(1)) - U(8) + U(7) -
                       U(6) + U(5) + U(4) - U(3) + U(2) - U(1)) * 0.0625f;
    const float ey = -((SYZ(8) + SYZ(7) - SYZ(6) - SYZ(5) + SYZ(4) + SYZ(3) - SYZ(2) - SYZ(1)) - V(8) - V(7) +
                       V(6) + V(5) + V(4) + V(3) - V(2) - V(1)) * 0.0625f;
    const float ez = -((SZZ(8) + SZZ(7) + SZZ(6) + SZZ(5) - SZZ(4) - SZZ(3) - SZZ(2) - SZZ(1))) * 0.03125f;
#undef SXX
#undef SYY
#undef SZZ
#undef SXY
#undef SYZ
#undef SXZ
#undef U
#undef V
#undef W
    float xpr = 0.f, ypr = 0.f, zpr = 0.f;
    if (!fc_form) {
        xpr = 4.f * xf * xf - 4.f * xf + 1.f;
        ypr = 4.f * yf * yf - 4.f * yf + 1.f;
        zpr = 4.f * zf * zf - 4.f * zf + 1.f;
    }
    mom[1] += cx * (1.f - xpr) + cy * (1.f - ypr) + cz * (1.f - zpr);
    mom[2] += dx * (1.f - xpr) + dy * (1.f - ypr) + dz * (1.f - zpr);
    mom[3] += ex * (1.f - xpr) + ey * (1.f - ypr) + ez * (1.f - zpr);
}

static inline int is_solid_code(uint8_t c) { return c == ORC_SOLID || c == ORC_SOLID_FORCE; }

void orc_extract_f32(const orc_params* pc, const orc_refine* r, const float* fc, float* M)
{
    const int nx = pc->nx, ny = pc->ny, nz = pc->nz;
    const size_t N = (size_t)nx * ny * nz;
    const int ncx = r->nf[0] / 2 + 1, ncy = r->nf[1] / 2 + 1, ncz = r->nf[2] / 2 + 1;
    const size_t NC = (size_t)ncx * ncy * ncz;
    for (int k = 0; k < ncz; ++k)
        for (int j = 0; j < ncy; ++j)
            for (int i = 0; i < ncx; ++i) {
                const size_t c = (size_t)(r->a[0] + i) + (size_t)nx * ((r->a[1] + j) + (size_t)ny * (r->a[2] + k));
                float f[19], mom[9];
                for (int q = 0; q < 19; ++q) f[q] = fc[q * N + c];
                orc_phys_moments_f32(f, mom);
                const size_t o = (size_t)i + (size_t)ncx * (j + (size_t)ncy * k);
                for (int m = 0; m < 9; ++m) M[m * NC + o] = mom[m];
            }
}

void orc_interp_cf_f32(const orc_refine* r, const float* Mold, const float* Mnew, float alpha,
                       const uint8_t* fmask, float* ff)
{
    const int nfx = r->nf[0], nfy = r->nf[1], nfz = r->nf[2];
    const size_t NF = (size_t)nfx * nfy * nfz;
    const int ncx = nfx / 2 + 1, ncy = nfy / 2 + 1, ncz = nfz / 2 + 1;
    const size_t NC = (size_t)ncx * ncy * ncz;
    for (int k = 0; k < nfz; ++k)
        for (int j = 0; j < nfy; ++j)
            for (int i = 0; i < nfx; ++i) {
                const int edge = !(i >= 2 && i < nfx - 2 && j >= 2 && j < nfy - 2 && k >= 2 && k < nfz - 2);
                if (!edge) continue;
                const size_t fcell = (size_t)i + (size_t)nfx * (j + (size_t)nfy * k);
                if (fmask && is_solid_code(fmask[fcell])) continue;
                const int xm = i / 2, ym = j / 2, zm = k / 2;
                const float xf = (i & 1) ? 0.75f : 0.25f, yf = (j & 1) ? 0.75f : 0.25f,
                            zf = (k & 1) ? 0.75f : 0.25f;
                float c[8][9];
                for (int n = 0; n < 8; ++n) {
                    const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
                    const size_t o = (size_t)(xm + dx) + (size_t)ncx * ((ym + dy) + (size_t)ncy * (zm + dz));
                    for (int m = 0; m < 9; ++m)
                        c[n][m] = alpha == 0.f ? Mold[m * NC + o]
                                               : (1.f - alpha) * Mold[m * NC + o] + alpha * Mnew[m * NC + o];
                }
                float mom[9];
                for (int m = 0; m < 9; ++m) {
                    /* trilinear_interp (:165-175) */
                    mom[m] = c[0][m] * (1.f - xf) * (1.f - yf) * (1.f - zf) + c[1][m] * xf * (1.f - yf) * (1.f - zf) +
                             c[2][m] * (1.f - xf) * yf * (1.f - zf) + c[3][m] * xf * yf * (1.f - zf) +
                             c[4][m] * (1.f - xf) * (1.f - yf) * zf + c[5][m] * xf * (1.f - yf) * zf +
                             c[6][m] * (1.f - xf) * yf * zf + c[7][m] * xf * yf * zf;
                }
                orc_order2_correction_f32(c, mom, xf, yf, zf, 0);
                float f[19];
                orc_invert_phys_les_cf_f32(f, mom, (float)r->omega_c, (float)r->cs, r->eq_kind);
                for (int q = 0; q < 19; ++q) ff[q * NF + fcell] = f[q];
            }
}

void orc_interp_fc_f32(const orc_params* pc, const orc_refine* r, const float* ff, const uint8_t* cmask, float* fc)
{
    const int nx = pc->nx, ny = pc->ny, nz = pc->nz;
    const size_t N = (size_t)nx * ny * nz;
    const int nfx = r->nf[0], nfy = r->nf[1], nfz = r->nf[2];
    const size_t NF = (size_t)nfx * nfy * nfz;
    for (int z = r->a[2] + 2; z <= r->a[2] + nfz / 2 - 2; ++z)
        for (int y = r->a[1] + 2; y <= r->a[1] + nfy / 2 - 2; ++y)
            for (int x = r->a[0] + 2; x <= r->a[0] + nfx / 2 - 2; ++x) {
                const size_t cc = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
                if (cmask && is_solid_code(cmask[cc])) continue;
                const int xm = 2 * (x - r->a[0]) - 1, ym = 2 * (y - r->a[1]) - 1, zm = 2 * (z - r->a[2]) - 1;
                float c[8][9];
                for (int n = 0; n < 8; ++n) {
                    const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
                    const size_t o = (size_t)(xm + dx) + (size_t)nfx * ((ym + dy) + (size_t)nfy * (zm + dz));
                    float f[19];
                    for (int q = 0; q < 19; ++q) f[q] = ff[q * NF + o];
                    orc_phys_moments_f32(f, c[n]);
                }
                float mom[9];
                for (int m = 0; m < 9; ++m)
                    mom[m] = 0.125f * (c[0][m] + c[1][m] + c[2][m] + c[3][m] + c[4][m] + c[5][m] + c[6][m] + c[7][m]);
                orc_order2_correction_f32(c, mom, 0.f, 0.f, 0.f, 1);
                float f[19];
                orc_invert_phys_les_fc_f32(f, mom, (float)r->omega_f, (float)r->cs, r->eq_kind);
                for (int q = 0; q < 19; ++q) fc[q * N + cc] = f[q];
            }
}

void orc_refine_step_f32(const orc_params* pc, const orc_params* pf, const orc_refine* r,
                         const uint8_t* cmask, const uint8_t* fmask, const float* u_in,
                         float* ca, float* cb, float* fa, float* fb, float* Mold, float* Mnew,
                         double* fine_force)
{
    const size_t N = (size_t)pc->nx * pc->ny * pc->nz;
    const size_t NF = (size_t)pf->nx * pf->ny * pf->nz;
    const size_t NC = (size_t)(r->nf[0] / 2 + 1) * (r->nf[1] / 2 + 1) * (r->nf[2] / 2 + 1);
    orc_step_f32(pc, cmask, u_in, ca, cb, 0);
    memcpy(ca, cb, sizeof(float) * 19 * N);
    orc_extract_f32(pc, r, ca, Mnew);
    for (int s = 0; s < 2; ++s) {
        if (s == 0) orc_interp_cf_f32(r, Mold, Mnew, 0.f, fmask, fa);
        else if (r->temporal) orc_interp_cf_f32(r, Mold, Mnew, 0.5f, fmask, fa);
        /* the fine level's momentum exchange of its first sub-step
           (update_inn_LR with flag_F on i == 0, MRTtest/DynInlet3.cu:3183) */
        orc_step_f32(pf, fmask, 0, fa, fb, s == 0 ? fine_force : 0);
        memcpy(fa, fb, sizeof(float) * 19 * NF);
    }
    orc_interp_fc_f32(pc, r, fa, cmask, ca);
    memcpy(Mold, Mnew, sizeof(float) * 9 * NC);
}
