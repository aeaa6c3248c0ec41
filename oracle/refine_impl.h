/*
 * refine_impl.h -- body of the coarse-fine interface restatement, included
 * twice by refine_impl.c: R = float (suffix _f32) and R = double (_f64).
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into liblbm.so).
 *
 * Requires: R, FN(name), C_(literal) (the literal as an R) and SQRT_.
 * The float instance keeps the operation order and the float constants of
 * the reference's interface code; the double one is the same sequence in
 * double.
 */

void FN(orc_phys_moments)(const R f[19], R mom[9])
{
    mom[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] +
             f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    mom[1] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + f[10] - f[12] + f[15] - f[17];
    mom[2] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + f[11] - f[13] + f[16] - f[18];
    mom[3] = f[9] + f[10] + f[11] + f[12] + f[13] - f[14] - f[15] - f[16] - f[17] - f[18];
    mom[4] = C_(2.) * f[1] - f[2] + C_(2.) * f[3] - f[4] + f[5] + f[6] + f[7] + f[8] - f[9] + f[10] - C_(2.) * f[11] +
             f[12] - C_(2.) * f[13] - f[14] + f[15] - C_(2.) * f[16] + f[17] - C_(2.) * f[18];
    mom[5] = f[2] + f[4] + f[5] + f[6] + f[7] + f[8] - f[9] - f[10] - f[12] - f[14] - f[15] - f[17];
    mom[6] = f[5] - f[6] + f[7] - f[8];
    mom[7] = f[11] - f[13] - f[16] + f[18];
    mom[8] = f[10] - f[12] - f[15] + f[17];
}

/* StrainRate with omega = 1 and m1 = 0 as coded (:367-393) */
void FN(orc_strain_rate)(const R m[9], R S[6])
{
    const R omega = C_(1.), m1 = C_(0.);
    const R u = m[1], v = m[2], w = m[3];
    const R m9 = m[4] - (C_(2.) * u * u - (v * v + w * w));
    const R m11 = m[5] - (v * v - w * w);
    const R m13 = m[6] - (u * v);
    const R m14 = m[7] - (v * w);
    const R m15 = m[8] - (u * w);
    S[0] = -C_(0.026315789) * (m1 + C_(19.) * omega * m9);
    S[1] = -C_(0.013157895) * (C_(2.) * m1 - C_(19.) * omega * (m9 - C_(3.) * m11));
    S[2] = -C_(0.013157895) * (C_(2.) * m1 - C_(19.) * omega * (m9 + C_(3.) * m11));
    S[3] = -C_(1.5) * omega * m13;
    S[4] = -C_(1.5) * omega * m14;
    S[5] = -C_(1.5) * omega * m15;
}

/* moments with the non-equilibrium part of the second moments scaled by SF,
 * everything else at equilibrium, then InvertMoments.  InvertPhysicalMoments
 * (:437-481) takes the higher moments from mrt_meq (eps = +475/63 j^2,
 * 3pixx = piww = 0), which is not the equilibrium the LES-family collide
 * relaxes to (bgk_meq: eps = 3 rho - 11/2 j^2, 3pixx = -meq9/6, piww =
 * -meq11/2); eq_kind selects which one is used. */
static void FN(invert_phys)(R f[19], const R mom[9], R SF, int eq_kind)
{
    R m[19] = {0};
    const R u = mom[1], v = mom[2], w = mom[3];
    const R usq = u * u + v * v + w * w;
    m[0] = mom[0];
    m[1] = -C_(11.) * mom[0] + C_(19.) * usq;
    if (eq_kind == ORC_EQ_MRT_MEQ) {
        m[2] = C_(7.53968254) * usq;
    } else {
        const R m9eq = C_(2.) * u * u - (v * v + w * w), m11eq = v * v - w * w;
        m[2] = C_(3.) * mom[0] - C_(5.5) * usq;
        m[10] = eq_kind == ORC_EQ_BGK_MEQ ? -C_(0.5) * m9eq * (C_(1.0) / C_(3.0)) : -C_(0.5) * m9eq;
        m[12] = -C_(0.5) * m11eq;
    }
    m[3] = u; m[4] = -C_(0.666666667) * u;
    m[5] = v; m[6] = -C_(0.666666667) * v;
    m[7] = w; m[8] = -C_(0.666666667) * w;
    m[9] = mom[4] * SF + (C_(1.) - SF) * (C_(2.) * u * u - (v * v + w * w));
    m[11] = mom[5] * SF + (C_(1.) - SF) * (v * v - w * w);
    m[13] = mom[6] * SF + (C_(1.) - SF) * u * v;
    m[14] = mom[7] * SF + (C_(1.) - SF) * v * w;
    m[15] = mom[8] * SF + (C_(1.) - SF) * u * w;
    double M[19][19];
    orc_moment_matrix(M);
    static const int norm[19] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8};
    for (int i = 0; i < 19; ++i) {
        R s = 0;
        for (int k = 0; k < 19; ++k)
            if (M[k][i] != 0) s += ((R)M[k][i] / (R)norm[k]) * m[k];
        f[i] = s;
    }
}

static R FN(smag)(const R mom[9])
{
    R S[6];
    FN(orc_strain_rate)(mom, S);
    return SQRT_(C_(2.) * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + C_(2.) * S[3] * S[3] + C_(2.) * S[4] * S[4] +
                        C_(2.) * S[5] * S[5]));
}

/* InvertPhysicalMoments_LES_cf (:540-595) */
void FN(orc_invert_phys_les_cf)(R f[19], const R mom[9], R omega_c, R cs, int eq_kind)
{
    R tau_c = C_(1.) / omega_c;
    const R Smag_c = FN(smag)(mom);
    const R tau_f = C_(2.) * tau_c - C_(0.5) + C_(1.5) * Smag_c * cs;
    const R omega_f = C_(1.) / tau_f;
    tau_c = tau_c + Smag_c * cs;
    omega_c = C_(1.) / tau_c;
    const R SF = (C_(0.5) * omega_c * (C_(1.) - omega_f)) / ((C_(1.) - omega_c) * omega_f);
    FN(invert_phys)(f, mom, SF, eq_kind);
}

/* InvertPhysicalMoments_LES_fc (:483-538) */
void FN(orc_invert_phys_les_fc)(R f[19], const R mom[9], R omega_f, R cs, int eq_kind)
{
    R tau_f = C_(1.) / omega_f;
    const R Smag_f = FN(smag)(mom);
    R tau_c = tau_f + C_(0.5) + C_(12.) * Smag_f * cs;
    tau_c *= C_(0.5);
    const R omega_c = C_(1.) / tau_c;
    tau_f = tau_f + Smag_f * cs;
    omega_f = C_(1.) / tau_f;
    const R SF = (C_(1.) - omega_c) * omega_f / (C_(0.5) * omega_c * (C_(1.) - omega_f));
    FN(invert_phys)(f, mom, SF, eq_kind);
}

/* ORDER-2 correction of the interpolated momentum (:2312-2357, FC form
 * :2477-2532 with xpr = ypr = zpr = 0).  Corners 1..8 = (x,y,z) bits
 * 000,100,010,110,001,101,011,111 (x fastest). */
void FN(orc_order2_correction)(const R c[8][9], R mom[9], R xf, R yf, R zf, int fc_form)
{
    R Sc[8][6];
    for (int k = 0; k < 8; ++k) FN(orc_strain_rate)(c[k], Sc[k]);
#define SXX(k) Sc[(k)-1][0]
#define SYY(k) Sc[(k)-1][1]
#define SZZ(k) Sc[(k)-1][2]
#define SXY(k) Sc[(k)-1][3]
#define SYZ(k) Sc[(k)-1][4]
#define SXZ(k) Sc[(k)-1][5]
#define U(k) c[(k)-1][1]
#define V(k) c[(k)-1][2]
#define W(k) c[(k)-1][3]
    const R cx = -((SXX(8) - SXX(7) + SXX(6) - SXX(5) + SXX(4) - SXX(3) + SXX(2) - SXX(1))) * C_(0.03125);
    const R cy = -((SXY(8) + SXY(7) - SXY(6) - SXY(5) + SXY(4) + SXY(3) - SXY(2) - SXY(1)) - V(8) + V(7) +
                       V(6) - V(5) - V(4) + V(3) + V(2) - V(1)) * C_(0.0625);
    const R cz = -((SXZ(8) + SXZ(7) + SXZ(6) + SXZ(5) - SXZ(4) - SXZ(3) - SXZ(2) - SXZ(1)) - W(8) + W(7) -
                       W(6) + W(5) + W(4) - W(3) + W(2) - W(1)) * C_(0.0625);
    const R dx = -((SXY(8) - SXY(7) + SXY(6) - SXY(5) + SXY(4) - SXY(3) + SXY(2) - SXY(1)) - U(8) + U(7) +
                       U(6) - U(5) - U(4) + U(3) + U(2) - U(1)) * C_(0.0625);
    const R dy = -((SYY(8) + SYY(7) - SYY(6) - SYY(5) + SYY(4) + SYY(3) - SYY(2) - SYY(1))) * C_(0.03125);
    const R dz = -((SYZ(8) + SYZ(7) + SYZ(6) + SYZ(5) - SYZ(4) - SYZ(3) - SYZ(2) - SYZ(1)) - W(8) - W(7) +
                       W(6) + W(5) + W(4) + W(3) - W(2) - W(1)) * C_(0.0625);
    const R ex = -((SXZ(8) - SXZ(7) + SXZ(6) - SXZ(5) + SXZ(4) - SXZ(3) + SXZ(2) - SXZ(1)) - U(8) + U(7) -
                       U(6) + U(5) + U(4) - U(3) + U(2) - U(1)) * C_(0.0625);
    const R ey = -((SYZ(8) + SYZ(7) - SYZ(6) - SYZ(5) + SYZ(4) + SYZ(3) - SYZ(2) - SYZ(1)) - V(8) - V(7) +
                       V(6) + V(5) + V(4) + V(3) - V(2) - V(1)) * C_(0.0625);
    const R ez = -((SZZ(8) + SZZ(7) + SZZ(6) + SZZ(5) - SZZ(4) - SZZ(3) - SZZ(2) - SZZ(1))) * C_(0.03125);
#undef SXX
#undef SYY
#undef SZZ
#undef SXY
#undef SYZ
#undef SXZ
#undef U
#undef V
#undef W
    R xpr = C_(0.), ypr = C_(0.), zpr = C_(0.);
    if (!fc_form) {
        xpr = C_(4.) * xf * xf - C_(4.) * xf + C_(1.);
        ypr = C_(4.) * yf * yf - C_(4.) * yf + C_(1.);
        zpr = C_(4.) * zf * zf - C_(4.) * zf + C_(1.);
    }
    mom[1] += cx * (C_(1.) - xpr) + cy * (C_(1.) - ypr) + cz * (C_(1.) - zpr);
    mom[2] += dx * (C_(1.) - xpr) + dy * (C_(1.) - ypr) + dz * (C_(1.) - zpr);
    mom[3] += ex * (C_(1.) - xpr) + ey * (C_(1.) - ypr) + ez * (C_(1.) - zpr);
}

static inline int FN(is_solid_code)(uint8_t c) { return c == ORC_SOLID || c == ORC_SOLID_FORCE; }

void FN(orc_extract)(const orc_params* pc, const orc_refine* r, const R* fc, R* M)
{
    const int nx = pc->nx, ny = pc->ny, nz = pc->nz;
    const size_t N = (size_t)nx * ny * nz;
    const int ncx = r->nf[0] / 2 + 1, ncy = r->nf[1] / 2 + 1, ncz = r->nf[2] / 2 + 1;
    const size_t NC = (size_t)ncx * ncy * ncz;
    for (int k = 0; k < ncz; ++k)
        for (int j = 0; j < ncy; ++j)
            for (int i = 0; i < ncx; ++i) {
                const size_t c = (size_t)(r->a[0] + i) + (size_t)nx * ((r->a[1] + j) + (size_t)ny * (r->a[2] + k));
                R f[19], mom[9];
                for (int q = 0; q < 19; ++q) f[q] = fc[q * N + c];
                FN(orc_phys_moments)(f, mom);
                const size_t o = (size_t)i + (size_t)ncx * (j + (size_t)ncy * k);
                for (int m = 0; m < 9; ++m) M[m * NC + o] = mom[m];
            }
}

void FN(orc_interp_cf)(const orc_refine* r, const R* Mold, const R* Mnew, R alpha,
                       const uint8_t* fmask, R* ff)
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
                if (fmask && FN(is_solid_code)(fmask[fcell])) continue;
                const int xm = i / 2, ym = j / 2, zm = k / 2;
                const R xf = (i & 1) ? C_(0.75) : C_(0.25), yf = (j & 1) ? C_(0.75) : C_(0.25),
                            zf = (k & 1) ? C_(0.75) : C_(0.25);
                R c[8][9];
                for (int n = 0; n < 8; ++n) {
                    const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
                    const size_t o = (size_t)(xm + dx) + (size_t)ncx * ((ym + dy) + (size_t)ncy * (zm + dz));
                    for (int m = 0; m < 9; ++m)
                        c[n][m] = alpha == C_(0.) ? Mold[m * NC + o]
                                               : (C_(1.) - alpha) * Mold[m * NC + o] + alpha * Mnew[m * NC + o];
                }
                R mom[9];
                for (int m = 0; m < 9; ++m) {
                    /* trilinear_interp (:165-175) */
                    mom[m] = c[0][m] * (C_(1.) - xf) * (C_(1.) - yf) * (C_(1.) - zf) + c[1][m] * xf * (C_(1.) - yf) * (C_(1.) - zf) +
                             c[2][m] * (C_(1.) - xf) * yf * (C_(1.) - zf) + c[3][m] * xf * yf * (C_(1.) - zf) +
                             c[4][m] * (C_(1.) - xf) * (C_(1.) - yf) * zf + c[5][m] * xf * (C_(1.) - yf) * zf +
                             c[6][m] * (C_(1.) - xf) * yf * zf + c[7][m] * xf * yf * zf;
                }
                FN(orc_order2_correction)(c, mom, xf, yf, zf, 0);
                R f[19];
                FN(orc_invert_phys_les_cf)(f, mom, (R)r->omega_c, (R)r->cs, r->eq_kind);
                for (int q = 0; q < 19; ++q) ff[q * NF + fcell] = f[q];
            }
}

void FN(orc_interp_fc)(const orc_params* pc, const orc_refine* r, const R* ff, const uint8_t* cmask, R* fc)
{
    const int nx = pc->nx, ny = pc->ny, nz = pc->nz;
    const size_t N = (size_t)nx * ny * nz;
    const int nfx = r->nf[0], nfy = r->nf[1], nfz = r->nf[2];
    const size_t NF = (size_t)nfx * nfy * nfz;
    for (int z = r->a[2] + 2; z <= r->a[2] + nfz / 2 - 2; ++z)
        for (int y = r->a[1] + 2; y <= r->a[1] + nfy / 2 - 2; ++y)
            for (int x = r->a[0] + 2; x <= r->a[0] + nfx / 2 - 2; ++x) {
                const size_t cc = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
                if (cmask && FN(is_solid_code)(cmask[cc])) continue;
                const int xm = 2 * (x - r->a[0]) - 1, ym = 2 * (y - r->a[1]) - 1, zm = 2 * (z - r->a[2]) - 1;
                R c[8][9];
                for (int n = 0; n < 8; ++n) {
                    const int dx = n & 1, dy = (n >> 1) & 1, dz = (n >> 2) & 1;
                    const size_t o = (size_t)(xm + dx) + (size_t)nfx * ((ym + dy) + (size_t)nfy * (zm + dz));
                    R f[19];
                    for (int q = 0; q < 19; ++q) f[q] = ff[q * NF + o];
                    FN(orc_phys_moments)(f, c[n]);
                }
                R mom[9];
                for (int m = 0; m < 9; ++m)
                    mom[m] = C_(0.125) * (c[0][m] + c[1][m] + c[2][m] + c[3][m] + c[4][m] + c[5][m] + c[6][m] + c[7][m]);
                FN(orc_order2_correction)(c, mom, C_(0.), C_(0.), C_(0.), 1);
                R f[19];
                FN(orc_invert_phys_les_fc)(f, mom, (R)r->omega_f, (R)r->cs, r->eq_kind);
                for (int q = 0; q < 19; ++q) fc[q * N + cc] = f[q];
            }
}

void FN(orc_refine_step)(const orc_params* pc, const orc_params* pf, const orc_refine* r,
                         const uint8_t* cmask, const uint8_t* fmask, const R* u_in,
                         R* ca, R* cb, R* fa, R* fb, R* Mold, R* Mnew,
                         double* fine_force)
{
    const size_t N = (size_t)pc->nx * pc->ny * pc->nz;
    const size_t NF = (size_t)pf->nx * pf->ny * pf->nz;
    const size_t NC = (size_t)(r->nf[0] / 2 + 1) * (r->nf[1] / 2 + 1) * (r->nf[2] / 2 + 1);
    FN(orc_step)(pc, cmask, u_in, ca, cb, 0);
    memcpy(ca, cb, sizeof(R) * 19 * N);
    FN(orc_extract)(pc, r, ca, Mnew);
    for (int s = 0; s < 2; ++s) {
        if (s == 0) FN(orc_interp_cf)(r, Mold, Mnew, C_(0.), fmask, fa);
        else if (r->temporal) FN(orc_interp_cf)(r, Mold, Mnew, C_(0.5), fmask, fa);
        /* the fine level's momentum exchange of its first sub-step
           (update_inn_LR with flag_F on i == 0, MRTtest/DynInlet3.cu:3183) */
        FN(orc_step)(pf, fmask, 0, fa, fb, s == 0 ? fine_force : 0);
        memcpy(fa, fb, sizeof(R) * 19 * NF);
    }
    FN(orc_interp_fc)(pc, r, fa, cmask, ca);
    memcpy(Mold, Mnew, sizeof(R) * 9 * NC);
}
