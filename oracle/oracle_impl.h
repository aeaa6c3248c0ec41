/*
 * oracle_impl.h -- body of the CPU restatement, included twice by oracle.c:
 * once with R=float (suffix _f32) and once with R=double (suffix _f64).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into liblbm.so).
 *
 * Requires: R (arithmetic type), FN(name) (suffixing macro), MT / MIT (the
 * moment matrix and its inverse as R tables).
 */

/* m = M f with the d'Humieres 2002 basis of MRTtest/DynInlet3.cu:255-276
 * (Moments).  The integer rows come from the defining polynomials, built in
 * orc_moment_matrix(); the sums run left to right over i like the reference. */
void FN(orc_moments)(const R f[19], R m[19])
{
    for (int k = 0; k < 19; ++k) {
        R s = 0;
        for (int n = 0; n < g_rn[k]; ++n) {
            const int i = g_rnz[k][n];
            s += MT[k][i] * f[i];
        }
        m[k] = s;
    }
}

/* f = M^-1 m (InvertMoments, MRTtest/DynInlet3.cu:409-435): because the rows
 * of M are orthogonal, M^-1 = M^T diag(1/|row_k|^2). */
static void FN(invert_moments)(const R m[19], R f[19])
{
    for (int i = 0; i < 19; ++i) {
        R s = 0;
        for (int n = 0; n < g_cn[i]; ++n) {
            const int k = g_cnz[i][n];
            s += MIT[i][k] * m[k];
        }
        f[i] = s;
    }
}

/* BGK equilibrium w_i(rho + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u) with u the momentum
 * (CPU_3D/CPU_3D.cpp:135-153, MRTtest/DynInlet3.cu:694-712). */
void FN(orc_feq_bgk)(R rho, R u, R v, R w, R f[19])
{
    R usqr = u * u + v * v + w * w;
    for (int i = 0; i < 19; ++i) {
        R cu = (R)g_c[i][0] * u + (R)g_c[i][1] * v + (R)g_c[i][2] * w;
        f[i] = (R)g_w[i] * (rho + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usqr);
    }
}

/* initialize (MRTtest/DynInlet3.cu:1186-1216): mrt_meq (:329-345) then
 * InvertMoments.  mrt_meq leaves m10, m12, m16-18 at zero. */
void FN(orc_feq_mrt)(R rho, R u, R v, R w, R f[19])
{
    R m[19] = {0};
    R usqr = u * u + v * v + w * w;
    m[0] = rho;
    m[1] = (R)-11 * rho + (R)19 * usqr;
    m[2] = (R)(475.0 / 63.0) * usqr;
    m[3] = u; m[4] = (R)(-2.0 / 3.0) * u;
    m[5] = v; m[6] = (R)(-2.0 / 3.0) * v;
    m[7] = w; m[8] = (R)(-2.0 / 3.0) * w;
    m[9] = (R)2 * u * u - (v * v + w * w);
    m[11] = v * v - w * w;
    m[13] = u * v; m[14] = v * w; m[15] = u * w;
    FN(invert_moments)(m, f);
}

/* Equilibrium moments used by the LES-family collide: bgk_meq
 * (MRTtest/DynInlet3.cu:347-365).  Note meq[10] = -0.5*meq[9]/3 as coded
 * (the projection of the BGK equilibrium would give -0.5*meq[9]). */
static void FN(bgk_meq)(R rho, R u, R v, R w, R meq[19])
{
    R usqr = u * u + v * v + w * w;
    for (int k = 0; k < 19; ++k) meq[k] = 0;
    meq[0] = rho;
    meq[1] = (R)-11 * rho + (R)19 * usqr;
    meq[2] = (R)3 * rho - (R)5.5 * usqr;
    meq[3] = u; meq[4] = (R)(-2.0 / 3.0) * u;
    meq[5] = v; meq[6] = (R)(-2.0 / 3.0) * v;
    meq[7] = w; meq[8] = (R)(-2.0 / 3.0) * w;
    meq[9] = (R)2 * u * u - (v * v + w * w);
    meq[10] = (R)-0.5 * meq[9] * (R)(1.0 / 3.0);
    meq[11] = v * v - w * w;
    meq[12] = (R)-0.5 * meq[11];
    meq[13] = u * v; meq[14] = v * w; meq[15] = u * w;
}

/* Smagorinsky effective rate (MRTtest/DynInlet3.cu:628-735): Pi_neq from the
 * BGK equilibrium, Q = sqrt(Pi:Pi), tau = tau0/2 + sqrt(tau0^2 + 18 sqrt2 CS Q)/2. */
void FN(orc_les_omega)(const orc_params* p, const R f[19], R* omega_out, R* q_out)
{
    R rho = 0, u = 0, v = 0, w = 0;
    for (int i = 0; i < 19; ++i) {
        rho += f[i];
        u += (R)g_c[i][0] * f[i];
        v += (R)g_c[i][1] * f[i];
        w += (R)g_c[i][2] * f[i];
    }
    R feq[19];
    FN(orc_feq_bgk)(rho, u, v, w, feq);
    R pi[3][3] = {{0}};
    for (int i = 0; i < 19; ++i) {
        R d = f[i] - feq[i];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                pi[a][b] += (R)(g_c[i][a] * g_c[i][b]) * d;
    }
    R q = (R)sqrt((double)(pi[0][0] * pi[0][0] + pi[1][1] * pi[1][1] + pi[2][2] * pi[2][2] +
                           (R)2 * pi[0][1] * pi[0][1] + (R)2 * pi[1][2] * pi[1][2] +
                           (R)2 * pi[0][2] * pi[0][2]));
    R tau0 = (R)1 / (R)p->omega;
    R tau = (R)0.5 * tau0 +
            (R)0.5 * (R)sqrt((double)(tau0 * tau0 + (R)18 * (R)p->cs * (R)sqrt(2.0) * q));
    if (omega_out) *omega_out = (R)1 / tau;
    if (q_out) *q_out = q;
}

/* AddForce (MRTtest/DynInlet3.cu:229-253): f_i -= 3 w_i c_iy dpdy. */
static void FN(add_force)(R f[19], R dpdy)
{
    for (int i = 0; i < 19; ++i)
        if (g_c[i][1]) f[i] -= (R)3 * (R)g_w[i] * (R)g_c[i][1] * dpdy;
}

/* One-node collide.
 *  ORC_BGK:       CPU_3D/CPU_3D.cpp:116-160 (+LES omega and AddForce when the
 *                 LES-family options are set, as DynInlet3.cu:802-830 does)
 *  ORC_MRT_CPU3D: CPU_3D/CPU_3D.cpp:229-279, f -= M^-1 S (m - meq) with
 *                 meq = (.., -11rho+19j^2, -475/63 j^2, -2/3 j, .., 2jx^2-jy^2-jz^2, 0,
 *                 jy^2-jz^2, 0, jxjy, jyjz, jxjz, 0,0,0) and S = 1 except
 *                 s9 = s11 = s15 = omega (m13, m14 at rate 1, as coded at :246-247)
 *  ORC_MRT_LES:   MRTtest/DynInlet3.cu:759-789, m* = meq except m9,m11,m13,m14,m15
 *                 relaxed at the (LES) rate, then InvertMoments and AddForce
 *  ORC_MRT_DHUM:  as ORC_MRT_LES with the d'Humieres rates of :774-783 on the
 *                 other non-conserved moments */
void FN(orc_collide)(const orc_params* p, R f[19])
{
    R omega = (R)p->omega;
    int les = p->cs > 0 && p->collision != ORC_MRT_CPU3D;
    if (les) FN(orc_les_omega)(p, f, &omega, 0);

    if (p->collision == ORC_BGK) {
        R rho = 0, u = 0, v = 0, w = 0;
        for (int i = 0; i < 19; ++i) {
            rho += f[i];
            u += (R)g_c[i][0] * f[i];
            v += (R)g_c[i][1] * f[i];
            w += (R)g_c[i][2] * f[i];
        }
        R feq[19];
        FN(orc_feq_bgk)(rho, u, v, w, feq);
        for (int i = 0; i < 19; ++i) f[i] -= omega * (f[i] - feq[i]);
        if (p->dpdy != 0) FN(add_force)(f, (R)p->dpdy);
        return;
    }

    R m[19];
    FN(orc_moments)(f, m);
    R rho = m[0], u = m[3], v = m[5], w = m[7];

    if (p->collision == ORC_MRT_CPU3D) {
        R usqr = u * u + v * v + w * w;
        R meq[19] = {0};
        meq[0] = rho;
        meq[1] = (R)-11 * rho + (R)19 * usqr;
        meq[2] = (R)(-475.0 / 63.0) * usqr;
        meq[3] = u; meq[4] = (R)(-2.0 / 3.0) * u;
        meq[5] = v; meq[6] = (R)(-2.0 / 3.0) * v;
        meq[7] = w; meq[8] = (R)(-2.0 / 3.0) * w;
        meq[9] = (R)2 * u * u - (v * v + w * w);
        meq[11] = v * v - w * w;
        meq[13] = u * v; meq[14] = v * w; meq[15] = u * w;
        R d[19];
        for (int k = 0; k < 19; ++k) {
            R s = 1;
            if (k == 0 || k == 3 || k == 5 || k == 7) s = 0;
            if (k == 9 || k == 11 || k == 15) s = omega;
            d[k] = s * (m[k] - meq[k]);
        }
        R df[19];
        FN(invert_moments)(d, df);
        for (int i = 0; i < 19; ++i) f[i] -= df[i];
        return;
    }

    R meq[19];
    FN(bgk_meq)(rho, u, v, w, meq);
    for (int k = 0; k < 19; ++k) {
        R s;
        if (k == 0 || k == 3 || k == 5 || k == 7) continue;   /* conserved */
        if (k == 9 || k == 11 || k == 13 || k == 14 || k == 15) s = omega;
        else if (p->collision == ORC_MRT_DHUM) s = (R)p->s_dhum[k];
        else s = 1;
        meq[k] = m[k] - s * (m[k] - meq[k]);
    }
    FN(invert_moments)(meq, f);
    if (p->dpdy != 0) FN(add_force)(f, (R)p->dpdy);
}

/* West_Extrap (MRTtest/DynInlet3.cu:973-1017): rho from the neighbour's sum,
 * prescribed (u, 0, 0), every non-conserved moment copied from the neighbour. */
void FN(orc_west_extrap)(R f[19], R u)
{
    R m[19];
    FN(orc_moments)(f, m);
    m[3] = u; m[5] = 0; m[7] = 0;
    FN(invert_moments)(m, f);
}

/* South_Extrap (MRTtest/DynInlet3.cu:881-924: j = (0, v, 0)) and its
 * four-argument form of bgk_channel1/Re3000_Cyl2/DynInlet3.cu (j = (u, v, w),
 * the im 70 plane): rho from the neighbour's sum, j prescribed, every
 * non-conserved moment copied from the neighbour. */
void FN(orc_vel_extrap)(R f[19], R u, R v, R w)
{
    R m[19];
    FN(orc_moments)(f, m);
    m[3] = u; m[5] = v; m[7] = w;
    FN(invert_moments)(m, f);
}

/* East_Extrap (MRTtest/DynInlet3.cu:926-971) and North_Extrap (:834-879):
 * rho prescribed, momentum and non-conserved moments copied from the
 * neighbour. */
void FN(orc_east_extrap)(R f[19], R rho)
{
    R m[19];
    FN(orc_moments)(f, m);
    m[0] = rho;
    FN(invert_moments)(m, f);
}

/* DirichletWest_Regularized (CPU_3D/CPU_3D.cpp:974-1054), kept as coded:
 * velocity (0, u_lid, 0); feq11/feq13 use (v+w)*(u+w) and (-v+w)*(u+w);
 * the unknown f8 is set to feq6 + f6 - feq6. */
void FN(orc_lid_west_reg)(R f[19], R u_lid)
{
    R u = 0, v = u_lid, w = 0;
    R usqr = u * u + v * v + w * w;
    R rho = u + (f[0] + f[2] + f[4] + f[9] + f[11] + f[13] + f[14] + f[16] + f[18] +
                 (R)2 * (f[3] + f[6] + f[7] + f[12] + f[17]));
    R feq[19];
    FN(orc_feq_bgk)(rho, u, v, w, feq);
    feq[11] = (R)g_w[11] * (rho + (R)3 * (v + w) + (R)4.5 * (v + w) * (u + w) - (R)1.5 * usqr);
    feq[13] = (R)g_w[13] * (rho + (R)3 * (-v + w) + (R)4.5 * (-v + w) * (u + w) - (R)1.5 * usqr);
    f[1] = feq[1] + f[3] - feq[3];
    f[5] = feq[5] + f[7] - feq[7];
    f[8] = feq[6] + f[6] - feq[6];
    f[10] = feq[10] + f[17] - feq[17];
    f[15] = feq[15] + f[12] - feq[12];
    R pi[3][3] = {{0}};
    for (int i = 0; i < 19; ++i) {
        R d = f[i] - feq[i];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                pi[a][b] += (R)(g_c[i][a] * g_c[i][b]) * d;
    }
    /* f_i = feq_i + w_i/(2 cs^4) Q_i : Pi with Q_i = c_i c_i - I/3 */
    for (int i = 0; i < 19; ++i) {
        R s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                R qab = (R)(g_c[i][a] * g_c[i][b]) - (a == b ? (R)(1.0 / 3.0) : (R)0);
                s += qab * pi[a][b];
            }
        f[i] = feq[i] + (R)4.5 * (R)g_w[i] * s;
    }
}

/* Symmetry planes x = 0 (xsymmetry_bot, MRTtest/DynInlet3.cu:1020-1075) and
 * x = nx-1 (xsymmetry_top, :1077-1137): copy statements in the reference's
 * order, edges and corners of the plane as coded. */
void FN(orc_xsym_bot)(R f[19], int y, int z, int ny, int nz)
{
    if (y == 0 && z == 0) {
        f[2] = f[4]; f[13] = f[18]; f[11] = f[18]; f[16] = f[18]; f[6] = f[7]; f[9] = f[14]; f[12] = f[17];
    } else if (y == 0 && z == nz - 1) {
        f[4] = f[2]; f[11] = f[13]; f[18] = f[13]; f[16] = f[13]; f[6] = f[7]; f[14] = f[9]; f[17] = f[12];
    } else if (y == ny - 1 && z == 0) {
        f[4] = f[2]; f[11] = f[16]; f[18] = f[16]; f[13] = f[16]; f[7] = f[6]; f[9] = f[14]; f[12] = f[17];
    } else if (y == ny - 1 && z == nz - 1) {
        f[4] = f[2]; f[16] = f[11]; f[18] = f[11]; f[13] = f[11]; f[7] = f[6]; f[14] = f[9]; f[17] = f[12];
    } else if (y == 0) {
        f[2] = f[4]; f[11] = f[13]; f[16] = f[18]; f[8] = f[5];
    } else if (y == ny - 1) {
        f[4] = f[2]; f[13] = f[11]; f[18] = f[16]; f[5] = f[8];
    }
    f[1] = f[3]; f[5] = f[6]; f[8] = f[7]; f[10] = f[12]; f[15] = f[17];
}

void FN(orc_xsym_top)(R f[19], int y, int z, int ny, int nz)
{
    if (y == 0 && z == 0) {
        f[2] = f[4]; f[13] = f[18]; f[11] = f[18]; f[16] = f[18]; f[5] = f[8]; f[9] = f[14]; f[10] = f[15];
    } else if (y == 0 && z == nz - 1) {
        f[2] = f[4]; f[11] = f[13]; f[18] = f[13]; f[16] = f[13]; f[5] = f[8]; f[14] = f[9]; f[15] = f[10];
    } else if (y == ny - 1 && z == 0) {
        f[4] = f[2]; f[18] = f[16]; f[11] = f[16]; f[13] = f[16]; f[8] = f[5]; f[9] = f[14]; f[10] = f[15];
    } else if (y == ny - 1 && z == nz - 1) {
        f[4] = f[2]; f[13] = f[11]; f[16] = f[11]; f[18] = f[11]; f[8] = f[5]; f[14] = f[9]; f[15] = f[10];
    } else if (y == 0) {
        f[2] = f[4]; f[11] = f[13]; f[16] = f[18]; f[5] = f[8];
    } else if (y == ny - 1) {
        f[4] = f[2]; f[13] = f[11]; f[18] = f[16]; f[8] = f[5];
    }
    f[3] = f[1]; f[6] = f[5]; f[7] = f[8]; f[12] = f[10]; f[17] = f[15];
}

static inline int FN(is_solid)(uint8_t code)
{
    return code == ORC_SOLID || code == ORC_SOLID_FORCE;
}

/* One fused pull step fin -> fout: CPU_3D/CPU_3D.cpp:1317-1443 (pull with
 * clamped neighbours, node bounce-back, boundaries(), collide) merged with the
 * node dispatch of MRTtest/DynInlet3.cu:1612-1840 (Extrap inlet/outlet read
 * the unstreamed neighbour; momentum exchange on im 10). */
void FN(orc_step)(const orc_params* p, const uint8_t* mask, const R* u_in,
                  const R* fin, R* fout, double force[3])
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t N = (size_t)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
    double F0 = 0, F1 = 0, F2 = 0;
    orc_params p_noforce = *p; /* rows past DYNY1 */
    p_noforce.dpdy = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : F0, F1, F2)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double F[3] = {0, 0, 0};
                const size_t cell = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
                const uint8_t code = mask ? mask[cell] : ORC_FLUID;
                if (p->halfway && FN(is_solid)(code)) {
                    for (int i = 0; i < 19; ++i) fout[i * N + cell] = fin[i * N + cell];
                    continue;
                }
                R f[19];
                const int xyz[3] = {x, y, z};
                for (int i = 0; i < 19; ++i) {
                    int s[3];
                    int wall = 0;
                    for (int a = 0; a < 3; ++a) {
                        s[a] = xyz[a] - g_c[i][a];
                        /* the precursor's periodic rows 0 .. y_wrap
                           (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1131-1132) */
                        if (a == 1 && p->y_wrap >= 0) {
                            if (y == p->y_wrap && g_c[i][1] < 0) { s[1] = 0; continue; }
                            if (y == 0 && g_c[i][1] > 0) { s[1] = p->y_wrap; continue; }
                        }
                        if (s[a] < 0 || s[a] >= n[a]) {
                            if (p->bc[a] == ORC_PERIODIC) s[a] = (s[a] + n[a]) % n[a];
                            else if (p->bc[a] == ORC_CLAMP) s[a] = s[a] < 0 ? 0 : n[a] - 1;
                            else wall = 1;
                        }
                    }
                    int bounce = wall;
                    int src_force = 0;
                    size_t sc = 0;
                    if (!wall) {
                        sc = (size_t)s[0] + (size_t)nx * (s[1] + (size_t)ny * s[2]);
                        if (p->halfway && mask && FN(is_solid)(mask[sc])) {
                            bounce = 1;
                            src_force = mask[sc] == ORC_SOLID_FORCE;
                        }
                    }
                    if (bounce) {
                        const int o = g_opp[i];
                        f[i] = fin[o * N + cell];
                        if (src_force)
                            for (int a = 0; a < 3; ++a) F[a] += 2.0 * g_c[o][a] * (double)f[i];
                    } else {
                        f[i] = fin[i * N + sc];
                    }
                }
                if (!p->halfway && FN(is_solid)(code)) {
                    if (code == ORC_SOLID_FORCE) {
                        double sx = 0, sy = 0, sz = 0;
                        for (int i = 0; i < 19; ++i) {
                            sx += 2.0 * g_c[i][0] * (double)f[i];
                            sy += 2.0 * g_c[i][1] * (double)f[i];
                            sz += 2.0 * g_c[i][2] * (double)f[i];
                        }
                        F[0] += sx; F[1] += sy; F[2] += sz;
                    }
                    fout[cell] = f[0];
                    for (int i = 1; i < 19; ++i) fout[g_opp[i] * N + cell] = f[i];
                    F0 += F[0]; F1 += F[1]; F2 += F[2];
                    continue;
                }
                if (code == ORC_INLET_WEST) {
                    const int xn = x + 1 < nx ? x + 1 : nx - 1;
                    const size_t nc = (size_t)xn + (size_t)nx * (y + (size_t)ny * z);
                    for (int i = 0; i < 19; ++i) f[i] = fin[i * N + nc];
                    FN(orc_west_extrap)(f, u_in ? u_in[(size_t)z * ny + y] : (R)0);
                } else if (code == ORC_OUTLET_EAST) {
                    const int xn = x > 0 ? x - 1 : 0;
                    const size_t nc = (size_t)xn + (size_t)nx * (y + (size_t)ny * z);
                    for (int i = 0; i < 19; ++i) f[i] = fin[i * N + nc];
                    FN(orc_east_extrap)(f, (R)1);
                } else if (code == ORC_LID_WEST_REG) {
                    FN(orc_lid_west_reg)(f, (R)p->u_lid);
                } else if (code == ORC_SYM_XMIN) {
                    FN(orc_xsym_bot)(f, y, z, ny, nz);
                } else if (code == ORC_SYM_XMAX) {
                    FN(orc_xsym_top)(f, y, z, ny, nz);
                } else if (code >= ORC_OUTLET_NORTH && code <= ORC_VEL_FROM_BELOW) {
                    /* update_inn's y-streamwise conditions (MRTtest/DynInlet3.cu:1711-1724;
                       bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1220-1239) on the
                       unstreamed rows y-1 / y+1 (f_mem wraps y, :177-185) */
                    const size_t bc_ = (size_t)x + (size_t)nx * ((y > 0 ? y - 1 : ny - 1) + (size_t)ny * z);
                    const size_t ac_ = (size_t)x + (size_t)nx * ((y + 1 < ny ? y + 1 : 0) + (size_t)ny * z);
                    if (code == ORC_OUTLET_NORTH) {
                        for (int i = 0; i < 19; ++i) f[i] = fin[i * N + bc_];
                        FN(orc_east_extrap)(f, (R)1);
                    } else if (code == ORC_INLET_SOUTH) {
                        for (int i = 0; i < 19; ++i) f[i] = fin[i * N + ac_];
                        FN(orc_vel_extrap)(f, (R)0, p->v_in ? (R)p->v_in[(size_t)z * nx + x] : (R)0, (R)0);
                    } else if (code == ORC_RECYCLE) {
                        R rho = 0;
                        for (int i = 0; i < 19; ++i) rho += fin[i * N + ac_];
                        for (int i = 0; i < 19; ++i) f[i] = fin[i * N + bc_];
                        FN(orc_east_extrap)(f, rho);
                    } else {
                        R j[3] = {0, 0, 0};
                        for (int i = 0; i < 19; ++i)
                            for (int a = 0; a < 3; ++a) j[a] += (R)g_c[i][a] * fin[i * N + bc_];
                        for (int i = 0; i < 19; ++i) f[i] = fin[i * N + ac_];
                        FN(orc_vel_extrap)(f, j[0], j[1], j[2]);
                    }
                } else if (code >= ORC_GHOST_W && code <= ORC_GHOST_DYN) {
                    /* the ghost copies follow the collide and overwrite it with
                       an unstreamed node (MRTtest/DynInlet3.cu:1747-1771) */
                    int sx = x, sy = y;
                    if (code == ORC_GHOST_W) sx = nx - 2;
                    else if (code == ORC_GHOST_E) sx = 1;
                    else sy = p->ghost_y[code - ORC_GHOST_S];
                    const size_t sc_ = (size_t)sx + (size_t)nx * (sy + (size_t)ny * z);
                    for (int i = 0; i < 19; ++i) fout[i * N + cell] = fin[i * N + sc_];
                    F0 += F[0]; F1 += F[1]; F2 += F[2];
                    continue;
                }
                FN(orc_collide)((p->dyn_y1 >= 0 && y > p->dyn_y1) ? &p_noforce : p, f);
                for (int i = 0; i < 19; ++i) fout[i * N + cell] = f[i];
                F0 += F[0]; F1 += F[1]; F2 += F[2];
            }
    if (force) {
        force[0] = F0;
        force[1] = F1;
        force[2] = F2;
    }
}

/* Time loop with A/B ping-pong (CPU_3D/CPU_3D.cpp:1317-1443, two sweeps per
 * iteration).  The final state is left in fa; forces[3*t+a] holds step t. */
void FN(orc_run)(const orc_params* p, const uint8_t* mask, const R* u_in,
                 R* fa, R* fb, int nsteps, double* forces)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    R* a = fa;
    R* b = fb;
    for (int t = 0; t < nsteps; ++t) {
        FN(orc_step)(p, mask, u_in, a, b, forces ? forces + 3 * (size_t)t : 0);
        R* tmp = a; a = b; b = tmp;
    }
    if (a != fa) memcpy(fa, a, sizeof(R) * 19 * N);
}

/* Node macroscopic fields as the reference writes them (WriteResults,
 * MRTtest/DynInlet3.cu:2601-2694): rho = sum f, (u,v,w) = momentum sums. */
void FN(orc_macro)(const orc_params* p, const R* f, R* rho, R* u, R* v, R* w)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    for (size_t c = 0; c < N; ++c) {
        R r = 0, a = 0, b = 0, d = 0;
        for (int i = 0; i < 19; ++i) {
            R fi = f[i * N + c];
            r += fi;
            a += (R)g_c[i][0] * fi;
            b += (R)g_c[i][1] * fi;
            d += (R)g_c[i][2] * fi;
        }
        if (rho) rho[c] = r;
        if (u) u[c] = a;
        if (v) v[c] = b;
        if (w) w[c] = d;
    }
}

/* Running velocity statistics after one step (vel_av / vel_fluc,
 * MRTtest/DynInlet3.cu:1139-1148, :1159-1172, applied after mrt_collide in
 * update_inn, :1773-1799; LR forms :1150-1157, :1174-1184 with weight w =
 * LRFACTOR).  f holds the step's output: at a non-solid node these are its
 * post-collision populations.  stat: six planes of nx*ny*nz floats (mean u,
 * v, w; mean squared deviations); mode 1 updates the means, mode 2 the
 * deviations against the frozen means; n = t - START of the phase. */
void FN(orc_stats_update)(const orc_params* p, const uint8_t* mask, const R* f, float* stat, int mode,
                          float n, float w)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    for (size_t c = 0; c < N; ++c) {
        if (mask && FN(is_solid)(mask[c])) continue;
        const R* g = f + c;
#define F_(i) g[(size_t)(i) * N]
        const float u[3] = {
            (float)(F_(1) - F_(3) + F_(5) - F_(6) - F_(7) + F_(8) + F_(10) - F_(12) + F_(15) - F_(17)),
            (float)(F_(2) - F_(4) + F_(5) + F_(6) - F_(7) - F_(8) + F_(11) - F_(13) + F_(16) - F_(18)),
            (float)(F_(9) + F_(10) + F_(11) + F_(12) + F_(13) - F_(14) - F_(15) - F_(16) - F_(17) - F_(18))};
#undef F_
        for (int a = 0; a < 3; ++a) {
            float* mean = stat + (size_t)a * N + c;
            if (mode == 1) {
                *mean = (*mean * n + u[a] * w) / (n + w);
            } else {
                float* dev = stat + (size_t)(3 + a) * N + c;
                const float d = (u[a] - *mean) * (u[a] - *mean);
                *dev = (*dev * n + d * w) / (n + w);
            }
        }
    }
}
