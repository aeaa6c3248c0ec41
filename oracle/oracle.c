/*
 * oracle.c -- CPU restatement of the reference's D3Q19 path (float and double).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never does.
 *
 * Pinning (what each function is checked against) is listed in the
 * docstrings of tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* D3Q19 velocities in the reference's order: MRTtest/DynInlet3.cu:603-605
 * (moment sums) and :1625-1671 (pull offsets). */
static const int g_c[19][3] = {
    {0, 0, 0},   {1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {1, 1, 0},  {-1, 1, 0},
    {-1, -1, 0}, {1, -1, 0}, {0, 0, 1},  {1, 0, 1},  {0, 1, 1},  {-1, 0, 1}, {0, -1, 1},
    {0, 0, -1},  {1, 0, -1}, {0, 1, -1}, {-1, 0, -1}, {0, -1, -1}};

/* Weights of the BGK equilibrium (CPU_3D/CPU_3D.cpp:135-153). */
static const double g_w[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 18,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 18,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

/* Opposite directions: the bounce-back permutation of MRTtest/DynInlet3.cu:1688-1705. */
static const int g_opp[19] = {0, 3, 4, 1, 2, 7, 8, 5, 6, 14, 17, 18, 15, 16, 9, 12, 13, 10, 11};

static int g_M[19][19];
static int g_norm[19];
/* sparse rows and columns of M, and M^-1 = M^T diag(1/norm) in both precisions */
static int g_rnz[19][19], g_rn[19];  /* nonzero i of row k */
static int g_cnz[19][19], g_cn[19];  /* nonzero k of column i */
static float g_Mf[19][19], g_Mif[19][19];
static double g_Md[19][19], g_Mid[19][19];

/* Moment matrix rows from the d'Humieres 2002 polynomials in c (rho, e, eps,
 * jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz). */
static void build_matrix(void)
{
    for (int i = 0; i < 19; ++i) {
        int cx = g_c[i][0], cy = g_c[i][1], cz = g_c[i][2];
        int c2 = cx * cx + cy * cy + cz * cz;
        int row[19];
        row[0] = 1;
        row[1] = 19 * c2 - 30;
        row[2] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
        row[3] = cx;
        row[4] = (5 * c2 - 9) * cx;
        row[5] = cy;
        row[6] = (5 * c2 - 9) * cy;
        row[7] = cz;
        row[8] = (5 * c2 - 9) * cz;
        row[9] = 3 * cx * cx - c2;
        row[10] = (3 * c2 - 5) * (3 * cx * cx - c2);
        row[11] = cy * cy - cz * cz;
        row[12] = (3 * c2 - 5) * (cy * cy - cz * cz);
        row[13] = cx * cy;
        row[14] = cy * cz;
        row[15] = cx * cz;
        row[16] = cx * (cy * cy - cz * cz);
        row[17] = cy * (cz * cz - cx * cx);
        row[18] = cz * (cx * cx - cy * cy);
        for (int k = 0; k < 19; ++k) g_M[k][i] = row[k];
    }
    for (int k = 0; k < 19; ++k) {
        int s = 0;
        for (int i = 0; i < 19; ++i) s += g_M[k][i] * g_M[k][i];
        g_norm[k] = s;
    }
    for (int k = 0; k < 19; ++k) {
        g_rn[k] = 0;
        for (int i = 0; i < 19; ++i)
            if (g_M[k][i]) g_rnz[k][g_rn[k]++] = i;
    }
    for (int i = 0; i < 19; ++i) {
        g_cn[i] = 0;
        for (int k = 0; k < 19; ++k)
            if (g_M[k][i]) g_cnz[i][g_cn[i]++] = k;
    }
    for (int k = 0; k < 19; ++k)
        for (int i = 0; i < 19; ++i) {
            g_Mf[k][i] = (float)g_M[k][i];
            g_Md[k][i] = (double)g_M[k][i];
            g_Mif[i][k] = (float)g_M[k][i] / (float)g_norm[k];
            g_Mid[i][k] = (double)g_M[k][i] / (double)g_norm[k];
        }
}

__attribute__((constructor)) static void orc_init_tables(void) { build_matrix(); }

void orc_tables(int c[19][3], double w[19], int opp[19])
{
    for (int i = 0; i < 19; ++i) {
        for (int a = 0; a < 3; ++a) c[i][a] = g_c[i][a];
        w[i] = g_w[i];
        opp[i] = g_opp[i];
    }
}

void orc_moment_matrix(double M[19][19])
{
    for (int k = 0; k < 19; ++k)
        for (int i = 0; i < 19; ++i) M[k][i] = g_M[k][i];
}

void orc_default_params(orc_params* p)
{
    memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 1;
    p->omega = 1.0;
    /* d'Humieres 2002 rates (MRTtest/DynInlet3.cu:774-783) */
    p->s_dhum[1] = 1.19;
    p->s_dhum[2] = 1.4;
    p->s_dhum[4] = p->s_dhum[6] = p->s_dhum[8] = 1.2;
    p->s_dhum[10] = p->s_dhum[12] = 1.4;
    p->s_dhum[16] = p->s_dhum[17] = p->s_dhum[18] = 1.98;
    p->dyn_y1 = -1;
    p->y_wrap = -1;
    p->ghost_y[0] = p->ghost_y[1] = p->ghost_y[2] = 0;
}

#define FN(name) name##_f32
#define R float
#define MT g_Mf
#define MIT g_Mif
#include "oracle_impl.h"
#undef R
#undef FN
#undef MT
#undef MIT

#define FN(name) name##_f64
#define R double
#define MT g_Md
#define MIT g_Mid
#include "oracle_impl.h"
#undef R
#undef FN
#undef MT
#undef MIT

/* splitmix64 finaliser over (seed, counter); the top 32 bits are returned. */
uint32_t orc_hash_u32(uint64_t seed, uint64_t counter)
{
    uint64_t x = seed ^ (counter * 0x9E3779B97F4A7C15ULL);
    x += 0x632BE59BD9B4E019ULL;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return (uint32_t)(x >> 32);
}

/* Inlet profile u_in[z*ny + y] = u0 + amp*(2r-1), r = (h >> 8) 2^-24 in [0,1).
 * The reference's inlet is uniform (PoisProf3D returns UMAX,
 * MRTtest/DynInlet3.cu:102-110); the seeded perturbation has no counterpart. */
void orc_inlet_profile(int ny, int nz, float u0, float amp, uint64_t seed, float* u_in)
{
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            uint64_t cell = (uint64_t)z * (uint64_t)ny + (uint64_t)y;
            uint32_t h = orc_hash_u32(seed, cell);
            float r = (float)(h >> 8) * (1.0f / 16777216.0f);
            float t = 2.0f * r - 1.0f;
            float d = amp * t;
            u_in[cell] = u0 + d;
        }
}
