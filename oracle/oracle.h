/*
 * oracle.h -- CPU restatement of the reference's D3Q19 stream-collide path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (liblbm.so) never links or calls it.
 *
 * Every function restates an algorithm of the reference; the file:line it
 * follows is cited at the definition in oracle_impl.h.  Two precisions are
 * built from the same source: *_f32 (float arithmetic, like the reference)
 * and *_f64 (double arithmetic, the FP64 yardstick).
 *
 * Layout: populations are dense SoA, f[q][z][y][x] (x fastest), 19 planes of
 * nx*ny*nz values each.  Masks are uint8 node codes in [z][y][x] order.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* collision operators */
enum {
    ORC_BGK = 0,        /* CPU_3D/CPU_3D.cpp:116-160 (also MRTtest/DynInlet3.cu:802-829) */
    ORC_MRT_CPU3D = 1,  /* CPU_3D/CPU_3D.cpp:164-279: m13,m14 relaxed at rate 1 */
    ORC_MRT_LES = 2,    /* MRTtest/DynInlet3.cu:599-832: m9,m11,m13,m14,m15 at omega */
    ORC_MRT_DHUM = 3    /* MRTtest/DynInlet3.cu:774-783 d'Humieres 2002 rates */
};

/* per-axis treatment of pulls that leave the domain */
enum {
    ORC_PERIODIC = 0,   /* wrap (f_mem y-wrap, MRTtest/DynInlet3.cu:177-185) */
    ORC_CLAMP = 1,      /* clamp to edge (dmin/dmax, CPU_3D/CPU_3D.cpp:104-113) */
    ORC_WALL = 2        /* halfway (link) bounce-back at the face */
};

/* node codes (uint8) and the reference `im` code each one stands for */
enum {
    ORC_FLUID = 0,          /* im 0 */
    ORC_SOLID = 1,          /* im 1: bounce-back wall */
    ORC_SOLID_FORCE = 2,    /* im 10: bounce-back wall, momentum exchange summed */
    ORC_INLET_WEST = 3,     /* im 400: West_Extrap velocity inlet */
    ORC_OUTLET_EAST = 4,    /* im 300: East_Extrap rho=1 outlet */
    ORC_LID_WEST_REG = 5,   /* CPU_3D im 53: DirichletWest_Regularized */
    ORC_SYM_XMIN = 6,       /* im 26: xsymmetry_bot */
    ORC_SYM_XMAX = 7,       /* im 25: xsymmetry_top */
    ORC_OUTLET_NORTH = 8,   /* im 100: North_Extrap rho = 1 from the unstreamed y-1 node */
    ORC_INLET_SOUTH = 9,    /* im 200: South_Extrap (0, v_in, 0) from the unstreamed y+1 node */
    ORC_RECYCLE = 10,       /* im 80: rho of y+1, moments of y-1 (North_Extrap(f(y-1), rho(y+1))) */
    ORC_VEL_FROM_BELOW = 11, /* im 70: velocity of y-1, moments of y+1 (South_Extrap(f(y+1), j(y-1))) */
    ORC_GHOST_W = 12,       /* im 50..54: copies of an unstreamed node */
    ORC_GHOST_E = 13,
    ORC_GHOST_S = 14,
    ORC_GHOST_N = 15,
    ORC_GHOST_DYN = 16
};

typedef struct {
    int nx, ny, nz;
    int bc[3];          /* ORC_PERIODIC / ORC_CLAMP / ORC_WALL for x, y, z */
    int collision;      /* ORC_BGK ... */
    int halfway;        /* 0: node (full-way) bounce-back of the reference;
                           1: link-wise halfway bounce-back */
    double omega;       /* base rate 1/tau0 (s_nu) */
    double cs;          /* Smagorinsky constant as the reference's CS (= Cs^2); 0 = off */
    double dpdy;        /* AddForce pressure gradient (LES-family operators only) */
    double u_lid;       /* UMAX used by DirichletWest_Regularized */
    double s_dhum[19];  /* rates for ORC_MRT_DHUM (entries 9,11,13,14,15 ignored: omega) */
    int dyn_y1;         /* rows y > dyn_y1 get no body force, "if(y>DYNY1) dpdy = 0.f"
                           (MRTtest/DynInlet3.cu:1744); -1: every row gets it */
    int y_wrap;         /* rows 0 .. y_wrap periodic in y among themselves
                           (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1131-1132); -1: none */
    int ghost_y[3];     /* source rows of the im 52 / 53 / 54 ghost copies */
    const float* v_in;  /* south inlet profile [z*nx + x] (NULL: 0) */
} orc_params;

void orc_default_params(orc_params* p);

/* D3Q19 tables and the moment matrix built from its defining polynomials */
void orc_tables(int c[19][3], double w[19], int opp[19]);
void orc_moment_matrix(double M[19][19]);

/* ---- float restatement ---- */
void orc_moments_f32(const float f[19], float m[19]);
void orc_collide_f32(const orc_params* p, float f[19]);
void orc_xsym_bot_f32(float f[19], int y, int z, int ny, int nz);
void orc_xsym_top_f32(float f[19], int y, int z, int ny, int nz);
void orc_les_omega_f32(const orc_params* p, const float f[19], float* omega_out, float* q_out);
void orc_west_extrap_f32(float f[19], float u);
void orc_vel_extrap_f32(float f[19], float u, float v, float w);
void orc_east_extrap_f32(float f[19], float rho);
void orc_lid_west_reg_f32(float f[19], float u_lid);
void orc_feq_bgk_f32(float rho, float u, float v, float w, float f[19]);
void orc_feq_mrt_f32(float rho, float u, float v, float w, float f[19]);
void orc_step_f32(const orc_params* p, const uint8_t* mask, const float* u_in,
                  const float* fin, float* fout, double force[3]);
void orc_run_f32(const orc_params* p, const uint8_t* mask, const float* u_in,
                 float* fa, float* fb, int nsteps, double* forces);
void orc_macro_f32(const orc_params* p, const float* f, float* rho, float* u, float* v, float* w);
void orc_stats_update_f32(const orc_params* p, const uint8_t* mask, const float* f, float* stat, int mode,
                          float n, float w);

/* ---- double restatement ---- */
void orc_moments_f64(const double f[19], double m[19]);
void orc_collide_f64(const orc_params* p, double f[19]);
void orc_xsym_bot_f64(double f[19], int y, int z, int ny, int nz);
void orc_xsym_top_f64(double f[19], int y, int z, int ny, int nz);
void orc_les_omega_f64(const orc_params* p, const double f[19], double* omega_out, double* q_out);
void orc_west_extrap_f64(double f[19], double u);
void orc_vel_extrap_f64(double f[19], double u, double v, double w);
void orc_east_extrap_f64(double f[19], double rho);
void orc_lid_west_reg_f64(double f[19], double u_lid);
void orc_feq_bgk_f64(double rho, double u, double v, double w, double f[19]);
void orc_feq_mrt_f64(double rho, double u, double v, double w, double f[19]);
void orc_step_f64(const orc_params* p, const uint8_t* mask, const double* u_in,
                  const double* fin, double* fout, double force[3]);
void orc_run_f64(const orc_params* p, const uint8_t* mask, const double* u_in,
                 double* fa, double* fb, int nsteps, double* forces);
void orc_macro_f64(const orc_params* p, const double* f, double* rho, double* u, double* v, double* w);
void orc_stats_update_f64(const orc_params* p, const uint8_t* mask, const double* f, float* stat, int mode,
                          float n, float w);

/* seeded inlet perturbation: counter-based hash of (seed, cell), no state */
uint32_t orc_hash_u32(uint64_t seed, uint64_t counter);
void orc_inlet_profile(int ny, int nz, float u0, float amp, uint64_t seed, float* u_in);

/* ---- 2:1 local refinement (MRTtest/DynInlet3.cu:2202-2560, :3180-3245) ----
 * Fine node (i,j,k) sits at coarse coordinates (a_x + 0.25 + 0.5 i, ...); the
 * coarse footprint is nodes a .. a + n_f/2 per axis (n_f/2 + 1 of them).
 * Physical moments (PhysicalMoments, :396-407): rho, jx, jy, jz, 3pxx-row,
 * pww-row, pxy, pyz, pxz; stored as 9 planes over the footprint. */
typedef struct {
    int a[3];           /* footprint origin (coarse integer coordinates) */
    int nf[3];          /* fine nodes per axis (even) */
    double omega_c, omega_f;
    double cs;          /* reference CS, used by the LES-aware rescale */
    int temporal;       /* 1: frame at t+1/2 for the second sub-step */
    int eq_kind;        /* higher-moment equilibrium of the rebuilt populations:
                           ORC_EQ_BGK_MEQ, ORC_EQ_BGK_PROJ or ORC_EQ_MRT_MEQ */
} orc_refine;

enum {
    ORC_EQ_BGK_MEQ = 0,  /* bgk_meq (:347-365): the LES-family collide's own equilibrium */
    ORC_EQ_BGK_PROJ = 1, /* the BGK equilibrium's projection (3pixx = -3pxx/2): BGK's own */
    ORC_EQ_MRT_MEQ = 2   /* mrt_meq (:329-345), as InvertPhysicalMoments codes it (:437-481) */
};

void orc_phys_moments_f32(const float f[19], float mom[9]);
void orc_strain_rate_f32(const float mom[9], float S[6]);
void orc_invert_phys_les_cf_f32(float f[19], const float mom[9], float omega_c, float cs, int eq_kind);
void orc_invert_phys_les_fc_f32(float f[19], const float mom[9], float omega_f, float cs, int eq_kind);
/* order-2 correction (:2312-2357): adds to mom[1..3]; fx etc. in {0.25, 0.75}
 * for CF, or cf_weights=0 (the FC form, weights 1) */
void orc_order2_correction_f32(const float corner_mom[8][9], float mom[9], float xf, float yf, float zf,
                               int fc_form);
/* extract the footprint moments of a coarse lattice (dense) */
void orc_extract_f32(const orc_params* pc, const orc_refine* r, const float* fc, float* M);
/* coarse -> fine frame (2 fine nodes thick on all six faces) */
void orc_interp_cf_f32(const orc_refine* r, const float* Mold, const float* Mnew, float alpha,
                       const uint8_t* fmask, float* ff);
/* fine -> coarse nodes a+2 .. a+n_f/2-2 */
void orc_interp_fc_f32(const orc_params* pc, const orc_refine* r, const float* ff, const uint8_t* cmask,
                       float* fc);
/* one coarse step with two fine sub-steps; Mold holds the coarse moments at
 * the step's start and is replaced by those at its end; fine_force (may be
 * NULL) receives the fine level's momentum exchange of the first sub-step */
void orc_refine_step_f32(const orc_params* pc, const orc_params* pf, const orc_refine* r,
                         const uint8_t* cmask, const uint8_t* fmask, const float* u_in,
                         float* ca, float* cb, float* fa, float* fb, float* Mold, float* Mnew,
                         double* fine_force);
/* the same interface in double (the FP64 yardstick of the refined path) */
void orc_phys_moments_f64(const double f[19], double mom[9]);
void orc_strain_rate_f64(const double mom[9], double S[6]);
void orc_invert_phys_les_cf_f64(double f[19], const double mom[9], double omega_c, double cs, int eq_kind);
void orc_invert_phys_les_fc_f64(double f[19], const double mom[9], double omega_f, double cs, int eq_kind);
void orc_order2_correction_f64(const double corner_mom[8][9], double mom[9], double xf, double yf, double zf,
                               int fc_form);
void orc_extract_f64(const orc_params* pc, const orc_refine* r, const double* fc, double* M);
void orc_interp_cf_f64(const orc_refine* r, const double* Mold, const double* Mnew, double alpha,
                       const uint8_t* fmask, double* ff);
void orc_interp_fc_f64(const orc_params* pc, const orc_refine* r, const double* ff, const uint8_t* cmask,
                       double* fc);
void orc_refine_step_f64(const orc_params* pc, const orc_params* pf, const orc_refine* r,
                         const uint8_t* cmask, const uint8_t* fmask, const double* u_in,
                         double* ca, double* cb, double* fa, double* fb, double* Mold, double* Mnew,
                         double* fine_force);

#ifdef __cplusplus
}
#endif
#endif
