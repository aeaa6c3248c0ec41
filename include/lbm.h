/*
 * lbm.h -- C ABI of liblbm, the MI355X (gfx950) D3Q19 stream-collide and
 * local-refinement library.
 *
 * The reference has no library boundary: its hot path is a set of
 * __global__ kernels launched by one main() per program, configured by
 * compile-time #defines and an ImageFcn() cell classifier.  Each entry point
 * below names the reference interface it replaces (file:line, relative to
 * the reference tree).
 *
 * Conventions
 *  - Every int-returning call returns 0 on success and a negative LBM_E*
 *    code on failure; lbm_last_error() then holds a message.  No C++
 *    exception crosses this boundary.
 *  - The library owns all device memory; the caller owns host buffers.
 *    Sizes come from lbm_query_sizes().
 *  - A context is single-threaded: the caller serialises calls on it.  Each
 *    context owns one HIP stream; lbm_step() is asynchronous and every
 *    lbm_get_*() call synchronises that stream.
 *  - Host arrays are dense, x fastest: scalar fields are [z][y][x],
 *    populations [q][z][y][x] with the reference's D3Q19 ordering
 *    (MRTtest/DynInlet3.cu:603-605, :1625-1671).
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define LBM_OK 0
#define LBM_EINVAL (-22)    /* bad argument or configuration */
#define LBM_ENOMEM (-12)    /* device or host allocation failed */
#define LBM_EDEVICE (-5)    /* HIP runtime error, or no usable gfx950 device */
#define LBM_ESTATE (-71)    /* call not valid in the current state */
#define LBM_EIO (-2)        /* file could not be read or written */

/* collision model: MODEL / SmagLES #defines (MRTtest/DynInlet3.cu:52-56) */
#define LBM_MODEL_BGK 0           /* CPU_3D/CPU_3D.cpp:116-160; DynInlet3.cu:802-829 */
#define LBM_MODEL_MRT_CPU3D 1     /* CPU_3D/CPU_3D.cpp:164-279 (m13, m14 at rate 1) */
#define LBM_MODEL_MRT_LES 2       /* MRTtest/DynInlet3.cu:599-832 */
#define LBM_MODEL_MRT_DHUMIERES 3 /* d'Humieres rates, MRTtest/DynInlet3.cu:774-783 */

/* what a pull that leaves the domain along an axis does */
#define LBM_BC_PERIODIC 0 /* wrap (f_mem y-wrap, MRTtest/DynInlet3.cu:177-185) */
#define LBM_BC_CLAMP 1    /* clamp to edge (dmin/dmax, CPU_3D/CPU_3D.cpp:104-113) */
#define LBM_BC_WALL 2     /* halfway (link) bounce-back wall at the face */

/* population storage in HBM */
#define LBM_STORE_F32 0  /* reference layout precision */
#define LBM_STORE_F16 1  /* f_i - f_i^eq(1, (umax,0,0)) in IEEE half, FP32 collide (= f_i - w_i when umax = 0,
                            and with the y-streamwise LBM_GEOM_PRECURSOR) */
#define LBM_STORE_BF16 2 /* the same offset in bfloat16, FP32 collide */
#define LBM_STORE_F64 3  /* FP64 storage and arithmetic (the FP64 yardstick) */

/* geometry generators standing in for ImageFcn() */
#define LBM_GEOM_NONE 0    /* all fluid */
#define LBM_GEOM_CHANNEL 1 /* west velocity inlet (im 400), east rho=1 outlet (im 300),
                              square obstacle |x-X0|<R && |y-Y0|<R over all z
                              (LaminarChannelCylinder1/LRMultiGPU6.cu:90-129) */
#define LBM_GEOM_CPU3D 2   /* CPU_3D/CPU_3D.cpp:39-75: circular cylinder (im 10),
                              regularised moving-lid x=0 plane (im 53), walls (im 1) */
#define LBM_GEOM_PRECURSOR 3 /* y-streamwise channel with a periodic inflow precursor
                              (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:75-94): trip
                              obstacle 2 (im 1) while t < trip_until, obstacle 1
                              (im 10), node walls at x = 0 and x = nx-1 (im 1), the
                              recycling plane y = dyn_y1+1 (im 80) and the north
                              rho = 1 outlet y = ny-1 (im 100) */

/* node codes of lbm_set_geometry()/lbm_get_mask(), with the reference im code */
#define LBM_NODE_FLUID 0        /* im 0 */
#define LBM_NODE_SOLID 1        /* im 1: bounce-back wall */
#define LBM_NODE_SOLID_FORCE 2  /* im 10: bounce-back wall, momentum exchange summed */
#define LBM_NODE_INLET_WEST 3   /* im 400: West_Extrap (MRTtest/DynInlet3.cu:973-1017) */
#define LBM_NODE_OUTLET_EAST 4  /* im 300: East_Extrap (MRTtest/DynInlet3.cu:926-971) */
#define LBM_NODE_LID_WEST_REG 5 /* im 53: DirichletWest_Regularized (CPU_3D/CPU_3D.cpp:974) */
#define LBM_NODE_SYM_XMIN 6     /* im 26: xsymmetry_bot, x = 0 symmetry plane (MRTtest/DynInlet3.cu:1020-1075) */
#define LBM_NODE_SYM_XMAX 7     /* im 25: xsymmetry_top, x = nx-1 symmetry plane (MRTtest/DynInlet3.cu:1077-1137) */
#define LBM_NODE_OUTLET_NORTH 8 /* im 100: North_Extrap rho = 1 from the unstreamed y-1 node
                                   (MRTtest/DynInlet3.cu:834-879, dispatch :1711-1716) */
#define LBM_NODE_INLET_SOUTH 9  /* im 200: South_Extrap, velocity (0, v_in, 0) and the other
                                   moments from the unstreamed y+1 node (:881-924, :1717-1724) */
#define LBM_NODE_RECYCLE 10     /* im 80: rho from the unstreamed y+1 node, every other moment
                                   from y-1 (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1231-1239) */
#define LBM_NODE_VEL_FROM_BELOW 11 /* im 70: velocity from the unstreamed y-1 node, every other
                                   moment from y+1 (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1220-1230) */
#define LBM_NODE_GHOST_W 12     /* im 50: copy of the unstreamed node (nx-2, y, z)
                                   (MRTtest/DynInlet3.cu:1747-1751) */
#define LBM_NODE_GHOST_E 13     /* im 51: copy of (1, y, z) (:1752-1756) */
#define LBM_NODE_GHOST_S 14     /* im 52: copy of (x, DYNY1-1, z) (:1757-1761) */
#define LBM_NODE_GHOST_N 15     /* im 53 of MRTtest (not CPU_3D's 53): copy of (x, DYNY2, z) (:1762-1766) */
#define LBM_NODE_GHOST_DYN 16   /* im 54: copy of (x, 1, z) (:1767-1771) */
#define LBM_NODE_MAX 16

/* initial state of lbm_init_equilibrium() */
#define LBM_INIT_BGK 0 /* explicit BGK feq (CPU_3D/CPU_3D.cpp:1218-1254) */
#define LBM_INIT_MRT 1 /* mrt_meq + InvertMoments (MRTtest/DynInlet3.cu:1186-1216) */

/*
 * Configuration.  Field names follow the reference's #defines and .inputs
 * echo keys (MRTtest/DynInlet3.cu:10-60, WriteInputs :2805-2833).
 */
typedef struct lbm_config {
    int nx, ny, nz;        /* "Base domain size" XDIM x YDIM x ZDIM */
    int bc[3];             /* LBM_BC_* per axis x, y, z */
    int model;             /* LBM_MODEL_* ("MODEL") */
    int halfway;           /* 0: node (full-way) bounce-back as the reference;
                              1: link-wise halfway bounce-back */
    int storage;           /* LBM_STORE_* */
    float omega;           /* "omega": 1/tau0 */
    float cs;              /* "CS" (reference meaning: Cs^2); 0 disables LES */
    float dpdy;            /* "DPDY": AddForce body force (MRTtest/DynInlet3.cu:229-253) */
    float umax;            /* "UMAX": inlet / lid velocity */
    float inlet_amp;       /* amplitude of the seeded inlet perturbation (0 = uniform) */
    uint64_t seed;         /* seed of that perturbation */
    float s_dhum[19];      /* d'Humieres rates for LBM_MODEL_MRT_DHUMIERES (indices
                              9,11,13,14,15 are replaced by the LES rate) */
    int geometry;          /* LBM_GEOM_* used when lbm_set_geometry(mask=NULL) */
    float obst_x, obst_y, obst_z, obst_r; /* "Obst1 location" / "Obst1 radius" */
    int startf;            /* "STARTF": first step whose force is recorded */
    int force_steps;       /* capacity of the per-step force history (0 = off) */
    /* local refinement ("REFINEMENT", "LR domain size", "LR location") */
    int refinement;        /* 0 off, 1 one 2:1 fine patch */
    int lr_nx, lr_ny, lr_nz;      /* fine nodes of the patch */
    float lr_x0, lr_y0, lr_z0;    /* position of fine node 0 in coarse units */
    int temporal_interp;   /* 1: the frame is refilled before the second fine
                              sub-step with the coarse moments interpolated to
                              t+1/2; 0: filled once per coarse step, as the
                              reference (MRTtest/DynInlet3.cu:3180-3245) */
    int interface_eq;      /* equilibrium of the higher moments that InterpCF/
                              InterpFC rebuild: 0 the collide's own (default);
                              1 mrt_meq as InvertPhysicalMoments codes it
                              (MRTtest/DynInlet3.cu:437-481) */
    /* dynamic-inlet precursor (MRTtest/DynInlet3.cu:1744, :2561-2600,
       :3147-3170): the body force acts only on nodes with y <= dyn_y1
       ("if(y>DYNY1) dpdy = 0.f"); after every coarse step t with
       t % ctrl_every == 0 and t > ctrl_start, the mean v over the plane
       y = dyn_y1 (AverageV: sum over all x and z, divided by (nx-2) nz)
       drives the P-controller dpdy += (Av_V - UMAX)/UMAX * kp * |DPDY|. */
    int dyn_y1;            /* DYNY1; -1: the force acts everywhere and the
                              plane is y = ny - 1 */
    float kp;              /* "KP" (0: the plane mean is recorded, dpdy held) */
    int ctrl_every;        /* 100 as the reference; 0 switches the control off */
    int ctrl_start;        /* 10000 as the reference */
    /* the precursor form of bgk_channel1/Re3000_Cyl2/DynInlet3.cu: with
       dyn_mode = 1 rows 0 .. dyn_y1 are periodic in y among themselves (row
       dyn_y1 pulls its +y neighbours from row 0 and row 0 its -y ones from
       row dyn_y1, :1131-1132) and the body force acts on every row (the
       "if(y>DYNY1) dpdy = 0.f" of MRTtest is commented out there, :1247);
       dyn_mode = 0 is MRTtest's form (no wrap, force up to dyn_y1). */
    int dyn_mode;
    int dyn_y2;            /* DYNY2: source row of the im 53 ghost copy (MRTtest/DynInlet3.cu:1765) */
    float obst2_x, obst2_y, obst2_z, obst2_r; /* "Obst2 location" / "Obst2 radius": the trip obstacle */
    int trip_until;        /* the trip obstacle exists for steps t < trip_until (ImageFcn's
                              "t < 5000", MRTtest/DynInlet3.cu:75-78); 0: never */
    int trip_z;            /* z plane of the trip's second, offset block (x-X2-3, y-Y2-3);
                              -1: none */
} lbm_config;

typedef struct lbm_sizes {
    int nx, ny, nz;        /* nodes of the level */
    int nx_pad;            /* padded x stride in HBM */
    size_t cells;          /* nx*ny*nz */
    size_t device_bytes;   /* device memory held for the level */
    int force_steps;       /* recorded force steps so far */
} lbm_sizes;

typedef struct lbm_ctx lbm_ctx;

/* Fill a config with the defaults of MRTtest/DynInlet3.cu:10-60 that apply
 * to one GPU (periodic box, BGK, FP32, no refinement). */
void lbm_config_default(lbm_config* cfg);

/* Replaces main()'s allocation and setup (MRTtest/DynInlet3.cu:2834-2990):
 * validates cfg, selects the device, allocates the A/B lattices. */
int lbm_create(const lbm_config* cfg, lbm_ctx** out);

/* Replaces ImageFcn()/ImageFcnLR() (MRTtest/DynInlet3.cu:62-94).  mask is a
 * dense [z][y][x] array of LBM_NODE_* codes, or NULL to generate it from
 * cfg->geometry. */
int lbm_set_geometry(lbm_ctx* ctx, const uint8_t* mask);

/* Mask as classified (codes LBM_NODE_*) for level 0 (coarse) or 1 (fine). */
int lbm_get_mask(lbm_ctx* ctx, int level, uint8_t* mask);

/* Replaces initialize()/initializeLR() (MRTtest/DynInlet3.cu:1186-1251) and
 * CPU_3D's initial fill (CPU_3D/CPU_3D.cpp:1218-1254).  rho, u, v, w: dense
 * fields, or NULL for rho=1 and u=(umax,0,0) -- (0,umax,0) with
 * LBM_GEOM_PRECURSOR, whose flow runs along y as initialize's does
 * (MRTtest/DynInlet3.cu:1200-1203).  kind = LBM_INIT_BGK fills
 * every node (solids included, as CPU_3D); LBM_INIT_MRT gives solids u=0. */
int lbm_init_equilibrium(lbm_ctx* ctx, int kind, const float* rho, const float* u,
                         const float* v, const float* w);

/* Taylor-Green initial state (no reference counterpart):
 * u = u0 cos(kx) sin(ky), v = -u0 sin(kx) cos(ky), w = 0, rho = 1, k = 2 pi/nx,
 * BGK equilibrium populations. */
int lbm_init_taylor_green(lbm_ctx* ctx, float u0);

/* Replaces one iteration of main()'s time loop (MRTtest/DynInlet3.cu:3119-3250):
 * update_inn/top/bot (pull, BC dispatch, collide, bounce-back, momentum
 * exchange), then, with refinement, two fine sub-steps with InterpCF and
 * InterpFC.  Asynchronous. */
int lbm_step(lbm_ctx* ctx, int nsteps);

/* As lbm_step, and the elapsed time of those launches on the context's
 * stream, from HIP events (synchronises). */
int lbm_step_timed(lbm_ctx* ctx, int nsteps, float* elapsed_ms);

/* Node fields as WriteResults computes them (MRTtest/DynInlet3.cu:2611-2620):
 * rho = sum f, (u,v,w) = momentum sums.  Any pointer may be NULL. */
int lbm_get_moments(lbm_ctx* ctx, int level, double* rho, double* u, double* v, double* w);

/* Populations, dense [q][z][y][x], converted to float or double. */
int lbm_get_populations(lbm_ctx* ctx, int level, float* f);
int lbm_get_populations_f64(lbm_ctx* ctx, int level, double* f);
int lbm_set_populations(lbm_ctx* ctx, int level, const float* f);
int lbm_set_populations_f64(lbm_ctx* ctx, int level, const double* f);

/* Momentum-exchange force per recorded step, FxFyFz (3 doubles per step),
 * un-normalised (WriteForces, :2788-2797, divides by UMAX^2 ZDIM OBSTR1).
 * Replaces the FX/FY/FZ atomicAdd arrays (:1819-1838).  Row s holds step
 * t = STARTF + s for s >= 1; steps t > STARTF are recorded, as main() sets
 * flag_F (:3130, :3183), so row 0 stays zero.  Without refinement the
 * coarse level's force is recorded; with refinement only the fine level's,
 * on the first of its two sub-steps, in fine lattice units (the reference
 * turns the coarse recording off under REFINEMENT, :3130, records
 * update_inn_LR's on sub-step 0, :3183, and WriteForces divides that by a
 * further LRLEVEL^2, :2790). */
int lbm_get_forces(lbm_ctx* ctx, double* fxyz, int max_steps, int* nsteps);

/* Global mass and momentum (sum rho, sum u, sum v, sum w) with FP64 sums in a
 * fixed order (deterministic).  No reference counterpart. */
int lbm_get_mass_momentum(lbm_ctx* ctx, int level, double out[4]);

/* lbm_get_mass_momentum for callers that cannot allocate device memory
 * (code under HIP graph capture, tight monitoring loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The same kernels and
 * the same bits as lbm_get_mass_momentum, but the four FP64 sums go to the
 * caller's DEVICE buffer out_dev[4], enqueued on `stream` (a hipStream_t;
 * NULL: the context's stream) without synchronising, and the call allocates
 * nothing: the per-workgroup partials live in a static device array of
 * 64 KiB (2048 workgroups x 4 sums; every lattice the library accepts needs
 * at most that many, and a larger need is refused with LBM_EINVAL) that each
 * GPU holds for the life of the process and every context on that GPU
 * shares.  Calls that overlap on one GPU (several contexts or streams) race
 * on it; serialise them. */
int lbm_mass_momentum_noalloc(lbm_ctx* ctx, int level, double* out_dev, void* stream);

/* Tecplot ASCII output as WriteResults / WriteResultsLR
 * (MRTtest/DynInlet3.cu:2601-2787): <casename>.dat (coarse level, POINT zone,
 * X,Y,Z,u,v,w,rho, plus velAv[0..2], ufluc, vfluc, wfluc, Smag with running
 * statistics), <casename>_LR.dat (fine level in coarse units; with
 * statistics velAv[0..1], ufluc, vfluc, Smag as WriteResultsLR) and
 * <casename>_slice.dat (coarse plane z = 1, fine plane k = 3). */
int lbm_write_tecplot(lbm_ctx* ctx, const char* casename);

int lbm_query_sizes(lbm_ctx* ctx, int level, lbm_sizes* out);

/* History of the dynamic-inlet control: one row (t, Av_V, dpdy after the
 * update) per control event, oldest first; the reference writes the rows
 * with t % 1000 == 0 to CASENAME.vel (MRTtest/DynInlet3.cu:3170).  With
 * rows == NULL only *nrows is set. */
int lbm_get_inlet_control(lbm_ctx* ctx, double* rows, int max_rows, int* nrows);

/* AverageV (MRTtest/DynInlet3.cu:2561-2600) on the current coarse lattice:
 * the sum of v = sum_i c_iy f_i over the plane y (all x, all z) in FP64, and
 * the reference's mean, that sum / ((nx - 2) nz).  No reference
 * counterpart as an entry point. */
int lbm_plane_mean_v(lbm_ctx* ctx, int y, double* sum, double* mean);

/* Body force currently applied (DPDY as the P-controller left it). */
int lbm_get_dpdy(lbm_ctx* ctx, float* dpdy);

/* Steps taken so far (the reference's t). */
int lbm_get_time(lbm_ctx* ctx, long long* t);

const char* lbm_last_error(lbm_ctx* ctx);

void lbm_destroy(lbm_ctx* ctx);

/*
 * Running velocity statistics, fused into the stream-collide (replaces
 * vel_av / vel_fluc and their LR forms, MRTtest/DynInlet3.cu:1139-1184 and
 * :1773-1799, :2153-2170, switched by VELAV / START_VELAV / START_VELFLUC,
 * :58-60).  From step start_av each collided node keeps the running mean of
 * its post-collision velocity sum c_i f_i; from step start_fluc the mean is
 * frozen and the running mean of (u - mean)^2 is kept.  The fine level
 * weighs each sub-step by LRFACTOR = 1/2 as the reference does.
 */
int lbm_stats_enable(lbm_ctx* ctx, long long start_av, long long start_fluc);
/* mean[3][nz][ny][nx] (u, v, w), dev[3][nz][ny][nx]; either may be NULL */
int lbm_stats_get(lbm_ctx* ctx, int level, float* mean, float* dev);

/*
 * Learned eddy-viscosity closure (no reference counterpart: the reference's
 * closure is constant-Cs Smagorinsky, MRTtest/DynInlet3.cu:628-735).  A per
 * node MLP 7 -> 64 -> 64 -> 64 -> 1 (GELU, softplus output) with BF16 weights
 * and FP32 accumulation on the matrix cores; trained by its own HIP forward,
 * backward and Adam kernels; inference fused into the collide.
 * Parameter vector (FP32, 8897): W1[7][64] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] w4[64] b4.
 */
typedef struct lbm_mlp lbm_mlp;

int lbm_mlp_create(uint64_t seed, lbm_mlp** out);
void lbm_mlp_destroy(lbm_mlp* m);
const char* lbm_mlp_last_error(lbm_mlp* m);
int lbm_mlp_num_params(void);
int lbm_mlp_set_params(lbm_mlp* m, const float* params);
int lbm_mlp_get_params(lbm_mlp* m, float* params);
/* input standardisation (x - mu)/sd, the training range [lo, hi] of each raw
 * input (outside it the collide uses the Smagorinsky rate) and the output
 * scale: nu_t = nu_scale * softplus(z) */
int lbm_mlp_set_norm(lbm_mlp* m, const float* mu, const float* sd, const float* lo, const float* hi,
                     float nu_scale);
/* training data on the device: X[n][7] raw inputs, y[n] targets (nu_t) */
int lbm_mlp_set_data(lbm_mlp* m, const float* X, const float* y, long long n);
/* gradient of the loss mean((softplus(z) - y/nu_scale)^2) (the squared
 * error of nu_t in units of nu_scale) over samples [offset, offset+count),
 * FP64 loss reduction; grad (host, 8897) may be NULL */
int lbm_mlp_gradient(lbm_mlp* m, long long offset, int count, float* grad, double* loss);
/* Adam over consecutive batches of the data set in its stored order
 * (cycling; shuffle before lbm_mlp_set_data); loss_hist[steps] (may be NULL) */
int lbm_mlp_train(lbm_mlp* m, int steps, int batch, float lr, double* loss_hist);
/* inference on raw inputs X[n][7] -> nu[n] */
int lbm_mlp_forward(lbm_mlp* m, const float* X, long long n, float* nu);

/* attach (or, with m = NULL, detach) the closure to an MRT-LES context: its
 * collide uses tau = tau0 + 3 nu_t, nu_t capped at nu_cap */
int lbm_set_closure(lbm_ctx* ctx, lbm_mlp* m, float nu_cap);
/* evaluate the network every `every` steps (default 1); the steps between
 * evaluations collide at the per-node rates of the last one */
int lbm_set_closure_interval(lbm_ctx* ctx, int every);
/* where the network's eddy viscosity goes (no counterpart in the reference):
 * 0 (default) tau = tau0 + 3 nu_t; 1 tau = tau_Smagorinsky + 3 nu_t, the
 * network as a learned correction on top of the configured CS */
int lbm_set_closure_base(lbm_ctx* ctx, int smagorinsky);
/* the per-node collision rate of the last closure step, dense [nz][ny][nx]
 * (coarse level): > 0 the network's, < 0 minus the Smagorinsky rate a node
 * fell back to (inputs outside the training range); a diagnostic of the
 * closure, which has no counterpart in the reference */
int lbm_get_closure_rates(lbm_ctx* ctx, float* rates);
/* what lbm_closure_samples takes as the eddy-viscosity target: 0 (the
 * default) the filtered field's subfilter stress alone; 1 also the run's own
 * Smagorinsky stress, filtered (the filtered equations of an LES carry both,
 * so a coarser run's closure has to stand in for both).  No counterpart in
 * the reference. */
int lbm_set_closure_target(lbm_ctx* ctx, int with_model_stress);
/* (with 1, lbm_closure_samples refuses runs whose tau0 lies within 0.1 of 1:
 * the stored post-collision stress, (1 - omega) times the one that set a
 * node's rate, then no longer shows that rate) */
/* training samples from a context's current state, box-filtered by 2 (the
 * solver's own finer run): features X[n][7] and the SGS eddy viscosity
 * y[n] = max(0, -tau_dev:S/(2 S:S)) (with lbm_set_closure_target(ctx, 1):
 * max(0, (-tau_dev:S + 2 <nu_f S_f>:S)/(2 S:S))) of a hashed fraction `keep` of the
 * filtered cells, in cell order.  When more than max_samples cells qualify
 * the call returns LBM_EINVAL with *count set to the number needed (a
 * truncated selection would depend on the order of the GPU's atomics). */
int lbm_closure_samples(lbm_ctx* ctx, uint64_t seed, float keep, float* X, float* y, int max_samples,
                        int* count);

#ifdef __cplusplus
}
#endif
#endif
