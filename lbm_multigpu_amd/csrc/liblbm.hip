// liblbm.hip -- host side of the C ABI declared in include/lbm.h.
//
// Replaces the reference's main() for the hot path (MRTtest/DynInlet3.cu:
// 2834-3335): allocation, geometry classification, initialisation, the time
// loop of update kernels, force collection and output.  One context owns one
// HIP stream on the current device; all launches go to that stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lbm.h"
#include "kernels.h"
#include "refine.h"
#include "closure.h"

using namespace lbm;

namespace {

thread_local std::string g_last_error;

#ifndef LBM_FINE_OVERLAP
#define LBM_FINE_OVERLAP 1 // first fine sub-step in the coarse stream-collide's launch
#endif

struct Level {
    int nx = 0, ny = 0, nz = 0, nxp = 0, nsegw = 0;
    long long qs = 0;
    size_t cells = 0;
    int storage = LBM_STORE_F32;
    size_t esize = 4;
    void* fa = nullptr; // current state
    void* fb = nullptr; // next state
    unsigned char* mask = nullptr; // device, padded
    std::vector<unsigned char> host_mask; // dense codes
    float* stat = nullptr;         // running statistics, 6 planes of qs
    PopShift shift{};              // 16-bit storage offsets (equilibrium of the inflow)
    int blocks = 0;    // workgroups of the one-node-per-lane kernels (Shape::NW row segments each)
    size_t bytes = 0;
    int* fslot = nullptr;  // device: force partial slot per stream-collide workgroup, -1 without force nodes
    int nfslot = 0;        // workgroups holding a force node (partials one recorded step writes)
};

} // namespace

struct lbm_ctx {
    lbm_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Level lv[2];
    int nlevels = 1;
    float* u_in = nullptr;     // west inlet profile [nz*ny]
    float* v_in = nullptr;     // south inlet profile [nz*nx]
    bool user_mask = false;    // the coarse mask came from the caller (no generator, no trip switch)
    bool trip_on = false;      // the classified mask holds the trip obstacle
    double* fpart = nullptr;   // force partials: a batch of steps' wave partials, or a closure step's block partials
    size_t fpart_cap = 0;      // doubles allocated in fpart
    double* fstage = nullptr;  // first-stage sums of the closure's block partials (k_reduce_chunks)
    // recorded steps whose wave partials wait in fpart for one batched
    // reduction: fb_n steps from history row fb_first, fb_w wave slots each
    int fb_n = 0, fb_w = 0;
    long long fb_first = 0;
    double* fhist = nullptr;   // [force_steps*3]
    double* red = nullptr;     // small reduction output
    double* mpart = nullptr;   // mass partials
    double* scratch = nullptr; // this device's address of the static scratch array (lbm_mass_momentum_noalloc)
    long long t = 0;
    int force_recorded = 0;
    long long stat_av = -1, stat_fluc = -1; // statistics phases (off while stat_av < 0)
    bool geometry_set = false;
    bool initialised = false;
    // local refinement
    RefineGeom rg{};
    // interface data, in the interface's arithmetic type (double for FP64
    // storage, float otherwise)
    void* Mold = nullptr;      // footprint moments at the step's start
    void* Mnew = nullptr;      // and at its end
    int fc_nb[3] = {0, 0, 0};  // FC workgroups per axis (FC_TX x FC_TY x FC_TZ coarse nodes each)
    void* Pold = nullptr;      // CF data (moments + strain rates, 16 values per node) at the step's start
    void* Pnew = nullptr;      // at its end (the next step's Pold)
    void* Phalf = nullptr;     // at its midpoint (temporal interpolation)
    // learned closure
    lbm_mlp* mlp = nullptr;    // not owned
    float* dwall = nullptr;    // wall distance per node (padded layout)
    float nu_cap = 0.05f;
    int closure_every = 1;     // network refresh interval (steps)
    int closure_smag_base = 0; // the network's nu_t adds to the Smagorinsky tau (1) or to tau0 (0)
    int closure_target = 0;    // lbm_closure_samples: 1 adds the run's own model stress to the target
    float dpdy = 0.f;          // body force as the inlet control leaves it
    std::vector<double> ctrl;  // control history, rows (t, Av_V, dpdy)
    long long closure_t0 = 0;  // step of the first refresh
    float* omega_field = nullptr; // per-node rate between refreshes (padded layout)
    std::string err;
};

struct lbm_mlp {
    float* params = nullptr;   // FP32 master parameters
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* grad = nullptr;
    short* wbf = nullptr;      // BF16 copies for the matrix cores (training)
    MlpChain* chain = nullptr; // BF16 weights arranged for the inference chain
    float* X = nullptr;        // device training data
    float* y = nullptr;
    long long n = 0;
    float* partials = nullptr; // per-tile gradients
    double* lossp = nullptr;
    double* dloss = nullptr;
    int max_tiles = 0;
    lbm::MlpNorm nrm{};
    long long t = 0;           // Adam step
    hipStream_t stream = nullptr;
    std::string err;
};

static int refine_setup(lbm_ctx* ctx);
static void classify_fine(lbm_ctx* ctx, bool user_mask);
static int refine_extract(lbm_ctx* ctx, const void* Mold, void* M, void* Pnew, void* Phalf);
static int refine_cf(lbm_ctx* ctx, const void* P, hipStream_t stream);
static int refine_fc(lbm_ctx* ctx);
static int lbm_internal_fine_step(lbm_ctx* ctx, float omega_f, double* fpart);
static int set_geometry_at(lbm_ctx* ctx, const uint8_t* mask, long long t);
static int flush_forces(lbm_ctx* ctx);

static int fail(lbm_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    g_last_error = msg;
    return code;
}

// Device memory of lbm_mass_momentum_noalloc: the per-workgroup FP64
// partials of one mass/momentum launch (at most MM_GRID_MAX workgroups x 4
// sums = 64 KiB).  Every GPU holds its own copy of this array at its own
// address for the life of the process; host code uses that address
// (noalloc_scratch), never the array's name, which on the host is only a
// stand-in.  Each call overwrites the partial slots it reduces, so nothing
// has to start at zero; calls that overlap on one GPU race on it.
constexpr int NOALLOC_SCRATCH_DOUBLES = MM_GRID_MAX * 4;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

// hipGetSymbolAddress of `scratch` once per device, with that device
// current; the address is kept per device for every later call.
static int noalloc_scratch(lbm_ctx* ctx, int device, double** out)
{
    static std::mutex mu;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, LBM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[device] = static_cast<double*>(p);
    }
    *out = addr[device];
    return LBM_OK;
}

#define HIP_TRY(ctx, expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, LBM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static size_t elem_size(int storage)
{
    switch (storage) {
    case LBM_STORE_F32: return 4;
    case LBM_STORE_F64: return 8;
    default: return 2;
    }
}

extern "C" void lbm_config_default(lbm_config* c)
{
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 64;
    c->bc[0] = c->bc[1] = c->bc[2] = LBM_BC_PERIODIC;
    c->model = LBM_MODEL_BGK;
    c->halfway = 0;
    c->storage = LBM_STORE_F32;
    c->omega = 1.0f;
    c->cs = 0.0f;
    c->dpdy = 0.0f;
    c->umax = 0.0f;
    c->inlet_amp = 0.0f;
    c->seed = 0;
    // d'Humieres 2002 rates (MRTtest/DynInlet3.cu:774-783)
    c->s_dhum[1] = 1.19f;
    c->s_dhum[2] = 1.4f;
    c->s_dhum[4] = c->s_dhum[6] = c->s_dhum[8] = 1.2f;
    c->s_dhum[10] = c->s_dhum[12] = 1.4f;
    c->s_dhum[16] = c->s_dhum[17] = c->s_dhum[18] = 1.98f;
    c->geometry = LBM_GEOM_NONE;
    c->startf = 0;
    c->force_steps = 0;
    c->refinement = 0;
    c->temporal_interp = 1;
    // dynamic-inlet control as main() runs it (MRTtest/DynInlet3.cu:3147)
    c->dyn_y1 = -1;
    c->kp = 0.0f;
    c->ctrl_every = 100;
    c->ctrl_start = 10000;
    c->dyn_mode = 0;
    c->dyn_y2 = 1; // DYNY2 of the precursor programs (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:26)
    c->trip_until = 0;
    c->trip_z = -1;
}

// ---------------------------------------------------------------- helpers

// splitmix64 finaliser over (seed, counter), top 32 bits.
static uint32_t hash_u32(uint64_t seed, uint64_t counter)
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

// u_in[z*ny+y] = u0 + amp*(2r-1), r = (h>>8) 2^-24 (separate statements: no
// contraction, so the table is bit-identical to any IEEE FP32 evaluation).
static void inlet_profile(int ny, int nz, float u0, float amp, uint64_t seed, float* out)
{
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            const uint64_t cell = (uint64_t)z * (uint64_t)ny + (uint64_t)y;
            const uint32_t h = hash_u32(seed, cell);
            volatile float r = (float)(h >> 8) * (1.0f / 16777216.0f);
            volatile float t = 2.0f * r - 1.0f;
            volatile float d = amp * t;
            out[cell] = u0 + d;
        }
}

extern "C" uint32_t lbm_hash_u32(uint64_t seed, uint64_t counter) { return hash_u32(seed, counter); }

extern "C" void lbm_inlet_profile(int ny, int nz, float u0, float amp, uint64_t seed, float* out)
{
    inlet_profile(ny, nz, u0, amp, seed, out);
}

// Padded row length: whole 64-node wavefront segments.
static int padded_nx(int nx, int) { return ((nx + 63) / 64) * 64; }

// Kernels address a direction plane with 32-bit byte offsets.
// Extra elements between consecutive direction planes.  A wavefront reads
// and writes the same cell offset in all 19 planes at once; when the plane
// size is a multiple of a large power of two (256 x 128 x 128 nodes: 16 MiB
// in FP32) those 38 streams fall on few HBM channels.  7 x 64 elements
// staggers the plane starts by 7 x 256 B (FP32) and keeps every row on a
// wavefront boundary; measured +8 % FP32 on 256 x 128 x 128, +1.4 % FP32
// and +3 % FP16 on 512 x 128 x 128, any non-zero pad doing about as well.
#ifndef LBM_PLANE_PAD
#define LBM_PLANE_PAD 448
#endif
static_assert(LBM_PLANE_PAD % 64 == 0, "plane padding keeps rows on wavefront boundaries");

static bool plane_fits(int nx, int ny, int nz, int storage)
{
    return ((long long)padded_nx(nx, storage) * ny * nz + LBM_PLANE_PAD) * elem_size(storage) <= (1LL << 32) - 64;
}

static int validate(const lbm_config* c, std::string& why)
{
    if (c->nx < 2 || c->ny < 1 || c->nz < 1) { why = "domain must be at least 2x1x1"; return 0; }
    if ((long long)c->nx * c->ny * c->nz > (1LL << 30)) { why = "domain exceeds 2^30 nodes"; return 0; }
    for (int a = 0; a < 3; ++a)
        if (c->bc[a] < 0 || c->bc[a] > 2) { why = "bc must be LBM_BC_PERIODIC/CLAMP/WALL"; return 0; }
    if (c->model < 0 || c->model > 3) { why = "unknown model"; return 0; }
    if (c->storage < 0 || c->storage > 3) { why = "unknown storage"; return 0; }
    if (!plane_fits(c->nx, c->ny, c->nz, c->storage)) { why = "a padded direction plane exceeds 4 GiB"; return 0; }
    if (!(c->omega > 0.0f && c->omega < 2.0f)) { why = "omega must lie in (0, 2) (tau > 0.5)"; return 0; }
    if (c->cs < 0.0f) { why = "CS must be >= 0"; return 0; }
    if (c->cs > 0.0f && c->model == LBM_MODEL_MRT_CPU3D) { why = "MRT_CPU3D has no LES"; return 0; }
    if (c->force_steps < 0 || c->startf < 0) { why = "force window must be >= 0"; return 0; }
    if (c->geometry < 0 || c->geometry > 3) { why = "unknown geometry"; return 0; }
    if (c->dyn_y1 < -1 || c->dyn_y1 >= c->ny) { why = "DYNY1 must be -1 or a row of the domain"; return 0; }
    if (c->dyn_mode < 0 || c->dyn_mode > 1) { why = "the dynamic-inlet mode must be 0 or 1"; return 0; }
    if (c->dyn_mode == 1 && c->dyn_y1 < 1) { why = "the periodic precursor (mode 1) needs DYNY1 >= 1"; return 0; }
    if (c->dyn_y2 < 0 || c->dyn_y2 >= c->ny) { why = "DYNY2 must be a row of the domain"; return 0; }
    if (c->trip_until < 0) { why = "the trip window must be >= 0"; return 0; }
    if (c->geometry == LBM_GEOM_PRECURSOR && (c->dyn_y1 < 0 || c->dyn_y1 + 2 >= c->ny))
        { why = "the precursor geometry needs 0 <= DYNY1 < ny - 2 (recycling plane DYNY1+1 below the outlet)"; return 0; }
    if (c->ctrl_every < 0 || c->ctrl_start < 0) { why = "the inlet control interval and start must be >= 0"; return 0; }
    if (c->kp != 0.0f && !(c->umax > 0.0f)) { why = "the inlet P-control (KP) needs UMAX > 0"; return 0; }
    if (c->model == LBM_MODEL_MRT_DHUMIERES)
        for (int k : {1, 2, 4, 6, 8, 10, 12, 16, 17, 18})
            if (!(c->s_dhum[k] > 0.0f && c->s_dhum[k] < 2.0f)) { why = "d'Humieres rates must lie in (0,2)"; return 0; }
    if (c->refinement) {
        const int nf[3] = {c->lr_nx, c->lr_ny, c->lr_nz};
        const float x0[3] = {c->lr_x0, c->lr_y0, c->lr_z0};
        const int n[3] = {c->nx, c->ny, c->nz};
        for (int a = 0; a < 3; ++a) {
            if (nf[a] < 12 || nf[a] % 2) { why = "LR domain size must be even and >= 12 per axis"; return 0; }
            const float ai = x0[a] - 0.25f;
            if (ai != std::floor(ai)) { why = "LR location must be an integer + 0.25 per axis"; return 0; }
            if (ai < 1.0f || ai + nf[a] / 2 > n[a] - 2) { why = "the refinement patch must lie inside the domain"; return 0; }
        }
        if (c->model == LBM_MODEL_MRT_CPU3D) { why = "refinement needs an LES-family or BGK model"; return 0; }
        if (!plane_fits(nf[0], nf[1], nf[2], c->storage)) { why = "a padded fine-level plane exceeds 4 GiB"; return 0; }
    }
    return 1;
}

static int alloc_level(lbm_ctx* ctx, Level& L, int nx, int ny, int nz, int storage)
{
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.nxp = padded_nx(nx, storage);    // rows padded to whole wavefront segments
    L.nsegw = L.nxp / 64;              // 64-node wavefront segments per row
    L.qs = (long long)L.nxp * ny * nz + LBM_PLANE_PAD;
    L.cells = (size_t)nx * ny * nz;
    L.storage = storage;
    L.esize = elem_size(storage);
    // 16-bit formats store f_i - f_i^eq(1, (UMAX, 0, 0)): in the bulk of an
    // inflow-driven run the stored deviations are small, so the 11 (FP16) or
    // 8 (BF16) significant bits go to the flow's variation, not its mean
    {
        // (the offset table is keyed on c_x: y-streamwise geometries keep the plain f_i - w_i)
        const float U = ctx->cfg.geometry == LBM_GEOM_PRECURSOR ? 0.f : ctx->cfg.umax;
        for (int i = 0; i < Q; ++i) {
            const float w = float(WNUM[i]) * (1.0f / 36.0f);
            const float cu = float(CX[i]) * U;
            L.shift.v[PopShift::slot(i)] = w * (1.f + 3.f * cu + 4.5f * cu * cu - 1.5f * U * U);
        }
    }
    const int nw = nwave_of_storage(L.storage);
    const long long segs = (long long)L.nsegw * ny * nz;
    L.blocks = (int)((segs + nw - 1) / nw);
    const size_t lattice = (size_t)Q * (size_t)L.qs * L.esize;
    HIP_TRY(ctx, hipMalloc(&L.fa, lattice));
    HIP_TRY(ctx, hipMalloc(&L.fb, lattice));
    HIP_TRY(ctx, hipMalloc(&L.mask, (size_t)L.qs));
    HIP_TRY(ctx, hipMemsetAsync(L.fa, 0, lattice, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(L.fb, 0, lattice, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(L.mask, 0, (size_t)L.qs, ctx->stream));
    HIP_TRY(ctx, hipMalloc(&L.fslot, (size_t)L.blocks * sizeof(int)));
    HIP_TRY(ctx, hipMemsetAsync(L.fslot, 0xff, (size_t)L.blocks * sizeof(int), ctx->stream));
    L.host_mask.assign(L.cells, 0);
    L.bytes = 2 * lattice + (size_t)L.qs;
    return LBM_OK;
}

static void free_level(Level& L)
{
    if (L.fa) (void)hipFree(L.fa);
    if (L.fb) (void)hipFree(L.fb);
    if (L.mask) (void)hipFree(L.mask);
    if (L.stat) (void)hipFree(L.stat);
    if (L.fslot) (void)hipFree(L.fslot);
    L.fa = L.fb = nullptr;
    L.mask = nullptr;
    L.stat = nullptr;
    L.fslot = nullptr;
}

extern "C" int lbm_create(const lbm_config* cfg, lbm_ctx** out)
{
    if (!cfg || !out) return fail(nullptr, LBM_EINVAL, "lbm_create: null argument");
    *out = nullptr;
    std::string why;
    if (!validate(cfg, why)) return fail(nullptr, LBM_EINVAL, "lbm_create: " + why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, LBM_EDEVICE, "lbm_create: no HIP device visible");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(nullptr, LBM_EDEVICE, "lbm_create: hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, LBM_EDEVICE,
                    std::string("lbm_create: liblbm is built for gfx950, device is ") + prop.gcnArchName);

    lbm_ctx* ctx = new lbm_ctx();
    ctx->cfg = *cfg;
    ctx->device = dev;
    ctx->dpdy = cfg->dpdy;
    int rc;
    auto bail = [&](int code) {
        lbm_destroy(ctx);
        return code;
    };
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, LBM_EDEVICE, "lbm_create: stream creation failed"));
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
        return bail(fail(nullptr, LBM_EDEVICE, "lbm_create: event creation failed"));
    if ((rc = alloc_level(ctx, ctx->lv[0], cfg->nx, cfg->ny, cfg->nz, cfg->storage)) != LBM_OK) {
        g_last_error = ctx->err;
        return bail(rc);
    }
    Level& L = ctx->lv[0];
    std::vector<float> prof((size_t)cfg->ny * cfg->nz), vprof((size_t)cfg->nx * cfg->nz);
    inlet_profile(cfg->ny, cfg->nz, cfg->umax, cfg->inlet_amp, cfg->seed, prof.data());
    // the south inlet's table: the same generator over (x, z)
    inlet_profile(cfg->nx, cfg->nz, cfg->umax, cfg->inlet_amp, cfg->seed, vprof.data());
    if (cfg->refinement) {
        ctx->nlevels = 2;
        if ((rc = alloc_level(ctx, ctx->lv[1], cfg->lr_nx, cfg->lr_ny, cfg->lr_nz, cfg->storage)) != LBM_OK) {
            g_last_error = ctx->err;
            return bail(rc);
        }
        if ((rc = refine_setup(ctx)) != LBM_OK) {
            g_last_error = ctx->err;
            return bail(rc);
        }
    }
    // force partials are sized when the geometry (and so the force slots) is
    // known; the closure's first-stage sums cover one per workgroup of level 0
    if (hipMalloc(&ctx->u_in, prof.size() * sizeof(float)) != hipSuccess ||
        hipMalloc(&ctx->v_in, vprof.size() * sizeof(float)) != hipSuccess ||
        hipMalloc(&ctx->fstage, (size_t)(L.blocks / REDUCE_CHUNK + 1) * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->mpart, (size_t)std::max(L.blocks, MM_GRID_MAX) * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->red, 8 * sizeof(double)) != hipSuccess)
        return bail(fail(nullptr, LBM_ENOMEM, "lbm_create: device allocation failed"));
    if ((rc = noalloc_scratch(nullptr, dev, &ctx->scratch)) != LBM_OK) return bail(rc);
    if (cfg->force_steps > 0) {
        if (hipMalloc(&ctx->fhist, (size_t)cfg->force_steps * 3 * sizeof(double)) != hipSuccess)
            return bail(fail(nullptr, LBM_ENOMEM, "lbm_create: force history allocation failed"));
        (void)hipMemsetAsync(ctx->fhist, 0, (size_t)cfg->force_steps * 3 * sizeof(double), ctx->stream);
    }
    (void)hipMemcpyAsync(ctx->u_in, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream);
    (void)hipMemcpyAsync(ctx->v_in, vprof.data(), vprof.size() * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return bail(fail(nullptr, LBM_EDEVICE, "lbm_create: setup failed on the device"));
    *out = ctx;
    return LBM_OK;
}

extern "C" void lbm_destroy(lbm_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int l = 0; l < 2; ++l) free_level(ctx->lv[l]);
    for (void* p : {(void*)ctx->u_in, (void*)ctx->v_in, (void*)ctx->fpart, (void*)ctx->fstage, (void*)ctx->fhist, (void*)ctx->red, (void*)ctx->mpart,
                    (void*)ctx->Mold, (void*)ctx->Mnew, (void*)ctx->Pold, (void*)ctx->Pnew, (void*)ctx->Phalf,
                    (void*)ctx->dwall,
                    (void*)ctx->omega_field})
        if (p) (void)hipFree(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* lbm_last_error(lbm_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

// ---------------------------------------------------------------- geometry

// ImageFcn stand-ins (see LBM_GEOM_* in lbm.h), at step t.
static unsigned char classify(const lbm_config& c, int x, int y, int z, long long t)
{
    if (c.geometry == LBM_GEOM_CHANNEL) {
        // LaminarChannelCylinder1/LRMultiGPU6.cu:90-129: obstacle first, then
        // the face chain y walls -> x inlet/outlet -> z walls; node walls only
        // where the axis is not a halfway wall.
        unsigned char v = LBM_NODE_FLUID;
        if (std::fabs((float)x - c.obst_x) < c.obst_r && std::fabs((float)y - c.obst_y) < c.obst_r)
            v = LBM_NODE_SOLID_FORCE;
        const bool ywall = c.bc[1] == LBM_BC_CLAMP, zwall = c.bc[2] == LBM_BC_CLAMP;
        if (ywall && (y == 0 || y == c.ny - 1)) v = LBM_NODE_SOLID;
        else if (x == 0) v = LBM_NODE_INLET_WEST;
        else if (x == c.nx - 1) v = LBM_NODE_OUTLET_EAST;
        else if (zwall && (z == 0 || z == c.nz - 1)) v = LBM_NODE_SOLID;
        return v;
    }
    if (c.geometry == LBM_GEOM_CPU3D) {
        // CPU_3D/CPU_3D.cpp:39-75 (float arithmetic as there)
        const float dx = (float)x - c.obst_x, dy = (float)y - c.obst_y;
        if ((dx * dx + dy * dy) < c.obst_r * c.obst_r) return LBM_NODE_SOLID_FORCE;
        if (x == 0) return LBM_NODE_LID_WEST_REG;
        if (x == c.nx - 1 || y == 0 || y == c.ny - 1 || z == 0 || z == c.nz - 1) return LBM_NODE_SOLID;
        return LBM_NODE_FLUID;
    }
    if (c.geometry == LBM_GEOM_PRECURSOR) {
        // bgk_channel1/Re3000_Cyl2/DynInlet3.cu:75-94 in its order: the trip
        // blocks while t < trip_until, obstacle 1, the x walls, then the
        // recycling plane and the north outlet
        unsigned char v = LBM_NODE_FLUID;
        if (t < c.trip_until) {
            if (std::fabs((float)x - c.obst2_x) < c.obst2_r && std::fabs((float)y - c.obst2_y) < c.obst2_r)
                v = LBM_NODE_SOLID;
            if (std::fabs((float)x - c.obst2_x - 3.f) < c.obst2_r && std::fabs((float)y - c.obst2_y - 3.f) < c.obst2_r &&
                z == c.trip_z)
                v = LBM_NODE_SOLID;
        }
        if (std::fabs((float)x - c.obst_x) < c.obst_r && std::fabs((float)y - c.obst_y) < c.obst_r)
            v = LBM_NODE_SOLID_FORCE;
        if (x == 0) v = LBM_NODE_SOLID;
        else if (x == c.nx - 1) v = LBM_NODE_SOLID;
        else if (y == c.dyn_y1 + 1) v = LBM_NODE_RECYCLE;
        else if (y == c.ny - 1) v = LBM_NODE_OUTLET_NORTH;
        return v;
    }
    return LBM_NODE_FLUID;
}

// Neighbour coordinate along one axis under the level's rule; false when the
// pull leaves the domain through a wall face.  ywrap: the precursor's
// periodic rows 0 .. ywrap (y axis only; -1: none).
static bool nb_coord(int c, int d, int n, int bc, int ywrap, int& out)
{
    if (ywrap >= 0) {
        if (c == ywrap && d > 0) { out = 0; return true; }
        if (c == 0 && d < 0) { out = ywrap; return true; }
    }
    int s = c + d;
    if (s < 0 || s >= n) {
        if (bc == LBM_BC_PERIODIC) s = (s + n) % n;
        else if (bc == LBM_BC_CLAMP) s = s < 0 ? 0 : n - 1;
        else return false;
    }
    out = s;
    return true;
}

// Device mask for one level: codes plus the NEAR bit on nodes with a link
// that bounces (out through a wall face, or from a solid node in link mode),
// and NEAR_SOLID where one of those links comes from a solid node.
static void build_device_mask(const Level& L, const int bc[3], int halfway, int ywrap, std::vector<unsigned char>& dev)
{
    dev.assign((size_t)L.qs, 0);
    const int n[3] = {L.nx, L.ny, L.nz};
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x) {
                const size_t c = (size_t)x + (size_t)L.nx * (y + (size_t)L.ny * z);
                unsigned char code = L.host_mask[c];
                bool near = false, near_solid = false;
                const int p[3] = {x, y, z};
                for (int i = 1; i < Q && !near_solid; ++i) {
                    const int cc[3] = {CX[i], CY[i], CZ[i]};
                    int s[3];
                    bool out = false;
                    for (int a = 0; a < 3; ++a)
                        if (!nb_coord(p[a], -cc[a], n[a], bc[a], a == 1 ? ywrap : -1, s[a])) out = true;
                    if (out) near = true;
                    else if (halfway && !is_solid(code)) {
                        const size_t sc = (size_t)s[0] + (size_t)L.nx * (s[1] + (size_t)L.ny * s[2]);
                        if (is_solid(L.host_mask[sc])) near = near_solid = true;
                    }
                }
                dev[(size_t)x + (size_t)L.nxp * (y + (size_t)L.ny * z)] =
                    code | (near ? NODE_NEAR : 0) | (near_solid ? NODE_NEAR_SOLID : 0);
            }
}

// Force slots of one level's stream-collide workgroups: a workgroup gets the
// next slot (in workgroup order) when one of its nodes is a force-summed
// solid or pulls from one, else -1; returns the slot count.  Only those
// workgroups write a momentum-exchange partial, and the reduction sums just
// them (deterministic, fixed order).
static int build_force_slots(const Level& L, const int bc[3], int ywrap, std::vector<int>& slot)
{
    const int nw = nwave_of_storage(L.storage);
    std::vector<char> has((size_t)L.blocks, 0);
    const int n[3] = {L.nx, L.ny, L.nz};
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x) {
                if (L.host_mask[(size_t)x + (size_t)L.nx * (y + (size_t)L.ny * z)] != LBM_NODE_SOLID_FORCE) continue;
                const int p[3] = {x, y, z};
                for (int i = 0; i < Q; ++i) {
                    const int cc[3] = {CX[i], CY[i], CZ[i]};
                    int s[3];
                    bool in = true;
                    for (int a = 0; a < 3; ++a)
                        if (!nb_coord(p[a], cc[a], n[a], bc[a], a == 1 ? ywrap : -1, s[a])) in = false;
                    if (!in) continue;
                    // the node that pulls along c_i from this solid (or the solid itself)
                    const long long seg = ((long long)s[2] * L.ny + s[1]) * L.nsegw + s[0] / 64;
                    has[(size_t)(seg / nw)] = 1;
                }
            }
    slot.assign((size_t)L.blocks, -1);
    int k = 0;
    for (int b = 0; b < L.blocks; ++b)
        if (has[(size_t)b]) slot[(size_t)b] = k++;
    return k;
}

// The rows 0 .. ywrap of the precursor wrap (dyn_mode 1), or -1.
static int y_wrap_of(const lbm_config& c) { return c.dyn_mode == 1 ? c.dyn_y1 : -1; }

static int upload_level_mask(lbm_ctx* ctx, Level& L, const int bc[3], int ywrap)
{
    std::vector<unsigned char> dev;
    build_device_mask(L, bc, ctx->cfg.halfway, ywrap, dev);
    HIP_TRY(ctx, hipMemcpyAsync(L.mask, dev.data(), dev.size(), hipMemcpyHostToDevice, ctx->stream));
    std::vector<int> slot;
    L.nfslot = build_force_slots(L, bc, ywrap, slot);
    HIP_TRY(ctx, hipMemcpyAsync(L.fslot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // both copies read host vectors that end with this call
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

// Classify (from the generator at step t, or the caller's mask) and upload
// every level's mask and force slots.
static int set_geometry_at(lbm_ctx* ctx, const uint8_t* mask, long long t)
{
    Level& L = ctx->lv[0];
    const lbm_config& c = ctx->cfg;
    if (int rc = flush_forces(ctx)) return rc; // the pending partials follow the old force slots
    if (mask) {
        for (size_t i = 0; i < L.cells; ++i)
            if (mask[i] > LBM_NODE_MAX)
                return fail(ctx, LBM_EINVAL, "lbm_set_geometry: unknown node code " + std::to_string(mask[i]));
        for (size_t i = 0; i < L.cells; ++i)
            if (mask[i] == LBM_NODE_GHOST_S && c.dyn_y1 < 1)
                return fail(ctx, LBM_EINVAL, "lbm_set_geometry: the im 52 ghost copy reads row DYNY1-1: DYNY1 must be >= 1");
        std::memcpy(L.host_mask.data(), mask, L.cells);
    } else {
        for (int z = 0; z < L.nz; ++z)
            for (int y = 0; y < L.ny; ++y)
                for (int x = 0; x < L.nx; ++x)
                    L.host_mask[(size_t)x + (size_t)L.nx * (y + (size_t)L.ny * z)] = classify(c, x, y, z, t);
    }
    ctx->user_mask = mask != nullptr;
    ctx->trip_on = !mask && c.geometry == LBM_GEOM_PRECURSOR && t < c.trip_until;
    if (int rc = upload_level_mask(ctx, L, c.bc, y_wrap_of(c))) return rc;
    if (ctx->nlevels > 1) {
        Level& F = ctx->lv[1];
        classify_fine(ctx, mask != nullptr);
        const int fbc[3] = {LBM_BC_CLAMP, LBM_BC_CLAMP, LBM_BC_CLAMP};
        if (int rc = upload_level_mask(ctx, F, fbc, -1)) return rc;
    }
    ctx->geometry_set = true;
    return LBM_OK;
}

extern "C" int lbm_set_geometry(lbm_ctx* ctx, const uint8_t* mask)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_set_geometry: null context");
    return set_geometry_at(ctx, mask, ctx->t);
}

extern "C" int lbm_get_mask(lbm_ctx* ctx, int level, uint8_t* mask)
{
    if (!ctx || !mask) return fail(ctx, LBM_EINVAL, "lbm_get_mask: null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_get_mask: bad level");
    const Level& L = ctx->lv[level];
    std::memcpy(mask, L.host_mask.data(), L.cells);
    return LBM_OK;
}

// ---------------------------------------------------------------- dispatch

template <class F>
static int with_storage(int storage, F&& fn)
{
    switch (storage) {
    case LBM_STORE_F32: return fn((float*)nullptr);
    case LBM_STORE_F64: return fn((double*)nullptr);
    case LBM_STORE_F16: return fn((Half16*)nullptr);
    case LBM_STORE_BF16: return fn((BHalf16*)nullptr);
    }
    return LBM_EINVAL;
}

static int grid_for(size_t n)
{
    size_t g = (n + 255) / 256;
    return (int)std::min<size_t>(g, 65536);
}

static int init_level(lbm_ctx* ctx, int level, int kind, const float* rho, const float* u, const float* v,
                      const float* w, float u0)
{
    Level& L = ctx->lv[level];
    float* d[4] = {nullptr, nullptr, nullptr, nullptr};
    const float* h[4] = {rho, u, v, w};
    for (int k = 0; k < 4; ++k)
        if (h[k]) {
            HIP_TRY(ctx, hipMalloc(&d[k], L.cells * sizeof(float)));
            HIP_TRY(ctx, hipMemcpyAsync(d[k], h[k], L.cells * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        }
    int rc = with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(k_init<S>, dim3(grid_for(L.cells)), dim3(256), 0, ctx->stream, (S*)L.fa, (S*)L.fb,
                           L.qs, L.nx, L.ny, L.nz, L.nxp, L.mask, d[0], d[1], d[2], d[3], u0,
                           (int)(ctx->cfg.geometry == LBM_GEOM_PRECURSOR), kind, L.shift);
        return LBM_OK;
    });
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; ++k)
        if (d[k]) (void)hipFree(d[k]);
    return rc;
}

// A new initial state restarts the closure's refresh cycle at t = 0 and
// clears the per-node rates the previous state's network left.
static int reset_closure_state(lbm_ctx* ctx)
{
    ctx->closure_t0 = 0;
    if (ctx->omega_field) {
        HIP_TRY(ctx, hipMemsetAsync(ctx->omega_field, 0, (size_t)ctx->lv[0].qs * sizeof(float), ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return LBM_OK;
}

extern "C" int lbm_init_equilibrium(lbm_ctx* ctx, int kind, const float* rho, const float* u, const float* v,
                                    const float* w)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_init_equilibrium: null context");
    if (kind != LBM_INIT_BGK && kind != LBM_INIT_MRT)
        return fail(ctx, LBM_EINVAL, "lbm_init_equilibrium: kind must be LBM_INIT_BGK or LBM_INIT_MRT");
    if (!ctx->geometry_set) {
        int rc = lbm_set_geometry(ctx, nullptr);
        if (rc) return rc;
    }
    int rc = init_level(ctx, 0, kind, rho, u, v, w, ctx->cfg.umax);
    if (rc) return rc;
    if ((rc = reset_closure_state(ctx))) return rc;
    if (ctx->nlevels > 1) {
        rc = init_level(ctx, 1, kind, nullptr, nullptr, nullptr, nullptr, ctx->cfg.umax);
        if (rc) return rc;
        rc = refine_extract(ctx, nullptr, ctx->Mold, ctx->Pold, nullptr);
        if (rc) return rc;
    }
    ctx->t = 0;
    ctx->force_recorded = 0;
    ctx->fb_n = 0;
    ctx->dpdy = ctx->cfg.dpdy;
    ctx->ctrl.clear();
    ctx->initialised = true;
    return LBM_OK;
}

extern "C" int lbm_init_taylor_green(lbm_ctx* ctx, float u0)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_init_taylor_green: null context");
    if (!ctx->geometry_set) {
        int rc = lbm_set_geometry(ctx, nullptr);
        if (rc) return rc;
    }
    int rc = init_level(ctx, 0, 2, nullptr, nullptr, nullptr, nullptr, u0);
    if (rc) return rc;
    if ((rc = reset_closure_state(ctx))) return rc;
    ctx->t = 0;
    ctx->force_recorded = 0;
    ctx->fb_n = 0;
    ctx->dpdy = ctx->cfg.dpdy;
    ctx->ctrl.clear();
    ctx->initialised = true;
    return LBM_OK;
}

static StepParams step_params(lbm_ctx* ctx, const Level& L, float omega)
{
    const lbm_config& c = ctx->cfg;
    StepParams p;
    p.fin = L.fa;
    p.fout = L.fb;
    p.mask = L.mask;
    p.u_in = ctx->u_in;
    p.fpart = nullptr;
    p.qs = L.qs;
    p.nx = L.nx; p.ny = L.ny; p.nz = L.nz; p.nxp = L.nxp;
    p.nsegw = L.nsegw;
    p.bcx = c.bc[0]; p.bcy = c.bc[1]; p.bcz = c.bc[2];
    p.halfway = c.halfway;
    p.omega = omega;
    p.cs = c.cs;
    // body force: the control's current DPDY; on the fine level scaled by
    // LRFACTOR = 1/2 (a force density in lattice units halves under the
    // 2:1 acoustic scaling).  The reference's fine-level kernels pass the
    // constant LRFACTOR itself as mrt_collide's dpdy argument
    // (MRTtest/DynInlet3.cu:1918, :2024, :2151), which AddForce would apply
    // as a body force of 0.5 per fine step; that is not reproduced.
    const bool fine = &L != &ctx->lv[0];
    p.dpdy = fine ? 0.5f * ctx->dpdy : ctx->dpdy;
    // rows past DYNY1 get no force (:1744); fine row j sits at coarse y = lr_y0 + j/2.
    // The periodic precursor (dyn_mode 1) forces every row, as bgk_channel1 does.
    p.dyn_y1 = INT_MAX;
    if (c.dyn_y1 >= 0 && c.dyn_mode == 0)
        p.dyn_y1 = fine ? (int)std::floor(2.0f * ((float)c.dyn_y1 - c.lr_y0)) : c.dyn_y1;
    p.v_in = ctx->v_in;
    p.fslot = nullptr;
    p.y_wrap = fine ? -1 : y_wrap_of(c);
    // ghost sources: im 52 row DYNY1-1, im 53 row DYNY2, im 54 row 1 (clamped into the level)
    const int gy[3] = {c.dyn_y1 - 1, c.dyn_y2, 1};
    for (int k = 0; k < 3; ++k) p.ghost_y[k] = std::min(std::max(gy[k], 0), L.ny - 1);
    p.u_lid = c.umax;
    p.s1 = c.s_dhum[1]; p.s2 = c.s_dhum[2]; p.s4 = c.s_dhum[4]; p.s10 = c.s_dhum[10]; p.s16 = c.s_dhum[16];
    p.omega_node = nullptr;
    p.sh = L.shift;
    p.fp_n[0] = p.fp_n[1] = p.fp_n[2] = 0; // no footprint extraction (coarse_step sets it)
    p.fp_a[0] = p.fp_a[1] = p.fp_a[2] = 0;
    p.fp_mold = nullptr;
    p.fp_m = p.fp_pnew = p.fp_phalf = nullptr;
    p.xcd_chunk = xcd_chunk_of(L.storage, L.nsegw);
    // the level's stream-collide launches take L.blocks workgroups
    p.map = node_map_of(L.nsegw, L.ny, p.xcd_chunk, L.blocks);
    // both lattices within the 256 MB last-level cache: normal stores, so the
    // next step's loads can hit it (128x96x96 FP32: 31 vs 25 GLUPS); beyond
    // it, nontemporal stores (512x128x128 FP32: 35.3 vs 32.6 GLUPS)
    p.nt_store = 2 * Q * L.qs * (long long)L.esize > (256LL << 20);
    p.whole_ok = Q * L.qs * (long long)L.esize < (4LL << 30);
    p.stat = L.stat;
    p.stat_mode = 0;
    p.stat_n = 0.f;
    p.stat_w = 1.f;
    if (L.stat && ctx->stat_av >= 0 && ctx->t >= ctx->stat_av) {
        const bool dev = ctx->t >= ctx->stat_fluc;
        p.stat_mode = dev ? 2 : 1;
        p.stat_n = (float)(ctx->t - (dev ? ctx->stat_fluc : ctx->stat_av));
        p.stat_w = (&L == &ctx->lv[0]) ? 1.f : 0.5f; // LRFACTOR on the fine level
    }
    return p;
}

template <class S, bool NT, bool WHOLE>
static void launch_sc_kernel(int model, const StepParams& p, int blocks, hipStream_t s)
{
    const dim3 blk(Shape<S>::THREADS), grd(blocks);
    switch (model) {
    case LBM_MODEL_BGK: hipLaunchKernelGGL((k_stream_collide<MODEL_BGK, S, NT, WHOLE>), grd, blk, 0, s, p); break;
    case LBM_MODEL_MRT_CPU3D:
        hipLaunchKernelGGL((k_stream_collide<MODEL_MRT_CPU3D, S, NT, WHOLE>), grd, blk, 0, s, p);
        break;
    case LBM_MODEL_MRT_LES: hipLaunchKernelGGL((k_stream_collide<MODEL_MRT_LES, S, NT, WHOLE>), grd, blk, 0, s, p); break;
    default: hipLaunchKernelGGL((k_stream_collide<MODEL_MRT_DHUM, S, NT, WHOLE>), grd, blk, 0, s, p); break;
    }
}

// The whole-lattice load descriptor serves FP32 lattices below 4 GiB (it
// measured no gain for the other formats).
template <class S, bool NT>
static void launch_sc_policy(int model, const StepParams& p, int blocks, hipStream_t s)
{
    if constexpr (std::is_same_v<S, float>) {
        if (p.whole_ok) {
            launch_sc_kernel<S, NT, true>(model, p, blocks, s);
            return;
        }
    }
    launch_sc_kernel<S, NT, false>(model, p, blocks, s);
}

template <class S>
static void launch_sc(int model, const StepParams& p, int blocks, hipStream_t s)
{
    if (p.nt_store) launch_sc_policy<S, true>(model, p, blocks, s);
    else launch_sc_policy<S, false>(model, p, blocks, s);
}

// Level a (blocks_a workgroups) and level b (blocks_b) in one launch.
template <class S, bool WHOLE>
static void launch_sc_two_kernel(int model, const StepParams& a, int blocks_a, const StepParams& b, int blocks_b,
                                 hipStream_t s)
{
    const dim3 blk(Shape<S>::THREADS), grd(blocks_a + blocks_b);
    switch (model) {
    case LBM_MODEL_BGK:
        hipLaunchKernelGGL((k_stream_collide_two<MODEL_BGK, S, WHOLE>), grd, blk, 0, s, a, b, blocks_a);
        break;
    case LBM_MODEL_MRT_CPU3D:
        hipLaunchKernelGGL((k_stream_collide_two<MODEL_MRT_CPU3D, S, WHOLE>), grd, blk, 0, s, a, b, blocks_a);
        break;
    case LBM_MODEL_MRT_LES:
        hipLaunchKernelGGL((k_stream_collide_two<MODEL_MRT_LES, S, WHOLE>), grd, blk, 0, s, a, b, blocks_a);
        break;
    default:
        hipLaunchKernelGGL((k_stream_collide_two<MODEL_MRT_DHUM, S, WHOLE>), grd, blk, 0, s, a, b, blocks_a);
        break;
    }
}

template <class S>
static void launch_sc_two(int model, const StepParams& a, int blocks_a, const StepParams& b, int blocks_b,
                          hipStream_t s)
{
    if constexpr (std::is_same_v<S, float>) {
        if (a.whole_ok && b.whole_ok) {
            launch_sc_two_kernel<S, true>(model, a, blocks_a, b, blocks_b, s);
            return;
        }
    }
    launch_sc_two_kernel<S, false>(model, a, blocks_a, b, blocks_b, s);
}

// Launches the fused closure collide over level L; returns its workgroup
// count (the number of force partials it wrote).
template <class S>
static long long launch_closure(lbm_ctx* ctx, const StepParams& p, const Level& L, float* omega_out)
{
    const long long segs = (long long)L.nsegw * L.ny * L.nz;
    const int blocks = (int)((segs + CLOSURE_NW - 1) / CLOSURE_NW);
    ClosureArgs ca;
    ca.omega_out = omega_out;
    ca.chain = ctx->mlp->chain;
    ca.dwall = ctx->dwall;
    ca.nrm = ctx->mlp->nrm;
    ca.nu_cap = ctx->nu_cap;
    ca.smag_base = ctx->closure_smag_base;
    const dim3 blk(64 * CLOSURE_NW), grd(blocks);
    if (ctx->cfg.model == LBM_MODEL_MRT_DHUMIERES)
        hipLaunchKernelGGL((k_stream_collide_closure<MODEL_MRT_DHUM, S>), grd, blk, 0, ctx->stream, p, ca);
    else
        hipLaunchKernelGGL((k_stream_collide_closure<MODEL_MRT_LES, S>), grd, blk, 0, ctx->stream, p, ca);
    return blocks;
}

// Force partials, replacing the FX/FY/FZ atomicAdd of the reference
// (MRTtest/DynInlet3.cu:1819-1838) with deterministic FP64 sums.  A recorded
// stream-collide step writes one partial per wavefront of each workgroup
// that holds a force node; FORCE_BATCH consecutive steps collect in fpart
// and one launch (k_reduce_steps, a workgroup per step, fixed order) sums
// them into the history.  The batch is flushed when full and before
// anything reads the history or changes the slot layout.
constexpr int FORCE_BATCH = 32;
constexpr size_t FORCE_BATCH_BYTES = 256u << 20; // the batch shrinks for larger slot counts

// fpart holds at least n doubles (synchronises before replacing it).
static int ensure_fpart(lbm_ctx* ctx, size_t n)
{
    if (ctx->fpart_cap >= n) return LBM_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->fpart) (void)hipFree(ctx->fpart);
    ctx->fpart = nullptr;
    ctx->fpart_cap = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->fpart, n * sizeof(double)));
    ctx->fpart_cap = n;
    return LBM_OK;
}

// Steps of one batch for w wave slots per step.
static int force_batch_steps(int w)
{
    const size_t per = (size_t)std::max(w, 1) * 3 * sizeof(double);
    return (int)std::max<size_t>(1, std::min<size_t>(FORCE_BATCH, FORCE_BATCH_BYTES / per));
}

// Reduce the pending batch into the history.
static int flush_forces(lbm_ctx* ctx)
{
    if (ctx->fb_n == 0) return LBM_OK;
    hipLaunchKernelGGL(k_reduce_steps, dim3(ctx->fb_n), dim3(256), 0, ctx->stream, (const double*)ctx->fpart,
                       (long long)ctx->fb_w, ctx->fhist + ctx->fb_first * 3);
    ctx->fb_n = 0;
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

// Partials buffer of the recorded step t of a level with w wave slots per
// step: the next position of the batch (flushing first when the batch is
// full, or does not continue it).
static int batch_slot(lbm_ctx* ctx, long long t, int w, double** out)
{
    const long long row = t - ctx->cfg.startf;
    if (ctx->fb_n > 0 && (w != ctx->fb_w || row != ctx->fb_first + ctx->fb_n || ctx->fb_n >= force_batch_steps(w)))
        if (int rc = flush_forces(ctx)) return rc;
    if (int rc = ensure_fpart(ctx, (size_t)force_batch_steps(w) * std::max(w, 1) * 3)) return rc;
    if (ctx->fb_n == 0) {
        ctx->fb_first = row;
        ctx->fb_w = w;
    }
    *out = ctx->fpart + (size_t)ctx->fb_n * w * 3;
    ctx->fb_n += 1;
    ctx->force_recorded = (int)std::max<long long>(ctx->force_recorded, row + 1);
    return LBM_OK;
}

// Deterministic FP64 sum of nparts block partials (3 components) into out,
// for the closure collide's one partial per workgroup: chunks of
// REDUCE_CHUNK partials per workgroup, then one workgroup over the chunk sums.
static void reduce_force_partials(lbm_ctx* ctx, long long nparts, double* out)
{
    if (nparts <= REDUCE_CHUNK) { // (no partials: the sum writes zeros)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, ctx->stream, (const double*)ctx->fpart, nparts,
                           3, out);
        return;
    }
    const long long chunks = (nparts + REDUCE_CHUNK - 1) / REDUCE_CHUNK;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)chunks), dim3(256), 0, ctx->stream, ctx->fpart, nparts, 3,
                       ctx->fstage);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->fstage, chunks, 3, out);
}

// Whether step t records its force: steps t > STARTF into slot t - STARTF
// (main() sets flag_F for t > STARTF, MRTtest/DynInlet3.cu:3130, :3183).
static bool records_force(const lbm_ctx* ctx, long long t)
{
    const lbm_config& c = ctx->cfg;
    return c.force_steps > 0 && t > c.startf && (t - c.startf) < c.force_steps;
}

// The coarse level's step parameters; with refinement its footprint nodes
// also write the moments at the step's end and the CF data of this step's
// second fine sub-step and of the next step's first.
static StepParams coarse_params(lbm_ctx* ctx)
{
    StepParams p = step_params(ctx, ctx->lv[0], ctx->cfg.omega);
    if (ctx->nlevels > 1) {
        for (int a = 0; a < 3; ++a) {
            p.fp_a[a] = ctx->rg.a[a];
            p.fp_n[a] = ctx->rg.nc[a];
        }
        p.fp_mold = ctx->Mold;
        p.fp_m = ctx->Mnew;
        p.fp_pnew = ctx->Pnew;
        p.fp_phalf = ctx->cfg.temporal_interp ? ctx->Phalf : nullptr;
    }
    return p;
}

// A fine sub-step's parameters (update_inn_LR, MRTtest/DynInlet3.cu:2051-2201):
// clamped axes at the fine rate; with fpart (the batch position of the first
// sub-step of a recorded coarse step, flag_F at :3183) the fine level's
// momentum exchange goes to the force history.
static StepParams fine_params(lbm_ctx* ctx, float omega_f, double* fpart)
{
    Level& F = ctx->lv[1];
    StepParams p = step_params(ctx, F, omega_f);
    p.bcx = p.bcy = p.bcz = LBM_BC_CLAMP;
    if (fpart) {
        p.fpart = fpart;
        p.fslot = F.fslot;
    }
    return p;
}

// One coarse-level stream-collide launch (update_inn/top/bot on one GPU),
// plus the deterministic force reduction when this step is recorded.
static int coarse_step(lbm_ctx* ctx)
{
    Level& L = ctx->lv[0];
    const lbm_config& c = ctx->cfg;
    StepParams p = coarse_params(ctx);
    // with refinement the fine level records the force, as main() turns the
    // coarse flag_F off under REFINEMENT (MRTtest/DynInlet3.cu:3130)
    const bool rec = ctx->nlevels == 1 && records_force(ctx, ctx->t);
    const bool closure = ctx->mlp && (ctx->closure_every <= 1 || (ctx->t - ctx->closure_t0) % ctx->closure_every == 0);
    if (rec && !closure) {
        // wave partials of the force workgroups, batched
        if (int rc = batch_slot(ctx, ctx->t, L.nfslot * nwave_of_storage(L.storage), &p.fpart)) return rc;
        p.fslot = L.fslot;
    } else if (rec) {
        // the closure collide writes one partial per workgroup, reduced at once
        if (int rc = flush_forces(ctx)) return rc;
        if (int rc = ensure_fpart(ctx, (size_t)L.blocks * 3)) return rc;
        p.fpart = ctx->fpart;
    }
    long long nparts = 0;
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        if (!ctx->mlp) {
            launch_sc<S>(c.model, p, L.blocks, ctx->stream);
        } else if (ctx->closure_every <= 1) {
            nparts = launch_closure<S>(ctx, p, L, ctx->omega_field);
        } else if (closure) {
            nparts = launch_closure<S>(ctx, p, L, ctx->omega_field); // refresh: network + rates for the next steps
        } else {
            StepParams q = p;
            q.omega_node = ctx->omega_field;
            launch_sc<S>(c.model, q, L.blocks, ctx->stream);
        }
        return 0;
    });
    if (rec && closure) {
        const long long row = ctx->t - c.startf;
        reduce_force_partials(ctx, nparts, ctx->fhist + row * 3);
        ctx->force_recorded = (int)std::max<long long>(ctx->force_recorded, row + 1);
    }
    std::swap(L.fa, L.fb);
    return LBM_OK;
}

// One fine sub-step launch on its own (fine_params).
static int lbm_internal_fine_step(lbm_ctx* ctx, float omega_f, double* fpart)
{
    Level& F = ctx->lv[1];
    const StepParams p = fine_params(ctx, omega_f, fpart);
    with_storage(F.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        launch_sc<S>(ctx->cfg.model, p, F.blocks, ctx->stream);
        return 0;
    });
    std::swap(F.fa, F.fb);
    return LBM_OK;
}

// ---------------------------------------------------------------- refinement

static int refine_setup(lbm_ctx* ctx)
{
    const lbm_config& c = ctx->cfg;
    RefineGeom& g = ctx->rg;
    const int nf[3] = {c.lr_nx, c.lr_ny, c.lr_nz};
    const float x0[3] = {c.lr_x0, c.lr_y0, c.lr_z0};
    for (int a = 0; a < 3; ++a) {
        g.a[a] = (int)std::lround(x0[a] - 0.25f);
        g.nf[a] = nf[a];
        g.nc[a] = nf[a] / 2 + 1;
    }
    ctx->fc_nb[0] = (nf[0] / 2 - 3 + FC_TX - 1) / FC_TX;
    ctx->fc_nb[1] = (nf[1] / 2 - 3 + FC_TY - 1) / FC_TY;
    ctx->fc_nb[2] = (nf[2] / 2 - 3 + FC_TZ - 1) / FC_TZ;
    g.omega_c = c.omega;
    // omegaLR as main() computes it (MRTtest/DynInlet3.cu:2870)
    g.omega_f = 2.0f / (1.0f + 2.0f * (2.0f / c.omega - 1.0f));
    g.cs = c.cs;
    g.eq_kind = c.interface_eq ? EQ_MRT_MEQ : (c.model == LBM_MODEL_BGK ? EQ_BGK_PROJ : EQ_BGK_MEQ);
    const size_t NC = (size_t)g.nc[0] * g.nc[1] * g.nc[2];
    // the interface computes in double for FP64 storage (the FP64 yardstick
    // end to end), in float otherwise
    const size_t rsz = c.storage == LBM_STORE_F64 ? sizeof(double) : sizeof(float);
    HIP_TRY(ctx, hipMalloc(&ctx->Mold, 9 * NC * rsz));
    HIP_TRY(ctx, hipMalloc(&ctx->Mnew, 9 * NC * rsz));
    HIP_TRY(ctx, hipMalloc(&ctx->Pold, 16 * NC * rsz));
    HIP_TRY(ctx, hipMalloc(&ctx->Pnew, 16 * NC * rsz));
    HIP_TRY(ctx, hipMalloc(&ctx->Phalf, 16 * NC * rsz));
    return LBM_OK;
}

// Fine-level node codes: the geometry generator evaluated at the fine node's
// coarse coordinates (ImageFcnLR, MRTtest/DynInlet3.cu:62-70), or, for a
// caller-supplied coarse mask, the code of the nearest coarse node.
static void classify_fine(lbm_ctx* ctx, bool user_mask)
{
    const lbm_config& c = ctx->cfg;
    Level& L = ctx->lv[0];
    Level& F = ctx->lv[1];
    for (int k = 0; k < F.nz; ++k)
        for (int j = 0; j < F.ny; ++j)
            for (int i = 0; i < F.nx; ++i) {
                const float X = c.lr_x0 + 0.5f * i, Y = c.lr_y0 + 0.5f * j, Z = c.lr_z0 + 0.5f * k;
                unsigned char v = LBM_NODE_FLUID;
                if (user_mask || (c.geometry != LBM_GEOM_CHANNEL && c.geometry != LBM_GEOM_PRECURSOR)) {
                    const int xi = std::min(std::max((int)std::lround(X), 0), L.nx - 1);
                    const int yi = std::min(std::max((int)std::lround(Y), 0), L.ny - 1);
                    const int zi = std::min(std::max((int)std::lround(Z), 0), L.nz - 1);
                    v = L.host_mask[(size_t)xi + (size_t)L.nx * (yi + (size_t)L.ny * zi)];
                    if (!is_solid(v)) v = LBM_NODE_FLUID;
                } else if (std::fabs(X - c.obst_x) < c.obst_r && std::fabs(Y - c.obst_y) < c.obst_r) {
                    v = LBM_NODE_SOLID_FORCE;
                }
                F.host_mask[(size_t)i + (size_t)F.nx * (j + (size_t)F.ny * k)] = v;
            }
}

static int refine_extract(lbm_ctx* ctx, const void* Mold, void* M, void* Pnew, void* Phalf)
{
    Level& L = ctx->lv[0];
    const size_t NC = (size_t)ctx->rg.nc[0] * ctx->rg.nc[1] * ctx->rg.nc[2];
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        using R = typename Arith<S>::R;
        hipLaunchKernelGGL(k_extract<S>, dim3((unsigned)((NC + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const S*)L.fa, L.qs, L.nxp, L.ny, ctx->rg, (const R*)Mold, (R*)M, (R*)Pnew, (R*)Phalf,
                           L.shift);
        return 0;
    });
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

static int refine_cf(lbm_ctx* ctx, const void* P, hipStream_t stream)
{
    Level& F = ctx->lv[1];
    with_storage(F.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        using R = typename Arith<S>::R;
        const long long n = frame_nodes(ctx->rg.nf);
        hipLaunchKernelGGL(k_interp_cf<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, (S*)F.fa,
                           F.qs, F.nxp, F.mask, (const R*)P, ctx->rg, F.shift);
        return 0;
    });
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

// The restriction together with the CF pass of the next step's first fine
// sub-step (from P) in one launch (k_interp_fc_cf).
static int refine_fc_cf(lbm_ctx* ctx, const void* P)
{
    Level& L = ctx->lv[0];
    Level& F = ctx->lv[1];
    const int nb = ctx->fc_nb[0] * ctx->fc_nb[1] * ctx->fc_nb[2];
    const int nb_cf = (int)((frame_nodes(ctx->rg.nf) + 255) / 256);
    with_storage(F.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        using R = typename Arith<S>::R;
        hipLaunchKernelGGL(k_interp_fc_cf<S>, dim3(nb_cf + nb), dim3(256), 0, ctx->stream, (const S*)F.fa, (S*)F.fa,
                           F.qs, F.nxp, F.mask, (const R*)P, (S*)L.fa, L.qs, L.nxp, L.ny, L.mask, ctx->rg,
                           ctx->fc_nb[0], ctx->fc_nb[1], nb_cf, F.shift);
        return 0;
    });
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

static int refine_fc(lbm_ctx* ctx)
{
    Level& L = ctx->lv[0];
    Level& F = ctx->lv[1];
    const int nb = ctx->fc_nb[0] * ctx->fc_nb[1] * ctx->fc_nb[2];
    with_storage(F.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(k_interp_fc<S>, dim3(nb), dim3(256), 0, ctx->stream, (const S*)F.fa, F.qs, F.nxp,
                           (S*)L.fa, L.qs, L.nxp, L.ny, L.mask, ctx->rg, ctx->fc_nb[0], ctx->fc_nb[1], F.shift);
        return 0;
    });
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

// AverageV on the current coarse lattice: FP64 sum of v over the plane y.
static int plane_v(lbm_ctx* ctx, int y, double* sum)
{
    Level& L = ctx->lv[0];
    const int nb = L.nsegw * L.nz; // <= 4 * L.blocks partials fit in mpart
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(k_plane_v<S>, dim3(nb), dim3(64), 0, ctx->stream, (const S*)L.fa, L.qs, L.nx, L.nxp,
                           L.ny, L.nsegw, y, ctx->mpart, L.shift);
        return 0;
    });
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->mpart, (long long)nb, 1,
                       ctx->red);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(sum, ctx->red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

// The reference's mean of that sum, in FP32 as main() forms it (:3160).
static float plane_mean(const Level& L, double sum)
{
    return (float)sum / (float)((L.nx - 2) * L.nz);
}

// Dynamic-inlet P-control after coarse step t (main(), MRTtest/DynInlet3.cu:
// 3147-3170): AverageV on the plane y = DYNY1, then dpdy += diff*KP*|DPDY|
// with diff = (Av_V - UMAX)/UMAX, in FP32.  DPDY is the configured initial
// force.  The new dpdy acts from the next step on.
static bool control_due(const lbm_ctx* ctx, long long t)
{
    const lbm_config& c = ctx->cfg;
    return c.ctrl_every > 0 && t % c.ctrl_every == 0 && t > c.ctrl_start;
}

static int inlet_control(lbm_ctx* ctx, long long t)
{
    const lbm_config& c = ctx->cfg;
    if (!control_due(ctx, t)) return LBM_OK;
    const Level& L = ctx->lv[0];
    double sum = 0;
    int rc = plane_v(ctx, c.dyn_y1 >= 0 ? c.dyn_y1 : L.ny - 1, &sum);
    if (rc) return rc;
    const float av = plane_mean(L, sum);
    if (c.kp != 0.0f) {
        const float diff = (av - c.umax) / c.umax;
        ctx->dpdy += diff * c.kp * std::fabs(c.dpdy);
    }
    ctx->ctrl.insert(ctx->ctrl.end(), {(double)t, (double)av, (double)ctx->dpdy});
    return LBM_OK;
}

extern "C" int lbm_get_inlet_control(lbm_ctx* ctx, double* rows, int max_rows, int* nrows)
{
    if (!ctx || !nrows) return fail(ctx, LBM_EINVAL, "lbm_get_inlet_control: null argument");
    const int n = (int)(ctx->ctrl.size() / 3);
    *nrows = n;
    if (rows) std::copy(ctx->ctrl.begin(), ctx->ctrl.begin() + 3 * (size_t)std::min(n, std::max(max_rows, 0)), rows);
    return LBM_OK;
}

extern "C" int lbm_plane_mean_v(lbm_ctx* ctx, int y, double* sum, double* mean)
{
    if (!ctx || !sum || !mean) return fail(ctx, LBM_EINVAL, "lbm_plane_mean_v: null argument");
    if (!ctx->initialised) return fail(ctx, LBM_ESTATE, "lbm_plane_mean_v: call lbm_init_equilibrium first");
    const Level& L = ctx->lv[0];
    if (y < 0 || y >= L.ny) return fail(ctx, LBM_EINVAL, "lbm_plane_mean_v: y outside the domain");
    int rc = plane_v(ctx, y, sum);
    if (rc) return rc;
    *mean = plane_mean(L, *sum);
    return LBM_OK;
}

extern "C" int lbm_get_dpdy(lbm_ctx* ctx, float* dpdy)
{
    if (!ctx || !dpdy) return fail(ctx, LBM_EINVAL, "lbm_get_dpdy: null argument");
    *dpdy = ctx->dpdy;
    return LBM_OK;
}

extern "C" int lbm_step(lbm_ctx* ctx, int nsteps)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_step: null context");
    if (!ctx->initialised) return fail(ctx, LBM_ESTATE, "lbm_step: call lbm_init_equilibrium first");
    if (nsteps < 0) return fail(ctx, LBM_EINVAL, "lbm_step: nsteps < 0");
    const lbm_config& c = ctx->cfg;
    // the trip obstacle exists for t < trip_until (ImageFcn(x, y, z, t)): a
    // step at which the generated mask changes
    auto geometry_switch = [&](long long t) {
        return !ctx->user_mask && c.geometry == LBM_GEOM_PRECURSOR && ctx->trip_on != (t < c.trip_until);
    };
    // the CF pass of the current step's first fine sub-step already launched
    // (at the end of the previous step of this call)
    bool cf_done = false;
    for (int s = 0; s < nsteps; ++s) {
        if (geometry_switch(ctx->t))
            if (int rc = set_geometry_at(ctx, nullptr, ctx->t)) return rc;
        if (ctx->nlevels > 1) {
            // coarse t -> t+1, the dynamic-inlet control; fine sub-steps
            // t -> t+1/2 -> t+1 (main(), MRTtest/DynInlet3.cu:3147-3245), then
            // the fine -> coarse restriction.  The first sub-step reads only
            // the fine level and the frame CF fills from the coarse data
            // packed one step earlier (Pold), so it runs in the same launch
            // as the coarse stream-collide (which packs this step's CF data),
            // its workgroups after the coarse ones.  On a control step the
            // body force it uses changes after the coarse step, so it runs
            // on its own after the control.  The CF pass of the next step's
            // first sub-step is launched before this step's restriction,
            // which it does not touch (the restriction reads fine nodes off
            // the frame and writes coarse ones; CF writes only the frame).
            const bool temporal = ctx->cfg.temporal_interp;
            // (16-bit storage: the two-level kernel would not fit the 64
            // registers those kernels are held to, so its levels run apart)
            const bool together = LBM_FINE_OVERLAP && !control_due(ctx, ctx->t) && !is16_storage(ctx->lv[0].storage);
            Level& L = ctx->lv[0];
            Level& F = ctx->lv[1];
            double* fpart = nullptr;
            if (records_force(ctx, ctx->t))
                if (int rc = batch_slot(ctx, ctx->t, F.nfslot * nwave_of_storage(F.storage), &fpart)) return rc;
            int rc = LBM_OK;
            if (!cf_done && (rc = refine_cf(ctx, ctx->Pold, ctx->stream))) return rc;
            if (together) {
                StepParams pa = coarse_params(ctx);
                StepParams pb = fine_params(ctx, ctx->rg.omega_f, fpart);
                with_storage(L.storage, [&](auto* tag) {
                    using S = std::remove_pointer_t<decltype(tag)>;
                    if constexpr (sizeof(S) != 2) launch_sc_two<S>(c.model, pa, L.blocks, pb, F.blocks, ctx->stream);
                    return 0;
                });
                std::swap(L.fa, L.fb);
                std::swap(F.fa, F.fb);
            } else {
                if ((rc = coarse_step(ctx))) return rc;
                if ((rc = inlet_control(ctx, ctx->t))) return rc;
                if ((rc = lbm_internal_fine_step(ctx, ctx->rg.omega_f, fpart))) return rc;
            }
            // (the coarse stream-collide wrote Mnew, Pnew and Phalf)
            if (temporal) rc = refine_cf(ctx, ctx->Phalf, ctx->stream);
            if (!rc) rc = lbm_internal_fine_step(ctx, ctx->rg.omega_f, nullptr);
            if (rc) return rc;
            std::swap(ctx->Mold, ctx->Mnew);
            std::swap(ctx->Pold, ctx->Pnew);
            cf_done = LBM_FINE_OVERLAP && s + 1 < nsteps && !geometry_switch(ctx->t + 1);
            if ((rc = cf_done ? refine_fc_cf(ctx, ctx->Pold) : refine_fc(ctx))) return rc;
        } else {
            if (int rc = coarse_step(ctx)) return rc;
            if (int rc = inlet_control(ctx, ctx->t)) return rc;
        }
        ctx->t += 1;
    }
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

extern "C" int lbm_step_timed(lbm_ctx* ctx, int nsteps, float* elapsed_ms)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_step_timed: null context");
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = lbm_step(ctx, nsteps);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    return LBM_OK;
}

// ---------------------------------------------------------------- readback

extern "C" int lbm_get_moments(lbm_ctx* ctx, int level, double* rho, double* u, double* v, double* w)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_get_moments: null context");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_get_moments: bad level");
    Level& L = ctx->lv[level];
    double* d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, 4 * L.cells * sizeof(double)));
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(k_macro<S>, dim3(grid_for(L.cells)), dim3(256), 0, ctx->stream, (const S*)L.fa, L.qs,
                           L.nx, L.ny, L.nz, L.nxp, d, d + L.cells, d + 2 * L.cells, d + 3 * L.cells, L.shift);
        return 0;
    });
    double* out[4] = {rho, u, v, w};
    hipError_t e = hipGetLastError();
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
        if (out[k])
            e = hipMemcpyAsync(out[k], d + k * L.cells, L.cells * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, LBM_EDEVICE, std::string("lbm_get_moments: ") + hipGetErrorString(e));
    return LBM_OK;
}

template <class T>
static int get_pops(lbm_ctx* ctx, int level, T* f, const char* who)
{
    if (!ctx || !f) return fail(ctx, LBM_EINVAL, std::string(who) + ": null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, std::string(who) + ": bad level");
    Level& L = ctx->lv[level];
    T* d = nullptr;
    const size_t n = (size_t)Q * L.cells;
    HIP_TRY(ctx, hipMalloc(&d, n * sizeof(T)));
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_pack<S, T>), dim3(grid_for(n)), dim3(256), 0, ctx->stream, (const S*)L.fa, L.qs,
                           L.nx, L.ny, L.nz, L.nxp, d, L.shift);
        return 0;
    });
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(f, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, LBM_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
    return LBM_OK;
}

template <class T>
static int set_pops(lbm_ctx* ctx, int level, const T* f, const char* who)
{
    if (!ctx || !f) return fail(ctx, LBM_EINVAL, std::string(who) + ": null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, std::string(who) + ": bad level");
    if (!ctx->geometry_set) {
        int rc = lbm_set_geometry(ctx, nullptr);
        if (rc) return rc;
    }
    Level& L = ctx->lv[level];
    T* d = nullptr;
    const size_t n = (size_t)Q * L.cells;
    HIP_TRY(ctx, hipMalloc(&d, n * sizeof(T)));
    hipError_t e = hipMemcpyAsync(d, f, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        with_storage(L.storage, [&](auto* tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL((k_unpack<S, T>), dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, L.qs, L.nx,
                               L.ny, L.nz, L.nxp, (S*)L.fa, (S*)L.fb, L.shift);
            return 0;
        });
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, LBM_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
    if (level == 0) {
        ctx->initialised = true;
        if (ctx->nlevels > 1) {
            int rc = refine_extract(ctx, nullptr, ctx->Mold, ctx->Pold, nullptr);
            if (rc) return rc;
        }
    }
    return LBM_OK;
}

extern "C" int lbm_get_populations(lbm_ctx* ctx, int level, float* f)
{
    return get_pops<float>(ctx, level, f, "lbm_get_populations");
}
extern "C" int lbm_get_populations_f64(lbm_ctx* ctx, int level, double* f)
{
    return get_pops<double>(ctx, level, f, "lbm_get_populations_f64");
}
extern "C" int lbm_set_populations(lbm_ctx* ctx, int level, const float* f)
{
    return set_pops<float>(ctx, level, f, "lbm_set_populations");
}
extern "C" int lbm_set_populations_f64(lbm_ctx* ctx, int level, const double* f)
{
    return set_pops<double>(ctx, level, f, "lbm_set_populations_f64");
}

extern "C" int lbm_get_forces(lbm_ctx* ctx, double* fxyz, int max_steps, int* nsteps)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_get_forces: null context");
    if (int rc = flush_forces(ctx)) return rc;
    const int n = std::min(ctx->force_recorded, std::max(max_steps, 0));
    if (nsteps) *nsteps = ctx->force_recorded;
    if (n > 0 && fxyz) {
        HIP_TRY(ctx, hipMemcpyAsync(fxyz, ctx->fhist, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LBM_OK;
}

// FP64 mass and momentum sums of one level, enqueued on stream s: the
// grid-stride partial kernel (mm_grid workgroups, one partial each, into
// part) and one workgroup summing them in a fixed order into out_dev[4].
// Deterministic: the same lattice gives the same bits whatever part is.
static int mass_momentum_launch(lbm_ctx* ctx, const Level& L, double* part, double* out_dev, hipStream_t s)
{
    const int G = mm_grid(L.nsegw, L.ny, L.nz);
    with_storage(L.storage, [&](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(k_mass_momentum<S>, dim3(G), dim3(64 * MM_NW), 0, s, (const S*)L.fa, L.mask, L.qs, L.nx,
                           L.ny, L.nxp, L.nsegw, L.nz, part, L.shift);
        return 0;
    });
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, (const double*)part, (long long)G, 4, out_dev);
    HIP_TRY(ctx, hipGetLastError());
    return LBM_OK;
}

extern "C" int lbm_get_mass_momentum(lbm_ctx* ctx, int level, double out[4])
{
    if (!ctx || !out) return fail(ctx, LBM_EINVAL, "lbm_get_mass_momentum: null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_get_mass_momentum: bad level");
    int rc = mass_momentum_launch(ctx, ctx->lv[level], ctx->mpart, ctx->red, ctx->stream);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(out, ctx->red, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, LBM_EDEVICE, std::string("lbm_get_mass_momentum: ") + hipGetErrorString(e));
    return LBM_OK;
}

// lbm_get_mass_momentum without any allocation or synchronisation: the same
// kernels, with the partials in this device's static scratch array and the
// four sums written to the caller's device buffer.
extern "C" int lbm_mass_momentum_noalloc(lbm_ctx* ctx, int level, double* out_dev, void* stream)
{
    if (!ctx || !out_dev) return fail(ctx, LBM_EINVAL, "lbm_mass_momentum_noalloc: null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_mass_momentum_noalloc: bad level");
    const Level& L = ctx->lv[level];
    const long long need = 4LL * mm_grid(L.nsegw, L.ny, L.nz);
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return fail(ctx, LBM_EINVAL, "lbm_mass_momentum_noalloc: " + std::to_string(need) +
                                         " partials exceed the scratch array of " +
                                         std::to_string(NOALLOC_SCRATCH_DOUBLES));
    if (!ctx->scratch) return fail(ctx, LBM_ESTATE, "lbm_mass_momentum_noalloc: no scratch address for this device");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    return mass_momentum_launch(ctx, L, ctx->scratch, out_dev, s);
}

extern "C" int lbm_query_sizes(lbm_ctx* ctx, int level, lbm_sizes* out)
{
    if (!ctx || !out) return fail(ctx, LBM_EINVAL, "lbm_query_sizes: null argument");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_query_sizes: bad level");
    const Level& L = ctx->lv[level];
    out->nx = L.nx; out->ny = L.ny; out->nz = L.nz;
    out->nx_pad = L.nxp;
    out->cells = L.cells;
    out->device_bytes = L.bytes;
    out->force_steps = ctx->force_recorded;
    return LBM_OK;
}

extern "C" int lbm_get_time(lbm_ctx* ctx, long long* t)
{
    if (!ctx || !t) return fail(ctx, LBM_EINVAL, "lbm_get_time: null argument");
    *t = ctx->t;
    return LBM_OK;
}

// Tecplot POINT zone as CPU_3D writes output.dat (CPU_3D/CPU_3D.cpp:1590-1614)
// and WriteResults its first seven columns (MRTtest/DynInlet3.cu:2604-2620).
// The "Smag" column of WriteResults (MRTtest/DynInlet3.cu:2637-2660): the
// second-order non-equilibrium moments of a node (m9, m11, m13-15 less their
// equilibrium parts, with the momentum sums as velocity) turned into the
// stress components PI and |PI| = sqrt(2 PI:PI).
static float smag_column(const float f[19])
{
    float m[19];
    for (int k = 0; k < 19; ++k) {
        float a = 0.f;
        for (int i = 0; i < 19; ++i) a += (float)mrow(k, i) * f[i];
        m[k] = a;
    }
    const float u = m[3], v = m[5], w = m[7];
    m[9] -= 2.f * u * u - (v * v + w * w);
    m[11] -= v * v - w * w;
    m[13] -= u * v;
    m[14] -= v * w;
    m[15] -= u * w;
    const float p11 = -0.5f * m[9];
    const float p22 = -(-38.f * m[9] - 3.f * m[11]) / 76.f;
    const float p33 = -(-38.f * m[9] + 3.f * m[11]) / 76.f;
    const float p12 = -1.5f * m[13], p23 = -1.5f * m[14], p13 = -1.5f * m[15];
    return std::sqrt(2.f * (p11 * p11 + p22 * p22 + p33 * p33 + 2.f * p12 * p12 + 2.f * p23 * p23 + 2.f * p13 * p13));
}

// Tecplot POINT zone per level (WriteResults, MRTtest/DynInlet3.cu:2601-2694,
// and WriteResultsLR for the fine patch in coarse coordinates).  Without
// statistics the columns are X,Y,Z,u,v,w,rho; with them the reference's full
// row: ... velAv[0..2], ufluc, vfluc, wfluc, Smag.
extern "C" int lbm_write_tecplot(lbm_ctx* ctx, const char* casename)
{
    if (!ctx || !casename) return fail(ctx, LBM_EINVAL, "lbm_write_tecplot: null argument");
    // the slice file (WriteResults/WriteResultsLR's outputslice): the coarse
    // plane z = 1 (:2672) and the fine plane k = 3 (:2762), one zone each
    const std::string slice_name = std::string(casename) + "_slice.dat";
    std::ofstream slice(slice_name);
    if (!slice) return fail(ctx, LBM_EIO, "lbm_write_tecplot: cannot open " + slice_name);
    for (int level = 0; level < ctx->nlevels; ++level) {
        const Level& L = ctx->lv[level];
        std::vector<double> r(L.cells), u(L.cells), v(L.cells), w(L.cells);
        int rc = lbm_get_moments(ctx, level, r.data(), u.data(), v.data(), w.data());
        if (rc) return rc;
        const bool stats = L.stat != nullptr;
        std::vector<float> mean, dev, pops;
        if (stats) {
            mean.resize(3 * L.cells);
            dev.resize(3 * L.cells);
            pops.resize((size_t)Q * L.cells);
            rc = lbm_stats_get(ctx, level, mean.data(), dev.data());
            if (!rc) rc = lbm_get_populations(ctx, level, pops.data());
            if (rc) return rc;
        }
        // columns: X,Y,Z,u,v,w,rho, then with statistics the coarse level's
        // velAv[0..2], ufluc, vfluc, wfluc, Smag (WriteResults, :2605) or the
        // fine level's velAv[0..1], ufluc, vfluc, Smag (WriteResultsLR, :2699)
        const int nstat = level == 0 ? 3 : 2;
        auto header = [&](std::ostream& out, int K) {
            out << "VARIABLES = \"X\",\"Y\",\"Z\",\"u\",\"v\",\"w\",\"rho\"";
            if (stats) {
                out << (level == 0 ? ",\"velAv[0]\",\"velAv[1]\",\"velAv[2]\",\"ufluc\",\"vfluc\",\"wfluc\",\"Smag\""
                                   : ",\"velAv[0]\",\"velAv[1]\",\"ufluc\",\"vfluc\",\"Smag\"");
            }
            out << "\n";
            out << "ZONE F=POINT, I=" << L.nx << ", J=" << L.ny << ", K=" << K << "\n";
        };
        auto row = [&](std::ostream& out, int x, int y, int z) {
            const size_t c = (size_t)x + (size_t)L.nx * (y + (size_t)L.ny * z);
            if (level == 0) out << x << ", " << y << ", " << z << ", ";
            else {
                // fine node position in coarse units (WriteResultsLR)
                out << ctx->cfg.lr_x0 + 0.5f * x << ", " << ctx->cfg.lr_y0 + 0.5f * y << ", "
                    << ctx->cfg.lr_z0 + 0.5f * z << ", ";
            }
            out << (float)u[c] << "," << (float)v[c] << "," << (float)w[c] << "," << (float)r[c];
            if (stats) {
                float f[19];
                for (int i = 0; i < Q; ++i) f[i] = pops[(size_t)i * L.cells + c];
                for (int a = 0; a < nstat; ++a) out << "," << mean[a * L.cells + c];
                for (int a = 0; a < nstat; ++a) out << "," << dev[a * L.cells + c];
                out << "," << smag_column(f);
            }
            out << "\n";
        };
        std::string name = std::string(casename) + (level == 0 ? ".dat" : "_LR.dat");
        std::ofstream out(name);
        if (!out) return fail(ctx, LBM_EIO, "lbm_write_tecplot: cannot open " + name);
        header(out, L.nz);
        for (int z = 0; z < L.nz; ++z)
            for (int y = 0; y < L.ny; ++y)
                for (int x = 0; x < L.nx; ++x) row(out, x, y, z);
        if (!out) return fail(ctx, LBM_EIO, "lbm_write_tecplot: write failed for " + name);
        const int zs = std::min(level == 0 ? 1 : 3, L.nz - 1);
        header(slice, 1);
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x) row(slice, x, y, zs);
    }
    if (!slice) return fail(ctx, LBM_EIO, "lbm_write_tecplot: write failed for " + slice_name);
    return LBM_OK;
}

// ---------------------------------------------------------------- statistics

// Running velocity statistics (VELAV, START_VELAV, START_VELFLUC,
// MRTtest/DynInlet3.cu:58-60): enabling allocates and zeroes the six
// per-node arrays of every level.
extern "C" int lbm_stats_enable(lbm_ctx* ctx, long long start_av, long long start_fluc)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_stats_enable: null context");
    if (start_av < 0 || start_fluc < start_av)
        return fail(ctx, LBM_EINVAL, "lbm_stats_enable: need 0 <= start_av <= start_fluc");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    for (int l = 0; l < ctx->nlevels; ++l) {
        Level& L = ctx->lv[l];
        const size_t bytes = (size_t)6 * L.qs * sizeof(float);
        if (!L.stat) HIP_TRY(ctx, hipMalloc(&L.stat, bytes));
        HIP_TRY(ctx, hipMemsetAsync(L.stat, 0, bytes, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stat_av = start_av;
    ctx->stat_fluc = start_fluc;
    return LBM_OK;
}

// mean[3][nz][ny][nx] (u v w) and dev[3][nz][ny][nx] (mean squared
// deviations); either may be NULL
extern "C" int lbm_stats_get(lbm_ctx* ctx, int level, float* mean, float* dev)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_stats_get: null context");
    if (level < 0 || level >= ctx->nlevels) return fail(ctx, LBM_EINVAL, "lbm_stats_get: no such level");
    Level& L = ctx->lv[level];
    if (!L.stat) return fail(ctx, LBM_ESTATE, "lbm_stats_get: statistics not enabled");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float> plane((size_t)L.qs);
    for (int k = 0; k < 6; ++k) {
        float* out = k < 3 ? mean : dev;
        if (!out) continue;
        HIP_TRY(ctx, hipMemcpy(plane.data(), L.stat + (size_t)k * L.qs, plane.size() * 4, hipMemcpyDeviceToHost));
        float* o = out + (size_t)(k % 3) * L.cells;
        for (int z = 0; z < L.nz; ++z)
            for (int y = 0; y < L.ny; ++y)
                std::memcpy(o + (size_t)L.nx * (y + (size_t)L.ny * z), plane.data() + (size_t)L.nxp * (y + (size_t)L.ny * z),
                            (size_t)L.nx * 4);
    }
    return LBM_OK;
}

// ---------------------------------------------------------------- closure

extern "C" int lbm_get_closure_rates(lbm_ctx* ctx, float* rates)
{
    if (!ctx || !rates) return fail(ctx, LBM_EINVAL, "lbm_get_closure_rates: null argument");
    if (!ctx->omega_field) return fail(ctx, LBM_ESTATE, "lbm_get_closure_rates: no closure attached");
    const Level& L = ctx->lv[0];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float> plane((size_t)L.qs);
    HIP_TRY(ctx, hipMemcpy(plane.data(), ctx->omega_field, plane.size() * 4, hipMemcpyDeviceToHost));
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            std::memcpy(rates + (size_t)L.nx * (y + (size_t)L.ny * z), plane.data() + (size_t)L.nxp * (y + (size_t)L.ny * z),
                        (size_t)L.nx * 4);
    return LBM_OK;
}

// Distance (cells) from each node to the nearest solid node or wall face:
// solid nodes 0, faces of LBM_BC_WALL axes half a cell outside the domain,
// then a two-pass 3-D chamfer transform (steps 1, sqrt2, sqrt3).
static std::vector<float> wall_distance(const Level& L, const int bc[3])
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    std::vector<float> d(L.cells);
    const float big = 1e9f;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t c = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
                float v = big;
                if (is_solid(L.host_mask[c])) v = 0.f;
                if (bc[0] == LBM_BC_WALL) v = std::min(v, std::min(x + 0.5f, nx - 0.5f - x));
                if (bc[1] == LBM_BC_WALL) v = std::min(v, std::min(y + 0.5f, ny - 0.5f - y));
                if (bc[2] == LBM_BC_WALL) v = std::min(v, std::min(z + 0.5f, nz - 0.5f - z));
                d[c] = v;
            }
    const float w[4] = {0.f, 1.f, 1.41421356f, 1.73205081f};
    for (int pass = 0; pass < 2; ++pass) {
        const int s = pass == 0 ? 1 : -1;
        for (int zi = 0; zi < nz; ++zi)
            for (int yi = 0; yi < ny; ++yi)
                for (int xi = 0; xi < nx; ++xi) {
                    const int x = pass == 0 ? xi : nx - 1 - xi;
                    const int y = pass == 0 ? yi : ny - 1 - yi;
                    const int z = pass == 0 ? zi : nz - 1 - zi;
                    const size_t c = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
                    float v = d[c];
                    for (int dz = -1; dz <= 0; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue; // causal half
                                const int X = x + s * dx, Y = y + s * dy, Z = z + s * dz;
                                if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
                                const float nv = d[(size_t)X + (size_t)nx * (Y + (size_t)ny * Z)] +
                                                 w[std::abs(dx) + std::abs(dy) + std::abs(dz)];
                                v = std::min(v, nv);
                            }
                    d[c] = v;
                }
    }
    return d;
}

static int mfail(lbm_mlp* m, int code, const std::string& msg)
{
    if (m) m->err = msg;
    g_last_error = msg;
    return code;
}

#define MLP_TRY(m, expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return mfail(m, LBM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int lbm_mlp_create(uint64_t seed, lbm_mlp** out)
{
    if (!out) return mfail(nullptr, LBM_EINVAL, "lbm_mlp_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return mfail(nullptr, LBM_EDEVICE, "lbm_mlp_create: no HIP device visible");
    lbm_mlp* m = new lbm_mlp();
    auto bail = [&](int rc) { lbm_mlp_destroy(m); return rc; };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(mfail(nullptr, LBM_EDEVICE, "lbm_mlp_create: stream creation failed"));
    const size_t P = MLP_NPARAM;
    const size_t nwbf = MLP_INP * MLP_H + 2 * MLP_H * MLP_H;
    if (hipMalloc(&m->params, P * 4) || hipMalloc(&m->m, P * 4) || hipMalloc(&m->v, P * 4) ||
        hipMalloc(&m->grad, P * 4) || hipMalloc(&m->wbf, nwbf * 2) || hipMalloc(&m->dloss, 8) ||
        hipMalloc(&m->chain, sizeof(MlpChain)))
        return bail(mfail(nullptr, LBM_ENOMEM, "lbm_mlp_create: device allocation failed"));
    // Glorot-uniform weights from the counter hash, zero biases
    std::vector<float> p(P, 0.f);
    auto fill = [&](int off, int fin, int fout, uint64_t salt) {
        const float a = std::sqrt(6.f / (fin + fout));
        for (int i = 0; i < fin * fout; ++i) {
            const float r = (float)(hash_u32(seed, salt * 1000003ULL + i) >> 8) * (1.0f / 16777216.0f);
            p[off + i] = a * (2.f * r - 1.f);
        }
    };
    fill(OFF_W1, MLP_IN, MLP_H, 1);
    fill(OFF_W2, MLP_H, MLP_H, 2);
    fill(OFF_W3, MLP_H, MLP_H, 3);
    fill(OFF_W4, MLP_H, 1, 4);
    MlpNorm n{};
    for (int k = 0; k < MLP_IN; ++k) { n.mu[k] = 0.f; n.inv_sd[k] = 1.f; n.lo[k] = -1e30f; n.hi[k] = 1e30f; }
    n.nu_scale = 1.f;
    m->nrm = n;
    *out = m;
    int rc = lbm_mlp_set_params(m, p.data());
    if (rc) { *out = nullptr; return bail(rc); }
    return LBM_OK;
}

extern "C" void lbm_mlp_destroy(lbm_mlp* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* q : {(void*)m->params, (void*)m->m, (void*)m->v, (void*)m->grad, (void*)m->wbf, (void*)m->chain, (void*)m->X,
                    (void*)m->y, (void*)m->partials, (void*)m->lossp, (void*)m->dloss})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" const char* lbm_mlp_last_error(lbm_mlp* m) { return m ? m->err.c_str() : g_last_error.c_str(); }

extern "C" int lbm_mlp_num_params(void) { return MLP_NPARAM; }

extern "C" int lbm_mlp_set_params(lbm_mlp* m, const float* params)
{
    if (!m || !params) return mfail(m, LBM_EINVAL, "lbm_mlp_set_params: null argument");
    const size_t P = MLP_NPARAM;
    MLP_TRY(m, hipMemcpyAsync(m->params, params, P * 4, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(m, hipMemsetAsync(m->m, 0, P * 4, m->stream));
    MLP_TRY(m, hipMemsetAsync(m->v, 0, P * 4, m->stream));
    const int nwbf = MLP_INP * MLP_H + 2 * MLP_H * MLP_H;
    hipLaunchKernelGGL(k_mlp_sync_bf16, dim3((nwbf + 255) / 256), dim3(256), 0, m->stream, m->params, m->wbf);
    hipLaunchKernelGGL(k_mlp_pack_chain, dim3(1), dim3(64), 0, m->stream, m->params, m->chain);
    MLP_TRY(m, hipGetLastError());
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    m->t = 0;
    return LBM_OK;
}

extern "C" int lbm_mlp_get_params(lbm_mlp* m, float* params)
{
    if (!m || !params) return mfail(m, LBM_EINVAL, "lbm_mlp_get_params: null argument");
    MLP_TRY(m, hipMemcpyAsync(params, m->params, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return LBM_OK;
}

extern "C" int lbm_mlp_set_norm(lbm_mlp* m, const float* mu, const float* sd, const float* lo, const float* hi,
                                float nu_scale)
{
    if (!m || !mu || !sd || !lo || !hi) return mfail(m, LBM_EINVAL, "lbm_mlp_set_norm: null argument");
    if (!(nu_scale > 0.f)) return mfail(m, LBM_EINVAL, "lbm_mlp_set_norm: nu_scale must be > 0");
    for (int k = 0; k < MLP_IN; ++k) {
        if (!(sd[k] > 0.f)) return mfail(m, LBM_EINVAL, "lbm_mlp_set_norm: sd must be > 0");
        m->nrm.mu[k] = mu[k];
        m->nrm.inv_sd[k] = 1.f / sd[k];
        m->nrm.lo[k] = lo[k];
        m->nrm.hi[k] = hi[k];
    }
    m->nrm.nu_scale = nu_scale;
    return LBM_OK;
}

extern "C" int lbm_mlp_set_data(lbm_mlp* m, const float* X, const float* y, long long n)
{
    if (!m || !X || !y || n <= 0) return mfail(m, LBM_EINVAL, "lbm_mlp_set_data: bad argument");
    if (m->X) (void)hipFree(m->X);
    if (m->y) (void)hipFree(m->y);
    m->X = m->y = nullptr;
    MLP_TRY(m, hipMalloc(&m->X, (size_t)n * MLP_IN * 4));
    MLP_TRY(m, hipMalloc(&m->y, (size_t)n * 4));
    MLP_TRY(m, hipMemcpyAsync(m->X, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(m, hipMemcpyAsync(m->y, y, (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    m->n = n;
    return LBM_OK;
}

static int mlp_grad_launch(lbm_mlp* m, long long offset, int count)
{
    const int tiles = (count + 63) / 64;
    if (tiles > m->max_tiles) {
        if (m->partials) (void)hipFree(m->partials);
        if (m->lossp) (void)hipFree(m->lossp);
        m->partials = nullptr;
        m->lossp = nullptr;
        MLP_TRY(m, hipMalloc(&m->partials, (size_t)tiles * MLP_NPARAM * 4));
        MLP_TRY(m, hipMalloc(&m->lossp, (size_t)tiles * 8));
        m->max_tiles = tiles;
    }
    hipLaunchKernelGGL(k_mlp_grad, dim3(tiles), dim3(64), 0, m->stream, m->X, m->y, (long long)count, offset,
                       m->params, m->wbf, m->nrm, 1.f / (float)count, m->partials, m->lossp);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->partials, tiles,
                       m->grad, m->lossp, m->dloss);
    MLP_TRY(m, hipGetLastError());
    return LBM_OK;
}

extern "C" int lbm_mlp_gradient(lbm_mlp* m, long long offset, int count, float* grad, double* loss)
{
    if (!m || !m->X) return mfail(m, LBM_ESTATE, "lbm_mlp_gradient: no data set");
    if (offset < 0 || count <= 0 || offset + count > m->n) return mfail(m, LBM_EINVAL, "lbm_mlp_gradient: bad range");
    int rc = mlp_grad_launch(m, offset, count);
    if (rc) return rc;
    if (grad) MLP_TRY(m, hipMemcpyAsync(grad, m->grad, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
    double l = 0;
    MLP_TRY(m, hipMemcpyAsync(&l, m->dloss, 8, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    if (loss) *loss = l / count;
    return LBM_OK;
}

// Adam on mean-squared error over consecutive batches of the data set
// (cycling); loss_hist[i] receives the batch loss of step i (FP64 sums).
extern "C" int lbm_mlp_train(lbm_mlp* m, int steps, int batch, float lr, double* loss_hist)
{
    if (!m || !m->X) return mfail(m, LBM_ESTATE, "lbm_mlp_train: no data set");
    if (steps < 0 || batch <= 0 || batch > m->n || !(lr > 0.f)) return mfail(m, LBM_EINVAL, "lbm_mlp_train: bad argument");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const long long nb = m->n / batch;
    for (int s = 0; s < steps; ++s) {
        const long long offset = (m->t % nb) * batch;
        int rc = mlp_grad_launch(m, offset, batch);
        if (rc) return rc;
        m->t += 1;
        const float bc1 = 1.f - std::pow(b1, (float)m->t), bc2 = 1.f - std::pow(b2, (float)m->t);
        hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->params, m->m, m->v,
                           m->grad, lr, b1, b2, eps, bc1, bc2, m->wbf);
        if (loss_hist) {
            double l = 0;
            MLP_TRY(m, hipMemcpyAsync(&l, m->dloss, 8, hipMemcpyDeviceToHost, m->stream));
            MLP_TRY(m, hipStreamSynchronize(m->stream));
            loss_hist[s] = l / batch;
        }
    }
    hipLaunchKernelGGL(k_mlp_pack_chain, dim3(1), dim3(64), 0, m->stream, m->params, m->chain);
    MLP_TRY(m, hipGetLastError());
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return LBM_OK;
}

extern "C" int lbm_mlp_forward(lbm_mlp* m, const float* X, long long n, float* nu)
{
    if (!m || !X || !nu || n <= 0) return mfail(m, LBM_EINVAL, "lbm_mlp_forward: bad argument");
    float *dx = nullptr, *dn = nullptr;
    MLP_TRY(m, hipMalloc(&dx, (size_t)n * MLP_IN * 4));
    MLP_TRY(m, hipMalloc(&dn, (size_t)n * 4));
    hipError_t e = hipMemcpyAsync(dx, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, dx, n,
                           m->chain, m->nrm, dn);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(nu, dn, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    (void)hipFree(dn);
    if (e != hipSuccess) return mfail(m, LBM_EDEVICE, std::string("lbm_mlp_forward: ") + hipGetErrorString(e));
    return LBM_OK;
}

extern "C" int lbm_set_closure(lbm_ctx* ctx, lbm_mlp* m, float nu_cap)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_set_closure: null context");
    if (!m) { ctx->mlp = nullptr; return LBM_OK; }
    if (ctx->cfg.model != LBM_MODEL_MRT_LES && ctx->cfg.model != LBM_MODEL_MRT_DHUMIERES)
        return fail(ctx, LBM_EINVAL, "lbm_set_closure: the closure replaces the LES rate of an MRT-LES model");
    if (ctx->nlevels > 1) return fail(ctx, LBM_EINVAL, "lbm_set_closure: not combined with refinement");
    if (!ctx->geometry_set) {
        int rc = lbm_set_geometry(ctx, nullptr);
        if (rc) return rc;
    }
    Level& L = ctx->lv[0];
    const std::vector<float> d = wall_distance(L, ctx->cfg.bc);
    std::vector<float> pad((size_t)L.qs, 1.f);
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x)
                pad[(size_t)x + (size_t)L.nxp * (y + (size_t)L.ny * z)] = d[(size_t)x + (size_t)L.nx * (y + (size_t)L.ny * z)];
    if (!ctx->dwall) HIP_TRY(ctx, hipMalloc(&ctx->dwall, pad.size() * 4));
    HIP_TRY(ctx, hipMemcpy(ctx->dwall, pad.data(), pad.size() * 4, hipMemcpyHostToDevice));
    // the per-node rate field: the rate each node was last relaxed at (the
    // closure's strain estimate) and the held rates between refreshes
    if (!ctx->omega_field) HIP_TRY(ctx, hipMalloc(&ctx->omega_field, (size_t)L.qs * sizeof(float)));
    HIP_TRY(ctx, hipMemset(ctx->omega_field, 0, (size_t)L.qs * sizeof(float)));
    ctx->mlp = m;
    ctx->nu_cap = nu_cap > 0.f ? nu_cap : 0.05f;
    ctx->closure_t0 = ctx->t;
    return LBM_OK;
}

// Evaluate the network every `every` steps; the steps in between collide at
// the per-node rates of the last evaluation (Smagorinsky where the inputs
// were outside the training range).
extern "C" int lbm_set_closure_interval(lbm_ctx* ctx, int every)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_set_closure_interval: null context");
    if (every < 1) return fail(ctx, LBM_EINVAL, "lbm_set_closure_interval: every must be >= 1");
    if (!ctx->omega_field) {
        Level& L = ctx->lv[0];
        HIP_TRY(ctx, hipMalloc(&ctx->omega_field, (size_t)L.qs * sizeof(float)));
        HIP_TRY(ctx, hipMemset(ctx->omega_field, 0, (size_t)L.qs * sizeof(float)));
    }
    ctx->closure_every = every;
    ctx->closure_t0 = ctx->t;
    return LBM_OK;
}

// Where the network's eddy viscosity goes: on top of tau0 (the plain
// learned closure) or on top of the Smagorinsky relaxation time of the
// configured CS (a learned correction to Smagorinsky).
extern "C" int lbm_set_closure_target(lbm_ctx* ctx, int with_model_stress)
{
    if (!ctx || (with_model_stress != 0 && with_model_stress != 1))
        return fail(ctx, LBM_EINVAL, "lbm_set_closure_target: bad argument");
    ctx->closure_target = with_model_stress;
    return LBM_OK;
}

extern "C" int lbm_set_closure_base(lbm_ctx* ctx, int smagorinsky)
{
    if (!ctx) return fail(nullptr, LBM_EINVAL, "lbm_set_closure_base: null context");
    if (smagorinsky != 0 && smagorinsky != 1) return fail(ctx, LBM_EINVAL, "lbm_set_closure_base: 0 or 1");
    ctx->closure_smag_base = smagorinsky;
    return LBM_OK;
}

// Training samples from this context's current state: box filter by 2,
// features and SGS eddy viscosity at a hashed selection (fraction `keep`) of
// the coarse cells, written in cell order.
extern "C" int lbm_closure_samples(lbm_ctx* ctx, uint64_t seed, float keep, float* X, float* y, int max_samples,
                                   int* count)
{
    if (!ctx || !X || !y || !count || max_samples <= 0) return fail(ctx, LBM_EINVAL, "lbm_closure_samples: bad argument");
    if (ctx->nlevels > 1) return fail(ctx, LBM_EINVAL, "lbm_closure_samples: single-level runs only");
    if (!ctx->geometry_set) {
        int rc = lbm_set_geometry(ctx, nullptr);
        if (rc) return rc;
    }
    Level& L = ctx->lv[0];
    if (L.nx % 2 || L.ny % 2 || L.nz % 2) return fail(ctx, LBM_EINVAL, "lbm_closure_samples: even domain needed");
    const int cx = L.nx / 2, cy = L.ny / 2, cz = L.nz / 2;
    const size_t NC = (size_t)cx * cy * cz;
    // coarse wall distance: the fine distance at the coarse cell centre, in coarse cells
    const std::vector<float> dfine = wall_distance(L, ctx->cfg.bc);
    std::vector<float> dc(NC);
    for (int k = 0; k < cz; ++k)
        for (int j = 0; j < cy; ++j)
            for (int i = 0; i < cx; ++i) {
                float s = 0.f;
                for (int d = 0; d < 8; ++d)
                    s += dfine[(size_t)(2 * i + (d & 1)) + (size_t)L.nx * ((2 * j + ((d >> 1) & 1)) + (size_t)L.ny * (2 * k + ((d >> 2) & 1)))];
                dc[(size_t)i + (size_t)cx * (j + (size_t)cy * k)] = 0.125f * s * 0.5f;
            }
    float *bx = nullptr, *dd = nullptr, *dX = nullptr, *dy = nullptr;
    unsigned int* di = nullptr;
    unsigned long long* dcount = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)bx, (void*)dd, (void*)dX, (void*)dy, (void*)di, (void*)dcount})
            if (q) (void)hipFree(q);
    };
    const int model = ctx->closure_target && ctx->cfg.cs > 0.f;
    // the post-collision stress is (1 - omega) times the one that set a
    // node's rate: near tau = 1 it no longer shows the rate
    if (model && std::fabs(1.f / ctx->cfg.omega - 1.f) < 0.1f)
        return fail(ctx, LBM_EINVAL, "lbm_closure_samples: the model-stress target needs |tau - 1| >= 0.1");
    if (hipMalloc(&bx, NC * (model ? 16 : 10) * 4) || hipMalloc(&dd, NC * 4) || hipMalloc(&dX, (size_t)max_samples * MLP_IN * 4) ||
        hipMalloc(&dy, (size_t)max_samples * 4) || hipMalloc(&di, (size_t)max_samples * 4) || hipMalloc(&dcount, 8)) {
        cleanup();
        return fail(ctx, LBM_ENOMEM, "lbm_closure_samples: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(dd, dc.data(), NC * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(dcount, 0, 8, ctx->stream);
    if (e == hipSuccess) {
        with_storage(L.storage, [&](auto* tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_box_filter<S>, dim3((unsigned)((NC + 255) / 256)), dim3(256), 0, ctx->stream,
                               (const S*)L.fa, L.qs, L.nxp, L.nx, L.ny, L.nz, L.mask, bx, L.shift, ctx->cfg.omega,
                               ctx->cfg.cs, model);
            return 0;
        });
        const uint32_t kp = (uint32_t)std::min(4294967295.0, std::max(0.0, (double)keep * 4294967296.0));
        hipLaunchKernelGGL(k_closure_features, dim3((unsigned)((NC + 255) / 256)), dim3(256), 0, ctx->stream, bx, cx, cy,
                           cz, dd, seed, kp, dX, dy, di, dcount, (unsigned long long)max_samples, model);
        e = hipGetLastError();
    }
    unsigned long long n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, dcount, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && n > (unsigned long long)max_samples) {
        // the kernel kept the first max_samples candidates in atomic order,
        // which is not deterministic: refuse, and report how many are needed
        cleanup();
        *count = (int)std::min<unsigned long long>(n, INT_MAX);
        return fail(ctx, LBM_EINVAL, "lbm_closure_samples: " + std::to_string(n) + " samples qualify, more than max_samples " +
                                         std::to_string(max_samples) + " (*count holds the number needed)");
    }
    const int got = (int)std::min<unsigned long long>(n, (unsigned long long)max_samples);
    std::vector<float> hx((size_t)got * MLP_IN), hy(got);
    std::vector<unsigned int> hi(got);
    if (e == hipSuccess && got > 0) {
        e = hipMemcpy(hx.data(), dX, hx.size() * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hy.data(), dy, hy.size() * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hi.data(), di, hi.size() * 4, hipMemcpyDeviceToHost);
    }
    cleanup();
    if (e != hipSuccess) return fail(ctx, LBM_EDEVICE, std::string("lbm_closure_samples: ") + hipGetErrorString(e));
    std::vector<int> order(got);
    for (int i = 0; i < got; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return hi[a] < hi[b]; });
    for (int i = 0; i < got; ++i) {
        for (int k = 0; k < MLP_IN; ++k) X[(size_t)i * MLP_IN + k] = hx[(size_t)order[i] * MLP_IN + k];
        y[i] = hy[order[i]];
    }
    *count = got;
    return LBM_OK;
}
