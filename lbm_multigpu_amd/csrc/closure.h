// closure.h -- learned eddy-viscosity closure: training kernels (forward,
// backward, deterministic gradient reduction, Adam), the box-filter sampler
// that builds training data from a finer run, and the stream-collide kernel
// with the closure's inference fused into the collide.
//
// Inputs (7, raw, normalised by the stored mean/std before the network):
//   x0 = log |S|          magnitude of the strain rate
//   x1 = 3 sqrt(6) det(S/|S|)   shape (in [-1, 1])
//   x2..x4 = S_xy, S_xz, S_yz over |S|   normalised shear components
//   x5 = log d_wall       distance to the nearest wall or solid node (cells)
//   x6 = |u|              local speed
// S is the strain rate in lattice units of the grid the model runs on: in the
// solver S = -3 omega0/(2 rho) Pi_neq with Pi_neq = P - rho/3 I - j j (the
// second moment of f - feq that the Smagorinsky rate uses,
// MRTtest/DynInlet3.cu:722-729); in the sampler, central differences of the
// box-filtered velocity.  Output: nu_t = nu_scale * softplus(z); the collide
// then uses tau = 1/omega0 + 3 nu_t.  Outside the training range of any
// input the node falls back to the Smagorinsky rate.
#pragma once

#include "kernels.h"
#include "mlp.h"

namespace lbm {

__device__ __forceinline__ void strain_features(float sxx, float syy, float szz, float sxy, float sxz, float syz,
                                                float dwall, float speed, float x[MLP_IN])
{
    const float tr = (sxx + syy + szz) * (1.f / 3.f);
    const float a = sxx - tr, b = syy - tr, c = szz - tr;
    const float n2 = a * a + b * b + c * c + 2.f * (sxy * sxy + sxz * sxz + syz * syz);
    const float n = sqrtf(fmaxf(n2, 1e-30f));
    const float inv = 1.f / n;
    const float A = a * inv, B = b * inv, Cc = c * inv, D = sxy * inv, E = sxz * inv, F = syz * inv;
    const float det = A * (B * Cc - F * F) - D * (D * Cc - F * E) + E * (D * F - B * E);
    x[0] = logf(n + 1e-12f);
    x[1] = fminf(fmaxf(7.348469228f * det, -1.f), 1.f);
    x[2] = D;
    x[3] = E;
    x[4] = F;
    x[5] = logf(fmaxf(dwall, 0.5f));
    x[6] = speed;
}

// ------------------------------------------------------------ training

// One wavefront per workgroup, 64 samples: forward (z stored in BF16), loss
// (softplus(zo) - y/nu_scale)^2 (the target in units of nu_scale, which keeps
// the gradients O(1) for Adam), backward through the three layers, per-tile
// gradients to partials[tile][MLP_NPARAM] and the FP64 loss to lossp[tile].
__global__ void __launch_bounds__(64) k_mlp_grad(const float* X, const float* y, long long n, long long offset,
                                                 const float* params, const short* wbf, MlpNorm nrm, float inv_n,
                                                 float* partials, double* lossp)
{
    __shared__ MlpLds s;
    __shared__ short xb[64 * MLP_INP];
    __shared__ short z1[64 * MLP_H], z2[64 * MLP_H], z3[64 * MLP_H];
    __shared__ short da[64 * MLP_H], db[64 * MLP_H];
    __shared__ float dzo[64];
    const int lane = threadIdx.x;
    float* part = partials + (long long)blockIdx.x * MLP_NPARAM;
    mlp_stage(s, wbf, params);
    const long long sidx = offset + (long long)blockIdx.x * 64 + lane;
    const bool valid = sidx < offset + n;
#pragma unroll
    for (int k = 0; k < MLP_INP; ++k) {
        float v = 0.f;
        if (valid && k < MLP_IN) v = (X[sidx * MLP_IN + k] - nrm.mu[k]) * nrm.inv_sd[k];
        xb[lane * MLP_INP + k] = bf(v);
    }
    __syncthreads();
    wave_gemm<64, 64, 16>([&](int m, int k) { return xb[m * MLP_INP + k]; },
                          [&](int k, int c) { return s.w1[k * MLP_H + c]; },
                          [&](int m, int c, float v) { z1[m * MLP_H + c] = bf(v + s.b1[c]); });
    __syncthreads();
    wave_gemm<64, 64, 64>([&](int m, int k) { return bf(gelu(fb(z1[m * MLP_H + k]))); },
                          [&](int k, int c) { return s.w2[k * MLP_H + c]; },
                          [&](int m, int c, float v) { z2[m * MLP_H + c] = bf(v + s.b2[c]); });
    __syncthreads();
    wave_gemm<64, 64, 64>([&](int m, int k) { return bf(gelu(fb(z2[m * MLP_H + k]))); },
                          [&](int k, int c) { return s.w3[k * MLP_H + c]; },
                          [&](int m, int c, float v) { z3[m * MLP_H + c] = bf(v + s.b3[c]); });
    __syncthreads();
    float zo = s.b4;
    for (int i = 0; i < MLP_H; ++i) zo += fb(bf(gelu(fb(z3[lane * MLP_H + i])))) * s.w4[i];
    const float r = valid ? softplus(zo) - y[sidx] / nrm.nu_scale : 0.f;
    const double loss = valid ? (double)r * (double)r : 0.0;
    const float g = valid ? 2.f * r * sigmoid(zo) * inv_n : 0.f;
    dzo[lane] = g;
    __syncthreads();
    // output layer: dw4[i] = sum_s h3[s][i] dzo[s], db4 = sum_s dzo[s]
    {
        float acc = 0.f;
        for (int t = 0; t < 64; ++t) acc += fb(bf(gelu(fb(z3[t * MLP_H + lane])))) * dzo[t];
        part[OFF_W4 + lane] = acc;
        const float sb = wave_sum(g);
        if (lane == 0) part[OFF_B4] = sb;
    }
    // dZ3[s][i] = dzo[s] w4[i] GELU'(z3[s][i])
    for (int i = 0; i < MLP_H; ++i) da[lane * MLP_H + i] = bf(g * s.w4[i] * gelu_grad(fb(z3[lane * MLP_H + i])));
    __syncthreads();
    // layer 3: dW3 = h2^T dZ3, db3, dZ2 = (dZ3 W3^T) * GELU'(z2)
    wave_gemm<64, 64, 64>([&](int m, int k) { return bf(gelu(fb(z2[k * MLP_H + m]))); },
                          [&](int k, int c) { return da[k * MLP_H + c]; },
                          [&](int m, int c, float v) { part[OFF_W3 + m * MLP_H + c] = v; });
    {
        float acc = 0.f;
        for (int t = 0; t < 64; ++t) acc += fb(da[t * MLP_H + lane]);
        part[OFF_B3 + lane] = acc;
    }
    wave_gemm<64, 64, 64>([&](int m, int k) { return da[m * MLP_H + k]; },
                          [&](int k, int c) { return s.w3[c * MLP_H + k]; },
                          [&](int m, int c, float v) { db[m * MLP_H + c] = bf(v * gelu_grad(fb(z2[m * MLP_H + c]))); });
    __syncthreads();
    // layer 2
    wave_gemm<64, 64, 64>([&](int m, int k) { return bf(gelu(fb(z1[k * MLP_H + m]))); },
                          [&](int k, int c) { return db[k * MLP_H + c]; },
                          [&](int m, int c, float v) { part[OFF_W2 + m * MLP_H + c] = v; });
    {
        float acc = 0.f;
        for (int t = 0; t < 64; ++t) acc += fb(db[t * MLP_H + lane]);
        part[OFF_B2 + lane] = acc;
    }
    wave_gemm<64, 64, 64>([&](int m, int k) { return db[m * MLP_H + k]; },
                          [&](int k, int c) { return s.w2[c * MLP_H + k]; },
                          [&](int m, int c, float v) { da[m * MLP_H + c] = bf(v * gelu_grad(fb(z1[m * MLP_H + c]))); });
    __syncthreads();
    // layer 1: dW1 = x^T dZ1 (rows >= 7 are the padding), db1
    wave_gemm<16, 64, 64>([&](int m, int k) { return xb[k * MLP_INP + m]; },
                          [&](int k, int c) { return da[k * MLP_H + c]; },
                          [&](int m, int c, float v) {
                              if (m < MLP_IN) part[OFF_W1 + m * MLP_H + c] = v;
                          });
    {
        float acc = 0.f;
        for (int t = 0; t < 64; ++t) acc += fb(da[t * MLP_H + lane]);
        part[OFF_B1 + lane] = acc;
    }
    const double ls = wave_sum(loss);
    if (lane == 0) lossp[blockIdx.x] = ls;
}

// Deterministic sum of the per-tile gradients (fixed order per parameter).
__global__ void k_mlp_reduce(const float* partials, int ntiles, float* grad, const double* lossp, double* loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) {
        float s = 0.f;
        for (int t = 0; t < ntiles; ++t) s += partials[(long long)t * MLP_NPARAM + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        double s = 0.0;
        for (int t = 0; t < ntiles; ++t) s += lossp[t];
        *loss = s;
    }
}

// Adam on the FP32 master parameters, then the BF16 copies used by the
// matrix cores (W1 padded to [16][64]).
__global__ void k_mlp_adam(float* params, float* m, float* v, const float* grad, float lr, float b1, float b2,
                           float eps, float bc1, float bc2, short* wbf)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    params[p] = w;
    if (p < OFF_B1) wbf[p] = bf(w); // W1 rows 0..6 of the padded [16][64]
    else if (p >= OFF_W2 && p < OFF_B2) wbf[MLP_INP * MLP_H + (p - OFF_W2)] = bf(w);
    else if (p >= OFF_W3 && p < OFF_B3) wbf[MLP_INP * MLP_H + MLP_H * MLP_H + (p - OFF_W3)] = bf(w);
}

__global__ void k_mlp_sync_bf16(const float* params, short* wbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = MLP_INP * MLP_H;
    if (i < n1) {
        const int k = i / MLP_H;
        wbf[i] = k < MLP_IN ? bf(params[OFF_W1 + i]) : (short)0;
    } else if (i < n1 + MLP_H * MLP_H) {
        wbf[i] = bf(params[OFF_W2 + (i - n1)]);
    } else if (i < n1 + 2 * MLP_H * MLP_H) {
        wbf[i] = bf(params[OFF_W3 + (i - n1 - MLP_H * MLP_H)]);
    }
}

// Inference on raw inputs X[n][7] -> nu[n] (256 threads: 4 wavefronts of 64),
// the same register chain the fused collide runs (launch bound as there).
__global__ void __launch_bounds__(256, 2) k_mlp_forward(const float* X, long long n, const MlpChain* chain,
                                                     MlpNorm nrm, float* nu)
{
    __shared__ MlpChain s;
    chain_stage(s, chain);
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = i < n ? (X[i * MLP_IN + k] - nrm.mu[k]) * nrm.inv_sd[k] : 0.f;
    __syncthreads();
    const float zo = mlp_chain_infer(s, mlp_pack_inputs(x));
    if (i < n) nu[i] = nrm.nu_scale * softplus(zo);
}

// ------------------------------------------------------------ sampler

// Box filter of the fine run by 2: velocity, products and solid fraction of
// the 8 fine nodes under coarse cell (X,Y,Z).  Output planes (coarse dense):
// 0..2 u, 3..8 uu uv uw vv vw ww, 9 solid fraction; with model = 1 also
// 10..15 nu_f S_f (xx yy zz xy xz yz; minus half the fine run's own
// Smagorinsky stress), at each fine node from its stored post-collision
// populations:
// Pi' = P - rho/3 I - j j is (1 - omega) times the pre-collision
// non-equilibrium stress Pi that set the node's rate (les_omega: tau =
// tau0/2 + sqrt(tau0^2 + 18 sqrt2 CS |Pi|)/2), so tau solves
// tau = tau0/2 + sqrt(tau0^2 + 18 sqrt2 CS |Pi'| tau / |tau - 1|)/2 (a few
// fixed-point steps from tau0; the trace of Pi', which relaxes at another
// rate, is left out), nu_f = (tau - tau0)/3 and S_f = -3 Pi'_dev /
// (2 rho (tau - 1)) (S = -3 Pi / (2 rho tau)).
template <class S>
__global__ void k_box_filter(const S* f, long long qs, int nxp, int nx, int ny, int nz, const unsigned char* mask,
                             float* out, PopShift sh, float omega0, float cs, int model)
{
    using R = typename Arith<S>::R;
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const long long NC = (long long)cx * cy * cz;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= NC) return;
    const int X = int(c % cx), Y = int((c / cx) % cy), Z = int(c / ((long long)cx * cy));
    float acc[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float tau0 = 1.f / omega0;
    for (int d = 0; d < 8; ++d) {
        const int x = 2 * X + (d & 1), y = 2 * Y + ((d >> 1) & 1), z = 2 * Z + ((d >> 2) & 1);
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        if (is_solid(mask[cell])) { acc[9] += 1.f; continue; }
        R g[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) g[i] = Store<S>::to_real(f[i * qs + cell], i, sh);
        R sp[NP], dp[NP];
        pair_split(g, sp, dp);
        const Macro<R> m = macro_from_pairs(g[0], sp, dp);
        const float U = float(m.jx), V = float(m.jy), W = float(m.jz);
        acc[0] += U; acc[1] += V; acc[2] += W;
        acc[3] += U * U; acc[4] += U * V; acc[5] += U * W; acc[6] += V * V; acc[7] += V * W; acc[8] += W * W;
        if (model && cs > 0.f) {
            const float third = float(m.rho) * (1.f / 3.f);
            float pxx = float(m.pxx) - third - U * U, pyy = float(m.pyy) - third - V * V;
            float pzz = float(m.pzz) - third - W * W;
            const float pxy = float(m.pxy) - U * V, pxz = float(m.pxz) - U * W, pyz = float(m.pyz) - V * W;
            const float tr = (pxx + pyy + pzz) * (1.f / 3.f);
            pxx -= tr; pyy -= tr; pzz -= tr;
            const float qpost = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.f * (pxy * pxy + pxz * pxz + pyz * pyz));
            const float k = float(18.0 * 1.4142135623730951) * cs * qpost;
            float tau = tau0;
            for (int it = 0; it < 8; ++it)
                tau = 0.5f * tau0 + 0.5f * sqrtf(tau0 * tau0 + k * tau / fmaxf(fabsf(tau - 1.f), 0.05f));
            const float nuf = (tau - tau0) * (1.f / 3.f);
            const float sc = nuf * -1.5f / (float(m.rho) * (tau - 1.f)); // nu_f S_f = sc Pi'_dev
            acc[10] += sc * pxx; acc[11] += sc * pyy; acc[12] += sc * pzz;
            acc[13] += sc * pxy; acc[14] += sc * pxz; acc[15] += sc * pyz;
        }
    }
    const int nout = model ? 16 : 10;
    for (int k = 0; k < nout; ++k) out[k * NC + c] = acc[k] * 0.125f;
}

// Features and target at coarse cell c from the box-filtered fields:
// S from central differences of u (coarse cells), tau_ij = <u_i u_j> - <u_i><u_j>,
// nu_t = max(0, -tau_dev:S / (2 S:S)).  With model = 1 the fine run's own
// model stress, filtered, joins the subfilter stress the coarse closure has
// to carry: nu_t = max(0, (-tau_dev:S + 2 <nu_f S_f>:S) / (2 S:S)) (the
// filtered fine equations hold div(2 <nu_f S_f>) beside div(tau)).  Cells
// touching a solid are skipped; selected cells are those whose hash(seed, c)
// falls under `keep`.
__global__ void k_closure_features(const float* bx, int cx, int cy, int cz, const float* dwall, uint64_t seed,
                                   uint32_t keep, float* X, float* y, unsigned int* cidx, unsigned long long* count,
                                   unsigned long long cap, int model)
{
    const long long NC = (long long)cx * cy * cz;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= NC) return;
    const int i = int(c % cx), j = int((c / cx) % cy), k = int(c / ((long long)cx * cy));
    if (i < 1 || j < 1 || k < 1 || i > cx - 2 || j > cy - 2 || k > cz - 2) return;
    // hash select (splitmix64 over (seed, c))
    uint64_t h = seed ^ ((uint64_t)c * 0x9E3779B97F4A7C15ULL);
    h += 0x632BE59BD9B4E019ULL;
    h ^= h >> 30; h *= 0xBF58476D1CE4E5B9ULL; h ^= h >> 27; h *= 0x94D049BB133111EBULL; h ^= h >> 31;
    if ((uint32_t)(h >> 32) >= keep) return;
    auto at3 = [&](int q, int a, int b, int d) { return bx[q * NC + a + (long long)cx * (b + (long long)cy * d)]; };
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk)
                if (at3(9, i + di, j + dj, k + dk) > 0.f) return;
    float g[3][3];
    for (int q = 0; q < 3; ++q) {
        g[q][0] = 0.5f * (at3(q, i + 1, j, k) - at3(q, i - 1, j, k));
        g[q][1] = 0.5f * (at3(q, i, j + 1, k) - at3(q, i, j - 1, k));
        g[q][2] = 0.5f * (at3(q, i, j, k + 1) - at3(q, i, j, k - 1));
    }
    const float sxx = g[0][0], syy = g[1][1], szz = g[2][2];
    const float sxy = 0.5f * (g[0][1] + g[1][0]), sxz = 0.5f * (g[0][2] + g[2][0]), syz = 0.5f * (g[1][2] + g[2][1]);
    const float u = at3(0, i, j, k), v = at3(1, i, j, k), w = at3(2, i, j, k);
    float t[6] = {at3(3, i, j, k) - u * u, at3(4, i, j, k) - u * v, at3(5, i, j, k) - u * w,
                  at3(6, i, j, k) - v * v, at3(7, i, j, k) - v * w, at3(8, i, j, k) - w * w};
    const float tr = (t[0] + t[3] + t[5]) * (1.f / 3.f);
    t[0] -= tr; t[3] -= tr; t[5] -= tr;
    const float ss = sxx * sxx + syy * syy + szz * szz + 2.f * (sxy * sxy + sxz * sxz + syz * syz);
    const float ts = t[0] * sxx + t[3] * syy + t[5] * szz + 2.f * (t[1] * sxy + t[2] * sxz + t[4] * syz);
    float ms = 0.f;
    if (model)
        ms = at3(10, i, j, k) * sxx + at3(11, i, j, k) * syy + at3(12, i, j, k) * szz +
             2.f * (at3(13, i, j, k) * sxy + at3(14, i, j, k) * sxz + at3(15, i, j, k) * syz);
    const float nut = ss > 1e-20f ? fmaxf(0.f, (2.f * ms - ts) / (2.f * ss)) : 0.f;
    const unsigned long long slot = atomicAdd(count, 1ULL);
    if (slot >= cap) return;
    float x[MLP_IN];
    strain_features(sxx, syy, szz, sxy, sxz, syz, dwall[c], sqrtf(u * u + v * v + w * w), x);
#pragma unroll
    for (int q = 0; q < MLP_IN; ++q) X[slot * MLP_IN + q] = x[q];
    y[slot] = nut;
    cidx[slot] = (unsigned int)c; // slots come from an atomic: the host sorts by cell
}

// ------------------------------------------------------------ fused closure

struct ClosureArgs {
    const MlpChain* chain; // weights in the register-chain arrangement
    const float* dwall;    // per node (padded layout)
    MlpNorm nrm;
    float nu_cap;       // upper bound on nu_t (stability)
    int smag_base;      // 1: nu_t adds to the Smagorinsky relaxation time, 0: to tau0
    // per-node rate field: read first, as the rate the node was last relaxed
    // at (|value|; 0: none yet, the base rate), then overwritten with this
    // step's rate -- the network's (> 0, held by the steps until the next
    // refresh) or the Smagorinsky rate negated (< 0: those steps recompute it)
    float* omega_out;
};

// Stream-collide with the learned closure: pull and boundary conditions per
// node, features from the pre-collision moments, the MLP for the 64 nodes of
// each wavefront on the matrix cores (register chain; all lanes take part,
// nodes that are not collided feed dummy inputs), then the MRT collide at
// tau = tau0 + 3 nu_t, or the Smagorinsky rate where an input leaves the
// training range.
// Four wavefronts per workgroup for every storage type: the 19.5 KB of
// staged weights are shared by all of them, and the LDS is what bounds the
// kernel's occupancy.  The launch bound of at least two waves per SIMD caps
// the registers at 256, so the compiler keeps the matrix-core accumulators
// in ordinary VGPRs (no accumulator-register copy per hidden unit).
constexpr int CLOSURE_NW = 4;
template <int MODEL, class S>
__global__ void __launch_bounds__(64 * CLOSURE_NW, 2) k_stream_collide_closure(StepParams p, ClosureArgs ca)
{
    using R = typename Arith<S>::R;
    __shared__ MlpChain s;
    constexpr int NW = CLOSURE_NW;
    chain_stage(s, ca.chain);
    const int lb = logical_block();
    int x, y, z;
    const bool row_ok = wave_node_map<NW>(lb, p.map, p.nz, x, y, z);
    const bool active = row_ok && x < p.nx;
    const int cell = active ? x + p.nxp * (y + p.ny * z) : 0;
    R fx = 0, fy = 0, fz = 0;
    R f[Q];
    int act = ACT_SKIP;
    if (active) act = node_pull<S>(p, x, y, z, cell, f, fx, fy, fz);
    float feat[MLP_IN] = {0, 0, 0, 0, 0, 0, 0};
    bool in_range = false;
    R omega_les = R(p.omega);
    if (act == ACT_COLLIDE) {
        R sp[NP], dp[NP];
        pair_split(f, sp, dp);
        const Macro<R> m = macro_from_pairs(f[0], sp, dp);
        const float rho = float(m.rho);
        const float third = rho * (1.f / 3.f);
        // S = -3 omega/(2 rho) Pi_neq with the rate that shaped Pi_neq: the
        // node's last applied one (tau0 + 3 nu_t), the base rate before any
        const float wlast = ca.omega_out ? fabsf(*at(ca.omega_out, cell)) : 0.f;
        const float k = -1.5f * (wlast > 0.f ? wlast : p.omega) / rho;
        const float sxx = k * (float(m.pxx) - third - float(m.jx * m.jx));
        const float syy = k * (float(m.pyy) - third - float(m.jy * m.jy));
        const float szz = k * (float(m.pzz) - third - float(m.jz * m.jz));
        const float sxy = k * (float(m.pxy) - float(m.jx * m.jy));
        const float sxz = k * (float(m.pxz) - float(m.jx * m.jz));
        const float syz = k * (float(m.pyz) - float(m.jy * m.jz));
        const float speed = sqrtf(float(m.jx * m.jx + m.jy * m.jy + m.jz * m.jz));
        strain_features(sxx, syy, szz, sxy, sxz, syz, ca.dwall[cell], speed, feat);
        in_range = true;
#pragma unroll
        for (int q = 0; q < MLP_IN; ++q) in_range = in_range && feat[q] >= ca.nrm.lo[q] && feat[q] <= ca.nrm.hi[q];
        if (p.cs > 0.f) omega_les = les_omega(m, R(p.omega), R(p.cs));
    }
    float xin[MLP_IN];
#pragma unroll
    for (int q = 0; q < MLP_IN; ++q) xin[q] = (feat[q] - ca.nrm.mu[q]) * ca.nrm.inv_sd[q];
    __syncthreads(); // the staged weights
    const float zo = mlp_chain_infer(s, mlp_pack_inputs(xin));
    if (act != ACT_SKIP) {
        if (act == ACT_COLLIDE) {
            CollideParams<R> cp = collide_params<R>(p, y);
            if (in_range) {
                const float nut = fminf(ca.nrm.nu_scale * softplus(zo), ca.nu_cap);
                const float wr = 1.f / ((ca.smag_base ? 1.f / float(omega_les) : 1.f / p.omega) + 3.f * nut);
                cp.omega = R(wr);
                if (ca.omega_out) *at(ca.omega_out, cell) = wr;
            } else {
                cp.omega = omega_les;
                if (ca.omega_out) *at(ca.omega_out, cell) = -float(omega_les);
            }
            cp.cs = R(0);
            collide<MODEL>(f, cp);
            if (p.stat_mode) stats_update(p, cell, f);
        } else if (act == ACT_COPY && p.stat_mode) {
            stats_update(p, cell, f);
        }
        node_store<S>(p, cell, f);
    }
    if (p.fpart) block_force_partial<NW>(p, lb, fx, fy, fz);
}

} // namespace lbm
