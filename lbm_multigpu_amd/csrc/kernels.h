// kernels.h -- HIP kernels of the D3Q19 stream-collide path for gfx950.
//
// One thread per node; each 64-lane wavefront owns 64 consecutive x of one
// (y, z) row and a workgroup takes consecutive row segments in memory order
// (wave_node; four for FP32, so each of the 19 population loads and stores
// of a workgroup is one contiguous 1 KB stretch of a plane); only the 10
// directions with c_x != 0 are offset by one element.  (A 64x4 shape -- four rows of 256 B per plane -- and an XCD-slab
// renumbering of the workgroups both measured 8-10 % below this on a
// 19-plane copy with the same pull pattern; the renumbering stays available
// as a knob.)
#pragma once

#include "collide.h"
#include "fastdiv.h"
#include "storage.h"

namespace lbm {

enum Bc : int { BC_PERIODIC = 0, BC_CLAMP = 1, BC_WALL = 2 };

// Tuning knobs (compile-time; the defaults are the measured best on MI355X)
#ifndef LBM_NWAVE
#define LBM_NWAVE 4 // wavefronts per workgroup, FP32 and FP64 populations
#endif
#ifndef LBM_NWAVE16
#define LBM_NWAVE16 1 // wavefronts per workgroup, 16-bit populations
#endif
#ifndef LBM_XCD_REMAP
#define LBM_XCD_REMAP 0
#endif
#ifndef LBM_XCD_CHUNK
#define LBM_XCD_CHUNK 4 // consecutive stream-collide workgroups per XCD, FP32/FP64
#endif
#ifndef LBM_XCD_CHUNK16
#define LBM_XCD_CHUNK16 7 // the same for the 16-bit formats
#endif
#ifndef LBM_MINW32
#define LBM_MINW32 1 // minimum wavefronts per SIMD for the FP32/FP64 stream-collide registers
#endif

// Workgroup shape per population type: NW wavefronts of 64 lanes on NW
// consecutive row segments, one node per lane.  FP32 streams best with four
// (one contiguous 1 KB stretch per plane and workgroup); the 16-bit
// kernels, with twice the vector work per byte, run best with one (measured on the 512x128x128 channel:
// 35.1 vs 32.2 GLUPS FP32, 46.2 vs 41.7 GLUPS FP16).  The 16-bit kernels
// fit 64 VGPRs without spilling, so they are held to 8 wavefronts per SIMD
// (56.3 vs 54.4 GLUPS at 7); the FP32 kernel at 8 would spill.
#ifndef LBM_MINW16
#define LBM_MINW16 8 // minimum wavefronts per SIMD for the 16-bit stream-collide registers
#endif
template <class S>
struct Shape {
    static constexpr int NW = sizeof(S) == 2 ? LBM_NWAVE16 : LBM_NWAVE;
    static constexpr int THREADS = 64 * NW;
    // minimum wavefronts per SIMD the compiler must fit the registers to
    static constexpr int MIN_WAVES = sizeof(S) == 2 ? LBM_MINW16 : (sizeof(S) == 4 ? LBM_MINW32 : 1);
};
__host__ inline bool is16_storage(int storage) { return storage == 1 || storage == 2; }
__host__ inline int nwave_of_storage(int storage) { return is16_storage(storage) ? LBM_NWAVE16 : LBM_NWAVE; }
// consecutive stream-collide workgroups per XCD (logical_block) on a level
// with nsegw 64-node segments per row: only rows longer than a workgroup's
// stretch have workgroup boundaries inside them
__host__ inline int xcd_chunk_of(int storage, int nsegw)
{
    return nsegw > nwave_of_storage(storage) ? (is16_storage(storage) ? LBM_XCD_CHUNK16 : LBM_XCD_CHUNK) : 1;
}

// wavefront index within the workgroup and lane within the wavefront
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Node of this lane: wavefronts enumerate the 64-node row segments in memory
// order (segment fastest, then y, then z) and workgroup lb takes NW
// consecutive ones, so every workgroup streams one contiguous stretch of
// each plane whatever nx is.  Returns false past the last row.
template <int NW>
__device__ __forceinline__ bool wave_node(int lb, int nsegw, int ny, int nz, int& x, int& y, int& z)
{
    // wave-uniform: the divisions run on the scalar unit, not per lane
    const int g = __builtin_amdgcn_readfirstlane(lb * NW + wave_id());
    const int row = g / nsegw;
    x = (g - row * nsegw) * 64 + lane_id();
    y = row % ny;
    z = row / ny;
    return z < nz;
}

// The stream-collide's wave-to-node map with its divisors precomputed:
// segment index g -> row = g / nsegw, z = row / ny.
struct NodeMap {
    FastDiv seg, ny;
    // XCD-chunked renumbering (logical_block_of): window of 8 * chunk ids,
    // and the ids below win_end, which are whole windows
    FastDiv win;
    int chunk, win_end;
};

__host__ inline NodeMap node_map_of(int nsegw, int ny, int chunk, int blocks)
{
    NodeMap m;
    m.seg = FastDiv::of((uint32_t)nsegw);
    m.ny = FastDiv::of((uint32_t)ny);
    m.chunk = chunk;
    m.win = FastDiv::of((uint32_t)(8 * chunk));
    m.win_end = chunk > 1 ? (blocks / (8 * chunk)) * (8 * chunk) : 0;
    return m;
}

struct StepParams {
    const void* fin;
    void* fout;
    const unsigned char* mask; // node codes, padded layout (every level has one; 0 = fluid)
    const float* u_in;         // west inlet profile [z*ny + y]
    double* fpart;             // momentum-exchange partials [blocks*3], or nullptr
    long long qs;              // elements per direction plane
    int nx, ny, nz, nxp;
    int nsegw;                 // 64-node wavefront segments per padded row (nxp / 64)
    int bcx, bcy, bcz;
    int halfway;
    float omega, cs, dpdy, u_lid;
    float s1, s2, s4, s10, s16;
    // running velocity statistics: 6 planes of qs floats (mean u v w, then
    // mean squared deviations); mode 0 off, 1 mean, 2 deviations
    float* stat;
    int stat_mode;
    float stat_n, stat_w; // samples so far (t - start) and this sample's weight
    // per-node collision rate left by the last learned-closure step (padded
    // layout): > 0 replaces the LES rate, <= 0 means the Smagorinsky rate
    const float* omega_node;
    PopShift sh;   // 16-bit storage offsets
    int dyn_y1;    // rows y > dyn_y1 get no body force (INT_MAX: all rows get it)
    const float* v_in;  // south inlet profile [z*nx + x] (NODE_INLET_SOUTH)
    const int* fslot;   // force partial slot of each workgroup (-1: no force node in it);
                        // nullptr: slot = workgroup
    int y_wrap;         // rows 0 .. y_wrap periodic in y among themselves (-1: no wrap)
    int ghost_y[3];     // source rows of the im 52 / 53 / 54 ghost copies
    // coarse level under refinement: nodes of the footprint box fp_a ..
    // fp_a + fp_n - 1 also write their post-collision physical moments and
    // CF data (fp_n[0] = 0: none; footprint_extract)
    int fp_a[3], fp_n[3];
    // (all in the interface's arithmetic type: double for FP64 storage, float otherwise)
    const void* fp_mold;  // moments at the step's start (temporal midpoint)
    void* fp_m;           // moments at the step's end [9][footprint]
    void* fp_pnew;        // CF data of the end state, 16 values per node
    void* fp_phalf;       // CF data of the midpoint, or nullptr
    int xcd_chunk; // consecutive workgroups per XCD (logical_block)
    int nt_store;  // nontemporal population stores (lattices larger than the MALL)
    int whole_ok;  // the input lattice's 19 planes span less than 4 GiB (node_load's WHOLE path)
    NodeMap map;   // the wave-to-node map's divisors (nsegw, ny, XCD window)
};

// XCD-aware renumbering.  The hardware id hw goes to XCD hw % 8.
//  - LBM_XCD_REMAP: XCD j takes the logical block range [j*G/8, (j+1)*G/8)
//    (bijective for any G);
//  - chunk = C > 1: within each window of 8C hardware ids, XCD j
//    takes the C consecutive logical blocks from j*C, so the 8 XCDs still
//    sweep the lattice together but each streams C workgroups' stretches
//    back to back (the x-shifted loads' straddling lines then stay in one
//    XCD's L2).  Ids past the last whole window map to themselves.
__device__ __forceinline__ int logical_block_of(int hw, int G, int chunk)
{
    if (LBM_XCD_REMAP) {
        const int per = G >> 3, rem = G & 7, xcd = hw & 7;
        return xcd * per + (xcd < rem ? xcd : rem) + (hw >> 3);
    }
    if (chunk > 1) { // wave-uniform: scalar arithmetic
        const int W = 8 * chunk;
        if (hw < (G / W) * W) {
            const int r = hw % W;
            return hw - r + (r & 7) * chunk + (r >> 3);
        }
    }
    return hw;
}
__device__ __forceinline__ int logical_block(int chunk = 1)
{
    return logical_block_of(blockIdx.x, gridDim.x, chunk);
}
// The same renumbering with the window's divisor precomputed (NodeMap).
__device__ __forceinline__ int logical_block_map(int hw, const NodeMap& m)
{
    if (LBM_XCD_REMAP) return logical_block_of(hw, (int)gridDim.x, m.chunk);
    if (m.chunk > 1 && hw < m.win_end) {
        const int W = 8 * m.chunk;
        const int r = hw - (int)m.win.div((uint32_t)hw) * W;
        return hw - r + (r & 7) * m.chunk + (r >> 3);
    }
    return hw;
}

// wave_node with the divisions by nsegw and ny done as FastDiv products.
template <int NW>
__device__ __forceinline__ bool wave_node_map(int lb, const NodeMap& m, int nz, int& x, int& y, int& z)
{
    const uint32_t g = (uint32_t)__builtin_amdgcn_readfirstlane(lb * NW + wave_id());
    const uint32_t row = m.seg.div(g);
    x = (int)(g - row * m.seg.d) * 64 + lane_id();
    const uint32_t zz = m.ny.div(row);
    y = (int)(row - zz * m.ny.d);
    z = (int)zz;
    return z < nz;
}

__device__ __forceinline__ void axis_nb(int c, int n, int bc, int& m, int& p, bool& om, bool& op)
{
    om = (c == 0);
    op = (c == n - 1);
    if (bc == BC_PERIODIC) {
        m = om ? n - 1 : c - 1;
        p = op ? 0 : c + 1;
        om = op = false;
    } else {
        m = om ? 0 : c - 1;
        p = op ? n - 1 : c + 1;
        if (bc != BC_WALL) om = op = false;
    }
}

// Element idx of a direction plane: plane base (wave-uniform, SGPRs) plus a
// 32-bit unsigned byte offset, so loads and stores use the SGPR-base +
// VGPR-offset addressing form instead of per-lane 64-bit address math.
template <class S>
__device__ __forceinline__ const S* at(const S* plane, int idx)
{
    return reinterpret_cast<const S*>(reinterpret_cast<const char*>(plane) + (uint32_t)idx * (uint32_t)sizeof(S));
}
template <class S>
__device__ __forceinline__ S* at(S* plane, int idx)
{
    return reinterpret_cast<S*>(reinterpret_cast<char*>(plane) + (uint32_t)idx * (uint32_t)sizeof(S));
}

template <class R>
__device__ __forceinline__ R wave_sum(R v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

enum NodeAction : int { ACT_SKIP = 0, ACT_STORE = 1, ACT_COLLIDE = 2, ACT_COPY = 3 };

// Neighbour rows along y: the axis rule, then the precursor wrap of
// bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1131-1132 (row y_wrap pulls its +y
// neighbours from row 0, row 0 its -y ones from row y_wrap).  y is
// wave-uniform, so this is scalar work.
__device__ __forceinline__ void y_nb(const StepParams& p, int y, int& m, int& pl, bool& om, bool& op)
{
    axis_nb(y, p.ny, p.bcy, m, pl, om, op);
    if (p.y_wrap >= 0) {
        if (y == p.y_wrap) { pl = 0; op = false; }
        if (y == 0) { m = p.y_wrap; om = false; }
    }
}

// The 19 pulled populations of node (x, y, z) as stored (raw words, decoded
// once the boundary handling has replaced what it replaces): neighbour rows
// per axis (periodic, clamped or walled), then one load per direction.
// WHOLE (FP32 lattices below 4 GiB, p.whole_ok): buffer loads from one
// descriptor over the whole input lattice, the direction plane's and source
// row's byte offset in a scalar register and the lane's x (one of three per
// node) as the vector offset, so no load needs a vector address add (the
// 512x128x128 channel steps 1.1 % faster).  Otherwise a per-plane base in
// scalar registers and a 32-bit vector offset per load.
template <class S, bool WHOLE = false>
__device__ __forceinline__ void node_load(const StepParams& p, int x, int y, int z, typename Store<S>::Raw r[Q])
{
    const S* __restrict__ fin = static_cast<const S*>(p.fin);
    int xm, xp, ym, yp, zm, zp;
    bool oxm, oxp, oym, oyp, ozm, ozp;
    axis_nb(x, p.nx, p.bcx, xm, xp, oxm, oxp);
    y_nb(p, y, ym, yp, oym, oyp);
    axis_nb(z, p.nz, p.bcz, zm, zp, ozm, ozp);
    const int Y[3] = {ym, y, yp};
    const int Z[3] = {zm, z, zp};
    const int X[3] = {xm, x, xp};
    if constexpr (WHOLE) {
        const __amdgpu_buffer_rsrc_t rs = plane_rsrc(const_cast<S*>(fin), 0xffffffffu);
        const uint32_t vo[3] = {(uint32_t)xm * (uint32_t)sizeof(S), (uint32_t)x * (uint32_t)sizeof(S),
                                (uint32_t)xp * (uint32_t)sizeof(S)};
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const uint32_t so =
                (uint32_t)(i * p.qs + p.nxp * (Y[1 - CY[i]] + p.ny * Z[1 - CZ[i]])) * (uint32_t)sizeof(S);
            r[i] = Store<S>::load_buf(rs, vo[1 - CX[i]], so);
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int src = p.nxp * (Y[1 - CY[i]] + p.ny * Z[1 - CZ[i]]) + X[1 - CX[i]];
            r[i] = Store<S>::load_raw(at(fin + i * p.qs, src));
        }
    }
}

// The unstreamed populations of node `cell` (its state at the step's start).
template <class S>
__device__ __forceinline__ void node_unstreamed(const StepParams& p, int cell, typename Arith<S>::R f[Q])
{
    const S* __restrict__ fin = static_cast<const S*>(p.fin);
    typename Store<S>::Raw r[Q]; // every load issued before the first decode
#pragma unroll
    for (int i = 0; i < Q; ++i) r[i] = Store<S>::load_raw(at(fin + i * p.qs, cell));
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = Store<S>::decode(r[i], i, p.sh);
}

// Node codes at and above NODE_OUTLET_NORTH: the y-streamwise Extrap family
// and the ghost copies, one branch off the common path.
//   im 100 North_Extrap(f, 1) on the node below      (MRTtest/DynInlet3.cu:1711-1716)
//   im 200 South_Extrap(f, v_in) on the node above   (:1717-1724)
//   im  80 rho of the node above, moments of the node below
//          (bgk_channel1/Re3000_Cyl2/DynInlet3.cu:1231-1239)
//   im  70 velocity of the node below, moments of the node above (:1220-1230)
//   im 50-54 copies of an unstreamed node            (MRTtest/DynInlet3.cu:1747-1771)
// The node whose moments are kept is read into f; the Extrap correction
// f_i += (rho* - rho)/19 (density form) or f_i += c_i . (j* - j)/10
// (velocity form) then needs only the differences, which for im 80 / im 70
// start from the kept node's negated sums and take the other node's values
// as they are read, so at most three sums stay live beside f.  Rows y +- 1
// wrap periodically, as the reference's f_mem does (MRTtest/DynInlet3.cu:177-185).
template <class S>
__device__ __forceinline__ int node_rare(const StepParams& p, int x, int y, int z, unsigned char base,
                                         typename Arith<S>::R f[Q])
{
    using R = typename Arith<S>::R;
    const int row = p.nxp * p.ny * z;
    const int below = x + p.nxp * (y > 0 ? y - 1 : p.ny - 1) + row;
    const int above = x + p.nxp * (y + 1 < p.ny ? y + 1 : 0) + row;
    const bool ghost = base >= NODE_GHOST_W;
    int sx = x, sy = y;
    if (base == NODE_GHOST_W) sx = p.nx - 2;
    else if (base == NODE_GHOST_E) sx = 1;
    else if (base == NODE_GHOST_S) sy = p.ghost_y[0];
    else if (base == NODE_GHOST_N) sy = p.ghost_y[1];
    else if (base == NODE_GHOST_DYN) sy = p.ghost_y[2];
    const bool density = base == NODE_OUTLET_NORTH || base == NODE_RECYCLE;
    node_unstreamed<S>(p, ghost ? sx + p.nxp * sy + row : (density ? below : above), f);
    if (ghost) return ACT_COPY;
    // differences to the prescribed values: d0 = rho* - rho (density form),
    // (dx, dy, dz) = j* - j (velocity form)
    R d0 = 0, dx = 0, dy = 0, dz = 0;
    if (density) {
#pragma unroll
        for (int i = 0; i < Q; ++i) d0 -= f[i];
        if (base == NODE_OUTLET_NORTH) d0 += R(1);
    } else {
        R jx, jy, jz;
        momentum(f, jx, jy, jz);
        dx = -jx; dy = -jy; dz = -jz;
        if (base == NODE_INLET_SOUTH) dy += R(p.v_in[z * p.nx + x]);
    }
    if (base == NODE_RECYCLE || base == NODE_VEL_FROM_BELOW) {
        const S* __restrict__ fin = static_cast<const S*>(p.fin);
        const int other = base == NODE_RECYCLE ? above : below;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = Store<S>::load(at(fin + i * p.qs, other), i, p.sh);
            if (density) d0 += v;
            else {
                if (CX[i]) dx += R(CX[i]) * v;
                if (CY[i]) dy += R(CY[i]) * v;
                if (CZ[i]) dz += R(CZ[i]) * v;
            }
        }
    }
    if (density) {
        const R a = d0 * R(1.0 / 19.0);
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] += a;
    } else {
        const R ax = dx * R(0.1), ay = dy * R(0.1), az = dz * R(0.1);
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] += R(CX[i]) * ax + R(CY[i]) * ay + R(CZ[i]) * az;
    }
    return ACT_COLLIDE;
}

// Boundary handling of a pulled node (the work of update_inn before
// mrt_collide, MRTtest/DynInlet3.cu:1612-1771, and of the CPU sweep
// CPU_3D/CPU_3D.cpp:1320-1420), given its node code.  Leaves in f the
// populations to collide (ACT_COLLIDE), or the reversed ones of a node
// bounce-back solid (ACT_STORE), or nothing (ACT_SKIP: a solid under link
// bounce-back).
template <class S>
__device__ __forceinline__ int node_fix(const StepParams& p, int x, int y, int z, int cell, unsigned char code,
                                        typename Store<S>::Raw r[Q], typename Arith<S>::R f[Q],
                                        typename Arith<S>::R& fx, typename Arith<S>::R& fy,
                                        typename Arith<S>::R& fz)
{
    using R = typename Arith<S>::R;
    const S* __restrict__ fin = static_cast<const S*>(p.fin);
    const long long qs = p.qs;
    const unsigned char base = code & NODE_CODE_MASK;
    // a solid under link bounce-back is inert; it is dropped only after the
    // decode below, which every path passes, so the compiler cannot move the
    // pulled loads behind the wait for the node code
    const bool inert = p.halfway && is_solid(base);

    // links bounced from a force-summed solid, as bits: the momentum exchange
    // is summed after the loads, off their dependency chain
    unsigned flinks = 0;
    if ((code & NODE_NEAR) && !inert) {
        // links that bounce: out of the domain through a wall face, or (link
        // mode) from a solid node; the population comes back reversed from
        // this node's own previous post-collision value.  The neighbours'
        // codes are read only where a solid neighbour exists, all of them
        // requested before any is tested (each load on its own would wait a
        // memory round trip), and the bounced words replace the pulled ones
        // before anything is decoded, so those loads are in flight together.
        const bool solid_nb = p.halfway && (code & NODE_NEAR_SOLID);
        int xm, xp, ym, yp, zm, zp;
        bool oxm, oxp, oym, oyp, ozm, ozp;
        axis_nb(x, p.nx, p.bcx, xm, xp, oxm, oxp);
        y_nb(p, y, ym, yp, oym, oyp);
        axis_nb(z, p.nz, p.bcz, zm, zp, ozm, ozp);
        const int Y[3] = {ym, y, yp};
        const int Z[3] = {zm, z, zp};
        const int X[3] = {xm, x, xp};
        unsigned bounce = 0;
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            const bool out = (CX[i] > 0 && oxm) || (CX[i] < 0 && oxp) || (CY[i] > 0 && oym) ||
                             (CY[i] < 0 && oyp) || (CZ[i] > 0 && ozm) || (CZ[i] < 0 && ozp);
            if (out) bounce |= 1u << i;
        }
        if (solid_nb) {
            // every neighbour index is inside the level (axis_nb clamps or wraps)
            unsigned char sc[Q];
#pragma unroll
            for (int i = 1; i < Q; ++i)
                sc[i] = *at(p.mask, p.nxp * (Y[1 - CY[i]] + p.ny * Z[1 - CZ[i]]) + X[1 - CX[i]]);
            const unsigned out = bounce;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (out & (1u << i)) continue;
                const unsigned char b = sc[i] & NODE_CODE_MASK;
                if (is_solid(b)) bounce |= 1u << i;
                if (b == NODE_SOLID_FORCE) flinks |= 1u << i;
            }
        }
        // (through per-plane buffer descriptors instead, the descriptors'
        // scalar registers spill: 140 vs 126 us per 16-bit channel step)
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if (bounce & (1u << i)) r[i] = Store<S>::load_raw(at(fin + OPP[i] * qs, cell));
        // a bounced word comes from plane OPP[i] and decodes with that
        // plane's 16-bit offset, which differs from plane i's where c_x != 0
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CX[i] == 0) f[i] = Store<S>::decode(r[i], i, p.sh);
            else f[i] = Store<S>::decode_shift(r[i], (bounce & (1u << i)) ? p.sh[OPP[i]] : p.sh[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = Store<S>::decode(r[i], i, p.sh);
    }
    if (inert) return ACT_SKIP;
    if (p.fpart && flinks) {
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if (flinks & (1u << i)) {
                fx += R(2 * CX[OPP[i]]) * f[i];
                fy += R(2 * CY[OPP[i]]) * f[i];
                fz += R(2 * CZ[OPP[i]]) * f[i];
            }
    }

    if (is_solid(base)) {
        // node bounce-back (MRTtest/DynInlet3.cu:1673-1705): the pulled values
        // are stored reversed (a register permutation, so every store still
        // writes plane i from a wave-uniform base); momentum exchange
        // 2 sum c_i f_i on im 10 nodes (:1675-1687)
        if (p.fpart && base == NODE_SOLID_FORCE) {
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                if (CX[i]) fx += R(2 * CX[i]) * f[i];
                if (CY[i]) fy += R(2 * CY[i]) * f[i];
                if (CZ[i]) fz += R(2 * CZ[i]) * f[i];
            }
        }
        R g[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) g[i] = f[OPP[i]];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = g[i];
        return ACT_STORE;
    }

    if (base == NODE_INLET_WEST || base == NODE_OUTLET_EAST) {
        // Extrap conditions start from the unstreamed neighbour (:1725-1746)
        const int nb = base == NODE_INLET_WEST ? (x + 1 < p.nx ? cell + 1 : cell)
                                               : (x > 0 ? cell - 1 : cell);
        node_unstreamed<S>(p, nb, f);
        if (base == NODE_INLET_WEST) west_extrap(f, R(p.u_in[z * p.ny + y]));
        else east_extrap(f, R(1));
    } else if (base == NODE_LID_WEST_REG) {
        lid_west_reg(f, R(p.u_lid));
    } else if (base == NODE_SYM_XMIN) {
        xsym_bot(f, y, z, p.ny, p.nz);
    } else if (base == NODE_SYM_XMAX) {
        xsym_top(f, y, z, p.ny, p.nz);
    } else if (base >= NODE_OUTLET_NORTH) {
        return node_rare<S>(p, x, y, z, base, f);
    }
    return ACT_COLLIDE;
}

// Pull stream + bounce-back + boundary conditions for one node.  The 19
// loads are issued together with the node code's, not after it: waiting for
// the code first would put a second memory round trip on every wavefront's
// path.  A solid under link bounce-back then loads values it never uses (in
// bounds; solids are a small share of the nodes).
template <class S, bool WHOLE = false>
__device__ __forceinline__ int node_pull(const StepParams& p, int x, int y, int z, int cell,
                                         typename Arith<S>::R f[Q], typename Arith<S>::R& fx,
                                         typename Arith<S>::R& fy, typename Arith<S>::R& fz)
{
    const unsigned char code = *at(p.mask, cell);
    typename Store<S>::Raw r[Q];
    node_load<S, WHOLE>(p, x, y, z, r);
    return node_fix<S>(p, x, y, z, cell, code, r, f, fx, fy, fz);
}

template <class R>
__device__ __forceinline__ CollideParams<R> collide_params(const StepParams& p, int y)
{
    CollideParams<R> cp;
    cp.omega = R(p.omega);
    cp.cs = R(p.cs);
    cp.dpdy = y > p.dyn_y1 ? R(0) : R(p.dpdy);
    cp.s1 = R(p.s1); cp.s2 = R(p.s2); cp.s4 = R(p.s4); cp.s10 = R(p.s10); cp.s16 = R(p.s16);
    return cp;
}

// Store one node's 19 populations: buffer stores from per-plane resource
// descriptors, so the 19 stores share one 32-bit offset register (global
// stores took a 64-bit per-lane address add each); out-of-range offsets are
// dropped by the hardware.  AUX is the cache policy (kStoreNT or 0).
template <class S, int AUX>
__device__ __forceinline__ void node_store_policy(const StepParams& p, int cell, const typename Arith<S>::R f[Q])
{
    S* __restrict__ fout = static_cast<S*>(p.fout);
    const uint32_t bytes = (uint32_t)(p.qs * (long long)sizeof(S)), off = (uint32_t)cell * (uint32_t)sizeof(S);
#pragma unroll
    for (int i = 0; i < Q; ++i)
        buffer_put<AUX>(Store<S>::raw(f[i], i, p.sh), plane_rsrc(fout + i * p.qs, bytes), off);
}

// The same with the level's policy chosen at run time (p.nt_store).
template <class S>
__device__ __forceinline__ void node_store(const StepParams& p, int cell, const typename Arith<S>::R f[Q])
{
    if (p.nt_store) node_store_policy<S, kStoreNT>(p, cell, f);
    else node_store_policy<S, 0>(p, cell, f);
}

// Running statistics of the post-collision velocity sum c_i f_i at one node:
// vel_av / vel_fluc (MRTtest/DynInlet3.cu:1139-1148, :1159-1172) on the
// coarse level, weight 1; vel_avLR / vel_flucLR (:1150-1157, :1174-1184) on
// the fine level, weight LRFACTOR.  mean <- (mean n + u w) / (n + w); from the
// deviation phase on the mean is frozen and dev <- (dev n + (u - mean)^2 w) /
// (n + w).  Sums in the reference's order, no contraction.
template <class R>
__device__ __forceinline__ void stats_update(const StepParams& p, int cell, const R f[Q])
{
#pragma clang fp contract(off)
    const float u[3] = {float(f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + f[10] - f[12] + f[15] - f[17]),
                        float(f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + f[11] - f[13] + f[16] - f[18]),
                        float(f[9] + f[10] + f[11] + f[12] + f[13] - f[14] - f[15] - f[16] - f[17] - f[18])};
    const float n = p.stat_n, w = p.stat_w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float* mean = at(p.stat + a * p.qs, cell);
        if (p.stat_mode == 1) {
            *mean = (*mean * n + u[a] * w) / (n + w);
        } else {
            float* dev = at(p.stat + (3 + a) * p.qs, cell);
            const float m = *mean;
            const float d = (u[a] - m) * (u[a] - m);
            *dev = (*dev * n + d * w) / (n + w);
        }
    }
}

// Collide a pulled node per its action (plus its running statistics); the
// body force acts only up to row y = dyn_y1 (MRTtest/DynInlet3.cu:1744).
template <int MODEL, class S>
__device__ __forceinline__ void node_collide(const StepParams& p, int y, int cell, int act, typename Arith<S>::R f[Q])
{
    using R = typename Arith<S>::R;
    if (act != ACT_COLLIDE) return;
    CollideParams<R> cp = collide_params<R>(p, y);
    if (p.omega_node) {
        const float w = *at(p.omega_node, cell);
        if (w > 0.f) {
            cp.omega = R(w);
            cp.cs = R(0);
        }
    }
    collide<MODEL, R, sizeof(S) == 2>(f, cp);
    if (p.stat_mode) stats_update(p, cell, f);
}

// Collide (per the node's action) and store one pulled node.
template <int MODEL, class S, int AUX>
__device__ __forceinline__ void node_finish(const StepParams& p, int y, int cell, int act, typename Arith<S>::R f[Q])
{
    if (act == ACT_SKIP) return;
    node_collide<MODEL, S>(p, y, cell, act, f);
    // a ghost copy's statistics see the copied populations, as update_inn's
    // vel_av follows the copies (MRTtest/DynInlet3.cu:1747-1779)
    if (act == ACT_COPY && p.stat_mode) stats_update(p, cell, f);
    if constexpr (AUX < 0) node_store<S>(p, cell, f); // the level's policy at run time
    else node_store_policy<S, AUX>(p, cell, f);
}

// Footprint moments of a coarse node under refinement (defined with the
// interface kernels).
template <class S, class R>
__device__ void footprint_extract(const StepParams& p, int x, int y, int z, int cell, int act, R f[Q]);

// Fused pull stream + boundary dispatch + collide + bounce-back for one node.
template <int MODEL, class S, int AUX, bool WHOLE = false>
__device__ __forceinline__ void node_update(const StepParams& p, int x, int y, int z,
                                            typename Arith<S>::R& fx, typename Arith<S>::R& fy,
                                            typename Arith<S>::R& fz)
{
    using R = typename Arith<S>::R;
    const int cell = x + p.nxp * (y + p.ny * z);
    R f[Q];
    const int act = node_pull<S, WHOLE>(p, x, y, z, cell, f, fx, fy, fz);
    node_finish<MODEL, S, AUX>(p, y, cell, act, f);
    if (p.fp_n[0] > 0) footprint_extract<S>(p, x, y, z, cell, act, f);
}

// Momentum-exchange partial of one workgroup: wavefront sums, then the
// wavefronts in a fixed order (deterministic), in FP64.
template <int NW, class R>
__device__ __forceinline__ void block_force_partial(const StepParams& p, int lb, R fx, R fy, R fz)
{
    __shared__ double red[3][NW];
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane_id() == 0) {
        red[0][wave_id()] = double(fx);
        red[1][wave_id()] = double(fy);
        red[2][wave_id()] = double(fz);
    }
    __syncthreads();
    if (threadIdx.x < 3 && threadIdx.y == 0) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[threadIdx.x][w];
        p.fpart[(long long)lb * 3 + threadIdx.x] = s;
    }
}

// Momentum-exchange partial of one wavefront of a force-holding workgroup:
// its FP64 sum (fixed shuffle tree) into wave slot slot*NW + wave of this
// step's partials.  No barrier: the other wavefronts of the workgroup need
// not wait for the slow near-solid one.
template <int NW, class R>
__device__ __forceinline__ void wave_force_partial(const StepParams& p, int slot, R fx, R fy, R fz)
{
    const double sx = wave_sum(double(fx)), sy = wave_sum(double(fy)), sz = wave_sum(double(fz));
    if (lane_id() == 0) {
        double* o = p.fpart + ((long long)slot * NW + wave_id()) * 3;
        o[0] = sx;
        o[1] = sy;
        o[2] = sz;
    }
}

// Force of each step of a batch: workgroup s sums step s's n wave partials
// (3 components) in a fixed order (strided per thread, then a fixed tree)
// into out[3 s + c].
__global__ void __launch_bounds__(256) k_reduce_steps(const double* part, long long n, double* out)
{
    __shared__ double sh[3][256];
    const double* q = part + (long long)blockIdx.x * n * 3;
    double a = 0, b = 0, c = 0;
    for (long long i = threadIdx.x; i < n; i += 256) {
        a += q[i * 3];
        b += q[i * 3 + 1];
        c += q[i * 3 + 2];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[(long long)blockIdx.x * 3 + threadIdx.x] = sh[threadIdx.x][0];
}

// The stream-collide step; NT selects nontemporal population stores (the
// level's p.nt_store, as a template argument so each variant is straight-line).
template <int MODEL, class S, int AUX, bool WHOLE = false>
__device__ __forceinline__ void sc_block(const StepParams& p, int lb)
{
    using R = typename Arith<S>::R;
    constexpr int NW = Shape<S>::NW;
    // only workgroups holding a force node write a partial; the slot is read
    // (a scalar load) before the pull, so its latency hides under the loads
    const int slot = p.fpart ? (p.fslot ? p.fslot[lb] : lb) : -1;
    R fx = 0, fy = 0, fz = 0;
    int x, y, z;
    const bool row_ok = wave_node_map<NW>(lb, p.map, p.nz, x, y, z);
    if (row_ok && x < p.nx) node_update<MODEL, S, AUX, WHOLE>(p, x, y, z, fx, fy, fz);
    if (slot >= 0) wave_force_partial<NW>(p, slot, fx, fy, fz);
}

// WHOLE: the whole-lattice load descriptor (node_load), for FP32 lattices
// below 4 GiB.
template <int MODEL, class S, bool NT, bool WHOLE = false>
__global__ void __launch_bounds__(Shape<S>::THREADS, Shape<S>::MIN_WAVES) k_stream_collide(StepParams p)
{
    sc_block<MODEL, S, NT ? kStoreNT : 0, WHOLE>(p, logical_block_map(blockIdx.x, p.map));
}

// Two levels in one launch: workgroups below blocks_a update level a (the
// coarse level), the rest level b (a fine sub-step that reads nothing level
// a writes).  The hardware hands out workgroup ids in order, so b's
// workgroups fill the compute units a's last workgroups leave idle, with no
// second stream and no cross-stream wait.  The store policy is each level's
// own, chosen at run time (p.nt_store; as template arguments it measured
// 0.5 % faster for 2.5 times the code).
template <int MODEL, class S, bool WHOLE = false>
__global__ void __launch_bounds__(Shape<S>::THREADS, Shape<S>::MIN_WAVES)
    k_stream_collide_two(StepParams a, StepParams b, int blocks_a)
{
    const int hw = blockIdx.x;
    if (hw < blocks_a) sc_block<MODEL, S, -1, WHOLE>(a, logical_block_map(hw, a.map));
    else sc_block<MODEL, S, -1, WHOLE>(b, hw - blocks_a);
}

// Deterministic FP64 reduction of ncomp-component partials (fixed order:
// strided per thread, then a fixed tree), one workgroup of 256 threads.
__global__ void __launch_bounds__(256) k_reduce_partials(const double* part, long long n, int ncomp,
                                                        double* out)
{
    __shared__ double sh[256];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0;
        for (long long i = threadIdx.x; i < n; i += 256) s += part[i * ncomp + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sh[0];
        __syncthreads();
    }
}

// First stage of the two-stage reduction of many partials: workgroup b sums
// partials [b*REDUCE_CHUNK, (b+1)*REDUCE_CHUNK) in a fixed order (strided
// per thread, then a fixed tree) into out[b*ncomp + c]; k_reduce_partials
// then sums the gridDim.x results.  A single workgroup over the 32 K force
// partials of a 512x128x128 step took 93 us; the two stages spread the
// reads over many CUs.
constexpr int REDUCE_CHUNK = 512;
__global__ void __launch_bounds__(256) k_reduce_chunks(const double* part, long long n, int ncomp, double* out)
{
    __shared__ double sh[256];
    const long long base = (long long)blockIdx.x * REDUCE_CHUNK;
    const long long end = base + REDUCE_CHUNK < n ? base + REDUCE_CHUNK : n;
    for (int c = 0; c < ncomp; ++c) {
        double s = 0;
        for (long long i = base + threadIdx.x; i < end; i += 256) s += part[i * ncomp + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(long long)blockIdx.x * ncomp + c] = sh[0];
        __syncthreads();
    }
}

// AverageV (MRTtest/DynInlet3.cu:2561-2600): v = sum_i c_iy f_i at each node
// of the plane y, summed in the reference's order, then over the 64-node
// row segment of the wavefront in FP64; one partial per (segment, z).
template <class S>
__global__ void __launch_bounds__(64) k_plane_v(const S* f, long long qs, int nx, int nxp, int ny, int nsegw, int y,
                                                double* part, PopShift sh)
{
    using R = typename Arith<S>::R;
    const int seg = blockIdx.x % nsegw, z = blockIdx.x / nsegw;
    const int x = seg * 64 + lane_id();
    double v = 0.0;
    if (x < nx) {
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        R g[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) g[i] = R(Store<S>::to_real(f[i * qs + cell], i, sh));
        v = double(g[2] - g[4] + g[5] + g[6] - g[7] - g[8] + g[11] - g[13] + g[16] - g[18]);
    }
    v = wave_sum(v);
    if (lane_id() == 0) part[blockIdx.x] = v;
}

// FP64 sums of (rho, jx, jy, jz) over non-solid nodes, one partial per
// workgroup.  MM_NW-wavefront workgroups stride over the groups of MM_NW row
// segments (workgroup g takes groups g, g + G, g + 2G, ... for a launch of G
// workgroups), each lane accumulating its nodes in that order; then the
// wavefront sums and the wavefronts in a fixed order.  The partial count is
// the launch's G whatever the lattice size, and the result is deterministic
// for a given G.
constexpr int MM_NW = 4;
constexpr int MM_GRID_MAX = 2048; // workgroups (partials) of one mass/momentum launch
__host__ inline int mm_grid(int nsegw, int ny, int nz)
{
    const long long groups = ((long long)nsegw * ny * nz + MM_NW - 1) / MM_NW;
    return (int)(groups < MM_GRID_MAX ? groups : MM_GRID_MAX);
}
template <class S>
__global__ void __launch_bounds__(64 * MM_NW) k_mass_momentum(const S* f, const unsigned char* mask,
                                                              long long qs, int nx, int ny, int nxp,
                                                              int nsegw, int nz, double* part, PopShift sh)
{
    using R = typename Arith<S>::R;
    const long long groups = ((long long)nsegw * ny * nz + MM_NW - 1) / MM_NW;
    double acc[4] = {0, 0, 0, 0};
    for (long long g = blockIdx.x; g < groups; g += gridDim.x) {
        int x, y, z;
        if (!wave_node<MM_NW>((int)g, nsegw, ny, nz, x, y, z) || x >= nx) continue;
        const int cell = x + nxp * (y + ny * z);
        if (mask && is_solid(mask[cell])) continue;
        R r = 0, a = 0, b = 0, c = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = Store<S>::to_real(f[i * qs + cell], i, sh);
            r += v;
            if (CX[i]) a += R(CX[i]) * v;
            if (CY[i]) b += R(CY[i]) * v;
            if (CZ[i]) c += R(CZ[i]) * v;
        }
        acc[0] += double(r); acc[1] += double(a); acc[2] += double(b); acc[3] += double(c);
    }
    __shared__ double red[4][MM_NW];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane_id() == 0) red[k][wave_id()] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int w = 0; w < MM_NW; ++w) s += red[threadIdx.x][w];
        part[(long long)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// Node fields as WriteResults computes them: rho = sum f, (u,v,w) = momentum.
template <class S>
__global__ void k_macro(const S* f, long long qs, int nx, int ny, int nz, int nxp, double* rho,
                        double* u, double* v, double* w, PopShift sh)
{
    using R = typename Arith<S>::R;
    const long long n = (long long)nx * ny * nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = int(c % nx);
        const long long r = c / nx;
        const int y = int(r % ny);
        const int z = int(r / ny);
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        R s = 0, a = 0, b = 0, d = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R val = Store<S>::to_real(f[i * qs + cell], i, sh);
            s += val;
            if (CX[i]) a += R(CX[i]) * val;
            if (CY[i]) b += R(CY[i]) * val;
            if (CZ[i]) d += R(CZ[i]) * val;
        }
        rho[c] = s; u[c] = a; v[c] = b; w[c] = d;
    }
}

// Dense <-> padded population copies with conversion.
template <class S, class T>
__global__ void k_pack(const S* f, long long qs, int nx, int ny, int nz, int nxp, T* out, PopShift sh)
{
    const long long n = (long long)nx * ny * nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n * Q;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = int(t / n);
        const long long c = t % n;
        const int x = int(c % nx);
        const long long r = c / nx;
        const int y = int(r % ny);
        const int z = int(r / ny);
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        out[t] = T(Store<S>::to_real(f[i * qs + cell], i, sh));
    }
}

template <class S, class T>
__global__ void k_unpack(const T* in, long long qs, int nx, int ny, int nz, int nxp, S* f, S* f2, PopShift sh)
{
    using R = typename Arith<S>::R;
    const long long n = (long long)nx * ny * nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n * Q;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = int(t / n);
        const long long c = t % n;
        const int x = int(c % nx);
        const long long r = c / nx;
        const int y = int(r % ny);
        const int z = int(r / ny);
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        const S v = Store<S>::from_real(R(in[t]), i, sh);
        f[i * qs + cell] = v;
        if (f2) f2[i * qs + cell] = v;
    }
}

// Equilibrium initialisation of both lattices.  Fields are dense; nullptr
// fields mean rho = 1, u = (u0, 0, 0), or (0, u0, 0) with ydir (the
// y-streamwise cases' initialize, MRTtest/DynInlet3.cu:1200-1203).
//  kind 0: BGK feq at every node, solids included (CPU_3D/CPU_3D.cpp:1218-1254)
//  kind 1: mrt_meq + InvertMoments, solids at u = 0 (initialize,
//          MRTtest/DynInlet3.cu:1186-1216)
//  kind 2: Taylor-Green (u0 cos kx sin ky, -u0 sin kx cos ky, 0), BGK feq.
template <class S>
__global__ void k_init(S* fa, S* fb, long long qs, int nx, int ny, int nz, int nxp,
                       const unsigned char* mask, const float* rho_in, const float* u_in,
                       const float* v_in, const float* w_in, float u0, int ydir, int kind, PopShift sh)
{
    using R = typename Arith<S>::R;
    const long long n = (long long)nx * ny * nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = int(c % nx);
        const long long r = c / nx;
        const int y = int(r % ny);
        const int z = int(r / ny);
        const long long cell = x + (long long)nxp * (y + (long long)ny * z);
        R rho = rho_in ? R(rho_in[c]) : R(1);
        R u = u_in ? R(u_in[c]) : (ydir ? R(0) : R(u0));
        R v = v_in ? R(v_in[c]) : (ydir ? R(u0) : R(0));
        R w = w_in ? R(w_in[c]) : R(0);
        if (kind == 2) {
            const double kx = 2.0 * 3.14159265358979323846 * x / nx;
            const double ky = 2.0 * 3.14159265358979323846 * y / nx;
            u = R(double(u0) * cos(kx) * sin(ky));
            v = R(-double(u0) * sin(kx) * cos(ky));
            w = R(0);
            rho = R(1);
        }
        if (kind == 1 && mask && is_solid(mask[cell])) { u = v = w = R(0); }
        R feq[Q];
        if (kind == 1) {
            // mrt_meq (MRTtest/DynInlet3.cu:329-345) then InvertMoments
            const R usqr = u * u + v * v + w * w;
            R m[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) m[k] = R(0);
            m[0] = rho;
            m[1] = R(-11) * rho + R(19) * usqr;
            m[2] = R(475.0 / 63.0) * usqr;
            m[3] = u; m[4] = R(-2.0 / 3.0) * u;
            m[5] = v; m[6] = R(-2.0 / 3.0) * v;
            m[7] = w; m[8] = R(-2.0 / 3.0) * w;
            m[9] = R(2) * u * u - (v * v + w * w);
            m[11] = v * v - w * w;
            m[13] = u * v; m[14] = v * w; m[15] = u * w;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                R s = 0;
#pragma unroll
                for (int k = 0; k < Q; ++k)
                    if (mrow(k, i) != 0) s += (R(mrow(k, i)) / R(NORM[k])) * m[k];
                feq[i] = s;
            }
        } else {
            feq_bgk(rho, u, v, w, feq);
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const S sv = Store<S>::from_real(feq[i], i, sh);
            fa[i * qs + cell] = sv;
            fb[i * qs + cell] = sv;
        }
    }
}

} // namespace lbm
