// storage.h -- how populations sit in HBM.
//
// Layout: SoA, one plane per direction, plane q holds f_q[z][y][x] with the
// x stride padded to a multiple of 64 elements so every 64-lane wavefront
// row access starts on a 256-byte boundary (FP32); consecutive planes sit
// 448 elements further apart than a plane's nodes need, which staggers the
// 19 planes over the HBM channels.  Four element types:
//   float      FP32 populations, FP32 arithmetic (reference precision)
//   double     FP64 populations and arithmetic (the FP64 yardstick)
//   Half16     f_i - w_i as IEEE binary16, FP32 arithmetic
//   BHalf16    f_i - w_i as bfloat16 (upper half of an FP32), FP32 arithmetic
// The 16-bit types store the shifted deviation: the lattice weight is added
// back in FP32 after the load and subtracted before the store.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19.h"

#ifndef LBM_NT_LOAD
#define LBM_NT_LOAD 0 // nontemporal population loads
#endif
#ifndef LBM_NT_STORE
#define LBM_NT_STORE 1 // nontemporal population stores
#endif

namespace lbm {

template <class T>
__device__ __forceinline__ T ld(const T* p)
{
    if (LBM_NT_LOAD) return __builtin_nontemporal_load(p);
    return *p;
}

template <class T>
__device__ __forceinline__ void st(T* p, T v)
{
    if (LBM_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

struct Half16 { uint16_t bits; };
struct BHalf16 { uint16_t bits; };

template <class S> struct Arith { using R = float; };
template <> struct Arith<double> { using R = double; };

// Offsets the 16-bit formats subtract before rounding: s_i = f_i^eq(rho = 1,
// u = (U, 0, 0)), the equilibrium of the inflow (w_i when U = 0, the plain
// shifted deviation f_i - w_i).  Passed by value in kernel arguments, so a
// load costs one add with a scalar operand.  Ignored by FP32/FP64 storage.
// Seven distinct values cover the 19 directions (rest; |c| = 1 with c_x =
// +1, -1, 0; |c| = sqrt2 with c_x = +1, -1, 0), which keeps the scalar
// registers they occupy low.
struct PopShift {
    float v[7];
    __host__ __device__ static constexpr int slot(int i)
    {
        return i == 0 ? 0 : (WNUM[i] == 2 ? (CX[i] > 0 ? 1 : (CX[i] < 0 ? 2 : 3)) : (CX[i] > 0 ? 4 : (CX[i] < 0 ? 5 : 6)));
    }
    __host__ __device__ float operator[](int i) const { return v[slot(i)]; }
};

template <class S>
struct Store;

// Buffer store of one stored word at byte offset off of a direction plane:
// the plane's resource descriptor sits in scalar registers and every lane
// supplies a 32-bit offset, so the 19 stores of a node share one offset
// register (global stores there took a 64-bit per-lane address add each).
// Descriptor word 3 is the raw-buffer format of gfx9-family targets.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(void* plane, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(plane, 0, (int)bytes, 0x00020000);
}
// Cache policy of the stream-collide's stores: nontemporal (aux 2) when the
// two lattices exceed the 256 MB last-level cache (MALL), normal (aux 0)
// when they fit and the next step's loads can hit it.
constexpr int kStoreNT = 2;
template <int AUX>
__device__ __forceinline__ void buffer_put(uint16_t w, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b16(w, r, (int)off, 0, AUX);
}
template <int AUX>
__device__ __forceinline__ void buffer_put(uint32_t w, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(w, r, (int)off, 0, AUX);
}
template <int AUX>
__device__ __forceinline__ void buffer_put(u32x2 w, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)off, 0, AUX);
}

template <>
struct Store<float> {
    using Raw = float;
    __device__ static __forceinline__ Raw load_raw(const float* p) { return ld(p); }
    // buffer load at byte offset voff + soff of the descriptor's range
    __device__ static __forceinline__ Raw load_buf(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    }
    __device__ static __forceinline__ float decode(Raw r, int, const PopShift&) { return r; }
    __device__ static __forceinline__ float decode_shift(Raw r, float) { return r; }
    // the value a store of v leaves in memory
    __device__ static __forceinline__ float stored(float v, int, const PopShift&) { return v; }
    __device__ static __forceinline__ float load(const float* p, int, const PopShift&) { return ld(p); }
    __device__ static __forceinline__ void store(float* p, float v, int, const PopShift&) { st(p, v); }
    __device__ static __forceinline__ uint32_t raw(float v, int, const PopShift&) { return __builtin_bit_cast(uint32_t, v); }
    __host__ __device__ static float to_real(float v, int, const PopShift&) { return v; }
    __host__ __device__ static float from_real(float v, int, const PopShift&) { return v; }
};

template <>
struct Store<double> {
    using Raw = double;
    __device__ static __forceinline__ Raw load_raw(const double* p) { return ld(p); }
    __device__ static __forceinline__ double decode(Raw r, int, const PopShift&) { return r; }
    __device__ static __forceinline__ double decode_shift(Raw r, float) { return r; }
    __device__ static __forceinline__ double stored(double v, int, const PopShift&) { return v; }
    __device__ static __forceinline__ double load(const double* p, int, const PopShift&) { return ld(p); }
    __device__ static __forceinline__ void store(double* p, double v, int, const PopShift&) { st(p, v); }
    __device__ static __forceinline__ u32x2 raw(double v, int, const PopShift&) { return __builtin_bit_cast(u32x2, v); }
    __host__ __device__ static double to_real(double v, int, const PopShift&) { return v; }
    __host__ __device__ static double from_real(double v, int, const PopShift&) { return v; }
};

__host__ __device__ inline float half_bits_to_float(uint16_t h)
{
    __half x;
    __builtin_memcpy(&x, &h, 2);
    return __half2float(x);
}

__host__ __device__ inline uint16_t float_to_half_bits(float v)
{
    __half x = __float2half_rn(v);
    uint16_t h;
    __builtin_memcpy(&h, &x, 2);
    return h;
}

__host__ __device__ inline float bf16_bits_to_float(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// round to nearest even (inputs here are finite)
__host__ __device__ inline uint16_t float_to_bf16_bits(float v)
{
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// The FP16 load and store each fold the conversion and the offset into one
// mixed-precision FMA (v_fma_mix_f32: f16 source times 1.0 plus the offset;
// v_fma_mixlo_f16: the f32 value minus the offset, rounded once to f16),
// instead of a conversion plus an add.  The offsets are kernel arguments, so
// they sit in scalar registers.
template <>
struct Store<Half16> {
    using Raw = unsigned; // the stored bits, zero-extended
    __device__ static __forceinline__ Raw load_raw(const Half16* p) { return ld(&p->bits); }
    __device__ static __forceinline__ float decode(Raw h, int I, const PopShift& sh)
    {
        float v;
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(v) : "v"(h), "s"(sh[I]));
        return v;
    }
    // the same with a per-lane offset
    __device__ static __forceinline__ float decode_shift(Raw h, float shift)
    {
        float v;
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(v) : "v"(h), "v"(shift));
        return v;
    }
    __device__ static __forceinline__ float load(const Half16* p, int I, const PopShift& sh)
    {
        return decode(load_raw(p), I, sh);
    }
    // the value a store of v leaves in memory (rounded to the 16-bit grid)
    __device__ static __forceinline__ float stored(float v, int I, const PopShift& sh)
    {
        return decode(raw(v, I, sh), I, sh);
    }
    __device__ static __forceinline__ uint16_t raw(float v, int I, const PopShift& sh)
    {
        unsigned o;
        asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2" : "=v"(o) : "v"(v), "s"(sh[I]));
        return (uint16_t)o;
    }
    __device__ static __forceinline__ void store(Half16* p, float v, int I, const PopShift& sh)
    {
        st(&p->bits, raw(v, I, sh));
    }
    __host__ __device__ static float to_real(Half16 v, int i, const PopShift& sh)
    {
        return half_bits_to_float(v.bits) + sh[i];
    }
    __host__ __device__ static Half16 from_real(float v, int i, const PopShift& sh)
    {
        return Half16{float_to_half_bits(v - sh[i])};
    }
};

template <>
struct Store<BHalf16> {
    using Raw = unsigned; // the stored bits, zero-extended
    __device__ static __forceinline__ Raw load_raw(const BHalf16* p) { return ld(&p->bits); }
    __device__ static __forceinline__ float decode(Raw h, int I, const PopShift& sh)
    {
        return bf16_bits_to_float((uint16_t)h) + sh[I];
    }
    __device__ static __forceinline__ float decode_shift(Raw h, float shift)
    {
        return bf16_bits_to_float((uint16_t)h) + shift;
    }
    __device__ static __forceinline__ float load(const BHalf16* p, int I, const PopShift& sh)
    {
        return decode(load_raw(p), I, sh);
    }
    // the value a store of v leaves in memory (rounded to the 16-bit grid)
    __device__ static __forceinline__ float stored(float v, int I, const PopShift& sh)
    {
        return decode(raw(v, I, sh), I, sh);
    }
    __device__ static __forceinline__ uint16_t raw(float v, int I, const PopShift& sh)
    {
        // v_cvt_pk_bf16_f32: round to nearest even, as float_to_bf16_bits
        return __builtin_bit_cast(uint16_t, (__bf16)(v - sh[I]));
    }
    __device__ static __forceinline__ void store(BHalf16* p, float v, int I, const PopShift& sh)
    {
        st(&p->bits, raw(v, I, sh));
    }
    __host__ __device__ static float to_real(BHalf16 v, int i, const PopShift& sh)
    {
        return bf16_bits_to_float(v.bits) + sh[i];
    }
    __host__ __device__ static BHalf16 from_real(float v, int i, const PopShift& sh)
    {
        return BHalf16{float_to_bf16_bits(v - sh[i])};
    }
};

} // namespace lbm
