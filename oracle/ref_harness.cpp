/*
 * ref_harness.cpp -- builds the reference's own CPU solver, CPU_3D/CPU_3D.cpp,
 * from where it lies under the reference tree (the Makefile passes its folder
 * with -I) into oracle/_ref/libcpu3d_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: the reference is used to pin the oracle and as
 * the CPU baseline; nothing of it is copied here, and the product never
 * loads it.  The harness renames the reference's main() (never called here:
 * the whole program is built unchanged as _ref/cpu3d_ref) and exposes its
 * per-node functions through a C ABI.
 */
#define main cpu3d_reference_main
#include "CPU_3D.cpp"
#undef main

#define F19(f) f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10], f[11], f[12], \
               f[13], f[14], f[15], f[16], f[17], f[18]

extern "C" {

void ref_bgk_collide(float* f, float omega) { bgk_collide(F19(f), omega); }

void ref_mrt_collide(float* f, float omega) { mrt_collide(F19(f), omega); }

void ref_boundaries(float* f, int y, int z, int im) { boundaries(F19(f), y, z, im); }

int ref_image(int x, int y, int z) { return ImageFcn(x, y, z); }

int ref_dmin(int a, int b) { return dmin(a, b); }

int ref_dmax(int a) { return dmax(a); }

/* grid and case constants of the reference build */
void ref_case(int* dims, float* consts)
{
    dims[0] = XDIM;
    dims[1] = YDIM;
    dims[2] = ZDIM;
    dims[3] = TMAX;
    consts[0] = UMAX;
    consts[1] = RE;
    consts[2] = OBSTR;
    const float CharLength = OBSTR * 2.0f; // as main() computes omega (:1221, :1299)
    consts[3] = 1.0f / (3.0f * (UMAX * CharLength / RE) + 0.5f);
}

}
