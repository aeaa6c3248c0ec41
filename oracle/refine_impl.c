/*
 * refine_impl.c -- CPU restatement of the reference's coarse-fine interface
 * (float and double).  TEST INFRASTRUCTURE ONLY: the checker of the HIP
 * refinement path.
 *
 * Follows MRTtest/DynInlet3.cu: PhysicalMoments :396-407, StrainRate :367-393,
 * InvertPhysicalMoments_LES_fc :483-538, _LES_cf :540-595, InterpCF
 * :2202-2382 (trilinear :165-175, ORDER-2 correction :2312-2357), InterpFC
 * :2383-2560, and the sub-step sequence of main() :3180-3245.  The frame is
 * two fine nodes thick on all six faces (the reference's patch spans z
 * periodically, so it has x/y faces only); the coarse footprint moments are
 * optionally interpolated to t+1/2 for the second sub-step (the reference
 * reuses the state of time t).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define FN(name) name##_f32
#define R float
#define C_(x) ((float)(x))
#define SQRT_ sqrtf
#include "refine_impl.h"
#undef FN
#undef R
#undef C_
#undef SQRT_

#define FN(name) name##_f64
#define R double
#define C_(x) ((double)(x))
#define SQRT_ sqrt
#include "refine_impl.h"
#undef FN
#undef R
#undef C_
#undef SQRT_
