// flow_tools.hpp -- device reduction behind the adaptive CFL controller.
#pragma once
#include "runtime.hpp"

namespace tarang {

// max over npts points of sum_i |comp_i| / dx_i (ncomp 1..3), or, with dx == nullptr
// and one component, the plain maximum of comp_0.  FP64 arithmetic; NaN propagates;
// npts == 0 gives 0.  Synchronises the context stream.
double cfl_frequency(Context* c, int ncomp, const void* const* comps, size_t npts, const double* dx, DType dt);

// the same maximum written to the device double out_dev on the context stream, with no
// allocation and no synchronisation: the partial maxima use a fixed per-GPU device array
// shared by every call on that GPU (calls that overlap on one GPU race)
void cfl_frequency_noalloc(Context* c, int ncomp, const void* const* comps, size_t npts, const double* dx, DType dt,
                           double* out_dev);

// *dflag = 1 (on the context stream, no synchronisation) if any of the nreal real values
// at p is NaN or infinite; *dflag is left untouched otherwise.  dflag is a device int.
void flag_nonfinite(Context* c, const void* p, size_t nreal, DType dt, int* dflag);

}  // namespace tarang
