// diagnostics.hpp -- trajectory diagnostics reduced on the device: energy and enstrophy of
// a periodic field, the Nusselt number of a Rayleigh-Benard channel.  FP64 accumulation,
// fixed reduction order (the same state gives the same bits).
#pragma once
#include "runtime.hpp"

namespace tarang {

// From the coefficients (nx/2+1, ny) of a real periodic field (unnormalised rfft along x,
// FFT along y; ny = 1 in 1-D), by Parseval over the half spectrum:
//   vorticity q:  E = 1/2 <u^2 + v^2> = 1/2 sum |q_k|^2 / |k|^2,   Z = 1/2 <q^2>
//   otherwise f:  E = 1/2 <f^2>,                                   Z = 1/2 <|grad f|^2>
// (the reference reduces 1/2 mean(u^2 + v^2) and 1/2 mean(q^2) on the grid on the device,
// examples/2d_decaying_turbulence_gpu.jl:84-92).  Synchronises the context stream.
void spectral_energy(Context* c, const void* coeff, DType dt, int nx, int ny, double k0x, double k0y, bool vorticity,
                     double* E, double* Z);

// Nu = <uz b - kappa dz b>_V / (kappa dB / Lz) from mode-major channel state rows
// (f[mode * N + k], mode = m + (nx/2+1) n): the horizontal mean exactly by Parseval, the
// z mean by Clenshaw-Curtis quadrature on the Gauss-Lobatto points, the conduction term
// from the wall values of the mean profile.  ny = 1 in 2-D.  Synchronises the stream.
double channel_nusselt(Context* c, const void* b, const void* uz, DType dt, int nx, int ny, int N, double Lz,
                       double kappa, double dB);

}  // namespace tarang
