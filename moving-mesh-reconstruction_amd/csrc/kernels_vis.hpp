// kernels_vis.hpp -- downsampling and column-density projection of the reference's analysis tools
// (from1024to512.f90, column_density.f90) on the device.
#pragma once
#include "kernels.hpp"

namespace mmh {

// One 2x2x2 block-sum coarsening (from1024to512.f90:54-66, subroutine coarsen): out (h^3) from in
// (n^3, n = 2h), the eight values summed in Fortran array-section order (i fastest).
__global__ void k_coarsen2(float* __restrict__ out, const float* __restrict__ in, int h) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)h * h * h) return;
  const int i = (int)(q % h), j = (int)((q / h) % h), k = (int)(q / ((i64)h * h));
  const int n = 2 * h;
  float s = 0.f;
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) s = s + in[ix3(2 * i + a, 2 * j + b, 2 * k + c, n)];
  out[q] = s;
}

// x = x / sum * ng^3 in FP32 with the FP64 total (mean 1).
__global__ void k_scale_mean1(float* __restrict__ x, const double* __restrict__ sum, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  x[p] = (float)((double)x[p] / sum[0] * (double)n3);
}

// Column density through the deformed mesh (column_density.f90:35-66): the z displacement of the
// cell faces, dspz = (4-point mean of def on plane k+1) - (same on plane k), sets the deformed
// column [kk1, kk2] of the cells k1..k1+nplot-1 at (i, j); proj is the density integrated over it,
// partial end cells weighted by their covered fraction, divided by nplot.  Fortran indices
// (1-based) in the arithmetic; planes outside 1..ng wrap periodically.
__global__ void k_column_density(float* __restrict__ proj, const float* __restrict__ den, const float* __restrict__ def,
                                 int ng, int nplot) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)ng * ng) return;
  const int i = (int)(q % ng) + 1, j = (int)(q / ng) + 1;  // 1-based
  auto w = [&](int a) { return ((a - 1) % ng + ng) % ng; };  // 1-based index -> 0-based, periodic
  auto D = [&](int a, int b, int c) { return def[ix3(w(a), w(b), w(c), ng)]; };
  auto dspz = [&](int a, int b, int c) {
    const int ap = a % ng + 1, bp = b % ng + 1, cp = c % ng + 1;
    float v = D(a, b, cp) + D(ap, b, cp) + D(a, bp, cp) + D(ap, bp, cp) - D(a, b, c) - D(ap, b, c) - D(a, bp, c) -
              D(ap, bp, c);
    return v / 4.f;
  };
  auto den_ = [&](int c) { return den[ix3(i - 1, j - 1, w(c), ng)]; };
  const int k1 = ng / 2, k2 = k1 + nplot - 1;
  const int jm = ((j - 2) % ng + ng) % ng + 1, im = ((i - 2) % ng + ng) % ng + 1, k1m = ((k1 - 2) % ng + ng) % ng + 1;
  const float kk1 = (float)(k1 - 1) + (dspz(im, jm, k1m) + dspz(i, jm, k1m) + dspz(im, j, k1m) + dspz(i, j, k1m)) / 4.f;
  const float kk2 = (float)k2 + (dspz(im, jm, k2) + dspz(i, jm, k2) + dspz(im, j, k2) + dspz(i, j, k2)) / 4.f;
  const int nk1 = (int)floorf(kk1) + 1, nk2 = (int)floorf(kk2) + 1;
  const float r1 = (float)nk1 - kk1, r2 = kk2 - (float)nk2 + 1.f;
  float s1 = 0.f, s2 = 0.f;
  for (int c = nk1 + 1; c <= nk2 - 1; ++c) s1 = s1 + den_(c);  // sum(den(i,j,nk1+1:nk2-1))
  for (int c = nk2; c <= nk1; ++c) s2 = s2 + den_(c);          // sum(den(i,j,nk2:nk1)), empty unless folded
  const float v = r1 * den_(nk1) + s1 + r2 * den_(nk2) - s2;
  proj[q] = v / (float)nplot;
}

}  // namespace mmh
