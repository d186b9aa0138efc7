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

// Column density through the deformed mesh (column_density.f90:26-57): the z displacement of the
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

// gridfigure.f90:41-131 on the device.  The coarse mesh (nf^3 deformation potential) gets
// cell-corner displacements in t = tmp_dsp(3, 0:nf, 0:nf, 0:nf) (offset c + 3(a + E(b + E d)),
// E = nf + 1, the reference's column-major storage), a ghost layer, and a trilinear interpolation
// onto the ng^3 density grid (ratio = ng / nf).  Fortran (1-based) indices in the arithmetic.

// Displacements of the coarse cell corners (gridfigure.f90:42-56): ratio/4 times the sum of the
// four def differences across the cell along each axis; t is zero beforehand (the ghost layer).
__global__ void k_grid_corner_dsp(float* __restrict__ t, const float* __restrict__ def, int nf, int ratio) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)nf * nf * nf) return;
  const int i = (int)(q % nf) + 1, j = (int)((q / nf) % nf) + 1, k = (int)(q / ((i64)nf * nf)) + 1;
  const int E = nf + 1;
  const int ip = i % nf + 1, jp = j % nf + 1, kp = k % nf + 1;
  auto D = [&](int a, int b, int c) { return def[ix3(a - 1, b - 1, c - 1, nf)]; };
  const float s4 = (float)ratio * 1.f / 4.f;
  float* o = t + 3 * ((i64)i + (i64)E * ((i64)j + (i64)E * k));
  o[0] = s4 * (D(ip, j, k) - D(i, j, k) + D(ip, jp, k) - D(i, jp, k) + D(ip, j, kp) - D(i, j, kp) + D(ip, jp, kp) -
               D(i, jp, kp));
  o[1] = s4 * (D(i, jp, k) - D(i, j, k) + D(ip, jp, k) - D(ip, j, k) + D(i, jp, kp) - D(i, j, kp) + D(ip, jp, kp) -
               D(ip, j, kp));
  o[2] = s4 * (D(i, j, kp) - D(i, j, k) + D(ip, j, kp) - D(ip, j, k) + D(i, jp, kp) - D(i, jp, k) + D(ip, jp, kp) -
               D(ip, jp, k));
}

// The reference's dsp_proj.dat (gridfigure.f90:57-59): tmp_dsp(1:2, :, :, nf/2:nf/2+3) before the
// ghost layer is filled.
__global__ void k_grid_dsp_slab(float* __restrict__ out, const float* __restrict__ t, int nf) {
  const int E = nf + 1;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)E * E * 4) return;
  const int a = (int)(q % E), b = (int)((q / E) % E), d = (int)(q / ((i64)E * E)) + nf / 2;
  const float* s = t + 3 * ((i64)a + (i64)E * ((i64)b + (i64)E * d));
  out[2 * q] = s[0];
  out[2 * q + 1] = s[1];
}

// Ghost layer (gridfigure.f90:61-67).  tmp = tmp_dsp shifted by one along all three axes at once
// (cshift over the 0:nf extents), and the faces i = 0 (j, k >= 1), k = 0 (i >= 1) and j = 0
// (i, k >= 1) take tmp's values: the ghost at (a, b, d) holds the corner (a-1, b-1, d-1) mod E of
// the unfilled array, which is zero wherever that corner is itself a ghost.  One thread per ghost
// cell of the three faces; every read is of an interior corner, which no thread writes.
__global__ void k_grid_ghosts(float* __restrict__ t, int nf) {
  const int E = nf + 1;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)E * E) return;
  const int x = (int)(q % E), y = (int)(q / E);
  auto get = [&](int a, int b, int d) -> float3 {
    a = (a + E - 1) % E; b = (b + E - 1) % E; d = (d + E - 1) % E;
    if (a == 0 || b == 0 || d == 0) return make_float3(0.f, 0.f, 0.f);
    const float* s = t + 3 * ((i64)a + (i64)E * ((i64)b + (i64)E * d));
    return make_float3(s[0], s[1], s[2]);
  };
  auto put = [&](int a, int b, int d, float3 v) {
    float* o = t + 3 * ((i64)a + (i64)E * ((i64)b + (i64)E * d));
    o[0] = v.x; o[1] = v.y; o[2] = v.z;
  };
  // face i = 0: (0, b, d), b, d in 1..nf
  float3 f0 = make_float3(0.f, 0.f, 0.f), f1 = f0, f2 = f0;
  const bool w0 = x >= 1 && y >= 1, w1 = x >= 1, w2 = x >= 1 && y >= 1;
  if (w0) f0 = get(0, x, y);
  // face k = 0: (a, b, 0), a in 1..nf, b in 0..nf
  if (w1) f1 = get(x, y, 0);
  // face j = 0: (a, 0, d), a in 1..nf, d in 1..nf
  if (w2) f2 = get(x, 0, y);
  if (w0) put(0, x, y, f0);
  if (w1) put(x, y, 0, f1);
  if (w2) put(x, 0, y, f2);
}

// Trilinear interpolation of component c at fine point (i, j, k) (gridfigure.f90:73-94): eight
// corners in the reference's order; corner index nf + 1 reaches the next row's ghost through the
// linear offset, as the reference's storage does.
__device__ inline float grid_interp(const float* __restrict__ t, int nf, int ratio, int c, int i, int j, int k) {
  const i64 E = nf + 1;
  const float ib = (float)(i % ratio) / (float)ratio, jb = (float)(j % ratio) / (float)ratio,
              kb = (float)(k % ratio) / (float)ratio;
  const i64 ip = (i64)floorf((float)i / (float)ratio) + 1 - (i64)floorf(ib);
  const i64 jp = (i64)floorf((float)j / (float)ratio) + 1 - (i64)floorf(jb);
  const i64 kp = (i64)floorf((float)k / (float)ratio) + 1 - (i64)floorf(kb);
  auto T = [&](i64 a, i64 b, i64 d) { return t[c + 3 * (a + E * (b + E * d))]; };
  const float p111 = ib * jb * kb, p211 = (1 - ib) * jb * kb, p121 = ib * (1 - jb) * kb, p112 = ib * jb * (1 - kb);
  const float p221 = (1 - ib) * (1 - jb) * kb, p212 = (1 - ib) * jb * (1 - kb), p122 = ib * (1 - jb) * (1 - kb);
  const float p222 = (1 - ib) * (1 - jb) * (1 - kb);
  float v = 0.f;
  v = v + T(ip - 1, jp - 1, kp - 1) * p111;
  v = v + T(ip, jp - 1, kp - 1) * p211;
  v = v + T(ip - 1, jp, kp - 1) * p121;
  v = v + T(ip - 1, jp - 1, kp) * p112;
  v = v + T(ip, jp, kp - 1) * p221;
  v = v + T(ip, jp - 1, kp) * p212;
  v = v + T(ip - 1, jp, kp) * p122;
  v = v + T(ip, jp, kp) * p222;
  return v;
}

// New (x, y) positions of the coarse nodes in the fine mid-plane (gridfigure.f90:95-97,121):
// pos(c, mi, mj) = (i or j) + 0.5 + dsp(c, ratio mi, ratio mj, ng/2).
__global__ void k_grid_positions(float* __restrict__ pos, const float* __restrict__ t, int nf, int ratio, int ng) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)nf * nf) return;
  const int mi = (int)(q % nf) + 1, mj = (int)(q / nf) + 1;
  const int i = ratio * mi, j = ratio * mj, k = ng / 2;
  pos[2 * q] = (float)i + 0.5f + grid_interp(t, nf, ratio, 0, i, j, k);
  pos[2 * q + 1] = (float)j + 0.5f + grid_interp(t, nf, ratio, 1, i, j, k);
}

// Column density over the interpolated z displacement (gridfigure.f90:104-120), den already scaled
// to mean 1; the formula of k_column_density with dsp(3) from grid_interp.
__global__ void k_grid_column(float* __restrict__ proj, const float* __restrict__ den, const float* __restrict__ t, int nf,
                              int ratio, int ng, int nplot) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)ng * ng) return;
  const int i = (int)(q % ng) + 1, j = (int)(q / ng) + 1;
  auto w = [&](int a) { return ((a - 1) % ng + ng) % ng; };
  auto den_ = [&](int c) { return den[ix3(i - 1, j - 1, w(c), ng)]; };
  const int k1 = ng / 2, k2 = k1 + nplot - 1;
  const int jm = ((j - 2) % ng + ng) % ng + 1, im = ((i - 2) % ng + ng) % ng + 1, k1m = ((k1 - 2) % ng + ng) % ng + 1;
  auto dz = [&](int a, int b, int c) { return grid_interp(t, nf, ratio, 2, a, b, c); };
  const float kk1 = (float)(k1 - 1) + (dz(im, jm, k1m) + dz(i, jm, k1m) + dz(im, j, k1m) + dz(i, j, k1m)) / 4.f;
  const float kk2 = (float)k2 + (dz(im, jm, k2) + dz(i, jm, k2) + dz(im, j, k2) + dz(i, j, k2)) / 4.f;
  const int nk1 = (int)floorf(kk1) + 1, nk2 = (int)floorf(kk2) + 1;
  const float r1 = (float)nk1 - kk1, r2 = kk2 - (float)nk2 + 1.f;
  float s1 = 0.f, s2 = 0.f;
  for (int c = nk1 + 1; c <= nk2 - 1; ++c) s1 = s1 + den_(c);
  for (int c = nk2; c <= nk1; ++c) s2 = s2 + den_(c);
  proj[q] = (r1 * den_(nk1) + s1 + r2 * den_(nk2) - s2) / (float)nplot;
}

}  // namespace mmh
