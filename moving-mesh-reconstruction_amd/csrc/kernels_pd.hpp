// kernels_pd.hpp -- particle to grid density assignment of haloden.f90 (mmh's input generation
// from N-body particles): nearest-grid-point counting plus a nearest-particle fill of empty cells,
// each particle's unit mass shared equally between its own cell and the empty cells it fills.
#pragma once
#include "kernels.hpp"

namespace mmh {

// Position of particle p in grid units and its cell (haloden.f90:84-99): x = modulo(x, nf) * ng / nf,
// cell = floor(x) (0-based; clamped to ng-1 where the scaled position rounds up to ng).
__global__ void k_pd_cells(float* __restrict__ xs, int* __restrict__ cell, int* __restrict__ idx,
                           const float* __restrict__ xp, i64 np, int ng, float nf) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int c[3];
  for (int d = 0; d < 3; ++d) {
    float x = xp[3 * p + d];
    x = x - floorf(x / nf) * nf;  // modulo(x, real(nf))
    x = x * (float)ng / nf;
    xs[3 * p + d] = x;
    int ci = (int)floorf(x);
    c[d] = ci < 0 ? 0 : (ci > ng - 1 ? ng - 1 : ci);
  }
  cell[p] = c[0] + ng * (c[1] + ng * c[2]);
  idx[p] = (int)p;
}

// [start, end) of every non-empty cell in the cell-sorted particle list.
__global__ void k_pd_ranges(int* __restrict__ start, int* __restrict__ end, const int* __restrict__ key, i64 np) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int k = key[q];
  if (q == 0 || key[q - 1] != k) start[k] = (int)q;
  if (q == np - 1 || key[q + 1] != k) end[k] = (int)q + 1;
}

// Nearest particle of every empty cell (haloden.f90:119-167): cubes of growing half-width r
// around the cell are scanned in (k, j, i) order, each cell's particles in the reference's
// linked-list order (newest first, i.e. decreasing index); the first particle with the smallest
// periodic squared distance to the cell centre wins, and the scan stops after the first cube that
// holds any particle.  near[c] = particle index, or -1 for a non-empty cell; wp[p] counts cells.
__global__ void k_pd_nearest(int* __restrict__ near, int* __restrict__ wp, const int* __restrict__ start,
                             const int* __restrict__ end, const int* __restrict__ order, const float* __restrict__ xs,
                             int ng, int rmax) {
  const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (i64)ng * ng * ng) return;
  if (end[c] > start[c]) { near[c] = -1; return; }
  const int i = (int)(c % ng), j = (int)((c / ng) % ng), k = (int)(c / ((i64)ng * ng));
  // 1-based cell (i+1, j+1, k+1) has centre (i+1, j+1, k+1) - 0.5
  const float gx = (float)(i + 1) - 0.5f, gy = (float)(j + 1) - 0.5f, gz = (float)(k + 1) - 0.5f;
  const float half = (float)(ng / 2), fng = (float)ng;
  float r2min = (float)(1000 * ng * ng);
  int best = -1;
  for (int r = 1; r <= rmax && best < 0; ++r) {
    for (int ks = k - r; ks <= k + r; ++ks)
      for (int js = j - r; js <= j + r; ++js)
        for (int is = i - r; is <= i + r; ++is) {
          const int kk = ((ks % ng) + ng) % ng, jj = ((js % ng) + ng) % ng, ii = ((is % ng) + ng) % ng;
          const i64 cc = ii + (i64)ng * (jj + (i64)ng * kk);
          for (int q = end[cc] - 1; q >= start[cc]; --q) {
            const int pp = order[q];
            float d[3] = {xs[3 * (i64)pp] - gx, xs[3 * (i64)pp + 1] - gy, xs[3 * (i64)pp + 2] - gz};
            for (int a = 0; a < 3; ++a) {
              const float t = d[a] + half;  // modulo(dpos + ng/2, real(ng)) - ng/2
              d[a] = (t - floorf(t / fng) * fng) - half;
            }
            const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < r2min) { r2min = r2; best = pp; }
          }
        }
  }
  near[c] = best;
  if (best >= 0) atomicAdd(&wp[best], 1);
}

// Density of every cell (haloden.f90:170-181): a non-empty cell sums 1/wp of its particles in
// increasing index order; an empty cell holds 1/wp of the particle that filled it.
__global__ void k_pd_density(float* __restrict__ den, const int* __restrict__ near, const int* __restrict__ wp,
                             const int* __restrict__ start, const int* __restrict__ end, const int* __restrict__ order,
                             i64 ncell) {
  const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float s = 0.f;
  if (near[c] >= 0) s = s + 1.f / (float)wp[near[c]];
  else
    for (int q = start[c]; q < end[c]; ++q) s = s + 1.f / (float)wp[order[q]];
  den[c] = s;
}

// den = real(den, 8) / sum * float(ng)**3 (haloden.f90:187)
__global__ void k_pd_normalise(float* __restrict__ den, const double* __restrict__ sum, float ng3, i64 ncell) {
  const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  den[c] = (float)((double)den[c] / sum[0] * (double)ng3);
}

}  // namespace mmh
