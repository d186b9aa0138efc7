// kernels_march.hpp -- residual and operator application of a level by k-marching workgroups.
//
// The cell-lockstep kernels (k_resid, k_apply_dot) read every phi value about 19 times and every
// metric value about 3 times; the L2 of an XCD holds only a few rows of a 512^3 level, so the
// reads of the k +- 1 planes come from HBM again and the kernels move about 3.8x the bytes they
// need.  Here a workgroup owns a TI x TJ column of points and walks KS consecutive planes of it:
//   - phi of four consecutive planes (with a one-point halo in i and j) sits in an LDS ring, and
//     each step loads one new plane, so phi is fetched from HBM about once;
//   - G3, G5 and G6 at the thread's own point of planes k-1, k, k+1 are carried in registers
//     (each step loads plane k+1), so the k-direction metric reads are not repeated;
//   - the in-plane neighbours' metric values are the ones the neighbouring threads load in the
//     same step, so those reads hit the caches.
// Every point evaluates the expression of stencil19 term for term (the same operands in the same
// order), so r and q are bit-identical to the cell-lockstep kernels'.  The per-block results
// (max |r|, or the sum of p*q over the thread's points in plane order and then the halving tree)
// are written at the block's tile index.
#pragma once
#include "kernels.hpp"

namespace mmh {

// Grid: (n/TI) * (n/TJ) * (n/KS) workgroups of TI*TJ threads, tiles in i fastest, then j, then
// the k segment (XCD-remapped so that neighbouring tiles share an L2); gridDim.y = B grids of a
// batch when batched.  MODE 0: out = rhs - L phi (out may be null), bres = max |out|.  MODE 1:
// out = L phi, bres = sum phi * out.  Requires n % TI == n % TJ == n % KS == 0, TI >= 2, TJ >= 2,
// and 2 (TI + TJ) + 4 <= TI * TJ (one halo point per thread at most).
template <class T, int MODE, int TI, int TJ>
__global__ __launch_bounds__(TI * TJ) void k_op_march(T* __restrict__ out, double* __restrict__ bres,
                                                      const T* __restrict__ phi, const T* __restrict__ rhs,
                                                      const T* __restrict__ G, int n, T di2, int KS, int batched) {
  constexpr int NT = TI * TJ, RI = TI + 2, RJ = TJ + 2, PL = RI * RJ, NH = PL - NT;
  static_assert(NH <= NT, "one halo point per thread at most");
  __shared__ T ring[4][PL];
  __shared__ double red[NT];
  const i64 n3 = (i64)n * n * n;
  if (batched) {
    const i64 y = blockIdx.y;
    if (out) out += y * n3;
    phi += y * n3;
    if (MODE == 0) rhs += y * n3;
    G += y * 6 * n3;
    bres += y * gridDim.x;
  }
  const int m = n - 1, h = n >> 1;
  const i64 hh = (i64)h * h, h3 = hh * h;
  const int tilesI = n / TI, tilesJ = n / TJ;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % tilesI), tj = (int)((b / tilesI) % tilesJ), seg = (int)(b / ((unsigned)tilesI * tilesJ));
  const int i0 = ti * TI, j0 = tj * TJ, k0 = seg * KS;
  const int tid = threadIdx.x, tx = tid % TI, ty = tid / TI;
  const int i = i0 + tx, j = j0 + ty;
  // sub-lattice offset = IJ(i, j) + K(k): the i/j part (class bits a + 2b and the column) ...
  auto IJ = [&](int x, int y) -> i64 { return (i64)((x & 1) + 2 * (y & 1)) * h3 + (x >> 1) + (i64)h * (y >> 1); };
  // ... and the k part (class bit 4e and the plane)
  auto K = [&](int z) -> i64 {
    z &= m;
    return (i64)(4 * (z & 1)) * h3 + hh * (z >> 1);
  };
  const i64 oc = IJ(i, j), oip = IJ((i + 1) & m, j), oim = IJ((i - 1) & m, j), ojp = IJ(i, (j + 1) & m),
            ojm = IJ(i, (j - 1) & m);
  // this thread's halo point (tid < NH): top row, bottom row, then the left/right columns
  int hx = 0, hy = 0;
  if (tid < NH) {
    if (tid < RI) {
      hx = tid - 1;
      hy = -1;
    } else if (tid < 2 * RI) {
      hx = tid - RI - 1;
      hy = TJ;
    } else {
      const int r = tid - 2 * RI;
      hy = r >> 1;
      hx = (r & 1) ? TI : -1;
    }
  }
  const i64 oh = IJ((i0 + hx) & m, (j0 + hy) & m);
  const int own = (ty + 1) * RI + tx + 1, hslot = (hy + 1) * RI + hx + 1;
  const T* G1 = G;
  const T* G2 = G + n3;
  const T* G3 = G + 2 * n3;
  const T* G4 = G + 3 * n3;
  const T* G5 = G + 4 * n3;
  const T* G6 = G + 5 * n3;

  // prologue: phi of planes k0-1, k0, k0+1; G3, G5, G6 of planes k0-1 and k0 at the own point
#pragma unroll 1
  for (int z = k0 - 1; z <= k0 + 1; ++z) {
    const i64 kz = K(z);
    T* s = ring[z & 3];
    s[own] = phi[oc + kz];
    if (tid < NH) s[hslot] = phi[oh + kz];
  }
  const i64 km0 = K(k0 - 1), kc0 = K(k0);
  T g3m = G3[oc + km0], g5m = G5[oc + km0], g6m = G6[oc + km0];
  T g3c = G3[oc + kc0], g5c = G5[oc + kc0], g6c = G6[oc + kc0];
  __syncthreads();

  double acc = 0.0;
#pragma unroll 1
  for (int t = 0; t < KS; ++t) {
    const int k = k0 + t;
    const i64 kc = K(k), kp = K(k + 1);
    // the phi plane two ahead, stored after this step's reads of the ring
    T nxo = T(0), nxh = T(0);
    const bool more = t + 1 < KS;
    if (more) {
      const i64 k2 = K(k + 2);
      nxo = phi[oc + k2];
      if (tid < NH) nxh = phi[oh + k2];
    }
    const T g3p = G3[oc + kp], g5p = G5[oc + kp], g6p = G6[oc + kp];
    const T g1c = G1[oc + kc], g4c = G4[oc + kc];
    const T g1ip = G1[oip + kc], g1im = G1[oim + kc];
    const T g4jp = G4[ojp + kc], g4jm = G4[ojm + kc];
    const T g2jp = G2[ojp + kc], g2jm = G2[ojm + kc], g2ip = G2[oip + kc], g2im = G2[oim + kc];
    const T g5jp = G5[ojp + kc], g5jm = G5[ojm + kc];
    const T g3ip = G3[oip + kc], g3im = G3[oim + kc];
    const T* sm = ring[(k - 1) & 3];
    const T* sc = ring[k & 3];
    const T* sp = ring[(k + 1) & 3];
#define PH_(s, dx, dy) s[own + (dy) * RI + (dx)]
    // stencil19 (mmh/gauss1.f90:954-978), operands and order unchanged
    T b11p = (g1ip + g1c) / T(2);
    T b11m = (g1im + g1c) / T(2);
    T b22p = (g4jp + g4c) / T(2);
    T b22m = (g4jm + g4c) / T(2);
    T b33p = (g6p + g6c) / T(2);
    T b33m = (g6m + g6c) / T(2);
    T tdiag = -(b11p + b11m + b22p + b22m + b33p + b33m);
    T t1 = (b11p * PH_(sc, 1, 0) + b11m * PH_(sc, -1, 0) + b22p * PH_(sc, 0, 1) + b22m * PH_(sc, 0, -1) +
            b33p * PH_(sp, 0, 0) + b33m * PH_(sm, 0, 0));
    T t2 = ((PH_(sc, 1, 1) - PH_(sc, -1, 1)) * g2jp
          - (PH_(sc, 1, -1) - PH_(sc, -1, -1)) * g2jm
          + (PH_(sc, 1, 1) - PH_(sc, 1, -1)) * g2ip
          - (PH_(sc, -1, 1) - PH_(sc, -1, -1)) * g2im
          + (PH_(sp, 0, 1) - PH_(sp, 0, -1)) * g5p
          - (PH_(sm, 0, 1) - PH_(sm, 0, -1)) * g5m
          + (PH_(sp, 0, 1) - PH_(sm, 0, 1)) * g5jp
          - (PH_(sp, 0, -1) - PH_(sm, 0, -1)) * g5jm
          + (PH_(sp, 1, 0) - PH_(sp, -1, 0)) * g3p
          - (PH_(sm, 1, 0) - PH_(sm, -1, 0)) * g3m
          + (PH_(sp, 1, 0) - PH_(sm, 1, 0)) * g3ip
          - (PH_(sp, -1, 0) - PH_(sm, -1, 0)) * g3im) / T(4);
    const T pc = PH_(sc, 0, 0);
#undef PH_
    if (MODE == 0) {
      const T rv = rhs[oc + kc] - (t1 + t2 + tdiag * pc) * di2;
      if (out) out[oc + kc] = rv;
      acc = dmax(acc, (double)fabs(rv));
    } else {
      const T v = (t1 + t2 + tdiag * pc) * di2;
      out[oc + kc] = v;
      acc = acc + (double)pc * (double)v;
    }
    // slot (k + 2) & 3 was last read in step k - 1, which every thread has left
    if (more) {
      T* s = ring[(k + 2) & 3];
      s[own] = nxo;
      if (tid < NH) s[hslot] = nxh;
    }
    g3m = g3c; g3c = g3p;
    g5m = g5c; g5c = g5p;
    g6m = g6c; g6c = g6p;
    __syncthreads();
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = MODE == 0 ? dmax(red[tid], red[tid + s]) : red[tid] + red[tid + s];
    __syncthreads();
  }
  if (tid == 0) bres[b] = red[0];
}

}  // namespace mmh
