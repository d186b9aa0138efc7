// kernels_march.hpp -- residual and operator application of a level by k-marching workgroups.
//
// The cell-lockstep kernels (k_resid, k_apply_dot) read every phi value about 19 times and every
// metric value about 3 times; the L2 of an XCD holds only a few rows of a 512^3 level, so the
// reads of the k +- 1 planes come from HBM again and the kernels move about 3.8x the bytes they
// need.  Here a workgroup owns a TI x TJ column of points and walks KS consecutive planes of it:
//   - phi of four consecutive planes (with a one-point halo in i and j) sits in an LDS ring, and
//     each step loads one new plane, so phi is fetched from HBM about once;
//   - G3, G5 and G6 at the thread's own point of planes k-1, k, k+1 are carried in registers,
//     so the k-direction metric reads are not repeated;
//   - G1..G5 of the current plane (own points plus the halo values the in-plane neighbours need)
//     are staged in LDS, so each metric value is loaded once;
//   - loads run one plane ahead of the arithmetic (software pipeline), so a load has a whole
//     step to arrive.
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
template <class T, int MODE, int TI, int TJ, class TG = T>
__global__ __launch_bounds__(TI * TJ) void k_op_march(T* __restrict__ out, double* __restrict__ bres,
                                                      const T* __restrict__ phi, const T* __restrict__ rhs,
                                                      const TG* __restrict__ G, int n, T di2, int KS, int batched, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  constexpr int NT = TI * TJ, RI = TI + 2, RJ = TJ + 2, PL = RI * RJ, NH = PL - NT;
  static_assert(NH <= NT, "one halo point per thread at most");
  static_assert(4 * PL * sizeof(T) >= NT * sizeof(double), "the block reduction reuses the phi ring");
  __shared__ __align__(16) T ring[4][PL];
  __shared__ T gs[2][5 * PL];  // G1..G5 of the current plane, double-buffered
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
  // Loads and stores address a plane-uniform base (array + K(k), scalar registers) plus this
  // thread's 32-bit byte offset of the i/j part: IJ < 4 (n/2)^3, so IJ * sizeof(T) < 2^32 for
  // n <= 1024.
  auto BO = [&](i64 e) -> uint32_t { return (uint32_t)(e * (i64)sizeof(T)); };
  const uint32_t oc = BO(IJ(i, j)), oip = BO(IJ((i + 1) & m, j)), oim = BO(IJ((i - 1) & m, j)),
                 ojp = BO(IJ(i, (j + 1) & m)), ojm = BO(IJ(i, (j - 1) & m));
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
  const uint32_t oh = BO(IJ((i0 + hx) & m, (j0 + hy) & m));
  auto LD = [](const T* base, uint32_t off) -> T { return *(const T*)((const char*)base + off); };
  // the metric's byte offsets (its element type TG may be narrower than T)
  auto BOG = [&](i64 e) -> uint32_t { return (uint32_t)(e * (i64)sizeof(TG)); };
  const uint32_t ocg = BOG(IJ(i, j)), ohg = BOG(IJ((i0 + hx) & m, (j0 + hy) & m));
  auto LDG = [](const TG* base, uint32_t off) -> T { return gl<T>((const TG*)((const char*)base + off), 0); };
  const int own = (ty + 1) * RI + tx + 1, hslot = (hy + 1) * RI + hx + 1;
  const TG* G1 = G;
  const TG* G2 = G + n3;
  const TG* G3 = G + 2 * n3;
  const TG* G4 = G + 3 * n3;
  const TG* G5 = G + 4 * n3;
  const TG* G6 = G + 5 * n3;

  // halo points in the left/right columns supply G1, G2, G3 (i-neighbours); in the top/bottom rows
  // G2, G4, G5 (j-neighbours); corners no metric value
  const bool hcol = tid < NH && hy >= 0 && hy < TJ;
  const bool hrow = tid < NH && !hcol && hx >= 0 && hx < TI;

  // Software pipeline, one plane ahead: step k writes the G1..G5 plane k and the phi plane k+1
  // that step k-1 loaded into registers, then loads G of plane k+1, phi of plane k+2 and G3/G5/G6
  // of plane k+2 at the own point, so a load has a whole step to arrive.
  T nphi = T(0), nphih = T(0);                      // phi of the plane to be stored next
  T n1 = T(0), n2 = T(0), n4 = T(0), nh0 = T(0), nh1 = T(0), nh2 = T(0);  // G of the next plane
#pragma unroll 1
  for (int z = k0 - 1; z <= k0; ++z) {
    const T* pz = phi + K(z);
    T* s = ring[z & 3];
    s[own] = LD(pz, oc);
    if (tid < NH) s[hslot] = LD(pz, oh);
  }
  auto load_next = [&](int kn, int kphi) {  // G1/G2/G4 and halo G of plane kn, phi of plane kphi
    const i64 q = K(kn);
    n1 = LDG(G1 + q, ocg);
    n2 = LDG(G2 + q, ocg);
    n4 = LDG(G4 + q, ocg);
    if (hcol) {
      nh0 = LDG(G1 + q, ohg);
      nh1 = LDG(G2 + q, ohg);
      nh2 = LDG(G3 + q, ohg);
    } else if (hrow) {
      nh0 = LDG(G2 + q, ohg);
      nh1 = LDG(G4 + q, ohg);
      nh2 = LDG(G5 + q, ohg);
    }
    const T* pz = phi + K(kphi);
    nphi = LD(pz, oc);
    if (tid < NH) nphih = LD(pz, oh);
  };
  load_next(k0, k0 + 1);
  T rhs_c = MODE == 0 ? LD(rhs + K(k0), oc) : T(0);  // rhs of the current plane, loaded a step ahead
  const i64 km0 = K(k0 - 1), kc0 = K(k0), kp0 = K(k0 + 1);
  T g3m = LDG(G3 + km0, ocg), g5m = LDG(G5 + km0, ocg), g6m = LDG(G6 + km0, ocg);
  T g3c = LDG(G3 + kc0, ocg), g5c = LDG(G5 + kc0, ocg), g6c = LDG(G6 + kc0, ocg);
  T g3p = LDG(G3 + kp0, ocg), g5p = LDG(G5 + kp0, ocg), g6p = LDG(G6 + kp0, ocg);

  double acc = 0.0;
#pragma unroll 1
  for (int t = 0; t < KS; ++t) {
    const int k = k0 + t;
    const i64 kc = K(k);
    // plane k of G1..G5 into buffer k & 1 (last read in step k - 2); phi plane k + 1 into ring slot
    // (k + 1) & 3 (last read in step k - 2)
    T* gb = &gs[k & 1][0];
    gb[0 * PL + own] = n1;
    gb[1 * PL + own] = n2;
    gb[2 * PL + own] = g3c;
    gb[3 * PL + own] = n4;
    gb[4 * PL + own] = g5c;
    if (hcol) {
      gb[0 * PL + hslot] = nh0;
      gb[1 * PL + hslot] = nh1;
      gb[2 * PL + hslot] = nh2;
    } else if (hrow) {
      gb[1 * PL + hslot] = nh0;
      gb[3 * PL + hslot] = nh1;
      gb[4 * PL + hslot] = nh2;
    }
    {
      T* s = ring[(k + 1) & 3];
      s[own] = nphi;
      if (tid < NH) s[hslot] = nphih;
    }
    const T g1c = n1, g4c = n4;
    // loads for the next step
    T g3n = T(0), g5n = T(0), g6n = T(0), rhs_n = T(0);
    if (t + 1 < KS) {
      load_next(k + 1, k + 2);
      const i64 k2 = K(k + 2);
      g3n = LDG(G3 + k2, ocg);
      g5n = LDG(G5 + k2, ocg);
      g6n = LDG(G6 + k2, ocg);
      if (MODE == 0) rhs_n = LD(rhs + K(k + 1), oc);
    }
    __syncthreads();
    const T* sm = ring[(k - 1) & 3];
    const T* sc = ring[k & 3];
    const T* sp = ring[(k + 1) & 3];
#define PH_(s, dx, dy) s[own + (dy) * RI + (dx)]
#define GG_(comp, dx, dy) gb[(comp) * PL + own + (dy) * RI + (dx)]
    const T g1ip = GG_(0, 1, 0), g1im = GG_(0, -1, 0);
    const T g4jp = GG_(3, 0, 1), g4jm = GG_(3, 0, -1);
    const T g2jp = GG_(1, 0, 1), g2jm = GG_(1, 0, -1), g2ip = GG_(1, 1, 0), g2im = GG_(1, -1, 0);
    const T g5jp = GG_(4, 0, 1), g5jm = GG_(4, 0, -1);
    const T g3ip = GG_(2, 1, 0), g3im = GG_(2, -1, 0);
#undef GG_
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
      const T rv = rhs_c - (t1 + t2 + tdiag * pc) * di2;
      if (out) *(T*)((char*)(out + kc) + oc) = rv;
      acc = dmax(acc, (double)fabs(rv));
    } else {
      const T v = (t1 + t2 + tdiag * pc) * di2;
      *(T*)((char*)(out + kc) + oc) = v;
      acc = acc + (double)pc * (double)v;
    }
    g3m = g3c; g3c = g3p; g3p = g3n;
    g5m = g5c; g5c = g5p; g5p = g5n;
    g6m = g6c; g6c = g6p; g6p = g6n;
    rhs_c = rhs_n;
  }
  __syncthreads();  // the last step's reads of the ring are done: it becomes the reduction's scratch
  acc = tree_reduce<NT, MODE == 0 ? 1 : 0>(acc, reinterpret_cast<double*>(&ring[0][0]));
  if (tid == 0) bres[b] = acc;
}

}  // namespace mmh
