// kernels_wave.hpp -- one whole 8-colour Gauss-Seidel sweep per launch (k-wavefront, LDS tiles).
//
// The colour-per-launch smoother (k_colour) streams HBM near its peak but moves about five times
// the algorithmic bytes: every launch updates one sub-lattice and re-reads phi of six others and
// the metric at the point and its face neighbours.  Here a workgroup owns an (i, j) tile and a
// segment of k, keeps four planes of phi in LDS and applies all eight colours to its segment in
// one pass, so phi, rhs and G are read about once per sweep.
//
// Order.  A sweep updates colours 0..7 (col = 4 kpar + 2 io + jo; reversed for the symmetric
// V-cycle's post-smoothing).  The planes of the first parity p1 depend only on the old values of
// the other planes, the planes of the second parity p2 on the new values of the p1 planes, and
// within a plane the four in-plane colours follow each other.  So the segment is walked as a
// wavefront: p1 plane t+1 (all four colours), then p2 plane t (all four colours), which applies
// exactly the operations of the colour-by-colour sweep in an order with the same dependencies;
// every update is the same expression on the same values, so the result is bit-identical.
//
// Halos.  Colour number s of the sweep (s = 0..7 in update order) is computed on the tile grown
// by EXT_I[s] points in i and EXT_J[s] in j, the cone of points that later colours read; old phi is
// loaded on the tile grown by 4 and 6.  The segment computes one extra p1 plane above its top
// (the next segment's first), read-only.  Old phi comes from `in`, new phi goes to `out` (two
// buffers), so concurrent workgroups never read each other's results.
#pragma once
#include "kernels.hpp"

namespace mmh {

constexpr int WAVE_HI = 4, WAVE_HJ = 6;
__device__ constexpr int WAVE_EXT_I[8] = {3, 3, 2, 2, 1, 1, 0, 0};
__device__ constexpr int WAVE_EXT_J[8] = {5, 4, 3, 2, 3, 2, 1, 0};

// The update of relaxplane (mmh/gauss1.f90:954-980) with phi and G behind accessors; the
// expression is the one of stencil19, term for term.
template <class T, class PHI, class GF>
__device__ __forceinline__ T wave_update(PHI ph, GF g, T rhs_dx2) {
  T b11p = (g(0, 1, 0, 0) + g(0, 0, 0, 0)) / T(2);
  T b11m = (g(0, -1, 0, 0) + g(0, 0, 0, 0)) / T(2);
  T b22p = (g(3, 0, 1, 0) + g(3, 0, 0, 0)) / T(2);
  T b22m = (g(3, 0, -1, 0) + g(3, 0, 0, 0)) / T(2);
  T b33p = (g(5, 0, 0, 1) + g(5, 0, 0, 0)) / T(2);
  T b33m = (g(5, 0, 0, -1) + g(5, 0, 0, 0)) / T(2);
  T tdiag = -(b11p + b11m + b22p + b22m + b33p + b33m);
  T t1 = (b11p * ph(1, 0, 0) + b11m * ph(-1, 0, 0) + b22p * ph(0, 1, 0) + b22m * ph(0, -1, 0) + b33p * ph(0, 0, 1) +
          b33m * ph(0, 0, -1));
  T t2 = ((ph(1, 1, 0) - ph(-1, 1, 0)) * g(1, 0, 1, 0)
        - (ph(1, -1, 0) - ph(-1, -1, 0)) * g(1, 0, -1, 0)
        + (ph(1, 1, 0) - ph(1, -1, 0)) * g(1, 1, 0, 0)
        - (ph(-1, 1, 0) - ph(-1, -1, 0)) * g(1, -1, 0, 0)
        + (ph(0, 1, 1) - ph(0, -1, 1)) * g(4, 0, 0, 1)
        - (ph(0, 1, -1) - ph(0, -1, -1)) * g(4, 0, 0, -1)
        + (ph(0, 1, 1) - ph(0, 1, -1)) * g(4, 0, 1, 0)
        - (ph(0, -1, 1) - ph(0, -1, -1)) * g(4, 0, -1, 0)
        + (ph(1, 0, 1) - ph(-1, 0, 1)) * g(2, 0, 0, 1)
        - (ph(1, 0, -1) - ph(-1, 0, -1)) * g(2, 0, 0, -1)
        + (ph(1, 0, 1) - ph(1, 0, -1)) * g(2, 1, 0, 0)
        - (ph(-1, 0, 1) - ph(-1, 0, -1)) * g(2, -1, 0, 0)) / T(4);
  return (rhs_dx2 - t1 - t2) / tdiag;
}

// One sweep of the level grid phi (sub-lattice layout) from `in` into `out`.  Grid: tilesI *
// tilesJ * (n / (2 K)) workgroups; K planes of each parity per segment; LDS 4 planes of
// (TI + 8)(TJ + 12) values.  Requires n % TI == n % TJ == n % (2 K) == 0.
template <class T, int TI, int TJ, bool REV, int WAVE_P>
__global__ __launch_bounds__(512) void k_sweep_wave(T* __restrict__ out, const T* __restrict__ in,
                                                    const T* __restrict__ rhs, const T* __restrict__ G, int n, T dx2,
                                                    int K) {
  constexpr int RI = TI + 2 * WAVE_HI, RJ = TJ + 2 * WAVE_HJ, PL = RI * RJ;
  extern __shared__ unsigned char wave_smem[];
  T* lds = reinterpret_cast<T*>(wave_smem);
  const int m = n - 1, h = n >> 1;
  const i64 h3 = (i64)h * h * h, n3 = (i64)n * n * n;
  const int tilesI = n / TI, tilesJ = n / TJ;
  int b = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int ti = b % tilesI;
  b /= tilesI;
  const int tj = b % tilesJ;
  const int seg = b / tilesJ;
  const int ilo = ti * TI - WAVE_HI, jlo = tj * TJ - WAVE_HJ;  // both even
  const int p1 = REV ? 1 : 0;
  const int s0 = seg * 2 * K + p1;  // first plane of parity p1
  const int nt = blockDim.x, tid = threadIdx.x;

  auto gidx = [&](int i, int j, int k) -> i64 { return lx<1>(i & m, j & m, k & m, n); };
  // old phi of plane k into LDS slot s, class by class so that reads are unit-stride
  auto load = [&](int s, int k) {
    T* dst = lds + s * PL;
    const int kk = k & m;
    for (int e = tid; e < PL; e += nt) {
      const int a = e % (RI / 2), r2 = e / (RI / 2);
      const int io = r2 & 1, lj = r2 >> 1;
      const int li = 2 * a + io;
      dst[lj * RI + li] = in[gidx(ilo + li, jlo + lj, kk)];
    }
  };
  auto store = [&](int s, int k) {
    const T* src = lds + s * PL;
    const int kk = k & m;
    for (int e = tid; e < TI * TJ; e += nt) {
      const int a = e % (TI / 2), r2 = e / (TI / 2);
      const int io = r2 & 1, lj = WAVE_HJ + (r2 >> 1);
      const int li = WAVE_HI + 2 * a + io;
      out[gidx(ilo + li, jlo + lj, kk)] = src[lj * RI + li];
    }
  };
  // the four colours of plane k (slot s, neighbours in slots sm = k-1 and sp = k+1); step0 is the
  // sweep position of its first colour (0 for the p1 planes, 4 for the p2 planes)
  auto plane = [&](int s, int sm, int sp, int k, int step0) {
    const int kk = k & m;
    for (int o = 0; o < 4; ++o) {
      const int st = step0 + o;
      const int io = REV ? 1 - (o >> 1) : (o >> 1), jo = REV ? 1 - (o & 1) : (o & 1);
      const int ei = WAVE_EXT_I[st], ej = WAVE_EXT_J[st];
      const int lis = WAVE_HI - ei, lie = WAVE_HI + TI - 1 + ei;
      const int ljs = WAVE_HJ - ej, lje = WAVE_HJ + TJ - 1 + ej;
      const int li0 = lis + ((io - lis) & 1), lj0 = ljs + ((jo - ljs) & 1);
      const int ci = (lie - li0) / 2 + 1, cj = (lje - lj0) / 2 + 1;
      const int cnt = ci * cj;
      // points of one colour are independent: a thread computes WAVE_P of them before storing
      // any, so their global loads can be in flight together
      for (int base = tid; base < cnt; base += nt * WAVE_P) {
        T v[WAVE_P];
        int at[WAVE_P];
#pragma unroll
        for (int r = 0; r < WAVE_P; ++r) {
          const int e = base + r * nt;
          at[r] = -1;
          if (e < cnt) {
            const int li = li0 + 2 * (e % ci), lj = lj0 + 2 * (e / ci);
            const int i = ilo + li, j = jlo + lj;
            auto ph = [&](int di, int dj, int dk) -> T {
              const int sl = dk == 0 ? s : (dk < 0 ? sm : sp);
              return lds[sl * PL + (lj + dj) * RI + (li + di)];
            };
            auto g = [&](int comp, int di, int dj, int dk) -> T {
              return G[(i64)comp * n3 + gidx(i + di, j + dj, kk + dk)];
            };
            v[r] = wave_update<T>(ph, g, rhs[gidx(i, j, kk)] * dx2);
            at[r] = lj * RI + li;
          }
        }
#pragma unroll
        for (int r = 0; r < WAVE_P; ++r)
          if (at[r] >= 0) lds[s * PL + at[r]] = v[r];
      }
      __syncthreads();
    }
  };
  (void)h3;
  // prologue: p1 plane 0 from the old p2 planes on either side
  int sA = 0, sB = 1, sC = 2, sD = 3;  // p1(t), p2(t), p1(t+1), p2(t+1)
  load(sA, s0);
  load(sB, s0 + 1);
  load(sC, s0 - 1);
  __syncthreads();
  plane(sA, sC, sB, s0, 0);
  for (int t = 0; t < K; ++t) {
    const int k1 = s0 + 2 * t;  // p1(t); p2(t) = k1 + 1; p1(t+1) = k1 + 2; p2(t+1) = k1 + 3
    load(sC, k1 + 2);
    load(sD, k1 + 3);
    __syncthreads();
    plane(sC, sB, sD, k1 + 2, 0);
    plane(sB, sA, sC, k1 + 1, 4);
    store(sA, k1);
    store(sB, k1 + 1);
    __syncthreads();
    const int oa = sA, ob = sB;
    sA = sC;
    sB = sD;
    sC = oa;
    sD = ob;
  }
}

}  // namespace mmh
