// kernels_gal.hpp -- Galerkin coarse operators for the Krylov solve's preconditioner.
//
// The reference rebuilds each coarse level's operator from restricted fields (metric of the
// restricted deformation and density).  With a lognormal density the coefficient contrast is
// large and that operator is a poor model of the fine one on coarse modes: the PCG iteration count
// grows with the grid (37 at 256^3, 46 at 512^3, 56 at 1024^3 for sigma = 2).  The Galerkin
// operator A_{l+1} = R A_l P, with P the trilinear prolongation of the V-cycle (k_prolong) and
// R = P^T / 8 its full-weighting restriction (k_restrict_resid), is the coarse operator the
// fine one induces; with it the iteration count is about half and no longer grows.  It is a
// 27-point stencil: 27 coefficient fields per level, field e = (ex+1) + 3(ey+1) + 9(ez+1) holding
// the coefficient of offset (ex, ey, ez), each field in the level's sub-lattice layout.
//
// Building is local: a coarse row sums, over the 27 fine points f of the full-weighting box and the
// stencil offsets d of the fine operator at f, R(C, f) a(f, d) P(f + d, C + e).  Per axis the
// fine offset o = delta + d from 2C feeds coarse offset e with weight 1 when o = 2e and 1/2 when
// |o - 2e| = 1.  The loops run over compile-time offsets, so the 27 sums stay in registers.
#pragma once
#include "kernels.hpp"

namespace mmh {

// Coefficients of the fine 19-point operator L phi = (t1 + t2 + tdiag phi) * di2 (stencil19,
// mmh/gauss1.f90:954-978) at point (i, j, k), read off the expression: faces from the face
// averages of G11/G22/G33, edges from the cross terms of G12/G23/G13.
// SM: n a power of two with n^3 < 2^31, indices in 32-bit shifts and masks (the same values).
template <class T, bool SM = false>
__device__ __forceinline__ void coef19(const T* __restrict__ G, int n, int i, int j, int k, double di2, double a[19]) {
  const i64 n3 = (i64)n * n * n;
  const int m = n - 1;
  auto g = [&](int comp, int di, int dj, int dk) -> double {
    return (double)G[(i64)comp * n3 + lx<1, SM>((i + di) & m, (j + dj) & m, (k + dk) & m, n)];
  };
  const double g1 = g(0, 0, 0, 0), g4 = g(3, 0, 0, 0), g6 = g(5, 0, 0, 0);
  const double b11p = (g(0, 1, 0, 0) + g1) / 2, b11m = (g(0, -1, 0, 0) + g1) / 2;
  const double b22p = (g(3, 0, 1, 0) + g4) / 2, b22m = (g(3, 0, -1, 0) + g4) / 2;
  const double b33p = (g(5, 0, 0, 1) + g6) / 2, b33m = (g(5, 0, 0, -1) + g6) / 2;
  a[0] = -(b11p + b11m + b22p + b22m + b33p + b33m);
  a[1] = b11p; a[2] = b11m; a[3] = b22p; a[4] = b22m; a[5] = b33p; a[6] = b33m;
  const double g2jp = g(1, 0, 1, 0), g2jm = g(1, 0, -1, 0), g2ip = g(1, 1, 0, 0), g2im = g(1, -1, 0, 0);
  a[7] = (g2jp + g2ip) / 4;    // (+1, +1, 0)
  a[8] = -(g2jp + g2im) / 4;   // (-1, +1, 0)
  a[9] = -(g2jm + g2ip) / 4;   // (+1, -1, 0)
  a[10] = (g2jm + g2im) / 4;   // (-1, -1, 0)
  const double g5kp = g(4, 0, 0, 1), g5km = g(4, 0, 0, -1), g5jp = g(4, 0, 1, 0), g5jm = g(4, 0, -1, 0);
  a[11] = (g5kp + g5jp) / 4;   // (0, +1, +1)
  a[12] = -(g5kp + g5jm) / 4;  // (0, -1, +1)
  a[13] = -(g5km + g5jp) / 4;  // (0, +1, -1)
  a[14] = (g5km + g5jm) / 4;   // (0, -1, -1)
  const double g3kp = g(2, 0, 0, 1), g3km = g(2, 0, 0, -1), g3ip = g(2, 1, 0, 0), g3im = g(2, -1, 0, 0);
  a[15] = (g3kp + g3ip) / 4;   // (+1, 0, +1)
  a[16] = -(g3kp + g3im) / 4;  // (-1, 0, +1)
  a[17] = -(g3km + g3ip) / 4;  // (+1, 0, -1)
  a[18] = (g3km + g3im) / 4;   // (-1, 0, -1)
#pragma unroll
  for (int q = 0; q < 19; ++q) a[q] *= di2;
}

__device__ __forceinline__ constexpr int st27_index(int ex, int ey, int ez) { return (ex + 1) + 3 * (ey + 1) + 9 * (ez + 1); }

// Adds R(C, f) a(f, d) P(f + d, C + e) for one fine point f = 2C + delta and one stencil offset d,
// whose fine offset from 2C is o = delta + d in [-2, 2]^3 (compile-time).  Per axis: o even ->
// e = o/2 (weight 1); o odd -> e = (o-1)/2 and (o+1)/2 (weight 1/2 each).
template <int OX, int OY, int OZ> __device__ __forceinline__ void rap_add(double (&acc)[27], double w) {
  constexpr int ex0 = (OX - (OX & 1)) / 2, ey0 = (OY - (OY & 1)) / 2, ez0 = (OZ - (OZ & 1)) / 2;
  constexpr int nx = (OX & 1) ? 2 : 1, ny = (OY & 1) ? 2 : 1, nz = (OZ & 1) ? 2 : 1;
  constexpr double ww = ((OX & 1) ? 0.5 : 1.0) * ((OY & 1) ? 0.5 : 1.0) * ((OZ & 1) ? 0.5 : 1.0);
  const double v = w * ww;
  acc[st27_index(ex0, ey0, ez0)] += v;
  if constexpr (nx == 2) acc[st27_index(ex0 + 1, ey0, ez0)] += v;
  if constexpr (ny == 2) acc[st27_index(ex0, ey0 + 1, ez0)] += v;
  if constexpr (nz == 2) acc[st27_index(ex0, ey0, ez0 + 1)] += v;
  if constexpr (nx == 2 && ny == 2) acc[st27_index(ex0 + 1, ey0 + 1, ez0)] += v;
  if constexpr (nx == 2 && nz == 2) acc[st27_index(ex0 + 1, ey0, ez0 + 1)] += v;
  if constexpr (ny == 2 && nz == 2) acc[st27_index(ex0, ey0 + 1, ez0 + 1)] += v;
  if constexpr (nx == 2 && ny == 2 && nz == 2) acc[st27_index(ex0 + 1, ey0 + 1, ez0 + 1)] += v;
}

// the offsets of the 19-point stencil as compile-time triples, in the order of coef19
template <int S> struct St19;
#define ST19_(S, X, Y, Z) \
  template <> struct St19<S> { static constexpr int x = X, y = Y, z = Z; };
ST19_(0, 0, 0, 0) ST19_(1, 1, 0, 0) ST19_(2, -1, 0, 0) ST19_(3, 0, 1, 0) ST19_(4, 0, -1, 0) ST19_(5, 0, 0, 1)
ST19_(6, 0, 0, -1) ST19_(7, 1, 1, 0) ST19_(8, -1, 1, 0) ST19_(9, 1, -1, 0) ST19_(10, -1, -1, 0) ST19_(11, 0, 1, 1)
ST19_(12, 0, -1, 1) ST19_(13, 0, 1, -1) ST19_(14, 0, -1, -1) ST19_(15, 1, 0, 1) ST19_(16, -1, 0, 1)
ST19_(17, 1, 0, -1) ST19_(18, -1, 0, -1)
#undef ST19_

// one fine point fd (delta = (fd%3-1, fd/3%3-1, fd/9-1)) of a level-1 row: its 19 coefficients
template <int FD, int S>
__device__ __forceinline__ void rap19_terms(double (&acc)[27], const double (&a)[19], double wr) {
  if constexpr (S < 19) {
    rap_add<FD % 3 - 1 + St19<S>::x, (FD / 3) % 3 - 1 + St19<S>::y, FD / 9 - 1 + St19<S>::z>(acc, wr * a[S]);
    rap19_terms<FD, S + 1>(acc, a, wr);
  }
}
template <class T, int FD>
__device__ __forceinline__ void rap19_points(double (&acc)[27], const T* __restrict__ G, int n, int ci, int cj, int ck,
                                             double di2) {
  if constexpr (FD < 27) {
    constexpr int dx = FD % 3 - 1, dy = (FD / 3) % 3 - 1, dz = FD / 9 - 1;
    constexpr double wr = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0) / 8.0;
    const int m = n - 1;
    double a[19];
    coef19<T, true>(G, n, (2 * ci + dx) & m, (2 * cj + dy) & m, (2 * ck + dz) & m, di2, a);
    rap19_terms<FD, 0>(acc, a, wr);
    rap19_points<T, FD + 1>(acc, G, n, ci, cj, ck, di2);
  }
}

// Level 1 from the fine 19-point operator: one thread per coarse point q (column-major coarse
// order), A1 (27 fields of (n/2)^3, sub-lattice layout of the coarse level).  n a power of two,
// n <= 1024 (32-bit indices).
template <class T>
__global__ __launch_bounds__(256) void k_rap19(double* __restrict__ A1, const T* __restrict__ G, int n, double di2,
                                               int batched = 0) {
  const int h = n >> 1;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 h3 = (i64)h * h * h;
  if (batched) {  // grid y of gridDim.y: A1 at y 27 (n/2)^3, G at y 6 n^3
    A1 += (i64)blockIdx.y * 27 * h3;
    G += (i64)blockIdx.y * 6 * ((i64)n * n * n);
  }
  if (q >= h3) return;
  int ci, cj, ck;
  point_of<0, true>(q, h, ci, cj, ck);
  double acc[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) acc[e] = 0.0;
  rap19_points<T, 0>(acc, G, n, ci, cj, ck, di2);
  const i64 p = lx<1>(ci, cj, ck, h);
#pragma unroll
  for (int e = 0; e < 27; ++e) A1[(i64)e * h3 + p] = acc[e];
}

template <int FD, int S>
__device__ __forceinline__ void rap27_terms(double (&acc)[27], const double* __restrict__ A, i64 n3, i64 f, double wr) {
  if constexpr (S < 27) {
    rap_add<FD % 3 - 1 + S % 3 - 1, (FD / 3) % 3 - 1 + (S / 3) % 3 - 1, FD / 9 - 1 + S / 9 - 1>(acc, wr * A[(i64)S * n3 + f]);
    rap27_terms<FD, S + 1>(acc, A, n3, f, wr);
  }
}
template <int FD>
__device__ __forceinline__ void rap27_points(double (&acc)[27], const double* __restrict__ A, int n, int ci, int cj, int ck) {
  if constexpr (FD < 27) {
    constexpr int dx = FD % 3 - 1, dy = (FD / 3) % 3 - 1, dz = FD / 9 - 1;
    constexpr double wr = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0) / 8.0;
    const int m = n - 1;
    const i64 f = lx<1>((2 * ci + dx) & m, (2 * cj + dy) & m, (2 * ck + dz) & m, n);
    rap27_terms<FD, 0>(acc, A, (i64)n * n * n, f, wr);
    rap27_points<FD + 1>(acc, A, n, ci, cj, ck);
  }
}

// Level l+1 from the 27-point level l (size n): one thread per coarse point.
__global__ __launch_bounds__(256) void k_rap27(double* __restrict__ Ac, const double* __restrict__ A, int n,
                                               int batched = 0) {
  const int h = n >> 1;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 h3 = (i64)h * h * h;
  if (batched) {  // grid y of gridDim.y: Ac at y 27 (n/2)^3, A at y 27 n^3
    Ac += (i64)blockIdx.y * 27 * h3;
    A += (i64)blockIdx.y * 27 * ((i64)n * n * n);
  }
  if (q >= h3) return;
  int ci, cj, ck;
  point_of<0>(q, h, ci, cj, ck);
  double acc[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) acc[e] = 0.0;
  rap27_points<0>(acc, A, n, ci, cj, ck);
  const i64 p = lx<1>(ci, cj, ck, h);
#pragma unroll
  for (int e = 0; e < 27; ++e) Ac[(i64)e * h3 + p] = acc[e];
}

// sum over the 26 off-centre offsets of A(p, e) phi(p + e) at point (i, j, k) of a 27-point level.
// With MASK, neighbours in colour classes b (4 (k&1) + 2 (i&1) + (j&1)) whose bit of live is clear
// hold zeros: their terms are left out, coefficient and value unread (the sum keeps its value; a
// zero may differ in sign).
template <class T, bool MASK = false, class TG = T, bool SM = false>
__device__ __forceinline__ T gal_offsum(const T* __restrict__ phi, const TG* __restrict__ A, int n, int i, int j, int k,
                                       i64 p, unsigned live = 0xFFu) {
  const int m = n - 1;
  const i64 n3 = (i64)n * n * n;
  const unsigned cc = 4u * (k & 1) + 2u * (i & 1) + (j & 1);
  T s = T(0);
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    if (e == 13) continue;
    const int ex = e % 3 - 1, ey = (e / 3) % 3 - 1, ez = e / 9 - 1;
    if (MASK && !((live >> (cc ^ (4u * (ez != 0) + 2u * (ex != 0) + (ey != 0)))) & 1u)) continue;
    s = s + gl<T>(A, (i64)e * n3 + p) * phi[lx<1, SM>((i + ex) & m, (j + ey) & m, (k + ez) & m, n)];
  }
  return s;
}

// Gauss-Seidel update of point q of colour class (kpar, io, jo) on a 27-point level (the 8
// classes decouple the 27-point stencil as they do the 19-point one).
template <class T, bool MASK = false, class TG = T, bool SM = false>
__device__ __forceinline__ void gal_colour_point(T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ A,
                                                 int n, int kpar, int io, int jo, i64 q, unsigned live = 0xFFu) {
  const int h = n >> 1;
  int i, j, k;
  if constexpr (SM) {  // n^3 < 2^31, n a power of two: shifts and masks
    const int qq = (int)q, lg = __builtin_ctz(h);
    i = 2 * (qq & (h - 1)) + io; j = 2 * ((qq >> lg) & (h - 1)) + jo; k = 2 * (qq >> (2 * lg)) + kpar;
  } else {
    i = 2 * (int)(q % h) + io; j = 2 * (int)((q / h) % h) + jo; k = 2 * (int)(q / ((i64)h * h)) + kpar;
  }
  const i64 p = lx<1, SM>(i, j, k, n);
  const i64 n3 = (i64)n * n * n;
  phi[p] = (rhs[p] - gal_offsum<T, MASK, TG, SM>(phi, A, n, i, j, k, p, live)) / gl<T>(A, 13 * n3 + p);
}

template <class T, class TG = T, bool SM = false>
__device__ __forceinline__ T gal_resid_point(const T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ A,
                                            int n, int i, int j, int k) {
  const i64 p = lx<1, SM>(i, j, k, n);
  const i64 n3 = (i64)n * n * n;
  return rhs[p] - (gal_offsum<T, false, TG, SM>(phi, A, n, i, j, k, p) + gl<T>(A, 13 * n3 + p) * phi[p]);
}

// The launched kernels work on one sub-lattice class at a time (a colour launch's points, or one
// class of a residual launch), so the classes of all 26 neighbours, and with them every array base
// below, are uniform over the workgroup; a lane adds its 32-bit byte offset within a class array
// (the load's scalar-base-plus-offset form, no 64-bit address arithmetic per load).  The products
// are summed in the order of gal_offsum (offsets e = 0..26, the centre left out), so the results
// are those of gal_colour_point and gal_resid_point, bit for bit.
// Sum over the 26 off-centre offsets at the point (qi, qj, qk) of class c (in-class offset o0 =
// qi + h qj + h^2 qk); Ac = A + c h^3 (the point's class in each coefficient field), h a power of two.
// TG = bf16x2: the 27 fields in 14 words (fields 2f, 2f + 1 in word f), read as words.  The centre
// coefficient comes back in a13.
template <class T, bool MASK, class TG>
__device__ __forceinline__ T gal_offsum_cls(const T* __restrict__ phi, const TG* __restrict__ Ac, int h, int c, int qi,
                                           int qj, int qk, unsigned o0, unsigned live, T& a13) {
  const int hm = h - 1;
  const i64 h3 = (i64)h * h * h, n3 = 8 * h3;
  const int io = c & 1, jo = (c >> 1) & 1, ko = c >> 2;
  // in-class coordinates of the neighbours at offsets -1, 0, +1 along each axis (x + 2 qx + o
  // lies in class x + o mod 2), scaled by the class array's strides
  const unsigned X0 = io ? qi : (qi - 1) & hm, X2 = io ? (qi + 1) & hm : qi;
  const unsigned Y0 = (unsigned)h * (jo ? qj : (qj - 1) & hm), Y1 = (unsigned)h * qj,
                 Y2 = (unsigned)h * (jo ? (qj + 1) & hm : qj);
  const unsigned hh = (unsigned)h * h;
  const unsigned Z0 = hh * (ko ? qk : (qk - 1) & hm), Z1 = hh * qk, Z2 = hh * (ko ? (qk + 1) & hm : qk);
  const unsigned X[3] = {X0, (unsigned)qi, X2}, Y[3] = {Y0, Y1, Y2}, Z[3] = {Z0, Z1, Z2};
  const unsigned cc = 4u * ko + 2u * io + jo;  // colour index of the point's class
  // All loads first, then the sum.  With MASK a left-out term still loads, from the point's own
  // centre coefficient and value (read by the update anyway, so no extra memory traffic), and is
  // skipped in the sum: loads behind a branch would each be waited for before the next.
  T av[27], vv[27];
  bool use[27];
  use[13] = false;
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    if (e == 13) continue;
    const int ex = e % 3 - 1, ey = (e / 3) % 3 - 1, ez = e / 9 - 1;
    use[e] = !MASK || ((live >> (cc ^ (4u * (ez != 0) + 2u * (ex != 0) + (ey != 0)))) & 1u);
    const int cn = use[e] ? c ^ ((ex != 0) | ((ey != 0) << 1) | ((ez != 0) << 2)) : c;
    const unsigned on = use[e] ? X[ex + 1] + Y[ey + 1] + Z[ez + 1] : o0;
    vv[e] = *(const T*)((const char*)(phi + (i64)cn * h3) + (uint32_t)(on * (unsigned)sizeof(T)));
  }
  if constexpr (std::is_same<TG, bf16x2>::value) {
    // word f unless neither of its fields is used (then the centre word 6, read anyway)
    unsigned w[14];
#pragma unroll
    for (int f = 0; f < 14; ++f) {
      const bool need = !MASK || use[2 * f] || (2 * f + 1 < 27 && use[2 * f + 1]);
      const TG* af = Ac + (need ? (i64)f * n3 : 6 * n3);
      w[f] = *(const unsigned*)((const char*)af + (uint32_t)(o0 * 4u));
    }
#pragma unroll
    for (int e = 0; e < 27; ++e)
      if (e != 13) av[e] = (e & 1) ? bf16_hi(w[e >> 1]) : bf16_lo(w[e >> 1]);
    a13 = bf16_hi(w[6]);
  } else {
#pragma unroll
    for (int e = 0; e < 27; ++e) {
      if (e == 13) continue;
      const TG* ae = Ac + (use[e] ? (i64)e * n3 : 13 * n3);
      av[e] = gl<T>((const TG*)((const char*)ae + (uint32_t)(o0 * (unsigned)sizeof(TG))), 0);
    }
    a13 = gl<T>((const TG*)((const char*)(Ac + 13 * n3) + (uint32_t)(o0 * (unsigned)sizeof(TG))), 0);
  }
  T s = T(0);
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    if (e == 13) continue;
    if (MASK && !use[e]) continue;
    s = s + av[e] * vv[e];
  }
  return s;
}

// Batched launches: gridDim.y = B grids stored one after another (bstride = n^3 values; A holds
// 27 n^3 per grid).  MASK: the first sweep from phi = 0 (live: classes updated earlier in it).
template <class T, bool MASK = false, class TG = T>
__global__ __launch_bounds__(256) void k_gal_colour(T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ A,
                                                    int n, int kpar, int io, int jo, i64 bstride = 0,
                                                    unsigned live = 0xFFu, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const int h = n >> 1, lg = __builtin_ctz(h);
  const i64 h3 = (i64)h * h * h;
  constexpr int NF = std::is_same<TG, bf16x2>::value ? 14 : 27;  // coefficient arrays per grid
  phi += (i64)blockIdx.y * bstride;
  rhs += (i64)blockIdx.y * bstride;
  A += (i64)blockIdx.y * NF * bstride;
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const int q = (int)(blk * blockDim.x + threadIdx.x);  // in-class offset, h^3 <= 2^27
  if (q >= h3) return;
  const int c = io | (jo << 1) | (kpar << 2);
  const int qi = q & (h - 1), qj = (q >> lg) & (h - 1), qk = q >> (2 * lg);
  const TG* Ac = A + (i64)c * h3;
  const uint32_t ob = (uint32_t)q * (uint32_t)sizeof(T);
  T* pc = (T*)((char*)(phi + (i64)c * h3) + ob);
  const T rv = *(const T*)((const char*)(rhs + (i64)c * h3) + ob);
  T a13;
  const T s = gal_offsum_cls<T, MASK, TG>(phi, Ac, h, c, qi, qj, qk, (unsigned)q, live, a13);
  *pc = (rv - s) / a13;
}

// Workgroups of k_gal_resid for an n^3 level: gal_resid_blocks(n) / 8 per sub-lattice class.
inline unsigned gal_resid_blocks(int n) {
  const i64 h = n >> 1;
  return 8u * (unsigned)((h * h * h + 255) >> 8);
}

// r = rhs - A phi at every point (one thread per point, sub-lattice order; the workgroups of one
// class come together, gal_resid_blocks); r stored.
template <class T, class TG = T>
__global__ __launch_bounds__(256) void k_gal_resid(T* __restrict__ r, const T* __restrict__ phi, const T* __restrict__ rhs,
                                                   const TG* __restrict__ A, int n, i64 bstride = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  r += (i64)blockIdx.y * bstride;
  phi += (i64)blockIdx.y * bstride;
  rhs += (i64)blockIdx.y * bstride;
  A += (i64)blockIdx.y * (std::is_same<TG, bf16x2>::value ? 14 : 27) * bstride;
  const int h = n >> 1, lg = __builtin_ctz(h);
  const i64 h3 = (i64)h * h * h;
  const unsigned gx = (unsigned)((h3 + 255) >> 8);
  const int c = (int)(blockIdx.x / gx);
  const int q = (int)((blockIdx.x - (unsigned)c * gx) * blockDim.x + threadIdx.x);
  if (q >= h3) return;
  const int qi = q & (h - 1), qj = (q >> lg) & (h - 1), qk = q >> (2 * lg);
  const TG* Ac = A + (i64)c * h3;
  const uint32_t ob = (uint32_t)q * (uint32_t)sizeof(T);
  T a13;
  const T s = gal_offsum_cls<T, false, TG>(phi, Ac, h, c, qi, qj, qk, (unsigned)q, 0xFFu, a13);
  const T pv = *(const T*)((const char*)(phi + (i64)c * h3) + ob);
  const T rv = *(const T*)((const char*)(rhs + (i64)c * h3) + ob);
  *(T*)((char*)(r + (i64)c * h3) + ob) = rv - (s + a13 * pv);
}

}  // namespace mmh
