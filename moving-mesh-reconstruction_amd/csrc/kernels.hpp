// kernels.hpp -- gfx950 device kernels of the moving-mesh reconstruction solve.
//
// Arrays are column-major (x fastest), offset = i + n*(j + n*k), periodic in all axes.  Every
// arithmetic statement keeps the association order of the reference Fortran statement it
// implements (cited per kernel); the library is built with -ffp-contract=off, so FP64 and FP32
// results are bit-identical to a sequential evaluation of those statements.  Global sums are
// FP64 with a fixed blocked tree order (k_sum_stage), so they are deterministic run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmh {

typedef int64_t i64;

// 16-bit storage of preconditioner coefficients (bfloat16: the top half of an FP32 word).  The
// arithmetic stays in T; a coefficient array of type TG is read through gl<T>, which widens bf16
// exactly and is the plain load when TG is T.
struct bf16 {
  unsigned short b;
};
// Two bfloat16 coefficients in one word: field 2f in the low half, field 2f + 1 in the high half.
struct bf16x2 {
  unsigned w;
};
__device__ __forceinline__ float bf16_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(unsigned w) { return __uint_as_float(w & 0xFFFF0000u); }
template <class T, class TG> __device__ __forceinline__ T gl(const TG* __restrict__ p, i64 o) { return p[o]; }
template <> __device__ __forceinline__ float gl<float, bf16>(const bf16* __restrict__ p, i64 o) {
  return __uint_as_float((unsigned)p[o].b << 16);
}
// Round to the nearest bf16 (ties to even) of the FP32 rounding of a double.
__device__ __forceinline__ bf16 to_bf16(double x) {
  unsigned u = __float_as_uint((float)x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return bf16{(unsigned short)(u >> 16)};
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ i64 ix3(int i, int j, int k, int n) { return (i64)i + (i64)n * ((i64)j + (i64)n * (i64)k); }
template <class T> __device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }

template <class T> __device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }

// Workgroups are dealt to the 8 XCDs round-robin (id mod 8), each XCD with its own L2.  Remapping
// the id so that XCD x receives the contiguous range [x*nb/8, (x+1)*nb/8) keeps spatially adjacent
// tiles, which share halo and neighbour-plane data, in one L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
  constexpr unsigned X = 8;
  return (nb % X) ? b : (b % X) * (nb / X) + b / X;
}

constexpr int RB = 256;  // reduction block
constexpr int RPT = 8;   // elements per thread per reduction block

// The halving tree s[t] = op(s[t], s[t + h]), h = NT/2 .. 1, over the values a of the NT threads
// of a workgroup (op 0: sum, op 1: maximum); the result is returned in thread 0.  Levels h >= 64
// pair values of different wavefronts through s[NT] in shared memory, one barrier each; levels
// h < 64 pair lanes t and t + h of wavefront 0 with __shfl_down, the same operands in the same
// order as in shared memory, so the result is bit-identical to the tree run wholly in s.
template <int NT, int OP> __device__ __forceinline__ double tree_reduce(double a, double* s) {
  static_assert(NT >= 128 && (NT & (NT - 1)) == 0, "power of two of at least two wavefronts");
  const int t = threadIdx.x;
  s[t] = a;
  __syncthreads();
#pragma unroll
  for (int h = NT / 2; h > 64; h >>= 1) {
    if (t < h) s[t] = OP == 0 ? s[t] + s[t + h] : dmax(s[t], s[t + h]);
    __syncthreads();
  }
  if (t < 64) {
    a = OP == 0 ? s[t] + s[t + 64] : dmax(s[t], s[t + 64]);
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) {
      const double o = __shfl_down(a, h);
      a = OP == 0 ? a + o : dmax(a, o);
    }
  }
  return a;
}

// Maximum of a over the RB threads of a workgroup, stored by thread 0 to *out: a butterfly over
// the 64 lanes of each wavefront, then one barrier and the RB/64 wavefront values.  A maximum does
// not depend on the order it is taken in, so this equals the halving tree over shared memory.
__device__ __forceinline__ void block_max_store(double a, double* out) {
  __shared__ double w[RB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a = dmax(a, __shfl_xor(a, off));
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = w[0];
#pragma unroll
    for (int v = 1; v < RB / 64; ++v) m = dmax(m, w[v]);
    *out = m;
  }
}

// ------------------------------------------------------------------------------------------------
// Deterministic FP64 reductions.  Block b covers elements [b*RB*RPT, (b+1)*RB*RPT); thread t adds
// elements t, t+RB, ... of the block in order; the RB per-thread values are combined by a halving
// tree.  Repeated on the block results until one value is left.
// op 0: sum of x; op 1: max of |x|.
// Batched launches (gridDim.y = B): item y reads x + y*xstride and writes out + y*gridDim.x.
template <class T, int OP>
__global__ __launch_bounds__(RB) void k_reduce_stage(double* __restrict__ out, const T* __restrict__ x, i64 n,
                                                    i64 xstride = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  __shared__ double s[RB];
  const int t = threadIdx.x;
  x += (i64)blockIdx.y * xstride;
  out += (i64)blockIdx.y * gridDim.x;
  const i64 base = (i64)blockIdx.x * (RB * RPT);
  double a = 0.0;
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    i64 q = base + (i64)m * RB + t;
    if (q < n) {
      double v = (double)x[q];
      if (OP == 0) a = a + v;
      else a = dmax(a, fabs(v));
    }
  }
  a = tree_reduce<RB, OP>(a, s);
  if (t == 0) out[blockIdx.x] = a;
}

template <class T> __global__ void k_fill(T* __restrict__ a, T v, i64 n, i64 bstride = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  a += (i64)blockIdx.y * bstride;
  i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

template <class T> __global__ void k_copy_strided(T* __restrict__ dst, const T* __restrict__ src, i64 n, int stride) {
  i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q * stride];
}

// ------------------------------------------------------------------------------------------------
// Array layouts of multigrid level data.  L = 0: column-major (the reference's order).  L = 1:
// sub-lattice order -- the 8 parity classes c = i%2 + 2*(j%2) + 4*(k%2) stored one after another,
// each (n/2)^3 in column-major order of (i/2, j/2, k/2).  A colour sweep touches one class and
// reads six others, all at unit stride in this order.
// SM: the level has fewer than 2^31 points per array (the one-workgroup cycle's levels), so the
// offset is formed in 32-bit arithmetic (the same offset, fewer instructions).
template <int L, bool SM = false> __device__ __forceinline__ i64 lx(int i, int j, int k, int n) {
  if constexpr (SM) {
    if (L == 0) return (i64)(i + n * (j + n * k));
    const int h = n >> 1;
    const int c = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
    return (i64)(c * (h * h * h) + (i >> 1) + h * ((j >> 1) + h * (k >> 1)));
  }
  if (L == 0) return ix3(i, j, k, n);
  const int h = n >> 1;
  const int c = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
  return (i64)c * ((i64)h * h * h) + (i >> 1) + (i64)h * ((j >> 1) + (i64)h * (k >> 1));
}

// Point of linear thread index p (0 <= p < n^3) such that consecutive p are consecutive in memory.
template <int L, bool SM = false> __device__ __forceinline__ void point_of(i64 p, int n, int& i, int& j, int& k) {
  if constexpr (SM) {  // n a power of two, n^3 < 2^31: shifts and masks
    const int q = (int)p;
    if (L == 0) {
      const int lg = __builtin_ctz(n);
      i = q & (n - 1); j = (q >> lg) & (n - 1); k = q >> (2 * lg);
      return;
    }
    const int h = n >> 1, lg = __builtin_ctz(h), h3 = h * h * h;
    const int c = q >> (3 * lg), r = q & (h3 - 1);
    i = 2 * (r & (h - 1)) + (c & 1);
    j = 2 * ((r >> lg) & (h - 1)) + ((c >> 1) & 1);
    k = 2 * (r >> (2 * lg)) + ((c >> 2) & 1);
    return;
  }
  if (L == 0) {
    i = (int)(p % n); j = (int)((p / n) % n); k = (int)(p / ((i64)n * n));
    return;
  }
  const int h = n >> 1;
  const i64 h3 = (i64)h * h * h;
  const int c = (int)(p / h3);
  const i64 q = p - (i64)c * h3;
  i = 2 * (int)(q % h) + (c & 1);
  j = 2 * (int)((q / h) % h) + ((c >> 1) & 1);
  k = 2 * (int)(q / ((i64)h * h)) + ((c >> 2) & 1);
}

template <class T, int TO>
__global__ void k_relayout(T* __restrict__ dst, const T* __restrict__ src, int n, int comps, int batched = 0) {
  // TO = 1: column-major -> sub-lattice; TO = 0: sub-lattice -> column-major; comps fields of n^3
  // (batched: grid y of gridDim.y at y comps n^3)
  const i64 n3 = (i64)n * n * n;
  if (batched) {
    dst += (i64)blockIdx.y * comps * n3;
    src += (i64)blockIdx.y * comps * n3;
  }
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3 * comps) return;
  const int f = (int)(p / n3);
  int i, j, k;
  point_of<1>(p - f * n3, n, i, j, k);
  const i64 a = f * n3 + lx<1>(i, j, k, n), b = f * n3 + ix3(i, j, k, n);
  if (TO == 1) dst[a] = src[b];
  else dst[b] = src[a];
}

// ------------------------------------------------------------------------------------------------
// Metric G = rho_box * adj(A)/det (mmh/gauss1.f90:1128-1205, cbajraw, triad branch) or
// adj(A) adj(A)^T / det (lsquare, 1173-1181); A = I + Hess(def)/dx^2; det <= 1e-5 -> identity.
// def and u are column-major; G is 6 fields (G11, G12, G13, G22, G23, G33) in layout L.
// Batched (gridDim.y = B, batched = 1): grid y at def + y n^3, u + y nu n^3, G + y 6 n^3.
// Metric at point (i, j, k) from its def and u neighbourhoods: dp(di, dj, dk) reads def, up(di, dj,
// dk) reads u (component 0 of nu), st(f, v) stores field f; the expressions of cbajraw.
template <class T, class DP, class UP, class ST>
__device__ __forceinline__ void metric_point(DP dp, UP up, ST st, T dx, int lsquare) {
  const T di2 = T(1) / (dx * dx);
  const T d0 = dp(0, 0, 0);
  T phixx = (dp(1, 0, 0) - T(2) * d0 + dp(-1, 0, 0)) * di2;
  T phiyy = (dp(0, 1, 0) - T(2) * d0 + dp(0, -1, 0)) * di2;
  T phizz = (dp(0, 0, 1) - T(2) * d0 + dp(0, 0, -1)) * di2;
  T phixy = (dp(1, 1, 0) - dp(-1, 1, 0) - dp(1, -1, 0) + dp(-1, -1, 0)) * di2 / T(4);
  T phiyz = (dp(0, 1, 1) - dp(0, 1, -1) - dp(0, -1, 1) + dp(0, -1, -1)) * di2 / T(4);
  T phixz = (dp(1, 0, 1) - dp(-1, 0, 1) - dp(1, 0, -1) + dp(-1, 0, -1)) * di2 / T(4);
  T a11 = T(1) + phixx, a12 = phixy, a13 = phixz, a22 = T(1) + phiyy, a23 = phiyz, a33 = T(1) + phizz;
  T det = a11 * a22 * a33 + T(2) * a12 * a13 * a23 - (a13 * a13) * a22 - a11 * (a23 * a23) - (a12 * a12) * a33;
  T b11 = a22 * a33 - a23 * a23;
  T b12 = a13 * a23 - a12 * a33;
  T b13 = a12 * a23 - a13 * a22;
  T b22 = a11 * a33 - a13 * a13;
  T b23 = a12 * a13 - a11 * a23;
  T b33 = a11 * a22 - a12 * a12;
  if (det <= T(1.e-5)) { b11 = 1; b12 = 0; b13 = 0; b22 = 1; b23 = 0; b33 = 1; det = 1; }
  if (lsquare) {
    st(0, (b11 * b11 + b12 * b12 + b13 * b13) / det);
    st(1, (b11 * b12 + b12 * b22 + b13 * b23) / det);
    st(2, (b11 * b13 + b12 * b23 + b13 * b33) / det);
    st(3, (b12 * b12 + b22 * b22 + b23 * b23) / det);
    st(4, (b12 * b13 + b22 * b23 + b23 * b33) / det);
    st(5, (b13 * b13 + b23 * b23 + b33 * b33) / det);
    return;
  }
  T urg = up(-1, -1, -1);
  bool first = true;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (first) { first = false; continue; }
        urg = urg + up(a - 1, b - 1, c - 1);
      }
  urg = urg / T(27);
  st(0, urg * b11 / det);
  st(1, urg * b12 / det);
  st(2, urg * b13 / det);
  st(3, urg * b22 / det);
  st(4, urg * b23 / det);
  st(5, urg * b33 / det);
}

template <class T, int L>
__global__ __launch_bounds__(256) void k_metric(T* __restrict__ G, const T* __restrict__ def, const T* __restrict__ u,
                                                int nu, int n, T dx, int lsquare, int batched = 0) {
  const i64 n3 = (i64)n * n * n;
  if (batched) {
    const i64 y = blockIdx.y;
    G += y * 6 * n3;
    def += y * n3;
    u += y * nu * n3;
  }
  const i64 pp = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (pp >= n3) return;
  // threads walk the column-major inputs in order; the store goes to layout L
  if ((n & (n - 1)) == 0 && (n * n) % 256 == 0 && n <= 1024) {
    // A workgroup's 256 points lie in one plane k (n a power of two, n^2 a multiple of 256): the
    // plane bases of def, u and G are scalars and a lane adds its 32-bit byte offset in the plane.
    const int lg = __builtin_ctz(n), m = n - 1;
    const int k = (int)(((i64)blockIdx.x * blockDim.x) >> (2 * lg));
    const int r = (int)(pp - ((i64)k << (2 * lg)));
    const int i = r & m, j = r >> lg;
    const int im = (i - 1) & m, ip = (i + 1) & m, jm = (j - 1) & m, jp = (j + 1) & m;
    const int is[3] = {im, i, ip}, js[3] = {jm, j, jp};
    auto plane = [&](int dk) { return (i64)((k + dk) & m) << (2 * lg); };
    auto dp = [&](int di, int dj, int dk) -> T {
      const uint32_t o = (uint32_t)(is[di + 1] + (js[dj + 1] << lg)) * (uint32_t)sizeof(T);
      return *(const T*)((const char*)(def + plane(dk)) + o);
    };
    auto up = [&](int di, int dj, int dk) -> T {
      const uint32_t o = (uint32_t)nu * (uint32_t)(is[di + 1] + (js[dj + 1] << lg)) * (uint32_t)sizeof(T);
      return *(const T*)((const char*)(u + (i64)nu * plane(dk)) + o);
    };
    // store offset: layout 1 splits the plane into classes by (i & 1, j & 1) (class bit 2 = k & 1 and
    // the class plane k >> 1 are uniform); layout 0 is the plane itself
    const int h = n >> 1;
    const i64 h3 = (i64)h * h * h;
    const T* gbase = G + (L == 1 ? (i64)(4 * (k & 1)) * h3 + (i64)h * h * (k >> 1) : ((i64)k << (2 * lg)));
    const uint32_t so = L == 1 ? (uint32_t)(((i & 1) + 2 * (j & 1)) * (uint32_t)h3 + (i >> 1) + (uint32_t)h * (j >> 1)) *
                                     (uint32_t)sizeof(T)
                               : (uint32_t)r * (uint32_t)sizeof(T);
    auto st = [&](int f, T v) { *(T*)((char*)(gbase + f * n3) + so) = v; };
    metric_point<T>(dp, up, st, dx, lsquare);
    return;
  }
  int i, j, k;
  point_of<0>(pp, n, i, j, k);
  const i64 p = lx<L>(i, j, k, n);
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
  const int is[3] = {im, i, ip}, js[3] = {jm, j, jp}, ks[3] = {km, k, kp};
  auto dp = [&](int di, int dj, int dk) -> T { return def[ix3(is[di + 1], js[dj + 1], ks[dk + 1], n)]; };
  auto up = [&](int di, int dj, int dk) -> T { return u[(i64)nu * ix3(is[di + 1], js[dj + 1], ks[dk + 1], n)]; };
  auto st = [&](int f, T v) { G[f * n3 + p] = v; };
  metric_point<T>(dp, up, st, dx, lsquare);
}

// The 19-point symmetric curvilinear operator at one point (mmh/gauss1.f90:954-978), on arrays in
// layout L.  DELTA = true (the FP32 preconditioner of the mixed solve only) returns in t1 the face
// sum over (phi_nb - phi_c) instead of over phi_nb, so that t1 + tdiag phi_c is formed from
// differences: in FP32 the update then loses precision to |grad phi| rather than to |phi|.
// MASK = true (with DELTA = false): bit b of live says whether colour class b (4 (k&1) + 2 (i&1) +
// (j&1)) of phi may be non-zero; values of the other classes are taken as 0 without being loaded,
// and so are the G2/G3/G5 factors of cross terms whose phi difference is then 0.  The result has
// the value the unmasked stencil gives when those classes hold zeros (a zero may differ in sign).
template <class T, int L, bool DELTA = false, bool MASK = false, class TG = T, bool SM = false>
__device__ __forceinline__ void stencil19(const T* __restrict__ phi, const TG* __restrict__ G, i64 n3, int n, int i, int j,
                                          int k, T& t1, T& t2, T& tdiag, unsigned live = 0xFFu) {
  const TG* G1 = G;
  const TG* G2 = G + n3;
  const TG* G3 = G + 2 * n3;
  const TG* G4 = G + 3 * n3;
  const TG* G5 = G + 4 * n3;
  const TG* G6 = G + 5 * n3;
  const int m = n - 1;
  const int ip = (i + 1) & m, im = (i - 1) & m, jp = (j + 1) & m, jm = (j - 1) & m, kp = (k + 1) & m, km = (k - 1) & m;
  const i64 c = lx<L, SM>(i, j, k, n);
#define X_(a, b, cc) lx<L, SM>(a, b, cc, n)
  T b11p = (gl<T>(G1, X_(ip, j, k)) + gl<T>(G1, c)) / T(2);
  T b11m = (gl<T>(G1, X_(im, j, k)) + gl<T>(G1, c)) / T(2);
  T b22p = (gl<T>(G4, X_(i, jp, k)) + gl<T>(G4, c)) / T(2);
  T b22m = (gl<T>(G4, X_(i, jm, k)) + gl<T>(G4, c)) / T(2);
  T b33p = (gl<T>(G6, X_(i, j, kp)) + gl<T>(G6, c)) / T(2);
  T b33m = (gl<T>(G6, X_(i, j, km)) + gl<T>(G6, c)) / T(2);
  tdiag = -(b11p + b11m + b22p + b22m + b33p + b33m);
  if constexpr (MASK) {
    static_assert(!DELTA, "the masked stencil is the plain form");
    const unsigned cc = 4u * (k & 1) + 2u * (i & 1) + (j & 1);
    const bool fx = (live >> (cc ^ 2u)) & 1u, fy = (live >> (cc ^ 1u)) & 1u, fz = (live >> (cc ^ 4u)) & 1u;
    const bool fxy = (live >> (cc ^ 3u)) & 1u, fyz = (live >> (cc ^ 5u)) & 1u, fxz = (live >> (cc ^ 6u)) & 1u;
    auto P = [&](bool on, i64 a) { return on ? phi[a] : T(0); };
    auto F = [&](bool on, const TG* g, i64 a) { return on ? gl<T>(g, a) : T(0); };
    t1 = (b11p * P(fx, X_(ip, j, k)) + b11m * P(fx, X_(im, j, k)) + b22p * P(fy, X_(i, jp, k)) +
          b22m * P(fy, X_(i, jm, k)) + b33p * P(fz, X_(i, j, kp)) + b33m * P(fz, X_(i, j, km)));
    t2 = ((P(fxy, X_(ip, jp, k)) - P(fxy, X_(im, jp, k))) * F(fxy, G2, X_(i, jp, k))
        - (P(fxy, X_(ip, jm, k)) - P(fxy, X_(im, jm, k))) * F(fxy, G2, X_(i, jm, k))
        + (P(fxy, X_(ip, jp, k)) - P(fxy, X_(ip, jm, k))) * F(fxy, G2, X_(ip, j, k))
        - (P(fxy, X_(im, jp, k)) - P(fxy, X_(im, jm, k))) * F(fxy, G2, X_(im, j, k))
        + (P(fyz, X_(i, jp, kp)) - P(fyz, X_(i, jm, kp))) * F(fyz, G5, X_(i, j, kp))
        - (P(fyz, X_(i, jp, km)) - P(fyz, X_(i, jm, km))) * F(fyz, G5, X_(i, j, km))
        + (P(fyz, X_(i, jp, kp)) - P(fyz, X_(i, jp, km))) * F(fyz, G5, X_(i, jp, k))
        - (P(fyz, X_(i, jm, kp)) - P(fyz, X_(i, jm, km))) * F(fyz, G5, X_(i, jm, k))
        + (P(fxz, X_(ip, j, kp)) - P(fxz, X_(im, j, kp))) * F(fxz, G3, X_(i, j, kp))
        - (P(fxz, X_(ip, j, km)) - P(fxz, X_(im, j, km))) * F(fxz, G3, X_(i, j, km))
        + (P(fxz, X_(ip, j, kp)) - P(fxz, X_(ip, j, km))) * F(fxz, G3, X_(ip, j, k))
        - (P(fxz, X_(im, j, kp)) - P(fxz, X_(im, j, km))) * F(fxz, G3, X_(im, j, k))) / T(4);
    return;
  }
  if (DELTA) {
    const T pc = phi[c];
    t1 = (b11p * (phi[X_(ip, j, k)] - pc) + b11m * (phi[X_(im, j, k)] - pc) + b22p * (phi[X_(i, jp, k)] - pc) +
          b22m * (phi[X_(i, jm, k)] - pc) + b33p * (phi[X_(i, j, kp)] - pc) + b33m * (phi[X_(i, j, km)] - pc));
  } else {
    t1 = (b11p * phi[X_(ip, j, k)] + b11m * phi[X_(im, j, k)] + b22p * phi[X_(i, jp, k)] +
          b22m * phi[X_(i, jm, k)] + b33p * phi[X_(i, j, kp)] + b33m * phi[X_(i, j, km)]);
  }
  t2 = ((phi[X_(ip, jp, k)] - phi[X_(im, jp, k)]) * gl<T>(G2, X_(i, jp, k))
      - (phi[X_(ip, jm, k)] - phi[X_(im, jm, k)]) * gl<T>(G2, X_(i, jm, k))
      + (phi[X_(ip, jp, k)] - phi[X_(ip, jm, k)]) * gl<T>(G2, X_(ip, j, k))
      - (phi[X_(im, jp, k)] - phi[X_(im, jm, k)]) * gl<T>(G2, X_(im, j, k))
      + (phi[X_(i, jp, kp)] - phi[X_(i, jm, kp)]) * gl<T>(G5, X_(i, j, kp))
      - (phi[X_(i, jp, km)] - phi[X_(i, jm, km)]) * gl<T>(G5, X_(i, j, km))
      + (phi[X_(i, jp, kp)] - phi[X_(i, jp, km)]) * gl<T>(G5, X_(i, jp, k))
      - (phi[X_(i, jm, kp)] - phi[X_(i, jm, km)]) * gl<T>(G5, X_(i, jm, k))
      + (phi[X_(ip, j, kp)] - phi[X_(im, j, kp)]) * gl<T>(G3, X_(i, j, kp))
      - (phi[X_(ip, j, km)] - phi[X_(im, j, km)]) * gl<T>(G3, X_(i, j, km))
      + (phi[X_(ip, j, kp)] - phi[X_(ip, j, km)]) * gl<T>(G3, X_(ip, j, k))
      - (phi[X_(im, j, kp)] - phi[X_(im, j, km)]) * gl<T>(G3, X_(im, j, k))) / T(4);
#undef X_
}

// Gauss-Seidel update of point q (0 <= q < (n/2)^3) of the colour class (kpar, io, jo), in place
// (relaxplane, mmh/gauss1.f90:954-980).
template <class T, int L, bool DELTA = false, bool MASK = false, class TG = T, bool SM = false>
__device__ __forceinline__ void colour_point(T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ G,
                                             int n, T dx2, int kpar, int io, int jo, i64 q, unsigned live = 0xFFu) {
  const int h = n >> 1;
  int i, j, k;
  if constexpr (SM) {
    const int qq = (int)q, lg = __builtin_ctz(h);
    i = 2 * (qq & (h - 1)) + io; j = 2 * ((qq >> lg) & (h - 1)) + jo; k = 2 * (qq >> (2 * lg)) + kpar;
  } else {
    i = 2 * (int)(q % h) + io; j = 2 * (int)((q / h) % h) + jo; k = 2 * (int)(q / ((i64)h * h)) + kpar;
  }
  const i64 n3 = (i64)n * n * n;
  T t1, t2, tdiag;
  stencil19<T, L, DELTA, MASK, TG, SM>(phi, G, n3, n, i, j, k, t1, t2, tdiag, live);
  const i64 c = lx<L, SM>(i, j, k, n);
  if (DELTA) phi[c] = phi[c] + (rhs[c] * dx2 - t1 - t2) / tdiag;
  else phi[c] = (rhs[c] * dx2 - t1 - t2) / tdiag;
}

// One colour class of the 8-colour Gauss-Seidel sweep: points with (k, i, j) parity (kpar, io, jo)
// (relaxplane, mmh/gauss1.f90:910-980).  One thread per point of the colour; in layout 1 the
// threads of a workgroup own consecutive points of the class and every stencil access is
// unit-stride.
// Batched launches: gridDim.y = B grids of the same size stored one after another (bstride = n^3
// values; G holds 6 n^3 per grid).  remap = 0 keeps the hardware's block order (a tuning choice).
// MASK: the first sweep from phi = 0, where live holds the classes updated earlier in the sweep.
template <class T, int L, bool DELTA = false, bool MASK = false, class TG = T>
__global__ __launch_bounds__(256) void k_colour(T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ G, int n,
                                                T dx2, int kpar, int io, int jo, i64 bstride = 0, int remap = 1,
                                                unsigned live = 0xFFu, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const int h = n >> 1;
  phi += (i64)blockIdx.y * bstride;
  rhs += (i64)blockIdx.y * bstride;
  G += (i64)blockIdx.y * 6 * bstride;
  const unsigned blk = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const i64 q = (i64)blk * blockDim.x + threadIdx.x;
  if (q >= (i64)h * h * h) return;
  colour_point<T, L, DELTA, MASK, TG>(phi, rhs, G, n, dx2, kpar, io, jo, q, live);
}

// The colour class of k_colour (sub-lattice layout) with each thread updating S consecutive points
// of its (i, j) column, k = 2 kk + kpar for kk = S*seg .. S*seg + S-1.  Points of one class never
// read each other, so the order is free; what the column order buys is reuse: the k+1 values one
// point loads (phi of the z, yz and xz neighbour classes, G6, G5 and G3 of the z class) are the
// k-1 values of the next point, carried in registers instead of fetched again.  Each point's
// arithmetic is stencil19's, term for term and in the same order (the build contracts no
// multiply-add), so the results are bit-identical to k_colour's, MASK and DELTA forms included.
// Occupancy the register allocator aims at: FP32 values (4-byte loads, low register need) up to 5
// waves per SIMD, so that the scheduler keeps some 20-30 loads of a wave in flight rather than
// waiting on each one to reach 8 waves; FP64 is faster at the compiler's own choice.
template <class T> struct ColWaves { static constexpr int value = sizeof(T) == 4 ? 5 : 8; };
template <class T, bool DELTA, bool MASK, int S, class TG = T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, ColWaves<T>::value))) void k_colour_col(T* __restrict__ phi, const T* __restrict__ rhs,
                                                    const TG* __restrict__ G, int n, T dx2, int kpar, int io, int jo,
                                                    unsigned live = 0xFFu, int seg0 = 0, int nseg = 0) {
  static_assert(!(DELTA && MASK), "the masked stencil is the plain form");
  const int h = n >> 1, m = n - 1, hh = h * h;
  const i64 h3 = (i64)hh * h, n3 = 8 * h3;
  // Block b (XCD-remapped) holds 256 columns of one segment of S class planes, colour_col_blocks(n)
  // blocks per segment: the planes are uniform over the block, so every array base below, plane
  // offset included, is a scalar, and a lane adds only its 32-bit offset within the plane.
  const unsigned gx = (unsigned)(hh + 255) >> 8;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int seg = (int)(b / gx);
  const int r = (int)(b - (unsigned)seg * gx) * 256 + (int)threadIdx.x;
  if (r >= hh || seg >= (nseg ? nseg : h / S)) return;
  const int qi = r % h, qj = r / h, kk0 = (seg + seg0) * S;
  const int i = 2 * qi + io, j = 2 * qj + jo;
  const int xp = ((i + 1) & m) >> 1, xm = ((i - 1) & m) >> 1, yp = ((j + 1) & m) >> 1, ym = ((j - 1) & m) >> 1;
  // offsets within a class plane: the column and its in-plane neighbours
  const unsigned o0 = qi + h * qj, oxp = xp + h * qj, oxm = xm + h * qj, oyp = qi + h * yp, oym = qi + h * ym;
  const unsigned opp = xp + h * yp, omp = xm + h * yp, opm = xp + h * ym, omm = xm + h * ym;
  // sub-lattice classes (i&1 | (j&1)<<1 | (k&1)<<2) of the point and of its neighbours
  const int c = io | (jo << 1) | (kpar << 2);
  const i64 ox = (c ^ 1) * h3, oy = (c ^ 2) * h3, oz = (c ^ 4) * h3, oc = (i64)c * h3;
  const TG* G1 = G;
  const TG* G2 = G + n3;
  const TG* G3 = G + 2 * n3;
  const TG* G4 = G + 3 * n3;
  const TG* G5 = G + 4 * n3;
  const TG* G6 = G + 5 * n3;
  const T* px = phi + ox;
  const T* py = phi + oy;
  const T* pz = phi + oz;
  const T* pxy = phi + (c ^ 3) * h3;
  const T* pyz = phi + (c ^ 6) * h3;
  const T* pxz = phi + (c ^ 5) * h3;
  T* pc_ = phi + oc;
  const T* rc_ = rhs + oc;
  // which neighbour classes may be non-zero (live is indexed by colour 4 kpar + 2 io + jo)
  const unsigned cc = 4u * kpar + 2u * io + jo;
  const bool fx = !MASK || ((live >> (cc ^ 2u)) & 1u), fy = !MASK || ((live >> (cc ^ 1u)) & 1u);
  const bool fz = !MASK || ((live >> (cc ^ 4u)) & 1u), fxy = !MASK || ((live >> (cc ^ 3u)) & 1u);
  const bool fyz = !MASK || ((live >> (cc ^ 5u)) & 1u), fxz = !MASK || ((live >> (cc ^ 6u)) & 1u);
  // value at offset o of class plane kq of an array: the plane base is a scalar and the lane adds
  // its offset in bytes as a 32-bit value (o * 8 < 2^32 for n <= 2048), the load's own
  // scalar-base-plus-offset form
  auto V = [&](const T* a, int kq, unsigned o) -> T {
    return *(const T*)((const char*)(a + (i64)hh * kq) + (uint32_t)(o * (unsigned)sizeof(T)));
  };
  auto W = [&](const TG* g, int kq, unsigned o) -> T {
    return gl<T>((const TG*)((const char*)(g + (i64)hh * kq) + (uint32_t)(o * (unsigned)sizeof(TG))), 0);
  };
  auto P = [&](bool on, const T* a, int kq, unsigned o) { return on ? V(a, kq, o) : T(0); };
  auto F = [&](bool on, const TG* g, int kq, unsigned o) { return on ? W(g, kq, o) : T(0); };
  // the k-1 plane values of the first point
  const int km0 = (((2 * kk0 + kpar) - 1) & m) >> 1;
  T g6m = W(G6 + oz, km0, o0), pzm = P(fz, pz, km0, o0);
  T pyzpm = P(fyz, pyz, km0, oyp), pyzmm = P(fyz, pyz, km0, oym);
  T g5m = F(fyz, G5 + oz, km0, o0);
  T pxzpm = P(fxz, pxz, km0, oxp), pxzmm = P(fxz, pxz, km0, oxm);
  T g3m = F(fxz, G3 + oz, km0, o0);
  T res[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int kk = kk0 + s;
    const int kpl = (((2 * kk + kpar) + 1) & m) >> 1;
    const T g6p = W(G6 + oz, kpl, o0), pzp = P(fz, pz, kpl, o0);
    const T pyzpp = P(fyz, pyz, kpl, oyp), pyzmp = P(fyz, pyz, kpl, oym);
    const T g5p = F(fyz, G5 + oz, kpl, o0);
    const T pxzpp = P(fxz, pxz, kpl, oxp), pxzmp = P(fxz, pxz, kpl, oxm);
    const T g3p = F(fxz, G3 + oz, kpl, o0);
    const T b11p = (W(G1 + ox, kk, oxp) + W(G1 + oc, kk, o0)) / T(2);
    const T b11m = (W(G1 + ox, kk, oxm) + W(G1 + oc, kk, o0)) / T(2);
    const T b22p = (W(G4 + oy, kk, oyp) + W(G4 + oc, kk, o0)) / T(2);
    const T b22m = (W(G4 + oy, kk, oym) + W(G4 + oc, kk, o0)) / T(2);
    const T b33p = (g6p + W(G6 + oc, kk, o0)) / T(2);
    const T b33m = (g6m + W(G6 + oc, kk, o0)) / T(2);
    const T tdiag = -(b11p + b11m + b22p + b22m + b33p + b33m);
    const T pxp = P(fx, px, kk, oxp), pxm = P(fx, px, kk, oxm);
    const T pyp = P(fy, py, kk, oyp), pym = P(fy, py, kk, oym);
    T t1;
    if constexpr (DELTA) {
      const T pc = V(pc_, kk, o0);
      t1 = (b11p * (pxp - pc) + b11m * (pxm - pc) + b22p * (pyp - pc) + b22m * (pym - pc) + b33p * (pzp - pc) +
            b33m * (pzm - pc));
    } else {
      t1 = (b11p * pxp + b11m * pxm + b22p * pyp + b22m * pym + b33p * pzp + b33m * pzm);
    }
    const T pxypp = P(fxy, pxy, kk, opp), pxymp = P(fxy, pxy, kk, omp);
    const T pxypm = P(fxy, pxy, kk, opm), pxymm = P(fxy, pxy, kk, omm);
    const T t2 = ((pxypp - pxymp) * F(fxy, G2 + oy, kk, oyp)
                - (pxypm - pxymm) * F(fxy, G2 + oy, kk, oym)
                + (pxypp - pxypm) * F(fxy, G2 + ox, kk, oxp)
                - (pxymp - pxymm) * F(fxy, G2 + ox, kk, oxm)
                + (pyzpp - pyzmp) * g5p
                - (pyzpm - pyzmm) * g5m
                + (pyzpp - pyzpm) * F(fyz, G5 + oy, kk, oyp)
                - (pyzmp - pyzmm) * F(fyz, G5 + oy, kk, oym)
                + (pxzpp - pxzmp) * g3p
                - (pxzpm - pxzmm) * g3m
                + (pxzpp - pxzpm) * F(fxz, G3 + ox, kk, oxp)
                - (pxzmp - pxzmm) * F(fxz, G3 + ox, kk, oxm)) / T(4);
    if constexpr (DELTA) res[s] = V(pc_, kk, o0) + (V(rc_, kk, o0) * dx2 - t1 - t2) / tdiag;
    else res[s] = (V(rc_, kk, o0) * dx2 - t1 - t2) / tdiag;
    if constexpr (MASK) *(T*)((char*)(pc_ + (i64)hh * kk) + (uint32_t)(o0 * (unsigned)sizeof(T))) = res[s];
    g6m = g6p; pzm = pzp; pyzpm = pyzpp; pyzmm = pyzmp; g5m = g5p; pxzpm = pxzpp; pxzmm = pxzmp; g3m = g3p;
  }
  // Stores after all loads: the compiler cannot tell that they never alias the other classes' phi,
  // and would otherwise hold the next point's loads behind each store.  The masked form stores
  // point by point: its hoisted loads would need more registers than the kernel can have (with
  // FP32 values and 16-bit coefficients too: 156 VGPRs and three times the time per launch).
  if constexpr (!MASK) {
#pragma unroll
    for (int s = 0; s < S; ++s) *(T*)((char*)(pc_ + (i64)hh * (kk0 + s)) + (uint32_t)(o0 * (unsigned)sizeof(T))) = res[s];
  }
}
// Workgroups of k_colour_col per segment of S class planes of an n^3 level.
inline unsigned colour_col_blocks(int n) {
  const int h = n >> 1;
  return (unsigned)(h * h + 255) >> 8;
}

// Residual rhs - L[phi] at point (i, j, k) (calcresid, mmh/gauss1.f90:683-709).
template <class T, int L, bool DELTA = false, class TG = T, bool SM = false>
__device__ __forceinline__ T resid_point(const T* __restrict__ phi, const T* __restrict__ rhs, const TG* __restrict__ G,
                                        int n, T di2, int i, int j, int k) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = lx<L, SM>(i, j, k, n);
  T t1, t2, tdiag;
  stencil19<T, L, DELTA, false, TG, SM>(phi, G, n3, n, i, j, k, t1, t2, tdiag);
  return DELTA ? rhs[p] - (t1 + t2) * di2 : rhs[p] - (t1 + t2 + tdiag * phi[p]) * di2;
}

// Residual r = rhs - L[phi] (calcresid, mmh/gauss1.f90:683-709); optional store; per-block max|r|.
// In layout 1 a thread owns one 2x2x2 cell and evaluates its 8 points, one per sub-lattice, so
// the 8 class arrays and their neighbours are read in lockstep at unit stride.
template <class T, int L, bool DELTA = false, class TG = T>
__global__ __launch_bounds__(RB) void k_resid(T* __restrict__ r, double* __restrict__ bmax, const T* __restrict__ phi,
                                              const T* __restrict__ rhs, const TG* __restrict__ G, int n, T di2,
                                              int batched = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const i64 n3 = (i64)n * n * n;
  if (batched) {
    const i64 y = blockIdx.y;
    if (r) r += y * n3;
    phi += y * n3;
    rhs += y * n3;
    G += y * 6 * n3;
    bmax += y * gridDim.x;
  }
  // XCD-aware tile order; the block result is stored at the tile's index so the reduction order
  // does not depend on the remap
  const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
  const i64 pp = (i64)tile * RB + threadIdx.x;
  double a = 0.0;
  const int h = n >> 1;
  if (L == 1 && pp < (i64)h * h * h) {
    const int ci = (int)(pp % h), cj = (int)((pp / h) % h), ck = (int)(pp / ((i64)h * h));
#pragma unroll 1
    for (int c = 0; c < 8; ++c) {
      const int i = 2 * ci + (c & 1), j = 2 * cj + ((c >> 1) & 1), k = 2 * ck + (c >> 2);
      const i64 p = lx<L>(i, j, k, n);
      const T rv = resid_point<T, L, DELTA, TG>(phi, rhs, G, n, di2, i, j, k);
      if (r) r[p] = rv;
      a = dmax(a, (double)fabs(rv));
    }
  } else if (L == 0 && pp < n3) {
    int i, j, k;
    point_of<L>(pp, n, i, j, k);
    const i64 p = lx<L>(i, j, k, n);
    const T rv = resid_point<T, L, false, TG>(phi, rhs, G, n, di2, i, j, k);
    if (r) r[p] = rv;
    a = (double)fabs(rv);
  }
  block_max_store(a, bmax + tile);
}

// q = L[p] (the operator of k_resid: (t1 + t2 + tdiag*p)/dx^2) and per-block sums of p*q in FP64,
// block b covering elements [b*RB, (b+1)*RB) combined by the halving tree (deterministic).
template <class T, int L, bool DELTA = false>
__global__ __launch_bounds__(RB) void k_apply_dot(T* __restrict__ q, double* __restrict__ bsum, const T* __restrict__ p,
                                                  const T* __restrict__ G, int n, T di2, int batched = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  __shared__ double s[RB];
  const i64 n3 = (i64)n * n * n;
  if (batched) {
    const i64 y = blockIdx.y;
    q += y * n3;
    p += y * n3;
    G += y * 6 * n3;
    bsum += y * gridDim.x;
  }
  // XCD-aware tile order; the block result is stored at the tile's index so the reduction order
  // does not depend on the remap
  const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
  const i64 pp = (i64)tile * RB + threadIdx.x;
  double a = 0.0;
  const int h = n >> 1;
  if (L == 1 && pp < (i64)h * h * h) {
    const int ci = (int)(pp % h), cj = (int)((pp / h) % h), ck = (int)(pp / ((i64)h * h));
#pragma unroll 1
    for (int cc = 0; cc < 8; ++cc) {
      const int i = 2 * ci + (cc & 1), j = 2 * cj + ((cc >> 1) & 1), k = 2 * ck + (cc >> 2);
      const i64 c = lx<L>(i, j, k, n);
      T t1, t2, tdiag;
      stencil19<T, L, DELTA>(p, G, n3, n, i, j, k, t1, t2, tdiag);
      T v = DELTA ? (t1 + t2) * di2 : (t1 + t2 + tdiag * p[c]) * di2;
      q[c] = v;
      a = a + (double)p[c] * (double)v;
    }
  } else if (L == 0 && pp < n3) {
    int i, j, k;
    point_of<L>(pp, n, i, j, k);
    const i64 c = lx<L>(i, j, k, n);
    T t1, t2, tdiag;
    stencil19<T, L>(p, G, n3, n, i, j, k, t1, t2, tdiag);
    T v = (t1 + t2 + tdiag * p[c]) * di2;
    q[c] = v;
    a = (double)p[c] * (double)v;
  }
  a = tree_reduce<RB, 0>(a, s);
  if (threadIdx.x == 0) bsum[tile] = a;
}

// Per-block sums of a*b (FP64), same blocking as k_apply_dot.
template <class T>
__global__ __launch_bounds__(RB) void k_dot(double* __restrict__ bsum, const T* __restrict__ a, const T* __restrict__ b, i64 n,
                                            i64 bstride = 0) {
  __shared__ double s[RB];
  a += (i64)blockIdx.y * bstride;
  b += (i64)blockIdx.y * bstride;
  bsum += (i64)blockIdx.y * gridDim.x;
  const i64 p = (i64)blockIdx.x * RB + threadIdx.x;
  const double v = tree_reduce<RB, 0>(p < n ? (double)a[p] * (double)b[p] : 0.0, s);
  if (threadIdx.x == 0) bsum[blockIdx.x] = v;
}

// z with its mean removed as k_zerosum_guard (mode 0) would store it, when zsum (the FP64 sum of
// z) is given; z itself otherwise.
template <class TZ> __device__ __forceinline__ TZ z_centred(const TZ* __restrict__ z, i64 p, const double* zsum, i64 n) {
  if (!zsum) return z[p];
  const double m = zsum[0] / (double)n;
  return (TZ)((double)z[p] - m);
}

// Per-block sums of z*r and z*ro in FP64 (z, r and ro widened exactly from their storage types, the
// mean of z removed on the fly when zsum is given): the two dot products of one flexible-CG
// iteration in one pass over z.  Blocking of reduce(): block b covers [b*RB*RPT, (b+1)*RB*RPT),
// thread t adds elements t, t+RB, ... of it in order, then the halving tree.  Batched (gridDim.y =
// B): grid y reads at y*bstride and writes its block sums at y*gridDim.x (zsum applies to
// unbatched launches only).
template <class TZ, class TR = double>
__global__ __launch_bounds__(RB) void k_dot2(double* __restrict__ b1, double* __restrict__ b2, const TZ* __restrict__ z,
                                             const TR* __restrict__ r, const TR* __restrict__ ro, i64 n,
                                             const double* __restrict__ zsum = nullptr, i64 bstride = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  __shared__ double s1[RB], s2[RB];
  const int t = threadIdx.x;
  const i64 y = blockIdx.y;
  z += y * bstride; r += y * bstride; ro += y * bstride;
  b1 += y * gridDim.x; b2 += y * gridDim.x;
  const i64 base = (i64)blockIdx.x * (RB * RPT);
  double a1 = 0.0, a2 = 0.0;
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const i64 p = base + (i64)m * RB + t;
    if (p < n) {
      const double zz = (double)z_centred<TZ>(z, p, zsum, n);
      a1 = a1 + zz * (double)r[p];
      a2 = a2 + zz * (double)ro[p];
    }
  }
  a1 = tree_reduce<RB, 0>(a1, s1);
  a2 = tree_reduce<RB, 0>(a2, s2);
  if (t == 0) {
    b1[blockIdx.x] = a1;
    b2[blockIdx.x] = a2;
  }
}

// Per-block sums of z, z*r, z*ro, r and ro (FP64, the inputs widened exactly; blocking of k_dot2):
// one pass that gives the dot products of z with its mean removed, z.r - (sum z / n) sum r, and that
// mean itself, where z's sum would otherwise take a pass of its own.
template <class TZ, class TR>
__global__ __launch_bounds__(RB) void k_dot5(double* __restrict__ bz, double* __restrict__ b1, double* __restrict__ b2,
                                             double* __restrict__ br, double* __restrict__ bro, const TZ* __restrict__ z,
                                             const TR* __restrict__ r, const TR* __restrict__ ro, i64 n) {
  __shared__ double s[RB];
  const int t = threadIdx.x;
  const i64 base = (i64)blockIdx.x * (RB * RPT);
  double az = 0.0, a1 = 0.0, a2 = 0.0, ar = 0.0, aro = 0.0;
#pragma unroll
  for (int m = 0; m < RPT; ++m) {
    const i64 p = base + (i64)m * RB + t;
    if (p < n) {
      const double zz = (double)z[p], rr = (double)r[p], rro = (double)ro[p];
      az = az + zz;
      a1 = a1 + zz * rr;
      a2 = a2 + zz * rro;
      ar = ar + rr;
      aro = aro + rro;
    }
  }
  az = tree_reduce<RB, 0>(az, s);
  if (t == 0) bz[blockIdx.x] = az;
  __syncthreads();
  a1 = tree_reduce<RB, 0>(a1, s);
  if (t == 0) b1[blockIdx.x] = a1;
  __syncthreads();
  a2 = tree_reduce<RB, 0>(a2, s);
  if (t == 0) b2[blockIdx.x] = a2;
  __syncthreads();
  ar = tree_reduce<RB, 0>(ar, s);
  if (t == 0) br[blockIdx.x] = ar;
  __syncthreads();
  aro = tree_reduce<RB, 0>(aro, s);
  if (t == 0) bro[blockIdx.x] = aro;
}

// r -= alpha q (the previous r kept in rold, or rounded to FP32 in rold32); per-block max|r|.  With
// r32, the new r is also stored rounded to FP32 (the right-hand side of an FP32 preconditioner).
// x is advanced by the direction update (k_cg_dir), one pass later.
__global__ __launch_bounds__(RB) void k_cg_update(double* __restrict__ r, double* __restrict__ rold,
                                                  const double* __restrict__ q, double alpha, double* __restrict__ bmax,
                                                  i64 n, float* __restrict__ r32 = nullptr,
                                                  float* __restrict__ rold32 = nullptr) {
  const i64 i = (i64)blockIdx.x * RB + threadIdx.x;
  double a = 0.0;
  if (i < n) {
    const double ro = r[i];
    if (rold) rold[i] = ro;
    if (rold32) rold32[i] = (float)ro;
    const double rn = ro - alpha * q[i];
    r[i] = rn;
    if (r32) r32[i] = (float)rn;
    a = fabs(rn);
  }
  block_max_store(a, bmax + blockIdx.x);
}

// x += alpha p (the last step's update of the iterate).
__global__ void k_axpy(double* __restrict__ x, const double* __restrict__ p, double alpha, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + alpha * p[i];
}

// x += alpha p (with x), then p = z + beta p (z centred on the fly when zsum is given); with
// accumulate = 0, p = z.  One pass for the iterate's update and the new direction.
template <class TZ>
__global__ void k_cg_dir(double* __restrict__ x, double* __restrict__ p, const TZ* __restrict__ z, double alpha,
                         double beta, i64 n, const double* __restrict__ zsum = nullptr, int accumulate = 1) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double zz = (double)z_centred<TZ>(z, i, zsum, n);
  const double pi = p[i];
  if (x) x[i] = x[i] + alpha * pi;
  p[i] = accumulate ? zz + beta * pi : zz;
}

// CG updates in the working precision T with FP64 scalars (the FP32 Krylov solver).
template <class T>
__global__ __launch_bounds__(RB) void k_cg_update_t(T* __restrict__ x, T* __restrict__ r, T* __restrict__ rold,
                                                    const T* __restrict__ p, const T* __restrict__ q, double alpha,
                                                    double* __restrict__ bmax, i64 n) {
  const i64 i = (i64)blockIdx.x * RB + threadIdx.x;
  const T al = (T)alpha;
  double a = 0.0;
  if (i < n) {
    x[i] = x[i] + al * p[i];
    const T ro = r[i];
    rold[i] = ro;
    const T rn = ro - al * q[i];
    r[i] = rn;
    a = fabs((double)rn);
  }
  block_max_store(a, bmax + blockIdx.x);
}

template <class T> __global__ void k_cg_dir_t(T* __restrict__ p, const T* __restrict__ z, double beta, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = z[i] + (T)beta * p[i];
}

// Batched CG updates: grid y uses alpha[y] / beta[y] (0 once it has converged, which freezes x and r).
__global__ __launch_bounds__(RB) void k_cg_update_b(double* __restrict__ r, double* __restrict__ rold,
                                                    const double* __restrict__ q, const double* __restrict__ alpha,
                                                    double* __restrict__ bmax, i64 n, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const i64 y = blockIdx.y;
  r += y * n; rold += y * n; q += y * n;
  const double al = alpha[y];
  const i64 i = (i64)blockIdx.x * RB + threadIdx.x;
  double a = 0.0;
  if (i < n) {
    const double ro = r[i];
    rold[i] = ro;
    const double rn = ro - al * q[i];
    r[i] = rn;
    a = fabs(rn);
  }
  block_max_store(a, bmax + y * gridDim.x + blockIdx.x);
}

// x += alpha p, p = z + beta p per grid (alpha = beta = 0 freezes a converged grid).
__global__ void k_cg_dir_b(double* __restrict__ x, double* __restrict__ p, const double* __restrict__ z,
                           const double* __restrict__ alpha, const double* __restrict__ beta, i64 n, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const i64 y = blockIdx.y;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pi = p[y * n + i];
  x[y * n + i] = x[y * n + i] + alpha[y] * pi;
  p[y * n + i] = z[y * n + i] + beta[y] * pi;
}

// Batched set-up, one thread per grid: rn = max|rhs| of grid y (rnmax[y]) gives the residual
// scale rden = rn (1 where rn = 0) and the active flags rn > 0; the iteration count starts at 0.
__global__ void k_batch_flags(int B, const double* __restrict__ rnmax, double* __restrict__ rden, int* __restrict__ active,
                              int* __restrict__ iters, int* __restrict__ act0) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= B) return;
  const double rn = rnmax[y];
  rden[y] = rn > 0 ? rn : 1.0;
  active[y] = rn > 0 ? 1 : 0;
  act0[y] = active[y];
  iters[y] = 0;
}

// x = 0 for the grids that start inactive (L u = 0: the mean-free solution is 0); gridDim.y = B.
__global__ void k_zero_inactive(double* __restrict__ x, const int* __restrict__ active, i64 n) {
  const i64 y = blockIdx.y;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !active[y]) x[y * n + i] = 0.0;
}

// Per-grid CG scalars, one thread per grid.  step 0: alpha = rz / pq for active grids, else 0.
// step 1: relres = rmax / rden and the iteration count advance for active grids; a grid whose
// relres <= tol becomes inactive.  step 2: beta = (rz_new - zro) / rz, rz = rz_new (0 for inactive grids).
// A breakdown (alpha not finite: p.Lp = 0, or a preconditioner that produced no finite z) freezes
// the grid where it is: it stays active, so it is reported as not converged, but every kernel of
// the iteration skips it (act0 = 0) and its x and r keep their last finite values.
__global__ void k_cg_scalars(int step, int B, int it, double tol, double* __restrict__ alpha, double* __restrict__ beta,
                             double* __restrict__ rz, const double* __restrict__ pq, const double* __restrict__ rmax,
                             const double* __restrict__ rden, const double* __restrict__ rznew,
                             const double* __restrict__ zro, int* __restrict__ active, int* __restrict__ iters,
                             double* __restrict__ relres, int* __restrict__ act0) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= B) return;
  if (step == 0) {
    const double a = active[y] ? rz[y] / pq[y] : 0.0;
    const bool ok = isfinite(a);
    alpha[y] = ok ? a : 0.0;
    act0[y] = active[y] && ok;  // the grids this iteration works on (a grid converging in it still updates x)
  } else if (step == 1) {
    if (act0[y]) {
      const double rr = rmax[y] / rden[y];
      relres[y] = rr;
      iters[y] = iters[y] + 1;
      if (rr <= tol) active[y] = 0;
    }
  } else {
    if (active[y]) {
      beta[y] = (rznew[y] - zro[y]) / rz[y];
      rz[y] = rznew[y];
    } else {
      beta[y] = 0.0;
    }
  }
}

// Full-weighting restriction of the residual as accumulated by calcresid
// (mmh/gauss1.f90:819-859): coarse plane ck gets a main part from fine planes 2ck, 2ck+1 and a
// share from fine plane 2ck-1; plane ck0 accumulates main then share, every other plane share
// then main (the order in which r1lax's plane pipeline reaches them).  r and rc in layout L.
// restrict_point computes coarse point q (column-major order of the coarse grid).
template <class T, int L, bool SM = false>
__device__ __forceinline__ void restrict_point(T* __restrict__ rc, const T* __restrict__ r, int n, int ck0, i64 q) {
  const int h = n >> 1;
  // coarse points in column-major order: in layout 1 the 27 fine reads of consecutive threads are
  // then consecutive within each fine class array
  int ci, cj, ck;
  point_of<0, SM>(q, h, ci, cj, ck);
  const int kb = 2 * ck, kbp = 2 * ck + 1, ks = wrapi(2 * ck - 1, n);
  const int ib = 2 * ci, ibp = wrapi(ib + 1, n), ibm = wrapi(ib - 1, n);
  const int jb = 2 * cj, jbp = wrapi(jb + 1, n), jbm = wrapi(jb - 1, n);
#define R_(a, b, c) r[lx<L, SM>(a, b, c, n)]
  T acc = 0;
  if (ck == ck0) {
    acc = acc + R_(ib, jb, kb) / T(8)
      + (R_(ibp, jb, kb) + R_(ibm, jb, kb) + R_(ib, jbp, kb) + R_(ib, jbm, kb) + R_(ib, jb, kbp)) / T(16)
      + (R_(ibp, jbp, kb) + R_(ibp, jbm, kb) + R_(ibm, jbp, kb) + R_(ibm, jbm, kb) + R_(ibp, jb, kbp) + R_(ibm, jb, kbp) + R_(ib, jbp, kbp) + R_(ib, jbm, kbp)) / T(32)
      + (R_(ibp, jbp, kbp) + R_(ibp, jbm, kbp) + R_(ibm, jbp, kbp) + R_(ibm, jbm, kbp)) / T(64);
    acc = acc + R_(ib, jb, ks) / T(16)
      + (R_(ibp, jb, ks) + R_(ibm, jb, ks) + R_(ib, jbp, ks) + R_(ib, jbm, ks)) / T(32)
      + (R_(ibp, jbp, ks) + R_(ibp, jbm, ks) + R_(ibm, jbp, ks) + R_(ibm, jbm, ks)) / T(64);
  } else {
    acc = acc + R_(ib, jb, ks) / T(16)
      + (R_(ibp, jb, ks) + R_(ibm, jb, ks) + R_(ib, jbp, ks) + R_(ib, jbm, ks)) / T(32)
      + (R_(ibp, jbp, ks) + R_(ibp, jbm, ks) + R_(ibm, jbp, ks) + R_(ibm, jbm, ks)) / T(64);
    acc = acc + R_(ib, jb, kb) / T(8)
      + (R_(ibp, jb, kb) + R_(ibm, jb, kb) + R_(ib, jbp, kb) + R_(ib, jbm, kb) + R_(ib, jb, kbp)) / T(16)
      + (R_(ibp, jbp, kb) + R_(ibp, jbm, kb) + R_(ibm, jbp, kb) + R_(ibm, jbm, kb) + R_(ibp, jb, kbp) + R_(ibm, jb, kbp) + R_(ib, jbp, kbp) + R_(ib, jbm, kbp)) / T(32)
      + (R_(ibp, jbp, kbp) + R_(ibp, jbm, kbp) + R_(ibm, jbp, kbp) + R_(ibm, jbm, kbp)) / T(64);
  }
#undef R_
  rc[lx<L, SM>(ci, cj, ck, h)] = acc;
}

// restrict_point of the sub-lattice layout for the launched kernel (n a power of two): the 27 fine
// points lie in fixed classes (each index 2c, 2c + 1 or 2c - 1 has a known parity) at in-class
// coordinates c or c - 1 per axis, so every read is a class base (uniform) plus a lane's 32-bit byte
// offset.  The same terms in the same order as restrict_point.
template <class T>
__device__ __forceinline__ void restrict_point_cls(T* __restrict__ rc, const T* __restrict__ r, int n, int ck0, int q) {
  const int h = n >> 1, hm = h - 1, lg = __builtin_ctz(h);
  const i64 h3 = (i64)h * h * h;
  const int ci = q & hm, cj = (q >> lg) & hm, ck = q >> (2 * lg);
  const unsigned hh = (unsigned)h * h;
  const unsigned X0 = ci, X1 = (ci - 1) & hm, Y0 = (unsigned)h * cj, Y1 = (unsigned)h * ((cj - 1) & hm),
                 Z0 = hh * ck, Z1 = hh * ((ck - 1) & hm);
  // fine index 2c: class bit 0, coordinate c; 2c + 1: bit 1, c; 2c - 1: bit 1, c - 1
#define IB 0, 0
#define IBP 1, 0
#define IBM 1, 1
#define RR_(px, xs, py, ys, pz, zs)                                                          \
  (*(const T*)((const char*)(r + (i64)((px) | ((py) << 1) | ((pz) << 2)) * h3) +           \
               (uint32_t)(((xs) ? X1 : X0) + ((ys) ? Y1 : Y0) + ((zs) ? Z1 : Z0)) * (uint32_t)sizeof(T)))
#define R_(...) RR_(__VA_ARGS__)
  T acc = 0;
  if (ck == ck0) {
    acc = acc + R_(IB, IB, IB) / T(8)
      + (R_(IBP, IB, IB) + R_(IBM, IB, IB) + R_(IB, IBP, IB) + R_(IB, IBM, IB) + R_(IB, IB, IBP)) / T(16)
      + (R_(IBP, IBP, IB) + R_(IBP, IBM, IB) + R_(IBM, IBP, IB) + R_(IBM, IBM, IB) + R_(IBP, IB, IBP) + R_(IBM, IB, IBP) + R_(IB, IBP, IBP) + R_(IB, IBM, IBP)) / T(32)
      + (R_(IBP, IBP, IBP) + R_(IBP, IBM, IBP) + R_(IBM, IBP, IBP) + R_(IBM, IBM, IBP)) / T(64);
    acc = acc + R_(IB, IB, IBM) / T(16)
      + (R_(IBP, IB, IBM) + R_(IBM, IB, IBM) + R_(IB, IBP, IBM) + R_(IB, IBM, IBM)) / T(32)
      + (R_(IBP, IBP, IBM) + R_(IBP, IBM, IBM) + R_(IBM, IBP, IBM) + R_(IBM, IBM, IBM)) / T(64);
  } else {
    acc = acc + R_(IB, IB, IBM) / T(16)
      + (R_(IBP, IB, IBM) + R_(IBM, IB, IBM) + R_(IB, IBP, IBM) + R_(IB, IBM, IBM)) / T(32)
      + (R_(IBP, IBP, IBM) + R_(IBP, IBM, IBM) + R_(IBM, IBP, IBM) + R_(IBM, IBM, IBM)) / T(64);
    acc = acc + R_(IB, IB, IB) / T(8)
      + (R_(IBP, IB, IB) + R_(IBM, IB, IB) + R_(IB, IBP, IB) + R_(IB, IBM, IB) + R_(IB, IB, IBP)) / T(16)
      + (R_(IBP, IBP, IB) + R_(IBP, IBM, IB) + R_(IBM, IBP, IB) + R_(IBM, IBM, IB) + R_(IBP, IB, IBP) + R_(IBM, IB, IBP) + R_(IB, IBP, IBP) + R_(IB, IBM, IBP)) / T(32)
      + (R_(IBP, IBP, IBP) + R_(IBP, IBM, IBP) + R_(IBM, IBP, IBP) + R_(IBM, IBM, IBP)) / T(64);
  }
#undef R_
#undef RR_
#undef IB
#undef IBP
#undef IBM
  *(T*)((char*)rc + (uint32_t)lx<1, true>(ci, cj, ck, h) * (uint32_t)sizeof(T)) = acc;
}

template <class T, int L>
__global__ __launch_bounds__(256) void k_restrict_resid(T* __restrict__ rc, const T* __restrict__ r, int n, int ck0,
                                                        int batched = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const int h = n >> 1;
  if (batched) {
    rc += (i64)blockIdx.y * h * h * h;
    r += (i64)blockIdx.y * n * n * n;
  }
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)h * h * h) return;
  if constexpr (L == 1) restrict_point_cls<T>(rc, r, n, ck0, (int)q);
  else restrict_point<T, L>(rc, r, n, ck0, q);
}

// 27-point weights 1/8, 1/16, 1/32, 1/64 around fine point (i2,j2,k2): restrict/restrictu
// (mmh/mgutil.f90:106-112, 153-161) and nsmooth (mmh/limiter.f90:705-711) share term order.
template <class T>
__device__ __forceinline__ T box27(const T* __restrict__ u, int st, int n, int i2, int j2, int k2) {
  const int ip = wrapi(i2 + 1, n), im = wrapi(i2 - 1, n), jp = wrapi(j2 + 1, n), jm = wrapi(j2 - 1, n), kp = wrapi(k2 + 1, n), km = wrapi(k2 - 1, n);
#define U_(a, b, c) u[(i64)st * ix3(a, b, c, n)]
  return U_(i2, j2, k2) / T(8)
    + (U_(ip, j2, k2) + U_(im, j2, k2) + U_(i2, jp, k2) + U_(i2, jm, k2) + U_(i2, j2, kp) + U_(i2, j2, km)) / T(16)
    + (U_(ip, jp, k2) + U_(ip, jm, k2) + U_(im, jp, k2) + U_(im, jm, k2) + U_(ip, j2, kp) + U_(ip, j2, km)
       + U_(im, j2, kp) + U_(im, j2, km) + U_(i2, jp, kp) + U_(i2, jp, km) + U_(i2, jm, kp) + U_(i2, jm, km)) / T(32)
    + (U_(ip, jp, kp) + U_(ip, jp, km) + U_(ip, jm, kp) + U_(ip, jm, km) + U_(im, jp, kp) + U_(im, jp, km)
       + U_(im, jm, kp) + U_(im, jm, km)) / T(64);
#undef U_
}

template <class T>
__global__ __launch_bounds__(256) void k_restrict27(T* __restrict__ half, const T* __restrict__ full, int n, int stride,
                                                    int batched = 0) {
  const int h = n >> 1;
  if (batched) {  // grid y of gridDim.y: half at y (n/2)^3, full at y stride n^3
    half += (i64)blockIdx.y * h * h * h;
    full += (i64)blockIdx.y * stride * n * n * n;
  }
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)h * h * h) return;
  const int i = (int)(q % h), j = (int)((q / h) % h), k = (int)(q / ((i64)h * h));
  half[q] = box27(full, stride, n, 2 * i, 2 * j, 2 * k);
}

template <class T>
__global__ __launch_bounds__(256) void k_smooth27(T* __restrict__ out, const T* __restrict__ in, int n) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (i64)n * n * n) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  out[p] = box27(in, 1, n, i, j, k);
}

// Trilinear prolongation (mmh/mgutil.f90:38-71 inject adds; 2-34 prolong overwrites); u (n^3) and
// uh ((n/2)^3) in layout L.
// Trilinear prolongation of fine 2x2x2 cell pp (column-major cell order) in sub-lattice layout:
// its 8 points interpolate from the coarse block {ci, ci+1} x {cj, cj+1} x {ck, ck+1}; each
// point's 8-term sum keeps the order below.  add: u += s, else u = s.
template <class T, bool SM = false>
__device__ __forceinline__ void prolong_cell(T* __restrict__ u, const T* __restrict__ uh, int n, int add, i64 pp) {
  const int h = n >> 1;
  int ci, cj, ck;
  point_of<0, SM>(pp, h, ci, cj, ck);
  const int c1 = SM ? ((ci + 1) & (h - 1)) : (ci + 1) % h, d1 = SM ? ((cj + 1) & (h - 1)) : (cj + 1) % h,
            e1 = SM ? ((ck + 1) & (h - 1)) : (ck + 1) % h;
  T v[2][2][2];
  v[0][0][0] = uh[lx<1, SM>(ci, cj, ck, h)]; v[0][0][1] = uh[lx<1, SM>(ci, cj, e1, h)];
  v[0][1][0] = uh[lx<1, SM>(ci, d1, ck, h)]; v[0][1][1] = uh[lx<1, SM>(ci, d1, e1, h)];
  v[1][0][0] = uh[lx<1, SM>(c1, cj, ck, h)]; v[1][0][1] = uh[lx<1, SM>(c1, cj, e1, h)];
  v[1][1][0] = uh[lx<1, SM>(c1, d1, ck, h)]; v[1][1][1] = uh[lx<1, SM>(c1, d1, e1, h)];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int a = c & 1, b = (c >> 1) & 1, e = c >> 2;  // i = 2ci+a: low index ci, high index ci+a
    const T s = (v[0][0][0] + v[0][0][e] + v[0][b][0] + v[0][b][e]
                 + v[a][b][0] + v[a][b][e] + v[a][0][0] + v[a][0][e]) / T(8);
    const i64 p = lx<1, SM>(2 * ci + a, 2 * cj + b, 2 * ck + e, n);
    u[p] = add ? u[p] + s : s;
  }
}

// prolong_cell for the launched kernel (n a power of two): fine point (2ci + a, 2cj + b, 2ck + e)
// is element pp = ci + h (cj + h ck) of fine class a + 2b + 4e, a class base (uniform) plus the
// lane's 32-bit byte offset; the coarse reads take 32-bit offsets too.  Same values and order.
template <class T>
__device__ __forceinline__ void prolong_cell_cls(T* __restrict__ u, const T* __restrict__ uh, int n, int add, int pp) {
  const int h = n >> 1, hm = h - 1, lg = __builtin_ctz(h);
  const i64 h3 = (i64)h * h * h;
  const int ci = pp & hm, cj = (pp >> lg) & hm, ck = pp >> (2 * lg);
  const int c1 = (ci + 1) & hm, d1 = (cj + 1) & hm, e1 = (ck + 1) & hm;
  auto H = [&](int a, int b, int c) -> T {
    return *(const T*)((const char*)uh + (uint32_t)lx<1, true>(a, b, c, h) * (uint32_t)sizeof(T));
  };
  T v[2][2][2];
  v[0][0][0] = H(ci, cj, ck); v[0][0][1] = H(ci, cj, e1);
  v[0][1][0] = H(ci, d1, ck); v[0][1][1] = H(ci, d1, e1);
  v[1][0][0] = H(c1, cj, ck); v[1][0][1] = H(c1, cj, e1);
  v[1][1][0] = H(c1, d1, ck); v[1][1][1] = H(c1, d1, e1);
  const uint32_t ob = (uint32_t)pp * (uint32_t)sizeof(T);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int a = c & 1, b = (c >> 1) & 1, e = c >> 2;  // i = 2ci+a: low index ci, high index ci+a
    const T s = (v[0][0][0] + v[0][0][e] + v[0][b][0] + v[0][b][e]
                 + v[a][b][0] + v[a][b][e] + v[a][0][0] + v[a][0][e]) / T(8);
    T* up = (T*)((char*)(u + (i64)c * h3) + ob);
    *up = add ? *up + s : s;
  }
}

template <class T, int L>
__global__ __launch_bounds__(256) void k_prolong(T* __restrict__ u, const T* __restrict__ uh, int n, int add,
                                                 int batched = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const i64 pp = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const int h = n >> 1;
  if (batched) {
    u += (i64)blockIdx.y * n * n * n;
    uh += (i64)blockIdx.y * h * h * h;
  }
  if constexpr (L == 1) {
    if (pp >= (i64)h * h * h) return;
    prolong_cell_cls<T>(u, uh, n, add, (int)pp);
    return;
  }
  if (pp >= (i64)n * n * n) return;
  int i, j, k;
  point_of<L>(pp, n, i, j, k);
  const i64 p = lx<L>(i, j, k, n);
  const int ibl = i >> 1, ibh = ((i + 1) >> 1) % h, jbl = j >> 1, jbh = ((j + 1) >> 1) % h, kbl = k >> 1, kbh = ((k + 1) >> 1) % h;
#define H_(a, b, c) uh[lx<L>(a, b, c, h)]
  T s = (H_(ibl, jbl, kbl) + H_(ibl, jbl, kbh) + H_(ibl, jbh, kbl) + H_(ibl, jbh, kbh)
       + H_(ibh, jbh, kbl) + H_(ibh, jbh, kbh) + H_(ibh, jbl, kbl) + H_(ibh, jbl, kbh)) / T(8);
#undef H_
  u[p] = add ? u[p] + s : s;
}

// Unclamped sqrt(g) of the mesh (mmh/gauss1.f90:1283-1295, rgzerosum).
template <class T>
__global__ __launch_bounds__(256) void k_det(T* __restrict__ det, const T* __restrict__ def, int n, T dx, int batched = 0) {
  if (batched) {  // grid y of gridDim.y at y n^3
    det += (i64)blockIdx.y * n * n * n;
    def += (i64)blockIdx.y * n * n * n;
  }
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (i64)n * n * n) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
  const T di2 = T(1) / (dx * dx);
  const T d0 = def[p];
  T phixx = (def[ix3(ip, j, k, n)] - T(2) * d0 + def[ix3(im, j, k, n)]) * di2;
  T phiyy = (def[ix3(i, jp, k, n)] - T(2) * d0 + def[ix3(i, jm, k, n)]) * di2;
  T phizz = (def[ix3(i, j, kp, n)] - T(2) * d0 + def[ix3(i, j, km, n)]) * di2;
  T phixy = (def[ix3(ip, jp, k, n)] - def[ix3(im, jp, k, n)] - def[ix3(ip, jm, k, n)] + def[ix3(im, jm, k, n)]) * di2 / T(4);
  T phiyz = (def[ix3(i, jp, kp, n)] - def[ix3(i, jp, km, n)] - def[ix3(i, jm, kp, n)] + def[ix3(i, jm, km, n)]) * di2 / T(4);
  T phixz = (def[ix3(ip, j, kp, n)] - def[ix3(im, j, kp, n)] - def[ix3(ip, j, km, n)] + def[ix3(im, j, km, n)]) * di2 / T(4);
  T a11 = T(1) + phixx, a12 = phixy, a13 = phixz, a22 = T(1) + phiyy, a23 = phiyz, a33 = T(1) + phizz;
  det[p] = a11 * a22 * a33 + T(2) * a12 * a13 * a23 - (a13 * a13) * a22 - a11 * (a23 * a23) - (a12 * a12) * a33;
}

// rhs -= dfact*det with dfact = (T)(asum/dsum) read from device scalars (mmh/gauss1.f90:1348,1404).
template <class T>
__global__ void k_rgz_update(T* __restrict__ arr, const T* __restrict__ det, const double* __restrict__ sums, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const T dfact = (T)(sums[1] / sums[0]);
  arr[p] = arr[p] - dfact * det[p];
}

// k_rgz_update for gridDim.y = B grids of n3 values, grid y with its sums dsum[y] and asum[y].
template <class T>
__global__ void k_rgz_update_b(T* __restrict__ arr, const T* __restrict__ det, const double* __restrict__ dsum,
                               const double* __restrict__ asum, i64 n3) {
  const i64 y = blockIdx.y;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const T dfact = (T)(asum[y] / dsum[y]);
  arr[y * n3 + p] = arr[y * n3 + p] - dfact * det[y * n3 + p];
}

// dst (gridDim.y = B grids of n values side by side) from B separate arrays: dst[y n + p] =
// src[y][p stride] (stride > 1 picks one component of an interleaved array).
template <class T>
__global__ void k_gather(T* __restrict__ dst, const T* const* __restrict__ src, i64 n, int stride) {
  const i64 y = blockIdx.y;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[y * n + p] = src[y][p * stride];
}

// The inverse of k_gather with stride 1: dst[y][p] = src[y n + p].
template <class T> __global__ void k_scatter(T* const* __restrict__ dst, const T* __restrict__ src, i64 n) {
  const i64 y = blockIdx.y;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[y][p] = src[y * n + p];
}

// zerosum (mmh/gauss1.f90:1450-1460) or, when the level's divergence guard fired, szero.
// mode 0: always zerosum; mode 1: guard rinf[1] > rinf[0] (mg level); mode 2: the 4^3 guard
// relaxerr > 1.001*r1 and r1 > 1e-5 (mmh/multigrid.f90:113).  guard_out records the decision.
template <class T>
__global__ void k_zerosum_guard(T* __restrict__ a, const double* __restrict__ sum, const double* __restrict__ r1re, i64 n3,
                                int mode, int* __restrict__ guard_out, int batched = 0, const int* __restrict__ act = nullptr) {
  if (act && !act[blockIdx.y]) return;  // a grid of a batch that has converged
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (batched) {  // mode 0 only: item y removes the mean sum[y] / n3 of its grid
    a += (i64)blockIdx.y * n3;
    sum += blockIdx.y;
  }
  bool g = false;
  if (mode == 1) g = (T)r1re[1] > (T)r1re[0];
  else if (mode == 2) g = ((T)r1re[1] > (T)r1re[0] * T(1.001) && (T)r1re[0] > T(1.e-5));
  if (p == 0 && guard_out) *guard_out = g ? 1 : 0;
  if (p >= n3) return;
  if (g) { a[p] = T(0); return; }
  const double m = sum[0] / (double)n3;
  a[p] = (T)((double)a[p] - m);
}

// usum + conditional zerosum at the multigrid entry (mmh/multigrid.f90:41-46).
template <class T>
__global__ void k_entry_zerosum(T* __restrict__ a, const double* __restrict__ sum, i64 n3, int* __restrict__ flag) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const T asum = (T)sum[0];
  const bool z = asum > T(1.e-2);
  if (p == 0 && flag) *flag = z ? 1 : 0;
  if (p >= n3 || !z) return;
  const double m = sum[0] / (double)n3;
  a[p] = (T)((double)a[p] - m);
}


// Per-block diagnostics of the outer iteration (mmh/main.f90:97-106), FP64: sum defp^2, sum def^2,
// sum (u1-1)^2, min u1, max u1, count(u1 < 0).  Block b covers [b*RB*RPT, (b+1)*RB*RPT).
template <class T>
__global__ __launch_bounds__(RB) void k_iter_moments(double* __restrict__ out, const T* __restrict__ defp,
                                                     const T* __restrict__ def, const T* __restrict__ u, int nu, i64 n3) {
  __shared__ double s[6][RB];
  double a0 = 0, a1 = 0, a2 = 0, mn = 1e300, mx = -1e300, ng = 0;
  const i64 base = (i64)blockIdx.x * RB * RPT;
  for (int r = 0; r < RPT; ++r) {
    const i64 p = base + (i64)r * RB + threadIdx.x;
    if (p < n3) {
      const double dp = (double)defp[p], d = (double)def[p], w = (double)u[(i64)nu * p];
      a0 += dp * dp;
      a1 += d * d;
      a2 += (w - 1.0) * (w - 1.0);
      mn = w < mn ? w : mn;
      mx = w > mx ? w : mx;
      ng += (w < 0.0) ? 1.0 : 0.0;
    }
  }
  s[0][threadIdx.x] = a0; s[1][threadIdx.x] = a1; s[2][threadIdx.x] = a2;
  s[3][threadIdx.x] = mn; s[4][threadIdx.x] = mx; s[5][threadIdx.x] = ng;
  __syncthreads();
  for (int h = RB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const int t = threadIdx.x;
      s[0][t] += s[0][t + h]; s[1][t] += s[1][t + h]; s[2][t] += s[2][t + h];
      s[3][t] = s[3][t + h] < s[3][t] ? s[3][t + h] : s[3][t];
      s[4][t] = s[4][t + h] > s[4][t] ? s[4][t + h] : s[4][t];
      s[5][t] += s[5][t + h];
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[(i64)blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace mmh
