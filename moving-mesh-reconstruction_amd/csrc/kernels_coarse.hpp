// kernels_coarse.hpp -- the coarse end of the symmetric preconditioning V-cycle in one workgroup.
//
// Below 64^3 a V-cycle level is a few thousand points or fewer, and each of its ~40 launches
// (8 colours per sweep, residual, restriction, fill, prolongation, sums, mean removal) takes a few
// microseconds plus the dispatch gap: the 4^3..32^3 levels cost about as much as the 256^3 level.
// Here one workgroup of CW threads runs the whole cycle from the top coarse level down to 4^3
// and back, with a barrier between dependent phases.  The level arrays stay in global memory
// (they fit in the L2).  Every phase applies the per-point functions of the multi-launch kernels
// (colour_point, resid_point, restrict_point, prolong_cell) and the sums follow reduce()'s
// blocked tree order exactly, so the result is bit-identical to the launch-by-launch cycle.
// Two schedules: the Krylov preconditioner's (no guard, mean removal on every level) and the
// reference V-cycle's (mgtemplate.f90 / mg4, multigrid.f90:94-121), whose levels also record the
// residual norms r1 (after the pre-sweeps) and relaxerr (after the post-sweeps) and apply the
// divergence guard, szero instead of zerosum, as k_zerosum_guard does.
#pragma once
#include "kernels.hpp"
#include "kernels_gal.hpp"

namespace mmh {

constexpr int CW = 512;        // threads of the coarse-cycle workgroup (1024 and 256 measured slower)
constexpr int CLEV_MAX = 8;    // levels one coarse cycle may hold
constexpr int CTOP_MAX = 32;   // largest level run this way (its sums have <= 16 blocks)

__device__ __forceinline__ i64 cube_d(int n) { return (i64)n * n * n; }

// Levels 0..nl-1 of the coarse cycle (level 0 the finest of them, level nl-1 the 4^3 level), in
// sub-lattice layout.  With mixed, T = float down to 8^3 and the 4^3 level runs in FP64 on the d*
// arrays (its T arrays receive the converted result), as in the multi-launch mixed cycle.
template <class T> struct CoarseCycle {
  T* u[CLEV_MAX];
  T* rhs[CLEV_MAX];
  T* r[CLEV_MAX];
  const T* G[CLEV_MAX];
  const T* A[CLEV_MAX];    // Galerkin operator of the level, or null (19-point operator from G)
  double* scal[CLEV_MAX];  // per level: [0..2] as the multi-launch cycle's scalars (sum in [2])
  int* guard[CLEV_MAX];
  T dx[CLEV_MAX];
  int n[CLEV_MAX];
  int sw[CLEV_MAX];       // sweeps each way on the level (pre == post)
  int stage[CLEV_MAX];    // k_coarse_cycle_lds: the level's u, rhs and G are held in LDS for the cycle
  int nl;
  int coarse;             // sweeps each way on 4^3
  int ref;                // reference schedule: residual norms and the divergence guard
  int top;                // level 0 is the hierarchy's finest level (no guard there)
  int symmetric;          // post-smoothing in reverse colour order
  int mixed;
  double* du;             // FP64 4^3 level (mixed)
  double* drhs;
  const double* dG;
  const double* dA;
  double* dscal;
  int* dguard;
  double ddx;
};

// One stage of reduce() (k_reduce_stage, op 0) over x[0..n) by the whole workgroup: block b's
// value, over elements [b*RB*RPT, (b+1)*RB*RPT), is formed by one group of RB threads exactly as
// k_reduce_stage forms it, four blocks at a time.  Block values go to out[b] (shared memory).
template <class S> __device__ void cw_stage(double* out, const S* x, i64 n, double (*s)[RB]) {
  const int t = threadIdx.x % RB, g = threadIdx.x / RB;
  constexpr int NG = CW / RB;
  const i64 ch = (i64)RB * RPT;
  const int nb = (int)((n + ch - 1) / ch);
  for (int b0 = 0; b0 < nb; b0 += NG) {
    const int b = b0 + g;
    double a = 0.0;
    if (b < nb) {
      const i64 base = (i64)b * ch;
#pragma unroll
      for (int m = 0; m < RPT; ++m) {
        const i64 q = base + (i64)m * RB + t;
        if (q < n) {
          const double v = (double)x[q];
          a = a + v;
        }
      }
    }
    s[g][t] = a;
    __syncthreads();
    for (int h = RB / 2; h > 0; h >>= 1) {
      if (t < h) s[g][t] = s[g][t] + s[g][t + h];
      __syncthreads();
    }
    if (t == 0 && b < nb) out[b] = s[g][0];
    __syncthreads();
  }
}

// reduce(x, n, op 0): the stages until one value is left; returns it (every thread).
template <class S> __device__ double cw_sum(const S* x, i64 n, double* blk, double* blk2, double (*s)[RB]) {
  const i64 ch = (i64)RB * RPT;
  i64 nb = (n + ch - 1) / ch;
  cw_stage<S>(blk, x, n, s);
  double* a = blk;
  double* b = blk2;
  while (nb > 1) {
    const i64 nb2 = (nb + ch - 1) / ch;
    cw_stage<double>(b, a, nb, s);
    double* tmp = a;
    a = b;
    b = tmp;
    nb = nb2;
  }
  return a[0];
}

// ns Gauss-Seidel sweeps of a level (colour order reversed when rev), one barrier per colour.
template <class T>
__device__ void cw_sweeps(T* u, const T* rhs, const T* G, const T* A, int n, T dx, int ns, bool rev) {
  const T dx2 = dx * dx;
  const i64 h3 = cube_d(n / 2);
  for (int sw = 0; sw < ns; ++sw)
    for (int cc = 0; cc < 8; ++cc) {
      const int col = rev ? 7 - cc : cc;
      const int kp = col >> 2, io = (col >> 1) & 1, jo = col & 1;
      if (A)
        for (i64 q = threadIdx.x; q < h3; q += CW)
          gal_colour_point<T, false, T, true>(u, rhs, A, n, kp, io, jo, q);
      else
        for (i64 q = threadIdx.x; q < h3; q += CW)
          colour_point<T, 1, false, false, T, true>(u, rhs, G, n, dx2, kp, io, jo, q);
      __syncthreads();
    }
}

// Maximum of a over the workgroup (every thread gets it); a maximum does not depend on order.
__device__ double cw_max(double a, double* wm) {
  for (int off = 32; off > 0; off >>= 1) a = dmax(a, __shfl_xor(a, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = a;
  __syncthreads();
  double m = wm[0];
  for (int v = 1; v < CW / 64; ++v) m = dmax(m, wm[v]);
  __syncthreads();
  return m;
}

// max |rhs - L u| of a level (the residual norm of relax, stored nowhere).
template <class T>
__device__ double cw_resid_norm(const T* u, const T* rhs, const T* G, const T* A, int n, T dx, double* wm) {
  const T di2 = T(1) / (dx * dx);
  const i64 n3 = cube_d(n);
  double a = 0.0;
  for (i64 p = threadIdx.x; p < n3; p += CW) {
    int i, j, k;
    point_of<1, true>(p, n, i, j, k);
    const T rv = A ? gal_resid_point<T, T, true>(u, rhs, A, n, i, j, k)
                   : resid_point<T, 1, false, T, true>(u, rhs, G, n, di2, i, j, k);
    a = dmax(a, (double)fabs(rv));
  }
  return cw_max(a, wm);
}

// zerosum (k_zerosum_guard): u -= sum(u)/n^3 with the sum in reduce()'s order (stored where the
// multi-launch cycle keeps it), or u = 0 when the guard of mode 1 (relaxerr > r1) or mode 2
// (relaxerr > 1.001 r1 and r1 > 1e-5) fires; mode 0 has no guard.
template <class T>
__device__ void cw_zerosum(T* u, int n, double* scal, int* guard, int mode, double* blk, double* blk2, double (*s)[RB]) {
  const i64 n3 = cube_d(n);
  const double sum = cw_sum<T>(u, n3, blk, blk2, s);
  bool g = false;
  if (mode == 1) g = (T)scal[1] > (T)scal[0];
  else if (mode == 2) g = ((T)scal[1] > (T)scal[0] * T(1.001) && (T)scal[0] > T(1.e-5));
  if (threadIdx.x == 0) {
    scal[2] = sum;
    if (guard) *guard = g ? 1 : 0;
  }
  const double m = sum / (double)n3;
  for (i64 p = threadIdx.x; p < n3; p += CW) u[p] = g ? T(0) : (T)((double)u[p] - m);
  __syncthreads();
}

// The 4^3 level: coarse sweeps forward, then forward or reversed (with the reference schedule each
// followed by the residual norm, r1 then relaxerr), then the mean removal (guard mode 2 when the
// level is not the finest).
template <class T>
__device__ void cw_bottom(T* u, const T* rhs, const T* G, const T* A, T dx, int nc, bool rev, double* scal, int* guard,
                          int ref, int mode, double* blk, double* blk2, double (*s)[RB], double* wm) {
  cw_sweeps<T>(u, rhs, G, A, 4, dx, nc, false);
  if (ref) {
    const double r1 = cw_resid_norm<T>(u, rhs, G, A, 4, dx, wm);
    if (threadIdx.x == 0) scal[0] = r1;
  }
  cw_sweeps<T>(u, rhs, G, A, 4, dx, nc, rev);
  if (ref) {
    const double re = cw_resid_norm<T>(u, rhs, G, A, 4, dx, wm);
    if (threadIdx.x == 0) scal[1] = re;
  }
  __syncthreads();
  cw_zerosum<T>(u, 4, scal, guard, mode, blk, blk2, s);
}

// The cycle, run by the whole workgroup (k_coarse_cycle, and the preconditioner of the
// one-workgroup-per-grid Krylov solve); blk/blk2 hold a level's block sums (n^3 / 2048 of them).
template <class T>
__device__ void cw_cycle(const CoarseCycle<T>& C, double (*s)[RB], double* blk, double* blk2, double* wm) {
  // down: pre-smoothing, residual (and r1), restriction into the next level's rhs, zero next u
  for (int l = 0; l + 1 < C.nl; ++l) {
    const int n = C.n[l];
    cw_sweeps<T>(C.u[l], C.rhs[l], C.G[l], C.A[l], n, C.dx[l], C.sw[l], false);
    const T di2 = T(1) / (C.dx[l] * C.dx[l]);
    const i64 n3 = cube_d(n), h3 = cube_d(n / 2);
    double rmax = 0.0;
    for (i64 p = threadIdx.x; p < n3; p += CW) {
      int i, j, k;
      point_of<1, true>(p, n, i, j, k);
      const T rv = C.A[l] ? gal_resid_point<T, T, true>(C.u[l], C.rhs[l], C.A[l], n, i, j, k)
                          : resid_point<T, 1, false, T, true>(C.u[l], C.rhs[l], C.G[l], n, di2, i, j, k);
      C.r[l][p] = rv;
      rmax = dmax(rmax, (double)fabs(rv));
    }
    __syncthreads();
    if (C.ref) {
      rmax = cw_max(rmax, wm);
      if (threadIdx.x == 0) C.scal[l][0] = rmax;
    }
    const int pre = C.sw[l];
    const int nrmax = min(4, max(1, n / 4));
    const int nrl = pre > nrmax ? (pre % nrmax ? pre % nrmax : nrmax) : pre;
    for (i64 q = threadIdx.x; q < h3; q += CW) {
      restrict_point<T, 1, true>(C.rhs[l + 1], C.r[l], n, nrl % (n / 2), q);
      C.u[l + 1][q] = T(0);
    }
    __syncthreads();
  }
  // the 4^3 level (guard mode 2 in the reference schedule unless it is the finest level)
  const int lb = C.nl - 1;
  const int mode_b = (C.ref && !(C.top && lb == 0)) ? 2 : 0;
  if (C.mixed) {
    for (int p = threadIdx.x; p < 64; p += CW) {
      C.drhs[p] = (double)C.rhs[lb][p];
      C.du[p] = 0.0;
    }
    __syncthreads();
    cw_bottom<double>(C.du, C.drhs, C.dG, C.dA, C.ddx, C.coarse, C.symmetric, C.dscal, C.dguard, C.ref, mode_b, blk,
                      blk2, s, wm);
    for (int p = threadIdx.x; p < 64; p += CW) C.u[lb][p] = (T)C.du[p];
    __syncthreads();
  } else {
    cw_bottom<T>(C.u[lb], C.rhs[lb], C.G[lb], C.A[lb], C.dx[lb], C.coarse, C.symmetric, C.scal[lb], C.guard[lb],
                 C.ref, mode_b, blk, blk2, s, wm);
  }
  // up: prolongation, post-smoothing (and relaxerr), mean removal or the guard's szero
  for (int l = C.nl - 2; l >= 0; --l) {
    const int n = C.n[l];
    const i64 h3 = cube_d(n / 2);
    for (i64 pp = threadIdx.x; pp < h3; pp += CW) prolong_cell<T, true>(C.u[l], C.u[l + 1], n, 1, pp);
    __syncthreads();
    cw_sweeps<T>(C.u[l], C.rhs[l], C.G[l], C.A[l], n, C.dx[l], C.sw[l], C.symmetric != 0);
    if (C.ref) {
      const double re = cw_resid_norm<T>(C.u[l], C.rhs[l], C.G[l], C.A[l], n, C.dx[l], wm);
      if (threadIdx.x == 0) C.scal[l][1] = re;
      __syncthreads();
    }
    const int mode = (C.ref && !(C.top && l == 0)) ? 1 : 0;
    cw_zerosum<T>(C.u[l], n, C.scal[l], C.guard[l], mode, blk, blk2, s);
  }
}

template <class T> __global__ __launch_bounds__(CW) void k_coarse_cycle(CoarseCycle<T> C) {
  __shared__ double s[CW / RB][RB];
  __shared__ double blk[64], blk2[64];  // block values of a level's sums (32^3: 16 blocks)
  __shared__ double wm[CW / 64];
  cw_cycle<T>(C, s, blk, blk2, wm);
}

// The same cycle with its smallest levels held in LDS.  The reference schedule sweeps the coarse
// levels hundreds of times (nu = 4 NG/N: 128, 256 and 32 + 32 sweeps on 16^3, 8^3 and 4^3 at
// NG = 512), some 3600 barrier-separated colour phases per cycle, each of which waits on the
// latency of its loads; from LDS that wait is a fraction of an L2 round trip.  The levels marked
// in stage[] (chosen on the host by stage_levels: smallest first, while their u, rhs and G, or
// Galerkin A, fit in CLDS_BYTES) are copied in, the cycle runs on the copies with the same
// per-point functions (bit-identical results), and their u is copied back.  Residual arrays,
// per-level scalars and guard flags stay in global memory.
constexpr int CLDS_BYTES = 148 * 1024;

template <class T> void stage_levels(CoarseCycle<T>& C) {
  size_t off = 0;
  for (int l = C.nl - 1; l >= 0; --l) {
    const size_t b = (C.A[l] ? 29 : 8) * (size_t)C.n[l] * C.n[l] * C.n[l] * sizeof(T);
    C.stage[l] = ((C.A[l] || C.G[l]) && off + b <= (size_t)CLDS_BYTES) ? 1 : 0;
    if (C.stage[l]) off += b;
  }
}

template <class T> __global__ __launch_bounds__(CW) void k_coarse_cycle_lds(CoarseCycle<T> Cg) {
  __shared__ double s[CW / RB][RB];
  __shared__ double blk[64], blk2[64];
  __shared__ double wm[CW / 64];
  __shared__ __attribute__((aligned(16))) unsigned char pool[CLDS_BYTES];
  __shared__ CoarseCycle<T> C;  // the level table, with the staged levels pointing into pool
  if (threadIdx.x == 0) {
    C = Cg;
    size_t off = 0;
    for (int l = Cg.nl - 1; l >= 0; --l) {
      if (!Cg.stage[l]) continue;
      const i64 m3 = cube_d(Cg.n[l]);
      T* b = reinterpret_cast<T*>(pool + off);
      C.u[l] = b;
      C.rhs[l] = b + m3;
      if (Cg.A[l]) C.A[l] = b + 2 * m3;  // 27 coefficient fields
      else C.G[l] = b + 2 * m3;
      off += (Cg.A[l] ? 29 : 8) * (size_t)m3 * sizeof(T);
    }
  }
  __syncthreads();
  for (int l = 0; l < Cg.nl; ++l) {
    if (!Cg.stage[l]) continue;
    const i64 m3 = cube_d(Cg.n[l]);
    T* u = C.u[l];
    T* rh = C.rhs[l];
    for (i64 t = threadIdx.x; t < m3; t += CW) {
      u[t] = Cg.u[l][t];
      rh[t] = Cg.rhs[l][t];
    }
    if (Cg.A[l]) {
      T* a = const_cast<T*>(C.A[l]);
      for (i64 t = threadIdx.x; t < 27 * m3; t += CW) a[t] = Cg.A[l][t];
    } else {
      T* g = const_cast<T*>(C.G[l]);
      for (i64 t = threadIdx.x; t < 6 * m3; t += CW) g[t] = Cg.G[l][t];
    }
  }
  __syncthreads();
  cw_cycle<T>(C, s, blk, blk2, wm);
  for (int l = 0; l < Cg.nl; ++l) {
    if (!Cg.stage[l]) continue;
    const i64 m3 = cube_d(Cg.n[l]);
    for (i64 t = threadIdx.x; t < m3; t += CW) Cg.u[l][t] = C.u[l][t];
  }
}

// ---------------------------------------------------------------------------------------------
// The whole flexible-PCG solve of one grid in one workgroup (batched solve, grids up to 64^3):
// workgroup g solves grid g of a stack of B equal-size grids to its own tolerance, with the
// preconditioner of the multi-launch batched solve (one symmetric V-cycle, cw_cycle, from the
// finest level down) and the same recurrence (Polak-Ribiere beta, FP64 throughout).  No grid waits
// for another: the launch replaces the batched iteration loop, its launches and its host round
// trips.  Sums are blocked like reduce()'s but differ in order from the multi-launch solve's, so
// the iterates agree with it to rounding, not bit for bit.

// sum of a[p] * b[p] over n3 points by the workgroup (every thread gets it): thread partials in
// point order, then the workgroup's halving tree.
__device__ double cw_dot(const double* a, const double* b, i64 n3, double (*s)[RB]) {
  double acc = 0.0;
  for (i64 p = threadIdx.x; p < n3; p += CW) acc = acc + a[p] * b[p];
  double* f = &s[0][0];
  f[threadIdx.x] = acc;
  __syncthreads();
  for (int h = CW / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) f[threadIdx.x] = f[threadIdx.x] + f[threadIdx.x + h];
    __syncthreads();
  }
  const double r = f[0];
  __syncthreads();
  return r;
}

// The levels of a stack of B equal-size grids (batched solve), grid y of level l at
// u[l] + y m_l^3 (G: 6 m_l^3 per grid, A: 27 m_l^3), as the batched launches store them.
struct StackLevels {
  int nl, coarse, symmetric;
  double* u[CLEV_MAX];
  double* rhs[CLEV_MAX];
  double* r[CLEV_MAX];
  const double* G[CLEV_MAX];
  const double* A[CLEV_MAX];
  int m[CLEV_MAX], sw[CLEV_MAX];
  double dx[CLEV_MAX];
  double* scal;  // 3 CLEV_MAX per grid: the cycle's per-level scalars
};

// Thread 0 writes grid g's cycle table (Krylov schedule: no guard, mean removal on every level).
__device__ void stack_table(CoarseCycle<double>& C, const StackLevels& L, int g) {
  if (threadIdx.x != 0) return;
  C.nl = L.nl;
  for (int l = 0; l < L.nl; ++l) {
    const i64 m3 = cube_d(L.m[l]);
    C.u[l] = L.u[l] + (i64)g * m3;
    C.rhs[l] = L.rhs[l] + (i64)g * m3;
    C.r[l] = L.r[l] + (i64)g * m3;
    C.G[l] = L.G[l] ? L.G[l] + (i64)g * 6 * m3 : nullptr;
    C.A[l] = L.A[l] ? L.A[l] + (i64)g * 27 * m3 : nullptr;
    C.scal[l] = L.scal + (i64)g * 3 * CLEV_MAX + 3 * l;
    C.guard[l] = nullptr;
    C.dx[l] = L.dx[l];
    C.n[l] = L.m[l];
    C.sw[l] = L.sw[l];
  }
  C.coarse = L.coarse;
  C.ref = 0;
  C.top = 0;
  C.symmetric = L.symmetric;
  C.mixed = 0;
}

// The coarse end of the batched preconditioning cycle: workgroup g runs grid g's levels (from the
// table's level 0 down to 4^3 and back) as k_coarse_cycle runs a single grid's, bit-identical to
// the batched launches; grids with act[g] = 0 (converged) are skipped.
__global__ __launch_bounds__(CW) void k_coarse_stack(StackLevels L, const int* __restrict__ act) {
  __shared__ double s[CW / RB][RB];
  __shared__ double blk[64], blk2[64];
  __shared__ double wm[CW / 64];
  // the level table once per workgroup in LDS (in registers it would take a scratch copy per lane)
  __shared__ CoarseCycle<double> C;
  if (act && !act[blockIdx.x]) return;
  stack_table(C, L, blockIdx.x);
  __syncthreads();
  cw_cycle<double>(C, s, blk, blk2, wm);
}

struct PcgStack {
  int n, B, max_it;
  double tol;
  StackLevels L;                        // level 0: z (u) and the residual (rhs) of the finest level
  double *x, *b, *cr, *ro, *p, *q, *z;  // finest-level vectors
  const double* rden;
  int *active, *iters;
  double* relres;
};

__global__ __launch_bounds__(CW) void k_pcg_stack(PcgStack P) {
  __shared__ double s[CW / RB][RB];
  __shared__ double blk[256], blk2[256];  // up to 64^3: 128 blocks
  __shared__ double wm[CW / 64];
  __shared__ CoarseCycle<double> C;
  const int g = blockIdx.x;
  const int n = P.n;
  const i64 n3 = cube_d(n);
  const i64 o = (i64)g * n3;
  double *x = P.x + o, *b = P.b + o, *r = P.cr + o, *ro = P.ro + o, *p = P.p + o, *q = P.q + o, *z = P.z + o;
  stack_table(C, P.L, g);
  for (i64 t = threadIdx.x; t < n3; t += CW) {
    x[t] = 0.0;
    r[t] = b[t];
  }
  __syncthreads();
  if (!P.active[g]) {  // converged from the start (the projected right-hand side vanishes)
    if (threadIdx.x == 0) P.iters[g] = 0;
    return;
  }
  auto precond = [&] {
    for (i64 t = threadIdx.x; t < n3; t += CW) z[t] = 0.0;
    __syncthreads();
    cw_cycle<double>(C, s, blk, blk2, wm);
  };
  precond();
  double rz = cw_dot(z, r, n3, s);
  for (i64 t = threadIdx.x; t < n3; t += CW) p[t] = z[t];
  __syncthreads();
  const double dx2 = P.L.dx[0] * P.L.dx[0];
  for (int it = 0; it < P.max_it; ++it) {
    // q = L p (the 19-point operator of the finest level)
    for (i64 t = threadIdx.x; t < n3; t += CW) {
      int i, j, k;
      point_of<1, true>(t, n, i, j, k);
      double t1, t2, td;
      stencil19<double, 1, false, false, double, true>(p, C.G[0], n3, n, i, j, k, t1, t2, td);
      q[t] = (t1 + t2 + td * p[t]) / dx2;
    }
    __syncthreads();
    const double pq = cw_dot(p, q, n3, s);
    const double alpha = rz / pq;
    // breakdown (p.Lp = 0, or no finite z): x and r keep their last finite values and the grid
    // stays active, reported as not converged
    if (!isfinite(alpha)) break;
    double rmax = 0.0;
    for (i64 t = threadIdx.x; t < n3; t += CW) {
      x[t] = x[t] + alpha * p[t];
      ro[t] = r[t];
      r[t] = r[t] - alpha * q[t];
      rmax = dmax(rmax, fabs(r[t]));
    }
    __syncthreads();
    rmax = cw_max(rmax, wm);
    const double rr = rmax / P.rden[g];
    if (threadIdx.x == 0) {
      P.relres[g] = rr;
      P.iters[g] = it + 1;
    }
    if (rr <= P.tol) {
      if (threadIdx.x == 0) P.active[g] = 0;
      break;
    }
    precond();
    const double rzn = cw_dot(z, r, n3, s);
    const double zro = cw_dot(z, ro, n3, s);
    const double beta = (rzn - zro) / rz;
    for (i64 t = threadIdx.x; t < n3; t += CW) p[t] = z[t] + beta * p[t];
    __syncthreads();
    rz = rzn;
  }
}

}  // namespace mmh
