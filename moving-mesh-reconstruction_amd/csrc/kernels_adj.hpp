// kernels_adj.hpp -- power spectrum binning and the adjoint of the first outer iteration's mesh
// solve (gradient of binned reconstructed power with respect to the input density).
//
// Forward map at def = 0 (the first outer iteration, mmh/main.f90:95-98 with mmh/limiter.f90):
//   c   = clip(1 - rho, +-2) / 30                          limiter.f90:157-165
//   r   = S^4 c   (nsmooth x3, eflux = tmp exactly at def = 0, nsmooth x1)   limiter.f90:171-392
//   b   = r - mean(r)                                      rgzerosum, det = 1 at def = 0
//   phi = mean-free solution of L(rhobar) phi = b          G = rhobar I at def = 0 (cbajraw triad)
//   delta = -lap_spectral(S^2 phi)                         limiter.f90:421-423, reconstruct.f90
//   P_b = sum_{k in bin b} |delta_k|^2 / (N_b N^3),  loss = sum_b w_b P_b
// The expansion limiter, the compression gate and the +-1/dtaumesh clamp cannot fire at def = 0
// (det = 1, all eigenvalues 1, |S^3 c| <= 1/15); the compression limiter after the solve is checked
// at run time and the gradient is refused where it fires.
#pragma once
#include "kernels.hpp"

namespace mmh {

// Bin of the packed spectral entry (ic, j, k) of an NR real FFT (ic = n/2 is the speq plane):
// |k| = 2 pi |m| / n with m = (ic, signed j, signed k); bins are [e_b, e_{b+1}) on log-spaced edges,
// the last bin closed.  Returns -1 outside [e_0, e_nb].  mult = modes the entry stands for.
__device__ __forceinline__ int spec_bin(int ic, int j, int k, int n, const double* __restrict__ edges, int nb,
                                        int& mult) {
  const int h = n / 2;
  const double mx = ic, my = (j <= h) ? j : j - n, mz = (k <= h) ? k : k - n;
  const double kk = 2.0 * M_PI * sqrt(mx * mx + my * my + mz * mz) / (double)n;
  mult = (ic == 0 || ic == h) ? 1 : 2;
  if (kk < edges[0] * (1.0 - 1e-12) || kk > edges[nb] * (1.0 + 1e-12)) return -1;
  int b = 0;
  while (b < nb - 1 && kk >= edges[b + 1]) ++b;
  return b;
}

// Per-block sums over the packed spectrum (data, speq) of mult*|F|^2 and mult per bin:
// out[block][0..nb) power, out[block][nb..2nb) mode counts.
__global__ __launch_bounds__(256) void k_bin_power(double* __restrict__ out, const double* __restrict__ data,
                                                   const double* __restrict__ speq, int n,
                                                   const double* __restrict__ edges, int nb) {
  __shared__ double s[256];
  const int h = n / 2;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1, mult = 0;
  double pw = 0.0;
  if (q < (i64)(h + 1) * n * n) {
    const int ic = (int)(q % (h + 1));
    const i64 r = q / (h + 1);
    const int j = (int)(r % n), k = (int)(r / n);
    const double* re = ic < h ? &data[ix3(2 * ic, j, k, n)] : &speq[2 * ((i64)j + (i64)n * k)];
    b = spec_bin(ic, j, k, n, edges, nb, mult);
    pw = (double)mult * (re[0] * re[0] + re[1] * re[1]);
  }
  for (int bb = 0; bb < nb; ++bb) {
    for (int part = 0; part < 2; ++part) {
      s[threadIdx.x] = (b == bb) ? (part == 0 ? pw : (double)mult) : 0.0;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) out[(i64)blockIdx.x * 2 * nb + part * nb + bb] = s[0];
      __syncthreads();
    }
  }
}

// Multiplies every packed entry by sym[bin] (0 outside the bins): the real, even symbol of the
// loss's second-derivative filter.
__global__ void k_bin_filter(double* __restrict__ data, double* __restrict__ speq, int n,
                             const double* __restrict__ edges, int nb, const double* __restrict__ sym) {
  const int h = n / 2;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)(h + 1) * n * n) return;
  const int ic = (int)(q % (h + 1));
  const i64 r = q / (h + 1);
  const int j = (int)(r % n), k = (int)(r / n);
  double* re = ic < h ? &data[ix3(2 * ic, j, k, n)] : &speq[2 * ((i64)j + (i64)n * k)];
  int mult;
  const int b = spec_bin(ic, j, k, n, edges, nb, mult);
  const double f = b >= 0 ? sym[b] : 0.0;
  re[0] = re[0] * f;
  re[1] = re[1] * f;
}

// rho (N^3) -> the reference's fluid array u(5,N,N,N) with u(1) = rho and u(2:5) = 0.
template <class T> __global__ void k_to_aos(T* __restrict__ u, const T* __restrict__ rho, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  u[5 * p] = rho[p];
  u[5 * p + 1] = 0;
  u[5 * p + 2] = 0;
  u[5 * p + 3] = 0;
  u[5 * p + 4] = 0;
}

// Sensitivity of lambda^T L(G) phi to the metric at def = 0, where G11 = G22 = G33 = rhobar and
// the off-diagonal coefficients vanish.  The operator is a sum over faces f = (p, p+e_d) of
// -b_f (lambda_{p+e_d} - lambda_p)(phi_{p+e_d} - phi_p) / dx^2 with b_f the mean of the two points'
// coefficient (relaxplane, mmh/gauss1.f90:954-968), so
//   g[q] = -d(lambda^T L phi)/d rhobar[q]
//        = 1/2 sum_d [ (l_{q+d} - l_q)(f_{q+d} - f_q) + (l_q - l_{q-d})(f_q - f_{q-d}) ] / dx^2.
// Column-major fields.
__global__ void k_adj_rhobar(double* __restrict__ g, const double* __restrict__ lam, const double* __restrict__ phi,
                             int n, double dx) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (i64)n * n * n) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int m = n - 1;
  const i64 nb[6] = {ix3((i + 1) & m, j, k, n), ix3((i - 1) & m, j, k, n), ix3(i, (j + 1) & m, k, n),
                     ix3(i, (j - 1) & m, k, n), ix3(i, j, (k + 1) & m, n), ix3(i, j, (k - 1) & m, n)};
  const double l0 = lam[p], f0 = phi[p];
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 6; ++d) s += (lam[nb[d]] - l0) * (phi[nb[d]] - f0);
  g[p] = 0.5 * s / (dx * dx);
}

// Adjoint of rhobar = (sum of u1 over the 27-point box) / 27 (cbajraw, mmh/gauss1.f90:1185-1197):
// the same box sum of the sensitivities, / 27.
__global__ void k_box27_adj(double* __restrict__ out, const double* __restrict__ g, int n) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (i64)n * n * n) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int m = n - 1;
  double s = 0.0;
  for (int c = -1; c <= 1; ++c)
    for (int b = -1; b <= 1; ++b)
      for (int a = -1; a <= 1; ++a) s += g[ix3((i + a) & m, (j + b) & m, (k + c) & m, n)];
  out[p] = s / 27.0;
}

// Gradient through the right-hand side: d c / d rho = -1 where |1 - rho| <= 2 (clip inactive;
// limiter.f90:164 clips only when |t| > 2), then / 30 / dtold.  grad += that times sm = S^4 lambda.
__global__ void k_rhs_adj(double* __restrict__ grad, const double* __restrict__ sm, const double* __restrict__ rho,
                          i64 n3, double dtold) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const double t = 1.0 - rho[p];
  const double d = fabs(t) > 2.0 ? 0.0 : -(sm[p] / 30.0 / dtold);
  grad[p] = grad[p] + d;
}

}  // namespace mmh
