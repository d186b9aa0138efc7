// kernels_adj.hpp -- power spectrum binning and the adjoint of an outer iteration's mesh solve
// (gradient of binned reconstructed power with respect to the input density and to the mesh).
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
#include "kernels_lim.hpp"

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

// Sensitivities of lambda^T L phi (the 19-point curvilinear operator, relaxplane/calcresid,
// mmh/gauss1.f90:954-980, dx = 1) to the six stored metric components at a point p:
//   W_aa = -1/2 [ (l(p+a) - l(p))(f(p+a) - f(p)) + (l(p-a) - l(p))(f(p-a) - f(p)) ]
//   W_ab = -1/4 [ D_b l D_a f + D_a l D_b f ],  D_a g = g(p+a) - g(p-a)        (a != b)
// (G_aa enters the two face coefficients beside p; G_ab the cross terms of the four neighbours
// of p in the (a, b) plane).  With A = I + H(def) (cbajraw's, mmh/gauss1.f90:1128-1147) and
// G = rhobar adj(A) / det (1185-1204; G = rhobar I where det <= 1e-5, 1162-1171), the metric part
// of dJ = lambda^T (db - dL phi) is -sum_p W : dG, which gives
//   g_rhobar(p) = -W : (adj(A) / det)                    (W : I where the tangle clamp is on)
//   S(p)        = rhobar A^-1 Wt A^-1 + gdet adj(A)      (0 + gdet adj(A) where it is on)
// with Wt the symmetric matrix of W (halves off the diagonal) and gdet(p) = s (ld / dsum - l(p))
// the sensitivity to rgzerosum's unclamped det (b = r - det s, s = sum r / dsum, ld = sum l det,
// mmh/gauss1.f90:1259-1407).  S holds dJ/dA (xx, yy, zz, xy, xz, yz: the off-diagonal entries
// counted once for the pair).  Column-major fields.
__global__ void k_adj_metric(double* __restrict__ grhobar, double* __restrict__ S, const double* __restrict__ lam,
                             const double* __restrict__ phi, const double* __restrict__ def,
                             const double* __restrict__ rho, int n, const double* __restrict__ sums) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int m = n - 1;
  auto L = [&](int a, int b, int c) { return lam[ix3((i + a) & m, (j + b) & m, (k + c) & m, n)]; };
  auto F = [&](int a, int b, int c) { return phi[ix3((i + a) & m, (j + b) & m, (k + c) & m, n)]; };
  const double l0 = lam[p], f0 = phi[p];
  const double wxx = -0.5 * ((L(1, 0, 0) - l0) * (F(1, 0, 0) - f0) + (L(-1, 0, 0) - l0) * (F(-1, 0, 0) - f0));
  const double wyy = -0.5 * ((L(0, 1, 0) - l0) * (F(0, 1, 0) - f0) + (L(0, -1, 0) - l0) * (F(0, -1, 0) - f0));
  const double wzz = -0.5 * ((L(0, 0, 1) - l0) * (F(0, 0, 1) - f0) + (L(0, 0, -1) - l0) * (F(0, 0, -1) - f0));
  const double dlx = L(1, 0, 0) - L(-1, 0, 0), dly = L(0, 1, 0) - L(0, -1, 0), dlz = L(0, 0, 1) - L(0, 0, -1);
  const double dfx = F(1, 0, 0) - F(-1, 0, 0), dfy = F(0, 1, 0) - F(0, -1, 0), dfz = F(0, 0, 1) - F(0, 0, -1);
  const double wxy = -0.25 * (dly * dfx + dlx * dfy);
  const double wxz = -0.25 * (dlz * dfx + dlx * dfz);
  const double wyz = -0.25 * (dlz * dfy + dly * dfz);
  const Hs<double> h = hess(def, n, i, j, k);
  const double a11 = 1.0 + h.xx, a12 = h.xy, a13 = h.xz, a22 = 1.0 + h.yy, a23 = h.yz, a33 = 1.0 + h.zz;
  const double det = a11 * a22 * a33 + 2.0 * a12 * a13 * a23 - a13 * a13 * a22 - a11 * a23 * a23 - a12 * a12 * a33;
  // adj(A) (cofactors, symmetric)
  const double c11 = a22 * a33 - a23 * a23, c12 = a13 * a23 - a12 * a33, c13 = a12 * a23 - a13 * a22;
  const double c22 = a11 * a33 - a13 * a13, c23 = a12 * a13 - a11 * a23, c33 = a11 * a22 - a12 * a12;
  double rb = 0.0;
  for (int c = -1; c <= 1; ++c)
    for (int b = -1; b <= 1; ++b)
      for (int a = -1; a <= 1; ++a) rb += rho[ix3((i + a) & m, (j + b) & m, (k + c) & m, n)];
  rb = rb / 27.0;
  const double s = sums[0], ld_over_dsum = sums[1];
  const double gdet = s * (ld_over_dsum - l0);
  double sxx = gdet * c11, syy = gdet * c22, szz = gdet * c33;
  double sxy = 2.0 * gdet * c12, sxz = 2.0 * gdet * c13, syz = 2.0 * gdet * c23;
  if (det <= 1.e-5) {
    grhobar[p] = -(wxx + wyy + wzz);
  } else {
    grhobar[p] = -(wxx * c11 + wyy * c22 + wzz * c33 + wxy * c12 + wxz * c13 + wyz * c23) / det;
    // M = A^-1 Wt A^-1 with A^-1 = adj / det
    const double Wt[3][3] = {{wxx, 0.5 * wxy, 0.5 * wxz}, {0.5 * wxy, wyy, 0.5 * wyz}, {0.5 * wxz, 0.5 * wyz, wzz}};
    const double Ai[3][3] = {{c11, c12, c13}, {c12, c22, c23}, {c13, c23, c33}};
    double T1[3][3], M[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) T1[a][b] = Ai[a][0] * Wt[0][b] + Ai[a][1] * Wt[1][b] + Ai[a][2] * Wt[2][b];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) M[a][b] = T1[a][0] * Ai[0][b] + T1[a][1] * Ai[1][b] + T1[a][2] * Ai[2][b];
    const double f = rb / (det * det);
    sxx += f * M[0][0]; syy += f * M[1][1]; szz += f * M[2][2];
    sxy += 2.0 * f * M[0][1]; sxz += 2.0 * f * M[0][2]; syz += 2.0 * f * M[1][2];
  }
  S[p] = sxx; S[n3 + p] = syy; S[2 * n3 + p] = szz; S[3 * n3 + p] = sxy; S[4 * n3 + p] = sxz; S[5 * n3 + p] = syz;
}

// g_def += the transpose of A's difference stencils applied to S (each stencil of hess is its own
// transpose under the periodic wrap): d(def) enters A_aa through (f(+a) - 2 f + f(-a)) and A_ab
// through (f(+a+b) - f(-a+b) - f(+a-b) + f(-a-b)) / 4.
__global__ void k_adj_hess(double* __restrict__ gdef, const double* __restrict__ S, int n) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int m = n - 1;
  auto V = [&](int c, int a, int b, int d) { return S[(i64)c * n3 + ix3((i + a) & m, (j + b) & m, (k + d) & m, n)]; };
  double g = (V(0, 1, 0, 0) - 2.0 * V(0, 0, 0, 0) + V(0, -1, 0, 0)) + (V(1, 0, 1, 0) - 2.0 * V(1, 0, 0, 0) + V(1, 0, -1, 0)) +
             (V(2, 0, 0, 1) - 2.0 * V(2, 0, 0, 0) + V(2, 0, 0, -1));
  g += (V(3, 1, 1, 0) - V(3, -1, 1, 0) - V(3, 1, -1, 0) + V(3, -1, -1, 0)) / 4.0;
  g += (V(4, 1, 0, 1) - V(4, -1, 0, 1) - V(4, 1, 0, -1) + V(4, -1, 0, -1)) / 4.0;
  g += (V(5, 0, 1, 1) - V(5, 0, -1, 1) - V(5, 0, 1, -1) + V(5, 0, -1, -1)) / 4.0;
  gdef[p] = gdef[p] + g;
}

// x <- x - mean: mean = sums[0] / n3.
__global__ void k_sub_mean(double* __restrict__ x, const double* __restrict__ sums, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) x[p] = x[p] - sums[0] / (double)n3;
}

// x <- x - v[0].
__global__ void k_sub_val(double* __restrict__ x, const double* __restrict__ v, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) x[p] = x[p] - v[0];
}

// out = a * b elementwise.
__global__ void k_mul(double* __restrict__ out, const double* __restrict__ a, const double* __restrict__ b, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) out[p] = a[p] * b[p];
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
