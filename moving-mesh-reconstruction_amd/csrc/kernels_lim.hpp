// kernels_lim.hpp -- grid limiter (calcdefp), relaxing-TVD transport and FFT kernels.
// Same conventions as kernels.hpp: column-major, periodic, reference association order,
// built with -ffp-contract=off.
#pragma once
#include "kernels.hpp"

namespace mmh {

template <class T> struct Hs { T xx, yy, zz, xy, yz, xz; };

// Hessian stencil without grid scaling (mmh/limiter.f90:268-273, mmh/relaxing.f90:876-881).
template <class T> __device__ __forceinline__ Hs<T> hess(const T* __restrict__ f, int n, int i, int j, int k) {
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
  const T f0 = f[ix3(i, j, k, n)];
  Hs<T> h;
  h.xx = (f[ix3(ip, j, k, n)] - T(2) * f0 + f[ix3(im, j, k, n)]);
  h.yy = (f[ix3(i, jp, k, n)] - T(2) * f0 + f[ix3(i, jm, k, n)]);
  h.zz = (f[ix3(i, j, kp, n)] - T(2) * f0 + f[ix3(i, j, km, n)]);
  h.xy = (f[ix3(ip, jp, k, n)] - f[ix3(im, jp, k, n)] - f[ix3(ip, jm, k, n)] + f[ix3(im, jm, k, n)]) / T(4);
  h.yz = (f[ix3(i, jp, kp, n)] - f[ix3(i, jp, km, n)] - f[ix3(i, jm, kp, n)] + f[ix3(i, jm, km, n)]) / T(4);
  h.xz = (f[ix3(ip, j, kp, n)] - f[ix3(im, j, kp, n)] - f[ix3(ip, j, km, n)] + f[ix3(im, j, km, n)]) / T(4);
  return h;
}

// b = adj(A)/det, det, A = I + H, no clamp (mmh/limiter.f90:274-286, mmh/relaxing.f90:882-901).
// B planes: 0..5 = b11,b12,b13,b22,b23,b33, 6 = det.
template <class T>
__global__ __launch_bounds__(256) void k_binv(T* __restrict__ B, const T* __restrict__ def, int n) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  Hs<T> h = hess(def, n, i, j, k);
  T a11 = T(1) + h.xx, a12 = h.xy, a13 = h.xz, a22 = T(1) + h.yy, a23 = h.yz, a33 = T(1) + h.zz;
  T det = a11 * a22 * a33 + T(2) * a12 * a13 * a23 - (a13 * a13) * a22 - a11 * (a23 * a23) - (a12 * a12) * a33;
  B[0 * n3 + p] = (a22 * a33 - a23 * a23) / det;
  B[1 * n3 + p] = (a13 * a23 - a12 * a33) / det;
  B[2 * n3 + p] = (a12 * a23 - a13 * a22) / det;
  B[3 * n3 + p] = (a11 * a33 - a13 * a13) / det;
  B[4 * n3 + p] = (a12 * a13 - a11 * a23) / det;
  B[5 * n3 + p] = (a11 * a22 - a12 * a12) / det;
  B[6 * n3 + p] = det;
}

// Smallest eigenvalue of symmetric C by the trigonometric method; returns 1/lambda_min
// (mmh/limiter.f90:346-358).
template <class T>
__device__ __forceinline__ T compression_of(T c11, T c12, T c13, T c22, T c23, T c33, T pi) {
  T trace = c11 + c22 + c33;
  T d11 = c11 - trace / T(3);
  T d22 = c22 - trace / T(3);
  T a = -(d11 * d11 + c12 * c12 + c13 * c13 + d11 * d22 + d22 * d22 + c23 * c23);
  T b = -(d11 * (c12 * c12)) + (d11 * d11) * d22 - (c12 * c12) * d22 + (c13 * c13) * d22 + d11 * (d22 * d22)
        - T(2) * c12 * c13 * c23 + d11 * (c23 * c23);
  if (fabs(a) < T(1.e-20)) a = T(1.e-20);
  T am = T(2) * sqrt(fabs(-a) / T(3));
  T arg = T(3) * b / a / am;
  arg = dmax(T(-1.), dmin(T(1.), arg));
  T theta = acos(arg) / T(3);
  T r1 = am * cos(theta);
  T r2 = am * cos(theta + T(2) * pi / T(3.));
  T r3 = am * cos(theta + T(4) * pi / T(3.));
  T t3 = trace / T(3);
  T l1 = r1 + t3, l2 = r2 + t3, l3 = r3 + t3;
  return T(1) / dmin(dmin(l1, l2), l3);
}

// Step 1 of calcdefp: clipped, scaled mass mismatch (mmh/limiter.f90:134-169).
template <class T>
__global__ void k_mismatch(T* __restrict__ tmp, const T* __restrict__ u, int nf, i64 n3, T dtold) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const T u1 = u[(i64)nf * p];
  T t1 = T(1.0f) - u1;
  const T xhigh = 2;
  if (fabs(t1) > xhigh) t1 = copysign(xhigh, t1);
  tmp[p] = t1 / T(30) / dtold;
}

// Step 3 of calcdefp: the eflux of plane k-1 from the metric and eigenvalues of plane k
// (mmh/limiter.f90:238-381), including the expansion limiter (det > 10) and the compression gate.
// stat: per-block [nexpand, detmax, detmin, fluxmax, fluxmin] combined by the host.
template <class T>
__global__ __launch_bounds__(256) void k_eflux(T* __restrict__ eflux, const T* __restrict__ tmp, const T* __restrict__ B,
                                               const T* __restrict__ def, const T* __restrict__ u, int nf, int n, T dtold,
                                               T dtaumesh, T cmpmaxglobal, T pi, double* __restrict__ bstat,
                                               T* __restrict__ gate = nullptr) {
  __shared__ double s[5][256];
  const i64 n3 = (i64)n * n * n;
  const i64 pk = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  double st[5] = {0.0, 0.0, 1.0, 0.0, 0.0};
  if (pk < n3) {
    const int i = (int)(pk % n), j = (int)((pk / n) % n), k = (int)(pk / ((i64)n * n));
    const int km = wrapi(k - 1, n), kmm = wrapi(k - 2, n);
    const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n);
    const i64 pkm = ix3(i, j, km, n);
    const T det = B[6 * n3 + pk];
    const T b13 = B[2 * n3 + pk], b23 = B[4 * n3 + pk], b33 = B[5 * n3 + pk];
#define BM(c, a, b2, kk) B[(i64)(c) * n3 + ix3(a, b2, kk, n)]
#define UU(c, pp) u[(i64)nf * (pp) + (c)]
    i64 q;
    q = ix3(ip, j, km, n);
    T s1p = (BM(0, ip, j, km) * UU(1, q) + BM(1, ip, j, km) * UU(2, q) + BM(2, ip, j, km) * UU(3, q)) / UU(0, q);
    q = ix3(i, jp, km, n);
    T s2p = (BM(1, i, jp, km) * UU(1, q) + BM(3, i, jp, km) * UU(2, q) + BM(4, i, jp, km) * UU(3, q)) / UU(0, q);
    T s3p = (b13 * UU(1, pk) + b23 * UU(2, pk) + b33 * UU(3, pk)) / UU(0, pk);
    q = ix3(im, j, km, n);
    T s1m = (BM(0, im, j, km) * UU(1, q) + BM(1, im, j, km) * UU(2, q) + BM(2, im, j, km) * UU(3, q)) / UU(0, q);
    q = ix3(i, jm, km, n);
    T s2m = (BM(1, i, jm, km) * UU(1, q) + BM(3, i, jm, km) * UU(2, q) + BM(4, i, jm, km) * UU(3, q)) / UU(0, q);
    q = ix3(i, j, kmm, n);
    T s3m = (BM(2, i, j, kmm) * UU(1, q) + BM(4, i, j, kmm) * UU(2, q) + BM(5, i, j, kmm) * UU(3, q)) / UU(0, q);
#undef BM
#undef UU
    T ef = (s1p + s2p + s3p - s1m - s2m - s3m) / T(2);
    ef = ef + tmp[pkm];
    // d eflux / d tmp with every gate frozen at its branch (the adjoint's use; see gate below)
    T dm = T(1);
    if (det > T(10.0f)) {
      dm = ef > T(0) ? -T(2) : (ef < T(0) ? T(2) : T(0));
      ef = -T(2) * fabs(ef);
      st[0] = 1.0;
    }
    Hs<T> h = hess(def, n, i, j, k);
    T compression = compression_of(T(1) + h.xx, h.xy, h.xz, T(1) + h.yy, h.yz, T(1) + h.zz, pi);
    T cmpmax = cmpmaxglobal * T(1);
    T tlim = T(3) * (T(1) / cmpmax - T(1) / compression) / dtold;
    if (tlim > T(0)) {
      if (!(ef > T(0))) dm = T(0);
      ef = dmax(T(0.), ef);
    }
    st[1] = (double)det;
    st[2] = (double)det;
    if (!(ef < T(1.) / dtaumesh && ef > -T(1.) / dtaumesh)) dm = T(0);
    ef = dmin(T(1.) / dtaumesh, ef);
    ef = dmax(-T(1.) / dtaumesh, ef);
    eflux[pkm] = ef;
    // gate: the derivative of this eflux with respect to the smoothed mismatch tmp[pkm] at fixed
    // def (the expansion limiter's -2|.|, the compression gate's max(0, .) and the clamp, each
    // taken on the branch the forward pass took; the e.v terms vanish with u(2:4) = 0)
    if (gate) gate[pkm] = dm;
    st[3] = (double)(ef * dtold);
    st[4] = (double)(ef * dtold);
  }
  for (int c = 0; c < 5; ++c) s[c][threadIdx.x] = st[c];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] = dmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] = dmin(s[2][threadIdx.x], s[2][threadIdx.x + h]);
      s[3][threadIdx.x] = dmax(s[3][threadIdx.x], s[3][threadIdx.x + h]);
      s[4][threadIdx.x] = dmin(s[4][threadIdx.x], s[4][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 5; ++c) bstat[5 * (i64)blockIdx.x + c] = s[c][0];
}

// Step 5 of calcdefp: compression limiter source (mmh/limiter.f90:425-533).  lim = tlim where
// lambda_min(I + H(def) + dtold*H(defp)) < 1/cmpmax; per-block [ncompress, cmax].
template <class T>
__global__ __launch_bounds__(256) void k_complim(T* __restrict__ lim, const T* __restrict__ defp, const T* __restrict__ def,
                                                 int n, T dtold, T cmpmaxglobal, T pi, double* __restrict__ bstat) {
  __shared__ double s[2][256];
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  double c0 = 0.0, c1 = 0.0;
  if (p < n3) {
    const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
    Hs<T> hp = hess(defp, n, i, j, k), hd = hess(def, n, i, j, k);
    T pxx = hp.xx * dtold, pyy = hp.yy * dtold, pzz = hp.zz * dtold;
    T pxy = hp.xy * dtold, pyz = hp.yz * dtold, pxz = hp.xz * dtold;
    pxx = pxx + hd.xx; pyy = pyy + hd.yy; pzz = pzz + hd.zz;
    pxy = pxy + hd.xy; pyz = pyz + hd.yz; pxz = pxz + hd.xz;
    T compression = compression_of(T(1) + pxx, pxy, pxz, T(1) + pyy, pyz, T(1) + pzz, pi);
    T cmpmax = cmpmaxglobal * T(1);
    T tlim = T(3) * (T(1) / cmpmax - T(1) / compression) / dtold;
    T v = T(0.);
    if (tlim > T(0)) { v = tlim; c0 = 1.0; }
    lim[p] = v;
    c1 = (double)compression;
  }
  s[0][threadIdx.x] = c0;
  s[1][threadIdx.x] = c1;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] = dmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { bstat[2 * (i64)blockIdx.x] = s[0][0]; bstat[2 * (i64)blockIdx.x + 1] = s[1][0]; }
}

template <class T> __global__ void k_axpy_scale(T* __restrict__ y, const T* __restrict__ x, int n, i64 n3) {
  // y += ((x*2)/n/n/n)  (mmh/limiter.f90:579,592)
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  T v = x[p] * T(2.) / T(n) / T(n) / T(n);
  y[p] = y[p] + v;
}

template <class T> __global__ void k_add(T* __restrict__ y, const T* __restrict__ x, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) y[p] = y[p] + x[p];
}

// ------------------------------------------------------------------------------------------------
// FFT: radix-2 Stockham per line in LDS.  One workgroup per line; twiddles W[q] = exp(s*2*pi*i*q/L)
// come from a host table so that rounding matches the CPU restatement exactly.
template <class T> struct C2 { T r, i; };

template <class T>
__global__ void k_fft_lines(C2<T>* __restrict__ z, const C2<T>* __restrict__ W, int L, int d0, int d1, int axis) {
  extern __shared__ unsigned char smem_raw[];
  C2<T>* a = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* b = a + L;
  const i64 l = blockIdx.x;
  i64 base, st;
  if (axis == 0) { base = l * (i64)d0; st = 1; }
  else if (axis == 1) { base = (l % d0) + (l / d0) * (i64)d0 * d1; st = d0; }
  else { base = l; st = (i64)d0 * d1; }
  for (int q = threadIdx.x; q < L; q += blockDim.x) a[q] = z[base + q * st];
  __syncthreads();
  int p = 1;
  while (p < L) {
    const int stride = L / (2 * p);
    for (int i = threadIdx.x; i < L / 2; i += blockDim.x) {
      const int k = i & (p - 1);
      const C2<T> w = W[k * stride];
      const C2<T> u0 = a[i], u1 = a[i + L / 2];
      T tr = w.r * u1.r - w.i * u1.i;
      T ti = w.r * u1.i + w.i * u1.r;
      const int jj = 2 * i - k;
      b[jj].r = u0.r + tr; b[jj].i = u0.i + ti;
      b[jj + p].r = u0.r - tr; b[jj + p].i = u0.i - ti;
    }
    __syncthreads();
    C2<T>* t = a; a = b; b = t;
    p <<= 1;
  }
  for (int q = threadIdx.x; q < L; q += blockDim.x) z[base + q * st] = a[q];
}

// rlft3 recombination between the half-length complex transform and the real spectrum
// (mmh/recipes12.f90:100-127).  Pass A handles k1 = 0 with the Nyquist plane speq (needs the
// previous speq, sold); pass B the pairs (i1, n1/2 - i1), 1 <= i1 < n1/4.
template <class T>
__global__ void k_rlft_k0(C2<T>* __restrict__ z, C2<T>* __restrict__ sq, const C2<T>* __restrict__ sold, int h1, int n2, int n3,
                          T c2i) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)n2 * n3) return;
  const int i2 = (int)(q % n2), i3 = (int)(q / n2);
  const int j2 = i2 == 0 ? 0 : n2 - i2, j3 = i3 == 0 ? 0 : n3 - i3;
  C2<T> X = z[(i64)h1 * ((i64)i2 + (i64)n2 * i3)], Y = sold[j2 + (i64)n2 * j3];
  C2<T> s{X.r + Y.r, X.i - Y.i}, d{X.r - Y.r, X.i + Y.i};
  C2<T> H1{T(0.5) * s.r, T(0.5) * s.i};
  C2<T> H2{-c2i * d.i, c2i * d.r};
  z[(i64)h1 * ((i64)i2 + (i64)n2 * i3)] = C2<T>{H1.r + H2.r, H1.i + H2.i};
  sq[j2 + (i64)n2 * j3] = C2<T>{H1.r - H2.r, -(H1.i - H2.i)};
}

template <class T>
__global__ void k_rlft_pairs(C2<T>* __restrict__ z, const C2<T>* __restrict__ wtab, int h1, int n2, int n3, T c2i) {
  // one thread per (i1, i2, i3) with 1 <= i1 < h1 - i1
  const int npair = (h1 - 1) / 2;  // number of i1 values with i1 < h1 - i1, i1 >= 1
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)npair * n2 * n3) return;
  const int i1 = 1 + (int)(q % npair);
  const i64 r = q / npair;
  const int i2 = (int)(r % n2), i3 = (int)(r / n2);
  const int j1 = h1 - i1, j2 = i2 == 0 ? 0 : n2 - i2, j3 = i3 == 0 ? 0 : n3 - i3;
  const T wr = wtab[i1].r, wi = wtab[i1].i;
  C2<T>& Xr = z[(i64)i1 + (i64)h1 * ((i64)i2 + (i64)n2 * i3)];
  C2<T>& Yr = z[(i64)j1 + (i64)h1 * ((i64)j2 + (i64)n2 * j3)];
  C2<T> X = Xr, Y = Yr;
  C2<T> s{X.r + Y.r, X.i - Y.i}, d{X.r - Y.r, X.i + Y.i};
  C2<T> H1{T(0.5) * s.r, T(0.5) * s.i};
  C2<T> H2{-c2i * d.i, c2i * d.r};
  C2<T> WH{wr * H2.r - wi * H2.i, wr * H2.i + wi * H2.r};
  Xr = C2<T>{H1.r + WH.r, H1.i + WH.i};
  Yr = C2<T>{H1.r - WH.r, -(H1.i - WH.i)};
}

// Spectral operations on the rlft3 layout: data (n x n x n reals = complex h x n x n) + speq (n x n).
// mode 0: divide by the 7-point symbol 2cos kx + 2cos ky + 2cos kz - 6 (mmh/limiter.f90:560-575)
// from host cosine tables; mode 1: multiply by |k|^2 (mmh/reconstruct.f90:14-36);
// mode 2/3/4: multiply by i k_x, i k_y, i k_z (mmh/gradient.f90:16-26, 34-48, 56-70).
template <class T>
__global__ void k_spectral(T* __restrict__ data, T* __restrict__ speq, int n, int mode, const T* __restrict__ cosx,
                           const T* __restrict__ cosy, const T* __restrict__ cosz, T pi) {
  const int h = n / 2;
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (i64)(h + 1) * n * n) return;
  const int ic = (int)(q % (h + 1));
  const i64 r = q / (h + 1);
  const int j = (int)(r % n), k = (int)(r / n);
  T* re = ic < h ? &data[ix3(2 * ic, j, k, n)] : &speq[2 * ((i64)j + (i64)n * k)];
  if (mode == 0) {
    T ak = T(2) * cosx[ic] + T(2) * cosy[j] + T(2) * cosz[k] - T(6);
    if (fabs(ak) < T(1e-10)) ak = 1;
    re[0] = re[0] / ak;
    re[1] = re[1] / ak;
  } else if (mode == 1) {
    T k3 = (k + 1 <= h + 1) ? T(k) : T(k - n);
    k3 = k3 / T(n) * T(2.) * pi;
    T k2 = (j + 1 <= h + 1) ? T(j) : T(j - n);
    k2 = k2 / T(n) * T(2.) * pi;
    T k1 = T(ic);
    k1 = k1 / T(n) * T(2.) * pi;
    T f = k1 * k1 + k2 * k2 + k3 * k3;
    re[0] = re[0] * f;
    re[1] = re[1] * f;
  } else {
    T km;
    if (mode == 2) km = T(ic);
    else if (mode == 3) km = (j + 1 <= h + 1) ? T(j) : T(j - n);
    else km = (k + 1 <= h + 1) ? T(k) : T(k - n);
    T A = re[0], Bv = re[1];
    T cr = A * T(0) - Bv * km, ci = A * km + Bv * T(0);
    cr = cr / T(n) * T(2.) * pi;
    ci = ci / T(n) * T(2.) * pi;
    re[0] = cr;
    re[1] = ci;
  }
}

template <class T> __global__ void k_scale2(T* __restrict__ out, const T* __restrict__ in, int n, i64 n3) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) out[p] = in[p] * T(2.) / T(n) / T(n) / T(n);
}

// ------------------------------------------------------------------------------------------------
// relaxing-TVD (mmh/relaxing.f90), whole-volume two-stage form.
// cflkeuler (1067-1139): freezing speed c.
template <class T>
__global__ __launch_bounds__(256) void k_cfl(T* __restrict__ c, const T* __restrict__ B, const T* __restrict__ defp,
                                             const T* __restrict__ u, int n) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
  const T gamma = T(5.) / T(3.);
  T pdx = (defp[ix3(ip, j, k, n)] - defp[ix3(im, j, k, n)]) / T(2);
  T pdy = (defp[ix3(i, jp, k, n)] - defp[ix3(i, jm, k, n)]) / T(2);
  T pdz = (defp[ix3(i, j, kp, n)] - defp[ix3(i, j, km, n)]) / T(2);
  T b11 = B[0 * n3 + p], b12 = B[1 * n3 + p], b13 = B[2 * n3 + p], b22 = B[3 * n3 + p], b23 = B[4 * n3 + p], b33 = B[5 * n3 + p];
  T det = B[6 * n3 + p];
  const T* ul = u + 5 * p;
  T vx = ul[1] / ul[0], vy = ul[2] / ul[0], vz = ul[3] / ul[0];
  T rho = ul[0] / det;
  T kin = rho * (vx * vx + vy * vy + vz * vz) / T(2);
  T ed = ul[4] / det - kin;
  T pressure = (gamma - T(1)) * ed;
  T cs = sqrt(gamma * fabs(pressure / rho));
  T v1 = vx - pdx, v2 = vy - pdy, v3 = vz - pdz;
  T vv1 = b11 * v1 + b12 * v2 + b13 * v3;
  T vv2 = b12 * v1 + b22 * v2 + b23 * v3;
  T vv3 = b13 * v1 + b23 * v2 + b33 * v3;
  T blim1 = cs * sqrt(b11 * b11 + b12 * b12 + b13 * b13) + fabs(vv1);
  T blim2 = cs * sqrt(b12 * b12 + b22 * b22 + b23 * b23) + fabs(vv2);
  T blim3 = cs * sqrt(b13 * b13 + b23 * b23 + b33 * b33) + fabs(vv3);
  c[p] = dmax(dmax(blim1, blim2), blim3);
}

// flux (910-996) with fluxeuler (999-1063): V[(d*5 + l)*n3 + p] = (vt_d / c) for component l;
// stored already divided by c, which is how dpv consumes it (326, 400, 444).
template <class T>
__global__ __launch_bounds__(256) void k_tvd_flux(T* __restrict__ V, const T* __restrict__ u, const T* __restrict__ defp,
                                                  const T* __restrict__ B, const T* __restrict__ c, int n) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const T gamma = T(5.) / T(3.);
  const T rg = B[6 * n3 + p];
  const T* ul = u + 5 * p;
  T rho = ul[0] / rg, mx = ul[1] / rg, my = ul[2] / rg, mz = ul[3] / rg;
  T vx = mx / rho, vy = my / rho, vz = mz / rho;
  T engy = ul[4] / rg;
  T kinetic = rho * (vx * vx + vy * vy + vz * vz) / T(2);
  T pressure = (gamma - T(1)) * (engy - kinetic);
  pressure = dmax(T(0.), pressure);
  T econv = engy + pressure;
  const T v1[3][5] = {{mx, mx * vx + pressure, my * vx, mz * vx, econv * vx},
                      {my, mx * vy, my * vy + pressure, mz * vy, econv * vy},
                      {mz, mx * vz, my * vz, mz * vz + pressure, econv * vz}};
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), kp = wrapi(k + 1, n), km = wrapi(k - 1, n);
  T pdxg = (defp[ix3(ip, j, k, n)] - defp[ix3(im, j, k, n)]) / T(2) / rg;
  T pdyg = (defp[ix3(i, jp, k, n)] - defp[ix3(i, jm, k, n)]) / T(2) / rg;
  T pdzg = (defp[ix3(i, j, kp, n)] - defp[ix3(i, j, km, n)]) / T(2) / rg;
  T b11 = B[0 * n3 + p] * rg, b12 = B[1 * n3 + p] * rg, b13 = B[2 * n3 + p] * rg;
  T b22 = B[3 * n3 + p] * rg, b23 = B[4 * n3 + p] * rg, b33 = B[5 * n3 + p] * rg;
  const T cp = c[p];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    T ulv = ul[l];
    T vtt1 = v1[0][l] - ulv * pdxg;
    T vtt2 = v1[1][l] - ulv * pdyg;
    T vtt3 = v1[2][l] - ulv * pdzg;
    V[(0 * 5 + l) * n3 + p] = (b11 * vtt1 + b12 * vtt2 + b13 * vtt3) / cp;
    V[(1 * 5 + l) * n3 + p] = (b12 * vtt1 + b22 * vtt2 + b23 * vtt3) / cp;
    V[(2 * 5 + l) * n3 + p] = (b13 * vtt1 + b23 * vtt2 + b33 * vtt3) / cp;
  }
}

template <class T> __device__ __forceinline__ T minmod_tvd(T a, T b) {
  // minmod (wphalfx 514-516) with wslopelim = 1.5, wbee = 1 (mmh/relaxgl.f90:8)
  const T wslopelim = T(1.5), wbee = T(1);
  T aa = fabs(a), ab = fabs(b);
  T x = dmin(aa, ab), y = dmax(aa, ab);
  T mm2 = dmin(wslopelim * x, wbee * y + (T(1) - wbee) * x);
  T sa = signbit(a) ? T(-1) : T(1), sb = signbit(b) ? T(-1) : T(1);
  return (sa + sb) * mm2 / T(2);
}

template <class T> __device__ __forceinline__ T wphalf_tvd(T um, T vm, T u0, T v0, T up, T vp, T upp, T vpp) {
  T a = (u0 + v0) / T(2);
  T b = (up - vp) / T(2);
  T da = (up + vp) / T(2) - a;
  T dam = a - (um + vm) / T(2);
  T ap = a + minmod_tvd(dam, da) / T(2);
  T db = (upp - vpp) / T(2) - b;
  T dbm = b - (u0 - v0) / T(2);
  T bp = b - minmod_tvd(db, dbm) / T(2);
  return ap - bp;
}

// dpv (269-496) + stepu/stepuu (774-837): unew = ubase - res(u, V, c)*dt.  unew may alias ubase
// (the full step updates u in place); each thread reads and writes only its own point of them.
template <class T>
__global__ __launch_bounds__(256) void k_tvd_step(T* unew, const T* ubase, const T* __restrict__ u,
                                                  const T* __restrict__ V, const T* __restrict__ c, int n, T dt) {
  const i64 n3 = (i64)n * n * n;
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  const int ip = wrapi(i + 1, n), im = wrapi(i - 1, n), ipp = wrapi(i + 2, n), imm = wrapi(i - 2, n);
  const int jp = wrapi(j + 1, n), jm = wrapi(j - 1, n), jpp = wrapi(j + 2, n), jmm = wrapi(j - 2, n);
  const int kp = wrapi(k + 1, n), km = wrapi(k - 1, n), kpp = wrapi(k + 2, n), kmm = wrapi(k - 2, n);
  const T ck = c[p], ckp = c[ix3(i, j, kp, n)], ckm = c[ix3(i, j, km, n)];
  const T cip = c[ix3(ip, j, k, n)], cim = c[ix3(im, j, k, n)];
  const T cjp = c[ix3(i, jp, k, n)], cjm = c[ix3(i, jm, k, n)];
  for (int l = 0; l < 5; ++l) {
    const T* vx = V + (0 * 5 + l) * n3;
    const T* vy = V + (1 * 5 + l) * n3;
    const T* vz = V + (2 * 5 + l) * n3;
#define UU(a, b, cc) u[5 * ix3(a, b, cc, n) + l]
    T flz1 = wphalf_tvd(UU(i, j, kmm), vz[ix3(i, j, kmm, n)], UU(i, j, km), vz[ix3(i, j, km, n)], UU(i, j, k), vz[p],
                        UU(i, j, kp), vz[ix3(i, j, kp, n)]);
    T flz2 = wphalf_tvd(UU(i, j, km), vz[ix3(i, j, km, n)], UU(i, j, k), vz[p], UU(i, j, kp), vz[ix3(i, j, kp, n)],
                        UU(i, j, kpp), vz[ix3(i, j, kpp, n)]);
    T r = (ckp + ck) / T(2) * flz2 - (ck + ckm) / T(2) * flz1;
    T flxi = wphalf_tvd(UU(im, j, k), vx[ix3(im, j, k, n)], UU(i, j, k), vx[p], UU(ip, j, k), vx[ix3(ip, j, k, n)],
                        UU(ipp, j, k), vx[ix3(ipp, j, k, n)]);
    T flxim = wphalf_tvd(UU(imm, j, k), vx[ix3(imm, j, k, n)], UU(im, j, k), vx[ix3(im, j, k, n)], UU(i, j, k), vx[p],
                         UU(ip, j, k), vx[ix3(ip, j, k, n)]);
    T fp = (cip + ck) / T(2) * flxi;
    T fm = (ck + cim) / T(2) * flxim;
    r = r + fp - fm;
    T flyj = wphalf_tvd(UU(i, jm, k), vy[ix3(i, jm, k, n)], UU(i, j, k), vy[p], UU(i, jp, k), vy[ix3(i, jp, k, n)],
                        UU(i, jpp, k), vy[ix3(i, jpp, k, n)]);
    T flyjm = wphalf_tvd(UU(i, jmm, k), vy[ix3(i, jmm, k, n)], UU(i, jm, k), vy[ix3(i, jm, k, n)], UU(i, j, k), vy[p],
                         UU(i, jp, k), vy[ix3(i, jp, k, n)]);
    r = r + (cjp + ck) / T(2) * flyj - (ck + cjm) / T(2) * flyjm;
#undef UU
    unew[5 * p + l] = ubase[5 * p + l] - r * dt;
  }
}

}  // namespace mmh
