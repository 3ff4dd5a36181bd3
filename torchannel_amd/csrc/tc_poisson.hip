// FFT + wall-normal tridiagonal pressure Poisson solve (projection_fft.py:3-165).
//
// Spectral layout: S[m][j][k] complex, m = kx in [0, nx/2] (real-to-half transform along x),
// j = ky index in [0, ny) (complex transform along y), k in [0, nz) contiguous.
//   P1 k_fwd_x : div (or div/dt computed from the velocities) -> x half spectra, two z-rows per
//                complex transform (real/imag packing)
//   P2 k_fft_y : forward complex FFT along y
//   P3 k_poisson_pcr : one wavefront per (kx, ky) system, PCR along z, (0,0) gauge-fixed in f64
//   P4 k_fft_y : inverse complex FFT along y
//   P5 k_inv_x : half spectra -> real rows, 1/(nx*ny), p written with every ghost image
#include "tc_ensemble.h"
#include "tc_fft.h"
#include "tc_internal.h"
#include "tc_poisson.h"

namespace tc {

// ---------------------------------------------------------------------------------------------
// P1: real rows along x -> half spectra.  Block = (y row j, chunk of KZ = 2*pairs z-rows).
// from_vel: rhs = div(q)/dt with div of utils.py:235-240 on the post-CN velocities (whose
// ghost columns are not stored: x-face 0 is the image of face nx, y-face 0 of face ny, and the
// w walls are zero).  Otherwise rhs is read from a (nx, ny, nz_in) array (planes of nz_in).
// ---------------------------------------------------------------------------------------------
template <typename T, bool FROM_VEL>
__device__ __forceinline__ void fwd_x_body(Dims d, Metric<T> mt, const T* __restrict__ u,
                                               const T* __restrict__ v, const T* __restrict__ w,
                                               const T* __restrict__ rin, int nz_in,
                                               const double* __restrict__ scal, const FftDesc& fx,
                                               const double2* __restrict__ twx, cplx<T>* __restrict__ S,
                                               int pairs,
    int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = d.nx, ny = d.ny;
  const int KZ = 2 * pairs;
  const int nch = (nz_in + KZ - 1) / KZ;
  const int j = blk / nch, k0 = (blk - j * nch) * KZ;
  const int ldb = nx + 1;
  const int lkz = 31 - __clz(KZ);   // KZ is a power of two (checked by the launchers)
  cplx<T>* twl = Z + pairs * ldb;
  fft_stage_twiddles<T>(twl, twx, nx);
  T rdt = T(1);
  if (FROM_VEL) rdt = (T)(1.0 / scal[TC_S_DT]);
  for (int t = threadIdx.x; t < nx * KZ; t += 256) {
    const int kk = t & (KZ - 1), i = t >> lkz;
    const int k = k0 + kk;
    T val = T(0);
    if (k < nz_in) {
      if constexpr (FROM_VEL) {
        const int iul = (i == 0) ? nx : i;
        const int jvl = (j == 0) ? ny : j;
        const T ur = u[(i + 1) * d.sxu + (j + 1) * d.syu + k + 1];
        const T ul = u[iul * d.sxu + (j + 1) * d.syu + k + 1];
        const T vr = v[(i + 1) * d.sxv + (j + 1) * d.syv + k + 1];
        const T vl = v[(i + 1) * d.sxv + jvl * d.syv + k + 1];
        const int cw = (i + 1) * d.sxw + (j + 1) * d.syw;
        const T wt = (k + 1 < d.nz) ? w[cw + k + 1] : T(0);
        const T wb = (k > 0) ? w[cw + k] : T(0);
        val = ((ur - ul) * mt.rdx + (vr - vl) * mt.rdy + (wt - wb) * mt.rdzf[k]) * rdt;
      } else {
        val = rin[((size_t)i * ny + j) * nz_in + k];
      }
    }
    T* zf = reinterpret_cast<T*>(&Z[(kk >> 1) * ldb + i]);
    zf[kk & 1] = val;
  }
  __syncthreads();
  lds_fft<T>(Z, ldb, pairs, fx, twl, -1);
  const int nkx = nx / 2 + 1;
  for (int t = threadIdx.x; t < nkx * KZ; t += 256) {
    const int kk = t & (KZ - 1), m = t >> lkz;
    const int k = k0 + kk;
    if (k >= nz_in) continue;
    const cplx<T> z1 = Z[(kk >> 1) * ldb + m];
    const cplx<T> z2 = Z[(kk >> 1) * ldb + (m == 0 ? 0 : nx - m)];
    cplx<T> o;
    if ((kk & 1) == 0) o = {T(0.5) * (z1.re + z2.re), T(0.5) * (z1.im - z2.im)};
    else o = {T(0.5) * (z1.im + z2.im), T(-0.5) * (z1.re - z2.re)};
    S[((size_t)m * ny + j) * nz_in + k] = o;
  }
}

template <typename T, bool FROM_VEL>
__global__ __launch_bounds__(256) void k_fwd_x(Dims d, Metric<T> mt, const T* __restrict__ u,
                                               const T* __restrict__ v, const T* __restrict__ w,
                                               const T* __restrict__ rin, int nz_in,
                                               const double* __restrict__ scal, FftDesc fx,
                                               const double2* __restrict__ twx, cplx<T>* __restrict__ S,
                                               int pairs) {
  fwd_x_body<T, FROM_VEL>(d, mt, u, v, w, rin, nz_in, scal, fx, twx, S, pairs, xcd_remap(blockIdx.x, gridDim.x));
}

// P2 / P4: complex FFT along y for every (kx, z) column.  Block = (kx, chunk of KZ z-rows).
template <typename T>
__device__ __forceinline__ void fft_y_body(int nkx, int ny, int nz_in, const FftDesc& fy,
                                               const double2* __restrict__ twy,
                                               cplx<T>* __restrict__ S, int KZ, int sign,
    int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* Y = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nch = (nz_in + KZ - 1) / KZ;
  const int m = blk / nch, k0 = (blk - m * nch) * KZ;
  const int ldb = ny + 1;
  const int lkz = 31 - __clz(KZ);   // KZ is a power of two (checked by the launchers)
  cplx<T>* twl = Y + KZ * ldb;
  fft_stage_twiddles<T>(twl, twy, ny);
  for (int t = threadIdx.x; t < ny * KZ; t += 256) {
    const int kk = t & (KZ - 1), j = t >> lkz;
    const int k = k0 + kk;
    Y[kk * ldb + j] = k < nz_in ? S[((size_t)m * ny + j) * nz_in + k] : cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T>(Y, ldb, KZ, fy, twl, sign);
  for (int t = threadIdx.x; t < ny * KZ; t += 256) {
    const int kk = t & (KZ - 1), j = t >> lkz;
    const int k = k0 + kk;
    if (k < nz_in) S[((size_t)m * ny + j) * nz_in + k] = Y[kk * ldb + j];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fft_y(int nkx, int ny, int nz_in, FftDesc fy,
                                               const double2* __restrict__ twy,
                                               cplx<T>* __restrict__ S, int KZ, int sign) {
  fft_y_body<T>(nkx, ny, nz_in, fy, twy, S, KZ, sign, xcd_remap(blockIdx.x, gridDim.x));
}

// P5: half spectra -> real rows (irfft semantics: the imaginary parts of the kx = 0 and Nyquist
// bins are dropped), scaled by 1/(nx*ny), written to p with the ghost images of
// projection_fft.py:150-163 (x copies, y copies, Neumann z walls).
template <typename T>
__device__ __forceinline__ void inv_x_body(Dims d, const FftDesc& fx, const double2* __restrict__ twx,
                                               const cplx<T>* __restrict__ S, T* __restrict__ p,
                                               T scale, int pairs,
    int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem_raw);
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int KZ = 2 * pairs;
  const int nch = (nz + KZ - 1) / KZ;
  const int j = blk / nch, k0 = (blk - j * nch) * KZ;
  const int ldb = nx + 1;
  const int h = nx / 2;
  const int lpr = 31 - __clz(pairs);   // pairs is a power of two (checked by the launchers)
  cplx<T>* twl = Z + pairs * ldb;
  fft_stage_twiddles<T>(twl, twx, nx);
  for (int t = threadIdx.x; t < (h + 1) * pairs; t += 256) {
    const int q = t & (pairs - 1), m = t >> lpr;
    const int ka = k0 + 2 * q, kb = ka + 1;
    cplx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
    if (ka < nz) A = S[((size_t)m * ny + j) * nz + ka];
    if (kb < nz) B = S[((size_t)m * ny + j) * nz + kb];
    if (m == 0 || m == h) { A.im = T(0); B.im = T(0); }
    Z[q * ldb + m] = {A.re - B.im, A.im + B.re};
    if (m > 0 && m < h) Z[q * ldb + nx - m] = {A.re + B.im, B.re - A.im};
  }
  __syncthreads();
  lds_fft<T>(Z, ldb, pairs, fx, twl, +1);
  for (int t = threadIdx.x; t < nx * KZ; t += 256) {
    const int kk = t & (KZ - 1), i = t >> (lpr + 1);
    const int k = k0 + kk;
    if (k >= nz) continue;
    const cplx<T> z = Z[(kk >> 1) * ldb + i];
    const T val = ((kk & 1) ? z.im : z.re) * scale;
    const int I = i + 1, J = j + 1, K = k + 1;
    p[I * d.sxp + J * d.syp + K] = val;
    if (I == 1 || I == nx || J == 1 || J == ny || K == 1 || K == nz) {
      // every combination of the cell and its images along each axis, the cell itself excepted
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int X = a == 0 ? I : (a == 1 ? (I == nx ? 0 : -1) : (I == 1 ? nx + 1 : -1));
        if (X < 0) continue;
#pragma unroll
        for (int bq = 0; bq < 3; ++bq) {
          const int Y = bq == 0 ? J : (bq == 1 ? (J == ny ? 0 : -1) : (J == 1 ? ny + 1 : -1));
          if (Y < 0) continue;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int Zc = c == 0 ? K : (c == 1 ? (K == 1 ? 0 : -1) : (K == nz ? nz + 1 : -1));
            if (Zc < 0 || (a | bq | c) == 0) continue;
            p[X * d.sxp + Y * d.syp + Zc] = val;
          }
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_inv_x(Dims d, FftDesc fx, const double2* __restrict__ twx,
                                               const cplx<T>* __restrict__ S, T* __restrict__ p,
                                               T scale, int pairs) {
  inv_x_body<T>(d, fx, twx, S, p, scale, pairs, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------------------------
// General PCR (tridiag.py:40-58) on registers: a, b, c, d rows m = r*64 + lane; rows >= n are
// padding (a = c = d = 0, b = 1), the same values the reference pads out-of-range rows with.
// ---------------------------------------------------------------------------------------------
template <int ST, int R, typename U, bool CPLX>
__device__ __forceinline__ void pcr_stage(U (&a)[R], U (&b)[R], U (&c)[R], U (&dr)[R], U (&di)[R], int n,
                                          int lane) {
  constexpr int s = 1 << ST;
  if (s >= n) return;
  U adn[R], aup[R], bdn[R], bup[R], cdn[R], cup[R], rdn[R], rup[R], idn[R], iup[R];
  shift_dn<s, R, U>(a, adn, U(0), lane);
  shift_up<s, R, U>(a, aup, U(0), lane);
  shift_dn<s, R, U>(b, bdn, U(1), lane);
  shift_up<s, R, U>(b, bup, U(1), lane);
  shift_dn<s, R, U>(c, cdn, U(0), lane);
  shift_up<s, R, U>(c, cup, U(0), lane);
  shift_dn<s, R, U>(dr, rdn, U(0), lane);
  shift_up<s, R, U>(dr, rup, U(0), lane);
  if constexpr (CPLX) {
    shift_dn<s, R, U>(di, idn, U(0), lane);
    shift_up<s, R, U>(di, iup, U(0), lane);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const U al = -a[r] / bdn[r];
    const U ga = -c[r] / bup[r];
    b[r] = b[r] + al * cdn[r] + ga * aup[r];
    dr[r] = dr[r] + al * rdn[r] + ga * rup[r];
    if constexpr (CPLX) di[r] = di[r] + al * idn[r] + ga * iup[r];
    a[r] = al * adn[r];
    c[r] = ga * cup[r];
  }
}

template <int R, typename U, bool CPLX>
__device__ __forceinline__ void pcr_general(U (&a)[R], U (&b)[R], U (&c)[R], U (&dr)[R], U (&di)[R], int n,
                                            int lane) {
  pcr_stage<0, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<1, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<2, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<3, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<4, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<5, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<6, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<7, R, U, CPLX>(a, b, c, dr, di, n, lane);
  pcr_stage<8, R, U, CPLX>(a, b, c, dr, di, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    dr[r] = dr[r] / b[r];
    if constexpr (CPLX) di[r] = di[r] / b[r];
  }
}

// P3: one wall-normal Thomas solve per (kx, ky) system (projection_fft.py:184-215), the
// reference's algorithm: PCR loses ~4 digits on the ill-conditioned low modes of this operator,
// Thomas does not.  The factorisation depends only on the grid and (kx, ky), so c'_k and 1/pivot_k
// are computed once per plan (extended precision on the host) and stored [k][system]; per step
// each system is one forward and one backward sweep.  A block stages `sb` consecutive systems
// (one contiguous chunk of S) in LDS; lane l of wave 0 sweeps system s0 + l.  The singular
// (0,0) system is gauge-fixed (first unknown pinned to zero) and swept in float64.
template <typename T>
__device__ __forceinline__ void thomas_body(int nsys, int nz, int sb, const double* __restrict__ cl,
                                                        const T* __restrict__ cp, const T* __restrict__ iden,
                                                        const double* __restrict__ cp0,
                                                        const double* __restrict__ iden0,
                                                        cplx<T>* __restrict__ S,
    int blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* col = reinterpret_cast<cplx<T>*>(smem_raw);
  const int s0 = blk * sb;
  const int ns = min(sb, nsys - s0);
  const int ld = nz + 1;
  const cplx<T>* src = S + (size_t)s0 * nz;
  for (int e = threadIdx.x; e < ns * nz; e += 256) {
    const int sl = e / nz, k = e - sl * nz;
    col[sl * ld + k] = src[e];
  }
  __syncthreads();
  if ((int)threadIdx.x < ns) {
    const int sl = threadIdx.x, s = s0 + sl;
    cplx<T>* c = col + sl * ld;
    if (s == 0) {
      // gauge-fixed (0,0) mode in float64 (forward values kept in a float64 LDS row): row 0 is
      // x0 = 0, row 1 is decoupled from it
      double* fw = reinterpret_cast<double*>(col + sb * ld);
      double pr = 0.0, pi = 0.0;
      fw[0] = 0.0; fw[nz] = 0.0;
      for (int k = 1; k < nz; ++k) {
        const double a = (k == 1) ? 0.0 : cl[k];
        const double id = iden0[k];
        const cplx<T> d = c[k];
        pr = ((double)d.re - a * pr) * id;
        pi = ((double)d.im - a * pi) * id;
        fw[k] = pr; fw[nz + k] = pi;
      }
      double xr = pr, xi = pi;
      c[nz - 1] = {(T)xr, (T)xi};
      for (int k = nz - 2; k >= 1; --k) {
        const double q = cp0[k];
        xr = fw[k] - q * xr;
        xi = fw[nz + k] - q * xi;
        c[k] = {(T)xr, (T)xi};
      }
      c[0] = {T(0), T(0)};
    } else {
      // the sweeps walk k with the table loads issued a chunk of 8 rows ahead of their use
      constexpr int CH = 8;
      cplx<T> dp = {T(0), T(0)};
      T idb[CH], idn[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) idb[q] = q < nz ? iden[(size_t)q * nsys + s] : T(0);
      for (int k0 = 0; k0 < nz; k0 += CH) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int kn = k0 + CH + q;
          idn[q] = kn < nz ? iden[(size_t)kn * nsys + s] : T(0);
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int k = k0 + q;
          if (k < nz) {
            const T a = k > 0 ? (T)cl[k] : T(0);
            const cplx<T> d = c[k];
            dp = {(d.re - a * dp.re) * idb[q], (d.im - a * dp.im) * idb[q]};
            c[k] = dp;
          }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) idb[q] = idn[q];
      }
      cplx<T> x = dp;
      T cpb[CH], cpn[CH];
      const int top = nz - 2;
#pragma unroll
      for (int q = 0; q < CH; ++q) cpb[q] = top - q >= 0 ? cp[(size_t)(top - q) * nsys + s] : T(0);
      for (int k0 = top; k0 >= 0; k0 -= CH) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int kn = k0 - CH - q;
          cpn[q] = kn >= 0 ? cp[(size_t)kn * nsys + s] : T(0);
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int k = k0 - q;
          if (k >= 0) {
            const cplx<T> d = c[k];
            x = {d.re - cpb[q] * x.re, d.im - cpb[q] * x.im};
            c[k] = x;
          }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) cpb[q] = cpn[q];
      }
    }
  }
  __syncthreads();
  cplx<T>* dst = S + (size_t)s0 * nz;
  for (int e = threadIdx.x; e < ns * nz; e += 256) {
    const int sl = e / nz, k = e - sl * nz;
    dst[e] = col[sl * ld + k];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_poisson_thomas(int nsys, int nz, int sb, const double* __restrict__ cl,
                                                        const T* __restrict__ cp, const T* __restrict__ iden,
                                                        const double* __restrict__ cp0,
                                                        const double* __restrict__ iden0,
                                                        cplx<T>* __restrict__ S) {
  thomas_body<T>(nsys, nz, sb, cl, cp, iden, cp0, iden0, S, xcd_remap(blockIdx.x, gridDim.x));
}

// Ensemble forms of P1-P5: global block -> (channel, block of that channel) via prefix sums.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_x_ens(const EnsChan<T>* __restrict__ ch, const int* __restrict__ start,
                                                   int n) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[q];   // fields are passed to the body by value
  fwd_x_body<T, true>(c.d, c.mt, c.su, c.sv, c.sw, nullptr, c.d.nz, c.scal, c.fx, c.twx, c.spec, c.pairs,
                      gb - start[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fft_y_ens(const EnsChan<T>* __restrict__ ch, const int* __restrict__ start,
                                                   int n, int sign) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[q];   // fields are passed to the body by value
  fft_y_body<T>(c.nkx, c.d.ny, c.d.nz, c.fy, c.twy, c.spec, c.kzy, sign, gb - start[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_poisson_thomas_ens(const EnsChan<T>* __restrict__ ch,
                                                            const int* __restrict__ start, int n) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[q];   // fields are passed to the body by value
  thomas_body<T>(c.nkx * c.d.ny, c.d.nz, c.sb, c.cl, c.th_cp, c.th_iden, c.th_cp0, c.th_iden0, c.spec,
                 gb - start[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_inv_x_ens(const EnsChan<T>* __restrict__ ch, const int* __restrict__ start,
                                                   int n) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[q];   // fields are passed to the body by value
  const T scale = (T)(1.0 / ((double)c.d.nx * (double)c.d.ny));
  inv_x_body<T>(c.d, c.fx, c.twx, c.spec, c.p, scale, c.pairs, gb - start[q]);
}


// Generic batched PCR for the stand-alone tridiagonal entry points.
template <typename T, int R, bool CPLX>
__global__ __launch_bounds__(256) void k_pcr_general(int n, int batch, const T* __restrict__ A,
                                                     const T* __restrict__ B, const T* __restrict__ C,
                                                     long long abc_stride, const T* __restrict__ D,
                                                     T* __restrict__ X) {
  const int sys = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (sys >= batch) return;
  const int lane = lane_id();
  const T* a0 = A + abc_stride * sys;
  const T* b0 = B + abc_stride * sys;
  const T* c0 = C + abc_stride * sys;
  T a[R], b[R], c[R], dr[R], di[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) {
      a[r] = m > 0 ? a0[m] : T(0);
      b[r] = b0[m];
      c[r] = m < n - 1 ? c0[m] : T(0);
      if (CPLX) {
        dr[r] = D[2 * ((size_t)sys * n + m)];
        di[r] = D[2 * ((size_t)sys * n + m) + 1];
      } else {
        dr[r] = D[(size_t)sys * n + m];
        di[r] = T(0);
      }
    } else {
      a[r] = T(0); b[r] = T(1); c[r] = T(0); dr[r] = T(0); di[r] = T(0);
    }
  }
  pcr_general<R, T, CPLX>(a, b, c, dr, di, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) {
      if (CPLX) {
        X[2 * ((size_t)sys * n + m)] = dr[r];
        X[2 * ((size_t)sys * n + m) + 1] = di[r];
      } else {
        X[(size_t)sys * n + m] = dr[r];
      }
    }
  }
}

// Batched Thomas sweep (projection_fft.py:184-215) for the stand-alone solve_tridiagonal: one
// thread per system, rows of a system contiguous; c' goes to the caller's workspace.
template <typename T, bool CPLX>
__global__ __launch_bounds__(256) void k_thomas_general(int n, int batch, const T* __restrict__ A,
                                                        const T* __restrict__ B, const T* __restrict__ C,
                                                        long long abc_stride, const T* __restrict__ D,
                                                        T* __restrict__ X, T* __restrict__ work) {
  const int sys = blockIdx.x * 256 + threadIdx.x;
  if (sys >= batch) return;
  const T* a = A + abc_stride * sys;
  const T* b = B + abc_stride * sys;
  const T* c = C + abc_stride * sys;
  T* cp = work + (size_t)sys * n;
  const size_t base = (size_t)sys * n;
  T pr, pi = T(0);
  cp[0] = c[0] / b[0];
  if (CPLX) { pr = D[2 * base] / b[0]; pi = D[2 * base + 1] / b[0]; X[2 * base] = pr; X[2 * base + 1] = pi; }
  else { pr = D[base] / b[0]; X[base] = pr; }
  for (int m = 1; m < n; ++m) {
    const T den = b[m] - a[m] * cp[m - 1];
    if (m < n - 1) cp[m] = c[m] / den;
    if (CPLX) {
      pr = (D[2 * (base + m)] - a[m] * pr) / den;
      pi = (D[2 * (base + m) + 1] - a[m] * pi) / den;
      X[2 * (base + m)] = pr; X[2 * (base + m) + 1] = pi;
    } else {
      pr = (D[base + m] - a[m] * pr) / den;
      X[base + m] = pr;
    }
  }
  for (int m = n - 2; m >= 0; --m) {
    if (CPLX) {
      pr = X[2 * (base + m)] - cp[m] * pr;
      pi = X[2 * (base + m) + 1] - cp[m] * pi;
      X[2 * (base + m)] = pr; X[2 * (base + m) + 1] = pi;
    } else {
      pr = X[base + m] - cp[m] * pr;
      X[base + m] = pr;
    }
  }
}

int thomas_batched(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                   long long s, const void* d, void* x, void* work, hipStream_t st) {
  const int blocks = (batch + 255) / 256;
#define TC_TH(TT, CC)                                                                            \
  hipLaunchKernelGGL((k_thomas_general<TT, CC>), dim3(blocks), dim3(256), 0, st, n, batch,        \
                     (const TT*)a, (const TT*)b, (const TT*)c, s, (const TT*)d, (TT*)x, (TT*)work)
  if (dtype == TC_F64) { if (cplx_d) TC_TH(double, true); else TC_TH(double, false); }
  else { if (cplx_d) TC_TH(float, true); else TC_TH(float, false); }
#undef TC_TH
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
static bool factor_fft(int n, FftDesc& f) {
  f.n = n;
  f.npass = 0;
  int m = n;
  while (m % 4 == 0 && f.npass < TC_FFT_MAXPASS) { f.radix[f.npass++] = 4; m /= 4; }
  while (m % 2 == 0 && f.npass < TC_FFT_MAXPASS) { f.radix[f.npass++] = 2; m /= 2; }
  while (m % 3 == 0 && f.npass < TC_FFT_MAXPASS) { f.radix[f.npass++] = 3; m /= 3; }
  while (m % 5 == 0 && f.npass < TC_FFT_MAXPASS) { f.radix[f.npass++] = 5; m /= 5; }
  int pp = 1;
  for (int q = 0; q < f.npass; ++q) {
    f.mag_m[q] = fft_magic((unsigned)(n / f.radix[q]));
    f.mag_p[q] = fft_magic((unsigned)pp);
    pp *= f.radix[q];
  }
  return m == 1;
}

static int pow2_floor(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

int poisson_init(Poisson* pl, const Grid* g, const double* dzf_host, const double* dzc_host) {
  pl->g = g;
  const Dims& d = g->d;
  pl->nkx = d.nx / 2 + 1;
  if (!factor_fft(d.nx, pl->fx) || !factor_fft(d.ny, pl->fy)) return TC_ERR_SHAPE;
  if (d.nx > TC_FFT_ELEMS || d.ny > TC_FFT_ELEMS) return TC_ERR_SHAPE;
  pl->pairs = std::min(8, pow2_floor(std::max(1, TC_FFT_ELEMS / d.nx)));
  pl->pairs = std::min(pl->pairs, pow2_floor(std::max(1, (d.nz + 1) / 2)));
  pl->kzy = std::min(16, pow2_floor(std::max(1, TC_FFT_ELEMS / d.ny)));
  pl->kzy = std::min(pl->kzy, pow2_floor(d.nz));
  // twiddles exp(-2 pi i t / N) in float64
  auto make_tw = [](int n, double2** out) -> int {
    std::vector<double2> h(n);
    for (int t = 0; t < n; ++t) {
      const double ang = -2.0 * M_PI * (double)t / (double)n;
      h[t].x = std::cos(ang);
      h[t].y = std::sin(ang);
    }
    if (hipMalloc(out, sizeof(double2) * n) != hipSuccess) return TC_ERR_NOMEM;
    if (hipMemcpy(*out, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess) return TC_ERR_NOMEM;
    return TC_OK;
  };
  int rc;
  if ((rc = make_tw(d.nx, &pl->twx)) != TC_OK) return rc;
  if ((rc = make_tw(d.ny, &pl->twy)) != TC_OK) return rc;
  // wall-normal rows (projection_fft.py:44-49) and modified wavenumbers (26-39)
  const int nz = d.nz;
  std::vector<double> cl(nz), cr(nz);
  for (int k = 0; k < nz; ++k) {
    cl[k] = 1.0 / (dzc_host[k] * dzf_host[k]);
    cr[k] = 1.0 / (dzc_host[k + 1] * dzf_host[k]);
  }
  const int nkx = pl->nkx, ny = d.ny, nx = d.nx;
  std::vector<double> k2((size_t)nkx * ny);
  const double two_pi = 2.0 * M_PI;
  for (int m = 0; m < nkx; ++m) {
    const double kx = (double)m / ((double)nx * (g->dx / two_pi));
    const double kxm = (2.0 / g->dx) * std::sin(kx * g->dx / 2.0);
    for (int j = 0; j < ny; ++j) {
      const int jj = j <= ny / 2 ? j : ny - j;
      const double ky = (double)jj / ((double)ny * (g->dy / two_pi));
      const double kym = (2.0 / g->dy) * std::sin(ky * g->dy / 2.0);
      k2[(size_t)m * ny + j] = kxm * kxm + kym * kym;
    }
  }
  if (hipMalloc(&pl->cl, sizeof(double) * nz) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->cr, sizeof(double) * nz) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->K2, sizeof(double) * k2.size()) != hipSuccess) return TC_ERR_NOMEM;
  hipMemcpy(pl->cl, cl.data(), sizeof(double) * nz, hipMemcpyHostToDevice);
  hipMemcpy(pl->cr, cr.data(), sizeof(double) * nz, hipMemcpyHostToDevice);
  hipMemcpy(pl->K2, k2.data(), sizeof(double) * k2.size(), hipMemcpyHostToDevice);
  // Thomas factorisation of every (kx, ky) system of projection_fft.py:56-82 (rows built in
  // float64 exactly as the reference builds tri_a/b/c, factorised in long double), [k][system]
  const int nsys = nkx * ny;
  std::vector<double> cp64((size_t)nz * nsys), id64((size_t)nz * nsys);
  std::vector<double> cp0(nz), id0(nz);
  for (int sy = 0; sy < nsys; ++sy) {
    const bool gauge = sy == 0;
    const double K2 = k2[sy];
    long double cprev = 0.0L;
    for (int k = 0; k < nz; ++k) {
      const double a = (k >= 1 && !(gauge && k == 1)) ? cl[k] : 0.0;
      double b = -((cl[k] + cr[k]) + K2);
      if (k == 0) b += cl[0];
      if (k == nz - 1) b += cr[nz - 1];
      double c = k <= nz - 2 ? cr[k] : 0.0;
      if (gauge && k == 0) { b = 1.0; c = 0.0; }
      const long double den = (long double)b - (long double)a * cprev;
      const long double cpk = (long double)c / den;
      cp64[(size_t)k * nsys + sy] = (double)cpk;
      id64[(size_t)k * nsys + sy] = (double)(1.0L / den);
      if (gauge) { cp0[k] = (double)cpk; id0[k] = (double)(1.0L / den); }
      cprev = cpk;
    }
  }
  const size_t esz = g->dtype == TC_F64 ? 16 : 8;
  const size_t rsz = esz / 2;
  if (hipMalloc(&pl->th_cp, rsz * cp64.size()) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->th_iden, rsz * id64.size()) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->th_cp0, sizeof(double) * nz) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->th_iden0, sizeof(double) * nz) != hipSuccess) return TC_ERR_NOMEM;
  if (g->dtype == TC_F64) {
    hipMemcpy(pl->th_cp, cp64.data(), 8 * cp64.size(), hipMemcpyHostToDevice);
    hipMemcpy(pl->th_iden, id64.data(), 8 * id64.size(), hipMemcpyHostToDevice);
  } else {
    std::vector<float> f1(cp64.begin(), cp64.end()), f2(id64.begin(), id64.end());
    hipMemcpy(pl->th_cp, f1.data(), 4 * f1.size(), hipMemcpyHostToDevice);
    hipMemcpy(pl->th_iden, f2.data(), 4 * f2.size(), hipMemcpyHostToDevice);
  }
  hipMemcpy(pl->th_cp0, cp0.data(), sizeof(double) * nz, hipMemcpyHostToDevice);
  hipMemcpy(pl->th_iden0, id0.data(), sizeof(double) * nz, hipMemcpyHostToDevice);
  pl->sb = std::max(1, std::min(64, (int)(67584 / ((size_t)(nz + 1) * esz))));
  if (hipMalloc(&pl->spec, esz * (size_t)nkx * ny * nz) != hipSuccess) return TC_ERR_NOMEM;
  // statistics workspace: 6 planes (u, v, w at the two sampled heights), float64
  if (hipMalloc(&pl->planes, sizeof(double) * (size_t)nx * ny * 6) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->plane_spec, 16 * (size_t)nkx * ny * 6) != hipSuccess) return TC_ERR_NOMEM;
  return TC_OK;
}

void poisson_free(Poisson* pl) {
  hipFree(pl->twx); hipFree(pl->twy); hipFree(pl->cl); hipFree(pl->cr); hipFree(pl->K2);
  hipFree(pl->spec); hipFree(pl->planes); hipFree(pl->plane_spec);
  hipFree(pl->th_cp); hipFree(pl->th_iden); hipFree(pl->th_cp0); hipFree(pl->th_iden0);
}

template <typename T>
int launch_fwd_x_t(const Poisson* pl, bool from_vel, const void* u, const void* v, const void* w,
                   const void* rin, int nz_in, const double* scal, void* spec, int pairs, hipStream_t st) {
  const Grid* g = pl->g;
  const Dims& d = g->d;
  const int KZ = 2 * pairs;
  if (pairs < 1 || (pairs & (pairs - 1)) != 0) return TC_ERR_ARG;
  const int blocks = d.ny * ((nz_in + KZ - 1) / KZ);
  const size_t lds = sizeof(cplx<T>) * ((size_t)pairs * (d.nx + 1) + d.nx);
  const Metric<T> mt = g->metric<T>();
  if (from_vel)
    hipLaunchKernelGGL((k_fwd_x<T, true>), dim3(blocks), dim3(256), lds, st, d, mt, (const T*)u,
                       (const T*)v, (const T*)w, (const T*)rin, nz_in, scal, pl->fx, pl->twx,
                       (cplx<T>*)spec, pairs);
  else
    hipLaunchKernelGGL((k_fwd_x<T, false>), dim3(blocks), dim3(256), lds, st, d, mt, (const T*)u,
                       (const T*)v, (const T*)w, (const T*)rin, nz_in, scal, pl->fx, pl->twx,
                       (cplx<T>*)spec, pairs);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
int launch_fft_y_t(const Poisson* pl, void* spec, int nz_in, int kz, int sign, hipStream_t st) {
  const Dims& d = pl->g->d;
  if (kz < 1 || (kz & (kz - 1)) != 0) return TC_ERR_ARG;
  const int blocks = pl->nkx * ((nz_in + kz - 1) / kz);
  const size_t lds = sizeof(cplx<T>) * ((size_t)kz * (d.ny + 1) + d.ny);
  hipLaunchKernelGGL((k_fft_y<T>), dim3(blocks), dim3(256), lds, st, pl->nkx, d.ny, nz_in, pl->fy,
                     pl->twy, (cplx<T>*)spec, kz, sign);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// One stage of the Poisson solve: 0 = P1 forward x, 1 = P2 forward y, 2 = P3 PCR, 3 = P4 inverse y,
// 4 = P5 inverse x + ghosts.
template <typename T>
static int poisson_phase_t(const Poisson* pl, int ph, bool from_vel, const void* u, const void* v,
                           const void* w, const void* rin, const double* scal, void* p, hipStream_t st) {
  const Grid* g = pl->g;
  const Dims& d = g->d;
  switch (ph) {
    case 0: return launch_fwd_x_t<T>(pl, from_vel, u, v, w, rin, d.nz, scal, pl->spec, pl->pairs, st);
    case 1: return launch_fft_y_t<T>(pl, pl->spec, d.nz, pl->kzy, -1, st);
    case 2: {
      const int nsys = pl->nkx * d.ny;
      const int sb = pl->g->thomas_sb > 0 ? pl->g->thomas_sb : pl->sb;
      const int blocks = (nsys + sb - 1) / sb;
      const size_t lds = sizeof(cplx<T>) * (size_t)sb * (d.nz + 1) + 2 * sizeof(double) * d.nz;
      if (lds > 160 * 1024) return TC_ERR_ARG;
      hipLaunchKernelGGL((k_poisson_thomas<T>), dim3(blocks), dim3(256), lds, st, nsys, d.nz, sb, pl->cl,
                         (const T*)pl->th_cp, (const T*)pl->th_iden, pl->th_cp0, pl->th_iden0,
                         (cplx<T>*)pl->spec);
      TC_CHECK_LAUNCH();
      return TC_OK;
    }
    case 3: return launch_fft_y_t<T>(pl, pl->spec, d.nz, pl->kzy, +1, st);
    case 4: {
      const int KZ = 2 * pl->pairs;
      const int blocks5 = d.ny * ((d.nz + KZ - 1) / KZ);
      const size_t lds = sizeof(cplx<T>) * ((size_t)pl->pairs * (d.nx + 1) + d.nx);
      const T scale = (T)(1.0 / ((double)d.nx * (double)d.ny));
      hipLaunchKernelGGL((k_inv_x<T>), dim3(blocks5), dim3(256), lds, st, d, pl->fx, pl->twx,
                         (const cplx<T>*)pl->spec, (T*)p, scale, pl->pairs);
      TC_CHECK_LAUNCH();
      return TC_OK;
    }
    default: return TC_ERR_ARG;
  }
}

int poisson_phase(const Poisson* pl, int ph, bool from_vel, const void* u, const void* v, const void* w,
                  const void* rin, const double* scal, void* p, hipStream_t st) {
  if (pl->g->dtype == TC_F64) return poisson_phase_t<double>(pl, ph, from_vel, u, v, w, rin, scal, p, st);
  return poisson_phase_t<float>(pl, ph, from_vel, u, v, w, rin, scal, p, st);
}

int poisson_solve(const Poisson* pl, bool from_vel, const void* u, const void* v, const void* w,
                  const void* rin, const double* scal, void* p, hipStream_t st) {
  for (int ph = 0; ph < 5; ++ph) {
    const int rc = poisson_phase(pl, ph, from_vel, u, v, w, rin, scal, p, st);
    if (rc != TC_OK) return rc;
  }
  return TC_OK;
}

// Plane FFTs for the statistics: planes (nx, ny, 6) float64 -> folded spectra inputs.
int plane_fft(const Poisson* pl, hipStream_t st) {
  int rc;
  const int pairs = std::min(2, pow2_floor(std::max(1, TC_FFT_ELEMS / pl->g->d.nx)));
  if ((rc = launch_fwd_x_t<double>(pl, false, nullptr, nullptr, nullptr, pl->planes, 6, nullptr,
                                   pl->plane_spec, pairs, st)) != TC_OK)
    return rc;
  const int kz = std::min(pow2_floor(std::max(1, TC_FFT_ELEMS / pl->g->d.ny)), 4);
  return launch_fft_y_t<double>(pl, pl->plane_spec, 6, kz, -1, st);
}

template <typename T, int R, bool CPLX>
static void launch_pcr_gen(int n, int batch, const void* a, const void* b, const void* c, long long s,
                           const void* d, void* x, hipStream_t st) {
  hipLaunchKernelGGL((k_pcr_general<T, R, CPLX>), dim3((batch + 3) / 4), dim3(256), 0, st, n, batch,
                     (const T*)a, (const T*)b, (const T*)c, s, (const T*)d, (T*)x);
}

template <typename T, bool CPLX>
static int pcr_dispatch_r(int n, int batch, const void* a, const void* b, const void* c, long long s,
                          const void* d, void* x, hipStream_t st) {
  const int R = (n + TC_WAVE - 1) / TC_WAVE;
  switch (R) {
    case 1: launch_pcr_gen<T, 1, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    case 2: launch_pcr_gen<T, 2, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    case 3: launch_pcr_gen<T, 3, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    case 4: launch_pcr_gen<T, 4, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    case 5: case 6: launch_pcr_gen<T, 6, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    case 7: case 8: launch_pcr_gen<T, 8, CPLX>(n, batch, a, b, c, s, d, x, st); break;
    default: return TC_ERR_UNSUPPORTED;
  }
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int pcr_batched(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                long long s, const void* d, void* x, hipStream_t st) {
  if (dtype == TC_F64)
    return cplx_d ? pcr_dispatch_r<double, true>(n, batch, a, b, c, s, d, x, st)
                  : pcr_dispatch_r<double, false>(n, batch, a, b, c, s, d, x, st);
  return cplx_d ? pcr_dispatch_r<float, true>(n, batch, a, b, c, s, d, x, st)
                : pcr_dispatch_r<float, false>(n, batch, a, b, c, s, d, x, st);
}

}  // namespace tc

namespace tc {

// Geometry of one channel's Poisson launches (block counts and dynamic LDS), shared by the
// single-channel launchers above and the ensemble.
void poisson_geometry(const Poisson* pl, int tsz, int* blocks, size_t* lds) {
  const Dims& d = pl->g->d;
  const size_t csz = 2 * (size_t)tsz;
  const int KZ = 2 * pl->pairs;
  const int sb = pl->g->thomas_sb > 0 ? pl->g->thomas_sb : pl->sb;
  blocks[0] = d.ny * ((d.nz + KZ - 1) / KZ);
  lds[0] = csz * ((size_t)pl->pairs * (d.nx + 1) + d.nx);
  blocks[1] = pl->nkx * ((d.nz + pl->kzy - 1) / pl->kzy);
  lds[1] = csz * ((size_t)pl->kzy * (d.ny + 1) + d.ny);
  blocks[2] = (pl->nkx * d.ny + sb - 1) / sb;
  lds[2] = csz * (size_t)sb * (d.nz + 1) + 2 * sizeof(double) * d.nz;
  blocks[3] = blocks[0];
  lds[3] = lds[0];
}

template <typename T>
static int ens_poisson_t(const void* dch, int n, const int* const* start, const int* blocks, const size_t* lds,
                         hipStream_t st) {
  const EnsChan<T>* c = (const EnsChan<T>*)dch;
  hipLaunchKernelGGL((k_fwd_x_ens<T>), dim3(blocks[0]), dim3(256), lds[0], st, c, start[0], n);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_fft_y_ens<T>), dim3(blocks[1]), dim3(256), lds[1], st, c, start[1], n, -1);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_poisson_thomas_ens<T>), dim3(blocks[2]), dim3(256), lds[2], st, c, start[2], n);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_fft_y_ens<T>), dim3(blocks[1]), dim3(256), lds[1], st, c, start[1], n, +1);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_inv_x_ens<T>), dim3(blocks[3]), dim3(256), lds[3], st, c, start[3], n);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int ens_launch_poisson(int dtype, const void* dch, int n, const int* const* start, const int* blocks,
                       const size_t* lds, hipStream_t st) {
  for (int q = 0; q < 4; ++q)
    if (lds[q] > 160 * 1024) return TC_ERR_ARG;
  if (dtype == TC_F64) return ens_poisson_t<double>(dch, n, start, blocks, lds, st);
  return ens_poisson_t<float>(dch, n, start, blocks, lds, st);
}

}  // namespace tc
