// Device bodies of the Poisson passes (P1-P5), shared by their single-channel and ensemble
// kernels (tc_poisson.hip) and by the persistent ensemble kernel (tc_step.hip).
#pragma once
#include "tc_fft.h"
#include "tc_internal.h"

namespace tc {

// ---------------------------------------------------------------------------------------------
// elements each thread loads before storing any, in the staging loops of P1, P2/P4 and P5
#define FFT_U 4

// LDS of the x and y FFT blocks: `rows` rows of fft_ld<T>(n) entries, the n twiddles, one
// spare entry.
template <typename T>
__host__ __device__ __forceinline__ size_t fft_lds(int rows, int n) {
  return sizeof(cplx<T>) * ((size_t)rows * fft_ld<T>(n) + n + 1);
}

// P1: real rows along x -> half spectra.  Block = (y row j, chunk of KZ = 2*pairs z-rows).
// from_vel: rhs = div(q)/dt with div of utils.py:235-240 on the post-CN velocities (whose
// ghost columns are not stored: x-face 0 is the image of face nx, y-face 0 of face ny, and the
// w walls are zero).  Otherwise rhs is read from a (nx, ny, nz_in) array (planes of nz_in).
// ---------------------------------------------------------------------------------------------
template <typename T, bool FROM_VEL, int ELEMS = TC_FFT_ELEMS>
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
  const int ldb = fft_ld<T>(nx);
  const int lkz = 31 - __clz(KZ);   // KZ is a power of two (checked by the launchers)
  cplx<T>* twl = Z + pairs * ldb;
  fft_stage_twiddles<T>(twl, twx, nx);
  T rdt = T(1);
  if (FROM_VEL) rdt = (T)(1.0 / scal[TC_S_DT]);
  // FFT_U elements per thread are loaded before any is stored: indices are clamped into range
  // (rows past nz_in read row nz_in - 1 and are zeroed, slots past the end store to the spare
  // entry) so that no load sits behind a branch
  const int tot = nx * KZ;
  T* zr = reinterpret_cast<T*>(Z);
  const int spare = 2 * (pairs * ldb + nx);
  for (int t0 = threadIdx.x; t0 < tot; t0 += 256 * FFT_U) {
    T val[FFT_U];
#pragma unroll
    for (int q = 0; q < FFT_U; ++q) {
      const int t = min(t0 + 256 * q, tot - 1);
      const int kk = t & (KZ - 1), i = t >> lkz;
      const int k = min(k0 + kk, nz_in - 1);
      if constexpr (FROM_VEL) {
        const int iul = (i == 0) ? nx : i;
        const int jvl = (j == 0) ? ny : j;
        const T ur = u[(i + 1) * d.sxu + (j + 1) * d.syu + k + 1];
        const T ul = u[iul * d.sxu + (j + 1) * d.syu + k + 1];
        const T vr = v[(i + 1) * d.sxv + (j + 1) * d.syv + k + 1];
        const T vl = v[(i + 1) * d.sxv + jvl * d.syv + k + 1];
        const int cw = (i + 1) * d.sxw + (j + 1) * d.syw;
        const T wt = w[cw + k + 1], wb = w[cw + k];
        const T wtv = (k + 1 < d.nz) ? wt : T(0), wbv = (k > 0) ? wb : T(0);
        val[q] = ((ur - ul) * mt.rdx + (vr - vl) * mt.rdy + (wtv - wbv) * mt.rdzf[k]) * rdt;
      } else {
        val[q] = rin[((size_t)i * ny + j) * nz_in + k];
      }
    }
#pragma unroll
    for (int q = 0; q < FFT_U; ++q) {
      const int t = t0 + 256 * q;
      const int kk = t & (KZ - 1), i = t >> lkz;
      const T x = (k0 + kk < nz_in) ? val[q] : T(0);
      zr[t < tot ? 2 * ((kk >> 1) * ldb + fpad<T>(i)) + (kk & 1) : spare] = x;
    }
  }
  __syncthreads();
  lds_fft<T, ELEMS>(Z, ldb, pairs, fx, twl, -1);
  const int nkx = nx / 2 + 1;
  for (int t = threadIdx.x; t < nkx * KZ; t += 256) {
    const int kk = t & (KZ - 1), m = t >> lkz;
    const int k = k0 + kk;
    if (k >= nz_in) continue;
    const cplx<T> z1 = Z[(kk >> 1) * ldb + fpad<T>(m)];
    const cplx<T> z2 = Z[(kk >> 1) * ldb + fpad<T>(m == 0 ? 0 : nx - m)];
    cplx<T> o;
    if ((kk & 1) == 0) o = {T(0.5) * (z1.re + z2.re), T(0.5) * (z1.im - z2.im)};
    else o = {T(0.5) * (z1.im + z2.im), T(-0.5) * (z1.re - z2.re)};
    S[((size_t)m * ny + j) * nz_in + k] = o;
  }
}

template <typename T, bool FROM_VEL, int ELEMS>
__global__ __launch_bounds__(256) void k_fwd_x(Dims d, Metric<T> mt, const T* __restrict__ u,
                                               const T* __restrict__ v, const T* __restrict__ w,
                                               const T* __restrict__ rin, int nz_in,
                                               const double* __restrict__ scal, FftDesc fx,
                                               const double2* __restrict__ twx, cplx<T>* __restrict__ S,
                                               int pairs) {
  fwd_x_body<T, FROM_VEL, ELEMS>(d, mt, u, v, w, rin, nz_in, scal, fx, twx, S, pairs,
                                 xcd_remap(blockIdx.x, gridDim.x));
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
  const int ldb = fft_ld<T>(ny);
  const int lkz = 31 - __clz(KZ);   // KZ is a power of two (checked by the launchers)
  cplx<T>* twl = Y + KZ * ldb;
  fft_stage_twiddles<T>(twl, twy, ny);
  cplx<T>* Sm = S + (size_t)m * ny * nz_in + k0;
  const int nk = min(KZ, nz_in - k0);
  if (fy.npass >= 2) {
    // first pass straight from the spectrum, last pass straight back to it
    fft_through<T, TC_FFT_ELEMS>(
        Y, ldb, lkz, fy, twl,
        [&](int b, int x) { return b < nk ? Sm[(size_t)x * nz_in + b] : cplx<T>{T(0), T(0)}; },
        [&](int b, int x, const cplx<T>& val) {
          if (b < nk) Sm[(size_t)x * nz_in + b] = val;
        },
        sign);
    return;
  }
  // loads ahead of stores, as in fwd_x_body
  const int tot = ny * KZ;
  const int spare = KZ * ldb + ny;
  for (int t0 = threadIdx.x; t0 < tot; t0 += 256 * FFT_U) {
    cplx<T> val[FFT_U];
#pragma unroll
    for (int q = 0; q < FFT_U; ++q) {
      const int t = min(t0 + 256 * q, tot - 1);
      const int kk = t & (KZ - 1), jj = t >> lkz;
      val[q] = S[((size_t)m * ny + jj) * nz_in + min(k0 + kk, nz_in - 1)];
    }
#pragma unroll
    for (int q = 0; q < FFT_U; ++q) {
      const int t = t0 + 256 * q;
      const int kk = t & (KZ - 1), jj = t >> lkz;
      Y[t < tot ? kk * ldb + fpad<T>(jj) : spare] = (k0 + kk < nz_in) ? val[q] : cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  lds_fft<T>(Y, ldb, KZ, fy, twl, sign);
  for (int t = threadIdx.x; t < ny * KZ; t += 256) {
    const int kk = t & (KZ - 1), j = t >> lkz;
    const int k = k0 + kk;
    if (k < nz_in) S[((size_t)m * ny + j) * nz_in + k] = Y[kk * ldb + fpad<T>(j)];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fft_y(int nkx, int ny, int nz_in, FftDesc fy,
                                               const double2* __restrict__ twy,
                                               cplx<T>* __restrict__ S, int KZ, int sign) {
  fft_y_body<T>(nkx, ny, nz_in, fy, twy, S, KZ, sign, xcd_remap(blockIdx.x, gridDim.x));
}

// Input slots t0 + 256 q (q < FFT_U) of the transform of block task (j, k0): bins ka = k0 + 2 qq
// and ka + 1 of spectrum column (m, j), slot t = m * pairs + qq.  Slots past the end and rows
// past nz are clamped into range (the staging zeroes them), so no load sits behind a branch.
template <typename T>
__device__ __forceinline__ void inv_x_fetch(const cplx<T>* __restrict__ S, int ny, int nz, int j, int k0,
                                            int lpr, int pairs, int tot, int t0, cplx<T> (&va)[FFT_U],
                                            cplx<T> (&vb)[FFT_U]) {
#pragma unroll
  for (int q = 0; q < FFT_U; ++q) {
    const int t = min(t0 + 256 * q, tot - 1);
    const int m = t >> lpr, ka = k0 + 2 * (t & (pairs - 1));
    const cplx<T>* col = S + ((size_t)m * ny + j) * nz;
    if constexpr (sizeof(T) == 8) {
      // the spectrum's last read of the step: non-temporal in float64 (measured: step -1 %;
      // float32 keeps plain loads)
      typedef T v2 __attribute__((ext_vector_type(2)));
      const v2 a2 = __builtin_nontemporal_load(reinterpret_cast<const v2*>(col + min(ka, nz - 1)));
      const v2 b2 = __builtin_nontemporal_load(reinterpret_cast<const v2*>(col + min(ka + 1, nz - 1)));
      va[q] = {a2.x, a2.y};
      vb[q] = {b2.x, b2.y};
    } else {
      va[q] = col[min(ka, nz - 1)];
      vb[q] = col[min(ka + 1, nz - 1)];
    }
  }
}

// Slots of inv_x_fetch into the block's LDS sequences: pair qq, input x = A + iB at bin m and its
// conjugate-symmetric image at nx - m (rows ka >= nz are zero, the imaginary parts of bins 0 and
// nx/2 are dropped); slots past the end go to the spare entry.
template <typename T>
__device__ __forceinline__ void inv_x_stage(cplx<T>* Z, int ldb, int nx, int nz, int k0, int lpr, int pairs,
                                            int tot, int spare, int t0, const cplx<T> (&va)[FFT_U],
                                            const cplx<T> (&vb)[FFT_U]) {
  const int h = nx / 2;
#pragma unroll
  for (int q = 0; q < FFT_U; ++q) {
    const int t = t0 + 256 * q;
    const int qq = t & (pairs - 1), m = t >> lpr;
    const int ka = k0 + 2 * qq, kb = ka + 1;
    cplx<T> A = ka < nz ? va[q] : cplx<T>{T(0), T(0)};
    cplx<T> B = kb < nz ? vb[q] : cplx<T>{T(0), T(0)};
    if (m == 0 || m == h) { A.im = T(0); B.im = T(0); }
    const bool in = t < tot;
    Z[in ? qq * ldb + fpad<T>(m) : spare] = {A.re - B.im, A.im + B.re};
    if (m > 0 && m < h) Z[in ? qq * ldb + fpad<T>(nx - m) : spare] = {A.re + B.im, B.re - A.im};
  }
}

// Rows of the transformed sequences to p, two z-rows per thread: pair qq of x-row i holds row
// ka = k0 + 2 qq in its real part and row ka + 1 in its imaginary part, written together as one
// two-element vector (element-aligned: row ka + 1 sits at an odd offset in the column), so the
// pairs of consecutive lanes cover one x-row's KZ rows.  Each axis has the cell's index and at
// most one image (x: face nx -> 0, face 1 -> nx + 1; y likewise; z: Neumann copies 1 -> 0,
// nz -> nz + 1), so a value goes to at most 8 places.
template <typename T>
__device__ __forceinline__ void inv_x_store(const Dims& d, const cplx<T>* Z, int ldb, T* __restrict__ p,
                                            T scale, int j, int k0, int lpr) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  typedef v2 v2e __attribute__((aligned(sizeof(T))));
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int pairs = 1 << lpr;
  const int J = j + 1;
  const int Y1 = J == ny ? 0 : (J == 1 ? ny + 1 : -1);
  const int sxp = d.sxp, syp = d.syp;
  for (int t = threadIdx.x; t < nx * pairs; t += 256) {
    const int qq = t & (pairs - 1), i = t >> lpr;
    const int ka = k0 + 2 * qq;
    if (ka >= nz) continue;
    const cplx<T> z = Z[qq * ldb + fpad<T>(i)];
    const T a = z.re * scale, b = z.im * scale;
    const bool two = ka + 1 < nz;
    const int I = i + 1, K = ka + 1;
    const int X1 = I == nx ? 0 : (I == 1 ? nx + 1 : -1);
    // the column (xi, yj): rows K and K + 1, or row K alone at an odd nz's last row
    auto put = [&](int xi, int yj) {
      T* q = p + xi * sxp + yj * syp + K;
      if (two) *reinterpret_cast<v2e*>(q) = v2{a, b};
      else q[0] = a;
    };
    put(I, J);
    if (X1 >= 0) put(X1, J);
    if (Y1 >= 0) {
      put(I, Y1);
      if (X1 >= 0) put(X1, Y1);
    }
    // Neumann wall images: row 1 -> 0 (from the pair's first row), row nz -> nz + 1 (from its
    // second row, or its first where the pair is the odd last row)
    const bool bot = K == 1;
    const bool top = two ? K + 1 == nz : K == nz;
    if (bot || top) {
      const int zk = bot ? 0 : nz + 1;
      const T zv = bot ? a : (two ? b : a);
      p[I * sxp + J * syp + zk] = zv;
      if (X1 >= 0) p[X1 * sxp + J * syp + zk] = zv;
      if (Y1 >= 0) {
        p[I * sxp + Y1 * syp + zk] = zv;
        if (X1 >= 0) p[X1 * sxp + Y1 * syp + zk] = zv;
      }
      if (bot && two && K + 1 == nz) {   // nz = 2: the pair's second row is also the top row
        p[I * sxp + J * syp + nz + 1] = b;
        if (X1 >= 0) p[X1 * sxp + J * syp + nz + 1] = b;
        if (Y1 >= 0) {
          p[I * sxp + Y1 * syp + nz + 1] = b;
          if (X1 >= 0) p[X1 * sxp + Y1 * syp + nz + 1] = b;
        }
      }
    }
  }
}

// P5: half spectra -> real rows along x (the x half of projection_fft.py:139's irfft2: the
// imaginary parts of the kx = 0 and Nyquist bins are dropped), scaled by 1/(nx*ny), written to p
// with the ghost images of projection_fft.py:146-163 (x copies, y copies, Neumann z walls).  Block
// task = (y row j, KZ = 2*pairs z-rows).  Two real z-rows ka, ka + 1 travel as one complex
// sequence (pair qq): input x of the pair is A + iB at bin m = x for x <= nx/2 and its
// conjugate-symmetric image at bin m = nx - x above, with A, B the rows' bins; output x holds row
// ka in its real and row ka + 1 in its imaginary part.
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
  const int ldb = fft_ld<T>(nx);
  const int lpr = 31 - __clz(pairs);   // pairs is a power of two (checked by the launchers)
  cplx<T>* twl = Z + pairs * ldb;
  // loads ahead of stores, as in fwd_x_body; the twiddles are staged while the first slots' bins
  // are in flight
  const int tot = (nx / 2 + 1) * pairs;
  const int spare = pairs * ldb + nx;
  cplx<T> va[FFT_U], vb[FFT_U];
  inv_x_fetch<T>(S, ny, nz, j, k0, lpr, pairs, tot, threadIdx.x, va, vb);
  fft_stage_twiddles<T>(twl, twx, nx);
  for (int t0 = threadIdx.x; t0 < tot; t0 += 256 * FFT_U) {
    if (t0 != (int)threadIdx.x) inv_x_fetch<T>(S, ny, nz, j, k0, lpr, pairs, tot, t0, va, vb);
    inv_x_stage<T>(Z, ldb, nx, nz, k0, lpr, pairs, tot, spare, t0, va, vb);
  }
  __syncthreads();
  lds_fft<T>(Z, ldb, pairs, fx, twl, +1);
  inv_x_store<T>(d, Z, ldb, p, scale, j, k0, lpr);
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

// P3: one wall-normal solve per (kx, ky) system (projection_fft.py:184-215): forward
// y_k = (d_k - a_k y_{k-1}) / pivot_k, backward x_k = y_k - c'_k x_{k+1}, with
// pivot_k = b_k - a_k c'_{k-1} and c'_k = c_k / pivot_k (projection_fft.py:56-82, 184-200).
// The factorisation is the one sequential recurrence of the solve that does not involve the
// data: each lane recomputes it in float64 over its own R rows from the c' of the row before
// its first row, which the plan stores per (system, lane) (factorised on the host in extended
// precision), so the per-system table is L values instead of nz.  The two data recurrences are
// first-order linear, evaluated here as scans of affine maps y -> A y + B, whose
// composition is associative: a system is a group of L = 16 lanes (one DPP row; L = 32, two rows,
// above 128 rows), lane g holds rows g*R .. g*R+R-1, composes its R maps in registers, and the L
// lane maps are combined by a Hillis-Steele scan over row_shr (forward) / row_shl (backward)
// moves, with one more stage across the two rows of a 32-lane group (row_bcast:15 forward,
// lanes 16 / 48 through scalar registers backward; readlane only reads, every lane of the
// wavefront executes it).  Rows L*R > nz are identity maps.  The products of |a_k / pivot_k| <= 1
// and |c'_k| <= 1 (diagonally dominant rows) keep every partial map bounded, so the scan carries
// the sequential sweep's round-off.  The singular (0,0) system is gauge-fixed (row 0 pinned to
// zero, row 1 decoupled from it) and solved in float64.
#define TH_THREADS 256

template <int CTRL, typename U>
__device__ __forceinline__ void th_combine(U& a, cplx<U>& b) {
  const U a2 = dpp_row<CTRL>(U(1), a);
  const U br = dpp_row<CTRL>(U(0), b.re), bi = dpp_row<CTRL>(U(0), b.im);
  b = {a * br + b.re, a * bi + b.im};
  a = a * a2;
}

// RM consecutive elements from / to a 16-byte aligned address, as 16-byte vector accesses
template <typename E, int RM>
__device__ __forceinline__ void vload_rows(const E* __restrict__ p, E (&x)[RM]) {
  constexpr int PER = 16 / (int)sizeof(E);
  static_assert(PER >= 1 && RM % PER == 0, "whole 16-byte vectors");
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int v = 0; v < RM / PER; ++v) {
    const uint4 w = q[v];
    __builtin_memcpy(&x[v * PER], &w, 16);
  }
}
template <typename E, int RM>
__device__ __forceinline__ void vstore_rows(E* __restrict__ p, const E (&x)[RM]) {
  constexpr int PER = 16 / (int)sizeof(E);
  static_assert(PER >= 1 && RM % PER == 0, "whole 16-byte vectors");
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int v = 0; v < RM / PER; ++v) {
    uint4 w;
    __builtin_memcpy(&w, &x[v * PER], 16);
    q[v] = w;
  }
}

// VEC: nz == L * RM (R == RM, every lane holds RM rows of the system) and every lane's first row
// sits on a 16-byte boundary of each array, so a lane's rows move as 16-byte vectors (from `lrow`,
// the lane's first row in the spectrum or in an LDS copy of it) instead of RM strided element
// accesses (from `col`, the system's row 0).
// Rows (projection_fft.py:44-82, as the plan builds them on the host): a_k = cl[k] (0 for k = 0),
// b_k = -((cl[k] + cr[k]) + K2) with the Neumann corrections +cl[0] at k = 0 and +cr[nz-1] at
// k = nz - 1, c_k = cr[k] (0 for the last row).  The gauge-fixed (0,0) system (gauge) pins row 0
// (b = 1, c = 0, and y_0 = 0 whatever d_0) and decouples row 1 from it (a_1 = 0).  cb: c' of the
// row before the lane's first row (0 for lane 0).
template <typename U, typename T, int RM, bool VEC>
__device__ __forceinline__ void thomas_rows(cplx<T>* __restrict__ col, cplx<T>* __restrict__ lrow,
                                            const double* __restrict__ cl, const double* __restrict__ cr,
                                            double k2, double cb, int nz, int R, int L, int g, bool gauge,
                                            bool store) {
  const int wl = lane_id();
  const int m0 = g * R;
  U A[RM], C[RM];
  cplx<U> B[RM];
  cplx<T> dv[RM];
  // row m (mm = m clamped into the system) -> the forward map y -> A y + B and the backward
  // factor C; rows past nz are identity maps
  double cprev = cb;
  auto row = [&](int r, cplx<T> d, double clv, double crv) {
    const int m = m0 + r, mm = min(m, nz - 1);
    const double a = (mm == 0 || (gauge && mm == 1)) ? 0.0 : clv;
    double b = -((clv + crv) + k2);
    if (mm == 0) b += clv;
    if (mm == nz - 1) b += crv;
    double c = mm <= nz - 2 ? crv : 0.0;
    if (gauge && mm == 0) {
      b = 1.0;
      c = 0.0;
    }
    const double rpiv = 1.0 / (b - a * cprev);
    const double cpk = c * rpiv;
    cprev = cpk;
    const double be = (gauge && mm == 0) ? 0.0 : rpiv;
    const bool in = m < nz;
    C[r] = (U)cpk;
    A[r] = in ? (U)(-a * be) : U(1);
    const U bu = (U)be;
    B[r] = in ? cplx<U>{bu * (U)d.re, bu * (U)d.im} : cplx<U>{U(0), U(0)};
  };
  if constexpr (VEC) {
    double clv[RM], crv[RM];
    vload_rows<cplx<T>, RM>(lrow, dv);
    vload_rows<double, RM>(cl + m0, clv);
    vload_rows<double, RM>(cr + m0, crv);
#pragma unroll
    for (int r = 0; r < RM; ++r) row(r, dv[r], clv[r], crv[r]);
  } else {
    // every load is issued up front and unconditionally (rows past nz read row nz - 1)
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      if (r < R) {
        const int mm = min(m0 + r, nz - 1);
        row(r, col[mm], cl[mm], cr[mm]);
      }
    }
  }
  // forward: lane-local prefix maps, then the scan of the 16 lane maps
  U pa = U(1);
  cplx<U> pb = {U(0), U(0)};
#pragma unroll
  for (int r = 0; r < RM; ++r) {
    if (r < R) {
      pb = {A[r] * pb.re + B[r].re, A[r] * pb.im + B[r].im};
      pa = A[r] * pa;
      A[r] = pa;
      B[r] = pb;
    }
  }
  th_combine<TC_DPP_ROW_SHR(1)>(pa, pb);
  th_combine<TC_DPP_ROW_SHR(2)>(pa, pb);
  th_combine<TC_DPP_ROW_SHR(4)>(pa, pb);
  th_combine<TC_DPP_ROW_SHR(8)>(pa, pb);
  if (L == 32) {
    // rows 1 and 3 take the prefix of rows 0 and 2 (their last lanes); the A part is not needed
    const U br = dpp_row<TC_DPP_ROW_BCAST15, 0xa>(U(0), pb.re), bi = dpp_row<TC_DPP_ROW_BCAST15, 0xa>(U(0), pb.im);
    pb = {pa * br + pb.re, pa * bi + pb.im};
  }
  U ein_r = dpp_row<TC_DPP_ROW_SHR(1)>(U(0), pb.re), ein_i = dpp_row<TC_DPP_ROW_SHR(1)>(U(0), pb.im);
  if (L == 32) {
    // the first lane of the group's second row continues from the last lane of its first row
    const U r15 = readlane(pb.re, 15), i15 = readlane(pb.im, 15), r47 = readlane(pb.re, 47), i47 = readlane(pb.im, 47);
    if ((wl & 31) == 16) {
      ein_r = wl < 32 ? r15 : r47;
      ein_i = wl < 32 ? i15 : i47;
    }
  }
  // y rows; backward maps x -> -c'_k x + y_k (rows past nz: x = 0)
#pragma unroll
  for (int r = 0; r < RM; ++r) {
    if (r < R) {
      const int m = m0 + r, mm = min(m, nz - 1);
      const bool in = m < nz;
      const cplx<U> y = {A[r] * ein_r + B[r].re, A[r] * ein_i + B[r].im};
      A[r] = in ? -C[r] : U(0);
      B[r] = in ? y : cplx<U>{U(0), U(0)};
    }
  }
  U qa = U(1);
  cplx<U> qb = {U(0), U(0)};
#pragma unroll
  for (int r = RM - 1; r >= 0; --r) {
    if (r < R) {
      qb = {A[r] * qb.re + B[r].re, A[r] * qb.im + B[r].im};
      qa = A[r] * qa;
      A[r] = qa;
      B[r] = qb;
    }
  }
  th_combine<TC_DPP_ROW_SHL(1)>(qa, qb);
  th_combine<TC_DPP_ROW_SHL(2)>(qa, qb);
  th_combine<TC_DPP_ROW_SHL(4)>(qa, qb);
  th_combine<TC_DPP_ROW_SHL(8)>(qa, qb);
  U sr = U(0), si = U(0);
  if (L == 32) {
    // rows 0 and 2 continue into the suffix of rows 1 and 3 (held by lanes 16 / 48)
    const U r16 = readlane(qb.re, 16), i16 = readlane(qb.im, 16), r48 = readlane(qb.re, 48), i48 = readlane(qb.im, 48);
    sr = wl < 32 ? r16 : r48;
    si = wl < 32 ? i16 : i48;
    if ((wl & 16) == 0) qb = {qa * sr + qb.re, qa * si + qb.im};
  }
  U xin_r = dpp_row<TC_DPP_ROW_SHL(1)>(U(0), qb.re), xin_i = dpp_row<TC_DPP_ROW_SHL(1)>(U(0), qb.im);
  if (L == 32 && (wl & 31) == 15) {
    xin_r = sr;
    xin_i = si;
  }
  if (!store) return;
  if constexpr (VEC) {
#pragma unroll
    for (int r = 0; r < RM; ++r) dv[r] = {(T)(A[r] * xin_r + B[r].re), (T)(A[r] * xin_i + B[r].im)};
    vstore_rows<cplx<T>, RM>(lrow, dv);
  } else {
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      if (r < R) {
        const int m = m0 + r;
        if (m < nz) col[m] = {(T)(A[r] * xin_r + B[r].re), (T)(A[r] * xin_i + B[r].im)};
      }
    }
  }
}

// lanes per system (16, or 32 above 128 rows) and rows per lane of a system of nz rows
__host__ __device__ __forceinline__ int thomas_lanes(int nz) { return nz <= 128 ? 16 : 32; }
__host__ __device__ __forceinline__ int thomas_rows_per_lane(int nz) {
  const int L = thomas_lanes(nz);
  const int r = (nz + L - 1) / L;
  return r <= 1 ? 1 : r <= 2 ? 2 : r <= 4 ? 4 : 8;
}

// LDS segment of one lane in the staged form: its RM rows plus 16 bytes, so that the 16-byte row
// reads and writes of 8 consecutive lanes fall in distinct banks.
template <typename T, int RM>
__host__ __device__ constexpr int thomas_seg() { return RM * (int)sizeof(cplx<T>) + 16; }

// Block = sb systems of L lanes each; R <= RM rows per lane.  cl, cr: the rows' couplings
// (float64, [nz]); K2: kx~^2 + ky~^2 per system; cb: [system][L] boundary c'.
// STG (with VEC): the block's sb systems, one contiguous stretch of the spectrum, are copied
// into LDS and back with 16-byte accesses of consecutive lanes, so that each wavefront-wide load
// or store covers 1 KiB of consecutive memory; the lanes then read their rows from LDS.
template <typename T, int RM, bool VEC = false, bool STG = false>
__device__ __forceinline__ void thomas_body(int nsys, int nz, int R, int L, int sb, const double* __restrict__ cl,
                                            const double* __restrict__ cr, const double* __restrict__ K2,
                                            const double* __restrict__ cb, cplx<T>* __restrict__ S, int blk) {
  static_assert(!STG || VEC, "the staged form moves whole 16-byte row vectors");
  const int lane = threadIdx.x & (L - 1);
  const int s = blk * sb + (int)(threadIdx.x / L);
  const bool live = s < nsys;
  const int sc = live ? s : nsys - 1;
  cplx<T>* col = S + (size_t)sc * nz;
  const double k2 = K2[sc];
  const double c0 = cb[(size_t)sc * L + lane];
  const bool gauge = sc == 0;
  if constexpr (STG) {
    constexpr int E = (int)sizeof(cplx<T>);
    constexpr int PER = 16 / E;   // rows per 16-byte unit
    constexpr int SEG = thomas_seg<T, RM>();
    constexpr int UPT = RM / PER;   // units per thread: the block's sb * L * RM rows / (sb * L) threads
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg = smem_raw;   // sb * L * SEG bytes
    const int nth = sb * L;
    const int nlive = min(sb, nsys - blk * sb);
    const int nunits = nlive * nz / PER;
    const int lnz = 31 - __clz(nz);   // nz = L * RM, a power of two
    constexpr int lrm = RM == 1 ? 0 : RM == 2 ? 1 : RM == 4 ? 2 : 3;
    uint4* gsrc = reinterpret_cast<uint4*>(S + (size_t)blk * sb * nz);
    // unit u holds rows e = u*PER .. of system e >> lnz, row k = e & (nz - 1): lane k >> lrm of
    // that system, position k & (RM - 1) in its segment
    auto lds_unit = [&](int u) -> uint4* {
      const int e = u * PER;
      const int sl = e >> lnz, k = e & (nz - 1);
      return reinterpret_cast<uint4*>(stg + (sl * L + (k >> lrm)) * SEG + (k & (RM - 1)) * E);
    };
    uint4 w[UPT];
#pragma unroll
    for (int q = 0; q < UPT; ++q) w[q] = gsrc[min((int)threadIdx.x + q * nth, nunits - 1)];
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int u = threadIdx.x + q * nth;
      if (u < nunits) *lds_unit(u) = w[q];
    }
    __syncthreads();
    cplx<T>* lrow = reinterpret_cast<cplx<T>*>(stg + threadIdx.x * SEG);
    if (sizeof(T) == sizeof(double) || !gauge)
      thomas_rows<T, T, RM, true>(col, lrow, cl, cr, k2, c0, nz, R, L, lane, gauge, live);
    else
      thomas_rows<double, T, RM, true>(col, lrow, cl, cr, k2, c0, nz, R, L, lane, true, live);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int u = threadIdx.x + q * nth;
      if (u < nunits) gsrc[u] = *lds_unit(u);
    }
  } else {
    cplx<T>* lrow = col + lane * R;
    // the gauge-fixed system runs in float64 in both precisions
    if (sizeof(T) == sizeof(double) || !gauge)
      thomas_rows<T, T, RM, VEC>(col, lrow, cl, cr, k2, c0, nz, R, L, lane, gauge, live);
    else
      thomas_rows<double, T, RM, VEC>(col, lrow, cl, cr, k2, c0, nz, R, L, lane, true, live);
  }
}

template <typename T, int RM, bool VEC, bool STG>
__global__ __launch_bounds__(256) void k_poisson_thomas(int nsys, int nz, int sb, const double* __restrict__ cl,
                                                        const double* __restrict__ cr,
                                                        const double* __restrict__ K2,
                                                        const double* __restrict__ cb,
                                                        cplx<T>* __restrict__ S) {
  thomas_body<T, RM, VEC, STG>(nsys, nz, RM, thomas_lanes(nz), sb, cl, cr, K2, cb, S,
                               xcd_remap(blockIdx.x, gridDim.x));
}

}  // namespace tc
