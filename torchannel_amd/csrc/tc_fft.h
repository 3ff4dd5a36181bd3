// LDS-resident batched complex FFT (Stockham autosort, radices 8, 4, 2, 3, 5; N = 2^a 3^b 5^c).
//
// A block of 256 threads transforms `batch` sequences of length N held in LDS (sequence b at
// buf[b*ldb + fpad<T>(0..N))), batch*N <= the pass's element budget (TC_FFT_ELEMS unless stated).  Pass with radix R and span p (p = product of the
// radices already applied): butterfly idx in [0, N/R), k = idx mod p, inputs x[idx + r*N/R],
// twiddles exp(sign*2*pi*i*r*k/(p*R)), outputs at (idx-k)*R + k + r*p.
#pragma once
#include "tc_common.h"

#define TC_FFT_ELEMS 2048
// budget of the forward x transform of long rows (nx > 256): twice the z rows per block, so each
// x row is read in segments twice as long; its passes hold twice the registers
#define TC_FFT_ELEMS_WIDE 4096
// float32 rows of more than 256: twice again, so each x row is read in whole 128-byte lines
#define TC_FFT_ELEMS_WIDE2 8192
#define TC_FFT_MAXPASS 12

// LDS slot of element x of a sequence.  float64 complex (16 bytes): one pad slot after every 8
// elements, so that the stride-R stores of a first pass (outputs idx*R + r) fall into distinct
// banks (measured at 192x192x128: y transforms 8 % and the inverse x transform 5 % faster);
// float32 complex keeps the plain layout, which measured faster than the padded one.  A sequence of n elements spans fft_ld<T>(n) slots
// (the extra slot staggers consecutive sequences).
template <typename T>
__device__ __forceinline__ int fpad(int x) {
  if constexpr (sizeof(T) == 8) return x + (x >> 3);
  else return x;
}
template <typename T>
__host__ __device__ __forceinline__ int fft_ld(int n) {
  if constexpr (sizeof(T) == 8) return n + (n >> 3) + 1;
  else return n + 1;
}

struct FftDesc {
  int n;
  int npass;
  int radix[TC_FFT_MAXPASS];
  // per pass: ceil(2^32 / d) for d = n / radix and for d = p (the product of the radices already
  // applied), so that x / d = mulhi(x, magic) for every x < 2^16 the pass divides
  unsigned mag_m[TC_FFT_MAXPASS];
  unsigned mag_p[TC_FFT_MAXPASS];
};

// x / d for x < 2^16, 2 <= d < 2^16, with magic = ceil(2^32 / d): the rounding error of magic is
// below d / 2^32 per unit of x, so x * magic / 2^32 stays below the next integer.  d = 1 (whose
// magic does not fit 32 bits) is encoded as magic 0.
__device__ __forceinline__ int fdiv(int x, unsigned magic) {
  return magic == 0u ? x : (int)__umulhi((unsigned)x, magic);
}

static inline unsigned fft_magic(unsigned d) { return d <= 1u ? 0u : (unsigned)(((1ull << 32) + d - 1) / d); }

template <typename T, int R>
__device__ __forceinline__ void fft_butterfly(cplx<T> (&x)[R], int sign) {
  if constexpr (R == 2) {
    const cplx<T> a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd inputs, then X[k], X[k+4] = E[k] +- w^k O[k] with
    // w = exp(sign*2*pi*i/8): w = c (1 + s i), w^2 = s i, w^3 = c (-1 + s i), c = 1/sqrt(2)
    cplx<T> e[4] = {x[0], x[2], x[4], x[6]};
    cplx<T> o[4] = {x[1], x[3], x[5], x[7]};
    fft_butterfly<T, 4>(e, sign);
    fft_butterfly<T, 4>(o, sign);
    const T c = T(0.70710678118654752440084436210485);
    const T sg = sign < 0 ? T(-1) : T(1);
    const cplx<T> o1 = {c * (o[1].re - sg * o[1].im), c * (o[1].im + sg * o[1].re)};
    const cplx<T> o2 = {-sg * o[2].im, sg * o[2].re};
    const cplx<T> o3 = {-c * (o[3].re + sg * o[3].im), c * (sg * o[3].re - o[3].im)};
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);
    x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);
    x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);
    x[7] = csub(e[3], o3);
  } else if constexpr (R == 4) {
    const cplx<T> a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    const cplx<T> c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    // s*i*d with s = sign
    const cplx<T> sid = sign < 0 ? cplx<T>{d.im, -d.re} : cplx<T>{-d.im, d.re};
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, sid);
    x[3] = csub(b, sid);
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410229341718282);    // cos(2 pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);   // cos(4 pi/5)
    const T s1 = T(0.95105651629515357211643933337938);    // sin(2 pi/5)
    const T s2 = T(0.58778525229247312916870595463907);    // sin(4 pi/5)
    const cplx<T> t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    const cplx<T> t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
    const cplx<T> a1 = {x[0].re + c1 * t1.re + c2 * t2.re, x[0].im + c1 * t1.im + c2 * t2.im};
    const cplx<T> a2 = {x[0].re + c2 * t1.re + c1 * t2.re, x[0].im + c2 * t1.im + c1 * t2.im};
    const cplx<T> b1 = {s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im};
    const cplx<T> b2 = {s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im};
    // sign*i*b
    const cplx<T> ib1 = sign < 0 ? cplx<T>{b1.im, -b1.re} : cplx<T>{-b1.im, b1.re};
    const cplx<T> ib2 = sign < 0 ? cplx<T>{b2.im, -b2.re} : cplx<T>{-b2.im, b2.re};
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, ib1);
    x[4] = csub(a1, ib1);
    x[2] = cadd(a2, ib2);
    x[3] = csub(a2, ib2);
  } else {  // R == 3
    const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
    const cplx<T> t1 = cadd(x[1], x[2]);
    const cplx<T> t2 = {x[0].re - T(0.5) * t1.re, x[0].im - T(0.5) * t1.im};
    const cplx<T> df = csub(x[1], x[2]);
    const T sh = sign < 0 ? -h : h;
    const cplx<T> rot = {-sh * df.im, sh * df.re};  // i*sign*(sqrt3/2)*(x1-x2)
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, rot);
    x[2] = csub(t2, rot);
  }
}

template <typename T, int R, int ELEMS>
__device__ __forceinline__ void fft_pass(cplx<T>* buf, int ldb, int batch, int N, int p, unsigned mag_m,
                                         unsigned mag_p, const cplx<T>* __restrict__ twl, int sign) {
  constexpr int MB = (ELEMS / R + 255) / 256;
  const int M = N / R;
  const int total = batch * M;
  const int tstride = N / (p * R);
  cplx<T> x[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < total) {
      const int bb = fdiv(t, mag_m), idx = t - bb * M;
      const cplx<T>* base = buf + bb * ldb;
#pragma unroll
      for (int mm = 0; mm < R; ++mm) x[q][mm] = base[fpad<T>(idx + mm * M)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < total) {
      const int bb = fdiv(t, mag_m), idx = t - bb * M;
      const int k = idx - fdiv(idx, mag_p) * p;
      // the first pass (p = 1) has k = 0: every twiddle is 1
      if (p > 1) {
#pragma unroll
        for (int mm = 1; mm < R; ++mm) {
          const cplx<T> wv = twl[mm * k * tstride];
          const cplx<T> wc = {wv.re, sign < 0 ? wv.im : -wv.im};
          x[q][mm] = cmul(x[q][mm], wc);
        }
      }
      fft_butterfly<T, R>(x[q], sign);
      cplx<T>* base = buf + bb * ldb;
      const int j0 = (idx - k) * R + k;
#pragma unroll
      for (int mm = 0; mm < R; ++mm) base[fpad<T>(j0 + mm * p)] = x[q][mm];
    }
  }
  __syncthreads();
}

// First pass (span p = 1: no twiddles) of a batch of 1 << lb sequences, fed from registers:
// ld(b, x) returns input x of sequence b.  Threads take the sequence index fastest, so a caller
// whose sequences are adjacent in memory reads contiguous segments.  Its outputs go to LDS in the
// layout of fft_pass; the closing barrier also publishes whatever the block staged before.
template <typename T, int R, int ELEMS, typename LD>
__device__ __forceinline__ void fft_pass_first(cplx<T>* buf, int ldb, int lb, int N, LD ld, int sign) {
  constexpr int MB = (ELEMS / R + 255) / 256;
  const int M = N / R;
  const int total = M << lb;
  cplx<T> x[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < total) {
      const int b = t & ((1 << lb) - 1), idx = t >> lb;
#pragma unroll
      for (int mm = 0; mm < R; ++mm) x[q][mm] = ld(b, idx + mm * M);
    }
  }
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < total) {
      const int b = t & ((1 << lb) - 1), idx = t >> lb;
      fft_butterfly<T, R>(x[q], sign);
      cplx<T>* base = buf + b * ldb;
#pragma unroll
      for (int mm = 0; mm < R; ++mm) base[fpad<T>(idx * R + mm)] = x[q][mm];
    }
  }
  __syncthreads();
}

// Last pass (span p = N / R: k = idx, twiddle stride 1, outputs at idx + r*p), its outputs handed
// to st(b, x, value) instead of LDS; sequence index fastest, as in fft_pass_first.  Nothing is
// written to LDS, so no barrier follows.
template <typename T, int R, int ELEMS, typename ST>
__device__ __forceinline__ void fft_pass_last(const cplx<T>* buf, int ldb, int lb, int N,
                                              const cplx<T>* __restrict__ twl, ST st, int sign) {
  constexpr int MB = (ELEMS / R + 255) / 256;
  const int M = N / R;
  const int total = M << lb;
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int t = threadIdx.x + q * 256;
    if (t < total) {
      const int b = t & ((1 << lb) - 1), idx = t >> lb;
      const cplx<T>* base = buf + b * ldb;
      cplx<T> x[R];
#pragma unroll
      for (int mm = 0; mm < R; ++mm) x[mm] = base[fpad<T>(idx + mm * M)];
#pragma unroll
      for (int mm = 1; mm < R; ++mm) {
        const cplx<T> wv = twl[mm * idx];
        const cplx<T> wc = {wv.re, sign < 0 ? wv.im : -wv.im};
        x[mm] = cmul(x[mm], wc);
      }
      fft_butterfly<T, R>(x, sign);
#pragma unroll
      for (int mm = 0; mm < R; ++mm) st(b, idx + mm * M, x[mm]);
    }
  }
}

// Whole transform of a batch of 1 << lb sequences (f.npass >= 2) with the first pass reading
// through ld and the last writing through st, so the inputs and outputs never pass through LDS
// on their own: two LDS round trips and four barriers fewer than staging + lds_fft + unstaging.
template <typename T, int ELEMS, typename LD, typename ST>
__device__ void fft_through(cplx<T>* buf, int ldb, int lb, const FftDesc& f, const cplx<T>* __restrict__ twl,
                            LD ld, ST st, int sign) {
  const int batch = 1 << lb;
  switch (f.radix[0]) {
    case 8: fft_pass_first<T, 8, ELEMS>(buf, ldb, lb, f.n, ld, sign); break;
    case 4: fft_pass_first<T, 4, ELEMS>(buf, ldb, lb, f.n, ld, sign); break;
    case 2: fft_pass_first<T, 2, ELEMS>(buf, ldb, lb, f.n, ld, sign); break;
    case 3: fft_pass_first<T, 3, ELEMS>(buf, ldb, lb, f.n, ld, sign); break;
    default: fft_pass_first<T, 5, ELEMS>(buf, ldb, lb, f.n, ld, sign); break;
  }
  int p = f.radix[0];
  for (int ps = 1; ps + 1 < f.npass; ++ps) {
    const int R = f.radix[ps];
    if (R == 8) fft_pass<T, 8, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 4) fft_pass<T, 4, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 2) fft_pass<T, 2, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 3) fft_pass<T, 3, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else fft_pass<T, 5, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    p *= R;
  }
  switch (f.radix[f.npass - 1]) {
    case 8: fft_pass_last<T, 8, ELEMS>(buf, ldb, lb, f.n, twl, st, sign); break;
    case 4: fft_pass_last<T, 4, ELEMS>(buf, ldb, lb, f.n, twl, st, sign); break;
    case 2: fft_pass_last<T, 2, ELEMS>(buf, ldb, lb, f.n, twl, st, sign); break;
    case 3: fft_pass_last<T, 3, ELEMS>(buf, ldb, lb, f.n, twl, st, sign); break;
    default: fft_pass_last<T, 5, ELEMS>(buf, ldb, lb, f.n, twl, st, sign); break;
  }
}

// Twiddles exp(-2*pi*i*t/N), t in [0, N), copied from the float64 table into LDS in the
// working type (every thread of the block takes part; the caller's next barrier publishes them).
template <typename T>
__device__ __forceinline__ void fft_stage_twiddles(cplx<T>* twl, const double2* __restrict__ tw, int N) {
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const double2 v = tw[t];
    twl[t] = {(T)v.x, (T)v.y};
  }
}

// batch * f.n <= ELEMS
template <typename T, int ELEMS = TC_FFT_ELEMS>
__device__ void lds_fft(cplx<T>* buf, int ldb, int batch, const FftDesc& f, const cplx<T>* __restrict__ twl,
                        int sign) {
  int p = 1;
  for (int ps = 0; ps < f.npass; ++ps) {
    const int R = f.radix[ps];
    if (R == 8) fft_pass<T, 8, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 4) fft_pass<T, 4, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 2) fft_pass<T, 2, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else if (R == 3) fft_pass<T, 3, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    else fft_pass<T, 5, ELEMS>(buf, ldb, batch, f.n, p, f.mag_m[ps], f.mag_p[ps], twl, sign);
    p *= R;
  }
}
