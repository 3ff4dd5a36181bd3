// FFT + wall-normal tridiagonal pressure Poisson solve (projection_fft.py:3-165).
//
// Spectral layout: S[m][j][k] complex, m = kx in [0, nx/2] (real-to-half transform along x),
// j = ky index in [0, ny) (complex transform along y), k in [0, nz) contiguous.
//   P1 k_fwd_x : div (or div/dt computed from the velocities) -> x half spectra, two z-rows per
//                complex transform (real/imag packing)
//   P2 k_fft_y : forward complex FFT along y
//   P3 k_poisson_thomas : one Thomas solve per (kx, ky) system along z (the two recurrences as
//                affine-map scans over 16 or 32 lanes, factors precomputed per plan), (0,0)
//                gauge-fixed in f64
//   P4 k_fft_y : inverse complex FFT along y
//   P5 k_inv_x : half spectra -> real rows, 1/(nx*ny), p written with every ghost image
#include <mutex>

#include "tc_ensemble.h"
#include "tc_fft.h"
#include "tc_internal.h"
#include "tc_poisson.h"
#include "tc_poisson_dev.h"

namespace tc {
// P1-P5 bodies and their single-channel kernels: tc_poisson_dev.h

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

template <typename T, int RM>
__global__ __launch_bounds__(256) void k_poisson_thomas_ens(const EnsChan<T>* __restrict__ ch,
                                                            const int* __restrict__ start, int n) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[q];   // fields are passed to the body by value
  const int L = thomas_lanes(c.d.nz);
  thomas_body<T, RM>(c.nkx * c.d.ny, c.d.nz, thomas_rows_per_lane(c.d.nz), L, TH_THREADS / L, c.cl, c.cr, c.K2,
                     c.th_cb, c.spec, gb - start[q]);
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

// Allocation-free form of the batched Thomas sweep: c' lives in this file-scope device array
// instead of a caller workspace.  Each GPU holds its own copy for the life of the process; every
// call on one GPU shares it (calls that overlap on one GPU race), and it keeps the last call's
// contents (c' is written before it is read, so nothing needs clearing).
static __device__ __align__(256) double scratch[TC_THOMAS_SCRATCH_WORDS];

// Device address of `scratch` per device ordinal, looked up once per device with that device
// current.  In host code the array's name is only a host-side stand-in, never a device address.
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[TC_MAX_DEVICES] = {};

static int scratch_address(void** out) {
  int dev = -1;
  TC_HIP_RC(hipGetDevice(&dev));
  if (dev < 0 || dev >= TC_MAX_DEVICES) return TC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TC_HIP_RC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return TC_ERR_UNSUPPORTED;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TC_OK;
}

int thomas_batched_static(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                          long long s, const void* d, void* x, hipStream_t st) {
  const size_t words = dtype == TC_F64 ? (size_t)TC_THOMAS_SCRATCH_WORDS : 2 * (size_t)TC_THOMAS_SCRATCH_WORDS;
  if ((size_t)n * (size_t)batch > words) return TC_ERR_SCRATCH;
  void* work = nullptr;
  const int rc = scratch_address(&work);
  if (rc != TC_OK) return rc;
  return thomas_batched(n, batch, dtype, cplx_d, a, b, c, s, d, x, work, st);
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
static bool factor_fft(int n, FftDesc& f) {
  f.n = n;
  f.npass = 0;
  int m = n;
  while (m % 8 == 0 && f.npass < TC_FFT_MAXPASS) { f.radix[f.npass++] = 8; m /= 8; }
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
  if ((d.nx & 1) || (d.ny & 1)) return TC_ERR_SHAPE;
  if (!factor_fft(d.nx, pl->fx) || !factor_fft(d.ny, pl->fy)) return TC_ERR_SHAPE;
  if (d.nx > TC_FFT_ELEMS || d.ny > TC_FFT_ELEMS) return TC_ERR_SHAPE;
  // z rows per x / y transform block: up to 8 pairs / 16 rows within the element budget, and
  // for short rows more, up to 32 pairs / 64 rows, while eight 256-thread blocks (a CU's 32
  // wavefronts) still fit the 160 KiB of LDS, i.e. 20 KiB per block: the spectrum's rows are then
  // read and written in longer runs (48-channel float32 ensemble: step -7 %; doubling the rows
  // past that bound, float64 at 96x96x64: 15 -> 30 KiB per block, step +3 %)
  auto rows_of = [&](int n, int base_cap, int max_cap, int zcap) {
    int r = std::min(base_cap, pow2_floor(std::max(1, TC_FFT_ELEMS / n)));
    r = std::min(r, zcap);
    while (r * 2 <= max_cap && r * 2 * n <= TC_FFT_ELEMS && r * 2 <= zcap) {
      const size_t l = g->dtype == TC_F64 ? fft_lds<double>(r * 2, n) : fft_lds<float>(r * 2, n);
      if (l > 20480) break;
      r *= 2;
    }
    return r;
  };
  pl->pairs = rows_of(d.nx, 8, 32, pow2_floor(std::max(1, (d.nz + 1) / 2)));
  pl->kzy = rows_of(d.ny, 16, 64, pow2_floor(d.nz));
  // the forward x transform of the solve takes the wide budget where that doubles the z rows per
  // block within 64 KiB of LDS: at nx = 384 each x row is then read in segments of 16 z rows
  // instead of 8 (the inverse transform and the y transforms keep the narrow budget, which needs
  // half the registers)
  pl->pairs_fx = pl->pairs;
  pl->wide_fx = 0;
  {
    const int zcap = pow2_floor(std::max(1, (d.nz + 1) / 2));
    int pw = std::min(8, pow2_floor(std::max(1, TC_FFT_ELEMS_WIDE / d.nx)));
    pw = std::min(pw, zcap);
    const size_t lw = g->dtype == TC_F64 ? fft_lds<double>(pw, d.nx) : fft_lds<float>(pw, d.nx);
    if (pw > pl->pairs && lw <= 65536) {
      pl->pairs_fx = pw;
      pl->wide_fx = 1;
    }
    // float32 rows of more than 256: 16 pairs, each x row read in 32-row (128-byte) runs
    // (384x384x256: P1 202 -> 190 us; at 192x192x128 the same budget was slower, 21.5 -> 24.7 us,
    // and these rows keep the narrow one)
    if (g->dtype == TC_F32 && d.nx > 256) {
      const int p2w = std::min(std::min(16, pow2_floor(std::max(1, TC_FFT_ELEMS_WIDE2 / d.nx))), zcap);
      if (p2w > pl->pairs_fx && fft_lds<float>(p2w, d.nx) <= 65536) {
        pl->pairs_fx = p2w;
        pl->wide_fx = 2;
      }
    }
  }
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
  // float64 exactly as the reference builds tri_a/b/c, and as k_poisson_thomas rebuilds them,
  // factorised in long double); kept: c' of the row before each lane's first row, [system][L]
  const int nsys = nkx * ny;
  const int TL = thomas_lanes(nz), TR = thomas_rows_per_lane(nz);
  std::vector<double> cb((size_t)nsys * TL, 0.0);
  for (int sy = 0; sy < nsys; ++sy) {
    const bool gauge = sy == 0;
    const double K2 = k2[sy];
    long double cprev = 0.0L;
    for (int k = 0; k < nz; ++k) {
      if (k % TR == 0) cb[(size_t)sy * TL + k / TR] = (double)cprev;
      const double a = (k >= 1 && !(gauge && k == 1)) ? cl[k] : 0.0;
      double b = -((cl[k] + cr[k]) + K2);
      if (k == 0) b += cl[0];
      if (k == nz - 1) b += cr[nz - 1];
      double c = k <= nz - 2 ? cr[k] : 0.0;
      if (gauge && k == 0) { b = 1.0; c = 0.0; }
      const long double den = (long double)b - (long double)a * cprev;
      cprev = (long double)c / den;
    }
  }
  if (hipMalloc(&pl->th_cb, sizeof(double) * cb.size()) != hipSuccess) return TC_ERR_NOMEM;
  hipMemcpy(pl->th_cb, cb.data(), sizeof(double) * cb.size(), hipMemcpyHostToDevice);
  const size_t esz = g->dtype == TC_F64 ? 16 : 8;
  if (hipMalloc(&pl->spec, esz * (size_t)nkx * ny * nz) != hipSuccess) return TC_ERR_NOMEM;
  // statistics workspace: 6 planes (u, v, w at the two sampled heights), float64
  if (hipMalloc(&pl->planes, sizeof(double) * (size_t)nx * ny * 6) != hipSuccess) return TC_ERR_NOMEM;
  if (hipMalloc(&pl->plane_spec, 16 * (size_t)nkx * ny * 6) != hipSuccess) return TC_ERR_NOMEM;
  return TC_OK;
}

void poisson_free(Poisson* pl) {
  hipFree(pl->twx); hipFree(pl->twy); hipFree(pl->cl); hipFree(pl->cr); hipFree(pl->K2);
  hipFree(pl->spec); hipFree(pl->planes); hipFree(pl->plane_spec);
  hipFree(pl->th_cb);
}

// pairs * nx must fit the budget the kernel is built for (wide 0: TC_FFT_ELEMS, 1:
// TC_FFT_ELEMS_WIDE, 2: TC_FFT_ELEMS_WIDE2, float32 only)
template <typename T>
int launch_fwd_x_t(const Poisson* pl, bool from_vel, const void* u, const void* v, const void* w,
                   const void* rin, int nz_in, const double* scal, void* spec, int pairs, int wide,
                   hipStream_t st) {
  const Grid* g = pl->g;
  const Dims& d = g->d;
  const int KZ = 2 * pairs;
  if (pairs < 1 || (pairs & (pairs - 1)) != 0) return TC_ERR_ARG;
  if (wide == 2 && sizeof(T) != 4) return TC_ERR_ARG;
  if (pairs * d.nx > (wide == 2 ? TC_FFT_ELEMS_WIDE2 : wide == 1 ? TC_FFT_ELEMS_WIDE : TC_FFT_ELEMS))
    return TC_ERR_ARG;
  const int blocks = d.ny * ((nz_in + KZ - 1) / KZ);
  const size_t lds = fft_lds<T>(pairs, d.nx);
  const Metric<T> mt = g->metric<T>();
#define TC_FWDX(FV, EL)                                                                               \
  hipLaunchKernelGGL((k_fwd_x<T, FV, EL>), dim3(blocks), dim3(256), lds, st, d, mt, (const T*)u,     \
                     (const T*)v, (const T*)w, (const T*)rin, nz_in, scal, pl->fx, pl->twx, (cplx<T>*)spec, \
                     pairs)
  if (from_vel) {
    if (wide == 2) {
      if constexpr (sizeof(T) == 4) TC_FWDX(true, TC_FFT_ELEMS_WIDE2);
    } else if (wide == 1) {
      TC_FWDX(true, TC_FFT_ELEMS_WIDE);
    } else {
      TC_FWDX(true, TC_FFT_ELEMS);
    }
  } else {
    if (wide == 2) {
      if constexpr (sizeof(T) == 4) TC_FWDX(false, TC_FFT_ELEMS_WIDE2);
    } else if (wide == 1) {
      TC_FWDX(false, TC_FFT_ELEMS_WIDE);
    } else {
      TC_FWDX(false, TC_FFT_ELEMS);
    }
  }
#undef TC_FWDX
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
int launch_fft_y_t(const Poisson* pl, void* spec, int nz_in, int kz, int sign, hipStream_t st) {
  const Dims& d = pl->g->d;
  if (kz < 1 || (kz & (kz - 1)) != 0) return TC_ERR_ARG;
  const int blocks = pl->nkx * ((nz_in + kz - 1) / kz);
  const size_t lds = fft_lds<T>(kz, d.ny);
  hipLaunchKernelGGL((k_fft_y<T>), dim3(blocks), dim3(256), lds, st, pl->nkx, d.ny, nz_in, pl->fy,
                     pl->twy, (cplx<T>*)spec, kz, sign);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// One stage of the Poisson solve: 0 = P1 forward x, 1 = P2 forward y, 2 = P3 Thomas, 3 = P4 inverse y,
// 4 = P5 inverse x + ghosts.
template <typename T>
static int poisson_phase_t(const Poisson* pl, int ph, bool from_vel, const void* u, const void* v,
                           const void* w, const void* rin, const double* scal, void* p, hipStream_t st) {
  const Grid* g = pl->g;
  const Dims& d = g->d;
  switch (ph) {
    case 0:
      return launch_fwd_x_t<T>(pl, from_vel, u, v, w, rin, d.nz, scal, pl->spec, pl->pairs_fx, pl->wide_fx, st);
    case 1: return launch_fft_y_t<T>(pl, pl->spec, d.nz, pl->kzy, -1, st);
    case 2: {
      const int nsys = pl->nkx * d.ny;
      const int L = thomas_lanes(d.nz);
      const int sb = std::min(pl->g->thomas_sb > 0 ? pl->g->thomas_sb : 16, TH_THREADS / L);
      const int blocks = (nsys + sb - 1) / sb;
      const int rm = thomas_rows_per_lane(d.nz);
      // 16-byte row vectors when every lane holds rm whole rows that start on 16-byte boundaries
      // (all arrays are [system][nz] with 256-byte aligned bases)
      const bool vec = d.nz == L * rm && (rm * (int)sizeof(T)) % 16 == 0;
      // staged through LDS (tuning knob) where the rows move as vectors
      const bool stg = vec && pl->g->thomas_lds == 1;
      switch (rm) {
#define TC_TH(RM, V, S)                                                                                 \
  hipLaunchKernelGGL((k_poisson_thomas<T, RM, V, S>), dim3(blocks), dim3(sb * L),                      \
                     (S ? (size_t)sb * L * thomas_seg<T, RM>() : 0), st, nsys, d.nz, sb, pl->cl, pl->cr, \
                     pl->K2, pl->th_cb, (cplx<T>*)pl->spec)
        case 1: TC_TH(1, false, false); break;
        case 2:
          if constexpr (sizeof(T) == 8) {
            if (stg) TC_TH(2, true, true);
            else if (vec) TC_TH(2, true, false);
            else TC_TH(2, false, false);
          } else {
            TC_TH(2, false, false);
          }
          break;
        case 4:
          if (stg) TC_TH(4, true, true);
          else if (vec) TC_TH(4, true, false);
          else TC_TH(4, false, false);
          break;
        case 8:
          if (stg) TC_TH(8, true, true);
          else if (vec) TC_TH(8, true, false);
          else TC_TH(8, false, false);
          break;
#undef TC_TH
        default: return TC_ERR_UNSUPPORTED;
      }
      TC_CHECK_LAUNCH();
      return TC_OK;
    }
    case 3: return launch_fft_y_t<T>(pl, pl->spec, d.nz, pl->kzy, +1, st);
    case 4: {
      const int KZ = 2 * pl->pairs;
      const int tasks = d.ny * ((d.nz + KZ - 1) / KZ);
      const size_t lds = fft_lds<T>(pl->pairs, d.nx);
      const T scale = (T)(1.0 / ((double)d.nx * (double)d.ny));
      hipLaunchKernelGGL((k_inv_x<T>), dim3(tasks), dim3(256), lds, st, d, pl->fx, pl->twx,
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
                                   pl->plane_spec, pairs, false, st)) != TC_OK)
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
  const int KZ = 2 * pl->pairs;
  blocks[0] = d.ny * ((d.nz + KZ - 1) / KZ);
  lds[0] = tsz == 8 ? fft_lds<double>(pl->pairs, d.nx) : fft_lds<float>(pl->pairs, d.nx);
  blocks[1] = pl->nkx * ((d.nz + pl->kzy - 1) / pl->kzy);
  lds[1] = tsz == 8 ? fft_lds<double>(pl->kzy, d.ny) : fft_lds<float>(pl->kzy, d.ny);
  const int tsb = TH_THREADS / thomas_lanes(d.nz);
  blocks[2] = (pl->nkx * d.ny + tsb - 1) / tsb;
  lds[2] = 0;
  blocks[3] = blocks[0];
  lds[3] = lds[0];
}

template <typename T>
static int ens_poisson_t(const void* dch, int n, const int* const* start, const int* blocks, const size_t* lds,
                         int th_rm, hipStream_t st) {
  const EnsChan<T>* c = (const EnsChan<T>*)dch;
  hipLaunchKernelGGL((k_fwd_x_ens<T>), dim3(blocks[0]), dim3(256), lds[0], st, c, start[0], n);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_fft_y_ens<T>), dim3(blocks[1]), dim3(256), lds[1], st, c, start[1], n, -1);
  TC_CHECK_LAUNCH();
  switch (th_rm) {
#define TC_TH(RM)                                                                                       \
  case RM:                                                                                              \
    hipLaunchKernelGGL((k_poisson_thomas_ens<T, RM>), dim3(blocks[2]), dim3(256), 0, st, c, start[2], n); \
    break;
    TC_TH(1) TC_TH(2) TC_TH(4) TC_TH(8)
#undef TC_TH
    default: return TC_ERR_UNSUPPORTED;
  }
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_fft_y_ens<T>), dim3(blocks[1]), dim3(256), lds[1], st, c, start[1], n, +1);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_inv_x_ens<T>), dim3(blocks[3]), dim3(256), lds[3], st, c, start[3], n);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int ens_launch_poisson(int dtype, const void* dch, int n, const int* const* start, const int* blocks,
                       const size_t* lds, int th_rm, hipStream_t st) {
  for (int q = 0; q < 4; ++q)
    if (lds[q] > 160 * 1024) return TC_ERR_ARG;
  if (dtype == TC_F64) return ens_poisson_t<double>(dch, n, start, blocks, lds, th_rm, st);
  return ens_poisson_t<float>(dch, n, start, blocks, lds, th_rm, st);
}

int thomas_rows(int nz) { return thomas_rows_per_lane(nz); }

}  // namespace tc
