// Shared device helpers for the torChannel IMEX kernels (gfx950 / CDNA4).
//
// Array conventions (reference index space, z-contiguous C order, 0-based):
//   u (nx+1, ny+2, nz+2)   v (nx+2, ny+1, nz+2)   w (nx+2, ny+2, nz+1)   p (nx+2, ny+2, nz+2)
// initflow.py:6-9 allocates these shapes; every kernel here indexes them the same way.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/torchannel.h"

#define TC_WAVE 64
#define TC_MAX_NZ 256        // rows per wall-normal column held by one wavefront (4 per lane)
#define TC_MAX_STAGES 8      // ceil(log2(TC_MAX_NZ))
#define TC_TAB_STRIDE (TC_MAX_NZ * (2 * TC_MAX_STAGES + 1))

#define TC_CHECK_LAUNCH()                                      \
  do {                                                         \
    hipError_t e_ = hipGetLastError();                         \
    if (e_ != hipSuccess) return -(int)e_;                     \
  } while (0)

// Return a failed HIP call's status as the library's negative code.
#define TC_HIP_RC(x)                                           \
  do {                                                         \
    hipError_t e_ = (x);                                       \
    if (e_ != hipSuccess) return -(int)e_;                     \
  } while (0)

#define TC_MAX_DEVICES 64     // device ordinals with a per-device static scratch address
#define TC_THOMAS_SCRATCH_WORDS TC_THOMAS_STATIC_MAX_F64

// Extents and strides of the four ghosted arrays.
struct Dims {
  int nx, ny, nz;
  int top;               // +1: free-slip top wall (ghost = +interior), -1: no-slip (ghost = -interior)
  int syu, sxu;          // u strides (elements)
  int syv, sxv;
  int syw, sxw;
  int syp, sxp;
};

// pitch 0: each array dense (column stride = its own column length, the torch.zeros layout);
// pitch > 0: every array's columns `pitch` elements apart (the padded layout of
// tc_grid_set_pitch, pitch >= nz + 2).
static inline Dims make_dims(int nx, int ny, int nz, int top_bc, int pitch = 0) {
  Dims d;
  d.nx = nx; d.ny = ny; d.nz = nz;
  d.top = top_bc == TC_TOP_NEUMANN ? 1 : -1;
  const int cu = pitch > 0 ? pitch : nz + 2, cw = pitch > 0 ? pitch : nz + 1;
  d.syu = cu; d.sxu = (ny + 2) * cu;
  d.syv = cu; d.sxv = (ny + 1) * cu;
  d.syw = cw; d.sxw = (ny + 2) * cw;
  d.syp = cu; d.sxp = (ny + 2) * cu;
  return d;
}

// Elements an array spans from its first entry to one past its last column's stride: u (nx+1
// planes), v, w, p (nx+2 planes).  Flat passes over a whole array (copies, fills, scaling) cover
// this many elements, padding included.
static inline size_t field_span(const Dims& d, int comp) {
  switch (comp) {
    case 0: return (size_t)(d.nx + 1) * d.sxu;
    case 1: return (size_t)(d.nx + 2) * d.sxv;
    case 2: return (size_t)(d.nx + 2) * d.sxw;
    default: return (size_t)(d.nx + 2) * d.sxp;
  }
}

// Per-grid wall-normal metric, precomputed once (tc_grid_create) in float64 and in the
// kernel's working type.  Index ranges: rdzf/dzf [0, nz), rdzc/dzc [0, nz].
template <typename T>
struct Metric {
  const T* rdzf;         // 1/dz_f
  const T* rdzc;         // 1/dz_c
  const double* dzf;     // dz_f (float64; coefficient tables are built from these)
  const double* dzc;     // dz_c
  const double* wbulk;   // dx*dy*dz_f[k] (float64 bulk weights, solver.py:231)
  T rdx, rdy;            // 1/dx, 1/dy
  T kxx, kyy;            // nu/dx^2, nu/dy^2 (operators.py:1064-1067)
  double nu;
};

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD walks a contiguous range
// of logical tiles and its private L2 sees the neighbouring columns of the tiles it runs.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, rem = nb & 7;
  const int x = b & 7, idx = b >> 3;
  return x < rem ? x * (q + 1) + idx : rem * (q + 1) + (x - rem) * q + idx;
}

// Column order of the per-column kernels: wave g -> (component, i, j), component fastest, the
// (i, j) plane walked in TC_TI x TC_TJ tiles so that the i+-1 and j+-1 neighbour columns a
// stencil reads were loaded by waves only a tile earlier (still in the XCD's L2).
#define TC_TI 8
#define TC_TJ 8
__host__ __device__ __forceinline__ int tiled_waves(int nxp, int nyp) {
  return 3 * ((nxp + TC_TI - 1) / TC_TI) * TC_TI * ((nyp + TC_TJ - 1) / TC_TJ) * TC_TJ;
}
__device__ __forceinline__ bool tiled_column(int g, int nxp, int nyp, int& comp, int& i, int& j) {
  comp = g % 3;
  const int pos = g / 3;
  const int tj_n = (nyp + TC_TJ - 1) / TC_TJ;
  const int tile = pos / (TC_TI * TC_TJ), w = pos - tile * (TC_TI * TC_TJ);
  i = (tile / tj_n) * TC_TI + w / TC_TJ;
  j = (tile - (tile / tj_n) * tj_n) * TC_TJ + (w - (w / TC_TJ) * TC_TJ);
  return i < nxp && j < nyp;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (TC_WAVE - 1); }

// DPP lane moves within rows of 16 lanes (GFX9 encodings): row_shr:n takes lane l - n, row_shl:n
// lane l + n; lanes whose source falls outside their row keep `old`.
#define TC_DPP_ROW_SHL(n) (0x100 + (n))
#define TC_DPP_ROW_SHR(n) (0x110 + (n))

// row_bcast:15 (every lane of the rows enabled by the row mask takes lane 15 of the row below)
#define TC_DPP_ROW_BCAST15 0x142

template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ float dpp_row(float old, float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), CTRL, ROWMASK, 0xf, false));
}

template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp_row(double old, double x) {
  const long long o = __double_as_longlong(old), v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// value of lane l (uniform l) for every lane
__device__ __forceinline__ float readlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double readlane(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, l), hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, TC_WAVE);
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, TC_WAVE));
  return x;
}

// Rows of a wall-normal column live in registers: row m = r*64 + lane, r in [0, R).
// shift_dn: out[r] = row (m - s), shift_up: out[r] = row (m + s); rows outside [0, 64R) -> pad.
// S is a compile-time shift so all register indexing stays static.
template <int S, int R, typename T>
__device__ __forceinline__ void shift_dn(const T (&x)[R], T (&out)[R], T pad, int lane) {
  if constexpr (S % TC_WAVE == 0) {
    constexpr int q = S / TC_WAVE;
#pragma unroll
    for (int r = 0; r < R; ++r) out[r] = (r - q >= 0) ? x[(r - q >= 0) ? r - q : 0] : pad;
  } else {
    constexpr int q = S / TC_WAVE, s = S % TC_WAVE;
    T t[R];
#pragma unroll
    for (int r = 0; r < R; ++r) t[r] = __shfl(x[r], (lane - s) & (TC_WAVE - 1), TC_WAVE);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ra = r - q, rb = r - q - 1;            // same wrap row / previous wrap row
      const T a = ra >= 0 ? t[ra >= 0 ? ra : 0] : pad;
      const T b = rb >= 0 ? t[rb >= 0 ? rb : 0] : pad;
      out[r] = lane >= s ? a : b;
    }
  }
}

template <int S, int R, typename T>
__device__ __forceinline__ void shift_up(const T (&x)[R], T (&out)[R], T pad, int lane) {
  if constexpr (S % TC_WAVE == 0) {
    constexpr int q = S / TC_WAVE;
#pragma unroll
    for (int r = 0; r < R; ++r) out[r] = (r + q < R) ? x[(r + q < R) ? r + q : 0] : pad;
  } else {
    constexpr int q = S / TC_WAVE, s = S % TC_WAVE;
    T t[R];
#pragma unroll
    for (int r = 0; r < R; ++r) t[r] = __shfl(x[r], (lane + s) & (TC_WAVE - 1), TC_WAVE);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ra = r + q, rb = r + q + 1;
      const T a = ra < R ? t[ra < R ? ra : 0] : pad;
      const T b = rb < R ? t[rb < R ? rb : 0] : pad;
      out[r] = lane + s < TC_WAVE ? a : b;
    }
  }
}

// Neighbour exchange along a column through a per-wavefront LDS row buffer: rows live at
// buf[64R + m] between two zero pads of 64R rows, so row m+-s (s <= 64R) is one ds_read with an
// immediate offset and out-of-range rows read zero.  LDS operations of one wavefront execute in
// order; the wave barriers keep the compiler from moving them across each other.
template <int R, typename T>
__device__ __forceinline__ void rows_zero_pads(T* buf, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    buf[r * TC_WAVE + lane] = T(0);
    buf[2 * TC_WAVE * R + r * TC_WAVE + lane] = T(0);
  }
}

template <int R, typename T>
__device__ __forceinline__ void rows_put(T* buf, const T (&x)[R], int lane) {
  // the fences order this wavefront's LDS accesses for the compiler (one lane reads what its
  // neighbours wrote); a wavefront's LDS instructions execute in order, so they emit no code
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int r = 0; r < R; ++r) buf[TC_WAVE * R + r * TC_WAVE + lane] = x[r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int OFF, int R, typename T>
__device__ __forceinline__ void rows_get(const T* buf, T (&out)[R], int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = buf[TC_WAVE * R + r * TC_WAVE + lane + OFF];
}

template <int ST, int R, typename T>
__device__ __forceinline__ void pcr_shared_stage_lds(T (&dv)[R], T* buf, const double* __restrict__ tab, int n,
                                                     int lane) {
  constexpr int s = 1 << ST;
  if (s >= n) return;
  rows_put<R, T>(buf, dv, lane);
  T dn[R], up[R];
  rows_get<-s, R, T>(buf, dn, lane);
  rows_get<s, R, T>(buf, up, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    const T al = (T)tab[(2 * ST) * TC_MAX_NZ + m];
    const T ga = (T)tab[(2 * ST + 1) * TC_MAX_NZ + m];
    dv[r] = dv[r] + al * dn[r] + ga * up[r];
  }
}

// PCR with shared tables (see pcr_shared_solve) using the LDS row buffer for the exchanges.
template <int R, typename T>
__device__ __forceinline__ void pcr_shared_solve_lds(T (&dv)[R], T* buf, const double* __restrict__ tab, int n,
                                                     int lane) {
  pcr_shared_stage_lds<0, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<1, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<2, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<3, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<4, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<5, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<6, R, T>(dv, buf, tab, n, lane);
  pcr_shared_stage_lds<7, R, T>(dv, buf, tab, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    dv[r] = dv[r] * (T)tab[2 * TC_MAX_STAGES * TC_MAX_NZ + m];   // reciprocal diagonal
  }
}

// Right-hand side d = q + beta*D2(q) of the CN system for one u/v column (operators.py:627-644):
// rows m = 0..n-1 hold cell k = m+1; the wall closures are the interior stencil with the
// reflected neighbours X[-1] = -X[0] and X[n] = top*X[n-1].
template <int R, typename T>
__device__ __forceinline__ void cn_rhs_cell(T (&x)[R], T* rows, const Metric<T>& mt, int n, int top, T beta,
                                            int lane) {
  T dn[R], up[R];
  rows_put<R, T>(rows, x, lane);
  rows_get<-1, R, T>(rows, dn, lane);
  rows_get<1, R, T>(rows, up, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m == 0) dn[r] = -x[r];
    if (m == n - 1) up[r] = top > 0 ? x[r] : -x[r];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) {
      const T lap = ((up[r] - x[r]) * mt.rdzc[m + 1] - (x[r] - dn[r]) * mt.rdzc[m]) * mt.rdzf[m];
      x[r] = x[r] + beta * lap;
    } else {
      x[r] = T(0);
    }
  }
}

// Same for a w column (operators.py:788-803): row m holds face m+1 but its explicit half is
// centred on face m, with the wall value W[0] (`wall`) below row 0, zero below that, and the
// top row's upper difference taken against zero.
template <int R, typename T>
__device__ __forceinline__ void cn_rhs_face(T (&x)[R], T* rows, const Metric<T>& mt, int n, T wall, T beta,
                                            int lane) {
  T d1[R], d2[R];
  rows_put<R, T>(rows, x, lane);
  rows_get<-1, R, T>(rows, d1, lane);
  rows_get<-2, R, T>(rows, d2, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m == 0) d1[r] = wall;
    if (m == 1) d2[r] = wall;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) {
      const T xp = (m == n - 1) ? T(0) : x[r];
      const T lap = ((xp - d1[r]) * mt.rdzf[m + 1] - (d1[r] - d2[r]) * mt.rdzf[m]) * mt.rdzc[m + 1];
      x[r] = x[r] + beta * lap;
    } else {
      x[r] = T(0);
    }
  }
}

// Parallel cyclic reduction of one column whose matrix is shared by every column (the CN
// systems), so the whole a/b/c reduction is a precomputed table: per stage st, alpha[m] at
// tab[st*2*MAXNZ + m], gamma[m] at tab[(st*2+1)*MAXNZ + m]; the reciprocal of the
// final diagonal b at tab[2*STAGES*MAXNZ + m].  Rows >= n carry alpha = gamma = 0, b = 1, d = 0.
// Recurrence of tridiag.py:50-58 restricted to d: d += alpha*d[m-s] + gamma*d[m+s]; x = d/b, as
// d times the table's reciprocal of b.
template <int ST, int R, typename T>
__device__ __forceinline__ void pcr_shared_stage(T (&dv)[R], const double* __restrict__ tab, int n, int lane) {
  constexpr int s = 1 << ST;
  if (s >= n) return;
  T dn[R], up[R];
  shift_dn<s, R, T>(dv, dn, T(0), lane);
  shift_up<s, R, T>(dv, up, T(0), lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    const T al = (T)tab[(2 * ST) * TC_MAX_NZ + m];
    const T ga = (T)tab[(2 * ST + 1) * TC_MAX_NZ + m];
    dv[r] = dv[r] + al * dn[r] + ga * up[r];
  }
}

template <int R, typename T>
__device__ __forceinline__ void pcr_shared_solve(T (&dv)[R], const double* __restrict__ tab, int n, int lane) {
  pcr_shared_stage<0, R, T>(dv, tab, n, lane);
  pcr_shared_stage<1, R, T>(dv, tab, n, lane);
  pcr_shared_stage<2, R, T>(dv, tab, n, lane);
  pcr_shared_stage<3, R, T>(dv, tab, n, lane);
  pcr_shared_stage<4, R, T>(dv, tab, n, lane);
  pcr_shared_stage<5, R, T>(dv, tab, n, lane);
  pcr_shared_stage<6, R, T>(dv, tab, n, lane);
  pcr_shared_stage<7, R, T>(dv, tab, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    dv[r] = dv[r] * (T)tab[2 * TC_MAX_STAGES * TC_MAX_NZ + m];   // reciprocal diagonal
  }
}

// Complex number in the working precision.
template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
