// Roofline calibration sweeps (no counterpart in the reference): the column access pattern of
// the explicit + CN kernel without its arithmetic, so that the kernel's achieved bandwidth can be
// set against what that pattern itself reaches on the device, next to the 8 TB/s HBM peak.
//
// Twelve arrays shaped like p ((nx+2) x (ny+2) x (nz+2), z contiguous, columns `pitch` elements
// apart): A0..A2 stand for u, v, w,
// P0..P2 for the AB2 history (read), O0..O2 for the rhs and S0..S2 for the star velocities
// (written).  Blocks of 4 wavefronts own one component and a band of 4 consecutive i; each
// wavefront marches its (component, i) column along j over jb columns, fetching one column ahead
// and storing two columns per step, in the order and XCD mapping of k_explicit_cn.
//   pattern 0: per column the kernel's eight input columns (own j+1, i-1, i+1, history, and two
//              columns of each other component) and two output columns;
//   pattern 1: only the compulsory streams: own j+1 and history in, two columns out;
//   pattern 2: the three-component pencils' pattern: blocks of 4 wavefronts own a band of 4
//              consecutive i for all three components; each wavefront marches its i along j,
//              per step loading the nine columns (components 0-2 at i-1, i, i+1) of plane j+1
//              and the three history columns of (i, j), and storing six columns.
#include "tc_internal.h"

namespace tc {

template <int R, typename T>
__device__ __forceinline__ void cal_load(const T* __restrict__ col, T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = r * TC_WAVE + lane + 1;
    x[r] = k <= n ? col[k] : T(0);
  }
}

template <int R, typename T>
__device__ __forceinline__ void cal_store(T* __restrict__ col, const T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = r * TC_WAVE + lane + 1;
    if (k <= n) col[k] = x[r];
  }
}

template <typename T, int R>
struct CalIn {
  T N[R], W[R], E[R], rp[R], b0[R], b1[R], c0[R], c1[R];
};

template <typename T, int R, int PAT>
__global__ __launch_bounds__(256) void k_calib_columns(int nx, int ny, int nz, int jb, int pitch, const T* __restrict__ A0,
                                                       const T* __restrict__ A1, const T* __restrict__ A2,
                                                       const T* __restrict__ P0, const T* __restrict__ P1,
                                                       const T* __restrict__ P2, T* __restrict__ O0,
                                                       T* __restrict__ O1, T* __restrict__ O2, T* __restrict__ S0,
                                                       T* __restrict__ S1, T* __restrict__ S2) {
  const long long sy = pitch, sx = (long long)(ny + 2) * pitch;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int comp = b % 3, band = b / 3;
  const int nbj = (ny + jb - 1) / jb;
  const int bi = band / nbj, bj = band - bi * nbj;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int i = 1 + bi * 4 + wv;
  if (i > nx) return;
  const int j0 = 1 + bj * jb, j1 = min(j0 + jb, ny + 1);
  const T* A = comp == 0 ? A0 : comp == 1 ? A1 : A2;
  const T* B = comp == 0 ? A1 : comp == 1 ? A2 : A0;
  const T* C = comp == 0 ? A2 : comp == 1 ? A0 : A1;
  const T* P = comp == 0 ? P0 : comp == 1 ? P1 : P2;
  T* O = comp == 0 ? O0 : comp == 1 ? O1 : O2;
  T* S = comp == 0 ? S0 : comp == 1 ? S1 : S2;
  const long long oi = (long long)i * sx;
  const T* Ai = A + oi;
  T cS[R], cC[R];
  cal_load<R, T>(Ai + (long long)(j0 - 1) * sy, cS, nz, lane);
  cal_load<R, T>(Ai + (long long)j0 * sy, cC, nz, lane);
  CalIn<T, R> L;
  auto fetch = [&](int j) {
    const long long cj = (long long)j * sy;
    cal_load<R, T>(Ai + cj + sy, L.N, nz, lane);
    cal_load<R, T>(P + oi + cj, L.rp, nz, lane);
    if (PAT == 0) {
      cal_load<R, T>(Ai - sx + cj, L.W, nz, lane);
      cal_load<R, T>(Ai + sx + cj, L.E, nz, lane);
      cal_load<R, T>(B + oi + cj, L.b0, nz, lane);
      cal_load<R, T>(B + oi + sx + cj, L.b1, nz, lane);
      cal_load<R, T>(C + oi + cj, L.c0, nz, lane);
      cal_load<R, T>(C + oi + sx + cj, L.c1, nz, lane);
    }
  };
  fetch(j0);
  for (int j = j0; j < j1; ++j) {
    const long long cj = (long long)j * sy;
    T x[R], y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      T s = L.N[r] - T(2) * cC[r] + cS[r] + T(0.5) * L.rp[r];
      if (PAT == 0) s = s + (L.W[r] + L.E[r]) * (L.b0[r] + L.b1[r]) - L.c0[r] * L.c1[r];
      x[r] = s;
      y[r] = s * T(0.25) + cC[r];
      cS[r] = cC[r];
      cC[r] = L.N[r];
    }
    if (j + 1 < j1) fetch(j + 1);
    cal_store<R, T>(O + oi + cj, x, nz, lane);
    cal_store<R, T>(S + oi + cj, y, nz, lane);
  }
}


template <typename T, int R>
__global__ __launch_bounds__(256) void k_calib_uvw(int nx, int ny, int nz, int jb, int pitch, const T* __restrict__ A0,
                                                   const T* __restrict__ A1, const T* __restrict__ A2,
                                                   const T* __restrict__ P0, const T* __restrict__ P1,
                                                   const T* __restrict__ P2, T* __restrict__ O0,
                                                   T* __restrict__ O1, T* __restrict__ O2, T* __restrict__ S0,
                                                   T* __restrict__ S1, T* __restrict__ S2) {
  const long long sy = pitch, sx = (long long)(ny + 2) * pitch;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int nbj = (ny + jb - 1) / jb;
  const int bi = b / nbj, bj = b - bi * nbj;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int i = 1 + bi * 4 + wv;
  if (i > nx) return;
  const int j0 = 1 + bj * jb, j1 = min(j0 + jb, ny + 1);
  const long long oi = (long long)i * sx;
  const T* A[3] = {A0, A1, A2};
  const T* P[3] = {P0, P1, P2};
  T* O[3] = {O0, O1, O2};
  T* S[3] = {S0, S1, S2};
  // c: plane j (component q at i-1, i, i+1 in c[3q..3q+2]); nx9: plane j+1; hs: plane j-1 at i
  T c[9][R], nx9[9][R], hp[3][R], hs[3][R];
  auto plane = [&](int j, T (&x)[9][R]) {
    const long long cj = (long long)j * sy;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cal_load<R, T>(A[q] + oi - sx + cj, x[3 * q], nz, lane);
      cal_load<R, T>(A[q] + oi + cj, x[3 * q + 1], nz, lane);
      cal_load<R, T>(A[q] + oi + sx + cj, x[3 * q + 2], nz, lane);
    }
  };
  plane(j0, c);
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) hs[q][r] = c[3 * q + 1][r];
  for (int j = j0; j < j1; ++j) {
    const long long cj = (long long)j * sy;
    plane(j + 1, nx9);
#pragma unroll
    for (int q = 0; q < 3; ++q) cal_load<R, T>(P[q] + oi + cj, hp[q], nz, lane);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int a = (q + 1) % 3, e = (q + 2) % 3;
      T x[R], y[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T v = nx9[3 * q + 1][r] - T(2) * c[3 * q + 1][r] + hs[q][r] + T(0.5) * hp[q][r];
        v = v + (c[3 * q][r] + c[3 * q + 2][r]) * (c[3 * a + 1][r] + c[3 * a + 2][r]) - c[3 * e + 1][r] * c[3 * e + 2][r];
        x[r] = v;
        y[r] = v * T(0.25) + c[3 * q + 1][r];
      }
      cal_store<R, T>(O[q] + oi + cj, x, nz, lane);
      cal_store<R, T>(S[q] + oi + cj, y, nz, lane);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (q % 3 == 1) hs[q / 3][r] = c[q][r];
        c[q][r] = nx9[q][r];
      }
  }
}

template <typename T, int R>
static void calib_launch(int pattern, int blocks, hipStream_t st, int nx, int ny, int nz, int jb, int pitch,
                         void* const* a) {
  if (pattern == 2)
    hipLaunchKernelGGL((k_calib_uvw<T, R>), dim3(blocks / 3), dim3(256), 0, st, nx, ny, nz, jb, pitch, (const T*)a[0],
                       (const T*)a[1], (const T*)a[2], (const T*)a[3], (const T*)a[4], (const T*)a[5], (T*)a[6],
                       (T*)a[7], (T*)a[8], (T*)a[9], (T*)a[10], (T*)a[11]);
  else if (pattern == 0)
    hipLaunchKernelGGL((k_calib_columns<T, R, 0>), dim3(blocks), dim3(256), 0, st, nx, ny, nz, jb, pitch, (const T*)a[0],
                       (const T*)a[1], (const T*)a[2], (const T*)a[3], (const T*)a[4], (const T*)a[5], (T*)a[6],
                       (T*)a[7], (T*)a[8], (T*)a[9], (T*)a[10], (T*)a[11]);
  else
    hipLaunchKernelGGL((k_calib_columns<T, R, 1>), dim3(blocks), dim3(256), 0, st, nx, ny, nz, jb, pitch, (const T*)a[0],
                       (const T*)a[1], (const T*)a[2], (const T*)a[3], (const T*)a[4], (const T*)a[5], (T*)a[6],
                       (T*)a[7], (T*)a[8], (T*)a[9], (T*)a[10], (T*)a[11]);
}

}  // namespace tc

using namespace tc;

extern "C" int tc_calib_columns(int dtype, int pattern, int nx, int ny, int nz, int jb, int pitch,
                                void* const* arrays, void* stream) {
  if (!arrays || nx < 1 || ny < 1 || nz < 1 || nz > TC_MAX_NZ || jb < 1 || pattern < 0 || pattern > 2)
    return TC_ERR_ARG;
  if (pitch == 0) pitch = nz + 2;
  if (pitch < nz + 2) return TC_ERR_ARG;
  for (int q = 0; q < 12; ++q)
    if (!arrays[q]) return TC_ERR_ARG;
  if (dtype != TC_F64 && dtype != TC_F32) return TC_ERR_DTYPE;
  const int R = (nz + TC_WAVE - 1) / TC_WAVE;
  const int blocks = 3 * ((nx + 3) / 4) * ((ny + jb - 1) / jb);
  hipStream_t st = (hipStream_t)stream;
#define TC_CAL(TT)                                                              \
  switch (R) {                                                                  \
    case 1: calib_launch<TT, 1>(pattern, blocks, st, nx, ny, nz, jb, pitch, arrays); break; \
    case 2: calib_launch<TT, 2>(pattern, blocks, st, nx, ny, nz, jb, pitch, arrays); break; \
    case 3: calib_launch<TT, 3>(pattern, blocks, st, nx, ny, nz, jb, pitch, arrays); break; \
    default: calib_launch<TT, 4>(pattern, blocks, st, nx, ny, nz, jb, pitch, arrays); break; \
  }
  if (dtype == TC_F64) { TC_CAL(double) } else { TC_CAL(float) }
#undef TC_CAL
  TC_CHECK_LAUNCH();
  return TC_OK;
}
