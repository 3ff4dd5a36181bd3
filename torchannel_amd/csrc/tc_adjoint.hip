// Discrete adjoint of the fixed-dt IMEX step (float64).  No counterpart in the reference: it gives
// the gradient of the time-averaged wall shear stress with respect to a body-force field f added
// to rhs_u next to the scalar forcing (the F term of solver.py:626).
//
// Forward step n (solver.py:588-708, as tc_solver_step runs it):
//   R_n     = E(q_n) + F_n + f          E = compute_momentum_rhs_fused_imex (GPU branch)
//   q*      = q_n + dt (c1 R_n - c0 R_{n-1})        (c1, c0) = (1.5, 0.5), (1, 0) on the first step
//   q_{n+1} = P C q*                    C = Crank-Nicolson per column, P = FFT projection
//   F_{n+1} = F_n + 0.1 (U_b - b.u_{n+1}) / dt
// The state is the interior of q (ghost entries are images of interior ones, apply_bc_all).
// Reverse step n, given the cotangents of (q_{n+1}, F_{n+1}, R_n):
//   lq'    += -0.1/dt b lF                                           (forcing update)
//   lq*     = C^T P^T lq'  =  W C P W^{-1} lq'   (u, v)
//                          =  B_w^T W A_w^{-1} P W^{-1} lq'  (w)
//     with W the cell heights dz_f (u, v rows) or face spacings dz_c (w rows): the discrete
//     divergence is -W^{-1} G^T W, so P^T = W P W^{-1}; the cell CN operator is W^{-1}-symmetric,
//     so C^T = W C W^{-1}; the w CN keeps its shifted explicit half B_w (operators.py:788-803)
//     and only the implicit part is symmetrised.
//   lR_n    = lR_acc + c1 dt lq*,  lR_acc <- -c0 dt lq*,  lF += sum(lR_n.u),  grad f += lR_n.u
//   lq_n    = lq* + Img^T E'(q_n)^T lR_n  (+ dJ/du_n)
// Img^T folds the cotangent of every ghost entry onto the interior entry it copies.
#include <algorithm>

#include "tc_internal.h"
#include "tc_poisson.h"

namespace tc {

// ---------------------------------------------------------------------------------------------
// A1: y = W^{-1} (lq' - 0.1/dt b lF) on interior entries, into the star buffers.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_pre(Dims d, const double* __restrict__ dzf,
                                                 const double* __restrict__ dzc,
                                                 const double* __restrict__ wbulk, double inv_volume,
                                                 double rdt, const double* __restrict__ ascal,
                                                 const double* __restrict__ lu, const double* __restrict__ lv,
                                                 const double* __restrict__ lw, double* __restrict__ yu,
                                                 double* __restrict__ yv, double* __restrict__ yw) {
  const int comp = blockIdx.y;
  const int nzc = comp == 2 ? d.nz - 1 : d.nz;
  const long long tot = (long long)d.nx * d.ny * nzc;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= tot) return;
  const int k = 1 + (int)(t % nzc);
  const long long c = t / nzc;
  const int j = 1 + (int)(c % d.ny), i = 1 + (int)(c / d.ny);
  if (comp == 0) {
    const long long o = (long long)i * d.sxu + (long long)j * d.syu + k;
    const double fu = -0.1 * rdt * ascal[TC_A_LF] * wbulk[k - 1] * inv_volume;
    yu[o] = (lu[o] + fu) / dzf[k - 1];
  } else if (comp == 1) {
    const long long o = (long long)i * d.sxv + (long long)j * d.syv + k;
    yv[o] = lv[o] / dzf[k - 1];
  } else {
    const long long o = (long long)i * d.sxw + (long long)j * d.syw + k;
    yw[o] = lw[o] / dzc[k];
  }
}

// ---------------------------------------------------------------------------------------------
// A4: weighted CN adjoint of one interior column + the AB2 / forcing / body-force bookkeeping.
//   u, v: x = W * CN(z)                    (cell tables, explicit half included)
//   w:    x = B_w^T (W * A_w^{-1} z)        (face tables, implicit part only)
// then lq* = x, lR_n = lR_acc + c1 dt x, lR_acc = -c0 dt x, grad_f += lR_n (u), partial sums of
// lR_n over the u interior (for lF).
// ---------------------------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void k_adj_cn(Dims d, Metric<double> mt, const double* __restrict__ zu,
                                                const double* __restrict__ zv, const double* __restrict__ zw,
                                                const double* __restrict__ tab_cell,
                                                const double* __restrict__ tab_face, double dt, double theta,
                                                double c1, double c0, double* __restrict__ qs_u,
                                                double* __restrict__ qs_v, double* __restrict__ qs_w,
                                                double* __restrict__ lr_u, double* __restrict__ lr_v,
                                                double* __restrict__ lr_w, double* __restrict__ acc_u,
                                                double* __restrict__ acc_v, double* __restrict__ acc_w,
                                                double* __restrict__ grad, double* __restrict__ partials) {
  __shared__ double lds_rows[4][3 * TC_WAVE * R];
  __shared__ double red[4];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int g = blockIdx.x * 4 + wv;
  const int ncol = d.nx * d.ny;
  const int comp = g / ncol;
  double part = 0.0;
  if (comp < 3) {
    const int c = g - comp * ncol;
    const int i = 1 + c / d.ny, j = 1 + c % d.ny;
    double* rows = lds_rows[wv];
    rows_zero_pads<R, double>(rows, lane);
    const double beta = (1.0 - theta) * dt * mt.nu;
    long long base;
    const double* z;
    double *qs, *lr, *acc;
    int n;
    if (comp == 0) { base = (long long)i * d.sxu + (long long)j * d.syu; z = zu; qs = qs_u; lr = lr_u; acc = acc_u; n = d.nz; }
    else if (comp == 1) { base = (long long)i * d.sxv + (long long)j * d.syv; z = zv; qs = qs_v; lr = lr_v; acc = acc_v; n = d.nz; }
    else { base = (long long)i * d.sxw + (long long)j * d.syw; z = zw; qs = qs_w; lr = lr_w; acc = acc_w; n = d.nz - 1; }
    double x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * TC_WAVE + lane + 1;
      x[r] = k <= n ? z[base + k] : 0.0;
    }
    if (comp < 2) {
      cn_rhs_cell<R, double>(x, rows, mt, n, d.top, beta, lane);
      pcr_shared_solve_lds<R, double>(x, rows, tab_cell, n, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = r * TC_WAVE + lane;
        x[r] = m < n ? x[r] * mt.dzf[m] : 0.0;
      }
    } else {
      pcr_shared_solve_lds<R, double>(x, rows, tab_face, n, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = r * TC_WAVE + lane;
        x[r] = m < n ? x[r] * mt.dzc[m + 1] : 0.0;
      }
      // B_w^T: B_w row m = x[m] + beta rdzc[m+1] (rdzf[m+1] x'[m] - (rdzf[m+1]+rdzf[m]) x[m-1]
      //        + rdzf[m] x[m-2]), x'[n-1] = 0
      double x1[R], x2[R];
      rows_put<R, double>(rows, x, lane);
      rows_get<1, R, double>(rows, x1, lane);
      rows_get<2, R, double>(rows, x2, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = r * TC_WAVE + lane;
        if (m < n) {
          const double cc0 = m != n - 1 ? mt.rdzc[m + 1] * mt.rdzf[m + 1] : 0.0;
          const double cc1 = m + 1 <= n - 1 ? -mt.rdzc[m + 2] * (mt.rdzf[m + 2] + mt.rdzf[m + 1]) : 0.0;
          const double cc2 = m + 2 <= n - 1 ? mt.rdzc[m + 3] * mt.rdzf[m + 2] : 0.0;
          x[r] = x[r] + beta * (cc0 * x[r] + cc1 * x1[r] + cc2 * x2[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = r * TC_WAVE + lane + 1;
      if (k <= n) {
        const long long o = base + k;
        const double l = acc[o] + c1 * dt * x[r];
        qs[o] = x[r];
        lr[o] = l;
        acc[o] = -c0 * dt * x[r];
        if (comp == 0) {
          grad[o] += l;
          part += l;
        }
      }
    }
  }
  part = wave_sum(part);
  if (lane == 0) red[wv] = part;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// lF += sum of the k_adj_cn partials (one block, fixed order).
__global__ __launch_bounds__(1024) void k_adj_lamf(const double* __restrict__ partials, int nb,
                                                   double* __restrict__ ascal) {
  __shared__ double acc[1024];
  double s = 0.0;
  for (int q = threadIdx.x; q < nb; q += 1024) s += partials[q];
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) acc[threadIdx.x] += acc[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) ascal[TC_A_LF] += acc[0];
}

// ---------------------------------------------------------------------------------------------
// A5: E'(q)^T lR on every entry (ghost columns included) of rows 1..nz (u, v) / 1..nz-1 (w).
// E is the GPU-branch explicit RHS: u on faces 1..nx-1, v on faces 1..ny-1, w on faces 1..nz-1
// (operators.py:1078-1223).  Each output entry gathers, for every flux that reads it, the flux
// cotangent (difference of the rhs cotangents of the two cells the flux separates, times the
// metric) times the partial derivative of the flux.  Wall-normal ghost rows of u and v get no
// cotangent: the only fluxes reading them carry the wall value of w, which is zero.
// ---------------------------------------------------------------------------------------------
struct AdjIn {
  Dims d;
  const double *u, *v, *w, *lu, *lv, *lw, *rdzf, *rdzc;
  double rdx, rdy, kxx, kyy;
  __device__ __forceinline__ double U(int i, int j, int k) const { return u[(long long)i * d.sxu + (long long)j * d.syu + k]; }
  __device__ __forceinline__ double V(int i, int j, int k) const { return v[(long long)i * d.sxv + (long long)j * d.syv + k]; }
  __device__ __forceinline__ double W(int i, int j, int k) const { return w[(long long)i * d.sxw + (long long)j * d.syw + k]; }
  __device__ __forceinline__ double LU(int i, int j, int k) const {
    return (i >= 1 && i <= d.nx - 1 && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz)
               ? lu[(long long)i * d.sxu + (long long)j * d.syu + k] : 0.0;
  }
  __device__ __forceinline__ double LV(int i, int j, int k) const {
    return (i >= 1 && i <= d.nx && j >= 1 && j <= d.ny - 1 && k >= 1 && k <= d.nz)
               ? lv[(long long)i * d.sxv + (long long)j * d.syv + k] : 0.0;
  }
  __device__ __forceinline__ double LW(int i, int j, int k) const {
    return (i >= 1 && i <= d.nx && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz - 1)
               ? lw[(long long)i * d.sxw + (long long)j * d.syw + k] : 0.0;
  }
  __device__ __forceinline__ double RZF(int k) const { return (k >= 0 && k < d.nz) ? rdzf[k] : 0.0; }
  __device__ __forceinline__ double RZC(int k) const { return (k >= 0 && k <= d.nz) ? rdzc[k] : 0.0; }
};

__device__ double adj_u_at(const AdjIn& a, int I, int J, int K) {
  const Dims& d = a.d;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double g = a.kxx * (a.LU(I + 1, J, K) + a.LU(I - 1, J, K) - 2.0 * a.LU(I, J, K)) +
             a.kyy * (a.LU(I, J + 1, K) + a.LU(I, J - 1, K) - 2.0 * a.LU(I, J, K));
  // u momentum, x flux ((u_i + u_{i+1})/2)^2
  for (int i = I - 1; i <= I; ++i)
    if (i >= 0 && i <= nx - 1) {
      const double mu = a.rdx * (a.LU(i + 1, J, K) - a.LU(i, J, K));
      g += 0.5 * (a.U(i, J, K) + a.U(i + 1, J, K)) * mu;
    }
  // u momentum, y flux: v interpolated to the x face times u interpolated to the y face
  if (I >= 1 && I <= nx - 1)
    for (int j = J - 1; j <= J; ++j)
      if (j >= 0 && j <= ny) {
        const double mu = a.rdy * (a.LU(I, j + 1, K) - a.LU(I, j, K));
        g += 0.25 * (a.V(I, j, K) + a.V(I + 1, j, K)) * mu;
      }
  // u momentum, z flux
  if (I >= 1 && I <= nx - 1 && J >= 1 && J <= ny)
    for (int k = K - 1; k <= K; ++k) {
      const double mu = a.RZF(k) * a.LU(I, J, k + 1) - a.RZF(k - 1) * a.LU(I, J, k);
      g += 0.25 * (a.W(I, J, k) + a.W(I + 1, J, k)) * mu;
    }
  // v momentum, x flux: u interpolated to the y face times v interpolated to the x face
  for (int j = J - 1; j <= J; ++j)
    if (j >= 1 && j <= ny - 1 && I >= 0 && I <= nx) {
      const double mu = a.rdx * (a.LV(I + 1, j, K) - a.LV(I, j, K));
      g += 0.25 * (a.V(I, j, K) + a.V(I + 1, j, K)) * mu;
    }
  // w momentum, x flux: u interpolated to the z face times w interpolated to the x face
  for (int k = K - 1; k <= K; ++k)
    if (k >= 1 && k <= nz - 1 && J >= 1 && J <= ny && I >= 0 && I <= nx) {
      const double mu = a.rdx * (a.LW(I + 1, J, k) - a.LW(I, J, k));
      g += 0.25 * (a.W(I, J, k) + a.W(I + 1, J, k)) * mu;
    }
  return g;
}

__device__ double adj_v_at(const AdjIn& a, int I, int J, int K) {
  const Dims& d = a.d;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double g = a.kxx * (a.LV(I + 1, J, K) + a.LV(I - 1, J, K) - 2.0 * a.LV(I, J, K)) +
             a.kyy * (a.LV(I, J + 1, K) + a.LV(I, J - 1, K) - 2.0 * a.LV(I, J, K));
  // u momentum, y flux
  for (int i = I - 1; i <= I; ++i)
    if (i >= 1 && i <= nx - 1 && J >= 0 && J <= ny) {
      const double mu = a.rdy * (a.LU(i, J + 1, K) - a.LU(i, J, K));
      g += 0.25 * (a.U(i, J, K) + a.U(i, J + 1, K)) * mu;
    }
  // v momentum, x flux
  for (int i = I - 1; i <= I; ++i)
    if (i >= 0 && i <= nx && J >= 1 && J <= ny - 1) {
      const double mu = a.rdx * (a.LV(i + 1, J, K) - a.LV(i, J, K));
      g += 0.25 * (a.U(i, J, K) + a.U(i, J + 1, K)) * mu;
    }
  // v momentum, y flux ((v_j + v_{j+1})/2)^2
  if (I >= 1 && I <= nx)
    for (int j = J - 1; j <= J; ++j)
      if (j >= 0 && j <= ny - 1) {
        const double mu = a.rdy * (a.LV(I, j + 1, K) - a.LV(I, j, K));
        g += 0.5 * (a.V(I, j, K) + a.V(I, j + 1, K)) * mu;
      }
  // v momentum, z flux
  if (I >= 1 && I <= nx && J >= 1 && J <= ny - 1)
    for (int k = K - 1; k <= K; ++k) {
      const double mu = a.RZF(k) * a.LV(I, J, k + 1) - a.RZF(k - 1) * a.LV(I, J, k);
      g += 0.25 * (a.W(I, J, k) + a.W(I, J + 1, k)) * mu;
    }
  // w momentum, y flux
  for (int k = K - 1; k <= K; ++k)
    if (k >= 1 && k <= nz - 1 && I >= 1 && I <= nx && J >= 0 && J <= ny) {
      const double mu = a.rdy * (a.LW(I, J + 1, k) - a.LW(I, J, k));
      g += 0.25 * (a.W(I, J, k) + a.W(I, J + 1, k)) * mu;
    }
  return g;
}

__device__ double adj_w_at(const AdjIn& a, int I, int J, int K) {
  const Dims& d = a.d;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double g = a.kxx * (a.LW(I + 1, J, K) + a.LW(I - 1, J, K) - 2.0 * a.LW(I, J, K)) +
             a.kyy * (a.LW(I, J + 1, K) + a.LW(I, J - 1, K) - 2.0 * a.LW(I, J, K));
  // u momentum, z flux
  if (J >= 1 && J <= ny)
    for (int i = I - 1; i <= I; ++i)
      if (i >= 1 && i <= nx - 1) {
        const double mu = a.RZF(K) * a.LU(i, J, K + 1) - a.RZF(K - 1) * a.LU(i, J, K);
        g += 0.25 * (a.U(i, J, K) + a.U(i, J, K + 1)) * mu;
      }
  // v momentum, z flux
  if (I >= 1 && I <= nx)
    for (int j = J - 1; j <= J; ++j)
      if (j >= 1 && j <= ny - 1) {
        const double mu = a.RZF(K) * a.LV(I, j, K + 1) - a.RZF(K - 1) * a.LV(I, j, K);
        g += 0.25 * (a.V(I, j, K) + a.V(I, j, K + 1)) * mu;
      }
  // w momentum, x flux
  if (J >= 1 && J <= ny)
    for (int i = I - 1; i <= I; ++i)
      if (i >= 0 && i <= nx) {
        const double mu = a.rdx * (a.LW(i + 1, J, K) - a.LW(i, J, K));
        g += 0.25 * (a.U(i, J, K) + a.U(i, J, K + 1)) * mu;
      }
  // w momentum, y flux
  if (I >= 1 && I <= nx)
    for (int j = J - 1; j <= J; ++j)
      if (j >= 0 && j <= ny) {
        const double mu = a.rdy * (a.LW(I, j + 1, K) - a.LW(I, j, K));
        g += 0.25 * (a.V(I, j, K) + a.V(I, j, K + 1)) * mu;
      }
  // w momentum, z flux ((w_k + w_{k+1})/2)^2 / dz_c
  if (I >= 1 && I <= nx && J >= 1 && J <= ny)
    for (int k = K - 1; k <= K; ++k)
      if (k >= 0 && k <= nz - 1) {
        const double mu = a.RZC(k + 1) * a.LW(I, J, k + 1) - a.RZC(k) * a.LW(I, J, k);
        g += 0.5 * (a.W(I, J, k) + a.W(I, J, k + 1)) * mu;
      }
  return g;
}

__global__ __launch_bounds__(256) void k_adj_rhs(AdjIn a, double* __restrict__ eu, double* __restrict__ ev,
                                                 double* __restrict__ ew) {
  const Dims& d = a.d;
  const int comp = blockIdx.y;
  const int ni = comp == 0 ? d.nx + 1 : d.nx + 2;
  const int nj = comp == 1 ? d.ny + 1 : d.ny + 2;
  const int nk = comp == 2 ? d.nz - 1 : d.nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int K = 1 + (int)(t % nk);
  const long long c = t / nk;
  const int J = (int)(c % nj), I = (int)(c / nj);
  if (comp == 0) eu[(long long)I * d.sxu + (long long)J * d.syu + K] = adj_u_at(a, I, J, K);
  else if (comp == 1) ev[(long long)I * d.sxv + (long long)J * d.syv + K] = adj_v_at(a, I, J, K);
  else ew[(long long)I * d.sxw + (long long)J * d.syw + K] = adj_w_at(a, I, J, K);
}

// ---------------------------------------------------------------------------------------------
// A6: lq_n = lq* + Img^T(E'^T lR) (+ dJ/du_n).  Interior entry (I, J) collects the entries of
// which it is the periodic source (apply_bc_all, solver.py:43-79): u face nx also feeds face 0,
// u/w row ny feeds row 0 and row 1 feeds row ny+1, v/w face 1 feeds face nx+1, v row ny feeds
// row 0.  The objective tau_w = c (mean u[:,:,1] + mean u[:,:,nz]) (means over whole planes,
// ghosts included, utils.py:187-204) contributes c/((nx+1)(ny+2)) per image.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_fold(Dims d, const double* __restrict__ qs_u,
                                                  const double* __restrict__ qs_v, const double* __restrict__ qs_w,
                                                  const double* __restrict__ eu, const double* __restrict__ ev,
                                                  const double* __restrict__ ew, double obj_coef, int obj_top,
                                                  double* __restrict__ lu, double* __restrict__ lv,
                                                  double* __restrict__ lw) {
  const int comp = blockIdx.y;
  const int nzc = comp == 2 ? d.nz - 1 : d.nz;
  const long long tot = (long long)d.nx * d.ny * nzc;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= tot) return;
  const int k = 1 + (int)(t % nzc);
  const long long c = t / nzc;
  const int J = 1 + (int)(c % d.ny), I = 1 + (int)(c / d.ny);
  int xs[2], ys[2], nxs = 1, nys = 1;
  xs[0] = I;
  ys[0] = J;
  if (comp == 0) {
    if (I == d.nx) xs[nxs++] = 0;
    if (J == d.ny) ys[nys++] = 0;
    else if (J == 1) ys[nys++] = d.ny + 1;
  } else if (comp == 1) {
    if (I == d.nx) xs[nxs++] = 0;
    else if (I == 1) xs[nxs++] = d.nx + 1;
    if (J == d.ny) ys[nys++] = 0;
  } else {
    if (I == d.nx) xs[nxs++] = 0;
    else if (I == 1) xs[nxs++] = d.nx + 1;
    if (J == d.ny) ys[nys++] = 0;
    else if (J == 1) ys[nys++] = d.ny + 1;
  }
  int sx, sy;
  const double *q, *e;
  double* l;
  if (comp == 0) { sx = d.sxu; sy = d.syu; q = qs_u; e = eu; l = lu; }
  else if (comp == 1) { sx = d.sxv; sy = d.syv; q = qs_v; e = ev; l = lv; }
  else { sx = d.sxw; sy = d.syw; q = qs_w; e = ew; l = lw; }
  const long long o = (long long)I * sx + (long long)J * sy + k;
  double s = 0.0;
  for (int a = 0; a < nxs; ++a)
    for (int b = 0; b < nys; ++b) s += e[(long long)xs[a] * sx + (long long)ys[b] * sy + k];
  double val = q[o] + s;
  if (comp == 0 && obj_coef != 0.0 && (k == 1 || (obj_top && k == d.nz))) val += obj_coef * (double)(nxs * nys);
  l[o] = val;
}

// ---------------------------------------------------------------------------------------------
// Objective: scale * (sum of u over the k=1 plane [+ the k=nz plane]) added into out[0]
// (partials + one-block fold).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_tauw_part(Dims d, const double* __restrict__ u, int top,
                                                       double* __restrict__ partials) {
  __shared__ double red[4];
  const int np = (d.nx + 1) * (d.ny + 2);
  double s = 0.0;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < np; t += gridDim.x * 256) {
    const int i = t / (d.ny + 2), j = t - i * (d.ny + 2);
    const long long o = (long long)i * d.sxu + (long long)j * d.syu;
    s += u[o + 1];
    if (top) s += u[o + d.nz];
  }
  s = wave_sum(s);
  if (lane_id() == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_adj_tauw_fold(const double* __restrict__ partials, int nb, double scale,
                                                       double* __restrict__ out) {
  __shared__ double acc[256];
  double s = 0.0;
  for (int q = threadIdx.x; q < nb; q += 256) s += partials[q];
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) acc[threadIdx.x] += acc[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += acc[0] * scale;
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
static inline dim3 grid3(long long n) { return dim3((unsigned)((n + 255) / 256), 3); }

int adj_launch_pre(const Grid* g, double rdt, double inv_volume, const double* ascal, void* const* lq,
                     void* const* ys, hipStream_t st) {
  const Dims& d = g->d;
  hipLaunchKernelGGL(k_adj_pre, grid3((long long)d.nx * d.ny * d.nz), dim3(256), 0, st, d, g->dzf64,
                     g->dzc64, g->wbulk, inv_volume, rdt, ascal, (const double*)lq[0], (const double*)lq[1],
                     (const double*)lq[2], (double*)ys[0], (double*)ys[1], (double*)ys[2]);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int adj_cn_blocks(const Grid* g) { return (3 * g->d.nx * g->d.ny + 3) / 4; }

int adj_launch_cn(const Grid* g, void* const* z, const double* tab_cell, const double* tab_face, double dt,
                  double c1, double c0, void* const* qs, void* const* lr, void* const* acc, void* grad,
                  double* partials, hipStream_t st) {
  const Dims& d = g->d;
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
  const Metric<double> mt = g->metric<double>();
  const int blocks = adj_cn_blocks(g);
#define TC_ACN(RR)                                                                                       \
  hipLaunchKernelGGL((k_adj_cn<RR>), dim3(blocks), dim3(256), 0, st, d, mt, (const double*)z[0],         \
                     (const double*)z[1], (const double*)z[2], tab_cell, tab_face, dt, 0.5, c1, c0,       \
                     (double*)qs[0], (double*)qs[1], (double*)qs[2], (double*)lr[0], (double*)lr[1],       \
                     (double*)lr[2], (double*)acc[0], (double*)acc[1], (double*)acc[2], (double*)grad,     \
                     partials)
  switch (R) {
    case 1: TC_ACN(1); break;
    case 2: TC_ACN(2); break;
    case 3: TC_ACN(3); break;
    case 4: TC_ACN(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_ACN
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int adj_launch_lamf(const double* partials, int nb, double* ascal, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_lamf, dim3(1), dim3(1024), 0, st, partials, nb, ascal);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int adj_launch_rhs(const Grid* g, const void* u, const void* v, const void* w, void* const* lr, void* const* e,
                   hipStream_t st) {
  const Dims& d = g->d;
  AdjIn a;
  a.d = d;
  a.u = (const double*)u; a.v = (const double*)v; a.w = (const double*)w;
  a.lu = (const double*)lr[0]; a.lv = (const double*)lr[1]; a.lw = (const double*)lr[2];
  a.rdzf = g->rdzf64; a.rdzc = g->rdzc64;
  a.rdx = 1.0 / g->dx; a.rdy = 1.0 / g->dy;
  a.kxx = g->nu * (1.0 / (g->dx * g->dx)); a.kyy = g->nu * (1.0 / (g->dy * g->dy));
  const long long n = (long long)(d.nx + 2) * (d.ny + 2) * d.nz;
  hipLaunchKernelGGL(k_adj_rhs, grid3(n), dim3(256), 0, st, a, (double*)e[0], (double*)e[1], (double*)e[2]);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int adj_launch_fold(const Grid* g, void* const* qs, void* const* e, double obj_coef, int obj_top, void* const* lq,
                    hipStream_t st) {
  const Dims& d = g->d;
  hipLaunchKernelGGL(k_adj_fold, grid3((long long)d.nx * d.ny * d.nz), dim3(256), 0, st, d,
                     (const double*)qs[0], (const double*)qs[1], (const double*)qs[2], (const double*)e[0],
                     (const double*)e[1], (const double*)e[2], obj_coef, obj_top, (double*)lq[0],
                     (double*)lq[1], (double*)lq[2]);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int adj_tauw_blocks(const Grid* g) {
  const int np = (g->d.nx + 1) * (g->d.ny + 2);
  return std::min(256, (np + 255) / 256);
}

int adj_launch_tauw(const Grid* g, const void* u, int top, double scale, double* partials, double* out,
                    hipStream_t st) {
  const int nb = adj_tauw_blocks(g);
  hipLaunchKernelGGL(k_adj_tauw_part, dim3(nb), dim3(256), 0, st, g->d, (const double*)u, top, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_adj_tauw_fold, dim3(1), dim3(256), 0, st, partials, nb, scale, out);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

}  // namespace tc
