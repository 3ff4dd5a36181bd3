// Stand-alone operators of the IMEX path and the device-side reductions: boundary conditions,
// divergence, projection, bulk velocity, CFL rate (with the adaptive-dt rule), wall shear and the
// turbulence-statistics accumulation.  Reductions are two-pass: blocks write per-block partials
// and a one-block kernel folds them in a fixed order, so results do not depend on scheduling.
#include "tc_internal.h"
#include "tc_poisson.h"

namespace tc {

__device__ __forceinline__ double block_reduce(double v, bool is_max, double* sh) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  v = is_max ? wave_max(v) : wave_sum(v);
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    r = sh[0];
    for (int q = 1; q < nw; ++q) r = is_max ? fmax(r, sh[q]) : r + sh[q];
  }
  __syncthreads();
  return r;
}

// Folds nb per-block partials (sum or max) in a fixed order inside one block.
__device__ __forceinline__ double fold_partials(const double* __restrict__ partials, int nb, int stride,
                                                int off, bool is_max, double* sh) {
  double s = is_max ? -1.0e308 : 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    const double x = partials[(size_t)t * stride + off];
    s = is_max ? fmax(s, x) : s + x;
  }
  return block_reduce(s, is_max, sh);
}

#define TC_FIN_BULK 0
#define TC_FIN_CFL 1
#define TC_FIN_UTAU 2

// Second pass of the stand-alone reductions: one block folds the per-block partials written by
// k_bulk / k_cfl / k_u_tau and writes the result (and, for the CFL rate with update_dt, the
// adaptive-dt rule of solver.py:483-494 and 811-812).
__global__ __launch_bounds__(256) void k_finalize(int kind, const double* __restrict__ partials, int nb,
                                                  double* __restrict__ out, double* __restrict__ scal,
                                                  double a0, double a1, double a2, int flag) {
  __shared__ double sh[4];
  if (kind == TC_FIN_BULK) {
    const double tot = fold_partials(partials, nb, 1, 0, false, sh);
    if (threadIdx.x == 0) *out = tot * a0;
  } else if (kind == TC_FIN_CFL) {
    const double rate = fold_partials(partials, nb, 1, 0, true, sh);
    if (threadIdx.x == 0) {
      *out = rate;
      if (flag) {
        const double dti = rate < 1e-10 ? 1.0 : rate;
        double dt_new = a0 * (1.0 / dti);
        dt_new = fmin(fmax(dt_new, a1), a2);
        const double dt = scal[TC_S_DT];
        if (fabs(dt_new - dt) / dt > 0.05) scal[TC_S_DT] = dt_new;
      }
    }
  } else {
    const double tb = fold_partials(partials, nb, 2, 0, false, sh);
    const double tt = fold_partials(partials, nb, 2, 1, false, sh);
    if (threadIdx.x == 0) {
      // a0 = nu, a1 = dist, a2 = plane size; flag = free-slip top (bottom wall only)
      const double ub = sqrt(fabs(a0 * (tb / a2) / a1));
      *out = flag ? ub : 0.5 * (ub + sqrt(fabs(a0 * (tt / a2) / a1)));
    }
  }
}

// ---------------------------------------------------------------------------------------------
// apply_bc_all (solver.py:23-79).  After the reference's x, y, z sequence every ghost entry holds
// the periodic image / wall reflection of an interior entry; each wavefront fills one column
// from its source column.  Interior columns only receive their wall-normal ghosts.
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_apply_bc(Dims d, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = lane_id();
  const int comp = g % 3, col = g / 3;
  const int Y = d.ny + 2;
  const int i = col / Y, j = col % Y;
  if (i >= d.nx + 2) return;
  if (comp == 0 && i > d.nx) return;
  if (comp == 1 && j > d.ny) return;
  const int si = (i == 0) ? d.nx : (i == d.nx + 1 ? 1 : i);
  const int sj = (j == 0) ? d.ny : (j == d.ny + 1 ? 1 : j);
  const bool ghost_col = si != i || sj != j;
  if (comp < 2) {
    T* q = comp == 0 ? u : v;
    const int sx = comp == 0 ? d.sxu : d.sxv, sy = comp == 0 ? d.syu : d.syv;
    const int dst = i * sx + j * sy, src = si * sx + sj * sy;
    for (int k = 1 + lane; k <= d.nz; k += TC_WAVE) {
      const T val = q[src + k];
      if (ghost_col) q[dst + k] = val;
      if (k == 1) q[dst] = -val;
      if (k == d.nz) q[dst + d.nz + 1] = d.top > 0 ? val : -val;
    }
  } else {
    const int dst = i * d.sxw + j * d.syw, src = si * d.sxw + sj * d.syw;
    if (ghost_col)
      for (int k = 1 + lane; k < d.nz; k += TC_WAVE) w[dst + k] = w[src + k];
    if (lane == 0) { w[dst] = T(0); w[dst + d.nz] = T(0); }
  }
}

// compute_divergence (utils.py:235-240), reading the arrays as stored.
template <typename T>
__global__ __launch_bounds__(256) void k_divergence(Dims d, Metric<T> mt, const T* __restrict__ u,
                                                    const T* __restrict__ v, const T* __restrict__ w,
                                                    T* __restrict__ div) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long N = (long long)d.nx * d.ny * d.nz;
  if (t >= N) return;
  const int k = (int)(t % d.nz);
  const long long ij = t / d.nz;
  const int j = (int)(ij % d.ny), i = (int)(ij / d.ny);
  const int cu = (i + 1) * d.sxu + (j + 1) * d.syu + k + 1;
  const int cv = (i + 1) * d.sxv + (j + 1) * d.syv + k + 1;
  const int cw = (i + 1) * d.sxw + (j + 1) * d.syw + k + 1;
  div[t] = (u[cu] - u[cu - d.sxu]) * mt.rdx + (v[cv] - v[cv - d.syv]) * mt.rdy + (w[cw] - w[cw - 1]) * mt.rdzf[k];
}

// compute_momentum_rhs_fused_v2 (operators.py:817-1034): advection plus the full 3-D diffusion
// (the z second derivative explicit too), on new arrays that are zero outside the interiors
// u: i = 1..nx-1, j = 1..ny, k = 1..nz; v: i = 1..nx, j = 1..ny-1, k = 1..nz;
// w: i = 1..nx, j = 1..ny, k = 1..nz-1.  One thread per output entry, the three arrays in turn.
// The w z-terms divide by dz_f / dz_c as the reference does, with its control volume dz_c[k-1]
// for the w second derivative (operators.py:1021-1026).
template <typename T>
__global__ __launch_bounds__(256) void k_rhs_v2(Dims d, Metric<T> mt, const T* __restrict__ u,
                                                const T* __restrict__ v, const T* __restrict__ w,
                                                T* __restrict__ ru, T* __restrict__ rv, T* __restrict__ rw) {
  const long long nu_ = (long long)(d.nx + 1) * d.sxu, nv_ = (long long)(d.nx + 2) * d.sxv;
  const long long nw_ = (long long)(d.nx + 2) * d.sxw;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const T h = T(0.5), nu = (T)mt.nu;
  if (t < nu_) {
    const int i = (int)(t / d.sxu), r = (int)(t - (long long)i * d.sxu), j = r / d.syu, k = r - j * d.syu;
    T val = T(0);
    if (i >= 1 && i <= d.nx - 1 && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz) {
      const int c = (int)t, X = d.sxu, Y = d.syu;
      const int cv = i * d.sxv + j * d.syv + k, cw = i * d.sxw + j * d.syw + k;
      const T ur = h * (u[c] + u[c + X]), ul = h * (u[c - X] + u[c]);
      const T duudx = (ur * ur - ul * ul) * mt.rdx;
      const T vt = h * (v[cv] + v[cv + d.sxv]), ut = h * (u[c] + u[c + Y]);
      const T vb = h * (v[cv - d.syv] + v[cv + d.sxv - d.syv]), ub = h * (u[c - Y] + u[c]);
      const T dvudy = (vt * ut - vb * ub) * mt.rdy;
      const T wt = h * (w[cw] + w[cw + d.sxw]), uzt = h * (u[c] + u[c + 1]);
      const T wb = h * (w[cw - 1] + w[cw + d.sxw - 1]), uzb = h * (u[c - 1] + u[c]);
      const T dwudz = (wt * uzt - wb * uzb) * mt.rdzf[k - 1];
      const T adv = duudx + dvudy + dwudz;
      const T dxx = (u[c + X] - T(2) * u[c] + u[c - X]) * mt.kxx;
      const T dyy = (u[c + Y] - T(2) * u[c] + u[c - Y]) * mt.kyy;
      const T dzz = nu * ((u[c + 1] - u[c]) * mt.rdzc[k] - (u[c] - u[c - 1]) * mt.rdzc[k - 1]) * mt.rdzf[k - 1];
      val = dxx + dyy + dzz - adv;
    }
    ru[t] = val;
    return;
  }
  t -= nu_;
  if (t < nv_) {
    const int i = (int)(t / d.sxv), r = (int)(t - (long long)i * d.sxv), j = r / d.syv, k = r - j * d.syv;
    T val = T(0);
    if (i >= 1 && i <= d.nx && j >= 1 && j <= d.ny - 1 && k >= 1 && k <= d.nz) {
      const int c = (int)t, X = d.sxv, Y = d.syv;
      const int cu = i * d.sxu + j * d.syu + k, cw = i * d.sxw + j * d.syw + k;
      const T ur = h * (u[cu] + u[cu + d.syu]), vxr = h * (v[c] + v[c + X]);
      const T ul = h * (u[cu - d.sxu] + u[cu - d.sxu + d.syu]), vxl = h * (v[c - X] + v[c]);
      const T duvdx = (ur * vxr - ul * vxl) * mt.rdx;
      const T vtp = h * (v[c] + v[c + Y]), vbt = h * (v[c - Y] + v[c]);
      const T dvvdy = (vtp * vtp - vbt * vbt) * mt.rdy;
      const T wt = h * (w[cw] + w[cw + d.syw]), vzt = h * (v[c] + v[c + 1]);
      const T wb = h * (w[cw - 1] + w[cw + d.syw - 1]), vzb = h * (v[c - 1] + v[c]);
      const T dwvdz = (wt * vzt - wb * vzb) * mt.rdzf[k - 1];
      const T adv = duvdx + dvvdy + dwvdz;
      const T dxx = (v[c + X] - T(2) * v[c] + v[c - X]) * mt.kxx;
      const T dyy = (v[c + Y] - T(2) * v[c] + v[c - Y]) * mt.kyy;
      const T dzz = nu * ((v[c + 1] - v[c]) * mt.rdzc[k] - (v[c] - v[c - 1]) * mt.rdzc[k - 1]) * mt.rdzf[k - 1];
      val = dxx + dyy + dzz - adv;
    }
    rv[t] = val;
    return;
  }
  t -= nv_;
  if (t < nw_) {
    const int i = (int)(t / d.sxw), r = (int)(t - (long long)i * d.sxw), j = r / d.syw, k = r - j * d.syw;
    T val = T(0);
    if (i >= 1 && i <= d.nx && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz - 1) {
      const int c = (int)t, X = d.sxw, Y = d.syw;
      const int cu = i * d.sxu + j * d.syu + k, cv = i * d.sxv + j * d.syv + k;
      const T uzr = h * (u[cu] + u[cu + 1]), wxr = h * (w[c] + w[c + X]);
      const T uzl = h * (u[cu - d.sxu] + u[cu - d.sxu + 1]), wxl = h * (w[c - X] + w[c]);
      const T duwdx = (uzr * wxr - uzl * wxl) * mt.rdx;
      const T vzt = h * (v[cv] + v[cv + 1]), wyt = h * (w[c] + w[c + Y]);
      const T vzb = h * (v[cv - d.syv] + v[cv - d.syv + 1]), wyb = h * (w[c - Y] + w[c]);
      const T dvwdy = (vzt * wyt - vzb * wyb) * mt.rdy;
      const T wtp = h * (w[c] + w[c + 1]), wbt = h * (w[c - 1] + w[c]);
      const T dwwdz = (wtp * wtp - wbt * wbt) / (T)mt.dzc[k];
      const T adv = duwdx + dvwdy + dwwdz;
      const T dxx = (w[c + X] - T(2) * w[c] + w[c - X]) * mt.kxx;
      const T dyy = (w[c + Y] - T(2) * w[c] + w[c - Y]) * mt.kyy;
      const T dzz = nu * ((w[c + 1] - w[c]) / (T)mt.dzf[k] - (w[c] - w[c - 1]) / (T)mt.dzf[k - 1]) / (T)mt.dzc[k - 1];
      val = dxx + dyy + dzz - adv;
    }
    rw[t] = val;
  }
}

// max |compute_divergence| (utils.py:235-240) over the interior: per-block maxima (the maximum
// does not depend on the order), folded by k_finalize.
template <typename T>
__global__ __launch_bounds__(256) void k_div_max(Dims d, Metric<T> mt, const T* __restrict__ u,
                                                 const T* __restrict__ v, const T* __restrict__ w,
                                                 double* __restrict__ partials) {
  __shared__ double sh[4];
  const long long N = (long long)d.nx * d.ny * d.nz;
  double best = 0.0;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long long)gridDim.x * 256) {
    const int k = (int)(t % d.nz);
    const long long ij = t / d.nz;
    const int j = (int)(ij % d.ny), i = (int)(ij / d.ny);
    const int cu = (i + 1) * d.sxu + (j + 1) * d.syu + k + 1;
    const int cv = (i + 1) * d.sxv + (j + 1) * d.syv + k + 1;
    const int cw = (i + 1) * d.sxw + (j + 1) * d.syw + k + 1;
    const T dv = (u[cu] - u[cu - d.sxu]) * mt.rdx + (v[cv] - v[cv - d.syv]) * mt.rdy + (w[cw] - w[cw - 1]) * mt.rdzf[k];
    best = fmax(best, (double)fabs(dv));
  }
  const double bm = block_reduce(best, true, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = bm;
}

// One row of the device time series: [step, time, u_bulk, u_tau, F, dt, max|div|, CFL rate] from
// the scalar block and the two values the reductions left in rows[1] (u_tau) and rows[2]
// (max|div|); rows[0] counts the rows written, rows beyond `capacity` are dropped.
__global__ void k_diag_row(const double* __restrict__ scal, double* __restrict__ rows, int capacity) {
  if (threadIdx.x != 0) return;
  const int c = (int)rows[0];
  if (c < capacity) {
    double* r = rows + 8 * (c + 1);
    r[0] = scal[TC_S_STEP];
    r[1] = scal[TC_S_TIME];
    r[2] = scal[TC_S_UBULK];
    r[3] = rows[1];
    r[4] = scal[TC_S_F];
    r[5] = scal[TC_S_DT];
    r[6] = rows[2];
    r[7] = scal[TC_S_CFL];
  }
  rows[0] = (double)(c + 1);
}

// project_velocity (projection.py:154-163) on the interior, in place.
template <typename T>
__global__ __launch_bounds__(256) void k_project(Dims d, Metric<T> mt, T* __restrict__ u, T* __restrict__ v,
                                                 T* __restrict__ w, const T* __restrict__ p,
                                                 double dt_host, const double* __restrict__ dt_dev) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long N = (long long)d.nx * d.ny * d.nz;
  if (t >= N) return;
  const T dt = (T)(dt_dev ? *dt_dev : dt_host);
  const int k = (int)(t % d.nz) + 1;
  const long long ij = t / d.nz;
  const int j = (int)(ij % d.ny) + 1, i = (int)(ij / d.ny) + 1;
  const int cp = i * d.sxp + j * d.syp + k;
  const T pc = p[cp];
  u[i * d.sxu + j * d.syu + k] -= dt * ((p[cp + d.sxp] - pc) * mt.rdx);
  v[i * d.sxv + j * d.syv + k] -= dt * ((p[cp + d.syp] - pc) * mt.rdy);
  if (k < d.nz) w[i * d.sxw + j * d.syw + k] -= dt * ((p[cp + 1] - pc) * mt.rdzc[k]);
}

// compute_bulk_velocity (utils.py:220-223) with an arbitrary strided float64 weight view.
template <typename T>
__global__ __launch_bounds__(256) void k_bulk(Dims d, const T* __restrict__ u, const double* __restrict__ wgt,
                                              long long ws0, long long ws1, long long ws2, double inv_vol,
                                              double* __restrict__ partials) {
  __shared__ double sh[4];
  const long long N = (long long)d.nx * d.ny * d.nz;
  double s = 0.0;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long long)gridDim.x * 256) {
    const int k = (int)(t % d.nz);
    const long long ij = t / d.nz;
    const int j = (int)(ij % d.ny), i = (int)(ij / d.ny);
    s += (double)u[(i + 1) * d.sxu + (j + 1) * d.syu + k + 1] * wgt[i * ws0 + j * ws1 + k * ws2];
  }
  const double bs = block_reduce(s, false, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// compute_cfl_fused (operators.py:1262-1331): per-cell convective rates in x, y, z; maximum.
// With update_dt, the fold (k_finalize) also applies compute_cfl_dt (solver.py:483-494) and the 5%
// rule (solver.py:811-812) to the scalar block.  Column pencils: a block holds TC_CFL_WAVES
// consecutive i, one wavefront per i marching j over jb columns, lane rows k = r*64 + lane + 1
// (r < R; rows past nz read row nz and are masked out).  Each step loads only the columns it has
// not seen (u(i, j+1), u(i-1, j+1), v(i, j), v(i+1, j), w(i, j+1), w(i+1, j), with their k+1 or
// k-1 rows) and carries the rest from the previous step in registers, so every column comes from
// HBM about once per pencil.  The maximum does not depend on the order, so the result is the same
// for any decomposition.
#define TC_CFL_WAVES 4
template <typename T, int R>
__global__ __launch_bounds__(64 * TC_CFL_WAVES) void k_cfl(Dims d, Metric<T> mt, const T* __restrict__ u,
                                                           const T* __restrict__ v, const T* __restrict__ w,
                                                           int jb, double* __restrict__ partials) {
  __shared__ double sh[TC_CFL_WAVES];
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nbj = (d.ny + jb - 1) / jb;
  const int bi = blockIdx.x / nbj, bj = blockIdx.x - bi * nbj;
  const int i = 1 + bi * TC_CFL_WAVES + wv;
  const int j0 = 1 + bj * jb, j1 = min(j0 + jb, d.ny + 1);
  const T q = T(0.25);
  double best = 0.0;
  if (i <= d.nx) {
    const T* ui = u + (long long)i * d.sxu;
    const T* um = ui - d.sxu;
    const T* vi = v + (long long)i * d.sxv;
    const T* ve = vi + d.sxv;
    const T* wi = w + (long long)i * d.sxw;
    const T* we = wi + d.sxw;
    int k[R];
    T rzf[R], rzc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      k[r] = min(r * TC_WAVE + lane + 1, d.nz);
      rzf[r] = mt.rdzf[k[r] - 1];
      rzc[r] = mt.rdzc[k[r]];
    }
    // carried: u(i, j), u(i-1, j) with their k+1 rows; v(i, j-1) with its k+1 row, v(i+1, j-1);
    // w(i, j) with its k-1 row
    T uA[R], uA1[R], uB[R], uB1[R], vP[R], vP1[R], vQ[R], wA[R], wA0[R];
    {
      const long long cu = (long long)j0 * d.syu, cv = (long long)(j0 - 1) * d.syv, cw = (long long)j0 * d.syw;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int kk = k[r];
        uA[r] = ui[cu + kk]; uA1[r] = ui[cu + kk + 1];
        uB[r] = um[cu + kk]; uB1[r] = um[cu + kk + 1];
        vP[r] = vi[cv + kk]; vP1[r] = vi[cv + kk + 1]; vQ[r] = ve[cv + kk];
        wA[r] = wi[cw + kk]; wA0[r] = wi[cw + kk - 1];
      }
    }
    for (int j = j0; j < j1; ++j) {
      const long long cu = (long long)(j + 1) * d.syu, cv = (long long)j * d.syv;
      const long long cwn = (long long)(j + 1) * d.syw, cwe = (long long)j * d.syw;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int kk = k[r];
        const T uN = ui[cu + kk], uN1 = ui[cu + kk + 1], umN = um[cu + kk], umN1 = um[cu + kk + 1];
        const T vC = vi[cv + kk], vC1 = vi[cv + kk + 1], vE = ve[cv + kk];
        const T wN = wi[cwn + kk], wN0 = wi[cwn + kk - 1], wE = we[cwe + kk], wE0 = we[cwe + kk - 1];
        const T ex = fabs(uA[r]) * mt.rdx +
                     q * fabs(vC + vP[r] + vE + vQ[r]) * mt.rdy +
                     q * fabs(wA[r] + wA0[r] + wE + wE0) * rzf[r];
        const T ey = q * fabs(uA[r] + uN + umN + uB[r]) * mt.rdx +
                     fabs(vC) * mt.rdy +
                     q * fabs(wA[r] + wN + wN0 + wA0[r]) * rzf[r];
        // rows k = nz read the ghost row nz + 1 of u and v here; their ez is not used
        const T ez = q * fabs(uA[r] + uB[r] + uB1[r] + uA1[r]) * mt.rdx +
                     q * fabs(vC + vP[r] + vP1[r] + vC1) * mt.rdy +
                     fabs(wA[r]) * rzc[r];
        double m = fmax((double)ex, (double)ey);
        if (kk < d.nz) m = fmax(m, (double)ez);
        if (r * TC_WAVE + lane + 1 <= d.nz) best = fmax(best, m);
        uA[r] = uN; uA1[r] = uN1; uB[r] = umN; uB1[r] = umN1;
        vP[r] = vC; vP1[r] = vC1; vQ[r] = vE;
        wA[r] = wN; wA0[r] = wN0;
      }
    }
  }
  const double bm = block_reduce(best, true, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = bm;
}

// compute_u_tau (utils.py:187-204): means over the whole k=1 and k=nz planes of u (ghosts too).
template <typename T>
__global__ __launch_bounds__(256) void k_u_tau(Dims d, const T* __restrict__ u, double* __restrict__ partials) {
  __shared__ double sh[4];
  const int P = (d.nx + 1) * (d.ny + 2);
  double sb = 0.0, st = 0.0;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < P; t += gridDim.x * 256) {
    sb += (double)u[t * d.syu + 1];
    st += (double)u[t * d.syu + d.nz];
  }
  const double b1 = block_reduce(sb, false, sh);
  const double b2 = block_reduce(st, false, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = b1;
    partials[2 * blockIdx.x + 1] = b2;
  }
}

// ---------------------------------------------------------------------------------------------
// Turbulence statistics (turbstats.py:114-228).  Pass A: xy means of the cell-centred u, v, w per
// z.  Pass B: xy means of the fluctuation products, and the fluctuation planes at the two
// sampled heights (top-wall w negated).  Both passes: block b accumulates columns b, b+nb, ...
// with one thread per z row, then k_stats_fold folds the per-block rows in order.
// ---------------------------------------------------------------------------------------------
template <typename T, bool PASS_B>
__global__ __launch_bounds__(256) void k_stats(Dims d, const T* __restrict__ u, const T* __restrict__ v,
                                               const T* __restrict__ w, double* __restrict__ means,
                                               double* __restrict__ prof, double* __restrict__ planes,
                                               int kb, int kt, double* __restrict__ partials) {
  const int k = threadIdx.x;
  const int nz = d.nz, nx = d.nx, ny = d.ny;
  constexpr int NQ = PASS_B ? 4 : 3;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  double U = 0.0, V = 0.0, W = 0.0;
  if (PASS_B && k < nz) { U = means[k]; V = means[nz + k]; W = means[2 * nz + k]; }
  if (k < nz) {
    for (int c = blockIdx.x; c < nx * ny; c += gridDim.x) {
      const int i = c / ny, j = c % ny;
      const double uc = 0.5 * ((double)u[i * d.sxu + (j + 1) * d.syu + k + 1] + (double)u[(i + 1) * d.sxu + (j + 1) * d.syu + k + 1]);
      const double vc = 0.5 * ((double)v[(i + 1) * d.sxv + j * d.syv + k + 1] + (double)v[(i + 1) * d.sxv + (j + 1) * d.syv + k + 1]);
      const double wc = 0.5 * ((double)w[(i + 1) * d.sxw + (j + 1) * d.syw + k] + (double)w[(i + 1) * d.sxw + (j + 1) * d.syw + k + 1]);
      if constexpr (!PASS_B) {
        acc[0] += uc; acc[1] += vc; acc[2] += wc;
      } else {
        const double uf = uc - U, vf = vc - V, wf = wc - W;
        acc[0] += uf * uf; acc[1] += vf * vf; acc[2] += wf * wf; acc[3] += uf * wf;
        if (k == kb - 1) {
          double* pl = planes + ((size_t)i * ny + j) * 6;
          pl[0] = uf; pl[1] = vf; pl[2] = wf;
        }
        if (k == kt - 1) {
          double* pl = planes + ((size_t)i * ny + j) * 6;
          pl[3] = uf; pl[4] = vf; pl[5] = -wf;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) partials[((size_t)blockIdx.x * NQ + q) * nz + k] = acc[q];
  }
}

// Folds the per-block z rows of k_stats in block order (one thread per z row).
template <bool PASS_B>
__global__ __launch_bounds__(256) void k_stats_fold(int nx, int ny, int nz, int nb, const double* __restrict__ partials,
                                                    double* __restrict__ means, double* __restrict__ prof) {
  constexpr int NQ = PASS_B ? 4 : 3;
  const int k = threadIdx.x;
  if (k >= nz) return;
  const double inv = 1.0 / ((double)nx * (double)ny);
  double tot[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) tot[q] = 0.0;
  for (int b = 0; b < nb; ++b)
#pragma unroll
    for (int q = 0; q < NQ; ++q) tot[q] += partials[((size_t)b * NQ + q) * nz + k];
  if constexpr (!PASS_B) {
    means[k] = tot[0] * inv; means[nz + k] = tot[1] * inv; means[2 * nz + k] = tot[2] * inv;
  } else {
    prof[k] += means[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) prof[(q + 1) * nz + k] += tot[q] * inv;
  }
}

// Folded, wall-averaged spectra (turbstats.py:176-225) from the plane spectra (nkx, ny, 6): our
// half spectrum keeps kx >= 0 and all ky, so the reference's E(-kx, ky) is E(kx, ny-ky) here.
__global__ __launch_bounds__(256) void k_spectra(int nx, int ny, const double2* __restrict__ ps,
                                                 double* __restrict__ spec) {
  const int hx = nx / 2, hy = ny / 2;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= hx * hy) return;
  const int a = t / hy, bq = t % hy;
  const int kx = a + 1, ky = bq + 1;
  const double nrm = ((double)nx * (double)ny) * ((double)nx * (double)ny);
  double e[2][4];
  const int jy[2] = {ky, ny - ky};
  for (int s = 0; s < 2; ++s) {
    const double2* z = ps + ((size_t)kx * ny + jy[s]) * 6;
    double eb[4], et[4];
    for (int side = 0; side < 2; ++side) {
      const double2 U = z[3 * side], V = z[3 * side + 1], W = z[3 * side + 2];
      double* o = side == 0 ? eb : et;
      o[0] = (U.x * U.x + U.y * U.y) / nrm;
      o[1] = (V.x * V.x + V.y * V.y) / nrm;
      o[2] = (W.x * W.x + W.y * W.y) / nrm;
      o[3] = (U.x * W.x + U.y * W.y) / nrm;
    }
    for (int q = 0; q < 4; ++q) e[s][q] = 0.5 * (eb[q] + et[q]);
  }
  for (int q = 0; q < 4; ++q) spec[(size_t)q * hx * hy + t] += e[0][q] + e[1][q];
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
int launch_apply_bc(const Grid* gr, void* u, void* v, void* w, hipStream_t st) {
  const Dims& d = gr->d;
  const int blocks = (3 * (d.nx + 2) * (d.ny + 2) + 3) / 4;
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_apply_bc<double>, dim3(blocks), dim3(256), 0, st, d, (double*)u, (double*)v, (double*)w);
  else
    hipLaunchKernelGGL(k_apply_bc<float>, dim3(blocks), dim3(256), 0, st, d, (float*)u, (float*)v, (float*)w);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_divergence(const Grid* gr, const void* u, const void* v, const void* w, void* div, hipStream_t st) {
  const Dims& d = gr->d;
  const long long N = (long long)d.nx * d.ny * d.nz;
  const int blocks = (int)((N + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_divergence<double>, dim3(blocks), dim3(256), 0, st, d, gr->metric<double>(),
                       (const double*)u, (const double*)v, (const double*)w, (double*)div);
  else
    hipLaunchKernelGGL(k_divergence<float>, dim3(blocks), dim3(256), 0, st, d, gr->metric<float>(),
                       (const float*)u, (const float*)v, (const float*)w, (float*)div);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_project(const Grid* gr, void* u, void* v, void* w, const void* p, double dt_host,
                   const double* dt_dev, hipStream_t st) {
  const Dims& d = gr->d;
  const long long N = (long long)d.nx * d.ny * d.nz;
  const int blocks = (int)((N + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_project<double>, dim3(blocks), dim3(256), 0, st, d, gr->metric<double>(),
                       (double*)u, (double*)v, (double*)w, (const double*)p, dt_host, dt_dev);
  else
    hipLaunchKernelGGL(k_project<float>, dim3(blocks), dim3(256), 0, st, d, gr->metric<float>(),
                       (float*)u, (float*)v, (float*)w, (const float*)p, dt_host, dt_dev);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// j columns per CFL pencil: at least 8, and few enough bands that the block count fits nblocks_max
static int cfl_jb(const Grid* gr, int nblocks_max) {
  const int nbi = (gr->d.nx + TC_CFL_WAVES - 1) / TC_CFL_WAVES;
  const int nbj_max = std::max(1, nblocks_max / nbi);
  return std::max(8, (gr->d.ny + nbj_max - 1) / nbj_max);
}

static int cfl_blocks(const Grid* gr, int nblocks_max) {
  const int jb = cfl_jb(gr, nblocks_max);
  return ((gr->d.nx + TC_CFL_WAVES - 1) / TC_CFL_WAVES) * ((gr->d.ny + jb - 1) / jb);
}

int launch_cfl(const Grid* gr, const void* u, const void* v, const void* w, double* scal, double* partials,
               int update_dt, double cfl_target, double dt_min, double dt_max,
               int nblocks_max, hipStream_t st) {
  const int jb = cfl_jb(gr, nblocks_max);
  const int blocks = cfl_blocks(gr, nblocks_max);
  if (blocks > nblocks_max) return TC_ERR_SHAPE;
  const int R = (gr->d.nz + TC_WAVE - 1) / TC_WAVE;
#define TC_CFL(TT, RR)                                                                                  \
  hipLaunchKernelGGL((k_cfl<TT, RR>), dim3(blocks), dim3(64 * TC_CFL_WAVES), 0, st, gr->d, gr->metric<TT>(), \
                     (const TT*)u, (const TT*)v, (const TT*)w, jb, partials)
  if (gr->dtype == TC_F64) {
    switch (R) {
      case 1: TC_CFL(double, 1); break;
      case 2: TC_CFL(double, 2); break;
      case 3: TC_CFL(double, 3); break;
      case 4: TC_CFL(double, 4); break;
      default: return TC_ERR_UNSUPPORTED;
    }
  } else {
    switch (R) {
      case 1: TC_CFL(float, 1); break;
      case 2: TC_CFL(float, 2); break;
      case 3: TC_CFL(float, 3); break;
      case 4: TC_CFL(float, 4); break;
      default: return TC_ERR_UNSUPPORTED;
    }
  }
#undef TC_CFL
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, TC_FIN_CFL, (const double*)partials, blocks,
                     scal + TC_S_CFL, scal, cfl_target, dt_min, dt_max, update_dt);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_bulk(const Grid* gr, const void* u, const double* wgt, long long ws0, long long ws1,
                long long ws2, double inv_vol, double* out, double* partials,
                int nblocks_max, hipStream_t st) {
  const long long N = (long long)gr->d.nx * gr->d.ny * gr->d.nz;
  const int blocks = (int)std::min<long long>(nblocks_max, (N + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_bulk<double>, dim3(blocks), dim3(256), 0, st, gr->d, (const double*)u, wgt, ws0,
                       ws1, ws2, inv_vol, partials);
  else
    hipLaunchKernelGGL(k_bulk<float>, dim3(blocks), dim3(256), 0, st, gr->d, (const float*)u, wgt, ws0, ws1,
                       ws2, inv_vol, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, TC_FIN_BULK, (const double*)partials, blocks, out,
                     (double*)nullptr, inv_vol, 0.0, 0.0, 0);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_rhs_v2(const Grid* gr, const void* u, const void* v, const void* w, void* ru, void* rv, void* rw,
                  hipStream_t st) {
  const Dims& d = gr->d;
  const long long n = (long long)(d.nx + 1) * d.sxu + (long long)(d.nx + 2) * d.sxv + (long long)(d.nx + 2) * d.sxw;
  const int blocks = (int)((n + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_rhs_v2<double>, dim3(blocks), dim3(256), 0, st, d, gr->metric<double>(), (const double*)u,
                       (const double*)v, (const double*)w, (double*)ru, (double*)rv, (double*)rw);
  else
    hipLaunchKernelGGL(k_rhs_v2<float>, dim3(blocks), dim3(256), 0, st, d, gr->metric<float>(), (const float*)u,
                       (const float*)v, (const float*)w, (float*)ru, (float*)rv, (float*)rw);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_div_max(const Grid* gr, const void* u, const void* v, const void* w, double* out, double* partials,
                   int nblocks_max, hipStream_t st) {
  const long long N = (long long)gr->d.nx * gr->d.ny * gr->d.nz;
  const int blocks = (int)std::min<long long>(nblocks_max, (N + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_div_max<double>, dim3(blocks), dim3(256), 0, st, gr->d, gr->metric<double>(),
                       (const double*)u, (const double*)v, (const double*)w, partials);
  else
    hipLaunchKernelGGL(k_div_max<float>, dim3(blocks), dim3(256), 0, st, gr->d, gr->metric<float>(),
                       (const float*)u, (const float*)v, (const float*)w, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, TC_FIN_CFL, (const double*)partials, blocks, out,
                     (double*)nullptr, 0.0, 0.0, 0.0, 0);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_diag_row(const double* scal, double* rows, int capacity, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_row, dim3(1), dim3(64), 0, st, scal, rows, capacity);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_u_tau(const Grid* gr, const void* u, double dist, double* out, double* partials,
                 int nblocks_max, hipStream_t st) {
  const int P = (gr->d.nx + 1) * (gr->d.ny + 2);
  const int blocks = std::min(nblocks_max / 2, (P + 255) / 256);
  if (gr->dtype == TC_F64)
    hipLaunchKernelGGL(k_u_tau<double>, dim3(blocks), dim3(256), 0, st, gr->d, (const double*)u, partials);
  else
    hipLaunchKernelGGL(k_u_tau<float>, dim3(blocks), dim3(256), 0, st, gr->d, (const float*)u, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, TC_FIN_UTAU, (const double*)partials, blocks, out,
                     (double*)nullptr, gr->nu, dist, (double)P, gr->d.top > 0 ? 1 : 0);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
static int stats_t(const Poisson* pl, const void* u, const void* v, const void* w, int kb, int kt,
                   double* prof, double* spec, double* means, double* partials,
                   int nblocks, hipStream_t st) {
  const Grid* gr = pl->g;
  const Dims& d = gr->d;
  hipLaunchKernelGGL((k_stats<T, false>), dim3(nblocks), dim3(256), 0, st, d, (const T*)u, (const T*)v,
                     (const T*)w, means, prof, pl->planes, kb, kt, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_stats_fold<false>, dim3(1), dim3(256), 0, st, d.nx, d.ny, d.nz, nblocks,
                     (const double*)partials, means, prof);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_stats<T, true>), dim3(nblocks), dim3(256), 0, st, d, (const T*)u, (const T*)v,
                     (const T*)w, means, prof, pl->planes, kb, kt, partials);
  TC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_stats_fold<true>, dim3(1), dim3(256), 0, st, d.nx, d.ny, d.nz, nblocks,
                     (const double*)partials, means, prof);
  TC_CHECK_LAUNCH();
  int rc = plane_fft(pl, st);
  if (rc != TC_OK) return rc;
  const int n = (d.nx / 2) * (d.ny / 2);
  hipLaunchKernelGGL(k_spectra, dim3((n + 255) / 256), dim3(256), 0, st, d.nx, d.ny,
                     (const double2*)pl->plane_spec, spec);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_stats(const Poisson* pl, const void* u, const void* v, const void* w, int kb, int kt,
                 double* prof, double* spec, double* means, double* partials,
                 int nblocks, hipStream_t st) {
  if (pl->g->dtype == TC_F64)
    return stats_t<double>(pl, u, v, w, kb, kt, prof, spec, means, partials, nblocks, st);
  return stats_t<float>(pl, u, v, w, kb, kt, prof, spec, means, partials, nblocks, st);
}

}  // namespace tc
