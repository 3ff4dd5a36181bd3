// Explicit stencil + AB2 + Crank-Nicolson kernel, and projection + boundary + bulk kernel.
//
// One 64-lane wavefront owns one wall-normal column of one velocity component; 4 wavefronts
// per 256-thread block.  Column order is (i, j, component) with the component fastest so the
// u, v and w columns at the same (i, j) -- which read each other's neighbours -- run together on
// the same XCD and share its L2.  Rows of a column sit in registers (row m = r*64 + lane) and the
// wall-normal tridiagonal solve is a wavefront-level parallel cyclic reduction with lane shuffles.
#include "tc_internal.h"

namespace tc {

// ---------------------------------------------------------------------------------------------
// Crank-Nicolson coefficient tables (operators.py:581-620 for u/v, 754-781 for w), reduced by
// the PCR recurrence of tridiag.py:40-56 once per dt.  Block 0: cell rows (u, v; n = nz);
// block 1: face rows (w; n = nz-1).  One thread per row, neighbours through LDS.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(TC_MAX_NZ) void k_cn_tables(const double* __restrict__ dzf,
                                                          const double* __restrict__ dzc, int nz,
                                                          int top, double nu, double theta,
                                                          double* __restrict__ scal, double dt_set,
                                                          double* __restrict__ tab_cell,
                                                          double* __restrict__ tab_face) {
  __shared__ double sa[TC_MAX_NZ], sb[TC_MAX_NZ], sc[TC_MAX_NZ];
  const int m = threadIdx.x;
  const bool face = blockIdx.x == 1;
  const int n = face ? nz - 1 : nz;
  double* tab = face ? tab_face : tab_cell;
  const double dt = dt_set > 0.0 ? dt_set : scal[TC_S_DT];
  const double alpha = theta * dt * nu;
  double a = 0.0, b = 1.0, c = 0.0;
  if (m < n) {
    if (!face) {
      const double dl = dzc[m], dr = dzc[m + 1], dc = dzf[m];
      const double lo = -alpha / (dl * dc);
      const double ce = 1.0 + alpha * (1.0 / dl + 1.0 / dr) / dc;
      const double up = -alpha / (dr * dc);
      a = lo; b = ce; c = up;
      if (m == 0) { b = ce - lo; a = 0.0; }
      if (m == n - 1) { b = (top > 0) ? ce + up : ce - up; c = 0.0; }
    } else {
      const double dl = dzf[m], dr = dzf[m + 1], dv = dzc[m + 1];
      a = -alpha / (dl * dv);
      b = 1.0 + alpha * (1.0 / dl + 1.0 / dr) / dv;
      c = -alpha / (dr * dv);
      if (m == 0) a = 0.0;
      if (m == n - 1) c = 0.0;
    }
  }
  for (int st = 0; st < TC_MAX_STAGES; ++st) {
    const int s = 1 << st;
    double al = 0.0, ga = 0.0;
    sa[m] = a; sb[m] = b; sc[m] = c;
    __syncthreads();
    if (s < n && m < n) {
      const double bdn = m - s >= 0 ? sb[m - s] : 1.0, bup = m + s < n ? sb[m + s] : 1.0;
      const double cdn = m - s >= 0 ? sc[m - s] : 0.0, aup = m + s < n ? sa[m + s] : 0.0;
      const double adn = m - s >= 0 ? sa[m - s] : 0.0, cup = m + s < n ? sc[m + s] : 0.0;
      al = -a / bdn;
      ga = -c / bup;
      b = b + al * cdn + ga * aup;
      a = al * adn;
      c = ga * cup;
    }
    __syncthreads();
    tab[(2 * st) * TC_MAX_NZ + m] = al;
    tab[(2 * st + 1) * TC_MAX_NZ + m] = ga;
  }
  tab[2 * TC_MAX_STAGES * TC_MAX_NZ + m] = m < n ? b : 1.0;
  if (dt_set > 0.0 && blockIdx.x == 0 && m == 0) scal[TC_S_DT] = dt_set;
}

// ---------------------------------------------------------------------------------------------
// Explicit right-hand sides, compute_momentum_rhs_fused_imex (operators.py:1078-1223).
// cpu_branch widens the xy-Laplacian to u-face nx / v-face ny with the periodic wrap of
// diffusion_xy_u/v (operators.py:485-495, 516-521).
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T rhs_u_at(const Dims& d, const Metric<T>& mt, const T* __restrict__ u,
                                      const T* __restrict__ v, const T* __restrict__ w, int i, int j,
                                      int k, bool adv) {
  const int c = i * d.sxu + j * d.syu + k;
  const T h = T(0.5);
  const T uc = u[c], uw = u[c - d.sxu];
  const T ue = (i == d.nx) ? u[d.sxu + j * d.syu + k] : u[c + d.sxu];
  const T us = u[c - d.syu], un = u[c + d.syu];
  const T lap = (ue - T(2) * uc + uw) * mt.kxx + (un - T(2) * uc + us) * mt.kyy;
  if (!adv) return lap;
  const T ub = u[c - 1], ut = u[c + 1];
  const int cv = i * d.sxv + j * d.syv + k;
  const T v00 = v[cv], v10 = v[cv + d.sxv], v0m = v[cv - d.syv], v1m = v[cv + d.sxv - d.syv];
  const int cw = i * d.sxw + j * d.syw + k;
  const T w00 = w[cw], w10 = w[cw + d.sxw], w0m = w[cw - 1], w1m = w[cw + d.sxw - 1];
  const T ur = h * (uc + ue), ul = h * (uw + uc);
  const T fx = (ur * ur - ul * ul) * mt.rdx;
  const T fy = (h * (v00 + v10) * (h * (uc + un)) - h * (v0m + v1m) * (h * (us + uc))) * mt.rdy;
  const T fz = (h * (w00 + w10) * (h * (uc + ut)) - h * (w0m + w1m) * (h * (ub + uc))) * mt.rdzf[k - 1];
  return lap - (fx + fy + fz);
}

template <typename T>
__device__ __forceinline__ T rhs_v_at(const Dims& d, const Metric<T>& mt, const T* __restrict__ u,
                                      const T* __restrict__ v, const T* __restrict__ w, int i, int j,
                                      int k, bool adv) {
  const int c = i * d.sxv + j * d.syv + k;
  const T h = T(0.5);
  const T vc = v[c], vw = v[c - d.sxv], ve = v[c + d.sxv];
  const T vs = v[c - d.syv];
  const T vn = (j == d.ny) ? v[i * d.sxv + d.syv + k] : v[c + d.syv];
  const T lap = (ve - T(2) * vc + vw) * mt.kxx + (vn - T(2) * vc + vs) * mt.kyy;
  if (!adv) return lap;
  const T vb = v[c - 1], vt = v[c + 1];
  const int cu = i * d.sxu + j * d.syu + k;
  const T u00 = u[cu], u01 = u[cu + d.syu], um0 = u[cu - d.sxu], um1 = u[cu - d.sxu + d.syu];
  const int cw = i * d.sxw + j * d.syw + k;
  const T w00 = w[cw], w01 = w[cw + d.syw], w0m = w[cw - 1], w1m = w[cw + d.syw - 1];
  const T fx = (h * (u00 + u01) * (h * (vc + ve)) - h * (um0 + um1) * (h * (vw + vc))) * mt.rdx;
  const T vr = h * (vc + vn), vl = h * (vs + vc);
  const T fy = (vr * vr - vl * vl) * mt.rdy;
  const T fz = (h * (w00 + w01) * (h * (vc + vt)) - h * (w0m + w1m) * (h * (vb + vc))) * mt.rdzf[k - 1];
  return lap - (fx + fy + fz);
}

template <typename T>
__device__ __forceinline__ T rhs_w_at(const Dims& d, const Metric<T>& mt, const T* __restrict__ u,
                                      const T* __restrict__ v, const T* __restrict__ w, int i, int j,
                                      int k) {
  const int c = i * d.sxw + j * d.syw + k;
  const T h = T(0.5);
  const T wc = w[c], ww = w[c - d.sxw], we = w[c + d.sxw], ws = w[c - d.syw], wn = w[c + d.syw];
  const T wb = w[c - 1], wt = w[c + 1];
  const int cu = i * d.sxu + j * d.syu + k;
  const T u00 = u[cu], u01 = u[cu + 1], um0 = u[cu - d.sxu], um1 = u[cu - d.sxu + 1];
  const int cv = i * d.sxv + j * d.syv + k;
  const T v00 = v[cv], v01 = v[cv + 1], vm0 = v[cv - d.syv], vm1 = v[cv - d.syv + 1];
  const T lap = (we - T(2) * wc + ww) * mt.kxx + (wn - T(2) * wc + ws) * mt.kyy;
  const T fx = (h * (u00 + u01) * (h * (wc + we)) - h * (um0 + um1) * (h * (ww + wc))) * mt.rdx;
  const T fy = (h * (v00 + v01) * (h * (wc + wn)) - h * (vm0 + vm1) * (h * (ws + wc))) * mt.rdy;
  const T wr = h * (wc + wt), wl = h * (wb + wc);
  const T fz = (wr * wr - wl * wl) * mt.rdzc[k];
  return lap - (fx + fy + fz);
}

// Right-hand side d = q + beta*D2(q) of the CN system for one u/v column (operators.py:627-644):
// rows m = 0..n-1 hold cell k = m+1; the wall closures are the interior stencil with the
// reflected neighbours X[-1] = -X[0] and X[n] = top*X[n-1].
template <int R, typename T>
__device__ __forceinline__ void cn_rhs_cell(T (&x)[R], const Metric<T>& mt, int n, int top, T beta, int lane) {
  T dn[R], up[R];
  shift_dn<1, R, T>(x, dn, T(0), lane);
  shift_up<1, R, T>(x, up, T(0), lane);
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
__device__ __forceinline__ void cn_rhs_face(T (&x)[R], const Metric<T>& mt, int n, T wall, T beta, int lane) {
  T d1[R], d2[R];
  shift_dn<1, R, T>(x, d1, T(0), lane);
  shift_dn<2, R, T>(x, d2, T(0), lane);
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

// ---------------------------------------------------------------------------------------------
// K1: explicit RHS (+ forcing) -> rhs_curr; AB2 -> CN solve -> star (post-CN) velocities.
// mode bits: TC_K1_RHS write rhs arrays; TC_K1_STEP do AB2 + CN; TC_K1_FIRST Euler start;
// TC_K1_FORCE add the scalar forcing scal[TC_S_F] (and body force) to every rhs_u entry.
// ---------------------------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(256) void k_explicit_cn(
    Dims d, Metric<T> mt, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
    const T* __restrict__ pu, const T* __restrict__ pv, const T* __restrict__ pw, T* __restrict__ ru,
    T* __restrict__ rv, T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv, T* __restrict__ sw,
    const double* __restrict__ tab_cell, const double* __restrict__ tab_face,
    const double* __restrict__ scal, const T* __restrict__ body, int mode, int branch, double theta) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b * 4 + (threadIdx.x >> 6);
  const int lane = lane_id();
  int comp, i, j;
  if (!tiled_column(g, d.nx + 2, d.ny + 2, comp, i, j)) return;
  if (comp == 0 && i > d.nx) return;
  if (comp == 1 && j > d.ny) return;

  const bool cpu = branch == TC_BRANCH_CPU;
  const bool inner_ij = i >= 1 && i <= d.nx && j >= 1 && j <= d.ny;
  const double dt = scal[TC_S_DT];
  const T Fu = (comp == 0 && (mode & TC_K1_FORCE)) ? (T)scal[TC_S_F] : T(0);

  T* rout;
  const T *q, *rp;
  T* sout;
  int base, Kc, n;
  bool lap_ok, adv_ok;
  if (comp == 0) {
    base = i * d.sxu + j * d.syu; Kc = d.nz + 2; n = d.nz;
    rout = ru; q = u; rp = pu; sout = su;
    adv_ok = inner_ij && i <= d.nx - 1;
    lap_ok = adv_ok || (cpu && inner_ij);
  } else if (comp == 1) {
    base = i * d.sxv + j * d.syv; Kc = d.nz + 2; n = d.nz;
    rout = rv; q = v; rp = pv; sout = sv;
    adv_ok = inner_ij && j <= d.ny - 1;
    lap_ok = adv_ok || (cpu && inner_ij);
  } else {
    base = i * d.sxw + j * d.syw; Kc = d.nz + 1; n = d.nz - 1;
    rout = rw; q = w; rp = pw; sout = sw;
    adv_ok = inner_ij;
    lap_ok = inner_ij;
  }

  T rhs[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    const int k = m + 1;
    T val = T(0);
    if (m < n && lap_ok) {
      if (comp == 0) val = rhs_u_at<T>(d, mt, u, v, w, i, j, k, adv_ok);
      else if (comp == 1) val = rhs_v_at<T>(d, mt, u, v, w, i, j, k, adv_ok);
      else val = rhs_w_at<T>(d, mt, u, v, w, i, j, k);
    }
    val = val + Fu;
    if (comp == 0 && body != nullptr && (mode & TC_K1_FORCE) && m < n) val = val + body[base + k];
    rhs[r] = val;
    if ((mode & TC_K1_RHS) && m < n) rout[base + k] = val;
  }
  if (mode & TC_K1_RHS) {
    // wall-normal ghost entries of the rhs column carry only the forcing (solver.py:626)
    if (lane == 0) rout[base] = Fu + ((comp == 0 && body && (mode & TC_K1_FORCE)) ? body[base] : T(0));
    if (lane == 1) rout[base + Kc - 1] = Fu + ((comp == 0 && body && (mode & TC_K1_FORCE)) ? body[base + Kc - 1] : T(0));
  }
  if (!(mode & TC_K1_STEP) || !inner_ij) return;

  // AB2 (solver.py:639-646): q* = q + dt*rhs (first step) or q + dt*(1.5 rhs - 0.5 rhs_prev)
  T x[R];
  const bool first = (mode & TC_K1_FIRST) != 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    T val = T(0);
    if (m < n) {
      const T qv = q[base + m + 1];
      const T inc = first ? rhs[r] : (T(1.5) * rhs[r] - T(0.5) * rp[base + m + 1]);
      val = qv + (T)dt * inc;
    }
    x[r] = val;
  }
  const T beta = (T)((1.0 - theta) * dt * mt.nu);
  if (comp < 2) {
    cn_rhs_cell<R, T>(x, mt, n, d.top, beta, lane);
    pcr_shared_solve<R, T>(x, tab_cell, n, lane);
  } else {
    cn_rhs_face<R, T>(x, mt, n, T(0), beta, lane);
    pcr_shared_solve<R, T>(x, tab_face, n, lane);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) sout[base + m + 1] = x[r];
  }
}

// Stand-alone CN solve of one component (solve_implicit_diffusion_u/v/w, operators.py:563-811):
// out's interior columns are replaced by the CN solution of q's columns; out's other entries
// are left untouched (the caller passes a clone of q).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_cn_only(Dims d, Metric<T> mt, int comp, const T* __restrict__ q,
                                                 T* __restrict__ out, const double* __restrict__ tab,
                                                 const double* __restrict__ scal, double theta) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = lane_id();
  const int i = 1 + g / d.ny, j = 1 + g % d.ny;
  if (i > d.nx) return;
  int base, n;
  if (comp == 0) { base = i * d.sxu + j * d.syu; n = d.nz; }
  else if (comp == 1) { base = i * d.sxv + j * d.syv; n = d.nz; }
  else { base = i * d.sxw + j * d.syw; n = d.nz - 1; }
  const double dt = scal[TC_S_DT];
  T x[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    x[r] = m < n ? q[base + m + 1] : T(0);
  }
  const T beta = (T)((1.0 - theta) * dt * mt.nu);
  if (theta < 1.0) {
    if (comp < 2) cn_rhs_cell<R, T>(x, mt, n, d.top, beta, lane);
    else cn_rhs_face<R, T>(x, mt, n, q[base], beta, lane);
  }
  pcr_shared_solve<R, T>(x, tab, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m < n) out[base + m + 1] = x[r];
  }
}

// ---------------------------------------------------------------------------------------------
// K5: q = q* - dt*grad(p) (projection.py:154-163) written with every ghost set to its periodic
// image / wall reflection (apply_bc_all, solver.py:43-79), plus the bulk-velocity partial sums
// (utils.py:220-223).  The last block to finish reduces the partials in a fixed order and
// updates the forcing controller (solver.py:701-708) and the simulated time.
// ---------------------------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(256) void k_project_bc(
    Dims d, Metric<T> mt, const T* __restrict__ su, const T* __restrict__ sv, const T* __restrict__ sw,
    const T* __restrict__ p, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
    double* __restrict__ scal, double* __restrict__ partials, int flags) {
  __shared__ double red[4];
  const int nb = gridDim.x;
  const int b = xcd_remap(blockIdx.x, nb);
  const int wv = threadIdx.x >> 6;
  const int g = b * 4 + wv;
  const int lane = lane_id();
  int comp, i, j;
  const bool in_grid = tiled_column(g, d.nx + 2, d.ny + 2, comp, i, j);
  const T dt = (T)scal[TC_S_DT];
  double part = 0.0;
  bool bad = false;
  bool live = in_grid && !(comp == 0 && i > d.nx) && !(comp == 1 && j > d.ny);
  if (live) {
    const int si = (i == 0) ? d.nx : (i == d.nx + 1 ? 1 : i);
    const int sj = (j == 0) ? d.ny : (j == d.ny + 1 ? 1 : j);
    const int pb = si * d.sxp + sj * d.syp;
    if (comp < 2) {
      const int dst = comp == 0 ? i * d.sxu + j * d.syu : i * d.sxv + j * d.syv;
      const int src = comp == 0 ? si * d.sxu + sj * d.syu : si * d.sxv + sj * d.syv;
      const T* s = comp == 0 ? su : sv;
      T* o = comp == 0 ? u : v;
      const int pstep = comp == 0 ? d.sxp : d.syp;
      const T rd = comp == 0 ? mt.rdx : mt.rdy;
      const bool bulk = comp == 0 && i >= 1 && j >= 1 && j <= d.ny;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = r * TC_WAVE + lane;
        if (m < d.nz) {
          const int k = m + 1;
          const T val = s[src + k] - dt * ((p[pb + pstep + k] - p[pb + k]) * rd);
          bad |= !isfinite(val);
          o[dst + k] = val;
          if (m == 0) o[dst] = -val;
          if (m == d.nz - 1) o[dst + k + 1] = d.top > 0 ? val : -val;
          if (bulk) part += (double)val * mt.wbulk[m];
        }
      }
    } else {
      const int dst = i * d.sxw + j * d.syw, src = si * d.sxw + sj * d.syw;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = r * TC_WAVE + lane;
        if (m < d.nz - 1) {
          const int k = m + 1;
          const T val = sw[src + k] - dt * ((p[pb + k + 1] - p[pb + k]) * mt.rdzc[k]);
          bad |= !isfinite(val);
          w[dst + k] = val;
        }
        if (m == 0) { w[dst] = T(0); w[dst + d.nz] = T(0); }
      }
    }
  }
  // device-side non-finite flag (the reference scans u, v, w, p on the host at save points,
  // solver.py:862-867); every writer stores the same value, so no atomic is needed
  if (__any(bad) && lane == 0) scal[TC_S_NONFINITE] = 1.0;
  if (!(flags & TC_K5_BULK)) return;
  part = wave_sum(part);
  if (lane == 0) red[wv] = part;
  __syncthreads();
  if (threadIdx.x == 0) partials[b] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Folds the per-block bulk partials of k_project_bc in a fixed order (one block), then the
// forcing controller F += 0.1*(U_b - u_b)/dt (solver.py:701-708), time and step count.
__global__ __launch_bounds__(256) void k_step_finalize(const double* __restrict__ partials, int nb,
                                                       double* __restrict__ scal, double inv_volume,
                                                       int flags) {
  __shared__ double acc[256];
  double s = 0.0;
  for (int t = threadIdx.x; t < nb; t += 256) s += partials[t];
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) acc[threadIdx.x] += acc[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double ub = acc[0] * inv_volume;
    const double dtd = scal[TC_S_DT];
    scal[TC_S_UBULK] = ub;
    if (flags & TC_K5_FORCING) {
      scal[TC_S_F] = scal[TC_S_F] + (scal[TC_S_UTARGET] - ub) / dtd * 0.1;
      scal[TC_S_TIME] = scal[TC_S_TIME] + dtd;
      scal[TC_S_STEP] = scal[TC_S_STEP] + 1.0;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
template <typename T>
static int launch_k1_t(const Grid* gr, const void* u, const void* v, const void* w, const void* pu,
                       const void* pv, const void* pw, void* ru, void* rv, void* rw, void* su,
                       void* sv, void* sw, const double* tab_cell, const double* tab_face,
                       const double* scal, const void* body, int mode, int branch, double theta,
                       hipStream_t st) {
  const Dims& d = gr->d;
  const int waves = tiled_waves(d.nx + 2, d.ny + 2);
  const int blocks = (waves + 3) / 4;
  const Metric<T> mt = gr->metric<T>();
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
#define TC_K1(RR)                                                                                 \
  hipLaunchKernelGGL((k_explicit_cn<T, RR>), dim3(blocks), dim3(256), 0, st, d, mt, (const T*)u,   \
                     (const T*)v, (const T*)w, (const T*)pu, (const T*)pv, (const T*)pw, (T*)ru,   \
                     (T*)rv, (T*)rw, (T*)su, (T*)sv, (T*)sw, tab_cell, tab_face, scal,            \
                     (const T*)body, mode, branch, theta)
  switch (R) {
    case 1: TC_K1(1); break;
    case 2: TC_K1(2); break;
    case 3: TC_K1(3); break;
    case 4: TC_K1(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_K1
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_k1(const Grid* gr, const void* u, const void* v, const void* w, const void* pu,
              const void* pv, const void* pw, void* ru, void* rv, void* rw, void* su, void* sv,
              void* sw, const double* tab_cell, const double* tab_face, const double* scal,
              const void* body, int mode, int branch, double theta, hipStream_t st) {
  if (gr->dtype == TC_F64)
    return launch_k1_t<double>(gr, u, v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, tab_cell, tab_face,
                               scal, body, mode, branch, theta, st);
  return launch_k1_t<float>(gr, u, v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, tab_cell, tab_face,
                            scal, body, mode, branch, theta, st);
}

int launch_tables(const Grid* gr, double* scal, double dt_set, double theta, double* tab_cell,
                  double* tab_face, hipStream_t st) {
  hipLaunchKernelGGL(k_cn_tables, dim3(2), dim3(TC_MAX_NZ), 0, st, gr->dzf64, gr->dzc64, gr->d.nz,
                     gr->d.top, gr->nu, theta, scal, dt_set, tab_cell, tab_face);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
static int launch_cn_only_t(const Grid* gr, int comp, const void* q, void* out, const double* tab,
                            const double* scal, double theta, hipStream_t st) {
  const Dims& d = gr->d;
  const int blocks = (d.nx * d.ny + 3) / 4;
  const Metric<T> mt = gr->metric<T>();
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
#define TC_CN(RR)                                                                          \
  hipLaunchKernelGGL((k_cn_only<T, RR>), dim3(blocks), dim3(256), 0, st, d, mt, comp,       \
                     (const T*)q, (T*)out, tab, scal, theta)
  switch (R) {
    case 1: TC_CN(1); break;
    case 2: TC_CN(2); break;
    case 3: TC_CN(3); break;
    case 4: TC_CN(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_CN
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int launch_cn_only(const Grid* gr, int comp, const void* q, void* out, const double* tab,
                   const double* scal, double theta, hipStream_t st) {
  if (gr->dtype == TC_F64) return launch_cn_only_t<double>(gr, comp, q, out, tab, scal, theta, st);
  return launch_cn_only_t<float>(gr, comp, q, out, tab, scal, theta, st);
}

template <typename T>
static int launch_k5_t(const Grid* gr, const void* su, const void* sv, const void* sw, const void* p,
                       void* u, void* v, void* w, double* scal, double* partials, unsigned* counter,
                       double inv_volume, int flags, hipStream_t st) {
  const Dims& d = gr->d;
  const int waves = tiled_waves(d.nx + 2, d.ny + 2);
  const int blocks = (waves + 3) / 4;
  const Metric<T> mt = gr->metric<T>();
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
#define TC_K5(RR)                                                                                \
  hipLaunchKernelGGL((k_project_bc<T, RR>), dim3(blocks), dim3(256), 0, st, d, mt, (const T*)su,  \
                     (const T*)sv, (const T*)sw, (const T*)p, (T*)u, (T*)v, (T*)w, scal, partials, \
                     flags)
  switch (R) {
    case 1: TC_K5(1); break;
    case 2: TC_K5(2); break;
    case 3: TC_K5(3); break;
    case 4: TC_K5(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_K5
  TC_CHECK_LAUNCH();
  if (flags & TC_K5_BULK) {
    hipLaunchKernelGGL(k_step_finalize, dim3(1), dim3(256), 0, st, partials, blocks, scal, inv_volume, flags);
    TC_CHECK_LAUNCH();
  }
  return TC_OK;
}

int k5_blocks(const Grid* gr) { return (tiled_waves(gr->d.nx + 2, gr->d.ny + 2) + 3) / 4; }

int launch_k5(const Grid* gr, const void* su, const void* sv, const void* sw, const void* p, void* u,
              void* v, void* w, double* scal, double* partials, unsigned* counter, double inv_volume,
              int flags, hipStream_t st) {
  if (gr->dtype == TC_F64)
    return launch_k5_t<double>(gr, su, sv, sw, p, u, v, w, scal, partials, counter, inv_volume, flags, st);
  return launch_k5_t<float>(gr, su, sv, sw, p, u, v, w, scal, partials, counter, inv_volume, flags, st);
}

}  // namespace tc
