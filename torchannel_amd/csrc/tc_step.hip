// Explicit stencil + AB2 + Crank-Nicolson kernel, and projection + boundary + bulk kernel.
//
// One 64-lane wavefront owns one wall-normal column of one velocity component; 4 wavefronts
// per 256-thread block.  Column order is (i, j, component) with the component fastest so the
// u, v and w columns at the same (i, j) -- which read each other's neighbours -- run together on
// the same XCD and share its L2.  Rows of a column sit in registers (row m = r*64 + lane) and the
// wall-normal tridiagonal solve is a wavefront-level parallel cyclic reduction with lane shuffles.
#include "tc_ensemble.h"
#include "tc_internal.h"
#include "tc_poisson_dev.h"

namespace tc {

// ---------------------------------------------------------------------------------------------
// Crank-Nicolson coefficient tables (operators.py:581-620 for u/v, 754-781 for w), reduced by
// the PCR recurrence of tridiag.py:40-56 once per dt.  Block 0: cell rows (u, v; n = nz);
// block 1: face rows (w; n = nz-1).  One thread per row, neighbours through LDS.  The last table row
// holds the reciprocal of each row's final diagonal, so every solve ends with a multiplication
// (tridiag.py:58 divides; the two differ by at most one rounding).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void cn_tables_body(const double* __restrict__ dzf, const double* __restrict__ dzc,
                                               int nz, int top, double nu, double theta,
                                               double* __restrict__ scal, double dt_set,
                                               double* __restrict__ tab_cell, double* __restrict__ tab_face,
                                               bool face) {
  __shared__ double sa[TC_MAX_NZ], sb[TC_MAX_NZ], sc[TC_MAX_NZ];
  const int m = threadIdx.x;
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
  tab[2 * TC_MAX_STAGES * TC_MAX_NZ + m] = m < n ? 1.0 / b : 1.0;
  if (dt_set > 0.0 && !face && m == 0) scal[TC_S_DT] = dt_set;
}

__global__ __launch_bounds__(TC_MAX_NZ) void k_cn_tables(const double* __restrict__ dzf,
                                                          const double* __restrict__ dzc, int nz,
                                                          int top, double nu, double theta,
                                                          double* __restrict__ scal, double dt_set,
                                                          double* __restrict__ tab_cell,
                                                          double* __restrict__ tab_face) {
  cn_tables_body(dzf, dzc, nz, top, nu, theta, scal, dt_set, tab_cell, tab_face, blockIdx.x == 1);
}

// Ensemble form: blocks 2c, 2c+1 build channel c's cell and face tables.
template <typename T>
__global__ __launch_bounds__(TC_MAX_NZ) void k_cn_tables_ens(const EnsChan<T>* __restrict__ ch, double dt_set) {
  const EnsChan<T>& c = ch[blockIdx.x >> 1];
  cn_tables_body(c.mt.dzf, c.mt.dzc, c.d.nz, c.d.top, c.nu, c.theta, c.scal, dt_set, c.tabs,
                 c.tabs + TC_TAB_STRIDE, (blockIdx.x & 1) != 0);
}

// ---------------------------------------------------------------------------------------------
// Explicit right-hand sides, compute_momentum_rhs_fused_imex (operators.py:1078-1223).
// Each function receives wavefront-uniform column pointers (scalar registers) and the lane's row
// index k, so every load is a scalar base plus one shared per-lane offset.
// cpu_branch widens the xy-Laplacian to u-face nx / v-face ny with the periodic wrap of
// diffusion_xy_u/v (operators.py:485-495, 516-521): the caller passes the wrapped column.
// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// K1: explicit RHS (+ forcing) -> rhs_curr; AB2 -> CN solve -> star (post-CN) velocities.
// mode bits: TC_K1_RHS write rhs arrays; TC_K1_STEP do AB2 + CN; TC_K1_FIRST Euler start;
// TC_K1_FORCE add the scalar forcing scal[TC_S_F] (and body force) to every rhs_u entry.
//
// Work decomposition: a block of TC_K1_WAVES wavefronts owns one component and a band of
// TC_K1_WAVES consecutive i; each wavefront marches one (component, i) pencil along j over `jb`
// columns, carrying in registers the columns the next j reuses (the j-1 / j neighbours of the
// stencil), so each column of u, v, w is fetched about once per pencil instead of once per
// stencil tap.  The CN reduction tables of the block's component are staged in LDS once per
// block; wall-normal neighbour values (k +- 1, PCR strides < 64) go through a per-wavefront LDS
// row buffer, PCR strides that are multiples of 64 rows are register moves.
// Explicit terms: compute_momentum_rhs_fused_imex (operators.py:1078-1223); the CPU-branch
// extension of the xy-Laplacian to u-face nx / v-face ny with the periodic wrap of
// diffusion_xy_u/v (operators.py:485-495, 516-521); AB2 solver.py:639-646; CN rows
// operators.py:627-644 (u, v) and 788-803 (w, shifted explicit half).
// ---------------------------------------------------------------------------------------------
#define TC_K1_WAVES 4

template <int R, typename T>
__device__ __forceinline__ void col_load(const T* __restrict__ col, T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    x[r] = (int)k <= n ? col[k] : T(0);
  }
}

template <int R, typename T>
__device__ __forceinline__ void col_store(T* __restrict__ col, const T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    if ((int)k <= n) col[k] = x[r];
  }
}

// The AB2 history (rhs) is written by one step's K1 and read by the next step's K1, with every
// other kernel of the step in between, and is dead once read.  Its loads are non-temporal, so the
// lines leave the last-level cache first; in float64 its stores are non-temporal too (the float64
// state cannot stay cached anyway; this keeps q*, p and the velocities there), while the float32
// state fits the 256 MB cache once the dead history leaves it, so its stores stay plain.
// Measured over whole steps at 192x192x128: float64 387 -> 363 us, float32 201 -> 195 us.
template <int R, typename T>
__device__ __forceinline__ void col_load_hist(const T* __restrict__ col, T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    x[r] = (int)k <= n ? __builtin_nontemporal_load(col + k) : T(0);
  }
}

template <int R, typename T>
__device__ __forceinline__ void col_store_hist(T* __restrict__ col, const T (&x)[R], int n, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    if ((int)k <= n) {
      if constexpr (sizeof(T) == 8) __builtin_nontemporal_store(x[r], col + k);
      else col[k] = x[r];
    }
  }
}

// dn = x one row below (row 0 takes `lo`), up = x one row above (row n-1 takes `hi`)
template <int R, typename T>
__device__ __forceinline__ void rows_pm(T* buf, const T (&x)[R], T (&dn)[R], T (&up)[R], T lo, T hi, int n,
                                        int lane) {
  rows_put<R, T>(buf, x, lane);
  rows_get<-1, R, T>(buf, dn, lane);
  rows_get<1, R, T>(buf, up, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    if (m == 0) dn[r] = lo;
    if (m == n - 1) up[r] = hi;
  }
}

template <int R, typename T>
__device__ __forceinline__ void rows_dn(T* buf, const T (&x)[R], T (&dn)[R], T lo, int lane) {
  rows_put<R, T>(buf, x, lane);
  rows_get<-1, R, T>(buf, dn, lane);
  if (lane == 0) dn[0] = lo;
}

// rows_pm without the top override and rows_up without any: for columns whose row n is held by
// the register rows themselves (the caller's row n is the buffer's next row)
template <int R, typename T>
__device__ __forceinline__ void rows_pm_lo(T* buf, const T (&x)[R], T (&dn)[R], T (&up)[R], T lo, int lane) {
  rows_put<R, T>(buf, x, lane);
  rows_get<-1, R, T>(buf, dn, lane);
  rows_get<1, R, T>(buf, up, lane);
  if (lane == 0) dn[0] = lo;
}

template <int R, typename T>
__device__ __forceinline__ void rows_up_in(T* buf, const T (&x)[R], T (&up)[R], int lane) {
  rows_put<R, T>(buf, x, lane);
  rows_get<1, R, T>(buf, up, lane);
}

template <int R, typename T>
__device__ __forceinline__ void rows_up(T* buf, const T (&x)[R], T (&up)[R], T hi, int n, int lane) {
  rows_put<R, T>(buf, x, lane);
  rows_get<1, R, T>(buf, up, lane);
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r * TC_WAVE + lane == n - 1) up[r] = hi;
}

// PCR stages a column of at most 64*R rows can need (stride 1 << ST below the row count), and
// the block's LDS copy of the shared reduction table: k_cn_tables' rows 0 .. 2*NST-1 (alpha,
// gamma per stage) and its reciprocal diagonal as row 2*NST, row stride 64*R.
template <int R>
__host__ __device__ constexpr int k1_nst() { return R == 1 ? 6 : R == 2 ? 7 : 8; }
template <int R>
__host__ __device__ constexpr int k1_tab_len() { return (2 * k1_nst<R>() + 1) * TC_WAVE * R; }
// Table entry e of the LDS layout, read from k_cn_tables' layout.
template <int R>
__device__ __forceinline__ double k1_tab_src(const double* __restrict__ tab, int e) {
  const int row = e / (TC_WAVE * R), m = e - row * (TC_WAVE * R);
  return tab[(row < 2 * k1_nst<R>() ? row : 2 * TC_MAX_STAGES) * TC_MAX_NZ + m];
}

// One PCR stage with the block's LDS copy of the shared reduction table.
template <int ST, int R, typename T>
__device__ __forceinline__ void pcr_stage_tab(T (&dv)[R], T* buf, const T* tabl, int n, int lane) {
  constexpr int s = 1 << ST;
  if constexpr (ST >= k1_nst<R>()) return;
  if (s >= n) return;
  T dn[R], up[R];
  if constexpr (s >= TC_WAVE) {
    constexpr int q = s / TC_WAVE;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      dn[r] = r >= q ? dv[r - q] : T(0);
      up[r] = r + q < R ? dv[r + q] : T(0);
    }
  } else {
    rows_put<R, T>(buf, dv, lane);
    rows_get<-s, R, T>(buf, dn, lane);
    rows_get<s, R, T>(buf, up, lane);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    const T al = tabl[(2 * ST) * TC_WAVE * R + m];
    const T ga = tabl[(2 * ST + 1) * TC_WAVE * R + m];
    dv[r] = dv[r] + al * dn[r] + ga * up[r];
  }
}

template <int R, typename T>
__device__ __forceinline__ void pcr_solve_tab(T (&dv)[R], T* buf, const T* tabl, int n, int lane) {
  pcr_stage_tab<0, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<1, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<2, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<3, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<4, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<5, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<6, R, T>(dv, buf, tabl, n, lane);
  pcr_stage_tab<7, R, T>(dv, buf, tabl, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + lane;
    dv[r] = dv[r] * tabl[2 * k1_nst<R>() * TC_WAVE * R + m];
  }
}

// Per-pencil constants of one wavefront.
template <typename T, int R>
struct K1Ctx {
  Dims d;
  Metric<T> mt;
  T* rows;              // this wavefront's LDS row buffer (3*64*R)
  const T* tabl;        // block's LDS table
  int lane, mode;
  bool cpu;
  T dt, beta, F;
  const T* body;
};

// RHS column write (interior + the two wall-normal ghost rows); returns whether the column
// steps, with the AB2 predictor q + dt*(1.5 rhs - 0.5 rp) (forward Euler on the first step) in x.
template <typename T, int R>
__device__ __forceinline__ bool k1_rhs_ab2(const K1Ctx<T, R>& c, T (&rhs)[R], const T (&q)[R], const T (&rp)[R],
                                           T* rout, const T* bcol, int n, int Kc, bool inner, T (&x)[R]) {
  const int lane = c.lane;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    T val = rhs[r] + c.F;
    if (bcol != nullptr && (int)k <= n) val = val + bcol[k];
    rhs[r] = val;
  }
  if (c.mode & TC_K1_RHS) {
    col_store_hist<R, T>(rout, rhs, n, lane);
    // wall-normal ghost entries of the rhs column carry only the forcing (solver.py:626)
    if (lane == 0) rout[0] = c.F + (bcol != nullptr ? bcol[0] : T(0));
    if (lane == 1) rout[Kc - 1] = c.F + (bcol != nullptr ? bcol[Kc - 1] : T(0));
  }
  if (!(c.mode & TC_K1_STEP) || !inner) return false;
  const bool first = (c.mode & TC_K1_FIRST) != 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    T val = T(0);
    if ((int)k <= n) {
      const T inc = first ? rhs[r] : (T(1.5) * rhs[r] - T(0.5) * rp[r]);
      val = q[r] + c.dt * inc;
    }
    x[r] = val;
  }
  return true;
}

// CN right-hand side d = x + beta*D2(x) with the rows' metric held in registers, the PCR solve
// and the store of the star column:
// cells (u, v): rz0 = 1/dz_c[m], rz1 = 1/dz_c[m+1], rz2 = 1/dz_f[m], reflected neighbours;
// faces (w): rz0 = 1/dz_f[m], rz1 = 1/dz_f[m+1], rz2 = 1/dz_c[m+1], explicit half centred on
// face m with the wall value below row 0 and zero above the top row.
template <typename T, int R, bool FACE>
__device__ __forceinline__ void k1_cn_store(const K1Ctx<T, R>& c, T (&x)[R], T* sout, int n, const T (&rz0)[R],
                                            const T (&rz1)[R], const T (&rz2)[R], T wall) {
  const int lane = c.lane;
  T dn[R], up[R];
  if constexpr (!FACE) {
    rows_put<R, T>(c.rows, x, lane);
    rows_get<-1, R, T>(c.rows, dn, lane);
    rows_get<1, R, T>(c.rows, up, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = r * TC_WAVE + lane;
      if (m == 0) dn[r] = -x[r];
      if (m == n - 1) up[r] = c.d.top > 0 ? x[r] : -x[r];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = r * TC_WAVE + lane;
      if (m < n) {
        const T lap = ((up[r] - x[r]) * rz1[r] - (x[r] - dn[r]) * rz0[r]) * rz2[r];
        x[r] = x[r] + c.beta * lap;
      } else {
        x[r] = T(0);
      }
    }
  } else {
    rows_put<R, T>(c.rows, x, lane);
    rows_get<-1, R, T>(c.rows, dn, lane);
    rows_get<-2, R, T>(c.rows, up, lane);       // two rows below
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = r * TC_WAVE + lane;
      if (m == 0) dn[r] = wall;
      if (m == 1) up[r] = wall;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = r * TC_WAVE + lane;
      if (m < n) {
        const T xp = (m == n - 1) ? T(0) : x[r];
        const T lap = ((xp - dn[r]) * rz1[r] - (dn[r] - up[r]) * rz0[r]) * rz2[r];
        x[r] = x[r] + c.beta * lap;
      } else {
        x[r] = T(0);
      }
    }
  }
  pcr_solve_tab<R, T>(x, c.rows, c.tabl, n, lane);
  col_store<R, T>(sout, x, n, lane);
}

template <typename T, int R, bool FACE>
__device__ __forceinline__ void k1_finish(const K1Ctx<T, R>& c, T (&rhs)[R], const T (&q)[R], const T (&rp)[R],
                                          T* rout, T* sout, const T* bcol, int n, int Kc, bool inner,
                                          const T (&rz0)[R], const T (&rz1)[R], const T (&rz2)[R],
                                          T wall) {
  T x[R];
  if (k1_rhs_ab2<T, R>(c, rhs, q, rp, rout, bcol, n, Kc, inner, x))
    k1_cn_store<T, R, FACE>(c, x, sout, n, rz0, rz1, rz2, wall);
}

// Ghost column of an rhs array: the forcing (and body force) on every entry, no star column.
template <typename T, int R>
__device__ __forceinline__ void k1_ghost_col(const K1Ctx<T, R>& c, T* rout, const T* bcol, int n, int Kc) {
  if (!(c.mode & TC_K1_RHS)) return;
  T val[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + c.lane + 1);
    val[r] = c.F;
    if (bcol != nullptr && (int)k <= n) val[r] = val[r] + bcol[k];
  }
  col_store<R, T>(rout, val, n, c.lane);
  if (c.lane == 0) rout[0] = c.F + (bcol != nullptr ? bcol[0] : T(0));
  if (c.lane == 1) rout[Kc - 1] = c.F + (bcol != nullptr ? bcol[Kc - 1] : T(0));
}

// Wall-normal metric of the rows this lane holds.
template <typename T, int R, bool FACE>
__device__ __forceinline__ void k1_metric(const K1Ctx<T, R>& c, int n, T (&rz0)[R], T (&rz1)[R], T (&rz2)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int m = r * TC_WAVE + c.lane;
    if (!FACE) {
      rz0[r] = m < n ? c.mt.rdzc[m] : T(0);
      rz1[r] = m < n ? c.mt.rdzc[m + 1] : T(0);
      rz2[r] = m < n ? c.mt.rdzf[m] : T(0);
    } else {
      rz0[r] = m < n ? c.mt.rdzf[m] : T(0);
      rz1[r] = m < n ? c.mt.rdzf[m + 1] : T(0);
      rz2[r] = m < n ? c.mt.rdzc[m + 1] : T(0);
    }
  }
}

// Columns of a pencil that carry no stencil (ghost columns, and the GPU-branch u-face nx /
// v-face ny whose explicit terms are zero): rhs = forcing, AB2 + CN on the inner ones.
template <typename T, int R, bool FACE>
__device__ __forceinline__ void k1_plain_cols(const K1Ctx<T, R>& c, const T* qi, const T* rpi, T* ri, T* si,
                                              const T* bi, int sy, int ja, int jb, int jin_lo, int jin_hi, int n,
                                              int Kc, const T (&rz0)[R], const T (&rz1)[R], const T (&rz2)[R]) {
  for (int j = ja; j < jb; ++j) {
    const long long cj = (long long)j * sy;
    const T* bcol = bi != nullptr ? bi + cj : nullptr;
    const bool inner = j >= jin_lo && j <= jin_hi;
    if (!inner || !(c.mode & TC_K1_STEP)) {
      k1_ghost_col<T, R>(c, ri + cj, bcol, n, Kc);
      continue;
    }
    T rhs[R], q[R], rp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rhs[r] = T(0);
    col_load<R, T>(qi + cj, q, n, c.lane);
    if (!(c.mode & TC_K1_FIRST)) col_load<R, T>(rpi + cj, rp, n, c.lane);
    k1_finish<T, R, FACE>(c, rhs, q, rp, ri + cj, si + cj, bcol, n, Kc, true, rz0, rz1, rz2, T(0));
  }
}

// Column inputs of one u stencil evaluation at (i, j), fetched one column ahead.
template <typename T, int R>
struct UIn {
  T uN[R], uW[R], uE[R], v00[R], v10[R], w00[R], w10[R], rp[R];
  T ub0, ubn, wb0, wb1;     // u(i,j) wall-normal ghosts, w(i,j)[0], w(i+1,j)[0]
};

// u pencil: columns (i, j0..j1-1) of u and rhs_u.
template <typename T, int R>
__device__ __forceinline__ void k1_pencil_u(const K1Ctx<T, R>& c, int i, int j0, int j1, const T* __restrict__ u,
                                            const T* __restrict__ v, const T* __restrict__ w,
                                            const T* __restrict__ pu, T* __restrict__ ru, T* __restrict__ su) {
  const Dims& d = c.d;
  const int lane = c.lane, n = d.nz, Kc = d.nz + 2;
  const T h = T(0.5);
  const bool i_in = i >= 1 && i <= d.nx;
  const bool adv = i_in && i <= d.nx - 1;
  const bool lap_ok = adv || (c.cpu && i_in);
  const bool step = (c.mode & TC_K1_STEP) != 0, ab2 = step && !(c.mode & TC_K1_FIRST);
  T rz0[R], rz1[R], rz2[R];
  k1_metric<T, R, false>(c, n, rz0, rz1, rz2);
  const long long oi = (long long)i * d.sxu;
  const T* ui = u + oi;
  const T* bi = c.body != nullptr ? c.body + oi : nullptr;
  const int jlo = max(j0, 1), jhi = min(j1 - 1, d.ny);
  if (!lap_ok || !i_in || jlo > jhi) {
    k1_plain_cols<T, R, false>(c, ui, pu + oi, ru + oi, su + oi, bi, d.syu, j0, j1, i_in ? 1 : 1 << 30,
                               d.ny, n, Kc, rz0, rz1, rz2);
    return;
  }
  k1_plain_cols<T, R, false>(c, ui, pu + oi, ru + oi, su + oi, bi, d.syu, j0, jlo, 1, 0, n, Kc, rz0, rz1, rz2);
  const T* uw = ui - d.sxu;
  const T* ue = (i == d.nx) ? u + d.sxu : ui + d.sxu;        // periodic wrap of the CPU branch
  const T* vi = v + (long long)i * d.sxv;
  const T* wi = w + (long long)i * d.sxw;
  auto fetch = [&](int j, UIn<T, R>& L) {
    const long long cu = (long long)j * d.syu;
    col_load<R, T>(ui + cu + d.syu, L.uN, n, lane);
    col_load<R, T>(uw + cu, L.uW, n, lane);
    col_load<R, T>(ue + cu, L.uE, n, lane);
    if (ab2) col_load_hist<R, T>(pu + oi + cu, L.rp, n, lane);
    if (adv) {
      const T* w0 = wi + (long long)j * d.syw;
      col_load<R, T>(vi + (long long)j * d.syv, L.v00, n, lane);
      col_load<R, T>(vi + d.sxv + (long long)j * d.syv, L.v10, n, lane);
      col_load<R, T>(w0, L.w00, n, lane);
      col_load<R, T>(w0 + d.sxw, L.w10, n, lane);
      L.ub0 = ui[cu];
      L.ubn = ui[cu + n + 1];
      L.wb0 = w0[0];
      L.wb1 = w0[d.sxw];
    }
  };
  T uS[R], uC[R], v0m[R], v1m[R];
  col_load<R, T>(ui + (long long)(jlo - 1) * d.syu, uS, n, lane);
  col_load<R, T>(ui + (long long)jlo * d.syu, uC, n, lane);
  if (adv) {
    col_load<R, T>(vi + (long long)(jlo - 1) * d.syv, v0m, n, lane);
    col_load<R, T>(vi + d.sxv + (long long)(jlo - 1) * d.syv, v1m, n, lane);
  }
  UIn<T, R> A;
  fetch(jlo, A);
  for (int j = jlo; j <= jhi; ++j) {
    const long long cu = (long long)j * d.syu;
    T rhs[R], q[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      rhs[r] = (A.uE[r] - T(2) * uC[r] + A.uW[r]) * c.mt.kxx + (A.uN[r] - T(2) * uC[r] + uS[r]) * c.mt.kyy;
    if (adv) {
      T ub[R], ut[R], w0m[R], w1m[R];
      rows_pm<R, T>(c.rows, uC, ub, ut, A.ub0, A.ubn, n, lane);
      rows_dn<R, T>(c.rows, A.w00, w0m, A.wb0, lane);
      rows_dn<R, T>(c.rows, A.w10, w1m, A.wb1, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T ur = h * (uC[r] + A.uE[r]), ul = h * (A.uW[r] + uC[r]);
        const T fx = (ur * ur - ul * ul) * c.mt.rdx;
        const T fy = (h * (A.v00[r] + A.v10[r]) * (h * (uC[r] + A.uN[r])) - h * (v0m[r] + v1m[r]) * (h * (uS[r] + uC[r]))) * c.mt.rdy;
        const T fz = (h * (A.w00[r] + A.w10[r]) * (h * (uC[r] + ut[r])) - h * (w0m[r] + w1m[r]) * (h * (ub[r] + uC[r]))) * rz2[r];
        rhs[r] = rhs[r] - (fx + fy + fz);
        v0m[r] = A.v00[r];
        v1m[r] = A.v10[r];
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      q[r] = uC[r];
      uS[r] = uC[r];
      uC[r] = A.uN[r];
    }
    // the next column's loads are issued once this column's inputs are dead, so they are in
    // flight during the CN solve
    T x[R];
    const bool cn = k1_rhs_ab2<T, R>(c, rhs, q, A.rp, ru + oi + cu, bi != nullptr ? bi + cu : nullptr, n, Kc,
                                     true, x);
    if (j < jhi) fetch(j + 1, A);
    if (cn) k1_cn_store<T, R, false>(c, x, su + oi + cu, n, rz0, rz1, rz2, T(0));
  }
  k1_plain_cols<T, R, false>(c, ui, pu + oi, ru + oi, su + oi, bi, d.syu, jhi + 1, j1, 1, 0, n, Kc, rz0, rz1, rz2);
}

template <typename T, int R>
struct VIn {
  T vN[R], vW[R], vE[R], u01[R], um1[R], w01[R], rp[R];
  T vb0, vbn, wb1;          // v(i,j) wall-normal ghosts, w(i,j+1)[0]
};

// v pencil: columns (i, j0..j1-1) of v and rhs_v.
template <typename T, int R>
__device__ __forceinline__ void k1_pencil_v(const K1Ctx<T, R>& c, int i, int j0, int j1, const T* __restrict__ u,
                                            const T* __restrict__ v, const T* __restrict__ w,
                                            const T* __restrict__ pv, T* __restrict__ rv, T* __restrict__ sv) {
  const Dims& d = c.d;
  const int lane = c.lane, n = d.nz, Kc = d.nz + 2;
  const T h = T(0.5);
  const bool i_in = i >= 1 && i <= d.nx;
  const bool step = (c.mode & TC_K1_STEP) != 0, ab2 = step && !(c.mode & TC_K1_FIRST);
  T rz0[R], rz1[R], rz2[R];
  k1_metric<T, R, false>(c, n, rz0, rz1, rz2);
  const long long oi = (long long)i * d.sxv;
  const T* vi = v + oi;
  // stencil columns: j = 1..ny-1 (both branches), plus j = ny in the CPU branch
  const int jst = c.cpu ? d.ny : d.ny - 1;
  const int jlo = max(j0, 1), jhi = min(j1 - 1, jst);
  if (!i_in || jlo > jhi) {
    k1_plain_cols<T, R, false>(c, vi, pv + oi, rv + oi, sv + oi, nullptr, d.syv, j0, j1, i_in ? 1 : 1 << 30,
                               d.ny, n, Kc, rz0, rz1, rz2);
    return;
  }
  k1_plain_cols<T, R, false>(c, vi, pv + oi, rv + oi, sv + oi, nullptr, d.syv, j0, jlo, 1, 0, n, Kc, rz0, rz1, rz2);
  const T* ui = u + (long long)i * d.sxu;
  const T* wi = w + (long long)i * d.sxw;
  auto fetch = [&](int j, VIn<T, R>& L) {
    const long long cv = (long long)j * d.syv;
    const T* vn = (j == d.ny) ? vi + d.syv : vi + cv + d.syv;    // periodic wrap of the CPU branch
    col_load<R, T>(vn, L.vN, n, lane);
    col_load<R, T>(vi - d.sxv + cv, L.vW, n, lane);
    col_load<R, T>(vi + d.sxv + cv, L.vE, n, lane);
    if (ab2) col_load_hist<R, T>(pv + oi + cv, L.rp, n, lane);
    if (j < d.ny) {
      const T* w1 = wi + (long long)(j + 1) * d.syw;
      col_load<R, T>(ui + (long long)(j + 1) * d.syu, L.u01, n, lane);
      col_load<R, T>(ui - d.sxu + (long long)(j + 1) * d.syu, L.um1, n, lane);
      col_load<R, T>(w1, L.w01, n, lane);
      L.vb0 = vi[cv];
      L.vbn = vi[cv + n + 1];
      L.wb1 = w1[0];
    }
  };
  T vS[R], vC[R], u00[R], um0[R], w00[R], w0m[R];
  col_load<R, T>(vi + (long long)(jlo - 1) * d.syv, vS, n, lane);
  col_load<R, T>(vi + (long long)jlo * d.syv, vC, n, lane);
  col_load<R, T>(ui + (long long)jlo * d.syu, u00, n, lane);
  col_load<R, T>(ui - d.sxu + (long long)jlo * d.syu, um0, n, lane);
  {
    const T* w0 = wi + (long long)jlo * d.syw;
    col_load<R, T>(w0, w00, n, lane);
    rows_dn<R, T>(c.rows, w00, w0m, w0[0], lane);
  }
  VIn<T, R> A;
  fetch(jlo, A);
  for (int j = jlo; j <= jhi; ++j) {
    const long long cv = (long long)j * d.syv;
    T rhs[R], q[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      rhs[r] = (A.vE[r] - T(2) * vC[r] + A.vW[r]) * c.mt.kxx + (A.vN[r] - T(2) * vC[r] + vS[r]) * c.mt.kyy;
    if (j < d.ny) {
      T vb[R], vt[R], w1m[R];
      rows_pm<R, T>(c.rows, vC, vb, vt, A.vb0, A.vbn, n, lane);
      rows_dn<R, T>(c.rows, A.w01, w1m, A.wb1, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T fx = (h * (u00[r] + A.u01[r]) * (h * (vC[r] + A.vE[r])) - h * (um0[r] + A.um1[r]) * (h * (A.vW[r] + vC[r]))) * c.mt.rdx;
        const T vr = h * (vC[r] + A.vN[r]), vl = h * (vS[r] + vC[r]);
        const T fy = (vr * vr - vl * vl) * c.mt.rdy;
        const T fz = (h * (w00[r] + A.w01[r]) * (h * (vC[r] + vt[r])) - h * (w0m[r] + w1m[r]) * (h * (vb[r] + vC[r]))) * rz2[r];
        rhs[r] = rhs[r] - (fx + fy + fz);
        u00[r] = A.u01[r];
        um0[r] = A.um1[r];
        w00[r] = A.w01[r];
        w0m[r] = w1m[r];
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      q[r] = vC[r];
      vS[r] = vC[r];
      vC[r] = A.vN[r];
    }
    T x[R];
    const bool cn = k1_rhs_ab2<T, R>(c, rhs, q, A.rp, rv + oi + cv, nullptr, n, Kc, true, x);
    if (j < jhi) fetch(j + 1, A);
    if (cn) k1_cn_store<T, R, false>(c, x, sv + oi + cv, n, rz0, rz1, rz2, T(0));
  }
  k1_plain_cols<T, R, false>(c, vi, pv + oi, rv + oi, sv + oi, nullptr, d.syv, jhi + 1, j1, 1, d.ny, n, Kc, rz0,
                             rz1, rz2);
}

template <typename T, int R>
struct WIn {
  T wN[R], wW[R], wE[R], u00[R], um0[R], v00[R], rp[R];
  T wb0, wbn, ubn, umbn, vbn;   // w(i,j) walls, top rows of u(i,j), u(i-1,j), v(i,j)
};

// w pencil: columns (i, j0..j1-1) of w and rhs_w (rows k = 1..nz-1).
template <typename T, int R>
__device__ __forceinline__ void k1_pencil_w(const K1Ctx<T, R>& c, int i, int j0, int j1, const T* __restrict__ u,
                                            const T* __restrict__ v, const T* __restrict__ w,
                                            const T* __restrict__ pw, T* __restrict__ rw, T* __restrict__ sw) {
  const Dims& d = c.d;
  const int lane = c.lane, n = d.nz - 1, Kc = d.nz + 1;
  const T h = T(0.5);
  const bool i_in = i >= 1 && i <= d.nx;
  const bool step = (c.mode & TC_K1_STEP) != 0, ab2 = step && !(c.mode & TC_K1_FIRST);
  T rz0[R], rz1[R], rz2[R];
  k1_metric<T, R, true>(c, n, rz0, rz1, rz2);
  const long long oi = (long long)i * d.sxw;
  const T* wi = w + oi;
  const int jlo = max(j0, 1), jhi = min(j1 - 1, d.ny);
  if (!i_in || jlo > jhi) {
    k1_plain_cols<T, R, true>(c, wi, pw + oi, rw + oi, sw + oi, nullptr, d.syw, j0, j1, 1 << 30, 0, n, Kc, rz0,
                              rz1, rz2);
    return;
  }
  k1_plain_cols<T, R, true>(c, wi, pw + oi, rw + oi, sw + oi, nullptr, d.syw, j0, jlo, 1, 0, n, Kc, rz0, rz1, rz2);
  const T* ui = u + (long long)i * d.sxu;
  const T* vi = v + (long long)i * d.sxv;
  auto fetch = [&](int j, WIn<T, R>& L) {
    const long long cw = (long long)j * d.syw;
    const T* u0 = ui + (long long)j * d.syu;
    const T* v0 = vi + (long long)j * d.syv;
    col_load<R, T>(wi + cw + d.syw, L.wN, n, lane);
    col_load<R, T>(wi - d.sxw + cw, L.wW, n, lane);
    col_load<R, T>(wi + d.sxw + cw, L.wE, n, lane);
    col_load<R, T>(u0, L.u00, n, lane);
    col_load<R, T>(u0 - d.sxu, L.um0, n, lane);
    col_load<R, T>(v0, L.v00, n, lane);
    if (ab2) col_load_hist<R, T>(pw + oi + cw, L.rp, n, lane);
    L.wb0 = wi[cw];
    L.wbn = wi[cw + n + 1];
    L.ubn = u0[n + 1];
    L.umbn = u0[n + 1 - d.sxu];
    L.vbn = v0[n + 1];
  };
  T wS[R], wC[R], vm0[R], vm1[R];
  col_load<R, T>(wi + (long long)(jlo - 1) * d.syw, wS, n, lane);
  col_load<R, T>(wi + (long long)jlo * d.syw, wC, n, lane);
  {
    const T* vp = vi + (long long)(jlo - 1) * d.syv;
    col_load<R, T>(vp, vm0, n, lane);
    rows_up<R, T>(c.rows, vm0, vm1, vp[n + 1], n, lane);
  }
  WIn<T, R> A;
  fetch(jlo, A);
  for (int j = jlo; j <= jhi; ++j) {
    const long long cw = (long long)j * d.syw;
    T rhs[R], q[R], wb[R], wt[R], u01[R], um1[R], v01[R];
    rows_pm<R, T>(c.rows, wC, wb, wt, A.wb0, A.wbn, n, lane);
    rows_up<R, T>(c.rows, A.u00, u01, A.ubn, n, lane);
    rows_up<R, T>(c.rows, A.um0, um1, A.umbn, n, lane);
    rows_up<R, T>(c.rows, A.v00, v01, A.vbn, n, lane);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T lap = (A.wE[r] - T(2) * wC[r] + A.wW[r]) * c.mt.kxx + (A.wN[r] - T(2) * wC[r] + wS[r]) * c.mt.kyy;
      const T fx = (h * (A.u00[r] + u01[r]) * (h * (wC[r] + A.wE[r])) - h * (A.um0[r] + um1[r]) * (h * (A.wW[r] + wC[r]))) * c.mt.rdx;
      const T fy = (h * (A.v00[r] + v01[r]) * (h * (wC[r] + A.wN[r])) - h * (vm0[r] + vm1[r]) * (h * (wS[r] + wC[r]))) * c.mt.rdy;
      const T wr = h * (wC[r] + wt[r]), wl = h * (wb[r] + wC[r]);
      const T fz = (wr * wr - wl * wl) * rz2[r];
      rhs[r] = lap - (fx + fy + fz);
      vm0[r] = A.v00[r];
      vm1[r] = v01[r];
      q[r] = wC[r];
      wS[r] = wC[r];
      wC[r] = A.wN[r];
    }
    T x[R];
    const bool cn = k1_rhs_ab2<T, R>(c, rhs, q, A.rp, rw + oi + cw, nullptr, n, Kc, true, x);
    if (j < jhi) fetch(j + 1, A);
    if (cn) k1_cn_store<T, R, true>(c, x, sw + oi + cw, n, rz0, rz1, rz2, T(0));
  }
  k1_plain_cols<T, R, true>(c, wi, pw + oi, rw + oi, sw + oi, nullptr, d.syw, jhi + 1, j1, 1, 0, n, Kc, rz0, rz1, rz2);
}

// ---------------------------------------------------------------------------------------------
// Three-component pencils.  One wavefront owns one i and marches j for u, v and w together: per
// step it loads the nine columns u, v, w at i-1, i, i+1 of a plane (four of them a plane ahead) and
// the three history columns, so each column that another component's stencil reads is fetched
// once per wavefront (12 column loads per (i, j) instead of the per-component pencils' 22).  The
// explicit terms are the per-component pencils' expressions, evaluated on the same values, so the
// results are bit-identical to theirs.  Interior only (1 <= i <= nx-1 and 1 <= j0, j1-1 <= ny-1,
// where all three components have stencils in both rhs branches); the caller runs the
// per-component pencils elsewhere.  Every column is held with nz rows (w's row nz is the top
// face), which changes no value the w stencil keeps: its rows past n = nz-1 are dropped.
// ---------------------------------------------------------------------------------------------
template <typename T, int R>
struct K3Next {
  T uE[R], vW[R], vE[R], wW[R], wE[R];   // plane j+1: columns only step j+1 reads
  T uN[R], umN[R], vN[R], wN[R];         // plane j+2: u(i), u(i-1), v(i), w(i)
  T pu[R], pv[R], pw[R];                 // history at j+1
  T ub0, ubn, vb0, vbn, wEb0;                          // ghost rows at j+1
  T wNb0;                                // w(i, j+2)[0]
};

// met: the block's LDS copy of the wall-normal metric of the lane rows (k1_metric: cell rz0, rz1,
// rz2, then face rz0, rz1, rz2, each 64*R values), read where used instead of held in registers.
template <int Q, typename T, int R>
__device__ __forceinline__ void k3_met(const T* met, T (&x)[R], int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = met[(Q * R + r) * TC_WAVE + lane];
}

// Registers the three-component march starts from: the plane j0-1 columns step j0 reads and the
// plane j0 columns u(i-1), u(i), v(i), w(i) with w's ghost row (loaded by k3_start), and the
// fetched inputs of step j0 (loaded by the march itself).
template <typename T, int R>
struct K3Start {
  T uS[R], vS[R], v1m[R], wS[R];
  T uC[R], uW[R], vC[R], wC[R];
  T wC0;
  K3Next<T, R> A;
};

// Loads of the fetched step's inputs that are not yet in registers (columns at (i, j) of the
// fetched step, history included) and of the plane one ahead.
template <typename T, int R>
__device__ __forceinline__ void k3_fetch_now(const Dims& d, const T* uj, const T* vj, const T* wj, const T* puj,
                                             const T* pvj, const T* pwj, bool ab2, int lane, K3Next<T, R>& L) {
  const int n = d.nz;
  col_load<R, T>(uj + d.sxu, L.uE, n, lane);
  col_load<R, T>(vj - d.sxv, L.vW, n, lane);
  col_load<R, T>(vj + d.sxv, L.vE, n, lane);
  col_load<R, T>(wj - d.sxw, L.wW, n, lane);
  col_load<R, T>(wj + d.sxw, L.wE, n, lane);
  if (ab2) {
    col_load_hist<R, T>(puj, L.pu, n, lane);
    col_load_hist<R, T>(pvj, L.pv, n, lane);
    col_load_hist<R, T>(pwj, L.pw, n, lane);
  }
}
template <typename T, int R>
__device__ __forceinline__ void k3_fetch_ahead(const Dims& d, const T* uj, const T* vj, const T* wj, int lane,
                                               K3Next<T, R>& L) {
  const int n = d.nz;
  col_load<R, T>(uj + d.syu, L.uN, n, lane);
  col_load<R, T>(uj - d.sxu + d.syu, L.umN, n, lane);
  col_load<R, T>(vj + d.syv, L.vN, n, lane);
  col_load<R, T>(wj + d.syw, L.wN, n, lane);
  L.ub0 = uj[0];
  L.ubn = uj[n + 1];
  L.vb0 = vj[0];
  L.vbn = vj[n + 1];
  L.wEb0 = wj[d.sxw];
  L.wNb0 = wj[d.syw];
}

// The march's first loads for column (i, j0), the two planes it starts from: issued by the kernel
// before its block stages the CN tables and metric into LDS, so that the two latencies overlap
// (measured at 192x192x128: K1 -1.2 % FP64, -1.5 % FP32; 384x384x256 FP32 -1.6 %).  Issuing step
// j0's fetch here too kept 21 columns live across the staging and took the float64 two-row
// instance from 146 to 209 VGPRs.
template <typename T, int R>
__device__ __forceinline__ void k3_start(const Dims& d, int i, int j0, const T* __restrict__ u,
                                         const T* __restrict__ v, const T* __restrict__ w, int lane,
                                         K3Start<T, R>& S) {
  const int n = d.nz;
  const long long ou = (long long)i * d.sxu, ov = (long long)i * d.sxv, ow = (long long)i * d.sxw;
  const T* uj = u + ou + (long long)j0 * d.syu;
  const T* vj = v + ov + (long long)j0 * d.syv;
  const T* wj = w + ow + (long long)j0 * d.syw;
  col_load<R, T>(uj - d.syu, S.uS, n, lane);
  col_load<R, T>(vj - d.syv, S.vS, n, lane);
  col_load<R, T>(vj + d.sxv - d.syv, S.v1m, n, lane);
  col_load<R, T>(wj - d.syw, S.wS, n, lane);
  col_load<R, T>(uj, S.uC, n, lane);
  col_load<R, T>(uj - d.sxu, S.uW, n, lane);
  col_load<R, T>(vj, S.vC, n, lane);
  col_load<R, T>(wj, S.wC, n, lane);
  S.wC0 = wj[0];
}

template <typename T, int R>
__device__ __forceinline__ void k1_pencil_uvw(const K1Ctx<T, R>& cu, const K1Ctx<T, R>& cv, const K1Ctx<T, R>& cw,
                                              const T* met, int i, int j0, int j1, const T* __restrict__ u,
                                              const T* __restrict__ v, const T* __restrict__ w,
                                              const T* __restrict__ pu, const T* __restrict__ pv,
                                              const T* __restrict__ pw, T* __restrict__ ru, T* __restrict__ rv,
                                              T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv,
                                              T* __restrict__ sw, K3Start<T, R>& S) {
  const Dims& d = cu.d;
  const int lane = cu.lane, n = d.nz, nw = d.nz - 1;
  const T h = T(0.5);
  const bool step = (cu.mode & TC_K1_STEP) != 0, ab2 = step && !(cu.mode & TC_K1_FIRST);
  const long long ou = (long long)i * d.sxu, ov = (long long)i * d.sxv, ow = (long long)i * d.sxw;
  const long long sxu = d.sxu, sxv = d.sxv, sxw = d.sxw, syu = d.syu, syv = d.syv, syw = d.syw;
  const T* ui = u + ou;
  const T* vi = v + ov;
  const T* wi = w + ow;
  // Column pointers carried along the march (each advanced by its array's column stride per
  // step, not recomputed from j): uj, vj, wj and the history columns at the (i, j) whose inputs
  // are being fetched; the rhs and star outputs at the (i, j) being finished.
  const T* uj = ui + (long long)j0 * syu;
  const T* vj = vi + (long long)j0 * syv;
  const T* wj = wi + (long long)j0 * syw;
  const T* puj = pu + ou + (long long)j0 * syu;
  const T* pvj = pv + ov + (long long)j0 * syv;
  const T* pwj = pw + ow + (long long)j0 * syw;
  T* ruj = ru + ou + (long long)j0 * syu;
  T* rvj = rv + ov + (long long)j0 * syv;
  T* rwj = rw + ow + (long long)j0 * syw;
  T* suj = su + ou + (long long)j0 * syu;
  T* svj = sv + ov + (long long)j0 * syv;
  T* swj = sw + ow + (long long)j0 * syw;
  const T* bj = cu.body != nullptr ? cu.body + ou + (long long)j0 * syu : nullptr;
  // the registers k3_start loaded (planes j0-1 and j0), then step j0's fetch
  T (&uS)[R] = S.uS;
  T (&vS)[R] = S.vS;
  T (&v1m)[R] = S.v1m;
  T (&wS)[R] = S.wS;
  T (&uC)[R] = S.uC;
  T (&uW)[R] = S.uW;
  T (&vC)[R] = S.vC;
  T (&wC)[R] = S.wC;
  T& wC0 = S.wC0;
  K3Next<T, R>& A = S.A;
  k3_fetch_ahead<T, R>(d, uj, vj, wj, lane, A);
  k3_fetch_now<T, R>(d, uj, vj, wj, puj, pvj, pwj, ab2, lane, A);
  // v(i, j0-1) one row up
  T vm1[R];
  rows_up_in<R, T>(cu.rows, vS, vm1, lane);
  for (int j = j0; j < j1; ++j) {
    T xu[R], xv[R], xw[R];
    bool cn_u, cn_v, cn_w;
    {
      // u (k1_pencil_u): v00 = v(i,j), v10 = v(i+1,j), v0m = v(i,j-1), v1m = v(i+1,j-1), w00 = w(i,j),
      // w10 = w(i+1,j)
      T rhs[R], ub[R], ut[R], w0m[R], w1m[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        rhs[r] = (A.uE[r] - T(2) * uC[r] + uW[r]) * cu.mt.kxx + (A.uN[r] - T(2) * uC[r] + uS[r]) * cu.mt.kyy;
      rows_pm<R, T>(cu.rows, uC, ub, ut, A.ub0, A.ubn, n, lane);
      rows_dn<R, T>(cu.rows, wC, w0m, wC0, lane);
      rows_dn<R, T>(cu.rows, A.wE, w1m, A.wEb0, lane);
      T rc2[R];
      k3_met<2, T, R>(met, rc2, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T ur = h * (uC[r] + A.uE[r]), ul = h * (uW[r] + uC[r]);
        const T fx = (ur * ur - ul * ul) * cu.mt.rdx;
        const T fy = (h * (vC[r] + A.vE[r]) * (h * (uC[r] + A.uN[r])) - h * (vS[r] + v1m[r]) * (h * (uS[r] + uC[r]))) * cu.mt.rdy;
        const T fz = (h * (wC[r] + A.wE[r]) * (h * (uC[r] + ut[r])) - h * (w0m[r] + w1m[r]) * (h * (ub[r] + uC[r]))) * rc2[r];
        rhs[r] = rhs[r] - (fx + fy + fz);
      }
      cn_u = k1_rhs_ab2<T, R>(cu, rhs, uC, A.pu, ruj, bj, n, n + 2, true, xu);
      // v (k1_pencil_v): u00 = u(i,j), u01 = u(i,j+1), um0 = u(i-1,j), um1 = u(i-1,j+1), w00 = w(i,j),
      // w01 = w(i,j+1), w0m = w(i,j) one row down (as above)
      T vb[R], vt[R], w1mv[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        rhs[r] = (A.vE[r] - T(2) * vC[r] + A.vW[r]) * cv.mt.kxx + (A.vN[r] - T(2) * vC[r] + vS[r]) * cv.mt.kyy;
      rows_pm<R, T>(cu.rows, vC, vb, vt, A.vb0, A.vbn, n, lane);
      rows_dn<R, T>(cu.rows, A.wN, w1mv, A.wNb0, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T fx = (h * (uC[r] + A.uN[r]) * (h * (vC[r] + A.vE[r])) - h * (uW[r] + A.umN[r]) * (h * (A.vW[r] + vC[r]))) * cv.mt.rdx;
        const T vr = h * (vC[r] + A.vN[r]), vl = h * (vS[r] + vC[r]);
        const T fy = (vr * vr - vl * vl) * cv.mt.rdy;
        const T fz = (h * (wC[r] + A.wN[r]) * (h * (vC[r] + vt[r])) - h * (w0m[r] + w1mv[r]) * (h * (vb[r] + vC[r]))) * rc2[r];
        rhs[r] = rhs[r] - (fx + fy + fz);
      }
      cn_v = k1_rhs_ab2<T, R>(cv, rhs, vC, A.pv, rvj, nullptr, n, n + 2, true, xv);
      // w (k1_pencil_w, rows k = 1..nz-1): u00 = u(i,j), um0 = u(i-1,j), v00 = v(i,j), vm0 = v(i,j-1),
      // each with its row above; the pencil's top overrides (row nz of w, u and v) are the rows
      // these columns hold in registers, so the plain exchanges give the same values
      T wb[R], wt[R], u01[R], um1[R], v01[R];
      rows_pm_lo<R, T>(cu.rows, wC, wb, wt, wC0, lane);
      rows_up_in<R, T>(cu.rows, uC, u01, lane);
      rows_up_in<R, T>(cu.rows, uW, um1, lane);
      rows_up_in<R, T>(cu.rows, vC, v01, lane);
      T rf2[R];
      k3_met<5, T, R>(met, rf2, lane);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T lap = (A.wE[r] - T(2) * wC[r] + A.wW[r]) * cw.mt.kxx + (A.wN[r] - T(2) * wC[r] + wS[r]) * cw.mt.kyy;
        const T fx = (h * (uC[r] + u01[r]) * (h * (wC[r] + A.wE[r])) - h * (uW[r] + um1[r]) * (h * (A.wW[r] + wC[r]))) * cw.mt.rdx;
        const T fy = (h * (vC[r] + v01[r]) * (h * (wC[r] + A.wN[r])) - h * (vS[r] + vm1[r]) * (h * (wS[r] + wC[r]))) * cw.mt.rdy;
        const T wr = h * (wC[r] + wt[r]), wl = h * (wb[r] + wC[r]);
        const T fz = (wr * wr - wl * wl) * rf2[r];
        rhs[r] = lap - (fx + fy + fz);
      }
      cn_w = k1_rhs_ab2<T, R>(cw, rhs, wC, A.pw, rwj, nullptr, nw, n + 1, true, xw);
      // carry plane j as plane j-1 and plane j+1 as plane j
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uS[r] = uC[r];
        vS[r] = vC[r];
        v1m[r] = A.vE[r];
        wS[r] = wC[r];
        vm1[r] = v01[r];
        uC[r] = A.uN[r];
        uW[r] = A.umN[r];
        vC[r] = A.vN[r];
        wC[r] = A.wN[r];
      }
      wC0 = A.wNb0;
    }
    // the next step's loads are in flight during the three CN solves
    if (j + 1 < j1) {
      uj += syu;
      vj += syv;
      wj += syw;
      puj += syu;
      pvj += syv;
      pwj += syw;
      k3_fetch_ahead<T, R>(d, uj, vj, wj, lane, A);
      k3_fetch_now<T, R>(d, uj, vj, wj, puj, pvj, pwj, ab2, lane, A);
    }
    if (cn_u || cn_v) {
      T a0[R], a1[R], a2[R];
      k3_met<0, T, R>(met, a0, lane);
      k3_met<1, T, R>(met, a1, lane);
      k3_met<2, T, R>(met, a2, lane);
      if (cn_u) k1_cn_store<T, R, false>(cu, xu, suj, n, a0, a1, a2, T(0));
      if (cn_v) k1_cn_store<T, R, false>(cv, xv, svj, n, a0, a1, a2, T(0));
    }
    if (cn_w) {
      T a0[R], a1[R], a2[R];
      k3_met<3, T, R>(met, a0, lane);
      k3_met<4, T, R>(met, a1, lane);
      k3_met<5, T, R>(met, a2, lane);
      k1_cn_store<T, R, true>(cw, xw, swj, nw, a0, a1, a2, T(0));
    }
    ruj += syu;
    rvj += syv;
    rwj += syw;
    suj += syu;
    svj += syv;
    swj += syw;
    if (bj != nullptr) bj += syu;
  }
}

// Pencil bands: a band is TC_K1_WAVES consecutive i (one per wavefront) by jb consecutive j, on
// the common (nx+2) x (ny+2) column plane; block b serves component b % 3 of band b / 3, so the
// three components of one band run back to back on the same XCD.
__host__ __device__ __forceinline__ int k1_bands_i(int nx) { return (nx + 2 + TC_K1_WAVES - 1) / TC_K1_WAVES; }
__host__ __device__ __forceinline__ int k1_bands_j(int ny, int jb) { return (ny + 2 + jb - 1) / jb; }

// LDS of the per-component pencils' block: the CN table, then one row buffer per wavefront.
template <typename T, int R>
__host__ __device__ constexpr size_t k1_block_lds() {
  return sizeof(T) * ((size_t)k1_tab_len<R>() + (size_t)TC_K1_WAVES * 3 * TC_WAVE * R);
}

// FULL: the launcher has checked nz == 64*R, so every lane's rows lie inside the u and v columns
// (and all but the last lane's last row inside the w column); with nz a compile-time constant the
// bounds guards of the column loads and stores fold away.
// DYN: the table and row buffers sit at `smem` (k1_block_lds<T, R>() bytes of the caller's LDS)
// instead of in static LDS of their own, for a kernel that runs several bodies in turn.
template <typename T, int R, bool FULL = false, bool DYN = false>
__device__ __forceinline__ void explicit_cn_body(
    Dims d, Metric<T> mt, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
    const T* __restrict__ pu, const T* __restrict__ pv, const T* __restrict__ pw, T* __restrict__ ru,
    T* __restrict__ rv, T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv, T* __restrict__ sw,
    const double* __restrict__ tab_cell, const double* __restrict__ tab_face,
    const double* __restrict__ scal, const T* __restrict__ body, int mode, int branch, double theta, int jb,
    int b, T* smem = nullptr) {
  T* lds_tab;
  T* lds_rows0;
  if constexpr (DYN) {
    lds_tab = smem;
    lds_rows0 = smem + k1_tab_len<R>();
  } else {
    __shared__ T s_tab[k1_tab_len<R>()];
    __shared__ T s_rows[TC_K1_WAVES][3 * TC_WAVE * R];
    lds_tab = s_tab;
    lds_rows0 = &s_rows[0][0];
  }
  if constexpr (FULL) d.nz = TC_WAVE * R;
  const int comp = b % 3, band = b / 3;
  const int nbj = k1_bands_j(d.ny, jb);
  const int bi = band / nbj, bj = band - bi * nbj;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  if (mode & TC_K1_STEP) {
    const double* tab = comp < 2 ? tab_cell : tab_face;
    for (int e = threadIdx.x; e < k1_tab_len<R>(); e += 64 * TC_K1_WAVES) lds_tab[e] = (T)k1_tab_src<R>(tab, e);
    __syncthreads();
  }
  const int i = bi * TC_K1_WAVES + wv;
  const int imax = comp == 0 ? d.nx : d.nx + 1;
  const int jmax = comp == 1 ? d.ny : d.ny + 1;
  const int j0 = bj * jb;
  const int j1 = min(j0 + jb, jmax + 1);
  if (i > imax || j0 > jmax) return;
  K1Ctx<T, R> c;
  c.d = d;
  c.mt = mt;
  c.rows = lds_rows0 + wv * (3 * TC_WAVE * R);
  c.tabl = lds_tab;
  c.lane = lane;
  c.mode = mode;
  c.cpu = branch == TC_BRANCH_CPU;
  const double dt = scal[TC_S_DT];
  c.dt = (T)dt;
  c.beta = (T)((1.0 - theta) * dt * mt.nu);
  c.F = (comp == 0 && (mode & TC_K1_FORCE)) ? (T)scal[TC_S_F] : T(0);
  c.body = (comp == 0 && (mode & TC_K1_FORCE)) ? body : nullptr;
  rows_zero_pads<R, T>(c.rows, lane);
  if (comp == 0) k1_pencil_u<T, R>(c, i, j0, j1, u, v, w, pu, ru, su);
  else if (comp == 1) k1_pencil_v<T, R>(c, i, j0, j1, u, v, w, pv, rv, sv);
  else k1_pencil_w<T, R>(c, i, j0, j1, u, v, w, pw, rw, sw);
}

template <typename T, int R, bool FULL>
__global__ __launch_bounds__(64 * TC_K1_WAVES) void k_explicit_cn(
    Dims d, Metric<T> mt, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
    const T* __restrict__ pu, const T* __restrict__ pv, const T* __restrict__ pw, T* __restrict__ ru,
    T* __restrict__ rv, T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv, T* __restrict__ sw,
    const double* __restrict__ tab_cell, const double* __restrict__ tab_face,
    const double* __restrict__ scal, const T* __restrict__ body, int mode, int branch, double theta, int jb) {
  explicit_cn_body<T, R, FULL>(d, mt, u, v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, tab_cell, tab_face, scal, body,
                               mode, branch, theta, jb, xcd_remap(blockIdx.x, gridDim.x));
}


// Three-component form (TC_TUNE_K1_TILE = 2): a block of W wavefronts owns a band of W
// consecutive i by jb consecutive j for all three components, with both CN tables in LDS;
// interior bands march k1_pencil_uvw, the others run the per-component pencils in turn.
// W = 8 where a wavefront fits 128 VGPRs (one row per lane, or float32 with two), so that two
// blocks give a CU 4 wavefronts per SIMD; else W = 4 (more VGPRs, three blocks per CU).  Capping
// the float64 two-row instance at 128 VGPRs spills inside the march (measured: 146 -> 248 us at
// 192x192x128), so it keeps its 152 VGPRs and 3 wavefronts per SIMD; likewise capping the float32
// two-row instance at 64 VGPRs for 8 wavefronts per SIMD spills (82 -> 157 us), so it keeps 79.
template <typename T, int R>
__host__ __device__ constexpr int k3_waves() { return (R == 1 || (sizeof(T) == 4 && R == 2)) ? 8 : 4; }
template <typename T, int R>
__host__ __device__ constexpr int k3_min_waves() { return k3_waves<T, R>() == 8 ? 4 : 1; }
__host__ __device__ __forceinline__ int k3_bands_i(int nx, int W) { return (nx + 2 + W - 1) / W; }

template <typename T, int R, bool FULL>
__global__ __launch_bounds__((64 * k3_waves<T, R>()), (k3_min_waves<T, R>())) void k_explicit_cn_uvw(
    Dims d, Metric<T> mt, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
    const T* __restrict__ pu, const T* __restrict__ pv, const T* __restrict__ pw, T* __restrict__ ru,
    T* __restrict__ rv, T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv, T* __restrict__ sw,
    const double* __restrict__ tab_cell, const double* __restrict__ tab_face,
    const double* __restrict__ scal, const T* __restrict__ body, int mode, int branch, double theta, int jb) {
  constexpr int W = k3_waves<T, R>();
  __shared__ T lds_tab[2][k1_tab_len<R>()];
  __shared__ T lds_rows[W][3 * TC_WAVE * R];
  __shared__ T lds_met[6 * TC_WAVE * R];
  if constexpr (FULL) d.nz = TC_WAVE * R;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int nbj = k1_bands_j(d.ny, jb);
  const int bi = b / nbj, bj = b - bi * nbj;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int i = bi * W + wv;
  const int j0 = bj * jb, jend = j0 + jb;
  const bool inner = i >= 1 && i <= d.nx - 1 && j0 >= 1 && jend - 1 <= d.ny - 1;
  // both CN tables and the metric of k1_metric (cell rows: n = nz, face rows: n = nz - 1) into
  // LDS; an interior wavefront issues its first column loads (k3_start) before taking part
  auto stage = [&]() {
    if (mode & TC_K1_STEP) {
      for (int e = threadIdx.x; e < k1_tab_len<R>(); e += 64 * W) {
        lds_tab[0][e] = (T)k1_tab_src<R>(tab_cell, e);
        lds_tab[1][e] = (T)k1_tab_src<R>(tab_face, e);
      }
    }
    for (int e = threadIdx.x; e < TC_WAVE * R; e += 64 * W) {
      const int nc = d.nz, nf = d.nz - 1;
      lds_met[e] = e < nc ? mt.rdzc[e] : T(0);
      lds_met[TC_WAVE * R + e] = e < nc ? mt.rdzc[e + 1] : T(0);
      lds_met[2 * TC_WAVE * R + e] = e < nc ? mt.rdzf[e] : T(0);
      lds_met[3 * TC_WAVE * R + e] = e < nf ? mt.rdzf[e] : T(0);
      lds_met[4 * TC_WAVE * R + e] = e < nf ? mt.rdzf[e + 1] : T(0);
      lds_met[5 * TC_WAVE * R + e] = e < nf ? mt.rdzc[e + 1] : T(0);
    }
  };
  auto ctx = [&](K1Ctx<T, R>& cu) {
    cu.d = d;
    cu.mt = mt;
    cu.rows = lds_rows[wv];
    cu.tabl = lds_tab[0];
    cu.lane = lane;
    cu.mode = mode;
    cu.cpu = branch == TC_BRANCH_CPU;
    const double dt = scal[TC_S_DT];
    cu.dt = (T)dt;
    cu.beta = (T)((1.0 - theta) * dt * mt.nu);
    cu.F = (mode & TC_K1_FORCE) ? (T)scal[TC_S_F] : T(0);
    cu.body = (mode & TC_K1_FORCE) ? body : nullptr;
  };
  K3Start<T, R> S;
  if (inner) k3_start<T, R>(d, i, j0, u, v, w, lane, S);
  stage();
  __syncthreads();
  if (i > d.nx + 1 || j0 > d.ny + 1) return;
  K1Ctx<T, R> cu;
  ctx(cu);
  K1Ctx<T, R> cv = cu;
  cv.F = T(0);
  cv.body = nullptr;
  K1Ctx<T, R> cw = cv;
  cw.tabl = lds_tab[1];
  rows_zero_pads<R, T>(cu.rows, lane);
  if (inner) {
    k1_pencil_uvw<T, R>(cu, cv, cw, lds_met, i, j0, jend, u, v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, S);
  } else {
    if (i <= d.nx) k1_pencil_u<T, R>(cu, i, j0, min(jend, d.ny + 2), u, v, w, pu, ru, su);
    if (j0 <= d.ny) k1_pencil_v<T, R>(cv, i, j0, min(jend, d.ny + 1), u, v, w, pv, rv, sv);
    k1_pencil_w<T, R>(cw, i, j0, min(jend, d.ny + 2), u, v, w, pw, rw, sw);
  }
}

// ---------------------------------------------------------------------------------------------
// K1, tiled form.  The same explicit terms, forcing, AB2 and CN solve as the pencils above
// (every expression is the pencils' own, so results are bit-identical), with the neighbour
// columns staged through LDS instead of each wavefront fetching them:
//
// * a block of TC_KT_WAVES = 12 wavefronts owns the three components of a band of TC_KT_I = 4
//   consecutive i (wavefront w: component w / 4, i = i0 + w % 4) and all of them march along j
//   together over the band's j range;
// * each j-plane's columns of u, v and w for i0-1 .. i0+4 (the band plus one halo column each
//   side, every row including the wall-normal ghosts) are read from HBM once, by the owning
//   wavefront (the edge wavefronts also read the halo columns), into a two-plane LDS ring;
// * a wavefront reads its i+-1 / j+1 / other-component neighbours, and their k+-1 rows, from
//   the ring (LDS entry e of a column is global row k = e), so the only HBM loads per (i, j)
//   and component are its own column, the AB2 history and the halo share;
// * plane j+2 and the next AB2 history are fetched while step j computes; the ring holds three
//   planes, so one block barrier per j step suffices: plane j+1 is written into the slot of plane
//   j-2, which every wavefront finished reading before it reached the barrier of step j-1.
// The pencils' per-wavefront row buffer still serves the CN right-hand side and PCR exchanges.
// ---------------------------------------------------------------------------------------------
#define TC_KT_I 4
#define TC_KT_WAVES (3 * TC_KT_I)

template <typename T, int R>
struct KT {
  static constexpr int CW = TC_WAVE * R + 2;                 // LDS column: global rows 0..64R+1
  static constexpr int SLOTS = TC_KT_I + 2;                  // i0-1 .. i0+4
  static constexpr int PLANE = 3 * SLOTS * CW;
  static constexpr int TAB = k1_tab_len<R>();
  static constexpr int ROWS = 3 * TC_WAVE * R;
  static constexpr int NPL = 3;                              // planes in the ring
  static constexpr size_t BYTES = sizeof(T) * ((size_t)NPL * PLANE + 2 * TAB + (size_t)TC_KT_WAVES * ROWS);
};

// Source column of ring slot (component c, global i) for plane p, or null when the slot has no
// reader: u at i = nx+1 is u(1) (the periodic wrap the CPU branch uses for u-face nx); v at
// p = ny+1 is v(i, 1) in the CPU branch (its wrap for v-face ny).
template <typename T>
__device__ __forceinline__ const T* kt_src(const Dims& d, int c, int i, int p, bool cpu, const T* u, const T* v,
                                           const T* w, int& kc) {
  if (i < 0 || p < 0) return nullptr;
  if (c == 0) {
    kc = d.nz + 2;
    if (i == d.nx + 1 && cpu) i = 1;
    if (i > d.nx || p > d.ny + 1) return nullptr;
    return u + (long long)i * d.sxu + (long long)p * d.syu;
  }
  if (c == 1) {
    kc = d.nz + 2;
    if (p == d.ny + 1 && cpu) p = 1;
    if (i > d.nx + 1 || p > d.ny) return nullptr;
    return v + (long long)i * d.sxv + (long long)p * d.syv;
  }
  kc = d.nz + 1;
  if (i > d.nx + 1 || p > d.ny + 1) return nullptr;
  return w + (long long)i * d.sxw + (long long)p * d.syw;
}

// A column in registers on its way to the ring: rows 1..64R, plus row 0 (lane 0) and row 64R+1
// (lane 1) in g; rows past the column read zero.
template <typename T, int R>
struct KTCol {
  T x[R];
  T g;
};

template <typename T, int R>
__device__ __forceinline__ void kt_fetch(const T* col, int kc, KTCol<T, R>& c, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = r * TC_WAVE + lane + 1;
    c.x[r] = k < kc ? col[k] : T(0);
  }
  const int kg = lane == 0 ? 0 : TC_WAVE * R + 1;
  c.g = (lane < 2 && kg < kc) ? col[kg] : T(0);
}

template <typename T, int R>
__device__ __forceinline__ void kt_put(T* colL, const KTCol<T, R>& c, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) colL[r * TC_WAVE + lane + 1] = c.x[r];
  if (lane == 0) colL[0] = c.g;
  if (lane == 1) colL[TC_WAVE * R + 1] = c.g;
}

// rows m = r*64 + lane of an LDS column at global row m + OFF (OFF = 0: k-1, 1: k, 2: k+1)
template <int OFF, typename T, int R>
__device__ __forceinline__ void kt_rows(const T* colL, T (&x)[R], int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = colL[r * TC_WAVE + lane + OFF];
}

// Columns this wavefront puts into the ring for one plane: its own slot, and the halo slot of
// the band's edge wavefronts.
template <typename T, int R>
struct KTLoads {
  KTCol<T, R> own, halo;
  bool has_own, has_halo;
};

template <typename T, int R>
__device__ __forceinline__ void kt_load_plane(const Dims& d, int comp, int i, int il, int i0, int p, bool cpu,
                                              const T* u, const T* v, const T* w, KTLoads<T, R>& L, int lane) {
  int kc = 0;
  const T* s = kt_src<T>(d, comp, i, p, cpu, u, v, w, kc);
  L.has_own = s != nullptr;
  if (L.has_own) kt_fetch<T, R>(s, kc, L.own, lane);
  L.has_halo = false;
  if (il == 0 || il == TC_KT_I - 1) {
    const int ih = il == 0 ? i0 - 1 : i0 + TC_KT_I;
    const T* h = kt_src<T>(d, comp, ih, p, cpu, u, v, w, kc);
    L.has_halo = h != nullptr;
    if (L.has_halo) kt_fetch<T, R>(h, kc, L.halo, lane);
  }
}

template <typename T, int R>
__device__ __forceinline__ void kt_store_plane(T* plane, int comp, int il, const KTLoads<T, R>& L, int lane) {
  constexpr int CW = KT<T, R>::CW;
  T* base = plane + comp * KT<T, R>::SLOTS * CW;
  if (L.has_own) kt_put<T, R>(base + (il + 1) * CW, L.own, lane);
  if (L.has_halo) kt_put<T, R>(base + (il == 0 ? 0 : TC_KT_I + 1) * CW, L.halo, lane);
}

template <typename T, int R>
__device__ __forceinline__ const T* kt_col(const T* plane, int c, int slot) {
  return plane + (c * KT<T, R>::SLOTS + slot) * KT<T, R>::CW;
}

// Column kinds of a step: 0 none (no such column), 1 ghost (rhs = forcing only), 2 inner
// without explicit terms (rhs = forcing, AB2 + CN), 3 stencil.
template <int COMP>
__device__ __forceinline__ int kt_kind(const Dims& d, int i, int j, bool cpu) {
  if (COMP == 0) {
    if (i > d.nx || j > d.ny + 1) return 0;
    const bool i_in = i >= 1, j_in = j >= 1 && j <= d.ny;
    if (!i_in || !j_in) return 1;
    return (i <= d.nx - 1 || cpu) ? 3 : 2;
  }
  if (COMP == 1) {
    if (i > d.nx + 1 || j > d.ny) return 0;
    const bool i_in = i >= 1 && i <= d.nx, j_in = j >= 1;
    if (!i_in || !j_in) return 1;
    return (j <= (cpu ? d.ny : d.ny - 1)) ? 3 : 2;
  }
  if (i > d.nx + 1 || j > d.ny + 1) return 0;
  const bool i_in = i >= 1 && i <= d.nx, j_in = j >= 1 && j <= d.ny;
  return (i_in && j_in) ? 3 : 1;
}

// One component's march over the band's j range.  `ring` holds the LDS planes; plane p sits in
// slot p % 3.
template <typename T, int R, int COMP>
__device__ __forceinline__ void kt_march(const K1Ctx<T, R>& c, int i, int il, int i0, int ja, int j0, int j1,
                                         T* ring, const T* __restrict__ u, const T* __restrict__ v,
                                         const T* __restrict__ w, const T* __restrict__ pq,
                                         T* __restrict__ rq, T* __restrict__ sq) {
  const Dims& d = c.d;
  const int lane = c.lane;
  constexpr bool FACE = COMP == 2;
  const int n = FACE ? d.nz - 1 : d.nz;
  const int Kc = FACE ? d.nz + 1 : d.nz + 2;
  const long long sx = COMP == 0 ? d.sxu : COMP == 1 ? d.sxv : d.sxw;
  const long long sy = COMP == 0 ? d.syu : COMP == 1 ? d.syv : d.syw;
  const T h = T(0.5);
  const bool step = (c.mode & TC_K1_STEP) != 0, ab2 = step && !(c.mode & TC_K1_FIRST);
  T rz0[R], rz1[R], rz2[R];
  k1_metric<T, R, FACE>(c, n, rz0, rz1, rz2);
  const long long oi = (long long)i * sx;
  const T* bi = (COMP == 0 && c.body != nullptr) ? c.body + oi : nullptr;
  constexpr int PL = KT<T, R>::PLANE;
  // own column at j-1 (uS / vS / wS) and the carried other-component rows of plane j-1:
  // u: v(i, j-1), v(i+1, j-1); w: v(i, j-1) and its row k+1
  T qS[R], ca[R], cb[R];
#pragma unroll
  for (int r = 0; r < R; ++r) { qS[r] = T(0); ca[r] = T(0); cb[r] = T(0); }
  KTLoads<T, R> L;
  kt_load_plane<T, R>(d, COMP, i, il, i0, ja, c.cpu, u, v, w, L, lane);
  kt_store_plane<T, R>(ring + (ja % 3) * PL, COMP, il, L, lane);
  if (ja + 1 <= j1) kt_load_plane<T, R>(d, COMP, i, il, i0, ja + 1, c.cpu, u, v, w, L, lane);
  T rp[R], rpn[R];
#pragma unroll
  for (int r = 0; r < R; ++r) { rp[r] = T(0); rpn[r] = T(0); }
  int kind = kt_kind<COMP>(d, i, ja, c.cpu);
  if (ab2 && ja >= j0 && kind >= 2) col_load_hist<R, T>(pq + oi + (long long)ja * sy, rp, n, lane);
  __syncthreads();
  for (int j = ja; j < j1; ++j) {
    const T* Pj = ring + (j % 3) * PL;
    T* Pn = ring + ((j + 1) % 3) * PL;
    if (j + 1 <= j1) kt_store_plane<T, R>(Pn, COMP, il, L, lane);
    __syncthreads();
    if (j + 2 <= j1) kt_load_plane<T, R>(d, COMP, i, il, i0, j + 2, c.cpu, u, v, w, L, lane);
    const int kind_n = kt_kind<COMP>(d, i, j + 1, c.cpu);
    if (ab2 && j + 1 < j1 && j + 1 >= j0 && kind_n >= 2)
      col_load_hist<R, T>(pq + oi + (long long)(j + 1) * sy, rpn, n, lane);
    const T* own = kt_col<T, R>(Pj, COMP, il + 1);
    T qC[R];
    if (kind != 0) kt_rows<1, T, R>(own, qC, lane);
    const long long cj = (long long)j * sy;
    if (j >= j0 && kind != 0) {
      const T* bcol = bi != nullptr ? bi + cj : nullptr;
      // ghost columns, and in rhs-only launches the inner columns without explicit terms, carry
      // the forcing alone (k1_plain_cols)
      if (kind == 1 || (kind == 2 && !step)) {
        k1_ghost_col<T, R>(c, rq + oi + cj, bcol, n, Kc);
      } else {
        T rhs[R];
#pragma unroll
        for (int r = 0; r < R; ++r) rhs[r] = T(0);
        if (kind == 3) {
          T qN[R], qW[R], qE[R];
          kt_rows<1, T, R>(kt_col<T, R>(Pn, COMP, il + 1), qN, lane);
          kt_rows<1, T, R>(kt_col<T, R>(Pj, COMP, il), qW, lane);
          kt_rows<1, T, R>(kt_col<T, R>(Pj, COMP, il + 2), qE, lane);
#pragma unroll
          for (int r = 0; r < R; ++r)
            rhs[r] = (qE[r] - T(2) * qC[r] + qW[r]) * c.mt.kxx + (qN[r] - T(2) * qC[r] + qS[r]) * c.mt.kyy;
          if constexpr (COMP == 0) {
            if (i <= d.nx - 1) {                                   // advection (not the CPU wrap face)
              T ub[R], ut[R], v00[R], v10[R], w00[R], w10[R], w0m[R], w1m[R];
              kt_rows<0, T, R>(own, ub, lane);
              kt_rows<2, T, R>(own, ut, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 1), v00, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 2), v10, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 2, il + 1), w00, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 2, il + 2), w10, lane);
              kt_rows<0, T, R>(kt_col<T, R>(Pj, 2, il + 1), w0m, lane);
              kt_rows<0, T, R>(kt_col<T, R>(Pj, 2, il + 2), w1m, lane);
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const T ur = h * (qC[r] + qE[r]), ul = h * (qW[r] + qC[r]);
                const T fx = (ur * ur - ul * ul) * c.mt.rdx;
                const T fy = (h * (v00[r] + v10[r]) * (h * (qC[r] + qN[r])) - h * (ca[r] + cb[r]) * (h * (qS[r] + qC[r]))) * c.mt.rdy;
                const T fz = (h * (w00[r] + w10[r]) * (h * (qC[r] + ut[r])) - h * (w0m[r] + w1m[r]) * (h * (ub[r] + qC[r]))) * rz2[r];
                rhs[r] = rhs[r] - (fx + fy + fz);
              }
            }
          } else if constexpr (COMP == 1) {
            if (j < d.ny) {
              T vb[R], vt[R], u00[R], um0[R], u01[R], um1[R], w00[R], w0m[R], w01[R], w1m[R];
              kt_rows<0, T, R>(own, vb, lane);
              kt_rows<2, T, R>(own, vt, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 0, il + 1), u00, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 0, il), um0, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pn, 0, il + 1), u01, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pn, 0, il), um1, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pj, 2, il + 1), w00, lane);
              kt_rows<0, T, R>(kt_col<T, R>(Pj, 2, il + 1), w0m, lane);
              kt_rows<1, T, R>(kt_col<T, R>(Pn, 2, il + 1), w01, lane);
              kt_rows<0, T, R>(kt_col<T, R>(Pn, 2, il + 1), w1m, lane);
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const T fx = (h * (u00[r] + u01[r]) * (h * (qC[r] + qE[r])) - h * (um0[r] + um1[r]) * (h * (qW[r] + qC[r]))) * c.mt.rdx;
                const T vr = h * (qC[r] + qN[r]), vl = h * (qS[r] + qC[r]);
                const T fy = (vr * vr - vl * vl) * c.mt.rdy;
                const T fz = (h * (w00[r] + w01[r]) * (h * (qC[r] + vt[r])) - h * (w0m[r] + w1m[r]) * (h * (vb[r] + qC[r]))) * rz2[r];
                rhs[r] = rhs[r] - (fx + fy + fz);
              }
            }
          } else {
            T wb[R], wt[R], u00[R], um0[R], u01[R], um1[R], v00[R], v01[R];
            kt_rows<0, T, R>(own, wb, lane);
            kt_rows<2, T, R>(own, wt, lane);
            kt_rows<1, T, R>(kt_col<T, R>(Pj, 0, il + 1), u00, lane);
            kt_rows<1, T, R>(kt_col<T, R>(Pj, 0, il), um0, lane);
            kt_rows<2, T, R>(kt_col<T, R>(Pj, 0, il + 1), u01, lane);
            kt_rows<2, T, R>(kt_col<T, R>(Pj, 0, il), um1, lane);
            kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 1), v00, lane);
            kt_rows<2, T, R>(kt_col<T, R>(Pj, 1, il + 1), v01, lane);
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const T lap = (qE[r] - T(2) * qC[r] + qW[r]) * c.mt.kxx + (qN[r] - T(2) * qC[r] + qS[r]) * c.mt.kyy;
              const T fx = (h * (u00[r] + u01[r]) * (h * (qC[r] + qE[r])) - h * (um0[r] + um1[r]) * (h * (qW[r] + qC[r]))) * c.mt.rdx;
              const T fy = (h * (v00[r] + v01[r]) * (h * (qC[r] + qN[r])) - h * (ca[r] + cb[r]) * (h * (qS[r] + qC[r]))) * c.mt.rdy;
              const T wr = h * (qC[r] + wt[r]), wl = h * (wb[r] + qC[r]);
              const T fz = (wr * wr - wl * wl) * rz2[r];
              rhs[r] = lap - (fx + fy + fz);
            }
          }
        }
        T x[R];
        if (k1_rhs_ab2<T, R>(c, rhs, qC, rp, rq + oi + cj, bcol, n, Kc, true, x))
          k1_cn_store<T, R, FACE>(c, x, sq + oi + cj, n, rz0, rz1, rz2, T(0));
      }
    }
    // carry plane j into step j+1
    if (kind != 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) qS[r] = qC[r];
      if constexpr (COMP == 0) {
        kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 1), ca, lane);
        kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 2), cb, lane);
      } else if constexpr (COMP == 2) {
        kt_rows<1, T, R>(kt_col<T, R>(Pj, 1, il + 1), ca, lane);
        kt_rows<2, T, R>(kt_col<T, R>(Pj, 1, il + 1), cb, lane);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) rp[r] = rpn[r];
    kind = kind_n;
  }
}

// FULL: nz == 64*R, checked by the launcher (as for k_explicit_cn).
template <typename T, int R, bool FULL>
__global__ __launch_bounds__(64 * TC_KT_WAVES) void k_explicit_cn_tile(
    Dims d, Metric<T> mt, const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
    const T* __restrict__ pu, const T* __restrict__ pv, const T* __restrict__ pw, T* __restrict__ ru,
    T* __restrict__ rv, T* __restrict__ rw, T* __restrict__ su, T* __restrict__ sv, T* __restrict__ sw,
    const double* __restrict__ tab_cell, const double* __restrict__ tab_face,
    const double* __restrict__ scal, const T* __restrict__ body, int mode, int branch, double theta, int jb) {
  extern __shared__ __align__(16) unsigned char kt_lds[];
  if constexpr (FULL) d.nz = TC_WAVE * R;
  T* ring = reinterpret_cast<T*>(kt_lds);
  T* tabs = ring + KT<T, R>::NPL * KT<T, R>::PLANE;          // cell table, then face table
  T* rows = tabs + 2 * KT<T, R>::TAB;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int nbj = k1_bands_j(d.ny, jb);
  const int bi = b / nbj, bj = b - bi * nbj;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int comp = wv / TC_KT_I, il = wv - comp * TC_KT_I;
  const int i0 = bi * TC_KT_I, i = i0 + il;
  const int j0 = bj * jb, j1 = min(j0 + jb, d.ny + 2);
  const int ja = j0 > 0 ? j0 - 1 : 0;
  if (mode & TC_K1_STEP) {
    for (int e = threadIdx.x; e < KT<T, R>::TAB; e += 64 * TC_KT_WAVES) {
      tabs[e] = (T)k1_tab_src<R>(tab_cell, e);
      tabs[KT<T, R>::TAB + e] = (T)k1_tab_src<R>(tab_face, e);
    }
  }
  K1Ctx<T, R> c;
  c.d = d;
  c.mt = mt;
  c.rows = rows + wv * KT<T, R>::ROWS;
  c.tabl = comp < 2 ? tabs : tabs + KT<T, R>::TAB;
  c.lane = lane;
  c.mode = mode;
  c.cpu = branch == TC_BRANCH_CPU;
  const double dt = scal[TC_S_DT];
  c.dt = (T)dt;
  c.beta = (T)((1.0 - theta) * dt * mt.nu);
  c.F = (comp == 0 && (mode & TC_K1_FORCE)) ? (T)scal[TC_S_F] : T(0);
  c.body = (comp == 0 && (mode & TC_K1_FORCE)) ? body : nullptr;
  rows_zero_pads<R, T>(c.rows, lane);
  // the tables and pads are visible after the march's first barrier
  if (comp == 0) kt_march<T, R, 0>(c, i, il, i0, ja, j0, j1, ring, u, v, w, pu, ru, su);
  else if (comp == 1) kt_march<T, R, 1>(c, i, il, i0, ja, j0, j1, ring, u, v, w, pv, rv, sv);
  else kt_march<T, R, 2>(c, i, il, i0, ja, j0, j1, ring, u, v, w, pw, rw, sw);
}

// Ensemble form over the channels chans[0..n) (all of row count R); start = prefix sums of
// their block counts.  rhs set `cur` receives the new rhs, `prev` holds the previous one.
template <typename T, int R>
__global__ __launch_bounds__(64 * TC_K1_WAVES) void k_explicit_cn_ens(const EnsChan<T>* __restrict__ ch,
                                                                      const int* __restrict__ chans,
                                                                      const int* __restrict__ start, int n,
                                                                      int cur, int prev, int mode, int branch) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[chans[q]];   // fields are passed to the body by value
  T* const* rc = c.rs[cur];
  T* const* rp = c.rs[prev < 0 ? cur : prev];
  explicit_cn_body<T, R>(c.d, c.mt, c.u, c.v, c.w, rp[0], rp[1], rp[2], rc[0], rc[1], rc[2], c.su, c.sv, c.sw,
                         c.tabs, c.tabs + TC_TAB_STRIDE, c.scal, c.body, mode, branch, c.theta, c.jb,
                         gb - start[q]);
}

// Stand-alone CN solve of one component (solve_implicit_diffusion_u/v/w, operators.py:563-811):
// out's interior columns are replaced by the CN solution of q's columns; out's other entries
// are left untouched (the caller passes a clone of q).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_cn_only(Dims d, Metric<T> mt, int comp, const T* __restrict__ q,
                                                 T* __restrict__ out, const double* __restrict__ tab,
                                                 const double* __restrict__ scal, double theta) {
  __shared__ T lds_rows[4][3 * TC_WAVE * R];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = blockIdx.x * 4 + wv;
  const int lane = lane_id();
  const int i = 1 + g / d.ny, j = 1 + g % d.ny;
  if (i > d.nx) return;
  long long base;
  int n;
  if (comp == 0) { base = (long long)i * d.sxu + (long long)j * d.syu; n = d.nz; }
  else if (comp == 1) { base = (long long)i * d.sxv + (long long)j * d.syv; n = d.nz; }
  else { base = (long long)i * d.sxw + (long long)j * d.syw; n = d.nz - 1; }
  const T* qc = q + base;
  T* oc = out + base;
  const double dt = scal[TC_S_DT];
  T x[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    x[r] = (int)k <= n ? qc[k] : T(0);
  }
  T* rows = lds_rows[wv];
  rows_zero_pads<R, T>(rows, lane);
  const T beta = (T)((1.0 - theta) * dt * mt.nu);
  if (theta < 1.0) {
    if (comp < 2) cn_rhs_cell<R, T>(x, rows, mt, n, d.top, beta, lane);
    else cn_rhs_face<R, T>(x, rows, mt, n, qc[0], beta, lane);
  }
  pcr_shared_solve_lds<R, T>(x, rows, tab, n, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    if ((int)k <= n) oc[k] = x[r];
  }
}

// ---------------------------------------------------------------------------------------------
// K5: q = q* - dt*grad(p) (projection.py:154-163) written with every ghost set to its periodic
// image / wall reflection (apply_bc_all, solver.py:43-79), plus the per-block partial sums of the
// bulk velocity (utils.py:220-223), folded by k_step_finalize.
// Work decomposition as in K1: a block of TC_K5_WAVES wavefronts owns one component and a band of
// TC_K5_WAVES consecutive i; each wavefront walks TC_K5_JB consecutive j columns (ghost columns
// included, each reading its periodic source) with the next column's loads issued before the
// current one is written, so a wavefront keeps two columns of loads in flight.  The walk is
// specialised per component (its pointers are computed on straight-line paths).
// ---------------------------------------------------------------------------------------------
#define TC_K5_WAVES 4
#define TC_K5_JB 8
__host__ __device__ __forceinline__ int k5_bands_i(int nx) { return (nx + 2 + TC_K5_WAVES - 1) / TC_K5_WAVES; }
__host__ __device__ __forceinline__ int k5_bands_j(int ny) { return (ny + 2 + TC_K5_JB - 1) / TC_K5_JB; }

template <typename T, int R>
struct K5In {
  T s[R], a[R], b[R];      // star column; u, v: p and its +x / +y neighbour; w: p[k] and p[k+1]
};

// Loads of column (i, j) of component COMP: the star column and the two pressure columns the
// gradient needs, from the periodic source (si, sj) of a ghost column.
template <typename T, int R, int COMP>
__device__ __forceinline__ void k5_fetch(const Dims& d, int i, int j, const T* __restrict__ s,
                                         const T* __restrict__ p, int lane, K5In<T, R>& L) {
  const int si = (i == 0) ? d.nx : (i == d.nx + 1 ? 1 : i);
  const int sj = (j == 0) ? d.ny : (j == d.ny + 1 ? 1 : j);
  const T* pc = p + (long long)si * d.sxp + (long long)sj * d.syp;
  constexpr bool W = COMP == 2;
  const int n = W ? d.nz - 1 : d.nz;
  const long long sx = COMP == 0 ? d.sxu : (COMP == 1 ? d.sxv : d.sxw);
  const long long sy = COMP == 0 ? d.syu : (COMP == 1 ? d.syv : d.syw);
  const long long po = COMP == 0 ? d.sxp : (COMP == 1 ? d.syp : 1);
  const T* sc = s + si * sx + sj * sy;
  const T* pb = pc + po;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
    const bool ok = (int)k <= n;
    L.s[r] = ok ? sc[k] : T(0);
    L.a[r] = ok ? pc[k] : T(0);
    L.b[r] = ok ? pb[k] : T(0);
  }
}

template <typename T, int R, int COMP>
__device__ __forceinline__ void k5_pencil(const Dims& d, const Metric<T>& mt, int i, int j0, int j1,
                                          const T* __restrict__ s, const T* __restrict__ p, T* __restrict__ q,
                                          T dt, int lane, double& part, bool& bad) {
  const T rd = COMP == 0 ? mt.rdx : mt.rdy;
  const long long sx = COMP == 0 ? d.sxu : (COMP == 1 ? d.sxv : d.sxw);
  const long long sy = COMP == 0 ? d.syu : (COMP == 1 ? d.syv : d.syw);
  K5In<T, R> A, B;
  k5_fetch<T, R, COMP>(d, i, j0, s, p, lane, A);
  for (int j = j0; j < j1; ++j) {
    const bool more = j + 1 < j1;
    if (more) k5_fetch<T, R, COMP>(d, i, j + 1, s, p, lane, B);
    T* o = q + i * sx + j * sy;
    if constexpr (COMP < 2) {
      const bool bulk = COMP == 0 && i >= 1 && j >= 1 && j <= d.ny;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
        if ((int)k <= d.nz) {
          const T val = A.s[r] - dt * ((A.b[r] - A.a[r]) * rd);
          bad |= !isfinite(val);
          o[k] = val;
          if (k == 1) o[0] = -val;
          if ((int)k == d.nz) o[k + 1] = d.top > 0 ? val : -val;
          if (bulk) part += (double)val * mt.wbulk[k - 1];
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned k = (unsigned)(r * TC_WAVE + lane + 1);
        if ((int)k < d.nz) {
          const T val = A.s[r] - dt * ((A.b[r] - A.a[r]) * mt.rdzc[k]);
          bad |= !isfinite(val);
          o[k] = val;
        }
        if (k == 1) { o[0] = T(0); o[d.nz] = T(0); }
      }
    }
    if (more) A = B;
  }
}

template <typename T, int R>
__device__ __forceinline__ void project_bc_body(
    Dims d, Metric<T> mt, const T* __restrict__ su, const T* __restrict__ sv, const T* __restrict__ sw,
    const T* __restrict__ p, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
    double* __restrict__ scal, double* __restrict__ partials, int flags, int b) {
  __shared__ double red[TC_K5_WAVES];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = lane_id();
  const int comp = b % 3, band = b / 3;
  const int nbj = k5_bands_j(d.ny);
  const int bi = band / nbj, bj = band - bi * nbj;
  const int i = bi * TC_K5_WAVES + wv;
  const int imax = comp == 0 ? d.nx : d.nx + 1;
  const int jmax = comp == 1 ? d.ny : d.ny + 1;
  const int j0 = bj * TC_K5_JB;
  const int j1 = min(j0 + TC_K5_JB, jmax + 1);
  const T dt = (T)scal[TC_S_DT];
  double part = 0.0;
  bool bad = false;
  if (i <= imax && j0 <= jmax) {
    if (comp == 0) k5_pencil<T, R, 0>(d, mt, i, j0, j1, su, p, u, dt, lane, part, bad);
    else if (comp == 1) k5_pencil<T, R, 1>(d, mt, i, j0, j1, sv, p, v, dt, lane, part, bad);
    else k5_pencil<T, R, 2>(d, mt, i, j0, j1, sw, p, w, dt, lane, part, bad);
  }
  // device-side non-finite flag (the reference scans u, v, w, p on the host at save points,
  // solver.py:862-867); every writer stores the same value, so no atomic is needed
  if (__any(bad) && lane == 0) scal[TC_S_NONFINITE] = 1.0;
  if (!(flags & TC_K5_BULK)) return;
  part = wave_sum(part);
  if (lane == 0) red[wv] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < TC_K5_WAVES; ++q) tot += red[q];
    partials[b] = tot;
  }
}

template <typename T, int R>
__global__ __launch_bounds__(64 * TC_K5_WAVES) void k_project_bc(
    Dims d, Metric<T> mt, const T* __restrict__ su, const T* __restrict__ sv, const T* __restrict__ sw,
    const T* __restrict__ p, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
    double* __restrict__ scal, double* __restrict__ partials, int flags) {
  project_bc_body<T, R>(d, mt, su, sv, sw, p, u, v, w, scal, partials, flags, xcd_remap(blockIdx.x, gridDim.x));
}

template <typename T, int R>
__global__ __launch_bounds__(64 * TC_K5_WAVES) void k_project_bc_ens(const EnsChan<T>* __restrict__ ch,
                                                                     const int* __restrict__ chans,
                                                                     const int* __restrict__ start, int n,
                                                                     int flags) {
  const int gb = xcd_remap(blockIdx.x, gridDim.x);
  const int q = ens_chan(start, n, gb);
  const EnsChan<T>& c = ch[chans[q]];   // fields are passed to the body by value
  project_bc_body<T, R>(c.d, c.mt, c.su, c.sv, c.sw, c.p, c.u, c.v, c.w, c.scal, c.partials, flags, gb - start[q]);
}

// Folds the per-block bulk partials of k_project_bc in a fixed order (one block), then the
// forcing controller F += 0.1*(U_b - u_b)/dt (solver.py:701-708), time and step count.
__device__ __forceinline__ void step_finalize_body(const double* __restrict__ partials, int nb,
                                                   double* __restrict__ scal, double inv_volume, int flags) {
  __shared__ double acc[1024];
  double s = 0.0;
  const int t = threadIdx.x;
  int q = t;
  for (; q + 3 * 1024 < nb; q += 4 * 1024) {
    const double a0 = partials[q], a1 = partials[q + 1024], a2 = partials[q + 2048], a3 = partials[q + 3072];
    s += (a0 + a1) + (a2 + a3);
  }
  for (; q < nb; q += 1024) s += partials[q];
  acc[t] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) acc[t] += acc[t + o];
    __syncthreads();
  }
  if (t == 0) {
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

__global__ __launch_bounds__(1024) void k_step_finalize(const double* __restrict__ partials, int nb,
                                                        double* __restrict__ scal, double inv_volume,
                                                        int flags) {
  step_finalize_body(partials, nb, scal, inv_volume, flags);
}

// Ensemble form: block c folds channel c.
template <typename T>
__global__ __launch_bounds__(1024) void k_step_finalize_ens(const EnsChan<T>* __restrict__ ch, int flags) {
  const EnsChan<T>& c = ch[blockIdx.x];
  step_finalize_body(c.partials, c.k5_blocks, c.scal, c.inv_volume, flags);
}

// ---------------------------------------------------------------------------------------------
// Persistent ensemble form: one launch advances every channel of an ensemble `steps` IMEX steps.
//
// The work is a list of items (channel, phase, block of that phase), phases K1, P1 .. P5, K5 and
// the bulk fold, ordered step-major, then phase-major, then by channel.  Every block of the
// launch takes the next item from a ticket counter, waits until its channel has finished the
// phase before it (K1: the previous step's fold), runs the same device body as the per-phase
// kernels with that block index, and counts the item done.  Results are therefore bit-identical
// to the per-phase kernels and to stepping each channel alone; what changes is that a channel's
// next phase starts as soon as its own previous phase is done, while others are still running.
//
// Progress: an item waits only for items of smaller tickets, which blocks that are already
// running hold; the oldest unfinished item waits for nothing, so every wait ends.  Each wait is
// also bounded (spin_limit polls, about a second): past it the block sets the error word and goes
// on, and every later wait of the launch returns at once, so a defect shows as TC_ERR_TIMEOUT from
// tc_ensemble_check, never as a hung device.
// Memory ordering: each wave releases its stores at agent scope before the block's barrier (a
// barrier alone does not wait for other waves' stores), then one thread counts the item; a
// waiting thread acquires at agent scope before the block's barrier.
// sync: [0] ticket, [1] blocks exited, [2] error word, [3] unused, [4 + 8c + phase] items of
// that phase the channel has finished in this launch; the last block out zeroes [0, 1] and the
// counts for the next launch.
// ---------------------------------------------------------------------------------------------
#define TC_PH_K1 0
#define TC_PH_P1 1
#define TC_PH_P2 2
#define TC_PH_P3 3
#define TC_PH_P4 4
#define TC_PH_P5 5
#define TC_PH_K5 6
#define TC_PH_FIN 7
#define TC_NPH 8

// step_finalize_body's fold with 256 threads: the same 1024 partial sums (thread t of 256 forms
// those of t, t+256, t+512, t+768) and the same tree, so the result is bit-identical.  acc: 1024
// doubles of LDS.
__device__ __forceinline__ void step_finalize_body_256(const double* __restrict__ partials, int nb,
                                                       double* __restrict__ scal, double inv_volume, int flags,
                                                       double* acc) {
  for (int m = 0; m < 4; ++m) {
    const int t = threadIdx.x + 256 * m;
    double s = 0.0;
    int q = t;
    for (; q + 3 * 1024 < nb; q += 4 * 1024) {
      const double a0 = partials[q], a1 = partials[q + 1024], a2 = partials[q + 2048], a3 = partials[q + 3072];
      s += (a0 + a1) + (a2 + a3);
    }
    for (; q < nb; q += 1024) s += partials[q];
    acc[t] = s;
  }
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    for (int t = threadIdx.x; t < o; t += 256) acc[t] += acc[t + o];
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

// RMAX: largest rows per lane of a channel's explicit + CN and projection columns; THRM: rows per
// lane template of the wall-normal Poisson solve (the ensemble's largest).
template <typename T, int RMAX, int THRM>
__global__ __launch_bounds__(256) void k_ensemble_persistent(const EnsChan<T>* __restrict__ ch,
                                                             const int2* __restrict__ items, int per_step,
                                                             int steps, const int* __restrict__ nbt,
                                                             int* __restrict__ sync, int nchan, int first,
                                                             int newest0, int branch, int spin_limit) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_item;
  const int total = per_step * steps;
  const int cur0 = first ? 0 : 1 - newest0;
  for (;;) {
    if (threadIdx.x == 0) s_item = __hip_atomic_fetch_add(sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int g = s_item;
    if (g >= total) break;
    const int st = g / per_step;
    const int2 it = items[g - st * per_step];
    const int ph = it.x >> 16, c = it.x & 0xffff, lb = it.y;
    if (threadIdx.x == 0) {
      const int dep = ph == TC_PH_K1 ? TC_PH_FIN : ph - 1;
      const int target = ph == TC_PH_K1 ? st : (st + 1) * nbt[c * TC_NPH + dep];
      if (target > 0) {
        const int* ctr = sync + 4 + c * TC_NPH + dep;
        int spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins > spin_limit) {
            __hip_atomic_store(sync + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          // once any wait has overrun, no block waits any more: the launch drains at once
          if ((spins & 255) == 0 && __hip_atomic_load(sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
    }
    __syncthreads();
    const EnsChan<T>& cc = ch[c];   // fields are passed to the bodies by value
    const int R = (cc.d.nz + TC_WAVE - 1) / TC_WAVE;
    switch (ph) {
      case TC_PH_K1: {
        const int cur = cur0 ^ (st & 1);
        const int prev = (first && st == 0) ? -1 : 1 - cur;
        const int mode = TC_K1_RHS | TC_K1_STEP | TC_K1_FORCE | ((first && st == 0) ? TC_K1_FIRST : 0);
        T* const* rc = cc.rs[cur];
        T* const* rp = cc.rs[prev < 0 ? cur : prev];
#define TC_PK1(RR)                                                                                              \
  explicit_cn_body<T, RR, false, true>(cc.d, cc.mt, cc.u, cc.v, cc.w, rp[0], rp[1], rp[2], rc[0], rc[1], rc[2], \
                                       cc.su, cc.sv, cc.sw, cc.tabs, cc.tabs + TC_TAB_STRIDE, cc.scal, cc.body,  \
                                       mode, branch, cc.theta, cc.jb, lb, (T*)smem_raw)
        if (R == 1) TC_PK1(1);
        else if constexpr (RMAX >= 2) {
          if (R == 2) TC_PK1(2);
          else if constexpr (RMAX >= 3) {
            if (R == 3) TC_PK1(3);
            else if constexpr (RMAX >= 4) TC_PK1(4);
          }
        }
#undef TC_PK1
        break;
      }
      case TC_PH_P1:
        fwd_x_body<T, true>(cc.d, cc.mt, cc.su, cc.sv, cc.sw, nullptr, cc.d.nz, cc.scal, cc.fx, cc.twx, cc.spec,
                            cc.pairs, lb);
        break;
      case TC_PH_P2:
        fft_y_body<T>(cc.nkx, cc.d.ny, cc.d.nz, cc.fy, cc.twy, cc.spec, cc.kzy, -1, lb);
        break;
      case TC_PH_P3: {
        const int L = thomas_lanes(cc.d.nz);
        thomas_body<T, THRM>(cc.nkx * cc.d.ny, cc.d.nz, thomas_rows_per_lane(cc.d.nz), L, TH_THREADS / L, cc.cl,
                             cc.cr, cc.K2, cc.th_cb, cc.spec, lb);
        break;
      }
      case TC_PH_P4:
        fft_y_body<T>(cc.nkx, cc.d.ny, cc.d.nz, cc.fy, cc.twy, cc.spec, cc.kzy, +1, lb);
        break;
      case TC_PH_P5:
        inv_x_body<T>(cc.d, cc.fx, cc.twx, cc.spec, cc.p, (T)(1.0 / ((double)cc.d.nx * (double)cc.d.ny)), cc.pairs,
                      lb);
        break;
      case TC_PH_K5: {
#define TC_PK5(RR)                                                                                          \
  project_bc_body<T, RR>(cc.d, cc.mt, cc.su, cc.sv, cc.sw, cc.p, cc.u, cc.v, cc.w, cc.scal, cc.partials,    \
                         TC_K5_BULK | TC_K5_FORCING, lb)
        if (R == 1) TC_PK5(1);
        else if constexpr (RMAX >= 2) {
          if (R == 2) TC_PK5(2);
          else if constexpr (RMAX >= 3) {
            if (R == 3) TC_PK5(3);
            else if constexpr (RMAX >= 4) TC_PK5(4);
          }
        }
#undef TC_PK5
        break;
      }
      default:
        step_finalize_body_256(cc.partials, cc.k5_blocks, cc.scal, cc.inv_volume, TC_K5_BULK | TC_K5_FORCING,
                               (double*)smem_raw);
        break;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(sync + 4 + c * TC_NPH + ph, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) {
    const int out = __hip_atomic_fetch_add(sync + 1, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (out == (int)gridDim.x - 1) {
      for (int q = 0; q < TC_NPH * nchan; ++q) sync[4 + q] = 0;
      sync[0] = 0;
      sync[1] = 0;
    }
  }
}

// Dynamic LDS of the persistent kernel: the largest of the Poisson passes' LDS (lds), the
// per-component pencils' table and row buffers and the fold's 1024 partial sums.
size_t ens_persistent_lds(int dtype, int rmax, const size_t* lds) {
  size_t b = 1024 * sizeof(double);
  for (int q = 0; q < 4; ++q) b = std::max(b, lds[q]);
  const size_t k1 = dtype == TC_F64 ? (rmax <= 1 ? k1_block_lds<double, 1>() : rmax <= 2 ? k1_block_lds<double, 2>()
                                                  : rmax <= 3 ? k1_block_lds<double, 3>() : k1_block_lds<double, 4>())
                                    : (rmax <= 1 ? k1_block_lds<float, 1>() : rmax <= 2 ? k1_block_lds<float, 2>()
                                                  : rmax <= 3 ? k1_block_lds<float, 3>() : k1_block_lds<float, 4>());
  return std::max(b, k1);
}

template <typename T, int RMAX, int THRM>
static int ens_persistent_t(const void* dch, const int2* items, int per_step, int steps, const int* nbt, int* sync,
                            int nchan, int first, int newest0, int branch, size_t lds, int* grid, hipStream_t st) {
  auto kern = k_ensemble_persistent<T, RMAX, THRM>;
  if (*grid <= 0) {
    int dev = 0, ncu = 0, per = 0;
    TC_HIP_RC(hipGetDevice(&dev));
    TC_HIP_RC(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TC_HIP_RC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, lds));
    if (per < 1) return TC_ERR_UNSUPPORTED;
    *grid = per * ncu;
  }
  const int blocks = std::max(1, std::min(*grid, per_step * steps));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, st, (const EnsChan<T>*)dch, items, per_step, steps, nbt,
                     sync, nchan, first, newest0, branch, 1 << 20);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// One persistent launch of `steps` steps.  The instance follows the ensemble's largest column:
// rmax rows per lane (1, 2 or up to 4) and the Poisson solve's rows-per-lane template (4 up to
// 64 rows, else 8).  *grid: resident blocks, found on the first call (0) and kept by the caller.
int ens_launch_persistent(int dtype, int rmax, int thrm, const void* dch, const int2* items, int per_step, int steps,
                          const int* nbt, int* sync, int nchan, int first, int newest0, int branch, size_t lds,
                          int* grid, hipStream_t st) {
  if (lds > 160 * 1024) return TC_ERR_ARG;
#define TC_EP(TT, RM, TH)                                                                                     \
  return ens_persistent_t<TT, RM, TH>(dch, items, per_step, steps, nbt, sync, nchan, first, newest0, branch, lds,   \
                                      grid, st)
  if (dtype == TC_F64) {
    if (rmax <= 1 && thrm <= 4) TC_EP(double, 1, 4);
    if (rmax <= 2) TC_EP(double, 2, 8);
    TC_EP(double, 4, 8);
  }
  if (rmax <= 1 && thrm <= 4) TC_EP(float, 1, 4);
  if (rmax <= 2) TC_EP(float, 2, 8);
  TC_EP(float, 4, 8);
#undef TC_EP
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
// The tiled K1 needs KT<T, R>::BYTES of LDS per block; TC_ERR_UNSUPPORTED when the device's
// per-block limit is smaller (the caller then launches the pencils).
template <typename T, int R>
static int launch_tile(int blocks, hipStream_t st, Dims d, Metric<T> mt, const T* u, const T* v, const T* w,
                       const T* pu, const T* pv, const T* pw, T* ru, T* rv, T* rw, T* su, T* sv, T* sw,
                       const double* tab_cell, const double* tab_face, const double* scal, const T* body,
                       int mode, int branch, double theta, int jb) {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      lim = 0;
    limit = lim;
  }
  const size_t bytes = KT<T, R>::BYTES;
  if ((long long)bytes > (long long)limit) return TC_ERR_UNSUPPORTED;
  if (d.nz == TC_WAVE * R)
    hipLaunchKernelGGL((k_explicit_cn_tile<T, R, true>), dim3(blocks), dim3(64 * TC_KT_WAVES), bytes, st, d, mt, u,
                       v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, tab_cell, tab_face, scal, body, mode, branch, theta,
                       jb);
  else
    hipLaunchKernelGGL((k_explicit_cn_tile<T, R, false>), dim3(blocks), dim3(64 * TC_KT_WAVES), bytes, st, d, mt, u,
                       v, w, pu, pv, pw, ru, rv, rw, su, sv, sw, tab_cell, tab_face, scal, body, mode, branch, theta,
                       jb);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
static int launch_k1_t(const Grid* gr, const void* u, const void* v, const void* w, const void* pu,
                       const void* pv, const void* pw, void* ru, void* rv, void* rw, void* su,
                       void* sv, void* sw, const double* tab_cell, const double* tab_face,
                       const double* scal, const void* body, int mode, int branch, double theta,
                       hipStream_t st) {
  const Dims& d = gr->d;
  const int jb = gr->k1_pencil();
  const Metric<T> mt = gr->metric<T>();
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
  if (gr->k1_tiled()) {
    const int tblocks = k1_bands_i(d.nx) * k1_bands_j(d.ny, jb);
    int rc = TC_ERR_UNSUPPORTED;
#define TC_KTL(RR)                                                                                          \
  rc = launch_tile<T, RR>(tblocks, st, d, mt, (const T*)u, (const T*)v, (const T*)w, (const T*)pu,            \
                          (const T*)pv, (const T*)pw, (T*)ru, (T*)rv, (T*)rw, (T*)su, (T*)sv, (T*)sw, tab_cell, \
                          tab_face, scal, (const T*)body, mode, branch, theta, jb)
    switch (R) {
      case 1: TC_KTL(1); break;
      case 2: TC_KTL(2); break;
      case 3: if constexpr (sizeof(T) == 4) TC_KTL(3); break;
      case 4: if constexpr (sizeof(T) == 4) TC_KTL(4); break;
      default: break;
    }
#undef TC_KTL
    if (rc != TC_ERR_UNSUPPORTED) return rc;      // else the block's LDS does not fit: pencils
  }
  if (gr->k1_form() == 2) {
    // three-component pencils: static LDS (both CN tables and the row buffers) within 64 KiB
    const bool full = d.nz == TC_WAVE * R;
    bool launched = true;
#define TC_K3(RR, F)                                                                                    \
  hipLaunchKernelGGL((k_explicit_cn_uvw<T, RR, F>),                                                    \
                     dim3(k3_bands_i(d.nx, k3_waves<T, RR>()) * k1_bands_j(d.ny, jb)),                 \
                     dim3(64 * k3_waves<T, RR>()), 0, st, d, mt,                                        \
                     (const T*)u, (const T*)v, (const T*)w, (const T*)pu, (const T*)pv, (const T*)pw,     \
                     (T*)ru, (T*)rv, (T*)rw, (T*)su, (T*)sv, (T*)sw, tab_cell, tab_face, scal,          \
                     (const T*)body, mode, branch, theta, jb)
    switch (R) {
      case 1: if (full) TC_K3(1, true); else TC_K3(1, false); break;
      case 2: if (full) TC_K3(2, true); else TC_K3(2, false); break;
      case 3: if constexpr (sizeof(T) == 4) { if (full) TC_K3(3, true); else TC_K3(3, false); } else launched = false; break;
      case 4: if constexpr (sizeof(T) == 4) { if (full) TC_K3(4, true); else TC_K3(4, false); } else launched = false; break;
      default: launched = false; break;
    }
#undef TC_K3
    if (launched) {
      TC_CHECK_LAUNCH();
      return TC_OK;
    }
  }
  const int blocks = 3 * k1_bands_i(d.nx) * k1_bands_j(d.ny, jb);
#define TC_K1(RR, F)                                                                                \
  hipLaunchKernelGGL((k_explicit_cn<T, RR, F>), dim3(blocks), dim3(64 * TC_K1_WAVES), 0, st, d, mt,  \
                     (const T*)u, (const T*)v, (const T*)w, (const T*)pu, (const T*)pv,               \
                     (const T*)pw, (T*)ru, (T*)rv, (T*)rw, (T*)su, (T*)sv, (T*)sw, tab_cell,          \
                     tab_face, scal, (const T*)body, mode, branch, theta, jb)
  const bool full = d.nz == TC_WAVE * R;
  switch (R) {
    case 1: if (full) TC_K1(1, true); else TC_K1(1, false); break;
    case 2: if (full) TC_K1(2, true); else TC_K1(2, false); break;
    case 3: if (full) TC_K1(3, true); else TC_K1(3, false); break;
    case 4: if (full) TC_K1(4, true); else TC_K1(4, false); break;
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
                       void* u, void* v, void* w, double* scal, double* partials,
                       double inv_volume, int flags, hipStream_t st) {
  const Dims& d = gr->d;
  const int blocks = k5_blocks(gr);
  const Metric<T> mt = gr->metric<T>();
  const int R = (d.nz + TC_WAVE - 1) / TC_WAVE;
#define TC_K5(RR)                                                                                \
  hipLaunchKernelGGL((k_project_bc<T, RR>), dim3(blocks), dim3(64 * TC_K5_WAVES), 0, st, d, mt, (const T*)su,  \
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
    hipLaunchKernelGGL(k_step_finalize, dim3(1), dim3(1024), 0, st, partials, blocks, scal, inv_volume, flags);
    TC_CHECK_LAUNCH();
  }
  return TC_OK;
}

int k5_blocks(const Grid* gr) { return 3 * k5_bands_i(gr->d.nx) * k5_bands_j(gr->d.ny); }

int launch_k5(const Grid* gr, const void* su, const void* sv, const void* sw, const void* p, void* u,
              void* v, void* w, double* scal, double* partials, double inv_volume,
              int flags, hipStream_t st) {
  if (gr->dtype == TC_F64)
    return launch_k5_t<double>(gr, su, sv, sw, p, u, v, w, scal, partials, inv_volume, flags, st);
  return launch_k5_t<float>(gr, su, sv, sw, p, u, v, w, scal, partials, inv_volume, flags, st);
}

}  // namespace tc

// ---------------------------------------------------------------------------------------------
// Ensemble launchers (one launch per phase over every channel of the batch)
// ---------------------------------------------------------------------------------------------
namespace tc {

int ens_launch_tables(int dtype, const void* dch, int n, double dt_set, hipStream_t st) {
  if (dtype == TC_F64)
    hipLaunchKernelGGL((k_cn_tables_ens<double>), dim3(2 * n), dim3(TC_MAX_NZ), 0, st,
                       (const EnsChan<double>*)dch, dt_set);
  else
    hipLaunchKernelGGL((k_cn_tables_ens<float>), dim3(2 * n), dim3(TC_MAX_NZ), 0, st,
                       (const EnsChan<float>*)dch, dt_set);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

template <typename T>
static int ens_k1_t(int R, const void* dch, const int* chans, const int* start, int ng, int blocks, int cur,
                    int prev, int mode, int branch, hipStream_t st) {
  const EnsChan<T>* c = (const EnsChan<T>*)dch;
#define TC_EK1(RR)                                                                                         \
  hipLaunchKernelGGL((k_explicit_cn_ens<T, RR>), dim3(blocks), dim3(64 * TC_K1_WAVES), 0, st, c, chans,   \
                     start, ng, cur, prev, mode, branch)
  switch (R) {
    case 1: TC_EK1(1); break;
    case 2: TC_EK1(2); break;
    case 3: TC_EK1(3); break;
    case 4: TC_EK1(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_EK1
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int ens_launch_k1(int dtype, int R, const void* dch, const int* chans, const int* start, int ng, int blocks,
                  int cur, int prev, int mode, int branch, hipStream_t st) {
  if (dtype == TC_F64) return ens_k1_t<double>(R, dch, chans, start, ng, blocks, cur, prev, mode, branch, st);
  return ens_k1_t<float>(R, dch, chans, start, ng, blocks, cur, prev, mode, branch, st);
}

template <typename T>
static int ens_k5_t(int R, const void* dch, const int* chans, const int* start, int ng, int blocks, int flags,
                    hipStream_t st) {
  const EnsChan<T>* c = (const EnsChan<T>*)dch;
#define TC_EK5(RR)                                                                                         \
  hipLaunchKernelGGL((k_project_bc_ens<T, RR>), dim3(blocks), dim3(64 * TC_K5_WAVES), 0, st, c, chans,    \
                     start, ng, flags)
  switch (R) {
    case 1: TC_EK5(1); break;
    case 2: TC_EK5(2); break;
    case 3: TC_EK5(3); break;
    case 4: TC_EK5(4); break;
    default: return TC_ERR_UNSUPPORTED;
  }
#undef TC_EK5
  TC_CHECK_LAUNCH();
  return TC_OK;
}

int ens_launch_k5(int dtype, int R, const void* dch, const int* chans, const int* start, int ng, int blocks,
                  int flags, hipStream_t st) {
  if (dtype == TC_F64) return ens_k5_t<double>(R, dch, chans, start, ng, blocks, flags, st);
  return ens_k5_t<float>(R, dch, chans, start, ng, blocks, flags, st);
}

int ens_launch_finalize(int dtype, const void* dch, int n, int flags, hipStream_t st) {
  if (dtype == TC_F64)
    hipLaunchKernelGGL((k_step_finalize_ens<double>), dim3(n), dim3(1024), 0, st, (const EnsChan<double>*)dch,
                       flags);
  else
    hipLaunchKernelGGL((k_step_finalize_ens<float>), dim3(n), dim3(1024), 0, st, (const EnsChan<float>*)dch,
                       flags);
  TC_CHECK_LAUNCH();
  return TC_OK;
}

// Blocks of the per-component pencils (the form the ensemble kernels run).
int k1_blocks(const Grid* g) {
  const int jb = g->k1_pencil_of(0);
  return 3 * k1_bands_i(g->d.nx) * k1_bands_j(g->d.ny, jb);
}

}  // namespace tc
