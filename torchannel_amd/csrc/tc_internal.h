// Library-internal declarations: the grid object, scalar-block layout, kernel mode bits and the
// launchers shared between the per-operator entry points and the fused step.
#pragma once
#include "tc_common.h"

// Layout of the float64 scalar block a solver keeps in device memory (length TC_NSCAL).
#define TC_S_DT 0        // current time step
#define TC_S_F 1         // bulk forcing (lagged P-controller state, solver.py:707-708)
#define TC_S_UBULK 2     // bulk velocity after the last step
#define TC_S_TIME 3      // simulated time
#define TC_S_CFL 4       // last convective rate (1/dt) from the CFL reduction
#define TC_S_UTARGET 5   // target bulk velocity U_bulk
#define TC_S_STEP 6      // steps taken (as a double)
#define TC_S_NONFINITE 7   // set to 1 by the projection kernel when a non-finite velocity appears
// adjoint scalar block
#define TC_A_LF 0        // cotangent of the forcing F
#define TC_A_NSCAL 4

// k_explicit_cn modes
#define TC_K1_RHS 1
#define TC_K1_STEP 2
#define TC_K1_FIRST 4
#define TC_K1_FORCE 8

// k_project_bc flags
#define TC_K5_BULK 1
#define TC_K5_FORCING 2

namespace tc {

// Host-side grid object: extents, spacings and device copies of the wall-normal metric.
struct Grid {
  Dims d;
  int dtype;
  double dx, dy, nu;
  // float64 device tables
  double* dzf64 = nullptr;   // nz
  double* dzc64 = nullptr;   // nz+1
  double* rdzf64 = nullptr;
  double* rdzc64 = nullptr;
  double* wbulk = nullptr;   // dx*dy*dz_f
  // float32 copies of the reciprocals
  float* rdzf32 = nullptr;
  float* rdzc32 = nullptr;
  // scratch for the stand-alone reductions and CN tables
  double* scratch = nullptr;      // partial sums / maxima
  int scratch_len = 0;
  double* tabs = nullptr;         // CN tables: cell, face
  double* scal = nullptr;         // scalar block for stand-alone ops
  int k1_jb = 0;                  // j-columns per K1 pencil (0: the default below; set by the tuner)
  int k1_tile = -1;               // K1 form: 0 pencils, 1 tiled (LDS-staged neighbour columns),
                                  // 2 three-component pencils, -1 the default below
  int thomas_sb = 0;              // systems per wall-normal Thomas block (0: plan default)
  int thomas_lds = 1;             // wall-normal Thomas form: 0 direct row loads, 1 staged through LDS
  int pitch = 0;                  // column stride of every array (0: dense, tc_grid_set_pitch)
  mutable bool has_solver = false;  // a solver's workspace was laid out for this pitch

  // Elements ahead of an array's first entry in a buffer the library lays out itself (the
  // solver's post-CN velocities, the adjoint's arrays): with a padded pitch, row k = 1 of every
  // column then starts on a 128-byte line when the buffer starts on one.
  int lead() const { return pitch > 0 ? 128 / (dtype == TC_F64 ? 8 : 4) - 1 : 0; }

  // K1 form: the tuned value, else the three-component pencils where they are built for the row
  // count (float64 up to 128 rows, float32 up to 256) and a lane holds two or more rows (measured
  // 6-8 % faster than the per-component pencils at 192x192x128 in both precisions and
  // 384x384x256 in float32), else the per-component pencils.
  int k1_form() const {
    if (k1_tile >= 0) return k1_tile;
    const int rows = (d.nz + 63) / 64;
    return rows >= 2 && rows <= (dtype == TC_F64 ? 2 : 4) ? 2 : 0;
  }
  bool k1_tiled() const { return k1_form() == 1; }
  // K1 pencil length of a form: the tuned value, else for the per-component pencils about ny/24
  // in steps of 4, within [8, 16] (ny = 192 fastest at 8, ny = 384 at 16, by 4-6 % over 8), and
  // for the three-component pencils 8 in float64, in float32 4 below ny = 256 and 6 from there
  // (192x192x128: 8 beats 4 by 2 % in float64, 4 beats 8 by 4 % in float32; 384x384x256 float32:
  // 6 beats 4 and 16 by 3-4 %).
  int k1_pencil_of(int form) const {
    if (k1_jb > 0) return k1_jb;
    if (form == 2) return dtype == TC_F64 ? 8 : (d.ny >= 256 ? 6 : 4);
    return std::min(16, std::max(8, 4 * ((d.ny + 48) / 96)));
  }
  int k1_pencil() const { return k1_pencil_of(k1_form()); }

  template <typename T>
  Metric<T> metric() const;
};

template <>
inline Metric<double> Grid::metric<double>() const {
  Metric<double> m;
  m.rdzf = rdzf64; m.rdzc = rdzc64; m.dzf = dzf64; m.dzc = dzc64; m.wbulk = wbulk;
  m.rdx = 1.0 / dx; m.rdy = 1.0 / dy;
  m.kxx = nu * (1.0 / (dx * dx)); m.kyy = nu * (1.0 / (dy * dy)); m.nu = nu;
  return m;
}

template <>
inline Metric<float> Grid::metric<float>() const {
  Metric<float> m;
  m.rdzf = rdzf32; m.rdzc = rdzc32; m.dzf = dzf64; m.dzc = dzc64; m.wbulk = wbulk;
  m.rdx = (float)(1.0 / dx); m.rdy = (float)(1.0 / dy);
  m.kxx = (float)(nu * (1.0 / (dx * dx))); m.kyy = (float)(nu * (1.0 / (dy * dy))); m.nu = nu;
  return m;
}

// tc_step.hip
int launch_tables(const Grid* gr, double* scal, double dt_set, double theta, double* tab_cell,
                  double* tab_face, hipStream_t st);
int launch_k1(const Grid* gr, const void* u, const void* v, const void* w, const void* pu,
              const void* pv, const void* pw, void* ru, void* rv, void* rw, void* su, void* sv,
              void* sw, const double* tab_cell, const double* tab_face, const double* scal,
              const void* body, int mode, int branch, double theta, hipStream_t st);
int launch_cn_only(const Grid* gr, int comp, const void* q, void* out, const double* tab,
                   const double* scal, double theta, hipStream_t st);
int launch_k5(const Grid* gr, const void* su, const void* sv, const void* sw, const void* p, void* u,
              void* v, void* w, double* scal, double* partials, double inv_volume,
              int flags, hipStream_t st);
int k5_blocks(const Grid* gr);

// tc_ops.hip
int launch_apply_bc(const Grid* gr, void* u, void* v, void* w, hipStream_t st);
int launch_divergence(const Grid* gr, const void* u, const void* v, const void* w, void* div,
                      hipStream_t st);
int launch_project(const Grid* gr, void* u, void* v, void* w, const void* p, double dt_host,
                   const double* dt_dev, hipStream_t st);
int launch_bulk(const Grid* gr, const void* u, const double* wgt, long long ws0, long long ws1,
                long long ws2, double inv_vol, double* out, double* partials,
                int nblocks_max, hipStream_t st);
int launch_rhs_v2(const Grid* gr, const void* u, const void* v, const void* w, void* ru, void* rv, void* rw,
                  hipStream_t st);
int launch_div_max(const Grid* gr, const void* u, const void* v, const void* w, double* out, double* partials,
                   int nblocks_max, hipStream_t st);
int launch_diag_row(const double* scal, double* rows, int capacity, hipStream_t st);
int launch_u_tau(const Grid* gr, const void* u, double dist, double* out, double* partials,
                 int nblocks_max, hipStream_t st);
int launch_cfl(const Grid* gr, const void* u, const void* v, const void* w, double* scal,
               double* partials, int update_dt, double cfl_target, double dt_min,
               double dt_max, int nblocks_max, hipStream_t st);

// tc_adjoint.hip (float64)
int adj_launch_pre(const Grid* g, double rdt, double inv_volume, const double* ascal, void* const* lq,
                   void* const* ys, hipStream_t st);
int adj_cn_blocks(const Grid* g);
int adj_launch_cn(const Grid* g, void* const* z, const double* tab_cell, const double* tab_face, double dt,
                  double c1, double c0, void* const* qs, void* const* lr, void* const* acc, void* grad,
                  double* partials, hipStream_t st);
int adj_launch_lamf(const double* partials, int nb, double* ascal, hipStream_t st);
int adj_launch_rhs(const Grid* g, const void* u, const void* v, const void* w, void* const* lr, void* const* e,
                   hipStream_t st);
int adj_launch_fold(const Grid* g, void* const* qs, void* const* e, double obj_coef, int obj_top, void* const* lq,
                    hipStream_t st);
int adj_tauw_blocks(const Grid* g);
int adj_launch_tauw(const Grid* g, const void* u, int top, double scale, double* partials, double* out,
                    hipStream_t st);

}  // namespace tc
