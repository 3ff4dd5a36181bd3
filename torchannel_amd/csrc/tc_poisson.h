// Poisson plan (host object) shared by the stand-alone solve, the fused step and the statistics.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "tc_fft.h"
#include "tc_internal.h"

namespace tc {

struct Poisson {
  const Grid* g = nullptr;
  int nkx = 0;                 // nx/2 + 1
  FftDesc fx, fy;
  double2* twx = nullptr;      // exp(-2 pi i t/nx)
  double2* twy = nullptr;
  double* cl = nullptr;        // 1/(dz_c[k] dz_f[k])     (projection_fft.py:48)
  double* cr = nullptr;        // 1/(dz_c[k+1] dz_f[k])   (projection_fft.py:49)
  double* K2 = nullptr;        // kx~^2 + ky~^2 per (m, j) (projection_fft.py:30-39)
  void* spec = nullptr;        // S[nkx][ny][nz] complex, working precision
  int pairs = 1;               // z-row pairs per x-transform block
  int pairs_fx = 1;            // ... of the solve's forward x transform
  int wide_fx = 0;             // that transform's budget: 0 TC_FFT_ELEMS, 1 TC_FFT_ELEMS_WIDE,
                               // 2 TC_FFT_ELEMS_WIDE2 (float32 only)
  int kzy = 1;                 // z-rows per y-transform block
  double* th_cb = nullptr;     // Thomas c' of the row before each lane's first row of the
                               // wall-normal solve, [system][lanes] (0 for lane 0)
  double* planes = nullptr;    // statistics: (nx, ny, 6) float64
  void* plane_spec = nullptr;  // statistics: (nkx, ny, 6) complex float64
};

int poisson_init(Poisson* pl, const Grid* g, const double* dzf_host, const double* dzc_host);
void poisson_free(Poisson* pl);
int poisson_solve(const Poisson* pl, bool from_vel, const void* u, const void* v, const void* w,
                  const void* rin, const double* scal, void* p, hipStream_t st);
int poisson_phase(const Poisson* pl, int ph, bool from_vel, const void* u, const void* v, const void* w,
                  const void* rin, const double* scal, void* p, hipStream_t st);
int plane_fft(const Poisson* pl, hipStream_t st);
int thomas_batched(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                   long long s, const void* d, void* x, void* work, hipStream_t st);
// The same sweep with c' in the library's static device scratch (no workspace argument, no
// allocation); batch*n above the scratch capacity returns TC_ERR_SCRATCH.
int thomas_batched_static(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                          long long s, const void* d, void* x, hipStream_t st);
int pcr_batched(int n, int batch, int dtype, int cplx_d, const void* a, const void* b, const void* c,
                long long s, const void* d, void* x, hipStream_t st);

}  // namespace tc
