// Ensembles of independent channels advanced by one launch per phase: every per-channel kernel
// body is entered either from its single-channel wrapper or from a batched wrapper that maps the
// global block index to (channel, block of that channel) through a prefix sum of block counts.
#pragma once
#include "tc_fft.h"
#include "tc_internal.h"

namespace tc {

// Everything one channel's step kernels read, in device memory (one entry per channel).
template <typename T>
struct EnsChan {
  Dims d;
  Metric<T> mt;
  T *u, *v, *w, *p, *su, *sv, *sw;
  T* rs[2][3];               // the two rhs sets of the AB2 history
  double* scal;              // the channel's scalar block (TC_S_*)
  double* tabs;              // CN tables: cell, face (TC_TAB_STRIDE each)
  double* partials;          // bulk partials of the projection kernel
  const T* body;             // optional body force (shaped like u) or null
  // Poisson plan
  FftDesc fx, fy;
  const double2 *twx, *twy;
  cplx<T>* spec;
  const double *cl, *cr, *K2;
  const double* th_cb;
  int pairs, kzy, nkx;
  // launch shapes and scalars
  int jb, k5_blocks;
  double inv_volume;
  double nu, theta;
  int top;
};

// Channel owning global block gb: the largest c with start[c] <= gb (start[0] = 0).
__device__ __forceinline__ int ens_chan(const int* __restrict__ start, int nchan, int gb) {
  int lo = 0, hi = nchan - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= gb) lo = mid;
    else hi = mid - 1;
  }
  // gb is the same for the whole block, so lo is too; saying so lets the descriptor loads that
  // follow go through scalar registers
  return __builtin_amdgcn_readfirstlane(lo);
}

}  // namespace tc

namespace tc {
struct Poisson;
int ens_launch_tables(int dtype, const void* dch, int n, double dt_set, hipStream_t st);
int ens_launch_k1(int dtype, int R, const void* dch, const int* chans, const int* start, int ng, int blocks,
                  int cur, int prev, int mode, int branch, hipStream_t st);
int ens_launch_k5(int dtype, int R, const void* dch, const int* chans, const int* start, int ng, int blocks,
                  int flags, hipStream_t st);
int ens_launch_finalize(int dtype, const void* dch, int n, int flags, hipStream_t st);
int k1_blocks(const Grid* g);
// blocks[4] / lds[4]: forward x FFT, y FFT (both directions), Thomas, inverse x FFT
void poisson_geometry(const Poisson* pl, int tsz, int* blocks, size_t* lds);
int ens_launch_poisson(int dtype, const void* dch, int n, const int* const* start, const int* blocks,
                       const size_t* lds, int th_rm, hipStream_t st);
int thomas_rows(int nz);   // rows per lane of the wall-normal Poisson solve
// persistent form (tc_step.hip): phases of a work item, its LDS and its launch
#define TC_ENS_NPH 8
size_t ens_persistent_lds(int dtype, int rmax, const size_t* lds);
int ens_launch_persistent(int dtype, int rmax, int thrm, const void* dch, const int2* items, int per_step, int steps,
                          const int* nbt, int* sync, int nchan, int first, int newest0, int branch, size_t lds,
                          int* grid, hipStream_t st);
}  // namespace tc
