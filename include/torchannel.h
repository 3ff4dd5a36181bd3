/* torchannel.h -- C ABI of the MI355X (gfx950) IMEX time step of torChannel.
 *
 * Every entry point takes plain device pointers to the reference's ghosted staggered arrays
 * (z-contiguous C order, the layout torch.zeros(nx+1, ny+2, nz+2) gives):
 *
 *   u (nx+1, ny+2, nz+2)   v (nx+2, ny+1, nz+2)   w (nx+2, ny+2, nz+1)   p (nx+2, ny+2, nz+2)
 *   div (nx, ny, nz)
 *
 * in float64 (TC_F64) or float32 (TC_F32), a hipStream_t passed as `void*`, and returns 0 on
 * success, a negative hipError_t on a HIP failure, or a TC_ERR_* code.  tc_strerror() maps
 * codes to text.  The library allocates only its own grid tables, Poisson plans, graphs and the
 * adjoint's and ensembles' buffers (freed by the matching *_destroy); a solver's scratch comes
 * from a workspace the caller allocates (tc_solver_workspace_size).  It never allocates or frees
 * caller memory.  Calls that share a grid, Poisson plan or solver object must be issued on one
 * stream at a time.
 *
 * Reference citations are to gmcavallazzi/torChannel (file:line).
 */
#ifndef TORCHANNEL_H
#define TORCHANNEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- codes and enums ------------------------------------------------------------------- */
#define TC_OK 0
#define TC_ERR_ARG (-1000)          /* invalid argument */
#define TC_ERR_SHAPE (-1001)        /* unsupported grid size */
#define TC_ERR_DTYPE (-1002)        /* unsupported dtype */
#define TC_ERR_UNSUPPORTED (-1003)  /* size outside the compiled kernel range */
#define TC_ERR_NOMEM (-1004)        /* library workspace allocation failed */
#define TC_ERR_STATE (-1005)        /* AB2 history parity differs from the captured graph's */
#define TC_ERR_SCRATCH (-1006)      /* batch*n above the static scratch of tc_thomas_batched_static */
#define TC_ERR_TIMEOUT (-1007)      /* a persistent ensemble launch's dependency wait overran its bound */
#define TC_ERR_DISABLED (-1008)     /* a code path that is built but switched off */

#define TC_TOP_DIRICHLET 0          /* no-slip top wall  (solver.py:53-54) */
#define TC_TOP_NEUMANN 1            /* free-slip top wall (solver.py:51-52) */

#define TC_BRANCH_GPU 0             /* fused explicit RHS (solver.py:593-600) */
#define TC_BRANCH_CPU 1             /* separate RHS kernels (solver.py:601-622) */

#define TC_F64 0
#define TC_F32 1

#define TC_NSCAL 8                  /* length of a solver's float64 scalar block */

typedef struct tc_grid tc_grid;
typedef struct tc_poisson tc_poisson;
typedef struct tc_solver tc_solver;

const char* tc_strerror(int code);
int tc_version(void);

/* ---- grid -------------------------------------------------------------------------------
 * Extents, spacings, viscosity and the wall-normal metric (host float64 arrays dz_f[nz],
 * dz_c[nz+1] as generate_grid returns them, utils.py:5-45).  Replaces the (nx, ny, nz, dx, dy,
 * dz_c, dz_f, nu, top_wall_bc_type) argument tuple threaded through operators.py:1038-1044 and
 * the ChannelFlow attributes of solver.py:111-234.  nz in [4, 256]; nx, ny even, of the form
 * 2^a 3^b 5^c, in [4, 1536] for the Poisson plan and the solver. */
int tc_grid_create(int nx, int ny, int nz, double dx, double dy, double nu, int top_bc, int dtype,
                   const double* dz_f_host, const double* dz_c_host, tc_grid** out);
int tc_grid_destroy(tc_grid* g);

/* Launch-shape knobs of a grid's kernels (no counterpart in the reference: its kernels are
 * TorchScript-fused).  Results do not depend on them; torchannel_amd/autotune.py picks them by
 * timing.  TC_TUNE_K1_JB: j-columns per explicit/CN pencil, 1..ny+2 (0 restores the default:
 * for the per-component pencils about ny/24 in steps of 4 within [8, 16]; for the
 * three-component pencils 8 in float64, in float32 4 below ny = 256 and 6 from there).
 * TC_TUNE_THOMAS_SB: (kx, ky) systems per block of the wall-normal Poisson solve (16 lanes each,
 * 32 above 128 rows): 4, 8 or 16, at most 256 threads per block (0 restores the default, 16).
 * TC_TUNE_K1_TILE: form of the explicit + CN kernel: 0 one component per block, each wavefront
 * fetching its neighbour columns; 1 the three components of a band in one block with the
 * neighbour columns staged through LDS (float64 up to 128 rows, float32 up to 256; larger
 * columns, or a device with less LDS per block, run the first form); 2 one wavefront per i
 * marching u, v and w together (float64 up to 128 rows, float32 up to 256, else the first form).
 * Default (-1, or never set): 2 where it is built and a lane holds two or more rows, else 0.
 * TC_TUNE_THOMAS_LDS: form of the wall-normal Poisson solve: 0 each lane loads its own rows; 1 (the
 * default) a block's systems are copied into LDS and back with consecutive 16-byte accesses
 * (where the rows are 16-byte vectors; otherwise the first form runs). */
#define TC_TUNE_K1_JB 1
#define TC_TUNE_THOMAS_SB 2
#define TC_TUNE_K1_TILE 3
#define TC_TUNE_THOMAS_LDS 4
int tc_grid_set_tuning(tc_grid* g, int key, int value);
int tc_grid_get_tuning(const tc_grid* g, int key, int* value);

/* Column stride of the arrays a grid's calls take (no counterpart in the reference, whose arrays
 * are always dense torch.zeros tensors, initflow.py:6-9).  pitch 0 (the default): each array is
 * dense, its columns nz+2 (u, v, p) or nz+1 (w) elements apart.  pitch >= nz+2: the columns of
 * u, v, w and p all sit `pitch` elements apart, with the same (i, j, k) indexing; element
 * [i][j][k] is at i*(columns per plane)*pitch + j*pitch + k.  With pitch a multiple of 128 bytes
 * and row k = 1 of the first column on a 128-byte line, every column's interior starts on a
 * line, which the column kernels stream faster (ChannelFlow's compute.layout 'padded').  Buffers the
 * library lays out itself (a solver's post-CN velocities, the adjoint's arrays) follow the
 * pitch.  Set it before a solver is created on the grid (TC_ERR_STATE after). */
int tc_grid_set_pitch(tc_grid* g, int pitch);
int tc_grid_get_pitch(const tc_grid* g, int* pitch);

/* apply_bc_all (solver.py:23-79): periodic x/y images, then wall reflections, in place. */
int tc_apply_bc(const tc_grid* g, void* u, void* v, void* w, void* stream);

/* compute_momentum_rhs_fused_imex (operators.py:1037-1225) for branch TC_BRANCH_GPU, or
 * diffusion_xy_* - advection_* (operators.py:126-247, 474-548) for TC_BRANCH_CPU.  Writes
 * every entry of ru, rv, rw (zero outside the interior). */
int tc_rhs_imex(const tc_grid* g, int branch, const void* u, const void* v, const void* w,
                void* ru, void* rv, void* rw, void* stream);
/* compute_momentum_rhs_fused_v2 (operators.py:817-1034; the explicit right-hand side of the
 * reference's FE scheme, solver.py:498-534): advection and the full 3-D diffusion, including
 * its w control-volume quirk (dz_c[k-1], operators.py:1021-1026).  Writes every entry of ru,
 * rv, rw (zero outside the interiors). */
int tc_rhs_fused_v2(const tc_grid* g, const void* u, const void* v, const void* w, void* ru, void* rv,
                    void* rw, void* stream);

/* solve_implicit_diffusion_u / _v / _w (operators.py:563-811) for comp 0 / 1 / 2: the interior
 * columns of q_out receive the theta-scheme solution for q_in; other entries of q_out are not
 * written (the reference returns a clone, operators.py:580).  dt is a host value, or, when
 * dt_dev is non-null, read from device memory (a 0-D tensor, solver.py:659). */
int tc_cn_solve(const tc_grid* g, int comp, double dt, const double* dt_dev, double theta,
                const void* q_in, void* q_out, void* stream);

/* pcr_solve (tridiag.py:33-58) / solve_tridiagonal_batch (operators.py:551-560) /
 * solve_tridiagonal (projection_fft.py:167-215): batch systems of n rows,
 * a[i] x[i-1] + b[i] x[i] + c[i] x[i+1] = d[i].  a, b, c are real with a batch stride of 0
 * (shared rows) or n; d and x are (batch, n), complex interleaved when is_complex.
 * n in [1, 512]. */
int tc_pcr_batched(int n, int batch, int dtype, int is_complex, const void* a, const void* b,
                   const void* c, int64_t abc_stride, const void* d, void* x, void* stream);

/* solve_tridiagonal (projection_fft.py:167-215): the same systems as tc_pcr_batched solved by the
 * reference's Thomas sweep, one thread per system; work holds batch*n reals (c'). */
int tc_thomas_batched(int n, int batch, int dtype, int is_complex, const void* a, const void* b,
                      const void* c, int64_t abc_stride, const void* d, void* x, void* work,
                      void* stream);

/* Allocation-free form of tc_thomas_batched.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same sweep and results as tc_thomas_batched, with c' in
 * a static device array of the library (TC_THOMAS_STATIC_MAX_F64 float64 words = 32 MiB per GPU,
 * held until the process ends) instead of a `work` argument: batch*n at most
 * TC_THOMAS_STATIC_MAX_F64 in float64 and twice that in float32, else TC_ERR_SCRATCH.  Calls on
 * one GPU share that array, so calls that overlap on one GPU (several streams or threads) race.
 * No hipMalloc, hipFree or allocating library call is made. */
#define TC_THOMAS_STATIC_MAX_F64 (1 << 22)
int tc_thomas_batched_static(int n, int batch, int dtype, int is_complex, const void* a, const void* b,
                             const void* c, int64_t abc_stride, const void* d, void* x, void* stream);

/* compute_divergence (utils.py:226-240). */
int tc_divergence(const tc_grid* g, const void* u, const void* v, const void* w, void* div,
                  void* stream);

/* project_velocity (projection.py:143-165), in place on the interior; dt from dt_dev when
 * non-null, else the host value. */
int tc_project(const tc_grid* g, void* u, void* v, void* w, const void* p, double dt,
               const double* dt_dev, void* stream);

/* compute_bulk_velocity (utils.py:207-223): sum(u[1:nx+1,1:ny+1,1:nz+1] * wgt) / total_volume
 * with wgt a float64 (nx, ny, nz) view given by element strides (0 for expanded dims).
 * Result (float64) to *out_dev. */
int tc_bulk_velocity(const tc_grid* g, const void* u, const double* wgt, int64_t ws0, int64_t ws1,
                     int64_t ws2, double total_volume, double* out_dev, void* stream);

/* compute_cfl_fused (operators.py:1232-1333): the largest convective rate, to *out_dev. */
int tc_cfl_rate(const tc_grid* g, const void* u, const void* v, const void* w, double* out_dev,
                void* stream);

/* compute_u_tau (utils.py:176-204) with dist = z_c[1]; result to *out_dev. */
int tc_u_tau(const tc_grid* g, const void* u, double z_c1, double* out_dev, void* stream);

/* ---- FFT + tridiagonal pressure Poisson solve ------------------------------------------
 * initialize_fft_solver (projection_fft.py:3-99): plan with modified wavenumbers, per-mode
 * tridiagonal rows and the spectral workspace.  tc_poisson_solve is solve_poisson_fft
 * (projection_fft.py:101-165): p (ghosts included) from rhs (nx, ny, nz).  The singular
 * (kx,ky)=(0,0) mode is gauge-fixed in float64 (its first unknown pinned to zero); p is defined
 * up to an additive constant, as in the reference.  The wall-normal systems are solved by the
 * reference's Thomas sweep with the factorisation precomputed per plan. */
int tc_poisson_create(const tc_grid* g, tc_poisson** out);
int tc_poisson_destroy(tc_poisson* pl);
int tc_poisson_solve(tc_poisson* pl, const void* rhs, void* p, void* stream);

/* ---- turbulence statistics (turbstats.py:101-228) ------------------------------------------
 * Adds one sample to float64 running sums: prof[5*nz] = U, uu, vv, ww, uw profiles and
 * spec[4*(nx/2)*(ny/2)] = E_uu, E_vv, E_ww, E_uw folded 2-D spectra sampled at the planes
 * k_bot, k_top (1-based cell indices, turbstats.py:64-69). */
int tc_stats_accumulate(tc_poisson* pl, const void* u, const void* v, const void* w, int k_bot,
                        int k_top, double* prof, double* spec, void* stream);

/* ---- fused IMEX step (ChannelFlow.step_imex, solver.py:580-710) ----------------------------
 * The solver owns its workspaces and advances caller-owned state in place:
 *   u, v, w, p   ghosted fields (updated every step, ghosts included);
 *   rhs_a/rhs_b  two sets of three rhs arrays shaped like u, v, w (AB2 history, swapped each
 *                step exactly like rhs_*_prev / rhs_*_curr, solver.py:629-651);
 *   scal         float64[TC_NSCAL] device scalar block: dt, forcing F, u_bulk, time, cfl rate,
 *                U_bulk target, step count, non-finite flag (see TC_S_* in the library sources).
 * body (optional, shaped like u) is an additive body force next to F in rhs_u.
 * Volume V = Lx*Ly*Lz (solver.py:234).  branch selects the explicit-RHS semantics of the
 * reference's device branches (TC_BRANCH_GPU: solver.py:593-600, TC_BRANCH_CPU: 601-622).
 * workspace: caller-allocated device memory of at least tc_solver_workspace_size bytes, 256-byte
 * aligned, alive and unshared until tc_solver_destroy (it holds the post-CN velocities, the CN
 * tables and the reduction partials; the solver itself allocates only its capture stream and
 * graphs, and the grid and Poisson plan hold the rest). */
int tc_solver_workspace_size(const tc_grid* g, size_t* bytes);
int tc_solver_create(const tc_grid* g, tc_poisson* pl, void* u, void* v, void* w, void* p,
                     void* const* rhs_a, void* const* rhs_b, double* scal, double volume,
                     int branch, void* workspace, size_t workspace_bytes, tc_solver** out);
int tc_solver_destroy(tc_solver* s);
/* One IMEX step.  dt > 0 sets scal[dt] first; dt <= 0 keeps the device value.  first != 0 is
 * the forward-Euler start (no rhs history).  Returns which rhs set now holds the newest rhs
 * (0 = rhs_a, 1 = rhs_b) through *newest (may be null). */
int tc_solver_step(tc_solver* s, double dt, int first, const void* body, int* newest, void* stream);
/* ChannelFlow.__init__'s start (solver.py:268-276, 342-369): when rescale != 0, u is scaled to
 * the target bulk velocity (skipped on restart, solver.py:271); then div -> Poisson with dt = 1
 * -> project -> boundary conditions. */
int tc_solver_project_initial(tc_solver* s, int rescale, void* stream);
/* The rescale alone (solver.py:271-276): u *= U_bulk / u_b when |u_b| > 1e-9, so that a caller
 * can save the rescaled, not yet projected state (fields_init.npz, solver.py:337-340) before
 * tc_solver_project_initial(s, 0, ..). */
int tc_solver_rescale(tc_solver* s, void* stream);
/* Adaptive dt on the device (solver.py:465-496 and the 5% rule of 807-812): reads the state,
 * writes the new dt into scal when it changed by more than 5%. */
int tc_solver_update_dt(tc_solver* s, double cfl_target, double dt_min, double dt_max, void* stream);
/* Device time-series row of run_simulation's n_out diagnostics (solver.py:823-842, which reads
 * them on the host each time): appends [step, time, u_bulk, u_tau, F, dt, max|div|, CFL rate] as
 * row rows[0] of the device buffer `rows` (rows[0] counts the rows; row r occupies
 * rows[8 (r + 1) .. 8 (r + 1) + 7]; rows past `capacity` are counted but not written).  u_tau is
 * compute_u_tau (utils.py:176-204) with wall distance z_c1, max|div| is over
 * compute_divergence (utils.py:226-240).  No host synchronisation. */
int tc_solver_diagnostics(tc_solver* s, double z_c1, double* rows, int capacity, void* stream);
/* Capture `steps` IMEX steps (even; AB2 history must exist) into a HIP graph, with the adaptive
 * dt update before every `dt_interval`-th step of the chunk when dt_interval > 0.  Replay
 * advances the state by that many steps on `stream`; it returns TC_ERR_STATE when the newest rhs
 * set is not the one it was at capture (an odd number of eager steps in between). */
int tc_solver_capture(tc_solver* s, int steps, int dt_interval, double cfl_target, double dt_min,
                      double dt_max, void* stream);
int tc_solver_replay(tc_solver* s, void* stream);
/* Launch one kernel of the step `reps` times back to back, for per-kernel timing (no reference
 * counterpart): 0 CN tables, 1 explicit+AB2+CN, 2 forward x FFT (with divergence), 3 forward y
 * FFT, 4 per-mode tridiagonal solve, 5 inverse y FFT, 6 inverse x FFT (+ p ghosts), 7 projection + boundary +
 * bulk, 8 CFL reduction.  Overwrites the step's intermediates; needs a step taken first. */
int tc_solver_phase(tc_solver* s, int phase, int reps, void* stream);
/* Set which rhs set holds the newest rhs (0 or 1), or -1 so the next step is a forward-Euler
 * start: restoring a checkpointed state (rhs arrays included) into the solver's buffers.  The
 * reference keeps this as the rhs_prev / rhs_curr swap of solver.py:649-651. */
int tc_solver_set_newest(tc_solver* s, int newest);

/* ---- roofline calibration ------------------------------------------------------------------
 * No counterpart in the reference.  Sweeps twelve arrays shaped like p ((nx+2) x (ny+2) x
 * (nz+2), columns `pitch` elements apart, 0 meaning nz+2; arrays[0..5] read, arrays[6..11]
 * written) in the column order of the explicit + CN
 * kernel without its arithmetic: pattern 0 moves the per-component pencils' eight input and two
 * output columns per column step (neighbour columns included), pattern 1 only the compulsory
 * streams (one input column and the AB2 history in, two columns out), pattern 2 the
 * three-component pencils' nine input, three history and six output columns per (i, j) step.
 * Timed beside the kernel, it gives the bandwidth each access pattern reaches on the device.
 * nz <= 256. */
int tc_calib_columns(int dtype, int pattern, int nx, int ny, int nz, int jb, int pitch, void* const* arrays,
                     void* stream);

/* ---- discrete adjoint --------------------------------------------------------------------
 * No counterpart in the reference.  Gradient of J = sum_n weight_n * tau_w(u_n) with respect to
 * a body-force field f (shaped like u) added to rhs_u next to the forcing (solver.py:626), for a
 * float64 GPU-branch solver stepped at fixed dt with tc_solver_step(.., body = f, ..).
 * tau_w(u) = nu/z_c[1] * (mean u[:,:,1] + mean u[:,:,nz]) / 2 over whole planes, ghosts
 * included (the linear core of compute_u_tau, utils.py:187-204); the free-slip top keeps the
 * bottom term only, nu/z_c[1] * mean u[:,:,1].
 * Reverse sweep: tc_adjoint_begin(dt, weight of u_N, grad) then, for n = N-1 .. 0,
 * tc_adjoint_step(q_n, first = (n == 0), weight of u_n, grad) with q_n the full state arrays at
 * the start of forward step n.  grad (shaped like u) receives dJ/df. */
typedef struct tc_adjoint tc_adjoint;
int tc_adjoint_create(tc_solver* s, double wall_dist, tc_adjoint** out);
int tc_adjoint_destroy(tc_adjoint* a);
/* j_dev[0] += weight * tau_w(u) (device float64 scalar). */
int tc_adjoint_objective(tc_adjoint* a, const void* u, double weight, double* j_dev, void* stream);
int tc_adjoint_begin(tc_adjoint* a, double dt, double obj_weight, void* grad, void* stream);
int tc_adjoint_step(tc_adjoint* a, const void* u, const void* v, const void* w, int first, double obj_weight,
                    void* grad, void* stream);

/* ---- ensembles ---------------------------------------------------------------------------
 * No counterpart in the reference (one channel per process there).  Independent channels of
 * any sizes (each a tc_solver with its own grid and Poisson plan; one dtype and one rhs branch)
 * advance together: every phase of the step is one launch over all channels, blocks mapped to
 * (channel, block) by prefix sums, so small channels share the GPU instead of each launching
 * its own under-filled kernels.  Results are bit-identical to tc_solver_step on each channel.
 * The channels' launch-shape knobs are read at creation.  dt > 0 sets every channel's dt, else
 * each keeps its device dt; `first` takes the forward-Euler start (histories must be in step). */
typedef struct tc_ensemble tc_ensemble;
int tc_ensemble_create(tc_solver* const* solvers, int n, tc_ensemble** out);
int tc_ensemble_destroy(tc_ensemble* e);
int tc_ensemble_step(tc_ensemble* e, double dt, int first, void* stream);
/* `steps` (even) ensemble steps captured into one HIP graph; replay advances every channel.
 * Step and replay return TC_ERR_STATE when a channel's AB2 history is no longer in step with
 * the ensemble's (a channel stepped on its own in between), or, for replay, when the newest rhs
 * set differs from the one at capture. */
int tc_ensemble_capture(tc_ensemble* e, int steps, void* stream);
int tc_ensemble_replay(tc_ensemble* e, void* stream);
/* Form of tc_ensemble_step and of the captured chunks.  TC_ENS_PHASED (the default): one launch
 * per phase over all channels.  TC_ENS_PERSISTENT: one persistent launch per step (per chunk
 * when captured: a chunk is the CN-table build and one launch that advances every channel all of
 * the chunk's steps); its blocks take (channel, phase, block) items from a ticket counter and
 * each item waits only for its own channel's previous phase, so a channel moves on while others
 * finish theirs.  Same device code per item.  The persistent form is built but switched off
 * (TC_ERR_DISABLED): on gfx950 its first run made no progress for minutes, cause not established.
 * Setting the form drops a captured graph. */
#define TC_ENS_PHASED 0
#define TC_ENS_PERSISTENT 1
int tc_ensemble_set_form(tc_ensemble* e, int form);
/* Waits for `stream`, then reports TC_ERR_TIMEOUT (and clears it) if a persistent launch since
 * the last check had a dependency wait overrun its bound (its results are then not to be used);
 * TC_OK otherwise.  Every wait is bounded, so such a launch still ends. */
int tc_ensemble_check(tc_ensemble* e, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TORCHANNEL_H */
