/*
 * azeban_hip.h -- C ABI of the MI355X (gfx950) spectral-viscosity Euler RHS and
 * SSP-RK3 time stepper.
 *
 * The entry points replace the device seams of TobiasRohner/azeban's
 * IncompressibleEuler<Dim, SmoothCutoff1D, NoForcing>::dudt + SSP_RK3::integrate
 * path.  Reference interfaces each one stands in for are cited per declaration
 * (paths relative to the reference repository root).
 *
 * Conventions
 *  - All arrays are device pointers unless the name ends in _host.
 *  - Fourier state u_hat: interleaved complex double (double2 / azeban complex_t),
 *    shape (n_vars, N, ..., N/2+1), row-major, components outermost
 *    (include/azeban/grid.hpp:126-133).
 *  - FFTs are unnormalised; forward is e^{-ikx}, backward e^{+ikx}
 *    (src/azeban/operations/fftwfft.cpp:321-350).
 *  - Every function returns AZB_OK (0) or a negative error code; the message is
 *    available from azb_last_error().  Calls are asynchronous on the solver's
 *    stream unless documented as synchronising.
 */
#ifndef AZEBAN_HIP_H_
#define AZEBAN_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AZB_OK 0
#define AZB_ERR_ARG -1
#define AZB_ERR_HIP -2
#define AZB_ERR_UNSUPPORTED -3
#define AZB_ERR_ALLOC -4

/* precision modes */
#define AZB_FP64 0        /* everything in FP64 (the drop-in path) */
#define AZB_FP32_MIXED 1  /* FFTs and products in FP32; RK state, reductions, SV weights FP64 */

typedef struct azb_solver azb_solver;

/* SmoothCutoff1D{eps, s, k0}: nu(k) = -eps k^{2s} (1 - exp(-(|k|/k0)^18))
 * (include/azeban/equations/spectral_viscosity.hpp:56-77). */
typedef struct {
  double eps;
  double s;
  double k0;
} azb_sv_params;

/* Error string of the last failing call on this thread. */
const char* azb_last_error(void);

/* Build an equation + integrator for a dim-D grid of N points per axis padded to
 * M = 3N/2 (the default of include/azeban/grid.hpp:44-49 and
 * include/azeban/grid_factory.hpp:28-92).  Allocates the device workspace and
 * the RK3 stage buffers; this stands in for the IncompressibleEuler constructor
 * (include/azeban/equations/incompressible_euler.hpp:56-96) plus the SSP_RK3
 * constructor (include/azeban/evolution/ssp_rk3.hpp:36-39).  stream is the
 * hipStream_t all of the solver's work is queued on; NULL selects the default
 * (null) stream. */
int azb_create(int dim, int N, int M, int n_vars, int precision, azb_sv_params visc,
               void* stream, azb_solver** out);
/* The same for `nsamples` independent solves of one grid size, stacked as
 * (nsamples, n_vars, N, ..., N/2+1): every kernel launch covers all samples, and
 * each sample keeps its own max|u|, dt, time and completion flag on the device.
 * This batches azeban's sequential sample loop (src/azeban/run_from_config.cpp:81-102)
 * into one launch per kernel.  Per-sample results are identical to solving each
 * sample alone. */
int azb_create_batch(int dim, int N, int M, int n_vars, int precision, azb_sv_params visc,
                     int nsamples, void* stream, azb_solver** out);
int azb_destroy(azb_solver* s);
int azb_num_samples(const azb_solver* s);
/* Device bytes azb_create_batch would allocate (admission control). */
size_t azb_bytes_estimate(int dim, int N, int precision, int nsamples);

/* Bytes of device memory held by the solver (workspace + stage buffers + twiddles). */
size_t azb_workspace_bytes(const azb_solver* s);
/* The solver's HIP stream (hipStream_t). */
void* azb_stream(const azb_solver* s);

/* dudt_hat = L(u_hat): one RHS evaluation, IncompressibleEuler::dudt
 * (include/azeban/equations/incompressible_euler.hpp:103-120).  dudt may alias
 * u_hat.  Also records max|u| / M^d for azb_dt. */
int azb_dudt(azb_solver* s, const void* u_hat, void* dudt_hat);

/* dudt_hat = L(u_hat) like azb_dudt, with no solver handle and no allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The intermediates live in a
 * file-scope device array of 96 MiB that each GPU holds until the process ends
 * (its twiddle tables are written on the first call on a device), so only
 * 2D N <= 1024 and 3D N <= 64 fit (larger grids: AZB_ERR_UNSUPPORTED).  Queued on
 * `stream` (NULL = null stream).  The first call on a device looks up the array's
 * address and uploads the twiddle tables with a synchronous copy, so it must not
 * be made under graph capture; later calls only launch kernels.  Every call on one
 * GPU uses the same array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be ordered by the caller.  Results are bit-identical to
 * azb_dudt's. */
int azb_dudt_noalloc(int dim, int N, int precision, azb_sv_params visc, const void* u_hat, void* dudt_hat,
                     void* stream);

/* C / (N u_max) of the latest RHS evaluation, IncompressibleEuler::dt
 * (incompressible_euler.hpp:122-124).  Synchronising. */
int azb_dt(azb_solver* s, double C, double* dt_host);
/* max|u| / M^d of the latest RHS evaluation.  Synchronising. */
int azb_u_max(azb_solver* s, double* u_max_host);

/* One SSP-RK3 step in place on u_hat with dt = min(C/(N u_max(stage 1)), max_dt),
 * SSP_RK3::integrate (include/azeban/evolution/ssp_rk3.hpp:48-64).  dt is taken on
 * the device; dt_host (if not NULL) receives it (synchronising). */
int azb_rk3_step(azb_solver* s, void* u_hat, double max_dt, double C, double* dt_host);

/* Simulation::simulate_until (src/azeban/simulation.cpp:89-102): while time < t_end,
 * take an RK3 step with max_dt = min(t_end - time, C*2/(eps (pi N)^2)).  The loop
 * runs on the device; the host checks completion every `chunk` steps.  On return
 * *time_inout = t_end (the reference sets time_ = t after its loop) and *steps_out
 * holds the number of steps taken.  max_steps > 0 stops early (then *time_inout is
 * the time reached).  Synchronising. */
int azb_simulate_until(azb_solver* s, void* u_hat, double t_end, double C, double* time_inout,
                       int max_steps, int chunk, int* steps_out);

/* ---- device-driven time loop, piecewise (for graph capture and overlap) ----
 * azb_loop_begin sets every sample's time (times0_host[i], or 0 if NULL) and target
 * t_end (synchronising); azb_enqueue_steps queues n steps without synchronising (a
 * step past t_end does nothing on the device, so extra steps are harmless);
 * azb_loop_read copies each sample's time and step count back (synchronising). */
int azb_loop_begin(azb_solver* s, double t_end, double C, const double* times0_host);
/* (Re)start one sample of a batched solver without touching the others and without
 * synchronising: queued on the solver's stream, so it takes effect after the steps
 * already queued there and before the next ones.  Sets that sample's time to time0,
 * its target to t_end and its step count to 0.  Every sample of a solver shares one
 * CFL number C (AZB_ERR_ARG otherwise).  This is the per-request start of the
 * continuous-batching server: each request is one Simulation::simulate_until
 * (src/azeban/simulation.cpp:89-102) in a slot of a batched solver. */
int azb_loop_begin_sample(azb_solver* s, int sample, double t_end, double C, double time0);
int azb_enqueue_steps(azb_solver* s, void* u_hat, int nsteps);
int azb_loop_read(azb_solver* s, double* times_host, int* steps_host);
/* Non-blocking variant: azb_loop_probe queues a copy of every sample's time and
 * step count into pinned host memory; azb_loop_probe_result sets *ready = 0 until
 * that copy has landed, then *ready = 1 and fills the arrays. */
int azb_loop_probe(azb_solver* s);
int azb_loop_probe_result(azb_solver* s, int* ready, double* times_host, int* steps_host);

/* ---- hipGraph capture of work queued on a stream (solvers created on it) ---- */
typedef struct azb_graph azb_graph;
int azb_capture_begin(void* stream);
int azb_capture_end(void* stream, azb_graph** out);
int azb_graph_launch(azb_graph* g, void* stream);
int azb_graph_destroy(azb_graph* g);
/* non-blocking HIP stream at default priority, for capture and overlap */
int azb_stream_create(void** stream);
int azb_stream_destroy(void* stream);
int azb_stream_synchronize(void* stream);
/* Work queued on `waiter` after this call starts only once everything queued on
 * `waitee` so far has finished (an event record + wait; inside a capture this joins
 * the two streams' branches of the graph).  Does not block the host. */
int azb_stream_wait(void* waiter, void* waitee);

/* ---- per-kernel-family device timing (hipEvent pairs on the solver's stream),
 * the analogue of azeban's ProfileDevice records (include/azeban/profiler.hpp:35-135). */
#define AZB_PROF_INV_AX0 0  /* inverse complex pass along axis 0, N -> M rows */
#define AZB_PROF_INV_AX1 1  /* inverse complex pass along axis 1 (3D), N -> M rows */
#define AZB_PROF_ROW 2      /* last axis: c2r, u_i u_j products, max|u|, r2c */
#define AZB_PROF_FWD_AX1 3  /* forward complex pass along axis 1 (3D), M -> N rows */
#define AZB_PROF_FWD_AX0 4  /* forward complex pass along axis 0, M -> N rows */
#define AZB_PROF_KSPACE 5   /* ik.B_hat, Leray, viscosity, RK stage update */
#define AZB_PROF_FAMILIES 6
int azb_profile_enable(azb_solver* s, int on);
/* accumulated milliseconds and launch counts per family since the last reset;
 * synchronising */
int azb_profile_read(azb_solver* s, double* ms_host, long* launches_host, int n);
int azb_profile_reset(azb_solver* s);

/* ---- autotuning of the kernel launch shapes --------------------------------
 * Per kernel family (AZB_PROF_*): the column passes take tiles of TS0, TS0/2 or
 * TS0/4 columns (tsdiv 1, 2, 4; TS0 = one 128-byte line), the row kernel P0 or
 * P0/2 line pairs per workgroup (pdiv 1, 2) where it has two shapes (3D N = 64,
 * and FP32 at 3D N = 128; every other row kernel has one).  azb_autotune times
 * every candidate on the device with the given state (reps RHS evaluations each,
 * output into the solver's own buffers) and keeps the fastest per family.
 * Results do not depend on the choice.  Synchronising. */
int azb_autotune(azb_solver* s, const void* u_hat, int reps);
int azb_set_tuning(azb_solver* s, const int* tsdiv, int pdiv);
int azb_get_tuning(const azb_solver* s, int* tsdiv, int* pdiv);
/* Radix plans of the column passes (per family; the row and k-space entries are 0).
 * Padded lengths 384, 768 and 1536 have three: plan 0 = 8.4.12 / 8.8.12 / 8.8.24,
 * plan 1 = 12.4.8 / 12.8.8 / 24.8.8, plan 2 = 4.4.24 / 4.8.24 / 8.24.8; the other
 * lengths one.  Unlike the launch shapes, the plan changes the rounding (results
 * agree to round-off, not bit for bit), so solves that must agree bit for bit (a
 * batch and its single solves) use the same plans.  azb_autotune_plans times each
 * plan on the device per family (reps RHS evaluations into the solver's own
 * buffers) and keeps the fastest.  The reference's nearest counterpart is its FFT
 * size choice from a benchmark file (src/azeban/operations/fft.cpp:63-108).
 * Synchronising. */
int azb_num_plans(const azb_solver* s);
int azb_set_plans(azb_solver* s, const int* plans);
int azb_get_plans(const azb_solver* s, int* plans);
int azb_autotune_plans(azb_solver* s, const void* u_hat, int reps);

/* ---- standalone operations (test seams) -------------------------------------- */

/* Rank-dim R2C, unnormalised, batch `howmany`, contiguous per batch entry:
 * in real (howmany, n, .., n) -> out complex (howmany, n, .., n/2+1).
 * precision selects double or float arrays.  FFTWFFT_R2C/CUFFT_R2C::forward
 * (src/azeban/operations/fftwfft.cpp:258-262, src/azeban/cuda/operations/cufft.cpp:134-140). */
int azb_fft_r2c(int dim, int n, int howmany, int precision, const void* in, void* out, void* stream);
/* Rank-dim C2R, unnormalised; complex passes over the leading axes (in place, the
 * input is overwritten as FFTW's and cuFFT's c2r may do), then c2r over the last
 * axis, which ignores Im at k_last = 0 and n/2.  FFTWFFT_R2C/CUFFT_R2C::backward
 * (fftwfft.cpp:265-269, cufft.cpp:143-149). */
int azb_fft_c2r(int dim, int n, int howmany, int precision, void* in, void* out, void* stream);

/* The same transforms from a caller-owned work area, for FFT<Dim>::set_work_area /
 * get_work_area_size (include/azeban/operations/fft.hpp:88-90,
 * src/azeban/cuda/operations/cufft.cpp:129-140,251-262).  The work area of a length-n
 * transform holds its twiddle table: azb_fft_work_bytes gives its size (0 for a length
 * or precision that is not built), azb_fft_work_init fills it (synchronising), and
 * transforms of the same n and precision may share one.  The _ws transforms allocate
 * nothing and do not synchronise (queued on `stream`). */
size_t azb_fft_work_bytes(int n, int precision);
int azb_fft_work_init(void* work, int n, int precision);
int azb_fft_r2c_ws(int dim, int n, int howmany, int precision, const void* in, void* out, void* work,
                   void* stream);
int azb_fft_c2r_ws(int dim, int n, int howmany, int precision, void* in, void* out, void* work,
                   void* stream);

/* One component N -> M zero padding, copy_to_padded
 * (src/azeban/operations/copy_to_padded.cpp:301-328,
 *  include/azeban/operations/copy_to_padded_functions.hpp:11-63). */
int azb_copy_to_padded(int dim, int N, int M, const void* in, void* out, void* stream);

/* One component M -> N truncation, the inverse of azb_copy_to_padded: copy_from_padded
 * (src/azeban/operations/copy_from_padded.cpp:28-86,
 *  include/azeban/operations/copy_from_padded_functions.hpp:11-27).  Used to write
 * downsampled output (src/azeban/io/netcdf_sample_writer.cpp:143-163). */
int azb_copy_from_padded(int dim, int M, int N, const void* in, void* out, void* stream);

/* x <- a x over n doubles, scale (include/azeban/operations/scale.hpp); the physical
 * output path scales u_hat by 1/N^dim before the inverse transform
 * (src/azeban/run_from_config.cpp:93-98). */
int azb_scale(double a, void* x, long n, void* stream);

/* y <- a x + b y over n complex doubles, axpby (include/azeban/operations/axpby.hpp:28-56). */
int azb_axpby(double a, const void* x, double b, void* y, long n, void* stream);

/* max |x| over n doubles, max_norm (include/azeban/operations/norm.hpp:72-91).  Synchronising. */
int azb_max_norm(const void* x, long n, double* out_host, void* stream);

/* Leray projection in place on u_hat (dim components on the N-grid), k = 0 set to 0
 * (src/azeban/operations/leray.cpp:28-116; used by Initializer::initialize,
 * include/azeban/init/initializer.hpp:68-75). */
int azb_leray(int dim, int N, void* u_hat, void* stream);

/* Energy spectrum E(K), K = 0..N/2, over L-infinity shells of u_hat into spec_host
 * (N/2+1 doubles), energy_spectrum (src/azeban/operations/energy_spectrum.cpp:10-50,
 * include/azeban/operations/spectrum.hpp:33-118).  Synchronising. */
int azb_energy_spectrum(int dim, int N, const void* u_hat, double* spec_host, void* stream);

/* ---- memory helpers (so callers need no HIP headers) ------------------------- */
int azb_malloc(void** ptr, size_t bytes);
int azb_free(void* ptr);
int azb_memcpy_h2d(void* dst, const void* src_host, size_t bytes);
int azb_memcpy_d2h(void* dst_host, const void* src, size_t bytes);
int azb_memcpy_d2d(void* dst, const void* src, size_t bytes);
int azb_memcpy_h2d_async(void* dst, const void* src_host, size_t bytes, void* stream);
int azb_memcpy_d2h_async(void* dst_host, const void* src, size_t bytes, void* stream);
int azb_synchronize(azb_solver* s);
int azb_device_synchronize(void);

/* Achievable HBM bandwidth on the current device: a streaming copy (one 16-byte
 * load and non-temporal store per thread) between two fresh buffers of gib GiB
 * (0 < gib <= 32), reps times, timed with events; *gbs_out = bytes read + written
 * per second / 1e9.  Synchronising.  No counterpart in the reference: the measured
 * roofline beside the 8 TB/s nominal one. */
int azb_copy_bandwidth(double gib, int reps, double* gbs_out);

#ifdef __cplusplus
}
#endif

#endif /* AZEBAN_HIP_H_ */
