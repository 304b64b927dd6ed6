/*
 * fdfault_hip -- C ABI of the MI355X (gfx950) time-stepping core of fdfault.
 *
 * The reference solver (edaub/fdfault) has no library seam: its hot path is
 * the C++ member function domain::do_rk_stage (src/domain.hpp:26,
 * src/domain.cpp:228-305), driven by problem::solve (src/problem.cpp:136-192)
 * and reached from the `fdfault <file.in>` executable (src/main.cpp:8-72).
 * Each entry point below states which reference interface it replaces.  All
 * entry points return 0 on success and a negative code on error; the message
 * the reference would have printed before MPI_Abort is available from
 * fdf_last_error().  A null handle, or a null pointer where a buffer, box or
 * path is required, is such an error, not a crash.  No C++ exception crosses
 * this boundary.  Handles are independent; one host thread per handle.
 */
#ifndef FDFAULT_HIP_H
#define FDFAULT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdf_s fdf_t;

/* precision of the bulk fields f/df; fault state is always FP64 */
enum { FDF_FP64 = 0, FDF_FP32 = 1 };

typedef struct fdf_opts {
  int precision;     /* FDF_FP64 (default) or FDF_FP32                          */
  int device;        /* HIP device ordinal                                      */
  int write_outputs; /* 1: open the .in output list and front files (like the
                        reference executable); 0: state stays on the device   */
  int quiet;         /* 1: no stdout progress lines                             */
  int use_graph;     /* 1: capture the per-step launch sequence in a hipGraph   */
  int exact_fp64;    /* 1: every FP64 stencil keeps the reference's per-point
                        association (bit-faithful); 0: Cartesian homogeneous
                        blocks without dissipation (3-D, and 2-D mode II
                        elastic) run on the tiled kernels (block-constant
                        metric, fused multiply-adds; within 1e-12 per stage
                        and 1e-10 over long runs)                             */
} fdf_opts;

typedef struct fdf_info {
  int nt, nstages, ndim, mode;
  int nx[3];
  int nfields;       /* evolved fields (RK updated)                             */
  int nfields_total; /* including plastic fields                                */
  int nifaces, sbporder, plastic, nblocks, step, precision;
  double dt, ttot, cfl, t;
  int tiled_blocks;  /* blocks updated by a rectilinear tiled kernel           */
  int row_blocks;    /* of which with row-shaped tiles (bulk_rows3d)          */
  int curv_blocks;   /* 2-D mode II blocks on the curvilinear tiled kernel      */
  int wide_blocks;   /* of the tiled and curvilinear blocks, those whose kernel
                        addresses each field through its own buffer
                        descriptor (nfields N elements above 4 GiB)           */
  int fused_faults;  /* interfaces whose friction solve runs inside the stage's
                        bulk launch (2-D single problems with one
                        slip-weakening interface along x between two tiled
                        blocks); the others have a fault launch of their own
                        before the bulk launch                               */
} fdf_info;

/* Replaces problem::problem(const char*) (src/problem.cpp:16-84): parses the
 * .in file, builds grid/metric/normals, chooses dt, uploads to the device,
 * performs the t=0 interface solve, initial outputs and front. */
int fdf_open(const char* in_path, const fdf_opts* opts, fdf_t** out);

/* Replaces nsteps iterations of the loop in problem::solve
 * (src/problem.cpp:151-177): all RK stages via do_rk_stage, plasticity,
 * outputs and front update. */
int fdf_step(fdf_t* h, int nsteps);

/* Replaces one call of domain::do_rk_stage(dt, stage, t, rk)
 * (src/domain.cpp:228-305) for time step `istep` (t = istep*dt); no
 * end-of-step work.  Used by kernel-level parity tests. */
int fdf_stage(fdf_t* h, int istep, int stage);

/* Whole-run equivalent of `fdfault <file.in>` (src/main.cpp:8-72). */
int fdf_run(const char* in_path, const fdf_opts* opts);

/* Read back a strided box of one field of f (the reference's outputunit
 * reads f through friend access, src/outputunit.cpp:791): indices lo..hi
 * inclusive with stride st, C order, as FP64. */
int fdf_read_field(fdf_t* h, int field, const int lo[3], const int hi[3], const int st[3], double* host);

/* Whole arrays, FP64 on the host: which = 0 f, 1 df, 2 x, 3 metric, 4 jac
 * (fields::f/df/x/metric/jac, src/fields.hpp:55-62).  n = element count. */
int fdf_read_array(fdf_t* h, int which, double* host, long n);
int fdf_write_array(fdf_t* h, int which, const double* host, long n);

/* Interface arrays (interface::u/v/s/sn/ux/vx/sx/state, src/interface.hpp:64-71,
 * plus friction::du/dux/dstate): which = 0 U,1 V,2 S,3 Sn,4 Ux,5 Vx,6 Sx,
 * 7 state,8 du,9 dux,10 dstate.  Returns the element count (host may be NULL
 * to query it) or a negative code. */
long fdf_read_iface(fdf_t* h, int iface, int which, double* host, long n);

/* Rupture-front times of a frictional interface (front::tvals,
 * src/front.cpp:103-125); returns the element count or a negative code. */
long fdf_read_front(fdf_t* h, int iface, double* host, long n);

/* State snapshot and resume: no counterpart in the reference, which has no
 * restart (its output files are truncated at open, src/outputunit.cpp:580).
 * fdf_save_state writes the evolving state after the last completed step --
 * f, df and the plastic fields in the bulk precision, every interface's
 * slip, slip rate, tractions, state, their RK increments and the rupture
 * front, and the step count -- to a binary file; fdf_load_state reads one
 * into a handle opened on the same problem with the same bulk precision, so
 * that stepping on gives bit-identical results to the run that wrote it.
 * Output units and the STZ proposal network are not part of the state. */
int fdf_save_state(fdf_t* h, const char* path);
int fdf_load_state(fdf_t* h, const char* path);

/* Flushes and closes output files and writes front files
 * (outputlist::close_list, frontlist::write_list; src/problem.cpp:182-187). */
int fdf_finish(fdf_t* h);

/* Problem summary (problem/domain members: nt, dt, ttot, cfl, sizes). */
int fdf_get_info(fdf_t* h, fdf_info* info);

/* Host-only set-up check: parses the .in file and builds grid, metric,
 * Jacobian and dt on the host without touching a GPU (the reference does this
 * in block::set_grid, src/block.cpp:769-1046, and problem::set_time_step,
 * src/problem.cpp:94-134).  Any output pointer may be NULL; sizes are
 * ndim*N, ndim*ndim*N and N with N = nx*ny*nz. */
int fdf_host_setup(const char* in_path, fdf_info* info, double* x, double* metric, double* jac);
const char* fdf_last_error(void);
void fdf_close(fdf_t* h);

/* Average device time (ms) of the bulk stencil kernel over the launches since
 * the last call, measured with HIP events on the stream it runs on; count of
 * launches in *nlaunch.  Measurement hook for the benchmark; no counterpart
 * in the reference. */
double fdf_bulk_kernel_ms(fdf_t* h, int* nlaunch);
/* The same for the fault (interface) launches of one stage, all interfaces
 * together; count of stages in *nstages. */
double fdf_fault_kernel_ms(fdf_t* h, int* nstages);

/* ---- Ensembles ------------------------------------------------------------
 * No counterpart in the reference, which runs problem::solve
 * (src/problem.cpp:136-192) once per .in file, one process each.  A batch
 * steps n independent 2-D mode II elastic problems together on one GPU: each
 * instance is set up as fdf_open would (same outputs and front files), then
 * every RK stage of every instance runs as one fault launch and one
 * persistent bulk launch.  Instances must share sbporder and rkorder; grid
 * sizes and time steps may differ.  The bulk kernel's tile width, x-chunk and
 * launch mode are chosen at open by timing candidates (cached in
 * $FDF_TUNE_CACHE or ~/.cache/fdfault_amd_tune.txt; FDF_BATCH_TW/FDF_BATCH_CX
 * fix the tile). */
typedef struct fdf_batch_s fdf_batch_t;

int fdf_open_batch(const char* const* in_paths, int n, const fdf_opts* opts, fdf_batch_t** out);
/* nsteps time steps of every instance (problem::solve's loop body, per
 * instance: stages, outputs and fronts). */
int fdf_batch_step(fdf_batch_t* h, int nsteps);
int fdf_batch_size(fdf_batch_t* h);
/* Per-instance readback, with fdf_get_info / fdf_read_iface /
 * fdf_read_front / fdf_read_array semantics for instance `inst`. */
int fdf_batch_get_info(fdf_batch_t* h, int inst, fdf_info* info);
long fdf_batch_read_iface(fdf_batch_t* h, int inst, int iface, int which, double* host, long n);
long fdf_batch_read_front(fdf_batch_t* h, int inst, int iface, double* host, long n);
int fdf_batch_read_array(fdf_batch_t* h, int inst, int which, double* host, long n);
/* The autotuner's choice: tile width (y points per workgroup), x planes per
 * chunk, the timed bulk-stage time (ms) and whether it came from the cache. */
int fdf_batch_tuning(fdf_batch_t* h, int* tile_width, int* chunk, double* stage_ms, int* from_cache);
/* The autotuner's launch mode of the bulk kernel: 1 a persistent grid of the
 * resident workgroups walking the work list, 0 one workgroup per work item
 * (FDF_BATCH_PERSIST fixes it). */
int fdf_batch_persistent(fdf_batch_t* h);
/* fdf_finish for every instance. */
int fdf_batch_finish(fdf_batch_t* h);
void fdf_batch_close(fdf_batch_t* h);

/* ---- Learned proposal for the STZ root ------------------------------------
 * No counterpart in the reference, whose STZ law solves for the friction
 * coefficient with a bracketed Newton iteration from the bracket midpoint
 * (stz::calc_mu, src/stz.cpp:263-356, and solve_newton,
 * src/utilities.cpp:89-180).  A 4 -> 64 -> 64 -> 1 tanh network maps
 * (phi, sn, chi, eta) -- locked shear traction, normal stress, effective
 * temperature, impedance -- to log(V); the friction coefficient implied by
 * that slip rate starts the same safeguarded Newton iteration, which must
 * reach |f| <= 1e-12 |phi|, else the plain bracketed solve runs.  The
 * parameter vector (fdf_mlp_param_count() floats) is
 * W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3 in_mean[4] in_std[4] out_mean
 * out_std. */
int fdf_mlp_param_count(void);
/* Train on n log rows (phi, sn, chi, eta, log V) with the GPU kernels
 * (standardised data, Adam, minibatches of `batch`, learning rate decaying to
 * lr/10); writes the parameters and each epoch's mean loss. */
int fdf_mlp_train(const double* rows, long n, int epochs, int batch, float lr, unsigned seed, float* params,
                  float* epoch_loss);
/* Gradient (4545 floats) and mean loss of one batch of n normalised samples,
 * X[n][4], T[n], computed by the GPU training kernels. */
int fdf_mlp_grad(const float* params, const float* X, const float* T, long n, float* grad, float* loss);
/* The gradient of fdf_mlp_grad without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight training or serving loops).  params, X[n][4], T[n],
 * grad[4545] and loss[1] (mean loss; may be NULL) are device pointers on
 * `device`; the kernels are queued on `stream` (a hipStream_t, NULL for the
 * null stream) and the call returns without waiting.  Working memory is a
 * fixed 4.36 MB array that each GPU holds until the process ends, so n is
 * limited to 1..4096 (larger batches return an error); calls that overlap on
 * one GPU share that array and race. */
int fdf_mlp_grad_noalloc(int device, const float* params, const float* X, const float* T, long n, float* grad,
                         float* loss, void* stream);
/* Normalised network outputs for n normalised inputs (GPU). */
int fdf_mlp_predict(const float* params, const float* X, long n, float* Y);
/* The same with the evaluation the fault kernel uses: the hidden 64 x 64
 * layer on the matrix cores with FP16 operands and FP32 accumulation. */
int fdf_mlp_predict_mfma(const float* params, const float* X, long n, float* Y);
/* Use (params != NULL, n = fdf_mlp_param_count()) or stop using the network
 * in this problem's STZ solves. */
int fdf_stz_set_model(fdf_t* h, const float* params, int n);
/* Log the next `capacity` STZ root solves as rows (phi, sn, chi, eta, log V);
 * 0 stops logging.  fdf_stz_log_read copies up to cap rows and returns the
 * number logged. */
int fdf_stz_log(fdf_t* h, long capacity);
long fdf_stz_log_read(fdf_t* h, double* out, long cap);
/* Counters since the last call: root solves, Newton iterations, solves
 * finished from the network's proposal, fallbacks to the plain solve. */
int fdf_stz_stats(fdf_t* h, unsigned long long* out4);

#ifdef __cplusplus
}
#endif
#endif
