/* C ABI of libinductioneq_hip.so -- the MI355X (gfx950) implementation of the
 * InductionEq right-hand side, Runge-Kutta stages and CG divergence cleaning.
 *
 * The reference drives one OpenCL program per run through MatCL (an external MATLAB
 * mex bridge): a build call with a source-module list and a -D settings string, then
 * run calls that execute a list of kernel names, in order, on one argument list.
 * This header keeps that contract: ieq_program_build takes the same settings string
 * and module list, ieq_run takes the same kernel names with the same argument order.
 * Arguments are device pointers; every entry point returns IEQ_OK (0) or a negative
 * error code, with the message available from ieq_last_error.
 *
 * Device field layout: a vector field (the reference's REAL4 array) is stored as
 * three planes of `stride` REALs (x, y, z components; the reference's unused fourth
 * lane is not stored); a scalar field (REAL array) is one plane; the node index is
 * ix + NX*(iy + NY*iz) as in the reference (include/utils.h:66-68).  `stride` is the
 * node count rounded up to 64 and is reported by ieq_program_get_info.  `time` is a
 * device REAL[2] with time[0] the time at the start of a step and time[1] the time
 * of the current stage (kernel_time_integrator.cl:51-55).
 */
#ifndef INDUCTIONEQ_H
#define INDUCTIONEQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ieq_program ieq_program;

enum {
  IEQ_OK = 0,
  IEQ_ERR_SETTINGS = -1,    /* invalid settings or module list (the reference's #error / build failure) */
  IEQ_ERR_KERNEL = -2,      /* kernel name not in the program (CL_INVALID_KERNEL_NAME) */
  IEQ_ERR_ARGS = -3,        /* wrong argument count, null pointer or NDRange outside the grid */
  IEQ_ERR_HIP = -4,         /* HIP runtime error */
  IEQ_ERR_UNSUPPORTED = -5  /* option combination this library does not provide */
};

/* Parsed program settings (read-only view). */
typedef struct {
  int real_bytes;     /* 8 = double, 4 = float (REAL) */
  int opset;          /* 0 2ndOrder, 1 4thOrder, 2 6thOrder, 3 2ndOrderExtended, 4 4thOrderExtended, 5 6thOrderExtended */
  int order;          /* ORDER */
  int num_bounds;     /* NUM_BOUNDS */
  int stencil_width;  /* STENCIL_WIDTH */
  int nx, ny, nz;     /* NODES_X/Y/Z */
  double dx, dy, dz, xmin, ymin, zmin, dt;
  int periodic;       /* USE_PERIODIC */
  int hall;           /* USE_HALL */
  int dissipation;    /* 0 none, 1 high order, 2 first order, 3 adaptive */
  int form_uibj, form_source, form_ujbi;
  int testcase;       /* 0 rotation_2D, 1 rotation_3D, 2 hall_periodic, 3 hall_travelling_wave, 4 confined_domain,
                         5 alfven_periodic_2D, 6 gaussian_rotation_3D, 7 uniform_flow */
  int laplace;        /* 0 WS_D0, 1 NS_D0, 2 WS_LN, -1 no divergence-cleaning module */
  int bnodes;         /* BNODES */
  int w_size;         /* W_SIZE */
  double ho_factor, mp2min, mp2max, cmin, cmax, ux, uy, uz;
  long long num_nodes, num_nodes_pad, stride, num_groups;
} ieq_program_info;

/* Build a program.  Replaces MatCL's build call cl_run_kernel(device, kernel_path_list,
 * settings) (matlab/+induction_eq/compute_numerical_solution.m:23; module list and
 * settings assembled at compute_numerical_solution_setup.m:57-137, initialize.m:63-77,
 * divergence_cleaning/initialize.m:105-154).  `settings` is the generate_settings
 * string (matlab/generate_settings.m:3-30); `modules` are the source names of
 * kernel_path_list with or without directory and extension ("4thOrder",
 * "../include/rotation_2D.h", "kernel_time_integrator", ...).  Does not touch the GPU. */
int ieq_program_build(ieq_program **prog, const char *settings, const char *const *modules, int nmod);

/* Release a program and the HIP objects it created. */
void ieq_program_release(ieq_program *prog);

/* Message of the last error on `prog` (or of the last failed build when prog is NULL). */
const char *ieq_last_error(const ieq_program *prog);

/* Parsed settings of a program. */
int ieq_program_get_info(const ieq_program *prog, ieq_program_info *info);

/* Stream on which all work of `prog` is enqueued (a hipStream_t; NULL = default stream). */
int ieq_set_stream(ieq_program *prog, void *hip_stream);

/* Wait for all work enqueued on the program's stream. */
int ieq_synchronize(ieq_program *prog);

/* Device allocation helpers (zero-filled).  Sizes: ieq_vector_field_bytes /
 * ieq_scalar_field_bytes for one field of this program. */
int ieq_alloc(ieq_program *prog, size_t bytes, void **dptr);
int ieq_free(ieq_program *prog, void *dptr);
size_t ieq_vector_field_bytes(const ieq_program *prog);
size_t ieq_scalar_field_bytes(const ieq_program *prog);

/* Host <-> device copies in the reference host layout: `n` nodes of AoS REAL4
 * (x, y, z, w) <-> a device vector field; `n` REALs <-> a device scalar field.
 * Replace MatCL's implicit copy-in / write-back of MATLAB arrays
 * (matlab/+induction_eq/initialize.m:88-93). Downloads write w = 0. */
int ieq_upload_real4(ieq_program *prog, void *dst, const void *host_aos, size_t n);
int ieq_download_real4(ieq_program *prog, void *host_aos, const void *src, size_t n);
int ieq_upload_real(ieq_program *prog, void *dst, const void *host, size_t n);
int ieq_download_real(ieq_program *prog, void *host, const void *src, size_t n);

/* Run a list of reference kernels, in order, on one argument list.  Replaces MatCL's
 * run call t = cl_run_kernel(device, names, g_range, l_range, args..., 0)
 * (compute_numerical_solution.m:39-41,95-97; clean_div_CG.m:31-79).  Kernel names and
 * argument orders are those of kernel/kernel_time_integrator.cl (95 stage, time,
 * velocity and curl kernels), kernel_init.cl, kernel_operator.cl,
 * kernel_div_cleaning.cl, kernel_CG.cl and kernel_dot_product.cl.  `g` is the NDRange,
 * `l` the local size (only the reductions use it: one output per l[0] nodes).
 * Scalars such as the CG coefficient are 1-element device buffers (kernel_CG.cl:6-7).
 * When kernel_seconds is non-NULL the call waits for the kernels and stores their time.
 * A list that is one integrator's RK_Step repeated k times over the whole grid (the
 * reference time loop's repmat(RK_Step, 1, k), compute_numerical_solution.m:91-108)
 * runs as ieq_step(integrator, k) would: on a periodic grid with uniform u and rho the
 * fused stage kernels advance b1, the other field registers, curlB_rho and time to the
 * same state as the listed kernels.  ieq_last_step_path reports which path ran. */
int ieq_run(ieq_program *prog, const char *const *names, int n, const uint32_t g[3], const uint32_t l[3], void *const *args,
            int nargs, double *kernel_seconds);

/* A host array argument of ieq_run_host: `n` nodes of AoS REAL4 (vector != 0; 4*n REALs in the
 * reference's host layout) or `n` REALs (vector == 0). */
typedef struct {
  void *host;
  size_t n;
  int vector;
} ieq_host_arg;

/* MatCL's run call on host arrays, with its copy semantics: every argument is copied to a
 * device field (AoS REAL4 -> SoA), the kernel list runs as in ieq_run, and every argument is
 * written back in place, as MatCL writes kernel arguments back into the MATLAB arrays
 * (compute_numerical_solution.m:39-41,95-97; initialize.m:88-93).  kernel_seconds receives
 * the kernels' time only (MatCL's return value), not the copies.  Device fields are
 * allocated per call; for long runs keep fields on the device and use ieq_run. */
int ieq_run_host(ieq_program *prog, const char *const *names, int n, const uint32_t g[3], const uint32_t l[3],
                 const ieq_host_arg *args, int nargs, double *kernel_seconds);

/* Advance `nsteps` full time steps of `integrator` (SSPRK33, SSPRK104,
 * KennedyCarpenterLewis2R54C, CalvoFrancoRandez2R64, CarpenterKennedy2N54,
 * ToulorgeDesmet2N84F) with the integrator's reference argument list
 * (b1, b2[, b3], curlB_rho, u, rho, time).  Equivalent to ieq_run on the repeated
 * RK_Step list of compute_numerical_solution_setup.m:140-371 (block loop of
 * compute_numerical_solution.m:91-108); when the grid is periodic and u and rho are
 * uniform it runs the fused stage kernels instead of the reference kernel sequence.
 * Uniformity is checked on the device, which costs one host synchronisation per call.
 * On return b1 and time hold the same state as after the reference schedule; b2/b3
 * and curlB_rho are work registers whose contents are unspecified.
 * `variable_u` requests calc_u before each stage (I_RunOps('variable_u')). */
int ieq_step(ieq_program *prog, const char *integrator, long long nsteps, int variable_u, void *const *args, int nargs,
             double *kernel_seconds);

/* ieq_step for a caller that knows u and rho are uniform: the fields hold u[0..2] and rho at
 * every node, and these values are used without reading the fields, so the call enqueues
 * its kernels without any host synchronisation.  Same results as ieq_step on such fields.
 * No counterpart in the reference (its coefficients are always fields). */
int ieq_step_uniform(ieq_program *prog, const char *integrator, long long nsteps, const double u[3], double rho,
                     void *const *args, int nargs, double *kernel_seconds);

/* Path taken by the last successful ieq_run / ieq_step / ieq_step_uniform on prog: 1 the fused
 * stage kernels, 2 the reference kernel sequence with its stage kernels run as z-marching tiles
 * (whole-grid NDRange, no Hall term, no adaptive dissipation), 0 the reference kernel sequence
 * with one work-item per node, -1 none (or the call failed).  Setting IEQ_NO_FUSED=1 in the
 * environment keeps every call on the reference kernel sequence, IEQ_NO_TILED=1 keeps its stage
 * kernels at one work-item per node. */
int ieq_last_step_path(const ieq_program *prog);

/* FP64 diagnostics of the vector field b at time `time`, whatever REAL is, in one pass with every
 * product and sum in double: out[0] = the SBP-norm energy sum fac |B|^2 (norm2,
 * kernel_operator.cl:11-44), out[1] = sqrt(sum fac (div_D B)^2) (calc_div + norm2_S,
 * kernel_operator.cl:98-155), out[2] = sqrt(sum fac |B - b_analytical(time)|^2) (analytical_b +
 * norm2_diff, kernel_init.cl:36-48; NaN for a testcase without an analytic field), fac =
 * prod_d dx_d / M_INV[row_d].  Replaces the host loop of reductions of
 * compute_numerical_solution.m:69-89,218-250, which forms these sums in REAL.  The partial sums
 * are added in a fixed order (deterministic).  Waits for the result. */
int ieq_diagnostics(ieq_program *prog, const void *b, double time, double out[3]);

/* ieq_diagnostics for callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No counterpart in the reference.  Same
 * kernels and results as ieq_diagnostics, but nothing is allocated and nothing waits: the three
 * results go to out_dev (3 doubles of device memory), enqueued on the program's stream.  The
 * partial sums live in a static device array of 1024 x 3 doubles (24 KiB per GPU, held until the
 * process ends); the partial pass uses at most 1024 blocks whatever the grid size, so every grid
 * fits, and every call overwrites the entries it reads.  Calls that overlap on one GPU (several
 * streams or host threads at once) race on that array. */
int ieq_diagnostics_noalloc(ieq_program *prog, const void *b, double time, double *out_dev);

/* Result of ieq_clean_div_cg. */
typedef struct {
  double divergence_norm_initial; /* sqrt(r.r), r = div B on the BNODES interior (clean_div_CG.m:36-44) */
  double divergence_norm_final;   /* SBP-norm of div B after the projection (clean_div_CG.m:164-167) */
  double residual_norm;           /* sqrt(r.r) when the loop stopped */
  int iterations;
  double seconds;                 /* device time of the CG loop */
} ieq_cg_result;

/* Divergence cleaning of the vector field b (in place) with the program's -Laplacian
 * (USE_LAPLACE_*, BNODES).  Replaces the host-driven loop of
 * matlab/+divergence_cleaning/clean_div_CG.m:31-86 (and +induction_eq/clean_div.m):
 * conjugate gradients on -Lap phi = div B with the same update sequence (u = beta*u + r,
 * alpha = r.r / u.q, beta = r.r / old, old = 1 initially), stopping when
 * r.r <= abs_threshold^2 (the reference's test) or, when rel_tolerance > 0, when
 * |r| <= rel_tolerance * |r0|, or after max_iterations; then B += grad phi
 * (B -= grad_adj phi for WS_LN) over the whole grid.  alpha and beta stay on the
 * device and the dot products accumulate in FP64 for both REAL types.  phi receives
 * the potential (scalar field). */
int ieq_clean_div_cg(ieq_program *prog, void *b, void *phi, double abs_threshold, double rel_tolerance, int max_iterations,
                     ieq_cg_result *result);

/* One member of an ensemble: an independent periodic solve with its own grid, spacing,
 * time step, uniform density (Hall coefficient 1/rho) and uniform velocity.  Fields are
 * device vector fields in this library's layout with component stride
 * ieq_member_stride(nx, ny, nz).  b3 is the third register of three-register schemes
 * and the ping-pong register of the 2N schemes; emf is a work register. */
typedef struct {
  int nx, ny, nz;
  double dx, dy, dz, dt;
  double rho;
  double ux, uy, uz;
  void *b1, *b2, *b3, *emf;
} ieq_member;

typedef struct ieq_ensemble ieq_ensemble;

/* Component stride (elements) of a member field: nx*ny*nz rounded up to 64. */
long long ieq_member_stride(int nx, int ny, int nz);

/* Create an ensemble of independent periodic solves sharing the program's settings
 * (REAL, operator order, USE_HALL, dissipation, flux forms; the program must be
 * periodic) and one integrator (SSPRK33, CarpenterKennedy2N54 or ToulorgeDesmet2N84F).
 * No counterpart in the reference (one program per solve there).  Each RK stage of the
 * whole ensemble is one persistent launch per fused kernel over a tile list covering
 * every member; on the first ieq_ensemble_step the tile shape, z-chunk and tile order are
 * chosen by timing the candidates (member fields are saved and restored around it). */
int ieq_ensemble_create(ieq_ensemble **ens, ieq_program *prog, const ieq_member *members, int nmembers, const char *integrator);

/* Advance every member by nsteps steps of its own dt, in periods of one step (two for the
 * 2N schemes, which must get an even nsteps).  b1 of every member holds the state on return. */
int ieq_ensemble_step(ieq_ensemble *ens, long long nsteps, double *kernel_seconds);

/* Chosen launch configuration: tile shape code (4: 64 x 4 cells, 8: 64 x 8, 32: 32 x 8 with the EMF
 * in its own launch; 1: the one-kernel Hall stage, 32 x 8 in FP64 and 64 x 8 in FP32; 2: the
 * one-kernel Hall stage with 16 x 16 tiles in FP64 and 32 x 8 in FP32),
 * z-chunk, order (0 member order, 1 largest first; +2 when the step is replayed from a
 * hipGraph rather than launched directly, whichever timed faster). */
int ieq_ensemble_config(const ieq_ensemble *ens, int *tile_shape, int *z_chunk, int *order);

void ieq_ensemble_release(ieq_ensemble *ens);

/* Discrete adjoint of a periodic SSPRK33 solve of the program (fused stage arithmetic,
 * uniform velocity u and Hall coefficient c_hall = 1/rho given at run time, HO or no
 * dissipation).  No counterpart in the reference; it differentiates the time loop of
 * compute_numerical_solution.m:91-108 run with the SSPRK33 kernels
 * (kernel_time_integrator.cl:224-303).  `snapshots` vector fields are kept for
 * optimal binomial checkpointing of the nsteps steps (stage inputs are recomputed from the
 * step start); snapshots >= nsteps keeps every step's three stage inputs (3 nsteps + 1
 * vector fields, when they fit in half the free device memory) and recomputes nothing. */
typedef struct ieq_adjoint ieq_adjoint;
int ieq_adjoint_create(ieq_adjoint **adj, ieq_program *prog, long long nsteps, int snapshots);

/* bT = B(T) after nsteps steps from b0 (vector fields of the program); stores the snapshots. */
int ieq_adjoint_forward(ieq_adjoint *adj, const void *b0, void *bT, double c_hall, const double u[3], double *kernel_seconds);

/* Given lambda_T = dL/dB(T): grad_b0 = dL/dB0, grads = {dL/dc_hall, dL/du_x, dL/du_y, dL/du_z}
 * (FP64, accumulated in a fixed order).  Uses the state of the last ieq_adjoint_forward. */
int ieq_adjoint_backward(ieq_adjoint *adj, const void *lambda_T, void *grad_b0, double grads[4], double *kernel_seconds);

/* Forward steps taken (forward sweep plus recomputation) and adjoint steps of the last calls. */
int ieq_adjoint_counts(const ieq_adjoint *adj, long long *forward_steps, long long *adjoint_steps);

void ieq_adjoint_release(ieq_adjoint *adj);

#ifdef __cplusplus
}
#endif

#endif /* INDUCTIONEQ_H */
