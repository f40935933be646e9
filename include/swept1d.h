/*
 * swept1d.h — C ABI of libswept1d, the MI355X (gfx950) 1D swept-rule solver.
 *
 * The reference (Niemeyer-Research-Group/1DSweptCUDA) has no library API: its
 * boundary is the process (argv + .dat/.txt files) of the executables
 * bin/{Heat,KS,Euler}{Single,Double}Out (SweptSource/Makefile:17-35).  This
 * header is the C entry surface those executables are rebuilt on; each entry
 * point names the reference code it replaces.  Plain C types only: pointers,
 * sizes, integers, doubles.
 *
 * Conventions
 *   - Every function returns 0 (SW1D_OK) on success and a negative SW1D_E*
 *     code on failure; sw1d_last_error() gives the message (thread-local).
 *     The library never calls exit().
 *   - The caller owns every host buffer; a plan owns its device memory and
 *     one HIP stream.  One host thread may use a plan at a time.
 *   - State layout on the host: heat and KS take nx REALs; Euler takes nx
 *     cells of three REALs (rho, rho*u, E) interleaved, as the reference's
 *     REALthree arrays (Euler1D_SweptShared.cu:1175-1179).
 *   - REAL is double for SW1D_F64 plans and float for SW1D_F32 plans.
 */
#ifndef SWEPT1D_H
#define SWEPT1D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW1D_ABI_VERSION 1

enum {
    SW1D_OK = 0,
    SW1D_EINVAL = -1,     /* bad argument (validation messages of main(), Heat:700-740) */
    SW1D_EDEVICE = -2,    /* HIP runtime error (cudaGetLastError path, Heat:796-802) */
    SW1D_ENOMEM = -3,     /* device or pinned allocation failed */
    SW1D_EUNSTABLE = -4,  /* Euler dt/dx > 0.21 (Euler:1168-1172) */
    SW1D_ENOTSUP = -5     /* combination the build does not provide */
};

/* Equations: one reference program each. */
enum { SW1D_HEAT = 0, SW1D_KS = 1, SW1D_EULER = 2 };
/* Precision: the -DREAL=double build (Makefile:4) or the float default. */
enum { SW1D_F64 = 0, SW1D_F32 = 1 };
/* Algorithm = argv[6] of the reference: 0 classic, 1 shared swept,
 * 2 hybrid CPU+GPU swept (heat, Euler) / register swept (KS).
 * 3 (library only, no reference counterpart): overlapped trapezoid tiles —
 * swept up-triangles widened by a redundant halo, any nx, same bits. */
enum { SW1D_CLASSIC = 0, SW1D_SWEPT = 1, SW1D_ALT = 2, SW1D_TRAPEZOID = 3 };
/* Boundary conditions.  DEFAULT picks the reference's: heat insulated
 * (mirror, Heat:140-142), KS periodic (KS:199), Euler Dirichlet ghost cells
 * (Euler:405-416).  PERIODIC heat is the synthetic periodic-heat workload. */
enum { SW1D_BC_DEFAULT = 0, SW1D_BC_MIRROR = 1, SW1D_BC_PERIODIC = 2, SW1D_BC_DIRICHLET = 3 };

typedef struct sw1d_problem {
    int32_t equation;    /* SW1D_HEAT / SW1D_KS / SW1D_EULER */
    int32_t precision;   /* SW1D_F64 / SW1D_F32 */
    int32_t algorithm;   /* SW1D_CLASSIC / SW1D_SWEPT / SW1D_ALT */
    int32_t bc;          /* SW1D_BC_* */
    int64_t nx;          /* spatial points, argv[1] */
    int32_t tpb;         /* threads per block, argv[2]: the swept node width */
    int32_t node_width;  /* 0: derive from tpb; else points per swept node (64*2^k) */
    double dt;           /* time step, argv[3] */
    double dx;           /* 0: reference value (heat ds 0.001, KS 0.5, Euler 1/(nx-2)) */
    double alpha;        /* heat diffusivity; 0: reference 8.418e-5 (Heat:56) */
    int32_t tile_depth;  /* SW1D_TRAPEZOID: stages per launch in units of 32/R (0: 1) */
    int32_t reserved;    /* must be 0 */
} sw1d_problem;

typedef struct sw1d_plan sw1d_plan;

/* Library / device ------------------------------------------------------- */
int sw1d_abi_version(void);
const char* sw1d_last_error(void);
/* Runs the lane-shift self check the swept kernels depend on.  1 = pass. */
int sw1d_device_check(void);

/* Plan lifecycle ---------------------------------------------------------- */
/* Replaces the setup of main() + the cudaMalloc block of classicWrapper /
 * sweptWrapper (Heat:471-481, 523-538; KS:596-605, 660-672; Euler:839-849,
 * 901-921): validates the problem, computes the constants exactly as the
 * reference main() does (Heat:719-728, KS:917-926, Euler:1127-1154) and
 * allocates the device state. */
int sw1d_plan_create(sw1d_plan** plan, const sw1d_problem* problem);
int sw1d_plan_destroy(sw1d_plan* plan);

/* H2D of the initial condition (Heat:481/538, KS:605/672, Euler:849/921). */
int sw1d_set_state(sw1d_plan* plan, const void* host);
/* Synchronous D2H of the current state (Heat:499/514/623/686 …). */
int sw1d_get_state(sw1d_plan* plan, void* host);
/* Device-to-device forms for callers that keep the state in HBM
 * (same layout as the host forms). */
int sw1d_set_state_device(sw1d_plan* plan, const void* device_ptr);
int sw1d_get_state_device(sw1d_plan* plan, void* device_ptr);

/* Advance the state by nsteps time steps with the plan's algorithm.  Replaces
 * the kernel loops of classicWrapper (Heat:491-510, KS:612-631,
 * Euler:859-888) and sweptWrapper (Heat:543-682, KS:753-851,
 * Euler:927-1100) between two outputs.  Asynchronous on the plan's stream
 * for the GPU-only algorithms; the hybrid algorithm synchronises once per
 * CPU node.  Results are bitwise the same for every algorithm. */
int sw1d_advance(sw1d_plan* plan, int64_t nsteps);
int sw1d_synchronize(sw1d_plan* plan);
/* Advance and return the elapsed device time in milliseconds (hipEvents on
 * the plan's stream; replaces the cudaEvent pair of Heat:774-794). */
int sw1d_advance_timed(sw1d_plan* plan, int64_t nsteps, double* elapsed_ms);

/* Allocation-free advance -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Advances a caller-owned device state in place by nsteps
 * steps and gives the bits of sw1d_plan_create + sw1d_set_state_device +
 * sw1d_advance + sw1d_get_state_device, without a plan: the two state
 * buffers live in a library-owned static device workspace of
 * SW1D_NOALLOC_MAX_BYTES on each GPU (held until the process ends), and the
 * call makes no hipMalloc/hipFree, creates no stream, event or graph.
 *   d_state  device pointer, layout of sw1d_set_state (read, then overwritten)
 *   stream   hipStream_t as void* (NULL: the null stream).  The call only
 *            enqueues work on it; the caller synchronises.  The first call on
 *            a device resolves the workspace and runs the lane-shift check
 *            synchronously, so it must not be under stream capture.
 * Limit: 2 * values_per_point * nx * sizeof(REAL) <= SW1D_NOALLOC_MAX_BYTES
 * (64 MiB: heat and KS FP64 up to 2^21 points, Euler FP64 up to 699050
 * cells); larger problems return SW1D_EINVAL.  The hybrid CPU+GPU algorithm
 * (heat/Euler SW1D_ALT) returns SW1D_ENOTSUP.  All calls on one GPU share its
 * workspace, which keeps the last call's contents: calls that overlap on one
 * GPU (other streams or host threads) race. */
#define SW1D_NOALLOC_MAX_BYTES (64ull << 20)
int sw1d_advance_noalloc(const sw1d_problem* problem, void* d_state, int64_t nsteps, void* stream);

/* Plan facts ----------------------------------------------------------------
 * info[0] node width (points per swept node), info[1] steps advanced by one
 * swept tile launch, info[2] values stored per point (2 heat and KS, 6 Euler),
 * info[3] device bytes held by the plan, info[4] steps per reference cycle
 * (the t_fullstep of sweptWrapper / classicWrapper divided by dt). */
int sw1d_plan_info(const sw1d_plan* plan, int64_t info[5]);
/* The REAL constants the kernels use, widened to double:
 * heat {Fo, 1-2Fo}; KS {1/(4dx), 1/dx^2, 1/dx^4, dt, dt/2};
 * Euler {gam, mgam, dt/dx, rhoL, mL, EL, rhoR, mR, ER}. */
int sw1d_plan_constants(const sw1d_plan* plan, double* out, int32_t capacity);
/* hipStream_t of the plan, as void*. */
int sw1d_plan_stream(const sw1d_plan* plan, void** stream);

/* Per-kernel timing: when enabled, every launch of the plan is bracketed by
 * hipEvents on the plan's stream.  kind: 0 classic stage, 1 up-triangle,
 * 2 down-triangle, 3 diamond, 4 split (boundary-straddling) diamond. */
int sw1d_set_profiling(sw1d_plan* plan, int32_t enabled);
int sw1d_kernel_stats(sw1d_plan* plan, int32_t kind, int64_t* launches, double* total_ms);

/* Schedule ------------------------------------------------------------------
 * The reference's time bookkeeping (Heat:483-510, 541-682; KS:609-631,
 * 676-851; Euler:855-888, 924-1098): given the argv values it yields the
 * step counts at which the program writes a .dat row and the time it prints
 * there.  out_steps/out_times receive up to capacity rows (the last row is the
 * final state, row count returned in *count).  Returns SW1D_EINVAL if
 * capacity is too small (count then holds the needed size).
 * kind: 0 heat classic, 1 heat swept/hybrid, 2 KS classic, 3 KS swept,
 * 4 KS register, 5 Euler classic, 6 Euler swept/hybrid. */
int sw1d_schedule(int32_t kind, double dt, double tf_arg, double freq, int32_t tpb,
                  double dt_sched, int64_t* out_steps, double* out_times,
                  int64_t capacity, int64_t* count);

/* Seeded initial conditions -------------------------------------------------
 * The reference's initFun (KS1D_SweptShared.cu:68-72, 945; the heat IC of
 * Heat1D_SweptShared.cu:60-66 is replaced by a periodic one) extended with the
 * synthetic workloads' seeded terms, drawn on the host from splitmix64 so that
 * every host gives the same values.  Computed in double, then rounded to REAL.
 *   SW1D_IC_KS_NOISY      u_k = 2cos(19 pi x/128) + param * N(0,1), x = k * length
 *                         (length = dx; Box-Muller, cosine branch)
 *   SW1D_IC_HEAT_FOURIER  T_j = 1 + sum_{k=1..4} a_k sin(2 pi k x/L), a_k ~ U[0,1),
 *                         x = j L / n (length = L; param unused)
 * Host only: needs no GPU. */
enum { SW1D_IC_KS_NOISY = 0, SW1D_IC_HEAT_FOURIER = 1 };
int sw1d_initial_condition(int32_t kind, int32_t precision, int64_t n, double length, double param,
                           uint64_t seed, void* host);

/* Batched solves -------------------------------------------------------------
 * No reference counterpart (the reference solves one problem per process):
 * many independent periodic heat / KS problems of any sizes, advanced together
 * by one persistent kernel launch, replayed from a captured hipGraph.
 * Problem i: equation[i] (SW1D_HEAT periodic or SW1D_KS), nx[i] >= 128 points,
 * time step dt[i], grid spacing dx[i] (heat diffusivity 8.418e-5), tile width
 * tile_points[i] output points per trapezoid tile (128, 256, 512 or 1024;
 * 0 picks 512).  The state is the concatenation of the problems' fields in index
 * order, REAL per the precision.  Every problem's result is bitwise the
 * single-problem result. */
typedef struct sw1d_batch sw1d_batch;
int sw1d_batch_create(sw1d_batch** batch, int32_t precision, int64_t count,
                      const int32_t* equation, const int64_t* nx, const double* dt,
                      const double* dx, const int32_t* tile_points);
int sw1d_batch_destroy(sw1d_batch* batch);
int sw1d_batch_set_state(sw1d_batch* batch, const void* host);
int sw1d_batch_get_state(sw1d_batch* batch, void* host);
int sw1d_batch_advance(sw1d_batch* batch, int64_t nsteps);
int sw1d_batch_advance_timed(sw1d_batch* batch, int64_t nsteps, double* elapsed_ms);
/* Waves per workgroup of the persistent launch, the tiles a workgroup works on
 * at once (the diamonds per workgroup of the reference's tuning): 2, 4
 * (default) or 8.  Results do not depend on it. */
int sw1d_batch_set_workgroup(sw1d_batch* batch, int32_t waves);
/* info[0] total points, info[1] persistent workgroups, info[2] problems,
 * info[3] device bytes */
int sw1d_batch_info(const sw1d_batch* batch, int64_t info[4]);

/* Learned KS surrogate -------------------------------------------------------
 * No reference counterpart.  A periodic 1D CNN (6 layers, 64 channels, kernel
 * 5, GELU) predicting the KS change over K solver steps; layers 2-5 run on
 * BF16 MFMA with FP32 accumulation, forward and backward; trained on the GPU
 * with Adam.  Inputs and outputs are in normalised units: x = u / in_scale,
 * y = du / out_scale.  Parameters are one FP32 vector of
 * sw1d_nn_param_count() values. */
typedef struct sw1d_nn sw1d_nn;
const char* sw1d_nn_last_error(void);
int sw1d_nn_param_count(void);
/* Checks the MFMA operand layout the kernels assume.  1 = pass. */
int sw1d_nn_device_check(void);
/* nx a multiple of 64 in [64, 2^24] (multiples of 256 use the fused inference
 * kernels), 1 <= max_batch <= 65535 samples per forward / training step. */
int sw1d_nn_create(sw1d_nn** nn, int32_t nx, int32_t max_batch, uint64_t seed);
int sw1d_nn_destroy(sw1d_nn* nn);
int sw1d_nn_set_scales(sw1d_nn* nn, double in_scale, double out_scale);
/* Inference kernels (forward and surrogate steps): 0 = automatic (one
 * whole-network launch at nx = 1024, per-layer launches for other multiples
 * of 256, the training kernels otherwise), 1 = per-layer launches (nx a
 * multiple of 256), 2 = the whole-network kernel (nx = 1024).  Every choice
 * gives the same bits. */
int sw1d_nn_set_inference(sw1d_nn* nn, int32_t mode);
int sw1d_nn_get_params(sw1d_nn* nn, float* host);
int sw1d_nn_set_params(sw1d_nn* nn, const float* host);
int sw1d_nn_forward(sw1d_nn* nn, const float* x_host, int32_t batch, float* y_host);
/* mean-square loss against t_host and its gradient (grad_host may be NULL) */
int sw1d_nn_gradient(sw1d_nn* nn, const float* x_host, const float* t_host, int32_t batch,
                     double* loss, float* grad_host);
/* Adam over device-resident [nsamples][nx] FP32 inputs / targets */
int sw1d_nn_train(sw1d_nn* nn, const float* d_inputs, const float* d_targets, int64_t nsamples,
                  int32_t batch, int64_t iters, float lr, uint64_t seed, double* last_loss);
/* Training set from an FP64 KS batch of the network's nx: the batch is
 * advanced nsnap times by k_steps; pairs (u_s / in_scale, (u_{s+1} - u_s) /
 * out_scale) are kept on the device, in_scale / out_scale = RMS of u and of
 * the K-step change (written to scales[0..1] and set on the network). */
int sw1d_nn_dataset(sw1d_nn* nn, sw1d_batch* batch, int32_t nsnap, int32_t k_steps,
                    int64_t* nsamples, double* scales);
/* Copies stored pairs [first, first+count) to the host: sample s*B + b is
 * snapshot s of problem b of the batch, nx floats each (x = u_s / in_scale,
 * t = (u_{s+1} - u_s) / out_scale, each computed in double, rounded once).
 * x_host or t_host may be NULL. */
int sw1d_nn_dataset_samples(sw1d_nn* nn, int64_t first, int64_t count, float* x_host, float* t_host);
/* Adam on the first nfit samples of the stored training set */
int sw1d_nn_fit(sw1d_nn* nn, int64_t nfit, int32_t batch, int64_t iters, float lr, uint64_t seed,
                double* last_loss);
/* relative L2 error of the predicted change over stored samples [first, first+count) */
int sw1d_nn_eval(sw1d_nn* nn, int64_t first, int64_t count, double* rel_err);
/* Advance every problem of an FP64 KS batch (all of the network's nx) by
 * surrogate_steps network steps of k_steps solver steps each.  Every
 * judge_every-th network step the plain solver advances the same states by
 * k_steps steps; a problem whose relative L2 difference ||u_net - u_plain|| /
 * ||u_plain|| exceeds tol keeps the plain solver's state (fallback).
 * stats: {network steps, judged steps, problem fallbacks, max judged error}. */
int sw1d_surrogate_advance(sw1d_nn* nn, sw1d_batch* batch, int64_t surrogate_steps,
                           int32_t k_steps, int32_t judge_every, double tol, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* SWEPT1D_H */
