/*
 * neurite_hip.h -- C ABI of the MI355X-native Neurite cable-equation solvers.
 *
 * Plain C types only (pointers, sizes, status codes); no HIP or torch types cross
 * this boundary. Every entry point returns an int status (NRT_OK == 0) and, on
 * failure, leaves a message retrievable with nrt_last_error(handle) (or
 * nrt_last_error(NULL) for errors raised before a handle exists).
 *
 * Reference interfaces replaced (Neurite, ModelDBRepository/168861):
 *   inc/solver.h:12-24          Solver::update / calculate / snapshot / ~Solver
 *   inc/solverEx-gpu.h:79-85    SolverExGPU constructor + methods
 *   inc/solverIm-gpu.h:111-118  SolverImGPU constructor + methods
 *   inc/cudaException.h:331-357 cudaException (status + message here)
 *   src/solverIm-gpu.cpp:440-473,766-822  cuBLAS/cuSPARSE calls (csrmv, dot, axpy)
 *   src/init.cpp:138-298, inc/discretization.h:50-238  setup chain (strain, HH init)
 *
 * Units are the reference's SI units: volts, seconds, metres, ohms, farads,
 * siemens, amperes.
 */
#ifndef NEURITE_HIP_H
#define NEURITE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define NRT_OK                 0
#define NRT_ERR_INVALID        1  /* bad argument / inconsistent arrays          */
#define NRT_ERR_HIP            2  /* HIP runtime error (cudaException analogue)  */
#define NRT_ERR_NOMEM          3  /* device allocation failed                    */
#define NRT_ERR_TOPOLOGY       4  /* topology the solver cannot represent        */
#define NRT_ERR_NOT_CONVERGED  5  /* Krylov hit max_iter (reference: exit(0))     */
#define NRT_ERR_NUMERIC        6  /* NaN / Inf in a Krylov scalar or residual    */
#define NRT_ERR_UNSUPPORTED    7  /* option combination not built                */
#define NRT_ERR_ADMISSION      8  /* not enough free HBM to admit a request      */
#define NRT_ERR_NO_DEVICE      9  /* no usable GPU                               */

/* ------------------------------------------------------------- options */
#define NRT_EXPLICIT 0   /* FTCS + forward-Euler HH gates (solverEx-*.cpp) */
#define NRT_IMPLICIT 1   /* backward Euler + Krylov (solverIm-*.cpp)       */

#define NRT_FP64          0  /* plain form: every step in FP64                      */
#define NRT_FP32_REFINE   1  /* FP32 solves inside FP64 iterative refinement: the
                                tree solve (NRT_DIRECT), GMRES's Arnoldi work
                                (SpMV, basis, compensated dots) or BiCGSTAB's
                                preconditioner (NRT_PRECOND_TREE) in FP32, and
                                the HH rate functions in FP32; the residual
                                b - A x, the update of V and the gate state stay
                                FP64                                              */

#define NRT_BICGSTAB 0   /* solverIm-gpu.cpp:725-843 */
#define NRT_GMRES    1   /* solverIm-cpu.cpp:192-272 (restarted GMRES)       */
#define NRT_DIRECT   2   /* exact tree partition solve (no Krylov)           */

#define NRT_PRECOND_NONE 0  /* reference: unpreconditioned                    */
#define NRT_PRECOND_TREE 1  /* BiCGSTAB only: the tree solve of the step's
                               matrix as a right preconditioner (FP64: exact,
                               one iteration; FP32_REFINE: the FP32 tree solve,
                               a few); batches and the server accept it       */

#define NRT_TOL_ABSOLUTE 0  /* ||b - A x||_2 <= tol      (reference: 1e-14)  */
#define NRT_TOL_RELATIVE 1  /* ||b - A x||_2 <= tol*||b||_2                  */

#define NRT_X0_PREVIOUS 0   /* Krylov starts from V of the previous step      */
#define NRT_X0_RHS      1   /* Krylov starts from b (solverIm-gpu.cpp:169,769) */
#define NRT_X0_ZERO     2   /* Krylov starts from 0 (solverIm-cpu.cpp:205-215) */

/* Per-element material description (discretization.cpp:19-97 defaults). */
typedef struct nrt_material {
  int32_t kind;          /* 0 = passive cable (Cable), 1 = Hodgkin-Huxley      */
  int32_t layers_my;     /* myelin layers (Cable, 45 by default)                */
  double rho_a;          /* axial resistivity  [ohm m]      (1.87)             */
  double rho_m;          /* membrane resistivity [ohm m]   (2.5e9)             */
  double per_m;          /* membrane permittivity [F/m]    (4e-11)             */
  double thickness;      /* membrane thickness [m]         (4e-9)              */
  double rest_pot;       /* resting potential [V]          (-0.0655)           */
  double thickness_my;   /* myelin layer thickness [m]     (18e-9)             */
  double per_my;         /* myelin permittivity [F/m]      (1.08e-10)          */
  double ro_my;          /* myelin resistivity [ohm m]     (4.44e6)            */
  double sigma_Na;       /* HH sodium conductivity         (4.8e-6)            */
  double sigma_K;        /* HH potassium conductivity      (1.44e-6)           */
  double sigma_L;        /* HH leak conductivity           (1.2e-8)            */
  double dE_Na;          /* E_Na0 = rest_pot + dE_Na       (0.115)             */
  double dE_K;           /* E_K0  = rest_pot + dE_K        (-0.012)            */
  double n_damage;       /* damage exponent                (2)                 */
} nrt_material;

/*
 * Raw geometry of one neurite: what a space.cpp builder produces before the
 * setup chain runs (unstrained sizes, per-element strain, topology).
 * Element 0 is the fictitious first element; terminals[0] is conventionally 0.
 */
typedef struct nrt_geometry {
  int32_t num_elem;
  const int32_t* mother;     /* [N] (-1 for element 0)                          */
  const int32_t* dr;         /* [N] right daughter (== self for terminals)      */
  const int32_t* dl;         /* [N] left daughter (0 when none)                 */
  const uint8_t* branching;  /* [N] 1 at branching points                       */
  const double* diameter;    /* [N] unstrained diameter [m]                     */
  const double* dx0;         /* [N] unstrained element length [m]               */
  const double* epsilon;     /* [N] axial strain                                */
  const double* input;       /* [N] stimulus current when active [A]            */
  const int32_t* material;   /* [N] index into materials                        */
  int32_t num_materials;
  const nrt_material* materials;
  int32_t myelinate;         /* cable_init_constants flag (params.HH_model)     */
  int32_t num_term;
  const int32_t* terminals;  /* [num_term]                                      */
  double E_MAX;              /* damage threshold for reversal potentials        */
} nrt_geometry;

/*
 * Derived per-element state: exactly the arrays the reference GPU solver
 * constructors copy out of the discretization objects
 * (solverEx-gpu.cpp:327-399, solverIm-gpu.cpp:527-595) plus the topology the
 * CPU solver reads through element[] (solverEx-cpu.cpp:92-116).
 * Caller-owned; read only during nrt_create.
 */
typedef struct nrt_problem {
  int32_t num_elem;
  double dT;                 /* time step [s]                                   */
  const double* potential0;  /* [N] potential[0] : V at the current step       */
  const double* potential1;  /* [N] potential[1] : the second buffer            */
  const double* dx;          /* [N] strained element length                     */
  const double* c_m;         /* [N] membrane capacitance per length             */
  const double* r_a;         /* [N] axial resistance per length                 */
  const double* W;           /* [N] W (CT: -1/r_m; HH: overwritten each step)   */
  const double* K;           /* [N] K (CT: Vr/r_m; HH: overwritten each step)   */
  const double* input;       /* [N] input_on values                             */
  const int32_t* mother;     /* [N]                                             */
  const int32_t* dr;         /* [N]                                             */
  const int32_t* dl;         /* [N]                                             */
  const uint8_t* branching;  /* [N]                                             */
  int32_t num_term;
  const int32_t* terminals;  /* [num_term]                                      */
  int32_t num_hh;
  const int32_t* hh_elem;    /* [num_hh] element index of each HH element       */
  const double* hh_m;        /* [num_hh] gates                                  */
  const double* hh_h;
  const double* hh_n;
  const double* E_Na;        /* [num_hh] reversal potentials                    */
  const double* E_K;
  const double* E_L;
  const double* g_Na;        /* [num_hh] maximal conductances (per length)      */
  const double* g_K;
  const double* G_L;
  const double* ls_Na;       /* [num_hh] left shifts                            */
  const double* ls_K;
  const double* rest_pot;    /* [num_hh] resting potential                      */
} nrt_problem;

/* Device-computed setup (nrt_setup) output buffers, caller-owned host arrays. */
typedef struct nrt_setup_out {
  double* dx; double* c_m; double* r_a; double* r_m; double* W; double* K;
  double* crit_dT;           /* [N] critical time step per element              */
  double* rest;              /* [N] resting potential per element               */
  int32_t* hh_elem;          /* [num_hh] (num_hh = number of HH-kind elements)  */
  double* hh_m; double* hh_h; double* hh_n;
  double* E_Na; double* E_K; double* E_L;
  double* g_Na; double* g_K; double* G_L;
  double* ls_Na; double* ls_K; double* hh_rest;
  double dT_crit;            /* min crit_dT over elements 1..N-2                */
  int32_t num_hh;
} nrt_setup_out;

typedef struct nrt_options {
  int32_t kind;         /* NRT_EXPLICIT / NRT_IMPLICIT                          */
  int32_t precision;    /* NRT_FP64 / NRT_FP32_REFINE                           */
  int32_t krylov;       /* NRT_BICGSTAB / NRT_GMRES / NRT_DIRECT                */
  int32_t restart;      /* GMRES restart length, 1..512 (reference CPU: 150)   */
  int32_t tol_mode;     /* NRT_TOL_ABSOLUTE / NRT_TOL_RELATIVE                  */
  int32_t x0_mode;      /* NRT_X0_PREVIOUS / NRT_X0_RHS / NRT_X0_ZERO (Krylov
                           only; the tree solve's refinement starts from V)     */
  int32_t check_every;  /* BiCGSTAB convergence test period (reference: 30)    */
  int32_t max_iter;     /* 0 -> N (reference MAX_ITER)                          */
  int32_t device;       /* -1 -> $idGPU or 0                                    */
  int32_t num_blocks;   /* 0 -> autotuned grid size for persistent kernels     */
  double tol;           /* Krylov tolerance (see tol_mode; FP64 forms)          */
  double refine_tol;    /* FP32_REFINE: outer FP64 residual tolerance, relative
                           and Jacobi-scaled (D = diag(A)). Tree solve, max
                           norms (a bound per element, whatever N):
                           |D^-1 (b - A x)|_inf <= refine_tol |D^-1 b|_inf;
                           GMRES, 2-norms (its residual estimate):
                           |D^-1 (b - A x)|_2 <= refine_tol |D^-1 b|_2        */
  int32_t refine_max;   /* FP32_REFINE tree solve: maximum refinement sweeps
                           (GMRES: each restart cycle is one sweep; cycles are
                           bounded by max_iter Arnoldi steps in total)          */
  int32_t precond;      /* NRT_PRECOND_NONE / NRT_PRECOND_TREE                  */
  int32_t chunk;        /* tree solve: chain chunk length at level 0 (0 -> auto) */
  int32_t chunk_hi;     /* tree solve: chunk length of the reduced levels
                           (0 -> chunk if set, else auto)                        */
  int32_t batch_threads;/* batch (implicit): workgroup size per tree, 128, 256,
                           512 or 1024 (0 -> autotuned on the first launches,
                           for each size class of trees separately)              */
  int32_t tree_groups;  /* tree solve, local mode: this many workgroups of 512
                           threads, each owning a contiguous share of the rows
                           (its chunks and final nodes), one or two grid
                           barriers per solve (0 -> auto: tree solve from 16384 rows with
                           num_blocks 0, N/2900 groups (FP32: N/3400) up to 64; 1 -> off;
                           ignored in a batch)                                    */
  int32_t reserved[2];
} nrt_options;

typedef struct nrt_stats {
  int64_t steps;            /* time steps taken                                 */
  int64_t krylov_iters;     /* Krylov iterations summed over steps              */
  int64_t refine_sweeps;    /* FP32_REFINE: refinement sweeps summed over steps */
  int32_t last_iters;       /* Krylov iterations of the last step               */
  int32_t last_status;      /* status of the last step                          */
  double last_residual;     /* ||b - A x||_2 at exit of the last step
                               (FP32_REFINE: ||D^-1 (b - A x)||, the norm of
                               refine_tol's test)                               */
  double last_rhs_norm;     /* ||b||_2 of the last step (FP32_REFINE: ||D^-1 b||,
                               as last_residual)                                */
  double kernel_ms;         /* device time of the dominant kernel, last run     */
  int64_t kernel_launches;  /* launches of the dominant kernel, last run        */
  int64_t matrix_nnz;       /* implicit: non-zeros of the assembled CSR         */
  int32_t num_blocks;       /* persistent kernels: workgroups in the grid       */
  int32_t restarts;         /* Krylov breakdown restarts summed over steps      */
  int32_t tree_form;        /* tree solve launch: NRT_TREE_BLOCK, _GRID, _LOCAL,
                               _LOCAL_LDS (group-resident), -1 otherwise        */
  int32_t reserved;
} nrt_stats;

/* nrt_stats.tree_form */
#define NRT_TREE_BLOCK      0  /* one workgroup                                  */
#define NRT_TREE_GRID       1  /* level-synchronous grid (a barrier per level)   */
#define NRT_TREE_LOCAL      2  /* local mode, levels in global memory            */
#define NRT_TREE_LOCAL_LDS  3  /* local mode, each group's levels in its LDS     */

typedef struct nrt_handle nrt_handle;

/* Library / device */
int nrt_version(void);
int nrt_device_count(int* count);
const char* nrt_last_error(const nrt_handle* h);
const char* nrt_status_string(int status);
void nrt_default_options(nrt_options* opt);

/* Setup chain on the device: geometry -> derived arrays (init.cpp:138-298). */
int nrt_setup(const nrt_geometry* g, int device, nrt_setup_out* out);

/*
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nrt_setup over arrays the caller already holds on the
 * device, enqueued on `stream` (a hipStream_t; NULL: the default stream) without
 * waiting, allocating nothing. Every array pointer of `g` (diameter, dx0,
 * epsilon, material, materials: num_materials nrt_material records) and of
 * `out` is a device pointer; the topology fields of `g` are not read. The
 * per-element outputs take num_elem values; the HH outputs and hh_elem take the
 * first hh_capacity HH elements. On completion *num_hh_out (device) holds the
 * number of HH elements (-1 if an element names a material out of range; then
 * the outputs are incomplete) and *dT_crit_out (device) the critical step, as
 * nrt_setup reports them; with *num_hh_out > hh_capacity the HH outputs hold
 * only the first hh_capacity. At most 67,108,864 elements: the kernels keep their
 * per-block counts in a fixed per-device scratch array of 1 MiB + 256 B, held
 * until the process ends. Calls that overlap on one device race on that array.
 */
int nrt_setup_noalloc(const nrt_geometry* g, int device, void* stream,
                      const nrt_setup_out* out, int32_t hh_capacity, double* dT_crit_out,
                      int32_t* num_hh_out);

/* Solver lifecycle: the Solver interface of inc/solver.h. */
int nrt_create(const nrt_problem* p, const nrt_options* opt, nrt_handle** out);
int nrt_update(nrt_handle* h, int input_active);      /* Solver::update    */
int nrt_calculate(nrt_handle* h);                     /* Solver::calculate */
int nrt_snapshot(nrt_handle* h, double* host_out);    /* Solver::snapshot  */
int nrt_destroy(nrt_handle* h);                       /* Solver::~Solver   */

/*
 * Device-resident time loop (neurite.cpp:224-255 without host round trips):
 * runs steps i = first_step .. first_step+num_steps-1; the stimulus is on for
 * active_first <= i <= active_last; every trace_every steps (counting as the
 * reference's GL_cont) V[trace_elem] is recorded into trace_out.
 * trace_out may be NULL when trace_every <= 0. Returns the number of samples in
 * *num_samples.
 */
int nrt_run(nrt_handle* h, int64_t first_step, int64_t num_steps,
            int64_t active_first, int64_t active_last,
            int32_t trace_elem, int64_t trace_every,
            double* trace_out, int64_t* num_samples);
int nrt_get_stats(nrt_handle* h, nrt_stats* st);
/* Synchronise the handle's stream (for timing by the caller). */
int nrt_sync(nrt_handle* h);

/* ------------------------------------------------------------------ batches
 * No counterpart in the reference, which advances one neurite per process
 * (neurite.cpp:185-255). A batch advances many independent neurites ("slots")
 * together, each slot being exactly the problem a handle would solve alone:
 *   implicit: each nrt_batch_step moves every active slot by up to k steps, a
 *     workgroup per tree (two or more for the largest), each with its own tree
 *     solve and its own refinement scalars (NRT_DIRECT, FP64 or FP32_REFINE); the
 *     slots of each size class (< 4,096 rows, < 16,384, the rest) run as a launch
 *     of their own, concurrently, while every class fills at least half the
 *     device, else all in one launch;
 *   explicit: one thread per element over all slots, one launch per step, the k
 *     steps of a call captured once as a HIP graph and replayed.
 * Slots keep their whole state on the device and join or leave between calls
 * (continuous batching). A slot counts its own steps from 1: its stimulus is on
 * for steps a0..a1 and it completes after num_steps. Probe traces follow
 * nrt_run: with trace_every >= 0, V[trace_elem] after every (trace_every+1)-th
 * step. Calls on one batch are not re-entrant. */
typedef struct nrt_batch nrt_batch;

typedef struct nrt_slot_info {
  int32_t N;               /* elements                                          */
  int32_t status;          /* NRT_OK, or the slot's first failure               */
  int64_t steps_done;
  int64_t num_steps;
  int64_t inner_iters;     /* tree solves / refinement sweeps (implicit)        */
  int64_t trace_samples;   /* probe samples recorded so far                     */
  int64_t nnz;             /* matrix entries (implicit)                         */
} nrt_slot_info;

typedef struct nrt_batch_stats {
  int64_t launches;          /* step launches (graph replays count each step) */
  int64_t slot_steps;        /* sum over slots of the steps advanced          */
  int64_t compartment_steps; /* sum over slots of N x steps advanced          */
  double kernel_ms;          /* device time of the step launches              */
  int32_t active_slots;      /* slots with steps left                         */
  int32_t threads;           /* implicit: workgroup size of the size class
                                holding most compartments                     */
  int64_t device_bytes;      /* device memory held by the slots               */
  int32_t class_threads[3];  /* implicit: workgroup size per size class (slots
                                of < 4,096 rows, < 16,384, the rest), tuned by
                                timing the class's launches 2 to 5 (its first
                                also factors its slots; classes that share a
                                launch run with the size of the class holding
                                the most rows)                                 */
  int32_t reserved;
} nrt_batch_stats;

int nrt_batch_create(const nrt_options* opt, nrt_batch** out);
/* Admits a problem as a new slot, starting at the next step; fails with
   NRT_ERR_ADMISSION when free device memory does not cover its footprint. */
int nrt_batch_add(nrt_batch* b, const nrt_problem* p, int64_t num_steps, int64_t a0,
                  int64_t a1, int32_t trace_elem, int64_t trace_every, int32_t* slot);
int nrt_batch_step(nrt_batch* b, int64_t k);
int nrt_batch_snapshot(nrt_batch* b, int32_t slot, double* host_out);
int nrt_batch_trace(nrt_batch* b, int32_t slot, double* host_out, int64_t capacity,
                    int64_t* count);
int nrt_batch_slot_info(nrt_batch* b, int32_t slot, nrt_slot_info* info);
int nrt_batch_remove(nrt_batch* b, int32_t slot);
int nrt_batch_get_stats(nrt_batch* b, nrt_batch_stats* out);

/* Launch shapes of an implicit batch, per size class. nrt_batch_tune times trial
 * batches built from a caller's sample of problems (each class: its problems,
 * largest first, up to rows_per_class compartments) over every pair of
 *   plan:   (level-0 chunk, reduced-level chunk) in {auto, 8/8, 12/8, 16/8, 24/8,
 *           16/5, 16/12} (auto: the rule nrt_options.chunk = 0 applies), and
 *   threads in {128, 256, 512, 1024} (workgroup size),
 * one untimed launch and one timed launch of `steps` steps each, and keeps the
 * pair with the least kernel time per compartment-step. Slots added afterwards are
 * planned with their class's chunk lengths and launched with its workgroup size
 * (slots already in the batch keep their plans). Classes with no problem in the
 * sample keep the defaults (chunks by the auto rule, workgroup size tuned on the
 * first launches). An explicit chunk or batch_threads in the batch's options
 * restricts the candidates to that value. */
typedef struct nrt_batch_tune_result {
  int32_t threads[3];            /* chosen workgroup size per size class (0: no
                                    sample in that class)                        */
  int32_t chunk[3];              /* chosen level-0 chunk length (0: auto rule)    */
  int32_t chunk_hi[3];           /* chosen reduced-level chunk length (0: auto)   */
  int32_t slots[3];              /* sample problems timed per class               */
  int64_t rows[3];               /* their compartments                            */
  double ns_per_row_step[3];     /* kernel time per compartment-step, chosen pair */
  double auto_ns_per_row_step[3];/* the same with auto chunks and 256 threads     */
  int32_t trials;                /* (plan, threads) pairs timed                   */
  int32_t reserved;
  double seconds;                /* wall time of the tuning                       */
} nrt_batch_tune_result;
int nrt_batch_tune(nrt_batch* b, const nrt_problem* const* sample, int32_t n, int64_t steps,
                   int64_t rows_per_class, nrt_batch_tune_result* out);
/* Sets the launch shape of one size class (0: < 4,096 rows, 1: < 16,384, 2: the
 * rest) as nrt_batch_tune would: workgroup size 128, 256, 512 or 1024 (0: tuned on
 * the first launches), chunk lengths (0: the auto rule). A shape tuned once can so
 * be kept and reused by later batches of the same kind of trees. */
int nrt_batch_set_shape(nrt_batch* b, int32_t size_class, int32_t threads, int32_t chunk,
                        int32_t chunk_hi);
const char* nrt_batch_last_error(const nrt_batch* b);
int nrt_batch_destroy(nrt_batch* b);

/* ------------------------------------------------------------------ serving
 * No counterpart in the reference. A server runs a loop thread over one batch:
 * submitted simulations queue in FIFO order, are admitted as device memory
 * allows (free HBM from hipMemGetInfo against each request's footprint), keep
 * their state on the device, and advance together steps_per_launch steps at a
 * time; requests that arrive meanwhile join at the next launch (continuous
 * batching) and completed ones leave, their final V and probe trace kept until
 * fetched. A request above big_threshold elements runs on its own grid-wide
 * handle, stepped in the same rhythm on the same stream. Submission, polling and
 * fetching are thread-safe. Latency = completion - submission. */
typedef struct nrt_server nrt_server;

typedef struct nrt_server_config {
  int32_t steps_per_launch;  /* k: steps per launch; implicit requests of N
                                elements take max(1, min(8k, k*16384/N)) (default
                                16)                                              */
  int32_t big_threshold;     /* elements above which a request gets its own
                                grid-wide handle instead of a batch slot (default
                                2^30: none; batch slots above 24,576 rows already
                                run on several workgroups)                       */
  int32_t max_big;           /* such requests running at once (default 4)        */
  int32_t keep_results;      /* keep V and trace until fetched (default 1)       */
  int32_t builders;          /* threads building admitted requests' device state
                                (1..16, default 4)                               */
} nrt_server_config;

typedef struct nrt_request_result {
  int64_t id;
  int32_t status;            /* NRT_OK, or why the request failed               */
  int32_t N;
  int64_t num_steps;
  int64_t steps_done;
  double submit_s;           /* seconds since the server was created            */
  double start_s;            /* admission                                        */
  double finish_s;           /* completion                                       */
  int64_t trace_samples;
} nrt_request_result;

typedef struct nrt_server_stats {
  int64_t submitted, completed, failed;
  int64_t queued, running;    /* at the start of the last loop iteration        */
  int64_t launches;           /* loop iterations that stepped work              */
  int64_t compartment_steps;  /* of completed requests                          */
  double wall_s;              /* since creation                                  */
  double p50_ms, p99_ms, mean_ms;  /* request latency                            */
  double batch_kernel_ms;     /* device time of the batched launches            */
  /* loop thread wall time by phase (ms): enqueueing launches (with slot-table
     rebuilds), admitting requests (slot construction), waiting for the device,
     retiring completed requests (result copies, slot release)                  */
  double launch_ms, admit_ms, wait_ms, retire_ms;  /* admit_ms: unused (builders) */
} nrt_server_stats;

void nrt_default_server_config(nrt_server_config* c);
int nrt_server_create(const nrt_options* opt, const nrt_server_config* cfg, nrt_server** out);
int nrt_server_submit(nrt_server* s, const nrt_problem* p, int64_t num_steps, int64_t a0,
                      int64_t a1, int32_t trace_elem, int64_t trace_every, int64_t* id);
/* Completed requests not polled before, up to capacity. */
int nrt_server_poll(nrt_server* s, nrt_request_result* out, int32_t capacity, int32_t* count);
/* Copies a completed request's final V [N] and trace (up to trace_capacity) and
   releases them; returns the request's status. */
int nrt_server_fetch(nrt_server* s, int64_t id, double* v_out, double* trace_out,
                     int64_t trace_capacity);
/* Blocks until nothing is queued or running (timeout_s < 0: no limit). */
int nrt_server_wait_idle(nrt_server* s, double timeout_s);
int nrt_server_get_stats(nrt_server* s, nrt_server_stats* out);
const char* nrt_server_last_error(const nrt_server* s);
/* Finishes the queued and running requests, then stops the loop. */
int nrt_server_destroy(nrt_server* s);

/* ------------------------------------------------- scenarios (the driver)
 * What the reference main does before a solver exists (neurite.cpp:84-184): the
 * command-line tokens (input.cpp:25-149; solver flags init.cpp:40-55), the
 * geometry builder named by the `simulation` environment variable
 * (init.cpp:84-130; builders and the neuron loader space.cpp:37-798), the
 * mechanical model's macroscopic strain from `input_data.txt`
 * (mechanical_model.cpp:21-99, init.cpp:320-333) and the measurement point
 * (init.cpp:339-358). Host work only. The driver `neurite_hip` (built beside the
 * library) feeds the geometry to nrt_setup and runs the time loop.
 *
 * Tokens beyond the reference's (ignored by it): `steps N` caps the time loop,
 * `method bicgstab|gmres|direct`, `prec fp64|mixed`, `restart M`, `tol T` choose
 * the implicit solve, `vout FILE` writes the final V of every element. */
typedef struct nrt_scenario nrt_scenario;
typedef struct nrt_scenario_info {
  nrt_geometry geometry;       /* raw geometry; arrays owned by the scenario      */
  int32_t explicit_mode;       /* token flags e / i / g                           */
  int32_t implicit_mode;
  int32_t viewer;              /* g: the OpenGL viewer (not part of this build)   */
  int32_t mes_point;           /* measurement element (init.cpp:339-358)          */
  double tspan, xspan;         /* after the builder (xspan strained)              */
  double dt_factor, nu, epsilon, E_MAX;
  double init_time, end_input, input_current;  /* stimulus window and amplitude   */
  double dist_mes;             /* measurement distance (after the strain)         */
  double epsilon_macro;        /* mechanical model (0 without strain)             */
  const char* out_path;        /* `out` token (NULL when absent)                  */
  const char* log_path;        /* `log` token                                     */
  const char* builder_log;     /* lines the builder writes to the log             */
  int64_t max_steps;           /* `steps` (0: the full run)                       */
  int32_t method;              /* `method` (-1: the reference's BiCGSTAB)         */
  int32_t precision;           /* `prec`                                          */
  int32_t restart;             /* `restart` (0: default)                          */
  double tol;                  /* `tol` (0: default)                              */
  const char* vout_path;       /* `vout`                                          */
} nrt_scenario_info;

/* argv as main's (argv[0] is the program); simulation / working_directory NULL
   -> the environment variables of those names. */
int nrt_scenario_create(int argc, const char* const* argv, const char* simulation,
                        const char* working_directory, nrt_scenario** out);
int nrt_scenario_get(const nrt_scenario* s, nrt_scenario_info* out);
const char* nrt_scenario_last_error(void);
void nrt_scenario_destroy(nrt_scenario* s);

#ifdef __cplusplus
}
#endif
#endif /* NEURITE_HIP_H */
