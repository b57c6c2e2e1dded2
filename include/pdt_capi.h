/* pdt_capi.h -- C ABI of the MI355X TV/TGV primal-dual solver library
 * (libpdt_hip.so).
 *
 * This is the boundary below the toolbox's C++ classes and Python modules:
 * plain C types, opaque handles, int status codes.  Each entry point names the
 * reference interface it replaces (paths relative to the primal-dual-toolbox
 * source tree).
 *
 * Conventions
 *  - Arrays use the reference's iu layout: size {s0, s1[, s2]} with dim 0
 *    fastest (element (x, y, c) at x + s0 * (y + s1 * c)).  A numpy array of
 *    shape (s2, s1, s0) in C order has the same memory image.
 *  - Complex data is interleaved (re, im) pairs of the real type.
 *  - Every call is blocking: it finishes its device work before returning.
 *    A handle is not thread-safe; distinct handles may be used concurrently.
 *  - Status 0 is success; on failure pdt_last_error() returns the message
 *    (thread-local), mirroring the IuException texts of the reference.
 */
#ifndef PDT_CAPI_H
#define PDT_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDT_API __attribute__((visibility("default")))

/* ---- status codes ---------------------------------------------------- */
enum {
  PDT_OK = 0,
  PDT_ERR_ARG = 1,     /* invalid argument / null handle                     */
  PDT_ERR_SIZE = 2,    /* IU_SIZE_CHECK failures (mricartesianoperator.cu:177-179) */
  PDT_ERR_STATE = 3,   /* "No operator is set!", "No noisy data is set!",
                          "There is no result yet!" (tgvmrioptimizer.cu:96-110,
                          tgvoptimizerwithop.cu:41-47)                        */
  PDT_ERR_CONST = 4,   /* constant count/type (operatorbase.h:242-256,276-287,493-502) */
  PDT_ERR_HIP = 5,     /* HIP runtime error (IU_CUDA_CHECK)                   */
  PDT_ERR_UNSUPPORTED = 6
};

/* ---- element types ---------------------------------------------------- */
enum {
  PDT_F32 = 0,  /* float    (iu float)   */
  PDT_C64 = 1,  /* complex<float>  (iu float2)  */
  PDT_F64 = 2,  /* double   (iu double)  */
  PDT_C128 = 3  /* complex<double> (iu double2) */
};

/* ---- operators (src/operator) ----------------------------------------- */
enum {
  PDT_OP_CARTESIAN = 0,              /* MriCartesianOperator (mricartesianoperator.h:31-70)            */
  PDT_OP_CARTESIAN_REMOVE_ROOS = 1,  /* MriCartesianRemoveROOSOperator (mricartesianoperator.h:79-118) */
  PDT_OP_SAMPLING = 2                /* MriSamplingOperator (mrisamplingoperator.h)                     */
};

/* ---- solvers (src/optimizer) ------------------------------------------ */
enum {
  PDT_SOLVER_TV = 0,      /* TvOptimizer (tvoptimizer.h:63-121)                 */
  PDT_SOLVER_TGV = 1,     /* TgvOptimizer (tgvoptimizer.h)                      */
  PDT_SOLVER_TGV_MRI = 2, /* TgvMriOptimizer (tgvmrioptimizer.h:75-150)         */
  PDT_SOLVER_TV_OP = 3,   /* TvOptimizerWithOp (tvoptimizerwithop.h)            */
  PDT_SOLVER_TGV_OP = 4   /* TgvOptimizerWithOp (tgvoptimizerwithop.h)          */
};

typedef struct pdt_array pdt_array;
typedef struct pdt_op pdt_op;
typedef struct pdt_solver pdt_solver;

/* Union of TvParameters / TgvParameters / TgvMriParameters
 * (tvoptimizer.h:32-61, tgvoptimizer.h:34-67, tgvmrioptimizer.h:33-66) plus
 * fields that have no reference counterpart (marked NEW). */
typedef struct pdt_params {
  double lambda;      /* data weight (TV/TGV/with-op); unused by TGV-MRI (fixed 1) */
  double alpha0;      /* TGV second-order weight                               */
  double alpha1;      /* TGV first-order weight                                */
  double reduction;   /* TGV-MRI alpha continuation factor                     */
  uint32_t max_iter;  /* the solvers run max_iter + 1 iterations               */
  uint32_t check;     /* print interval when verbose                           */
  double gap_tol;     /* NEW: stop once rel. primal-dual gap <= gap_tol (0 = off) */
  uint32_t gap_every; /* NEW: evaluate the gap every gap_every iterations (0 = never) */
  int32_t dual_fp16;  /* NEW: store TGV dual fields p, q as fp16 (TGV denoising)  */
  int32_t use_graph;  /* NEW: replay iteration chunks through a hipGraph (default 1) */
  double radius_u;    /* NEW: partial-gap ball radius for u (0 = 0.01 ||A^*f||)  */
  double radius_v;    /* NEW: partial-gap ball radius for v (0 = sqrt(8) radius_u) */
} pdt_params;

typedef struct pdt_stats {
  uint32_t iters;     /* iterations run by the last solve                      */
  uint32_t n_gap;     /* gap evaluations in the last solve                     */
  double primal;      /* last primal energy P                                  */
  double dual;        /* last dual value D                                     */
  double gap;         /* P - D                                                 */
  double rel_gap;     /* (P - D) / |P|                                         */
  double ms;          /* device time of the iteration loop (hipEvent), ms      */
  int32_t converged;  /* 1 if gap_tol was reached                              */
  double ms_to_stop;  /* NEW: in a batched solve, host time from the start of the
                         loop to the end of the chunk in which this slice stopped;
                         otherwise equal to ms                                  */
} pdt_stats;

/* ---- library ------------------------------------------------------------ */
PDT_API const char* pdt_last_error(void);
PDT_API int pdt_version(void);
/* Selects the HIP device for subsequent handle creation. */
PDT_API int pdt_set_device(int device);

/* ---- device arrays: replace iu::LinearDeviceMemory<T, N> ----------------- */
PDT_API int pdt_array_create(int dtype, int ndim, const int64_t* size, pdt_array** out);
/* Wraps existing device memory (not owned, e.g. a torch tensor). */
PDT_API int pdt_array_wrap(int dtype, int ndim, const int64_t* size, void* device_ptr,
                           pdt_array** out);
PDT_API int pdt_array_upload(pdt_array* a, const void* host, size_t bytes);
PDT_API int pdt_array_download(const pdt_array* a, void* host, size_t bytes);
PDT_API int pdt_array_info(const pdt_array* a, int* dtype, int* ndim, int64_t* size3);
PDT_API void* pdt_array_device_ptr(pdt_array* a);
PDT_API int pdt_array_destroy(pdt_array* a);

/* ---- operators: replace OperatorBase<In,Out> and the MRI operators -------- */
/* dtype: PDT_C64 or PDT_C128 (the reference instantiates float2/double2,
 * mricartesianoperator.cu:436-444). */
PDT_API int pdt_op_create(int kind, int dtype, pdt_op** out);
/* OperatorBase::addConstant (operatorbase.h:238-257).  Order matters:
 * Cartesian ops take [0] coil sensitivities (complex, 3D {W, Hk, C}) and
 * [1] mask (real, 2D {W, Hk}); the sampling op takes [0] mask. */
PDT_API int pdt_op_add_constant(pdt_op* op, const void* host, int dtype, int ndim,
                                const int64_t* size);
PDT_API int pdt_op_num_constants(const pdt_op* op, int* have, int* required);
/* OperatorBase::forward / adjoint (operatorbase.h:297-312): dst must be
 * pre-allocated at the right size; sizes and constants are checked first. */
PDT_API int pdt_op_forward(pdt_op* op, const pdt_array* src, pdt_array* dst);
PDT_API int pdt_op_adjoint(pdt_op* op, const pdt_array* src, pdt_array* dst);
/* NEW, no counterpart in the reference: pdt_op_forward (forward != 0) or
 * pdt_op_adjoint (forward == 0) without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  src and dst as for those calls.
 * Enqueues on `stream` (a hipStream_t; NULL = the null stream) and returns
 * without waiting.  The intermediates live in one array of
 * PDT_NOALLOC_SCRATCH_BYTES per GPU, part of the library's code object and
 * held until the process ends: every call on a GPU shares it, so calls that
 * overlap on one GPU race; each call rewrites everything it reads.
 * PDT_ERR_STATE until one pdt_op_forward, pdt_op_adjoint or
 * pdt_op_adjointness has prepared the operator (that call allocates);
 * PDT_ERR_SIZE when the intermediates (2 C H nc + ceil(C / 4) H W complex
 * values for C coils, H x W images and nc sampled lines) exceed the array. */
#define PDT_NOALLOC_SCRATCH_BYTES (32u << 20)
PDT_API int pdt_op_apply_noalloc(pdt_op* op, int forward, const pdt_array* src, pdt_array* dst, void* stream);
/* getInputSize / getOutputSize (mricartesianoperator.cu:183-194,296-308). */
PDT_API int pdt_op_input_size(pdt_op* op, int out_ndim, const int64_t* out_size,
                              int* in_ndim, int64_t* in_size);
PDT_API int pdt_op_output_size(pdt_op* op, int in_ndim, const int64_t* in_size,
                               int* out_ndim, int64_t* out_size);
/* adjointnessCheck (mricartesianoperator.cu:267-282): random u, p from the
 * reference's seeded generator; returns <Au, p> and <u, A^*p>. */
PDT_API int pdt_op_adjointness(pdt_op* op, double* lhs_re, double* lhs_im, double* rhs_re,
                               double* rhs_im);
PDT_API int pdt_op_name(const pdt_op* op, char* buf, size_t len);
PDT_API int pdt_op_destroy(pdt_op* op);

/* ---- solvers: replace Tv/Tgv/TgvMri/TvWithOp/TgvWithOp optimizers --------- */
/* dtype of the primal image: PDT_F32/PDT_F64 (denoising only) or
 * PDT_C64/PDT_C128. */
PDT_API int pdt_solver_create(int kind, int dtype, pdt_solver** out);
PDT_API int pdt_params_default(int kind, pdt_params* p);
PDT_API int pdt_solver_get_params(const pdt_solver* s, pdt_params* p);
PDT_API int pdt_solver_set_params(pdt_solver* s, const pdt_params* p);
/* setOperator (tgvmrioptimizer.cu:56-60): shared ownership (ref-counted). */
PDT_API int pdt_solver_set_operator(pdt_solver* s, pdt_op* op);
/* setNoisyData (tgvmrioptimizer.cu:82-91): copies f to the solver and resets
 * the dual variable r to 0. */
PDT_API int pdt_solver_set_noisy_data(pdt_solver* s, const pdt_array* f);
/* setInput0 (tgvmrioptimizer.cu:63-79): u0 copied to u and u_bar, duals reset
 * to 0.  v0 (NEW, may be NULL) warm-starts v = v_bar ({W, H, 2} planes x, y);
 * with v0 = NULL this is the reference behaviour. */
PDT_API int pdt_solver_set_input0(pdt_solver* s, const pdt_array* u0, const pdt_array* v0);
/* solve (tvoptimizer.cu:77-125 etc.): runs max_iter + 1 iterations, or stops
 * early at gap_tol when the gap is enabled. */
PDT_API int pdt_solver_solve(pdt_solver* s, int verbose);
/* getResult: borrowed handle to u, valid until the next solve/set_input0 or
 * destroy.  STATE error before the first solve (the with-op solvers throw
 * "There is no result yet!", tvoptimizerwithop.cu:41-47). */
PDT_API int pdt_solver_get_result(pdt_solver* s, pdt_array** u);
/* NEW: v of the TGV solvers ({W, H, 2}), borrowed. */
PDT_API int pdt_solver_get_v(pdt_solver* s, pdt_array** v);
PDT_API int pdt_solver_get_stats(const pdt_solver* s, pdt_stats* st);
PDT_API int pdt_solver_destroy(pdt_solver* s);

/* NEW: per-kernel timing of a configured solver.  A hipGraph of `reps`
 * iterations (stages in solve order) is timed with a hipEvent pair on the
 * solver's stream; that time per iteration is divided among the stages in
 * proportion to event pairs around their direct launches in iteration order
 * (an event pair adds the same dispatch delay to every launch).  Reported
 * per kernel: the duration of one launch in the loop and the algorithmic HBM
 * bytes of one launch (every distinct element the launch reads or writes,
 * once, at storage precision).  Stages that launch nothing for this solver
 * (the separate primal kernel when every slice has one coil group) are
 * omitted; *n receives the number of entries written.  A single MRI slice
 * that runs paired iterations (pdt_tiles.unpaired) reports their two
 * launches, "k_cols_dual" and "k_rows_adj_next"; otherwise the stages
 * ("k_rows_fwd", "k_cols", "k_rows_adj", "k_mri_primal"; denoising:
 * "k_tv_dual"/"k_tv_primal", "k_tgv_primal"/"k_tgv_dual", or one
 * "k_tv_block"/"k_tgv_block" entry per launch of the temporally blocked
 * path).  Runs from the solver's current state and puts its iterates back
 * at the end. */
typedef struct pdt_kernel_prof {
  char name[32];
  double ms;
  double bytes;
} pdt_kernel_prof;
PDT_API int pdt_solver_profile(pdt_solver* s, int reps, pdt_kernel_prof* out, int cap, int* n);
/* NEW: pdt_solver_profile over the launches of a batch (pdt_batch_solve):
 * per kernel, the duration of one launch over every slice and the
 * algorithmic bytes of that launch summed over the slices. */
PDT_API int pdt_batch_profile(pdt_solver** solvers, int count, int reps, pdt_kernel_prof* out, int cap, int* n);

/* NEW: tile shapes and launch shapes of the fused kernels, one choice per
 * problem shape (kept in a process-wide cache, persisted to the file named by
 * the environment variable PDT_TUNE_CACHE, if set). */
typedef struct pdt_tiles {
  int32_t cb;      /* coils per LDS chunk of the forward row pass (MRI)           */
  int32_t cg;      /* columns per group of the column pass (MRI)                 */
  int32_t rpi;     /* rows per work item of the denoising passes                 */
  int32_t cb_a;    /* coils per LDS chunk of the adjoint row pass (MRI)          */
  int32_t wpc[4];  /* workgroups per CU, 0 = no cap: forward row, column,
                      adjoint row and denoising passes                          */
  int32_t tb_x, tb_y, tb_k; /* ROF or TGV denoising, one slice: tb_k
                      iterations per launch on tb_x x tb_y tiles held in LDS
                      (tb_k 0: one iteration per pair of launches)               */
  int32_t dual_px; /* TGV denoising: pixels per thread of the dual step, 1 or 2
                      (2 needs even widths; 0: the default, 2 where possible)   */
  int32_t unpaired; /* MRI, one slice: 0 (default) = an iteration without the
                      gap evaluation in two launches where the adjoint row pass
                      holds every coil (the column pass also does the dual
                      step, the adjoint row pass also the next iteration's
                      forward row FFTs); 1 = always the three passes         */
} pdt_tiles;

/* NEW: autotuner.  Times every candidate tile shape (the fields of
 * pdt_tiles; for a single ROF or TGV denoising slice also temporally blocked
 * launches) on the
 * configured solvers (one batch) with hipEvents, per iteration; then, with
 * those tiles, caps of 1, 2, 3, 4 or 6 workgroups per CU (3, 4 or 6 for
 * denoising) on each pass in turn, each pass timed alone, keeping a cap only
 * where it beats no cap by 2 %.  Stores the choice in the tile cache and
 * returns it (out may be NULL) with the time per iteration it gives.  The
 * timing runs iterate on the solvers' buffers; their iterates are copied
 * aside first and put back at the end, so a solve after tuning starts where
 * it would have without it. */
PDT_API int pdt_autotune(pdt_solver** solvers, int count, int reps, pdt_tiles* out, double* ms_per_iteration);

/* NEW: a manual tile choice for the shape of these solvers (one batch),
 * stored in the same tile cache as pdt_autotune's result: cb, cg, rpi, cb_a
 * <= 0 take the default (cb_a: every coil of a row in one adjoint workgroup
 * where a single slice's straight-line adjoint path fits, else cb), values
 * above a shape's limit are clamped; wpc in
 * [0, 6]; tb_k <= 0 turns the blocked denoising launches off, and tiles
 * that do not fit the LDS are not used.  The next solve of the same shape
 * uses it. */
PDT_API int pdt_set_tiles(pdt_solver** solvers, int count, const pdt_tiles* tiles);

/* NEW: solve many configured solvers (same kind and dtype) together, with one
 * persistent work-list launch per half-iteration for the whole batch.  Each
 * solver keeps its own parameters, sizes, coils and stopping. */
PDT_API int pdt_batch_solve(pdt_solver** solvers, int count, int verbose);

/* NEW (no counterpart in the reference, which solves one problem per call,
 * tgvmrioptimizer.cu:94-203): solves count configured solvers, each alone
 * as pdt_solver_solve would, `streams` of them at a time (1..16): host
 * worker threads (the calling thread among them) take the next unsolved
 * solver in the given order and run its solve on the solver's own HIP
 * stream, so the launches of different slices overlap on the GPU.  Solvers
 * may differ in kind, dtype, size and parameters, and may share an
 * operator; each one's iterates, stop and statistics are those of its own
 * pdt_solver_solve (its device time is the elapsed time of its stream,
 * which the slices running beside it share).  Null or repeated solvers are
 * refused before any solve starts; after a failure no further solver is
 * started and the first error is returned. */
#define PDT_MAX_CONCURRENT 16
PDT_API int pdt_solve_concurrent(pdt_solver** solvers, int count, int streams, int verbose);

/* ---- NEW: learned warm start ----------------------------------------------
 * A conv3x3 network (zero padding, stride 1, ReLU between layers) that maps
 * the zero-filled reconstruction A^*f -- 2 channels (re, im) divided by its
 * RMS -- to 6 channels: a residual for u (re, im) and v (re/im of v_x, v_y).
 * The convolutions run as bf16 MFMA implicit GEMMs with fp32 accumulation;
 * parameters, gradients and Adam state are fp32.  Supported shapes: cin 2,
 * width 32, cout 6, depth >= 2 layers.  Parameter order: per layer
 * W[tap][co][ci] (tap = 3 ky + kx), then b[co].  Tensors are flat PDT_F32
 * arrays in NHWC order (channel fastest) of n images of h x w. */
typedef struct pdt_cnn pdt_cnn;
typedef struct pdt_adam {
  double lr, beta1, beta2, eps;
} pdt_adam;
PDT_API int pdt_cnn_create(int cin, int width, int depth, int cout, pdt_cnn** out);
PDT_API int pdt_cnn_destroy(pdt_cnn* net);
PDT_API int pdt_cnn_num_params(const pdt_cnn* net, size_t* n);
/* Uploads the parameters and resets the Adam moments and step count. */
PDT_API int pdt_cnn_set_params(pdt_cnn* net, const float* host, size_t n);
PDT_API int pdt_cnn_get_params(const pdt_cnn* net, float* host, size_t n);
/* y = net(x); x holds n h w cin floats, y n h w cout. */
PDT_API int pdt_cnn_forward(pdt_cnn* net, const pdt_array* x, pdt_array* y, int n, int h, int w);
/* Loss 1/(2m) sum (net(x) - t)^2 over the m = n h w cout outputs and its
 * gradient, kept for pdt_cnn_adam_step and copied to grad_host when given. */
PDT_API int pdt_cnn_loss_grad(pdt_cnn* net, const pdt_array* x, const pdt_array* t, int n, int h, int w,
                              double* loss, float* grad_host, size_t n_grad);
/* One Adam update (Kingma & Ba, Algorithm 1) with the last gradient. */
PDT_API int pdt_cnn_adam_step(pdt_cnn* net, const pdt_adam* opt);
/* pdt_cnn_loss_grad followed by pdt_cnn_adam_step. */
PDT_API int pdt_cnn_train_step(pdt_cnn* net, const pdt_array* x, const pdt_array* t, int n, int h, int w,
                               const pdt_adam* opt, double* loss);
/* Warm start of a complex64 TGV-MRI / TGV-with-operator solver: u0 = A^*f,
 * s = RMS(u0), (r_u, v') = net(u0 / s); sets u = u0 + s r_u and v = s v'
 * (p, q zeroed) as pdt_solver_set_input0 would. */
PDT_API int pdt_solver_warm_start(pdt_solver* s, pdt_cnn* net);

#ifdef __cplusplus
}
#endif
#endif /* PDT_CAPI_H */
