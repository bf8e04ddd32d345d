/*
 * chai_mi355x.h -- C ABI of the MI355X (gfx950) back end for Chai.
 *
 * This library replaces, behind Chai's DeviceBase plug point
 * (runtime/DeviceBase.hpp:55-60, runtime/DeviceMap.cpp:26-36), the OpenCL
 * code emitters and their dispatch/caching layers:
 *   autotune/Code*, autotune/EvergreenMatmulMM/MV  -> chai_hip_gemm / chai_hip_gemv
 *   enqueue/Enqueue*, memmgr/MemTransJIT           -> chai_hip_trace_* / chai_hip_rng_fill / chai_hip_reduce
 *   autotune/AutoTuneBench/EM, StandardEM           -> chai_hip_tune_*
 *   autotune/AutoTuneMemo, enqueue/JustInTimeMemo   -> chai_hip_memo_* / chai_hip_jit_*
 *   vendor/OCLdevice (buffers, queue, timing)       -> chai_hip_init / alloc / h2d / d2h / timer
 *
 * Conventions (kept from the reference):
 *   - Every entry point returns an int status: CHAI_OK (0) or a negative code.
 *     No exception crosses the ABI.  A failure maps to sub_evaluate() returning
 *     false, which makes the scheduler re-queue the trace on another device
 *     (runtime/Scheduler.cpp:221-254).  chai_hip_last_error() describes the
 *     last failure of the calling thread.
 *   - Precision codes are Chai's PrecType codes (misc/PrecType.hpp:18-21).
 *   - Arrays are dense row-major; "slots" are independent instances stacked
 *     one after another (a VectorTrace's gathered traces, common/VectorTrace.cpp).
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Entry points enqueue work and return; call chai_hip_sync() or
 *     synchronize the stream before reading results on the host.
 *   - Device pointers come from chai_hip_alloc() or any HIP allocation.
 */
#ifndef CHAI_MI355X_H
#define CHAI_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
enum {
    CHAI_OK = 0,
    CHAI_ERR_ARG = -1,        /* invalid argument / unsupported combination */
    CHAI_ERR_HIP = -2,        /* HIP runtime error                          */
    CHAI_ERR_NOMEM = -3,      /* allocation failed                          */
    CHAI_ERR_STATE = -4,      /* library not initialised / bad handle       */
    CHAI_ERR_IO = -5          /* journal file could not be read or written  */
};

/* PrecType codes (misc/PrecType.hpp:18-21) */
enum { CHAI_U32 = 0, CHAI_I32 = 1, CHAI_F32 = 4, CHAI_F64 = 8 };

/* RNG variants (api/chai.h:412-416) */
enum { CHAI_RNG_PHILOX = 1230, CHAI_RNG_THREEFRY = 1231 };

/* reduction op codes = Chai bytecodes (bytecode/ByteCodes.hpp:256-258) */
enum { CHAI_OP_DOT = 130, CHAI_OP_MEAN = 131, CHAI_OP_SUM = 132 };

/* ------------------------------------------------------------- lifecycle */
/* Replaces OCLinit + Translator::sub_initDevice (runtime/Translator.cpp:353-364). */
int chai_hip_init(int device_ordinal);
int chai_hip_shutdown(void);
const char* chai_hip_last_error(void);
/* Device properties: CU count, LDS bytes per CU, clock (kHz), arch name.
 * Replaces OCLdevice::maxWorkGroupSize and the OpenCL device queries behind
 * it (vendor/OCLdevice.hpp:544, vendor/OCLdevice.cpp:892-895). */
int chai_hip_device_info(int* cu_count, int* lds_per_cu, int* clock_khz,
                         char* arch_name, size_t arch_len);

/* ---------------------------------------------------------------- memory */
/* Replaces the OCLdevice buffer heap (vendor/OCLdevice.cpp) and
 * MemTrans host<->device transfers (memmgr/MemTrans.hpp). */
int chai_hip_alloc(size_t bytes, void** out);
int chai_hip_free(void* ptr);
int chai_hip_h2d(void* dst, const void* src, size_t bytes, void* stream);
int chai_hip_d2h(void* dst, const void* src, size_t bytes, void* stream);
int chai_hip_memset(void* dst, int value, size_t bytes, void* stream);
int chai_hip_sync(void* stream);

/* ----------------------------------------------------------------- timing */
/* hipEvent timer on a stream; replaces gettimeofday timing
 * (autotune/AutoTuneBench.cpp:116-123).  Handles are opaque. */
int chai_hip_timer_create(void** timer);
int chai_hip_timer_start(void* timer, void* stream);
int chai_hip_timer_stop(void* timer, void* stream);
int chai_hip_timer_elapsed_ms(void* timer, float* ms);   /* waits for stop */
int chai_hip_timer_destroy(void* timer);

/* -------------------------------------------------------------------- RNG */
/* Fills out[slots][H][W] with Chai's counter-based stream.
 * Replaces Enqueue::visit(AstRNGuniform/AstRNGnormal) (enqueue/EnqueueAST.cpp:496-546)
 * and the Random123 text emitters (Random123/R123Philox.cpp:288-520,
 * Random123/R123Threefry.cpp:330-520).
 *   variant : CHAI_RNG_PHILOX | CHAI_RNG_THREEFRY
 *   prec    : CHAI_U32 | CHAI_I32 | CHAI_F32 | CHAI_F64
 *   vec_len : 1, 2, 4 (32-bit) or 1, 2 (64-bit); 0 = PrecType default (4 / 2)
 *   seed    : generator seed of this call (api/RngFun.cpp:24-30)
 *   normal  : Box-Muller normals (CHAI_F32 / CHAI_F64 only)
 *   has_limits, lo, hi : uniform in [lo, hi] as u*(hi-lo)+lo
 */
int chai_hip_rng_fill(int variant, int prec, int vec_len, uint64_t seed,
                      size_t W, size_t H, size_t slots, int normal,
                      int has_limits, double lo, double hi,
                      void* out, void* stream);

/* Self-check of the f32 normals' arithmetic; no counterpart in the reference.
 * Box-Muller's log and sqrt in the f32 streams use short forms valid for
 * the uniforms a stream word yields; this runs all 2^32 words through them
 * and through the general IEEE forms and writes the number of words whose
 * log (mismatches[0]) or sqrt (mismatches[1]) differ in any bit: both zero
 * on a correct device.  Takes about 20 ms; allocates 16 bytes. */
int chai_hip_rng_check_normals(uint64_t* mismatches, void* stream);

/* ------------------------------------------------------------- reductions */
/* out[s] = sum / mean of in[s][0..n) (interp/InterpAccum.cpp:13-70 semantics:
 * accumulation is exact-in-FP64-combine, result in prec's type).  Replaces the
 * serial one-work-item reduction of enqueue/EnqueueReduce.cpp:45-103.
 * out_f64 (nullable) also receives the FP64 value before rounding. */
int chai_hip_reduce(int op, int prec, const void* in, size_t n, size_t slots,
                    void* out, double* out_f64, void* stream);

/* chai_hip_reduce without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The FP64
 * block partials live in a fixed device array of 65536 doubles (512 KiB per
 * GPU, held until the process ends).  A call needs blocks x slots of them,
 * with ceil(n / 4096) blocks per slot capped at 2048 / min(slots, 8): any
 * call with up to 256 slots fits, and a call that does not fit returns
 * CHAI_ERR_ARG before anything is launched.  Same kernels and
 * block counts as chai_hip_reduce, so results are identical to it.  Calls
 * on one GPU share the array: calls that overlap in time on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
int chai_hip_reduce_noalloc(int op, int prec, const void* in, size_t n, size_t slots,
                            void* out, double* out_f64, void* stream);

/* ------------------------------------------------------------------- GEMM */
/* Kernel family selector.  AUTO picks F64 when any operand is f64, INT for
 * two integer operands, and for f32 or mixed int / f32 operands the split-
 * BF16 form within twice the FP32-FMA kernel's error: BF16X6 below
 * K = 16384, BF16X3 from there (BF16X6 also where op(A) has more than
 * 2^25 - 256 rows or op(B) more columns: BF16X3 takes at most 2^25 padded
 * rows / columns per call and refuses larger ones). */
enum {
    CHAI_GEMM_AUTO = 0,
    CHAI_GEMM_FMA = 1,      /* plain FP32 VALU FMA                        */
    CHAI_GEMM_BF16X3 = 2,   /* split-BF16 three-product MFMA, FP32 accum  */
    CHAI_GEMM_F64 = 3,      /* FP64 MFMA                                  */
    CHAI_GEMM_BF16X6 = 5,   /* split-BF16 with three pieces per operand, six
                               MFMA products (all of order <= 2), FP32
                               accumulate: FP32-FMA-level error at twice the
                               BF16X3 MFMA work                              */
    CHAI_GEMM_INT = 4,      /* integer x integer, the interpreter's exact
                               semantics (interp/InterpMatmulBase.hpp:49-76):
                               products wrap in the promoted 32-bit type, sum
                               in k order as a double accumulator would,
                               x86 conversion to C's type; AUTO picks it for
                               u32 / i32 operand pairs                       */
    CHAI_GEMM_GEMV = 6      /* the matrix-vector kernels (chai_hip_gemv) for
                               N == 1 with contiguous x and y (ldb == 1 unless
                               trans_b, ldc == 1): HBM streaming of A, each
                               product in the operands' promoted type summed
                               in FP64, as the interpreter's matvec
                               (interp/InterpMatmulBase.hpp:37-46).  Its tile
                               variants split each dot product over waves or
                               row chunks and vary the loads in flight       */
};

/* One GEMM call (or a batch of identical-shape instances).  Mirrors the
 * exogenous parameters of autotune/KernelParams.hpp:16-371 (precA/B/C, M, N, K,
 * transA/B, general, batching, sameDataA/B) and StmtMatmul's
 * C = op(A) op(B) / C = alpha op(A) op(B) + beta C (api/MatmulFun.cpp:26-54). */
typedef struct chai_gemm_desc {
    int prec_a, prec_b, prec_c;     /* CHAI_U32/I32/F32/F64                     */
    int trans_a, trans_b;           /* op(X) = X^T when set                     */
    size_t M, N, K;                 /* op(A) is M x K, op(B) is K x N           */
    size_t lda, ldb, ldc;           /* row strides of the stored arrays (elems) */
    int general;                    /* 1: C = alpha AB + beta C                 */
    double alpha, beta;
    size_t batch;                   /* instances (0 or 1 = single)              */
    int same_a, same_b;             /* batch shares A / B (sameDataA/B)         */
    size_t stride_a, stride_b, stride_c; /* elements between instances          */
    int mode;                       /* CHAI_GEMM_*                              */
    int variant;                    /* tile variant; -1 = tuned choice / default */
} chai_gemm_desc;

/* The tile-variant search space of a kernel family (autotune/KernelParams.hpp's
 * endogenous parameters: work-group shape, inner blocking).  params receives
 * {BM, BN, BK, waves_m, waves_n}; for CHAI_GEMM_GEMV {split, loads in flight,
 * 0, 0, 0}: waves per row of A (for op(A) = A^T: row chunks sized so that
 * about 1024 x split blocks run) and 16-byte loads in flight per lane (the
 * MV generator's g / bW space, autotune/EvergreenMatmulMV.cpp:189-264). */
int chai_hip_gemm_num_variants(int mode, int* count);
int chai_hip_gemm_variant_params(int mode, int variant, int* params);

/* Replaces Enqueue::visit(StmtMatmul/StmtMatmulAuto) and the Evergreen MM
 * kernels (enqueue/EnqueueMatmul.cpp:24-161, autotune/EvergreenMatmulMM.cpp:1077-1630). */
int chai_hip_gemm(const chai_gemm_desc* desc, const void* A, const void* B, void* C, void* stream);

/* Elementwise epilogue fused into the store of an f32 GEMM result c (after
 * alpha / beta).  CHAI_EPI_AFFINE_MAX computes max(c * scale + shift, lo) with
 * each operation rounded in f32: the statements `D = C * 0.5f + 1;
 * E = max(D, 0)` of a trace that follows a matmul, which the reference runs
 * as a second kernel over C (enqueue/EnqueueIdSpace.cpp:19-320). */
enum { CHAI_EPI_NONE = 0, CHAI_EPI_AFFINE_MAX = 1 };
typedef struct chai_epilogue {
    int kind;                       /* CHAI_EPI_*                               */
    float scale, shift, lo;
} chai_epilogue;

/* chai_hip_gemm with an epilogue (FMA / BF16X3 families; epi may be NULL). */
int chai_hip_gemm_epilogue(const chai_gemm_desc* desc, const void* A, const void* B, void* C,
                           const chai_epilogue* epi, void* stream);

/* --------------------------------------------------------- ragged batches */
/* The replacement for gathering identical traces from many host threads into
 * one VectorTrace (runtime/Scheduler.cpp:48-307, common/VectorTrace.cpp:9-100,
 * enqueue/EnqueueMatmulAuto.cpp:57-68 setBatching): instances of
 * C = A B + epilogue with DIFFERENT shapes run as one persistent
 * split-BF16x3 launch: tile variants 0..7 (the dense path's shapes) after one
 * split pre-pass launch, variants 8..9 reading the f32 operands directly with
 * no pre-pass (the split happens inside each tile; they need row strides
 * lda, ldb below 2^23).  Results are identical for every variant.  The
 * persistent kernel pulls output tiles from eight per-XCD queues, each
 * filled largest-instance-first
 * (LPT by M N K) and balanced by work; a block drains its own XCD's queue, then
 * steals from the others.  f32 operands, row-major, no transposes. */
typedef struct chai_ragged_item {
    const float* A;                 /* M x K, row stride lda                    */
    const float* B;                 /* K x N, row stride ldb                    */
    float* C;                       /* M x N, row stride ldc                    */
    size_t M, N, K;
    size_t lda, ldb, ldc;
} chai_ragged_item;

/* Build a plan: validates the items, allocates the tile queues and, for
 * variants 0..7, the split workspace (the plan owns them), picks the tile
 * variant (-1 = the one the ragged tuner chose last, else variant 0). */
int chai_hip_ragged_plan(const chai_ragged_item* items, size_t count, const chai_epilogue* epi,
                         int variant, void** plan);
/* Run the plan on `stream` (asynchronous).  use_graph = 1 captures the
 * launches (pre-pass or counter reset, then the persistent kernel) into a
 * hipGraph on the first call and replays it afterwards. */
int chai_hip_ragged_run(void* plan, int use_graph, void* stream);
/* info[0..6]: output tiles, persistent blocks, workspace bytes, useful flops
 * (sum 2MNK), padded flops (tiles x BM x BN x 2Kp), variant, split blocks
 * (0 for the f32-fed variants). */
int chai_hip_ragged_info(void* plan, uint64_t* info);
/* hipEvent durations (ms) of the last run's split pre-pass (or counter
 * reset) and persistent kernel (not available for graph replays: returns
 * CHAI_ERR_STATE). */
int chai_hip_ragged_last_timing(void* plan, float* split_ms, float* core_ms);
int chai_hip_ragged_release(void* plan);
/* Number of ragged tile variants (valid variant arguments are 0 .. n - 1). */
int chai_hip_ragged_num_variants(int* n);
/* Time every tile variant on this batch (one warm-up and `trials` runs each)
 * and remember the fastest for chai_hip_ragged_plan(..., -1, ...). */
int chai_hip_ragged_tune(const chai_ragged_item* items, size_t count, const chai_epilogue* epi,
                         int trials, int* best_variant, float* best_ms, void* stream);

/* Thread-safe batcher: any number of host threads submit items; the thread
 * whose item fills the batch (`expected` items) runs it on the batcher's own
 * stream, and every submitter returns when the batches holding its items have
 * finished (Scheduler::wait semantics, runtime/Scheduler.cpp:307).  When
 * timeout_ms > 0, a submitter that has waited that long for a batch that is
 * not full runs it as it is (the reference's boss loop likewise runs what it
 * has gathered after a wait, runtime/Scheduler.cpp:48-100); timeout_ms = 0
 * waits for a full batch.  A batch with the same items as the previous one
 * replays its hipGraph.  variant -1 tunes the tile variant on the first batch. */
int chai_hip_batcher_create(size_t expected, const chai_epilogue* epi, int variant, int use_graph,
                            int timeout_ms, void** batcher);
/* Queue `count` items (one host thread's trace) and wait for them.  `stream`
 * is the stream whose queued work produces the items' A and B (NULL = the
 * null stream): the submit records an event on it and the batch's own stream
 * waits for that event before the GEMMs read the operands, so a submitter
 * need not synchronise its producer first. */
int chai_hip_batcher_submit(void* batcher, const chai_ragged_item* items, size_t count, void* stream);
/* run whatever is queued now (for a last, partial batch) */
int chai_hip_batcher_flush(void* batcher);
/* stats[0..4]: batches run, items run, plans built, graph replays, variant */
int chai_hip_batcher_stats(void* batcher, uint64_t* stats);
int chai_hip_batcher_destroy(void* batcher);

/* ------------------------------------------------------------- autotuning */
/* Journal memo (autotune/AutoTuneMemo.cpp:16-318): lines "<device> <kernel>
 * <state|microseconds>" appended as each step is entered; on open, kernels
 * left in COMPILE_KERNEL / START_TIMER / BEFORE_CHECK are blacklisted
 * (they crashed or hung).  The file name convention of the reference is
 * journal<DEVICE>.<VENDOR> (vendor/OCLhacks.cpp:483-491). */
int chai_hip_memo_open(const char* journal_path, int device);
int chai_hip_memo_close(void);
int chai_hip_memo_save(const char* path);
/* status: 1 good (timed), 0 unknown, -1 blacklisted; us = first bench time */
int chai_hip_memo_kernel_status(const char* kernel, int* status, long* us);
/* append one journal value (state code < 0, or a time in microseconds) */
int chai_hip_memo_record(const char* kernel, long value);
/* Time one tile variant (AutoTuneBench::runTime, autotune/AutoTuneBench.cpp:
 * 17-251): memo hit returns the journaled time; otherwise one warm-up and
 * `trials` timed runs of the core kernel (hipEvents), best time journaled.
 * core_ms = -1 for a blacklisted variant. */
int chai_hip_gemm_time_variant(const chai_gemm_desc* desc, const void* A, const void* B, void* C,
                               int variant, int trials, float* core_ms, void* stream);
/* Exhaustive search over the family's variants (StandardEM::searchLoop /
 * AutoTuneEM, autotune/StandardEM.cpp:69-121); the winner is recorded for
 * the shape's exogenous key (enqueue/JustInTimeMemo.cpp:92-114) and used by
 * chai_hip_gemm whenever desc->variant < 0. */
int chai_hip_gemm_tune(const chai_gemm_desc* desc, const void* A, const void* B, void* C, int trials,
                       int* best_variant, float* best_ms, void* stream);
/* record a variant for the shape's exogenous key (e.g. a model prediction) */
int chai_hip_gemm_set_variant(const chai_gemm_desc* desc, int variant);

/* ------------------------------------------------- learned tile-cost model */
/* A 3 x 64 ReLU MLP (FP32) that predicts the core-kernel time of a tile
 * variant of the BF16X3, BF16X6, FMA, F64 and GEMV families from
 * CHAI_GEMM_FEATURES features, trained on the tuner's own
 * timings with the library's HIP forward / backward / update kernels.  It has
 * no counterpart in the reference, whose only parameter choice is the timed
 * search (autotune/StandardEM.cpp:69-121); that search stays the fallback and
 * measures the model's regret.  X is [B][inputs] row-major f32, Y / out are
 * [B] f32, all device pointers.  The loss is mean_b (out_b - Y_b)^2. */
enum { CHAI_GEMM_FEATURES = 25 };
enum { CHAI_MLP_SGD = 0, CHAI_MLP_ADAM = 1 };
int chai_hip_mlp_create(int inputs, uint64_t seed, void** mlp);
int chai_hip_mlp_num_params(void* mlp, size_t* count);
/* parameter order: W1 [inputs][64], b1, W2 [64][64], b2, W3 [64][64], b3, w4 [64], b4 */
int chai_hip_mlp_get_params(void* mlp, float* host, size_t count);
int chai_hip_mlp_set_params(void* mlp, const float* host, size_t count);
int chai_hip_mlp_forward(void* mlp, const float* X, size_t B, float* out, void* stream);
/* grad (device, num_params floats) = d loss / d params */
int chai_hip_mlp_backward(void* mlp, const float* X, const float* Y, size_t B, float* grad, void* stream);
int chai_hip_mlp_update(void* mlp, const float* grad, int kind, float lr, void* stream);
/* backward + update without materialising the gradient */
int chai_hip_mlp_train_step(void* mlp, const float* X, const float* Y, size_t B, int kind, float lr,
                            void* stream);
int chai_hip_mlp_release(void* mlp);
/* The model input of (shape, tile variant) for the family desc->mode names
 * (AUTO resolves as in chai_hip_gemm): CHAI_GEMM_FEATURES floats. */
int chai_hip_gemm_features(const chai_gemm_desc* desc, int variant, float* out);
/* One inference over every variant of the shape; the lowest predicted time
 * wins and is recorded for the shape's exogenous key (as
 * chai_hip_gemm_set_variant).  predicted (nullable) receives each variant's
 * model output. */
int chai_hip_gemm_predict_variant(const chai_gemm_desc* desc, void* mlp, int* variant, float* predicted,
                                  void* stream);

/* ----------------------------------------------------------------- traces */
/* A trace is what one VectorTrace hands a device (common/VectorTrace.cpp):
 * variables (arrays) and statements "var = expression", each expression a
 * postfix (RPN) list of nodes whose codes are Chai's own bytecodes
 * (bytecode/ByteCodes.hpp:21-338) plus CHAI_TRACE_VAR for a variable read.
 * chai_hip_trace_compile groups statements into fused kernels the way the
 * reference JIT kernelizes (jitkrnl/WorkTrace.cpp:62-370,
 * jitkrnl/GroupTogether.cpp:178-201): elementwise statements over one index
 * space fuse into one launch of the trace VM kernel; a sum/mean/dot fuses
 * into the kernel producing its operand; matmul runs the GEMM/GEMV family;
 * temporaries used inside one fused kernel never touch memory; gather1_floor /
 * gather2_floor (bytecodes 40 / 41, node order data, x subscript[, y
 * subscript]) read a variable at computed subscripts with the interpreter's
 * wrap (interp/InterpGather.hpp:20-52), and a statement gathering from a
 * variable the current kernel computes starts a new kernel.  An OUTPUT or
 * TEMP variable may be assigned by several statements (Chai arrays are
 * mutable); each read sees the latest assignment before it, and reading such
 * a variable before its first assignment is refused.  Reductions and matmuls
 * below the root of a statement, matmul operands that are expressions and
 * gather data that is an expression are split into statements of their own
 * (jitstmt/MakeStmt.cpp:579-726).  The compiled handle keeps its own copy of
 * the trace: the caller may free it after compiling.  A trace the
 * compiler cannot place on the GPU returns CHAI_ERR_ARG, so the caller falls
 * back to the interpreter (runtime/Scheduler.cpp:221-254). */
enum { CHAI_TRACE_VAR = 0x10000 };
enum { CHAI_VAR_INPUT = 0, CHAI_VAR_OUTPUT = 1, CHAI_VAR_TEMP = 2 };

typedef struct chai_trace_var {
    uint32_t prec;                 /* CHAI_U32/I32/F32/F64                    */
    uint32_t role;                 /* CHAI_VAR_INPUT / OUTPUT / TEMP          */
    uint64_t W, H, slots;          /* dense [slots][H][W]                     */
} chai_trace_var;

typedef struct chai_trace_node {
    uint32_t code;                 /* bytecode or CHAI_TRACE_VAR              */
    uint32_t arg;                  /* VAR: variable index; rng: variant;
                                      index1/2: dimension (0 = x, 1 = y)      */
    uint32_t vec_len;              /* rng vector length (0 = default)        */
    uint32_t flags;                /* transpose of a matmul operand (bit 0)   */
    double s0, s1;                 /* scalar_* value; rng limits lo, hi;
                                      matmulG alpha, beta                     */
    uint64_t seed;                 /* rng seed                                */
    uint64_t W, H, slots;          /* dims of data-creating leaves            */
} chai_trace_node;

typedef struct chai_trace_stmt {
    uint32_t lhs;                  /* variable assigned                       */
    uint32_t first, count;         /* node range in postfix order            */
    uint32_t pad;
} chai_trace_stmt;

typedef struct chai_trace {
    const chai_trace_var* vars;   uint32_t num_vars;
    const chai_trace_stmt* stmts; uint32_t num_stmts;
    const chai_trace_node* nodes; uint32_t num_nodes;
} chai_trace;

/* Compile (validate, type, group, emit VM programs).  Replaces MakeStmt /
 * WorkTrace / Enqueue text emission and the per-trace JustInTimeMemo
 * (enqueue/JustInTimeMemo.cpp:31-334): compiled programs are cached for the
 * life of the process, keyed by the trace's full description (variables with
 * their shapes, statements, nodes) and the engine (VM or JIT); a repeated
 * trace shares the compiled program of the first. */
int chai_hip_trace_compile(const chai_trace* trace, void** handle);
/* Counters of that cache: compiles answered from it, compiles that did not
 * find their trace in it, programs held.  JustInTimeMemo's lookup-or-build
 * (enqueue/JustInTimeMemo.cpp:59-88) made observable. */
int chai_hip_trace_cache_stats(uint64_t* hits, uint64_t* misses, uint64_t* entries);
/* buffers[v] = device pointer of variable v (INPUT and OUTPUT required;
 * TEMP may be NULL, the library allocates one when it must be stored). */
int chai_hip_trace_launch(void* handle, void* const* buffers, void* stream);
/* number of kernels the trace compiled into, and of VM instructions (the
 * statement / kernel split the reference makes in WorkTrace::together,
 * jitkrnl/WorkTrace.cpp:62-370) */
int chai_hip_trace_info(void* handle, int* num_kernels, int* num_instructions);
int chai_hip_trace_release(void* handle);
/* Fused kernels run either on the pre-compiled trace VM kernel or as
 * straight-line HIP generated from the same VM program and compiled at run
 * time with hipRTC (default; env CHAI_TRACE_JIT=0 or set_jit(0) selects the
 * VM).  A failed JIT compile falls back to the VM. */
int chai_hip_trace_set_jit(int on);
/* Generated source (or the compile log when the JIT failed) of kernel k;
 * replaces the kernel text the reference prints to its journal
 * (AutoTuneMemo::printSource, autotune/AutoTuneMemo.cpp:248). */
int chai_hip_trace_kernel_source(void* handle, int kernel, char* buf, size_t len, int* jitted);
/* Front end only (no device): the JIT source kernel k of a trace would get
 * (the same text as chai_hip_trace_kernel_source, produced without a GPU). */
int chai_hip_trace_emit_source(const chai_trace* trace, int kernel, char* buf, size_t len);

/* Front end only (no device): fill in prec, W, H and slots of every
 * variable that a statement assigns and whose W is 0, from the statement's
 * expression, in statement order (the dimension rules of the reference's
 * API, api/ArrayFun.cpp:45-118, and interpreter, interp/InterpMatmulBase.cpp:
 * 79-118).  A trace lowered from a VectorTrace declares only its inputs; this
 * gives every statement's result its shape before the trace is compiled. */
int chai_hip_trace_infer(chai_trace_var* vars, uint32_t num_vars, const chai_trace_stmt* stmts,
                         uint32_t num_stmts, const chai_trace_node* nodes, uint32_t num_nodes);

/* Front end only (no device): the checks of chai_hip_trace_compile (parse,
 * type, split, group) without compiling kernels; CHAI_OK when the GPU back
 * end takes the trace, else CHAI_ERR_ARG and chai_hip_last_error() says why
 * (the answer the reference's JIT gives by failing, which sends the trace to
 * the interpreter, runtime/Scheduler.cpp:221-254). */
int chai_hip_trace_check(const chai_trace* trace);

/* ------------------------------------------------------- inline OpenCL C */
/* The ProgramCL hook (api/ProgramCL.cpp:133-376, api/chai.h:703-771): user
 * OpenCL C kernels launched on the back end's arrays.  Replaces the OpenCL
 * build and enqueue of source kernels (vendor/OCLdevice.cpp:940-968,
 * enqueue/EnqueueAST.cpp:432-489, memmgr/MemTransJIT.cpp:161-260).  The text
 * is compiled as OpenCL C 1.2 for gfx950 by ROCm's code-object manager;
 * argument kinds and sizes come from the code object's kernel metadata.
 * Argument kinds follow ProgramCL's ARRAY_BASE / LOCAL_MEM / SCALAR
 * (api/chai.h:708-713). */
enum { CHAI_CL_ARG_ARRAY = 0, CHAI_CL_ARG_LOCAL = 1, CHAI_CL_ARG_SCALAR = 2, CHAI_CL_ARG_OTHER = 3 };

typedef struct chai_cl_arg {
    uint32_t kind;                 /* CHAI_CL_ARG_ARRAY / LOCAL / SCALAR      */
    uint32_t prec;                 /* scalar precision (CHAI_U32 ... F64)     */
    union {
        void* ptr;                 /* ARRAY: device buffer                    */
        uint64_t count;            /* LOCAL: bytes of __local memory          */
        uint32_t u; int32_t i; float f; double d;   /* SCALAR                 */
    } v;
} chai_cl_arg;

typedef struct chai_cl_arg_info {
    uint32_t kind;                 /* CHAI_CL_ARG_*                           */
    uint32_t size;                 /* bytes in the kernel argument segment    */
    char type_name[64];            /* OpenCL type as written ("float*", "uint") */
    char name[64];                 /* parameter name                          */
} chai_cl_arg_info;

/* Compile (or find in the cache: same text and options) a program.  On a
 * compile error returns CHAI_ERR_ARG and chai_hip_last_error() holds the
 * compiler log (ProgramCL's build failure, vendor/OCLdevice.cpp:940-968).
 * Needs no device: compiling works without a GPU.  Handles stay valid until
 * the process ends. */
int chai_hip_cl_build(const char* text, size_t len, const char* options, void** program);
int chai_hip_cl_build_log(void* program, char* buf, size_t len);
/* The program's gfx950 executable code object: what the OpenCL build of
 * OCLkernel::buildCL (vendor/OCLdevice.cpp:940-968) produces and keeps inside
 * the OpenCL runtime.  *size receives its length; up to len bytes are copied
 * to buf when buf is not NULL.  Needs no device. */
int chai_hip_cl_code_object(void* program, void* buf, size_t len, size_t* size);
/* explicit arguments of kernel `name` (ProgramCL::parseKernelArgs,
 * api/ProgramCL.cpp:37-131, from metadata instead of the text) */
int chai_hip_cl_kernel_args(void* program, const char* name, int* num_args, chai_cl_arg_info* info, int max_info);
/* Launch kernel `name` over a 1- or 2-D work space (ProgramCL::Kernel's
 * operator(), api/ProgramCL.cpp:327-342).  global sizes count work items, as
 * in OpenCL, and must be multiples of the local sizes (local NULL or 0: 64
 * when it divides, else 1).  Arguments are checked against the kernel's
 * metadata (kind, and size for scalars); a mismatch returns CHAI_ERR_ARG
 * before anything is launched.  __local arguments get their byte count, as
 * clSetKernelArg(k, i, bytes, NULL) does. */
int chai_hip_cl_launch(void* program, const char* name, const chai_cl_arg* args, uint32_t nargs, uint32_t dims,
                       const uint64_t* global, const uint64_t* local, void* stream);
/* A kernel handle (ProgramCL::Kernel, api/ProgramCL.cpp:185-200), valid as
 * long as its program, for inline-kernel statements of traces. */
int chai_hip_cl_kernel(void* program, const char* name, void** kernel);
/* The checks of chai_hip_cl_launch without a launch (no device needed;
 * array pointers may be NULL).  Compiles the __local wrapper if needed. */
int chai_hip_cl_check(void* kernel, const chai_cl_arg* args, uint32_t nargs, uint32_t dims,
                      const uint64_t* global, const uint64_t* local);

/* Inline kernels inside traces: the statement the reference records for
 * (progCL, "name", args...)(global, local) (api/ProgramCL.cpp:202-219,
 * bytecode kernel_from_opencl = 50).  Nodes, in argument order: a
 * CHAI_TRACE_VAR node per array argument, a scalar_u32/i32/f32/f64 node
 * (bytecodes 5-8, value in s0) per scalar argument, a CHAI_TRACE_CL_LOCAL
 * node (W = bytes) per __local argument; then the root node: code 50,
 * seed = kernel handle from chai_hip_cl_kernel, arg = argument count,
 * vec_len = work dimensions (1 or 2), W/H = global sizes, s0/s1 = local
 * sizes.  The statement's lhs is its first array argument.  The kernel
 * reads and writes its array arguments in place (inputs included, as the
 * reference's kernels write the arrays they are given) and runs as a kernel
 * of its own between the fused kernels before and after it; it does not
 * count as an assignment of lhs. */
enum { CHAI_TRACE_CL_LOCAL = 0x10001 };

/* ------------------------------------------------------------------- GEMV */
/* y = op(A) x or y = alpha op(A) x + beta y (Evergreen MV,
 * autotune/EvergreenMatmulMV.cpp:884-1403): A is M x N row-major (op = A^T
 * when trans_a, then y has N entries).  HBM-streaming kernel, one wavefront
 * per row block with cross-lane reductions; accumulation in FP32 per lane and
 * FP64 across lanes. */
int chai_hip_gemv(int prec_a, int prec_x, int prec_y, int trans_a, size_t M, size_t N,
                  const void* A, size_t lda, const void* x, void* y, int general,
                  double alpha, double beta, size_t batch, int same_a,
                  size_t stride_a, size_t stride_x, size_t stride_y, void* stream);

/* Per-kernel timing (the replacement for AutoTuneBench's timer hooks,
 * autotune/AutoTuneBench.cpp:116-123): when enabled, chai_hip_gemm records
 * hipEvents on its launch stream around the operand pre-pass (split / pack)
 * and around the core GEMM kernel; chai_hip_gemm_last_timing waits for the
 * calling thread's most recent profiled call and returns both durations.
 * chai_hip_gemm_profile_mean returns their means over the thread's last
 * `count` profiled calls (at most 64 are kept; enabling starts over), so a
 * run of back-to-back calls is timed without host waits in between. */
int chai_hip_profile_enable(int on);
int chai_hip_gemm_last_timing(float* prep_ms, float* core_ms);
int chai_hip_gemm_profile_mean(int count, float* prep_ms, float* core_ms);

#ifdef __cplusplus
}
#endif
#endif /* CHAI_MI355X_H */
