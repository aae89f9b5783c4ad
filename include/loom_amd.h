/* loom_amd.h — C ABI of the MI355X (gfx950) execution backend for Loom
 * kernel_blank IR.
 *
 * The reference exposes its backends as C++ class templates (no FFI):
 *   simd::compact_function<K>  (src/vm/simd.hpp:4329-4335, 3837-4008)
 *   simd::hybrid_function<K>   (src/vm/simd.hpp:4310-4317, 4255-4279)
 *   simd::hybrid_stream<K>     (src/vm/simd.hpp:4319-4327, 4281-4291,
 *                               facade 2797-2903)
 *   mp::Shop                   (src/vm/mp.hpp:316-385)
 * The C++ facade in include/loom_vm/vm/hip.hpp keeps those call shapes and
 * reaches the GPU only through the entry points below.  Kernels are compiled
 * ahead of time into the library (or a module linked with it) and are
 * addressed by index; nothing of C++ crosses this boundary.
 *
 * All entry points return 0 on success and a nonzero status otherwise (a HIP
 * error code, or LOOM_E_* below); loom_last_error() describes the last
 * failure of the calling thread.
 */
#ifndef LOOM_AMD_H
#define LOOM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOOM_MAX_IO 32
#define LOOM_MAX_BLOCKS 64
#define LOOM_NUM_COUNTERS 16

enum loom_status
{
    LOOM_OK          = 0,
    LOOM_E_ARG       = 1000, /* invalid argument */
    LOOM_E_KERNEL    = 1001, /* unknown kernel index or name */
    LOOM_E_NODEVICE  = 1002, /* no GPU or library not initialised */
    LOOM_E_RUNTIME   = 1003, /* a kernel raised an error flag */
    LOOM_E_CAPACITY  = 1004, /* a size exceeds a configured capacity */
    LOOM_E_STATE     = 1005  /* call not valid in the object's state */
};

/* scalar type codes of kernel inputs/outputs */
enum loom_type
{
    LOOM_U8 = 0, LOOM_U16, LOOM_U32, LOOM_U64,
    LOOM_I8, LOOM_I16, LOOM_I32, LOOM_I64,
    LOOM_F32, LOOM_F64
};

enum loom_mode
{
    LOOM_MODE_COMPACT = 0, /* predicated wave64, per-wave reconvergence */
    LOOM_MODE_HYBRID  = 1  /* persistent, in-wave refill + label-queue regroup */
};

/* counters reported by a run (indices into loom_run_stats.counters) */
enum loom_counter
{
    LOOM_C_ACTIVE_OPS = 0, /* Σ popcount(EXEC)·ops(B) */
    LOOM_C_SLOT_OPS,       /* Σ 64·ops(B): active-lane % = ACTIVE/SLOT */
    LOOM_C_DISPATCHES,     /* wave block dispatches */
    LOOM_C_THREAD_STEPS,   /* thread-block-steps */
    LOOM_C_SPILL_ROWS,     /* threads spilled to regroup queues */
    LOOM_C_SPILL_BYTES,
    LOOM_C_FILL_ROWS,      /* threads filled from regroup queues */
    LOOM_C_FILL_BYTES,
    LOOM_C_ADMITTED,       /* threads admitted from the input */
    LOOM_C_REFILLS,        /* in-wave refill events */
    LOOM_C_BATCHES,        /* batches claimed */
    LOOM_C_SPILL_FAILED,   /* lanes kept because a queue was full */
    LOOM_C_ACTIVE_FLOPS    /* Σ popcount(EXEC)·flops(B), fma = 2 */
};

/* error flag bits of loom_run_stats.errors */
#define LOOM_ERR_ASSERT_FP   1u /* ops::assert_fp saw NaN or Inf */
#define LOOM_ERR_STEP_BUDGET 2u /* a wave exceeded max_steps */

typedef struct loom_kernel_info
{
    char     name[64];
    uint64_t hash;        /* structural hash of the lowered kernel */
    int32_t  n_inputs, n_outputs;
    int32_t  input_type[LOOM_MAX_IO], output_type[LOOM_MAX_IO];
    int32_t  n_blocks, n_registers, index_bytes;
    uint32_t state_size;      /* Σ register widths incl. PC (core.hpp:1027) */
    uint32_t queue_row_bytes; /* Σ over labels of 4 + live-in bytes */
    uint16_t block_ops[LOOM_MAX_BLOCKS];        /* #Compute + 1 */
    uint32_t block_live_bytes[LOOM_MAX_BLOCKS]; /* 4 + live-in bytes */
    uint16_t block_flops[LOOM_MAX_BLOCKS];      /* floating-point ops */
} loom_kernel_info;

typedef struct loom_params
{
    uint32_t threshold;      /* block-population threshold in lanes (0..63):
                                a wave stops when its best block has no more
                                (reference UTR, simd.hpp:3270-3281) */
    uint32_t live_threshold; /* hybrid: stop when live lanes <= this (0..63) */
    uint32_t refill;         /* hybrid: admit new threads into free lanes */
    uint32_t regroup;        /* hybrid: LOOM_REGROUP_NONE, LOOM_REGROUP_QUEUES
                                (spill survivors to per-label queues for the
                                next pass) or LOOM_REGROUP_WORKGROUP (sort the
                                threads of a workgroup by block through LDS
                                between run phases; reconv 0 or 2) */
    uint32_t reconv;         /* reconvergence policy: 0 = successor-major
                                (reference simd.hpp:3974-4005), 1 = topological
                                (loop heads wait for their bodies), 2 = sweep
                                (every block in topological order under its
                                mask, empty blocks skipped) */
    uint32_t waves_per_cu;   /* hybrid persistent grid; 0 = occupancy limit */
    uint32_t collect_stats;  /* fill loom_run_stats.counters */
    uint64_t queue_capacity; /* hybrid rows per label queue; 0 = automatic */
    uint64_t max_steps;      /* per-wave block dispatch budget; 0 = 2^40.
                                Checked after every dispatch (reconv 0, 1)
                                or after every sweep (reconv 2) */
    uint32_t register_file;  /* compact: LOOM_REGFILE_VGPR (registers in
                                VGPRs for the thread's life) or
                                LOOM_REGFILE_MEMORY (one column per register
                                in device memory, the reference's SoA
                                register file, bve.hpp:494-693; each block
                                loads what it reads and stores what stays
                                live, so VGPRs hold one block's working set;
                                sweep policy only) */
    uint32_t regroup_phase;  /* LOOM_REGROUP_WORKGROUP: dispatches of a run
                                phase between two sorts (0: 64) */
    uint32_t regroup_waves;  /* LOOM_REGROUP_WORKGROUP: waves per workgroup,
                                16 (one workgroup per CU) or 8 (two per CU;
                                one sorts while the other runs); 0: 16 */
} loom_params;
#define LOOM_REGFILE_VGPR 0
#define LOOM_REGFILE_MEMORY 1
#define LOOM_REGROUP_NONE 0
#define LOOM_REGROUP_QUEUES 1
#define LOOM_REGROUP_WORKGROUP 2
/* A NULL loom_params* means: threshold 0, refill on, regroup off, no stats,
 * successor-major for compact, sweep with live_threshold 48 for hybrid and
 * batched launches. */

typedef struct loom_run_stats
{
    double   ms_total;      /* whole run on the stream (HIP events) */
    double   ms_main;       /* Σ duration of the main kernel launches */
    uint32_t main_launches; /* launches of the main kernel */
    uint32_t passes;        /* hybrid passes */
    uint32_t errors;        /* LOOM_ERR_* bits */
    uint32_t grid_waves;    /* waves of the (persistent) grid */
    uint64_t counters[LOOM_NUM_COUNTERS];
} loom_run_stats;

/* ---- device and registry ------------------------------------------------ */
/* Replaces nothing in the reference (CPU backends need no device); mirrors
 * the CUDA test backend's implicit device selection (src/vm/cuda.hpp:34-116). */
int         loom_init(int device);
const char* loom_last_error(void);
int         loom_device_info(int* compute_units, int* wave_size,
                             size_t* global_mem, char* name, size_t name_len);
/* Kernel lookup: the reference binds kernels at compile time through the
 * template parameter K (simd.hpp:3742-3748); here by name or index. */
int loom_kernel_count(void);
int loom_kernel_info_get(int index, loom_kernel_info* out);
int loom_kernel_find(const char* name);
/* Text listing of the lowered kernel (blocks, cost, live-in registers,
 * successors, instructions): the reference IR inspector's block view
 * (src/vm/dbg.hpp:1183-1202).  Writes at most len bytes including the
 * terminating NUL; *needed receives the full size.  Needs no device. */
int loom_kernel_dump(int kernel, char* buf, size_t len, size_t* needed);
/* VGPRs and scratch bytes per lane of a compiled kernel (mode: compact or
 * hybrid pass kernel; policy: loom_params.reconv, or'ed with
 * LOOM_RESOURCES_REGROUP for the hybrid pass kernel that regroups, which a
 * run with loom_params.regroup set launches; runs without it launch a pass
 * kernel compiled without the queue paths; or'ed with LOOM_RESOURCES_MEMORY
 * for the compact kernel whose register file is in device memory). */
#define LOOM_RESOURCES_REGROUP 0x100
#define LOOM_RESOURCES_MEMORY 0x200
#define LOOM_RESOURCES_WORKGROUP 0x400 /* the workgroup-regrouping kernel (reconv 0 or 2) */
#define LOOM_RESOURCES_WORKGROUP_8 0x800 /* the same with 8-wave workgroups */
int loom_kernel_resources(int kernel, int mode, int policy, int* vgprs, int* scratch_bytes);
/* Register r (0 <= r < n_registers) of a kernel: its name (NUL-terminated,
 * at most len bytes), first type (loom_type), width in bytes, number of
 * types (> 1: a union register, one name used with several same-size types,
 * src/vm/core.hpp:993-1003) and LOOM_REG_* flags — the register list of the
 * reference inspector's info view (src/vm/dbg.hpp:1183-1200).  Any output
 * pointer may be NULL.  Needs no device. */
#define LOOM_REG_INPUT   1 /* a kernel input */
#define LOOM_REG_OUTPUT  2 /* a kernel output */
#define LOOM_REG_WRITTEN 4 /* the destination of some instruction */
int loom_kernel_register(int kernel, int r, char* name, size_t len, int* type, int* width,
                         int* ntypes, int* flags);
/* Per-thread step trace, the GPU counterpart of the reference's
 * step-by-step inspector (src/vm/dbg.hpp:1011-1140; one row per block as
 * print_state, 920-935).  Threads [0, n) of the device input columns run one
 * block at a time, at most max_steps blocks each, with the same lowered block
 * code as the compact and hybrid forms.  After every block one row of
 * LOOM_TRACE_WORDS(n_registers) 64-bit words is written to `rows` (device):
 *   word 0      block run (low 32 bits) | next block (high 32 bits,
 *               LOOM_TRACE_HALT when the thread halted)
 *   word 1      error bits so far (LOOM_ERR_ASSERT_FP)
 *   word 2 + r  bits of register r, zero-extended; 0 before its first write.
 * Thread t's rows start at row t * max_steps; steps[t] (device) receives the
 * number of rows written for t (fewer than max_steps only if t halted).
 * rows_bytes must hold n * max_steps rows; n < 2^20, 1 <= max_steps <= 2^16.
 * Returns when the rows are written. */
#define LOOM_TRACE_WORDS(n_registers) (2 + (n_registers))
#define LOOM_TRACE_HALT 0xFFFFu
int loom_kernel_trace(int kernel, const void* const* inputs, uint64_t n, uint32_t max_steps,
                      uint64_t* rows, size_t rows_bytes, uint32_t* steps);

/* ---- device memory (the reference owns SoA columns in bve::Memory,
 * src/vm/bve.hpp:494-693; the CUDA backend in DeviceAllocator,
 * src/vm/cuda.hpp:86-116) --------------------------------------------------- */
int loom_malloc(void** ptr, size_t bytes);
int loom_free(void* ptr);
int loom_memcpy_h2d(void* dst, const void* src, size_t bytes);
int loom_memcpy_d2h(void* dst, const void* src, size_t bytes);
int loom_memcpy_d2d(void* dst, const void* src, size_t bytes);
int loom_memset(void* dst, int value, size_t bytes);
int loom_synchronize(void);
/* Fill `count` elements of `elem_bytes` (1, 2, 4 or 8) bytes at `dst` with
 * the element at `pattern`: the broadcast of a scalar argument, which the
 * reference applies when a function argument is a scalar instead of a range
 * (src/vm/bve.hpp:1125-1133).  A kernel, not a memset. */
int loom_fill(void* dst, const void* pattern, size_t elem_bytes, uint64_t count);
/* Background copies on a second stream (default priority) that overlap
 * kernels on the main stream; copies complete in the order they were
 * issued.  Each call returns a ticket: loom_copy_status reports whether the
 * copy (and every earlier one) completed, loom_copy_wait blocks until it
 * has.  The caller keeps both buffers alive until then; host buffers should
 * be pinned (loom_host_alloc) for the copy to run asynchronously.  Copies are
 * not ordered after main-stream work still running: runs return once their
 * kernels finished, so copy a run's outputs after it returned, and start a
 * run on copied inputs after their ticket completed.  Replaces the CUDA test
 * backend's staging copies (src/vm/cuda.hpp:1468-1578). */
enum loom_copy_kind
{
    LOOM_COPY_H2D = 0,
    LOOM_COPY_D2H = 1
};
int loom_copy_async(void* dst, const void* src, size_t bytes, int kind, uint64_t* ticket);
int loom_copy_status(uint64_t ticket, int* done);
int loom_copy_wait(uint64_t ticket);
/* Pinned host memory for staging (the reference CUDA backend stages through
 * pinned buffers, src/vm/cuda.hpp:1041-1055). */
int loom_host_alloc(void** ptr, size_t bytes);
int loom_host_free(void* ptr);

/* ---- function execution --------------------------------------------------
 * Replaces compact_function<K>::operator() (simd.hpp:3837-4008) and
 * hybrid_function<K>::operator() (simd.hpp:4262-4278): runs every thread of
 * the inputs to HALT.  `inputs`/`outputs` hold one device column per kernel
 * input/output (structure of arrays, element type per loom_kernel_info).
 * Outputs are written in input order (the reference hybrid returns them in
 * its sorted order; input order is the stricter contract). */
int loom_run(int kernel, int mode, const void* const* inputs,
             void* const* outputs, uint64_t n, const loom_params* params,
             loom_run_stats* stats);
/* loom_run without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The hybrid form's
 * control block and regroup queues live in the library's static device
 * scratch array of 16 MiB (one per GPU, held until the process ends); a
 * regrouping run whose queues (2 x queue capacity x queue_row_bytes) do not
 * fit is rejected with LOOM_E_CAPACITY.  Passes of a regrouping run are timed
 * only as a whole (ms_main = ms_total).  Every call on one GPU shares the
 * scratch array: calls that overlap on one GPU (from several threads) race
 * and must be serialised by the caller.  Inputs, outputs and other
 * arguments as loom_run, on the default context. */
int loom_run_noalloc(int kernel, int mode, const void* const* inputs, void* const* outputs,
                     uint64_t n, const loom_params* params, loom_run_stats* stats);
/* The same with host columns (copies in and out; device buffers cached). */
int loom_run_host(int kernel, int mode, const void* const* inputs,
                  void* const* outputs, uint64_t n, const loom_params* params,
                  loom_run_stats* stats);

/* ---- execution contexts ------------------------------------------------------
 * Several host threads feeding one GPU: the reference runs many worker
 * threads, each stepping its own streams of every kernel, while the main
 * thread exchanges inputs and outputs (src/vm/mp.hpp:61-131 Context, 158-190
 * Worker, 316-385 Shop).  A context is a host thread's own lane to the
 * device: a kernel stream and a copy stream (default priority), its error
 * flags, counters and workspaces.  Calls on different contexts run
 * concurrently and their kernels overlap on the device; calls on one
 * context serialise on its lock.  The entry points without a context
 * argument use the default context made by loom_init, and are safe to call
 * from several threads (they serialise). */
typedef struct loom_context loom_context;
int loom_context_create(loom_context** ctx);
int loom_context_destroy(loom_context* ctx);
/* loom_run, loom_run_host, loom_batch_run, loom_fill, loom_copy_*,
 * loom_synchronize on context ctx (NULL: the default context). */
int loom_ctx_run(loom_context* ctx, int kernel, int mode, const void* const* inputs,
                 void* const* outputs, uint64_t n, const loom_params* params,
                 loom_run_stats* stats);
int loom_ctx_run_host(loom_context* ctx, int kernel, int mode, const void* const* inputs,
                      void* const* outputs, uint64_t n, const loom_params* params,
                      loom_run_stats* stats);
int loom_ctx_fill(loom_context* ctx, void* dst, const void* pattern, size_t elem_bytes,
                  uint64_t count);
struct loom_job;
int loom_ctx_batch_run(loom_context* ctx, int set, const struct loom_job* jobs, int njobs,
                       const loom_params* params, loom_run_stats* stats);
int loom_ctx_copy_async(loom_context* ctx, void* dst, const void* src, size_t bytes, int kind,
                        uint64_t* ticket);
int loom_ctx_copy_status(loom_context* ctx, uint64_t ticket, int* done);
int loom_ctx_copy_wait(loom_context* ctx, uint64_t ticket);
int loom_ctx_synchronize(loom_context* ctx);

/* ---- device-resident stream ----------------------------------------------------
 * Replaces hybrid_stream<K> / the concurrent stream facade
 * (src/vm/simd.hpp:4319-4327, 2797-2903; bve::MemoryFacade and
 * exchange_from, src/vm/bve.hpp:1028-1229).  `slots` thread slots whose
 * registers stay in device memory (structure of arrays, one column per
 * register) between steps:
 *   set_inputs  starts threads 0..m-1 (m = min(n, slots)) in the slots and
 *               empties the rest;
 *   step        one launch: every live slot runs under loom_params.reconv
 *               until at most live_threshold lanes of its wave are live
 *               (0: to HALT) or max_steps dispatches were made; threads keep
 *               their registers between steps;
 *   sort        reports whether an exchange is available: at least 64 slots
 *               halted since the last exchange, or some halted and none live;
 *   exchange    for m = min(n, halted) finished threads, in the order they
 *               halted: their outputs (and sequence numbers, if ids is not
 *               NULL) go to the host columns, and the next m inputs start
 *               in their slots; only those m rows cross PCIe;
 *   drain       the outputs of up to cap finished threads, whose slots become
 *               empty (the reference's get_output_zipped at the end);
 *   state       live threads (done: live == 0) and finished threads waiting.
 * Threads are numbered in the order they were given: set_inputs numbers
 * them 0.., every exchange continues the numbering.  The reference returns
 * a stream's outputs unordered; the sequence numbers let a caller restore
 * input order.  Each stream owns a HIP stream; calls on one stream
 * serialise on its lock. */
typedef struct loom_stream loom_stream;
typedef struct loom_stream_counters
{
    uint64_t steps, exchanges;
    uint64_t rows_in, rows_out;    /* threads started / outputs returned */
    uint64_t bytes_h2d, bytes_d2h; /* bytes copied in / out by this stream */
} loom_stream_counters;
int loom_stream_create(int kernel, uint64_t slots, const loom_params* params,
                       loom_stream** stream);
int loom_stream_destroy(loom_stream* stream);
int loom_stream_set_inputs(loom_stream* stream, const void* const* inputs, uint64_t n,
                           uint64_t* taken);
int loom_stream_step(loom_stream* stream, loom_run_stats* stats);
int loom_stream_sort(loom_stream* stream, int* exchange_available);
int loom_stream_exchange(loom_stream* stream, const void* const* inputs, uint64_t n,
                         void* const* outputs, uint64_t* ids, uint64_t* exchanged);
int loom_stream_drain(loom_stream* stream, void* const* outputs, uint64_t* ids, uint64_t cap,
                      uint64_t* drained);
int loom_stream_state(loom_stream* stream, uint64_t* live, uint64_t* halted);
int loom_stream_counters_get(loom_stream* stream, loom_stream_counters* counters);

/* ---- batched launches ------------------------------------------------------
 * Replaces the multi-kernel Shop (src/vm/mp.hpp:316-385: bind_inputs, the
 * workers' round-robin over per-kernel streams, finalize): the jobs of one
 * call run in ONE persistent launch; every job's kernel must belong to the
 * compiled batch set.  Jobs read/write device columns in [begin, end). */
#define LOOM_MAX_JOBS 64

typedef struct loom_job
{
    int32_t     kernel; /* global kernel index, a member of the batch set */
    int32_t     pad;
    uint64_t    begin, end;
    const void* inputs[LOOM_MAX_IO];
    void*       outputs[LOOM_MAX_IO];
} loom_job;

int loom_batch_set_count(void);
int loom_batch_set_find(const char* name);
/* global kernel indices of the members; returns their count (< 0: error) */
int loom_batch_set_members(int set, int* kernels, int max);
int loom_batch_run(int set, const loom_job* jobs, int njobs,
                   const loom_params* params, loom_run_stats* stats);

/* A batch captured once into a HIP graph and replayed (no counterpart in
 * the CPU reference).  The jobs are fixed at capture time. */
typedef struct loom_graph loom_graph;
int loom_batch_graph_create(int set, const loom_job* jobs, int njobs,
                            const loom_params* params, loom_graph** graph);
int loom_graph_launch(loom_graph* graph, loom_run_stats* stats);
int loom_graph_destroy(loom_graph* graph);

/* ---- serving -----------------------------------------------------------------
 * Continuous batching of kernel launches (requests).  No counterpart in the
 * reference, whose nearest relative is the Shop's exchange loop that admits
 * new inputs into a stream as old ones halt (src/vm/mp.hpp:61-131, 316-385).
 * A request is admitted FIFO when its device columns fit the HBM budget and
 * a job slot is free; every step is ONE batch launch that advances all
 * resident requests of the batch set's kernels by a slice of their threads
 * (step_threads shared among them), so requests that arrive while others
 * run join the next step and share the device with them.  A request of any
 * other registered kernel runs its slice of the step as a hybrid launch of
 * its own, after the batch launch, on the same stream.  A request completes
 * at the end of the step that ran its last slice; its outputs are copied
 * back and its device memory is released. */
typedef struct loom_server loom_server;

typedef struct loom_server_config
{
    int32_t     batch_set;
    uint32_t    max_resident; /* requests resident at once, 1..LOOM_MAX_JOBS */
    uint64_t    hbm_budget;   /* bytes of device columns of resident requests */
    uint64_t    step_threads; /* threads advanced per step, shared by residents */
    loom_params params;       /* parameters of the batch launches and of the
                                 launches of kernels outside the set */
} loom_server_config;

typedef struct loom_request
{
    int32_t     kernel;                /* global kernel index (any registered kernel) */
    uint32_t    broadcast;             /* bit i: input i is scalars[i], not a column */
    uint64_t    n;                     /* threads */
    const void* inputs[LOOM_MAX_IO];   /* host columns of n elements */
    void*       outputs[LOOM_MAX_IO];  /* host columns of n elements */
    uint64_t    scalars[LOOM_MAX_IO];  /* broadcast values (low bytes used) */
    double      arrival_ms;            /* arrival, relative to the run's start */
    /* set by the server */
    double      admit_ms;              /* admission (inputs resident) */
    double      done_ms;               /* completion (outputs on the host) */
    uint32_t    steps;                 /* steps the request took part in */
    int32_t     status;                /* 0, or the error code of the request */
    int32_t     complete;              /* set (release) when done_ms and status are final */
    int32_t     reserved;
} loom_request;

typedef struct loom_server_stats
{
    double   makespan_ms;  /* run start .. last completion */
    double   gpu_ms;       /* sum of the batch launches' durations */
    uint64_t steps;        /* batch launches */
    uint64_t threads;      /* threads of completed requests */
    uint32_t peak_resident;
    uint32_t pad;
    uint64_t peak_bytes;   /* most device bytes held by resident requests */
} loom_server_stats;

int loom_server_create(const loom_server_config* config, loom_server** server);
/* Serve `count` requests sorted by arrival_ms.  realtime != 0 replays the
 * arrival times against the wall clock (idle until the next arrival);
 * realtime == 0 treats every request as arrived at the start. */
int loom_server_run(loom_server* server, loom_request* requests, uint64_t count,
                    int realtime, loom_server_stats* stats);
/* Online serving: the loop runs on a host thread of its own and requests
 * are submitted from any host thread while it runs (the reference Shop's
 * main thread exchanging while the workers compute, src/vm/mp.hpp:349-369).
 * A submitted request's arrival is the time of its submission; the caller
 * keeps the request and its host columns alive until it is complete
 * (loom_server_poll / loom_server_wait; wait returns its status).
 * loom_server_stop stops accepting, completes every submitted request and
 * joins the loop.  loom_server_run and the online loop are exclusive. */
int loom_server_start(loom_server* server);
int loom_server_submit(loom_server* server, loom_request* request);
int loom_server_poll(loom_server* server, const loom_request* request, int* done);
int loom_server_wait(loom_server* server, const loom_request* request);
int loom_server_stop(loom_server* server, loom_server_stats* stats);
int loom_server_destroy(loom_server* server);

#ifdef __cplusplus
}
#endif

#endif /* LOOM_AMD_H */
