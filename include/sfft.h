/*
 * sfft.h -- C ABI of the MI355X transform engine (libsfft.so).
 *
 * The reference library is the header-only C++ template signalsmith-fft.h;
 * it has no C ABI.  These entry points are what a binding for its transform
 * path needs; each one names the reference interface it stands in for.
 * The C++ drop-in facade include/signalsmith-fft.h is built on top of them.
 *
 * Conventions
 *  - Complex data is interleaved (re, im): 8 bytes per value in F32, 16 in F64.
 *  - A batch is `batch` transforms stored back to back (stride n complex
 *    values for C2C, n reals / n/2 complex values for the real kinds).
 *  - All data pointers passed to sfft_execute* are DEVICE pointers (HBM).
 *  - Transforms are out of place, or exactly in place (in == out with the
 *    same byte layout) for plans whose sfft_plan_info.in_place is 1 (one-pass
 *    and chirp-z plans); any other overlap of input and output, and in place
 *    with a two-pass plan, returns SFFT_ERR_ALIASING (the reference silently
 *    corrupts in place).  An out-of-place call never writes `in`.
 *  - The inverse is unnormalised: ifft(fft(x)) = n * x (reference
 *    signalsmith-fft.h:295-315, no 1/N scaling).
 *  - Streams are hipStream_t passed as void*; NULL is the default stream.
 *  - Every function returns SFFT_OK (0) or a negative error code;
 *    sfft_last_error() gives a message for the calling thread.
 *  - A plan is immutable after creation and may be shared, but executions
 *    of one plan on several streams at once are allowed only for plans whose
 *    sfft_plan_info.workspace_bytes is 0 (every plan this library builds).
 */
#ifndef SFFT_H
#define SFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFFT_VERSION_MAJOR 0
#define SFFT_VERSION_MINOR 1

/* status codes */
#define SFFT_OK 0
#define SFFT_ERR_INVALID_ARG (-1)
#define SFFT_ERR_UNSUPPORTED_SIZE (-2)
#define SFFT_ERR_OUT_OF_MEMORY (-3)
#define SFFT_ERR_DEVICE (-4)
#define SFFT_ERR_ALIASING (-5)
#define SFFT_ERR_NOT_READY (-6)
#define SFFT_ERR_BUDGET (-7)

/* transform kinds */
#define SFFT_C2C 0           /* FFT<V>                       (signalsmith-fft.h:69-387)  */
#define SFFT_R2C 1           /* RealFFT<V>                   (signalsmith-fft.h:393-503) */
#define SFFT_R2C_HALFSHIFT 2 /* RealFFT<V, halfFreqShift> = ModifiedRealFFT<V> (:389-391, :505-508) */

/* precisions (the arithmetic type of the transform) */
#define SFFT_F32 0 /* FFT<float>  */
#define SFFT_F64 1 /* FFT<double> */
/* 16-bit HBM storage, float arithmetic, complex-to-complex only (no
 * counterpart in the reference): complex values are interleaved IEEE half /
 * bfloat16 pairs (4 bytes).  With the plan hint mfma = 1 the radix-16 stages
 * run on the matrix cores with split hi+lo operands and float accumulation;
 * by default they run as float butterflies (faster on MI355X). */
#define SFFT_F16 2
#define SFFT_BF16 3
/* FP32 HBM storage (complex64: the same buffers as FFT<float>) with FP64
 * twiddles, LDS and arithmetic, complex-to-complex only.  The reference's
 * FFT<float> computes its twiddle phases and products in float
 * (signalsmith-fft.h:123-124, :198-200, :217-286); here every twiddle
 * product and butterfly is formed in FP64, so the only FP32 roundings are
 * the stored output (and the stored intermediate of two-pass lengths):
 * relative L2 error about 4e-8 against FFT<double>, where F32 sits at
 * 1.2e-7 to 1.9e-7.  Costs FP64 registers and LDS (a smaller one-pass range). */
#define SFFT_F32X 4

/* directions */
#define SFFT_FORWARD 0 /* fft()  */
#define SFFT_INVERSE 1 /* ifft() */

typedef struct sfft_plan_s *sfft_plan;

typedef struct {
    size_t n;               /* transform length (reals for the real kinds) */
    int kind, precision;
    int passes;             /* HBM passes per transform (1 = LDS resident) */
    int nradices;           /* radices of the first pass, in stage order */
    int radices[16];
    int threads_per_transform, transforms_per_block;
    size_t workspace_bytes; /* device workspace per transform of an in-place call
                               (chirp-z lengths: every call); 0: none */
    size_t table_bytes;     /* device twiddle tables held by the plan */
    int mfma_stages;        /* radix-16 stages run as MFMA contractions (16-bit storage) */
    size_t chirp_length;    /* lengths with prime factors > 13: the plan is a chirp-z
                               convolution through FFTs of this length (0: direct) */
    int in_place;           /* 1: in == out (same layout) is allowed (every plan:
                               two-pass lengths run it through the workspace) */
} sfft_plan_info;

/* ---- library ---- */

/* Initialise the HIP device (optional; the first plan does it). */
int sfft_init(void);
/* Message for the last error on this thread ("" if none). */
const char *sfft_last_error(void);
/* Device properties the planner uses: CU count, LDS bytes per block. */
int sfft_device_info(int *cu_count, size_t *lds_bytes, char *arch, size_t arch_len);

/* ---- size helpers: integer-exact restatements of the reference ---- */

/* FFT<V>::sizeMinimum / sizeMaximum (signalsmith-fft.h:327-348): fast size
 * 2^a * {1,3,9} >= n / <= n. */
size_t sfft_size_minimum(size_t n);
size_t sfft_size_maximum(size_t n);
/* RealFFT<V>::sizeMinimum / sizeMaximum (signalsmith-fft.h:403-408),
 * including the reference's quirks (sizeMaximum may exceed n). */
size_t sfft_real_size_minimum(size_t n);
size_t sfft_real_size_maximum(size_t n);
/* 1 if the engine plans this length for the kind in F32, else 0.  Every
 * length up to 2^27 is planned, as the reference accepts every length
 * (signalsmith-fft.h:187-215 handles any factor): lengths whose factors are
 * all <= 13 run directly (up to 2^28), the others as a chirp-z convolution.
 * F64 plans reach 2^26 directly and 2^25 by chirp-z (two-pass column tiles
 * of 16-byte values must fit one CU's LDS); beyond, sfft_plan_create
 * returns SFFT_ERR_UNSUPPORTED_SIZE. */
int sfft_size_supported(size_t n, int kind);

/* ---- plans ---- */

/* FFT<V>::FFT(size) / setSize(size) and RealFFT<V,flags>::setSize
 * (signalsmith-fft.h:350-363, :410-435): factorise, choose radices, build the
 * twiddle tables on the device in FP64 and round them once.  Plans are
 * cached by (n, kind, precision); creating the same plan again is cheap. */
int sfft_plan_create(sfft_plan *plan, size_t n, int kind, int precision);
int sfft_plan_destroy(sfft_plan plan);
int sfft_plan_get_info(sfft_plan plan, sfft_plan_info *info);

/* Planner overrides (used by the autotuner; no counterpart in the reference,
 * whose plan is fixed by signalsmith-fft.h:93-185). */
typedef struct {
    int variant;          /* compile-time planned kernel schedule: -1 default,
                             -2 none (runtime-planned kernel), k >= 0 that variant */
    int nradices;         /* > 0: radix order of the runtime-planned pass */
    int radices[16];
    int transforms_per_block; /* 0: planner's choice */
    size_t split;         /* two-pass N1 (0: planner's choice, 1: force one pass) */
    int elements_per_thread;  /* 16 or 32 (0: planner's choice) */
    int mfma;             /* 16-bit storage: 1 = radix-16 stages on MFMA; -1 (default) or
                             0 = float butterflies (measured faster) */
} sfft_plan_hints;

/* A plan built with hints (not cached, not used by sfft_plan_create). */
int sfft_plan_create_hinted(sfft_plan *plan, size_t n, int kind, int precision,
                            const sfft_plan_hints *hints);
/* Number of compiled fixed-plan schedules for (n, kind, precision). */
int sfft_plan_variants(size_t n, int kind, int precision);

/* ---- autotuner (no counterpart in the reference) ----
 * sfft_autotune times the planner's candidates for (n, kind, precision) --
 * compiled schedule variants, radix multisets and their order, transforms
 * per block, elements per thread, two-pass splits and column tile widths,
 * MFMA radix-16 stages for 16-bit storage -- over `batch` transforms of synthetic data in
 * HBM (0: about 1 GB per launch), forward and inverse, and makes the
 * fastest the plan that sfft_plan_create returns for that key from then on.
 * Tuned choices are saved / loaded as JSON keyed by the GPU architecture and
 * the HIP runtime version; a file from another device or runtime is ignored
 * when check_device is set.  The environment variable SFFT_TUNE_FILE names a
 * file loaded (checked) at the first plan creation. */
typedef struct {
    sfft_plan_hints hints;  /* the winner */
    double ms;              /* per launch, mean of forward and inverse */
    double gbs;             /* algorithmic bytes (input + output) per second / 1e9 */
    size_t batch;           /* transforms per timed launch */
    int candidates;         /* candidates timed */
} sfft_tune_result;
int sfft_autotune(size_t n, int kind, int precision, size_t batch, sfft_tune_result *result);
int sfft_tune_save(const char *path);
int sfft_tune_load(const char *path, int check_device, size_t *loaded);
int sfft_tune_clear(void);

/* FFT<V>::fft / ifft and RealFFT<V,flags>::fft / ifft
 * (signalsmith-fft.h:374-386, :446-502) over a batch of device buffers.
 *   C2C:       in, out: batch x n complex
 *   R2C fwd:   in: batch x n reals;  out: batch x n/2 complex,
 *              out[0] = (DC, Nyquist) packed (not for the half-shift kind)
 *   R2C inv:   in: batch x n/2 complex; out: batch x n reals
 * batch == 0 is a no-op.  The call is asynchronous on `stream`.
 * in == out transforms in place (the reference accepts any iterator pair,
 * signalsmith-fft.h:55-67, but corrupts in place); other overlaps return
 * SFFT_ERR_ALIASING.  Chirp-z lengths, and two-pass lengths in place, need
 * a device workspace (sfft_plan_workspace_size): sfft_execute allocates it
 * stream-ordered (hipMallocAsync, released after the call). */
int sfft_execute(sfft_plan plan, const void *in, void *out, size_t batch, int direction,
                 void *stream);

/* Bytes of device workspace an execute call of `batch` transforms needs
 * (in_place: in == out); 0 if none. */
int sfft_plan_workspace_size(sfft_plan plan, size_t batch, int in_place, size_t *bytes);
/* sfft_execute with a caller-owned device workspace of at least
 * sfft_plan_workspace_size bytes (may be NULL when that is 0): the call
 * allocates nothing.  Calls that share a workspace must not overlap. */
int sfft_execute_workspace(sfft_plan plan, const void *in, void *out, size_t batch,
                           int direction, void *workspace, size_t workspace_bytes,
                           void *stream);

/* sfft_execute that allocates nothing on the device.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The workspace a
 * call needs (chirp-z lengths; two-pass lengths in place) comes from a
 * library-owned static device array of SFFT_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends; a call needing more returns SFFT_ERR_BUDGET.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race. */
#define SFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int sfft_execute_noalloc(sfft_plan plan, const void *in, void *out, size_t batch, int direction,
                         void *stream);

/* Same as sfft_execute with an explicit distance (in elements of the
 * input / output type) between consecutive transforms of the batch. */
int sfft_execute_strided(sfft_plan plan, const void *in, size_t in_dist, void *out,
                         size_t out_dist, size_t batch, int direction, void *stream);

/* ---- uneven batches: many sizes, one persistent launch ---- */

typedef struct sfft_batch_s *sfft_batch;

typedef struct {
    size_t n;          /* transform length */
    int kind;          /* SFFT_C2C, SFFT_R2C or SFFT_R2C_HALFSHIFT */
    int direction;     /* SFFT_FORWARD / SFFT_INVERSE */
    size_t count;      /* transforms of this shape, back to back */
    size_t in_offset;  /* byte offset of the first input in the input arena */
    size_t out_offset; /* byte offset of the first output in the output arena */
} sfft_batch_item;

/* Builds the work list and the per-size plans (no device work per item at
 * execute time beyond the one launch).  Items of chirp-z lengths (prime
 * factors above 13) run as their plan's own launches after the persistent
 * launch, in the same captured graph, through a workspace the batch owns. */
int sfft_batch_create(sfft_batch *batch, const sfft_batch_item *items, size_t count,
                      int precision);
int sfft_batch_destroy(sfft_batch batch);
/* One persistent launch over every item: in_arena/out_arena are device
 * base pointers for the item offsets.  Launches of one batch share its work
 * queue, so they are serialised on the device: a launch waits for the
 * previous launch of the same batch (on any stream, from any thread). */
int sfft_batch_execute(sfft_batch batch, const void *in_arena, void *out_arena, void *stream);
/* Same work as one sfft_execute per item group (the plain per-size loop). */
int sfft_batch_execute_per_size(sfft_batch batch, const void *in_arena, void *out_arena,
                                void *stream);

/* ---- serving: continuous batching of transform requests ----
 * No counterpart in the reference (a single-threaded library).  Requests of
 * any kind, size and batch are admitted in arrival order as in-flight slots
 * and HBM budget (inputs, outputs and chirp-z workspaces) free up.  The small
 * requests admitted together (up to 16 MiB of input + output each) run merged
 * in one persistent launch; larger ones run their plan's own kernels; all of
 * them overlap on several device streams, and a waiter per stream retires
 * them as their completion events fire.  Each result equals the plain
 * sfft_execute of the same request bitwise.  Inputs are synthesised in HBM
 * from the request's seed (stand-in for the client's upload); outputs stay in
 * the server's pool. */

typedef struct sfft_server_s *sfft_server;

typedef struct {
    size_t hbm_budget_bytes; /* pool for request buffers (0: 64 GiB) */
    int max_inflight;        /* requests on the device at once (0: 64) */
    int streams;             /* device streams (0: 4) */
    int precision;           /* SFFT_F32 / SFFT_F64 */
    int plain;               /* 1: serve one request at a time, synchronously */
} sfft_server_config;

typedef struct {
    uint64_t id;
    size_t n;
    int kind;      /* SFFT_C2C or SFFT_R2C */
    int direction; /* SFFT_FORWARD / SFFT_INVERSE */
    size_t batch;
    uint64_t seed; /* input: U[-1, 1) from this seed */
    int retain;    /* 1: keep the output until sfft_server_release */
} sfft_request;

typedef struct {
    uint64_t id;
    int status;             /* SFFT_OK or an error code */
    double t_arrive, t_admit, t_done; /* seconds, server clock */
    size_t bytes;           /* algorithmic bytes of the request */
} sfft_completion;

int sfft_server_create(sfft_server *server, const sfft_server_config *config);
int sfft_server_destroy(sfft_server server);
/* Queue a request (never blocks on the device; SFFT_ERR_BUDGET if a single
 * request exceeds the whole budget). */
int sfft_server_submit(sfft_server server, const sfft_request *request);
/* Completed requests since the last poll (up to max). */
int sfft_server_poll(sfft_server server, sfft_completion *out, size_t max, size_t *count);
/* Block until every submitted request has completed. */
int sfft_server_drain(sfft_server server);
/* Copy a retained output to host memory / free it. */
int sfft_server_fetch(sfft_server server, uint64_t id, void *host_dst, size_t bytes);
int sfft_server_release(sfft_server server, uint64_t id);
/* Seconds on the server clock. */
double sfft_server_now(sfft_server server);

/* Built-in load generator: Poisson arrivals at `rate` req/s for `duration`
 * seconds; each request is C2C forward, C2C inverse or R2C forward (1/3
 * each), a fast size log-uniform in [n_min, n_max], batch uniform in
 * [1, batch_max], inputs from per-request seeds; returns when all have
 * completed.  records: one per request (up to max_records). */
typedef struct {
    double rate, duration;
    size_t n_min, n_max, batch_max;
    uint64_t seed;
} sfft_poisson_config;
int sfft_server_run_poisson(sfft_server server, const sfft_poisson_config *config,
                            sfft_completion *records, size_t max_records, size_t *count);

/* ---- memory helpers for host-side bindings (the drop-in header uses them to
 * stage host iterators; no counterpart in the reference, which is host-only) ---- */

#define SFFT_COPY_H2D 1
#define SFFT_COPY_D2H 2
#define SFFT_COPY_D2D 3

int sfft_malloc_device(void **ptr, size_t bytes);
int sfft_free_device(void *ptr);
/* page-locked host memory (staging for asynchronous copies), mapped for
 * device access and coherent */
int sfft_malloc_host(void **ptr, size_t bytes);
int sfft_free_host(void *ptr);
int sfft_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int sfft_stream_synchronize(void *stream);
/* Device address of page-locked host memory from sfft_malloc_host: kernels
 * may read and write it directly (over the host link). */
int sfft_host_device_pointer(void **device_ptr, void *host_ptr);
/* Streams (non-blocking, default priority) and events, so that bindings can
 * overlap staging copies with host work. */
int sfft_stream_create(void **stream);
int sfft_stream_destroy(void *stream);
int sfft_event_create(void **event);
int sfft_event_destroy(void *event);
int sfft_event_record(void *event, void *stream);
int sfft_event_synchronize(void *event);

/* ---- synthetic inputs (seeded, counter-based splitmix64) ---- */

/* dist: 0 = U[-1, 1), 1 = N(0, 1) (Box-Muller); count values of the
 * precision's real type (complex data: 2 per element). */
int sfft_fill(void *dst, size_t count, int precision, int dist, uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SFFT_H */
