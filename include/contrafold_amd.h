/*
 * contrafold_amd.h -- C ABI of the MI355X (gfx950) CONTRAfold fold engine.
 *
 * The reference (atlasbioinfo/Contrafold_GPU) has no C layer: its boundary is
 * the Python package gpu_contrafold (gpu_contrafold/__init__.py:17-31).  Each
 * entry point below names the reference function whose work it performs; the
 * Python drop-in (contrafold_gpu_amd/__init__.py) binds them with ctypes.
 *
 * Conventions
 *   - All calls return an int status: CF_OK (0) or a negative CF_E* code, and
 *     never throw.  cf_last_error() gives the message of the last failure.
 *   - Sequences are passed as one char buffer plus n+1 offsets (sequence b is
 *     seqs[offsets[b] .. offsets[b+1])).  Characters are upper-cased, T is U,
 *     anything outside ACGU is N, which never pairs and scores 0
 *     (gpu_contrafold/cpu.py:23-29, 138-144).
 *   - masks: NULL, or one byte per base in the same layout; nonzero = forced
 *     unpaired (gpu_contrafold/gpu.py:84 tests fo[i] == 0).
 *   - Parameter tables: 2577 doubles in the order of the dict returned by
 *     gpu_contrafold/cpu.py:load (cpu.py:132-135): bp[5][5] stack[5][5][5][5]
 *     tm[5][5][5][5] hc[5][5] dl[5][5][5] dr[5][5][5] hp_cum[31] cs[31][31]
 *     i11[5][5] b01[5] mb mu mp eu ep.
 *   - Results come back in input order.  Structures are dot-bracket bytes
 *     ('.', '(', ')'), L bytes per sequence, concatenated in input order.
 *   - A context owns one HIP stream and the device parameter tables; a batch
 *     owns the device-resident DP tables of its sequences, so logZ -> bpp ->
 *     MEA / samples / gradient reuse them across cf_batch_run calls.  Calls on
 *     one context must not overlap.
 */
#ifndef CONTRAFOLD_AMD_H
#define CONTRAFOLD_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cf_ctx cf_ctx;
typedef struct cf_batch cf_batch;
typedef struct cf_server cf_server;

enum {
    CF_OK = 0,
    CF_EINVAL = -1,    /* bad argument (maps to ValueError) */
    CF_ENOMEM = -2,    /* device memory / admission budget (MemoryError) */
    CF_EDEVICE = -3,   /* HIP runtime or kernel failure (RuntimeError) */
    CF_ENOPARAMS = -4, /* cf_set_params not called */
    CF_ESTATE = -5     /* result requested that the batch has not computed */
};

/* work flags for cf_batch_run */
enum {
    CF_LOGZ = 1,    /* inside + exterior chain: gpu.logZ_batch (gpu.py:480-517) */
    CF_BPP = 2,     /* + outside + posterior: cpu.bpp (cpu.py:926-929) */
    CF_MEA = 4,     /* + MEA decode per gamma: cpu.mea (cpu.py:932-941) */
    CF_MFE = 8,     /* FP64 Viterbi + traceback: cpu.mfe (cpu.py:505-519) */
    CF_SAMPLE = 16, /* Boltzmann traceback: gpu.sample_batch (gpu.py:520-583) */
    CF_GRAD = 32,   /* expected counts = d logZ / d table (no reference counterpart) */
    CF_LOGZ64 = 64, /* FP64 inside with exact log1p/exp: cpu.logZ (cpu.py:165-287) */
    CF_GRAPH = 256, /* replay the run as a captured hipGraph (steady-state repeats) */
    CF_TIMING = 512,   /* record per-kernel device times (cf_batch_kernel_ms) */
    CF_RECOMPUTE = 1024 /* recompute stages even if their tables are resident */
};

int cf_version(void);
int cf_table_size(void); /* 2577 */

/* Context: gpu.py allocates per call; here the device, stream and parameter
 * tables live in a context.  device = HIP device ordinal. */
int cf_create(int device, cf_ctx **out);
void cf_destroy(cf_ctx *ctx);
const char *cf_last_error(const cf_ctx *ctx);

/* Upload parameter tables (replaces the per-call cuda.to_device of
 * gpu.py:500-514 and the FP32 cast of cpu._params_f32, cpu.py:898-910). */
int cf_set_params(cf_ctx *ctx, const double *tables);

/* Length buckets: sequences with L <= edges[k] (and above edges[k-1]) run in
 * one launch per kernel with team width threads[k] (a multiple of 64 in
 * 64..1024): 64 packs several sequences into a workgroup, one wave each
 * (cf_set_packing); wider teams give a sequence the whole workgroup (lower
 * latency for long folds).  Defaults are set at cf_create; the autotuner
 * overrides them. */
int cf_set_tuning(cf_ctx *ctx, const int *edges, const int *threads, int n_buckets);

/* Packing (no counterpart in the reference, which launches one thread block
 * per sequence, gpu.py:508): one-wave teams per workgroup, 1..16, of the
 * packed (width 64) inside and outside sweeps; fewer are used when LDS is
 * short.  Results do not depend on it.  Defaults 8 and 4; the autotuner
 * overrides them. */
int cf_set_packing(cf_ctx *ctx, int teams_inside, int teams_outside);

/* Split teams (no counterpart in the reference, whose kernel gives a fold
 * one thread block, gpu.py:55-184, 508): in a wide bucket (team width > 64)
 * folds of at least min_length nt run their inside, outside and posterior
 * stages on wgs_per_fold workgroups of 256 threads -- one wave per SIMD on
 * wgs_per_fold CUs -- that meet at a barrier per anti-diagonal, which cuts a
 * long fold's latency.  Results are bit-identical to the one-workgroup team.
 * A launch is kept to two workgroups per CU in total (fewer parts per fold
 * otherwise); if the parts of a fold are still not resident together the
 * run fails with CF_EDEVICE after a bounded wait instead of hanging.
 * wgs_per_fold = 1 (or min_length = 0) turns splitting off, the default. */
int cf_set_split(cf_ctx *ctx, int min_length, int wgs_per_fold);

/* Batch: encode, pack and upload n sequences (gpu.py:482-498). keys: NULL or
 * per-sequence 64-bit ids for the sampler's random streams (default: index). */
int cf_batch_create(cf_ctx *ctx, int n, const char *seqs, const int64_t *offsets, const uint8_t *masks,
                    const uint64_t *keys, cf_batch **out);
/* Run the stages in flags (dependencies are added: BPP implies LOGZ, ...).
 * gammas: n_gamma MEA trade-offs (cpu.mea gamma).  n_samples / seed: sampler.
 * grad_weights: NULL (all 1) or one weight per sequence; CF_GRAD returns
 * sum_b w_b * d logZ_b / d table.  Synchronous. */
int cf_batch_run(cf_batch *b, unsigned flags, const float *gammas, int n_gamma, int n_samples, uint64_t seed,
                 const float *grad_weights);
int cf_batch_logz(const cf_batch *b, float *out);                /* n floats */
int cf_batch_bpp(const cf_batch *b, int index, float *out);      /* L*L, 0-based upper triangle */
int cf_batch_mea(const cf_batch *b, int gamma_index, char *out); /* sum L bytes */
int cf_batch_mfe(const cf_batch *b, char *out, double *scores);  /* sum L bytes; scores may be NULL */
int cf_batch_samples(const cf_batch *b, char *out); /* n_samples * sum L bytes: seq-major, sample-minor */
int cf_batch_grad(const cf_batch *b, double *out);  /* 2577 doubles */
int cf_batch_logz64(const cf_batch *b, double *out); /* n doubles (CF_LOGZ64) */
/* Copy one DP table of sequence `index` (introspection for parity checks):
 * which = 0 FC, 1 FM, 2 FM1, 3 FM2, 4 FCo, 5 FMo, 6 FM1o, 7 FM2o, 8 POST
 * ((L+1)(L+2)/2 floats, diagonal-major: cell (i, j) at d*(L+1) - d*(d-1)/2 + i
 * with d = j - i; POST cell (a, b) holds P(a pairs b+1)), 9 F5, 10 F5o (L+2). */
int cf_batch_table(const cf_batch *b, int index, int which, float *out);
/* Device milliseconds of the last cf_batch_run (HIP events on the batch stream). */
float cf_batch_last_ms(const cf_batch *b);
/* Per-kernel device milliseconds of the last CF_TIMING run; names in
 * cf_kernel_name(k), k < cf_kernel_count().  In a run with logZ, bpp and
 * samples the sampler runs on a lane stream beside the outside sweep (and the
 * start of the posterior), so those stages overlap; its time runs from the
 * outside sweep's start to its own end. */
int cf_kernel_count(void);
const char *cf_kernel_name(int k);
int cf_batch_kernel_ms(const cf_batch *b, float *out);
int64_t cf_batch_device_bytes(const cf_batch *b);
void cf_batch_destroy(cf_batch *b);

/* Device bytes a fold of length L needs for the given flags (admission). */
int64_t cf_fold_bytes(int L, unsigned flags, int n_gamma, int n_samples);

/* Fold without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same work and results as
 * cf_batch_create + cf_batch_run(flags) + cf_batch_logz / cf_batch_mea (index
 * 0) + cf_batch_destroy, for flags CF_LOGZ or CF_MEA (one gamma), but every
 * device buffer -- inputs, DP tables, results -- is carved out of one fixed
 * 8 MiB device array that the library holds on each GPU until the process
 * ends; a call whose tables do not fit (about 8 folds of 200 nt, or one of
 * 590 nt, with CF_MEA) is rejected with CF_EINVAL.  The call makes no
 * hipMalloc / hipFree and is synchronous on the context stream.  All calls
 * on one GPU share the array: calls that overlap on one GPU race. */
int cf_fold_noalloc(cf_ctx *ctx, int n, const char *seqs, const int64_t *offsets, const uint8_t *masks,
                    unsigned flags, float gamma, float *out_logz, char *out_mea);

/* Diagnostic, no counterpart in the reference: one device-to-device copy
 * kernel over `bytes` (a positive multiple of 16; read once, written once) on
 * the context stream, synchronous.  It is the known-size reference against
 * which hardware traffic counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) are
 * scaled when the fold kernels' HBM bytes are measured. */
int cf_copy_probe(cf_ctx *ctx, int64_t bytes);

/* ---------------------------------------------------------------- serving
 * Continuous batching (no counterpart in the reference, which has no server).
 * Requests are admitted while their O(L^2) table bytes fit the HBM budget;
 * admitted requests are launched as batches on a pool of streams, so work
 * that arrives while others run overlaps with it on the device. */
/* n_streams (1..32) slots, each with its own stream.  HIP maps streams onto
 * GPU_MAX_HW_QUEUES hardware queues (default 4) and a queue runs in order, so
 * set that environment variable >= n_streams before the process first uses the
 * GPU, or short folds wait behind long ones that share their queue.  Measured
 * on the serving mix: 16 slots; 32 oversubscribe the hardware queues (p50 rose
 * from ~80 to ~450 ms), and one stream per length class with two batches in
 * flight serialised each class and fell behind the load. */
int cf_server_create(cf_ctx *ctx, int64_t hbm_budget, int max_batch, int n_streams, cf_server **out);
/* kind: CF_MEA, CF_BPP, CF_SAMPLE or CF_LOGZ; returns a request id (>= 0) or
 * a negative status.  seed: the request's 64-bit sampling-stream key (the same
 * key gives the same samples whatever else is in flight).  arrival_us: time
 * on the server clock (microseconds since cf_server_create) for latency
 * accounting; 0 = now. */
int64_t cf_server_submit(cf_server *s, const char *seq, int L, const uint8_t *mask, unsigned kind, float gamma,
                         int n_samples, uint64_t seed, double arrival_us);
/* Drive the scheduler once (admit, launch, retire); returns the number of
 * completed requests copied into ids (up to cap). */
int cf_server_poll(cf_server *s, int64_t *ids, int cap);
/* Result of a completed request.  structure: L bytes (MEA) / n_samples*L
 * bytes (SAMPLE) / L*L float32 posteriors (BPP) / unused (LOGZ); logz: logZ;
 * latency_us: completion - arrival. */
int cf_server_result(cf_server *s, int64_t id, char *structure, float *logz, double *latency_us);
int cf_server_pending(const cf_server *s);
/* Counters since cf_server_create, up to n doubles in this order: batches
 * launched, folds launched, host microseconds spent launching, host
 * microseconds spent retiring, summed batch residence (launch to retirement,
 * microseconds), largest queue length seen.  Returns how many were written. */
int cf_server_stats(const cf_server *s, double *out, int n);
/* Replay a recorded arrival trace in real time (arrival_us sorted, relative
 * to the start) and fill per-request latencies (us) and, if wait_us is not
 * NULL, the part of each latency spent queued before the request's batch was
 * launched.  Returns CF_OK. */
int cf_server_replay(cf_server *s, int n, const char *seqs, const int64_t *offsets, const uint8_t *masks,
                     const unsigned *kinds, const float *gammas, const int *n_samples, const double *arrival_us,
                     uint64_t seed, double *latency_us, double *wait_us, double *wall_s);
void cf_server_destroy(cf_server *s);

#ifdef __cplusplus
}
#endif
#endif
