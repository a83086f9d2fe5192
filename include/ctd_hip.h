/*
 * ctd_hip.h -- C ABI of the MI355X (gfx950) trial-factoring engine.
 *
 * This is the thin boundary the C++ host (CLI, number_sequence plugins,
 * checkpointing) calls into.  Signatures hold only plain integers, pointers
 * and sizes.  Every entry point returns CTD_OK (0) or a negative CTD_E_* code;
 * ctd_hip_last_error() then describes the failure.
 *
 * Each declaration names the reference interface it replaces (paths are under
 * the reference's src/).
 */
#ifndef CTD_HIP_H
#define CTD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTD_HIP_ABI_VERSION 2

enum {
    CTD_OK = 0,
    CTD_E_INVALID = -1,   /* bad argument (host validation) */
    CTD_E_DEVICE = -2,    /* HIP runtime / kernel error */
    CTD_E_NOMEM = -3,     /* device allocation failed */
    CTD_E_UNSUPPORTED = -4,
    CTD_E_BUSY = -5       /* serving queue full */
};

/* Sequence ids.  The values are the persisted order of num_seq_id
 * (num_seq_id.hpp:16-27); CTD_SEQ_PEANO is a test-only sequence S(n) = n
 * (tests/peano.hpp:127-137) and is never persisted. */
enum {
    CTD_SEQ_MERSENNE = 0,
    CTD_SEQ_PROTH = 1,
    CTD_SEQ_RIESEL = 2,
    CTD_SEQ_SMARANDACHE = 3,
    CTD_SEQ_FIBONACCI = 4,
    CTD_SEQ_LUCAS = 5,
    CTD_SEQ_PROTH_FIXED_N = 6,
    CTD_SEQ_RIESEL_FIXED_N = 7,
    CTD_SEQ_TRIBONACCI = 8,
    CTD_SEQ_PEANO = 255
};

/* One trial-factoring job: the fields of trial_factoring_options
 * (trial_factoring_options.hpp:25-39) plus the parsed sequence parameter of
 * num_seq_spec (num_seq_dispatch.hpp:119-168). */
typedef struct ctd_job {
    uint32_t seq_id;        /* CTD_SEQ_* */
    uint32_t max_factors;   /* max_factors_per_number (>= 1) */
    uint64_t param;         /* k (proth, riesel), n (fixed-n), base (smarandache, 2..64) */
    uint64_t n0, n1;        /* index range [n0, n1) */
    uint64_t f0, f1;        /* prime range [f0, f1); f0 already moved past a resumed prime */
} ctd_job;

/* Host callbacks around the segment loop of device_trial_factor
 * (trial_factoring.cuh:437-488).  Any member may be NULL. */
typedef struct ctd_callbacks {
    void *user;
    /* After each completed prime batch: upper bound of the batch and the last
     * prime processed (progress::update, progress.cpp:70-104). */
    void (*progress)(void *user, uint64_t bound, uint64_t last_prime);
    /* Nonzero when a checkpoint is due (checkpoint_manager::due, checkpoint.cpp:166-170). */
    int (*checkpoint_due)(void *user);
    /* The host counts/factors arrays were refreshed and hold every factor of
     * every prime <= last_prime (trial_factoring.cuh:455-465). */
    void (*checkpoint)(void *user, uint64_t last_prime);
} ctd_callbacks;

/* Launch shape overrides: --grid-size / --block-size (command_line.cpp:148-164)
 * and --segment-length (trial_factoring_options.cpp:31).  Zero = autotuned. */
typedef struct ctd_run_opts {
    uint64_t batch_length;  /* integers of the prime range per device batch */
    int32_t grid_size;
    int32_t block_size;
    uint32_t chunk_length;  /* indices per work item (index chunking) */
    uint32_t flags;         /* CTD_RUN_* */
} ctd_run_opts;

#define CTD_RUN_GENERIC_ONLY 1u /* force the generic 64-bit Montgomery kernels */
#define CTD_RUN_NO_FP64 2u      /* integer forms only (no FP64 residues below 2^48) */
#define CTD_RUN_SM_FULL_STEP 4u /* Smarandache FP64 form: the general 7-instruction step only (no short step) */
#define CTD_RUN_STEPPED 8u      /* fixed-n families: step every index (riesel_proth_fixed_n.hpp:142-151)
                                   instead of solving k = -+2^-n (mod p) once per prime */
#define CTD_RUN_SM_NO_SHORT 16u /* Smarandache FP64 form: no short step (where r b^digits is exact the
                                   shifted step, or with CTD_RUN_SM_NO_SHIFT the 6-instruction
                                   exact-product step; else the general step); test hook */
#define CTD_RUN_SM_NO_SHIFT 32u /* Smarandache FP64 form: no shifted step (the short, exact-product or
                                   general step, as CTD_RUN_SM_NO_SHORT allows); test hook */
#define CTD_RUN_SM_NO_PAIR 64u  /* Smarandache: one prime per lane instead of two interleaved; test hook */

typedef struct ctd_hip_engine ctd_hip_engine;

/* Opens the engine on a device (cudaSetDevice + device properties,
 * trial_factoring.cuh:349-363). */
int ctd_hip_open(int device, ctd_hip_engine **out);
/* Releases every device resource of the engine. */
void ctd_hip_close(ctd_hip_engine *eng);
/* Message of the last failure on this engine (or of the last failed open when
 * eng is NULL).  The string stays valid until the next call on the engine. */
const char *ctd_hip_last_error(const ctd_hip_engine *eng);
/* Device name, total memory and compute-unit count
 * (print_gpu_device_info, trial_factoring.cuh:241-252). */
int ctd_hip_device_info(ctd_hip_engine *eng, char *name, size_t name_len,
                        uint64_t *total_mem, int32_t *cu_count);
/* ABI version of the loaded library (CTD_HIP_ABI_VERSION). */
int ctd_hip_abi_version(void);

/* Prime supply: primes of [start, end) in ascending order, generated by the
 * device sieve.  Replaces siever::sieve (siever.hpp:179-237) plus the H2D
 * copy of device_prime_data::copy_from_host (prime_data.hpp:165-196).
 * Writes at most cap primes to out (host memory) and the total to *count. */
int ctd_hip_sieve(ctd_hip_engine *eng, uint64_t start, uint64_t end,
                  uint64_t *out, uint64_t cap, uint64_t *count);

/* Plain single-job form: replaces device_trial_factor<Seq>
 * (trial_factoring.cuh:336-513) for the built-in families.
 * counts[n1-n0] and factors[(n1-n0)*max_factors] are host arrays in the
 * factors_buffer layout (factors_buffer.hpp:77-86); they are read (resume
 * state) and written back.  *last_prime receives the last prime processed
 * (0 if none).  opts and cb may be NULL. */
int ctd_hip_trial_factor(ctd_hip_engine *eng, const ctd_job *job,
                         const ctd_run_opts *opts, const ctd_callbacks *cb,
                         uint32_t *counts, uint64_t *factors, uint64_t *last_prime);

/* Same as ctd_hip_trial_factor with counts/factors already in device memory
 * (d_counts, d_factors) and no host copies: the form timed with its buffers
 * resident in HBM.  stream is a hipStream_t (NULL = the engine's stream). */
int ctd_hip_trial_factor_device(ctd_hip_engine *eng, const ctd_job *job,
                                const ctd_run_opts *opts, uint32_t *d_counts,
                                uint64_t *d_factors, uint64_t *last_prime,
                                void *stream);

/* Allocation-free form of ctd_hip_trial_factor_device: the same arguments and
 * results, and no device memory is allocated or freed during the call.  No
 * counterpart in the reference; it is for callers that cannot allocate:
 * code being captured into a HIP graph, tight serving loops, harnesses that
 * forbid hipMalloc.  Its sieving base, prime stream, segment counts and
 * scalars live in one fixed device array of CTD_NOALLOC_WORKSPACE_BYTES
 * (about 46 MiB) per GPU, which each GPU the library runs on holds until the
 * process ends; ctd_hip_open looks up its address.  Limits: f1 <=
 * CTD_NOALLOC_F1_MAX = 2^48 (the array holds the sieving primes below 2^24;
 * larger windows return CTD_E_UNSUPPORTED), and device batches of at most
 * 2^25 integers (opts->batch_length above that is capped).  Every call on a
 * GPU uses that GPU's one array and leaves its contents behind: calls that
 * overlap on one GPU (from two threads, or on two streams) race and must not
 * be made. */
#define CTD_NOALLOC_F1_MAX ((uint64_t)1 << 48)
#define CTD_NOALLOC_WORKSPACE_BYTES ((uint64_t)46611968)
#define CTD_MAX_DEVICES 64
int ctd_hip_trial_factor_device_noalloc(ctd_hip_engine *eng, const ctd_job *job,
                                        const ctd_run_opts *opts, uint32_t *d_counts,
                                        uint64_t *d_factors, uint64_t *last_prime,
                                        void *stream);

/* Statistics of the last run on this engine: primes tested (odd primes of the
 * window), residue updates (primes x indices), kernel milliseconds of the
 * trial-division kernels and of the sieve kernels (HIP events). */
typedef struct ctd_run_stats {
    uint64_t primes;
    uint64_t updates;
    double tf_ms;
    double sieve_ms;
    double total_ms;
    uint32_t tf_launches;
    uint32_t sieve_launches;
} ctd_run_stats;
int ctd_hip_last_stats(const ctd_hip_engine *eng, ctd_run_stats *out);

/* ---- Batched form ------------------------------------------------------
 * Many jobs of any family, index range and prime window in one hipGraph
 * launch.  Each job's prime window is cut into tiles; a tile is sieved into
 * one of `streams` resident prime slots and propagated over the job's
 * indices by that family's residue kernel, while the other slots sieve and
 * propagate other tiles.  The launch shape of each family (workgroup size,
 * waves per CU, index-chunk split) is chosen by a timed probe on first use and
 * cached in the engine.  No reference counterpart: the reference runs one job
 * per process (trial_factoring.cuh:336-513 per job); results per job are the
 * same as ctd_hip_trial_factor's. */
#define CTD_BATCH_HOST_BUFFERS 1u /* counts/factors are host arrays (copied in and back) */
#define CTD_BATCH_NO_GRAPH 2u     /* issue the tiles eagerly on the streams (no graph) */
#define CTD_BATCH_NO_AUTOTUNE 4u  /* fixed shape: 256-lane groups, one item per lane, no extra chunking */
#define CTD_BATCH_NO_FP64 8u      /* integer residue forms only */
#define CTD_BATCH_STEPPED 16u     /* fixed-n families step every index (as CTD_RUN_STEPPED) */

typedef struct ctd_batch_opts {
    uint64_t tile_length;  /* integers of a job's prime window per tile (0 = 2^28) */
    uint32_t streams;      /* tiles in flight = prime slots (0 = 4, at most 16) */
    uint32_t flags;        /* CTD_BATCH_* */
} ctd_batch_opts;

typedef struct ctd_batch_stats {
    uint64_t primes;       /* odd primes tested over all jobs */
    uint64_t updates;      /* sum over jobs of primes x indices */
    double total_ms;       /* device time of the batch launch(es), HIP events */
    double build_ms;       /* host time to capture and instantiate the graph(s) */
    double autotune_ms;    /* host time spent in launch-shape probes during this call */
    uint32_t tiles;
    uint32_t graphs;       /* graph launches (tiles are split into graphs of <= 1024 tiles) */
    uint32_t tuned;        /* launch shapes probed during this call */
    uint32_t reserved;
} ctd_batch_stats;

/* counts[j] / factors[j]: job j's factors buffer (device pointers, or host
 * arrays with CTD_BATCH_HOST_BUFFERS), accumulated into like
 * ctd_hip_trial_factor.  stream is a hipStream_t (NULL = the engine's). */
int ctd_hip_run_batch(ctd_hip_engine *eng, const ctd_job *jobs, uint32_t njobs,
                      uint32_t *const *counts, uint64_t *const *factors,
                      const ctd_batch_opts *opts, ctd_batch_stats *stats, void *stream);

/* Launch shape chosen for a job by the batch autotuner (probing it if needed):
 * block (lanes per workgroup: 64 or 256), primes per lane (work items each
 * lane takes in turn: 1, 2 or 4; the reference runs B = 1..3 primes per
 * thread, trial_factoring.cuh:472-481) and the index-chunk split (1, 2 or 4).
 * Test and report hook. */
int ctd_hip_batch_shape(ctd_hip_engine *eng, const ctd_job *job, uint32_t *block,
                        uint32_t *primes_per_lane, uint32_t *chunk_split);

/* ---- Serving form -------------------------------------------------------
 * Continuous batching: jobs are submitted while others run.  A scheduler
 * thread admits them while their factors buffers fit in a device-memory
 * budget, cuts their prime windows into tiles and interleaves the tiles of the
 * admitted jobs on `streams` streams, each with a resident prime slot; the
 * sieving base stays resident.  Each job's factors buffer is the one
 * ctd_hip_trial_factor would produce.  No reference counterpart: the reference
 * CLI runs one job per process.  While a server is open, the engine handle
 * must not be used for other calls. */
typedef struct ctd_hip_server ctd_hip_server;

typedef struct ctd_server_opts {
    uint64_t memory_budget; /* bytes of job buffers, counted in power-of-two size classes:
                               device factors buffers of admitted jobs, and separately the
                               pinned staging buffers of admitted and uncollected jobs
                               (0 = 8 GiB each); cached buffers are freed to stay under it */
    uint64_t tile_length;   /* integers of a prime window per tile (0 = 2^26) */
    uint32_t streams;       /* streams = resident prime slots (0 = 3, at most 16) */
    uint32_t queue_limit;   /* jobs waiting for admission before submit returns CTD_E_BUSY (0 = 65536) */
    uint32_t flags;         /* CTD_SERVE_* */
    uint32_t depth;         /* estimated work each stream keeps in flight, in tiles
                               of 2^35 VALU lane-instructions (0 = 2, at most 16) */
} ctd_server_opts;
#define CTD_SERVE_FRESH_BUFFERS 1u /* jobs start from empty factors buffers: the host
                                      buffers are only written, never read */
#define CTD_SERVE_FIFO 2u          /* tiles go to the oldest admitted job (default: places
                                      alternate between the oldest job and the one with the
                                      least remaining work) */
#define CTD_SERVE_OWN_QUEUES 4u    /* the server's streams are created with a CU mask of every
                                      CU, which gives each a hardware queue of its own: latency
                                      no longer depends on how the process's other streams share
                                      the GPU_MAX_HW_QUEUES queues (default: non-blocking streams,
                                      which HIP places on those queues with the others).  Such
                                      streams are blocking streams: work on the null stream waits
                                      for them and they for it */

typedef struct ctd_server_stats {
    uint64_t submitted, admitted, completed, failed;
    uint64_t tiles;         /* tiles issued */
    uint64_t primes;        /* odd primes tested by completed jobs */
    uint64_t updates;       /* primes x indices of completed jobs */
    uint32_t active_max;    /* largest number of admitted jobs at once */
    uint32_t pending_max;   /* largest queue of jobs waiting for admission */
    uint32_t pending;       /* jobs waiting for admission now */
    uint32_t reserved;
    uint64_t device_bytes_peak; /* most device memory the job-buffer pool held (in use + cached) */
    uint64_t pinned_bytes_peak; /* most pinned host memory the staging pool held */
    double scheduler_cpu_ms;    /* CPU time of the scheduler thread so far */
    uint64_t scheduler_passes;  /* scheduler passes so far (each follows a wake-up or progress) */
} ctd_server_stats;

int ctd_hip_server_open(ctd_hip_engine *eng, const ctd_server_opts *opts, ctd_hip_server **out);
/* counts/factors: the job's host factors buffer (read at admission, written
 * when the job is collected); it must stay valid until the job is done.  The
 * device results go to a pinned staging buffer of the server; the thread that
 * collects the job (poll reporting done, wait, or close) copies them into
 * counts/factors.  Returns CTD_E_BUSY when queue_limit jobs are already
 * waiting. */
int ctd_hip_server_submit(ctd_hip_server *srv, const ctd_job *job, uint32_t *counts,
                          uint64_t *factors, uint64_t *ticket);
/* *state: 0 queued, 1 admitted, 2 done (the results are in the job's
 * buffers), negative: the CTD_E_* code of a failure.  *latency_ms: submit to
 * the moment the results were back in host memory, once done. */
int ctd_hip_server_poll(ctd_hip_server *srv, uint64_t ticket, int *state, double *latency_ms);
/* Blocks until the job is done (or failed). */
int ctd_hip_server_wait(ctd_hip_server *srv, uint64_t ticket, double *latency_ms);
int ctd_hip_server_stats(ctd_hip_server *srv, ctd_server_stats *out);
/* Completes every submitted job, then releases the server. */
int ctd_hip_server_close(ctd_hip_server *srv);

/* Test hook of the residue kernels: for each odd modulus primes[i] (host
 * array), runs the same kernel path as ctd_hip_trial_factor over [n0, n1) and
 * writes the plain residue S(n1-1) mod primes[i] to out[i] (host array).
 * flags takes the CTD_RUN_* values.  Hits are recorded into scratch buffers
 * and discarded.  No reference counterpart: it exposes the per-prime value
 * that trial_div_by_odd_primes (trial_factoring.cuh:156-228) keeps in
 * registers, so every step of the propagation can be checked, not only the
 * rare zero residues. */
int ctd_hip_residues(ctd_hip_engine *eng, const ctd_job *job, const uint64_t *primes,
                     uint64_t nprimes, uint64_t *out, uint32_t flags);

/* VALU throughput probe: lane-instructions per second of one instruction
 * (op: 0 v_add_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32, 3 v_mad_u64_u32,
 * 4 v_mad_u32_u24, 5 v_lshl_add_u64, 6 v_cndmask_b32, 7 v_alignbit_b32,
 * 8 v_min_u32, 9 v_lshlrev_b64, 10 v_add_f64, 11 v_fma_f64, 12 v_mul_f64,
 * 13 v_max_f64, 14 v_cmp_eq_u32, 15 v_cmp_eq_f64, 16 v_cmp_eq_u64 (compares
 * writing a lane mask to scalar registers), 17 v_xor_b32, 18 v_sub_u32,
 * 19 v_add_f32) over independent chains on every CU.  The
 * measured integer peak the roofline is quoted against.  No reference
 * counterpart: the reference publishes no device peak. */
int ctd_hip_valu_probe(ctd_hip_engine *eng, int op, uint32_t iters, double *lane_ops_per_s);

#ifdef __cplusplus
}
#endif
#endif
