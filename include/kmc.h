/* kmc.h -- C ABI of the MI355X k-mer counting engine (libkmc.so).
 *
 * Drop-in boundary for the counting path of KmerTreeCounter (tree_v4).  The
 * reference has no FFI: its path is a set of template classes wired together in
 * process_main<N> (src/main.cpp:111-312).  These entry points replace, one to
 * one, the seams the reference host code calls there:
 *
 *   kc_create        <- construction of ParserThreadPool / ClassifierThreadPool /
 *                       SchedulerThreadPool / KmerTree (src/main.cpp:195-215)
 *   kc_sample_block  <- FastqPreParser::parse_and_push + FastqPrefixCounter::
 *                       count_prefixes (Stage 0, src/main.cpp:140-171,
 *                       src/FastqPrefixCounter.h:31-82)
 *   kc_plan          <- prefix clamp + lpt + calculate_bloom_filter_capacity
 *                       (src/main.cpp:173-189, :38-108)
 *   kc_push_block    <- the SPMCRingMemoryPool consumer side: one sequence-only
 *                       block {data, length} as FastqReader publishes it
 *                       (src/FastqReader.h:367-388, src/RingMemoryPool.h:219-237,
 *                       content_type src/definition.h:59-63)
 *   kc_finish        <- parser/classifier/scheduler join + final_drain_parallel
 *                       (src/main.cpp:242-264)
 *   kc_emit_firsts   <- the ExportWriter<N> input: first-occurrence keys that go
 *                       to low.bin (src/ExportWriter.h:400-497)
 *   kc_emit_root     <- FinalDrainWriter::open/write/close per root id
 *                       (src/FinalDrainWriter.h:31-91): ExportRecord<N> records
 *                       (src/definition.h:159-163) whose key has key.data[0]>>56 == root_id
 *   kc_stats         <- "Total read k-mer count" and the phase timers printed by
 *                       process_main (src/main.cpp:266-309)
 *   kc_destroy       <- teardown at the end of process_main
 *
 * Conventions.  Every entry point returns 0 on success and a negative kc_status
 * on failure and never throws.  The engine owns all device memory.  Blocks passed
 * to kc_push_block are copied before the call returns, so the caller may reuse
 * them at once.  Emit callbacks lend buffers that are valid only during the
 * call.  One host thread drives a handle; handles are independent.
 *
 * Key layout: N = 1 (k <= 32), 2 (33 <= k <= 64) or 4 (65 <= k <= 128)
 * little-endian u64 words per key (src/main.cpp:579-595), word 0 most
 * significant, 2 bits per base MSB-first, left aligned (src/kmer.h:12-21).
 * Canonical key = min(pack(reverse(w)), pack(complement(w))) for a window w,
 * as GetKmer computes it (src/GetKmer.h:58-129); at k = 96 word 3 holds the 32
 * bases before the window, as GetKmer leaves it.
 */
#ifndef KMC_H
#define KMC_H

#include <stdint.h>

#ifndef KC_API
#define KC_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kc_status {
    KC_OK = 0,
    KC_ERR_ARG = -1,         /* invalid argument (k out of range, null pointer, ...) */
    KC_ERR_HIP = -2,         /* a HIP runtime call failed */
    KC_ERR_NOMEM = -3,       /* device or pinned host allocation failed */
    KC_ERR_STATE = -4,       /* call out of order (e.g. emit before finish) */
    KC_ERR_NODEV = -5,       /* no usable GPU */
    KC_ERR_CAPACITY = -6     /* a device table could not grow enough */
} kc_status;

typedef enum kc_partition {
    KC_PARTITION_PREFIX = 0, /* fixed 256 root prefixes, then the next key bits */
    KC_PARTITION_LEARNED = 1 /* learned CDF model -> balanced key-range buckets */
} kc_partition;

#define KC_SUB_BITS_NONE 0xFFFFFFFFu

/* Repeat counting inside a bucket.  Both give the same merged table (the
 * low.bin keys plus the root records).  BLOOM (the default) runs the
 * reference's first-occurrence Bloom test (three bits per key) and
 * counts only the repeats, so a Bloom false positive moves a key's first
 * occurrence into its record, as in the reference.  EXACT counts every
 * occurrence in the bucket's table and splits each key into one low.bin entry
 * and a root record of the remaining count (the reference's split when its
 * filter reports no false positive); its table must hold every distinct key
 * of a pass, not only the repeated ones, so it takes more passes. */
typedef enum kc_counter {
    KC_COUNTER_BLOOM = 0,
    KC_COUNTER_EXACT = 1,
    KC_COUNTER_SORT = 2      /* the exact counter's split from a sort: each pass of a bucket's keys is
                                sorted in LDS (stable LSD radix over the bits that vary) and counted by
                                a run-length reduce (sort_and_export_leaf, src/NewKmerTree.h:1170-1250) */
} kc_counter;

typedef struct kc_config {
    uint32_t k;               /* k-mer length, 1..128 */
    uint32_t partition;       /* kc_partition */
    uint32_t min_count;       /* root record filter on the repeat count, inclusive */
    uint32_t max_count;       /*   (append_export_record, src/NewKmerTree.h:1159-1168); 0 = UINT32_MAX */
    uint64_t expected_bytes;  /* hint: total sequence bytes to be pushed (0 = grow on demand) */
    uint32_t sub_bits;        /* fine fan-out bits below the root prefix: 0 = automatic, 1..10,
                                 or KC_SUB_BITS_NONE for no sub-partition */
    uint32_t device;          /* HIP device ordinal */
    uint32_t classifier_num;  /* LPT owner count reported by kc_plan (reference: (n-2-n/8)/2) */
    uint32_t pass_keys;       /* Bloom counter: keys per counting pass of a bucket (0 = default; larger
                                 buckets take several hash-range passes).  Exact counter: distinct keys
                                 a pass aims to hold (0 = the table size) */
    uint32_t counter;         /* kc_counter */
    uint32_t table_slots;     /* exact counter: count-table slots per bucket and pass (0 = the whole
                                 LDS table, 10240 for k <= 32 and 6144 above); sort counter: keys a
                                 pass stages (0 = 8192 / N); smaller values narrow each pass and are
                                 meant for tuning and tests */
    uint32_t heavy_keys;      /* a bucket of more keys than this (0 = 1Mi) is counted in chunks by the
                                 whole GPU instead of one workgroup; its keys are written as root
                                 records with their whole count (no low.bin entry) */
    uint32_t learned_min_skew_pct; /* KC_PARTITION_LEARNED: the network is trained only when the largest
                                 of the Stage-0 sample's 4096 12-bit key prefixes holds more than this
                                 percentage of their mean distinct keys (0 = 250).  Canonical keys of a
                                 uniform genome have density 2(1 - x) over the key range, so their
                                 prefixes reach about 200% with no skew of the input; at or below the
                                 threshold the run keeps the prefix partition (kc_stats.learned_fallback) */
    uint64_t max_round_keys;  /* keys one pass over the device holds (0 = as many as the free device
                                 memory allows).  An input with more positions is counted in
                                 bucket-range rounds: contiguous root-bucket ranges, one pass each;
                                 the fewest rounds that fit, with ranges as even as the root buckets
                                 allow so the key buffers leave room for the rounds' outputs */
    uint32_t graph;           /* 1: capture one pass of the pipeline in a HIP graph and replay it
                                 (one-round runs; the graph is captured again when buffers or
                                 options change); the stats then carry the total time only */
    uint32_t round_outputs;   /* bucket-range rounds: 0 = each round's outputs stay in device memory
                              (one buffer sized from the first round) while they fit, the rest go
                              to host memory; 1 = all to host memory */
    uint64_t est_file_size;   /* Stage 0: the reference's estimated FASTQ size, input of the Bloom
                                 sizing (FastqPreReader::get_estimated_raw_fastq_file_size: the size
                                 on disk, four times it for gzip; src/main.cpp:88-108); 0 = expected_bytes,
                                 or the sample size when that is 0 too */
    uint32_t learned_max_fine; /* KC_PARTITION_LEARNED: fine buckets of one root bucket, at most (0 = 1024,
                                 the largest).  Each root takes its share of the fine buckets by the
                                 learned CDF's mass on its key range */
    uint32_t learned_steps;   /* KC_PARTITION_LEARNED: training steps of the network in kc_plan (0 = 1000;
                                 8000 balance the 4096 learned buckets of the sample to about 1.15 of
                                 their mean, 1000 to about 3-5, with the same counting speed) */
} kc_config;

typedef struct kc_stats {
    uint32_t k;                  /* k-mer length of the handle */
    uint32_t key_words;          /* N: u64 words per key */
    uint64_t total_kmers;        /* K: every window of k valid bases, counted once */
    uint64_t firsts;             /* keys reported first by the Bloom filter (low.bin) */
    uint64_t records;            /* root records after the min/max filter */
    uint64_t records_unfiltered; /* distinct keys seen again (before the filter) */
    uint64_t spilled_keys;       /* distinct keys whose count lived in the global spill table */
    uint64_t sequence_bytes;     /* bytes resident on the device (blocks + separators) */
    uint64_t fine_buckets;       /* work units of the counting kernel */
    uint64_t max_bucket_keys;    /* largest fine bucket, in keys */
    double mean_bucket_keys;     /* K / fine_buckets */
    double ms_extract_count;     /* device time per phase of the last kc_finish/kc_count_resident */
    double ms_extract_scatter;
    double ms_sub_count;
    double ms_sub_scatter;
    double ms_count;             /* the counting kernel; heavy-bucket chunks are ms_heavy */
    double ms_compact;           /* spill emit and finalize */
    double ms_total;             /* first kernel start to last kernel end */
    double ms_h2d;               /* host -> device sequence copies (pushes) */
    uint64_t stage0_sample_kmers;
    uint64_t prefix_counts[256]; /* Stage-0 root prefix histogram, after the >= 1000 clamp */
    uint64_t bloom_capacity[256];/* reference Bloom sizing (words) from the Stage-0 sample */
    uint8_t prefix_owners[256];  /* LPT owner of each root prefix */
    /* learned partitioner (kc_plan with KC_PARTITION_LEARNED) */
    uint64_t learned_train_keys;        /* sorted Stage-0 keys the model was trained on */
    double learned_train_ms;            /* sample extraction + training + boundary derivation */
    double learned_loss_first;          /* training loss at the first and last step */
    double learned_loss_last;
    double learned_sample_max_over_mean;/* largest of the 4096 learned buckets / mean, on the sample */
    double prefix_sample_max_over_mean; /* largest of the 4096 12-bit key prefixes / mean, on the sample */
    double learned_sample_distinct_max_over_mean; /* the same two, counting distinct keys per bucket */
    double prefix_sample_distinct_max_over_mean;
    uint64_t extra_passes;              /* counting passes beyond one per bucket, summed over buckets */
    uint64_t heavy_buckets;             /* buckets counted in chunks by the whole GPU */
    uint64_t rounds;                    /* bucket-range rounds of the last run (1 = one pass) */
    uint64_t learned_split_buckets;     /* learned buckets the network left empty, refilled from a merged neighbour */
    double bucket4096_max_over_mean;    /* keys of the largest of 4096 bucket groups over the mean, on the
                                           counted input: the 4096 learned buckets, or the 4096 12-bit key
                                           prefixes with the prefix partition (0 when the fan-out is below
                                           4 bits) */
    uint64_t hist_tail_keys;            /* last kc_histogram: keys whose count is above max_freq (not in it) */
    uint64_t hist_tail_kmers;           /*   and the k-mers they hold (the sum of their counts) */
    uint64_t learned_fallback;          /* kc_plan with KC_PARTITION_LEARNED kept the prefix partition: 1 the
                                           sample's prefixes were balanced (learned_min_skew_pct), 2 the sample
                                           had fewer distinct keys than buckets */
    double ms_heavy;                    /* device time of the heavy-bucket chunks (k_heavy); ms_count is the
                                           counting kernel alone */
    uint64_t rounds_in_device;          /* bucket-range rounds whose outputs stayed in device memory */
    uint64_t round_key_cap;             /* keys the key buffers held in the last run */
} kc_stats;

typedef struct kc_engine kc_engine;

/* Records callback: `records` holds n ExportRecord<N> structs (N key words, u32
 * count, zero padding to 8 bytes: 16 bytes at N=1, 24 at N=2). */
typedef int (*kc_records_cb)(void* user, const void* records, uint64_t n);
/* Count histogram of the counted results, without writing low.bin and the
 * root files (replaces `histogram_tool precise|approximate`,
 * tool/histogram_tool.cpp:445-527 and :551-592).  hist[f - min_freq] for f in
 * [min_freq, max_freq], as the tool writes it (tool/histogram_tool.cpp:425-442):
 * PRECISE counts every k-mer at its merged count (with the tool's rule that a
 * record of count r < min_freq - 1 is skipped); APPROXIMATE counts the root
 * records only. */
typedef enum kc_histogram_mode {
    KC_HISTOGRAM_PRECISE = 0,
    KC_HISTOGRAM_APPROXIMATE = 1
} kc_histogram_mode;
KC_API int kc_histogram(kc_engine* h, uint32_t mode, uint32_t min_freq, uint32_t max_freq, int64_t* hist);

/* Replaces FastqReader<N>::read's line state machine (src/FastqReader.h:140-308):
 * `data` is the next `len` bytes of a FASTQ file, cut anywhere (a record or a
 * line may span calls).  The bytes are copied to the device as they are and
 * the sequence lines (line index mod 4 == 1 from the first byte of the first
 * call), each with its '\n', are appended to the resident sequence by device
 * kernels.  Whole lines are kept together, which gives the k-mers of the
 * reference's 32 KiB blocks for every k except 96, whose keys carry the
 * bases before the window and so depend on where the reference cuts its
 * blocks: k = 96 returns KC_ERR_ARG (push the host reader's blocks instead).
 * kc_push_block after a FASTQ chunk that ends inside a record returns
 * KC_ERR_STATE. */
KC_API int kc_push_fastq(kc_engine* h, const char* data, uint64_t len);

/* Keys callback: `keys` holds n keys of N u64 words each. */
typedef int (*kc_keys_cb)(void* user, const uint64_t* keys, uint64_t n);

KC_API int kc_create(const kc_config* cfg, kc_engine** out);
KC_API int kc_sample_block(kc_engine* h, const char* data, uint64_t len);
KC_API int kc_plan(kc_engine* h);

/* The learned partition's boundaries, no reference counterpart: 4097
 * nondecreasing words on key word 0 (bucket j holds the keys whose word 0 w
 * has b[j] <= w < b[j+1]; b[0] = 0, b[4096] = ~0).
 * kc_get_partition_bounds copies them into bounds (cap >= 4097) and sets *n to
 * 4097, or sets *n to 0 when the handle has no learned boundaries (no kc_plan,
 * a prefix handle, or a plan that kept the prefix partition).
 * kc_set_partition_bounds makes them the partition of a KC_PARTITION_LEARNED
 * handle without training, so a partition trained once serves later jobs on
 * like input; the next kc_plan replaces them.  Any valid boundaries count
 * exactly; only the bucket balance depends on them.  KC_ERR_ARG for n != 4097,
 * b[0] != 0, b[4096] != ~0 or a decreasing pair; KC_ERR_STATE on a prefix
 * handle. */
KC_API int kc_get_partition_bounds(kc_engine* h, uint64_t* bounds, uint32_t cap, uint32_t* n);
KC_API int kc_set_partition_bounds(kc_engine* h, const uint64_t* bounds, uint32_t n);
KC_API int kc_push_block(kc_engine* h, const char* data, uint64_t len);
KC_API int kc_finish(kc_engine* h);
KC_API int kc_emit_firsts(kc_engine* h, kc_keys_cb cb, void* user);
KC_API int kc_emit_root(kc_engine* h, uint32_t root_id, kc_records_cb cb, void* user);
KC_API int kc_stats_get(kc_engine* h, kc_stats* out);
KC_API int kc_destroy(kc_engine* h);

/* Re-runs the whole device pipeline on the sequence already resident in HBM
 * (no host transfer inside).  Used to time the counting path with inputs in
 * HBM; results replace those of the previous run.  No reference counterpart:
 * the reference cannot separate ingest from counting. */
KC_API int kc_count_resident(kc_engine* h);

/* Autotuner (no counterpart in the reference; its CPU pipeline has no such
 * parameters): times one pass over the resident input for each candidate --
 * the Bloom, exact and sort counters, each with its automatic fine fan-out and
 * one bit less (tile and bucket size), two heavy-bucket thresholds -- keeps the
 * fastest in the handle's configuration and leaves its results in place.
 * A candidate that does not fit the device (KC_ERR_NOMEM or KC_ERR_CAPACITY,
 * e.g. a spill table that cannot grow) is skipped and counted in `skipped`;
 * the call fails only when no candidate fits. */
typedef struct kc_tune_result {
    uint32_t sub_bits;    /* chosen fine fan-out bits */
    uint32_t counter;     /* chosen kc_counter */
    uint32_t heavy_keys;  /* chosen heavy-bucket threshold */
    uint32_t candidates;  /* configurations timed */
    double best_ms;       /* device time of the chosen one */
    uint32_t learned_max_fine; /* chosen cap on a root's fine buckets (learned partition; 0 = the default) */
    uint32_t skipped;     /* candidates skipped because they did not fit the device (KC_ERR_NOMEM / KC_ERR_CAPACITY) */
} kc_tune_result;
KC_API int kc_autotune(kc_engine* h, kc_tune_result* out);

/* kc_count_noalloc: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): one pass of the counting pipeline of
 * kc_finish / kc_count_resident over a sequence already in device memory
 * (`d_seq`, `seq_len` bytes, the same '\n'-separated sequence lines
 * kc_push_block takes), written to the caller's device buffers, with no
 * allocation at all: every work buffer is carved out of one static 64 MiB
 * device array that each GPU the library runs on holds until the process
 * ends.  Limits: seq_len <= kc_noalloc_max_bytes(k) (2 MiB at k <= 32, 1 MiB
 * at k <= 64, 512 KiB above; larger inputs return KC_ERR_CAPACITY), the prefix
 * partition over the 256 root buckets (partition PREFIX, sub_bits 0 or
 * NONE), and cfg's k, device, min_count, max_count, counter, pass_keys and
 * table_slots; the other fields are ignored.  Outputs: `d_firsts` receives
 * res->firsts keys (N words each, the low.bin key set), `d_records` res->records
 * ExportRecord<N> structs grouped by root: the records of root r are
 * [root_off[0][r], root_off[0][r+1]) followed by [root_off[1][r],
 * root_off[1][r+1]).  firsts_cap / records_cap are the keys / records those
 * buffers hold (every position yields at most one key); too small returns
 * KC_ERR_CAPACITY with the needed counts in *res.  `stream` is a hipStream_t
 * (NULL: the default stream); the call returns after the pass has finished.
 * Calls on one GPU share its array and keep the last call's contents; they
 * are serialized by a per-device lock inside one process.  Errors: the codes
 * of kc_status, with the text in kc_last_error(NULL) on the calling thread. */
typedef struct kc_noalloc_result {
    uint64_t total_kmers;        /* K of the input */
    uint64_t firsts;             /* keys written to d_firsts */
    uint64_t records;            /* records written to d_records (after the min/max filter) */
    uint64_t records_unfiltered; /* distinct keys seen again (before the filter) */
    uint64_t spilled_keys;       /* distinct keys whose count lived in the spill table */
    uint64_t root_off[2][257];   /* root ranges of the records, see above */
} kc_noalloc_result;
KC_API int kc_count_noalloc(const kc_config* cfg, const char* d_seq, uint64_t seq_len, uint64_t* d_firsts,
                            uint64_t firsts_cap, void* d_records, uint64_t records_cap, kc_noalloc_result* res,
                            void* stream);
/* Largest seq_len kc_count_noalloc takes for k (0 if k is out of range). */
KC_API uint64_t kc_noalloc_max_bytes(uint32_t k);

/* Free and total device memory of a device in bytes (hipMemGetInfo); for
 * choosing max_round_keys, and for tests that check nothing stays allocated.
 * No reference counterpart. */
KC_API int kc_device_memory(uint32_t device, uint64_t* free_bytes, uint64_t* total_bytes);

/* Number of u64 words per key for k (1, 2 or 4), 0 if k is 0 or above 128. */
KC_API uint32_t kc_words_for_k(uint32_t k);

/* Human-readable text of the last error on this handle (never NULL); with
 * h == NULL, of the calling thread's last kc_count_noalloc. */
KC_API const char* kc_last_error(kc_engine* h);

/* Test hook, no reference counterpart: copies up to `cap` bytes of the
 * resident sequence (what the pushed blocks or FASTQ chunks left in HBM after
 * kc_finish) to `out`; *len = its full length. */
KC_API int kc_debug_resident(kc_engine* h, char* out, uint64_t cap, uint64_t* len);

/* kc_debug_copy: one plain device copy of `bytes` (a multiple of 16) with a
 * 16-byte-per-lane kernel; *ms = its device time.  Reads `bytes` and writes
 * `bytes` of HBM: the known-size reference for hardware traffic counters. */
KC_API int kc_debug_copy(uint64_t bytes, uint32_t device, float* ms);

/* Fault hook for tests, no reference counterpart (the reference exits on the
 * same failures: src/FastqReader.h:78-90, src/ExportWriter.h:595-599, or throws
 * std::bad_alloc, src/BloomFilter.h:45-48).  kind 1: device and pinned-host
 * allocations of the engine; kind 2: its host<->device and device copies.
 * The nth such call from now on fails once (nth <= 0: none), as if the runtime
 * had refused it; the entry point that made it returns KC_ERR_NOMEM or
 * KC_ERR_HIP, frees what it allocated and leaves the handle usable.  Returns
 * the calls of that kind made since the hook was last set (-1: bad kind).
 * KMC_FAULT_ALLOC=n / KMC_FAULT_COPY=n in the environment set it at the first
 * kc_create of the process. */
KC_API int64_t kc_debug_fault(uint32_t kind, int64_t nth);

/* Test hooks of the learned partitioner, no reference counterpart.
 * kc_debug_gemm_bf16: C[M][N] = A[M][K] . Bt[N][K]^T on the MFMA product
 *   (A, Bt as bf16 bit patterns; M, N multiples of 64, K of 16).
 * kc_debug_mlp_grad: loss and gradient (fp32, parameter layout of the model)
 *   of a batch (nb a multiple of 64) with the device kernels. */
KC_API int kc_debug_gemm_bf16(const uint16_t* A, const uint16_t* Bt, float* C, uint32_t M, uint32_t N, uint32_t K);
KC_API int kc_debug_mlp_grad(const float* params, const uint64_t* keys, const uint32_t* labels, uint32_t nb,
                             float* grad, float* loss);
/* kc_debug_learned_fit: train on n sorted key words (steps, lr, batch, seed as
 * in kc_plan) and return the 4097 bucket boundaries and the loss every 50 steps.
 * flags bit 0: train on the distinct keys; bit 1: boundaries from the expected
 * bucket instead of the most likely one. */
KC_API int kc_debug_learned_fit(const uint64_t* sorted, uint64_t n, uint32_t steps, float lr, uint32_t batch,
                                uint64_t seed, uint32_t flags, uint64_t* bounds, float* loss_curve, uint32_t loss_cap);
/* kc_debug_learned_routing: the fine buckets a learned partition routes by
 * (host only, no device call).  From the 4097 learned boundaries (bounds[0] =
 * 0, bounds[4096] = UINT64_MAX, nondecreasing): root r (the keys' top byte)
 * gets fine buckets [fine_base[r], fine_base[r+1]) (257 entries, 1..1024 per
 * root, about nf_target in all) bounded by fine_bounds[fine_base[r] - r ...]
 * (fine_base[256] - 256 of them, cap at least that); lut: 256 x 4096 cells,
 * cell c of root r = number of r's bounds <= (r << 56) + (c << 44). */
KC_API int kc_debug_learned_routing(const uint64_t* bounds, uint32_t nf_target, uint32_t* fine_base,
                                    uint64_t* fine_bounds, uint64_t cap, uint16_t* lut);

#ifdef __cplusplus
}
#endif

#endif /* KMC_H */
