/* kmc_host.h -- host side of the drop-in: FASTQ chunking, output files and the
 * `Tree`-equivalent driver, all calling the engine through include/kmc.h.
 */
#ifndef KMC_HOST_H
#define KMC_HOST_H

#include <stdint.h>

#include "kmc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Block sink: one sequence-only block as FastqReader publishes it. */
typedef int (*kmc_block_cb)(void* user, const char* data, uint64_t len);

/* FASTQ bytes -> sequence-only blocks of at most block_size bytes
 * (replaces FastqReader<N>::read, src/FastqReader.h:140-290: 4-line records,
 * only the sequence line kept and '\n'-terminated, a line split at a full
 * block carries its last k-1 bytes -- or the bytes after the block's last
 * '\n', if nearer -- into the next block, :338-365).  `fastq` is the whole
 * input in memory; kmc_fastq_blocks_file streams a file (plain or gzip, the
 * reference's gzip detection, :92-104) in 2 MiB reads.  Returns 0 or the first
 * nonzero callback / I/O status. */
KC_API int kmc_fastq_blocks(const char* fastq, uint64_t len, uint32_t k, uint64_t block_size,
                            kmc_block_cb cb, void* user);
KC_API int kmc_fastq_blocks_file(const char* path, uint32_t k, uint64_t block_size,
                                 uint64_t max_bytes, kmc_block_cb cb, void* user);

/* Stage-0 sample of a FASTQ file as the reference's FastqPreReader<N>::pre_read
 * reads it (src/FastqPreReader.h:128-300): min(size, 256 MiB) bytes, as three
 * record-aligned parts of a plain file or the head of a gzip stream, cut into
 * sequence-only blocks like kmc_fastq_blocks.  *est_file_size (may be NULL)
 * receives the reference's estimated FASTQ size
 * (get_estimated_raw_fastq_file_size: the size on disk, four times it for
 * gzip), the Bloom sizing input of kc_plan. */
KC_API int kmc_stage0_sample_file(const char* path, uint32_t k, uint64_t block_size, kmc_block_cb cb, void* user,
                                  uint64_t* est_file_size);

/* low.bin record packing (src/ExportWriter.h:522-550): per key k/32 raw
 * little-endian words, then the top ceil(2*(k%32)/8) bytes of the masked tail
 * word.  out needs n * kmc_packed_bytes(k) bytes. */
KC_API uint64_t kmc_packed_bytes(uint32_t k);
KC_API void kmc_pack_low(const uint64_t* keys, uint64_t n, uint32_t k, uint8_t* out);

/* One file in the low.bin format (ExportWriter, src/ExportWriter.h:522-550)
 * from n keys in host memory, and one root file (FinalDrainWriter,
 * src/FinalDrainWriter.h:31-91) from n ExportRecord<N> structs; both are
 * written by up to 8 threads, each packing and pwrite-ing its slice at its
 * offset (the writers kmc_write_outputs uses per emitted chunk). */
KC_API int kmc_write_low_keys(const char* path, const uint64_t* keys, uint64_t n, uint32_t k);
KC_API int kmc_write_root_records(const char* path, const void* records, uint64_t n, uint32_t k);

/* Write tmp_dir/low.bin and tmp_dir/root_{0..255}.bin from finished engine
 * results (ExportWriter + FinalDrainWriter formats; all 256 root files are
 * created, empty ones included, as the reference does). */
KC_API int kmc_write_outputs(kc_engine* h, const char* tmp_dir);

/* The `Tree` pipeline (src/main.cpp:111-312) on the GPU engine: Stage 0 on the
 * reference's sample of the file, then the whole file, then outputs in tmp_dir.
 * Fills *stats.  Prints the reference's summary lines when verbose. */
typedef struct kmc_tree_opts {
    const char* fastq_path;
    const char* tmp_dir;
    uint32_t k;
    uint32_t n_thread;      /* reference thread budget; sets the LPT owner count */
    uint32_t min_count;
    uint32_t max_count;     /* 0 = UINT32_MAX */
    uint32_t partition;     /* kc_partition */
    uint32_t verbose;
    uint64_t sample_bytes;  /* Stage-0 sample: 0 = the reference's (kmc_stage0_sample_file); otherwise
                               the first sample_bytes of the file */
    uint32_t counter;       /* kc_counter */
    uint32_t histogram_only; /* 1: skip low.bin and the root files (needs histogram_path) */
    const char* histogram_path; /* if set: write the precise count histogram here, as histogram_tool does */
    uint32_t hist_min;      /* histogram range (0 = 1 and 10000, the tool's defaults) */
    uint32_t hist_max;
    uint32_t host_parse;    /* 1: the host reader's blocks (kc_push_block); 0: raw FASTQ chunks parsed on
                               the device (kc_push_fastq).  k = 96 always uses the host reader. */
} kmc_tree_opts;

KC_API int kmc_tree_run(const kmc_tree_opts* opts, kc_stats* stats, double* wall_ms);

/* Histogram file as histogram_tool writes it (tool/histogram_tool.cpp:425-442):
 * one line `f<TAB>n` for each f in [min_freq, max_freq]; hist has
 * max_freq - min_freq + 1 entries (kc_histogram's output). */
KC_API int kmc_write_histogram(const char* path, const int64_t* hist, uint32_t min_freq, uint32_t max_freq);

#ifdef __cplusplus
}
#endif

#endif /* KMC_HOST_H */
