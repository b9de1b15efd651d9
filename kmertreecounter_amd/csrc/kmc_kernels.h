// kmc_kernels.h -- launch interface of the counting kernels (host <-> device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_device.h"

namespace kmc {

constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, workgroup i -> XCD i % 8

constexpr unsigned long long kErrSpillFull = 1ULL;
constexpr unsigned long long kErrRecFull = 2ULL;  // overflow record list full
constexpr unsigned long long kErrRecOut = 4ULL;   // records exceed the final record buffer
constexpr unsigned long long kErrHistFull = 8ULL; // histogram record table: a probe ran out
constexpr unsigned long long kErrFirstOut = 16ULL; // first occurrences exceed the final firsts buffer

// P0/P1: one 16-byte word of sequence per thread, kThreads*16 positions per tile.
// kStageRounds: the scatter stages its sorted tile in LDS in this many rounds
// (two halve its LDS, 36 KB instead of 70, so a third workgroup fits on a CU);
// kWaves: waves per SIMD its register allocation must allow (k <= 32: 6, at
// 80 VGPRs; k <= 64: 3, at 148; 8 waves would spill)
template <int NW> struct ExtractCfg;
template <> struct ExtractCfg<1> {
    static constexpr int kThreads = 512;
    static constexpr int kStageRounds = 2;
    static constexpr int kWaves = 6;
};
template <> struct ExtractCfg<2> { static constexpr int kThreads = 256; static constexpr int kStageRounds = 2; static constexpr int kWaves = 3; };
// 256 threads at least: the bucket histogram and its scan take one thread per bucket
template <> struct ExtractCfg<4> { static constexpr int kThreads = 256; static constexpr int kStageRounds = 2; static constexpr int kWaves = 1; };  // (1: no limit)

// P2: kTile keys of one root bucket per tile; the scatter (k_sub) takes one
// tile per workgroup of kThreads, the count pass (k_sub_count) several per
// workgroup of kCountThreads.  A root has up to kMaxFine fine buckets.
template <int NW> struct SubCfg;
template <> struct SubCfg<1> { static constexpr int kThreads = 1024; static constexpr int kCountThreads = 512; static constexpr int kTile = 8192; };
template <> struct SubCfg<2> { static constexpr int kThreads = 1024; static constexpr int kCountThreads = 512; static constexpr int kTile = 4096; };
template <> struct SubCfg<4> { static constexpr int kThreads = 1024; static constexpr int kCountThreads = 512; static constexpr int kTile = 2048; };
constexpr uint32_t kMaxSubBits = 10;

// P3: one workgroup per fine bucket (1 per CU).  LDS: Bloom words + count
// table; buckets above kPassKeys keys are processed in hash-range passes.
template <int NW> struct CountCfg;
// Bloom counter: one 1024-thread workgroup per fine bucket, two workgroups per
// CU (8 waves per SIMD, so at most 64 VGPRs and 80 KiB of LDS each): a 16 KiB
// filter and a count table of 5120 (k <= 32) or 3072 (k <= 64) slots.  The
// kernel is bound by the latency of its LDS and memory round trips: with one
// workgroup per CU (a 32 KiB filter and twice the table) it ran 25-30% slower
// in spite of needing fewer hash-range passes.  A pass aims at about 70% table
// load; kPassKeys bounds the occurrences of a bucket's first pass.
template <> struct CountCfg<1> {
    static constexpr int kThreads = 1024;
    static constexpr int kWaves = 8;  // waves per SIMD
    static constexpr uint32_t kBloomWords = 4096;  // 32-bit filter words (16 KiB)
    static constexpr uint32_t kTableSlots = 5120;
    static constexpr uint32_t kPassKeys = 73728;
    static constexpr uint32_t kPassDistinct = 3600;  // target distinct keys per pass
    static constexpr uint32_t kProbe = 64;
    static constexpr int kLoads = 2;  // key loads in flight per thread
    static constexpr uint32_t kExactSlots = 10240;  // exact counter: 120 KiB of keys + counts
    static constexpr int kExactLoads = 4;  // exact counter and k_heavy (one workgroup per CU)
};
template <> struct CountCfg<2> {
    static constexpr int kThreads = 1024;
    static constexpr int kWaves = 8;
    static constexpr uint32_t kBloomWords = 4096;  // 32-bit filter words (16 KiB)
    static constexpr uint32_t kTableSlots = 3072;
    static constexpr uint32_t kPassKeys = 30720;
    static constexpr uint32_t kPassDistinct = 2150;
    static constexpr uint32_t kProbe = 64;
    static constexpr int kLoads = 1;
    static constexpr uint32_t kExactSlots = 6144;
    static constexpr int kExactLoads = 4;
};
// k 65..128: a table slot is five 63-bit words plus its count (44 bytes); one
// workgroup per CU
template <> struct CountCfg<4> {
    static constexpr int kThreads = 1024;
    static constexpr int kWaves = 4;
    static constexpr uint32_t kBloomWords = 8192;  // 32-bit filter words (32 KiB)
    static constexpr uint32_t kTableSlots = 2048;
    static constexpr uint32_t kPassKeys = 24576;
    static constexpr uint32_t kPassDistinct = 1250;
    static constexpr uint32_t kProbe = 64;
    static constexpr int kLoads = 3;
    static constexpr uint32_t kExactSlots = 2560;
    static constexpr int kExactLoads = 4;
};

// Sort counter: one 1024-thread workgroup per fine bucket; a pass stages kKeys
// keys (two LDS buffers of kKeys keys for the radix sort: 128 KiB) and sorts
// them by 6-bit digits.
constexpr int kSortThreads = 1024;
constexpr uint32_t kSortDigitBits = 6;
constexpr uint32_t kSortDigits = 1u << kSortDigitBits;
template <int NW> struct SortCfg { static constexpr uint32_t kKeys = 8192 / NW; };

// probes into the global spill table before a key is declared unplaceable
constexpr uint32_t kSpillProbe = 1024;

// ExportRecord<N> size (src/definition.h:159-163): N key words + u32 count, padded to 8
template <int NW> struct RecBytes { static constexpr uint32_t value = 8 * NW + 8; };

// device counters (d_ctr)
enum : int { kCtrFirst = 0, kCtrRec, kCtrSpilled, kCtrErr, kCtrMaxBucket, kCtrDistinct, kCtrOvf, kCtrSpillDistinct,
             kCtrExtraPasses, kCtrHeavyChunks, kCtrHeavyNext, kCtrHeavyBuckets, kCtrCount };

struct CountArgs {
    uint64_t* firsts;                   // bucket f: keys at [fine_off[f], fine_off[f+1])
    uint8_t* rec_tmp;                   // bucket f: records at byte fine_off[f]*8*NW, n_f*8*NW bytes
    uint32_t* nfirst;                   // [nf]
    uint32_t* nrec;                     // [nf] records kept in the bucket's slice
    uint32_t* ndist;                    // [nf] distinct repeated keys in LDS tables (before the filter)
    uint8_t* ovf;                       // overflow records (slice full, spilled keys)
    unsigned long long* ovf_count;
    uint64_t ovf_cap;
    unsigned long long* spill_keys;     // NW x spill_cap words (structure of arrays)
    uint32_t* spill_cnt;
    uint64_t spill_cap;                 // power of two
    unsigned long long* spilled;        // distinct keys claimed in the spill table
    unsigned long long* spill_distinct; // spilled keys found by k_spill_emit
    unsigned long long* error_flags;
    uint32_t min_count;
    uint32_t max_count;
    uint64_t pass_keys;
    unsigned long long* extra_passes;   // passes beyond the first, summed over buckets
    uint32_t table_slots;               // exact counter: LDS slots used (0 = all)
    // Buckets of more than heavy_keys keys are not counted by one workgroup:
    // their workgroup lists heavy_chunk-key chunks of them and k_heavy counts
    // the chunks across the whole GPU into the global table.
    uint64_t heavy_keys;                // 0 = never
    uint32_t heavy_chunk;
    unsigned long long* heavy_list;     // [heavy_cap] x (lo, hi) key ranges
    uint64_t heavy_cap;
    unsigned long long* heavy_count;    // chunks listed
    unsigned long long* heavy_next;     // chunks taken by k_heavy
    unsigned long long* heavy_buckets;  // buckets handed over
    const uint32_t* order;              // [nf] bucket of each workgroup, largest first (null: blockIdx)
};

struct SegArgs {
    uint32_t nf;
    const unsigned long long* fine_off;
    const uint32_t* nfirst;
    const uint32_t* nrec;
    const uint32_t* ndist;
    unsigned long long* foff;           // [nf+1]
    unsigned long long* roff;           // [nf+1]
    unsigned long long* root_ovf;       // [256] overflow records per root
    unsigned long long* root_main;      // [256] in-bucket records per root
    unsigned long long* root_off;       // [257] root ranges of the in-bucket records
    unsigned long long* ovf_off;        // [257] root ranges of the overflow records (after all in-bucket ones)
    unsigned long long* ovf_cursor;     // [256]
    unsigned long long* ctr;
    uint64_t rec_out_cap;               // records the final record buffer holds
    uint64_t firsts_out_cap;            // keys the final firsts buffer holds
    unsigned long long* part;           // [2 x ceil(nf / 1024)] firsts and records per chunk of 1024 buckets
};

template <int NW>
hipError_t launch_extract(bool scatter, const uint8_t* seq, uint64_t padded_len, uint32_t k, const Route& rt,
                          unsigned long long* counts, unsigned long long* cursor, uint64_t* out, hipStream_t st);
hipError_t launch_scan_coarse(const unsigned long long* counts, unsigned long long* coarse_off, unsigned long long* cursor,
                              unsigned long long* tile_start, uint32_t sub_tile, hipStream_t st,
                              uint32_t rlo = 0, uint32_t rhi = 256);
// fine bucket offsets and cursors; part: kScanGroups scratch words
constexpr uint32_t kScanGroups = 256;  // up to 262144 fine buckets (sbits 10)
hipError_t launch_scan_fine(const unsigned long long* counts, uint32_t nf, unsigned long long* fine_off,
                            unsigned long long* cursor, unsigned long long* part, hipStream_t st);
template <int NW>
hipError_t launch_sub(bool scatter, uint64_t max_tiles, const uint64_t* A, const unsigned long long* coarse_off,
                      const unsigned long long* tile_start, const Route& rt, unsigned long long* counts,
                      unsigned long long* cursor, uint64_t* B, hipStream_t st);
// counter: the kc_counter values
enum : uint32_t { kCounterBloom = 0, kCounterExact = 1, kCounterSort = 2 };
// order[nf]: the counting kernels' bucket order (args.order), largest first
// order[0..nf): the fine buckets by size class, largest first; scratch: 512
// zeroed 32-bit words
hipError_t launch_bucket_order(uint32_t nf, const unsigned long long* fine_off, uint32_t* order, uint32_t* scratch,
                               hipStream_t st);
template <int NW>
hipError_t launch_count(uint32_t nf, uint32_t counter, const uint64_t* B, const unsigned long long* fine_off,
                        const CountArgs& args, hipStream_t st);
template <int NW>
hipError_t launch_heavy(uint32_t workgroups, const uint64_t* B, const CountArgs& args, hipStream_t st);
template <int NW>
hipError_t launch_spill_emit(const CountArgs& args, hipStream_t st);
template <int NW>
hipError_t launch_finalize(const SegArgs& sa, const CountArgs& ca, const uint64_t* firsts_in, uint64_t* firsts_out,
                           const uint8_t* rec_in, uint8_t* rec_out, hipStream_t st);
hipError_t launch_debug_copy(const void* src, void* dst, uint64_t n16, hipStream_t st);
// bytes of p set to v by kernels (the captured pipeline graph has no memset nodes)
hipError_t launch_fill(void* p, uint64_t bytes, uint8_t v, hipStream_t st);

// FASTQ chunk -> sequence lines appended at state[0] (state[1]: line phase at
// the chunk's first byte); scratch of fastq_tile_scratch_bytes(n) bytes.
uint64_t fastq_tile_scratch_bytes(uint64_t max_chunk);
hipError_t launch_fastq_parse(const uint8_t* raw, uint64_t n, void* scratch, unsigned long long* state, uint8_t* seq,
                              hipStream_t st);

// Count histogram of the resident results (precise or approximate mode).
constexpr uint32_t kHistProbe = 4096;
struct HistArgs {
    long long* hist;                    // [maxf - minf + 1], zeroed by the caller
    uint32_t minf, maxf;
    uint32_t precise;                   // 1: records and low.bin keys; 0: records only
    unsigned long long* keys;           // NW x cap words (structure of arrays), kEmpty = free
    uint32_t* vals;                     // [cap] record counts
    uint64_t cap;                       // power of two, at least twice the records
    unsigned long long* error_flags;    // kErrHistFull (8) when a probe runs out
    unsigned long long* tail;           // [2]: keys whose merged count is above maxf, and their k-mers
};
template <int NW>
hipError_t launch_histogram(const uint8_t* recs, uint64_t nrec, const uint64_t* firsts, uint64_t nfirst, const HistArgs& a,
                            hipStream_t st);

}  // namespace kmc
