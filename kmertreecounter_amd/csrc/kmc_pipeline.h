// kmc_pipeline.h -- one pass of the counting pipeline over a resident
// sequence, on buffers the caller provides: the engine's own (grown on
// demand) or the static scratch of the no-allocation entry point.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_kernels.h"

namespace kmc {

// small arrays (unsigned long long units): per-XCD coarse counts and cursors,
// coarse offsets, P2 tile starts, root histograms and ranges of the records,
// scan scratch, the counting order's size-class counts and cursors (2 x 256
// 32-bit words), the output scan's chunk sums (2 x 256 words: 2^18 buckets)
constexpr uint64_t kSmCounts = 0, kSmCursor = 8 * 256, kSmCoarseOff = 16 * 256, kSmTileStart = kSmCoarseOff + 257,
                   kSmRootOvf = kSmTileStart + 257, kSmRootMain = kSmRootOvf + 256, kSmRootOff = kSmRootMain + 256,
                   kSmOvfOff = kSmRootOff + 257, kSmOvfCursor = kSmOvfOff + 257, kSmScanPart = kSmOvfCursor + 256,
                   kSmOrder = kSmScanPart + kScanGroups, kSmSegPart = kSmOrder + 256, kSmTotal = kSmSegPart + 512;

// positions per extraction tile, keys per sub-partition tile, bytes per
// ExportRecord<N>, words per key in the open-addressing tables (k > 64 keys
// take five 63-bit words)
inline uint64_t ext_tile(uint32_t nw)
{
    return (nw == 1 ? ExtractCfg<1>::kThreads : nw == 2 ? ExtractCfg<2>::kThreads : ExtractCfg<4>::kThreads) * 16ULL;
}
inline uint64_t sub_tile(uint32_t nw)
{
    return nw == 1 ? SubCfg<1>::kTile : nw == 2 ? SubCfg<2>::kTile : SubCfg<4>::kTile;
}
inline uint64_t rec_bytes(uint32_t nw) { return 8ULL * nw + 8; }
inline uint64_t table_words(uint32_t nw) { return nw == 4 ? (uint64_t)TableWords<4>::value : nw; }

struct PipeBufs {
    uint8_t* seq;                    // padded bytes; [seq_len, padded) is filled with '\n'
    uint64_t seq_len, padded;
    uint64_t* A;                     // key_cap x NW words each
    uint64_t* B;
    uint8_t* R;                      // key_cap x 8NW bytes of per-bucket record slices
    uint64_t key_cap;
    unsigned long long* small;       // kSmTotal words
    unsigned long long* ctr;         // kCtrCount words
    unsigned long long* fcounts;     // kXcds x nf (with a sub-partition)
    unsigned long long* fcursor;     // kXcds x nf (with a sub-partition)
    unsigned long long* fine_off;    // nf + 1 (with a sub-partition)
    unsigned long long* foff;        // nf + 1
    unsigned long long* roff;        // nf + 1
    uint32_t* bucket_u32;            // 4 x nf: nfirst, nrec, ndist per bucket, counting order
    uint8_t* ovf;                    // ovf_cap overflow records
    uint64_t ovf_cap;
    unsigned long long* spill_keys;  // TableWords x spill_cap words
    uint32_t* spill_cnt;             // spill_cap
    uint64_t spill_cap;              // power of two
    bool clear_spill;                // the spill table may hold a previous pass's keys
    unsigned long long* heavy;       // heavy_cap x (lo, hi)
    uint64_t heavy_cap;
    uint64_t* firsts_out;            // final first-occurrence keys (null: the P3 key input)
    uint8_t* recs_out;               // final root-grouped records (null: the per-bucket firsts buffer)
    uint64_t rec_out_cap;            // records recs_out holds
    uint64_t firsts_out_cap;         // keys firsts_out holds
};

struct PipeOpts {
    uint32_t k;
    bool sub;                        // fine buckets below the roots (P2 runs): prefix sbits > 0, or learned
    uint32_t nf;                     // fine buckets (256 without a sub-partition)
    Route rt;
    uint32_t min_count, max_count;   // record filter (append_export_record)
    uint64_t pass_keys;              // 0: the counter's default
    uint32_t counter;                // kc_counter
    uint32_t table_slots;
    uint64_t heavy_keys;
    uint32_t heavy_chunk;
    uint32_t heavy_workgroups;       // persistent k_heavy grid (one per CU)
    // bucket-range rounds: the per-XCD root histogram of the whole input
    // (kXcds x 256, the count pass's own layout), computed once before the
    // rounds; the pass copies it instead of running the count pass again
    const unsigned long long* root_counts;
};

// Where a pass leaves its outputs: firsts, records, and the P3 key input.
template <int NW>
inline uint64_t* pipe_firsts(const PipeBufs& b, bool sub)
{
    return b.firsts_out ? b.firsts_out : (sub ? b.B : b.A);
}
template <int NW>
inline uint8_t* pipe_records(const PipeBufs& b, bool sub)
{
    return b.recs_out ? b.recs_out : reinterpret_cast<uint8_t*>(sub ? b.A : b.B);
}

// last error text of kc_count_noalloc on the calling thread
const char* noalloc_last_error();

// Enqueues one pass on `st`.  graph: being captured (clears are fill kernels,
// no events); otherwise ev[0..9] and ev[11] bracket the phases: P0-P2 kernels
// ev[2p]..ev[2p+1], the counting kernel ev[8]..ev[9], the heavy-bucket chunks
// ev[9]..ev[11], then the spill emit and finalize (ev may be null).
template <int NW>
hipError_t enqueue_pipeline(const PipeBufs& b, const PipeOpts& o, hipStream_t st, bool graph, hipEvent_t* ev);

}  // namespace kmc
