// kmc_kernels.h -- launch interface of the counting kernels (host <-> device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_device.h"

namespace kmc {

constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, workgroup i -> XCD i % 8

constexpr unsigned long long kErrSpillFull = 1ULL;
constexpr unsigned long long kErrRecFull = 2ULL;

// P0/P1: one 16-byte word of sequence per thread, kThreads*16 positions per tile.
template <int NW> struct ExtractCfg;
template <> struct ExtractCfg<1> { static constexpr int kThreads = 512; };
template <> struct ExtractCfg<2> { static constexpr int kThreads = 256; };

// P2: kTile keys of one coarse bucket per workgroup.
template <int NW> struct SubCfg;
template <> struct SubCfg<1> { static constexpr int kThreads = 512; static constexpr int kTile = 8192; };
template <> struct SubCfg<2> { static constexpr int kThreads = 256; static constexpr int kTile = 4096; };

// P3: one workgroup per fine bucket.  LDS: Bloom words + count table + firsts buffer.
template <int NW> struct CountCfg;
template <> struct CountCfg<1> {
    static constexpr int kThreads = 512;
    static constexpr uint32_t kBloomWords = 2048;
    static constexpr uint32_t kTableSlots = 4096;
    static constexpr uint32_t kFirstBuf = 1024;
    static constexpr uint32_t kProbe = 64;
};
template <> struct CountCfg<2> {
    static constexpr int kThreads = 512;
    static constexpr uint32_t kBloomWords = 2048;
    static constexpr uint32_t kTableSlots = 2048;
    static constexpr uint32_t kFirstBuf = 1024;
    static constexpr uint32_t kProbe = 64;
};

// ExportRecord<N> size (src/definition.h:159-163): N key words + u32 count, padded to 8
template <int NW> struct RecBytes { static constexpr uint32_t value = 8 * NW + 8; };

struct CountArgs {
    uint64_t* firsts;                 // first-occurrence keys, appended
    unsigned long long* first_count;
    uint8_t* recs;                    // ExportRecord<N> array, appended
    unsigned long long* rec_count;
    uint64_t rec_cap;
    unsigned long long* spill_keys;   // NW x spill_cap words (structure of arrays)
    uint32_t* spill_cnt;
    uint64_t spill_cap;               // power of two
    unsigned long long* spilled;      // distinct keys claimed in the spill table
    unsigned long long* error_flags;
    unsigned long long* max_bucket;
    unsigned long long* distinct_rep; // distinct repeated keys (before the min/max filter)
    uint32_t min_count;
    uint32_t max_count;
};

template <int NW>
hipError_t launch_extract(bool scatter, const uint8_t* seq, uint64_t padded_len, uint32_t k, const Route& rt,
                          unsigned long long* counts, unsigned long long* cursor, uint64_t* out, hipStream_t st);
hipError_t launch_scan_coarse(const unsigned long long* counts, unsigned long long* coarse_off, unsigned long long* cursor,
                              unsigned long long* tile_start, uint32_t sub_tile, hipStream_t st);
hipError_t launch_scan_fine(const unsigned long long* counts, uint32_t nf, unsigned long long* fine_off,
                            unsigned long long* cursor, hipStream_t st);
template <int NW>
hipError_t launch_sub(bool scatter, uint64_t max_tiles, const uint64_t* A, const unsigned long long* coarse_off,
                      const unsigned long long* tile_start, const Route& rt, unsigned long long* counts,
                      unsigned long long* cursor, uint64_t* B, hipStream_t st);
template <int NW>
hipError_t launch_count(uint32_t nf, const uint64_t* B, const unsigned long long* fine_off, const CountArgs& args,
                        hipStream_t st);
template <int NW>
hipError_t launch_spill_emit(const CountArgs& args, hipStream_t st);
template <int NW>
hipError_t launch_group_records(bool scatter, const uint8_t* recs, const unsigned long long* n_ptr, uint64_t cap,
                                unsigned long long* root_counts, unsigned long long* root_cursor, uint8_t* out,
                                hipStream_t st);
hipError_t launch_rec_scan(const unsigned long long* root_counts, unsigned long long* root_off,
                           unsigned long long* root_cursor, hipStream_t st);
template <int NW>
hipError_t launch_debug_hash(const uint64_t* keys, uint64_t n, uint64_t* out, hipStream_t st);

}  // namespace kmc
