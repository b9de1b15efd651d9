// kmc_plan.h -- Stage 0: sample keys -> partition plan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/kmc.h"

namespace kmc {

struct PlanResult {
    uint64_t sample_kmers = 0;
    uint64_t prefix_counts[256] = {};   // after the >= 1000 clamp
    uint64_t raw_prefix_counts[256] = {};
    uint64_t bloom_capacity[256] = {};
    uint8_t owners[256] = {};
    uint32_t sbits = 0;
    // learned partition: 4097 bucket boundaries on key word 0 (bucket j is
    // [b[j], b[j+1])); empty for a prefix plan
    std::vector<uint64_t> learned_b;
    std::vector<float> loss_curve;      // training loss every 50 steps
    uint64_t train_keys = 0;            // sorted sample keys the model saw
    uint32_t split_buckets = 0;         // empty learned buckets refilled by split_merged_buckets
    uint32_t fallback = 0;              // learned plan kept the prefix partition: 1 balanced sample, 2 too few keys
    double train_ms = 0.0;
    double sample_max_over_mean_learned = 0.0;  // over the 4096 learned buckets, on the sample
    double sample_max_over_mean_prefix = 0.0;   // over the 4096 12-bit prefixes, on the sample
    double sample_distinct_learned = 0.0;       // the same, counting distinct keys per bucket
    double sample_distinct_prefix = 0.0;
};

// Reference Stage-0 steps restated for the GPU host (src/main.cpp:173-189).
void clamp_prefix_counts(uint64_t* counts256);                                  // :175-184
void lpt_assign(const uint64_t* counts256, uint32_t classifiers, uint8_t* owners); // :38-86
void bloom_capacity(const uint64_t* counts256, uint64_t est_file_size, uint64_t* caps256); // :88-108

int stage0_plan(uint32_t k, const kc_config& cfg, const std::vector<char>& sample, hipStream_t st,
                PlanResult& out, std::string& err);
// Train the CDF model on the sample's keys and fill out.learned_b (and its
// stats).  A sample with fewer distinct keys than buckets leaves learned_b
// empty: the run then uses the prefix partition.
// The network is trained only when the sample's largest 12-bit prefix holds
// more than min_skew times their mean distinct keys.
int learned_plan(uint32_t k, const std::vector<char>& sample, double min_skew, uint32_t train_steps, hipStream_t st,
                 PlanResult& out,
                 std::string& err);

}  // namespace kmc
