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
    // learned partition: 255 coarse boundaries followed by 256 x 255 fine
    // boundaries on key word 0 (empty when the plan is a prefix plan)
    std::vector<uint64_t> bounds;
    double train_loss = 0.0;
};

// Reference Stage-0 steps restated for the GPU host (src/main.cpp:173-189).
void clamp_prefix_counts(uint64_t* counts256);                                  // :175-184
void lpt_assign(const uint64_t* counts256, uint32_t classifiers, uint8_t* owners); // :38-86
void bloom_capacity(const uint64_t* counts256, uint64_t est_file_size, uint64_t* caps256); // :88-108

int stage0_plan(uint32_t k, const kc_config& cfg, const std::vector<char>& sample, hipStream_t st,
                PlanResult& out, std::string& err);

}  // namespace kmc
