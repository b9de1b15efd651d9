// kmc_plan.hip -- Stage 0 on the GPU: the sampled sequence is run through the
// extraction kernel in count mode to get the 256 root-prefix histogram, then
// the reference's clamp / LPT / Bloom sizing run on the host.
#include "kmc_plan.h"

#include <algorithm>
#include <cstring>
#include <numeric>

#include "kmc_kernels.h"

namespace kmc {

void clamp_prefix_counts(uint64_t* c)
{
    for (int i = 0; i < 256; ++i)
        if (c[i] < 1000) c[i] = 1000;
}

void lpt_assign(const uint64_t* counts, uint32_t classifiers, uint8_t* owners)
{
    // prefixes by count, largest first, each to the currently least-loaded owner
    if (classifiers == 0) classifiers = 1;
    std::vector<uint32_t> order(256);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return counts[a] > counts[b]; });
    std::vector<uint64_t> load(classifiers, 0);
    for (uint32_t p : order) {
        uint32_t best = 0;
        for (uint32_t c = 1; c < classifiers; ++c)
            if (load[c] < load[best]) best = c;
        owners[p] = (uint8_t)best;
        load[best] += counts[p];
    }
}

void bloom_capacity(const uint64_t* counts, uint64_t est_file_size, uint64_t* caps)
{
    // words = max(bit_ceil(file_size/3 * share * 4.81 / 64), 2^15)
    const uint64_t est_total = est_file_size / 3;
    uint64_t total = 0;
    for (int i = 0; i < 256; ++i) total += counts[i];
    for (int i = 0; i < 256; ++i) {
        const double ratio = total ? (double)counts[i] / (double)total : 0.0;
        const uint64_t est = (uint64_t)((double)est_total * ratio * 4.81 / 64);
        uint64_t c = 1;
        while (c < est) c <<= 1;
        caps[i] = std::max<uint64_t>(c, 1ULL << 15);
    }
}

int stage0_plan(uint32_t k, const kc_config& cfg, const std::vector<char>& sample, hipStream_t st,
                PlanResult& out, std::string& err)
{
    out = PlanResult();
    const uint32_t nw = k <= 32 ? 1u : 2u;
    const uint64_t tile = (nw == 1 ? ExtractCfg<1>::kThreads : ExtractCfg<2>::kThreads) * 16ULL;
    const uint64_t padded = ((sample.size() + tile - 1) / tile) * tile;
    if (padded) {
        uint8_t* d_seq = nullptr;
        unsigned long long* d_counts = nullptr;
        auto cleanup = [&]() { hipFree(d_seq); hipFree(d_counts); };
        if (hipMalloc(&d_seq, padded + 256) != hipSuccess || hipMalloc(&d_counts, sizeof(unsigned long long) * kXcds * 256) != hipSuccess) {
            cleanup();
            err = "stage 0: device allocation failed";
            return KC_ERR_NOMEM;
        }
        hipError_t e = hipMemcpyAsync(d_seq, sample.data(), sample.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && padded > sample.size()) e = hipMemsetAsync(d_seq + sample.size(), '\n', padded - sample.size(), st);
        if (e == hipSuccess) e = hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * kXcds * 256, st);
        Route rt{};
        if (e == hipSuccess)
            e = nw == 1 ? launch_extract<1>(false, d_seq, padded, k, rt, d_counts, nullptr, nullptr, st)
                        : launch_extract<2>(false, d_seq, padded, k, rt, d_counts, nullptr, nullptr, st);
        std::vector<unsigned long long> h(kXcds * 256);
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_counts, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        cleanup();
        if (e != hipSuccess) {
            err = std::string("stage 0: ") + hipGetErrorString(e);
            return KC_ERR_HIP;
        }
        for (int x = 0; x < kXcds; ++x)
            for (int b = 0; b < 256; ++b) out.raw_prefix_counts[b] += h[(size_t)x * 256 + b];
    }
    for (int b = 0; b < 256; ++b) {
        out.sample_kmers += out.raw_prefix_counts[b];
        out.prefix_counts[b] = out.raw_prefix_counts[b];
    }
    clamp_prefix_counts(out.prefix_counts);
    lpt_assign(out.prefix_counts, cfg.classifier_num ? cfg.classifier_num : 1, out.owners);
    bloom_capacity(out.prefix_counts, cfg.expected_bytes ? cfg.expected_bytes : sample.size(), out.bloom_capacity);
    return KC_OK;
}

}  // namespace kmc
