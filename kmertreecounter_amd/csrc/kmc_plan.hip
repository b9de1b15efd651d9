// kmc_plan.hip -- Stage 0 on the GPU: the sampled sequence is run through the
// extraction kernel in count mode to get the 256 root-prefix histogram, then
// the reference's clamp / LPT / Bloom sizing run on the host.
#include "kmc_plan.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>

#include <chrono>
#include <cstdio>
#include <cstdlib>

#include <rocprim/device/device_radix_sort.hpp>

#include "kmc_fault.h"
#include "kmc_kernels.h"
#include "kmc_learned.h"

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

namespace {

uint32_t key_words(uint32_t k) { return k <= 32 ? 1u : k <= 64 ? 2u : 4u; }

__device__ __forceinline__ uint64_t splitmix64_dev(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Word 0 of the sample keys whose hash is below keep_below (every key with
// keep_all), appended to out: one counter add per wave.  A key is kept with
// all its occurrences, so the kept multiset does not depend on the order the
// extraction wrote the keys in.
__global__ void __launch_bounds__(256)
k_keep_by_hash(const uint64_t* __restrict__ keys, uint64_t n, uint32_t nw, uint64_t keep_below, uint32_t keep_all,
               uint64_t* __restrict__ out, unsigned long long* __restrict__ count)
{
    const uint32_t lane = threadIdx.x & 63u;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        uint64_t w0 = 0;
        bool keep = false;
        if (i < n) {
            w0 = keys[i * nw];
            keep = keep_all || splitmix64_dev(w0) < keep_below;
        }
        const uint64_t bal = __ballot(keep);
        unsigned long long at = 0;
        if (lane == 0 && bal) at = atomicAdd(count, (unsigned long long)__popcll(bal));
        at = __shfl(at, 0, 64);
        if (keep) out[at + __popcll(bal & ((1ULL << lane) - 1ULL))] = w0;
    }
}

uint64_t extract_tile(uint32_t nw)
{
    return (nw == 1 ? ExtractCfg<1>::kThreads : nw == 2 ? ExtractCfg<2>::kThreads : ExtractCfg<4>::kThreads) * 16ULL;
}

hipError_t extract(uint32_t nw, bool scatter, const uint8_t* seq, uint64_t padded, uint32_t k, const Route& rt,
                   unsigned long long* counts, unsigned long long* cursor, uint64_t* out, hipStream_t st)
{
    switch (nw) {
    case 1: return launch_extract<1>(scatter, seq, padded, k, rt, counts, cursor, out, st);
    case 2: return launch_extract<2>(scatter, seq, padded, k, rt, counts, cursor, out, st);
    default: return launch_extract<4>(scatter, seq, padded, k, rt, counts, cursor, out, st);
    }
}

}  // namespace

int stage0_plan(uint32_t k, const kc_config& cfg, const std::vector<char>& sample, hipStream_t st,
                PlanResult& out, std::string& err)
{
    out = PlanResult();
    const uint32_t nw = key_words(k);
    const uint64_t tile = extract_tile(nw);
    const uint64_t padded = ((sample.size() + tile - 1) / tile) * tile;
    if (padded) {
        uint8_t* d_seq = nullptr;
        unsigned long long* d_counts = nullptr;
        auto cleanup = [&]() { hipFree(d_seq); hipFree(d_counts); };
        if (dmalloc(&d_seq, padded + 256) != hipSuccess || dmalloc(&d_counts, sizeof(unsigned long long) * kXcds * 256) != hipSuccess) {
            cleanup();
            err = "stage 0: device allocation failed";
            return KC_ERR_NOMEM;
        }
        hipError_t e = dcopy_async(d_seq, sample.data(), sample.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && padded > sample.size()) e = hipMemsetAsync(d_seq + sample.size(), '\n', padded - sample.size(), st);
        if (e == hipSuccess) e = hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * kXcds * 256, st);
        Route rt{};
        rt.rhi = 256;  // every coarse bucket
        if (e == hipSuccess) e = extract(nw, false, d_seq, padded, k, rt, d_counts, nullptr, nullptr, st);
        std::vector<unsigned long long> h(kXcds * 256);
        if (e == hipSuccess) e = dcopy_async(h.data(), d_counts, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st);
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
    const uint64_t est = cfg.est_file_size ? cfg.est_file_size : cfg.expected_bytes ? cfg.expected_bytes : sample.size();
    bloom_capacity(out.prefix_counts, est, out.bloom_capacity);
    return KC_OK;
}

namespace {

// max/mean of bucket sizes, buckets given by sorted lower bounds over sorted keys
double imbalance(const std::vector<uint64_t>& sorted, const std::vector<uint64_t>& lower_bounds)
{
    const size_t nb = lower_bounds.size();
    std::vector<uint64_t> pos(nb + 1);
    for (size_t j = 0; j < nb; ++j)
        pos[j] = std::lower_bound(sorted.begin(), sorted.end(), lower_bounds[j]) - sorted.begin();
    pos[nb] = sorted.size();
    uint64_t mx = 0;
    for (size_t j = 0; j < nb; ++j) mx = std::max<uint64_t>(mx, pos[j + 1] - pos[j]);
    const double mean = (double)sorted.size() / (double)nb;
    return mean > 0 ? (double)mx / mean : 0.0;
}

}  // namespace

int learned_plan(uint32_t k, const std::vector<char>& sample, double min_skew, uint32_t train_steps, hipStream_t st,
                 PlanResult& out, std::string& err)
{
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t nw = key_words(k);
    const uint64_t tile = extract_tile(nw);
    const uint64_t padded = ((sample.size() + tile - 1) / tile) * tile;
    if (padded == 0) { err = "learned plan: empty Stage-0 sample"; return KC_ERR_STATE; }
    // sample keys on the device (root-grouped, the order does not matter here)
    uint8_t* d_seq = nullptr;
    unsigned long long* d_small = nullptr;
    uint64_t* d_keys = nullptr;
    uint64_t* d_kept = nullptr;
    void* d_tmp = nullptr;
    auto cleanup = [&]() { hipFree(d_seq); hipFree(d_small); hipFree(d_keys); hipFree(d_kept); hipFree(d_tmp); };
    if (dmalloc(&d_seq, padded + 256) != hipSuccess || dmalloc(&d_small, sizeof(unsigned long long) * (16 * 256 + 600)) != hipSuccess ||
        dmalloc(&d_keys, padded * 8ULL * nw) != hipSuccess) {
        cleanup();
        err = "learned plan: device allocation failed";
        return KC_ERR_NOMEM;
    }
    unsigned long long* counts = d_small;
    unsigned long long* cursor = d_small + 8 * 256;
    unsigned long long* coff = d_small + 16 * 256;
    unsigned long long* tstart = coff + 257;
    Route rt{};
    rt.rhi = 256;  // every coarse bucket
    hipError_t e = dcopy_async(d_seq, sample.data(), sample.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && padded > sample.size()) e = hipMemsetAsync(d_seq + sample.size(), '\n', padded - sample.size(), st);
    if (e == hipSuccess) e = hipMemsetAsync(d_small, 0, sizeof(unsigned long long) * (16 * 256 + 600), st);
    if (e == hipSuccess) e = extract(nw, false, d_seq, padded, k, rt, counts, cursor, nullptr, st);
    if (e == hipSuccess) e = launch_scan_coarse(counts, coff, cursor, tstart, 4096, st);
    if (e == hipSuccess) e = extract(nw, true, d_seq, padded, k, rt, counts, cursor, d_keys, st);
    unsigned long long nkeys = 0;
    if (e == hipSuccess) e = dcopy_async(&nkeys, coff + 256, sizeof(nkeys), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    // word 0 of the keys, at most about 4M of them, sorted (on the device: the
    // sample's 24M keys at workload 3 took 120 ms to copy and filter on the
    // host).  A key is kept by its hash, so the sample does not depend on the
    // order the extraction wrote the keys in, and a kept key keeps all its
    // occurrences.
    const uint64_t cap = 4ULL << 20;
    const uint64_t stride = nkeys > cap ? (nkeys + cap - 1) / cap : 1;
    // kept: hash below 2^64 / stride (about one key in `stride`)
    const uint64_t keep_below = stride == 1 ? ~0ULL : ~0ULL / stride;
    unsigned long long nkept = 0;
    size_t tmp_bytes = 0;
    if (e == hipSuccess && nkeys) e = dmalloc(&d_kept, 2 * 8ULL * nkeys);  // kept keys, then sorted
    if (e == hipSuccess && nkeys) e = hipMemsetAsync(d_small, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess && nkeys) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, (nkeys + 255) / 256);
        hipLaunchKernelGGL(k_keep_by_hash, dim3(grid), dim3(256), 0, st, d_keys, (uint64_t)nkeys, nw, keep_below,
                           stride == 1 ? 1u : 0u, d_kept, d_small);
        e = hipGetLastError();
    }
    if (e == hipSuccess && nkeys) e = dcopy_async(&nkept, d_small, sizeof(nkept), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && nkeys) e = hipStreamSynchronize(st);
    if (e == hipSuccess && nkept)
        e = rocprim::radix_sort_keys(nullptr, tmp_bytes, d_kept, d_kept + nkeys, (size_t)nkept, 0, 64, st);
    if (e == hipSuccess && nkept) e = dmalloc(&d_tmp, tmp_bytes);
    if (e == hipSuccess && nkept)
        e = rocprim::radix_sort_keys(d_tmp, tmp_bytes, d_kept, d_kept + nkeys, (size_t)nkept, 0, 64, st);
    std::vector<uint64_t> sorted(nkept);
    if (e == hipSuccess && nkept) e = dcopy_async(sorted.data(), d_kept + nkeys, 8ULL * nkept, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        cleanup();
        err = std::string("learned plan: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? KC_ERR_NOMEM : KC_ERR_HIP;
    }
    auto lap = [&, last = t0](const char* what) mutable {
        if (!std::getenv("KMC_TIMING")) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[kmc] learned plan: %s %.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    };
    lap("sample keys, sorted");
    MlpTrainOpts o;
    if (train_steps) o.steps = train_steps;
    // the network learns the quantiles of the keys it is shown
    std::vector<uint64_t> uniq(sorted);
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    const std::vector<uint64_t>& train = o.distinct ? uniq : sorted;
    std::vector<uint64_t> pb(kMlpC);
    for (uint32_t j = 0; j < kMlpC; ++j) pb[j] = (uint64_t)j << 52;
    out.sample_max_over_mean_prefix = imbalance(sorted, pb);
    out.sample_distinct_prefix = imbalance(uniq, pb);
    if (uniq.size() < kMlpC || out.sample_distinct_prefix <= min_skew) {
        // fewer distinct sample keys than buckets (nothing to learn), or 12-bit
        // prefixes that already hold about as many distinct keys each (nothing
        // to gain, and learned routing costs more than prefix bits): the run
        // keeps the prefix partition (learned_b stays empty)
        out.fallback = uniq.size() < kMlpC ? 2u : 1u;
        cleanup();
        out.train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return KC_OK;
    }
    e = dcopy(d_keys, train.data(), train.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        err = std::string("learned plan: ") + hipGetErrorString(e);
        return KC_ERR_HIP;
    }
    std::vector<float> params;
    int rc = mlp_train(d_keys, train.size(), o, st, params, out.loss_curve, err);
    lap("train");
    // predictions of every s-th training key (about 1M keys): the isotonic fit
    // places each boundary within s keys of where all the keys would put it
    const size_t s = std::max<size_t>(1, train.size() >> 20);
    std::vector<uint64_t> grid;
    grid.reserve(train.size() / s + 1);
    for (size_t i = 0; i < train.size(); i += s) grid.push_back(train[i]);
    if (!rc && s > 1 && dcopy(d_keys, grid.data(), grid.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        err = "learned plan: copy of the prediction keys failed";
        rc = -2;
    }
    std::vector<float> pred;
    if (!rc) rc = mlp_predict(params.data(), d_keys, grid.size(), o.expect, st, pred, err);
    lap("predict");
    cleanup();
    if (rc) return KC_ERR_HIP;
    boundaries_from_predictions(grid, pred, out.learned_b);
    lap("boundaries");
    out.split_buckets = split_merged_buckets(uniq, out.learned_b);
    out.train_keys = train.size();
    std::vector<uint64_t> lb(out.learned_b.begin(), out.learned_b.begin() + kMlpC);
    out.sample_max_over_mean_learned = imbalance(sorted, lb);
    out.sample_distinct_learned = imbalance(uniq, lb);
    out.train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return KC_OK;
}

}  // namespace kmc
