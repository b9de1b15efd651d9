// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" shim, written for this project, that is compiled against
// the reference's own headers where they lie (the reference repository's src/
// and include/ directories; nothing from them is copied here).  The resulting
// oracle/_ref/libref_harness.so lets the tests check the plain-C oracle
// (oracle/kmc_oracle.c) against the reference implementation itself:
//   ref_extract      -> GetKmer<N>::get_next_one + canonical_kmer (src/GetKmer.h:58-82)
//   ref_xxh3_128     -> XXH3_128bits (include/xxhash.h v0.8.3)
//   ref_bloom_classify -> BloomFilter<N>::insert (src/BloomFilter.h:88-118), the
//                       sequential twin of ConcurrentBloomFilter with the same probe
// It is built by oracle/Makefile into oracle/_ref/ and is never part of the product.
#include <cstdint>
#include <cstring>

#include "GetKmer.h"
#include "BloomFilter.h"

namespace {

template <uint32_t N>
uint64_t extract_n(const char* data, uint64_t len, uint32_t k, uint64_t* out, uint64_t cap)
{
    GetKmer<N> g(k);
    g.clear();
    uint64_t n = 0;
    for (uint64_t i = 0; i < len; ++i) {
        if (g.get_next_one(static_cast<unsigned char>(data[i]))) {
            if (n >= cap) return UINT64_MAX;
            std::memcpy(out + n * N, g.canonical_kmer.data.data(), N * sizeof(uint64_t));
            ++n;
        }
    }
    return n;
}

template <uint32_t N>
uint64_t bloom_n(const uint64_t* keys, uint64_t n, uint64_t cap_words, uint64_t* filter_out, uint8_t* is_first)
{
    BloomFilter<N> bf(cap_words);
    uint64_t firsts = 0;
    for (uint64_t i = 0; i < n; ++i) {
        kmer<N> key;
        std::memcpy(key.data.data(), keys + i * N, N * sizeof(uint64_t));
        const bool first = bf.insert(key);
        if (is_first) is_first[i] = first ? 1 : 0;
        firsts += first ? 1 : 0;
    }
    if (filter_out) std::memcpy(filter_out, bf.get_filter_bins(), cap_words * sizeof(uint64_t));
    return firsts;
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) uint64_t ref_extract(const char* data, uint64_t len, uint32_t k,
                                                            uint64_t* out, uint64_t cap)
{
    if (k <= 32) return extract_n<1>(data, len, k, out, cap);
    if (k <= 64) return extract_n<2>(data, len, k, out, cap);
    if (k <= 128) return extract_n<4>(data, len, k, out, cap);
    return UINT64_MAX;
}

__attribute__((visibility("default"))) void ref_xxh3_128(const void* data, uint64_t len, uint64_t* lo, uint64_t* hi)
{
    const XXH128_hash_t h = XXH3_128bits(data, static_cast<size_t>(len));
    *lo = h.low64;
    *hi = h.high64;
}

__attribute__((visibility("default"))) uint64_t ref_bloom_classify(const uint64_t* keys, uint64_t n, uint32_t nw,
                                                                   uint64_t cap_words, uint64_t* filter_out,
                                                                   uint8_t* is_first)
{
    if (nw == 1) return bloom_n<1>(keys, n, cap_words, filter_out, is_first);
    if (nw == 2) return bloom_n<2>(keys, n, cap_words, filter_out, is_first);
    if (nw == 4) return bloom_n<4>(keys, n, cap_words, filter_out, is_first);
    return UINT64_MAX;
}

}  // extern "C"
