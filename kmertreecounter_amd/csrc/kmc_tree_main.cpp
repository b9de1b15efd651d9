// kmc_tree -- drop-in for the reference `Tree` command (src/main.cpp:314-374):
//
//   kmc_tree <fastq_file> <k_len> <n_thread> <memory_limit_gb> [map_capacity] [min_count] [max_count]
//
// Same positional arguments and checks; writes ./tmp/low.bin and
// ./tmp/root_{0..255}.bin (./tmp/ must exist, as for the reference) for
// histogram_tool.  The counting runs on the GPU.  KMC_PARTITION=learned picks
// the learned partitioner (default: the 256 root-prefix partition);
// KMC_COUNTER=exact the exact counter, KMC_COUNTER=sort the radix-sort
// counter (default: the reference's Bloom first-occurrence split); KMC_HISTOGRAM=<file> also writes histogram_tool's
// precise histogram computed on the device (KMC_HISTOGRAM_ONLY=1: no tmp/ files).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "kmc_host.h"

static int usage(const char* argv0)
{
    std::fprintf(stderr, "Usage: %s <fastq_file> <k_len> <n_thread> <memory_limit_gb> [map_capacity] [min_count] [max_count]\n", argv0);
    return 1;
}

static bool parse_u64(const char* s, unsigned long long& v)
{
    if (!s || !*s) return false;
    char* end = nullptr;
    v = std::strtoull(s, &end, 10);
    return end && *end == '\0';
}

int main(int argc, char** argv)
{
    if (argc < 5 || argc > 8) return usage(argv[0]);
    unsigned long long k = 0, n_thread = 0, mem = 0, map_cap = 1024, min_c = 1, max_c = 0xFFFFFFFFULL;
    if (!parse_u64(argv[2], k) || !parse_u64(argv[3], n_thread) || !parse_u64(argv[4], mem)) return usage(argv[0]);
    if (argc >= 6 && !parse_u64(argv[5], map_cap)) return usage(argv[0]);
    if (argc >= 7 && !parse_u64(argv[6], min_c)) return usage(argv[0]);
    if (argc >= 8 && !parse_u64(argv[7], max_c)) return usage(argv[0]);
    if (n_thread < 6) {
        std::fprintf(stderr, "n_thread must be >= 6\n");
        return -1;
    }
    if (map_cap <= 1 || map_cap >= 16ULL * 1024 * 1024 || max_c < min_c || max_c > 0xFFFFFFFFULL) return usage(argv[0]);
    if (k == 0 || k > 128) {  // src/main.cpp:579-595
        std::fprintf(stderr, "k_len must be <= 128\n");
        return 1;
    }
    (void)mem;  // device memory is sized from the input, not from this budget
    const char* part = std::getenv("KMC_PARTITION");
    kmc_tree_opts o{};
    o.fastq_path = argv[1];
    o.tmp_dir = "./tmp/";
    o.k = (uint32_t)k;
    o.n_thread = (uint32_t)n_thread;
    o.min_count = (uint32_t)min_c;
    o.max_count = (uint32_t)max_c;
    o.partition = (part && std::strcmp(part, "learned") == 0) ? KC_PARTITION_LEARNED : KC_PARTITION_PREFIX;
    // KMC_HISTOGRAM=<file>: also write histogram_tool's precise histogram (1..10000)
    // from the device tables; KMC_HISTOGRAM_ONLY=1 skips the tmp/ files
    o.histogram_path = std::getenv("KMC_HISTOGRAM");
    const char* hist_only = std::getenv("KMC_HISTOGRAM_ONLY");
    o.histogram_only = (o.histogram_path && hist_only && std::strcmp(hist_only, "1") == 0) ? 1u : 0u;
    const char* counter = std::getenv("KMC_COUNTER");
    o.counter = KC_COUNTER_BLOOM;
    if (counter && std::strcmp(counter, "exact") == 0) o.counter = KC_COUNTER_EXACT;
    if (counter && std::strcmp(counter, "sort") == 0) o.counter = KC_COUNTER_SORT;
    const char* parse = std::getenv("KMC_PARSE");
    o.host_parse = (parse && std::strcmp(parse, "host") == 0) ? 1u : 0u;
    o.verbose = 1;
    kc_stats st;
    double wall = 0;
    const int rc = kmc_tree_run(&o, &st, &wall);
    return rc == 0 ? 0 : 1;
}
