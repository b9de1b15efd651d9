// host_driver.cpp -- drives the host side of the drop-in (include/kmc_host.h,
// include/kmc_synth.h) with no GPU call, for builds under AddressSanitizer +
// UndefinedBehaviorSanitizer and ThreadSanitizer (tests/sanitize/Makefile):
// the generator's worker threads, the FASTQ chunker and file reader (plain
// and gzip), the Stage-0 sampler, low.bin packing and the multi-threaded
// low.bin / root-file / histogram writers.  Every output is checked against a
// plain recomputation; the exit status is 0 only if all agree.
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "kmc_host.h"
#include "kmc_synth.h"

namespace {

int failures = 0;
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

std::vector<char> read_file(const std::string& p)
{
    std::vector<char> v;
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return v;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) v.insert(v.end(), buf, buf + n);
    std::fclose(f);
    return v;
}

struct Blocks {
    std::vector<char> bytes;
    uint64_t count = 0;
};

int collect(void* user, const char* data, uint64_t len)
{
    Blocks* b = static_cast<Blocks*>(user);
    b->bytes.insert(b->bytes.end(), data, data + len);
    b->count++;
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    // generator, several threads, FASTQ and sequence-only
    kcs_spec spec{};
    spec.config = 2;
    spec.threads = 8;
    spec.seed = 5;
    spec.scale = 400.0;
    spec.format = 0;
    char* fq = nullptr;
    uint64_t fq_len = 0, nreads = 0;
    CHECK(kcs_generate(&spec, &fq, &fq_len, &nreads) == 0);
    spec.format = 1;
    char* seq = nullptr;
    uint64_t seq_len = 0;
    CHECK(kcs_generate(&spec, &seq, &seq_len, &nreads) == 0);
    // the chunker keeps exactly the sequence lines (blocks large enough not to cut a line)
    Blocks whole;
    CHECK(kmc_fastq_blocks(fq, fq_len, 31, 1ULL << 30, collect, &whole) == 0);
    CHECK(whole.bytes.size() == seq_len && std::memcmp(whole.bytes.data(), seq, seq_len) == 0);
    Blocks small;
    CHECK(kmc_fastq_blocks(fq, fq_len, 31, 4096, collect, &small) == 0);
    CHECK(small.count > 100);
    // file reader, plain and gzip, and the Stage-0 sampler
    const std::string plain = dir + "/sanitize_in.fq", gz = dir + "/sanitize_in.fq.gz";
    {
        FILE* f = std::fopen(plain.c_str(), "wb");
        CHECK(f && std::fwrite(fq, 1, fq_len, f) == fq_len);
        if (f) std::fclose(f);
        gzFile g = gzopen(gz.c_str(), "wb6");
        CHECK(g && gzwrite(g, fq, (unsigned)fq_len) == (int)fq_len);
        if (g) gzclose(g);
    }
    for (const std::string& p : {plain, gz}) {
        Blocks b;
        CHECK(kmc_fastq_blocks_file(p.c_str(), 31, 4096, 0, collect, &b) == 0);
        CHECK(b.bytes == small.bytes);
        Blocks s;
        uint64_t est = 0;
        CHECK(kmc_stage0_sample_file(p.c_str(), 31, 32768, collect, &s, &est) == 0);
        CHECK(!s.bytes.empty() && est > 0);
    }
    // low.bin packing and the threaded writers (more than 1 MiB: 8 threads)
    std::mt19937_64 rng(7);
    for (uint32_t k : {31u, 33u, 63u, 100u}) {
        const uint32_t nw = k <= 32 ? 1 : (k <= 64 ? 2 : 4);
        const uint64_t n = 300000;
        std::vector<uint64_t> keys(n * nw);
        for (auto& x : keys) x = rng();
        std::vector<uint8_t> packed(n * kmc_packed_bytes(k));
        kmc_pack_low(keys.data(), n, k, packed.data());
        const std::string lp = dir + "/sanitize_low.bin";
        CHECK(kmc_write_low_keys(lp.c_str(), keys.data(), n, k) == 0);
        const std::vector<char> got = read_file(lp);
        CHECK(got.size() == packed.size() && std::memcmp(got.data(), packed.data(), packed.size()) == 0);
        const uint64_t rb = 8ULL * nw + 8;
        std::vector<uint64_t> recs(n * rb / 8);
        for (auto& x : recs) x = rng();
        const std::string rp = dir + "/sanitize_root.bin";
        CHECK(kmc_write_root_records(rp.c_str(), recs.data(), n, k) == 0);
        const std::vector<char> rgot = read_file(rp);
        CHECK(rgot.size() == n * rb && std::memcmp(rgot.data(), recs.data(), n * rb) == 0);
    }
    std::vector<int64_t> hist(100);
    for (size_t i = 0; i < hist.size(); ++i) hist[i] = (int64_t)(i * i);
    const std::string hp = dir + "/sanitize_hist.tsv";
    CHECK(kmc_write_histogram(hp.c_str(), hist.data(), 1, 100) == 0);
    const std::vector<char> h = read_file(hp);
    CHECK(h.size() > 100 && std::string(h.begin(), h.begin() + 4) == "1\t0\n");
    kcs_free(fq);
    kcs_free(seq);
    std::remove(plain.c_str());
    std::remove(gz.c_str());
    if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
    else std::printf("host driver ok\n");
    return failures ? 1 : 0;
}
