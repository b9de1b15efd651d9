// kmc_host.cpp -- FASTQ chunking, output files and the Tree-equivalent driver.
//
// Everything here runs on the host and reaches the GPU only through the C ABI
// of include/kmc.h.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "kmc_host.h"

namespace {

// Sequence-only block builder with the reference reader's semantics
// (src/FastqReader.h:140-388).  The output does not depend on how the input
// is split into feed() calls.
class FastqChunker {
public:
    FastqChunker(uint32_t k, uint64_t block_size, kmc_block_cb cb, void* user)
        : overlap_(k > 1 ? k - 1 : 0), block_size_(block_size), cb_(cb), user_(user), block_(block_size) {}

    int feed(const char* data, uint64_t len)
    {
        uint64_t pos = 0;
        while (pos < len) {
            if (state_ != kSeq) {
                const void* nl = std::memchr(data + pos, '\n', len - pos);
                if (!nl) return 0;  // rest of a non-sequence line
                pos = (uint64_t)((const char*)nl - data) + 1;
                state_ = (state_ + 1) & 3;
                continue;
            }
            const void* nl = std::memchr(data + pos, '\n', len - pos);
            const uint64_t seq_len = nl ? (uint64_t)((const char*)nl - (data + pos)) : len - pos;
            uint64_t copied = 0;
            while (copied < seq_len) {
                if (fill_ == block_size_) {
                    // full block inside a line: keep the tail that the next
                    // block needs to complete windows (:338-365)
                    take_carry();
                    int rc = publish();
                    if (rc) return rc;
                    std::memcpy(block_.data(), carry_, carry_len_);
                    fill_ = carry_len_;
                    last_nl_ = kNone;
                }
                const uint64_t n = std::min(seq_len - copied, block_size_ - fill_);
                std::memcpy(block_.data() + fill_, data + pos + copied, n);
                fill_ += n;
                copied += n;
            }
            pos += seq_len;
            if (!nl) return 0;  // the line continues in the next feed()
            if (fill_ == block_size_) {
                int rc = publish();
                if (rc) return rc;
            }
            last_nl_ = fill_;
            block_[fill_++] = '\n';
            pos++;
            state_ = kPlus;
        }
        return 0;
    }

    int finish() { return publish(); }

    // back to the header line (the Stage-0 sampler's jump to a record start)
    void reset_line_state() { state_ = 0; }

private:
    static constexpr int kSeq = 1, kPlus = 2;
    static constexpr uint64_t kNone = ~0ULL;

    void take_carry()
    {
        carry_len_ = 0;
        if (overlap_ == 0 || fill_ == 0) return;
        const uint64_t keep = fill_ < overlap_ ? fill_ : overlap_;
        const uint64_t start = fill_ - keep;
        if (last_nl_ != kNone && last_nl_ >= start) {
            carry_len_ = fill_ - (last_nl_ + 1);
            std::memcpy(carry_, block_.data() + last_nl_ + 1, carry_len_);
            return;
        }
        std::memcpy(carry_, block_.data() + start, keep);
        carry_len_ = keep;
    }

    int publish()
    {
        int rc = 0;
        if (fill_ > 0) rc = cb_(user_, block_.data(), fill_);
        fill_ = 0;
        last_nl_ = kNone;
        return rc;
    }

    uint64_t overlap_, block_size_;
    kmc_block_cb cb_;
    void* user_;
    std::vector<char> block_;
    uint64_t fill_ = 0;
    uint64_t last_nl_ = kNone;
    int state_ = 0;
    char carry_[128];
    uint64_t carry_len_ = 0;
};

constexpr uint64_t kReadChunk = 2ULL << 20;  // FASTQ_FILE_CHUNK_SIZE, src/definition.h:90

struct FileSource {
    int fd = -1;
    gzFile gz = nullptr;
    ~FileSource()
    {
        if (gz) gzclose(gz);
        if (fd >= 0) ::close(fd);
    }
    int open(const char* path)
    {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return -1;
        unsigned char magic[2] = {0, 0};
        const ssize_t n = ::read(fd, magic, 2);
        if (n == 2 && magic[0] == 0x1F && magic[1] == 0x8B) {
            gz = gzopen(path, "rb");
            if (!gz) return -1;
        } else if (::lseek(fd, 0, SEEK_SET) != 0) {
            return -1;
        }
        return 0;
    }
    int64_t read(char* buf, uint64_t cap)
    {
        if (gz) return gzread(gz, buf, (unsigned)cap);
        return ::read(fd, buf, cap);
    }
};

bool pwrite_all(int fd, const void* p, uint64_t n, uint64_t off)
{
    const char* c = static_cast<const char*>(p);
    while (n) {
        const ssize_t w = ::pwrite(fd, c, n, (off_t)off);
        if (w <= 0) return false;
        c += w;
        n -= (uint64_t)w;
        off += (uint64_t)w;
    }
    return true;
}

// Output writers: each emitted chunk is split over kWriteThreads threads that
// pack (low.bin) and pwrite their slice at its file offset, so the page-cache
// copies run in parallel while the engine copies the next chunk from HBM.
constexpr uint32_t kWriteThreads = 8;
constexpr uint64_t kParallelBytes = 1ULL << 20;  // smaller chunks are written by the calling thread

template <typename F>
bool parallel_slices(uint64_t n, uint64_t bytes_per_item, F&& f)
{
    const uint32_t t = (n * bytes_per_item < kParallelBytes) ? 1u : kWriteThreads;
    if (t == 1) return f(0, n, 0);
    std::vector<std::thread> th;
    std::vector<char> ok(t, 1);
    const uint64_t per = (n + t - 1) / t;
    for (uint32_t i = 0; i < t; ++i) {
        const uint64_t b = std::min(n, i * per), e = std::min(n, b + per);
        th.emplace_back([&, i, b, e] { ok[i] = f(b, e, i) ? 1 : 0; });
    }
    for (auto& x : th) x.join();
    for (char c : ok)
        if (!c) return false;
    return true;
}

struct LowWriter {
    int fd = -1;
    uint32_t k = 0;
    uint32_t nw = 1;
    uint64_t pb = 0;
    uint64_t at = 0;  // keys written so far
    std::vector<uint8_t> buf[kWriteThreads];
};

int low_cb(void* user, const uint64_t* keys, uint64_t n)
{
    LowWriter* w = (LowWriter*)user;
    const bool ok = parallel_slices(n, w->pb, [&](uint64_t b, uint64_t e, uint32_t t) {
        std::vector<uint8_t>& buf = w->buf[t];
        const uint64_t per = (4ULL << 20) / w->pb;
        for (uint64_t i = b; i < e; i += per) {
            const uint64_t m = std::min(per, e - i);
            buf.resize(m * w->pb);
            kmc_pack_low(keys + i * w->nw, m, w->k, buf.data());
            if (!pwrite_all(w->fd, buf.data(), buf.size(), (w->at + i) * w->pb)) return false;
        }
        return true;
    });
    w->at += n;
    return ok ? 0 : -1;
}

struct RootWriter {
    int fd = -1;
    uint64_t rb = 0;  // bytes per ExportRecord<N>
    uint64_t at = 0;  // records written so far
};

int root_cb(void* user, const void* recs, uint64_t n)
{
    RootWriter* w = (RootWriter*)user;
    // ExportRecord<N> bytes straight from the device buffer (padding is zero)
    const char* p = static_cast<const char*>(recs);
    const bool ok = parallel_slices(n, w->rb, [&](uint64_t b, uint64_t e, uint32_t) {
        return pwrite_all(w->fd, p + b * w->rb, (e - b) * w->rb, (w->at + b) * w->rb);
    });
    w->at += n;
    return ok ? 0 : -1;
}

struct PushCtx {
    kc_engine* h;
    uint64_t bytes = 0;
    int (*fn)(kc_engine*, const char*, uint64_t);
};

int push_cb(void* user, const char* data, uint64_t len)
{
    PushCtx* c = (PushCtx*)user;
    c->bytes += len;
    return c->fn(c->h, data, len);
}

// The whole file, raw, to the device parser in 16 MiB reads.
int push_fastq_file(kc_engine* h, const char* path)
{
    FileSource src;
    if (src.open(path) != 0) return KC_ERR_ARG;
    std::vector<char> buf(8 * kReadChunk);
    for (;;) {
        const int64_t n = src.read(buf.data(), buf.size());
        if (n < 0) return KC_ERR_ARG;
        if (n == 0) return KC_OK;
        int rc = kc_push_fastq(h, buf.data(), (uint64_t)n);
        if (rc) return rc;
    }
}

// Stage-0 sample of a FASTQ file, as FastqPreReader<N>::pre_read reads it
// (src/FastqPreReader.h:110-165, :128-300): need = min(size on disk, 256 MiB).
// Plain files: three parts of need/3 bytes at 0, size/2 - part/2 and
// size - part, each but the first moved to the next "\n@" (:330-347), read
// max(2 MiB, part) bytes at a time.  The read that completes a part's bytes
// resets the line state and moves to the next part before its own bytes are
// parsed (:193-207), so the third part is parsed on from the second part's
// line state, and a sequence line cut at a part's end runs on into the next
// part's first sequence line.  Gzip: 2 MiB reads of the decompressed stream
// while fewer than need bytes were read.  The estimated file size is the size
// on disk, four times that for gzip (:300-310).
int stage0_sample(const char* path, uint32_t k, uint64_t block_size, kmc_block_cb cb, void* user, uint64_t* est_size)
{
    struct stat sb;
    if (::stat(path, &sb) != 0) return KC_ERR_ARG;
    const uint64_t fsize = (uint64_t)sb.st_size;
    FileSource src;
    if (src.open(path) != 0) return KC_ERR_ARG;
    const bool gz = src.gz != nullptr;
    if (est_size) *est_size = gz ? fsize * 4 : fsize;
    const uint64_t need = std::min<uint64_t>(fsize, 256ULL << 20);
    const uint64_t part = need / 3 ? need / 3 : 1;
    const uint64_t chunk = (part > kReadChunk && !gz) ? part : kReadChunk;
    FastqChunker ch(k, block_size, cb, user);
    std::vector<char> buf(chunk);
    uint64_t have = 0;
    if (gz) {
        while (have < need) {
            const int64_t n = src.read(buf.data(), chunk);
            if (n < 0) return KC_ERR_ARG;
            if (n == 0) break;
            have += (uint64_t)n;
            const int rc = ch.feed(buf.data(), (uint64_t)n);
            if (rc) return rc;
        }
        return ch.finish();
    }
    const uint64_t offs[3] = {0, fsize / 2 >= part / 2 ? fsize / 2 - part / 2 : 0, fsize >= part ? fsize - part : 0};
    // the first "\n@" in the 64 KiB from `start` on (its '@'), else `start`
    auto record_start = [&](uint64_t start) -> uint64_t {
        char look[65536];
        const ssize_t n = ::pread(src.fd, look, sizeof(look), (off_t)start);
        if (n <= 1) return start;
        for (ssize_t i = 1; i < n; ++i)
            if (look[i - 1] == '\n' && look[i] == '@') return start + (uint64_t)i;
        return start;
    };
    uint64_t pos = 0, seg_read = 0;
    int seg = 0;
    for (;;) {
        uint64_t n = 0;
        while (n < chunk) {  // one read() of a regular file: full unless at the end
            const ssize_t r = ::pread(src.fd, buf.data() + n, chunk - n, (off_t)(pos + n));
            if (r < 0) return KC_ERR_ARG;
            if (r == 0) break;
            n += (uint64_t)r;
        }
        if (n == 0) break;
        pos += n;
        seg_read += n;
        have += n;
        if (seg_read >= part && seg < 2) {
            ch.reset_line_state();
            ++seg;
            seg_read = 0;
            pos = record_start(offs[seg]);
        }
        const int rc = ch.feed(buf.data(), n);
        if (rc) return rc;
        if (have >= need) break;
    }
    return ch.finish();
}

}  // namespace

extern "C" {

int kmc_stage0_sample_file(const char* path, uint32_t k, uint64_t block_size, kmc_block_cb cb, void* user,
                           uint64_t* est_file_size)
{
    if (!path || !cb || block_size <= 128 || k == 0 || k > 128) return KC_ERR_ARG;
    return stage0_sample(path, k, block_size, cb, user, est_file_size);
}

int kmc_fastq_blocks(const char* fastq, uint64_t len, uint32_t k, uint64_t block_size, kmc_block_cb cb, void* user)
{
    if ((!fastq && len) || !cb || block_size <= 128 || k == 0 || k > 128) return KC_ERR_ARG;
    FastqChunker ch(k, block_size, cb, user);
    int rc = ch.feed(fastq, len);
    if (rc) return rc;
    return ch.finish();
}

int kmc_fastq_blocks_file(const char* path, uint32_t k, uint64_t block_size, uint64_t max_bytes, kmc_block_cb cb,
                          void* user)
{
    if (!path || !cb || block_size <= 128 || k == 0 || k > 128) return KC_ERR_ARG;
    FileSource src;
    if (src.open(path) != 0) return KC_ERR_ARG;
    FastqChunker ch(k, block_size, cb, user);
    std::vector<char> buf(kReadChunk);
    uint64_t total = 0;
    for (;;) {
        uint64_t want = kReadChunk;
        if (max_bytes && total + want > max_bytes) want = max_bytes - total;
        if (want == 0) break;
        const int64_t n = src.read(buf.data(), want);
        if (n < 0) return KC_ERR_ARG;
        if (n == 0) break;
        total += (uint64_t)n;
        int rc = ch.feed(buf.data(), (uint64_t)n);
        if (rc) return rc;
    }
    return ch.finish();
}

uint64_t kmc_packed_bytes(uint32_t k)
{
    const uint64_t full = k / 32, tail_bits = 2ULL * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    return full * 8 + tail_bytes;
}

void kmc_pack_low(const uint64_t* keys, uint64_t n, uint32_t k, uint8_t* out)
{
    const uint32_t nw = k <= 32 ? 1 : (k <= 64 ? 2 : 4);
    const uint64_t full = k / 32, tail_bits = 2ULL * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    const uint64_t mask = tail_bytes ? (~0ULL) << (64 - tail_bits) : 0;
    uint8_t* p = out;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t* key = keys + i * nw;
        std::memcpy(p, key, full * 8);  // little-endian host
        p += full * 8;
        if (tail_bytes) {
            const uint64_t t = key[full] & mask;
            std::memcpy(p, reinterpret_cast<const uint8_t*>(&t) + (8 - tail_bytes), tail_bytes);
            p += tail_bytes;
        }
    }
}

int kmc_write_outputs(kc_engine* h, const char* tmp_dir)
{
    if (!h || !tmp_dir) return KC_ERR_ARG;
    kc_stats st;
    int src = kc_stats_get(h, &st);
    if (src) return src;
    const uint32_t k = st.k;
    std::string dir(tmp_dir);
    if (!dir.empty() && dir.back() != '/') dir.push_back('/');
    LowWriter lw;
    lw.k = k;
    lw.nw = k <= 32 ? 1 : (k <= 64 ? 2 : 4);
    lw.pb = kmc_packed_bytes(k);
    lw.fd = ::open((dir + "low.bin").c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (lw.fd < 0) return KC_ERR_ARG;
    int rc = kc_emit_firsts(h, low_cb, &lw);
    ::close(lw.fd);
    if (rc) return rc;
    for (uint32_t r = 0; r < 256; ++r) {
        RootWriter rw;
        rw.fd = ::open((dir + "root_" + std::to_string(r) + ".bin").c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        rw.rb = 8ULL * lw.nw + 8;
        if (rw.fd < 0) return KC_ERR_ARG;
        rc = kc_emit_root(h, r, root_cb, &rw);
        ::close(rw.fd);
        if (rc) return rc;
    }
    return KC_OK;
}

int kmc_write_low_keys(const char* path, const uint64_t* keys, uint64_t n, uint32_t k)
{
    if (!path || (!keys && n) || k == 0 || k > 128) return KC_ERR_ARG;
    LowWriter lw;
    lw.k = k;
    lw.nw = k <= 32 ? 1 : (k <= 64 ? 2 : 4);
    lw.pb = kmc_packed_bytes(k);
    lw.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (lw.fd < 0) return KC_ERR_ARG;
    const int rc = n ? low_cb(&lw, keys, n) : 0;
    ::close(lw.fd);
    return rc ? KC_ERR_ARG : KC_OK;
}

int kmc_write_root_records(const char* path, const void* records, uint64_t n, uint32_t k)
{
    if (!path || (!records && n) || k == 0 || k > 128) return KC_ERR_ARG;
    RootWriter rw;
    rw.rb = 8ULL * (k <= 32 ? 1 : (k <= 64 ? 2 : 4)) + 8;
    rw.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (rw.fd < 0) return KC_ERR_ARG;
    const int rc = n ? root_cb(&rw, records, n) : 0;
    ::close(rw.fd);
    return rc ? KC_ERR_ARG : KC_OK;
}

int kmc_write_histogram(const char* path, const int64_t* hist, uint32_t min_freq, uint32_t max_freq)
{
    // `f\tn\n` for f = min_freq .. max_freq (tool/histogram_tool.cpp:425-442)
    if (!path || !hist || max_freq < min_freq) return KC_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) {
        std::fprintf(stderr, "failed to open output file: %s\n", path);
        return KC_ERR_ARG;
    }
    for (uint64_t i = 0; i <= (uint64_t)max_freq - min_freq; ++i)
        std::fprintf(f, "%llu\t%lld\n", (unsigned long long)(min_freq + i), (long long)hist[i]);
    return std::fclose(f) == 0 ? KC_OK : KC_ERR_ARG;
}

int kmc_tree_run(const kmc_tree_opts* o, kc_stats* stats, double* wall_ms)
{
    if (!o || !o->fastq_path || !o->tmp_dir || !kc_words_for_k(o->k)) return KC_ERR_ARG;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    struct stat sb;
    if (::stat(o->fastq_path, &sb) != 0) {
        std::fprintf(stderr, "Failed to open file\n");
        return KC_ERR_ARG;
    }

    // Stage 0 first, into host memory: the reference's sample (or the first
    // sample_bytes) and its estimated file size
    struct Sample {
        std::vector<char> bytes;
        std::vector<uint64_t> lens;
    } sample;
    auto keep = [](void* user, const char* data, uint64_t len) -> int {
        Sample* sp = static_cast<Sample*>(user);
        sp->bytes.insert(sp->bytes.end(), data, data + len);
        sp->lens.push_back(len);
        return 0;
    };
    uint64_t est_file = 0;
    int rc = o->sample_bytes ? kmc_fastq_blocks_file(o->fastq_path, o->k, 32ULL << 10, o->sample_bytes, keep, &sample)
                             : kmc_stage0_sample_file(o->fastq_path, o->k, 32ULL << 10, keep, &sample, &est_file);
    if (rc) return rc;
    if (!est_file) est_file = (uint64_t)sb.st_size;
    // sequence bytes per FASTQ byte, from the head of the file: the resident
    // sequence buffer is sized from it (it still grows if the estimate is short)
    uint64_t head_seq = 0;
    auto count_seq = [](void* user, const char*, uint64_t len) -> int {
        *static_cast<uint64_t*>(user) += len;
        return 0;
    };
    const uint64_t head = 4ULL << 20;
    rc = kmc_fastq_blocks_file(o->fastq_path, o->k, 32ULL << 10, head, count_seq, &head_seq);
    if (rc) return rc;
    const double ratio = std::min(1.0, (double)head_seq / (double)std::max<uint64_t>(1, std::min(head, est_file)));

    kc_config cfg{};
    cfg.k = o->k;
    cfg.partition = o->partition;
    cfg.counter = o->counter;
    cfg.min_count = o->min_count ? o->min_count : 1;
    cfg.max_count = o->max_count;
    cfg.expected_bytes = (uint64_t)(ratio * (double)est_file * 1.02) + (1ULL << 20);
    cfg.est_file_size = est_file;
    // classifier thread count of the reference split (src/main.cpp:114-117)
    const uint32_t n = o->n_thread ? o->n_thread : 8;
    const uint32_t parsers = n / 8 > 0 ? n / 8 : 1;
    const uint32_t budget = n > 2 + parsers ? n - 2 - parsers : 1;
    cfg.classifier_num = budget / 2 ? budget / 2 : 1;
    kc_engine* h = nullptr;
    rc = kc_create(&cfg, &h);
    if (rc) return rc;
    {
        uint64_t at = 0;
        for (uint64_t len : sample.lens) {
            if (!rc) rc = kc_sample_block(h, sample.bytes.data() + at, len);
            at += len;
        }
        Sample().bytes.swap(sample.bytes);
    }
    if (!rc) rc = kc_plan(h);
    const auto t1 = clk::now();
    PushCtx pc{h, 0, kc_push_block};
    if (!rc) {
        if (o->host_parse || o->k == 96)
            rc = kmc_fastq_blocks_file(o->fastq_path, o->k, 32ULL << 10, 0, push_cb, &pc);
        else
            rc = push_fastq_file(h, o->fastq_path);
    }
    if (!rc) rc = kc_finish(h);
    const auto t2 = clk::now();
    if (!rc && !o->histogram_only) rc = kmc_write_outputs(h, o->tmp_dir);
    if (!rc && o->histogram_path) {
        // the histogram_tool result straight from the device tables
        const uint32_t lo = o->hist_min ? o->hist_min : 1, hi = o->hist_max ? o->hist_max : 10000;
        if (hi < lo) {
            rc = KC_ERR_ARG;
        } else {
            std::vector<int64_t> hist((size_t)hi - lo + 1);
            rc = kc_histogram(h, KC_HISTOGRAM_PRECISE, lo, hi, hist.data());
            if (!rc) rc = kmc_write_histogram(o->histogram_path, hist.data(), lo, hi);
        }
    }
    const auto t3 = clk::now();
    kc_stats st{};
    if (!rc) rc = kc_stats_get(h, &st);
    if (rc) {
        std::fprintf(stderr, "kmc_tree: %s (status %d)\n", kc_last_error(h), rc);
        kc_destroy(h);
        return rc;
    }
    auto us = [](clk::duration d) { return (long long)std::chrono::duration_cast<std::chrono::microseconds>(d).count(); };
    if (o->verbose) {
        // Stage-0 plan lines as the reference prints them (src/main.cpp:77-82, :106, :186)
        uint64_t sum = 0, max_cap = 0;
        std::vector<uint64_t> loads(cfg.classifier_num, 0);
        for (int p = 0; p < 256; ++p) {
            sum += st.prefix_counts[p];
            max_cap = std::max<uint64_t>(max_cap, st.bloom_capacity[p]);
            if (st.prefix_owners[p] < loads.size()) loads[st.prefix_owners[p]] += st.prefix_counts[p];
        }
        std::printf("Average prefix count: %llu\n", (unsigned long long)(sum / 256));
        for (uint32_t c = 0; c < cfg.classifier_num; ++c)
            std::printf("Classifier %u total prefix count: %llu\n", c, (unsigned long long)loads[c]);
        std::printf("Max Bloom Filter capacity: %llu\n", (unsigned long long)max_cap);
        std::printf("Total read k-mer count: %llu\n", (unsigned long long)st.total_kmers);
        std::printf("Init elapsed us: %lld\n", us(t1 - t0));
        std::printf("Mid elapsed us: %lld\n", us(t2 - t1));
        std::printf("Final elapsed us: %lld\n", us(t3 - t2));
        std::printf("Total elapsed us: %lld\n", us(t3 - t0));
        std::printf("Device counting ms: %.3f (k-mers/s %.3e)\n", st.ms_total,
                    st.ms_total > 0 ? st.total_kmers / (st.ms_total * 1e-3) : 0.0);
        std::printf("High frequency unique k-mer : %llu\n", (unsigned long long)st.records);
    }
    if (stats) *stats = st;
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t3 - t0).count();
    kc_destroy(h);
    return KC_OK;
}

}  // extern "C"
