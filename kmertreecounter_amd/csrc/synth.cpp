// synth.cpp -- deterministic synthetic FASTQ for the five k-mer counting workloads.
//
// Every random draw comes from SplitMix64 streams keyed by (seed, purpose,
// index), so a read's bytes depend only on the seed and its index: the output
// is identical for any thread count and on any x86-64 host.
//
// Workloads (genome length G, read count R at scale 1; `scale` divides G and R):
//   1  G=2 Mbp i.i.d. uniform; 400k x 150 bp; 1% substitutions; no N; quality 'I'.
//   2  G=100 Mbp at 45% GC; 10% of G is 40 families x 50 copies of 5 kbp repeat
//      units (each copy 2% diverged from its family consensus); 20M x 150 bp;
//      0.5% substitutions; each base N with p=0.001; Phred uniform 20..40.
//   3  G=100 Mbp at 65% GC plus 5% low-complexity tracts (1 kbp slots filled
//      with a period 1..6 unit); 20M x 150 bp; 0.5% substitutions.
//   4  G=200 Mbp uniform; 30x of 150 bp reads plus 10x of long reads with
//      lognormal length (median 8 kbp, sigma 0.8, clipped 500..50000) carrying
//      5% errors (60% substitution, 20% insertion, 20% deletion).
//   5  200 genomes x 5 Mbp, GC uniform in [35,65]%, lognormal abundance
//      (sigma 2.0); 20 Gbp of 150 bp reads; 0.5% substitutions (k=63 workload).
// Reads start uniformly, come from either strand with probability 1/2, are
// named @r{index}, have a bare '+' line and never wrap.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kmc_synth.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t purpose, uint64_t index)
        : s(mix64(seed ^ mix64(purpose * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL) ^ mix64(index + 0xD1B54A32D192ED03ULL))) {}
    inline uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ULL;
        return mix64(s);
    }
    // uniform in [0, n)
    inline uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    inline double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    inline double normal()
    {
        double u1 = unit(), u2 = unit();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

enum Purpose : uint64_t {
    P_GENOME = 1, P_REPEAT_SLOTS, P_FAMILY, P_COPY, P_MICRO_SLOTS, P_MICRO, P_READ,
    P_LONG_LEN, P_LONG_READ, P_ABUND, P_GC
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

inline uint8_t comp_char(uint8_t c)
{
    switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return c;
    }
}

inline int char_code(uint8_t c)
{
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    default: return 3;
    }
}

// one base with GC fraction gc (A=T=(1-gc)/2, C=G=gc/2)
inline char gc_base(Rng& r, double gc)
{
    const double u = r.unit();
    const double at = (1.0 - gc) * 0.5;
    if (u < at) return 'A';
    if (u < at + gc * 0.5) return 'C';
    if (u < at + gc) return 'G';
    return 'T';
}

inline char substitute(Rng& r, char c)
{
    const int code = char_code((uint8_t)c);
    return kBases[(code + 1 + (int)r.below(3)) & 3];
}

template <class F>
void parallel_for(uint64_t n, uint32_t threads, F&& f)
{
    if (threads <= 1 || n < 2) { f(0, n, 0u); return; }
    std::vector<std::thread> ts;
    const uint64_t per = (n + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        const uint64_t a = std::min(n, t * per), b = std::min(n, a + per);
        if (a >= b) break;
        ts.emplace_back([=, &f] { f(a, b, t); });
    }
    for (auto& t : ts) t.join();
}

struct Workload {
    uint32_t config = 1;
    uint64_t genome_len = 0;      // per genome
    uint32_t n_genomes = 1;
    uint64_t n_short = 0;         // short reads
    uint32_t short_len = 150;
    uint64_t long_bases = 0;      // total long-read bases (config 4)
    double sub_rate = 0.0;
    double n_rate = 0.0;
    int phred_random = 0;
};

Workload make_workload(const kcs_spec& s)
{
    Workload w;
    w.config = s.config;
    const double sc = s.scale >= 1.0 ? s.scale : 1.0;
    auto scaled = [&](double v) { return (uint64_t)std::llround(v / sc); };
    switch (s.config) {
    case 1: w.genome_len = scaled(2e6); w.n_short = scaled(4e5); w.sub_rate = 0.01; break;
    case 2: w.genome_len = scaled(1e8); w.n_short = scaled(2e7); w.sub_rate = 0.005; w.n_rate = 0.001; w.phred_random = 1; break;
    case 3: w.genome_len = scaled(1e8); w.n_short = scaled(2e7); w.sub_rate = 0.005; break;
    case 4: w.genome_len = scaled(2e8); w.n_short = scaled(4e7); w.long_bases = scaled(2e9); w.sub_rate = 0.005; break;
    case 5: w.genome_len = scaled(5e6); w.n_genomes = 200; w.n_short = scaled(2e10 / 150.0); w.sub_rate = 0.005; break;
    default: break;
    }
    if (s.genome_len) w.genome_len = s.genome_len;
    if (s.n_reads) w.n_short = s.n_reads;
    if (w.genome_len < 1000) w.genome_len = 1000;
    return w;
}

// ---------------------------------------------------------------- genomes
std::vector<std::string> build_genomes(const Workload& w, uint64_t seed, uint32_t threads)
{
    std::vector<std::string> g(w.n_genomes);
    std::vector<double> gcs(w.n_genomes, 0.5);
    if (w.config == 2) gcs[0] = 0.45;
    if (w.config == 3) gcs[0] = 0.65;
    if (w.config == 5) {
        for (uint32_t i = 0; i < w.n_genomes; ++i) {
            Rng r(seed, P_GC, i);
            gcs[i] = 0.35 + 0.30 * r.unit();
        }
    }
    const uint64_t CH = 1ULL << 16;
    for (uint32_t gi = 0; gi < w.n_genomes; ++gi) {
        std::string& s = g[gi];
        s.resize(w.genome_len);
        const uint64_t chunks = (w.genome_len + CH - 1) / CH;
        const double gc = gcs[gi];
        parallel_for(chunks, threads, [&](uint64_t a, uint64_t b, uint32_t) {
            for (uint64_t c = a; c < b; ++c) {
                Rng r(seed, P_GENOME, ((uint64_t)gi << 40) | c);
                const uint64_t e = std::min(w.genome_len, (c + 1) * CH);
                for (uint64_t p = c * CH; p < e; ++p) s[p] = gc == 0.5 ? kBases[r.next() >> 62] : gc_base(r, gc);
            }
        });
    }
    if (w.config == 2) {
        // 40 families x 50 copies of 5 kbp placed in 5 kbp slots (10% of G)
        const uint64_t unit = 5000, fams = 40, copies = 50;
        const uint64_t slots = w.genome_len / unit;
        const uint64_t need = std::min<uint64_t>(fams * copies, slots / 10 ? slots / 10 : 1);
        std::vector<std::string> cons(fams);
        for (uint64_t f = 0; f < fams; ++f) {
            Rng r(seed, P_FAMILY, f);
            cons[f].resize(unit);
            for (uint64_t p = 0; p < unit; ++p) cons[f][p] = gc_base(r, 0.45);
        }
        std::vector<uint64_t> idx(slots);
        for (uint64_t i = 0; i < slots; ++i) idx[i] = i;
        Rng rs(seed, P_REPEAT_SLOTS, 0);
        for (uint64_t i = 0; i < need && i + 1 < slots; ++i) std::swap(idx[i], idx[i + rs.below(slots - i)]);
        std::string& s = g[0];
        for (uint64_t j = 0; j < need; ++j) {
            Rng r(seed, P_COPY, j);
            const std::string& c = cons[j % fams];
            const uint64_t base = idx[j] * unit;
            for (uint64_t p = 0; p < unit; ++p) s[base + p] = (r.unit() < 0.02) ? substitute(r, c[p]) : c[p];
        }
    }
    if (w.config == 3) {
        // 5% of G in 1 kbp low-complexity slots, each a repeated unit of period 1..6
        const uint64_t unit = 1000;
        const uint64_t slots = w.genome_len / unit;
        const uint64_t need = slots / 20;
        std::vector<uint64_t> idx(slots);
        for (uint64_t i = 0; i < slots; ++i) idx[i] = i;
        Rng rs(seed, P_MICRO_SLOTS, 0);
        for (uint64_t i = 0; i < need && i + 1 < slots; ++i) std::swap(idx[i], idx[i + rs.below(slots - i)]);
        std::string& s = g[0];
        for (uint64_t j = 0; j < need; ++j) {
            Rng r(seed, P_MICRO, j);
            const uint32_t period = 1 + (uint32_t)r.below(6);
            char u[6];
            for (uint32_t q = 0; q < period; ++q) u[q] = gc_base(r, 0.65);
            const uint64_t base = idx[j] * unit;
            for (uint64_t p = 0; p < unit; ++p) s[base + p] = u[p % period];
        }
    }
    return g;
}

// ---------------------------------------------------------------- reads
struct ReadCtx {
    const Workload* w;
    const std::vector<std::string>* genomes;
    const std::vector<double>* cdf;      // genome choice (config 5)
    const std::vector<uint32_t>* long_lens;
    uint64_t seed;
    int fastq;
};

void emit_record(std::string& out, uint64_t idx, const char* seq, uint64_t len, const char* qual, int fastq)
{
    if (fastq) {
        out.push_back('@');
        out.push_back('r');
        out += std::to_string(idx);
        out.push_back('\n');
    }
    out.append(seq, len);
    out.push_back('\n');
    if (fastq) {
        out.append("+\n");
        out.append(qual, len);
        out.push_back('\n');
    }
}

void gen_short(const ReadCtx& c, uint64_t i, std::string& out, std::string& seq, std::string& qual)
{
    const Workload& w = *c.w;
    Rng r(c.seed, P_READ, i);
    uint32_t gi = 0;
    if (c.cdf && c.cdf->size() > 1) {
        const double u = r.unit();
        gi = (uint32_t)(std::upper_bound(c.cdf->begin(), c.cdf->end(), u) - c.cdf->begin());
        if (gi >= c.cdf->size()) gi = (uint32_t)c.cdf->size() - 1;
    }
    const std::string& g = (*c.genomes)[gi];
    const uint64_t L = w.short_len;
    const uint64_t start = r.below(g.size() - L + 1);
    const bool rc = (r.next() >> 63) != 0;
    seq.resize(L);
    qual.resize(L);
    for (uint64_t p = 0; p < L; ++p) {
        seq[p] = rc ? (char)comp_char((uint8_t)g[start + L - 1 - p]) : g[start + p];
    }
    for (uint64_t p = 0; p < L; ++p) {
        if (w.sub_rate > 0 && r.unit() < w.sub_rate) seq[p] = substitute(r, seq[p]);
        if (w.n_rate > 0 && r.unit() < w.n_rate) seq[p] = 'N';
        qual[p] = w.phred_random ? (char)(33 + 20 + (int)r.below(21)) : 'I';
    }
    emit_record(out, i, seq.data(), L, qual.data(), c.fastq);
}

void gen_long(const ReadCtx& c, uint64_t i, uint64_t name_idx, std::string& out, std::string& seq, std::string& qual)
{
    const std::string& g = (*c.genomes)[0];
    Rng r(c.seed, P_LONG_READ, i);
    const uint64_t want = (*c.long_lens)[i];
    const uint64_t span = std::min<uint64_t>(want, g.size());
    uint64_t pos = r.below(g.size() - span + 1);
    const bool rc = (r.next() >> 63) != 0;
    seq.clear();
    while (seq.size() < want && pos < g.size()) {
        if (r.unit() < 0.05) {
            const double t = r.unit();
            if (t < 0.6) { seq.push_back(substitute(r, g[pos])); pos++; }
            else if (t < 0.8) { seq.push_back(kBases[r.next() >> 62]); }
            else { pos++; }
        } else {
            seq.push_back(g[pos]);
            pos++;
        }
    }
    if (rc) {
        std::reverse(seq.begin(), seq.end());
        for (auto& ch : seq) ch = (char)comp_char((uint8_t)ch);
    }
    qual.assign(seq.size(), 'I');
    emit_record(out, name_idx, seq.data(), seq.size(), qual.data(), c.fastq);
}

}  // namespace

extern "C" {

int kcs_generate(const kcs_spec* spec, char** out, uint64_t* out_len, uint64_t* n_reads_out)
{
    if (!spec || !out || !out_len || spec->config < 1 || spec->config > 5) return -1;
    const Workload w = make_workload(*spec);
    const uint32_t threads = spec->threads ? spec->threads : 1;
    std::vector<std::string> genomes = build_genomes(w, spec->seed, threads);

    std::vector<double> cdf;
    if (w.n_genomes > 1) {
        std::vector<double> a(w.n_genomes);
        double sum = 0;
        for (uint32_t i = 0; i < w.n_genomes; ++i) {
            Rng r(spec->seed, P_ABUND, i);
            a[i] = std::exp(2.0 * r.normal());
            sum += a[i];
        }
        double acc = 0;
        cdf.resize(w.n_genomes);
        for (uint32_t i = 0; i < w.n_genomes; ++i) { acc += a[i] / sum; cdf[i] = acc; }
    }
    std::vector<uint32_t> long_lens;
    if (w.long_bases) {
        Rng r(spec->seed, P_LONG_LEN, 0);
        uint64_t total = 0;
        while (total < w.long_bases) {
            double len = std::exp(std::log(8000.0) + 0.8 * r.normal());
            len = std::min(50000.0, std::max(500.0, len));
            long_lens.push_back((uint32_t)len);
            total += (uint64_t)len;
        }
    }
    ReadCtx ctx{&w, &genomes, cdf.empty() ? nullptr : &cdf, &long_lens, spec->seed, spec->format == 0 ? 1 : 0};

    const uint64_t n_short = w.n_short, n_long = long_lens.size();
    const uint64_t total = n_short + n_long;
    // reads are produced in fixed index order in per-piece buffers, then concatenated
    const uint64_t pieces = std::max<uint64_t>(1, std::min<uint64_t>(total, (uint64_t)threads * 8));
    std::vector<std::string> bufs(pieces);
    std::atomic<uint64_t> next{0};
    auto worker = [&]() {
        std::string seq, qual;
        for (;;) {
            const uint64_t p = next.fetch_add(1);
            if (p >= pieces) break;
            const uint64_t a = total * p / pieces, b = total * (p + 1) / pieces;
            std::string& o = bufs[p];
            o.reserve((b - a) * (spec->format == 0 ? 2 * w.short_len + 24 : w.short_len + 1));
            for (uint64_t i = a; i < b; ++i) {
                if (i < n_short) gen_short(ctx, i, o, seq, qual);
                else gen_long(ctx, i - n_short, i, o, seq, qual);
            }
        }
    };
    std::vector<std::thread> ts;
    for (uint32_t t = 1; t < threads; ++t) ts.emplace_back(worker);
    worker();
    for (auto& t : ts) t.join();

    uint64_t len = 0;
    for (auto& b : bufs) len += b.size();
    char* buf = (char*)std::malloc(len ? len : 1);
    if (!buf) return -2;
    std::vector<uint64_t> offs(pieces + 1, 0);
    for (uint64_t p = 0; p < pieces; ++p) offs[p + 1] = offs[p] + bufs[p].size();
    parallel_for(pieces, threads, [&](uint64_t a, uint64_t b, uint32_t) {
        for (uint64_t p = a; p < b; ++p) {
            std::memcpy(buf + offs[p], bufs[p].data(), bufs[p].size());
            std::string().swap(bufs[p]);
        }
    });
    *out = buf;
    *out_len = len;
    if (n_reads_out) *n_reads_out = total;
    return 0;
}

void kcs_free(char* p) { std::free(p); }

}  // extern "C"
