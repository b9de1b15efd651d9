// kmc_engine.hip -- host side of the engine: device memory, block staging,
// the pipeline driver and the C ABI declared in include/kmc.h.
#include <hip/hip_runtime.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/kmc.h"
#include "kmc_fault.h"
#include "kmc_kernels.h"
#include "kmc_learned.h"
#include "kmc_pipeline.h"
#include "kmc_plan.h"

using namespace kmc;

static_assert(kCounterBloom == KC_COUNTER_BLOOM && kCounterExact == KC_COUNTER_EXACT && kCounterSort == KC_COUNTER_SORT,
              "kernel counter ids are the kc_counter values");

namespace kmc {

namespace {
// fault hook: calls seen per kind since the hook was last set, and the call
// (1-based) that fails; 0 = none
std::atomic<int64_t> g_fault_calls[3];
std::atomic<int64_t> g_fault_at[3];
}  // namespace

bool fault_hit(uint32_t kind)
{
    const int64_t c = g_fault_calls[kind].fetch_add(1, std::memory_order_relaxed) + 1;
    const int64_t at = g_fault_at[kind].load(std::memory_order_relaxed);
    return at > 0 && c == at;
}

template <int NW>
hipError_t enqueue_pipeline(const PipeBufs& b, const PipeOpts& o, hipStream_t st, bool graph, hipEvent_t* ev)
{
#define PIPE_CHECK(expr)                     \
    do {                                     \
        const hipError_t e_ = (expr);        \
        if (e_ != hipSuccess) return e_;     \
    } while (0)
    const bool sub = o.sub;
    const uint32_t nf = o.nf;
    unsigned long long* counts = b.small + kSmCounts;
    unsigned long long* cursor = b.small + kSmCursor;
    unsigned long long* coarse_off = b.small + kSmCoarseOff;
    unsigned long long* tile_start = b.small + kSmTileStart;
    uint32_t* d_nfirst = b.bucket_u32;
    uint32_t* d_nrec = d_nfirst + nf;
    uint32_t* d_ndist = d_nrec + nf;
    // keys grouped by fine bucket (P3 input) and the per-bucket first-occurrence slices
    uint64_t* src = sub ? b.B : b.A;
    uint64_t* slices = sub ? b.A : b.B;
    auto fill = [&](void* p, uint64_t bytes, uint8_t v) -> hipError_t {
        return graph ? launch_fill(p, bytes, v, st) : hipMemsetAsync(p, v, bytes, st);
    };
    auto mark = [&](int e) -> hipError_t { return ev ? hipEventRecord(ev[e], st) : hipSuccess; };
    if (b.padded > b.seq_len) PIPE_CHECK(fill(b.seq + b.seq_len, b.padded - b.seq_len, '\n'));
    PIPE_CHECK(fill(b.small, sizeof(unsigned long long) * kSmTotal, 0));
    PIPE_CHECK(fill(b.ctr, sizeof(unsigned long long) * kCtrCount, 0));
    // a graph cannot skip the clear, so a captured run always clears the spill table
    if (b.clear_spill || graph) {
        PIPE_CHECK(fill(b.spill_keys, b.spill_cap * 8ULL * TableWords<NW>::value, 0xFF));
        PIPE_CHECK(fill(b.spill_cnt, b.spill_cap * 4ULL, 0));
    }
    if (sub) PIPE_CHECK(fill(b.fcounts, sizeof(unsigned long long) * kXcds * nf, 0));

    // events bracket the main kernel of each phase: ev[2p] .. ev[2p+1]; the
    // heavy-bucket chunks run between ev[9] and ev[11]
    PIPE_CHECK(mark(0));
    if (o.root_counts)
        PIPE_CHECK(hipMemcpyAsync(counts, o.root_counts, sizeof(unsigned long long) * kXcds * 256,
                                  hipMemcpyDeviceToDevice, st));
    else
        PIPE_CHECK(launch_extract<NW>(false, b.seq, b.padded, o.k, o.rt, counts, cursor, nullptr, st));
    PIPE_CHECK(mark(1));
    PIPE_CHECK(launch_scan_coarse(counts, coarse_off, cursor, tile_start, (uint32_t)sub_tile(NW), st, o.rt.rlo, o.rt.rhi));
    PIPE_CHECK(mark(2));
    PIPE_CHECK(launch_extract<NW>(true, b.seq, b.padded, o.k, o.rt, counts, cursor, b.A, st));
    PIPE_CHECK(mark(3));
    const unsigned long long* fine_off = coarse_off;
    if (sub) {
        const uint64_t max_tiles = b.padded / sub_tile(NW) + 256;
        PIPE_CHECK(mark(4));
        PIPE_CHECK(launch_sub<NW>(false, max_tiles, b.A, coarse_off, tile_start, o.rt, b.fcounts, nullptr, nullptr, st));
        PIPE_CHECK(mark(5));
        PIPE_CHECK(launch_scan_fine(b.fcounts, nf, b.fine_off, b.fcursor, b.small + kSmScanPart, st));
        PIPE_CHECK(mark(6));
        PIPE_CHECK(launch_sub<NW>(true, max_tiles, b.A, coarse_off, tile_start, o.rt, b.fcounts, b.fcursor, b.B, st));
        PIPE_CHECK(mark(7));
        fine_off = b.fine_off;
    } else {
        for (int e = 4; e < 8; ++e) PIPE_CHECK(mark(e));
    }
    CountArgs a{};
    a.firsts = slices;
    a.rec_tmp = b.R;
    a.nfirst = d_nfirst;
    a.nrec = d_nrec;
    a.ndist = d_ndist;
    a.ovf = b.ovf;
    a.ovf_count = b.ctr + kCtrOvf;
    a.ovf_cap = b.ovf_cap;
    a.spill_keys = b.spill_keys;
    a.spill_cnt = b.spill_cnt;
    a.spill_cap = b.spill_cap;
    a.spilled = b.ctr + kCtrSpilled;
    a.spill_distinct = b.ctr + kCtrSpillDistinct;
    a.error_flags = b.ctr + kCtrErr;
    a.min_count = o.min_count;
    a.max_count = o.max_count;
    // the exact and sort counters read 0 as their own default (their table or stage size)
    a.pass_keys = o.pass_keys ? o.pass_keys : (o.counter != KC_COUNTER_BLOOM ? 0u : CountCfg<NW>::kPassKeys);
    a.extra_passes = b.ctr + kCtrExtraPasses;
    a.table_slots = o.table_slots;
    a.heavy_keys = o.heavy_keys;
    a.heavy_chunk = o.heavy_chunk;
    a.heavy_list = b.heavy;
    a.heavy_cap = b.heavy_cap;
    a.heavy_count = b.ctr + kCtrHeavyChunks;
    a.heavy_next = b.ctr + kCtrHeavyNext;
    a.heavy_buckets = b.ctr + kCtrHeavyBuckets;
    a.order = d_ndist + nf;  // largest buckets first
    PIPE_CHECK(launch_bucket_order(nf, fine_off, const_cast<uint32_t*>(a.order),
                                   reinterpret_cast<uint32_t*>(b.small + kSmOrder), st));
    PIPE_CHECK(mark(8));
    PIPE_CHECK(launch_count<NW>(nf, o.counter, src, fine_off, a, st));
    PIPE_CHECK(mark(9));
    PIPE_CHECK(launch_heavy<NW>(o.heavy_workgroups, src, a, st));
    PIPE_CHECK(mark(11));
    PIPE_CHECK(launch_spill_emit<NW>(a, st));
    SegArgs sa{};
    sa.nf = nf;
    sa.fine_off = fine_off;
    sa.nfirst = d_nfirst;
    sa.nrec = d_nrec;
    sa.ndist = d_ndist;
    sa.foff = b.foff;
    sa.roff = b.roff;
    sa.root_ovf = b.small + kSmRootOvf;
    sa.root_main = b.small + kSmRootMain;
    sa.root_off = b.small + kSmRootOff;
    sa.ovf_off = b.small + kSmOvfOff;
    sa.ovf_cursor = b.small + kSmOvfCursor;
    sa.ctr = b.ctr;
    sa.rec_out_cap = b.rec_out_cap;
    sa.firsts_out_cap = b.firsts_out_cap;
    sa.part = b.small + kSmSegPart;
    // final firsts into the P3 key input (free now), final records into the slices buffer
    PIPE_CHECK(launch_finalize<NW>(sa, a, slices, pipe_firsts<NW>(b, sub), b.R, pipe_records<NW>(b, sub), st));
    return hipSuccess;
#undef PIPE_CHECK
}

template hipError_t enqueue_pipeline<1>(const PipeBufs&, const PipeOpts&, hipStream_t, bool, hipEvent_t*);
template hipError_t enqueue_pipeline<2>(const PipeBufs&, const PipeOpts&, hipStream_t, bool, hipEvent_t*);
template hipError_t enqueue_pipeline<4>(const PipeBufs&, const PipeOpts&, hipStream_t, bool, hipEvent_t*);

}  // namespace kmc

namespace {

constexpr uint64_t kStageBytes = 64ULL << 20;   // pinned staging buffer for pushed blocks
constexpr uint64_t kEmitChunk = 32ULL << 20;    // pinned bounce buffer for emits
constexpr uint64_t kFineTarget = 32768;         // keys per fine bucket the fan-out aims for


}  // namespace

// What a captured pipeline graph depends on: buffers, sizes and options.  A
// run whose key differs captures the graph again.
struct GraphKey {
    const void* ptrs[18];
    uint64_t vals[18];
};

struct kc_engine {
    kc_config cfg{};
    uint32_t k = 0, nw = 0;
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;

    // resident sequence
    uint8_t* d_seq = nullptr;
    uint64_t seq_len = 0, seq_alloc = 0;
    char* h_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_busy[2] = {false, false};
    int cur_stage = 0;
    uint64_t stage_len = 0;
    double h2d_ms = 0;
    // FASTQ chunks (kc_push_fastq): staged like blocks, parsed on the device
    bool stage_fastq = false;           // the current stage holds FASTQ bytes
    uint8_t* d_raw = nullptr;           // one staged chunk as pushed
    void* d_fq_scratch = nullptr;       // per-tile spans, phases and offsets
    unsigned long long* d_fq_state = nullptr;  // [sequence length, line phase]
    bool fq_state_live = false;         // d_fq_state holds the current stream state
    uint64_t fq_unsettled = 0;          // FASTQ bytes parsed since seq_len was last read back
    uint32_t fq_phase = 0;              // line phase after the last settled chunk

    // Stage 0
    std::vector<char> sample;
    bool planned = false;
    PlanResult plan;

    // work buffers: A/B hold keys (K_cap x N words), R the per-bucket record
    // slices (K_cap x N words of bytes); the final firsts and records reuse them
    uint64_t key_cap = 0;  // keys
    uint64_t* d_A = nullptr;
    uint64_t* d_B = nullptr;
    uint8_t* d_R = nullptr;
    // bucket-range rounds: device buffer the rounds' outputs are copied into
    // while they fit (sized from the first round), kept across runs
    uint8_t* d_arena = nullptr;
    uint64_t arena_cap = 0;
    uint64_t rounds_in_device = 0;  // rounds of the last run whose outputs are in the arena
    unsigned long long* d_small = nullptr;  // coarse counts/cursors/offsets/tile starts, root offsets
    unsigned long long* d_fcounts = nullptr;
    unsigned long long* d_fcursor = nullptr;
    unsigned long long* d_fine_off = nullptr;
    unsigned long long* d_foff = nullptr;
    unsigned long long* d_roff = nullptr;
    uint32_t* d_bucket_u32 = nullptr;       // nfirst, nrec, ndist per fine bucket; counting order
    uint32_t nf_alloc = 0;
    uint8_t* d_ovf = nullptr;               // overflow records
    uint64_t ovf_cap = 0;
    unsigned long long* d_spill_keys = nullptr;
    uint32_t* d_spill_cnt = nullptr;
    uint64_t spill_cap = 0;
    bool spill_dirty = true;
    unsigned long long* d_ctr = nullptr;
    uint64_t* d_bounds = nullptr;  // learned: each root's fine bounds, up to 256 x (kMaxFine - 1)
    uint32_t bounds_sbits = ~0u;   // fan-out (256 << bits fine buckets in all) the uploaded tables were built for
    uint32_t bounds_max_fine = 0;  // and the cap on one root's fine buckets
    uint32_t learned_nf = 0;       // fine buckets of the uploaded learned tables
    uint8_t* d_lut = nullptr;      // learned: routing lookup tables and fine bases (kRouteLutBytes)

    hipEvent_t ev[12] = {};

    // results
    bool have_results = false;
    uint64_t K = 0, n_first = 0, n_rec = 0, n_distinct_rep = 0, n_spilled = 0, max_bucket = 0, extra_passes = 0;
    uint64_t heavy_buckets = 0;
    uint64_t hist_tail_keys = 0, hist_tail_kmers = 0;  // last kc_histogram: keys above max_freq, their k-mers
    std::vector<uint64_t> occ4096;                     // keys per group of fine buckets (4096 groups), last run
    // bucket-range rounds: each round's outputs in host memory (empty when the
    // results of a single round stay on the device)
    // Host memory kept across runs: pinned (hipHostMalloc), so device results
    // copy into it at full link speed; pageable with huge pages if pinning fails.
    struct HostBuf {
        uint8_t* p = nullptr;
        uint64_t cap = 0;
        bool pinned = false;
        uint64_t touched = 0;  // leading bytes already written once (their pages are mapped)
        HostBuf() = default;
        HostBuf(const HostBuf&) = delete;
        HostBuf(HostBuf&& o) noexcept : p(o.p), cap(o.cap), pinned(o.pinned), touched(o.touched)
        {
            o.p = nullptr;
            o.cap = 0;
            o.touched = 0;
        }
        ~HostBuf() { release(); }
        void release()
        {
            if (p && pinned) (void)hipHostFree(p);
            else std::free(p);
            p = nullptr;
            cap = 0;
            touched = 0;
        }
        bool ensure(uint64_t n)
        {
            if (n <= cap) return true;
            release();
            constexpr uint64_t kAlign = 2ULL << 20;
            const uint64_t sz = std::max<uint64_t>(kAlign, (n + kAlign - 1) / kAlign * kAlign);
            void* q = nullptr;
            if (dhost_malloc(&q, sz, hipHostMallocDefault) == hipSuccess && q) {
                p = static_cast<uint8_t*>(q);
                pinned = true;
            } else {
                (void)hipGetLastError();  // clear the failed pinning before the pageable fallback
                p = static_cast<uint8_t*>(std::aligned_alloc(kAlign, sz));
                pinned = false;
                if (p) madvise(p, sz, MADV_HUGEPAGE);
            }
            cap = p ? sz : 0;
            return p != nullptr;
        }
    };
    struct RoundOut {
        HostBuf firsts;  // n_first x NW words
        HostBuf recs;    // n_rec records
        // outputs kept in the device arena (null: in the host buffers)
        const uint64_t* d_firsts = nullptr;
        const uint8_t* d_recs = nullptr;
        uint64_t n_first = 0, n_rec = 0;
        std::vector<unsigned long long> root_off;  // 2 x 257, as root_off
        const uint64_t* first_words() const { return reinterpret_cast<const uint64_t*>(firsts.p); }
        bool on_device() const { return d_firsts != nullptr || d_recs != nullptr; }
    };
    std::vector<RoundOut> rounds;  // the first rounds_used hold the last run's outputs
    uint32_t rounds_used = 0;
    uint32_t n_rounds = 0;
    hipGraphExec_t graph_exec = nullptr;  // captured single-round pipeline (kc_config.graph)
    GraphKey graph_key{};
    unsigned long long* d_heavy = nullptr;  // heavy-bucket chunk list, heavy_cap x (lo, hi)
    uint64_t heavy_cap = 0;
    uint32_t cu_count = 0;
    uint32_t nf_last = 0;
    const uint64_t* res_firsts = nullptr;
    const uint8_t* res_recs = nullptr;
    std::vector<unsigned long long> root_off;
    double phase_ms[8] = {};  // P0, P1, P2a, P2b, P3, P4, total, heavy buckets
    char* h_bounce = nullptr;
    char* h_bounce2 = nullptr;             // second emit buffer: the next chunk copies while the callback runs
    hipEvent_t emit_ev[2] = {nullptr, nullptr};

    int fail(int code, const std::string& what)
    {
        err = what;
        return code;
    }
    int hip_fail(hipError_t e, const char* where)
    {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return KC_ERR_HIP;
    }
};

#define HIPX(expr)                                              \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return h->hip_fail(e_, #expr);    \
    } while (0)

namespace {


uint64_t pow2_at_least(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

int ensure_seq_capacity(kc_engine* h, uint64_t need_len)
{
    // allocation covers need_len rounded up to a whole extraction tile plus slack
    const uint64_t tile = ext_tile(h->nw);
    const uint64_t want = ((need_len + tile - 1) / tile) * tile + 256;
    if (want <= h->seq_alloc) return KC_OK;
    uint64_t alloc = std::max<uint64_t>(want, h->seq_alloc * 2);
    uint8_t* nd = nullptr;
    hipError_t e = dmalloc(&nd, alloc);
    if (e != hipSuccess) {
        alloc = want;
        e = dmalloc(&nd, alloc);
        if (e != hipSuccess) return h->fail(KC_ERR_NOMEM, "device allocation for the sequence buffer failed");
    }
    if (h->d_seq) {
        // parsed FASTQ output not yet read back is at most fq_unsettled bytes
        const uint64_t used = std::min<uint64_t>(h->seq_alloc, h->seq_len + h->fq_unsettled);
        hipError_t ce = hipStreamSynchronize(h->st);
        if (ce == hipSuccess) ce = dcopy(nd, h->d_seq, used, hipMemcpyDeviceToDevice);
        if (ce != hipSuccess) {
            (void)hipFree(nd);  // the old buffer stays in place
            return h->hip_fail(ce, "copy of the sequence buffer");
        }
        (void)hipFree(h->d_seq);
    }
    h->d_seq = nd;
    h->seq_alloc = alloc;
    return KC_OK;
}

int flush_stage(kc_engine* h)
{
    if (h->stage_len == 0) return KC_OK;
    int rc = ensure_seq_capacity(h, h->seq_len + h->stage_len);
    if (rc) return rc;
    const int s = h->cur_stage;
    HIPX(dcopy_async(h->d_seq + h->seq_len, h->h_stage[s], h->stage_len, hipMemcpyHostToDevice, h->st));
    HIPX(hipEventRecord(h->stage_ev[s], h->st));
    h->stage_busy[s] = true;
    h->seq_len += h->stage_len;
    h->stage_len = 0;
    h->cur_stage ^= 1;
    const int n = h->cur_stage;
    if (h->stage_busy[n]) {
        HIPX(hipEventSynchronize(h->stage_ev[n]));
        h->stage_busy[n] = false;
    }
    return KC_OK;
}

// The staged FASTQ chunk goes to d_raw and its sequence lines are appended to
// d_seq by the parse kernels; the sequence length stays on the device until
// settle_fastq reads it back.
int flush_fastq(kc_engine* h)
{
    if (h->stage_len == 0) return KC_OK;
    if (!h->d_raw || !h->d_fq_scratch || !h->d_fq_state) {
        dfree(h->d_raw);
        dfree(h->d_fq_scratch);
        dfree(h->d_fq_state);
        if (dmalloc(&h->d_raw, kStageBytes + 64) != hipSuccess ||
            dmalloc(&h->d_fq_scratch, fastq_tile_scratch_bytes(kStageBytes)) != hipSuccess ||
            dmalloc(&h->d_fq_state, 2 * sizeof(unsigned long long)) != hipSuccess) {
            dfree(h->d_raw);
            dfree(h->d_fq_scratch);
            dfree(h->d_fq_state);
            return h->fail(KC_ERR_NOMEM, "device allocation for FASTQ parsing failed");
        }
    }
    // each raw byte yields at most one sequence byte
    int rc = ensure_seq_capacity(h, h->seq_len + h->fq_unsettled + h->stage_len);
    if (rc) return rc;
    if (!h->fq_state_live) {
        const unsigned long long init[2] = {h->seq_len, h->fq_phase};
        HIPX(hipStreamSynchronize(h->st));
        HIPX(dcopy(h->d_fq_state, init, sizeof(init), hipMemcpyHostToDevice));
        h->fq_state_live = true;
    }
    const int s = h->cur_stage;
    HIPX(dcopy_async(h->d_raw, h->h_stage[s], h->stage_len, hipMemcpyHostToDevice, h->st));
    HIPX(hipEventRecord(h->stage_ev[s], h->st));
    h->stage_busy[s] = true;
    HIPX(launch_fastq_parse(h->d_raw, h->stage_len, h->d_fq_scratch, h->d_fq_state, h->d_seq, h->st));
    h->fq_unsettled += h->stage_len;
    h->stage_len = 0;
    h->cur_stage ^= 1;
    const int n = h->cur_stage;
    if (h->stage_busy[n]) {
        HIPX(hipEventSynchronize(h->stage_ev[n]));
        h->stage_busy[n] = false;
    }
    return KC_OK;
}

int flush_any(kc_engine* h) { return h->stage_fastq ? flush_fastq(h) : flush_stage(h); }

// Reads the sequence length and line phase back after FASTQ chunks were parsed.
int settle_fastq(kc_engine* h)
{
    if (!h->fq_unsettled) return KC_OK;
    unsigned long long st[2] = {0, 0};
    HIPX(hipStreamSynchronize(h->st));
    HIPX(dcopy(st, h->d_fq_state, sizeof(st), hipMemcpyDeviceToHost));
    h->seq_len = st[0];
    h->fq_phase = (uint32_t)st[1];
    h->fq_unsettled = 0;
    return KC_OK;
}

int append_bytes(kc_engine* h, const char* data, uint64_t len)
{
    while (len) {
        const uint64_t room = kStageBytes - h->stage_len;
        const uint64_t n = std::min(room, len);
        std::memcpy(h->h_stage[h->cur_stage] + h->stage_len, data, n);
        h->stage_len += n;
        data += n;
        len -= n;
        if (h->stage_len == kStageBytes) {
            int rc = flush_any(h);
            if (rc) return rc;
        }
    }
    return KC_OK;
}

void free_arena(kc_engine* h)
{
    dfree(h->d_arena);
    h->arena_cap = 0;
    h->rounds_used = 0;  // round outputs kept there are gone
    h->have_results = false;
}

void free_work(kc_engine* h)
{
    dfree(h->d_A);
    dfree(h->d_B);
    dfree(h->d_R);
    h->key_cap = 0;
    free_arena(h);
}

void free_fine(kc_engine* h)
{
    dfree(h->d_fcounts);
    dfree(h->d_fcursor);
    dfree(h->d_fine_off);
    dfree(h->d_foff);
    dfree(h->d_roff);
    dfree(h->d_bucket_u32);
    h->nf_alloc = 0;
}

void free_spill(kc_engine* h)
{
    dfree(h->d_spill_keys);
    dfree(h->d_spill_cnt);
    h->spill_cap = 0;
    h->spill_dirty = true;
}


// Buckets above heavy_keys keys are counted in chunks by the whole GPU.
uint64_t heavy_keys(const kc_engine* h) { return h->cfg.heavy_keys ? h->cfg.heavy_keys : (1ULL << 20); }
uint32_t heavy_chunk_keys(const kc_engine* h)
{
    const uint64_t c = heavy_keys(h) / 4;
    return (uint32_t)std::min<uint64_t>(1ULL << 18, std::max<uint64_t>(1024, c));
}

// Device results to a host buffer kept across runs.  Pinned buffers, and
// pageable ones whose pages were written before, take one copy (measured 53
// and 17 GB/s).  Into fresh pageable pages a plain copy ran at 2.8 GB/s (page
// faults), so those go in 64 MiB chunks through the two pinned staging buffers,
// chunk c + 1 copied from HBM while 8 threads copy chunk c out (faulting the
// pages in parallel).  The staging buffers are idle once pushes are flushed.
int copy_to_host(kc_engine* h, kc_engine::HostBuf& hb, const void* src, uint64_t bytes)
{
    if (bytes == 0) return KC_OK;
    void* dst = hb.p;
    if (hb.pinned || hb.touched >= bytes) {
        HIPX(dcopy_async(dst, src, bytes, hipMemcpyDeviceToHost, h->st));
        HIPX(hipStreamSynchronize(h->st));
        return KC_OK;
    }
    hb.touched = std::max(hb.touched, bytes);
    constexpr uint32_t kThreads = 8;
    char* buf[2] = {h->h_stage[0], h->h_stage[1]};
    const uint64_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
    auto issue = [&](uint64_t c) -> hipError_t {
        const uint64_t off = c * kStageBytes, m = std::min<uint64_t>(kStageBytes, bytes - off);
        hipError_t e = dcopy_async(buf[c & 1], static_cast<const char*>(src) + off, m, hipMemcpyDeviceToHost, h->st);
        if (e == hipSuccess) e = hipEventRecord(h->emit_ev[c & 1], h->st);
        return e;
    };
    HIPX(issue(0));
    for (uint64_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks) HIPX(issue(c + 1));
        HIPX(hipEventSynchronize(h->emit_ev[c & 1]));
        const uint64_t off = c * kStageBytes, m = std::min<uint64_t>(kStageBytes, bytes - off);
        const uint64_t per = (m + kThreads - 1) / kThreads;
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < kThreads; ++t) {
            const uint64_t b = std::min(m, t * per), e = std::min(m, b + per);
            if (e > b)
                th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + off + b, buf[c & 1] + b, e - b); });
        }
        for (auto& x : th) x.join();
    }
    return KC_OK;
}

int ensure_small(kc_engine* h);

int ensure_work(kc_engine* h, uint64_t padded_len, uint32_t nf)
{
    const uint64_t kb = 8ULL * h->nw;
    if (h->key_cap < padded_len) {
        if (std::getenv("KMC_TIMING"))
            std::fprintf(stderr, "[kmc] key buffers %llu -> %llu keys\n", (unsigned long long)h->key_cap,
                         (unsigned long long)padded_len);
        free_work(h);
        if (dmalloc(&h->d_A, padded_len * kb) != hipSuccess || dmalloc(&h->d_B, padded_len * kb) != hipSuccess ||
            dmalloc(&h->d_R, padded_len * kb) != hipSuccess) {
            free_work(h);
            return h->fail(KC_ERR_NOMEM, "device allocation for the key buffers failed");
        }
        h->key_cap = padded_len;
    }
    // Every buffer group is allocated whole or not at all: a failed call leaves
    // null pointers and zero capacities, so the next call allocates again.
    auto nomem = [&](const char* what) { return h->fail(KC_ERR_NOMEM, std::string("device allocation for ") + what + " failed"); };
    if (int rc = ensure_small(h)) return rc;
    if (h->nf_alloc < nf) {
        free_fine(h);
        if (dmalloc(&h->d_fcounts, sizeof(unsigned long long) * kXcds * nf) != hipSuccess ||
            dmalloc(&h->d_fcursor, sizeof(unsigned long long) * kXcds * nf) != hipSuccess ||
            dmalloc(&h->d_fine_off, sizeof(unsigned long long) * (nf + 1)) != hipSuccess ||
            dmalloc(&h->d_foff, sizeof(unsigned long long) * (nf + 1)) != hipSuccess ||
            dmalloc(&h->d_roff, sizeof(unsigned long long) * (nf + 1)) != hipSuccess ||
            dmalloc(&h->d_bucket_u32, sizeof(uint32_t) * 4 * nf) != hipSuccess) {
            free_fine(h);
            return nomem("the fine-bucket arrays");
        }
        h->nf_alloc = nf;
    }
    if (h->ovf_cap == 0) {
        const uint64_t cap = 1ULL << 20;
        dfree(h->d_ovf);
        if (dmalloc(&h->d_ovf, cap * rec_bytes(h->nw)) != hipSuccess) return nomem("the overflow records");
        h->ovf_cap = cap;
    }
    // every listed chunk but the last of a bucket is full: at most K/chunk + nf chunks
    const uint64_t want_heavy = h->key_cap / heavy_chunk_keys(h) + nf + 1;
    if (h->heavy_cap < want_heavy) {
        dfree(h->d_heavy);
        h->heavy_cap = 0;
        if (dmalloc(&h->d_heavy, sizeof(unsigned long long) * 2 * want_heavy) != hipSuccess) return nomem("the heavy-bucket list");
        h->heavy_cap = want_heavy;
    }
    const uint64_t want_spill = std::min<uint64_t>(1ULL << 26, std::max<uint64_t>(1ULL << 16, pow2_at_least(h->key_cap / 512)));
    if (h->spill_cap < want_spill) {
        free_spill(h);
        if (dmalloc(&h->d_spill_keys, want_spill * 8ULL * table_words(h->nw)) != hipSuccess ||
            dmalloc(&h->d_spill_cnt, want_spill * 4ULL) != hipSuccess) {
            free_spill(h);
            return nomem("the spill table");
        }
        h->spill_cap = want_spill;
    }
    return KC_OK;
}

uint32_t learned_max_fine(const kc_engine* h)
{
    return h->cfg.learned_max_fine ? std::min<uint32_t>(h->cfg.learned_max_fine, kMaxFine) : kMaxFine;
}

uint32_t choose_sbits(const kc_engine* h, uint64_t padded_len)
{
    if (h->cfg.sub_bits == KC_SUB_BITS_NONE) return 0;
    if (h->cfg.sub_bits) return std::min<uint32_t>(kMaxSubBits, h->cfg.sub_bits);
    const double per_coarse = (double)padded_len / 256.0;
    uint32_t sb = 0;
    while (sb < 8 && per_coarse / (double)(1u << sb) > (double)kFineTarget) sb++;
    // Past 256 fine buckets per coarse bucket the sub-partition scatter writes
    // short runs and runs at about half speed, so it is only worth it where the
    // counting kernel would otherwise pass over most buckets several times
    // (measured: workload 4 -> 9 bits, workload 5 -> 10 bits, workload 2 stays at 8).
    double pass_keys = h->nw == 1 ? CountCfg<1>::kPassKeys : h->nw == 2 ? CountCfg<2>::kPassKeys : CountCfg<4>::kPassKeys;
    // the sort counter stages at most 8192 / N keys per pass: its buckets aim at 3/4 of that
    if (h->cfg.counter == KC_COUNTER_SORT) pass_keys = 0.75 * (8192.0 / h->nw);
    while (sb >= 8 && sb < kMaxSubBits && per_coarse / (double)(1u << sb) > pass_keys) sb++;
    return sb;
}

// Keys one round can hold: every position when the three key buffers fit the
// device (or max_round_keys), else what the free device memory allows.
template <int NW>
uint64_t round_capacity(const kc_engine* h, uint64_t padded)
{
    uint64_t cap = padded;
    if (h->cfg.max_round_keys) cap = std::min<uint64_t>(cap, h->cfg.max_round_keys);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const uint64_t per_key = 3ULL * 8 * NW;              // A, B and R
        const uint64_t held = h->key_cap * per_key + h->arena_cap;  // buffers this engine already holds
        const uint64_t reserve = 8ULL << 30;                 // tables, spill, histogram, outputs in flight
        const uint64_t avail = free_b + held > reserve ? free_b + held - reserve : 0;
        cap = std::min<uint64_t>(cap, avail / per_key);
    }
    return std::max<uint64_t>(cap, 1024);
}

template <int NW>
GraphKey graph_key(const kc_engine* h, const Route& rt, bool sub, uint32_t nf, uint64_t padded)
{
    GraphKey g{};
    const void* p[18] = {h->d_seq, h->d_A, h->d_B, h->d_R, h->d_small, h->d_ctr, h->d_spill_keys, h->d_spill_cnt,
                         h->d_ovf, h->d_heavy, h->d_fcounts, h->d_fcursor, h->d_fine_off, h->d_foff, h->d_roff,
                         h->d_bucket_u32, h->d_bounds, h->d_lut};
    const uint64_t v[18] = {padded, h->seq_len, (uint64_t)sub | ((uint64_t)rt.sbits << 1), nf, h->key_cap, h->spill_cap, h->ovf_cap, h->heavy_cap,
                            h->cu_count, rt.mode, h->k, h->cfg.min_count, h->cfg.max_count, h->cfg.pass_keys,
                            h->cfg.counter, h->cfg.table_slots, h->cfg.heavy_keys, h->bounds_sbits};
    for (int i = 0; i < 18; ++i) {
        g.ptrs[i] = p[i];
        g.vals[i] = v[i];
    }
    return g;
}

template <int NW>
int run_round_t(kc_engine* h, const Route& rt, bool sub, uint32_t nf, uint64_t padded,
                const unsigned long long* root_counts = nullptr);

// Contiguous root-bucket ranges of at most `cap` keys each (counts: 256 root
// bucket sizes); returns the number of ranges, 0 if one bucket exceeds cap.
uint32_t split_ranges(const uint64_t* counts, uint64_t cap, std::vector<std::pair<uint32_t, uint32_t>>* out)
{
    if (out) out->clear();
    uint64_t acc = 0;
    uint32_t start = 0, n = 0;
    for (uint32_t b = 0; b < 256; ++b) {
        if (counts[b] > cap) return 0;
        if (acc + counts[b] > cap) {
            if (out) out->emplace_back(start, b);
            ++n;
            start = b;
            acc = 0;
        }
        acc += counts[b];
    }
    if (out) out->emplace_back(start, 256u);
    return n + 1;
}

// Copies a round's outputs from HBM into the device arena if they fit there
// (set out.d_*), else into the round's host buffers.
template <int NW>
int keep_round_outputs(kc_engine* h, kc_engine::RoundOut& out, uint64_t& arena_used)
{
    const uint64_t fb = h->n_first * 8ULL * NW, rb = h->n_rec * rec_bytes(NW);
    const uint64_t fb_al = (fb + 255) / 256 * 256, rb_al = (rb + 255) / 256 * 256;
    out.d_firsts = nullptr;
    out.d_recs = nullptr;
    if (!h->cfg.round_outputs && h->d_arena && arena_used + fb_al + rb_al <= h->arena_cap) {
        uint8_t* f = h->d_arena + arena_used;
        uint8_t* r = f + fb_al;
        // stream order keeps the next round's kernels behind these copies
        if (fb) HIPX(dcopy_async(f, h->res_firsts, fb, hipMemcpyDeviceToDevice, h->st));
        if (rb) HIPX(dcopy_async(r, h->res_recs, rb, hipMemcpyDeviceToDevice, h->st));
        out.d_firsts = reinterpret_cast<const uint64_t*>(f);
        out.d_recs = r;
        arena_used += fb_al + rb_al;
        return KC_OK;
    }
    if (!out.firsts.ensure(fb) || !out.recs.ensure(rb)) return h->fail(KC_ERR_NOMEM, "host memory for a round's outputs");
    int rc = copy_to_host(h, out.firsts, h->res_firsts, fb);
    if (!rc) rc = copy_to_host(h, out.recs, h->res_recs, rb);
    return rc;
}

int ensure_small(kc_engine* h)
{
    if (h->d_small && h->d_ctr) return KC_OK;
    dfree(h->d_small);
    dfree(h->d_ctr);
    // kSmTotal words cleared by every pass, then the rounds' root histogram
    if (dmalloc(&h->d_small, sizeof(unsigned long long) * (kSmTotal + kXcds * 256)) != hipSuccess ||
        dmalloc(&h->d_ctr, sizeof(unsigned long long) * kCtrCount) != hipSuccess) {
        dfree(h->d_small);
        dfree(h->d_ctr);
        return h->fail(KC_ERR_NOMEM, "device allocation for the bucket counters failed");
    }
    return KC_OK;
}

template <int NW>
int run_pipeline_t(kc_engine* h)
{
    hipStream_t st = h->st;
    const uint64_t tile = ext_tile(h->nw);
    const uint64_t padded = ((h->seq_len + tile - 1) / tile) * tile;
    const uint32_t sbits = choose_sbits(h, padded);
    const bool learned = h->cfg.partition == KC_PARTITION_LEARNED && !h->plan.learned_b.empty();
    if (learned && (h->bounds_sbits != sbits || h->bounds_max_fine != learned_max_fine(h))) {
        // the roots' fine buckets in proportion to the learned CDF's mass on
        // their key ranges, 256 << sbits in all (about)
        LearnedRouting lr;
        learned_routing(h->plan.learned_b, 256u << sbits, learned_max_fine(h), lr);
        if (!h->d_bounds && dmalloc(&h->d_bounds, sizeof(uint64_t) * 256 * (kMaxFine - 1)) != hipSuccess)
            return h->fail(KC_ERR_NOMEM, "device allocation for the learned bounds failed");
        if (!lr.bounds.empty())
            HIPX(dcopy(h->d_bounds, lr.bounds.data(), sizeof(uint64_t) * lr.bounds.size(), hipMemcpyHostToDevice));
        std::vector<uint8_t> tables(kRouteLutBytes);
        std::memcpy(tables.data(), lr.lut.data(), sizeof(uint16_t) * lr.lut.size());
        std::memcpy(tables.data() + kRouteFineBaseOffset, lr.fine_base.data(), sizeof(uint32_t) * 257);
        if (!h->d_lut && dmalloc(&h->d_lut, kRouteLutBytes) != hipSuccess)
            return h->fail(KC_ERR_NOMEM, "device allocation for the routing tables failed");
        HIPX(dcopy(h->d_lut, tables.data(), kRouteLutBytes, hipMemcpyHostToDevice));
        h->learned_nf = lr.fine_base[256];
        h->bounds_sbits = sbits;
        h->bounds_max_fine = learned_max_fine(h);
    }
    // the learned partition always splits its roots (P2 runs, whatever the fan-out)
    const bool sub = learned || sbits > 0;
    const uint32_t nf = learned ? h->learned_nf : 256u << sbits;
    const uint64_t cap = round_capacity<NW>(h, padded > 0 ? padded : tile);
    const bool one_round = padded <= cap;
    if (one_round) {
        if (h->d_arena) free_arena(h);  // round outputs of an earlier, larger input
        const int rc = ensure_work(h, padded > 0 ? padded : tile, nf);
        if (rc) return rc;
    } else {
        // the key buffers are sized after the root histogram, below
        const int rc = ensure_small(h);
        if (rc) return rc;
    }
    h->occ4096.clear();
    h->rounds_used = 0;
    h->n_rounds = 1;
    h->rounds_in_device = 0;

    Route rt{};
    rt.mode = learned ? 1u : 0u;
    rt.sbits = learned ? 0u : sbits;
    rt.nf = nf;
    rt.fine_base = learned ? reinterpret_cast<const uint32_t*>(h->d_lut + kRouteFineBaseOffset) : nullptr;
    rt.fine_bounds = learned ? h->d_bounds : nullptr;
    rt.lut = learned ? reinterpret_cast<const uint16_t*>(h->d_lut) : nullptr;
    rt.rlo = 0;
    rt.rhi = 256;
    if (one_round) return run_round_t<NW>(h, rt, sub, nf, padded);

    // Bucket-range rounds: more positions than one round holds.  The root
    // histogram splits the 256 root buckets into the fewest contiguous ranges
    // of at most `cap` keys, evened out to the smallest capacity that still
    // needs no more ranges; the key buffers are sized for that, which leaves
    // device memory for the rounds' outputs.  Each round extracts the whole
    // sequence again but keeps its own buckets (key ranges of rounds are
    // disjoint, so outputs, stats and histograms combine).
    unsigned long long* counts = h->d_small + kSmCounts;
    if (padded > h->seq_len) HIPX(hipMemsetAsync(h->d_seq + h->seq_len, '\n', padded - h->seq_len, st));
    HIPX(hipMemsetAsync(h->d_small, 0, sizeof(unsigned long long) * kSmTotal, st));
    HIPX(launch_extract<NW>(false, h->d_seq, padded, h->k, rt, counts, nullptr, nullptr, st));
    std::vector<unsigned long long> hc((size_t)kXcds * 256);
    HIPX(dcopy_async(hc.data(), counts, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost, st));
    // kept past the words a pass clears: every round's pass copies it instead
    // of running the count pass over the whole sequence again
    unsigned long long* saved_counts = h->d_small + kSmTotal;
    HIPX(dcopy_async(saved_counts, counts, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToDevice, st));
    HIPX(hipStreamSynchronize(st));
    uint64_t root[256];
    for (uint32_t b = 0; b < 256; ++b) {
        root[b] = 0;
        for (int x = 0; x < kXcds; ++x) root[b] += hc[(size_t)x * 256 + b];
    }
    const uint32_t nr = split_ranges(root, cap, nullptr);
    if (nr == 0) return h->fail(KC_ERR_CAPACITY, "one root bucket holds more keys than a round");
    uint64_t lo = *std::max_element(root, root + 256), hi = cap;
    while (lo < hi) {  // smallest capacity with at most nr ranges
        const uint64_t mid = lo + (hi - lo) / 2;
        if (split_ranges(root, mid, nullptr) <= nr) hi = mid;
        else lo = mid + 1;
    }
    // buffers already held are kept when they are large enough
    uint64_t need = lo;
    if (h->key_cap >= need && h->key_cap <= cap) need = h->key_cap;
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    split_ranges(root, need, &ranges);
    int rc = ensure_work(h, need, nf);
    if (rc) return rc;
    uint64_t K = 0, nfirst = 0, nrec = 0, spilled = 0, maxb = 0, dist = 0, xp = 0, heavy = 0;
    float phase[8] = {};
    // KMC_TIMING=1: host wall time of each round's count and output copy on stderr
    const bool timing = std::getenv("KMC_TIMING") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto lap = [&](const char* what, uint32_t i) {
        if (!timing) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[kmc] round %u %s %.1f ms\n", i, what,
                     std::chrono::duration<double, std::milli>(now - t_last).count());
        t_last = now;
    };
    lap("ranges", 0);
    uint64_t arena_used = 0;
    for (const auto& r : ranges) {
        rt.rlo = r.first;
        rt.rhi = r.second;
        rc = run_round_t<NW>(h, rt, sub, nf, padded, saved_counts);
        if (rc) return rc;
        lap("count", h->rounds_used);
        if (h->rounds_used == 0 && !h->cfg.round_outputs) {
            // the arena is sized from the first round's outputs (rounds of even
            // key counts have similar outputs), within the free device memory
            // less a reserve for tables that may grow
            const uint64_t first = h->n_first * 8ULL * NW + h->n_rec * rec_bytes(NW) + 512;
            const uint64_t want = (uint64_t)((double)first * (double)ranges.size() * 1.15) + (64ULL << 20);
            size_t free_b = 0, total_b = 0;
            if (h->arena_cap < want && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
                const uint64_t reserve = 8ULL << 30;
                const uint64_t avail = free_b + h->arena_cap > reserve ? free_b + h->arena_cap - reserve : 0;
                uint64_t size = std::min(want, avail) / 256 * 256;
                // KMC_MAX_ARENA_BYTES: a smaller arena (tests of rounds split between HBM and host memory)
                if (const char* m = std::getenv("KMC_MAX_ARENA_BYTES")) size = std::min<uint64_t>(size, std::strtoull(m, nullptr, 10));
                // a larger arena when it holds every round, else only for a real
                // gain: reallocating tens of GB costs seconds (measured 2.2-2.5 s
                // for about 70 GB, every pass while the free memory moved a little)
                if (size > h->arena_cap && (size >= want || size > h->arena_cap + h->arena_cap / 8)) {
                    dfree(h->d_arena);
                    h->arena_cap = 0;
                    if (size >= (1ULL << 20) && dmalloc(&h->d_arena, size) == hipSuccess) h->arena_cap = size;
                    else (void)hipGetLastError();  // no arena: the outputs go to host memory
                }
            }
        }
        if (h->rounds.size() <= h->rounds_used) h->rounds.emplace_back();
        kc_engine::RoundOut& out = h->rounds[h->rounds_used];
        out.n_first = h->n_first;
        out.n_rec = h->n_rec;
        out.root_off = h->root_off;
        rc = keep_round_outputs<NW>(h, out, arena_used);
        if (rc) return rc;
        lap(out.on_device() ? "kept in HBM" : "copy", h->rounds_used);
        h->rounds_in_device += out.on_device() ? 1 : 0;
        ++h->rounds_used;
        K += h->K;
        nfirst += h->n_first;
        nrec += h->n_rec;
        spilled += h->n_spilled;
        maxb = std::max<uint64_t>(maxb, h->max_bucket);
        dist += h->n_distinct_rep;
        xp += h->extra_passes;
        heavy += h->heavy_buckets;
        for (int p = 0; p < 8; ++p) phase[p] += h->phase_ms[p];
    }
    HIPX(hipStreamSynchronize(st));  // the last device-to-device copies
    h->K = K;
    h->n_first = nfirst;
    h->n_rec = nrec;
    h->n_spilled = spilled;
    h->max_bucket = maxb;
    h->n_distinct_rep = dist;
    h->extra_passes = xp;
    h->heavy_buckets = heavy;
    for (int p = 0; p < 8; ++p) h->phase_ms[p] = phase[p];
    h->res_firsts = nullptr;
    h->res_recs = nullptr;
    h->n_rounds = (uint32_t)ranges.size();
    h->have_results = true;
    return KC_OK;
}

template <int NW>
int run_round_t(kc_engine* h, const Route& rt, bool sub, uint32_t nf, uint64_t padded,
                const unsigned long long* root_counts)
{
    hipStream_t st = h->st;
    unsigned long long* coarse_off = h->d_small + kSmCoarseOff;
    for (int attempt = 0; attempt < 6; ++attempt) {
        PipeBufs pb{};
        pb.seq = h->d_seq;
        pb.seq_len = h->seq_len;
        pb.padded = padded;
        pb.A = h->d_A;
        pb.B = h->d_B;
        pb.R = h->d_R;
        pb.key_cap = h->key_cap;
        pb.small = h->d_small;
        pb.ctr = h->d_ctr;
        pb.fcounts = h->d_fcounts;
        pb.fcursor = h->d_fcursor;
        pb.fine_off = h->d_fine_off;
        pb.foff = h->d_foff;
        pb.roff = h->d_roff;
        pb.bucket_u32 = h->d_bucket_u32;
        pb.ovf = h->d_ovf;
        pb.ovf_cap = h->ovf_cap;
        pb.spill_keys = h->d_spill_keys;
        pb.spill_cnt = h->d_spill_cnt;
        pb.spill_cap = h->spill_cap;
        pb.clear_spill = h->spill_dirty;
        pb.heavy = h->d_heavy;
        pb.heavy_cap = h->heavy_cap;
        pb.rec_out_cap = (h->key_cap * 8ULL * NW) / rec_bytes(NW);
        pb.firsts_out_cap = h->key_cap;
        PipeOpts po{};
        po.k = h->k;
        po.sub = sub;
        po.nf = nf;
        po.rt = rt;
        po.min_count = h->cfg.min_count ? h->cfg.min_count : 1u;
        po.max_count = h->cfg.max_count ? h->cfg.max_count : 0xFFFFFFFFu;
        po.pass_keys = h->cfg.pass_keys;
        po.counter = h->cfg.counter;
        po.table_slots = h->cfg.table_slots;
        po.heavy_keys = heavy_keys(h);
        po.heavy_chunk = heavy_chunk_keys(h);
        po.heavy_workgroups = h->cu_count;
        po.root_counts = root_counts;
        auto enqueue = [&](bool graph) -> int {
            const hipError_t e = enqueue_pipeline<NW>(pb, po, st, graph, graph ? nullptr : h->ev);
            return e == hipSuccess ? KC_OK : h->hip_fail(e, "enqueue of the counting pipeline");
        };
        const bool use_graph = h->cfg.graph && rt.rlo == 0 && rt.rhi == 256;
        if (use_graph) {
            const GraphKey key = graph_key<NW>(h, rt, sub, nf, padded);
            if (!h->graph_exec || std::memcmp(&key, &h->graph_key, sizeof(key)) != 0) {
                if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
                h->graph_exec = nullptr;
                HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                const int erc = enqueue(true);
                hipGraph_t g = nullptr;
                const hipError_t ee = hipStreamEndCapture(st, &g);
                if (erc) {
                    if (g) hipGraphDestroy(g);
                    return erc;
                }
                HIPX(ee);
                const hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                HIPX(ie);
                h->graph_key = key;
            }
            HIPX(hipEventRecord(h->ev[0], st));
            // the spill table counts as written until this run's counters are read back
            h->spill_dirty = true;
            HIPX(hipGraphLaunch(h->graph_exec, st));
            HIPX(hipEventRecord(h->ev[10], st));
        } else {
            const int erc = enqueue(false);
            h->spill_dirty = true;
            if (erc) return erc;
            HIPX(hipEventRecord(h->ev[10], st));
        }
        HIPX(hipStreamSynchronize(st));

        unsigned long long ctr[kCtrCount];
        HIPX(dcopy(ctr, h->d_ctr, sizeof(ctr), hipMemcpyDeviceToHost));
        unsigned long long coff256 = 0;
        HIPX(dcopy(&coff256, coarse_off + 256, sizeof(coff256), hipMemcpyDeviceToHost));
        h->spill_dirty = ctr[kCtrSpilled] != 0;
        if (ctr[kCtrErr] & kErrSpillFull) {
            if (h->spill_cap >= (1ULL << 31)) return h->fail(KC_ERR_CAPACITY, "spill table cannot grow further");
            const uint64_t ncap = std::min<uint64_t>(1ULL << 31, std::max<uint64_t>(h->spill_cap * 4, pow2_at_least(ctr[kCtrSpilled] * 4)));
            free_spill(h);
            if (dmalloc(&h->d_spill_keys, ncap * 8ULL * TableWords<NW>::value) != hipSuccess || dmalloc(&h->d_spill_cnt, ncap * 4ULL) != hipSuccess) {
                free_spill(h);
                return h->fail(KC_ERR_NOMEM, "spill table growth failed");
            }
            h->spill_cap = ncap;
            continue;
        }
        if (ctr[kCtrErr] & kErrRecFull) {
            const uint64_t ncap = std::max<uint64_t>(h->ovf_cap * 4, ctr[kCtrOvf] + 1024);
            dfree(h->d_ovf);
            h->ovf_cap = 0;
            if (dmalloc(&h->d_ovf, ncap * rec_bytes(NW)) != hipSuccess)
                return h->fail(KC_ERR_NOMEM, "overflow record buffer growth failed");
            h->ovf_cap = ncap;
            continue;
        }
        if (ctr[kCtrErr] & kErrFirstOut)
            return h->fail(KC_ERR_CAPACITY, "first occurrences exceed the key buffer");
        if (ctr[kCtrErr] & kErrRecOut)
            return h->fail(KC_ERR_CAPACITY, "root records exceed the record buffer (too many overflow records)");
        h->K = coff256;
        h->n_first = ctr[kCtrFirst];
        h->n_rec = ctr[kCtrRec];
        h->n_spilled = ctr[kCtrSpilled];
        h->max_bucket = ctr[kCtrMaxBucket];
        h->n_distinct_rep = ctr[kCtrDistinct] + ctr[kCtrSpillDistinct];
        h->extra_passes = ctr[kCtrExtraPasses];
        h->heavy_buckets = ctr[kCtrHeavyBuckets];
        h->nf_last = nf;
        if (sub && nf >= 4096) {
            // keys per group of consecutive fine buckets (fine bucket f in group
            // f * 4096 / nf): the 12-bit prefixes, or about the learned buckets
            std::vector<unsigned long long> fo(nf + 1);
            HIPX(dcopy(fo.data(), h->d_fine_off, sizeof(unsigned long long) * (nf + 1), hipMemcpyDeviceToHost));
            if (h->occ4096.size() != 4096) h->occ4096.assign(4096, 0);
            for (uint32_t f = 0; f < nf; ++f) h->occ4096[(uint32_t)(((uint64_t)f * 4096) / nf)] += fo[f + 1] - fo[f];
        }
        h->res_firsts = pipe_firsts<NW>(pb, sub);
        h->res_recs = pipe_records<NW>(pb, sub);
        h->root_off.assign(2 * 257, 0);
        HIPX(dcopy(h->root_off.data(), h->d_small + kSmRootOff, sizeof(unsigned long long) * 2 * 257, hipMemcpyDeviceToHost));
        // P0, P1, P2a, P2b, P3 kernels; P4 = spill emit + finalize; total;
        // heavy-bucket chunks (a graph replay records the total only)
        const int pairs[8][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {8, 9}, {11, 10}, {0, 10}, {9, 11}};
        float ms = 0;
        for (int p = 0; p < 8; ++p) {
            if (use_graph && p != 6) {
                h->phase_ms[p] = 0.f;
                continue;
            }
            HIPX(hipEventElapsedTime(&ms, h->ev[pairs[p][0]], h->ev[pairs[p][1]]));
            h->phase_ms[p] = ms;
        }
        h->have_results = true;
        return KC_OK;
    }
    return h->fail(KC_ERR_CAPACITY, "device tables did not converge after growth");
}

int run_pipeline(kc_engine* h)
{
    h->have_results = false;
    switch (h->nw) {
    case 1: return run_pipeline_t<1>(h);
    case 2: return run_pipeline_t<2>(h);
    default: return run_pipeline_t<4>(h);
    }
}

// n items of `isz` bytes from device memory to the callback in pinned chunks
// of at most kEmitChunk bytes; chunk c + 1 is copied while the callback
// handles chunk c.
template <typename F>
int emit_chunks(kc_engine* h, const uint8_t* src, uint64_t n, uint64_t isz, F&& f)
{
    if (n == 0) return KC_OK;
    const uint64_t per = kEmitChunk / isz;
    char* buf[2] = {h->h_bounce, h->h_bounce2};
    auto issue = [&](uint64_t c) -> hipError_t {
        const uint64_t i = c * per, m = std::min(per, n - i);
        hipError_t e = dcopy_async(buf[c & 1], src + i * isz, m * isz, hipMemcpyDeviceToHost, h->st);
        if (e == hipSuccess) e = hipEventRecord(h->emit_ev[c & 1], h->st);
        return e;
    };
    const uint64_t chunks = (n + per - 1) / per;
    HIPX(issue(0));
    for (uint64_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks) HIPX(issue(c + 1));
        HIPX(hipEventSynchronize(h->emit_ev[c & 1]));
        const uint64_t m = std::min(per, n - c * per);
        const int rc = f(buf[c & 1], m);
        if (rc) {
            HIPX(hipStreamSynchronize(h->st));
            return h->fail(rc, "emit callback returned an error");
        }
    }
    return KC_OK;
}

}  // namespace

extern "C" {

uint32_t kc_words_for_k(uint32_t k)
{
    if (k == 0 || k > 128) return 0;
    return k <= 32 ? 1u : k <= 64 ? 2u : 4u;
}

const char* kc_last_error(kc_engine* h)
{
    if (!h) return noalloc_last_error();
    return h->err.c_str();
}

int kc_device_memory(uint32_t device, uint64_t* free_bytes, uint64_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return KC_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_ERR_NODEV;
    if ((int)device >= ndev) return KC_ERR_ARG;
    if (hipSetDevice((int)device) != hipSuccess) return KC_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return KC_ERR_HIP;
    *free_bytes = f;
    *total_bytes = t;
    return KC_OK;
}

int64_t kc_debug_fault(uint32_t kind, int64_t nth)
{
    if (kind != kFaultAlloc && kind != kFaultCopy) return -1;
    g_fault_at[kind].store(nth > 0 ? nth : 0, std::memory_order_relaxed);
    return g_fault_calls[kind].exchange(0, std::memory_order_relaxed);
}

int kc_create(const kc_config* cfg, kc_engine** out)
{
    // KMC_FAULT_ALLOC=n / KMC_FAULT_COPY=n: the n-th allocation / copy after the
    // first kc_create of the process fails (the hook of kc_debug_fault)
    static std::once_flag env_hook;
    std::call_once(env_hook, [] {
        if (const char* v = std::getenv("KMC_FAULT_ALLOC")) kc_debug_fault(kFaultAlloc, std::atoll(v));
        if (const char* v = std::getenv("KMC_FAULT_COPY")) kc_debug_fault(kFaultCopy, std::atoll(v));
    });
    if (!cfg || !out) return KC_ERR_ARG;
    *out = nullptr;
    const uint32_t nw = kc_words_for_k(cfg->k);
    if (!nw) return KC_ERR_ARG;
    if (cfg->partition > 1 || (cfg->sub_bits > kMaxSubBits && cfg->sub_bits != KC_SUB_BITS_NONE)) return KC_ERR_ARG;
    if (cfg->counter > KC_COUNTER_SORT || (cfg->table_slots && cfg->table_slots < 16)) return KC_ERR_ARG;
    if (cfg->max_count && cfg->min_count > cfg->max_count) return KC_ERR_ARG;
    if (cfg->learned_max_fine > kMaxFine) return KC_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_ERR_NODEV;
    if ((int)cfg->device >= ndev) return KC_ERR_ARG;
    kc_engine* h = new (std::nothrow) kc_engine();
    if (!h) return KC_ERR_NOMEM;
    h->cfg = *cfg;
    h->k = cfg->k;
    h->nw = nw;
    h->device = (int)cfg->device;
    auto bail = [&](int code) { kc_destroy(h); return code; };
    if (hipSetDevice(h->device) != hipSuccess) return bail(KC_ERR_HIP);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || cus <= 0)
        return bail(KC_ERR_HIP);
    h->cu_count = (uint32_t)cus;  // k_heavy: one persistent workgroup per CU
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return bail(KC_ERR_HIP);
    for (int i = 0; i < 2; ++i) {
        if (dhost_malloc((void**)&h->h_stage[i], kStageBytes, hipHostMallocDefault) != hipSuccess) return bail(KC_ERR_NOMEM);
        if (hipEventCreateWithFlags(&h->stage_ev[i], hipEventDisableTiming) != hipSuccess) return bail(KC_ERR_HIP);
    }
    if (dhost_malloc((void**)&h->h_bounce, kEmitChunk, hipHostMallocDefault) != hipSuccess) return bail(KC_ERR_NOMEM);
    if (dhost_malloc((void**)&h->h_bounce2, kEmitChunk, hipHostMallocDefault) != hipSuccess) return bail(KC_ERR_NOMEM);
    for (int i = 0; i < 2; ++i)
        if (hipEventCreateWithFlags(&h->emit_ev[i], hipEventDisableTiming) != hipSuccess) return bail(KC_ERR_HIP);
    for (auto& e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(KC_ERR_HIP);
    if (cfg->expected_bytes) {
        if (ensure_seq_capacity(h, cfg->expected_bytes + cfg->expected_bytes / 64 + 4096) != KC_OK) return bail(KC_ERR_NOMEM);
    }
    *out = h;
    return KC_OK;
}

int kc_destroy(kc_engine* h)
{
    if (!h) return KC_OK;
    hipSetDevice(h->device);
    if (h->st) hipStreamSynchronize(h->st);
    hipFree(h->d_seq);
    hipFree(h->d_raw);
    hipFree(h->d_fq_scratch);
    hipFree(h->d_fq_state);
    free_work(h);
    hipFree(h->d_small);
    hipFree(h->d_ctr);
    hipFree(h->d_fcounts);
    hipFree(h->d_fcursor);
    hipFree(h->d_fine_off);
    hipFree(h->d_foff);
    hipFree(h->d_roff);
    hipFree(h->d_bucket_u32);
    hipFree(h->d_ovf);
    hipFree(h->d_spill_keys);
    hipFree(h->d_spill_cnt);
    hipFree(h->d_bounds);
    if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
    hipFree(h->d_lut);
    hipFree(h->d_heavy);
    for (int i = 0; i < 2; ++i) {
        if (h->h_stage[i]) hipHostFree(h->h_stage[i]);
        if (h->stage_ev[i]) hipEventDestroy(h->stage_ev[i]);
    }
    if (h->h_bounce) hipHostFree(h->h_bounce);
    if (h->h_bounce2) hipHostFree(h->h_bounce2);
    for (int i = 0; i < 2; ++i)
        if (h->emit_ev[i]) hipEventDestroy(h->emit_ev[i]);
    for (auto& e : h->ev)
        if (e) hipEventDestroy(e);
    if (h->st) hipStreamDestroy(h->st);
    delete h;
    return KC_OK;
}

int kc_sample_block(kc_engine* h, const char* data, uint64_t len)
{
    if (!h || (!data && len)) return KC_ERR_ARG;
    if (!len) return KC_OK;
    h->sample.insert(h->sample.end(), data, data + len);
    if (h->sample.back() != '\n') h->sample.push_back('\n');
    return KC_OK;
}

int kc_plan(kc_engine* h)
{
    if (!h) return KC_ERR_ARG;
    HIPX(hipSetDevice(h->device));
    int rc = stage0_plan(h->k, h->cfg, h->sample, h->st, h->plan, h->err);
    if (rc) return rc;
    if (h->cfg.partition == KC_PARTITION_LEARNED) {
        const double min_skew = (h->cfg.learned_min_skew_pct ? h->cfg.learned_min_skew_pct : 250u) / 100.0;
        rc = learned_plan(h->k, h->sample, min_skew, h->cfg.learned_steps, h->st, h->plan, h->err);
        if (rc) return rc;
        h->bounds_sbits = ~0u;  // routing bounds are rebuilt for the run's fan-out
    }
    h->planned = true;
    return KC_OK;
}

int kc_get_partition_bounds(kc_engine* h, uint64_t* bounds, uint32_t cap, uint32_t* n)
{
    if (!h || !n) return KC_ERR_ARG;
    const std::vector<uint64_t>& b = h->plan.learned_b;
    const bool have = h->cfg.partition == KC_PARTITION_LEARNED && b.size() == kLearnedBounds;
    if (have && (!bounds || cap < kLearnedBounds)) return KC_ERR_ARG;
    *n = have ? kLearnedBounds : 0u;
    if (have) std::memcpy(bounds, b.data(), sizeof(uint64_t) * kLearnedBounds);
    return KC_OK;
}

int kc_set_partition_bounds(kc_engine* h, const uint64_t* bounds, uint32_t n)
{
    if (!h || !bounds || n != kLearnedBounds) return KC_ERR_ARG;
    if (h->cfg.partition != KC_PARTITION_LEARNED)
        return h->fail(KC_ERR_STATE, "partition bounds need a KC_PARTITION_LEARNED handle");
    if (bounds[0] != 0 || bounds[n - 1] != ~0ULL) return h->fail(KC_ERR_ARG, "bounds must run from 0 to ~0");
    for (uint32_t i = 1; i < n; ++i)
        if (bounds[i] < bounds[i - 1]) return h->fail(KC_ERR_ARG, "bounds must be nondecreasing");
    h->plan.learned_b.assign(bounds, bounds + n);
    h->plan.fallback = 0;
    h->bounds_sbits = ~0u;  // routing tables are rebuilt at the next run
    return KC_OK;
}

int kc_push_block(kc_engine* h, const char* data, uint64_t len)
{
    if (!h || (!data && len)) return KC_ERR_ARG;
    if (!len) return KC_OK;
    HIPX(hipSetDevice(h->device));
    h->have_results = false;
    if (h->stage_fastq) {
        int rc = flush_fastq(h);
        if (!rc) rc = settle_fastq(h);
        if (rc) return rc;
        // a FASTQ stream stopped inside a record would run into the block's lines
        if (h->fq_phase != 0) return h->fail(KC_ERR_STATE, "kc_push_block after a FASTQ chunk that ends inside a record");
        h->stage_fastq = false;
        h->fq_state_live = false;  // seq_len moves on without the device state
    }
    int rc = append_bytes(h, data, len);
    if (rc) return rc;
    // a block ending inside a line is followed by one that starts with its carry:
    // a separator keeps windows from spanning the two (each block is parsed from a
    // cleared state, src/FastqParser.h:193)
    if (data[len - 1] != '\n') {
        const char nl = '\n';
        rc = append_bytes(h, &nl, 1);
        if (rc) return rc;
    }
    return KC_OK;
}

int kc_push_fastq(kc_engine* h, const char* data, uint64_t len)
{
    if (!h || (!data && len)) return KC_ERR_ARG;
    if (!len) return KC_OK;
    if (h->k == 96)
        return h->fail(KC_ERR_ARG, "k = 96 keys depend on the reader's 32 KiB block seams: push the host reader's blocks");
    HIPX(hipSetDevice(h->device));
    h->have_results = false;
    if (!h->stage_fastq) {
        int rc = flush_stage(h);
        if (rc) return rc;
        h->stage_fastq = true;
    }
    return append_bytes(h, data, len);
}

int kc_debug_resident(kc_engine* h, char* out, uint64_t cap, uint64_t* len)
{
    if (!h || !len || (!out && cap)) return KC_ERR_ARG;
    HIPX(hipSetDevice(h->device));
    if (h->stage_len || h->fq_unsettled) return h->fail(KC_ERR_STATE, "blocks pushed since the last kc_finish");
    *len = h->seq_len;
    const uint64_t n = std::min<uint64_t>(cap, h->seq_len);
    if (n) {
        HIPX(hipStreamSynchronize(h->st));
        HIPX(dcopy(out, h->d_seq, n, hipMemcpyDeviceToHost));
    }
    return KC_OK;
}

int kc_finish(kc_engine* h)
{
    if (!h) return KC_ERR_ARG;
    HIPX(hipSetDevice(h->device));
    hipEvent_t t0, t1;
    HIPX(hipEventCreate(&t0));
    HIPX(hipEventCreate(&t1));
    HIPX(hipEventRecord(t0, h->st));
    int rc = flush_any(h);
    if (!rc) rc = settle_fastq(h);
    if (rc) return rc;
    HIPX(hipEventRecord(t1, h->st));
    HIPX(hipEventSynchronize(t1));
    float ms = 0;
    hipEventElapsedTime(&ms, t0, t1);
    h->h2d_ms += ms;
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    for (int i = 0; i < 2; ++i) {
        if (h->stage_busy[i]) {
            HIPX(hipEventSynchronize(h->stage_ev[i]));
            h->stage_busy[i] = false;
        }
    }
    return run_pipeline(h);
}

int kc_count_resident(kc_engine* h)
{
    if (!h) return KC_ERR_ARG;
    HIPX(hipSetDevice(h->device));
    if (h->stage_len || h->fq_unsettled) return h->fail(KC_ERR_STATE, "blocks pushed since the last kc_finish");
    return run_pipeline(h);
}

int kc_autotune(kc_engine* h, kc_tune_result* out)
{
    if (!h) return KC_ERR_ARG;
    HIPX(hipSetDevice(h->device));
    if (h->stage_len || h->fq_unsettled) return h->fail(KC_ERR_STATE, "blocks pushed since the last kc_finish");
    if (!h->seq_len) return h->fail(KC_ERR_STATE, "no resident input: push blocks and call kc_finish first");
    // candidates: for each counter (Bloom, exact, sort) its automatic fine
    // fan-out and one bit less, two heavy-bucket thresholds, and with a learned
    // partition and the Bloom counter two caps on a root's fine buckets (1024:
    // heavy roots split as finely as the rest, on the wide sub-partition path;
    // 256); each is timed on one pass over the resident input
    const kc_config saved = h->cfg;
    const uint64_t tile = ext_tile(h->nw);
    const uint64_t padded = ((h->seq_len + tile - 1) / tile) * tile;
    const uint32_t counters[3] = {KC_COUNTER_BLOOM, KC_COUNTER_EXACT, KC_COUNTER_SORT};
    const uint32_t heavies[2] = {0u, 1u << 18};
    double best = 1e300;
    kc_config best_cfg = saved;
    uint32_t tried = 0, skipped = 0;
    int skip_rc = KC_OK;
    for (uint32_t c : counters) {
        kc_config probe = saved;
        probe.sub_bits = 0;
        probe.counter = c;
        h->cfg = probe;
        const uint32_t sb_auto = choose_sbits(h, padded);
        std::vector<uint32_t> sbs{sb_auto};
        if (sb_auto > 1) sbs.push_back(sb_auto - 1);
        if (saved.sub_bits && saved.sub_bits != KC_SUB_BITS_NONE && saved.sub_bits != sb_auto) sbs.push_back(saved.sub_bits);
        const bool learned = saved.partition == KC_PARTITION_LEARNED && !h->plan.learned_b.empty();
        std::vector<uint32_t> caps{saved.learned_max_fine};
        if (learned && c == KC_COUNTER_BLOOM) caps.push_back(saved.learned_max_fine == 256u ? 0u : 256u);
        for (uint32_t sb : sbs)
            for (uint32_t hv : heavies)
                for (uint32_t cap : caps) {
                    kc_config t = saved;
                    t.sub_bits = sb ? sb : KC_SUB_BITS_NONE;
                    t.counter = c;
                    t.heavy_keys = hv;
                    t.learned_max_fine = cap;
                    h->cfg = t;
                    const int rc = run_pipeline(h);
                    // a candidate that does not fit the device (its spill
                    // table or buffers cannot grow) is skipped; any other
                    // error ends the tuning
                    if (rc == KC_ERR_NOMEM || rc == KC_ERR_CAPACITY) {
                        ++skipped;
                        skip_rc = rc;
                        continue;
                    }
                    if (rc) {
                        h->cfg = saved;
                        return rc;
                    }
                    ++tried;
                    if (h->phase_ms[6] < best) {
                        best = h->phase_ms[6];
                        best_cfg = t;
                    }
                }
    }
    if (!tried) {
        h->cfg = saved;
        return h->fail(skip_rc, "autotune: no candidate fits the device");
    }
    h->cfg = best_cfg;
    const int rc = run_pipeline(h);  // results of the chosen configuration
    if (rc) return rc;
    if (out) {
        out->skipped = skipped;
        out->sub_bits = best_cfg.sub_bits == KC_SUB_BITS_NONE ? 0u : best_cfg.sub_bits;
        out->counter = best_cfg.counter;
        out->heavy_keys = best_cfg.heavy_keys ? best_cfg.heavy_keys : (uint32_t)(1u << 20);
        out->candidates = tried;
        out->best_ms = best;
        out->learned_max_fine = best_cfg.learned_max_fine;
    }
    return KC_OK;
}

int kc_emit_firsts(kc_engine* h, kc_keys_cb cb, void* user)
{
    if (!h || !cb) return KC_ERR_ARG;
    if (!h->have_results) return h->fail(KC_ERR_STATE, "no results: call kc_finish first");
    HIPX(hipSetDevice(h->device));
    const uint64_t kb = 8ULL * h->nw;
    const uint64_t per = kEmitChunk / kb;
    if (h->rounds_used) {  // bucket-range rounds: outputs in the device arena or host memory
        for (uint32_t ri = 0; ri < h->rounds_used; ++ri) {
            const auto& r = h->rounds[ri];
            const uint64_t n = r.n_first;
            if (r.on_device()) {
                const int rc = emit_chunks(h, reinterpret_cast<const uint8_t*>(r.d_firsts), n, kb,
                                           [&](const void* p, uint64_t m) {
                                               return cb(user, static_cast<const uint64_t*>(p), m);
                                           });
                if (rc) return rc;
                continue;
            }
            for (uint64_t i = 0; i < n; i += per) {
                int rc = cb(user, r.first_words() + i * h->nw, std::min(per, n - i));
                if (rc) return h->fail(rc, "keys callback returned an error");
            }
        }
        return KC_OK;
    }
    return emit_chunks(h, reinterpret_cast<const uint8_t*>(h->res_firsts), h->n_first, kb, [&](const void* p, uint64_t n) {
        return cb(user, static_cast<const uint64_t*>(p), n);
    });
}

int kc_emit_root(kc_engine* h, uint32_t root_id, kc_records_cb cb, void* user)
{
    if (!h || !cb || root_id > 255) return KC_ERR_ARG;
    if (!h->have_results) return h->fail(KC_ERR_STATE, "no results: call kc_finish first");
    HIPX(hipSetDevice(h->device));
    const uint64_t rb = rec_bytes(h->nw);
    const uint64_t per = kEmitChunk / rb;
    if (h->rounds_used) {  // bucket-range rounds: outputs in the device arena or host memory
        for (uint32_t ri = 0; ri < h->rounds_used; ++ri) {
            const auto& r = h->rounds[ri];
            for (int part = 0; part < 2; ++part) {
                const uint64_t b = r.root_off[part * 257 + root_id], e = r.root_off[part * 257 + root_id + 1];
                if (r.on_device()) {
                    const int rc = emit_chunks(h, r.d_recs + b * rb, e - b, rb,
                                               [&](const void* p, uint64_t m) { return cb(user, p, m); });
                    if (rc) return rc;
                    continue;
                }
                for (uint64_t i = b; i < e; i += per) {
                    int rc = cb(user, r.recs.p + i * rb, std::min(per, e - i));
                    if (rc) return h->fail(rc, "records callback returned an error");
                }
            }
        }
        return KC_OK;
    }
    // in-bucket records of the root, then its overflow records
    for (int part = 0; part < 2; ++part) {
        const uint64_t b = h->root_off[part * 257 + root_id], e = h->root_off[part * 257 + root_id + 1];
        int rc = emit_chunks(h, h->res_recs + b * rb, e - b, rb, [&](const void* p, uint64_t n) { return cb(user, p, n); });
        if (rc) return rc;
    }
    return KC_OK;
}

namespace {

// Histogram of one set of device-resident results into hist (added, not
// stored): the record table doubles until no probe runs out.
int histogram_pass(kc_engine* h, bool precise, uint32_t min_freq, uint32_t max_freq, const uint8_t* recs, uint64_t nrec,
                   const uint64_t* firsts, uint64_t nfirst, void* scratch, uint64_t scratch_bytes, int64_t* hist)
{
    const uint64_t nb = (uint64_t)max_freq - min_freq + 1;
    long long* d_hist = nullptr;
    unsigned long long* d_err = nullptr;
    unsigned long long* d_keys = nullptr;
    uint32_t* d_vals = nullptr;
    bool own_table = false;  // the table lives in the free key buffer when it fits there
    auto free_table = [&]() {
        if (own_table) { hipFree(d_keys); hipFree(d_vals); }
        d_keys = nullptr; d_vals = nullptr; own_table = false;
    };
    auto release = [&]() {
        hipFree(d_hist); hipFree(d_err); free_table();
        d_hist = nullptr; d_err = nullptr;
    };
    if (dmalloc(&d_hist, sizeof(long long) * nb) != hipSuccess || dmalloc(&d_err, sizeof(unsigned long long) * 3) != hipSuccess) {
        release();
        return h->fail(KC_ERR_NOMEM, "histogram allocation failed");
    }
    // the record table is at most half full; a probe that runs out doubles it
    uint64_t cap = precise ? pow2_at_least(std::max<uint64_t>(1024, 2 * nrec)) : 0;
    std::vector<long long> part(nb);
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (precise) {
            free_table();
            const uint64_t kbytes = cap * 8ULL * table_words(h->nw);
            if (scratch && kbytes + cap * 4ULL <= scratch_bytes) {
                d_keys = static_cast<unsigned long long*>(scratch);
                d_vals = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(scratch) + kbytes);
            } else {
                own_table = true;
                if (dmalloc(&d_keys, kbytes) != hipSuccess || dmalloc(&d_vals, cap * 4ULL) != hipSuccess) {
                    release();
                    return h->fail(KC_ERR_NOMEM, "histogram table allocation failed");
                }
            }
        }
        hipError_t e = hipMemsetAsync(d_hist, 0, sizeof(long long) * nb, h->st);
        if (e == hipSuccess) e = hipMemsetAsync(d_err, 0, sizeof(unsigned long long) * 3, h->st);
        if (e == hipSuccess && precise) e = hipMemsetAsync(d_keys, 0xFF, cap * 8ULL * table_words(h->nw), h->st);
        HistArgs a{};
        a.hist = d_hist;
        a.minf = min_freq;
        a.maxf = max_freq;
        a.precise = precise ? 1u : 0u;
        a.keys = d_keys;
        a.vals = d_vals;
        a.cap = cap;
        a.error_flags = d_err;
        a.tail = d_err + 1;
        if (e == hipSuccess)
            e = h->nw == 1   ? launch_histogram<1>(recs, nrec, firsts, nfirst, a, h->st)
                : h->nw == 2 ? launch_histogram<2>(recs, nrec, firsts, nfirst, a, h->st)
                             : launch_histogram<4>(recs, nrec, firsts, nfirst, a, h->st);
        unsigned long long errt[3] = {0, 0, 0};
        if (e == hipSuccess) e = dcopy_async(errt, d_err, sizeof(errt), hipMemcpyDeviceToHost, h->st);
        if (e == hipSuccess) e = hipStreamSynchronize(h->st);
        if (e != hipSuccess) {
            release();
            return h->fail(KC_ERR_HIP, std::string("histogram: ") + hipGetErrorString(e));
        }
        const unsigned long long err = errt[0];
        if (err & kErrHistFull) {
            cap *= 2;
            continue;
        }
        e = dcopy(part.data(), d_hist, sizeof(long long) * nb, hipMemcpyDeviceToHost);
        release();
        if (e != hipSuccess) return h->fail(KC_ERR_HIP, std::string("histogram: ") + hipGetErrorString(e));
        for (uint64_t b = 0; b < nb; ++b) hist[b] += part[b];
        h->hist_tail_keys += errt[1];
        h->hist_tail_kmers += errt[2];
        return KC_OK;
    }
    release();
    return h->fail(KC_ERR_CAPACITY, "histogram table did not converge");
}

}  // namespace

int kc_histogram(kc_engine* h, uint32_t mode, uint32_t min_freq, uint32_t max_freq, int64_t* hist)
{
    if (!h || !hist || min_freq == 0 || max_freq < min_freq || mode > KC_HISTOGRAM_APPROXIMATE) return KC_ERR_ARG;
    if (!h->have_results) return h->fail(KC_ERR_STATE, "no results: call kc_finish first");
    HIPX(hipSetDevice(h->device));
    const bool precise = mode == KC_HISTOGRAM_PRECISE;
    std::fill(hist, hist + ((uint64_t)max_freq - min_freq + 1), (int64_t)0);
    h->hist_tail_keys = 0;
    h->hist_tail_kmers = 0;
    const uint64_t buf_bytes = h->key_cap * 8ULL * h->nw;
    if (!h->rounds_used)  // results in A and B; the record slices (R) are free
        return histogram_pass(h, precise, min_freq, max_freq, h->res_recs, h->n_rec, h->res_firsts, h->n_first, h->d_R,
                              buf_bytes, hist);
    // bucket-range rounds: the key ranges are disjoint, so the rounds' histograms
    // add up; a round kept in the device arena is read there with R as the
    // table, a round in host memory goes back into R and A with B as the table
    for (uint32_t ri = 0; ri < h->rounds_used; ++ri) {
        const auto& r = h->rounds[ri];
        if (r.on_device()) {
            const int rc = histogram_pass(h, precise, min_freq, max_freq, r.d_recs, r.n_rec, r.d_firsts, r.n_first,
                                          h->d_R, buf_bytes, hist);
            if (rc) return rc;
            continue;
        }
        if (r.n_rec) HIPX(dcopy(h->d_R, r.recs.p, r.n_rec * rec_bytes(h->nw), hipMemcpyHostToDevice));
        if (r.n_first) HIPX(dcopy(h->d_A, r.firsts.p, r.n_first * 8ULL * h->nw, hipMemcpyHostToDevice));
        const int rc = histogram_pass(h, precise, min_freq, max_freq, reinterpret_cast<const uint8_t*>(h->d_R), r.n_rec,
                                      h->d_A, r.n_first, h->d_B, buf_bytes, hist);
        if (rc) return rc;
    }
    return KC_OK;
}

int kc_stats_get(kc_engine* h, kc_stats* s)
{
    if (!h || !s) return KC_ERR_ARG;
    std::memset(s, 0, sizeof(*s));
    s->k = h->k;
    s->key_words = h->nw;
    s->total_kmers = h->K;
    s->firsts = h->n_first;
    s->records = h->n_rec;
    s->records_unfiltered = h->n_distinct_rep;
    s->spilled_keys = h->n_spilled;
    s->sequence_bytes = h->seq_len;
    s->fine_buckets = h->nf_last;
    s->max_bucket_keys = h->max_bucket;
    s->extra_passes = h->extra_passes;
    s->heavy_buckets = h->heavy_buckets;
    s->rounds = h->n_rounds;
    s->hist_tail_keys = h->hist_tail_keys;
    s->hist_tail_kmers = h->hist_tail_kmers;
    if (h->occ4096.size() == 4096 && h->K) {
        const uint64_t mx = *std::max_element(h->occ4096.begin(), h->occ4096.end());
        s->bucket4096_max_over_mean = (double)mx / ((double)h->K / 4096.0);
    }
    s->mean_bucket_keys = h->nf_last ? (double)h->K / h->nf_last : 0.0;
    s->ms_extract_count = h->phase_ms[0];
    s->ms_extract_scatter = h->phase_ms[1];
    s->ms_sub_count = h->phase_ms[2];
    s->ms_sub_scatter = h->phase_ms[3];
    s->ms_count = h->phase_ms[4];
    s->ms_compact = h->phase_ms[5];
    s->ms_total = h->phase_ms[6];
    s->ms_heavy = h->phase_ms[7];
    s->rounds_in_device = h->rounds_in_device;
    s->round_key_cap = h->key_cap;
    s->ms_h2d = h->h2d_ms;
    if (h->planned) {
        s->learned_train_keys = h->plan.train_keys;
        s->learned_fallback = h->plan.fallback;
        s->learned_split_buckets = h->plan.split_buckets;
        s->learned_sample_distinct_max_over_mean = h->plan.sample_distinct_learned;
        s->prefix_sample_distinct_max_over_mean = h->plan.sample_distinct_prefix;
        s->learned_train_ms = h->plan.train_ms;
        if (!h->plan.loss_curve.empty()) {
            s->learned_loss_first = h->plan.loss_curve.front();
            s->learned_loss_last = h->plan.loss_curve.back();
        }
        s->learned_sample_max_over_mean = h->plan.sample_max_over_mean_learned;
        s->prefix_sample_max_over_mean = h->plan.sample_max_over_mean_prefix;
        s->stage0_sample_kmers = h->plan.sample_kmers;
        for (int i = 0; i < 256; ++i) {
            s->prefix_counts[i] = h->plan.prefix_counts[i];
            s->bloom_capacity[i] = h->plan.bloom_capacity[i];
            s->prefix_owners[i] = h->plan.owners[i];
        }
    }
    return KC_OK;
}

int kc_debug_copy(uint64_t bytes, uint32_t device, float* ms)
{
    if (bytes == 0 || bytes % 16) return KC_ERR_ARG;
    if (hipSetDevice((int)device) != hipSuccess) return KC_ERR_NODEV;
    void *a = nullptr, *b = nullptr;
    if (dmalloc(&a, bytes) != hipSuccess) return KC_ERR_NOMEM;
    if (dmalloc(&b, bytes) != hipSuccess) { hipFree(a); return KC_ERR_NOMEM; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = KC_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = KC_ERR_HIP;
    if (!rc && hipMemset(a, 1, bytes) != hipSuccess) rc = KC_ERR_HIP;
    if (!rc && hipEventRecord(e0, nullptr) != hipSuccess) rc = KC_ERR_HIP;
    if (!rc && launch_debug_copy(a, b, bytes / 16, nullptr) != hipSuccess) rc = KC_ERR_HIP;
    if (!rc && hipEventRecord(e1, nullptr) != hipSuccess) rc = KC_ERR_HIP;
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = KC_ERR_HIP;
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = KC_ERR_HIP;
    if (ms) *ms = t;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}

int kc_debug_gemm_bf16(const uint16_t* A, const uint16_t* Bt, float* C, uint32_t M, uint32_t N, uint32_t K)
{
    if (!A || !Bt || !C) return KC_ERR_ARG;
    std::string err;
    const int rc = mlp_debug_gemm(A, Bt, C, M, N, K, err);
    return rc == -1 ? KC_ERR_ARG : (rc ? KC_ERR_HIP : KC_OK);
}

int kc_debug_mlp_grad(const float* params, const uint64_t* keys, const uint32_t* labels, uint32_t nb, float* grad,
                      float* loss)
{
    if (!params || !keys || !labels || !grad || !loss) return KC_ERR_ARG;
    for (uint32_t i = 0; i < nb; ++i)
        if (labels[i] >= kMlpC) return KC_ERR_ARG;
    std::string err;
    const int rc = mlp_debug_loss_grad(params, keys, labels, nb, grad, loss, err);
    return rc == -1 ? KC_ERR_ARG : (rc ? KC_ERR_HIP : KC_OK);
}

int kc_debug_learned_fit(const uint64_t* sorted, uint64_t n, uint32_t steps, float lr, uint32_t batch, uint64_t seed,
                         uint32_t flags, uint64_t* bounds, float* loss_curve, uint32_t loss_cap)
{
    if (!sorted || !bounds || n < 64 || batch == 0 || batch % 64 || steps == 0 || flags > 3) return KC_ERR_ARG;
    MlpTrainOpts o;
    o.steps = steps;
    o.lr = lr;
    o.batch = batch;
    o.seed = seed;
    o.distinct = (flags & 1u) != 0;
    o.expect = (flags & 2u) != 0;
    std::string err;
    const int rc = mlp_debug_fit(sorted, n, o, bounds, loss_curve, loss_curve ? loss_cap : 0, err);
    return rc ? KC_ERR_HIP : KC_OK;
}

int kc_debug_learned_routing(const uint64_t* bounds, uint32_t nf_target, uint32_t* fine_base, uint64_t* fine_bounds,
                             uint64_t cap, uint16_t* lut)
{
    if (!bounds || !fine_base || !lut || nf_target == 0 || nf_target > 256u * kMaxFine) return KC_ERR_ARG;
    if (bounds[0] != 0 || bounds[kLearnedBounds - 1] != ~0ULL) return KC_ERR_ARG;
    for (uint32_t i = 1; i < kLearnedBounds; ++i)
        if (bounds[i] < bounds[i - 1]) return KC_ERR_ARG;
    LearnedRouting lr;
    learned_routing(std::vector<uint64_t>(bounds, bounds + kLearnedBounds), nf_target, kMaxFine, lr);
    if (lr.bounds.size() > cap || (!fine_bounds && !lr.bounds.empty())) return KC_ERR_CAPACITY;
    std::memcpy(fine_base, lr.fine_base.data(), sizeof(uint32_t) * 257);
    if (!lr.bounds.empty()) std::memcpy(fine_bounds, lr.bounds.data(), sizeof(uint64_t) * lr.bounds.size());
    std::memcpy(lut, lr.lut.data(), sizeof(uint16_t) * lr.lut.size());
    return KC_OK;
}

}  // extern "C"
