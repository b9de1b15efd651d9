// kmc_noalloc.hip -- kc_count_noalloc: one pass of the counting pipeline over
// a device-resident sequence into the caller's output buffers, with every
// work buffer carved out of one static device array, so the call allocates
// nothing (no hipMalloc, no hipFree, no allocating library routine).
//
// The array lives in this file's code object; each GPU holds its own copy
// until the process ends, at the address hipGetSymbolAddress returns with
// that GPU current.  Host code only ever uses that address, never the
// array's name.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/kmc.h"
#include "kmc_pipeline.h"

using namespace kmc;

namespace {

constexpr uint64_t kScratchBytes = 64ULL << 20;  // the whole static array
constexpr uint64_t kKeyBytes = 16ULL << 20;      // each of the three key buffers A, B, R
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_dev_mu[kMaxDevices];      // one call at a time per device (they share its array)
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device
thread_local std::string g_err;         // kc_last_error(NULL) of the calling thread

int fail(int code, const std::string& what)
{
    g_err = what;
    return code;
}

// Key positions (padded to a whole extraction tile) the scratch holds at N
// key words: each key buffer holds kKeyBytes.
uint64_t max_positions(uint32_t nw) { return kKeyBytes / (8ULL * nw); }

uint64_t align256(uint64_t v) { return (v + 255) & ~255ULL; }

// Byte offsets of the work buffers inside the scratch for N key words.
struct Layout {
    uint64_t seq, A, B, R, small, ctr, foff, roff, bucket, ovf, spill_keys, spill_cnt, heavy, end;
    uint64_t key_cap, ovf_cap, spill_cap, heavy_cap, heavy_chunk;
};

Layout layout(uint32_t nw)
{
    Layout L{};
    const uint32_t nf = 256;
    L.key_cap = max_positions(nw);
    L.ovf_cap = L.key_cap / 8;
    L.spill_cap = 1ULL << 16;
    L.heavy_chunk = 1u << 18;  // heavy buckets (over 1Mi keys) cannot occur below key_cap keys
    L.heavy_cap = L.key_cap / L.heavy_chunk + nf + 1;
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = align256(at + bytes);
        return o;
    };
    L.seq = take(L.key_cap + 256);
    L.A = take(L.key_cap * 8ULL * nw);
    L.B = take(L.key_cap * 8ULL * nw);
    L.R = take(L.key_cap * 8ULL * nw);
    L.small = take(sizeof(unsigned long long) * kSmTotal);
    L.ctr = take(sizeof(unsigned long long) * kCtrCount);
    L.foff = take(sizeof(unsigned long long) * (nf + 1));
    L.roff = take(sizeof(unsigned long long) * (nf + 1));
    L.bucket = take(sizeof(uint32_t) * 4 * nf);
    L.ovf = take(L.ovf_cap * rec_bytes(nw));
    L.spill_keys = take(L.spill_cap * 8ULL * table_words(nw));
    L.spill_cnt = take(L.spill_cap * 4ULL);
    L.heavy = take(sizeof(unsigned long long) * 2 * L.heavy_cap);
    L.end = at;
    return L;
}

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_base(int dev, uint8_t** out)
{
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(KC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_ptr[dev]);
    return KC_OK;
}

template <int NW>
int count_noalloc_t(const kc_config& cfg, int dev, const char* d_seq, uint64_t seq_len, uint64_t* d_firsts,
                    uint64_t firsts_cap, void* d_records, uint64_t records_cap, kc_noalloc_result* res,
                    hipStream_t st)
{
    const Layout L = layout(NW);
    const uint64_t tile = ext_tile(NW);
    const uint64_t padded = (seq_len + tile - 1) / tile * tile;
    if (padded > L.key_cap)
        return fail(KC_ERR_CAPACITY, "input longer than kc_noalloc_max_bytes(k)");
    uint8_t* base = nullptr;
    int rc = scratch_base(dev, &base);
    if (rc) return rc;
    PipeBufs b{};
    b.seq = base + L.seq;
    b.seq_len = seq_len;
    b.padded = padded;
    b.A = reinterpret_cast<uint64_t*>(base + L.A);
    b.B = reinterpret_cast<uint64_t*>(base + L.B);
    b.R = base + L.R;
    b.key_cap = L.key_cap;
    b.small = reinterpret_cast<unsigned long long*>(base + L.small);
    b.ctr = reinterpret_cast<unsigned long long*>(base + L.ctr);
    b.foff = reinterpret_cast<unsigned long long*>(base + L.foff);
    b.roff = reinterpret_cast<unsigned long long*>(base + L.roff);
    b.bucket_u32 = reinterpret_cast<uint32_t*>(base + L.bucket);
    b.ovf = base + L.ovf;
    b.ovf_cap = L.ovf_cap;
    b.spill_keys = reinterpret_cast<unsigned long long*>(base + L.spill_keys);
    b.spill_cnt = reinterpret_cast<uint32_t*>(base + L.spill_cnt);
    b.spill_cap = L.spill_cap;
    b.clear_spill = true;  // the array keeps the previous call's contents
    b.heavy = reinterpret_cast<unsigned long long*>(base + L.heavy);
    b.heavy_cap = L.heavy_cap;
    b.firsts_out = d_firsts;
    b.recs_out = static_cast<uint8_t*>(d_records);
    b.firsts_out_cap = firsts_cap;
    b.rec_out_cap = records_cap;
    PipeOpts o{};
    o.k = cfg.k;
    o.sub = false;
    o.nf = 256;
    o.rt.mode = 0;
    o.rt.sbits = 0;
    o.rt.nf = 256;
    o.rt.rlo = 0;
    o.rt.rhi = 256;
    o.min_count = cfg.min_count ? cfg.min_count : 1u;
    o.max_count = cfg.max_count ? cfg.max_count : 0xFFFFFFFFu;
    o.pass_keys = cfg.pass_keys;
    o.counter = cfg.counter;
    o.table_slots = cfg.table_slots;
    o.heavy_keys = 1ULL << 20;
    o.heavy_chunk = (uint32_t)L.heavy_chunk;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return fail(KC_ERR_HIP, "hipDeviceGetAttribute(multiprocessor count)");
    o.heavy_workgroups = (uint32_t)cus;
    // the input is copied into the scratch, whose tail the pipeline pads with '\n'
    hipError_t e = seq_len ? hipMemcpyAsync(b.seq, d_seq, seq_len, hipMemcpyDeviceToDevice, st) : hipSuccess;
    if (e == hipSuccess) e = enqueue_pipeline<NW>(b, o, st, /*graph=*/true, nullptr);
    unsigned long long ctr[kCtrCount];
    unsigned long long smalls[3 * 257];  // coarse offsets, then the two root ranges of the records
    if (e == hipSuccess) e = hipMemcpyAsync(ctr, b.ctr, sizeof(ctr), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(smalls, b.small + kSmCoarseOff, sizeof(unsigned long long) * 257, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(smalls + 257, b.small + kSmRootOff, sizeof(unsigned long long) * 2 * 257, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(KC_ERR_HIP, std::string("kc_count_noalloc: ") + hipGetErrorString(e));
    std::memset(res, 0, sizeof(*res));
    res->total_kmers = smalls[256];
    res->firsts = ctr[kCtrFirst];
    res->records = ctr[kCtrRec];
    res->records_unfiltered = ctr[kCtrDistinct] + ctr[kCtrSpillDistinct];
    res->spilled_keys = ctr[kCtrSpilled];
    for (int p = 0; p < 2; ++p)
        for (int r = 0; r < 257; ++r) res->root_off[p][r] = smalls[257 + p * 257 + r];
    const unsigned long long err = ctr[kCtrErr];
    if (err & kErrSpillFull) return fail(KC_ERR_CAPACITY, "the scratch spill table is full (input too large for the static scratch)");
    if (err & kErrRecFull) return fail(KC_ERR_CAPACITY, "the scratch overflow-record list is full");
    if (err & kErrFirstOut) return fail(KC_ERR_CAPACITY, "first occurrences exceed firsts_cap (res->firsts holds the count)");
    if (err & kErrRecOut) return fail(KC_ERR_CAPACITY, "records exceed records_cap (res->records holds the count)");
    return KC_OK;
}

}  // namespace

namespace kmc {
const char* noalloc_last_error() { return g_err.c_str(); }
}  // namespace kmc

extern "C" {

uint64_t kc_noalloc_max_bytes(uint32_t k)
{
    const uint32_t nw = kc_words_for_k(k);
    if (!nw) return 0;
    return max_positions(nw);
}

int kc_count_noalloc(const kc_config* cfg, const char* d_seq, uint64_t seq_len, uint64_t* d_firsts, uint64_t firsts_cap,
                     void* d_records, uint64_t records_cap, kc_noalloc_result* res, void* stream)
{
    g_err.clear();
    if (!cfg || !res || (!d_seq && seq_len) || !d_firsts || !d_records) return fail(KC_ERR_ARG, "null argument");
    const uint32_t nw = kc_words_for_k(cfg->k);
    if (!nw) return fail(KC_ERR_ARG, "k must be in [1, 128]");
    if (cfg->partition != KC_PARTITION_PREFIX) return fail(KC_ERR_ARG, "kc_count_noalloc uses the prefix partition only");
    if (cfg->sub_bits != 0 && cfg->sub_bits != KC_SUB_BITS_NONE)
        return fail(KC_ERR_ARG, "kc_count_noalloc counts the 256 root buckets (sub_bits 0)");
    if (cfg->counter > KC_COUNTER_SORT || (cfg->table_slots && cfg->table_slots < 16))
        return fail(KC_ERR_ARG, "bad counter or table_slots");
    if (cfg->max_count && cfg->min_count > cfg->max_count) return fail(KC_ERR_ARG, "min_count above max_count");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(KC_ERR_NODEV, "no GPU");
    if ((int)cfg->device >= ndev || (int)cfg->device >= kMaxDevices) return fail(KC_ERR_ARG, "bad device ordinal");
    const int dev = (int)cfg->device;
    if (hipSetDevice(dev) != hipSuccess) return fail(KC_ERR_HIP, "hipSetDevice");
    if (layout(nw).end > kScratchBytes) return fail(KC_ERR_CAPACITY, "scratch layout exceeds the static array");
    std::lock_guard<std::mutex> lock(g_dev_mu[dev]);
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (nw) {
    case 1: return count_noalloc_t<1>(*cfg, dev, d_seq, seq_len, d_firsts, firsts_cap, d_records, records_cap, res, st);
    case 2: return count_noalloc_t<2>(*cfg, dev, d_seq, seq_len, d_firsts, firsts_cap, d_records, records_cap, res, st);
    default: return count_noalloc_t<4>(*cfg, dev, d_seq, seq_len, d_firsts, firsts_cap, d_records, records_cap, res, st);
    }
}

}  // extern "C"
