// kmc_kernels.hip -- the counting path as hand-written HIP for gfx950.
//
// Pipeline over the sequence bytes resident in HBM (one pass of each kernel):
//   P0  k_extract<count>    bytes -> canonical keys -> coarse-bucket histogram
//   S0  k_scan_coarse       bucket offsets, per-XCD write cursors, tile lists
//   P1  k_extract<scatter>  bytes -> keys, LDS counting sort by coarse bucket,
//                           coalesced runs into A (bucket-major, XCD-minor)
//   P2a k_sub_count         per root bucket: fine-bucket histogram
//   S1  k_scan_fine_local, k_scan_fine_add
//   P2b k_sub               A -> B grouped by fine bucket
//   P3  k_count             one workgroup per fine bucket: LDS Bloom filter
//                           (first occurrences -> firsts) and LDS open-
//                           addressing count table for repeats (hash-range
//                           passes for large buckets, a global spill table
//                           behind it); writes into the bucket's own slices
//   P3b k_spill_emit        records of keys that lived in the spill table
//   P4  k_seg_part/k_seg_apply/k_compact per-bucket outputs made contiguous: firsts in one
//                           array, records grouped by root id (word 0 >> 56)
// Workgroup-to-XCD: workgroup i runs on XCD i % 8, so write cursors are kept
// per XCD and each XCD fills its own sub-range of a bucket (partial cache lines
// merge in that XCD's L2 instead of crossing XCDs).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "kmc_device.h"
#include "kmc_kernels.h"

namespace kmc {

// ------------------------------------------------------------------ helpers

// exclusive scan of one value per thread across the workgroup; returns the
// exclusive prefix, total in *total.  scratch: >= blockDim/64 + 1 entries.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* scratch, T* total)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nwaves = (blockDim.x + kWave - 1) / kWave;
    T x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        T y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) scratch[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int w = 0; w < nwaves; ++w) {
            T t = scratch[w];
            scratch[w] = acc;
            acc += t;
        }
        scratch[nwaves] = acc;
    }
    __syncthreads();
    T r = scratch[wid] + x - v;
    *total = scratch[nwaves];
    __syncthreads();
    return r;
}

// Waits until every vector memory operation this wave has issued is complete
// (s_waitcnt vmcnt(0); the LDS/scalar and export counters are left alone).
// One counter covers both loads and stores on CDNA, and a wait for loads has
// to wait for every store issued after them whenever the number of those
// stores varies (stores under a lane condition).  A kernel that prefetches
// its next keys and then issues such stores waits here, before the stores,
// for the prefetch only; the first use of the prefetched keys then needs no
// wait at all.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Workgroup copy of n units, four loads in flight per thread before their
// stores (one load then its store per step waits out the load latency).
template <typename U>
__device__ __forceinline__ void copy_units(const U* __restrict__ src, U* __restrict__ dst, uint64_t n)
{
    const uint64_t T = blockDim.x;
    uint64_t i = threadIdx.x;
    for (; i + 3 * T < n; i += 4 * T) {
        const U a = src[i], b = src[i + T], c = src[i + 2 * T], d = src[i + 3 * T];
        dst[i] = a;
        dst[i + T] = b;
        dst[i + 2 * T] = c;
        dst[i + 3 * T] = d;
    }
    for (; i < n; i += T) dst[i] = src[i];
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    return (uint32_t)__popcll(mask & ((1ULL << lane) - 1ULL));
}

template <int NW>
__device__ __forceinline__ Key<NW> load_key(const uint64_t* p, uint64_t i)
{
    Key<NW> k;
    if constexpr (NW == 1) {
        k.w[0] = p[i];
    } else {
#pragma unroll
        for (int h = 0; h < NW / 2; ++h) {
            const ulonglong2 v = reinterpret_cast<const ulonglong2*>(p)[i * (NW / 2) + h];
            k.w[2 * h] = v.x;
            k.w[2 * h + 1] = v.y;
        }
    }
    return k;
}

template <int NW>
__device__ __forceinline__ void store_key(uint64_t* p, uint64_t i, const Key<NW>& k)
{
    if constexpr (NW == 1) {
        p[i] = k.w[0];
    } else {
#pragma unroll
        for (int h = 0; h < NW / 2; ++h) {
            ulonglong2 v;
            v.x = k.w[2 * h];
            v.y = k.w[2 * h + 1];
            reinterpret_cast<ulonglong2*>(p)[i * (NW / 2) + h] = v;
        }
    }
}

// Learned fine buckets of N keys inside root r (keys lo = r << 56 <= w0):
// the root's table is indexed by (w0 - lo) >> kLutShift, then the bounds
// inside the cell are stepped over (nb: the root's bounds, s_fb padded with
// UINT64_MAX up to index nb + 1).  Every table read of the N keys is issued
// before the first result is used (a key's reads depend on each other,
// different keys' do not).  Lanes without a key pass lo, whose cell is 0.
template <int N>
__device__ __forceinline__ void fine_of_all(const uint64_t* s_fb, const uint16_t* s_lut, uint64_t lo, uint32_t nb,
                                            const uint64_t* w0, uint32_t* f)
{
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] = s_lut[(w0[i] - lo) >> kLutShift];
    uint64_t b1[N], b2[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        b1[i] = s_fb[f[i]];
        b2[i] = s_fb[f[i] + 1];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (f[i] < nb && b1[i] <= w0[i]) {
            uint32_t x = f[i] + 1;
            if (x < nb && b2[i] <= w0[i]) {
                ++x;
                while (x < nb && s_fb[x] <= w0[i]) ++x;
            }
            f[i] = x;
        }
    }
}

// ------------------------------------------------------------------ P0 / P1

// Window ending at position i (0..15) of the thread's 16-byte word.  c/v hold
// the code/validity words from H words before the thread's word up to it
// (c[H] is the thread's own word).
template <int NW>
__device__ __forceinline__ bool window_key(const uint32_t* c, const uint32_t* v, int i, uint32_t k, Key<NW>& key);

// The 16 windows ending in a thread's word.  k > 32: window_key per window.
template <int NW>
struct Windows {
    const uint32_t* c;
    const uint32_t* v;
    uint32_t k;
    __device__ __forceinline__ Windows(const uint32_t* c_, const uint32_t* v_, uint32_t k_) : c(c_), v(v_), k(k_) {}
    __device__ __forceinline__ bool key(int i, Key<NW>& out) const { return window_key<NW>(c, v, i, k, out); }
    __device__ __forceinline__ bool root(int i, uint32_t& r) const;
};

// The root bucket of a window without its key.  The canonical key is the
// smaller of seq = pack(reverse(window)) and rev = pack(complement(window)),
// so its top byte is the smaller of their top bytes (when those are equal the
// key's is the same either way): seq's holds the window's last four bases,
// newest first, rev's the complements of its first four (for k < 4 only the
// top 2k bits, as in the key).
__device__ __forceinline__ uint32_t rev4(uint32_t b)  // reverse the four 2-bit groups of a byte
{
    b = __builtin_bitreverse32(b) >> 24;
    return ((b >> 1) & 0x55u) | ((b & 0x55u) << 1);
}

template <int NW>
__device__ __forceinline__ bool Windows<NW>::root(int i, uint32_t& r) const
{
    Key<NW> key;
    const bool ok = window_key<NW>(c, v, i, k, key);
    r = (uint32_t)(key.w[0] >> 56);
    return ok;
}

// 33 <= k <= 64: the window's forward code string F (2k bits, oldest base on
// top) as in window_key<2>, then its last and first four bases.
template <>
__device__ __forceinline__ bool Windows<2>::root(int i, uint32_t& r) const
{
    const int s = 2 * (15 - i);
    const uint64_t x0 = ((uint64_t)c[3] << 32) | c[4];
    const uint64_t x1 = ((uint64_t)c[1] << 32) | c[2];
    const uint64_t x2 = c[0];
    const uint64_t flo = s ? ((x0 >> s) | (x1 << (64 - s))) : x0;
    const uint64_t fhi = s ? ((x1 >> s) | (x2 << (64 - s))) : x1;
    const int t = 15 - i;
    const uint64_t vlo = ((uint64_t)v[1] << 48) | ((uint64_t)v[2] << 32) | ((uint64_t)v[3] << 16) | (uint64_t)v[4];
    const uint64_t vhi = v[0];
    const uint64_t W = t ? ((vlo >> t) | (vhi << (64 - t))) : vlo;
    const uint64_t km = (k == 64) ? ~0ULL : ((1ULL << k) - 1ULL);
    const uint32_t seq8 = rev4((uint32_t)flo & 0xFFu);
    const uint32_t p = 2 * k - 8;  // bit of F where its first four bases start, 58..120
    const uint32_t first = (uint32_t)(p >= 64 ? (fhi >> (p - 64)) : ((fhi << (64 - p)) | (flo >> p))) & 0xFFu;
    const uint32_t rev8 = ~first & 0xFFu;
    r = seq8 < rev8 ? seq8 : rev8;
    return (W & km) == km;
}


// k <= 32, one key word.  The reference keeps, left aligned,
//   seq = pack(reverse(window))    (src/GetKmer.h:71-73, shift right, new base on top)
//   rev = pack(complement(window)) (src/GetKmer.h:75-77, shift left, complement at the low slot)
// and takes the smaller (src/GetKmer.h:181-206).  Per thread, once: the
// 96-bit code string c0:c1:c2 (base i of c2 at bits 2(15-i)+1..2(15-i)), the
// same words with their bases in reverse order (r2:r1:r0, base i of c2 at bits
// 64+2i+1..64+2i), and a mask of the windows whose k bases are all valid (the
// validity string ANDed with itself shifted, doubling the span up to k).
// Per window: rev is a constant funnel shift of the forward string moved to
// the top and complemented, seq a constant funnel shift of the reversed
// string, both cut to the top 2k bits.
template <>
struct Windows<1> {
    uint32_t c0, c1, c2, r0, r1, r2;
    uint32_t okm;    // bit 15 - i: the window ending at base i is all valid
    uint32_t shl;    // 64 - 2k
    uint64_t top;    // the top 2k bits
    __device__ __forceinline__ Windows(const uint32_t* c, const uint32_t* v, uint32_t k)
    {
        c0 = c[0];
        c1 = c[1];
        c2 = c[2];
        r0 = rev16(c0);
        r1 = rev16(c1);
        r2 = rev16(c2);
        // base i of the thread's word at bit 15 - i, the 32 bases before it above
        uint64_t p = ((uint64_t)v[0] << 32) | ((uint64_t)v[1] << 16) | (uint64_t)v[2];
        uint32_t span = 1;  // bit b of p: bases at bits b .. b + span - 1 all valid
        while (2 * span <= k) {
            p &= p >> span;
            span *= 2;
        }
        p &= p >> (k - span);  // k - span <= span: the two spans cover k bases
        okm = (uint32_t)p & 0xFFFFu;
        shl = 64 - 2 * k;
        top = ~0ULL << shl;
    }
    __device__ __forceinline__ bool key(int i, Key<1>& out) const
    {
        const uint32_t s = 2 * (15 - i);  // 0..30, a constant once unrolled
        const uint64_t fwd = ((uint64_t)funnel32(c0, c1, s) << 32) | funnel32(c1, c2, s);
        const uint64_t rev = ~(fwd << shl) & top;
        const uint32_t sr = 2 * i + 2;  // 2..32
        const uint64_t seq = (sr == 32 ? (((uint64_t)r2 << 32) | r1)
                                       : (((uint64_t)funnel32(r2, r1, sr) << 32) | funnel32(r1, r0, sr))) & top;
        out.w[0] = seq < rev ? seq : rev;
        return (okm >> (15 - i)) & 1u;
    }
    // the key's top byte only (see Windows<NW>::root): seq's from the reversed
    // string's top word, rev's from the forward string's first four bases
    __device__ __forceinline__ bool root(int i, uint32_t& r) const
    {
        const uint32_t s = 2 * (15 - i);
        const uint64_t fwd = ((uint64_t)funnel32(c0, c1, s) << 32) | funnel32(c1, c2, s);
        const uint32_t mask8 = (uint32_t)(top >> 56);
        const uint32_t rev8 = (uint32_t)(~(fwd << shl) >> 56) & mask8;
        const uint32_t sr = 2 * i + 2;
        const uint32_t shi = sr == 32 ? r2 : funnel32(r2, r1, sr);
        const uint32_t seq8 = (shi >> 24) & mask8;
        r = seq8 < rev8 ? seq8 : rev8;
        return (okm >> (15 - i)) & 1u;
    }
};

template <>
__device__ __forceinline__ bool window_key<2>(const uint32_t* c, const uint32_t* v, int i, uint32_t k, Key<2>& key)
{
    // 160-bit code string c[0]:..:c[4]
    const int s = 2 * (15 - i);
    const uint64_t x0 = ((uint64_t)c[3] << 32) | c[4];
    const uint64_t x1 = ((uint64_t)c[1] << 32) | c[2];
    const uint64_t x2 = c[0];
    const uint64_t flo = s ? ((x0 >> s) | (x1 << (64 - s))) : x0;
    const uint64_t fhi = s ? ((x1 >> s) | (x2 << (64 - s))) : x1;
    // 80-bit validity string v[0]:..:v[4]
    const int t = 15 - i;
    const uint64_t vlo = ((uint64_t)v[1] << 48) | ((uint64_t)v[2] << 32) | ((uint64_t)v[3] << 16) | (uint64_t)v[4];
    const uint64_t vhi = v[0];
    const uint64_t W = t ? ((vlo >> t) | (vhi << (64 - t))) : vlo;
    const uint64_t km = (k == 64) ? ~0ULL : ((1ULL << k) - 1ULL);
    const bool ok = (W & km) == km;
    key = canon2(fhi, flo, k);
    return ok;
}

template <>
__device__ __forceinline__ bool window_key<4>(const uint32_t* c, const uint32_t* v, int i, uint32_t k, Key<4>& key)
{
    // 288-bit code string c[0]:..:c[8]; S = the 128 bases ending at base i of c[8]
    const int s = 2 * (15 - i);
    const uint64_t x0 = ((uint64_t)c[7] << 32) | c[8];
    const uint64_t x1 = ((uint64_t)c[5] << 32) | c[6];
    const uint64_t x2 = ((uint64_t)c[3] << 32) | c[4];
    const uint64_t x3 = ((uint64_t)c[1] << 32) | c[2];
    const uint64_t x4 = c[0];
    const uint64_t s3 = s ? ((x0 >> s) | (x1 << (64 - s))) : x0;
    const uint64_t s2 = s ? ((x1 >> s) | (x2 << (64 - s))) : x1;
    const uint64_t s1 = s ? ((x2 >> s) | (x3 << (64 - s))) : x2;
    const uint64_t s0 = s ? ((x3 >> s) | (x4 << (64 - s))) : x3;
    // 144-bit validity string v[0]:..:v[8], same alignment (one bit per base)
    const int t = 15 - i;
    const uint64_t vl = ((uint64_t)v[5] << 48) | ((uint64_t)v[6] << 32) | ((uint64_t)v[7] << 16) | (uint64_t)v[8];
    const uint64_t vm = ((uint64_t)v[1] << 48) | ((uint64_t)v[2] << 32) | ((uint64_t)v[3] << 16) | (uint64_t)v[4];
    const uint64_t vh = v[0];
    const uint64_t wlo = t ? ((vl >> t) | (vm << (64 - t))) : vl;
    const uint64_t whi = t ? ((vm >> t) | (vh << (64 - t))) : vm;
    const uint32_t kh = k - 64;  // 1..64
    const uint64_t hm = kh == 64 ? ~0ULL : ((1ULL << kh) - 1ULL);
    const bool ok = wlo == ~0ULL && (whi & hm) == hm;
    // k = 96: the 32 bases before the window, zeroed before the current run
    uint64_t pre = 0;
    if (k == 96) {
        const uint32_t vb = (uint32_t)(whi >> 32);  // bit 0: the base just before the window
        const uint32_t run = (uint32_t)__ffs(~vb);  // 1 + valid bases before the window, 0 if all 32 are
        const uint32_t keep = run ? run - 1 : 32;
        pre = keep >= 32 ? s0 : (s0 & ((1ULL << (2 * keep)) - 1ULL));
    }
    key = canon4(s0, s1, s2, s3, k, pre);
    return ok;
}

// Root bucket of a key: its top byte, for both partitions (the learned
// partition splits roots into fine buckets in P2).
template <int NW, bool SCATTER>
__global__ void __launch_bounds__(ExtractCfg<NW>::kThreads) __attribute__((amdgpu_waves_per_eu(ExtractCfg<NW>::kWaves)))
k_extract(const uint8_t* __restrict__ seq, uint64_t padded_len, uint32_t k, Route rt,
          unsigned long long* __restrict__ counts, unsigned long long* __restrict__ cursor,
          uint64_t* __restrict__ out, uint64_t tile0)
{
    constexpr int T = ExtractCfg<NW>::kThreads;
    constexpr int H = (NW == 1) ? 2 : (NW == 2 ? 4 : 8);  // words before a thread's own for a k-base window
    constexpr int TILE = T * 16;
    // The encoded tile is dead once every thread has its keys (a barrier
    // later), so it shares LDS with the staged keys.
    constexpr size_t kStage = 2 * (T + H) * sizeof(uint32_t);
    constexpr int kRounds = ExtractCfg<NW>::kStageRounds;
    constexpr uint32_t kHalf = TILE / kRounds;  // slots staged per round
    constexpr size_t kKeys = SCATTER ? kHalf * sizeof(Key<NW>) : 0;
    // the scatter's staging must let ExtractCfg::kWaves waves per SIMD share a
    // CU's 160 KiB of LDS (workgroups of T threads, T / 256 waves per SIMD each)
    static_assert(!SCATTER || ExtractCfg<NW>::kWaves == 1 ||
                      (size_t)(ExtractCfg<NW>::kWaves * 256 / T) * (kKeys + 4096 + 64) <= 160 * 1024,
                  "k_extract's staging does not allow the configured waves per SIMD");
    __shared__ __align__(16) uint8_t s_raw[kStage > kKeys ? kStage : kKeys];
    uint32_t* s_code = reinterpret_cast<uint32_t*>(s_raw);
    uint32_t* s_valid = s_code + (T + H);
    Key<NW>* s_keys = reinterpret_cast<Key<NW>*>(s_raw);
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_off[256];
    __shared__ unsigned long long s_dst[256];  // global index of the tile's slot 0 for bucket b
    __shared__ uint32_t s_scan[T / kWave + 1];

    const int t = threadIdx.x;
    // tile0: first tile of this launch (long inputs are launched in pieces)
    const uint64_t base = (tile0 + blockIdx.x) * TILE;
    const uint32_t xcd = blockIdx.x % kXcds;

    // 16 bytes per thread, aligned; the buffer is padded with '\n' to a whole tile
    {
        const uint4 q = reinterpret_cast<const uint4*>(seq + base)[t];
        uint32_t cd, vd;
        encode16(q, cd, vd);
        s_code[H + t] = cd;
        s_valid[H + t] = vd;
    }
    if (t < H) {
        const int64_t p = (int64_t)base - (int64_t)(H - t) * 16;
        uint32_t cd = 0, vd = 0;
        if (p >= 0) encode16(*reinterpret_cast<const uint4*>(seq + p), cd, vd);
        s_code[t] = cd;
        s_valid[t] = vd;
    }
    if (t < 256) s_hist[t] = 0;
    __syncthreads();

    uint32_t c[H + 1], v[H + 1];
#pragma unroll
    for (int j = 0; j <= H; ++j) {
        c[j] = s_code[t + j];
        v[j] = s_valid[t + j];
    }
    (void)padded_len;
    const Windows<NW> win(c, v, k);

    if constexpr (!SCATTER) {
        // the count pass needs each key's root only (its top byte)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t r;
            if (win.root(i, r)) atomicAdd(&s_hist[r], 1u);
        }
        __syncthreads();
        if (t < 256) {
            const uint32_t h = s_hist[t];
            if (h) atomicAdd(&counts[(size_t)xcd * 256 + t], (unsigned long long)h);
        }
    } else {
        Key<NW> keys[16];
        uint32_t rank[16];
        uint32_t okm = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            Key<NW> key;
            if (win.key(i, key)) okm |= 1u << i;
            keys[i] = key;
        }
        // the scatter of a bucket-range round keeps its own buckets only, then
        // the histogram atomics (after every key is computed)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t bk = (uint32_t)(keys[i].w[0] >> 56);
            if (bk < rt.rlo || bk >= rt.rhi) okm &= ~(1u << i);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (okm & (1u << i)) rank[i] = atomicAdd(&s_hist[(uint32_t)(keys[i].w[0] >> 56)], 1u);
        __syncthreads();
        uint32_t h = (t < 256) ? s_hist[t] : 0u;
        // the cursor reservation is issued first and its result stored only
        // after the scan and the staging: its round trip overlaps them (the
        // barriers in between wait for LDS operations only)
        unsigned long long gb = 0;
        if (h) gb = atomicAdd(&cursor[(size_t)xcd * 256 + t], (unsigned long long)h);
        uint32_t total;
        const uint32_t off = block_exclusive_scan<uint32_t>(h, s_scan, &total);
        if (t < 256) s_off[t] = off;
        __syncthreads();
        // bucket b's keys go to gb + (slot - off): one base per bucket (the
        // difference wraps modulo 2^64 and the sum with slot is exact)
        if (h) s_dst[t] = gb - (unsigned long long)off;
        // the sorted tile is staged and stored kHalf slots at a time
#pragma unroll
        for (int round = 0; round < kRounds; ++round) {
            const uint32_t r0 = (uint32_t)round * kHalf;
            if (round > 0) {
                if (total <= r0) break;  // uniform: total is the workgroup's
                __syncthreads();         // the previous round's stores have read their keys
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint32_t slot = s_off[(uint32_t)(keys[i].w[0] >> 56)] + rank[i];
                if ((okm & (1u << i)) && slot - r0 < kHalf) s_keys[slot - r0] = keys[i];
            }
            __syncthreads();
            const uint32_t end = min(total, r0 + kHalf);
            for (uint32_t slot = r0 + t; slot < end; slot += T) {
                const Key<NW> key = s_keys[slot - r0];
                store_key<NW>(out, s_dst[(uint32_t)(key.w[0] >> 56)] + slot, key);
            }
        }
    }
}

// ------------------------------------------------------------------ scans

// counts[XCDS][256] -> coarse_off[257], cursor[XCDS][256], tile_start[257]
__global__ void __launch_bounds__(256)
k_scan_coarse(const unsigned long long* __restrict__ counts, unsigned long long* __restrict__ coarse_off,
              unsigned long long* __restrict__ cursor, unsigned long long* __restrict__ tile_start,
              uint32_t sub_tile, uint32_t rlo, uint32_t rhi)
{
    // buckets outside the round [rlo, rhi) get no keys
    __shared__ unsigned long long s_scan[8];
    const int b = threadIdx.x;
    const bool in = (uint32_t)b >= rlo && (uint32_t)b < rhi;
    unsigned long long tot = 0;
    unsigned long long per[kXcds];
#pragma unroll
    for (int x = 0; x < kXcds; ++x) {
        per[x] = in ? counts[(size_t)x * 256 + b] : 0ULL;
        tot += per[x];
    }
    unsigned long long all;
    const unsigned long long off = block_exclusive_scan<unsigned long long>(tot, s_scan, &all);
    coarse_off[b] = off;
    if (b == 255) coarse_off[256] = all;
    unsigned long long run = off;
#pragma unroll
    for (int x = 0; x < kXcds; ++x) {
        cursor[(size_t)x * 256 + b] = run;
        run += per[x];
    }
    const unsigned long long tiles = (tot + sub_tile - 1) / sub_tile;
    unsigned long long tall;
    const unsigned long long toff = block_exclusive_scan<unsigned long long>(tiles, s_scan, &tall);
    tile_start[b] = toff;
    if (b == 255) tile_start[256] = tall;
}

// counts[XCDS][nf] -> fine_off[nf+1], cursor[XCDS][nf], in two launches of
// ceil(nf / 1024) workgroups (at most kScanGroups).  First: workgroup g sums
// fine buckets [1024 g, 1024 g + 1024) over the XCDs (one bucket per thread,
// coalesced), writes each bucket's offset inside the group to fine_off and the
// group's total to part[g].
__global__ void __launch_bounds__(1024)
k_scan_fine_local(const unsigned long long* __restrict__ counts, uint32_t nf, unsigned long long* __restrict__ fine_off,
                  unsigned long long* __restrict__ part)
{
    __shared__ unsigned long long s_scan[17];
    const uint32_t f = blockIdx.x * 1024 + threadIdx.x;
    unsigned long long sum = 0;
    if (f < nf) {
#pragma unroll
        for (int x = 0; x < kXcds; ++x) sum += counts[(size_t)x * nf + f];
    }
    unsigned long long all;
    const unsigned long long off = block_exclusive_scan<unsigned long long>(sum, s_scan, &all);
    if (f < nf) fine_off[f] = off;
    if (threadIdx.x == 0) part[blockIdx.x] = all;
}

// Second: each group adds the totals of the groups before it, then writes the
// per-XCD cursors of its buckets; the last group writes fine_off[nf].
__global__ void __launch_bounds__(1024)
k_scan_fine_add(const unsigned long long* __restrict__ counts, uint32_t nf, unsigned long long* __restrict__ fine_off,
                unsigned long long* __restrict__ cursor, const unsigned long long* __restrict__ part)
{
    __shared__ unsigned long long s_base;
    const uint32_t g = blockIdx.x;
    if (threadIdx.x < kWave) {
        unsigned long long v = 0;
        for (uint32_t j = threadIdx.x; j < g; j += kWave) v += part[j];
#pragma unroll
        for (int off = kWave / 2; off; off >>= 1) v += __shfl_down(v, off, kWave);
        if (threadIdx.x == 0) s_base = v;
    }
    __syncthreads();
    const uint32_t f = g * 1024 + threadIdx.x;
    if (f < nf) {
        unsigned long long run = s_base + fine_off[f];
        fine_off[f] = run;
#pragma unroll
        for (int x = 0; x < kXcds; ++x) {
            cursor[(size_t)x * nf + f] = run;
            run += counts[(size_t)x * nf + f];
        }
    }
    if (g == gridDim.x - 1 && threadIdx.x == 0) fine_off[nf] = s_base + part[g];
}

// ------------------------------------------------------------------ P2

// Root bucket of tile tid: last r with tile_start[r] <= tid (tid < tile_start[256]).
__device__ __forceinline__ uint32_t root_of_tile(const unsigned long long* tile_start, unsigned long long tid)
{
    uint32_t lo = 0, hi = 256;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_start[mid] <= tid) lo = mid; else hi = mid;
    }
    return lo;
}

// P2a: fine bucket counts per XCD slot.  The scatter pass (k_sub) reserves
// each tile's runs on the cursor of the XCD slot its workgroup gets, (tile %
// piece) % 8 with `piece` tiles per scatter launch, so each tile is counted
// into that slot.  A workgroup takes kSubCountTiles tiles 8 apart (one slot
// when piece is a multiple of 8) and keeps one LDS histogram while root and
// slot stay the same: one global add per fine bucket for up to kSubCountTiles
// tiles instead of one per tile, and the root's learned tables loaded once.
constexpr uint32_t kSubCountTiles = 8;

template <int NW>
__global__ void __launch_bounds__(SubCfg<NW>::kCountThreads)
k_sub_count(const uint64_t* __restrict__ A, const unsigned long long* __restrict__ coarse_off,
            const unsigned long long* __restrict__ tile_start, Route rt, unsigned long long* __restrict__ counts,
            uint64_t piece, uint64_t group0)
{
    constexpr int T = SubCfg<NW>::kCountThreads;
    constexpr int TILE = SubCfg<NW>::kTile;
    constexpr int PER = TILE / T;
    constexpr uint32_t NB = kMaxFine;
    __shared__ uint64_t s_fb[NB + 2];
    __shared__ __align__(16) uint16_t s_lut[kLutCells];
    __shared__ uint32_t s_hist[NB];

    const int t = threadIdx.x;
    const unsigned long long ntiles = tile_start[256];
    // tiles g0 + x + 8 j, j < kSubCountTiles: g0 = 8 kSubCountTiles (wg / 8), x = wg % 8
    const unsigned long long wg = group0 + blockIdx.x;
    const unsigned long long first = (wg / kXcds) * (kXcds * kSubCountTiles) + (wg % kXcds);
    if (first >= ntiles) return;
    const bool learned = rt.mode == 1;
    const uint32_t nf = rt.nf;
    for (uint32_t i = t; i < NB; i += T) s_hist[i] = 0;
    uint32_t r = root_of_tile(tile_start, first);
    uint32_t cur_r = ~0u, cur_x = 0, fbase = 0, nsub = 0;
    auto flush = [&]() {
        for (uint32_t i = t; i < nsub; i += T) {
            const uint32_t h = s_hist[i];
            if (h) {
                atomicAdd(&counts[(size_t)cur_x * nf + fbase + i], (unsigned long long)h);
                s_hist[i] = 0;
            }
        }
    };
    for (uint32_t j = 0; j < kSubCountTiles; ++j) {
        const unsigned long long tid = first + (unsigned long long)kXcds * j;
        if (tid >= ntiles) break;
        while (tile_start[r + 1] <= tid) ++r;
        const uint32_t x = (uint32_t)((tid % piece) % kXcds);
        if (r != cur_r || x != cur_x) {
            __syncthreads();  // the last tile's histogram atomics are done
            flush();
            cur_r = r;
            cur_x = x;
            fbase = learned ? rt.fine_base[r] : (r << rt.sbits);
            nsub = learned ? rt.fine_base[r + 1] - fbase : (1u << rt.sbits);
            if (learned) {
                // the root's bounds and two pads (fine_of_all reads up to nb + 1)
                const uint32_t nb = nsub - 1u;
                const uint64_t* fb = rt.fine_bounds + (fbase - r);
                for (uint32_t i = t; i < nb + 2; i += T) s_fb[i] = i < nb ? fb[i] : ~0ULL;
                const uint4* tab = reinterpret_cast<const uint4*>(rt.lut + (size_t)r * kLutCells);
                for (int i = t; i < (int)(kLutCells * sizeof(uint16_t) / 16); i += T) reinterpret_cast<uint4*>(s_lut)[i] = tab[i];
            }
            __syncthreads();
        }
        const unsigned long long beg = coarse_off[r] + (tid - tile_start[r]) * (unsigned long long)TILE;
        const unsigned long long endr = coarse_off[r + 1];
        const uint32_t n = (uint32_t)((endr - beg) < (unsigned long long)TILE ? (endr - beg) : TILE);
        const uint64_t flo = (uint64_t)r << 56;
        // chunks of at most 8 keys per thread (fewer registers, more waves)
        constexpr int C = PER < 8 ? PER : 8;
        constexpr int G = C < 4 ? C : 4;
#pragma unroll 1
        for (int c0 = 0; c0 < PER; c0 += C) {
            uint64_t w[C];
#pragma unroll
            for (int q = 0; q < C; ++q) {
                const uint32_t idx = t + (c0 + q) * T;
                w[q] = idx < n ? A[(beg + idx) * NW] : flo;  // word 0 routes
            }
            uint32_t sb[C];
            if (learned) {
#pragma unroll
                for (int g = 0; g < C; g += G) fine_of_all<G>(s_fb, s_lut, flo, nsub - 1u, w + g, sb + g);
            } else {
#pragma unroll
                for (int q = 0; q < C; ++q) sb[q] = (uint32_t)((w[q] >> (56 - rt.sbits)) & (nsub - 1u));
            }
#pragma unroll
            for (int q = 0; q < C; ++q)
                if (t + (c0 + q) * T < n) atomicAdd(&s_hist[sb[q]], 1u);
        }
    }
    __syncthreads();
    flush();
}

// P2b: tiles of kTile keys inside one root bucket of A, scattered to their
// fine buckets in B.  A root's fine buckets are its prefix bits (sbits of
// them) or, learned, its share of the learned partition's fine buckets (up to
// kMaxFine: u16 bucket indexes).  The sorted tile is staged in two rounds of
// half a tile each (the keys of slots [0, TILE/2), stored, then the rest):
// 56 KB of LDS, so two workgroups share a CU and one's stores overlap the
// other's loads and routing.
template <int NW>
__global__ void __launch_bounds__(SubCfg<NW>::kThreads)
k_sub(const uint64_t* __restrict__ A, const unsigned long long* __restrict__ coarse_off,
      const unsigned long long* __restrict__ tile_start, Route rt, unsigned long long* __restrict__ cursor,
      uint64_t* __restrict__ B, uint64_t tile0)
{
    constexpr int T = SubCfg<NW>::kThreads;
    constexpr int TILE = SubCfg<NW>::kTile;
    constexpr int PER = TILE / T;
    constexpr uint32_t NB = kMaxFine;                 // fine buckets per root, at most
    constexpr uint32_t E = (NB + T - 1) / T;           // histogram entries per thread in the scan
    constexpr int SR = 2;
    constexpr uint32_t STAGE = TILE / SR;
    // the root's fine bounds (NB + 1 entries, padded with UINT64_MAX; NB + 2
    // slots keep the table 16-byte aligned) and table are dead once every
    // thread has routed its keys (a barrier later): they share LDS with the
    // staged keys
    constexpr size_t kFbBytes = (NB + 2) * sizeof(uint64_t);
    static_assert(kFbBytes % 16 == 0, "the table is copied in 16-byte units");
    constexpr size_t kStage = kFbBytes + kLutCells * sizeof(uint16_t);
    constexpr size_t kKeys = STAGE * sizeof(Key<NW>);
    __shared__ __align__(16) uint8_t s_raw[kStage > kKeys ? kStage : kKeys];
    uint64_t* s_fb = reinterpret_cast<uint64_t*>(s_raw);
    uint16_t* s_lut = reinterpret_cast<uint16_t*>(s_raw + kFbBytes);
    Key<NW>* s_keys = reinterpret_cast<Key<NW>*>(s_raw);
    __shared__ uint32_t s_hist[NB];
    __shared__ uint32_t s_off[NB];
    __shared__ unsigned long long s_dst[NB];  // global index of the tile's slot 0 for fine bucket f
    __shared__ uint32_t s_scan[T / kWave + 1];
    __shared__ uint16_t s_bkt[STAGE];

    const int t = threadIdx.x;
    const unsigned long long tid = tile0 + blockIdx.x;  // tile0: first tile of this launch
    const unsigned long long ntiles = tile_start[256];
    if (tid >= ntiles) return;
    const uint32_t r = root_of_tile(tile_start, tid);
    const bool learned = rt.mode == 1;
    const uint32_t fbase = learned ? rt.fine_base[r] : (r << rt.sbits);
    const uint32_t nsub = learned ? rt.fine_base[r + 1] - fbase : (1u << rt.sbits);
    const unsigned long long beg = coarse_off[r] + (tid - tile_start[r]) * (unsigned long long)TILE;
    const unsigned long long endr = coarse_off[r + 1];
    const uint32_t n = (uint32_t)((endr - beg) < (unsigned long long)TILE ? (endr - beg) : TILE);
    const uint32_t xcd = blockIdx.x % kXcds;
    const uint32_t nf = rt.nf;
    const uint32_t nb = nsub - 1u;  // the root's bounds (learned)
    const uint64_t flo = (uint64_t)r << 56;

    for (uint32_t i = t; i < NB; i += T) s_hist[i] = 0;
    if (learned) {
        // the root's bounds and two pads (fine_of_all reads up to nb + 1)
        const uint64_t* fb = rt.fine_bounds + (fbase - r);
        for (uint32_t i = t; i < nb + 2; i += T) s_fb[i] = i < nb ? fb[i] : ~0ULL;
        const uint4* tab = reinterpret_cast<const uint4*>(rt.lut + (size_t)r * kLutCells);
        for (int i = t; i < (int)(kLutCells * sizeof(uint16_t) / 16); i += T) reinterpret_cast<uint4*>(s_lut)[i] = tab[i];
    }
    __syncthreads();

    Key<NW> keys[PER];
    uint32_t sb[PER];
    uint32_t rank[PER];
    // every load issued before the first key is routed
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t idx = t + j * T;
        if (idx < n) keys[j] = load_key<NW>(A, beg + idx);
    }
    // one wait for all of them on every path (vmcnt(0), expcnt and lgkmcnt
    // left open): the loads sit in branches, and without a wait that
    // dominates the later uses each use of a key would wait for every global
    // operation in flight, the cursor reservations below included
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (learned) {
        // learned fine buckets looked up four keys at a time (fine_of_all;
        // with 512 or 1024 fine buckets a quarter of the table cells hold a
        // bound: measured 57.7 -> 53.5 ms at 6.8e9 keys)
        constexpr int G = PER < 4 ? PER : 4;
#pragma unroll
        for (int g = 0; g < PER; g += G) {
            uint64_t w[G];
            uint32_t f[G];
#pragma unroll
            for (int u = 0; u < G; ++u) w[u] = (t + (g + u) * T < n) ? keys[g + u].w[0] : flo;
            fine_of_all<G>(s_fb, s_lut, flo, nb, w, f);
#pragma unroll
            for (int u = 0; u < G; ++u) sb[g + u] = f[u];
        }
    } else {
#pragma unroll
        for (int j = 0; j < PER; ++j) sb[j] = (uint32_t)((keys[j].w[0] >> (56 - rt.sbits)) & (nsub - 1u));
    }
    // routed first, then the atomics (as in k_extract's scatter)
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (t + j * T < n) rank[j] = atomicAdd(&s_hist[sb[j]], 1u);
    __syncthreads();

    // entries [E t, E t + E) per thread: local sums, then one workgroup scan
    uint32_t h[E];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t q = 0; q < E; ++q) {
        const uint32_t i = E * t + q;
        h[q] = i < nsub ? s_hist[i] : 0u;
        mine += h[q];
    }
    // cursor reservations issued before the scan, stored after the staging
    // (their round trips overlap both, as in k_extract)
    unsigned long long gb[E];
#pragma unroll
    for (uint32_t q = 0; q < E; ++q) {
        const uint32_t i = E * t + q;
        gb[q] = 0;
        if (h[q]) gb[q] = atomicAdd(&cursor[(size_t)xcd * nf + fbase + i], (unsigned long long)h[q]);
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan<uint32_t>(mine, s_scan, &total);
    uint32_t offq[E];
#pragma unroll
    for (uint32_t q = 0; q < E; ++q) {
        const uint32_t i = E * t + q;
        offq[q] = off;
        if (i < nsub) s_off[i] = off;
        off += h[q];
    }
    __syncthreads();
    // fine bucket f's keys go to gb + (slot - off), as in k_extract
#pragma unroll
    for (uint32_t q = 0; q < E; ++q)
        if (h[q]) s_dst[E * t + q] = gb[q] - (unsigned long long)offq[q];
    uint32_t slot[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) slot[j] = (t + j * T < n) ? s_off[sb[j]] + rank[j] : ~0u;
#pragma unroll
    for (int h2 = 0; h2 < SR; ++h2) {
        const uint32_t s0 = h2 * STAGE;
        if (h2) __syncthreads();  // the previous round's stores have read the stage
        // prefix fine buckets are key bits: no bucket index per slot then
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (slot[j] - s0 < STAGE) {  // slots of this round (~0u: no key)
                s_keys[slot[j] - s0] = keys[j];
                if (learned) s_bkt[slot[j] - s0] = (uint16_t)sb[j];
            }
        }
        __syncthreads();
        const uint32_t end = min(total, s0 + STAGE);
        if (learned) {
            for (uint32_t sl = s0 + t; sl < end; sl += T)
                store_key<NW>(B, s_dst[s_bkt[sl - s0]] + sl, s_keys[sl - s0]);
        } else {
            for (uint32_t sl = s0 + t; sl < end; sl += T) {
                const Key<NW> key = s_keys[sl - s0];
                const uint32_t f = (uint32_t)((key.w[0] >> (56 - rt.sbits)) & (nsub - 1u));
                store_key<NW>(B, s_dst[f] + sl, key);
            }
        }
    }
}

// ------------------------------------------------------------------ P3

// Runs of equal keys in consecutive lanes of a wave (the lanes read consecutive
// keys of a bucket, and a microsatellite's repeated key stays clustered through
// both scatters).  Each lane compares its key with the previous lane's through
// a DPP row shift (row_shr:1, within rows of 16 lanes), so runs are cut at row
// starts: lanes 0, 16, 32 and 48 always head a run.  Invalid lanes (past the
// end of the range) are only ever the top lanes of a step, so a valid lane's
// previous lane is valid.  Returns this lane's run length if it heads a run,
// 0 if it continues one or is invalid.
template <int NW>
__device__ __forceinline__ uint32_t run_head_length(const Key<NW>& key, bool valid)
{
    constexpr int kRowShr1 = 0x111;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    // the key's words XORed with the previous lane's (zero where equal); a
    // row's first lane always heads a run
    uint32_t diff = (lane & 15u) ? 0u : 1u;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int lo = (int)(uint32_t)key.w[w], hi = (int)(uint32_t)(key.w[w] >> 32);
        const int plo = __builtin_amdgcn_update_dpp(0, lo, kRowShr1, 0xF, 0xF, true);
        const int phi = __builtin_amdgcn_update_dpp(0, hi, kRowShr1, 0xF, 0xF, true);
        diff |= (uint32_t)(plo ^ lo) | (uint32_t)(phi ^ hi);
    }
    const bool same = valid && diff == 0u;
    const uint64_t heads = __ballot(!same);  // invalid lanes head empty runs, so no run crosses them
    if (heads == ~0ULL) return valid ? 1u : 0u;  // no repeated neighbours in the wave (the usual case)
    if (!valid || same) return 0;
    const uint64_t above = lane == 63 ? 0ULL : (heads >> (lane + 1));
    return above ? (uint32_t)__ffsll((long long)above) : 64u - lane;
}



// Two-word-safe open addressing insert.  Slot s holds key words w[0..NW) that
// only ever change from kEmpty to a value (each word claimed by CAS in order),
// so every thread inserting the same key walks the same permanent slot states
// and lands on the same slot.  Returns the slot or -1 after `limit` probes.
template <int NW, typename WordPtr>
__device__ __forceinline__ int64_t table_find_or_claim(WordPtr words, uint64_t stride, uint64_t cap_mask,
                                                       uint64_t start, uint32_t limit, const Key<NW>& key,
                                                       bool& claimed)
{
    claimed = false;
    uint64_t s = start & cap_mask;
    for (uint32_t p = 0; p < limit; ++p) {
        bool match = true;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (!match) break;
            unsigned long long* slotw = (unsigned long long*)&words[(uint64_t)w * stride + s];
            unsigned long long cur = *slotw;
            if (cur == kEmpty) {
                cur = atomicCAS(slotw, (unsigned long long)kEmpty, (unsigned long long)key.w[w]);
                if (cur == kEmpty) {
                    cur = key.w[w];
                    if (w == NW - 1) claimed = true;
                }
            }
            match = (cur == key.w[w]);
        }
        if (match) return (int64_t)s;
        s = (s + 1) & cap_mask;
    }
    return -1;
}

template <int NW>
__device__ __forceinline__ void write_record(uint8_t* recs, uint64_t i, const Key<NW>& key, uint32_t count)
{
    uint64_t* p = reinterpret_cast<uint64_t*>(recs + i * RecBytes<NW>::value);
#pragma unroll
    for (int w = 0; w < NW; ++w) p[w] = key.w[w];
    p[NW] = (uint64_t)count;  // count in the low 4 bytes, zero padding above
}

// Overflow list append (records that did not fit their bucket's region, and
// records of spilled keys).  Rare path: one global atomic per record.
template <int NW>
__device__ __forceinline__ void append_overflow(const CountArgs& a, const Key<NW>& key, uint32_t count)
{
    const unsigned long long pos = atomicAdd(a.ovf_count, 1ULL);
    if (pos < a.ovf_cap) write_record<NW>(a.ovf, pos, key, count);
    else atomicOr(a.error_flags, (unsigned long long)kErrRecFull);
}

// 64-bit hash of a key: the key words folded into one, then the SplitMix64
// finalizer.  The global tables start their probes at its low bits (every
// kernel that inserts into a table uses this one function, so a key always
// walks the same probe sequence).
template <int NW>
__device__ __forceinline__ uint64_t key_mix(const Key<NW>& k)
{
    uint64_t z = k.w[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) z ^= ((k.w[w] >> 7) ^ (uint64_t)w) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Per-key hash of the Bloom counter: the key words folded as in key_mix, then
// one multiply and xorshift (the full SplitMix64 finalizer measured 6% slower
// in k_count; this hash only places keys in the filter, the LDS table and the
// hash-range passes, so it changes no count).
template <int NW>
__device__ __forceinline__ uint64_t count_mix(const Key<NW>& k)
{
    uint64_t z = k.w[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) z ^= ((k.w[w] >> 7) ^ (uint64_t)w) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 29)) * 0xBF58476D1CE4E5B9ULL;
    return z ^ (z >> 32);
}

// Three bits of a 32-bit filter word from hash bits 13..27 (the reference
// sets three bits of a 64-bit word, src/BloomFilter.h:248-258; 32-bit words,
// twice as many in the same LDS, measured 3% faster in k_count).
__device__ __forceinline__ uint32_t bloom_mask32(uint64_t h)
{
    return (1u << ((h >> 13) & 31)) | (1u << ((h >> 18) & 31)) | (1u << ((h >> 23) & 31));
}

// Adds `add` occurrences of key to the global table.
template <int NW>
__device__ __forceinline__ void global_add(const CountArgs& args, const Key<NW>& key, uint32_t add)
{
    bool claimed;
    const uint32_t glimit = args.spill_cap < kSpillProbe ? (uint32_t)args.spill_cap : kSpillProbe;
    constexpr int TW = TableWords<NW>::value;
    const int64_t g = table_find_or_claim<TW>(args.spill_keys, args.spill_cap, args.spill_cap - 1, key_mix<NW>(key),
                                             glimit, to_table<NW>(key), claimed);
    if (g >= 0) {
        atomicAdd(&args.spill_cnt[g], add);
        if (claimed) atomicAdd(args.spilled, 1ULL);
    } else {
        atomicOr(args.error_flags, (unsigned long long)kErrSpillFull);
    }
}

// A bucket of more than args.heavy_keys keys is handed to k_heavy: thread 0
// lists its chunks; the bucket's own slices stay empty.  Returns true when the
// workgroup is done with the bucket.
__device__ __forceinline__ bool hand_over_heavy(const CountArgs& args, uint32_t f, unsigned long long lo,
                                                unsigned long long hi)
{
    const unsigned long long n = hi - lo;
    if (!args.heavy_keys || n <= args.heavy_keys) return false;
    if (threadIdx.x == 0) {
        const unsigned long long ch = args.heavy_chunk;
        const unsigned long long nch = (n + ch - 1) / ch;
        const unsigned long long base = atomicAdd(args.heavy_count, nch);
        for (unsigned long long c = 0; c < nch; ++c) {
            if (base + c >= args.heavy_cap) {  // sized so that this cannot happen
                atomicOr(args.error_flags, (unsigned long long)kErrSpillFull);
                break;
            }
            args.heavy_list[2 * (base + c)] = lo + c * ch;
            args.heavy_list[2 * (base + c) + 1] = min(hi, lo + (c + 1) * ch);
        }
        args.nfirst[f] = 0;
        args.nrec[f] = 0;
        args.ndist[f] = 0;
        atomicAdd(args.heavy_buckets, 1ULL);
    }
    return true;
}

// Open addressing over the first ts (any size) of TS LDS slots (word w of
// slot s at words[w * TS + s]), same claim protocol as table_find_or_claim;
// returns the slot or -1 after `limit` probes.  Tri: the p-th step moves p
// slots ahead (offsets 0, 1, 3, 6, ... from the home slot; needs limit <= ts),
// which builds no long runs of occupied slots the way +1 steps do: in a
// 5120-slot table filled to 3600 keys, the longest probe among 64 lanes
// averages 9 slots instead of 16 (simulated), and a wave waits for its
// longest probe.
template <int NW, uint32_t TS, bool Tri = false>
__device__ __forceinline__ int32_t lds_find_or_claim(unsigned long long* words, uint32_t ts, uint32_t start,
                                                     uint32_t limit, const Key<NW>& key, bool& claimed)
{
    claimed = false;
    uint32_t s = start;
    for (uint32_t p = 0; p < limit; ++p) {
        bool match = true;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (!match) break;
            unsigned long long* slotw = &words[(uint32_t)w * TS + s];
            unsigned long long cur = *slotw;
            if (cur == kEmpty) {
                cur = atomicCAS(slotw, (unsigned long long)kEmpty, (unsigned long long)key.w[w]);
                if (cur == kEmpty) {
                    cur = key.w[w];
                    if (w == NW - 1) claimed = true;
                }
            }
            match = (cur == key.w[w]);
        }
        if (match) return (int32_t)s;
        if constexpr (Tri) {
            s += p + 1;
            if (s >= ts) s -= ts;
        } else if (++s == ts) {
            s = 0;
        }
    }
    return -1;
}

// One workgroup per fine bucket f, keys B[fine_off[f], fine_off[f+1]).
// A bucket larger than pass_keys is processed in hash-range passes (range of
// the top 16 hash bits), each with a fresh LDS Bloom filter and count table, so
// the table holds the repeated keys of one range only.  The first range is
// sized from the bucket's occurrences, each later one from the distinct keys
// the previous range held: a bucket dominated by a few very frequent keys
// finishes in one or two passes.  First occurrences are
// written to the bucket's own slice of `firsts`; root records to the bucket's
// own slice of `rec_tmp` (n_f*8*NW bytes); no global atomics on the hot path.
template <int NW>
__global__ void __launch_bounds__(CountCfg<NW>::kThreads) __attribute__((amdgpu_waves_per_eu(CountCfg<NW>::kWaves)))
k_count(const uint64_t* __restrict__ B, const unsigned long long* __restrict__ fine_off, CountArgs args)
{
    constexpr int T = CountCfg<NW>::kThreads;
    constexpr uint32_t BW = CountCfg<NW>::kBloomWords;
    constexpr uint32_t TS = CountCfg<NW>::kTableSlots;
    __shared__ uint32_t s_bloom[BW];
    constexpr int TW = TableWords<NW>::value;
    __shared__ unsigned long long s_tkey[TW * TS];
    __shared__ uint32_t s_tcnt[TS];
    __shared__ uint32_t s_nfirst;
    __shared__ uint32_t s_nrec;
    __shared__ uint32_t s_dist;
    __shared__ uint32_t s_dpass;
    __shared__ uint32_t s_abort;
    __shared__ uint32_t s_hstart;
    __shared__ uint32_t s_hwidth;
    __shared__ uint32_t s_passes;
    // the waves per SIMD the configuration asks for must fit the CU's 160 KiB
    // of LDS: two 1024-thread workgroups per CU for 8 waves
    static_assert(CountCfg<NW>::kWaves * 64 * 4 / T == 1 ||
                      (CountCfg<NW>::kWaves * 64 * 4 / T) * (BW * 4 + TW * TS * 8 + TS * 4 + 64) <= 160 * 1024,
                  "k_count's LDS does not allow the configured waves per SIMD");
    static_assert(CountCfg<NW>::kProbe <= TS, "triangular probing wraps at most once");

    const int t = threadIdx.x;
    const uint32_t lane = t & (kWave - 1);
    const uint32_t f = args.order ? args.order[blockIdx.x] : blockIdx.x;
    const unsigned long long lo = fine_off[f], hi = fine_off[f + 1];
    const unsigned long long n = hi - lo;
    const unsigned long long rec_cap = (n * 8ULL * NW) / RecBytes<NW>::value;
    uint8_t* rec_region = args.rec_tmp + lo * 8ULL * NW;
    if (hand_over_heavy(args, f, lo, hi)) return;
    if (t == 0) {
        s_nfirst = 0;
        s_nrec = 0;
        s_dist = 0;
        s_hstart = 0;
        // first hash range sized by occurrences (a conservative bound on the
        // distinct keys); later ranges by the distinct density measured so far
        s_hwidth = n > args.pass_keys ? (uint32_t)max(1ULL, (65536ULL * args.pass_keys) / n) : 65536u;
        s_passes = 0;
    }
    constexpr int kLoads = CountCfg<NW>::kLoads;
    const unsigned long long iters = (n + (unsigned long long)T * kLoads - 1) / ((unsigned long long)T * kLoads);

    for (;;) {
        for (uint32_t i = t; i < BW; i += T) s_bloom[i] = 0u;
        for (uint32_t i = t; i < TW * TS; i += T) s_tkey[i] = kEmpty;
        for (uint32_t i = t; i < TS; i += T) s_tcnt[i] = 0u;
        // a full spill table voids the run (the host regrows it and runs again):
        // skip the remaining work
        if (t == 0) {
            s_abort = (__atomic_load_n(args.error_flags, __ATOMIC_RELAXED) & kErrSpillFull) ? 1u : 0u;
            s_dpass = 0;
        }
        __syncthreads();
        if (s_abort) break;
        const uint32_t hstart = s_hstart;
        const uint32_t hend = min(65536u, hstart + s_hwidth);
        // kLoads keys per thread and step, and the next step's keys are loaded
        // before this step's are processed: the loads stay in flight while
        // the current keys are hashed and counted
        Key<NW> nb[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const unsigned long long i = lo + (unsigned long long)u * T + t;
            if (i < hi) nb[u] = load_key<NW>(B, i);
        }
        // every step's keys are complete when the step starts (the first
        // step's here, the others before the previous step's stores)
        wait_vmem();
        for (unsigned long long it = 0; it < iters; ++it) {
            Key<NW> kb[kLoads];
#pragma unroll
            for (int u = 0; u < kLoads; ++u) kb[u] = nb[u];
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const unsigned long long i = lo + ((it + 1) * kLoads + u) * T + t;
                if (i < hi) nb[u] = load_key<NW>(B, i);
            }
            uint64_t fbal[kLoads];  // first occurrences of each key slot, one reservation per step
            // per key slot: run length (a run of r equal keys is one occurrence
            // followed by r - 1 repeats: the head lane tests the filter once and
            // counts the rest), hash (bits 0..27 filter word and bits, 32..47
            // home slot scaled to the table size, 48..63 pass range) and
            // filter bits (0: not tested)
            uint32_t run[kLoads], fmask[kLoads], fold[kLoads];
            uint64_t hk[kLoads];
            // k <= 64: the step's filter atomics issued back to back; k > 64
            // (more key registers) one key at a time
            constexpr bool kBatch = NW <= 2;
            auto filter = [&](int u) {
                const unsigned long long i = lo + (it * kLoads + u) * T + t;
                run[u] = run_head_length<NW>(kb[u], i < hi);
                hk[u] = count_mix<NW>(kb[u]);
                const uint32_t hb = (uint32_t)(hk[u] >> 48);
                fmask[u] = (run[u] && hb >= hstart && hb < hend) ? bloom_mask32(hk[u]) : 0u;
                // test-and-set of the key's bits (src/BloomFilter.h:225-234; an
                // empty mask leaves the word as it is); the reference's relaxed
                // pre-read is left out: it costs more than the atomics it saves here
                fold[u] = atomicOr(&s_bloom[(uint32_t)(hk[u] & (BW - 1))], fmask[u]);
            };
            if constexpr (kBatch) {
#pragma unroll
                for (int u = 0; u < kLoads; ++u) filter(u);
            }
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                if constexpr (!kBatch) filter(u);
                bool first = false;
                if (fmask[u]) {
                    first = (fold[u] & fmask[u]) != fmask[u];
                    const uint32_t add = run[u] - (first ? 1u : 0u);
                    if (add) {
                        bool claimed;
                        const int32_t s = lds_find_or_claim<TW, TS, true>(
                            s_tkey, TS, (uint32_t)((((hk[u] >> 32) & 0xFFFFu) * TS) >> 16), CountCfg<NW>::kProbe,
                            to_table<NW>(kb[u]), claimed);
                        if (s >= 0) atomicAdd(&s_tcnt[s], add);
                        else global_add<NW>(args, kb[u], add);
                    }
                }
                fbal[u] = __ballot(first);
            }
            // first occurrences of the step: one LDS atomic per wave, consecutive slots
            uint32_t nf = 0;
#pragma unroll
            for (int u = 0; u < kLoads; ++u) nf += (uint32_t)__popcll(fbal[u]);
            uint32_t wbase = 0;
            if (lane == 0 && nf) wbase = atomicAdd(&s_nfirst, nf);
            wbase = __builtin_amdgcn_readlane(wbase, 0);
            // the next step's keys (loaded a step ago) are waited for before
            // this step's stores are issued, not after them
            wait_vmem();
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                if ((fbal[u] >> lane) & 1ULL) store_key<NW>(args.firsts, lo + wbase + lanes_below(fbal[u]), kb[u]);
                wbase += (uint32_t)__popcll(fbal[u]);
            }
        }
        __syncthreads();
        // root records of this pass (append_export_record filter on the repeat
        // count, src/NewKmerTree.h:1159-1168)
        uint32_t dist = 0;
        for (uint32_t s0 = 0; s0 < TS; s0 += T) {
            const uint32_t s = s0 + t;
            const uint32_t cnt = s_tcnt[s];
            dist += cnt ? 1u : 0u;
            const bool keep = cnt && cnt >= args.min_count && cnt <= args.max_count;
            const uint64_t bal = __ballot(keep);
            uint32_t wbase = 0;
            if (lane == 0 && bal) wbase = atomicAdd(&s_nrec, (uint32_t)__popcll(bal));
            wbase = __shfl(wbase, 0, kWave);
            if (keep) {
                Key<NW> key;
                {
                    Key<TW> tk;
#pragma unroll
                    for (int w = 0; w < TW; ++w) tk.w[w] = s_tkey[(uint32_t)w * TS + s];
                    key = from_table<NW>(tk);
                }
                const uint64_t pos = wbase + lanes_below(bal);
                if (pos < rec_cap) write_record<NW>(rec_region, pos, key, cnt);
                else append_overflow<NW>(args, key, cnt);
            }
        }
        if (dist) atomicAdd(&s_dpass, dist);
        __syncthreads();
        if (t == 0) {
            const uint32_t d = s_dpass;
            s_dist += d;
            s_passes++;
            s_hstart = hend;
            // next range: about kPassDistinct distinct keys at the density just seen
            const uint32_t width = hend - hstart;
            uint64_t next = d ? ((uint64_t)width * CountCfg<NW>::kPassDistinct) / d : 65536ULL;
            s_hwidth = (uint32_t)(next < 1 ? 1 : (next > 65536 ? 65536 : next));
        }
        __syncthreads();
        if (s_hstart >= 65536u) break;
    }
    if (t == 0) {
        args.nfirst[f] = s_nfirst;
        args.nrec[f] = (uint32_t)(s_nrec < rec_cap ? s_nrec : rec_cap);
        args.ndist[f] = s_dist;
        if (s_passes > 1) atomicAdd(args.extra_passes, (unsigned long long)(s_passes - 1));
    }
}

// Records of keys that lived in the global spill table go to the overflow
// list.  Fixed grid, grid-stride; returns at once when nothing spilled.
// 32-bit mix of a key for the exact counter (multiply-xorshift; the final
// three steps are the "lowbias32" integer finalizer): bits 31..16 pick the
// hash range of a pass, bits 15..0 the first table slot.
template <int NW>
__device__ __forceinline__ uint32_t count_hash(const Key<NW>& k)
{
    uint32_t x = (uint32_t)(k.w[0] >> 32) * 0x9E3779B1u ^ (uint32_t)k.w[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) x = (x ^ (uint32_t)(k.w[w] >> 32)) * 0x85EBCA77u ^ (uint32_t)k.w[w];
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Exact counter: every occurrence is counted in the bucket's LDS table; a key
// seen c times leaves one first occurrence (the low.bin key) and, for c >= 2,
// the root record (key, c - 1) under the min/max filter -- the split the
// reference produces when its Bloom filter reports no false positive
// (src/FastqClassifier.h:241-349, src/NewKmerTree.h:1159-1168).
// A pass takes the keys whose hash range (bits 31..16) lies in
// [hstart, hend).  The first pass starts with the whole range, later ones with
// a width sized from the density the previous pass emitted.  hend can shrink
// while the pass runs: once, when the claimed slots reach `shrink_at`, to the
// width whose remaining keys are expected to fit; and to a key's range when
// its probe finds no slot.  Only keys below the final hend are emitted, the
// rest are counted again from scratch by the next pass.  A key in range hstart itself is in every pass's
// final range, so only such a key goes to the global spill table (whose
// records keep their full count, with no first occurrence).
template <int NW>
__global__ void __launch_bounds__(CountCfg<NW>::kThreads)
k_count_exact(const uint64_t* __restrict__ B, const unsigned long long* __restrict__ fine_off, CountArgs args)
{
    constexpr int T = CountCfg<NW>::kThreads;
    constexpr uint32_t TS = CountCfg<NW>::kExactSlots;
    constexpr int kLoads = CountCfg<NW>::kExactLoads;
    constexpr int TW = TableWords<NW>::value;
    __shared__ unsigned long long s_tkey[TW * TS];
    __shared__ uint32_t s_tcnt[TS];
    __shared__ uint32_t s_nfirst, s_nrec, s_dist, s_dpass, s_kept, s_claims, s_abort, s_hstart, s_hend, s_width, s_passes;

    const int t = threadIdx.x;
    const uint32_t lane = t & (kWave - 1);
    const uint32_t f = args.order ? args.order[blockIdx.x] : blockIdx.x;
    const unsigned long long lo = fine_off[f], hi = fine_off[f + 1];
    const unsigned long long n = hi - lo;
    const unsigned long long rec_cap = (n * 8ULL * NW) / RecBytes<NW>::value;
    uint8_t* rec_region = args.rec_tmp + lo * 8ULL * NW;
    if (hand_over_heavy(args, f, lo, hi)) return;
    // slots in use: the whole table unless a smaller one is asked for
    const uint32_t ts = args.table_slots ? min(args.table_slots, TS) : TS;
    // distinct keys a pass aims to hold: the table, or fewer when asked
    const uint32_t cap = (args.pass_keys && args.pass_keys < ts) ? (uint32_t)args.pass_keys : ts;
    const uint32_t shrink_at = cap / 2 ? cap / 2 : 1u;
    if (t == 0) {
        s_nfirst = 0;
        s_nrec = 0;
        s_dist = 0;
        s_hstart = 0;
        s_width = 65536u;
        s_passes = 0;
    }
    const unsigned long long iters = (n + (unsigned long long)T * kLoads - 1) / ((unsigned long long)T * kLoads);

    for (;;) {
#pragma unroll
        for (int w = 0; w < TW; ++w)
            for (uint32_t i = t; i < ts; i += T) s_tkey[(uint32_t)w * TS + i] = kEmpty;
        for (uint32_t i = t; i < ts; i += T) s_tcnt[i] = 0u;
        if (t == 0) {
            s_abort = (__atomic_load_n(args.error_flags, __ATOMIC_RELAXED) & kErrSpillFull) ? 1u : 0u;
            s_dpass = 0;
            s_kept = 0;
            s_claims = 0;
            s_hend = min(65536u, s_hstart + s_width);
        }
        __syncthreads();
        if (s_abort) break;
        const uint32_t hstart = s_hstart;
        for (unsigned long long it = 0; it < iters; ++it) {
            Key<NW> kb[kLoads];
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const unsigned long long i = lo + (it * kLoads + u) * T + t;
                if (i < hi) kb[u] = load_key<NW>(B, i);
            }
            // a stale (larger) hend only admits keys that the emission drops
            const uint32_t hend = __atomic_load_n(&s_hend, __ATOMIC_RELAXED);
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const unsigned long long i = lo + (it * kLoads + u) * T + t;
                // the head of a run of equal keys counts the whole run
                const uint32_t run = run_head_length<NW>(kb[u], i < hi);
                if (!run) continue;
                const uint32_t x = count_hash<NW>(kb[u]);
                const uint32_t hb = x >> 16;
                if (hb < hstart || hb >= hend) continue;
                bool claimed;
                const int32_t s = lds_find_or_claim<TW, TS>(s_tkey, ts, ((x & 0xFFFFu) * ts) >> 16,
                                                              CountCfg<NW>::kProbe, to_table<NW>(kb[u]), claimed);
                if (s >= 0) {
                    atomicAdd(&s_tcnt[s], run);
                    if (claimed) {
                        const uint32_t c = atomicAdd(&s_claims, 1u) + 1u;
                        if (c == shrink_at) {
                            // Keys are met at least as fast as the scan advances, so
                            // c / phi bounds the range's distinct keys; keep the part
                            // of the range whose remaining keys fit the free slots.
                            const double phi = (double)(it + 1) / (double)iters;
                            const double dest = (double)c / phi;
                            const double room = 0.9 * cap - c;
                            const uint32_t cur = __atomic_load_n(&s_hend, __ATOMIC_RELAXED);
                            if (dest > 0.9 * cap) {
                                const double w = (double)(cur - hstart) * room / (dest - c);
                                const uint32_t wi = w < 1.0 ? 1u : (uint32_t)w;
                                atomicMin(&s_hend, hstart + wi);
                            }
                        }
                    }
                } else if (hb > hstart) {
                    atomicMin(&s_hend, hb);  // defer this key's range to a later pass
                } else {
                    global_add<NW>(args, kb[u], run);
                }
            }
        }
        __syncthreads();
        const uint32_t hend = s_hend;
        uint32_t dist = 0, kept = 0;
        for (uint32_t s0 = 0; s0 < ts; s0 += T) {
            const uint32_t s = s0 + t;
            uint32_t cnt = 0;
            Key<NW> key;
            if (s < ts) {
                cnt = s_tcnt[s];
                if (cnt) {
                    Key<TW> tk;
#pragma unroll
                    for (int w = 0; w < TW; ++w) tk.w[w] = s_tkey[(uint32_t)w * TS + s];
                    key = from_table<NW>(tk);
                    if ((count_hash<NW>(key) >> 16) >= hend) cnt = 0;  // counted again next pass
                }
            }
            const uint64_t bf = __ballot(cnt != 0);
            uint32_t fbase = 0;
            if (lane == 0 && bf) fbase = atomicAdd(&s_nfirst, (uint32_t)__popcll(bf));
            fbase = __shfl(fbase, 0, kWave);
            if (cnt) store_key<NW>(args.firsts, lo + fbase + lanes_below(bf), key);
            dist += cnt >= 2 ? 1u : 0u;
            kept += cnt ? 1u : 0u;
            const uint32_t r = cnt - 1;
            const bool keep = cnt >= 2 && r >= args.min_count && r <= args.max_count;
            const uint64_t bal = __ballot(keep);
            uint32_t wbase = 0;
            if (lane == 0 && bal) wbase = atomicAdd(&s_nrec, (uint32_t)__popcll(bal));
            wbase = __shfl(wbase, 0, kWave);
            if (keep) {
                const uint64_t pos = wbase + lanes_below(bal);
                if (pos < rec_cap) write_record<NW>(rec_region, pos, key, r);
                else append_overflow<NW>(args, key, r);
            }
        }
        if (dist) atomicAdd(&s_dpass, dist);
        if (kept) atomicAdd(&s_kept, kept);
        __syncthreads();
        if (t == 0) {
            s_dist += s_dpass;
            s_passes++;
            s_hstart = hend;
            // next range: about 60% of the table at the density this pass emitted
            const uint64_t w = ((uint64_t)(hend - hstart) * (cap - cap * 2 / 5)) / (s_kept ? s_kept : 1u);
            s_width = (uint32_t)(w < 1 ? 1 : (w > 65536 ? 65536 : w));
        }
        __syncthreads();
        if (s_hstart >= 65536u) break;
    }
    if (t == 0) {
        args.nfirst[f] = s_nfirst;
        args.nrec[f] = (uint32_t)(s_nrec < rec_cap ? s_nrec : rec_cap);
        args.ndist[f] = s_dist;
        if (s_passes > 1) atomicAdd(args.extra_passes, (unsigned long long)(s_passes - 1));
    }
}

// ------------------------------------------------------------------ sort counter

// Sort counter (KC_COUNTER_SORT): the keys of a hash-range pass of the bucket
// are staged in LDS, sorted by a stable LSD radix sort over the bits that vary
// among them (6-bit digits; the bucket's shared prefix bits are skipped, as
// sort/ExportRecordRadixSort.h:19-24 skips its 20 known prefix bits), and
// counted by a run-length reduce over the sorted keys, as sort_and_export_leaf
// does for a leaf (src/NewKmerTree.h:1170-1250).  The split is the exact
// counter's: each distinct key leaves one first occurrence and, if it was seen
// c >= 2 times, the root record (key, c - 1) under the min/max filter.
// Passes take the keys whose top 16 hash bits lie in [hstart, hend), sized to
// fill about 3/4 of the stage; a pass that overflows is narrowed and read
// again, and a single hash value with more occurrences than the stage holds
// (one key repeated that often) is counted in the global spill table (its
// record keeps the whole count, with no first occurrence).
template <int NW>
__global__ void __launch_bounds__(kSortThreads)
k_count_sort(const uint64_t* __restrict__ B, const unsigned long long* __restrict__ fine_off, CountArgs args)
{
    constexpr int T = kSortThreads;
    constexpr uint32_t NWV = T / kWave;              // waves
    constexpr uint32_t S = SortCfg<NW>::kKeys;       // keys a pass stages
    constexpr uint32_t P = S / T;                    // keys per thread and rank step
    constexpr uint32_t NC = kSortDigits * P * NWV;   // rank counters: (digit, step, wave)
    static_assert(NC % T == 0, "counters split evenly over the threads");
    __shared__ unsigned long long s_k0[NW * S];      // word w of entry i at [w * S + i]
    __shared__ unsigned long long s_k1[NW * S];
    __shared__ uint16_t s_cnt[NC];
    __shared__ uint32_t s_scan[NWV + 1];
    __shared__ unsigned long long s_or[NW], s_and[NW];
    __shared__ uint32_t s_m, s_nfirst, s_nrec, s_dist, s_dpass, s_passes, s_hstart, s_width, s_abort;

    const int t = threadIdx.x;
    const uint32_t lane = t & (kWave - 1), wave = t / kWave;
    const uint32_t f = args.order ? args.order[blockIdx.x] : blockIdx.x;
    const unsigned long long lo = fine_off[f], hi = fine_off[f + 1];
    const unsigned long long n = hi - lo;
    const unsigned long long rec_cap = (n * 8ULL * NW) / RecBytes<NW>::value;
    uint8_t* rec_region = args.rec_tmp + lo * 8ULL * NW;
    if (hand_over_heavy(args, f, lo, hi)) return;
    // keys a pass may stage (a smaller stage is for tests); passes aim at 3/4 of it
    const uint32_t cap = args.table_slots ? min(args.table_slots, S) : S;
    const uint32_t aim = cap - cap / 4;
    if (t == 0) {
        s_nfirst = 0;
        s_nrec = 0;
        s_dist = 0;
        s_passes = 0;
        s_hstart = 0;
        s_width = n <= cap ? 65536u : (uint32_t)max(1ULL, (65536ULL * aim) / n);
    }
    const uint64_t below = (1ULL << lane) - 1ULL;

    for (;;) {
        if (t == 0) {
            s_m = 0;
            s_dpass = 0;
            s_abort = (__atomic_load_n(args.error_flags, __ATOMIC_RELAXED) & kErrSpillFull) ? 1u : 0u;
        }
        if (t < NW) {
            s_or[t] = 0ULL;
            s_and[t] = ~0ULL;
        }
        __syncthreads();
        if (s_abort) break;
        const uint32_t hstart = s_hstart;
        const uint32_t hend = min(65536u, hstart + s_width);
        // stage the pass's keys (wave-compacted appends)
        for (unsigned long long b0 = lo; b0 < hi; b0 += T) {
            const unsigned long long i = b0 + t;
            Key<NW> key;
            bool in = false;
            if (i < hi) {
                key = load_key<NW>(B, i);
                const uint32_t hb = (uint32_t)(count_mix<NW>(key) >> 48);
                in = hb >= hstart && hb < hend;
            }
            const uint64_t bal = __ballot(in);
            uint32_t wbase = 0;
            if (lane == 0 && bal) wbase = atomicAdd(&s_m, (uint32_t)__popcll(bal));
            wbase = __shfl(wbase, 0, kWave);
            const uint32_t pos = wbase + (uint32_t)__popcll(bal & below);
            if (in && pos < cap) {
#pragma unroll
                for (int w = 0; w < NW; ++w) s_k0[(uint32_t)w * S + pos] = key.w[w];
            }
        }
        __syncthreads();
        const uint32_t m = s_m;
        if (m > cap) {
            if (hend - hstart > 1) {
                // too many keys in the range: the same start, a narrower range
                if (t == 0) s_width = max(1u, (uint32_t)(((uint64_t)(hend - hstart) * aim) / m));
                __syncthreads();
                continue;
            }
            // one hash value with more keys than the stage: the global table
            // counts them (runs of equal keys in a wave row added once)
            for (unsigned long long b0 = lo; b0 < hi; b0 += T) {
                const unsigned long long i = b0 + t;
                Key<NW> key;
                if (i < hi) key = load_key<NW>(B, i);
                const uint32_t run = run_head_length<NW>(key, i < hi);
                if (run && (uint32_t)(count_mix<NW>(key) >> 48) == hstart) global_add<NW>(args, key, run);
            }
            __syncthreads();
            if (t == 0) {
                s_passes++;
                s_hstart = hend;
                s_width = 1;
            }
            __syncthreads();
            if (s_hstart >= 65536u) break;
            continue;
        }
        // bits that vary among the staged keys
        {
            unsigned long long vo[NW], va[NW];
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                vo[w] = 0ULL;
                va[w] = ~0ULL;
            }
            for (uint32_t i = t; i < m; i += T) {
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const unsigned long long x = s_k0[(uint32_t)w * S + i];
                    vo[w] |= x;
                    va[w] &= x;
                }
            }
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                if (vo[w]) atomicOr(&s_or[w], vo[w]);
                if (va[w] != ~0ULL) atomicAnd(&s_and[w], va[w]);
            }
        }
        __syncthreads();
        // stable LSD radix sort: least significant word first, 6-bit digits
        unsigned long long* src = s_k0;
        unsigned long long* dst = s_k1;
        for (int w = NW - 1; w >= 0; --w) {
            const unsigned long long vary = s_or[w] ^ s_and[w];
            if (!vary) continue;
            const uint32_t lob = (uint32_t)__ffsll((long long)vary) - 1u;
            const uint32_t hib = 63u - (uint32_t)__clzll((long long)vary);
            for (uint32_t sh = lob; sh <= hib; sh += kSortDigitBits) {
                for (uint32_t c = t; c < NC; c += T) s_cnt[c] = 0;
                __syncthreads();
                // rank among the wave's lanes of equal digit (ballot per digit bit);
                // entry i = j * T + t, so (digit, j, wave, lane) order is entry order
                uint32_t dg[P], rk[P];
#pragma unroll
                for (uint32_t j = 0; j < P; ++j) {
                    const uint32_t i = j * T + t;
                    const bool valid = i < m;
                    const uint32_t d = valid ? (uint32_t)((src[(uint32_t)w * S + i] >> sh) & (kSortDigits - 1)) : 0u;
                    uint64_t peers = __ballot(valid);
#pragma unroll
                    for (uint32_t bit = 0; bit < kSortDigitBits; ++bit) {
                        const uint64_t bb = __ballot((d >> bit) & 1u);
                        peers &= ((d >> bit) & 1u) ? bb : ~bb;
                    }
                    dg[j] = d;
                    rk[j] = (uint32_t)__popcll(peers & below);
                    if (valid && rk[j] == 0) s_cnt[(d * P + j) * NWV + wave] = (uint16_t)__popcll(peers);
                }
                __syncthreads();
                // exclusive scan of the counters in (digit, step, wave) order
                {
                    constexpr uint32_t PER = NC / T;
                    uint32_t v[PER], sum = 0;
#pragma unroll
                    for (uint32_t q = 0; q < PER; ++q) {
                        v[q] = s_cnt[t * PER + q];
                        sum += v[q];
                    }
                    uint32_t total;
                    uint32_t run = block_exclusive_scan<uint32_t>(sum, s_scan, &total);
#pragma unroll
                    for (uint32_t q = 0; q < PER; ++q) {
                        s_cnt[t * PER + q] = (uint16_t)run;
                        run += v[q];
                    }
                }
                __syncthreads();
#pragma unroll
                for (uint32_t j = 0; j < P; ++j) {
                    const uint32_t i = j * T + t;
                    if (i < m) {
                        const uint32_t pos = (uint32_t)s_cnt[(dg[j] * P + j) * NWV + wave] + rk[j];
#pragma unroll
                        for (int ww = 0; ww < NW; ++ww) dst[(uint32_t)ww * S + pos] = src[(uint32_t)ww * S + i];
                    }
                }
                __syncthreads();
                unsigned long long* x = src;
                src = dst;
                dst = x;
            }
        }
        // run-length reduce: run heads (entries that differ from their
        // predecessor), their positions listed in entry order in `dst`
        uint32_t* heads = reinterpret_cast<uint32_t*>(dst);
        {
            uint32_t hf[P], nh = 0;
#pragma unroll
            for (uint32_t q = 0; q < P; ++q) {
                const uint32_t i = t * P + q;  // blocked: each thread's entries are consecutive
                bool h = false;
                if (i < m) {
                    h = i == 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) h = h || src[(uint32_t)w * S + i] != src[(uint32_t)w * S + i - (i ? 1 : 0)];
                }
                hf[q] = h ? 1u : 0u;
                nh += hf[q];
            }
            uint32_t total;
            uint32_t at = block_exclusive_scan<uint32_t>(nh, s_scan, &total);
#pragma unroll
            for (uint32_t q = 0; q < P; ++q)
                if (hf[q]) heads[at++] = t * P + q;
            if (t == 0) heads[total] = m;
            __syncthreads();
            // one run per lane: first occurrence, and the record of c - 1
            uint32_t dist = 0;
            for (uint32_t h0 = 0; h0 < total; h0 += T) {
                const uint32_t hidx = h0 + t;
                const bool valid = hidx < total;
                uint32_t c = 0;
                Key<NW> key;
                if (valid) {
                    const uint32_t i = heads[hidx];
                    c = heads[hidx + 1] - i;
#pragma unroll
                    for (int w = 0; w < NW; ++w) key.w[w] = src[(uint32_t)w * S + i];
                }
                const uint64_t bf = __ballot(valid);
                uint32_t fbase = 0;
                if (lane == 0 && bf) fbase = atomicAdd(&s_nfirst, (uint32_t)__popcll(bf));
                fbase = __shfl(fbase, 0, kWave);
                if (valid) store_key<NW>(args.firsts, lo + fbase + (uint32_t)__popcll(bf & below), key);
                dist += c >= 2 ? 1u : 0u;
                const uint32_t r = c - 1;
                const bool keep = valid && c >= 2 && r >= args.min_count && r <= args.max_count;
                const uint64_t bal = __ballot(keep);
                uint32_t wbase = 0;
                if (lane == 0 && bal) wbase = atomicAdd(&s_nrec, (uint32_t)__popcll(bal));
                wbase = __shfl(wbase, 0, kWave);
                if (keep) {
                    const uint64_t pos = wbase + (uint32_t)__popcll(bal & below);
                    if (pos < rec_cap) write_record<NW>(rec_region, pos, key, r);
                    else append_overflow<NW>(args, key, r);
                }
            }
            if (dist) atomicAdd(&s_dpass, dist);
        }
        __syncthreads();
        if (t == 0) {
            s_dist += s_dpass;
            s_passes++;
            s_hstart = hend;
            // next range: about `aim` keys at the density this pass held
            const uint64_t wd = ((uint64_t)(hend - hstart) * aim) / (m ? m : 1u);
            s_width = (uint32_t)(wd < 1 ? 1 : (wd > 65536 ? 65536 : wd));
        }
        __syncthreads();
        if (s_hstart >= 65536u) break;
    }
    if (t == 0) {
        args.nfirst[f] = s_nfirst;
        args.nrec[f] = (uint32_t)(s_nrec < rec_cap ? s_nrec : rec_cap);
        args.ndist[f] = s_dist;
        if (s_passes > 1) atomicAdd(args.extra_passes, (unsigned long long)(s_passes - 1));
    }
}

// Chunks of the buckets handed over by the counting kernels, taken from a
// work counter by a persistent grid (one workgroup per CU).  A chunk is
// counted exactly in an LDS table (runs of equal keys in a wave counted once),
// then the table is added into the global table; k_spill_emit writes those
// keys' records with their whole count (no first occurrence, the outcome a
// Bloom false positive gives in the reference).  Every workgroup leaves the
// loop at the same chunk index, once the listed chunks are taken.
template <int NW>
__global__ void __launch_bounds__(CountCfg<NW>::kThreads)
k_heavy(const uint64_t* __restrict__ B, CountArgs args)
{
    constexpr int T = CountCfg<NW>::kThreads;
    constexpr uint32_t TS = CountCfg<NW>::kExactSlots;
    constexpr int kLoads = CountCfg<NW>::kExactLoads;
    constexpr int TW = TableWords<NW>::value;
    __shared__ unsigned long long s_tkey[TW * TS];
    __shared__ uint32_t s_tcnt[TS];
    __shared__ unsigned long long s_chunk;
    const int t = threadIdx.x;
    const unsigned long long nchunks = min((unsigned long long)args.heavy_cap, *args.heavy_count);
    for (;;) {
        if (t == 0) s_chunk = atomicAdd(args.heavy_next, 1ULL);
        __syncthreads();
        const unsigned long long c = s_chunk;
        if (c >= nchunks) break;
        for (uint32_t i = t; i < TW * TS; i += T) s_tkey[i] = kEmpty;
        for (uint32_t i = t; i < TS; i += T) s_tcnt[i] = 0u;
        __syncthreads();
        const unsigned long long lo = args.heavy_list[2 * c], hi = args.heavy_list[2 * c + 1];
        const unsigned long long iters = (hi - lo + (unsigned long long)T * kLoads - 1) / ((unsigned long long)T * kLoads);
        for (unsigned long long it = 0; it < iters; ++it) {
            Key<NW> kb[kLoads];
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const unsigned long long i = lo + (it * kLoads + u) * T + t;
                if (i < hi) kb[u] = load_key<NW>(B, i);
            }
#pragma unroll
            for (int u = 0; u < kLoads; ++u) {
                const unsigned long long i = lo + (it * kLoads + u) * T + t;
                const uint32_t run = run_head_length<NW>(kb[u], i < hi);
                if (!run) continue;
                const uint32_t x = count_hash<NW>(kb[u]);
                bool claimed;
                const int32_t s = lds_find_or_claim<TW, TS>(s_tkey, TS, ((x & 0xFFFFu) * TS) >> 16,
                                                              CountCfg<NW>::kProbe, to_table<NW>(kb[u]), claimed);
                if (s >= 0) atomicAdd(&s_tcnt[s], run);
                else global_add<NW>(args, kb[u], run);
            }
        }
        __syncthreads();
        for (uint32_t s = t; s < TS; s += T) {
            const uint32_t cnt = s_tcnt[s];
            if (!cnt) continue;
            Key<NW> key;
            {
                Key<TW> tk;
#pragma unroll
                for (int w = 0; w < TW; ++w) tk.w[w] = s_tkey[(uint32_t)w * TS + s];
                key = from_table<NW>(tk);
            }
            global_add<NW>(args, key, cnt);
        }
        __syncthreads();  // the table is cleared for the next chunk
    }
}

// Each workgroup owns a contiguous range of slots.  A first sweep counts the
// records each thread keeps; one scan and one global reservation per workgroup
// place them; a second sweep writes them.  (Reserving once per 256 slots put
// tens of thousands of atomics on the one overflow counter: 0.79 ms for an
// 8Mi-slot table, most of it waiting on that address.)
template <int NW>
__global__ void __launch_bounds__(256)
k_spill_emit(CountArgs args)
{
    __shared__ uint32_t s_scan[256 / kWave + 1];
    __shared__ uint32_t s_dist;
    __shared__ unsigned long long s_base;
    if (*args.spilled == 0) return;
    const uint64_t per = (args.spill_cap + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * per;
    const uint64_t hi = lo + per < args.spill_cap ? lo + per : args.spill_cap;
    if (threadIdx.x == 0) s_dist = 0;
    uint32_t mine = 0, dist = 0;
    for (uint64_t s = lo + threadIdx.x; s < hi; s += 256) {
        const uint32_t cnt = args.spill_cnt[s];
        dist += cnt != 0;
        mine += cnt != 0 && cnt >= args.min_count && cnt <= args.max_count;
    }
    uint32_t total = 0;
    const uint32_t off = block_exclusive_scan(mine, s_scan, &total);  // its barriers order the s_dist clear
    if (dist) atomicAdd(&s_dist, dist);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base = total ? atomicAdd(args.ovf_count, (unsigned long long)total) : 0ULL;
        if (s_dist) atomicAdd(args.spill_distinct, (unsigned long long)s_dist);
    }
    __syncthreads();
    if (!mine) return;
    uint64_t pos = s_base + off;
    constexpr int TW = TableWords<NW>::value;
    for (uint64_t s = lo + threadIdx.x; s < hi; s += 256) {
        const uint32_t cnt = args.spill_cnt[s];
        if (!(cnt != 0 && cnt >= args.min_count && cnt <= args.max_count)) continue;
        Key<TW> tk;
#pragma unroll
        for (int w = 0; w < TW; ++w) tk.w[w] = args.spill_keys[(uint64_t)w * args.spill_cap + s];
        if (pos < args.ovf_cap) write_record<NW>(args.ovf, pos, from_table<NW>(tk), cnt);
        else atomicOr(args.error_flags, (unsigned long long)kErrRecFull);
        ++pos;
    }
}

// ------------------------------------------------------------------ P4

// overflow records: per-root histogram (single workgroup, the list is short)
template <int NW>
__global__ void __launch_bounds__(256)
k_ovf_hist(const uint8_t* __restrict__ ovf, const unsigned long long* __restrict__ n_ptr, uint64_t cap,
           unsigned long long* __restrict__ root_ovf)
{
    // root histogram of the overflow records: LDS per workgroup, then one
    // global add per (workgroup, root); root_ovf starts zeroed
    __shared__ uint32_t s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t n = *n_ptr < cap ? *n_ptr : cap;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint64_t w0 = *reinterpret_cast<const uint64_t*>(ovf + i * RecBytes<NW>::value);
        atomicAdd(&s_h[w0 >> 56], 1u);
    }
    __syncthreads();
    if (s_h[threadIdx.x]) atomicAdd(&root_ovf[threadIdx.x], (unsigned long long)s_h[threadIdx.x]);
}

// Exclusive scans of the per-bucket outputs, in two launches over chunks of
// 1024 buckets (one workgroup walking every bucket took 0.32 ms at 65536
// buckets): the chunks' sums (and the distinct and largest-bucket counters),
// then each chunk's scan behind the sums of the chunks before it.  foff/roff
// run over buckets in key-range order; the overflow records' root ranges
// follow all in-bucket records.
__global__ void __launch_bounds__(1024)
k_seg_part(SegArgs a)
{
    __shared__ unsigned long long s_scan[1024 / kWave + 1];
    __shared__ unsigned long long s_max;
    const uint32_t t = threadIdx.x;
    const uint32_t f = blockIdx.x * 1024u + t;
    if (t == 0) s_max = 0;
    unsigned long long sf = 0, sr = 0, sd = 0, mx = 0;
    if (f < a.nf) {
        sf = a.nfirst[f];
        sr = a.nrec[f];
        sd = a.ndist[f];
        mx = a.fine_off[f + 1] - a.fine_off[f];
    }
    unsigned long long tf, tr, td;
    (void)block_exclusive_scan<unsigned long long>(sf, s_scan, &tf);
    (void)block_exclusive_scan<unsigned long long>(sr, s_scan, &tr);
    (void)block_exclusive_scan<unsigned long long>(sd, s_scan, &td);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
    }
    // largest bucket: one LDS atomic per wave, one global one per workgroup
    // (the barriers of the scans above order s_max's clear before these)
    if ((t & (kWave - 1)) == 0 && mx) atomicMax(&s_max, mx);
    __syncthreads();
    if (t == 0) {
        if (s_max) atomicMax(&a.ctr[kCtrMaxBucket], s_max);
        a.part[2 * blockIdx.x] = tf;
        a.part[2 * blockIdx.x + 1] = tr;
        if (td) atomicAdd(&a.ctr[kCtrDistinct], td);
    }
}

__global__ void __launch_bounds__(1024)
k_seg_apply(SegArgs a, uint32_t nchunks)
{
    __shared__ unsigned long long s_scan[1024 / kWave + 1];
    __shared__ unsigned long long s_base[2];
    const uint32_t t = threadIdx.x;
    const uint32_t g = blockIdx.x;
    // this chunk's bases and the totals, from the chunk sums (nchunks <= 1024)
    const unsigned long long pf = t < nchunks ? a.part[2 * t] : 0ULL;
    const unsigned long long pr = t < nchunks ? a.part[2 * t + 1] : 0ULL;
    unsigned long long tf, tr;
    const unsigned long long xf = block_exclusive_scan<unsigned long long>(pf, s_scan, &tf);
    const unsigned long long xr = block_exclusive_scan<unsigned long long>(pr, s_scan, &tr);
    if (t == g) {
        s_base[0] = xf;
        s_base[1] = xr;
    }
    __syncthreads();
    const uint32_t f = g * 1024u + t;
    const unsigned long long vf = f < a.nf ? a.nfirst[f] : 0ULL;
    const unsigned long long vr = f < a.nf ? a.nrec[f] : 0ULL;
    unsigned long long cf, cr;
    const unsigned long long ef = block_exclusive_scan<unsigned long long>(vf, s_scan, &cf);
    const unsigned long long er = block_exclusive_scan<unsigned long long>(vr, s_scan, &cr);
    if (f < a.nf) {
        a.foff[f] = s_base[0] + ef;
        a.roff[f] = s_base[1] + er;
    }
    if (g == nchunks - 1 && t == 0) {
        a.foff[a.nf] = tf;
        a.roff[a.nf] = tr;
    }
    if (g == 0) {  // uniform: the workgroup's barriers below are all taken
        const unsigned long long o = t < 256 ? a.root_ovf[t] : 0ULL;
        unsigned long long all;
        const unsigned long long ooff = block_exclusive_scan<unsigned long long>(o, s_scan, &all);
        if (t < 256) {
            a.ovf_off[t] = tr + ooff;
            a.ovf_cursor[t] = tr + ooff;
            if (t == 255) a.ovf_off[256] = tr + all;
        }
        if (t == 0) {
            a.ctr[kCtrFirst] = tf;
            a.ctr[kCtrRec] = tr + all;
            if (tr + all > a.rec_out_cap) atomicOr(&a.ctr[kCtrErr], kErrRecOut);
            if (tf > a.firsts_out_cap) atomicOr(&a.ctr[kCtrErr], kErrFirstOut);
        }
    }
}

// Segmented copies of bucket f: its firsts to the contiguous firsts array,
// then (second kernel) its records to [roff[f], roff[f] + nrec[f]), grouped
// by root when the bucket's key range spans several roots.
template <int NW>
__global__ void __launch_bounds__(256)
k_compact_firsts(SegArgs a, const uint64_t* __restrict__ firsts_in, uint64_t* __restrict__ firsts_out)
{
    const uint32_t f = blockIdx.x;
    if (a.ctr[kCtrErr] & kErrFirstOut) return;
    const unsigned long long lo = a.fine_off[f];
    const uint32_t nfst = a.nfirst[f];
    const unsigned long long fdst = a.foff[f];
    if constexpr (NW == 1)
        copy_units(firsts_in + lo, firsts_out + fdst, nfst);
    else  // keys of two or four words: 16-byte units
        copy_units(reinterpret_cast<const ulonglong2*>(firsts_in) + lo * (NW / 2),
                   reinterpret_cast<ulonglong2*>(firsts_out) + fdst * (NW / 2), (uint64_t)nfst * (NW / 2));
}

template <int NW>
__global__ void __launch_bounds__(256)
k_compact_recs(SegArgs a, const uint8_t* __restrict__ rec_in, uint8_t* __restrict__ rec_out)
{
    constexpr uint32_t RW = RecBytes<NW>::value / 8;
    __shared__ uint32_t s_min, s_max;
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_scan[5];
    const uint32_t t = threadIdx.x;
    const uint32_t f = blockIdx.x;
    const uint32_t nr = a.nrec[f];
    if (nr == 0 || (a.ctr[kCtrErr] & kErrRecOut)) return;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(rec_in) + a.fine_off[f] * NW;
    uint64_t* dst = reinterpret_cast<uint64_t*>(rec_out) + a.roff[f] * RW;
    if (t == 0) { s_min = 255; s_max = 0; }
    s_off[t] = 0;
    __syncthreads();
    uint32_t lmin = 255, lmax = 0;
    for (uint32_t i = t; i < nr; i += 256) {
        const uint32_t r = (uint32_t)(src[(size_t)i * RW] >> 56);
        lmin = r < lmin ? r : lmin;
        lmax = r > lmax ? r : lmax;
    }
    atomicMin(&s_min, lmin);
    atomicMax(&s_max, lmax);
    __syncthreads();
    const uint32_t rmin = s_min, rmax = s_max;
    if (rmin == rmax) {
        copy_units(src, dst, (uint64_t)nr * RW);
        if (t == 0) atomicAdd(&a.root_main[rmin], (unsigned long long)nr);
        return;
    }
    for (uint32_t i = t; i < nr; i += 256) atomicAdd(&s_off[(uint32_t)(src[(size_t)i * RW] >> 56) - rmin], 1u);
    __syncthreads();
    const uint32_t c = s_off[t];
    uint32_t tot;
    const uint32_t off = block_exclusive_scan<uint32_t>(c, s_scan, &tot);
    s_off[t] = off;
    if (c) atomicAdd(&a.root_main[rmin + t], (unsigned long long)c);
    __syncthreads();
    for (uint32_t i = t; i < nr; i += 256) {
        const uint32_t r = (uint32_t)(src[(size_t)i * RW] >> 56) - rmin;
        const uint32_t pos = atomicAdd(&s_off[r], 1u);
#pragma unroll
        for (uint32_t w = 0; w < RW; ++w) dst[(size_t)pos * RW + w] = src[(size_t)i * RW + w];
    }
}

// root ranges of the in-bucket records
__global__ void __launch_bounds__(256)
k_root_scan(SegArgs a)
{
    __shared__ unsigned long long s_scan[5];
    unsigned long long all;
    const unsigned long long off = block_exclusive_scan<unsigned long long>(a.root_main[threadIdx.x], s_scan, &all);
    a.root_off[threadIdx.x] = off;
    if (threadIdx.x == 255) a.root_off[256] = all;
}

// Overflow records into their root's tail range: tiles of 2048 records are
// ranked by root in LDS, and each root's part of a tile reserves its range
// with one global add.
template <int NW>
__global__ void __launch_bounds__(256)
k_ovf_scatter(const uint8_t* __restrict__ ovf, const unsigned long long* __restrict__ n_ptr, uint64_t cap,
              unsigned long long* __restrict__ ovf_cursor, uint8_t* __restrict__ rec_out,
              const unsigned long long* __restrict__ ctr)
{
    constexpr uint32_t RW = RecBytes<NW>::value / 8;
    constexpr uint32_t kPer = 8;
    constexpr uint32_t kTile = 256 * kPer;
    __shared__ uint32_t s_h[256];
    __shared__ unsigned long long s_base[256];
    if (ctr[kCtrErr] & kErrRecOut) return;
    const uint64_t n = *n_ptr < cap ? *n_ptr : cap;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kTile; t0 < n; t0 += (uint64_t)gridDim.x * kTile) {
        s_h[threadIdx.x] = 0;
        __syncthreads();
        uint32_t root[kPer], rank[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint64_t i = t0 + j * 256 + threadIdx.x;
            if (i < n) {
                root[j] = (uint32_t)(*reinterpret_cast<const uint64_t*>(ovf + i * (RW * 8)) >> 56);
                rank[j] = atomicAdd(&s_h[root[j]], 1u);
            }
        }
        __syncthreads();
        const uint32_t h = s_h[threadIdx.x];
        if (h) s_base[threadIdx.x] = atomicAdd(&ovf_cursor[threadIdx.x], (unsigned long long)h);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint64_t i = t0 + j * 256 + threadIdx.x;
            if (i < n) {
                const uint64_t* src = reinterpret_cast<const uint64_t*>(ovf + i * (RW * 8));
                uint64_t* d = reinterpret_cast<uint64_t*>(rec_out + (s_base[root[j]] + rank[j]) * (RW * 8));
#pragma unroll
                for (uint32_t w = 0; w < RW; ++w) d[w] = src[w];
            }
        }
        __syncthreads();  // s_h and s_base are reused by the next tile
    }
}

// ------------------------------------------------------------------ launchers

#define KMC_LAUNCH_CHECK() do { hipError_t e_ = hipGetLastError(); if (e_ != hipSuccess) return e_; } while (0)

// Workgroups of T threads per launch of a tile-per-workgroup kernel: a grid
// dimension must stay below 2^32 threads, so inputs of 64 GiB and more of
// sequence are launched in pieces (2^31 threads each).  KMC_MAX_GRID_TILES
// lowers the piece size (tests of the piece logic).
uint64_t launch_piece_tiles(uint32_t T)
{
    uint64_t piece = (1ULL << 31) / T;
    if (const char* v = std::getenv("KMC_MAX_GRID_TILES")) {
        const unsigned long long n = std::strtoull(v, nullptr, 10);
        if (n > 0 && n < piece) piece = n;
    }
    return piece;
}

template <int NW>
hipError_t launch_extract(bool scatter, const uint8_t* seq, uint64_t padded_len, uint32_t k, const Route& rt,
                          unsigned long long* counts, unsigned long long* cursor, uint64_t* out, hipStream_t st)
{
    constexpr int T = ExtractCfg<NW>::kThreads;
    const uint64_t tiles = padded_len / (T * 16);
    const uint64_t piece = launch_piece_tiles(T);
    for (uint64_t t0 = 0; t0 < tiles; t0 += piece) {
        const dim3 g((uint32_t)std::min<uint64_t>(piece, tiles - t0));
        if (scatter)
            hipLaunchKernelGGL((k_extract<NW, true>), g, dim3(T), 0, st, seq, padded_len, k, rt, counts, cursor, out, t0);
        else
            hipLaunchKernelGGL((k_extract<NW, false>), g, dim3(T), 0, st, seq, padded_len, k, rt, counts, cursor, out, t0);
        KMC_LAUNCH_CHECK();
    }
    return hipSuccess;
}

hipError_t launch_scan_coarse(const unsigned long long* counts, unsigned long long* coarse_off, unsigned long long* cursor,
                              unsigned long long* tile_start, uint32_t sub_tile, hipStream_t st, uint32_t rlo, uint32_t rhi)
{
    hipLaunchKernelGGL(k_scan_coarse, dim3(1), dim3(256), 0, st, counts, coarse_off, cursor, tile_start, sub_tile, rlo, rhi);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_scan_fine(const unsigned long long* counts, uint32_t nf, unsigned long long* fine_off,
                            unsigned long long* cursor, unsigned long long* part, hipStream_t st)
{
    const uint32_t groups = (nf + 1023) / 1024;
    if (groups == 0 || groups > kScanGroups) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_fine_local, dim3(groups), dim3(1024), 0, st, counts, nf, fine_off, part);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_fine_add, dim3(groups), dim3(1024), 0, st, counts, nf, fine_off, cursor, part);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template <int NW>
hipError_t launch_sub(bool scatter, uint64_t max_tiles, const uint64_t* A, const unsigned long long* coarse_off,
                      const unsigned long long* tile_start, const Route& rt, unsigned long long* counts,
                      unsigned long long* cursor, uint64_t* B, hipStream_t st)
{
    if (rt.mode == 0 && (rt.sbits == 0 || rt.sbits > kMaxSubBits)) return hipErrorInvalidValue;
    // tiles per scatter launch: the count pass files each tile under the XCD
    // slot the scatter gives it
    const uint64_t piece = launch_piece_tiles(SubCfg<NW>::kThreads);
    if (!scatter) {
        // workgroups of kSubCountTiles tiles 8 apart: ceil(tiles / (8 G)) * 8
        const uint64_t span = (uint64_t)kXcds * kSubCountTiles;
        const uint64_t groups = (max_tiles + span - 1) / span * kXcds;
        const uint64_t gpiece = std::max<uint64_t>(kXcds, launch_piece_tiles(SubCfg<NW>::kCountThreads) / kXcds * kXcds);
        for (uint64_t g0 = 0; g0 < groups; g0 += gpiece) {
            const dim3 g((uint32_t)std::min<uint64_t>(gpiece, groups - g0));
            hipLaunchKernelGGL((k_sub_count<NW>), g, dim3(SubCfg<NW>::kCountThreads), 0, st, A, coarse_off, tile_start,
                               rt, counts, piece, g0);
            KMC_LAUNCH_CHECK();
        }
        return hipSuccess;
    }
    for (uint64_t t0 = 0; t0 < max_tiles; t0 += piece) {
        const dim3 g((uint32_t)std::min<uint64_t>(piece, max_tiles - t0));
        hipLaunchKernelGGL((k_sub<NW>), g, dim3(SubCfg<NW>::kThreads), 0, st, A, coarse_off, tile_start, rt, cursor, B, t0);
        KMC_LAUNCH_CHECK();
    }
    return hipSuccess;
}

// Workgroup order of the counting kernels: buckets by size class, largest
// first, so the largest buckets start first and the small ones fill the end
// of the grid instead of a large bucket starting last and running alone (a
// workgroup owns one CU for its whole bucket).  Classes are quarter powers of
// two; the order within a class is arbitrary and changes no result.  Two
// launches over 1024 buckets per workgroup: the class counts (one LDS atomic
// per wave and class, one global add per workgroup and class), then each
// workgroup scans the 256 counts, reserves its share of each class behind a
// per-class global cursor and places its buckets there (one workgroup doing it
// all took 0.25 ms at 65536 buckets).
__device__ __forceinline__ uint32_t size_class(unsigned long long n)
{
    if (n == 0) return 255;  // empty buckets last
    const uint32_t top = 63 - __clzll(n);
    const uint32_t frac = top >= 2 ? (uint32_t)(n >> (top - 2)) & 3u : (uint32_t)(n << (2 - top)) & 3u;
    return 254 - min(254u, 4 * top + frac);
}

// The wave's lanes grouped by class: for each class present, fn(class, lanes)
// runs once on every lane of the wave (uniform arguments).
template <typename Fn>
__device__ __forceinline__ void for_each_class(uint32_t c, bool valid, Fn&& fn)
{
    uint64_t pending = __ballot(valid);
    while (pending) {
        const uint32_t lead = (uint32_t)__ffsll((unsigned long long)pending) - 1;
        const uint32_t lc = (uint32_t)__builtin_amdgcn_readlane((int)c, (int)lead);
        const uint64_t same = __ballot(c == lc) & pending;
        fn(lc, same);
        pending &= ~same;
    }
}

__global__ void __launch_bounds__(1024)
k_order_count(const unsigned long long* __restrict__ fine_off, uint32_t nf, uint32_t* __restrict__ ghist)
{
    __shared__ uint32_t s_cnt[256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & (kWave - 1);
    if (t < 256) s_cnt[t] = 0;
    __syncthreads();
    const uint32_t f = blockIdx.x * 1024u + t;
    const uint32_t c = f < nf ? size_class(fine_off[f + 1] - fine_off[f]) : 0u;
    for_each_class(c, f < nf, [&](uint32_t lc, uint64_t same) {
        if (lane == 0) atomicAdd(&s_cnt[lc], (uint32_t)__popcll(same));
    });
    __syncthreads();
    if (t < 256 && s_cnt[t]) atomicAdd(&ghist[t], s_cnt[t]);
}

__global__ void __launch_bounds__(1024)
k_order_place(const unsigned long long* __restrict__ fine_off, uint32_t nf, const uint32_t* __restrict__ ghist,
              uint32_t* __restrict__ gcursor, uint32_t* __restrict__ order)
{
    __shared__ uint32_t s_cnt[256];   // the workgroup's class counts, then its cursors
    __shared__ uint32_t s_base[256];  // where the workgroup's buckets of a class start
    __shared__ uint32_t s_scan[1024 / kWave + 1];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & (kWave - 1);
    if (t < 256) s_cnt[t] = 0;
    uint32_t total;
    const uint32_t off = block_exclusive_scan<uint32_t>(t < 256 ? ghist[t] : 0u, s_scan, &total);
    const uint32_t f = blockIdx.x * 1024u + t;
    const uint32_t c = f < nf ? size_class(fine_off[f + 1] - fine_off[f]) : 0u;
    for_each_class(c, f < nf, [&](uint32_t lc, uint64_t same) {
        if (lane == 0) atomicAdd(&s_cnt[lc], (uint32_t)__popcll(same));
    });
    __syncthreads();
    // one global reservation per class and workgroup (a reservation per wave
    // made every wave of the grid wait on the same few cursors)
    if (t < 256) {
        const uint32_t n = s_cnt[t];
        s_base[t] = n ? off + atomicAdd(&gcursor[t], n) : 0u;
        s_cnt[t] = 0;
    }
    __syncthreads();
    for_each_class(c, f < nf, [&](uint32_t lc, uint64_t same) {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&s_cnt[lc], (uint32_t)__popcll(same));
        at = (uint32_t)__builtin_amdgcn_readlane((int)at, 0);
        if ((same >> lane) & 1ULL) order[s_base[lc] + at + lanes_below(same)] = f;
    });
}

hipError_t launch_bucket_order(uint32_t nf, const unsigned long long* fine_off, uint32_t* order, uint32_t* scratch,
                               hipStream_t st)
{
    const uint32_t grid = (nf + 1023) / 1024;
    hipLaunchKernelGGL(k_order_count, dim3(grid), dim3(1024), 0, st, fine_off, nf, scratch);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_order_place, dim3(grid), dim3(1024), 0, st, fine_off, nf, scratch, scratch + 256, order);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template <int NW>
hipError_t launch_count(uint32_t nf, uint32_t counter, const uint64_t* B, const unsigned long long* fine_off,
                        const CountArgs& args, hipStream_t st)
{
    if (counter == kCounterSort)
        hipLaunchKernelGGL((k_count_sort<NW>), dim3(nf), dim3(kSortThreads), 0, st, B, fine_off, args);
    else if (counter == kCounterExact)
        hipLaunchKernelGGL((k_count_exact<NW>), dim3(nf), dim3(CountCfg<NW>::kThreads), 0, st, B, fine_off, args);
    else
        hipLaunchKernelGGL((k_count<NW>), dim3(nf), dim3(CountCfg<NW>::kThreads), 0, st, B, fine_off, args);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template <int NW>
hipError_t launch_heavy(uint32_t workgroups, const uint64_t* B, const CountArgs& args, hipStream_t st)
{
    hipLaunchKernelGGL((k_heavy<NW>), dim3(workgroups), dim3(CountCfg<NW>::kThreads), 0, st, B, args);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template <int NW>
hipError_t launch_spill_emit(const CountArgs& args, hipStream_t st)
{
    hipLaunchKernelGGL((k_spill_emit<NW>), dim3(1024), dim3(256), 0, st, args);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template <int NW>
hipError_t launch_finalize(const SegArgs& sa, const CountArgs& ca, const uint64_t* firsts_in, uint64_t* firsts_out,
                           const uint8_t* rec_in, uint8_t* rec_out, hipStream_t st)
{
    hipLaunchKernelGGL((k_ovf_hist<NW>), dim3(256), dim3(256), 0, st, ca.ovf, ca.ovf_count, ca.ovf_cap, sa.root_ovf);
    KMC_LAUNCH_CHECK();
    const uint32_t nchunks = (sa.nf + 1023) / 1024;
    hipLaunchKernelGGL(k_seg_part, dim3(nchunks), dim3(1024), 0, st, sa);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_seg_apply, dim3(nchunks), dim3(1024), 0, st, sa, nchunks);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_compact_firsts<NW>), dim3(sa.nf), dim3(256), 0, st, sa, firsts_in, firsts_out);
    KMC_LAUNCH_CHECK();
    // records go into the buffer the firsts were just copied out of
    hipLaunchKernelGGL((k_compact_recs<NW>), dim3(sa.nf), dim3(256), 0, st, sa, rec_in, rec_out);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_root_scan, dim3(1), dim3(256), 0, st, sa);
    KMC_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_ovf_scatter<NW>), dim3(256), dim3(256), 0, st, ca.ovf, ca.ovf_count, ca.ovf_cap, sa.ovf_cursor, rec_out, sa.ctr);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}

template hipError_t launch_extract<1>(bool, const uint8_t*, uint64_t, uint32_t, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_extract<2>(bool, const uint8_t*, uint64_t, uint32_t, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_extract<4>(bool, const uint8_t*, uint64_t, uint32_t, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_sub<1>(bool, uint64_t, const uint64_t*, const unsigned long long*, const unsigned long long*, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_sub<2>(bool, uint64_t, const uint64_t*, const unsigned long long*, const unsigned long long*, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_sub<4>(bool, uint64_t, const uint64_t*, const unsigned long long*, const unsigned long long*, const Route&, unsigned long long*, unsigned long long*, uint64_t*, hipStream_t);
template hipError_t launch_count<1>(uint32_t, uint32_t, const uint64_t*, const unsigned long long*, const CountArgs&, hipStream_t);
template hipError_t launch_count<2>(uint32_t, uint32_t, const uint64_t*, const unsigned long long*, const CountArgs&, hipStream_t);
template hipError_t launch_count<4>(uint32_t, uint32_t, const uint64_t*, const unsigned long long*, const CountArgs&, hipStream_t);
template hipError_t launch_heavy<1>(uint32_t, const uint64_t*, const CountArgs&, hipStream_t);
template hipError_t launch_heavy<2>(uint32_t, const uint64_t*, const CountArgs&, hipStream_t);
template hipError_t launch_heavy<4>(uint32_t, const uint64_t*, const CountArgs&, hipStream_t);
template hipError_t launch_spill_emit<1>(const CountArgs&, hipStream_t);
template hipError_t launch_spill_emit<2>(const CountArgs&, hipStream_t);
template hipError_t launch_spill_emit<4>(const CountArgs&, hipStream_t);
template hipError_t launch_finalize<1>(const SegArgs&, const CountArgs&, const uint64_t*, uint64_t*, const uint8_t*, uint8_t*, hipStream_t);
template hipError_t launch_finalize<2>(const SegArgs&, const CountArgs&, const uint64_t*, uint64_t*, const uint8_t*, uint8_t*, hipStream_t);
template hipError_t launch_finalize<4>(const SegArgs&, const CountArgs&, const uint64_t*, uint64_t*, const uint8_t*, uint8_t*, hipStream_t);
// ------------------------------------------------------------------ histogram

// Count histogram of the resident results, histogram_tool's two modes
// (tool/histogram_tool.cpp:445-527): hist[f - minf] for f in [minf, maxf].
// Bins below minf + kHistLds are summed in LDS and added once per workgroup.
constexpr uint32_t kHistLds = 4096;

__device__ __forceinline__ void hist_add(int* s_h, long long* hist, uint32_t minf, uint32_t maxf, uint64_t f, int d)
{
    if (f < minf || f > maxf) return;
    const uint64_t b = f - minf;
    if (b < kHistLds) atomicAdd(&s_h[b], d);
    else atomicAdd((unsigned long long*)&hist[b], (unsigned long long)(long long)d);
}

__device__ __forceinline__ void hist_flush(const int* s_h, long long* hist, uint32_t minf, uint32_t maxf)
{
    const uint64_t nb = (uint64_t)maxf - minf + 1;
    for (uint32_t b = threadIdx.x; b < kHistLds && b < nb; b += blockDim.x)
        if (s_h[b]) atomicAdd((unsigned long long*)&hist[b], (unsigned long long)(long long)s_h[b]);
}

// Records: each one is counted at its count r (precise mode skips a record
// with r + 1 < minf, tool/HighFrequencyInsertThreadPool.h:150-177) and, in
// precise mode, its key is entered into the table with r.
template <int NW>
__global__ void __launch_bounds__(256)
k_hist_records(const uint8_t* __restrict__ recs, uint64_t nrec, HistArgs a)
{
    __shared__ int s_h[kHistLds];
    for (uint32_t b = threadIdx.x; b < kHistLds; b += 256) s_h[b] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nrec; i += (uint64_t)gridDim.x * 256) {
        const uint64_t* r = reinterpret_cast<const uint64_t*>(recs + i * RecBytes<NW>::value);
        const uint32_t c = (uint32_t)r[NW];
        if (a.precise && (uint64_t)c + 1 < a.minf) continue;
        hist_add(s_h, a.hist, a.minf, a.maxf, c, 1);
        if (c > a.maxf) {  // beyond the histogram: kept in the tail totals
            atomicAdd(&a.tail[0], 1ULL);
            atomicAdd(&a.tail[1], (unsigned long long)c);
        }
        if (!a.precise) continue;
        Key<NW> key;
#pragma unroll
        for (int w = 0; w < NW; ++w) key.w[w] = r[w];
        bool claimed;
        const int64_t s = table_find_or_claim<TableWords<NW>::value>(a.keys, a.cap, a.cap - 1, key_mix<NW>(key),
                                                                    kHistProbe, to_table<NW>(key), claimed);
        if (s >= 0) a.vals[s] = c;
        else atomicOr(a.error_flags, kErrHistFull);
    }
    __syncthreads();
    hist_flush(s_h, a.hist, a.minf, a.maxf);
}

// First-occurrence keys (precise mode, tool/LowFrequencyQueryThreadPool.h:184-232):
// a key found with r moves one unit from bin r to bin r + 1, a key not found
// adds one to bin 1.
template <int NW>
__global__ void __launch_bounds__(256)
k_hist_firsts(const uint64_t* __restrict__ firsts, uint64_t nfirst, HistArgs a)
{
    __shared__ int s_h[kHistLds];
    for (uint32_t b = threadIdx.x; b < kHistLds; b += 256) s_h[b] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nfirst; i += (uint64_t)gridDim.x * 256) {
        constexpr int TW = TableWords<NW>::value;
        const Key<NW> key = load_key<NW>(firsts, i);
        const Key<TW> tkey = to_table<NW>(key);
        // probe until the key or an empty slot (the table is at most half full)
        uint64_t s = key_mix<NW>(key) & (a.cap - 1);
        bool found = false;
        for (uint32_t p = 0; p < kHistProbe; ++p) {
            const unsigned long long w0 = a.keys[s];
            if (w0 == kEmpty) break;
            bool match = w0 == tkey.w[0];
#pragma unroll
            for (int w = 1; w < TW; ++w) match = match && a.keys[(uint64_t)w * a.cap + s] == tkey.w[w];
            if (match) { found = true; break; }
            s = (s + 1) & (a.cap - 1);
        }
        if (found) {
            const uint32_t c = a.vals[s];
            hist_add(s_h, a.hist, a.minf, a.maxf, c, -1);
            hist_add(s_h, a.hist, a.minf, a.maxf, (uint64_t)c + 1, 1);
            // a key of merged count c + 1 above maxf is in the tail (moved there
            // from bin maxf, or already there at c)
            if (c > a.maxf) {
                atomicAdd(&a.tail[1], 1ULL);
            } else if (c == a.maxf) {
                atomicAdd(&a.tail[0], 1ULL);
                atomicAdd(&a.tail[1], (unsigned long long)c + 1);
            }
        } else {
            hist_add(s_h, a.hist, a.minf, a.maxf, 1, 1);
        }
    }
    __syncthreads();
    hist_flush(s_h, a.hist, a.minf, a.maxf);
}

template <int NW>
hipError_t launch_histogram(const uint8_t* recs, uint64_t nrec, const uint64_t* firsts, uint64_t nfirst, const HistArgs& a,
                            hipStream_t st)
{
    hipLaunchKernelGGL((k_hist_records<NW>), dim3(2048), dim3(256), 0, st, recs, nrec, a);
    KMC_LAUNCH_CHECK();
    if (a.precise) {
        hipLaunchKernelGGL((k_hist_firsts<NW>), dim3(2048), dim3(256), 0, st, firsts, nfirst, a);
        KMC_LAUNCH_CHECK();
    }
    return hipSuccess;
}

template hipError_t launch_histogram<1>(const uint8_t*, uint64_t, const uint64_t*, uint64_t, const HistArgs&, hipStream_t);
template hipError_t launch_histogram<2>(const uint8_t*, uint64_t, const uint64_t*, uint64_t, const HistArgs&, hipStream_t);
template hipError_t launch_histogram<4>(const uint8_t*, uint64_t, const uint64_t*, uint64_t, const HistArgs&, hipStream_t);

// Fills used instead of memset nodes in the captured pipeline graph.
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t* __restrict__ p, uint64_t n, uint8_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill(void* p, uint64_t bytes, uint8_t v, hipStream_t st)
{
    if (!bytes) return hipSuccess;
    const uint64_t aligned = (reinterpret_cast<uintptr_t>(p) % 4 == 0) ? bytes / 4 : 0;
    if (aligned) {
        const uint32_t v4 = 0x01010101u * v;
        const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, (aligned + 255) / 256);
        hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, st, reinterpret_cast<uint32_t*>(p), aligned, v4);
        KMC_LAUNCH_CHECK();
    }
    const uint64_t rest = bytes - aligned * 4;
    if (rest) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, (rest + 255) / 256);
        hipLaunchKernelGGL(k_fill_u8, dim3(grid), dim3(256), 0, st, reinterpret_cast<uint8_t*>(p) + aligned * 4, rest, v);
        KMC_LAUNCH_CHECK();
    }
    return hipSuccess;
}

// Plain copy of n16 16-byte words: the traffic reference that HBM counters
// of the pipeline kernels are scaled by (reads 16*n16 bytes, writes 16*n16).
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

hipError_t launch_debug_copy(const void* src, void* dst, uint64_t n16, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(256), 0, st, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(dst), n16);
    KMC_LAUNCH_CHECK();
    return hipSuccess;
}


}  // namespace kmc
