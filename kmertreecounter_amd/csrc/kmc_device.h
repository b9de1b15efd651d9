// kmc_device.h -- device-side building blocks shared by the engine kernels:
// base encoding, canonical k-mer construction, table key encoding and
// routing tables.  gfx950 (CDNA4), wave64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmc {

constexpr int kWave = 64;
// empty table word: tables store keys encoded so that no stored word is all ones (TableKey)
constexpr uint64_t kEmpty = ~0ULL;

template <int NW>
struct Key {
    uint64_t w[NW];
};

// ---------------------------------------------------------------- encoding
// ASCII -> 2-bit code: A/a=0, C/c=1, G/g=2, T/t/U/u=3 (reference table,
// src/GetKmer.h:22-36).  ((c >> 1) & 3) maps A,C,G,T to 0,1,3,2; xor with
// its high bit swaps G and T.  Validity is the case-folded byte being one of
// A C G T U (src/FastqParser.h:194-236); everything else resets a window.
//
// Four bytes per 32-bit operation: the codes of all four bytes at once, and
// validity as "the case-folded byte is the letter its code stands for" (a
// byte lookup of A C G T by code with v_perm_b32; for code 3, U is folded
// onto T by clearing bit 0), then an exact zero-byte test.  Outputs: the
// four codes in the low byte (first byte in bits 7..6) and a nibble of
// invalid flags (first byte in bit 3).
__device__ __forceinline__ void encode_word(uint32_t w, uint32_t& codes, uint32_t& bad)
{
    const uint32_t t = (w >> 1) & 0x03030303u;
    const uint32_t cd = t ^ ((t >> 1) & 0x01010101u);  // code of each byte, 0..3
    codes = (cd << 6) | (cd >> 4) | (cd >> 14) | (cd >> 24);
    constexpr uint32_t kLetters = 0x54474341u;  // 'T' 'G' 'C' 'A' in bytes 3..0
    const uint32_t expect = __builtin_amdgcn_perm(kLetters, kLetters, cd);
    const uint32_t u_as_t = cd & (cd >> 1) & 0x01010101u;  // bit 0 of code-3 bytes
    const uint32_t d = (w & 0xDFDFDFDFu & ~u_as_t) ^ expect;  // zero byte: valid
    const uint32_t nz = (((d & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | d) & 0x80808080u;
    const uint32_t g = nz >> 7;  // bytes 0..3 -> bits 0, 8, 16, 24
    bad = ((g << 3) | (g >> 6) | (g >> 15) | (g >> 24)) & 0xFu;
}

// 16 sequence bytes -> 16 codes (MSB-first, first byte in bits 31..30) and a
// 16-bit validity mask (first byte in bit 15).
__device__ __forceinline__ void encode16(uint4 q, uint32_t& code, uint32_t& valid)
{
    uint32_t c0, c1, c2, c3, b0, b1, b2, b3;
    encode_word(q.x, c0, b0);
    encode_word(q.y, c1, b1);
    encode_word(q.z, c2, b2);
    encode_word(q.w, c3, b3);
    // low bytes of c0..c3 into bytes 3..0 (v_perm_b32: selectors 4..7 pick the
    // first operand's bytes, 0..3 the second's, 0x0C a zero byte)
    code = __builtin_amdgcn_perm(c0, c1, 0x04000C0Cu) | __builtin_amdgcn_perm(c2, c3, 0x0C0C0400u);
    valid = ~((b0 << 12) | (b1 << 8) | (b2 << 4) | b3) & 0xFFFFu;
}

// reverse the order of the 16 two-bit groups of x
__device__ __forceinline__ uint32_t rev16(uint32_t x)
{
    x = __builtin_bitreverse32(x);
    return ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
}

// bits [s, s + 32) of the 64-bit value hi:lo (s in 0..31; v_alignbit_b32)
__device__ __forceinline__ uint32_t funnel32(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
}

// reverse the order of the 32 two-bit groups of x
__device__ __forceinline__ uint64_t rev2(uint64_t x)
{
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
}

// 33 <= k <= 64: F = (fhi, flo) is a 2k-bit value
__device__ __forceinline__ Key<2> canon2(uint64_t fhi, uint64_t flo, uint32_t k)
{
    const uint32_t hb = 2 * k - 64;  // bits of fhi in use, 2..64
    const uint64_t hmask = hb == 64 ? ~0ULL : (1ULL << hb) - 1ULL;
    fhi &= hmask;
    const uint64_t chi = ~fhi & hmask, clo = ~flo;
    const uint32_t sh = 128 - 2 * k;  // 0..62
    const uint64_t rhi = sh ? (chi << sh) | (clo >> (64 - sh)) : chi;
    const uint64_t rlo = clo << sh;
    const uint64_t shi = rev2(flo), slo = rev2(fhi);
    const bool seq_less = (shi < rhi) || (shi == rhi && slo < rlo);
    Key<2> r;
    r.w[0] = seq_less ? shi : rhi;
    r.w[1] = seq_less ? slo : rlo;
    return r;
}

// 65 <= k <= 128: S = (s0, s1, s2, s3) holds the codes of the last 128 bases
// ending at the window's last base (oldest on top, 2 bits each).  As in
// canon2 (and the reference, src/GetKmer.h:71-77), seq = pack(reverse(window))
// and rev = pack(complement(window)), left aligned in four words.  For k = 96 the reference never masks word 2 of
// seq (kmer_last_index 2, remainder 0, src/GetKmer.h:38-47,:71-73), so the
// bases shifted out of the window stay in word 3 of seq: the 32 bases before
// the window, newest first, zero where they precede the current run.  `pre`
// holds those 32 codes (oldest on top), already zeroed before the run start.
__device__ __forceinline__ Key<4> canon4(uint64_t s0, uint64_t s1, uint64_t s2, uint64_t s3, uint32_t k, uint64_t pre)
{
    // forward window F: the low 2k bits of S
    const uint32_t bits = 2 * k;                     // 130..256
    uint64_t f0 = bits >= 256 ? s0 : (bits > 192 ? s0 & ((1ULL << (bits - 192)) - 1ULL) : 0ULL);
    uint64_t f1 = bits >= 192 ? s1 : s1 & ((1ULL << (bits - 128)) - 1ULL);
    const uint64_t f2 = s2, f3 = s3;
    // rev = ~F (2k bits) shifted to the top of 256 bits
    const uint64_t m0 = bits >= 256 ? ~0ULL : (bits > 192 ? (1ULL << (bits - 192)) - 1ULL : 0ULL);
    const uint64_t m1 = bits >= 192 ? ~0ULL : (1ULL << (bits - 128)) - 1ULL;
    uint64_t c0 = ~f0 & m0, c1 = ~f1 & m1, c2 = ~f2, c3 = ~f3;
    const uint32_t sh = 256 - bits;                  // 0..126
    uint64_t r0, r1, r2, r3;
    if (sh >= 64) {  // 130 <= 2k <= 192: F fits words 1..3
        const uint32_t t = sh - 64;
        r0 = t ? (c1 << t) | (c2 >> (64 - t)) : c1;
        r1 = t ? (c2 << t) | (c3 >> (64 - t)) : c2;
        r2 = c3 << t;
        r3 = 0;
    } else {
        r0 = sh ? (c0 << sh) | (c1 >> (64 - sh)) : c0;
        r1 = sh ? (c1 << sh) | (c2 >> (64 - sh)) : c1;
        r2 = sh ? (c2 << sh) | (c3 >> (64 - sh)) : c2;
        r3 = c3 << sh;
    }
    // seq = the 128 two-bit groups of F in reverse order (newest base on top)
    const uint64_t q0 = rev2(f3), q1 = rev2(f2), q2 = rev2(f1);
    const uint64_t q3 = k == 96 ? rev2(pre) : rev2(f0);
    const bool seq_less = q0 != r0 ? q0 < r0 : (q1 != r1 ? q1 < r1 : (q2 != r2 ? q2 < r2 : q3 < r3));
    Key<4> r;
    r.w[0] = seq_less ? q0 : r0;
    r.w[1] = seq_less ? q1 : r1;
    r.w[2] = seq_less ? q2 : r2;
    r.w[3] = seq_less ? q3 : r3;
    return r;
}

template <int NW>
__device__ __forceinline__ bool key_eq(const Key<NW>& a, const Key<NW>& b)
{
    bool e = true;
#pragma unroll
    for (int i = 0; i < NW; ++i) e = e && (a.w[i] == b.w[i]);
    return e;
}

// ---------------------------------------------------------------- table keys
// Open-addressing tables mark a free word with kEmpty (all ones), so a stored
// word must never be all ones.  k <= 32: a canonical key word never is.
// k <= 64: only k = 64 reaches it, and only with the key (~0, 0) (first 32
// bases A, last 32 T); it is stored as (~0 - 1, 1), which no canonical key of
// k <= 64 can be.  k <= 128: any word can be all ones, so the 256 key bits
// are stored as five chunks of at most 63 bits.
template <int NW> struct TableWords { static constexpr int value = NW; };
template <> struct TableWords<4> { static constexpr int value = 5; };

template <int NW>
__device__ __forceinline__ Key<TableWords<NW>::value> to_table(const Key<NW>& k)
{
    if constexpr (NW == 1) {
        return k;
    } else if constexpr (NW == 2) {
        Key<2> t = k;
        if (k.w[0] == ~0ULL && k.w[1] == 0ULL) { t.w[0] = ~0ULL - 1; t.w[1] = 1ULL; }
        return t;
    } else {
        Key<5> t;
        t.w[0] = k.w[0] >> 1;
        t.w[1] = ((k.w[0] & 1ULL) << 62) | (k.w[1] >> 2);
        t.w[2] = ((k.w[1] & 3ULL) << 61) | (k.w[2] >> 3);
        t.w[3] = ((k.w[2] & 7ULL) << 60) | (k.w[3] >> 4);
        t.w[4] = k.w[3] & 15ULL;
        return t;
    }
}

template <int NW>
__device__ __forceinline__ Key<NW> from_table(const Key<TableWords<NW>::value>& t)
{
    if constexpr (NW == 1) {
        return t;
    } else if constexpr (NW == 2) {
        Key<2> k = t;
        if (t.w[0] == ~0ULL - 1 && t.w[1] == 1ULL) { k.w[0] = ~0ULL; k.w[1] = 0ULL; }
        return k;
    } else {
        Key<4> k;
        k.w[0] = (t.w[0] << 1) | (t.w[1] >> 62);
        k.w[1] = (t.w[1] << 2) | (t.w[2] >> 61);
        k.w[2] = (t.w[2] << 3) | (t.w[3] >> 60);
        k.w[3] = (t.w[3] << 4) | t.w[4];
        return k;
    }
}

// ---------------------------------------------------------------- routing
// Root bucket (0..255) and fine bucket.  Roots are the top 8 key bits (the
// reference's root prefix, ROOT_BASES=4, src/definition.h:28) for both
// partitions.  Fine buckets:
//   prefix: the next sbits bits (fine index r << sbits | bits).
//   learned: root r has nsub_r = fine_base[r+1] - fine_base[r] fine buckets
//           (1..kMaxFine), bounded by nsub_r - 1 ascending keys taken from the
//           learned CDF model (fine index fine_base[r] + number of r's bounds
//           <= key).
// learned routing tables: per root a table of kLutCells u16 cells over the
// root's key range, cell c = the root's bounds <= (r << 56) + (c << 44), then
// the 257 fine bases (u32)
constexpr uint32_t kLutCells = 4096;
constexpr uint32_t kLutShift = 44;  // a root spans 2^56 keys: 4096 cells of 2^44
constexpr uint32_t kMaxFine = 1024;  // fine buckets of one root, at most
constexpr size_t kRouteLutBytes = 256 * (size_t)kLutCells * 2 + 257 * 4;
constexpr size_t kRouteFineBaseOffset = 256 * (size_t)kLutCells * 2;

struct Route {
    uint32_t mode;   // 0 prefix, 1 learned
    uint32_t sbits;  // prefix: fine bits 0..kMaxSubBits (learned: unused)
    uint32_t nf;     // fine buckets in all (prefix: 256 << sbits)
    const uint32_t* fine_base;    // learned: [257]
    const uint64_t* fine_bounds;  // learned: root r's bounds at fine_bounds[fine_base[r] - r]
    const uint16_t* lut;          // learned: 256 x kLutCells cells
    // bucket-range round: the scatter keeps the keys of root buckets [rlo, rhi)
    uint32_t rlo, rhi;
};

}  // namespace kmc
