/*
 * kmc_oracle.c -- CPU restatement of the KmerTreeCounter (tree_v4) counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it.  The product (kmertreecounter_amd, libkmc.so) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Reference paths are relative to the reference repository root.
 *
 * Key layout (reference src/kmer.h:12-21): N = 1, 2 or 4 u64 words per key,
 * word 0 most significant, bases packed MSB-first 2 bits each, left aligned,
 * low unused bits zero.  Keys compare as N-word big-endian numbers
 * (defaulted operator<=> over std::array, src/kmer.h:21).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_fastq_blocks   - every-offset split invariance + reference Tree K,
 *                        and the product reader byte-for-byte.
 *   orc_extract        - brute force per-window closed form, strand symmetry,
 *                        reference GetKmer via oracle/_ref harness, reference
 *                        Tree merged tables (tests/golden).
 *   orc_count_sorted   - Python sorted()/Counter on small inputs, sum == K.
 *   orc_xxh3_128       - reference xxhash.h v0.8.3 via oracle/_ref harness
 *                        (vectors committed in tests/golden/harness_vectors.json).
 *   orc_bloom_*        - no-false-negative and order-independence invariants,
 *                        reference ConcurrentBloomFilter via oracle/_ref harness.
 *   orc_pack_low/unpack- reference ExportReader KAT inputs (k=33, k=64),
 *                        reference Tree low.bin round trip.
 *   orc_histogram_*    - reference histogram_tool KATs ({1,1,1,0}, {1,0,0},
 *                        {0,1,0}, {1,-1,1,-1}) + closed form from the table.
 *   orc_pre_read, orc_prefix_counts, orc_clamp_prefix_counts, orc_lpt,
 *   orc_bloom_capacity - the Stage-0 lines the reference Tree prints (average
 *                        clamped prefix count, every classifier's LPT load,
 *                        largest Bloom capacity) on plain, gzip and small
 *                        inputs (tests/golden/stage0_runs.json); prefix
 *                        counts also against numpy bincount and sum == K;
 *                        hand-computed LPT and capacity cases.  The
 *                        reference's unstable std::sort leaves the owner of
 *                        equal-count prefixes open, never the set of loads.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_API __attribute__((visibility("default")))

static inline uint32_t words_for_k(uint32_t k)
{
    /* src/main.cpp:579-595: k<=32 -> 1 word, k<=64 -> 2, k<=128 -> 4 */
    return k <= 32 ? 1u : (k <= 64 ? 2u : 4u);
}

ORC_API uint32_t orc_words_for_k(uint32_t k) { return words_for_k(k); }

/* ------------------------------------------------------------------------- */
/* A1: FastqReader block chunking (src/FastqReader.h:140-388).               */
/* ------------------------------------------------------------------------- */

typedef struct {
    char* out;            /* concatenated block bytes */
    uint64_t out_cap;
    uint64_t out_len;
    uint64_t* lens;       /* block lengths */
    uint64_t lens_cap;
    uint64_t n_blocks;
    uint64_t block_size;
    /* current block */
    uint64_t cur_start;   /* offset of current block in out */
    uint64_t write_size;
    uint64_t last_nl;     /* UINT64_MAX == none */
    int has_block;
    char carry[128];
    uint64_t carry_len;
    int overflow;
} orc_blocks_t;

#define ORC_NO_NL UINT64_MAX

static void ob_acquire(orc_blocks_t* b)
{
    /* acquire_block, src/FastqReader.h:310-336 */
    b->cur_start = b->out_len;
    b->write_size = 0;
    b->last_nl = ORC_NO_NL;
    b->has_block = 1;
    if (b->carry_len > 0) {
        if (b->cur_start + b->carry_len > b->out_cap) { b->overflow = 1; return; }
        memcpy(b->out + b->cur_start, b->carry, b->carry_len);
        b->write_size = b->carry_len;
        b->carry_len = 0;
    }
}

static void ob_publish(orc_blocks_t* b)
{
    /* publish_current_block, src/FastqReader.h:367-388: empty blocks are returned */
    if (!b->has_block) return;
    if (b->write_size > 0) {
        if (b->n_blocks >= b->lens_cap) { b->overflow = 1; }
        else {
            b->lens[b->n_blocks++] = b->write_size;
            b->out_len = b->cur_start + b->write_size;
        }
    }
    b->has_block = 0;
    b->write_size = 0;
    b->last_nl = ORC_NO_NL;
}

static void ob_store_overlap(orc_blocks_t* b, uint64_t overlap)
{
    /* store_overlap_from_block_end, src/FastqReader.h:338-365 */
    b->carry_len = 0;
    if (overlap == 0 || b->write_size == 0) return;
    uint64_t keep_window = b->write_size < overlap ? b->write_size : overlap;
    uint64_t start = b->write_size - keep_window;
    const char* blk = b->out + b->cur_start;
    if (b->last_nl != ORC_NO_NL && b->last_nl >= start) {
        uint64_t keep = b->write_size - (b->last_nl + 1);
        if (keep > 0) { memcpy(b->carry, blk + b->last_nl + 1, keep); b->carry_len = keep; }
        return;
    }
    memcpy(b->carry, blk + start, keep_window);
    b->carry_len = keep_window;
}

/*
 * The reader's line state machine over one buffer of FASTQ bytes
 * (src/FastqReader.h:140-308).  The 4-line state machine (:293-308) keeps only
 * line 2 of each record; every sequence line ends with '\n' inside the block;
 * a line split at a full block carries its last k-1 bytes (or the bytes after
 * the last '\n', if nearer) into the next block.  *state (0 header,
 * 1 sequence, 2 plus, 3 quality) persists across buffers, as it does across
 * the reference's 2 MiB reads.  Returns 0, or -1 when out/lens are too small.
 */
static int ob_feed(orc_blocks_t* b, const char* fq, uint64_t fq_len, int* state, uint64_t overlap)
{
    uint64_t pos = 0;
    while (pos < fq_len) {
        if (!b->has_block) ob_acquire(b);
        if (b->overflow) return -1;
        if (*state != 1) {
            const char* nl = memchr(fq + pos, '\n', fq_len - pos);
            if (!nl) { pos = fq_len; continue; }
            pos = (uint64_t)(nl - fq) + 1;
            *state = (*state + 1) & 3;
            continue;
        }
        const char* nl = memchr(fq + pos, '\n', fq_len - pos);
        uint64_t seq_len = nl ? (uint64_t)(nl - (fq + pos)) : fq_len - pos;
        uint64_t copied = 0;
        while (copied < seq_len) {
            if (b->write_size == b->block_size) {
                ob_store_overlap(b, overlap);
                ob_publish(b);
                ob_acquire(b);
                if (b->overflow) return -1;
            }
            uint64_t room = b->block_size - b->write_size;
            uint64_t n = seq_len - copied < room ? seq_len - copied : room;
            if (b->cur_start + b->write_size + n > b->out_cap) return -1;
            memcpy(b->out + b->cur_start + b->write_size, fq + pos + copied, n);
            b->write_size += n;
            copied += n;
        }
        pos += seq_len;
        if (!nl) continue;
        if (b->write_size == b->block_size) {
            b->carry_len = 0;
            ob_publish(b);
            ob_acquire(b);
            if (b->overflow) return -1;
        }
        if (b->cur_start + b->write_size + 1 > b->out_cap) return -1;
        b->last_nl = b->write_size;
        b->out[b->cur_start + b->write_size++] = '\n';
        pos++;
        *state = (*state + 1) & 3;
    }
    return 0;
}

static int ob_finish(orc_blocks_t* b, uint64_t* n_blocks, uint64_t* out_len)
{
    if (!b->has_block) ob_acquire(b);
    ob_publish(b);
    if (b->overflow) return -1;
    *n_blocks = b->n_blocks;
    *out_len = b->out_len;
    return 0;
}

static void ob_init(orc_blocks_t* b, uint64_t block_size, char* out, uint64_t out_cap, uint64_t* lens, uint64_t lens_cap)
{
    memset(b, 0, sizeof(*b));
    b->out = out; b->out_cap = out_cap; b->lens = lens; b->lens_cap = lens_cap;
    b->block_size = block_size; b->last_nl = ORC_NO_NL;
}

/*
 * Convert FASTQ bytes into the reader's sequence-only blocks
 * (FastqReader<N>::read, src/FastqReader.h:140-290).  The reference reads the
 * file in 2 MiB chunks; the output does not depend on the chunk size, so the
 * whole input is one buffer here.
 */
ORC_API int orc_fastq_blocks(const char* fq, uint64_t fq_len, uint32_t k, uint64_t block_size,
                             char* out, uint64_t out_cap, uint64_t* lens, uint64_t lens_cap,
                             uint64_t* n_blocks, uint64_t* out_len)
{
    orc_blocks_t b;
    ob_init(&b, block_size, out, out_cap, lens, lens_cap);
    int state = 0;
    if (ob_feed(&b, fq, fq_len, &state, k > 1 ? (uint64_t)(k - 1) : 0)) return -1;
    return ob_finish(&b, n_blocks, out_len);
}

/* find_next_record_start (src/FastqPreReader.h:330-347): the first "\n@" in
 * the 64 KiB from `start` on; the record starts at its '@' (a '\n' at
 * `start` - 1 is not looked at).  No match: `start` itself. */
static uint64_t pre_record_start(const char* fq, uint64_t fq_len, uint64_t start)
{
    if (start >= fq_len) return start;
    uint64_t n = fq_len - start < 65536 ? fq_len - start : 65536;
    if (n <= 1) return start;
    for (uint64_t i = 1; i < n; ++i)
        if (fq[start + i - 1] == '\n' && fq[start + i] == '@') return start + i;
    return start;
}

/*
 * Stage-0 sample blocks: FastqPreReader<N>::pre_read (src/FastqPreReader.h:
 * 110-165, :128-300).  `fq` is the whole (decompressed) input.  need =
 * min(file_size, 256 MiB), where file_size is the size on disk (the compressed
 * size for gzip, as fstat gives it, :79-86).
 *   plain: three parts of need/3 bytes (at least 1) at 0, file_size/2 -
 *     part/2 and file_size - part (clamped at 0), each but the first moved to
 *     the next "\n@".  Reads are max(2 MiB, part) bytes (:150-154); a read that
 *     completes a part's byte count resets the line state to the header and
 *     seeks to the next part BEFORE its own bytes are parsed (:193-207), so
 *     the state is not reset between the second and third parts, and a
 *     sequence line cut at a part's end runs on into the next part's first
 *     sequence line.  Reading stops once need bytes were read (:209-212).
 *   gzip: 2 MiB reads from the start while fewer than need bytes were read;
 *     the read that finds need reached is dropped (:180-190).
 * Blocks are cut as in orc_fastq_blocks.  Returns 0 or -1 (capacity).
 */
ORC_API int orc_pre_read(const char* fq, uint64_t fq_len, uint64_t file_size, int is_gz, uint32_t k,
                         uint64_t block_size, char* out, uint64_t out_cap, uint64_t* lens, uint64_t lens_cap,
                         uint64_t* n_blocks, uint64_t* out_len)
{
    const uint64_t cap256 = 256ULL << 20;
    const uint64_t need = file_size < cap256 ? file_size : cap256;
    uint64_t part = need / 3;
    if (part == 0) part = 1;
    uint64_t chunk = 2ULL << 20;
    if (part > chunk && !is_gz) chunk = part;
    const uint64_t overlap = k > 1 ? (uint64_t)(k - 1) : 0;
    orc_blocks_t b;
    ob_init(&b, block_size, out, out_cap, lens, lens_cap);
    int state = 0;
    uint64_t pos = 0, have = 0;
    if (is_gz) {
        while (have < need && pos < fq_len) {
            const uint64_t n = fq_len - pos < chunk ? fq_len - pos : chunk;
            have += n;
            if (ob_feed(&b, fq + pos, n, &state, overlap)) return -1;
            pos += n;
        }
        return ob_finish(&b, n_blocks, out_len);
    }
    uint64_t offs[3];
    offs[0] = 0;
    offs[1] = file_size / 2 >= part / 2 ? file_size / 2 - part / 2 : 0;
    offs[2] = file_size >= part ? file_size - part : 0;
    int seg = 0;
    uint64_t seg_read = 0;
    for (;;) {
        if (pos >= fq_len) break;
        const uint64_t n = fq_len - pos < chunk ? fq_len - pos : chunk;
        const uint64_t at = pos;
        pos += n;
        seg_read += n;
        have += n;
        if (seg_read >= part && seg < 2) {
            state = 0;
            ++seg;
            seg_read = 0;
            pos = pre_record_start(fq, fq_len, offs[seg]);
        }
        if (ob_feed(&b, fq + at, n, &state, overlap)) return -1;
        if (have >= need) break;
    }
    return ob_finish(&b, n_blocks, out_len);
}

/* ------------------------------------------------------------------------- */
/* A3/A4: canonical k-mer extraction (src/GetKmer.h:19-129,                   */
/* src/FastqParser.h:182-324).                                                */
/* ------------------------------------------------------------------------- */

static inline int base_code(unsigned char c)
{
    /* char_to_2bit table, src/GetKmer.h:22-36 */
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return -1;
    }
}

typedef struct {
    uint32_t k, nw, last;
    uint64_t back_mask, rev_shift;
    uint64_t seq[4], rev[4];
    uint64_t have;
} orc_roll_t;

static void roll_init(orc_roll_t* r, uint32_t k)
{
    memset(r, 0, sizeof(*r));
    r->k = k;
    r->nw = words_for_k(k);
    r->last = (k - 1) / 32;                /* kmer_last_index, src/GetKmer.h:19 */
    uint32_t rem = k % 32;
    if (rem) {                             /* src/GetKmer.h:38-47 */
        r->back_mask = ((1ULL << (2 * rem)) - 1ULL) << (64 - 2 * rem);
        r->rev_shift = 2 * (32 - rem);
    } else {
        r->back_mask = ~0ULL;
        r->rev_shift = 0;
    }
}

static inline void roll_clear(orc_roll_t* r)
{
    memset(r->seq, 0, sizeof(r->seq));
    memset(r->rev, 0, sizeof(r->rev));
    r->have = 0;
}

/* One get_next_one step (src/GetKmer.h:58-82) with a valid base.  Returns 1
 * when a full window is available; the canonical key is min(seq, rev). */
static inline int roll_push(orc_roll_t* r, uint64_t code)
{
    const uint32_t n = r->nw;
    /* seq: shift right one base, new base at the top of word 0 */
    for (int i = (int)n - 1; i > 0; --i) r->seq[i] = (r->seq[i] >> 2) | (r->seq[i - 1] << 62);
    r->seq[0] >>= 2;
    r->seq[0] |= code << 62;
    r->seq[r->last] &= r->back_mask;
    /* rev: shift left one base, complement at the low valid slot */
    for (uint32_t i = 0; i + 1 < n; ++i) r->rev[i] = (r->rev[i] << 2) | (r->rev[i + 1] >> 62);
    r->rev[n - 1] <<= 2;
    r->rev[r->last] |= (code ^ 3ULL) << r->rev_shift;
    r->have++;
    return r->have >= r->k;
}

static inline int words_less(const uint64_t* a, const uint64_t* b, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i) {
        if (a[i] != b[i]) return a[i] < b[i];
    }
    return 0;
}

static inline void roll_key(const orc_roll_t* r, uint64_t* out)
{
    /* canonicalize, src/GetKmer.h:181-206: seq if seq < rev else rev */
    const uint64_t* src = words_less(r->seq, r->rev, r->nw) ? r->seq : r->rev;
    for (uint32_t i = 0; i < r->nw; ++i) out[i] = src[i];
}

/*
 * Extract canonical keys from a byte stream, resetting the state on every
 * byte that is not A/C/G/T/U (either case) -- this includes '\n', 'N', '\r'
 * (FastqParser.h:211-324 + GetKmer::clear).  Writes nw words per key.
 * Returns the number of keys, or UINT64_MAX if cap (in keys) is exceeded.
 */
ORC_API uint64_t orc_extract(const char* data, uint64_t len, uint32_t k, uint64_t* out, uint64_t cap)
{
    if (k == 0 || k > 128) return UINT64_MAX;
    orc_roll_t r;
    roll_init(&r, k);
    uint64_t n = 0;
    for (uint64_t i = 0; i < len; ++i) {
        int c = base_code((unsigned char)data[i]);
        if (c < 0) { roll_clear(&r); continue; }
        if (roll_push(&r, (uint64_t)c)) {
            if (n >= cap) return UINT64_MAX;
            roll_key(&r, out + n * r.nw);
            n++;
        }
    }
    return n;
}

/* Count keys only (same rule as orc_extract). */
ORC_API uint64_t orc_count_kmers(const char* data, uint64_t len, uint32_t k)
{
    uint64_t run = 0, n = 0;
    for (uint64_t i = 0; i < len; ++i) {
        if (base_code((unsigned char)data[i]) < 0) { run = 0; continue; }
        if (++run >= k) n++;
    }
    return n;
}

/* Extraction over reader blocks: each block starts from a cleared state
 * (FastqParser::parse calls get_kmer.clear() per block, src/FastqParser.h:193). */
ORC_API uint64_t orc_extract_blocks(const char* data, const uint64_t* lens, uint64_t n_blocks,
                                    uint32_t k, uint64_t* out, uint64_t cap)
{
    uint64_t off = 0, n = 0;
    const uint32_t nw = words_for_k(k);
    for (uint64_t b = 0; b < n_blocks; ++b) {
        uint64_t got = orc_extract(data + off, lens[b], k, out + n * nw, cap - n);
        if (got == UINT64_MAX) return UINT64_MAX;
        n += got;
        off += lens[b];
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* Exact counting: sort + run-length (the result the reference's tree + final */
/* drain produce, src/NewKmerTree.h:1170-1266, merged with low.bin).          */
/* ------------------------------------------------------------------------- */

/* LSD radix sort over nw-word keys, 16-bit digits, least significant word last. */
static int radix_sort_words(uint64_t* keys, uint64_t n, uint32_t nw)
{
    if (n < 2) return 0;
    uint64_t* tmp = (uint64_t*)malloc(n * nw * sizeof(uint64_t));
    if (!tmp) return -1;
    uint64_t* src = keys;
    uint64_t* dst = tmp;
    uint64_t* cnt = (uint64_t*)malloc(65536 * sizeof(uint64_t));
    if (!cnt) { free(tmp); return -1; }
    for (int w = (int)nw - 1; w >= 0; --w) {
        for (int shift = 0; shift < 64; shift += 16) {
            memset(cnt, 0, 65536 * sizeof(uint64_t));
            for (uint64_t i = 0; i < n; ++i) cnt[(src[i * nw + w] >> shift) & 0xFFFF]++;
            /* skip passes where every key has the same digit */
            int trivial = 0;
            for (uint32_t d = 0; d < 65536; ++d) if (cnt[d] == n) { trivial = 1; break; }
            if (trivial) continue;
            uint64_t sum = 0;
            for (uint32_t d = 0; d < 65536; ++d) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
            for (uint64_t i = 0; i < n; ++i) {
                uint64_t d = (src[i * nw + w] >> shift) & 0xFFFF;
                uint64_t p = cnt[d]++;
                memcpy(dst + p * nw, src + i * nw, nw * sizeof(uint64_t));
            }
            uint64_t* t = src; src = dst; dst = t;
        }
    }
    if (src != keys) memcpy(keys, src, n * nw * sizeof(uint64_t));
    free(cnt);
    free(tmp);
    return 0;
}

ORC_API int orc_sort_keys(uint64_t* keys, uint64_t n, uint32_t nw) { return radix_sort_words(keys, n, nw); }

/* Sorts keys in place, then writes unique keys (in place, front of the array)
 * and their counts.  Returns the number of distinct keys. */
ORC_API uint64_t orc_count_sorted(uint64_t* keys, uint64_t n, uint32_t nw, uint32_t* counts)
{
    if (n == 0) return 0;
    if (radix_sort_words(keys, n, nw) != 0) return UINT64_MAX;
    uint64_t u = 0;
    uint64_t i = 0;
    while (i < n) {
        uint64_t j = i + 1;
        while (j < n && memcmp(keys + j * nw, keys + i * nw, nw * 8) == 0) j++;
        if (u != i) memmove(keys + u * nw, keys + i * nw, nw * 8);
        counts[u] = (uint32_t)(j - i);
        u++;
        i = j;
    }
    return u;
}

/* ------------------------------------------------------------------------- */
/* A7: XXH3-128 for 8-, 16- and 32-byte inputs, seed 0, default secret.     */
/* Restated from the published XXH3 algorithm as vendored by the reference   */
/* (include/xxhash.h v0.8.3: XXH3_len_4to8_128b :6900, XXH3_len_9to16_128b   */
/* :6927, XXH3_len_17to128_128b with XXH128_mix32B, XXH3_avalanche :4616).   */
/* Only the secret bytes these paths read (offsets 0..63 of XXH3_kSecret,    */
/* :4398) are needed.                                                        */
/* The secret bytes and the XXH64/XXH3 prime constants below are those of    */
/* xxHash, Copyright (C) 2012-2023 Yann Collet, BSD 2-Clause License.         */
/* ------------------------------------------------------------------------- */

static const uint8_t kSecret0to16[16] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
};

static const uint8_t kSecret16to64[48] = {
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
};

static inline uint64_t le64(const uint8_t* p)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

#define P64_1 0x9E3779B185EBCA87ULL
#define P64_2 0xC2B2AE3D27D4EB4FULL
#define P32_2 0x85EBCA77ULL
#define PMX1 0x165667919E3779F9ULL
#define PMX2 0x9FB21C651E98DF25ULL
#define P64_4 0x85EBCA77C2B2AE63ULL

static inline void mul128(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi)
{
    unsigned __int128 p = (unsigned __int128)a * b;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
}

static inline uint64_t xs(uint64_t v, int s) { return v ^ (v >> s); }
static inline uint64_t aval(uint64_t h) { h = xs(h, 37); h *= PMX1; return xs(h, 32); }
static inline uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }

/* words are the key's u64 words in memory order (data[0], data[1]); the hash
 * input is their little-endian bytes, exactly as XXH3_128bits(&k_mer, 8N)
 * sees them (src/BloomFilter.h:209). */
ORC_API int orc_xxh3_128(const uint64_t* words, uint32_t nw, uint64_t* out_lo, uint64_t* out_hi)
{
    if (nw == 1) {
        const uint64_t len = 8;
        const uint64_t bitflip = le64(kSecret16to64 + 0) ^ le64(kSecret16to64 + 8);
        const uint64_t keyed = words[0] ^ bitflip;
        uint64_t lo, hi;
        mul128(keyed, P64_1 + (len << 2), &lo, &hi);
        hi += lo << 1;
        lo ^= hi >> 3;
        lo = xs(lo, 35);
        lo *= PMX2;
        lo = xs(lo, 28);
        hi = aval(hi);
        *out_lo = lo; *out_hi = hi;
        return 0;
    }
    if (nw == 2) {
        const uint64_t len = 16;
        const uint64_t bitflipl = le64(kSecret16to64 + 16) ^ le64(kSecret16to64 + 24);
        const uint64_t bitfliph = le64(kSecret16to64 + 32) ^ le64(kSecret16to64 + 40);
        const uint64_t in_lo = words[0];
        uint64_t in_hi = words[1];
        uint64_t mlo, mhi;
        mul128(in_lo ^ in_hi ^ bitflipl, P64_1, &mlo, &mhi);
        mlo += (len - 1) << 54;
        in_hi ^= bitfliph;
        mhi += in_hi + (uint64_t)(uint32_t)in_hi * (P32_2 - 1);
        mlo ^= bswap64(mhi);
        uint64_t hlo, hhi;
        mul128(mlo, P64_2, &hlo, &hhi);
        hhi += mhi * P64_2;
        *out_lo = aval(hlo);
        *out_hi = aval(hhi);
        return 0;
    }
    if (nw == 4) {
        /* 32 bytes: one XXH128_mix32B over (bytes 0..15, bytes 16..31) */
        const uint64_t len = 32;
        uint64_t alo = len * P64_1, ahi = 0, l, h;
        mul128(words[0] ^ le64(kSecret0to16), words[1] ^ le64(kSecret0to16 + 8), &l, &h);   /* mix16B(in1, s+0) */
        alo += l ^ h;
        alo ^= words[2] + words[3];
        mul128(words[2] ^ le64(kSecret16to64), words[3] ^ le64(kSecret16to64 + 8), &l, &h); /* mix16B(in2, s+16) */
        ahi += l ^ h;
        ahi ^= words[0] + words[1];
        *out_lo = aval(alo + ahi);
        *out_hi = (uint64_t)0 - aval(alo * P64_1 + ahi * P64_4 + len * P64_2);
        return 0;
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* A7/A8: first-occurrence Bloom filter (src/BloomFilter.h:142-259).          */
/* One 64-bit word per key, three bits (h1 + i*h2) & 63 with h2 = high|1.     */
/* ------------------------------------------------------------------------- */

ORC_API void orc_bloom_probe(const uint64_t* key, uint32_t nw, uint64_t cap_words,
                             uint64_t* word_idx, uint64_t* mask)
{
    uint64_t lo = 0, hi = 0;
    orc_xxh3_128(key, nw, &lo, &hi);
    const uint64_t h2 = hi | 1ULL;                 /* prepare_insert :207-214 */
    uint64_t m = 0;
    for (uint64_t i = 0; i < 3; ++i) m |= 1ULL << ((lo + i * h2) & 63); /* :248-258 */
    *word_idx = lo & (cap_words - 1);
    *mask = m;
}

/* Sequential inserts in array order.  is_first[i] = 1 iff the i-th key's bits
 * were not all set before its insert (insert_prepared, :225-234). */
ORC_API uint64_t orc_bloom_classify(const uint64_t* keys, uint64_t n, uint32_t nw, uint64_t cap_words,
                                    uint64_t* filter, uint8_t* is_first)
{
    uint64_t firsts = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t w, m;
        orc_bloom_probe(keys + i * nw, nw, cap_words, &w, &m);
        uint64_t old = filter[w];
        int first = (old & m) != m;
        filter[w] = old | m;
        if (is_first) is_first[i] = (uint8_t)first;
        firsts += (uint64_t)first;
    }
    return firsts;
}

/* ------------------------------------------------------------------------- */
/* A11: low.bin packing (src/ExportWriter.h:522-550) and unpacking           */
/* (tool/ExportReader.h).                                                     */
/* ------------------------------------------------------------------------- */

ORC_API uint64_t orc_packed_bytes(uint32_t k)
{
    uint64_t full = k / 32, tail_bits = 2ULL * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    return full * 8 + tail_bytes;
}

ORC_API void orc_pack_low(const uint64_t* keys, uint64_t n, uint32_t k, uint32_t nw, uint8_t* out)
{
    const uint64_t full = k / 32, tail_bits = 2ULL * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    const uint64_t mask = tail_bytes ? (~0ULL) << (64 - tail_bits) : 0;
    uint8_t* p = out;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t* key = keys + i * nw;
        for (uint64_t w = 0; w < full; ++w) {
            for (int b = 0; b < 8; ++b) *p++ = (uint8_t)(key[w] >> (8 * b));
        }
        if (tail_bytes) {
            uint64_t t = key[full] & mask;
            /* bytes 8-tail_bytes .. 7 of the little-endian word */
            for (uint64_t b = 8 - tail_bytes; b < 8; ++b) *p++ = (uint8_t)(t >> (8 * b));
        }
    }
}

ORC_API void orc_unpack_low(const uint8_t* in, uint64_t n, uint32_t k, uint32_t nw, uint64_t* keys)
{
    const uint64_t full = k / 32, tail_bits = 2ULL * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    const uint8_t* p = in;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t* key = keys + i * nw;
        for (uint32_t w = 0; w < nw; ++w) key[w] = 0;
        for (uint64_t w = 0; w < full; ++w) {
            uint64_t v = 0;
            for (int b = 0; b < 8; ++b) v |= (uint64_t)(*p++) << (8 * b);
            key[w] = v;
        }
        if (tail_bytes) {
            uint64_t v = 0;
            for (uint64_t b = 8 - tail_bytes; b < 8; ++b) v |= (uint64_t)(*p++) << (8 * b);
            key[full] = v;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* A12: root records, ExportRecord<N> (src/definition.h:159-163): nw key      */
/* words, u32 count, zero tail padding to 8-byte alignment.                   */
/* ------------------------------------------------------------------------- */

ORC_API uint64_t orc_record_bytes(uint32_t nw) { return (uint64_t)nw * 8 + 8; }

/* ------------------------------------------------------------------------- */
/* Merge engine outputs into the (key, count) table: count = [key in low] +  */
/* root record count (semantics of tool/HighFrequencyInsertThreadPool.h:     */
/* 150-177 + tool/LowFrequencyQueryThreadPool.h:184-232).                    */
/* firsts: nf keys; recs: nr records of orc_record_bytes(nw) bytes.           */
/* Output arrays need room for nf + nr entries.  Returns distinct count, or   */
/* UINT64_MAX when a key repeats inside low.bin or inside the records (which  */
/* the file contract forbids).                                                */
/* ------------------------------------------------------------------------- */

ORC_API uint64_t orc_merge_outputs(const uint64_t* firsts, uint64_t nf, const uint8_t* recs, uint64_t nr,
                                   uint32_t nw, uint64_t* out_keys, uint32_t* out_counts)
{
    const uint64_t rb = orc_record_bytes(nw);
    /* tagged copies: nw words + tag word (count, with bit 63 marking a low key) */
    const uint32_t tw = nw + 1;
    uint64_t n = nf + nr;
    uint64_t* t = (uint64_t*)malloc((n ? n : 1) * tw * sizeof(uint64_t));
    if (!t) return UINT64_MAX;
    for (uint64_t i = 0; i < nf; ++i) {
        memcpy(t + i * tw, firsts + i * nw, nw * 8);
        t[i * tw + nw] = 1ULL << 63;
    }
    for (uint64_t i = 0; i < nr; ++i) {
        const uint8_t* r = recs + i * rb;
        memcpy(t + (nf + i) * tw, r, nw * 8);
        uint32_t c;
        memcpy(&c, r + nw * 8, 4);
        t[(nf + i) * tw + nw] = c;
    }
    /* sort by key words (LSD over key words only, tag travels along) */
    if (n > 1) {
        uint64_t* tmp = (uint64_t*)malloc(n * tw * sizeof(uint64_t));
        uint64_t* cnt = (uint64_t*)malloc(65536 * sizeof(uint64_t));
        if (!tmp || !cnt) { free(t); free(tmp); free(cnt); return UINT64_MAX; }
        uint64_t* src = t; uint64_t* dst = tmp;
        for (int w = (int)nw - 1; w >= 0; --w) {
            for (int shift = 0; shift < 64; shift += 16) {
                memset(cnt, 0, 65536 * sizeof(uint64_t));
                for (uint64_t i = 0; i < n; ++i) cnt[(src[i * tw + w] >> shift) & 0xFFFF]++;
                uint64_t sum = 0;
                for (uint32_t d = 0; d < 65536; ++d) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
                for (uint64_t i = 0; i < n; ++i) {
                    uint64_t d = (src[i * tw + w] >> shift) & 0xFFFF;
                    memcpy(dst + (cnt[d]++) * tw, src + i * tw, tw * 8);
                }
                uint64_t* x = src; src = dst; dst = x;
            }
        }
        if (src != t) memcpy(t, src, n * tw * 8);
        free(tmp); free(cnt);
    }
    uint64_t u = 0, i = 0;
    while (i < n) {
        uint64_t j = i;
        int lows = 0, recsn = 0;
        uint64_t c = 0;
        while (j < n && memcmp(t + j * tw, t + i * tw, nw * 8) == 0) {
            uint64_t tag = t[j * tw + nw];
            if (tag >> 63) { lows++; c += 1; } else { recsn++; c += tag; }
            j++;
        }
        if (lows > 1 || recsn > 1) { free(t); return UINT64_MAX; }
        memcpy(out_keys + u * nw, t + i * tw, nw * 8);
        out_counts[u] = (uint32_t)c;
        u++;
        i = j;
    }
    free(t);
    return u;
}

/* ------------------------------------------------------------------------- */
/* A14: histogram_tool precise (tool/histogram_tool.cpp:445-527,              */
/* tool/HighFrequencyInsertThreadPool.h:150-177,                              */
/* tool/LowFrequencyQueryThreadPool.h:184-232).  hist has max-min+1 entries,  */
/* index f-min.  The hash map of the tool is a sorted array here.             */

static uint32_t g_cmp_nw;
static int cmp_rec(const void* a, const void* b)
{
    const uint64_t* x = (const uint64_t*)a;
    const uint64_t* y = (const uint64_t*)b;
    for (uint32_t i = 0; i < g_cmp_nw; ++i) {
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    }
    return 0;
}

ORC_API int orc_histogram_precise(const uint8_t* recs, uint64_t nr, const uint64_t* low, uint64_t nl,
                                  uint32_t nw, uint32_t minf, uint32_t maxf, int64_t* hist, int query_only)
{
    const uint64_t rb = orc_record_bytes(nw);
    const uint32_t tw = nw + 1;
    uint64_t* table = (uint64_t*)malloc((nr ? nr : 1) * tw * sizeof(uint64_t));
    if (!table) return -1;
    uint64_t ins = 0;
    const uint64_t hsize = (uint64_t)maxf - minf + 1;
    memset(hist, 0, hsize * sizeof(int64_t));
    for (uint64_t i = 0; i < nr; ++i) {
        const uint8_t* r = recs + i * rb;
        uint32_t c;
        memcpy(&c, r + nw * 8, 4);
        if ((uint64_t)c + 1 < minf) continue;                 /* :161-165 */
        if (!query_only && c >= minf && c <= maxf) hist[c - minf] += 1;  /* :167-170 */
        memcpy(table + ins * tw, r, nw * 8);
        table[ins * tw + nw] = c;
        ins++;
    }
    g_cmp_nw = nw;
    qsort(table, ins, tw * 8, cmp_rec);
    for (uint64_t i = 0; i < nl; ++i) {
        const uint64_t* key = low + i * nw;
        uint64_t lo = 0, hi = ins;
        int found = 0;
        uint64_t c = 0;
        while (lo < hi) {
            uint64_t mid = (lo + hi) / 2;
            int s = cmp_rec(table + mid * tw, key);
            if (s == 0) { found = 1; c = table[mid * tw + nw]; break; }
            if (s < 0) lo = mid + 1; else hi = mid;
        }
        if (found) {                                          /* :212-223 */
            if (c >= minf && c <= maxf) hist[c - minf] -= 1;
            if (c + 1 >= minf && c + 1 <= maxf) hist[c + 1 - minf] += 1;
        } else if (1 >= minf && 1 <= maxf) {                  /* :224-227 */
            hist[1 - minf] += 1;
        }
    }
    free(table);
    return 0;
}

/* histogram_tool approximate: the root records' counts only
 * (tool/ApproximateHighFrequency.h:133-148, record counts outside
 * [minf, maxf] dropped). */
ORC_API void orc_histogram_approximate(const uint8_t* recs, uint64_t nr, uint32_t nw, uint32_t minf, uint32_t maxf,
                                       int64_t* hist)
{
    const uint64_t rb = orc_record_bytes(nw);
    memset(hist, 0, ((uint64_t)maxf - minf + 1) * sizeof(int64_t));
    for (uint64_t i = 0; i < nr; ++i) {
        uint32_t c;
        memcpy(&c, recs + i * rb + nw * 8, 4);
        if (c >= minf && c <= maxf) hist[c - minf] += 1;
    }
}

/* Closed form: histogram of a merged (key, count) table. */
ORC_API void orc_histogram_from_counts(const uint32_t* counts, uint64_t n, uint32_t minf, uint32_t maxf, int64_t* hist)
{
    const uint64_t hsize = (uint64_t)maxf - minf + 1;
    memset(hist, 0, hsize * sizeof(int64_t));
    for (uint64_t i = 0; i < n; ++i) {
        if (counts[i] >= minf && counts[i] <= maxf) hist[counts[i] - minf]++;
    }
}

/* ------------------------------------------------------------------------- */
/* A13: Stage 0 (src/FastqPrefixCounter.h:31-82, src/main.cpp:38-108,175-184) */
/* ------------------------------------------------------------------------- */

/* root prefix = top 8 bits of word 0 (ROOT_BASES = 4, src/definition.h:28) */
ORC_API void orc_prefix_counts(const uint64_t* keys, uint64_t n, uint32_t nw, uint64_t* counts256)
{
    memset(counts256, 0, 256 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) counts256[keys[i * nw] >> 56]++;
}

/* clamp to >= 1000 (src/main.cpp:175-184); returns the average */
ORC_API uint64_t orc_clamp_prefix_counts(uint64_t* counts256)
{
    uint64_t sum = 0;
    for (int i = 0; i < 256; ++i) {
        if (counts256[i] < 1000) counts256[i] = 1000;
        sum += counts256[i];
    }
    return sum / 256;
}

/* LPT (src/main.cpp:38-86): prefixes by count descending, each to the least
 * loaded classifier (ties -> lowest index, as the priority queue's
 * comparator orders by count only; the reference's std::sort is unstable, so
 * equal-count prefix order is not pinned).  Here equal counts keep ascending
 * prefix order. */
ORC_API void orc_lpt(const uint64_t* counts256, uint32_t classifiers, uint8_t* owners, uint64_t* loads)
{
    uint32_t order[256];
    for (int i = 0; i < 256; ++i) order[i] = (uint32_t)i;
    for (int i = 1; i < 256; ++i) {           /* stable insertion sort, descending */
        uint32_t x = order[i];
        int j = i - 1;
        while (j >= 0 && counts256[order[j]] < counts256[x]) { order[j + 1] = order[j]; j--; }
        order[j + 1] = x;
    }
    for (uint32_t c = 0; c < classifiers; ++c) loads[c] = 0;
    for (int i = 0; i < 256; ++i) {
        uint32_t best = 0;
        for (uint32_t c = 1; c < classifiers; ++c) if (loads[c] < loads[best]) best = c;
        owners[order[i]] = (uint8_t)best;
        loads[best] += counts256[order[i]];
    }
}

static uint64_t bit_ceil64(uint64_t v)
{
    if (v <= 1) return 1;
    return 1ULL << (64 - __builtin_clzll(v - 1));
}

/* calculate_bloom_filter_capacity (src/main.cpp:88-108), floor 2^15
 * (src/definition.h:171). */
ORC_API void orc_bloom_capacity(const uint64_t* counts256, uint64_t est_file_size, uint64_t* caps256)
{
    const uint64_t est_total = est_file_size / 3;
    uint64_t total = 0;
    for (int i = 0; i < 256; ++i) total += counts256[i];
    for (int i = 0; i < 256; ++i) {
        double ratio = (double)counts256[i] / (double)total;
        uint64_t est = (uint64_t)((double)est_total * ratio * 4.81 / 64);
        uint64_t c = bit_ceil64(est);
        caps256[i] = c > (1ULL << 15) ? c : (1ULL << 15);
    }
}

/* ------------------------------------------------------------------------- */
/* A16 (learned partitioner, no reference code): the MLP that maps a key to   */
/* one of 4096 CDF buckets.  fp64 restatement of the loss and its gradient,   */
/* for checking the device training kernels.                                   */
/*   input x_j = +1 if bit (63-j) of key word 0 is set else -1 (64 inputs)    */
/*   h1 = relu(W1 x + b1) (64), h2 = relu(W2 h1 + b2) (64),                   */
/*   z = W3 h2 + b3 (4096), loss = mean over the batch of -log softmax(z)[y]   */
/* params layout (floats): W1[64][64] b1[64] W2[64][64] b2[64] W3[4096][64]   */
/* b3[4096]; 274560 values.  Gradients use the same layout.                   */
/* ------------------------------------------------------------------------- */

#define MLP_IN 64
#define MLP_H 64
#define MLP_C 4096
#define MLP_OFF_B1 (MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_C * MLP_H)
#define MLP_NPARAM (MLP_OFF_B3 + MLP_C)

ORC_API uint32_t orc_mlp_nparam(void) { return MLP_NPARAM; }

ORC_API double orc_mlp_loss_grad(const double* p, const uint64_t* keys, const uint32_t* labels, uint32_t nb,
                                 double* grad)
{
    if (grad) memset(grad, 0, sizeof(double) * MLP_NPARAM);
    const double* W1 = p;
    const double* b1 = p + MLP_OFF_B1;
    const double* W2 = p + MLP_OFF_W2;
    const double* b2 = p + MLP_OFF_B2;
    const double* W3 = p + MLP_OFF_W3;
    const double* b3 = p + MLP_OFF_B3;
    double* z = (double*)malloc(sizeof(double) * MLP_C);
    double loss = 0.0;
    for (uint32_t s = 0; s < nb; ++s) {
        double x[MLP_IN], a1[MLP_H], h1[MLP_H], a2[MLP_H], h2[MLP_H];
        for (int j = 0; j < MLP_IN; ++j) x[j] = ((keys[s] >> (63 - j)) & 1ULL) ? 1.0 : -1.0;
        for (int i = 0; i < MLP_H; ++i) {
            double acc = b1[i];
            for (int j = 0; j < MLP_IN; ++j) acc += W1[i * MLP_IN + j] * x[j];
            a1[i] = acc;
            h1[i] = acc > 0 ? acc : 0;
        }
        for (int i = 0; i < MLP_H; ++i) {
            double acc = b2[i];
            for (int j = 0; j < MLP_H; ++j) acc += W2[i * MLP_H + j] * h1[j];
            a2[i] = acc;
            h2[i] = acc > 0 ? acc : 0;
        }
        double mx = -1e300;
        for (int c = 0; c < MLP_C; ++c) {
            double acc = b3[c];
            for (int j = 0; j < MLP_H; ++j) acc += W3[c * MLP_H + j] * h2[j];
            z[c] = acc;
            if (acc > mx) mx = acc;
        }
        double se = 0;
        for (int c = 0; c < MLP_C; ++c) se += exp(z[c] - mx);
        loss += log(se) + mx - z[labels[s]];
        if (!grad) continue;
        double dh2[MLP_H] = {0}, da2[MLP_H], dh1[MLP_H] = {0}, da1[MLP_H];
        for (int c = 0; c < MLP_C; ++c) {
            const double dz = (exp(z[c] - mx) / se - (c == (int)labels[s] ? 1.0 : 0.0)) / nb;
            grad[MLP_OFF_B3 + c] += dz;
            for (int j = 0; j < MLP_H; ++j) {
                grad[MLP_OFF_W3 + c * MLP_H + j] += dz * h2[j];
                dh2[j] += dz * W3[c * MLP_H + j];
            }
        }
        for (int i = 0; i < MLP_H; ++i) da2[i] = a2[i] > 0 ? dh2[i] : 0;
        for (int i = 0; i < MLP_H; ++i) {
            grad[MLP_OFF_B2 + i] += da2[i];
            for (int j = 0; j < MLP_H; ++j) {
                grad[MLP_OFF_W2 + i * MLP_H + j] += da2[i] * h1[j];
                dh1[j] += da2[i] * W2[i * MLP_H + j];
            }
        }
        for (int i = 0; i < MLP_H; ++i) da1[i] = a1[i] > 0 ? dh1[i] : 0;
        for (int i = 0; i < MLP_H; ++i) {
            grad[MLP_OFF_B1 + i] += da1[i];
            for (int j = 0; j < MLP_IN; ++j) grad[i * MLP_IN + j] += da1[i] * x[j];
        }
    }
    free(z);
    return loss / nb;
}
