// kmc_fastq.hip -- FASTQ records parsed on the device.
//
// The reference reader (src/FastqReader.h:140-308) walks the file line by
// line with a four-state machine (header, sequence, plus, quality) and copies
// every sequence line, with its terminating '\n', into the parser's blocks.
// Here a raw FASTQ chunk is copied to HBM as it is and three launches move
// its sequence lines to the end of the resident sequence buffer:
//   k_fq_tiles  per 4096-byte tile: newlines and bytes per line phase
//   k_fq_scan   one workgroup: each tile's line phase and output offset from
//               the running state (line phase, sequence length)
//   k_fq_write  per tile: the bytes of sequence lines to their offsets
// A byte belongs to the line started by the newlines before it; it is copied
// when that line's phase (line index mod 4) is 1.  The state lives in device
// memory, so a record may span chunks and no chunk needs a host round trip.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_kernels.h"

namespace kmc {

namespace {

constexpr int kFqThreads = 256;
constexpr uint32_t kFqTile = kFqThreads * 16;

// Bytes of a span counted by line phase relative to the span's first byte
// (m[r]: bytes after r mod 4 newlines of the span), and the span's newlines
// mod 4.  Spans compose left to right: b's relative phases shift by a's
// newlines.
struct FqSpan {
    uint32_t nl;
    uint32_t m[4];
};

__device__ __forceinline__ FqSpan fq_compose(const FqSpan& a, const FqSpan& b)
{
    FqSpan c;
    c.nl = (a.nl + b.nl) & 3u;
#pragma unroll
    for (int r = 0; r < 4; ++r) c.m[r] = a.m[r] + b.m[(r - a.nl) & 3u];
    return c;
}

// The 16 bytes of thread t of tile `tile` (fewer at the end of the chunk).
__device__ __forceinline__ void fq_load(const uint8_t* raw, uint64_t n, uint64_t at, uint8_t b[16], uint32_t& valid)
{
    if (at + 16 <= n) {
        const uint4 q = *reinterpret_cast<const uint4*>(raw + at);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 16; ++j) b[j] = (uint8_t)(w[j >> 2] >> (8 * (j & 3)));
        valid = 16;
    } else {
        valid = at < n ? (uint32_t)(n - at) : 0u;
#pragma unroll
        for (int j = 0; j < 16; ++j) b[j] = (uint32_t)j < valid ? raw[at + j] : (uint8_t)0;
    }
}

__device__ __forceinline__ FqSpan fq_thread_span(const uint8_t b[16], uint32_t valid)
{
    // four 8-bit counters in one word (at most 16 bytes per thread)
    uint32_t packed = 0, nl = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if ((uint32_t)j < valid) {
            packed += 1u << (8u * (nl & 3u));
            nl += b[j] == '\n' ? 1u : 0u;
        }
    }
    FqSpan s;
    s.nl = nl & 3u;
#pragma unroll
    for (int r = 0; r < 4; ++r) s.m[r] = (packed >> (8 * r)) & 0xFFu;
    return s;
}

// Inclusive composition scan of one span per thread over the workgroup
// (Hillis-Steele in LDS); returns the exclusive prefix and the total.
template <int T>
__device__ __forceinline__ FqSpan fq_block_scan(const FqSpan& v, FqSpan* s_buf, FqSpan* total)
{
    const int t = threadIdx.x;
    s_buf[t] = v;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        FqSpan x = s_buf[t];
        if (t >= off) x = fq_compose(s_buf[t - off], x);
        __syncthreads();
        s_buf[t] = x;
        __syncthreads();
    }
    FqSpan ex{};
    if (t > 0) ex = s_buf[t - 1];
    *total = s_buf[T - 1];
    __syncthreads();
    return ex;
}

__global__ void __launch_bounds__(kFqThreads) k_fq_tiles(const uint8_t* __restrict__ raw, uint64_t n, FqSpan* __restrict__ tiles)
{
    __shared__ FqSpan s_buf[kFqThreads];
    uint8_t b[16];
    uint32_t valid;
    fq_load(raw, n, (uint64_t)blockIdx.x * kFqTile + (uint64_t)threadIdx.x * 16, b, valid);
    FqSpan total;
    fq_block_scan<kFqThreads>(fq_thread_span(b, valid), s_buf, &total);
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

// One workgroup of 1024 threads; tiles <= 1024 * kFqPerThread.  state[0]: the
// sequence length, state[1]: the line phase at the chunk's first byte; both
// are advanced past the chunk.
constexpr int kFqScanThreads = 1024;
constexpr uint32_t kFqPerThread = 16;

__global__ void __launch_bounds__(kFqScanThreads)
k_fq_scan(const FqSpan* __restrict__ tiles, uint32_t ntiles, unsigned long long* __restrict__ state,
          uint32_t* __restrict__ tile_phase, unsigned long long* __restrict__ tile_off)
{
    __shared__ FqSpan s_buf[kFqScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t t0 = t * kFqPerThread;
    FqSpan acc{};
    for (uint32_t j = 0; j < kFqPerThread; ++j)
        if (t0 + j < ntiles) acc = fq_compose(acc, tiles[t0 + j]);
    FqSpan total;
    FqSpan pre = fq_block_scan<kFqScanThreads>(acc, s_buf, &total);
    const unsigned long long off0 = state[0];
    const uint32_t ph0 = (uint32_t)state[1] & 3u;
    // output bytes of a span entered at phase p: those at relative phase (1 - p) mod 4
    for (uint32_t j = 0; j < kFqPerThread; ++j) {
        const uint32_t i = t0 + j;
        if (i >= ntiles) break;
        tile_phase[i] = (ph0 + pre.nl) & 3u;
        tile_off[i] = off0 + pre.m[(1u - ph0) & 3u];
        pre = fq_compose(pre, tiles[i]);
    }
    __syncthreads();
    if (t == 0) {
        state[0] = off0 + total.m[(1u - ph0) & 3u];
        state[1] = (ph0 + total.nl) & 3u;
    }
}

__global__ void __launch_bounds__(kFqThreads)
k_fq_write(const uint8_t* __restrict__ raw, uint64_t n, const uint32_t* __restrict__ tile_phase,
           const unsigned long long* __restrict__ tile_off, uint8_t* __restrict__ seq)
{
    __shared__ FqSpan s_buf[kFqThreads];
    uint8_t b[16];
    uint32_t valid;
    fq_load(raw, n, (uint64_t)blockIdx.x * kFqTile + (uint64_t)threadIdx.x * 16, b, valid);
    const FqSpan mine = fq_thread_span(b, valid);
    FqSpan total;
    const FqSpan pre = fq_block_scan<kFqThreads>(mine, s_buf, &total);
    const uint32_t tph = tile_phase[blockIdx.x];
    uint32_t ph = (tph + pre.nl) & 3u;  // phase of the thread's first byte
    unsigned long long o = tile_off[blockIdx.x] + pre.m[(1u - tph) & 3u];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if ((uint32_t)j < valid) {
            if (ph == 1u) seq[o++] = b[j];
            ph = (ph + (b[j] == '\n' ? 1u : 0u)) & 3u;
        }
    }
}

}  // namespace

uint64_t fastq_tile_scratch_bytes(uint64_t max_chunk)
{
    const uint64_t tiles = (max_chunk + kFqTile - 1) / kFqTile;
    return tiles * (sizeof(FqSpan) + sizeof(uint32_t) + sizeof(unsigned long long)) + 64;
}

hipError_t launch_fastq_parse(const uint8_t* raw, uint64_t n, void* scratch, unsigned long long* state, uint8_t* seq,
                              hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint64_t tiles = (n + kFqTile - 1) / kFqTile;
    if (tiles > (uint64_t)kFqScanThreads * kFqPerThread) return hipErrorInvalidValue;
    unsigned long long* d_off = static_cast<unsigned long long*>(scratch);
    FqSpan* d_tiles = reinterpret_cast<FqSpan*>(d_off + tiles);
    uint32_t* d_phase = reinterpret_cast<uint32_t*>(d_tiles + tiles);
    hipLaunchKernelGGL(k_fq_tiles, dim3((uint32_t)tiles), dim3(kFqThreads), 0, st, raw, n, d_tiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fq_scan, dim3(1), dim3(kFqScanThreads), 0, st, d_tiles, (uint32_t)tiles, state, d_phase, d_off);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fq_write, dim3((uint32_t)tiles), dim3(kFqThreads), 0, st, raw, n, d_phase, d_off, seq);
    return hipGetLastError();
}

}  // namespace kmc
