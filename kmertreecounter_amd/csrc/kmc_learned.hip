// kmc_learned.hip -- the learned partitioner: an MLP CDF model over canonical
// keys, trained on the Stage-0 sample with hand-written HIP kernels, whose
// predictions are turned into monotone key boundaries for the partition.
//
// Model: x_j = +1/-1 from bit (63-j) of key word 0 (64 inputs) -> relu(64) ->
// relu(64) -> 4096 logits; target = the key's quantile bucket in the sorted
// sample; loss = softmax cross entropy.  Weights are kept in fp32 (Adam master
// copy) and used as bf16 in the matrix products, which accumulate in fp32:
// the three 4096-wide products run on MFMA (v_mfma_f32_32x32x16_bf16), the two
// 64x64 layers on fp32 FMA.  Routing does not evaluate the network per key:
// the network's bucket predictions on the sorted sample give 4095 monotone
// boundaries, and keys are routed by binary search over them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kmc_fault.h"
#include "kmc_learned.h"

namespace kmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }
__device__ __forceinline__ float from_bf16(__bf16 x) { return (float)x; }

// ------------------------------------------------------------------ GEMM
// C[M][N] = sum_k A[M][K] * Bt[N][K] (both operands K-contiguous, bf16),
// fp32 output.  One wave per 32x32 tile, four waves per workgroup (64x64).
// MFMA operand layout (32x32x16): lane l supplies A[l%32][8*(l/32)+i] and
// B[k=8*(l/32)+i][n=l%32] = Bt[l%32][8*(l/32)+i], i = 0..7; the accumulator
// element r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32].
__global__ void __launch_bounds__(256)
k_gemm_nt(const __bf16* __restrict__ A, const __bf16* __restrict__ Bt, float* __restrict__ C, uint32_t M, uint32_t N,
          uint32_t K, uint32_t lda, uint32_t ldb, uint32_t ldc, uint32_t ksplit)
{
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t m0 = blockIdx.y * 64 + (wave / 2) * 32;
    const uint32_t n0 = blockIdx.x * 64 + (wave % 2) * 32;
    if (m0 >= M || n0 >= N) return;
    // split-K: blockIdx.z takes K range [kb, kb + K/ksplit) and writes its
    // partial product to slab z of C (slabs M*ldc apart), summed afterwards
    const uint32_t kchunk = K / ksplit;
    const uint32_t kb = blockIdx.z * kchunk, ke = kb + kchunk;
    C += (size_t)blockIdx.z * M * ldc;
    const __bf16* a = A + (size_t)(m0 + lane % 32) * lda + 8 * (lane / 32);
    const __bf16* b = Bt + (size_t)(n0 + lane % 32) * ldb + 8 * (lane / 32);
    f32x16 acc = {};
    for (uint32_t k0 = kb; k0 < ke; k0 += 16) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a + k0);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t row = m0 + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        C[(size_t)row * ldc + n0 + lane % 32] = acc[r];
    }
}

// out[i] = sum over slabs s (in order) of in[s * stride + i]; out may be
// slab 0 itself (each element is read before it is written, by one thread)
__global__ void __launch_bounds__(256)
k_sum_slabs(const float* in, uint32_t nslab, size_t stride, uint32_t n, float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.f;
#pragma unroll 4
    for (uint32_t sl = 0; sl < nslab; ++sl) acc += in[sl * stride + i];
    out[i] = acc;
}

// ------------------------------------------------------------------ layers

__device__ __forceinline__ float xin(uint64_t key, int j) { return ((key >> (63 - j)) & 1ULL) ? 1.0f : -1.0f; }

// Samples per workgroup in the hidden-layer kernels (two per wave group: 256
// workgroups for a batch of 2048).
constexpr uint32_t kSpb = 8;

// Hidden layers for kSpb samples per workgroup: thread (unit i = tid % 64,
// sample group tid / 64) keeps row i of W1 and then of W2 in registers; a1
// of the samples is exchanged through LDS.  h2 is written in both layouts the
// products need (rows for the logits, columns for dW3).  nb is a multiple of kSpb.
__global__ void __launch_bounds__(256)
k_mlp_hidden(const uint64_t* __restrict__ keys, uint32_t nb, const float* __restrict__ params,
             const __bf16* __restrict__ w1b, const __bf16* __restrict__ w2b, float* __restrict__ h1,
             __bf16* __restrict__ h2, __bf16* __restrict__ h2t, uint32_t ldt)
{
    constexpr uint32_t kPer = kSpb / 4;
    __shared__ float s_a[kSpb][kMlpH + 1];
    const uint32_t unit = threadIdx.x % 64, grp = threadIdx.x / 64;
    const uint32_t s0 = blockIdx.x * kSpb;
    float w[kMlpIn];
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) w[j] = from_bf16(w1b[unit * kMlpIn + j]);
    float bias = params[kMlpOffB1 + unit];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t sl = grp + 4 * r;
        const uint64_t key = keys[s0 + sl];
        float acc = bias;
#pragma unroll
        for (int j = 0; j < kMlpIn; ++j) acc += ((key >> (63 - j)) & 1ULL) ? w[j] : -w[j];
        const float a = acc > 0.f ? acc : 0.f;
        h1[(size_t)(s0 + sl) * kMlpH + unit] = a;
        s_a[sl][unit] = a;
    }
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) w[j] = from_bf16(w2b[unit * kMlpH + j]);
    bias = params[kMlpOffB2 + unit];
    __syncthreads();
    float out[kPer];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t sl = grp + 4 * r;
        float acc = bias;
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) acc += w[j] * s_a[sl][j];
        out[r] = acc > 0.f ? acc : 0.f;
        h2[(size_t)(s0 + sl) * kMlpH + unit] = to_bf16(out[r]);
    }
    if (!h2t) return;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) s_a[grp + 4 * r][unit] = out[r];
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < kSpb * kMlpH; idx += 256) {
        const uint32_t u = idx / kSpb, sl = idx % kSpb;
        h2t[(size_t)u * ldt + s0 + sl] = to_bf16(s_a[sl][u]);
    }
}

// Reductions over a 256-thread workgroup: within each wave by shuffles, then
// the four wave results through LDS (s4 holds four floats; every thread gets
// the result).
__device__ __forceinline__ float block256_max(float x, float* s4)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
    __syncthreads();  // s4 may still be read from the previous reduction
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = x;
    __syncthreads();
    return fmaxf(fmaxf(s4[0], s4[1]), fmaxf(s4[2], s4[3]));
}
__device__ __forceinline__ float block256_sum(float x, float* s4)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) x += __shfl_xor(x, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = x;
    __syncthreads();
    return (s4[0] + s4[1]) + (s4[2] + s4[3]);
}

// Row softmax + cross entropy: one workgroup per sample, the row held in
// registers.  dz = (p - onehot)/nb in bf16 (row layout); loss summed into *loss.
__global__ void __launch_bounds__(256)
k_softmax_ce(const float* __restrict__ logits, const float* __restrict__ b3, const uint32_t* __restrict__ labels,
             uint32_t nb, __bf16* __restrict__ dz, float* __restrict__ row_loss)
{
    constexpr int kPer = kMlpC / 256;
    __shared__ float s4[4];
    const uint32_t row = blockIdx.x;
    const float* z = logits + (size_t)row * kMlpC;
    float v[kPer];
    float mx = -3.0e38f;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t c = threadIdx.x + 256 * r;
        v[r] = z[c] + b3[c];
        mx = fmaxf(mx, v[r]);
    }
    mx = block256_max(mx, s4);
    float se = 0.f;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        v[r] = __expf(v[r] - mx);
        se += v[r];
    }
    se = block256_sum(se, s4);
    const uint32_t y = labels[row];
    const float inv = 1.0f / se, invb = 1.0f / (float)nb;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t c = threadIdx.x + 256 * r;
        dz[(size_t)row * kMlpC + c] = to_bf16((v[r] * inv - (c == y ? 1.0f : 0.0f)) * invb);
    }
    // one loss term per row (same-address atomics from every row serialised
    // in the L2 and took most of this kernel's time); k_step_end sums them
    if (threadIdx.x == 0) row_loss[row] = (__logf(se) + mx - (z[y] + b3[y])) * invb;
}

// 64x64 tile transpose of a bf16 matrix: out[c][r] = in[r][c]
__global__ void __launch_bounds__(256)
k_transpose_bf16(const __bf16* __restrict__ in, __bf16* __restrict__ out, uint32_t rows, uint32_t cols)
{
    __shared__ __bf16 tile[64][65];
    const uint32_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += 256) {
        const uint32_t r = i / 64, c = i % 64;
        tile[r][c] = in[(size_t)(r0 + r) * cols + c0 + c];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += 256) {
        const uint32_t c = i / 64, r = i % 64;
        out[(size_t)(c0 + c) * rows + r0 + r] = tile[r][c];
    }
}

// db3[c] = sum over the batch of dz[b][c] (from the transposed dz)
__global__ void __launch_bounds__(256)
k_row_sum_bf16(const __bf16* __restrict__ m, uint32_t cols, float* __restrict__ out)
{
    __shared__ float s4[4];
    const __bf16* row = m + (size_t)blockIdx.x * cols;
    float acc = 0.f;
    for (uint32_t i = threadIdx.x; i < cols; i += 256) acc += from_bf16(row[i]);
    acc = block256_sum(acc, s4);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// Backward through the hidden layers, kSpb samples per workgroup (nb a
// multiple of kSpb): da2 = dh2 * [a2 > 0], da1 = (W2^T da2) * [a1 > 0].
__global__ void __launch_bounds__(256)
k_mlp_back_hidden(const float* __restrict__ dh2, const __bf16* __restrict__ h2, const float* __restrict__ h1,
                  const __bf16* __restrict__ w2b, uint32_t nb, float* __restrict__ da2, float* __restrict__ da1)
{
    constexpr uint32_t kPer = kSpb / 4;
    __shared__ float s_d[kSpb][kMlpH + 1];
    const uint32_t unit = threadIdx.x % 64, grp = threadIdx.x / 64;
    const uint32_t s0 = blockIdx.x * kSpb;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t sl = grp + 4 * r;
        const size_t e = (size_t)(s0 + sl) * kMlpH + unit;
        const float d = from_bf16(h2[e]) > 0.f ? dh2[e] : 0.f;
        da2[e] = d;
        s_d[sl][unit] = d;
    }
    float w[kMlpH];  // column `unit` of W2: w[i] = W2[i][unit]
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) w[i] = from_bf16(w2b[i * kMlpH + unit]);
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t sl = grp + 4 * r;
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < kMlpH; ++i) acc += w[i] * s_d[sl][i];
        const size_t e = (size_t)(s0 + sl) * kMlpH + unit;
        da1[e] = h1[e] > 0.f ? acc : 0.f;
    }
}

// Gradients of W1, b1, W2, b2, one workgroup per kSgChunk samples writing
// its partial sums to part[blockIdx.x][kSmallGrads] (summed in order afterwards).
// Output o: [0, 4096) dW2[i][j]; [4096, 8192) dW1[i][j]; then db2, db1.
constexpr uint32_t kSmallGrads = 2 * kMlpH * kMlpH + 2 * kMlpH;
constexpr uint32_t kSgChunk = 8;  // 256 workgroups for a batch of 2048
__global__ void __launch_bounds__(256)
k_mlp_small_grads(const float* __restrict__ da2, const float* __restrict__ h1, const float* __restrict__ da1,
                  const uint64_t* __restrict__ keys, float* __restrict__ part)
{
    __shared__ float s_da2[kSgChunk][kMlpH], s_h1[kSgChunk][kMlpH], s_da1[kSgChunk][kMlpH];
    __shared__ uint64_t s_key[kSgChunk];
    const uint32_t s0 = blockIdx.x * kSgChunk;
    for (uint32_t idx = threadIdx.x; idx < kSgChunk * kMlpH; idx += 256) {
        const uint32_t sl = idx / kMlpH, u = idx % kMlpH;
        const size_t e = (size_t)(s0 + sl) * kMlpH + u;
        s_da2[sl][u] = da2[e];
        s_h1[sl][u] = h1[e];
        s_da1[sl][u] = da1[e];
    }
    if (threadIdx.x < kSgChunk) s_key[threadIdx.x] = keys[s0 + threadIdx.x];
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * kSmallGrads;
    for (uint32_t o = threadIdx.x; o < kSmallGrads; o += 256) {
        float acc = 0.f;
        if (o < kMlpH * kMlpH) {
            const uint32_t i = o / kMlpH, j = o % kMlpH;
#pragma unroll 8
            for (uint32_t b = 0; b < kSgChunk; ++b) acc += s_da2[b][i] * s_h1[b][j];
        } else if (o < 2 * kMlpH * kMlpH) {
            const uint32_t u = o - kMlpH * kMlpH, i = u / kMlpIn, j = u % kMlpIn;
#pragma unroll 8
            for (uint32_t b = 0; b < kSgChunk; ++b) acc += ((s_key[b] >> (63 - j)) & 1ULL) ? s_da1[b][i] : -s_da1[b][i];
        } else if (o < 2 * kMlpH * kMlpH + kMlpH) {
            const uint32_t i = o - 2 * kMlpH * kMlpH;
            for (uint32_t b = 0; b < kSgChunk; ++b) acc += s_da2[b][i];
        } else {
            const uint32_t i = o - 2 * kMlpH * kMlpH - kMlpH;
            for (uint32_t b = 0; b < kSgChunk; ++b) acc += s_da1[b][i];
        }
        out[o] = acc;
    }
}

// grad[W1, b1, W2, b2] from the per-workgroup partials of k_mlp_small_grads:
// 64 outputs per workgroup (consecutive, so each partial row is read
// coalesced), the partials split over the four waves, eight loads in flight
// per thread, the quarters combined through LDS; all in a fixed order
__global__ void __launch_bounds__(256)
k_small_grads_reduce(const float* __restrict__ part, uint32_t nblk, float* __restrict__ grad)
{
    __shared__ float s_q[4][64];
    const uint32_t lane = threadIdx.x % 64, q = threadIdx.x / 64;
    const uint32_t o = blockIdx.x * 64 + lane;
    const uint32_t per = (nblk + 3) / 4, b0 = q * per, b1 = min(nblk, b0 + per);
    float a8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (o < kSmallGrads) {
        uint32_t b = b0;
        for (; b + 8 <= b1; b += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) a8[u] += part[(size_t)(b + u) * kSmallGrads + o];
        }
        for (; b < b1; ++b) a8[(b - b0) & 7] += part[(size_t)b * kSmallGrads + o];
    }
    s_q[q][lane] = ((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]));
    __syncthreads();
    if (q != 0 || o >= kSmallGrads) return;
    const float acc = (s_q[0][lane] + s_q[1][lane]) + (s_q[2][lane] + s_q[3][lane]);
    uint32_t dst;
    if (o < kMlpH * kMlpH) dst = kMlpOffW2 + o;
    else if (o < 2 * kMlpH * kMlpH) dst = o - kMlpH * kMlpH;
    else if (o < 2 * kMlpH * kMlpH + kMlpH) dst = kMlpOffB2 + (o - 2 * kMlpH * kMlpH);
    else dst = kMlpOffB1 + (o - 2 * kMlpH * kMlpH - kMlpH);
    grad[dst] = acc;
}

// Adam on the fp32 master weights, then the bf16 copies the products read:
// W1, W2 (row major), W3 [C][H] and W3^T [H][C].
__global__ void __launch_bounds__(256)
k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, uint32_t n,
       const uint32_t* __restrict__ step, uint32_t steps, float lr0, float b1, float b2, float eps,
       __bf16* __restrict__ w1b, __bf16* __restrict__ w2b, __bf16* __restrict__ w3b, __bf16* __restrict__ w3tb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // step size with cosine decay to zero, and Adam's bias corrections, from the step counter
    const uint32_t t = *step;
    const float lr = lr0 * 0.5f * (1.0f + cosf(3.14159265f * (float)t / (float)steps));
    const float c1 = 1.f - powf(b1, (float)(t + 1)), c2 = 1.f - powf(b2, (float)(t + 1));
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    write_bf16_copy(i, pi, w1b, w2b, w3b, w3tb);
}

// End of a step: the batch loss (the rows' terms summed in a fixed order) into
// loss[*step], then the step counter moves on (the next step reads it).  One
// workgroup of 256 threads.
__global__ void __launch_bounds__(256)
k_step_end(const float* __restrict__ row_loss, uint32_t nb, float* __restrict__ loss, uint32_t* __restrict__ step)
{
    __shared__ float s4[4];
    float acc = 0.f;
    for (uint32_t i = threadIdx.x; i < nb; i += 256) acc += row_loss[i];
    acc = block256_sum(acc, s4);
    if (threadIdx.x == 0) {
        const uint32_t t = step ? *step : 0u;
        loss[t] = acc;
        if (step) *step = t + 1;
    }
}

__global__ void __launch_bounds__(256) k_zero_f32(float* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void __launch_bounds__(256)
k_to_bf16_copies(const float* __restrict__ p, uint32_t n, __bf16* __restrict__ w1b, __bf16* __restrict__ w2b,
                 __bf16* __restrict__ w3b, __bf16* __restrict__ w3tb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) write_bf16_copy(i, p[i], w1b, w2b, w3b, w3tb);
}

// Batch of sorted-sample indices from a counter-based hash; label = quantile bucket
__global__ void __launch_bounds__(256)
k_gather_batch(const uint64_t* __restrict__ sorted, uint64_t n, uint64_t seed, const uint32_t* __restrict__ step,
               uint32_t nb, uint64_t* __restrict__ keys, uint32_t* __restrict__ labels)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    uint64_t z = seed * 0x100000001B3ULL + *step + 0x9E3779B97F4A7C15ULL * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const uint64_t idx = __umul64hi(z, n);
    keys[i] = sorted[idx];
    labels[i] = (uint32_t)((idx * (uint64_t)kMlpC) / n);
}

// Expected bucket of each row: sum_c c * softmax(logits + b3)_c.  A model that
// spreads its probability over neighbouring buckets still places the key
// between them, which the argmax would not.
__global__ void __launch_bounds__(256)
k_expect(const float* __restrict__ logits, const float* __restrict__ b3, uint32_t rows, float* __restrict__ out)
{
    constexpr int kPer = kMlpC / 256;
    __shared__ float s_red[256];
    __shared__ float s_red2[256];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    const float* z = logits + (size_t)row * kMlpC;
    float v[kPer];
    float mx = -3.0e38f;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t c = threadIdx.x + 256 * r;
        v[r] = z[c] + b3[c];
        mx = fmaxf(mx, v[r]);
    }
    s_red[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] = fmaxf(s_red[threadIdx.x], s_red[threadIdx.x + w]);
        __syncthreads();
    }
    mx = s_red[0];
    __syncthreads();
    float se = 0.f, sc = 0.f;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const float e = __expf(v[r] - mx);
        se += e;
        sc += e * (float)(threadIdx.x + 256 * r);
    }
    s_red[threadIdx.x] = se;
    s_red2[threadIdx.x] = sc;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_red[threadIdx.x] += s_red[threadIdx.x + w];
            s_red2[threadIdx.x] += s_red2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[row] = s_red2[0] / s_red[0];
}

// Most likely bucket of each row (lowest index on ties), as a float.
__global__ void __launch_bounds__(256)
k_argmax(const float* __restrict__ logits, const float* __restrict__ b3, uint32_t rows, float* __restrict__ out)
{
    __shared__ float s_v[256];
    __shared__ uint32_t s_i[256];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    float best = -3.0e38f;
    uint32_t bi = 0;
    for (uint32_t c = threadIdx.x; c < kMlpC; c += 256) {
        const float v = logits[(size_t)row * kMlpC + c] + b3[c];
        if (v > best) { best = v; bi = c; }
    }
    s_v[threadIdx.x] = best;
    s_i[threadIdx.x] = bi;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const float o = s_v[threadIdx.x + w];
            const uint32_t oi = s_i[threadIdx.x + w];
            if (o > s_v[threadIdx.x] || (o == s_v[threadIdx.x] && oi < s_i[threadIdx.x])) {
                s_v[threadIdx.x] = o;
                s_i[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[row] = (float)s_i[0];
}

// ------------------------------------------------------------------ host side

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
    template <typename T> T* as() { return reinterpret_cast<T*>(p); }
    hipError_t alloc(size_t bytes) { return dmalloc(&p, bytes); }
};

#define LCHK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) { err = std::string(#expr) + ": " + hipGetErrorString(e_); return -2; } } while (0)

}  // namespace

struct MlpWork {
    uint32_t cap = 0;  // batch capacity (multiple of 64)
    DevBuf params, grad, m, v, w1b, w2b, w3b, w3tb;
    DevBuf keys, labels, h1, h2, h2t, logits, dz, dzt, dh2, da2, da1, loss, row_loss, part, step;
    uint32_t loss_slots = 1;  // training: one loss per step

    // split-K partial slabs and per-workgroup partial gradients
    static constexpr uint32_t kSplitW3 = 4, kSplitDh2 = 16;

    int init(uint32_t batch, std::string& err, bool train = true, uint32_t steps = 1)
    {
        cap = batch;
        loss_slots = steps ? steps : 1;
        LCHK(params.alloc(sizeof(float) * kMlpNParam));
        LCHK(grad.alloc(sizeof(float) * kMlpNParam));
        LCHK(m.alloc(sizeof(float) * kMlpNParam));
        LCHK(v.alloc(sizeof(float) * kMlpNParam));
        LCHK(w1b.alloc(2 * kMlpH * kMlpIn));
        LCHK(w2b.alloc(2 * kMlpH * kMlpH));
        LCHK(w3b.alloc(2 * kMlpC * kMlpH));
        LCHK(w3tb.alloc(2 * kMlpC * kMlpH));
        LCHK(keys.alloc(8ULL * batch));
        LCHK(labels.alloc(4ULL * batch));
        LCHK(h1.alloc(4ULL * batch * kMlpH));
        LCHK(h2.alloc(2ULL * batch * kMlpH));
        LCHK(h2t.alloc(2ULL * batch * kMlpH));
        LCHK(logits.alloc(4ULL * batch * kMlpC));
        LCHK(dz.alloc(2ULL * batch * kMlpC));
        LCHK(dzt.alloc(2ULL * batch * kMlpC));
        LCHK(loss.alloc(4ULL * loss_slots));
        LCHK(row_loss.alloc(4ULL * batch));
        LCHK(step.alloc(4));
        if (!train) return 0;
        LCHK(dh2.alloc(4ULL * kSplitDh2 * batch * kMlpH));
        LCHK(da2.alloc(4ULL * batch * kMlpH));
        LCHK(da1.alloc(4ULL * batch * kMlpH));
        const size_t np = std::max<size_t>((size_t)kSplitW3 * kMlpC * kMlpH, (size_t)(batch / kSgChunk) * kSmallGrads);
        LCHK(part.alloc(4 * np));
        return 0;
    }
};

// C = A . Bt^T; with ksplit > 1 (K a multiple of 16 * ksplit) slab z of C
// (slabs M*ldc apart) receives the product over the z-th K slice
static hipError_t gemm_nt(const __bf16* A, const __bf16* Bt, float* C, uint32_t M, uint32_t N, uint32_t K, uint32_t lda,
                          uint32_t ldb, uint32_t ldc, hipStream_t st, uint32_t ksplit = 1)
{
    if (ksplit == 0 || K % (16 * ksplit)) return hipErrorInvalidValue;
    dim3 grid((N + 63) / 64, (M + 63) / 64, ksplit);
    hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, st, A, Bt, C, M, N, K, lda, ldb, ldc, ksplit);
    return hipGetLastError();
}

// loss and gradient of the batch already in w.keys / w.labels (nb <= cap,
// multiple of 64); each row's loss term goes to row_loss (k_step_end sums them)
static int mlp_loss_grad(MlpWork& w, uint32_t nb, hipStream_t st, std::string& err)
{
    hipLaunchKernelGGL(k_mlp_hidden, dim3(nb / kSpb), dim3(256), 0, st, w.keys.as<uint64_t>(), nb,
                       w.params.as<float>(), w.w1b.as<__bf16>(), w.w2b.as<__bf16>(), w.h1.as<float>(),
                       w.h2.as<__bf16>(), w.h2t.as<__bf16>(), nb);
    LCHK(hipGetLastError());
    // logits[nb][C] = h2[nb][H] . W3[C][H]^T
    LCHK(gemm_nt(w.h2.as<__bf16>(), w.w3b.as<__bf16>(), w.logits.as<float>(), nb, kMlpC, kMlpH, kMlpH, kMlpH, kMlpC, st));
    hipLaunchKernelGGL(k_softmax_ce, dim3(nb), dim3(256), 0, st, w.logits.as<float>(),
                       w.params.as<float>() + kMlpOffB3, w.labels.as<uint32_t>(), nb, w.dz.as<__bf16>(),
                       w.row_loss.as<float>());
    LCHK(hipGetLastError());
    hipLaunchKernelGGL(k_transpose_bf16, dim3(kMlpC / 64, nb / 64), dim3(256), 0, st, w.dz.as<__bf16>(),
                       w.dzt.as<__bf16>(), nb, kMlpC);
    LCHK(hipGetLastError());
    float* g = w.grad.as<float>();
    // dW3[C][H] = dz^T[C][nb] . h2^T[H][nb]^T, split over the batch
    const uint32_t sw3 = nb % (16 * MlpWork::kSplitW3) ? 1 : MlpWork::kSplitW3;
    LCHK(gemm_nt(w.dzt.as<__bf16>(), w.h2t.as<__bf16>(), w.part.as<float>(), kMlpC, kMlpH, nb, nb, nb, kMlpH, st, sw3));
    hipLaunchKernelGGL(k_sum_slabs, dim3(kMlpC * kMlpH / 256), dim3(256), 0, st, w.part.as<float>(), sw3,
                       (size_t)kMlpC * kMlpH, (uint32_t)(kMlpC * kMlpH), g + kMlpOffW3);
    LCHK(hipGetLastError());
    hipLaunchKernelGGL(k_row_sum_bf16, dim3(kMlpC), dim3(256), 0, st, w.dzt.as<__bf16>(), nb, g + kMlpOffB3);
    LCHK(hipGetLastError());
    // dh2[nb][H] = dz[nb][C] . W3^T[H][C]^T, split over the 4096 buckets
    LCHK(gemm_nt(w.dz.as<__bf16>(), w.w3tb.as<__bf16>(), w.dh2.as<float>(), nb, kMlpH, kMlpC, kMlpC, kMlpC, kMlpH, st,
                 MlpWork::kSplitDh2));
    // the slabs are summed into slab 0 (element-parallel, fixed order)
    hipLaunchKernelGGL(k_sum_slabs, dim3(nb * kMlpH / 256), dim3(256), 0, st, w.dh2.as<float>(), MlpWork::kSplitDh2,
                       (size_t)nb * kMlpH, nb * kMlpH, w.dh2.as<float>());
    LCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_back_hidden, dim3(nb / kSpb), dim3(256), 0, st, w.dh2.as<float>(), w.h2.as<__bf16>(),
                       w.h1.as<float>(), w.w2b.as<__bf16>(), nb, w.da2.as<float>(), w.da1.as<float>());
    LCHK(hipGetLastError());
    // the per-workgroup partials reuse the dW3 slabs, which are consumed above
    hipLaunchKernelGGL(k_mlp_small_grads, dim3(nb / kSgChunk), dim3(256), 0, st, w.da2.as<float>(), w.h1.as<float>(),
                       w.da1.as<float>(), w.keys.as<uint64_t>(), w.part.as<float>());
    LCHK(hipGetLastError());
    hipLaunchKernelGGL(k_small_grads_reduce, dim3((kSmallGrads + 63) / 64), dim3(256), 0, st, w.part.as<float>(),
                       nb / kSgChunk, g);
    LCHK(hipGetLastError());
    return 0;
}

static int upload_params(MlpWork& w, const float* params, hipStream_t st, std::string& err)
{
    LCHK(dcopy_async(w.params.p, params, sizeof(float) * kMlpNParam, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_to_bf16_copies, dim3((kMlpNParam + 255) / 256), dim3(256), 0, st, w.params.as<float>(),
                       (uint32_t)kMlpNParam, w.w1b.as<__bf16>(), w.w2b.as<__bf16>(), w.w3b.as<__bf16>(),
                       w.w3tb.as<__bf16>());
    LCHK(hipGetLastError());
    return 0;
}

void mlp_init_params(uint64_t seed, std::vector<float>& p)
{
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a SplitMix64 stream, biases zero
    p.assign(kMlpNParam, 0.f);
    uint64_t s = seed ^ 0x5DEECE66DULL;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    auto unif = [&](float a) { return (float)(((next() >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
    for (uint32_t i = 0; i < kMlpH * kMlpIn; ++i) p[i] = unif(1.0f / std::sqrt((float)kMlpIn));
    for (uint32_t i = 0; i < kMlpH * kMlpH; ++i) p[kMlpOffW2 + i] = unif(1.0f / std::sqrt((float)kMlpH));
    for (uint32_t i = 0; i < (uint32_t)kMlpC * kMlpH; ++i) p[kMlpOffW3 + i] = unif(1.0f / std::sqrt((float)kMlpH));
}

int mlp_debug_loss_grad(const float* params, const uint64_t* keys, const uint32_t* labels, uint32_t nb, float* grad,
                        float* loss, std::string& err)
{
    if (nb == 0 || nb % 64) { err = "batch must be a positive multiple of 64"; return -1; }
    MlpWork w;
    if (w.init(nb, err)) return -3;
    hipStream_t st = nullptr;
    if (upload_params(w, params, st, err)) return -2;
    LCHK(dcopy(w.keys.p, keys, 8ULL * nb, hipMemcpyHostToDevice));
    LCHK(dcopy(w.labels.p, labels, 4ULL * nb, hipMemcpyHostToDevice));
    LCHK(hipMemset(w.grad.p, 0, sizeof(float) * kMlpNParam));
    LCHK(hipMemset(w.loss.p, 0, 4));
    if (mlp_loss_grad(w, nb, st, err)) return -2;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(256), 0, st, w.row_loss.as<float>(), nb, w.loss.as<float>(),
                       (uint32_t*)nullptr);
    LCHK(hipGetLastError());
    LCHK(hipDeviceSynchronize());
    LCHK(dcopy(grad, w.grad.p, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost));
    LCHK(dcopy(loss, w.loss.p, 4, hipMemcpyDeviceToHost));
    return 0;
}

int mlp_debug_gemm(const uint16_t* A, const uint16_t* Bt, float* C, uint32_t M, uint32_t N, uint32_t K, std::string& err)
{
    if (M % 64 || N % 64 || K % 16 || !M || !N || !K) { err = "M, N multiples of 64 and K of 16 required"; return -1; }
    DevBuf a, b, c;
    LCHK(a.alloc(2ULL * M * K));
    LCHK(b.alloc(2ULL * N * K));
    LCHK(c.alloc(4ULL * M * N));
    LCHK(dcopy(a.p, A, 2ULL * M * K, hipMemcpyHostToDevice));
    LCHK(dcopy(b.p, Bt, 2ULL * N * K, hipMemcpyHostToDevice));
    LCHK(gemm_nt(a.as<__bf16>(), b.as<__bf16>(), c.as<float>(), M, N, K, K, K, N, nullptr));
    LCHK(hipDeviceSynchronize());
    LCHK(dcopy(C, c.p, 4ULL * M * N, hipMemcpyDeviceToHost));
    return 0;
}

// One training step on `st`: batch, loss and gradient, Adam, step counter.
// Everything that changes from step to step is read from the device (the
// step counter), so a run of steps can be captured once and replayed.
static int train_step(MlpWork& w, const uint64_t* d_sorted, uint64_t n, const MlpTrainOpts& o, hipStream_t st,
                      std::string& err)
{
    const uint32_t* step = w.step.as<uint32_t>();
    hipLaunchKernelGGL(k_gather_batch, dim3((o.batch + 255) / 256), dim3(256), 0, st, d_sorted, n, o.seed, step,
                       o.batch, w.keys.as<uint64_t>(), w.labels.as<uint32_t>());
    LCHK(hipGetLastError());
    if (mlp_loss_grad(w, o.batch, st, err)) return -2;
    hipLaunchKernelGGL(k_adam, dim3((kMlpNParam + 255) / 256), dim3(256), 0, st, w.params.as<float>(),
                       w.grad.as<float>(), w.m.as<float>(), w.v.as<float>(), (uint32_t)kMlpNParam, step, o.steps, o.lr,
                       0.9f, 0.999f, 1e-8f, w.w1b.as<__bf16>(), w.w2b.as<__bf16>(), w.w3b.as<__bf16>(),
                       w.w3tb.as<__bf16>());
    LCHK(hipGetLastError());
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(256), 0, st, w.row_loss.as<float>(), o.batch, w.loss.as<float>(),
                       w.step.as<uint32_t>());
    LCHK(hipGetLastError());
    return 0;
}

int mlp_train(const uint64_t* d_sorted, uint64_t n, const MlpTrainOpts& o, hipStream_t st, std::vector<float>& params,
              std::vector<float>& loss_curve, std::string& err)
{
    MlpWork w;
    if (w.init(o.batch, err, true, o.steps)) return -3;
    mlp_init_params(o.seed, params);
    if (upload_params(w, params.data(), st, err)) return -2;
    hipLaunchKernelGGL(k_zero_f32, dim3((kMlpNParam + 255) / 256), dim3(256), 0, st, w.m.as<float>(), (uint32_t)kMlpNParam);
    hipLaunchKernelGGL(k_zero_f32, dim3((kMlpNParam + 255) / 256), dim3(256), 0, st, w.v.as<float>(), (uint32_t)kMlpNParam);
    hipLaunchKernelGGL(k_zero_f32, dim3((o.steps + 255) / 256), dim3(256), 0, st, w.loss.as<float>(), o.steps);
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(1), 0, st, w.step.as<float>(), 1u);  // step counter 0
    LCHK(hipGetLastError());
    // runs of kGraphSteps steps are captured once and replayed (each step is
    // about 16 short kernels: launched one by one, the launches took about
    // half of the training time); the steps left over run eagerly
    // (KMC_TRAIN_GRAPH=0: every step launched eagerly, for profilers that do
    // not follow graph capture)
    constexpr uint32_t kGraphSteps = 50;
    const char* tg = std::getenv("KMC_TRAIN_GRAPH");
    const uint32_t runs = (tg && tg[0] == '0') ? 0u : o.steps / kGraphSteps;
    uint32_t done = 0;
    if (runs > 1) {
        hipStream_t cs = nullptr;
        LCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        // the capture stream starts after what `st` has queued, and `st` waits for the replays
        hipEvent_t ev = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ev, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, ev, 0);
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        int rc = 0;
        if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (uint32_t i = 0; i < kGraphSteps && !rc; ++i) rc = train_step(w, d_sorted, n, o, cs, err);
            const hipError_t ee = hipStreamEndCapture(cs, &g);
            if (!rc && ee != hipSuccess) e = ee;
        }
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        for (uint32_t r = 0; r < runs && !rc && e == hipSuccess; ++r) e = hipGraphLaunch(ge, cs);
        if (!rc && e == hipSuccess) e = hipEventRecord(ev, cs);
        if (!rc && e == hipSuccess) e = hipStreamWaitEvent(st, ev, 0);
        if (e == hipSuccess) e = hipStreamSynchronize(cs);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        if (ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
        if (e != hipSuccess) {
            err = std::string("training graph: ") + hipGetErrorString(e);
            return -2;
        }
        done = runs * kGraphSteps;
    }
    for (uint32_t step = done; step < o.steps; ++step)
        if (train_step(w, d_sorted, n, o, st, err)) return -2;
    // the loss every 50 steps and at the last step
    std::vector<float> all(o.steps);
    LCHK(dcopy_async(all.data(), w.loss.p, sizeof(float) * o.steps, hipMemcpyDeviceToHost, st));
    LCHK(dcopy_async(params.data(), w.params.p, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    loss_curve.clear();
    for (uint32_t step = 0; step < o.steps; ++step)
        if (step % 50 == 0 || step + 1 == o.steps) loss_curve.push_back(all[step]);
    return 0;
}

int mlp_predict(const float* params, const uint64_t* d_keys, uint64_t n, bool expect, hipStream_t st,
                std::vector<float>& pred, std::string& err)
{
    const uint32_t chunk = 8192;
    MlpWork w;
    if (w.init(chunk, err, false)) return -3;
    if (upload_params(w, params, st, err)) return -2;
    DevBuf out;
    LCHK(out.alloc(4ULL * (n + chunk)));
    for (uint64_t i = 0; i < n; i += chunk) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, n - i);
        const uint32_t mp = (m + 63) / 64 * 64;
        LCHK(hipMemsetAsync(w.keys.p, 0, 8ULL * mp, st));
        LCHK(dcopy_async(w.keys.p, d_keys + i, 8ULL * m, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_mlp_hidden, dim3(mp / kSpb), dim3(256), 0, st, w.keys.as<uint64_t>(), mp,
                           w.params.as<float>(), w.w1b.as<__bf16>(), w.w2b.as<__bf16>(), w.h1.as<float>(),
                           w.h2.as<__bf16>(), (__bf16*)nullptr, mp);
        LCHK(hipGetLastError());
        LCHK(gemm_nt(w.h2.as<__bf16>(), w.w3b.as<__bf16>(), w.logits.as<float>(), mp, kMlpC, kMlpH, kMlpH, kMlpH,
                     kMlpC, st));
        if (expect)
            hipLaunchKernelGGL(k_expect, dim3(m), dim3(256), 0, st, w.logits.as<float>(),
                               w.params.as<float>() + kMlpOffB3, m, out.as<float>() + i);
        else
            hipLaunchKernelGGL(k_argmax, dim3(m), dim3(256), 0, st, w.logits.as<float>(),
                               w.params.as<float>() + kMlpOffB3, m, out.as<float>() + i);
        LCHK(hipGetLastError());
    }
    pred.resize(n);
    LCHK(dcopy_async(pred.data(), out.p, 4ULL * n, hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    return 0;
}

int mlp_debug_fit(const uint64_t* sorted, uint64_t n, const MlpTrainOpts& o, uint64_t* bounds, float* loss_curve,
                  uint32_t loss_cap, std::string& err)
{
    std::vector<uint64_t> keys(sorted, sorted + n);
    if (o.distinct) keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    DevBuf d;
    LCHK(d.alloc(8ULL * keys.size()));
    LCHK(dcopy(d.p, keys.data(), 8ULL * keys.size(), hipMemcpyHostToDevice));
    std::vector<float> params, curve;
    int rc = mlp_train(d.as<uint64_t>(), keys.size(), o, nullptr, params, curve, err);
    if (rc) return rc;
    std::vector<float> pred;
    rc = mlp_predict(params.data(), d.as<uint64_t>(), keys.size(), o.expect, nullptr, pred, err);
    if (rc) return rc;
    std::vector<uint64_t> b;
    boundaries_from_predictions(keys, pred, b);
    std::memcpy(bounds, b.data(), sizeof(uint64_t) * b.size());
    for (uint32_t i = 0; i < loss_cap && i < curve.size(); ++i) loss_curve[i] = curve[i];
    return 0;
}

void boundaries_from_predictions(const std::vector<uint64_t>& sorted, const std::vector<float>& pred,
                                 std::vector<uint64_t>& b)
{
    // Isotonic (non-decreasing) least-squares fit of the predicted bucket
    // against key order, by pooling adjacent violators: a stray prediction is
    // averaged into its neighbours instead of moving every later boundary.
    const size_t n = sorted.size();
    std::vector<double> val;
    std::vector<uint64_t> len;
    val.reserve(1024);
    len.reserve(1024);
    for (size_t i = 0; i < n; ++i) {
        double v = pred[i];
        uint64_t l = 1;
        while (!val.empty() && val.back() > v) {
            v = (val.back() * len.back() + v * l) / (double)(len.back() + l);
            l += len.back();
            val.pop_back();
            len.pop_back();
        }
        val.push_back(v);
        len.push_back(l);
    }
    b.assign(kMlpC + 1, ~0ULL);
    b[0] = 0;
    uint32_t next = 1;
    size_t i = 0;
    for (size_t blk = 0; blk < val.size() && next < kMlpC; ++blk) {
        // a pooled block's keys belong to bucket round(val); its first key opens
        // every bucket up to that one
        const uint32_t q = (uint32_t)std::min<double>(kMlpC - 1, std::floor(val[blk] + 0.5));
        while (next <= q && next < kMlpC) b[next++] = sorted[i];
        i += len[blk];
    }
}

uint32_t split_merged_buckets(const std::vector<uint64_t>& keys, std::vector<uint64_t>& b)
{
    const size_t n = keys.size();
    std::vector<size_t> pos(kMlpC + 1);
    for (uint32_t j = 0; j < kMlpC; ++j) pos[j] = std::lower_bound(keys.begin(), keys.end(), b[j]) - keys.begin();
    pos[kMlpC] = n;
    const double mean = (double)n / kMlpC;
    uint32_t split = 0;
    for (uint32_t e = 0; e < kMlpC; ++e) {
        if (pos[e + 1] != pos[e]) continue;
        const size_t below = e > 0 ? pos[e] - pos[e - 1] : 0;
        const size_t above = e + 1 < kMlpC ? pos[e + 2] - pos[e + 1] : 0;
        if (2.0 * (double)std::max(below, above) < 3.0 * mean) continue;
        if (below >= above) {
            // bucket e takes the upper half of bucket e - 1
            const size_t m = pos[e - 1] + below / 2;
            b[e] = keys[m];
            pos[e] = m;
        } else {
            // bucket e takes the lower half of bucket e + 1
            const size_t m = pos[e + 1] + above / 2;
            b[e + 1] = keys[m];
            pos[e + 1] = m;
        }
        ++split;
    }
    return split;
}

namespace {

// The learned CDF: the piecewise-linear function through (b[j], j / 4096),
// j = 0..4096, with b[4096] read as 2^64; a run of equal boundaries (empty
// buckets) is a step.  In bucket units (0..4096).
long double learned_cdf(const std::vector<uint64_t>& b, long double x)
{
    const long double top = 18446744073709551616.0L;  // 2^64
    if (x >= top) return (long double)kMlpC;
    const uint64_t xi = (uint64_t)x;
    uint32_t j = (uint32_t)(std::upper_bound(b.begin(), b.begin() + kMlpC, xi) - b.begin()) - 1u;
    const long double lo = (long double)b[j];
    const long double hi = j + 1 < kMlpC ? (long double)b[j + 1] : top;
    return (long double)j + (hi > lo ? (x - lo) / (hi - lo) : 0.0L);
}

// Its inverse: the smallest key whose CDF reaches y (0 <= y < 4096).
long double learned_quantile(const std::vector<uint64_t>& b, long double y)
{
    const long double top = 18446744073709551616.0L;
    uint32_t j = (uint32_t)std::floor(y);
    if (j >= kMlpC) return top;
    const long double lo = (long double)b[j];
    const long double hi = j + 1 < kMlpC ? (long double)b[j + 1] : top;
    return lo + (y - (long double)j) * (hi - lo);
}

}  // namespace

void learned_routing(const std::vector<uint64_t>& b, uint32_t nf_target, uint32_t max_fine, LearnedRouting& out)
{
    max_fine = std::max<uint32_t>(1u, std::min<uint32_t>(max_fine, kMaxFine));
    const long double span = 72057594037927936.0L;  // 2^56 keys per root
    out.fine_base.assign(257, 0);
    out.bounds.clear();
    out.lut.assign(256 * (size_t)kLutCells, 0);
    std::vector<uint32_t> nsub(256);
    std::vector<long double> f_lo(257);
    for (uint32_t r = 0; r <= 256; ++r) f_lo[r] = learned_cdf(b, (long double)r * span);
    for (uint32_t r = 0; r < 256; ++r) {
        const long double mass = (f_lo[r + 1] - f_lo[r]) / (long double)kMlpC;
        const long double want = std::floor(mass * (long double)nf_target + 0.5L);
        nsub[r] = (uint32_t)std::min<long double>((long double)max_fine, std::max<long double>(1.0L, want));
        out.fine_base[r + 1] = out.fine_base[r] + nsub[r];
    }
    out.bounds.reserve(out.fine_base[256] - 256);
    for (uint32_t r = 0; r < 256; ++r) {
        const uint64_t lo = (uint64_t)r << 56;
        const uint64_t last = r < 255 ? lo + ((1ULL << 56) - 1) : ~0ULL;
        const size_t first = out.bounds.size();
        uint64_t prev = lo;
        for (uint32_t s = 1; s < nsub[r]; ++s) {
            // key at the CDF level s / nsub of the way through the root's mass
            const long double y = f_lo[r] + (f_lo[r + 1] - f_lo[r]) * (long double)s / (long double)nsub[r];
            long double q = std::ceil(learned_quantile(b, y));
            uint64_t v = q >= (long double)last ? last : (uint64_t)q;
            v = std::max<uint64_t>(v, std::max<uint64_t>(prev, lo + 1));  // inside the root, nondecreasing
            out.bounds.push_back(v);
            prev = v;
        }
        const uint64_t* fb = out.bounds.data() + first;
        const uint32_t nb = nsub[r] - 1;
        uint16_t* t = out.lut.data() + (size_t)r * kLutCells;
        for (uint32_t c = 0; c < kLutCells; ++c)
            t[c] = (uint16_t)(std::upper_bound(fb, fb + nb, lo + ((uint64_t)c << kLutShift)) - fb);
    }
}

}  // namespace kmc
