// Learned warm start: a 24 -> 128 -> 128 -> 2 MLP on MFMA (bf16 operands, fp32 accumulation).
//
// Parameters are kept in fp32 (master copy, Adam moments) and mirrored as bf16 for the MFMA
// kernels. Activations between layers are rounded to bf16; every sum is accumulated in fp32.
// Device layouts (row-major, output-major weights):
//   W1 [128][32] (inputs 24..31 zero), b1[128], W2 [128][128], b2[128], W3 [16][128] (rows 2..15
//   zero), b3[16] (2..15 zero).
// The 16x16x16 bf16 MFMA operand layout used throughout (lane l, i = l % 16, q = l / 16):
//   A fragment: A[i][4q .. 4q+3];  B fragment: B[4q .. 4q+3][i];  C: C[4q + r][i], r = 0..3.
#include "agx_internal.h"
#include "mlp_internal.h"

namespace agx {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ inline float4_t mfma16(short4_t a, short4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 16x16x32: lane (i, q) holds 8 k-slots of A row i and of B column i. Which k a slot stands for
// depends only on (q, slot), so any assignment shared by A and B gives the same product; here slots
// 0..3 are k = 4q .. 4q+3 of the first 16-block and slots 4..7 the same k of the second, which is
// how two 16x16 output tiles of the previous layer already sit in the registers.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
__device__ inline float4_t mfma32(short4_t aLo, short4_t aHi, short4_t bLo, short4_t bHi, float4_t c) {
    const bf16x8_t a = __builtin_bit_cast(bf16x8_t, __builtin_shufflevector(aLo, aHi, 0, 1, 2, 3, 4, 5, 6, 7));
    const bf16x8_t b = __builtin_bit_cast(bf16x8_t, __builtin_shufflevector(bLo, bHi, 0, 1, 2, 3, 4, 5, 6, 7));
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline short4_t frag4(const uint16_t* p) {   // 4 consecutive bf16
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    short4_t s;
    s[0] = (short)(v.x & 0xFFFFu); s[1] = (short)(v.x >> 16);
    s[2] = (short)(v.y & 0xFFFFu); s[3] = (short)(v.y >> 16);
    return s;
}

// ---------------------------------------------------------------- forward
// Transposed formulation: each layer computes H_next^T = W H^T with the weights as the MFMA's A
// operand and the activations as its B operand. A 16x16 output tile then sits in the registers as
// C[unit 4q + r][sample i], which is exactly the B fragment (k = 4q .. 4q+3, column i) the next
// layer needs for that block of 16 input units: activations never leave the registers. Each wave
// runs 16 samples per tile; the block's weights sit in LDS with rows padded so that the 8-byte
// fragment reads of one half-wave fall into distinct banks.
constexpr uint32_t W1_STRIDE = MLP_INP + 8;   // 40 bf16 = 20 dwords: rows 20 dwords apart
constexpr uint32_t WH_STRIDE = MLP_H + 8;     // 136 bf16 = 68 dwords: rows 4 banks apart
constexpr uint32_t FORWARD_LDS_BYTES =
    2 * (MLP_H * W1_STRIDE + MLP_H * WH_STRIDE + MLP_OUTP * WH_STRIDE) + 4 * (2 * MLP_H + MLP_OUTP);
static_assert(FORWARD_LDS_BYTES * 3 <= 160 * 1024, "three forward blocks share a CU's LDS");

__device__ inline short4_t pack4(float a, float b, float c, float d) {
    short4_t s;
    s[0] = (short)bf16Bits(a); s[1] = (short)bf16Bits(b); s[2] = (short)bf16Bits(c); s[3] = (short)bf16Bits(d);
    return s;
}

template <bool TRAIN>
__global__ __launch_bounds__(256) void kMlpForward(const uint16_t* __restrict__ x, uint32_t n, const uint32_t* nDev,
                                                   MlpDevice p, float* __restrict__ y, uint16_t* __restrict__ h1Out,
                                                   uint16_t* __restrict__ h2Out) {
    __shared__ __align__(16) uint16_t w1[MLP_H * W1_STRIDE];
    __shared__ __align__(16) uint16_t w2[MLP_H * WH_STRIDE];
    __shared__ __align__(16) uint16_t w3[MLP_OUTP * WH_STRIDE];
    __shared__ float bias[2 * MLP_H + MLP_OUTP];
    if (nDev) n = min(n, *nDev);   // rows counted on the device; n is then the capacity
    if (n == 0) return;
    // stage the weights: 4 bf16 (8 bytes) per copy
    for (uint32_t k = threadIdx.x; k < MLP_H * MLP_INP / 4; k += blockDim.x) {
        const uint32_t r = k / (MLP_INP / 4), c = 4 * (k % (MLP_INP / 4));
        *reinterpret_cast<uint2*>(&w1[r * W1_STRIDE + c]) = *reinterpret_cast<const uint2*>(&p.w1b[r * MLP_INP + c]);
    }
    for (uint32_t k = threadIdx.x; k < MLP_H * MLP_H / 4; k += blockDim.x) {
        const uint32_t r = k / (MLP_H / 4), c = 4 * (k % (MLP_H / 4));
        *reinterpret_cast<uint2*>(&w2[r * WH_STRIDE + c]) = *reinterpret_cast<const uint2*>(&p.w2b[r * MLP_H + c]);
    }
    for (uint32_t k = threadIdx.x; k < MLP_OUTP * MLP_H / 4; k += blockDim.x) {
        const uint32_t r = k / (MLP_H / 4), c = 4 * (k % (MLP_H / 4));
        *reinterpret_cast<uint2*>(&w3[r * WH_STRIDE + c]) = *reinterpret_cast<const uint2*>(&p.w3b[r * MLP_H + c]);
    }
    for (uint32_t k = threadIdx.x; k < MLP_H; k += blockDim.x) { bias[k] = p.b1[k]; bias[MLP_H + k] = p.b2[k]; }
    for (uint32_t k = threadIdx.x; k < MLP_OUTP; k += blockDim.x) bias[2 * MLP_H + k] = p.b3[k];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t i = lane & 15u, q = lane >> 4;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t stride = gridDim.x * 4;
    // the next tile's features are loaded while this one computes (one wave per SIMD hides little)
    auto loadX = [&](uint32_t tile, short4_t* xf) {
        const uint32_t r = tile * 16 + i;
        for (uint32_t kb = 0; kb < 2; kb++) {
            if (tile < tiles && r < n) xf[kb] = frag4(x + (size_t)r * MLP_INP + kb * 16 + 4 * q);
            else xf[kb] = short4_t{ 0, 0, 0, 0 };
        }
    };
    short4_t xn[2];
    loadX(blockIdx.x * 4 + wid, xn);
    for (uint32_t t = blockIdx.x * 4 + wid; t < tiles; t += stride) {
        const uint32_t row = t * 16 + i;   // this lane's sample (B-operand column)
        const bool valid = row < n;
        short4_t xb[2] = { xn[0], xn[1] };
        loadX(t + stride, xn);
        // layer 1: h1^T [128 x 16] = W1 [128 x 32] x^T, one 16x16x32 product per 16 units
        short4_t h1[MLP_H / 16];
        for (uint32_t ob = 0; ob < MLP_H / 16; ob++) {
            const uint16_t* wr = w1 + (ob * 16 + i) * W1_STRIDE + 4 * q;
            float4_t c = { 0.f, 0.f, 0.f, 0.f };
            c = mfma32(frag4(wr), frag4(wr + 16), xb[0], xb[1], c);
            const float* b = bias + ob * 16 + 4 * q;
            h1[ob] = pack4(fmaxf(c[0] + b[0], 0.0f), fmaxf(c[1] + b[1], 0.0f), fmaxf(c[2] + b[2], 0.0f),
                           fmaxf(c[3] + b[3], 0.0f));
        }
        // layer 2: h2^T = W2 h1^T; k-block kb spans the unit blocks 2 kb and 2 kb + 1
        short4_t h2[MLP_H / 16];
        for (uint32_t ob = 0; ob < MLP_H / 16; ob++) {
            const uint16_t* wr = w2 + (ob * 16 + i) * WH_STRIDE + 4 * q;
            float4_t c = { 0.f, 0.f, 0.f, 0.f };
            for (uint32_t kb = 0; kb < MLP_H / 32; kb++)
                c = mfma32(frag4(wr + 32 * kb), frag4(wr + 32 * kb + 16), h1[2 * kb], h1[2 * kb + 1], c);
            const float* b = bias + MLP_H + ob * 16 + 4 * q;
            h2[ob] = pack4(fmaxf(c[0] + b[0], 0.0f), fmaxf(c[1] + b[1], 0.0f), fmaxf(c[2] + b[2], 0.0f),
                           fmaxf(c[3] + b[3], 0.0f));
        }
        // layer 3: y^T [16 x 16] = W3 (rows 2..15 zero) h2^T; lane (i, q = 0) holds outputs 0, 1 of sample i
        float4_t c = { 0.f, 0.f, 0.f, 0.f };
        {
            const uint16_t* wr = w3 + i * WH_STRIDE + 4 * q;
            for (uint32_t kb = 0; kb < MLP_H / 32; kb++)
                c = mfma32(frag4(wr + 32 * kb), frag4(wr + 32 * kb + 16), h2[2 * kb], h2[2 * kb + 1], c);
        }
        if (q == 0 && valid)
            *reinterpret_cast<float2*>(y + (size_t)row * MLP_OUT) =
                make_float2(c[0] + bias[2 * MLP_H], c[1] + bias[2 * MLP_H + 1]);
        if (TRAIN && valid) {
            for (uint32_t ob = 0; ob < MLP_H / 16; ob++) {
                *reinterpret_cast<short4_t*>(h1Out + (size_t)row * MLP_H + ob * 16 + 4 * q) = h1[ob];
                *reinterpret_cast<short4_t*>(h2Out + (size_t)row * MLP_H + ob * 16 + 4 * q) = h2[ob];
            }
        }
    }
}

// ---------------------------------------------------------------- backward
// Mean squared error over the batch: dY = 2 (Y - T) / B for rows with weight 1. Then
// dW3 = dY^T H2, db3 = sum dY, dZ2 = (dY W3) * [H2 > 0] (bf16), db2 = sum dZ2 (fp32).
// One thread per (sample, hidden unit); fp32 partial sums per block, added atomically.
__global__ __launch_bounds__(256) void kMlpHead(const float* __restrict__ y, const float* __restrict__ t, uint32_t n,
                                                const uint16_t* __restrict__ h2, MlpDevice p, MlpGrad g,
                                                uint16_t* __restrict__ dz2, float* __restrict__ loss) {
    __shared__ float sW3[MLP_OUT][MLP_H];
    __shared__ float sB3[MLP_OUT];
    __shared__ float sB2[MLP_H];
    __shared__ float sLoss;
    for (uint32_t k = threadIdx.x; k < MLP_OUT * MLP_H; k += blockDim.x) sW3[k / MLP_H][k % MLP_H] = 0.0f;
    for (uint32_t k = threadIdx.x; k < MLP_H; k += blockDim.x) sB2[k] = 0.0f;
    if (threadIdx.x < MLP_OUT) sB3[threadIdx.x] = 0.0f;
    if (threadIdx.x == 0) sLoss = 0.0f;
    __syncthreads();
    const float inv = 2.0f / (float)n;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n * MLP_H; e += gridDim.x * blockDim.x) {
        const uint32_t s = e / MLP_H, j = e % MLP_H;
        float dy[MLP_OUT];
        for (uint32_t o = 0; o < MLP_OUT; o++) {
            const float diff = y[(size_t)s * MLP_OUT + o] - t[(size_t)s * MLP_OUT + o];
            dy[o] = inv * diff;
            if (j == 0) atomicAdd(&sLoss, diff * diff / (float)n);
            if (j == 0) atomicAdd(&sB3[o], dy[o]);
        }
        const float hv = bf16Value(h2[(size_t)s * MLP_H + j]);
        float dh = 0.0f;
        for (uint32_t o = 0; o < MLP_OUT; o++) {
            atomicAdd(&sW3[o][j], dy[o] * hv);
            dh += dy[o] * bf16Value(p.w3b[o * MLP_H + j]);
        }
        const float dz = hv > 0.0f ? dh : 0.0f;
        dz2[(size_t)s * MLP_H + j] = bf16Bits(dz);
        atomicAdd(&sB2[j], dz);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < MLP_OUT * MLP_H; k += blockDim.x) {
        const float v = sW3[k / MLP_H][k % MLP_H];
        if (v != 0.0f) atomicAdd(&g.w3[k], v);
    }
    for (uint32_t k = threadIdx.x; k < MLP_H; k += blockDim.x) if (sB2[k] != 0.0f) atomicAdd(&g.b2[k], sB2[k]);
    if (threadIdx.x < MLP_OUT) atomicAdd(&g.b3[threadIdx.x], sB3[threadIdx.x]);
    if (threadIdx.x == 0) atomicAdd(loss, sLoss);
}

// C[M x N] += sum_k A(m, k) B(k, n) over this block's slice of k; A(m, k) = a[m * am + k * ak],
// B(k, n) = b[k * bk + n * bn] (bf16). One wave per 16 x 16 tile of C; the k range is split over
// gridDim.y slices (multiples of 16) and the partial tiles are added atomically in fp32.
__global__ __launch_bounds__(64) void kGemmAccum(const uint16_t* __restrict__ a, uint32_t am, uint32_t ak,
                                                 const uint16_t* __restrict__ b, uint32_t bk, uint32_t bn, uint32_t M,
                                                 uint32_t N, uint32_t K, uint32_t kSlice, float* __restrict__ c) {
    const uint32_t lane = threadIdx.x;
    const uint32_t i = lane & 15u, q = lane >> 4;
    const uint32_t tilesN = N / 16;
    const uint32_t tm = blockIdx.x / tilesN, tn = blockIdx.x % tilesN;
    const uint32_t k0 = blockIdx.y * kSlice, k1 = min(K, k0 + kSlice);
    float4_t acc = { 0.f, 0.f, 0.f, 0.f };
    for (uint32_t kb = k0; kb < k1; kb += 16) {
        short4_t fa, fb;
        for (uint32_t r = 0; r < 4; r++) {
            const uint32_t k = kb + 4 * q + r;
            fa[r] = (short)a[(size_t)(tm * 16 + i) * am + (size_t)k * ak];
            fb[r] = (short)b[(size_t)k * bk + (size_t)(tn * 16 + i) * bn];
        }
        acc = mfma16(fa, fb, acc);
    }
    for (uint32_t r = 0; r < 4; r++) atomicAdd(&c[(size_t)(tm * 16 + 4 * q + r) * N + tn * 16 + i], acc[r]);
}

// dZ1 = dH1 * [H1 > 0] (bf16) and db1 += sum dZ1 (fp32).
__global__ __launch_bounds__(256) void kMlpRelu1(const float* __restrict__ dh1, const uint16_t* __restrict__ h1, uint32_t n,
                                                 uint16_t* __restrict__ dz1, float* __restrict__ gb1) {
    __shared__ float sB1[MLP_H];
    for (uint32_t k = threadIdx.x; k < MLP_H; k += blockDim.x) sB1[k] = 0.0f;
    __syncthreads();
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n * MLP_H; e += gridDim.x * blockDim.x) {
        const float dz = bf16Value(h1[e]) > 0.0f ? dh1[e] : 0.0f;
        dz1[e] = bf16Bits(dz);
        atomicAdd(&sB1[e % MLP_H], dz);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < MLP_H; k += blockDim.x) if (sB1[k] != 0.0f) atomicAdd(&gb1[k], sB1[k]);
}

// Adam on the fp32 master parameters, then the bf16 mirrors.
__global__ __launch_bounds__(256) void kMlpAdam(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m,
                                                float* __restrict__ v, uint32_t count, float lr, float b1, float b2,
                                                float eps, float c1, float c2, const uint8_t* __restrict__ trainable) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        if (!trainable[k]) continue;
        const float gk = grad[k];
        const float mk = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    }
}

__global__ __launch_bounds__(256) void kMlpMirror(const float* __restrict__ w, uint16_t* __restrict__ wb, uint32_t count) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) wb[k] = bf16Bits(w[k]);
}

inline uint32_t gridFor(uint64_t n, uint32_t per = 256, uint32_t cap = 4096) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + per - 1) / per, cap));
}

}  // namespace

hipError_t launchMlpMirror(hipStream_t st, MlpDevice& p) {
    hipLaunchKernelGGL(kMlpMirror, dim3(gridFor(MLP_PARAMS_DEV)), dim3(256), 0, st, p.master, p.mirror, (uint32_t)MLP_PARAMS_DEV);
    return hipGetLastError();
}

hipError_t launchMlpForward(hipStream_t st, const MlpDevice& p, const uint16_t* x, uint32_t n, float* y, uint16_t* h1,
                            uint16_t* h2, const uint32_t* nDev) {
    if (n == 0) return hipSuccess;
    // each block stages the weights once (about 49 KB of LDS) and then loops over 16-row tiles
    const uint32_t blocks = gridFor((n + 15) / 16, 4, MLP_FORWARD_BLOCKS);
    if (h1) hipLaunchKernelGGL(kMlpForward<true>, dim3(blocks), dim3(256), 0, st, x, n, nDev, p, y, h1, h2);
    else hipLaunchKernelGGL(kMlpForward<false>, dim3(blocks), dim3(256), 0, st, x, n, nDev, p, y, h1, h2);
    return hipGetLastError();
}

// Gradients of the batch loss into g (cleared here). n must be a multiple of 16.
hipError_t launchMlpBackward(hipStream_t st, const MlpDevice& p, MlpGrad& g, const uint16_t* x, const float* y,
                             const float* t, uint32_t n, const uint16_t* h1, const uint16_t* h2, MlpScratch& s,
                             float* loss) {
    hipError_t e = launchFillU32(st, reinterpret_cast<uint32_t*>(g.all), 0u, MLP_PARAMS_DEV);
    if (e != hipSuccess) return e;
    e = launchFillU32(st, reinterpret_cast<uint32_t*>(s.dh1), 0u, (size_t)n * MLP_H);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kMlpHead, dim3(gridFor((uint64_t)n * MLP_H, 256, 1024)), dim3(256), 0, st, y, t, n, h2, p, g, s.dz2,
                       loss);
    const uint32_t slice = 256;
    const uint32_t slices = (n + slice - 1) / slice;
    // dW2[i][j] += sum_s dZ2[s][i] H1[s][j]
    hipLaunchKernelGGL(kGemmAccum, dim3((MLP_H / 16) * (MLP_H / 16), slices), dim3(64), 0, st, s.dz2, 1u, (uint32_t)MLP_H, h1,
                       (uint32_t)MLP_H, 1u, (uint32_t)MLP_H, (uint32_t)MLP_H, n, slice, g.w2);
    // dH1[s][j] += sum_i dZ2[s][i] W2[i][j]
    hipLaunchKernelGGL(kGemmAccum, dim3((n / 16) * (MLP_H / 16), 1), dim3(64), 0, st, s.dz2, (uint32_t)MLP_H, 1u, p.w2b,
                       (uint32_t)MLP_H, 1u, n, (uint32_t)MLP_H, (uint32_t)MLP_H, (uint32_t)MLP_H, s.dh1);
    hipLaunchKernelGGL(kMlpRelu1, dim3(gridFor((uint64_t)n * MLP_H, 256, 1024)), dim3(256), 0, st, s.dh1, h1, n, s.dz1, g.b1);
    // dW1[i][k] += sum_s dZ1[s][i] X[s][k]
    hipLaunchKernelGGL(kGemmAccum, dim3((MLP_H / 16) * (MLP_INP / 16), slices), dim3(64), 0, st, s.dz1, 1u, (uint32_t)MLP_H, x,
                       (uint32_t)MLP_INP, 1u, (uint32_t)MLP_H, (uint32_t)MLP_INP, n, slice, g.w1);
    return hipGetLastError();
}

hipError_t launchMlpAdam(hipStream_t st, MlpDevice& p, const MlpGrad& g, MlpAdam& a, float lr) {
    a.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)a.step), c2 = 1.0f - std::pow(b2, (float)a.step);
    hipLaunchKernelGGL(kMlpAdam, dim3(gridFor(MLP_PARAMS_DEV)), dim3(256), 0, st, p.master, g.all, a.m, a.v,
                       (uint32_t)MLP_PARAMS_DEV, lr, b1, b2, 1e-8f, c1, c2, a.trainable);
    return launchMlpMirror(st, p);
}

}  // namespace agx

namespace agx {
namespace {

__device__ inline uint32_t hashMix(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void kMlpGather(const uint16_t* __restrict__ sx, const float* __restrict__ st,
                                                  const uint32_t* __restrict__ count, uint64_t cap, uint32_t n,
                                                  uint32_t seed, uint16_t* __restrict__ x, float* __restrict__ t) {
    const uint64_t avail = min((uint64_t)*count, cap);
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n * MLP_INP; e += gridDim.x * blockDim.x) {
        const uint32_t b = e / MLP_INP, k = e % MLP_INP;
        const uint64_t row = avail ? (uint64_t)hashMix(hashMix(seed) ^ (b * 0x9E3779B9u)) % avail : 0u;
        x[e] = avail ? sx[row * MLP_INP + k] : (uint16_t)0;
        if (k < MLP_OUT) t[b * MLP_OUT + k] = avail ? st[row * MLP_OUT + k] : 0.0f;
    }
}

}  // namespace

hipError_t launchMlpGatherBatch(hipStream_t st, const MlpSamples& smp, uint32_t n, uint32_t seed, MlpScratch& s) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)n * MLP_INP + 255) / 256, 4096);
    hipLaunchKernelGGL(kMlpGather, dim3(std::max(1u, blocks)), dim3(256), 0, st, smp.x, smp.t, smp.count, smp.cap, n, seed,
                       s.x, s.t);
    return hipGetLastError();
}

}  // namespace agx
