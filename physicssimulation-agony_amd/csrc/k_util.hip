// Fill and scan kernels shared by the phases.
#include "agx_internal.h"

namespace agx {

namespace {

__global__ void kFillU32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Block-wide exclusive scan of one value per thread (blockDim.x == 256).
__device__ inline uint32_t blockExclusiveScan256(uint32_t v, uint32_t& total) {
    __shared__ uint32_t waveSums[4];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) waveSums[wid] = x;
    __syncthreads();
    uint32_t waveBase = 0;
    for (uint32_t k = 0; k < wid; k++) waveBase += waveSums[k];
    total = waveSums[0] + waveSums[1] + waveSums[2] + waveSums[3];
    __syncthreads();
    return waveBase + x - v;
}

constexpr uint32_t SCAN_ITEMS = 4;                  // per thread per tile
constexpr uint32_t SCAN_TILE = 256 * SCAN_ITEMS;

// Pass 1: sum of each block's contiguous segment.
__global__ __launch_bounds__(256) void kScanReduce(const uint32_t* in, const uint32_t* countPtr, uint32_t countImm,
                                                   uint32_t* blockSums) {
    const uint32_t n = countPtr ? *countPtr : countImm;
    const uint32_t nb = gridDim.x;
    const uint32_t seg = (n + nb - 1) / nb;
    const uint32_t lo = blockIdx.x * seg;
    const uint32_t hi = min(n, lo + seg);
    uint32_t s = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) s += in[i];
    for (uint32_t off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ uint32_t ws[4];
    if ((threadIdx.x & 63u) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Pass 2: exclusive scan of the block sums (one block).
__global__ __launch_bounds__(256) void kScanTop(uint32_t* blockSums, uint32_t nb, uint32_t* totalPtr) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? blockSums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan256(v, tot);
        if (i < nb) blockSums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && totalPtr) *totalPtr = carry;
}

// Pass 3: rescan each segment tile by tile and add the block offset.
__global__ __launch_bounds__(256) void kScanDown(const uint32_t* in, uint32_t* out, const uint32_t* countPtr,
                                                 uint32_t countImm, const uint32_t* blockSums) {
    const uint32_t n = countPtr ? *countPtr : countImm;
    const uint32_t nb = gridDim.x;
    const uint32_t seg = (n + nb - 1) / nb;
    const uint32_t lo = blockIdx.x * seg;
    const uint32_t hi = min(n, lo + seg);
    uint32_t carry = blockSums[blockIdx.x];
    for (uint32_t base = lo; base < hi; base += SCAN_TILE) {
        uint32_t v[SCAN_ITEMS];
        uint32_t local = 0;
        const uint32_t first = base + threadIdx.x * SCAN_ITEMS;
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            const uint32_t i = first + k;
            v[k] = i < hi ? in[i] : 0u;
            local += v[k];
        }
        uint32_t tot;
        uint32_t ex = blockExclusiveScan256(local, tot);
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            const uint32_t i = first + k;
            if (i < hi) out[i] = carry + ex;
            ex += v[k];
        }
        carry += tot;
    }
}

}  // namespace

hipError_t launchFillU32(hipStream_t st, uint32_t* p, uint32_t v, size_t n) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(kFillU32, dim3(blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launchFillBytes(hipStream_t st, void* p, size_t bytes) {
    return launchFillU32(st, static_cast<uint32_t*>(p), 0u, bytes / 4);
}

hipError_t launchScan(hipStream_t st, const uint32_t* in, uint32_t* out, const uint32_t* countPtr,
                      uint32_t* totalPtr, uint64_t countHost, uint32_t* blockScratch, uint32_t blockCap) {
    uint32_t nb = (uint32_t)std::min<uint64_t>((countHost + 4 * SCAN_TILE - 1) / (4 * SCAN_TILE), blockCap);
    if (nb == 0) nb = 1;
    const uint32_t imm = (uint32_t)countHost;
    hipLaunchKernelGGL(kScanReduce, dim3(nb), dim3(256), 0, st, in, countPtr, imm, blockScratch);
    hipLaunchKernelGGL(kScanTop, dim3(1), dim3(256), 0, st, blockScratch, nb, totalPtr);
    hipLaunchKernelGGL(kScanDown, dim3(nb), dim3(256), 0, st, in, out, countPtr, imm, blockScratch);
    return hipGetLastError();
}

}  // namespace agx
