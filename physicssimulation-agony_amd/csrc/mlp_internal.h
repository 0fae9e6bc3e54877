// Internal declarations of the learned warm start (MLP 24 -> 128 -> 128 -> 2).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

struct agx_mlp;

namespace agx {

constexpr uint32_t MLP_IN = 24;      // features
constexpr uint32_t MLP_INP = 32;     // features padded to two MFMA k-blocks
constexpr uint32_t MLP_H = 128;      // hidden width
constexpr uint32_t MLP_OUT = 2;      // normal and tangent impulse (in units of the point's weight impulse)
constexpr uint32_t MLP_OUTP = 16;    // outputs padded to one MFMA tile
// Device parameter block (fp32 master, same order mirrored as bf16 for the weights):
// W1 [128][32], b1 [128], W2 [128][128], b2 [128], W3 [16][128], b3 [16].
constexpr uint32_t MLP_OFF_W1 = 0;
constexpr uint32_t MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_INP;
constexpr uint32_t MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr uint32_t MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr uint32_t MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr uint32_t MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUTP * MLP_H;
constexpr uint32_t MLP_PARAMS_DEV = MLP_OFF_B3 + MLP_OUTP;
// Host layout (agx.h): W1 [128][24], b1 [128], W2 [128][128], b2 [128], W3 [2][128], b3 [2].
constexpr uint32_t MLP_PARAMS_HOST = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;
constexpr uint32_t MLP_BATCH_MAX = 8192;
constexpr uint32_t MLP_FORWARD_BLOCKS = 256;   // one 4-wave block per CU (the weights stay in registers)

// Round to nearest even (NaN kept quiet); the oracle rounds the same way.
__host__ __device__ inline uint16_t bf16Bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16Value(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

struct MlpDevice {
    float* master = nullptr;      // MLP_PARAMS_DEV fp32
    uint16_t* mirror = nullptr;   // MLP_PARAMS_DEV bf16 (biases unused)
    const float* b1 = nullptr;
    const float* b2 = nullptr;
    const float* b3 = nullptr;
    const uint16_t* w1b = nullptr;
    const uint16_t* w2b = nullptr;
    const uint16_t* w3b = nullptr;
};

struct MlpGrad {
    float* all = nullptr;         // MLP_PARAMS_DEV
    float* w1 = nullptr;
    float* b1 = nullptr;
    float* w2 = nullptr;
    float* b2 = nullptr;
    float* w3 = nullptr;
    float* b3 = nullptr;
};

struct MlpAdam {
    float* m = nullptr;
    float* v = nullptr;
    uint8_t* trainable = nullptr;
    uint32_t step = 0;
};

struct MlpScratch {
    uint16_t* x = nullptr;        // [B][32]
    float* t = nullptr;           // [B][2]
    float* y = nullptr;           // [B][2]
    uint16_t* h1 = nullptr;       // [B][128]
    uint16_t* h2 = nullptr;
    uint16_t* dz1 = nullptr;
    uint16_t* dz2 = nullptr;
    float* dh1 = nullptr;         // [B][128]
    uint32_t* rows = nullptr;     // [B] sample rows of the batch
};

// Training samples gathered from solver runs (features bf16, targets fp32).
struct MlpSamples {
    uint16_t* x = nullptr;        // [cap][32]
    float* t = nullptr;           // [cap][2]
    uint32_t* count = nullptr;    // device counter (rows claimed; may exceed cap)
    uint64_t cap = 0;
};

hipError_t launchMlpMirror(hipStream_t st, MlpDevice& p);
// Forward pass over n rows (h1/h2 non-null: keep the hidden activations for training). With nDev
// the row count is read on the device (min(n, *nDev)) and n is the capacity.
hipError_t launchMlpForward(hipStream_t st, const MlpDevice& p, const uint16_t* x, uint32_t n, float* y, uint16_t* h1,
                            uint16_t* h2, const uint32_t* nDev = nullptr);
hipError_t launchMlpBackward(hipStream_t st, const MlpDevice& p, MlpGrad& g, const uint16_t* x, const float* y,
                             const float* t, uint32_t n, const uint16_t* h1, const uint16_t* h2, MlpScratch& s,
                             float* loss);
hipError_t launchMlpAdam(hipStream_t st, MlpDevice& p, const MlpGrad& g, MlpAdam& a, float lr);
// Random batch of n sample rows (hash of seed and index) gathered into s.x / s.t.
hipError_t launchMlpGatherBatch(hipStream_t st, const MlpSamples& smp, uint32_t n, uint32_t seed, MlpScratch& s);

const MlpDevice* mlpDevice(const agx_mlp* m);
const MlpSamples* mlpSamples(const agx_mlp* m);
hipStream_t mlpStream(const agx_mlp* m);

// Learned warm start inside a world's step (between constraint prep and the warm start, and
// between the sweeps and the write-back of the accumulators).
struct DeviceWorld;
hipError_t launchLearnBeforeWarm(hipStream_t st, DeviceWorld& w, const MlpDevice* mlp, uint32_t predict, uint32_t gated,
                                 uint32_t collect, uint32_t newOnly, float gx, float gy, float dt);
hipError_t launchLearnCollect(hipStream_t st, DeviceWorld& w, const MlpSamples& smp, uint32_t newOnly, float gx,
                              float gy, float dt);
hipError_t launchGateUpdate(hipStream_t st, DeviceWorld& w, uint32_t toTarget);

}  // namespace agx
