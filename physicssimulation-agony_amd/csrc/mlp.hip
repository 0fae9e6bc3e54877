// Host side of the learned warm start: the agx_mlp object and its C ABI (include/agx.h).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "agx_internal.h"
#include "mlp_internal.h"

struct agx_mlp {
    hipStream_t stream = nullptr;
    agx::MlpDevice p;
    agx::MlpGrad g;
    agx::MlpAdam adam;
    agx::MlpScratch s;
    agx::MlpSamples smp;
    float* loss = nullptr;
    std::vector<void*> allocations;
};

namespace {

thread_local std::string g_mlpError;

int mfail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_mlpError = buf;
    agx::setLastError(buf);
    return code;
}

#define MLP_CHECK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return mfail(AGX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); } while (0)

template <typename T>
hipError_t mAlloc(agx_mlp* m, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    m->allocations.push_back(q);
    *p = static_cast<T*>(q);
    return hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
}

// Host layout <-> device layout of the parameters.
void hostToDevice(const float* h, std::vector<float>& d) {
    d.assign(agx::MLP_PARAMS_DEV, 0.0f);
    size_t k = 0;
    for (uint32_t o = 0; o < agx::MLP_H; o++)
        for (uint32_t i = 0; i < agx::MLP_IN; i++) d[agx::MLP_OFF_W1 + o * agx::MLP_INP + i] = h[k++];
    for (uint32_t o = 0; o < agx::MLP_H; o++) d[agx::MLP_OFF_B1 + o] = h[k++];
    for (uint32_t o = 0; o < agx::MLP_H * agx::MLP_H; o++) d[agx::MLP_OFF_W2 + o] = h[k++];
    for (uint32_t o = 0; o < agx::MLP_H; o++) d[agx::MLP_OFF_B2 + o] = h[k++];
    for (uint32_t o = 0; o < agx::MLP_OUT; o++)
        for (uint32_t i = 0; i < agx::MLP_H; i++) d[agx::MLP_OFF_W3 + o * agx::MLP_H + i] = h[k++];
    for (uint32_t o = 0; o < agx::MLP_OUT; o++) d[agx::MLP_OFF_B3 + o] = h[k++];
}

void deviceToHost(const std::vector<float>& d, float* h) {
    size_t k = 0;
    for (uint32_t o = 0; o < agx::MLP_H; o++)
        for (uint32_t i = 0; i < agx::MLP_IN; i++) h[k++] = d[agx::MLP_OFF_W1 + o * agx::MLP_INP + i];
    for (uint32_t o = 0; o < agx::MLP_H; o++) h[k++] = d[agx::MLP_OFF_B1 + o];
    for (uint32_t o = 0; o < agx::MLP_H * agx::MLP_H; o++) h[k++] = d[agx::MLP_OFF_W2 + o];
    for (uint32_t o = 0; o < agx::MLP_H; o++) h[k++] = d[agx::MLP_OFF_B2 + o];
    for (uint32_t o = 0; o < agx::MLP_OUT; o++)
        for (uint32_t i = 0; i < agx::MLP_H; i++) h[k++] = d[agx::MLP_OFF_W3 + o * agx::MLP_H + i];
    for (uint32_t o = 0; o < agx::MLP_OUT; o++) h[k++] = d[agx::MLP_OFF_B3 + o];
}

void padFeatures(const float* x, uint64_t n, std::vector<uint16_t>& out) {
    out.assign(n * agx::MLP_INP, 0);
    for (uint64_t r = 0; r < n; r++)
        for (uint32_t i = 0; i < agx::MLP_IN; i++) out[r * agx::MLP_INP + i] = agx::bf16Bits(x[r * agx::MLP_IN + i]);
}

int upload(agx_mlp* m, const std::vector<float>& dev) {
    MLP_CHECK(hipMemcpy(m->p.master, dev.data(), dev.size() * sizeof(float), hipMemcpyHostToDevice));
    MLP_CHECK(agx::launchMlpMirror(m->stream, m->p));
    MLP_CHECK(hipStreamSynchronize(m->stream));
    return AGX_OK;
}

}  // namespace

namespace agx {
const MlpDevice* mlpDevice(const agx_mlp* m) { return m ? &m->p : nullptr; }
const MlpSamples* mlpSamples(const agx_mlp* m) { return m ? &m->smp : nullptr; }
hipStream_t mlpStream(const agx_mlp* m) { return m ? m->stream : nullptr; }
}  // namespace agx

extern "C" {

int agx_mlp_create(uint32_t seed, uint64_t sample_capacity, agx_mlp** out) {
    if (!out) return mfail(AGX_ERR_INVALID_ARGUMENT, "out is required");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return mfail(AGX_ERR_NO_DEVICE, "no HIP device");
    agx_mlp* m = new (std::nothrow) agx_mlp();
    if (!m) return mfail(AGX_ERR_CAPACITY, "out of host memory");
    auto bail = [&](int rc) { agx_mlp_destroy(m); return rc; };
    if (agx::deviceStream(&m->stream) != hipSuccess) return bail(mfail(AGX_ERR_HIP, "stream"));
    const uint32_t P = agx::MLP_PARAMS_DEV, B = agx::MLP_BATCH_MAX;
#define MA(p, n) do { if (mAlloc(m, &(p), (n)) != hipSuccess) return bail(mfail(AGX_ERR_HIP, "alloc " #p)); } while (0)
    MA(m->p.master, P); MA(m->p.mirror, P); MA(m->g.all, P); MA(m->adam.m, P); MA(m->adam.v, P); MA(m->adam.trainable, P);
    MA(m->s.x, (size_t)B * agx::MLP_INP); MA(m->s.t, (size_t)B * agx::MLP_OUT); MA(m->s.y, (size_t)B * agx::MLP_OUT);
    MA(m->s.h1, (size_t)B * agx::MLP_H); MA(m->s.h2, (size_t)B * agx::MLP_H); MA(m->s.dz1, (size_t)B * agx::MLP_H);
    MA(m->s.dz2, (size_t)B * agx::MLP_H); MA(m->s.dh1, (size_t)B * agx::MLP_H); MA(m->s.rows, B);
    MA(m->loss, 1);
    m->smp.cap = std::max<uint64_t>(sample_capacity, 16);
    MA(m->smp.x, m->smp.cap * agx::MLP_INP); MA(m->smp.t, m->smp.cap * agx::MLP_OUT); MA(m->smp.count, 1);
#undef MA
    m->p.b1 = m->p.master + agx::MLP_OFF_B1; m->p.b2 = m->p.master + agx::MLP_OFF_B2; m->p.b3 = m->p.master + agx::MLP_OFF_B3;
    m->p.w1b = m->p.mirror + agx::MLP_OFF_W1; m->p.w2b = m->p.mirror + agx::MLP_OFF_W2; m->p.w3b = m->p.mirror + agx::MLP_OFF_W3;
    m->g.w1 = m->g.all + agx::MLP_OFF_W1; m->g.b1 = m->g.all + agx::MLP_OFF_B1; m->g.w2 = m->g.all + agx::MLP_OFF_W2;
    m->g.b2 = m->g.all + agx::MLP_OFF_B2; m->g.w3 = m->g.all + agx::MLP_OFF_W3; m->g.b3 = m->g.all + agx::MLP_OFF_B3;
    // trainable mask: the padding stays zero
    std::vector<uint8_t> mask(P, 0);
    for (uint32_t o = 0; o < agx::MLP_H; o++)
        for (uint32_t i = 0; i < agx::MLP_IN; i++) mask[agx::MLP_OFF_W1 + o * agx::MLP_INP + i] = 1;
    for (uint32_t k = 0; k < agx::MLP_H; k++) { mask[agx::MLP_OFF_B1 + k] = 1; mask[agx::MLP_OFF_B2 + k] = 1; }
    for (uint32_t k = 0; k < agx::MLP_H * agx::MLP_H; k++) mask[agx::MLP_OFF_W2 + k] = 1;
    for (uint32_t k = 0; k < agx::MLP_OUT * agx::MLP_H; k++) mask[agx::MLP_OFF_W3 + k] = 1;
    for (uint32_t k = 0; k < agx::MLP_OUT; k++) mask[agx::MLP_OFF_B3 + k] = 1;
    if (hipMemcpy(m->adam.trainable, mask.data(), P, hipMemcpyHostToDevice) != hipSuccess) return bail(mfail(AGX_ERR_HIP, "mask"));
    // He initialisation (normal, std sqrt(2 / fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> h(agx::MLP_PARAMS_HOST, 0.0f);
    size_t k = 0;
    auto fill = [&](size_t count, uint32_t fanIn) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)fanIn));
        for (size_t c = 0; c < count; c++) h[k++] = nd(rng);
    };
    fill((size_t)agx::MLP_H * agx::MLP_IN, agx::MLP_IN); k += agx::MLP_H;
    fill((size_t)agx::MLP_H * agx::MLP_H, agx::MLP_H); k += agx::MLP_H;
    fill((size_t)agx::MLP_OUT * agx::MLP_H, agx::MLP_H); k += agx::MLP_OUT;
    std::vector<float> dev;
    hostToDevice(h.data(), dev);
    if (int rc = upload(m, dev)) return bail(rc);
    *out = m;
    return AGX_OK;
}

int agx_mlp_destroy(agx_mlp* m) {
    if (!m) return AGX_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);   // the device's shared stream stays
    for (void* p : m->allocations) (void)hipFree(p);
    delete m;
    return AGX_OK;
}

int agx_mlp_get_params(agx_mlp* m, float* out, uint64_t cap) {
    if (!m || !out) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    MLP_CHECK(hipDeviceSynchronize());   // worlds may be running the network on their streams
    if (cap < agx::MLP_PARAMS_HOST) return mfail(AGX_ERR_CAPACITY, "need %u parameters", agx::MLP_PARAMS_HOST);
    MLP_CHECK(hipStreamSynchronize(m->stream));
    std::vector<float> dev(agx::MLP_PARAMS_DEV);
    MLP_CHECK(hipMemcpy(dev.data(), m->p.master, dev.size() * sizeof(float), hipMemcpyDeviceToHost));
    deviceToHost(dev, out);
    return AGX_OK;
}

int agx_mlp_set_params(agx_mlp* m, const float* in, uint64_t n) {
    if (!m || !in) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    MLP_CHECK(hipDeviceSynchronize());   // worlds may be running the network on their streams
    if (n != agx::MLP_PARAMS_HOST) return mfail(AGX_ERR_INVALID_ARGUMENT, "expected %u parameters", agx::MLP_PARAMS_HOST);
    std::vector<float> dev;
    hostToDevice(in, dev);
    m->adam.step = 0;
    MLP_CHECK(hipMemset(m->adam.m, 0, agx::MLP_PARAMS_DEV * sizeof(float)));
    MLP_CHECK(hipMemset(m->adam.v, 0, agx::MLP_PARAMS_DEV * sizeof(float)));
    return upload(m, dev);
}

int agx_mlp_forward(agx_mlp* m, const float* x, uint64_t n, float* y) {
    if (!m || (n && (!x || !y))) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<uint16_t> xb;
    for (uint64_t base = 0; base < n; base += agx::MLP_BATCH_MAX) {
        const uint32_t c = (uint32_t)std::min<uint64_t>(agx::MLP_BATCH_MAX, n - base);
        padFeatures(x + base * agx::MLP_IN, c, xb);
        MLP_CHECK(hipMemcpyAsync(m->s.x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, m->stream));
        MLP_CHECK(agx::launchMlpForward(m->stream, m->p, m->s.x, c, m->s.y, nullptr, nullptr, nullptr));
        MLP_CHECK(hipMemcpyAsync(y + base * agx::MLP_OUT, m->s.y, (size_t)c * agx::MLP_OUT * sizeof(float),
                                 hipMemcpyDeviceToHost, m->stream));
        MLP_CHECK(hipStreamSynchronize(m->stream));
    }
    return AGX_OK;
}

int agx_mlp_gradients(agx_mlp* m, const float* x, const float* t, uint64_t n, float* grad, float* loss) {
    if (!m || !x || !t || !grad) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0 || n % 16 != 0 || n > agx::MLP_BATCH_MAX)
        return mfail(AGX_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 16 up to %u", agx::MLP_BATCH_MAX);
    std::vector<uint16_t> xb;
    padFeatures(x, n, xb);
    MLP_CHECK(hipMemcpyAsync(m->s.x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, m->stream));
    MLP_CHECK(hipMemcpyAsync(m->s.t, t, n * agx::MLP_OUT * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_CHECK(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
    MLP_CHECK(agx::launchMlpForward(m->stream, m->p, m->s.x, (uint32_t)n, m->s.y, m->s.h1, m->s.h2, nullptr));
    MLP_CHECK(agx::launchMlpBackward(m->stream, m->p, m->g, m->s.x, m->s.y, m->s.t, (uint32_t)n, m->s.h1, m->s.h2, m->s,
                                     m->loss));
    std::vector<float> dev(agx::MLP_PARAMS_DEV);
    MLP_CHECK(hipMemcpyAsync(dev.data(), m->g.all, dev.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    float l = 0.0f;
    MLP_CHECK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_CHECK(hipStreamSynchronize(m->stream));
    deviceToHost(dev, grad);
    if (loss) *loss = l;
    return AGX_OK;
}

int agx_mlp_add_samples(agx_mlp* m, const float* x, const float* t, uint64_t n) {
    if (!m || (n && (!x || !t))) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    MLP_CHECK(hipDeviceSynchronize());   // worlds may be running the network on their streams
    uint32_t count = 0;
    MLP_CHECK(hipMemcpy(&count, m->smp.count, 4, hipMemcpyDeviceToHost));
    const uint64_t have = std::min<uint64_t>(count, m->smp.cap);
    if (have + n > m->smp.cap) return mfail(AGX_ERR_CAPACITY, "sample buffer holds %llu", (unsigned long long)m->smp.cap);
    std::vector<uint16_t> xb;
    padFeatures(x, n, xb);
    MLP_CHECK(hipMemcpy(m->smp.x + have * agx::MLP_INP, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    MLP_CHECK(hipMemcpy(m->smp.t + have * agx::MLP_OUT, t, n * agx::MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
    const uint32_t next = (uint32_t)(have + n);
    MLP_CHECK(hipMemcpy(m->smp.count, &next, 4, hipMemcpyHostToDevice));
    return AGX_OK;
}

int agx_mlp_sample_count(agx_mlp* m, uint64_t* count) {
    if (!m || !count) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    MLP_CHECK(hipDeviceSynchronize());
    uint32_t c = 0;
    MLP_CHECK(hipMemcpy(&c, m->smp.count, 4, hipMemcpyDeviceToHost));
    *count = std::min<uint64_t>(c, m->smp.cap);
    return AGX_OK;
}

int agx_mlp_read_samples(agx_mlp* m, float* x, float* t, uint64_t offset, uint64_t n) {
    if (!m || (n && (!x || !t))) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t have = 0;
    if (int rc = agx_mlp_sample_count(m, &have)) return rc;
    if (offset + n > have) return mfail(AGX_ERR_CAPACITY, "only %llu samples", (unsigned long long)have);
    std::vector<uint16_t> xb(n * agx::MLP_INP);
    MLP_CHECK(hipMemcpy(xb.data(), m->smp.x + offset * agx::MLP_INP, xb.size() * 2, hipMemcpyDeviceToHost));
    MLP_CHECK(hipMemcpy(t, m->smp.t + offset * agx::MLP_OUT, n * agx::MLP_OUT * sizeof(float), hipMemcpyDeviceToHost));
    for (uint64_t r = 0; r < n; r++)
        for (uint32_t i = 0; i < agx::MLP_IN; i++) x[r * agx::MLP_IN + i] = agx::bf16Value(xb[r * agx::MLP_INP + i]);
    return AGX_OK;
}

int agx_mlp_clear_samples(agx_mlp* m) {
    if (!m) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    MLP_CHECK(hipDeviceSynchronize());
    MLP_CHECK(hipMemset(m->smp.count, 0, 4));
    return AGX_OK;
}

int agx_mlp_train(agx_mlp* m, uint32_t iterations, uint32_t batch, float learning_rate, uint32_t seed, float* mean_loss) {
    if (!m) return mfail(AGX_ERR_INVALID_ARGUMENT, "null argument");
    if (batch == 0 || batch % 16 != 0 || batch > agx::MLP_BATCH_MAX)
        return mfail(AGX_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 16 up to %u", agx::MLP_BATCH_MAX);
    uint64_t have = 0;
    if (int rc = agx_mlp_sample_count(m, &have)) return rc;
    if (have == 0) return mfail(AGX_ERR_STATE, "no training samples");
    float* lossSum = nullptr;
    MLP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&lossSum), sizeof(float), m->stream));
    MLP_CHECK(hipMemsetAsync(lossSum, 0, sizeof(float), m->stream));
    for (uint32_t it = 0; it < iterations; it++) {
        MLP_CHECK(agx::launchMlpGatherBatch(m->stream, m->smp, batch, seed * 0x9E3779B9u + it, m->s));
        MLP_CHECK(agx::launchMlpForward(m->stream, m->p, m->s.x, batch, m->s.y, m->s.h1, m->s.h2, nullptr));
        MLP_CHECK(agx::launchMlpBackward(m->stream, m->p, m->g, m->s.x, m->s.y, m->s.t, batch, m->s.h1, m->s.h2, m->s,
                                         lossSum));
        MLP_CHECK(agx::launchMlpAdam(m->stream, m->p, m->g, m->adam, learning_rate));
    }
    float l = 0.0f;
    MLP_CHECK(hipMemcpyAsync(&l, lossSum, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_CHECK(hipFreeAsync(lossSum, m->stream));
    MLP_CHECK(hipStreamSynchronize(m->stream));
    if (mean_loss) *mean_loss = iterations ? l / (float)iterations : 0.0f;
    return AGX_OK;
}

}  // extern "C"
