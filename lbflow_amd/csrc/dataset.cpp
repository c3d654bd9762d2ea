// dataset.cpp -- C ABI of the warm-start network's training set and of plain
// device buffers (include/lbflow.h, lbf_dataset_* and lbf_buffer_*), so that
// generating, scaling, augmenting and feeding training data runs on the
// engine's own HIP runtime: steady fields go from a solved batch straight
// into the training set, and each minibatch is gathered, with one of the 16
// lattice symmetries applied, into the buffers the network trains on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbflow.h"

namespace lbf {
hipError_t launch_data_sumsq(const uint8_t* mask, const float* f, uint64_t nvox, double nu_over_f, double inv_f,
                             double* partial, hipStream_t s);
hipError_t launch_data_normalise(const uint8_t* mask, float* f, uint64_t nvox, float nu_over_f, float force,
                                 const float* rms, hipStream_t s);
hipError_t launch_data_gather(const uint8_t* masks, const float* t, const uint32_t* idx, uint32_t ls, uint32_t n,
                              uint32_t sym, const float* mul, uint8_t* mo, float* to, hipStream_t s);
constexpr uint32_t DATA_BLOCKS = 1024;
}  // namespace lbf

using namespace lbf;

struct lbf_dataset {
    uint32_t side = 0, ls = 0, count = 0;
    uint64_t s3 = 0;
    hipStream_t stream = nullptr;
    uint8_t* masks = nullptr;        // [count][s3]
    float* targets = nullptr;        // [count][s3][4]
    double* partial = nullptr;       // [DATA_BLOCKS][4]
    // minibatch staging: masks, targets and sample indices of up to `stage_n` samples
    uint8_t* stage_mask = nullptr;
    float* stage_target = nullptr;
    uint32_t* stage_idx = nullptr;
    uint32_t stage_n = 0;
    double rms[4] = {1.0, 1.0, 1.0, 1.0};
    bool normalised = false;
    std::string err;
};

namespace {

thread_local std::string g_data_error;

int dfail(lbf_dataset* d, const std::string& m)
{
    if (d) d->err = m;
    else g_data_error = m;
    return -1;
}

#define DHIP(d, expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return dfail(d, std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

// Signs and RMS ratios of the output channels of symmetry `sym` (data_gather).
void symmetry_mul(const lbf_dataset* d, uint32_t sym, float* mul)
{
    float sign[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (sym & 1u) sign[0] = -sign[0];
    if (sym & 2u) sign[1] = -sign[1];
    if (sym & 8u) {
        sign[0] = -sign[0];
        sign[1] = -sign[1];
        sign[3] = -sign[3];
    }
    if (sym & 4u) {
        // channels are normalised per component: carry u_y <-> u_z across their RMS
        mul[0] = sign[1] * float(d->rms[1] / d->rms[0]);
        mul[1] = sign[0] * float(d->rms[0] / d->rms[1]);
    } else {
        mul[0] = sign[0];
        mul[1] = sign[1];
    }
    mul[2] = sign[2];
    mul[3] = sign[3];
}

int stage(lbf_dataset* d, uint32_t n)
{
    if (n <= d->stage_n) return 0;
    if (d->stage_mask) (void)hipFree(d->stage_mask);
    if (d->stage_target) (void)hipFree(d->stage_target);
    if (d->stage_idx) (void)hipFree(d->stage_idx);
    d->stage_mask = nullptr;
    d->stage_target = nullptr;
    d->stage_idx = nullptr;
    d->stage_n = 0;
    DHIP(d, hipMalloc(reinterpret_cast<void**>(&d->stage_mask), size_t(n) * d->s3));
    DHIP(d, hipMalloc(reinterpret_cast<void**>(&d->stage_target), size_t(n) * d->s3 * 16));
    DHIP(d, hipMalloc(reinterpret_cast<void**>(&d->stage_idx), size_t(n) * 4));
    d->stage_n = n;
    return 0;
}

// Gathers samples idx[0..n) mapped by `sym` into the staging buffers.
int gather(lbf_dataset* d, const uint32_t* idx, uint32_t n, uint32_t sym)
{
    if (sym > 15) return dfail(d, "lbf_dataset: symmetry outside 0..15");
    for (uint32_t j = 0; j < n; ++j)
        if (idx[j] >= d->count) return dfail(d, "lbf_dataset: sample index outside the data set");
    if (stage(d, n)) return -1;
    float mul[4];
    symmetry_mul(d, sym, mul);
    DHIP(d, hipMemcpyAsync(d->stage_idx, idx, size_t(n) * 4, hipMemcpyHostToDevice, d->stream));
    DHIP(d, launch_data_gather(d->masks, d->targets, d->stage_idx, d->ls, n, sym, mul, d->stage_mask,
                               d->stage_target, d->stream));
    DHIP(d, hipStreamSynchronize(d->stream));   // idx is host memory of the caller
    return 0;
}

}  // namespace

extern "C" {

int lbf_buffer_alloc(uint64_t bytes, void** ptr)
{
    if (!ptr) return -1;
    *ptr = nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return -1;
    if (hipMemset(p, 0, bytes ? bytes : 16) != hipSuccess) {
        (void)hipFree(p);
        return -1;
    }
    *ptr = p;
    return 0;
}

void lbf_buffer_free(void* ptr)
{
    if (ptr) (void)hipFree(ptr);
}

int lbf_buffer_upload(void* dst, const void* host, uint64_t bytes)
{
    if (!dst || !host) return -1;
    return hipMemcpy(dst, host, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int lbf_buffer_download(void* host, const void* src, uint64_t bytes)
{
    if (!host || !src) return -1;
    return hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

const char* lbf_dataset_error(const lbf_dataset* d) { return d ? d->err.c_str() : g_data_error.c_str(); }

void lbf_dataset_destroy(lbf_dataset* d)
{
    if (!d) return;
    for (void* p : {static_cast<void*>(d->masks), static_cast<void*>(d->targets), static_cast<void*>(d->partial),
                    static_cast<void*>(d->stage_mask), static_cast<void*>(d->stage_target),
                    static_cast<void*>(d->stage_idx)})
        if (p) (void)hipFree(p);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int lbf_dataset_create(uint32_t side, uint32_t count, lbf_dataset** out)
{
    g_data_error.clear();
    if (!out) return dfail(nullptr, "lbf_dataset_create: null argument");
    *out = nullptr;
    if (side < 8 || (side & (side - 1)) || side > 256)
        return dfail(nullptr, "lbf_dataset_create: the side must be a power of two in [8, 256]");
    if (count == 0) return dfail(nullptr, "lbf_dataset_create: no samples");
    lbf_dataset* d = new lbf_dataset();
    d->side = side;
    while ((1u << d->ls) < side) ++d->ls;
    d->count = count;
    d->s3 = uint64_t(side) * side * side;
    const bool ok = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) == hipSuccess &&
                    hipMalloc(reinterpret_cast<void**>(&d->masks), size_t(count) * d->s3) == hipSuccess &&
                    hipMalloc(reinterpret_cast<void**>(&d->targets), size_t(count) * d->s3 * 16) == hipSuccess &&
                    hipMalloc(reinterpret_cast<void**>(&d->partial), size_t(DATA_BLOCKS) * 4 * 8) == hipSuccess &&
                    hipMemset(d->masks, 0, size_t(count) * d->s3) == hipSuccess &&
                    hipMemset(d->targets, 0, size_t(count) * d->s3 * 16) == hipSuccess;
    if (!ok) {
        lbf_dataset_destroy(d);
        return dfail(nullptr, "HIP error: cannot allocate the data set (no GPU, or out of memory)");
    }
    *out = d;
    return 0;
}

int lbf_dataset_set_masks(lbf_dataset* d, uint32_t first, uint32_t n, const uint8_t* masks)
{
    if (!d || !masks) return -1;
    if (uint64_t(first) + n > d->count) return dfail(d, "lbf_dataset_set_masks: samples outside the data set");
    DHIP(d, hipMemcpy(d->masks + size_t(first) * d->s3, masks, size_t(n) * d->s3, hipMemcpyHostToDevice));
    return 0;
}

int lbf_dataset_set_targets(lbf_dataset* d, uint32_t first, uint32_t n, const float* targets)
{
    if (!d || !targets) return -1;
    if (uint64_t(first) + n > d->count) return dfail(d, "lbf_dataset_set_targets: samples outside the data set");
    DHIP(d, hipMemcpy(d->targets + size_t(first) * d->s3 * 4, targets, size_t(n) * d->s3 * 16,
                      hipMemcpyHostToDevice));
    return 0;
}

int lbf_dataset_take_fields(lbf_dataset* d, uint32_t first, lbf_batch* b, uint32_t crop0, uint32_t n)
{
    if (!d || !b) return -1;
    if (uint64_t(first) + n > d->count) return dfail(d, "lbf_dataset_take_fields: samples outside the data set");
    for (uint32_t c = 0; c < n; ++c) {
        uint64_t sz[3] = {0, 0, 0};
        if (lbf_batch_crop_size(b, crop0 + c, sz) != 3 || sz[0] != d->side || sz[1] != d->side || sz[2] != d->side)
            return dfail(d, "lbf_dataset_take_fields: crop " + std::to_string(crop0 + c) + " is not a " +
                                std::to_string(d->side) + "^3 crop of the batch");
        if (lbf_batch_crop_fields(b, crop0 + c, d->targets + size_t(first + c) * d->s3 * 4))
            return dfail(d, std::string("lbf_dataset_take_fields: ") + lbf_batch_error(b));
    }
    return 0;
}

int lbf_dataset_normalise(lbf_dataset* d, double tau, double force, int fit, double* rms)
{
    if (!d || !rms) return -1;
    if (d->normalised) return dfail(d, "lbf_dataset_normalise: the targets are normalised already");
    if (!(tau > 0.5) || force == 0.0 || !std::isfinite(force))
        return dfail(d, "lbf_dataset_normalise: tau must exceed 0.5 and the force must be finite and non-zero");
    const double nu = (tau - 0.5) / 3.0;
    const uint64_t nvox = uint64_t(d->count) * d->s3;
    if (fit) {
        DHIP(d, launch_data_sumsq(d->masks, d->targets, nvox, nu / force, 1.0 / force, d->partial, d->stream));
        std::vector<double> part(size_t(DATA_BLOCKS) * 4);
        DHIP(d, hipMemcpyAsync(part.data(), d->partial, part.size() * 8, hipMemcpyDeviceToHost, d->stream));
        DHIP(d, hipStreamSynchronize(d->stream));
        std::vector<uint8_t> m(size_t(d->count) * d->s3);
        DHIP(d, hipMemcpy(m.data(), d->masks, m.size(), hipMemcpyDeviceToHost));
        uint64_t pores = 0;
        for (uint8_t x : m) pores += x != 0;
        if (pores == 0) return dfail(d, "lbf_dataset_normalise: no pore voxel to fit the scaling on");
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
            for (uint32_t k = 0; k < DATA_BLOCKS; ++k) s += part[size_t(k) * 4 + c];
            rms[c] = std::sqrt(s / double(pores));
            if (!(rms[c] > 0.0)) return dfail(d, "lbf_dataset_normalise: a target channel is 0 everywhere");
        }
    }
    const float r32[4] = {float(rms[0]), float(rms[1]), float(rms[2]), float(rms[3])};
    DHIP(d, launch_data_normalise(d->masks, d->targets, nvox, float(nu / force), float(force), r32, d->stream));
    DHIP(d, hipStreamSynchronize(d->stream));
    for (int c = 0; c < 4; ++c) d->rms[c] = rms[c];
    d->normalised = true;
    return 0;
}

int lbf_dataset_sample(lbf_dataset* d, uint32_t i, uint32_t sym, uint8_t* mask, float* target)
{
    if (!d) return -1;
    if (gather(d, &i, 1, sym)) return -1;
    if (mask) DHIP(d, hipMemcpy(mask, d->stage_mask, d->s3, hipMemcpyDeviceToHost));
    if (target) DHIP(d, hipMemcpy(target, d->stage_target, d->s3 * 16, hipMemcpyDeviceToHost));
    return 0;
}

int lbf_dataset_device(const lbf_dataset* d, uint8_t** masks, float** targets)
{
    if (!d) return -1;
    if (masks) *masks = d->masks;
    if (targets) *targets = d->targets;
    return 0;
}

int lbf_net_train_samples(lbf_net* n, lbf_dataset* d, const uint32_t* idx, uint32_t count, uint32_t sym,
                          const lbf_adam* opt, double* loss)
{
    if (!n || !d || !idx || !opt) return -1;
    if (lbf_net_side(n) != d->side) return dfail(d, "lbf_net_train_samples: the network's side is not the data set's");
    if (gather(d, idx, count, sym)) return -1;
    if (lbf_net_train_step(n, d->stage_mask, d->stage_target, count, opt, loss))
        return dfail(d, std::string("lbf_net_train_samples: ") + lbf_net_error(n));
    return 0;
}

}  // extern "C"
