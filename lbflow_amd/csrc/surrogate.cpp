// surrogate.cpp -- the learned warm-start network (include/lbflow.h, lbf_net_*).
//
// A 4-level periodic 3D U-Net that maps a crop's fluid mask to its steady
// (u, rho): per level two 3x3x3 convolutions with ReLU, 2x average pooling
// down, nearest upsampling up, skip connections by channel concatenation,
// and a linear 1x1x1 head.  Channels 32/64/128/256, 5.8 M parameters.
// Every convolution (forward, input gradient, weight gradient) runs on the
// BF16 MFMA kernels of kernels_conv.hip with FP32 accumulation; the
// optimizer keeps FP32 master weights (Adam) and refreshes the BF16 copies
// and their transposed, tap-mirrored twins after each step.
//
// Activations are channels-last BF16.  Concatenations are free: an encoder
// writes its output straight into the channel range of the decoder's input
// buffer that the skip connection occupies, and the upsampling writes the
// other range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lbflow.h"
#include "conv_types.h"

namespace lbf {
uint32_t conv_wgrad_splits(int cin, int cout, int taps, uint32_t nvox);
hipError_t launch_conv_forward(int cin, int cout, int taps, const ConvGeom& g, const __bf16* in, const __bf16* w,
                               const float* bias, void* out, int act, bool out_f32, hipStream_t s);
hipError_t launch_conv_wgrad(int cin, int cout, int taps, const ConvGeom& g, const __bf16* in, const __bf16* dz,
                             float* partial, uint32_t splits, float* dw, float* bpart, float* db, hipStream_t s);
bool conv_wgrad_fuses_bias(int cin, int cout, int taps, uint32_t log_s);
bool conv_shape_supported(int cin, int cout, int taps);
hipError_t launch_net_input(const uint8_t* mask, uint32_t nvox, __bf16* x0, hipStream_t s);
hipError_t launch_net_pool(Chan in, Chan out, uint32_t C, uint32_t ls_out, uint32_t nvox_out, hipStream_t s);
hipError_t launch_net_upsample(Chan in, Chan out, uint32_t C, uint32_t ls_out, uint32_t nvox_out, hipStream_t s);
hipError_t launch_net_enc_grad(Chan dpool, Chan dskip, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox,
                               hipStream_t s);
hipError_t launch_net_up_grad(Chan dup, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox, hipStream_t s);
hipError_t launch_net_loss(const float* out, const float* target, const float* w, uint32_t nvox, float scale,
                           __bf16* dout, double* partial, double* loss, hipStream_t s);
hipError_t launch_net_bias_grad(Chan dz, uint32_t C, uint32_t nvox, float* partial, uint32_t blocks, float* db,
                                hipStream_t s);
hipError_t launch_net_adam(float* p, float* m, float* v, const float* g, uint32_t n, float lr, float b1, float b2,
                           float eps, float c1, float c2, __bf16* pb, hipStream_t s);
hipError_t launch_net_head4(const float* out16, uint32_t nvox, float* out4, hipStream_t s);
hipError_t launch_net_head_fields(const float* out16, const uint8_t* mask, uint32_t nvox, const float* scale,
                                  const float* offset, float* out4, hipStream_t s);
hipError_t launch_net_to_bf16(const float* p, uint32_t n, __bf16* pb, hipStream_t s);
hipError_t launch_net_transpose_w(const __bf16* w, uint32_t cin, uint32_t cout, uint32_t taps, uint32_t kp,
                                  uint32_t kpt, __bf16* wt, hipStream_t s);
}  // namespace lbf

using namespace lbf;

namespace {

constexpr int NLAYERS = 15;
constexpr int HEAD = 14;
constexpr int HEAD_PAD = 16;   // the head's 4 outputs padded to one MFMA column tile
// forward split over input channels: at most 512 workgroups x 256 voxels x 64
// output channels of FP32 partials
constexpr uint32_t KPART_FLOATS = 512u * 256u * 64u;

// cin, cout, taps, level
constexpr int kLayers[NLAYERS][4] = {
    {8, 32, 27, 0},    {32, 32, 27, 0},    {32, 64, 27, 1},   {64, 64, 27, 1},   {64, 128, 27, 2},
    {128, 128, 27, 2}, {128, 256, 27, 3},  {256, 256, 27, 3}, {384, 128, 27, 2}, {128, 128, 27, 2},
    {192, 64, 27, 1},  {64, 64, 27, 1},    {96, 32, 27, 0},   {32, 32, 27, 0},   {32, HEAD_PAD, 1, 0}};

struct Layer {
    int cin = 0, cout = 0, taps = 0, level = 0, real_out = 0;
    uint32_t kp = 0, kpt = 0;
    float *w = nullptr, *b = nullptr;          // FP32 master [cout][kp], [cout]
    float *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    __bf16 *wb = nullptr, *wt = nullptr;       // BF16 copy, transposed-mirrored copy [cin][kpt]
};

}  // namespace

struct lbf_net {
    uint32_t side = 0, ls = 0, cap = 0;
    hipStream_t stream = nullptr;
    Layer L[NLAYERS];
    std::vector<void*> allocations;
    std::string err;
    uint64_t step = 0;
    // activations
    __bf16 *x0, *e1a, *cat1, *d1a, *d1, *p1, *e2a, *cat2, *d2a, *d2, *p2, *e3a, *cat3, *d3a, *d3, *p3, *e4a, *e4;
    float* out16;
    // gradients
    __bf16 *dout, *gA[4], *gB[4], *gcat1, *gcat2, *gcat3, *gp1, *gp2, *gp3;
    float *wpart, *bpart;
    float* kpart;                  // forward split over input channels: FP32 partials
    double *lpart, *loss;
    size_t wpart_floats = 0;
    // warm starts: capacity pore masks and their physical fields, allocated on first use
    uint8_t* warm_mask = nullptr;
    float* warm_fields = nullptr;
};

namespace {

thread_local std::string g_net_error;

int nfail(lbf_net* n, const std::string& m)
{
    if (n) n->err = m;
    else g_net_error = m;
    return -1;
}

#define NHIP(n, expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return nfail(n, std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T> int nalloc(lbf_net* n, T** p, size_t count)
{
    void* q = nullptr;
    const size_t bytes = count * sizeof(T);
    if (hipMalloc(&q, bytes ? bytes : 16) != hipSuccess) return nfail(n, "HIP error: hipMalloc failed");
    n->allocations.push_back(q);
    if (hipMemset(q, 0, bytes ? bytes : 16) != hipSuccess) return nfail(n, "HIP error: hipMemset failed");
    *p = static_cast<T*>(q);
    return 0;
}

uint32_t nvox_at(const lbf_net* n, int level, uint32_t count) { return count * (1u << (3 * (n->ls - level))); }

ConvGeom geom(const lbf_net* n, int level, uint32_t count, uint32_t in_stride, uint32_t out_stride)
{
    ConvGeom g{n->ls - uint32_t(level), nvox_at(n, level, count), in_stride, out_stride};
    g.kpart = n->kpart;
    g.kpart_floats = KPART_FLOATS;
    return g;
}

Chan ch(__bf16* p, uint32_t stride, uint32_t off = 0) { return Chan{p, stride, off}; }

int refresh_bf16(lbf_net* n)
{
    for (Layer& l : n->L) {
        NHIP(n, launch_net_to_bf16(l.w, uint32_t(l.cout) * l.kp, l.wb, n->stream));
        NHIP(n, launch_net_transpose_w(l.wb, l.cin, l.cout, l.taps, l.kp, l.kpt, l.wt, n->stream));
    }
    return 0;
}

// conv + bias + activation of layer i
int fwd(lbf_net* n, int i, uint32_t count, const __bf16* in, uint32_t in_stride, void* out, uint32_t out_stride,
        bool relu, bool f32 = false)
{
    const Layer& l = n->L[i];
    NHIP(n, launch_conv_forward(l.cin, l.cout, l.taps, geom(n, l.level, count, in_stride, out_stride), in, l.wb, l.b,
                                out, relu ? ACT_RELU : ACT_NONE, f32, n->stream));
    return 0;
}

// weight and bias gradient of layer i from its input and its pre-activation gradient
int wgrad(lbf_net* n, int i, uint32_t count, const __bf16* in, uint32_t in_stride, const __bf16* dz,
          uint32_t dz_stride)
{
    const Layer& l = n->L[i];
    const ConvGeom g = geom(n, l.level, count, in_stride, dz_stride);
    uint32_t splits = conv_wgrad_splits(l.cin, l.cout, l.taps, g.nvox);
    if (size_t(splits) * l.cout * l.kp > n->wpart_floats) splits = uint32_t(n->wpart_floats / (size_t(l.cout) * l.kp));
    if (splits < 1) splits = 1;
    // the tiled weight-gradient kernels reduce the bias gradient on the way
    if (conv_wgrad_fuses_bias(l.cin, l.cout, l.taps, g.log_s)) {
        NHIP(n, launch_conv_wgrad(l.cin, l.cout, l.taps, g, in, dz, n->wpart, splits, l.gw, n->bpart, l.gb,
                                  n->stream));
        return 0;
    }
    NHIP(n, launch_conv_wgrad(l.cin, l.cout, l.taps, g, in, dz, n->wpart, splits, l.gw, nullptr, nullptr,
                              n->stream));
    uint32_t blocks = g.nvox / 1024;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    NHIP(n, launch_net_bias_grad(ch(const_cast<__bf16*>(dz), dz_stride), l.cout, g.nvox, n->bpart, blocks, l.gb,
                                 n->stream));
    return 0;
}

// input gradient of layer i: the forward kernel on the transposed-mirrored
// weights; with a gate, the ReLU gradient of the layer that produced the
// input (dx = gate > 0 ? dx : 0) is applied as dx is stored
int dgrad(lbf_net* n, int i, uint32_t count, const __bf16* dz, uint32_t dz_stride, __bf16* dx, uint32_t dx_stride,
          const __bf16* gate = nullptr, uint32_t gate_stride = 0)
{
    const Layer& l = n->L[i];
    ConvGeom g = geom(n, l.level, count, dz_stride, dx_stride);
    g.gate = gate;
    g.gate_stride = gate_stride;
    NHIP(n, launch_conv_forward(l.cout, l.cin, l.taps, g, dz, l.wt, nullptr, dx, gate ? ACT_RELU_GATE : ACT_NONE,
                                false, n->stream));
    return 0;
}

int forward(lbf_net* n, const uint8_t* mask, uint32_t c)
{
    const uint32_t v0 = nvox_at(n, 0, c), v1 = nvox_at(n, 1, c), v2 = nvox_at(n, 2, c), v3 = nvox_at(n, 3, c);
    const uint32_t l0 = n->ls;
    hipStream_t s = n->stream;
    NHIP(n, launch_net_input(mask, v0, n->x0, s));
    if (fwd(n, 0, c, n->x0, 8, n->e1a, 32, true) || fwd(n, 1, c, n->e1a, 32, n->cat1 + 64, 96, true)) return -1;
    NHIP(n, launch_net_pool(ch(n->cat1, 96, 64), ch(n->p1, 32), 32, l0 - 1, v1, s));
    if (fwd(n, 2, c, n->p1, 32, n->e2a, 64, true) || fwd(n, 3, c, n->e2a, 64, n->cat2 + 128, 192, true)) return -1;
    NHIP(n, launch_net_pool(ch(n->cat2, 192, 128), ch(n->p2, 64), 64, l0 - 2, v2, s));
    if (fwd(n, 4, c, n->p2, 64, n->e3a, 128, true) || fwd(n, 5, c, n->e3a, 128, n->cat3 + 256, 384, true)) return -1;
    NHIP(n, launch_net_pool(ch(n->cat3, 384, 256), ch(n->p3, 128), 128, l0 - 3, v3, s));
    if (fwd(n, 6, c, n->p3, 128, n->e4a, 256, true) || fwd(n, 7, c, n->e4a, 256, n->e4, 256, true)) return -1;
    NHIP(n, launch_net_upsample(ch(n->e4, 256), ch(n->cat3, 384, 0), 256, l0 - 2, v2, s));
    if (fwd(n, 8, c, n->cat3, 384, n->d3a, 128, true) || fwd(n, 9, c, n->d3a, 128, n->d3, 128, true)) return -1;
    NHIP(n, launch_net_upsample(ch(n->d3, 128), ch(n->cat2, 192, 0), 128, l0 - 1, v1, s));
    if (fwd(n, 10, c, n->cat2, 192, n->d2a, 64, true) || fwd(n, 11, c, n->d2a, 64, n->d2, 64, true)) return -1;
    NHIP(n, launch_net_upsample(ch(n->d2, 64), ch(n->cat1, 96, 0), 64, l0, v0, s));
    if (fwd(n, 12, c, n->cat1, 96, n->d1a, 32, true) || fwd(n, 13, c, n->d1a, 32, n->d1, 32, true)) return -1;
    return fwd(n, HEAD, c, n->d1, 32, n->out16, HEAD_PAD, false, true);
}

int backward(lbf_net* n, uint32_t c)
{
    const uint32_t v0 = nvox_at(n, 0, c), v1 = nvox_at(n, 1, c), v2 = nvox_at(n, 2, c), v3 = nvox_at(n, 3, c);
    const uint32_t l0 = n->ls;
    hipStream_t s = n->stream;
    __bf16 **A = n->gA, **B = n->gB;
    // head and level-0 decoder
    if (wgrad(n, 14, c, n->d1, 32, n->dout, HEAD_PAD) || dgrad(n, 14, c, n->dout, HEAD_PAD, A[0], 32, n->d1, 32))
        return -1;
    if (wgrad(n, 13, c, n->d1a, 32, A[0], 32) || dgrad(n, 13, c, A[0], 32, B[0], 32, n->d1a, 32)) return -1;
    if (wgrad(n, 12, c, n->cat1, 96, B[0], 32) || dgrad(n, 12, c, B[0], 32, n->gcat1, 96)) return -1;
    // level-1 decoder
    NHIP(n, launch_net_up_grad(ch(n->gcat1, 96, 0), ch(n->d2, 64), ch(A[1], 64), 64, l0 - 1, v1, s));
    if (wgrad(n, 11, c, n->d2a, 64, A[1], 64) || dgrad(n, 11, c, A[1], 64, B[1], 64, n->d2a, 64)) return -1;
    if (wgrad(n, 10, c, n->cat2, 192, B[1], 64) || dgrad(n, 10, c, B[1], 64, n->gcat2, 192)) return -1;
    // level-2 decoder
    NHIP(n, launch_net_up_grad(ch(n->gcat2, 192, 0), ch(n->d3, 128), ch(A[2], 128), 128, l0 - 2, v2, s));
    if (wgrad(n, 9, c, n->d3a, 128, A[2], 128) || dgrad(n, 9, c, A[2], 128, B[2], 128, n->d3a, 128)) return -1;
    if (wgrad(n, 8, c, n->cat3, 384, B[2], 128) || dgrad(n, 8, c, B[2], 128, n->gcat3, 384)) return -1;
    // bottom
    NHIP(n, launch_net_up_grad(ch(n->gcat3, 384, 0), ch(n->e4, 256), ch(A[3], 256), 256, l0 - 3, v3, s));
    if (wgrad(n, 7, c, n->e4a, 256, A[3], 256) || dgrad(n, 7, c, A[3], 256, B[3], 256, n->e4a, 256)) return -1;
    if (wgrad(n, 6, c, n->p3, 128, B[3], 256) || dgrad(n, 6, c, B[3], 256, n->gp3, 128)) return -1;
    // level-2 encoder
    NHIP(n, launch_net_enc_grad(ch(n->gp3, 128), ch(n->gcat3, 384, 256), ch(n->cat3, 384, 256), ch(A[2], 128), 128,
                                l0 - 2, v2, s));
    if (wgrad(n, 5, c, n->e3a, 128, A[2], 128) || dgrad(n, 5, c, A[2], 128, B[2], 128, n->e3a, 128)) return -1;
    if (wgrad(n, 4, c, n->p2, 64, B[2], 128) || dgrad(n, 4, c, B[2], 128, n->gp2, 64)) return -1;
    // level-1 encoder
    NHIP(n, launch_net_enc_grad(ch(n->gp2, 64), ch(n->gcat2, 192, 128), ch(n->cat2, 192, 128), ch(A[1], 64), 64,
                                l0 - 1, v1, s));
    if (wgrad(n, 3, c, n->e2a, 64, A[1], 64) || dgrad(n, 3, c, A[1], 64, B[1], 64, n->e2a, 64)) return -1;
    if (wgrad(n, 2, c, n->p1, 32, B[1], 64) || dgrad(n, 2, c, B[1], 64, n->gp1, 32)) return -1;
    // level-0 encoder
    NHIP(n, launch_net_enc_grad(ch(n->gp1, 32), ch(n->gcat1, 96, 64), ch(n->cat1, 96, 64), ch(A[0], 32), 32, l0, v0,
                                s));
    if (wgrad(n, 1, c, n->e1a, 32, A[0], 32) || dgrad(n, 1, c, A[0], 32, B[0], 32, n->e1a, 32)) return -1;
    return wgrad(n, 0, c, n->x0, 8, B[0], 32);
}

int loss_and_grad(lbf_net* n, const uint8_t* mask, const float* target, uint32_t c, const float* w4, double* loss)
{
    if (c == 0 || c > n->cap) return nfail(n, "lbf_net: sample count outside [1, capacity]");
    if (forward(n, mask, c)) return -1;
    const uint32_t v0 = nvox_at(n, 0, c);
    NHIP(n, launch_net_loss(n->out16, target, w4, v0, 1.0f / float(v0), n->dout, n->lpart, n->loss, n->stream));
    if (backward(n, c)) return -1;
    double l = 0.0;
    NHIP(n, hipMemcpyAsync(&l, n->loss, sizeof l, hipMemcpyDeviceToHost, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    if (loss) *loss = l / double(v0);
    return 0;
}

uint64_t real_params(const Layer& l) { return uint64_t(l.real_out) * (uint64_t(l.taps) * l.cin + 1); }

}  // namespace

extern "C" {

const char* lbf_net_error(const lbf_net* n) { return n ? n->err.c_str() : g_net_error.c_str(); }

void lbf_net_destroy(lbf_net* n)
{
    if (!n) return;
    for (void* p : n->allocations) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

int lbf_net_create(uint32_t side, uint32_t capacity, uint64_t seed, lbf_net** out)
{
    g_net_error.clear();
    if (!out) return nfail(nullptr, "lbf_net_create: null argument");
    *out = nullptr;
    if (side < 8 || side > 256 || (side & (side - 1)) || capacity == 0)
        return nfail(nullptr, "lbf_net_create: side must be a power of two in [8, 256] and capacity >= 1");
    uint32_t ls = 0;
    while ((1u << ls) < side) ++ls;
    const uint64_t v0 = uint64_t(capacity) << (3 * ls);
    // voxel indices and per-launch thread counts (voxels x channels / 8, at
    // most 96 channels at the finest level) stay 32-bit; element offsets are 64-bit
    if (v0 >= (uint64_t(1) << 31) || v0 * 96 / 8 >= (uint64_t(1) << 31))
        return nfail(nullptr, "lbf_net_create: batch too large for 32-bit voxel indexing");
    lbf_net* n = new lbf_net();
    n->side = side;
    n->ls = ls;
    n->cap = capacity;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        lbf_net_destroy(n);
        return nfail(nullptr, "HIP error: cannot create a stream (no GPU?)");
    }
    std::mt19937_64 rng(seed);
    int rc = 0;
    size_t wmax = 0;
    for (int i = 0; i < NLAYERS; ++i) {
        Layer& l = n->L[i];
        l.cin = kLayers[i][0];
        l.cout = kLayers[i][1];
        l.taps = kLayers[i][2];
        l.level = kLayers[i][3];
        l.real_out = i == HEAD ? 4 : l.cout;
        l.kp = uint32_t(conv_kp(l.cin, l.taps));
        l.kpt = uint32_t(conv_kp(l.cout, l.taps));
        const size_t nw = size_t(l.cout) * l.kp;
        rc |= nalloc(n, &l.w, nw);
        rc |= nalloc(n, &l.gw, nw);
        rc |= nalloc(n, &l.mw, nw);
        rc |= nalloc(n, &l.vw, nw);
        rc |= nalloc(n, &l.b, size_t(l.cout));
        rc |= nalloc(n, &l.gb, size_t(l.cout));
        rc |= nalloc(n, &l.mb, size_t(l.cout));
        rc |= nalloc(n, &l.vb, size_t(l.cout));
        rc |= nalloc(n, &l.wb, nw);
        rc |= nalloc(n, &l.wt, size_t(l.cin) * l.kpt);
        if (rc) break;
        // He-normal initialisation of the real weights, zero bias and padding
        std::vector<float> hw(nw, 0.0f);
        const double sd = i == HEAD ? 0.1 / std::sqrt(double(l.cin)) : std::sqrt(2.0 / (double(l.taps) * l.cin));
        std::normal_distribution<double> nd(0.0, sd);
        for (int co = 0; co < l.real_out; ++co)
            for (int k = 0; k < l.taps * l.cin; ++k) hw[size_t(co) * l.kp + k] = float(nd(rng));
        if (hipMemcpy(l.w, hw.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
        // weight-gradient partials at the largest batch
        const size_t splits = conv_wgrad_splits(l.cin, l.cout, l.taps, uint32_t(size_t(capacity) << (3 * (ls - l.level))));
        wmax = std::max(wmax, splits * nw);
    }
    const uint32_t v[4] = {uint32_t(v0), uint32_t(v0 >> 3), uint32_t(v0 >> 6), uint32_t(v0 >> 9)};
    if (!rc) {
        rc |= nalloc(n, &n->x0, size_t(v[0]) * 8);
        rc |= nalloc(n, &n->e1a, size_t(v[0]) * 32);
        rc |= nalloc(n, &n->cat1, size_t(v[0]) * 96);
        rc |= nalloc(n, &n->d1a, size_t(v[0]) * 32);
        rc |= nalloc(n, &n->d1, size_t(v[0]) * 32);
        rc |= nalloc(n, &n->out16, size_t(v[0]) * HEAD_PAD);
        rc |= nalloc(n, &n->p1, size_t(v[1]) * 32);
        rc |= nalloc(n, &n->e2a, size_t(v[1]) * 64);
        rc |= nalloc(n, &n->cat2, size_t(v[1]) * 192);
        rc |= nalloc(n, &n->d2a, size_t(v[1]) * 64);
        rc |= nalloc(n, &n->d2, size_t(v[1]) * 64);
        rc |= nalloc(n, &n->p2, size_t(v[2]) * 64);
        rc |= nalloc(n, &n->e3a, size_t(v[2]) * 128);
        rc |= nalloc(n, &n->cat3, size_t(v[2]) * 384);
        rc |= nalloc(n, &n->d3a, size_t(v[2]) * 128);
        rc |= nalloc(n, &n->d3, size_t(v[2]) * 128);
        rc |= nalloc(n, &n->p3, size_t(v[3]) * 128);
        rc |= nalloc(n, &n->e4a, size_t(v[3]) * 256);
        rc |= nalloc(n, &n->e4, size_t(v[3]) * 256);
        rc |= nalloc(n, &n->dout, size_t(v[0]) * HEAD_PAD);
        const uint32_t width[4] = {32, 64, 128, 256};
        for (int l = 0; l < 4; ++l) {
            rc |= nalloc(n, &n->gA[l], size_t(v[l]) * width[l]);
            rc |= nalloc(n, &n->gB[l], size_t(v[l]) * width[l]);
        }
        rc |= nalloc(n, &n->gcat1, size_t(v[0]) * 96);
        rc |= nalloc(n, &n->gcat2, size_t(v[1]) * 192);
        rc |= nalloc(n, &n->gcat3, size_t(v[2]) * 384);
        rc |= nalloc(n, &n->gp1, size_t(v[1]) * 32);
        rc |= nalloc(n, &n->gp2, size_t(v[2]) * 64);
        rc |= nalloc(n, &n->gp3, size_t(v[3]) * 128);
        n->wpart_floats = wmax;
        rc |= nalloc(n, &n->wpart, wmax);
        rc |= nalloc(n, &n->bpart, size_t(4096) * 256);
        rc |= nalloc(n, &n->kpart, KPART_FLOATS);
        rc |= nalloc(n, &n->lpart, size_t(v[0] / 256 + 1));
        rc |= nalloc(n, &n->loss, 1);
    }
    if (rc || refresh_bf16(n) || hipStreamSynchronize(n->stream) != hipSuccess) {
        std::string e = n->err.empty() ? "HIP error while creating the network" : n->err;
        lbf_net_destroy(n);
        return nfail(nullptr, e);
    }
    *out = n;
    return 0;
}

uint64_t lbf_net_param_count(const lbf_net* n)
{
    if (!n) return 0;
    uint64_t s = 0;
    for (const Layer& l : n->L) s += real_params(l);
    return s;
}

int lbf_net_get_params(lbf_net* n, float* host)
{
    if (!n || !host) return -1;
    NHIP(n, hipStreamSynchronize(n->stream));
    float* o = host;
    for (const Layer& l : n->L) {
        std::vector<float> w(size_t(l.cout) * l.kp), b(l.cout);
        NHIP(n, hipMemcpy(w.data(), l.w, w.size() * 4, hipMemcpyDeviceToHost));
        NHIP(n, hipMemcpy(b.data(), l.b, b.size() * 4, hipMemcpyDeviceToHost));
        for (int co = 0; co < l.real_out; ++co)
            for (int k = 0; k < l.taps * l.cin; ++k) *o++ = w[size_t(co) * l.kp + k];
        for (int co = 0; co < l.real_out; ++co) *o++ = b[co];
    }
    return 0;
}

int lbf_net_set_params(lbf_net* n, const float* host)
{
    if (!n || !host) return -1;
    NHIP(n, hipStreamSynchronize(n->stream));
    const float* p = host;
    for (Layer& l : n->L) {
        std::vector<float> w(size_t(l.cout) * l.kp, 0.0f), b(l.cout, 0.0f);
        for (int co = 0; co < l.real_out; ++co)
            for (int k = 0; k < l.taps * l.cin; ++k) w[size_t(co) * l.kp + k] = *p++;
        for (int co = 0; co < l.real_out; ++co) b[co] = *p++;
        NHIP(n, hipMemcpy(l.w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        NHIP(n, hipMemcpy(l.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
    }
    if (refresh_bf16(n)) return -1;
    NHIP(n, hipStreamSynchronize(n->stream));
    return 0;
}

uint32_t lbf_net_side(const lbf_net* n) { return n ? n->side : 0; }

int lbf_net_forward(lbf_net* n, const uint8_t* mask, uint32_t count, float* out)
{
    if (!n || !mask || !out) return -1;
    if (count == 0 || count > n->cap) return nfail(n, "lbf_net: sample count outside [1, capacity]");
    if (forward(n, mask, count)) return -1;
    NHIP(n, launch_net_head4(n->out16, nvox_at(n, 0, count), out, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    return 0;
}

int lbf_batch_warm_start(lbf_batch* b, lbf_net* n, const uint8_t* masks, uint32_t count, const float* scale,
                         const float* offset, uint32_t relax)
{
    if (!b || !n || !masks || !scale || !offset) return -1;
    const uint64_t s3 = uint64_t(n->side) * n->side * n->side;
    for (uint32_t c = 0; c < count; ++c) {
        uint64_t sz[3] = {0, 0, 0};
        if (lbf_batch_crop_size(b, c, sz) != 3 || sz[0] != n->side || sz[1] != n->side || sz[2] != n->side)
            return nfail(n, "lbf_batch_warm_start: crop " + std::to_string(c) + " is not a " +
                                std::to_string(n->side) + "^3 D3Q19 crop of the batch");
    }
    if (!n->warm_mask && (nalloc(n, &n->warm_mask, size_t(n->cap) * s3) ||
                          nalloc(n, &n->warm_fields, size_t(n->cap) * s3 * 4)))
        return -1;
    for (uint32_t c0 = 0; c0 < count; c0 += n->cap) {
        const uint32_t k = std::min(n->cap, count - c0);
        NHIP(n, hipMemcpyAsync(n->warm_mask, masks + size_t(c0) * s3, size_t(k) * s3, hipMemcpyHostToDevice,
                               n->stream));
        if (forward(n, n->warm_mask, k)) return -1;
        NHIP(n, launch_net_head_fields(n->out16, n->warm_mask, uint32_t(k * s3), scale, offset, n->warm_fields,
                                       n->stream));
        NHIP(n, hipStreamSynchronize(n->stream));
        for (uint32_t j = 0; j < k; ++j)
            if (lbf_batch_set_fields(b, c0 + j, n->warm_fields + size_t(j) * s3 * 4))
                return nfail(n, std::string("lbf_batch_warm_start: ") + lbf_batch_error(b));
    }
    if (relax && lbf_batch_relax(b, relax)) return nfail(n, std::string("lbf_batch_warm_start: ") + lbf_batch_error(b));
    return 0;
}

int lbf_net_gradients(lbf_net* n, const uint8_t* mask, const float* target, uint32_t count, const float* weights,
                      double* loss, float* grad_host)
{
    if (!n || !mask || !target || !weights) return -1;
    if (loss_and_grad(n, mask, target, count, weights, loss)) return -1;
    if (grad_host) {
        float* o = grad_host;
        for (const Layer& l : n->L) {
            std::vector<float> w(size_t(l.cout) * l.kp), b(l.cout);
            NHIP(n, hipMemcpy(w.data(), l.gw, w.size() * 4, hipMemcpyDeviceToHost));
            NHIP(n, hipMemcpy(b.data(), l.gb, b.size() * 4, hipMemcpyDeviceToHost));
            for (int co = 0; co < l.real_out; ++co)
                for (int k = 0; k < l.taps * l.cin; ++k) *o++ = w[size_t(co) * l.kp + k];
            for (int co = 0; co < l.real_out; ++co) *o++ = b[co];
        }
    }
    return 0;
}

int lbf_net_train_step(lbf_net* n, const uint8_t* mask, const float* target, uint32_t count, const lbf_adam* opt,
                       double* loss)
{
    if (!n || !mask || !target || !opt) return -1;
    if (loss_and_grad(n, mask, target, count, opt->weights, loss)) return -1;
    n->step++;
    const float c1 = float(1.0 - std::pow(double(opt->beta1), double(n->step)));
    const float c2 = float(1.0 - std::pow(double(opt->beta2), double(n->step)));
    for (Layer& l : n->L) {
        const uint32_t nw = uint32_t(l.cout) * l.kp;
        NHIP(n, launch_net_adam(l.w, l.mw, l.vw, l.gw, nw, opt->lr, opt->beta1, opt->beta2, opt->eps, c1, c2, l.wb,
                                n->stream));
        NHIP(n, launch_net_adam(l.b, l.mb, l.vb, l.gb, uint32_t(l.cout), opt->lr, opt->beta1, opt->beta2, opt->eps,
                                c1, c2, nullptr, n->stream));
        NHIP(n, launch_net_transpose_w(l.wb, l.cin, l.cout, l.taps, l.kp, l.kpt, l.wt, n->stream));
    }
    NHIP(n, hipStreamSynchronize(n->stream));
    return 0;
}

int lbf_net_timed_train(lbf_net* n, const uint8_t* mask, const float* target, uint32_t count, const lbf_adam* opt,
                        uint32_t steps, double* ms, double* flops_per_step)
{
    if (!n || !ms) return -1;
    hipEvent_t e0, e1;
    NHIP(n, hipEventCreate(&e0));
    NHIP(n, hipEventCreate(&e1));
    NHIP(n, hipEventRecord(e0, n->stream));
    for (uint32_t k = 0; k < steps; ++k)
        if (lbf_net_train_step(n, mask, target, count, opt, nullptr)) return -1;
    NHIP(n, hipEventRecord(e1, n->stream));
    NHIP(n, hipEventSynchronize(e1));
    float t = 0.0f;
    NHIP(n, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = double(t) / steps;
    if (flops_per_step) {
        // forward + input gradient + weight gradient: 3 GEMMs of 2 * voxels * cout * taps * cin each
        // (the first layer has no input gradient)
        double f = 0.0;
        for (int i = 0; i < NLAYERS; ++i) {
            const Layer& l = n->L[i];
            const double per = 2.0 * nvox_at(n, l.level, count) * l.real_out * double(l.taps) * l.cin;
            f += per * (i == 0 ? 2.0 : 3.0);
        }
        *flops_per_step = f;
    }
    return 0;
}

// --- single convolutions (kernel tests) -----------------------------------
int lbf_conv3d_forward(int cin, int cout, int taps, uint32_t side, uint32_t samples, const void* in,
                       uint32_t in_stride, const void* w, const float* bias, void* out, uint32_t out_stride, int relu,
                       int out_f32)
{
    if (!conv_shape_supported(cin, cout, taps) || side == 0 || (side & (side - 1)))
        return nfail(nullptr, "lbf_conv3d_forward: unsupported shape");
    uint32_t ls = 0;
    while ((1u << ls) < side) ++ls;
    const ConvGeom g{ls, samples << (3 * ls), in_stride, out_stride};
    NHIP(nullptr, launch_conv_forward(cin, cout, taps, g, static_cast<const __bf16*>(in),
                                      static_cast<const __bf16*>(w), bias, out, relu ? ACT_RELU : ACT_NONE,
                                      out_f32 != 0, nullptr));
    NHIP(nullptr, hipDeviceSynchronize());
    return 0;
}

int lbf_conv3d_wgrad(int cin, int cout, int taps, uint32_t side, uint32_t samples, const void* in,
                     uint32_t in_stride, const void* dz, uint32_t dz_stride, uint32_t splits, float* dw)
{
    if (!conv_shape_supported(cin, cout, taps) || side == 0 || (side & (side - 1)) || splits == 0)
        return nfail(nullptr, "lbf_conv3d_wgrad: unsupported shape");
    uint32_t ls = 0;
    while ((1u << ls) < side) ++ls;
    const ConvGeom g{ls, samples << (3 * ls), in_stride, dz_stride};
    float* partial = nullptr;
    const size_t nw = size_t(cout) * conv_kp(cin, taps);
    NHIP(nullptr, hipMalloc(&partial, nw * splits * 4));
    const hipError_t e = launch_conv_wgrad(cin, cout, taps, g, static_cast<const __bf16*>(in),
                                           static_cast<const __bf16*>(dz), partial, splits, dw, nullptr, nullptr,
                                           nullptr);
    const hipError_t e2 = hipDeviceSynchronize();
    (void)hipFree(partial);
    NHIP(nullptr, e);
    NHIP(nullptr, e2);
    return 0;
}

}  // extern "C"
