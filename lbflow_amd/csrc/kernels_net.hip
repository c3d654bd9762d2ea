// kernels_net.hip -- the memory-bound parts of the flow surrogate: input
// encoding, 2x average pooling and nearest upsampling (with their
// gradients), ReLU gradients, the weighted squared-error loss, bias
// gradients, Adam, and the transposed/tap-mirrored weight copy that turns
// the input gradient into a forward convolution.
//
// All activation tensors are channels-last BF16 with a row stride and a
// channel offset (a tensor may be a channel range of a concatenation
// buffer).  One thread handles 8 channels (16 bytes) of one voxel.
#include "conv_types.h"

namespace lbf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 ld8(const Chan& t, uint32_t v, uint32_t c)
{
    return *reinterpret_cast<const bf16x8*>(t.p + size_t(v) * t.stride + t.off + c);
}
__device__ __forceinline__ void st8(const Chan& t, uint32_t v, uint32_t c, const float* x)
{
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = __bf16(x[j]);
    *reinterpret_cast<bf16x8*>(t.p + size_t(v) * t.stride + t.off + c) = o;
}

// Voxel (z,y,x) of side 2^ls inside sample n  <->  parent (z/2,y/2,x/2) of side 2^(ls-1).
__device__ __forceinline__ uint32_t parent_of(uint32_t v, uint32_t ls)
{
    const uint32_t m = (1u << ls) - 1u;
    const uint32_t n = v >> (3 * ls);
    const uint32_t x = v & m, y = (v >> ls) & m, z = (v >> (2 * ls)) & m;
    const uint32_t lc = ls - 1;
    return (n << (3 * lc)) | ((z >> 1) << (2 * lc)) | ((y >> 1) << lc) | (x >> 1);
}
__device__ __forceinline__ uint32_t child_of(uint32_t V, uint32_t lc, int j)
{
    const uint32_t m = (1u << lc) - 1u;
    const uint32_t n = V >> (3 * lc);
    const uint32_t X = V & m, Y = (V >> lc) & m, Z = (V >> (2 * lc)) & m;
    const uint32_t lf = lc + 1;
    const uint32_t z = 2 * Z + ((j >> 2) & 1), y = 2 * Y + ((j >> 1) & 1), x = 2 * X + (j & 1);
    return (n << (3 * lf)) | (z << (2 * lf)) | (y << lf) | x;
}

// x0[v] = (fluid, 0 x 7)
__global__ void net_input(const uint8_t* __restrict__ mask, uint32_t nvox, __bf16* __restrict__ x0)
{
    const uint32_t v = blockIdx.x * 256u + threadIdx.x;
    if (v >= nvox) return;
    float x[8] = {mask[v] ? 1.0f : 0.0f, 0, 0, 0, 0, 0, 0, 0};
    st8(Chan{x0, 8, 0}, v, 0, x);
}

// out[V][c] = mean of the 8 children of V.  ls = log2 side of the output.
__global__ void net_pool(Chan in, Chan out, uint32_t C, uint32_t ls, uint32_t nvox_out)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t V = t / (C / 8), c = (t % (C / 8)) * 8;
    if (V >= nvox_out) return;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 8; ++j) {
        const bf16x8 a = ld8(in, child_of(V, ls, j), c);
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += float(a[k]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] *= 0.125f;
    st8(out, V, c, s);
}

// out[v][c] = in[parent(v)][c].  ls = log2 side of the output.
__global__ void net_upsample(Chan in, Chan out, uint32_t C, uint32_t ls, uint32_t nvox_out)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t v = t / (C / 8), c = (t % (C / 8)) * 8;
    if (v >= nvox_out) return;
    *reinterpret_cast<bf16x8*>(out.p + size_t(v) * out.stride + out.off + c) = ld8(in, parent_of(v, ls), c);
}

// Encoder output feeding both a pooling and a skip connection:
// dz[v][c] = (dpool[parent(v)][c] / 8 + dskip[v][c]) * (y[v][c] > 0).  ls = log2 side of v.
__global__ void net_enc_grad(Chan dpool, Chan dskip, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t v = t / (C / 8), c = (t % (C / 8)) * 8;
    if (v >= nvox) return;
    const bf16x8 a = ld8(dpool, parent_of(v, ls), c), b = ld8(dskip, v, c), yy = ld8(y, v, c);
    float o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = float(yy[k]) > 0.0f ? 0.125f * float(a[k]) + float(b[k]) : 0.0f;
    st8(dz, v, c, o);
}

// Output that was upsampled: dz[V][c] = (sum of dup over the 8 children) * (y[V][c] > 0).
// ls = log2 side of V (the coarse level).
__global__ void net_up_grad(Chan dup, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t V = t / (C / 8), c = (t % (C / 8)) * 8;
    if (V >= nvox) return;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 8; ++j) {
        const bf16x8 a = ld8(dup, child_of(V, ls, j), c);
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += float(a[k]);
    }
    const bf16x8 yy = ld8(y, V, c);
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = float(yy[k]) > 0.0f ? s[k] : 0.0f;
    st8(dz, V, c, s);
}

// Loss sum_v sum_c<4 w_c (out - target)^2 / nvox and its gradient into the
// 16-channel head gradient (channels 4..15 get 0).  Per-block FP64 partials.
__global__ void __launch_bounds__(256)
net_loss(const float* __restrict__ out, const float* __restrict__ target, float w0, float w1, float w2, float w3,
         uint32_t nvox, float scale, __bf16* __restrict__ dout, double* __restrict__ partial)
{
    __shared__ double red[4];
    const uint32_t v = blockIdx.x * 256u + threadIdx.x;
    double l = 0.0;
    if (v < nvox) {
        const float w[4] = {w0, w1, w2, w3};
        float g[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) g[c] = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float d = out[size_t(v) * 16 + c] - target[size_t(v) * 4 + c];
            l += double(w[c]) * double(d) * double(d);
            g[c] = 2.0f * w[c] * d * scale;
        }
        bf16x8 a, b;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a[k] = __bf16(g[k]);
            b[k] = __bf16(g[8 + k]);
        }
        *reinterpret_cast<bf16x8*>(dout + size_t(v) * 16) = a;
        *reinterpret_cast<bf16x8*>(dout + size_t(v) * 16 + 8) = b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Fixed-order sum of n doubles by one workgroup.
__global__ void __launch_bounds__(256)
net_sum(const double* __restrict__ partial, uint32_t n, double* __restrict__ out)
{
    __shared__ double red[4];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += partial[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// Per-block channel sums of dz over `per` voxels: partial[block][c].  The 256
// threads split into 256/C voxel lanes per channel (C <= 256, C | 256),
// reduced through LDS in a fixed order.
__global__ void __launch_bounds__(256)
net_bias_partial(Chan dz, uint32_t C, uint32_t nvox, uint32_t per, float* __restrict__ partial)
{
    __shared__ float red[256];
    const uint32_t c = threadIdx.x % C, lanes = 256u / C, g = threadIdx.x / C;
    const uint32_t v0 = blockIdx.x * per;
    const uint32_t v1 = v0 + per < nvox ? v0 + per : nvox;
    float s = 0.0f;
    for (uint32_t v = v0 + g; v < v1; v += lanes) s += float(dz.p[size_t(v) * dz.stride + dz.off + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (uint32_t k = 0; k < lanes; ++k) t += red[k * C + threadIdx.x];
        partial[size_t(blockIdx.x) * C + threadIdx.x] = t;
    }
}

// Same sum with 16-byte loads of 8 channels (C, the channel offset and the
// row stride multiples of 8): thread t reduces channel group t % (C/8) over
// the voxel lanes t / (C/8), then the lanes are combined through LDS in a
// fixed order.
__global__ void __launch_bounds__(256)
net_bias_partial8(Chan dz, uint32_t C, uint32_t nvox, uint32_t per, float* __restrict__ partial)
{
    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
    __shared__ float red[256 * 8];
    const uint32_t groups = C / 8u, c8 = threadIdx.x % groups, lanes = 256u / groups, g = threadIdx.x / groups;
    const uint32_t v0 = blockIdx.x * per;
    const uint32_t v1 = v0 + per < nvox ? v0 + per : nvox;
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t v = v0 + g; v < v1; v += lanes) {
        const bf16x8 x = *reinterpret_cast<const bf16x8*>(dz.p + size_t(v) * dz.stride + dz.off + 8 * c8);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] += float(x[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[(g * groups + c8) * 8 + j] = s[j];
    __syncthreads();
    if (threadIdx.x < C) {
        const uint32_t c = threadIdx.x;
        float t = 0.0f;
        for (uint32_t k = 0; k < lanes; ++k) t += red[(k * groups + c / 8) * 8 + c % 8];
        partial[size_t(blockIdx.x) * C + c] = t;
    }
}

// db[c] = sum over blocks of partial[b][c]: one workgroup per channel, strided
// partial sums then a fixed-order tree.
__global__ void __launch_bounds__(256)
net_bias_final(const float* __restrict__ partial, uint32_t C, uint32_t blocks, float* __restrict__ db)
{
    __shared__ float red[256];
    const uint32_t c = blockIdx.x;
    float s = 0.0f;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256) s += partial[size_t(b) * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// Adam (bias-corrected) on FP32 master parameters; optionally refreshes the BF16 copy.
__global__ void __launch_bounds__(256)
net_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g, uint32_t n,
         float lr, float b1, float b2, float eps, float c1, float c2, __bf16* __restrict__ pb)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = __bf16(pi);
}

// out4[v][c] = out16[v][c], c < 4 (the head's real channels).
__global__ void net_head4(const float* __restrict__ out16, uint32_t nvox, float* __restrict__ out4)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= nvox * 4u) return;
    out4[i] = out16[size_t(i >> 2) * 16 + (i & 3u)];
}

// Physical fields of a prediction: out4[v][c] = out16[v][c] scale[c] +
// offset[c], the velocity channels (c < 3) zero where mask[v] is 0 (solid).
__global__ void net_head_fields(const float* __restrict__ out16, const uint8_t* __restrict__ mask, uint32_t nvox,
                                float4 scale, float4 offset, float* __restrict__ out4)
{
    const uint32_t v = blockIdx.x * 256u + threadIdx.x;
    if (v >= nvox) return;
    const float* p = out16 + size_t(v) * 16;
    const float m = mask[v] ? 1.0f : 0.0f;
    float4 r;
    r.x = m * (p[0] * scale.x + offset.x);
    r.y = m * (p[1] * scale.y + offset.y);
    r.z = m * (p[2] * scale.z + offset.z);
    r.w = p[3] * scale.w + offset.w;
    reinterpret_cast<float4*>(out4)[v] = r;
}

// FP32 -> BF16 copy.
__global__ void net_to_bf16(const float* __restrict__ p, uint32_t n, __bf16* __restrict__ pb)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) pb[i] = __bf16(p[i]);
}

// wt[ci][tap' * cout + co] = w[co][tap * cin + ci], tap' = taps - 1 - tap (mirrored
// offset), both zero-padded to their KP.
__global__ void net_transpose_w(const __bf16* __restrict__ w, uint32_t cin, uint32_t cout, uint32_t taps,
                                uint32_t kp, uint32_t kpt, __bf16* __restrict__ wt)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;   // index into wt
    if (i >= cin * kpt) return;
    const uint32_t ci = i / kpt, k = i % kpt;
    __bf16 val = __bf16(0.0f);
    if (k < taps * cout) {
        const uint32_t tp = k / cout, co = k % cout;
        const uint32_t tap = taps - 1 - tp;
        val = w[size_t(co) * kp + tap * cin + ci];
    }
    wt[i] = val;
}

static inline dim3 grid_for(uint64_t threads) { return dim3(uint32_t((threads + 255) / 256)); }

hipError_t launch_net_input(const uint8_t* mask, uint32_t nvox, __bf16* x0, hipStream_t s)
{
    hipLaunchKernelGGL(net_input, grid_for(nvox), dim3(256), 0, s, mask, nvox, x0);
    return hipGetLastError();
}
hipError_t launch_net_pool(Chan in, Chan out, uint32_t C, uint32_t ls_out, uint32_t nvox_out, hipStream_t s)
{
    hipLaunchKernelGGL(net_pool, grid_for(uint64_t(nvox_out) * (C / 8)), dim3(256), 0, s, in, out, C, ls_out, nvox_out);
    return hipGetLastError();
}
hipError_t launch_net_upsample(Chan in, Chan out, uint32_t C, uint32_t ls_out, uint32_t nvox_out, hipStream_t s)
{
    hipLaunchKernelGGL(net_upsample, grid_for(uint64_t(nvox_out) * (C / 8)), dim3(256), 0, s, in, out, C, ls_out,
                       nvox_out);
    return hipGetLastError();
}
hipError_t launch_net_enc_grad(Chan dpool, Chan dskip, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox,
                               hipStream_t s)
{
    hipLaunchKernelGGL(net_enc_grad, grid_for(uint64_t(nvox) * (C / 8)), dim3(256), 0, s, dpool, dskip, y, dz, C, ls,
                       nvox);
    return hipGetLastError();
}
hipError_t launch_net_up_grad(Chan dup, Chan y, Chan dz, uint32_t C, uint32_t ls, uint32_t nvox, hipStream_t s)
{
    hipLaunchKernelGGL(net_up_grad, grid_for(uint64_t(nvox) * (C / 8)), dim3(256), 0, s, dup, y, dz, C, ls, nvox);
    return hipGetLastError();
}
hipError_t launch_net_loss(const float* out, const float* target, const float* w, uint32_t nvox, float scale,
                           __bf16* dout, double* partial, double* loss, hipStream_t s)
{
    const dim3 grid = grid_for(nvox);
    hipLaunchKernelGGL(net_loss, grid, dim3(256), 0, s, out, target, w[0], w[1], w[2], w[3], nvox, scale, dout,
                       partial);
    hipLaunchKernelGGL(net_sum, dim3(1), dim3(256), 0, s, partial, grid.x, loss);
    return hipGetLastError();
}
hipError_t launch_net_bias_grad(Chan dz, uint32_t C, uint32_t nvox, float* partial, uint32_t blocks, float* db,
                                hipStream_t s)
{
    const uint32_t per = (nvox + blocks - 1) / blocks;
    if (C % 8 == 0 && 256 % (C / 8) == 0 && dz.off % 8 == 0 && dz.stride % 8 == 0)
        hipLaunchKernelGGL(net_bias_partial8, dim3(blocks), dim3(256), 0, s, dz, C, nvox, per, partial);
    else
        hipLaunchKernelGGL(net_bias_partial, dim3(blocks), dim3(256), 0, s, dz, C, nvox, per, partial);
    hipLaunchKernelGGL(net_bias_final, dim3(C), dim3(256), 0, s, partial, C, blocks, db);
    return hipGetLastError();
}
hipError_t launch_net_adam(float* p, float* m, float* v, const float* g, uint32_t n, float lr, float b1, float b2,
                           float eps, float c1, float c2, __bf16* pb, hipStream_t s)
{
    hipLaunchKernelGGL(net_adam, grid_for(n), dim3(256), 0, s, p, m, v, g, n, lr, b1, b2, eps, c1, c2, pb);
    return hipGetLastError();
}
hipError_t launch_net_head4(const float* out16, uint32_t nvox, float* out4, hipStream_t s)
{
    hipLaunchKernelGGL(net_head4, grid_for(uint64_t(nvox) * 4), dim3(256), 0, s, out16, nvox, out4);
    return hipGetLastError();
}
hipError_t launch_net_head_fields(const float* out16, const uint8_t* mask, uint32_t nvox, const float* scale,
                                  const float* offset, float* out4, hipStream_t s)
{
    const float4 sc = make_float4(scale[0], scale[1], scale[2], scale[3]);
    const float4 of = make_float4(offset[0], offset[1], offset[2], offset[3]);
    hipLaunchKernelGGL(net_head_fields, grid_for(nvox), dim3(256), 0, s, out16, mask, nvox, sc, of, out4);
    return hipGetLastError();
}
hipError_t launch_net_to_bf16(const float* p, uint32_t n, __bf16* pb, hipStream_t s)
{
    hipLaunchKernelGGL(net_to_bf16, grid_for(n), dim3(256), 0, s, p, n, pb);
    return hipGetLastError();
}
hipError_t launch_net_transpose_w(const __bf16* w, uint32_t cin, uint32_t cout, uint32_t taps, uint32_t kp,
                                  uint32_t kpt, __bf16* wt, hipStream_t s)
{
    hipLaunchKernelGGL(net_transpose_w, grid_for(uint64_t(cin) * kpt), dim3(256), 0, s, w, cin, cout, taps, kp, kpt,
                       wt);
    return hipGetLastError();
}

}  // namespace lbf
