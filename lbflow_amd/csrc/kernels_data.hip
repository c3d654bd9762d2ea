// kernels_data.hip -- the warm-start network's training set on the device:
// per-channel sums of squares of the scaled steady fields (target scaling),
// the normalisation itself, and the minibatch gather that applies one of
// the 16 lattice symmetries on the way.
//
// A sample is a cubic crop of side 2^ls, voxels in z,y,x order: a pore
// mask byte and 4 floats (u_z, u_y, u_x, rho) per voxel.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbf {

constexpr uint32_t DATA_BLOCKS = 1024;   // partial sums of the scaling reduction

// raw target of channel c: u nu / F (c < 3) or (rho - 1) / F (c = 3), as
// FP64 from the FP32 field; the per-block sums of its square over pore
// voxels go to partial[block][4] (summed in block order on the host).
__global__ void __launch_bounds__(256) data_sumsq(const uint8_t* __restrict__ mask, const float* __restrict__ f,
                                                  uint64_t nvox, double nu_over_f, double inv_f,
                                                  double* __restrict__ partial)
{
    __shared__ double lds[256][4];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t v = uint64_t(blockIdx.x) * 256u + threadIdx.x; v < nvox; v += uint64_t(gridDim.x) * 256u) {
        if (!mask[v]) continue;
        const float4 x = reinterpret_cast<const float4*>(f)[v];
        const double r[4] = {double(x.x) * nu_over_f, double(x.y) * nu_over_f, double(x.z) * nu_over_f,
                             (double(x.w) - 1.0) * inv_f};
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += r[c] * r[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) lds[threadIdx.x][c] = acc[c];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < 4; ++c) lds[threadIdx.x][c] += lds[threadIdx.x + h][c];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[blockIdx.x * 4u + threadIdx.x] = lds[0][threadIdx.x];
}

// In place: t_c = raw_c / rms_c on pore voxels, 0 on solid ones, in FP32 as
// (u * (nu / F)) / rms and ((rho - 1) / F) / rms.
__global__ void __launch_bounds__(256) data_normalise(const uint8_t* __restrict__ mask, float* __restrict__ f,
                                                      uint64_t nvox, float nu_over_f, float force, float4 rms)
{
    for (uint64_t v = uint64_t(blockIdx.x) * 256u + threadIdx.x; v < nvox; v += uint64_t(gridDim.x) * 256u) {
        float4 x = reinterpret_cast<float4*>(f)[v];
        const float m = mask[v] ? 1.0f : 0.0f;
        x.x = (x.x * nu_over_f) / rms.x * m;
        x.y = (x.y * nu_over_f) / rms.y * m;
        x.z = (x.z * nu_over_f) / rms.z * m;
        x.w = ((x.w - 1.0f) / force) / rms.w * m;
        reinterpret_cast<float4*>(f)[v] = x;
    }
}

// Minibatch gather: output sample j is stored sample idx[j] mapped by
// symmetry `sym` (bit 0 flips z, bit 1 flips y, bit 3 flips x, bit 2 swaps
// y and z after the flips).  Output voxel (z, y, x) reads the source voxel
// (fz(y), fy(z), fx(x)) under the swap, (fz(z), fy(y), fx(x)) without it,
// f. flipping its axis when its bit is set.  Output channel c is source
// channel perm[c] times mul[c] (perm swaps u_z and u_y under bit 2; mul
// carries the signs and the RMS ratio of the swapped channels).
__device__ __forceinline__ void gather_voxel(const uint8_t* __restrict__ masks, const float* __restrict__ t,
                                             const uint32_t* __restrict__ idx, uint32_t ls, uint64_t v, uint32_t sym,
                                             float4 mul, uint8_t* __restrict__ mo, float* __restrict__ to)
{
    const uint32_t s = 1u << ls, m = s - 1u;
    const uint64_t j = v >> (3 * ls);
    const uint32_t local = uint32_t(v & ((uint64_t(1) << (3 * ls)) - 1u));
    const uint32_t x = local & m, y = (local >> ls) & m, z = local >> (2 * ls);
    uint32_t zs = (sym & 4u) ? y : z, ys = (sym & 4u) ? z : y, xs = x;
    if (sym & 1u) zs = m - zs;
    if (sym & 2u) ys = m - ys;
    if (sym & 8u) xs = m - xs;
    const uint64_t src = (uint64_t(idx[j]) << (3 * ls)) | (uint64_t(zs) << (2 * ls)) | (uint64_t(ys) << ls) | xs;
    mo[v] = masks[src];
    const float4 a = reinterpret_cast<const float4*>(t)[src];
    float4 o;
    if (sym & 4u) {
        o.x = a.y * mul.x;
        o.y = a.x * mul.y;
    } else {
        o.x = a.x * mul.x;
        o.y = a.y * mul.y;
    }
    o.z = a.z * mul.z;
    o.w = a.w * mul.w;
    reinterpret_cast<float4*>(to)[v] = o;
}

__global__ void __launch_bounds__(256) data_gather(const uint8_t* __restrict__ masks, const float* __restrict__ t,
                                                   const uint32_t* __restrict__ idx, uint32_t ls, uint64_t nvox,
                                                   uint32_t sym, float4 mul, uint8_t* __restrict__ mo,
                                                   float* __restrict__ to)
{
    for (uint64_t v = uint64_t(blockIdx.x) * 256u + threadIdx.x; v < nvox; v += uint64_t(gridDim.x) * 256u)
        gather_voxel(masks, t, idx, ls, v, sym, mul, mo, to);
}

// A launch's threads along x must stay below 2^32 (16384 crops of 64^3 are
// 2^32 voxels): grids are capped and the kernels loop over the rest.
constexpr uint64_t MAX_BLOCKS = 1u << 20;
static dim3 grid_of(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return dim3(uint32_t(b < MAX_BLOCKS ? b : MAX_BLOCKS));
}

hipError_t launch_data_sumsq(const uint8_t* mask, const float* f, uint64_t nvox, double nu_over_f, double inv_f,
                             double* partial, hipStream_t s)
{
    hipLaunchKernelGGL(data_sumsq, dim3(DATA_BLOCKS), dim3(256), 0, s, mask, f, nvox, nu_over_f, inv_f, partial);
    return hipGetLastError();
}

hipError_t launch_data_normalise(const uint8_t* mask, float* f, uint64_t nvox, float nu_over_f, float force,
                                 const float* rms, hipStream_t s)
{
    if (nvox == 0) return hipSuccess;
    hipLaunchKernelGGL(data_normalise, grid_of(nvox), dim3(256), 0, s, mask, f, nvox, nu_over_f, force,
                       make_float4(rms[0], rms[1], rms[2], rms[3]));
    return hipGetLastError();
}

hipError_t launch_data_gather(const uint8_t* masks, const float* t, const uint32_t* idx, uint32_t ls, uint32_t n,
                              uint32_t sym, const float* mul, uint8_t* mo, float* to, hipStream_t s)
{
    const uint64_t nvox = uint64_t(n) << (3 * ls);
    if (nvox == 0) return hipSuccess;
    hipLaunchKernelGGL(data_gather, grid_of(nvox), dim3(256), 0, s, masks, t, idx, ls, nvox, sym,
                       make_float4(mul[0], mul[1], mul[2], mul[3]), mo, to);
    return hipGetLastError();
}

}  // namespace lbf
