// kernels_reduce.hip -- device reductions of a report step.
//
// Replaces the reference's report sums and their MPI_Allreduce
// (solver/src/lbm/scsp.cpp:325-341, colorfluid.cpp:841-867) and the
// convergence sums of Simulator::updateConvergence (simulator.cpp:263-312).
// FP64 accumulators; wavefront (64-lane) shuffles, then LDS across the
// block's waves, then a single-block pass over the block partials in a
// fixed order, so the result is deterministic run to run.  Summation order
// differs from the reference's sequential loop, so sums agree with it to
// rounding (about 1e-15 relative), not bit for bit.
#include <mutex>
#include <vector>

#include "device.h"

namespace lbf {

// At most this many partial blocks per report: the block partials of any
// lattice fit [REPORT_MAX_BLOCKS][NREP] doubles.
constexpr uint32_t REPORT_MAX_BLOCKS = 4096;

// Device memory of lbf_report_static: the block partials followed by the
// final sums, [REPORT_MAX_BLOCKS][NREP] + [NREP] doubles (512 KiB), one copy
// per GPU for the life of the process.  Host code never uses this name as an
// address: report_scratch() looks up each GPU's copy.
constexpr uint64_t REPORT_SCRATCH_DOUBLES = uint64_t(REPORT_MAX_BLOCKS) * NREP + NREP;
static __device__ __align__(256) double scratch[REPORT_SCRATCH_DOUBLES];

enum Slot {
    S_VEL = 0, S_RHO_R0, S_RHO_R1, S_FIRST0, S_FIRST1, S_LAST0, S_LAST1, S_CNT_FIRST, S_CNT_LAST,
    S_DU2, S_U2, S_COUNT, S_DA2, S_CHI_PREV, S_CHI_ADJ, S_NONFINITE
};
static_assert(S_NONFINITE < NREP, "slots");

__device__ __forceinline__ bool is_max_slot(int k) { return k == S_CHI_PREV || k == S_CHI_ADJ; }

__device__ __forceinline__ double wave_reduce(double v, bool maxop)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(v, off, 64);
        v = maxop ? fmax(v, o) : v + o;
    }
    return v;
}

__global__ void __launch_bounds__(256) report_partial(ReportArgs a)
{
    __shared__ double lds[4][NREP];
    double acc[NREP];
#pragma unroll
    for (int k = 0; k < NREP; ++k) acc[k] = 0.0;
    const uint64_t N = a.g.N;
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < N; i += stride) {
        const uint32_t desc = a.desc[i];
        const int kind = a.lab[desc >> DESC_LABEL_SHIFT].kind;
        const bool space = kind & KIND_FLUID;
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        // coordinate along the report axis (reference order, slowest first:
        // z y x in 3D, y x in 2D), selected without a runtime-indexed array
        const int ax = a.D == 3 ? a.axis : a.axis + 1;
        const int along = ax == 0 ? z : (ax == 1 ? y : x);
        const double r0 = a.rho0[i];
        const double r1 = a.color ? a.rho1[i] : 0.0;
        if (space) {
            if (uint32_t(along) >= a.first && uint32_t(along) <= a.last) {
                acc[S_VEL] += a.u[uint64_t(a.axis) * N + i];
                acc[S_RHO_R0] += r0;
                acc[S_RHO_R1] += r1;
            }
            if (uint32_t(along) == a.first) { acc[S_FIRST0] += r0; acc[S_FIRST1] += r1; acc[S_CNT_FIRST] += 1.0; }
            if (uint32_t(along) == a.last) { acc[S_LAST0] += r0; acc[S_LAST1] += r1; acc[S_CNT_LAST] += 1.0; }
        }
        if (a.conv) {
            for (int d = 0; d < a.D; ++d) {
                const uint64_t k = uint64_t(d) * N + i;
                const double v = a.u[k];
                if (!isfinite(v)) acc[S_NONFINITE] += 1.0;
                const double delta = v - a.uprev[k];
                acc[S_DU2] += delta * delta;
                acc[S_U2] += v * v;
                acc[S_COUNT] += 1.0;
                if (a.adjacent) {
                    // single-phase PB/VB sites: the reference copies the adjacent
                    // velocity after the boundary has set this step's value, which
                    // calcMacroVar keeps (scsp.cpp:79-83); colour PB/VB sites carry
                    // their boundary velocity in the history buffer itself
                    const double adj = ((kind & KIND_ZOUHE) && !a.color) ? v : a.uadj[k];
                    if (!isfinite(adj)) acc[S_NONFINITE] += 1.0;
                    acc[S_DA2] += (v - adj) * (v - adj);
                }
                a.uprev[k] = v;
            }
            if (a.color) {
                const double rt = r0 + r1;
                const double chi = rt > 0.0 ? (r0 - r1) / rt : 0.0;
                if (!isfinite(chi)) acc[S_NONFINITE] += 1.0;
                acc[S_CHI_PREV] = fmax(acc[S_CHI_PREV], fabs(chi - a.chiprev[i]));
                if (a.adjacent) {
                    const double ca = a.chiadj[i];
                    if (!isfinite(ca)) acc[S_NONFINITE] += 1.0;
                    acc[S_CHI_ADJ] = fmax(acc[S_CHI_ADJ], fabs(chi - ca));
                }
                a.chiprev[i] = chi;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NREP; ++k) {
        const double v = wave_reduce(acc[k], is_max_slot(k));
        if (lane == 0) lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NREP) {
        const int k = threadIdx.x;
        double v = lds[0][k];
        for (int w = 1; w < 4; ++w) v = is_max_slot(k) ? fmax(v, lds[w][k]) : v + lds[w][k];
        a.partial[uint64_t(blockIdx.x) * NREP + k] = v;
    }
}

__global__ void __launch_bounds__(256) report_final(const double* partial, uint32_t blocks, double* out)
{
    __shared__ double lds[4][NREP];
    double acc[NREP];
#pragma unroll
    for (int k = 0; k < NREP; ++k) acc[k] = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256) {
#pragma unroll
        for (int k = 0; k < NREP; ++k) {
            const double v = partial[uint64_t(b) * NREP + k];
            acc[k] = is_max_slot(k) ? fmax(acc[k], v) : acc[k] + v;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NREP; ++k) {
        const double v = wave_reduce(acc[k], is_max_slot(k));
        if (lane == 0) lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NREP) {
        const int k = threadIdx.x;
        double v = lds[0][k];
        for (int w = 1; w < 4; ++w) v = is_max_slot(k) ? fmax(v, lds[w][k]) : v + lds[w][k];
        out[k] = v;
    }
}

uint32_t report_blocks(uint64_t N)
{
    const uint64_t want = (N + 255) / 256;
    return uint32_t(want < REPORT_MAX_BLOCKS ? want : REPORT_MAX_BLOCKS);
}

// Device address of the current GPU's copy of `scratch` and its size in
// doubles.  The address is looked up once per GPU (hipGetSymbolAddress with
// that GPU current) and kept.
hipError_t report_scratch(double** ptr, uint64_t* capacity)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    static std::mutex lock;
    static std::vector<double*> address;
    std::lock_guard<std::mutex> hold(lock);
    if (size_t(dev) >= address.size()) address.resize(size_t(dev) + 1, nullptr);
    if (!address[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        address[dev] = static_cast<double*>(p);
    }
    *ptr = address[dev];
    *capacity = REPORT_SCRATCH_DOUBLES;
    return hipSuccess;
}

hipError_t launch_report(const ReportArgs& a, double* out, hipStream_t s)
{
    const uint32_t blocks = report_blocks(a.g.N);
    hipLaunchKernelGGL(report_partial, dim3(blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(report_final, dim3(1), dim3(256), 0, s, a.partial, blocks, out);
    return hipGetLastError();
}

// Convert FP32 shifted populations to FP64 f = h + shift_q (checkpoint layout).
__global__ void __launch_bounds__(256) unshift_f32(const float* h, double* f, const double* shift,
                                                   uint64_t N, uint32_t Q, uint32_t nplanes)
{
    const uint64_t total = N * nplanes;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < total; k += uint64_t(gridDim.x) * 256u) {
        const uint32_t plane = uint32_t(k / N);
        f[k] = double(h[k]) + shift[plane];
    }
}

__global__ void __launch_bounds__(256) shift_f64(const double* f, float* h, const double* shift,
                                                 uint64_t N, uint32_t Q, uint32_t nplanes)
{
    const uint64_t total = N * nplanes;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < total; k += uint64_t(gridDim.x) * 256u) {
        const uint32_t plane = uint32_t(k / N);
        h[k] = float(f[k] - shift[plane]);
    }
}

hipError_t launch_unshift(const float* h, double* f, const double* shift, uint64_t N, uint32_t Q,
                          uint32_t nplanes, hipStream_t s)
{
    hipLaunchKernelGGL(unshift_f32, dim3(2048), dim3(256), 0, s, h, f, shift, N, Q, nplanes);
    return hipGetLastError();
}

hipError_t launch_shift(const double* f, float* h, const double* shift, uint64_t N, uint32_t Q,
                        uint32_t nplanes, hipStream_t s)
{
    hipLaunchKernelGGL(shift_f64, dim3(2048), dim3(256), 0, s, f, h, shift, N, Q, nplanes);
    return hipGetLastError();
}

}  // namespace lbf

namespace lbf {

// Plain streaming copy, 16 bytes per lane per iteration: the achievable HBM
// bandwidth reference for roofline fractions and counter calibration.
__global__ void __launch_bounds__(256) copy_stream(const float4* __restrict__ src, float4* __restrict__ dst,
                                                  uint64_t n)
{
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < n; k += uint64_t(gridDim.x) * 256u)
        dst[k] = src[k];
}

// One 16-byte element per thread over the whole array (a launch of n / 256
// workgroups): on the MI355X this reaches about 6.2 TB/s for 1 GiB, where a
// grid-stride loop over 65536 workgroups reaches about 5.4 TB/s.
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s)
{
    const uint64_t n = bytes / 16;
    const uint64_t want = (n + 255) / 256;
    const uint32_t blocks = uint32_t(want < (uint64_t(1) << 31) ? want : (uint64_t(1) << 31));
    hipLaunchKernelGGL(copy_stream, dim3(blocks), dim3(256), 0, s, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n);
    return hipGetLastError();
}

}  // namespace lbf
