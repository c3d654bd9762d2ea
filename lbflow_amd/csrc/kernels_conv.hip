// kernels_conv.hip -- 3D convolution of the flow surrogate on BF16 MFMA.
//
// Periodic 3x3x3 (and 1x1x1) convolutions over channels-last BF16 volumes as
// implicit GEMMs on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//
//   forward / input gradient   out[v][co] = sum_k A[v][k] W[co][k],
//                              A[v][k] = in[nbr(v, tap(k))][ci(k)]
//   weight gradient            dW[co][k] = sum_v dz[v][co] A[v][k]
//
// MFMA operand layout (16x16x32): lane l holds row (l % 16) of A and column
// (l % 16) of B, each with the 8 consecutive k of group l / 16; the 16x16
// FP32 result gives lane l column (l % 16), rows 4 (l / 16) .. +3.  With
// channels-last activations and k = tap * Cin + ci, one lane's 8 k are 8
// consecutive channels of one voxel: a single 16-byte load.
//
// Forward: a workgroup of 4 waves covers 256 voxels x BN output channels;
// each wave owns 64 voxels (4 row tiles) x BN.  A is loaded straight from
// memory into MFMA registers (neighbour gather, 16 B per lane); the weight
// panel is staged through LDS (128 k per stage, rows padded by 16 B so the 16
// lanes of a group read 16 distinct bank quads).  The input gradient is the
// same kernel on the transposed, tap-mirrored weights.
//
// Weight gradient: the reduction runs over voxels, so both operands need 8
// consecutive voxels per lane; 128-voxel stages of dz and of the gathered
// input are fetched into registers one stage ahead, transposed into LDS
// ([channel][voxel] rows, padded to 272 B), and every voxel split writes its
// own FP32 partial, summed in split order by a second kernel (deterministic).
#include "conv_types.h"

namespace lbf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Periodic neighbour of voxel v at tap (dz,dy,dx) = (tap/9-1, tap/3%3-1, tap%3-1).
template <int TAPS>
__device__ __forceinline__ uint32_t nbr(uint32_t v, int tap, uint32_t ls)
{
    if constexpr (TAPS == 1) {
        return v;
    } else {
        const uint32_t m = (1u << ls) - 1u;
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const uint32_t x = (v + uint32_t(dx)) & m;
        const uint32_t y = ((v >> ls) + uint32_t(dy)) & m;
        const uint32_t z = ((v >> (2 * ls)) + uint32_t(dz)) & m;
        return (v & ~((1u << (3 * ls)) - 1u)) | (z << (2 * ls)) | (y << ls) | x;
    }
}

__device__ __forceinline__ bf16x8 bf16_zero()
{
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = __bf16(0.0f);
    return z;
}

// ---------------------------------------------------------------------------
// forward (and input gradient)
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TAPS, int BN, int ACT, bool BIAS, bool OUTF32>
__global__ void __launch_bounds__(256)
conv_fwd(ConvGeom g, const __bf16* __restrict__ in, const __bf16* __restrict__ w, const float* __restrict__ bias,
         void* __restrict__ out)
{
    constexpr int KP = conv_kp(CIN, TAPS);
    constexpr int KREAL = TAPS * CIN;
    constexpr int KC = 128;
    constexpr int ROW = KC + 8;
    constexpr int NT = BN / 16;
    static_assert(CIN % 8 == 0 && BN % 16 == 0, "channel counts must be multiples of 8 / 16");
    __shared__ __attribute__((aligned(16))) __bf16 sw[BN * ROW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const uint32_t vbase = blockIdx.x * 256u + uint32_t(wave) * 64u;
    const int co0 = blockIdx.y * BN;
    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t vox[4];
    bool live[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        vox[i] = vbase + 16u * i + r16;
        live[i] = vox[i] < g.nvox;
    }

    constexpr int NSTAGE = (KP + KC - 1) / KC;
    for (int st = 0; st < NSTAGE; ++st) {
        const int kc = st * KC;
        __syncthreads();
        for (int idx = threadIdx.x; idx < BN * (KC / 8); idx += 256) {
            const int row = idx / (KC / 8), c8 = idx % (KC / 8);
            const int k = kc + 8 * c8;
            bf16x8 val = bf16_zero();
            if (k < KP && co0 + row < COUT)
                val = *reinterpret_cast<const bf16x8*>(w + size_t(co0 + row) * KP + k);
            *reinterpret_cast<bf16x8*>(sw + row * ROW + 8 * c8) = val;
        }
        // gather the A operands of all K steps of this stage first, so that
        // their loads are in flight together and overlap the LDS sync
        bf16x8 a[KC / CONV_K][4];
#pragma unroll
        for (int s = 0; s < KC / CONV_K; ++s) {
            const int k = kc + CONV_K * s + 8 * grp;
            if (k < KREAL) {
                const int tap = k / CIN, ci = k - tap * CIN;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[s][i] = live[i] ? *reinterpret_cast<const bf16x8*>(in + size_t(nbr<TAPS>(vox[i], tap, g.log_s)) *
                                                                            g.in_stride + ci)
                                      : bf16_zero();
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) a[s][i] = bf16_zero();
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < KC / CONV_K; ++s) {
            if (kc + CONV_K * s >= KP) break;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(sw + (16 * j + r16) * ROW + CONV_K * s + 8 * grp);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s][i], b, acc[i][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int co = co0 + 16 * j + r16;
        if (co >= COUT) continue;
        const float bb = BIAS ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t v = vbase + 16u * i + 4u * grp + r;
                if (v >= g.nvox) continue;
                float x = acc[i][j][r] + bb;
                if (ACT == ACT_RELU) x = x > 0.0f ? x : 0.0f;
                if (OUTF32) static_cast<float*>(out)[size_t(v) * g.out_stride + co] = x;
                else static_cast<__bf16*>(out)[size_t(v) * g.out_stride + co] = __bf16(x);
            }
    }
}

// ---------------------------------------------------------------------------
// weight gradient
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TAPS, int BM, int BN>
__global__ void __launch_bounds__(256)
conv_wgrad(ConvGeom g, const __bf16* __restrict__ in, const __bf16* __restrict__ dz, float* __restrict__ partial,
           uint32_t vox_per_split)
{
    constexpr int KP = conv_kp(CIN, TAPS);
    constexpr int KREAL = TAPS * CIN;
    constexpr int VC = 128;                        // voxels per LDS stage (4 K steps)
    constexpr int ROW = VC + 8;                    // bf16 per LDS row: 272 B, 16 lanes hit 16 bank quads
    constexpr int WM = BM >= 32 ? 2 : 1, WN = 4 / WM;
    constexpr int MT = BM / WM / 16, NT = BN / WN / 16;
    constexpr int NA = (BM / 8) * VC / 256, NB = (BN / 8) * VC / 256;   // 16-byte loads per thread per stage
    static_assert(MT >= 1 && NT >= 1 && NA >= 1 && NB >= 1, "tile too small for the wave layout");
    __shared__ __attribute__((aligned(16))) __bf16 sa[BM * ROW];
    __shared__ __attribute__((aligned(16))) __bf16 sb[BN * ROW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const int wm = wave / WN, wn = wave % WN;
    const int k0 = blockIdx.x * BN, co0 = blockIdx.y * BM;
    const uint32_t v_begin = blockIdx.z * vox_per_split;
    const uint32_t v_end = v_begin + vox_per_split < g.nvox ? v_begin + vox_per_split : g.nvox;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // one stage's operands in registers: item idx = vv + VC * c8 (voxel vv, 8 channels c8)
    bf16x8 ra[NA], rb[NB];
    auto fetch = [&](uint32_t v0) {
#pragma unroll
        for (int n = 0; n < NA; ++n) {
            const int idx = threadIdx.x + 256 * n;
            const int vv = idx % VC, c8 = idx / VC;
            const uint32_t v = v0 + vv;
            ra[n] = (v < v_end && co0 + 8 * c8 < COUT)
                        ? *reinterpret_cast<const bf16x8*>(dz + size_t(v) * g.out_stride + co0 + 8 * c8)
                        : bf16_zero();
        }
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            const int idx = threadIdx.x + 256 * n;
            const int vv = idx % VC, k8 = idx / VC;
            const uint32_t v = v0 + vv;
            const int k = k0 + 8 * k8;
            bf16x8 val = bf16_zero();
            if (v < v_end && k < KREAL) {
                const int tap = k / CIN, ci = k - tap * CIN;
                val = *reinterpret_cast<const bf16x8*>(in + size_t(nbr<TAPS>(v, tap, g.log_s)) * g.in_stride + ci);
            }
            rb[n] = val;
        }
    };
    if (v_begin < v_end) fetch(v_begin);
    for (uint32_t v0 = v_begin; v0 < v_end; v0 += VC) {
        __syncthreads();   // the previous stage's MFMAs are done with the LDS tiles
        // transpose the fetched stage into [channel][voxel] rows
#pragma unroll
        for (int n = 0; n < NA; ++n) {
            const int idx = threadIdx.x + 256 * n;
            const int vv = idx % VC, c8 = idx / VC;
#pragma unroll
            for (int j = 0; j < 8; ++j) sa[(8 * c8 + j) * ROW + vv] = ra[n][j];
        }
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            const int idx = threadIdx.x + 256 * n;
            const int vv = idx % VC, k8 = idx / VC;
#pragma unroll
            for (int j = 0; j < 8; ++j) sb[(8 * k8 + j) * ROW + vv] = rb[n][j];
        }
        __syncthreads();
        if (v0 + VC < v_end) fetch(v0 + VC);   // next stage's loads overlap this stage's MFMAs
#pragma unroll
        for (int s = 0; s < VC / CONV_K; ++s) {
            bf16x8 a[MT], b[NT];
#pragma unroll
            for (int i = 0; i < MT; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(sa + (wm * (BM / WM) + 16 * i + r16) * ROW + CONV_K * s + 8 * grp);
#pragma unroll
            for (int j = 0; j < NT; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(sb + (wn * (BN / WN) + 16 * j + r16) * ROW + CONV_K * s + 8 * grp);
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int k = k0 + wn * (BN / WN) + 16 * j + r16;
            if (k >= KP) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + wm * (BM / WM) + 16 * i + 4 * grp + r;
                if (co < COUT) partial[(size_t(blockIdx.z) * COUT + co) * KP + k] = acc[i][j][r];
            }
        }
}

// dW[i] = sum over splits of partial[s][i], in split order.
__global__ void __launch_bounds__(256)
sum_splits(const float* __restrict__ partial, uint32_t n, uint32_t splits, float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < splits; ++k) s += partial[size_t(k) * n + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TAPS>
struct ConvLaunch {
    static constexpr int BN_FWD = COUT >= 64 ? 64 : COUT;
    static constexpr int BM_W = COUT >= 64 ? 64 : COUT;
    static constexpr int BN_W = 64;

    static void forward(const ConvGeom& g, const __bf16* in, const __bf16* w, const float* bias, void* out, int act,
                        bool out_f32, hipStream_t s)
    {
        const dim3 grid((g.nvox + 255u) / 256u, (COUT + BN_FWD - 1) / BN_FWD);
        if (out_f32 && act == ACT_RELU)
            hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT_RELU, true, true>), grid, dim3(256), 0, s, g,
                               in, w, bias, out);
        else if (out_f32)
            hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT_NONE, true, true>), grid, dim3(256), 0, s, g,
                               in, w, bias, out);
        else if (act == ACT_RELU)
            hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT_RELU, true, false>), grid, dim3(256), 0, s, g,
                               in, w, bias, out);
        else if (bias)
            hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT_NONE, true, false>), grid, dim3(256), 0, s, g,
                               in, w, bias, out);
        else
            hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT_NONE, false, false>), grid, dim3(256), 0, s, g,
                               in, w, bias, out);
    }

    static void wgrad(const ConvGeom& g, const __bf16* in, const __bf16* dz, float* partial, uint32_t splits,
                      float* dw, hipStream_t s)
    {
        constexpr int KP = conv_kp(CIN, TAPS);
        const uint32_t per = ((g.nvox + splits - 1) / splits + 127u) / 128u * 128u;
        const dim3 grid((KP + BN_W - 1) / BN_W, (COUT + BM_W - 1) / BM_W, splits);
        hipLaunchKernelGGL((conv_wgrad<CIN, COUT, TAPS, BM_W, BN_W>), grid, dim3(256), 0, s, g, in, dz, partial, per);
        const uint32_t n = uint32_t(COUT) * KP;
        hipLaunchKernelGGL(sum_splits, dim3((n + 255) / 256), dim3(256), 0, s, partial, n, splits, dw);
    }
};

// Runtime dispatch over the layer shapes the surrogate uses.
#define LBF_CONV_SHAPES(X) \
    X(8, 32, 27)           \
    X(32, 32, 27)          \
    X(32, 64, 27)          \
    X(64, 64, 27)          \
    X(64, 128, 27)         \
    X(128, 128, 27)        \
    X(128, 256, 27)        \
    X(256, 256, 27)        \
    X(384, 128, 27)        \
    X(192, 64, 27)         \
    X(96, 32, 27)          \
    X(32, 16, 1)           \
    X(16, 32, 1)           \
    X(32, 96, 27)          \
    X(64, 192, 27)         \
    X(128, 384, 27)        \
    X(256, 128, 27)        \
    X(128, 64, 27)         \
    X(64, 32, 27)

bool conv_shape_supported(int cin, int cout, int taps)
{
#define X(a, b, c) if (cin == a && cout == b && taps == c) return true;
    LBF_CONV_SHAPES(X)
#undef X
    return false;
}

hipError_t launch_conv_forward(int cin, int cout, int taps, const ConvGeom& g, const __bf16* in, const __bf16* w,
                               const float* bias, void* out, int act, bool out_f32, hipStream_t s)
{
#define X(a, b, c)                                                                       \
    if (cin == a && cout == b && taps == c) {                                            \
        ConvLaunch<a, b, c>::forward(g, in, w, bias, out, act, out_f32, s);              \
        return hipGetLastError();                                                        \
    }
    LBF_CONV_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_conv_wgrad(int cin, int cout, int taps, const ConvGeom& g, const __bf16* in, const __bf16* dz,
                             float* partial, uint32_t splits, float* dw, hipStream_t s)
{
#define X(a, b, c)                                                                       \
    if (cin == a && cout == b && taps == c) {                                            \
        ConvLaunch<a, b, c>::wgrad(g, in, dz, partial, splits, dw, s);                   \
        return hipGetLastError();                                                        \
    }
    LBF_CONV_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace lbf
