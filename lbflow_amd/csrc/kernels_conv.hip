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
// panel is staged through LDS (128 k per stage, rows padded by 32 B: 72
// dwords, conflict-free, see LDS_NOTE).  The input gradient is the
// same kernel on the transposed, tap-mirrored weights.
//
// Weight gradient: the reduction runs over voxels, so both operands need 8
// consecutive voxels per lane; the dz and input tiles are staged in LDS in
// their global layout (voxel rows of channels) and read transposed with
// ds_read_b64_tr_b16, and every voxel split writes its own FP32 partial,
// summed in split order by a second kernel (deterministic).
#include <type_traits>

#include "conv_types.h"
#include "device.h"

// LDS_NOTE.  Every MFMA operand read from LDS is a 16-byte ds_read_b128 in
// which the 16 lanes of a row group read 16 rows (voxels or channels) at a
// fixed row stride and the 4 groups sit 16 bytes apart.  Measured on the
// MI355X with SQ_LDS_BANK_CONFLICT: a row stride of 4 (mod 8) dwords (80 B,
// 112 B, 528 B, ...) costs about 4 conflict cycles per read, 0 (mod 16)
// dwords more; strides of 8 (mod 16) dwords (96 B, 288 B, 544 B, ...) and
// 2 or 6 (mod 8) are conflict-free.  The tiles below pad their rows to 8
// (mod 16) dwords, keeping 16-byte alignment.

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
    constexpr int ROW = KC + 16;                   // 72 dwords per LDS row (see LDS_NOTE)
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
                if (ACT == ACT_RELU_GATE) x = float(g.gate[size_t(v) * g.gate_stride + co]) > 0.0f ? x : 0.0f;
                if (OUTF32) static_cast<float*>(out)[size_t(v) * g.out_stride + co] = x;
                else static_cast<__bf16*>(out)[size_t(v) * g.out_stride + co] = __bf16(x);
            }
    }
}

// ---------------------------------------------------------------------------
// forward (and input gradient), 3x3x3 with Cin a multiple of 32: input tile
// with its one-voxel halo staged in LDS once per 32-channel chunk and reused
// by all 27 taps (the direct kernel above gathers every tap from L2).
// A workgroup covers a BZ x BY x BX block of 256 voxels (BX = min(side, 32));
// the halo tile is (BZ+2)(BY+2)(BX+2) voxels x 32 channels, 96 B per voxel
// (conflict-free reads, see LDS_NOTE).  Each
// wave owns 64 voxels (4 row tiles) x BN output channels; the weight
// fragments come from global memory (the [Cout][KP] panel of a layer is
// shared by every workgroup and stays in the caches).
// ---------------------------------------------------------------------------
constexpr int HALO_VS = 48;          // bf16 per halo voxel in LDS: 24 dwords (see LDS_NOTE)
constexpr int HALO_MAX = 34 * 6 * 4; // largest halo tile (side >= 32: 32 x 4 x 2 block)

struct HaloShape {
    int lbx, lby, lbz;               // log2 of the block extent along x, y, z
    int hx, hy, hz;                  // halo tile extents
};

__device__ __forceinline__ HaloShape halo_shape(int ls)
{
    HaloShape h;
    h.lbx = ls < 5 ? ls : 5;
    const int rem = 8 - h.lbx;
    h.lby = ls < (rem + 1) / 2 ? ls : (rem + 1) / 2;
    h.lbz = rem - h.lby;
    h.hx = (1 << h.lbx) + 2;
    h.hy = (1 << h.lby) + 2;
    h.hz = (1 << h.lbz) + 2;
    return h;
}

template <int CIN, int COUT, int BN, int ACT, bool BIAS, bool OUTF32>
__global__ void __launch_bounds__(256)
conv_fwd_halo(ConvGeom g, const __bf16* __restrict__ in, const __bf16* __restrict__ w,
              const float* __restrict__ bias, void* __restrict__ out)
{
    constexpr int KP = conv_kp(CIN, 27);
    constexpr int NT = BN / 16;
    constexpr int CC = 32;
    static_assert(CIN % CC == 0 && BN % 16 == 0, "Cin must be a multiple of 32");
    __shared__ __attribute__((aligned(16))) __bf16 tile[HALO_MAX * HALO_VS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const int ls = int(g.log_s);
    const uint32_t S = 1u << ls, m = S - 1u;
    const HaloShape h = halo_shape(ls);
    const int hv_n = h.hx * h.hy * h.hz;
    // block -> sample and block origin
    const uint32_t bps = 1u << (3 * ls - 8);                    // blocks per sample
    const uint32_t smp = blockIdx.x / bps, b = blockIdx.x % bps;
    const uint32_t nbx = S >> h.lbx, nby = S >> h.lby;
    const uint32_t x0 = (b % nbx) << h.lbx, y0 = ((b / nbx) % nby) << h.lby, z0 = (b / (nbx * nby)) << h.lbz;
    const size_t sbase = size_t(smp) << (3 * ls);
    const int co0 = blockIdx.y * BN;
    // this lane's A voxels (4 row tiles) as halo-tile indices
    int hv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lv = wave * 64 + 16 * i + r16;
        const int lx = lv & ((1 << h.lbx) - 1), ly = (lv >> h.lbx) & ((1 << h.lby) - 1), lz = lv >> (h.lbx + h.lby);
        hv[i] = ((lz + 1) * h.hy + (ly + 1)) * h.hx + (lx + 1);
    }
    const int oy = h.hx, oz = h.hx * h.hy;
    // the lane's four A rows as LDS pointers at tap (0, 0, 0); a tap adds a
    // uniform offset
    const __bf16* arow[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) arow[i] = tile + hv[i] * HALO_VS + 8 * grp;
    // staging: thread t copies 8-channel group q of halo voxels t/4, t/4 + 64,
    // ...; its first voxel's halo coordinates once, then steps of 64 voxels
    // as (+ d64x, + d64y) with carries, no division per voxel (d64y + 1 <= hy
    // for every side 8 .. 256, so one carry into z suffices); the voxel's
    // offset inside its sample fits 32 bits
    const int sq = threadIdx.x & 3, sv0 = threadIdx.x >> 2;
    const int sx0 = sv0 % h.hx, sy0 = (sv0 / h.hx) % h.hy, sz0 = sv0 / (h.hx * h.hy);
    const int d64x = 64 % h.hx, d64y = 64 / h.hx;
    const __bf16* in_s = in + sbase * g.in_stride + 8 * sq;
    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const __bf16* wrow[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) wrow[j] = w + size_t(co0 + 16 * j + r16) * KP + 8 * grp;

    // split over input channels (gridDim.z > 1): this workgroup's chunks
    const int cps = (CIN / CC) / int(gridDim.z);
    const int c_begin = int(blockIdx.z) * cps * CC, c_end = c_begin + cps * CC;
    for (int c0 = c_begin; c0 < c_end; c0 += CC) {
        __syncthreads();   // the previous chunk's MFMAs are done with the tile
        {
            // 4 voxels per trip: their 4 loads are in flight together
            const __bf16* src = in_s + c0;
            int hx = sx0, hy = sy0, hz = sz0;
            for (int v = sv0; v < hv_n; v += 256) {
                bf16x8 r[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (v + 64 * u < hv_n) {
                        const uint32_t x = (x0 + uint32_t(hx) - 1u) & m, y = (y0 + uint32_t(hy) - 1u) & m,
                                       z = (z0 + uint32_t(hz) - 1u) & m;
                        const uint32_t lv = (z << (2 * ls)) | (y << ls) | x;
                        r[u] = *reinterpret_cast<const bf16x8*>(src + uint64_t(lv) * g.in_stride);
                    }
                    hx += d64x;
                    hy += d64y;
                    if (hx >= h.hx) {
                        hx -= h.hx;
                        ++hy;
                    }
                    if (hy >= h.hy) {
                        hy -= h.hy;
                        ++hz;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (v + 64 * u < hv_n) *reinterpret_cast<bf16x8*>(tile + (v + 64 * u) * HALO_VS + 8 * sq) = r[u];
            }
        }
        __syncthreads();
        bf16x8 bnext[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) bnext[j] = *reinterpret_cast<const bf16x8*>(wrow[j] + c0);
        // 9 (dz, dy) rows of 3 taps, 3 rows per loop trip (a trip per row
        // rotates the accumulators through AGPR moves at the back edge);
        // the next tap's weight fragments are loaded one tap ahead
#pragma unroll 3
        for (int t9 = 0; t9 < 9; ++t9) {
            const int row = (t9 / 3 - 1) * oz + (t9 % 3 - 1) * oy;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int tap = 3 * t9 + dx;
                bf16x8 bcur[NT];
#pragma unroll
                for (int j = 0; j < NT; ++j) bcur[j] = bnext[j];
                if (tap < 26) {
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        bnext[j] = *reinterpret_cast<const bf16x8*>(wrow[j] + (tap + 1) * CIN + c0);
                }
                bf16x8 a[4];
                const int toff = (row + dx - 1) * HALO_VS;   // uniform
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(arow[i] + toff);
#pragma unroll
                for (int j = 0; j < NT; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bcur[j], acc[i][j], 0, 0, 0);
            }
        }
    }
    // output rows 4 grp .. 4 grp + 3 of a row tile are 4 consecutive voxels
    // along x (the block extent along x is at least 8): one element offset
    // per row tile, then + r out_stride
    if (gridDim.z > 1) {
        // split: this split's FP32 partial, summed by conv_ksplit_sum
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lv = wave * 64 + 16 * i + 4 * grp;
            const uint32_t lx = lv & ((1 << h.lbx) - 1), ly = (lv >> h.lbx) & ((1 << h.lby) - 1),
                           lz = lv >> (h.lbx + h.lby);
            const size_t v = sbase + ((size_t(z0 + lz) << (2 * ls)) | (size_t(y0 + ly) << ls) | (x0 + lx));
            float* kb = g.kpart + (size_t(blockIdx.z) * g.nvox + v) * COUT + co0 + r16;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                if (co0 + 16 * j + r16 >= COUT) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) kb[r * COUT + 16 * j] = acc[i][j][r];
            }
        }
        return;
    }
    using OutT = typename std::conditional<OUTF32, float, __bf16>::type;
    OutT* obase[4];
    const __bf16* gbase[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lv = wave * 64 + 16 * i + 4 * grp;
        const uint32_t lx = lv & ((1 << h.lbx) - 1), ly = (lv >> h.lbx) & ((1 << h.lby) - 1),
                       lz = lv >> (h.lbx + h.lby);
        const size_t v = sbase + ((size_t(z0 + lz) << (2 * ls)) | (size_t(y0 + ly) << ls) | (x0 + lx));
        obase[i] = static_cast<OutT*>(out) + v * g.out_stride + co0 + r16;
        if constexpr (ACT == ACT_RELU_GATE) gbase[i] = g.gate + v * g.gate_stride + co0 + r16;
    }
    const uint32_t os = g.out_stride, gs = g.gate_stride;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        if (co0 + 16 * j + r16 >= COUT) continue;
        const float bb = BIAS ? bias[co0 + 16 * j + r16] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float x = acc[i][j][r] + bb;
                if (ACT == ACT_RELU) x = x > 0.0f ? x : 0.0f;
                if constexpr (ACT == ACT_RELU_GATE) x = float(gbase[i][r * gs + 16 * j]) > 0.0f ? x : 0.0f;
                obase[i][r * os + 16 * j] = OutT(x);
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
    constexpr int ROW = VC + 16;                   // bf16 per LDS row: 72 dwords (see LDS_NOTE)
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

// Forward split over input channels: out = act(sum of the splits' FP32
// partials in split order + bias), one thread per voxel and 4 channels.
template <int ACT, bool BIAS, bool OUTF32>
__global__ void __launch_bounds__(256)
conv_ksplit_sum(ConvGeom g, uint32_t cout, uint32_t ks, const float* __restrict__ bias, void* __restrict__ out)
{
    const uint32_t q = cout / 4;
    const uint64_t t = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (t >= uint64_t(g.nvox) * q) return;
    const uint32_t v = uint32_t(t / q), c = uint32_t(t % q) * 4;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < ks; ++k) {
        const float4 p = *reinterpret_cast<const float4*>(g.kpart + (uint64_t(k) * g.nvox + v) * cout + c);
        s[0] += p.x;
        s[1] += p.y;
        s[2] += p.z;
        s[3] += p.w;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t co = c + e;
        float x = s[e] + (BIAS ? bias[co] : 0.0f);
        if (ACT == ACT_RELU) x = x > 0.0f ? x : 0.0f;
        if (ACT == ACT_RELU_GATE) x = float(g.gate[size_t(v) * g.gate_stride + co]) > 0.0f ? x : 0.0f;
        if (OUTF32) static_cast<float*>(out)[size_t(v) * g.out_stride + co] = x;
        else static_cast<__bf16*>(out)[size_t(v) * g.out_stride + co] = __bf16(x);
    }
}

// ---------------------------------------------------------------------------
// weight gradient, 3x3x3 with Cin a multiple of 32: a workgroup owns 32
// output channels x 16 input channels x all 27 taps and runs over the
// 256-voxel blocks of its voxel split (the direct kernel above re-gathers
// every tap of the input for every 64-wide k tile).  Per block it stages, in
// their global layout (voxel rows of channels) by 16-byte copies,
//   dz     [256 voxels][32 channels]                 (rows of PD bf16)
//   input  [halo voxels, x fastest][16 channels]     (rows of PX bf16)
// and reads both as MFMA operands, K along the voxels, with the transpose
// loads of tr_rows8: the 8 consecutive voxels of a lane group are 8
// consecutive rows, and a tap's shift along x is a shift by whole rows.  One
// MFMA reduces 32 voxels; wave w accumulates the taps w, w+4, ... for both
// 16-channel output tiles.  The next block's operands are fetched into
// registers while the current block runs on MFMA.
// ---------------------------------------------------------------------------
typedef short short4v __attribute__((ext_vector_type(4)));

// Lane li (0..15) of each 16-lane group receives column col + li of rows
// v0 .. v0 + 7 (v0 per group) of a row-major LDS tile with a pitch of P
// bf16: ds_read_b64_tr_b16 hands lane li element li % 4 of the 8 bytes read
// by lane 4 k + li / 4 of its group, k = 0..3, so lane li reads row
// v0 + li / 4 at column col + 4 (li % 4); a second read covers rows +4..+7.
__device__ __forceinline__ bf16x8 tr_rows8(const __bf16* tile, int v0, int P, int col, int li)
{
    const __bf16* p = tile + (v0 + (li >> 2)) * P + col + 4 * (li & 3);
    const short4v lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) short4v*)(p));
    const short4v hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) short4v*)(p + 4 * P));
    bf16x8 r;
    __builtin_memcpy(&r, &lo, 8);
    __builtin_memcpy(reinterpret_cast<char*>(&r) + 8, &hi, 8);
    return r;
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
conv_wgrad_halo(ConvGeom g, const __bf16* __restrict__ in, const __bf16* __restrict__ dz,
                float* __restrict__ partial, uint32_t blocks_per_split, float* __restrict__ bpart)
{
    constexpr int KP = conv_kp(CIN, 27);
    constexpr int CC = 16, BM = 32;
    constexpr int TW = 7;                                  // taps per wave (27 over 4 waves)
    constexpr int NCH = CIN / CC, NCO = COUT / BM;
    // row pitches (bf16): 16-byte aligned rows, padded off a multiple of 64 B
    constexpr int PD = BM + 8, PX = CC + 8;
    constexpr int NDZ = 256 * (BM / 8) / 256;              // dz chunks (16 B) per thread
    constexpr int NXI = (HALO_MAX * (CC / 8) + 255) / 256; // input chunks per thread
    static_assert(CIN % 32 == 0 && COUT % BM == 0, "shape");
    __shared__ __attribute__((aligned(16))) __bf16 sdz[256 * PD];
    __shared__ __attribute__((aligned(16))) __bf16 sx[HALO_MAX * PX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    // workgroups of one voxel split (all channel tiles) run on one XCD
    const uint32_t wg = xcd_block();
    const uint32_t split = wg / (NCH * NCO), rest = wg % (NCH * NCO);
    const int c0 = int(rest % NCH) * CC, co0 = int(rest / NCH) * BM;
    const int ls = int(g.log_s);
    const uint32_t S = 1u << ls, m = S - 1u;
    const HaloShape h = halo_shape(ls);
    const int bx = 1 << h.lbx;
    const int nxi = h.hx * h.hy * h.hz * (CC / 8);         // input chunks: (halo voxel, 8-channel group)
    const uint32_t nblocks = g.nvox >> 8;
    const uint32_t b_begin = split * blocks_per_split;
    const uint32_t b_end = b_begin + blocks_per_split < nblocks ? b_begin + blocks_per_split : nblocks;
    const uint32_t bps = 1u << (3 * ls - 8);
    const uint32_t nbx = S >> h.lbx, nby = S >> h.lby;
    f32x4 acc[TW][2];
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[t][0] = acc[t][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // one block's operands in registers
    bf16x8 rdz[NDZ], rx[NXI];
    auto fetch = [&](uint32_t blk) {
        const uint32_t smp = blk / bps, b = blk % bps;
        const uint32_t x0 = (b % nbx) << h.lbx, y0 = ((b / nbx) % nby) << h.lby, z0 = (b / (nbx * nby)) << h.lbz;
        const size_t sbase = size_t(smp) << (3 * ls);
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int t = threadIdx.x + 256 * n;
            const int lv = t >> 2, c8 = t & 3;
            const uint32_t lx = lv & (bx - 1), ly = (lv >> h.lbx) & ((1 << h.lby) - 1), lz = lv >> (h.lbx + h.lby);
            const size_t v = sbase + ((size_t(z0 + lz) << (2 * ls)) | (size_t(y0 + ly) << ls) | (x0 + lx));
            rdz[n] = *reinterpret_cast<const bf16x8*>(dz + v * g.out_stride + co0 + 8 * c8);
        }
#pragma unroll
        for (int n = 0; n < NXI; ++n) {
            const int t = threadIdx.x + 256 * n;
            if (t < nxi) {
                const int hv = t >> 1, c8 = t & 1;
                const int hx = hv % h.hx, row = hv / h.hx, hy = row % h.hy, hz = row / h.hy;
                const uint32_t x = (x0 + uint32_t(hx) - 1u) & m, y = (y0 + uint32_t(hy) - 1u) & m,
                               z = (z0 + uint32_t(hz) - 1u) & m;
                const size_t v = sbase + ((size_t(z) << (2 * ls)) | (size_t(y) << ls) | x);
                rx[n] = *reinterpret_cast<const bf16x8*>(in + v * g.in_stride + c0 + 8 * c8);
            }
        }
    };
    // bias gradient (bpart, workgroups of the first input-channel chunk):
    // every dz chunk of this thread holds channels co0 + 8 (tid % 4) .. + 7
    const bool bias_wg = bpart != nullptr && c0 == 0;
    float bs[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto store = [&]() {
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int t = threadIdx.x + 256 * n;
            *reinterpret_cast<bf16x8*>(sdz + (t >> 2) * PD + 8 * (t & 3)) = rdz[n];
            if (bias_wg)
#pragma unroll
                for (int j = 0; j < 8; ++j) bs[j] += float(rdz[n][j]);
        }
#pragma unroll
        for (int n = 0; n < NXI; ++n) {
            const int t = threadIdx.x + 256 * n;
            if (t < nxi) *reinterpret_cast<bf16x8*>(sx + (t >> 1) * PX + 8 * (t & 1)) = rx[n];
        }
    };

    if (b_begin < b_end) fetch(b_begin);
    for (uint32_t blk = b_begin; blk < b_end; ++blk) {
        __syncthreads();   // the previous block's MFMAs are done with the tiles
        store();
        __syncthreads();
        if (blk + 1 < b_end) fetch(blk + 1);   // the next block's loads overlap this block's MFMAs
#pragma unroll 1
        for (int r = 0; r < 8; ++r) {
            // this lane group's 8 voxels: block positions 32 r + 8 grp .. + 7 (one x run)
            const int lv = 32 * r + 8 * grp;
            const int lx = lv & (bx - 1), ly = (lv >> h.lbx) & ((1 << h.lby) - 1), lz = lv >> (h.lbx + h.lby);
            const bf16x8 a0 = tr_rows8(sdz, lv, PD, 0, r16);
            const bf16x8 a1 = tr_rows8(sdz, lv, PD, 16, r16);
#pragma unroll
            for (int tt = 0; tt < TW; ++tt) {
                const int tap = wave + 4 * tt;
                if (tap >= 27) break;
                const int dzt = tap / 9 - 1, dyt = (tap / 3) % 3 - 1, dxt = tap % 3 - 1;
                const int row = (lz + dzt + 1) * h.hy + (ly + dyt + 1);
                const bf16x8 bv = tr_rows8(sx, row * h.hx + lx + dxt + 1, PX, 0, r16);
                acc[tt][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bv, acc[tt][0], 0, 0, 0);
                acc[tt][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bv, acc[tt][1], 0, 0, 0);
            }
        }
    }
    if (bias_wg) {
        // this split's bias partial: the threads of one 8-channel group in
        // thread order, through the (now idle) dz tile
        __syncthreads();
        float* red = reinterpret_cast<float*>(sdz);
#pragma unroll
        for (int j = 0; j < 8; ++j) red[threadIdx.x * 8 + j] = bs[j];
        __syncthreads();
        if (threadIdx.x < BM) {
            float t = 0.0f;
            for (int k = threadIdx.x >> 3; k < int(blockDim.x); k += 4) t += red[k * 8 + (threadIdx.x & 7)];
            bpart[size_t(split) * COUT + co0 + threadIdx.x] = t;
        }
    }
    // C[co][ci]: lane holds column ci = c0 + r16, rows co = 4 grp .. + 3 of each 16-row tile
#pragma unroll
    for (int tt = 0; tt < TW; ++tt) {
        const int tap = wave + 4 * tt;
        if (tap >= 27) break;
        const int k = tap * CIN + c0 + r16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int co = co0 + 16 * i + 4 * grp + rr;
                partial[(size_t(split) * COUT + co) * KP + k] = acc[tt][i][rr];
            }
    }
}

// ---------------------------------------------------------------------------
// weight gradient for sides >= 32, where a block is 2 x 4 x-rows of 32
// voxels and one MFMA's 32 voxels are one x-row: three waves, one per tap
// dx.  Per output z-layer a wave keeps the dz^T fragments of its 4 rows in
// registers and loads the input fragment of each of the 18 halo rows it
// touches once; that fragment serves every tap dy whose output row lies in
// the layer (up to 3 taps x 2 output-channel tiles).
// Both tiles are staged in their global layout, voxel rows of channels, by
// 16-byte copies, and read as MFMA operands (K along the voxels) with the
// transpose loads ds_read_b64_tr_b16 (tr_rows8); a shift of the taps along x
// is a shift by whole rows, so one copy of the input halo serves all three.
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(2)))
conv_wgrad_rows(ConvGeom g, const __bf16* __restrict__ in, const __bf16* __restrict__ dz,
                float* __restrict__ partial, uint32_t blocks_per_split, float* __restrict__ bpart)
{
    constexpr int KP = conv_kp(CIN, 27);
    constexpr int CC = 16, BM = 32, NTH = 192;
    constexpr int NCH = CIN / CC, NCO = COUT / BM;
    // halo of the 2 x 4 x 32 block: 4 z-layers x HY rows of HX voxels
    constexpr int HY = 6, HX = 34, HV = 4 * HY * HX;
    // row pitches (bf16): 16-byte aligned rows, padded off a multiple of 64 B
    constexpr int PD = BM + 8, PX = CC + 8;
    constexpr int NDZ = (256 * (BM / 8) + NTH - 1) / NTH, NXI = (HV * (CC / 8) + NTH - 1) / NTH;
    static_assert(CIN % 32 == 0 && COUT % BM == 0, "shape");
    __shared__ __attribute__((aligned(16))) __bf16 sdz[256 * PD];
    __shared__ __attribute__((aligned(16))) __bf16 sx[HV * PX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;   // wave = dx + 1
    const int r16 = lane & 15, grp = lane >> 4;
    const uint32_t wg = xcd_block();
    const uint32_t split = wg / (NCH * NCO), rest = wg % (NCH * NCO);
    const int c0 = int(rest % NCH) * CC, co0 = int(rest / NCH) * BM;
    const int ls = int(g.log_s);
    const uint32_t S = 1u << ls, m = S - 1u;
    const uint32_t nblocks = g.nvox >> 8;
    const uint32_t b_begin = split * blocks_per_split;
    const uint32_t b_end = b_begin + blocks_per_split < nblocks ? b_begin + blocks_per_split : nblocks;
    const uint32_t bps = 1u << (3 * ls - 8);
    const uint32_t nbx = S >> 5, nby = S >> 2;
    f32x4 acc[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t][0] = acc[t][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // bias gradient (bpart, workgroups of the first input-channel chunk):
    // every dz chunk of this thread holds channels co0 + 8 (tid % 4) .. + 7
    const bool bias_wg = bpart != nullptr && c0 == 0;
    float bs[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

    auto load_tiles = [&](uint32_t blk) {
        const uint32_t smp = blk / bps, b = blk % bps;
        const uint32_t x0 = (b % nbx) << 5, y0 = ((b / nbx) % nby) << 2, z0 = (b / (nbx * nby)) << 1;
        const size_t sbase = size_t(smp) << (3 * ls);
        // every load first, then every LDS store
        bf16x8 rdz[NDZ], rx[NXI];
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int t = threadIdx.x + NTH * n;
            if (t < 256 * (BM / 8)) {
                const int lv = t >> 2, c8 = t & 3;
                const uint32_t lx = lv & 31, ly = (lv >> 5) & 3, lz = lv >> 7;
                const size_t v = sbase + ((size_t(z0 + lz) << (2 * ls)) | (size_t(y0 + ly) << ls) | (x0 + lx));
                rdz[n] = *reinterpret_cast<const bf16x8*>(dz + v * g.out_stride + co0 + 8 * c8);
            }
        }
#pragma unroll
        for (int n = 0; n < NXI; ++n) {
            const int t = threadIdx.x + NTH * n;
            if (t < HV * (CC / 8)) {
                const int hv = t >> 1, c8 = t & 1;
                const int hx = hv % HX, row = hv / HX, hy = row % HY, hz = row / HY;
                const uint32_t x = (x0 + uint32_t(hx) - 1u) & m, y = (y0 + uint32_t(hy) - 1u) & m,
                               z = (z0 + uint32_t(hz) - 1u) & m;
                const size_t v = sbase + ((size_t(z) << (2 * ls)) | (size_t(y) << ls) | x);
                rx[n] = *reinterpret_cast<const bf16x8*>(in + v * g.in_stride + c0 + 8 * c8);
            }
        }
#pragma unroll
        for (int n = 0; n < NDZ; ++n) {
            const int t = threadIdx.x + NTH * n;
            if (t < 256 * (BM / 8)) {
                *reinterpret_cast<bf16x8*>(sdz + (t >> 2) * PD + 8 * (t & 3)) = rdz[n];
                if (bias_wg)
#pragma unroll
                    for (int j = 0; j < 8; ++j) bs[j] += float(rdz[n][j]);
            }
        }
#pragma unroll
        for (int n = 0; n < NXI; ++n) {
            const int t = threadIdx.x + NTH * n;
            if (t < HV * (CC / 8)) *reinterpret_cast<bf16x8*>(sx + (t >> 1) * PX + 8 * (t & 1)) = rx[n];
        }
    };

    for (uint32_t blk = b_begin; blk < b_end; ++blk) {
        __syncthreads();   // the previous block's MFMAs are done with the tiles
        load_tiles(blk);
        __syncthreads();
        // one output z-layer at a time: its 4 rows' dz^T fragments in
        // registers, the 3 x 6 halo rows it touches loaded once each; this
        // wave's input fragment starts at halo x = 8 grp + dx + 1
#pragma unroll
        for (int lz = 0; lz < 2; ++lz) {
            bf16x8 a[4][2];
#pragma unroll
            for (int ly = 0; ly < 4; ++ly) {
                a[ly][0] = tr_rows8(sdz, 32 * (4 * lz + ly) + 8 * grp, PD, 0, r16);
                a[ly][1] = tr_rows8(sdz, 32 * (4 * lz + ly) + 8 * grp, PD, 16, r16);
            }
#pragma unroll
            for (int dzt = -1; dzt <= 1; ++dzt) {
                const int hz = lz + dzt + 1;
#pragma unroll
                for (int hy = 0; hy < HY; ++hy) {
                    const int rho = hz * HY + hy;
                    const bf16x8 bv = tr_rows8(sx, rho * HX + 8 * grp + wave, PX, 0, r16);
#pragma unroll
                    for (int dyt = -1; dyt <= 1; ++dyt) {
                        const int ly = hy - dyt - 1;
                        if (ly < 0 || ly > 3) continue;
                        const int t9 = (dzt + 1) * 3 + (dyt + 1);
                        acc[t9][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ly][0], bv, acc[t9][0], 0, 0, 0);
                        acc[t9][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ly][1], bv, acc[t9][1], 0, 0, 0);
                    }
                }
            }
        }
    }
    if (bias_wg) {
        // this split's bias partial: the threads of one 8-channel group in
        // thread order, through the (now idle) dz tile
        __syncthreads();
        float* red = reinterpret_cast<float*>(sdz);
#pragma unroll
        for (int j = 0; j < 8; ++j) red[threadIdx.x * 8 + j] = bs[j];
        __syncthreads();
        if (threadIdx.x < BM) {
            float t = 0.0f;
            for (int k = threadIdx.x >> 3; k < int(blockDim.x); k += 4) t += red[k * 8 + (threadIdx.x & 7)];
            bpart[size_t(split) * COUT + co0 + threadIdx.x] = t;
        }
    }
#pragma unroll
    for (int t9 = 0; t9 < 9; ++t9) {
        const int tap = t9 * 3 + wave;    // (dz+1)*9 + (dy+1)*3 + (dx+1)
        const int k = tap * CIN + c0 + r16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int co = co0 + 16 * i + 4 * grp + rr;
                partial[(size_t(split) * COUT + co) * KP + k] = acc[t9][i][rr];
            }
    }
}

// dW[i] = sum over splits of partial[s][i] in a fixed order (deterministic).
// Up to 64 splits: in split order.  More: the splits are summed in groups of
// SPLIT_GROUP consecutive ones in parallel (each group's sum stored in the
// group's first row, a row only that thread reads), then the group sums in
// order.
constexpr uint32_t SPLIT_GROUP = 32;

__global__ void __launch_bounds__(256)
sum_split_groups(float* __restrict__ partial, uint32_t n, uint32_t splits)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k0 = blockIdx.y * SPLIT_GROUP, k1 = k0 + SPLIT_GROUP < splits ? k0 + SPLIT_GROUP : splits;
    float s = 0.0f;
    for (uint32_t k = k0; k < k1; ++k) s += partial[size_t(k) * n + i];
    partial[size_t(k0) * n + i] = s;
}

__global__ void __launch_bounds__(256)
sum_splits(const float* __restrict__ partial, uint32_t n, uint32_t splits, uint32_t step, float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < splits; k += step) s += partial[size_t(k) * n + i];
    out[i] = s;
}

// db[c] = sum over splits of bpart[split][c], in split order.
__global__ void __launch_bounds__(256)
sum_bias_splits(const float* __restrict__ bpart, uint32_t cout, uint32_t splits, float* __restrict__ db)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= cout) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < splits; ++k) s += bpart[size_t(k) * cout + c];
    db[c] = s;
}

static void launch_sum_splits(float* partial, uint32_t n, uint32_t splits, float* dw, hipStream_t s)
{
    const uint32_t blocks = (n + 255) / 256;
    if (splits > 64) {
        hipLaunchKernelGGL(sum_split_groups, dim3(blocks, (splits + SPLIT_GROUP - 1) / SPLIT_GROUP), dim3(256), 0, s,
                           partial, n, splits);
        hipLaunchKernelGGL(sum_splits, dim3(blocks), dim3(256), 0, s, partial, n, splits, SPLIT_GROUP, dw);
    } else {
        hipLaunchKernelGGL(sum_splits, dim3(blocks), dim3(256), 0, s, partial, n, splits, 1u, dw);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TAPS>
struct ConvLaunch {
    static constexpr int BN_FWD = COUT >= 64 ? 64 : COUT;
    // the halo kernel: 3x3x3 with Cin a multiple of 32; Cout <= 96 in one block
    static constexpr bool HALO = TAPS == 27 && CIN % 32 == 0;
    static constexpr int BN_HALO = COUT <= 64 ? COUT : (COUT % 64 == 0 ? 64 : 48);

    template <int ACT, bool BIAS, bool OUTF32>
    static void fwd(const ConvGeom& g, const __bf16* in, const __bf16* w, const float* bias, void* out, hipStream_t s)
    {
        if constexpr (HALO) {
            // a 256-voxel block must fit one sample: side >= 8
            if (g.log_s >= 3) {
                // with a workspace, a grid that fills less than one round of
                // two workgroups per CU (512) is split over input-channel
                // chunks, doubled while the chunks divide evenly
                const uint32_t wgs = (g.nvox / 256u) * ((COUT + BN_HALO - 1) / BN_HALO);
                constexpr uint32_t NCHUNK = CIN / 32;
                uint32_t ks = 1;
                if (g.kpart)
                    while (NCHUNK % (2 * ks) == 0 && wgs * 2 * ks <= 512u &&
                           size_t(2 * ks) * g.nvox * COUT <= g.kpart_floats)
                        ks *= 2;
                const dim3 grid(g.nvox / 256u, (COUT + BN_HALO - 1) / BN_HALO, ks);
                hipLaunchKernelGGL((conv_fwd_halo<CIN, COUT, BN_HALO, ACT, BIAS, OUTF32>), grid, dim3(256), 0, s, g,
                                   in, w, bias, out);
                if (ks > 1) {
                    const uint64_t n = uint64_t(g.nvox) * (COUT / 4);
                    hipLaunchKernelGGL((conv_ksplit_sum<ACT, BIAS, OUTF32>), dim3(uint32_t((n + 255) / 256)), dim3(256),
                                       0, s, g, uint32_t(COUT), ks, bias, out);
                }
                return;
            }
        }
        const dim3 grid((g.nvox + 255u) / 256u, (COUT + BN_FWD - 1) / BN_FWD);
        hipLaunchKernelGGL((conv_fwd<CIN, COUT, TAPS, BN_FWD, ACT, BIAS, OUTF32>), grid, dim3(256), 0, s, g, in, w,
                           bias, out);
    }
    static constexpr int BM_W = COUT >= 64 ? 64 : COUT;
    static constexpr int BN_W = 64;

    static void forward(const ConvGeom& g, const __bf16* in, const __bf16* w, const float* bias, void* out, int act,
                        bool out_f32, hipStream_t s)
    {
        if (out_f32 && act == ACT_RELU) fwd<ACT_RELU, true, true>(g, in, w, bias, out, s);
        else if (out_f32) fwd<ACT_NONE, true, true>(g, in, w, bias, out, s);
        else if (act == ACT_RELU) fwd<ACT_RELU, true, false>(g, in, w, bias, out, s);
        else if (act == ACT_RELU_GATE) fwd<ACT_RELU_GATE, false, false>(g, in, w, bias, out, s);
        else if (bias) fwd<ACT_NONE, true, false>(g, in, w, bias, out, s);
        else fwd<ACT_NONE, false, false>(g, in, w, bias, out, s);
    }

    static constexpr bool HALO_W = HALO && COUT % 32 == 0;

    // the tiled kernels also reduce the bias gradient into db when given one
    static bool fuses_bias(uint32_t log_s) { return HALO_W && log_s >= 3; }

    static void wgrad(const ConvGeom& g, const __bf16* in, const __bf16* dz, float* partial, uint32_t splits,
                      float* dw, float* bpart, float* db, hipStream_t s)
    {
        constexpr int KP = conv_kp(CIN, TAPS);
        if constexpr (HALO_W) {
            if (g.log_s >= 3) {
                const uint32_t nblocks = g.nvox >> 8;
                const uint32_t per = (nblocks + splits - 1) / splits;
                float* bp = db ? bpart : nullptr;
                if (g.log_s >= 5)
                    hipLaunchKernelGGL((conv_wgrad_rows<CIN, COUT>), dim3((CIN / 16) * (COUT / 32) * splits),
                                       dim3(192), 0, s, g, in, dz, partial, per, bp);
                else
                    hipLaunchKernelGGL((conv_wgrad_halo<CIN, COUT>), dim3((CIN / 16) * (COUT / 32) * splits),
                                       dim3(256), 0, s, g, in, dz, partial, per, bp);
                launch_sum_splits(partial, uint32_t(COUT) * KP, splits, dw, s);
                if (db)
                    hipLaunchKernelGGL(sum_bias_splits, dim3((COUT + 255) / 256), dim3(256), 0, s, bpart,
                                       uint32_t(COUT), splits, db);
                return;
            }
        }
        const uint32_t per = ((g.nvox + splits - 1) / splits + 127u) / 128u * 128u;
        const dim3 grid((KP + BN_W - 1) / BN_W, (COUT + BM_W - 1) / BM_W, splits);
        hipLaunchKernelGGL((conv_wgrad<CIN, COUT, TAPS, BM_W, BN_W>), grid, dim3(256), 0, s, g, in, dz, partial, per);
        launch_sum_splits(partial, uint32_t(COUT) * KP, splits, dw, s);
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

// Voxel splits of a weight gradient: enough workgroups to fill the chip, few
// enough partials that the fixed-order split sum stays cheap.
uint32_t conv_wgrad_splits(int cin, int cout, int taps, uint32_t nvox)
{
    const uint32_t max_splits = nvox / 256 > 0 ? nvox / 256 : 1;
    uint32_t splits;
    if (taps == 27 && cin % 32 == 0 && cout % 32 == 0 && nvox >= 512) {
        // at most 1024 workgroups: two full rounds of two per CU (a count
        // just above 1024 would run a third, nearly empty round)
        const uint32_t base = uint32_t(cin / 16) * uint32_t(cout / 32);
        splits = 1024u / base;
        if (splits > 256u) splits = 256u;
    } else {
        // the direct kernel: about 2048 workgroups (the 1x1x1 head has a
        // single 64 x 64 tile, so its voxels carry all the parallelism)
        const uint32_t kb = (uint32_t(conv_kp(cin, taps)) + 63) / 64, cb = (uint32_t(cout) + 63) / 64;
        splits = 2048u / (kb * cb);
        if (splits > 1024u) splits = 1024u;
    }
    if (splits > max_splits) splits = max_splits;
    return splits < 1 ? 1 : splits;
}

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

// db (optional, with bpart: splits x cout floats): the bias gradient, reduced
// by the weight-gradient kernel where conv_wgrad_fuses_bias says so
hipError_t launch_conv_wgrad(int cin, int cout, int taps, const ConvGeom& g, const __bf16* in, const __bf16* dz,
                             float* partial, uint32_t splits, float* dw, float* bpart, float* db, hipStream_t s)
{
#define X(a, b, c)                                                                       \
    if (cin == a && cout == b && taps == c) {                                            \
        ConvLaunch<a, b, c>::wgrad(g, in, dz, partial, splits, dw, bpart, db, s);        \
        return hipGetLastError();                                                        \
    }
    LBF_CONV_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

bool conv_wgrad_fuses_bias(int cin, int cout, int taps, uint32_t log_s)
{
#define X(a, b, c) if (cin == a && cout == b && taps == c) return ConvLaunch<a, b, c>::fuses_bias(log_s);
    LBF_CONV_SHAPES(X)
#undef X
    return false;
}

}  // namespace lbf
