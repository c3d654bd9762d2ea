// kernels_batch.hip -- many independent single-phase crops in one launch.
//
// Every crop's fluid sites get slots in one shared slot space, padded per
// crop to whole 256-slot tiles, so a tile belongs to exactly one crop and a
// workgroup's crop parameters are uniform (scalar loads).  The step kernel is
// persistent: a fixed grid walks a device-resident list of active tiles, so
// crops that have converged drop out by rewriting that list, and the same
// captured hipGraph keeps working.  Per crop: periodic wrap inside the crop,
// halfway bounce-back on BB/SD, Guo force, SRT/MRT collision on FP32 shifted
// populations (collide_f32.h: SCSP::equilSRT/equilMRT, scsp.cpp:374-464).
// The per-crop convergence residuals are reduced per tile in FP64 and summed
// per crop in tile order (deterministic).
#include <type_traits>

#include "batch_types.h"
#include "stream.h"

namespace lbf {

// One fluid slot of one crop: pull (bounce-back where the upstream site is
// solid), macro, optional prescribed-velocity sweep, collision, store.
// Populations are blocked per 64 slots ([slot / 64][q][slot % 64], stream.h
// pop_index); a crop starts on a whole tile, so its slots form whole blocks
// and a crop-local offset plus the crop's block base addresses the batch.
template <int Q, bool MRT, bool INIT, typename O>
__device__ __forceinline__ void batch_site(const BatchParams& p, const float* __restrict__ fin,
                                           float* __restrict__ fout, const CropDev& cr, uint32_t slot,
                                           uint32_t s, uint32_t dsc, const uint2* __restrict__ rank,
                                           const LabelF* __restrict__ lab,
                                           float* __restrict__ ucur, float* __restrict__ uadj,
                                           const float* __restrict__ upres, DevStatus* __restrict__ status)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const LabelF lb = lab[dsc >> DESC_LABEL_SHIFT];
    const float F[3] = {lb.force[0], lb.force[1], D == 3 ? lb.force[2] : 0.0f};
    const uint32_t local = slot - cr.slot_base;
    const O base = O(cr.slot_base >> 6) * O(Q * 64);
    float h[Q];
    if (INIT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float ef = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) ef += float(L::e[q][d]) * F[d];
            h[q] = float(-1.5 * L::w[q]) * ef;
        }
    } else {
        // bit op(q) of the descriptor: x - e_q is solid and bounces back into
        // op(q); otherwise the crop's rank index gives the upstream slot
        const Stencil stn = make_stencil(cr.g, s);
        O off[Q];
        h[0] = ld(fin, base + pop_index<Q, true, O>(0u, 0, local));
        pull_offsets<Q, SPARSE, O, true>(rank + cr.rank_base, dsc & DESC_LINK_MASK, local, stn, cr.g.nx, 0u, off);
#pragma unroll
        for (int q = 1; q < Q; ++q) h[q] = ld(fin, base + off[q]);
    }
    float u[D];
    const float drho = macro_f32<Q>(h, p.inv_rho0, lb.half_rho, u);
    if (!isfinite(drho)) raise_status(status, 0, p.t);
    if (p.flags & 4u) {
        // Consistent initialisation in the spirit of Mei, Luo, Lallemand &
        // d'Humieres (Computers & Fluids 35 (2006) 855-862): the streamed
        // populations get the momentum of the prescribed velocity (added as
        // 3 w_q e_q . dj, which carries momentum dj and no mass) before the
        // ordinary collision, so the non-equilibrium parts and the density
        // settle to the given flow.  With MRT the momentum moments do not
        // relax, so the velocity is imposed on the populations themselves
        // rather than only through the equilibrium.
        float dj[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float up = upres[uint64_t(d) * p.S + slot];
            dj[d] = p.relax.rho0 * (up - u[d]);
            u[d] = up;
        }
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            float ej = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) ej += float(L::e[q][d]) * dj[d];
            h[q] += float(3.0 * L::w[q]) * ej;
        }
    }
    if (p.flags & 1u) {
#pragma unroll
        for (int d = 0; d < D; ++d) ucur[uint64_t(d) * p.S + slot] = u[d];
    }
    if (p.flags & 2u) {
#pragma unroll
        for (int d = 0; d < D; ++d) uadj[uint64_t(d) * p.S + slot] = u[d];
    }
    collide_f32<Q, MRT>(h, drho, u, F, p.relax);
    bool finite = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        finite = finite && isfinite(h[q]);
        st_pop<true>(fout, base + pop_index<Q, true, O>(0u, q, local), h[q]);
    }
    if (!finite) raise_status(status, 4, p.t);
}

// Persistent step over the active tiles.  The work unit is UNIT slots of one
// tile (TILE / UNIT units per tile) and a workgroup of BS threads walks a
// unit in UNIT / BS passes; the autotuner picks BS, UNIT and the grid.
template <int Q, bool MRT, bool INIT, bool WIDE, uint32_t BS, uint32_t UNIT>
__global__ void __launch_bounds__(BS)
batch_step(BatchParams p, const float* __restrict__ fin, float* __restrict__ fout,
           const uint32_t* __restrict__ tiles, const uint32_t* __restrict__ ntiles,
           const uint32_t* __restrict__ tile_crop, const CropDev* __restrict__ crops,
           const uint32_t* __restrict__ site, const uint32_t* __restrict__ desc, const uint2* __restrict__ rank,
           const LabelF* __restrict__ lab, float* __restrict__ ucur, float* __restrict__ uadj,
           const float* __restrict__ upres, DevStatus* __restrict__ status)
{
    static_assert(TILE % UNIT == 0 && UNIT % BS == 0, "a unit is whole passes of the block inside one tile");
    using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    constexpr uint32_t PER_TILE = TILE / UNIT;
    const uint32_t n = *ntiles * PER_TILE;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t tile = tiles[k / PER_TILE];
        const CropDev cr = crops[tile_crop[tile]];
#pragma unroll
        for (uint32_t j = 0; j < UNIT; j += BS) {
            const uint32_t slot = tile * TILE + (k % PER_TILE) * UNIT + j + threadIdx.x;
            const uint32_t s = site[slot];
            if (s != NO_SITE)
                batch_site<Q, MRT, INIT, O>(p, fin, fout, cr, slot, s, desc[slot], rank, lab, ucur, uadj, upres,
                                            status);
        }
    }
}

// Per tile: sum (u - u_prev)^2, sum u^2, sum (u - u_adj)^2, sum u_axis,
// then u_prev = u.  One workgroup per active tile (persistent loop).
__global__ void __launch_bounds__(TILE)
batch_report(uint64_t S, int D, int axis, const uint32_t* __restrict__ tiles,
             const uint32_t* __restrict__ ntiles, const uint32_t* __restrict__ site,
             const float* __restrict__ ucur, float* __restrict__ uprev, const float* __restrict__ uadj,
             double* __restrict__ partial)
{
    __shared__ double lds[TILE / 64][4];
    const uint32_t n = *ntiles;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t tile = tiles[k];
        const uint32_t slot = tile * TILE + threadIdx.x;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (site[slot] != NO_SITE) {
            for (int d = 0; d < D; ++d) {
                const uint64_t i = uint64_t(d) * S + slot;
                const double v = ucur[i], pv = uprev[i];
                const double a = uadj[i];
                acc[0] += (v - pv) * (v - pv);
                acc[1] += v * v;
                acc[2] += (v - a) * (v - a);
                if (d == axis) acc[3] += v;
                uprev[i] = float(v);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double v = acc[j];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            acc[j] = v;
        }
        if ((threadIdx.x & 63) == 0)
            for (int j = 0; j < 4; ++j) lds[threadIdx.x >> 6][j] = acc[j];
        __syncthreads();
        if (threadIdx.x < 4) {
            double v = 0.0;
            for (int w = 0; w < int(TILE / 64); ++w) v += lds[w][threadIdx.x];
            partial[uint64_t(tile) * 4 + threadIdx.x] = v;
        }
        __syncthreads();
    }
}

// Per crop: sum its tiles' partials in tile order.
__global__ void __launch_bounds__(64)
batch_crop_sums(uint32_t ncrops, const CropDev* __restrict__ crops, const double* __restrict__ partial,
                double* __restrict__ out)
{
    const uint32_t c = blockIdx.x;
    if (c >= ncrops) return;
    const CropDev cr = crops[c];
    for (int j = 0; j < 4; ++j) {
        double v = 0.0;
        for (uint32_t t = cr.tile_begin + threadIdx.x; t < cr.tile_end; t += 64) v += partial[uint64_t(t) * 4 + j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x == 0) out[uint64_t(c) * 4 + j] = v;
    }
}

// Macro fields of one crop after its last report: fields[site][0..D-1] = u of
// the last fields step, fields[site][D] = rho0 + sum h (collision conserves
// mass, so the post-collision sum is the macro density).  One workgroup per tile.
template <int Q>
__global__ void __launch_bounds__(TILE)
batch_fields(uint64_t S, float rho0, uint32_t tile_begin, const uint32_t* __restrict__ site,
             const float* __restrict__ f, const float* __restrict__ ucur, float* __restrict__ fields)
{
    constexpr int D = Lattice<Q>::D;
    const uint32_t slot = (tile_begin + blockIdx.x) * TILE + threadIdx.x;
    const uint32_t s = site[slot];
    if (s == NO_SITE) return;
    float drho = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) drho += f[pop_index<Q, true, uint64_t>(0u, q, slot)];
#pragma unroll
    for (int d = 0; d < D; ++d) fields[uint64_t(s) * (D + 1) + d] = ucur[uint64_t(d) * S + slot];
    fields[uint64_t(s) * (D + 1) + D] = rho0 + drho;
}

// Warm start of one crop: the t = 0 state from given fields, i.e. the
// equilibrium of (rho, u - F/(2 rho0)) followed by the t = 0 macro step and
// collision, as the initial branch does for the rest state.  Sites outside the
// clusters that carry the steady flow (flow8 = 0; their steady velocity is 0)
// start from the rest state instead.
template <int Q, bool MRT>
__global__ void __launch_bounds__(TILE)
batch_warm(BatchParams p, uint32_t tile_begin, const uint32_t* __restrict__ site, const uint8_t* __restrict__ lab8,
           const uint8_t* __restrict__ flow8, const LabelF* __restrict__ lab, const float* __restrict__ fields,
           float* __restrict__ f, float* __restrict__ upres)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t slot = (tile_begin + blockIdx.x) * TILE + threadIdx.x;
    const uint32_t s = site[slot];
    if (s == NO_SITE) return;
    const LabelF lb = lab[lab8[slot]];
    const float F[3] = {lb.force[0], lb.force[1], D == 3 ? lb.force[2] : 0.0f};
    float ue[D];
    float usq = 0.0f;
    // sites off the flowing clusters start at rest, as in the cold start
    const bool flows = flow8[slot] != 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float ud = flows ? fields[uint64_t(s) * (D + 1) + d] : 0.0f;
        upres[uint64_t(d) * p.S + slot] = ud;
        ue[d] = ud - lb.half_rho[d];
        usq += ue[d] * ue[d];
    }
    const float rho0 = p.relax.rho0;
    const float dr = flows ? fields[uint64_t(s) * (D + 1) + D] - rho0 : 0.0f;
    float h[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        float eu = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) eu += float(L::e[q][d]) * ue[d];
        h[q] = float(L::w[q]) * (dr + rho0 * (eu * (3.0f + 4.5f * eu) - 1.5f * usq));
    }
    float u[D];
    const float drho = macro_f32<Q>(h, p.inv_rho0, lb.half_rho, u);
    collide_f32<Q, MRT>(h, drho, u, F, p.relax);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[pop_index<Q, true, uint64_t>(0u, q, slot)] = h[q];
}

// a = (a + b) / 2 over n floats: the average of two consecutive states, which
// removes the populations' period-2 component.
__global__ void __launch_bounds__(256) batch_average(float* __restrict__ a, const float* __restrict__ b, uint64_t n)
{
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i < n) a[i] = 0.5f * (a[i] + b[i]);
}

hipError_t launch_batch_average(float* a, const float* b, uint64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(batch_average, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_batch_fields(int Q, uint64_t S, float rho0, uint32_t tile_begin, uint32_t ntiles,
                               const uint32_t* site, const float* f, const float* ucur, float* fields, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
    if (Q == 19)
        hipLaunchKernelGGL(batch_fields<19>, dim3(ntiles), dim3(TILE), 0, s, S, rho0, tile_begin, site, f, ucur,
                           fields);
    else
        hipLaunchKernelGGL(batch_fields<9>, dim3(ntiles), dim3(TILE), 0, s, S, rho0, tile_begin, site, f, ucur,
                           fields);
    return hipGetLastError();
}

hipError_t launch_batch_warm(int Q, bool mrt, const BatchParams& p, uint32_t tile_begin, uint32_t ntiles,
                             const uint32_t* site, const uint8_t* lab8, const uint8_t* flow8, const LabelF* lab,
                             const float* fields, float* f, float* upres, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
    const dim3 g(ntiles), b(TILE);
    if (Q == 19) {
        if (mrt) hipLaunchKernelGGL((batch_warm<19, true>), g, b, 0, s, p, tile_begin, site, lab8, flow8, lab, fields, f, upres);
        else hipLaunchKernelGGL((batch_warm<19, false>), g, b, 0, s, p, tile_begin, site, lab8, flow8, lab, fields, f, upres);
    } else {
        if (mrt) hipLaunchKernelGGL((batch_warm<9, true>), g, b, 0, s, p, tile_begin, site, lab8, flow8, lab, fields, f, upres);
        else hipLaunchKernelGGL((batch_warm<9, false>), g, b, 0, s, p, tile_begin, site, lab8, flow8, lab, fields, f, upres);
    }
    return hipGetLastError();
}

template <int Q, bool MRT, bool INIT, bool WIDE, uint32_t BS, uint32_t UNIT>
static void launch_batch_s(uint32_t grid, const BatchParams& p, const BatchBuffers& b, hipStream_t s)
{
    hipLaunchKernelGGL((batch_step<Q, MRT, INIT, WIDE, BS, UNIT>), dim3(grid), dim3(BS), 0, s, p,
                       static_cast<const float*>(b.fin), static_cast<float*>(b.fout), b.tiles, b.ntiles, b.tile_crop,
                       static_cast<const CropDev*>(b.crops), b.site, b.desc, b.rank, b.lab, b.ucur, b.uadj, b.upres,
                       b.status);
}

template <int Q, bool MRT, bool INIT, bool WIDE>
static hipError_t launch_batch_w(uint32_t grid, const BatchParams& p, const BatchBuffers& b, BatchShape sh,
                                 hipStream_t s)
{
    // the initial step runs once, in the default shape
    if constexpr (INIT) {
        launch_batch_s<Q, MRT, INIT, WIDE, 256, 256>(grid, p, b, s);
        return hipSuccess;
    }
    if (sh.block == 256 && sh.unit == 256) launch_batch_s<Q, MRT, INIT, WIDE, 256, 256>(grid, p, b, s);
    else if (sh.block == 128 && sh.unit == 256) launch_batch_s<Q, MRT, INIT, WIDE, 128, 256>(grid, p, b, s);
    else if (sh.block == 128 && sh.unit == 128) launch_batch_s<Q, MRT, INIT, WIDE, 128, 128>(grid, p, b, s);
    else if (sh.block == 64 && sh.unit == 256) launch_batch_s<Q, MRT, INIT, WIDE, 64, 256>(grid, p, b, s);
    else if (sh.block == 64 && sh.unit == 64) launch_batch_s<Q, MRT, INIT, WIDE, 64, 64>(grid, p, b, s);
    else return hipErrorInvalidValue;
    return hipSuccess;
}

template <int Q, bool MRT, bool INIT>
static hipError_t launch_batch_q(uint32_t grid, const BatchParams& p, const BatchBuffers& b, BatchShape sh, bool wide,
                                 hipStream_t s)
{
    return wide ? launch_batch_w<Q, MRT, INIT, true>(grid, p, b, sh, s)
                : launch_batch_w<Q, MRT, INIT, false>(grid, p, b, sh, s);
}

hipError_t launch_batch_step(int Q, bool mrt, bool init, uint32_t grid, BatchShape sh, const BatchParams& p,
                             const BatchBuffers& b, hipStream_t s)
{
    const bool wide = p.S * uint64_t(Q) * 4u >= (uint64_t(1) << 32);
    hipError_t e;
    if (Q == 19) {
        if (mrt) e = init ? launch_batch_q<19, true, true>(grid, p, b, sh, wide, s)
                          : launch_batch_q<19, true, false>(grid, p, b, sh, wide, s);
        else e = init ? launch_batch_q<19, false, true>(grid, p, b, sh, wide, s)
                      : launch_batch_q<19, false, false>(grid, p, b, sh, wide, s);
    } else {
        if (mrt) e = init ? launch_batch_q<9, true, true>(grid, p, b, sh, wide, s)
                          : launch_batch_q<9, true, false>(grid, p, b, sh, wide, s);
        else e = init ? launch_batch_q<9, false, true>(grid, p, b, sh, wide, s)
                      : launch_batch_q<9, false, false>(grid, p, b, sh, wide, s);
    }
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_batch_report(uint32_t grid, uint64_t S, int D, int axis, const BatchBuffers& b, uint32_t ncrops,
                               double* partial, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(batch_report, dim3(grid), dim3(TILE), 0, s, S, D, axis, b.tiles, b.ntiles,
                       b.site, b.ucur, b.uprev, b.uadj, partial);
    hipLaunchKernelGGL(batch_crop_sums, dim3(ncrops), dim3(64), 0, s, ncrops,
                       static_cast<const CropDev*>(b.crops), partial, out);
    return hipGetLastError();
}

}  // namespace lbf
