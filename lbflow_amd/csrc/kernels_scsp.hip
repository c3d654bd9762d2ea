// kernels_scsp.hip -- single-phase (scsp) time step on CDNA4.
//
// One fused pull kernel per step: stream (periodic wrap in the index, the
// reference's MPI halo on one GPU) + halfway bounce-back + macro variables +
// half-force shift + SRT/MRT-with-Guo collision, one thread per lattice site,
// SoA populations [q][slot] so every direction is a coalesced HBM stream.
//
// Reference operators (zhou1800/lbflow):
//   streaming + bounceBack   solver/src/lbm/simulator.cpp:73-111
//   calcMacroVar             solver/src/lbm/scsp.cpp:136-160
//   bodyForce                solver/src/lbm/scsp.cpp:162-175
//   equilSRT / equilMRT      solver/src/lbm/scsp.cpp:374-464
//   forceSource              solver/src/lbm/simulator.cpp:230-246
//   initialize               solver/src/lbm/scsp.cpp:116-134
//   validatePopulations      solver/src/lbm/simulator.cpp:327-333
//   zouHeBoundary (PB/VB)    solver/src/lbm/scsp.cpp:188-256 (zouhe.h)
//
// Layouts (template parameter):
//   DENSE_ALL    slot = site, every site streamed (solid sites are streamed
//                but never collided, exactly like the reference, so
//                checkpoints are byte-identical);
//   DENSE_FLUID  slot = site, solid sites skipped;
//   SPARSE       slot = fluid index: populations and descriptors of fluid
//                sites only, a sorted fluid-site list and a 32-site rank
//                index that turns a neighbour site into its slot, so solid
//                voxels cost neither HBM traffic nor lanes.
//
// FP64 kernels keep the reference's floating-point operation order (the file
// is built with -ffp-contract=off) and are bit-identical to it.  Terms whose
// lattice coefficient is 0 are skipped and +-1 coefficients become add/sub:
// both are exact in IEEE arithmetic except for the sign of an exact zero,
// which never reaches a stored value because every such sum is later added
// to a non-zero density.
//
// FP32 kernels store h_q = f_q - w_q rho0 ("shifted" populations); streaming
// and bounce-back commute with the shift (w_q = w_op(q)), rho = rho0 + sum h,
// and the MRT collision is done in moment space with the equilibrium and
// Guo-source moments evaluated from compile-time moment tables of the
// reference's transform matrix, so no large constant is carried in single
// precision and the arithmetic is about 400 FP32 operations per site.
#include <type_traits>

#include "collide_f32.h"
#include "device.h"
#include "stream.h"
#include "zouhe.h"

namespace lbf {

constexpr uint32_t F_FIELDS = 1, F_ADJ = 2;
constexpr double C1 = 3.0, C2 = 4.5, C3 = 1.5;   // 3/c^2, 4.5/c^4, 1.5/c^2 with c = 1


// ---------------------------------------------------------------------------
// site / slot resolution
// ---------------------------------------------------------------------------
template <int LAYOUT>
__device__ __forceinline__ bool resolve(const ScspParams& p, const uint32_t* __restrict__ list, uint32_t& site,
                                        uint32_t& slot)
{
    uint32_t k = xcd_block() * 256u + threadIdx.x;
    if constexpr (LAYOUT == SPARSE) {
        if (k >= p.nfluid) return false;
        site = list[k];
    } else {
        if (p.tile_chunks) {
            // block b -> (column t, layer z, chunk j of the column)
            const uint32_t b = k >> 8, T = p.tile_chunks;
            const uint32_t j = b % T, r = b / T, z = r % p.g.nz, t = r / p.g.nz;
            k = ((z * p.layer_chunks + t * T + j) << 8) | threadIdx.x;
        }
        if (k >= p.g.N) return false;
        site = k;
    }
    slot = k;
    return true;
}

// ---------------------------------------------------------------------------
// FP64, reference operation order.
// ---------------------------------------------------------------------------
template <int Q, bool MRT, bool INIT, int LAYOUT, bool WIDE>
__global__ void __launch_bounds__(256)
scsp_step_f64(ScspParams p, const double* __restrict__ fin, double* __restrict__ fout,
              const uint32_t* __restrict__ desc, const uint32_t* __restrict__ list,
              const uint2* __restrict__ rank, const DevLabel* __restrict__ lab,
              const double* __restrict__ AT, const double* __restrict__ B, double* __restrict__ rho_out,
              double* __restrict__ u_out, double* __restrict__ uadj, DevStatus* __restrict__ status,
              const DevZouHe* __restrict__ zh)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    // The dense MRT operators are staged in LDS once per workgroup and read
    // with uniform addresses (broadcast), before any thread leaves.
    __shared__ double sAT[MRT ? Q * Q : 1], sB[MRT ? Q * Q : 1];
    if constexpr (MRT) {
        for (int k = threadIdx.x; k < Q * Q; k += 256) {
            sAT[k] = AT[k];
            sB[k] = B[k];
        }
        __syncthreads();
    }
    uint32_t i, slot;
    if (!resolve<LAYOUT>(p, list, i, slot)) return;
    const uint32_t P = LAYOUT == SPARSE ? p.nfluid : uint32_t(p.g.N);   // plane length
    constexpr bool BLK = LAYOUT == SPARSE;   // sparse: blocks of 64 slots
    const uint32_t dsc = desc[LAYOUT == SPARSE ? slot : i];
    const DevLabel* lb = lab + (dsc >> DESC_LABEL_SHIFT);
    const int kind = lb->kind;
    double f[Q];
    Stencil stn{};
    if (!INIT) stn = make_stencil(p.g, i);

    if (!(kind & KIND_FLUID)) {
        if (INIT || LAYOUT != DENSE_ALL) return;
        // Solid sites are streamed but never collided, like the reference's
        // whole-array streaming; their contents only matter for checkpoints.
        bool finite = true;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            f[q] = ld(fin, O(q) * P + (q == 0 ? slot : upstream<Q>(stn, q)));
            finite = finite && isfinite(f[q]);
            st(fout, O(q) * P + slot, f[q]);
        }
        if (!finite) raise_status(status, 4, p.t);
        return;
    }

    const double force[3] = {lb->force[0], lb->force[1], D == 3 ? lb->force[2] : 0.0};
    const bool forced = kind & KIND_FORCED;
    if (INIT) {
        // SCSP::initialize: f = w rho0 - 1.5 w (e.F)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ef = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) ef += force[d];
                else if (L::e[q][d] == -1) ef += -force[d];
            }
            f[q] = p.idist[q];
            f[q] -= 1.5 * L::w[q] * ef;
        }
    } else {
        f[0] = ld(fin, pop_index<Q, BLK, O>(P, 0, slot));
        O off[Q];
        pull_offsets<Q, LAYOUT, O, BLK>(rank, dsc, slot, stn, p.g.nx, P, off);
#pragma unroll
        for (int q = 1; q < Q; ++q) f[q] = ld(fin, off[q]);
    }

    double rho, u[D];
    if (kind & KIND_ZOUHE) {
        // PB/VB: density and velocity come from the boundary, not from
        // calcMacroVar; at t = 0 they are initialize's rho0 and 0.
        if (INIT) {
            rho = p.dens0;
#pragma unroll
            for (int d = 0; d < D; ++d) u[d] = 0.0;
        } else {
            zouhe_f64<Q>(f, zh[dsc >> DESC_LABEL_SHIFT], rho, u);
        }
    } else {
        // calcMacroVar: rho = sum_q f_q (q ascending); u = (sum +-f) / rho0
        rho = f[0];
#pragma unroll
        for (int q = 1; q < Q; ++q) rho += f[q];
        if (!isfinite(rho) || rho <= 0.0) raise_status(status, 0, p.t);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double s = 0;
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                if (L::e[q][d] == 1) s += f[q];
                else if (L::e[q][d] == -1) s -= f[q];
            }
            s /= p.dens0;
            u[d] = s;
        }
        // bodyForce: u += (F/2)/rho0 on GV sites
        if (kind & KIND_GV) {
#pragma unroll
            for (int d = 0; d < D; ++d) u[d] += lb->half_rho[d];
        }
    }
    if (p.flags & F_FIELDS) {
        rho_out[i] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) u_out[uint64_t(d) * p.g.N + i] = u[d];
    }
    if (p.flags & F_ADJ) {
#pragma unroll
        for (int d = 0; d < D; ++d) uadj[uint64_t(d) * p.g.N + i] = u[d];
    }

    // collision
    double usq = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) usq += u[d] * u[d];
    if constexpr (!MRT) {
        double uf = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) uf += u[d] * force[d];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ui = 0, ef = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) { ui += u[d]; ef += force[d]; }
                else if (L::e[q][d] == -1) { ui += -u[d]; ef += -force[d]; }
            }
            const double feq = L::w[q] * (rho + p.dens0 * (C1 * ui + C2 * ui * ui - C3 * usq));
            f[q] += (feq - f[q]) * p.omega;
            // forceSource: G_q = w_q (3 (e.F - u.F) + 9 (e.u)(e.F))
            if (forced) f[q] += p.omh * (L::w[q] * (3.0 * (ef - uf) + 9.0 * ui * ef));
        }
    } else {
        double r[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ui = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) ui += u[d];
                else if (L::e[q][d] == -1) ui += -u[d];
            }
            const double feq = L::w[q] * (rho + p.dens0 * (C1 * ui + C2 * ui * ui - C3 * usq));
            r[q] = feq - f[q];
        }
        double ff[Q];
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) s += sAT[q * Q + qq] * r[q];
            ff[qq] = s;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] += ff[q];
        if (forced) {
            double uf = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) uf += u[d] * force[d];
            double src[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double eu = 0.0, ef = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    if (L::e[q][d] == 1) { eu += u[d]; ef += force[d]; }
                    else if (L::e[q][d] == -1) { eu += -u[d]; ef += -force[d]; }
                }
                src[q] = L::w[q] * (3.0 * (ef - uf) + 9.0 * eu * ef);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < Q; ++j) s = s + src[j] * sB[q * Q + j];
                f[q] += s;
            }
        }
    }
    bool finite = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        finite = finite && isfinite(f[q]);
        st(fout, pop_index<Q, BLK, O>(P, q, slot), f[q]);
    }
    if (!finite) raise_status(status, 4, p.t);
}

// ---------------------------------------------------------------------------
// FP32 shifted populations.
// ---------------------------------------------------------------------------
template <int Q, bool MRT, bool INIT, int LAYOUT, bool WIDE>
__global__ void __launch_bounds__(256)
scsp_step_f32(ScspParams p, const float* __restrict__ fin, float* __restrict__ fout,
              const uint32_t* __restrict__ desc, const uint32_t* __restrict__ list,
              const uint2* __restrict__ rank, const LabelF* __restrict__ lab, double* __restrict__ rho_out,
              double* __restrict__ u_out, double* __restrict__ uadj, DevStatus* __restrict__ status,
              const DevZouHe* __restrict__ zh)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    uint32_t i, slot;
    if (!resolve<LAYOUT>(p, list, i, slot)) return;
    const uint32_t P = LAYOUT == SPARSE ? p.nfluid : uint32_t(p.g.N);
    constexpr bool BLK = LAYOUT == SPARSE;   // sparse: blocks of 64 slots
    const uint32_t dsc = desc[LAYOUT == SPARSE ? slot : i];
    const LabelF lb = lab[dsc >> DESC_LABEL_SHIFT];
    float h[Q];
    Stencil stn{};
    if (!INIT) stn = make_stencil(p.g, i);

    if (!(lb.kind & KIND_FLUID)) {
        if (INIT || LAYOUT != DENSE_ALL) return;
#pragma unroll
        for (int q = 0; q < Q; ++q)
            st(fout, O(q) * P + slot, ld(fin, O(q) * P + (q == 0 ? slot : upstream<Q>(stn, q))));
        return;
    }
    const float F[3] = {lb.force[0], lb.force[1], D == 3 ? lb.force[2] : 0.0f};
    if (INIT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float ef = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) ef += float(L::e[q][d]) * F[d];
            h[q] = float(-1.5 * L::w[q]) * ef;
        }
    } else {
        h[0] = ld(fin, pop_index<Q, BLK, O>(P, 0, slot));
        O off[Q];
        pull_offsets<Q, LAYOUT, O, BLK>(rank, dsc, slot, stn, p.g.nx, P, off);
#pragma unroll
        for (int q = 1; q < Q; ++q) h[q] = ld(fin, off[q]);
    }
    // macro on shifted populations: rho = rho0 + sum h, j = sum e h
    // (PB/VB: from the boundary; at t = 0 every h is 0, so drho = 0, u = 0)
    float u[D];
    float drho;
    if (!INIT && (lb.kind & KIND_ZOUHE)) {
        zouhe_f32<Q>(h, zh[dsc >> DESC_LABEL_SHIFT], p.rho0_32, drho, u);
    } else {
        drho = macro_f32<Q>(h, p.inv_rho0, lb.half_rho, u);
        if (!isfinite(drho) || p.dens0 + double(drho) <= 0.0) raise_status(status, 0, p.t);
    }
    if (p.flags & F_FIELDS) {
        rho_out[i] = p.dens0 + double(drho);
#pragma unroll
        for (int d = 0; d < D; ++d) u_out[uint64_t(d) * p.g.N + i] = double(u[d]);
    }
    if (p.flags & F_ADJ) {
#pragma unroll
        for (int d = 0; d < D; ++d) uadj[uint64_t(d) * p.g.N + i] = double(u[d]);
    }
    const RelaxF32 relax{p.rho0_32, p.omega32, p.omh32, p.c1_nu, p.c2_nu, p.c1_q, p.c2_q};
    collide_f32<Q, MRT>(h, drho, u, F, relax);
    bool finite = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        finite = finite && isfinite(h[q]);
        st_pop<BLK>(fout, pop_index<Q, BLK, O>(P, q, slot), h[q]);
    }
    if (!finite) raise_status(status, 4, p.t);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int Q, bool MRT, bool INIT, int LAYOUT, bool WIDE>
static void launch_scsp_w(const ScspParams& p, const ScspBuffers& b, bool fp32, hipStream_t s)
{
    const uint64_t n = LAYOUT == SPARSE ? p.nfluid : p.g.N;
    const dim3 grid(uint32_t((n + 255) / 256)), block(256);
    if (n == 0) return;
    if (fp32)
        hipLaunchKernelGGL((scsp_step_f32<Q, MRT, INIT, LAYOUT, WIDE>), grid, block, 0, s, p,
                           static_cast<const float*>(b.fin), static_cast<float*>(b.fout), b.desc, b.list, b.rank,
                           b.lab32, b.rho, b.u, b.uadj, b.status, b.zh);
    else
        hipLaunchKernelGGL((scsp_step_f64<Q, MRT, INIT, LAYOUT, WIDE>), grid, block, 0, s, p,
                           static_cast<const double*>(b.fin), static_cast<double*>(b.fout), b.desc, b.list,
                           b.rank, b.lab, b.AT, b.B, b.rho, b.u, b.uadj, b.status, b.zh);
}

template <int Q, bool MRT, bool INIT, int LAYOUT>
static void launch_scsp_l(const ScspParams& p, const ScspBuffers& b, bool fp32, hipStream_t s)
{
    // sparse: the last block of 64 slots is allocated whole
    const uint64_t P = LAYOUT == SPARSE ? (uint64_t(p.nfluid) + 63) / 64 * 64 : p.g.N;
    const uint64_t bytes = P * uint64_t(Q) * (fp32 ? 4u : 8u);
    if (bytes >= (uint64_t(1) << 32)) launch_scsp_w<Q, MRT, INIT, LAYOUT, true>(p, b, fp32, s);
    else launch_scsp_w<Q, MRT, INIT, LAYOUT, false>(p, b, fp32, s);
}

template <int Q, bool MRT, bool INIT>
static void launch_scsp_q(const ScspParams& p, const ScspBuffers& b, bool fp32, int layout, hipStream_t s)
{
    if (layout == DENSE_ALL) launch_scsp_l<Q, MRT, INIT, DENSE_ALL>(p, b, fp32, s);
    else if (layout == DENSE_FLUID) launch_scsp_l<Q, MRT, INIT, DENSE_FLUID>(p, b, fp32, s);
    else launch_scsp_l<Q, MRT, INIT, SPARSE>(p, b, fp32, s);
}

hipError_t launch_scsp(int Q, bool mrt, bool init, bool fp32, int layout, const ScspParams& p,
                       const ScspBuffers& b, hipStream_t s)
{
    if (Q == 19) {
        if (mrt) init ? launch_scsp_q<19, true, true>(p, b, fp32, layout, s)
                      : launch_scsp_q<19, true, false>(p, b, fp32, layout, s);
        else init ? launch_scsp_q<19, false, true>(p, b, fp32, layout, s)
                  : launch_scsp_q<19, false, false>(p, b, fp32, layout, s);
    } else {
        if (mrt) init ? launch_scsp_q<9, true, true>(p, b, fp32, layout, s)
                      : launch_scsp_q<9, true, false>(p, b, fp32, layout, s);
        else init ? launch_scsp_q<9, false, true>(p, b, fp32, layout, s)
                  : launch_scsp_q<9, false, false>(p, b, fp32, layout, s);
    }
    return hipGetLastError();
}

}  // namespace lbf
