// kernels_color32.hip -- FP32 colour-gradient time step.
//
// Same two-launch structure and operators as kernels_color.hip (reference:
// ColorFluid::calcMacroVar / bodyForce / collision / colorGradient,
// solver/src/lbm/colorfluid.cpp:238-570, 736-750), on FP32 state:
//   planes [0, Q):  h_q = fT_q - s_q, the total population shifted by phase
//                   0's rest state s_q = ai0_q rho_0 (h is the small
//                   deviation that carries the momentum);
//   planes [Q, 2Q): f0_q, the phase-0 population (the colour).
// rho = rho_0 + sum h, rho0 = sum f0, rho1 = rho - rho0.  The collision runs
// in moment space with the reference's transform M and per-site rates from
// tau (the non-uniform branch, colorfluid.cpp:374-386; with equal tau it is
// the uniform A = M^-1 S M form); the shifted equilibrium moments are
// (rho0 - rho_0) M ai0 + rho1 M ai1 + rho M w (3eu + 4.5 (eu)^2 - 1.5 u^2).
// Perturbation and recolouring act on the unshifted total fT = h + s.
// Solid sites are not streamed: the populations live either in the dense
// layout with solid slots skipped (DENSE_FLUID) or on fluid slots only
// (SPARSE, pulled through the rank index of stream.h); the indicator, the
// wall tables and the macro fields stay per site.  Macro fields for output
// and reports are written in FP64, so the component-mass and flux reductions
// run in FP64.
//
// Per-site sums (rho_0 = sum f0, rho = sum h + sum float(s_q)) and the
// recolouring share are formed in FP64 from the FP32 populations; the
// equilibrium's rest part is written without cancelling order-one terms.
// FP32 tolerance: the negative-component check uses 1e-5 rho (the FP64
// kernel's 1e-10 rho is below the rounding of the stored FP32 populations).
#include <type_traits>

#include "collide_f32.h"
#include "device.h"
#include "stream.h"
#include "zouhe.h"

namespace lbf {

namespace {

constexpr float SOLID_N32 = 86.0f;   // colorfluid.cpp:49
constexpr float NEG_TOL32 = 1e-5f;
constexpr float SQRT2_32 = 1.41421356f;
// CSF in FP32: the reference cuts the normal field at |g| <= 1e-9 and the
// curvature at |g|^2 <= 1e-18 (colorfluid.cpp:583, 592), cuts chosen for
// FP64.  The FP32 populations leave the indicator about 1e-7 off +-1 after
// the capillary shift, so pure-phase sites get |g| of order 1e-8 and, past
// the FP64 cut, a random unit normal (or, for the gradient stencil, a
// unit force direction) where FP64 has exactly 0.  The FP32 engine cuts at
// |g| <= 1e-5 instead, ten times that noise floor.
constexpr double CSF_GRAD_CUT32 = 1e-5;

template <int LAYOUT>
__device__ __forceinline__ bool resolve32(const ColorArgs& a, uint32_t& site, uint32_t& slot)
{
    const uint32_t k = xcd_block() * 256u + threadIdx.x;
    if (k >= a.P) return false;
    slot = k;
    site = LAYOUT == SPARSE ? a.list[k] : k;
    return true;
}

// PB/VB sites apply the Zou-He boundary to the pulled populations and return
// its density deviation and velocity in dzh / uzh.
struct NoPre {
    template <typename O> __device__ void operator()(const O*) const {}
};

// PRE (sparse): called with the pull offsets before the populations are
// loaded (the step issues its indicator gathers there).
template <int Q, bool INIT, int LAYOUT, typename O, bool OWN = false, typename PRE = NoPre>
__device__ __forceinline__ void load_color32(const ColorArgs& a, uint32_t site, uint32_t slot, uint32_t desc,
                                             const DevLabel& lab, float* h, float* f0, double& r0, double& dr,
                                             float& dzh, float* uzh, PRE pre = PRE())
{
    using L = Lattice<Q>;
    r0 = 0.0;
    dr = 0.0;
    if (INIT && OWN) {
        // the site's own slots of fin32, not streamed (CSF initialisation)
        constexpr bool BLK = LAYOUT == SPARSE;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            h[q] = ld(a.fin32, pop_index<2 * Q, BLK, O>(a.P, q, slot));
            f0[q] = ld(a.fin32, pop_index<2 * Q, BLK, O>(a.P, Q + q, slot));
            r0 += double(f0[q]);
            dr += double(h[q]);
        }
        return;
    }
    if (INIT) {
        if constexpr (LAYOUT == SPARSE && !std::is_same<PRE, NoPre>::value) {
            // only the neighbours' offsets (the populations are set, not pulled)
            O off[Q];
            pull_offsets<Q, LAYOUT, O, true, 2 * Q>(a.rank, desc & DESC_LINK_MASK, slot, make_stencil(a.g, site),
                                                    a.g.nx, a.P, off);
            pre(off);
        }
        // colorfluid.cpp:214-236: f[phase] = ai_phase rho_phase - 1.5 w (e.F), other phase 0
        const int ph = lab.phase;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double f = (ph == 0 ? a.dist0[q] : a.dist1[q]) - 1.5 * L::w[q] * lab.ef[q];
            h[q] = float(f - a.dist0r[q]);
            f0[q] = ph == 0 ? float(f) : 0.0f;
            r0 += double(f0[q]);
            dr += double(h[q]);
        }
        return;
    }
    // sparse: blocks of 64 slots holding the 2Q planes (stream.h pop_index)
    constexpr bool BLK = LAYOUT == SPARSE;
    const O P = a.P, f0plane = BLK ? O(Q * 64) : O(Q) * P;
    const Stencil stn = make_stencil(a.g, site);
    O off[Q];
    off[0] = pop_index<2 * Q, BLK, O>(a.P, 0, slot);
    pull_offsets<Q, LAYOUT, O, BLK, 2 * Q>(a.rank, desc & DESC_LINK_MASK, slot, stn, a.g.nx, uint32_t(P), off);
    pre(off);
    // The boundary path is chosen before the loads, so that plain sites can
    // fold their populations into the sums as they arrive.
    const bool film_links = desc & DESC_FILM;
    if (!film_links && !(lab.kind & (KIND_FILM | KIND_ZOUHE))) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            h[q] = ld(a.fin32, off[q]);
            f0[q] = ld(a.fin32, off[q] + f0plane);
            r0 += double(f0[q]);
            dr += double(h[q]);
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            h[q] = ld(a.fin32, off[q]);
            f0[q] = ld(a.fin32, off[q] + f0plane);
        }
        if (film_links) {
            // A phase that bounces back at a link into a film of the other
            // phase takes its own outgoing population of the opposite
            // direction; the other phase keeps the pulled one.  With
            // fT = h + s and f1 = fT - f0:
            //   phase 0 back: f0 = f0_loc, h = f0_loc + (h_up - f0_up);
            //   phase 1 back: f0 = f0_up,  h = f0_up + (h_loc - f0_loc).
            const uint64_t film = a.film[site];
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                const int oq = L::op[q];
                const bool back0 = (film >> oq) & 1u, back1 = (film >> (32 + oq)) & 1u;
                if (back0 || back1) {
                    const O loc = pop_index<2 * Q, BLK, O>(a.P, oq, slot);
                    const float h_loc = ld(a.fin32, loc), f0_loc = ld(a.fin32, loc + f0plane);
                    if (back0) {
                        h[q] = f0_loc + (h[q] - f0[q]);
                        f0[q] = f0_loc;
                    } else {
                        h[q] = f0[q] + (h_loc - f0_loc);
                    }
                }
            }
        }
        if (lab.kind & KIND_FILM) {
            // clearHalfFilm (colorfluid.cpp:752-764): phase-0 film: f1 = 0,
            // so fT = f0; phase-1 film: f0 = 0, so fT = f1 = h + s - f0
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (lab.phase == 0) {
                    h[q] = f0[q] - a.dist0f[q];
                } else {
                    h[q] = h[q] - f0[q];
                    f0[q] = 0.0f;
                }
            }
        }
        if (lab.kind & KIND_ZOUHE)
            zouhe_color_f32<Q>(h, f0, a.zh[desc >> DESC_LABEL_SHIFT], lab.phase, a.rho_reff, a.dist0, dzh, uzh);
        r0 = 0.0;
        dr = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            r0 += double(f0[q]);
            dr += double(h[q]);
        }
    }
}

// Colour gradient of a fluid site from the FP32 indicator (ColorFluid::
// colorGradient, colorfluid.cpp:435-570): bulk stencil, or at wall sites the
// one-sided stencil, the renormalized or tanh reconstruction and the
// contact-angle rotation.  Returns |g|.
//
// dens(q) is the indicator at x + e_q (q >= 1; SOLID_N32 on solid sites)
// and dens(0) the site's own; wi is the site's wall entry (0: bulk).
template <int Q, typename DN>
__device__ __forceinline__ float color_gradient_f32(const ColorArgs& a, int32_t wi, DN dens, float* g)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    g[0] = g[1] = g[2] = 0.0f;
    float norm;
    if (!wi) {
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < L::NREL; ++k) {
                const int q = L::reld[d][k];
                const float dnq = dens(q);
                if (L::e[q][d] == 1) g[d] += float(L::w[q]) * dnq;
                else g[d] -= float(L::w[q]) * dnq;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] *= 3.0f;
        float n2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) n2 += g[d] * g[d];
        norm = sqrtf(n2);
    } else {
        const float dn0 = dens(0);
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < L::NREL; ++k) {
                const int q = L::reld[d][k];
                const float dnq = dens(q);
                if (dnq == SOLID_N32) continue;
                if (L::e[q][d] == 1) g[d] += float(L::w[q]) * (dnq - dn0);
                else g[d] -= float(L::w[q]) * (dnq - dn0);
            }
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] *= 3.0f;
        float n2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) n2 += g[d] * g[d];
        norm = sqrtf(n2);
        const WallF32& wl = a.walls32[wi];
        if (a.renormalized) {
            // wallGradient renormalized (colorfluid.cpp:487-492): g = inverse g / 3
            const float* gi = a.ginv32 + size_t(wi) * 9;
            float r[3];
#pragma unroll
            for (int c = 0; c < D; ++c) {
                float t = 0.0f;
#pragma unroll
                for (int b = 0; b < D; ++b) t += g[b] * gi[c * D + b];
                r[c] = t;
            }
#pragma unroll
            for (int c = 0; c < D; ++c) g[c] = r[c] * (1.0f / 3.0f);
            n2 = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) n2 += g[d] * g[d];
            norm = sqrtf(n2);
        }
        if (a.tanh_wall) {
            // wallGradient tanh: the reconstruction of color_tanh_wall_f32
#pragma unroll
            for (int d = 0; d < D; ++d) g[d] = float(a.tanhg[uint64_t(wi) * 3 + d]);
            n2 = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) n2 += g[d] * g[d];
            norm = sqrtf(n2);
        }
        const float cA = wl.cosca, sA = wl.sinca;
        if constexpr (D == 3) {
            // the inverse frame is the transpose: Ti[r][c] = T[c][r]
            float T[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) T[k] = wl.tib[k];
            float gn[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) gn[r] = T[r * 3] * g[0] + T[r * 3 + 1] * g[1] + T[r * 3 + 2] * g[2];
            // RotateTx3D(-gn0, gn1) (fn.cpp:130-146)
            float c = 1.0f, sn = 0.0f;
            const float so = -gn[0], sa = gn[1];
            const float hyp = sqrtf(so * so + sa * sa);
            if (hyp != 0.0f) {
                c = sa / hyp;
                sn = so / hyp;
            }
            const float tx[9] = {c, -sn, 0.0f, sn, c, 0.0f, 0.0f, 0.0f, 1.0f};
            const float hv[3] = {0.0f, norm * sA, norm * cA};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float s = 0.0f;
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const float tic = T[r] * tx[cc] + T[3 + r] * tx[3 + cc] + T[6 + r] * tx[6 + cc];
                    s += tic * hv[cc];
                }
                g[r] = s;
            }
        } else {
            const float gn0 = wl.tib[0] * g[0] + wl.tib[1] * g[1];
            const float hv0 = gn0 >= 0.0f ? norm * sA : -norm * sA, hv1 = norm * cA;
            const float g0 = wl.tib[0] * hv0 + wl.tib[2] * hv1;   // the inverse frame is the transpose
            const float g1 = wl.tib[1] * hv0 + wl.tib[3] * hv1;
            g[0] = g0;
            g[1] = g1;
        }
    }

    return norm;
}

// Capillary force of a VD/GV site (ColorFluid::calcCapillaryForce,
// colorfluid.cpp:591-617), in FP64 from the stored gradient and the
// curvature field of the fluid neighbours (the FP32 indicator marks solids).
template <int Q>
__device__ __forceinline__ bool capillary_force_f32(const ColorArgs& a, uint32_t i, int x, int y, int z, float* F)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint64_t N = a.g.N;
    const bool gc = a.curv_gradient;
    double grad[3] = {0.0, 0.0, 0.0}, field[3] = {0.0, 0.0, 0.0};
    double curvature = 0.0, norm2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        grad[d] = a.capgrad[uint64_t(d) * N + i];
        field[d] = a.capfield[uint64_t(d) * N + i];
        norm2 += grad[d] * grad[d];
    }
    const bool curved = norm2 > CSF_GRAD_CUT32 * CSF_GRAD_CUT32;
    if (curved) {
        const int32_t w = a.wall[i];
        uint32_t nb[Q];
        bool fluid[Q];
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            nb[q] = push_target<Q>(a.g, x, y, z, q);
            fluid[q] = a.densN32[nb[q]] != SOLID_N32;
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double moment[3] = {0.0, 0.0, 0.0}, derivative[3];
            const double center = field[c];
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                if (!fluid[q]) continue;
                const double difference = a.capfield[uint64_t(c) * N + nb[q]] - center;
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    if (L::e[q][b] == 1) moment[b] += L::w[q] * difference;
                    else if (L::e[q][b] == -1) moment[b] -= L::w[q] * difference;
                }
            }
            if (w) {
                const double* inv = a.walls[w].ginv;
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    double t = 0.0;
#pragma unroll
                    for (int b = 0; b < D; ++b) t = t + moment[b] * inv[r * D + b];
                    derivative[r] = t;
                }
            } else {
#pragma unroll
                for (int b = 0; b < D; ++b) derivative[b] = moment[b] * 3.0;
            }
#pragma unroll
            for (int b = 0; b < D; ++b) {
                const double na = gc ? grad[c] / sqrt(norm2) : center;
                const double nbv = gc ? grad[b] / sqrt(norm2) : field[b];
                curvature += (na * nbv - (c == b ? 1.0 : 0.0)) * derivative[b];
            }
        }
    }
    bool finite = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double direction = gc && curved ? grad[d] / sqrt(norm2) : grad[d];
        F[d] = float(.5 * a.sigma * curvature * direction);
        finite = finite && isfinite(F[d]);
    }
    return finite;
}

}  // namespace

template <int Q, bool INIT, int LAYOUT, typename O, bool OWN = false>
__global__ void __launch_bounds__(256) color_density_f32(ColorArgs a)
{
    uint32_t site, slot;
    if (!resolve32<LAYOUT>(a, site, slot)) return;
    const uint32_t desc = a.desc[slot];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    if (!(lab.kind & KIND_FLUID)) return;
    float h[Q], f0[Q], dzh, uzh[Lattice<Q>::D];
    double r0, dr;
    load_color32<Q, INIT, LAYOUT, O, OWN>(a, site, slot, desc, lab, h, f0, r0, dr, dzh, uzh);
    const double rho = dr + a.rest32;
    const float chi = float((2.0 * r0 - rho) / rho);   // (rho0 - rho1) / rho
    // the sparse step reads the indicator by slot; the CSF and tanh-wall
    // passes read it per voxel
    if (LAYOUT == SPARSE) a.dens_slot[slot] = chi;
    if (LAYOUT != SPARSE || a.csf || a.tanh_wall) a.densN32[site] = chi;
}

template <int Q, bool MRT, bool INIT, int LAYOUT, typename O, bool CSF = false, bool OWN = false>
__global__ void __launch_bounds__(256) color_step_f32(ColorArgs a)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    constexpr int H = L::NHALF;
    constexpr MomentTables<Q> MT = kMT<Q>;
    uint32_t i, slot;
    if (!resolve32<LAYOUT>(a, i, slot)) return;
    const uint64_t N = a.g.N;
    const uint32_t desc = a.desc[slot];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    const int kind = lab.kind;
    if (!(kind & KIND_FLUID)) return;

    float h[Q], f0[Q], dzh = 0.0f, uzh[D];
    double r0d, drd;
    // sparse: the colour gradient from the indicator per slot, at the
    // neighbours the pull has just located (x + e_q is the pull source of
    // op(q); solid where link q bounces back)
    constexpr bool SLOT_GRADIENT = LAYOUT == SPARSE && !CSF;
    float g[3] = {0.0f, 0.0f, 0.0f};
    float norm = 0.0f;
    if constexpr (SLOT_GRADIENT) {
        // The indicator at the 18 neighbours is gathered with the pull
        // offsets, before the populations are loaded, and the gradient is
        // formed once both have arrived: one memory round trip for the site
        // (a gradient formed before the population loads waited on its own
        // gathers and on the wall entry first).  Neighbour x + e_q is the
        // pull source of op(q), its slot recovered from the blocked offset
        // (block, lane); solid where link q bounces back.
        const uint32_t links = desc & DESC_LINK_MASK;
        const int32_t wi = a.wall_slot[slot];
        float dn[Q];
        auto gather = [&](const O* off) __attribute__((always_inline)) {
            dn[0] = a.dens_slot[slot];
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                const O o = off[L::op[q]];
                const uint32_t m = (uint32_t(o / O(2 * Q * 64)) << 6) | uint32_t(o & O(63));
                dn[q] = ((links >> q) & 1u) ? SOLID_N32 : a.dens_slot[m];
            }
        };
        load_color32<Q, INIT, LAYOUT, O, OWN>(a, i, slot, desc, lab, h, f0, r0d, drd, dzh, uzh, gather);
        norm = color_gradient_f32<Q>(a, wi, [&](int q) __attribute__((always_inline)) { return dn[q]; }, g);
    } else {
        load_color32<Q, INIT, LAYOUT, O, OWN>(a, i, slot, desc, lab, h, f0, r0d, drd, dzh, uzh);
    }
    int cx, cy, cz;
    site_coords(a.g, i, cx, cy, cz);

    // calcMacroVar (colorfluid.cpp:238-290).  The site's density is the sum
    // of its totals fT_q = h_q + float(s_q), formed in FP64 from the FP32
    // populations (rho = sum h + sum float(s_q)); rho_1 = rho - rho_0.
    const double rtd = drd + a.rest32, r1d = rtd - r0d;
    const float r0 = float(r0d), rt = float(rtd), r1 = float(r1d);
    if (!isfinite(r0d) || !isfinite(rtd) || rtd <= 0.0) raise_status(a.status, 0, a.t);
    else if (r0d < -NEG_TOL32 * rtd || r1d < -NEG_TOL32 * rtd) raise_status(a.status, 1, a.t);
    const float inv_rt = 1.0f / rt;
    const float dn = (r0 - r1) * inv_rt;
    float tau = (1.0f + dn) * a.relax0f + (1.0f - dn) * a.relax1f;
    if (a.harmonic) {
        const float phi = r0 * inv_rt;
        tau = 0.5f + 3.0f / (phi / a.nu0f + (1.0f - phi) / a.nu1f);
    }
    if (!isfinite(tau) || tau <= 0.5f) raise_status(a.status, 2, a.t);
    // bodyForce (colorfluid.cpp:736-750): under CSF first (F_cap/2)/rho on
    // VD/GV sites, then (F/2)/rho on GV sites
    float fcap[3] = {0.0f, 0.0f, 0.0f};
    if constexpr (CSF) {
        if (!capillary_force_f32<Q>(a, i, cx, cy, cz, fcap)) raise_status(a.status, 5, a.t);
    }
    float u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        float s = 0.0f;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (L::e[q][d] == 1) s += h[q];
            else if (L::e[q][d] == -1) s -= h[q];
        }
        u[d] = s * inv_rt;
        if constexpr (CSF) u[d] += 0.5f * fcap[d] * inv_rt;
        if (kind & KIND_GV) u[d] += float(lab.half[d]) * inv_rt;
        if (!isfinite(u[d])) raise_status(a.status, 3, a.t);
    }
    if (a.flags & 1u) {
        a.rho0[i] = r0d;
        a.rho1[i] = r1d;
#pragma unroll
        for (int d = 0; d < D; ++d) a.u[uint64_t(d) * N + i] = double(u[d]);
    }
    // adjacent report history; PB/VB sites as in kernels_color.hip
    if (a.flags & 2u) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj[uint64_t(d) * N + i] = double(u[d]);
        a.chiadj[i] = (kind & KIND_ZOUHE) ? r0d - r1d : double(dn);
    }
    if (!INIT && (a.flags & 1u) && (kind & KIND_ZOUHE)) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj_prev[uint64_t(d) * N + i] = double(uzh[d]);
        const double rzh = a.rho_ref + double(dzh);
        const double num = a.chiadj_prev[i];
        a.chiadj_prev[i] = rzh > 0.0 ? num / rzh : 0.0;
    }

    // collision of the total population (colorfluid.cpp:292-386), on h;
    // under CSF the force is the label's plus the capillary force
    bool forced = kind & KIND_FORCED;
    float F[3] = {0.0f, 0.0f, 0.0f};
    if (forced)
#pragma unroll
        for (int d = 0; d < D; ++d) F[d] = float(lab.force[d]);
    if constexpr (CSF) {
        float f2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            F[d] += fcap[d];
            f2 += F[d] * F[d];
        }
        forced = f2 != 0.0f;
    }
    float usq = 0.0f, uF = 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        usq += u[d] * u[d];
        uF += u[d] * F[d];
    }
    // Rest part of the shifted equilibrium, r0 ai0 + r1 ai1 - float(s_q), as
    // (rho - rho_0) ai0 + r1 (ai1 - ai0) + (s_q - float(s_q)): no term of
    // order one that cancels, so FP32 keeps its digits, and its sum over q is
    // exactly sum h (the SRT collision adds no mass).
    const float drest = float(rtd - a.rho_ref);
    if constexpr (!MRT) {
        const float inv_tau = 1.0f / tau, cg = 1.0f - 0.5f * inv_tau;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float eu = 0.0f, ef = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) { eu += u[d]; ef += F[d]; }
                else if (L::e[q][d] == -1) { eu -= u[d]; ef -= F[d]; }
            }
            const float w = float(L::w[q]);
            const float heq = drest * a.ai0f[q] + r1 * a.daif[q] + a.rest_errf[q] +
                              rt * w * (eu * (3.0f + 4.5f * eu) - 1.5f * usq);
            h[q] += (heq - h[q]) * inv_tau;
            if (forced) h[q] += cg * w * (3.0f * (ef - uF) + 9.0f * eu * ef);
        }
    } else {
        const float su = 1.0f / tau;
        const float sq = 8.0f * (2.0f - su) / (8.0f - su);
        float sp[H], sm[H];
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const int q = L::half[j];
            sp[j] = h[q] + h[L::op[q]];
            sm[j] = h[q] - h[L::op[q]];
        }
        float y[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            float mg = 0.0f;
            if (forced) {
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    if (MT.T1[k][p] != 0.0) mg += float(3.0 * MT.T1[k][p]) * F[p];
#pragma unroll
                    for (int b = 0; b < D; ++b)
                        if (MT.T2[k][p][b] != 0.0) mg += float(9.0 * MT.T2[k][p][b]) * u[p] * F[b];
                }
                if (MT.T0[k] != 0.0) mg -= float(3.0 * MT.T0[k]) * uF;
            }
            if (MT.rate[k] == 0) {
                y[k] = mg * float(1.0 / MT.n2[k]);
                continue;
            }
            float m = 0.0f;
            if (MT.parity[k] > 0) {
                if (L::M[k][0] != 0) m += float(L::M[k][0]) * h[0];
#pragma unroll
                for (int j = 0; j < H; ++j)
                    if (L::M[k][L::half[j]] != 0) m += float(L::M[k][L::half[j]]) * sp[j];
            } else {
#pragma unroll
                for (int j = 0; j < H; ++j)
                    if (L::M[k][L::half[j]] != 0) m += float(L::M[k][L::half[j]]) * sm[j];
            }
            float meq = 0.0f;
#pragma unroll
            for (int p = 0; p < D; ++p) {
                if (MT.T1[k][p] != 0.0) meq += float(3.0 * MT.T1[k][p]) * u[p];
#pragma unroll
                for (int b = 0; b < D; ++b)
                    if (MT.T2[k][p][b] != 0.0) meq += float(4.5 * MT.T2[k][p][b]) * u[p] * u[b];
            }
            if (MT.T0[k] != 0.0) meq -= float(1.5 * MT.T0[k]) * usq;
            meq = meq * rt + drest * a.mai0f[k] + r1 * a.mdaif[k] + a.mrest_errf[k];
            const float s = MT.rate[k] == 1 ? su : sq;
            y[k] = (s * (meq - m) + (1.0f - 0.5f * s) * mg) * float(1.0 / MT.n2[k]);
        }
        {
            float e0 = 0.0f;
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (MT.parity[k] > 0 && L::M[k][0] != 0) e0 += float(L::M[k][0]) * y[k];
            h[0] += e0;
        }
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const int q = L::half[j];
            float ev = 0.0f, od = 0.0f;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (L::M[k][q] == 0) continue;
                if (MT.parity[k] > 0) ev += float(L::M[k][q]) * y[k];
                else od += float(L::M[k][q]) * y[k];
            }
            h[q] += ev + od;
            h[L::op[q]] += ev - od;
        }
    }

    // colour gradient: the stored capillary gradient under CSF (no
    // perturbation, colorfluid.cpp:389-411), else colorGradient (sparse:
    // computed above)
    if constexpr (CSF) {
        float n2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            g[d] = float(a.capgrad[uint64_t(d) * N + i]);
            n2 += g[d] * g[d];
        }
        norm = sqrtf(n2);
    } else if constexpr (!SLOT_GRADIENT) {
        norm = color_gradient_f32<Q>(a, a.wall[i], [&](int q) {
            return a.densN32[q == 0 ? i : push_target<Q>(a.g, cx, cy, cz, q)];
        }, g);
    }

    // perturbation + recolouring (colorfluid.cpp:389-428).  Phase 0 takes the
    // share r0 / rho of the total fT_q = h_q + float(s_q) of each direction,
    // formed in FP64: rho is the sum of exactly those totals, so that
    // sum_q f0_q = r0 up to unbiased rounding of the stored values (near 1 an
    // FP32 rho has only 6e-8 resolution, and with rho_0 + sum h in place of
    // that sum the phase-0 mass grew by a factor 1 + 1.5e-8 per step).
    const double dens0T64 = r0d / rtd;
    const float dens0T = float(dens0T64);
    bool finite = true;
    if (norm > 1e-9f) {
        const float b01 = a.betaf * dens0T * r1 * inv_rt;
        const float akn = a.akf * norm / tau;
        const float inv_norm = 1.0f / norm;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float angle = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) angle += g[d];
                else if (L::e[q][d] == -1) angle -= g[d];
            }
            angle *= inv_norm;
            if constexpr (!CSF) h[q] += akn * (float(L::w[q]) * angle * angle - a.bif[q]);
            if (L::diagonal(q)) angle *= (1.0f / SQRT2_32);
            const float feq0 = r0 * a.ai0f[q] + r1 * a.ai1f[q];
            const float fo = float(fma(dens0T64, double(h[q]) + a.dist0r[q], double(b01 * angle * feq0)));
            finite = finite && isfinite(h[q]) && isfinite(fo);
            st_pop<LAYOUT == SPARSE>(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, q, slot), h[q]);
            st_pop<LAYOUT == SPARSE>(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, Q + q, slot), fo);
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float fo = float(dens0T64 * (double(h[q]) + a.dist0r[q]));
            finite = finite && isfinite(h[q]) && isfinite(fo);
            st_pop<LAYOUT == SPARSE>(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, q, slot), h[q]);
            st_pop<LAYOUT == SPARSE>(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, Q + q, slot), fo);
        }
    }
    if (!finite) raise_status(a.status, 4, a.t);
}

// wallGradient tanh at the wall sites (colorfluid.cpp:494-518), from the FP32
// colour indicator but in FP64 (the clamp at 1 - 1e-8 needs it): atanh(chi)
// fitted as affine over the fluid neighbours, tanh continued over the full
// stencil, the bulk gradient of that stored per wall entry.
template <int Q, int LAYOUT>
__global__ void __launch_bounds__(256) color_tanh_wall_f32(ColorArgs a)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    uint32_t site, slot;
    if (!resolve32<LAYOUT>(a, site, slot)) return;
    const int32_t w = a.wall[site];
    if (!w || !(a.lab[a.desc[slot] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) return;
    int x, y, z;
    site_coords(a.g, site, x, y, z);
    const DevWall& wl = a.walls[w];
    const double center = tanh_wall_transform(double(a.densN32[site]));
    double moment[3] = {0.0, 0.0, 0.0}, slope[3], g[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const float nb = a.densN32[push_target<Q>(a.g, x, y, z, q)];
        if (nb == SOLID_N32) continue;
        const double difference = tanh_wall_transform(double(nb)) - center;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) moment[d] += L::w[q] * difference;
            else if (L::e[q][d] == -1) moment[d] -= L::w[q] * difference;
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < D; ++b) t = t + moment[b] * wl.ginv[c * D + b];
        slope[c] = t;
    }
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        double position = center;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) position += slope[d];
            else if (L::e[q][d] == -1) position += -slope[d];
        }
        const double color = tanh(position);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) g[d] += 3.0 * L::w[q] * color;
            else if (L::e[q][d] == -1) g[d] -= 3.0 * L::w[q] * color;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) a.tanhg[uint64_t(w) * 3 + d] = g[d];
}

// CSF, first half of calcCapillaryForce (colorfluid.cpp:577-586): colour
// gradient and curvature field of every VD/GV site, stored in FP64 per site.
template <int Q, int LAYOUT>
__global__ void __launch_bounds__(256) color_capillary_f32(ColorArgs a)
{
    constexpr int D = Lattice<Q>::D;
    uint32_t site, slot;
    if (!resolve32<LAYOUT>(a, site, slot)) return;
    if (!(a.lab[a.desc[slot] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) return;
    const uint64_t N = a.g.N;
    int x, y, z;
    site_coords(a.g, site, x, y, z);
    float g[3];
    const float norm = color_gradient_f32<Q>(a, a.wall[site], [&](int q) {
        return a.densN32[q == 0 ? site : push_target<Q>(a.g, x, y, z, q)];
    }, g);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        a.capgrad[uint64_t(d) * N + site] = double(g[d]);
        a.capfield[uint64_t(d) * N + site] =
            a.curv_gradient ? double(g[d]) : (norm > float(CSF_GRAD_CUT32) ? double(g[d] / norm) : 0.0);
    }
}

// CSF initialisation (ColorFluid::initializeCapillaryForce,
// colorfluid.cpp:619-634): the t = 0 populations of initialize() with
// f_k -= 1.5 w (e . F) rho_k / rho, written to the site's own slots of fout32.
template <int Q, int LAYOUT, typename O>
__global__ void __launch_bounds__(256) color_csf_init_f32(ColorArgs a)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    uint32_t site, slot;
    if (!resolve32<LAYOUT>(a, site, slot)) return;
    const DevLabel& lab = a.lab[a.desc[slot] >> DESC_LABEL_SHIFT];
    if (!(lab.kind & KIND_FLUID)) return;
    const int ph = lab.phase;
    double f0[Q], f1[Q], r0 = 0.0, r1 = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double f = (ph == 0 ? a.dist0[q] : a.dist1[q]) - 1.5 * L::w[q] * lab.ef[q];
        f0[q] = ph == 0 ? f : 0.0;
        f1[q] = ph == 0 ? 0.0 : f;
        r0 += f0[q];
        r1 += f1[q];
    }
    const double rt = r0 + r1;
    int x, y, z;
    site_coords(a.g, site, x, y, z);
    float F[3] = {0.0f, 0.0f, 0.0f};
    if (!capillary_force_f32<Q>(a, site, x, y, z, F)) raise_status(a.status, 5, a.t);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double ef = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) ef += double(F[d]);
            else if (L::e[q][d] == -1) ef -= double(F[d]);
        }
        const double g0 = f0[q] - 1.5 * L::w[q] * ef * r0 / rt;
        const double g1 = f1[q] - 1.5 * L::w[q] * ef * r1 / rt;
        st(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, q, slot), float(g0 + g1 - a.dist0r[q]));
        st(a.fout32, pop_index<2 * Q, LAYOUT == SPARSE, O>(a.P, Q + q, slot), float(g0));
    }
}

template <int Q, bool MRT, bool INIT, int LAYOUT, typename O>
static hipError_t launch_color32_l(const ColorArgs& a, hipStream_t s)
{
    const dim3 grid(uint32_t((a.P + 255) / 256)), block(256);
    if (a.P == 0) return hipSuccess;
    if (a.csf) {
        // CSF (colorfluid.cpp:124-137): density, [tanh walls,] gradient and
        // curvature field, step; t = 0: the capillary shift of the initial
        // populations, then density, field and collision in place on them
        hipLaunchKernelGGL((color_density_f32<Q, INIT, LAYOUT, O>), grid, block, 0, s, a);
        if (a.tanh_wall) hipLaunchKernelGGL((color_tanh_wall_f32<Q, LAYOUT>), grid, block, 0, s, a);
        hipLaunchKernelGGL((color_capillary_f32<Q, LAYOUT>), grid, block, 0, s, a);
        if (!INIT) {
            hipLaunchKernelGGL((color_step_f32<Q, MRT, false, LAYOUT, O, true>), grid, block, 0, s, a);
            return hipGetLastError();
        }
        hipLaunchKernelGGL((color_csf_init_f32<Q, LAYOUT, O>), grid, block, 0, s, a);
        ColorArgs b = a;
        b.fin32 = a.fout32;
        hipLaunchKernelGGL((color_density_f32<Q, true, LAYOUT, O, true>), grid, block, 0, s, b);
        if (b.tanh_wall) hipLaunchKernelGGL((color_tanh_wall_f32<Q, LAYOUT>), grid, block, 0, s, b);
        hipLaunchKernelGGL((color_capillary_f32<Q, LAYOUT>), grid, block, 0, s, b);
        hipLaunchKernelGGL((color_step_f32<Q, MRT, true, LAYOUT, O, true, true>), grid, block, 0, s, b);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((color_density_f32<Q, INIT, LAYOUT, O>), grid, block, 0, s, a);
    if (a.tanh_wall) hipLaunchKernelGGL((color_tanh_wall_f32<Q, LAYOUT>), grid, block, 0, s, a);
    hipLaunchKernelGGL((color_step_f32<Q, MRT, INIT, LAYOUT, O>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <int Q, bool MRT, bool INIT>
static hipError_t launch_color32_q(const ColorArgs& a, bool sparse, hipStream_t s)
{
    // 32-bit element offsets while both plane sets fit below 4 GiB (sparse:
    // the last block of 64 slots is allocated whole)
    const uint64_t P = sparse ? (uint64_t(a.P) + 63) / 64 * 64 : a.P;
    const bool wide = uint64_t(2) * Q * P * 4u >= (uint64_t(1) << 32);
    if (sparse)
        return wide ? launch_color32_l<Q, MRT, INIT, SPARSE, uint64_t>(a, s)
                    : launch_color32_l<Q, MRT, INIT, SPARSE, uint32_t>(a, s);
    return wide ? launch_color32_l<Q, MRT, INIT, DENSE_FLUID, uint64_t>(a, s)
                : launch_color32_l<Q, MRT, INIT, DENSE_FLUID, uint32_t>(a, s);
}

hipError_t launch_color32(int Q, bool mrt, bool init, bool sparse, const ColorArgs& a, hipStream_t s)
{
    if (Q == 19) {
        if (mrt) return init ? launch_color32_q<19, true, true>(a, sparse, s)
                             : launch_color32_q<19, true, false>(a, sparse, s);
        return init ? launch_color32_q<19, false, true>(a, sparse, s) : launch_color32_q<19, false, false>(a, sparse, s);
    }
    if (mrt) return init ? launch_color32_q<9, true, true>(a, sparse, s) : launch_color32_q<9, true, false>(a, sparse, s);
    return init ? launch_color32_q<9, false, true>(a, sparse, s) : launch_color32_q<9, false, false>(a, sparse, s);
}

}  // namespace lbf
