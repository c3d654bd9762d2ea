// kernels_color.hip -- two-phase colour-gradient time step on CDNA4.
//
// The colour gradient of a site needs the colour indicator rhoN of its 18
// neighbours *after* streaming, so a step is two launches:
//   A (color_density): pull both phases, write rhoN = (rho0-rho1)/rho per site
//     (solid sites keep the reference's sentinel 86);
//   B (color_step): pull both phases again, macro + tau + body force,
//     SRT/MRT collision of the total population, colour gradient from the
//     rhoN stencil with the geometric contact-angle rotation at wall sites,
//     perturbation and recolouring, write both phases.
// Streaming is the same pull with periodic wrap and halfway bounce-back as
// the single-phase kernel.  tanh walls add a wall pass after A
// (color_tanh_wall); the CSF capillary force adds a pass between A and B
// (color_capillary) that writes the colour gradient and curvature field.
//
// Reference operators (zhou1800/lbflow):
//   initialize              solver/src/lbm/colorfluid.cpp:214-236
//   calcMacroVar            solver/src/lbm/colorfluid.cpp:238-290
//   bodyForce               solver/src/lbm/colorfluid.cpp:736-750
//   collision               solver/src/lbm/colorfluid.cpp:292-433
//   colorGradient (legacy / renormalized / tanh walls)  solver/src/lbm/colorfluid.cpp:435-570
//   calcCapillaryForce, initializeCapillaryForce (CSF)  solver/src/lbm/colorfluid.cpp:572-634
//   RotateTx3D etc.         solver/src/algorithm/fn.cpp:118-164
//   zouHeBoundary (PB/VB)   solver/src/lbm/colorfluid.cpp:638-734 (zouhe.h)
//   clearHalfFilm (HF)      solver/src/lbm/colorfluid.cpp:752-764
//
// FP64, reference operation order (built with -ffp-contract=off); zero
// lattice coefficients are skipped exactly as in kernels_scsp.hip.
#include "device.h"
#include "zouhe.h"

namespace lbf {

constexpr double SOLID_DENS_N = 86.0;   // colorfluid.cpp:49
constexpr double SQRT2 = 1.4142135623730951;  // sqrt(2.0), Qnorm of diagonal links

namespace {

// Pull (or initialise) both phases; returns the total fT_q = f0_q + f1_q and
// the phase densities summed in direction order.  PB/VB sites apply the
// Zou-He boundary to the pulled populations first and return its density and
// velocity in rzh / uzh.  OWN (with INIT): the site's own populations in fin,
// not streamed (the capillary initialisation's second macro pass).
template <int Q, bool INIT, bool OWN = false>
__device__ __forceinline__ void load_color(const ColorArgs& a, uint32_t i, uint32_t desc,
                                           const DevLabel& lab, double* fT, double& r0, double& r1,
                                           double& rzh, double* uzh)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    const uint64_t N = a.g.N;
    if (INIT && OWN) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double f0 = a.fin[uint64_t(q) * N + i];
            const double f1 = a.fin[(uint64_t(Q) + q) * N + i];
            if (q == 0) { r0 = f0; r1 = f1; }
            else { r0 += f0; r1 += f1; }
            fT[q] = f0 + f1;
        }
        return;
    }
    if (INIT) {
        const int ph = lab.phase;
        const double* dist = ph == 0 ? a.dist0 : a.dist1;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double f = dist[q];
            f -= 1.5 * L::w[q] * lab.ef[q];
            const double f0 = ph == 0 ? f : 0.0;
            const double f1 = ph == 0 ? 0.0 : f;
            if (q == 0) { r0 = f0; r1 = f1; }
            else { r0 += f0; r1 += f1; }
            fT[q] = f0 + f1;
        }
        return;
    }
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    const uint32_t links = desc & DESC_LINK_MASK;
    if ((lab.kind & (KIND_ZOUHE | KIND_FILM)) || (desc & DESC_FILM)) {
        // Boundary sites, in the reference's order: bounce-back at solid
        // links, then each phase bounces back at links into a film of the
        // other phase, the film empties the other phase (clearHalfFilm,
        // colorfluid.cpp:752-764), then the Zou-He boundary.
        const uint64_t film = (desc & DESC_FILM) ? a.film[i] : 0;
        double f0[Q], f1[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            uint64_t src;
            int qs = q;
            if (q == 0) src = i;
            else if ((links >> L::op[q]) & 1u) { src = i; qs = L::op[q]; }
            else src = pull_source<Q>(a.g, x, y, z, q);
            const bool back0 = q != 0 && ((film >> L::op[q]) & 1u);
            const bool back1 = q != 0 && ((film >> (32 + L::op[q])) & 1u);
            f0[q] = a.fin[uint64_t(back0 ? L::op[q] : qs) * N + (back0 ? i : src)];
            f1[q] = a.fin[(uint64_t(Q) + (back1 ? L::op[q] : qs)) * N + (back1 ? i : src)];
        }
        if (lab.kind & KIND_FILM) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (lab.phase == 0) f1[q] = 0.0;
                else f0[q] = 0.0;
            }
        }
        if (lab.kind & KIND_ZOUHE) zouhe_color_f64<Q>(f0, f1, a.zh[desc >> DESC_LABEL_SHIFT], lab.phase, rzh, uzh);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (q == 0) { r0 = f0[q]; r1 = f1[q]; }
            else { r0 += f0[q]; r1 += f1[q]; }
            fT[q] = f0[q] + f1[q];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint64_t src;
        int qs = q;
        if (q == 0) src = i;
        else if ((links >> L::op[q]) & 1u) { src = i; qs = L::op[q]; }
        else src = pull_source<Q>(a.g, x, y, z, q);
        const double f0 = a.fin[uint64_t(qs) * N + src];
        const double f1 = a.fin[(uint64_t(Q) + qs) * N + src];
        if (q == 0) { r0 = f0; r1 = f1; }
        else { r0 += f0; r1 += f1; }
        fT[q] = f0 + f1;
    }
}

// Colour gradient of a fluid site from the rhoN stencil (ColorFluid::
// colorGradient, colorfluid.cpp:435-570): bulk stencil, or at wall sites the
// one-sided stencil over fluid neighbours, the renormalized or tanh
// reconstruction, and the geometric contact-angle rotation.  Returns |g|.
template <int Q>
__device__ __forceinline__ double color_gradient_f64(const ColorArgs& a, uint32_t i, int x, int y, int z,
                                                     double* g)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    g[0] = g[1] = g[2] = 0.0;
    double norm;
    const int32_t w = a.wall[i];
    if (!w) {
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < L::NREL; ++k) {
                const int q = L::reld[d][k];
                const double dnq = a.densN[push_target<Q>(a.g, x, y, z, q)];
                if (L::e[q][d] == 1) g[d] += L::w[q] * dnq;
                else g[d] -= L::w[q] * dnq;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] *= 3.0;
        double n2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
        norm = sqrt(n2);
    } else {
        const double dn0 = a.densN[i];
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < L::NREL; ++k) {
                const int q = L::reld[d][k];
                const double dnq = a.densN[push_target<Q>(a.g, x, y, z, q)];
                if (dnq == SOLID_DENS_N) continue;
                if (L::e[q][d] == 1) g[d] += L::w[q] * (dnq - dn0);
                else g[d] -= L::w[q] * (dnq - dn0);
            }
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] *= 3.0;
        double n2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
        norm = sqrt(n2);
        const DevWall& wl = a.walls[w];
        if (a.renormalized) {
            // wallGradient renormalized (colorfluid.cpp:487-492): g = inverse g / 3
            double r[3];
#pragma unroll
            for (int c = 0; c < D; ++c) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < D; ++b) t = t + g[b] * wl.ginv[c * D + b];
                r[c] = t;
            }
#pragma unroll
            for (int c = 0; c < D; ++c) g[c] = r[c] / 3.0;
            n2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
            norm = sqrt(n2);
        }
        if (a.tanh_wall) {
            // wallGradient tanh: the reconstruction of color_tanh_wall
#pragma unroll
            for (int d = 0; d < D; ++d) g[d] = a.tanhg[uint64_t(w) * 3 + d];
            n2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) n2 = n2 + g[d] * g[d];
            norm = sqrt(n2);
        }
        const double cA = wl.cosca, sA = wl.sinca;
        if constexpr (D == 3) {
            const double* T = wl.tib;
            const double* Ti = wl.tibinv;
            double gn[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) s = s + T[r * 3 + k] * g[k];
                gn[r] = s;
            }
            // RotateTx3D(-gn0, gn1)
            double c = 1, sn = 0;
            const double so = -gn[0], sa = gn[1];
            const double hyp = sqrt(so * so + sa * sa);
            if (hyp != 0) { c = sa / hyp; sn = so / hyp; }
            const double tx[9] = {c, -sn, 0, sn, c, 0, 0, 0, 1};
            double ti[9];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) s = s + Ti[r * 3 + k] * tx[k * 3 + cc];
                    ti[r * 3 + cc] = s;
                }
            const double h[3] = {0, norm * sA, norm * cA};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) s = s + ti[r * 3 + k] * h[k];
                g[r] = s;
            }
        } else {
            const double* T = wl.tib;
            const double* Ti = wl.tibinv;
            double gn0 = 0.0;
            gn0 = gn0 + T[0] * g[0];
            gn0 = gn0 + T[1] * g[1];
            double h[2];
            if (gn0 >= 0) { h[0] = norm * sA; h[1] = norm * cA; }
            else { h[0] = -norm * sA; h[1] = norm * cA; }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                double s = 0.0;
                s = s + Ti[r * 2 + 0] * h[0];
                s = s + Ti[r * 2 + 1] * h[1];
                g[r] = s;
            }
        }
    }

    return norm;
}

// Capillary force of a VD/GV site (ColorFluid::calcCapillaryForce,
// colorfluid.cpp:591-617) from the site's stored gradient and the curvature
// field of its fluid neighbours: kappa = sum_ab (n_a n_b - delta_ab) d_b c_a,
// with d_b c_a from the inverse moment at wall sites, else 3 sum w e dc;
// F = sigma/2 kappa direction.  Returns false if F is not finite.
template <int Q>
__device__ __forceinline__ bool capillary_force(const ColorArgs& a, uint32_t i, int x, int y, int z, double* F)
{
#pragma clang fp contract(off)
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
    if (norm2 > 1e-18) {
        const int32_t w = a.wall[i];
        uint32_t nb[Q];
        bool fluid[Q];
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            nb[q] = push_target<Q>(a.g, x, y, z, q);
            fluid[q] = a.densN[nb[q]] != SOLID_DENS_N;
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
        const double direction = gc && norm2 > 1e-18 ? grad[d] / sqrt(norm2) : grad[d];
        F[d] = .5 * a.sigma * curvature * direction;
        finite = finite && isfinite(F[d]);
    }
    return finite;
}

}  // namespace

template <int Q, bool INIT, bool OWN = false>
__global__ void __launch_bounds__(256) color_density(ColorArgs a)
{
#pragma clang fp contract(off)
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    if (!(lab.kind & KIND_FLUID)) return;
    double fT[Q], r0, r1, rzh, uzh[Lattice<Q>::D];
    load_color<Q, INIT, OWN>(a, i, desc, lab, fT, r0, r1, rzh, uzh);
    const double rt = r0 + r1;
    a.densN[i] = (r0 - r1) / rt;
}

// CSF, first half of calcCapillaryForce (colorfluid.cpp:577-586): the colour
// gradient of every VD/GV site and its curvature field (the unit normal, or
// the gradient for curvatureStencil "gradient").
template <int Q>
__global__ void __launch_bounds__(256) color_capillary(ColorArgs a)
{
#pragma clang fp contract(off)
    constexpr int D = Lattice<Q>::D;
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint64_t N = a.g.N;
    const uint32_t desc = a.desc[i];
    if (!(a.lab[desc >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) return;
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    double g[3];
    const double norm = color_gradient_f64<Q>(a, i, x, y, z, g);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        a.capgrad[uint64_t(d) * N + i] = g[d];
        a.capfield[uint64_t(d) * N + i] = a.curv_gradient ? g[d] : (norm > 1e-9 ? g[d] / norm : 0.0);
    }
}

// wallGradient tanh (colorfluid.cpp:494-518), at every wall site: atanh(chi)
// fitted as affine over the fluid neighbours (slope = inverse moment), tanh
// continued over the full stencil, and the bulk gradient of that, stored per
// wall entry for the gradient of the capillary and step passes.  A pass of its
// own keeps the transcendental loops out of the step kernel's registers.
template <int Q>
__global__ void __launch_bounds__(256) color_tanh_wall(ColorArgs a)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const int32_t w = a.wall[i];
    if (!w || !(a.lab[a.desc[i] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) return;
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    const DevWall& wl = a.walls[w];
    const double dn0 = a.densN[i];
    double g[3];
    const double center = tanh_wall_transform(dn0);
    double moment[3] = {0.0, 0.0, 0.0}, slope[3];
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const double nb = a.densN[push_target<Q>(a.g, x, y, z, q)];
        if (nb == SOLID_DENS_N) continue;
        const double difference = tanh_wall_transform(nb) - center;
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
    for (int d = 0; d < D; ++d) g[d] = 0.0;
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

// CSF initialisation (ColorFluid::initializeCapillaryForce,
// colorfluid.cpp:619-634): the t = 0 populations of initialize() shifted by
// the capillary force, f_k -= 1.5 w (e . F) rho_k / rho, written to the
// site's own slots of fout.
template <int Q>
__global__ void __launch_bounds__(256) color_csf_init(ColorArgs a)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint64_t N = a.g.N;
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    if (!(lab.kind & KIND_FLUID)) return;
    const int ph = lab.phase;
    const double* dist = ph == 0 ? a.dist0 : a.dist1;
    double f0[Q], f1[Q], r0, r1;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double f = dist[q];
        f -= 1.5 * L::w[q] * lab.ef[q];
        f0[q] = ph == 0 ? f : 0.0;
        f1[q] = ph == 0 ? 0.0 : f;
        if (q == 0) { r0 = f0[q]; r1 = f1[q]; }
        else { r0 += f0[q]; r1 += f1[q]; }
    }
    const double rt = r0 + r1;
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    double F[3] = {0.0, 0.0, 0.0};
    if (!capillary_force<Q>(a, i, x, y, z, F)) raise_status(a.status, 5, a.t);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double ef = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) ef += F[d];
            else if (L::e[q][d] == -1) ef += -F[d];
        }
        f0[q] -= 1.5 * L::w[q] * ef * r0 / rt;
        f1[q] -= 1.5 * L::w[q] * ef * r1 / rt;
        a.fout[uint64_t(q) * N + i] = f0[q];
        a.fout[(uint64_t(Q) + q) * N + i] = f1[q];
    }
}

template <int Q, bool MRT, bool INIT, bool STREAM_SOLIDS, bool CSF = false, bool OWN = false>
__global__ void __launch_bounds__(256) color_step(ColorArgs a)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint64_t N = a.g.N;
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    const int kind = lab.kind;

    if (!(kind & KIND_FLUID)) {
        if (INIT || !STREAM_SOLIDS) return;
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        bool finite = true;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint64_t src = q == 0 ? i : pull_source<Q>(a.g, x, y, z, q);
            const double f0 = a.fin[uint64_t(q) * N + src];
            const double f1 = a.fin[(uint64_t(Q) + q) * N + src];
            finite = finite && isfinite(f0) && isfinite(f1);
            a.fout[uint64_t(q) * N + i] = f0;
            a.fout[(uint64_t(Q) + q) * N + i] = f1;
        }
        if (!finite) raise_status(a.status, 4, a.t);
        return;
    }

    double fT[Q], r0, r1, rzh = 0.0, uzh[D];
    load_color<Q, INIT, OWN>(a, i, desc, lab, fT, r0, r1, rzh, uzh);

    // calcMacroVar
    const double rt = r0 + r1;
    if (!isfinite(r0) || !isfinite(r1) || !isfinite(rt) || rt <= 0.0) raise_status(a.status, 0, a.t);
    else if (r0 < -1e-10 * rt || r1 < -1e-10 * rt) raise_status(a.status, 1, a.t);
    const double dn = (r0 - r1) / rt;
    double tau = (1 + dn) * a.relax0 + (1 - dn) * a.relax1;
    if (a.harmonic) {
        const double phi = r0 / rt;
        tau = 0.5 + 3.0 / (phi / a.nu0 + (1.0 - phi) / a.nu1);
    }
    if (!isfinite(tau) || tau <= 0.5) raise_status(a.status, 2, a.t);
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double s = 0;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (L::e[q][d] == 1) s += fT[q];
            else if (L::e[q][d] == -1) s -= fT[q];
        }
        s /= rt;
        u[d] = s;
        if (!isfinite(s)) raise_status(a.status, 3, a.t);
    }
    // bodyForce (colorfluid.cpp:736-750): under CSF first u += (F_cap/2)/rho
    // on VD/GV sites, then u += (F/2)/rho on GV sites
    double fcap[3] = {0.0, 0.0, 0.0};
    if constexpr (CSF) {
        int cx, cy, cz;
        site_coords(a.g, i, cx, cy, cz);
        if (!capillary_force<Q>(a, i, cx, cy, cz, fcap)) raise_status(a.status, 5, a.t);
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] += .5 * fcap[d] / rt;
    }
    if (kind & KIND_GV) {
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] += lab.half[d] / rt;
    }
    if (a.flags & 1u) {
        a.rho0[i] = r0;
        a.rho1[i] = r1;
#pragma unroll
        for (int d = 0; d < D; ++d) a.u[uint64_t(d) * N + i] = u[d];
    }
    // adjacent report history (colorfluid.cpp:166-171): copied after the
    // boundaries, so a PB/VB site carries its Zou-He velocity and
    // (rho0 - rho1 of the previous step) / (Zou-He density).  The previous
    // step keeps the numerator; this step's field pass completes it.
    if (a.flags & 2u) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj[uint64_t(d) * N + i] = u[d];
        a.chiadj[i] = (kind & KIND_ZOUHE) ? r0 - r1 : (r0 - r1) / rt;
    }
    if (!INIT && (a.flags & 1u) && (kind & KIND_ZOUHE)) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj_prev[uint64_t(d) * N + i] = uzh[d];
        const double num = a.chiadj_prev[i];
        a.chiadj_prev[i] = rzh > 0.0 ? num / rzh : 0.0;
    }

    // collision of the total population; under CSF the site's force is the
    // label's plus the capillary force (colorfluid.cpp:326-333)
    bool forced = kind & KIND_FORCED;
    double force[3] = {lab.force[0], lab.force[1], D == 3 ? lab.force[2] : 0.0};
    if constexpr (CSF) {
        double f2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            force[d] += fcap[d];
            f2 = f2 + force[d] * force[d];
        }
        forced = f2 != 0.0;
    }
    double usq = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) usq += u[d] * u[d];
    double src[Q];
    if (forced) {
        double uf = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) uf += u[d] * force[d];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double eu = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) eu += u[d];
                else if (L::e[q][d] == -1) eu += -u[d];
            }
            double ef;
            if constexpr (CSF) {
                ef = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    if (L::e[q][d] == 1) ef += force[d];
                    else if (L::e[q][d] == -1) ef += -force[d];
                }
            } else {
                ef = lab.ef[q];
            }
            src[q] = L::w[q] * (3.0 * (ef - uf) + 9.0 * eu * ef);
        }
    }
    double feq0[Q];   // feq_u0: rest equilibrium of the two phases
    if constexpr (!MRT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ui = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) ui += u[d];
                else if (L::e[q][d] == -1) ui += -u[d];
            }
            feq0[q] = r0 * a.ai0[q] + r1 * a.ai1[q];
            const double feq = feq0[q] + rt * L::w[q] * (3.0 * ui + 4.5 * ui * ui - 1.5 * usq);
            fT[q] += (feq - fT[q]) / tau;
            if (forced) fT[q] += (1.0 - 0.5 / tau) * src[q];
        }
    } else {
        double res[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ui = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) ui += u[d];
                else if (L::e[q][d] == -1) ui += -u[d];
            }
            feq0[q] = r0 * a.ai0[q] + r1 * a.ai1[q];
            const double feq = feq0[q] + rt * L::w[q] * (3.0 * ui + 4.5 * ui * ui - 1.5 * usq);
            res[q] = feq - fT[q];
        }
        double corr[Q];
        if (a.uniform) {
            const double* __restrict__ A = a.A;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < Q; ++j) s = s + res[j] * A[q * Q + j];
                corr[q] = s;
            }
            if (forced) {
                const double* __restrict__ B = a.B;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    double s = 0.0;
#pragma unroll
                    for (int j = 0; j < Q; ++j) s = s + src[j] * B[q * Q + j];
                    corr[q] += s;
                }
            }
        } else {
            // per-site rates S(tau) in moment space (colorfluid.cpp:379-383)
            const double su = 1.0 / tau;
            const double sq = 8.0 * (2.0 - su) / (8.0 - su);
            double m[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    if (L::M[k][q] == 1) s = s + res[q];
                    else if (L::M[k][q] == -1) s = s + -res[q];
                    else if (L::M[k][q] != 0) s = s + res[q] * double(L::M[k][q]);
                }
                double rate;
                if constexpr (Q == 9) {
                    rate = (k == 0 || k == 3 || k == 5) ? 0.0 : ((k == 4 || k == 6) ? sq : su);
                } else {
                    rate = (k == 0 || k == 3 || k == 5 || k == 7) ? 0.0
                         : ((k == 4 || k == 6 || k == 8 || k >= 16) ? sq : su);
                }
                m[k] = s * rate;
                if (forced) {
                    double g = 0.0;
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        if (L::M[k][q] == 1) g = g + src[q];
                        else if (L::M[k][q] == -1) g = g + -src[q];
                        else if (L::M[k][q] != 0) g = g + src[q] * double(L::M[k][q]);
                    }
                    m[k] += g * (1.0 - 0.5 * rate);
                }
            }
            const double* __restrict__ MI = a.MInv;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < Q; ++k)
                    if (L::M[k][q] != 0) s = s + m[k] * MI[q * Q + k];
                corr[q] = s;
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) fT[q] += corr[q];
    }

    // colour gradient: the stored capillary gradient under CSF (no
    // perturbation, colorfluid.cpp:389-411), else colorGradient
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    double g[3] = {0.0, 0.0, 0.0};
    double norm;
    if constexpr (CSF) {
        double n2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            g[d] = a.capgrad[uint64_t(d) * N + i];
            n2 = n2 + g[d] * g[d];
        }
        norm = sqrt(n2);
    } else {
        norm = color_gradient_f64<Q>(a, i, x, y, z, g);
    }

    // perturbation + recolouring
    const double dens0T = r0 / rt;
    bool finite = true;
    if (norm > 1e-9) {
        const double b01 = a.beta * dens0T * r1 / rt;
        const double akn = a.ak * norm / tau;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double angle = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) angle += g[d];
                else if (L::e[q][d] == -1) angle -= g[d];
            }
            angle /= norm;
            if constexpr (!CSF) fT[q] += akn * (L::w[q] * angle * angle - a.bi[q]);
            if (L::diagonal(q)) angle /= SQRT2;
            const double f0 = dens0T * fT[q] + b01 * angle * feq0[q];
            const double f1 = fT[q] - f0;
            finite = finite && isfinite(f0) && isfinite(f1);
            a.fout[uint64_t(q) * N + i] = f0;
            a.fout[(uint64_t(Q) + q) * N + i] = f1;
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double f0 = dens0T * fT[q];
            const double f1 = fT[q] - f0;
            finite = finite && isfinite(f0) && isfinite(f1);
            a.fout[uint64_t(q) * N + i] = f0;
            a.fout[(uint64_t(Q) + q) * N + i] = f1;
        }
    }
    if (!finite) raise_status(a.status, 4, a.t);
}

template <int Q, bool MRT, bool INIT>
static hipError_t launch_color_q(const ColorArgs& a, bool stream_solids, hipStream_t s)
{
    const dim3 grid(uint32_t((a.g.N + 255) / 256)), block(256);
    auto tanh_pass = [&](const ColorArgs& x) {
        if (x.tanh_wall) hipLaunchKernelGGL((color_tanh_wall<Q>), grid, block, 0, s, x);
    };
    if (a.csf) {
        // CSF (colorfluid.cpp:124-137): a step is density, gradient/curvature
        // field, step.  t = 0: initialize + calcMacroVar, the capillary shift
        // of the populations, calcMacroVar again, bodyForce and collision on
        // the shifted populations in place.
        hipLaunchKernelGGL((color_density<Q, INIT>), grid, block, 0, s, a);
        tanh_pass(a);
        hipLaunchKernelGGL((color_capillary<Q>), grid, block, 0, s, a);
        if (!INIT) {
            if (stream_solids) hipLaunchKernelGGL((color_step<Q, MRT, false, true, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((color_step<Q, MRT, false, false, true>), grid, block, 0, s, a);
            return hipGetLastError();
        }
        hipLaunchKernelGGL((color_csf_init<Q>), grid, block, 0, s, a);
        ColorArgs b = a;
        b.fin = a.fout;
        hipLaunchKernelGGL((color_density<Q, true, true>), grid, block, 0, s, b);
        tanh_pass(b);
        hipLaunchKernelGGL((color_capillary<Q>), grid, block, 0, s, b);
        hipLaunchKernelGGL((color_step<Q, MRT, true, false, true, true>), grid, block, 0, s, b);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((color_density<Q, INIT>), grid, block, 0, s, a);
    tanh_pass(a);
    if (stream_solids && !INIT) hipLaunchKernelGGL((color_step<Q, MRT, INIT, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((color_step<Q, MRT, INIT, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_color(int Q, bool mrt, bool init, bool stream_solids, const ColorArgs& a, hipStream_t s)
{
    if (Q == 19) {
        if (mrt) return init ? launch_color_q<19, true, true>(a, stream_solids, s)
                             : launch_color_q<19, true, false>(a, stream_solids, s);
        return init ? launch_color_q<19, false, true>(a, stream_solids, s)
                    : launch_color_q<19, false, false>(a, stream_solids, s);
    }
    if (mrt) return init ? launch_color_q<9, true, true>(a, stream_solids, s)
                         : launch_color_q<9, true, false>(a, stream_solids, s);
    return init ? launch_color_q<9, false, true>(a, stream_solids, s)
                : launch_color_q<9, false, false>(a, stream_solids, s);
}

}  // namespace lbf
