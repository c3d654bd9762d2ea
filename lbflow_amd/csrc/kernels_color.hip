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
// the single-phase kernel.
//
// Reference operators (zhou1800/lbflow, default "legacy"/"perturbation"/
// "cosine" options):
//   initialize              solver/src/lbm/colorfluid.cpp:214-236
//   calcMacroVar            solver/src/lbm/colorfluid.cpp:238-290
//   bodyForce               solver/src/lbm/colorfluid.cpp:736-750
//   collision               solver/src/lbm/colorfluid.cpp:292-433
//   colorGradient (legacy / renormalized walls)  solver/src/lbm/colorfluid.cpp:435-570
//   RotateTx3D etc.         solver/src/algorithm/fn.cpp:118-164
//   zouHeBoundary (PB/VB)   solver/src/lbm/colorfluid.cpp:638-734 (zouhe.h)
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
// velocity in rzh / uzh.
template <int Q, bool INIT>
__device__ __forceinline__ void load_color(const ColorArgs& a, uint32_t i, uint32_t desc,
                                           const DevLabel& lab, double* fT, double& r0, double& r1,
                                           double& rzh, double* uzh)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    const uint64_t N = a.g.N;
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

}  // namespace

template <int Q, bool INIT>
__global__ void __launch_bounds__(256) color_density(ColorArgs a)
{
#pragma clang fp contract(off)
    const uint32_t i = xcd_block() * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    if (!(lab.kind & KIND_FLUID)) return;
    double fT[Q], r0, r1, rzh, uzh[Lattice<Q>::D];
    load_color<Q, INIT>(a, i, desc, lab, fT, r0, r1, rzh, uzh);
    const double rt = r0 + r1;
    a.densN[i] = (r0 - r1) / rt;
}

template <int Q, bool MRT, bool INIT, bool STREAM_SOLIDS>
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
    load_color<Q, INIT>(a, i, desc, lab, fT, r0, r1, rzh, uzh);

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
    // bodyForce: u += (F/2)/rho on GV sites
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

    // collision of the total population
    const bool forced = kind & KIND_FORCED;
    double usq = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) usq += u[d] * u[d];
    double src[Q];
    if (forced) {
        double uf = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) uf += u[d] * lab.force[d];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double eu = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) eu += u[d];
                else if (L::e[q][d] == -1) eu += -u[d];
            }
            const double ef = lab.ef[q];
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

    // colour gradient (legacy branch)
    int x, y, z;
    site_coords(a.g, i, x, y, z);
    double g[3] = {0.0, 0.0, 0.0};
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
            fT[q] += akn * (L::w[q] * angle * angle - a.bi[q]);
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
    hipLaunchKernelGGL((color_density<Q, INIT>), grid, block, 0, s, a);
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
