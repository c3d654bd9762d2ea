// kernels_scsp.hip -- single-phase (scsp) time step on CDNA4.
//
// One fused pull kernel per step: stream (periodic wrap in the index, the
// reference's MPI halo on one GPU) + halfway bounce-back + macro variables +
// half-force shift + SRT/MRT-with-Guo collision, one thread per lattice site,
// SoA populations [q][site] so every direction is a coalesced HBM stream.
//
// Reference operators (zhou1800/lbflow):
//   streaming + bounceBack   solver/src/lbm/simulator.cpp:73-111
//   calcMacroVar             solver/src/lbm/scsp.cpp:136-160
//   bodyForce                solver/src/lbm/scsp.cpp:162-175
//   equilSRT / equilMRT      solver/src/lbm/scsp.cpp:374-464
//   forceSource              solver/src/lbm/simulator.cpp:230-246
//   initialize               solver/src/lbm/scsp.cpp:116-134
//   validatePopulations      solver/src/lbm/simulator.cpp:327-333
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
// and the collision acts on h directly, so no large constant is carried in
// single precision.  They use FMA contraction.
#include "device.h"

namespace lbf {

constexpr uint32_t F_FIELDS = 1, F_ADJ = 2;
constexpr double C1 = 3.0, C2 = 4.5, C3 = 1.5;   // 3/c^2, 4.5/c^4, 1.5/c^2 with c = 1

template <int Q>
__device__ __forceinline__ void pull_populations_f64(const ScspArgs& a, const double* fin, uint32_t i,
                                                     uint32_t links, int x, int y, int z, double* f)
{
    using L = Lattice<Q>;
    const uint64_t N = a.g.N;
    f[0] = fin[i];
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const int o = L::op[q];
        // bounce-back (simulator.cpp:103-111): if the neighbour x - e_q is
        // solid the site is in ibS[op(q)] and receives its own outgoing f_op(q).
        if ((links >> o) & 1u) f[q] = fin[uint64_t(o) * N + i];
        else f[q] = fin[uint64_t(q) * N + pull_source<Q>(a.g, x, y, z, q)];
    }
}

// ---------------------------------------------------------------------------
// FP64, reference operation order.
// ---------------------------------------------------------------------------
template <int Q, bool MRT, bool INIT, bool STREAM_SOLIDS>
__global__ void __launch_bounds__(256) scsp_step_f64(ScspArgs a)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint64_t N = a.g.N;
    const double* __restrict__ fin = static_cast<const double*>(a.fin);
    double* __restrict__ fout = static_cast<double*>(a.fout);
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    const int kind = lab.kind;
    double f[Q];

    if (!(kind & KIND_FLUID)) {
        if (INIT || !STREAM_SOLIDS) return;
        // Solid sites are streamed but never collided, like the reference's
        // whole-array streaming; their contents only matter for checkpoints.
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        pull_populations_f64<Q>(a, fin, i, 0u, x, y, z, f);
        bool finite = true;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            finite = finite && isfinite(f[q]);
            fout[uint64_t(q) * N + i] = f[q];
        }
        if (!finite) raise_status(a.status, 4, a.t);
        return;
    }

    const bool forced = kind & KIND_FORCED;
    if (INIT) {
        // SCSP::initialize: f = w rho0 - 1.5 w (e.F)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            f[q] = a.idist[q];
            f[q] -= 1.5 * L::w[q] * lab.ef[q];
        }
    } else {
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        pull_populations_f64<Q>(a, fin, i, desc & DESC_LINK_MASK, x, y, z, f);
    }

    // calcMacroVar: rho = sum_q f_q (q ascending); u = (sum +-f) / rho0
    double rho = f[0];
#pragma unroll
    for (int q = 1; q < Q; ++q) rho += f[q];
    if (!isfinite(rho) || rho <= 0.0) raise_status(a.status, 0, a.t);
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double s = 0;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (L::e[q][d] == 1) s += f[q];
            else if (L::e[q][d] == -1) s -= f[q];
        }
        s /= a.dens0;
        u[d] = s;
    }
    // bodyForce: u += (F/2)/rho0 on GV sites
    if (kind & KIND_GV) {
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] += lab.half_rho[d];
    }
    if (a.flags & F_FIELDS) {
        a.rho[i] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) a.u[uint64_t(d) * N + i] = u[d];
    }
    if (a.flags & F_ADJ) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj[uint64_t(d) * N + i] = u[d];
    }

    // collision
    double usq = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) usq += u[d] * u[d];
    double src[Q];
    if (forced) {
        // forceSource: G_q = w_q (3 (e.F - u.F) + 9 (e.u)(e.F))
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
    if constexpr (!MRT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double ui = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) ui += u[d];
                else if (L::e[q][d] == -1) ui += -u[d];
            }
            const double feq = L::w[q] * (rho + a.dens0 * (C1 * ui + C2 * ui * ui - C3 * usq));
            f[q] += (feq - f[q]) * a.omega;
            if (forced) f[q] += a.omh * src[q];
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
            const double feq = L::w[q] * (rho + a.dens0 * (C1 * ui + C2 * ui * ui - C3 * usq));
            r[q] = feq - f[q];
        }
        const double* __restrict__ AT = a.AT;
        double ff[Q];
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) s += AT[q * Q + qq] * r[q];
            ff[qq] = s;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] += ff[q];
        if (forced) {
            const double* __restrict__ B = a.B;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < Q; ++j) s = s + src[j] * B[q * Q + j];
                f[q] += s;
            }
        }
    }
    bool finite = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        finite = finite && isfinite(f[q]);
        fout[uint64_t(q) * N + i] = f[q];
    }
    if (!finite) raise_status(a.status, 4, a.t);
}

// ---------------------------------------------------------------------------
// FP32 shifted populations.
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ void pull_populations_f32(const ScspArgs& a, const float* fin, uint32_t i,
                                                     uint32_t links, int x, int y, int z, float* h)
{
    using L = Lattice<Q>;
    const uint64_t N = a.g.N;
    h[0] = fin[i];
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const int o = L::op[q];
        if ((links >> o) & 1u) h[q] = fin[uint64_t(o) * N + i];
        else h[q] = fin[uint64_t(q) * N + pull_source<Q>(a.g, x, y, z, q)];
    }
}

// y = Op . v for an operator with the lattice parity symmetry
// Op[op q][op p] = Op[q][p], stored pair-split: for the rest direction and
// each representative q of an opposite pair, the even row
// [Op[q][0], (Op[q][p]+Op[q][op p])/2 ...] and for each representative the
// odd row [(Op[q][p]-Op[q][op p])/2 ...]; then y_q = E_q + O_q and
// y_op(q) = E_q - O_q.  181 FMAs instead of 361 for D3Q19.
template <int Q>
__device__ __forceinline__ void apply_split(const float* __restrict__ C, const float* v, float* y)
{
    using L = Lattice<Q>;
    constexpr int H = L::NHALF;
    float s[H], d[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const int p = L::half[k];
        s[k] = v[p] + v[L::op[p]];
        d[k] = v[p] - v[L::op[p]];
    }
    // rows: 0, then half[k]
#pragma unroll
    for (int r = 0; r <= H; ++r) {
        const float* row = C + r * (H + 1);
        float e = row[0] * v[0];
#pragma unroll
        for (int k = 0; k < H; ++k) e = fmaf(row[k + 1], s[k], e);
        if (r == 0) {
            y[0] = e;
        } else {
            const float* orow = C + (H + 1) * (H + 1) + (r - 1) * H;
            float o = 0.0f;
#pragma unroll
            for (int k = 0; k < H; ++k) o = fmaf(orow[k], d[k], o);
            const int q = L::half[r - 1];
            y[q] = e + o;
            y[L::op[q]] = e - o;
        }
    }
}

template <int Q, bool MRT, bool INIT, bool STREAM_SOLIDS>
__global__ void __launch_bounds__(256) scsp_step_f32(ScspArgs a)
{
    using L = Lattice<Q>;
    constexpr int D = L::D;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.g.N) return;
    const uint64_t N = a.g.N;
    const float* __restrict__ fin = static_cast<const float*>(a.fin);
    float* __restrict__ fout = static_cast<float*>(a.fout);
    const uint32_t desc = a.desc[i];
    const DevLabel& lab = a.lab[desc >> DESC_LABEL_SHIFT];
    const int kind = lab.kind;
    float h[Q];

    if (!(kind & KIND_FLUID)) {
        if (INIT || !STREAM_SOLIDS) return;
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        pull_populations_f32<Q>(a, fin, i, 0u, x, y, z, h);
#pragma unroll
        for (int q = 0; q < Q; ++q) fout[uint64_t(q) * N + i] = h[q];
        return;
    }
    const bool forced = kind & KIND_FORCED;
    const float rho0 = float(a.dens0);
    if (INIT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) h[q] = float(-1.5 * L::w[q] * lab.ef[q]);
    } else {
        int x, y, z;
        site_coords(a.g, i, x, y, z);
        pull_populations_f32<Q>(a, fin, i, desc & DESC_LINK_MASK, x, y, z, h);
    }
    // macro on shifted populations: rho = rho0 + sum h, j = sum e h
    float drho = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) drho += h[q];
    float u[D];
    const float inv_rho0 = float(1.0 / a.dens0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        float s = 0.0f;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (L::e[q][d] == 1) s += h[q];
            else if (L::e[q][d] == -1) s -= h[q];
        }
        u[d] = s * inv_rho0;
        if (kind & KIND_GV) u[d] += float(lab.half_rho[d]);
    }
    const double rho_d = a.dens0 + double(drho);
    if (!isfinite(drho) || rho_d <= 0.0) raise_status(a.status, 0, a.t);
    if (a.flags & F_FIELDS) {
        a.rho[i] = rho_d;
#pragma unroll
        for (int d = 0; d < D; ++d) a.u[uint64_t(d) * N + i] = double(u[d]);
    }
    if (a.flags & F_ADJ) {
#pragma unroll
        for (int d = 0; d < D; ++d) a.uadj[uint64_t(d) * N + i] = double(u[d]);
    }
    float usq = 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) usq = fmaf(u[d], u[d], usq);
    // shifted equilibrium h_eq = w (drho + rho0 (3 eu + 4.5 eu^2 - 1.5 u^2))
    float r[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        float ui = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (L::e[q][d] == 1) ui += u[d];
            else if (L::e[q][d] == -1) ui -= u[d];
        }
        const float poly = fmaf(ui, fmaf(4.5f, ui, 3.0f), -1.5f * usq);
        const float heq = float(L::w[q]) * fmaf(rho0, poly, drho);
        r[q] = heq - h[q];
    }
    float src[Q];
    if (forced) {
        float uf = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) uf = fmaf(u[d], float(lab.force[d]), uf);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float eu = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) eu += u[d];
                else if (L::e[q][d] == -1) eu -= u[d];
            }
            const float ef = float(lab.ef[q]);
            src[q] = float(L::w[q]) * fmaf(9.0f * eu, ef, 3.0f * (ef - uf));
        }
    }
    if constexpr (!MRT) {
        const float om = float(a.omega), omh = float(a.omh);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            h[q] = fmaf(om, r[q], h[q]);
            if (forced) h[q] = fmaf(omh, src[q], h[q]);
        }
    } else {
        float dh[Q];
        apply_split<Q>(a.A32, r, dh);
#pragma unroll
        for (int q = 0; q < Q; ++q) h[q] += dh[q];
        if (forced) {
            apply_split<Q>(a.B32, src, dh);
#pragma unroll
            for (int q = 0; q < Q; ++q) h[q] += dh[q];
        }
    }
    bool finite = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        finite = finite && isfinite(h[q]);
        fout[uint64_t(q) * N + i] = h[q];
    }
    if (!finite) raise_status(a.status, 4, a.t);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int Q, bool MRT, bool INIT>
static hipError_t launch_scsp_q(const ScspArgs& a, bool fp32, bool stream_solids, hipStream_t s)
{
    const dim3 grid(uint32_t((a.g.N + 255) / 256)), block(256);
    if (fp32) {
        if (stream_solids) hipLaunchKernelGGL((scsp_step_f32<Q, MRT, INIT, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((scsp_step_f32<Q, MRT, INIT, false>), grid, block, 0, s, a);
    } else {
        if (stream_solids) hipLaunchKernelGGL((scsp_step_f64<Q, MRT, INIT, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((scsp_step_f64<Q, MRT, INIT, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_scsp(int Q, bool mrt, bool init, bool fp32, bool stream_solids, const ScspArgs& a,
                       hipStream_t s)
{
    if (Q == 19) {
        if (mrt) return init ? launch_scsp_q<19, true, true>(a, fp32, stream_solids, s)
                             : launch_scsp_q<19, true, false>(a, fp32, stream_solids, s);
        return init ? launch_scsp_q<19, false, true>(a, fp32, stream_solids, s)
                    : launch_scsp_q<19, false, false>(a, fp32, stream_solids, s);
    }
    if (mrt) return init ? launch_scsp_q<9, true, true>(a, fp32, stream_solids, s)
                         : launch_scsp_q<9, true, false>(a, fp32, stream_solids, s);
    return init ? launch_scsp_q<9, false, true>(a, fp32, stream_solids, s)
                : launch_scsp_q<9, false, false>(a, fp32, stream_solids, s);
}

}  // namespace lbf
