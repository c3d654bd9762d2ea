// collide_f32.h -- single-phase SRT / MRT + Guo collision on FP32 shifted
// populations h_q = f_q - w_q rho0, shared by the single-domain and the
// batched kernels.
//
// Reference operators: SCSP::equilSRT / equilMRT (solver/src/lbm/scsp.cpp:
// 374-464), Simulator::forceSource (simulator.cpp:230-246), the MRT
// transform and rates of simulator.cpp:162-216 (legacy spectrum).
//
// The MRT collision runs in moment space with the reference's transform M:
// m = M h (pair sums of opposite directions, even/odd rows), equilibrium and
// Guo-source moments from the compile-time moment tables T0 = M w,
// T1 = M (w e), T2 = M (w e e) (only non-zero terms are emitted), relaxation
// y_k = (s_k (m_eq,k - m_k) + (1 - s_k/2) m_G,k) / |M_k|^2 and back-transform
// h += M^T y.  Conserved moments (s_k = 0) carry only the source term.
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.h"

namespace lbf {

// Moment tables of the transform M: for row k, T0[k] = sum_q M_kq w_q,
// T1[k][a] = sum_q M_kq w_q e_qa, T2[k][a][b] = sum_q M_kq w_q e_qa e_qb,
// n2[k] = sum_q M_kq^2, parity[k] = +1 (row even under q -> op(q)) or -1,
// rate[k] = 0 (conserved), 1 (s_nu = 1/tau) or 2 (s_q).
template <int Q> struct MomentTables {
    double T0[Q];
    double T1[Q][3];
    double T2[Q][3][3];
    double n2[Q];
    int parity[Q];
    int rate[Q];
    constexpr MomentTables() : T0{}, T1{}, T2{}, n2{}, parity{}, rate{}
    {
        using L = Lattice<Q>;
        for (int k = 0; k < Q; ++k) {
            bool even = true, odd = true;
            for (int q = 0; q < Q; ++q) {
                const double m = L::M[k][q], w = L::w[q];
                T0[k] += m * w;
                n2[k] += m * m;
                for (int a = 0; a < L::D; ++a) {
                    T1[k][a] += m * w * L::e[q][a];
                    for (int b = 0; b < L::D; ++b) T2[k][a][b] += m * w * L::e[q][a] * L::e[q][b];
                }
                even = even && L::M[k][L::op[q]] == L::M[k][q];
                odd = odd && L::M[k][L::op[q]] == -L::M[k][q];
            }
            parity[k] = even ? 1 : (odd ? -1 : 0);
        }
        if constexpr (Q == 9) {
            const int r[9] = {0, 1, 1, 0, 2, 0, 2, 1, 1};
            for (int k = 0; k < 9; ++k) rate[k] = r[k];
        } else {
            const int r[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
            for (int k = 0; k < 19; ++k) rate[k] = r[k];
        }
    }
};
template <int Q> inline constexpr MomentTables<Q> kMT{};

// Relaxation constants of one collision.
struct RelaxF32 {
    float rho0, omega, omh;            // SRT: 1/tau, 1 - 1/(2 tau)
    float c1_nu, c2_nu, c1_q, c2_q;    // MRT: s and 1 - s/2 of the two non-conserved rate classes
};

// Macro values of shifted populations: drho = sum h, u = (sum e h)/rho0 + F/(2 rho0).
// Every product-sum below is an explicit fmaf and nothing else is contracted,
// so the rounding is fixed by this source alone: kernels that inline these
// functions (single domain, batched crops) produce identical bits.
template <int Q>
__device__ __forceinline__ float macro_f32(const float* h, float inv_rho0, const float* half_rho, float* u)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    float drho = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) drho += h[q];
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        float s = 0.0f;
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (L::e[q][d] == 1) s += h[q];
            else if (L::e[q][d] == -1) s -= h[q];
        }
        u[d] = fmaf(s, inv_rho0, half_rho[d]);
    }
    return drho;
}

template <int Q, bool MRT>
__device__ __forceinline__ void collide_f32(float* h, float drho, const float* u, const float* F, const RelaxF32& r)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    constexpr int D = L::D;
    constexpr int H = L::NHALF;
    constexpr MomentTables<Q> MT = kMT<Q>;
    float usq = 0.0f, uF = 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        usq = fmaf(u[d], u[d], usq);
        uF = fmaf(u[d], F[d], uF);
    }
    const float rho0 = r.rho0;
    if constexpr (!MRT) {
        // h += omega (h_eq - h) + (1 - omega/2) G,  h_eq = w (drho + rho0 (3eu + 4.5eu^2 - 1.5u^2))
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float eu = 0.0f, ef = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[q][d] == 1) { eu += u[d]; ef += F[d]; }
                else if (L::e[q][d] == -1) { eu -= u[d]; ef -= F[d]; }
            }
            const float w = float(L::w[q]);
            const float quad = fmaf(eu, fmaf(4.5f, eu, 3.0f), -1.5f * usq);
            const float heq = w * fmaf(rho0, quad, drho);
            const float G = w * fmaf(9.0f * eu, ef, 3.0f * (ef - uF));
            h[q] = fmaf(r.omega, heq - h[q], fmaf(r.omh, G, h[q]));
        }
    } else {
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
            // Guo source moment: M G = 3 T1.F - 3 (u.F) T0 + 9 u.T2.F
            float mg = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (MT.T1[k][a] != 0.0) mg = fmaf(float(3.0 * MT.T1[k][a]), F[a], mg);
#pragma unroll
                for (int b = 0; b < D; ++b)
                    if (MT.T2[k][a][b] != 0.0) mg = fmaf(float(9.0 * MT.T2[k][a][b]) * u[a], F[b], mg);
            }
            if (MT.T0[k] != 0.0) mg = fmaf(float(-3.0 * MT.T0[k]), uF, mg);
            if (MT.rate[k] == 0) {
                y[k] = mg * float(1.0 / MT.n2[k]);
                continue;
            }
            float m = 0.0f;
            if (MT.parity[k] > 0) {
                if (L::M[k][0] != 0) m = fmaf(float(L::M[k][0]), h[0], m);
#pragma unroll
                for (int j = 0; j < H; ++j)
                    if (L::M[k][L::half[j]] != 0) m = fmaf(float(L::M[k][L::half[j]]), sp[j], m);
            } else {
#pragma unroll
                for (int j = 0; j < H; ++j)
                    if (L::M[k][L::half[j]] != 0) m = fmaf(float(L::M[k][L::half[j]]), sm[j], m);
            }
            // shifted equilibrium moment: drho T0 + rho0 (3 T1.u + 4.5 u.T2.u - 1.5 u^2 T0)
            float meq = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (MT.T1[k][a] != 0.0) meq = fmaf(float(3.0 * MT.T1[k][a]), u[a], meq);
#pragma unroll
                for (int b = 0; b < D; ++b)
                    if (MT.T2[k][a][b] != 0.0) meq = fmaf(float(4.5 * MT.T2[k][a][b]) * u[a], u[b], meq);
            }
            if (MT.T0[k] != 0.0) meq = fmaf(float(-1.5 * MT.T0[k]), usq, meq);
            meq *= rho0;
            if (MT.T0[k] != 0.0) meq = fmaf(float(MT.T0[k]), drho, meq);
            const float c1 = MT.rate[k] == 1 ? r.c1_nu : r.c1_q;
            const float c2 = MT.rate[k] == 1 ? r.c2_nu : r.c2_q;
            y[k] = fmaf(c1, meq - m, c2 * mg) * float(1.0 / MT.n2[k]);
        }
        {
            float e0 = 0.0f;
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (MT.parity[k] > 0 && L::M[k][0] != 0) e0 = fmaf(float(L::M[k][0]), y[k], e0);
            h[0] += e0;
        }
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const int q = L::half[j];
            float ev = 0.0f, od = 0.0f;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (L::M[k][q] == 0) continue;
                if (MT.parity[k] > 0) ev = fmaf(float(L::M[k][q]), y[k], ev);
                else od = fmaf(float(L::M[k][q]), y[k], od);
            }
            h[q] += ev + od;
            h[L::op[q]] += ev - od;
        }
    }
}

}  // namespace lbf
