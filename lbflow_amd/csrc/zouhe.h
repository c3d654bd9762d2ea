// zouhe.h -- Zou-He pressure (PB) and velocity (VB) boundaries of the
// single-phase model, applied to one site's pulled populations before its
// collision (SCSP::zouHeBoundary, solver/src/lbm/scsp.cpp:188-256; called
// after streaming and bounce-back, scsp.cpp:79).
//
// The boundary direction dir (1..2D) is a property of the label, so the
// tables are selected by a switch and every population index stays a
// compile-time constant (the populations live in registers).
//
// FP64 keeps the reference's operation order: the known-population sum
// 2 sum_{Qrel[op dir]} f + sum_{Qprep[dir]} f, then the boundary density and
// velocity, then for every unknown q in Qrel[dir]
//   f_q = f_op(q) + sign wiT_q rho u_axis - (sum_{Qprep[dir]} +-f) / 2
// with wiT = 6 w / latticeSpeed (simulator.cpp:28), latticeSpeed 1.
//
// FP32 does the same on shifted populations h = f - w rho0.  The weights of
// the known sum add up to 1 (2/6 + 2/3), so sum f = sum h + rho0; the +-
// tangential sum of the rest-state parts cancels pairwise; the unknown update
// therefore reads h_q = h_op(q) + sign 6 w_q rho u_axis - (sum +-h) / 2.
//
// The colour model (ColorFluid::zouHeBoundary, colorfluid.cpp:638-734) runs
// the same construction on the total population f0 + f1 and gives each
// unknown to the label's phase (the other phase gets 0); its density and
// velocity are then recomputed by calcMacroVar like at any fluid site and
// only reach the adjacent report history.  The FP32 colour planes hold
// h = fT - ai0 rho_0, whose known-sum weights also add up to 1, so the
// shifted form is the single-phase one with rho0 = rho_0.
#pragma once

#include "device.h"

namespace lbf {

template <int Q, int DIR>
__device__ __forceinline__ void zouhe_f64_dir(double* f, const DevZouHe& z, double& rho, double* u)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    using Z = ZouHe<Q>;
    constexpr int D = L::D;
    constexpr int AX = Z::axis[DIR], SG = Z::sign[DIR], OPD = L::op[DIR];
    double sumF = 0;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) sumF += f[Z::rel[OPD][r]] + f[Z::rel[OPD][r]];
#pragma unroll
    for (int r = 0; r < Z::NPERP; ++r) sumF += f[Z::perp[DIR][r]];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = 0.0;
    if (z.type == ZH_VB) {
        const double velocity = SG == 1 ? z.value : -z.value;
        const double vel_1 = 1.0 / (1.0 - velocity);
        u[AX] = velocity;
        rho = sumF * vel_1;
    } else {
        rho = z.value;
        sumF /= z.value;
        sumF = sumF - 1.0;
        u[AX] = SG == 1 ? -sumF : sumF;
    }
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[DIR][r];
        double fq = f[L::op[q]];
        fq += double(SG) * (L::w[q] * 6.0) * rho * u[AX];
        double s = 0;
#pragma unroll
        for (int n = 0; n < Z::NPERP; ++n) {
            const int qn = Z::perp[DIR][n];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[qn][d] == 0 || L::e[q][d] == 0) continue;
                if (L::e[qn][d] == L::e[q][d]) s += f[qn];
                else s -= f[qn];
            }
        }
        fq -= s * 0.5;
        f[q] = fq;
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_f64(double* f, const DevZouHe& z, double& rho, double* u)
{
    switch (z.dir) {
    case 1: zouhe_f64_dir<Q, 1>(f, z, rho, u); break;
    case 2: zouhe_f64_dir<Q, 2>(f, z, rho, u); break;
    case 3: zouhe_f64_dir<Q, 3>(f, z, rho, u); break;
    case 4: zouhe_f64_dir<Q, 4>(f, z, rho, u); break;
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) zouhe_f64_dir<Q, 5>(f, z, rho, u);
            else zouhe_f64_dir<Q, 6>(f, z, rho, u);
        }
        break;
    }
}

template <int Q, int DIR>
__device__ __forceinline__ void zouhe_color_f64_dir(double* f0, double* f1, const DevZouHe& z, int phs,
                                                    double& rhoT, double* u)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    using Z = ZouHe<Q>;
    constexpr int D = L::D;
    constexpr int AX = Z::axis[DIR], SG = Z::sign[DIR], OPD = L::op[DIR];
    double sumF = 0;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[OPD][r];
        sumF += f0[q] + f0[q] + f1[q] + f1[q];
    }
#pragma unroll
    for (int r = 0; r < Z::NPERP; ++r) {
        const int q = Z::perp[DIR][r];
        sumF += f0[q] + f1[q];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = 0.0;
    if (z.type == ZH_VB) {
        const double velocity = SG == 1 ? z.value : -z.value;
        const double vel_1 = 1.0 / (1.0 - velocity);
        u[AX] = velocity;
        rhoT = sumF * vel_1;
    } else {
        rhoT = z.value;
        sumF /= z.value;
        sumF = sumF - 1.0;
        u[AX] = SG == 1 ? -sumF : sumF;
    }
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[DIR][r];
        double fT = f0[L::op[q]] + f1[L::op[q]];
        fT += double(SG) * (L::w[q] * 6.0) * rhoT * u[AX];
        double s = 0;
#pragma unroll
        for (int n = 0; n < Z::NPERP; ++n) {
            const int qn = Z::perp[DIR][n];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[qn][d] == 0 || L::e[q][d] == 0) continue;
                if (L::e[qn][d] == L::e[q][d]) s += f0[qn] + f1[qn];
                else s -= f0[qn] + f1[qn];
            }
        }
        fT -= s * 0.5;
        f0[q] = phs == 0 ? fT : 0.0;
        f1[q] = phs == 0 ? 0.0 : fT;
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_color_f64(double* f0, double* f1, const DevZouHe& z, int phs, double& rhoT,
                                                double* u)
{
    switch (z.dir) {
    case 1: zouhe_color_f64_dir<Q, 1>(f0, f1, z, phs, rhoT, u); break;
    case 2: zouhe_color_f64_dir<Q, 2>(f0, f1, z, phs, rhoT, u); break;
    case 3: zouhe_color_f64_dir<Q, 3>(f0, f1, z, phs, rhoT, u); break;
    case 4: zouhe_color_f64_dir<Q, 4>(f0, f1, z, phs, rhoT, u); break;
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) zouhe_color_f64_dir<Q, 5>(f0, f1, z, phs, rhoT, u);
            else zouhe_color_f64_dir<Q, 6>(f0, f1, z, phs, rhoT, u);
        }
        break;
    }
}

template <int Q, int DIR>
__device__ __forceinline__ void zouhe_f32_dir(float* h, const DevZouHe& z, float rho0, float& drho, float* u)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    using Z = ZouHe<Q>;
    constexpr int D = L::D;
    constexpr int AX = Z::axis[DIR], SG = Z::sign[DIR], OPD = L::op[DIR];
    float sumH = 0.0f;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) sumH += h[Z::rel[OPD][r]] + h[Z::rel[OPD][r]];
#pragma unroll
    for (int r = 0; r < Z::NPERP; ++r) sumH += h[Z::perp[DIR][r]];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = 0.0f;
    if (z.type == ZH_VB) {
        u[AX] = z.a;
        drho = (sumH + z.b) * z.c;
    } else {
        const float s = (sumH + z.b) * z.c;
        u[AX] = SG == 1 ? -s : s;
        drho = -z.b;
    }
    const float rho = rho0 + drho;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[DIR][r];
        float s = 0.0f;
#pragma unroll
        for (int n = 0; n < Z::NPERP; ++n) {
            const int qn = Z::perp[DIR][n];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[qn][d] == 0 || L::e[q][d] == 0) continue;
                if (L::e[qn][d] == L::e[q][d]) s += h[qn];
                else s -= h[qn];
            }
        }
        h[q] = h[L::op[q]] + float(SG * 6.0 * L::w[q]) * rho * u[AX] - 0.5f * s;
    }
}

// Colour FP32: the unknowns of the total (shifted) population as in the
// single-phase form, then the phase-0 population of each unknown: fT if the
// label is phase 0, else 0 (fT = h + ai0 rho_0).
template <int Q, int DIR>
__device__ __forceinline__ void zouhe_color_f32_dir(float* h, float* f0, const DevZouHe& z, int phs, float rho0,
                                                    const double* dist0, float& drho, float* u)
{
    using Z = ZouHe<Q>;
    zouhe_f32_dir<Q, DIR>(h, z, rho0, drho, u);
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[DIR][r];
        f0[q] = phs == 0 ? h[q] + float(dist0[q]) : 0.0f;
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_color_f32(float* h, float* f0, const DevZouHe& z, int phs, float rho0,
                                                const double* dist0, float& drho, float* u)
{
    switch (z.dir) {
    case 1: zouhe_color_f32_dir<Q, 1>(h, f0, z, phs, rho0, dist0, drho, u); break;
    case 2: zouhe_color_f32_dir<Q, 2>(h, f0, z, phs, rho0, dist0, drho, u); break;
    case 3: zouhe_color_f32_dir<Q, 3>(h, f0, z, phs, rho0, dist0, drho, u); break;
    case 4: zouhe_color_f32_dir<Q, 4>(h, f0, z, phs, rho0, dist0, drho, u); break;
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) zouhe_color_f32_dir<Q, 5>(h, f0, z, phs, rho0, dist0, drho, u);
            else zouhe_color_f32_dir<Q, 6>(h, f0, z, phs, rho0, dist0, drho, u);
        }
        break;
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_f32(float* h, const DevZouHe& z, float rho0, float& drho, float* u)
{
    switch (z.dir) {
    case 1: zouhe_f32_dir<Q, 1>(h, z, rho0, drho, u); break;
    case 2: zouhe_f32_dir<Q, 2>(h, z, rho0, drho, u); break;
    case 3: zouhe_f32_dir<Q, 3>(h, z, rho0, drho, u); break;
    case 4: zouhe_f32_dir<Q, 4>(h, z, rho0, drho, u); break;
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) zouhe_f32_dir<Q, 5>(h, z, rho0, drho, u);
            else zouhe_f32_dir<Q, 6>(h, z, rho0, drho, u);
        }
        break;
    }
}

}  // namespace lbf
