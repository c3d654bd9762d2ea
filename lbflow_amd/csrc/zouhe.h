// zouhe.h -- Zou-He pressure (PB) and velocity (VB) boundaries of the
// single-phase model, applied to one site's pulled populations before its
// collision (SCSP::zouHeBoundary, solver/src/lbm/scsp.cpp:188-256; called
// after streaming and bounce-back, scsp.cpp:79).
//
// The boundary direction dir (1..2D) is a property of the label, so the
// tables are selected by a switch and every population index stays a
// compile-time constant (the populations live in registers).  The arms of
// the switch only compute: they return the new unknown populations with
// their directions (ZhSite), and one write-back over compile-time indices
// stores them.  Arms that each ended by storing to different elements of
// the population array would be merged by the compiler into one store
// through a pointer chosen per arm, which moves those elements to scratch
// memory for the whole kernel.
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

// Unknown populations of one PB/VB site in the order of ZouHe::rel[dir],
// with the boundary density (FP32: its deviation from rho0) and the axial
// velocity.
template <int Q, typename T>
struct ZhSite {
    T val[ZouHe<Q>::NREL];
    int idx[ZouHe<Q>::NREL];
    T rho;
    T uax;
    int axis;
};

template <int Q, typename T>
__device__ __forceinline__ void zh_store(T* f, const ZhSite<Q, T>& z)
{
#pragma unroll
    for (int q = 1; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < ZouHe<Q>::NREL; ++r)
            if (z.idx[r] == q) f[q] = z.val[r];
}

template <int Q, typename T>
__device__ __forceinline__ void zh_velocity(const ZhSite<Q, T>& z, T* u)
{
#pragma unroll
    for (int d = 0; d < Lattice<Q>::D; ++d) u[d] = d == z.axis ? z.uax : T(0);
}

// Total population of the site: f (single-phase) or f0 + f1 (colour).
template <int Q, int DIR, typename T, typename Pop>
__device__ __forceinline__ ZhSite<Q, T> zouhe_unknowns_f64(Pop pop, const DevZouHe& z)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    using Z = ZouHe<Q>;
    constexpr int D = L::D;
    constexpr int AX = Z::axis[DIR], SG = Z::sign[DIR], OPD = L::op[DIR];
    ZhSite<Q, T> o;
    T sumF = 0;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) sumF += pop.twice(Z::rel[OPD][r]);
#pragma unroll
    for (int r = 0; r < Z::NPERP; ++r) sumF += pop(Z::perp[DIR][r]);
    if (z.type == ZH_VB) {
        const T velocity = SG == 1 ? z.value : -z.value;
        const T vel_1 = 1.0 / (1.0 - velocity);
        o.uax = velocity;
        o.rho = sumF * vel_1;
    } else {
        o.rho = z.value;
        sumF /= z.value;
        sumF = sumF - 1.0;
        o.uax = SG == 1 ? -sumF : sumF;
    }
    o.axis = AX;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) {
        const int q = Z::rel[DIR][r];
        T fq = pop(L::op[q]);
        fq += T(SG) * (L::w[q] * 6.0) * o.rho * o.uax;
        T s = 0;
#pragma unroll
        for (int n = 0; n < Z::NPERP; ++n) {
            const int qn = Z::perp[DIR][n];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (L::e[qn][d] == 0 || L::e[q][d] == 0) continue;
                if (L::e[qn][d] == L::e[q][d]) s += pop(qn);
                else s -= pop(qn);
            }
        }
        fq -= s * 0.5;
        o.val[r] = fq;
        o.idx[r] = q;
    }
    return o;
}

// Single-phase populations: sumF adds f + f for the reflected set.
struct PopF64 {
    const double* f;
    __device__ double operator()(int q) const { return f[q]; }
    __device__ double twice(int q) const { return f[q] + f[q]; }
};

// Colour populations: the total f0 + f1, the reflected set as f0 + f0 + f1 + f1
// (colorfluid.cpp:652-657 sums both phases' terms in that order).
struct PopColorF64 {
    const double* f0;
    const double* f1;
    __device__ double operator()(int q) const { return f0[q] + f1[q]; }
    __device__ double twice(int q) const { return f0[q] + f0[q] + f1[q] + f1[q]; }
};

template <int Q, typename Pop>
__device__ __forceinline__ ZhSite<Q, double> zouhe_dispatch_f64(Pop pop, const DevZouHe& z)
{
    switch (z.dir) {
    case 1: return zouhe_unknowns_f64<Q, 1, double>(pop, z);
    case 2: return zouhe_unknowns_f64<Q, 2, double>(pop, z);
    case 3: return zouhe_unknowns_f64<Q, 3, double>(pop, z);
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) return zouhe_unknowns_f64<Q, 5, double>(pop, z);
            if (z.dir == 6) return zouhe_unknowns_f64<Q, 6, double>(pop, z);
        }
        return zouhe_unknowns_f64<Q, 4, double>(pop, z);
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_f64(double* f, const DevZouHe& z, double& rho, double* u)
{
    const ZhSite<Q, double> o = zouhe_dispatch_f64<Q>(PopF64{f}, z);
    zh_store<Q>(f, o);
    rho = o.rho;
    zh_velocity<Q>(o, u);
}

// Colour: each unknown total goes to the label's phase, the other phase gets 0.
template <int Q>
__device__ __forceinline__ void zouhe_color_f64(double* f0, double* f1, const DevZouHe& z, int phs,
                                                double& rhoT, double* u)
{
    const ZhSite<Q, double> o = zouhe_dispatch_f64<Q>(PopColorF64{f0, f1}, z);
#pragma unroll
    for (int q = 1; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < ZouHe<Q>::NREL; ++r)
            if (o.idx[r] == q) {
                f0[q] = phs == 0 ? o.val[r] : 0.0;
                f1[q] = phs == 0 ? 0.0 : o.val[r];
            }
    rhoT = o.rho;
    zh_velocity<Q>(o, u);
}

template <int Q, int DIR>
__device__ __forceinline__ ZhSite<Q, float> zouhe_unknowns_f32(const float* h, const DevZouHe& z, float rho0)
{
#pragma clang fp contract(off)
    using L = Lattice<Q>;
    using Z = ZouHe<Q>;
    constexpr int D = L::D;
    constexpr int AX = Z::axis[DIR], SG = Z::sign[DIR], OPD = L::op[DIR];
    ZhSite<Q, float> o;
    float sumH = 0.0f;
#pragma unroll
    for (int r = 0; r < Z::NREL; ++r) sumH += h[Z::rel[OPD][r]] + h[Z::rel[OPD][r]];
#pragma unroll
    for (int r = 0; r < Z::NPERP; ++r) sumH += h[Z::perp[DIR][r]];
    if (z.type == ZH_VB) {
        o.uax = z.a;
        o.rho = (sumH + z.b) * z.c;
    } else {
        const float s = (sumH + z.b) * z.c;
        o.uax = SG == 1 ? -s : s;
        o.rho = -z.b;
    }
    o.axis = AX;
    const float rho = rho0 + o.rho;
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
        o.val[r] = h[L::op[q]] + float(SG * 6.0 * L::w[q]) * rho * o.uax - 0.5f * s;
        o.idx[r] = q;
    }
    return o;
}

template <int Q>
__device__ __forceinline__ ZhSite<Q, float> zouhe_dispatch_f32(const float* h, const DevZouHe& z, float rho0)
{
    switch (z.dir) {
    case 1: return zouhe_unknowns_f32<Q, 1>(h, z, rho0);
    case 2: return zouhe_unknowns_f32<Q, 2>(h, z, rho0);
    case 3: return zouhe_unknowns_f32<Q, 3>(h, z, rho0);
    default:
        if constexpr (Q == 19) {
            if (z.dir == 5) return zouhe_unknowns_f32<Q, 5>(h, z, rho0);
            if (z.dir == 6) return zouhe_unknowns_f32<Q, 6>(h, z, rho0);
        }
        return zouhe_unknowns_f32<Q, 4>(h, z, rho0);
    }
}

template <int Q>
__device__ __forceinline__ void zouhe_f32(float* h, const DevZouHe& z, float rho0, float& drho, float* u)
{
    const ZhSite<Q, float> o = zouhe_dispatch_f32<Q>(h, z, rho0);
    zh_store<Q>(h, o);
    drho = o.rho;
    zh_velocity<Q>(o, u);
}

// Colour FP32: the unknowns of the total (shifted) population as in the
// single-phase form, then the phase-0 population of each unknown: fT if the
// label is phase 0, else 0 (fT = h + ai0 rho_0).
template <int Q>
__device__ __forceinline__ void zouhe_color_f32(float* h, float* f0, const DevZouHe& z, int phs, float rho0,
                                                const double* dist0, float& drho, float* u)
{
    const ZhSite<Q, float> o = zouhe_dispatch_f32<Q>(h, z, rho0);
#pragma unroll
    for (int q = 1; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < ZouHe<Q>::NREL; ++r)
            if (o.idx[r] == q) {
                h[q] = o.val[r];
                f0[q] = phs == 0 ? o.val[r] + float(dist0[q]) : 0.0f;
            }
    drho = o.rho;
    zh_velocity<Q>(o, u);
}

}  // namespace lbf
