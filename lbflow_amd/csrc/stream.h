// stream.h -- pull streaming with halfway bounce-back, shared by the
// single-phase and the FP32 colour kernels: per-site wrapped coordinate
// terms, the element offset each direction pulls from in the dense and the
// sparse (fluid sites only) layouts, and uniform-base loads/stores.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace lbf {

enum : int { DENSE_ALL = 0, DENSE_FLUID = 1, SPARSE = 2 };

// Wrapped coordinate terms of a site: x-1, x, x+1 and the matching row and
// plane offsets (site index units), computed once per site so that every
// neighbour index is two integer adds.
struct Stencil {
    uint32_t x[3], y[3], z[3];   // [0] = -1, [1] = 0, [2] = +1
};

__device__ __forceinline__ Stencil make_stencil(const GridDev& g, uint32_t site)
{
    int x, y, z;
    site_coords(g, site, x, y, z);
    Stencil s;
    s.x[1] = uint32_t(x);
    s.x[0] = x == 0 ? g.nx - 1 : uint32_t(x - 1);
    s.x[2] = uint32_t(x) + 1 == g.nx ? 0u : uint32_t(x + 1);
    s.y[1] = uint32_t(y) * g.nx;
    s.y[0] = (y == 0 ? g.ny - 1 : uint32_t(y - 1)) * g.nx;
    s.y[2] = (uint32_t(y) + 1 == g.ny ? 0u : uint32_t(y + 1)) * g.nx;
    s.z[1] = uint32_t(z) * g.nxy;
    s.z[0] = (z == 0 ? g.nz - 1 : uint32_t(z - 1)) * g.nxy;
    s.z[2] = (uint32_t(z) + 1 == g.nz ? 0u : uint32_t(z + 1)) * g.nxy;
    return s;
}

// Site x - e_q (pull source of direction q).
template <int Q>
__device__ __forceinline__ uint32_t upstream(const Stencil& s, int q)
{
    using L = Lattice<Q>;
    int ex, ey, ez;
    if constexpr (Q == 9) {
        ez = 0; ey = L::e[q][0]; ex = L::e[q][1];
    } else {
        ez = L::e[q][0]; ey = L::e[q][1]; ex = L::e[q][2];
    }
    return s.z[1 - ez] + s.y[1 - ey] + s.x[1 - ex];
}

// Element offset (from the start of the population array) of the value that
// streams into direction q of this slot: the upstream neighbour's f_q, or
// this slot's own f_op(q) when the link bounces back.
//
// Sparse layout: the slot of a fluid site is its rank among the fluid sites
// in z,y,x order, read from a 32-site rank index {fluid sites before the
// word, fluid bit mask of the word}: 8 bytes per 32 sites (4 MB at 256^3),
// small enough to stay in L2, instead of a 4-byte map entry per site.
template <int Q, int LAYOUT, typename O>
__device__ __forceinline__ O pull_offset(const uint2* __restrict__ rank, uint32_t links, uint32_t slot,
                                         const Stencil& st, uint32_t P, int q)
{
    using L = Lattice<Q>;
    // site x - e_q solid <=> site in ibS[op(q)] (simulator.cpp:103-111)
    if ((links >> L::op[q]) & 1u) return O(L::op[q]) * P + slot;
    const uint32_t src = upstream<Q>(st, q);
    if constexpr (LAYOUT == SPARSE) {
        const uint2 r = rank[src >> 5];
        return O(q) * P + (r.x + uint32_t(__popc(r.y & ((1u << (src & 31u)) - 1u))));
    } else {
        return O(q) * P + src;
    }
}

// Element index of plane q of slot s in an array of NPL planes.  The dense
// layouts store plane after plane (q P + s).  The sparse layout (BLK) stores
// blocks of 64 slots, each holding its NPL planes one after another
// ([s / 64][q][s % 64]): a wave's 64 lanes still read and write 64
// consecutive values per plane, while the values one site pulls, and the
// ones it writes, lie within a few nearby blocks instead of NPL planes P
// values apart (fewer DRAM pages and TLB entries per wave).
template <int NPL, bool BLK, typename O>
__device__ __forceinline__ O pop_index(uint32_t P, int q, uint32_t s)
{
    if constexpr (BLK) return O(s >> 6) * O(NPL * 64) + O(q * 64) + O(s & 63u);
    else return O(q) * P + s;
}

// Offsets of the values streaming into directions 1..Q-1 (off[0] unused).
// Sparse layout: the pull sources of one site lie in 9 rows (dz, dy) at
// columns x-1, x, x+1.  A fluid site's slot is its rank among the fluid
// sites in z,y,x order, so inside a row the column x-1 has rank(x) - 1 and
// the column x+1 has rank(x) + [x is fluid] (rank counts across word
// boundaries).  One rank lookup per row serves the whole row: 8 lookups for
// D3Q19 instead of 18 (the site's own row needs none).  Columns that wrap
// around the row end are looked up on their own.
template <int Q, int LAYOUT, typename O, bool BLK = false, int NPL = Q>
__device__ __forceinline__ void pull_offsets(const uint2* __restrict__ rank, uint32_t links, uint32_t slot,
                                             const Stencil& st, uint32_t nx, uint32_t P, O* off)
{
    using L = Lattice<Q>;
    if constexpr (LAYOUT != SPARSE) {
#pragma unroll
        for (int q = 1; q < Q; ++q) off[q] = pull_offset<Q, LAYOUT, O>(rank, links, slot, st, P, q);
    } else {
        auto rank_of = [&](uint32_t site, uint32_t& fluid) {
            const uint2 r = rank[site >> 5];
            const uint32_t bit = site & 31u;
            fluid = (r.y >> bit) & 1u;
            return r.x + uint32_t(__popc(r.y & ((1u << bit) - 1u)));
        };
        uint32_t rr[3][3], rf[3][3];   // rank and fluid flag of column x in row (dz, dy)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                rr[a][b] = 0;
                rf[a][b] = 0;
                if (a == 1 && b == 1) {          // the site's own row: its own slot
                    rr[a][b] = slot;
                    rf[a][b] = 1u;
                    continue;
                }
                bool used = false;               // does any pull source lie in this row?
#pragma unroll
                for (int q = 1; q < Q; ++q) {
                    const int ez = Q == 9 ? 0 : L::e[q][0], ey = Q == 9 ? L::e[q][0] : L::e[q][1];
                    used = used || (1 - ez == a && 1 - ey == b);
                }
                if (used) rr[a][b] = rank_of(st.z[a] + st.y[b] + st.x[1], rf[a][b]);
            }
        const uint32_t x = st.x[1];
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            int ez, ey, ex;
            if constexpr (Q == 9) {
                ez = 0; ey = L::e[q][0]; ex = L::e[q][1];
            } else {
                ez = L::e[q][0]; ey = L::e[q][1]; ex = L::e[q][2];
            }
            const int a = 1 - ez, b = 1 - ey;
            uint32_t m;
            if (ex == 0) {
                m = rr[a][b];
            } else if (ex == 1) {        // source column x - 1
                uint32_t fl;
                m = x == 0 ? rank_of(st.z[a] + st.y[b] + st.x[0], fl) : rr[a][b] - 1u;
            } else {                     // source column x + 1
                uint32_t fl;
                m = x + 1 == nx ? rank_of(st.z[a] + st.y[b] + st.x[2], fl) : rr[a][b] + rf[a][b];
            }
            // site x - e_q solid <=> site in ibS[op(q)] (simulator.cpp:103-111)
            const bool bounce = (links >> L::op[q]) & 1u;
            off[q] = pop_index<NPL, BLK, O>(P, bounce ? L::op[q] : q, bounce ? slot : m);
        }
    }
}

// Load/store with a uniform base and an element offset.  With a 32-bit
// offset type (population array below 4 GiB) the address is an SGPR base plus
// a zero-extended 32-bit VGPR byte offset; the 64-bit type serves larger
// arrays.
template <typename T, typename O> __device__ __forceinline__ T ld(const T* __restrict__ base, O elem)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + elem * O(sizeof(T)));
}
template <typename T, typename O> __device__ __forceinline__ void st(T* __restrict__ base, O elem, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + elem * O(sizeof(T))) = v;
}
// A step's FP32 population store.  In the blocked layout (blocks of 64
// slots) a wave's 64 stores of one direction fill whole lines that the next
// step reads back once; the nontemporal hint made the 256^3 sparse FP32 step
// 5 % and the sparse colour pair 3 % faster, and the dense all-sites step
// 43 % slower (its solid sites' copies interleave with fluid writes), so the
// dense layouts keep plain stores.  Nontemporal loads were slower in every
// layout.
template <bool BLK, typename T, typename O> __device__ __forceinline__ void st_pop(T* __restrict__ base, O elem, T v)
{
    if constexpr (BLK)
        __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + elem * O(sizeof(T))));
    else
        st(base, elem, v);
}

}  // namespace lbf
