// device.h -- structures shared by the HIP kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lattice.h"

namespace lbf {

// Site kinds of the label table.
enum : int32_t {
    KIND_FLUID = 0x1,    // macro + collision site (VD, GV)
    KIND_FORCED = 0x2,   // GV with a non-zero force vector
    KIND_GV = 0x4,       // GV: body force shifts the velocity
    KIND_ZOUHE = 0x8,    // PB / VB: density and velocity from the Zou-He boundary
    KIND_FILM = 0x10,    // HF: half-permeable film of the label's phase
};

// Zou-He boundary of one label (SCSP::zouHeBoundary, scsp.cpp:188-256).
// FP64 uses type/dir/value; FP32 works on shifted populations h = f - w rho0:
//   VB: drho = (sum_h + b) c, u_axis = a   (a = velocity with the sign of
//       dir, b = rho0 a, c = 1 / (1 - a));
//   PB: s = (sum_h + b) c, u_axis = -+s, drho = -b   (b = rho0 - rho_b,
//       c = 1 / rho_b).
struct DevZouHe {
    int32_t type;        // 1 VB, 2 PB
    int32_t dir;         // axial lattice direction 1..2D
    double value;        // VB speed / PB density
    float a, b, c, pad;
};
enum : int32_t { ZH_VB = 1, ZH_PB = 2 };

// One entry of the 256-entry device label table.
struct DevLabel {
    int32_t kind;
    int32_t phase;
    double force[3];     // Qp[dir] * value, reference component order
    double half_rho[3];  // (force * 0.5) / rho0           (scsp body force)
    double half[3];      // force * 0.5                     (color body force)
    double ef[19];       // e_q . force (Guo source; initial shift)
};

// Per wall entry of the colour model: sin/cos of the mean contact angle and
// the rotations TibNorm / TibNormInv (colorfluid.cpp:1111-1129).
struct DevWall {
    double sinca, cosca;
    double tib[9];
    double tibinv[9];
    double ginv[9];      // renormalized wall gradient: inverse fluid-link moment (colorfluid.cpp:917-957)
};

// A wall entry in FP32 for the FP32 colour kernels: the frame and the
// contact angle as the kernels use them (the float of each FP64 value); 48
// bytes, three 16-byte loads.  The inverse frame is the transpose of the
// frame, bit for bit (the host checks it when it packs the entries), and the
// renormalizing inverse, read only with renormalized walls, is an array of
// its own (ColorArgs::ginv32).
struct __attribute__((aligned(16))) WallF32 {
    float tib[9];
    float sinca, cosca;
    float pad;
};

// Device failure record.
struct DevStatus {
    uint32_t bits;
    uint32_t pad;
    unsigned long long first[6];   // first step per bit (LBF_FAIL_*), ~0 = never
};

struct DevFastDiv {
    uint32_t d, m, s;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, DevFastDiv f)
{
    const uint32_t t = __umulhi(n, f.m);
    return uint32_t((uint64_t(t) + n) >> f.s);
}

struct GridDev {
    uint32_t nx, ny, nz, nxy;
    uint64_t N;
    DevFastDiv div_nx, div_nxy;
};

// Workgroup index remapped so that each of the 8 XCDs (each with its own L2)
// processes one contiguous range of the grid: the hardware hands workgroup b
// to XCD b % 8, and neighbouring sites then share that XCD's L2 lines.
// A bijection on [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    const uint32_t per = nb >> 3, rem = nb & 7u, x = b & 7u;
    return x * per + (x < rem ? x : rem) + (b >> 3);
}

__device__ __forceinline__ void site_coords(const GridDev& g, uint32_t i, int& x, int& y, int& z)
{
    const uint32_t zz = fdiv(i, g.div_nxy);
    const uint32_t r = i - zz * g.nxy;
    const uint32_t yy = fdiv(r, g.div_nx);
    z = int(zz);
    y = int(yy);
    x = int(r - yy * g.nx);
}

__device__ __forceinline__ int wrap(int c, int n)
{
    return c < 0 ? c + n : (c >= n ? c - n : c);
}

// Pull source of direction q at (x,y,z) for lattice Q (components slowest first).
template <int Q>
__device__ __forceinline__ uint32_t pull_source(const GridDev& g, int x, int y, int z, int q)
{
    using L = Lattice<Q>;
    int ex, ey, ez;
    if constexpr (Q == 9) {
        ez = 0; ey = L::e[q][0]; ex = L::e[q][1];
    } else {
        ez = L::e[q][0]; ey = L::e[q][1]; ex = L::e[q][2];
    }
    const int xs = ex ? wrap(x - ex, int(g.nx)) : x;
    const int ys = ey ? wrap(y - ey, int(g.ny)) : y;
    const int zs = ez ? wrap(z - ez, int(g.nz)) : z;
    return uint32_t(zs) * g.nxy + uint32_t(ys) * g.nx + uint32_t(xs);
}

// Neighbour x + e_q (gradient stencil).
template <int Q>
__device__ __forceinline__ uint32_t push_target(const GridDev& g, int x, int y, int z, int q)
{
    using L = Lattice<Q>;
    int ex, ey, ez;
    if constexpr (Q == 9) {
        ez = 0; ey = L::e[q][0]; ex = L::e[q][1];
    } else {
        ez = L::e[q][0]; ey = L::e[q][1]; ex = L::e[q][2];
    }
    const int xs = ex ? wrap(x + ex, int(g.nx)) : x;
    const int ys = ey ? wrap(y + ey, int(g.ny)) : y;
    const int zs = ez ? wrap(z + ez, int(g.nz)) : z;
    return uint32_t(zs) * g.nxy + uint32_t(ys) * g.nx + uint32_t(xs);
}

// wallGradient tanh: atanh of the colour clamped to (-1, 1) by 1e-8
// (colorfluid.cpp:497-499), in FP64 for both engines (in FP32 the clamp
// would round to +-1 and atanh to infinity)
__device__ __forceinline__ double tanh_wall_transform(double color)
{
    const double hi = 1.0 - 1e-8, lo = -1.0 + 1e-8;
    const double c = color < hi ? color : hi;
    return atanh(lo < c ? c : lo);
}

__device__ __forceinline__ void raise_status(DevStatus* st, int bit, unsigned long long t)
{
    atomicOr(&st->bits, 1u << bit);
    atomicMin(&st->first[bit], t);
}

// Single-precision label entry of the FP32 kernels.
struct LabelF {
    int32_t kind;
    float force[3];
    float half_rho[3];
    float pad;
};

// Scalar parameters of the single-phase step kernels (by value).
struct ScspParams {
    GridDev g;
    uint32_t nfluid;         // SPARSE layout: number of fluid sites
    uint32_t flags;          // bit0 fields, bit1 adjacent
    unsigned long long t;    // step being computed (failure records)
    double dens0, omega, omh; // omh = 1 - 0.5 omega
    double idist[19];        // w_q * rho0
    float inv_rho0, rho0_32, omega32, omh32;
    float c1_nu, c2_nu, c1_q, c2_q;   // FP32 MRT: s and 1 - s/2 of the two non-conserved rate classes
    // dense layouts: workgroups walk columns of tile_chunks 256-site chunks
    // through all z layers (layer_chunks chunks per layer), so the z +- 1
    // rows a workgroup pulls were just read by its neighbours; 0: site order
    uint32_t tile_chunks, layer_chunks;
};

// Device buffers of the single-phase step kernels.
struct ScspBuffers {
    const void* fin;
    void* fout;
    const uint32_t* desc;    // descriptor per site (SPARSE: per slot)
    const uint32_t* list;    // SPARSE: fluid site of each slot
    const uint2* rank;       // SPARSE: per 32 sites {fluid sites before them, fluid bit mask}
    const DevLabel* lab;
    const LabelF* lab32;
    const double* AT;        // A transposed (MInvSMT of scsp.cpp:425)
    const double* B;         // force matrix (scsp.cpp:426-427)
    double* rho;             // fields: density
    double* u;               // fields: velocity [d][site]
    double* uadj;            // adjacent velocity history
    DevStatus* status;
    const DevZouHe* zh;      // Zou-He boundary per label
};

struct ColorArgs {
    GridDev g;
    const double* fin;       // [phase][q][site]
    double* fout;
    const uint32_t* desc;
    const DevLabel* lab;
    const int32_t* wall;
    const DevWall* walls;
    double* densN;           // colour indicator with solid sentinel
    const double* A;         // uniform-tau collision matrix (row-major)
    const double* B;         // uniform-tau force matrix
    const double* MInv;      // inverse transform (row-major)
    double* rho0;
    double* rho1;
    double* u;
    double* uadj;
    double* chiadj;
    // adjacent history read by this step's report (written by the previous
    // step); PB/VB sites put their boundary values there on field steps
    double* uadj_prev;
    double* chiadj_prev;
    const DevZouHe* zh;      // Zou-He boundary per label
    int32_t renormalized;    // wallGradient renormalized (colorfluid.cpp:487-492)
    DevStatus* status;
    unsigned long long t;
    uint32_t flags;
    int32_t mrt, uniform, harmonic;
    double relax0, relax1;   // tau_k / 2
    double nu0, nu1;         // (tau_k - 0.5) / 3
    double ak, beta;
    double ai0[19], ai1[19], bi[19];
    double dist0[19], dist1[19];   // ai_k * rho_k
    // FP32 form: planes [0,Q) hold h = fT - dist0 (the total population shifted by
    // phase 0's rest state), planes [Q,2Q) hold f0; densN32 is the FP32 indicator.
    const float* fin32;
    float* fout32;
    float* densN32;
    double mai0[19], mai1[19];     // M ai_k: rest-state moments of each phase
    double rho_ref;                // rho_0, the density of the shift s_q = ai0_q rho_0
    // sum over q of float(s_q): the density of the FP32 shift as the kernels
    // add it back (fT_q = h_q + float(s_q)); it differs from rho_0 by about
    // 1.5e-8, which recolouring must not turn into a mass transfer
    double rest32;
    double dai[19], mdai[19];         // ai1 - ai0 and M (ai1 - ai0)
    double rest_err[19], mrest_err[19];   // s_q - float(s_q) and M of it
    // FP32 layouts: DENSE_FLUID (slot = site) or SPARSE (slot = fluid index)
    const uint32_t* list;          // SPARSE: fluid site of each slot
    float* dens_slot;              // SPARSE: the FP32 indicator per fluid slot
    // the FP32 kernels' constants as they use them (the float of each FP64
    // value above, double(float(dist0)) for the FP64 recolouring), so that
    // they enter the arithmetic as scalar operands instead of being
    // converted per site
    float ai0f[19], ai1f[19], bif[19], dist0f[19], daif[19], rest_errf[19];
    float mai0f[19], mdaif[19], mrest_errf[19];
    double dist0r[19];
    float relax0f, relax1f, nu0f, nu1f, betaf, akf, rho_reff;
    const WallF32* walls32;        // the wall entries in FP32
    const float* ginv32;           // renormalized walls: ginv per wall entry in FP32 (9 floats)
    const int32_t* wall_slot;      // SPARSE: the wall entry per fluid slot
    const uint2* rank;             // SPARSE: rank index (see stream.h)
    uint32_t P;                    // slots per plane
    // Half-permeable films (simulator.cpp:418-421, colorfluid.cpp:156-159):
    // per site, bit q = phase 0 bounces back at link q (the q-neighbour is a
    // film of phase 1), bit 32 + q = phase 1 bounces back (film of phase 0);
    // read only where the descriptor has DESC_FILM.
    const uint64_t* film;
    // wallGradient tanh (colorfluid.cpp:494-518): 1 = tanh reconstruction at
    // walls, computed per wall entry into tanhg [wall][3] by its own pass
    int32_t tanh_wall;
    double* tanhg;
    // capillaryModel csf (colorfluid.cpp:572-634): per site [d][site] the colour
    // gradient and the curvature field of the VD/GV sites (0 elsewhere)
    int32_t csf, curv_gradient;
    double* capgrad;
    double* capfield;
    double sigma;
};

struct ReportArgs {
    GridDev g;
    const uint32_t* desc;
    const DevLabel* lab;
    const double* rho0;
    const double* rho1;
    const double* u;
    double* uprev;
    const double* uadj;
    double* chiprev;
    const double* chiadj;
    double* partial;         // [blocks][NREP]
    int32_t D, axis, color, conv, adjacent, space_mask_unused;
    uint32_t first, last;
};
constexpr int NREP = 16;

}  // namespace lbf
