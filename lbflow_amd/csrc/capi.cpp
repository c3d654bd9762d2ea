// capi.cpp -- implementation of include/lbflow.h on one MI355X (HIP runtime).
//
// The context owns every device buffer.  Populations live in two SoA
// buffers [phase][q][site] that swap roles each step (pull scheme: read A,
// write B).  Everything the kernels read per site is a 32-bit descriptor
// (bounce-back link bits + geometry label), a 256-entry label table and,
// for the colour model, a wall index with per-wall rotation tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/lbflow.h"
#include "device.h"
#include "setup.h"

namespace lbf {
hipError_t launch_scsp(int Q, bool mrt, bool init, bool fp32, int layout, const ScspParams& p,
                       const ScspBuffers& b, hipStream_t s);
hipError_t launch_color(int Q, bool mrt, bool init, bool stream_solids, const ColorArgs& a, hipStream_t s);
hipError_t launch_color32(int Q, bool mrt, bool init, bool sparse, const ColorArgs& a, hipStream_t s);
hipError_t launch_report(const ReportArgs& a, double* out, hipStream_t s);
uint32_t report_blocks(uint64_t N);
hipError_t report_scratch(double** ptr, uint64_t* capacity);
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
}  // namespace lbf

using namespace lbf;

struct lbf_ctx {
    lbf_desc d{};
    int D = 0, Q = 0, NP = 1;
    uint64_t N = 0, nfluid = 0;
    std::vector<LabelInfo> labels;
    Links links;
    MrtOperators ops;
    std::vector<DevLabel> hlab;
    std::vector<double> shift;      // FP32: per plane shift (phase*Q + q)
    std::vector<uint32_t> list;     // SPARSE: fluid site of each slot
    hipStream_t stream = nullptr;
    int device = 0;                 // the GPU current at lbf_create, which holds every buffer
    void* f[2] = {nullptr, nullptr};
    int cur = 0;
    int layout = 0;                 // 0 dense (all sites streamed), 1 dense fluid-only, 2 sparse
    uint64_t P = 0;                 // slots per population plane
    // sparse layout: populations in blocks of 64 slots, [s / 64][plane][s % 64]
    // (stream.h pop_index); the dense layouts store plane after plane, q P + s
    bool blocked = false;
    uint64_t pop_elems = 0;         // elements of one population buffer
    size_t elem = 8;
    uint32_t* d_desc = nullptr;
    uint32_t* d_list = nullptr;
    uint2* d_rank = nullptr;         // sparse: per 32 sites {fluid sites before, fluid mask}
    uint32_t* d_desc_s = nullptr;    // sparse: descriptor per slot
    DevLabel* d_lab = nullptr;
    LabelF* d_lab32 = nullptr;
    DevZouHe* d_zh = nullptr;
    double *d_AT = nullptr, *d_A = nullptr, *d_B = nullptr, *d_MInv = nullptr;
    float *d_A32 = nullptr, *d_B32 = nullptr;
    int32_t* d_wall = nullptr;
    DevWall* d_walls = nullptr;
    WallF32* d_walls32 = nullptr;   // FP32 colour: the wall entries in FP32
    float* d_ginv32 = nullptr;      // FP32 colour, renormalized walls: ginv per wall entry
    int32_t* d_wall_slot = nullptr; // FP32 colour, sparse: wall entry per fluid slot
    float* d_dens_slot = nullptr;   // FP32 colour, sparse: indicator per fluid slot
    uint64_t* d_film = nullptr;      // per-site film links, only with HF labels
    double *d_capgrad = nullptr, *d_capfield = nullptr;   // CSF gradient and curvature field
    double* d_tanhg = nullptr;       // tanh wall gradient per wall entry
    double* d_densN = nullptr;
    double *d_rho0 = nullptr, *d_rho1 = nullptr, *d_u = nullptr, *d_uadj = nullptr, *d_uprev = nullptr;
    double *d_chiadj = nullptr, *d_chiprev = nullptr;
    DevStatus* d_status = nullptr;
    double *d_partial = nullptr, *d_out = nullptr;
    GridDev g{};
    uint64_t t = 0;
    bool has_prev = false;
    uint64_t dev_bytes = 0;
    std::vector<void*> allocations;
    std::string err;
    std::string kernel_name;
};

namespace {

thread_local std::string g_create_error;

int fail(lbf_ctx* c, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return -1;
}

#define HIP_CHECK(ctx, expr)                                                                   \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(ctx, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

template <typename T> int dalloc(lbf_ctx* c, T** p, size_t bytes)
{
    void* q = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(c, std::string("HIP error: hipMalloc failed: ") + hipGetErrorString(e));
    c->allocations.push_back(q);
    c->dev_bytes += bytes;
    *p = static_cast<T*>(q);
    return 0;
}

template <typename T> int upload(lbf_ctx* c, T* dst, const void* src, size_t bytes)
{
    HIP_CHECK(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

// Qp of the reference (gridspace.cpp:85-87, :112-116): unit vector of a link.
double unit_component(int Q, int q, int d)
{
    const int e = Q == 9 ? Lattice<9>::e[q][d] : Lattice<19>::e[q][d];
    if (e == 0) return 0.0;
    const bool diag = Q == 9 ? Lattice<9>::diagonal(q) : Lattice<19>::diagonal(q);
    const double sq = std::sqrt(2) / 2;
    if (!diag) return double(e);
    return e > 0 ? sq : -sq;
}

int lattice_e(int Q, int q, int d) { return Q == 9 ? Lattice<9>::e[q][d] : Lattice<19>::e[q][d]; }

// Pair-split form of an operator with the lattice parity symmetry (see
// apply_split in kernels_scsp.hip).
std::vector<float> pair_split(int Q, const std::vector<double>& op)
{
    const int H = Q == 9 ? Lattice<9>::NHALF : Lattice<19>::NHALF;
    const int* half = Q == 9 ? Lattice<9>::half : Lattice<19>::half;
    const int* opp = Q == 9 ? Lattice<9>::op : Lattice<19>::op;
    std::vector<float> c(size_t((H + 1) * (H + 1) + H * H));
    for (int r = 0; r <= H; ++r) {
        const int q = r == 0 ? 0 : half[r - 1];
        c[r * (H + 1)] = float(op[q * Q + 0]);
        for (int k = 0; k < H; ++k) {
            const int p = half[k];
            c[r * (H + 1) + k + 1] = float(0.5 * (op[q * Q + p] + op[q * Q + opp[p]]));
        }
    }
    for (int r = 1; r <= H; ++r) {
        const int q = half[r - 1];
        for (int k = 0; k < H; ++k) {
            const int p = half[k];
            c[(H + 1) * (H + 1) + (r - 1) * H + k] = float(0.5 * (op[q * Q + p] - op[q * Q + opp[p]]));
        }
    }
    return c;
}

}  // namespace

extern "C" {

const char* lbf_last_error(const lbf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

const char* lbf_status_message(uint32_t bit)
{
    switch (bit) {
    case LBF_FAIL_DENSITY: return "Non-finite or non-positive fluid density.";
    case LBF_FAIL_NEGATIVE: return "Negative component density outside roundoff tolerance.";
    case LBF_FAIL_TAU: return "Invalid local relaxation time.";
    case LBF_FAIL_VELOCITY: return "Non-finite fluid velocity.";
    case LBF_FAIL_POPULATION: return "Non-finite post-collision population; no restart will be written.";
    case LBF_FAIL_CAPILLARY: return "Non-finite capillary force.";
    default: return "";
    }
}

void lbf_destroy(lbf_ctx* c)
{
    if (!c) return;
    for (void* p : c->allocations) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lbf_create(const lbf_desc* desc, const lbf_label* labels, const uint8_t* geometry, lbf_ctx** out)
{
    g_create_error.clear();
    if (!desc || !labels || !geometry || !out) return fail(nullptr, "lbf_create: null argument");
    *out = nullptr;
    const lbf_desc& d = *desc;
    if (!((d.D == 2 && d.Q == 9) || (d.D == 3 && d.Q == 19)))
        return fail(nullptr, "Only D2Q9 and D3Q19 are validated for simulation.");
    if (d.model != LBF_MODEL_SCSP && d.model != LBF_MODEL_COLOR)
        return fail(nullptr, "Only scsp and color fluid models are implemented.");
    if (d.equil != LBF_SRT && d.equil != LBF_MRT) return fail(nullptr, "Equilibrium model must be SRT or MRT.");
    if (d.precision != LBF_FP64 && d.precision != LBF_FP32) return fail(nullptr, "lbf_create: unknown precision");
    const int NP = d.model == LBF_MODEL_COLOR ? 2 : 1;
    uint64_t N = 1;
    for (int k = 0; k < d.D; ++k) {
        if (d.sz[k] == 0) return fail(nullptr, "Invalid grid dimensions.");
        N *= d.sz[k];
    }
    if (N >= (uint64_t(1) << 32)) return fail(nullptr, "lbf_create: more than 2^32-1 sites on one device");
    for (int k = 0; k < NP; ++k)
        if (!std::isfinite(d.tau[k]) || d.tau[k] <= 0.5 || !std::isfinite(d.density[k]) || d.density[k] <= 0.0)
            return fail(nullptr, "Densities must be positive; relaxation times must exceed 0.5.");

    lbf_ctx* c = new lbf_ctx();
    c->d = d;
    c->D = d.D;
    c->Q = d.Q;
    c->NP = NP;
    c->N = N;
    c->elem = d.precision == LBF_FP32 ? 4 : 8;
    c->labels.resize(256);
    for (int l = 0; l < 256; ++l) {
        c->labels[l].type = labels[l].type;
        c->labels[l].dir = labels[l].dir;
        c->labels[l].phase = labels[l].phase;
        c->labels[l].value = labels[l].value;
    }
    // Labels used by the geometry must be supported by the kernels.
    {
        bool used[256] = {false};
        for (uint64_t i = 0; i < N; ++i) used[geometry[i]] = true;
        for (int l = 0; l < 256; ++l) {
            if (!used[l] || !labels[l].type) continue;
            const int t = labels[l].type;
            if (t == LBF_AB) {
                lbf_destroy(c);
                return fail(nullptr, "lbf_create: label " + std::to_string(l) +
                                         " uses a boundary type (AB) this engine does not implement");
            }
            if (t == LBF_HF && d.model != LBF_MODEL_COLOR) {
                // the reference's single-phase model never initialises or
                // collides HF sites (scsp.cpp:116-175, 374-464)
                lbf_destroy(c);
                return fail(nullptr, "lbf_create: label " + std::to_string(l) +
                                         " is a half-permeable film (HF), which this engine supports in the colour model only");
            }
            if ((t == LBF_PB || t == LBF_VB) && (labels[l].dir < 1 || labels[l].dir > 2 * d.D)) {
                lbf_destroy(c);
                return fail(nullptr, "Pressure and velocity boundaries require an axial direction.");
            }
            if (t == LBF_PB && !(labels[l].value > 0.0)) {
                lbf_destroy(c);
                return fail(nullptr, "Pressure-boundary density must be positive.");
            }
            if ((t == LBF_VD || t == LBF_GV || t == LBF_PB || t == LBF_VB || t == LBF_HF) && labels[l].phase >= NP) {
                lbf_destroy(c);
                return fail(nullptr, "Boundary phase is outside the configured fluid model.");
            }
        }
    }
    std::string msg = build_links(d.D, d.Q, d.sz, c->labels.data(), geometry, d.model == LBF_MODEL_COLOR,
                                  c->links);
    // the wall moment inverse serves renormalized and tanh walls and CSF (colorfluid.cpp:917-920)
    if (msg.empty() && d.model == LBF_MODEL_COLOR &&
        (d.wall_gradient == LBF_WALL_RENORMALIZED || d.wall_gradient == LBF_WALL_TANH ||
         d.capillary == LBF_CAPILLARY_CSF))
        msg = wall_gradient_inverse(d.D, d.Q, d.sz, c->labels.data(), geometry, c->links);
    if (!msg.empty()) {
        lbf_destroy(c);
        return fail(nullptr, msg);
    }
    if (d.wall_gradient != LBF_WALL_LEGACY && d.wall_gradient != LBF_WALL_RENORMALIZED &&
        d.wall_gradient != LBF_WALL_TANH) {
        lbf_destroy(c);
        return fail(nullptr, "wallGradient must be legacy, renormalized or tanh in this engine.");
    }
    if (d.capillary == LBF_CAPILLARY_CSF) {
        // closed flow only, as the reference requires without pressure
        // reservoirs (fileprm.cpp validation of capillaryModel csf)
        bool closed = d.model == LBF_MODEL_COLOR;
        for (uint64_t i = 0; closed && i < N; ++i) {
            const int t = labels[geometry[i]].type;
            closed = t == LBF_VD || t == LBF_GV || t == LBF_SD || t == LBF_BB;
        }
        if (!closed) {
            lbf_destroy(c);
            return fail(nullptr, "CSF requires closed flow or guarded flat ghost/flux pressure reservoirs.");
        }
    } else if (d.capillary != LBF_CAPILLARY_PERTURBATION) {
        lbf_destroy(c);
        return fail(nullptr, "capillaryModel must be perturbation or csf.");
    }

    // label table (simulator.cpp:218-228 forceForIndex)
    c->hlab.assign(256, DevLabel{});
    for (int l = 0; l < 256; ++l) {
        DevLabel& L = c->hlab[l];
        std::memset(&L, 0, sizeof L);
        const LabelInfo& li = c->labels[l];
        if (li.type == T_VD || li.type == T_GV) L.kind = KIND_FLUID;
        if (li.type == T_PB || li.type == T_VB) L.kind = KIND_FLUID | KIND_ZOUHE;
        if (li.type == T_HF) L.kind = KIND_FLUID | KIND_FILM;
        L.phase = li.phase;
        if (li.type == T_GV) {
            if (li.dir < 0 || li.dir >= d.Q || !std::isfinite(li.value)) {
                lbf_destroy(c);
                return fail(nullptr, "Invalid body-force direction or magnitude.");
            }
            L.kind |= KIND_GV;
            double ff = 0.0;
            for (int k = 0; k < d.D; ++k) {
                L.force[k] = unit_component(d.Q, li.dir, k) * li.value;
                ff = ff + L.force[k] * L.force[k];
            }
            if (ff != 0.0) L.kind |= KIND_FORCED;
            for (int k = 0; k < d.D; ++k) {
                L.half[k] = L.force[k] * 0.5;
                L.half_rho[k] = L.half[k] / d.density[0];
            }
            for (int q = 0; q < d.Q; ++q) {
                double ef = 0.0;
                for (int k = 0; k < d.D; ++k) {
                    const int e = lattice_e(d.Q, q, k);
                    if (e == 1) ef += L.force[k];
                    else if (e == -1) ef += -L.force[k];
                }
                L.ef[q] = ef;
            }
        }
    }
    bool all_bounce = true;   // every non-fluid label bounces back (BB or SD)
    for (uint64_t i = 0; i < N; ++i) {
        if (c->hlab[geometry[i]].kind & KIND_FLUID) c->list.push_back(uint32_t(i));
        else if (!(c->labels[geometry[i]].type & (T_BB | T_SD))) all_bounce = false;
    }
    c->nfluid = c->list.size();
    // Layout: the FP64 colour model and checkpoint-exact runs stream every
    // site; otherwise only fluid sites are stored (sparse) unless the dense
    // fluid-only layout is asked for.  The FP32 colour model is sparse when
    // every non-fluid site bounces back, else dense with solid sites skipped.
    if (NP == 2 && d.precision != LBF_FP32) c->layout = 0;
    else if (NP == 2) c->layout = (all_bounce && !(d.flags & LBF_LAYOUT_DENSE)) ? 2 : 1;
    else if (d.flags & LBF_STREAM_SOLIDS) c->layout = 0;
    else c->layout = (d.flags & LBF_LAYOUT_DENSE) ? 1 : 2;
    c->P = c->layout == 2 ? c->nfluid : N;
    c->blocked = c->layout == 2;
    c->pop_elems = uint64_t(NP) * d.Q * (c->blocked ? (c->P + 63) / 64 * 64 : c->P);

    build_mrt(d.Q, d.tau[0], c->ops);

    // grid
    c->g.nx = uint32_t(d.sz[d.D - 1]);
    c->g.ny = uint32_t(d.sz[d.D - 2]);
    c->g.nz = d.D == 3 ? uint32_t(d.sz[0]) : 1u;
    c->g.nxy = c->g.nx * c->g.ny;
    c->g.N = N;
    const FastDiv fx = make_fastdiv(c->g.nx), fxy = make_fastdiv(c->g.nxy);
    c->g.div_nx = DevFastDiv{fx.d, fx.m, fx.s};
    c->g.div_nxy = DevFastDiv{fxy.d, fxy.m, fxy.s};

    // FP32 shift per plane
    c->shift.assign(size_t(NP) * d.Q, 0.0);
    if (d.precision == LBF_FP32)
        for (int q = 0; q < d.Q; ++q)
            c->shift[q] = (d.Q == 9 ? Lattice<9>::w[q] : Lattice<19>::w[q]) * d.density[0];

    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        lbf_destroy(c);
        return fail(nullptr, "HIP error: cannot create a stream (no GPU?)");
    }
    (void)hipGetDevice(&c->device);
    const size_t pop = c->pop_elems * c->elem;
    const size_t Q2 = size_t(d.Q) * d.Q;
    int rc = 0;
    rc |= dalloc(c, &c->f[0], pop);
    rc |= dalloc(c, &c->f[1], pop);
    rc |= dalloc(c, &c->d_desc, N * 4);
    rc |= dalloc(c, &c->d_lab, 256 * sizeof(DevLabel));
    rc |= dalloc(c, &c->d_lab32, 256 * sizeof(LabelF));
    rc |= dalloc(c, &c->d_zh, 256 * sizeof(DevZouHe));
    if (c->layout == 2) {
        rc |= dalloc(c, &c->d_list, c->nfluid * 4);
        rc |= dalloc(c, &c->d_rank, (N + 31) / 32 * sizeof(uint2));
        rc |= dalloc(c, &c->d_desc_s, c->nfluid * 4);
    }
    rc |= dalloc(c, &c->d_AT, Q2 * 8);
    rc |= dalloc(c, &c->d_A, Q2 * 8);
    rc |= dalloc(c, &c->d_B, Q2 * 8);
    rc |= dalloc(c, &c->d_MInv, Q2 * 8);
    rc |= dalloc(c, &c->d_A32, Q2 * 4);
    rc |= dalloc(c, &c->d_B32, Q2 * 4);
    rc |= dalloc(c, &c->d_rho0, N * 8);
    rc |= dalloc(c, &c->d_u, size_t(d.D) * N * 8);
    // adjacent history, two buffers by step parity: step t writes buffer t & 1,
    // the report of step t reads buffer (t - 1) & 1
    rc |= dalloc(c, &c->d_uadj, 2 * size_t(d.D) * N * 8);
    rc |= dalloc(c, &c->d_uprev, size_t(d.D) * N * 8);
    rc |= dalloc(c, &c->d_status, sizeof(DevStatus));
    rc |= dalloc(c, &c->d_partial, size_t(report_blocks(N)) * NREP * 8);
    rc |= dalloc(c, &c->d_out, NREP * 8);
    if (NP == 2) {
        rc |= dalloc(c, &c->d_rho1, N * 8);
        rc |= dalloc(c, &c->d_chiadj, 2 * N * 8);
        rc |= dalloc(c, &c->d_chiprev, N * 8);
        rc |= dalloc(c, &c->d_densN, N * 8);
        rc |= dalloc(c, &c->d_wall, N * 4);
        rc |= dalloc(c, &c->d_walls, c->links.nwall * sizeof(DevWall));
        if (c->elem == 4) {
            rc |= dalloc(c, &c->d_walls32, c->links.nwall * sizeof(WallF32));
            if (!c->links.ginv.empty()) rc |= dalloc(c, &c->d_ginv32, c->links.nwall * 9 * sizeof(float));
            if (c->layout == 2) {
                rc |= dalloc(c, &c->d_dens_slot, c->nfluid * sizeof(float));
                rc |= dalloc(c, &c->d_wall_slot, c->nfluid * sizeof(int32_t));
            }
        }
        if (!c->links.film.empty()) rc |= dalloc(c, &c->d_film, N * sizeof(uint64_t));
        if (d.wall_gradient == LBF_WALL_TANH) rc |= dalloc(c, &c->d_tanhg, c->links.nwall * 3 * 8);
        if (d.capillary == LBF_CAPILLARY_CSF) {
            rc |= dalloc(c, &c->d_capgrad, size_t(d.D) * N * 8);
            rc |= dalloc(c, &c->d_capfield, size_t(d.D) * N * 8);
        }
    }
    if (rc) {
        std::string e = c->err;
        lbf_destroy(c);
        return fail(nullptr, e);
    }
    rc |= upload(c, c->d_desc, c->links.desc.data(), N * 4);
    rc |= upload(c, c->d_lab, c->hlab.data(), 256 * sizeof(DevLabel));
    {
        // Zou-He boundaries (scsp.cpp:188-256); FP32 constants as in DevZouHe
        std::vector<DevZouHe> zh(256);
        for (int l = 0; l < 256; ++l) {
            std::memset(&zh[l], 0, sizeof(DevZouHe));
            const LabelInfo& li = c->labels[l];
            if ((li.type != T_PB && li.type != T_VB) || li.dir < 1 || li.dir > 2 * d.D) continue;
            const int sgn = d.D == 3 ? ZouHe<19>::sign[li.dir] : ZouHe<9>::sign[li.dir];
            const double rho0 = d.density[0];
            zh[l].dir = li.dir;
            zh[l].value = li.value;
            if (li.type == T_VB) {
                const double v = sgn == 1 ? li.value : -li.value;
                zh[l].type = ZH_VB;
                zh[l].a = float(v);
                zh[l].b = float(rho0 * v);
                zh[l].c = float(1.0 / (1.0 - v));
            } else {
                zh[l].type = ZH_PB;
                zh[l].b = float(rho0 - li.value);
                zh[l].c = float(1.0 / li.value);
            }
        }
        rc |= upload(c, c->d_zh, zh.data(), 256 * sizeof(DevZouHe));
    }
    rc |= upload(c, c->d_AT, c->ops.AT.data(), Q2 * 8);
    rc |= upload(c, c->d_A, c->ops.A.data(), Q2 * 8);
    rc |= upload(c, c->d_B, c->ops.B.data(), Q2 * 8);
    rc |= upload(c, c->d_MInv, c->ops.MInv.data(), Q2 * 8);
    {
        std::vector<float> a32 = pair_split(d.Q, c->ops.A), b32 = pair_split(d.Q, c->ops.B);
        rc |= upload(c, c->d_A32, a32.data(), a32.size() * 4);
        rc |= upload(c, c->d_B32, b32.data(), b32.size() * 4);
        std::vector<LabelF> l32(256);
        for (int l = 0; l < 256; ++l) {
            std::memset(&l32[l], 0, sizeof(LabelF));
            l32[l].kind = c->hlab[l].kind;
            for (int k = 0; k < d.D; ++k) {
                l32[l].force[k] = float(c->hlab[l].force[k]);
                l32[l].half_rho[k] = float(c->hlab[l].half_rho[k]);
            }
        }
        rc |= upload(c, c->d_lab32, l32.data(), 256 * sizeof(LabelF));
    }
    if (c->layout == 2) {
        // Every non-fluid site must bounce back (BB/SD), so a pull from it is a
        // link of the descriptor and the rank index only ever sees fluid sites.
        if (!all_bounce) {
            lbf_destroy(c);
            return fail(nullptr, "lbf_create: sparse layout needs every non-fluid label to be BB or SD");
        }
        std::vector<uint2> rank((N + 31) / 32, uint2{0u, 0u});
        uint32_t before = 0;
        for (uint64_t w = 0; w < rank.size(); ++w) {
            rank[w].x = before;
            for (uint64_t i = w * 32; i < std::min<uint64_t>(N, w * 32 + 32); ++i)
                if (c->hlab[geometry[i]].kind & KIND_FLUID) {
                    rank[w].y |= 1u << (i - w * 32);
                    ++before;
                }
        }
        std::vector<uint32_t> desc_s(c->nfluid);
        for (uint64_t k = 0; k < c->nfluid; ++k) desc_s[k] = c->links.desc[c->list[k]];
        rc |= upload(c, c->d_list, c->list.data(), c->nfluid * 4);
        rc |= upload(c, c->d_rank, rank.data(), rank.size() * sizeof(uint2));
        rc |= upload(c, c->d_desc_s, desc_s.data(), c->nfluid * 4);
    }
    if (NP == 2) {
        if (c->elem == 4) {
            std::vector<float> dn(N, 86.0f);   // solid sentinel, colorfluid.cpp:49
            rc |= upload(c, c->d_densN, dn.data(), N * 4);
        } else {
            std::vector<double> dn(N, 86.0);
            rc |= upload(c, c->d_densN, dn.data(), N * 8);
        }
        rc |= upload(c, c->d_wall, c->links.wall.data(), N * 4);
        std::vector<DevWall> walls(c->links.nwall);
        const int DD = d.D * d.D;
        for (uint64_t w = 0; w < c->links.nwall; ++w) {
            std::memset(&walls[w], 0, sizeof(DevWall));
            walls[w].sinca = c->links.sinca[w];
            walls[w].cosca = c->links.cosca[w];
            for (int k = 0; k < DD; ++k) {
                walls[w].tib[k] = c->links.tib[w * DD + k];
                walls[w].tibinv[k] = c->links.tibinv[w * DD + k];
                if (!c->links.ginv.empty()) walls[w].ginv[k] = c->links.ginv[w * DD + k];
            }
        }
        rc |= upload(c, c->d_walls, walls.data(), walls.size() * sizeof(DevWall));
        if (c->d_walls32) {
            // The FP32 kernels take the inverse frame as the transpose of the
            // frame: tibinv = txInv tzInv multiplies the transposes of tz and
            // tx (setup.cpp, rotation(-o, a) = (c, -s) exactly), the same
            // products summed in the same order, so the two are equal (as
            // values: where a rotation degenerates to the identity, s = 0 and
            // -s = -0 trade places, and a zero may differ in sign).
            const int D = d.D;
            bool transpose = true;
            for (uint64_t w = 0; w < c->links.nwall && transpose; ++w)
                for (int r = 0; r < D; ++r)
                    for (int k = 0; k < D; ++k)
                        transpose = transpose && walls[w].tibinv[r * D + k] == walls[w].tib[k * D + r];
            if (!transpose) rc |= fail(c, "wall frames: the inverse frame is not the transpose of the frame");
            std::vector<WallF32> w32(c->links.nwall);
            std::vector<float> g32(c->d_ginv32 ? c->links.nwall * 9 : 0);
            for (uint64_t w = 0; w < c->links.nwall; ++w) {
                std::memset(&w32[w], 0, sizeof(WallF32));
                w32[w].sinca = float(walls[w].sinca);
                w32[w].cosca = float(walls[w].cosca);
                for (int k = 0; k < 9; ++k) {
                    w32[w].tib[k] = float(walls[w].tib[k]);
                    if (!g32.empty()) g32[w * 9 + k] = float(walls[w].ginv[k]);
                }
            }
            rc |= upload(c, c->d_walls32, w32.data(), w32.size() * sizeof(WallF32));
            if (c->d_ginv32) rc |= upload(c, c->d_ginv32, g32.data(), g32.size() * sizeof(float));
            if (c->d_wall_slot) {
                std::vector<int32_t> ws(c->nfluid);
                for (uint64_t s = 0; s < c->nfluid; ++s) ws[s] = c->links.wall[c->list[s]];
                rc |= upload(c, c->d_wall_slot, ws.data(), ws.size() * sizeof(int32_t));
            }
        }
        if (c->d_film) rc |= upload(c, c->d_film, c->links.film.data(), N * sizeof(uint64_t));
    }
    if (rc) {
        std::string e = c->err;
        lbf_destroy(c);
        return fail(nullptr, e);
    }
    for (void* p : {c->f[0], c->f[1]}) (void)hipMemset(p, 0, pop);
    (void)hipMemset(c->d_rho0, 0, N * 8);
    (void)hipMemset(c->d_u, 0, size_t(d.D) * N * 8);
    (void)hipMemset(c->d_uadj, 0, 2 * size_t(d.D) * N * 8);
    (void)hipMemset(c->d_uprev, 0, size_t(d.D) * N * 8);
    if (NP == 2) {
        (void)hipMemset(c->d_rho1, 0, N * 8);
        (void)hipMemset(c->d_chiadj, 0, 2 * N * 8);
        (void)hipMemset(c->d_chiprev, 0, N * 8);
        // only VD/GV sites are written; the others stay 0 (colorfluid.cpp:73-75)
        if (c->d_capgrad) (void)hipMemset(c->d_capgrad, 0, size_t(d.D) * N * 8);
        if (c->d_capfield) (void)hipMemset(c->d_capfield, 0, size_t(d.D) * N * 8);
    }
    DevStatus st{};
    for (auto& v : st.first) v = ~0ull;
    (void)hipMemcpy(c->d_status, &st, sizeof st, hipMemcpyHostToDevice);
    if (hipDeviceSynchronize() != hipSuccess) {
        lbf_destroy(c);
        return fail(nullptr, "HIP error during context set-up");
    }
    if (NP == 2) c->kernel_name = d.precision == LBF_FP32 ? "color_step_f32" : "color_step";
    else c->kernel_name = d.precision == LBF_FP32 ? "scsp_step_f32" : "scsp_step_f64";
    static const char* layout_names[3] = {"dense", "dense-fluid", "sparse"};
    c->kernel_name += std::string(" (") + layout_names[c->layout] + ")";
    *out = c;
    return 0;
}

// Element index of plane q, slot s in a population buffer (stream.h pop_index).
static size_t pop_elem(const lbf_ctx* c, size_t q, uint64_t s)
{
    if (c->blocked) return size_t(s >> 6) * (size_t(c->NP) * c->Q * 64) + q * 64 + size_t(s & 63);
    return q * c->P + s;
}

static void reset_status(lbf_ctx* c)
{
    DevStatus st{};
    for (auto& v : st.first) v = ~0ull;
    (void)hipMemcpyAsync(c->d_status, &st, sizeof st, hipMemcpyHostToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
}

static ScspParams scsp_params(lbf_ctx* c, uint32_t flags)
{
    ScspParams a{};
    a.g = c->g;
    a.nfluid = uint32_t(c->nfluid);
    a.flags = flags;
    a.dens0 = c->d.density[0];
    a.omega = 1.0 / c->d.tau[0];
    a.omh = 1.0 - 0.5 * a.omega;
    for (int q = 0; q < c->Q; ++q) a.idist[q] = (c->Q == 9 ? Lattice<9>::w[q] : Lattice<19>::w[q]) * c->d.density[0];
    a.inv_rho0 = float(1.0 / c->d.density[0]);
    a.rho0_32 = float(c->d.density[0]);
    // dense 3D lattices whose layers are whole 256-site chunks: columns of 32
    // chunks through all layers
    if (c->layout != 2 && c->D == 3 && c->g.nxy % 256 == 0 && (c->g.nxy / 256) % 32 == 0) {
        a.tile_chunks = 32;
        a.layer_chunks = c->g.nxy / 256;
    }
    a.omega32 = float(a.omega);
    a.omh32 = float(a.omh);
    const double su = 1.0 / c->d.tau[0], sq = 8.0 * (2.0 - su) / (8.0 - su);
    a.c1_nu = float(su);
    a.c2_nu = float(1.0 - 0.5 * su);
    a.c1_q = float(sq);
    a.c2_q = float(1.0 - 0.5 * sq);
    return a;
}

static ScspBuffers scsp_buffers(lbf_ctx* c, bool init)
{
    ScspBuffers b{};
    b.fin = c->f[c->cur];
    b.fout = init ? c->f[c->cur] : c->f[c->cur ^ 1];
    b.desc = c->layout == 2 ? c->d_desc_s : c->d_desc;
    b.list = c->d_list;
    b.rank = c->d_rank;
    b.zh = c->d_zh;
    b.lab = c->d_lab;
    b.lab32 = c->d_lab32;
    b.AT = c->d_AT;
    b.B = c->d_B;
    b.rho = c->d_rho0;
    b.u = c->d_u;
    const uint64_t step = init ? 0 : c->t + 1;
    b.uadj = c->d_uadj + (step & 1) * uint64_t(c->D) * c->N;
    b.status = c->d_status;
    return b;
}

static ColorArgs color_args(lbf_ctx* c, uint32_t flags)
{
    ColorArgs a{};
    a.g = c->g;
    a.desc = c->d_desc;
    a.lab = c->d_lab;
    a.wall = c->d_wall;
    a.walls = c->d_walls;
    a.film = c->d_film;
    a.tanh_wall = c->d.wall_gradient == LBF_WALL_TANH;
    a.tanhg = c->d_tanhg;
    a.csf = c->d.capillary == LBF_CAPILLARY_CSF;
    a.curv_gradient = c->d.curvature == LBF_CURVATURE_GRADIENT;
    a.capgrad = c->d_capgrad;
    a.capfield = c->d_capfield;
    a.sigma = c->d.sigma;
    a.renormalized = c->d.wall_gradient == LBF_WALL_RENORMALIZED;
    a.densN = c->d_densN;
    a.A = c->d_A;
    a.B = c->d_B;
    a.MInv = c->d_MInv;
    a.rho0 = c->d_rho0;
    a.rho1 = c->d_rho1;
    a.u = c->d_u;
    a.status = c->d_status;
    a.zh = c->d_zh;
    a.flags = flags;
    a.mrt = c->d.equil == LBF_MRT;
    a.uniform = c->d.tau[0] == c->d.tau[1];
    a.harmonic = c->d.harmonic;
    a.relax0 = c->d.tau[0] / 2.0;
    a.relax1 = c->d.tau[1] / 2.0;
    a.nu0 = (c->d.tau[0] - 0.5) / 3.0;
    a.nu1 = (c->d.tau[1] - 0.5) / 3.0;
    a.ak = 2.25 * c->d.sigma;
    a.beta = c->d.beta;
    // initPhi / getBi (colorfluid.cpp:1131-1168)
    double alpha[2];
    alpha[0] = c->d.alpha0;
    alpha[1] = 1 - c->d.density[0] / c->d.density[1] * (1 - alpha[0]);
    for (int q = 0; q < c->Q; ++q) {
        for (int ph = 0; ph < 2; ++ph) {
            double v;
            if (c->Q == 9) v = q == 0 ? alpha[ph] : (q <= 4 ? (1 - alpha[ph]) / 5 : (1 - alpha[ph]) / 20);
            else v = q == 0 ? alpha[ph] : (q <= 6 ? (1 - alpha[ph]) / 12 : (1 - alpha[ph]) / 24);
            (ph == 0 ? a.ai0 : a.ai1)[q] = v;
            (ph == 0 ? a.dist0 : a.dist1)[q] = v * c->d.density[ph];
        }
        if (c->Q == 9) {
            a.bi[q] = q == 0 ? -4.0 / 27 : (q <= 4 ? 2.0 / 27 : 5.0 / 108);
        } else {
            const double x = 2, c2 = 1.0;
            const double b0 = -(2.0 + 2.0 * x) / (3.0 * x + 12.0) * c2;
            const double b1 = x / (6.0 * x + 24.0) * c2;
            const double b7 = 1.0 / (6.0 * x + 24.0) * c2;
            a.bi[q] = q == 0 ? b0 : (q <= 6 ? b1 : b7);
        }
    }
    // FP32 form: rest-state moments M ai_k and the shift density
    for (int k = 0; k < c->Q; ++k) {
        double m0 = 0.0, m1 = 0.0;
        for (int q = 0; q < c->Q; ++q) {
            const double mk = c->Q == 9 ? Lattice<9>::M[k][q] : Lattice<19>::M[k][q];
            m0 += mk * a.ai0[q];
            m1 += mk * a.ai1[q];
        }
        a.mai0[k] = m0;
        a.mai1[k] = m1;
    }
    a.rho_ref = c->d.density[0];
    a.rest32 = 0.0;
    for (int q = 0; q < c->Q; ++q) {
        a.rest32 += double(float(a.dist0[q]));
        a.dai[q] = a.ai1[q] - a.ai0[q];
        a.rest_err[q] = a.dist0[q] - double(float(a.dist0[q]));
    }
    for (int k = 0; k < c->Q; ++k) {
        double m = 0.0, e = 0.0;
        for (int q = 0; q < c->Q; ++q) {
            const double mk = c->Q == 9 ? Lattice<9>::M[k][q] : Lattice<19>::M[k][q];
            m += mk * a.dai[q];
            e += mk * a.rest_err[q];
        }
        a.mdai[k] = m;
        a.mrest_err[k] = e;
    }
    for (int q = 0; q < 19; ++q) {
        a.ai0f[q] = float(a.ai0[q]);
        a.ai1f[q] = float(a.ai1[q]);
        a.bif[q] = float(a.bi[q]);
        a.dist0f[q] = float(a.dist0[q]);
        a.dist0r[q] = double(float(a.dist0[q]));
        a.daif[q] = float(a.dai[q]);
        a.rest_errf[q] = float(a.rest_err[q]);
        a.mai0f[q] = float(a.mai0[q]);
        a.mdaif[q] = float(a.mdai[q]);
        a.mrest_errf[q] = float(a.mrest_err[q]);
    }
    a.relax0f = float(a.relax0);
    a.relax1f = float(a.relax1);
    a.nu0f = float(a.nu0);
    a.nu1f = float(a.nu1);
    a.betaf = float(a.beta);
    a.akf = float(a.ak);
    a.rho_reff = float(a.rho_ref);
    return a;
}

static int run_step(lbf_ctx* c, bool init, uint32_t flags)
{
    const bool stream_solids = c->d.flags & LBF_STREAM_SOLIDS;
    const bool mrt = c->d.equil == LBF_MRT;
    hipError_t e;
    if (c->NP == 1) {
        ScspParams p = scsp_params(c, flags);
        p.t = init ? 0 : c->t + 1;
        e = launch_scsp(c->Q, mrt, init, c->d.precision == LBF_FP32, c->layout, p, scsp_buffers(c, init),
                        c->stream);
    } else {
        ColorArgs a = color_args(c, flags);
        a.t = init ? 0 : c->t + 1;
        a.uadj = c->d_uadj + (a.t & 1) * uint64_t(c->D) * c->N;
        a.chiadj = c->d_chiadj + (a.t & 1) * c->N;
        a.uadj_prev = c->d_uadj + ((a.t + 1) & 1) * uint64_t(c->D) * c->N;
        a.chiadj_prev = c->d_chiadj + ((a.t + 1) & 1) * c->N;
        if (c->elem == 4) {
            a.fin32 = static_cast<const float*>(c->f[c->cur]);
            a.fout32 = static_cast<float*>(init ? c->f[c->cur] : c->f[c->cur ^ 1]);
            a.densN32 = reinterpret_cast<float*>(c->d_densN);
            a.dens_slot = c->d_dens_slot;
            a.walls32 = c->d_walls32;
            a.ginv32 = c->d_ginv32;
            a.wall_slot = c->d_wall_slot;
            a.list = c->d_list;
            a.rank = c->d_rank;
            a.P = uint32_t(c->P);
            if (c->layout == 2) a.desc = c->d_desc_s;
            e = launch_color32(c->Q, mrt, init, c->layout == 2, a, c->stream);
        } else {
            a.fin = static_cast<const double*>(c->f[c->cur]);
            a.fout = static_cast<double*>(init ? c->f[c->cur] : c->f[c->cur ^ 1]);
            e = launch_color(c->Q, mrt, init, stream_solids, a, c->stream);
        }
    }
    if (e != hipSuccess) return fail(c, std::string("HIP error: kernel launch: ") + hipGetErrorString(e));
    if (!init) {
        c->cur ^= 1;
        c->t++;
    }
    return 0;
}

int lbf_init(lbf_ctx* c)
{
    if (!c) return -1;
    reset_status(c);
    c->t = 0;
    c->has_prev = false;
    const size_t pop = c->pop_elems * c->elem;
    HIP_CHECK(c, hipMemsetAsync(c->f[c->cur], 0, pop, c->stream));
    if (run_step(c, true, 3u)) return -1;   // fields + adjacent history of a report at t = 1
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lbf_step(lbf_ctx* c, uint32_t flags)
{
    if (!c) return -1;
    if (run_step(c, false, flags & 3u)) return -1;
    return 0;
}

int lbf_advance(lbf_ctx* c, uint64_t n)
{
    if (!c) return -1;
    for (uint64_t k = 0; k < n; ++k)
        if (run_step(c, false, 0u)) return -1;
    return 0;
}

uint64_t lbf_time(const lbf_ctx* c) { return c ? c->t : 0; }

int lbf_synchronize(lbf_ctx* c)
{
    if (!c) return -1;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lbf_status(lbf_ctx* c, uint32_t* bits, uint64_t* first)
{
    if (!c) return -1;
    DevStatus st{};
    HIP_CHECK(c, hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (bits) *bits = st.bits;
    if (first)
        for (int k = 0; k < 6; ++k) first[k] = st.first[k];
    return 0;
}

int lbf_fields(lbf_ctx* c, double* rho0, double* rho1, double* u)
{
    if (!c) return -1;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (rho0) HIP_CHECK(c, hipMemcpy(rho0, c->d_rho0, c->N * 8, hipMemcpyDeviceToHost));
    if (rho1 && c->NP == 2) HIP_CHECK(c, hipMemcpy(rho1, c->d_rho1, c->N * 8, hipMemcpyDeviceToHost));
    if (u) HIP_CHECK(c, hipMemcpy(u, c->d_u, size_t(c->D) * c->N * 8, hipMemcpyDeviceToHost));
    return 0;
}

int lbf_get_populations(lbf_ctx* c, double* out)
{
    if (!c || !out) return -1;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const size_t planes = size_t(c->NP) * c->Q;
    const size_t n = c->pop_elems;
    if (c->elem == 8 && c->layout == 0) {
        HIP_CHECK(c, hipMemcpy(out, c->f[c->cur], n * 8, hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<double> h64(c->elem == 8 ? n : 0);
    std::vector<float> h32(c->elem == 4 ? n : 0);
    if (c->elem == 8) HIP_CHECK(c, hipMemcpy(h64.data(), c->f[c->cur], n * 8, hipMemcpyDeviceToHost));
    else HIP_CHECK(c, hipMemcpy(h32.data(), c->f[c->cur], n * 4, hipMemcpyDeviceToHost));
    std::memset(out, 0, planes * c->N * sizeof(double));
    const std::vector<uint32_t>& desc = c->links.desc;
    if (c->NP == 2 && c->elem == 4) {
        // device planes: h = fT - float(s) (Q planes), then f0 (Q planes);
        // checkpoint: f0, f1 (the kernels add the shift back as float(s))
        const ColorArgs a = color_args(c, 0u);
        const int Q = c->Q;
        for (int q = 0; q < Q; ++q)
            for (uint64_t s = 0; s < c->P; ++s) {
                const uint64_t i = c->layout == 2 ? c->list[s] : s;
                if (!(c->hlab[desc[i] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) continue;
                const double h = h32[pop_elem(c, q, s)], f0 = h32[pop_elem(c, Q + q, s)];
                out[size_t(q) * c->N + i] = f0;
                out[size_t(Q + q) * c->N + i] = h + double(float(a.dist0[q])) - f0;
            }
        return 0;
    }
    for (size_t plane = 0; plane < planes; ++plane)
        for (uint64_t s = 0; s < c->P; ++s) {
            const uint64_t i = c->layout == 2 ? c->list[s] : s;
            // solid sites that are never streamed keep the value 0
            if (c->layout == 1 && !(c->hlab[desc[i] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) continue;
            const size_t k = pop_elem(c, plane, s);
            out[plane * c->N + i] = c->elem == 8 ? h64[k] : double(h32[k]) + c->shift[plane];
        }
    return 0;
}

int lbf_set_populations(lbf_ctx* c, const double* in, uint64_t t)
{
    if (!c || !in) return -1;
    // Steps queued on the context's (non-blocking) stream may still be writing
    // the buffers replaced here: wait for them, then upload on the same stream.
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const size_t planes = size_t(c->NP) * c->Q;
    const size_t n = c->pop_elems;
    if (c->elem == 8 && c->layout == 0) {
        HIP_CHECK(c, hipMemcpyAsync(c->f[c->cur], in, n * 8, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<double> h64(c->elem == 8 ? n : 0);
        std::vector<float> h32(c->elem == 4 ? n : 0);
        if (c->NP == 2 && c->elem == 4) {
            const ColorArgs a = color_args(c, 0u);
            const int Q = c->Q;
            const std::vector<uint32_t>& desc = c->links.desc;
            for (int q = 0; q < Q; ++q)
                for (uint64_t s = 0; s < c->P; ++s) {
                    const uint64_t i = c->layout == 2 ? c->list[s] : s;
                    if (!(c->hlab[desc[i] >> DESC_LABEL_SHIFT].kind & KIND_FLUID)) continue;
                    const double f0 = in[size_t(q) * c->N + i], f1 = in[size_t(Q + q) * c->N + i];
                    h32[pop_elem(c, q, s)] = float(f0 + f1 - double(float(a.dist0[q])));
                    h32[pop_elem(c, Q + q, s)] = float(f0);
                }
        } else
        for (size_t plane = 0; plane < planes; ++plane)
            for (uint64_t s = 0; s < c->P; ++s) {
                const double v = in[plane * c->N + (c->layout == 2 ? c->list[s] : s)];
                const size_t k = pop_elem(c, plane, s);
                if (c->elem == 8) h64[k] = v;
                else h32[k] = float(v - c->shift[plane]);
            }
        if (c->elem == 8)
            HIP_CHECK(c, hipMemcpyAsync(c->f[c->cur], h64.data(), n * 8, hipMemcpyHostToDevice, c->stream));
        else
            HIP_CHECK(c, hipMemcpyAsync(c->f[c->cur], h32.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));   // the staging vectors end here
    }
    // A restarted run holds no macro history (Simulator::validateRestart resets it,
    // simulator.cpp:382-383; the macro arrays are fresh zeros until the first step).
    HIP_CHECK(c, hipMemsetAsync(c->d_rho0, 0, c->N * 8, c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->d_u, 0, size_t(c->D) * c->N * 8, c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->d_uadj, 0, 2 * size_t(c->D) * c->N * 8, c->stream));
    if (c->NP == 2) {
        HIP_CHECK(c, hipMemsetAsync(c->d_rho1, 0, c->N * 8, c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->d_chiadj, 0, 2 * c->N * 8, c->stream));
    }
    reset_status(c);   // synchronises the stream
    c->t = t;
    c->has_prev = false;
    return 0;
}

}  // extern "C"

// One report's reductions with the block partials in `partial` and the sums
// in `sums` (device memory of at least report_blocks(N) * NREP and NREP
// doubles).
static int report_into(lbf_ctx* c, const lbf_report_req* req, lbf_report_out* out, double* partial, double* sums)
{
    if (req->axis < 0 || req->axis >= c->D) return fail(c, "lbf_report: axis outside the lattice");
    ReportArgs a{};
    a.g = c->g;
    a.desc = c->d_desc;
    a.lab = c->d_lab;
    a.rho0 = c->d_rho0;
    a.rho1 = c->NP == 2 ? c->d_rho1 : c->d_rho0;
    a.u = c->d_u;
    a.uprev = c->d_uprev;
    // the adjacent history of a report at step t was written by step t - 1
    const uint64_t adjacent_buffer = (c->t + 1) & 1;
    a.uadj = c->d_uadj + adjacent_buffer * uint64_t(c->D) * c->N;
    a.chiprev = c->d_chiprev;
    a.chiadj = c->NP == 2 ? c->d_chiadj + adjacent_buffer * c->N : nullptr;
    a.partial = partial;
    a.D = c->D;
    a.axis = req->axis;
    a.color = c->NP == 2;
    a.conv = req->convergence ? 1 : 0;
    a.adjacent = req->adjacent ? 1 : 0;
    a.first = uint32_t(req->first);
    a.last = uint32_t(req->last);
    HIP_CHECK(c, launch_report(a, sums, c->stream));
    double r[NREP];
    HIP_CHECK(c, hipMemcpyAsync(r, sums, sizeof r, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    std::memset(out, 0, sizeof *out);
    out->vel_range = r[0];
    out->rho_range[0] = r[1];
    out->rho_range[1] = r[2];
    out->rho_slice[0][0] = r[3];
    out->rho_slice[0][1] = r[4];
    out->rho_slice[1][0] = r[5];
    out->rho_slice[1][1] = r[6];
    out->count_slice[0] = r[7];
    out->count_slice[1] = r[8];
    for (int k = 0; k < 4; ++k) out->conv[k] = r[9 + k];
    out->phase_change[0] = r[13];
    out->phase_change[1] = r[14];
    out->nonfinite = r[15] != 0.0;
    out->has_previous = req->convergence ? (c->has_prev ? 1 : 0) : 0;
    if (req->convergence) c->has_prev = true;
    return 0;
}

extern "C" {

int lbf_report(lbf_ctx* c, const lbf_report_req* req, lbf_report_out* out)
{
    if (!c || !req || !out) return -1;
    return report_into(c, req, out, c->d_partial, c->d_out);
}

int lbf_report_static(lbf_ctx* c, const lbf_report_req* req, lbf_report_out* out)
{
    if (!c || !req || !out) return -1;
    int dev = 0;
    HIP_CHECK(c, hipGetDevice(&dev));
    if (dev != c->device) return fail(c, "lbf_report_static: the context's GPU is not the current device");
    double* scratch = nullptr;
    uint64_t capacity = 0;
    HIP_CHECK(c, report_scratch(&scratch, &capacity));
    const uint64_t blocks = report_blocks(c->N);
    if (blocks * NREP + NREP > capacity)
        return fail(c, "lbf_report_static: " + std::to_string(blocks) + " partial blocks exceed the static scratch of " +
                           std::to_string(capacity) + " doubles");
    return report_into(c, req, out, scratch, scratch + blocks * NREP);
}

uint64_t lbf_links(lbf_ctx* c, uint32_t* linkmask, int32_t* wall, double* normal, double* sinca, double* cosca)
{
    if (!c) return 0;
    const Links& L = c->links;
    if (linkmask)
        for (uint64_t i = 0; i < c->N; ++i) linkmask[i] = L.desc[i] & DESC_LINK_MASK;
    if (wall) std::memcpy(wall, L.wall.data(), c->N * 4);
    if (normal) std::memcpy(normal, L.normal.data(), L.normal.size() * 8);
    if (sinca) std::memcpy(sinca, L.sinca.data(), L.sinca.size() * 8);
    if (cosca) std::memcpy(cosca, L.cosca.data(), L.cosca.size() * 8);
    return L.nwall;
}

uint64_t lbf_sites(const lbf_ctx* c) { return c ? c->N : 0; }
uint64_t lbf_device_bytes(const lbf_ctx* c) { return c ? c->dev_bytes : 0; }
const char* lbf_step_kernel(const lbf_ctx* c) { return c ? c->kernel_name.c_str() : ""; }

double lbf_step_bytes(const lbf_ctx* c)
{
    if (!c) return 0.0;
    const double pops = 2.0 * c->NP * c->Q * double(c->elem);   // read + write per streamed site
    // the colour model also writes and reads its indicator once per site
    const double indicator = c->NP == 2 ? 2.0 * double(c->elem) : 0.0;
    if (c->NP == 2 && c->layout == 0) return double(c->N) * (pops + indicator + 4.0);
    if (c->NP == 2 && c->layout == 1) return double(c->nfluid) * (pops + indicator) + double(c->N) * 4.0;
    if (c->NP == 2) return double(c->nfluid) * (pops + indicator + 8.0) + double((c->N + 31) / 32) * 8.0;
    if (c->layout == 0) return double(c->N) * (pops + 4.0);       // + descriptor
    if (c->layout == 1) return double(c->nfluid) * pops + double(c->N) * 4.0;
    // sparse: per fluid site its populations, its list entry and its
    // descriptor, plus one pass over the rank index (8 bytes per 32 sites)
    return double(c->nfluid) * (pops + 8.0) + double((c->N + 31) / 32) * 8.0;
}

int lbf_timed_advance(lbf_ctx* c, uint64_t n, double* ms)
{
    if (!c || !ms) return -1;
    hipEvent_t e0, e1;
    HIP_CHECK(c, hipEventCreate(&e0));
    HIP_CHECK(c, hipEventCreate(&e1));
    HIP_CHECK(c, hipEventRecord(e0, c->stream));
    for (uint64_t k = 0; k < n; ++k)
        if (run_step(c, false, 0u)) return -1;
    HIP_CHECK(c, hipEventRecord(e1, c->stream));
    HIP_CHECK(c, hipEventSynchronize(e1));
    float f = 0.0f;
    HIP_CHECK(c, hipEventElapsedTime(&f, e0, e1));
    *ms = double(f);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

double lbf_copy_bandwidth(uint64_t bytes, int iters)
{
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double result = -1.0;
    bytes &= ~uint64_t(15);
    if (bytes == 0 || iters <= 0) return -1.0;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        g_create_error = "lbf_copy_bandwidth: HIP set-up failed";
    } else {
        (void)hipMemsetAsync(a, 0, bytes, s);
        (void)launch_copy(a, b, bytes, s);   // warm-up
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < iters; ++k) (void)launch_copy(k % 2 ? b : a, k % 2 ? a : b, bytes, s);
        (void)hipEventRecord(e1, s);
        float ms = 0.0f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0)
            result = 2.0 * double(bytes) * iters / (double(ms) * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return result;
}

}  // extern "C"
