// batch.cpp -- C ABI of the batched-crop engine (include/lbflow.h, lbf_batch_*).
//
// Hundreds of independent single-phase crops share one slot space and one
// persistent step kernel (kernels_batch.hip).  A run captures `interval`
// steps plus the per-crop convergence reduction in one hipGraph, replays it,
// applies the reference's stop rule per crop on the host (velocity residual
// and adjacent residual below the tolerance for three consecutive reports,
// Simulator::updateConvergence, simulator.cpp:248-325) and drops converged
// crops from the device-resident active-tile list.  The autotuner times the
// step's launch shape (threads per workgroup, slots per work unit) and
// persistent grid size (workgroups per CU) on the actual batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbflow.h"
#include "batch_types.h"
#include "setup.h"

namespace lbf {
hipError_t launch_batch_step(int Q, bool mrt, bool init, uint32_t grid, BatchShape sh, const BatchParams& p,
                             const BatchBuffers& b, hipStream_t s);
hipError_t launch_batch_fields(int Q, uint64_t S, float rho0, uint32_t tile_begin, uint32_t ntiles,
                               const uint32_t* site, const float* f, const float* ucur, float* fields, hipStream_t s);
hipError_t launch_batch_warm(int Q, bool mrt, const BatchParams& p, uint32_t tile_begin, uint32_t ntiles,
                             const uint32_t* site, const uint8_t* lab8, const uint8_t* flow8, const LabelF* lab,
                             const float* fields, float* f, float* upres, hipStream_t s);
hipError_t launch_batch_average(float* a, const float* b, uint64_t n, hipStream_t s);
hipError_t launch_batch_report(uint32_t grid, uint64_t S, int D, int axis, const BatchBuffers& b, uint32_t ncrops,
                               double* partial, double* out, hipStream_t s);
}  // namespace lbf

using namespace lbf;

struct lbf_batch {
    lbf_batch_desc d{};
    int D = 0, Q = 0;
    uint32_t ncrops = 0;
    uint64_t S = 0, T = 0, sites = 0, fluid = 0;
    std::vector<CropDev> crops;
    std::vector<uint64_t> crop_sites;
    std::vector<double> crop_force;      // FP64 force of the crop's first GV label, per axis
    std::vector<uint32_t> tile_crop;
    hipStream_t stream = nullptr;
    void* f[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* d_tiles = nullptr;
    uint32_t* d_ntiles = nullptr;
    uint32_t* d_tile_crop = nullptr;
    CropDev* d_crops = nullptr;
    uint32_t* d_site = nullptr;
    uint8_t* d_lab8 = nullptr;
    uint32_t* d_desc = nullptr;      // label and bounce-back links per slot (kernels_batch.hip)
    uint8_t* d_flow8 = nullptr;      // 1 on fluid sites of clusters that carry the steady flow
    uint2* d_rank = nullptr;
    LabelF* d_lab = nullptr;
    float *d_ucur = nullptr, *d_uprev = nullptr, *d_uadj = nullptr, *d_upres = nullptr;
    DevStatus* d_status = nullptr;
    double *d_partial = nullptr, *d_out = nullptr;
    double* h_out = nullptr;         // pinned host copy of d_out, read after each graph replay
    std::vector<void*> allocations;
    uint32_t grid = 0, cus = 0;
    uint32_t report_grid = 0;        // persistent grid of the report reduction (256-thread tiles)
    BatchShape shape;                // launch shape of the step (autotuned)
    std::string err;
    uint64_t t = 0;
};

namespace {

thread_local std::string g_batch_error;

// A captured graph and its executable, released on every exit path.
struct CapturedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CapturedGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int bfail(lbf_batch* b, const std::string& m)
{
    if (b) b->err = m;
    else g_batch_error = m;
    return -1;
}

#define BHIP(b, expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return bfail(b, std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T> int balloc(lbf_batch* b, T** p, size_t bytes)
{
    void* q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess) return bfail(b, "HIP error: hipMalloc failed");
    b->allocations.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

RelaxF32 relax_of(double tau, double rho0)
{
    const double su = 1.0 / tau, sq = 8.0 * (2.0 - su) / (8.0 - su);
    return RelaxF32{float(rho0), float(su), float(1.0 - 0.5 * su), float(su), float(1.0 - 0.5 * su), float(sq),
                    float(1.0 - 0.5 * sq)};
}

BatchBuffers buffers(lbf_batch* b, bool swap)
{
    BatchBuffers x{};
    x.fin = b->f[b->cur];
    x.fout = swap ? b->f[b->cur ^ 1] : b->f[b->cur];
    x.tiles = b->d_tiles;
    x.ntiles = b->d_ntiles;
    x.tile_crop = b->d_tile_crop;
    x.crops = b->d_crops;
    x.site = b->d_site;
    x.lab8 = b->d_lab8;
    x.desc = b->d_desc;
    x.rank = b->d_rank;
    x.lab = b->d_lab;
    x.ucur = b->d_ucur;
    x.uprev = b->d_uprev;
    x.uadj = b->d_uadj;
    x.upres = b->d_upres;
    x.status = b->d_status;
    return x;
}

BatchParams params(lbf_batch* b, uint32_t flags)
{
    BatchParams p{};
    p.S = b->S;
    p.flags = flags;
    p.t = b->t;
    p.inv_rho0 = float(1.0 / b->d.density);
    p.relax = relax_of(b->d.tau, b->d.density);
    return p;
}

int set_active(lbf_batch* b, const std::vector<char>& active)
{
    std::vector<uint32_t> tiles;
    for (uint64_t t = 0; t < b->T; ++t)
        if (active[b->tile_crop[t]]) tiles.push_back(uint32_t(t));
    const uint32_t n = uint32_t(tiles.size());
    if (n) BHIP(b, hipMemcpy(b->d_tiles, tiles.data(), n * 4, hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_ntiles, &n, 4, hipMemcpyHostToDevice));
    return 0;
}

// One step on the stream: persistent grid over the active tiles.
int step(lbf_batch* b, uint32_t flags, bool init)
{
    const BatchParams p = params(b, flags);
    hipError_t e = launch_batch_step(b->Q, b->d.equil == LBF_MRT, init, init ? b->report_grid : b->grid, b->shape, p,
                                     buffers(b, !init), b->stream);
    if (e != hipSuccess) return bfail(b, std::string("HIP error: ") + hipGetErrorString(e));
    if (!init) {
        b->cur ^= 1;
        b->t++;
    }
    return 0;
}

// Fluid sites that carry the steady flow: those whose cluster (fluid sites
// joined by lattice links, periodic) wraps around the periodic box along an
// axis the body force has a component on.  In every other cluster the steady
// velocity is exactly 0 (a uniform force there is balanced by pressure), and
// a warm start must not put flow into them: a pocket whose links all bounce
// back would keep a prescribed momentum forever, flipping its sign each step.
// Union-find with each site's winding (box crossings) relative to its root;
// a link that closes a cycle with a non-zero net winding marks a wrap.
std::vector<uint8_t> flowing_sites(const uint8_t* geom, const uint64_t* sz, int D, int Q, const LabelF* lab,
                                   const double* force)
{
    const uint32_t nx = uint32_t(sz[D - 1]), ny = uint32_t(sz[D - 2]), nz = D == 3 ? uint32_t(sz[0]) : 1u;
    const uint64_t N = uint64_t(nx) * ny * nz;
    std::vector<int32_t> parent(N);
    std::vector<int16_t> off(N * 3, 0);        // winding of a site relative to its parent (z, y, x)
    std::vector<uint8_t> wrap(N, 0);           // at roots: bit a = the cluster wraps along axis a
    for (uint64_t i = 0; i < N; ++i) parent[i] = int32_t(i);
    auto fluid = [&](uint64_t i) { return (lab[geom[i]].kind & KIND_FLUID) != 0; };
    // root of i and the winding of i relative to it (path compression keeps windings)
    auto find = [&](uint64_t i, int* w) {
        w[0] = w[1] = w[2] = 0;
        uint64_t r = i;
        while (uint64_t(parent[r]) != r) {
            for (int a = 0; a < 3; ++a) w[a] += off[r * 3 + a];
            r = uint64_t(parent[r]);
        }
        // compress: point every node on the path at r with its total winding
        int acc[3] = {w[0], w[1], w[2]};
        uint64_t c = i;
        while (uint64_t(parent[c]) != r && c != r) {
            const uint64_t nxt = uint64_t(parent[c]);
            int here[3];
            for (int a = 0; a < 3; ++a) here[a] = off[c * 3 + a];
            for (int a = 0; a < 3; ++a) off[c * 3 + a] = int16_t(acc[a]);
            parent[c] = int32_t(r);
            for (int a = 0; a < 3; ++a) acc[a] -= here[a];
            c = nxt;
        }
        return r;
    };
    const int H = Q == 9 ? Lattice<9>::NHALF : Lattice<19>::NHALF;
    const int* half = Q == 9 ? Lattice<9>::half : Lattice<19>::half;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const uint64_t i = (uint64_t(z) * ny + y) * nx + x;
                if (!fluid(i)) continue;
                for (int j = 0; j < H; ++j) {
                    const int q = half[j];
                    int e[3] = {0, 0, 0};   // (z, y, x)
                    for (int a = 0; a < D; ++a)
                        e[3 - D + a] = Q == 9 ? Lattice<9>::e[q][a] : Lattice<19>::e[q][a];
                    const int64_t c[3] = {int64_t(z) + e[0], int64_t(y) + e[1], int64_t(x) + e[2]};
                    const int64_t n[3] = {nz, ny, nx};
                    int cross[3];
                    int64_t t[3];
                    for (int a = 0; a < 3; ++a) {
                        cross[a] = c[a] < 0 ? -1 : (c[a] >= n[a] ? 1 : 0);
                        t[a] = c[a] - cross[a] * n[a];
                    }
                    const uint64_t k = (uint64_t(t[0]) * ny + uint64_t(t[1])) * nx + uint64_t(t[2]);
                    if (!fluid(k)) continue;
                    int wi[3], wk[3];
                    const uint64_t ri = find(i, wi), rk = find(k, wk);
                    // image of k reached from i lies `cross` boxes away
                    int d[3];
                    for (int a = 0; a < 3; ++a) d[a] = wi[a] + cross[a] - wk[a];
                    if (ri == rk) {
                        for (int a = 0; a < 3; ++a)
                            if (d[a] != 0) wrap[ri] |= uint8_t(1u << a);
                    } else {
                        parent[rk] = int32_t(ri);
                        for (int a = 0; a < 3; ++a) off[rk * 3 + a] = int16_t(d[a]);
                        wrap[ri] |= wrap[rk];
                    }
                }
            }
    uint8_t need = 0;
    for (int a = 0; a < D; ++a)
        if (force[a] != 0.0) need |= uint8_t(1u << (3 - D + a));
    std::vector<uint8_t> flow(N, 0);
    for (uint64_t i = 0; i < N; ++i) {
        if (!fluid(i)) continue;
        int w[3];
        flow[i] = (wrap[find(i, w)] & need) ? 1 : 0;
    }
    return flow;
}

}  // namespace

extern "C" {

const char* lbf_batch_error(const lbf_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

void lbf_batch_destroy(lbf_batch* b)
{
    if (!b) return;
    for (void* p : b->allocations) (void)hipFree(p);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int lbf_batch_create(const lbf_batch_desc* desc, const lbf_label* labels, const lbf_crop* crops, lbf_batch** out)
{
    g_batch_error.clear();
    if (!desc || !labels || !crops || !out) return bfail(nullptr, "lbf_batch_create: null argument");
    *out = nullptr;
    const lbf_batch_desc& d = *desc;
    if (!((d.D == 2 && d.Q == 9) || (d.D == 3 && d.Q == 19)))
        return bfail(nullptr, "Only D2Q9 and D3Q19 are validated for simulation.");
    if (!std::isfinite(d.tau) || d.tau <= 0.5 || !(d.density > 0.0))
        return bfail(nullptr, "Densities must be positive; relaxation times must exceed 0.5.");
    if (d.ncrops == 0) return bfail(nullptr, "lbf_batch_create: no crops");
    lbf_batch* b = new lbf_batch();
    b->d = d;
    b->D = d.D;
    b->Q = d.Q;
    b->ncrops = d.ncrops;
    // label table: fluid = VD/GV, solids = BB/SD; nothing else is batched
    std::vector<LabelF> lab(256);
    std::vector<double> force64(size_t(256) * 3, 0.0);
    for (int l = 0; l < 256; ++l) {
        std::memset(&lab[l], 0, sizeof(LabelF));
        const int t = labels[l].type;
        if (t == LBF_VD || t == LBF_GV) lab[l].kind = KIND_FLUID;
        if (t == LBF_GV) {
            lab[l].kind |= KIND_GV;
            for (int k = 0; k < d.D; ++k) {
                const int e = d.Q == 9 ? Lattice<9>::e[labels[l].dir][k] : Lattice<19>::e[labels[l].dir][k];
                const bool diag = d.Q == 9 ? Lattice<9>::diagonal(labels[l].dir) : Lattice<19>::diagonal(labels[l].dir);
                const double unit = e == 0 ? 0.0 : (diag ? (e > 0 ? std::sqrt(2) / 2 : -std::sqrt(2) / 2) : double(e));
                const double F = unit * labels[l].value;
                force64[size_t(l) * 3 + k] = F;
                lab[l].force[k] = float(F);
                lab[l].half_rho[k] = float(F * 0.5 / d.density);
            }
        }
    }
    // slot space
    std::vector<uint32_t> site;
    std::vector<uint8_t> lab8, flow8;
    std::vector<uint32_t> descs;
    std::vector<uint2> rank;
    b->crops.resize(d.ncrops);
    b->crop_sites.resize(d.ncrops);
    b->crop_force.assign(size_t(d.ncrops) * 3, 0.0);
    for (uint32_t c = 0; c < d.ncrops; ++c) {
        const lbf_crop& cr = crops[c];
        uint64_t N = 1;
        for (int k = 0; k < d.D; ++k) N *= cr.sz[k];
        const uint32_t nx = uint32_t(cr.sz[d.D - 1]), ny = uint32_t(cr.sz[d.D - 2]);
        const uint32_t nz = d.D == 3 ? uint32_t(cr.sz[0]) : 1u;
        CropDev cd{};
        cd.g.nx = nx;
        cd.g.ny = ny;
        cd.g.nz = nz;
        cd.g.nxy = nx * ny;
        cd.g.N = N;
        const FastDiv fx = make_fastdiv(nx), fxy = make_fastdiv(nx * ny);
        cd.g.div_nx = DevFastDiv{fx.d, fx.m, fx.s};
        cd.g.div_nxy = DevFastDiv{fxy.d, fxy.m, fxy.s};
        cd.slot_base = uint32_t(site.size());
        cd.rank_base = uint32_t(rank.size());
        rank.resize(rank.size() + (N + 31) / 32, uint2{0u, 0u});
        cd.tile_begin = uint32_t(site.size() / TILE);
        uint32_t nf = 0;
        bool have_force = false;
        for (uint64_t i = 0; i < N; ++i) {
            const uint8_t v = cr.geometry[i];
            const int t = labels[v].type;
            if (!t) {
                lbf_batch_destroy(b);
                return bfail(nullptr, "map::at");
            }
            if (lab[v].kind & KIND_FLUID) {
                uint2& r = rank[cd.rank_base + i / 32];
                r.y |= 1u << (i % 32);
                ++nf;
                site.push_back(uint32_t(i));
                lab8.push_back(v);
                flow8.push_back(0);
                if (!have_force && (lab[v].kind & KIND_GV)) {
                    for (int k = 0; k < d.D; ++k) b->crop_force[size_t(c) * 3 + k] = force64[size_t(v) * 3 + k];
                    have_force = true;
                }
            } else if (t != LBF_BB && t != LBF_SD) {
                lbf_batch_destroy(b);
                return bfail(nullptr, "lbf_batch_create: batched crops support VD/GV fluid and BB/SD solids only");
            }
        }
        uint32_t before = 0;
        for (size_t w = cd.rank_base; w < rank.size(); ++w) {
            rank[w].x = before;
            before += uint32_t(__builtin_popcount(rank[w].y));
        }
        // per fluid slot: its label and bit op(q) where the upstream site
        // x - e_q (periodic in the crop) is solid
        for (uint64_t k = cd.slot_base; k < site.size(); ++k) {
            const uint32_t i = site[k];
            const int64_t x = i % nx, y = (i / nx) % ny, z = i / (uint64_t(nx) * ny);
            uint32_t w = uint32_t(lab8[k]) << DESC_LABEL_SHIFT;
            for (int q = 1; q < d.Q; ++q) {
                int64_t ex, ey, ez;
                if (d.Q == 9) {
                    ez = 0; ey = Lattice<9>::e[q][0]; ex = Lattice<9>::e[q][1];
                } else {
                    ez = Lattice<19>::e[q][0]; ey = Lattice<19>::e[q][1]; ex = Lattice<19>::e[q][2];
                }
                const uint64_t ux = uint64_t((x - ex + nx) % nx), uy = uint64_t((y - ey + ny) % ny),
                               uz = uint64_t((z - ez + nz) % nz);
                const uint64_t up = (uz * ny + uy) * nx + ux;
                if (!((rank[cd.rank_base + up / 32].y >> (up % 32)) & 1u))
                    w |= 1u << (d.Q == 9 ? Lattice<9>::op[q] : Lattice<19>::op[q]);
            }
            descs.push_back(w);
        }
        while (site.size() % TILE) {
            site.push_back(NO_SITE);
            lab8.push_back(0);
            flow8.push_back(0);
            descs.push_back(0);
        }
        cd.nfluid = nf;
        cd.tile_end = uint32_t(site.size() / TILE);
        b->crops[c] = cd;
        b->crop_sites[c] = N;
        b->sites += N;
        b->fluid += nf;
        for (uint32_t t = cd.tile_begin; t < cd.tile_end; ++t) b->tile_crop.push_back(c);
    }
    // flowing clusters of every crop, crops in parallel (disjoint slot ranges)
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t)
            pool.emplace_back([&, t]() {
                for (uint32_t c = t; c < d.ncrops; c += nt) {
                    const std::vector<uint8_t> flow = flowing_sites(crops[c].geometry, crops[c].sz, d.D, d.Q,
                                                                    lab.data(), &b->crop_force[size_t(c) * 3]);
                    const CropDev& cd = b->crops[c];
                    for (uint64_t k = cd.slot_base; k < uint64_t(cd.slot_base) + cd.nfluid; ++k)
                        flow8[k] = flow[site[k]];
                }
            });
        for (auto& th : pool) th.join();
    }
    if (site.size() >= (uint64_t(1) << 32) || rank.size() >= (uint64_t(1) << 31)) {
        lbf_batch_destroy(b);
        return bfail(nullptr, "lbf_batch_create: batch too large for 32-bit slot indices");
    }
    b->S = site.size();
    b->T = b->S / TILE;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        lbf_batch_destroy(b);
        return bfail(nullptr, "HIP error: cannot create a stream (no GPU?)");
    }
    hipDeviceProp_t prop{};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    b->cus = uint32_t(prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256);
    b->grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(b->T, uint64_t(b->cus) * 8)));
    b->report_grid = b->grid;
    // populations blocked per 64 slots, [slot / 64][q][slot % 64]; S is whole tiles
    const size_t pop = size_t(b->Q) * b->S * 4;
    int rc = 0;
    rc |= balloc(b, &b->f[0], pop);
    rc |= balloc(b, &b->f[1], pop);
    rc |= balloc(b, &b->d_tiles, b->T * 4);
    rc |= balloc(b, &b->d_ntiles, 4);
    rc |= balloc(b, &b->d_tile_crop, b->T * 4);
    rc |= balloc(b, &b->d_crops, b->crops.size() * sizeof(CropDev));
    rc |= balloc(b, &b->d_site, b->S * 4);
    rc |= balloc(b, &b->d_lab8, b->S);
    rc |= balloc(b, &b->d_desc, b->S * 4);
    rc |= balloc(b, &b->d_flow8, b->S);
    rc |= balloc(b, &b->d_rank, rank.size() * sizeof(uint2));
    rc |= balloc(b, &b->d_lab, 256 * sizeof(LabelF));
    rc |= balloc(b, &b->d_ucur, size_t(b->D) * b->S * 4);
    rc |= balloc(b, &b->d_uprev, size_t(b->D) * b->S * 4);
    rc |= balloc(b, &b->d_uadj, size_t(b->D) * b->S * 4);
    rc |= balloc(b, &b->d_upres, size_t(b->D) * b->S * 4);
    rc |= balloc(b, &b->d_status, sizeof(DevStatus));
    rc |= balloc(b, &b->d_partial, b->T * 4 * 8);
    rc |= balloc(b, &b->d_out, size_t(b->ncrops) * 4 * 8);
    if (!rc && hipHostMalloc(reinterpret_cast<void**>(&b->h_out), size_t(b->ncrops) * 4 * 8) != hipSuccess)
        rc = bfail(b, "HIP error: hipHostMalloc failed");
    if (rc) {
        std::string e = b->err;
        lbf_batch_destroy(b);
        return bfail(nullptr, e);
    }
    auto up = [&](void* dst, const void* src, size_t n) { return hipMemcpy(dst, src, n, hipMemcpyHostToDevice); };
    bool ok = up(b->d_tile_crop, b->tile_crop.data(), b->T * 4) == hipSuccess &&
              up(b->d_crops, b->crops.data(), b->crops.size() * sizeof(CropDev)) == hipSuccess &&
              up(b->d_site, site.data(), b->S * 4) == hipSuccess && up(b->d_lab8, lab8.data(), b->S) == hipSuccess &&
              up(b->d_desc, descs.data(), b->S * 4) == hipSuccess &&
              up(b->d_flow8, flow8.data(), b->S) == hipSuccess &&
              up(b->d_rank, rank.data(), rank.size() * sizeof(uint2)) == hipSuccess &&
              up(b->d_lab, lab.data(), 256 * sizeof(LabelF)) == hipSuccess;
    DevStatus st{};
    for (auto& v : st.first) v = ~0ull;
    ok = ok && up(b->d_status, &st, sizeof st) == hipSuccess;
    for (float* p : {b->d_ucur, b->d_uprev, b->d_uadj, b->d_upres})
        ok = ok && hipMemset(p, 0, size_t(b->D) * b->S * 4) == hipSuccess;
    ok = ok && hipMemset(b->f[0], 0, pop) == hipSuccess && hipMemset(b->f[1], 0, pop) == hipSuccess;
    std::vector<char> all(b->ncrops, 1);
    if (!ok || set_active(b, all)) {
        lbf_batch_destroy(b);
        return bfail(nullptr, "HIP error while uploading the batch");
    }
    *out = b;
    return 0;
}

uint64_t lbf_batch_time(const lbf_batch* b) { return b ? b->t : 0; }

uint64_t lbf_batch_sites(const lbf_batch* b, uint64_t* fluid)
{
    if (!b) return 0;
    if (fluid) *fluid = b->fluid;
    return b->sites;
}

int lbf_batch_init(lbf_batch* b)
{
    if (!b) return -1;
    std::vector<char> all(b->ncrops, 1);
    if (set_active(b, all)) return -1;
    b->t = 0;
    b->cur = 0;
    BHIP(b, hipMemsetAsync(b->d_uprev, 0, size_t(b->D) * b->S * 4, b->stream));
    if (step(b, 0u, true)) return -1;
    BHIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int lbf_batch_advance(lbf_batch* b, uint64_t n)
{
    if (!b) return -1;
    for (uint64_t k = 0; k < n; ++k)
        if (step(b, 0u, false)) return -1;
    BHIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int lbf_batch_tune(lbf_batch* b, uint32_t steps, uint32_t* best_per_cu, double* best_ms)
{
    if (!b) return -1;
    if (steps == 0) steps = 1;
    static const uint32_t per_cu[] = {1, 2, 4, 8, 16, 32, 64, 128};
    static const BatchShape shapes[] = {{256, 256}, {128, 256}, {128, 128}, {64, 256}, {64, 64}};
    double best = 1e300;
    uint32_t pick = 8;
    BatchShape pick_shape{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = bfail(b, "HIP error: hipEventCreate failed");
    for (const BatchShape& sh : shapes) {
        // work units of the whole batch: the largest useful grid
        const uint64_t units = b->T * (TILE / sh.unit);
        for (uint32_t per : per_cu) {
            if (rc) break;
            b->shape = sh;
            b->grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(units, uint64_t(b->cus) * per)));
            rc = step(b, 0u, false) || step(b, 0u, false) ? -1 : 0;   // warm-up, keeps buffer parity
            if (!rc && hipEventRecord(e0, b->stream) != hipSuccess) rc = bfail(b, "HIP error: hipEventRecord failed");
            for (uint32_t k = 0; k < steps && !rc; ++k)
                if (step(b, 0u, false)) rc = -1;
            float ms = 0.0f;
            if (!rc && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = bfail(b, "HIP error: timing events failed");
            if (!rc && ms / steps < best) {
                best = ms / steps;
                pick = per;
                pick_shape = sh;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return -1;
    b->shape = pick_shape;
    b->grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(b->T * (TILE / pick_shape.unit),
                                                                 uint64_t(b->cus) * pick)));
    if (best_per_cu) *best_per_cu = pick;
    if (best_ms) *best_ms = best;
    return 0;
}

int lbf_batch_shape(const lbf_batch* b, uint32_t* threads, uint32_t* unit, uint32_t* workgroups)
{
    if (!b) return -1;
    if (threads) *threads = b->shape.block;
    if (unit) *unit = b->shape.unit;
    if (workgroups) *workgroups = b->grid;
    return 0;
}

int lbf_batch_run(lbf_batch* b, int axis, double tol, uint32_t interval, uint64_t max_steps,
                  lbf_crop_result* results, double* seconds)
{
    if (!b || !results) return -1;
    if (axis < 0 || axis >= b->D) return bfail(b, "lbf_batch_run: axis outside the lattice");
    if (interval < 2) interval = 2;
    if (interval % 2) ++interval;   // keeps the population buffers in the same role per replay
    const auto start = std::chrono::steady_clock::now();
    // capture: interval steps (the last two keep adjacent and current velocity) + reduction
    CapturedGraph cg;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const uint64_t t0 = b->t;
    for (uint32_t k = 0; k < interval; ++k) {
        const uint32_t flags = k == interval - 2 ? 2u : (k == interval - 1 ? 1u : 0u);
        if (step(b, flags, false)) {
            (void)hipStreamEndCapture(b->stream, &cg.graph);
            return -1;
        }
    }
    hipError_t e = launch_batch_report(b->report_grid, b->S, b->D, axis, buffers(b, false), b->ncrops, b->d_partial,
                                       b->d_out, b->stream);
    BHIP(b, hipStreamEndCapture(b->stream, &cg.graph));
    BHIP(b, e);
    BHIP(b, hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0));
    b->t = t0;   // capture advanced the host counters without running anything
    std::vector<char> active(b->ncrops, 1);
    if (set_active(b, active)) return -1;
    std::vector<int> stable(b->ncrops, 0);
    std::vector<char> has_prev(b->ncrops, 0);
    const double* out = b->h_out;
    for (uint32_t c = 0; c < b->ncrops; ++c) std::memset(&results[c], 0, sizeof(lbf_crop_result));
    uint32_t remaining = b->ncrops;
    const double nu = (b->d.tau - 0.5) / 3.0;
    while (remaining && b->t < max_steps) {
        BHIP(b, hipGraphLaunch(cg.exec, b->stream));
        b->t += interval;
        BHIP(b, hipMemcpyAsync(b->h_out, b->d_out, size_t(b->ncrops) * 4 * 8, hipMemcpyDeviceToHost, b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
        bool changed = false;
        for (uint32_t c = 0; c < b->ncrops; ++c) {
            if (!active[c]) continue;
            const double du2 = out[c * 4 + 0], u2 = out[c * 4 + 1], da2 = out[c * 4 + 2], ux = out[c * 4 + 3];
            const double count = double(b->crop_sites[c]) * b->D;
            const double res = u2 > 0.0 ? std::sqrt(du2 / u2) : 0.0, rms = std::sqrt(du2 / count);
            const double ares = u2 > 0.0 ? std::sqrt(da2 / u2) : 0.0, arms = std::sqrt(da2 / count);
            lbf_crop_result& r = results[c];
            r.steps = b->t;
            r.residual = res;
            r.adjacent_residual = ares;
            const double F = b->crop_force[size_t(c) * 3 + axis];
            r.k_darcy = F != 0.0 ? nu * (ux / double(b->crop_sites[c])) / F : 0.0;
            if (!std::isfinite(du2) || !std::isfinite(u2) || !std::isfinite(da2))
                return bfail(b, "Non-finite velocity in convergence report.");
            if (has_prev[c]) {
                const bool ok = (rms <= 1e-12 || res < tol) && (arms <= 1e-12 || ares < tol);
                stable[c] = ok ? stable[c] + 1 : 0;
            }
            has_prev[c] = 1;
            if (stable[c] >= 3) {
                r.converged = 1;
                active[c] = 0;
                --remaining;
                changed = true;
            }
        }
        if (changed && remaining && set_active(b, active)) return -1;
    }
    // Only the failure bits are reported: steps replayed from the captured graph
    // carry their capture-time step numbers, so DevStatus.first is not a step here.
    DevStatus st{};
    BHIP(b, hipMemcpy(&st, b->d_status, sizeof st, hipMemcpyDeviceToHost));
    if (st.bits) return bfail(b, "Non-finite post-collision population in the batch.");
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return 0;
}

int lbf_batch_crop_size(const lbf_batch* b, uint32_t c, uint64_t* sz)
{
    if (!b || !sz || c >= b->ncrops) return 0;
    const GridDev& g = b->crops[c].g;
    const uint64_t n[3] = {g.nz, g.ny, g.nx};
    for (int k = 0; k < b->D; ++k) sz[k] = n[3 - b->D + k];
    return b->D;
}

int lbf_batch_crop_fields(lbf_batch* b, uint32_t c, float* fields)
{
    if (!b || !fields || c >= b->ncrops) return -1;
    const CropDev& cd = b->crops[c];
    BHIP(b, hipMemsetAsync(fields, 0, b->crop_sites[c] * (b->D + 1) * sizeof(float), b->stream));
    BHIP(b, launch_batch_fields(b->Q, b->S, float(b->d.density), cd.tile_begin, cd.tile_end - cd.tile_begin,
                                b->d_site, static_cast<const float*>(b->f[b->cur]), b->d_ucur, fields, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int lbf_batch_set_fields(lbf_batch* b, uint32_t c, const float* fields)
{
    if (!b || !fields || c >= b->ncrops) return -1;
    const CropDev& cd = b->crops[c];
    const BatchParams p = params(b, 0u);
    BHIP(b, launch_batch_warm(b->Q, b->d.equil == LBF_MRT, p, cd.tile_begin, cd.tile_end - cd.tile_begin, b->d_site,
                              b->d_lab8, b->d_flow8, b->d_lab, fields, static_cast<float*>(b->f[b->cur]), b->d_upres,
                              b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int lbf_batch_relax(lbf_batch* b, uint32_t steps)
{
    if (!b) return -1;
    for (uint32_t k = 0; k < steps; ++k)
        if (step(b, 4u, false)) return -1;
    // Two free steps, each averaged with the state before it.  A start that is
    // not smooth on the lattice scale (any prediction) excites a global mode of
    // the lattice Boltzmann iteration that flips sign every step and does not
    // decay; it never lets the adjacent-velocity residual drop.  The average of
    // two consecutive states cancels it.
    const uint64_t n = uint64_t(b->Q) * b->S;
    for (int k = 0; k < 2; ++k) {
        if (step(b, 0u, false)) return -1;
        BHIP(b, launch_batch_average(static_cast<float*>(b->f[b->cur]), static_cast<const float*>(b->f[b->cur ^ 1]),
                                     n, b->stream));
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int lbf_batch_crop_populations(lbf_batch* b, uint32_t c, double* out)
{
    if (!b || !out || c >= b->ncrops) return -1;
    BHIP(b, hipStreamSynchronize(b->stream));
    const CropDev& cd = b->crops[c];
    const uint64_t N = b->crop_sites[c];
    const uint64_t n = uint64_t(cd.tile_end - cd.tile_begin) * TILE;
    std::vector<uint32_t> site(n);
    BHIP(b, hipMemcpy(site.data(), b->d_site + cd.slot_base, n * 4, hipMemcpyDeviceToHost));
    std::memset(out, 0, size_t(b->Q) * N * sizeof(double));
    // the crop's slots are whole blocks of 64: one contiguous range of Q n values
    std::vector<float> hq(size_t(b->Q) * n);
    BHIP(b, hipMemcpy(hq.data(), static_cast<const float*>(b->f[b->cur]) + uint64_t(cd.slot_base) * b->Q,
                      hq.size() * 4, hipMemcpyDeviceToHost));
    for (int q = 0; q < b->Q; ++q) {
        const double shift = (b->Q == 9 ? Lattice<9>::w[q] : Lattice<19>::w[q]) * b->d.density;
        for (uint64_t k = 0; k < n; ++k)
            if (site[k] != NO_SITE)
                out[uint64_t(q) * N + site[k]] = double(hq[(k >> 6) * (uint64_t(b->Q) * 64) + uint64_t(q) * 64 + (k & 63)]) + shift;
    }
    return 0;
}

}  // extern "C"
