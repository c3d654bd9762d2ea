// setup.cpp -- host construction of operators, links and wall tables.
// Compiled without FMA contraction (-ffp-contract=off) like the reference.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lattice.h"

namespace lbf {

namespace {

// Row-major dense products with the reference's summation order
// (solver/src/algorithm/fn.cpp:29-116): every inner product starts at 0.0
// and runs over the shared index in ascending order.
std::vector<double> transpose(const std::vector<double>& a, int rows, int cols)
{
    std::vector<double> t(size_t(rows) * cols);
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) t[size_t(j) * rows + i] = a[size_t(i) * cols + j];
    return t;
}

std::vector<double> matmul(const std::vector<double>& a, const std::vector<double>& b, int m, int p,
                           int n)
{
    std::vector<double> bt = transpose(b, p, n), r(size_t(m) * n);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < p; ++k) s = s + a[size_t(i) * p + k] * bt[size_t(j) * p + k];
            r[size_t(i) * n + j] = s;
        }
    return r;
}

std::vector<double> scale_columns(const std::vector<double>& a, const double* d, int m)
{
    std::vector<double> r(size_t(m) * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) r[size_t(i) * m + j] = a[size_t(i) * m + j] * d[j];
    return r;
}

// Inverse of a matrix with orthogonal rows: column j of the inverse is row j
// scaled by 1/|row j|^2 (fn.cpp:102-116).
std::vector<double> inverse_orthogonal_rows(const std::vector<double>& a, int m, int n)
{
    std::vector<double> r(size_t(m) * n);
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s = s + a[size_t(j) * n + k] * a[size_t(j) * n + k];
        s = 1 / s;
        for (int i = 0; i < n; ++i) r[size_t(m) * i + j] = a[size_t(j) * n + i] * s;
    }
    return r;
}

// Planar rotation taking (a, o) onto the first axis (fn.cpp:118-164).
void rotation(double o, double a, double& c, double& s)
{
    c = 1;
    s = 0;
    const double h = std::sqrt(o * o + a * a);
    if (h != 0) {
        c = a / h;
        s = o / h;
    }
}

template <int Q> std::vector<double> transform_matrix()
{
    std::vector<double> m(static_cast<size_t>(Q) * Q);
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) m[size_t(i) * Q + j] = Lattice<Q>::M[i][j];
    return m;
}

}  // namespace

void mrt_rates(int Q, double tau, double* S)
{
    const double su = 1.0 / tau;
    const double sq = 8.0 * (2.0 - su) / (8.0 - su);
    if (Q == 9) {
        const double s[9] = {0, su, su, 0, sq, 0, sq, su, su};
        std::memcpy(S, s, sizeof s);
    } else {
        const double s[19] = {0, su, su, 0, sq, 0, sq, 0, sq, su, su, su, su, su, su, su, sq, sq, sq};
        std::memcpy(S, s, sizeof s);
    }
}

void build_mrt(int Q, double tau, MrtOperators& out)
{
    out.Q = Q;
    out.M = Q == 9 ? transform_matrix<9>() : transform_matrix<19>();
    out.MInv = inverse_orthogonal_rows(out.M, Q, Q);
    out.S.assign(Q, 0.0);
    mrt_rates(Q, tau, out.S.data());
    std::vector<double> h(Q);
    for (int k = 0; k < Q; ++k) h[k] = 1.0 - 0.5 * out.S[k];
    out.A = matmul(scale_columns(out.MInv, out.S.data(), Q), out.M, Q, Q, Q);
    out.AT = transpose(out.A, Q, Q);
    out.B = matmul(scale_columns(out.MInv, h.data(), Q), out.M, Q, Q, Q);
}

uint64_t neighbour(int D, const uint64_t* sz, uint64_t i, const int* e)
{
    uint64_t stride[3] = {1, 1, 1};
    for (int d = D - 1; d > 0; --d) stride[d - 1] = stride[d] * sz[d];
    uint64_t r = 0;
    for (int d = 0; d < D; ++d) {
        const int64_t c = int64_t((i / stride[d]) % sz[d]) + e[d];
        const int64_t n = int64_t(sz[d]);
        r += uint64_t(((c % n) + n) % n) * stride[d];
    }
    return r;
}

std::string build_links(int D, int Q, const uint64_t* sz, const LabelInfo* labels,
                        const uint8_t* geometry, bool rotations, Links& out)
{
    static constexpr double pi = 3.1415926535897932384626433832795;
    uint64_t N = 1;
    for (int d = 0; d < D; ++d) N *= sz[d];
    // Simulator ctor (simulator.cpp:30): every label of the geometry is defined.
    for (uint64_t i = 0; i < N; ++i)
        if (!labels[geometry[i]].type) return "map::at";

    // Sites grouped by label in ascending order (ROI::getRawIdx, roi.h:267-280):
    // wall entries are numbered in this order, as in buildIb.
    std::vector<uint64_t> start(257, 0), order(N);
    for (uint64_t i = 0; i < N; ++i) start[geometry[i] + 1]++;
    for (int l = 0; l < 256; ++l) start[l + 1] += start[l];
    {
        std::vector<uint64_t> fill(start.begin(), start.end() - 1);
        for (uint64_t i = 0; i < N; ++i) order[fill[geometry[i]]++] = i;
    }
    out.desc.assign(N, 0);
    out.wall.assign(N, 0);
    for (uint64_t i = 0; i < N; ++i) out.desc[i] = uint32_t(geometry[i]) << DESC_LABEL_SHIFT;
    out.normal.assign(size_t(D), 0.0);
    out.sinca.assign(1, 0.0);
    out.cosca.assign(1, 0.0);

    const int wallTypes = T_VD | T_PB | T_VB | T_GV | T_HF;
    uint64_t stride[3] = {1, 1, 1};
    for (int d = D - 1; d > 0; --d) stride[d - 1] = stride[d] * sz[d];
    int32_t c = 1;
    for (int l = 0; l < 256; ++l) {
        if (!(labels[l].type & wallTypes)) continue;
        for (uint64_t k = start[l]; k < start[l + 1]; ++k) {
            const uint64_t i = order[k];
            double norm[3] = {0.0, 0.0, 0.0}, ca = 0;
            int nCA = 0;
            uint32_t bits = 0;
            uint64_t film = 0;
            for (int q = 1; q < Q; ++q) {
                int e[3];
                for (int d = 0; d < D; ++d) e[d] = Q == 9 ? Lattice<9>::e[q][d] : Lattice<19>::e[q][d];
                const uint8_t nb = geometry[neighbour(D, sz, i, e)];
                const int type = labels[nb].type;
                if (type == T_BB || type == T_SD) {
                    bits |= 1u << q;
                    for (int d = 0; d < D; ++d) norm[d] -= e[d];
                    ca += labels[nb].value;
                    nCA++;
                } else if (type == T_HF) {
                    // ibF[phase][q] (simulator.cpp:418-421): the other phase bounces back
                    film |= uint64_t(1) << (labels[nb].phase == 1 ? q : 32 + q);
                }
            }
            if (film) {
                if (out.film.empty()) out.film.assign(N, 0);
                out.film[i] = film;
                out.desc[i] |= DESC_FILM;
            }
            if (bits) {
                out.desc[i] |= bits;
                ca /= nCA;
                out.wall[i] = c++;
                for (int d = 0; d < D; ++d) out.normal.push_back(norm[d]);
                out.sinca.push_back(std::sin(ca / 180.0 * pi));
                out.cosca.push_back(std::cos(ca / 180.0 * pi));
            }
        }
    }
    out.nwall = uint64_t(c);

    out.tib.assign(size_t(c) * D * D, 0.0);
    out.tibinv.assign(size_t(c) * D * D, 0.0);
    if (!rotations) return "";
    for (int32_t w = 1; w < c; ++w) {
        const double* n = &out.normal[size_t(w) * D];
        if (D == 3) {
            double cx, sx, cxi, sxi, cz, sz_, czi, szi;
            rotation(n[0], n[1], cx, sx);
            rotation(-n[0], n[1], cxi, sxi);
            const double val = std::sqrt(n[0] * n[0] + n[1] * n[1]);
            rotation(val, n[2], cz, sz_);
            rotation(-val, n[2], czi, szi);
            const std::vector<double> tx = {cx, -sx, 0, sx, cx, 0, 0, 0, 1};
            const std::vector<double> txInv = {cxi, -sxi, 0, sxi, cxi, 0, 0, 0, 1};
            const std::vector<double> tz = {1, 0, 0, 0, cz, -sz_, 0, sz_, cz};
            const std::vector<double> tzInv = {1, 0, 0, 0, czi, -szi, 0, szi, czi};
            const std::vector<double> t = matmul(tz, tx, 3, 3, 3);
            const std::vector<double> ti = matmul(txInv, tzInv, 3, 3, 3);
            std::memcpy(&out.tib[size_t(w) * 9], t.data(), 9 * sizeof(double));
            std::memcpy(&out.tibinv[size_t(w) * 9], ti.data(), 9 * sizeof(double));
        } else {
            double c2, s2, c2i, s2i;
            rotation(n[0], n[1], c2, s2);
            rotation(-n[0], n[1], c2i, s2i);
            const double t[4] = {c2, -s2, s2, c2}, ti[4] = {c2i, -s2i, s2i, c2i};
            std::memcpy(&out.tib[size_t(w) * 4], t, sizeof t);
            std::memcpy(&out.tibinv[size_t(w) * 4], ti, sizeof ti);
        }
    }
    return "";
}

std::string wall_gradient_inverse(int D, int Q, const uint64_t* sz, const LabelInfo* labels,
                                  const uint8_t* geometry, Links& out)
{
    uint64_t N = 1;
    for (int d = 0; d < D; ++d) N *= sz[d];
    out.ginv.assign(size_t(out.nwall) * D * D, 0.0);
    for (uint64_t i = 0; i < N; ++i) {
        if (!out.wall[i]) continue;
        double aug[3][6] = {};
        for (int q = 1; q < Q; ++q) {
            int e[3];
            for (int d = 0; d < D; ++d) e[d] = Q == 9 ? Lattice<9>::e[q][d] : Lattice<19>::e[q][d];
            const int type = labels[geometry[neighbour(D, sz, i, e)]].type;
            if (type == T_SD || type == T_BB) continue;
            const double w = Q == 9 ? Lattice<9>::w[q] : Lattice<19>::w[q];
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) aug[a][b] += w * e[a] * e[b];
        }
        for (int a = 0; a < D; ++a) aug[a][D + a] = 1.0;
        for (int col = 0; col < D; ++col) {
            int pivot = col;
            for (int row = col + 1; row < D; ++row)
                if (std::abs(aug[row][col]) > std::abs(aug[pivot][col])) pivot = row;
            if (std::abs(aug[pivot][col]) < 1e-12)
                return "Rank-deficient fluid wall stencil; reconstructed derivatives are unsupported here.";
            for (int k = 0; k < 2 * D; ++k) std::swap(aug[col][k], aug[pivot][k]);
            const double diagonal = aug[col][col];
            for (int k = 0; k < 2 * D; ++k) aug[col][k] /= diagonal;
            for (int row = 0; row < D; ++row) {
                if (row == col) continue;
                const double scale = aug[row][col];
                for (int k = 0; k < 2 * D; ++k) aug[row][k] -= scale * aug[col][k];
            }
        }
        double* inv = &out.ginv[size_t(out.wall[i]) * D * D];
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) inv[a * D + b] = aug[a][D + b];
    }
    return "";
}

FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d) ++s;
    f.s = s;
    // m = floor(2^32 * (2^s - d) / d) + 1
    const unsigned __int128 num = (unsigned __int128)((uint64_t(1) << s) - d) << 32;
    f.m = uint32_t(uint64_t(num / d) + 1);
    return f;
}

}  // namespace lbf
