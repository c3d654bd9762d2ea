// lattice.h -- D2Q9 / D3Q19 lattice tables shared by host setup and HIP kernels.
//
// Direction numbering, components and auxiliary lists follow the reference
// exactly (solver/src/grid/gridspace.cpp:82-96 for D2Q9, :106-143 for D3Q19):
// components are stored in the reference's order, slowest axis first, i.e.
// (y,x) for D2Q9 and (z,y,x) for D3Q19, so q = 1 is +x in both.  The MRT
// transform matrices are the reference's (solver/src/lbm/simulator.cpp:162-197)
// and the weights are solver/src/lbm/simulator.cpp:438-463.  The numbering
// has to be identical because the checkpoint layout [phase][q][site] and the
// accumulation orders of the FP64 mode depend on it.
#pragma once

#if defined(__HIPCC__)
#define LBF_HD __host__ __device__
#else
#define LBF_HD
#endif

namespace lbf {

template <int Q> struct Lattice;

template <> struct Lattice<9> {
    static constexpr int D = 2, Q = 9, NREL = 6, NHALF = 4;
    // e[q][d], d in (y, x) order
    static constexpr int e[9][2] = {{0, 0}, {0, 1}, {1, 0}, {0, -1}, {-1, 0},
                                    {1, 1}, {1, -1}, {-1, -1}, {-1, 1}};
    static constexpr int op[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    // QrelD: directions with a non-zero component along axis d, reference order
    static constexpr int reld[2][6] = {{2, 5, 6, 4, 7, 8}, {1, 5, 8, 3, 6, 7}};
    // one representative of every opposite pair (qa.half)
    static constexpr int half[4] = {1, 2, 5, 6};
    static constexpr double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    static constexpr int M[9][9] = {
        {1, 1, 1, 1, 1, 1, 1, 1, 1},      {-4, -1, -1, -1, -1, 2, 2, 2, 2},
        {4, -2, -2, -2, -2, 1, 1, 1, 1},  {0, 1, 0, -1, 0, 1, -1, -1, 1},
        {0, -2, 0, 2, 0, 1, -1, -1, 1},   {0, 0, 1, 0, -1, 1, 1, -1, -1},
        {0, 0, -2, 0, 2, 1, 1, -1, -1},   {0, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 1, -1, 1, -1}};
    LBF_HD static constexpr bool diagonal(int q) { return q > 4; }
};

template <> struct Lattice<19> {
    static constexpr int D = 3, Q = 19, NREL = 10, NHALF = 9;
    // e[q][d], d in (z, y, x) order
    static constexpr int e[19][3] = {
        {0, 0, 0},  {0, 0, 1},  {0, 0, -1}, {0, 1, 0},   {0, -1, 0}, {1, 0, 0},  {-1, 0, 0},
        {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},
        {-1, 0, -1}, {1, 1, 0}, {1, -1, 0}, {-1, 1, 0},  {-1, -1, 0}};
    static constexpr int op[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    static constexpr int reld[3][10] = {{5, 11, 12, 15, 16, 6, 13, 14, 17, 18},
                                        {3, 8, 7, 15, 17, 4, 9, 10, 16, 18},
                                        {1, 7, 9, 11, 13, 2, 8, 10, 12, 14}};
    static constexpr int half[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
    static constexpr double w[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                     1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                     1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                     1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    static constexpr int M[19][19] = {
        {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {-30, -11, -11, -11, -11, -11, -11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8},
        {12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, -4, 4, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
        {0, 0, 0, 0, 0, -4, 4, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
        {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
        {0, -4, -4, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
        {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
        {0, 0, 0, -2, -2, 2, 2, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, -1, -1, 1, 1}};
    LBF_HD static constexpr bool diagonal(int q) { return q > 6; }
};

// Zou-He tables of the axial directions dir = 1..2D (gridspace.cpp:79-85
// D2Q9, :123-137 D3Q19): rel[dir] = directions with a component along dir
// (Qrel), perp[dir] = directions without a component along dir's axis
// (Qprep), axis[dir] = that axis in reference component order (Qaxis),
// sign[dir] = its sign (Qsign).  List order is the reference's: the FP64
// boundary sums run in it.
template <int Q> struct ZouHe;

template <> struct ZouHe<9> {
    static constexpr int NREL = 3, NPERP = 3;
    static constexpr int rel[5][3] = {{0, 0, 0}, {1, 5, 8}, {2, 5, 6}, {3, 6, 7}, {4, 7, 8}};
    static constexpr int perp[5][3] = {{0, 0, 0}, {0, 2, 4}, {0, 1, 3}, {0, 2, 4}, {0, 1, 3}};
    static constexpr int axis[5] = {0, 1, 0, 1, 0};
    static constexpr int sign[5] = {0, 1, 1, -1, -1};
};

template <> struct ZouHe<19> {
    static constexpr int NREL = 5, NPERP = 9;
    static constexpr int rel[7][5] = {{0, 0, 0, 0, 0},    {1, 7, 9, 11, 13},  {2, 8, 10, 12, 14},
                                      {3, 8, 7, 15, 17},  {4, 9, 10, 16, 18}, {5, 11, 12, 15, 16},
                                      {6, 13, 14, 17, 18}};
    static constexpr int perp[7][9] = {{0, 0, 0, 0, 0, 0, 0, 0, 0},     {0, 3, 4, 5, 6, 15, 16, 17, 18},
                                       {0, 3, 4, 5, 6, 15, 16, 17, 18}, {0, 1, 2, 5, 6, 11, 12, 13, 14},
                                       {0, 1, 2, 5, 6, 11, 12, 13, 14}, {0, 1, 2, 3, 4, 7, 8, 9, 10},
                                       {0, 1, 2, 3, 4, 7, 8, 9, 10}};
    static constexpr int axis[7] = {0, 2, 2, 1, 1, 0, 0};
    static constexpr int sign[7] = {0, 1, -1, 1, -1, 1, -1};
};

// Label type bits (solver/inc/file/fileprm.h:44-53).
enum : int { T_VD = 0x1, T_GV = 0x2, T_PB = 0x4, T_VB = 0x8, T_AB = 0x10, T_BB = 0x20,
             T_HF = 0x40, T_SD = 0x80 };

// Per-site descriptor word: bits 1..Q-1 = bounce-back links (site in ibS[q]),
// bit 19 = film links present, bits 24..31 = geometry label.
constexpr unsigned DESC_LABEL_SHIFT = 24;
constexpr unsigned DESC_LINK_MASK = 0x7ffffu;
// Bit 19: the site has a link into a half-permeable film site of the other
// phase for one of its phases (the per-site film word holds which).
constexpr unsigned DESC_FILM = 1u << 19;

}  // namespace lbf
