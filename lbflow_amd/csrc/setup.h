// setup.h -- host-side construction of everything the kernels read:
// MRT operators, per-site link descriptors, wall tables and the label table.
// All FP64 arithmetic here keeps the reference's operation order so that
// device results can be compared bit for bit.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lbf {

struct LabelInfo {
    int type = 0;      // T_* bit, 0 = undefined
    int dir = 0;
    int phase = 0;
    double value = 0.0;
};

// Dense 19x19 (or 9x9) operators of the MRT collision
// (solver/src/lbm/simulator.cpp:162-216, scsp.cpp:420-427, colorfluid.cpp:315-322).
struct MrtOperators {
    int Q = 0;
    std::vector<double> M, MInv, S, A, AT, B;   // row-major Q*Q (S: Q)
};
void build_mrt(int Q, double tau, MrtOperators& out);
// Relaxation rates S(tau) of Simulator::getMRTS (legacy spectrum).
void mrt_rates(int Q, double tau, double* S);

// Boundary links (Simulator::buildIb, simulator.cpp:386-436) and the wall
// rotation matrices (ColorFluid::calcTibNorm, colorfluid.cpp:1111-1129).
struct Links {
    std::vector<uint32_t> desc;     // per site: link bits | label << 24
    std::vector<int32_t> wall;      // per site: ibI (0 = not a wall site)
    std::vector<double> normal;     // per wall entry: D components
    std::vector<double> sinca, cosca;
    std::vector<double> tib, tibinv;  // per wall entry: D*D
    std::vector<double> ginv;       // per wall entry: D*D, renormalized wall gradient
    std::vector<uint64_t> film;     // per site: film links (ibF), empty without HF labels
    uint64_t nwall = 1;             // entries incl. the unused entry 0
};

// Inverse of the fluid-link second moment sum_q w_q e_q e_q^T of every wall
// site (ColorFluid::calcWallGradientInverse, colorfluid.cpp:917-957) into
// out.ginv.  Returns an empty string on success, else the reference's message.
std::string wall_gradient_inverse(int D, int Q, const uint64_t* sz, const LabelInfo* labels,
                                  const uint8_t* geometry, Links& out);

// Validates the labels used by the geometry and builds the link tables.
// Returns an empty string on success, else the reference's message.
std::string build_links(int D, int Q, const uint64_t* sz, const LabelInfo* labels,
                        const uint8_t* geometry, bool rotations, Links& out);

// Periodic neighbour of site i along lattice vector e (components slowest first).
uint64_t neighbour(int D, const uint64_t* sz, uint64_t i, const int* e);

// Magic-number unsigned division by a runtime constant (round-up method).
struct FastDiv {
    uint32_t d = 1, m = 1, s = 0;
};
FastDiv make_fastdiv(uint32_t d);

}  // namespace lbf
