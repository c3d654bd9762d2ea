// lbsim.cpp -- drop-in replacement for LBFlow's `LBFlow <input.prm>` process.
//
// Process contract of the reference (solver/apps/LBSim/lbsim.cpp:16-74):
// stdout markers "Model constructing...Done!", "Simulation started!",
// "Initializing...Done!", "Step N - Time elapsed: ...", report lines,
// "LBFlowTiming {...}", "Simulation finished!"; VTI snapshots
// outPath.%010d.vti at t = 0 and every writeInterval; FP64 checkpoints
// outPath.%010d.rst + .rst.meta at every restartInterval and at
// convergence; any error prints "<rank> Failed to excute: <what>" and exits
// with a non-zero status.  The time-step loop orchestration (cadence,
// reports, convergence, checkpoints) is SCSP::Run / ColorFluid::Run
// (solver/src/lbm/scsp.cpp:46-114, colorfluid.cpp:109-212); every per-site
// operator runs on the GPU behind include/lbflow.h.
//
// Options beyond the reference (ignored by it, so one input file serves
// both): input.prm keys `enginePrecision "fp64"|"fp32";` and
// `engineLayout "exact"|"sparse";`, or command-line flags --fp32, --sparse.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbflow.h"
#include "prm.h"

namespace lbsim {

using std::size_t;
using std::string;

double now()
{
    using clock = std::chrono::steady_clock;
    static const clock::time_point origin = clock::now();
    return std::chrono::duration<double>(clock::now() - origin).count();
}

// helper/logger.cpp:18-75
class Logger {
public:
    explicit Logger(std::ostream& os) : out_(os), timer_(now()) {}
    void log(const string& s, bool showTimer = true)
    {
        out_ << s;
        if (showTimer) out_ << "Time elapsed: " << human(now() - timer_) << std::endl;
        if (autoReset_) timer_ = now();
        out_.flush();
    }
    void reset(bool autoReset = false)
    {
        autoReset_ = autoReset;
        timer_ = now();
    }
    static string human(double time)
    {
        const size_t t = static_cast<size_t>(time);
        const size_t days = t / 86400, hh = (t % 86400) / 3600, mm = (t % 3600) / 60, ss = t % 60;
        const double left = time - static_cast<double>(t) + ss;
        std::ostringstream os;
        os.precision(6);
        os << days << "d," << std::setw(2) << std::setfill('0') << hh << ":" << std::setw(2) << std::setfill('0')
           << mm << ":" << std::setw(9) << std::setfill('0') << std::setiosflags(std::ios::fixed) << left;
        return os.str();
    }

private:
    std::ostream& out_;
    double timer_;
    bool autoReset_ = false;
};

string numbered(const string& base, size_t t, const char* ext)
{
    std::ostringstream os;
    os << base << "." << std::setw(10) << std::setfill('0') << t << ext;
    return os.str();
}

void append(const string& path, const void* data, size_t bytes)
{
    std::FILE* f = std::fopen(path.c_str(), "ab");
    if (!f) throw std::runtime_error("FileRaw::Save() - Can not open file for output.");
    if (bytes && std::fwrite(data, 1, bytes, f) != bytes) {
        std::fclose(f);
        throw std::runtime_error("FileRaw::Save() - Failed to save to file");
    }
    std::fclose(f);
}

// Sub-box of one rank of an R-rank reference run (sizes in reference order,
// slowest first): GridMc::initCoreDistribution and initLclGs
// (grid/gridmc.cpp:55-95) with the rank's core coordinates from CalcI2P
// (grid.h:153-161).  The prime factors of R, largest first (ALG::Prime,
// algorithm/fn.cpp:14-27), each split the axis with the largest remaining
// extent; along an axis split in c parts of a length L = n / c, the last
// n % c parts hold one site more.
struct RankBox {
    size_t start[3] = {0, 0, 0}, size[3] = {1, 1, 1};
};

std::vector<RankBox> rank_boxes(size_t D, const std::vector<size_t>& sz, size_t R)
{
    std::vector<size_t> primes;
    for (size_t v = R, i = 2; v > 1;) {
        if (v % i == 0) {
            primes.insert(primes.begin(), i);
            v /= i;
        } else {
            ++i;
        }
    }
    std::vector<double> rest(sz.begin(), sz.end());
    std::vector<size_t> cores(D, 1);
    for (size_t prime : primes) {
        const size_t k = size_t(std::max_element(rest.begin(), rest.end()) - rest.begin());
        if (rest[k] < double(prime)) throw std::invalid_argument("too many processors");
        cores[k] *= prime;
        rest[k] /= double(prime);
    }
    std::vector<RankBox> boxes(R);
    for (size_t r = 0; r < R; ++r) {
        size_t rem = r;
        for (size_t d = D; d-- > 0;) {
            const size_t c = rem % cores[d];
            rem /= cores[d];
            const size_t mode = cores[d] - sz[d] % cores[d], len = sz[d] / cores[d];
            boxes[r].start[d] = c * len + (mode >= c ? 0 : c - mode);
            boxes[r].size[d] = len + (mode > c ? 0 : 1);
        }
    }
    return boxes;
}

// VTI writer (file/filevti.cpp:39-172): header, one Piece per rank of an
// R-rank reference run (R = 1: the whole grid), appended raw blocks
// [uint64 byte length][Float64 ...] per array and rank in rank order, scalars
// in name order, then vectors with interleaved (x, y, z) components.  Piece
// offsets follow FileVti::Init: the sites of the ranks before this one,
// offRaw = sum_d start_d stride_d prod_{d' < d} size_d', plus one length word
// per earlier rank.
class Vti {
public:
    struct Array {
        string name;
        std::vector<const double*> parts;   // scalar: 1 part; vector: components x, y(, z)
    };
    Vti(string base, size_t D, const std::vector<size_t>& sz, size_t ranks = 1)
        : base_(std::move(base)), D_(D), sz_(sz), boxes_(rank_boxes(D, sz, ranks))
    {
        N_ = 1;
        for (size_t n : sz) N_ *= n;
    }
    void save(size_t t, const std::vector<Array>& scalars, const std::vector<Array>& vectors) const
    {
        const size_t R = boxes_.size();
        std::ostringstream head;
        head << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">"
             << "<ImageData WholeExtent=\" ";
        for (size_t d = D_; d > 0; --d) head << "0 " << sz_[d - 1] - 1 << " ";
        for (size_t d = 3 - D_; d > 0; --d) head << "0 0 ";
        head << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">";
        string text = head.str();
        std::vector<size_t> stride(D_, 1);
        for (size_t d = D_ - 1; d > 0; --d) stride[d - 1] = stride[d] * sz_[d];
        for (size_t r = 0; r < R; ++r) {
            const RankBox& b = boxes_[r];
            std::ostringstream piece;
            piece << "<Piece Extent=\"";
            for (size_t d = D_; d > 0; --d) piece << b.start[d - 1] << " " << b.start[d - 1] + b.size[d - 1] - 1 << " ";
            for (size_t d = 3 - D_; d > 0; --d) piece << "0 0 ";
            piece << "\"><PointData Scalars=\"";
            for (const auto& a : scalars) piece << a.name << " ";
            piece << "\" Vectors=\"";
            for (const auto& a : vectors) piece << a.name << " ";
            piece << "\" >";
            size_t offRaw = 0, lclN = 1;
            for (size_t d = 0; d < D_; ++d) {
                offRaw += b.start[d] * stride[d] * lclN;
                lclN *= b.size[d];
            }
            size_t offset = offRaw * sizeof(double) + r * sizeof(uint64_t);
            for (const auto& a : scalars) {
                piece << "<DataArray type=\"Float64\" Name=\"" << a.name << "\" format=\"appended\" offset=\""
                      << offset << "\"/>";
                offset += N_ * sizeof(double) + R * sizeof(uint64_t);
            }
            if (!vectors.empty())
                offset = offRaw * sizeof(double) * vectors.front().parts.size() + r * sizeof(uint64_t) +
                         N_ * scalars.size() * sizeof(double) + R * scalars.size() * sizeof(uint64_t);
            for (const auto& a : vectors) {
                piece << "<DataArray type=\"Float64\" Name=\"" << a.name << "\" NumberOfComponents=\""
                      << a.parts.size() << "\" format=\"appended\" offset=\"" << offset << "\"/>";
                offset += N_ * sizeof(double) * 3 + R * sizeof(uint64_t);
            }
            piece << "</PointData><CellData></CellData></Piece>";
            text += piece.str();
        }
        text += "</ImageData><AppendedData encoding=\"raw\">_";
        const string path = numbered(base_, t, ".vti");
        std::remove(path.c_str());
        std::vector<char> body(text.begin(), text.end());
        auto put = [&body](const void* p, size_t n) {
            const char* c = static_cast<const char*>(p);
            body.insert(body.end(), c, c + n);
        };
        // one rank's sub-box of an array, local sites in z, y, x order, components interleaved
        auto block = [&](const Array& a, const RankBox& b) {
            const size_t nc = a.parts.size();
            size_t lclN = 1;
            for (size_t d = 0; d < D_; ++d) lclN *= b.size[d];
            const uint64_t len = nc * lclN * sizeof(double);
            put(&len, sizeof len);
            std::vector<double> out;
            out.reserve(nc * lclN);
            const size_t nz = D_ == 3 ? b.size[0] : 1, ny = b.size[D_ - 2], nx = b.size[D_ - 1];
            const size_t z0 = D_ == 3 ? b.start[0] : 0, y0 = b.start[D_ - 2], x0 = b.start[D_ - 1];
            const size_t sy = sz_[D_ - 1], sz = D_ == 3 ? sz_[1] * sz_[2] : 0;
            for (size_t z = 0; z < nz; ++z)
                for (size_t y = 0; y < ny; ++y)
                    for (size_t x = 0; x < nx; ++x) {
                        const size_t i = (z0 + z) * sz + (y0 + y) * sy + (x0 + x);
                        for (size_t c = 0; c < nc; ++c) out.push_back(a.parts[c][i]);
                    }
            put(out.data(), len);
        };
        for (const auto& a : scalars)
            for (const RankBox& b : boxes_) block(a, b);
        for (const auto& a : vectors)
            for (const RankBox& b : boxes_) block(a, b);
        const char* foot = "</AppendedData></VTKFile>";
        put(foot, std::strlen(foot));
        append(path, body.data(), body.size());
    }

private:
    string base_;
    size_t D_, N_;
    std::vector<size_t> sz_;
    std::vector<RankBox> boxes_;
};

class Simulation {
public:
    explicit Simulation(const Params& p, const std::vector<string>& flags) : p_(p)
    {
        p_.validate();
        color_ = p_.fluidModel() == "color";
        NP_ = color_ ? 2 : 1;
        D_ = p_.dimension().size();
        Q_ = p_.directions();
        sz_ = p_.dimension();
        N_ = 1;
        for (size_t n : sz_) N_ *= n;
        // geometry (FileRaw<uint8_t>::Load, file/fileraw.h:232-291)
        std::FILE* f = std::fopen(p_.gridFile().c_str(), "rb");
        if (!f) throw std::runtime_error("FileRaw::Load() - Can not open file for input.");
        geometry_.assign(N_, 0);
        const size_t got = std::fread(geometry_.data(), 1, N_, f);
        std::fclose(f);
        if (got != N_) throw std::runtime_error("FileRaw::Load() - Failed to read from raw file");
        for (uint8_t v : geometry_) p_.label(v);   // every label used must be defined (simulator.cpp:30)
        if (p_.reportDir() != 0) {
            const size_t axis = Params::axisOf(D_, p_.reportDir());
            for (size_t v : p_.reportSlice())
                if (v >= sz_[axis]) throw std::invalid_argument("Report slice is outside the selected axis.");
        }
        unsupported();

        string precision = p_.enginePrecision(), layout = p_.engineLayout(), ranks = p_.outputRanks();
        const string ranksFlag = "--output-ranks=";
        for (const string& fl : flags) {
            if (fl == "--fp32") precision = "fp32";
            else if (fl == "--fp64") precision = "fp64";
            else if (fl == "--sparse") layout = "sparse";
            else if (fl == "--exact") layout = "exact";
            else if (fl.compare(0, ranksFlag.size(), ranksFlag) == 0) ranks = fl.substr(ranksFlag.size());
            else throw std::invalid_argument("Unknown option " + fl);
        }
        {
            size_t used = 0;
            unsigned long r = 0;
            try {
                r = std::stoul(ranks, &used);
            } catch (const std::exception&) {
                used = 0;
            }
            if (used != ranks.size() || r == 0 || r > 65536)
                throw std::invalid_argument("outputRanks must be a rank count of at least 1.");
            outRanks_ = r;
            rank_boxes(D_, sz_, outRanks_);   // an R-rank reference run fails here too (gridmc.cpp:67-68)
        }
        lbf_desc d{};
        d.D = int32_t(D_);
        d.Q = int32_t(Q_);
        d.model = color_ ? LBF_MODEL_COLOR : LBF_MODEL_SCSP;
        d.equil = p_.equilModel() == "MRT" ? LBF_MRT : LBF_SRT;
        d.precision = precision == "fp32" ? LBF_FP32 : LBF_FP64;
        d.flags = layout == "sparse" ? 0 : LBF_STREAM_SOLIDS;
        d.harmonic = color_ && p_.viscosityInterpolation() == "harmonic";
        d.wall_gradient = !color_ ? LBF_WALL_LEGACY
                          : p_.wallGradient() == "renormalized" ? LBF_WALL_RENORMALIZED
                          : p_.wallGradient() == "tanh" ? LBF_WALL_TANH : LBF_WALL_LEGACY;
        d.capillary = color_ && p_.capillaryModel() == "csf" ? LBF_CAPILLARY_CSF : LBF_CAPILLARY_PERTURBATION;
        d.curvature = color_ && p_.curvatureStencil() == "gradient" ? LBF_CURVATURE_GRADIENT : LBF_CURVATURE_NORMAL;
        for (size_t k = 0; k < D_; ++k) d.sz[k] = sz_[k];
        for (size_t k = 0; k < NP_; ++k) {
            d.tau[k] = p_.relaxationTime()[k];
            d.density[k] = p_.density()[k];
        }
        if (color_) {
            d.alpha0 = p_.alpha0();
            d.sigma = p_.surfaceTension();
            d.beta = p_.interfaceThickness();
        }
        std::vector<lbf_label> labels(256);
        for (const auto& kv : p_.labels()) {
            lbf_label& l = labels[kv.first];
            l.type = kv.second.type;
            l.dir = int32_t(kv.second.dir);
            l.phase = int32_t(kv.second.phase);
            l.value = kv.second.value;
        }
        if (lbf_create(&d, labels.data(), geometry_.data(), &ctx_)) throw std::runtime_error(lbf_last_error(nullptr));
        fluid_ = 0;
        const int active = color_ ? (VD | GV | VB | PB | HF) : (VD | GV | VB | PB);
        for (uint8_t v : geometry_)
            if (p_.label(v).type & active) fluid_ += 1.0;
        rho0_.assign(N_, 0.0);
        if (color_) rho1_.assign(N_, 0.0);
        u_.assign(D_ * N_, 0.0);
    }

    ~Simulation() { lbf_destroy(ctx_); }

    void run(std::ostream& log);

private:
    enum { INIT_TIME, NUMERICAL_TIME, FIELD_TIME, CHECKPOINT_TIME, TIMING_CATEGORIES };

    void unsupported() const
    {
        // Operators outside this engine: the remaining experimental options of
        // ColorFluid (tangent/halfway/ghost walls, wall recolouring, phase
        // flux, projection recolouring, the liu2015 spectrum, wall slip).
        auto need = [](const string& key, const string& value, const string& want) {
            if (value != want)
                throw std::invalid_argument(key + " \"" + value + "\" is not implemented by this engine (only \"" +
                                            want + "\").");
        };
        if (color_) {
            const string wg = p_.wallGradient();
            if (wg != "legacy" && wg != "renormalized" && wg != "tanh")
                throw std::invalid_argument("wallGradient \"" + wg +
                                            "\" is not implemented by this engine (only \"legacy\", \"renormalized\", "
                                            "\"tanh\").");
            need("wallRecoloring", p_.wallRecoloring(), "legacy");
            need("phaseBoundary", p_.phaseBoundary(), "bounce");
            need("recoloringModel", p_.recoloringModel(), "cosine");
        }
        need("mrtSpectrum", p_.mrtSpectrum(), "legacy");
        if (p_.wallSlipLength() > 0.0) throw std::invalid_argument("wallSlipLength is not implemented by this engine.");
    }

    // Failure records of the device, thrown at the point of the reference's loop
    // where the failing operator would have thrown (macro before the output of a
    // step, collision after it).
    void failures(size_t t, bool afterCollision)
    {
        uint32_t bits = 0;
        uint64_t first[6];
        if (lbf_status(ctx_, &bits, first)) throw std::runtime_error(lbf_last_error(ctx_));
        if (!bits) return;
        uint64_t tmin = UINT64_MAX;
        for (int k = 0; k < 6; ++k)
            if (bits & (1u << k)) tmin = std::min(tmin, first[k]);
        // bits in the order the reference's step throws them: calcMacroVar
        // (0-3), bodyForce's capillary force (5), before the output; then the
        // collision's population check (4)
        static const int order[6] = {0, 1, 2, 3, 5, 4};
        const bool macroAt = [&] {
            for (int j = 0; j < 5; ++j)
                if ((bits & (1u << order[j])) && first[order[j]] == tmin) return true;
            return false;
        }();
        if (tmin > t) return;
        if (!afterCollision && tmin == t && !macroAt) return;   // collision of step t still ahead
        for (int k : order)
            if ((bits & (1u << k)) && first[k] == tmin) {
                if (k == 0)
                    throw std::runtime_error(color_ ? "Non-finite or non-positive total fluid density."
                                                    : "Non-finite or non-positive fluid density.");
                throw std::runtime_error(lbf_status_message(1u << k));
            }
    }

    void sync(double* seconds)
    {
        const double b = now();
        if (lbf_synchronize(ctx_)) throw std::runtime_error(lbf_last_error(ctx_));
        seconds[NUMERICAL_TIME] += now() - b;
    }

    void saveFields(size_t t)
    {
        if (lbf_fields(ctx_, rho0_.data(), color_ ? rho1_.data() : nullptr, u_.data()))
            throw std::runtime_error(lbf_last_error(ctx_));
        std::vector<Vti::Array> scalars, vectors;
        Vti::Array vel{color_ ? "v" : "velocity", {}};
        for (size_t d = D_; d != 0; --d) vel.parts.push_back(u_.data() + (d - 1) * N_);
        if (color_) {
            scalars.push_back({"r0", {rho0_.data()}});
            scalars.push_back({"r1", {rho1_.data()}});
        } else {
            scalars.push_back({"density", {rho0_.data()}});
        }
        vectors.push_back(vel);
        Vti(p_.outPath(), D_, sz_, outRanks_).save(t, scalars, vectors);
        fieldsAt_ = t;
    }

    void fetchFields(size_t t)
    {
        if (fieldsAt_ == t) return;
        if (lbf_fields(ctx_, rho0_.data(), color_ ? rho1_.data() : nullptr, u_.data()))
            throw std::runtime_error(lbf_last_error(ctx_));
        fieldsAt_ = t;
    }

    void saveRst(size_t t)
    {
        // SCSP::saveRst / ColorFluid::saveRst + Simulator::prepareRestart / saveRestartMetadata
        const string path = numbered(p_.outPath(), t, ".rst");
        {
            std::ofstream data(path, std::ios::binary | std::ios::trunc);
            if (!data) throw std::runtime_error("Cannot create restart population file.");
        }
        std::remove((path + ".meta").c_str());
        std::vector<double> pops(NP_ * Q_ * N_);
        if (lbf_get_populations(ctx_, pops.data())) throw std::runtime_error(lbf_last_error(ctx_));
        append(path, pops.data(), pops.size() * sizeof(double));
        std::ofstream out(path + ".meta", std::ios::trunc);
        out << "DRLAB_OUTGOING_V1\n" << t << ' ' << NP_ << ' ' << D_ << ' ' << Q_ << '\n';
        for (size_t n : sz_) out << n << ' ';
        out << '\n'
            << "NEXT_STREAM " << t + 1 << '\n'
            << "INITIAL_COLLISION_INCLUDED " << t + 1 << '\n'
            << "CONVERGENCE_HISTORY_RESET_ON_LOAD\n";
        if (!out) throw std::runtime_error("Cannot write restart metadata.");
    }

    void loadRst(size_t t)
    {
        // Simulator::validateRestart (simulator.cpp:361-384) + FileRaw<double> loads
        const string path = numbered(p_.outPath(), t, ".rst");
        {
            std::ifstream in(path + ".meta");
            string format;
            size_t st = 0, sp = 0, sd = 0, sq = 0;
            in >> format >> st >> sp >> sd >> sq;
            if (!in || format != "DRLAB_OUTGOING_V1" || st != t || sp != NP_ || sd != D_ || sq != Q_)
                throw std::runtime_error(
                    "Missing or incompatible restart metadata; historical unlabeled checkpoints are not accepted.");
            for (size_t n : sz_) {
                size_t stored = 0;
                in >> stored;
                if (!in || n != stored) throw std::runtime_error("Restart grid dimensions differ.");
            }
            std::ifstream data(path, std::ios::binary | std::ios::ate);
            const size_t expected = N_ * Q_ * NP_ * sizeof(double);
            if (!data || data.tellg() != static_cast<std::streamoff>(expected))
                throw std::runtime_error("Restart population file has an invalid length.");
        }
        std::vector<double> pops(NP_ * Q_ * N_);
        std::FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("FileRaw::Load() - Can not open file for input.");
        const size_t got = std::fread(pops.data(), sizeof(double), pops.size(), f);
        std::fclose(f);
        if (got != pops.size()) throw std::runtime_error("FileRaw::Load() - Failed to read from raw file");
        if (lbf_set_populations(ctx_, pops.data(), t)) throw std::runtime_error(lbf_last_error(ctx_));
        hasReportVelocity_ = false;
        stableReports_ = 0;
    }

    bool report(std::ostream& os);
    bool updateConvergence(std::ostream& os, const lbf_report_out& r, bool scalarsStable, double tol);
    double kPar(size_t d) const;

    const Params& p_;
    bool color_ = false;
    size_t NP_ = 1, D_ = 0, Q_ = 0, N_ = 0;
    size_t outRanks_ = 1;   // ranks of the reference run whose VTI layout is written
    std::vector<size_t> sz_;
    std::vector<uint8_t> geometry_;
    lbf_ctx* ctx_ = nullptr;
    double fluid_ = 0.0;
    std::vector<double> rho0_, rho1_, u_;
    size_t fieldsAt_ = SIZE_MAX;
    bool hasReportVelocity_ = false;
    size_t stableReports_ = 0;
    double KOld_ = 0.0;
    double SOld_[2] = {0.0, 0.0};
};

double Simulation::kPar(size_t d) const
{
    // Simulator::getKPar (simulator.cpp:151-160)
    const double viscosity = (p_.relaxationTime()[0] - 0.5) * p_.latticeSpeed() / 3;
    const double len = double(p_.reportSlice().back() - p_.reportSlice().front());
    double area = 1;
    for (size_t k = 0; k < D_; ++k)
        if (k != d) area *= sz_[k];
    return len * viscosity / area * std::pow(p_.latticeLen(), 2) * 1.0e12;
}

bool Simulation::updateConvergence(std::ostream& os, const lbf_report_out& r, bool scalarsStable, double tol)
{
    // Simulator::updateConvergence (simulator.cpp:248-325)
    if (!r.has_previous) {
        hasReportVelocity_ = true;
        stableReports_ = 0;
        return false;
    }
    if (r.nonfinite) throw std::runtime_error("Non-finite velocity in convergence report.");
    const double g0 = r.conv[0], g1 = r.conv[1], g2 = r.conv[2], g3 = r.conv[3];
    if (!std::isfinite(g0) || !std::isfinite(g1) || !std::isfinite(g3) || g2 <= 0.0)
        throw std::runtime_error("Invalid global velocity residual.");
    const double rms = std::sqrt(g0 / g2);
    const double residual = g1 > 0.0 ? std::sqrt(g0 / g1) : 0.0;
    const bool velocityStable = rms <= 1e-12 || residual < tol;
    const double adjRms = std::sqrt(g3 / g2);
    const double adjResidual = g1 > 0.0 ? std::sqrt(g3 / g1) : 0.0;
    const bool adjacentStable = adjRms <= 1e-12 || adjResidual < tol;
    const double pc0 = color_ ? r.phase_change[0] : 0.0, pc1 = color_ ? r.phase_change[1] : 0.0;
    stableReports_ =
        scalarsStable && velocityStable && adjacentStable && std::max(pc0, pc1) < tol ? stableReports_ + 1 : 0;
    os << " - velocityResidual = " << residual << ", velocityRmsChange = " << rms
       << ", adjacentVelocityResidual = " << adjResidual << ", adjacentVelocityRmsChange = " << adjRms
       << ", maxPhaseChange = " << pc0 << ", adjacentMaxPhaseChange = " << pc1
       << ", stableReports = " << stableReports_ << "/3" << std::endl;
    return stableReports_ >= 3;
}

bool Simulation::report(std::ostream& os)
{
    const size_t dir = p_.reportDir();
    if (dir == 0) return false;
    try {
        const size_t d = Params::axisOf(D_, dir);
        const auto& slices = p_.reportSlice();
        const auto fl = std::minmax_element(slices.begin(), slices.end());
        const double maxErr = p_.stopError();
        lbf_report_req req{};
        req.axis = int32_t(d);
        req.convergence = 1;
        req.adjacent = 1;
        req.first = *fl.first;
        req.last = *fl.second;
        lbf_report_out r{};
        if (lbf_report(ctx_, &req, &r)) throw std::runtime_error(lbf_last_error(ctx_));
        if (!hasReportVelocity_) r.has_previous = 0;
        const double Q = r.vel_range / double(*fl.second - *fl.first + 1);
        if (!color_) {
            // SCSP::report (scsp.cpp:304-372)
            const double dP = (r.rho_slice[0][0] / r.count_slice[0] - r.rho_slice[1][0] / r.count_slice[1]) / 3.0;
            if (!std::isfinite(dP) || !std::isfinite(Q)) throw std::runtime_error("Non-finite single-phase report state.");
            const bool drop = std::abs(dP) > 1e-14;
            const double K = drop ? kPar(d) * Q / dP : 0.0;
            const double errK = drop ? std::abs(K - KOld_) / std::max(std::abs(K), 1e-30) : 0.0;
            if (!std::isfinite(K) || !std::isfinite(errK)) throw std::runtime_error("Non-finite apparent permeability.");
            KOld_ = K;
            os << std::setprecision(6) << std::scientific << " - dP = " << dP << "(mu/lu/ts^2)," << " Q = " << Q
               << "(lu^3/ts)" << std::endl
               << " - K = " << K << "(um^2), " << std::fixed << "errK = " << errK * 100.0 << "%" << std::endl;
            if (!drop) os << " - pressure-based permeability unavailable: zero pressure drop" << std::endl;
            return updateConvergence(os, r, !drop || errK < maxErr, maxErr);
        }
        // ColorFluid::report (colorfluid.cpp:814-915)
        const double dens0[2] = {p_.density()[0], p_.density()[1]};
        const double pressureFactor = Q_ == 9 ? 3.0 / 5.0 : 1.0 / 2.0;
        const double a0 = p_.alpha0();
        const double a1 = 1 - p_.density()[0] / p_.density()[1] * (1 - a0);
        const double r2p[2] = {(1.0 - a0) * pressureFactor, (1.0 - a1) * pressureFactor};
        const double stc[2] = {r.rho_range[0] / dens0[0], r.rho_range[1] / dens0[1]};
        const double stcT = stc[0] + stc[1];
        const double S[2] = {stc[0] / stcT, stc[1] / stcT};
        const double errS[2] = {std::abs(S[0] - SOld_[0]), std::abs(S[1] - SOld_[1])};
        SOld_[0] = S[0];
        SOld_[1] = S[1];
        const double dP = (r.rho_slice[0][0] * r2p[0] + r.rho_slice[0][1] * r2p[1]) / r.count_slice[0] -
                          (r.rho_slice[1][0] * r2p[0] + r.rho_slice[1][1] * r2p[1]) / r.count_slice[1];
        if (!std::isfinite(S[0]) || !std::isfinite(S[1]) || !std::isfinite(errS[0]) || !std::isfinite(errS[1]) ||
            !std::isfinite(dP) || !std::isfinite(Q))
            throw std::runtime_error("Non-finite two-phase report state.");
        const bool drop = std::abs(dP) > 1e-14;
        const double K = drop ? kPar(d) * Q / dP : 0.0;
        const double errK = drop ? std::abs(K - KOld_) / std::max(std::abs(K), 1e-30) : 0.0;
        if (!std::isfinite(K) || !std::isfinite(errK)) throw std::runtime_error("Non-finite apparent permeability.");
        KOld_ = K;
        os << std::setprecision(6) << std::fixed << " - S0 = " << S[0] * 100 << "%, " << "dS0 = " << errS[0] * 100
           << "%" << std::endl
           << " - S1 = " << S[1] * 100 << "%, " << "dS1 = " << errS[1] * 100 << "%" << std::endl
           << std::scientific << " - dP = " << dP << "(mu/lu/ts^2), " << "Q = " << Q << "(lu^3/ts)" << std::endl
           << " - K = " << K << "(um^2), " << std::fixed << "errK = " << errK * 100.0 << "%" << std::endl;
        if (!drop) os << " - pressure-based permeability unavailable: zero pressure drop" << std::endl;
        const bool scalarsStable = std::max(errS[0], errS[1]) < maxErr && (!drop || errK < maxErr);
        return updateConvergence(os, r, scalarsStable, maxErr);
    } catch (std::exception& e) {
        os << e.what();
        throw;
    }
}

void Simulation::run(std::ostream& logProc)
{
    const double runStart = now();
    double seconds[TIMING_CATEGORIES] = {};
    size_t completed = 0;
    Logger procLog(logProc);
    procLog.reset(true);
    procLog.log("Initializing...", false);
    const size_t t0 = p_.startTime(), tEnd = p_.endTime();
    const size_t tStep = p_.writeInterval(), tRpt = p_.reportInterval(), tRst = p_.restartInterval();
    const bool reporting = p_.reportDir() != 0;
    if (t0 != 0) {
        const double b = now();
        loadRst(t0);
        seconds[CHECKPOINT_TIME] += now() - b;
    } else {
        double b = now();
        if (lbf_init(ctx_)) throw std::runtime_error(lbf_last_error(ctx_));
        seconds[INIT_TIME] += now() - b;
        failures(0, false);
        b = now();
        saveFields(0);
        seconds[FIELD_TIME] += now() - b;
        failures(0, true);
    }
    procLog.log("Done!\t");
    for (size_t t = t0 + 1; t <= tEnd; ++t) {
        bool converged = false;
        double b = now();
        const bool output = t % tStep == 0, rpt = t % tRpt == 0;
        uint32_t flags = (output || rpt) ? uint32_t(LBF_STEP_FIELDS) : 0u;
        // the velocity of this step is the adjacent history of a report at t + 1
        if (reporting && ((t + 1) % tStep == 0 || (t + 1) % tRpt == 0)) flags |= LBF_STEP_ADJACENT;
        if (lbf_step(ctx_, flags)) throw std::runtime_error(lbf_last_error(ctx_));
        seconds[NUMERICAL_TIME] += now() - b;
        if (output || rpt) {
            sync(seconds);
            failures(t, false);
            b = now();
            std::stringstream ss;
            ss << "Step " << t << " - ";
            procLog.log(ss.str());
            if (output) saveFields(t);
            converged = report(logProc);
            seconds[FIELD_TIME] += now() - b;
            failures(t, true);
        }
        if (converged || t % tRst == 0) {
            sync(seconds);
            failures(t, true);
            b = now();
            saveRst(t);
            seconds[CHECKPOINT_TIME] += now() - b;
        }
        ++completed;
        if (converged) break;
    }
    sync(seconds);
    failures(tEnd, true);
    // Simulator::writeTiming (simulator.cpp:40-71)
    std::ostringstream summary;
    summary << std::setprecision(17) << "LBFlowTiming {\"schema\":1,\"ranks\":" << 1
            << ",\"completed_steps\":" << completed << ",\"fluid_nodes\":" << fluid_ << ",\"total_voxels\":" << N_
            << ",\"initialization_seconds_max\":" << seconds[INIT_TIME]
            << ",\"numerical_seconds_max\":" << seconds[NUMERICAL_TIME]
            << ",\"fields_diagnostics_seconds_max\":" << seconds[FIELD_TIME]
            << ",\"checkpoint_seconds_max\":" << seconds[CHECKPOINT_TIME]
            << ",\"run_seconds_max\":" << now() - runStart << "}\n";
    logProc << summary.str();
}

}  // namespace lbsim

int main(int argc, char* argv[])
{
    using namespace lbsim;
    const int thisId = 0;
    try {
        if (argc < 2) {
            std::cerr << argv[0] << ": you must specify an input filename." << std::endl;
            return 1;
        }
        Logger logger(std::cout);
        logger.log("Model constructing...", false);
        Params input(argv[1]);
        const string model = input.fluidModel();
        if (model == "shanchen") {
            input.validate();   // the reference constructs the simulator (and validates) first
            throw std::invalid_argument("Main(): Fluid model shanchen currently not supported.");
        }
        if (model != "color" && model != "scsp")
            throw std::invalid_argument(
                "Main(): Unsupport fluid model (only color model and single components single phase model are "
                "supported at this time)");
        std::vector<std::string> flags(argv + 2, argv + argc);
        Simulation sim(input, flags);
        logger.log("Done!\n");
        logger.reset();
        logger.log("Simulation started!\n", false);
        sim.run(std::cout);
        logger.log("Simulation finished!\nTotal ");
    } catch (std::exception& e) {
        std::cout << thisId << " Failed to excute: " << e.what() << std::endl;
        std::cout.flush();
        return EXIT_FAILURE;
    }
    return 0;
}
