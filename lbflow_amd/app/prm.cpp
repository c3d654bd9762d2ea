// prm.cpp -- input.prm reader and validator (see prm.h).
#include "prm.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>

namespace lbsim {

namespace {

// Value parsers with the reference's stream semantics (fileprm.cpp:491-561):
// leading whitespace skipped, trailing ';' left unread, quoted strings read
// between the first pair of double quotes.
bool read(const std::string& s, std::size_t* v)
{
    std::istringstream is{s};
    is >> *v;
    return !is.fail();
}
bool read(const std::string& s, double* v)
{
    std::istringstream is{s};
    is >> *v;
    return !is.fail();
}
bool read(const std::string& s, std::string* v)
{
    v->clear();
    std::istringstream is{s};
    char c;
    while (is.get(c) && c != '"') {
    }
    while (is.get(c) && c != '"') v->push_back(c);
    return !is.fail();
}
bool read(const std::string& s, std::vector<std::size_t>* v)
{
    v->clear();
    std::istringstream is{s};
    while (is) {
        int i;
        is >> i;
        if (is.fail()) break;
        v->push_back(static_cast<std::size_t>(i));
    }
    return !v->empty();
}
bool read(const std::string& s, std::vector<double>* v)
{
    v->clear();
    std::istringstream is{s};
    while (is) {
        double d;
        is >> d;
        if (is.fail()) break;
        v->push_back(d);
    }
    return !v->empty();
}
bool read(const std::string& s, std::vector<std::string>* v)
{
    v->clear();
    std::istringstream is{s};
    while (is) {
        std::string item;
        char c;
        while (is.get(c) && c != '"') {
        }
        while (is.get(c) && c != '"') item.push_back(c);
        if (is.fail()) break;
        v->push_back(item);
    }
    return !v->empty();
}

// Optional value: kept only when it parses (FilePrm::parse(str, T**)).
template <typename T> void optional(const std::string& s, std::unique_ptr<T>& v)
{
    std::unique_ptr<T> t(new T());
    if (read(s, t.get())) v = std::move(t);
    else v.reset();
}

double number(const std::string& s, bool must)
{
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) {
        if (must) throw std::invalid_argument("IdxHead::str2d() - double value not defined");
        return 0;
    }
    return v;
}

std::size_t unsignedNumber(const std::string& s, bool must)
{
    char* end = nullptr;
    const std::size_t v = std::strtoul(s.c_str(), &end, 10);
    if (end == s.c_str()) {
        if (must) throw std::invalid_argument("IdxHead::str2ul() - unsigned long value not defined");
        return 0;
    }
    return v;
}

// One label definition (FilePrm::IDXVAL, fileprm.cpp:633-670).
Label makeLabel(const std::vector<std::string>& v)
{
    if (v.size() < 4) throw std::invalid_argument("FilePrm::IDX::IDX() - insufficient val size.");
    Label l;
    const std::string& t = v[0];
    if (t == "VD") {
        l.type = VD;
        l.phase = unsignedNumber(v[3], true);
    } else if (t == "GV" || t == "PB" || t == "VB") {
        l.type = t == "GV" ? GV : (t == "PB" ? PB : VB);
        l.value = number(v[1], true);
        l.dir = unsignedNumber(v[2], true);
        l.phase = unsignedNumber(v[3], true);
    } else if (t == "AB") {
        l.type = AB;
    } else if (t == "BB") {
        l.type = BB;
    } else if (t == "SD") {
        l.type = SD;
        l.value = number(v[1], true);
    } else if (t == "HF") {
        l.type = HF;
        l.phase = unsignedNumber(v[3], true);
    } else {
        throw std::invalid_argument("FilePrm::IDXVAL::IDXVAL() - unrecognized index descriptor type");
    }
    return l;
}

}  // namespace

Params::Params(const std::string& filename)
{
    {
        std::FILE* f = std::fopen(filename.c_str(), "r");
        if (!f) throw std::runtime_error("FilePrm::Load() - Input file does not exist.");
        std::fclose(f);
    }
    std::ifstream in(filename);
    if (!in.is_open()) throw std::runtime_error("FilePrm::Load() - Failed to load input file");
    for (std::string key; in >> key;) {
        std::string rest;
        std::getline(in, rest);
        if (std::isalnum(static_cast<unsigned char>(key[0]))) raw_[key] = rest;
    }
    // keys are visited in sorted order, as the reference's std::map
    for (const auto& kv : raw_) {
        const std::string& k = kv.first;
        const std::string& v = kv.second;
        if (k == "gridFile") optional(v, gridFile_);
        else if (k == "dimension") read(v, &dimension_);
        else if (k == "directions") optional(v, directions_);
        else if (k == "fileType") optional(v, fileType_);
        else if (k == "latticeLen") optional(v, latticeLen_);
        else if (k == "startTime") optional(v, tStart_);
        else if (k == "endTime") optional(v, tEnd_);
        else if (k == "writeInterval") optional(v, tStep_);
        else if (k == "reportInterval") optional(v, tRpt_);
        else if (k == "restartInterval") optional(v, tRst_);
        else if (k == "stopError") optional(v, stopError_);
        else if (k == "outPath") optional(v, outPath_);
        else if (k == "logPath") optional(v, logPath_);
        else if (k == "reportDir") optional(v, reportDir_);
        else if (k == "reportSlice") read(v, &reportSlice_);
        else if (k == "fluidModel") optional(v, fluidModel_);
        else if (k == "equilModel") optional(v, equilModel_);
        else if (k == "rexlationTime") read(v, &relaxationTime_);
        else if (k == "latticeSpeed") optional(v, latticeSpeed_);
        else if (k == "alpha0") optional(v, alpha0_);
        else if (k == "density") read(v, &density_);
        else if (k == "phase") read(v, &phase_);
        else if (k == "contactAngle") optional(v, cAngle_);
        else if (k == "surfaceTension") optional(v, sTension_);
        else if (k == "interfaceThickness") optional(v, iThickness_);
        else if (k == "components") read(v, &components_);
        else {
            // numeric keys define geometry labels; other keys are selectors read on demand
            char* end = nullptr;
            const std::size_t value = std::strtoul(k.c_str(), &end, 10);
            if (end == k.c_str()) continue;
            std::vector<std::string> fields;
            if (read(v, &fields)) labels_[value] = makeLabel(fields);
        }
    }
}

std::string Params::option(const char* key, const char* fallback) const
{
    const auto it = raw_.find(key);
    if (it == raw_.end()) return fallback;
    std::string out;
    if (!read(it->second, &out)) throw std::invalid_argument(std::string(key) + " must be a quoted string.");
    return out;
}

double Params::wallSlipLength() const
{
    const auto it = raw_.find("wallSlipLength");
    if (it == raw_.end()) return 0.0;
    double v;
    if (!read(it->second, &v) || !std::isfinite(v) || v < 0.0)
        throw std::invalid_argument("wallSlipLength must be a finite nonnegative number.");
    return v;
}

const Label& Params::label(std::size_t value) const { return labels_.at(value); }

std::set<std::size_t> Params::labelsOf(int mask) const
{
    std::set<std::size_t> out;
    for (const auto& kv : labels_)
        if (kv.second.type & mask) out.insert(kv.first);
    return out;
}

std::size_t Params::axisOf(std::size_t D, std::size_t dir)
{
    // qa.axis of gridspace.cpp:95 (D2Q9) and :142 (D3Q19)
    static const std::size_t a2[5] = {0, 1, 0, 1, 0};
    static const std::size_t a3[7] = {0, 2, 2, 1, 1, 0, 0};
    return D == 2 ? a2[dir] : a3[dir];
}

void Params::validate() const
{
    const std::size_t ndim = dimension_.size();
    const std::size_t q = directions();
    if (!((ndim == 2 && q == 9) || (ndim == 3 && q == 19)))
        throw std::invalid_argument("Only D2Q9 and D3Q19 are validated for simulation.");
    std::size_t count = 1;
    for (std::size_t n : dimension_) {
        if (n == 0 || n > std::numeric_limits<std::size_t>::max() / count)
            throw std::invalid_argument("Invalid grid dimensions.");
        count *= n;
    }
    if (fileType() != "uint8") throw std::invalid_argument("Simulation geometry must use uint8 labels.");
    if (endTime() < startTime() || endTime() == std::numeric_limits<std::size_t>::max() || writeInterval() == 0 ||
        reportInterval() == 0 || restartInterval() == 0)
        throw std::invalid_argument("Invalid time range or zero output/report/restart interval.");
    if (!std::isfinite(stopError()) || stopError() <= 0.0 || !std::isfinite(latticeLen()) || latticeLen() <= 0.0)
        throw std::invalid_argument("Invalid convergence tolerance or lattice length.");
    if (latticeSpeed() != 1.0) throw std::invalid_argument("This research implementation requires latticeSpeed=1.");
    const std::string& model = fluidModel();
    if (model != "scsp" && model != "color")
        throw std::invalid_argument("Only scsp and color fluid models are implemented.");
    if (equilModel() != "SRT" && equilModel() != "MRT")
        throw std::invalid_argument("Equilibrium model must be SRT or MRT.");
    const std::size_t phases = model == "color" ? 2 : 1;
    if (relaxationTime_.size() != phases || density_.size() != phases)
        throw std::invalid_argument("Density and relaxation-time counts must match the fluid model.");
    for (std::size_t k = 0; k < phases; ++k)
        if (!std::isfinite(relaxationTime_[k]) || relaxationTime_[k] <= 0.5 || !std::isfinite(density_[k]) ||
            density_[k] <= 0.0)
            throw std::invalid_argument("Densities must be positive; relaxation times must exceed 0.5.");
    const std::string recoloring = recoloringModel(), spectrum = mrtSpectrum();
    const bool equalDensity = phases == 2 && density_[0] == density_[1];
    if (wallSlipLength() > 0.0) {
        if (model != "color" || ndim != 2 || equilModel() != "MRT" || spectrum != "legacy" ||
            recoloring != "projection" || !equalDensity || relaxationTime_[0] != 1.0 || relaxationTime_[1] != 1.0 ||
            capillaryModel() != "csf" || wallGradient() != "ghost" || phaseBoundary() != "flux")
            throw std::invalid_argument(
                "Prescribed slip initially requires neutral ghost/flux projection/legacy-MRT color D2Q9 at equal tau1.");
        for (const auto& kv : labels_) {
            const int t = kv.second.type;
            if (t != VD && t != GV && t != SD && t != BB)
                throw std::invalid_argument("Prescribed slip initially supports closed VD/GV with flat SD/BB walls.");
        }
    }
    if (recoloring != "cosine" && recoloring != "projection")
        throw std::invalid_argument("recoloringModel must be cosine or projection.");
    if (spectrum != "legacy" && spectrum != "liu2015")
        throw std::invalid_argument("mrtSpectrum must be legacy or liu2015.");
    if (recoloring != "cosine" || spectrum != "legacy") {
        if (model != "color" || ndim != 2 || !equalDensity || std::abs(alpha0() - 4.0 / 9.0) > 1e-14)
            throw std::invalid_argument("Published core controls require equal-density standard-weight color D2Q9.");
        if (spectrum == "liu2015" && equilModel() != "MRT")
            throw std::invalid_argument("The published MRT spectrum requires equilModel MRT.");
        if (spectrum == "liu2015" && capillaryModel() != "csf")
            throw std::invalid_argument(
                "The published MRT spectrum requires capillaryModel csf; its perturbation tension mapping is not supported.");
        if (recoloring == "projection" && wallRecoloring() != "legacy")
            throw std::invalid_argument("Projection recoloring cannot also use the cosine-derived wall multiplier.");
    }
    if (model == "color") {
        const bool pressureCsf = capillaryModel() == "csf" && !labelsOf(PB).empty();
        if (pressureCsf && (ndim != 2 || wallGradient() != "ghost" || phaseBoundary() != "flux"))
            throw std::invalid_argument("Pressure CSF requires flat D2Q9 ghost walls and phaseBoundary flux.");
        const std::string boundary = phaseBoundary();
        if (boundary != "bounce" && boundary != "flux")
            throw std::invalid_argument("phaseBoundary must be bounce or flux.");
        const std::string wg = wallGradient();
        if (boundary == "flux" || wg == "tangent" || wg == "halfway" || wg == "ghost") {
            if (ndim != 2 || latticeSpeed() != 1.0 || !equalDensity || std::abs(alpha0() - 4.0 / 9.0) > 1e-14)
                throw std::invalid_argument(
                    "Experimental wall phase/gradient operators require equal-density, standard-weight, unit-speed D2Q9.");
            if (boundary == "flux" && wallRecoloring() != "legacy")
                throw std::invalid_argument(
                    "Wall phase flux requires legacy recoloring; do not also apply the tangent multiplier.");
            if (boundary == "flux" && wg != "renormalized" && wg != "tangent" && wg != "halfway" && wg != "ghost")
                throw std::invalid_argument(
                    "Wall phase flux requires renormalized, tangent, halfway or ghost wall gradients.");
            for (const auto& kv : labels_) {
                const int t = kv.second.type;
                if (t != VD && t != GV && t != SD && t != BB && !(pressureCsf && t == PB))
                    throw std::invalid_argument(
                        "Wall operators require closed flow or guarded ghost/flux pressure CSF.");
            }
        }
        const std::string wr = wallRecoloring();
        if (wr != "legacy" && wr != "tangential")
            throw std::invalid_argument("wallRecoloring must be legacy or tangential.");
        if (wr == "tangential") {
            if (wg != "renormalized")
                throw std::invalid_argument("Tangential wall recoloring requires wallGradient renormalized.");
            if (ndim != 2 || !equalDensity || std::abs(alpha0() - 4.0 / 9.0) > 1e-14)
                throw std::invalid_argument(
                    "Tangential wall recoloring requires equal-density standard-weight D2Q9.");
            for (const auto& kv : labels_) {
                const int t = kv.second.type;
                if (t != VD && t != GV && t != SD && t != BB)
                    throw std::invalid_argument(
                        "Tangential wall recoloring initially supports only closed VD/GV flow with SD/BB walls.");
            }
        }
        const std::string capillary = capillaryModel();
        const std::string curvature = curvatureStencil();
        if (curvature != "normal" && curvature != "gradient")
            throw std::invalid_argument("curvatureStencil must be normal or gradient.");
        if (curvature == "gradient" && capillary != "csf")
            throw std::invalid_argument("The gradient curvature stencil requires capillaryModel csf.");
        if (capillary != "perturbation" && capillary != "csf")
            throw std::invalid_argument("capillaryModel must be perturbation or csf.");
        if (capillary == "csf") {
            if (latticeSpeed() != 1.0)
                throw std::invalid_argument("Experimental CSF requires latticeSpeed=1 for Guo forcing.");
            const double standard = ndim == 2 ? 4.0 / 9.0 : 1.0 / 3.0;
            if (!equalDensity || std::abs(alpha0() - standard) > 1e-14)
                throw std::invalid_argument(
                    "Experimental CSF requires equal reference densities and standard pressure weights.");
            for (const auto& kv : labels_) {
                const int t = kv.second.type;
                if (t != VD && t != GV && t != SD && t != BB && !(pressureCsf && t == PB))
                    throw std::invalid_argument(
                        "CSF requires closed flow or guarded flat ghost/flux pressure reservoirs.");
            }
        }
        const std::string visc = viscosityInterpolation();
        if (visc != "arithmetic" && visc != "harmonic")
            throw std::invalid_argument("viscosityInterpolation must be arithmetic or harmonic.");
        if (wg != "legacy" && wg != "renormalized" && wg != "tanh" && wg != "tangent" && wg != "halfway" &&
            wg != "ghost")
            throw std::invalid_argument("wallGradient must be legacy, renormalized, tanh, tangent, halfway or ghost.");
        const double a0 = alpha0();
        const double a1 = 1.0 - density_[0] / density_[1] * (1.0 - a0);
        if (!std::isfinite(a0) || a0 < 0.0 || a0 >= 1.0 || !std::isfinite(a1) || a1 < 0.0 || a1 >= 1.0 ||
            !std::isfinite(surfaceTension()) || surfaceTension() < 0.0 || !std::isfinite(interfaceThickness()) ||
            interfaceThickness() < 0.0 || interfaceThickness() > 1.0)
            throw std::invalid_argument("Invalid phase weights, surface tension, or recoloring beta.");
    }
    if (reportDir() >= q) throw std::invalid_argument("Report direction is outside the lattice.");
    if (reportDir() != 0) {
        if (reportSlice_.size() < 2) throw std::invalid_argument("At least two report slices are required.");
        if (*std::max_element(reportSlice_.begin(), reportSlice_.end()) >=
            *std::max_element(dimension_.begin(), dimension_.end()))
            throw std::invalid_argument("Report slice is outside the grid.");
    }
    for (const auto& kv : labels_) {
        if (kv.first > 255) throw std::invalid_argument("Geometry labels must fit in uint8.");
        const Label& v = kv.second;
        if (!std::isfinite(v.value)) throw std::invalid_argument("Non-finite boundary value.");
        if ((v.type == VD || v.type == GV || v.type == PB || v.type == VB || v.type == HF) && v.phase >= phases)
            throw std::invalid_argument("Boundary phase is outside the configured fluid model.");
        if ((v.type == GV || v.type == PB || v.type == VB) && (v.dir == 0 || v.dir >= q))
            throw std::invalid_argument("Boundary direction must be a non-rest lattice direction.");
        if (v.type == PB && v.value <= 0.0) throw std::invalid_argument("Pressure-boundary density must be positive.");
        if ((v.type == PB || v.type == VB) && v.dir > 2 * ndim)
            throw std::invalid_argument("Pressure and velocity boundaries require an axial direction.");
        if (v.type == SD && (v.value < 0.0 || v.value > 180.0))
            throw std::invalid_argument("Solid contact angle must be between 0 and 180 degrees.");
        if (v.type == AB) throw std::invalid_argument("Axisymmetric AB boundary handling is not implemented.");
    }
}

}  // namespace lbsim
