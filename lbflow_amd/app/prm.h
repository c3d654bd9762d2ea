// prm.h -- the input.prm contract of LBFlow.
//
// Same grammar, getters, defaults and validation messages as the
// reference's FilePrm (solver/src/file/fileprm.cpp:20-688,
// solver/inc/file/fileprm.h): one `key value;` entry per line, keys must start
// with an alphanumeric character, strings are quoted, arrays are
// space-separated, numeric keys are geometry labels
// `<uint8> "TYPE" "value" "dir" "phase"`, dimensions are given slowest axis
// first.  Messages are the reference's because callers (harness scripts,
// tests) match them.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace lbsim {

enum LabelType { VD = 0x1, GV = 0x2, PB = 0x4, VB = 0x8, AB = 0x10, BB = 0x20, HF = 0x40, SD = 0x80 };

struct Label {
    int type = VD;
    double value = 0.0;
    std::size_t dir = 0;
    std::size_t phase = 0;
};

class Params {
public:
    explicit Params(const std::string& filename);

    // FilePrm::ValidateSimulation (fileprm.cpp:136-295).
    void validate() const;

    const std::string& gridFile() const { return need(gridFile_); }
    const std::vector<std::size_t>& dimension() const { return dimension_; }
    std::size_t directions() const { return need(directions_); }
    const std::string& fileType() const { return need(fileType_); }
    double latticeLen() const { return need(latticeLen_); }
    std::size_t startTime() const { return need(tStart_); }
    std::size_t endTime() const { return need(tEnd_); }
    std::size_t writeInterval() const { return need(tStep_); }
    std::size_t reportInterval() const { return need(tRpt_); }
    std::size_t restartInterval() const { return need(tRst_); }
    double stopError() const { return need(stopError_); }
    const std::string& outPath() const { return need(outPath_); }
    std::size_t reportDir() const { return need(reportDir_); }
    const std::vector<std::size_t>& reportSlice() const { return reportSlice_; }
    const std::string& fluidModel() const { return need(fluidModel_); }
    const std::string& equilModel() const { return need(equilModel_); }
    const std::vector<double>& relaxationTime() const { return relaxationTime_; }
    double latticeSpeed() const { return need(latticeSpeed_); }
    double alpha0() const { return need(alpha0_); }
    const std::vector<double>& density() const { return density_; }
    double surfaceTension() const { return need(sTension_); }
    double interfaceThickness() const { return need(iThickness_); }

    // optional selectors with defaults (fileprm.cpp:46-134)
    std::string wallGradient() const { return option("wallGradient", "legacy"); }
    std::string viscosityInterpolation() const { return option("viscosityInterpolation", "arithmetic"); }
    std::string capillaryModel() const { return option("capillaryModel", "perturbation"); }
    std::string wallRecoloring() const { return option("wallRecoloring", "legacy"); }
    std::string curvatureStencil() const { return option("curvatureStencil", "normal"); }
    std::string phaseBoundary() const { return option("phaseBoundary", "bounce"); }
    std::string recoloringModel() const { return option("recoloringModel", "cosine"); }
    std::string mrtSpectrum() const { return option("mrtSpectrum", "legacy"); }
    double wallSlipLength() const;
    // engine selectors; the reference ignores unknown keys, so the same file runs on both
    std::string enginePrecision() const { return option("enginePrecision", "fp64"); }
    std::string engineLayout() const { return option("engineLayout", "exact"); }
    // VTI files laid out as an R-rank reference run writes them (one piece per rank)
    std::string outputRanks() const { return option("outputRanks", "1"); }

    const Label& label(std::size_t value) const;            // throws std::out_of_range "map::at"
    const std::map<std::size_t, Label>& labels() const { return labels_; }
    std::set<std::size_t> labelsOf(int typeMask) const;

    static std::size_t axisOf(std::size_t D, std::size_t dir);   // GridSpace::Qaxis

private:
    template <typename T> static const T& need(const std::unique_ptr<T>& v)
    {
        if (!v) throw std::runtime_error("the value may not be defined in parameter file.");
        return *v;
    }
    std::string option(const char* key, const char* fallback) const;

    std::map<std::string, std::string> raw_;
    std::unique_ptr<std::string> gridFile_, fileType_, outPath_, logPath_, fluidModel_, equilModel_;
    std::unique_ptr<std::size_t> directions_, tStart_, tEnd_, tStep_, tRpt_, tRst_, reportDir_;
    std::unique_ptr<double> latticeLen_, stopError_, latticeSpeed_, alpha0_, cAngle_, sTension_, iThickness_;
    std::vector<std::size_t> dimension_, reportSlice_;
    std::vector<double> relaxationTime_, density_;
    std::vector<std::string> phase_, components_;
    std::map<std::size_t, Label> labels_;
};

}  // namespace lbsim
