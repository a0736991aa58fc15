// C++ host face of the MI355X time loop: the reference's ForwardSolver
// interface (src/ForwardSolver/ForwardSolver.hpp:26-70, factory
// ForwardSolverFactory.cpp:4-66) over the C ABI of libwavehip.so, plus the
// pieces of the reference's input/output the time loop needs: the KEY=VALUE
// configuration file (Configuration.cpp:45-87), source / receiver text files
// (AcquisitionSettings.hpp:24-100, 238-270), MatrixMarket vectors and
// seismograms (IO.hpp:21-164, Seismogram.cpp:82-156), the source signals
// (Sources.cpp:99-220, 231-249: synthetic shapes 1..9 or rows of a signal file)
// and the seismogram resampling (Common.hpp:202-233).
// Errors throw std::runtime_error, as the reference throws.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wave_abi.h"

namespace wavesim {

class Configuration {
  public:
    Configuration() = default;
    explicit Configuration(const std::string &filename) { readFromFile(filename, true); }
    void readFromFile(const std::string &filename, bool overwrite);
    void add(const std::string &key, const std::string &value, bool overwrite = true);
    bool has(const std::string &key) const;
    std::string getString(const std::string &key) const;
    double getDouble(const std::string &key) const;
    long getInt(const std::string &key) const;
    double getDouble(const std::string &key, double fallback) const { return has(key) ? getDouble(key) : fallback; }
    long getInt(const std::string &key, long fallback) const { return has(key) ? getInt(key) : fallback; }

  private:
    std::map<std::string, std::string> map_;
};

struct SourceRow {
    int shot, x, y, z, type, wavelet, shape;
    double fc, amp, tshift;
    long row;  // data line of the source file, from 0 (AcquisitionSettings.hpp:68-102)
};
struct ReceiverRow {
    int x, y, z, type;
};

std::vector<SourceRow> readSources(const std::string &path);
std::vector<ReceiverRow> readReceivers(const std::string &path);
// synthetic signal of wavelet shape 1..9 (Sources.cpp:157-220) from the library
std::vector<float> sourceSignal(int shape, int NT, double DT, double fc, double amp, double tshift);
std::vector<float> ricker(int NT, double DT, double fc, double amp, double tshift);
// row `row` of a MatrixMarket matrix with `cols` columns, "array" or "coordinate"
// (IO.hpp:172-195, DenseStorage::readFromFile of one row)
std::vector<float> readMatrixRowMtx(const std::string &path, long row, size_t cols);

// MatrixMarket "array real general" (values column-major)
std::vector<float> readVectorMtx(const std::string &path, size_t expected);
void writeMatrixMtx(const std::string &path, const std::vector<float> &rowMajor, int rows, int cols);
// linear resampling of traces [ntr][NT] from DT to seismoDT
std::vector<float> resampleTraces(const std::vector<float> &data, int ntr, int NT, double ratio, int &NTout);

// One solver handle: a grid, a model and nshots acquisitions on the GPU.
class ForwardSolver {
  public:
    static std::unique_ptr<ForwardSolver> create(const std::string &dimension, const std::string &type);
    ~ForwardSolver();
    ForwardSolver(const ForwardSolver &) = delete;
    ForwardSolver &operator=(const ForwardSolver &) = delete;

    // initForwardSolver: grid, DT, derivative order, boundaries from the configuration
    void initForwardSolver(const Configuration &config, int nshots, int stepsPerGraph = 50, int tileY = 0);
    // prepareForModelling: model vectors in grid order (x fastest, then z, then depth y)
    void prepareForModelling(const std::vector<float> &m0, const std::vector<float> &m1, const std::vector<float> &m2);
    // viscoelastic: relaxation frequencies (1..4 mechanisms) and tauP, tauS per cell, before
    // prepareForModelling (Modelparameter::Viscoelastic, Viscoelastic.cpp:160-260)
    void setRelaxation(const std::vector<double> &freqs, const std::vector<float> &tauP, const std::vector<float> &tauS);
    void setAcquisition(int shot, const std::vector<long long> &srcIndex, const std::vector<int> &srcType,
                        const std::vector<float> &srcSignal, const std::vector<long long> &recIndex,
                        const std::vector<int> &recType);
    void resetCPML();
    void resetWavefields() { resetCPML(); }
    void run(int t) { run(t, t + 1); }
    void run(int tStart, int tEnd);
    void synchronize();
    // Simulation.cpp:519: throws if a wavefield value or seismogram sample is Inf or NaN
    void checkFinite();
    const wv_config &config() const { return cfg_; }
    double lastRunMs();
    // time every kernel variant for nsteps steps, keep the fastest (resets the wavefields)
    int tune(int nsteps);
    std::vector<float> seismograms(int shot, int nrec);  // [nrec][NT]
    int NT() const { return cfg_.NT; }
    int NX() const { return cfg_.NX; }
    int NY() const { return cfg_.NY; }
    int NZ() const { return cfg_.NZ; }
    wv_solver *handle() { return h_; }

  private:
    ForwardSolver() = default;
    wv_solver *h_ = nullptr;
    wv_config cfg_{};
};

void check(int rc);

// CheckParameter::checkNumericalArtefactsAndInstabilities (CheckParameter.hpp:166-246):
// throws when the CFL criterion or vp/vs >= sqrt(2) fails; returns the dispersion
// warning ("" when the criterion holds), which the reference only prints
std::string checkNumericalArtefactsAndInstabilities(const std::string &dimension, const std::string &type,
                                                    const wv_config &cfg, const std::vector<float> &m0,
                                                    const std::vector<float> &m1, const std::vector<float> &m2,
                                                    double fcMax);

}  // namespace wavesim
