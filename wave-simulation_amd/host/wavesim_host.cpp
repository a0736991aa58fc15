// C++ host face of the MI355X time loop (declarations and references in wavesim.hpp).
#include "wavesim.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace wavesim {

void check(int rc)
{
    if (rc != 0) throw std::runtime_error(std::string("wavesim: ") + wv_last_error());
}

// ---- configuration (Configuration.cpp:15-87): lines KEY=VALUE, '#' starts a
// comment, keys are lower-cased, values are read with >> (first token).
void Configuration::add(const std::string &key, const std::string &value, bool overwrite)
{
    if (!map_.count(key) || overwrite) map_[key] = value;
}

void Configuration::readFromFile(const std::string &filename, bool overwrite)
{
    std::ifstream in(filename);
    if (!in.good()) throw std::runtime_error("Configuration file " + filename + " was not found");
    std::string line;
    while (std::getline(in, line)) {
        size_t end = line.size();
        const size_t hash = line.find('#');
        if (hash != std::string::npos) {
            if (hash == 0) continue;
            end = hash;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos || eq >= end) continue;
        std::string name = line.substr(0, eq);
        for (auto &c : name) c = (char)std::tolower((unsigned char)c);
        add(name, line.substr(eq + 1, end - eq - 1), overwrite);
    }
}

static std::string lowerKey(const std::string &k)
{
    std::string r = k;
    for (auto &c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}

bool Configuration::has(const std::string &key) const { return map_.count(lowerKey(key)) != 0; }

std::string Configuration::getString(const std::string &key) const
{
    auto it = map_.find(lowerKey(key));
    if (it == map_.end()) throw std::runtime_error("Parameter " + key + " not found in the configuration");
    std::istringstream is(it->second);
    std::string v;
    is >> v;
    return v;
}

double Configuration::getDouble(const std::string &key) const
{
    std::istringstream is(getString(key));
    double v;
    if (!(is >> v)) throw std::runtime_error("Parameter " + key + " is not a number");
    return v;
}

long Configuration::getInt(const std::string &key) const
{
    std::istringstream is(getString(key));
    long v;
    if (!(is >> v)) throw std::runtime_error("Parameter " + key + " is not an integer");
    return v;
}

// ---- acquisition files (AcquisitionSettings.hpp:55-100, 238-270)
static std::vector<std::vector<std::string>> readRows(const std::string &path, size_t minCols)
{
    std::ifstream in(path);
    if (!in.good()) throw std::runtime_error("cannot open " + path);
    std::vector<std::vector<std::string>> rows;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream is(line);
        std::vector<std::string> t;
        std::string w;
        while (is >> w) t.push_back(w);
        if (t.empty() || t[0][0] == '#') continue;
        if (t.size() < minCols) throw std::runtime_error("short line in " + path);
        rows.push_back(t);
    }
    return rows;
}

std::vector<SourceRow> readSources(const std::string &path)
{
    std::vector<SourceRow> out;
    for (auto &t : readRows(path, 10))
        out.push_back({std::stoi(t[0]), std::stoi(t[1]), std::stoi(t[2]), std::stoi(t[3]), std::stoi(t[4]),
                       std::stoi(t[5]), std::stoi(t[6]), std::stod(t[7]), std::stod(t[8]), std::stod(t[9]),
                       (long)out.size()});
    return out;
}

std::vector<ReceiverRow> readReceivers(const std::string &path)
{
    std::vector<ReceiverRow> out;
    for (auto &t : readRows(path, 4)) out.push_back({std::stoi(t[0]), std::stoi(t[1]), std::stoi(t[2]), std::stoi(t[3])});
    return out;
}

std::vector<float> sourceSignal(int shape, int NT, double DT, double fc, double amp, double tshift)
{
    std::vector<float> w((size_t)std::max(NT, 1));
    check(wv_source_signal(shape, NT, DT, fc, amp, tshift, w.data()));
    return w;
}

std::vector<float> ricker(int NT, double DT, double fc, double amp, double tshift)
{
    return sourceSignal(1, NT, DT, fc, amp, tshift);
}

// ---- MatrixMarket (IO.hpp:21-164)
std::vector<float> readVectorMtx(const std::string &path, size_t expected)
{
    std::ifstream in(path);
    if (!in.good()) throw std::runtime_error("cannot open " + path);
    std::string line;
    bool header = false;
    size_t rows = 0, cols = 0;
    std::vector<float> v;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream is(line);
        if (!header) {
            is >> rows >> cols;
            header = true;
            v.reserve(rows * std::max<size_t>(cols, 1));
            continue;
        }
        double x;
        while (is >> x) v.push_back((float)x);
    }
    if (v.size() != expected) throw std::runtime_error(path + ": expected " + std::to_string(expected) + " values");
    return v;
}

std::vector<float> readMatrixRowMtx(const std::string &path, long row, size_t cols)
{
    std::ifstream in(path);
    if (!in.good()) throw std::runtime_error("cannot open " + path);
    std::string line;
    bool banner = false, coordinate = false, header = false;
    long R = 0, C = 0, nnz = 0;
    size_t k = 0;
    std::vector<float> out;
    while (std::getline(in, line)) {
        if (!banner && line.rfind("%%MatrixMarket", 0) == 0) {
            banner = true;
            coordinate = line.find("coordinate") != std::string::npos;
            continue;
        }
        if (line.empty() || line[0] == '%') continue;
        std::istringstream is(line);
        if (!header) {
            is >> R >> C;
            if (coordinate) is >> nnz;
            if (!is || R <= 0 || C <= 0) throw std::runtime_error(path + ": bad MatrixMarket size line");
            if (row < 0 || row >= R)
                throw std::runtime_error(path + ": row " + std::to_string(row) + " of a matrix with " + std::to_string(R) + " rows");
            if ((size_t)C != cols) throw std::runtime_error(path + ": Source signal has invalid length");
            out.assign(cols, 0.f);
            header = true;
            continue;
        }
        if (coordinate) {
            long i, j;
            double v;
            if (is >> i >> j >> v) {
                if (i < 1 || i > R || j < 1 || j > C) throw std::runtime_error(path + ": entry out of range");
                if (i - 1 == row) out[(size_t)(j - 1)] = (float)v;
            }
        } else {
            double v;
            while (is >> v) {
                if ((long)(k % (size_t)R) == row && k / (size_t)R < cols) out[k / (size_t)R] = (float)v;
                k++;
            }
        }
    }
    if (!header) throw std::runtime_error(path + ": no MatrixMarket size line");
    if (!coordinate && k != (size_t)R * (size_t)C)
        throw std::runtime_error(path + ": expected " + std::to_string((size_t)R * C) + " values, read " + std::to_string(k));
    return out;
}

void writeMatrixMtx(const std::string &path, const std::vector<float> &rowMajor, int rows, int cols)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d %d\n", rows, cols);
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) std::fprintf(f, "%.9g\n", (double)rowMajor[(size_t)r * cols + c]);
    std::fclose(f);
}

// Common.hpp:202-233: sample k of the output lies at k*ratio input samples
std::vector<float> resampleTraces(const std::vector<float> &data, int ntr, int NT, double ratio, int &NTout)
{
    const float c = (float)ratio;
    NTout = (int)std::floor((float)(NT - 1) / c) + 1;
    std::vector<float> out((size_t)ntr * NTout, 0.f);
    for (int k = 0; k < NTout; k++) {
        const float rel = (float)k * c;
        const int i0 = (int)std::floor(rel);
        const float fr = std::fmod(rel, 1.f);
        for (int r = 0; r < ntr; r++) {
            float v = 0.f;
            if (i0 < NT) v += (1.f - fr) * data[(size_t)r * NT + i0];
            if (i0 + 1 < NT) v += fr * data[(size_t)r * NT + i0 + 1];
            out[(size_t)r * NTout + k] = v;
        }
    }
    return out;
}

// ---- solver
std::unique_ptr<ForwardSolver> ForwardSolver::create(const std::string &dimension, const std::string &type)
{
    std::unique_ptr<ForwardSolver> s(new ForwardSolver());
    check(wv_create(dimension.c_str(), type.c_str(), &s->h_));
    return s;
}

ForwardSolver::~ForwardSolver()
{
    if (h_) wv_destroy(h_);
}

void ForwardSolver::initForwardSolver(const Configuration &c, int nshots, int stepsPerGraph, int tileY)
{
    wv_config k{};
    k.NX = (int)c.getInt("NX");
    k.NY = (int)c.getInt("NY");
    k.NZ = (int)c.getInt("NZ", 1);
    k.DH = c.getDouble("DH");
    k.DT = c.getDouble("DT");
    // NT = (IndexType)(T/DT + 0.5) in the value type (Sources.cpp:26)
    k.NT = (int)((float)c.getDouble("T") / (float)k.DT + 0.5f);
    k.spatial_fd_order = (int)c.getInt("spatialFDorder");
    k.ops_mode = c.getInt("useStencilMatrix", 0) ? WV_OPS_STENCIL : WV_OPS_SPARSE;
    k.free_surface = (int)c.getInt("FreeSurface", 0);
    k.damping_boundary = (int)c.getInt("DampingBoundary", 0);
    k.boundary_width = (int)c.getInt("BoundaryWidth", 0);
    k.damping_coeff = c.getDouble("DampingCoeff", 8.0);
    k.npower = c.getDouble("NPower", 4.0);
    k.fc_cpml = c.getDouble("CenterFrequencyCPML", 5.0);
    k.vmax_cpml = c.getDouble("VMaxCPML", 3500.0);
    k.nshots = nshots;
    // kernel family: fused (single-pass where it exists, default), phased (two launches), strip (2D elastic
    // row-strip kernels), plain (reference sequence)
    const std::string path = c.has("solverPath") ? c.getString("solverPath") : "fused";
    if (path == "phased")
        k.path = WV_PATH_PHASED;
    else if (path == "strip")
        k.path = WV_PATH_STRIP;
    else if (path == "plain")
        k.path = WV_PATH_PLAIN;
    else if (path == "fused")
        k.path = WV_PATH_FUSED;
    else
        throw std::runtime_error("solverPath must be fused, phased, strip or plain");
    k.steps_per_graph = stepsPerGraph;
    k.tile_y = tileY;
    check(wv_init(h_, &k));
    cfg_ = k;
}

void ForwardSolver::prepareForModelling(const std::vector<float> &m0, const std::vector<float> &m1,
                                        const std::vector<float> &m2)
{
    check(wv_set_model(h_, m0.data(), m1.empty() ? nullptr : m1.data(), m2.data()));
}

void ForwardSolver::setRelaxation(const std::vector<double> &freqs, const std::vector<float> &tauP,
                                  const std::vector<float> &tauS)
{
    check(wv_set_relaxation(h_, (int)freqs.size(), freqs.data(), tauP.data(), tauS.data()));
}

void ForwardSolver::setAcquisition(int shot, const std::vector<long long> &srcIndex, const std::vector<int> &srcType,
                                   const std::vector<float> &srcSignal, const std::vector<long long> &recIndex,
                                   const std::vector<int> &recType)
{
    check(wv_set_acquisition(h_, shot, (int)srcIndex.size(), srcIndex.data(), srcType.data(), srcSignal.data(),
                             (int)recIndex.size(), recIndex.data(), recType.data()));
}

void ForwardSolver::resetCPML() { check(wv_reset(h_)); }
void ForwardSolver::run(int tStart, int tEnd) { check(wv_run(h_, tStart, tEnd)); }
void ForwardSolver::synchronize() { check(wv_synchronize(h_)); }
void ForwardSolver::checkFinite() { check(wv_check_finite(h_)); }

std::string checkNumericalArtefactsAndInstabilities(const std::string &dimension, const std::string &type,
                                                    const wv_config &cfg, const std::vector<float> &m0,
                                                    const std::vector<float> &m1, const std::vector<float> &m2,
                                                    double fcMax)
{
    int ok = 1;
    check(wv_check_parameters(dimension.c_str(), type.c_str(), &cfg, m0.empty() ? nullptr : m0.data(),
                              m1.empty() ? nullptr : m1.data(), m2.empty() ? nullptr : m2.data(), fcMax, &ok));
    return ok ? std::string() : std::string(wv_last_error());
}

int ForwardSolver::tune(int nsteps)
{
    int choice = -1;
    check(wv_tune(h_, nsteps, &choice));
    return choice;
}

double ForwardSolver::lastRunMs()
{
    double ms = 0;
    check(wv_last_run_ms(h_, &ms));
    return ms;
}

std::vector<float> ForwardSolver::seismograms(int shot, int nrec)
{
    std::vector<float> out((size_t)nrec * cfg_.NT);
    if (nrec) check(wv_get_seismograms(h_, shot, out.data()));
    return out;
}

}  // namespace wavesim
