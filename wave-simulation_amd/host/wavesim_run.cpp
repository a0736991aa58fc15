// Command-line driver: the shot loop of the reference's Simulation.cpp
// (Simulation.cpp:362-535) for the paths this library accelerates.
//
//   wavesim_run <configuration file> [KEY=VALUE ...]
//
// Reads the reference configuration (later KEY=VALUE arguments override it),
// the model (ModelRead=1: <ModelFilename>.{vp,vs,density}.mtx, TM-EM
// {epsilonr,sigma,mur}, viscoelastic also .tauP / .tauS with the relaxation
// mechanisms of the configuration; ModelRead=0: the homogeneous values of the
// configuration), the source file (<SourceFilename>.txt: synthetic wavelets of
// shapes 1..9, or wavelet types 2 / 3, rows of <SourceSignalFilename>.mtx) and
// receiver files (<ReceiverFilename>.txt, or .shot_<n>.txt with
// useReceiversPerShot=1); runs every shot on the GPU in one batched handle and
// writes <SeismogramFilename>.shot_<n>.<component>.mtx per shot and recorded
// component, resampled to seismoDT (Seismogram.cpp:82-156), after the trace
// normalisation normalizeTraces = 1 / 2 (Seismogram.cpp:215-234), as MatrixMarket
// (SeismogramFormat = 1) or SeismicUnix (SeismogramFormat = 4, <...>.su,
// SUIO.hpp:156-260).  Keys of this
// driver beyond the reference's: stepsPerGraph (time steps per captured graph,
// default 50), tileY (fixed tile height; default: autotuned) and solverPath
// (fused | phased | strip | plain).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>

#include "wavesim.hpp"

using namespace wavesim;

namespace {

// Seismogram::normalizeTrace (Seismogram.cpp:215-234): each trace [ntr][ns] divided by
// its maximum absolute value (mode 1) or its L2 norm (mode 2); a zero norm divides by 1.
void normalizeTraces(std::vector<float> &d, int ntr, int ns, int mode)
{
    if (mode <= 0) return;
    for (int r = 0; r < ntr; r++) {
        float *t = &d[(size_t)r * ns];
        double nrm = 0.0;
        if (mode == 1) {
            for (int i = 0; i < ns; i++) nrm = std::max(nrm, (double)std::fabs(t[i]));
        } else {
            for (int i = 0; i < ns; i++) nrm += (double)t[i] * t[i];
            nrm = std::sqrt(nrm);
        }
        float f = (float)nrm;
        if (f == 0.f) f = 1.f;
        for (int i = 0; i < ns; i++) t[i] /= f;
    }
}

// SUIO::writeSU (SUIO.hpp:156-260): per trace a 240-byte SEG-Y trace header (the SU
// layout; unset fields 0, counit 1) and ns native-endian FP32 samples.  Coordinates in
// millimetres (scalel = scalco = -3), the source position only when the shot has a
// single source (else the origin), angles as in the reference (reference point
// shifted by 800 m for swdep); arithmetic in FP32 like the reference's float build.
void writeSU(const std::string &fn, const std::vector<float> &d, int ntr, int ns, double outDT,
             const std::vector<const ReceiverRow *> &rec, int sx, int sy, int sz, float DH)
{
    std::ofstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + fn);
    const float XS = (float)sx * DH, YS = (float)sy * DH, ZS = (float)sz * DH;
    const float xshift = 800.f, yshift = 800.f, pi = 3.1415926f;
    const float dtus = (float)(outDT * 1000000.0);
    unsigned char h[240];
    auto i32 = [&](int off, int32_t v) { std::memcpy(h + off, &v, 4); };
    auto i16 = [&](int off, int16_t v) { std::memcpy(h + off, &v, 2); };
    auto f32 = [&](int off, float v) { std::memcpy(h + off, &v, 4); };
    for (int r = 0; r < ntr; r++) {
        std::memset(h, 0, sizeof(h));
        const float xr = (float)rec[r]->x * DH, yr = (float)rec[r]->y * DH, zr = (float)rec[r]->z * DH;
        const float x = xr - XS, y = yr - YS, z = zr - ZS;
        i32(0, r + 1);                                                                  // tracl
        i32(4, 1);                                                                      // tracr
        i32(16, 1);                                                                     // ep
        i32(20, ntr);                                                                   // cdp
        i16(28, 1);                                                                     // trid
        i32(36, (int32_t)std::lround(std::sqrt(x * x + y * y + z * z) * 1000.0));       // offset
        i32(40, (int32_t)std::lround(yr * 1000.0));                                     // gelev
        i32(48, (int32_t)std::lround(YS * 1000.0));                                     // sdepth
        i32(52, (int32_t)std::lround(std::atan2(-y, z) * 180 * 1000.0 / pi));           // gdel
        i32(60, (int32_t)std::lround((360.0f / (2.0f * pi)) * std::atan2(x - xshift, y - yshift) * 1000.0));  // swdep
        i32(64, (int32_t)std::lround(std::sqrt(z * z + y * y) * 1000.0));               // gwdep
        i16(68, -3);                                                                    // scalel
        i16(70, -3);                                                                    // scalco
        i32(72, (int32_t)std::lround(XS * 1000.0));                                     // sx
        i32(76, (int32_t)std::lround(ZS * 1000.0));                                     // sy
        i32(80, (int32_t)std::lround(xr * 1000.0));                                     // gx
        i32(84, (int32_t)std::lround(zr * 1000.0));                                     // gy
        i16(88, 1);                                                                     // counit
        const uint16_t nsu = (uint16_t)ns, dtu = (uint16_t)std::lround(dtus);
        std::memcpy(h + 114, &nsu, 2);                                                  // ns
        std::memcpy(h + 116, &dtu, 2);                                                  // dt [us]
        f32(180, (float)dtu * 1.0e-6f);                                                 // d1
        i32(204, ntr);                                                                  // ntr
        f.write((const char *)h, sizeof(h));
        f.write((const char *)&d[(size_t)r * ns], sizeof(float) * (size_t)ns);
    }
    if (!f) throw std::runtime_error("write failed: " + fn);
}

// One trace header of an SU file (the fields acquisition input reads; SEG-Y byte offsets:
// sdepth 48, scalel 68, scalco 70, sx 72, sy 76, gx 80, gy 84, gelev 40, ns 114).
struct SuHeader {
    int32_t gelev, sdepth, sx, sy, gx, gy;
    int16_t scalel, scalco;
    uint16_t ns;
};

// SUIO::readHeaderSU (SUIO.hpp:380-405): every trace header of an SU file (native endian)
std::vector<SuHeader> readSuHeaders(const std::string &fn)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("Error opening file: " + fn);
    std::vector<SuHeader> out;
    unsigned char h[240];
    while (f.read((char *)h, sizeof(h))) {
        SuHeader t;
        std::memcpy(&t.gelev, h + 40, 4);
        std::memcpy(&t.sdepth, h + 48, 4);
        std::memcpy(&t.scalel, h + 68, 2);
        std::memcpy(&t.scalco, h + 70, 2);
        std::memcpy(&t.sx, h + 72, 4);
        std::memcpy(&t.sy, h + 76, 4);
        std::memcpy(&t.gx, h + 80, 4);
        std::memcpy(&t.gy, h + 84, 4);
        std::memcpy(&t.ns, h + 114, 2);
        out.push_back(t);
        f.seekg((std::streamoff)t.ns * 4, std::ios::cur);
    }
    return out;
}

// SUIO::readSingleDataSU (SUIO.hpp:350-372): the samples of trace `trace` (every trace of the
// file holds the first header's ns samples)
std::vector<float> readSuTrace(const std::string &fn, long trace)
{
    std::ifstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("Error opening file: " + fn);
    unsigned char h[240];
    if (!f.read((char *)h, sizeof(h))) throw std::runtime_error("empty SU file: " + fn);
    uint16_t ns;
    std::memcpy(&ns, h + 114, 2);
    f.seekg((std::streamoff)(240 + 4 * (std::streamoff)ns) * trace, std::ios::cur);
    std::vector<float> d(ns);
    if (!f.read((char *)d.data(), 4 * (std::streamoff)ns)) throw std::runtime_error("short SU file: " + fn);
    return d;
}

// Acquisition components in the order of SeismogramTypeString (Acquisition.hpp:28-32): the SU
// acquisition files are <name>.p.su, .vx.su, .vy.su, .vz.su (the seismic names, also for EM)
const char *const kSuComp[4] = {"p", "vx", "vy", "vz"};

// suHandler::readAllSettingsFromSU for sources (suHandler.cpp:13-27, 63-90): every trace of the
// four component files is a source of that component at grid point
// x = int(sx 10^scalco / DH + 0.5), y = int(sdepth 10^scalel / DH + 0.5), z = int(sy 10^scalco / DH
// + 0.5) (value-type arithmetic); the settings the SU header has no field for stay zero: every
// source belongs to shot 0, reads row 0 of the signal file (wavelet type 3) and has amp 0.  All four
// files must exist (readHeaderSU throws otherwise).
std::vector<SourceRow> readSourcesSU(const std::string &base, float DH)
{
    std::vector<SourceRow> out;
    for (int c = 0; c < 4; c++) {
        for (const SuHeader &h : readSuHeaders(base + "." + kSuComp[c] + ".su")) {
            const float sco = std::pow(10.f, (float)h.scalco), sel = std::pow(10.f, (float)h.scalel);
            SourceRow r{};
            r.x = (int)((double)((float)h.sx * sco / DH) + 0.5);
            r.y = (int)((double)((float)h.sdepth * sel / DH) + 0.5);
            r.z = (int)((double)((float)h.sy * sco / DH) + 0.5);
            r.type = c + 1;
            r.wavelet = 3;
            out.push_back(r);
        }
    }
    return out;
}

// suHandler::readAllSettingsFromSU for receivers (suHandler.cpp:29-49, 92-113): as for sources
// with gx, gelev, gy and no rounding (truncation); missing component files are skipped, all four
// missing is an error.
std::vector<ReceiverRow> readReceiversSU(const std::string &base, float DH)
{
    std::vector<ReceiverRow> out;
    int missing = 0;
    for (int c = 0; c < 4; c++) {
        std::vector<SuHeader> hs;
        try {
            hs = readSuHeaders(base + "." + kSuComp[c] + ".su");
        } catch (const std::runtime_error &) {
            if (++missing == 4) throw std::runtime_error("No file with name: " + base + ".'comp'.su could be read");
            continue;
        }
        for (const SuHeader &h : hs) {
            const float sco = std::pow(10.f, (float)h.scalco), sel = std::pow(10.f, (float)h.scalel);
            ReceiverRow r{};
            r.x = (int)((float)h.gx * sco / DH);
            r.y = (int)((float)h.gelev * sel / DH);
            r.z = (int)((float)h.gy * sco / DH);
            r.type = c + 1;
            out.push_back(r);
        }
    }
    return out;
}

// Sources::readSignalFromFile (Sources.cpp:231-249): with initSourcesFromSU = 1 trace `row` of the
// SU file <SourceSignalFilename> as given, else row `row` of <SourceSignalFilename>.mtx
// (SeismogramFormat 1; IO.hpp:172-195); NT samples
std::vector<float> signalFromFile(const Configuration &config, long row, int NT)
{
    if (config.getInt("initSourcesFromSU", 0) != 0) {
        std::vector<float> d = readSuTrace(config.getString("SourceSignalFilename"), row);
        if ((long)d.size() != NT) throw std::runtime_error("Source signal has invalid length");
        return d;
    }
    const int format = (int)config.getInt("SeismogramFormat", 1);
    if (format == 2 || format == 3)
        throw std::runtime_error("source signal files in the lmf/frv formats are not read (SeismogramFormat 1 = mtx)");
    if (format != 1) throw std::runtime_error("Unexpected fileFormat option!");
    return readMatrixRowMtx(config.getString("SourceSignalFilename") + ".mtx", row, (size_t)NT);
}

// Visco-EM model (ViscoTMEM.cpp / ViscoEMEM.cpp init): tau of the permittivity (".tauEpsilon",
// relative) and of the conductivity (".tauSigmar", relative to tau_ref = 1 / (2 pi
// CenterFrequencyCPML), stored times tau_ref).  Constant models (ModelRead = 0) give the static
// epsilon_r and sigma directly; model files hold the "real effective" values at the reference
// frequency, which become the static ones (calculateType 2, ModelparameterEM.cpp:473-554):
//   a = mean_l w^2 t_l^2 / (1 + w^2 t_l^2), b = mean_l w^2 t_l / (1 + w^2 t_l^2)   (w = 2 pi fc)
//   eps_s = (eps_re - sig_re tauS) / (1 - a tauE - b tauE tauS), floored at eps0
//   sig_s = sig_re - b tauE eps_s, floored at 0
// in the value type.  m0 returns eps_s / eps0, m1 sig_s; tauP / tauS the two tau as the solver
// consumes them.
void viscoEmModel(const Configuration &config, size_t n, const std::vector<double> &freqs, std::vector<float> &m0,
                  std::vector<float> &m1, std::vector<float> &tauP, std::vector<float> &tauS)
{
    const float fc = (float)config.getDouble("CenterFrequencyCPML");
    const float tref = (float)(1.0 / (2.0 * M_PI * (double)fc));
    if (config.getInt("ModelRead", 0) == 1) {
        const std::string base = config.getString("ModelFilename");
        tauP = readVectorMtx(base + ".tauEpsilon.mtx", n);
        tauS = readVectorMtx(base + ".tauSigmar.mtx", n);
    } else {
        tauP.assign(n, (float)config.getDouble("tauEpsilon"));
        tauS.assign(n, (float)config.getDouble("tauSigmar"));
    }
    for (auto &v : tauS) v *= tref;
    if (config.getInt("ModelRead", 0) != 1) return;
    const float eps0 = 8.8541878176e-12f;
    const float w = (float)(2.0 * M_PI * (double)fc);
    float a = 0.f, b = 0.f;
    for (double f : freqs) {
        const float t = (float)(1.0 / (2.0 * M_PI * (double)(float)f));
        a += w * w * t * t / (1 + w * w * t * t);
        b += w * w * t / (1 + w * w * t * t);
    }
    a /= (float)freqs.size();
    b /= (float)freqs.size();
    for (size_t i = 0; i < n; i++) {
        const float ere = m0[i] * eps0, sre = m1[i];
        float at = a * tauP[i];
        at = 1 - at;
        float bt = b * tauP[i];
        bt *= tauS[i];
        at -= bt;
        float es = sre * tauS[i];
        es = ere - es;
        es /= at;
        if (es < eps0) es = eps0;
        float sb = b * tauP[i];
        sb *= es;
        float ss = sre - sb;
        if (ss < 0) ss = 0;
        m0[i] = es / eps0;
        m1[i] = ss;
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <configuration file> [KEY=VALUE ...]\n", argv[0]);
        return 2;
    }
    try {
        Configuration config(argv[1]);
        for (int a = 2; a < argc; a++) {
            std::string kv = argv[a];
            const size_t eq = kv.find('=');
            if (eq == std::string::npos) throw std::runtime_error("override must be KEY=VALUE: " + kv);
            std::string k = kv.substr(0, eq);
            for (auto &c : k) c = (char)std::tolower((unsigned char)c);
            config.add(k, kv.substr(eq + 1), true);
        }
        const std::string dim = config.getString("dimension"), type = config.getString("equationType");
        std::string tl = type;
        for (auto &c : tl) c = (char)std::tolower((unsigned char)c);
        const bool em = tl == "tmem" || tl == "emem" || tl == "viscotmem" || tl == "viscoemem";
        const bool vem = tl == "viscotmem" || tl == "viscoemem";
        const int NX = (int)config.getInt("NX"), NY = (int)config.getInt("NY");
        const int NZ = (dim == "2D" || dim == "2d") ? 1 : (int)config.getInt("NZ");
        config.add("nz", std::to_string(NZ), true);
        const size_t n = (size_t)NX * NY * NZ;

        // model
        std::vector<float> m0, m1, m2;
        const char *names[3] = {"vp", "vs", "density"};
        const char *names_em[3] = {"epsilonr", "sigma", "mur"};
        const char *keys[3] = {"velocityP", "velocityS", "rho"};
        const char *keys_em[3] = {"epsilonr", "sigma", "mur"};
        std::vector<float> *m[3] = {&m0, &m1, &m2};
        const bool acoustic = type == "acoustic";
        if (config.getInt("ModelRead", 0) == 1) {
            if (config.getInt("fileFormat", 1) != 1) throw std::runtime_error("only fileFormat=1 (mtx) is read");
            const std::string base = config.getString("ModelFilename");
            for (int q = 0; q < 3; q++) {
                if (acoustic && q == 1) continue;
                *m[q] = readVectorMtx(base + "." + (em ? names_em[q] : names[q]) + ".mtx", n);
            }
        } else {
            for (int q = 0; q < 3; q++) {
                if (acoustic && q == 1) continue;
                m[q]->assign(n, (float)config.getDouble(em ? keys_em[q] : keys[q]));
            }
        }
        // viscoelastic / visco-SH: relaxation mechanisms from the configuration, tauP / tauS from
        // the model files or the configuration (Viscoelastic.cpp:160-200, 260-290; visco-SH reads
        // tauS only, ViscoSH.cpp)
        const bool visco = tl == "viscoelastic" || tl == "viscosh" || vem;
        std::vector<double> freqs;
        std::vector<float> tauP, tauS;
        if (visco) {
            const long nm = config.getInt("numRelaxationMechanisms");
            const char *fk[4] = {"relaxationFrequency", "relaxationFrequency2", "relaxationFrequency3",
                                 "relaxationFrequency4"};
            for (long l = 0; l < nm; l++) freqs.push_back(l < 4 ? config.getDouble(fk[l]) : 0.0);
            if (vem) {
                viscoEmModel(config, n, freqs, m0, m1, tauP, tauS);
            } else if (config.getInt("ModelRead", 0) == 1) {
                const std::string base = config.getString("ModelFilename");
                if (tl != "viscosh") tauP = readVectorMtx(base + ".tauP.mtx", n);
                tauS = readVectorMtx(base + ".tauS.mtx", n);
            } else {
                if (tl != "viscosh") tauP.assign(n, (float)config.getDouble("tauP"));
                tauS.assign(n, (float)config.getDouble("tauS"));
            }
        }

        // acquisition: shots in order of first appearance in the source file
        const bool srcSU = config.getInt("initSourcesFromSU", 0) != 0, recSU = config.getInt("initReceiverFromSU", 0) != 0;
        const float DHf = (float)config.getDouble("DH");
        const std::vector<SourceRow> src = srcSU ? readSourcesSU(config.getString("SourceFilename"), DHf)
                                                 : readSources(config.getString("SourceFilename") + ".txt");
        std::vector<int> shotNumbers;
        for (auto &s : src)
            if (std::find(shotNumbers.begin(), shotNumbers.end(), s.shot) == shotNumbers.end())
                shotNumbers.push_back(s.shot);
        const int nshots = (int)shotNumbers.size();
        auto solver = ForwardSolver::create(dim, type);
        solver->initForwardSolver(config, nshots, (int)config.getInt("stepsPerGraph", 50),
                                  (int)config.getInt("tileY", 0));
        const int NT = solver->NT();
        const double DT = config.getDouble("DT");
        if (visco) solver->setRelaxation(freqs, tauP, tauS);
        solver->prepareForModelling(m0, m1, m2);
        auto lin = [&](int x, int y, int z) { return (long long)x + (long long)z * NX + (long long)y * NX * NZ; };
        std::vector<std::vector<ReceiverRow>> recs(nshots);
        for (int k = 0; k < nshots; k++) {
            std::vector<long long> si, ri;
            std::vector<int> st, rt;
            std::vector<float> sig;
            // per shot, as Simulation.cpp:389: CFL and vp/vs throw, dispersion is reported
            double fcMax = 0;
            for (auto &s : src)
                if (s.shot == shotNumbers[k]) fcMax = std::max(fcMax, s.fc);
            const std::string warn =
                checkNumericalArtefactsAndInstabilities(dim, type, solver->config(), m0, m1, m2, fcMax);
            if (!warn.empty()) std::printf("shot %d: %s\n", shotNumbers[k], warn.c_str());
            bool fromFile2 = false, fromFile3 = false;
            for (auto &s : src) {
                if (s.shot != shotNumbers[k]) continue;
                si.push_back(lin(s.x, s.y, s.z));
                st.push_back(s.type);
                std::vector<float> w;
                if (s.wavelet == 1) {
                    w = sourceSignal(s.shape, NT, DT, s.fc, s.amp, s.tshift);
                } else if (s.wavelet == 2 || s.wavelet == 3) {
                    (s.wavelet == 2 ? fromFile2 : fromFile3) = true;
                    if (fromFile2 && fromFile3) throw std::runtime_error("Combination of wavelet type 2 and 3 not supported");
                    w = signalFromFile(config, s.wavelet == 2 ? 0 : s.row, NT);
                    // copySignalsToSeismogramHandler keeps only the sign of the amplitude, held there as
                    // an integer: 0 <= amp < 1 gives -1 (Sources.cpp:37-45, 354-375)
                    if (!((long)s.amp > 0))
                        for (auto &v : w) v = -v;
                } else {
                    throw std::runtime_error("Unknown wavelet type ");
                }
                sig.insert(sig.end(), w.begin(), w.end());
            }
            const std::string rf = config.getString("ReceiverFilename") +
                                   (config.getInt("useReceiversPerShot", 0) ? ".shot_" + std::to_string(shotNumbers[k]) : "") +
                                   ".txt";
            recs[k] = recSU ? readReceiversSU(config.getString("ReceiverFilename"), DHf) : readReceivers(rf);
            for (auto &r : recs[k]) {
                ri.push_back(lin(r.x, r.y, r.z));
                rt.push_back(r.type);
            }
            solver->setAcquisition(k, si, st, sig, ri, rt);
        }

        // kernel variant: autotuned on this problem unless tileY is given
        if (!config.has("tileY")) solver->tune(std::min(NT, 100));
        solver->resetCPML();
        const auto w0 = std::chrono::steady_clock::now();
        solver->run(0, NT);
        solver->synchronize();
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
        solver->checkFinite();  // Simulation.cpp:519, over every shot of the batch

        // seismograms per shot and component (Acquisition.hpp:28-33 names)
        static const char *comp[5] = {"", "p", "vx", "vy", "vz"};
        static const char *comp_em[5] = {"", "ez", "ex", "ey", "hz"};
        const double seismoDT = config.getDouble("seismoDT", DT);
        const double ratio = seismoDT / DT;
        const std::string out = config.getString("SeismogramFilename");
        const int format = (int)config.getInt("SeismogramFormat", 1), norm = (int)config.getInt("normalizeTraces", 0);
        if (format != 1 && format != 4)
            throw std::runtime_error("SeismogramFormat " + std::to_string(format) + " is not written (1 = mtx, 4 = su)");
        if (norm < 0 || norm > 2)
            throw std::runtime_error("normalizeTraces " + std::to_string(norm) + " is not applied (0, 1, 2)");
        for (int k = 0; k < nshots; k++) {
            const int nr = (int)recs[k].size();
            std::vector<float> all = solver->seismograms(k, nr);
            // the shot's source position for the SU headers: its only source, or the origin
            int sxi = 0, syi = 0, szi = 0, nsrc = 0;
            for (auto &s2 : src)
                if (s2.shot == shotNumbers[k]) {
                    sxi = s2.x;
                    syi = s2.y;
                    szi = s2.z;
                    nsrc++;
                }
            if (nsrc != 1) sxi = syi = szi = 0;
            for (int ty = 1; ty <= 4; ty++) {
                std::vector<float> sel;
                std::vector<const ReceiverRow *> selr;
                int ntr = 0;
                for (int r = 0; r < nr; r++) {
                    if (recs[k][r].type != ty) continue;
                    sel.insert(sel.end(), all.begin() + (size_t)r * NT, all.begin() + (size_t)(r + 1) * NT);
                    selr.push_back(&recs[k][r]);
                    ntr++;
                }
                if (!ntr) continue;
                normalizeTraces(sel, ntr, NT, norm);  // on the time-loop samples, before resampling
                int NTo = NT;
                if (std::fabs(ratio - 1.0) > 1e-12) sel = resampleTraces(sel, ntr, NT, ratio, NTo);
                const std::string fn = out + ".shot_" + std::to_string(shotNumbers[k]) + "." +
                                       (em ? comp_em[ty] : comp[ty]) + (format == 4 ? ".su" : ".mtx");
                if (format == 4)
                    writeSU(fn, sel, ntr, NTo, seismoDT, selr, sxi, syi, szi, DHf);
                else
                    writeMatrixMtx(fn, sel, ntr, NTo);
                std::printf("writing %s\n", fn.c_str());
            }
        }
        std::printf("shots %d, time steps %d, time loop %.3f ms (GPU %.3f ms), %.4g grid-cell updates/s\n", nshots, NT,
                    wall * 1e3, solver->lastRunMs(), (double)n * NT * nshots / wall);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
