// Command-line driver: the shot loop of the reference's Simulation.cpp
// (Simulation.cpp:362-535) for the paths this library accelerates.
//
//   wavesim_run <configuration file> [KEY=VALUE ...]
//
// Reads the reference configuration (later KEY=VALUE arguments override it),
// the model (ModelRead=1: <ModelFilename>.{vp,vs,density}.mtx, TM-EM
// {epsilonr,sigma,mur}; ModelRead=0: the homogeneous values of the
// configuration), the source file (<SourceFilename>.txt, Ricker wavelets) and
// receiver files (<ReceiverFilename>.txt, or .shot_<n>.txt with
// useReceiversPerShot=1); runs every shot on the GPU in one batched handle and
// writes <SeismogramFilename>.shot_<n>.<component>.mtx per shot and recorded
// component, resampled to seismoDT (Seismogram.cpp:82-156).  Keys of this
// driver beyond the reference's: stepsPerGraph (time steps per captured graph,
// default 50), tileY (fixed tile height; default: autotuned) and solverPath
// (fused | phased | strip | plain).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <map>

#include "wavesim.hpp"

using namespace wavesim;

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
        const bool em = type == "tmem" || type == "TMEM";
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

        // acquisition: shots in order of first appearance in the source file
        const std::vector<SourceRow> src = readSources(config.getString("SourceFilename") + ".txt");
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
        solver->prepareForModelling(m0, m1, m2);
        auto lin = [&](int x, int y, int z) { return (long long)x + (long long)z * NX + (long long)y * NX * NZ; };
        std::vector<std::vector<ReceiverRow>> recs(nshots);
        for (int k = 0; k < nshots; k++) {
            std::vector<long long> si, ri;
            std::vector<int> st, rt;
            std::vector<float> sig;
            for (auto &s : src) {
                if (s.shot != shotNumbers[k]) continue;
                if (s.wavelet != 1 || s.shape != 1) throw std::runtime_error("only synthetic Ricker sources are generated");
                si.push_back(lin(s.x, s.y, s.z));
                st.push_back(s.type);
                std::vector<float> w = ricker(NT, DT, s.fc, s.amp, s.tshift);
                sig.insert(sig.end(), w.begin(), w.end());
            }
            const std::string rf = config.getString("ReceiverFilename") +
                                   (config.getInt("useReceiversPerShot", 0) ? ".shot_" + std::to_string(shotNumbers[k]) : "") +
                                   ".txt";
            recs[k] = readReceivers(rf);
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

        // seismograms per shot and component (Acquisition.hpp:28-33 names)
        static const char *comp[5] = {"", "p", "vx", "vy", "vz"};
        static const char *comp_em[5] = {"", "ez", "ex", "ey", "hz"};
        const double seismoDT = config.getDouble("seismoDT", DT);
        const double ratio = seismoDT / DT;
        const std::string out = config.getString("SeismogramFilename");
        for (int k = 0; k < nshots; k++) {
            const int nr = (int)recs[k].size();
            std::vector<float> all = solver->seismograms(k, nr);
            for (int ty = 1; ty <= 4; ty++) {
                std::vector<float> sel;
                int ntr = 0;
                for (int r = 0; r < nr; r++) {
                    if (recs[k][r].type != ty) continue;
                    sel.insert(sel.end(), all.begin() + (size_t)r * NT, all.begin() + (size_t)(r + 1) * NT);
                    ntr++;
                }
                if (!ntr) continue;
                int NTo = NT;
                if (std::fabs(ratio - 1.0) > 1e-12) sel = resampleTraces(sel, ntr, NT, ratio, NTo);
                const std::string fn =
                    out + ".shot_" + std::to_string(shotNumbers[k]) + "." + (em ? comp_em[ty] : comp[ty]) + ".mtx";
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
