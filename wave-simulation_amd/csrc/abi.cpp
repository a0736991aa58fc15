// C ABI of the MI355X forward solver (include/wave_abi.h): handle management,
// model/acquisition upload and the time loop driver.
#include "../../include/wave_abi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "handle.hpp"
#include "prep.hpp"

namespace {
thread_local std::string g_err;

std::string lower(const char *s)
{
    std::string r = s ? s : "";
    for (auto &c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}
}  // namespace

int wv_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

wv_solver::~wv_solver()
{
    if (st) (void)hipStreamSynchronize(st);
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    drop_graph();
    delete plain;
    delete e3;
    adjoint_destroy(adj);
    free_list(acq_allocs);
    free_list(allocs);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (st) (void)hipStreamDestroy(st);
}

int wv_solver::sync_acquisition()
{
    if (!acq_dirty) return 0;
    drop_graph();
    HIP_TRY(hipStreamSynchronize(st));
    free_list(acq_allocs);
    std::vector<int> sp(1, 0), rp(1, 0), stp, rtp;
    std::vector<long long> sc, rc;
    std::vector<float> sig;
    for (auto &s : shots) {
        for (size_t k = 0; k < s.sidx.size(); k++) {
            const long long g = s.sidx[k];
            sc.push_back(L.at((int)(g % L.NX), (int)(g / L.NX)));
            stp.push_back(s.stype[k]);
        }
        sig.insert(sig.end(), s.ssig.begin(), s.ssig.end());
        sp.push_back((int)sc.size());
        for (size_t k = 0; k < s.ridx.size(); k++) {
            const long long g = s.ridx[k];
            rc.push_back(L.at((int)(g % L.NX), (int)(g / L.NX)));
            rtp.push_back(s.rtype[k]);
        }
        rp.push_back((int)rc.size());
    }
    rec_total = (int)rc.size();
    rec_ptr_h = rp;
    // per (shot, tile) lists of the sources / receivers each tile owns, in
    // list order; for the single-pass kernel also the velocity-type sources
    // inside each tile's velocity region (tile + H halo)
    nbx = (L.NX + tile_x - 1) / tile_x;
    nby = (L.NY + tile_y - 1) / tile_y;
    const size_t nsb = (size_t)cfg.nshots * nbx * nby;
    std::vector<std::vector<int>> bs(nsb), br(nsb), bv(nsb);
    {
        int gs = 0, gr = 0;
        for (int sh = 0; sh < (int)shots.size(); sh++) {
            auto blk = [&](int bxx, int byy) { return ((size_t)sh * nby + byy) * nbx + bxx; };
            for (size_t k = 0; k < shots[sh].sidx.size(); k++, gs++) {
                const long long g = shots[sh].sidx[k];
                const int x = (int)(g % L.NX), y = (int)(g / L.NX);
                bs[blk(x / tile_x, y / tile_y)].push_back(gs);
                if (variant == V_SINGLE && (shots[sh].stype[k] == 2 || shots[sh].stype[k] == 3)) {
                    for (int byy = 0; byy < nby; byy++)
                        for (int bxx = 0; bxx < nbx; bxx++) {
                            const int xlo = bxx * tile_x - H, xhi = bxx * tile_x + tile_x + H;
                            const int ylo = byy * tile_y - H, yhi = byy * tile_y + tile_y + H;
                            if (x >= xlo && x < xhi && y >= ylo && y < yhi) bv[blk(bxx, byy)].push_back(gs);
                        }
                }
            }
            for (long long g : shots[sh].ridx) {
                const int x = (int)(g % L.NX), y = (int)(g / L.NX);
                br[blk(x / tile_x, y / tile_y)].push_back(gr++);
            }
        }
    }
    std::vector<int> bsp(1, 0), bsl, brp(1, 0), brl, bvp(1, 0), bvl;
    for (size_t b = 0; b < nsb; b++) {
        bsl.insert(bsl.end(), bs[b].begin(), bs[b].end());
        bsp.push_back((int)bsl.size());
        brl.insert(brl.end(), br[b].begin(), br[b].end());
        brp.push_back((int)brl.size());
        bvl.insert(bvl.end(), bv[b].begin(), bv[b].end());
        bvp.push_back((int)bvl.size());
    }
    int r;
    if ((r = upload(&d_bsrc_ptr, bsp, acq_allocs)) || (r = upload(&d_bsrc, bsl, acq_allocs)) ||
        (r = upload(&d_brec_ptr, brp, acq_allocs)) || (r = upload(&d_brec, brl, acq_allocs)) ||
        (r = upload(&d_vsrc_ptr, bvp, acq_allocs)) || (r = upload(&d_vsrc, bvl, acq_allocs)) ||
        (r = upload(&d_src_ptr, sp, acq_allocs)) || (r = upload(&d_rec_ptr, rp, acq_allocs)) ||
        (r = upload(&d_src_cell, sc, acq_allocs)) || (r = upload(&d_rec_cell, rc, acq_allocs)) ||
        (r = upload(&d_src_type, stp, acq_allocs)) || (r = upload(&d_rec_type, rtp, acq_allocs)) ||
        (r = upload(&d_src_sig, sig, acq_allocs)) ||
        (r = dalloc(&d_seis, (size_t)cfg.NT * std::max(rec_total, 1), acq_allocs)))
        return r;
    HIP_TRY(hipStreamSynchronize(st));
    acq_dirty = false;
    return 0;
}

// Enqueue `nsteps` time steps starting at the device step counter: one run
// segment (velocity-type sources of its last step are applied by the final
// flush, so the fields in memory are the reference state afterwards).  With
// `timed`, every kernel is bracketed by hipEvents and its time accumulated.
int wv_solver::enqueue_steps(int nsteps, bool timed)
{
    const wv::Coef2D C = coef2d();
    const wv::Taps T = taps();
    const wv::Acq A = acq();
    const int order = cfg.spatial_fd_order;
    constexpr int BATCH = 32;
    // timed: every kernel is launched with its own start/stop events (the
    // events are written by the dispatch itself, not by separate commands)
    if (timed && kev.empty()) {
        kev.resize(4 * BATCH + 2);
        for (auto &e : kev) HIP_TRY(hipEventCreate(&e));
    }
    if (variant == V_PERSIST) {
        // one cooperative launch for all nsteps (timed: its events, counted as nsteps H+E steps)
        if (nsteps <= 0) return 0;
        if (timed) HIP_TRY(hipEventRecord(kev[0], st));
        HIP_TRY(wv::launch_tm2d_persistent(order, tile_y, L, cfg.nshots, fields2d(0), C, T, cpml2d(0), A, nsteps, st));
        if (timed) {
            HIP_TRY(hipEventRecord(kev[1], st));
            HIP_TRY(hipEventSynchronize(kev[1]));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, kev[0], kev[1]));
            kms[0] += ms;
            klaunch[0] += nsteps;
        }
        wv::launch_advance(d_tbase, nsteps, st);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    const int nk = variant == V_SINGLE ? 1 : 2;
    int pending = 0;
    auto collect = [&](int n) -> int {
        HIP_TRY(hipEventSynchronize(kev[4 * (n - 1) + 2 * nk - 1]));
        for (int b = 0; b < n; b++)
            for (int k = 0; k < nk; k++) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, kev[4 * b + 2 * k], kev[4 * b + 2 * k + 1]));
                kms[k] += ms;
                klaunch[k] += 1;
            }
        return 0;
    };
    for (int j = 0; j < nsteps; j++) {
        hipEvent_t *e = timed ? &kev[4 * pending] : nullptr;
        hipEvent_t e0 = e ? e[0] : nullptr, e1 = e ? e[1] : nullptr, e2 = e ? e[2] : nullptr, e3 = e ? e[3] : nullptr;
        if (variant == V_SINGLE) {
            wv::launch_fused2d(order, tile_y * tile_load, L, cfg.nshots, fields2d(cur), fields2d(cur ^ 1), C, T, cpml2d(cur),
                               psinew(cur ^ 1), A, j, j > 0, st, e0, e1);
            cur ^= 1;
        } else {
            const wv::Fields2D F = fields2d(cur);
            const wv::Cpml2D P = cpml2d(cur);
            if (variant == V_STRIP && physics == wv::TMEM) {
                wv::launch_strip2d_tm(order, tile_y, L, cfg.nshots, F, C, T, P, A, j, j > 0, st, 0, e0, e1);
                wv::launch_strip2d_tm(order, tile_y, L, cfg.nshots, F, C, T, P, A, j, 0, st, 1, e2, e3);
            } else if (variant == V_STRIP) {
                const wv::Rec2D *rec = adjoint_record(this);
                wv::launch_strip2d(order, tile_y, L, cfg.nshots, F, C, T, P, A, j, j > 0, st, 0, e0, e1, nullptr, rec);
                wv::launch_strip2d(order, tile_y, L, cfg.nshots, F, C, T, P, A, j, 0, st, 1, e2, e3, nullptr, rec);
            } else if (visco && (physics == wv::TMEM || physics == wv::TEMEM)) {
                wv::launch_step2d(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, j > 0, st, 0, e0, e1);
                wv::launch_viscoem2d_e(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, visco2d(), st, e2, e3);
            } else if (visco) {
                wv::launch_step2d(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, j > 0, st, 0, e0, e1);
                wv::launch_visco2d_stress(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, visco2d(), st, e2,
                                          e3);
            } else {
                wv::launch_step2d(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, j > 0, st, 0, e0, e1);
                wv::launch_step2d(physics, order, tile_y, L, cfg.nshots, F, C, T, P, A, j, 0, st, 1, e2, e3);
            }
            if (cfg.damping_boundary == 1) wv::launch_abs_velocity2d(physics, L, cfg.nshots, F, C, st);
        }
        if (timed && ++pending == BATCH) {
            int rc = collect(pending);
            if (rc) return rc;
            pending = 0;
        }
    }
    if (timed && pending) {
        int rc = collect(pending);
        if (rc) return rc;
    }
    if (timed) HIP_TRY(hipEventRecord(kev[4 * BATCH], st));
    if (nsteps > 0) {
        const wv::Fields2D F = fields2d(cur);
        // the single-pass kernel defers the ABS damping of the velocities to the next read
        if (variant == V_SINGLE && cfg.damping_boundary == 1) wv::launch_abs_velocity2d(physics, L, cfg.nshots, F, C, st);
        wv::launch_flush2d(physics, L, F, A, nsteps - 1, st);
    }
    if (timed) {
        HIP_TRY(hipEventRecord(kev[4 * BATCH + 1], st));
        HIP_TRY(hipEventSynchronize(kev[4 * BATCH + 1]));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, kev[4 * BATCH], kev[4 * BATCH + 1]));
        kms[2] += ms;
        klaunch[2] += 1;
    }
    wv::launch_advance(d_tbase, nsteps, st);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Algorithmic HBM bytes of one launch of a kernel kind: fields of every shot
// read once and updated fields written once, shared coefficients read once,
// CPML memory variables read+written on their frame cells.
double wv_solver::kernel_bytes(int kind) const
{
    const double cells = (double)cfg.NX * cfg.NY, B = cfg.nshots;
    if (kind == 2) {
        double nsrc = 0;
        for (auto &sh : shots) nsrc += sh.sidx.size();
        return 12.0 * nsrc + 8.0 * rec_total;
    }
    if (kind >= 3) {
        // adjoint pass (2D elastic): per shot-cell field bytes, per cell coefficient bytes
        //   3 recomputed velocity phase: forward velocity phase + 2 recorded derivative sums written
        //   4 recomputed stress phase  : forward stress phase + 3 recorded derivative sums written
        //   5 adjoint stress step      : r lsxx lsyy lsxy, r+w lvx lvy, the 5 recorded sums (FP16 record or
        //                                FP32 segment buffers); per cell 3 coefficients and 5 FP64 gradient
        //                                updates (one per cell: the shots are summed in the block)
        //   6 adjoint velocity step    : r lvx lvy, r+w lsxx lsyy lsxy; per cell 2 coefficients
        const double hb = adjoint_record(this) ? 2.0 : 4.0;
        const double per_shot[4] = {4.0 * (7 + 2), 4.0 * (8 + 3), 4.0 * 7 + 5 * hb, 4.0 * 8};
        const double per_cell[4] = {4.0 * 2, 4.0 * 3, 4.0 * 3 + 8.0 * 5, 4.0 * 2};
        return cells * (B * per_shot[kind - 3] + per_cell[kind - 3]);
    }
    // {field reads+writes per shot-cell, coefficient reads per cell, psi per x-frame cell, psi per y-frame cell}
    static const int phased[5][2][4] = {
        {{5, 2, 1, 1}, {4, 1, 1, 1}},  // acoustic: velocity, pressure
        {{7, 2, 2, 2}, {8, 3, 2, 2}},  // elastic: velocity, stress
        {{5, 2, 1, 1}, {4, 2, 1, 1}},  // tmem: H, E
        {{4, 1, 1, 1}, {5, 2, 1, 1}},  // sh: vz, stresses
        {{4, 1, 1, 1}, {5, 4, 1, 1}},  // emem: hz, (ex, ey)
    };
    static const int single[4] = {10, 5, 4, 4};  // elastic single pass
    if ((variant == V_SINGLE || variant == V_PERSIST) && kind == 1) return 0.0;
    if (variant == V_PERSIST) {
        // one launch per step loop, counted per time step: the H and E updates together
        wv_solver *self = const_cast<wv_solver *>(this);
        self->variant = V_PHASED;
        const double b = kernel_bytes(0) + kernel_bytes(1);
        self->variant = V_PERSIST;
        return b;
    }
    const int *t = variant == V_SINGLE ? single : phased[physics][kind];
    double bytes = 4.0 * (cells * B * t[0] + cells * t[1]);
    if (cfg.damping_boundary == 2) {
        const double W = cfg.boundary_width;
        const double xf = (double)cfg.NY * 2 * W, yf = (double)cfg.NX * (cfg.free_surface ? W : 2 * W);
        bytes += 8.0 * B * (xf * t[2] + yf * t[3]);
    }
    if (cfg.damping_boundary == 1 && (kind == 1 || variant == V_SINGLE)) bytes += 4.0 * cells;
    // viscoelastic stress phase: the memory variables (3 L read and written per shot-cell) and
    // tauP, tauS, tauS_xy, 1 + L tauP, 1 + L tauS per cell
    // (visco-EM: the memory variables of ez, or of ex and ey, read and written per shot-cell and
    // their Cd per cell)
    if (visco && kind == 1) {
        if (physics == wv::TMEM || physics == wv::TEMEM) {
            const double g = physics == wv::TMEM ? 1.0 : 2.0;
            bytes += 4.0 * (cells * B * 2.0 * g * vL + cells * g * vL);
        } else {
            bytes += physics == wv::SH ? 4.0 * (cells * B * 4.0 * vL + cells * 3.0)
                                       : 4.0 * (cells * B * 6.0 * vL + cells * 5.0);
        }
    }
    return bytes;
}

extern "C" {

const char *wv_last_error(void) { return g_err.c_str(); }

int wv_create(const char *dimension, const char *type, wv_solver **out)
{
    if (!out) return fail(WV_EINVAL, "null output pointer");
    *out = nullptr;
    std::string d = lower(dimension), t = lower(type);
    if (d != "2d" && d != "3d") return fail(WV_EINVAL, "Unknown dimension " + d);
    static const char *known[] = {"acoustic", "elastic", "viscoelastic", "sh", "viscosh",
                                  "tmem", "emem", "viscotmem", "viscoemem"};
    bool ok = false;
    for (auto k : known) ok |= (t == k);
    if (!ok) return fail(WV_EINVAL, "Unkown type " + t);
    int phys;
    if (t == "acoustic")
        phys = wv::ACOUSTIC;
    else if (t == "elastic")
        phys = wv::ELASTIC;
    else if (t == "tmem")
        phys = wv::TMEM;
    else if (t == "sh")
        phys = wv::SH;
    else if (t == "emem")
        phys = wv::TEMEM;  // 2D: TE mode (hz, ex, ey); 3D: all six components
    else if (t == "viscoelastic")
        phys = wv::ELASTIC;  // the elastic step with memory variables in the stress phase
    else if (t == "viscosh" && d == "2d")
        phys = wv::SH;  // the SH step with memory variables in the stress phase
    else if (t == "viscotmem" && d == "2d")
        phys = wv::TMEM;  // the TM step with memory variables in the electric phase
    else if (t == "viscoemem")
        phys = wv::TEMEM;  // the TE (3D: EM) step with memory variables in the electric phase
    else
        return fail(WV_EINVAL, "equationType " + t + " is not provided by this solver");
    if (d == "3d" && phys != wv::ELASTIC && phys != wv::ACOUSTIC && phys != wv::TEMEM)
        return fail(WV_EINVAL, "3D " + t + " is not provided by this build");
    wv_solver *s = new wv_solver();
    s->dim = d == "3d" ? 3 : 2;
    s->physics = phys;
    s->visco = t == "viscoelastic" || t == "viscosh" || t == "viscotmem" || t == "viscoemem";
    *out = s;
    return WV_OK;
}

int wv_destroy(wv_solver *s)
{
    delete s;
    return WV_OK;
}

int wv_init(wv_solver *s, const wv_config *c)
{
    if (!s || !c) return fail(WV_EINVAL, "null argument");
    const wv_config &k = *c;
    if (s->dim == 2 && (k.NX < 2 || k.NY < 2 || k.NZ != 1))
        return fail(WV_EINVAL, "2D grid needs NX, NY >= 2 and NZ = 1");
    if (k.NT < 1) return fail(WV_EINVAL, "NT must be positive");
    if (!(k.DH > 0) || !(k.DT > 0)) return fail(WV_EINVAL, "DH and DT must be positive");
    const int N = k.spatial_fd_order;
    if (N < 2 || N > 12 || (N & 1)) return fail(WV_EINVAL, "spatialFDorder must be one of 2,4,6,8,10,12");
    if (k.ops_mode != WV_OPS_SPARSE && k.ops_mode != WV_OPS_STENCIL) return fail(WV_EINVAL, "bad ops_mode");
    if (k.free_surface != 0 && k.free_surface != 1) return fail(WV_EINVAL, "FreeSurface must be 0 or 1");
    if (k.free_surface && (s->physics == wv::TMEM || s->physics == wv::TEMEM))
        return fail(WV_EINVAL, "FreeSurface is not available for the EM solvers");
    if (k.damping_boundary < 0 || k.damping_boundary > 2) return fail(WV_EINVAL, "DampingBoundary must be 0, 1 or 2");
    if (k.damping_boundary && (k.boundary_width < 1 || 2 * k.boundary_width > std::min(k.NX, k.NY)))
        return fail(WV_EINVAL, "BoundaryWidth out of range");
    if (k.damping_boundary == 1 && !(k.damping_coeff > 0 && k.damping_coeff < 100))
        return fail(WV_EINVAL, "DampingCoeff must be in (0,100)");
    if (k.nshots < 1) return fail(WV_EINVAL, "nshots must be >= 1");
    if (k.path != WV_PATH_FUSED && k.path != WV_PATH_PLAIN && k.path != WV_PATH_PHASED && k.path != WV_PATH_STRIP &&
        k.path != WV_PATH_PERSISTENT)
        return fail(WV_EINVAL, "bad path");
    if (k.path == WV_PATH_PERSISTENT && (s->dim != 2 || s->physics != wv::TMEM || s->visco || k.damping_boundary == 1))
        return fail(WV_EINVAL, "the persistent kernel is 2D TM-EM without ABS only");
    if (k.path == WV_PATH_PERSISTENT && k.tile_y != 0 && k.tile_y != 8 && k.tile_y != 16 && k.tile_y != 32)
        return fail(WV_EINVAL, "tile_y of the persistent kernel must be 0 (default), 8, 16 or 32");
    if (k.path == WV_PATH_STRIP && (s->dim != 2 || (s->physics != wv::ELASTIC && s->physics != wv::TMEM) || s->visco))
        return fail(WV_EINVAL, "the row-strip kernels are 2D elastic and TM-EM only");
    if (s->visco && k.path == WV_PATH_PLAIN)
        return fail(WV_EINVAL, "the plain path is provided for acoustic, elastic and tmem");
    if (k.path == WV_PATH_STRIP && k.tile_y != 0 && k.tile_y != 8 && k.tile_y != 16 && k.tile_y != 32)
        return fail(WV_EINVAL, "tile_y of the row-strip kernels must be 0 (default), 8, 16 or 32");
    if (k.path == WV_PATH_PLAIN && k.nshots != 1) return fail(WV_EINVAL, "the plain path runs one shot per handle");
    if (s->dim == 3 && k.path == WV_PATH_PLAIN) return fail(WV_EINVAL, "the plain path is 2D only");
    if ((s->physics == wv::SH || s->physics == wv::TEMEM) && k.path == WV_PATH_PLAIN)
        return fail(WV_EINVAL, "the plain path is provided for acoustic, elastic and tmem");
    const bool single = s->dim == 2 && k.path == WV_PATH_FUSED && s->physics == wv::ELASTIC && !s->visco;
    if (s->dim == 2 && single && k.tile_y != 0 && std::abs(k.tile_y) != 8 && std::abs(k.tile_y) != 16 && std::abs(k.tile_y) != 24)
        return fail(WV_EINVAL, "tile_y of the single-pass kernel must be 0 (default), +-8, +-16 or +-24");
    if (s->dim == 2 && !single && k.path != WV_PATH_STRIP && k.tile_y != 0 && k.tile_y != 8 && k.tile_y != 16 &&
        k.tile_y != 32)
        return fail(WV_EINVAL, "tile_y of the two-launch kernels must be 0 (default), 8, 16 or 32");
    if (s->inited) return fail(WV_ESTATE, "solver already initialised");

    s->cfg = k;
    if (s->cfg.steps_per_graph < 0) s->cfg.steps_per_graph = 0;
    s->cfg.steps_per_graph += s->cfg.steps_per_graph & 1;  // graphs start and end in buffer 0
    HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&s->ev0));
    HIP_TRY(hipEventCreate(&s->ev1));

    if (s->dim == 3) {
        s->e3 = new wv::Elastic3D();
        int rc = s->e3->init(s->cfg, s->st, s->physics, s->visco);
        if (rc) return fail(rc, s->e3->error);
        s->shots.assign(k.nshots, ShotAcq());
        s->inited = true;
        return WV_OK;
    }

    if (k.path == WV_PATH_PLAIN) {
        s->plain = new wv::PlainSolver();
        int rc = s->plain->init(s->physics, k, s->st);
        if (rc) return fail(rc, s->plain->error);
        s->shots.assign(1, ShotAcq());
        s->inited = true;
        return WV_OK;
    }

    s->variant = single ? V_SINGLE
                        : (k.path == WV_PATH_STRIP ? V_STRIP : (k.path == WV_PATH_PERSISTENT ? V_PERSIST : V_PHASED));
    s->H = N / 2;
    if (s->variant == V_SINGLE) {
        s->tile_x = wv::fused2d_tile_x(N);
        s->tile_y = k.tile_y ? std::abs(k.tile_y) : 16;
        s->tile_load = k.tile_y < 0 ? -1 : 1;
    } else if (s->variant == V_STRIP || s->variant == V_PERSIST) {
        s->tile_x = wv::TILE_X;
        s->tile_y = k.tile_y ? k.tile_y : 16;
    } else {
        s->tile_x = wv::TILE_X;
        s->tile_y = k.tile_y ? k.tile_y : wv::TILE_Y;
    }
    wv::Layout2D &L = s->L;
    L.NX = k.NX;
    L.NY = k.NY;
    L.H = s->H;
    L.ypad = 2 * s->H;
    L.pitch = wv::XPAD + round_up(k.NX + 64, 64) + wv::XPAD;
    L.rows = L.ypad + round_up(k.NY, 32) + 32 + 2 * s->H;
    L.plane = (long long)L.pitch * L.rows;
    s->nfields = s->physics == wv::ELASTIC ? 5 : 3;
    // kernels index inside one field plane and one shot's CPML frame with 32-bit offsets
    if (L.plane >= (1LL << 31) || (long long)k.NY * k.NX >= (1LL << 30))
        return fail(WV_EINVAL, "grid too large for one field plane");
    if (s->variant == V_STRIP && !wv::strip2d_fits(L))
        return fail(WV_EINVAL, "grid too large for the row-strip kernels (field plane of 2 GiB or more)");

    int rc;
    // elastic keeps two field buffers so the tuner may switch between the
    // single-pass (ping-pong) and the two-launch kernels
    const int nbuf = s->physics == wv::ELASTIC ? 2 : 1;
    for (int b = 0; b < nbuf; b++)
        for (int f = 0; f < s->nfields; f++)
            if ((rc = s->dalloc(&s->field[b][f], (size_t)L.plane * k.nshots, s->allocs))) return rc;

    // derivative taps, positions [0, NXp) / [0, NYp) plus ORG on each side
    const int NXp = round_up(k.NX, 64) + 64, NYp = round_up(k.NY, 32) + 32;
    const int KT = 2 * s->H + 1;
    const bool stencil = k.ops_mode == WV_OPS_STENCIL;
    const float DH = (float)k.DH, DT = (float)k.DT;
    std::vector<float> t_xf, t_xb, t_yf, t_yb, t_yfv, t_ybv;
    wv::build_taps(t_xf, k.NX, NXp, N, s->H, true, stencil, DH, DT);
    wv::build_taps(t_xb, k.NX, NXp, N, s->H, false, stencil, DH, DT);
    wv::build_taps(t_yf, k.NY, NYp, N, s->H, true, stencil, DH, DT);
    wv::build_taps(t_yb, k.NY, NYp, N, s->H, false, stencil, DH, DT);
    if (k.free_surface == 1) {
        // velocity-type updates use the image operators (ForwardSolver2Delastic.cpp:168-202)
        // (SH: DybFreeSurface for syz, ForwardSolver2Dsh.cpp:141-146)
        wv::build_taps_free_surface(t_yfv, k.NY, NYp, N, s->H, true, DH, DT);
        if (s->physics == wv::ELASTIC || s->physics == wv::SH)
            wv::build_taps_free_surface(t_ybv, k.NY, NYp, N, s->H, false, DH, DT);
        else
            t_ybv = t_yb;
    } else {
        t_yfv = t_yf;
        t_ybv = t_yb;
    }
    const std::vector<float> *tt[6] = {&t_xf, &t_xb, &t_yf, &t_yb, &t_yfv, &t_ybv};
    for (int q = 0; q < 6; q++)
        if ((rc = s->upload_pos(&s->tap[q], *tt[q], KT, 0.f))) return rc;

    // CPML
    if (k.damping_boundary == 2) {
        const int W = k.boundary_width;
        std::vector<float> a, b, ah, bh;
        std::vector<int> idx;  // frame indices: the kernels compute them (frame_index)
        wv::cpml_axis(k.NX, NXp, W, false, k.npower, k.fc_cpml, k.vmax_cpml, k.DT, k.DH, a, b, ah, bh, idx);
        if ((rc = s->upload_pos(&s->cpx[0], a, 1, 0.f)) || (rc = s->upload_pos(&s->cpx[1], b, 1, 0.f)) ||
            (rc = s->upload_pos(&s->cpx[2], ah, 1, 0.f)) || (rc = s->upload_pos(&s->cpx[3], bh, 1, 0.f)))
            return rc;
        wv::cpml_axis(k.NY, NYp, W, k.free_surface != 0, k.npower, k.fc_cpml, k.vmax_cpml, k.DT, k.DH, a, b, ah, bh,
                      idx);
        if ((rc = s->upload_pos(&s->cpy[0], a, 1, 0.f)) || (rc = s->upload_pos(&s->cpy[1], b, 1, 0.f)) ||
            (rc = s->upload_pos(&s->cpy[2], ah, 1, 0.f)) || (rc = s->upload_pos(&s->cpy[3], bh, 1, 0.f)))
            return rc;
        s->psx_n = (size_t)k.nshots * k.NY * 2 * W;
        s->psy_n = (size_t)k.nshots * 2 * W * k.NX;
        // memory variables of the velocity-type update (ping-ponged by the single-pass kernel)
        // and of the stress-type update; acoustic/tmem need one of each per axis
        const int nv = s->physics == wv::ELASTIC ? 2 : 1;
        for (int b2 = 0; b2 < nbuf; b2++)
            for (int q = 0; q < nv; q++)
                if ((rc = s->dalloc(&s->psv[b2][q], s->psx_n, s->allocs)) ||
                    (rc = s->dalloc(&s->psv[b2][2 + q], s->psy_n, s->allocs)))
                    return rc;
        for (int q = 0; q < nv; q++)
            if ((rc = s->dalloc(&s->pss[q], s->psx_n, s->allocs)) ||
                (rc = s->dalloc(&s->pss[2 + q], s->psy_n, s->allocs)))
                return rc;
        if (s->physics != wv::ELASTIC) {
            // the phased acoustic/tmem kernels use psx[0]/psy[0] for the velocity-type
            // update and psx[1]/psy[1] for the stress-type update
            s->psv[0][1] = s->pss[0];
            s->psv[0][3] = s->pss[2];
        }
    }

    // ABS
    if (k.damping_boundary == 1) {
        const int W = k.boundary_width;
        std::vector<double> prof = wv::abs_profile(W, k.damping_coeff);
        std::vector<float> d((size_t)L.plane, 0.f);
        for (int y = 0; y < k.NY; y++)
            for (int x = 0; x < k.NX; x++) {
                const int dx = std::min(x, k.NX - 1 - x), dy = std::min(y, k.NY - 1 - y);
                const int mn = std::min(dx, dy);
                double v = 1.0;
                if (k.free_surface == 0) {
                    if (mn < W) v = prof[mn];
                } else if (y < W) {
                    if (dx < W) v = prof[dx];
                } else if (mn < W) {
                    v = prof[mn];
                }
                d[(size_t)L.at(x, y)] = (float)v;
            }
        if ((rc = s->upload(&s->damp, d, s->allocs))) return rc;
    }
    if ((rc = s->dalloc(&s->d_tbase, 1, s->allocs))) return rc;
    s->shots.assign(k.nshots, ShotAcq());
    HIP_TRY(hipStreamSynchronize(s->st));
    s->inited = true;
    return WV_OK;
}

int wv_set_model(wv_solver *s, const float *m0, const float *m1, const float *m2)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_set_model before wv_init");
    // (SH needs no velocityP: m0 may be null)
    if (!m2 || (s->physics != wv::SH && !m0) || (s->physics != wv::ACOUSTIC && !m1))
        return fail(WV_EINVAL, "missing model vector");
    if (s->plain) {
        int rc = s->plain->set_model(m0, m1, m2);
        if (rc) return fail(rc, s->plain->error);
        s->model_ok = true;
        return WV_OK;
    }
    if (s->e3) {
        int rc = s->e3->set_model(m0, m1, m2);
        if (rc) return fail(rc, s->e3->error);
        s->model_ok = true;
        return WV_OK;
    }
    if (s->visco && s->vL == 0) return fail(WV_ESTATE, "the viscoelastic model needs wv_set_relaxation first");
    const wv::Layout2D &L = s->L;
    const int NX = L.NX, NY = L.NY;
    auto gi = [&](int x, int y) { return (size_t)x + (size_t)y * NX; };
    // viscoelastic: the moduli given at the reference frequency CenterFrequencyCPML are scaled to
    // the relaxed ones, M / (1 + sum tauP) and mu / (1 + sum tauS) with
    // sum = sum_l w^2 ts_l^2 / (1 + w^2 ts_l^2), w = 2 pi fc, ts_l = 1 / (2 pi f_l)
    // (Viscoelastic.cpp:650-700; Bohlen 2002, eq. 12)
    double vsum = 0.0;
    if (s->visco) {
        const double w = 2.0 * M_PI * s->cfg.fc_cpml;
        for (int l = 0; l < s->vL; l++) {
            const double ts = 1.0 / (2.0 * M_PI * s->vfreq[l]);
            vsum += w * w * ts * ts / (1.0 + w * w * ts * ts);
        }
    }
    if (s->physics == wv::ELASTIC) {
        // kept for the chain rule of the adjoint gradient
        s->model_h[0].assign(m0, m0 + (size_t)NX * NY);
        s->model_h[1].assign(m1, m1 + (size_t)NX * NY);
        s->model_h[2].assign(m2, m2 + (size_t)NX * NY);
    }
    auto li = [&](int x, int y) { return (size_t)L.at(x, y); };
    std::vector<std::vector<float>> c(5, std::vector<float>((size_t)L.plane, 0.f));
    // visco-EM: relaxation times tau_l in the value type, the mean of 1 / tau_l and the
    // per-mechanism factor of Cd (ForwardSolverEM.cpp:100-157), Cd planes (TM: l; TE: l, 4 + l)
    const bool vem = s->visco && (s->physics == wv::TMEM || s->physics == wv::TEMEM);
    float vinv = 0.f, vcdl[4] = {0.f, 0.f, 0.f, 0.f};
    std::vector<std::vector<float>> vcd(8);
    if (vem) {
        const float dtf = (float)s->cfg.DT;
        for (int l = 0; l < s->vL; l++) {
            const float tau = (float)(1.0 / (2.0 * M_PI * (double)(float)s->vfreq[l]));
            vinv = (float)((double)vinv + 1.0 / (double)tau);
            float tv = (float)(1.0 / (1.0 + 0.5 * (double)dtf / (double)tau));
            float den = (float)s->vL * tau;
            den *= tau;
            tv /= den;
            vcdl[l] = tv;
            vcd[l].assign((size_t)L.plane, 0.f);
            if (s->physics == wv::TEMEM) vcd[4 + l].assign((size_t)L.plane, 0.f);
        }
        vinv /= (float)s->vL;
    }
    auto inv_avg = [&](const std::vector<double> &m, std::vector<float> &out, int ax) {
        // 1/((m_i + m_next)/2), last point 1/m_i, inf/nan -> 0 (Modelparameter.cpp:336-446, 633-639)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const bool last = ax == 0 ? (x + 1 >= NX) : (y + 1 >= NY);
                const double a = m[gi(x, y)];
                const double avg = last ? a : 0.5 * (a + m[ax == 0 ? gi(x + 1, y) : gi(x, y + 1)]);
                double v = 1.0 / avg;
                if (!std::isfinite(v)) v = 0.0;
                out[li(x, y)] = (float)v;
            }
    };
    std::vector<double> rho((size_t)NX * NY), M((size_t)NX * NY), mu((size_t)NX * NY);
    if (s->physics == wv::ACOUSTIC || s->physics == wv::ELASTIC) {
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y);
                const double vp = m0[i], r = m2[i];
                rho[i] = r;
                M[i] = r * vp * vp;
                if (s->visco) M[i] /= 1.0 + vsum * s->vtauP[i];
                if (s->physics == wv::ELASTIC) {
                    const double vs = m1[i];
                    double m = r * vs * vs;
                    if (s->visco) m /= 1.0 + vsum * s->vtauS[i];
                    mu[i] = std::max(m, 1.0);  // ModelparameterSeismic.cpp:424
                }
            }
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) c[0][li(x, y)] = (float)M[gi(x, y)];
        if (s->physics == wv::ACOUSTIC) {
            inv_avg(rho, c[1], 0);
            inv_avg(rho, c[2], 1);
        } else {
            for (int y = 0; y < NY; y++)
                for (int x = 0; x < NX; x++) c[1][li(x, y)] = (float)mu[gi(x, y)];
            inv_avg(rho, c[2], 0);
            inv_avg(rho, c[3], 1);
            // harmonic 4-point average of mu (Modelparameter.cpp:449-526, ModelparameterSeismic.cpp:421-433)
            for (int y = 0; y < NY; y++)
                for (int x = 0; x < NX; x++) {
                    const bool inx = x + 1 < NX, iny = y + 1 < NY;
                    double avg;
                    const double i00 = 1.0 / mu[gi(x, y)];
                    if (inx && iny)
                        avg = 0.25 * (i00 + 1.0 / mu[gi(x + 1, y)] + 1.0 / mu[gi(x, y + 1)] + 1.0 / mu[gi(x + 1, y + 1)]);
                    else if (inx)
                        avg = 0.5 * (i00 + 1.0 / mu[gi(x + 1, y)]);
                    else if (iny)
                        avg = 0.5 * (i00 + 1.0 / mu[gi(x, y + 1)]);
                    else
                        avg = i00;
                    double v = 1.0 / avg;
                    if (v < 4.0) v = 0.0;
                    c[4][li(x, y)] = (float)v;
                }
        }
    } else if (s->physics == wv::SH) {
        // SH (SH.cpp:413-433): 1/rho at the node (ModelparameterSeismic.cpp:164-172, in float),
        // mu_xz / mu_yz the harmonic 2-point averages of mu along x / y (calcAveragedSWaveModulus,
        // ModelparameterSeismic.cpp:421-433, with calcAverageMatrixX/Y, Modelparameter.cpp:336-408)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y);
                c[0][li(x, y)] = 1.f / m2[i];
                double m = (double)m2[i] * m1[i] * m1[i];
                if (s->visco) m /= 1.0 + vsum * s->vtauS[i];  // ViscoSH.cpp:555-575
                mu[i] = std::max(m, 1.0);
            }
        for (int ax = 0; ax < 2; ax++)
            for (int y = 0; y < NY; y++)
                for (int x = 0; x < NX; x++) {
                    const bool last = ax == 0 ? (x + 1 >= NX) : (y + 1 >= NY);
                    const double a = 1.0 / mu[gi(x, y)];
                    const double avg = last ? a : 0.5 * (a + 1.0 / mu[ax == 0 ? gi(x + 1, y) : gi(x, y + 1)]);
                    double v = 1.0 / avg;
                    if (v < 4.0) v = 0.0;
                    c[1 + ax][li(x, y)] = (float)v;
                }
    } else if (s->physics == wv::TEMEM) {
        // TE-EM (EMEM.cpp:374-395, ForwardSolver2Demem.cpp:80-92): 1/mu with the 4-point xy average
        // (Modelparameter.cpp:449-526, 633-639; invalid -> 0); epsilon and sigma averaged along x for
        // ex and along y for ey (2-point, the last point alone), then Ca, Cb in the reference's float
        // steps; the averages are formed in float as the float build's sparse products form them
        const double mu0 = 1.2566370614e-6;
        const float eps0f = 8.8541878176e-12f, dtf = (float)s->cfg.DT;
        std::vector<float> eps((size_t)NX * NY);
        for (size_t i = 0; i < eps.size(); i++) eps[i] = m0[i] * eps0f;
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const bool inx = x + 1 < NX, iny = y + 1 < NY;
                const double a = (double)m2[gi(x, y)] * mu0;
                double avg;
                if (inx && iny)
                    avg = 0.25 * (a + (double)m2[gi(x + 1, y)] * mu0 + (double)m2[gi(x, y + 1)] * mu0 +
                                  (double)m2[gi(x + 1, y + 1)] * mu0);
                else if (inx)
                    avg = 0.5 * (a + (double)m2[gi(x + 1, y)] * mu0);
                else if (iny)
                    avg = 0.5 * (a + (double)m2[gi(x, y + 1)] * mu0);
                else
                    avg = a;
                double v = 1.0 / avg;
                if (!std::isfinite(v)) v = 0.0;
                c[0][li(x, y)] = (float)v;
                for (int ax = 0; ax < 2; ax++) {
                    const bool last = ax == 0 ? !inx : !iny;
                    const size_t i = gi(x, y), j = last ? i : (ax == 0 ? gi(x + 1, y) : gi(x, y + 1));
                    float e = eps[i], sg = m1[i];
                    if (!last) {
                        e = 0.5f * eps[i] + 0.5f * eps[j];
                        sg = 0.5f * m1[i] + 0.5f * m1[j];
                    }
                    if (s->visco) {
                        // ViscoEMEM.cpp:409-435: both tau averaged as epsilon and sigma
                        float te = s->vtauP[i], ts = s->vtauS[i];
                        if (!last) {
                            te = 0.5f * s->vtauP[i] + 0.5f * s->vtauP[j];
                            ts = 0.5f * s->vtauS[i] + 0.5f * s->vtauS[j];
                        }
                        float cd[4];
                        wv::em_visco_coef(e, sg, te, ts, dtf, vinv, s->vL, vcdl, c[1 + 2 * ax][li(x, y)],
                                          c[2 + 2 * ax][li(x, y)], cd);
                        for (int l = 0; l < s->vL; l++) vcd[4 * ax + l][li(x, y)] = cd[l];
                        continue;
                    }
                    wv::em_ca_cb(e, sg, dtf, c[1 + 2 * ax][li(x, y)], c[2 + 2 * ax][li(x, y)]);
                }
            }
    } else {
        const double mu0 = 1.2566370614e-6;
        const float eps0f = 8.8541878176e-12f, dtf = (float)s->cfg.DT;
        std::vector<double> perm((size_t)NX * NY);
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y);
                perm[i] = (double)m2[i] * mu0;
                // Ca, Cb in the reference's float arithmetic, step by step as ForwardSolverEM.cpp:14-68
                // (epsilon = epsilon_r * eps0, TMEM.cpp:230-241): the time loop multiplies ez by Ca
                // every step, so a differently rounded Ca drifts from the float build's seismograms
                if (s->visco) {
                    // ForwardSolver2Dviscotmem.cpp:77-100: node values
                    float cd[4];
                    wv::em_visco_coef(m0[i] * eps0f, m1[i], s->vtauP[i], s->vtauS[i], dtf, vinv, s->vL, vcdl,
                                      c[0][li(x, y)], c[1][li(x, y)], cd);
                    for (int l = 0; l < s->vL; l++) vcd[l][li(x, y)] = cd[l];
                    continue;
                }
                wv::em_ca_cb(m0[i] * eps0f, m1[i], dtf, c[0][li(x, y)], c[1][li(x, y)]);
            }
        inv_avg(perm, c[2], 0);  // 1/mu averaged along x (TMEM.cpp:366-373)
        inv_avg(perm, c[3], 1);  // 1/mu averaged along y
    }
    int rc;
    for (int q = 0; q < 5; q++) {
        if (!s->coef[q]) {
            if ((rc = s->dalloc(&s->coef[q], (size_t)L.plane, s->allocs))) return rc;
        }
        HIP_TRY(hipMemcpyAsync(s->coef[q], c[q].data(), c[q].size() * sizeof(float), hipMemcpyHostToDevice, s->st));
    }
    if (vem) {
        for (int k = 0; k < 8; k++) {
            if (vcd[k].empty()) continue;
            if (!s->vcd[k] && (rc = s->dalloc(&s->vcd[k], (size_t)L.plane, s->allocs))) return rc;
            HIP_TRY(hipMemcpyAsync(s->vcd[k], vcd[k].data(), vcd[k].size() * sizeof(float), hipMemcpyHostToDevice,
                                   s->st));
        }
        HIP_TRY(hipStreamSynchronize(s->st));  // vcd is freed at the end of this call
    }
    if (s->visco && s->physics == wv::SH) {
        // visco-SH: tauS averaged along x and y with the 2-point matrices of mu_xz / mu_yz
        // (ViscoSH.cpp:450-460), 1 + L tauS (ForwardSolver2Dviscosh.cpp:70-90)
        std::vector<std::vector<float>> vc(3, std::vector<float>((size_t)L.plane, 0.f));
        const std::vector<float> &ts = s->vtauS;
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y), o = li(x, y);
                vc[0][o] = x + 1 < NX ? (float)(0.5 * ((double)ts[i] + ts[gi(x + 1, y)])) : ts[i];
                vc[1][o] = y + 1 < NY ? (float)(0.5 * ((double)ts[i] + ts[gi(x, y + 1)])) : ts[i];
                vc[2][o] = (float)(1.0 + (double)s->vL * ts[i]);
            }
        const int slot[3] = {5, 6, 4};
        for (int q = 0; q < 3; q++) {
            float *&d = s->vcoef[slot[q]];
            if (!d && (rc = s->dalloc(&d, (size_t)L.plane, s->allocs))) return rc;
            HIP_TRY(hipMemcpyAsync(d, vc[q].data(), vc[q].size() * sizeof(float), hipMemcpyHostToDevice, s->st));
            HIP_TRY(hipStreamSynchronize(s->st));  // vc is freed at the end of this block
        }
    }
    if (s->visco && s->physics == wv::ELASTIC) {
        // per cell tauP, tauS, tauS averaged with the 4-point matrix of mu_xy (Viscoelastic.cpp:
        // 562-575), 1 + L tauP, 1 + L tauS (ForwardSolver2Dviscoelastic.cpp:85-100)
        std::vector<std::vector<float>> vc(5, std::vector<float>((size_t)L.plane, 0.f));
        const std::vector<float> &tp = s->vtauP, &ts = s->vtauS;
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y), o = li(x, y);
                const bool inx = x + 1 < NX, iny = y + 1 < NY;
                double a;
                if (inx && iny)
                    a = 0.25 * ((double)ts[i] + ts[gi(x + 1, y)] + ts[gi(x, y + 1)] + ts[gi(x + 1, y + 1)]);
                else if (inx)
                    a = 0.5 * ((double)ts[i] + ts[gi(x + 1, y)]);
                else if (iny)
                    a = 0.5 * ((double)ts[i] + ts[gi(x, y + 1)]);
                else
                    a = ts[i];
                vc[0][o] = tp[i];
                vc[1][o] = ts[i];
                vc[2][o] = (float)a;
                vc[3][o] = (float)(1.0 + (double)s->vL * tp[i]);
                vc[4][o] = (float)(1.0 + (double)s->vL * ts[i]);
            }
        for (int q = 0; q < 5; q++) {
            if (!s->vcoef[q] && (rc = s->dalloc(&s->vcoef[q], (size_t)L.plane, s->allocs))) return rc;
            HIP_TRY(hipMemcpyAsync(s->vcoef[q], vc[q].data(), vc[q].size() * sizeof(float), hipMemcpyHostToDevice,
                                   s->st));
            HIP_TRY(hipStreamSynchronize(s->st));  // vc is freed at the end of this block
        }
    }
    if (s->physics == wv::ELASTIC && s->cfg.free_surface == 1) {
        // FreeSurfaceElastic.cpp:11-47: scales of the horizontal/vertical updates at y = 0;
        // viscoelastic (FreeSurfaceViscoelastic.cpp:10-60): with the unrelaxed moduli M (1 + L tauP),
        // mu (1 + L tauS), and per mechanism the updates of Rxx
        const int NXp = round_up(NX, 64) + 64;
        std::vector<float> h((size_t)NXp, 0.f), v((size_t)NXp, 0.f), rh((size_t)4 * NXp, 0.f),
            rv((size_t)4 * NXp, 0.f);
        for (int x = 0; x < NX; x++) {
            const size_t i = gi(x, 0);
            if (!s->visco) {
                const double Mx = M[i], t = Mx - 2.0 * mu[i];
                v[x] = (float)t;
                h[x] = (float)(-(t * t) / Mx);
                continue;
            }
            const double oP = 1.0 + (double)s->vL * s->vtauP[i], oS = 1.0 + (double)s->vL * s->vtauS[i];
            const double Mu = M[i] * oP, t = Mu - 2.0 * mu[i] * oS;
            v[x] = (float)t;
            h[x] = (float)(-(t * t) / Mu);
            const double t3 = 2.0 * mu[i] * s->vtauS[i] - M[i] * s->vtauP[i];
            const double ratio = 2.0 * mu[i] * oS / Mu - 1.0;
            for (int l = 0; l < s->vL; l++) {
                const double tau = 1.0 / (2.0 * M_PI * s->vfreq[l]);
                const double c2 = 1.0 / (1.0 + s->cfg.DT / (2.0 * tau));
                rv[(size_t)l * NXp + x] = (float)(t3 * c2 / tau);
                rh[(size_t)l * NXp + x] = (float)(t3 * ratio * c2 / tau);
            }
        }
        if (!s->fsH && ((rc = s->dalloc(&s->fsH, NXp, s->allocs)) || (rc = s->dalloc(&s->fsV, NXp, s->allocs))))
            return rc;
        HIP_TRY(hipMemcpyAsync(s->fsH, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipMemcpyAsync(s->fsV, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, s->st));
        if (s->visco)
            for (int l = 0; l < s->vL; l++) {
                if (!s->vfs[l] && ((rc = s->dalloc(&s->vfs[l], NXp, s->allocs)) ||
                                   (rc = s->dalloc(&s->vfs[4 + l], NXp, s->allocs))))
                    return rc;
                HIP_TRY(hipMemcpyAsync(s->vfs[l], rh.data() + (size_t)l * NXp, NXp * sizeof(float),
                                       hipMemcpyHostToDevice, s->st));
                HIP_TRY(hipMemcpyAsync(s->vfs[4 + l], rv.data() + (size_t)l * NXp, NXp * sizeof(float),
                                       hipMemcpyHostToDevice, s->st));
            }
        HIP_TRY(hipStreamSynchronize(s->st));  // h, v, rh, rv are freed at the end of this block
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    s->model_ok = true;
    return WV_OK;
}

int wv_set_relaxation(wv_solver *s, int nmech, const double *freqs, const float *tauP, const float *tauS)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_set_relaxation before wv_init");
    if (!s->visco) return fail(WV_EINVAL, "relaxation mechanisms belong to the viscoelastic solver");
    // Viscoelastic.cpp:585-596
    if (nmech < 1 || nmech > 4)
        return fail(WV_EINVAL, "The number of relaxation mechanisms should be >0 in an viscoelastic simulation (1..4)");
    if (!freqs || !tauS || (!tauP && s->physics != wv::SH)) return fail(WV_EINVAL, "missing relaxation argument");
    if (s->dim == 2 && (s->physics == wv::TMEM || s->physics == wv::TEMEM) && (!tauP || !tauS))
        return fail(WV_EINVAL, "visco-EM needs tau of the permittivity and of the conductivity");
    for (int l = 0; l < nmech; l++)
        if (!(freqs[l] > 0)) return fail(WV_EINVAL, "The relaxation frequency should be >=0 in an viscoelastic simulation");
    if (s->e3) {
        int rc = s->e3->set_relaxation(nmech, freqs, tauP, tauS);
        if (rc) return fail(rc, s->e3->error);
        s->vL = nmech;
        s->model_ok = false;
        return WV_OK;
    }
    const size_t n = (size_t)s->L.NX * s->L.NY;
    s->vfreq.assign(freqs, freqs + nmech);
    if (tauP)
        s->vtauP.assign(tauP, tauP + n);
    else
        s->vtauP.assign(n, 0.f);  // visco-SH has no tauP
    s->vtauS.assign(tauS, tauS + n);
    // per-mechanism constants in the reference's value type, step by step as
    // ForwardSolver2Dviscoelastic.cpp:50-60 (they multiply the memory variables every step)
    const float DTf = (float)s->cfg.DT;
    for (int l = 0; l < 4; l++) s->vitau[l] = s->vc1[l] = s->vc2[l] = 0.f;
    for (int l = 0; l < nmech; l++) {
        const float tau = (float)(1.0 / (2.0 * M_PI * (double)(float)freqs[l]));
        s->vitau[l] = (float)(1.0 / (double)tau);
        s->vc1[l] = (float)(1.0 - (double)DTf / (2.0 * (double)tau));
        s->vc2[l] = (float)(1.0 / (1.0 + (double)DTf / (2.0 * (double)tau)));
    }
    s->vdth = (float)((double)DTf / 2.0);
    const size_t fn = (size_t)s->L.plane * s->cfg.nshots;
    // Rxz, Ryz | Rxx, Ryy, Rxy | visco-EM: r of ez | r of ex, ey
    const int ngroups = s->physics == wv::SH ? 2 : (s->physics == wv::TMEM ? 1 : (s->physics == wv::TEMEM ? 2 : 3));
    // visco-EM (ForwardSolverEM.cpp:90-100): Cc_l = (1 - DT / (2 tau_l)) / (1 + DT / (2 tau_l)), DT
    for (int l = 0; l < 4; l++) s->vcc[l] = 0.f;
    for (int l = 0; l < nmech; l++) {
        const double tau = (double)(float)(1.0 / (2.0 * M_PI * (double)(float)freqs[l]));
        s->vcc[l] = (float)((1.0 - 0.5 * (double)DTf / tau) / (1.0 + 0.5 * (double)DTf / tau));
    }
    s->vdt = DTf;
    for (int l = 0; l < nmech; l++)
        for (int g = 0; g < ngroups; g++) {
            float *&p = s->vR[4 * g + l];
            if (!p) {
                int rc = s->dalloc(&p, fn, s->allocs);
                if (rc) return rc;
                HIP_TRY(hipMemsetAsync(p, 0, fn * sizeof(float), s->st));
            }
        }
    s->vL = nmech;
    s->model_ok = false;  // the moduli depend on tauP, tauS: wv_set_model follows
    HIP_TRY(hipStreamSynchronize(s->st));
    return WV_OK;
}

int wv_set_acquisition(wv_solver *s, int shot, int nsrc, const long long *src_index, const int *src_type,
                       const float *src_signal, int nrec, const long long *rec_index, const int *rec_type)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_set_acquisition before wv_init");
    if (shot < 0 || shot >= s->cfg.nshots) return fail(WV_EINVAL, "shot index out of range");
    if (nsrc < 0 || nrec < 0) return fail(WV_EINVAL, "negative count");
    if ((nsrc && (!src_index || !src_type || !src_signal)) || (nrec && (!rec_index || !rec_type)))
        return fail(WV_EINVAL, "missing acquisition array");
    const long long n = (long long)s->cfg.NX * s->cfg.NY * s->cfg.NZ;
    const int maxtype =
        s->physics == wv::TMEM ? 1 : (s->dim == 3 || s->physics == wv::SH || s->physics == wv::TEMEM ? 4 : 3);
    // SH has VZ sources and receivers only (FDTD2Dsh.hpp:25-27 throws for P, VX, VY); 2D TE-EM
    // EX, EY, HZ (FDTD2Demem.hpp:29-30 throws for EZ); 3D EM EZ, EX, EY, HZ
    const int mintype = s->physics == wv::SH ? 4 : (s->physics == wv::TEMEM && s->dim == 2 ? 2 : 1);
    for (int k = 0; k < nsrc; k++) {
        if (src_index[k] < 0 || src_index[k] >= n) return fail(WV_EINVAL, "source outside the grid");
        if (src_type[k] < mintype || src_type[k] > maxtype) return fail(WV_EINVAL, "unsupported source type");
    }
    for (int k = 0; k < nrec; k++) {
        if (rec_index[k] < 0 || rec_index[k] >= n) return fail(WV_EINVAL, "receiver outside the grid");
        if (rec_type[k] < mintype || rec_type[k] > maxtype) return fail(WV_EINVAL, "unsupported receiver type");
    }
    ShotAcq a;
    a.sidx.assign(src_index, src_index + nsrc);
    a.stype.assign(src_type, src_type + nsrc);
    a.ssig.assign(src_signal, src_signal + (size_t)nsrc * s->cfg.NT);
    a.ridx.assign(rec_index, rec_index + nrec);
    a.rtype.assign(rec_type, rec_type + nrec);
    if (s->plain) {
        int rc = s->plain->set_acquisition(a.sidx, a.stype, a.ssig, a.ridx, a.rtype);
        if (rc) return fail(rc, s->plain->error);
    }
    if (s->e3) {
        int rc = s->e3->set_acquisition(shot, a.sidx, a.stype, a.ssig, a.ridx, a.rtype);
        if (rc) return fail(rc, s->e3->error);
    }
    s->shots[shot] = std::move(a);
    s->acq_dirty = true;
    return WV_OK;
}

int wv_reset(wv_solver *s)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_reset before wv_init");
    if (s->plain) {
        int rc = s->plain->reset();
        return rc ? fail(rc, s->plain->error) : WV_OK;
    }
    if (s->e3) {
        int rc = s->e3->reset();
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    const size_t fn = (size_t)s->L.plane * s->cfg.nshots;
    for (int b = 0; b < 2; b++)
        for (int f = 0; f < s->nfields; f++)
            if (s->field[b][f]) HIP_TRY(hipMemsetAsync(s->field[b][f], 0, fn * sizeof(float), s->st));
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 4; q++)
            if (s->psv[b][q]) HIP_TRY(hipMemsetAsync(s->psv[b][q], 0, (q < 2 ? s->psx_n : s->psy_n) * 4, s->st));
    for (int q = 0; q < 4; q++)
        if (s->pss[q]) HIP_TRY(hipMemsetAsync(s->pss[q], 0, (q < 2 ? s->psx_n : s->psy_n) * 4, s->st));
    if (s->d_seis) HIP_TRY(hipMemsetAsync(s->d_seis, 0, (size_t)s->cfg.NT * std::max(s->rec_total, 1) * 4, s->st));
    for (int k = 0; k < 12; k++)
        if (s->vR[k]) HIP_TRY(hipMemsetAsync(s->vR[k], 0, fn * sizeof(float), s->st));
    return WV_OK;
}

int wv_run(wv_solver *s, int t_start, int t_end)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_run before wv_init");
    if (!s->model_ok) return fail(WV_ESTATE, "wv_run before wv_set_model");
    if (t_start < 0 || t_end > s->cfg.NT || t_start > t_end) return fail(WV_EINVAL, "time range out of [0, NT]");
    if (s->plain) {
        HIP_TRY(hipEventRecord(s->ev0, s->st));
        int rc = s->plain->run(t_start, t_end);
        if (rc) return fail(rc, s->plain->error);
        HIP_TRY(hipEventRecord(s->ev1, s->st));
        s->ran = true;
        return WV_OK;
    }
    if (s->e3) {
        HIP_TRY(hipEventRecord(s->ev0, s->st));
        int rc = s->e3->run(t_start, t_end, s->ktiming);
        if (rc) return fail(rc, s->e3->error);
        HIP_TRY(hipEventRecord(s->ev1, s->st));
        s->ran = true;
        return WV_OK;
    }
    int rc = s->sync_acquisition();
    if (rc) return rc;
    HIP_TRY(hipEventRecord(s->ev0, s->st));
    int t = t_start;
    auto eager = [&](int n) -> int {
        wv::launch_set(s->d_tbase, t, s->st);
        int r2 = s->enqueue_steps(n, s->ktiming);
        t += n;
        return r2;
    };
    const int K = s->ktiming ? 0 : s->cfg.steps_per_graph;
    const int ck = adjoint_every(s);  // checkpoint interval of the adjoint (0 = off)
    adjoint_run_begin(s, t_start);
    while (t < t_end) {
        if (ck && t % ck == 0 && (rc = adjoint_checkpoint(s, t))) return rc;
        int n = t_end - t;
        if (ck) n = std::min(n, ck - t % ck);
        // (the persistent kernel is one cooperative launch per segment already: no graph)
        if (K > 0 && n >= K + s->cur && s->variant != V_PERSIST) {
            if (s->cur) {  // graphs start in buffer 0
                if ((rc = eager(1))) return rc;
                continue;
            }
            if (!s->gexec || s->gsteps != K) {
                // the single-pass kernel ping-pongs two buffers: a graph must hold an even step count
                if (s->variant == V_SINGLE && (K & 1))
                    return fail(WV_ESTATE, "graph segments of the single-pass kernel need an even step count");
                s->drop_graph();
                hipGraph_t g;
                HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
                rc = s->enqueue_steps(K, false);
                hipError_t e = hipStreamEndCapture(s->st, &g);
                s->cur = 0;
                if (rc) return rc;
                if (e != hipSuccess) return fail(WV_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
                e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return fail(WV_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
                s->gsteps = K;
            }
            // the graph advances the device step counter itself
            wv::launch_set(s->d_tbase, t, s->st);
            const int reps = ck ? 1 : n / K;
            for (int q = 0; q < reps; q++) HIP_TRY(hipGraphLaunch(s->gexec, s->st));
            t += reps * K;
            continue;
        }
        if ((rc = eager(n))) return rc;
    }
    adjoint_run_done(s, t_end);
    HIP_TRY(hipEventRecord(s->ev1, s->st));
    s->ran = true;
    return WV_OK;
}

int wv_synchronize(wv_solver *s)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    HIP_TRY(hipStreamSynchronize(s->st));
    return WV_OK;
}

int wv_last_run_ms(wv_solver *s, double *ms)
{
    if (!s || !s->inited || !s->ran || !ms) return fail(WV_ESTATE, "no run recorded");
    HIP_TRY(hipEventSynchronize(s->ev1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, s->ev0, s->ev1));
    *ms = f;
    return WV_OK;
}

int wv_get_seismograms(wv_solver *s, int shot, float *out)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    if (shot < 0 || shot >= s->cfg.nshots || !out) return fail(WV_EINVAL, "bad shot or output");
    if (s->plain) {
        int rc = s->plain->get_seismograms(out);
        return rc ? fail(rc, s->plain->error) : WV_OK;
    }
    if (s->e3) {
        int rc = s->e3->get_seismograms(shot, out);
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    const int NT = s->cfg.NT;
    const int nrec = (int)s->shots[shot].ridx.size();
    if (nrec == 0) return WV_OK;
    if (s->acq_dirty || !s->d_seis) {
        std::fill(out, out + (size_t)nrec * NT, 0.f);
        return WV_OK;
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    std::vector<float> h((size_t)NT * s->rec_total);
    HIP_TRY(hipMemcpy(h.data(), s->d_seis, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    const int r0 = s->rec_ptr_h[shot];
    for (int r = 0; r < nrec; r++)
        for (int t = 0; t < NT; t++) out[(size_t)r * NT + t] = h[(size_t)t * s->rec_total + r0 + r];
    return WV_OK;
}

static int field_io(wv_solver *s, int shot, int field, float *out, const float *in)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    if (shot < 0 || shot >= s->cfg.nshots) return fail(WV_EINVAL, "bad shot");
    if (s->plain) {
        int rc = s->plain->field_io(field, out, in);
        return rc ? fail(rc, s->plain->error) : WV_OK;
    }
    if (s->e3) {
        int rc = s->e3->field_io(shot, field, out, in);
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    if (field < 0 || field >= s->nfields) return fail(WV_EINVAL, "bad field index");
    const wv::Layout2D &L = s->L;
    std::vector<float> h((size_t)L.plane);
    float *d = s->field[s->cur][field] + (size_t)shot * L.plane;
    HIP_TRY(hipStreamSynchronize(s->st));
    if (out) {
        HIP_TRY(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int y = 0; y < L.NY; y++)
            for (int x = 0; x < L.NX; x++) out[(size_t)y * L.NX + x] = h[(size_t)L.at(x, y)];
    } else {
        std::fill(h.begin(), h.end(), 0.f);
        for (int y = 0; y < L.NY; y++)
            for (int x = 0; x < L.NX; x++) h[(size_t)L.at(x, y)] = in[(size_t)y * L.NX + x];
        HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return WV_OK;
}

int wv_get_wavefield(wv_solver *s, int shot, int field, float *out)
{
    if (!out) return fail(WV_EINVAL, "null output");
    return field_io(s, shot, field, out, nullptr);
}

int wv_set_wavefield(wv_solver *s, int shot, int field, const float *in)
{
    if (!in) return fail(WV_EINVAL, "null input");
    return field_io(s, shot, field, nullptr, in);
}

int wv_tune(wv_solver *s, int nsteps, int *choice)
{
    if (!s || !s->inited || !s->model_ok) return fail(WV_ESTATE, "wv_tune needs an initialised solver with a model");
    if (nsteps < 2 || nsteps > s->cfg.NT) return fail(WV_EINVAL, "nsteps must be in [2, NT]");
    if (choice) *choice = -1;
    if (s->plain) return WV_OK;  // one candidate
    if (s->e3) {
        int rc = s->e3->tune(nsteps, choice);
        if (rc) return fail(rc, s->e3->error);
        if (choice) *choice += 6;
        return WV_OK;
    }
    // candidates: two-launch kernels with 8/16/32-row tiles; for elastic P-SV
    // also the single-pass kernel with register-staged tiles of 8/16/24 rows
    // and the row-strip kernels with 8/16/32-row blocks
    // (and, for TM-EM without ABS, the persistent kernel with 16-row tiles)
    const int cand_var[10] = {V_PHASED, V_PHASED, V_PHASED, V_SINGLE, V_SINGLE, V_SINGLE, V_STRIP, V_STRIP, V_STRIP,
                              V_PERSIST};
    const int cand_ty[10] = {8, 16, 32, 8, 16, 24, 8, 16, 32, 16};
    const int cand_id[10] = {0, 1, 2, 3, 4, 5, 9, 10, 11, 12};
    // (TM-EM: the two-launch and the row-strip kernels)
    const bool all2d = s->physics == wv::ELASTIC && s->cfg.path == WV_PATH_FUSED && !s->visco;
    const bool tm = s->physics == wv::TMEM && s->cfg.path == WV_PATH_FUSED && !s->visco;
    const int ncand = (all2d || tm) ? (wv::strip2d_fits(s->L) ? (tm && s->cfg.damping_boundary != 1 ? 10 : 9) : 6) : 3;
    float best = 1e30f;
    const int k0 = s->cfg.path == WV_PATH_STRIP ? 6 : 0, k1 = s->cfg.path == WV_PATH_STRIP ? 9 : ncand;
    // with checkpointing on, the adjoint recomputes with a two-launch pair: the
    // single-pass kernel (ping-pong buffers, even graph segments) is no candidate
    const int ck_every = adjoint_every(s);
    int best_k = ck_every && k0 == 0 ? 0 : k0;
    for (int k = k0; k < k1; k++) {
        if ((ck_every || s->physics != wv::ELASTIC) && cand_var[k] == V_SINGLE) continue;
        s->variant = cand_var[k];
        s->tile_x = s->variant == V_SINGLE ? wv::fused2d_tile_x(s->cfg.spatial_fd_order) : wv::TILE_X;
        s->tile_y = cand_ty[k];
        s->tile_load = 1;
        s->acq_dirty = true;
        s->cur = 0;
        int rc = wv_reset(s);
        if (rc) return rc;
        if ((rc = s->sync_acquisition())) return rc;
        wv::launch_set(s->d_tbase, 0, s->st);
        if ((rc = s->enqueue_steps(2, false))) return rc;  // warm-up
        wv::launch_set(s->d_tbase, 0, s->st);
        HIP_TRY(hipEventRecord(s->ev0, s->st));
        if ((rc = s->enqueue_steps(nsteps, false))) return rc;
        HIP_TRY(hipEventRecord(s->ev1, s->st));
        HIP_TRY(hipEventSynchronize(s->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        if (ms < best) {
            best = ms;
            best_k = k;
        }
    }
    s->variant = cand_var[best_k];
    s->tile_x = s->variant == V_SINGLE ? wv::fused2d_tile_x(s->cfg.spatial_fd_order) : wv::TILE_X;
    s->tile_y = cand_ty[best_k];
    s->acq_dirty = true;
    s->cur = 0;
    s->drop_graph();
    if (choice) *choice = cand_id[best_k];
    if (ck_every) {
        // the checkpoint / record layout depends on the kernel variant: lay it out
        // again for the variant kept (this also drops any earlier forward run)
        const int fp16 = adjoint_fp16(s);
        adjoint_invalidate_layout(s);
        return wv_set_checkpointing(s, ck_every, fp16);
    }
    return wv_reset(s);
}

int wv_set_kernel_timing(wv_solver *s, int on)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    if (s->plain) return fail(WV_EINVAL, "kernel timing is provided by the tiled kernels");
    s->ktiming = on != 0;
    for (int k = 0; k < 7; k++) {
        s->kms[k] = 0;
        s->klaunch[k] = 0;
    }
    for (int k = 0; k < 3; k++) {
        if (s->e3) {
            s->e3->kms[k] = 0;
            s->e3->klaunch[k] = 0;
        }
    }
    return WV_OK;
}

int wv_kernel_stats(wv_solver *s, int kind, const char **name, double *avg_ms, long long *launches,
                    double *bytes_per_launch)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    if (kind < 0 || kind > 6 || (s->e3 && kind > 2)) return fail(WV_EINVAL, "unknown kernel kind");
    if (s->plain) return fail(WV_EINVAL, "kernel timing is provided by the tiled kernels");
    static const char *names[5][3] = {{"ac2d_velocity", "ac2d_pressure", "flush2d"},
                                      {"el2d_velocity", "el2d_stress", "flush2d"},
                                      {"tm2d_h", "tm2d_e", "flush2d"},
                                      {"sh2d_v", "sh2d_s", "flush2d"},
                                      {"te2d_h", "te2d_e", "flush2d"}};
    static const char *adj_names[4] = {"velocity phase (recompute)", "stress phase (recompute)", "adj_el2d_stress",
                                       "adj_el2d_velocity"};
    HIP_TRY(hipStreamSynchronize(s->st));
    if (s->e3) {
        if (name) *name = s->e3->kernel_name(kind);
        if (launches) *launches = s->e3->klaunch[kind];
        if (avg_ms) *avg_ms = s->e3->klaunch[kind] ? s->e3->kms[kind] / s->e3->klaunch[kind] : 0.0;
        if (bytes_per_launch) *bytes_per_launch = s->e3->kernel_bytes(kind);
        return WV_OK;
    }
    if (name)
        *name = kind >= 3 ? adj_names[kind - 3]
                          : ((s->variant == V_SINGLE && kind < 2)   ? (kind == 0 ? "el2d_fused" : "-")
                             : (s->variant == V_PERSIST && kind < 2) ? (kind == 0 ? "tm2d_persistent" : "")
                             : (s->variant == V_STRIP && kind < 2 && s->physics == wv::TMEM)
                                 ? (kind == 0 ? "tm2d_h_strip" : "tm2d_e_strip")
                             : (s->variant == V_STRIP && kind < 2) ? (kind == 0 ? "el2d_velocity_strip" : "el2d_stress_strip")
                             : (s->visco && kind == 1)              ? (s->physics == wv::SH ? "vsh2d_s" : "ve2d_stress")
                                                                   : names[s->physics][kind]);
    if (launches) *launches = s->klaunch[kind];
    if (avg_ms) *avg_ms = s->klaunch[kind] ? s->kms[kind] / s->klaunch[kind] : 0.0;
    if (bytes_per_launch) *bytes_per_launch = s->kernel_bytes(kind);
    return WV_OK;
}

}  // extern "C"
