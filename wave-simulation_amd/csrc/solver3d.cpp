// Host driver of the 3D elastic solver: layout, coefficient preparation
// (Modelparameter.cpp:336-639 rules, FP64 then one rounding), acquisition
// lists, graph-captured run segments.
#include <cstring>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/wave_abi.h"
#include "elastic3d.hpp"
#include "prep.hpp"

namespace wv {

// Blocks per launch the y chunking aims at: about 48 per CU, so the last round of blocks
// leaves little of the GPU idle (measured on 320^3 x 8 shots: 2048 -> 10.9 ms per step,
// 12288 -> 10.1 ms; each chunk re-reads 2H planes of y halo).
constexpr int kBlocks3D = 12288;


namespace {
constexpr int ORG = 64;  // margin of position-indexed tables (see abi.cpp)
int round_up(int v, int m) { return (v + m - 1) / m * m; }
}  // namespace

#define HIP3(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return fail(WV_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// axis of CPML memory variable q (see Cpml3D; EM: 0 ezy 1 eyz 2 exz 3 ezx 4 eyx 5 exy 6 hzy 7 hyz
// 8 hxz 9 hzx 10 hyx 11 hxy)
int Elastic3D::psi_axis(int q) const
{
    static const int seismic[18] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 1, 0, 2, 0, 2, 1};
    static const int em[12] = {1, 2, 2, 0, 0, 1, 1, 2, 2, 0, 0, 1};
    return physics == TEMEM ? em[q] : seismic[q];
}

int Elastic3D::fail(int code, const std::string &m)
{
    error = m;
    return code;
}

Elastic3D::~Elastic3D()
{
    if (st) (void)hipStreamSynchronize(st);
    free_adjoint();
    drop_graph();
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    free_list(acq_allocs);
    free_list(allocs);
}

void Elastic3D::free_list(std::vector<void *> &l)
{
    for (void *p : l) (void)hipFree(p);
    l.clear();
}

void Elastic3D::drop_graph()
{
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gsteps = 0;
}

int Elastic3D::dalloc(float **p, size_t n)
{
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess)
        return fail(WV_ENOMEM, "hipMalloc failed (3D fields)");
    allocs.push_back(q);
    HIP3(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(float), st));
    *p = (float *)q;
    return 0;
}

template <typename T>
int Elastic3D::upload(T **p, const std::vector<T> &h, std::vector<void *> &list)
{
    void *q = nullptr;
    const size_t n = std::max<size_t>(h.size(), 1);
    if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return fail(WV_ENOMEM, "hipMalloc failed");
    list.push_back(q);
    HIP3(hipMemsetAsync(q, 0, n * sizeof(T), st));
    if (!h.empty()) {
        // pageable source: the copy is complete before the caller may free it
        HIP3(hipMemcpyAsync(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIP3(hipStreamSynchronize(st));
    }
    *p = (T *)q;
    return 0;
}

template <typename T>
int Elastic3D::upload_pos(T **p0, const std::vector<T> &h, int per_pos, T fill)
{
    std::vector<T> ext((size_t)(h.size() / per_pos + 2 * ORG) * per_pos, fill);
    std::copy(h.begin(), h.end(), ext.begin() + (size_t)ORG * per_pos);
    T *base = nullptr;
    int rc = upload(&base, ext, allocs);
    if (rc) return rc;
    *p0 = base + (size_t)ORG * per_pos;
    return 0;
}

int Elastic3D::init(const wv_config &k, hipStream_t stream, int phys, bool visco_)
{
    st = stream;
    physics = phys;
    visco = visco_ && (phys == ELASTIC || phys == TEMEM);
    nf = physics == ACOUSTIC ? 4 : (physics == TEMEM ? 6 : 9);
    nps = physics == ACOUSTIC ? 6 : (physics == TEMEM ? 12 : 18);
    if (k.NX < 2 || k.NY < 2 || k.NZ < 2) return fail(WV_EINVAL, "3D grid needs NX, NY, NZ >= 2");
    if (k.nshots < 1) return fail(WV_EINVAL, "nshots must be >= 1");
    NX = k.NX;
    NY = k.NY;
    NZ = k.NZ;
    NT = k.NT;
    order = k.spatial_fd_order;
    H = order / 2;
    nshots = k.nshots;
    damping = k.damping_boundary;
    free_surface = k.free_surface;
    W = k.boundary_width;
    DT = k.DT;
    DH = k.DH;
    fc_ref = k.fc_cpml;
    steps_per_graph = std::max(0, k.steps_per_graph);
    if (damping && 2 * W > std::min(NX, std::min(NY, NZ))) return fail(WV_EINVAL, "BoundaryWidth out of range");
    if (k.tile_y != 0 && k.tile_y != 4 && k.tile_y != 8 && k.tile_y != 16)
        return fail(WV_EINVAL, "tile_y of the 3D kernels must be 0 (default), 4, 8 or 16");
    rz = k.tile_y ? k.tile_y / 4 : 2;

    L.NX = NX;
    L.NY = NY;
    L.NZ = NZ;
    L.H = H;
    L.pitch = round_up(NX, TILE_X) + 2 * XPAD;
    L.zrows = H + round_up(NZ, 16) + H;
    L.slab = (long long)L.zrows * L.pitch;
    L.vol = (long long)(NY + 2 * H) * L.slab;
    // the kernels address one y plane of a field and one shot's CPML frame of each
    // axis through buffer resources with 32-bit byte offsets
    {
        const long long lim = 1LL << 31, w2 = damping == 2 ? 2LL * W : 0;
        if (L.slab * 4 >= lim || (long long)NY * NZ * w2 * 4 >= lim || w2 * NZ * NX * 4 >= lim ||
            (long long)NY * w2 * NX * 4 >= lim)
            return fail(WV_EINVAL, "grid too large for the 3D kernels (a y plane or a CPML frame of 2 GiB or more)");
    }
    int rc;
    // one plane of slack after the last shot: the kernels' y march issues the loads of the plane
    // after its chunk (2H + 1 planes beyond the last grid plane) before leaving the loop
    for (int f = 0; f < nf; f++)
        if ((rc = dalloc(&F[f], (size_t)L.vol * nshots + (size_t)L.slab))) return rc;

    const bool stencil = k.ops_mode == WV_OPS_STENCIL;
    ops_stencil = stencil;
    const int NXp = round_up(NX, 64) + 64, NYp = NY + 64, NZp = round_up(NZ, 16) + 64;
    const int K = 2 * H + 1;
    const float fDH = (float)DH, fDT = (float)DT;
    std::vector<float> t[8];
    build_taps(t[0], NX, NXp, order, H, true, stencil, fDH, fDT);
    build_taps(t[1], NX, NXp, order, H, false, stencil, fDH, fDT);
    build_taps(t[2], NY, NYp, order, H, true, stencil, fDH, fDT);
    build_taps(t[3], NY, NYp, order, H, false, stencil, fDH, fDT);
    if (free_surface == 1) {
        // velocity update uses the image operators for syy (Dyf) and sxy, syz (Dyb)
        build_taps_free_surface(t[4], NY, NYp, order, H, true, fDH, fDT);
        build_taps_free_surface(t[5], NY, NYp, order, H, false, fDH, fDT);
    } else {
        t[4] = t[2];
        t[5] = t[3];
    }
    build_taps(t[6], NZ, NZp, order, H, true, stencil, fDH, fDT);
    build_taps(t[7], NZ, NZp, order, H, false, stencil, fDH, fDT);
    for (int q = 0; q < 8; q++)
        if ((rc = upload_pos(&tap[q], t[q], K, 0.f))) return rc;

    if (damping == 2) {
        const int n[3] = {NX, NY, NZ}, np[3] = {NXp, NYp, NZp};
        for (int ax = 0; ax < 3; ax++) {
            std::vector<float> a, b, ah, bh;
            std::vector<int> id;
            cpml_axis(n[ax], np[ax], W, ax == 1 && free_surface != 0, k.npower, k.fc_cpml, k.vmax_cpml, DT, DH, a, b,
                      ah, bh, id);
            if ((rc = upload_pos(&cp[ax][0], a, 1, 0.f)) || (rc = upload_pos(&cp[ax][1], b, 1, 0.f)) ||
                (rc = upload_pos(&cp[ax][2], ah, 1, 0.f)) || (rc = upload_pos(&cp[ax][3], bh, 1, 0.f)) ||
                (rc = upload_pos(&idx[ax], id, 1, -1)))
                return rc;
        }
        ps_n[0] = (size_t)nshots * NY * NZ * 2 * W;
        ps_n[1] = (size_t)nshots * 2 * W * NZ * NX;
        ps_n[2] = (size_t)nshots * NY * 2 * W * NX;
        // axis of each memory variable (see Cpml3D)
        for (int q = 0; q < nps; q++)
            if ((rc = dalloc(&ps[q], ps_n[psi_axis(q)]))) return rc;
    }
    {
        // per-plane y tables of the two kernels: taps, CPML y index and coefficients (see Taps3D)
        std::vector<float> ay(NYp, 0.f), by(NYp, 0.f), ayh(NYp, 0.f), byh(NYp, 0.f);
        std::vector<int> yi(NYp, -1);
        if (damping == 2) {
            std::vector<float> a, b, ah, bh;
            std::vector<int> id;
            cpml_axis(NY, NYp, W, free_surface != 0, k.npower, k.fc_cpml, k.vmax_cpml, DT, DH, a, b, ah, bh, id);
            for (int y = 0; y < NYp; y++) {
                yi[y] = id[y];
                if (id[y] >= 0) {
                    ay[y] = a[y];
                    by[y] = b[y];
                    ayh[y] = ah[y];
                    byh[y] = bh[y];
                }
            }
        }
        const int E = 2 * K + 5;
        for (int v = 0; v < 2; v++) {
            const std::vector<float> &tb = v ? t[5] : t[3], &tf = v ? t[4] : t[2];
            std::vector<float> tab((size_t)NYp * E, 0.f);
            for (int y = 0; y < NYp; y++) {
                float *e = &tab[(size_t)y * E];
                for (int j = 0; j < K; j++) {
                    e[j] = tb[(size_t)y * K + j];
                    e[K + j] = tf[(size_t)y * K + j];
                }
                std::memcpy(&e[2 * K], &yi[y], sizeof(int));
                e[2 * K + 1] = ay[y];
                e[2 * K + 2] = by[y];
                e[2 * K + 3] = ayh[y];
                e[2 * K + 4] = byh[y];
            }
            if ((rc = upload_pos(&ytab[v], tab, E, 0.f))) return rc;
        }
    }
    if (damping == 1) {
        // ABS3D.cpp:150-215
        std::vector<double> prof = abs_profile(W, k.damping_coeff);
        std::vector<float> d((size_t)L.vol, 0.f);
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++) {
                    const int dx = std::min(x, NX - 1 - x), dy = std::min(y, NY - 1 - y), dz = std::min(z, NZ - 1 - z);
                    const int mn = std::min(dx, std::min(dy, dz));
                    double v = 1.0;
                    if (free_surface == 0) {
                        if (mn < W) v = prof[mn];
                    } else if (y < W) {
                        if (dz < W || dx < W) v = prof[std::min(dx, dz)];
                    } else if (mn < W) {
                        v = prof[mn];
                    }
                    d[(size_t)L.at(x, y, z)] = (float)v;
                }
        if ((rc = upload(&damp, d, allocs))) return rc;
    }
    {
        std::vector<int> zero(1, 0);
        if ((rc = upload(&d_tbase, zero, allocs))) return rc;
    }
    // chunks along y: enough blocks to fill the chip, chunks of at least 4H planes
    const int nxz = ((NX + TILE_X - 1) / TILE_X) * ((NZ + 4 * rz - 1) / (4 * rz)) * nshots;
    nyc = std::max(1, std::min((kBlocks3D + nxz - 1) / nxz, NY / std::max(4 * H, 8)));
    HIP3(hipEventCreate(&ev0));
    HIP3(hipEventCreate(&ev1));
    shots.assign(nshots, Shot());
    HIP3(hipStreamSynchronize(st));
    return 0;
}

int Elastic3D::set_rz(int r)
{
    rz = r;
    const int nxz = ((NX + TILE_X - 1) / TILE_X) * ((NZ + 4 * rz - 1) / (4 * rz)) * nshots;
    nyc = std::max(1, std::min((kBlocks3D + nxz - 1) / nxz, NY / std::max(4 * H, 8)));
    acq_dirty = true;
    drop_graph();
    return 0;
}

int Elastic3D::set_relaxation(int nmech, const double *freqs, const float *tauP, const float *tauS)
{
    const size_t n = (size_t)NX * NY * NZ;
    vfreq.assign(freqs, freqs + nmech);
    vtauP.assign(tauP, tauP + n);
    vtauS.assign(tauS, tauS + n);
    // per-mechanism constants in the reference's value type, step by step as
    // ForwardSolver3Dviscoelastic.cpp:55-66 (they multiply the memory variables every step)
    const float DTf = (float)DT;
    for (int l = 0; l < 4; l++) vitau[l] = vc1[l] = vc2[l] = 0.f;
    for (int l = 0; l < nmech; l++) {
        const float tau = (float)(1.0 / (2.0 * M_PI * (double)(float)freqs[l]));
        vitau[l] = (float)(1.0 / (double)tau);
        vc1[l] = (float)(1.0 - (double)DTf / (2.0 * (double)tau));
        vc2[l] = (float)(1.0 / (1.0 + (double)DTf / (2.0 * (double)tau)));
    }
    vdth = (float)((double)DTf / 2.0);
    // visco-EM (ForwardSolverEM.cpp:90-100): Cc_l = (1 - DT / (2 tau_l)) / (1 + DT / (2 tau_l)), DT
    for (int l = 0; l < 4; l++) vcc[l] = 0.f;
    for (int l = 0; l < nmech; l++) {
        const double tau = (double)(float)(1.0 / (2.0 * M_PI * (double)(float)freqs[l]));
        vcc[l] = (float)((1.0 - 0.5 * (double)DTf / tau) / (1.0 + 0.5 * (double)DTf / tau));
    }
    vdt = DTf;
    // memory-variable groups per mechanism (six stresses; EM: ex, ey, ez), laid out as the fields
    // (one plane of slack, as they)
    const int ngroups = physics == TEMEM ? 3 : 6;
    for (int l = 0; l < nmech; l++)
        for (int g = 0; g < ngroups; g++) {
            float *&p = vR[4 * g + l];
            if (!p) {
                int rc = dalloc(&p, (size_t)L.vol * nshots + (size_t)L.slab);
                if (rc) return rc;
            }
        }
    vL = nmech;
    model_ok = false;  // the moduli depend on tauP, tauS: set_model follows
    drop_graph();
    HIP3(hipStreamSynchronize(st));
    return 0;
}

Visco3D Elastic3D::visco3d() const
{
    Visco3D V{};
    V.L = vL;
    for (int k = 0; k < 24; k++) V.R[k] = vR[k];
    V.tauP = vcoef[0];
    V.tauS = vcoef[1];
    V.tSxy = vcoef[2];
    V.tSxz = vcoef[3];
    V.tSyz = vcoef[4];
    V.opP = vcoef[5];
    V.opS = vcoef[6];
    for (int l = 0; l < 4; l++) {
        V.fsRH[l] = vfs[l];
        V.fsRV[l] = vfs[4 + l];
        V.itau[l] = vitau[l];
        V.c1[l] = vc1[l];
        V.c2[l] = vc2[l];
    }
    V.dth = vdth;
    for (int k = 0; k < 12; k++) V.cd[k] = vcd[k];
    for (int l = 0; l < 4; l++) V.cc[l] = vcc[l];
    V.dt = vdt;
    return V;
}

int Elastic3D::set_model(const float *vp, const float *vs, const float *rho)
{
    const size_t n = (size_t)NX * NY * NZ;
    auto gi = [&](int x, int y, int z) { return (size_t)x + (size_t)z * NX + (size_t)y * NX * NZ; };
    const bool ac = physics == ACOUSTIC;
    if (visco && vL == 0) return fail(WV_ESTATE, "the viscoelastic model needs wv_set_relaxation first");
    if (physics == TEMEM) return set_model_em(vp, vs, rho);
    model_h[0].assign(vp, vp + n);
    if (!ac) model_h[1].assign(vs, vs + n);
    model_h[2].assign(rho, rho + n);
    // viscoelastic: the moduli given at the reference frequency CenterFrequencyCPML are scaled to
    // the relaxed ones, M / (1 + sum tauP) and mu / (1 + sum tauS) with
    // sum = sum_l w^2 ts_l^2 / (1 + w^2 ts_l^2), w = 2 pi fc, ts_l = 1 / (2 pi f_l)
    // (Viscoelastic.cpp:650-700)
    double vsum = 0.0;
    if (visco) {
        const double w = 2.0 * M_PI * fc_ref;
        for (int l = 0; l < vL; l++) {
            const double ts = 1.0 / (2.0 * M_PI * vfreq[l]);
            vsum += w * w * ts * ts / (1.0 + w * w * ts * ts);
        }
    }
    std::vector<double> r(n), M(n), mu(n, 1.0), imu(n, 1.0);
    for (size_t i = 0; i < n; i++) {
        r[i] = rho[i];
        M[i] = (double)rho[i] * vp[i] * vp[i];
        if (visco) M[i] /= 1.0 + vsum * vtauP[i];
        if (ac) continue;
        double m = (double)rho[i] * vs[i] * vs[i];
        if (visco) m /= 1.0 + vsum * vtauS[i];
        mu[i] = std::max(m, 1.0);  // ModelparameterSeismic.cpp:424
        imu[i] = 1.0 / mu[i];
    }
    std::vector<std::vector<float>> c(8, std::vector<float>((size_t)L.vol, 0.f));
    const int len[3] = {NX, NY, NZ};
    auto nb = [&](int x, int y, int z, int ax, int d) {
        int p[3] = {x, y, z};
        p[ax] += d;
        return gi(p[0], p[1], p[2]);
    };
    for (int y = 0; y < NY; y++)
        for (int z = 0; z < NZ; z++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y, z), o = (size_t)L.at(x, y, z);
                const int pos[3] = {x, y, z};
                c[3][o] = (float)M[i];
                c[4][o] = (float)mu[i];
                // 1/((rho_i + rho_next)/2), last point 1/rho_i (Modelparameter.cpp:336-446, 633-639)
                for (int ax = 0; ax < 3; ax++) {
                    const bool last = pos[ax] + 1 >= len[ax];
                    const double avg = last ? r[i] : 0.5 * (r[i] + r[nb(x, y, z, ax, 1)]);
                    double v = 1.0 / avg;
                    if (!std::isfinite(v)) v = 0.0;
                    c[ax][o] = (float)v;
                }
                // harmonic 4-point averages of mu in the xy, xz, yz planes
                // (Modelparameter.cpp:449-526, ModelparameterSeismic.cpp:421-433)
                static const int planes[3][2] = {{0, 1}, {0, 2}, {1, 2}};
                for (int pl = 0; pl < (ac ? 0 : 3); pl++) {
                    const int a0 = planes[pl][0], a1 = planes[pl][1];
                    const bool in0 = pos[a0] + 1 < len[a0], in1 = pos[a1] + 1 < len[a1];
                    double avg;
                    if (in0 && in1) {
                        int p[3] = {x, y, z};
                        p[a0]++;
                        p[a1]++;
                        avg = 0.25 * (imu[i] + imu[nb(x, y, z, a0, 1)] + imu[nb(x, y, z, a1, 1)] + imu[gi(p[0], p[1], p[2])]);
                    } else if (in0) {
                        avg = 0.5 * (imu[i] + imu[nb(x, y, z, a0, 1)]);
                    } else if (in1) {
                        avg = 0.5 * (imu[i] + imu[nb(x, y, z, a1, 1)]);
                    } else {
                        avg = imu[i];
                    }
                    double v = 1.0 / avg;
                    if (v < 4.0) v = 0.0;
                    c[5 + pl][o] = (float)v;
                }
            }
    int rc;
    for (int q = 0; q < (ac ? 4 : 8); q++) {
        if (!coef[q] && (rc = dalloc(&coef[q], (size_t)L.vol))) return rc;
        HIP3(hipMemcpyAsync(coef[q], c[q].data(), c[q].size() * sizeof(float), hipMemcpyHostToDevice, st));
    }
    HIP3(hipStreamSynchronize(st));  // c is reused below
    if (visco) {
        // per cell tauP, tauS, tauS averaged with the 4-point matrices of mu_xy, mu_xz, mu_yz
        // (Viscoelastic.cpp:562-576, Modelparameter.cpp:449-526), 1 + L tauP, 1 + L tauS
        // (ForwardSolver3Dviscoelastic.cpp:105-118)
        const int len[3] = {NX, NY, NZ};
        static const int planes[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int q = 0; q < 7; q++) std::fill(c[q].begin(), c[q].end(), 0.f);
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++) {
                    const size_t i = gi(x, y, z), o = (size_t)L.at(x, y, z);
                    const int pos[3] = {x, y, z};
                    c[0][o] = vtauP[i];
                    c[1][o] = vtauS[i];
                    for (int pl = 0; pl < 3; pl++) {
                        const int a0 = planes[pl][0], a1 = planes[pl][1];
                        const bool in0 = pos[a0] + 1 < len[a0], in1 = pos[a1] + 1 < len[a1];
                        int p0[3] = {x, y, z}, p1[3] = {x, y, z}, p2[3] = {x, y, z};
                        p0[a0]++;
                        p1[a1]++;
                        p2[a0]++;
                        p2[a1]++;
                        double a;
                        if (in0 && in1)
                            a = 0.25 * ((double)vtauS[i] + vtauS[gi(p0[0], p0[1], p0[2])] +
                                        vtauS[gi(p1[0], p1[1], p1[2])] + vtauS[gi(p2[0], p2[1], p2[2])]);
                        else if (in0)
                            a = 0.5 * ((double)vtauS[i] + vtauS[gi(p0[0], p0[1], p0[2])]);
                        else if (in1)
                            a = 0.5 * ((double)vtauS[i] + vtauS[gi(p1[0], p1[1], p1[2])]);
                        else
                            a = vtauS[i];
                        c[2 + pl][o] = (float)a;
                    }
                    c[5][o] = (float)(1.0 + (double)vL * vtauP[i]);
                    c[6][o] = (float)(1.0 + (double)vL * vtauS[i]);
                }
        for (int q = 0; q < 7; q++) {
            if (!vcoef[q] && (rc = dalloc(&vcoef[q], (size_t)L.vol))) return rc;
            HIP3(hipMemcpyAsync(vcoef[q], c[q].data(), c[q].size() * sizeof(float), hipMemcpyHostToDevice, st));
        }
        HIP3(hipStreamSynchronize(st));
    }
    if (free_surface == 1 && !ac) {
        // FreeSurfaceElastic.cpp:11-47: scales of the horizontal update at y = 0; viscoelastic
        // (FreeSurfaceViscoelastic.cpp:10-60): with the unrelaxed moduli M (1 + L tauP),
        // mu (1 + L tauS), and per mechanism the updates of Rxx, Rzz
        std::vector<float> h((size_t)L.slab, 0.f), v((size_t)L.slab, 0.f), rh((size_t)4 * L.slab, 0.f),
            rv((size_t)4 * L.slab, 0.f);
        for (int z = 0; z < NZ; z++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, 0, z), o = (size_t)(z + H) * L.pitch + x + XPAD;
                if (!visco) {
                    const double t = M[i] - 2.0 * mu[i];
                    v[o] = (float)t;
                    h[o] = (float)(-(t * t) / M[i]);
                    continue;
                }
                const double oP = 1.0 + (double)vL * vtauP[i], oS = 1.0 + (double)vL * vtauS[i];
                const double Mu = M[i] * oP, t = Mu - 2.0 * mu[i] * oS;
                v[o] = (float)t;
                h[o] = (float)(-(t * t) / Mu);
                const double t3 = 2.0 * mu[i] * vtauS[i] - M[i] * vtauP[i];
                const double ratio = 2.0 * mu[i] * oS / Mu - 1.0;
                for (int l = 0; l < vL; l++) {
                    const double tau = 1.0 / (2.0 * M_PI * vfreq[l]);
                    const double c2 = 1.0 / (1.0 + DT / (2.0 * tau));
                    rv[(size_t)l * L.slab + o] = (float)(t3 * c2 / tau);
                    rh[(size_t)l * L.slab + o] = (float)(t3 * ratio * c2 / tau);
                }
            }
        if (!fsH && ((rc = dalloc(&fsH, (size_t)L.slab)) || (rc = dalloc(&fsV, (size_t)L.slab)))) return rc;
        HIP3(hipMemcpyAsync(fsH, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
        HIP3(hipMemcpyAsync(fsV, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, st));
        for (int l = 0; visco && l < vL; l++) {
            if (!vfs[l] && ((rc = dalloc(&vfs[l], (size_t)L.slab)) || (rc = dalloc(&vfs[4 + l], (size_t)L.slab))))
                return rc;
            HIP3(hipMemcpyAsync(vfs[l], rh.data() + (size_t)l * L.slab, (size_t)L.slab * sizeof(float),
                                hipMemcpyHostToDevice, st));
            HIP3(hipMemcpyAsync(vfs[4 + l], rv.data() + (size_t)l * L.slab, (size_t)L.slab * sizeof(float),
                                hipMemcpyHostToDevice, st));
        }
        HIP3(hipStreamSynchronize(st));  // h, v, rh, rv are freed at the end of this block
    }
    HIP3(hipStreamSynchronize(st));
    model_ok = true;
    return 0;
}

// 3D EM (EMEM.cpp:378-395, ForwardSolver3Demem.cpp:85-100): m0 = relative permittivity, m1 =
// conductivity, m2 = relative permeability.  1/mu with the 4-point yz, xz, xy averages (invalid ->
// 0, Modelparameter.cpp:449-526, 633-639); epsilon and sigma with the 2-point averages along x, y, z
// (the last point alone), formed in float as the float build's sparse products form them, then Ca,
// Cb in the reference's float steps.  Visco-EM (ViscoEMEM.cpp:409-435, ForwardSolver3Dviscoemem.cpp:
// 78-135): both tau averaged the same way, Ca, Cb of the effective optical values and Cd_l.
int Elastic3D::set_model_em(const float *m0, const float *m1, const float *m2)
{
    auto gi = [&](int x, int y, int z) { return (size_t)x + (size_t)z * NX + (size_t)y * NX * NZ; };
    const size_t n = (size_t)NX * NY * NZ;
    const double mu0 = 1.2566370614e-6;
    const float eps0f = 8.8541878176e-12f, dtf = (float)DT;
    std::vector<float> eps(n);
    for (size_t i = 0; i < n; i++) eps[i] = m0[i] * eps0f;
    float vinv = 0.f, vcdl[4] = {0.f, 0.f, 0.f, 0.f};
    if (visco) {
        for (int l = 0; l < vL; l++) {
            const float tau = (float)(1.0 / (2.0 * M_PI * (double)(float)vfreq[l]));
            vinv = (float)((double)vinv + 1.0 / (double)tau);
            float tv = (float)(1.0 / (1.0 + 0.5 * (double)dtf / (double)tau));
            float den = (float)vL * tau;
            den *= tau;
            tv /= den;
            vcdl[l] = tv;
        }
        vinv /= (float)vL;
    }
    std::vector<std::vector<float>> c(9, std::vector<float>((size_t)L.vol, 0.f));
    std::vector<std::vector<float>> cd(visco ? 3 * vL : 0, std::vector<float>((size_t)L.vol, 0.f));
    const int len[3] = {NX, NY, NZ};
    static const int planes[3][2] = {{1, 2}, {0, 2}, {0, 1}};  // hx: yz, hy: xz, hz: xy
    for (int y = 0; y < NY; y++)
        for (int z = 0; z < NZ; z++)
            for (int x = 0; x < NX; x++) {
                const size_t i = gi(x, y, z), o = (size_t)L.at(x, y, z);
                const int pos[3] = {x, y, z};
                for (int pl = 0; pl < 3; pl++) {
                    const int a0 = planes[pl][0], a1 = planes[pl][1];
                    const bool in0 = pos[a0] + 1 < len[a0], in1 = pos[a1] + 1 < len[a1];
                    int p0[3] = {x, y, z}, p1[3] = {x, y, z}, p2[3] = {x, y, z};
                    p0[a0]++;
                    p1[a1]++;
                    p2[a0]++;
                    p2[a1]++;
                    auto mu = [&](const int *p) { return (double)m2[gi(p[0], p[1], p[2])] * mu0; };
                    const double m00 = (double)m2[i] * mu0;
                    double avg;
                    if (in0 && in1)
                        avg = 0.25 * (m00 + mu(p0) + mu(p1) + mu(p2));
                    else if (in0)
                        avg = 0.5 * (m00 + mu(p0));
                    else if (in1)
                        avg = 0.5 * (m00 + mu(p1));
                    else
                        avg = m00;
                    double v = 1.0 / avg;
                    if (!std::isfinite(v)) v = 0.0;
                    c[pl][o] = (float)v;
                }
                for (int ax = 0; ax < 3; ax++) {
                    const bool last = pos[ax] + 1 >= len[ax];
                    int pj[3] = {x, y, z};
                    pj[ax]++;
                    const size_t j = last ? i : gi(pj[0], pj[1], pj[2]);
                    float e = eps[i], sg = m1[i];
                    if (!last) {
                        e = 0.5f * eps[i] + 0.5f * eps[j];
                        sg = 0.5f * m1[i] + 0.5f * m1[j];
                    }
                    if (visco) {
                        float te = vtauP[i], ts = vtauS[i];
                        if (!last) {
                            te = 0.5f * vtauP[i] + 0.5f * vtauP[j];
                            ts = 0.5f * vtauS[i] + 0.5f * vtauS[j];
                        }
                        float cdv[4];
                        em_visco_coef(e, sg, te, ts, dtf, vinv, vL, vcdl, c[3 + 2 * ax][o], c[4 + 2 * ax][o], cdv);
                        for (int l = 0; l < vL; l++) cd[ax * vL + l][o] = cdv[l];
                    } else {
                        em_ca_cb(e, sg, dtf, c[3 + 2 * ax][o], c[4 + 2 * ax][o]);
                    }
                }
            }
    int rc;
    for (int q = 0; q < 3; q++) {
        if (!coef[q] && (rc = dalloc(&coef[q], (size_t)L.vol))) return rc;
        HIP3(hipMemcpyAsync(coef[q], c[q].data(), c[q].size() * sizeof(float), hipMemcpyHostToDevice, st));
    }
    for (int q = 0; q < 6; q++) {
        if (!emc[q] && (rc = dalloc(&emc[q], (size_t)L.vol))) return rc;
        HIP3(hipMemcpyAsync(emc[q], c[3 + q].data(), c[3 + q].size() * sizeof(float), hipMemcpyHostToDevice, st));
    }
    for (int ax = 0; visco && ax < 3; ax++)
        for (int l = 0; l < vL; l++) {
            float *&d = vcd[4 * ax + l];
            if (!d && (rc = dalloc(&d, (size_t)L.vol))) return rc;
            const std::vector<float> &h = cd[ax * vL + l];
            HIP3(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
        }
    HIP3(hipStreamSynchronize(st));  // c and cd are freed at the end of this call
    model_ok = true;
    return 0;
}

int Elastic3D::set_acquisition(int shot, const std::vector<long long> &sidx, const std::vector<int> &stype,
                               const std::vector<float> &ssig, const std::vector<long long> &ridx,
                               const std::vector<int> &rtype)
{
    Shot &s = shots[shot];
    s.sidx = sidx;
    s.stype = stype;
    s.ssig = ssig;
    s.ridx = ridx;
    s.rtype = rtype;
    acq_dirty = true;
    return 0;
}

int Elastic3D::nrec(int shot) const { return (int)shots[shot].ridx.size(); }

int Elastic3D::sync_acquisition()
{
    if (!acq_dirty) return 0;
    drop_graph();
    HIP3(hipStreamSynchronize(st));
    free_list(acq_allocs);
    const int TZ = 4 * rz, LY = (NY + nyc - 1) / nyc;
    nbx = (NX + TILE_X - 1) / TILE_X;
    nbz = (NZ + TZ - 1) / TZ;
    const size_t nsb = (size_t)nshots * nyc * nbz * nbx;
    std::vector<int> sp(1, 0), rp(1, 0), stp, rtp;
    std::vector<long long> sc, rc_;
    std::vector<float> sig;
    std::vector<std::vector<int>> bs(nsb), br(nsb);
    auto cell = [&](long long g, int &x, int &y, int &z) {
        x = (int)(g % NX);
        z = (int)((g / NX) % NZ);
        y = (int)(g / ((long long)NX * NZ));
    };
    int gs = 0, gr = 0;
    for (int sh = 0; sh < nshots; sh++) {
        const Shot &s = shots[sh];
        auto blk = [&](int x, int y, int z) {
            return (((size_t)sh * nyc + y / LY) * nbz + z / TZ) * nbx + x / TILE_X;
        };
        for (size_t k = 0; k < s.sidx.size(); k++, gs++) {
            int x, y, z;
            cell(s.sidx[k], x, y, z);
            sc.push_back(L.at(x, y, z));
            stp.push_back(s.stype[k]);
            bs[blk(x, y, z)].push_back(gs);
        }
        sig.insert(sig.end(), s.ssig.begin(), s.ssig.end());
        for (size_t k = 0; k < s.ridx.size(); k++, gr++) {
            int x, y, z;
            cell(s.ridx[k], x, y, z);
            rc_.push_back(L.at(x, y, z));
            rtp.push_back(s.rtype[k]);
            br[blk(x, y, z)].push_back(gr);
        }
        sp.push_back(gs);
        rp.push_back(gr);
    }
    rec_total = gr;
    rec_ptr_h = rp;
    std::vector<int> bsp(1, 0), bsl, brp(1, 0), brl;
    for (size_t b = 0; b < nsb; b++) {
        bsl.insert(bsl.end(), bs[b].begin(), bs[b].end());
        bsp.push_back((int)bsl.size());
        brl.insert(brl.end(), br[b].begin(), br[b].end());
        brp.push_back((int)brl.size());
    }
    int rc;
    std::vector<float> zero_seis((size_t)NT * std::max(rec_total, 1), 0.f);
    if ((rc = upload(&d_bsrc_ptr, bsp, acq_allocs)) || (rc = upload(&d_bsrc, bsl, acq_allocs)) ||
        (rc = upload(&d_brec_ptr, brp, acq_allocs)) || (rc = upload(&d_brec, brl, acq_allocs)) ||
        (rc = upload(&d_src_ptr, sp, acq_allocs)) || (rc = upload(&d_rec_ptr, rp, acq_allocs)) ||
        (rc = upload(&d_src_cell, sc, acq_allocs)) || (rc = upload(&d_rec_cell, rc_, acq_allocs)) ||
        (rc = upload(&d_src_type, stp, acq_allocs)) || (rc = upload(&d_rec_type, rtp, acq_allocs)) ||
        (rc = upload(&d_src_sig, sig, acq_allocs)) || (rc = upload(&d_seis, zero_seis, acq_allocs)))
        return rc;
    HIP3(hipStreamSynchronize(st));
    acq_dirty = false;
    return 0;
}

int Elastic3D::reset()
{
    for (int f = 0; f < nf; f++) HIP3(hipMemsetAsync(F[f], 0, (size_t)L.vol * nshots * sizeof(float), st));
    for (int q = 0; q < nps; q++)
        if (ps[q]) HIP3(hipMemsetAsync(ps[q], 0, ps_n[psi_axis(q)] * sizeof(float), st));
    if (d_seis) HIP3(hipMemsetAsync(d_seis, 0, (size_t)NT * std::max(rec_total, 1) * sizeof(float), st));
    for (int k = 0; k < 24; k++)
        if (vR[k]) HIP3(hipMemsetAsync(vR[k], 0, (size_t)L.vol * nshots * sizeof(float), st));
    return 0;
}

Fields3D Elastic3D::fields() const
{
    Fields3D f{};
    for (int k = 0; k < 9; k++) f.f[k] = F[k];
    return f;
}

Coef3D Elastic3D::coefs() const
{
    Coef3D c{};
    c.irx = coef[0];
    c.iry = coef[1];
    c.irz = coef[2];
    c.M = coef[3];
    c.mu = coef[4];
    c.muxy = coef[5];
    c.muxz = coef[6];
    c.muyz = coef[7];
    c.damp = damping == 1 ? damp : nullptr;
    c.fsH = free_surface == 1 ? fsH : nullptr;
    c.fsV = free_surface == 1 ? fsV : nullptr;
    c.fs = free_surface == 1;
    for (int k = 0; k < 6; k++) c.em[k] = emc[k];
    return c;
}

Taps3D Elastic3D::taps() const
{
    Taps3D t{};
    t.xf = tap[0];
    t.xb = tap[1];
    t.yf = tap[2];
    t.yb = tap[3];
    t.yf_v = tap[4];
    t.yb_v = tap[5];
    t.zf = tap[6];
    t.zb = tap[7];
    t.ys = ytab[0];
    t.yv = ytab[1];
    return t;
}

Cpml3D Elastic3D::cpml() const
{
    Cpml3D p{};
    p.on = damping == 2;
    p.W = W;
    p.ax = cp[0][0];
    p.bx = cp[0][1];
    p.axh = cp[0][2];
    p.bxh = cp[0][3];
    p.ay = cp[1][0];
    p.by = cp[1][1];
    p.ayh = cp[1][2];
    p.byh = cp[1][3];
    p.az = cp[2][0];
    p.bz = cp[2][1];
    p.azh = cp[2][2];
    p.bzh = cp[2][3];
    p.xi = idx[0];
    p.yi = idx[1];
    p.zi = idx[2];
    for (int q = 0; q < 18; q++) p.ps[q] = ps[q];
    return p;
}

Acq Elastic3D::acq() const
{
    Acq A{};
    A.nshots = nshots;
    A.NT = NT;
    A.nbx = nbx;
    A.nby = nbz;
    A.src_ptr = d_src_ptr;
    A.src_cell = d_src_cell;
    A.src_type = d_src_type;
    A.src_sig = d_src_sig;
    A.rec_ptr = d_rec_ptr;
    A.rec_cell = d_rec_cell;
    A.rec_type = d_rec_type;
    A.rec_total = rec_total;
    A.seis = d_seis;
    A.bsrc_ptr = d_bsrc_ptr;
    A.bsrc = d_bsrc;
    A.brec_ptr = d_brec_ptr;
    A.brec = d_brec;
    A.tbase = d_tbase;
    return A;
}

// nsteps time steps from the device step counter; the final flush applies the
// velocity-type sources / receivers of the last step.
int Elastic3D::enqueue(int nsteps, bool timed)
{
    const Fields3D Fd = fields();
    const Coef3D C = coefs();
    const Taps3D T = taps();
    const Cpml3D P = cpml();
    const Acq A = acq();
    constexpr int BATCH = 32;
    if (timed && kev.empty()) {
        kev.resize(4 * BATCH);
        for (auto &e : kev) HIP3(hipEventCreate(&e));
    }
    int pending = 0;
    auto collect = [&]() -> int {
        HIP3(hipEventSynchronize(kev[4 * (pending - 1) + 3]));
        for (int b = 0; b < pending; b++)
            for (int k = 0; k < 2; k++) {
                float ms = 0;
                HIP3(hipEventElapsedTime(&ms, kev[4 * b + 2 * k], kev[4 * b + 2 * k + 1]));
                kms[k] += ms;
                klaunch[k] += 1;
            }
        pending = 0;
        return 0;
    };
    const Visco3D V = visco3d();
    for (int j = 0; j < nsteps; j++) {
        hipEvent_t *e = timed ? &kev[4 * pending] : nullptr;
        hipEvent_t e0 = e ? e[0] : nullptr, e1 = e ? e[1] : nullptr, e2 = e ? e[2] : nullptr, e3 = e ? e[3] : nullptr;
        if (physics == ACOUSTIC) {
            launch_ac3d(0, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, j > 0, st, e0, e1);
            launch_ac3d(1, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, 0, st, e2, e3);
        } else if (physics == TEMEM) {
            launch_em3d(0, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, j > 0, st, e0, e1);
            launch_em3d(1, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, 0, st, e2, e3, visco ? &V : nullptr);
        } else {
            launch_el3d(0, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, j > 0, st, e0, e1);
            launch_el3d(1, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, 0, st, e2, e3, nullptr, visco ? &V : nullptr);
        }
        if (damping == 1) launch_abs_velocity3d(L, nshots, Fd, C, st);
        if (timed && ++pending == BATCH) {
            int rc = collect();
            if (rc) return rc;
        }
    }
    if (timed && pending) {
        int rc = collect();
        if (rc) return rc;
    }
    if (nsteps > 0) launch_flush3d(L, Fd, A, nsteps - 1, st, physics == TEMEM);
    launch_advance(d_tbase, nsteps, st);
    HIP3(hipGetLastError());
    return 0;
}

int Elastic3D::run_segment(int t_start, int t_end, bool timed)
{
    int rc;
    int t = t_start;
    const int K = timed ? 0 : steps_per_graph;
    if (K > 0 && t_end - t >= K) {
        if (!gexec || gsteps != K) {
            drop_graph();
            hipGraph_t g;
            HIP3(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = enqueue(K, false);
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) return rc;
            if (e != hipSuccess) return fail(WV_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(WV_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
            gsteps = K;
        }
        launch_set(d_tbase, t, st);
        while (t_end - t >= K) {
            HIP3(hipGraphLaunch(gexec, st));
            t += K;
        }
    }
    if (t_end > t) {
        launch_set(d_tbase, t, st);
        if ((rc = enqueue(t_end - t, timed))) return rc;
    }
    return 0;
}

int Elastic3D::run(int t_start, int t_end, bool timed)
{
    if (!model_ok) return fail(WV_ESTATE, "wv_run before wv_set_model");
    int rc = sync_acquisition();
    if (rc) return rc;
    const int every = adj3 ? checkpoint_interval() : 0;
    if (!every) return run_segment(t_start, t_end, timed);
    // checkpointed: segments end on multiples of the interval; the state at the start of
    // each segment (all steps before it done and flushed) is saved first
    for (int t = t_start; t < t_end;) {
        if (t % every == 0 && (rc = save_checkpoint(t / every))) return rc;
        const int next = std::min(t_end, (t / every + 1) * every);
        if ((rc = run_segment(t, next, timed))) return rc;
        t = next;
    }
    return 0;
}

int Elastic3D::get_seismograms(int shot, float *out)
{
    const int nr = nrec(shot);
    if (nr == 0) return 0;
    if (acq_dirty || !d_seis) {
        std::fill(out, out + (size_t)nr * NT, 0.f);
        return 0;
    }
    HIP3(hipStreamSynchronize(st));
    std::vector<float> h((size_t)NT * rec_total);
    HIP3(hipMemcpy(h.data(), d_seis, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    const int r0 = rec_ptr_h[shot];
    for (int r = 0; r < nr; r++)
        for (int t = 0; t < NT; t++) out[(size_t)r * NT + t] = h[(size_t)t * rec_total + r0 + r];
    return 0;
}

void Elastic3D::buffers(wv::StateBuffers &out) const
{
    for (int f = 0; f < nf; f++) out.add(F[f], (long long)L.vol * nshots);
    if (d_seis && !acq_dirty) out.add(d_seis, (long long)NT * rec_total);
}

int Elastic3D::field_io(int shot, int field, float *out, const float *in)
{
    if (field < 0 || field >= nf) return fail(WV_EINVAL, "bad field index");
    std::vector<float> h((size_t)L.vol);
    float *d = F[field] + (size_t)shot * L.vol;
    HIP3(hipStreamSynchronize(st));
    if (out) {
        HIP3(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++)
                    out[(size_t)x + (size_t)z * NX + (size_t)y * NX * NZ] = h[(size_t)L.at(x, y, z)];
    } else {
        std::fill(h.begin(), h.end(), 0.f);
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++)
                    h[(size_t)L.at(x, y, z)] = in[(size_t)x + (size_t)z * NX + (size_t)y * NX * NZ];
        HIP3(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

int Elastic3D::tune(int nsteps, int *choice)
{
    if (!model_ok) return fail(WV_ESTATE, "wv_tune needs a model");
    const int cand[3] = {1, 2, 4};
    float best = 1e30f;
    int best_k = 1;
    for (int k = 0; k < 3; k++) {
        set_rz(cand[k]);
        int rc;
        if ((rc = reset()) || (rc = sync_acquisition())) return rc;
        launch_set(d_tbase, 0, st);
        if ((rc = enqueue(2, false))) return rc;
        launch_set(d_tbase, 0, st);
        HIP3(hipEventRecord(ev0, st));
        if ((rc = enqueue(nsteps, false))) return rc;
        HIP3(hipEventRecord(ev1, st));
        HIP3(hipEventSynchronize(ev1));
        float ms = 0;
        HIP3(hipEventElapsedTime(&ms, ev0, ev1));
        if (ms < best) {
            best = ms;
            best_k = k;
        }
    }
    set_rz(cand[best_k]);
    if (choice) *choice = best_k;
    return reset();
}

// Algorithmic bytes of one launch: velocity = 6 stresses + 3 velocities read,
// 3 velocities written, 3 coefficients; stress = 3 velocities + 6 stresses
// read, 6 stresses written, 5 coefficients; CPML memory variables r+w on the
// frame cells of their axis.
double Elastic3D::kernel_bytes(int kind) const
{
    const double cells = (double)NX * NY * NZ, B = nshots;
    if (kind == 2) {
        double ns = 0;
        for (auto &s : shots) ns += s.sidx.size();
        return 12.0 * ns + 8.0 * rec_total;
    }
    // per shot-cell field reads + writes, per-cell coefficient reads, memory variables per frame cell
    // elastic: velocity 6 + 3 read, 3 written, 3 coefficients; stress 3 + 6 read, 6 written, 5 coefficients
    // acoustic: velocity 1 + 3 read, 3 written, 3 coefficients; pressure 3 + 1 read, 1 written, M (+ ABS)
    // EM: H 3 + 3 read, 3 written, 3 coefficients; E 3 + 3 read, 3 written, 6 coefficients (+ ABS);
    // visco-EM: the memory variables (3 L read and written per shot-cell) and 3 L Cd per cell
    const bool ac = physics == ACOUSTIC, em = physics == TEMEM;
    double fr = ac ? (kind == 0 ? 7 : 5) : (kind == 0 ? 12 : 15), co = ac ? (kind == 0 ? 3 : 1) : (kind == 0 ? 3 : 5);
    if (em) {
        fr = 9;
        co = kind == 0 ? 3 : 6 + (damping == 1 ? 1 : 0);
        if (visco && kind == 1) {
            fr += 6.0 * vL;
            co += 3.0 * vL;
        }
    }
    double bytes = 4.0 * (cells * B * fr + cells * co);
    if (damping == 2) {
        const double fx = (double)NY * NZ * 2 * W, fy = (double)NX * NZ * (free_surface ? W : 2 * W),
                     fz = (double)NX * NY * 2 * W;
        bytes += 8.0 * B * (ac ? 1 : (em ? 2 : 3)) * (fx + fy + fz);
    }
    return bytes;
}

const char *Elastic3D::kernel_name(int kind) const
{
    static const char *n[3] = {"el3d_velocity", "el3d_stress", "flush3d"};
    static const char *a[3] = {"ac3d_velocity", "ac3d_pressure", "flush3d"};
    static const char *e[3] = {"em3d_h", "em3d_e", "flush3d"};
    return physics == ACOUSTIC ? a[kind] : (physics == TEMEM ? e[kind] : n[kind]);
}

}  // namespace wv
