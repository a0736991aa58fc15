// Definition of the solver handle behind the C ABI (shared by abi.cpp and
// adjoint.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/wave_abi.h"
#include "elastic3d.hpp"
#include "plain.hpp"
#include "wave_internal.hpp"

// status + thread-local message returned through wv_last_error (abi.cpp)
int wv_fail(int code, const std::string &msg);
inline int fail(int code, const std::string &msg) { return wv_fail(code, msg); }

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(WV_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Position-indexed tables (taps, CPML coefficients, frame indices) are stored
// for positions [-ORG, n_padded + ORG) so kernels may index halo positions
// outside the grid; the pointer handed to a kernel points at position 0.
constexpr int ORG = 64;

struct ShotAcq {
    std::vector<long long> sidx;
    std::vector<int> stype;
    std::vector<float> ssig;
    std::vector<long long> ridx;
    std::vector<int> rtype;
};

enum Variant { V_SINGLE = 0, V_PHASED = 1, V_STRIP = 2, V_PERSIST = 3 };

struct AdjointState;

struct wv_solver {
    int dim = 2;
    int physics = wv::ELASTIC;
    wv_config cfg{};
    bool inited = false;
    bool model_ok = false;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;

    int variant = V_PHASED;
    int tile_x = wv::TILE_X, tile_y = wv::TILE_Y;  // tiling of the kernels in use
    int tile_load = 1;  // single-pass kernel: -1 direct global->LDS loads, 1 register-staged loads
    wv::Layout2D L{};
    int H = 0;
    int nfields = 0;
    int cur = 0;  // buffer holding the current state (single-pass kernel ping-pongs)
    std::vector<void *> allocs;

    // device buffers
    float *field[2][5] = {};
    float *coef[5] = {};
    float *damp = nullptr, *fsH = nullptr, *fsV = nullptr;
    float *tap[6] = {};  // position-0 pointers: xf xb yf yb yf_v yb_v
    // viscoelastic (2D): relaxation mechanisms (wv_set_relaxation) and their device data
    bool visco = false;
    int vL = 0;
    std::vector<double> vfreq;
    std::vector<float> vtauP, vtauS;  // grid order
    float *vR[12] = {};               // memory variables Rxx[4], Ryy[4], Rxy[4] (shot planes)
    float *vcoef[7] = {};             // tauP, tauS, tauS_xy, 1 + L tauP, 1 + L tauS, tauS_xz, tauS_yz
    float *vfs[8] = {};               // free-surface relaxation updates: horizontal[4], vertical[4]
    float vitau[4] = {}, vc1[4] = {}, vc2[4] = {}, vdth = 0.f;
    // visco-EM: Cd_l per cell (TM: node; TE: x-averaged at l, y-averaged at 4 + l), Cc_l, DT
    float *vcd[8] = {};
    float vcc[4] = {}, vdt = 0.f;
    float *cpx[4] = {}, *cpy[4] = {};
    float *psv[2][4] = {};  // ping-pong: velocity-derivative memory variables psx0 psx1 psy0 psy1
    float *pss[4] = {};     // stress-derivative memory variables psx2 psx3 psy2 psy3
    size_t psx_n = 0, psy_n = 0;

    // acquisition
    std::vector<ShotAcq> shots;
    bool acq_dirty = true;
    std::vector<void *> acq_allocs;
    int *d_src_ptr = nullptr, *d_src_type = nullptr, *d_rec_ptr = nullptr, *d_rec_type = nullptr;
    int *d_bsrc_ptr = nullptr, *d_bsrc = nullptr, *d_brec_ptr = nullptr, *d_brec = nullptr;
    int *d_vsrc_ptr = nullptr, *d_vsrc = nullptr;
    int nbx = 0, nby = 0;
    long long *d_src_cell = nullptr, *d_rec_cell = nullptr;
    float *d_src_sig = nullptr, *d_seis = nullptr;
    int rec_total = 0;
    std::vector<int> rec_ptr_h;
    int *d_tbase = nullptr;
    unsigned *d_flag = nullptr;  // wv_check_finite's result word

    // graph of steps_per_graph time steps (starts and ends in buffer 0)
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;

    // reference-sequence path
    wv::PlainSolver *plain = nullptr;
    // 3D elastic driver
    wv::Elastic3D *e3 = nullptr;
    // adjoint (checkpoints, adjoint state, gradients) and the host copy of
    // the model the gradient chain rule needs
    AdjointState *adj = nullptr;
    std::vector<float> model_h[3];

    // per-kernel timing (wv_set_kernel_timing)
    bool ktiming = false;
    std::vector<hipEvent_t> kev;
    // kinds: 0 velocity phase, 1 stress phase, 2 flush; adjoint pass: 3 recomputed
    // velocity phase, 4 recomputed stress phase, 5 adjoint stress step, 6 adjoint velocity step
    double kms[7] = {0, 0, 0, 0, 0, 0, 0};
    long long klaunch[7] = {0, 0, 0, 0, 0, 0, 0};

    ~wv_solver();
    template <typename T>
    int dalloc(T **p, size_t n, std::vector<void *> &list)
    {
        void *q = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return fail(WV_ENOMEM, "hipMalloc failed");
        if (hipMemsetAsync(q, 0, n * sizeof(T), st) != hipSuccess) return fail(WV_EDEVICE, "hipMemset failed");
        list.push_back(q);
        *p = (T *)q;
        return 0;
    }
    template <typename T>
    int upload(T **p, const std::vector<T> &h, std::vector<void *> &list)
    {
        int rc = dalloc(p, h.size(), list);
        if (rc) return rc;
        // pageable source: the copy is complete before the caller may free it
        if (!h.empty() && (hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess ||
                           hipStreamSynchronize(st) != hipSuccess))
            return fail(WV_EDEVICE, "hipMemcpy failed");
        return 0;
    }
    // upload a position-indexed table with ORG positions of `fill` on both sides
    template <typename T>
    int upload_pos(T **p0, const std::vector<T> &h, int per_pos, T fill)
    {
        std::vector<T> ext((size_t)(h.size() / per_pos + 2 * ORG) * per_pos, fill);
        std::copy(h.begin(), h.end(), ext.begin() + (size_t)ORG * per_pos);
        T *base = nullptr;
        int rc = upload(&base, ext, allocs);
        if (rc) return rc;
        *p0 = base + (size_t)ORG * per_pos;
        return 0;
    }
    void free_list(std::vector<void *> &list)
    {
        for (void *p : list) (void)hipFree(p);
        list.clear();
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gsteps = 0;
    }
    wv::Fields2D fields2d(int b) const
    {
        wv::Fields2D F{};
        for (int k = 0; k < 5; k++) F.f[k] = field[b][k];
        return F;
    }
    wv::Coef2D coef2d() const
    {
        wv::Coef2D C{};
        for (int k = 0; k < 5; k++) C.c[k] = coef[k];
        C.damp = cfg.damping_boundary == 1 ? damp : nullptr;
        C.fsH = fsH;
        C.fsV = fsV;
        C.free_surface = cfg.free_surface == 1;
        return C;
    }
    wv::Taps taps() const
    {
        wv::Taps T{};
        T.xf = tap[0];
        T.xb = tap[1];
        T.yf = tap[2];
        T.yb = tap[3];
        T.yf_v = tap[4];
        T.yb_v = tap[5];
        return T;
    }
    wv::Visco2D visco2d() const
    {
        wv::Visco2D V{};
        V.L = vL;
        for (int k = 0; k < 12; k++) V.R[k] = vR[k];
        V.tauP = vcoef[0];
        V.tauS = vcoef[1];
        V.tauSxy = vcoef[2];
        V.opP = vcoef[3];
        V.opS = vcoef[4];
        V.tauSxz = vcoef[5];
        V.tauSyz = vcoef[6];
        for (int l = 0; l < 4; l++) {
            V.fsRH[l] = vfs[l];
            V.fsRV[l] = vfs[4 + l];
            V.itau[l] = vitau[l];
            V.c1[l] = vc1[l];
            V.c2[l] = vc2[l];
        }
        V.dth = vdth;
        for (int k = 0; k < 8; k++) V.cd[k] = vcd[k];
        for (int l = 0; l < 4; l++) V.cc[l] = vcc[l];
        V.dt = vdt;
        return V;
    }
    wv::Cpml2D cpml2d(int b) const
    {
        wv::Cpml2D P{};
        P.on = cfg.damping_boundary == 2;
        P.W = cfg.boundary_width;
        P.ax = cpx[0];
        P.bx = cpx[1];
        P.axh = cpx[2];
        P.bxh = cpx[3];
        P.ay = cpy[0];
        P.by = cpy[1];
        P.ayh = cpy[2];
        P.byh = cpy[3];
        P.psx[0] = psv[b][0];
        P.psx[1] = psv[b][1];
        P.psy[0] = psv[b][2];
        P.psy[1] = psv[b][3];
        P.psx[2] = pss[0];
        P.psx[3] = pss[1];
        P.psy[2] = pss[2];
        P.psy[3] = pss[3];
        return P;
    }
    wv::PsiNew2D psinew(int b) const
    {
        wv::PsiNew2D N{};
        N.psx[0] = psv[b][0];
        N.psx[1] = psv[b][1];
        N.psy[0] = psv[b][2];
        N.psy[1] = psv[b][3];
        return N;
    }
    wv::Acq acq() const
    {
        wv::Acq A{};
        A.nshots = cfg.nshots;
        A.NT = cfg.NT;
        A.src_ptr = d_src_ptr;
        A.src_cell = d_src_cell;
        A.src_type = d_src_type;
        A.src_sig = d_src_sig;
        A.rec_ptr = d_rec_ptr;
        A.rec_cell = d_rec_cell;
        A.rec_type = d_rec_type;
        A.rec_total = rec_total;
        A.seis = d_seis;
        A.tbase = d_tbase;
        A.nbx = nbx;
        A.nby = nby;
        A.bsrc_ptr = d_bsrc_ptr;
        A.bsrc = d_bsrc;
        A.brec_ptr = d_brec_ptr;
        A.brec = d_brec;
        A.vsrc_ptr = d_vsrc_ptr;
        A.vsrc = d_vsrc;
        return A;
    }
    int enqueue_steps(int nsteps, bool timed);
    int sync_acquisition();
    double kernel_bytes(int kind) const;
};

// adjoint.cpp
void adjoint_destroy(AdjointState *a);
// checkpoint interval (0 = checkpointing off)
int adjoint_every(const wv_solver *s);
// checkpoint precision of the active setup (1 = FP16)
int adjoint_fp16(const wv_solver *s);
// forget the checkpoint / record layout so the next wv_set_checkpointing lays it out again
void adjoint_invalidate_layout(wv_solver *s);
// the per-step record the row-strip kernels write during a checkpointed run, or nullptr
const wv::Rec2D *adjoint_record(const wv_solver *s);
// called by wv_run: at its start, before step t (t a multiple of the interval), at its end
void adjoint_run_begin(wv_solver *s, int t_start);
int adjoint_checkpoint(wv_solver *s, int t);
void adjoint_run_done(wv_solver *s, int t_end);
