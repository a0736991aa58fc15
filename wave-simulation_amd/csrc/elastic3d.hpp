// 3D elastic staggered-grid solver (ForwardSolver3Delastic.cpp:120-416):
// device layout, kernel launchers and the host-side driver used by the C ABI.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../../include/wave_abi.h"
#include "wave_internal.hpp"

namespace wv {

// Padded 3D layout of one field of one shot.  Index order as the reference
// (x fastest, then z, then depth y slowest: Coordinates.cpp:687):
// element (x, y, z) lives at (y + H) * slab + (z + H) * pitch + x + XPAD.
// Padding is zero and never written; a tile of 64 columns plus an H-wide halo
// and up to 16 + H rows in z and H planes in y around the grid stay inside.
struct Layout3D {
    int NX, NY, NZ;
    int H;
    int pitch;       // floats per x row
    int zrows;       // rows per y plane (H + roundup(NZ, 16) + H)
    long long slab;  // floats per y plane
    long long vol;   // floats per field of one shot
    __host__ __device__ long long at(int x, int y, int z) const
    {
        return (long long)(y + H) * slab + (long long)(z + H) * pitch + (x + XPAD);
    }
};

struct Fields3D {
    float *f[9];  // vx vy vz sxx syy szz syz sxz sxy (EM: hx hy hz ex ey ez)
};

struct Coef3D {
    // 1/rho_x 1/rho_y 1/rho_z M mu mu_xy mu_xz mu_yz (shared by the shots of a batch)
    const float *irx, *iry, *irz, *M, *mu, *muxy, *muxz, *muyz;
    const float *damp;      // ABS factors (layout of one field) or nullptr
    const float *fsH, *fsV;  // free-surface scales on the y = 0 plane ((z + H) * pitch + x + XPAD) or nullptr
    int fs;                  // FreeSurface = 1 (acoustic: p = 0 on the y = 0 plane)
    // EM: irx, iry, irz hold 1/mu averaged in the yz, xz, xy planes (for hx, hy, hz) and em[] the
    // Ca, Cb of the x-, y- and z-averaged permittivity and conductivity (for ex, ey, ez)
    const float *em[6];
};

struct Taps3D {
    const float *xf, *xb, *yf, *yb, *yf_v, *yb_v, *zf, *zb;  // [pos][2H+1], see Taps
    // per y position, everything a plane of a kernel needs along y, [pos][2K + 5] (K = 2H+1):
    // K taps of the backward operator, K of the forward one, the CPML y frame index (int
    // bits, -1 outside), a_y, b_y, a_y(half), b_y(half) (0 outside the frame); ys: stress
    // update (Dyb, Dyf), yv: velocity update (its operators, image ones under a free surface)
    const float *ys, *yv;
};

// CPML coefficients per axis and compact frame indices (as Cpml2D).  Memory
// variables: x frame [shot][NY][NZ][2W], y frame [shot][2W][NZ][NX],
// z frame [shot][NY][2W][NX].  Index k of ps[] follows the derivative terms
// of the reference's run(): 0 sxx_x 1 sxy_y 2 sxz_z 3 sxy_x 4 syy_y 5 syz_z
// 6 sxz_x 7 syz_y 8 szz_z | 9 vx_x 10 vy_y 11 vz_z 12 vx_y 13 vy_x 14 vx_z
// 15 vz_x 16 vy_z 17 vz_y.
struct Cpml3D {
    int on;
    int W;
    const float *ax, *bx, *axh, *bxh;
    const float *ay, *by, *ayh, *byh;
    const float *az, *bz, *azh, *bzh;
    const int *xi, *yi, *zi;
    float *ps[18];
};

// What the forward kernels store, when asked, for the adjoint pass: per time step, the
// CPML-modified derivative sums the gradient correlates with, in the layout of the fields
// (all shots): r[0..2] the velocity-update sums of vx, vy, vz (before 1/rho), r[3..5] the
// normal strains Dxb vx, Dyb vy, Dzb vz, r[6..8] the shear sums of sxy, sxz, syz (before
// mu_xy, mu_xz, mu_yz).
struct Rec3D {
    float *r[9];
};

// Viscoelastic stress update (ForwardSolver3Dviscoelastic.cpp:270-444): per cell tauP, tauS, the
// 4-point averages of tauS with the matrices of mu_xy, mu_xz, mu_yz, 1 + L tauP, 1 + L tauS
// (layout of one field, shared by the shots), the memory variables R[6 * 4] (group g of
// xx yy zz yz xz xy, mechanism l at 4 g + l; layout of the fields), per mechanism the
// constants 1 / tau_l, c1 = 1 - DT / (2 tau_l), c2 = 1 / (1 + DT / (2 tau_l)), and the
// free-surface relaxation updates on the y = 0 plane (layout of Coef3D::fsH).
struct Visco3D {
    int L;
    const float *tauP, *tauS, *tSxy, *tSxz, *tSyz, *opP, *opS;
    float *R[24];
    const float *fsRH[4], *fsRV[4];
    float itau[4], c1[4], c2[4];
    float dth;
    // visco-EM: the memory variables of ex, ey, ez are R[4 g + l] (g = 0, 1, 2), their Cd_l per
    // cell (x-, y-, z-averaged, DT included) cd[4 g + l], Cc_l and DT
    const float *cd[12];
    float cc[4], dt;
};

// One time step = velocity launch + stress launch (+ ABS velocity damping).
// rz = rows in z per thread (tile height 4 * rz), nyc = chunks along y; R (optional)
// receives this step's record; V (optional) makes the stress launch the viscoelastic one.
void launch_el3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const Rec3D *R = nullptr,
                 const Visco3D *V = nullptr);
// 3D acoustic step (fields vx vy vz p, memory variables 0..5 of Cpml3D::ps): which = 0 velocity, 1 pressure
void launch_ac3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_abs_velocity3d(const Layout3D &L, int nshots, const Fields3D &F, const Coef3D &C, hipStream_t stream);
void launch_flush3d(const Layout3D &L, const Fields3D &F, const Acq &A, int tofs, hipStream_t stream, bool em = false);
// 3D EM step (ForwardSolver3Demem.cpp:117-235; fields hx hy hz ex ey ez, CPML memory variables
// 0..11 of Cpml3D::ps as the EM terms): which = 0 H update, 1 E update (+ ABS of E); V (optional)
// adds the memory variables of the visco-EM solver (ForwardSolver3Dviscoemem.cpp:170-305)
void launch_em3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const Visco3D *V = nullptr);

// Host driver of the 3D elastic solver (owned by a wv_solver handle).
class Elastic3D {
  public:
    ~Elastic3D();
    int init(const wv_config &cfg, hipStream_t stream, int physics = ELASTIC, bool visco = false);
    // viscoelastic (3D): relaxation mechanisms before set_model (tauP, tauS in grid order)
    int set_relaxation(int nmech, const double *freqs, const float *tauP, const float *tauS);
    int set_model(const float *vp, const float *vs, const float *rho);  // EM: eps_r, sigma, mu_r
    int set_acquisition(int shot, const std::vector<long long> &sidx, const std::vector<int> &stype,
                        const std::vector<float> &ssig, const std::vector<long long> &ridx,
                        const std::vector<int> &rtype);
    int reset();
    int run(int t_start, int t_end, bool timed);
    int get_seismograms(int shot, float *out);  // [nrec][NT] in receiver input order
    int field_io(int shot, int field, float *out, const float *in);
    int tune(int nsteps, int *choice);
    int set_rz(int rz);
    int nrec(int shot) const;
    double kernel_bytes(int kind) const;
    const char *kernel_name(int kind) const;
    // device buffers of the state (wavefields of all shots, seismograms) for wv_check_finite
    void buffers(wv::StateBuffers &out) const;

    // adjoint-state gradient (CPML or no absorbing frame, no free surface; adjoint3d_host.cpp):
    // checkpoints every `every` steps of the following runs (FP16 or FP32 fields), kept in
    // device memory or, when they do not fit next to the recomputation buffers, in pinned
    // host memory; budget_gib < 0: 3/4 of the free device memory
    int set_checkpointing(int every, int fp16, double budget_gib);
    int set_adjoint_source(int shot, const float *adj);  // [nrec][NT]
    int run_adjoint();
    int get_gradient(int param, double *out);  // 0 vp 1 vs 2 rho 3..10 coefficients
    // memory plan of the last set_checkpointing: sub-segment length, sub-checkpoints per
    // segment, whether the checkpoints sit in host memory
    void adjoint_plan(int *k2, int *nsub, int *on_host) const;
    double kms[3] = {0, 0, 0};
    long long klaunch[3] = {0, 0, 0};
    std::string error;

  private:
    struct Adj3State;
    Adj3State *adj3 = nullptr;
    std::vector<float> model_h[3];  // vp, vs, rho (chain rule of the gradient)
    int run_segment(int t_start, int t_end, bool timed);
    int set_model_em(const float *m0, const float *m1, const float *m2);
    int psi_axis(int q) const;
    int checkpoint_interval() const;
    int enqueue_rec(int nsteps, float *rec);
    int save_checkpoint(int k);
    int restore_checkpoint(int k);
    int save_state(float *dst);
    int restore_state(const float *src);
    size_t psi_floats() const;
    void free_adjoint();
    int fail(int code, const std::string &m);
    int dalloc(float **p, size_t n);
    template <typename T>
    int upload(T **p, const std::vector<T> &h, std::vector<void *> &list);
    template <typename T>
    int upload_pos(T **p0, const std::vector<T> &h, int per_pos, T fill);
    int sync_acquisition();
    int enqueue(int nsteps, bool timed);
    void drop_graph();
    void free_list(std::vector<void *> &l);
    Fields3D fields() const;
    Coef3D coefs() const;
    Taps3D taps() const;
    Cpml3D cpml() const;
    Acq acq() const;
    Visco3D visco3d() const;

    hipStream_t st = nullptr;
    int NX = 0, NY = 0, NZ = 0, NT = 0, H = 0, order = 0, nshots = 1;
    int damping = 0, free_surface = 0, W = 0, steps_per_graph = 0;
    bool ops_stencil = false;
    int physics = ELASTIC;
    // viscoelastic: relaxation mechanisms and their device data
    bool visco = false;
    int vL = 0;
    std::vector<double> vfreq;
    std::vector<float> vtauP, vtauS;  // grid order
    float *vR[24] = {};               // memory variables (Visco3D::R)
    float *vcoef[7] = {};             // tauP, tauS, tauS_xy, tauS_xz, tauS_yz, 1 + L tauP, 1 + L tauS
    float *vfs[8] = {};               // free-surface relaxation updates: horizontal[4], vertical[4]
    float vitau[4] = {}, vc1[4] = {}, vc2[4] = {}, vdth = 0.f;
    // EM: Ca, Cb of the x-, y-, z-averaged values (Coef3D::em); visco-EM: Cd_l per axis, Cc_l, DT
    float *emc[6] = {};
    float *vcd[12] = {};
    float vcc[4] = {}, vdt = 0.f;
    int nf = 9;    // fields: elastic vx vy vz sxx syy szz syz sxz sxy, acoustic vx vy vz p, EM hx hy hz ex ey ez
    int nps = 18;  // CPML memory variables (acoustic: p_x p_y p_z vx_x vy_y vz_z; EM: 12)
    double DT = 0, DH = 0, fc_ref = 0;  // fc_ref: CenterFrequencyCPML, the viscoelastic reference frequency
    int rz = 2, nyc = 1, nbx = 0, nbz = 0;
    Layout3D L{};
    std::vector<void *> allocs, acq_allocs;
    float *F[9] = {};
    float *coef[8] = {};
    float *damp = nullptr, *fsH = nullptr, *fsV = nullptr;
    float *tap[8] = {};
    float *ytab[2] = {};  // Taps3D::ys, Taps3D::yv
    float *cp[3][4] = {};
    int *idx[3] = {};
    float *ps[18] = {};
    size_t ps_n[3] = {0, 0, 0};
    bool model_ok = false;

    struct Shot {
        std::vector<long long> sidx, ridx;
        std::vector<int> stype, rtype;
        std::vector<float> ssig;
    };
    std::vector<Shot> shots;
    bool acq_dirty = true;
    int *d_src_ptr = nullptr, *d_src_type = nullptr, *d_rec_ptr = nullptr, *d_rec_type = nullptr;
    int *d_bsrc_ptr = nullptr, *d_bsrc = nullptr, *d_brec_ptr = nullptr, *d_brec = nullptr;
    long long *d_src_cell = nullptr, *d_rec_cell = nullptr;
    float *d_src_sig = nullptr, *d_seis = nullptr;
    int rec_total = 0;
    std::vector<int> rec_ptr_h;
    int *d_tbase = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;
    std::vector<hipEvent_t> kev;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

}  // namespace wv
