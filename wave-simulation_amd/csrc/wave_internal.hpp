// Internal declarations shared by the host runtime (abi.cpp, prep.cpp) and the
// HIP kernels.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>

#include "../../include/wave_abi.h"

namespace wv {

// The device buffers that hold a handle's state (wavefields of every shot, seismograms):
// what wv_check_finite and wv_check_finite_noalloc scan.  Fixed capacity, so listing them
// allocates nothing; a handle never has more (2D: 5 fields + seismograms; 3D: 9 + 1;
// reference-sequence path: 5 fields + 5 seismogram components).
struct StateBuffers {
    static constexpr int CAP = WV_NOALLOC_MAX_BUFFERS;
    const float *p[CAP];
    long long n[CAP];
    int k = 0;
    bool overflow = false;
    void add(const float *ptr, long long len)
    {
        if (!ptr || len <= 0) return;
        if (k == CAP) {
            overflow = true;
            return;
        }
        p[k] = ptr;
        n[k] = len;
        k++;
    }
};

enum Physics { ACOUSTIC = 0, ELASTIC = 1, TMEM = 2, SH = 3, TEMEM = 4 };

// Maximum half-width of a derivative stencil window (order 12 -> 6).
constexpr int MAX_H = 6;
// Tile of the 2D kernels: 64 columns (one wavefront) x tile_y rows, 256
// threads per block (4 wavefronts stacked in y, each thread tile_y/4 rows);
// TILE_Y is the default tile height of the two-launch kernels.
constexpr int TILE_X = 64;
constexpr int TILE_Y = 16;
constexpr int BLOCK_Y = 4;
// Left/right padding of a row in floats (keeps x = 0 on a 128-byte line).
constexpr int XPAD = 32;

// Padded 2D layout: element (x, y) of a field lives at
// (y + ypad) * pitch + (x + XPAD); the padding is zero and never written,
// so stencil taps that leave the grid read zeros.
struct Layout2D {
    int NX, NY;
    int H;            // taps reach [-H, H]
    int ypad;         // rows of zero padding above the grid (= H)
    int pitch;        // floats per row
    int rows;         // rows per field plane
    long long plane;  // floats per plane (one field of one shot)
    __host__ __device__ long long at(int x, int y) const { return (long long)(y + ypad) * pitch + (x + XPAD); }
};

// Tap tables: for every position along an axis, 2H+1 coefficients for the
// offsets -H..H (already multiplied by DT/DH).  Rows beyond the grid are zero.
struct Taps {
    const float *xf, *xb;      // [NXpad][2H+1]
    const float *yf, *yb;      // [NYpad][2H+1] plain operators
    const float *yf_v, *yb_v;  // y operators used in the velocity-type update
                               // (free-surface image versions when FreeSurface=1)
};

// Compact CPML frame index of position p on an axis of n positions (the rule of
// the host tables, cpml_axis): near side 0..W-1, far side W..2W-1, -1 outside
// the frame or on a skipped near side (the top under a free surface).  Kernels
// compute it rather than load it, so no load address waits for it.
__host__ __device__ inline int frame_index(int p, int n, int W, bool skip_near)
{
    if (p < 0 || p >= n) return -1;
    if (p < W) return skip_near ? -1 : p;
    return n - 1 - p < W ? 2 * W - n + p : -1;
}

// CPML coefficients along each axis (zero outside the frame); the compact frame
// index of a position is frame_index(), computed in the kernels.  Memory
// variables of the x frame are stored [shot][NY][2W], of the y frame
// [shot][2W][NX].
struct Cpml2D {
    int on;
    int W;
    const float *ax, *bx, *axh, *bxh;  // [NXpad]
    const float *ay, *by, *ayh, *byh;  // [NYpad]
    float *psx[4];                     // x-frame memory variables
    float *psy[4];                     // y-frame memory variables
};

struct Fields2D {
    float *f[5];  // acoustic: vx vy p ; elastic: vx vy sxx syy sxy ; tmem: hx hy ez ; sh: vz sxz syz
};

// write buffers of the CPML memory variables of the velocity derivatives
// (single-pass kernel: read from Cpml2D::psx/psy[0..1], write here)
struct PsiNew2D {
    float *psx[2];
    float *psy[2];
};

struct Coef2D {
    // acoustic: c0 = M, c1 = 1/rho_x, c2 = 1/rho_y
    // elastic : c0 = M, c1 = mu, c2 = 1/rho_x, c3 = 1/rho_y, c4 = mu_xy
    // tmem    : c0 = Ca, c1 = Cb, c2 = 1/mu_xz, c3 = 1/mu_yz
    // sh      : c0 = 1/rho, c1 = mu_xz, c2 = mu_yz
    // emem    : c0 = 1/mu_xy, c1 = Ca_x, c2 = Cb_x, c3 = Ca_y, c4 = Cb_y
    const float *c[5];
    const float *damp;  // ABS damping factors (padded layout) or nullptr
    const float *fsH;   // [NXpad] free-surface horizontal scale at y = 0 (elastic) or nullptr
    const float *fsV;   // [NXpad] free-surface vertical scale at y = 0 (elastic)
    int free_surface;
};

// Sources and receivers of all shots.  Sources of the stress-type fields (P,
// EZ) are injected and stress-type receivers gathered by the block that owns
// the cell, at the end of the stress-type kernel of step t.  Velocity-type
// sources (VX, VY) of step t are injected, and velocity-type receivers of
// step t gathered, by the owning block at the start of the velocity-type
// kernel of step t+1 (before anything reads the cell), or by the flush kernel
// after the last step of a run segment.
struct Acq {
    int nshots;
    int NT;
    int nbx, nby;               // blocks per shot of the tiled kernels
    const int *src_ptr;         // [nshots+1] sources of each shot
    const long long *src_cell;  // field offset (layout) of each source
    const int *src_type;        // 1..4
    const float *src_sig;       // [nsrc_total][NT]
    const int *rec_ptr;         // [nshots+1]
    const long long *rec_cell;
    const int *rec_type;
    int rec_total;
    float *seis;                // time-major [NT][rec_total]
    const int *bsrc_ptr;        // [nshots*nbx*nby+1] sources per (shot, block), in source order
    const int *bsrc;            // global source ids
    const int *brec_ptr;        // [nshots*nbx*nby+1]
    const int *brec;            // global receiver ids
    const int *vsrc_ptr;        // single-pass kernel: velocity-type sources per (shot, block)
    const int *vsrc;            //   covering the block's velocity region (tile + H halo)
    const int *tbase;           // device step counter base
};

// 2D viscoelastic and visco-SH stress updates (visco2d.hip): L relaxation mechanisms (1..4),
// their memory variables (shot planes in the field layout: P-SV Rxx[l] = R[l], Ryy[l] = R[4 + l],
// Rxy[l] = R[8 + l]; SH Rxz[l] = R[l], Ryz[l] = R[4 + l]), per cell tauP, tauS, tauS averaged
// like mu_xy (P-SV) or like mu_xz / mu_yz (SH), 1 + L tauP, 1 + L tauS (padded layout, one
// plane), per mechanism 1/tau_l, c1 = 1 - DT/(2 tau_l), c2 = 1/(1 + DT/(2 tau_l)), DT/2, and
// the relaxation updates of the P-SV free surface at y = 0 ([NXpad], or null).
struct Visco2D {
    int L;
    float *R[12];
    const float *tauP, *tauS, *tauSxy, *opP, *opS, *tauSxz, *tauSyz;
    const float *fsRH[4], *fsRV[4];
    float itau[4], c1[4], c2[4], dth;
    // visco-EM (ForwardSolver2Dviscotmem.cpp:172-199, ForwardSolver2Dviscoemem.cpp:175-218): memory
    // variables of the electric field in R[l] (TM: ez; TE: ex) and R[4 + l] (TE: ey), their per-cell
    // coefficients Cd_l in cd[l] / cd[4 + l] (Cd includes -DT), Cc_l, and DT
    const float *cd[8];
    float cc[4], dt;
};

// Forward quantities a 2D elastic step hands to the adjoint's gradient (all
// pointers null when not recording), per shot cell, same layout as the fields:
// the CPML-modified derivative sums of the step, h = g + (b psi + a g):
//   h[0] = h0 + h1 (vx update / (1/rho_x)),  h[1] = h2 + h3 (vy update / (1/rho_y)),
//   h[2] = h4 = CPML(Dxb vx),  h[3] = h5 = CPML(Dyb vy),  h[4] = h6 + h7 (sxy update / mu_xy).
// The adjoint step needs nothing else from the forward run.
struct Snap2D {
    float *h[5];
};

// The same five sums for every step of a run, kept for the adjoint pass
// instead of recomputing them (row-strip kernels only), in FP16 with one FP32
// scale per step, field, shot and strip of 64 columns x srows rows (the cells
// one wavefront of the row-strip kernel updates).  With
// fs = (t * 5 + f) * nshots + shot and q = (fs * nty + y / srows) * nbx + x / 64:
//   scale[q] = max|h| / 32768 over the strip (1 when all are below 1e-30)
//   h[(q * (srows / 2) + (y % srows) / 2) * 64 + x % 64] = FP16 pair (row y even:
//   low half, y + 1: high half) of h / scale,
// so a wavefront writes and reads aligned runs of 256 bytes.  h == nullptr:
// not recording.
struct Rec2D {
    unsigned *h;
    float *scale;
    int nbx, nty, srows;
};

// Launchers (kernels2d.hip).  All launches go to `stream`.
// which = 0: velocity-type phase (pend = apply the velocity-type sources /
// receivers of the previous step first); which = 1: stress-type phase.
// tile_y in {8, 16, 32}.  e0/e1 (optional) receive the kernel's own start/end timestamps.
void launch_step2d(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                   const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend,
                   hipStream_t stream, int which, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                   const Snap2D *snap = nullptr);
// velocity-type sources / receivers of step tbase + tofs (end of a run segment)
// Persistent TM-EM loop (kernels2d.hip): nsteps time steps of every shot in one cooperative launch
// of as many blocks as can be resident (grid barriers between the H and E updates); no ABS.
// grid_out (optional) receives the block count.  Returns the launch status.
hipError_t launch_tm2d_persistent(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                                  const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int nsteps,
                                  hipStream_t stream, int *grid_out = nullptr);
// Row-strip TM-EM step (strip2d.hip, blocks as the elastic row-strip kernels below): which = 0
// H update, 1 E update (+ ABS, EZ sources / receivers)
void launch_strip2d_tm(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                       const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t stream, int which,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Row-strip 2D elastic kernels (strip2d.hip): blocks of 64 columns x tile_y rows,
// tile_y in {8, 16, 32} (each of the 4 wavefronts owns tile_y/4 rows); same
// phases, acquisition lists and recorded quantities (snap) as launch_step2d.
// Needs planes below 2 GiB (strip2d_fits).
bool strip2d_fits(const Layout2D &L);
void launch_strip2d(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                    const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t stream, int which,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const Snap2D *snap = nullptr,
                    const Rec2D *rec = nullptr);
void launch_flush2d(int physics, const Layout2D &L, const Fields2D &F, const Acq &A, int tofs, hipStream_t stream);
void launch_abs_velocity2d(int physics, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                           hipStream_t stream);
// Single-pass 2D elastic step (fused2d.hip): reads step t-1 from F, writes step t to Fn.
// Output tiles are (64 - order) x |tile_y| cells, |tile_y| in {8, 16, 24}; tile_y < 0 loads the
// tile with direct global->LDS transfers, tile_y > 0 through registers.
int fused2d_tile_x(int order);
void launch_fused2d(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Fields2D &Fn,
                    const Coef2D &C, const Taps &T, const Cpml2D &P, const PsiNew2D &PN, const Acq &A, int tofs,
                    int pend, hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Visco-EM electric phase (kernels2d.hip: tm2d_e / te2d_e with memory variables); the magnetic
// phase is launch_step2d(TMEM or TEMEM, which = 0).
void launch_viscoem2d_e(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                        const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V,
                        hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Viscoelastic stress phase (visco2d.hip); the velocity phase is launch_step2d(ELASTIC or SH,
// which = 0).  physics ELASTIC: P-SV, SH: visco-SH.  tile_y in {8, 16, 32}.
void launch_visco2d_stress(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                           const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V,
                           hipStream_t stream, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_advance(int *tbase, int k, hipStream_t stream);
void launch_set(int *p, int v, hipStream_t stream);

}  // namespace wv
