// Adjoint (time-reversed) 2D elastic step and FP16 checkpoints: device
// structures and launchers (adjoint2d.hip).  The reference has no adjoint;
// this is the exact discrete adjoint of the forward step of kernels2d.hip.
#pragma once

#include "wave_internal.hpp"

namespace wv {

// adjoint wavefields (same padded layout and shot stride as the forward fields)
struct AdjFields {
    float *lvx, *lvy, *lsxx, *lsyy, *lsxy;
};

// adjoint of the CPML memory variables, same compact layout as Cpml2D::psx/psy.
// A kernel reads mx/my (values before its step) on its whole tile + halo and
// writes the updated values of its own cells to mxn/myn (ping-pong, so no
// block reads a value another block has already advanced).
struct AdjPsi {
    const float *mx[4], *my[4];
    float *mxn[4], *myn[4];
};

// forward quantities of one step read by the adjoint kernels (recorded by the
// forward kernels, Snap2D): the CPML-modified derivative sums of the step
// h[0] = h0 + h1, h[1] = h2 + h3, h[2] = h4, h[3] = h5, h[4] = h6 + h7.
struct FwdSnap {
    const float *h[5];
    // or, when hs != nullptr, the FP16 record of the whole run (Rec2D layout) at step t
    const unsigned *hs;
    const float *ss;
    int t, nbx, nty, srows, nshots;
};

// transposed derivative tables: (D^T w)[j] = sum_k t[j][k] w[j + k - H]
struct TapsT {
    const float *xf, *xb, *yf, *yb;
};

// gradients w.r.t. 1/rho_x, 1/rho_y, M, mu, mu_xy (FP64, one plane shared by all shots)
struct Grad {
    double *g[5];
};

// e0/e1 (optional) receive the kernel's own start/end timestamps
void launch_adj_stress(int order, int tile_y, const Layout2D &L, int nshots, const AdjFields &A, const AdjPsi &M,
                       const FwdSnap &S, const Coef2D &C, const Taps &T, const TapsT &TT, const Cpml2D &P,
                       const Grad &G, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_adj_velocity(int order, int tile_y, const Layout2D &L, int nshots, const AdjFields &A, const AdjPsi &M,
                         const FwdSnap &S, const Coef2D &C, const Taps &T, const TapsT &TT, const Cpml2D &P,
                         const Grad &G, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// adjoint sources of step t (time-major [NT][rec_total]) into the adjoint fields
void launch_adj_inject(const Layout2D &L, const AdjFields &A, const Acq &Q, const float *adj, int t, hipStream_t st);

// checkpoints: per (field, shot) scale = max|x| / 32768, values stored as FP16
void launch_ckpt_pack(const float *f, unsigned short *h, unsigned *maxbits, long long per_shot, int nshots,
                      hipStream_t st);
void launch_ckpt_unpack(const unsigned short *h, float *f, const unsigned *maxbits, long long per_shot, int nshots,
                        hipStream_t st);

}  // namespace wv
