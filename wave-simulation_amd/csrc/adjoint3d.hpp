// Adjoint (time-reversed) 3D elastic step: device structures and launcher
// (adjoint3d.hip).  The reference has no adjoint; this is the exact discrete
// adjoint of the forward step of elastic3d.hip.
#pragma once

#include <algorithm>

#include "elastic3d.hpp"

namespace wv {

// adjoint state of the batch, same layout as Fields3D / Cpml3D::ps:
// l: lambda of vx vy vz sxx syy szz syz sxz sxy; g: scratch for the nine
// CPML-transposed terms of one phase; m: adjoint CPML memory variables
struct Adj3D {
    float *l[9];
    float *g[9];
    float *m[18];
};

// gradient planes (FP64, layout of one shot's field, summed over the shots) w.r.t.
// 1/rho_x 1/rho_y 1/rho_z M mu mu_xy mu_xz mu_yz
struct Grad3D {
    double *g[8];
};

// transposed derivative tables per axis: (D^T w)[j] = sum_k t[j][k] w[j + k - H]
struct TapsT3D {
    const float *xf, *xb, *yf, *yb, *zf, *zb;
};

// One adjoint time step t: the adjoint sources of t (time-major [NT][rec_total]) at the
// receivers, then the transposes of the stress and velocity updates with the record R of
// forward step t, accumulating the gradient in G.
void launch_adj3d_step(int order, const Layout3D &L, int nshots, int nyc, const Adj3D &A, const Coef3D &C,
                       const Cpml3D &P, const Rec3D &R, const Grad3D &G, const TapsT3D &T, const Acq &Q,
                       const float *adj, int t, hipStream_t st);

}  // namespace wv
