// Host-side preparation of everything the kernels consume: derivative tap
// tables, CPML profiles, model coefficients.  Runs once per model / config.
#pragma once

#include <string>
#include <vector>

namespace wv {

// Exact staggered-grid Taylor coefficients for offsets (j - N/2 + 1/2), j < N,
// ordered as in the reference (Derivatives.cpp:2001-2042).
double fd_coefficient(int order, int j);

// Taps of one derivative operator along an axis of length n:
// out[(c)*(2H+1) + (k)] is the weight of point c + k - H in the derivative
// at position c, already multiplied by DT/DH.  padded_n >= n rows are written
// (rows >= n are zero).
//   forward=true : Dxf-type operator (result lives half a cell forward)
//   stencil=false: sparse operator with edge order reduction (Derivatives.cpp:129-1234)
//   stencil=true : StencilMatrix operator, points outside the axis dropped (FDTD2D.cpp:156-185)
void build_taps(std::vector<float> &out, int n, int padded_n, int order, int H, bool forward, bool stencil,
                float DH, float DT);

// Transposed tap table of a build_taps table: (D^T w)[j] = sum_k Tt[j][k] w[j + k - H] with
// Tt[j][k] = T[j + k - H][2H - k] (rows outside the table contribute nothing).
std::vector<float> transpose_taps(const std::vector<float> &t, int H);

// Free-surface image operators along depth (Derivatives.cpp:367-526, non-hybrid):
// forward=true -> DyfFreeSurface, false -> DybFreeSurface.
void build_taps_free_surface(std::vector<float> &out, int n, int padded_n, int order, int H, bool forward, float DH,
                             float DT);

// CPML damping profile a, b of width W (CPML.cpp:39-68), computed in FP64.
void cpml_profile(int W, double npower, double fc, double vmax, double DT, double DH, bool half,
                  std::vector<double> &a, std::vector<double> &b);

// Per-position CPML coefficients along an axis (CPML2D.cpp:228-270): positions
// within W of either end get the profile value at their distance; on the far
// side full and half profiles swap; the near side of the depth axis is left
// out when a free surface is on.  idx gets the compact frame index or -1.
void cpml_axis(int n, int padded_n, int W, bool skip_near_side, double npower, double fc, double vmax, double DT,
               double DH, std::vector<float> &a, std::vector<float> &b, std::vector<float> &ah, std::vector<float> &bh,
               std::vector<int> &idx);

// Ca = (1 - q) / (1 + q), Cb = 1 / (eps (1 + q)), q = sigma DT / (2 eps), in the float
// arithmetic and operation order of ForwardSolverEM.cpp:14-68 (the time loop multiplies ez
// by Ca every step, so a Ca rounded differently drifts from the reference float build).
void em_ca_cb(float eps, float sigma, float DT, float &ca, float &cb);
// Visco-EM coefficients of one cell (ForwardSolverEM.cpp:70-157) from the static permittivity
// eps_s [F/m], conductivity sig_s, tau of the permittivity tE and of the conductivity tS [s]:
// the effective optical values eps_opt = max((1 - tE) eps_s + sig_s tS, eps0) and
// sig_opt = tE inv_tau_mean eps_s + sig_s give Ca, Cb; cd[l] = tE cdl[l] eps_s (-DT) with
// cdl[l] = (1 / (1 + DT / (2 tau_l))) / (L tau_l^2).  Float steps of the reference's float build.
void em_visco_coef(float eps_s, float sig_s, float tE, float tS, float DT, float inv_tau_mean, int L,
                   const float *cdl, float &ca, float &cb, float *cd);

// ABS damping factor of one grid cell (ABS2D.cpp:110-178, ABS3D.cpp:150-215).
std::vector<double> abs_profile(int W, double damping_coeff);

}  // namespace wv
