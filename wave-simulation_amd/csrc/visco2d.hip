// 2D viscoelastic P-SV and visco-SH stress updates for gfx950 (ForwardSolver2Dviscoelastic.cpp:
// 190-300, ForwardSolver2Dviscosh.cpp:130-185).
//
// The velocity update of the viscoelastic solver is the elastic one
// (ForwardSolver2Dviscoelastic.cpp:140-190), so a viscoelastic step is the elastic velocity
// kernel followed by this stress kernel.  Same tile, LDS staging, CPML, free surface, ABS and
// acquisition as el2d_stress (kernels2d.hip); per cell it also carries the memory variables
// Rxx, Ryy, Rxy of each of the L relaxation mechanisms (generalised standard linear solid,
// Crank-Nicolson in the relaxation term), in the reference's order of operations:
//   u = M (vxx + vyy);  for l: g = u tauP / tau_l; sxx += dt/2 Rxx_l; Rxx_l = c1 Rxx_l - g; (same for yy)
//   sxx += (1 + L tauP) u, syy += (1 + L tauP) u
//   u = 2 mu vyy;  for l: Rxx_l = c2 (Rxx_l + u tauS / tau_l); sxx += dt/2 Rxx_l;  sxx -= (1 + L tauS) u
//   u = 2 mu vxx;  for l: Ryy_l = c2 (Ryy_l + u tauS / tau_l); syy += dt/2 Ryy_l;  syy -= (1 + L tauS) u
//   u = mu_xy (vxy + vyx);  for l: sxy += dt/2 Rxy_l; Rxy_l = c2 (c1 Rxy_l - u tauS_xy / tau_l);
//       sxy += dt/2 Rxy_l;  sxy += (1 + L tauS) u
// with c1 = 1 - dt / (2 tau_l), c2 = 1 / (1 + dt / (2 tau_l)), then the free-surface exchange
// of FreeSurface2Dviscoelastic.cpp:14-45 on the surface row, syy and Ryy zero there, and the
// ABS damping of the stresses.
#include "acq2d.hpp"

namespace wv {

namespace {

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void ve2d_stress(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                        int tofs, Visco2D V)
{
    using RT = RowTile<H, TY>;
    constexpr int R = TY / BLOCK_Y, K = 2 * H + 1;
    __shared__ float s_vx[RT::RH][RT::RW];
    __shared__ float s_vy[RT::RH][RT::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tx = threadIdx.x;
    const unsigned ux = threadIdx.x;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int NL = V.L;

    const long long corner = so + L.at(x0 - H, y0 - H);
    RT rvx, rvy;
    rvx.load(F.f[0] + corner, L.pitch, wy, tx);
    rvy.load(F.f[1] + corner, L.pitch, wy, tx);

    const Col col = column(L, P, x0 + tx);
    float cxf[K], cxb[K];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;

    float p2[R], p3[R], q2[R], q3[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        p2[r] = p3[r] = q2[r] = q3[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) {
            const long long pr = ((long long)shot * L.NY + y) * W2;
            p2[r] = (P.psx[2] + pr)[(unsigned)col.xi];
            p3[r] = (P.psx[3] + pr)[(unsigned)col.xi];
        }
        if (yi[r] >= 0 && col.in) {
            const long long pr = ((long long)shot * W2 + yi[r]) * L.NX + x0;
            q2[r] = (P.psy[2] + pr)[ux];
            q3[r] = (P.psy[3] + pr)[ux];
        }
    }
    rvx.store(s_vx, wy, tx);
    rvy.store(s_vy, wy, tx);
    __syncthreads();

    const bool fs = C.free_surface != 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyb = T.yb + (long long)y * K;
        const float *cyf = T.yf + (long long)y * K;
        const int Rr = ly + H;
        float vxx = dx<H>(s_vx, Rr, tx, cxb);
        float vyy = dy<H>(s_vy, Rr, tx, cyb);
        float u2 = dy<H>(s_vx, Rr, tx, cyf);
        float t2 = dx<H>(s_vy, Rr, tx, cxf);
        if (!col.in) continue;
        if (col.xi >= 0) {
            const long long pr = ((long long)shot * L.NY + y) * W2;
            (P.psx[2] + pr)[(unsigned)col.xi] = cpml(vxx, p2[r], col.ax, col.bx);
            (P.psx[3] + pr)[(unsigned)col.xi] = cpml(t2, p3[r], col.axh, col.bxh);
        }
        if (yi[r] >= 0) {
            const long long pr = ((long long)shot * W2 + yi[r]) * L.NX + x0;
            (P.psy[2] + pr)[ux] = cpml(vyy, q2[r], P.ay[y], P.by[y]);
            (P.psy[3] + pr)[ux] = cpml(u2, q3[r], P.ayh[y], P.byh[y]);
        }
        const long long crow = L.at(x0, y), row = so + crow;
        const float M = (C.c[0] + crow)[ux], mu = (C.c[1] + crow)[ux], mxy = (C.c[4] + crow)[ux];
        const float tP = (V.tauP + crow)[ux], tS = (V.tauS + crow)[ux], tSxy = (V.tauSxy + crow)[ux];
        const float oP = (V.opP + crow)[ux], oS = (V.opS + crow)[ux];
        float xx = (F.f[2] + row)[ux], yy = (F.f[3] + row)[ux], xy = (F.f[4] + row)[ux];
        float rxx[4], ryy[4], rxy[4];
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                rxx[l] = (V.R[l] + row)[ux];
                ryy[l] = (V.R[4 + l] + row)[ux];
                rxy[l] = (V.R[8 + l] + row)[ux];
            }
        const float dth = V.dth;
        float u = vxx + vyy;
        u *= M;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float g = V.itau[l] * u;
                g *= tP;
                xx += dth * rxx[l];
                rxx[l] *= V.c1[l];
                rxx[l] -= g;
                yy += dth * ryy[l];
                ryy[l] *= V.c1[l];
                ryy[l] -= g;
            }
        u *= oP;
        xx += u;
        yy += u;
        u = vyy * mu;
        u *= 2.f;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float g = V.itau[l] * u;
                g *= tS;
                rxx[l] += g;
                rxx[l] *= V.c2[l];
                xx += dth * rxx[l];
            }
        u *= oS;
        xx -= u;
        u = vxx * mu;
        u *= 2.f;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float g = V.itau[l] * u;
                g *= tS;
                ryy[l] += g;
                ryy[l] *= V.c2[l];
                yy += dth * ryy[l];
            }
        u *= oS;
        yy -= u;
        u2 += t2;
        u2 *= mxy;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                xy += dth * rxy[l];
                rxy[l] *= V.c1[l];
                float g = V.itau[l] * u2;
                g *= tSxy;
                rxy[l] -= g;
                rxy[l] *= V.c2[l];
                xy += dth * rxy[l];
            }
        u2 *= oS;
        xy += u2;
        if (fs && y == 0) {
            // FreeSurface2Dviscoelastic.cpp:14-45, then syy and every Ryy zero on the surface
#pragma unroll
            for (int l = 0; l < 4; l++)
                if (l < NL) xx -= dth * rxx[l];
            xx += C.fsH[col.x] * vxx;
            xx -= C.fsV[col.x] * vyy;
#pragma unroll
            for (int l = 0; l < 4; l++)
                if (l < NL) {
                    rxx[l] += V.fsRH[l][col.x] * vxx;
                    rxx[l] -= V.fsRV[l][col.x] * vyy;
                    xx += dth * rxx[l];
                    ryy[l] *= 0.f;
                }
            yy *= 0.f;
        }
        if (C.damp) {
            const float d = (C.damp + crow)[ux];
            xx *= d;
            yy *= d;
            xy *= d;
        }
        (F.f[2] + row)[ux] = xx;
        (F.f[3] + row)[ux] = yy;
        (F.f[4] + row)[ux] = xy;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                (V.R[l] + row)[ux] = rxx[l];
                (V.R[4 + l] + row)[ux] = ryy[l];
                (V.R[8 + l] + row)[ux] = rxy[l];
            }
    }
    block_acquisition<ELASTIC>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// visco-SH stress update (ForwardSolver2Dviscosh.cpp:130-185): sh2d_s (kernels2d.hip) with the
// memory variables Rxz, Ryz of each mechanism, in the reference's order:
//   u = mu_xz CPML(Dxf vz);  for l: sxz += dt/2 Rxz_l; Rxz_l = c2 (c1 Rxz_l - u tauS_xz / tau_l);
//       sxz += dt/2 Rxz_l;  sxz += (1 + L tauS) u;  the same for syz with Dyf, mu_yz, tauS_yz;
// then the ABS damping of the stresses.
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void vsh2d_s(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                    int tofs, Visco2D V)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_v[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    TileRegs<H, TY> rv;
    rv.load(F.f[0] + so, L, x0, y0, tid);
    float *__restrict__ sxz = F.f[1] + so;
    float *__restrict__ syz = F.f[2] + so;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int NL = V.L;
    float p1[R], q1[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        p1[r] = q1[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p1[r] = P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q1[r] = P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rv.store(s_v, tid);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyf = T.yf + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dx<H>(s_v, Rr, tx, cxf);
        float t = dy<H>(s_v, Rr, tx, cyf);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p1[r], col.axh, col.bxh);
        if (yi[r] >= 0) P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q1[r], P.ayh[y], P.byh[y]);
        const long long c = L.at(col.x, y), o = so + c;
        const float oS = V.opS[c], txz = V.tauSxz[c], tyz = V.tauSyz[c];
        const float dth = V.dth;
        float xz = F.f[1][o], yz = F.f[2][o];
        u *= C.c[1][c];
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float rr = V.R[l][o];
                xz += dth * rr;
                rr *= V.c1[l];
                float g = V.itau[l] * u;
                g *= txz;
                rr -= g;
                rr *= V.c2[l];
                xz += dth * rr;
                V.R[l][o] = rr;
            }
        u *= oS;
        xz += u;
        t *= C.c[2][c];
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float rr = V.R[4 + l][o];
                yz += dth * rr;
                rr *= V.c1[l];
                float g = V.itau[l] * t;
                g *= tyz;
                rr -= g;
                rr *= V.c2[l];
                yz += dth * rr;
                V.R[4 + l][o] = rr;
            }
        t *= oS;
        yz += t;
        if (C.damp) {
            const float d = C.damp[c];
            xz *= d;
            yz *= d;
        }
        sxz[c] = xz;
        syz[c] = yz;
    }
    block_acquisition<SH>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

template <typename Kern, typename... Args>
void go(Kern k, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0)
        hipExtLaunchKernelGGL(k, grid, block, 0, st, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, st, args...);
}

template <int H>
void launch_ty(int physics, int ty, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C, const Taps &T,
               const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    const dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + ty - 1) / ty, nshots), block(TILE_X, BLOCK_Y);
    if (physics == SH) {
        if (ty == 8)
            go(vsh2d_s<H, 8>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
        else if (ty == 32)
            go(vsh2d_s<H, 32>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
        else
            go(vsh2d_s<H, 16>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
    } else if (ty == 8)
        go(ve2d_stress<H, 8>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
    else if (ty == 32)
        go(ve2d_stress<H, 32>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
    else
        go(ve2d_stress<H, 16>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, V);
}

}  // namespace

void launch_visco2d_stress(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                           const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V,
                           hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: launch_ty<1>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 2: launch_ty<2>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 3: launch_ty<3>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 4: launch_ty<4>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 5: launch_ty<5>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 6: launch_ty<6>(physics, tile_y, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    default: break;
    }
}

}  // namespace wv
