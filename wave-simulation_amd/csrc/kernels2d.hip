// Tiled 2D staggered-grid kernels for gfx950: acoustic, elastic P-SV and TM-mode EM.
//
// One time step of the reference (ForwardSolver2Dacoustic.cpp:91-194,
// ForwardSolver2Delastic.cpp:117-294, ForwardSolver2Dtmem.cpp:108-171) is
// executed as two launches covering every shot of the batch:
//   phase 0: velocity-type update (vx, vy | hx, hy) from the stress-type fields;
//            first injects the velocity-type sources / gathers the
//            velocity-type receivers left pending by the previous step
//   phase 1: stress-type update (p | sxx, syy, sxy | ez) from the new
//            velocities, with CPML, free surface and ABS of the stress-type
//            fields fused; then injects the stress-type sources and gathers
//            the stress-type receivers of this step
// Each block stages its 64 x TY tile plus an H-wide halo of the fields it
// differentiates in LDS.  All global loads of a block (tile, centre values,
// coefficients, CPML memory variables) are issued before the first wait so the
// kernel streams at memory-level parallelism instead of serialised round trips.
// Derivative taps along x sit in registers (one column per lane), taps along y
// in scalar registers (one row per wavefront).
#include <algorithm>

#include <hip/hip_cooperative_groups.h>

#include "acq2d.hpp"

namespace wv {

namespace {


// ---------------------------------------------------------------------------
// elastic P-SV
// ---------------------------------------------------------------------------
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void el2d_velocity(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                          int tofs, int pend, Snap2D S)
{
    using RT = RowTile<H, TY>;
    constexpr int R = TY / BLOCK_Y, K = 2 * H + 1;
    __shared__ float s_xx[RT::RH][RT::RW];
    __shared__ float s_yy[RT::RH][RT::RW];
    __shared__ float s_xy[RT::RH][RT::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tx = threadIdx.x;
    const unsigned ux = threadIdx.x;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (pend) block_acquisition<ELASTIC>(A, F, so, shot_block(A, B), A.tbase[0] + tofs - 1, true);

    const long long corner = so + L.at(x0 - H, y0 - H);
    RT rxx, ryy, rxy;
    rxx.load(F.f[2] + corner, L.pitch, wy, tx);
    ryy.load(F.f[3] + corner, L.pitch, wy, tx);
    rxy.load(F.f[4] + corner, L.pitch, wy, tx);

    const Col col = column(L, P, x0 + tx);
    float cxf[K], cxb[K];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;

    float ovx[R], ovy[R], crx[R], cry[R], p0[R], p1[R], q0[R], q1[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long row = L.at(x0, y);  // wavefront-uniform
        ovx[r] = (F.f[0] + (so + row))[ux];
        ovy[r] = (F.f[1] + (so + row))[ux];
        crx[r] = (C.c[2] + row)[ux];
        cry[r] = (C.c[3] + row)[ux];
        p0[r] = p1[r] = q0[r] = q1[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) {
            const long long pr = ((long long)shot * L.NY + y) * W2;
            p0[r] = (P.psx[0] + pr)[(unsigned)col.xi];
            p1[r] = (P.psx[1] + pr)[(unsigned)col.xi];
        }
        if (yi[r] >= 0 && col.in) {
            const long long pr = ((long long)shot * W2 + yi[r]) * L.NX + x0;
            q0[r] = (P.psy[0] + pr)[ux];
            q1[r] = (P.psy[1] + pr)[ux];
        }
    }
    rxx.store(s_xx, wy, tx);
    ryy.store(s_yy, wy, tx);
    rxy.store(s_xy, wy, tx);
    __syncthreads();

#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyb = T.yb_v + (long long)y * K;
        const float *cyf = T.yf_v + (long long)y * K;
        const int Rr = ly + H;
        // vx += 1/rho_x * (CPML(Dxf sxx) + CPML(Dyb sxy))
        float u = dx<H>(s_xx, Rr, tx, cxf);
        float t = dy<H>(s_xy, Rr, tx, cyb);
        // vy += 1/rho_y * (CPML(Dxb sxy) + CPML(Dyf syy))
        float u2 = dx<H>(s_xy, Rr, tx, cxb);
        float t2 = dy<H>(s_yy, Rr, tx, cyf);
        if (!col.in) continue;
        const long long row = so + L.at(x0, y);
        if (col.xi >= 0) {
            const long long pr = ((long long)shot * L.NY + y) * W2;
            (P.psx[0] + pr)[(unsigned)col.xi] = cpml(u, p0[r], col.axh, col.bxh);
            (P.psx[1] + pr)[(unsigned)col.xi] = cpml(u2, p1[r], col.ax, col.bx);
        }
        if (yi[r] >= 0) {
            const long long pr = ((long long)shot * W2 + yi[r]) * L.NX + x0;
            (P.psy[0] + pr)[ux] = cpml(t, q0[r], P.ay[y], P.by[y]);
            (P.psy[1] + pr)[ux] = cpml(t2, q1[r], P.ayh[y], P.byh[y]);
        }
        u += t;
        u2 += t2;
        if (S.h[0]) {
            (S.h[0] + row)[ux] = u;
            (S.h[1] + row)[ux] = u2;
        }
        u *= crx[r];
        const float nvx = ovx[r] + u;
        u2 *= cry[r];
        const float nvy = ovy[r] + u2;
        (F.f[0] + row)[ux] = nvx;
        (F.f[1] + row)[ux] = nvy;
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void el2d_stress(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                        int tofs, int pend, Snap2D S)
{
    (void)pend;
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

    const long long corner = so + L.at(x0 - H, y0 - H);
    RT rvx, rvy;
    rvx.load(F.f[0] + corner, L.pitch, wy, tx);
    rvy.load(F.f[1] + corner, L.pitch, wy, tx);

    const Col col = column(L, P, x0 + tx);
    float cxf[K], cxb[K];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;

    float oxx[R], oyy[R], oxy[R], cM[R], cmu[R], cmxy[R], p2[R], p3[R], q2[R], q3[R], dmp[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long row = L.at(x0, y);  // wavefront-uniform
        oxx[r] = (F.f[2] + (so + row))[ux];
        oyy[r] = (F.f[3] + (so + row))[ux];
        oxy[r] = (F.f[4] + (so + row))[ux];
        cM[r] = (C.c[0] + row)[ux];
        cmu[r] = (C.c[1] + row)[ux];
        cmxy[r] = (C.c[4] + row)[ux];
        dmp[r] = C.damp ? (C.damp + row)[ux] : 1.f;
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
        float xx = oxx[r], yy = oyy[r], xy = oxy[r];
        float u = vxx + vyy;
        u *= cM[r];
        xx += u;
        yy += u;
        xx -= 2.f * (vyy * cmu[r]);
        yy -= 2.f * (vxx * cmu[r]);
        u2 += t2;
        if (S.h[2]) {
            const long long hr = so + L.at(x0, y);
            (S.h[2] + hr)[ux] = vxx;
            (S.h[3] + hr)[ux] = vyy;
            (S.h[4] + hr)[ux] = u2;
        }
        u2 *= cmxy[r];
        xy += u2;
        if (fs && y == 0) {
            // FreeSurface2Delastic.cpp:14-45, FreeSurface.cpp:14-21
            xx += C.fsH[col.x] * vxx;
            xx -= C.fsV[col.x] * vyy;
            yy *= 0.f;
        }
        xx *= dmp[r];
        yy *= dmp[r];
        xy *= dmp[r];
        const long long row = so + L.at(x0, y);
        (F.f[2] + row)[ux] = xx;
        (F.f[3] + row)[ux] = yy;
        (F.f[4] + row)[ux] = xy;
    }
    block_acquisition<ELASTIC>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// ---------------------------------------------------------------------------
// acoustic
// ---------------------------------------------------------------------------
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void ac2d_velocity(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                          int tofs, int pend)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_p[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    if (pend) block_acquisition<ACOUSTIC>(A, F, so, shot_block(A, B), A.tbase[0] + tofs - 1, true);

    TileRegs<H, TY> rp;
    rp.load(F.f[2] + so, L, x0, y0, tid);
    float *__restrict__ vx = F.f[0] + so;
    float *__restrict__ vy = F.f[1] + so;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float ovx[R], ovy[R], crx[R], cry[R], p0[R], q0[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        ovx[r] = vx[o];
        ovy[r] = vy[o];
        crx[r] = C.c[1][o];
        cry[r] = C.c[2][o];
        p0[r] = q0[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p0[r] = P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q0[r] = P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rp.store(s_p, tid);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyf = T.yf_v + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dx<H>(s_p, Rr, tx, cxf);
        float t = dy<H>(s_p, Rr, tx, cyf);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p0[r], col.axh, col.bxh);
        if (yi[r] >= 0) P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q0[r], P.ayh[y], P.byh[y]);
        const long long o = L.at(col.x, y);
        u *= crx[r];
        vx[o] = ovx[r] + u;
        t *= cry[r];
        vy[o] = ovy[r] + t;
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void ac2d_pressure(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                          int tofs, int pend)
{
    (void)pend;
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_vx[TL::RH][TL::RW];
    __shared__ float s_vy[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    TileRegs<H, TY> rvx, rvy;
    rvx.load(F.f[0] + so, L, x0, y0, tid);
    rvy.load(F.f[1] + so, L, x0, y0, tid);
    float *__restrict__ pr = F.f[2] + so;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float op[R], cM[R], p1[R], q1[R], dmp[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        op[r] = pr[o];
        cM[r] = C.c[0][o];
        dmp[r] = C.damp ? C.damp[o] : 1.f;
        p1[r] = q1[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p1[r] = P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q1[r] = P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rvx.store(s_vx, tid);
    rvy.store(s_vy, tid);
    __syncthreads();
    const bool fs = C.free_surface != 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dx<H>(s_vx, Rr, tx, cxb);
        float t = dy<H>(s_vy, Rr, tx, cyb);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p1[r], col.ax, col.bx);
        if (yi[r] >= 0) P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q1[r], P.ay[y], P.by[y]);
        u += t;
        u *= cM[r];
        float p = op[r] + u;
        p *= dmp[r];
        if (fs && y == 0) p *= 0.f;
        pr[L.at(col.x, y)] = p;
    }
    block_acquisition<ACOUSTIC>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// ---------------------------------------------------------------------------
// TM-mode EM
// ---------------------------------------------------------------------------
// LDS of one TM-EM H tile
template <int H, int TY>
struct TmHLds {
    float ez[Tile<H, TY>::RH][Tile<H, TY>::RW];
    float ty[TY][2 * H + 1];  // Dyf taps of the tile's rows
    float cy[TY][2];          // a_y(half), b_y(half) of the tile's rows (0 outside the frame)
};

// H update of tile B (called by tm2d_h with the block's own tile, and tile after tile by the
// persistent kernel; the LDS is reused only after a barrier)
template <int H, int TY>
__device__ __forceinline__ void tm_h_tile(const Blk &B, const Layout2D &L, const Fields2D &F, const Coef2D &C,
                                          const Taps &T, const Cpml2D &P, TmHLds<H, TY> &S)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    auto &s_ez = S.ez;
    auto &s_ty = S.ty;
    auto &s_cy = S.cy;
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    // per-row tables first, so waiting for them does not wait for the field loads
    constexpr int NTY = TY * (2 * H + 1), ITY = (NTY + NTHREADS - 1) / NTHREADS;
    float tyv[ITY], cyv = 0.f;
#pragma unroll
    for (int k = 0; k < ITY; k++) tyv[k] = T.yf[(long long)y0 * (2 * H + 1) + min(tid + k * NTHREADS, NTY - 1)];
    if (tid < 2 * TY && P.on) cyv = (tid & 1 ? P.byh : P.ayh)[y0 + (tid >> 1)];
    TileRegs<H, TY> rez;
    rez.load(F.f[2] + so, L, x0, y0, tid);
    float *__restrict__ hx = F.f[0] + so;
    float *__restrict__ hy = F.f[1] + so;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float ohx[R], ohy[R], cxz[R], cyz[R], p0[R], q0[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        ohx[r] = hx[o];
        ohy[r] = hy[o];
        cxz[r] = C.c[2][o];
        cyz[r] = C.c[3][o];
        p0[r] = q0[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p0[r] = P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q0[r] = P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rez.store(s_ez, tid);
#pragma unroll
    for (int k = 0; k < ITY; k++)
        if (tid + k * NTHREADS < NTY) (&s_ty[0][0])[tid + k * NTHREADS] = tyv[k];
    if (tid < 2 * TY) (&s_cy[0][0])[tid] = cyv;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int Rr = ly + H;
        float u = dy<H>(s_ez, Rr, tx, s_ty[ly]);
        float u2 = dx<H>(s_ez, Rr, tx, cxf);
        if (!col.in) continue;
        if (yi[r] >= 0) P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(u, q0[r], s_cy[ly][0], s_cy[ly][1]);
        if (col.xi >= 0) P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u2, p0[r], col.axh, col.bxh);
        // hx -= 1/mu_yz * CPML(Dyf ez) ; hy -= 1/mu_xz * (-CPML(Dxf ez))
        const long long o = L.at(col.x, y);
        u *= cyz[r];
        hx[o] = ohx[r] - u;
        float w = -u2;
        w *= cxz[r];
        hy[o] = ohy[r] - w;
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void tm2d_h(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend)
{
    (void)A;
    (void)tofs;
    (void)pend;
    __shared__ TmHLds<H, TY> S;
    tm_h_tile<H, TY>(xcd_block(), L, F, C, T, P, S);
}

// VE: visco TM-EM (ForwardSolver2Dviscotmem.cpp:172-199): per mechanism r_l = Cd_l ez + Cc_l r_l
// from the ez of the previous step, then ez = Ca ez + Cb ((Dxb hy - Dyb hx) - sum_l DT r_l)
// LDS of one TM-EM E tile
template <int H, int TY>
struct TmELds {
    float hx[Tile<H, TY>::RH][Tile<H, TY>::RW];
    float hy[Tile<H, TY>::RH][Tile<H, TY>::RW];
    float ty[TY][2 * H + 1];  // Dyb taps of the tile's rows
    float cy[TY][2];          // a_y, b_y of the tile's rows (0 outside the frame)
};

// E update of tile B at step tbase + tofs, then its EZ sources / receivers
template <int H, int TY, bool VE>
__device__ __forceinline__ void tm_e_tile(const Blk &B, const Layout2D &L, const Fields2D &F, const Coef2D &C,
                                          const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V,
                                          TmELds<H, TY> &S)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    auto &s_hx = S.hx;
    auto &s_hy = S.hy;
    auto &s_ty = S.ty;
    auto &s_cy = S.cy;
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    // per-row tables first, so waiting for them does not wait for the field loads
    constexpr int NTY = TY * (2 * H + 1), ITY = (NTY + NTHREADS - 1) / NTHREADS;
    float tyv[ITY], cyv = 0.f;
#pragma unroll
    for (int k = 0; k < ITY; k++) tyv[k] = T.yb[(long long)y0 * (2 * H + 1) + min(tid + k * NTHREADS, NTY - 1)];
    if (tid < 2 * TY && P.on) cyv = (tid & 1 ? P.by : P.ay)[y0 + (tid >> 1)];
    TileRegs<H, TY> rhx, rhy;
    rhx.load(F.f[0] + so, L, x0, y0, tid);
    rhy.load(F.f[1] + so, L, x0, y0, tid);
    float *__restrict__ ez = F.f[2] + so;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float oez[R], ca[R], cb[R], p1[R], q1[R], dmp[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        oez[r] = ez[o];
        ca[r] = C.c[0][o];
        cb[r] = C.c[1][o];
        dmp[r] = C.damp ? C.damp[o] : 1.f;
        p1[r] = q1[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p1[r] = P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q1[r] = P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rhx.store(s_hx, tid);
    rhy.store(s_hy, tid);
#pragma unroll
    for (int k = 0; k < ITY; k++)
        if (tid + k * NTHREADS < NTY) (&s_ty[0][0])[tid + k * NTHREADS] = tyv[k];
    if (tid < 2 * TY) (&s_cy[0][0])[tid] = cyv;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int Rr = ly + H;
        float u = dx<H>(s_hy, Rr, tx, cxb);
        float t = dy<H>(s_hx, Rr, tx, s_ty[ly]);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p1[r], col.ax, col.bx);
        if (yi[r] >= 0) P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q1[r], s_cy[ly][0], s_cy[ly][1]);
        u -= t;
        if constexpr (VE) {
            const long long c0 = L.at(col.x, y), o = so + c0;
#pragma unroll
            for (int l = 0; l < 4; l++)
                if (l < V.L) {
                    const float a = V.cc[l] * V.R[l][o];
                    const float b = V.cd[l][c0] * oez[r];
                    const float rn = b + a;
                    V.R[l][o] = rn;
                    u -= V.dt * rn;
                }
        }
        u *= cb[r];
        float e = ca[r] * oez[r] + u;
        e *= dmp[r];
        ez[L.at(col.x, y)] = e;
    }
    block_acquisition<TMEM>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

template <int H, int TY, bool VE = false>
__global__ __launch_bounds__(NTHREADS) void tm2d_e(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend, Visco2D V)
{
    (void)pend;
    __shared__ TmELds<H, TY> S;
    tm_e_tile<H, TY, VE>(xcd_block(), L, F, C, T, P, A, tofs, V, S);
}

// Persistent TM-EM step loop: one cooperative launch runs nsteps time steps of every shot.  The
// resident blocks take the tiles of the H update in turn, meet at a grid barrier, take the tiles of
// the E update (with its sources / receivers), meet again; between two tiles of one block the LDS
// is reused after a block barrier.  ABS is not supported here (the two-launch path damps H in a
// third launch).  Tiles: as the two-launch kernels' grid, numbered shot-major.
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void tm2d_persistent(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P,
                                                            Acq A, int nsteps, int nbx, int nby, int nshots)
{
    __shared__ TmHLds<H, TY> SH;
    __shared__ TmELds<H, TY> SE;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int ntiles = nbx * nby * nshots;
    for (int j = 0; j < nsteps; j++) {
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            Blk B;
            B.bx = t % nbx;
            B.by = (t / nbx) % nby;
            B.shot = t / (nbx * nby);
            tm_h_tile<H, TY>(B, L, F, C, T, P, SH);
            __syncthreads();
        }
        grid.sync();
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            Blk B;
            B.bx = t % nbx;
            B.by = (t / nbx) % nby;
            B.shot = t / (nbx * nby);
            tm_e_tile<H, TY, false>(B, L, F, C, T, P, A, j, Visco2D{}, SE);
            __syncthreads();
        }
        grid.sync();
    }
}

// ---------------------------------------------------------------------------
// SH (ForwardSolver2Dsh.cpp:101-190): phase 0 vz from the stresses, phase 1 the
// stresses from vz; VZ sources / receivers are velocity-type
// ---------------------------------------------------------------------------
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void sh2d_v(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_xz[TL::RH][TL::RW];
    __shared__ float s_yz[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    if (pend) block_acquisition<SH>(A, F, so, shot_block(A, B), A.tbase[0] + tofs - 1, true);
    TileRegs<H, TY> rxz, ryz;
    rxz.load(F.f[1] + so, L, x0, y0, tid);
    ryz.load(F.f[2] + so, L, x0, y0, tid);
    float *__restrict__ vz = F.f[0] + so;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float ov[R], cr[R], p0[R], q0[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        ov[r] = vz[o];
        cr[r] = C.c[0][o];
        p0[r] = q0[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
        if (col.xi >= 0 && y < L.NY) p0[r] = P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q0[r] = P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rxz.store(s_xz, tid);
    ryz.store(s_yz, tid);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        // Dyb of syz (the free-surface image operator when FreeSurface = 1)
        const float *cyb = T.yb_v + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dx<H>(s_xz, Rr, tx, cxb);
        float t = dy<H>(s_yz, Rr, tx, cyb);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p0[r], col.ax, col.bx);
        if (yi[r] >= 0) P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q0[r], P.ay[y], P.by[y]);
        u += t;
        u *= cr[r];
        vz[L.at(col.x, y)] = ov[r] + u;
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void sh2d_s(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend)
{
    (void)pend;
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
    float oxz[R], oyz[R], cxz[R], cyz[R], p1[R], q1[R], dmp[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        oxz[r] = sxz[o];
        oyz[r] = syz[o];
        cxz[r] = C.c[1][o];
        cyz[r] = C.c[2][o];
        dmp[r] = C.damp ? C.damp[o] : 1.f;
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
        const long long o = L.at(col.x, y);
        u *= cxz[r];
        sxz[o] = (oxz[r] + u) * dmp[r];
        t *= cyz[r];
        syz[o] = (oyz[r] + t) * dmp[r];
    }
    block_acquisition<SH>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// ---------------------------------------------------------------------------
// TE-EM (ForwardSolver2Demem.cpp:108-171): phase 0 hz from ex, ey (forward
// derivatives; CPMLEM2D eyx / exy on the half profiles), phase 1 ex, ey from hz
// (backward derivatives; hzy / hzx on the full profiles).  HZ sources and
// receivers are velocity-type (the field phase 0 writes), EX / EY stress-type.
// ---------------------------------------------------------------------------
template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void te2d_h(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_ex[TL::RH][TL::RW];
    __shared__ float s_ey[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    if (pend) block_acquisition<TEMEM>(A, F, so, shot_block(A, B), A.tbase[0] + tofs - 1, true);
    TileRegs<H, TY> rex, rey;
    rex.load(F.f[1] + so, L, x0, y0, tid);
    rey.load(F.f[2] + so, L, x0, y0, tid);
    float *__restrict__ hz = F.f[0] + so;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float oh[R], cm[R], p0[R], q0[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        oh[r] = hz[o];
        cm[r] = C.c[0][o];
        p0[r] = q0[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, 0) : -1;
        if (col.xi >= 0 && y < L.NY) p0[r] = P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q0[r] = P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rex.store(s_ex, tid);
    rey.store(s_ey, tid);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyf = T.yf + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dx<H>(s_ey, Rr, tx, cxf);
        float t = dy<H>(s_ex, Rr, tx, cyf);
        if (!col.in) continue;
        if (col.xi >= 0) P.psx[0][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(u, p0[r], col.axh, col.bxh);
        if (yi[r] >= 0) P.psy[0][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(t, q0[r], P.ayh[y], P.byh[y]);
        // hz -= 1/mu_xy * (CPML(Dxf ey) - CPML(Dyf ex))
        u -= t;
        u *= cm[r];
        hz[L.at(col.x, y)] = oh[r] - u;
    }
}

// VE: visco TE-EM (ForwardSolver2Dviscoemem.cpp:175-218): r_l = Cd_l e + Cc_l r_l for ex and ey
// (x- and y-averaged Cd), then ex = Ca_x ex + Cb_x (CPML(Dyb hz) - sum_l DT rx_l),
// ey = Ca_y ey + Cb_y (-CPML(Dxb hz) - sum_l DT ry_l)
template <int H, int TY, bool VE = false>
__global__ __launch_bounds__(NTHREADS) void te2d_e(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                   int tofs, int pend, Visco2D V)
{
    (void)pend;
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_h[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    TileRegs<H, TY> rh;
    rh.load(F.f[0] + so, L, x0, y0, tid);
    float *__restrict__ ex = F.f[1] + so;
    float *__restrict__ ey = F.f[2] + so;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    float oex[R], oey[R], p1[R], q1[R];
    int yi[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + wy + r * BLOCK_Y;
        const long long o = L.at(col.x, y);
        oex[r] = ex[o];
        oey[r] = ey[o];
        p1[r] = q1[r] = 0.f;
        yi[r] = P.on ? frame_index(y, L.NY, P.W, 0) : -1;
        if (col.xi >= 0 && y < L.NY) p1[r] = P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi];
        if (yi[r] >= 0 && col.in) q1[r] = P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x];
    }
    rh.store(s_h, tid);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const int Rr = ly + H;
        float u = dy<H>(s_h, Rr, tx, cyb);
        float w = dx<H>(s_h, Rr, tx, cxb);
        if (!col.in) continue;
        if (yi[r] >= 0) P.psy[1][((long long)shot * W2 + yi[r]) * L.NX + col.x] = cpml(u, q1[r], P.ay[y], P.by[y]);
        if (col.xi >= 0) P.psx[1][((long long)shot * L.NY + y) * W2 + col.xi] = cpml(w, p1[r], col.ax, col.bx);
        const long long o = L.at(col.x, y);
        const float dmp = C.damp ? C.damp[o] : 1.f;
        // ex = Ca_x ex + Cb_x CPML(Dyb hz) ; ey = Ca_y ey - Cb_y CPML(Dxb hz)
        if constexpr (VE) {
            const long long os = so + o;
            float v = -w;
#pragma unroll
            for (int l = 0; l < 4; l++)
                if (l < V.L) {
                    float a = V.cc[l] * V.R[l][os];
                    float b = V.cd[l][o] * oex[r];
                    const float rx = b + a;
                    a = V.cc[l] * V.R[4 + l][os];
                    b = V.cd[4 + l][o] * oey[r];
                    const float ry = b + a;
                    V.R[l][os] = rx;
                    V.R[4 + l][os] = ry;
                    u -= V.dt * rx;
                    v -= V.dt * ry;
                }
            u *= C.c[2][o];
            float e = C.c[1][o] * oex[r];
            e += u;
            ex[o] = e * dmp;
            v *= C.c[4][o];
            float f = C.c[3][o] * oey[r];
            f += v;
            ey[o] = f * dmp;
            continue;
        }
        u *= C.c[2][o];
        float e = C.c[1][o] * oex[r];
        e += u;
        ex[o] = e * dmp;
        w *= C.c[4][o];
        float f = C.c[3][o] * oey[r];
        f -= w;
        ey[o] = f * dmp;
    }
    block_acquisition<TEMEM>(A, F, so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// velocity-type sources / receivers of step t for every shot (end of a run segment)
template <int PHYS>
__global__ __launch_bounds__(256) void flush2d(Layout2D L, Fields2D F, Acq A, int tofs)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.plane;
    const int t = A.tbase[0] + tofs;
    if (t < 0 || t >= A.NT) return;
    if (threadIdx.x == 0) {
        const int s0 = A.src_ptr[shot], s1 = A.src_ptr[shot + 1];
        for (int ty = 2; ty <= 4; ty++)
            for (int s = s0; s < s1; s++) {
                if (A.src_type[s] != ty || !velocity_type_of<PHYS>(ty)) continue;
                F.f[velocity_field<PHYS>(ty)][so + A.src_cell[s]] += A.src_sig[(long long)s * A.NT + t];
            }
    }
    __syncthreads();
    const int r0 = A.rec_ptr[shot], r1 = A.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int ty = A.rec_type[r];
        if (!velocity_type_of<PHYS>(ty)) continue;
        A.seis[(long long)t * A.rec_total + r] = F.f[velocity_field<PHYS>(ty)][so + A.rec_cell[r]];
    }
}

// ABS of the velocity-type fields (b may be null: SH has one, vz; TE-EM one, hz)
__global__ void abs_velocity(long long n, float *a, float *b, const float *__restrict__ damp, long long plane)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = damp[i % plane];
        a[i] *= d;
        if (b) b[i] *= d;
    }
}

__global__ void advance_kernel(int *tbase, int k)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) tbase[0] += k;
}

__global__ void set_kernel(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// launch `k`, bracketed by e0/e1 at the kernel's own start and end when given
template <typename Kern, typename... Args>
void go(Kern k, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0)
        hipExtLaunchKernelGGL(k, grid, block, 0, st, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, st, args...);
}

template <int H, int TY>
void launch_h(int physics, int which, dim3 grid, const Layout2D &L, const Fields2D &F, const Coef2D &C, const Taps &T,
              const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
              const Snap2D &S)
{
    dim3 block(TILE_X, BLOCK_Y);
    if (physics == ELASTIC) {
        if (which == 0)
            go(el2d_velocity<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, S);
        else
            go(el2d_stress<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, S);
    } else if (physics == ACOUSTIC) {
        if (which == 0)
            go(ac2d_velocity<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
        else
            go(ac2d_pressure<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
    } else if (physics == SH) {
        if (which == 0)
            go(sh2d_v<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
        else
            go(sh2d_s<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
    } else if (physics == TEMEM) {
        if (which == 0)
            go(te2d_h<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
        else
            go(te2d_e<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, Visco2D{});
    } else {
        if (which == 0)
            go(tm2d_h<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
        else
            go(tm2d_e<H, TY>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, Visco2D{});
    }
}

template <int H>
void launch_vem_ty(int ty, int physics, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                   const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V, hipStream_t st,
                   hipEvent_t e0, hipEvent_t e1)
{
    const int TY = ty == 8 ? 8 : (ty == 32 ? 32 : 16);
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + TY - 1) / TY, nshots), block(TILE_X, BLOCK_Y);
    auto pick = [&](auto k) { go(k, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, 0, V); };
    if (physics == TEMEM) {
        if (TY == 8)
            pick(te2d_e<H, 8, true>);
        else if (TY == 32)
            pick(te2d_e<H, 32, true>);
        else
            pick(te2d_e<H, 16, true>);
    } else {
        if (TY == 8)
            pick(tm2d_e<H, 8, true>);
        else if (TY == 32)
            pick(tm2d_e<H, 32, true>);
        else
            pick(tm2d_e<H, 16, true>);
    }
}

template <int H>
void launch_ty2(int ty, int physics, int which, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                hipEvent_t e1, const Snap2D &S)
{
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + ty - 1) / ty, nshots);
    if (ty == 8)
        launch_h<H, 8>(physics, which, grid, L, F, C, T, P, A, tofs, pend, st, e0, e1, S);
    else if (ty == 32)
        launch_h<H, 32>(physics, which, grid, L, F, C, T, P, A, tofs, pend, st, e0, e1, S);
    else
        launch_h<H, 16>(physics, which, grid, L, F, C, T, P, A, tofs, pend, st, e0, e1, S);
}

}  // namespace

void launch_step2d(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                   const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend,
                   hipStream_t stream, int which, hipEvent_t e0, hipEvent_t e1, const Snap2D *snap)
{
    const Snap2D S = snap ? *snap : Snap2D{};
    switch (order / 2) {
    case 1: launch_ty2<1>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    case 2: launch_ty2<2>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    case 3: launch_ty2<3>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    case 4: launch_ty2<4>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    case 5: launch_ty2<5>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    case 6: launch_ty2<6>(tile_y, physics, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, S); break;
    default: break;
    }
}

namespace {
template <int H, int TY>
hipError_t tm_persist_ty(const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C, const Taps &T,
                         const Cpml2D &P, const Acq &A, int nsteps, hipStream_t st, int *grid_out)
{
    auto k = tm2d_persistent<H, TY>;
    int dev = 0, cus = 0, nb = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, NTHREADS, 0);
    if (e != hipSuccess) return e;
    int nbx = (L.NX + TILE_X - 1) / TILE_X, nby = (L.NY + TY - 1) / TY, ns = nshots, steps = nsteps;
    // every block of a cooperative launch must be resident at once
    const int grid = std::min(nbx * nby * ns, nb * cus);
    if (grid < 1) return hipErrorInvalidConfiguration;
    if (grid_out) *grid_out = grid;
    Layout2D L_ = L;
    Fields2D F_ = F;
    Coef2D C_ = C;
    Taps T_ = T;
    Cpml2D P_ = P;
    Acq A_ = A;
    void *args[] = {&L_, &F_, &C_, &T_, &P_, &A_, &steps, &nbx, &nby, &ns};
    return hipLaunchCooperativeKernel((const void *)k, dim3(grid), dim3(TILE_X, BLOCK_Y), args, 0, st);
}
template <int H>
hipError_t tm_persist_h(int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C, const Taps &T,
                        const Cpml2D &P, const Acq &A, int nsteps, hipStream_t st, int *grid_out)
{
    if (tile_y == 8) return tm_persist_ty<H, 8>(L, nshots, F, C, T, P, A, nsteps, st, grid_out);
    if (tile_y == 32) return tm_persist_ty<H, 32>(L, nshots, F, C, T, P, A, nsteps, st, grid_out);
    return tm_persist_ty<H, 16>(L, nshots, F, C, T, P, A, nsteps, st, grid_out);
}
}  // namespace

hipError_t launch_tm2d_persistent(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                                  const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int nsteps,
                                  hipStream_t stream, int *grid_out)
{
    switch (order / 2) {
    case 1: return tm_persist_h<1>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    case 2: return tm_persist_h<2>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    case 3: return tm_persist_h<3>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    case 4: return tm_persist_h<4>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    case 5: return tm_persist_h<5>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    case 6: return tm_persist_h<6>(tile_y, L, nshots, F, C, T, P, A, nsteps, stream, grid_out);
    default: return hipErrorInvalidValue;
    }
}

void launch_viscoem2d_e(int physics, int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F,
                        const Coef2D &C, const Taps &T, const Cpml2D &P, const Acq &A, int tofs, const Visco2D &V,
                        hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: launch_vem_ty<1>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 2: launch_vem_ty<2>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 3: launch_vem_ty<3>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 4: launch_vem_ty<4>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 5: launch_vem_ty<5>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    case 6: launch_vem_ty<6>(tile_y, physics, L, nshots, F, C, T, P, A, tofs, V, stream, e0, e1); break;
    default: break;
    }
}

void launch_flush2d(int physics, const Layout2D &L, const Fields2D &F, const Acq &A, int tofs, hipStream_t stream)
{
    if (physics == ELASTIC)
        flush2d<ELASTIC><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else if (physics == ACOUSTIC)
        flush2d<ACOUSTIC><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else if (physics == SH)
        flush2d<SH><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else if (physics == TEMEM)
        flush2d<TEMEM><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
}

void launch_abs_velocity2d(int physics, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                           hipStream_t stream)
{
    const long long n = (long long)nshots * L.plane;
    // SH: vz; TE-EM: hz (phase 0's field; ex, ey are damped by phase 1)
    abs_velocity<<<1024, 256, 0, stream>>>(n, F.f[0], (physics == SH || physics == TEMEM) ? nullptr : F.f[1], C.damp,
                                           L.plane);
}

void launch_advance(int *tbase, int k, hipStream_t stream) { advance_kernel<<<1, 64, 0, stream>>>(tbase, k); }
void launch_set(int *p, int v, hipStream_t stream) { set_kernel<<<1, 64, 0, stream>>>(p, v); }

}  // namespace wv
