// Single-pass 2D elastic P-SV time step for gfx950.
//
// One launch per time step, all shots of the batch.  A block owns an output
// tile of (64 - 2H) x TY cells and
//   A) loads, with direct global->LDS transfers (global_load_lds, no staging
//      registers, every load of the block in flight at once):
//        the stress fields of step t-1 on the tile plus a 2H halo,
//        the old velocities and 1/rho of the tile plus an H halo (one column
//        per lane), and the stress-update coefficients of the tile;
//   B) recomputes the velocities of step t on the tile plus the H halo into
//      LDS, writing to HBM (with the CPML memory variables) only the cells it
//      owns;
//   C) updates the stresses of its own cells from the LDS velocities, with
//      CPML, free surface and ABS fused;
// so every field and coefficient crosses HBM once per step (halo re-reads hit
// L2).  Fields and the CPML memory variables of the velocity derivatives are
// ping-ponged (read step t-1 from F, write step t to Fn): blocks recompute
// their neighbours' halo cells, so nothing a neighbour may still read is
// overwritten within the launch.  Velocity-type sources of step t-1 and the
// ABS damping of the velocities are applied to the old velocities before use
// (every block that recomputes a cell applies them identically); stress-type
// sources and receivers are handled by the owner at the end.  Tiles are mapped
// to blocks so that each XCD works on a contiguous range of tiles (its L2 sees
// the halos of its neighbours).  Offsets inside a field plane are 32-bit
// (planes are checked to be < 2^31 elements at init).
//
// Reference: ForwardSolver2Delastic.cpp:117-294 (update order), CPML.cpp:84-95,
// FreeSurface2Delastic.cpp:14-45, ABS2D.cpp:14-48, SourceReceiverImpl.cpp:12-37.
#include "wave_internal.hpp"

#include <cstdlib>

namespace wv {

namespace {

constexpr int NTH = TILE_X * BLOCK_Y;
typedef __attribute__((address_space(3))) void lds_void;

template <int H, int TY>
struct FT {
    static constexpr int TXO = TILE_X - 2 * H;  // output columns per tile
    static constexpr int VH = TY + 2 * H;       // velocity region rows
    static constexpr int SW = TILE_X + 2 * H;   // stress region columns
    static constexpr int SH = TY + 4 * H;       // stress region rows
    static constexpr int STOT = SW * SH;
    static constexpr int SCH = (STOT + TILE_X - 1) / TILE_X;  // 64-float chunks of the stress region
    static constexpr int SPAD = SCH * TILE_X;                 // LDS floats per stress field
};

__device__ __forceinline__ float cpml_step(float &d, float psi, float a, float b)
{
    float p = b * psi;
    p = p + a * d;
    d = d + p;
    return p;
}

__device__ __forceinline__ bool vel_type(int ty) { return ty == 2 || ty == 3; }

// one lane's float from global `g` into LDS at lds_row + lane (lds_row wave-uniform)
__device__ __forceinline__ void dma(const float *g, float *lds_row)
{
    __builtin_amdgcn_global_load_lds((const void *)g, (lds_void *)lds_row, 4, 0, 0);
}

template <int H, int TY, bool ABS, bool DMA>
__global__ __launch_bounds__(NTH) void el2d_fused(Layout2D L, Fields2D F, Fields2D Fn, Coef2D C, Taps T, Cpml2D P,
                                                  PsiNew2D PN, Acq A, int tofs, int pend, int nbx, int nby, int nshots)
{
    using G = FT<H, TY>;
    constexpr int K = 2 * H + 1;
    constexpr int NCF = ABS ? 4 : 3;
    constexpr int NOV = ABS ? 5 : 4;
    __shared__ float s_st[3][G::SPAD];         // stress region, flat [SH][SW]
    __shared__ float s_cf[NCF][TY][TILE_X];    // M, mu, mu_xy (, damping) of the tile rows
    __shared__ float s_ov[NOV][G::VH][TILE_X]; // old vx, vy, 1/rho_x, 1/rho_y (, damping) -> new vx, vy in 0, 1

    // XCD-aware tile order: blocks b, b+8, b+16, ... run on one XCD; give each
    // XCD a contiguous range of (spatial tile, shot) indices.
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int xcd = b & 7, kk = b >> 3, q = nblk >> 3, rem = nblk & 7;
    const int tile = xcd < rem ? xcd * (q + 1) + kk : rem * (q + 1) + (xcd - rem) * q + kk;
    const int shot = tile % nshots;
    const int sp = tile / nshots;
    const int by = sp / nbx, bx = sp - by * nbx;
    const int x0 = bx * G::TXO, y0 = by * TY;
    const long long so = (long long)shot * L.plane;
    const int sb = (shot * nby + by) * nbx + bx;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int W2 = 2 * P.W;
    const int pitch = L.pitch;
    const int xv = x0 - H + tx;                 // this lane's velocity column
    const int obase = (int)L.at(xv, y0 - H);    // offset of (xv, y0-H)

    // ---- A: stage the stress region, the velocity region and the tile coefficients in LDS
    {
        const int base = (int)L.at(x0 - 2 * H, y0 - 2 * H);
        const float *ov[NOV];
        ov[0] = F.f[0] + so;
        ov[1] = F.f[1] + so;
        ov[2] = C.c[2];
        ov[3] = C.c[3];
        if (ABS) ov[NOV - 1] = C.damp;
        const float *cf[NCF];
        cf[0] = C.c[0];
        cf[1] = C.c[1];
        cf[2] = C.c[4];
        if (ABS) cf[NCF - 1] = C.damp;
        const int cbase = obase + H * pitch;
        if constexpr (DMA) {
            // direct global->LDS transfers: no staging registers
#pragma unroll
            for (int f = 0; f < 3; f++) {
                const float *g = F.f[2 + f] + so + base;
                for (int c = wy; c < G::SCH; c += BLOCK_Y) {
                    int i = c * TILE_X + tx;
                    if (i >= G::STOT) i = 0;  // tail lanes fill the LDS padding
                    const int r = i / G::SW;
                    dma(g + r * pitch + (i - r * G::SW), &s_st[f][c * TILE_X]);
                }
            }
#pragma unroll
            for (int f = 0; f < NOV; f++)
                for (int j = wy; j < G::VH; j += BLOCK_Y) dma(ov[f] + obase + j * pitch, &s_ov[f][j][0]);
#pragma unroll
            for (int f = 0; f < NCF; f++)
                for (int j = wy; j < TY; j += BLOCK_Y) dma(cf[f] + cbase + j * pitch, &s_cf[f][j][0]);
        } else {
            // every load issued into registers before the first LDS write
            constexpr int SC = (G::SCH + BLOCK_Y - 1) / BLOCK_Y;   // stress chunks per wave
            constexpr int VR = (G::VH + BLOCK_Y - 1) / BLOCK_Y;    // velocity-region rows per wave
            constexpr int CR = TY / BLOCK_Y;                        // tile rows per wave
            float rs[3][SC], rv[NOV][VR], rc[NCF][CR];
#pragma unroll
            for (int f = 0; f < 3; f++) {
                const float *g = F.f[2 + f] + so + base;
#pragma unroll
                for (int k = 0; k < SC; k++) {
                    const int c = wy + k * BLOCK_Y;
                    int i = c * TILE_X + tx;
                    if (c >= G::SCH || i >= G::STOT) i = 0;
                    const int r = i / G::SW;
                    rs[f][k] = g[r * pitch + (i - r * G::SW)];
                }
            }
#pragma unroll
            for (int f = 0; f < NOV; f++)
#pragma unroll
                for (int k = 0; k < VR; k++) {
                    const int j = min(wy + k * BLOCK_Y, G::VH - 1);
                    rv[f][k] = ov[f][obase + j * pitch];
                }
#pragma unroll
            for (int f = 0; f < NCF; f++)
#pragma unroll
                for (int k = 0; k < CR; k++) rc[f][k] = cf[f][cbase + (wy + k * BLOCK_Y) * pitch];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < SC; k++) {
                    const int c = wy + k * BLOCK_Y;
                    if (c < G::SCH) s_st[f][c * TILE_X + tx] = rs[f][k];
                }
#pragma unroll
            for (int f = 0; f < NOV; f++)
#pragma unroll
                for (int k = 0; k < VR; k++) {
                    const int j = wy + k * BLOCK_Y;
                    if (j < G::VH) s_ov[f][j][tx] = rv[f][k];
                }
#pragma unroll
            for (int f = 0; f < NCF; f++)
#pragma unroll
                for (int k = 0; k < CR; k++) s_cf[f][wy + k * BLOCK_Y][tx] = rc[f][k];
        }
    }

    // ---- per-column state while the loads are in flight
    const bool xin = xv >= 0 && xv < L.NX;
    const bool own_lane = tx >= H && tx < TILE_X - H;
    const bool own_x = own_lane && xv < L.NX;
    float cxf[K], cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        cxf[k] = T.xf[xv * K + k];
        cxb[k] = T.xb[xv * K + k];
    }
    int xi = -1;
    float ax = 0.f, bxx = 0.f, axh = 0.f, bxh = 0.f;
    if (P.on && xin) {
        xi = frame_index(xv, L.NX, P.W, false);
        if (xi >= 0) {
            ax = P.ax[xv];
            bxx = P.bx[xv];
            axh = P.axh[xv];
            bxh = P.bxh[xv];
        }
    }
    if constexpr (DMA) __builtin_amdgcn_s_waitcnt(0);  // the direct LDS loads have landed
    __syncthreads();

    // ---- pending ABS damping and velocity-type sources / receivers of step t-1
    // (each (column, row) of the velocity region is touched by one thread only:
    //  the one that computes it in B)
    if (pend) {
        if (ABS)
            for (int j = wy; j < G::VH; j += BLOCK_Y) {
                const float d = s_ov[NOV - 1][j][tx];
                s_ov[0][j][tx] *= d;
                s_ov[1][j][tx] *= d;
            }
        const int tp = A.tbase[0] + tofs - 1;
        const int e0 = A.vsrc_ptr[sb], e1 = A.vsrc_ptr[sb + 1];
        for (int e = e0; e < e1; e++) {
            const int s = A.vsrc[e];
            const long long cell = A.src_cell[s];
            const int sy = (int)(cell / pitch) - L.ypad, sx = (int)(cell % pitch) - XPAD;
            const int lane = sx - (x0 - H), j = sy - (y0 - H);
            if (tx == lane && (j & (BLOCK_Y - 1)) == wy)
                s_ov[A.src_type[s] == 2 ? 0 : 1][j][tx] += A.src_sig[(long long)s * A.NT + tp];
        }
        const int r0 = A.brec_ptr[sb], r1 = A.brec_ptr[sb + 1];
        for (int e = r0; e < r1; e++) {
            const int rc = A.brec[e];
            const int ty = A.rec_type[rc];
            if (!vel_type(ty)) continue;
            const long long cell = A.rec_cell[rc];
            const int ry = (int)(cell / pitch) - L.ypad, rx = (int)(cell % pitch) - XPAD;
            const int lane = rx - (x0 - H), j = ry - (y0 - H);
            if (tx == lane && (j & (BLOCK_Y - 1)) == wy) A.seis[(long long)tp * A.rec_total + rc] = s_ov[ty - 2][j][tx];
        }
    }

    // ---- B: velocities of step t on the region [x0-H, x0+64-H) x [y0-H, y0+TY+H)
    const float(*s_xx)[G::SW] = (const float(*)[G::SW])s_st[0];
    const float(*s_yy)[G::SW] = (const float(*)[G::SW])s_st[1];
    const float(*s_xy)[G::SW] = (const float(*)[G::SW])s_st[2];
    float *__restrict__ nvxp = Fn.f[0] + so;
    float *__restrict__ nvyp = Fn.f[1] + so;
    const long long pxs = (long long)shot * L.NY * W2, pys = (long long)shot * W2 * L.NX;
    for (int j = wy; j < G::VH; j += BLOCK_Y) {
        const int yv = y0 - H + j;
        const bool yin = yv >= 0 && yv < L.NY;
        const float *cyb = T.yb_v + yv * K;
        const float *cyf = T.yf_v + yv * K;
        const int Sr = j + H;
        float u = 0.f, t = 0.f, u2 = 0.f, t2 = 0.f;
#pragma unroll
        for (int k = 0; k < K; k++) {
            u += cxf[k] * s_xx[Sr][tx + k];
            u2 += cxb[k] * s_xy[Sr][tx + k];
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            t += cyb[k] * s_xy[Sr + k - H][tx + H];
            t2 += cyf[k] * s_yy[Sr + k - H][tx + H];
        }
        float nvx = 0.f, nvy = 0.f;
        if (xin && yin) {
            const bool owner = own_lane && j >= H && j < H + TY;
            if (xi >= 0) {
                const int pi = yv * W2 + xi;
                const float a0 = cpml_step(u, P.psx[0][pxs + pi], axh, bxh);
                const float a1 = cpml_step(u2, P.psx[1][pxs + pi], ax, bxx);
                if (owner) {
                    PN.psx[0][pxs + pi] = a0;
                    PN.psx[1][pxs + pi] = a1;
                }
            }
            if (P.on) {
                const int yi = frame_index(yv, L.NY, P.W, C.free_surface);
                if (yi >= 0) {
                    const int pi = yi * L.NX + xv;
                    const float b0 = cpml_step(t, P.psy[0][pys + pi], P.ay[yv], P.by[yv]);
                    const float b1 = cpml_step(t2, P.psy[1][pys + pi], P.ayh[yv], P.byh[yv]);
                    if (owner) {
                        PN.psy[0][pys + pi] = b0;
                        PN.psy[1][pys + pi] = b1;
                    }
                }
            }
            u += t;
            u *= s_ov[2][j][tx];
            nvx = s_ov[0][j][tx] + u;
            u2 += t2;
            u2 *= s_ov[3][j][tx];
            nvy = s_ov[1][j][tx] + u2;
            if (owner) {
                const int o = obase + j * pitch;
                nvxp[o] = nvx;
                nvyp[o] = nvy;
            }
        }
        // new velocities replace the old ones this thread alone has read
        s_ov[0][j][tx] = nvx;
        s_ov[1][j][tx] = nvy;
    }
    __syncthreads();

    // ---- C: stresses of the owned cells
    if (own_x) {
        const float(*v_x)[TILE_X] = s_ov[0];
        const float(*v_y)[TILE_X] = s_ov[1];
        float *__restrict__ sxx = Fn.f[2] + so;
        float *__restrict__ syy = Fn.f[3] + so;
        float *__restrict__ sxy = Fn.f[4] + so;
        const bool fs = C.free_surface != 0;
        for (int js = wy; js < TY; js += BLOCK_Y) {
            const int y = y0 + js;
            if (y >= L.NY) break;
            const int j = js + H;
            const float *cyb = T.yb + y * K;
            const float *cyf = T.yf + y * K;
            float vxx = 0.f, vyy = 0.f, u2 = 0.f, t2 = 0.f;
#pragma unroll
            for (int k = 0; k < K; k++) {
                vxx += cxb[k] * v_x[j][tx + k - H];
                t2 += cxf[k] * v_y[j][tx + k - H];
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                vyy += cyb[k] * v_y[j + k - H][tx];
                u2 += cyf[k] * v_x[j + k - H][tx];
            }
            if (xi >= 0) {
                const int pi = y * W2 + xi;
                P.psx[2][pxs + pi] = cpml_step(vxx, P.psx[2][pxs + pi], ax, bxx);
                P.psx[3][pxs + pi] = cpml_step(t2, P.psx[3][pxs + pi], axh, bxh);
            }
            if (P.on) {
                const int yi = frame_index(y, L.NY, P.W, C.free_surface);
                if (yi >= 0) {
                    const int pi = yi * L.NX + xv;
                    P.psy[2][pys + pi] = cpml_step(vyy, P.psy[2][pys + pi], P.ay[y], P.by[y]);
                    P.psy[3][pys + pi] = cpml_step(u2, P.psy[3][pys + pi], P.ayh[y], P.byh[y]);
                }
            }
            const int sc = (j + H) * G::SW + tx + H;
            float xx = s_st[0][sc], yy = s_st[1][sc], xy = s_st[2][sc];
            float uu = vxx + vyy;
            uu *= s_cf[0][js][tx];
            xx += uu;
            yy += uu;
            xx -= 2.f * (vyy * s_cf[1][js][tx]);
            yy -= 2.f * (vxx * s_cf[1][js][tx]);
            u2 += t2;
            u2 *= s_cf[2][js][tx];
            xy += u2;
            if (fs && y == 0) {
                xx += C.fsH[xv] * vxx;
                xx -= C.fsV[xv] * vyy;
                yy *= 0.f;
            }
            if (ABS) {
                const float d = s_cf[NCF - 1][js][tx];
                xx *= d;
                yy *= d;
                xy *= d;
            }
            const int o = obase + j * pitch;
            sxx[o] = xx;
            syy[o] = yy;
            sxy[o] = xy;
        }
    }

    // ---- stress-type sources and receivers of step t owned by this tile
    const int s0 = A.bsrc_ptr[sb], s1 = A.bsrc_ptr[sb + 1];
    const int r0 = A.brec_ptr[sb], r1 = A.brec_ptr[sb + 1];
    if (s0 == s1 && r0 == r1) return;
    const int t = A.tbase[0] + tofs;
    __syncthreads();
    if (tid == 0) {
        for (int e = s0; e < s1; e++) {
            const int s = A.bsrc[e];
            if (A.src_type[s] != 1) continue;
            const float v = A.src_sig[(long long)s * A.NT + t];
            const long long c = so + A.src_cell[s];
            Fn.f[2][c] += v;
            Fn.f[3][c] += v;
        }
    }
    __syncthreads();
    for (int e = r0 + tid; e < r1; e += NTH) {
        const int rc = A.brec[e];
        if (A.rec_type[rc] != 1) continue;
        const long long c = so + A.rec_cell[rc];
        float v = Fn.f[2][c];
        v += Fn.f[3][c];
        v *= 0.5f;
        A.seis[(long long)t * A.rec_total + rc] = v;
    }
}

template <typename Kern, typename... Args>
void go(Kern k, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0)
        hipExtLaunchKernelGGL(k, grid, block, 0, st, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, st, args...);
}

template <int H, int TY>
void launch_abs(bool abs_on, bool dma_on, int nblk, const Layout2D &L, const Fields2D &F, const Fields2D &Fn,
                const Coef2D &C, const Taps &T, const Cpml2D &P, const PsiNew2D &PN, const Acq &A, int tofs, int pend,
                int nbx, int nby, int nshots, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    dim3 grid(nblk), block(TILE_X, BLOCK_Y);
    if (abs_on)
        go(el2d_fused<H, TY, true, false>, grid, block, st, e0, e1, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby,
           nshots);
    else if (dma_on)
        go(el2d_fused<H, TY, false, true>, grid, block, st, e0, e1, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby,
           nshots);
    else
        go(el2d_fused<H, TY, false, false>, grid, block, st, e0, e1, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby,
           nshots);
}

template <int H>
void launch_ty(int ty, int nblk, const Layout2D &L, const Fields2D &F, const Fields2D &Fn, const Coef2D &C,
               const Taps &T, const Cpml2D &P, const PsiNew2D &PN, const Acq &A, int tofs, int pend, int nbx, int nby,
               int nshots, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    const bool abs_on = C.damp != nullptr;
    const bool dma_on = ty < 0;
    ty = ty < 0 ? -ty : ty;
    if (ty == 8)
        launch_abs<H, 8>(abs_on, dma_on, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, st, e0, e1);
    else if (ty == 16)
        launch_abs<H, 16>(abs_on, dma_on, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, st, e0, e1);
    else
        launch_abs<H, 24>(abs_on, dma_on, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, st, e0, e1);
}

}  // namespace

int fused2d_tile_x(int order) { return TILE_X - order; }

void launch_fused2d(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Fields2D &Fn,
                    const Coef2D &C, const Taps &T, const Cpml2D &P, const PsiNew2D &PN, const Acq &A, int tofs,
                    int pend, hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    const int txo = TILE_X - order;
    const int nbx = (L.NX + txo - 1) / txo, nby = (L.NY + std::abs(tile_y) - 1) / std::abs(tile_y);
    const int nblk = nbx * nby * nshots;
    switch (order / 2) {
    case 1: launch_ty<1>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    case 2: launch_ty<2>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    case 3: launch_ty<3>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    case 4: launch_ty<4>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    case 5: launch_ty<5>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    case 6: launch_ty<6>(tile_y, nblk, L, F, Fn, C, T, P, PN, A, tofs, pend, nbx, nby, nshots, stream, e0, e1); break;
    default: break;
    }
}

}  // namespace wv
