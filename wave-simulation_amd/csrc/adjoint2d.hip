// Adjoint 2D elastic step for gfx950: the exact transpose of one forward step
// (el2d_velocity + el2d_stress, CPML or no absorbing frame, no free surface),
// run backward in time, with the gradient of the misfit with respect to the
// five model coefficients accumulated in FP64.
//
// Forward step t (kernels2d.hip), with h = CPML(g) = g + (b psi + a g):
//   vx += irx (h0 + h1)   h0 = CPML(Dxf sxx)  h1 = CPML(Dyb sxy)
//   vy += iry (h2 + h3)   h2 = CPML(Dxb sxy)  h3 = CPML(Dyf syy)
//   sxx += M (h4 + h5) - 2 mu h5 ; syy += M (h4 + h5) - 2 mu h4
//   sxy += muxy (h6 + h7) h4 = CPML(Dxb vx) h5 = CPML(Dyb vy) h6 = CPML(Dyf vx) h7 = CPML(Dxf vy)
// Adjoint of one CPML term, given the adjoint lh of h and m of the new psi:
//   lg = lh + a (m + lh) ; m <- b (m + lh)
// Backward step t: adjoint sources of step t, then adj_stress (transpose of
// the stress update; adds D^T lg to the adjoint velocities), then
// adj_velocity (transpose of the velocity update; adds D^T lg to the adjoint
// stresses).  Each kernel evaluates lg on its tile plus halo into LDS and
// applies the transposed stencils (position-dependent tables, so the
// edge-reduced rows transpose exactly).
#include <hip/hip_fp16.h>

#include "adjoint.hpp"
#include "tile2d.hpp"

namespace wv {

namespace {

// walk the region elements a thread owns (as TileRegs)
template <int H, int TY, typename Fn>
__device__ __forceinline__ void region_each(int tid, Fn fn)
{
    using TL = Tile<H, TY>;
#pragma unroll
    for (int k = 0; k < TL::IT; k++) {
        const int i = tid + k * NTHREADS;
        if (TL::TOT % NTHREADS == 0 || i < TL::TOT) {
            const int r = i / TL::RW;
            fn(k, r, i - r * TL::RW);
        }
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void adj_el2d_stress(Layout2D L, AdjFields A, AdjPsi M, FwdSnap S, Coef2D C,
                                                            Taps T, TapsT TT, Cpml2D P, Grad G)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_g4[TL::RH][TL::RW];
    __shared__ float s_g5[TL::RH][TL::RW];
    __shared__ float s_g6[TL::RH][TL::RW];
    __shared__ float s_g7[TL::RH][TL::RW];
    __shared__ float s_vx[TL::RH][TL::RW];
    __shared__ float s_vy[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int W2 = 2 * P.W;

    TileRegs<H, TY> lxx, lyy, lxy, fvx, fvy, cM, cmu, cmxy;
    lxx.load(A.lsxx + so, L, x0, y0, tid);
    lyy.load(A.lsyy + so, L, x0, y0, tid);
    lxy.load(A.lsxy + so, L, x0, y0, tid);
    fvx.load(S.f0 + so, L, x0, y0, tid);
    fvy.load(S.f1 + so, L, x0, y0, tid);
    cM.load(C.c[0], L, x0, y0, tid);
    cmu.load(C.c[1], L, x0, y0, tid);
    cmxy.load(C.c[4], L, x0, y0, tid);
    // adjoint CPML terms on the whole region
    region_each<H, TY>(tid, [&](int k, int r, int c) {
        const int x = x0 - H + c, y = y0 - H + r;
        const float ls = lxx.v[k] + lyy.v[k];
        const float h4 = cM.v[k] * ls - 2.f * cmu.v[k] * lyy.v[k];
        const float h5 = cM.v[k] * ls - 2.f * cmu.v[k] * lxx.v[k];
        const float h6 = cmxy.v[k] * lxy.v[k];
        float g4 = h4, g5 = h5, g6 = h6, g7 = h6;
        const bool in = x >= 0 && x < L.NX && y >= 0 && y < L.NY;
        if (P.on && in) {
            const int xi = P.xi[x], yi = P.yi[y];
            if (xi >= 0) {
                const long long pi = ((long long)shot * L.NY + y) * W2 + xi;
                g4 = h4 + P.ax[x] * (M.mx[2][pi] + h4);
                g7 = h6 + P.axh[x] * (M.mx[3][pi] + h6);
            }
            if (yi >= 0) {
                const long long pi = ((long long)shot * W2 + yi) * L.NX + x;
                g5 = h5 + P.ay[y] * (M.my[2][pi] + h5);
                g6 = h6 + P.ayh[y] * (M.my[3][pi] + h6);
            }
        }
        s_g4[r][c] = g4;
        s_g5[r][c] = g5;
        s_g6[r][c] = g6;
        s_g7[r][c] = g7;
    });
    fvx.store(s_vx, tid);
    fvy.store(s_vy, tid);

    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1], cxb[2 * H + 1], txf[2 * H + 1], txb[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    load_taps<H>(txf, TT.xf, col.x);
    load_taps<H>(txb, TT.xb, col.x);
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    __syncthreads();

#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        if (!col.in) continue;
        const int Rr = ly + H;
        const long long o = L.at(col.x, y);
        const float sxx = A.lsxx[so + o], syy = A.lsyy[so + o], sxy = A.lsxy[so + o];
        const float Mo = C.c[0][o], muo = C.c[1][o], mxyo = C.c[4][o];
        const float ls = sxx + syy;
        const float h4 = Mo * ls - 2.f * muo * syy;
        const float h5 = Mo * ls - 2.f * muo * sxx;
        const float h6 = mxyo * sxy;
        // forward CPML-modified derivatives of step t (el2d_stress)
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const float *cyf = T.yf + (long long)y * (2 * H + 1);
        float f4 = dx<H>(s_vx, Rr, tx, cxb);
        float f5 = dy<H>(s_vy, Rr, tx, cyb);
        float f6 = dy<H>(s_vx, Rr, tx, cyf);
        float f7 = dx<H>(s_vy, Rr, tx, cxf);
        const int yi = P.on ? P.yi[y] : -1;
        if (col.xi >= 0) {
            const long long pi = ((long long)shot * L.NY + y) * W2 + col.xi;
            (void)cpml(f4, S.px[2][pi], col.ax, col.bx);
            (void)cpml(f7, S.px[3][pi], col.axh, col.bxh);
            M.mxn[2][pi] = col.bx * (M.mx[2][pi] + h4);
            M.mxn[3][pi] = col.bxh * (M.mx[3][pi] + h6);
        }
        if (yi >= 0) {
            const long long pi = ((long long)shot * W2 + yi) * L.NX + col.x;
            (void)cpml(f5, S.py[2][pi], P.ay[y], P.by[y]);
            (void)cpml(f6, S.py[3][pi], P.ayh[y], P.byh[y]);
            M.myn[2][pi] = P.by[y] * (M.my[2][pi] + h5);
            M.myn[3][pi] = P.byh[y] * (M.my[3][pi] + h6);
        }
        atomicAdd(&G.g[2][o], (double)ls * ((double)f4 + (double)f5));
        atomicAdd(&G.g[3][o], -2.0 * (double)sxx * f5 - 2.0 * (double)syy * f4);
        atomicAdd(&G.g[4][o], (double)sxy * ((double)f6 + (double)f7));
        // lvx += Dxb^T lg4 + Dyf^T lg6 ; lvy += Dyb^T lg5 + Dxf^T lg7
        const float *tyf = TT.yf + (long long)y * (2 * H + 1);
        const float *tyb = TT.yb + (long long)y * (2 * H + 1);
        float dvx = dx<H>(s_g4, Rr, tx, txb);
        dvx += dy<H>(s_g6, Rr, tx, tyf);
        float dvy = dy<H>(s_g5, Rr, tx, tyb);
        dvy += dx<H>(s_g7, Rr, tx, txf);
        A.lvx[so + o] += dvx;
        A.lvy[so + o] += dvy;
    }
}

template <int H, int TY>
__global__ __launch_bounds__(NTHREADS) void adj_el2d_velocity(Layout2D L, AdjFields A, AdjPsi M, FwdSnap S, Coef2D C,
                                                              Taps T, TapsT TT, Cpml2D P, Grad G)
{
    using TL = Tile<H, TY>;
    constexpr int R = TY / BLOCK_Y;
    __shared__ float s_g0[TL::RH][TL::RW];
    __shared__ float s_g1[TL::RH][TL::RW];
    __shared__ float s_g2[TL::RH][TL::RW];
    __shared__ float s_g3[TL::RH][TL::RW];
    __shared__ float s_xx[TL::RH][TL::RW];
    __shared__ float s_yy[TL::RH][TL::RW];
    __shared__ float s_xy[TL::RH][TL::RW];
    const Blk B = xcd_block();
    const int shot = B.shot;
    const long long so = (long long)shot * L.plane;
    const int x0 = B.bx * TILE_X, y0 = B.by * TY;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int W2 = 2 * P.W;

    TileRegs<H, TY> lvx, lvy, crx, cry, fxx, fyy, fxy;
    lvx.load(A.lvx + so, L, x0, y0, tid);
    lvy.load(A.lvy + so, L, x0, y0, tid);
    crx.load(C.c[2], L, x0, y0, tid);
    cry.load(C.c[3], L, x0, y0, tid);
    fxx.load(S.f0 + so, L, x0, y0, tid);
    fyy.load(S.f1 + so, L, x0, y0, tid);
    fxy.load(S.f2 + so, L, x0, y0, tid);
    region_each<H, TY>(tid, [&](int k, int r, int c) {
        const int x = x0 - H + c, y = y0 - H + r;
        const float h0 = crx.v[k] * lvx.v[k];
        const float h2 = cry.v[k] * lvy.v[k];
        float g0 = h0, g1 = h0, g2 = h2, g3 = h2;
        const bool in = x >= 0 && x < L.NX && y >= 0 && y < L.NY;
        if (P.on && in) {
            const int xi = P.xi[x], yi = P.yi[y];
            if (xi >= 0) {
                const long long pi = ((long long)shot * L.NY + y) * W2 + xi;
                g0 = h0 + P.axh[x] * (M.mx[0][pi] + h0);
                g2 = h2 + P.ax[x] * (M.mx[1][pi] + h2);
            }
            if (yi >= 0) {
                const long long pi = ((long long)shot * W2 + yi) * L.NX + x;
                g1 = h0 + P.ay[y] * (M.my[0][pi] + h0);
                g3 = h2 + P.ayh[y] * (M.my[1][pi] + h2);
            }
        }
        s_g0[r][c] = g0;
        s_g1[r][c] = g1;
        s_g2[r][c] = g2;
        s_g3[r][c] = g3;
    });
    fxx.store(s_xx, tid);
    fyy.store(s_yy, tid);
    fxy.store(s_xy, tid);

    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1], cxb[2 * H + 1], txf[2 * H + 1], txb[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    load_taps<H>(txf, TT.xf, col.x);
    load_taps<H>(txb, TT.xb, col.x);
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    __syncthreads();

#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        if (!col.in) continue;
        const int Rr = ly + H;
        const long long o = L.at(col.x, y);
        const float vx = A.lvx[so + o], vy = A.lvy[so + o];
        const float h0 = C.c[2][o] * vx;
        const float h2 = C.c[3][o] * vy;
        // forward CPML-modified derivatives of step t (el2d_velocity)
        const float *cyb = T.yb_v + (long long)y * (2 * H + 1);
        const float *cyf = T.yf_v + (long long)y * (2 * H + 1);
        float f0 = dx<H>(s_xx, Rr, tx, cxf);
        float f1 = dy<H>(s_xy, Rr, tx, cyb);
        float f2 = dx<H>(s_xy, Rr, tx, cxb);
        float f3 = dy<H>(s_yy, Rr, tx, cyf);
        const int yi = P.on ? P.yi[y] : -1;
        if (col.xi >= 0) {
            const long long pi = ((long long)shot * L.NY + y) * W2 + col.xi;
            (void)cpml(f0, S.px[0][pi], col.axh, col.bxh);
            (void)cpml(f2, S.px[1][pi], col.ax, col.bx);
            M.mxn[0][pi] = col.bxh * (M.mx[0][pi] + h0);
            M.mxn[1][pi] = col.bx * (M.mx[1][pi] + h2);
        }
        if (yi >= 0) {
            const long long pi = ((long long)shot * W2 + yi) * L.NX + col.x;
            (void)cpml(f1, S.py[0][pi], P.ay[y], P.by[y]);
            (void)cpml(f3, S.py[1][pi], P.ayh[y], P.byh[y]);
            M.myn[0][pi] = P.by[y] * (M.my[0][pi] + h0);
            M.myn[1][pi] = P.byh[y] * (M.my[1][pi] + h2);
        }
        f0 += f1;
        f2 += f3;
        atomicAdd(&G.g[0][o], (double)vx * f0);
        atomicAdd(&G.g[1][o], (double)vy * f2);
        // lsxx += Dxf^T lg0 ; lsxy += Dyb^T lg1 + Dxb^T lg2 ; lsyy += Dyf^T lg3
        const float *tyf = TT.yf + (long long)y * (2 * H + 1);
        const float *tyb = TT.yb + (long long)y * (2 * H + 1);
        const float dxx = dx<H>(s_g0, Rr, tx, txf);
        float dxy = dy<H>(s_g1, Rr, tx, tyb);
        dxy += dx<H>(s_g2, Rr, tx, txb);
        const float dyy = dy<H>(s_g3, Rr, tx, tyf);
        A.lsxx[so + o] += dxx;
        A.lsxy[so + o] += dxy;
        A.lsyy[so + o] += dyy;
    }
}

__global__ __launch_bounds__(256) void adj_inject(Layout2D L, AdjFields A, Acq Q, const float *__restrict__ adj, int t)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.plane;
    const int r0 = Q.rec_ptr[shot], r1 = Q.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const float v = adj[(long long)t * Q.rec_total + r];
        const long long c = so + Q.rec_cell[r];
        const int ty = Q.rec_type[r];
        if (ty == 1) {
            atomicAdd(A.lsxx + c, 0.5f * v);
            atomicAdd(A.lsyy + c, 0.5f * v);
        } else if (ty == 2) {
            atomicAdd(A.lvx + c, v);
        } else {
            atomicAdd(A.lvy + c, v);
        }
    }
}

__global__ void ckpt_maxabs(const float *__restrict__ f, long long per_shot, unsigned *maxbits)
{
    const int shot = blockIdx.y;
    const float *p = f + (long long)shot * per_shot;
    float m = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per_shot;
         i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(p[i]));
    // wavefront maximum, then one atomic per wavefront (non-negative floats order like their bits)
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits + shot, __float_as_uint(m));
}

__device__ __forceinline__ float ckpt_scale(unsigned bits)
{
    const float m = __uint_as_float(bits);
    return m > 0.f ? m * (1.f / 32768.f) : 1.f;
}

__global__ void ckpt_pack(const float *__restrict__ f, __half *h, const unsigned *__restrict__ maxbits,
                          long long per_shot, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float s = ckpt_scale(maxbits[i / per_shot]);
        h[i] = __float2half_rn(f[i] / s);
    }
}

__global__ void ckpt_unpack(const __half *__restrict__ h, float *f, const unsigned *__restrict__ maxbits,
                            long long per_shot, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float s = ckpt_scale(maxbits[i / per_shot]);
        f[i] = __half2float(h[i]) * s;
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

struct AdjArgs {
    const AdjFields &A;
    const AdjPsi &M;
    const FwdSnap &S;
    const Coef2D &C;
    const Taps &T;
    const TapsT &TT;
    const Cpml2D &P;
    const Grad &G;
};

template <int H, int TY>
void adj_launch(int which, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
                hipEvent_t e1)
{
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + TY - 1) / TY, nshots), block(TILE_X, BLOCK_Y);
    if (which == 0)
        go(adj_el2d_stress<H, TY>, grid, block, st, e0, e1, L, a.A, a.M, a.S, a.C, a.T, a.TT, a.P, a.G);
    else
        go(adj_el2d_velocity<H, TY>, grid, block, st, e0, e1, L, a.A, a.M, a.S, a.C, a.T, a.TT, a.P, a.G);
}

template <int H>
void adj_ty(int which, int ty, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
            hipEvent_t e1)
{
    if (ty == 16)
        adj_launch<H, 16>(which, L, nshots, a, st, e0, e1);
    else
        adj_launch<H, 8>(which, L, nshots, a, st, e0, e1);
}

void adj_dispatch(int which, int order, int ty, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st,
                  hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: adj_ty<1>(which, ty, L, nshots, a, st, e0, e1); break;
    case 2: adj_ty<2>(which, ty, L, nshots, a, st, e0, e1); break;
    case 3: adj_ty<3>(which, ty, L, nshots, a, st, e0, e1); break;
    case 4: adj_ty<4>(which, ty, L, nshots, a, st, e0, e1); break;
    case 5: adj_ty<5>(which, ty, L, nshots, a, st, e0, e1); break;
    case 6: adj_ty<6>(which, ty, L, nshots, a, st, e0, e1); break;
    default: break;
    }
}

}  // namespace

void launch_adj_stress(int order, int tile_y, const Layout2D &L, int nshots, const AdjFields &A, const AdjPsi &M,
                       const FwdSnap &S, const Coef2D &C, const Taps &T, const TapsT &TT, const Cpml2D &P,
                       const Grad &G, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    adj_dispatch(0, order, tile_y, L, nshots, AdjArgs{A, M, S, C, T, TT, P, G}, st, e0, e1);
}

void launch_adj_velocity(int order, int tile_y, const Layout2D &L, int nshots, const AdjFields &A, const AdjPsi &M,
                         const FwdSnap &S, const Coef2D &C, const Taps &T, const TapsT &TT, const Cpml2D &P,
                         const Grad &G, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    adj_dispatch(1, order, tile_y, L, nshots, AdjArgs{A, M, S, C, T, TT, P, G}, st, e0, e1);
}

void launch_adj_inject(const Layout2D &L, const AdjFields &A, const Acq &Q, const float *adj, int t, hipStream_t st)
{
    adj_inject<<<Q.nshots, 256, 0, st>>>(L, A, Q, adj, t);
}

void launch_ckpt_pack(const float *f, unsigned short *h, unsigned *maxbits, long long per_shot, int nshots,
                      hipStream_t st)
{
    (void)hipMemsetAsync(maxbits, 0, sizeof(unsigned) * nshots, st);
    ckpt_maxabs<<<dim3(256, nshots), 256, 0, st>>>(f, per_shot, maxbits);
    ckpt_pack<<<2048, 256, 0, st>>>(f, (__half *)h, maxbits, per_shot, per_shot * nshots);
}

void launch_ckpt_unpack(const unsigned short *h, float *f, const unsigned *maxbits, long long per_shot, int nshots,
                        hipStream_t st)
{
    ckpt_unpack<<<2048, 256, 0, st>>>((const __half *)h, f, maxbits, per_shot, per_shot * nshots);
}

}  // namespace wv
