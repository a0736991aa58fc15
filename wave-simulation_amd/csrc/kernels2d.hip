// Fused 2D staggered-grid kernels for gfx950: acoustic, elastic P-SV and TM-mode EM.
//
// One time step of the reference (ForwardSolver2Dacoustic.cpp:91-194,
// ForwardSolver2Delastic.cpp:117-294, ForwardSolver2Dtmem.cpp:108-171) is
// executed as
//   phase 0: velocity-type update (vx, vy | hx, hy) from the stress-type fields
//   phase 1: stress-type update (p | sxx, syy, sxy | ez) from the new velocities,
//            with CPML, free surface and the stress part of ABS fused in
//   [ABS only: damping of the velocity-type fields]
//   epilogue: source injection and receiver gather for every shot.
// Each phase kernel reads every field and coefficient of its tile once from
// HBM (halo re-reads hit L2), keeps the derivative taps in registers (x) and
// scalar registers (y), and updates the CPML memory variables of frame cells
// only.  All shots of a batch are covered by one launch (blockIdx.z = shot).
#include "wave_internal.hpp"

namespace wv {

namespace {

constexpr int NTHREADS = TILE_X * BLOCK_Y;

template <int H>
struct Tile {
    static constexpr int RW = TILE_X + 2 * H;
    static constexpr int RH = TILE_Y + 2 * H;
};

// Stage the region [x0-H, x0+64+H) x [y0-H, y0+TILE_Y+H) of one field plane in LDS.
template <int H>
__device__ __forceinline__ void load_tile(float (*S)[Tile<H>::RW], const float *__restrict__ g, const Layout2D &L,
                                          int x0, int y0)
{
    constexpr int RW = Tile<H>::RW, RH = Tile<H>::RH;
    const int tid = threadIdx.y * TILE_X + threadIdx.x;
    const long long base = L.at(x0 - H, y0 - H);
    for (int i = tid; i < RW * RH; i += NTHREADS) {
        const int r = i / RW;
        const int c = i - r * RW;
        S[r][c] = g[base + (long long)r * L.pitch + c];
    }
}

// derivative along x at LDS row r, column tx (taps for offsets -H..H)
template <int H>
__device__ __forceinline__ float dx(const float (*S)[Tile<H>::RW], int r, int tx, const float *c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[r][tx + k];
    return s;
}

// derivative along y at LDS row r (center), column tx; c is wavefront-uniform
template <int H>
__device__ __forceinline__ float dy(const float (*S)[Tile<H>::RW], int r, int tx, const float *__restrict__ c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[r + k - H][tx + H];
    return s;
}

// psi = b*psi + a*d ; d += psi   (CPML.cpp:84-95)
__device__ __forceinline__ void cpml(float &d, float *__restrict__ psi, float a, float b)
{
    float p = b * (*psi);
    p = p + a * d;
    *psi = p;
    d = d + p;
}

struct Col {
    int x, xi;
    bool in;
    float ax, bx, axh, bxh;
};

__device__ __forceinline__ Col column(const Layout2D &L, const Cpml2D &P, int x)
{
    Col c;
    c.x = x;
    c.in = x < L.NX;
    c.xi = -1;
    c.ax = c.bx = c.axh = c.bxh = 0.f;
    if (P.on) {
        c.xi = P.xi[x];
        if (c.xi >= 0) {
            c.ax = P.ax[x];
            c.bx = P.bx[x];
            c.axh = P.axh[x];
            c.bxh = P.bxh[x];
        }
    }
    return c;
}

template <int H>
__device__ __forceinline__ void load_taps(float *dst, const float *__restrict__ tab, int pos)
{
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) dst[k] = tab[(long long)pos * (2 * H + 1) + k];
}

// ---------------------------------------------------------------------------
// elastic P-SV
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(NTHREADS) void el2d_velocity(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_xx[TL::RH][TL::RW];
    __shared__ float s_yy[TL::RH][TL::RW];
    __shared__ float s_xy[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_xx, F.f[2] + so, L, x0, y0);
    load_tile<H>(s_yy, F.f[3] + so, L, x0, y0);
    load_tile<H>(s_xy, F.f[4] + so, L, x0, y0);
    __syncthreads();

    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1], cxb[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);

    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyb = T.yb_v + (long long)y * (2 * H + 1);
        const float *cyf = T.yf_v + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;

        // vx += 1/rho_x * (CPML(Dxf sxx) + CPML(Dyb sxy))
        float u = dx<H>(s_xx, R, tx, cxf);
        float t = dy<H>(s_xy, R, tx, cyb);
        // vy += 1/rho_y * (CPML(Dxb sxy) + CPML(Dyf syy))
        float u2 = dx<H>(s_xy, R, tx, cxb);
        float t2 = dy<H>(s_yy, R, tx, cyf);
        if (!col.in) continue;
        if (col.xi >= 0) {
            const long long pi = ((long long)shot * L.NY + y) * W2 + col.xi;
            cpml(u, P.psx[0] + pi, col.axh, col.bxh);
            cpml(u2, P.psx[1] + pi, col.ax, col.bx);
        }
        if (yi >= 0) {
            const long long pi = ((long long)shot * W2 + yi) * L.NX + col.x;
            cpml(t, P.psy[0] + pi, P.ay[y], P.by[y]);
            cpml(t2, P.psy[1] + pi, P.ayh[y], P.byh[y]);
        }
        u += t;
        u *= C.c[2][o];
        F.f[0][so + o] = F.f[0][so + o] + u;
        u2 += t2;
        u2 *= C.c[3][o];
        F.f[1][so + o] = F.f[1][so + o] + u2;
    }
}

template <int H>
__global__ __launch_bounds__(NTHREADS) void el2d_stress(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_vx[TL::RH][TL::RW];
    __shared__ float s_vy[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_vx, F.f[0] + so, L, x0, y0);
    load_tile<H>(s_vy, F.f[1] + so, L, x0, y0);
    __syncthreads();

    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1], cxb[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const bool fs = C.free_surface != 0;
    const bool abs_on = C.damp != nullptr;

    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const float *cyf = T.yf + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;

        float vxx = dx<H>(s_vx, R, tx, cxb);
        float vyy = dy<H>(s_vy, R, tx, cyb);
        float u2 = dy<H>(s_vx, R, tx, cyf);
        float t2 = dx<H>(s_vy, R, tx, cxf);
        if (!col.in) continue;
        if (col.xi >= 0) {
            const long long pi = ((long long)shot * L.NY + y) * W2 + col.xi;
            cpml(vxx, P.psx[2] + pi, col.ax, col.bx);
            cpml(t2, P.psx[3] + pi, col.axh, col.bxh);
        }
        if (yi >= 0) {
            const long long pi = ((long long)shot * W2 + yi) * L.NX + col.x;
            cpml(vyy, P.psy[2] + pi, P.ay[y], P.by[y]);
            cpml(u2, P.psy[3] + pi, P.ayh[y], P.byh[y]);
        }
        const float M = C.c[0][o], mu = C.c[1][o];
        float sxx = F.f[2][so + o], syy = F.f[3][so + o], sxy = F.f[4][so + o];
        float u = vxx + vyy;
        u *= M;
        sxx += u;
        syy += u;
        sxx -= 2.f * (vyy * mu);
        syy -= 2.f * (vxx * mu);
        u2 += t2;
        u2 *= C.c[4][o];
        sxy += u2;
        if (fs && y == 0) {
            // FreeSurface2Delastic.cpp:14-45, FreeSurface.cpp:14-21
            sxx += C.fsH[col.x] * vxx;
            sxx -= C.fsV[col.x] * vyy;
            syy *= 0.f;
        }
        if (abs_on) {
            const float d = C.damp[o];
            sxx *= d;
            syy *= d;
            sxy *= d;
        }
        F.f[2][so + o] = sxx;
        F.f[3][so + o] = syy;
        F.f[4][so + o] = sxy;
    }
}

// ---------------------------------------------------------------------------
// acoustic
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(NTHREADS) void ac2d_velocity(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_p[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_p, F.f[2] + so, L, x0, y0);
    __syncthreads();
    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyf = T.yf_v + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;
        float u = dx<H>(s_p, R, tx, cxf);
        float t = dy<H>(s_p, R, tx, cyf);
        if (!col.in) continue;
        if (col.xi >= 0) cpml(u, P.psx[0] + ((long long)shot * L.NY + y) * W2 + col.xi, col.axh, col.bxh);
        if (yi >= 0) cpml(t, P.psy[0] + ((long long)shot * W2 + yi) * L.NX + col.x, P.ayh[y], P.byh[y]);
        u *= C.c[1][o];
        F.f[0][so + o] = F.f[0][so + o] + u;
        t *= C.c[2][o];
        F.f[1][so + o] = F.f[1][so + o] + t;
    }
}

template <int H>
__global__ __launch_bounds__(NTHREADS) void ac2d_pressure(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_vx[TL::RH][TL::RW];
    __shared__ float s_vy[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_vx, F.f[0] + so, L, x0, y0);
    load_tile<H>(s_vy, F.f[1] + so, L, x0, y0);
    __syncthreads();
    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const bool fs = C.free_surface != 0;
    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;
        float u = dx<H>(s_vx, R, tx, cxb);
        float t = dy<H>(s_vy, R, tx, cyb);
        if (!col.in) continue;
        if (col.xi >= 0) cpml(u, P.psx[1] + ((long long)shot * L.NY + y) * W2 + col.xi, col.ax, col.bx);
        if (yi >= 0) cpml(t, P.psy[1] + ((long long)shot * W2 + yi) * L.NX + col.x, P.ay[y], P.by[y]);
        u += t;
        u *= C.c[0][o];
        float p = F.f[2][so + o] + u;
        if (C.damp) p *= C.damp[o];
        if (fs && y == 0) p *= 0.f;
        F.f[2][so + o] = p;
    }
}

// ---------------------------------------------------------------------------
// TM-mode EM
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(NTHREADS) void tm2d_h(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_ez[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_ez, F.f[2] + so, L, x0, y0);
    __syncthreads();
    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxf[2 * H + 1];
    load_taps<H>(cxf, T.xf, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyf = T.yf + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;
        float u = dy<H>(s_ez, R, tx, cyf);
        float u2 = dx<H>(s_ez, R, tx, cxf);
        if (!col.in) continue;
        if (yi >= 0) cpml(u, P.psy[0] + ((long long)shot * W2 + yi) * L.NX + col.x, P.ayh[y], P.byh[y]);
        if (col.xi >= 0) cpml(u2, P.psx[0] + ((long long)shot * L.NY + y) * W2 + col.xi, col.axh, col.bxh);
        // hx -= 1/mu_yz * CPML(Dyf ez) ; hy -= 1/mu_xz * (-CPML(Dxf ez))
        u *= C.c[3][o];
        F.f[0][so + o] = F.f[0][so + o] - u;
        float w = -u2;
        w *= C.c[2][o];
        F.f[1][so + o] = F.f[1][so + o] - w;
    }
}

template <int H>
__global__ __launch_bounds__(NTHREADS) void tm2d_e(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P)
{
    using TL = Tile<H>;
    __shared__ float s_hx[TL::RH][TL::RW];
    __shared__ float s_hy[TL::RH][TL::RW];
    const int shot = blockIdx.z;
    const long long so = (long long)shot * L.plane;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y;
    load_tile<H>(s_hx, F.f[0] + so, L, x0, y0);
    load_tile<H>(s_hy, F.f[1] + so, L, x0, y0);
    __syncthreads();
    const int tx = threadIdx.x;
    const Col col = column(L, P, x0 + tx);
    float cxb[2 * H + 1];
    load_taps<H>(cxb, T.xb, col.x);
    const int W2 = 2 * P.W;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    for (int r = 0; r < TILE_Y / BLOCK_Y; r++) {
        const int ly = wy + r * BLOCK_Y;
        const int y = y0 + ly;
        if (y >= L.NY) break;
        const int yi = P.on ? P.yi[y] : -1;
        const float *cyb = T.yb + (long long)y * (2 * H + 1);
        const long long o = L.at(col.x, y);
        const int R = ly + H;
        float u = dx<H>(s_hy, R, tx, cxb);
        float t = dy<H>(s_hx, R, tx, cyb);
        if (!col.in) continue;
        if (col.xi >= 0) cpml(u, P.psx[1] + ((long long)shot * L.NY + y) * W2 + col.xi, col.ax, col.bx);
        if (yi >= 0) cpml(t, P.psy[1] + ((long long)shot * W2 + yi) * L.NX + col.x, P.ay[y], P.by[y]);
        u -= t;
        u *= C.c[1][o];
        float ez = C.c[0][o] * F.f[2][so + o] + u;
        if (C.damp) ez *= C.damp[o];
        F.f[2][so + o] = ez;
    }
}

// ---------------------------------------------------------------------------
// sources and receivers (SourceReceiverImplSeismic.cpp:34-105, FDTD2Delastic.cpp:12-49)
// ---------------------------------------------------------------------------
template <int PHYS>
__global__ __launch_bounds__(256) void epilogue2d(Layout2D L, Fields2D F, Acq A, int tofs)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.plane;
    const int t = A.tbase[0] + tofs;
    if (t < 0 || t >= A.NT) return;
    if (threadIdx.x == 0) {
        const int s0 = A.src_ptr[shot], s1 = A.src_ptr[shot + 1];
        for (int ty = 1; ty <= 4; ty++)
            for (int s = s0; s < s1; s++) {
                if (A.src_type[s] != ty) continue;
                const float v = A.src_sig[(long long)s * A.NT + t];
                const long long c = so + A.src_cell[s];
                if (PHYS == ELASTIC) {
                    if (ty == 1) {
                        F.f[2][c] += v;
                        F.f[3][c] += v;
                    } else if (ty == 2) {
                        F.f[0][c] += v;
                    } else if (ty == 3) {
                        F.f[1][c] += v;
                    }
                } else if (PHYS == ACOUSTIC) {
                    if (ty == 1) F.f[2][c] += v;
                    else if (ty == 2) F.f[0][c] += v;
                    else if (ty == 3) F.f[1][c] += v;
                } else {
                    if (ty == 1) F.f[2][c] += v;
                }
            }
    }
    __syncthreads();
    const int r0 = A.rec_ptr[shot], r1 = A.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const long long c = so + A.rec_cell[r];
        const int ty = A.rec_type[r];
        float v;
        if (PHYS == ELASTIC) {
            if (ty == 1) {
                v = F.f[2][c];
                v += F.f[3][c];
                v *= 0.5f;
            } else {
                v = ty == 2 ? F.f[0][c] : F.f[1][c];
            }
        } else if (PHYS == ACOUSTIC) {
            v = ty == 1 ? F.f[2][c] : (ty == 2 ? F.f[0][c] : F.f[1][c]);
        } else {
            v = F.f[2][c];
        }
        A.seis[(long long)t * A.rec_total + r] = v;
    }
}

__global__ void abs_velocity(long long n, float *a, float *b, const float *__restrict__ damp, long long plane)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = damp[i % plane];
        a[i] *= d;
        b[i] *= d;
    }
}

__global__ void advance_kernel(int *tbase, int k)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) tbase[0] += k;
}

template <int H>
void launch_h(int physics, int which, dim3 grid, const Layout2D &L, const Fields2D &F, const Coef2D &C,
              const Taps &T, const Cpml2D &P, hipStream_t st)
{
    dim3 block(TILE_X, BLOCK_Y);
    if (physics == ELASTIC) {
        if (which == 0)
            el2d_velocity<H><<<grid, block, 0, st>>>(L, F, C, T, P);
        else
            el2d_stress<H><<<grid, block, 0, st>>>(L, F, C, T, P);
    } else if (physics == ACOUSTIC) {
        if (which == 0)
            ac2d_velocity<H><<<grid, block, 0, st>>>(L, F, C, T, P);
        else
            ac2d_pressure<H><<<grid, block, 0, st>>>(L, F, C, T, P);
    } else {
        if (which == 0)
            tm2d_h<H><<<grid, block, 0, st>>>(L, F, C, T, P);
        else
            tm2d_e<H><<<grid, block, 0, st>>>(L, F, C, T, P);
    }
}

}  // namespace

void launch_step2d(int physics, int order, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                   const Taps &T, const Cpml2D &P, hipStream_t stream, int which)
{
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + TILE_Y - 1) / TILE_Y, nshots);
    switch (order / 2) {
    case 1: launch_h<1>(physics, which, grid, L, F, C, T, P, stream); break;
    case 2: launch_h<2>(physics, which, grid, L, F, C, T, P, stream); break;
    case 3: launch_h<3>(physics, which, grid, L, F, C, T, P, stream); break;
    case 4: launch_h<4>(physics, which, grid, L, F, C, T, P, stream); break;
    case 5: launch_h<5>(physics, which, grid, L, F, C, T, P, stream); break;
    case 6: launch_h<6>(physics, which, grid, L, F, C, T, P, stream); break;
    default: break;
    }
}

void launch_epilogue2d(int physics, const Layout2D &L, const Fields2D &F, const Acq &A, int tofs, hipStream_t stream)
{
    if (physics == ELASTIC)
        epilogue2d<ELASTIC><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else if (physics == ACOUSTIC)
        epilogue2d<ACOUSTIC><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else
        epilogue2d<TMEM><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
}

void launch_abs_velocity2d(int physics, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                           hipStream_t stream)
{
    (void)physics;
    const long long n = (long long)nshots * L.plane;
    abs_velocity<<<1024, 256, 0, stream>>>(n, F.f[0], F.f[1], C.damp, L.plane);
}

void launch_advance(int *tbase, int k, hipStream_t stream) { advance_kernel<<<1, 64, 0, stream>>>(tbase, k); }

}  // namespace wv
