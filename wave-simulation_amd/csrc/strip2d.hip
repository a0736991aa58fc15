// Row-strip 2D elastic P-SV and TM-EM kernels for gfx950.
//
// The same two-launch time step as the tiled kernels (velocity phase, then
// stress phase: ForwardSolver2Delastic.cpp:117-294) with another split of the
// work.  A block covers 64 columns x 4S rows; each wavefront owns S
// consecutive rows of those 64 columns, needs nothing from the other
// wavefronts of its block, and works through its rows in groups of U:
//   * a lane holds, in registers, its column of every field differentiated
//     along y over the rows of the current group plus an H-row window on
//     each side, so the y derivatives are register FMAs with wavefront-uniform
//     taps (half weights on interior rows, the position tables elsewhere);
//   * the fields differentiated along x go to a wavefront-private LDS strip
//     (U rows x (64 + 2H) columns) and are read back at the 2H+1 offsets with
//     the lane's own column taps; the wavefront orders its own LDS writes and
//     reads, so no workgroup barrier sits between loads and arithmetic;
//   * with several groups (U < S) the loads of group g+1 are issued before
//     group g is computed; the launcher uses one group per wavefront, which
//     measured faster on gfx950 (S = 2, 4, 8 for 8-, 16-, 32-row blocks);
//   * every global access to a field plane is a raw buffer access: one
//     128-bit resource per array in scalar registers, the lane's byte offset
//     and the row's byte offset in a scalar register, so loads and stores of
//     the planes carry no vector address arithmetic.
// Sources and receivers use the per-(shot, block) lists of the tiled kernels
// with 4S-row blocks.
#include <hip/hip_fp16.h>

#include "acq2d.hpp"

namespace wv {
namespace {

// byte offset of (x, y) in a padded plane (planes are below 2 GiB: checked by the launcher)
__device__ __forceinline__ int boff(const Layout2D &L, int x, int y) { return (int)(L.at(x, y) * 4); }

// general form: table taps for offsets -H..H
template <int K>
__device__ __forceinline__ float dtab(const float *v, const float *c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < K; k++) s += c[k] * v[k];
    return s;
}

// Per-array buffer resources and lane offsets of one wavefront's strip.
struct StripCtx {
    int shot, lane, x0, ya, xl, W2;
    unsigned vc, vh, vxi, vl;  // lane byte offsets: own column, x halo, x-frame column, y-frame column
    bool hl, xf;               // lane loads an x-halo column; lane is in the x frame
    long long so;
};

template <int H>
__device__ __forceinline__ StripCtx strip_ctx(const Layout2D &L, const Cpml2D &P, const Blk &B, int S, int wy,
                                              const Col &col)
{
    StripCtx c;
    c.shot = B.shot;
    c.so = (long long)B.shot * L.plane;
    c.lane = threadIdx.x;
    c.x0 = B.bx * TILE_X;
    c.ya = B.by * (S * BLOCK_Y) + wy * S;
    c.xl = c.x0 - H;
    c.W2 = 2 * P.W;
    c.vc = (unsigned)(H + c.lane) * 4u;
    c.hl = c.lane < 2 * H;
    c.vh = (unsigned)(c.lane < H ? c.lane : TILE_X + c.lane) * 4u;
    c.xf = col.xi >= 0;
    c.vxi = (unsigned)col.xi * 4u;
    c.vl = (unsigned)c.lane * 4u;
    return c;
}

// Record NF fields (f0, f0+1, ...) of the wavefront's strip (S rows x 64
// columns) for the adjoint pass (Rec2D): the maximum over the strip (a DPP
// wavefront reduction) sets each field's FP16 scale; rows are stored in pairs,
// one 32-bit word per lane and pair.
template <int S, int NF>
__device__ __forceinline__ void record_strips(const Rec2D &R, const Layout2D &L, int nshots, int t, int f0, int shot,
                                              int x0, int ya, int lane, const float (&h)[NF][S])
{
    static_assert(S % 2 == 0, "rows are recorded in pairs");
    const bool in = x0 + lane < L.NX;
    float m[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        m[f] = 0.f;
#pragma unroll
        for (int r = 0; r < S; r++)
            if (in && ya + r < L.NY) m[f] = fmaxf(m[f], fabsf(h[f][r]));
    }
#pragma unroll
    for (int f = 0; f < NF; f++) m[f] = wave_max_nonneg(m[f]);
    if (ya >= L.NY) return;  // a strip wholly below the grid (last block row) records nothing
#pragma unroll
    for (int f = 0; f < NF; f++) {
        // strips whose values are all below 1e-30 are stored as zeros (scale 1); above it the
        // scale stays a normal float, so h / scale is finite
        const float sc = m[f] > 1e-30f ? m[f] * (1.f / 32768.f) : 1.f, is = 1.f / sc;
        const long long fs = ((long long)t * 5 + f0 + f) * nshots + shot;
        const long long q = (fs * R.nty + ya / S) * R.nbx + x0 / TILE_X;
        if (lane == 0) R.scale[q] = sc;
        unsigned *w = R.h + q * (S / 2) * TILE_X + lane;
#pragma unroll
        for (int p = 0; p < S / 2; p++) {
            // cells outside the grid store zeros
            const float a = (in && ya + 2 * p < L.NY) ? h[f][2 * p] * is : 0.f;
            const float b = (in && ya + 2 * p + 1 < L.NY) ? h[f][2 * p + 1] * is : 0.f;
            // written once, read once much later: streamed past the caches
            __builtin_nontemporal_store((unsigned)__half_as_ushort(__float2half_rn(a)) |
                                            ((unsigned)__half_as_ushort(__float2half_rn(b)) << 16),
                                        w + p * TILE_X);
        }
    }
}

// y derivative at window row i (own row) with the row's table taps c (offsets -H..H)
template <int H, int N>
__device__ __forceinline__ float ydiff(const float (&w)[N], int i, const float *c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * w[i + k - H];
    return s;
}

// The y taps of this wavefront's S rows (two tables) and the rows' CPML y
// coefficients (a_y, b_y, a_y(half), b_y(half); 0 outside the frame), loaded
// into registers before the field loads and written to the wavefront's LDS
// slice after them, so waiting for them does not wait for the fields.
template <int H, int S>
struct YStage {
    static constexpr int K = 2 * H + 1, M = S * K, IT = (M + TILE_X - 1) / TILE_X;
    float a[IT], b[IT], cy[4];
    __device__ __forceinline__ void load(const float *t0, const float *t1, const Cpml2D &P, int ya, int lane)
    {
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int e = min(lane + it * TILE_X, M - 1);
            a[it] = t0[(long long)ya * K + e];
            b[it] = t1[(long long)ya * K + e];
        }
        const int y = ya + min(lane, S - 1);
        cy[0] = cy[1] = cy[2] = cy[3] = 0.f;
        if (P.on) {
            cy[0] = P.ay[y];
            cy[1] = P.by[y];
            cy[2] = P.ayh[y];
            cy[3] = P.byh[y];
        }
    }
    // taps to dst[2][M]; coefficients of row r to scy[r][o .. o+3]
    template <int NC>
    __device__ __forceinline__ void store(float (*dst)[M], float (*scy)[NC], int o, int lane) const
    {
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int e = lane + it * TILE_X;
            if (M % TILE_X == 0 || e < M) {
                dst[0][e] = a[it];
                dst[1][e] = b[it];
            }
        }
        if (lane < S)
#pragma unroll
            for (int j = 0; j < 4; j++) scy[lane][o + j] = cy[j];
    }
};

// velocity update of one row (ForwardSolver2Delastic.cpp:117-166):
// vx += 1/rho_x * (CPML(Dxf sxx) + CPML(Dyb sxy)); vy += 1/rho_y * (CPML(Dxb sxy) + CPML(Dyf syy))
template <int H, int S, int U>
__global__ __launch_bounds__(NTHREADS) void el2d_velocity_strip(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P,
                                                                Acq A, int tofs, int pend, Snap2D Sn, Rec2D Rc)
{
    constexpr int G = S / U, K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    static_assert(S % U == 0, "rows per wavefront must be a multiple of the group");
    __shared__ float s_xx[BLOCK_Y][U][RW];
    __shared__ float s_xy[BLOCK_Y][U][RW];
    __shared__ float s_ty[BLOCK_Y][2][S * K];
    __shared__ float s_cy[BLOCK_Y][S][4];  // per row: a_y, b_y, a_y(half), b_y(half) (0 outside the y frame)
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const Col col = column(L, P, B.bx * TILE_X + (int)threadIdx.x);
    const StripCtx c = strip_ctx<H>(L, P, B, S, wy, col);
    const int lane = c.lane;
    if (pend) block_acquisition<ELASTIC>(A, F, c.so, shot_block(A, B), A.tbase[0] + tofs - 1, true);

    const unsigned pb = (unsigned)(L.plane * 4);
    const auto bvx = brsrc(F.f[0] + c.so, pb), bvy = brsrc(F.f[1] + c.so, pb);
    const auto bxx = brsrc(F.f[2] + c.so, pb), byy = brsrc(F.f[3] + c.so, pb), bxy = brsrc(F.f[4] + c.so, pb);
    const auto bcx = brsrc(C.c[2], pb), bcy = brsrc(C.c[3], pb);
    float *__restrict__ px0 = P.psx[0] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ px1 = P.psx[1] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ py0 = P.psy[0] + (long long)c.shot * c.W2 * L.NX;
    float *__restrict__ py1 = P.psy[1] + (long long)c.shot * c.W2 * L.NX;
    const int xi = col.xi;
    const int hc = lane < H ? lane : TILE_X + lane;

    YStage<H, S> ys;
    ys.load(T.yb_v, T.yf_v, P, c.ya, lane);
    // column windows of the y-differentiated fields (rows ya-H .. ya+S+H-1) and per-row values
    float wxy[N], wyy[N];
    float axx[S], hxx[S], hxy[S], ovx[S], ovy[S], crx[S], cry[S], p0[S], p1[S], q0[S], q1[S];
    auto load_group = [&](int g) {
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r, ro = boff(L, c.xl, y), rw = boff(L, c.xl, y + H);
            wxy[r + 2 * H] = bload(bxy, c.vc, rw);
            wyy[r + 2 * H] = bload(byy, c.vc, rw);
            axx[r] = bload(bxx, c.vc, ro);
            hxx[r] = hxy[r] = 0.f;
            if (c.hl) {
                hxx[r] = bload(bxx, c.vh, ro);
                hxy[r] = bload(bxy, c.vh, ro);
            }
            ovx[r] = bload(bvx, c.vc, ro);
            ovy[r] = bload(bvy, c.vc, ro);
            crx[r] = bload(bcx, c.vc, ro);
            cry[r] = bload(bcy, c.vc, ro);
            p0[r] = p1[r] = q0[r] = q1[r] = 0.f;
            if (y < L.NY) {
                if (c.xf) {
                    p0[r] = px0[y * c.W2 + xi];
                    p1[r] = px1[y * c.W2 + xi];
                }
                const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
                if (yi >= 0 && col.in) {
                    q0[r] = py0[yi * L.NX + c.x0 + lane];
                    q1[r] = py1[yi * L.NX + c.x0 + lane];
                }
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 2 * H; i++) {
        const int rw = boff(L, c.xl, c.ya - H + i);
        wxy[i] = bload(bxy, c.vc, rw);
        wyy[i] = bload(byy, c.vc, rw);
    }
    load_group(0);
    float cxf[K], cxb[K];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    ys.store(s_ty[wy], s_cy[wy], 0, lane);
    float rh[2][S];
#pragma unroll
    for (int r = 0; r < S; r++) rh[0][r] = rh[1][r] = 0.f;

#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g + 1 < G) load_group(g + 1);
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i;
            s_xx[wy][i][H + lane] = axx[r];
            s_xy[wy][i][H + lane] = wxy[r + H];
            if (c.hl) {
                s_xx[wy][i][hc] = hxx[r];
                s_xy[wy][i][hc] = hxy[r];
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r;
            if (y >= L.NY) break;
            float u = dtab<K>(&s_xx[wy][i][lane], cxf);
            const float t = ydiff<H, N>(wxy, r + H, &s_ty[wy][0][r * K]);
            float u2 = dtab<K>(&s_xy[wy][i][lane], cxb);
            const float t2 = ydiff<H, N>(wyy, r + H, &s_ty[wy][1][r * K]);
            if (!col.in) continue;
            float tt = t, tt2 = t2;
            if (c.xf) {
                px0[y * c.W2 + xi] = cpml(u, p0[r], col.axh, col.bxh);
                px1[y * c.W2 + xi] = cpml(u2, p1[r], col.ax, col.bx);
            }
            const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
            if (yi >= 0) {
                const float *cy = s_cy[wy][r];
                py0[yi * L.NX + c.x0 + lane] = cpml(tt, q0[r], cy[0], cy[1]);
                py1[yi * L.NX + c.x0 + lane] = cpml(tt2, q1[r], cy[2], cy[3]);
            }
            u += tt;
            u2 += tt2;
            rh[0][r] = u;
            rh[1][r] = u2;
            const int ro = boff(L, c.xl, y);
            if (Sn.h[0]) {
                Sn.h[0][c.so + L.at(c.x0, y) + lane] = u;
                Sn.h[1][c.so + L.at(c.x0, y) + lane] = u2;
            }
            u *= crx[r];
            u2 *= cry[r];
            bstore(ovx[r] + u, bvx, c.vc, ro);
            bstore(ovy[r] + u2, bvy, c.vc, ro);
        }
        wave_sync();
    }
    if (Rc.h) {
        const int ts = A.tbase[0] + tofs;
        record_strips<S, 2>(Rc, L, A.nshots, ts, 0, c.shot, c.x0, c.ya, lane, rh);
    }
}

// stress update of one row (ForwardSolver2Delastic.cpp:204-294)
template <int H, int S, int U>
__global__ __launch_bounds__(NTHREADS) void el2d_stress_strip(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P,
                                                              Acq A, int tofs, int pend, Snap2D Sn, Rec2D Rc)
{
    (void)pend;
    constexpr int G = S / U, K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    static_assert(S % U == 0, "rows per wavefront must be a multiple of the group");
    __shared__ float s_vx[BLOCK_Y][U][RW];
    __shared__ float s_vy[BLOCK_Y][U][RW];
    __shared__ float s_ty[BLOCK_Y][2][S * K];
    __shared__ float s_cy[BLOCK_Y][S][5];  // per row: y-frame row (-1: none), a_y, b_y, a_y(half), b_y(half)
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const Col col = column(L, P, B.bx * TILE_X + (int)threadIdx.x);
    const StripCtx c = strip_ctx<H>(L, P, B, S, wy, col);
    const int lane = c.lane;

    const unsigned pb = (unsigned)(L.plane * 4);
    const auto bvx = brsrc(F.f[0] + c.so, pb), bvy = brsrc(F.f[1] + c.so, pb);
    const auto bxx = brsrc(F.f[2] + c.so, pb), byy = brsrc(F.f[3] + c.so, pb), bxy = brsrc(F.f[4] + c.so, pb);
    const auto bM = brsrc(C.c[0], pb), bmu = brsrc(C.c[1], pb), bmxy = brsrc(C.c[4], pb);
    float *__restrict__ px2 = P.psx[2] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ px3 = P.psx[3] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ py2 = P.psy[2] + (long long)c.shot * c.W2 * L.NX;
    float *__restrict__ py3 = P.psy[3] + (long long)c.shot * c.W2 * L.NX;
    const int xi = col.xi;
    const int hc = lane < H ? lane : TILE_X + lane;

    YStage<H, S> ys;
    ys.load(T.yb, T.yf, P, c.ya, lane);
    float wx[N], wyv[N];
    float hx[S], hy[S], oxx[S], oyy[S], oxy[S], cM[S], cmu[S], cmxy[S], dmp[S], p2[S], p3[S], q2[S], q3[S];
    auto load_group = [&](int g) {
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r, ro = boff(L, c.xl, y), rw = boff(L, c.xl, y + H);
            wx[r + 2 * H] = bload(bvx, c.vc, rw);
            wyv[r + 2 * H] = bload(bvy, c.vc, rw);
            hx[r] = hy[r] = 0.f;
            if (c.hl) {
                hx[r] = bload(bvx, c.vh, ro);
                hy[r] = bload(bvy, c.vh, ro);
            }
            oxx[r] = bload(bxx, c.vc, ro);
            oyy[r] = bload(byy, c.vc, ro);
            oxy[r] = bload(bxy, c.vc, ro);
            cM[r] = bload(bM, c.vc, ro);
            cmu[r] = bload(bmu, c.vc, ro);
            cmxy[r] = bload(bmxy, c.vc, ro);
            dmp[r] = 1.f;
            p2[r] = p3[r] = q2[r] = q3[r] = 0.f;
            if (y < L.NY) {
                if (C.damp) dmp[r] = C.damp[L.at(c.x0, y) + lane];
                if (c.xf) {
                    p2[r] = px2[y * c.W2 + xi];
                    p3[r] = px3[y * c.W2 + xi];
                }
                const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
                if (yi >= 0 && col.in) {
                    q2[r] = py2[yi * L.NX + c.x0 + lane];
                    q3[r] = py3[yi * L.NX + c.x0 + lane];
                }
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 2 * H; i++) {
        const int rw = boff(L, c.xl, c.ya - H + i);
        wx[i] = bload(bvx, c.vc, rw);
        wyv[i] = bload(bvy, c.vc, rw);
    }
    load_group(0);
    float cxf[K], cxb[K];
    load_taps<H>(cxf, T.xf, col.x);
    load_taps<H>(cxb, T.xb, col.x);
    ys.store(s_ty[wy], s_cy[wy], 1, lane);
    if (lane < S)
        s_cy[wy][lane][0] = (float)(P.on ? frame_index(c.ya + lane, L.NY, P.W, C.free_surface) : -1);
    const bool fs = C.free_surface != 0;
    float rh[3][S];
#pragma unroll
    for (int r = 0; r < S; r++) rh[0][r] = rh[1][r] = rh[2][r] = 0.f;

#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g + 1 < G) load_group(g + 1);
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i;
            s_vx[wy][i][H + lane] = wx[r + H];
            s_vy[wy][i][H + lane] = wyv[r + H];
            if (c.hl) {
                s_vx[wy][i][hc] = hx[r];
                s_vy[wy][i][hc] = hy[r];
            }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r;
            if (y >= L.NY) break;
            const float vxx = dtab<K>(&s_vx[wy][i][lane], cxb);
            const float vyy = ydiff<H, N>(wyv, r + H, &s_ty[wy][0][r * K]);
            float u2 = ydiff<H, N>(wx, r + H, &s_ty[wy][1][r * K]);
            const float t2 = dtab<K>(&s_vy[wy][i][lane], cxf);
            if (!col.in) continue;
            float a1 = vxx, a2 = t2, a3 = vyy;
            if (c.xf) {
                px2[y * c.W2 + xi] = cpml(a1, p2[r], col.ax, col.bx);
                px3[y * c.W2 + xi] = cpml(a2, p3[r], col.axh, col.bxh);
            }
            const float *cy = s_cy[wy][r];
            const int yi = (int)cy[0];
            if (yi >= 0) {
                py2[yi * L.NX + c.x0 + lane] = cpml(a3, q2[r], cy[1], cy[2]);
                py3[yi * L.NX + c.x0 + lane] = cpml(u2, q3[r], cy[3], cy[4]);
            }
            float xx = oxx[r], yy = oyy[r], xy = oxy[r];
            float u = a1 + a3;
            u *= cM[r];
            xx += u;
            yy += u;
            xx -= 2.f * (a3 * cmu[r]);
            yy -= 2.f * (a1 * cmu[r]);
            u2 += a2;
            rh[0][r] = a1;
            rh[1][r] = a3;
            rh[2][r] = u2;
            if (Sn.h[2]) {
                const long long hr = c.so + L.at(c.x0, y) + lane;
                Sn.h[2][hr] = a1;
                Sn.h[3][hr] = a3;
                Sn.h[4][hr] = u2;
            }
            u2 *= cmxy[r];
            xy += u2;
            if (fs && y == 0) {
                // FreeSurface2Delastic.cpp:14-45, FreeSurface.cpp:14-21
                xx += C.fsH[col.x] * a1;
                xx -= C.fsV[col.x] * a3;
                yy *= 0.f;
            }
            xx *= dmp[r];
            yy *= dmp[r];
            xy *= dmp[r];
            const int ro = boff(L, c.xl, y);
            bstore(xx, bxx, c.vc, ro);
            bstore(yy, byy, c.vc, ro);
            bstore(xy, bxy, c.vc, ro);
        }
        wave_sync();
    }
    if (Rc.h) {
        const int ts = A.tbase[0] + tofs;
        record_strips<S, 3>(Rc, L, A.nshots, ts, 2, c.shot, c.x0, c.ya, lane, rh);
    }
    block_acquisition<ELASTIC>(A, F, c.so, shot_block(A, B), A.tbase[0] + tofs, false);
}

// TM-EM H update of one row (ForwardSolver2Dtmem.cpp:108-135):
// hx -= 1/mu_yz * CPML(Dyf ez); hy -= 1/mu_xz * (-CPML(Dxf ez)).  The column window of ez serves the
// y derivative, its own rows (and the halo columns) the LDS strip of the x derivative.
template <int H, int S, int U>
__global__ __launch_bounds__(NTHREADS) void tm2d_h_strip(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                         int tofs, int pend)
{
    (void)A;
    (void)tofs;
    (void)pend;
    constexpr int G = S / U, K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    static_assert(S % U == 0, "rows per wavefront must be a multiple of the group");
    __shared__ float s_ez[BLOCK_Y][U][RW];
    __shared__ float s_ty[BLOCK_Y][2][S * K];
    __shared__ float s_cy[BLOCK_Y][S][4];  // per row: a_y, b_y, a_y(half), b_y(half) (0 outside the y frame)
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const Col col = column(L, P, B.bx * TILE_X + (int)threadIdx.x);
    const StripCtx c = strip_ctx<H>(L, P, B, S, wy, col);
    const int lane = c.lane;

    const unsigned pb = (unsigned)(L.plane * 4);
    const auto bhx = brsrc(F.f[0] + c.so, pb), bhy = brsrc(F.f[1] + c.so, pb), bez = brsrc(F.f[2] + c.so, pb);
    const auto bcx = brsrc(C.c[2], pb), bcy = brsrc(C.c[3], pb);  // 1/mu_xz (hy), 1/mu_yz (hx)
    float *__restrict__ px0 = P.psx[0] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ py0 = P.psy[0] + (long long)c.shot * c.W2 * L.NX;
    const int xi = col.xi;
    const int hc = lane < H ? lane : TILE_X + lane;

    YStage<H, S> ys;
    ys.load(T.yf, T.yb, P, c.ya, lane);
    float wez[N];
    float hez[S], ohx[S], ohy[S], cxz[S], cyz[S], p0[S], q0[S];
    auto load_group = [&](int g) {
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r, ro = boff(L, c.xl, y), rw = boff(L, c.xl, y + H);
            wez[r + 2 * H] = bload(bez, c.vc, rw);
            hez[r] = c.hl ? bload(bez, c.vh, ro) : 0.f;
            ohx[r] = bload(bhx, c.vc, ro);
            ohy[r] = bload(bhy, c.vc, ro);
            cxz[r] = bload(bcx, c.vc, ro);
            cyz[r] = bload(bcy, c.vc, ro);
            p0[r] = q0[r] = 0.f;
            if (y < L.NY) {
                if (c.xf) p0[r] = px0[y * c.W2 + xi];
                const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
                if (yi >= 0 && col.in) q0[r] = py0[yi * L.NX + c.x0 + lane];
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 2 * H; i++) wez[i] = bload(bez, c.vc, boff(L, c.xl, c.ya - H + i));
    load_group(0);
    float cxf[K];
    load_taps<H>(cxf, T.xf, col.x);
    ys.store(s_ty[wy], s_cy[wy], 0, lane);

#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g + 1 < G) load_group(g + 1);
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i;
            s_ez[wy][i][H + lane] = wez[r + H];
            if (c.hl) s_ez[wy][i][hc] = hez[r];
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r;
            if (y >= L.NY) break;
            float u = ydiff<H, N>(wez, r + H, &s_ty[wy][0][r * K]);
            float u2 = dtab<K>(&s_ez[wy][i][lane], cxf);
            if (!col.in) continue;
            if (c.xf) px0[y * c.W2 + xi] = cpml(u2, p0[r], col.axh, col.bxh);
            const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
            if (yi >= 0) {
                const float *cy = s_cy[wy][r];
                py0[yi * L.NX + c.x0 + lane] = cpml(u, q0[r], cy[2], cy[3]);
            }
            const int ro = boff(L, c.xl, y);
            u *= cyz[r];
            bstore(ohx[r] - u, bhx, c.vc, ro);
            float w = -u2;
            w *= cxz[r];
            bstore(ohy[r] - w, bhy, c.vc, ro);
        }
        wave_sync();
    }
}

// TM-EM E update of one row (ForwardSolver2Dtmem.cpp:137-171):
// ez = Ca ez + Cb (CPML(Dxb hy) - CPML(Dyb hx)), ABS damping, then the EZ sources / receivers.
template <int H, int S, int U>
__global__ __launch_bounds__(NTHREADS) void tm2d_e_strip(Layout2D L, Fields2D F, Coef2D C, Taps T, Cpml2D P, Acq A,
                                                         int tofs, int pend)
{
    (void)pend;
    constexpr int G = S / U, K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    static_assert(S % U == 0, "rows per wavefront must be a multiple of the group");
    __shared__ float s_hy[BLOCK_Y][U][RW];
    __shared__ float s_ty[BLOCK_Y][2][S * K];
    __shared__ float s_cy[BLOCK_Y][S][4];
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const Col col = column(L, P, B.bx * TILE_X + (int)threadIdx.x);
    const StripCtx c = strip_ctx<H>(L, P, B, S, wy, col);
    const int lane = c.lane;

    const unsigned pb = (unsigned)(L.plane * 4);
    const auto bhx = brsrc(F.f[0] + c.so, pb), bhy = brsrc(F.f[1] + c.so, pb), bez = brsrc(F.f[2] + c.so, pb);
    const auto bca = brsrc(C.c[0], pb), bcb = brsrc(C.c[1], pb);
    float *__restrict__ px1 = P.psx[1] + (long long)c.shot * L.NY * c.W2;
    float *__restrict__ py1 = P.psy[1] + (long long)c.shot * c.W2 * L.NX;
    const int xi = col.xi;
    const int hc = lane < H ? lane : TILE_X + lane;

    YStage<H, S> ys;
    ys.load(T.yb, T.yf, P, c.ya, lane);
    float whx[N];
    float ahy[S], hhy[S], oez[S], ca[S], cb[S], dmp[S], p1[S], q1[S];
    auto load_group = [&](int g) {
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r, ro = boff(L, c.xl, y), rw = boff(L, c.xl, y + H);
            whx[r + 2 * H] = bload(bhx, c.vc, rw);
            ahy[r] = bload(bhy, c.vc, ro);
            hhy[r] = c.hl ? bload(bhy, c.vh, ro) : 0.f;
            oez[r] = bload(bez, c.vc, ro);
            ca[r] = bload(bca, c.vc, ro);
            cb[r] = bload(bcb, c.vc, ro);
            dmp[r] = 1.f;
            p1[r] = q1[r] = 0.f;
            if (y < L.NY) {
                if (C.damp) dmp[r] = C.damp[L.at(c.x0, y) + lane];
                if (c.xf) p1[r] = px1[y * c.W2 + xi];
                const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
                if (yi >= 0 && col.in) q1[r] = py1[yi * L.NX + c.x0 + lane];
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 2 * H; i++) whx[i] = bload(bhx, c.vc, boff(L, c.xl, c.ya - H + i));
    load_group(0);
    float cxb[K];
    load_taps<H>(cxb, T.xb, col.x);
    ys.store(s_ty[wy], s_cy[wy], 0, lane);

#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g + 1 < G) load_group(g + 1);
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i;
            s_hy[wy][i][H + lane] = ahy[r];
            if (c.hl) s_hy[wy][i][hc] = hhy[r];
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < U; i++) {
            const int r = g * U + i, y = c.ya + r;
            if (y >= L.NY) break;
            float u = dtab<K>(&s_hy[wy][i][lane], cxb);
            float t = ydiff<H, N>(whx, r + H, &s_ty[wy][0][r * K]);
            if (!col.in) continue;
            if (c.xf) px1[y * c.W2 + xi] = cpml(u, p1[r], col.ax, col.bx);
            const int yi = P.on ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
            if (yi >= 0) {
                const float *cy = s_cy[wy][r];
                py1[yi * L.NX + c.x0 + lane] = cpml(t, q1[r], cy[0], cy[1]);
            }
            u -= t;
            u *= cb[r];
            float e = ca[r] * oez[r] + u;
            e *= dmp[r];
            bstore(e, bez, c.vc, boff(L, c.xl, y));
        }
        wave_sync();
    }
    block_acquisition<TMEM>(A, F, c.so, shot_block(A, B), A.tbase[0] + tofs, false);
}

template <typename Kern, typename... Args>
void go(Kern k, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0)
        hipExtLaunchKernelGGL(k, grid, block, 0, st, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, st, args...);
}

template <int H, int S, int U>
void launch_hs(int which, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C, const Taps &T,
               const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
               const Snap2D &Sn, const Rec2D &Rc)
{
    constexpr int TY = S * BLOCK_Y;
    const dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + TY - 1) / TY, nshots), block(TILE_X, BLOCK_Y);
    if (which == 0)
        go(el2d_velocity_strip<H, S, U>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, Sn, Rc);
    else
        go(el2d_stress_strip<H, S, U>, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend, Sn, Rc);
}

template <int H>
void launch_s(int tile_y, int which, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C, const Taps &T,
              const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
              const Snap2D &Sn, const Rec2D &Rc)
{
    // one group per wavefront: measured faster than marching with prefetch on gfx950
    if (tile_y == 8)
        launch_hs<H, 2, 2>(which, L, nshots, F, C, T, P, A, tofs, pend, st, e0, e1, Sn, Rc);
    else if (tile_y == 32)
        launch_hs<H, 8, 8>(which, L, nshots, F, C, T, P, A, tofs, pend, st, e0, e1, Sn, Rc);
    else
        launch_hs<H, 4, 4>(which, L, nshots, F, C, T, P, A, tofs, pend, st, e0, e1, Sn, Rc);
}

template <int H>
void launch_tm_s(int tile_y, int which, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                 const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                 hipEvent_t e1)
{
    const int S = tile_y == 8 ? 2 : (tile_y == 32 ? 8 : 4), TY = S * BLOCK_Y;
    const dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + TY - 1) / TY, nshots), block(TILE_X, BLOCK_Y);
    auto pick = [&](auto kh, auto ke) {
        if (which == 0)
            go(kh, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
        else
            go(ke, grid, block, st, e0, e1, L, F, C, T, P, A, tofs, pend);
    };
    if (S == 2)
        pick(tm2d_h_strip<H, 2, 2>, tm2d_e_strip<H, 2, 2>);
    else if (S == 8)
        pick(tm2d_h_strip<H, 8, 8>, tm2d_e_strip<H, 8, 8>);
    else
        pick(tm2d_h_strip<H, 4, 4>, tm2d_e_strip<H, 4, 4>);
}

}  // namespace

bool strip2d_fits(const Layout2D &L) { return L.plane * 4 < (1LL << 31); }

void launch_strip2d_tm(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                       const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t stream, int which,
                       hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: launch_tm_s<1>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 2: launch_tm_s<2>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 3: launch_tm_s<3>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 4: launch_tm_s<4>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 5: launch_tm_s<5>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 6: launch_tm_s<6>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    default: break;
    }
}

void launch_strip2d(int order, int tile_y, const Layout2D &L, int nshots, const Fields2D &F, const Coef2D &C,
                    const Taps &T, const Cpml2D &P, const Acq &A, int tofs, int pend, hipStream_t stream, int which,
                    hipEvent_t e0, hipEvent_t e1, const Snap2D *snap, const Rec2D *rec)
{
    const Snap2D Sn = snap ? *snap : Snap2D{};
    const Rec2D Rc = rec ? *rec : Rec2D{};
    switch (order / 2) {
    case 1: launch_s<1>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    case 2: launch_s<2>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    case 3: launch_s<3>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    case 4: launch_s<4>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    case 5: launch_s<5>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    case 6: launch_s<6>(tile_y, which, L, nshots, F, C, T, P, A, tofs, pend, stream, e0, e1, Sn, Rc); break;
    default: break;
    }
}

}  // namespace wv
