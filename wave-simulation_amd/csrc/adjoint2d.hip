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
// stresses).  Each kernel evaluates lg on its cells plus an H-wide halo and
// applies the transposed stencils (position-dependent tables, so the
// edge-reduced rows transpose exactly).  The gradient terms multiply the
// adjoint fields by the forward step's CPML-modified derivative sums, which
// the forward kernels recorded (FwdSnap), so no forward derivative is
// evaluated here.
#include <hip/hip_fp16.h>

#include "adjoint.hpp"
#include "tile2d.hpp"

namespace wv {

namespace {

// ---------------------------------------------------------------------------
// Strip form of the two adjoint steps.  A block
// covers 64 columns x S rows for W shots: wavefront w works on shot
// (group * W + w) over the block's cells and needs nothing from the other
// wavefronts until the gradient, so the W shots' contributions to a cell are
// summed in LDS (FP64) and reach the gradient planes as one FP64 atomic per
// cell and block instead of one per shot.  Within a wavefront: the
// y-transposed stencils read a register column of lg over rows y-H .. y+S+H-1
// (lg evaluated by the lane itself), the x-transposed ones a wavefront-private
// LDS strip of lg over columns x-H .. x+63+H (the halo columns evaluated by
// lanes < 2H); plane arrays are raw buffer accesses with scalar row offsets.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int aboff(const Layout2D &L, int x, int y) { return (int)(L.at(x, y) * 4); }

// x-frame CPML quantities at column x (lane-level): frame column, a, a(half), b, b(half)
struct XFrame {
    int xi;
    float a, ah, b, bh;
};
__device__ __forceinline__ XFrame xframe(const Layout2D &L, const Cpml2D &P, int x)
{
    XFrame f{-1, 0.f, 0.f, 0.f, 0.f};
    if (P.on && x >= 0 && x < L.NX) {
        f.xi = P.xi[x];
        if (f.xi >= 0) {
            f.a = P.ax[x];
            f.ah = P.axh[x];
            f.b = P.bx[x];
            f.bh = P.bxh[x];
        }
    }
    return f;
}

// Sum over the block's W wavefronts of per-cell FP64 contributions of NF
// gradient planes: each wavefront adds its values into LDS accumulators (LDS
// FP64 atomics) and counts itself in; the last wavefront to arrive adds the
// sums to the gradient planes (one global FP64 atomic per cell and plane), so
// no wavefront waits for the others.  init() runs before any wavefront adds.
template <int NF, int S, int W>
struct GradSum {
    double acc[NF][S][TILE_X];
    int arrived;
    __device__ __forceinline__ void init(int wy, int lane)
    {
        if (wy == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int r = 0; r < S; r++) acc[f][r][lane] = 0.0;
            if (lane == 0) arrived = 0;
        }
        __syncthreads();
    }
    __device__ __forceinline__ void add(const double (&c)[NF][S], bool act, double *const (&g)[NF],
                                        const Layout2D &L, int x0, int ya, int lane)
    {
        if (act && W > 1) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int r = 0; r < S; r++) atomicAdd(&acc[f][r][lane], c[f][r]);
        }
        const int x = x0 + lane;
        if (W == 1) {
            if (!act || x >= L.NX) return;
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int r = 0; r < S; r++)
                    if (ya + r < L.NY) atomicAdd(g[f] + L.at(x, ya + r), c[f][r]);
            return;
        }
        __threadfence_block();
        int last = 0;
        if (lane == 0) last = atomicAdd(&arrived, 1) == W - 1;
        last = __builtin_amdgcn_readfirstlane(last);
        if (!last) return;
        __threadfence_block();
        if (x >= L.NX) return;
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int r = 0; r < S; r++)
                if (ya + r < L.NY) atomicAdd(g[f] + L.at(x, ya + r), acc[f][r][lane]);
    }
};

// Recorded forward derivative sums of the strip's cells: FP32 segment buffers
// (recomputation) or the FP16 record of the run with one scale per record
// strip (srows rows x 64 columns), field and shot.  The NF scales a wavefront
// needs are fetched by one vector load (lane j: sum f0 + j / S, record strip
// r0 + j % S) and broadcast per row with readlane.
template <int NF, int S, bool REC>
struct FwdSums {
    float sc;
    int r0;
    __device__ __forceinline__ void init(const FwdSnap &F, const Layout2D &L, int f0, int shot, int x0, int ya,
                                         int lane)
    {
        if (!REC) return;
        r0 = ya / F.srows;
        const int rl = min(ya + S - 1, L.NY - 1) / F.srows;
        const int f = lane / S, q = lane - (lane / S) * S;
        sc = 1.f;
        if (lane < NF * S && r0 + q <= rl) {
            const long long fs = ((long long)F.t * 5 + f0 + f) * F.nshots + shot;
            sc = F.ss[(fs * F.nty + r0 + q) * F.nbx + x0 / TILE_X];
        }
    }
    // sum f0 + f at (shot, xc, yc) (yc wavefront-uniform, xc < NX), o = shot plane offset + at(xc, yc)
    __device__ __forceinline__ float get(const FwdSnap &F, const Layout2D &L, int f0, int f, int shot, int xc,
                                         int yc, long long o) const
    {
        if (!REC) return F.h[f0 + f][o];
        const long long fs = ((long long)F.t * 5 + f0 + f) * F.nshots + shot;
        const float h = __half2float(__ushort_as_half(F.hs[(fs * L.NY + yc) * L.NX + xc]));
        return h * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc), f * S + yc / F.srows - r0));
    }
};

// Whether the strip's window (rows ya-H .. ya+S+H-1, columns x0-H .. x0+63+H)
// touches a CPML frame; wavefront-uniform, so interior strips skip the
// memory-variable loads and stores entirely.
template <int H, int S>
__device__ __forceinline__ bool strip_in_frame(const Layout2D &L, const Cpml2D &P, int x0, int ya)
{
    return P.on && (x0 - H < P.W || x0 + TILE_X + H > L.NX - P.W || ya - H < P.W || ya + S + H > L.NY - P.W);
}

// Loads are issued branch-free before any arithmetic (CPML memory variables at
// clamped indices, the x-halo columns through an out-of-range offset on the
// lanes that have none, which the buffer unit answers with 0 without a memory
// access), so a wavefront waits for few memory round trips.
constexpr unsigned NO_LANE = 0x7ffffff0u;

template <int H, int S, int W, bool REC>
__global__ __launch_bounds__(TILE_X * W) void adj_el2d_stress_strip(Layout2D L, int nshots, AdjFields A, AdjPsi M,
                                                                   FwdSnap F, Coef2D C, TapsT TT, Cpml2D P, Grad G)
{
    constexpr int K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    __shared__ float s_g4[W][S][RW];
    __shared__ float s_g7[W][S][RW];
    __shared__ float s_ty[W][2][S * K];
    __shared__ GradSum<3, S, W> s_sum;
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
    const int shot = B.shot * W + wy;
    const bool act = shot < nshots;
    const int x0 = B.bx * TILE_X, ya = B.by * S, xl = x0 - H, x = x0 + lane;
    const long long so = (long long)(act ? shot : 0) * L.plane;
    const int W2 = 2 * P.W;
    s_sum.init(wy, lane);
    double c[3][S];
#pragma unroll
    for (int r = 0; r < S; r++) c[0][r] = c[1][r] = c[2][r] = 0.0;

    if (act) {
        const unsigned pb = (unsigned)(L.plane * 4);
        const auto bxx = brsrc(A.lsxx + so, pb), byy = brsrc(A.lsyy + so, pb), bxy = brsrc(A.lsxy + so, pb);
        const auto bM = brsrc(C.c[0], pb), bmu = brsrc(C.c[1], pb), bmxy = brsrc(C.c[4], pb);
        const unsigned vc = (unsigned)(H + lane) * 4u;
        const bool inx = x < L.NX;
        const int xc = min(x, L.NX - 1);
        const int hc = lane < H ? lane : TILE_X + lane;
        const unsigned vh = lane < 2 * H ? (unsigned)hc * 4u : NO_LANE;
        const bool fr = strip_in_frame<H, S>(L, P, x0, ya);
        XFrame fx{-1, 0.f, 0.f, 0.f, 0.f}, fh = fx;
        if (fr) {
            fx = xframe(L, P, x);
            fh = xframe(L, P, xl + hc);
        }
        const auto bvx = brsrc(A.lvx + so, pb), bvy = brsrc(A.lvy + so, pb);
        // own cells of the adjoint velocities (read-modify-write at the end)
        float ovx[S], ovy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, min(ya + r, L.NY - 1));
            ovx[r] = bload(bvx, vc, ro);
            ovy[r] = bload(bvy, vc, ro);
        }

        // window of the lane's column (rows ya-H .. ya+S+H-1) and the x-halo columns of the strip's rows
        float lxx[N], lyy[N], lxy[N], cM[N], cmu[N], cmxy[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int ro = aboff(L, xl, ya - H + i);
            lxx[i] = bload(bxx, vc, ro);
            lyy[i] = bload(byy, vc, ro);
            lxy[i] = bload(bxy, vc, ro);
            cM[i] = bload(bM, vc, ro);
            cmu[i] = bload(bmu, vc, ro);
            cmxy[i] = bload(bmxy, vc, ro);
        }
        float hxx[S], hyy[S], hxy[S], hM[S], hmu[S], hmxy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, ya + r);
            hxx[r] = bload(bxx, vh, ro);
            hyy[r] = bload(byy, vh, ro);
            hxy[r] = bload(bxy, vh, ro);
            hM[r] = bload(bM, vh, ro);
            hmu[r] = bload(bmu, vh, ro);
            hmxy[r] = bload(bmxy, vh, ro);
        }
        // adjoint memory variables (read buffers): y frame on the window rows, x frame on the strip's rows
        float my2[N], my3[N], mx2[S], mx3[S], hx2[S], hx3[S];
        int yiw[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int y = ya - H + i;
            yiw[i] = (fr && y >= 0 && y < L.NY) ? P.yi[y] : -1;
            my2[i] = my3[i] = 0.f;
        }
#pragma unroll
        for (int r = 0; r < S; r++) mx2[r] = mx3[r] = hx2[r] = hx3[r] = 0.f;
        if (fr) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const long long pi = ((long long)shot * W2 + max(yiw[i], 0)) * L.NX + xc;
                my2[i] = M.my[2][pi];
                my3[i] = M.my[3][pi];
            }
#pragma unroll
            for (int r = 0; r < S; r++) {
                const long long row = ((long long)shot * L.NY + min(ya + r, L.NY - 1)) * W2;
                mx2[r] = M.mx[2][row + max(fx.xi, 0)];
                mx3[r] = M.mx[3][row + max(fx.xi, 0)];
                hx2[r] = M.mx[2][row + max(fh.xi, 0)];
                hx3[r] = M.mx[3][row + max(fh.xi, 0)];
            }
        }
        // own cells: recorded forward derivative sums
        float f4[S], f5[S], f67[S];
        {
            FwdSums<3, S, REC> fs;
            fs.init(F, L, 2, shot, x0, ya, lane);
#pragma unroll
            for (int r = 0; r < S; r++) {
                const int yc = min(ya + r, L.NY - 1);
                const long long o = so + L.at(xc, yc);
                f4[r] = fs.get(F, L, 2, 0, shot, xc, yc, o);
                f5[r] = fs.get(F, L, 2, 1, shot, xc, yc, o);
                f67[r] = fs.get(F, L, 2, 2, shot, xc, yc, o);
            }
        }

        // lg5, lg6 on the window (registers); lg4, lg7 on the strip (LDS)
        float w5[N], w6[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float ls = lxx[i] + lyy[i];
            const float h5 = cM[i] * ls - 2.f * cmu[i] * lxx[i], h6 = cmxy[i] * lxy[i];
            float g5 = h5, g6 = h6;
            if (yiw[i] >= 0 && inx) {
                const int y = ya - H + i;
                g5 = h5 + P.ay[y] * (my2[i] + h5);
                g6 = h6 + P.ayh[y] * (my3[i] + h6);
            }
            w5[i] = g5;
            w6[i] = g6;
        }
#pragma unroll
        for (int r = 0; r < S; r++) {
            const bool iny = ya + r < L.NY;
            const int i = r + H;
            const float ls = lxx[i] + lyy[i];
            const float h4 = cM[i] * ls - 2.f * cmu[i] * lyy[i], h6 = cmxy[i] * lxy[i];
            const bool fxo = fx.xi >= 0 && iny;
            s_g4[wy][r][H + lane] = fxo ? h4 + fx.a * (mx2[r] + h4) : h4;
            s_g7[wy][r][H + lane] = fxo ? h6 + fx.ah * (mx3[r] + h6) : h6;
            const float hls = hxx[r] + hyy[r];
            const float hh4 = hM[r] * hls - 2.f * hmu[r] * hyy[r], hh6 = hmxy[r] * hxy[r];
            const bool fho = fh.xi >= 0 && iny;
            if (lane < 2 * H) {
                s_g4[wy][r][hc] = fho ? hh4 + fh.a * (hx2[r] + hh4) : hh4;
                s_g7[wy][r][hc] = fho ? hh6 + fh.ah * (hx3[r] + hh6) : hh6;
            }
        }
        {
            const float *__restrict__ ta = TT.yf + (long long)ya * K;
            const float *__restrict__ tb = TT.yb + (long long)ya * K;
            for (int e = lane; e < S * K; e += TILE_X) {
                s_ty[wy][0][e] = ta[e];
                s_ty[wy][1][e] = tb[e];
            }
        }
        float txf[K], txb[K];
        load_taps<H>(txf, TT.xf, x);
        load_taps<H>(txb, TT.xb, x);
        wave_sync();

#pragma unroll
        for (int r = 0; r < S; r++) {
            const int y = ya + r, i = r + H;
            // lvx += Dxb^T lg4 + Dyf^T lg6 ; lvy += Dyb^T lg5 + Dxf^T lg7
            float dvx = 0.f, dvy = 0.f;
#pragma unroll
            for (int k = 0; k < K; k++) dvx += txb[k] * s_g4[wy][r][lane + k];
            {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) s += s_ty[wy][0][r * K + k] * w6[r + k];
                dvx += s;
            }
#pragma unroll
            for (int k = 0; k < K; k++) dvy += s_ty[wy][1][r * K + k] * w5[r + k];
            {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) s += txf[k] * s_g7[wy][r][lane + k];
                dvy += s;
            }
            if (y >= L.NY || !inx) continue;
            const int ro = aboff(L, xl, y);
            bstore(ovx[r] + dvx, bvx, vc, ro);
            bstore(ovy[r] + dvy, bvy, vc, ro);
            // adjoint memory variables of the own cell (write buffers)
            const float ls = lxx[i] + lyy[i];
            const float h4 = cM[i] * ls - 2.f * cmu[i] * lyy[i], h5 = cM[i] * ls - 2.f * cmu[i] * lxx[i];
            const float h6 = cmxy[i] * lxy[i];
            if (fx.xi >= 0) {
                const long long pi = ((long long)shot * L.NY + y) * W2 + fx.xi;
                M.mxn[2][pi] = fx.b * (mx2[r] + h4);
                M.mxn[3][pi] = fx.bh * (mx3[r] + h6);
            }
            if (yiw[i] >= 0) {
                const long long pi = ((long long)shot * W2 + yiw[i]) * L.NX + x;
                M.myn[2][pi] = P.by[y] * (my2[i] + h5);
                M.myn[3][pi] = P.byh[y] * (my3[i] + h6);
            }
            // gradient terms with the forward step's recorded derivative sums
            c[0][r] = (double)ls * ((double)f4[r] + (double)f5[r]);
            c[1][r] = -2.0 * (double)lxx[i] * f5[r] - 2.0 * (double)lyy[i] * f4[r];
            c[2][r] = (double)lxy[i] * (double)f67[r];
        }
    }
    double *const g[3] = {G.g[2], G.g[3], G.g[4]};
    s_sum.add(c, act, g, L, x0, ya, lane);
}

template <int H, int S, int W, bool REC>
__global__ __launch_bounds__(TILE_X * W) void adj_el2d_velocity_strip(Layout2D L, int nshots, AdjFields A, AdjPsi M,
                                                                     FwdSnap F, Coef2D C, TapsT TT, Cpml2D P, Grad G)
{
    constexpr int K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    __shared__ float s_g0[W][S][RW];
    __shared__ float s_g2[W][S][RW];
    __shared__ float s_ty[W][2][S * K];
    __shared__ GradSum<2, S, W> s_sum;
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
    const int shot = B.shot * W + wy;
    const bool act = shot < nshots;
    const int x0 = B.bx * TILE_X, ya = B.by * S, xl = x0 - H, x = x0 + lane;
    const long long so = (long long)(act ? shot : 0) * L.plane;
    const int W2 = 2 * P.W;
    s_sum.init(wy, lane);
    double c[2][S];
#pragma unroll
    for (int r = 0; r < S; r++) c[0][r] = c[1][r] = 0.0;

    if (act) {
        const unsigned pb = (unsigned)(L.plane * 4);
        const auto bvx = brsrc(A.lvx + so, pb), bvy = brsrc(A.lvy + so, pb);
        const auto bcx = brsrc(C.c[2], pb), bcy = brsrc(C.c[3], pb);
        const unsigned vc = (unsigned)(H + lane) * 4u;
        const bool inx = x < L.NX;
        const int xc = min(x, L.NX - 1);
        const int hc = lane < H ? lane : TILE_X + lane;
        const unsigned vh = lane < 2 * H ? (unsigned)hc * 4u : NO_LANE;
        const bool fr = strip_in_frame<H, S>(L, P, x0, ya);
        XFrame fx{-1, 0.f, 0.f, 0.f, 0.f}, fh = fx;
        if (fr) {
            fx = xframe(L, P, x);
            fh = xframe(L, P, xl + hc);
        }
        const auto bxx = brsrc(A.lsxx + so, pb), byy = brsrc(A.lsyy + so, pb), bxy = brsrc(A.lsxy + so, pb);
        // own cells of the adjoint stresses (read-modify-write at the end)
        float oxx[S], oyy[S], oxy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, min(ya + r, L.NY - 1));
            oxx[r] = bload(bxx, vc, ro);
            oyy[r] = bload(byy, vc, ro);
            oxy[r] = bload(bxy, vc, ro);
        }

        float vx[N], vy[N], cx[N], cy[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int ro = aboff(L, xl, ya - H + i);
            vx[i] = bload(bvx, vc, ro);
            vy[i] = bload(bvy, vc, ro);
            cx[i] = bload(bcx, vc, ro);
            cy[i] = bload(bcy, vc, ro);
        }
        float hvx[S], hvy[S], hcx[S], hcy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, ya + r);
            hvx[r] = bload(bvx, vh, ro);
            hvy[r] = bload(bvy, vh, ro);
            hcx[r] = bload(bcx, vh, ro);
            hcy[r] = bload(bcy, vh, ro);
        }
        float my0[N], my1[N], mx0[S], mx1[S], hx0[S], hx1[S];
        int yiw[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int y = ya - H + i;
            yiw[i] = (fr && y >= 0 && y < L.NY) ? P.yi[y] : -1;
            my0[i] = my1[i] = 0.f;
        }
#pragma unroll
        for (int r = 0; r < S; r++) mx0[r] = mx1[r] = hx0[r] = hx1[r] = 0.f;
        if (fr) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const long long pi = ((long long)shot * W2 + max(yiw[i], 0)) * L.NX + xc;
                my0[i] = M.my[0][pi];
                my1[i] = M.my[1][pi];
            }
#pragma unroll
            for (int r = 0; r < S; r++) {
                const long long row = ((long long)shot * L.NY + min(ya + r, L.NY - 1)) * W2;
                mx0[r] = M.mx[0][row + max(fx.xi, 0)];
                mx1[r] = M.mx[1][row + max(fx.xi, 0)];
                hx0[r] = M.mx[0][row + max(fh.xi, 0)];
                hx1[r] = M.mx[1][row + max(fh.xi, 0)];
            }
        }
        float h01[S], h23[S];
        {
            FwdSums<2, S, REC> fs;
            fs.init(F, L, 0, shot, x0, ya, lane);
#pragma unroll
            for (int r = 0; r < S; r++) {
                const int yc = min(ya + r, L.NY - 1);
                const long long o = so + L.at(xc, yc);
                h01[r] = fs.get(F, L, 0, 0, shot, xc, yc, o);
                h23[r] = fs.get(F, L, 0, 1, shot, xc, yc, o);
            }
        }

        // lg1, lg3 on the window (registers); lg0, lg2 on the strip (LDS)
        float w1[N], w3[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float h0 = cx[i] * vx[i], h2 = cy[i] * vy[i];
            float g1 = h0, g3 = h2;
            if (yiw[i] >= 0 && inx) {
                const int y = ya - H + i;
                g1 = h0 + P.ay[y] * (my0[i] + h0);
                g3 = h2 + P.ayh[y] * (my1[i] + h2);
            }
            w1[i] = g1;
            w3[i] = g3;
        }
#pragma unroll
        for (int r = 0; r < S; r++) {
            const bool iny = ya + r < L.NY;
            const int i = r + H;
            const float h0 = cx[i] * vx[i], h2 = cy[i] * vy[i];
            const bool fxo = fx.xi >= 0 && iny;
            s_g0[wy][r][H + lane] = fxo ? h0 + fx.ah * (mx0[r] + h0) : h0;
            s_g2[wy][r][H + lane] = fxo ? h2 + fx.a * (mx1[r] + h2) : h2;
            const float hh0 = hcx[r] * hvx[r], hh2 = hcy[r] * hvy[r];
            const bool fho = fh.xi >= 0 && iny;
            if (lane < 2 * H) {
                s_g0[wy][r][hc] = fho ? hh0 + fh.ah * (hx0[r] + hh0) : hh0;
                s_g2[wy][r][hc] = fho ? hh2 + fh.a * (hx1[r] + hh2) : hh2;
            }
        }
        {
            const float *__restrict__ tb = TT.yb + (long long)ya * K;
            const float *__restrict__ ta = TT.yf + (long long)ya * K;
            for (int e = lane; e < S * K; e += TILE_X) {
                s_ty[wy][0][e] = tb[e];
                s_ty[wy][1][e] = ta[e];
            }
        }
        float txf[K], txb[K];
        load_taps<H>(txf, TT.xf, x);
        load_taps<H>(txb, TT.xb, x);
        wave_sync();

#pragma unroll
        for (int r = 0; r < S; r++) {
            const int y = ya + r, i = r + H;
            // lsxx += Dxf^T lg0 ; lsxy += Dyb^T lg1 + Dxb^T lg2 ; lsyy += Dyf^T lg3
            float dxx = 0.f, dxy = 0.f, dyy = 0.f;
#pragma unroll
            for (int k = 0; k < K; k++) dxx += txf[k] * s_g0[wy][r][lane + k];
#pragma unroll
            for (int k = 0; k < K; k++) dxy += s_ty[wy][0][r * K + k] * w1[r + k];
            {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) s += txb[k] * s_g2[wy][r][lane + k];
                dxy += s;
            }
#pragma unroll
            for (int k = 0; k < K; k++) dyy += s_ty[wy][1][r * K + k] * w3[r + k];
            if (y >= L.NY || !inx) continue;
            const int ro = aboff(L, xl, y);
            bstore(oxx[r] + dxx, bxx, vc, ro);
            bstore(oxy[r] + dxy, bxy, vc, ro);
            bstore(oyy[r] + dyy, byy, vc, ro);
            const float h0 = cx[i] * vx[i], h2 = cy[i] * vy[i];
            if (fx.xi >= 0) {
                const long long pi = ((long long)shot * L.NY + y) * W2 + fx.xi;
                M.mxn[0][pi] = fx.bh * (mx0[r] + h0);
                M.mxn[1][pi] = fx.b * (mx1[r] + h2);
            }
            if (yiw[i] >= 0) {
                const long long pi = ((long long)shot * W2 + yiw[i]) * L.NX + x;
                M.myn[0][pi] = P.by[y] * (my0[i] + h0);
                M.myn[1][pi] = P.byh[y] * (my1[i] + h2);
            }
            c[0][r] = (double)vx[i] * h01[r];
            c[1][r] = (double)vy[i] * h23[r];
        }
    }
    double *const g[2] = {G.g[0], G.g[1]};
    s_sum.add(c, act, g, L, x0, ya, lane);
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

__global__ __launch_bounds__(256) void ckpt_maxabs(const float *__restrict__ f, long long per_shot, unsigned *maxbits)
{
    __shared__ float s_m[4];
    const int shot = blockIdx.y;
    const float *p = f + (long long)shot * per_shot;
    float m = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per_shot;
         i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(p[i]));
    // wavefront maximum, block maximum, then one atomic per block (non-negative floats order like their bits)
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(maxbits + shot, __float_as_uint(fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]))));
}

__device__ __forceinline__ float ckpt_scale(unsigned bits)
{
    const float m = __uint_as_float(bits);
    return m > 1e-30f ? m * (1.f / 32768.f) : 1.f;  // normal scale: x / scale stays finite
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

template <int H, int S, int W, bool REC>
void adj_launch_swr(int which, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1)
{
    const dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NY + S - 1) / S, (nshots + W - 1) / W), block(TILE_X, W);
    if (which == 0)
        go(adj_el2d_stress_strip<H, S, W, REC>, grid, block, st, e0, e1, L, nshots, a.A, a.M, a.S, a.C, a.TT, a.P,
           a.G);
    else
        go(adj_el2d_velocity_strip<H, S, W, REC>, grid, block, st, e0, e1, L, nshots, a.A, a.M, a.S, a.C, a.TT, a.P,
           a.G);
}

template <int H, int S, int W>
void adj_launch_sw(int which, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
                   hipEvent_t e1)
{
    if (a.S.hs)
        adj_launch_swr<H, S, W, true>(which, L, nshots, a, st, e0, e1);
    else
        adj_launch_swr<H, S, W, false>(which, L, nshots, a, st, e0, e1);
}

// W shots per block: up to 8 (one wavefront each)
template <int H>
void adj_ty(int which, int ty, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
            hipEvent_t e1)
{
    (void)ty;
    if (nshots >= 8)
        adj_launch_sw<H, 4, 8>(which, L, nshots, a, st, e0, e1);
    else if (nshots >= 4)
        adj_launch_sw<H, 4, 4>(which, L, nshots, a, st, e0, e1);
    else if (nshots >= 2)
        adj_launch_sw<H, 4, 2>(which, L, nshots, a, st, e0, e1);
    else
        adj_launch_sw<H, 4, 1>(which, L, nshots, a, st, e0, e1);
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
    ckpt_maxabs<<<dim3(64, nshots), 256, 0, st>>>(f, per_shot, maxbits);
    ckpt_pack<<<2048, 256, 0, st>>>(f, (__half *)h, maxbits, per_shot, per_shot * nshots);
}

void launch_ckpt_unpack(const unsigned short *h, float *f, const unsigned *maxbits, long long per_shot, int nshots,
                        hipStream_t st)
{
    ckpt_unpack<<<2048, 256, 0, st>>>((const __half *)h, f, maxbits, per_shot, per_shot * nshots);
}

}  // namespace wv
