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
// Strip form of the two adjoint steps.  A block covers 64 columns x S rows
// for every shot: its W wavefronts take W shots at a time (wavefront w: shots
// w, w + W, ...), so the shot-independent windows (coefficients, taps) are
// loaded once per block and the shots' gradient contributions to a cell are
// summed in LDS (FP64) and reach the gradient planes as one FP64 atomic per
// cell, plane and launch.  Within a wavefront: the
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
    if (P.on) {
        // the frame index is computed (no load for later loads to wait on); the coefficient
        // tables are 0 outside the frame and padded around the grid
        f.xi = frame_index(x, L.NX, P.W, false);
        f.a = P.ax[x];
        f.ah = P.axh[x];
        f.b = P.bx[x];
        f.bh = P.bxh[x];
    }
    return f;
}

// CPML y coefficients (a_y, a_y(half), b_y, b_y(half)) of the N window rows from y0, staged in LDS by
// the block's first 4N threads before the block barrier
template <int N>
__device__ __forceinline__ void stage_ycpml(float (*dst)[4], const Cpml2D &P, int y0, int tid)
{
    if (P.on && tid < 4 * N) {
        const int i = tid >> 2, j = tid & 3;
        const float *t = j == 0 ? P.ay : j == 1 ? P.ayh : j == 2 ? P.by : P.byh;
        dst[i][j] = t[y0 + i];
    }
}

// Sum of per-cell FP64 gradient contributions of NF planes over every shot a
// block handles (its W wavefronts, each over several shot groups): LDS
// accumulators (FP64 LDS atomics), then, after a block barrier, one global
// FP64 atomic per cell and plane, the cells' rows and planes spread over the
// wavefronts.
template <int NF, int S, int W>
struct GradSum {
    double acc[NF][S][TILE_X];
    // before any wavefront adds (a block barrier must follow)
    __device__ __forceinline__ void zero(int wy, int lane)
    {
        for (int j = wy; j < NF * S; j += W) acc[j / S][j % S][lane] = 0.0;
    }
    // the contributions of one row r of the strip
    __device__ __forceinline__ void add_row(const double (&c)[NF], int r, int lane)
    {
#pragma unroll
        for (int f = 0; f < NF; f++) atomicAdd(&acc[f][r][lane], c[f]);
    }
    // after a block barrier that follows every add
    __device__ __forceinline__ void flush(double *const (&g)[NF], const Layout2D &L, int x0, int ya, int wy,
                                          int lane) const
    {
        const int x = x0 + lane;
        if (x >= L.NX) return;
        for (int j = wy; j < NF * S; j += W) {
            const int f = j / S, r = j % S;
            if (ya + r < L.NY) atomicAdd(g[f] + L.at(x, ya + r), acc[f][r][lane]);
        }
    }
};

// Recorded forward derivative sums of the strip's cells: FP32 segment buffers
// (recomputation) or the FP16 record of the run (Rec2D layout: row pairs of
// srows x 64 strips, one scale per strip, field and shot).  The NF scales a
// wavefront needs are fetched by one vector load (lane j: sum f0 + j / S,
// record strip r0 + j % S) and broadcast per row with readlane.
template <int NF, int S, bool REC>
struct FwdSums {
    static_assert(S % 2 == 0, "the record is read in row pairs");
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
    // sums f0 .. f0+NF-1 of rows ya .. ya+S-1 (ya even) of column xc = min(x, NX-1) of the
    // shot whose plane starts at so; rows past the grid get values that are not used
    __device__ __forceinline__ void fetch(const FwdSnap &F, const Layout2D &L, int f0, int shot, int x0, int xc,
                                          int lane, int ya, long long so, float (&out)[NF][S]) const
    {
        if (!REC) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int r = 0; r < S; r++) out[f][r] = F.h[f0 + f][so + L.at(xc, min(ya + r, L.NY - 1))];
            return;
        }
        const int sp = F.srows / 2;
        const long long pw = (long long)F.nty * F.nbx * sp * TILE_X;  // words of one (step, sum, shot) plane
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const long long fs = ((long long)F.t * 5 + f0 + f) * F.nshots + shot;
            const auto rb = brsrc(F.hs + fs * pw, (unsigned)(pw * 4));
#pragma unroll
            for (int p = 0; p < S / 2; p++) {
                const int yp = min(ya + 2 * p, (L.NY - 1) & ~1);  // even row of a pair inside the grid
                const int q = yp / F.srows;
                const int wo = ((q * F.nbx + x0 / TILE_X) * sp + (yp - q * F.srows) / 2) * TILE_X * 4;
                // read once: streamed past the caches
                const unsigned u = __builtin_amdgcn_raw_buffer_load_b32(rb, lane * 4, wo, 2);
                const int y0 = min(ya + 2 * p, L.NY - 1), y1 = min(ya + 2 * p + 1, L.NY - 1);
                const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc), f * S + y0 / F.srows - r0));
                const float s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc), f * S + y1 / F.srows - r0));
                out[f][2 * p] = __half2float(__ushort_as_half((unsigned short)(u & 0xffffu))) * s0;
                out[f][2 * p + 1] = __half2float(__ushort_as_half((unsigned short)(u >> 16))) * s1;
            }
        }
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

// Window rows y0 .. y0+N-1, columns xl .. xl+RW-1 of NC coefficient planes,
// which are the same for every shot of a block: the block's T threads load it
// once and store it column-major in LDS, so a lane reads its column's window
// with wide LDS reads.
template <int NC, int N, int RW, int T>
struct CoefWin {
    static constexpr int TOT = N * RW, IT = (TOT + T - 1) / T;
    float v[NC][IT];
    __device__ __forceinline__ void load(const float *const (&c)[NC], const Layout2D &L, int xl, int y0, int tid,
                                         unsigned pb)
    {
#pragma unroll
        for (int a = 0; a < NC; a++) {
            const auto r = brsrc(c[a], pb);
#pragma unroll
            for (int k = 0; k < IT; k++) {
                const int e = min(tid + k * T, TOT - 1);
                const int i = e / RW, col = e - i * RW;
                v[a][k] = bload(r, (unsigned)(L.at(xl + col, y0 + i) * 4), 0);
            }
        }
    }
    __device__ __forceinline__ void store(float (*s)[RW][N], int tid) const
    {
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
            for (int k = 0; k < IT; k++) {
                const int e = tid + k * T;
                if (TOT % T == 0 || e < TOT) {
                    const int i = e / RW, col = e - i * RW;
                    s[a][col][i] = v[a][k];
                }
            }
    }
};

// the y taps of the strip's S rows (two tables) into LDS, by the block's first S*K threads
template <int S, int K>
__device__ __forceinline__ void stage_ytaps_block(float (*dst)[S * K], const float *t0, const float *t1, int ya,
                                                  int tid)
{
    if (tid < S * K) {
        dst[0][tid] = t0[(long long)ya * K + tid];
        dst[1][tid] = t1[(long long)ya * K + tid];
    }
}

// Loads are issued branch-free before any arithmetic (CPML memory variables at
// clamped indices, the x-halo columns through an out-of-range offset on the
// lanes that have none, which the buffer unit answers with 0 without a memory
// access), so a wavefront waits for few memory round trips.
constexpr unsigned NO_LANE = 0x7ffffff0u;

template <int H, int S, int W, bool REC, bool LOOP>
__global__ __launch_bounds__(TILE_X * W) void adj_el2d_stress_strip(Layout2D L, int nshots, AdjFields A, AdjPsi M,
                                                                   FwdSnap F, Coef2D C, TapsT TT, Cpml2D P, Grad G)
{
    constexpr int K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    __shared__ float s_g4[W][S][RW];
    __shared__ float s_g7[W][S][RW];
    __shared__ float s_ty[2][S * K];
    __shared__ __align__(16) float s_cf[3][RW][N];
    __shared__ float s_cyw[N][4];  // CPML y coefficients of the window rows
    __shared__ GradSum<5, S, W> s_sum;
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x, tid = wy * TILE_X + lane;
    const int x0 = B.bx * TILE_X, ya = B.by * S, xl = x0 - H, x = x0 + lane;
    const int W2 = 2 * P.W;
    const unsigned pb = (unsigned)(L.plane * 4);
    // bytes of one shot's CPML memory variables: x frame [NY][2W], y frame [2W][NX]
    const unsigned pxb = (unsigned)(L.NY * W2 * 4), pyb = (unsigned)(W2 * L.NX * 4);

    // shot-independent: coefficient windows (M, mu, mu_xy), taps, the CPML frame of the strip
    CoefWin<3, N, RW, TILE_X * W> cw;
    {
        const float *const cp[3] = {C.c[0], C.c[1], C.c[4]};
        cw.load(cp, L, xl, ya - H, tid, pb);
    }
    stage_ytaps_block<S, K>(s_ty, TT.yf, TT.yb, ya, tid);
    stage_ycpml<N>(s_cyw, P, ya - H, tid);
    const unsigned vc = (unsigned)(H + lane) * 4u;
    const bool inx = x < L.NX;
    const int xc = min(x, L.NX - 1);
    const unsigned vxc = (unsigned)xc * 4u;
    const int hc = lane < H ? lane : TILE_X + lane;
    const int hcl = lane < 2 * H ? hc : 0;
    const unsigned vh = lane < 2 * H ? (unsigned)hc * 4u : NO_LANE;
    const bool fr = strip_in_frame<H, S>(L, P, x0, ya);
    XFrame fx{-1, 0.f, 0.f, 0.f, 0.f}, fh = fx;
    if (fr) {
        fx = xframe(L, P, x);
        fh = xframe(L, P, xl + hc);
    }
    int yiw[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int y = ya - H + i;
        yiw[i] = fr ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
    }
    float txf[K], txb[K];
    load_taps<H>(txf, TT.xf, x);
    load_taps<H>(txb, TT.xb, x);
    cw.store(s_cf, tid);
    s_sum.zero(wy, lane);
    block_sync_lds();

    // LOOP: the block takes every shot, wavefront wy shots wy, wy + W, ... (one shot group per
    // pass); otherwise the block takes shot group blockIdx.z only (fewer registers, more atomics)
    const int gbeg = LOOP ? 0 : B.shot * W, gend = LOOP ? nshots : gbeg + 1;
#pragma unroll 1
    for (int g0 = gbeg; g0 < gend; g0 += W) {
        const int shot = g0 + wy;
        const bool act = shot < nshots;
        const int sh = act ? shot : 0;  // wavefronts past the last shot load shot 0 and store nothing
        const long long so = (long long)sh * L.plane;
        const auto bxx = brsrc(A.lsxx + so, pb), byy = brsrc(A.lsyy + so, pb), bxy = brsrc(A.lsxy + so, pb);
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
        float lxx[N], lyy[N], lxy[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int ro = aboff(L, xl, ya - H + i);
            lxx[i] = bload(bxx, vc, ro);
            lyy[i] = bload(byy, vc, ro);
            lxy[i] = bload(bxy, vc, ro);
        }
        float hxx[S], hyy[S], hxy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, ya + r);
            hxx[r] = bload(bxx, vh, ro);
            hyy[r] = bload(byy, vh, ro);
            hxy[r] = bload(bxy, vh, ro);
        }
        // adjoint memory variables (read buffers): y frame on the window rows, x frame on the strip's rows
        float my2[N], my3[N], mx2[S], mx3[S], hx2[S], hx3[S];
#pragma unroll
        for (int i = 0; i < N; i++) my2[i] = my3[i] = 0.f;
#pragma unroll
        for (int r = 0; r < S; r++) mx2[r] = mx3[r] = hx2[r] = hx3[r] = 0.f;
        if (fr) {
            // buffer accesses: per-shot base, scalar row offset, one lane offset per frame
            const auto y2 = brsrc(M.my[2] + (long long)sh * W2 * L.NX, pyb);
            const auto y3 = brsrc(M.my[3] + (long long)sh * W2 * L.NX, pyb);
            const auto x2 = brsrc(M.mx[2] + (long long)sh * L.NY * W2, pxb);
            const auto x3 = brsrc(M.mx[3] + (long long)sh * L.NY * W2, pxb);
            const unsigned vfx = (unsigned)max(fx.xi, 0) * 4u, vfh = (unsigned)max(fh.xi, 0) * 4u;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int ro = max(yiw[i], 0) * L.NX * 4;
                my2[i] = bload(y2, vxc, ro);
                my3[i] = bload(y3, vxc, ro);
            }
#pragma unroll
            for (int r = 0; r < S; r++) {
                const int ro = min(ya + r, L.NY - 1) * W2 * 4;
                mx2[r] = bload(x2, vfx, ro);
                mx3[r] = bload(x3, vfx, ro);
                hx2[r] = bload(x2, vfh, ro);
                hx3[r] = bload(x3, vfh, ro);
            }
        }
        // own cells: the forward step's recorded derivative sums h0 + h1, h2 + h3 (velocity update),
        // h4, h5, h6 + h7 (stress update)
        float rs[5][S];
        {
            FwdSums<5, S, REC> fs;
            fs.init(F, L, 0, sh, x0, ya, lane);
            fs.fetch(F, L, 0, sh, x0, xc, lane, ya, so, rs);
        }
        float cM[N], cmu[N], cmxy[N], hM[S], hmu[S], hmxy[S];
#pragma unroll
        for (int i = 0; i < N; i++) {
            cM[i] = s_cf[0][H + lane][i];
            cmu[i] = s_cf[1][H + lane][i];
            cmxy[i] = s_cf[2][H + lane][i];
        }
#pragma unroll
        for (int r = 0; r < S; r++) {
            hM[r] = s_cf[0][hcl][H + r];
            hmu[r] = s_cf[1][hcl][H + r];
            hmxy[r] = s_cf[2][hcl][H + r];
        }

        // lg5, lg6 on the window (registers); lg4, lg7 on the strip (LDS)
        float w5[N], w6[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float ls = lxx[i] + lyy[i];
            const float h5 = cM[i] * ls - 2.f * cmu[i] * lxx[i], h6 = cmxy[i] * lxy[i];
            float g5 = h5, g6 = h6;
            if (yiw[i] >= 0 && inx) {
                g5 = h5 + s_cyw[i][0] * (my2[i] + h5);
                g6 = h6 + s_cyw[i][1] * (my3[i] + h6);
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
                for (int k = 0; k < K; k++) s += s_ty[0][r * K + k] * w6[r + k];
                dvx += s;
            }
#pragma unroll
            for (int k = 0; k < K; k++) dvy += s_ty[1][r * K + k] * w5[r + k];
            {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) s += txf[k] * s_g7[wy][r][lane + k];
                dvy += s;
            }
            if (!act || y >= L.NY || !inx) continue;
            const int ro = aboff(L, xl, y);
            const float nvx = ovx[r] + dvx, nvy = ovy[r] + dvy;
            bstore(nvx, bvx, vc, ro);
            bstore(nvy, bvy, vc, ro);
            // adjoint memory variables of the own cell (write buffers)
            const float ls = lxx[i] + lyy[i];
            const float h4 = cM[i] * ls - 2.f * cmu[i] * lyy[i], h5 = cM[i] * ls - 2.f * cmu[i] * lxx[i];
            const float h6 = cmxy[i] * lxy[i];
            if (fx.xi >= 0) {
                const int ro = y * W2 * 4;
                bstore(fx.b * (mx2[r] + h4), brsrc(M.mxn[2] + (long long)shot * L.NY * W2, pxb), fx.xi * 4, ro);
                bstore(fx.bh * (mx3[r] + h6), brsrc(M.mxn[3] + (long long)shot * L.NY * W2, pxb), fx.xi * 4, ro);
            }
            if (yiw[i] >= 0) {
                const int ro = yiw[i] * L.NX * 4;
                bstore(s_cyw[i][2] * (my2[i] + h5), brsrc(M.myn[2] + (long long)shot * W2 * L.NX, pyb), vxc, ro);
                bstore(s_cyw[i][3] * (my3[i] + h6), brsrc(M.myn[3] + (long long)shot * W2 * L.NX, pyb), vxc, ro);
            }
            // gradient terms with the forward step's recorded derivative sums: the velocity
            // update's (1/rho_x, 1/rho_y) take the adjoint velocities just formed, which the
            // adjoint velocity step of the same time step reads
            const double c[5] = {(double)nvx * rs[0][r], (double)nvy * rs[1][r],
                                 (double)ls * ((double)rs[2][r] + (double)rs[3][r]),
                                 -2.0 * (double)lxx[i] * rs[3][r] - 2.0 * (double)lyy[i] * rs[2][r],
                                 (double)lxy[i] * (double)rs[4][r]};
            s_sum.add_row(c, r, lane);
        }
        wave_sync();  // the next shot group rewrites this wavefront's LDS strips
    }
    block_sync_lds();
    double *const g[5] = {G.g[0], G.g[1], G.g[2], G.g[3], G.g[4]};
    s_sum.flush(g, L, x0, ya, wy, lane);
}

template <int H, int S, int W, bool REC, bool LOOP>
__global__ __launch_bounds__(TILE_X * W) void adj_el2d_velocity_strip(Layout2D L, int nshots, AdjFields A, AdjPsi M,
                                                                     FwdSnap F, Coef2D C, TapsT TT, Cpml2D P, Grad G)
{
    (void)F;  // the gradient terms of this step are summed by the adjoint stress step
    (void)G;
    constexpr int K = 2 * H + 1, N = S + 2 * H, RW = TILE_X + 2 * H;
    __shared__ float s_g0[W][S][RW];
    __shared__ float s_g2[W][S][RW];
    __shared__ float s_ty[2][S * K];
    __shared__ __align__(16) float s_cf[2][RW][N];
    __shared__ float s_cyw[N][4];  // CPML y coefficients of the window rows
    const Blk B = xcd_block();
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x, tid = wy * TILE_X + lane;
    const int x0 = B.bx * TILE_X, ya = B.by * S, xl = x0 - H, x = x0 + lane;
    const int W2 = 2 * P.W;
    const unsigned pb = (unsigned)(L.plane * 4);
    // bytes of one shot's CPML memory variables: x frame [NY][2W], y frame [2W][NX]
    const unsigned pxb = (unsigned)(L.NY * W2 * 4), pyb = (unsigned)(W2 * L.NX * 4);

    // shot-independent: coefficient windows (1/rho_x, 1/rho_y), taps, the CPML frame of the strip
    CoefWin<2, N, RW, TILE_X * W> cw;
    {
        const float *const cp[2] = {C.c[2], C.c[3]};
        cw.load(cp, L, xl, ya - H, tid, pb);
    }
    stage_ytaps_block<S, K>(s_ty, TT.yb, TT.yf, ya, tid);
    stage_ycpml<N>(s_cyw, P, ya - H, tid);
    const unsigned vc = (unsigned)(H + lane) * 4u;
    const bool inx = x < L.NX;
    const int xc = min(x, L.NX - 1);
    const unsigned vxc = (unsigned)xc * 4u;
    const int hc = lane < H ? lane : TILE_X + lane;
    const int hcl = lane < 2 * H ? hc : 0;
    const unsigned vh = lane < 2 * H ? (unsigned)hc * 4u : NO_LANE;
    const bool fr = strip_in_frame<H, S>(L, P, x0, ya);
    XFrame fx{-1, 0.f, 0.f, 0.f, 0.f}, fh = fx;
    if (fr) {
        fx = xframe(L, P, x);
        fh = xframe(L, P, xl + hc);
    }
    int yiw[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int y = ya - H + i;
        yiw[i] = fr ? frame_index(y, L.NY, P.W, C.free_surface) : -1;
    }
    float txf[K], txb[K];
    load_taps<H>(txf, TT.xf, x);
    load_taps<H>(txb, TT.xb, x);
    cw.store(s_cf, tid);
    block_sync_lds();

    // LOOP: the block takes every shot, wavefront wy shots wy, wy + W, ... (one shot group per
    // pass); otherwise the block takes shot group blockIdx.z only (fewer registers, more atomics)
    const int gbeg = LOOP ? 0 : B.shot * W, gend = LOOP ? nshots : gbeg + 1;
#pragma unroll 1
    for (int g0 = gbeg; g0 < gend; g0 += W) {
        const int shot = g0 + wy;
        const bool act = shot < nshots;
        const int sh = act ? shot : 0;  // wavefronts past the last shot load shot 0 and store nothing
        const long long so = (long long)sh * L.plane;
        const auto bvx = brsrc(A.lvx + so, pb), bvy = brsrc(A.lvy + so, pb);
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
        float vx[N], vy[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int ro = aboff(L, xl, ya - H + i);
            vx[i] = bload(bvx, vc, ro);
            vy[i] = bload(bvy, vc, ro);
        }
        float hvx[S], hvy[S];
#pragma unroll
        for (int r = 0; r < S; r++) {
            const int ro = aboff(L, xl, ya + r);
            hvx[r] = bload(bvx, vh, ro);
            hvy[r] = bload(bvy, vh, ro);
        }
        float my0[N], my1[N], mx0[S], mx1[S], hx0[S], hx1[S];
#pragma unroll
        for (int i = 0; i < N; i++) my0[i] = my1[i] = 0.f;
#pragma unroll
        for (int r = 0; r < S; r++) mx0[r] = mx1[r] = hx0[r] = hx1[r] = 0.f;
        if (fr) {
            // buffer accesses: per-shot base, scalar row offset, one lane offset per frame
            const auto y2 = brsrc(M.my[0] + (long long)sh * W2 * L.NX, pyb);
            const auto y3 = brsrc(M.my[1] + (long long)sh * W2 * L.NX, pyb);
            const auto x2 = brsrc(M.mx[0] + (long long)sh * L.NY * W2, pxb);
            const auto x3 = brsrc(M.mx[1] + (long long)sh * L.NY * W2, pxb);
            const unsigned vfx = (unsigned)max(fx.xi, 0) * 4u, vfh = (unsigned)max(fh.xi, 0) * 4u;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int ro = max(yiw[i], 0) * L.NX * 4;
                my0[i] = bload(y2, vxc, ro);
                my1[i] = bload(y3, vxc, ro);
            }
#pragma unroll
            for (int r = 0; r < S; r++) {
                const int ro = min(ya + r, L.NY - 1) * W2 * 4;
                mx0[r] = bload(x2, vfx, ro);
                mx1[r] = bload(x3, vfx, ro);
                hx0[r] = bload(x2, vfh, ro);
                hx1[r] = bload(x3, vfh, ro);
            }
        }
        float cx[N], cy[N], hcx[S], hcy[S];
#pragma unroll
        for (int i = 0; i < N; i++) {
            cx[i] = s_cf[0][H + lane][i];
            cy[i] = s_cf[1][H + lane][i];
        }
#pragma unroll
        for (int r = 0; r < S; r++) {
            hcx[r] = s_cf[0][hcl][H + r];
            hcy[r] = s_cf[1][hcl][H + r];
        }

        // lg1, lg3 on the window (registers); lg0, lg2 on the strip (LDS)
        float w1[N], w3[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const float h0 = cx[i] * vx[i], h2 = cy[i] * vy[i];
            float g1 = h0, g3 = h2;
            if (yiw[i] >= 0 && inx) {
                g1 = h0 + s_cyw[i][0] * (my0[i] + h0);
                g3 = h2 + s_cyw[i][1] * (my1[i] + h2);
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
        wave_sync();

#pragma unroll
        for (int r = 0; r < S; r++) {
            const int y = ya + r, i = r + H;
            // lsxx += Dxf^T lg0 ; lsxy += Dyb^T lg1 + Dxb^T lg2 ; lsyy += Dyf^T lg3
            float dxx = 0.f, dxy = 0.f, dyy = 0.f;
#pragma unroll
            for (int k = 0; k < K; k++) dxx += txf[k] * s_g0[wy][r][lane + k];
#pragma unroll
            for (int k = 0; k < K; k++) dxy += s_ty[0][r * K + k] * w1[r + k];
            {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) s += txb[k] * s_g2[wy][r][lane + k];
                dxy += s;
            }
#pragma unroll
            for (int k = 0; k < K; k++) dyy += s_ty[1][r * K + k] * w3[r + k];
            if (!act || y >= L.NY || !inx) continue;
            const int ro = aboff(L, xl, y);
            bstore(oxx[r] + dxx, bxx, vc, ro);
            bstore(oxy[r] + dxy, bxy, vc, ro);
            bstore(oyy[r] + dyy, byy, vc, ro);
            const float h0 = cx[i] * vx[i], h2 = cy[i] * vy[i];
            if (fx.xi >= 0) {
                const int ro = y * W2 * 4;
                bstore(fx.bh * (mx0[r] + h0), brsrc(M.mxn[0] + (long long)shot * L.NY * W2, pxb), fx.xi * 4, ro);
                bstore(fx.b * (mx1[r] + h2), brsrc(M.mxn[1] + (long long)shot * L.NY * W2, pxb), fx.xi * 4, ro);
            }
            if (yiw[i] >= 0) {
                const int ro = yiw[i] * L.NX * 4;
                bstore(s_cyw[i][2] * (my0[i] + h0), brsrc(M.myn[0] + (long long)shot * W2 * L.NX, pyb), vxc, ro);
                bstore(s_cyw[i][3] * (my1[i] + h2), brsrc(M.myn[1] + (long long)shot * W2 * L.NX, pyb), vxc, ro);
            }
        }
        wave_sync();  // the next shot group rewrites this wavefront's LDS strips
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

// The stress kernel walks all shot groups inside each block (one gradient atomic per cell
// and launch); the velocity kernel, whose loop form needs more registers than its
// occupancy allows, takes one shot group per block (grid z).
template <int H, int S, int W, bool REC>
void adj_launch_swr(int which, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1)
{
    const int nbx = (L.NX + TILE_X - 1) / TILE_X, nby = (L.NY + S - 1) / S;
    const dim3 block(TILE_X, W);
    if (which == 0)
        go(adj_el2d_stress_strip<H, S, W, REC, true>, dim3(nbx, nby), block, st, e0, e1, L, nshots, a.A, a.M, a.S,
           a.C, a.TT, a.P, a.G);
    else
        go(adj_el2d_velocity_strip<H, S, W, REC, false>, dim3(nbx, nby, (nshots + W - 1) / W), block, st, e0, e1, L,
           nshots, a.A, a.M, a.S, a.C, a.TT, a.P, a.G);
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

// W shots at a time per block (one wavefront each; a block walks all shots of its cells in groups of W)
template <int H>
void adj_ty(int which, int ty, const Layout2D &L, int nshots, const AdjArgs &a, hipStream_t st, hipEvent_t e0,
            hipEvent_t e1)
{
    (void)ty;
    if (nshots >= 4)
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
