// Adjoint (time-reversed) 3D elastic step for gfx950: the exact discrete
// adjoint of the forward step of elastic3d.hip (ForwardSolver3Delastic.cpp:120-416),
// which the reference does not have.  One adjoint time step t, after the adjoint
// sources of t are added at the receivers (adj3d_inject):
//   1. adj3d_terms<0>: per cell and shot, the stress-phase adjoint terms
//      g = CPML^T(coefficient * lambda_stress) of the nine velocity-derivative terms of
//      the stress update, advancing their adjoint CPML memory variables;
//   2. adj3d_stencil<0>: lambda_v += sum of the transposed derivatives of those terms;
//   3. adj3d_terms<1>: per cell, looping over the shots, the gradient contributions of the
//      step (FP64, one read-modify-write per cell and coefficient) and the velocity-phase
//      terms g = CPML^T(1/rho * lambda_v) of the nine stress-derivative terms;
//   4. adj3d_stencil<1>: lambda_stress += sum of their transposed derivatives.
// Steps 1 and 3 are pointwise (each cell's memory variables are read and written by
// its own thread), steps 2 and 4 are stencils over the terms with the per-position
// transposed tap tables, staged like the forward kernels: x-z tiles of 64 columns x
// 4*RZ rows in LDS for the x and z terms, register queues of 2H+1 planes for the y
// terms, marching along depth y.
#include "adjoint3d.hpp"
#include "buffer.hpp"

namespace wv {

namespace {

constexpr int NTH = TILE_X * BLOCK_Y;

// adjoint of the CPML update psi' = b psi + a d, d' = d + psi' for one cell:
// given lh = adjoint of d', returns the adjoint of d and advances m (the adjoint of psi)
__device__ __forceinline__ float cpml_t(float lh, float &m, float a, float b)
{
    const float s = m + lh;
    m = b * s;
    return lh + a * s;
}

struct CellFrames {
    long long ox, oy, oz;  // element offsets in the x, y, z frames (-1 outside)
    float ax, bx, axh, bxh, ay, by, ayh, byh, az, bz, azh, bzh;
};

__device__ __forceinline__ CellFrames frames(const Cpml3D &P, const Layout3D &L, int shot, int x, int y, int z)
{
    CellFrames f;
    f.ox = f.oy = f.oz = -1;
    f.ax = f.bx = f.axh = f.bxh = f.ay = f.by = f.ayh = f.byh = f.az = f.bz = f.azh = f.bzh = 0.f;
    if (!P.on) return f;
    const long long W2 = 2 * P.W;
    const int xi = P.xi[x], yi = P.yi[y], zi = P.zi[z];
    if (xi >= 0) {
        f.ox = (((long long)shot * L.NY + y) * L.NZ + z) * W2 + xi;
        f.ax = P.ax[x];
        f.bx = P.bx[x];
        f.axh = P.axh[x];
        f.bxh = P.bxh[x];
    }
    if (yi >= 0) {
        f.oy = (((long long)shot * W2 + yi) * L.NZ + z) * L.NX + x;
        f.ay = P.ay[y];
        f.by = P.by[y];
        f.ayh = P.ayh[y];
        f.byh = P.byh[y];
    }
    if (zi >= 0) {
        f.oz = (((long long)shot * L.NY + y) * W2 + zi) * L.NX + x;
        f.az = P.az[z];
        f.bz = P.bz[z];
        f.azh = P.azh[z];
        f.bzh = P.bzh[z];
    }
    return f;
}

// one term: g = CPML^T(lh) with the memory variable k of the frame at offset o (-1: none)
__device__ __forceinline__ float term(float lh, float *m, long long o, float a, float b)
{
    if (o < 0) return lh;
    float v = m[o];
    const float g = cpml_t(lh, v, a, b);
    m[o] = v;
    return g;
}

// Phase 0 (stress phase): g from lambda_sxx..lambda_sxy, M, mu, mu_xy, mu_xz, mu_yz.
// Phase 1 (velocity phase): gradient of the step, then g from lambda_vx..vz and 1/rho.
// g layout: [0..2] x terms, [3..5] y terms, [6..8] z terms (term j of an axis goes to
// output j of the stencil of that phase).
template <int PHASE>
__global__ __launch_bounds__(256) void adj3d_terms(Layout3D L, int nshots, Adj3D A, Coef3D C, Cpml3D P, Rec3D R,
                                                   Grad3D G)
{
    const long long ncell = (long long)L.NX * L.NZ * L.NY;
    const long long per = PHASE == 0 ? ncell * nshots : ncell;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per;
         i += (long long)gridDim.x * blockDim.x) {
        long long t = i;
        const int x = (int)(t % L.NX);
        t /= L.NX;
        const int z = (int)(t % L.NZ);
        t /= L.NZ;
        const int y = (int)(t % L.NY);
        const int shot0 = PHASE == 0 ? (int)(t / L.NY) : 0;
        const long long oc = L.at(x, y, z);
        if (PHASE == 0) {
            const int s = shot0;
            const long long o = (long long)s * L.vol + oc;
            const CellFrames f = frames(P, L, s, x, y, z);
            const float lxx = A.l[3][o], lyy = A.l[4][o], lzz = A.l[5][o];
            const float lyz = A.l[6][o], lxz = A.l[7][o], lxy = A.l[8][o];
            const float M = C.M[oc], mu = C.mu[oc];
            const float ls = lxx + lyy + lzz;
            const float hx = M * ls - 2.f * mu * (ls - lxx);
            const float hy = M * ls - 2.f * mu * (ls - lyy);
            const float hz = M * ls - 2.f * mu * (ls - lzz);
            const float sxy = C.muxy[oc] * lxy, sxz = C.muxz[oc] * lxz, syz = C.muyz[oc] * lyz;
            // x terms: 9 (Dxb vx), 13 (Dxf vy), 15 (Dxf vz)
            A.g[0][o] = term(hx, A.m[9], f.ox, f.ax, f.bx);
            A.g[1][o] = term(sxy, A.m[13], f.ox, f.axh, f.bxh);
            A.g[2][o] = term(sxz, A.m[15], f.ox, f.axh, f.bxh);
            // y terms: 12 (Dyf vx), 10 (Dyb vy), 17 (Dyf vz)
            A.g[3][o] = term(sxy, A.m[12], f.oy, f.ayh, f.byh);
            A.g[4][o] = term(hy, A.m[10], f.oy, f.ay, f.by);
            A.g[5][o] = term(syz, A.m[17], f.oy, f.ayh, f.byh);
            // z terms: 14 (Dzf vx), 16 (Dzf vy), 11 (Dzb vz)
            A.g[6][o] = term(sxz, A.m[14], f.oz, f.azh, f.bzh);
            A.g[7][o] = term(syz, A.m[16], f.oz, f.azh, f.bzh);
            A.g[8][o] = term(hz, A.m[11], f.oz, f.az, f.bz);
        } else {
            double gr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const float irx = C.irx[oc], iry = C.iry[oc], irz = C.irz[oc];
            for (int s = 0; s < nshots; s++) {
                const long long o = (long long)s * L.vol + oc;
                const float lvx = A.l[0][o], lvy = A.l[1][o], lvz = A.l[2][o];
                const float lxx = A.l[3][o], lyy = A.l[4][o], lzz = A.l[5][o];
                const float lyz = A.l[6][o], lxz = A.l[7][o], lxy = A.l[8][o];
                const double e0 = R.r[3][o], e1 = R.r[4][o], e2 = R.r[5][o];
                gr[0] += (double)lvx * R.r[0][o];
                gr[1] += (double)lvy * R.r[1][o];
                gr[2] += (double)lvz * R.r[2][o];
                gr[3] += ((double)lxx + lyy + lzz) * (e0 + e1 + e2);
                gr[4] += -2.0 * (lxx * (e1 + e2) + lyy * (e0 + e2) + lzz * (e0 + e1));
                gr[5] += (double)lxy * R.r[6][o];
                gr[6] += (double)lxz * R.r[7][o];
                gr[7] += (double)lyz * R.r[8][o];
                const CellFrames f = frames(P, L, s, x, y, z);
                const float hx = irx * lvx, hy = iry * lvy, hz = irz * lvz;
                // x terms: 0 (Dxf sxx), 3 (Dxb sxy), 6 (Dxb sxz)
                A.g[0][o] = term(hx, A.m[0], f.ox, f.axh, f.bxh);
                A.g[1][o] = term(hy, A.m[3], f.ox, f.ax, f.bx);
                A.g[2][o] = term(hz, A.m[6], f.ox, f.ax, f.bx);
                // y terms: 1 (Dyb sxy), 4 (Dyf syy), 7 (Dyb syz)
                A.g[3][o] = term(hx, A.m[1], f.oy, f.ay, f.by);
                A.g[4][o] = term(hy, A.m[4], f.oy, f.ayh, f.byh);
                A.g[5][o] = term(hz, A.m[7], f.oy, f.ay, f.by);
                // z terms: 2 (Dzb sxz), 5 (Dzb syz), 8 (Dzf szz)
                A.g[6][o] = term(hx, A.m[2], f.oz, f.az, f.bz);
                A.g[7][o] = term(hy, A.m[5], f.oz, f.az, f.bz);
                A.g[8][o] = term(hz, A.m[8], f.oz, f.azh, f.bzh);
            }
#pragma unroll
            for (int k = 0; k < 8; k++) G.g[k][oc] += gr[k];
        }
    }
}

// Outputs and tap tables of the two stencil phases.  Phase 0 writes lambda_vx, vy, vz
// (term j of every axis into output j); phase 1 writes lambda_sxx .. lambda_sxy: x terms
// into sxx, sxy, sxz, y terms into sxy, syy, syz, z terms into sxz, syz, szz.  FWD[a][j]:
// term j of axis a is the transpose of the forward-type (Dxf, Dyf, Dzf) operator.
template <int PHASE>
struct StencilMap;
template <>
struct StencilMap<0> {
    static constexpr int NOUT = 3, OUT0 = 0;
    static constexpr int OUT[3][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}};
    static constexpr bool FWD[3][3] = {{false, true, true}, {true, false, true}, {true, true, false}};
};
template <>
struct StencilMap<1> {
    static constexpr int NOUT = 6, OUT0 = 3;
    // output k = field OUT0 + k: 0 sxx 1 syy 2 szz 3 syz 4 sxz 5 sxy
    static constexpr int OUT[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
    static constexpr bool FWD[3][3] = {{true, false, false}, {false, true, false}, {false, false, true}};
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_res(const float *f, const Layout3D &L, int y)
{
    return brsrc(f + (long long)(y + L.H) * L.slab, (unsigned)(L.slab * 4));
}

// Transposed stencil phase: out[o] += sum over the terms of (D^T g).  Block = 64 x 4
// threads over a 64-column x TZ-row x-z tile of one shot, marching over a chunk of y.
template <int H, int RZ, int PHASE>
__global__ __launch_bounds__(NTH) void adj3d_stencil(Layout3D L, int nyc, Adj3D A, TapsT3D T)
{
    using M = StencilMap<PHASE>;
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1, RW = TILE_X + 2 * H, RH = TZ + 2 * H;
    __shared__ float sx[3][TZ][RW];       // x terms with an x halo
    __shared__ float sz[3][RH][TILE_X];   // z terms with a z halo
    __shared__ float szt[2][TZ][K];       // z taps (transposed) of the tile's rows: [0] backward, [1] forward
    const int nbx = gridDim.x;
    const int bx = blockIdx.x % nbx;
    const int bz = blockIdx.y;
    const int shot = blockIdx.z / nyc, yc = blockIdx.z - shot * nyc;
    const int LY = (L.NY + nyc - 1) / nyc;
    const int ys = yc * LY, ye = min(L.NY, ys + LY);
    const int x0 = bx * TILE_X, z0 = bz * TZ;
    const int tx = threadIdx.x, wz = threadIdx.y, tid = wz * TILE_X + tx;
    const int x = x0 + tx;
    const bool xin = x < L.NX;
    const long long so = (long long)shot * L.vol;

    float cxf[K], cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        cxf[k] = T.xf[(long long)x * K + k];
        cxb[k] = T.xb[(long long)x * K + k];
    }
    for (int e = tid; e < TZ * K; e += NTH) {
        const int r = e / K, k = e - r * K;
        szt[0][r][k] = T.zb[(long long)(z0 + r) * K + k];
        szt[1][r][k] = T.zf[(long long)(z0 + r) * K + k];
    }
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);

    // y queues of the three y terms, planes y-H .. y+H
    float q[3][K][RZ];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 2 * H; k++) {
            const auto b = plane_res(A.g[3 + j] + so, L, ys - H + k);
#pragma unroll
            for (int r = 0; r < RZ; r++) q[j][k + 1][r] = bload(b, oc[r], 0);
        }

    for (int y = ys; y < ye; y++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
#pragma unroll
            for (int k = 0; k < 2 * H; k++)
#pragma unroll
                for (int r = 0; r < RZ; r++) q[j][k][r] = q[j][k + 1][r];
            const auto b = plane_res(A.g[3 + j] + so, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) q[j][2 * H][r] = bload(b, oc[r], 0);
        }
        // x terms: rows z0 .. z0+TZ, columns x0-H .. x0+64+H
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const auto b = plane_res(A.g[j] + so, L, y);
            for (int e = tid; e < TZ * RW; e += NTH) {
                const int r = e / RW, c = e - r * RW;
                sx[j][r][c] = bload(b, (unsigned)(((z0 + r + L.H) * L.pitch + x0 - H + c + XPAD) * 4), 0);
            }
        }
        // z terms: rows z0-H .. z0+TZ+H, columns x0 .. x0+64
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const auto b = plane_res(A.g[6 + j] + so, L, y);
            for (int e = tid; e < RH * TILE_X; e += NTH) {
                const int r = e / TILE_X, c = e - r * TILE_X;
                sz[j][r][c] = bload(b, (unsigned)(((z0 - H + r + L.H) * L.pitch + x0 + c + XPAD) * 4), 0);
            }
        }
        // outputs of this plane (read before the barrier, so the loads overlap it)
        float acc[M::NOUT][RZ];
#pragma unroll
        for (int o = 0; o < M::NOUT; o++) {
            const auto b = plane_res(A.l[M::OUT0 + o] + so, L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) acc[o][r] = bload(b, oc[r], 0);
        }
        // y taps (transposed) of this plane, the same for every lane
        float cyf[K], cyb[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            cyf[k] = T.yf[(long long)y * K + k];
            cyb[k] = T.yb[(long long)y * K + k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float sxv = 0.f, syv = 0.f, szv = 0.f;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    sxv += (M::FWD[0][j] ? cxf[k] : cxb[k]) * sx[j][lz][tx + k];
                    syv += (M::FWD[1][j] ? cyf[k] : cyb[k]) * q[j][k][r];
                    szv += szt[M::FWD[2][j] ? 1 : 0][lz][k] * sz[j][lz + k][tx];
                }
                acc[M::OUT[0][j]][r] += sxv;
                acc[M::OUT[1][j]][r] += syv;
                acc[M::OUT[2][j]][r] += szv;
            }
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = z0 + wz + r * BLOCK_Y;
            if (!xin || z >= L.NZ) continue;
#pragma unroll
            for (int o = 0; o < M::NOUT; o++) bstore(acc[o][r], plane_res(A.l[M::OUT0 + o] + so, L, y), oc[r], 0);
        }
        __syncthreads();  // the next plane rewrites the LDS tiles
    }
}

// adjoint sources of step t (time-major [NT][rec_total]) at the receivers: P receivers
// recorded (sxx+syy+szz)/3, so each normal stress gets a third; VX, VY, VZ their field
__global__ __launch_bounds__(256) void adj3d_inject(Layout3D L, Adj3D A, Acq Q, const float *__restrict__ adj, int t)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.vol;
    const int r0 = Q.rec_ptr[shot], r1 = Q.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const float v = adj[(long long)t * Q.rec_total + r];
        const long long c = so + Q.rec_cell[r];
        const int ty = Q.rec_type[r];
        if (ty == 1) {
            const float w = v / 3.f;
            atomicAdd(A.l[3] + c, w);
            atomicAdd(A.l[4] + c, w);
            atomicAdd(A.l[5] + c, w);
        } else {
            atomicAdd(A.l[ty - 2] + c, v);
        }
    }
}

template <int H, int RZ>
void launch_stencil_h(int phase, const Layout3D &L, int nshots, int nyc, const Adj3D &A, const TapsT3D &T,
                      hipStream_t st)
{
    constexpr int TZ = RZ * BLOCK_Y;
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NZ + TZ - 1) / TZ, nyc * nshots), block(TILE_X, BLOCK_Y);
    if (phase == 0)
        hipLaunchKernelGGL((adj3d_stencil<H, RZ, 0>), grid, block, 0, st, L, nyc, A, T);
    else
        hipLaunchKernelGGL((adj3d_stencil<H, RZ, 1>), grid, block, 0, st, L, nyc, A, T);
}

}  // namespace

void launch_adj3d_step(int order, const Layout3D &L, int nshots, int nyc, const Adj3D &A, const Coef3D &C,
                       const Cpml3D &P, const Rec3D &R, const Grad3D &G, const TapsT3D &T, const Acq &Q,
                       const float *adj, int t, hipStream_t st)
{
    const long long ncell = (long long)L.NX * L.NY * L.NZ;
    const unsigned b0 = (unsigned)std::min<long long>((ncell * nshots + 255) / 256, 16384);
    const unsigned b1 = (unsigned)std::min<long long>((ncell + 255) / 256, 16384);
    adj3d_inject<<<Q.nshots, 256, 0, st>>>(L, A, Q, adj, t);
    adj3d_terms<0><<<b0, 256, 0, st>>>(L, nshots, A, C, P, R, G);
    for (int phase = 0; phase < 2; phase++) {
        if (phase == 1) adj3d_terms<1><<<b1, 256, 0, st>>>(L, nshots, A, C, P, R, G);
        switch (order / 2) {
        case 1: launch_stencil_h<1, 2>(phase, L, nshots, nyc, A, T, st); break;
        case 2: launch_stencil_h<2, 2>(phase, L, nshots, nyc, A, T, st); break;
        case 3: launch_stencil_h<3, 2>(phase, L, nshots, nyc, A, T, st); break;
        case 4: launch_stencil_h<4, 2>(phase, L, nshots, nyc, A, T, st); break;
        case 5: launch_stencil_h<5, 2>(phase, L, nshots, nyc, A, T, st); break;
        case 6: launch_stencil_h<6, 2>(phase, L, nshots, nyc, A, T, st); break;
        default: break;
        }
    }
}

}  // namespace wv
