// 3D elastic staggered-grid kernels for gfx950 (ForwardSolver3Delastic.cpp:120-416).
//
// One time step = two launches over every shot of the batch:
//   velocity: vx, vy, vz from the six stresses (9 CPML terms), after the
//             velocity-type sources / receivers left pending by the previous step
//   stress  : the six stresses from the new velocities (9 CPML terms), free
//             surface and ABS of the stresses fused, then the stress-type
//             sources (P into sxx, syy, szz) and receivers ((sxx+syy+szz)/3)
// Blocks tile the x-z plane (64 columns x 4*RZ rows, one wavefront per row
// group) and march along depth y over a chunk of planes.  Each plane's x-z
// tile plus its halo is staged in LDS for the x and z derivatives; the y
// derivatives come from a register queue of the 2H+1 planes around the
// current one, so every plane of a field is read from HBM once per block.
#include "elastic3d.hpp"

namespace wv {

namespace {

constexpr int NTH = TILE_X * BLOCK_Y;

// Thread share of the region [x0-HX, x0+64+HX) x [z0-HZ, z0+TZ+HZ) of one y plane.
template <int HX, int HZ, int TZ>
struct Region {
    static constexpr int RW = TILE_X + 2 * HX;
    static constexpr int RH = TZ + 2 * HZ;
    static constexpr int TOT = RW * RH;
    static constexpr int IT = (TOT + NTH - 1) / NTH;
    float v[IT];
    __device__ __forceinline__ void load(const float *__restrict__ g, const Layout3D &L, int x0, int y, int z0, int tid)
    {
        const long long base = L.at(x0 - HX, y, z0 - HZ);
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * NTH;
            if (TOT % NTH == 0 || i < TOT) {
                const int r = i / RW;
                const int c = i - r * RW;
                v[k] = g[base + (long long)r * L.pitch + c];
            }
        }
    }
    __device__ __forceinline__ void store(float (*S)[RW], int tid) const
    {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * NTH;
            if (TOT % NTH == 0 || i < TOT) {
                const int r = i / RW;
                const int c = i - r * RW;
                S[r][c] = v[k];
            }
        }
    }
};

// derivative along x at LDS row r (array with an H-wide x halo), column tx
template <int H, int RW>
__device__ __forceinline__ float ddx(const float (*S)[RW], int r, int tx, const float *c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[r][tx + k];
    return s;
}

// derivative along z at output row lz (array with an H-wide z halo), LDS column cx
template <int H, int RW>
__device__ __forceinline__ float ddz(const float (*S)[RW], int lz, int cx, const float *__restrict__ c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[lz + k][cx];
    return s;
}

// derivative along y from the register queue (q[k] = plane y - H + k)
template <int H, int RZ>
__device__ __forceinline__ float ddy(const float (&q)[2 * H + 1][RZ], int r, const float *__restrict__ c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * q[k][r];
    return s;
}

template <int H, int RZ>
__device__ __forceinline__ void shift(float (&q)[2 * H + 1][RZ])
{
#pragma unroll
    for (int k = 0; k < 2 * H; k++)
#pragma unroll
        for (int r = 0; r < RZ; r++) q[k][r] = q[k + 1][r];
}

// psi = b*psi + a*d ; d += psi   (CPML.cpp:84-95)
__device__ __forceinline__ float cpml(float &d, float psi, float a, float b)
{
    float p = b * psi;
    p = p + a * d;
    d = d + p;
    return p;
}

__device__ __forceinline__ bool velocity_type(int ty) { return ty >= 2; }

// Sources / receivers owned by block sb of one shot at step t, velocity-type
// (VX, VY, VZ) or stress-type (P).  Uniform per block.
__device__ void acquisition3d(const Acq &A, const Fields3D &F, long long so, int sb, int t, bool vel)
{
    const int s0 = A.bsrc_ptr[sb], s1 = A.bsrc_ptr[sb + 1];
    const int r0 = A.brec_ptr[sb], r1 = A.brec_ptr[sb + 1];
    if (s0 == s1 && r0 == r1) return;
    const int tid = threadIdx.y * TILE_X + threadIdx.x;
    __syncthreads();
    if (tid == 0) {
        // components in the reference's order P, VX, VY, VZ; each in list order
        for (int ty = 1; ty <= 4; ty++) {
            if (velocity_type(ty) != vel) continue;
            for (int e = s0; e < s1; e++) {
                const int s = A.bsrc[e];
                if (A.src_type[s] != ty) continue;
                const float v = A.src_sig[(long long)s * A.NT + t];
                const long long c = so + A.src_cell[s];
                if (ty == 1) {
                    F.f[3][c] += v;
                    F.f[4][c] += v;
                    F.f[5][c] += v;
                } else {
                    F.f[ty - 2][c] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int e = r0 + tid; e < r1; e += NTH) {
        const int r = A.brec[e];
        const int ty = A.rec_type[r];
        if (velocity_type(ty) != vel) continue;
        const long long c = so + A.rec_cell[r];
        float v;
        if (ty == 1) {
            v = F.f[3][c];
            v += F.f[4][c];
            v += F.f[5][c];
            v /= 3.f;
        } else {
            v = F.f[ty - 2][c];
        }
        A.seis[(long long)t * A.rec_total + r] = v;
    }
    __syncthreads();
}

struct Pos {
    int shot, x0, z0, ys, ye, sb;
};

// Block position.  Workgroups are dealt to the 8 XCDs round robin (block b
// on XCD b mod 8); renumbering them so that each XCD works on one contiguous
// range of (x tile, z tile, y chunk, shot) keeps shared halos in its L2.
__device__ __forceinline__ Pos position(const Layout3D &L, int TZ, int nyc, int LY, const Acq &A)
{
    const int nbx = gridDim.x, nbz = gridDim.y;
    const int nblk = nbx * nbz * gridDim.z;
    const int b = blockIdx.x + nbx * (blockIdx.y + nbz * blockIdx.z);
    const int xcd = b & 7, k = b >> 3, q = nblk >> 3, rem = nblk & 7;
    const int t = xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
    const int bx = t % nbx, u = t / nbx, bz = u % nbz, bzz = u / nbz;
    Pos p;
    p.shot = bzz / nyc;
    const int yc = bzz - p.shot * nyc;
    p.x0 = bx * TILE_X;
    p.z0 = bz * TZ;
    p.ys = yc * LY;
    p.ye = min(L.NY, p.ys + LY);
    p.sb = (bzz * A.nby + bz) * A.nbx + bx;
    return p;
}

template <int H, int RZ>
__global__ __launch_bounds__(NTH) void el3d_velocity(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                     int nyc, int LY, int tofs, int pend)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RX = Region<H, 0, TZ>;   // x derivatives only
    using RXZ = Region<H, H, TZ>;  // x and z derivatives
    using RZZ = Region<0, H, TZ>;  // z derivatives only
    __shared__ float s_xx[RX::RH][RX::RW];
    __shared__ float s_xy[RX::RH][RX::RW];
    __shared__ float s_xz[RXZ::RH][RXZ::RW];
    __shared__ float s_yz[RZZ::RH][RZZ::RW];
    __shared__ float s_zz[RZZ::RH][RZZ::RW];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (pend) acquisition3d(A, F, so, q.sb, A.tbase[0] + tofs - 1, true);

    const float *__restrict__ sxx = F.f[3] + so;
    const float *__restrict__ syy = F.f[4] + so;
    const float *__restrict__ szz = F.f[5] + so;
    const float *__restrict__ syz = F.f[6] + so;
    const float *__restrict__ sxz = F.f[7] + so;
    const float *__restrict__ sxy = F.f[8] + so;
    float *__restrict__ vx = F.f[0] + so;
    float *__restrict__ vy = F.f[1] + so;
    float *__restrict__ vz = F.f[2] + so;

    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxf[K], cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        cxf[k] = T.xf[(long long)x * K + k];
        cxb[k] = T.xb[(long long)x * K + k];
    }
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    float ax = 0.f, bx = 0.f, axh = 0.f, bxh = 0.f;
    if (xi >= 0) {
        ax = P.ax[x];
        bx = P.bx[x];
        axh = P.axh[x];
        bxh = P.bxh[x];
    }
    const int W2 = 2 * P.W;

    // y queues of the stresses differentiated along y: sxy (Dyb), syy (Dyf), syz (Dyb)
    float qxy[K][RZ], qyy[K][RZ], qyz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++)
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const long long o = L.at(x, q.ys - H + k, q.z0 + wz + r * BLOCK_Y);
            qxy[k + 1][r] = sxy[o];
            qyy[k + 1][r] = syy[o];
            qyz[k + 1][r] = syz[o];
        }

    for (int y = q.ys; y < q.ye; y++) {
        shift<H, RZ>(qxy);
        shift<H, RZ>(qyy);
        shift<H, RZ>(qyz);
        float ovx[RZ], ovy[RZ], ovz[RZ], crx[RZ], cry[RZ], crz[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = q.z0 + wz + r * BLOCK_Y;
            const long long on = L.at(x, y + H, z);
            qxy[2 * H][r] = sxy[on];
            qyy[2 * H][r] = syy[on];
            qyz[2 * H][r] = syz[on];
            const long long o = L.at(x, y, z);
            ovx[r] = vx[o];
            ovy[r] = vy[o];
            ovz[r] = vz[o];
            crx[r] = C.irx[o];
            cry[r] = C.iry[o];
            crz[r] = C.irz[o];
        }
        RX rxx, rxy;
        RXZ rxz;
        RZZ ryz, rzz;
        rxx.load(sxx, L, q.x0, y, q.z0, tid);
        rxy.load(sxy, L, q.x0, y, q.z0, tid);
        rxz.load(sxz, L, q.x0, y, q.z0, tid);
        ryz.load(syz, L, q.x0, y, q.z0, tid);
        rzz.load(szz, L, q.x0, y, q.z0, tid);
        const int yi = P.on ? P.yi[y] : -1;
        // CPML memory variables of this plane
        float p0[RZ], p3[RZ], p6[RZ], p1[RZ], p4[RZ], p7[RZ], p2[RZ], p5[RZ], p8[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = q.z0 + wz + r * BLOCK_Y;
            p0[r] = p3[r] = p6[r] = p1[r] = p4[r] = p7[r] = p2[r] = p5[r] = p8[r] = 0.f;
            if (z >= L.NZ) continue;
            if (xi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * L.NZ + z) * W2 + xi;
                p0[r] = P.ps[0][pi];
                p3[r] = P.ps[3][pi];
                p6[r] = P.ps[6][pi];
            }
            if (yi >= 0 && xin) {
                const long long pi = (((long long)q.shot * W2 + yi) * L.NZ + z) * L.NX + x;
                p1[r] = P.ps[1][pi];
                p4[r] = P.ps[4][pi];
                p7[r] = P.ps[7][pi];
            }
            const int zi = P.on ? P.zi[z] : -1;
            if (zi >= 0 && xin) {
                const long long pi = (((long long)q.shot * L.NY + y) * W2 + zi) * L.NX + x;
                p2[r] = P.ps[2][pi];
                p5[r] = P.ps[5][pi];
                p8[r] = P.ps[8][pi];
            }
        }
        rxx.store(s_xx, tid);
        rxy.store(s_xy, tid);
        rxz.store(s_xz, tid);
        ryz.store(s_yz, tid);
        rzz.store(s_zz, tid);
        __syncthreads();

        const float *cyb = T.yb_v + (long long)y * K;
        const float *cyf = T.yf_v + (long long)y * K;
        float ay = 0.f, by = 0.f, ayh = 0.f, byh = 0.f;
        if (yi >= 0) {
            ay = P.ay[y];
            by = P.by[y];
            ayh = P.ayh[y];
            byh = P.byh[y];
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            const float *czb = T.zb + (long long)z * K;
            const float *czf = T.zf + (long long)z * K;
            const int zi = P.on ? P.zi[z] : -1;
            float az = 0.f, bz = 0.f, azh = 0.f, bzh = 0.f;
            if (zi >= 0) {
                az = P.az[z];
                bz = P.bz[z];
                azh = P.azh[z];
                bzh = P.bzh[z];
            }
            // vx += 1/rho_x (Dxf sxx + Dyb sxy + Dzb sxz)
            float u = ddx<H>(s_xx, lz, tx, cxf);
            float t = ddy<H, RZ>(qxy, r, cyb);
            float w = ddz<H>(s_xz, lz, tx + H, czb);
            // vy += 1/rho_y (Dxb sxy + Dyf syy + Dzb syz)
            float u2 = ddx<H>(s_xy, lz, tx, cxb);
            float t2 = ddy<H, RZ>(qyy, r, cyf);
            float w2 = ddz<H>(s_yz, lz, tx, czb);
            // vz += 1/rho_z (Dxb sxz + Dyb syz + Dzf szz)
            float u3 = ddx<H>(s_xz, lz + H, tx, cxb);
            float t3 = ddy<H, RZ>(qyz, r, cyb);
            float w3 = ddz<H>(s_zz, lz, tx, czf);
            if (xi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * L.NZ + z) * W2 + xi;
                P.ps[0][pi] = cpml(u, p0[r], axh, bxh);
                P.ps[3][pi] = cpml(u2, p3[r], ax, bx);
                P.ps[6][pi] = cpml(u3, p6[r], ax, bx);
            }
            if (yi >= 0) {
                const long long pi = (((long long)q.shot * W2 + yi) * L.NZ + z) * L.NX + x;
                P.ps[1][pi] = cpml(t, p1[r], ay, by);
                P.ps[4][pi] = cpml(t2, p4[r], ayh, byh);
                P.ps[7][pi] = cpml(t3, p7[r], ay, by);
            }
            if (zi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * W2 + zi) * L.NX + x;
                P.ps[2][pi] = cpml(w, p2[r], az, bz);
                P.ps[5][pi] = cpml(w2, p5[r], az, bz);
                P.ps[8][pi] = cpml(w3, p8[r], azh, bzh);
            }
            const long long o = L.at(x, y, z);
            u += t;
            u += w;
            u *= crx[r];
            vx[o] = ovx[r] + u;
            u2 += t2;
            u2 += w2;
            u2 *= cry[r];
            vy[o] = ovy[r] + u2;
            u3 += t3;
            u3 += w3;
            u3 *= crz[r];
            vz[o] = ovz[r] + u3;
        }
        __syncthreads();
    }
}

template <int H, int RZ>
__global__ __launch_bounds__(NTH) void el3d_stress(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                   int nyc, int LY, int tofs)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RXZ = Region<H, H, TZ>;
    __shared__ float s_vx[RXZ::RH][RXZ::RW];
    __shared__ float s_vy[RXZ::RH][RXZ::RW];
    __shared__ float s_vz[RXZ::RH][RXZ::RW];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);

    const float *__restrict__ vx = F.f[0] + so;
    const float *__restrict__ vy = F.f[1] + so;
    const float *__restrict__ vz = F.f[2] + so;
    float *__restrict__ sxx = F.f[3] + so;
    float *__restrict__ syy = F.f[4] + so;
    float *__restrict__ szz = F.f[5] + so;
    float *__restrict__ syz = F.f[6] + so;
    float *__restrict__ sxz = F.f[7] + so;
    float *__restrict__ sxy = F.f[8] + so;

    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxf[K], cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        cxf[k] = T.xf[(long long)x * K + k];
        cxb[k] = T.xb[(long long)x * K + k];
    }
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    float ax = 0.f, bx = 0.f, axh = 0.f, bxh = 0.f;
    if (xi >= 0) {
        ax = P.ax[x];
        bx = P.bx[x];
        axh = P.axh[x];
        bxh = P.bxh[x];
    }
    const int W2 = 2 * P.W;

    // y queues: vx (Dyf), vy (Dyb), vz (Dyf)
    float qvx[K][RZ], qvy[K][RZ], qvz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++)
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const long long o = L.at(x, q.ys - H + k, q.z0 + wz + r * BLOCK_Y);
            qvx[k + 1][r] = vx[o];
            qvy[k + 1][r] = vy[o];
            qvz[k + 1][r] = vz[o];
        }

    for (int y = q.ys; y < q.ye; y++) {
        shift<H, RZ>(qvx);
        shift<H, RZ>(qvy);
        shift<H, RZ>(qvz);
        float o_xx[RZ], o_yy[RZ], o_zz[RZ], o_yz[RZ], o_xz[RZ], o_xy[RZ];
        float cM[RZ], cmu[RZ], cxy[RZ], cxz[RZ], cyz[RZ], dmp[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = q.z0 + wz + r * BLOCK_Y;
            const long long on = L.at(x, y + H, z);
            qvx[2 * H][r] = vx[on];
            qvy[2 * H][r] = vy[on];
            qvz[2 * H][r] = vz[on];
            const long long o = L.at(x, y, z);
            o_xx[r] = sxx[o];
            o_yy[r] = syy[o];
            o_zz[r] = szz[o];
            o_yz[r] = syz[o];
            o_xz[r] = sxz[o];
            o_xy[r] = sxy[o];
            cM[r] = C.M[o];
            cmu[r] = C.mu[o];
            cxy[r] = C.muxy[o];
            cxz[r] = C.muxz[o];
            cyz[r] = C.muyz[o];
            dmp[r] = C.damp ? C.damp[o] : 1.f;
        }
        RXZ rvx, rvy, rvz;
        rvx.load(vx, L, q.x0, y, q.z0, tid);
        rvy.load(vy, L, q.x0, y, q.z0, tid);
        rvz.load(vz, L, q.x0, y, q.z0, tid);
        const int yi = P.on ? P.yi[y] : -1;
        float p9[RZ], p13[RZ], p15[RZ], p10[RZ], p12[RZ], p17[RZ], p11[RZ], p14[RZ], p16[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = q.z0 + wz + r * BLOCK_Y;
            p9[r] = p13[r] = p15[r] = p10[r] = p12[r] = p17[r] = p11[r] = p14[r] = p16[r] = 0.f;
            if (z >= L.NZ) continue;
            if (xi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * L.NZ + z) * W2 + xi;
                p9[r] = P.ps[9][pi];
                p13[r] = P.ps[13][pi];
                p15[r] = P.ps[15][pi];
            }
            if (yi >= 0 && xin) {
                const long long pi = (((long long)q.shot * W2 + yi) * L.NZ + z) * L.NX + x;
                p10[r] = P.ps[10][pi];
                p12[r] = P.ps[12][pi];
                p17[r] = P.ps[17][pi];
            }
            const int zi = P.on ? P.zi[z] : -1;
            if (zi >= 0 && xin) {
                const long long pi = (((long long)q.shot * L.NY + y) * W2 + zi) * L.NX + x;
                p11[r] = P.ps[11][pi];
                p14[r] = P.ps[14][pi];
                p16[r] = P.ps[16][pi];
            }
        }
        rvx.store(s_vx, tid);
        rvy.store(s_vy, tid);
        rvz.store(s_vz, tid);
        __syncthreads();

        const float *cyb = T.yb + (long long)y * K;
        const float *cyf = T.yf + (long long)y * K;
        float ay = 0.f, by = 0.f, ayh = 0.f, byh = 0.f;
        if (yi >= 0) {
            ay = P.ay[y];
            by = P.by[y];
            ayh = P.ayh[y];
            byh = P.byh[y];
        }
        const bool surface = C.fsH && y == 0;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            const float *czb = T.zb + (long long)z * K;
            const float *czf = T.zf + (long long)z * K;
            const int zi = P.on ? P.zi[z] : -1;
            float az = 0.f, bz = 0.f, azh = 0.f, bzh = 0.f;
            if (zi >= 0) {
                az = P.az[z];
                bz = P.bz[z];
                azh = P.azh[z];
                bzh = P.bzh[z];
            }
            float vxx = ddx<H>(s_vx, lz + H, tx, cxb);
            float vyy = ddy<H, RZ>(qvy, r, cyb);
            float vzz = ddz<H>(s_vz, lz, tx + H, czb);
            float vxy = ddy<H, RZ>(qvx, r, cyf);
            float vyx = ddx<H>(s_vy, lz + H, tx, cxf);
            float vxz = ddz<H>(s_vx, lz, tx + H, czf);
            float vzx = ddx<H>(s_vz, lz + H, tx, cxf);
            float vyz = ddz<H>(s_vy, lz, tx + H, czf);
            float vzy = ddy<H, RZ>(qvz, r, cyf);
            if (xi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * L.NZ + z) * W2 + xi;
                P.ps[9][pi] = cpml(vxx, p9[r], ax, bx);
                P.ps[13][pi] = cpml(vyx, p13[r], axh, bxh);
                P.ps[15][pi] = cpml(vzx, p15[r], axh, bxh);
            }
            if (yi >= 0) {
                const long long pi = (((long long)q.shot * W2 + yi) * L.NZ + z) * L.NX + x;
                P.ps[10][pi] = cpml(vyy, p10[r], ay, by);
                P.ps[12][pi] = cpml(vxy, p12[r], ayh, byh);
                P.ps[17][pi] = cpml(vzy, p17[r], ayh, byh);
            }
            if (zi >= 0) {
                const long long pi = (((long long)q.shot * L.NY + y) * W2 + zi) * L.NX + x;
                P.ps[11][pi] = cpml(vzz, p11[r], az, bz);
                P.ps[14][pi] = cpml(vxz, p14[r], azh, bzh);
                P.ps[16][pi] = cpml(vyz, p16[r], azh, bzh);
            }
            // normal stresses (ForwardSolver3Delastic.cpp: M (vxx+vyy+vzz) - 2 mu (...))
            float s = vxx + vyy;
            s += vzz;
            s *= cM[r];
            float nxx = o_xx[r] + s, nyy = o_yy[r] + s, nzz = o_zz[r] + s;
            float m = vyy + vzz;
            m *= cmu[r];
            nxx -= 2.f * m;
            m = vxx + vzz;
            m *= cmu[r];
            nyy -= 2.f * m;
            m = vxx + vyy;
            m *= cmu[r];
            nzz -= 2.f * m;
            float sh = vxy + vyx;
            sh *= cxy[r];
            const float nxy = o_xy[r] + sh;
            sh = vxz + vzx;
            sh *= cxz[r];
            const float nxz = o_xz[r] + sh;
            sh = vyz + vzy;
            sh *= cyz[r];
            const float nyz = o_yz[r] + sh;
            if (surface) {
                // FreeSurface3Delastic: Syy = 0 at the surface, then the horizontal update exchange
                const long long fo = (long long)(z + L.H) * L.pitch + x + XPAD;
                nyy *= 0.f;
                float hs = vxx + vzz;
                const float a = C.fsH[fo] * hs;
                nxx += a;
                nzz += a;
                const float b = C.fsV[fo] * vyy;
                nxx -= b;
                nzz -= b;
            }
            const long long o = L.at(x, y, z);
            const float d = dmp[r];
            sxx[o] = nxx * d;
            syy[o] = nyy * d;
            szz[o] = nzz * d;
            sxy[o] = nxy * d;
            sxz[o] = nxz * d;
            syz[o] = nyz * d;
        }
        __syncthreads();
    }
    acquisition3d(A, F, so, q.sb, A.tbase[0] + tofs, false);
}

__global__ __launch_bounds__(256) void flush3d(Layout3D L, Fields3D F, Acq A, int tofs)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.vol;
    const int t = A.tbase[0] + tofs;
    if (t < 0 || t >= A.NT) return;
    if (threadIdx.x == 0) {
        const int s0 = A.src_ptr[shot], s1 = A.src_ptr[shot + 1];
        for (int ty = 2; ty <= 4; ty++)
            for (int s = s0; s < s1; s++) {
                if (A.src_type[s] != ty) continue;
                F.f[ty - 2][so + A.src_cell[s]] += A.src_sig[(long long)s * A.NT + t];
            }
    }
    __syncthreads();
    const int r0 = A.rec_ptr[shot], r1 = A.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int ty = A.rec_type[r];
        if (!velocity_type(ty)) continue;
        A.seis[(long long)t * A.rec_total + r] = F.f[ty - 2][so + A.rec_cell[r]];
    }
}

__global__ void abs_velocity3d(long long n, float *a, float *b, float *c, const float *__restrict__ damp, long long vol)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = damp[i % vol];
        a[i] *= d;
        b[i] *= d;
        c[i] *= d;
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

template <int H, int RZ>
void launch_h3(int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
               const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
               hipEvent_t e1)
{
    constexpr int TZ = RZ * BLOCK_Y;
    const int LY = (L.NY + nyc - 1) / nyc;
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NZ + TZ - 1) / TZ, nyc * nshots), block(TILE_X, BLOCK_Y);
    if (which == 0)
        go(el3d_velocity<H, RZ>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, pend);
    else
        go(el3d_stress<H, RZ>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs);
}

template <int H>
void launch_rz(int rz, int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
               const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
               hipEvent_t e1)
{
    if (rz == 1)
        launch_h3<H, 1>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
    else if (rz == 4)
        launch_h3<H, 4>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
    else
        launch_h3<H, 2>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
}

}  // namespace

void launch_el3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: launch_rz<1>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 2: launch_rz<2>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 3: launch_rz<3>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 4: launch_rz<4>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 5: launch_rz<5>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 6: launch_rz<6>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    default: break;
    }
}

void launch_abs_velocity3d(const Layout3D &L, int nshots, const Fields3D &F, const Coef3D &C, hipStream_t stream)
{
    const long long n = (long long)nshots * L.vol;
    abs_velocity3d<<<2048, 256, 0, stream>>>(n, F.f[0], F.f[1], F.f[2], C.damp, L.vol);
}

void launch_flush3d(const Layout3D &L, const Fields3D &F, const Acq &A, int tofs, hipStream_t stream)
{
    flush3d<<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
}

}  // namespace wv
