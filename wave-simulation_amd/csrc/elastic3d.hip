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
#include <type_traits>

#include "buffer.hpp"
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
    // rows of an LDS array that every thread's share fits in (surplus threads write padding)
    static constexpr int RHP = (IT * NTH + RW - 1) / RW;
    // the same without a guard (S has RHP rows), so no divergent branch sits between the loads
    // in flight and this store
    __device__ __forceinline__ void store_padded(float (*S)[RW], int tid) const
    {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * NTH;
            const int r = i / RW;
            const int c = i - r * RW;
            S[r][c] = v[k];
        }
    }
};

// Byte offsets, within one y plane, of a thread's share of a Region (the same
// for every plane, so they are computed once per block).
template <int HX, int HZ, int TZ>
struct RegionOff {
    using R = Region<HX, HZ, TZ>;
    unsigned o[R::IT];
    __device__ __forceinline__ void init(const Layout3D &L, int x0, int z0, int tid)
    {
#pragma unroll
        for (int k = 0; k < R::IT; k++) {
            const int i = min(tid + k * NTH, R::TOT - 1);  // surplus threads reload the last element
            const int r = i / R::RW, c = i - r * R::RW;
            o[k] = (unsigned)(((z0 - HZ + r + L.H) * L.pitch + (x0 - HX + c + XPAD)) * 4);
        }
    }
};

// One y plane of a field as a buffer resource (plane offsets stay 32-bit for any grid).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float *f, const Layout3D &L, int y)
{
    return brsrc(f + (long long)(y + L.H) * L.slab, (unsigned)(L.slab * 4));
}

// Buffer offset that the hardware answers with 0 (loads) or drops (stores):
// the CPML accesses of lanes outside a frame, without a branch.
constexpr unsigned NO_LANE = 0x7ffffff0u;

// CPML memory variables of one shot as buffers: x frame [NY][NZ][2W], y frame
// [2W][NZ][NX], z frame [NY][2W][NX] (per shot, so offsets stay 32-bit).
// Built once per kernel, before the plane loop.
struct PsiRes3D {
    __amdgpu_buffer_rsrc_t x[3], y[3], z[3];
    __device__ __forceinline__ void init(const Cpml3D &P, const Layout3D &L, int shot, int kx0, int kx1, int kx2,
                                         int ky0, int ky1, int ky2, int kz0, int kz1, int kz2)
    {
        const long long nx = (long long)L.NY * L.NZ * 2 * P.W, ny = (long long)2 * P.W * L.NZ * L.NX,
                        nz = (long long)L.NY * 2 * P.W * L.NX;
        const int kx[3] = {kx0, kx1, kx2}, ky[3] = {ky0, ky1, ky2}, kz[3] = {kz0, kz1, kz2};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            x[j] = brsrc(P.ps[kx[j]] + shot * nx, (unsigned)(nx * 4));
            y[j] = brsrc(P.ps[ky[j]] + shot * ny, (unsigned)(ny * 4));
            z[j] = brsrc(P.ps[kz[j]] + shot * nz, (unsigned)(nz * 4));
        }
    }
};
// lane byte offsets of cell (x, y, z) in the three frames (NO_LANE outside)
__device__ __forceinline__ unsigned psi_ox(const Layout3D &L, int W2, int xi, int y, int z, bool in)
{
    return (in && xi >= 0) ? (unsigned)(((y * L.NZ + z) * W2 + xi) * 4) : NO_LANE;
}
__device__ __forceinline__ unsigned psi_oy(const Layout3D &L, int yi, int x, int z, bool in)
{
    return (in && yi >= 0) ? (unsigned)(((yi * L.NZ + z) * L.NX + x) * 4) : NO_LANE;
}
__device__ __forceinline__ unsigned psi_oz(const Layout3D &L, int W2, int zi, int x, int y, bool in)
{
    return (in && zi >= 0) ? (unsigned)(((y * W2 + zi) * L.NX + x) * 4) : NO_LANE;
}

// This thread's element of plane y's y table (Taps3D::ys / yv layout, 2K + 5 entries): the
// tables are staged in LDS a plane ahead, one element per thread.
template <int K>
__device__ __forceinline__ float ytab_elem(const float *tab, int y, int tid)
{
    return tab[(long long)y * (2 * K + 5) + min(tid, 2 * K + 4)];
}

// z tables of the tile's TZ rows staged in LDS before the plane loop: taps of the backward and
// forward z operators, CPML z index and a_z, b_z, a_z(half), b_z(half) (0 outside the frame)
template <int TZ, int K>
struct ZTab {
    float t[2][TZ][K];
    float c[TZ][4];
    int zi[TZ];
    __device__ __forceinline__ void fill(const Taps3D &T, const Cpml3D &P, const Layout3D &L, int z0, int tid)
    {
        for (int e = tid; e < TZ * K; e += NTH) {
            const int r = e / K, k = e - r * K;
            t[0][r][k] = T.zb[(long long)(z0 + r) * K + k];
            t[1][r][k] = T.zf[(long long)(z0 + r) * K + k];
        }
        if (tid < TZ) {
            const int z = z0 + tid;
            const int zi_ = (P.on && z < L.NZ) ? P.zi[z] : -1;
            zi[tid] = zi_;
            c[tid][0] = zi_ >= 0 ? P.az[z] : 0.f;
            c[tid][1] = zi_ >= 0 ? P.bz[z] : 0.f;
            c[tid][2] = zi_ >= 0 ? P.azh[z] : 0.f;
            c[tid][3] = zi_ >= 0 ? P.bzh[z] : 0.f;
        }
    }
};

// Whether any cell of the block's x-z tile lies in the x or z CPML frame (block-uniform).
__device__ __forceinline__ bool tile_in_frame(const Layout3D &L, const Cpml3D &P, int x0, int z0, int TZ)
{
    return P.on && (x0 < P.W || x0 + TILE_X > L.NX - P.W || z0 < P.W || z0 + TZ > L.NZ - P.W);
}

template <int HX, int HZ, int TZ>
__device__ __forceinline__ void load_region(Region<HX, HZ, TZ> &g, const RegionOff<HX, HZ, TZ> &o, const float *f,
                                            const Layout3D &L, int y)
{
    const auto b = plane_rsrc(f, L, y);
#pragma unroll
    for (int k = 0; k < Region<HX, HZ, TZ>::IT; k++) g.v[k] = bload(b, o.o[k], 0);
}

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

// Acquisition modes of the 3D kernels: elastic, acoustic, EM (fields hx hy hz ex ey ez)
enum { ACQ_ELASTIC = 0, ACQ_ACOUSTIC = 1, ACQ_EM = 2 };

// Source / receiver types that belong to the first phase of a step (velocities, or H in EM:
// HZ) and are applied after the second phase has read that phase's fields, i.e. at the start of
// the next step's first launch or by the flush at the end of a run segment.
template <int MODE>
__device__ __forceinline__ bool velocity_type(int ty)
{
    return MODE == ACQ_EM ? ty == 4 : ty >= 2;
}
// field of a source / receiver type other than the elastic P (EM: EZ, EX, EY, HZ -> 5, 3, 4, 2;
// SourceReceiverImplEM.cpp:100-115; seismic: VX, VY, VZ -> 0, 1, 2, acoustic P -> 3)
template <int MODE>
__device__ __forceinline__ int type_field(int ty)
{
    if (MODE == ACQ_EM) return ty == 1 ? 5 : (ty == 4 ? 2 : ty + 1);
    return ty == 1 ? 3 : ty - 2;
}

// Sources / receivers owned by block sb of one shot at step t, velocity-type
// (VX, VY, VZ; EM: HZ) or stress-type (P: sxx, syy, szz in elastic, p = field 3 in acoustic;
// EM: EZ, EX, EY).  Uniform per block.
template <int MODE>
__device__ void acquisition3d(const Acq &A, const Fields3D &F, long long so, int sb, int t, bool vel)
{
    const int s0 = A.bsrc_ptr[sb], s1 = A.bsrc_ptr[sb + 1];
    const int r0 = A.brec_ptr[sb], r1 = A.brec_ptr[sb + 1];
    if (s0 == s1 && r0 == r1) return;
    const int tid = threadIdx.y * TILE_X + threadIdx.x;
    __syncthreads();
    if (tid == 0) {
        // components in the reference's order P (EZ), VX (EX), VY (EY), VZ (HZ); each in list order
        for (int ty = 1; ty <= 4; ty++) {
            if (velocity_type<MODE>(ty) != vel) continue;
            for (int e = s0; e < s1; e++) {
                const int s = A.bsrc[e];
                if (A.src_type[s] != ty) continue;
                const float v = A.src_sig[(long long)s * A.NT + t];
                const long long c = so + A.src_cell[s];
                if (ty == 1 && MODE == ACQ_ELASTIC) {
                    F.f[3][c] += v;
                    F.f[4][c] += v;
                    F.f[5][c] += v;
                } else {
                    F.f[type_field<MODE>(ty)][c] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int e = r0 + tid; e < r1; e += NTH) {
        const int r = A.brec[e];
        const int ty = A.rec_type[r];
        if (velocity_type<MODE>(ty) != vel) continue;
        const long long c = so + A.rec_cell[r];
        float v;
        if (ty == 1 && MODE == ACQ_ELASTIC) {
            v = F.f[3][c];
            v += F.f[4][c];
            v += F.f[5][c];
            v /= 3.f;
        } else {
            v = F.f[type_field<MODE>(ty)][c];
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

template <int H, int RZ, bool REC>
__global__ __launch_bounds__(NTH) void el3d_velocity(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                     int nyc, int LY, int tofs, int pend, Rec3D R)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RX = Region<H, 0, TZ>;   // x derivatives only
    using RXZ = Region<H, H, TZ>;  // x and z derivatives
    using RZZ = Region<0, H, TZ>;  // z derivatives only
    __shared__ float s_xx[RX::RHP][RX::RW];
    __shared__ float s_xy[RX::RHP][RX::RW];
    __shared__ float s_xz[RXZ::RHP][RXZ::RW];
    __shared__ float s_yz[RZZ::RHP][RZZ::RW];
    __shared__ float s_zz[RZZ::RHP][RZZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];  // the plane's y table in [0, 2K+5); every thread writes its slot
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (pend) acquisition3d<ACQ_ELASTIC>(A, F, so, q.sb, A.tbase[0] + tofs - 1, true);

    const float *sxx = F.f[3] + so, *syy = F.f[4] + so, *szz = F.f[5] + so;
    const float *syz = F.f[6] + so, *sxz = F.f[7] + so, *sxy = F.f[8] + so;
    float *vx = F.f[0] + so, *vy = F.f[1] + so, *vz = F.f[2] + so;

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
    // byte offsets within a plane: the thread's own cells and its shares of the regions
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, 0, TZ> ox;
    RegionOff<H, H, TZ> oxz;
    RegionOff<0, H, TZ> ozz;
    ox.init(L, q.x0, q.z0, tid);
    oxz.init(L, q.x0, q.z0, tid);
    ozz.init(L, q.x0, q.z0, tid);

    // y queues of the stresses differentiated along y: sxy (Dyb), syy (Dyf), syz (Dyb)
    float qxy[K][RZ], qyy[K][RZ], qyz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b0 = plane_rsrc(sxy, L, q.ys - H + k), b1 = plane_rsrc(syy, L, q.ys - H + k),
                   b2 = plane_rsrc(syz, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qxy[k + 1][r] = bload(b0, oc[r], 0);
            qyy[k + 1][r] = bload(b1, oc[r], 0);
            qyz[k + 1][r] = bload(b2, oc[r], 0);
        }
    }

    // everything a plane reads besides the CPML memory variables; the next plane's set is
    // loaded while the current one is computed
    struct PlaneIn {
        float nxy[RZ], nyy[RZ], nyz[RZ];  // queue entries at y + H
        float o[3][RZ];                   // vx vy vz at y
        float c[3][RZ];                   // 1/rho_x 1/rho_y 1/rho_z at y
        float yt;                         // this thread's element of the plane's y tables
    };
    const float *const cr[3] = {C.irx, C.iry, C.irz};
    const float *const vv[3] = {vx, vy, vz};
    RX rxx, rxy;
    RXZ rxz;
    RZZ ryz, rzz;
    auto load_plane = [&](PlaneIn &pi, int y) {
        // the y table first: the LDS store that waits for it then waits for nothing issued after it
        pi.yt = ytab_elem<K>(T.yv, y, tid);
        load_region(rxx, ox, sxx, L, y);
        load_region(rxy, ox, sxy, L, y);
        load_region(rxz, oxz, sxz, L, y);
        load_region(ryz, ozz, syz, L, y);
        load_region(rzz, ozz, szz, L, y);
        {
            const auto b0 = plane_rsrc(sxy, L, y + H), b1 = plane_rsrc(syy, L, y + H), b2 = plane_rsrc(syz, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                pi.nxy[r] = bload(b0, oc[r], 0);
                pi.nyy[r] = bload(b1, oc[r], 0);
                pi.nyz[r] = bload(b2, oc[r], 0);
            }
        }
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const auto b = plane_rsrc(vv[f], L, y), bc = plane_rsrc(cr[f], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                pi.o[f][r] = bload(b, oc[r], 0);
                pi.c[f][r] = bload(bc, oc[r], 0);
            }
        }
    };

    PlaneIn cur;
    load_plane(cur, q.ys);
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);  // ordered before its first use by the first plane's barrier
    PsiRes3D pr;  // 0 sxx_x, 3 sxy_x, 6 sxz_x | 1 sxy_y, 4 syy_y, 7 syz_y | 2 sxz_z, 5 syz_z, 8 szz_z
    pr.init(P, L, q.shot, 0, 3, 6, 1, 4, 7, 2, 5, 8);

    for (int y = q.ys; y < q.ye; y++) {
        shift<H, RZ>(qxy);
        shift<H, RZ>(qyy);
        shift<H, RZ>(qyz);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qxy[2 * H][r] = cur.nxy[r];
            qyy[2 * H][r] = cur.nyy[r];
            qyz[2 * H][r] = cur.nyz[r];
        }
        rxx.store_padded(s_xx, tid);
        rxy.store_padded(s_xy, tid);
        rxz.store_padded(s_xz, tid);
        ryz.store_padded(s_yz, tid);
        rzz.store_padded(s_zz, tid);
        s_yt[tid] = cur.yt;  // (no guard: a divergent store here would wait for every load)
        block_sync_lds();

        // CPML memory variables: tiles and planes outside every frame skip them (a uniform
        // branch); lanes outside a frame read 0 and drop their stores (no divergent branches, so
        // the next plane's loads stay in flight through the plane's arithmetic)
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        // (readfirstlane: a branch the compiler sees as uniform, so y and the plane resources stay scalar)
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        float p0[RZ], p3[RZ], p6[RZ], p1[RZ], p4[RZ], p7[RZ], p2[RZ], p5[RZ], p8[RZ];
        unsigned pox[RZ], poy[RZ], poz[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int z = q.z0 + wz + r * BLOCK_Y;
            const bool in = z < L.NZ && xin;
            p0[r] = p3[r] = p6[r] = p1[r] = p4[r] = p7[r] = p2[r] = p5[r] = p8[r] = 0.f;
            if (fr) {
                pox[r] = psi_ox(L, W2, xi, y, z, in);
                poy[r] = psi_oy(L, yi, x, z, in);
                poz[r] = psi_oz(L, W2, s_zt.zi[wz + r * BLOCK_Y], x, y, in);
                p0[r] = bload(pr.x[0], pox[r], 0);
                p3[r] = bload(pr.x[1], pox[r], 0);
                p6[r] = bload(pr.x[2], pox[r], 0);
                p1[r] = bload(pr.y[0], poy[r], 0);
                p4[r] = bload(pr.y[1], poy[r], 0);
                p7[r] = bload(pr.y[2], poy[r], 0);
                p2[r] = bload(pr.z[0], poz[r], 0);
                p5[r] = bload(pr.z[1], poz[r], 0);
                p8[r] = bload(pr.z[2], poz[r], 0);
            }
        }
        // (also after the last plane: plane ye and its queue plane ye + H lie inside the padded
        // volume, and a branch here would make the compiler wait for every load at the join)
        PlaneIn nxt;
        load_plane(nxt, y + 1);

        const float *cyb = s_yt, *cyf = s_yt + K;
        const bool yf = yi >= 0;
        const float ay = yf ? s_yt[2 * K + 1] : 0.f, by = yf ? s_yt[2 * K + 2] : 0.f;
        const float ayh = yf ? s_yt[2 * K + 3] : 0.f, byh = yf ? s_yt[2 * K + 4] : 0.f;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            const float *czb = s_zt.t[0][lz], *czf = s_zt.t[1][lz];
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
            if (fr) {
                const float az = s_zt.c[lz][0], bz = s_zt.c[lz][1], azh = s_zt.c[lz][2], bzh = s_zt.c[lz][3];
                // outside a frame a = b = 0 and psi = 0: the derivative is unchanged
                bstore(cpml(u, p0[r], axh, bxh), pr.x[0], pox[r], 0);
                bstore(cpml(u2, p3[r], ax, bx), pr.x[1], pox[r], 0);
                bstore(cpml(u3, p6[r], ax, bx), pr.x[2], pox[r], 0);
                bstore(cpml(t, p1[r], ay, by), pr.y[0], poy[r], 0);
                bstore(cpml(t2, p4[r], ayh, byh), pr.y[1], poy[r], 0);
                bstore(cpml(t3, p7[r], ay, by), pr.y[2], poy[r], 0);
                bstore(cpml(w, p2[r], az, bz), pr.z[0], poz[r], 0);
                bstore(cpml(w2, p5[r], az, bz), pr.z[1], poz[r], 0);
                bstore(cpml(w3, p8[r], azh, bzh), pr.z[2], poz[r], 0);
            }
            u += t;
            u += w;
            u2 += t2;
            u2 += w2;
            u3 += t3;
            u3 += w3;
            if (REC) {
                bstore(u, plane_rsrc(R.r[0] + so, L, y), oc[r], 0);
                bstore(u2, plane_rsrc(R.r[1] + so, L, y), oc[r], 0);
                bstore(u3, plane_rsrc(R.r[2] + so, L, y), oc[r], 0);
            }
            u *= cur.c[0][r];
            bstore(cur.o[0][r] + u, plane_rsrc(vx, L, y), oc[r], 0);
            u2 *= cur.c[1][r];
            bstore(cur.o[1][r] + u2, plane_rsrc(vy, L, y), oc[r], 0);
            u3 *= cur.c[2][r];
            bstore(cur.o[2][r] + u3, plane_rsrc(vz, L, y), oc[r], 0);
        }
        block_sync_lds();  // the next plane rewrites the LDS regions
        cur = nxt;
    }
}

// Viscoelastic stress update of one cell from its nine CPML-modified derivatives
// (ForwardSolver3Dviscoelastic.cpp:270-444, generalised standard linear solid with the
// relaxation term in Crank-Nicolson form, in the reference's order of operations):
//   u = M (vxx + vyy + vzz); for l: g = u tauP / tau_l; s_ii += dt/2 R_ii,l; R_ii,l = c1 R_ii,l - g
//   s_ii += (1 + L tauP) u                                        (ii = xx, yy, zz)
//   u = 2 mu (vyy + vzz); for l: R_xx,l = c2 (R_xx,l + u tauS / tau_l); s_xx += dt/2 R_xx,l
//   s_xx -= (1 + L tauS) u                      (yy from vxx + vzz, zz from vxx + vyy)
//   u = mu_ij (d_i v_j + d_j v_i); for l: s_ij += dt/2 R_ij,l;
//       R_ij,l = c2 (c1 R_ij,l - u tauS_ij / tau_l); s_ij += dt/2 R_ij,l;  s_ij += (1 + L tauS) u
// then on the free surface the exchange of FreeSurface3Dviscoelastic.cpp:17-75 for sxx, szz and
// their memory variables, syy and every Ryy zero; ABS damping of the stresses on the store.
// s[] = sxx syy szz syz sxz sxy (updated in place); o = byte offset of the cell in its plane.
__device__ __forceinline__ void ve3d_cell(const Visco3D &V, const Layout3D &L, long long so, int y, unsigned o,
                                          float (&s)[6], float M, float mu, float mxy, float mxz, float myz,
                                          float vxx, float vyy, float vzz, float dxy, float dxz, float dyz,
                                          const float *fsH, const float *fsV, long long fo)
{
    const int NL = V.L;
    const float tP = bload(plane_rsrc(V.tauP, L, y), o, 0), tS = bload(plane_rsrc(V.tauS, L, y), o, 0);
    const float oP = bload(plane_rsrc(V.opP, L, y), o, 0), oS = bload(plane_rsrc(V.opS, L, y), o, 0);
    const float tsh[3] = {bload(plane_rsrc(V.tSyz, L, y), o, 0), bload(plane_rsrc(V.tSxz, L, y), o, 0),
                          bload(plane_rsrc(V.tSxy, L, y), o, 0)};
    float R[6][4];
#pragma unroll
    for (int g = 0; g < 6; g++)
#pragma unroll
        for (int l = 0; l < 4; l++) R[g][l] = l < NL ? bload(plane_rsrc(V.R[4 * g + l] + so, L, y), o, 0) : 0.f;
    const float dth = V.dth;
    float u = vxx + vyy;
    u += vzz;
    u *= M;
#pragma unroll
    for (int l = 0; l < 4; l++)
        if (l < NL) {
            float g = V.itau[l] * u;
            g *= tP;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                s[c] += dth * R[c][l];
                R[c][l] *= V.c1[l];
                R[c][l] -= g;
            }
        }
    u *= oP;
    s[0] += u;
    s[1] += u;
    s[2] += u;
    const float da[3] = {vyy, vxx, vxx}, db[3] = {vzz, vzz, vyy};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        u = da[c] + db[c];
        u *= mu;
        u *= 2.f;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                float g = V.itau[l] * u;
                g *= tS;
                R[c][l] += g;
                R[c][l] *= V.c2[l];
                s[c] += dth * R[c][l];
            }
        u *= oS;
        s[c] -= u;
    }
    // shear: groups 3 yz, 4 xz, 5 xy
    const float dsh[3] = {dyz, dxz, dxy}, msh[3] = {myz, mxz, mxy};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        u = dsh[c] * msh[c];
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                s[3 + c] += dth * R[3 + c][l];
                R[3 + c][l] *= V.c1[l];
                float g = V.itau[l] * u;
                g *= tsh[c];
                R[3 + c][l] -= g;
                R[3 + c][l] *= V.c2[l];
                s[3 + c] += dth * R[3 + c][l];
            }
        u *= oS;
        s[3 + c] += u;
    }
    if (fsH && y == 0) {
        const float hs = vxx + vzz;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                s[0] -= dth * R[0][l];
                s[2] -= dth * R[2][l];
            }
        const float a = fsH[fo] * hs;
        s[0] += a;
        s[2] += a;
        const float b = fsV[fo] * vyy;
        s[0] -= b;
        s[2] -= b;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) {
                const float c = V.fsRH[l][fo] * hs;
                R[0][l] += c;
                R[2][l] += c;
                const float d = V.fsRV[l][fo] * vyy;
                R[0][l] -= d;
                R[2][l] -= d;
                s[0] += dth * R[0][l];
                s[2] += dth * R[2][l];
                R[1][l] *= 0.f;
            }
        s[1] *= 0.f;
    }
#pragma unroll
    for (int g = 0; g < 6; g++)
#pragma unroll
        for (int l = 0; l < 4; l++)
            if (l < NL) bstore(R[g][l], plane_rsrc(V.R[4 * g + l] + so, L, y), o, 0);
}

template <int H, int RZ, bool REC, bool VE = false>
__global__ __launch_bounds__(NTH) void el3d_stress(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                   int nyc, int LY, int tofs, Rec3D R, Visco3D V)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RXZ = Region<H, H, TZ>;
    __shared__ float s_vx[RXZ::RHP][RXZ::RW];
    __shared__ float s_vy[RXZ::RHP][RXZ::RW];
    __shared__ float s_vz[RXZ::RHP][RXZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];  // the plane's y table in [0, 2K+5); every thread writes its slot
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);

    const float *vx = F.f[0] + so, *vy = F.f[1] + so, *vz = F.f[2] + so;
    float *sxx = F.f[3] + so, *syy = F.f[4] + so, *szz = F.f[5] + so;
    float *syz = F.f[6] + so, *sxz = F.f[7] + so, *sxy = F.f[8] + so;

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
    // byte offsets within a plane: the thread's own cells and its share of the x-z region
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, H, TZ> ro;
    ro.init(L, q.x0, q.z0, tid);

    // y queues: vx (Dyf), vy (Dyb), vz (Dyf)
    float qvx[K][RZ], qvy[K][RZ], qvz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b0 = plane_rsrc(vx, L, q.ys - H + k), b1 = plane_rsrc(vy, L, q.ys - H + k),
                   b2 = plane_rsrc(vz, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qvx[k + 1][r] = bload(b0, oc[r], 0);
            qvy[k + 1][r] = bload(b1, oc[r], 0);
            qvz[k + 1][r] = bload(b2, oc[r], 0);
        }
    }

    // everything a plane reads besides the CPML memory variables; the next plane's set is
    // loaded while the current one is computed
    struct PlaneIn {
        float nvx[RZ], nvy[RZ], nvz[RZ];  // queue entries at y + H
        float o[6][RZ];                   // sxx syy szz syz sxz sxy at y
        float c[6][RZ];                   // M mu mu_xy mu_xz mu_yz damp at y
        float yt;                         // this thread's element of the plane's y tables
    };
    const float *const cf[5] = {C.M, C.mu, C.muxy, C.muxz, C.muyz};
    const float *const st[6] = {sxx, syy, szz, syz, sxz, sxy};
    auto load_plane = [&](PlaneIn &pi, RXZ &gx, RXZ &gy, RXZ &gz, int y) {
        // the y table first: the LDS store that waits for it then waits for nothing issued after it
        pi.yt = ytab_elem<K>(T.ys, y, tid);
        load_region(gx, ro, vx, L, y);
        load_region(gy, ro, vy, L, y);
        load_region(gz, ro, vz, L, y);
        {
            const auto b0 = plane_rsrc(vx, L, y + H), b1 = plane_rsrc(vy, L, y + H), b2 = plane_rsrc(vz, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                pi.nvx[r] = bload(b0, oc[r], 0);
                pi.nvy[r] = bload(b1, oc[r], 0);
                pi.nvz[r] = bload(b2, oc[r], 0);
            }
        }
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const auto b = plane_rsrc(st[f], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) pi.o[f][r] = bload(b, oc[r], 0);
        }
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const auto b = plane_rsrc(cf[f], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) pi.c[f][r] = bload(b, oc[r], 0);
        }
        if (C.damp) {
            const auto b = plane_rsrc(C.damp, L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) pi.c[5][r] = bload(b, oc[r], 0);
        } else {
#pragma unroll
            for (int r = 0; r < RZ; r++) pi.c[5][r] = 1.f;
        }
    };

    PlaneIn cur;
    RXZ rvx, rvy, rvz;
    load_plane(cur, rvx, rvy, rvz, q.ys);
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);  // ordered before its first use by the first plane's barrier
    PsiRes3D pr;  // 9 vx_x, 13 vy_x, 15 vz_x | 10 vy_y, 12 vx_y, 17 vz_y | 11 vz_z, 14 vx_z, 16 vy_z
    pr.init(P, L, q.shot, 9, 13, 15, 10, 12, 17, 11, 14, 16);
    const int W2 = 2 * P.W;

    for (int y = q.ys; y < q.ye; y++) {
        shift<H, RZ>(qvx);
        shift<H, RZ>(qvy);
        shift<H, RZ>(qvz);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qvx[2 * H][r] = cur.nvx[r];
            qvy[2 * H][r] = cur.nvy[r];
            qvz[2 * H][r] = cur.nvz[r];
        }
        rvx.store_padded(s_vx, tid);
        rvy.store_padded(s_vy, tid);
        rvz.store_padded(s_vz, tid);
        s_yt[tid] = cur.yt;  // (no guard: a divergent store here would wait for every load)
        block_sync_lds();

        // CPML memory variables: planes and tiles outside every frame take the body without them;
        // in the other, lanes outside a frame read 0 and drop their stores (no branches, so the
        // next plane's loads stay in flight through the plane's arithmetic)
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        // (readfirstlane: a branch the compiler sees as uniform, so y and the plane resources stay scalar)
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        {
            float p9[RZ], p13[RZ], p15[RZ], p10[RZ], p12[RZ], p17[RZ], p11[RZ], p14[RZ], p16[RZ];
            unsigned pox[RZ], poy[RZ], poz[RZ];
            int zir[RZ];
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int z = q.z0 + wz + r * BLOCK_Y;
                const bool in = z < L.NZ && xin;
                zir[r] = s_zt.zi[wz + r * BLOCK_Y];
                p9[r] = p13[r] = p15[r] = p10[r] = p12[r] = p17[r] = p11[r] = p14[r] = p16[r] = 0.f;
                if (fr) {
                    pox[r] = psi_ox(L, W2, xi, y, z, in);
                    poy[r] = psi_oy(L, yi, x, z, in);
                    poz[r] = psi_oz(L, W2, zir[r], x, y, in);
                    p9[r] = bload(pr.x[0], pox[r], 0);
                    p13[r] = bload(pr.x[1], pox[r], 0);
                    p15[r] = bload(pr.x[2], pox[r], 0);
                    p10[r] = bload(pr.y[0], poy[r], 0);
                    p12[r] = bload(pr.y[1], poy[r], 0);
                    p17[r] = bload(pr.y[2], poy[r], 0);
                    p11[r] = bload(pr.z[0], poz[r], 0);
                    p14[r] = bload(pr.z[1], poz[r], 0);
                    p16[r] = bload(pr.z[2], poz[r], 0);
                }
            }
            // (also after the last plane: plane ye and its queue plane ye + H lie inside the padded
            // volume, and a branch here would make the compiler wait for every load at the join)
            PlaneIn nxt;
            load_plane(nxt, rvx, rvy, rvz, y + 1);

            const float *cyb = s_yt, *cyf = s_yt + K;
            const bool yf = yi >= 0;
            const float ay = yf ? s_yt[2 * K + 1] : 0.f, by = yf ? s_yt[2 * K + 2] : 0.f;
            const float ayh = yf ? s_yt[2 * K + 3] : 0.f, byh = yf ? s_yt[2 * K + 4] : 0.f;
            const bool surface = C.fsH && y == 0;
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = wz + r * BLOCK_Y;
                const int z = q.z0 + lz;
                if (z >= L.NZ || !xin) continue;
                const float *czb = s_zt.t[0][lz], *czf = s_zt.t[1][lz];
                float vxx = ddx<H>(s_vx, lz + H, tx, cxb);
                float vyy = ddy<H, RZ>(qvy, r, cyb);
                float vzz = ddz<H>(s_vz, lz, tx + H, czb);
                float vxy = ddy<H, RZ>(qvx, r, cyf);
                float vyx = ddx<H>(s_vy, lz + H, tx, cxf);
                float vxz = ddz<H>(s_vx, lz, tx + H, czf);
                float vzx = ddx<H>(s_vz, lz + H, tx, cxf);
                float vyz = ddz<H>(s_vy, lz, tx + H, czf);
                float vzy = ddy<H, RZ>(qvz, r, cyf);
                if (fr) {
                    const float az = s_zt.c[lz][0], bz = s_zt.c[lz][1], azh = s_zt.c[lz][2], bzh = s_zt.c[lz][3];
                    // outside a frame a = b = 0 and psi = 0: the derivative is unchanged
                    bstore(cpml(vxx, p9[r], ax, bx), pr.x[0], pox[r], 0);
                    bstore(cpml(vyx, p13[r], axh, bxh), pr.x[1], pox[r], 0);
                    bstore(cpml(vzx, p15[r], axh, bxh), pr.x[2], pox[r], 0);
                    bstore(cpml(vyy, p10[r], ay, by), pr.y[0], poy[r], 0);
                    bstore(cpml(vxy, p12[r], ayh, byh), pr.y[1], poy[r], 0);
                    bstore(cpml(vzy, p17[r], ayh, byh), pr.y[2], poy[r], 0);
                    bstore(cpml(vzz, p11[r], az, bz), pr.z[0], poz[r], 0);
                    bstore(cpml(vxz, p14[r], azh, bzh), pr.z[1], poz[r], 0);
                    bstore(cpml(vyz, p16[r], azh, bzh), pr.z[2], poz[r], 0);
                }
                if (REC) {
                    bstore(vxx, plane_rsrc(R.r[3] + so, L, y), oc[r], 0);
                    bstore(vyy, plane_rsrc(R.r[4] + so, L, y), oc[r], 0);
                    bstore(vzz, plane_rsrc(R.r[5] + so, L, y), oc[r], 0);
                    bstore(vxy + vyx, plane_rsrc(R.r[6] + so, L, y), oc[r], 0);
                    bstore(vxz + vzx, plane_rsrc(R.r[7] + so, L, y), oc[r], 0);
                    bstore(vyz + vzy, plane_rsrc(R.r[8] + so, L, y), oc[r], 0);
                }
                if constexpr (VE) {
                    float sv[6] = {cur.o[0][r], cur.o[1][r], cur.o[2][r], cur.o[3][r], cur.o[4][r], cur.o[5][r]};
                    const long long fo = (long long)(z + L.H) * L.pitch + x + XPAD;
                    ve3d_cell(V, L, so, y, oc[r], sv, cur.c[0][r], cur.c[1][r], cur.c[2][r], cur.c[3][r], cur.c[4][r],
                              vxx, vyy, vzz, vxy + vyx, vxz + vzx, vyz + vzy, C.fsH, C.fsV, fo);
                    const float d = cur.c[5][r];
                    bstore(sv[0] * d, plane_rsrc(sxx, L, y), oc[r], 0);
                    bstore(sv[1] * d, plane_rsrc(syy, L, y), oc[r], 0);
                    bstore(sv[2] * d, plane_rsrc(szz, L, y), oc[r], 0);
                    bstore(sv[3] * d, plane_rsrc(syz, L, y), oc[r], 0);
                    bstore(sv[4] * d, plane_rsrc(sxz, L, y), oc[r], 0);
                    bstore(sv[5] * d, plane_rsrc(sxy, L, y), oc[r], 0);
                    continue;
                }
                // normal stresses (ForwardSolver3Delastic.cpp: M (vxx+vyy+vzz) - 2 mu (...))
                float s = vxx + vyy;
                s += vzz;
                s *= cur.c[0][r];
                float nxx = cur.o[0][r] + s, nyy = cur.o[1][r] + s, nzz = cur.o[2][r] + s;
                float m = vyy + vzz;
                m *= cur.c[1][r];
                nxx -= 2.f * m;
                m = vxx + vzz;
                m *= cur.c[1][r];
                nyy -= 2.f * m;
                m = vxx + vyy;
                m *= cur.c[1][r];
                nzz -= 2.f * m;
                float sh = vxy + vyx;
                sh *= cur.c[2][r];
                const float nxy = cur.o[5][r] + sh;
                sh = vxz + vzx;
                sh *= cur.c[3][r];
                const float nxz = cur.o[4][r] + sh;
                sh = vyz + vzy;
                sh *= cur.c[4][r];
                const float nyz = cur.o[3][r] + sh;
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
                const float d = cur.c[5][r];
                bstore(nxx * d, plane_rsrc(sxx, L, y), oc[r], 0);
                bstore(nyy * d, plane_rsrc(syy, L, y), oc[r], 0);
                bstore(nzz * d, plane_rsrc(szz, L, y), oc[r], 0);
                bstore(nxy * d, plane_rsrc(sxy, L, y), oc[r], 0);
                bstore(nxz * d, plane_rsrc(sxz, L, y), oc[r], 0);
                bstore(nyz * d, plane_rsrc(syz, L, y), oc[r], 0);
            }
            cur = nxt;
        }
        block_sync_lds();  // the next plane rewrites the LDS regions
    }
    acquisition3d<ACQ_ELASTIC>(A, F, so, q.sb, A.tbase[0] + tofs, false);
}

// ---------------------------------------------------------------------------
// 3D acoustic (ForwardSolver3Dacoustic.cpp:95-229).  Fields vx vy vz p (0..3),
// coefficients 1/rho_x 1/rho_y 1/rho_z M (Coef3D irx iry irz M), CPML memory
// variables 0 p_x, 1 p_y, 2 p_z (half-grid coefficients), 3 vx_x, 4 vy_y,
// 5 vz_z (full).  Same tiles and y march as the elastic kernels; a plane's
// loads are issued together before the plane's arithmetic.
// ---------------------------------------------------------------------------
template <int H, int RZ>
__global__ __launch_bounds__(NTH) void ac3d_velocity(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                     int nyc, int LY, int tofs, int pend)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RXZ = Region<H, H, TZ>;
    __shared__ float s_p[RXZ::RHP][RXZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (pend) acquisition3d<ACQ_ACOUSTIC>(A, F, so, q.sb, A.tbase[0] + tofs - 1, true);
    const float *pr = F.f[3] + so;
    float *const vv[3] = {F.f[0] + so, F.f[1] + so, F.f[2] + so};
    const float *const cr[3] = {C.irx, C.iry, C.irz};
    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxf[K];
#pragma unroll
    for (int k = 0; k < K; k++) cxf[k] = T.xf[(long long)x * K + k];
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    const float axh = xi >= 0 ? P.axh[x] : 0.f, bxh = xi >= 0 ? P.bxh[x] : 0.f;
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, H, TZ> ro;
    ro.init(L, q.x0, q.z0, tid);
    float qp[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b = plane_rsrc(pr, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) qp[k + 1][r] = bload(b, oc[r], 0);
    }
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);
    PsiRes3D ps;  // x frame: 0 p_x | y frame: 1 p_y | z frame: 2 p_z
    ps.init(P, L, q.shot, 0, 0, 0, 1, 1, 1, 2, 2, 2);
    const int W2 = 2 * P.W;
    RXZ rp;
    for (int y = q.ys; y < q.ye; y++) {
        const float yt = ytab_elem<K>(T.yv, y, tid);
        load_region(rp, ro, pr, L, y);
        float np[RZ], o[3][RZ], c[3][RZ];
        {
            const auto b = plane_rsrc(pr, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) np[r] = bload(b, oc[r], 0);
        }
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const auto b = plane_rsrc(vv[f], L, y), bc = plane_rsrc(cr[f], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                o[f][r] = bload(b, oc[r], 0);
                c[f][r] = bload(bc, oc[r], 0);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * H; k++)
#pragma unroll
            for (int r = 0; r < RZ; r++) qp[k][r] = qp[k + 1][r];
#pragma unroll
        for (int r = 0; r < RZ; r++) qp[2 * H][r] = np[r];
        rp.store_padded(s_p, tid);
        s_yt[tid] = yt;
        block_sync_lds();
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        const float *cyf = s_yt + K;
        const float ayh = yi >= 0 ? s_yt[2 * K + 3] : 0.f, byh = yi >= 0 ? s_yt[2 * K + 4] : 0.f;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            float u = ddx<H>(s_p, lz + H, tx, cxf);
            float t = ddy<H, RZ>(qp, r, cyf);
            float w = ddz<H>(s_p, lz, tx + H, s_zt.t[1][lz]);
            if (fr) {
                const unsigned ox = psi_ox(L, W2, xi, y, z, true), oy = psi_oy(L, yi, x, z, true),
                               oz = psi_oz(L, W2, s_zt.zi[lz], x, y, true);
                bstore(cpml(u, bload(ps.x[0], ox, 0), axh, bxh), ps.x[0], ox, 0);
                bstore(cpml(t, bload(ps.y[0], oy, 0), ayh, byh), ps.y[0], oy, 0);
                bstore(cpml(w, bload(ps.z[0], oz, 0), s_zt.c[lz][2], s_zt.c[lz][3]), ps.z[0], oz, 0);
            }
            u *= c[0][r];
            bstore(o[0][r] + u, plane_rsrc(vv[0], L, y), oc[r], 0);
            t *= c[1][r];
            bstore(o[1][r] + t, plane_rsrc(vv[1], L, y), oc[r], 0);
            w *= c[2][r];
            bstore(o[2][r] + w, plane_rsrc(vv[2], L, y), oc[r], 0);
        }
        block_sync_lds();
    }
}

template <int H, int RZ>
__global__ __launch_bounds__(NTH) void ac3d_pressure(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                                     int nyc, int LY, int tofs)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RX = Region<H, 0, TZ>;
    using RZZ = Region<0, H, TZ>;
    __shared__ float s_vx[RX::RHP][RX::RW];
    __shared__ float s_vz[RZZ::RHP][RZZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const float *vx = F.f[0] + so, *vy = F.f[1] + so, *vz = F.f[2] + so;
    float *pr = F.f[3] + so;
    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) cxb[k] = T.xb[(long long)x * K + k];
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    const float ax = xi >= 0 ? P.ax[x] : 0.f, bx = xi >= 0 ? P.bx[x] : 0.f;
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, 0, TZ> ox_;
    RegionOff<0, H, TZ> oz_;
    ox_.init(L, q.x0, q.z0, tid);
    oz_.init(L, q.x0, q.z0, tid);
    float qv[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b = plane_rsrc(vy, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) qv[k + 1][r] = bload(b, oc[r], 0);
    }
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);
    PsiRes3D ps;  // x frame: 3 vx_x | y frame: 4 vy_y | z frame: 5 vz_z
    ps.init(P, L, q.shot, 3, 3, 3, 4, 4, 4, 5, 5, 5);
    const int W2 = 2 * P.W;
    const bool surface_plane = C.fs != 0;
    RX rvx;
    RZZ rvz;
    for (int y = q.ys; y < q.ye; y++) {
        const float yt = ytab_elem<K>(T.ys, y, tid);
        load_region(rvx, ox_, vx, L, y);
        load_region(rvz, oz_, vz, L, y);
        float nv[RZ], op[RZ], cm[RZ], dmp[RZ];
        {
            const auto b = plane_rsrc(vy, L, y + H), bp = plane_rsrc(pr, L, y), bm = plane_rsrc(C.M, L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                nv[r] = bload(b, oc[r], 0);
                op[r] = bload(bp, oc[r], 0);
                cm[r] = bload(bm, oc[r], 0);
                dmp[r] = 1.f;
            }
            if (C.damp) {
                const auto bd = plane_rsrc(C.damp, L, y);
#pragma unroll
                for (int r = 0; r < RZ; r++) dmp[r] = bload(bd, oc[r], 0);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * H; k++)
#pragma unroll
            for (int r = 0; r < RZ; r++) qv[k][r] = qv[k + 1][r];
#pragma unroll
        for (int r = 0; r < RZ; r++) qv[2 * H][r] = nv[r];
        rvx.store_padded(s_vx, tid);
        rvz.store_padded(s_vz, tid);
        s_yt[tid] = yt;
        block_sync_lds();
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        const float *cyb = s_yt;
        const float ay = yi >= 0 ? s_yt[2 * K + 1] : 0.f, by = yi >= 0 ? s_yt[2 * K + 2] : 0.f;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            float u = ddx<H>(s_vx, lz, tx, cxb);
            float t = ddy<H, RZ>(qv, r, cyb);
            float w = ddz<H>(s_vz, lz, tx, s_zt.t[0][lz]);
            if (fr) {
                const unsigned ox = psi_ox(L, W2, xi, y, z, true), oy = psi_oy(L, yi, x, z, true),
                               oz = psi_oz(L, W2, s_zt.zi[lz], x, y, true);
                bstore(cpml(u, bload(ps.x[0], ox, 0), ax, bx), ps.x[0], ox, 0);
                bstore(cpml(t, bload(ps.y[0], oy, 0), ay, by), ps.y[0], oy, 0);
                bstore(cpml(w, bload(ps.z[0], oz, 0), s_zt.c[lz][0], s_zt.c[lz][1]), ps.z[0], oz, 0);
            }
            u += t;
            u += w;
            u *= cm[r];
            float p = op[r] + u;
            p *= dmp[r];
            if (surface_plane && y == 0) p *= 0.f;  // FreeSurface: setSurfaceZero(p)
            bstore(p, plane_rsrc(pr, L, y), oc[r], 0);
        }
        block_sync_lds();
    }
    acquisition3d<ACQ_ACOUSTIC>(A, F, so, q.sb, A.tbase[0] + tofs, false);
}

// ---------------------------------------------------------------------------
// 3D EM (ForwardSolver3Demem.cpp:117-235).  Fields hx hy hz ex ey ez (0..5),
// coefficients 1/mu_yz 1/mu_xz 1/mu_xy (Coef3D irx iry irz) and the Ca, Cb of the
// x-, y-, z-averaged permittivity and conductivity (Coef3D em).  CPML memory variables
// (CPMLEM3D.cpp:25-104): E derivatives on the half profiles 0 ezy 1 eyz 2 exz 3 ezx 4 eyx
// 5 exy; H derivatives on the full ones 6 hzy 7 hyz 8 hxz 9 hzx 11 hxy, except 10 hyx,
// which the reference gives the half x profile (CPMLEM3D.cpp:67-70).  Same tiles and y march
// as the elastic kernels: x-z regions of the three fields read in LDS, y queues of the two
// fields with a y derivative.
// ---------------------------------------------------------------------------
template <int H, int RZ>
__global__ __launch_bounds__(NTH) void em3d_h(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                              int nyc, int LY, int tofs, int pend)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RXZ = Region<H, H, TZ>;
    __shared__ float s_ex[RXZ::RHP][RXZ::RW];
    __shared__ float s_ey[RXZ::RHP][RXZ::RW];
    __shared__ float s_ez[RXZ::RHP][RXZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    // the HZ sources / receivers of the previous step (applied after its E update read hz)
    if (pend) acquisition3d<ACQ_EM>(A, F, so, q.sb, A.tbase[0] + tofs - 1, true);
    float *const hh[3] = {F.f[0] + so, F.f[1] + so, F.f[2] + so};
    const float *ex = F.f[3] + so, *ey = F.f[4] + so, *ez = F.f[5] + so;
    const float *const cm[3] = {C.irx, C.iry, C.irz};
    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxf[K];
#pragma unroll
    for (int k = 0; k < K; k++) cxf[k] = T.xf[(long long)x * K + k];
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    const float axh = xi >= 0 ? P.axh[x] : 0.f, bxh = xi >= 0 ? P.bxh[x] : 0.f;
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, H, TZ> ro;
    ro.init(L, q.x0, q.z0, tid);
    // y queues: ex (exy), ez (ezy)
    float qx[K][RZ], qz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b0 = plane_rsrc(ex, L, q.ys - H + k), b1 = plane_rsrc(ez, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qx[k + 1][r] = bload(b0, oc[r], 0);
            qz[k + 1][r] = bload(b1, oc[r], 0);
        }
    }
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);
    PsiRes3D ps;  // x frame: 3 ezx, 4 eyx | y frame: 0 ezy, 5 exy | z frame: 1 eyz, 2 exz
    ps.init(P, L, q.shot, 3, 4, 4, 0, 5, 5, 1, 2, 2);
    const int W2 = 2 * P.W;
    RXZ rx, ry, rzr;
    for (int y = q.ys; y < q.ye; y++) {
        const float yt = ytab_elem<K>(T.ys, y, tid);
        load_region(rx, ro, ex, L, y);
        load_region(ry, ro, ey, L, y);
        load_region(rzr, ro, ez, L, y);
        float nx[RZ], nz[RZ], o[3][RZ], c[3][RZ];
        {
            const auto b0 = plane_rsrc(ex, L, y + H), b1 = plane_rsrc(ez, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                nx[r] = bload(b0, oc[r], 0);
                nz[r] = bload(b1, oc[r], 0);
            }
        }
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const auto b = plane_rsrc(hh[f], L, y), bc = plane_rsrc(cm[f], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                o[f][r] = bload(b, oc[r], 0);
                c[f][r] = bload(bc, oc[r], 0);
            }
        }
        shift<H, RZ>(qx);
        shift<H, RZ>(qz);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qx[2 * H][r] = nx[r];
            qz[2 * H][r] = nz[r];
        }
        rx.store_padded(s_ex, tid);
        ry.store_padded(s_ey, tid);
        rzr.store_padded(s_ez, tid);
        s_yt[tid] = yt;
        block_sync_lds();
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        const float *cyf = s_yt + K;
        const float ayh = yi >= 0 ? s_yt[2 * K + 3] : 0.f, byh = yi >= 0 ? s_yt[2 * K + 4] : 0.f;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            const float *czf = s_zt.t[1][lz];
            float ezy = ddy<H, RZ>(qz, r, cyf);
            float exy = ddy<H, RZ>(qx, r, cyf);
            float ezx = ddx<H>(s_ez, lz + H, tx, cxf);
            float eyx = ddx<H>(s_ey, lz + H, tx, cxf);
            float eyz = ddz<H>(s_ey, lz, tx + H, czf);
            float exz = ddz<H>(s_ex, lz, tx + H, czf);
            if (fr) {
                const unsigned ox = psi_ox(L, W2, xi, y, z, true), oy = psi_oy(L, yi, x, z, true),
                               oz = psi_oz(L, W2, s_zt.zi[lz], x, y, true);
                const float azh = s_zt.c[lz][2], bzh = s_zt.c[lz][3];
                bstore(cpml(ezx, bload(ps.x[0], ox, 0), axh, bxh), ps.x[0], ox, 0);
                bstore(cpml(eyx, bload(ps.x[1], ox, 0), axh, bxh), ps.x[1], ox, 0);
                bstore(cpml(ezy, bload(ps.y[0], oy, 0), ayh, byh), ps.y[0], oy, 0);
                bstore(cpml(exy, bload(ps.y[1], oy, 0), ayh, byh), ps.y[1], oy, 0);
                bstore(cpml(eyz, bload(ps.z[0], oz, 0), azh, bzh), ps.z[0], oz, 0);
                bstore(cpml(exz, bload(ps.z[1], oz, 0), azh, bzh), ps.z[1], oz, 0);
            }
            // hx -= (ezy - eyz) / mu_yz ; hy -= (exz - ezx) / mu_xz ; hz -= (eyx - exy) / mu_xy
            float u = ezy - eyz;
            u *= c[0][r];
            bstore(o[0][r] - u, plane_rsrc(hh[0], L, y), oc[r], 0);
            u = exz - ezx;
            u *= c[1][r];
            bstore(o[1][r] - u, plane_rsrc(hh[1], L, y), oc[r], 0);
            u = eyx - exy;
            u *= c[2][r];
            bstore(o[2][r] - u, plane_rsrc(hh[2], L, y), oc[r], 0);
        }
        block_sync_lds();
    }
}

// E update: ex = Ca_x ex + Cb_x ((hzy - hyz) [- sum_l DT rx_l]), ey from (hxz - hzx), ez from
// (hyx - hxy); VE: first r_l = Cd_l e + Cc_l r_l with the e of the previous step
// (ForwardSolver3Dviscoemem.cpp:238-300); then the ABS damping of E and the EZ, EX, EY
// sources / receivers of the step.
template <int H, int RZ, bool VE>
__global__ __launch_bounds__(NTH) void em3d_e(Layout3D L, Fields3D F, Coef3D C, Taps3D T, Cpml3D P, Acq A,
                                              int nyc, int LY, int tofs, Visco3D V)
{
    constexpr int TZ = RZ * BLOCK_Y, K = 2 * H + 1;
    using RXZ = Region<H, H, TZ>;
    __shared__ float s_hx[RXZ::RHP][RXZ::RW];
    __shared__ float s_hy[RXZ::RHP][RXZ::RW];
    __shared__ float s_hz[RXZ::RHP][RXZ::RW];
    __shared__ ZTab<TZ, K> s_zt;
    __shared__ float s_yt[NTH];
    const Pos q = position(L, TZ, nyc, LY, A);
    const long long so = (long long)q.shot * L.vol;
    const int tid = threadIdx.y * TILE_X + threadIdx.x, tx = threadIdx.x;
    const int wz = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const float *hx = F.f[0] + so, *hy = F.f[1] + so, *hz = F.f[2] + so;
    float *const ee[3] = {F.f[3] + so, F.f[4] + so, F.f[5] + so};
    const int x = q.x0 + tx;
    const bool xin = x < L.NX;
    float cxb[K];
#pragma unroll
    for (int k = 0; k < K; k++) cxb[k] = T.xb[(long long)x * K + k];
    const int xi = (P.on && xin) ? P.xi[x] : -1;
    float ax = 0.f, bx = 0.f, axh = 0.f, bxh = 0.f;
    if (xi >= 0) {
        ax = P.ax[x];
        bx = P.bx[x];
        axh = P.axh[x];
        bxh = P.bxh[x];
    }
    unsigned oc[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) oc[r] = (unsigned)(((q.z0 + wz + r * BLOCK_Y + L.H) * L.pitch + x + XPAD) * 4);
    RegionOff<H, H, TZ> ro;
    ro.init(L, q.x0, q.z0, tid);
    // y queues: hx (hxy), hz (hzy)
    float qx[K][RZ], qz[K][RZ];
#pragma unroll
    for (int k = 0; k < 2 * H; k++) {
        const auto b0 = plane_rsrc(hx, L, q.ys - H + k), b1 = plane_rsrc(hz, L, q.ys - H + k);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qx[k + 1][r] = bload(b0, oc[r], 0);
            qz[k + 1][r] = bload(b1, oc[r], 0);
        }
    }
    const bool tfr = tile_in_frame(L, P, q.x0, q.z0, TZ);
    s_zt.fill(T, P, L, q.z0, tid);
    PsiRes3D ps;  // x frame: 9 hzx, 10 hyx | y frame: 6 hzy, 11 hxy | z frame: 7 hyz, 8 hxz
    ps.init(P, L, q.shot, 9, 10, 10, 6, 11, 11, 7, 8, 8);
    const int W2 = 2 * P.W;
    RXZ rx, ry, rzr;
    for (int y = q.ys; y < q.ye; y++) {
        // the y operators of the E update are the backward ones (Dyb in the first K entries)
        const float yt = ytab_elem<K>(T.ys, y, tid);
        load_region(rx, ro, hx, L, y);
        load_region(ry, ro, hy, L, y);
        load_region(rzr, ro, hz, L, y);
        float nx[RZ], nz[RZ], o[3][RZ], ca[3][RZ], cb[3][RZ], d[RZ];
        {
            const auto b0 = plane_rsrc(hx, L, y + H), b1 = plane_rsrc(hz, L, y + H);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                nx[r] = bload(b0, oc[r], 0);
                nz[r] = bload(b1, oc[r], 0);
            }
        }
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const auto b = plane_rsrc(ee[f], L, y), ba = plane_rsrc(C.em[2 * f], L, y),
                       bb = plane_rsrc(C.em[2 * f + 1], L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                o[f][r] = bload(b, oc[r], 0);
                ca[f][r] = bload(ba, oc[r], 0);
                cb[f][r] = bload(bb, oc[r], 0);
            }
        }
        if (C.damp) {
            const auto b = plane_rsrc(C.damp, L, y);
#pragma unroll
            for (int r = 0; r < RZ; r++) d[r] = bload(b, oc[r], 0);
        } else {
#pragma unroll
            for (int r = 0; r < RZ; r++) d[r] = 1.f;
        }
        shift<H, RZ>(qx);
        shift<H, RZ>(qz);
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            qx[2 * H][r] = nx[r];
            qz[2 * H][r] = nz[r];
        }
        rx.store_padded(s_hx, tid);
        ry.store_padded(s_hy, tid);
        rzr.store_padded(s_hz, tid);
        s_yt[tid] = yt;
        block_sync_lds();
        const int yi = __builtin_amdgcn_readfirstlane(P.on ? __float_as_int(s_yt[2 * K]) : -1);
        const bool fr = __builtin_amdgcn_readfirstlane((int)(tfr || yi >= 0)) != 0;
        const float *cyb = s_yt;
        const float ay = yi >= 0 ? s_yt[2 * K + 1] : 0.f, by = yi >= 0 ? s_yt[2 * K + 2] : 0.f;
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int lz = wz + r * BLOCK_Y;
            const int z = q.z0 + lz;
            if (z >= L.NZ || !xin) continue;
            const float *czb = s_zt.t[0][lz];
            float hzy = ddy<H, RZ>(qz, r, cyb);
            float hxy = ddy<H, RZ>(qx, r, cyb);
            float hzx = ddx<H>(s_hz, lz + H, tx, cxb);
            float hyx = ddx<H>(s_hy, lz + H, tx, cxb);
            float hyz = ddz<H>(s_hy, lz, tx + H, czb);
            float hxz = ddz<H>(s_hx, lz, tx + H, czb);
            if (fr) {
                const unsigned ox = psi_ox(L, W2, xi, y, z, true), oy = psi_oy(L, yi, x, z, true),
                               oz = psi_oz(L, W2, s_zt.zi[lz], x, y, true);
                const float az = s_zt.c[lz][0], bz = s_zt.c[lz][1];
                bstore(cpml(hzx, bload(ps.x[0], ox, 0), ax, bx), ps.x[0], ox, 0);
                bstore(cpml(hyx, bload(ps.x[1], ox, 0), axh, bxh), ps.x[1], ox, 0);
                bstore(cpml(hzy, bload(ps.y[0], oy, 0), ay, by), ps.y[0], oy, 0);
                bstore(cpml(hxy, bload(ps.y[1], oy, 0), ay, by), ps.y[1], oy, 0);
                bstore(cpml(hyz, bload(ps.z[0], oz, 0), az, bz), ps.z[0], oz, 0);
                bstore(cpml(hxz, bload(ps.z[1], oz, 0), az, bz), ps.z[1], oz, 0);
            }
            float u[3] = {hzy - hyz, hxz - hzx, hyx - hxy};
            if constexpr (VE) {
#pragma unroll
                for (int f = 0; f < 3; f++)
#pragma unroll
                    for (int l = 0; l < 4; l++)
                        if (l < V.L) {
                            const auto br = plane_rsrc(V.R[4 * f + l] + so, L, y);
                            const float a = V.cc[l] * bload(br, oc[r], 0);
                            const float b = bload(plane_rsrc(V.cd[4 * f + l], L, y), oc[r], 0) * o[f][r];
                            const float rn = b + a;
                            bstore(rn, br, oc[r], 0);
                            u[f] -= V.dt * rn;
                        }
            }
#pragma unroll
            for (int f = 0; f < 3; f++) {
                u[f] *= cb[f][r];
                float e = ca[f][r] * o[f][r] + u[f];
                bstore(e * d[r], plane_rsrc(ee[f], L, y), oc[r], 0);
            }
        }
        block_sync_lds();
    }
    acquisition3d<ACQ_EM>(A, F, so, q.sb, A.tbase[0] + tofs, false);
}

template <int MODE>
__global__ __launch_bounds__(256) void flush3d(Layout3D L, Fields3D F, Acq A, int tofs)
{
    const int shot = blockIdx.x;
    const long long so = (long long)shot * L.vol;
    const int t = A.tbase[0] + tofs;
    if (t < 0 || t >= A.NT) return;
    if (threadIdx.x == 0) {
        const int s0 = A.src_ptr[shot], s1 = A.src_ptr[shot + 1];
        for (int ty = 2; ty <= 4; ty++) {
            if (!velocity_type<MODE>(ty)) continue;
            for (int s = s0; s < s1; s++) {
                if (A.src_type[s] != ty) continue;
                F.f[type_field<MODE>(ty)][so + A.src_cell[s]] += A.src_sig[(long long)s * A.NT + t];
            }
        }
    }
    __syncthreads();
    const int r0 = A.rec_ptr[shot], r1 = A.rec_ptr[shot + 1];
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int ty = A.rec_type[r];
        if (!velocity_type<MODE>(ty)) continue;
        A.seis[(long long)t * A.rec_total + r] = F.f[type_field<MODE>(ty)][so + A.rec_cell[r]];
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
               hipEvent_t e1, const Rec3D *R, const Visco3D *V)
{
    constexpr int TZ = RZ * BLOCK_Y;
    const int LY = (L.NY + nyc - 1) / nyc;
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NZ + TZ - 1) / TZ, nyc * nshots), block(TILE_X, BLOCK_Y);
    const Rec3D r = R ? *R : Rec3D{};
    const Visco3D v = V ? *V : Visco3D{};
    if (which == 0) {
        if (R)
            go(el3d_velocity<H, RZ, true>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, pend, r);
        else
            go(el3d_velocity<H, RZ, false>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, pend, r);
    } else {
        if (V)
            go(el3d_stress<H, RZ, false, true>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, r, v);
        else if (R)
            go(el3d_stress<H, RZ, true>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, r, v);
        else
            go(el3d_stress<H, RZ, false>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, r, v);
    }
}

template <int H>
void launch_rz(int rz, int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
               const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
               hipEvent_t e1, const Rec3D *R, const Visco3D *V)
{
    if (rz == 1)
        launch_h3<H, 1>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, R, V);
    else if (rz == 4)
        launch_h3<H, 4>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, R, V);
    else
        launch_h3<H, 2>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, R, V);
}

}  // namespace

void launch_el3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0, hipEvent_t e1, const Rec3D *R, const Visco3D *V)
{
    switch (order / 2) {
    case 1: launch_rz<1>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    case 2: launch_rz<2>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    case 3: launch_rz<3>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    case 4: launch_rz<4>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    case 5: launch_rz<5>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    case 6: launch_rz<6>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, R, V); break;
    default: break;
    }
}

namespace {
template <int H, int RZ>
void launch_ac3d_h(int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
                   const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                   hipEvent_t e1)
{
    constexpr int TZ = RZ * BLOCK_Y;
    const int LY = (L.NY + nyc - 1) / nyc;
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NZ + TZ - 1) / TZ, nyc * nshots), block(TILE_X, BLOCK_Y);
    if (which == 0)
        go(ac3d_velocity<H, RZ>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, pend);
    else
        go(ac3d_pressure<H, RZ>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs);
}
template <int H>
void launch_ac3d_rz(int rz, int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
                    const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1)
{
    if (rz == 1)
        launch_ac3d_h<H, 1>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
    else if (rz == 4)
        launch_ac3d_h<H, 4>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
    else
        launch_ac3d_h<H, 2>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1);
}
}  // namespace

void launch_ac3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    switch (order / 2) {
    case 1: launch_ac3d_rz<1>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 2: launch_ac3d_rz<2>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 3: launch_ac3d_rz<3>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 4: launch_ac3d_rz<4>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 5: launch_ac3d_rz<5>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    case 6: launch_ac3d_rz<6>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1); break;
    default: break;
    }
}

namespace {
template <int H, int RZ>
void launch_em3d_h(int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
                   const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                   hipEvent_t e1, const Visco3D *V)
{
    constexpr int TZ = RZ * BLOCK_Y;
    const int LY = (L.NY + nyc - 1) / nyc;
    dim3 grid((L.NX + TILE_X - 1) / TILE_X, (L.NZ + TZ - 1) / TZ, nyc * nshots), block(TILE_X, BLOCK_Y);
    if (which == 0)
        go(em3d_h<H, RZ>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, pend);
    else if (V)
        go(em3d_e<H, RZ, true>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, *V);
    else
        go(em3d_e<H, RZ, false>, grid, block, st, e0, e1, L, F, C, T, P, A, nyc, LY, tofs, Visco3D{});
}
template <int H>
void launch_em3d_rz(int rz, int which, const Layout3D &L, int nshots, int nyc, const Fields3D &F, const Coef3D &C,
                    const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1, const Visco3D *V)
{
    if (rz == 1)
        launch_em3d_h<H, 1>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, V);
    else if (rz == 4)
        launch_em3d_h<H, 4>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, V);
    else
        launch_em3d_h<H, 2>(which, L, nshots, nyc, F, C, T, P, A, tofs, pend, st, e0, e1, V);
}
}  // namespace

void launch_em3d(int which, int order, int rz, const Layout3D &L, int nshots, int nyc, const Fields3D &F,
                 const Coef3D &C, const Taps3D &T, const Cpml3D &P, const Acq &A, int tofs, int pend,
                 hipStream_t stream, hipEvent_t e0, hipEvent_t e1, const Visco3D *V)
{
    switch (order / 2) {
    case 1: launch_em3d_rz<1>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    case 2: launch_em3d_rz<2>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    case 3: launch_em3d_rz<3>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    case 4: launch_em3d_rz<4>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    case 5: launch_em3d_rz<5>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    case 6: launch_em3d_rz<6>(rz, which, L, nshots, nyc, F, C, T, P, A, tofs, pend, stream, e0, e1, V); break;
    default: break;
    }
}

void launch_abs_velocity3d(const Layout3D &L, int nshots, const Fields3D &F, const Coef3D &C, hipStream_t stream)
{
    const long long n = (long long)nshots * L.vol;
    abs_velocity3d<<<2048, 256, 0, stream>>>(n, F.f[0], F.f[1], F.f[2], C.damp, L.vol);
}

void launch_flush3d(const Layout3D &L, const Fields3D &F, const Acq &A, int tofs, hipStream_t stream, bool em)
{
    if (em)
        flush3d<ACQ_EM><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
    else
        flush3d<ACQ_ELASTIC><<<A.nshots, 256, 0, stream>>>(L, F, A, tofs);
}

}  // namespace wv
