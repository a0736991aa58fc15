// Tile helpers shared by the 2D kernels (forward and adjoint): register-staged
// LDS tiles with an H-wide halo, derivative taps, CPML memory update.
#pragma once

#include "buffer.hpp"
#include "wave_internal.hpp"

namespace wv {
namespace {

constexpr int NTHREADS = TILE_X * BLOCK_Y;

template <int H, int TY>
struct Tile {
    static constexpr int RW = TILE_X + 2 * H;
    static constexpr int RH = TY + 2 * H;
    static constexpr int TOT = RW * RH;
    static constexpr int IT = (TOT + NTHREADS - 1) / NTHREADS;
};

// Registers holding one thread's share of the region
// [x0-H, x0+64+H) x [y0-H, y0+TY+H) of one field plane.
template <int H, int TY>
struct TileRegs {
    using T = Tile<H, TY>;
    float v[T::IT];
    __device__ __forceinline__ void load(const float *__restrict__ g, const Layout2D &L, int x0, int y0, int tid)
    {
        const long long base = L.at(x0 - H, y0 - H);
#pragma unroll
        for (int k = 0; k < T::IT; k++) {
            const int i = tid + k * NTHREADS;
            if (T::TOT % NTHREADS == 0 || i < T::TOT) {
                const int r = i / T::RW;
                const int c = i - r * T::RW;
                v[k] = g[base + (long long)r * L.pitch + c];
            }
        }
    }
    __device__ __forceinline__ void store(float (*S)[T::RW], int tid) const
    {
#pragma unroll
        for (int k = 0; k < T::IT; k++) {
            const int i = tid + k * NTHREADS;
            if (T::TOT % NTHREADS == 0 || i < T::TOT) {
                const int r = i / T::RW;
                const int c = i - r * T::RW;
                S[r][c] = v[k];
            }
        }
    }
};

// Rows [y0-H, y0+TY+H) x columns [x0-H, x0+64+H) of one field plane, loaded
// row by row: wavefront wy takes rows wy, wy+4, ...; lane l holds column l of
// the row and lanes < 2H also column 64+l.  Every address is a wavefront-
// uniform row pointer plus the lane, so the loads need no per-element index
// arithmetic.
template <int H, int TY>
struct RowTile {
    static constexpr int RW = TILE_X + 2 * H;
    static constexpr int RH = TY + 2 * H;
    static constexpr int NR = (RH + BLOCK_Y - 1) / BLOCK_Y;
    float a[NR], b[NR];
    // g: element (x0-H, y0-H) of the plane
    __device__ __forceinline__ void load(const float *__restrict__ g, int pitch, int wy, int lane)
    {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int r = wy + k * BLOCK_Y;
            if (RH % BLOCK_Y == 0 || r < RH) {
                const float *__restrict__ rp = g + (long long)r * pitch;
                a[k] = rp[(unsigned)lane];
                if (lane < 2 * H) b[k] = rp[TILE_X + (unsigned)lane];
            }
        }
    }
    __device__ __forceinline__ void store(float (*S)[RW], int wy, int lane) const
    {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int r = wy + k * BLOCK_Y;
            if (RH % BLOCK_Y == 0 || r < RH) {
                S[r][lane] = a[k];
                if (lane < 2 * H) S[r][TILE_X + lane] = b[k];
            }
        }
    }
};

// LDS written by a wavefront and read back by the same wavefront: orders the
// accesses without a workgroup barrier.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// derivative along x at LDS row r, column tx (taps for offsets -H..H)
template <int H, int RW>
__device__ __forceinline__ float dx(const float (*S)[RW], int r, int tx, const float *c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[r][tx + k];
    return s;
}

// derivative along y at LDS row r (center), column tx; c is wavefront-uniform
template <int H, int RW>
__device__ __forceinline__ float dy(const float (*S)[RW], int r, int tx, const float *__restrict__ c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) s += c[k] * S[r + k - H][tx + H];
    return s;
}

// psi = b*psi + a*d ; d += psi   (CPML.cpp:84-95)
__device__ __forceinline__ float cpml(float &d, float psi, float a, float b)
{
    float p = b * psi;
    p = p + a * d;
    d = d + p;
    return p;
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
        // the coefficient tables are 0 outside the frame and padded beyond the grid
        c.xi = frame_index(x, L.NX, P.W, false);
        c.ax = P.ax[x];
        c.bx = P.bx[x];
        c.axh = P.axh[x];
        c.bxh = P.bxh[x];
    }
    return c;
}

template <int H>
__device__ __forceinline__ void load_taps(float *dst, const float *__restrict__ tab, int pos)
{
#pragma unroll
    for (int k = 0; k <= 2 * H; k++) dst[k] = tab[(long long)pos * (2 * H + 1) + k];
}

// Tile and shot of this block.  Workgroups are dealt to the 8 XCDs round
// robin (block b runs on XCD b mod 8); renumbering them so that each XCD gets
// one contiguous range of (tile x, tile y, shot) keeps the halo rows and
// columns a block shares with its neighbours in that XCD's L2.
struct Blk {
    int bx, by, shot;
};

__device__ __forceinline__ Blk xcd_block()
{
    const int nbx = gridDim.x, nby = gridDim.y;
    const int nblk = nbx * nby * gridDim.z;
    const int b = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int xcd = b & 7, k = b >> 3, q = nblk >> 3, rem = nblk & 7;
    const int t = xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
    Blk r;
    r.bx = t % nbx;
    const int u = t / nbx;
    r.by = u % nby;
    r.shot = u / nby;
    return r;
}

}  // namespace
}  // namespace wv
