// Source injection and receiver gather by the block that owns the cells
// (2D tiled kernels).
#pragma once

#include "tile2d.hpp"

namespace wv {
namespace {

// velocity-type components (the field the first phase of a step writes): VX (2) and
// VY (3), VZ (4) in SH, HZ (4) in TE-EM; stress-type: P / EZ (1), EX (2) / EY (3) in TE-EM
__device__ __forceinline__ bool velocity_type(int ty) { return ty == 2 || ty == 3; }
template <int PHYS>
__device__ __forceinline__ bool velocity_type_of(int ty)
{
    return (PHYS == SH || PHYS == TEMEM) ? ty == 4 : velocity_type(ty);
}
// field of a component: vx, vy (fields 0, 1), P / EZ field 2; vz of SH and hz of TE-EM
// field 0, ex / ey of TE-EM fields 1 / 2
template <int PHYS>
__device__ __forceinline__ int velocity_field(int ty)
{
    if (PHYS == TEMEM) return ty == 4 ? 0 : ty - 1;
    return PHYS == SH ? 0 : (ty == 1 ? 2 : ty - 2);
}

// Sources and receivers owned by this block (shot-block sb) at step t, for
// either the velocity-type or the stress-type components.  Uniform per block.
template <int PHYS>
__device__ void block_acquisition(const Acq &A, const Fields2D &F, long long so, int sb, int t, bool vel)
{
    const int s0 = A.bsrc_ptr[sb], s1 = A.bsrc_ptr[sb + 1];
    const int r0 = A.brec_ptr[sb], r1 = A.brec_ptr[sb + 1];
    if (s0 == s1 && r0 == r1) return;
    const int tid = threadIdx.y * TILE_X + threadIdx.x;
    __syncthreads();  // this block's field stores are visible to its threads
    if (tid == 0) {
        // components in the reference's order P, VX, VY, VZ; each in list order
        for (int ty = 1; ty <= 4; ty++) {
            if (velocity_type_of<PHYS>(ty) != vel) continue;
            for (int e = s0; e < s1; e++) {
                const int s = A.bsrc[e];
                if (A.src_type[s] != ty) continue;
                const float v = A.src_sig[(long long)s * A.NT + t];
                const long long c = so + A.src_cell[s];
                if (PHYS == ELASTIC && ty == 1) {
                    F.f[2][c] += v;
                    F.f[3][c] += v;
                } else {
                    F.f[velocity_field<PHYS>(ty)][c] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int e = r0 + tid; e < r1; e += NTHREADS) {
        const int r = A.brec[e];
        const int ty = A.rec_type[r];
        if (velocity_type_of<PHYS>(ty) != vel) continue;
        const long long c = so + A.rec_cell[r];
        float v;
        if (PHYS == ELASTIC && ty == 1) {
            v = F.f[2][c];
            v += F.f[3][c];
            v *= 0.5f;
        } else if (ty == 1) {
            v = F.f[2][c];
        } else {
            v = F.f[velocity_field<PHYS>(ty)][c];
        }
        A.seis[(long long)t * A.rec_total + r] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ int shot_block(const Acq &A, const Blk &B) { return (B.shot * A.nby + B.by) * A.nbx + B.bx; }

}  // namespace
}  // namespace wv
