// Raw buffer loads and stores, an LDS-only block barrier and a wavefront
// maximum, for the gfx950 kernels (shared by the 2D and 3D kernel files).  Not
// part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace wv {
namespace {

// Raw buffer access: a 128-bit resource (base, size in bytes) in scalar
// registers, a per-lane 32-bit byte offset and a wavefront-uniform 32-bit byte
// offset (soffset).  Rows of a plane differ only in soffset, so a load or store
// needs no vector address arithmetic.  Word 3 = 0x00020000 is the gfx9 raw
// 32-bit-element resource format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, soff, 0);
}

// Block barrier that orders LDS accesses only: global loads and stores issued
// before it stay in flight (a full __syncthreads waits for them).
__device__ __forceinline__ void block_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Maximum of a non-negative float over the 64 lanes, returned to every lane, with
// DPP moves only (no LDS round trip): swap neighbours, swap pairs, mirror within 8,
// mirror within 16 (every lane of a row holds the row maximum), then broadcast
// lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3; lane 63 holds the result.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_max_step(float x)
{
    const float y =
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, ROWS, 0xF, false));
    return fmaxf(x, y);
}
__device__ __forceinline__ float wave_max_nonneg(float x)
{
    x = dpp_max_step<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
    x = dpp_max_step<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
    x = dpp_max_step<0x141, 0xF>(x);  // row_half_mirror
    x = dpp_max_step<0x140, 0xF>(x);  // row_mirror
    x = dpp_max_step<0x142, 0xA>(x);  // row_bcast:15 into rows 1, 3
    x = dpp_max_step<0x143, 0xC>(x);  // row_bcast:31 into rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

}  // namespace
}  // namespace wv
