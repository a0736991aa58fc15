// Raw buffer loads and stores, and an LDS-only block barrier, for the gfx950
// kernels (shared by the 2D and 3D kernel files).  Not part of the public ABI.
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

}  // namespace
}  // namespace wv
