// Raw buffer loads and stores for the padded-layout kernels: one 128-bit resource per array
// (built in the kernel), a 32-bit byte offset per lane, and the hardware range check: an
// offset at or past the array's end reads zero and drops a store, which is how the marching
// kernels keep their planes branch-free at the domain edges. Arrays must stay below 2 GiB.
#pragma once
#include <hip/hip_runtime.h>

namespace bufio {

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef float f4v __attribute__((ext_vector_type(4)));

// a byte offset that is past the end of every array these kernels address
constexpr unsigned kOOB = 0x80000000u;
// cache-policy bit "nt" (streaming): for arrays touched once per step
constexpr int kNT = 2;

__device__ __forceinline__ rsrc_t make(const void *p, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float load(rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

__device__ __forceinline__ float load_nt(rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kNT));
}

// (the whole vector is reinterpreted at once: a bit cast of one element of a vector value reads
// element 0 with this compiler)
__device__ __forceinline__ float4 load4(rsrc_t r, unsigned off) {
    const f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void store_nt(float v, rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, kNT);
}

// two consecutive floats (8-byte aligned offset): the two shots of one cell of an interleaved field
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 load2_nt(rsrc_t r, unsigned off) {
    const f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kNT));
    return make_float2(v.x, v.y);
}

__device__ __forceinline__ void store2_nt(float x, float y, rsrc_t r, unsigned off) {
    const f2v v = {x, y};
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), r, off, 0, kNT);
}

}  // namespace bufio
