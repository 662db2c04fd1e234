// Tagged 16-byte exchange between workgroups through uncached device memory (allocated with
// hipDeviceMallocUncached): loads and stores go to memory and are coherent across the XCDs
// without cache maintenance. Each 16-byte unit is written and read with one buffer access at
// system scope (sc0 sc1) and carries a tag derived from its contents and a sequence number,
// so a reader recognises stale or partly written data and simply reads again: no
// store-completion waits or separate flags. Buffer accesses, unlike volatile ones, do not wait
// for completion one by one, and the descriptor spans exactly the area in use, so an access
// out of range reads zeros (and fails its tag check) instead of touching other memory.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace elx {
namespace xchg {

typedef uint32_t u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t mix32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}
constexpr int kSysScope = 1 | 16;  // sc0 | sc1
// Workgroups on one XCD share its L2: stores go through the (write-through) vector L1 to L2,
// and loads at device scope (sc1) miss the L1 and read L2 (sc0 alone is workgroup scope and
// may hit a stale line of the reader's own L1). Ordinary device memory, coherent among the
// workgroups of that XCD only.
constexpr int kXcdLoad = 16;  // sc1
constexpr int kXcdStore = 0;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t area(const u4* base, uint32_t count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<u4*>(base), 0, (int)(count * 16u), 0x00020000);
}
template <int AUX = kSysScope>
__device__ __forceinline__ void st16(const u4* base, uint32_t count, uint32_t i, u4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, area(base, count), (int)(i * 16u), 0, AUX);
}
template <int AUX = kSysScope>
__device__ __forceinline__ u4 ld16(const u4* base, uint32_t count, uint32_t i) {
    return __builtin_amdgcn_raw_buffer_load_b128(area(base, count), (int)(i * 16u), 0, AUX);
}

}  // namespace xchg
}  // namespace elx
