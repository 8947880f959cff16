// Morton codes (x in bit 0, y in bit 1, z in bit 2 of every 3-bit group), host and device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hfmm {

__host__ __device__ inline uint64_t morton_expand(uint64_t v) {
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ inline uint64_t morton_compact(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v | (v >> 2)) & 0x10C30C30C30C30C3ull;
    v = (v | (v >> 4)) & 0x100F00F00F00F00Full;
    v = (v | (v >> 8)) & 0x1F0000FF0000FFull;
    v = (v | (v >> 16)) & 0x1F00000000FFFFull;
    v = (v | (v >> 32)) & 0x1FFFFFull;
    return v;
}

__host__ __device__ inline uint64_t morton_encode(uint64_t x, uint64_t y, uint64_t z) {
    return morton_expand(x) | (morton_expand(y) << 1) | (morton_expand(z) << 2);
}

__host__ __device__ inline void morton_decode(uint64_t c, int& x, int& y, int& z) {
    x = (int)morton_compact(c);
    y = (int)morton_compact(c >> 1);
    z = (int)morton_compact(c >> 2);
}

// the 26 neighbour offsets in the order the lists are built
__host__ __device__ inline void neighbour_offset(int d, int& dx, int& dy, int& dz) {
    const int e = d < 13 ? d : d + 1;   // skip (0, 0, 0)
    dx = e / 9 - 1;
    dy = (e / 3) % 3 - 1;
    dz = e % 3 - 1;
}

}  // namespace hfmm
