// morton.hpp -- 30-bit Morton codes of points of the unit sphere (10 bits
// per axis over [-1.1, 1.1]^3), shared by the LBVH build and the spatial row
// order of the apply.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace axis_hip {

__device__ inline std::uint32_t spread10(std::uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ inline std::uint32_t morton30(float x, float y, float z) {
    const float c[3] = {x, y, z};
    std::uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float t = (c[a] + 1.1f) * (1.0f / 2.2f);
        t = fminf(fmaxf(t, 0.0f), 1.0f);  // NaN -> 0
        q[a] = static_cast<std::uint32_t>(fminf(t * 1024.0f, 1023.0f));
    }
    return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

}  // namespace axis_hip
