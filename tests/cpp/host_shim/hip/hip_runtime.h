// Host-compilation stand-in for the HIP names the traversal core uses (uint4, __float_as_uint, and
// the C math functions the device headers declare), so that
// core_emulation.cpp can compile cubiquity_amd/csrc/traverse_core.h with g++ for a CPU test.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
struct uint4 {
    unsigned x, y, z, w;
};
inline uint32_t __float_as_uint(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
