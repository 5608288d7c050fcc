// Index helpers of the exact-root link pass (bf_rays.hip), host and
// device: the host instantiation lets a CPU test walk every index the pass
// forms (bf_adjoint_check.cpp, bfa_layout_check).
#pragma once

#include "bf_device.h"

namespace bf {

// The source coordinate the fluid step pulls from along one axis: wrapped
// across a periodic face, clamped at an outlet, none beyond a far field.
__host__ __device__ inline bool ray_source(const int bc[2], int s, int n, int& out) {
    if (s < 0) {
        if (bc[0] == BC_FARFIELD) return false;
        s = bc[0] == BC_PERIODIC ? n - 1 : 0;
    } else if (s >= n) {
        if (bc[1] == BC_FARFIELD) return false;
        s = bc[1] == BC_PERIODIC ? 0 : n - 1;
    }
    out = s;
    return true;
}

// Coordinates x along one axis whose pull along c lands on source coordinate
// s (at most two: the neighbour s + c, wrapped across a periodic face, and
// at an outlet the face cell itself, whose pull is clamped onto itself).
__host__ __device__ inline int pullers(const int bc[2], int s, int c, int n, int out[2]) {
    int k = 0;
    const int cand[4] = {s + c, s + c - n, s + c + n, s};
    for (int i = 0; i < 4; ++i) {
        const int x = cand[i];
        if (x < 0 || x >= n) continue;
        int src;
        if (!ray_source(bc, x - c, n, src) || src != s) continue;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen = seen || out[j] == x;
        if (!seen && k < 2) out[k++] = x;
    }
    return k;
}

__host__ __device__ inline size_t table_slot(const Grid& g, int x, int y, int z, int q) {
    const uint32_t lane = (uint32_t)(x & (TILE_X - 1)) + (uint32_t)(y & (TILE_Y - 1)) * (uint32_t)TILE_X;
    return ((size_t)tile_of(g, x, y, z) * 19u + (size_t)q) * TILE_CELLS + lane;
}

}  // namespace bf
