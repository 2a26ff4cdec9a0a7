// beamgrid.hpp -- per-cell beam reduction of the reference's uniform beam grid, query side.
//
// The reference may test only some of the beams registered in a crowded grid cell
// (Grid::intersect with mMaxBeamsInCell > 0, SmallUPBP/src/Beams/Grid.hxx:318-399,440-631) and
// weights each tested beam by 1 / pdf with pdf = max / n (reduceBeams, :558-631). A hit
// belongs to the cell holding its query point: Grid::intersect only accepts it while the
// query parameter is inside that cell's interval. Here the traversal finds hits as before
// and each hit is kept and weighted as the cell's reduction decides:
//   PRESAMPLE       kept iff the beam is among the cell's first max beams of a random
//                   permutation fixed at build time;
//   OFFSET          kept iff the beam's place in that permutation is within max of a random
//                   offset drawn per (query, cell);
//   RESAMPLE_FIXED  counted once per draw, of max draws with replacement per (query, cell),
//                   that picks its place in the cell's list (the reference's registration
//                   order, beams by decreasing index);
//   RESAMPLE        kept with probability pdf, drawn per (query, cell, beam).
// The random numbers are counter-based hashes of (seed, query, cell, draw), the same in the
// CPU restatement. A hit in a cell where its beam is not registered is dropped, as the
// reference never tests it there; cells with n <= max keep every hit (pdf 1).
#pragma once

#include <stdint.h>

#include "../../include/upbp.h"

namespace upbp {

struct RedDev {
    uint32_t max = 0;   // 0: no reduction
    uint32_t type = 0;  // UPBP_REDUCE_*
    uint64_t seed = 0;
    float lo[3] = {0.f, 0.f, 0.f}, inv[3] = {0.f, 0.f, 0.f};
    int32_t res[3] = {0, 0, 0};
    const uint32_t* off = nullptr;   // [cells + 1] start of each cell's beams
    const uint32_t* ids = nullptr;   // beam indices per cell, ascending
    const uint32_t* rank = nullptr;  // place of each entry in the cell's build-time permutation
};

__host__ __device__ inline uint64_t red_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The permutation key of beam b in cell c (build time).
__host__ __device__ inline uint32_t red_perm_key(uint64_t seed, uint32_t c, uint32_t b) {
    return (uint32_t)(red_mix(seed ^ red_mix(((uint64_t)c << 32) | b)) >> 32);
}

// Uniform in (0, 1) for (query key, cell, draw).
__host__ __device__ inline float red_uniform(uint64_t seed, uint64_t qkey, uint32_t c, uint32_t t) {
    const uint64_t h = red_mix(seed + red_mix(qkey ^ red_mix(((uint64_t)c << 32) | t)));
    return ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

// (uint)(u * n) as the reference picks a place, kept below n when the product rounds up.
__host__ __device__ inline uint32_t red_index(float u, uint32_t n) {
    const uint32_t i = (uint32_t)(u * (float)n);
    return i < n ? i : n - 1;
}

// A query's key: its ray (origin and direction bits).
__host__ __device__ inline uint64_t red_query_key(const uint32_t bits[6]) {
    uint64_t k = 0x243F6A8885A308D3ull;
    for (int i = 0; i < 6; ++i) k = red_mix(k ^ bits[i]);
    return k;
}

// Cell of a position: Grid::pdf's posToVoxel and index (Grid.hxx:71-116,409-415).
__host__ __device__ inline uint32_t red_cell(const RedDev& R, float x, float y, float z) {
    const float p[3] = {(x - R.lo[0]) * R.inv[0], (y - R.lo[1]) * R.inv[1], (z - R.lo[2]) * R.inv[2]};
    int32_t v[3];
    for (int a = 0; a < 3; ++a) {
        int32_t i = (int32_t)p[a];
        v[a] = i < 0 ? 0 : (i > R.res[a] - 1 ? R.res[a] - 1 : i);
    }
    return (uint32_t)(v[0] + R.res[0] * (v[1] + R.res[1] * v[2]));
}

// Selection pdf of the cell holding (x, y, z): Grid::pdf (Grid.hxx:409-415), max / n in a cell
// of n > max beams, else 1.
__device__ inline float red_cell_pdf(const RedDev& R, float x, float y, float z) {
    const uint32_t c = red_cell(R, x, y, z);
    const uint32_t n = R.off[c + 1] - R.off[c];
    return n <= R.max ? 1.0f : (float)R.max / (float)n;
}

// Times (0 = dropped) a hit of beam b whose query point is (x, y, z) is counted, and the
// selection pdf of its cell (each count weighs 1 / pdf).
__device__ inline uint32_t red_decide(const RedDev& R, uint64_t qkey, float x, float y, float z, uint32_t b,
                                      float& pdf) {
    pdf = 1.0f;
    const uint32_t c = red_cell(R, x, y, z);
    const uint32_t b0 = R.off[c], n = R.off[c + 1] - b0;
    if (n <= R.max) return 1u;
    pdf = (float)R.max / (float)n;
    uint32_t lo = 0, hi = n;  // the beam's entry among the cell's ascending ids
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (R.ids[b0 + mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo >= n || R.ids[b0 + lo] != b) return 0u;
    uint32_t mult = 0;
    if (R.type == UPBP_REDUCE_PRESAMPLE) {
        mult = R.rank[b0 + lo] < R.max ? 1u : 0u;
    } else if (R.type == UPBP_REDUCE_OFFSET) {
        const uint32_t o = red_index(red_uniform(R.seed, qkey, c, 0), n);
        mult = ((R.rank[b0 + lo] + n - o) % n) < R.max ? 1u : 0u;
    } else if (R.type == UPBP_REDUCE_RESAMPLE_FIXED) {
        const uint32_t place = n - 1 - lo;  // registration order: decreasing beam index
        for (uint32_t t = 0; t < R.max; ++t)
            if (red_index(red_uniform(R.seed, qkey, c, t), n) == place) ++mult;
    } else {
        mult = red_uniform(R.seed, qkey, c, b) < pdf ? 1u : 0u;
    }
    return mult;
}

}  // namespace upbp
