// beamgrid.hip -- the reference's uniform beam grid, built on the GPU for per-cell beam reduction.
//
// Grid::build and buildloop (SmallUPBP/src/Beams/Grid.hxx:167-279) register every beam in each
// cell overlapped by the boxes of its chopped segments (PhotonBeam::getSegmentAABB,
// PhotonBeam.hxx:186-198), once per (beam, cell); PhGrid's AABB (PhGrid.hxx:75-87) is the union
// of the full beams' boxes. Here one thread per beam walks the same chops and cell ranges with
// the same FP32 operations (no contraction) and emits (cell, beam) keys; a radix sort and a
// unique pass give each cell its beams in ascending index order, and a second sort by a hashed
// key gives each entry its place in the cell's random permutation (PRESAMPLE / OFFSET,
// reduceBeams :558-631). The traversal keeps its LBVH: only the reduction reads this grid.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "upbp_internal.hpp"

namespace upbp {
namespace {

constexpr int kGridBlock = 256;

struct GridDev {
    float lo[3], inv_cell[3];
    int32_t res[3];
    int64_t n_cells;
};

struct BeamGeo {
    float o[3], d[3], len, r0, dr, m[3];
};

__device__ __forceinline__ BeamGeo load_beam(const float* origin, const float* dir, const float* length,
                                             const float* r_start, const float* r_end, int64_t i) {
    BeamGeo g;
    for (int a = 0; a < 3; ++a) {
        g.o[a] = origin[3 * i + a];
        g.d[a] = dir[3 * i + a];
    }
    g.len = length[i];
    g.r0 = r_start[i];
    g.dr = (r_end[i] - g.r0) / g.len;  // mRadiusChange, PhBeams.cxx:108 (0 for a constant radius)
    // margins of the cone's box, PhotonBeam.hxx:74-80
    const float dx = g.d[0] * g.d[0], dy = g.d[1] * g.d[1], dz = g.d[2] * g.d[2];
    g.m[0] = sqrtf(dy + dz);
    g.m[1] = sqrtf(dx + dz);
    g.m[2] = sqrtf(dx + dy);
    return g;
}

// PhotonBeam::getSegmentAABB over the beam parameters [s0, s1] * length.
__device__ __forceinline__ void segment_box(const BeamGeo& g, float s0, float s1, float* lo, float* hi) {
    s0 = s0 * g.len;
    s1 = s1 * g.len;
    const float rs = g.r0 + s0 * g.dr, re = g.r0 + s1 * g.dr;
    for (int a = 0; a < 3; ++a) {
        const float bs = g.o[a] + g.d[a] * s0, be = g.o[a] + g.d[a] * s1;
        const float ms = rs * g.m[a], me = re * g.m[a];
        const float l0 = bs - ms, l1 = be - me, h0 = bs + ms, h1 = be + me;
        lo[a] = l0 < l1 ? l0 : l1;
        hi[a] = h0 > h1 ? h0 : h1;
    }
}

// Order-preserving map of a finite float to an unsigned word (for atomicMin / atomicMax).
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float key_float(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__global__ void k_grid_bounds(int64_t n, const float* origin, const float* dir, const float* length,
                              const float* r_start, const float* r_end, uint32_t* box) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BeamGeo g = load_beam(origin, dir, length, r_start, r_end, i);
    float lo[3], hi[3];
    segment_box(g, 0.0f, 1.0f, lo, hi);
    for (int a = 0; a < 3; ++a) {
        atomicMin(&box[a], float_key(lo[a]));
        atomicMax(&box[3 + a], float_key(hi[a]));
    }
}

// buildloop's walk of one beam (Grid.hxx:223-254): visit(cell) for every cell of every chop's
// box, in the reference's order, with its "idx < 0 continue" / "idx >= cells break" handling.
template <class Visit>
__device__ __forceinline__ void walk_beam(const GridDev& G, const BeamGeo& g, Visit visit) {
    float blo[3], bhi[3];
    segment_box(g, 0.0f, 1.0f, blo, bhi);
    float ext[3] = {bhi[0] - blo[0], bhi[1] - blo[1], bhi[2] - blo[2]};
    const float ad0 = fabsf(g.d[0]), ad1 = fabsf(g.d[1]), ad2 = fabsf(g.d[2]);
    const int ax = (ad0 > ad1) ? ((ad0 > ad2) ? 0 : 2) : ((ad1 > ad2) ? 1 : 2);
    const int chops = (int)(ext[ax] * G.inv_cell[ax]) + 1;
    const float inv_chops = 1.0f / (float)chops;
    const int64_t sh1 = G.res[0], sh2 = (int64_t)G.res[0] * G.res[1];
    for (int ch = 0; ch < chops; ++ch) {
        float lo[3], hi[3];
        segment_box(g, (float)ch * inv_chops, (float)(ch + 1) * inv_chops, lo, hi);
        int is[3], ie[3];
        for (int a = 0; a < 3; ++a) {
            is[a] = (int)((lo[a] - G.lo[a]) * G.inv_cell[a]);
            ie[a] = (int)((hi[a] - G.lo[a]) * G.inv_cell[a]);
        }
        int64_t id = (int64_t)is[0] + (int64_t)G.res[0] * ((int64_t)is[1] + (int64_t)G.res[1] * is[2]);
        for (int z = is[2]; z <= ie[2]; ++z, id += sh2) {
            int64_t idy = id;
            for (int y = is[1]; y <= ie[1]; ++y, idy += sh1) {
                int64_t idx = idy;
                for (int x = is[0]; x <= ie[0]; ++x, ++idx) {
                    if (idx < 0) continue;
                    if (idx >= G.n_cells) break;
                    visit(idx);
                }
            }
        }
    }
}

__global__ void k_grid_count(int64_t n, GridDev G, const float* origin, const float* dir, const float* length,
                             const float* r_start, const float* r_end, uint32_t* count,
                             unsigned long long* total) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BeamGeo g = load_beam(origin, dir, length, r_start, r_end, i);
    uint64_t c = 0;
    walk_beam(G, g, [&](int64_t) { ++c; });
    count[i] = c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c;
    atomicAdd(total, (unsigned long long)c);
}

__global__ void k_grid_emit(int64_t n, GridDev G, const float* origin, const float* dir, const float* length,
                            const float* r_start, const float* r_end, const uint32_t* offset, uint64_t* keys) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BeamGeo g = load_beam(origin, dir, length, r_start, r_end, i);
    uint64_t* out = keys + offset[i];
    walk_beam(G, g, [&](int64_t idx) { *out++ = ((uint64_t)idx << 32) | (uint64_t)i; });
}

// ids, the end of each cell's run, and the permutation keys of the entries.
__global__ void k_grid_lists(int64_t m, const uint64_t* uniq, uint64_t seed, uint32_t* ids, uint32_t* run_end,
                             uint64_t* perm_key, uint32_t* perm_val) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t k = uniq[j];
    const uint32_t c = (uint32_t)(k >> 32), b = (uint32_t)k;
    ids[j] = b;
    if (j + 1 == m || (uint32_t)(uniq[j + 1] >> 32) != c) run_end[c] = (uint32_t)(j + 1);
    if (perm_key) {
        perm_key[j] = ((uint64_t)c << 32) | red_perm_key(seed, c, b);
        perm_val[j] = (uint32_t)j;
    }
}

__global__ void k_grid_ranks(int64_t m, const uint64_t* sorted_key, const uint32_t* sorted_val,
                             const uint32_t* off, uint32_t* rank) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t c = (uint32_t)(sorted_key[k] >> 32);
    rank[sorted_val[k]] = (uint32_t)k - off[c];
}

__global__ void k_selection_pdfs(int64_t n, RedDev R, const float* pos, float* out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pdf = 1.0f;
    if (R.max != 0) {
        const uint32_t c = red_cell(R, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        const uint32_t cnt = R.off[c + 1] - R.off[c];
        if (cnt > R.max) pdf = (float)R.max / (float)cnt;
    }
    out[i] = pdf;
}

// ---- GridStats: the reference grid's walk per query ---------------------------------------

struct GridWalkDev {
    float lo[3], hi[3], cell[3], inv_cell[3];
    int32_t res[3];
    const uint32_t* off;  // [cells + 1]
    uint32_t max;         // cells holding more beams are reduced (overfull); 0: none
};

// BoundingBox3::intersect (SmallUPBP/src/Structs/BoundingBox.hxx:218-271): the SSE slab test,
// whose min / max against +-inf send a NaN slab value to the bound that keeps the box.
__device__ __forceinline__ bool grid_box_intersect(const GridWalkDev& G, const float o[3], const float inv[3],
                                                   float& tmin, float& tmax) {
    float mx[3], mn[3];
    for (int a = 0; a < 3; ++a) {
        const float l1 = (G.lo[a] - o[a]) * inv[a];
        const float l2 = (G.hi[a] - o[a]) * inv[a];
        const float l1a = (l1 < INFINITY) ? l1 : INFINITY, l2a = (l2 < INFINITY) ? l2 : INFINITY;
        const float l1b = (l1 > -INFINITY) ? l1 : -INFINITY, l2b = (l2 > -INFINITY) ? l2 : -INFINITY;
        mx[a] = (l1a > l2a) ? l1a : l2a;
        mn[a] = (l1b < l2b) ? l1b : l2b;
    }
    float lmax = mx[0], lmin = mn[0];
    if (mx[1] < lmax) lmax = mx[1];
    if (mn[1] > lmin) lmin = mn[1];
    if (mx[2] < lmax) lmax = mx[2];
    if (mn[2] > lmin) lmin = mn[2];
    if (!(lmax >= lmin)) return false;
    tmin = lmin;
    tmax = lmax;
    return true;
}

// Grid::intersect's 3D-DDA (SmallUPBP/src/Beams/Grid.hxx:318-399) in its FP32 operations,
// gathering only its GridStats (:365,378-382): the beams each crossed cell holds (max of an
// overfull one), the cells, the overfull cells. Every step moves one axis' cell index toward
// that axis' exit value, so a walk takes at most res[0] + res[1] + res[2] steps.
__device__ void grid_stats_walk(const GridWalkDev& G, const float o[3], const float d[3], float mint, float maxt,
                                uint32_t gs[3]) {
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    float t0, t1;
    if (!grid_box_intersect(G, o, inv, t0, t1)) return;
    if (t0 < mint) t0 = mint;
    if (t1 > maxt) t1 = maxt;
    if (!(t1 > t0)) return;
    float delta[3], enter[3], l[3];
    int ipos[3], step[3], check[3];
    const int64_t stride[3] = {1, (int64_t)G.res[0], (int64_t)G.res[0] * G.res[1]};
    for (int a = 0; a < 3; ++a) {
        delta[a] = G.cell[a] * inv[a];
        enter[a] = ((o[a] + d[a] * t0) - G.lo[a]) * G.inv_cell[a];
        const int v = (int)enter[a];
        ipos[a] = v < 0 ? 0 : (v > G.res[a] - 1 ? G.res[a] - 1 : v);
    }
    int64_t cell = (int64_t)ipos[0] + (int64_t)G.res[0] * ((int64_t)ipos[1] + (int64_t)G.res[1] * ipos[2]);
    int shift[3];
    for (int a = 0; a < 3; ++a) {
        if (d[a] >= 0.0f) {
            shift[a] = 1;
            check[a] = G.res[a];
            step[a] = 1;
        } else {
            shift[a] = 0;
            check[a] = -1;
            step[a] = -1;
            delta[a] = -delta[a];
        }
    }
    for (int a = 0; a < 3; ++a) l[a] = (((float)(ipos[a] + shift[a]) - enter[a]) * inv[a]) * G.cell[a];
    float t = t0;
    do {
        const uint32_t n = G.off[cell + 1] - G.off[cell];
        const bool over = G.max != 0 && n > G.max;
        gs[0] += over ? G.max : n;
        gs[1] += 1;
        gs[2] += over ? 1u : 0u;
        const int m = (l[0] < l[1]) ? ((l[0] < l[2]) ? 0 : 2) : ((l[1] < l[2]) ? 1 : 2);
        t += l[m];
        const float lm = l[m];
        for (int a = 0; a < 3; ++a) l[a] -= lm;
        l[m] = delta[m];
        cell += stride[m] * step[m];
        ipos[m] += step[m];
        if (ipos[m] == check[m]) break;
    } while (t < t1);
}

// One query per thread: the walk over its segments in a scattering medium, as
// PhotonBeamsEvaluator::evalBeamBeamEstimate (SmallUPBP/src/Beams/PhBeams.cxx:177-230) calls
// Grid::intersect, ending once a channel of the FP32 attenuation product (divided by the
// ray-sample pdf for SHORT queries) is not above 1e-20 (Rgb::isPositive).
__global__ void k_grid_stats(int64_t nq, GridWalkDev G, const float* origin, const float* direction,
                             const int64_t* offsets, const upbp_volume_segment* seg, const MediumDev* media,
                             int32_t n_media, int32_t short_q, uint32_t* out, uint32_t* err) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t gs[3] = {0u, 0u, 0u};
    const int64_t s0 = offsets[q], s1 = offsets[q + 1], sn = offsets[nq];
    if (!(s0 >= 0 && s0 <= s1 && s1 <= sn)) {
        atomicOr(err, kErrBadOffsets);
    } else {
        const float o[3] = {origin[3 * q], origin[3 * q + 1], origin[3 * q + 2]};
        const float d[3] = {direction[3 * q], direction[3 * q + 1], direction[3 * q + 2]};
        float att[3] = {1.0f, 1.0f, 1.0f};
        for (int64_t i = s0; i < s1; ++i) {
            const upbp_volume_segment& sg = seg[i];
            if (sg.medium_id < 0 || sg.medium_id >= n_media) {
                atomicOr(err, kErrBadMedium);
                break;
            }
            if (media[sg.medium_id].has_scattering) grid_stats_walk(G, o, d, sg.dist_min, sg.dist_max, gs);
            for (int c = 0; c < 3; ++c)
                att[c] = att[c] * (short_q ? sg.attenuation[c] / sg.ray_sample_pdf : sg.attenuation[c]);
            if (!(att[0] > 1e-20f && att[1] > 1e-20f && att[2] > 1e-20f)) break;
        }
    }
    out[3 * q] = gs[0];
    out[3 * q + 1] = gs[1];
    out[3 * q + 2] = gs[2];
}

// Per-beam origin, direction and length from the pieces (every valid beam has at least one,
// and all of a beam's pieces hold the same values, so concurrent writes agree).
__global__ void k_beams_from_pieces(int64_t np, const PieceDev* pieces, const uint32_t* piece_beam, float* origin,
                                    float* dir, float* length) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= np) return;
    const PieceDev pc = pieces[p];
    const int64_t b = piece_beam[p];
    origin[3 * b] = pc.p0.x;
    origin[3 * b + 1] = pc.p0.y;
    origin[3 * b + 2] = pc.p0.z;
    dir[3 * b] = pc.p1.x;
    dir[3 * b + 1] = pc.p1.y;
    dir[3 * b + 2] = pc.p1.z;
    length[b] = pc.p0.w;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kGridBlock - 1) / kGridBlock); }

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

}  // namespace

void beam_reduction_build(Ctx& ctx, int64_t n, const float* origin, const float* dir, const float* length,
                          const float* r_start, const float* r_end, const upbp_build_params& params, Beams& out,
                          bool always) {
    out.red = RedDev{};
    out.red_cells = 0;
    if (params.max_beams_in_cell == 0 && !always) return;
    UPBP_REQUIRE(params.reduction_type <= UPBP_REDUCE_RESAMPLE, UPBP_EINVAL,
                 "reduction_type must be one of UPBP_REDUCE_*");
    const int res = params.grid_resolution > 0 ? params.grid_resolution : 256;
    UPBP_REQUIRE(res <= 2048, UPBP_EINVAL, "grid_resolution must be in 0..2048");
    if (n == 0) return;
    hipStream_t st = ctx.stream;
    // grid box: union of the full beams' boxes (PhGrid.hxx:75-87)
    DevBuf box;
    box.ensure(6 * 4);
    const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    UPBP_HIP(hipMemcpyAsync(box.p, init, sizeof(init), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_grid_bounds, dim3(blocks_for(n)), dim3(kGridBlock), 0, st, n, origin, dir, length, r_start,
                       r_end, box.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    uint32_t words[6];
    UPBP_HIP(hipMemcpyAsync(words, box.p, sizeof(words), hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    // Grid::build's resolution and cell size (Grid.hxx:167-221), the same FP32 operations
    GridDev G;
    float ext[3];
    for (int a = 0; a < 3; ++a) {
        G.lo[a] = key_float(words[a]);
        ext[a] = key_float(words[3 + a]) - G.lo[a];
    }
    float emax = ext[0] > ext[1] ? ext[0] : ext[1];
    if (ext[2] > emax) emax = ext[2];
    const float size = emax / (float)res;
    UPBP_REQUIRE(size > 0.f && std::isfinite(size), UPBP_EINVAL, "beam reduction grid has an empty box");
    for (int a = 0; a < 3; ++a) {
        G.res[a] = (int)(unsigned)(ext[a] / size) + 1;
        const float cell = ext[a] / (float)G.res[a];
        G.inv_cell[a] = 1.0f / cell;
        out.grid_cell[a] = cell;
        out.grid_hi[a] = key_float(words[3 + a]);
    }
    G.n_cells = (int64_t)(G.res[0] + 1) * (G.res[1] + 1) * (G.res[2] + 1) + 1;
    UPBP_REQUIRE(G.n_cells < (int64_t)0x7FFFFFFF, UPBP_EINVAL, "beam reduction grid has too many cells");
    // registrations per beam, then the (cell, beam) keys
    DevBuf count, offset, sum;
    count.ensure(n * 4);
    offset.ensure(n * 4);
    sum.ensure(8);
    UPBP_HIP(hipMemsetAsync(sum.p, 0, 8, st));
    hipLaunchKernelGGL(k_grid_count, dim3(blocks_for(n)), dim3(kGridBlock), 0, st, n, G, origin, dir, length,
                       r_start, r_end, count.as<uint32_t>(), sum.as<unsigned long long>());
    UPBP_HIP(hipGetLastError());
    uint64_t total = 0;
    UPBP_HIP(hipMemcpyAsync(&total, sum.p, 8, hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    UPBP_REQUIRE(total < 0x7FFFFFFFull, UPBP_EINVAL, "too many beam registrations in the reduction grid");
    exclusive_scan_u32(ctx, count.as<uint32_t>(), offset.as<uint32_t>(), n);
    DevBuf keys, sorted, nsel;
    keys.ensure(total * 8 + 8);
    sorted.ensure(total * 8 + 8);
    nsel.ensure(8);
    hipLaunchKernelGGL(k_grid_emit, dim3(blocks_for(n)), dim3(kGridBlock), 0, st, n, G, origin, dir, length, r_start,
                       r_end, offset.as<uint32_t>(), keys.as<uint64_t>());
    UPBP_HIP(hipGetLastError());
    const int end_bit = 32 + bits_for((uint64_t)G.n_cells);
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                               (int)total, 0, end_bit, st));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceRadixSort::SortKeys(ctx.sort_tmp.p, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                               (int)total, 0, end_bit, st));
    // one entry per (cell, beam): a beam's chops may share cells (the reference's mDuplicates)
    bytes = 0;
    UPBP_HIP(hipcub::DeviceSelect::Unique(nullptr, bytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                                          nsel.as<int>(), (int)total, st));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceSelect::Unique(ctx.sort_tmp.p, bytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                                          nsel.as<int>(), (int)total, st));
    int32_t m32 = 0;
    UPBP_HIP(hipMemcpyAsync(&m32, nsel.p, 4, hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    const int64_t m = m32;
    const bool ranks = params.reduction_type == UPBP_REDUCE_PRESAMPLE || params.reduction_type == UPBP_REDUCE_OFFSET;
    DevBuf run_end, pkey, pval, pkey_s, pval_s;
    run_end.ensure(G.n_cells * 4);
    UPBP_HIP(hipMemsetAsync(run_end.p, 0, G.n_cells * 4, st));
    out.red_ids.ensure(m * 4 + 4);
    out.red_off.ensure((G.n_cells + 1) * 4);
    if (ranks) {
        pkey.ensure(m * 8 + 8);
        pval.ensure(m * 4 + 4);
    }
    if (m > 0)
        hipLaunchKernelGGL(k_grid_lists, dim3(blocks_for(m)), dim3(kGridBlock), 0, st, m, keys.as<uint64_t>(),
                           (uint64_t)(uint32_t)params.reduction_seed, out.red_ids.as<uint32_t>(),
                           run_end.as<uint32_t>(), ranks ? pkey.as<uint64_t>() : nullptr,
                           ranks ? pval.as<uint32_t>() : nullptr);
    UPBP_HIP(hipGetLastError());
    // off[0] = 0, off[c + 1] = end of the last non-empty cell <= c
    UPBP_HIP(hipMemsetAsync(out.red_off.p, 0, 4, st));
    {
        size_t b2 = 0;
        UPBP_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, b2, run_end.as<uint32_t>(), out.red_off.as<uint32_t>() + 1,
                                                   hipcub::Max(), (int)G.n_cells, st));
        ctx.sort_tmp.ensure(b2);
        UPBP_HIP(hipcub::DeviceScan::InclusiveScan(ctx.sort_tmp.p, b2, run_end.as<uint32_t>(),
                                                   out.red_off.as<uint32_t>() + 1, hipcub::Max(), (int)G.n_cells,
                                                   st));
    }
    out.red_rank.ensure(m * 4 + 4);
    if (ranks && m > 0) {
        pkey_s.ensure(m * 8);
        pval_s.ensure(m * 4);
        sort_pairs_u64(ctx, pkey.as<uint64_t>(), pkey_s.as<uint64_t>(), pval.as<uint32_t>(), pval_s.as<uint32_t>(), m,
                       64);
        hipLaunchKernelGGL(k_grid_ranks, dim3(blocks_for(m)), dim3(kGridBlock), 0, st, m, pkey_s.as<uint64_t>(),
                           pval_s.as<uint32_t>(), out.red_off.as<uint32_t>(), out.red_rank.as<uint32_t>());
        UPBP_HIP(hipGetLastError());
    }
    UPBP_HIP(hipStreamSynchronize(st));
    RedDev& R = out.red;
    R.max = params.max_beams_in_cell;
    R.type = params.reduction_type;
    R.seed = (uint64_t)(uint32_t)params.reduction_seed;
    for (int a = 0; a < 3; ++a) {
        R.lo[a] = G.lo[a];
        R.inv[a] = G.inv_cell[a];
        R.res[a] = G.res[a];
    }
    R.off = out.red_off.as<uint32_t>();
    R.ids = out.red_ids.as<uint32_t>();
    R.rank = out.red_rank.as<uint32_t>();
    out.red_cells = G.n_cells;
}

float beam_selection_pdf(const Beams& b, const float pos[3]) {
    if (b.red.max == 0 || !b.ctx) return 1.0f;
    const uint32_t c = red_cell(b.red, pos[0], pos[1], pos[2]);
    uint32_t w[2];
    UPBP_HIP(hipMemcpy(w, b.red.off + c, 8, hipMemcpyDeviceToHost));
    const uint32_t n = w[1] - w[0];
    return n <= b.red.max ? 1.0f : (float)b.red.max / (float)n;
}

void beam_selection_pdfs(const Beams& b, const float* pos, int64_t n, float* out, upbp_memory mem) {
    if (n == 0) return;
    Ctx& ctx = *b.ctx;
    hipStream_t st = ctx.stream;
    DevBuf in_buf, out_buf;
    const float* p = pos;
    float* o = out;
    if (mem == UPBP_MEM_HOST) {
        in_buf.ensure(n * 12);
        out_buf.ensure(n * 4);
        UPBP_HIP(hipMemcpyAsync(in_buf.p, pos, n * 12, hipMemcpyHostToDevice, st));
        p = in_buf.as<float>();
        o = out_buf.as<float>();
    }
    hipLaunchKernelGGL(k_selection_pdfs, dim3(blocks_for(n)), dim3(kGridBlock), 0, st, n, b.red, p, o);
    UPBP_HIP(hipGetLastError());
    if (mem == UPBP_MEM_HOST) UPBP_HIP(hipMemcpyAsync(out, o, n * 4, hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
}

void beam_grid_stats(Ctx& ctx, const Beams& b, const QueryView& q, uint32_t* out) {
    if (q.count == 0) return;
    hipStream_t st = ctx.stream;
    const Beams* g = &b;
    Beams tmp;
    if (b.red_cells == 0 && b.n > 0 && b.pieces > 0) {
        // no grid kept (no reduction): build one of the beams' resolution for this call
        DevBuf o, d, len;
        o.ensure(b.n * 12);
        d.ensure(b.n * 12);
        len.ensure(b.n * 4);
        hipLaunchKernelGGL(k_beams_from_pieces, dim3(blocks_for(b.pieces)), dim3(kGridBlock), 0, st, b.pieces,
                           b.piece.as<PieceDev>(), b.piece_beam.as<uint32_t>(), o.as<float>(), d.as<float>(),
                           len.as<float>());
        UPBP_HIP(hipGetLastError());
        upbp_build_params p{};
        p.grid_resolution = b.grid_resolution;
        p.max_beams_in_cell = 0;
        p.reduction_type = UPBP_REDUCE_RESAMPLE;  // no permutation ranks needed
        beam_reduction_build(ctx, b.n, o.as<float>(), d.as<float>(), len.as<float>(), b.r_start.as<float>(),
                             b.r_end.as<float>(), p, tmp, true);
        g = &tmp;
    }
    if (g->red_cells == 0) {
        // no beams: Grid::intersect misses the empty grid box
        launch_fill_u32(ctx, out, 0u, 3 * q.count);
        UPBP_HIP(hipStreamSynchronize(st));
        return;
    }
    GridWalkDev G;
    for (int a = 0; a < 3; ++a) {
        G.lo[a] = g->red.lo[a];
        G.hi[a] = g->grid_hi[a];
        G.cell[a] = g->grid_cell[a];
        G.inv_cell[a] = g->red.inv[a];
        G.res[a] = g->red.res[a];
    }
    G.off = g->red.off;
    G.max = b.red.max;
    DevBuf err;
    err.ensure(4);
    launch_fill_u32(ctx, err.as<uint32_t>(), 0u, 1);
    hipLaunchKernelGGL(k_grid_stats, dim3(blocks_for(q.count)), dim3(kGridBlock), 0, st, q.count, G, q.origin,
                       q.direction, q.offsets, q.segments, ctx.media.as<MediumDev>(), (int32_t)ctx.n_media,
                       (int32_t)(q.beam_type == UPBP_SHORT_BEAM), out, err.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    uint32_t e = 0;
    UPBP_HIP(hipMemcpyAsync(&e, err.p, 4, hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    UPBP_REQUIRE(!(e & kErrBadOffsets), UPBP_EINVAL, "segment_offsets out of order or out of range");
    UPBP_REQUIRE(!(e & kErrBadMedium), UPBP_EINVAL, "segment medium outside the media table");
}

}  // namespace upbp
