// lbvh.hip -- GPU linear BVH build (Morton codes, radix sort, Karras hierarchy,
// bottom-up refit, emission of 64-byte two-box nodes).
//
// Replaces the Embree BVH4 binned-SAH build over photon cubes
// (SmallUPBP/src/Bre/Bre.cxx:492-507, embree-2.0/embree/bvh4/bvh4.cpp:71) and
// the beam grid build (SmallUPBP/src/Beams/Grid.hxx:223-279). Culling only:
// node boxes are widened by a margin so that no item whose predicate accepts a
// query is ever culled; the hit set is then the brute-force hit set.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "upbp_internal.hpp"

namespace upbp {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

// bounds[0..2] = min item-box lo, [3..5] = max item-box hi,
// [6..8] = min centroid, [9..11] = max centroid (order-preserving uint encoding).
// Per-block partial bounds; k_bounds_final folds the partials (no global atomics:
// thousands of atomics on the same twelve words serialise).
__global__ __launch_bounds__(kBlock) void k_bounds(int64_t n, const float4* lo, const float4* hi,
                                                   uint32_t* partial) {
    uint32_t v[12];
    for (int k = 0; k < 12; ++k) v[k] = ((k / 3) % 2 == 0) ? 0xFFFFFFFFu : 0u;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float4 l = lo[i], h = hi[i];
        float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
        float ll[3] = {l.x, l.y, l.z}, hh[3] = {h.x, h.y, h.z};
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], f2ord(ll[a]));
            v[3 + a] = max(v[3 + a], f2ord(hh[a]));
            v[6 + a] = min(v[6 + a], f2ord(c[a]));
            v[9 + a] = max(v[9 + a], f2ord(c[a]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 12; ++k) {
            uint32_t o = __shfl_xor(v[k], off, 64);
            v[k] = ((k / 3) % 2 == 0) ? min(v[k], o) : max(v[k], o);
        }
    }
    __shared__ uint32_t sh[kBlock / 64][12];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; ++k) sh[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        uint32_t r = sh[0][k];
        for (int j = 1; j < (int)(blockDim.x / 64); ++j) r = ((k / 3) % 2 == 0) ? min(r, sh[j][k]) : max(r, sh[j][k]);
        partial[(size_t)blockIdx.x * 12 + k] = r;
    }
}

// One wave per bound component: its lanes stride over the per-block partials, then a wave
// reduction (a single thread per component took ~125 us over 1024 partials).
__global__ void k_bounds_final(int nb, const uint32_t* partial, uint32_t* b) {
    const int k = threadIdx.x / 64, lane = threadIdx.x & 63;
    if (k >= 12) return;  // whole waves
    const bool is_min = (k / 3) % 2 == 0;
    uint32_t r = is_min ? 0xFFFFFFFFu : 0u;
    for (int i = lane; i < nb; i += 64) {
        const uint32_t v = partial[(size_t)i * 12 + k];
        r = is_min ? min(r, v) : max(r, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)r, off, 64);
        r = is_min ? min(r, o) : max(r, o);
    }
    if (lane == 0) b[k] = r;
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(int64_t n, const float4* lo, const float4* hi,
                                                   const uint32_t* b, uint64_t* keys, uint32_t* vals) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 l = lo[i], h = hi[i];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint64_t q[3];
    for (int a = 0; a < 3; ++a) {
        float mn = ord2f(b[6 + a]), mx = ord2f(b[9 + a]);
        float ext = mx - mn;
        float u = ext > 0.f ? (c[a] - mn) / ext : 0.5f;
        u = fminf(fmaxf(u, 0.f), 1.f);
        q[a] = (uint64_t)(u * 2097151.0f);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = (uint32_t)i;
}

// Length of the common prefix of keys i and j (index-tie-broken), -1 out of range.
__device__ __forceinline__ int delta(const uint64_t* keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// Karras (HPG 2012) binary radix tree: internal node i in [0, n-2].
// child encoding: bit 31 set = leaf (sorted item index), clear = internal node.
__global__ __launch_bounds__(kBlock) void k_karras(int64_t n, const uint64_t* keys, uint32_t* left,
                                                   uint32_t* right, uint32_t* first, uint32_t* last,
                                                   uint32_t* leaf_parent, uint32_t* int_parent) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t lc, rc;
    if (lo == gamma) {
        lc = 0x80000000u | (uint32_t)gamma;
        leaf_parent[gamma] = (uint32_t)i;
    } else {
        lc = (uint32_t)gamma;
        int_parent[gamma] = (uint32_t)i;
    }
    if (hi == gamma + 1) {
        rc = 0x80000000u | (uint32_t)(gamma + 1);
        leaf_parent[gamma + 1] = (uint32_t)i;
    } else {
        rc = (uint32_t)(gamma + 1);
        int_parent[gamma + 1] = (uint32_t)i;
    }
    left[i] = lc;
    right[i] = rc;
    first[i] = (uint32_t)lo;
    last[i] = (uint32_t)hi;
}

// Level-synchronous refit. Every internal node's depth is found by walking its
// parent links; nodes are grouped by depth and merged one level per launch, deepest
// first, so a node's children are final before it reads them. Kernel boundaries
// order the levels: no atomics or device-scope fences, which on a multi-XCD GPU
// would write back and invalidate L2 at every level of every climb.
constexpr uint32_t kMaxDepth = 128;  // > Karras depth bound (63 key bits + 31 index bits)

__global__ __launch_bounds__(kBlock) void k_depth(int64_t ni, const uint32_t* int_parent, uint32_t* depth,
                                                  uint32_t* ids, uint32_t* err) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ni) return;
    uint32_t node = (uint32_t)i, d = 0;
    while (node != 0 && d < kMaxDepth) {
        node = int_parent[node];
        ++d;
    }
    if (d >= kMaxDepth) {
        atomicOr(err, 1u);
        d = kMaxDepth - 1;
    }
    depth[i] = d;
    ids[i] = (uint32_t)i;
}

// start[d] = first position of depth d in the depth-sorted node list
__global__ __launch_bounds__(kBlock) void k_level_starts(int64_t ni, const uint32_t* sorted_depth, uint32_t* start) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= ni) return;
    const uint32_t d = sorted_depth[j];
    if (j == 0 || sorted_depth[j - 1] != d) start[d] = (uint32_t)j;
}

__global__ __launch_bounds__(kBlock) void k_refit_level(int64_t count, const uint32_t* ids, const uint32_t* order,
                                                        const float4* lo, const float4* hi, const uint32_t* left,
                                                        const uint32_t* right, float4* ilo, float4* ihi) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t node = ids[t];
    float4 b0l, b0h, b1l, b1h;
    const uint32_t c0 = left[node], c1 = right[node];
    if (c0 & 0x80000000u) {
        const uint32_t it = order[c0 & 0x7FFFFFFFu];
        b0l = lo[it];
        b0h = hi[it];
    } else {
        b0l = ilo[c0];
        b0h = ihi[c0];
    }
    if (c1 & 0x80000000u) {
        const uint32_t it = order[c1 & 0x7FFFFFFFu];
        b1l = lo[it];
        b1h = hi[it];
    } else {
        b1l = ilo[c1];
        b1h = ihi[c1];
    }
    ilo[node] = make_float4(fminf(b0l.x, b1l.x), fminf(b0l.y, b1l.y), fminf(b0l.z, b1l.z), 0.f);
    ihi[node] = make_float4(fmaxf(b0h.x, b1h.x), fmaxf(b0h.y, b1h.y), fmaxf(b0h.z, b1h.z), 0.f);
}

__global__ __launch_bounds__(kBlock) void k_emit(int64_t n, int leaf_size, float eps,
                                                 const uint32_t* order, const float4* lo,
                                                 const float4* hi, const uint32_t* left,
                                                 const uint32_t* right, const uint32_t* first,
                                                 const uint32_t* last, const float4* ilo,
                                                 const float4* ihi, NodeDev* nodes) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    uint32_t ref[2], cnt[2];
    float4 bl[2], bh[2];
    uint32_t ch[2] = {left[i], right[i]};
    for (int c = 0; c < 2; ++c) {
        if (ch[c] & 0x80000000u) {
            uint32_t s = ch[c] & 0x7FFFFFFFu;
            uint32_t it = order[s];
            bl[c] = lo[it];
            bh[c] = hi[it];
            ref[c] = s;
            cnt[c] = 1;
        } else {
            uint32_t m = ch[c];
            bl[c] = ilo[m];
            bh[c] = ihi[m];
            uint32_t count = last[m] - first[m] + 1;
            if ((int)count <= leaf_size) {
                ref[c] = first[m];
                cnt[c] = count;
            } else {
                // the subtree's size and where its items start, for traversals that take a
                // small subtree as one leaf
                ref[c] = m;
                cnt[c] = kInner | (first[m] == m ? 0u : kSpanEnds) | min(count, kSpanMask - 1u);
            }
        }
        bl[c] = make_float4(bl[c].x - eps, bl[c].y - eps, bl[c].z - eps, 0.f);
        bh[c] = make_float4(bh[c].x + eps, bh[c].y + eps, bh[c].z + eps, 0.f);
    }
    NodeDev nd;
    nd.a = make_float4(bl[0].x, bl[0].y, bl[0].z, bh[0].x);
    nd.b = make_float4(bh[0].y, bh[0].z, bl[1].x, bl[1].y);
    nd.c = make_float4(bl[1].z, bh[1].x, bh[1].y, bh[1].z);
    nd.r = make_uint4(ref[0], ref[1], cnt[0], cnt[1]);
    nodes[i] = nd;
}

__global__ void k_iota(int64_t n, uint32_t* v) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)std::max<int64_t>(1, g);
}

}  // namespace

void sort_pairs_u64(Ctx& ctx, uint64_t* keys_in, uint64_t* keys_out, uint32_t* vals_in,
                    uint32_t* vals_out, int64_t n, int end_bit) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                                (int)n, 0, end_bit, ctx.stream));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx.sort_tmp.p, bytes, keys_in, keys_out, vals_in,
                                                vals_out, (int)n, 0, end_bit, ctx.stream));
}

void sort_pairs_u32(Ctx& ctx, uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_in,
                    uint32_t* vals_out, int64_t n, int end_bit) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                                (int)n, 0, end_bit, ctx.stream));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(ctx.sort_tmp.p, bytes, keys_in, keys_out, vals_in,
                                                vals_out, (int)n, 0, end_bit, ctx.stream));
}

size_t sort_pairs_u32_temp_bytes(int64_t n, int end_bit) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, end_bit,
                                                (hipStream_t)0));
    return bytes;
}

void sort_pairs_u32_in(Ctx& ctx, void* tmp, size_t tmp_bytes, uint32_t* keys_in, uint32_t* keys_out,
                       uint32_t* vals_in, uint32_t* vals_out, int64_t n, int end_bit) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                                (int)n, 0, end_bit, ctx.stream));
    UPBP_REQUIRE(bytes <= tmp_bytes, UPBP_EDEVICE, "radix sort needs more temporary storage than the workspace holds");
    UPBP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
                                                end_bit, ctx.stream));
}

void exclusive_scan_u32(Ctx& ctx, const uint32_t* in, uint32_t* out, int64_t n) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx.stream));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceScan::ExclusiveSum(ctx.sort_tmp.p, bytes, in, out, (int)n, ctx.stream));
}

void build_lbvh(Ctx& ctx, int64_t n, const float4* box_lo, const float4* box_hi, int leaf_size,
                Lbvh& out, DevBuf& out_order) {
    UPBP_REQUIRE(n > 0 && n < (int64_t)0x7FFFFFFF, UPBP_EINVAL, "LBVH item count out of range");
    hipStream_t st = ctx.stream;
    out.n_items = n;
    TmpBuf bounds(ctx.tmp), partial(ctx.tmp);
    bounds.ensure(12 * 4);
    const unsigned gb = std::min<unsigned>(grid_for(n), 1024u);
    partial.ensure((size_t)gb * 12 * 4);
    hipLaunchKernelGGL(k_bounds, dim3(gb), dim3(kBlock), 0, st, n, box_lo, box_hi, partial.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(12 * 64), 0, st, (int)gb, partial.as<uint32_t>(),
                       bounds.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    uint32_t hb[12];
    UPBP_HIP(hipMemcpyAsync(hb, bounds.p, sizeof(hb), hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    float maxabs = 0.f, maxext = 0.f;
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = ord2f(hb[a]);
        out.hi[a] = ord2f(hb[3 + a]);
        UPBP_REQUIRE(std::isfinite(out.lo[a]) && std::isfinite(out.hi[a]), UPBP_ENONFINITE,
                     "non-finite item bounds");
        maxabs = std::max(maxabs, std::max(std::fabs(out.lo[a]), std::fabs(out.hi[a])));
        maxext = std::max(maxext, out.hi[a] - out.lo[a]);
    }
    // Margin that keeps the float slab test conservative for every accepted item.
    const float eps = 1e-5f * std::max(std::max(maxabs, maxext), 1e-3f);

    out_order.ensure((size_t)n * 4);
    if (n <= leaf_size) {
        // one node: child 0 is a leaf with every item, child 1 absent
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, st, n, out_order.as<uint32_t>());
        NodeDev nd;
        nd.a = make_float4(out.lo[0] - eps, out.lo[1] - eps, out.lo[2] - eps, out.hi[0] + eps);
        nd.b = make_float4(out.hi[1] + eps, out.hi[2] + eps, 0.f, 0.f);
        nd.c = make_float4(0.f, 0.f, 0.f, 0.f);
        nd.r = make_uint4(0u, 0u, (uint32_t)n, kInvalid);
        out.nodes.ensure(sizeof(NodeDev));
        UPBP_HIP(hipMemcpyAsync(out.nodes.p, &nd, sizeof(nd), hipMemcpyHostToDevice, st));
        out.n_nodes = 1;
        UPBP_HIP(hipStreamSynchronize(st));
        return;
    }
    TmpBuf keys(ctx.tmp), keys2(ctx.tmp), vals(ctx.tmp);
    keys.ensure((size_t)n * 8);
    keys2.ensure((size_t)n * 8);
    vals.ensure((size_t)n * 4);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBlock), 0, st, n, box_lo, box_hi,
                       bounds.as<uint32_t>(), keys.as<uint64_t>(), vals.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    sort_pairs_u64(ctx, keys.as<uint64_t>(), keys2.as<uint64_t>(), vals.as<uint32_t>(),
                   out_order.as<uint32_t>(), n, 63);
    const int64_t ni = n - 1;
    TmpBuf left(ctx.tmp), right(ctx.tmp), first(ctx.tmp), last(ctx.tmp), leaf_parent(ctx.tmp),
        int_parent(ctx.tmp), ilo(ctx.tmp), ihi(ctx.tmp);
    left.ensure(ni * 4);
    right.ensure(ni * 4);
    first.ensure(ni * 4);
    last.ensure(ni * 4);
    leaf_parent.ensure(n * 4);
    int_parent.ensure(ni * 4);
    ilo.ensure(ni * 16);
    ihi.ensure(ni * 16);
    hipLaunchKernelGGL(k_karras, dim3(grid_for(ni)), dim3(kBlock), 0, st, n, keys2.as<uint64_t>(),
                       left.as<uint32_t>(), right.as<uint32_t>(), first.as<uint32_t>(),
                       last.as<uint32_t>(), leaf_parent.as<uint32_t>(), int_parent.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    {
        // group internal nodes by depth, then merge bottom-up one level per launch
        TmpBuf depth(ctx.tmp), depth_sorted(ctx.tmp), ids(ctx.tmp), ids_sorted(ctx.tmp), starts(ctx.tmp),
            err(ctx.tmp);
        depth.ensure(ni * 4);
        depth_sorted.ensure(ni * 4);
        ids.ensure(ni * 4);
        ids_sorted.ensure(ni * 4);
        starts.ensure(kMaxDepth * 4);
        err.ensure(4);
        launch_fill_u32(ctx, err.as<uint32_t>(), 0u, 1);
        hipLaunchKernelGGL(k_depth, dim3(grid_for(ni)), dim3(kBlock), 0, st, ni, int_parent.as<uint32_t>(),
                           depth.as<uint32_t>(), ids.as<uint32_t>(), err.as<uint32_t>());
        UPBP_HIP(hipGetLastError());
        sort_pairs_u32(ctx, depth.as<uint32_t>(), depth_sorted.as<uint32_t>(), ids.as<uint32_t>(),
                       ids_sorted.as<uint32_t>(), ni, 8);
        launch_fill_u32(ctx, starts.as<uint32_t>(), (uint32_t)ni, kMaxDepth);
        hipLaunchKernelGGL(k_level_starts, dim3(grid_for(ni)), dim3(kBlock), 0, st, ni,
                           depth_sorted.as<uint32_t>(), starts.as<uint32_t>());
        UPBP_HIP(hipGetLastError());
        uint32_t hs[kMaxDepth + 1], herr = 0, dmax = 0;
        UPBP_HIP(hipMemcpyAsync(hs, starts.p, kMaxDepth * 4, hipMemcpyDeviceToHost, st));
        UPBP_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, st));
        UPBP_HIP(hipMemcpyAsync(&dmax, depth_sorted.as<uint32_t>() + (ni - 1), 4, hipMemcpyDeviceToHost, st));
        UPBP_HIP(hipStreamSynchronize(st));
        UPBP_REQUIRE(herr == 0 && dmax < kMaxDepth, UPBP_EDEVICE, "LBVH deeper than expected");
        hs[dmax + 1] = (uint32_t)ni;
        for (int d = (int)dmax; d >= 0; --d) {
            const int64_t cnt = (int64_t)hs[d + 1] - hs[d];
            UPBP_REQUIRE(cnt > 0 && hs[d] < (uint32_t)ni, UPBP_EDEVICE, "LBVH level table inconsistent");
            hipLaunchKernelGGL(k_refit_level, dim3(grid_for(cnt)), dim3(kBlock), 0, st, cnt,
                               ids_sorted.as<uint32_t>() + hs[d], out_order.as<uint32_t>(), box_lo, box_hi,
                               left.as<uint32_t>(), right.as<uint32_t>(), ilo.as<float4>(), ihi.as<float4>());
            UPBP_HIP(hipGetLastError());
        }
    }
    out.nodes.ensure((size_t)ni * sizeof(NodeDev));
    out.n_nodes = ni;
    hipLaunchKernelGGL(k_emit, dim3(grid_for(ni)), dim3(kBlock), 0, st, n, leaf_size, eps,
                       out_order.as<uint32_t>(), box_lo, box_hi, left.as<uint32_t>(),
                       right.as<uint32_t>(), first.as<uint32_t>(), last.as<uint32_t>(),
                       ilo.as<float4>(), ihi.as<float4>(), out.nodes.as<NodeDev>());
    UPBP_HIP(hipGetLastError());
    UPBP_HIP(hipStreamSynchronize(st));
}

}  // namespace upbp
