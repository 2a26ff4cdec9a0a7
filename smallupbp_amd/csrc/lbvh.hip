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
__global__ void k_bounds_init(uint32_t* b) {
    int i = threadIdx.x;
    if (i < 12) b[i] = ((i / 3) % 2 == 0) ? 0xFFFFFFFFu : 0u;
}

__global__ __launch_bounds__(kBlock) void k_bounds(int64_t n, const float4* lo, const float4* hi,
                                                   uint32_t* b) {
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
    // wave reduction, one atomic per wave and value
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 12; ++k) {
            uint32_t o = __shfl_xor(v[k], off, 64);
            v[k] = ((k / 3) % 2 == 0) ? min(v[k], o) : max(v[k], o);
        }
    }
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 12; ++k) {
            if ((k / 3) % 2 == 0)
                atomicMin(&b[k], v[k]);
            else
                atomicMax(&b[k], v[k]);
        }
    }
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

// Bottom-up refit: every leaf climbs; the second thread to reach a node merges
// its two children (agent-scope acquire/release on the arrival counter).
__global__ __launch_bounds__(kBlock) void k_refit(int64_t n, const uint32_t* order, const float4* lo,
                                                  const float4* hi, const uint32_t* left,
                                                  const uint32_t* right, const uint32_t* leaf_parent,
                                                  const uint32_t* int_parent, uint32_t* arrive,
                                                  float4* ilo, float4* ihi) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t node = leaf_parent[k];
    while (true) {
        uint32_t prev = __hip_atomic_fetch_add(&arrive[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) return;
        float4 b0l, b0h, b1l, b1h;
        uint32_t c0 = left[node], c1 = right[node];
        if (c0 & 0x80000000u) {
            uint32_t it = order[c0 & 0x7FFFFFFFu];
            b0l = lo[it];
            b0h = hi[it];
        } else {
            b0l = ilo[c0];
            b0h = ihi[c0];
        }
        if (c1 & 0x80000000u) {
            uint32_t it = order[c1 & 0x7FFFFFFFu];
            b1l = lo[it];
            b1h = hi[it];
        } else {
            b1l = ilo[c1];
            b1h = ihi[c1];
        }
        float4 ml = make_float4(fminf(b0l.x, b1l.x), fminf(b0l.y, b1l.y), fminf(b0l.z, b1l.z), 0.f);
        float4 mh = make_float4(fmaxf(b0h.x, b1h.x), fmaxf(b0h.y, b1h.y), fmaxf(b0h.z, b1h.z), 0.f);
        ilo[node] = ml;
        ihi[node] = mh;
        if (node == 0) return;
        node = int_parent[node];
    }
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
                ref[c] = m;
                cnt[c] = 0;
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
    DevBuf bounds;
    bounds.ensure(12 * 4);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, bounds.as<uint32_t>());
    unsigned gb = std::min<unsigned>(grid_for(n), 2048u);
    hipLaunchKernelGGL(k_bounds, dim3(gb), dim3(kBlock), 0, st, n, box_lo, box_hi, bounds.as<uint32_t>());
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
    DevBuf keys, keys2, vals;
    keys.ensure((size_t)n * 8);
    keys2.ensure((size_t)n * 8);
    vals.ensure((size_t)n * 4);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBlock), 0, st, n, box_lo, box_hi,
                       bounds.as<uint32_t>(), keys.as<uint64_t>(), vals.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    sort_pairs_u64(ctx, keys.as<uint64_t>(), keys2.as<uint64_t>(), vals.as<uint32_t>(),
                   out_order.as<uint32_t>(), n, 63);
    const int64_t ni = n - 1;
    DevBuf left, right, first, last, leaf_parent, int_parent, arrive, ilo, ihi;
    left.ensure(ni * 4);
    right.ensure(ni * 4);
    first.ensure(ni * 4);
    last.ensure(ni * 4);
    leaf_parent.ensure(n * 4);
    int_parent.ensure(ni * 4);
    arrive.ensure(ni * 4);
    ilo.ensure(ni * 16);
    ihi.ensure(ni * 16);
    UPBP_HIP(hipMemsetAsync(arrive.p, 0, ni * 4, st));
    hipLaunchKernelGGL(k_karras, dim3(grid_for(ni)), dim3(kBlock), 0, st, n, keys2.as<uint64_t>(),
                       left.as<uint32_t>(), right.as<uint32_t>(), first.as<uint32_t>(),
                       last.as<uint32_t>(), leaf_parent.as<uint32_t>(), int_parent.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_refit, dim3(grid_for(n)), dim3(kBlock), 0, st, n, out_order.as<uint32_t>(),
                       box_lo, box_hi, left.as<uint32_t>(), right.as<uint32_t>(),
                       leaf_parent.as<uint32_t>(), int_parent.as<uint32_t>(), arrive.as<uint32_t>(),
                       ilo.as<float4>(), ihi.as<float4>());
    UPBP_HIP(hipGetLastError());
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
