// knn.hip -- k-th nearest neighbour squared distances on the GPU, for the
// KNN_RADIUS builds of photons (SmallUPBP/src/Bre/Bre.cxx:562-602) and beams
// (SmallUPBP/src/Beams/PhBeams.cxx:77-117).
//
// The reference answers each query with Jensen's kd-tree (Misc/KdTmpl.hxx:600-792):
// a bounded max-heap of k squared distances ((dx*dx) + dy*dy) + dz*dz with
// d = item - query, an item entering only below the heap maximum, the answer
// the heap maximum. For more than k items that is exactly the k-th smallest
// squared distance (the query's own item included), which is what these
// kernels compute; with n <= k items the caller asks for the n-th (largest).
//
// Layout: an LBVH over the points (degenerate boxes), points gathered into leaf
// order as float4. Queries are Morton-sorted so that one wavefront carries 64
// nearby queries; the wave walks ONE traversal (node records at wave-uniform
// addresses), each lane testing child boxes against its own current bound.
//  - k <= 16: each lane keeps its k best distances sorted in registers
//    (compile-time capacity K in {8, 16}; larger arrays leave registers), bound =
//    k-th best so far.
//  - k > 16: each lane bisects the float bit pattern of the answer; one
//    traversal per step counts the items with d2 <= mid and stops once k are found.
// Both are exact: a box's float distance never exceeds that of an item inside it
// (rounding is monotone and the box is the items' own coordinates widened), so
// pruning never drops an item that could change the answer.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "upbp_internal.hpp"

namespace upbp {

namespace {

constexpr int kBlock = 256;
constexpr int kWaveK = 64;
constexpr int kKnnStack = 128;  // >= LBVH depth (63 Morton bits + 31 index bits)
inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

__device__ __forceinline__ float d2_of(float px, float py, float pz, float qx, float qy, float qz) {
    const float dx = px - qx, dy = py - qy, dz = pz - qz;  // item - query, KdTmpl.hxx:604-610
    return (dx * dx + dy * dy) + dz * dz;
}

// Squared distance from q to a box, computed so that it never exceeds d2_of for
// an item inside the box: per axis the gap is (lo - q) or (q - hi) or 0.
__device__ __forceinline__ float box_d2(float qx, float qy, float qz, float lx, float ly, float lz, float hx,
                                        float hy, float hz) {
    const float gx = qx < lx ? lx - qx : (qx > hx ? qx - hx : 0.f);
    const float gy = qy < ly ? ly - qy : (qy > hy ? qy - hy : 0.f);
    const float gz = qz < lz ? lz - qz : (qz > hz ? qz - hz : 0.f);
    return (gx * gx + gy * gy) + gz * gz;
}

template <class T>
__device__ __forceinline__ T uni(T v) {
    return (T)__builtin_amdgcn_readfirstlane((int)v);
}

__global__ void k_point_boxes(int64_t n, const float* pts, const uint32_t* keep, float4* lo, float4* hi) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = keep ? keep[k] : k;
    const float4 p = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 0.f);
    lo[k] = p;
    hi[k] = p;
}

__global__ void k_gather_xyz(int64_t n, const float* pts, const uint32_t* keep, float* out) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = keep[k];
    out[3 * k] = pts[3 * i];
    out[3 * k + 1] = pts[3 * i + 1];
    out[3 * k + 2] = pts[3 * i + 2];
}

__global__ void k_point_gather(int64_t n, const uint32_t* order, const float4* in, float4* out) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < n) out[s] = in[order[s]];
}

// 30-bit Morton key of each query point inside the point bounds.
__global__ void k_point_keys(int64_t n, const float* q, float lx, float ly, float lz, float sx, float sy, float sz,
                             uint32_t* keys, uint32_t* idx) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float l[3] = {lx, ly, lz}, s[3] = {sx, sy, sz};
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
        float u = (q[3 * i + a] - l[a]) * s[a];
        u = __builtin_isfinite(u) ? fminf(fmaxf(u, 0.f), 1023.f) : 0.f;
        c[a] = (uint32_t)u;
    }
    uint32_t key = 0;
    for (int b = 9; b >= 0; --b)
        key = (key << 3) | (((c[0] >> b) & 1u) << 2) | (((c[1] >> b) & 1u) << 1) | ((c[2] >> b) & 1u);
    keys[i] = key;
    idx[i] = (uint32_t)i;
}

// Pushes/pops wave-uniform node indices; entry e lives in lane e % 64 of stk[e / 64].
struct WaveStack {
    uint32_t s0 = 0, s1 = 0;
    int sp = 0;
    __device__ __forceinline__ bool push(uint32_t v, int lane) {
        if (sp >= kKnnStack) return false;
        if (sp < kWaveK)
            s0 = (lane == sp) ? v : s0;
        else
            s1 = (lane == sp - kWaveK) ? v : s1;
        ++sp;
        return true;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        return sp < kWaveK ? (uint32_t)__builtin_amdgcn_readlane((int)s0, sp)
                           : (uint32_t)__builtin_amdgcn_readlane((int)s1, sp - kWaveK);
    }
};

// Walks the tree once for the wave. want(box_d2) says whether a lane still needs
// a box; visit(item) handles one leaf item for every lane that wanted the leaf.
// Returns false on stack overflow.
template <class Want, class Visit>
__device__ __forceinline__ bool walk(const NodeDev* nodes, float qx, float qy, float qz, Want want, Visit visit) {
    const int lane = threadIdx.x & (kWaveK - 1);
    WaveStack stk;
    uint32_t node = 0;
    while (true) {
        node = uni(node);
        const NodeDev* nd = nodes + node;
        const float4 a = nd->a, b = nd->b, c = nd->c;
        const uint4 rr = nd->r;
        const float d0 = rr.z != kInvalid ? box_d2(qx, qy, qz, a.x, a.y, a.z, a.w, b.x, b.y) : FLT_MAX;
        const float d1 = rr.w != kInvalid ? box_d2(qx, qy, qz, b.z, b.w, c.x, c.y, c.z, c.w) : FLT_MAX;
        const bool h0 = rr.z != kInvalid && want(d0);
        const bool h1 = rr.w != kInvalid && want(d1);
        const uint64_t b0 = __ballot(h0), b1 = __ballot(h1);
        if (b0 && !(rr.z & kInner))
            for (uint32_t i = rr.x; i < rr.x + rr.z; ++i) visit(i);
        if (b1 && !(rr.w & kInner))
            for (uint32_t i = rr.y; i < rr.y + rr.w; ++i) visit(i);
        const bool go0 = b0 && (rr.z & kInner), go1 = b1 && (rr.w & kInner);
        if (go0 && go1) {
            // nearer child first, by a vote of the lanes that want both
            const uint64_t both = b0 & b1;
            const uint64_t near0 = __ballot(h0 && h1 && d0 <= d1);
            const bool first0 = both == 0 ? (__popcll(b0) >= __popcll(b1)) : (2 * __popcll(near0) >= __popcll(both));
            if (!stk.push(first0 ? rr.y : rr.x, lane)) return false;
            node = first0 ? rr.x : rr.y;
        } else if (go0) {
            node = rr.x;
        } else if (go1) {
            node = rr.y;
        } else {
            if (stk.sp == 0) return true;
            node = stk.pop();
        }
    }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_knn_sorted(const NodeDev* __restrict__ nodes,
                                                       const float4* __restrict__ pts, int64_t nq,
                                                       const float* __restrict__ q, const uint32_t* __restrict__ perm,
                                                       int k, float init_d2, float* __restrict__ out,
                                                       uint32_t* __restrict__ overflow) {
    const int64_t slot = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
    const bool valid = slot < nq;
    const int64_t qi = valid ? (int64_t)perm[slot] : 0;
    const float qx = valid ? q[3 * qi] : 0.f, qy = valid ? q[3 * qi + 1] : 0.f, qz = valid ? q[3 * qi + 2] : 0.f;
    float best[K];
#pragma unroll
    for (int j = 0; j < K; ++j) best[j] = init_d2;
    float bound = valid ? init_d2 : -1.f;  // an invalid lane wants nothing
    const bool ok = walk(
        nodes, qx, qy, qz, [&](float bd) { return bd < bound; },
        [&](uint32_t i) {
            const float4 p = pts[i];
            const float d2 = d2_of(p.x, p.y, p.z, qx, qy, qz);
            if (d2 < bound) {
                float v = d2;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const float a = best[j];
                    best[j] = fminf(a, v);
                    v = fmaxf(a, v);
                }
                float nb = best[0];
#pragma unroll
                for (int j = 1; j < K; ++j) nb = (j == k - 1) ? best[j] : nb;
                bound = nb;
            }
        });
    if (!ok && (threadIdx.x & (kWaveK - 1)) == 0) atomicAdd(overflow, 1u);
    if (valid) out[qi] = bound;
}

// k > 16: per lane, the smallest float bit pattern v with #{d2 <= v} >= k.
__global__ __launch_bounds__(kBlock) void k_knn_bisect(const NodeDev* __restrict__ nodes,
                                                       const float4* __restrict__ pts, int64_t nq,
                                                       const float* __restrict__ q, const uint32_t* __restrict__ perm,
                                                       int k, float init_d2, float* __restrict__ out,
                                                       uint32_t* __restrict__ overflow) {
    const int64_t slot = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
    const bool valid = slot < nq;
    const int64_t qi = valid ? (int64_t)perm[slot] : 0;
    const float qx = valid ? q[3 * qi] : 0.f, qy = valid ? q[3 * qi + 1] : 0.f, qz = valid ? q[3 * qi + 2] : 0.f;
    // invariant: count(d2 <= hi) >= k (hi = the largest value an item may enter with),
    // count(d2 <= lo - 1) < k
    uint32_t lo = 0, hi = __float_as_uint(init_d2);
    bool ok = true;
    while (__ballot(valid && lo < hi)) {
        const bool active = valid && lo < hi;
        const uint32_t mid = lo + (hi - lo) / 2;
        const float v = __uint_as_float(mid);
        int cnt = 0;
        float lim = active ? v : -1.f;  // lanes drop out once k items are found
        ok = walk(
                 nodes, qx, qy, qz, [&](float bd) { return bd <= lim; },
                 [&](uint32_t i) {
                     const float4 p = pts[i];
                     const float d2 = d2_of(p.x, p.y, p.z, qx, qy, qz);
                     if (d2 <= lim) {
                         ++cnt;
                         if (cnt >= k) lim = -1.f;
                     }
                 }) &&
             ok;
        if (active) {
            if (cnt >= k)
                hi = mid;
            else
                lo = mid + 1;
        }
    }
    if (!ok && (threadIdx.x & (kWaveK - 1)) == 0) atomicAdd(overflow, 1u);
    if (valid) out[qi] = __uint_as_float(hi);
}

__global__ void k_photon_knn_radius(int64_t n, const uint32_t* keep, const float* d2, float radius, float* out) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = keep ? keep[k] : k;
    out[i] = radius * (2.0f * sqrtf(d2[k]));  // Bre.cxx:593
}

__global__ void k_beam_ends(int64_t n, const float* o, const float* d, const float* len, float* ends) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < 3; ++a) ends[3 * i + a] = o[3 * i + a] + d[3 * i + a] * len[i];  // Ray::target
}

__global__ void k_beam_knn_radius(int64_t n, const float* d2, float radius, float* out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = (2.0f * sqrtf(d2[i])) * radius;  // PhBeams.cxx:100,105
    out[i] = r < 0.001f ? 0.001f : r;                 // SMALLEST_RADIUS, PhBeams.cxx:72
}

}  // namespace

void KnnIndex::build(Ctx& ctx, int64_t n_points, const float* pts, const uint32_t* keep) {
    n = n_points;
    UPBP_REQUIRE(n > 0, UPBP_EINVAL, "k-nearest-neighbour radius needs at least one point");
    hipStream_t st = ctx.stream;
    DevBuf lo, hi, order;
    lo.ensure(n * 16);
    hi.ensure(n * 16);
    hipLaunchKernelGGL(k_point_boxes, dim3(grid_for(n)), dim3(kBlock), 0, st, n, pts, keep, lo.as<float4>(),
                       hi.as<float4>());
    UPBP_HIP(hipGetLastError());
    build_lbvh(ctx, n, lo.as<float4>(), hi.as<float4>(), 8, bvh, order);
    points.ensure(n * 16);
    hipLaunchKernelGGL(k_point_gather, dim3(grid_for(n)), dim3(kBlock), 0, st, n, order.as<uint32_t>(),
                       lo.as<float4>(), points.as<float4>());
    UPBP_HIP(hipGetLastError());
}

void KnnIndex::kth_d2(Ctx& ctx, int64_t nq, const float* q, int k, float* out) {
    UPBP_REQUIRE(k >= 1, UPBP_EINVAL, "knn must be >= 1");
    if (nq == 0) return;
    hipStream_t st = ctx.stream;
    const int keff = (int64_t)k < n ? k : (int)n;
    // query order: Morton keys within the point bounds
    DevBuf keys, keys2, idx, perm, ovf;
    keys.ensure(nq * 4);
    keys2.ensure(nq * 4);
    idx.ensure(nq * 4);
    perm.ensure(nq * 4);
    ovf.ensure(4);
    float sc[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = bvh.hi[a] - bvh.lo[a];
        sc[a] = ext > 0.f ? 1023.f / ext : 0.f;
    }
    hipLaunchKernelGGL(k_point_keys, dim3(grid_for(nq)), dim3(kBlock), 0, st, nq, q, bvh.lo[0], bvh.lo[1],
                       bvh.lo[2], sc[0], sc[1], sc[2], keys.as<uint32_t>(), idx.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    sort_pairs_u32(ctx, keys.as<uint32_t>(), keys2.as<uint32_t>(), idx.as<uint32_t>(), perm.as<uint32_t>(), nq, 30);
    launch_fill_u32(ctx, ovf.as<uint32_t>(), 0u, 1);
    // initial search radius sqrt(FLT_MAX), squared (KdTmpl.hxx:139 with Misc/Defs.hxx:48)
    const float init_r = std::sqrt(FLT_MAX);
    const float init_d2 = init_r * init_r;
    const NodeDev* nd = bvh.nodes.as<NodeDev>();
    const float4* pp = points.as<float4>();
    const uint32_t* pm = perm.as<uint32_t>();
    uint32_t* ov = ovf.as<uint32_t>();
    const dim3 g(grid_for(nq)), b(kBlock);
    if (keff <= 8)
        hipLaunchKernelGGL(k_knn_sorted<8>, g, b, 0, st, nd, pp, nq, q, pm, keff, init_d2, out, ov);
    else if (keff <= 16)
        hipLaunchKernelGGL(k_knn_sorted<16>, g, b, 0, st, nd, pp, nq, q, pm, keff, init_d2, out, ov);
    else
        hipLaunchKernelGGL(k_knn_bisect, g, b, 0, st, nd, pp, nq, q, pm, keff, init_d2, out, ov);
    UPBP_HIP(hipGetLastError());
    uint32_t h = 0;
    UPBP_HIP(hipMemcpyAsync(&h, ovf.p, 4, hipMemcpyDeviceToHost, st));
    UPBP_HIP(hipStreamSynchronize(st));
    UPBP_REQUIRE(h == 0, UPBP_EDEVICE, "k-nearest-neighbour traversal stack overflow");
}

void photon_knn_radius(Ctx& ctx, int64_t n_kept, const float* pos, const uint32_t* keep, float radius, int knn,
                       DevBuf& out_radius, int64_t n_in) {
    KnnIndex idx;
    idx.build(ctx, n_kept, pos, keep);
    DevBuf qpts, d2;
    const float* q = pos;
    if (keep) {
        // the queries are the kept photons themselves, in kept order
        qpts.ensure(n_kept * 12);
        hipLaunchKernelGGL(k_gather_xyz, dim3(grid_for(n_kept)), dim3(kBlock), 0, ctx.stream, n_kept, pos, keep,
                           qpts.as<float>());
        UPBP_HIP(hipGetLastError());
        q = qpts.as<float>();
    }
    d2.ensure(n_kept * 4);
    idx.kth_d2(ctx, n_kept, q, knn, d2.as<float>());
    out_radius.ensure(n_in * 4);
    launch_clear_f32(ctx, out_radius.as<float>(), n_in);
    hipLaunchKernelGGL(k_photon_knn_radius, dim3(grid_for(n_kept)), dim3(kBlock), 0, ctx.stream, n_kept, keep,
                       d2.as<float>(), radius, out_radius.as<float>());
    UPBP_HIP(hipGetLastError());
}

void beam_knn_radius(Ctx& ctx, int64_t n, const float* origin, const float* dir, const float* length, float radius,
                     int knn, DevBuf& r0, DevBuf& r1) {
    KnnIndex idx;
    idx.build(ctx, n, origin, nullptr);
    DevBuf ends, d2;
    ends.ensure(n * 12);
    d2.ensure(n * 4);
    r0.ensure(n * 4);
    r1.ensure(n * 4);
    hipLaunchKernelGGL(k_beam_ends, dim3(grid_for(n)), dim3(kBlock), 0, ctx.stream, n, origin, dir, length,
                       ends.as<float>());
    UPBP_HIP(hipGetLastError());
    idx.kth_d2(ctx, n, origin, knn, d2.as<float>());
    hipLaunchKernelGGL(k_beam_knn_radius, dim3(grid_for(n)), dim3(kBlock), 0, ctx.stream, n, d2.as<float>(), radius,
                       r0.as<float>());
    UPBP_HIP(hipGetLastError());
    idx.kth_d2(ctx, n, ends.as<float>(), knn, d2.as<float>());
    hipLaunchKernelGGL(k_beam_knn_radius, dim3(grid_for(n)), dim3(kBlock), 0, ctx.stream, n, d2.as<float>(), radius,
                       r1.as<float>());
    UPBP_HIP(hipGetLastError());
}

}  // namespace upbp
