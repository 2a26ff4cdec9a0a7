// build.hip -- photon and beam structure builds on the GPU.
//
// Photons (EmbreeBre::build, SmallUPBP/src/Bre/Bre.cxx:526-632): in-medium
// filter, radius (constant, per photon, or from the k-th nearest photon), r^2 = r*r, cube box pos +- r
// (Bre.cxx:174-177), LBVH, then photon records gathered into leaf order.
// Beams (PhotonBeamsEvaluator::build, SmallUPBP/src/Beams/PhBeams.cxx:66-148):
// radii as the reference (constant: start = end = r, change 0), then each beam
// is cut into pieces of about beam_piece_factor * r_max length (at most
// max_pieces_per_beam). Each piece gets the tight cone box of
// PhotonBeam::getSegmentAABB (PhotonBeam.hxx:186-198) and a half-open beam
// parameter range [t_lo, t_hi); the LBVH is built over pieces.
#include <algorithm>
#include <cmath>

#include "upbp_internal.hpp"

namespace upbp {

namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

__global__ void k_flags_u8(int64_t n, const uint8_t* in_medium, uint32_t* flags) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) flags[i] = in_medium[i] ? 1u : 0u;
}

__global__ void k_compact(int64_t n, const uint8_t* in_medium, const uint32_t* pos, uint32_t* keep) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && in_medium[i]) keep[pos[i]] = (uint32_t)i;
}

__global__ void k_photon_boxes(int64_t n, const uint32_t* keep, const float* position, const float* radius,
                               float const_radius, const float* power, const float* inc, float4* lo,
                               float4* hi, uint32_t* err) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    int64_t i = keep ? keep[k] : k;
    float px = position[3 * i], py = position[3 * i + 1], pz = position[3 * i + 2];
    float r = radius ? radius[i] : const_radius;
    uint32_t e = 0;
    if (!finite3(px, py, pz) || !__builtin_isfinite(r)) e |= kErrNonFinite;
    if (!finite3(power[3 * i], power[3 * i + 1], power[3 * i + 2])) e |= kErrNonFinite;
    if (inc && !finite3(inc[3 * i], inc[3 * i + 1], inc[3 * i + 2])) e |= kErrNonFinite;
    if (!(r > 0.f)) e |= kErrBadSegment;
    if (e) {
        atomicOr(err, e);
        px = py = pz = 0.f;
        r = 0.f;
    }
    lo[k] = make_float4(px - r, py - r, pz - r, 0.f);
    hi[k] = make_float4(px + r, py + r, pz + r, 0.f);
}

__global__ void k_photon_gather(int64_t n, const uint32_t* order, const uint32_t* keep,
                                const float* position, const float* radius, float const_radius,
                                const float* power, const float* inc, float4* pos_r2, float4* pw,
                                float4* inc_out, uint32_t* orig, float* radius_in) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = order[s];
    int64_t i = keep ? keep[k] : k;
    float r = radius ? radius[i] : const_radius;
    const float r2 = r * r;
    pos_r2[s] = make_float4(position[3 * i], position[3 * i + 1], position[3 * i + 2], r2);
    // kernel normalisation 1 / (r^2 pi / 2) of Bre.cxx:273, once per photon
    pw[s] = make_float4(power[3 * i], power[3 * i + 1], power[3 * i + 2], 1.0f / ((r2 * kPi) * 0.5f));
    if (inc)
        inc_out[s] = make_float4(inc[3 * i], inc[3 * i + 1], inc[3 * i + 2], 0.f);
    else
        inc_out[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    orig[s] = (uint32_t)i;
    radius_in[i] = r;
}

// new power in original input order -> leaf order
__global__ void k_photon_set_power(int64_t n, const uint32_t* orig, const float* power, float4* pw) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t i = orig[s];
    pw[s] = make_float4(power[3 * i], power[3 * i + 1], power[3 * i + 2], pw[s].w);
}

struct BeamIn {
    const float* origin;
    const float* direction;
    const float* length;
    const float* throughput;
    const uint32_t* flags;
    const int32_t* medium;
    const float* r0;
    const float* r1;
    float const_radius;
    float piece_factor;
    int max_pieces;
    int n_media;
};

__device__ __forceinline__ void beam_radii(const BeamIn& b, int64_t i, float& r0, float& r1, float& rmax,
                                           float& dr) {
    if (b.r0) {
        r0 = b.r0[i];
        r1 = b.r1 ? b.r1[i] : r0;
        rmax = fmaxf(r0, r1);
        dr = (r1 - r0) / b.length[i];  // mRadiusChange, PhBeams.cxx:108
    } else {
        r0 = r1 = rmax = b.const_radius;
        dr = 0.f;
    }
}

__global__ void k_beam_count(int64_t n, BeamIn b, uint32_t* count, uint32_t* err) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r0, r1, rmax, dr;
    beam_radii(b, i, r0, r1, rmax, dr);
    const float len = b.length[i];
    uint32_t e = 0;
    if (!finite3(b.origin[3 * i], b.origin[3 * i + 1], b.origin[3 * i + 2]) ||
        !finite3(b.direction[3 * i], b.direction[3 * i + 1], b.direction[3 * i + 2]) ||
        !finite3(b.throughput[3 * i], b.throughput[3 * i + 1], b.throughput[3 * i + 2]) ||
        !__builtin_isfinite(len) || !__builtin_isfinite(r0) || !__builtin_isfinite(r1))
        e |= kErrNonFinite;
    if (!(len > 0.f) || !(r0 > 0.f) || !(r1 > 0.f)) e |= kErrBadSegment;
    const uint32_t f = b.flags[i];
    if (f != UPBP_SHORT_BEAM && f != UPBP_LONG_BEAM) e |= kErrBadSegment;
    const int32_t m = b.medium[i];
    if (m < 0 || m >= b.n_media) e |= kErrBadMedium;
    uint32_t c = 1;
    if (e) {
        atomicOr(err, e);
    } else {
        float want = ceilf(len / (b.piece_factor * rmax));
        c = (uint32_t)fminf(fmaxf(want, 1.f), (float)b.max_pieces);
    }
    count[i] = c;
}

// Parameter of cut k of a beam cut into c pieces: exactly 0 at k = 0 and exactly len at k = c.
__device__ __forceinline__ float cut_at(float len, uint32_t k, uint32_t c) {
    return k >= c ? len : len * ((float)k / (float)c);
}

__global__ void k_beam_pieces(int64_t n, BeamIn b, const uint32_t* count, const uint32_t* offset,
                              PieceDev* pieces, float4* flux, uint32_t* piece_beam, float4* lo, float4* hi,
                              float* r_start, float* r_end) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r0, r1, rmax, dr;
    beam_radii(b, i, r0, r1, rmax, dr);
    r_start[i] = r0;
    r_end[i] = r1;
    const float ox = b.origin[3 * i], oy = b.origin[3 * i + 1], oz = b.origin[3 * i + 2];
    const float dx = b.direction[3 * i], dy = b.direction[3 * i + 1], dz = b.direction[3 * i + 2];
    const float len = b.length[i];
    // cone box margins r * sqrt(1 - d_a^2), PhotonBeam.hxx:74-80
    const float mx = sqrtf(dy * dy + dz * dz), my = sqrtf(dx * dx + dz * dz), mz = sqrtf(dx * dx + dy * dy);
    const uint32_t word = ((uint32_t)b.medium[i] & kMediumMask) | (b.flags[i] == UPBP_SHORT_BEAM ? kShortBeamBit : 0u);
    const uint32_t c = count[i];
    const uint32_t base = offset[i];
    const float d2o2 = dot3(dx, dy, dz, ox, oy, oz);  // dot(d2, O2) of PhotonBeam.hxx:278
    // The box of a piece holds every point within the beam's largest radius (1 % wider, for
    // the rounding of the closest-approach point) of its part of the beam; ownership of a
    // (query, beam) pair goes by that point's foot on the beam line (test_beam_beam), with
    // the first and last pieces owning whatever lies before and after them.
    const float rb = rmax * 1.01f;
    for (uint32_t k = 0; k < c; ++k) {
        const float t0 = cut_at(len, k, c), t1 = cut_at(len, k + 1, c);
        const float rs = rb, re = rb;
        const float sx = ox + dx * t0, sy = oy + dy * t0, sz = oz + dz * t0;
        const float ex = ox + dx * t1, ey = oy + dy * t1, ez = oz + dz * t1;
        const uint32_t p = base + k;
        lo[p] = make_float4(fminf(sx - rs * mx, ex - re * mx), fminf(sy - rs * my, ey - re * my),
                            fminf(sz - rs * mz, ez - re * mz), 0.f);
        hi[p] = make_float4(fmaxf(sx + rs * mx, ex + re * mx), fmaxf(sy + rs * my, ey + re * my),
                            fmaxf(sz + rs * mz, ez + re * mz), 0.f);
        PieceDev pc;
        pc.p0 = make_float4(ox, oy, oz, len);
        pc.p1 = make_float4(dx, dy, dz, r0);
        pc.p2 = make_float4(k == 0 ? -INFINITY : t0, k + 1 >= c ? INFINITY : t1, __uint_as_float(word), rmax * rmax);
        pc.p3 = make_float4(d2o2, dr, 0.f, 0.f);
        pieces[p] = pc;
        flux[p] = make_float4(b.throughput[3 * i], b.throughput[3 * i + 1], b.throughput[3 * i + 2], 0.f);
        piece_beam[p] = (uint32_t)i;
    }
}

__global__ void k_piece_gather(int64_t n, const uint32_t* order, const PieceDev* in, const float4* in_flux,
                               const uint32_t* in_beam, PieceDev* out, float4* out_flux, uint32_t* out_beam) {
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = order[s];
    out[s] = in[k];
    out_flux[s] = in_flux[k];
    out_beam[s] = in_beam[k];
}

// Copy host arrays to device staging buffers when needed.
template <class T>
const T* stage(Ctx& ctx, DevBuf& buf, const T* src, int64_t elems, upbp_memory mem) {
    if (!src) return nullptr;
    if (mem == UPBP_MEM_DEVICE) return src;
    buf.ensure((size_t)elems * sizeof(T));
    UPBP_HIP(hipMemcpyAsync(buf.p, src, (size_t)elems * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
    return buf.as<T>();
}

uint32_t read_u32(Ctx& ctx, const uint32_t* dev) {
    uint32_t v = 0;
    UPBP_HIP(hipMemcpyAsync(&v, dev, 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
    return v;
}

void check_err_word(uint32_t e) {
    UPBP_REQUIRE(!(e & kErrNonFinite), UPBP_ENONFINITE, "non-finite value in build input");
    UPBP_REQUIRE(!(e & kErrBadMedium), UPBP_EINVAL, "beam medium id outside the media table");
    UPBP_REQUIRE(!(e & kErrBadSegment), UPBP_EINVAL,
                 "invalid build input (radius or length <= 0, or beam flags not SHORT/LONG)");
}

}  // namespace

// Defaults of leaf size and piece length: the tuner's best on both the 1920x1080 workload
// (1M photons + 200k beams) and the 2M-beam stress workload; leaf 4 with 4 r pieces was
// 11-17 % slower on each.
constexpr int kDefaultLeaf = 8;
constexpr float kDefaultPieceFactor = 8.f;
// Largest leaf: the traversal checks a leaf's items against the wave's frustum one item per lane.
constexpr int kMaxLeaf = 64;

void bre_build(Ctx& ctx, const upbp_photon_input& in, const upbp_build_params& params, Bre& out) {
    UPBP_REQUIRE(in.count >= 0 && in.count < (int64_t)0x7FFFFFFF, UPBP_EINVAL, "photon count out of range");
    const bool knn = params.radius_calculation == UPBP_KNN_RADIUS;
    UPBP_REQUIRE(knn || params.radius_calculation == UPBP_CONSTANT_RADIUS, UPBP_EINVAL,
                 "radius_calculation must be UPBP_CONSTANT_RADIUS or UPBP_KNN_RADIUS");
    UPBP_REQUIRE(!knn || (params.knn >= 1 && !in.radius), UPBP_EINVAL,
                 "UPBP_KNN_RADIUS needs knn >= 1 and no per-photon radius array");
    UPBP_REQUIRE(in.radius || (params.radius > 0.f && std::isfinite(params.radius)), UPBP_EINVAL,
                 "photon radius must be > 0");
    const int leaf = params.leaf_size > 0 ? params.leaf_size : kDefaultLeaf;
    UPBP_REQUIRE(leaf >= 1 && leaf <= kMaxLeaf, UPBP_EINVAL, "leaf_size must be in 1..64");
    out.ctx = &ctx;
    out.n_in = in.count;
    out.n = 0;
    if (in.count == 0) return;
    UPBP_REQUIRE(in.position && in.power, UPBP_EINVAL, "photon position and power are required");
    const int64_t n_in = in.count;
    TmpBuf s_pos(ctx.tmp), s_inc(ctx.tmp), s_pow(ctx.tmp), s_rad(ctx.tmp), s_med(ctx.tmp);
    const float* pos = stage(ctx, s_pos, in.position, 3 * n_in, in.memory);
    const float* inc = stage(ctx, s_inc, in.inc_dir, 3 * n_in, in.memory);
    const float* pw = stage(ctx, s_pow, in.power, 3 * n_in, in.memory);
    const float* rad = stage(ctx, s_rad, in.radius, n_in, in.memory);
    const uint8_t* med = stage(ctx, s_med, in.in_medium, n_in, in.memory);
    hipStream_t st = ctx.stream;

    TmpBuf keep_buf(ctx.tmp);
    const uint32_t* keep = nullptr;
    int64_t n = n_in;
    if (med) {
        TmpBuf flags(ctx.tmp), pos_scan(ctx.tmp);
        flags.ensure(n_in * 4);
        pos_scan.ensure(n_in * 4);
        hipLaunchKernelGGL(k_flags_u8, dim3(grid_for(n_in)), dim3(kBlock), 0, st, n_in, med, flags.as<uint32_t>());
        exclusive_scan_u32(ctx, flags.as<uint32_t>(), pos_scan.as<uint32_t>(), n_in);
        uint32_t last_pos = read_u32(ctx, pos_scan.as<uint32_t>() + (n_in - 1));
        uint32_t last_flag = read_u32(ctx, flags.as<uint32_t>() + (n_in - 1));
        n = (int64_t)last_pos + last_flag;
        keep_buf.ensure(std::max<int64_t>(n, 1) * 4);
        hipLaunchKernelGGL(k_compact, dim3(grid_for(n_in)), dim3(kBlock), 0, st, n_in, med, pos_scan.as<uint32_t>(),
                           keep_buf.as<uint32_t>());
        keep = keep_buf.as<uint32_t>();
    }
    out.n = n;
    if (n == 0) return;
    TmpBuf knn_radius(ctx.tmp);
    if (knn) {
        // radius of each kept photon from its knn-th nearest kept photon (Bre.cxx:562-602)
        photon_knn_radius(ctx, n, pos, keep, params.radius, params.knn, knn_radius, n_in);
        rad = knn_radius.as<float>();
    }
    TmpBuf lo(ctx.tmp), hi(ctx.tmp), err(ctx.tmp), order(ctx.tmp);
    lo.ensure(n * 16);
    hi.ensure(n * 16);
    err.ensure(4);
    UPBP_HIP(hipMemsetAsync(err.p, 0, 4, st));
    hipLaunchKernelGGL(k_photon_boxes, dim3(grid_for(n)), dim3(kBlock), 0, st, n, keep, pos, rad, params.radius,
                       pw, inc, lo.as<float4>(), hi.as<float4>(), err.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    check_err_word(read_u32(ctx, err.as<uint32_t>()));
    build_lbvh(ctx, n, lo.as<float4>(), hi.as<float4>(), leaf, out.bvh, order);
    out.pos_r2.ensure(n * 16);
    out.power.ensure(n * 16);
    out.inc.ensure(n * 16);
    out.orig.ensure(n * 4);
    out.radius_in.ensure(n_in * 4);
    launch_clear_f32(ctx, out.radius_in.as<float>(), n_in);
    hipLaunchKernelGGL(k_photon_gather, dim3(grid_for(n)), dim3(kBlock), 0, st, n, order.as<uint32_t>(), keep,
                       pos, rad, params.radius, pw, inc, out.pos_r2.as<float4>(), out.power.as<float4>(),
                       out.inc.as<float4>(), out.orig.as<uint32_t>(), out.radius_in.as<float>());
    UPBP_HIP(hipGetLastError());
    UPBP_HIP(hipStreamSynchronize(st));
}

void bre_set_power(Bre& bre, const float* power_dev) {
    if (bre.n == 0) return;
    Ctx& ctx = *bre.ctx;
    hipLaunchKernelGGL(k_photon_set_power, dim3(grid_for(bre.n)), dim3(kBlock), 0, ctx.stream, bre.n,
                       bre.orig.as<uint32_t>(), power_dev, bre.power.as<float4>());
    UPBP_HIP(hipGetLastError());
}

void beams_build(Ctx& ctx, const upbp_beam_input& in, const upbp_build_params& params, Beams& out) {
    UPBP_REQUIRE(in.count >= 0 && in.count < (int64_t)0x7FFFFFFF, UPBP_EINVAL, "beam count out of range");
    const bool knn = params.radius_calculation == UPBP_KNN_RADIUS;
    UPBP_REQUIRE(knn || params.radius_calculation == UPBP_CONSTANT_RADIUS, UPBP_EINVAL,
                 "radius_calculation must be UPBP_CONSTANT_RADIUS or UPBP_KNN_RADIUS");
    UPBP_REQUIRE(!knn || (params.knn >= 1 && !in.start_radius && !in.end_radius), UPBP_EINVAL,
                 "UPBP_KNN_RADIUS needs knn >= 1 and no per-beam radius arrays");
    UPBP_REQUIRE(in.start_radius || (params.radius > 0.f && std::isfinite(params.radius)), UPBP_EINVAL,
                 "beam radius must be > 0");
    UPBP_REQUIRE(ctx.n_media > 0, UPBP_EINVAL, "upbp_set_media must be called before building beams");
    const int leaf = params.leaf_size > 0 ? params.leaf_size : kDefaultLeaf;
    UPBP_REQUIRE(leaf >= 1 && leaf <= kMaxLeaf, UPBP_EINVAL, "leaf_size must be in 1..64");
    const float factor = params.beam_piece_factor > 0.f ? params.beam_piece_factor : kDefaultPieceFactor;
    const int maxp = params.max_pieces_per_beam > 0 ? params.max_pieces_per_beam : 64;
    UPBP_REQUIRE(maxp <= 256 && std::isfinite(factor), UPBP_EINVAL, "max_pieces_per_beam must be in 1..256");
    out.ctx = &ctx;
    out.n = in.count;
    out.pieces = 0;
    out.grid_resolution = params.grid_resolution;
    if (in.count == 0) {
        beam_reduction_build(ctx, 0, nullptr, nullptr, nullptr, nullptr, nullptr, params, out);
        return;
    }
    UPBP_REQUIRE(in.origin && in.direction && in.length && in.throughput && in.flags && in.medium_id,
                 UPBP_EINVAL, "beam origin, direction, length, throughput, flags and medium_id are required");
    const int64_t n = in.count;
    TmpBuf s0(ctx.tmp), s1(ctx.tmp), s2(ctx.tmp), s3(ctx.tmp), s4(ctx.tmp), s5(ctx.tmp), s6(ctx.tmp), s7(ctx.tmp);
    BeamIn b;
    b.origin = stage(ctx, s0, in.origin, 3 * n, in.memory);
    b.direction = stage(ctx, s1, in.direction, 3 * n, in.memory);
    b.length = stage(ctx, s2, in.length, n, in.memory);
    b.throughput = stage(ctx, s3, in.throughput, 3 * n, in.memory);
    b.flags = stage(ctx, s4, in.flags, n, in.memory);
    b.medium = stage(ctx, s5, in.medium_id, n, in.memory);
    b.r0 = stage(ctx, s6, in.start_radius, n, in.memory);
    b.r1 = stage(ctx, s7, in.end_radius, n, in.memory);
    TmpBuf knn_r0(ctx.tmp), knn_r1(ctx.tmp);
    if (knn) {
        // start / end radius from the knn-th nearest beam origin (PhBeams.cxx:77-117)
        beam_knn_radius(ctx, n, b.origin, b.direction, b.length, params.radius, params.knn, knn_r0, knn_r1);
        b.r0 = knn_r0.as<float>();
        b.r1 = knn_r1.as<float>();
    }
    b.const_radius = params.radius;
    b.piece_factor = factor;
    b.max_pieces = maxp;
    b.n_media = ctx.n_media;
    hipStream_t st = ctx.stream;
    TmpBuf count(ctx.tmp), offset(ctx.tmp), err(ctx.tmp);
    count.ensure(n * 4);
    offset.ensure(n * 4);
    err.ensure(4);
    UPBP_HIP(hipMemsetAsync(err.p, 0, 4, st));
    hipLaunchKernelGGL(k_beam_count, dim3(grid_for(n)), dim3(kBlock), 0, st, n, b, count.as<uint32_t>(),
                       err.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    check_err_word(read_u32(ctx, err.as<uint32_t>()));
    exclusive_scan_u32(ctx, count.as<uint32_t>(), offset.as<uint32_t>(), n);
    const int64_t total = (int64_t)read_u32(ctx, offset.as<uint32_t>() + (n - 1)) +
                          read_u32(ctx, count.as<uint32_t>() + (n - 1));
    UPBP_REQUIRE(total < (int64_t)0x7FFFFFFF, UPBP_EINVAL, "too many beam pieces");
    TmpBuf pieces(ctx.tmp), flux(ctx.tmp), piece_beam(ctx.tmp), lo(ctx.tmp), hi(ctx.tmp), order(ctx.tmp);
    pieces.ensure(total * sizeof(PieceDev));
    flux.ensure(total * 16);
    piece_beam.ensure(total * 4);
    lo.ensure(total * 16);
    hi.ensure(total * 16);
    out.r_start.ensure(n * 4);
    out.r_end.ensure(n * 4);
    hipLaunchKernelGGL(k_beam_pieces, dim3(grid_for(n)), dim3(kBlock), 0, st, n, b, count.as<uint32_t>(),
                       offset.as<uint32_t>(), pieces.as<PieceDev>(), flux.as<float4>(), piece_beam.as<uint32_t>(),
                       lo.as<float4>(), hi.as<float4>(), out.r_start.as<float>(), out.r_end.as<float>());
    UPBP_HIP(hipGetLastError());
    build_lbvh(ctx, total, lo.as<float4>(), hi.as<float4>(), leaf, out.bvh, order);
    out.pieces = total;
    out.piece.ensure(total * sizeof(PieceDev));
    out.piece_flux.ensure(total * 16);
    out.piece_beam.ensure(total * 4);
    hipLaunchKernelGGL(k_piece_gather, dim3(grid_for(total)), dim3(kBlock), 0, st, total, order.as<uint32_t>(),
                       pieces.as<PieceDev>(), flux.as<float4>(), piece_beam.as<uint32_t>(), out.piece.as<PieceDev>(),
                       out.piece_flux.as<float4>(), out.piece_beam.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    beam_reduction_build(ctx, n, b.origin, b.direction, b.length, out.r_start.as<float>(), out.r_end.as<float>(),
                         params, out);
    UPBP_HIP(hipStreamSynchronize(st));
}

}  // namespace upbp
