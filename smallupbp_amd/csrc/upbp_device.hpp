// upbp_device.hpp -- device-side records and estimator mathematics for gfx950.
//
// Everything that decides a hit (the photon-disc test and the beam-beam
// closest-approach test) is written in the reference's operation order and the
// file is compiled with -ffp-contract=off, so the GPU hit decision is bit-identical
// to an IEEE FP32 evaluation of SmallUPBP/src/Bre/Bre.cxx:208-222 and
// SmallUPBP/src/Beams/PhotonBeam.hxx:253-295. Contributions may use FMA and the
// hardware exp2 (their error is far below the 1e-4 parity tolerance).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace upbp {

constexpr float kPi = 3.141592653589793238462643383279f;  // PI_F, Defs.hxx:74
constexpr float kInvPi = 1.f / kPi;                         // INV_PI_F, Defs.hxx:77
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kShortBeamBit = 1u << 24;  // piece word: medium id in bits 0-23, SHORT flag bit 24
constexpr uint32_t kMediumMask = 0x00FFFFFFu;
constexpr uint32_t kBadLightIndex = 16u;  // error bit: MIS light-vertex / path index out of range

// Device medium, 64 bytes. Derived like HomogeneousMedium's constructor
// (SmallUPBP/src/Scene/Media.hxx:129-159).
struct alignas(16) MediumDev {
    float st[3];       // sigma_t = sigma_a + sigma_s
    float min_comp;    // minimum positive sigma_t component
    float ss[3];       // sigma_s
    float g;           // mean cosine
    float sa[3];       // sigma_a
    float cont;        // continuation probability
    int32_t min_idx;   // index of min_comp
    int32_t has_scattering;
    int32_t isotropic; // |g| < 1e-3 (Media.hxx:111)
    float mfp;         // mean free path 1 / max sigma_t, or +inf (Media.hxx:158)
};
static_assert(sizeof(MediumDev) == 64, "MediumDev layout");

// LBVH node: two child boxes and references, 64 bytes, read with scalar loads.
//   a = (c0.lo.xyz, c0.hi.x)  b = (c0.hi.yz, c1.lo.xy)  c = (c1.lo.z, c1.hi.xyz)
//   r = (ref0, ref1, cnt0, cnt1): cnt == kInvalid = no child; cnt without kInner = leaf of cnt
//       items starting at ref; cnt with kInner = inner node ref whose subtree holds the
//       cnt & kSpanMask items (saturated below kSpanMask) that start at ref, or that end at ref
//       when kSpanEnds is set: the item range of Karras node i starts or ends at item i.
struct alignas(16) NodeDev {
    float4 a, b, c;
    uint4 r;
};
static_assert(sizeof(NodeDev) == 64, "NodeDev layout");
constexpr uint32_t kInner = 0x80000000u, kSpanEnds = 0x40000000u, kSpanMask = 0x3FFFFFFFu;

// Photon record in leaf order: pos_r2 = (pos.xyz, r^2), power = (rgb, 1 / (r^2 pi / 2)),
// inc = (WorldDirFix.xyz, unused) read only for anisotropic media.
// Beam piece record (64 bytes) in leaf order; the throughput lives in a separate
// array (piece_flux) because it is read only when some lane of the wave hits:
//   p0 = (origin.xyz, length)  p1 = (dir.xyz, start radius)
//   p2 = (t_lo, t_hi, word, r^2 max)  p3 = (dot(dir, origin), radius change, 0, 0)
struct alignas(16) PieceDev {
    float4 p0, p1, p2, p3;
};
static_assert(sizeof(PieceDev) == 64, "PieceDev layout");

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    // (x*x' + y*y') + z*z', the order of _mm_dp_ps(.., 0x71) (Vector3.hxx:518-526)
    return (ax * bx + ay * by) + az * bz;
}

// EmbreePhoton::testIntersectionBre (Bre.cxx:199-223).
__device__ __forceinline__ bool test_bre(float ox, float oy, float oz, float dx, float dy, float dz,
                                         float min_t, float max_t, float px, float py, float pz,
                                         float r2, float& t_out, float& s_out) {
    const float t = dot3(px - ox, py - oy, pz - oz, dx, dy, dz);
    if (t > min_t && t < max_t) {
        const float wx = (ox + dx * t) - px;
        const float wy = (oy + dy * t) - py;
        const float wz = (oz + dz * t) - pz;
        const float s = dot3(wx, wy, wz, wx, wy, wz);
        if (s <= r2) {
            t_out = t;
            s_out = s;
            return true;
        }
    }
    return false;
}

// Relative margin inside which the radius test falls back to the reference's exact
// square root and division: the hardware reciprocal square root is within 1 ulp, so
// |ad| * rsq(s2) is within a few ulp (< 1e-6) of fl(|ad| / fl(sqrt(s2))).
constexpr float kRadiusTie = 1e-5f;

// Kernels 1 - x (x = d^2 / r^2 of a hit) closer to 0 than this are evaluated with the
// reference's exact operations: near the radius the subtraction amplifies the few-ulp
// difference of the hardware reciprocal (square root) by 1 / (1 - x), which above this
// bound stays below 3e-5 relative.
constexpr float kGrazing = 0.02f;

// 1 - num / den, the 2D and 1D kernels' radial factor (Bre.cxx:272, PhotonBeam.hxx:364):
// through the hardware reciprocal, or the reference's IEEE division near the radius.
__device__ __forceinline__ float radial_kernel(float num, float den) {
    const float k = 1.0f - num * __builtin_amdgcn_rcpf(den);
    return k < kGrazing ? 1.0f - num / den : k;
}

// PhotonBeam::testIntersectionBeamBeam (PhotonBeam.hxx:238-296) with
// minT2 = 0 and maxT2 = length, the radius test of PhotonBeam::accumulate
// (radius = r0 + t2 * dr, hit unless radius < dist, PhotonBeam.hxx:330-334), and the
// piece's half-open range [t_lo, t_hi) that makes a beam cut into pieces count once.
// The range is tested on the foot of the query's closest-approach point on the beam line,
// tw = d2.(o1 - o2) + t1 q, not on the reference's t2 = ((t1 + d1.o1) - d1.o2) / q: for
// nearly perpendicular lines (q -> 0) that quotient carries an error of about 1e-7 / |q|
// of the scene scale, so its piece can lie far from the point the ray actually passes
// (a piece the traversal rightly culls). tw is well conditioned and the same for every
// piece of the beam, so exactly one piece owns each pair; the beam's first and last
// pieces own everything before and after them. The hit decision itself stays the
// reference's, on its own t2. d1o1 = dot3(d1, o1) and d2o2 = dot3(d2, o2) are passed in
// precomputed (same operations, same values). The radius test decides with the exact
// dist = |ad| / sqrt(s2) of the reference; dist and sin_theta are returned from the
// reciprocal square root unless the test needed the exact values.
__device__ __forceinline__ bool test_beam_beam(float o1x, float o1y, float o1z, float d1x, float d1y,
                                               float d1z, float d1o1, float min_t1, float max_t1, float o2x,
                                               float o2y, float o2z, float d2x, float d2y, float d2z,
                                               float d2o2, float len, float max_dist_sqr, float t_lo,
                                               float t_hi, float r0, float dr, float& dist, float& sin_theta,
                                               float& t1_out, float& t2_out, float& radius_out) {
    const float cx = d1y * d2z - d1z * d2y;
    const float cy = d1z * d2x - d1x * d2z;
    const float cz = d1x * d2y - d1y * d2x;
    const float s2 = dot3(cx, cy, cz, cx, cy, cz);
    const float ad = dot3(o2x - o1x, o2y - o1y, o2z - o1z, cx, cy, cz);
    if (ad * ad >= max_dist_sqr * s2) return false;
    const float q = dot3(d1x, d1y, d1z, d2x, d2y, d2z);
    const float q2m1 = q * q - 1.0f;
    if (q2m1 < 1e-5f && q2m1 > -1e-5f) return false;
    const float d1o2 = dot3(d1x, d1y, d1z, o2x, o2y, o2z);
    const float d2oo = dot3(d2x, d2y, d2z, o1x, o1y, o1z) - d2o2;
    const float t1 = ((d1o1 - d1o2) - q * d2oo) / q2m1;
    if (t1 <= min_t1 || t1 >= max_t1) return false;
    const float tw = d2oo + t1 * q;
    if (!(tw >= t_lo && tw < t_hi)) return false;
    const float t2 = ((t1 + d1o1) - d1o2) / q;
    if (t2 <= 0.0f || t2 >= len || __builtin_isnan(t2)) return false;
    const float radius = r0 + t2 * dr;
    const float aad = __builtin_fabsf(ad);
    const float rs = __builtin_amdgcn_rsqf(s2);
    float d = aad * rs, st = s2 * rs;
    // within 1 % of the radius (where the kernel 1 - d^2 / r^2 drops below ~kGrazing), or at
    // the tie, the reference's exact sqrt and division decide and give the returned values.
    // One exit for both outcomes: separate early returns make the compiler copy the five
    // outputs through every exit block.
    bool ok = true;
    if (!(radius > d * 1.0101f)) {
        ok = !(radius < d * (1.0f - kRadiusTie));
        st = __builtin_sqrtf(s2);
        d = aad / st;
        ok = ok && !(radius < d);
    }
    dist = d;
    sin_theta = st;
    t1_out = t1;
    t2_out = t2;
    radius_out = radius;
    return ok;
}

// exp(x) for x <= 0 through the hardware base-2 exponential.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// PhaseFunction::Pdf (PhaseFunction.hxx:52-80); cos_theta = -dot(gen, fix). It only
// scales contributions and weights (no hit decision), so dd^(-3/2) is the hardware
// reciprocal times the hardware reciprocal square root.
__device__ __forceinline__ float phase_pdf(float g, bool isotropic, float cos_theta) {
    const float uni = kInvPi * 0.25f;
    if (isotropic) return uni;
    const float sq = g * g;
    const float dd = (1.0f + sq) - (g + g) * cos_theta;
    return dd > 0.0f ? ((uni * (1.0f - sq)) * __builtin_amdgcn_rcpf(dd)) * __builtin_amdgcn_rsqf(dd) : 0.0f;
}

__device__ __forceinline__ uint32_t hash32(uint32_t i) {
    // "lowbias32" integer hash (C. Wellons) of i + 1; also used by the CPU checker
    uint32_t x = i + 1u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

}  // namespace upbp
