// traverse.hip -- query traversal kernels for gfx950 (wave64).
//
// One wavefront carries a bundle of 64 query beams (one per lane). The wave
// walks ONE traversal of the LBVH: every node record is fetched once per wave
// from a wave-uniform address (scalar loads), each lane slab-tests its own ray
// segment against the two child boxes, and __ballot decides where the wave
// goes. Leaf items (photons or beam pieces) are also fetched once per wave and
// tested by every lane against its own ray. The traversal stack holds
// wave-uniform node indices, one entry per lane register (two VGPRs = 128
// entries), pushed with a lane-select and popped with readlane.
//
// Per query the segment loop of EmbreeBre::evalBre (SmallUPBP/src/Bre/Bre.cxx:812-879,
// lite :895-964) and PhotonBeamsEvaluator::evalBeamBeamEstimate
// (SmallUPBP/src/Beams/PhBeams.cxx:177-241, lite :257-324) runs in registers:
// result += P * segment_result, P *= attenuation (/ pdf for SHORT queries), stop
// when a channel of P is <= 1e-20. Per hit the plain callbacks are evaluated:
// breIntersectFuncHomogeneous (Bre.cxx:239-277) and PhotonBeam::accumulate
// (SmallUPBP/src/Beams/PhotonBeam.hxx:315-367). The adjoint mode re-traverses
// and recomputes transmittance per hit instead of storing it.
#include <algorithm>

#include "upbp_internal.hpp"
#include "upbp_mis.hpp"

namespace upbp {

namespace {

constexpr int kWave = 64;
constexpr int kStackMax = 128;
constexpr int kEstPhotons = 0, kEstBeams = 1;
constexpr int kMaxLdsMedia = 8;  // adjoint: per-lane medium gradient slabs held in LDS

struct EvalArgs {
    // structure
    const NodeDev* __restrict__ nodes;
    const float4* __restrict__ pos_r2;
    const float4* __restrict__ power;
    const float4* __restrict__ inc;
    const uint32_t* __restrict__ orig;  // photon original index / piece -> beam index
    const PieceDev* __restrict__ pieces;
    const float4* __restrict__ piece_flux;
    // queries
    const float* __restrict__ origin;
    const float* __restrict__ direction;
    const int64_t* __restrict__ offsets;
    const upbp_volume_segment* __restrict__ segments;
    const uint32_t* __restrict__ perm;
    int64_t count;
    int32_t short_query;
    int32_t lite;
    // media
    const MediumDev* __restrict__ media;
    int32_t n_media;
    int32_t media_base;
    int32_t lds_media_n;  // adjoint: media with an LDS gradient slab (min(n_media, kMaxLdsMedia))
    // outputs
    float* __restrict__ out;
    uint32_t* __restrict__ sig;
    unsigned long long* __restrict__ stats;
    // adjoint
    const float* __restrict__ grad_out;
    float* __restrict__ grad_media;  // [M][6]
    float* __restrict__ grad_items;  // [items][3]
    // UPBP mode
    const upbp_query_mis* __restrict__ qmis;
    MisGlobals mis;
    // traversal starts here (0, or an instance's root in a batched plan's merged tree array)
    uint32_t root;
    // a wave with a frustum tests an inner node's subtree of at most this many items as one leaf
    uint32_t frustum_leaf;
    // per-cell beam reduction of the beam estimators (red.max == 0: none)
    RedDev red;
};

__device__ __forceinline__ float guard_inv(float d) {
    const float a = __builtin_fabsf(d) < 1e-20f ? __builtin_copysignf(1e-20f, d) : d;
    return __builtin_amdgcn_rcpf(a);
}

struct Ray {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;     // guarded reciprocal direction
    float oix, oiy, oiz;  // o * inv
    float d1o1;           // dot3(d, o), loop invariant of the beam test
    uint64_t qkey;        // per-cell beam reduction: key of the query ray (reduced kernels only)
};

// Wave mask of the lanes whose segment [tmin, tmax] meets the box:
// max(tn, tmin) <= min(tf, tmax) as three compares, each balloted on its own. min / max of
// tmin and tmax (values the compiler cannot prove canonical) would put a canonicalising max
// in front of each use at every node, and a ballot of a combined bool first selects it into
// a VGPR and compares it back; the slab values are fma results and need neither.
__device__ __forceinline__ uint64_t slab_mask(const Ray& r, float tmin, float tmax, float lx, float ly, float lz,
                                              float hx, float hy, float hz) {
    const float ax = __builtin_fmaf(lx, r.ix, -r.oix), bx = __builtin_fmaf(hx, r.ix, -r.oix);
    const float ay = __builtin_fmaf(ly, r.iy, -r.oiy), by = __builtin_fmaf(hy, r.iy, -r.oiy);
    const float az = __builtin_fmaf(lz, r.iz, -r.oiz), bz = __builtin_fmaf(hz, r.iz, -r.oiz);
    const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    return __ballot(tn <= tf) & __ballot(tn <= tmax) & __ballot(tf >= tmin);
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// A wave-uniform global pointer read from memory, made a scalar: the two halves pass
// through readfirstlane and the result is tagged as global memory, so loads through it
// use a scalar base (as kernel-argument pointers do) instead of a 64-bit VGPR pair.
// 64-bit readfirstlane from two 32-bit halves. readfirstlane returns int: each half
// goes through uint32_t before widening, or a low word with bit 31 set would
// sign-extend over the high word.
__device__ __forceinline__ uint64_t uni_u64(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

template <class T>
__device__ __forceinline__ T* uni_global(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)uni_u64((uint64_t)p);
}

__device__ __forceinline__ int64_t uni64(int64_t x) { return (int64_t)uni_u64((uint64_t)x); }

// Lane state of one segment being gathered.
struct SegState {
    float tmin, tmax;
    float st0, st1, st2;
    float ss0, ss1, ss2;
    float g;
    int32_t iso;
    int32_t kidx;
    int32_t medium;     // medium id (instance-local, for the beam medium test)
    float min_comp, mfp, cont;  // UPBP mode: min sigma_t component, mean free path, continuation
    float acc0, acc1, acc2;
    // adjoint per segment: w (upstream) * prefix, and accumulators
    float w0, w1, w2;
    float gs0, gs1, gs2;  // d/d sigma_s direct
    float gt0, gt1, gt2;  // d/d sigma_t
};

// A MODE of Forward or Signature with this bit set applies the per-cell beam reduction
// (a separate instantiation: the plain kernels keep their registers).
constexpr int kReduced = 8;

template <int MODE>
struct LaneOut {
    uint32_t hits = 0;
    uint32_t sig_n = 0, sig_h = 0;
    uint32_t err = 0;
    MisLane ml;               // UPBP mode only
    uint32_t qh = 0, qt = 0;  // UPBP mode: head and tail of the wave's hit queue (wave-uniform)
};

__device__ __forceinline__ float pick3(int k, float a, float b, float c) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// v + DPP-moved v; lanes outside row_mask add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
    const int m = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false);
    return v + __int_as_float(m);
}

// Sum over the 64 lanes of a wave with DPP row operations (all lanes must be active).
// quad_perm(1,0,3,2) and (2,3,0,1) sum quads, row_ror 4 and 8 sum rows of 16,
// row_bcast15 / row_bcast31 carry row sums upward; lane 63 ends with the total.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x124, 0xF>(v);  // row_ror:4
    v = dpp_add<0x128, 0xF>(v);  // row_ror:8
    v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// ---- leaf visitors ---------------------------------------------------------
// Exponent of the transmittance of channel c along `dist` with the SHORT
// normalisation T_c / T_k folded in: exp(-st_c d + st_k d) = T_c(d) / T_k(d).
__device__ __forceinline__ float neg_tau(float st_c, float st_k, float d, bool short_norm) {
    return short_norm ? (st_k * d - st_c * d) : -(st_c * d);
}

// Sums the three channels' item gradients over the wave (DPP) and adds them once.
__device__ __forceinline__ void add_item_gradient(float* grad, uint32_t o, float g0, float g1, float g2) {
    const float r0 = wave_sum_dpp(g0), r1 = wave_sum_dpp(g1), r2 = wave_sum_dpp(g2);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unsafeAtomicAdd(&grad[3 * o + 0], r0);
        unsafeAtomicAdd(&grad[3 * o + 1], r1);
        unsafeAtomicAdd(&grad[3 * o + 2], r2);
    }
}

__device__ __forceinline__ MisF merge_vertex_mis(const upbp_mis_params& P, float pdf_a_inv, float ray_pdf_inv,
                                                 float ratio, float bb1d_sel, uint32_t flags) {
    MisF m = {};
    m.pdf_a_inv = pdf_a_inv;
    m.ray_sample_pdf_inv = ray_pdf_inv;
    m.ray_sample_pdfs_ratio = ratio;
    m.surf = 0.f;
    m.pp3d = P.pp3d_factor;
    m.pb2d = P.pb2d_factor;
    m.bb1d = P.bb1d_factor;
    m.bb1d_sel = bb1d_sel;
    m.flags = flags;
    return m;
}

// breIntersectFuncHomogeneous2 after the disc test (Bre.cxx:305-463), per lane; the
// photon's light vertex is read by index (lanes hit different photons only through
// the same wave-uniform photon, so this is one gather per hit).
__device__ void photon_hit_mis(const EvalArgs& A, const Ray& r, SegState& s, const MisLane& ml, float t, float sq,
                               float r2, uint32_t lv_idx, bool short_q, uint32_t& err) {
    const MisGlobals& G = A.mis;
    const upbp_mis_params& P = G.p;
    if ((int64_t)lv_idx >= G.lv_count) {
        err |= kBadLightIndex;
        return;
    }
    // the vertex's record: its fields, and those of the walk's index 1 with its tail, in one read
    const PhotonRec rec = G.photon_recs[lv_idx];
    const uint32_t path_len = rec.d.x & kRecLengthMask;
    const uint32_t cam_len = (uint32_t)ml.cam_len;
    if ((path_len + cam_len > P.max_path_length) || (path_len + cam_len < P.min_path_length)) return;
    if (cam_len == 1 && t < P.min_dist_to_med) return;
    // The weights and contributions below need 1e-4 relative accuracy, not the reference's
    // rounding (hit decisions are all made before this point), so reciprocals and square
    // roots are the single hardware instructions.
    const float dt = t - s.tmin;
    float a0 = fast_exp(-s.st0 * dt), a1 = fast_exp(-s.st1 * dt), a2 = fast_exp(-s.st2 * dt);
    const float pdf = pick3(s.kidx, a0, a1, a2);
    if (short_q) {
        const float ip = __builtin_amdgcn_rcpf(pdf);
        a0 *= ip;
        a1 *= ip;
        a2 *= ip;
    }
    float ray_pdf = s.min_comp * pdf;
    float ray_rev = (ml.seg_flags & UPBP_ORIGIN_IN_MEDIUM) ? ray_pdf : pdf;
    const float ratio = __builtin_amdgcn_rcpf(s.min_comp);
    if (!(a0 > 1e-20f && a1 > 1e-20f && a2 > 1e-20f)) return;
    ray_pdf *= ml.pref_pdf;
    ray_rev *= ml.pref_rev;
    // PhaseFunction::Evaluate(-ray.dir, lightDirection): cos = -dot(gen, fix) = dot(dir_fix, d)
    const float lx = rec.b.x, ly = rec.b.y, lz = rec.b.z;
    const float cos_t = -dot3(lx, ly, lz, -r.dx, -r.dy, -r.dz);
    const float bsdf = phase_pdf(s.g, s.iso != 0, cos_t);
    const float sin_t = __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - cos_t * cos_t));
    if (bsdf <= 0.f) return;
    const float dir_pdf = bsdf * s.cont;
    const float rev_pdf = bsdf * rec.a.w;
    const float kernel = radial_kernel(sq, r2) * __builtin_amdgcn_rcpf((r2 * kPi) * 0.5f);
    if (!(kernel > 1e-20f)) return;
    const float u0 = (((rec.a.x * a0) * s.ss0) * bsdf) * kernel;
    const float u1 = (((rec.a.y * a1) * s.ss1) * bsdf) * kernel;
    const float u2 = (((rec.a.z * a2) * s.ss2) * bsdf) * kernel;
    if (u0 <= 0.f && u1 <= 0.f && u2 <= 0.f) return;
    const float dist_sq = t * t;
    const float ray_pdf_inv = __builtin_amdgcn_rcpf(ray_pdf);
    // Bre.cxx:407-415: the photon-beam evaluator's getBeamSelectionPdf(isectPt) when it reduces
    // its cells (A.red is then its grid), 1 when it does not, 0 without photon beams
    float sel = P.has_bb1d_beams ? 1.f : 0.f;
    if (P.has_bb1d_beams && A.red.max) sel = red_cell_pdf(A.red, r.ox + t * r.dx, r.oy + t * r.dy, r.oz + t * r.dz);
    const MisF mv = merge_vertex_mis(P, ml.last_pdf_w_inv * dist_sq * ray_pdf_inv, ray_pdf_inv, ratio, sel,
                                     s.mfp > P.bb1d_min_mfp ? UPBP_MIS_IN_MEDIUM_WITH_BEAMS : 0u);
    const float last = short_q ? __builtin_amdgcn_rcpf(ratio * mv.pb2d) : ray_pdf * __builtin_amdgcn_rcpf(mv.pb2d);
    const uint32_t tech = A.short_query ? (UPBP_SHORT_BEAM | P.techniques) : (UPBP_LONG_BEAM | P.techniques);
    const float w_cam = camera_weight_tail(ml.tail, mv, last, sin_t, rec.b.w, rec.c.x,
                                           rev_pdf * ray_rev * __builtin_amdgcn_rcpf(dist_sq),
                                           __builtin_amdgcn_rcpf(ray_rev), P.query_beam_type, P.photon_beam_type,
                                           tech);
    float w_light;
    if (!light_weight_tail_prec(rec, last, dir_pdf, tech, w_light)) {
        LightView lview{&G, nullptr, -1, 0.f};
        w_light = light_path_weight(lview, (int)rec.d.y, (int)path_len, last, dir_pdf, UPBP_PB2D, P.query_beam_type,
                                    P.photon_beam_type, tech, err);
    }
    const float w = __builtin_amdgcn_rcpf(w_light + w_cam);
    s.acc0 += w * u0;
    s.acc1 += w * u1;
    s.acc2 += w * u2;
}

// PhotonBeam::accumulate2 after the beam-beam test (PhotonBeam.hxx:393-605). The beam,
// its MIS record and its origin light vertex are wave-uniform (scalar loads).
// sel_pdf: the beam selection pdf of the hit's reduction cell (1 without reduction); mult: the
// times that cell tests the beam, each test adding the weighted contribution.
__device__ void beam_hit_mis(const EvalArgs& A, const Ray& r, SegState& s, const MisLane& ml, float dist,
                             float sin_theta, float t1, float t2, float radius, bool short_b, float4 flux,
                             float d2x, float d2y, float d2z, uint32_t beam_idx, float sel_pdf, uint32_t mult,
                             bool short_q, uint32_t& err) {
    const MisGlobals& G = A.mis;
    const upbp_mis_params& P = G.p;
    // the beam's record: its MIS fields, origin light vertex and light tail in one 64-byte read
    const BeamRec rec = G.beam_recs[beam_idx];
    if (rec.info & kRecBadVertex) {
        err |= kBadLightIndex;
        return;
    }
    const uint32_t path_len = rec.info & kRecLengthMask;
    const uint32_t cam_len = (uint32_t)ml.cam_len;
    if ((path_len + 1 + cam_len > P.max_path_length) || (path_len + 1 + cam_len < P.min_path_length)) return;
    if (cam_len == 1 && t1 < P.min_dist_to_med) return;
    // As in photon_hit_mis, everything below is weight and contribution arithmetic with a
    // 1e-4 budget: hardware reciprocals, and one exponential per channel for the query and
    // beam transmittances together (with the SHORT normalisations folded into the exponent).
    const float dq = t1 - s.tmin;
    const float stk = pick3(s.kidx, s.st0, s.st1, s.st2);
    const float pq = fast_exp(-stk * dq), pb = fast_exp(-stk * t2);
    const float a0 = fast_exp(neg_tau(s.st0, stk, dq, short_q) + neg_tau(s.st0, stk, t2, short_b));
    const float a1 = fast_exp(neg_tau(s.st1, stk, dq, short_q) + neg_tau(s.st1, stk, t2, short_b));
    const float a2 = fast_exp(neg_tau(s.st2, stk, dq, short_q) + neg_tau(s.st2, stk, t2, short_b));
    float rsp_q = s.min_comp * pq;
    float rsr_q = (ml.seg_flags & UPBP_ORIGIN_IN_MEDIUM) ? rsp_q : pq;
    float rsp_b = s.min_comp * pb;
    float rsr_b = (rec.info & kRecOriginInMedium) ? rsp_b : pb;
    const float ratio = __builtin_amdgcn_rcpf(s.min_comp);  // ratio_q = ratio_b
    if (!(a0 > 1e-20f && a1 > 1e-20f && a2 > 1e-20f)) return;
    rsp_q *= ml.pref_pdf;
    rsr_q *= ml.pref_rev;
    rsp_b *= rec.ray_sample_pdf;
    rsr_b *= rec.ray_sample_rev_pdf;
    // PhaseFunction::Evaluate(-ray.direction, -beam.direction): cos = -dot(-d2, -d1)
    const float cos_t = -dot3(-d2x, -d2y, -d2z, -r.dx, -r.dy, -r.dz);
    const float bsdf = phase_pdf(s.g, s.iso != 0, cos_t);
    if (bsdf <= 0.f) return;
    const float dir_pdf = bsdf * s.cont;
    const float rev_pdf = bsdf * s.cont;
    // 1D kernel (1 - dist^2/r^2) * 3 / (4 r sin(theta)) / selection pdf (PhotonBeam.hxx:487-494)
    const float k1 = radial_kernel(dist * dist, radius * radius) * 3.0f;
    const float kd = k1 * __builtin_amdgcn_rcpf((4.0f * radius) * sin_theta) * __builtin_amdgcn_rcpf(sel_pdf);
    const float u0 = ((((flux.x * s.ss0) * a0) * bsdf) * kd);
    const float u1 = ((((flux.y * s.ss1) * a1) * bsdf) * kd);
    const float u2 = ((((flux.z * s.ss2) * a2) * bsdf) * kd);
    if (u0 <= 0.f && u1 <= 0.f && u2 <= 0.f) return;
    const float dist_sq_q = t1 * t1;
    const float rsp_inv_q = __builtin_amdgcn_rcpf(rsp_q);
    const MisF mv = merge_vertex_mis(P, ml.last_pdf_w_inv * dist_sq_q * rsp_inv_q, rsp_inv_q, ratio, sel_pdf,
                                     UPBP_MIS_IN_MEDIUM_WITH_BEAMS);
    const float dist_sq_b = (rec.info & kRecFinite) ? t2 * t2 : 1.0f;
    const float rsp_inv_b = __builtin_amdgcn_rcpf(rsp_b);
    const MisF bv = merge_vertex_mis(P, rec.last_pdf_w_inv * dist_sq_b * rsp_inv_b, rsp_inv_b, ratio, sel_pdf,
                                     UPBP_MIS_IN_MEDIUM_WITH_BEAMS);
    const uint32_t tech = A.short_query ? (UPBP_SHORT_BEAM | P.techniques) : (UPBP_LONG_BEAM | P.techniques);
    const float sin_t = (tech & UPBP_NO_SINE_IN_WEIGHTS) ? 1.0f : sin_theta;
    // PhotonBeam.hxx:555-564: 1 / (the merged technique's pdf) per beam-type pair
    const float tp = mv.bb1d * sel_pdf * sin_t;
    float last;
    if (short_q)
        last = short_b ? __builtin_amdgcn_rcpf(ratio * ratio * tp) : rsp_b * __builtin_amdgcn_rcpf(ratio * tp);
    else
        last = short_b ? rsp_q * __builtin_amdgcn_rcpf(ratio * tp) : rsp_q * rsp_b * __builtin_amdgcn_rcpf(tp);
    const float w_cam = camera_weight_tail(ml.tail, mv, last, sin_t, rsp_inv_b, ratio,
                                           rev_pdf * rsr_q * __builtin_amdgcn_rcpf(dist_sq_q),
                                           __builtin_amdgcn_rcpf(rsr_q), P.query_beam_type, P.photon_beam_type, tech);
    const uint32_t lvf = rec.lv_flags;
    const bool on_light_delta = (lvf & UPBP_MIS_ON_LIGHT_SOURCE) && (lvf & UPBP_MIS_DELTA);
    const float ntl = on_light_delta
                          ? 0.f
                          : dir_pdf * rsr_b * __builtin_fabsf(rec.cos_theta_out) * __builtin_amdgcn_rcpf(dist_sq_b);
    const float rsr_inv_b = __builtin_amdgcn_rcpf(rsr_b);
    float w_light;
    if (!light_weight_tail_rec(rec, last, bv.pdf_a_inv, ntl, rsr_inv_b, tech, w_light)) {
        LightView lview{&G, &bv, (int64_t)rec.light_vertex, rsr_inv_b};
        w_light = light_path_weight(lview, rec.path_idx, (int)path_len + 1, last, ntl, UPBP_BB1D, P.query_beam_type,
                                    P.photon_beam_type, tech, err);
    }
    const float w = __builtin_amdgcn_rcpf(w_light + w_cam) * (float)mult;
    s.acc0 += w * u0;
    s.acc1 += w * u1;
    s.acc2 += w * u2;
}

// ---- UPBP hit queue ----------------------------------------------------------
// A hit's MIS weight costs a few hundred instructions, and when a wave tests an item only
// about 20 of its 64 lanes hit it (21.5 photon and 18.2 beam hits per item with any hit,
// 1920x1080 UPBP workload). Evaluated in place, the weight would run with two thirds of
// the lanes idle. So each wave queues its hits in LDS and evaluates them 64 at a time:
// a lane takes one queued hit, reads the owner lane's ray, segment and camera state with
// ds_bpermute, and the weighted contribution goes back to the owner's LDS accumulator
// with a float atomic. The queue drains completely before the segment ends, because the
// owner's segment state is what the weights read.
constexpr int kHq = 128;     // ring entries per wave: a ballot adds at most 64 to fewer than 64 queued
constexpr int kHqWaves = 4;  // waves per block of the UPBP launches (launch_mode clamps to 4)
constexpr int kHqWords = 7;  // words per queued hit
__shared__ uint32_t g_hq[kHqWaves][kHqWords][kHq];
__shared__ float g_hacc[kHqWaves][3][kWave];

// Position of this lane's entry among the set bits of m below it.
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ float from_lane(float v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ int32_t from_lane(int32_t v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ uint32_t from_lane(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, kWave); }

// The owner lane's state that photon_hit_mis / beam_hit_mis read (all lanes take part).
__device__ __forceinline__ void owner_state(const Ray& r, const SegState& s, const MisLane& ml, int owner, Ray& ro,
                                            SegState& so, MisLane& mo) {
    ro.dx = from_lane(r.dx, owner);
    ro.dy = from_lane(r.dy, owner);
    ro.dz = from_lane(r.dz, owner);
    so.tmin = from_lane(s.tmin, owner);
    so.st0 = from_lane(s.st0, owner);
    so.st1 = from_lane(s.st1, owner);
    so.st2 = from_lane(s.st2, owner);
    so.ss0 = from_lane(s.ss0, owner);
    so.ss1 = from_lane(s.ss1, owner);
    so.ss2 = from_lane(s.ss2, owner);
    so.g = from_lane(s.g, owner);
    so.iso = from_lane(s.iso, owner);
    so.kidx = from_lane(s.kidx, owner);
    so.min_comp = from_lane(s.min_comp, owner);
    so.mfp = from_lane(s.mfp, owner);
    so.cont = from_lane(s.cont, owner);
    so.acc0 = so.acc1 = so.acc2 = 0.f;
    mo.tail.c = from_lane(ml.tail.c, owner);
    mo.tail.co = from_lane(ml.tail.co, owner);
    mo.tail.fl = from_lane(ml.tail.fl, owner);
    mo.cam_len = from_lane(ml.cam_len, owner);
    mo.last_pdf_w_inv = from_lane(ml.last_pdf_w_inv, owner);
    mo.pref_pdf = from_lane(ml.pref_pdf, owner);
    mo.pref_rev = from_lane(ml.pref_rev, owner);
    mo.seg_flags = from_lane(ml.seg_flags, owner);
}

__device__ __forceinline__ void return_to_owner(int w, int owner, const SegState& so) {
    atomicAdd(&g_hacc[w][0][owner], so.acc0);
    atomicAdd(&g_hacc[w][1][owner], so.acc1);
    atomicAdd(&g_hacc[w][2][owner], so.acc2);
}

// Evaluates the first n (1..64) queued photon hits: [0] light vertex, [1] owner lane,
// [2] t, [3] squared distance, [4] r^2.
template <int MODE>
__device__ __forceinline__ void drain_photons(const EvalArgs& A, const Ray& r, const SegState& s,
                                              LaneOut<MODE>& lo, uint32_t n, bool short_q) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t pos = (lo.qh + (uint32_t)lane) & (kHq - 1);
    const bool act = (uint32_t)lane < n;
    const int owner = act ? (int)(g_hq[w][1][pos] & (kWave - 1)) : lane;
    Ray ro = {};
    SegState so;
    MisLane mo;
    owner_state(r, s, lo.ml, owner, ro, so, mo);
    if (A.red.max) {  // the hit point's reduction cell needs the owner's origin (wave-uniform branch)
        ro.ox = from_lane(r.ox, owner);
        ro.oy = from_lane(r.oy, owner);
        ro.oz = from_lane(r.oz, owner);
    }
    if (act) {
        photon_hit_mis(A, ro, so, mo, __uint_as_float(g_hq[w][2][pos]), __uint_as_float(g_hq[w][3][pos]),
                       __uint_as_float(g_hq[w][4][pos]), g_hq[w][0][pos], short_q, lo.err);
        return_to_owner(w, owner, so);
    }
    lo.qh += n;
}

// The same for beam hits: [0] piece, [1] owner lane | SHORT beam << 8 | times counted << 9,
// [2] dist, [3] sin(theta), [4] t1, [5] t2, [6] radius (reduced kernels: the cell's selection
// pdf; the radius r0 + t2 dr is then recomputed from the piece, the same operations as the
// beam test's).
template <int MODE>
__device__ __forceinline__ void drain_beams(const EvalArgs& A, const Ray& r, const SegState& s, LaneOut<MODE>& lo,
                                            uint32_t n, bool short_q) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t pos = (lo.qh + (uint32_t)lane) & (kHq - 1);
    const bool act = (uint32_t)lane < n;
    const uint32_t meta = g_hq[w][1][pos];
    const int owner = act ? (int)(meta & (kWave - 1)) : lane;
    Ray ro = {};
    SegState so;
    MisLane mo;
    owner_state(r, s, lo.ml, owner, ro, so, mo);
    if (act) {
        const uint32_t i = g_hq[w][0][pos];
        const float4 d2 = A.pieces[i].p1;
        const float t2 = __uint_as_float(g_hq[w][5][pos]);
        const float w6 = __uint_as_float(g_hq[w][6][pos]);
        float radius = w6, sel_pdf = 1.0f;
        if (MODE & kReduced) {
            radius = d2.w + t2 * A.pieces[i].p3.y;
            sel_pdf = w6;
        }
        beam_hit_mis(A, ro, so, mo, __uint_as_float(g_hq[w][2][pos]), __uint_as_float(g_hq[w][3][pos]),
                     __uint_as_float(g_hq[w][4][pos]), t2, radius, ((meta >> 8) & 1u) != 0u, A.piece_flux[i], d2.x,
                     d2.y, d2.z, A.orig[i], sel_pdf, meta >> 9, short_q, lo.err);
        return_to_owner(w, owner, so);
    }
    lo.qh += n;
}

// End of a segment: evaluates what is still queued and moves the owners' sums into s.acc.
template <int EST, int MODE>
__device__ __forceinline__ void upbp_segment_end(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                                 bool short_q) {
    const uint32_t n = uni(lo.qt - lo.qh);
    if (n) {
        if (EST == kEstPhotons)
            drain_photons<MODE>(A, r, s, lo, n, short_q);
        else
            drain_beams<MODE>(A, r, s, lo, n, short_q);
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    s.acc0 += g_hacc[w][0][lane];
    s.acc1 += g_hacc[w][1][lane];
    s.acc2 += g_hacc[w][2][lane];
    g_hacc[w][0][lane] = 0.f;
    g_hacc[w][1][lane] = 0.f;
    g_hacc[w][2][lane] = 0.f;
}

// Tests the photons start + j of a leaf for the bits j of items (ascending).
template <int MODE>
__device__ __forceinline__ void photon_leaf(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                            uint32_t start, uint64_t items, bool short_q) {
    while (items) {
        const uint32_t i = start + (uint32_t)__builtin_ctzll(items);
        items &= items - 1;
        const float4 p = A.pos_r2[i];  // scalar load, wave-uniform index
        float t, sq;
        const bool hit = test_bre(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s.tmin, s.tmax, p.x, p.y, p.z, p.w, t, sq);
        if ((MODE & 7) == (int)Mode::Upbp) {
            const uint64_t m = __ballot(hit);
            if (m) {
                const uint32_t lv_idx = A.orig[i];
                if (hit) {
                    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
                    const uint32_t pos = (lo.qt + lane_rank(m)) & (kHq - 1);
                    g_hq[w][0][pos] = lv_idx;
                    g_hq[w][1][pos] = (uint32_t)lane;
                    g_hq[w][2][pos] = __float_as_uint(t);
                    g_hq[w][3][pos] = __float_as_uint(sq);
                    g_hq[w][4][pos] = __float_as_uint(p.w);
                    lo.hits++;
                }
                lo.qt = uni(lo.qt + (uint32_t)__popcll(m));
                if (lo.qt - lo.qh >= (uint32_t)kWave) drain_photons<MODE>(A, r, s, lo, kWave, short_q);
            }
            continue;
        }
        // as in beam_leaf: only the adjoint needs the whole wave inside
        if ((MODE & 7) == (int)Mode::Adjoint ? __ballot(hit) != 0 : hit) {
            const float4 w = A.power[i];  // (rgb, 1 / (r^2 pi / 2))
            float gp0 = 0.f, gp1 = 0.f, gp2 = 0.f;
            if (hit) {
                const float dt = t - s.tmin;
                const float stk = pick3(s.kidx, s.st0, s.st1, s.st2);
                const float a0 = fast_exp(neg_tau(s.st0, stk, dt, short_q));
                const float a1 = fast_exp(neg_tau(s.st1, stk, dt, short_q));
                const float a2 = fast_exp(neg_tau(s.st2, stk, dt, short_q));
                float ph;
                if (s.iso) {
                    ph = kInvPi * 0.25f;
                } else {
                    const float4 in = A.inc[i];
                    ph = phase_pdf(s.g, false, dot3(in.x, in.y, in.z, r.dx, r.dy, r.dz));
                }
                // Epanechnikov kernel (1 - s/r^2) / (r^2 pi / 2), Bre.cxx:272-273; the hardware
                // reciprocal except near the radius (radial_kernel)
                const float kern = radial_kernel(sq, p.w);
                const float scale = (ph * kern) * w.w;
                // base_c = a_c * scale; contribution_c = power_c * base_c * sigma_s_c
                const float b0 = a0 * scale, b1 = a1 * scale, b2 = a2 * scale;
                const float c0 = w.x * b0 * s.ss0, c1 = w.y * b1 * s.ss1, c2 = w.z * b2 * s.ss2;
                if ((MODE & 7) == (int)Mode::Adjoint) {
                    // s.w = dL/dout * P (prefix transmittance of earlier segments)
                    gp0 = s.w0 * b0 * s.ss0;
                    gp1 = s.w1 * b1 * s.ss1;
                    gp2 = s.w2 * b2 * s.ss2;
                    s.gs0 += s.w0 * b0 * w.x;
                    s.gs1 += s.w1 * b1 * w.y;
                    s.gs2 += s.w2 * b2 * w.z;
                    s.gt0 -= s.w0 * c0 * dt;
                    s.gt1 -= s.w1 * c1 * dt;
                    s.gt2 -= s.w2 * c2 * dt;
                    // SHORT query: T_c / T_k, so channel k gains +dt from every channel
                    if (short_q) {
                        const float sum = (s.w0 * c0 + s.w1 * c1) + s.w2 * c2;
                        if (s.kidx == 0) s.gt0 += sum * dt;
                        else if (s.kidx == 1) s.gt1 += sum * dt;
                        else s.gt2 += sum * dt;
                    }
                    s.acc0 += s.w0 * c0;
                    s.acc1 += s.w1 * c1;
                    s.acc2 += s.w2 * c2;
                } else {
                    s.acc0 += c0;
                    s.acc1 += c1;
                    s.acc2 += c2;
                }
                lo.hits++;
                if ((MODE & 7) == (int)Mode::Signature) {
                    lo.sig_n++;
                    lo.sig_h += hash32(A.orig[i]);
                }
            }
            if ((MODE & 7) == (int)Mode::Adjoint) add_item_gradient(A.grad_items, A.orig[i], gp0, gp1, gp2);
        }
    }
}

// ---- wave frustum of rays from one point ------------------------------------
// When every active lane of a wave starts its ray at the same point (the camera rays of a
// frame), four planes through that point bound the rays' directions, and every ray lies on
// the inner side of each. A photon lying beyond one of the planes by more than its radius, or a
// beam piece whose segment does (widened by its box margin), is then farther than its radius
// from every ray of the wave: no lane can accept it, and its test is skipped for the whole
// wave. One lane per item of a leaf does that check, so a leaf costs one pass however many
// items it holds (which is why such a wave also takes small subtrees as one leaf). The planes
// are wave-uniform and live in SGPRs (VGPRs are what limits the beam kernel's occupancy); the
// vector instructions of the check take them as scalar operands.
struct Frustum {
    float n[4][3];     // inner unit normals
    float ox, oy, oz;  // the common origin
    float om;          // the origin's part of the items' rounding margin: 1e-6 |o|_1
};

__device__ __forceinline__ float uni_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

// Builds the wave's frustum for the lanes with a segment (s.tmin <= s.tmax); false (no
// culling) when their origins differ or their directions spread over more than 60 degrees
// from the mean. The bounds of the direction slopes are widened by 1e-5 relative, far above
// their rounding, so every ray stays inside its planes.
__device__ bool make_frustum(const Ray& r, const SegState& s, Frustum& F) {
    const bool act = s.tmin <= s.tmax;
    const uint64_t am = __ballot(act);
    if (am == 0) return false;
    if (__ballot(act && !(s.tmin >= 0.f)) != 0) return false;  // hits only at t >= 0 lie inside
    const int fl = __builtin_ctzll(am);
    const float ox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.ox), fl));
    const float oy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.oy), fl));
    const float oz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.oz), fl));
    if (__ballot(act && !(r.ox == ox && r.oy == oy && r.oz == oz)) != 0) return false;
    const float sx = wave_sum_dpp(act ? r.dx : 0.f), sy = wave_sum_dpp(act ? r.dy : 0.f),
                sz = wave_sum_dpp(act ? r.dz : 0.f);
    const float sn = __builtin_sqrtf((sx * sx + sy * sy) + sz * sz);
    if (!(sn > 0.f)) return false;
    const float mx = sx / sn, my = sy / sn, mz = sz / sn;  // mean direction
    const float dd = act ? (r.dx * mx + r.dy * my) + r.dz * mz : 1.f;
    if (__ballot(act && !(dd > 0.5f)) != 0) return false;
    // U across the mean direction (crossed with the axis it has the least of), V = M x U
    const float amx = __builtin_fabsf(mx), amy = __builtin_fabsf(my), amz = __builtin_fabsf(mz);
    float ux, uy, uz;
    if (amx <= amy && amx <= amz) {
        ux = 0.f, uy = mz, uz = -my;  // M x (1,0,0)
    } else if (amy <= amz) {
        ux = -mz, uy = 0.f, uz = mx;  // M x (0,1,0)
    } else {
        ux = my, uy = -mx, uz = 0.f;  // M x (0,0,1)
    }
    const float un = __builtin_sqrtf((ux * ux + uy * uy) + uz * uz);
    ux /= un, uy /= un, uz /= un;
    const float vx = my * uz - mz * uy, vy = mz * ux - mx * uz, vz = mx * uy - my * ux;
    const float u = act ? ((r.dx * ux + r.dy * uy) + r.dz * uz) / dd : 0.f;
    const float v = act ? ((r.dx * vx + r.dy * vy) + r.dz * vz) / dd : 0.f;
    float u0 = wave_min(act ? u : INFINITY), u1 = wave_max(act ? u : -INFINITY);
    float v0 = wave_min(act ? v : INFINITY), v1 = wave_max(act ? v : -INFINITY);
    u0 -= 1e-5f * (1.f + __builtin_fabsf(u0));
    u1 += 1e-5f * (1.f + __builtin_fabsf(u1));
    v0 -= 1e-5f * (1.f + __builtin_fabsf(v0));
    v1 += 1e-5f * (1.f + __builtin_fabsf(v1));
    // inner normals: d.n >= 0 for every ray direction d inside the slope bounds
    float n[4][3] = {{ux - u0 * mx, uy - u0 * my, uz - u0 * mz},
                     {u1 * mx - ux, u1 * my - uy, u1 * mz - uz},
                     {vx - v0 * mx, vy - v0 * my, vz - v0 * mz},
                     {v1 * mx - vx, v1 * my - vy, v1 * mz - vz}};
    // every lane holds the same values (wave reductions of the same inputs): take lane 0's
    for (int k = 0; k < 4; ++k) {
        const float nn = __builtin_sqrtf((n[k][0] * n[k][0] + n[k][1] * n[k][1]) + n[k][2] * n[k][2]);
        F.n[k][0] = uni_f(n[k][0] / nn);
        F.n[k][1] = uni_f(n[k][1] / nn);
        F.n[k][2] = uni_f(n[k][2] / nn);
    }
    F.ox = ox, F.oy = oy, F.oz = oz;
    F.om = uni_f(1e-6f * ((__builtin_fabsf(ox) + __builtin_fabsf(oy)) + __builtin_fabsf(oz)));
    return true;
}

// The pieces [start, start + cnt) of a leaf that some ray of the wave may accept: lane j checks
// piece start + j against the wave's frustum. An inner piece owns only pairs whose closest
// point on the beam line lies on its own segment, and that point is within the beam's radius
// of a ray, so the segment widened by 1.01 of the largest radius (the piece box's margin)
// cannot lie beyond a plane. The first and last piece of a beam also own the pairs whose
// closest point lies beyond the beam's ends (which the reference accepts only through the
// rounding of its t2 for nearly perpendicular lines): they are always kept, so culling drops
// no pair that the visited leaf (or subtree) boxes reach.
__device__ __forceinline__ uint64_t frustum_items(const EvalArgs& A, const Frustum& F, uint32_t start,
                                                  uint32_t cnt) {
    const int lane = threadIdx.x & (kWave - 1);
    bool out = false;
    if ((uint32_t)lane < cnt) {
        const float4 p0 = A.pieces[start + lane].p0, p1 = A.pieces[start + lane].p1,
                     p2 = A.pieces[start + lane].p2;
        const float ta = p2.x, tb = p2.y;
        // A margin only, so the hardware square root (within 1 ulp) does, and the plane
        // distances are fused multiply-adds: the 1e-6 relative margin is far above the error.
        const float rb = __builtin_amdgcn_sqrtf(p2.w) * 1.01f;
        const float ox = p0.x - F.ox, oy = p0.y - F.oy, oz = p0.z - F.oz;
        const float ax = __builtin_fmaf(p1.x, ta, ox), ay = __builtin_fmaf(p1.y, ta, oy),
                    az = __builtin_fmaf(p1.z, ta, oz);
        const float bx = __builtin_fmaf(p1.x, tb, ox), by = __builtin_fmaf(p1.y, tb, oy),
                    bz = __builtin_fmaf(p1.z, tb, oz);
        // rounding of the endpoints and of the plane distances: 1e-6 of the coordinates involved
        const float m = __builtin_fmaf(
            1e-6f, ((__builtin_fabsf(p0.x) + __builtin_fabsf(p0.y)) + __builtin_fabsf(p0.z)) + (p0.w + rb), rb + F.om);
        // the smallest over the planes of the larger endpoint distance: the segment lies beyond
        // some plane iff it is below -m
        float near = INFINITY;
        for (int k = 0; k < 4; ++k) {
            const float* n = F.n[k];
            const float da = __builtin_fmaf(n[2], az, __builtin_fmaf(n[1], ay, n[0] * ax));
            const float db = __builtin_fmaf(n[2], bz, __builtin_fmaf(n[1], by, n[0] * bx));
            near = fminf(near, fmaxf(da, db));
        }
        out = near < -m && __builtin_isfinite(ta) && __builtin_isfinite(tb);  // first and last pieces stay
    }
    return __ballot((uint32_t)lane < cnt && !out);
}

// The photons [start, start + cnt) of a leaf that some ray of the wave may accept: a photon is
// accepted at a point of a ray with t > min_t >= 0, inside the frustum, within its radius, so a
// photon farther than 1.01 of its radius beyond one of the planes is skipped.
__device__ __forceinline__ uint64_t frustum_photons(const EvalArgs& A, const Frustum& F, uint32_t start,
                                                    uint32_t cnt) {
    const int lane = threadIdx.x & (kWave - 1);
    bool out = false;
    if ((uint32_t)lane < cnt) {
        const float4 p = A.pos_r2[start + lane];
        const float cx = p.x - F.ox, cy = p.y - F.oy, cz = p.z - F.oz;
        // as for beam pieces: hardware square root and fused plane distances, inside the margin
        const float m = __builtin_fmaf(1e-6f, (__builtin_fabsf(p.x) + __builtin_fabsf(p.y)) + __builtin_fabsf(p.z),
                                       __builtin_fmaf(__builtin_amdgcn_sqrtf(p.w), 1.01f, F.om));
        float near = INFINITY;
        for (int k = 0; k < 4; ++k) {
            const float* n = F.n[k];
            near = fminf(near, __builtin_fmaf(n[2], cz, __builtin_fmaf(n[1], cy, n[0] * cx)));
        }
        out = near < -m;
    }
    return __ballot((uint32_t)lane < cnt && !out);
}

template <int MODE>
__device__ __forceinline__ void beam_leaf(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                          uint32_t start, uint64_t items, bool short_q) {
    // No software prefetch of the next piece: the rotation of two 64-byte records through
    // SGPRs cost more scalar moves and SGPR pressure than the latency it hid (measured 14 %).
    while (items) {
        const uint32_t i = start + (uint32_t)__builtin_ctzll(items);
        items &= items - 1;
        const PieceDev pc = A.pieces[i];
        const uint32_t word = __float_as_uint(pc.p2.z);
        const int32_t bmed = (int32_t)(word & kMediumMask);
        float dist, st, t1, t2, radius;
        // radius = mStartRadius (p1.w) + t2 * mRadiusChange (p3.y)
        const bool hit = (bmed == s.medium) &&
                         test_beam_beam(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.d1o1, s.tmin, s.tmax, pc.p0.x, pc.p0.y,
                                        pc.p0.z, pc.p1.x, pc.p1.y, pc.p1.z, pc.p3.x, pc.p0.w, pc.p2.w, pc.p2.x,
                                        pc.p2.y, pc.p1.w, pc.p3.y, dist, st, t1, t2, radius);
        if ((MODE & 7) == (int)Mode::Upbp) {
            // per-cell reduction: a hit is queued with the times its cell tests it and the
            // cell's pdf, in place of the radius (recomputed from the piece when drained)
            uint32_t red_n = 1;
            float red_pdf = 1.0f;
            if ((MODE & kReduced) && hit)
                red_n = red_decide(A.red, r.qkey, r.ox + r.dx * t1, r.oy + r.dy * t1, r.oz + r.dz * t1, A.orig[i],
                                   red_pdf);
            const bool queue = hit && red_n != 0u;
            const uint64_t m = __ballot(queue);
            if (m) {
                if (queue) {
                    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
                    const uint32_t pos = (lo.qt + lane_rank(m)) & (kHq - 1);
                    g_hq[w][0][pos] = i;
                    g_hq[w][1][pos] = (uint32_t)lane | ((word & kShortBeamBit) ? 256u : 0u) | (red_n << 9);
                    g_hq[w][2][pos] = __float_as_uint(dist);
                    g_hq[w][3][pos] = __float_as_uint(st);
                    g_hq[w][4][pos] = __float_as_uint(t1);
                    g_hq[w][5][pos] = __float_as_uint(t2);
                    g_hq[w][6][pos] = __float_as_uint((MODE & kReduced) ? red_pdf : radius);
                    lo.hits += red_n;
                }
                lo.qt = uni(lo.qt + (uint32_t)__popcll(m));
                if (lo.qt - lo.qh >= (uint32_t)kWave) drain_beams<MODE>(A, r, s, lo, kWave, short_q);
            }
            continue;
        }
        float red_w = 1.0f;
        uint32_t red_n = 1;
        if ((MODE & kReduced) && hit) {
            // the hit belongs to the reduction grid cell of its query point
            float pdf;
            red_n = red_decide(A.red, r.qkey, r.ox + r.dx * t1, r.oy + r.dy * t1, r.oz + r.dz * t1, A.orig[i], pdf);
            red_w = (1.0f / pdf) * (float)red_n;
        }
        // the adjoint reduces the item gradient over the wave, so every lane enters when one
        // hits; elsewhere the lanes that hit enter on their own (a ballot of this divergent
        // bool would first select it into a VGPR and compare it back)
        const bool take = hit && red_n != 0;
        if ((MODE & 7) == (int)Mode::Adjoint ? __ballot(take) != 0 : take) {
            const float4 flux = A.piece_flux[i];
            float gb0 = 0.f, gb1 = 0.f, gb2 = 0.f;
            if (hit && red_n != 0) {
                const bool short_b = (word & kShortBeamBit) != 0;
                const float dq = t1 - s.tmin;
                const float stk = pick3(s.kidx, s.st0, s.st1, s.st2);
                // T(t1 - tmin) [/ T_k] * T(t2) [/ T_k(t2)], one exponential per channel
                const float a0 = fast_exp(neg_tau(s.st0, stk, dq, short_q) + neg_tau(s.st0, stk, t2, short_b));
                const float a1 = fast_exp(neg_tau(s.st1, stk, dq, short_q) + neg_tau(s.st1, stk, t2, short_b));
                const float a2 = fast_exp(neg_tau(s.st2, stk, dq, short_q) + neg_tau(s.st2, stk, t2, short_b));
                // phase: Evaluate(ray.direction, beam direction), cos = -dot(d2, d1)
                const float ph = s.iso ? kInvPi * 0.25f
                                       : phase_pdf(s.g, false, -dot3(pc.p1.x, pc.p1.y, pc.p1.z, r.dx, r.dy, r.dz));
                // 1D kernel (1 - dist^2/r^2) * 3 / (4 r sin(theta)), PhotonBeam.hxx:364-365
                const float kern = radial_kernel(dist * dist, radius * radius);
                float scale = (ph * kern * 3.0f) * __builtin_amdgcn_rcpf((4.0f * radius) * st);
                if (MODE & kReduced) scale *= red_w;
                const float b0 = a0 * scale, b1 = a1 * scale, b2 = a2 * scale;
                const float c0 = flux.x * s.ss0 * b0, c1 = flux.y * s.ss1 * b1, c2 = flux.z * s.ss2 * b2;
                if ((MODE & 7) == (int)Mode::Adjoint) {
                    gb0 = s.w0 * b0 * s.ss0;
                    gb1 = s.w1 * b1 * s.ss1;
                    gb2 = s.w2 * b2 * s.ss2;
                    s.gs0 += s.w0 * b0 * flux.x;
                    s.gs1 += s.w1 * b1 * flux.y;
                    s.gs2 += s.w2 * b2 * flux.z;
                    const float g0 = s.w0 * c0, g1 = s.w1 * c1, g2 = s.w2 * c2;
                    // log C_c holds -sigma_t_c ((t1 - tmin) + t2) plus, for the SHORT
                    // normalisations by the channel-k transmittance, +sigma_t_k (t1 - tmin)
                    // (SHORT query) and +sigma_t_k t2 (SHORT beam)
                    const float L = dq + t2;
                    s.gt0 -= g0 * L;
                    s.gt1 -= g1 * L;
                    s.gt2 -= g2 * L;
                    const float Lk = (short_q ? dq : 0.f) + (short_b ? t2 : 0.f);
                    if (Lk != 0.f) {
                        const float sum = ((g0 + g1) + g2) * Lk;
                        if (s.kidx == 0) s.gt0 += sum;
                        else if (s.kidx == 1) s.gt1 += sum;
                        else s.gt2 += sum;
                    }
                    s.acc0 += g0;
                    s.acc1 += g1;
                    s.acc2 += g2;
                } else {
                    s.acc0 += c0;
                    s.acc1 += c1;
                    s.acc2 += c2;
                }
                lo.hits++;
                if ((MODE & 7) == (int)Mode::Signature) {
                    lo.sig_n += red_n;
                    lo.sig_h += red_n * hash32(A.orig[i]);
                }
            }
            if ((MODE & 7) == (int)Mode::Adjoint) add_item_gradient(A.grad_items, A.orig[i], gb0, gb1, gb2);
        }
    }
}

// ---- wave-cooperative traversal ------------------------------------------
// The items [start, start + cnt) (cnt <= 64) of a leaf, culled against the wave's frustum first
// when it has one.
template <int EST, int MODE>
__device__ __forceinline__ void visit_leaf(const EvalArgs& A, const Frustum& F, bool fr, const Ray& r, SegState& s,
                                           LaneOut<MODE>& lo, uint32_t start, uint32_t cnt, bool short_q,
                                           uint32_t active, uint64_t& tests) {
    const uint64_t all = cnt >= 64u ? ~0ull : ((1ull << cnt) - 1ull);
    if (EST == kEstPhotons) {
        const uint64_t items = fr ? frustum_photons(A, F, start, cnt) & all : all;
        tests += (uint64_t)__popcll(items) * active;
        photon_leaf<MODE>(A, r, s, lo, start, items, short_q);
    } else {
        const uint64_t items = fr ? frustum_items(A, F, start, cnt) & all : all;
        tests += (uint64_t)__popcll(items) * active;
        beam_leaf<MODE>(A, r, s, lo, start, items, short_q);
    }
}

template <int EST, int MODE>
__device__ __forceinline__ void traverse(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                         bool short_q, uint64_t& nodes_visited, uint64_t& tests,
                                         uint32_t& overflow) {
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t stk0 = 0, stk1 = 0;
    int sp = 0;
    uint32_t node = A.root;
    const uint32_t active = (uint32_t)__popcll(__ballot(s.tmin <= s.tmax));
    Frustum F;
    const bool fr = make_frustum(r, s, F);
    while (true) {
        node = uni(node);
        // One 64-byte scalar load for the whole node, then both slab tests without
        // short-circuit: testing the child references first would make the box loads
        // wait on them (three dependent scalar round trips per node instead of one).
        const NodeDev nd = A.nodes[node];
        const float4 a = nd.a, b = nd.b, c = nd.c;
        const uint4 rr = nd.r;
        nodes_visited++;
        const uint64_t m0 = slab_mask(r, s.tmin, s.tmax, a.x, a.y, a.z, a.w, b.x, b.y);
        const uint64_t m1 = slab_mask(r, s.tmin, s.tmax, b.z, b.w, c.x, c.y, c.z, c.w);
        // the wave-uniform child references are tested on the scalar side
        const bool any0 = m0 != 0 && rr.z != kInvalid;
        const bool any1 = m1 != 0 && rr.w != kInvalid;
        // A leaf; or, for a wave with a frustum, an inner node whose subtree holds at most
        // frustum_leaf items: the frustum check takes a whole leaf in one pass, so a small
        // subtree costs less as one leaf than as nodes (leaves stay small for the other waves).
        const uint32_t n0 = rr.z & kSpanMask, n1 = rr.w & kSpanMask;
        const bool leaf0 = any0 && (!(rr.z & kInner) || (fr && n0 <= A.frustum_leaf));
        const bool leaf1 = any1 && (!(rr.w & kInner) || (fr && n1 <= A.frustum_leaf));
        if (leaf0)
            visit_leaf<EST, MODE>(A, F, fr, r, s, lo, (rr.z & kSpanEnds) ? rr.x + 1u - n0 : rr.x, n0, short_q,
                                  active, tests);
        if (leaf1)
            visit_leaf<EST, MODE>(A, F, fr, r, s, lo, (rr.w & kSpanEnds) ? rr.y + 1u - n1 : rr.y, n1, short_q,
                                  active, tests);
        const bool go0 = any0 && !leaf0;
        const bool go1 = any1 && !leaf1;
        if (go0 && go1) {
            if (sp >= kStackMax) {
                overflow++;
                break;
            }
            if (sp < kWave)
                stk0 = (lane == sp) ? rr.y : stk0;
            else
                stk1 = (lane == sp - kWave) ? rr.y : stk1;
            ++sp;
            node = rr.x;
        } else if (go0) {
            node = rr.x;
        } else if (go1) {
            node = rr.y;
        } else {
            if (sp == 0) break;
            --sp;
            node = sp < kWave ? (uint32_t)__builtin_amdgcn_readlane((int)stk0, sp)
                              : (uint32_t)__builtin_amdgcn_readlane((int)stk1, sp - kWave);
        }
    }
}

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

__device__ __forceinline__ float s_tmin_of(const EvalArgs& A, int64_t i) { return A.segments[i].dist_min; }
__device__ __forceinline__ float s_tmax_of(const EvalArgs& A, int64_t i) { return A.segments[i].dist_max; }

// Processes the 64 queries of one tile (wave). `slot0` is the first lane slot.
template <int EST, int MODE>
__device__ void run_tile(const EvalArgs& A, int64_t slot0, unsigned long long* local_stats,
                         float* lds_media /* adjoint: [lds_media_n][6][64] for this wave or null */) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t slot = slot0 + lane;
    const bool valid = slot < A.count;
    int64_t qi = 0;
    if (valid) qi = A.perm ? (int64_t)A.perm[slot] : slot;
    uint32_t err = 0;
    Ray r = {};
    int64_t s0 = 0, s1 = 0;
    bool alive = false;
    if (valid) {
        r.ox = A.origin[3 * qi];
        r.oy = A.origin[3 * qi + 1];
        r.oz = A.origin[3 * qi + 2];
        r.dx = A.direction[3 * qi];
        r.dy = A.direction[3 * qi + 1];
        r.dz = A.direction[3 * qi + 2];
        s0 = A.offsets[qi];
        s1 = A.offsets[qi + 1];
        const int64_t sq = A.offsets[A.count];
        alive = true;
        if (!(s0 >= 0 && s0 <= s1 && s1 <= sq)) {
            err |= kErrBadOffsets;
            alive = false;
        }
        if (!finite3(r.ox, r.oy, r.oz) || !finite3(r.dx, r.dy, r.dz)) {
            err |= kErrNonFinite;
            alive = false;
        }
    }
    r.ix = guard_inv(r.dx);
    r.iy = guard_inv(r.dy);
    r.iz = guard_inv(r.dz);
    r.oix = r.ox * r.ix;
    r.oiy = r.oy * r.iy;
    r.oiz = r.oz * r.iz;
    r.d1o1 = dot3(r.dx, r.dy, r.dz, r.ox, r.oy, r.oz);
    if (MODE & kReduced) {
        const uint32_t bits[6] = {__float_as_uint(r.ox), __float_as_uint(r.oy), __float_as_uint(r.oz),
                                  __float_as_uint(r.dx), __float_as_uint(r.dy), __float_as_uint(r.dz)};
        r.qkey = red_query_key(bits);
    }
    int nseg = alive ? (int)(s1 - s0) : 0;
    int maxseg = nseg;
    for (int off = 32; off > 0; off >>= 1) maxseg = max(maxseg, __shfl_xor(maxseg, off, kWave));
    maxseg = (int)uni((uint32_t)maxseg);

    const bool short_q = A.short_query != 0;
    float P0 = 1.f, P1 = 1.f, P2 = 1.f;
    float R0 = 0.f, R1 = 0.f, R2 = 0.f;
    float W0 = 0.f, W1 = 0.f, W2 = 0.f;
    if ((MODE & 7) == (int)Mode::Adjoint && valid) {
        W0 = A.grad_out[3 * qi];
        W1 = A.grad_out[3 * qi + 1];
        W2 = A.grad_out[3 * qi + 2];
    }
    LaneOut<MODE> lo;
    if ((MODE & 7) == (int)Mode::Upbp) {
        // per-query AdditionalRayDataForMis (Bre.cxx:842-849 runs it per segment below)
        lo.ml.tail = CamTail{0.f, 0.f, 0u};
        lo.ml.cam_len = 1;
        lo.ml.ray_flags = 0;
        lo.ml.last_pdf_w_inv = 1.f;
        lo.ml.pref_pdf = 1.f;
        lo.ml.pref_rev = 1.f;
        lo.ml.seg_flags = UPBP_END_IN_MEDIUM;
        if (valid) {
            const upbp_query_mis qm = A.qmis[qi];
            if (qm.camera_path_length < 1 || qm.camera_offset < 0 ||
                qm.camera_offset + (int64_t)qm.camera_path_length >= A.mis.camera_count) {
                err |= kBadLightIndex;
                alive = false;
                nseg = 0;
            } else {
                lo.ml.cam_len = (int)qm.camera_path_length;
                lo.ml.tail = camera_tail(A.mis.camera + qm.camera_offset, lo.ml.cam_len, A.mis.p.query_beam_type,
                                         A.mis.p.photon_beam_type, A.mis.p.techniques);
                lo.ml.ray_flags = qm.ray_sampling_flags;
                lo.ml.last_pdf_w_inv = qm.last_pdf_w_inv;
            }
        }
    }
    if ((MODE & 7) == (int)Mode::Upbp) {
        const int w = threadIdx.x / kWave;
        g_hacc[w][0][lane] = 0.f;
        g_hacc[w][1][lane] = 0.f;
        g_hacc[w][2][lane] = 0.f;
    }
    uint64_t nodes_visited = 0, tests = 0;
    uint32_t overflow = 0;
    for (int k = 0; k < maxseg; ++k) {
        bool has = alive && k < nseg;
        SegState s;
        s.tmin = __builtin_inff();
        s.tmax = -__builtin_inff();
        s.st0 = s.st1 = s.st2 = 0.f;
        s.ss0 = s.ss1 = s.ss2 = 0.f;
        s.g = 0.f;
        s.iso = 1;
        s.kidx = 0;
        s.medium = -1;
        s.acc0 = s.acc1 = s.acc2 = 0.f;
        s.w0 = s.w1 = s.w2 = 0.f;
        s.gs0 = s.gs1 = s.gs2 = 0.f;
        s.gt0 = s.gt1 = s.gt2 = 0.f;
        const MediumDev* m = nullptr;
        float a0 = 1.f, a1 = 1.f, a2 = 1.f, pdf = 1.f;
        bool scatter = false;
        if (has) {
            const upbp_volume_segment* sg = A.segments + (s0 + k);
            const float tmin = sg->dist_min, tmax = sg->dist_max;
            const int32_t mid = sg->medium_id;
            a0 = sg->attenuation[0];
            a1 = sg->attenuation[1];
            a2 = sg->attenuation[2];
            pdf = sg->ray_sample_pdf;
            // A.n_media counts the media from A.media_base on (the whole table outside batched plans)
            if (mid < 0 || mid >= A.n_media) {
                err |= kErrBadMedium;
                has = false;
                alive = false;
            } else if (!__builtin_isfinite(tmin) || !__builtin_isfinite(tmax) || !finite3(a0, a1, a2) ||
                       !__builtin_isfinite(pdf)) {
                err |= kErrNonFinite;
                has = false;
                alive = false;
            } else if (tmin > tmax) {
                err |= kErrBadSegment;
                has = false;
                alive = false;
            } else {
                m = A.media + (A.media_base + mid);
                scatter = m->has_scattering != 0;
                s.medium = mid;
                if (scatter) {
                    s.tmin = tmin;
                    s.tmax = tmax;
                    s.st0 = m->st[0];
                    s.st1 = m->st[1];
                    s.st2 = m->st[2];
                    s.ss0 = m->ss[0];
                    s.ss1 = m->ss[1];
                    s.ss2 = m->ss[2];
                    s.g = m->g;
                    s.iso = m->isotropic;
                    s.kidx = m->min_idx;
                    s.min_comp = m->min_comp;
                    s.mfp = m->mfp;
                    s.cont = m->cont;
                }
                if ((MODE & 7) == (int)Mode::Upbp)
                    lo.ml.seg_flags = UPBP_END_IN_MEDIUM | (k == 0 ? lo.ml.ray_flags : 0u);
                if (A.lite) {
                    const float len = tmax - tmin;
                    a0 = fast_exp(-m->st[0] * len);
                    a1 = fast_exp(-m->st[1] * len);
                    a2 = fast_exp(-m->st[2] * len);
                }
            }
        }
        if ((MODE & 7) == (int)Mode::Adjoint) {
            s.w0 = W0 * P0;
            s.w1 = W1 * P1;
            s.w2 = W2 * P2;
        }
        if (__ballot(scatter)) {
            traverse<EST, MODE>(A, r, s, lo, short_q, nodes_visited, tests, overflow);
            if ((MODE & 7) == (int)Mode::Upbp) upbp_segment_end<EST, MODE>(A, r, s, lo, short_q);
        }
        if (has) {
            R0 += P0 * s.acc0;
            R1 += P1 * s.acc1;
            R2 += P2 * s.acc2;
            if ((MODE & 7) == (int)Mode::Adjoint && scatter) {
                // medium gradients of this segment plus the prefix terms of earlier segments
                const int mloc = A.media_base + s.medium;
                float* slab_m = (lds_media && s.medium < A.lds_media_n) ? lds_media : nullptr;
                if (slab_m) {
                    float* base = slab_m + (size_t)s.medium * 6 * kWave + lane;
                    base[0 * kWave] += s.gs0;
                    base[1 * kWave] += s.gs1;
                    base[2 * kWave] += s.gs2;
                    base[3 * kWave] += s.gt0;
                    base[4 * kWave] += s.gt1;
                    base[5 * kWave] += s.gt2;
                } else {
                    float* gm = A.grad_media + (size_t)mloc * 6;
                    unsafeAtomicAdd(&gm[0], s.gs0);
                    unsafeAtomicAdd(&gm[1], s.gs1);
                    unsafeAtomicAdd(&gm[2], s.gs2);
                    unsafeAtomicAdd(&gm[3], s.gt0);
                    unsafeAtomicAdd(&gm[4], s.gt1);
                    unsafeAtomicAdd(&gm[5], s.gt2);
                }
                // s.acc = sum_hits w * C for this segment (w already includes P)
                for (int j = 0; j < k; ++j) {
                    const upbp_volume_segment* sp = A.segments + (s0 + j);
                    const float len = sp->dist_max - sp->dist_min;
                    const int32_t mj = sp->medium_id;
                    float d0 = -s.acc0 * len, d1 = -s.acc1 * len, d2 = -s.acc2 * len;
                    // SHORT query: the segment's factor is attenuation over its pass-through pdf
                    // exp(-sigma_t_k len) when the medium samples distances (Media.hxx:294-323)
                    if (short_q && !A.lite) {
                        const MediumDev* mm = A.media + (A.media_base + mj);
                        if (mm->min_comp > 0.f && mm->has_scattering) {
                            const float sum = ((s.acc0 + s.acc1) + s.acc2) * len;
                            if (mm->min_idx == 0) d0 += sum;
                            else if (mm->min_idx == 1) d1 += sum;
                            else d2 += sum;
                        }
                    }
                    if (lds_media && mj < A.lds_media_n) {
                        float* base = lds_media + (size_t)mj * 6 * kWave + lane;
                        base[3 * kWave] += d0;
                        base[4 * kWave] += d1;
                        base[5 * kWave] += d2;
                    } else {
                        float* gm = A.grad_media + (size_t)(A.media_base + mj) * 6;
                        unsafeAtomicAdd(&gm[3], d0);
                        unsafeAtomicAdd(&gm[4], d1);
                        unsafeAtomicAdd(&gm[5], d2);
                    }
                }
            }
            if (A.lite || !short_q) {
                P0 *= a0;
                P1 *= a1;
                P2 *= a2;
            } else {
                P0 *= a0 / pdf;
                P1 *= a1 / pdf;
                P2 *= a2 / pdf;
            }
            if (!(P0 > 1e-20f && P1 > 1e-20f && P2 > 1e-20f)) alive = false;
            if ((MODE & 7) == (int)Mode::Upbp) {
                // ray-sampling pdf prefixes of the segment loop (Bre.cxx:871-873, lite :955-958)
                if (A.lite) {
                    float rev = 1.f, spdf = 1.f;
                    if (m->min_comp != 0.f && m->has_scattering) {
                        float att = fmaxf(__expf(-m->min_comp * (s_tmax_of(A, s0 + k) - s_tmin_of(A, s0 + k))), 1e-35f);
                        const float minatt = m->min_comp * att;
                        const uint32_t f = k == 0 ? lo.ml.ray_flags : 0u;
                        spdf = (f & UPBP_END_IN_MEDIUM) ? minatt : att;
                        rev = (f & UPBP_ORIGIN_IN_MEDIUM) ? minatt : att;
                    }
                    lo.ml.pref_pdf *= spdf;
                    lo.ml.pref_rev *= rev;
                } else {
                    lo.ml.pref_pdf *= pdf;
                    lo.ml.pref_rev *= A.segments[s0 + k].ray_sample_rev_pdf;
                }
            }
        }
    }
    err |= lo.err;
    if (valid && (MODE & 7) != (int)Mode::Adjoint) {
        if ((MODE & 7) == (int)Mode::Signature) {
            A.sig[2 * qi] = lo.sig_n;
            A.sig[2 * qi + 1] = lo.sig_h;
        } else {
            A.out[3 * qi] = R0;
            A.out[3 * qi + 1] = R1;
            A.out[3 * qi + 2] = R2;
        }
    }
    // per-wave counters
    // Reduced over the wave and made scalar: the running counters then live in SGPRs,
    // not in 64-bit VGPR pairs held across every tile of a persistent wave.
    uint32_t h = lo.hits;
    for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off, kWave);
    uint32_t e = err;
    for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off, kWave);
    local_stats[EST == kEstPhotons ? kStatPhotonHits : kStatBeamHits] += uni(h);
    local_stats[EST == kEstPhotons ? kStatPhotonCands : kStatBeamCands] += uni_u64(tests);
    local_stats[kStatNodes] += uni_u64(nodes_visited);
    local_stats[kStatError] |= uni(e);
    local_stats[kStatOverflow] += uni(overflow);
}

__device__ __forceinline__ void flush_stats(const EvalArgs& A, const unsigned long long* ls) {
    if ((threadIdx.x & (kWave - 1)) == 0) {
        for (int k = 0; k < kStatSlots; ++k) {
            if (k == kStatError) {
                if (ls[k]) atomicOr(&A.stats[k], ls[k]);
            } else if (ls[k]) {
                atomicAdd(&A.stats[k], ls[k]);
            }
        }
    }
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over 8 XCDs; remap so that each XCD works on one contiguous range of tiles
// (neighbouring tiles share nodes and leaves in that XCD's L2).
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nblk) {
    const int64_t q = nblk / 8, rmd = nblk % 8;
    const int64_t xcd = bid % 8, local = bid / 8;
    return xcd < rmd ? xcd * (q + 1) + local : rmd * (q + 1) + (xcd - rmd) * q + local;
}

// Forward kernels ask for 8 resident waves per SIMD (64 VGPRs): a few bytes of
// register spill cost less than the latency the extra wave hides (measured 3-4 %).
// UPBP kernels ask for 4 (128 VGPRs), which both fit without spilling.
template <int EST, int MODE>
__global__ __launch_bounds__(256, ((MODE & 7) == (int)Mode::Upbp ? 4 : (MODE & 7) == (int)Mode::Adjoint ? 1 : (MODE & kReduced) ? 6 : (MODE & 7) == (int)Mode::Forward ? 8 : 1)) void k_eval(
    EvalArgs A) {
    const int wave_in_block = threadIdx.x / kWave;
    const int waves_per_block = blockDim.x / kWave;
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t wave = blk * waves_per_block + wave_in_block;
    unsigned long long ls[kStatSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    // adjoint: per-lane medium gradient slabs, [lds_media_n][6][64] floats per wave, sized at launch
    extern __shared__ float lds_dyn[];
    float* my_lds = nullptr;
    if ((MODE & 7) == (int)Mode::Adjoint) {
        const int slab = A.lds_media_n * 6 * kWave;
        my_lds = lds_dyn + (size_t)wave_in_block * slab;
        for (int k = threadIdx.x & (kWave - 1); k < slab; k += kWave) my_lds[k] = 0.f;
    }
    if (wave * kWave < A.count) run_tile<EST, MODE>(A, wave * kWave, ls, my_lds);
    if ((MODE & 7) == (int)Mode::Adjoint) {
        // reduce the per-lane medium slabs of this wave and add them once
        const int lane = threadIdx.x & (kWave - 1);
        const int nm = A.lds_media_n;
        for (int e = 0; e < nm * 6; ++e) {
            const float v = wave_sum_dpp(my_lds[(size_t)e * kWave + lane]);
            if (lane == 0 && v != 0.f) unsafeAtomicAdd(&A.grad_media[(size_t)(A.media_base) * 6 + e], v);
        }
    }
    flush_stats(A, ls);
}

// Persistent batched kernel: waves take (instance, tile) items from a device
// counter until the queue drains; every wave exits when an item index passes
// total_tiles. Structures come from the plan's merged arrays through the kernel
// argument A (wave-uniform scalar loads, as in k_eval); only the instance's root
// node, query pointers and flags come from the instance table. Six waves per SIMD
// (at most 80 VGPRs) measured faster than the unconstrained allocation.
template <int EST>
__global__ __launch_bounds__(256, 6) void k_eval_batched(EvalArgs A, const InstanceDev* __restrict__ inst,
                                                            int n_inst, int64_t total_tiles, uint32_t* queue) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long ls[kStatSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    while (true) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(queue, 1u);
        item = uni((uint32_t)__shfl(item, 0, kWave));
        if ((int64_t)item >= total_tiles) break;
        // binary search of the instance owning this tile (uniform)
        int lo_i = 0, hi_i = n_inst - 1;
        while (lo_i < hi_i) {
            const int mid = (lo_i + hi_i + 1) >> 1;
            if (inst[mid].tile_begin <= (int64_t)item) lo_i = mid; else hi_i = mid - 1;
        }
        const InstanceDev& I = inst[lo_i];
        const bool has = EST == kEstPhotons ? I.has_photons : I.has_beams;
        if (!uni((uint32_t)has)) continue;
        EvalArgs B = A;
        B.root = uni(EST == kEstPhotons ? I.ph_root : I.bm_root);
        B.origin = uni_global(I.origin);
        B.direction = uni_global(I.direction);
        B.offsets = uni_global(I.offsets);
        B.segments = uni_global(I.segments);
        B.perm = I.perm ? uni_global(I.perm) : nullptr;
        B.count = uni64(I.count);
        B.short_query = (int32_t)uni((uint32_t)(I.beam_type == UPBP_SHORT_BEAM));
        B.lite = (int32_t)uni(I.lite);
        B.media_base = (int32_t)uni((uint32_t)I.media_base);
        // an instance's media are the table entries from media_base on: run_tile bounds every
        // instance-local medium id by this count (<= 0 rejects every segment)
        B.n_media = A.n_media - B.media_base;
        B.out = uni_global(EST == kEstPhotons ? I.out_photons : I.out_beams);
        const int64_t tile_begin = uni64(I.tile_begin);
        run_tile<EST, (int)Mode::Forward>(B, ((int64_t)item - tile_begin) * kWave, ls, nullptr);
    }
    flush_stats(A, ls);
}

// Rebases one instance's copied nodes into the merged array.
__global__ void k_rebase_nodes(NodeDev* nodes, int64_t n, uint32_t node_off, uint32_t item_off) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 r = nodes[i].r;
    if (r.z != kInvalid) r.x += (r.z & kInner) ? node_off : item_off;
    if (r.w != kInvalid) r.y += (r.w & kInner) ? node_off : item_off;
    nodes[i].r = r;
}

__global__ void k_clear_u64(unsigned long long* p, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}
__global__ void k_clear_f32(float* p, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_combine_media(const float* gm, const float* sign, int n, float* gs, float* ga) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const int m = i / 3, c = i % 3;
    const float direct = gm[6 * m + c], dt = gm[6 * m + 3 + c];
    if (gs) gs[i] += sign[6 * m + 3 + c] * (direct + dt);
    if (ga) ga[i] += sign[6 * m + c] * dt;
}

// Morton key of the first segment's entry point, for coherent query bundles.
__global__ void k_query_keys(int64_t n, const float* __restrict__ origin, const float* __restrict__ direction,
                             const int64_t* __restrict__ offsets, const upbp_volume_segment* __restrict__ seg,
                             float lx, float ly, float lz, float sx, float sy, float sz, uint32_t* keys,
                             uint32_t* idx) {
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t s0 = offsets[q], s1 = offsets[q + 1], sn = offsets[n];
    float t = 0.f;
    // first segment in a scattering-capable position: use its start
    if (s0 >= 0 && s0 < s1 && s1 <= sn) {
        const int64_t last = s1 - 1;
        t = seg[s0].dist_min;
        // skip a leading clear segment (camera outside the media) when there is a next one
        if (s0 < last) t = seg[s0 + 1].dist_min;
    }
    float p[3] = {origin[3 * q] + direction[3 * q] * t, origin[3 * q + 1] + direction[3 * q + 1] * t,
                  origin[3 * q + 2] + direction[3 * q + 2] * t};
    const float l[3] = {lx, ly, lz}, s[3] = {sx, sy, sz};
    uint32_t key = 0;
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
        float u = (p[a] - l[a]) * s[a];
        u = __builtin_isfinite(u) ? fminf(fmaxf(u, 0.f), 1023.f) : 0.f;
        c[a] = (uint32_t)u;
    }
    for (int b = 9; b >= 0; --b)
        key = (key << 3) | (((c[0] >> b) & 1u) << 2) | (((c[1] >> b) & 1u) << 1) | ((c[2] >> b) & 1u);
    keys[q] = key;
    idx[q] = (uint32_t)q;
}

inline unsigned grid_for(int64_t n, int block = 256) {
    return (unsigned)std::max<int64_t>(1, (n + block - 1) / block);
}

template <int EST>
void launch_mode(Ctx& ctx, Mode mode, const EvalArgs& A, unsigned blocks, unsigned threads) {
    switch (mode) {
        case Mode::Forward:
            if constexpr (EST == kEstBeams) {
                if (A.red.max != 0) {
                    hipLaunchKernelGGL((k_eval<EST, (int)Mode::Forward | kReduced>), dim3(blocks), dim3(threads), 0,
                                       ctx.stream, A);
                    break;
                }
            }
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Forward>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
            break;
        case Mode::Signature:
            if constexpr (EST == kEstBeams) {
                if (A.red.max != 0) {
                    hipLaunchKernelGGL((k_eval<EST, (int)Mode::Signature | kReduced>), dim3(blocks), dim3(threads), 0,
                                       ctx.stream, A);
                    break;
                }
            }
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Signature>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
            break;
        case Mode::Adjoint: {
            // LDS gradient slabs for the media actually present: fewer bytes, more resident waves
            const size_t lds = (size_t)(threads / kWave) * A.lds_media_n * 6 * kWave * sizeof(float);
            if constexpr (EST == kEstBeams) {
                if (A.red.max != 0) {
                    hipLaunchKernelGGL((k_eval<EST, (int)Mode::Adjoint | kReduced>), dim3(blocks), dim3(threads), lds,
                                       ctx.stream, A);
                    break;
                }
            }
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Adjoint>), dim3(blocks), dim3(threads), lds, ctx.stream, A);
            break;
        }
        case Mode::Upbp:
            if constexpr (EST == kEstBeams) {
                if (A.red.max != 0) {
                    hipLaunchKernelGGL((k_eval<EST, (int)Mode::Upbp | kReduced>), dim3(blocks), dim3(threads), 0,
                                       ctx.stream, A);
                    break;
                }
            }
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Upbp>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
            break;
    }
    UPBP_HIP(hipGetLastError());
}

EvalArgs base_args(Ctx& ctx, const QueryView& q) {
    EvalArgs A = {};
    A.origin = q.origin;
    A.direction = q.direction;
    A.offsets = q.offsets;
    A.segments = q.segments;
    A.perm = q.perm;
    A.count = q.count;
    A.short_query = q.beam_type == UPBP_SHORT_BEAM;
    A.lite = (int32_t)q.lite;
    A.media = ctx.media.as<MediumDev>();
    A.n_media = ctx.n_media;
    A.media_base = 0;
    A.lds_media_n = std::max(1, std::min(ctx.n_media, kMaxLdsMedia));
    A.stats = ctx.stats.as<unsigned long long>();
    A.frustum_leaf = (uint32_t)ctx.tuning.frustum_leaf;
    return A;
}

}  // namespace

// One thread per light vertex j starting a chain (path length 0): walks up the chain while
// the stored path lengths count 1, 2, ..., folding each vertex's terms into the tail of the
// next one. A vertex no chain reaches keeps the NaN below-marker of k_light_tails_fill.
__global__ void k_light_tails_fill(float4* __restrict__ tails, int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) tails[j] = make_float4(0.f, 0.f, 0.f, __builtin_nanf(""));
}

__global__ void k_light_tails(const upbp_light_vertex* __restrict__ lv, int64_t n, upbp_mis_params p,
                              float4* __restrict__ tails) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || lv[s].path_length != 0) return;
    const uint32_t tech = p.techniques;
    float below = 0.f;  // tail of the vertex under the current one
    for (int64_t m = s; m < n && lv[m].path_length == (uint32_t)(m - s); ++m) {
        const upbp_mis_data& d = lv[m].mis;
        float t_rest, t_ray;
        light_tail_terms(d, tech, p.query_beam_type, p.photon_beam_type, t_rest, t_ray);
        const bool spec = (d.flags & UPBP_MIS_SPECULAR) != 0;
        const bool next_spec = m > s ? (lv[m - 1].mis.flags & UPBP_MIS_SPECULAR) != 0 : spec;
        const float b = ((tech & UPBP_BPT) && !spec && !next_spec) ? 1.f : 0.f;
        tails[m] = make_float4(t_rest, t_ray, b, below);
        const float ray_rev = d.ray_sample_rev_pdf_inv;
        below = d.rev_pdf_a == 0.f ? 0.f : d.rev_pdf_a * ((t_rest + ray_rev * t_ray) + d.pdf_a_inv * (b + below));
    }
}

__global__ void k_beam_records(MisGlobals G, int64_t n, BeamRec* __restrict__ out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const upbp_beam_mis bm = G.beams[b];
    BeamRec r = {};
    r.light_vertex = bm.light_vertex;
    if (bm.light_vertex < 0 || (int64_t)bm.light_vertex >= G.lv_count) {
        r.info = kRecBadVertex;
        out[b] = r;
        return;
    }
    const upbp_light_vertex* __restrict__ lv = G.lv + bm.light_vertex;
    r.tail = G.tails ? G.tails[bm.light_vertex] : make_float4(0.f, 0.f, 0.f, __builtin_nanf(""));
    r.ray_sample_pdf = bm.ray_sample_pdf;
    r.ray_sample_rev_pdf = bm.ray_sample_rev_pdf;
    r.last_pdf_w_inv = bm.last_pdf_w_inv;
    r.cos_theta_out = lv->mis.cos_theta_out;
    r.pdf_a_inv = lv->mis.pdf_a_inv;
    r.lv_flags = lv->mis.flags;
    r.path_idx = lv->path_idx;
    const uint32_t len = lv->path_length;
    // light_weight_tail_bb1d's test that the tails apply: the vertex ends the stored chain
    const bool chain =
        G.tails && light_last_index(G, lv->path_idx, (int)len + 1) == (int64_t)bm.light_vertex + 1;
    r.info = (len < kRecLengthMask ? len : kRecLengthMask) | (lv->is_finite ? kRecFinite : 0u) |
             ((bm.ray_sampling_flags & UPBP_ORIGIN_IN_MEDIUM) ? kRecOriginInMedium : 0u) | (chain ? kRecTailChain : 0u);
    out[b] = r;
}

// light_weight_tail's inputs per light vertex v: everything but `last` and `ntl`.
__global__ void k_photon_records(MisGlobals G, PhotonRec* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.lv_count) return;
    const upbp_light_vertex* __restrict__ lv = G.lv + v;
    PhotonRec r;
    r.a = make_float4(lv->throughput[0], lv->throughput[1], lv->throughput[2], lv->continuation_prob);
    r.b = make_float4(lv->dir_fix[0], lv->dir_fix[1], lv->dir_fix[2], lv->mis.ray_sample_pdf_inv);
    const uint32_t len = lv->path_length;
    const bool has1 = len > 0;
    uint32_t info = len < kRecLengthMask ? len : kRecLengthMask;
    if (G.tails && light_last_index(G, lv->path_idx, (int)len) == v && v >= (has1 ? 1 : 0)) info |= kPRecChain;
    if (has1) info |= kPRecHas1;
    float cs = 0.f, tt = 0.f;
    if (info & kPRecChain) {
        const upbp_light_vertex* __restrict__ nx = has1 ? lv - 1 : lv;
        if (!(nx->mis.flags & UPBP_MIS_SPECULAR)) info |= kPRecNextPlain;
        if (has1) {
            const float4 t = G.tails[v - 1];
            if (!__builtin_isnan(t.w) && (int)nx->path_length == (int)len - 1) {
                info |= kPRecNextTail;
                cs = nx->mis.rev_pdf_a_without_bsdf;
                tt = (t.x + nx->mis.ray_sample_rev_pdf_inv * t.y) + nx->mis.pdf_a_inv * (t.z + t.w);
            }
        }
    }
    r.c = make_float4(lv->mis.ray_sample_pdfs_ratio, lv->mis.pdf_a_inv, cs, tt);
    r.d = make_uint4(info, (uint32_t)lv->path_idx, (uint32_t)v, 0u);
    out[v] = r;
}

void launch_photon_records(Ctx& ctx, MisDev& mis) {
    mis.photon_recs = nullptr;
    if (mis.lv_count <= 0) return;
    ctx.photon_recs.ensure((size_t)mis.lv_count * sizeof(PhotonRec));
    MisGlobals G = {};
    G.lv = mis.lv;
    G.lv_count = mis.lv_count;
    G.path_ends = mis.path_ends;
    G.path_count = mis.path_count;
    G.p = mis.params;
    G.tails = mis.tails;
    hipLaunchKernelGGL(k_photon_records, dim3(grid_for(mis.lv_count)), dim3(256), 0, ctx.stream, G,
                       ctx.photon_recs.as<PhotonRec>());
    UPBP_HIP(hipGetLastError());
    mis.photon_recs = ctx.photon_recs.p;
}

void launch_beam_records(Ctx& ctx, MisDev& mis, int64_t n_beams) {
    mis.beam_recs = nullptr;
    if (n_beams <= 0 || !mis.beams) return;
    ctx.beam_recs.ensure((size_t)n_beams * sizeof(BeamRec));
    MisGlobals G = {};
    G.lv = mis.lv;
    G.lv_count = mis.lv_count;
    G.path_ends = mis.path_ends;
    G.path_count = mis.path_count;
    G.beams = mis.beams;
    G.p = mis.params;
    G.tails = mis.tails;
    hipLaunchKernelGGL(k_beam_records, dim3(grid_for(n_beams)), dim3(256), 0, ctx.stream, G, n_beams,
                       ctx.beam_recs.as<BeamRec>());
    UPBP_HIP(hipGetLastError());
    mis.beam_recs = ctx.beam_recs.p;
}

void launch_light_tails(Ctx& ctx, MisDev& mis) {
    mis.tails = nullptr;
    if (mis.lv_count <= 0) return;
    ctx.light_tails.ensure((size_t)mis.lv_count * sizeof(float4));
    float4* t = ctx.light_tails.as<float4>();
    hipLaunchKernelGGL(k_light_tails_fill, dim3(grid_for(mis.lv_count)), dim3(256), 0, ctx.stream, t, mis.lv_count);
    UPBP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_light_tails, dim3(grid_for(mis.lv_count)), dim3(256), 0, ctx.stream, mis.lv, mis.lv_count,
                       mis.params, t);
    UPBP_HIP(hipGetLastError());
    mis.tails = t;
}

static void set_mis(EvalArgs& A, const MisDev* mis) {
    if (!mis) return;
    A.mis.tails = mis->tails;
    A.qmis = mis->queries;
    A.mis.camera = mis->camera;
    A.mis.camera_count = mis->camera_count;
    A.mis.lv = mis->lv;
    A.mis.lv_count = mis->lv_count;
    A.mis.path_ends = mis->path_ends;
    A.mis.path_count = mis->path_count;
    A.mis.beams = mis->beams;
    A.mis.beam_recs = static_cast<const BeamRec*>(mis->beam_recs);
    A.mis.photon_recs = static_cast<const PhotonRec*>(mis->photon_recs);
    A.mis.p = mis->params;
}

void launch_photons(Ctx& ctx, Mode mode, const PhotonArgs& ph, const QueryView& q, float* out, uint32_t* sig,
                    const AdjointOut* adj, const MisDev* mis, unsigned long long* stats) {
    if (q.count == 0) return;
    EvalArgs A = base_args(ctx, q);
    if (stats) A.stats = stats;
    set_mis(A, mis);
    A.nodes = ph.nodes;
    A.pos_r2 = ph.pos_r2;
    A.power = ph.power;
    A.inc = ph.inc;
    A.orig = ph.orig;
    A.red = ph.sel;
    A.out = out;
    A.sig = sig;
    if (adj) {
        A.grad_out = adj->grad_out;
        A.grad_media = adj->grad_media;
        A.grad_items = adj->grad_items;
    }
    const int want = ctx.tuning.photon_waves_per_block > 0 ? ctx.tuning.photon_waves_per_block
                                                           : ctx.tuning.waves_per_block;
    const int wpb = mode == Mode::Adjoint ? 4 : std::min(std::max(want, 1), 4);
    const int64_t waves = (q.count + kWave - 1) / kWave;
    launch_mode<kEstPhotons>(ctx, mode, A, (unsigned)((waves + wpb - 1) / wpb), (unsigned)(wpb * kWave));
}

void launch_beams(Ctx& ctx, Mode mode, const BeamArgs& bm, const QueryView& q, float* out, uint32_t* sig,
                  const AdjointOut* adj, const MisDev* mis, unsigned long long* stats) {
    if (q.count == 0) return;
    EvalArgs A = base_args(ctx, q);
    if (stats) A.stats = stats;
    set_mis(A, mis);
    A.nodes = bm.nodes;
    A.pieces = bm.pieces;
    A.piece_flux = bm.piece_flux;
    A.orig = bm.piece_beam;
    A.red = bm.red;
    A.out = out;
    A.sig = sig;
    if (adj) {
        A.grad_out = adj->grad_out;
        A.grad_media = adj->grad_media;
        A.grad_items = adj->grad_items;
    }
    const int wpb = mode == Mode::Adjoint ? 4 : std::min(std::max(ctx.tuning.waves_per_block, 1), 4);
    const int64_t waves = (q.count + kWave - 1) / kWave;
    launch_mode<kEstBeams>(ctx, mode, A, (unsigned)((waves + wpb - 1) / wpb), (unsigned)(wpb * kWave));
}

void launch_rebase_nodes(Ctx& ctx, NodeDev* nodes, int64_t n, uint32_t node_off, uint32_t item_off) {
    if (n <= 0 || (node_off == 0 && item_off == 0)) return;
    hipLaunchKernelGGL(k_rebase_nodes, dim3(grid_for(n)), dim3(256), 0, ctx.stream, nodes, n, node_off, item_off);
    UPBP_HIP(hipGetLastError());
}

void prepare_batched(Ctx& ctx, int waves_per_block) {
    const int wpb = std::min(std::max(waves_per_block, 1), 4);
    if (ctx.batched_wpb == wpb) return;
    int p = 0, b = 0;
    UPBP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &p, reinterpret_cast<const void*>(&k_eval_batched<kEstPhotons>), wpb * kWave, 0));
    UPBP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &b, reinterpret_cast<const void*>(&k_eval_batched<kEstBeams>), wpb * kWave, 0));
    ctx.batched_per_cu[0] = std::max(p, 1);
    ctx.batched_per_cu[1] = std::max(b, 1);
    ctx.batched_wpb = wpb;
}

void launch_batched(Ctx& ctx, const InstanceDev* inst, int n_inst, int64_t total_tiles, uint32_t* counters,
                    const MergedDev& merged, int which, int grid_blocks, int waves_per_block) {
    EvalArgs A = {};
    A.media = ctx.media.as<MediumDev>();
    A.n_media = ctx.n_media;
    A.stats = ctx.stats.as<unsigned long long>();
    A.frustum_leaf = (uint32_t)ctx.tuning.frustum_leaf;  // merged nodes keep node index == item index offsets
    if (which == kEstPhotons) {
        A.nodes = merged.ph_nodes;
        A.pos_r2 = merged.ph_pos_r2;
        A.power = merged.ph_power;
        A.inc = merged.ph_inc;
    } else {
        A.nodes = merged.bm_nodes;
        A.pieces = merged.bm_pieces;
        A.piece_flux = merged.bm_flux;
    }
    const unsigned threads = (unsigned)(std::min(std::max(waves_per_block, 1), 4) * kWave);
    uint32_t* queue = counters + which;
    if (grid_blocks <= 0) {
        // persistent grid: as many blocks as can be resident at once on every CU
        UPBP_REQUIRE(ctx.batched_wpb == (int)(threads / kWave) && ctx.batched_per_cu[which] > 0, UPBP_EDEVICE,
                     "batched occupancy not prepared");
        grid_blocks = ctx.num_cus * ctx.batched_per_cu[which];
    }
    if (which == kEstPhotons)
        hipLaunchKernelGGL(k_eval_batched<kEstPhotons>, dim3(grid_blocks), dim3(threads), 0, ctx.stream, A, inst,
                           n_inst, total_tiles, queue);
    else
        hipLaunchKernelGGL(k_eval_batched<kEstBeams>, dim3(grid_blocks), dim3(threads), 0, ctx.stream, A, inst, n_inst,
                           total_tiles, queue);
    UPBP_HIP(hipGetLastError());
}

void launch_query_keys(Ctx& ctx, const QueryView& q, const float* lo, const float* hi, uint32_t* keys,
                       uint32_t* idx) {
    float sc[3];
    for (int a = 0; a < 3; ++a) {
        const float e = hi[a] - lo[a];
        sc[a] = e > 0.f ? 1024.f / e : 0.f;
    }
    hipLaunchKernelGGL(k_query_keys, dim3(grid_for(q.count)), dim3(256), 0, ctx.stream, q.count, q.origin,
                       q.direction, q.offsets, q.segments, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2], keys, idx);
    UPBP_HIP(hipGetLastError());
}

void launch_clear_u64(Ctx& ctx, uint64_t* p, int64_t n) {
    hipLaunchKernelGGL(k_clear_u64, dim3(grid_for(n)), dim3(256), 0, ctx.stream, (unsigned long long*)p, n);
    UPBP_HIP(hipGetLastError());
}

void launch_clear_f32(Ctx& ctx, float* p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_clear_f32, dim3(grid_for(n)), dim3(256), 0, ctx.stream, p, n);
    UPBP_HIP(hipGetLastError());
}

void launch_fill_u32(Ctx& ctx, uint32_t* p, uint32_t v, int64_t n) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, ctx.stream, p, v, n);
    UPBP_HIP(hipGetLastError());
}

void launch_combine_media_grads(Ctx& ctx, const float* gm, const float* sign, int n_media, float* gs, float* ga) {
    hipLaunchKernelGGL(k_combine_media, dim3(grid_for(3 * n_media)), dim3(256), 0, ctx.stream, gm, sign, n_media, gs,
                       ga);
    UPBP_HIP(hipGetLastError());
}

// EmbreeBre::evalBreBruteForce (SmallUPBP/src/Bre/Bre.cxx:980-1029), the reference's debug
// estimate: one wave per query, the lanes taking every 64th kept photon; a photon inside the
// query's [min_t, max_t] disc test adds ((flux exp(-sigma_t t)) rho(d, inc)) / r^2, with the
// attenuation from the ray origin and the phase of the stored incoming direction as given
// (PhaseFunction::Evaluate(d, incDir)), and the wave's sum is scaled by sigma_s / pi.
__global__ __launch_bounds__(256) void k_bre_brute(int64_t nq, int64_t np, const float4* __restrict__ pos_r2,
                                                   const float4* __restrict__ power, const float4* __restrict__ inc,
                                                   const float* origin, const float* direction, const float* min_t,
                                                   const float* max_t, const int32_t* medium, const MediumDev* media,
                                                   int32_t n_media, float* out, uint32_t* err) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t q = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave;  // wave-uniform
    if (q >= nq) return;
    const int32_t mid = medium[q];
    if (mid < 0 || mid >= n_media) {
        if (lane == 0) {
            atomicOr(err, kErrBadMedium);
            out[3 * q] = out[3 * q + 1] = out[3 * q + 2] = 0.f;
        }
        return;
    }
    const MediumDev& M = media[mid];
    const float ox = origin[3 * q], oy = origin[3 * q + 1], oz = origin[3 * q + 2];
    const float dx = direction[3 * q], dy = direction[3 * q + 1], dz = direction[3 * q + 2];
    const float t0 = min_t[q], t1 = max_t[q];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int64_t i = lane; i < np; i += kWave) {
        const float4 p = pos_r2[i];
        float t, sq;
        if (test_bre(ox, oy, oz, dx, dy, dz, t0, t1, p.x, p.y, p.z, p.w, t, sq)) {
            const float4 f = power[i];
            const float4 in = inc[i];
            const float rho = phase_pdf(M.g, M.isotropic != 0, -dot3(in.x, in.y, in.z, dx, dy, dz));
            a0 += ((f.x * fast_exp(-t * M.st[0])) * rho) / p.w;
            a1 += ((f.y * fast_exp(-t * M.st[1])) * rho) / p.w;
            a2 += ((f.z * fast_exp(-t * M.st[2])) * rho) / p.w;
        }
    }
    a0 = wave_sum_dpp(a0);
    a1 = wave_sum_dpp(a1);
    a2 = wave_sum_dpp(a2);
    if (lane == 0) {
        out[3 * q] = (a0 * M.ss[0]) * kInvPi;
        out[3 * q + 1] = (a1 * M.ss[1]) * kInvPi;
        out[3 * q + 2] = (a2 * M.ss[2]) * kInvPi;
    }
}

void bre_eval_brute_force(Ctx& ctx, const Bre& b, const float* origin, const float* direction, const float* min_t,
                          const float* max_t, const int32_t* medium, int64_t count, float* out) {
    if (count == 0) return;
    DevBuf err;
    err.ensure(4);
    launch_fill_u32(ctx, err.as<uint32_t>(), 0u, 1);
    const int64_t threads = count * kWave;
    hipLaunchKernelGGL(k_bre_brute, dim3(grid_for(threads)), dim3(256), 0, ctx.stream, count, b.n,
                       b.pos_r2.as<float4>(), b.power.as<float4>(), b.inc.as<float4>(), origin, direction, min_t,
                       max_t, medium, ctx.media.as<MediumDev>(), (int32_t)ctx.n_media, out, err.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    uint32_t e = 0;
    UPBP_HIP(hipMemcpyAsync(&e, err.p, 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
    UPBP_REQUIRE(!(e & kErrBadMedium), UPBP_EINVAL, "query medium outside the media table");
}

}  // namespace upbp
