// upbp_mis.hpp -- device side of the UPBP multiple-importance-sampling weights.
//
// Restates, per hit and in registers, the MIS part of breIntersectFuncHomogeneous2
// (SmallUPBP/src/Bre/Bre.cxx:293-464) and PhotonBeam::accumulate2
// (SmallUPBP/src/Beams/PhotonBeam.hxx:386-607) with the two path-weight walks
// AccumulateCameraPathWeight / AccumulateLightPathWeight
// (SmallUPBP/src/Path/PathWeight.hxx:33-270). The reference writes the hit's MIS
// data into the caller's camera array and temporarily into the beam's origin light
// vertex; here those values are register overrides of the walks, so no input is
// ever written and concurrent lanes never race.
//
// Both walks are linear in their running product, so only their first one or two
// vertices depend on the hit. The camera walk is split once per query (camera_tail,
// camera_split) and the light walk once per light vertex (LightTail, light_split);
// each hit then costs O(1). light_path_weight, the full light walk, remains for light
// paths whose stored lengths do not count up by one.
#pragma once

#include "upbp_device.hpp"
#include "../../include/upbp.h"

namespace upbp {

struct MisF {
    float pdf_a_inv, rev_pdf_a, rev_pdf_a_without_bsdf, ray_sample_pdf_inv, ray_sample_rev_pdf_inv,
        ray_sample_pdfs_ratio, ray_sample_rev_pdfs_ratio, sin_theta, cos_theta_out, surf, pp3d, pb2d, bb1d, bb1d_sel;
    uint32_t flags;
};

__device__ __forceinline__ MisF mis_load(const upbp_mis_data* __restrict__ p) {
    const float4* v = reinterpret_cast<const float4*>(p);
    const float4 a = v[0], b = v[1], c = v[2], d = v[3];
    MisF m;
    m.pdf_a_inv = a.x;
    m.rev_pdf_a = a.y;
    m.rev_pdf_a_without_bsdf = a.z;
    m.ray_sample_pdf_inv = a.w;
    m.ray_sample_rev_pdf_inv = b.x;
    m.ray_sample_pdfs_ratio = b.y;
    m.ray_sample_rev_pdfs_ratio = b.z;
    m.sin_theta = b.w;
    m.cos_theta_out = c.x;
    m.surf = c.y;
    m.pp3d = c.z;
    m.pb2d = c.w;
    m.bb1d = d.x;
    m.bb1d_sel = d.y;
    m.flags = __float_as_uint(d.z);
    return m;
}

// Per photon beam, what its hits' MIS weights read of the beam's MIS record, its origin light
// vertex and that vertex's light-walk tail, gathered once per call (k_beam_records) so that a
// hit reads one 64-byte record instead of following beam -> light vertex -> path end -> tail.
struct alignas(16) BeamRec {
    float4 tail;                 // the origin light vertex's LightTail (NaN .w: none)
    float ray_sample_pdf;        // upbp_beam_mis.ray_sample_pdf
    float ray_sample_rev_pdf;    // upbp_beam_mis.ray_sample_rev_pdf
    float last_pdf_w_inv;        // upbp_beam_mis.last_pdf_w_inv
    float cos_theta_out;         // light vertex mis.cos_theta_out
    float pdf_a_inv;             // light vertex mis.pdf_a_inv
    uint32_t lv_flags;           // light vertex mis.flags
    uint32_t info;               // path length | kRec* bits
    int32_t light_vertex;        // upbp_beam_mis.light_vertex
    int32_t path_idx;            // light vertex path_idx
    uint32_t pad[3];
};
static_assert(sizeof(BeamRec) == 64, "BeamRec layout");
constexpr uint32_t kRecLengthMask = 0x00FFFFFFu;
constexpr uint32_t kRecFinite = 1u << 24;          // light vertex is_finite
constexpr uint32_t kRecOriginInMedium = 1u << 25;  // beam ray_sampling_flags & UPBP_ORIGIN_IN_MEDIUM
constexpr uint32_t kRecTailChain = 1u << 26;       // the vertex ends its path's stored chain (tail usable)
constexpr uint32_t kRecBadVertex = 1u << 27;       // light_vertex outside the light vertices

// Per light vertex used as a photon, what its PB2D hits' weights read of the vertex, of the
// vertex below it on its path and of that vertex's light tail, gathered once per call
// (k_photon_records): one 64-byte record per hit instead of the 112-byte vertex, the path end
// and the neighbour's fields and tail.
struct alignas(16) PhotonRec {
    float4 a;     // throughput.xyz, continuation_prob
    float4 b;     // dir_fix.xyz, mis.ray_sample_pdf_inv
    float4 c;     // mis.ray_sample_pdfs_ratio, mis.pdf_a_inv, next.rev_pdf_a_without_bsdf, next tail term
    uint4 d;      // path length | kPRec* bits, path_idx, the vertex index, unused
};
static_assert(sizeof(PhotonRec) == 64, "PhotonRec layout");
constexpr uint32_t kPRecChain = 1u << 24;       // the vertex ends its path's stored chain (tails apply)
constexpr uint32_t kPRecHas1 = 1u << 25;        // path length > 0: the walk has an index 1
constexpr uint32_t kPRecNextPlain = 1u << 26;   // the index-1 vertex is not specular
constexpr uint32_t kPRecNextTail = 1u << 27;    // the index-1 vertex's tail is usable

struct MisGlobals {
    const upbp_mis_data* __restrict__ camera;
    int64_t camera_count;
    const upbp_light_vertex* __restrict__ lv;
    int64_t lv_count;
    const int32_t* __restrict__ path_ends;
    int64_t path_count;
    const upbp_beam_mis* __restrict__ beams;
    upbp_mis_params p;
    // per light vertex, the light walk below it folded into four numbers (LightTail), or null
    const float4* __restrict__ tails;
    // per photon beam, its hits' MIS inputs (k_beam_records), or null
    const BeamRec* __restrict__ beam_recs;
    // per light vertex, its PB2D hits' MIS inputs (k_photon_records), or null
    const PhotonRec* __restrict__ photon_recs;
};

struct LightView {
    const MisGlobals* g;
    const MisF* beam_mis;       // BB1D virtual vertex (walk index 0) or null
    int64_t override_idx;       // vertex whose mRaySampleRevPdfInv is replaced, or -1
    float override_ray_rev_inv;
};

__device__ __forceinline__ MisF light_get(const LightView& l, int64_t i, uint32_t& err) {
    if (i < 0 || i >= l.g->lv_count) {
        err |= kBadLightIndex;
        MisF m = {};
        m.pdf_a_inv = 1.f;
        return m;
    }
    MisF m = mis_load(&l.g->lv[i].mis);
    if (i == l.override_idx) m.ray_sample_rev_pdf_inv = l.override_ray_rev_inv;
    return m;
}

// AccumulateLightPathWeight, PathWeight.hxx:145-270 (camera connection = false).
__device__ inline float light_path_weight(const LightView& l, int path_idx, int path_length, float last_rev,
                                          float next_to_last_partial, uint32_t evaluated, uint32_t qbt, uint32_t pbt,
                                          uint32_t tech, uint32_t& err) {
    float weight = 0.f, product = 1.f;
    int64_t last_index;
    if (path_idx == 0) {
        last_index = path_length;
    } else {
        if (path_idx - 1 < 0 || path_idx - 1 >= l.g->path_count) {
            err |= kBadLightIndex;
            return 0.f;
        }
        last_index = (int64_t)l.g->path_ends[path_idx - 1] + path_length;
    }
    MisF cur = (evaluated == UPBP_BB1D) ? *l.beam_mis : light_get(l, last_index, err);
    for (int index = 0; index <= path_length; ++index) {
        const MisF next = index < path_length ? light_get(l, last_index - index - 1, err) : cur;
        float rev = cur.rev_pdf_a, sin_t = cur.sin_theta, ray_rev = cur.ray_sample_rev_pdf_inv,
              ray_rev_ratio = cur.ray_sample_rev_pdfs_ratio;
        if (index == 0) {
            rev = last_rev;
            sin_t = 0.f;
            ray_rev = 0.f;
            ray_rev_ratio = 0.f;
        } else if (index == 1) {
            rev = (evaluated == UPBP_BB1D) ? next_to_last_partial : cur.rev_pdf_a_without_bsdf * next_to_last_partial;
        }
        if (rev == 0.f) break;
        product *= rev;
        const bool usable = index == 0 || !(cur.flags & UPBP_MIS_SPECULAR);
        if (index != 0 || evaluated == UPBP_BPT) {
            if ((tech & UPBP_SURF) && usable) weight += product * cur.surf;
            if (tech & UPBP_PP3D) weight += product * cur.pp3d;
            if (tech & UPBP_PB2D) weight += product * cur.pb2d * ((qbt & UPBP_LONG_BEAM) ? ray_rev : ray_rev_ratio);
            if ((tech & UPBP_BB1D) && !(tech & UPBP_BB1D_PREVIOUS) && (cur.flags & UPBP_MIS_IN_MEDIUM_WITH_BEAMS)) {
                if (tech & UPBP_NO_SINE_IN_WEIGHTS) sin_t = 1.f;
                const float a = (qbt & UPBP_LONG_BEAM) ? ray_rev : ray_rev_ratio;
                const float b = (pbt & UPBP_LONG_BEAM) ? cur.ray_sample_pdf_inv : cur.ray_sample_pdfs_ratio;
                weight += product * cur.bb1d * cur.bb1d_sel * a * sin_t * b;
            }
        }
        product *= cur.pdf_a_inv;
        if ((tech & UPBP_BPT) && usable && !(next.flags & UPBP_MIS_SPECULAR)) weight += product;
        if (tech & UPBP_PREVIOUS) break;
        cur = next;
    }
    return weight;
}

// ---- light walk in O(1) per hit ------------------------------------------
// The light walk is linear in its running product: from index 2 on, every term is the
// product so far times constants of the vertices below. So each light vertex j carries
// LightTail = (t_rest, t_ray, b, below):
//   terms of j itself when it sits at walk index >= 1 (usable = not specular):
//       t_rest + ray_rev(j) * t_ray, ray_rev(j) being its mRaySampleRevPdfInv (LONG query beams;
//       with SHORT ones the ratio terms are folded into t_rest and t_ray is 0);
//   b = 1 if the BPT term follows j (BPT enabled, j usable, its next vertex not specular);
//   below = the walk from j's next vertex downwards per unit product:
//       tail(m) = rev_pdf_a(m) * (T(m) + pdf_a_inv(m) * (b(m) + tail(m-1))), 0 where rev_pdf_a is 0,
//       and 0 below the light source; NaN when the light path in memory is not well formed
//       (path lengths not counting down to 0), where callers fall back to the full walk.
__device__ __forceinline__ void light_tail_terms(const upbp_mis_data& m, uint32_t tech, uint32_t qbt, uint32_t pbt,
                                                 float& t_rest, float& t_ray) {
    const bool usable = !(m.flags & UPBP_MIS_SPECULAR);
    const bool long_q = (qbt & UPBP_LONG_BEAM) != 0;
    t_rest = 0.f;
    t_ray = 0.f;
    if ((tech & UPBP_SURF) && usable) t_rest += m.surf_factor;
    if (tech & UPBP_PP3D) t_rest += m.pp3d_factor;
    if (tech & UPBP_PB2D) {
        if (long_q)
            t_ray += m.pb2d_factor;
        else
            t_rest += m.pb2d_factor * m.ray_sample_rev_pdfs_ratio;
    }
    if ((tech & UPBP_BB1D) && !(tech & UPBP_BB1D_PREVIOUS) && (m.flags & UPBP_MIS_IN_MEDIUM_WITH_BEAMS)) {
        const float sin_t = (tech & UPBP_NO_SINE_IN_WEIGHTS) ? 1.f : m.sin_theta;
        const float b = (pbt & UPBP_LONG_BEAM) ? m.ray_sample_pdf_inv : m.ray_sample_pdfs_ratio;
        const float k = m.bb1d_factor * m.bb1d_beam_selection_pdf * sin_t * b;
        if (long_q)
            t_ray += k;
        else
            t_rest += k * m.ray_sample_rev_pdfs_ratio;
    }
}

// Index of the walk's first light vertex (PathWeight.hxx:167), or -1 for a path index
// outside the path table.
__device__ __forceinline__ int64_t light_last_index(const MisGlobals& G, int path_idx, int path_length) {
    if (path_idx == 0) return path_length;
    if (path_idx - 1 < 0 || path_idx - 1 >= G.path_count) return -1;
    return (int64_t)G.path_ends[path_idx - 1] + path_length;
}

// AccumulateLightPathWeight for a merge at light vertex v (PB2D, PP3D, SURF: index 0 is v
// itself, index 1 its next vertex with rev = rev_pdf_a_without_bsdf * ntl) through the
// tails. Returns false where the full walk must run instead.
__device__ __forceinline__ bool light_weight_tail(const MisGlobals& G, int64_t v, int path_idx, int path_length,
                                                  float last, float ntl, uint32_t tech, float& w) {
    const bool has1 = path_length > 0;
    if (!G.tails || light_last_index(G, path_idx, path_length) != v || v < (has1 ? 1 : 0) || v >= G.lv_count)
        return false;
    const upbp_light_vertex* __restrict__ lv = G.lv + v;
    w = 0.f;
    if (last == 0.f) return true;  // the walk breaks at index 0
    const float p0 = last * lv->mis.pdf_a_inv;
    const upbp_light_vertex* __restrict__ nx = has1 ? lv - 1 : lv;
    if ((tech & UPBP_BPT) && !(nx->mis.flags & UPBP_MIS_SPECULAR)) w = p0;
    if ((tech & UPBP_PREVIOUS) || !has1) return true;
    const float4 t = G.tails[v - 1];
    if (__builtin_isnan(t.w) || (int)nx->path_length != path_length - 1) return false;
    const float r1 = nx->mis.rev_pdf_a_without_bsdf * ntl;
    if (r1 == 0.f) return true;
    w += p0 * r1 * ((t.x + nx->mis.ray_sample_rev_pdf_inv * t.y) + nx->mis.pdf_a_inv * (t.z + t.w));
    return true;
}

// light_weight_tail from the photon's record: the same tests and terms, without loads.
__device__ __forceinline__ bool light_weight_tail_prec(const PhotonRec& rec, float last, float ntl, uint32_t tech,
                                                       float& w) {
    const uint32_t info = rec.d.x;
    if (!(info & kPRecChain)) return false;
    w = 0.f;
    if (last == 0.f) return true;  // the walk breaks at index 0
    const float p0 = last * rec.c.y;
    if ((tech & UPBP_BPT) && (info & kPRecNextPlain)) w = p0;
    if ((tech & UPBP_PREVIOUS) || !(info & kPRecHas1)) return true;
    if (!(info & kPRecNextTail)) return false;
    const float r1 = rec.c.z * ntl;
    if (r1 == 0.f) return true;
    w += p0 * r1 * rec.c.w;
    return true;
}

// light_weight_tail_bb1d from the beam's record: the same tests and terms, without loads.
__device__ __forceinline__ bool light_weight_tail_rec(const BeamRec& rec, float last, float bv_pdf_a_inv, float ntl,
                                                      float ray_rev_b, uint32_t tech, float& w) {
    if (!(rec.info & kRecTailChain)) return false;
    w = 0.f;
    if (last == 0.f) return true;
    const float p0 = last * bv_pdf_a_inv;
    if ((tech & UPBP_BPT) && !(rec.lv_flags & UPBP_MIS_SPECULAR)) w = p0;
    if (tech & UPBP_PREVIOUS) return true;
    const float4 t = rec.tail;
    if (__builtin_isnan(t.w)) return false;
    if (ntl == 0.f) return true;
    w += p0 * ntl * ((t.x + ray_rev_b * t.y) + rec.pdf_a_inv * (t.z + t.w));
    return true;
}

// The same for BB1D, whose walk starts at the beam's virtual vertex bv (index 0) above the
// beam's origin light vertex b; index 1 is b with rev = ntl and its mRaySampleRevPdfInv
// replaced by ray_rev_b (PhotonBeam.hxx:548-591).
__device__ __forceinline__ bool light_weight_tail_bb1d(const MisGlobals& G, int64_t b, int path_idx,
                                                       int path_length_b, float last, float bv_pdf_a_inv, float ntl,
                                                       float ray_rev_b, uint32_t tech, float& w) {
    if (!G.tails || light_last_index(G, path_idx, path_length_b + 1) != b + 1 || b < 0 || b >= G.lv_count)
        return false;
    const upbp_light_vertex* __restrict__ lv = G.lv + b;
    w = 0.f;
    if (last == 0.f) return true;
    const float p0 = last * bv_pdf_a_inv;
    if ((tech & UPBP_BPT) && !(lv->mis.flags & UPBP_MIS_SPECULAR)) w = p0;
    if (tech & UPBP_PREVIOUS) return true;
    const float4 t = G.tails[b];
    if (__builtin_isnan(t.w)) return false;
    if (ntl == 0.f) return true;
    w += p0 * ntl * ((t.x + ray_rev_b * t.y) + lv->mis.pdf_a_inv * (t.z + t.w));
    return true;
}

// The light walk of a merge at light vertex v (PP3D, SURF) split into its two per-vertex
// constants: AccumulateLightPathWeight(last, ntl) = last * (k0 + k1 * ntl). Returns false
// where the tails do not apply (the full walk must run).
__device__ __forceinline__ bool light_split(const MisGlobals& G, int64_t v, int path_idx, int path_length,
                                            uint32_t tech, float& k0, float& k1) {
    const bool has1 = path_length > 0;
    if (!G.tails || light_last_index(G, path_idx, path_length) != v || v < (has1 ? 1 : 0) || v >= G.lv_count)
        return false;
    const upbp_light_vertex* __restrict__ lv = G.lv + v;
    const float p0 = lv->mis.pdf_a_inv;
    const upbp_light_vertex* __restrict__ nx = has1 ? lv - 1 : lv;
    k0 = ((tech & UPBP_BPT) && !(nx->mis.flags & UPBP_MIS_SPECULAR)) ? p0 : 0.f;
    k1 = 0.f;
    if ((tech & UPBP_PREVIOUS) || !has1) return true;
    const float4 t = G.tails[v - 1];
    if (__builtin_isnan(t.w) || (int)nx->path_length != path_length - 1) return false;
    k1 = p0 * nx->mis.rev_pdf_a_without_bsdf *
         ((t.x + nx->mis.ray_sample_rev_pdf_inv * t.y) + nx->mis.pdf_a_inv * (t.z + t.w));
    return !__builtin_isnan(k1);
}

// The camera walk of a merge (PP3D, SURF: the merged vertex is the camera sub-path's
// entry [len] itself, no override) split into three per-query constants:
//   AccumulateCameraPathWeight(last_rev, last_sin, ray_rev, ray_rev_ratio, ntl)
//     = last_rev * ((a0 + bb0 * s * b) + cn * ntl),
// s = last_sin (1 with NO_SINE_IN_WEIGHTS), b = ray_rev (LONG photon beams) or ray_rev_ratio.
// Every term past index 0 carries ntl once (through index 1's reverse pdf); with
// BB1D_PREVIOUS a non-specular vertex past index 0 drops the index-0 BB1D term.
struct CamSplit {
    float a0, bb0, cn;
};

__device__ inline CamSplit camera_split(const upbp_mis_data* __restrict__ base, int len, uint32_t qbt, uint32_t pbt,
                                        uint32_t tech) {
    float lin0 = 0.f, bb0 = 0.f, lin = 0.f, bb = 0.f, product = 1.f;
    bool drop0 = false;
    const bool long_q = (qbt & UPBP_LONG_BEAM) != 0;
    MisF cur = mis_load(base + len);
    for (int index = 0; index < len; ++index) {
        const MisF next = index < len - 1 ? mis_load(base + len - index - 1) : cur;
        const bool cur_spec = (cur.flags & UPBP_MIS_SPECULAR) != 0;
        if (index > 0 && (tech & UPBP_BB1D_PREVIOUS) && !cur_spec) {
            bb = 0.f;
            drop0 = true;
        }
        if (index > 0) product *= cur.rev_pdf_a;
        const bool usable = index == 0 || !cur_spec;
        const float a = long_q ? cur.ray_sample_pdf_inv : cur.ray_sample_pdfs_ratio;
        float t = 0.f;
        if ((tech & UPBP_SURF) && usable) t += product * cur.surf;
        if (tech & UPBP_PP3D) t += product * cur.pp3d;
        if (tech & UPBP_PB2D) t += product * cur.pb2d * a;
        if ((tech & UPBP_BB1D) && (cur.flags & UPBP_MIS_IN_MEDIUM_WITH_BEAMS)) {
            const float k = product * cur.bb1d * cur.bb1d_sel * a;
            if (index == 0) {
                bb0 = k;
            } else {
                const float sin_t = (tech & UPBP_NO_SINE_IN_WEIGHTS) ? 1.f : cur.sin_theta;
                const float b = (pbt & UPBP_LONG_BEAM) ? cur.ray_sample_rev_pdf_inv : cur.ray_sample_rev_pdfs_ratio;
                bb += k * sin_t * b;
            }
        }
        product *= cur.pdf_a_inv;
        if ((tech & UPBP_BPT) && usable && !(next.flags & UPBP_MIS_SPECULAR)) t += product;
        if (index == 0)
            lin0 += t;
        else
            lin += t;
        cur = next;
    }
    return CamSplit{lin0, drop0 ? 0.f : bb0, lin + bb};
}

// The camera walk of a PB2D / BB1D hit (AccumulateCameraPathWeight with the hit's merge
// vertex as entry [len] and the override of entry [len-1].mRaySampleRevPdfInv) split at
// index 1: everything from index 1 on is a per-query constant times ntl, except index 1's
// BB1D term with LONG photon beams, which is linear in the override.
//   weight = last_rev * (t0 + pdf0 * (b0 + ntl * (c + co * override)))
// with t0, pdf0 the merge vertex's index-0 terms and forward pdf.
struct CamTail {
    float c, co;
    uint32_t fl;  // kCamB0: the BPT term after index 0; kCamDrop0: index 0's BB1D term is dropped
};
constexpr uint32_t kCamB0 = 1u;
constexpr uint32_t kCamDrop0 = 2u;

__device__ inline CamTail camera_tail(const upbp_mis_data* __restrict__ base, int len, uint32_t qbt, uint32_t pbt,
                                      uint32_t tech) {
    CamTail ct = {0.f, 0.f, 0u};
    // index 0's next: entry [len-1], or the merge vertex itself (never specular) for len 1
    if ((tech & UPBP_BPT) && (len < 2 || !(base[len - 1].flags & UPBP_MIS_SPECULAR))) ct.fl |= kCamB0;
    if (len < 2) return ct;
    const bool long_q = (qbt & UPBP_LONG_BEAM) != 0;
    const bool long_p = (pbt & UPBP_LONG_BEAM) != 0;
    float lin = 0.f, bb = 0.f, bb_over = 0.f, product = 1.f;
    MisF cur = mis_load(base + len - 1);
    for (int index = 1; index < len; ++index) {
        const MisF next = index < len - 1 ? mis_load(base + len - index - 1) : cur;
        const bool cur_spec = (cur.flags & UPBP_MIS_SPECULAR) != 0;
        if ((tech & UPBP_BB1D_PREVIOUS) && !cur_spec) {
            bb = 0.f;
            bb_over = 0.f;
            ct.fl |= kCamDrop0;
        }
        product *= cur.rev_pdf_a;
        const bool usable = !cur_spec;
        const float a = long_q ? cur.ray_sample_pdf_inv : cur.ray_sample_pdfs_ratio;
        if ((tech & UPBP_SURF) && usable) lin += product * cur.surf;
        if (tech & UPBP_PP3D) lin += product * cur.pp3d;
        if (tech & UPBP_PB2D) lin += product * cur.pb2d * a;
        if ((tech & UPBP_BB1D) && (cur.flags & UPBP_MIS_IN_MEDIUM_WITH_BEAMS)) {
            const float sin_t = (tech & UPBP_NO_SINE_IN_WEIGHTS) ? 1.f : cur.sin_theta;
            const float k = product * cur.bb1d * cur.bb1d_sel * a * sin_t;
            if (index == 1 && long_p)
                bb_over += k;
            else
                bb += k * (long_p ? cur.ray_sample_rev_pdf_inv : cur.ray_sample_rev_pdfs_ratio);
        }
        product *= cur.pdf_a_inv;
        if ((tech & UPBP_BPT) && usable && !(next.flags & UPBP_MIS_SPECULAR)) lin += product;
        cur = next;
    }
    ct.c = lin + bb;
    ct.co = bb_over;
    return ct;
}

// camera_path_weight through a CamTail: last = the merge vertex (index 0).
__device__ __forceinline__ float camera_weight_tail(const CamTail& ct, const MisF& last, float last_rev,
                                                    float last_sin, float last_ray_rev_inv, float last_ray_rev_ratio,
                                                    float ntl, float override_ray_rev_inv, uint32_t qbt, uint32_t pbt,
                                                    uint32_t tech) {
    const float a0 = (qbt & UPBP_LONG_BEAM) ? last.ray_sample_pdf_inv : last.ray_sample_pdfs_ratio;
    float t0 = 0.f;
    if (tech & UPBP_SURF) t0 += last.surf;
    if (tech & UPBP_PP3D) t0 += last.pp3d;
    if (tech & UPBP_PB2D) t0 += last.pb2d * a0;
    if ((tech & UPBP_BB1D) && (last.flags & UPBP_MIS_IN_MEDIUM_WITH_BEAMS) && !(ct.fl & kCamDrop0)) {
        const float sin_t = (tech & UPBP_NO_SINE_IN_WEIGHTS) ? 1.f : last_sin;
        const float b = (pbt & UPBP_LONG_BEAM) ? last_ray_rev_inv : last_ray_rev_ratio;
        t0 += last.bb1d * last.bb1d_sel * a0 * sin_t * b;
    }
    const float b0 = (ct.fl & kCamB0) ? 1.f : 0.f;
    return last_rev * (t0 + last.pdf_a_inv * (b0 + ntl * (ct.c + ct.co * override_ray_rev_inv)));
}

// Per-lane query state of the UPBP callbacks: AdditionalRayDataForMis of the current segment.
struct MisLane {
    CamTail tail;  // camera_tail of the query's camera sub-path
    int cam_len;
    uint32_t ray_flags;    // raySamplingFlags of evalBre (first segment only)
    float last_pdf_w_inv;
    float pref_pdf, pref_rev;  // mRaySamplePdf, mRaySampleRevPdf of the segment loop
    uint32_t seg_flags;        // mRaySamplingFlags of the current segment
};

}  // namespace upbp
