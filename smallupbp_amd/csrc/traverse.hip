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
    // outputs
    float* __restrict__ out;
    uint32_t* __restrict__ sig;
    unsigned long long* __restrict__ stats;
    // adjoint
    const float* __restrict__ grad_out;
    float* __restrict__ grad_media;  // [M][6]
    float* __restrict__ grad_items;  // [items][3]
    // persistent batching
    const InstanceDev* __restrict__ inst;
    int32_t n_inst;
    int64_t total_tiles;
    uint32_t* __restrict__ queue;
    int32_t which;
};

__device__ __forceinline__ float guard_inv(float d) {
    const float a = __builtin_fabsf(d) < 1e-20f ? __builtin_copysignf(1e-20f, d) : d;
    return __builtin_amdgcn_rcpf(a);
}

struct Ray {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;     // guarded reciprocal direction
    float oix, oiy, oiz;  // o * inv
};

__device__ __forceinline__ bool slab(const Ray& r, float tmin, float tmax, float lx, float ly, float lz,
                                     float hx, float hy, float hz) {
    const float ax = __builtin_fmaf(lx, r.ix, -r.oix), bx = __builtin_fmaf(hx, r.ix, -r.oix);
    const float ay = __builtin_fmaf(ly, r.iy, -r.oiy), by = __builtin_fmaf(hy, r.iy, -r.oiy);
    const float az = __builtin_fmaf(lz, r.iz, -r.oiz), bz = __builtin_fmaf(hz, r.iz, -r.oiz);
    const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    return tn <= tf;
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane state of one segment being gathered.
struct SegState {
    float tmin, tmax;
    float st0, st1, st2;
    float ss0, ss1, ss2;
    float g;
    int32_t iso;
    int32_t kidx;
    int32_t medium;     // medium id (instance-local, for the beam medium test)
    float acc0, acc1, acc2;
    // adjoint per segment: w (upstream) * prefix, and accumulators
    float w0, w1, w2;
    float gs0, gs1, gs2;  // d/d sigma_s direct
    float gt0, gt1, gt2;  // d/d sigma_t
};

template <int MODE>
struct LaneOut {
    uint32_t hits = 0;
    uint32_t sig_n = 0, sig_h = 0;
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
template <int MODE>
__device__ __forceinline__ void photon_leaf(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                            uint32_t start, uint32_t cnt, bool short_q) {
    for (uint32_t i = start; i < start + cnt; ++i) {
        const float4 p = A.pos_r2[i];
        float t, sq;
        const bool hit = test_bre(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s.tmin, s.tmax, p.x, p.y, p.z, p.w, t, sq);
        float gp0 = 0.f, gp1 = 0.f, gp2 = 0.f;
        if (hit) {
            const float4 w = A.power[i];
            const float dt = t - s.tmin;
            float a0 = fast_exp(-s.st0 * dt), a1 = fast_exp(-s.st1 * dt), a2 = fast_exp(-s.st2 * dt);
            if (short_q) {
                const float ak = pick3(s.kidx, a0, a1, a2);
                a0 = a0 / ak;
                a1 = a1 / ak;
                a2 = a2 / ak;
            }
            float ph;
            if (s.iso) {
                ph = kInvPi * 0.25f;
            } else {
                const float4 in = A.inc[i];
                ph = phase_pdf(s.g, false, dot3(in.x, in.y, in.z, r.dx, r.dy, r.dz));
            }
            const float kern = 1.0f - sq / p.w;
            const float den = (p.w * kPi) * 0.5f;
            const float scale = (ph * kern) / den;
            // base_c = a_c * scale; contribution = power_c * base_c * sigma_s_c
            const float b0 = a0 * scale, b1 = a1 * scale, b2 = a2 * scale;
            const float c0 = w.x * b0 * s.ss0, c1 = w.y * b1 * s.ss1, c2 = w.z * b2 * s.ss2;
            if (MODE == (int)Mode::Adjoint) {
                // w holds dL/dout * P (prefix transmittance)
                gp0 = s.w0 * b0 * s.ss0;
                gp1 = s.w1 * b1 * s.ss1;
                gp2 = s.w2 * b2 * s.ss2;
                s.gs0 += s.w0 * b0 * w.x;
                s.gs1 += s.w1 * b1 * w.y;
                s.gs2 += s.w2 * b2 * w.z;
                s.gt0 -= s.w0 * c0 * dt;
                s.gt1 -= s.w1 * c1 * dt;
                s.gt2 -= s.w2 * c2 * dt;
                s.acc0 += s.w0 * c0;
                s.acc1 += s.w1 * c1;
                s.acc2 += s.w2 * c2;
            } else {
                s.acc0 += c0;
                s.acc1 += c1;
                s.acc2 += c2;
            }
            lo.hits++;
            if (MODE == (int)Mode::Signature) {
                lo.sig_n++;
                lo.sig_h += hash32(A.orig[i]);
            }
        }
        if (MODE == (int)Mode::Adjoint) {
            if (__ballot(hit)) {
                const float r0 = wave_sum_dpp(gp0), r1 = wave_sum_dpp(gp1), r2 = wave_sum_dpp(gp2);
                if ((threadIdx.x & (kWave - 1)) == 0) {
                    const uint32_t o = A.orig[i];
                    unsafeAtomicAdd(&A.grad_items[3 * o + 0], r0);
                    unsafeAtomicAdd(&A.grad_items[3 * o + 1], r1);
                    unsafeAtomicAdd(&A.grad_items[3 * o + 2], r2);
                }
            }
        }
    }
}

template <int MODE>
__device__ __forceinline__ void beam_leaf(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                          uint32_t start, uint32_t cnt, bool short_q) {
    for (uint32_t i = start; i < start + cnt; ++i) {
        const PieceDev pc = A.pieces[i];
        const uint32_t word = __float_as_uint(pc.p3.z);
        const int32_t bmed = (int32_t)(word & kMediumMask);
        float dist, st, t1, t2;
        bool hit = (bmed == s.medium) &&
                   test_beam_beam(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s.tmin, s.tmax, pc.p0.x, pc.p0.y, pc.p0.z,
                                  pc.p1.x, pc.p1.y, pc.p1.z, pc.p0.w, pc.p3.w, pc.p3.x, pc.p3.y, dist, st, t1, t2);
        float radius = 0.f;
        if (hit) {
            radius = pc.p1.w + t2 * pc.p2.w;  // mStartRadius + t2 * mRadiusChange
            hit = !(radius < dist);
        }
        float gb0 = 0.f, gb1 = 0.f, gb2 = 0.f;
        if (hit) {
            const float dq = t1 - s.tmin;
            float a0 = fast_exp(-s.st0 * dq), a1 = fast_exp(-s.st1 * dq), a2 = fast_exp(-s.st2 * dq);
            const float e0 = fast_exp(-s.st0 * t2), e1 = fast_exp(-s.st1 * t2), e2 = fast_exp(-s.st2 * t2);
            if (short_q) {
                const float ak = pick3(s.kidx, a0, a1, a2);
                a0 = a0 / ak;
                a1 = a1 / ak;
                a2 = a2 / ak;
            }
            a0 *= e0;
            a1 *= e1;
            a2 *= e2;
            const bool short_b = (word & kShortBeamBit) != 0;
            if (short_b) {
                const float bk = pick3(s.kidx, e0, e1, e2);
                a0 = a0 / bk;
                a1 = a1 / bk;
                a2 = a2 / bk;
            }
            // phase: Evaluate(ray.direction, beam direction): cos = -dot(d2, d1)
            const float ph = s.iso ? kInvPi * 0.25f
                                   : phase_pdf(s.g, false, -dot3(pc.p1.x, pc.p1.y, pc.p1.z, r.dx, r.dy, r.dz));
            const float kern = 1.0f - (dist * dist) / (radius * radius);
            const float scale = (ph * kern * 3.0f) / ((4.0f * radius) * st);
            const float b0 = a0 * scale, b1 = a1 * scale, b2 = a2 * scale;
            const float c0 = pc.p2.x * s.ss0 * b0, c1 = pc.p2.y * s.ss1 * b1, c2 = pc.p2.z * s.ss2 * b2;
            if (MODE == (int)Mode::Adjoint) {
                gb0 = s.w0 * b0 * s.ss0;
                gb1 = s.w1 * b1 * s.ss1;
                gb2 = s.w2 * b2 * s.ss2;
                s.gs0 += s.w0 * b0 * pc.p2.x;
                s.gs1 += s.w1 * b1 * pc.p2.y;
                s.gs2 += s.w2 * b2 * pc.p2.z;
                const float g0 = s.w0 * c0, g1 = s.w1 * c1, g2 = s.w2 * c2;
                // d log C_c / d sigma_t_c = -(t1 - tmin) - t2, except the SHORT-beam
                // normalisation channel k (its t2 terms cancel) ...
                const int k = s.kidx;
                s.gt0 -= g0 * (dq + ((short_b && k == 0) ? 0.f : t2));
                s.gt1 -= g1 * (dq + ((short_b && k == 1) ? 0.f : t2));
                s.gt2 -= g2 * (dq + ((short_b && k == 2) ? 0.f : t2));
                // ... which instead gains +t2 from every other channel of a SHORT beam
                if (short_b) {
                    const float other = (k == 0 ? 0.f : g0) + (k == 1 ? 0.f : g1) + (k == 2 ? 0.f : g2);
                    if (k == 0) s.gt0 += other * t2;
                    else if (k == 1) s.gt1 += other * t2;
                    else s.gt2 += other * t2;
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
            if (MODE == (int)Mode::Signature) {
                lo.sig_n++;
                lo.sig_h += hash32(A.orig[i]);
            }
        }
        if (MODE == (int)Mode::Adjoint) {
            if (__ballot(hit)) {
                const float r0 = wave_sum_dpp(gb0), r1 = wave_sum_dpp(gb1), r2 = wave_sum_dpp(gb2);
                if ((threadIdx.x & (kWave - 1)) == 0) {
                    const uint32_t o = A.orig[i];
                    unsafeAtomicAdd(&A.grad_items[3 * o + 0], r0);
                    unsafeAtomicAdd(&A.grad_items[3 * o + 1], r1);
                    unsafeAtomicAdd(&A.grad_items[3 * o + 2], r2);
                }
            }
        }
    }
}

// ---- wave-cooperative traversal ------------------------------------------
template <int EST, int MODE>
__device__ __forceinline__ void traverse(const EvalArgs& A, const Ray& r, SegState& s, LaneOut<MODE>& lo,
                                         bool short_q, uint64_t& nodes_visited, uint64_t& tests,
                                         uint32_t& overflow) {
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t stk0 = 0, stk1 = 0;
    int sp = 0;
    uint32_t node = 0;
    const uint32_t active = (uint32_t)__popcll(__ballot(s.tmin <= s.tmax));
    while (true) {
        node = uni(node);
        const NodeDev* nd = A.nodes + node;
        const float4 a = nd->a, b = nd->b, c = nd->c;
        const uint4 rr = nd->r;
        nodes_visited++;
        const bool h0 = (rr.z != kInvalid) && slab(r, s.tmin, s.tmax, a.x, a.y, a.z, a.w, b.x, b.y);
        const bool h1 = (rr.w != kInvalid) && slab(r, s.tmin, s.tmax, b.z, b.w, c.x, c.y, c.z, c.w);
        const bool any0 = __ballot(h0) != 0;
        const bool any1 = __ballot(h1) != 0;
        if (any0 && rr.z != 0) {
            tests += (uint64_t)rr.z * active;
            if (EST == kEstPhotons)
                photon_leaf<MODE>(A, r, s, lo, rr.x, rr.z, short_q);
            else
                beam_leaf<MODE>(A, r, s, lo, rr.x, rr.z, short_q);
        }
        if (any1 && rr.w != 0) {
            tests += (uint64_t)rr.w * active;
            if (EST == kEstPhotons)
                photon_leaf<MODE>(A, r, s, lo, rr.y, rr.w, short_q);
            else
                beam_leaf<MODE>(A, r, s, lo, rr.y, rr.w, short_q);
        }
        const bool go0 = any0 && rr.z == 0;
        const bool go1 = any1 && rr.w == 0;
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

// Processes the 64 queries of one tile (wave). `slot0` is the first lane slot.
template <int EST, int MODE>
__device__ void run_tile(const EvalArgs& A, int64_t slot0, unsigned long long* local_stats,
                         float* lds_media /* adjoint: [kMaxLdsMedia][6][64] for this wave or null */) {
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
    int nseg = alive ? (int)(s1 - s0) : 0;
    int maxseg = nseg;
    for (int off = 32; off > 0; off >>= 1) maxseg = max(maxseg, __shfl_xor(maxseg, off, kWave));
    maxseg = (int)uni((uint32_t)maxseg);

    const bool short_q = A.short_query != 0;
    float P0 = 1.f, P1 = 1.f, P2 = 1.f;
    float R0 = 0.f, R1 = 0.f, R2 = 0.f;
    float W0 = 0.f, W1 = 0.f, W2 = 0.f;
    if (MODE == (int)Mode::Adjoint && valid) {
        W0 = A.grad_out[3 * qi];
        W1 = A.grad_out[3 * qi + 1];
        W2 = A.grad_out[3 * qi + 2];
    }
    LaneOut<MODE> lo;
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
                }
                if (A.lite) {
                    const float len = tmax - tmin;
                    a0 = fast_exp(-m->st[0] * len);
                    a1 = fast_exp(-m->st[1] * len);
                    a2 = fast_exp(-m->st[2] * len);
                }
            }
        }
        if (MODE == (int)Mode::Adjoint) {
            s.w0 = W0 * P0;
            s.w1 = W1 * P1;
            s.w2 = W2 * P2;
        }
        if (__ballot(scatter)) traverse<EST, MODE>(A, r, s, lo, short_q, nodes_visited, tests, overflow);
        if (has) {
            R0 += P0 * s.acc0;
            R1 += P1 * s.acc1;
            R2 += P2 * s.acc2;
            if (MODE == (int)Mode::Adjoint && scatter) {
                // medium gradients of this segment plus the prefix terms of earlier segments
                const int mloc = A.media_base + s.medium;
                float* slab_m = (lds_media && s.medium < kMaxLdsMedia) ? lds_media : nullptr;
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
                    const float d0 = -s.acc0 * len, d1 = -s.acc1 * len, d2 = -s.acc2 * len;
                    if (lds_media && mj < kMaxLdsMedia) {
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
        }
    }
    if (valid && MODE != (int)Mode::Adjoint) {
        if (MODE == (int)Mode::Signature) {
            A.sig[2 * qi] = lo.sig_n;
            A.sig[2 * qi + 1] = lo.sig_h;
        } else {
            A.out[3 * qi] = R0;
            A.out[3 * qi + 1] = R1;
            A.out[3 * qi + 2] = R2;
        }
    }
    // per-wave counters
    uint32_t h = lo.hits;
    for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off, kWave);
    uint32_t e = err;
    for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off, kWave);
    if (lane == 0) {
        local_stats[EST == kEstPhotons ? kStatPhotonHits : kStatBeamHits] += h;
        local_stats[EST == kEstPhotons ? kStatPhotonCands : kStatBeamCands] += tests;
        local_stats[kStatNodes] += nodes_visited;
        local_stats[kStatError] |= e;
        local_stats[kStatOverflow] += overflow;
    }
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

template <int EST, int MODE>
__global__ __launch_bounds__(256) void k_eval(EvalArgs A) {
    const int wave_in_block = threadIdx.x / kWave;
    const int waves_per_block = blockDim.x / kWave;
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t wave = blk * waves_per_block + wave_in_block;
    unsigned long long ls[kStatSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    __shared__ float lds[MODE == (int)Mode::Adjoint ? 4 * kMaxLdsMedia * 6 * kWave : 1];
    float* my_lds = nullptr;
    if (MODE == (int)Mode::Adjoint) {
        my_lds = lds + (size_t)wave_in_block * kMaxLdsMedia * 6 * kWave;
        for (int k = threadIdx.x & (kWave - 1); k < kMaxLdsMedia * 6 * kWave; k += kWave) my_lds[k] = 0.f;
    }
    if (wave * kWave < A.count) run_tile<EST, MODE>(A, wave * kWave, ls, my_lds);
    if (MODE == (int)Mode::Adjoint) {
        // reduce the per-lane medium slabs of this wave and add them once
        const int lane = threadIdx.x & (kWave - 1);
        const int nm = min(A.n_media, kMaxLdsMedia);
        for (int e = 0; e < nm * 6; ++e) {
            const float v = wave_sum_dpp(my_lds[(size_t)e * kWave + lane]);
            if (lane == 0 && v != 0.f) unsafeAtomicAdd(&A.grad_media[(size_t)(A.media_base) * 6 + e], v);
        }
    }
    flush_stats(A, ls);
}

// Persistent batched kernel: waves take (instance, tile) items from a device
// counter until the queue drains; every wave exits when an item index passes
// total_tiles.
template <int EST>
__global__ __launch_bounds__(256) void k_eval_batched(EvalArgs A) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long ls[kStatSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    while (true) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(A.queue, 1u);
        item = uni((uint32_t)__shfl(item, 0, kWave));
        if ((int64_t)item >= A.total_tiles) break;
        // binary search of the instance owning this tile (uniform)
        int lo_i = 0, hi_i = A.n_inst - 1;
        while (lo_i < hi_i) {
            const int mid = (lo_i + hi_i + 1) >> 1;
            if (A.inst[mid].tile_begin <= (int64_t)item) lo_i = mid; else hi_i = mid - 1;
        }
        const InstanceDev& I = A.inst[lo_i];
        const bool has = EST == kEstPhotons ? I.has_photons : I.has_beams;
        if (!has) continue;
        EvalArgs B = A;
        B.nodes = EST == kEstPhotons ? I.ph_nodes : I.bm_nodes;
        B.pos_r2 = I.ph_pos_r2;
        B.power = I.ph_power;
        B.inc = I.ph_inc;
        B.pieces = I.bm_pieces;
        B.origin = I.origin;
        B.direction = I.direction;
        B.offsets = I.offsets;
        B.segments = I.segments;
        B.perm = I.perm;
        B.count = I.count;
        B.short_query = I.beam_type == UPBP_SHORT_BEAM;
        B.lite = (int32_t)I.lite;
        B.media_base = I.media_base;
        B.out = EST == kEstPhotons ? I.out_photons : I.out_beams;
        run_tile<EST, (int)Mode::Forward>(B, ((int64_t)item - I.tile_begin) * kWave, ls, nullptr);
    }
    flush_stats(A, ls);
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

__global__ void k_combine_media(const float* gm, int n, float* gs, float* ga) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const int m = i / 3, c = i % 3;
    const float direct = gm[6 * m + c], dt = gm[6 * m + 3 + c];
    if (gs) gs[i] += direct + dt;
    if (ga) ga[i] += dt;
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
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Forward>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
            break;
        case Mode::Signature:
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Signature>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
            break;
        case Mode::Adjoint:
            hipLaunchKernelGGL((k_eval<EST, (int)Mode::Adjoint>), dim3(blocks), dim3(threads), 0, ctx.stream, A);
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
    A.stats = ctx.stats.as<unsigned long long>();
    return A;
}

}  // namespace

void launch_photons(Ctx& ctx, Mode mode, const PhotonArgs& ph, const QueryView& q, float* out, uint32_t* sig,
                    const AdjointOut* adj) {
    if (q.count == 0) return;
    EvalArgs A = base_args(ctx, q);
    A.nodes = ph.nodes;
    A.pos_r2 = ph.pos_r2;
    A.power = ph.power;
    A.inc = ph.inc;
    A.orig = ph.orig;
    A.out = out;
    A.sig = sig;
    if (adj) {
        A.grad_out = adj->grad_out;
        A.grad_media = adj->grad_media;
        A.grad_items = adj->grad_items;
    }
    const int wpb = mode == Mode::Adjoint ? 4 : std::min(std::max(ctx.tuning.waves_per_block, 1), 4);
    const int64_t waves = (q.count + kWave - 1) / kWave;
    launch_mode<kEstPhotons>(ctx, mode, A, (unsigned)((waves + wpb - 1) / wpb), (unsigned)(wpb * kWave));
}

void launch_beams(Ctx& ctx, Mode mode, const BeamArgs& bm, const QueryView& q, float* out, uint32_t* sig,
                  const AdjointOut* adj) {
    if (q.count == 0) return;
    EvalArgs A = base_args(ctx, q);
    A.nodes = bm.nodes;
    A.pieces = bm.pieces;
    A.orig = bm.piece_beam;
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

void launch_batched(Ctx& ctx, const InstanceDev* inst, int n_inst, int64_t total_tiles, uint32_t* counters,
                    int which, int grid_blocks, int waves_per_block) {
    EvalArgs A = {};
    A.media = ctx.media.as<MediumDev>();
    A.n_media = ctx.n_media;
    A.stats = ctx.stats.as<unsigned long long>();
    A.inst = inst;
    A.n_inst = n_inst;
    A.total_tiles = total_tiles;
    A.queue = counters + which;
    A.which = which;
    const unsigned threads = (unsigned)(std::min(std::max(waves_per_block, 1), 4) * kWave);
    if (which == kEstPhotons)
        hipLaunchKernelGGL(k_eval_batched<kEstPhotons>, dim3(grid_blocks), dim3(threads), 0, ctx.stream, A);
    else
        hipLaunchKernelGGL(k_eval_batched<kEstBeams>, dim3(grid_blocks), dim3(threads), 0, ctx.stream, A);
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

void launch_combine_media_grads(Ctx& ctx, const float* gm, int n_media, float* gs, float* ga) {
    hipLaunchKernelGGL(k_combine_media, dim3(grid_for(3 * n_media)), dim3(256), 0, ctx.stream, gm, n_media, gs, ga);
    UPBP_HIP(hipGetLastError());
}

}  // namespace upbp
