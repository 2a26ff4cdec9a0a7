// lighttrace.hip -- light sub-path tracing in media on the GPU: the producer of the light
// vertices, their MisData and the photon beams that the estimators consume.
//
// Restates the light-tracing loop of UPBPRenderer::RunIteration
// (SmallUPBP/src/Renderers/UPBP.hxx:501-674) with GenerateLightSample (:1592-1662) for
// point lights (Scene/Lights.hxx:363-384), AddBeams (:1783-1870) and SampleScattering
// (:1908-1969) with the medium BSDF (Path/Bsdf.hxx:140-151,385-401) and
// PhaseFunction::Sample (Path/PhaseFunction.hxx:95-131). The media part of
// Scene::Intersect (Scene/Scene.hxx:393-531) becomes, for homogeneous media bounded by
// imaginary spheres and boxes: the ray is cut at every boundary crossing, each piece takes
// the medium of the innermost region containing its midpoint, and distance sampling follows
// HomogeneousMedium::SampleRay / RaySamplePdf (Scene/Media.hxx:240-323).
//
// One thread traces one path; paths are independent because each draws from its own
// counter-based SplitMix64 stream (path index, draw index). The crossings of a ray are not
// stored: the next one is found again from the regions each time (at most a few per ray),
// which keeps the thread's state in registers. A path writes into fixed slots (vertices:
// max_path_length, beams: (max_path_length - 1) rays of at most 2 R + 1 pieces); an
// exclusive scan of the per-path counts then packs vertices and beams in path order into
// the caller's arrays and gives mPathEnds. Slots are slot-major across paths (slot k of path p
// at k * path_count + p), so a wave's accesses to its paths' slot k are contiguous.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "upbp_internal.hpp"

namespace upbp {
namespace {

constexpr int kBlock = 128;
constexpr float kPiF = 3.141592653589793238462643383279f;
constexpr float kInvPiF = 1.f / kPiF;
constexpr float kFar = 1e36f;

struct LtScene {
    upbp_region regions[UPBP_MAX_REGIONS];
    int32_t n_regions;
    int32_t global_medium;
    int32_t n_lights;
};

// A beam in the path's slots: the upbp_beam_input fields plus its MIS record (light_vertex
// relative to the path until packed).
struct BeamRec {
    float o[3], d[3], len, thr[3];
    uint32_t flags;
    int32_t medium;
    upbp_beam_mis mis;
};

// ---- counter-based random numbers: draw k of path p ------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t base, k;
    __device__ float next() {
        k += 1;
        const uint64_t h = mix64(base + k * 0x9E3779B97F4A7C15ull);
        return ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);  // (0, 1), exact
    }
};

// ---- regions -------------------------------------------------------------------------------
__device__ __forceinline__ bool contains(const upbp_region& g, float x, float y, float z) {
    if (g.shape == UPBP_REGION_SPHERE) {
        const float dx = x - g.p0[0], dy = y - g.p0[1], dz = z - g.p0[2];
        return (dx * dx + dy * dy) + dz * dz <= g.p1[0] * g.p1[0];
    }
    return x >= g.p0[0] && x <= g.p1[0] && y >= g.p0[1] && y <= g.p1[1] && z >= g.p0[2] && z <= g.p1[2];
}

__device__ int32_t medium_at(const LtScene& S, float x, float y, float z) {
    int32_t m = S.global_medium;
    for (int i = 0; i < S.n_regions; ++i)
        if (contains(S.regions[i], x, y, z)) m = S.regions[i].medium_id;
    return m;
}

// Smallest boundary crossing of region g beyond t (kFar if none).
__device__ float next_crossing(const upbp_region& g, const float* o, const float* d, float t) {
    float c0 = -1.f, c1 = -1.f;
    if (g.shape == UPBP_REGION_SPHERE) {
        const float ox = o[0] - g.p0[0], oy = o[1] - g.p0[1], oz = o[2] - g.p0[2];
        const float b = (ox * d[0] + oy * d[1]) + oz * d[2];
        const float c = ((ox * ox + oy * oy) + oz * oz) - g.p1[0] * g.p1[0];
        const float disc = b * b - c;
        if (disc < 0.f) return kFar;
        const float s = __builtin_sqrtf(disc);
        c0 = -b - s;
        c1 = -b + s;
    } else {
        float tn = -INFINITY, tf = INFINITY;
        for (int a = 0; a < 3; ++a) {
            if (d[a] != 0.f) {
                float a0 = (g.p0[a] - o[a]) / d[a], a1 = (g.p1[a] - o[a]) / d[a];
                if (a0 > a1) {
                    const float x = a0;
                    a0 = a1;
                    a1 = x;
                }
                tn = fmaxf(tn, a0);
                tf = fminf(tf, a1);
            } else if (o[a] < g.p0[a] || o[a] > g.p1[a]) {
                return kFar;
            }
        }
        if (!(tn <= tf)) return kFar;
        c0 = tn;
        c1 = tf;
    }
    float best = kFar;
    if (c0 > 0.f && c0 > t) best = c0;
    if (c1 > 0.f && c1 > t && c1 < best) best = c1;
    return best;
}

// The piece of the ray starting at t: its end and medium (VolumeSegmentLite).
__device__ void piece_at(const LtScene& S, const float* o, const float* d, float t, float& t1, int32_t& med) {
    float b = kFar;
    for (int i = 0; i < S.n_regions; ++i) b = fminf(b, next_crossing(S.regions[i], o, d, t));
    t1 = b;
    const float mid = b < kFar ? 0.5f * (t + b) : t + 1.f;
    med = medium_at(S, o[0] + d[0] * mid, o[1] + d[1] * mid, o[2] + d[2] * mid);
}

// ---- sampling ----------------------------------------------------------------------------
// SampleUniformSphereW, Misc/Utils2.hxx:224-241
__device__ __forceinline__ void uniform_sphere(float u0, float u1, float* dir) {
    const float term1 = 2.f * kPiF * u0;
    const float term2 = 2.f * __builtin_sqrtf(u1 - u1 * u1);
    dir[0] = cosf(term1) * term2;
    dir[1] = sinf(term1) * term2;
    dir[2] = 1.f - 2.f * u1;
}

// PhaseFunction::Pdf with cos = -dot(gen, fix), PhaseFunction.hxx:52-80 (IEEE operations, as
// the CPU restatement; the directions it scores are decisions of the path)
__device__ __forceinline__ float phase_pdf_exact(float g, float cos_t) {
    const float uni = kInvPiF * 0.25f;
    if (fabsf(g) < 1e-3f) return uni;
    const float sq = g * g;
    const float dd = (1.f + sq) - (g + g) * cos_t;
    return dd > 0.f ? (uni * (1.f - sq)) / (dd * __builtin_sqrtf(dd)) : 0.f;
}

// PhaseFunction::Sample, PhaseFunction.hxx:95-131, frame Frame::SetFromZ(dir) (Frame.hxx:56-62)
__device__ void phase_sample(float g, const float* dir, float u0, float u1, float* gen, float& pdf, float& sin_t) {
    if (fabsf(g) < 1e-3f) {
        uniform_sphere(u0, u1, gen);
        pdf = kInvPiF * 0.25f;
        const float c = (gen[0] * dir[0] + gen[1] * dir[1]) + gen[2] * dir[2];
        sin_t = __builtin_sqrtf(fmaxf(0.f, 1.f - c * c));
        return;
    }
    const float zl = __builtin_sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
    const float z[3] = {dir[0] / zl, dir[1] / zl, dir[2] / zl};
    const bool alt = fabsf(z[0]) > 0.99f;
    const float tx[3] = {alt ? 0.f : 1.f, alt ? 1.f : 0.f, 0.f};
    float y[3] = {z[1] * tx[2] - z[2] * tx[1], z[2] * tx[0] - z[0] * tx[2], z[0] * tx[1] - z[1] * tx[0]};
    const float yl = __builtin_sqrtf((y[0] * y[0] + y[1] * y[1]) + y[2] * y[2]);
    y[0] /= yl;
    y[1] /= yl;
    y[2] /= yl;
    const float x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
    const float sq = g * g, two = g + g;
    const float sqrtt = (1.f - sq) / (1.f - g + two * u0);
    const float cos_t = (1.f + sq - sqrtt * sqrtt) / two;
    const float st = __builtin_sqrtf(fmaxf(0.f, 1.f - cos_t * cos_t));
    const float phi = 2.f * kPiF * u1;
    const float cp = cosf(phi) * st, sp = sinf(phi) * st;
    for (int a = 0; a < 3; ++a) gen[a] = (x[a] * cp + y[a] * sp) + z[a] * cos_t;
    const float dd = 1.f + sq - two * cos_t;
    pdf = dd > 0.f ? (kInvPiF * 0.25f * (1.f - sq)) / (dd * __builtin_sqrtf(dd)) : 0.f;
    sin_t = st;
}

__device__ __forceinline__ float max_beam_length(const MediumDev& m) {
    // HomogeneousMedium constructor, Media.hxx:153-156
    const float mx = fmaxf(m.st[0], fmaxf(m.st[1], m.st[2]));
    return mx > 0.f ? (float)(20.723265836946411 / (double)mx) : INFINITY;
}

// The slots of one path, slot-major across paths (slot k of path p at k * path_count + p),
// so that neighbouring threads' slots are neighbours in memory for the tracer's writes and
// the packing reads.
template <class T>
struct Slots {
    T* base;
    int64_t stride;
    __device__ T& operator[](int64_t k) const { return base[k * stride]; }
};

__device__ void push_beam(const Slots<BeamRec>& b, int32_t& nb, const float* o, const float* d, float t0, float len,
                          const float* thr, uint32_t flags, int32_t med, int32_t lv, uint32_t rsf, float rsp, float rsr,
                          float lpwi) {
    BeamRec r;
    for (int a = 0; a < 3; ++a) {
        r.o[a] = o[a] + d[a] * t0;
        r.d[a] = d[a];
        r.thr[a] = thr[a];
    }
    r.len = len;
    r.flags = flags;
    r.medium = med;
    r.mis.light_vertex = lv;
    r.mis.ray_sampling_flags = rsf;
    r.mis.ray_sample_pdf = rsp;
    r.mis.ray_sample_rev_pdf = rsr;
    r.mis.last_pdf_w_inv = lpwi;
    r.mis.pad[0] = r.mis.pad[1] = r.mis.pad[2] = 0u;
    b[nb++] = r;
}

__global__ __launch_bounds__(kBlock) void k_trace_paths(LtScene S, const upbp_point_light* __restrict__ lights,
                                                         const MediumDev* __restrict__ media,
                                                         upbp_light_trace_params P, int64_t path_count,
                                                         int32_t cap_v, int32_t cap_b,
                                                         upbp_light_vertex* __restrict__ sv, float* __restrict__ sp,
                                                         BeamRec* __restrict__ sb, uint32_t* __restrict__ nv_out,
                                                         uint32_t* __restrict__ nb_out) {
    const int64_t path = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (path >= path_count) return;
    const Slots<upbp_light_vertex> V{sv + path, path_count};
    const Slots<float3> X{reinterpret_cast<float3*>(sp) + path, path_count};
    const Slots<BeamRec> B{sb + path, path_count};
    Rng rng{mix64(P.seed ^ ((uint64_t)path * 0xD1B54A32D192ED03ull)), 0};
    int32_t nv = 0, nb = 0;
    // GenerateLightSample
    const float pick = 1.f / (float)S.n_lights;
    int32_t lid = (int32_t)(rng.next() * (float)S.n_lights);
    if (lid >= S.n_lights) lid = S.n_lights - 1;
    const float du0 = rng.next(), du1 = rng.next();
    (void)rng.next();  // position samples, unused by point lights
    (void)rng.next();
    const upbp_point_light L = lights[lid];
    float dir[3];
    uniform_sphere(du0, du1, dir);
    const float emission_pdf = (kInvPiF * 0.25f) * pick;
    const float direct_pdf_a = 1.f * pick;
    {
        upbp_light_vertex v = {};
        v.throughput[0] = v.throughput[1] = v.throughput[2] = 1.f;
        v.path_idx = (int32_t)path;
        v.is_finite = 1;
        v.mis.pdf_a_inv = 1.f / direct_pdf_a;
        v.mis.ray_sample_rev_pdf_inv = 1.f;
        v.mis.cos_theta_out = 1.f;
        v.mis.flags = UPBP_MIS_DELTA | UPBP_MIS_ON_LIGHT_SOURCE;  // delta light: rev pdf 0
        V[0] = v;
        X[0] = make_float3(L.position[0], L.position[1], L.position[2]);
        nv = 1;
    }
    float thr[3] = {L.intensity[0] / emission_pdf, L.intensity[1] / emission_pdf, L.intensity[2] / emission_pdf};
    float org[3] = {L.position[0], L.position[1], L.position[2]};
    float last_pdf_w_inv = direct_pdf_a / emission_pdf;
    bool origin_in_medium = false;
    const bool want_beams = P.beam_path_count > 0 && path < P.beam_path_count;
    const bool short_beams = P.photon_beam_type == UPBP_SHORT_BEAM;
    for (uint32_t pl = 1;; ++pl) {
        const uint32_t rsf0 = origin_in_medium ? UPBP_ORIGIN_IN_MEDIUM : 0u;
        const int32_t lv = nv - 1;
        // Scene::Intersect, media part, fused with the SHORT-beam AddBeams pass (which reads the
        // same sampled segments in order with running products)
        float ray_pdf = 1.f, ray_rev = 1.f, att0 = 1.f, att1 = 1.f, att2 = 1.f;
        float bt[3] = {thr[0], thr[1], thr[2]};
        bool scattered = false;
        float dist = 0.f;
        int32_t hit_med = -1;
        float t = 0.f;
        for (bool first = true; t < kFar; first = false) {
            float t1;
            int32_t med;
            piece_at(S, org, dir, t, t1, med);
            const MediumDev m = media[med];
            const uint32_t flags = (origin_in_medium && t == 0.f) ? UPBP_ORIGIN_IN_MEDIUM : 0u;
            float pdf = 1.f, rev = 1.f, tmax = t1;
            bool here = false;
            if (m.min_comp > 0.f && m.has_scattering) {
                const float u = rng.next();
                const float s = -logf(u) / m.min_comp;
                const float len = t1 - t;
                if (s < len) {
                    const float a = expf(-m.min_comp * s);
                    pdf = m.min_comp * a;
                    rev = (flags & UPBP_ORIGIN_IN_MEDIUM) ? pdf : a;
                    tmax = t + s;
                    here = true;
                } else {
                    const float a = expf(-m.min_comp * len);
                    pdf = a;
                    rev = (flags & UPBP_ORIGIN_IN_MEDIUM) ? m.min_comp * a : a;
                }
            }
            const float sa0 = expf(-m.st[0] * (tmax - t)), sa1 = expf(-m.st[1] * (tmax - t)),
                        sa2 = expf(-m.st[2] * (tmax - t));
            if (want_beams && short_beams) {
                if (m.has_scattering && (!P.merge_pb2d || m.mfp > P.bb1d_min_mfp))
                    push_beam(B, nb, org, dir, t, tmax - t, bt, UPBP_SHORT_BEAM, med, lv,
                              UPBP_END_IN_MEDIUM | (first ? rsf0 : 0u), ray_pdf, ray_rev, last_pdf_w_inv);
                bt[0] *= sa0 / pdf;
                bt[1] *= sa1 / pdf;
                bt[2] *= sa2 / pdf;
            }
            ray_pdf *= pdf;
            ray_rev *= rev;
            att0 *= sa0;
            att1 *= sa1;
            att2 *= sa2;
            if (here) {
                scattered = true;
                dist = tmax;
                hit_med = med;
                break;
            }
            t = t1;
        }
        // AddBeams, LONG: every piece of the ray to its end, with RaySamplePdf (no sampling)
        if (want_beams && !short_beams) {
            float rsp = 1.f, rsr = 1.f;
            int32_t last = -1;
            float tt = 0.f;
            for (bool first = true; tt < kFar; first = false) {
                float t1;
                int32_t med;
                piece_at(S, org, dir, tt, t1, med);
                const MediumDev m = media[med];
                const float len = t1 - tt;
                if (m.has_scattering && (!P.merge_pb2d || m.mfp > P.bb1d_min_mfp)) {
                    last = nb;
                    push_beam(B, nb, org, dir, tt, len, bt, UPBP_LONG_BEAM, med, lv,
                              UPBP_END_IN_MEDIUM | (first ? rsf0 : 0u), rsp, rsr, last_pdf_w_inv);
                }
                bt[0] *= expf(-m.st[0] * len);
                bt[1] *= expf(-m.st[1] * len);
                bt[2] *= expf(-m.st[2] * len);
                float spdf = 1.f, srev = 1.f;
                if (m.min_comp > 0.f && m.has_scattering) {
                    const float a = fmaxf(expf(-m.min_comp * len), 1e-35f);
                    const float minatt = m.min_comp * a;
                    spdf = a;
                    srev = ((first ? rsf0 : 0u) & UPBP_ORIGIN_IN_MEDIUM) ? minatt : a;
                }
                rsp *= spdf;
                rsr *= srev;
                tt = t1;
            }
            if (last >= 0) {  // the last beam stored is clamped to its medium's MaxBeamLength
                const float mbl = max_beam_length(media[B[last].medium]);
                if (B[last].len > mbl) B[last].len = mbl;
            }
        }
        if (!scattered) break;
        thr[0] *= att0 / ray_pdf;
        thr[1] *= att1 / ray_pdf;
        thr[2] *= att2 / ray_pdf;
        if (thr[0] <= 0.f && thr[1] <= 0.f && thr[2] <= 0.f) break;
        const MediumDev m = media[hit_med];
        const float hit[3] = {org[0] + dir[0] * dist, org[1] + dir[1] * dist, org[2] + dir[2] * dist};
        origin_in_medium = true;
        // store the vertex (UPBP.hxx:563-636)
        const float dist_sq = dist * dist;  // point lights are finite
        const float ray_pdf_inv = 1.f / ray_pdf;
        upbp_light_vertex v = {};
        for (int c = 0; c < 3; ++c) {
            v.throughput[c] = thr[c];
            v.dir_fix[c] = -dir[c];
        }
        X[nv] = make_float3(hit[0], hit[1], hit[2]);
        v.continuation_prob = m.cont;
        v.path_idx = (int32_t)path;
        v.path_length = pl;
        v.is_finite = 1;
        v.flags = UPBP_LV_CONNECTABLE | UPBP_LV_IN_MEDIUM;
        v.mis.pdf_a_inv = last_pdf_w_inv * dist_sq * ray_pdf_inv / 1.f;
        v.mis.rev_pdf_a = 1.f;
        v.mis.rev_pdf_a_without_bsdf = 1.f;
        v.mis.ray_sample_pdf_inv = ray_pdf_inv;
        v.mis.ray_sample_rev_pdf_inv = 1.f;
        v.mis.pp3d_factor = P.pp3d_factor;
        v.mis.pb2d_factor = P.pb2d_factor;
        v.mis.bb1d_factor = P.bb1d_factor;
        v.mis.bb1d_beam_selection_pdf = 1.f;
        v.mis.flags = (!P.merge_pb2d || m.mfp > P.bb1d_min_mfp) ? UPBP_MIS_IN_MEDIUM_WITH_BEAMS : 0u;
        v.mis.ray_sample_pdfs_ratio = 1.f / m.min_comp;
        v.mis.ray_sample_rev_pdfs_ratio = v.mis.ray_sample_pdfs_ratio;
        upbp_mis_data& pm = V[nv - 1].mis;
        pm.rev_pdf_a *= ray_rev / dist_sq;
        pm.rev_pdf_a_without_bsdf = pm.rev_pdf_a;
        pm.ray_sample_rev_pdf_inv = 1.f / ray_rev;
        const int32_t cur = nv;
        V[cur] = v;
        nv += 1;
        if (pl + 2 > P.max_path_length) break;
        // SampleScattering with the medium BSDF
        const float u0 = rng.next(), u1 = rng.next();
        (void)rng.next();  // component sample, unused in media
        float gen[3], pdf, sin_t;
        phase_sample(m.g, dir, u0, u1, gen, pdf, sin_t);
        const float f0 = m.ss[0] * pdf, f1 = m.ss[1] * pdf, f2 = m.ss[2] * pdf;
        if (f0 <= 0.f && f1 <= 0.f && f2 <= 0.f) break;
        float rev_pdf = phase_pdf_exact(m.g, (gen[0] * dir[0] + gen[1] * dir[1]) + gen[2] * dir[2]);
        const float cont = m.cont;
        if (cont == 0.f || (cont < 1.f && rng.next() > cont)) break;
        const float dir_pdf = pdf * cont;
        rev_pdf *= cont;
        const float dir_pdf_inv = 1.f / dir_pdf;
        thr[0] *= f0 * (1.f * dir_pdf_inv);
        thr[1] *= f1 * (1.f * dir_pdf_inv);
        thr[2] *= f2 * (1.f * dir_pdf_inv);
        last_pdf_w_inv = dir_pdf_inv;
        upbp_mis_data& cm = V[cur].mis;
        cm.rev_pdf_a *= 1.f;
        cm.rev_pdf_a_without_bsdf = cm.rev_pdf_a;
        cm.sin_theta = sin_t;
        cm.cos_theta_out = 1.f;
        V[cur - 1].mis.rev_pdf_a *= rev_pdf;
        for (int a = 0; a < 3; ++a) {
            org[a] = hit[a];
            dir[a] = gen[a];
        }
    }
    nv_out[path] = (uint32_t)nv;
    nb_out[path] = (uint32_t)nb;
}

// Packs path p's used slots at its scanned offsets (one thread per path; the slot-major
// layout makes a wave's reads of slot k contiguous) and writes mPathEnds.
__global__ void k_pack_paths(int64_t path_count, const upbp_light_vertex* __restrict__ sv, const float3* __restrict__ sp,
                             const BeamRec* __restrict__ sb, const uint32_t* __restrict__ nv,
                             const uint32_t* __restrict__ nb, const uint32_t* __restrict__ ov,
                             const uint32_t* __restrict__ ob, upbp_light_paths out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= path_count) return;
    const int64_t v0 = ov[p], b0 = ob[p];
    const uint32_t n_v = nv[p], n_b = nb[p];
    for (uint32_t k = 0; k < n_v; ++k) {
        out.vertices[v0 + k] = sv[k * path_count + p];
        const float3 x = sp[k * path_count + p];
        out.positions[3 * (v0 + k)] = x.x;
        out.positions[3 * (v0 + k) + 1] = x.y;
        out.positions[3 * (v0 + k) + 2] = x.z;
    }
    for (uint32_t k = 0; k < n_b; ++k) {
        const BeamRec r = sb[k * path_count + p];
        const int64_t d = b0 + k;
        for (int a = 0; a < 3; ++a) {
            out.beam_origin[3 * d + a] = r.o[a];
            out.beam_direction[3 * d + a] = r.d[a];
            out.beam_throughput[3 * d + a] = r.thr[a];
        }
        out.beam_length[d] = r.len;
        out.beam_flags[d] = r.flags;
        out.beam_medium[d] = r.medium;
        upbp_beam_mis m = r.mis;
        m.light_vertex += (int32_t)v0;
        out.beam_mis[d] = m;
    }
    out.path_ends[p] = (int32_t)(v0 + n_v);
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

}  // namespace

void trace_light_paths(Ctx& ctx, const upbp_region* regions, int32_t n_regions, const upbp_point_light* lights,
                       int32_t n_lights, const upbp_light_trace_params& P, int64_t path_count,
                       upbp_light_paths& out) {
    out.vertex_count = 0;
    out.beam_count = 0;
    if (path_count == 0) return;
    if (P.max_path_length < 2) {  // the reference skips light tracing (UPBP.hxx:501)
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(path_count)), dim3(kBlock), 0, ctx.stream, out.path_ends, 0,
                           path_count);
        UPBP_HIP(hipGetLastError());
        return;
    }
    LtScene S = {};
    for (int i = 0; i < n_regions; ++i) S.regions[i] = regions[i];
    S.n_regions = n_regions;
    S.global_medium = P.global_medium;
    S.n_lights = n_lights;
    const int32_t cap_v = (int32_t)P.max_path_length;
    const int32_t cap_b = (int32_t)((P.max_path_length - 1) * (2 * n_regions + 1));
    UPBP_REQUIRE(path_count * (int64_t)cap_v < ((int64_t)1 << 31) && path_count * (int64_t)cap_b < ((int64_t)1 << 31),
                 UPBP_EINVAL, "too many light paths for one call (path count x path length slots must fit 2^31)");
    // the slots stay with the context (grow-only), so repeated calls allocate nothing
    DevBuf &sv = ctx.lt_slots[0], &sp = ctx.lt_slots[1], &sb = ctx.lt_slots[2], &nv = ctx.lt_slots[3],
           &nb = ctx.lt_slots[4], &ov = ctx.lt_slots[5], &ob = ctx.lt_slots[6], &dl = ctx.lt_slots[7];
    dl.ensure((size_t)n_lights * sizeof(upbp_point_light));
    UPBP_HIP(hipMemcpyAsync(dl.p, lights, (size_t)n_lights * sizeof(upbp_point_light), hipMemcpyHostToDevice,
                            ctx.stream));
    sv.ensure((size_t)path_count * cap_v * sizeof(upbp_light_vertex));
    sp.ensure((size_t)path_count * cap_v * 3 * sizeof(float));
    sb.ensure((size_t)std::max<int64_t>(path_count * cap_b, 1) * sizeof(BeamRec));
    nv.ensure((size_t)path_count * 4);
    nb.ensure((size_t)path_count * 4);
    ov.ensure((size_t)path_count * 4);
    ob.ensure((size_t)path_count * 4);
    hipLaunchKernelGGL(k_trace_paths, dim3(blocks_for(path_count)), dim3(kBlock), 0, ctx.stream, S,
                       dl.as<upbp_point_light>(), ctx.media.as<MediumDev>(), P, path_count, cap_v, cap_b,
                       sv.as<upbp_light_vertex>(), sp.as<float>(), sb.as<BeamRec>(), nv.as<uint32_t>(),
                       nb.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    exclusive_scan_u32(ctx, nv.as<uint32_t>(), ov.as<uint32_t>(), path_count);
    exclusive_scan_u32(ctx, nb.as<uint32_t>(), ob.as<uint32_t>(), path_count);
    uint32_t tail[4];
    UPBP_HIP(hipMemcpyAsync(&tail[0], ov.as<uint32_t>() + (path_count - 1), 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipMemcpyAsync(&tail[1], nv.as<uint32_t>() + (path_count - 1), 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipMemcpyAsync(&tail[2], ob.as<uint32_t>() + (path_count - 1), 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipMemcpyAsync(&tail[3], nb.as<uint32_t>() + (path_count - 1), 4, hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
    out.vertex_count = (int64_t)tail[0] + tail[1];
    out.beam_count = (int64_t)tail[2] + tail[3];
    UPBP_REQUIRE(out.vertex_count <= out.vertex_capacity && out.beam_count <= out.beam_capacity, UPBP_EINVAL,
                 "light path output capacity too small (vertex_count / beam_count hold what is needed)");
    hipLaunchKernelGGL(k_pack_paths, dim3(blocks_for(path_count)), dim3(kBlock), 0, ctx.stream, path_count,
                       sv.as<upbp_light_vertex>(), sp.as<float3>(), sb.as<BeamRec>(), nv.as<uint32_t>(),
                       nb.as<uint32_t>(), ov.as<uint32_t>(), ob.as<uint32_t>(), out);
    UPBP_HIP(hipGetLastError());
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
}

}  // namespace upbp
