/* lighttrace.c -- CPU restatement of light sub-path tracing in media (TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU comparison use it).
 *
 * Follows UPBPRenderer's light-tracing loop (SmallUPBP/src/Renderers/UPBP.hxx:501-674) with
 * GenerateLightSample (:1592-1662) for point lights (Scene/Lights.hxx:363-384), AddBeams
 * (:1783-1870), SampleScattering (:1908-1969) with the medium BSDF (Path/Bsdf.hxx:140-151,
 * 385-401) and PhaseFunction::Sample (Path/PhaseFunction.hxx:95-131), and the media part of
 * Scene::Intersect (Scene/Scene.hxx:393-531) with HomogeneousMedium::SampleRay / RaySamplePdf
 * (Scene/Media.hxx:240-323), for homogeneous media bounded by imaginary spheres and boxes:
 * the ray is cut at every boundary crossing and each piece takes the medium of the innermost
 * region containing its midpoint. Random numbers: a counter-based SplitMix64 stream per path
 * (the same on the GPU), in the reference's draw order. FP32 throughout, the operations of
 * the GPU kernel. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "upbp_oracle.h"

#define LT_PI 3.141592653589793238462643383279f
#define LT_INV_PI (1.f / LT_PI)
#define LT_INF 1e36f

/* ---- counter-based random numbers: draw k of path p ------------------------------- */
static inline uint64_t lt_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t base;
    uint64_t k;
} lt_rng;

static inline lt_rng lt_rng_for(uint64_t seed, int64_t path) {
    lt_rng r;
    r.base = lt_mix(seed ^ ((uint64_t)path * 0xD1B54A32D192ED03ull));
    r.k = 0;
    return r;
}

/* uniform in (0, 1): 24 random bits + 1/2, exact in float */
static inline float lt_next(lt_rng* r) {
    r->k += 1;
    const uint64_t h = lt_mix(r->base + r->k * 0x9E3779B97F4A7C15ull);
    return ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

/* ---- regions and ray segments ------------------------------------------------------ */
static int lt_contains(const upbp_region* g, const float* p) {
    if (g->shape == UPBP_REGION_SPHERE) {
        const float dx = p[0] - g->p0[0], dy = p[1] - g->p0[1], dz = p[2] - g->p0[2];
        return (dx * dx + dy * dy) + dz * dz <= g->p1[0] * g->p1[0];
    }
    return p[0] >= g->p0[0] && p[0] <= g->p1[0] && p[1] >= g->p0[1] && p[1] <= g->p1[1] && p[2] >= g->p0[2] &&
           p[2] <= g->p1[2];
}

static int32_t lt_medium_at(const upbp_region* rg, int32_t n, int32_t global, const float* p) {
    int32_t m = global;
    for (int32_t i = 0; i < n; ++i)
        if (lt_contains(&rg[i], p)) m = rg[i].medium_id;
    return m;
}

/* Closest call of a path's branch decisions (test support): each comparison a < b that picks
 * a branch lowers *margin to |a - b| / scale, the relative distance of the two sides. A path
 * whose every margin is large takes the same branches under last-ulp differences of the
 * math library; only a path with a small margin can take another one. */
static inline void lt_margin(float* margin, float a, float b, float scale) {
    if (!margin) return;
    const float m = fabsf(a - b) / fmaxf(scale, 1e-30f);
    if (m < *margin) *margin = m;
}

/* positive distances at which the ray crosses the region's boundary */
static int lt_crossings(const upbp_region* g, const float* o, const float* d, float* t, float* margin) {
    int n = 0;
    if (g->shape == UPBP_REGION_SPHERE) {
        const float ox = o[0] - g->p0[0], oy = o[1] - g->p0[1], oz = o[2] - g->p0[2];
        const float b = (ox * d[0] + oy * d[1]) + oz * d[2];
        const float c = ((ox * ox + oy * oy) + oz * oz) - g->p1[0] * g->p1[0];
        const float disc = b * b - c;
        lt_margin(margin, disc, 0.f, fmaxf(b * b, fabsf(c)));
        if (disc < 0.f) return 0;
        const float s = sqrtf(disc);
        const float t0 = -b - s, t1 = -b + s;
        lt_margin(margin, t0, 0.f, fabsf(b) + s);
        lt_margin(margin, t1, 0.f, fabsf(b) + s);
        if (t0 > 0.f) t[n++] = t0;
        if (t1 > 0.f) t[n++] = t1;
        return n;
    }
    float tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        if (d[a] != 0.f) {
            float a0 = (g->p0[a] - o[a]) / d[a], a1 = (g->p1[a] - o[a]) / d[a];
            if (a0 > a1) {
                const float x = a0;
                a0 = a1;
                a1 = x;
            }
            tn = fmaxf(tn, a0);
            tf = fminf(tf, a1);
        } else if (o[a] < g->p0[a] || o[a] > g->p1[a]) {
            return 0;
        }
    }
    lt_margin(margin, tn, tf, fmaxf(fabsf(tn), fabsf(tf)));
    if (!(tn <= tf)) return 0;
    lt_margin(margin, tn, 0.f, fabsf(tn) + fabsf(tf));
    lt_margin(margin, tf, 0.f, fabsf(tn) + fabsf(tf));
    if (tn > 0.f) t[n++] = tn;
    if (tf > 0.f) t[n++] = tf;
    return n;
}

typedef struct {
    float t0, t1;
    int32_t medium;
} lt_piece;

/* VolumeSegmentLite list of the ray: [0, c1], [c1, c2], ..., [cn, LT_INF) */
static int lt_cut(const upbp_region* rg, int32_t nr, int32_t global, const float* o, const float* d, lt_piece* out,
                  float* margin) {
    float c[2 * LT_MAX_REGIONS];
    int n = 0;
    for (int32_t i = 0; i < nr; ++i) n += lt_crossings(&rg[i], o, d, c + n, margin);
    for (int i = 1; i < n; ++i) { /* insertion sort */
        const float x = c[i];
        int j = i - 1;
        while (j >= 0 && c[j] > x) {
            c[j + 1] = c[j];
            --j;
        }
        c[j + 1] = x;
    }
    int m = 0;
    float a = 0.f;
    for (int i = 0; i <= n; ++i) {
        const float b = i < n ? c[i] : LT_INF;
        if (i < n && b == a) continue; /* repeated crossing */
        const float mid = i < n ? 0.5f * (a + b) : a + 1.f;
        const float p[3] = {o[0] + d[0] * mid, o[1] + d[1] * mid, o[2] + d[2] * mid};
        out[m].t0 = a;
        out[m].t1 = b;
        out[m].medium = lt_medium_at(rg, nr, global, p);
        ++m;
        a = b;
    }
    return m;
}

/* ---- sampling ------------------------------------------------------------------------ */
/* SampleUniformSphereW, Misc/Utils2.hxx:224-241 */
static inline void lt_uniform_sphere(float u0, float u1, float* dir) {
    const float term1 = 2.f * LT_PI * u0;
    const float term2 = 2.f * sqrtf(u1 - u1 * u1);
    dir[0] = cosf(term1) * term2;
    dir[1] = sinf(term1) * term2;
    dir[2] = 1.f - 2.f * u1;
}

/* PhaseFunction::Pdf with cos = -dot(gen, fix) (PhaseFunction.hxx:52-80) */
static inline float lt_phase_pdf(float g, float cos_t) {
    const float uni = LT_INV_PI * 0.25f;
    if (fabsf(g) < 1e-3f) return uni;
    const float sq = g * g;
    const float dd = (1.f + sq) - (g + g) * cos_t;
    return dd > 0.f ? (uni * (1.f - sq)) / (dd * sqrtf(dd)) : 0.f;
}

/* PhaseFunction::Sample (PhaseFunction.hxx:95-131); dir = the incoming ray direction
 * (WorldDirFix = -dir), frame from Frame::SetFromZ(dir) (Path/Frame.hxx:56-62) */
static void lt_phase_sample(float g, const float* dir, float u0, float u1, float* gen, float* pdf, float* sin_t) {
    if (fabsf(g) < 1e-3f) {
        lt_uniform_sphere(u0, u1, gen);
        *pdf = LT_INV_PI * 0.25f;
        const float c = (gen[0] * dir[0] + gen[1] * dir[1]) + gen[2] * dir[2]; /* -dot(gen, fix) */
        *sin_t = sqrtf(fmaxf(0.f, 1.f - c * c));
        return;
    }
    const float zl = sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
    const float z[3] = {dir[0] / zl, dir[1] / zl, dir[2] / zl};
    const float tx[3] = {fabsf(z[0]) > 0.99f ? 0.f : 1.f, fabsf(z[0]) > 0.99f ? 1.f : 0.f, 0.f};
    float y[3] = {z[1] * tx[2] - z[2] * tx[1], z[2] * tx[0] - z[0] * tx[2], z[0] * tx[1] - z[1] * tx[0]};
    const float yl = sqrtf((y[0] * y[0] + y[1] * y[1]) + y[2] * y[2]);
    y[0] /= yl;
    y[1] /= yl;
    y[2] /= yl;
    const float x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
    const float sq = g * g, two = g + g;
    const float sqrtt = (1.f - sq) / (1.f - g + two * u0);
    const float cos_t = (1.f + sq - sqrtt * sqrtt) / two;
    const float st = sqrtf(fmaxf(0.f, 1.f - cos_t * cos_t));
    const float phi = 2.f * LT_PI * u1;
    const float cp = cosf(phi) * st, sp = sinf(phi) * st;
    for (int a = 0; a < 3; ++a) gen[a] = (x[a] * cp + y[a] * sp) + z[a] * cos_t;
    const float dd = 1.f + sq - two * cos_t;
    *pdf = dd > 0.f ? (LT_INV_PI * 0.25f * (1.f - sq)) / (dd * sqrtf(dd)) : 0.f;
    *sin_t = st;
}

/* ---- one light path ---------------------------------------------------------------- */
typedef struct {
    upbp_light_vertex* v; /* [cap_v] */
    float* pos;           /* [cap_v][3] */
    int32_t nv;
    float* b_o; /* [cap_b][3] */
    float* b_d;
    float* b_len;
    float* b_thr;
    uint32_t* b_flags;
    int32_t* b_med;
    upbp_beam_mis* b_mis; /* light_vertex = index inside the path */
    int32_t nb;
    float* margin; /* closest call of the path's branch decisions, or NULL */
} lt_out;

static void lt_push_beam(lt_out* o, const float* org, const float* dir, float len, const float* thr, uint32_t flags,
                         int32_t med, int32_t lv, uint32_t rsf, float rsp, float rsr, float lpwi) {
    const int32_t k = o->nb++;
    for (int a = 0; a < 3; ++a) {
        o->b_o[3 * k + a] = org[a];
        o->b_d[3 * k + a] = dir[a];
        o->b_thr[3 * k + a] = thr[a];
    }
    o->b_len[k] = len;
    o->b_flags[k] = flags;
    o->b_med[k] = med;
    memset(&o->b_mis[k], 0, sizeof(upbp_beam_mis));
    o->b_mis[k].light_vertex = lv;
    o->b_mis[k].ray_sampling_flags = rsf;
    o->b_mis[k].ray_sample_pdf = rsp;
    o->b_mis[k].ray_sample_rev_pdf = rsr;
    o->b_mis[k].last_pdf_w_inv = lpwi;
}

static void lt_trace_path(const or_medium* media, const upbp_region* rg, int32_t nr, const upbp_point_light* lights,
                          int32_t nl, const upbp_light_trace_params* P, int64_t path, lt_out* o) {
    lt_rng rng = lt_rng_for(P->seed, path);
    o->nv = 0;
    o->nb = 0;
    /* GenerateLightSample */
    const float pick = 1.f / (float)nl;
    int32_t lid = (int32_t)(lt_next(&rng) * (float)nl);
    if (lid >= nl) lid = nl - 1;
    const float du0 = lt_next(&rng), du1 = lt_next(&rng);
    (void)lt_next(&rng); /* position samples, unused by point lights */
    (void)lt_next(&rng);
    const upbp_point_light* L = &lights[lid];
    float dir[3];
    lt_uniform_sphere(du0, du1, dir);
    const float emission_pdf = (LT_INV_PI * 0.25f) * pick;
    const float direct_pdf_a = 1.f * pick;
    {
        upbp_light_vertex* v = &o->v[0];
        memset(v, 0, sizeof(*v));
        v->throughput[0] = v->throughput[1] = v->throughput[2] = 1.f;
        v->path_idx = (int32_t)path;
        v->path_length = 0;
        v->is_finite = 1;
        v->mis.pdf_a_inv = 1.f / direct_pdf_a;
        v->mis.rev_pdf_a = 0.f; /* delta light */
        v->mis.rev_pdf_a_without_bsdf = 0.f;
        v->mis.ray_sample_rev_pdf_inv = 1.f;
        v->mis.cos_theta_out = 1.f;
        v->mis.flags = UPBP_MIS_DELTA | UPBP_MIS_ON_LIGHT_SOURCE;
        for (int a = 0; a < 3; ++a) o->pos[a] = L->position[a];
        o->nv = 1;
    }
    float thr[3] = {L->intensity[0] / emission_pdf, L->intensity[1] / emission_pdf, L->intensity[2] / emission_pdf};
    float org[3] = {L->position[0], L->position[1], L->position[2]};
    float last_pdf_w_inv = direct_pdf_a / emission_pdf;
    int origin_in_medium = 0;
    const int want_beams = P->beam_path_count > 0 && path < P->beam_path_count;
    for (uint32_t pl = 1;; ++pl) {
        lt_piece seg[2 * LT_MAX_REGIONS + 1];
        const int ns = lt_cut(rg, nr, P->global_medium, org, dir, seg, o->margin);
        /* Scene::Intersect, media part: sample distances segment by segment */
        float vs_t0[2 * LT_MAX_REGIONS + 1], vs_t1[2 * LT_MAX_REGIONS + 1], vs_pdf[2 * LT_MAX_REGIONS + 1],
            vs_rev[2 * LT_MAX_REGIONS + 1], vs_att[2 * LT_MAX_REGIONS + 1][3];
        int32_t vs_med[2 * LT_MAX_REGIONS + 1];
        int nvs = 0, scattered = 0;
        for (int i = 0; i < ns; ++i) {
            const or_medium* m = &media[seg[i].medium];
            const uint32_t flags = (origin_in_medium && seg[i].t0 == 0.f) ? UPBP_ORIGIN_IN_MEDIUM : 0u;
            float pdf = 1.f, rev = 1.f, tmax = seg[i].t1;
            if (m->min_comp > 0.f && m->has_scattering) {
                const float u = lt_next(&rng);
                const float s = -logf(u) / m->min_comp;
                const float len = seg[i].t1 - seg[i].t0;
                lt_margin(o->margin, s, len, fmaxf(s, len));
                if (s < len) {
                    const float att = expf(-m->min_comp * s);
                    pdf = m->min_comp * att;
                    rev = (flags & UPBP_ORIGIN_IN_MEDIUM) ? pdf : att;
                    tmax = seg[i].t0 + s;
                    scattered = 1;
                } else {
                    const float att = expf(-m->min_comp * len);
                    pdf = att;
                    rev = (flags & UPBP_ORIGIN_IN_MEDIUM) ? m->min_comp * att : att;
                }
            }
            vs_t0[nvs] = seg[i].t0;
            vs_t1[nvs] = tmax;
            vs_pdf[nvs] = pdf;
            vs_rev[nvs] = rev;
            for (int c = 0; c < 3; ++c) vs_att[nvs][c] = expf(-m->sigma_t[c] * (tmax - seg[i].t0));
            vs_med[nvs] = seg[i].medium;
            ++nvs;
            if (scattered) break;
        }
        /* AddBeams */
        if (want_beams) {
            const uint32_t rsf0 = origin_in_medium ? UPBP_ORIGIN_IN_MEDIUM : 0u;
            float bt[3] = {thr[0], thr[1], thr[2]};
            float rsp = 1.f, rsr = 1.f;
            const int32_t lv = o->nv - 1;
            if (P->photon_beam_type == UPBP_SHORT_BEAM) {
                for (int i = 0; i < nvs; ++i) {
                    const or_medium* m = &media[vs_med[i]];
                    if (m->has_scattering && (!P->merge_pb2d || m->mfp > P->bb1d_min_mfp)) {
                        const float bo[3] = {org[0] + dir[0] * vs_t0[i], org[1] + dir[1] * vs_t0[i],
                                             org[2] + dir[2] * vs_t0[i]};
                        lt_push_beam(o, bo, dir, vs_t1[i] - vs_t0[i], bt, UPBP_SHORT_BEAM, vs_med[i], lv,
                                     UPBP_END_IN_MEDIUM | (i == 0 ? rsf0 : 0u), rsp, rsr, last_pdf_w_inv);
                    }
                    for (int c = 0; c < 3; ++c) bt[c] *= vs_att[i][c] / vs_pdf[i];
                    rsp *= vs_pdf[i];
                    rsr *= vs_rev[i];
                }
            } else {
                int added = 0;
                for (int i = 0; i < ns; ++i) {
                    const or_medium* m = &media[seg[i].medium];
                    const float len = seg[i].t1 - seg[i].t0;
                    if (m->has_scattering && (!P->merge_pb2d || m->mfp > P->bb1d_min_mfp)) {
                        const float bo[3] = {org[0] + dir[0] * seg[i].t0, org[1] + dir[1] * seg[i].t0,
                                             org[2] + dir[2] * seg[i].t0};
                        lt_push_beam(o, bo, dir, len, bt, UPBP_LONG_BEAM, seg[i].medium, lv,
                                     UPBP_END_IN_MEDIUM | (i == 0 ? rsf0 : 0u), rsp, rsr, last_pdf_w_inv);
                        added = 1;
                    }
                    for (int c = 0; c < 3; ++c) bt[c] *= expf(-m->sigma_t[c] * len);
                    float sp = 1.f, sr = 1.f;
                    if (m->min_comp > 0.f && m->has_scattering) {
                        const float att = fmaxf(expf(-m->min_comp * len), 1e-35f);
                        const float minatt = m->min_comp * att;
                        sp = att; /* no kEndInMedium for light rays */
                        sr = ((i == 0 ? rsf0 : 0u) & UPBP_ORIGIN_IN_MEDIUM) ? minatt : att;
                    }
                    rsp *= sp;
                    rsr *= sr;
                }
                if (added) { /* the last beam stored is clamped to its medium's MaxBeamLength */
                    const int32_t k = o->nb - 1;
                    const float mbl = media[o->b_med[k]].max_beam_length;
                    if (o->b_len[k] > mbl) o->b_len[k] = mbl;
                }
            }
        }
        if (!scattered) break;
        float ray_pdf = 1.f, ray_rev = 1.f, att[3] = {1.f, 1.f, 1.f};
        for (int i = 0; i < nvs; ++i) {
            ray_pdf *= vs_pdf[i];
            ray_rev *= vs_rev[i];
            for (int c = 0; c < 3; ++c) att[c] *= vs_att[i][c];
        }
        for (int c = 0; c < 3; ++c) thr[c] *= att[c] / ray_pdf;
        if (thr[0] <= 0.f && thr[1] <= 0.f && thr[2] <= 0.f) break;
        const float dist = vs_t1[nvs - 1];
        const int32_t mid = vs_med[nvs - 1];
        const or_medium* m = &media[mid];
        const float hit[3] = {org[0] + dir[0] * dist, org[1] + dir[1] * dist, org[2] + dir[2] * dist};
        origin_in_medium = 1;
        /* store the vertex */
        const float dist_sq = dist * dist; /* point lights are finite */
        const float ray_pdf_inv = 1.f / ray_pdf;
        upbp_light_vertex* v = &o->v[o->nv];
        memset(v, 0, sizeof(*v));
        for (int c = 0; c < 3; ++c) {
            v->throughput[c] = thr[c];
            v->dir_fix[c] = -dir[c];
            o->pos[3 * o->nv + c] = hit[c];
        }
        v->continuation_prob = m->continuation_prob;
        v->path_idx = (int32_t)path;
        v->path_length = pl;
        v->is_finite = 1;
        v->flags = UPBP_LV_CONNECTABLE | UPBP_LV_IN_MEDIUM;
        v->mis.pdf_a_inv = last_pdf_w_inv * dist_sq * ray_pdf_inv / 1.f;
        v->mis.rev_pdf_a = 1.f;
        v->mis.rev_pdf_a_without_bsdf = 1.f;
        v->mis.ray_sample_pdf_inv = ray_pdf_inv;
        v->mis.ray_sample_rev_pdf_inv = 1.f;
        v->mis.pp3d_factor = P->pp3d_factor;
        v->mis.pb2d_factor = P->pb2d_factor;
        v->mis.bb1d_factor = P->bb1d_factor;
        v->mis.bb1d_beam_selection_pdf = 1.f;
        v->mis.flags = (!P->merge_pb2d || m->mfp > P->bb1d_min_mfp) ? UPBP_MIS_IN_MEDIUM_WITH_BEAMS : 0u;
        v->mis.ray_sample_pdfs_ratio = 1.f / m->min_comp;
        v->mis.ray_sample_rev_pdfs_ratio = v->mis.ray_sample_pdfs_ratio;
        upbp_light_vertex* pv = &o->v[o->nv - 1];
        pv->mis.rev_pdf_a *= ray_rev / dist_sq;
        pv->mis.rev_pdf_a_without_bsdf = pv->mis.rev_pdf_a;
        pv->mis.ray_sample_rev_pdf_inv = 1.f / ray_rev;
        o->nv += 1;
        if (pl + 2 > P->max_path_length) break;
        /* SampleScattering with the medium BSDF */
        const float u0 = lt_next(&rng), u1 = lt_next(&rng);
        (void)lt_next(&rng); /* component sample, unused in media */
        float gen[3], pdf, sin_t;
        lt_phase_sample(m->mean_cosine, dir, u0, u1, gen, &pdf, &sin_t);
        const float f[3] = {m->sigma_s[0] * pdf, m->sigma_s[1] * pdf, m->sigma_s[2] * pdf};
        if (f[0] <= 0.f && f[1] <= 0.f && f[2] <= 0.f) break;
        float rev_pdf = lt_phase_pdf(m->mean_cosine, (gen[0] * dir[0] + gen[1] * dir[1]) + gen[2] * dir[2]);
        const float cont = m->continuation_prob;
        if (cont == 0.f) break;
        if (cont < 1.f) {
            const float u = lt_next(&rng);
            lt_margin(o->margin, u, cont, fabsf(cont));
            if (u > cont) break;
        }
        float dir_pdf = pdf * cont;
        rev_pdf *= cont;
        const float dir_pdf_inv = 1.f / dir_pdf;
        for (int c = 0; c < 3; ++c) thr[c] *= f[c] * (1.f * dir_pdf_inv);
        last_pdf_w_inv = dir_pdf_inv;
        v->mis.rev_pdf_a *= 1.f;
        v->mis.rev_pdf_a_without_bsdf = v->mis.rev_pdf_a;
        v->mis.sin_theta = sin_t;
        v->mis.cos_theta_out = 1.f;
        pv->mis.rev_pdf_a *= rev_pdf;
        for (int a = 0; a < 3; ++a) {
            org[a] = hit[a];
            dir[a] = gen[a];
        }
    }
}

int or_trace_light_paths(const upbp_medium* media, int32_t n_media, const upbp_region* regions, int32_t n_regions,
                         const upbp_point_light* lights, int32_t n_lights, const upbp_light_trace_params* P,
                         int64_t path_count, int threads, upbp_light_paths* out) {
    return or_trace_light_paths_margins(media, n_media, regions, n_regions, lights, n_lights, P, path_count, threads,
                                        out, NULL);
}

int or_trace_light_paths_margins(const upbp_medium* media, int32_t n_media, const upbp_region* regions,
                                 int32_t n_regions, const upbp_point_light* lights, int32_t n_lights,
                                 const upbp_light_trace_params* P, int64_t path_count, int threads,
                                 upbp_light_paths* out, float* margins) {
    if (!media || n_media < 1 || n_regions < 0 || n_regions > LT_MAX_REGIONS || (n_regions && !regions) ||
        !lights || n_lights < 1 || !P || path_count < 0 || !out || (path_count > 0 && !out->path_ends) ||
        P->global_medium < 0 || P->global_medium >= n_media)
        return UPBP_EINVAL;
    for (int32_t i = 0; i < n_regions; ++i)
        if (regions[i].medium_id < 0 || regions[i].medium_id >= n_media) return UPBP_EINVAL;
    if (P->max_path_length < 2) { /* the reference skips light tracing (UPBP.hxx:501) */
        out->vertex_count = 0;
        out->beam_count = 0;
        for (int64_t p = 0; p < path_count; ++p) out->path_ends[p] = 0;
        if (margins)
            for (int64_t p = 0; p < path_count; ++p) margins[p] = INFINITY;
        return 0;
    }
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    const int32_t cap_v = (int32_t)P->max_path_length;
    const int32_t cap_b = (int32_t)(P->max_path_length * (2 * n_regions + 1));
    upbp_light_vertex* sv = (upbp_light_vertex*)malloc(sizeof(upbp_light_vertex) * (size_t)(path_count * cap_v + 1));
    float* sp = (float*)malloc(sizeof(float) * 3 * (size_t)(path_count * cap_v + 1));
    float* sb = (float*)malloc(sizeof(float) * 10 * (size_t)(path_count * cap_b + 1));
    uint32_t* sf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(path_count * cap_b + 1));
    int32_t* smed = (int32_t*)malloc(sizeof(int32_t) * (size_t)(path_count * cap_b + 1));
    upbp_beam_mis* smis = (upbp_beam_mis*)malloc(sizeof(upbp_beam_mis) * (size_t)(path_count * cap_b + 1));
    int32_t* nv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(path_count + 1));
    int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(path_count + 1));
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t p = 0; p < path_count; ++p) {
        const int64_t bv = p * cap_v, bb = p * cap_b;
        lt_out o;
        o.v = sv + bv;
        o.pos = sp + 3 * bv;
        o.b_o = sb + 3 * bb;
        o.b_d = sb + 3 * (size_t)(path_count * cap_b) + 3 * bb;
        o.b_thr = sb + 6 * (size_t)(path_count * cap_b) + 3 * bb;
        o.b_len = sb + 9 * (size_t)(path_count * cap_b) + bb;
        o.b_flags = sf + bb;
        o.b_med = smed + bb;
        o.b_mis = smis + bb;
        o.margin = margins ? &margins[p] : NULL;
        if (margins) margins[p] = INFINITY;
        lt_trace_path(m, regions, n_regions, lights, n_lights, P, p, &o);
        nv[p] = o.nv;
        nb[p] = o.nb;
    }
    int64_t tv = 0, tb = 0;
    for (int64_t p = 0; p < path_count; ++p) {
        tv += nv[p];
        tb += nb[p];
    }
    out->vertex_count = tv;
    out->beam_count = tb;
    int r = 0;
    if (tv > out->vertex_capacity || tb > out->beam_capacity) {
        r = UPBP_EINVAL;
    } else {
        int64_t ov = 0, ob = 0;
        for (int64_t p = 0; p < path_count; ++p) {
            const int64_t bv = p * cap_v, bb = p * cap_b;
            memcpy(out->vertices + ov, sv + bv, sizeof(upbp_light_vertex) * (size_t)nv[p]);
            memcpy(out->positions + 3 * ov, sp + 3 * bv, sizeof(float) * 3 * (size_t)nv[p]);
            for (int32_t k = 0; k < nb[p]; ++k) {
                const int64_t s = bb + k, d = ob + k;
                for (int a = 0; a < 3; ++a) {
                    out->beam_origin[3 * d + a] = sb[3 * s + a];
                    out->beam_direction[3 * d + a] = sb[3 * (size_t)(path_count * cap_b) + 3 * s + a];
                    out->beam_throughput[3 * d + a] = sb[6 * (size_t)(path_count * cap_b) + 3 * s + a];
                }
                out->beam_length[d] = sb[9 * (size_t)(path_count * cap_b) + s];
                out->beam_flags[d] = sf[s];
                out->beam_medium[d] = smed[s];
                out->beam_mis[d] = smis[s];
                out->beam_mis[d].light_vertex += (int32_t)ov;
            }
            ov += nv[p];
            ob += nb[p];
            out->path_ends[p] = (int32_t)ov;
        }
    }
    free(m);
    free(sv);
    free(sp);
    free(sb);
    free(sf);
    free(smed);
    free(smis);
    free(nv);
    free(nb);
    return r;
}
