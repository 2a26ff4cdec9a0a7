/*
 * upbp_oracle.c -- CPU restatement of SmallUPBP's PB2D (BRE) and BB1D (photon
 * beam) estimators. TEST INFRASTRUCTURE: the checker of the GPU library and the
 * timed CPU baseline; never part of the product. See upbp_oracle.h.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp). The FP32 path
 * must not be compiled with FMA contraction or -ffast-math: it reproduces the
 * reference's SSE operation order (SmallUPBP/SmallUPBP.vcxproj:160-162 builds
 * with /fp:precise and SSE2).
 */
#include "upbp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* SmallUPBP/src/Misc/Defs.hxx:74,77 */
#define PI_F 3.141592653589793238462643383279f
#define INV_PI_F (1.f / PI_F)

/* ------------------------------------------------------------------------- */
/* Media, SmallUPBP/src/Scene/Media.hxx:129-159 */
void or_prepare_media(const upbp_medium* in, int32_t n, or_medium* out) {
    for (int32_t i = 0; i < n; ++i) {
        or_medium* m = &out[i];
        memset(m, 0, sizeof(*m));
        for (int c = 0; c < 3; ++c) {
            m->sigma_a[c] = fabsf(in[i].absorption[c]);
            m->sigma_s[c] = fabsf(in[i].scattering[c]);
            m->sigma_t[c] = m->sigma_a[c] + m->sigma_s[c];
        }
        m->mean_cosine = in[i].mean_cosine;
        m->continuation_prob = in[i].continuation_prob;
        /* min positive component, first index on ties, zero components skipped (Media.hxx:142-145) */
        m->min_comp = m->sigma_t[0];
        m->min_idx = 0;
        if (m->sigma_t[1] > 0 && (m->min_comp == 0 || m->sigma_t[1] < m->min_comp)) {
            m->min_comp = m->sigma_t[1];
            m->min_idx = 1;
        }
        if (m->sigma_t[2] > 0 && (m->min_comp == 0 || m->sigma_t[2] < m->min_comp)) {
            m->min_comp = m->sigma_t[2];
            m->min_idx = 2;
        }
        m->has_scattering = (m->sigma_s[0] > 0 || m->sigma_s[1] > 0 || m->sigma_s[2] > 0);
        m->has_attenuation = (m->sigma_t[0] > 0 || m->sigma_t[1] > 0 || m->sigma_t[2] > 0);
        float mx = fmaxf(m->sigma_t[0], fmaxf(m->sigma_t[1], m->sigma_t[2]));
        m->max_beam_length = m->has_attenuation ? (float)(-log(1e-9) / mx) : INFINITY;
        m->mfp = m->has_attenuation ? 1.0f / mx : INFINITY;
    }
}

/* ------------------------------------------------------------------------- */
/* Single-precision exp by the Cephes method (S. Moshier, Cephes Math Library
 * expf), as vectorised by J. Pommier's sse_mathfun exp_ps, which the reference
 * uses for RGB exponentials (Rgb.hxx:334-336 -> FastFloat.h:85-87 ->
 * sse_mathfun.h:250-300): clamp to +-88.3762626647949, split
 * x = n*ln2 + r with a two-part ln2 (0.693359375, -2.12194440e-4), a degree-5
 * minimax polynomial on r, then scale by 2^n. */
float or_exp_cephes(float x) {
    const float hi = 88.3762626647949f, lo = -88.3762626647949f;
    if (x > hi) x = hi;
    if (x < lo) x = lo;
    float fx = x * 1.44269504088896341f;
    fx = fx + 0.5f;
    float fl = floorf(fx);
    float tmp = fl * 0.693359375f;
    float z = fl * -2.12194440e-4f;
    x = x - tmp;
    x = x - z;
    z = x * x;
    float y = 1.9875691500E-4f;
    y = y * x;
    y = y + 1.3981999507E-3f;
    y = y * x;
    y = y + 8.3334519073E-3f;
    y = y * x;
    y = y + 4.1665795894E-2f;
    y = y * x;
    y = y + 1.6666665459E-1f;
    y = y * x;
    y = y + 5.0000001201E-1f;
    y = y * z;
    y = y + x;
    y = y + 1.0f;
    int n = (int)fl + 127;
    uint32_t bits = (uint32_t)n << 23;
    float p2;
    memcpy(&p2, &bits, 4);
    return y * p2;
}

uint32_t or_hash32(uint32_t i) {
    /* "lowbias32" integer hash (C. Wellons); i + 1 so that index 0 hashes non-trivially */
    uint32_t x = i + 1u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* ------------------------------------------------------------------------- */
/* Acceleration structures (culling only). */
typedef struct or_bvh_node {
    float lo[3], hi[3];
    int32_t first; /* leaf: first item; inner: index of the left child (right = first + 1) */
    int32_t count; /* leaf: item count > 0; inner: 0 */
} or_bvh_node;

struct or_photons {
    int64_t n;    /* kept photons */
    int64_t n_in; /* photons given to create (before the in_medium filter) */
    float* pos;
    float* r2;
    float* inc;
    float* power;
    int64_t* orig;
    or_bvh_node* nodes;
    int64_t n_nodes;
};

typedef struct or_grid {
    float lo[3], hi[3];
    float cell[3], inv_cell[3];
    int res[3];
    int64_t shift[3];
    uint32_t* cells;
    int64_t n_cells;
    uint32_t* ptrs;
    int64_t n_ptrs;
} or_grid;

struct or_beams {
    int64_t n;
    float* o;
    float* d;
    float* len;
    float* flux;
    uint32_t* flags;
    int32_t* medium;
    float* r0;
    float* r1;
    float* dr;
    float* r2max;
    float* margins;
    or_grid grid;
    /* per-cell beam reduction (Grid::reduceBeams, Grid.hxx:558-631); red_max 0 = none */
    uint32_t red_max, red_type;
    uint64_t red_seed;
};

/* ---- per-cell beam reduction -------------------------------------------------
 * The reference draws its selections from a sequential Rng; here every random number is a
 * counter-based hash of (seed, query ray, cell, draw), so a selection does not depend on
 * the order queries run in. Permutation of a crowded cell (PRESAMPLE / OFFSET): its entries
 * sorted by a hashed key of (seed, cell, beam), ties by beam index. */
static inline uint64_t red_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint32_t red_perm_key(uint64_t seed, uint32_t c, uint32_t b) {
    return (uint32_t)(red_mix(seed ^ red_mix(((uint64_t)c << 32) | b)) >> 32);
}
/* uniform in (0, 1): 24 hashed bits, centred */
static inline float red_uniform(uint64_t seed, uint64_t qkey, uint32_t c, uint32_t t) {
    uint64_t h = red_mix(seed + red_mix(qkey ^ red_mix(((uint64_t)c << 32) | t)));
    return ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}
/* (uint)(r * n), Grid.hxx:484,520 */
static inline uint32_t red_index(float u, uint32_t n) {
    uint32_t i = (uint32_t)(u * (float)n);
    return i < n ? i : n - 1;
}
static uint64_t red_query_key(const float* o, const float* d) {
    uint32_t bits[6];
    memcpy(bits, o, 12);
    memcpy(bits + 3, d, 12);
    uint64_t k = 0x243F6A8885A308D3ull;
    for (int i = 0; i < 6; ++i) k = red_mix(k ^ bits[i]);
    return k;
}
/* Grid::pdf's cell of a position: posToVoxel clamp and index (Grid.hxx:71-116,409-415) */
static uint32_t red_cell(const or_grid* g, const float* p) {
    int v[3];
    for (int a = 0; a < 3; ++a) {
        int i = (int)((p[a] - g->lo[a]) * g->inv_cell[a]);
        v[a] = i < 0 ? 0 : (i > g->res[a] - 1 ? g->res[a] - 1 : i);
    }
    return (uint32_t)(v[0] + g->res[0] * (v[1] + g->res[1] * v[2]));
}
/* place of beam b in the cell's permutation */
static uint32_t red_rank(const or_beams* bm, uint32_t c, uint32_t begin, uint32_t n, uint32_t b) {
    uint32_t kb = red_perm_key(bm->red_seed, c, b), r = 0;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t e = bm->grid.ptrs[begin + k];
        uint32_t ke = red_perm_key(bm->red_seed, c, e);
        if (ke < kb || (ke == kb && e < b)) ++r;
    }
    return r;
}
/* How many times the reduction tests the entry at `place` of crowded cell c's list (the
 * reference's registration order) for the query: intersectPresampled / intersectOffsetted /
 * intersectFixedSampled / intersectSampled, Grid.hxx:440-548. */
static uint32_t red_mult(const or_beams* bm, uint64_t qkey, uint32_t c, uint32_t begin, uint32_t n, uint32_t place,
                         float pdf) {
    const uint32_t mx = bm->red_max, b = bm->grid.ptrs[begin + place];
    if (bm->red_type == UPBP_REDUCE_PRESAMPLE) return red_rank(bm, c, begin, n, b) < mx;
    if (bm->red_type == UPBP_REDUCE_OFFSET) {
        uint32_t o = red_index(red_uniform(bm->red_seed, qkey, c, 0), n);
        return ((red_rank(bm, c, begin, n, b) + n - o) % n) < mx;
    }
    if (bm->red_type == UPBP_REDUCE_RESAMPLE_FIXED) {
        uint32_t m = 0;
        for (uint32_t t = 0; t < mx; ++t) m += red_index(red_uniform(bm->red_seed, qkey, c, t), n) == place;
        return m;
    }
    return red_uniform(bm->red_seed, qkey, c, b) < pdf;
}
/* The selection pdf of the cell holding p (Grid::pdf, Grid.hxx:409-415): max / n in a cell
 * of n > max beams, 1 otherwise and without reduction. */
static float red_cell_pdf(const or_beams* bm, const float* p) {
    if (!bm->red_max || !bm->grid.cells) return 1.0f;
    uint32_t c = red_cell(&bm->grid, p), n = bm->grid.cells[c + 1] - bm->grid.cells[c];
    return n <= bm->red_max ? 1.0f : (float)bm->red_max / (float)n;
}

/* A hit of beam b whose query point is p belongs to p's cell: times counted, pdf of the cell. */
static uint32_t red_decide_at(const or_beams* bm, uint64_t qkey, const float* p, uint32_t b, float* pdf) {
    const or_grid* g = &bm->grid;
    uint32_t c = red_cell(g, p), begin = g->cells[c], n = g->cells[c + 1] - begin;
    *pdf = 1.0f;
    if (n <= bm->red_max) return 1;
    *pdf = (float)bm->red_max / (float)n;
    for (uint32_t k = 0; k < n; ++k)
        if (g->ptrs[begin + k] == b) return red_mult(bm, qkey, c, begin, n, k, *pdf);
    return 0; /* never tested there */
}

/* Slab test of a ray parameter interval against a box (culling only). */
static inline int or_slab(const float* lo, const float* hi, const float* o, const float* inv,
                          float tn, float tf) {
    for (int a = 0; a < 3; ++a) {
        float t0 = (lo[a] - o[a]) * inv[a];
        float t1 = (hi[a] - o[a]) * inv[a];
        if (t0 != t0) t0 = -INFINITY;
        if (t1 != t1) t1 = INFINITY;
        float a0 = t0 < t1 ? t0 : t1, a1 = t0 < t1 ? t1 : t0;
        if (a0 > tn) tn = a0;
        if (a1 < tf) tf = a1;
    }
    return tn <= tf;
}

/* BoundingBox3::intersect, SmallUPBP/src/Structs/BoundingBox.hxx:218-271
 * (flipcode SSE ray-box test, NaN-filtered by min/max against +-inf). */
static int or_sse_box_intersect(const float* lo, const float* hi, const float* o, const float* inv,
                                float* tmin, float* tmax) {
    float lmax = INFINITY, lmin = -INFINITY;
    float mx[3], mn[3];
    for (int a = 0; a < 3; ++a) {
        float l1 = (lo[a] - o[a]) * inv[a];
        float l2 = (hi[a] - o[a]) * inv[a];
        /* _mm_min_ps(l, +inf) maps NaN to +inf; _mm_max_ps(l, -inf) maps NaN to -inf */
        float l1a = (l1 < INFINITY) ? l1 : INFINITY, l2a = (l2 < INFINITY) ? l2 : INFINITY;
        float l1b = (l1 > -INFINITY) ? l1 : -INFINITY, l2b = (l2 > -INFINITY) ? l2 : -INFINITY;
        mx[a] = (l1a > l2a) ? l1a : l2a;
        mn[a] = (l1b < l2b) ? l1b : l2b;
    }
    lmax = mx[0];
    lmin = mn[0];
    if (mx[1] < lmax) lmax = mx[1];
    if (mn[1] > lmin) lmin = mn[1];
    if (mx[2] < lmax) lmax = mx[2];
    if (mn[2] > lmin) lmin = mn[2];
    if (lmax >= lmin) {
        *tmin = lmin;
        *tmax = lmax;
        return 1;
    }
    return 0;
}

/* ---- binned-SAH BVH over photon cubes ----------------------------------- */
typedef struct bvh_build {
    const float* blo; /* [n][3] item boxes */
    const float* bhi;
    const float* cen; /* [n][3] centroids */
    int32_t* idx;
    or_bvh_node* nodes;
    int64_t n_nodes, cap;
    int leaf_size;
} bvh_build;

static int64_t bvh_alloc(bvh_build* b, int k) {
    if (b->n_nodes + k > b->cap) {
        b->cap = (b->cap * 2 > b->n_nodes + k) ? b->cap * 2 : b->n_nodes + k;
        b->nodes = (or_bvh_node*)realloc(b->nodes, (size_t)b->cap * sizeof(or_bvh_node));
    }
    int64_t r = b->n_nodes;
    b->n_nodes += k;
    return r;
}

static float box_area(const float* lo, const float* hi) {
    float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
    if (e0 < 0 || e1 < 0 || e2 < 0) return 0;
    return e0 * e1 + e1 * e2 + e2 * e0;
}

static void bvh_recurse(bvh_build* b, int64_t node, int32_t first, int32_t count, float eps) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t k = first; k < first + count; ++k) {
        int32_t i = b->idx[k];
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], b->blo[3 * i + a]);
            hi[a] = fmaxf(hi[a], b->bhi[3 * i + a]);
            clo[a] = fminf(clo[a], b->cen[3 * i + a]);
            chi[a] = fmaxf(chi[a], b->cen[3 * i + a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        b->nodes[node].lo[a] = lo[a] - eps;
        b->nodes[node].hi[a] = hi[a] + eps;
    }
    if (count <= b->leaf_size) {
        b->nodes[node].first = first;
        b->nodes[node].count = count;
        return;
    }
    /* 16-bin SAH on the widest centroid axis; median split when degenerate */
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > ext) {
            ext = chi[a] - clo[a];
            axis = a;
        }
    int32_t mid = first + count / 2;
    if (ext > 0) {
        enum { NB = 16 };
        int32_t cnt[NB] = {0};
        float blo[NB][3], bhi[NB][3];
        for (int k = 0; k < NB; ++k)
            for (int a = 0; a < 3; ++a) {
                blo[k][a] = INFINITY;
                bhi[k][a] = -INFINITY;
            }
        float scale = NB / ext;
        for (int32_t k = first; k < first + count; ++k) {
            int32_t i = b->idx[k];
            int bin = (int)((b->cen[3 * i + axis] - clo[axis]) * scale);
            if (bin >= NB) bin = NB - 1;
            if (bin < 0) bin = 0;
            cnt[bin]++;
            for (int a = 0; a < 3; ++a) {
                blo[bin][a] = fminf(blo[bin][a], b->blo[3 * i + a]);
                bhi[bin][a] = fmaxf(bhi[bin][a], b->bhi[3 * i + a]);
            }
        }
        float best = INFINITY;
        int best_split = -1;
        for (int s = 1; s < NB; ++s) {
            float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            int32_t nl = 0, nr = 0;
            for (int k = 0; k < s; ++k) {
                nl += cnt[k];
                for (int a = 0; a < 3; ++a) {
                    llo[a] = fminf(llo[a], blo[k][a]);
                    lhi[a] = fmaxf(lhi[a], bhi[k][a]);
                }
            }
            for (int k = s; k < NB; ++k) {
                nr += cnt[k];
                for (int a = 0; a < 3; ++a) {
                    rlo[a] = fminf(rlo[a], blo[k][a]);
                    rhi[a] = fmaxf(rhi[a], bhi[k][a]);
                }
            }
            if (nl == 0 || nr == 0) continue;
            float cost = box_area(llo, lhi) * nl + box_area(rlo, rhi) * nr;
            if (cost < best) {
                best = cost;
                best_split = s;
            }
        }
        if (best_split > 0) {
            int32_t i0 = first, i1 = first + count - 1;
            while (i0 <= i1) {
                int32_t i = b->idx[i0];
                int bin = (int)((b->cen[3 * i + axis] - clo[axis]) * scale);
                if (bin >= NB) bin = NB - 1;
                if (bin < 0) bin = 0;
                if (bin < best_split) {
                    i0++;
                } else {
                    int32_t t = b->idx[i0];
                    b->idx[i0] = b->idx[i1];
                    b->idx[i1] = t;
                    i1--;
                }
            }
            mid = i0;
        }
    }
    if (mid <= first || mid >= first + count) mid = first + count / 2;
    int64_t c = bvh_alloc(b, 2);
    b->nodes[node].first = (int32_t)c;
    b->nodes[node].count = 0;
    bvh_recurse(b, c, first, mid - first, eps);
    bvh_recurse(b, c + 1, mid, first + count - mid, eps);
}

static void photons_build_bvh(or_photons* ph, int leaf_size) {
    int64_t n = ph->n;
    if (n <= 0) return;
    float* blo = (float*)malloc((size_t)n * 12);
    float* bhi = (float*)malloc((size_t)n * 12);
    float* cen = (float*)malloc((size_t)n * 12);
    float maxc = 0;
    for (int64_t i = 0; i < n; ++i) {
        float r = sqrtf(ph->r2[i]);
        for (int a = 0; a < 3; ++a) {
            float p = ph->pos[3 * i + a];
            blo[3 * i + a] = p - r;
            bhi[3 * i + a] = p + r;
            cen[3 * i + a] = p;
            if (fabsf(p) + r > maxc) maxc = fabsf(p) + r;
        }
    }
    bvh_build b;
    memset(&b, 0, sizeof(b));
    b.blo = blo;
    b.bhi = bhi;
    b.cen = cen;
    b.leaf_size = leaf_size > 0 ? leaf_size : 4;
    b.idx = (int32_t*)malloc((size_t)n * 4);
    for (int64_t i = 0; i < n; ++i) b.idx[i] = (int32_t)i;
    bvh_alloc(&b, 1);
    bvh_recurse(&b, 0, 0, (int32_t)n, 1e-5f * (maxc > 1 ? maxc : 1));
    /* reorder photon arrays into leaf order so leaves are contiguous */
    float* np = (float*)malloc((size_t)n * 12);
    float* nr = (float*)malloc((size_t)n * 4);
    float* ni = (float*)malloc((size_t)n * 12);
    float* nw = (float*)malloc((size_t)n * 12);
    int64_t* no = (int64_t*)malloc((size_t)n * 8);
    for (int64_t k = 0; k < n; ++k) {
        int32_t i = b.idx[k];
        memcpy(&np[3 * k], &ph->pos[3 * i], 12);
        nr[k] = ph->r2[i];
        memcpy(&ni[3 * k], &ph->inc[3 * i], 12);
        memcpy(&nw[3 * k], &ph->power[3 * i], 12);
        no[k] = ph->orig[i];
    }
    free(ph->pos);
    free(ph->r2);
    free(ph->inc);
    free(ph->power);
    free(ph->orig);
    ph->pos = np;
    ph->r2 = nr;
    ph->inc = ni;
    ph->power = nw;
    ph->orig = no;
    ph->nodes = b.nodes;
    ph->n_nodes = b.n_nodes;
    free(b.idx);
    free(blo);
    free(bhi);
    free(cen);
}

or_photons* or_photons_create(int64_t n, const float* pos, const float* inc_dir, const float* power,
                              const float* radius, float const_radius, const uint8_t* in_medium,
                              int32_t leaf_size) {
    or_photons* ph = (or_photons*)calloc(1, sizeof(or_photons));
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
        if (!in_medium || in_medium[i]) m++;
    ph->n = m;
    ph->n_in = n;
    ph->pos = (float*)malloc((size_t)(m ? m : 1) * 12);
    ph->r2 = (float*)malloc((size_t)(m ? m : 1) * 4);
    ph->inc = (float*)malloc((size_t)(m ? m : 1) * 12);
    ph->power = (float*)malloc((size_t)(m ? m : 1) * 12);
    ph->orig = (int64_t*)malloc((size_t)(m ? m : 1) * 8);
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (in_medium && !in_medium[i]) continue;
        float r = radius ? radius[i] : const_radius;
        memcpy(&ph->pos[3 * k], &pos[3 * i], 12);
        ph->r2[k] = r * r; /* radiusSqr = aRadius*aRadius, Bre.cxx:148 */
        if (inc_dir)
            memcpy(&ph->inc[3 * k], &inc_dir[3 * i], 12);
        else
            ph->inc[3 * k] = ph->inc[3 * k + 1] = ph->inc[3 * k + 2] = 0.0f;
        memcpy(&ph->power[3 * k], &power[3 * i], 12);
        ph->orig[k] = i;
        k++;
    }
    if (leaf_size >= 0) photons_build_bvh(ph, leaf_size);
    return ph;
}

void or_photons_free(or_photons* ph) {
    if (!ph) return;
    free(ph->pos);
    free(ph->r2);
    free(ph->inc);
    free(ph->power);
    free(ph->orig);
    free(ph->nodes);
    free(ph);
}

int64_t or_photons_count(const or_photons* ph) { return ph->n; }

void or_photons_orig_index(const or_photons* ph, int64_t* out) {
    for (int64_t i = 0; i < ph->n; ++i) out[i] = ph->orig[i];
}

/* ---- beams: radii and the reference grid -------------------------------- */
/* PhotonBeam::getSegmentAABB, SmallUPBP/src/Beams/PhotonBeam.hxx:186-198 */
static void beam_segment_aabb(const or_beams* bm, int64_t i, float s0, float s1, float* lo, float* hi) {
    s0 *= bm->len[i];
    s1 *= bm->len[i];
    float rs = bm->r0[i] + s0 * bm->dr[i];
    float re = bm->r0[i] + s1 * bm->dr[i];
    for (int a = 0; a < 3; ++a) {
        float bs = bm->o[3 * i + a] + bm->d[3 * i + a] * s0;
        float be = bm->o[3 * i + a] + bm->d[3 * i + a] * s1;
        float ms = rs * bm->margins[3 * i + a], me = re * bm->margins[3 * i + a];
        float l0 = bs - ms, l1 = be - me, h0 = bs + ms, h1 = be + me;
        lo[a] = l0 < l1 ? l0 : l1;
        hi[a] = h0 > h1 ? h0 : h1;
    }
}

/* Grid::build + buildloop, SmallUPBP/src/Beams/Grid.hxx:167-279, and
 * AccelStruct::build (PhGrid.hxx:75-87) for the grid AABB. */
static void beams_build_grid(or_beams* bm, int res) {
    or_grid* g = &bm->grid;
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = INFINITY;
        g->hi[a] = -INFINITY;
    }
    for (int64_t i = 0; i < bm->n; ++i) {
        float lo[3], hi[3];
        beam_segment_aabb(bm, i, 0.0f, 1.0f, lo, hi);
        for (int a = 0; a < 3; ++a) {
            g->lo[a] = fminf(g->lo[a], lo[a]);
            g->hi[a] = fmaxf(g->hi[a], hi[a]);
        }
    }
    float ext[3] = {g->hi[0] - g->lo[0], g->hi[1] - g->lo[1], g->hi[2] - g->lo[2]};
    float emax = ext[0] > ext[1] ? ext[0] : ext[1];
    if (ext[2] > emax) emax = ext[2];
    float size = emax / (float)res;
    for (int a = 0; a < 3; ++a) g->res[a] = (int)(unsigned)(ext[a] / size) + 1;
    g->shift[0] = 1;
    g->shift[1] = g->res[0];
    g->shift[2] = (int64_t)g->res[0] * g->res[1];
    for (int a = 0; a < 3; ++a) {
        g->cell[a] = ext[a] / (float)g->res[a];
        g->inv_cell[a] = 1.0f / g->cell[a];
    }
    g->n_cells = (int64_t)(g->res[0] + 1) * (g->res[1] + 1) * (g->res[2] + 1) + 1;
    g->cells = (uint32_t*)calloc((size_t)g->n_cells, 4);
    uint32_t* dup = (uint32_t*)malloc((size_t)g->n_cells * 4);
    for (int64_t c = 0; c < g->n_cells; ++c) dup[c] = (uint32_t)bm->n;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            /* inclusive prefix sum (Grid.hxx:256-264), then Store decrements back */
            uint64_t acc = 0;
            for (int64_t c = 0; c < g->n_cells; ++c) {
                acc += g->cells[c];
                g->cells[c] = (uint32_t)acc;
                dup[c] = (uint32_t)bm->n;
            }
            g->n_ptrs = (int64_t)acc;
            g->ptrs = (uint32_t*)malloc((size_t)(acc ? acc : 1) * 4);
        }
        for (int64_t i = 0; i < bm->n; ++i) {
            float blo[3], bhi[3];
            beam_segment_aabb(bm, i, 0.0f, 1.0f, blo, bhi);
            float ext_b[3] = {bhi[0] - blo[0], bhi[1] - blo[1], bhi[2] - blo[2]};
            float ad[3] = {fabsf(bm->d[3 * i]), fabsf(bm->d[3 * i + 1]), fabsf(bm->d[3 * i + 2])};
            int max_axis = (ad[0] > ad[1]) ? ((ad[0] > ad[2]) ? 0 : 2) : ((ad[1] > ad[2]) ? 1 : 2);
            int chops = (int)(ext_b[max_axis] * g->inv_cell[max_axis]) + 1;
            float inv_chops = 1.0f / (float)chops;
            for (int ch = 0; ch < chops; ++ch) {
                float lo[3], hi[3];
                beam_segment_aabb(bm, i, (float)ch * inv_chops, (float)(ch + 1) * inv_chops, lo, hi);
                int is[3], ie[3];
                for (int a = 0; a < 3; ++a) {
                    is[a] = (int)((lo[a] - g->lo[a]) * g->inv_cell[a]);
                    ie[a] = (int)((hi[a] - g->lo[a]) * g->inv_cell[a]);
                }
                int64_t id = (int64_t)is[0] + (int64_t)g->res[0] * ((int64_t)is[1] + (int64_t)g->res[1] * is[2]);
                for (int z = is[2]; z <= ie[2]; ++z, id += g->shift[2]) {
                    int64_t idy = id;
                    for (int y = is[1]; y <= ie[1]; ++y, idy += g->shift[1]) {
                        int64_t idx = idy;
                        for (int x = is[0]; x <= ie[0]; ++x, ++idx) {
                            if (idx < 0) continue;
                            if (idx >= g->n_cells) break;
                            if (dup[idx] != (uint32_t)i) {
                                dup[idx] = (uint32_t)i;
                                if (pass == 0)
                                    g->cells[idx]++;
                                else
                                    g->ptrs[--g->cells[idx]] = (uint32_t)i;
                            }
                        }
                    }
                }
            }
        }
    }
    free(dup);
}

or_beams* or_beams_create(int64_t n, const float* origin, const float* dir, const float* length,
                          const float* throughput, const uint32_t* flags, const int32_t* medium,
                          const float* start_radius, const float* end_radius, float const_radius,
                          int32_t grid_res) {
    or_beams* bm = (or_beams*)calloc(1, sizeof(or_beams));
    int64_t m = n > 0 ? n : 1;
    bm->n = n;
    bm->o = (float*)malloc((size_t)m * 12);
    bm->d = (float*)malloc((size_t)m * 12);
    bm->len = (float*)malloc((size_t)m * 4);
    bm->flux = (float*)malloc((size_t)m * 12);
    bm->flags = (uint32_t*)malloc((size_t)m * 4);
    bm->medium = (int32_t*)malloc((size_t)m * 4);
    bm->r0 = (float*)malloc((size_t)m * 4);
    bm->r1 = (float*)malloc((size_t)m * 4);
    bm->dr = (float*)malloc((size_t)m * 4);
    bm->r2max = (float*)malloc((size_t)m * 4);
    bm->margins = (float*)malloc((size_t)m * 12);
    for (int64_t i = 0; i < n; ++i) {
        memcpy(&bm->o[3 * i], &origin[3 * i], 12);
        memcpy(&bm->d[3 * i], &dir[3 * i], 12);
        bm->len[i] = length[i];
        memcpy(&bm->flux[3 * i], &throughput[3 * i], 12);
        bm->flags[i] = flags[i];
        bm->medium[i] = medium[i];
        /* PhBeams.cxx:92-117 */
        if (start_radius) {
            float r0 = start_radius[i];
            float r1 = end_radius ? end_radius[i] : r0;
            float mx = r0 > r1 ? r0 : r1;
            bm->r0[i] = r0;
            bm->r1[i] = r1;
            bm->r2max[i] = mx * mx;
            bm->dr[i] = (r1 - r0) / length[i];
        } else {
            bm->r0[i] = bm->r1[i] = const_radius;
            bm->r2max[i] = const_radius * const_radius;
            bm->dr[i] = 0.0f;
        }
        /* margins of the cone AABB, PhotonBeam.hxx:74-80 */
        float dx = bm->d[3 * i] * bm->d[3 * i], dy = bm->d[3 * i + 1] * bm->d[3 * i + 1],
              dz = bm->d[3 * i + 2] * bm->d[3 * i + 2];
        bm->margins[3 * i] = sqrtf(dy + dz);
        bm->margins[3 * i + 1] = sqrtf(dx + dz);
        bm->margins[3 * i + 2] = sqrtf(dx + dy);
    }
    if (grid_res > 0 && n > 0) beams_build_grid(bm, grid_res);
    return bm;
}

int or_beams_set_reduction(or_beams* bm, uint32_t max_beams_in_cell, uint32_t reduction_type, int32_t seed) {
    if (max_beams_in_cell && (!bm->grid.cells || reduction_type > UPBP_REDUCE_RESAMPLE)) return UPBP_EINVAL;
    bm->red_max = max_beams_in_cell;
    bm->red_type = reduction_type;
    bm->red_seed = (uint64_t)(uint32_t)seed;
    return 0;
}

float or_beams_selection_pdf(const or_beams* bm, const float* pos) { return red_cell_pdf(bm, pos); }

void or_beams_free(or_beams* bm) {
    if (!bm) return;
    free(bm->o);
    free(bm->d);
    free(bm->len);
    free(bm->flux);
    free(bm->flags);
    free(bm->medium);
    free(bm->r0);
    free(bm->r1);
    free(bm->dr);
    free(bm->r2max);
    free(bm->margins);
    free(bm->grid.cells);
    free(bm->grid.ptrs);
    free(bm);
}

int64_t or_beams_count(const or_beams* bm) { return bm->n; }

void or_beams_grid_info(const or_beams* bm, int64_t* info6) {
    info6[0] = bm->grid.res[0];
    info6[1] = bm->grid.res[1];
    info6[2] = bm->grid.res[2];
    info6[3] = bm->grid.n_cells;
    info6[4] = bm->grid.n_ptrs;
    info6[5] = bm->n;
}

void or_beams_grid_geometry(const or_beams* bm, float* out9) {
    for (int a = 0; a < 3; ++a) {
        out9[a] = bm->grid.lo[a];
        out9[3 + a] = bm->grid.hi[a];
        out9[6 + a] = bm->grid.cell[a];
    }
}

int or_beams_grid_cell_counts(const or_beams* bm, uint32_t* out, int64_t n) {
    const or_grid* g = &bm->grid;
    if (!g->cells || n != (int64_t)g->res[0] * g->res[1] * g->res[2]) return UPBP_EINVAL;
    for (int64_t c = 0; c < n; ++c) out[c] = g->cells[c + 1] - g->cells[c];
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Vertex-merging hashed grid (SmallUPBP/src/Misc/HashGrid.hxx:38-230), FP32 throughout:
 * the grid decides which vertices a camera vertex visits, so it is restated in the
 * reference's operation order. */
struct or_hgrid {
    uint32_t type;          /* UPBP_PP3D (in-medium vertices) or UPBP_SURF (the others) */
    int64_t n_src;          /* light vertices given to the build */
    int64_t n;              /* vertices in the grid (MatchesType) */
    float radius, radius_sqr, cell_size, inv_cell_size;
    float bb_min[3], bb_max[3];
    int64_t cells;          /* mCellEnds.size() */
    int32_t* indices;       /* mIndices: light-vertex indices in bucket order */
    int32_t* cell_ends;     /* mCellEnds after the build: one past each bucket's last index */
    uint32_t* bucket;       /* bucket of each grid slot's vertex (brute-force check) */
    float* pos;             /* [n_src][3] */
};

/* GetCellIndex(x, y, z), HashGrid.hxx:195-206: unsigned 32-bit products, XOR, modulo. */
static inline int32_t pp3d_hash(const or_hgrid* g, int32_t x, int32_t y, int32_t z) {
    const uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    return (int32_t)(((ux * 73856093u) ^ (uy * 19349663u) ^ (uz * 83492791u)) % (uint32_t)g->cells);
}

/* GetCellIndex(Pos), HashGrid.hxx:208-218. */
static inline int32_t pp3d_cell_of(const or_hgrid* g, const float* p) {
    const float dx = p[0] - g->bb_min[0], dy = p[1] - g->bb_min[1], dz = p[2] - g->bb_min[2];
    return pp3d_hash(g, (int32_t)floorf(g->inv_cell_size * dx), (int32_t)floorf(g->inv_cell_size * dy),
                     (int32_t)floorf(g->inv_cell_size * dz));
}

/* UPBPLightVertex::MatchesType, UPBPLightVertex.hxx:74-77. */
static inline int grid_matches(const or_hgrid* g, const upbp_light_vertex* v) {
    const int in_medium = (v->flags & UPBP_LV_IN_MEDIUM) != 0;
    return g->type == UPBP_PP3D ? in_medium : !in_medium;
}

/* HashGrid::Build, HashGrid.hxx:46-128. */
or_hgrid* or_hgrid_create(int64_t n, const float* pos, const upbp_light_vertex* lv, float radius, int64_t cell_count,
                        uint32_t type) {
    if (n < 0 || cell_count < 1 || !(radius > 0.f) || (type != UPBP_PP3D && type != UPBP_SURF)) return NULL;
    or_hgrid* g = (or_hgrid*)calloc(1, sizeof(or_hgrid));
    g->type = type;
    g->n_src = n;
    g->radius = radius;
    g->radius_sqr = radius * radius;
    g->cell_size = radius * 2.f;
    g->inv_cell_size = 1.f / g->cell_size;
    g->cells = cell_count;
    for (int a = 0; a < 3; ++a) {
        g->bb_min[a] = 1e36f;
        g->bb_max[a] = -1e36f;
    }
    g->pos = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(g->pos, pos, sizeof(float) * 3 * (size_t)n);
    int64_t matched = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (!grid_matches(g, &lv[i])) continue;
        for (int a = 0; a < 3; ++a) {
            /* std::max(a, b) = (a < b) ? b : a; std::min(a, b) = (b < a) ? b : a */
            if (g->bb_max[a] < pos[3 * i + a]) g->bb_max[a] = pos[3 * i + a];
            if (pos[3 * i + a] < g->bb_min[a]) g->bb_min[a] = pos[3 * i + a];
        }
        matched++;
    }
    g->n = matched;
    g->indices = (int32_t*)malloc(sizeof(int32_t) * (size_t)(matched > 0 ? matched : 1));
    g->bucket = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(matched > 0 ? matched : 1));
    g->cell_ends = (int32_t*)calloc((size_t)cell_count, sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i)
        if (grid_matches(g, &lv[i])) g->cell_ends[pp3d_cell_of(g, pos + 3 * i)]++;
    int32_t sum = 0;
    for (int64_t c = 0; c < cell_count; ++c) {
        const int32_t t = g->cell_ends[c];
        g->cell_ends[c] = sum;
        sum += t;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (!grid_matches(g, &lv[i])) continue;
        const int32_t c = pp3d_cell_of(g, pos + 3 * i);
        const int32_t slot = g->cell_ends[c]++;
        g->indices[slot] = (int32_t)i;
        g->bucket[slot] = (uint32_t)c;
    }
    return g;
}

void or_hgrid_free(or_hgrid* g) {
    if (!g) return;
    free(g->indices);
    free(g->bucket);
    free(g->cell_ends);
    free(g->pos);
    free(g);
}

int64_t or_hgrid_count(const or_hgrid* g) { return g->n; }

void or_hgrid_info(const or_hgrid* g, float* bb6, int32_t* indices, int32_t* cell_ends) {
    if (bb6)
        for (int a = 0; a < 3; ++a) {
            bb6[a] = g->bb_min[a];
            bb6[3 + a] = g->bb_max[a];
        }
    if (indices && g->n > 0) memcpy(indices, g->indices, sizeof(int32_t) * (size_t)g->n);
    if (cell_ends) memcpy(cell_ends, g->cell_ends, sizeof(int32_t) * (size_t)g->cells);
}

/* The eight buckets a query visits, HashGrid::Process, HashGrid.hxx:130-186.
 * Returns 0 when the query lies outside the grid's bounding box. */
static int pp3d_buckets(const or_hgrid* g, const float* q, int32_t* h8) {
    float dmin[3], dmax[3];
    for (int a = 0; a < 3; ++a) {
        dmin[a] = q[a] - g->bb_min[a];
        dmax[a] = g->bb_max[a] - q[a];
    }
    for (int a = 0; a < 3; ++a) {
        if (dmin[a] < 0.f) return 0;
        if (dmax[a] < 0.f) return 0;
    }
    int32_t p[3], po[3];
    for (int a = 0; a < 3; ++a) {
        const float cp = g->inv_cell_size * dmin[a];
        const float cf = floorf(cp);
        p[a] = (int32_t)cf;
        const float fr = cp - cf;
        po[a] = p[a] + (fr < 0.5f ? -1 : +1);
    }
    for (int j = 0; j < 8; ++j) {
        /* case order of HashGrid.hxx:163-172: x varies slowest, z fastest */
        const int32_t x = (j & 4) ? po[0] : p[0], y = (j & 2) ? po[1] : p[1], z = (j & 1) ? po[2] : p[2];
        h8[j] = pp3d_hash(g, x, y, z);
    }
    return 1;
}

/* distSqr of HashGrid.hxx:178-179: (query - particle).square() = dot(d, d). */
static inline float pp3d_dist_sqr(const float* q, const float* p) {
    const float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    return (dx * dx + dy * dy) + dz * dz;
}

/* Calls visit(ctx, vertex) for every Process call of one query, in the grid's order
 * (accel = 1), or, as an independent statement of the same multiset (accel = 0), for
 * every grid vertex within the radius once per neighbour bucket equal to its own. */
typedef void (*pp3d_visit_fn)(void* ctx, int32_t vertex);
static int64_t pp3d_query(const or_hgrid* g, const float* q, int accel, pp3d_visit_fn visit, void* ctx) {
    int32_t h8[8];
    if (g->n == 0 || !pp3d_buckets(g, q, h8)) return 0;
    int64_t tested = 0;
    if (accel) {
        for (int j = 0; j < 8; ++j) {
            const int32_t c = h8[j];
            int32_t k = c == 0 ? 0 : g->cell_ends[c - 1];
            const int32_t e = g->cell_ends[c];
            for (; k < e; ++k) {
                const int32_t v = g->indices[k];
                tested++;
                if (pp3d_dist_sqr(q, g->pos + 3 * (int64_t)v) <= g->radius_sqr) visit(ctx, v);
            }
        }
    } else {
        for (int64_t k = 0; k < g->n; ++k) {
            const int32_t v = g->indices[k];
            tested++;
            if (!(pp3d_dist_sqr(q, g->pos + 3 * (int64_t)v) <= g->radius_sqr)) continue;
            for (int j = 0; j < 8; ++j)
                if ((uint32_t)h8[j] == g->bucket[k]) visit(ctx, v);
        }
    }
    return tested;
}

/* ------------------------------------------------------------------------- */
/* The estimator mathematics, in FP32 and FP64. */
#define REAL float
#define SUF(name) name##_f32
#define EXPRGB(x) or_exp_cephes(x)
#define EXP1(x) expf(x)
#define SQRT(x) sqrtf(x)
#define FABS(x) fabsf(x)
#define POW(x, y) powf(x, y)
#include "oracle_core.inc"
#undef REAL
#undef SUF
#undef EXPRGB
#undef EXP1
#undef SQRT
#undef FABS
#undef POW

#define REAL double
#define SUF(name) name##_f64
#define EXPRGB(x) exp(x)
#define EXP1(x) exp(x)
#define SQRT(x) sqrt(x)
#define FABS(x) fabs(x)
#define POW(x, y) pow(x, y)
#include "oracle_core.inc"
#undef REAL
#undef SUF
#undef EXPRGB
#undef EXP1
#undef SQRT
#undef FABS
#undef POW

float or_phase_pdf(const float fix[3], const float gen[3], float g, float* sin_theta) {
    if (sin_theta) {
        float c = -dot3_f32(gen, fix);
        *sin_theta = sqrtf(fmaxf(0.f, 1.f - c * c));
    }
    return phase_f32(fix, gen, g);
}

double or_camera_walk(const upbp_mis_data* cam, int len, double last_rev, double last_sin, double last_ray_rev_inv,
                      double last_ray_rev_ratio, double next_to_last_partial, uint32_t qbt, uint32_t pbt, uint32_t tech) {
    cam_view_f64 cv;
    cv.base = cam;
    cv.len = len;
    cv.last = mis_load_f64(&cam[len]);
    cv.prev_ray_rev_inv = len >= 1 ? cam[len - 1].ray_sample_rev_pdf_inv : 0.0;
    return camera_path_weight_f64(&cv, last_rev, last_sin, last_ray_rev_inv, last_ray_rev_ratio, next_to_last_partial,
                                  qbt, pbt, tech);
}

double or_light_walk(const upbp_light_vertex* lv, int64_t lv_count, const int32_t* path_ends, int64_t path_count,
                     int path_idx, int path_length, double last_rev, double next_to_last_partial, uint32_t evaluated,
                     uint32_t qbt, uint32_t pbt, uint32_t tech, int* err) {
    light_view_f64 l;
    l.lv = lv;
    l.lv_count = lv_count;
    l.path_ends = path_ends;
    l.path_count = path_count;
    l.beam_mis = NULL;
    l.override_idx = -1;
    l.override_ray_rev_inv = 0;
    l.err = err;
    return light_path_weight_f64(&l, path_idx, path_length, last_rev, next_to_last_partial, evaluated, qbt, pbt, tech);
}

int or_test_bre(const float o[3], const float d[3], float min_t, float max_t, const float p[3],
                float r2, float* t, float* s) {
    return test_bre_f32(o, d, min_t, max_t, p, r2, t, s);
}

int or_test_beam_beam(const float o1[3], const float d1[3], float min_t1, float max_t1,
                      const float o2[3], const float d2[3], float min_t2, float max_t2,
                      float max_dist_sqr, float* dist, float* sin_theta, float* t1, float* t2) {
    return test_beam_beam_f32(o1, d1, min_t1, max_t1, o2, d2, min_t2, max_t2, max_dist_sqr, dist,
                              sin_theta, t1, t2);
}

static int check_batch(const upbp_query_batch* q, int32_t n_media) {
    if (!q || q->count < 0) return UPBP_EINVAL;
    if (q->count > 0 && (!q->origin || !q->direction || !q->segment_offsets)) return UPBP_EINVAL;
    if (q->lite_segments && q->beam_type != UPBP_LONG_BEAM) return UPBP_EINVAL;
    if (n_media <= 0) return UPBP_EINVAL;
    /* the library's CSR rule: offsets non-decreasing inside [0, offsets[count]] */
    if (q->count > 0) {
        const int64_t end = q->segment_offsets[q->count];
        if (end < 0 || (end > 0 && !q->segments)) return UPBP_EINVAL;
        for (int64_t i = 0; i < q->count; ++i) {
            const int64_t s0 = q->segment_offsets[i], s1 = q->segment_offsets[i + 1];
            if (!(s0 >= 0 && s0 <= s1 && s1 <= end)) return UPBP_EINVAL;
        }
    }
    return 0;
}

int or_bre_eval(const or_photons* ph, const upbp_medium* media, int32_t n_media,
                const upbp_query_batch* q, int accel, int precision, int threads, float* out32,
                double* out64, uint32_t* sig, uint64_t* counters) {
    int e = check_batch(q, n_media);
    if (e) return e;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
    int r;
    if (precision == 64)
        r = bre_eval_all_f64(ph, m, n_media, q, accel, threads, out64, sig, counters);
    else
        r = bre_eval_all_f32(ph, m, n_media, q, accel, threads, out32, sig, counters);
    free(m);
    return r;
}

int or_beams_eval(const or_beams* bm, const upbp_medium* media, int32_t n_media,
                  const upbp_query_batch* q, int accel, int precision, int threads, float* out32,
                  double* out64, uint32_t* sig, uint64_t* counters) {
    int e = check_batch(q, n_media);
    if (e) return e;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
    int r;
    if (precision == 64)
        r = beams_eval_all_f64(bm, m, n_media, q, accel == 1 ? 0 : accel, threads, out64, sig, counters);
    else
        r = beams_eval_all_f32(bm, m, n_media, q, accel, threads, out32, sig, counters);
    free(m);
    return r;
}

/* EmbreeBre::evalBreBruteForce, Bre.cxx:980-1029, in float as the reference: every kept photon
 * against [min_t, max_t] with testIntersectionBre, each hit adding
 * ((flux exp(-t sigma_t)) PhaseFunction::Evaluate(d, incDir)) / r^2 (attenuation from the ray
 * origin, the stored incoming direction as given, no medium check of the photon), the sum times
 * sigma_s and INV_PI_F of the query's medium. */
int or_bre_eval_brute_force(const or_photons* ph, const upbp_medium* media, int32_t n_media, const float* origin,
                            const float* direction, const float* min_t, const float* max_t, const int32_t* medium,
                            int64_t count, int threads, float* out) {
    if (count < 0 || n_media <= 0 || (count > 0 && (!origin || !direction || !min_t || !max_t || !medium || !out)))
        return UPBP_EINVAL;
    for (int64_t q = 0; q < count; ++q)
        if (medium[q] < 0 || medium[q] >= n_media) return UPBP_EINVAL;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t q = 0; q < count; ++q) {
        const or_medium* md = &m[medium[q]];
        const float* o = &origin[3 * q];
        const float* d = &direction[3 * q];
        float acc[3] = {0.f, 0.f, 0.f};
        for (int64_t i = 0; i < ph->n; ++i) {
            float t, s;
            if (!test_bre_f32(o, d, min_t[q], max_t[q], &ph->pos[3 * i], ph->r2[i], &t, &s)) continue;
            const float rho = phase_f32(d, &ph->inc[3 * i], md->mean_cosine);
            for (int c = 0; c < 3; ++c) {
                const float att = or_exp_cephes(-t * md->sigma_t[c]);
                acc[c] += ((ph->power[3 * i + c] * att) * rho) / ph->r2[i];
            }
        }
        for (int c = 0; c < 3; ++c) out[3 * q + c] = (acc[c] * md->sigma_s[c]) * INV_PI_F;
    }
    free(m);
    return 0;
}

/* GridStats per query (GridStats.hxx:37-54): Grid::intersect's counts (Grid.hxx:365,378-382)
 * summed over the query's segments in a scattering medium, with the walk over segments of
 * PhotonBeamsEvaluator::evalBeamBeamEstimate (PhBeams.cxx:177-230): the float attenuation
 * product (divided by the ray-sample pdf for a SHORT query) ends the walk once a channel is
 * not above 1e-20 (Rgb::isPositive, Utils.hxx:109-111). */
int or_beam_grid_stats(const or_beams* bm, const upbp_medium* media, int32_t n_media, const upbp_query_batch* q,
                       int threads, uint32_t* out) {
    int e = check_batch(q, n_media);
    if (e) return e;
    if (!bm->grid.cells || !out) return UPBP_EINVAL;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
    const int short_q = q->beam_type == UPBP_SHORT_BEAM;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t qi = 0; qi < q->count; ++qi) {
        uint32_t gs[3] = {0, 0, 0};
        float att[3] = {1.0f, 1.0f, 1.0f};
        const float* o = &q->origin[3 * qi];
        const float* d = &q->direction[3 * qi];
        for (int64_t si = q->segment_offsets[qi]; si < q->segment_offsets[qi + 1]; ++si) {
            const upbp_volume_segment* sg = &q->segments[si];
            if (m[sg->medium_id].has_scattering)
                grid_walk_f32(bm, o, d, sg->dist_min, sg->dist_max, 0, NULL, NULL, gs);
            for (int c = 0; c < 3; ++c)
                att[c] = att[c] * (short_q ? sg->attenuation[c] / sg->ray_sample_pdf : sg->attenuation[c]);
            if (!(att[0] > 1e-20f && att[1] > 1e-20f && att[2] > 1e-20f)) break;
        }
        out[3 * qi] = gs[0];
        out[3 * qi + 1] = gs[1];
        out[3 * qi + 2] = gs[2];
    }
    free(m);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* FP64 analytic gradient of the plain estimators (LONG queries).
 * Forward (per query q, segment s of medium m, prefix P_s = prod_{j<s} A_j with
 * A_j = exp(-sigma_t[m_j] * (dist_max_j - dist_min_j))):
 *   photon hit:  C_c = P_c * flux_c * exp(-st_c * (t - tmin)) * ss_c * rho * K
 *   beam hit:    C_c = P_c * flux_c * ss_c * exp(-st_c * (t1 - tmin)) * Eb_c * rho * K1
 *                Eb_c = exp(-st_c t2), divided by exp(-st_k t2) for SHORT beams.
 * d/d sigma_s = direct + d/d sigma_t; d/d sigma_a = d/d sigma_t. */
typedef struct adj_acc {
    double* gs; /* [n_media][3] d/d sigma_s direct */
    double* gt; /* [n_media][3] d/d sigma_t */
} adj_acc;

/* Photon and beam hit decisions of the adjoint: FP64 (the FP64 forward's, for finite
 * differences of it) or FP32 (the reference's, hit sets identical to the FP32 forward and the GPU;
 * the hit's parameters are then the FP32 ones, widened to double). */
static int adj_photon_hit(int fp32, const double* o, const double* d, double tmin, double tmax, const double* p,
                          float r2, double* t, double* s) {
    if (!fp32) return test_bre_f64(o, d, tmin, tmax, p, (double)r2, t, s);
    const float of[3] = {(float)o[0], (float)o[1], (float)o[2]}, df[3] = {(float)d[0], (float)d[1], (float)d[2]};
    const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
    float tf, sf;
    if (!test_bre_f32(of, df, (float)tmin, (float)tmax, pf, r2, &tf, &sf)) return 0;
    *t = tf;
    *s = sf;
    return 1;
}

static int adj_beam_hit(int fp32, const double* o, const double* d, double tmin, double tmax, const or_beams* bm,
                        int64_t i, double* dist, double* st, double* t1, double* t2, double* radius) {
    if (!fp32) {
        const double o2[3] = {bm->o[3 * i], bm->o[3 * i + 1], bm->o[3 * i + 2]};
        const double d2[3] = {bm->d[3 * i], bm->d[3 * i + 1], bm->d[3 * i + 2]};
        if (!test_beam_beam_f64(o, d, tmin, tmax, o2, d2, 0.0, bm->len[i], bm->r2max[i], dist, st, t1, t2)) return 0;
        *radius = (double)bm->r0[i] + *t2 * (double)bm->dr[i];
        return !(*radius < *dist);
    }
    const float of[3] = {(float)o[0], (float)o[1], (float)o[2]}, df[3] = {(float)d[0], (float)d[1], (float)d[2]};
    float fd, fs, f1, f2;
    if (!test_beam_beam_f32(of, df, (float)tmin, (float)tmax, &bm->o[3 * i], &bm->d[3 * i], 0.f, bm->len[i],
                            bm->r2max[i], &fd, &fs, &f1, &f2))
        return 0;
    const float fr = bm->r0[i] + f2 * bm->dr[i];
    if (fr < fd) return 0;
    *dist = fd;
    *st = fs;
    *t1 = f1;
    *t2 = f2;
    *radius = fr;
    return 1;
}

int or_adjoint(const or_photons* ph, const or_beams* bm, const upbp_medium* media, int32_t n_media,
               const upbp_query_batch* q, const double* w_ph, const double* w_bm, int threads,
               double* grad_sigma_s, double* grad_sigma_a, double* grad_photon_power,
               double* grad_beam_power, int fp32_hits) {
    int e = check_batch(q, n_media);
    if (e) return e;
    const int short_q = q->beam_type == UPBP_SHORT_BEAM;
    or_medium* md = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, md);
    if (threads <= 0) threads = 1;
    int64_t nph = ph ? ph->n : 0, nbm = bm ? bm->n : 0;
    double* gs = (double*)calloc((size_t)n_media * 3, 8);
    double* gt = (double*)calloc((size_t)n_media * 3, 8);
    double* gph = (double*)calloc((size_t)(nph ? nph : 1) * 3, 8);
    double* gbm = (double*)calloc((size_t)(nbm ? nbm : 1) * 3, 8);
    int err = 0;
#pragma omp parallel num_threads(threads)
    {
        double* ls = (double*)calloc((size_t)n_media * 3, 8);
        double* lt = (double*)calloc((size_t)n_media * 3, 8);
#pragma omp for schedule(dynamic, 16)
        for (int64_t qi = 0; qi < q->count; ++qi) {
            double o[3], d[3];
            for (int a = 0; a < 3; ++a) {
                o[a] = q->origin[3 * qi + a];
                d[a] = q->direction[3 * qi + a];
            }
            double P[3] = {1, 1, 1};
            const uint64_t qkey = red_query_key(&q->origin[3 * qi], &q->direction[3 * qi]);
            int64_t s0 = q->segment_offsets[qi], s1 = q->segment_offsets[qi + 1];
            for (int64_t si = s0; si < s1; ++si) {
                const upbp_volume_segment* sg = &q->segments[si];
                if (sg->medium_id < 0 || sg->medium_id >= n_media) {
                    err = 1;
                    break;
                }
                const or_medium* m = &md[sg->medium_id];
                const int k = m->min_idx;
                double tmin = sg->dist_min, tmax = sg->dist_max;
                double g[3] = {0, 0, 0}; /* sum over hits of w_c * C_c in this segment */
                if (m->has_scattering) {
                    const double* w = w_ph ? &w_ph[3 * qi] : NULL;
                    for (int64_t i = 0; w && i < nph; ++i) {
                        double p[3] = {ph->pos[3 * i], ph->pos[3 * i + 1], ph->pos[3 * i + 2]};
                        double t, s;
                        if (!adj_photon_hit(fp32_hits, o, d, tmin, tmax, p, ph->r2[i], &t, &s)) continue;
                        double r2 = ph->r2[i];
                        double inc[3] = {-ph->inc[3 * i], -ph->inc[3 * i + 1], -ph->inc[3 * i + 2]};
                        double rho = phase_f64(d, inc, m->mean_cosine);
                        /* fp32_hits: the radial factor is also the reference's float expression
                         * (Bre.cxx:272-273); near the radius 1 - s / r^2 keeps only about
                         * 2^-24 / (1 - s / r^2) relative accuracy in float, so a double value
                         * would not be the reference's */
                        double radial = fp32_hits ? (double)(1.0f - (float)s / ph->r2[i]) : 1.0 - s / r2;
                        double K = radial / (r2 * (double)PI_F * 0.5);
                        /* SHORT query: T_c / T_k along t - tmin (Bre.cxx:257-263); the exponent
                         * -(sigma_t_c - sigma_t_k)(t - tmin) gives channel k the +(t - tmin) terms */
                        const double Ek = exp(-(double)m->sigma_t[k] * (t - tmin));
                        double sumC = 0;
                        for (int c = 0; c < 3; ++c) {
                            double E = exp(-(double)m->sigma_t[c] * (t - tmin));
                            if (short_q) E /= Ek;
                            double base = P[c] * E * rho * K; /* C_c = base * flux_c * ss_c */
                            double Cc = base * ph->power[3 * i + c] * m->sigma_s[c];
#pragma omp atomic
                            gph[3 * i + c] += w[c] * base * m->sigma_s[c];
                            ls[3 * sg->medium_id + c] += w[c] * base * ph->power[3 * i + c];
                            lt[3 * sg->medium_id + c] += w[c] * Cc * (-(t - tmin));
                            g[c] += w[c] * Cc;
                            sumC += w[c] * Cc;
                        }
                        if (short_q) lt[3 * sg->medium_id + k] += sumC * (t - tmin);
                    }
                    const double* wb = w_bm ? &w_bm[3 * qi] : NULL;
                    for (int64_t i = 0; wb && i < nbm; ++i) {
                        if (bm->medium[i] != sg->medium_id) continue;
                        double d2[3] = {bm->d[3 * i], bm->d[3 * i + 1], bm->d[3 * i + 2]};
                        double dist, st, t1, t2, radius;
                        if (!adj_beam_hit(fp32_hits, o, d, tmin, tmax, bm, i, &dist, &st, &t1, &t2, &radius)) continue;
                        /* per-cell reduction: the hit counts `mult` times, each weighted 1 / pdf */
                        double red_w = 1;
                        if (bm->red_max) {
                            float pq[3], pdf;
                            for (int a = 0; a < 3; ++a) pq[a] = (float)o[a] + (float)d[a] * (float)t1;
                            const uint32_t mult = red_decide_at(bm, qkey, pq, (uint32_t)i, &pdf);
                            if (!mult) continue;
                            red_w = (double)mult / (double)pdf;
                        }
                        double rho = phase_f64(d, d2, m->mean_cosine);
                        /* fp32_hits: the reference's float radial factor (PhotonBeam.hxx:364) */
                        const float fd = (float)dist, fr = (float)radius;
                        double radial = fp32_hits ? (double)(1.0f - fd * fd / (fr * fr))
                                                  : 1.0 - dist * dist / (radius * radius);
                        double K1 = radial * 3.0 / (4.0 * radius * st) * red_w;
                        int sb = (bm->flags[i] & UPBP_SHORT_BEAM) != 0;
                        double Eqk = exp(-(double)m->sigma_t[k] * (t1 - tmin));
                        double Ebk = exp(-(double)m->sigma_t[k] * t2);
                        double Cs[3], sumC = 0;
                        for (int c = 0; c < 3; ++c) {
                            double Eq = exp(-(double)m->sigma_t[c] * (t1 - tmin));
                            double Eb = exp(-(double)m->sigma_t[c] * t2);
                            if (short_q) Eq /= Eqk;
                            if (sb) Eb /= Ebk;
                            double base = P[c] * Eq * Eb * rho * K1;
                            double Cc = base * bm->flux[3 * i + c] * m->sigma_s[c];
                            Cs[c] = wb[c] * Cc;
#pragma omp atomic
                            gbm[3 * i + c] += wb[c] * base * m->sigma_s[c];
                            ls[3 * sg->medium_id + c] += wb[c] * base * bm->flux[3 * i + c];
                            lt[3 * sg->medium_id + c] += Cs[c] * (-(t1 - tmin) - t2);
                            g[c] += Cs[c];
                            sumC += Cs[c];
                        }
                        /* the SHORT normalisations by the channel-k transmittance (PhotonBeam.hxx:336-345) */
                        lt[3 * sg->medium_id + k] += sumC * ((short_q ? (t1 - tmin) : 0.0) + (sb ? t2 : 0.0));
                    }
                }
                /* prefix transmittance of the earlier segments: exp(-sigma_t len) each, divided for a
                 * SHORT query by the pass-through pdf exp(-sigma_t_k len) of a medium that samples
                 * distances (RaySamplePdf, Media.hxx:294-323) */
                for (int64_t sj = s0; sj < si; ++sj) {
                    const upbp_volume_segment* sp = &q->segments[sj];
                    const or_medium* mj = &md[sp->medium_id];
                    double len = (double)sp->dist_max - (double)sp->dist_min;
                    for (int c = 0; c < 3; ++c) lt[3 * sp->medium_id + c] += g[c] * (-len);
                    if (short_q && mj->min_comp > 0 && mj->has_scattering)
                        lt[3 * sp->medium_id + mj->min_idx] += (g[0] + g[1] + g[2]) * len;
                }
                for (int c = 0; c < 3; ++c)
                    P[c] *= short_q ? (double)sg->attenuation[c] / (double)sg->ray_sample_pdf
                                    : (double)sg->attenuation[c];
                if (!(P[0] > 1e-20 && P[1] > 1e-20 && P[2] > 1e-20)) break;
            }
        }
#pragma omp critical
        {
            for (int64_t k = 0; k < (int64_t)n_media * 3; ++k) {
                gs[k] += ls[k];
                gt[k] += lt[k];
            }
        }
        free(ls);
        free(lt);
    }
    /* the media use |sigma| (Media.hxx:136-140): d|x|/dx = sign(x) maps the gradient back to the
     * caller's coefficients */
    for (int64_t k = 0; k < (int64_t)n_media * 3; ++k) {
        const double sgn_s = copysign(1.0, (double)media[k / 3].scattering[k % 3]);
        const double sgn_a = copysign(1.0, (double)media[k / 3].absorption[k % 3]);
        if (grad_sigma_s) grad_sigma_s[k] = sgn_s * (gs[k] + gt[k]);
        if (grad_sigma_a) grad_sigma_a[k] = sgn_a * gt[k];
    }
    if (grad_photon_power && ph) {
        memset(grad_photon_power, 0, (size_t)ph->n_in * 3 * sizeof(double));
        for (int64_t i = 0; i < nph; ++i)
            for (int c = 0; c < 3; ++c) grad_photon_power[3 * ph->orig[i] + c] = gph[3 * i + c];
    }
    if (grad_beam_power)
        for (int64_t i = 0; i < nbm; ++i)
            for (int c = 0; c < 3; ++c) grad_beam_power[3 * i + c] = gbm[3 * i + c];
    free(gs);
    free(gt);
    free(gph);
    free(gbm);
    free(md);
    return err ? UPBP_EINVAL : 0;
}

int or_upbp_eval(const or_photons* ph, const or_beams* bm, const upbp_medium* media, int32_t n_media,
                 const upbp_query_batch* q, const upbp_mis_inputs* mis, int which, int accel, int precision,
                 int threads, float* out32, double* out64, uint64_t* counters) {
    int e = check_batch(q, n_media);
    if (e) return e;
    if (!mis || (which == 0 && !ph) || (which == 1 && (!bm || !mis->beams))) return UPBP_EINVAL;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
    int r;
    if (precision == 64)
        r = upbp_eval_all_f64(ph, bm, m, n_media, q, mis, which, accel, threads, out64, counters);
    else
        r = upbp_eval_all_f32(ph, bm, m, n_media, q, mis, which, accel, threads, out32, counters);
    free(m);
    return r;
}

int or_pp3d_eval(const or_hgrid* g, const upbp_medium* media, int32_t n_media, const upbp_pp3d_query* q, int64_t nq,
                 const upbp_mis_inputs* mis, uint32_t flags, int accel, int precision, int threads, float* out32,
                 double* out64, uint32_t* sig, uint64_t* counters) {
    if (!g || g->type != UPBP_PP3D || nq < 0 || (nq > 0 && !q) || !mis || n_media < 1 || !media) return UPBP_EINVAL;
    if (g->n_src > mis->light_vertex_count) return UPBP_EINVAL;
    or_medium* m = (or_medium*)malloc(sizeof(or_medium) * (size_t)n_media);
    or_prepare_media(media, n_media, m);
    if (threads <= 0) threads = 1;
    int r;
    if (precision == 64)
        r = pp3d_eval_all_f64(g, m, n_media, q, nq, mis, flags, accel, threads, out64, sig, counters);
    else
        r = pp3d_eval_all_f32(g, m, n_media, q, nq, mis, flags, accel, threads, out32, sig, counters);
    free(m);
    return r;
}

int or_surf_eval(const or_hgrid* g, const upbp_surf_query* q, int64_t nq, const upbp_mis_inputs* mis, uint32_t flags,
                 int accel, int precision, int threads, float* out32, double* out64, uint32_t* sig,
                 uint64_t* counters) {
    if (!g || g->type != UPBP_SURF || nq < 0 || (nq > 0 && !q) || !mis) return UPBP_EINVAL;
    if (g->n_src > mis->light_vertex_count) return UPBP_EINVAL;
    if (threads <= 0) threads = 1;
    if (precision == 64) return surf_eval_all_f64(g, q, nq, mis, flags, accel, threads, out64, sig, counters);
    return surf_eval_all_f32(g, q, nq, mis, flags, accel, threads, out32, sig, counters);
}

/* ------------------------------------------------------------------------- */
/* k-nearest-neighbour radii.
 *
 * SmallUPBP answers "k-th nearest neighbour" with Jensen's kd-tree query
 * (SmallUPBP/src/Misc/KdTmpl.hxx:600-792): a bounded max-heap of k squared
 * distances, each computed as ((dx*dx) + dy*dy) + dz*dz with d = item - query
 * (:604-610), an item entering only if its distance is strictly below the heap
 * maximum (:612), and the result read from dist2[1], the heap maximum. Plane
 * pruning (:751) never drops an item that could enter, so for more than k items
 * the result is exactly the k-th smallest squared distance, the query's own item
 * (distance 0) included. This restatement keeps the same bounded max-heap but
 * scans every item (brute force). With n <= k items the kd-tree never builds its
 * heap and dist2[1] is the first item it happened to visit; this restatement
 * returns the largest of the n distances instead, and so does the GPU library. */
static void knn_heap_push(float* heap, int* found, int k, float d2) {
    /* heap[1..k], max at heap[1] once full (1-based like KdTmpl.hxx) */
    if (*found < k) {
        int c = ++(*found);
        heap[c] = d2;
        while (c > 1 && heap[c / 2] < heap[c]) {
            float t = heap[c / 2];
            heap[c / 2] = heap[c];
            heap[c] = t;
            c /= 2;
        }
        return;
    }
    if (!(d2 < heap[1])) return;
    int p = 1;
    for (;;) {
        int j = 2 * p;
        if (j > k) break;
        if (j < k && heap[j] < heap[j + 1]) ++j;
        if (!(d2 < heap[j])) break;
        heap[p] = heap[j];
        p = j;
    }
    heap[p] = d2;
}

void or_knn_d2(const float* pts, int64_t n, const float* q, int64_t nq, int32_t k, int threads, float* out) {
    /* initial search radius MAX_FLOAT_SQUARE_ROOT (Misc/Defs.hxx:48), squared as KdTmpl.hxx:139 does */
    const float init_r = sqrtf(FLT_MAX);
    const float init_d2 = init_r * init_r;
    const int keff = (int64_t)k < n ? k : (int)n;
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        float* heap = (float*)malloc(sizeof(float) * (size_t)(keff + 2));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nq; ++i) {
            const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
            int found = 0;
            for (int64_t j = 0; j < n; ++j) {
                const float dx = pts[3 * j] - qx, dy = pts[3 * j + 1] - qy, dz = pts[3 * j + 2] - qz;
                const float d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 < init_d2) knn_heap_push(heap, &found, keff, d2);
            }
            out[i] = found > 0 ? heap[1] : init_d2;
        }
        free(heap);
    }
}

/* EmbreeBre::build with KNN_RADIUS (SmallUPBP/src/Bre/Bre.cxx:562-602): the
 * kd-tree holds the in-medium vertices; radius_i = radius * (2 * sqrt(d2_k)).
 * out[i] is written for in-medium photons only. */
void or_photon_knn_radius(int64_t n, const float* pos, const uint8_t* in_medium, float radius, int32_t knn,
                          int threads, float* out) {
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) m += (!in_medium || in_medium[i]) ? 1 : 0;
    float* pts = (float*)calloc(3 * (size_t)(m + 1), sizeof(float));
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + 1));
    float* d2 = (float*)malloc(sizeof(float) * (size_t)(m + 1));
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (in_medium && !in_medium[i]) continue;
        memcpy(&pts[3 * c], &pos[3 * i], 12);
        idx[c++] = i;
    }
    or_knn_d2(pts, m, pts, m, knn, threads, d2);
    for (int64_t j = 0; j < m; ++j) out[idx[j]] = radius * (2.0f * sqrtf(d2[j]));
    free(pts);
    free(idx);
    free(d2);
}

/* PhotonBeamsEvaluator::build with KNN_RADIUS (SmallUPBP/src/Beams/PhBeams.cxx:77-117):
 * the kd-tree holds the beam origins; each beam queries at its origin and at
 * origin + direction * length (Ray::target, Path/Ray.hxx:57-59). */
void or_beam_knn_radius(int64_t n, const float* origin, const float* dir, const float* length, float radius,
                        int32_t knn, int threads, float* r0, float* r1) {
    const float smallest = 0.001f; /* SMALLEST_RADIUS, PhBeams.cxx:72 */
    float* ends = (float*)malloc(sizeof(float) * 3 * (size_t)(n + 1));
    float* d2 = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) ends[3 * i + a] = origin[3 * i + a] + dir[3 * i + a] * length[i];
    or_knn_d2(origin, n, origin, n, knn, threads, d2);
    for (int64_t i = 0; i < n; ++i) {
        const float r = (2.0f * sqrtf(d2[i])) * radius;
        r0[i] = r < smallest ? smallest : r;
    }
    or_knn_d2(origin, n, ends, n, knn, threads, d2);
    for (int64_t i = 0; i < n; ++i) {
        const float r = (2.0f * sqrtf(d2[i])) * radius;
        r1[i] = r < smallest ? smallest : r;
    }
    free(ends);
    free(d2);
}
