// hashgrid.hip -- vertex merging on the GPU through the reference's hashed uniform
// grid (SmallUPBP/src/Misc/HashGrid.hxx:38-230) with the RangeQuery::Process weight
// (SmallUPBP/src/Renderers/UPBP.hxx:84-129): PP3D for camera vertices in media, SURF
// for camera vertices on surfaces.
//
// Build. One bounds pass over the grid's vertices (in-medium ones for PP3D, the others
// for SURF), one bucket key per vertex, a
// stable radix sort of vertex indices by bucket, and the bucket ends from the run
// boundaries of the sorted keys followed by a running maximum (empty buckets end
// where the previous one ended). The reference fills its buckets with a counting
// sort in ascending vertex order; a stable sort of ascending indices gives the same
// order inside every bucket. Each grid slot is one float4 (position, vertex index),
// so a distance test reads 16 bytes.
//
// Query. One lane per camera vertex, lanes ordered by a Morton key of the vertex
// position so that a wave's lanes walk nearby buckets. Each lane walks its eight
// buckets in the reference's order, visiting a bucket twice when two neighbour
// cells hash to it, and tests every slot with the reference's FP32 distance. The
// camera BSDF (a medium's phase function for PP3D, the surface's diffuse and Phong
// lobes for SURF), contribution and MIS weight of a vertex found within the radius
// are computed in registers with the shared path-weight walks.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "upbp_internal.hpp"
#include "upbp_mis.hpp"

namespace upbp {
namespace {

constexpr int kBlock = 256;
constexpr int kBoundsBlocks = 256;

// Grid constants of HashGrid::Build (HashGrid.hxx:52-58) and its bounding box.
struct GridDev {
    float lo[3], hi[3];
    float r2, inv_cell;
    uint32_t cells;
};

// GetCellIndex(x, y, z), HashGrid.hxx:195-206.
__device__ __forceinline__ uint32_t bucket_of(int32_t x, int32_t y, int32_t z, uint32_t cells) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) % cells;
}

// GetCellIndex(Pos), HashGrid.hxx:208-218.
__device__ __forceinline__ uint32_t bucket_of_pos(const GridDev& g, float x, float y, float z) {
    const float cx = floorf(g.inv_cell * (x - g.lo[0]));
    const float cy = floorf(g.inv_cell * (y - g.lo[1]));
    const float cz = floorf(g.inv_cell * (z - g.lo[2]));
    return bucket_of((int32_t)cx, (int32_t)cy, (int32_t)cz, g.cells);
}

// UPBPLightVertex::MatchesType (UPBPLightVertex.hxx:74-77): PP3D grids keep in-medium
// vertices, SURF grids the others.
__device__ __forceinline__ bool in_grid(const uint32_t* flags, int64_t stride, int64_t i, bool want_medium) {
    return ((flags[i * stride] & UPBP_LV_IN_MEDIUM) != 0) == want_medium;
}

// Per-block bounds of the grid's vertices: part[b] = (lo.xyz, hi.xyz, count, non-finite).
__global__ void k_grid_bounds(const float* __restrict__ pos, const uint32_t* __restrict__ flags, int64_t stride,
                              int64_t n, bool want_medium, float* __restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t cnt = 0, bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!in_grid(flags, stride, i, want_medium)) continue;
        const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if (!__builtin_isfinite(x) || !__builtin_isfinite(y) || !__builtin_isfinite(z)) {
            bad = 1;
            continue;
        }
        lo[0] = fminf(lo[0], x);
        lo[1] = fminf(lo[1], y);
        lo[2] = fminf(lo[2], z);
        hi[0] = fmaxf(hi[0], x);
        hi[1] = fmaxf(hi[1], y);
        hi[2] = fmaxf(hi[2], z);
        cnt++;
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
        }
        cnt += __shfl_xor(cnt, off, 64);
        bad |= __shfl_xor(bad, off, 64);
    }
    __shared__ float s[kBlock / 64][8];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) {
            s[w][a] = lo[a];
            s[w][3 + a] = hi[a];
        }
        s[w][6] = __uint_as_float(cnt);
        s[w][7] = __uint_as_float(bad);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[8];
        for (int k = 0; k < 8; ++k) r[k] = s[0][k];
        uint32_t c = __float_as_uint(r[6]), b = __float_as_uint(r[7]);
        for (int v = 1; v < (int)(blockDim.x >> 6); ++v) {
            for (int a = 0; a < 3; ++a) {
                r[a] = fminf(r[a], s[v][a]);
                r[3 + a] = fmaxf(r[3 + a], s[v][3 + a]);
            }
            c += __float_as_uint(s[v][6]);
            b |= __float_as_uint(s[v][7]);
        }
        r[6] = __uint_as_float(c);
        r[7] = __uint_as_float(b);
        for (int k = 0; k < 8; ++k) part[8 * blockIdx.x + k] = r[k];
    }
}

// Bucket key per vertex (the sentinel `cells` for vertices outside the grid) and its index.
__global__ void k_grid_keys(const float* __restrict__ pos, const uint32_t* __restrict__ flags, int64_t stride,
                            int64_t n, bool want_medium, GridDev g, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = in_grid(flags, stride, i, want_medium) ? bucket_of_pos(g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2])
                                                     : g.cells;
    vals[i] = (uint32_t)i;
}

// Slots in bucket order and the end of each non-empty bucket (run boundaries).
__global__ void k_grid_slots(const float* __restrict__ pos, const uint32_t* __restrict__ keys,
                             const uint32_t* __restrict__ vals, int64_t n_sel, float4* __restrict__ slots,
                             uint32_t* __restrict__ ends) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n_sel) return;
    const uint32_t v = vals[k];
    slots[k] = make_float4(pos[3 * (int64_t)v], pos[3 * (int64_t)v + 1], pos[3 * (int64_t)v + 2], __uint_as_float(v));
    const uint32_t key = keys[k];
    if (k + 1 == n_sel || keys[k + 1] != key) ends[key] = (uint32_t)(k + 1);
}

// 30-bit Morton key of a camera vertex inside the grid's box (clamped), for the lane order.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

template <class Query>
__global__ void k_query_morton(const Query* __restrict__ q, int64_t n, GridDev g, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = g.hi[a] - g.lo[a];
        float u = ext > 0.f ? (q[i].position[a] - g.lo[a]) / ext : 0.f;
        u = fminf(fmaxf(u, 0.f), 1.f);  // NaN -> 0
        c[a] = (uint32_t)(u * 1023.f);
    }
    keys[i] = (spread10(c[0]) << 2) | (spread10(c[1]) << 1) | spread10(c[2]);
    idx[i] = (uint32_t)i;
}

struct VmArgs {
    const float4* slots;
    const uint32_t* ends;
    const float4* rec;  // kRecWords float4 per slot, k_vm_records
    GridDev g;
    int64_t n_sel;
    const uint32_t* perm;
    int64_t count;
    const MediumDev* media;
    int32_t n_media;
    uint32_t flags;
    MisGlobals mis;
    float* out;
    uint32_t* sig;
    unsigned long long* stats;
    int32_t count_stats;  // the launch that adds to the call's counters (one per call)
};

constexpr float kEpsCosine = 1e-6f;  // EPS_COSINE, Defs.hxx:65
constexpr float kEpsPhong = 1e-3f;   // EPS_PHONG, Defs.hxx:68

__device__ __forceinline__ bool finite3v(const float* v) {
    return __builtin_isfinite(v[0]) && __builtin_isfinite(v[1]) && __builtin_isfinite(v[2]);
}

__device__ __forceinline__ bool camera_range_ok(const VmArgs& A, uint32_t len, int64_t off) {
    return (A.flags & UPBP_VM_NO_MIS) || (len >= 1 && off >= 0 && off + (int64_t)len < A.mis.camera_count);
}

// Per-slot merge record, gathered once per call in slot (bucket) order so that the
// vertices a query finds are read as contiguous 64-byte records instead of scattered
// light-vertex gathers:
//   [0] WorldDirFix xyz, mPathLength (bits)
//   [1] mThroughput rgb, continuation probability
//   [2] light-walk constants k0, k1 (light_split; k1 NaN: run the full walk),
//       mRaySamplePdfInv, mRaySamplePdfsRatio
//   [3] merging factor (mPP3DMisWeightFactor or mSurfMisWeightFactor), record flags (bits), -, -
constexpr int kRecWords = 4;
constexpr uint32_t kRecConnectable = 1u;  // mConnectable
constexpr uint32_t kRecBadIndex = 2u;     // the slot's vertex index is past the light vertices

__global__ void k_vm_records(VmArgs A, uint32_t type, float4* __restrict__ rec) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= A.n_sel) return;
    const MisGlobals& G = A.mis;
    const uint32_t v = __float_as_uint(A.slots[k].w);
    float4* r = rec + kRecWords * k;
    if ((int64_t)v >= G.lv_count) {
        r[0] = make_float4(0.f, 0.f, 0.f, 0.f);
        r[1] = make_float4(0.f, 0.f, 0.f, 0.f);
        r[2] = make_float4(0.f, 0.f, 0.f, 0.f);
        r[3] = make_float4(0.f, __uint_as_float(kRecBadIndex), 0.f, 0.f);
        return;
    }
    const upbp_light_vertex* __restrict__ lv = G.lv + v;
    float k0 = 0.f, k1 = 0.f;
    if (!(A.flags & UPBP_VM_NO_MIS) &&
        !light_split(G, (int64_t)v, lv->path_idx, (int)lv->path_length, G.p.techniques, k0, k1))
        k1 = __builtin_nanf("");
    r[0] = make_float4(lv->dir_fix[0], lv->dir_fix[1], lv->dir_fix[2], __uint_as_float(lv->path_length));
    r[1] = make_float4(lv->throughput[0], lv->throughput[1], lv->throughput[2], lv->continuation_prob);
    r[2] = make_float4(k0, k1, lv->mis.ray_sample_pdf_inv, lv->mis.ray_sample_pdfs_ratio);
    const float factor = type == UPBP_PP3D ? lv->mis.pp3d_factor : lv->mis.surf_factor;
    const uint32_t fl = (lv->flags & UPBP_LV_CONNECTABLE) ? kRecConnectable : 0u;
    r[3] = make_float4(factor, __uint_as_float(fl), 0.f, 0.f);
}

// The merge record of slot k if RangeQuery::Process merges with its vertex
// (UPBP.hxx:89-98): dir_fix and path length in rec[0] on success.
__device__ __forceinline__ bool vm_take(const VmArgs& A, uint32_t k, uint32_t cam_len, float4& r0, float4& r3,
                                        uint32_t& err) {
    r3 = A.rec[kRecWords * (int64_t)k + 3];
    const uint32_t fl = __float_as_uint(r3.y);
    if (fl & kRecBadIndex) {
        err |= kBadLightIndex;
        return false;
    }
    if (!(fl & kRecConnectable)) return false;
    r0 = A.rec[kRecWords * (int64_t)k];
    const uint32_t tot = __float_as_uint(r0.w) + cam_len;
    return !(tot > A.mis.p.max_path_length || tot < A.mis.p.min_path_length);
}

// The two path-weight walks of RangeQuery::Process (UPBP.hxx:120-127) and the weighted
// contribution (:128-129). f = camera BSDF value, dir_pdf / rev_pdf already multiplied by
// the camera / light continuation probabilities. Both walks are linear in their first
// reverse pdf 1 / factor: the light walk is last * (k0 + k1 * dir_pdf) from the slot's
// record, the camera walk last * ((a0 + bb0 * s * b) + cn * rev_pdf) from the query's split.
__device__ __forceinline__ void vm_accumulate(const VmArgs& A, uint32_t k, uint32_t v, const CamSplit& cs,
                                              float factor, float sin_t, float dir_pdf, float rev_pdf,
                                              uint32_t evaluated, float f0, float f1, float f2, float acc[3],
                                              uint32_t& err) {
    const float4 r1 = A.rec[kRecWords * (int64_t)k + 1];
    float w = 1.f;
    if (!(A.flags & UPBP_VM_NO_MIS)) {
        const MisGlobals& G = A.mis;
        const upbp_mis_params& P = G.p;
        const float4 r2 = A.rec[kRecWords * (int64_t)k + 2];
        const float inv = 1.f / factor;
        const float s = (P.techniques & UPBP_NO_SINE_IN_WEIGHTS) ? 1.f : sin_t;
        const float b = (P.photon_beam_type & UPBP_LONG_BEAM) ? r2.z : r2.w;
        const float w_cam = inv * ((cs.a0 + cs.bb0 * (s * b)) + cs.cn * rev_pdf);
        float w_light;
        if (!__builtin_isnan(r2.y)) {
            w_light = inv * (r2.x + r2.y * dir_pdf);
        } else {
            const upbp_light_vertex* __restrict__ lv = G.lv + v;
            LightView lview{&G, nullptr, -1, 0.f};
            w_light = light_path_weight(lview, lv->path_idx, (int)lv->path_length, inv, dir_pdf, evaluated,
                                        P.query_beam_type, P.photon_beam_type, P.techniques, err);
        }
        w = 1.f / (w_light + w_cam);
    }
    acc[0] += w * (f0 * r1.x);
    acc[1] += w * (f1 * r1.y);
    acc[2] += w * (f2 * r1.z);
}

// ---- PP3D: camera vertex in a medium -------------------------------------
struct Pp3dCam {
    MediumDev m;
};

__device__ __forceinline__ bool query_ok(const VmArgs& A, const upbp_pp3d_query& q, uint32_t& err) {
    bool ok = true;
    if (!finite3v(q.position) || !finite3v(q.dir_fix)) {
        err |= kErrNonFinite;
        ok = false;
    }
    if (q.medium < 0 || q.medium >= A.n_media) {
        err |= kErrBadMedium;
        ok = false;
    }
    if (!camera_range_ok(A, q.camera_path_length, q.camera_offset)) {
        err |= kBadLightIndex;
        ok = false;
    }
    return ok;
}

__device__ __forceinline__ Pp3dCam camera_of(const VmArgs& A, const upbp_pp3d_query& q) {
    return Pp3dCam{A.media[q.medium]};
}

// RangeQuery::Process with a medium camera BSDF (Bsdf.hxx:216-230): sigma_s * phase
// (fix = camera WorldDirFix, gen = light WorldDirFix), zero without scattering.
__device__ void vm_process(const VmArgs& A, const upbp_pp3d_query& q, const Pp3dCam& c, const CamSplit& cs,
                           uint32_t k, uint32_t v, float acc[3], uint32_t& err) {
    float4 r0, r3;
    if (!vm_take(A, k, q.camera_path_length, r0, r3, err)) return;
    const MediumDev& m = c.m;
    if (!m.has_scattering) return;
    const float cos_t = -dot3(r0.x, r0.y, r0.z, q.dir_fix[0], q.dir_fix[1], q.dir_fix[2]);
    const float sin_t = __builtin_sqrtf(fmaxf(0.f, 1.f - cos_t * cos_t));
    const float pdf = phase_pdf(m.g, m.isotropic != 0, cos_t);
    const float f0 = m.ss[0] * pdf, f1 = m.ss[1] * pdf, f2 = m.ss[2] * pdf;
    if (f0 <= 0.f && f1 <= 0.f && f2 <= 0.f) return;
    const float cont_l = A.rec[kRecWords * (int64_t)k + 1].w;
    // camera continuation on the direct pdf (UPBP.hxx:110), light continuation on the reverse (:116)
    vm_accumulate(A, k, v, cs, r3.x, sin_t, pdf * m.cont, pdf * cont_l, UPBP_PP3D, f0, f1, f2, acc, err);
}

// ---- SURF: camera vertex on a surface ------------------------------------
struct SurfCam {};

__device__ __forceinline__ bool query_ok(const VmArgs& A, const upbp_surf_query& q, uint32_t& err) {
    bool ok = true;
    if (!finite3v(q.position) || !finite3v(q.frame_x) || !finite3v(q.frame_y) || !finite3v(q.frame_z) ||
        !finite3v(q.local_dir_fix)) {
        err |= kErrNonFinite;
        ok = false;
    }
    if (!camera_range_ok(A, q.camera_path_length, q.camera_offset)) {
        err |= kBadLightIndex;
        ok = false;
    }
    return ok;
}

__device__ __forceinline__ SurfCam camera_of(const VmArgs&, const upbp_surf_query&) { return SurfCam{}; }

// RangeQuery::Process with a surface camera BSDF: Bsdf::Evaluate on a surface
// (Bsdf.hxx:176-215) = EvaluateDiffuse (:542-561) + EvaluatePhong (:563-600), with the
// light's WorldDirFix as the generated direction; sin(theta) is 0 on surfaces.
__device__ void vm_process(const VmArgs& A, const upbp_surf_query& q, const SurfCam&, const CamSplit& cs, uint32_t k,
                           uint32_t v, float acc[3], uint32_t& err) {
    float4 r0, r3;
    if (!vm_take(A, k, q.camera_path_length, r0, r3, err)) return;
    const float gx = r0.x, gy = r0.y, gz = r0.z;
    // Frame::ToLocal (Frame.hxx:69-72)
    const float lx = dot3(gx, gy, gz, q.frame_x[0], q.frame_x[1], q.frame_x[2]);
    const float ly = dot3(gx, gy, gz, q.frame_y[0], q.frame_y[1], q.frame_y[2]);
    const float lz = dot3(gx, gy, gz, q.frame_z[0], q.frame_z[1], q.frame_z[2]);
    const float fx = q.local_dir_fix[0], fy = q.local_dir_fix[1], fz = q.local_dir_fix[2];
    if (fabsf(lz) < kEpsCosine) return;  // too parallel with the tangent plane
    if (lz * fz < 0.f) return;           // the other side of the surface
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, dir_pdf = 0.f, rev_pdf = 0.f;
    const bool lobes = !(fz < kEpsCosine || lz < kEpsCosine);
    if (q.diff_prob != 0.f && lobes) {
        dir_pdf += q.diff_prob * fmaxf(0.f, lz * kInvPi);
        rev_pdf += q.diff_prob * fmaxf(0.f, fz * kInvPi);
        f0 += q.diffuse[0] * kInvPi;
        f1 += q.diffuse[1] * kInvPi;
        f2 += q.diffuse[2] * kInvPi;
    }
    if (q.phong_prob != 0.f && lobes) {
        // ReflectLocal(mLocalDirFix) = (-x, -y, z) (Utils2.hxx:374-377)
        const float dot_r = dot3(-fx, -fy, fz, lx, ly, lz);
        if (dot_r > kEpsPhong) {
            const float e = q.phong_exponent;
            const float pw = powf(dot_r, e);
            if (pw > 1e-20f) {  // Float::isPositive (Utils.hxx:109-111)
                // PowerCosHemispherePdfW (Utils2.hxx:78-86), symmetric
                // dot_r > EPS_PHONG, so powf(max(0, dot_r), e) is pw itself
                const float pdf_w = q.phong_prob * (((e + 1.f) * pw) * (kInvPi * 0.5f));
                dir_pdf += pdf_w;
                rev_pdf += pdf_w;
                f0 += ((q.phong[0] * (e + 2.f)) * 0.5f * kInvPi) * pw;
                f1 += ((q.phong[1] * (e + 2.f)) * 0.5f * kInvPi) * pw;
                f2 += ((q.phong[2] * (e + 2.f)) * 0.5f * kInvPi) * pw;
            }
        }
    }
    if (f0 <= 0.f && f1 <= 0.f && f2 <= 0.f) return;
    const float cont_l = A.rec[kRecWords * (int64_t)k + 1].w;
    vm_accumulate(A, k, v, cs, r3.x, 0.f, dir_pdf * q.continuation_prob, rev_pdf * cont_l, UPBP_SURF, f0, f1, f2,
                  acc, err);
}

// HashGrid::Process (HashGrid.hxx:130-186) over one lane's camera vertex.
template <class Query, bool SIG>
__global__ __launch_bounds__(kBlock) void k_vm_eval(VmArgs A, const Query* __restrict__ queries) {
    const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool valid = slot < A.count;
    uint32_t err = 0, processed = 0, sig_h = 0;
    uint64_t tested = 0;
    float acc[3] = {0.f, 0.f, 0.f};
    int64_t qi = 0;
    if (valid) {
        qi = A.perm ? (int64_t)A.perm[slot] : slot;
        const Query q = queries[qi];
        const bool ok = query_ok(A, q, err);
        float dmin[3], dmax[3];
        for (int a = 0; a < 3; ++a) {
            dmin[a] = q.position[a] - A.g.lo[a];
            dmax[a] = A.g.hi[a] - q.position[a];
        }
        if (ok && A.n_sel > 0 && !(dmin[0] < 0.f) && !(dmax[0] < 0.f) && !(dmin[1] < 0.f) && !(dmax[1] < 0.f) &&
            !(dmin[2] < 0.f) && !(dmax[2] < 0.f)) {
            const auto cam = camera_of(A, q);
            const CamSplit cs =
                (SIG || (A.flags & UPBP_VM_NO_MIS))
                    ? CamSplit{0.f, 0.f, 0.f}
                    : camera_split(A.mis.camera + q.camera_offset, (int)q.camera_path_length,
                                   A.mis.p.query_beam_type, A.mis.p.photon_beam_type, A.mis.p.techniques);
            int32_t p[3], po[3];
            for (int a = 0; a < 3; ++a) {
                const float cp = A.g.inv_cell * dmin[a];
                const float cf = floorf(cp);
                p[a] = (int32_t)cf;
                po[a] = p[a] + ((cp - cf) < 0.5f ? -1 : 1);
            }
            for (int j = 0; j < 8; ++j) {
                // case order of HashGrid.hxx:163-172: x slowest, z fastest
                const uint32_t c = bucket_of((j & 4) ? po[0] : p[0], (j & 2) ? po[1] : p[1], (j & 1) ? po[2] : p[2],
                                             A.g.cells);
                uint32_t k = c == 0 ? 0u : A.ends[c - 1];
                const uint32_t e = A.ends[c];
                tested += e - k;
                for (; k < e; ++k) {
                    const float4 s = A.slots[k];
                    const float dx = q.position[0] - s.x, dy = q.position[1] - s.y, dz = q.position[2] - s.z;
                    if ((dx * dx + dy * dy) + dz * dz <= A.g.r2) {
                        const uint32_t v = __float_as_uint(s.w);
                        processed++;
                        if (SIG)
                            sig_h += hash32(v);
                        else
                            vm_process(A, q, cam, cs, k, v, acc, err);
                    }
                }
            }
        }
        if (SIG) {
            A.sig[2 * qi] = processed;
            A.sig[2 * qi + 1] = sig_h;
        } else {
            A.out[3 * qi] = acc[0];
            A.out[3 * qi + 1] = acc[1];
            A.out[3 * qi + 2] = acc[2];
        }
    }
    // one atomic per wave for each counter
    uint64_t hp = processed, ht = tested;
    uint32_t he = err;
    for (int off = 32; off > 0; off >>= 1) {
        hp += __shfl_xor(hp, off, 64);
        ht += __shfl_xor(ht, off, 64);
        he |= __shfl_xor(he, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && A.count_stats) {
        if (hp) atomicAdd(A.stats + kStatPhotonHits, (unsigned long long)hp);
        if (ht) atomicAdd(A.stats + kStatPhotonCands, (unsigned long long)ht);
        if (he) atomicOr(A.stats + kStatError, (unsigned long long)he);
    }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

void inclusive_max_u32(Ctx& ctx, uint32_t* data, uint32_t* out, int64_t n) {
    size_t bytes = 0;
    UPBP_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, bytes, data, out, hipcub::Max(), (int)n, ctx.stream));
    ctx.sort_tmp.ensure(bytes);
    UPBP_HIP(hipcub::DeviceScan::InclusiveScan(ctx.sort_tmp.p, bytes, data, out, hipcub::Max(), (int)n, ctx.stream));
}

}  // namespace

void hash_grid_build(Ctx& ctx, const float* pos, const uint32_t* flags, int64_t stride, int64_t n, float radius,
                     int64_t cells, uint32_t type, HashGrid& out) {
    out.ctx = &ctx;
    out.type = type;
    out.n_src = n;
    out.cells = cells;
    out.radius = radius;
    const bool want_medium = type == UPBP_PP3D;
    // HashGrid::Build, HashGrid.hxx:52-55 (FP32 on the host, the same IEEE operations)
    out.r2 = radius * radius;
    const float cell_size = radius * 2.f;
    out.inv_cell = 1.f / cell_size;
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = 1e36f;
        out.hi[a] = -1e36f;
    }
    out.n = 0;
    out.ends.ensure((size_t)cells * 4);
    launch_fill_u32(ctx, out.ends.as<uint32_t>(), 0u, cells);
    if (n == 0) return;
    DevBuf part;
    part.ensure(kBoundsBlocks * 8 * sizeof(float));
    hipLaunchKernelGGL(k_grid_bounds, dim3(kBoundsBlocks), dim3(kBlock), 0, ctx.stream, pos, flags, stride, n,
                       want_medium, part.as<float>());
    UPBP_HIP(hipGetLastError());
    std::vector<float> hp(kBoundsBlocks * 8);
    UPBP_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(float), hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
    uint64_t count = 0;
    uint32_t bad = 0;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < kBoundsBlocks; ++b) {
        uint32_t c, e;
        std::memcpy(&c, &hp[8 * b + 6], 4);
        std::memcpy(&e, &hp[8 * b + 7], 4);
        count += c;
        bad |= e;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], hp[8 * b + a]);
            hi[a] = std::max(hi[a], hp[8 * b + 3 + a]);
        }
    }
    UPBP_REQUIRE(!bad, UPBP_ENONFINITE, "non-finite position of a light vertex of the grid");
    if (count == 0) return;
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = lo[a];
        out.hi[a] = hi[a];
        // cell coordinates must fit a 32-bit int (the reference converts them with int())
        UPBP_REQUIRE((double)(hi[a] - lo[a]) * (double)out.inv_cell < 1073741824.0, UPBP_EINVAL,
                     "merging radius too small for the extent of the light vertices (cell coordinates overflow)");
    }
    out.n = (int64_t)count;
    GridDev g;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = out.lo[a];
        g.hi[a] = out.hi[a];
    }
    g.r2 = out.r2;
    g.inv_cell = out.inv_cell;
    g.cells = (uint32_t)cells;
    DevBuf keys, keys2, vals, vals2;
    keys.ensure(n * 4);
    keys2.ensure(n * 4);
    vals.ensure(n * 4);
    vals2.ensure(n * 4);
    hipLaunchKernelGGL(k_grid_keys, dim3(blocks_for(n)), dim3(kBlock), 0, ctx.stream, pos, flags, stride, n,
                       want_medium, g, keys.as<uint32_t>(), vals.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    sort_pairs_u32(ctx, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<uint32_t>(), vals2.as<uint32_t>(), n,
                   bits_for((uint64_t)cells));
    out.slots.ensure((size_t)out.n * sizeof(float4));
    DevBuf runs;
    runs.ensure((size_t)cells * 4);
    launch_fill_u32(ctx, runs.as<uint32_t>(), 0u, cells);
    hipLaunchKernelGGL(k_grid_slots, dim3(blocks_for(out.n)), dim3(kBlock), 0, ctx.stream, pos, keys2.as<uint32_t>(),
                       vals2.as<uint32_t>(), out.n, out.slots.as<float4>(), runs.as<uint32_t>());
    UPBP_HIP(hipGetLastError());
    inclusive_max_u32(ctx, runs.as<uint32_t>(), out.ends.as<uint32_t>(), cells);
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
}

namespace {

template <class Query>
void vm_eval(Ctx& ctx, const HashGrid& grid, const Query* q, int64_t count, const MisDev& mis, uint32_t flags,
             float* out, uint32_t* sig) {
    if (count == 0) return;
    VmArgs A = {};
    A.slots = grid.slots.as<float4>();
    A.ends = grid.ends.as<uint32_t>();
    for (int a = 0; a < 3; ++a) {
        A.g.lo[a] = grid.lo[a];
        A.g.hi[a] = grid.hi[a];
    }
    A.g.r2 = grid.r2;
    A.g.inv_cell = grid.inv_cell;
    A.g.cells = (uint32_t)grid.cells;
    A.n_sel = grid.n;
    A.count = count;
    A.media = ctx.media.as<MediumDev>();
    A.n_media = ctx.n_media;
    A.flags = flags;
    A.mis.camera = mis.camera;
    A.mis.camera_count = mis.camera_count;
    A.mis.lv = mis.lv;
    A.mis.lv_count = mis.lv_count;
    A.mis.path_ends = mis.path_ends;
    A.mis.path_count = mis.path_count;
    A.mis.beams = nullptr;
    A.mis.beam_recs = nullptr;
    A.mis.photon_recs = nullptr;
    A.mis.p = mis.params;
    A.mis.tails = mis.tails;
    A.stats = ctx.stats.as<unsigned long long>();
    A.perm = nullptr;
    A.rec = nullptr;
    if (out && grid.n > 0) {
        ctx.vm_rec.ensure((size_t)grid.n * kRecWords * sizeof(float4));
        hipLaunchKernelGGL(k_vm_records, dim3(blocks_for(grid.n)), dim3(kBlock), 0, ctx.stream, A, grid.type,
                           ctx.vm_rec.as<float4>());
        UPBP_HIP(hipGetLastError());
        A.rec = ctx.vm_rec.as<float4>();
    }
    if (count > 64 && grid.n > 0) {
        ctx.qkeys.ensure(count * 4);
        ctx.qkeys2.ensure(count * 4);
        ctx.qperm.ensure(count * 4);
        ctx.qperm2.ensure(count * 4);
        hipLaunchKernelGGL(k_query_morton<Query>, dim3(blocks_for(count)), dim3(kBlock), 0, ctx.stream, q, count, A.g,
                           ctx.qkeys.as<uint32_t>(), ctx.qperm.as<uint32_t>());
        UPBP_HIP(hipGetLastError());
        sort_pairs_u32(ctx, ctx.qkeys.as<uint32_t>(), ctx.qkeys2.as<uint32_t>(), ctx.qperm.as<uint32_t>(),
                       ctx.qperm2.as<uint32_t>(), count, 30);
        A.perm = ctx.qperm2.as<uint32_t>();
    }
    if (out) {
        A.out = out;
        A.count_stats = 1;
        hipLaunchKernelGGL((k_vm_eval<Query, false>), dim3(blocks_for(count)), dim3(kBlock), 0, ctx.stream, A, q);
        UPBP_HIP(hipGetLastError());
    }
    if (sig) {
        A.sig = sig;
        A.count_stats = out ? 0 : 1;
        hipLaunchKernelGGL((k_vm_eval<Query, true>), dim3(blocks_for(count)), dim3(kBlock), 0, ctx.stream, A, q);
        UPBP_HIP(hipGetLastError());
    }
}

}  // namespace

void pp3d_eval(Ctx& ctx, const HashGrid& grid, const upbp_pp3d_query* q, int64_t count, const MisDev& mis,
               uint32_t flags, float* out, uint32_t* sig) {
    vm_eval(ctx, grid, q, count, mis, flags, out, sig);
}

void surf_eval(Ctx& ctx, const HashGrid& grid, const upbp_surf_query* q, int64_t count, const MisDev& mis,
               uint32_t flags, float* out, uint32_t* sig) {
    vm_eval(ctx, grid, q, count, mis, flags, out, sig);
}

}  // namespace upbp
