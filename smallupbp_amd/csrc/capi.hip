// capi.hip -- the extern "C" boundary declared in include/upbp.h.
//
// Every entry point catches all errors and returns an upbp_status; the message
// is kept in the context (upbp_ctx_last_error). Nothing here exits the process.
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "upbp_internal.hpp"

using namespace upbp;

struct upbp_ctx : Ctx {};
struct upbp_bre : Bre {};
struct upbp_beams : Beams {};
struct upbp_hash_grid : HashGrid {};

static_assert(sizeof(upbp_light_vertex) == 112, "upbp_light_vertex layout");
static_assert(sizeof(upbp_pp3d_query) == 48, "upbp_pp3d_query layout");
static_assert(sizeof(upbp_surf_query) == 128, "upbp_surf_query layout");

struct upbp_batch_plan {
    Ctx* ctx = nullptr;
    int n = 0;
    int64_t total_tiles = 0;
    std::vector<InstanceDev> host;
    DevBuf table;
    DevBuf counters;  // [2] queue counters (photons, beams)
    // copies of every instance's structures, merged (see MergedDev)
    DevBuf m_ph_nodes, m_ph_pos_r2, m_ph_power, m_ph_inc, m_bm_nodes, m_bm_pieces, m_bm_flux;
    MergedDev merged = {};
    std::vector<DevBuf*> perms;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_blocks = 0;
    int graph_wpb = 0;
    // the context's media generation the instances were validated against (and the graph captured with)
    uint64_t media_gen = 0;
    std::vector<int32_t> media_base;
    ~upbp_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (DevBuf* p : perms) delete p;
    }
};

// Fixed device region of upbp_eval_noalloc: the kernels' counters, the query sort's Morton keys
// and permutations for up to UPBP_NOALLOC_MAX_SORTED_QUERIES queries, and the radix sort's
// temporary storage. Every GPU the library is loaded on holds its own copy until the process
// ends. In host code `scratch` names a host-side stand-in: only the address that
// hipGetSymbolAddress returns for the current device (kept per device below) is ever used.
constexpr size_t kScratchBytes = (size_t)UPBP_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchStatsBytes = 256;
constexpr int64_t kScratchMaxSorted = UPBP_NOALLOC_MAX_SORTED_QUERIES;
static_assert(kStatSlots * 8 <= kScratchStatsBytes, "counters fit the scratch header");
static_assert(kScratchStatsBytes + 4 * (size_t)kScratchMaxSorted * 4 < kScratchBytes, "scratch keeps room for the sort");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_no_ctx_error;

// Device address of `scratch` per device index, looked up once per device with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

// Called by upbp_ctx_create with `device` current. Returns null (and the reason) on failure, so
// only the allocation-free entry point reports it.
void* scratch_address(int device, std::string& why) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            why = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// The workspace carved out of a device's scratch region: counters first, then the four
// [kScratchMaxSorted] key / permutation arrays, then the sort storage in what remains.
Workspace scratch_workspace(void* base) {
    Workspace w;
    char* p = static_cast<char*>(base);
    w.stats = reinterpret_cast<unsigned long long*>(p);
    p += kScratchStatsBytes;
    uint32_t** arrays[4] = {&w.keys, &w.keys2, &w.perm, &w.perm2};
    for (uint32_t** a : arrays) {
        *a = reinterpret_cast<uint32_t*>(p);
        p += (size_t)kScratchMaxSorted * 4;
    }
    w.max_sorted = kScratchMaxSorted;
    w.sort_tmp = p;
    w.sort_tmp_bytes = kScratchBytes - (size_t)(p - static_cast<char*>(base));
    return w;
}

upbp_status fail(Ctx* ctx, upbp_status st, const std::string& msg) {
    if (ctx)
        ctx->last_error = msg;
    else
        g_no_ctx_error = msg;
    return st;
}

template <class F>
upbp_status guarded(Ctx* ctx, F&& f) {
    try {
        // start from a clear HIP error state, so that launch checks report this call's errors only
        (void)hipGetLastError();
        if (ctx) UPBP_HIP(hipSetDevice(ctx->device));
        f();
        if (ctx) ctx->last_error.clear();
        return UPBP_OK;
    } catch (const Error& e) {
        return fail(ctx, e.status, e.message);
    } catch (const std::bad_alloc&) {
        return fail(ctx, UPBP_ENOMEM, "host allocation failed");
    } catch (...) {
        return fail(ctx, UPBP_EDEVICE, "unexpected exception");
    }
}

MediumDev derive_medium(const upbp_medium& in) {
    // HomogeneousMedium constructor, SmallUPBP/src/Scene/Media.hxx:129-159
    MediumDev m = {};
    for (int c = 0; c < 3; ++c) {
        m.sa[c] = std::fabs(in.absorption[c]);
        m.ss[c] = std::fabs(in.scattering[c]);
        m.st[c] = m.sa[c] + m.ss[c];
    }
    m.min_comp = m.st[0];
    m.min_idx = 0;
    if (m.st[1] > 0 && (m.min_comp == 0 || m.st[1] < m.min_comp)) {
        m.min_comp = m.st[1];
        m.min_idx = 1;
    }
    if (m.st[2] > 0 && (m.min_comp == 0 || m.st[2] < m.min_comp)) {
        m.min_comp = m.st[2];
        m.min_idx = 2;
    }
    m.g = in.mean_cosine;
    m.cont = in.continuation_prob;
    m.has_scattering = (m.ss[0] > 0 || m.ss[1] > 0 || m.ss[2] > 0) ? 1 : 0;
    m.isotropic = std::fabs(in.mean_cosine) < 1e-3f ? 1 : 0;
    const float mx = std::max(m.st[0], std::max(m.st[1], m.st[2]));
    m.mfp = mx > 0.f ? 1.0f / mx : INFINITY;
    return m;
}

template <class T>
const T* stage_q(Ctx& ctx, int slot, const T* src, int64_t elems, upbp_memory mem) {
    if (!src || mem == UPBP_MEM_DEVICE) return src;
    ctx.staging[slot].ensure((size_t)std::max<int64_t>(elems, 1) * sizeof(T));
    UPBP_HIP(hipMemcpyAsync(ctx.staging[slot].p, src, (size_t)elems * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
    return ctx.staging[slot].as<T>();
}

void check_batch(const Ctx& ctx, const upbp_query_batch* q) {
    UPBP_REQUIRE(q != nullptr, UPBP_EINVAL, "query batch is NULL");
    UPBP_REQUIRE(q->count >= 0, UPBP_EINVAL, "negative query count");
    UPBP_REQUIRE(q->count < ((int64_t)1 << 32), UPBP_EINVAL, "query count out of range");
    UPBP_REQUIRE(q->count == 0 || (q->origin && q->direction && q->segment_offsets), UPBP_EINVAL,
                 "query origin, direction and segment_offsets are required");
    UPBP_REQUIRE(q->beam_type == UPBP_SHORT_BEAM || q->beam_type == UPBP_LONG_BEAM, UPBP_EINVAL,
                 "query beam_type must be UPBP_SHORT_BEAM or UPBP_LONG_BEAM");
    UPBP_REQUIRE(!q->lite_segments || q->beam_type == UPBP_LONG_BEAM, UPBP_EINVAL,
                 "lite segments require LONG query beams (Bre.cxx:903)");
    UPBP_REQUIRE(q->order == UPBP_ORDER_SORTED || q->order == UPBP_ORDER_AS_GIVEN, UPBP_EINVAL, "bad query order");
    UPBP_REQUIRE(ctx.n_media > 0, UPBP_EINVAL, "upbp_set_media must be called before evaluation");
}

// The context's own workspace: its counters, and for a sorted batch of Q queries its key and
// permutation buffers and sort storage, grown to fit.
Workspace context_workspace(Ctx& ctx, const upbp_query_batch* q) {
    Workspace w;
    w.stats = ctx.stats.as<unsigned long long>();
    const int64_t Q = q->count;
    if (q->order == UPBP_ORDER_SORTED && Q > 64) {
        for (DevBuf* b : {&ctx.qkeys, &ctx.qkeys2, &ctx.qperm, &ctx.qperm2}) b->ensure((size_t)Q * 4);
        ctx.sort_tmp.ensure(sort_pairs_u32_temp_bytes(Q, 30));
        w.keys = ctx.qkeys.as<uint32_t>();
        w.keys2 = ctx.qkeys2.as<uint32_t>();
        w.perm = ctx.qperm.as<uint32_t>();
        w.perm2 = ctx.qperm2.as<uint32_t>();
        w.max_sorted = Q;
        w.sort_tmp = ctx.sort_tmp.p;
        w.sort_tmp_bytes = ctx.sort_tmp.bytes;
    }
    return w;
}

// Device view of a batch (copies host batches), with the sorted permutation when requested.
QueryView make_view(Ctx& ctx, const upbp_query_batch* q, const float* lo, const float* hi, const Workspace& ws) {
    QueryView v;
    v.count = q->count;
    v.beam_type = q->beam_type;
    v.lite = q->lite_segments;
    if (q->count == 0) return v;
    int64_t nseg = 0;
    if (q->memory == UPBP_MEM_HOST) {
        nseg = q->segment_offsets[q->count];
        UPBP_REQUIRE(nseg >= 0, UPBP_EINVAL, "segment_offsets[count] is negative");
        UPBP_REQUIRE(nseg == 0 || q->segments, UPBP_EINVAL, "segments pointer is NULL");
    }
    v.origin = stage_q(ctx, 0, q->origin, 3 * q->count, q->memory);
    v.direction = stage_q(ctx, 1, q->direction, 3 * q->count, q->memory);
    v.offsets = stage_q(ctx, 2, q->segment_offsets, q->count + 1, q->memory);
    v.segments = q->memory == UPBP_MEM_HOST ? stage_q(ctx, 3, q->segments, nseg, q->memory) : q->segments;
    if (q->order == UPBP_ORDER_SORTED && q->count > 64) {
        UPBP_REQUIRE(q->count <= ws.max_sorted && ws.keys && ws.keys2 && ws.perm && ws.perm2, UPBP_EINVAL,
                     "more queries to sort than the workspace holds");
        launch_query_keys(ctx, v, lo, hi, ws.keys, ws.perm);
        sort_pairs_u32_in(ctx, ws.sort_tmp, ws.sort_tmp_bytes, ws.keys, ws.keys2, ws.perm, ws.perm2, q->count, 30);
        v.perm = ws.perm2;
    }
    return v;
}

void scene_bounds(const Bre* bre, const Beams* bm, float* lo, float* hi) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    if (bre && bre->n > 0)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], bre->bvh.lo[a]);
            hi[a] = std::max(hi[a], bre->bvh.hi[a]);
        }
    if (bm && bm->pieces > 0)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], bm->bvh.lo[a]);
            hi[a] = std::max(hi[a], bm->bvh.hi[a]);
        }
    for (int a = 0; a < 3; ++a)
        if (!(lo[a] <= hi[a])) {
            lo[a] = 0.f;
            hi[a] = 1.f;
        }
}

void check_stats_errors(Ctx& ctx, const uint64_t* st) {
    const uint64_t e = st[kStatError];
    UPBP_REQUIRE(!(e & kErrBadOffsets), UPBP_EINVAL, "segment_offsets not monotone or outside [0, offsets[count]]");
    UPBP_REQUIRE(!(e & kErrBadMedium), UPBP_EINVAL, "segment medium_id outside the media table");
    UPBP_REQUIRE(!(e & kErrNonFinite), UPBP_ENONFINITE, "non-finite query origin, direction or segment value");
    UPBP_REQUIRE(!(e & kErrBadSegment), UPBP_EINVAL, "segment with dist_min > dist_max");
    UPBP_REQUIRE(!(e & kBadLightIndex), UPBP_EINVAL,
                 "UPBP inputs: camera range, light vertex, beam light vertex or path index out of range");
    UPBP_REQUIRE(st[kStatOverflow] == 0, UPBP_EDEVICE, "LBVH traversal stack overflow (tree deeper than 128)");
    (void)ctx;
}

void fill_stats(upbp_eval_stats* out, const uint64_t* st, float ms_ph, float ms_bm, float ms_other, uint32_t launches) {
    if (!out) return;
    out->photon_hits = st[kStatPhotonHits];
    out->beam_hits = st[kStatBeamHits];
    out->photon_candidates = st[kStatPhotonCands];
    out->beam_candidates = st[kStatBeamCands];
    out->nodes_visited = st[kStatNodes];
    out->kernel_ms_photons = ms_ph;
    out->kernel_ms_beams = ms_bm;
    out->kernel_ms_other = ms_other;
    out->launches = launches;
}

PhotonArgs photon_args(const Bre* b) {
    PhotonArgs a = {};
    a.nodes = b->bvh.nodes.as<NodeDev>();
    a.pos_r2 = b->pos_r2.as<float4>();
    a.power = b->power.as<float4>();
    a.inc = b->inc.as<float4>();
    a.orig = b->orig.as<uint32_t>();
    a.n = b->n;
    return a;
}

BeamArgs beam_args(const Beams* b) {
    BeamArgs a;
    a.nodes = b->bvh.nodes.as<NodeDev>();
    a.pieces = b->piece.as<PieceDev>();
    a.piece_flux = b->piece_flux.as<float4>();
    a.piece_beam = b->piece_beam.as<uint32_t>();
    a.pieces_n = b->pieces;
    a.red = b->red;
    return a;
}

// Shared driver of eval / signatures / adjoint.
struct RunOut {
    float* out_ph = nullptr;
    float* out_bm = nullptr;
    uint32_t* sig_ph = nullptr;
    uint32_t* sig_bm = nullptr;
};

// ws: where the call's counters, query keys and sort storage live (null: the context's own buffers).
void run_eval(Ctx& ctx, const Bre* bre, const Beams* bm, const upbp_query_batch* q, Mode mode, RunOut o,
              upbp_memory out_mem, upbp_eval_stats* stats, const AdjointOut* adj_ph, const AdjointOut* adj_bm,
              bool force_sync, const MisDev* mis = nullptr, const Workspace* ws = nullptr) {
    check_batch(ctx, q);
    if (bre) UPBP_REQUIRE(bre->ctx == &ctx, UPBP_EINVAL, "photon structure belongs to another context");
    if (bm) UPBP_REQUIRE(bm->ctx == &ctx, UPBP_EINVAL, "beam structure belongs to another context");
    const int64_t Q = q->count;
    const bool want_ph = bre && (mode == Mode::Adjoint ? adj_ph != nullptr : (o.out_ph || o.sig_ph));
    const bool want_bm = bm && (mode == Mode::Adjoint ? adj_bm != nullptr : (o.out_bm || o.sig_bm));
    const bool timing = stats != nullptr;
    if (timing) UPBP_HIP(hipEventRecord(ctx.ev[0], ctx.stream));
    float lo[3], hi[3];
    scene_bounds(bre, bm, lo, hi);
    Workspace own;
    if (!ws) {
        own = context_workspace(ctx, q);
        ws = &own;
    }
    QueryView v = make_view(ctx, q, lo, hi, *ws);
    launch_clear_u64(ctx, (uint64_t*)ws->stats, kStatSlots);
    // the UPBP light walks read per-vertex tails, built once here for both estimators
    MisDev mis_t;
    if (mis && mode == Mode::Upbp && Q > 0) {
        mis_t = *mis;
        launch_light_tails(ctx, mis_t);
        if (bm) launch_beam_records(ctx, mis_t, bm->n);
        if (bre) launch_photon_records(ctx, mis_t);
        mis = &mis_t;
    }
    // device outputs (temporary for host outputs)
    float* dph = o.out_ph;
    float* dbm = o.out_bm;
    uint32_t* sph = o.sig_ph;
    uint32_t* sbm = o.sig_bm;
    if (out_mem == UPBP_MEM_HOST && Q > 0) {
        if (dph) { ctx.staging[4].ensure(Q * 12); dph = ctx.staging[4].as<float>(); }
        if (dbm) { ctx.staging[5].ensure(Q * 12); dbm = ctx.staging[5].as<float>(); }
        if (sph) { ctx.staging[6].ensure(Q * 8); sph = ctx.staging[6].as<uint32_t>(); }
        if (sbm) { ctx.staging[7].ensure(Q * 8); sbm = ctx.staging[7].as<uint32_t>(); }
    }
    // Photons go to the side stream when beams also run, so the two kernels overlap; each
    // kernel's events are recorded on the stream it runs on.
    const bool overlap = want_ph && want_bm && bre->n > 0 && bm->pieces > 0 && Q > 0;
    hipStream_t main_stream = ctx.stream;
    struct StreamRestore {
        Ctx& c;
        hipStream_t s;
        ~StreamRestore() { c.stream = s; }
    } restore{ctx, main_stream};
    if (overlap) {
        UPBP_HIP(hipEventRecord(ctx.fork, main_stream));
        UPBP_HIP(hipStreamWaitEvent(ctx.side, ctx.fork, 0));
        ctx.stream = ctx.side;
    }
    if (timing) UPBP_HIP(hipEventRecord(ctx.ev[1], ctx.stream));
    uint32_t launches = 0;
    if (want_ph) {
        if (bre->n > 0) {
            PhotonArgs pa = photon_args(bre);
            if (mode == Mode::Upbp && bm) pa.sel = bm->red;
            launch_photons(ctx, mode, pa, v, dph, sph, adj_ph, mis, ws->stats);
            launches++;
        } else if (mode != Mode::Adjoint) {
            if (dph) launch_clear_f32(ctx, dph, 3 * Q);
            if (sph) launch_clear_f32(ctx, (float*)sph, 2 * Q);
        }
    }
    if (timing) UPBP_HIP(hipEventRecord(ctx.ev[2], ctx.stream));
    if (overlap) {
        UPBP_HIP(hipEventRecord(ctx.join, ctx.side));
        ctx.stream = main_stream;
    }
    if (timing) UPBP_HIP(hipEventRecord(ctx.ev[4], ctx.stream));
    if (want_bm) {
        if (bm->pieces > 0) {
            launch_beams(ctx, mode, beam_args(bm), v, dbm, sbm, adj_bm, mis, ws->stats);
            launches++;
        } else if (mode != Mode::Adjoint) {
            if (dbm) launch_clear_f32(ctx, dbm, 3 * Q);
            if (sbm) launch_clear_f32(ctx, (float*)sbm, 2 * Q);
        }
    }
    if (timing) UPBP_HIP(hipEventRecord(ctx.ev[3], ctx.stream));
    if (overlap) UPBP_HIP(hipStreamWaitEvent(ctx.stream, ctx.join, 0));
    if (out_mem == UPBP_MEM_HOST && Q > 0) {
        if (o.out_ph && want_ph) UPBP_HIP(hipMemcpyAsync(o.out_ph, dph, Q * 12, hipMemcpyDeviceToHost, ctx.stream));
        if (o.out_bm && want_bm) UPBP_HIP(hipMemcpyAsync(o.out_bm, dbm, Q * 12, hipMemcpyDeviceToHost, ctx.stream));
        if (o.sig_ph && want_ph) UPBP_HIP(hipMemcpyAsync(o.sig_ph, sph, Q * 8, hipMemcpyDeviceToHost, ctx.stream));
        if (o.sig_bm && want_bm) UPBP_HIP(hipMemcpyAsync(o.sig_bm, sbm, Q * 8, hipMemcpyDeviceToHost, ctx.stream));
    }
    const bool do_sync = ctx.sync || timing || force_sync || out_mem == UPBP_MEM_HOST || q->memory == UPBP_MEM_HOST;
    if (!do_sync) return;
    uint64_t st[kStatSlots];
    UPBP_HIP(hipMemcpyAsync(st, ws->stats, sizeof(st), hipMemcpyDeviceToHost, ctx.stream));
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
    float ms_other = 0.f, ms_ph = 0.f, ms_bm = 0.f;
    if (timing) {
        UPBP_HIP(hipEventElapsedTime(&ms_other, ctx.ev[0], ctx.ev[1]));
        UPBP_HIP(hipEventElapsedTime(&ms_ph, ctx.ev[1], ctx.ev[2]));
        UPBP_HIP(hipEventElapsedTime(&ms_bm, ctx.ev[4], ctx.ev[3]));
    }
    fill_stats(stats, st, ms_ph, ms_bm, ms_other, launches);
    check_stats_errors(ctx, st);
}

}  // namespace

extern "C" {

upbp_status upbp_ctx_create(int device, upbp_ctx** out_ctx) {
    if (!out_ctx) return fail(nullptr, UPBP_EINVAL, "out_ctx is NULL");
    *out_ctx = nullptr;
    upbp_ctx* ctx = nullptr;
    try {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) return fail(nullptr, UPBP_EDEVICE, "no HIP device available");
        if (device < 0 || device >= n) return fail(nullptr, UPBP_EINVAL, "device index out of range");
        UPBP_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        UPBP_HIP(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(nullptr, UPBP_EDEVICE, std::string("this library is built for gfx950 only, device is ") +
                                                   prop.gcnArchName);
        ctx = new upbp_ctx();
        ctx->device = device;
        ctx->num_cus = prop.multiProcessorCount;
        UPBP_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
        ctx->stream = ctx->own_stream;
        for (auto& ev : ctx->ev) UPBP_HIP(hipEventCreate(&ev));
        UPBP_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        UPBP_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
        UPBP_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
        ctx->stats.ensure(kStatSlots * 8);
        ctx->scratch = scratch_address(device, ctx->scratch_error);
        *out_ctx = ctx;
        return UPBP_OK;
    } catch (const Error& e) {
        delete ctx;
        return fail(nullptr, e.status, e.message);
    } catch (...) {
        delete ctx;
        return fail(nullptr, UPBP_EDEVICE, "context creation failed");
    }
}

namespace {
void ctx_retain(Ctx* ctx) { ctx->refs.fetch_add(1); }

// Frees the context once the caller and every structure built on it have let go.
void ctx_release(Ctx* ctx) {
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete static_cast<upbp_ctx*>(ctx);
    // a failed teardown call must not surface as the error of an unrelated later launch
    (void)hipGetLastError();
}
}  // namespace

upbp_status upbp_ctx_destroy(upbp_ctx* ctx) {
    if (!ctx) return UPBP_OK;
    ctx_release(ctx);
    return UPBP_OK;
}

const char* upbp_ctx_last_error(const upbp_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : g_no_ctx_error.c_str();
}

upbp_status upbp_ctx_set_stream(upbp_ctx* ctx, void* hip_stream) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return UPBP_OK;
}

void* upbp_ctx_get_stream(const upbp_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

upbp_status upbp_ctx_set_sync(upbp_ctx* ctx, int sync) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    ctx->sync = sync != 0;
    return UPBP_OK;
}

upbp_status upbp_set_media(upbp_ctx* ctx, const upbp_medium* media, int32_t count) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(count > 0 && count <= (1 << 24) && media, UPBP_EINVAL, "media must be a non-empty array");
        std::vector<MediumDev> m(count);
        std::vector<float> sign((size_t)count * 6);
        for (int i = 0; i < count; ++i) {
            for (int c = 0; c < 3; ++c)
                UPBP_REQUIRE(std::isfinite(media[i].absorption[c]) && std::isfinite(media[i].scattering[c]),
                             UPBP_ENONFINITE, "non-finite medium coefficient");
            UPBP_REQUIRE(media[i].mean_cosine >= -1.f && media[i].mean_cosine <= 1.f, UPBP_EINVAL,
                         "mean_cosine outside [-1, 1]");
            m[i] = derive_medium(media[i]);
            for (int c = 0; c < 3; ++c) {
                sign[6 * i + c] = std::copysign(1.0f, media[i].absorption[c]);
                sign[6 * i + 3 + c] = std::copysign(1.0f, media[i].scattering[c]);
            }
        }
        // a batched plan's graph may hold the old table's address: it re-captures when media_gen moves
        ctx->media_gen++;
        ctx->n_media = 0;
        ctx->media.ensure(count * sizeof(MediumDev));
        ctx->media_sign.ensure(sign.size() * sizeof(float));
        UPBP_HIP(hipMemcpyAsync(ctx->media.p, m.data(), count * sizeof(MediumDev), hipMemcpyHostToDevice, ctx->stream));
        UPBP_HIP(hipMemcpyAsync(ctx->media_sign.p, sign.data(), sign.size() * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
        ctx->media_host = m;
        ctx->media_sign_host = sign;
        ctx->n_media = count;
    });
}

upbp_status upbp_bre_build(upbp_ctx* ctx, const upbp_photon_input* photons, const upbp_build_params* params,
                           upbp_bre** out_bre, int64_t* out_num_photons) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    if (!photons || !params || !out_bre) return fail(ctx, UPBP_EINVAL, "photons, params and out_bre are required");
    *out_bre = nullptr;
    upbp_bre* b = new (std::nothrow) upbp_bre();
    if (!b) return fail(ctx, UPBP_ENOMEM, "host allocation failed");
    upbp_status st = guarded(ctx, [&] { bre_build(*ctx, *photons, *params, *b); });
    if (st != UPBP_OK) {
        delete b;
        return st;
    }
    ctx_retain(ctx);
    *out_bre = b;
    if (out_num_photons) *out_num_photons = b->n;
    return UPBP_OK;
}

upbp_status upbp_bre_destroy(upbp_bre* bre) {
    if (!bre) return UPBP_OK;
    Ctx* ctx = bre->ctx;
    (void)hipSetDevice(ctx->device);
    delete bre;
    ctx_release(ctx);
    return UPBP_OK;
}

upbp_status upbp_bre_set_power(upbp_bre* bre, const float* power, upbp_memory memory) {
    if (!bre) return fail(nullptr, UPBP_EINVAL, "bre is NULL");
    Ctx* ctx = bre->ctx;
    return guarded(ctx, [&] {
        if (bre->n == 0) return;
        UPBP_REQUIRE(power, UPBP_EINVAL, "power is NULL");
        const float* p = stage_q(*ctx, 0, power, 3 * bre->n_in, memory);
        bre_set_power(*bre, p);
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int64_t upbp_bre_count(const upbp_bre* bre) { return bre ? bre->n : 0; }

namespace {
// Copies n floats of a device buffer (or zeros when it is empty) to a caller buffer.
void copy_out_f32(Ctx& ctx, const DevBuf& src, int64_t n, float* dst, upbp_memory memory) {
    if (n <= 0) return;
    const hipMemcpyKind kind = memory == UPBP_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (src.p) {
        UPBP_HIP(hipMemcpyAsync(dst, src.p, (size_t)n * 4, kind, ctx.stream));
    } else if (memory == UPBP_MEM_DEVICE) {
        launch_clear_f32(ctx, dst, n);
    } else {
        std::fill(dst, dst + n, 0.f);
    }
    UPBP_HIP(hipStreamSynchronize(ctx.stream));
}
}  // namespace

upbp_status upbp_bre_radii(const upbp_bre* bre, float* radius, upbp_memory memory) {
    if (!bre) return fail(nullptr, UPBP_EINVAL, "bre is NULL");
    Ctx* ctx = bre->ctx;
    return guarded(ctx, [&] {
        UPBP_REQUIRE(radius || bre->n_in == 0, UPBP_EINVAL, "radius is NULL");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        // radius_in is empty when the build kept no photon: zeros then
        copy_out_f32(*ctx, bre->radius_in, bre->n_in, radius, memory);
    });
}

upbp_status upbp_beams_build(upbp_ctx* ctx, const upbp_beam_input* beams, const upbp_build_params* params,
                             upbp_beams** out_beams) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    if (!beams || !params || !out_beams) return fail(ctx, UPBP_EINVAL, "beams, params and out_beams are required");
    *out_beams = nullptr;
    upbp_beams* b = new (std::nothrow) upbp_beams();
    if (!b) return fail(ctx, UPBP_ENOMEM, "host allocation failed");
    upbp_status st = guarded(ctx, [&] { beams_build(*ctx, *beams, *params, *b); });
    if (st != UPBP_OK) {
        delete b;
        return st;
    }
    ctx_retain(ctx);
    *out_beams = b;
    return UPBP_OK;
}

upbp_status upbp_beams_destroy(upbp_beams* beams) {
    if (!beams) return UPBP_OK;
    Ctx* ctx = beams->ctx;
    (void)hipSetDevice(ctx->device);
    delete beams;
    ctx_release(ctx);
    return UPBP_OK;
}

int64_t upbp_beams_count(const upbp_beams* beams) { return beams ? beams->n : 0; }
int64_t upbp_beams_piece_count(const upbp_beams* beams) { return beams ? beams->pieces : 0; }

upbp_status upbp_beams_radii(const upbp_beams* beams, float* start_radius, float* end_radius,
                             upbp_memory memory) {
    if (!beams) return fail(nullptr, UPBP_EINVAL, "beams is NULL");
    Ctx* ctx = beams->ctx;
    return guarded(ctx, [&] {
        UPBP_REQUIRE((start_radius && end_radius) || beams->n == 0, UPBP_EINVAL,
                     "start_radius and end_radius are required");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        copy_out_f32(*ctx, beams->r_start, beams->n, start_radius, memory);
        copy_out_f32(*ctx, beams->r_end, beams->n, end_radius, memory);
    });
}

float upbp_beams_selection_pdf(const upbp_beams* beams, const float position[3]) {
    // 1 without reduction (the reference default sMaxBeamsInCell = 0, PhBeams.cxx:51)
    if (!beams || !position) return 1.0f;
    // a failed device read must not pass for "no reduction": NaN, with the reason in the context
    float pdf = NAN;
    const upbp_status st = guarded(beams->ctx, [&] { pdf = beam_selection_pdf(*beams, position); });
    return st == UPBP_OK ? pdf : NAN;
}

upbp_status upbp_beams_selection_pdfs(const upbp_beams* beams, const float* positions, int64_t count, float* out_pdf,
                                      upbp_memory memory) {
    if (!beams) return fail(nullptr, UPBP_EINVAL, "beams is NULL");
    Ctx* ctx = beams->ctx;
    return guarded(ctx, [&] {
        UPBP_REQUIRE(count >= 0 && count < ((int64_t)1 << 40), UPBP_EINVAL, "count out of range");
        UPBP_REQUIRE(count == 0 || (positions && out_pdf), UPBP_EINVAL, "positions and out_pdf are required");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        if (count == 0 || !ctx) return;
        beam_selection_pdfs(*beams, positions, count, out_pdf, memory);
    });
}

upbp_status upbp_bre_eval_brute_force(upbp_ctx* ctx, const upbp_bre* bre, const float* origin, const float* direction,
                                      const float* min_t, const float* max_t, const int32_t* medium_id, int64_t count,
                                      float* out, upbp_memory memory) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(bre != nullptr, UPBP_EINVAL, "bre is NULL");
        UPBP_REQUIRE(bre->ctx == ctx, UPBP_EINVAL, "bre belongs to another context");
        UPBP_REQUIRE(count >= 0 && count < ((int64_t)1 << 31), UPBP_EINVAL, "query count out of range");
        UPBP_REQUIRE(count == 0 || (origin && direction && min_t && max_t && medium_id && out), UPBP_EINVAL,
                     "origin, direction, min_t, max_t, medium_id and out are required");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        UPBP_REQUIRE(ctx->n_media > 0, UPBP_EINVAL, "upbp_set_media must be called before evaluation");
        if (count == 0) return;
        float* dout = out;
        DevBuf tmp;
        if (memory == UPBP_MEM_HOST) {
            tmp.ensure((size_t)count * 12);
            dout = tmp.as<float>();
        }
        if (bre->n == 0) {
            launch_clear_f32(*ctx, dout, 3 * count);
        } else {
            bre_eval_brute_force(*ctx, *bre, stage_q(*ctx, 0, origin, 3 * count, memory),
                                 stage_q(*ctx, 1, direction, 3 * count, memory), stage_q(*ctx, 2, min_t, count, memory),
                                 stage_q(*ctx, 3, max_t, count, memory), stage_q(*ctx, 4, medium_id, count, memory),
                                 count, dout);
        }
        if (memory == UPBP_MEM_HOST)
            UPBP_HIP(hipMemcpyAsync(out, dout, (size_t)count * 12, hipMemcpyDeviceToHost, ctx->stream));
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

upbp_status upbp_beams_grid_stats(upbp_ctx* ctx, const upbp_beams* beams, const upbp_query_batch* queries,
                                  uint32_t* out_stats, upbp_memory out_memory) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(beams != nullptr, UPBP_EINVAL, "beams is NULL");
        UPBP_REQUIRE(beams->ctx == ctx, UPBP_EINVAL, "beams belong to another context");
        check_batch(*ctx, queries);
        const int64_t Q = queries->count;
        UPBP_REQUIRE(Q == 0 || out_stats, UPBP_EINVAL, "out_stats is NULL");
        UPBP_REQUIRE(out_memory == UPBP_MEM_HOST || out_memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        if (Q == 0) return;
        // one thread per query in the given order: no sort, no workspace
        upbp_query_batch qb = *queries;
        qb.order = UPBP_ORDER_AS_GIVEN;
        const float none[3] = {0.f, 0.f, 0.f};
        const QueryView v = make_view(*ctx, &qb, none, none, Workspace{});
        DevBuf tmp;
        uint32_t* dout = out_stats;
        if (out_memory == UPBP_MEM_HOST) {
            tmp.ensure((size_t)Q * 12);
            dout = tmp.as<uint32_t>();
        }
        beam_grid_stats(*ctx, *beams, v, dout);
        if (out_memory == UPBP_MEM_HOST) {
            UPBP_HIP(hipMemcpyAsync(out_stats, dout, (size_t)Q * 12, hipMemcpyDeviceToHost, ctx->stream));
            UPBP_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

upbp_status upbp_eval(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams, const upbp_query_batch* queries,
                      float* out_photons, float* out_beams, upbp_memory out_memory, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(!bre || out_photons, UPBP_EINVAL, "out_photons is NULL but a photon structure was given");
        UPBP_REQUIRE(!beams || out_beams, UPBP_EINVAL, "out_beams is NULL but a beam structure was given");
        RunOut o;
        o.out_ph = bre ? out_photons : nullptr;
        o.out_bm = beams ? out_beams : nullptr;
        run_eval(*ctx, bre, beams, queries, Mode::Forward, o, out_memory, stats, nullptr, nullptr, false);
    });
}

upbp_status upbp_eval_noalloc(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                              const upbp_query_batch* queries, float* out_photons, float* out_beams,
                              upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(ctx->scratch != nullptr, UPBP_EDEVICE,
                     "the allocation-free region is unavailable on this device: " + ctx->scratch_error);
        UPBP_REQUIRE(queries != nullptr, UPBP_EINVAL, "query batch is NULL");
        UPBP_REQUIRE(queries->memory == UPBP_MEM_DEVICE, UPBP_EINVAL,
                     "upbp_eval_noalloc takes device-resident queries only (host queries need staging copies)");
        UPBP_REQUIRE(!(queries->order == UPBP_ORDER_SORTED && queries->count > kScratchMaxSorted), UPBP_EINVAL,
                     "upbp_eval_noalloc sorts at most UPBP_NOALLOC_MAX_SORTED_QUERIES (" +
                         std::to_string(kScratchMaxSorted) + ") queries; got " + std::to_string(queries->count) +
                         " (pass UPBP_ORDER_AS_GIVEN for larger batches)");
        UPBP_REQUIRE(!bre || out_photons, UPBP_EINVAL, "out_photons is NULL but a photon structure was given");
        UPBP_REQUIRE(!beams || out_beams, UPBP_EINVAL, "out_beams is NULL but a beam structure was given");
        RunOut o;
        o.out_ph = bre ? out_photons : nullptr;
        o.out_bm = beams ? out_beams : nullptr;
        const Workspace ws = scratch_workspace(ctx->scratch);
        run_eval(*ctx, bre, beams, queries, Mode::Forward, o, UPBP_MEM_DEVICE, stats, nullptr, nullptr, false, nullptr,
                 &ws);
    });
}

upbp_status upbp_eval_upbp(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                           const upbp_query_batch* queries, const upbp_mis_inputs* mis, float* out_photons,
                           float* out_beams, upbp_memory out_memory, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        check_batch(*ctx, queries);
        UPBP_REQUIRE(mis != nullptr, UPBP_EINVAL, "mis inputs are required");
        // the hit queue packs a reduced hit's test count into 23 bits; a cell tests at most
        // max_beams_in_cell beams, and reduction happens only in cells holding more
        UPBP_REQUIRE(!beams || beams->red.max < (1u << 23), UPBP_EINVAL,
                     "max_beams_in_cell must be below 2^23 for UPBP evaluation");
        UPBP_REQUIRE(!bre || out_photons, UPBP_EINVAL, "out_photons is NULL but a photon structure was given");
        // beams with out_beams NULL: the structure only gives PB2D hits their BB1D selection pdf
        const int64_t Q = queries->count;
        UPBP_REQUIRE(Q == 0 || mis->queries, UPBP_EINVAL, "mis->queries is required");
        UPBP_REQUIRE(mis->camera_count > 0 && mis->camera, UPBP_EINVAL, "mis->camera is required");
        UPBP_REQUIRE(mis->light_vertex_count >= 0 && (mis->light_vertex_count == 0 || mis->light_vertices),
                     UPBP_EINVAL, "mis->light_vertices is required");
        UPBP_REQUIRE(mis->path_count >= 0 && (mis->path_count == 0 || mis->path_ends), UPBP_EINVAL,
                     "mis->path_ends is required");
        UPBP_REQUIRE(!beams || !out_beams || beams->n == 0 || mis->beams, UPBP_EINVAL,
                     "mis->beams is required to evaluate beams");
        UPBP_REQUIRE(!bre || bre->n_in <= mis->light_vertex_count, UPBP_EINVAL,
                     "the photon structure must be built from the light vertices (photon i = light vertex i)");
        const upbp_mis_params& P = mis->params;
        UPBP_REQUIRE(P.query_beam_type == queries->beam_type, UPBP_EINVAL,
                     "params.query_beam_type must equal the batch beam_type");
        UPBP_REQUIRE(P.pb2d_factor != 0.f || !bre, UPBP_EINVAL, "pb2d_factor must be nonzero with photons");
        UPBP_REQUIRE(P.bb1d_factor != 0.f || !beams, UPBP_EINVAL, "bb1d_factor must be nonzero with beams");
        MisDev d;
        d.params = P;
        // staging slots 0..3 hold the query batch; 4..7 the outputs; MIS arrays get their own buffers
        DevBuf sq, sc, sl, sp, sb;
        auto stage_mis = [&](DevBuf& buf, const void* src, size_t bytes) -> const void* {
            if (!src || mis->memory == UPBP_MEM_DEVICE) return src;
            buf.ensure(std::max<size_t>(bytes, 16));
            UPBP_HIP(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
            return buf.p;
        };
        d.queries = (const upbp_query_mis*)stage_mis(sq, mis->queries, (size_t)Q * sizeof(upbp_query_mis));
        d.camera = (const upbp_mis_data*)stage_mis(sc, mis->camera, (size_t)mis->camera_count * sizeof(upbp_mis_data));
        d.camera_count = mis->camera_count;
        d.lv = (const upbp_light_vertex*)stage_mis(sl, mis->light_vertices,
                                                   (size_t)mis->light_vertex_count * sizeof(upbp_light_vertex));
        d.lv_count = mis->light_vertex_count;
        d.path_ends = (const int32_t*)stage_mis(sp, mis->path_ends, (size_t)mis->path_count * 4);
        d.path_count = mis->path_count;
        d.beams = (const upbp_beam_mis*)stage_mis(sb, out_beams ? mis->beams : nullptr,
                                                  beams ? (size_t)beams->n * sizeof(upbp_beam_mis) : 0);
        RunOut o;
        o.out_ph = bre ? out_photons : nullptr;
        o.out_bm = beams ? out_beams : nullptr;
        run_eval(*ctx, bre, beams, queries, Mode::Upbp, o, out_memory, stats, nullptr, nullptr, true, &d);
    });
}

upbp_status upbp_hash_grid_build(upbp_ctx* ctx, const float* positions, const upbp_light_vertex* vertices,
                                 int64_t count, float radius, int64_t cell_count, uint32_t type,
                                 upbp_memory memory, upbp_hash_grid** out_grid) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    if (!out_grid) return fail(ctx, UPBP_EINVAL, "out_grid is NULL");
    *out_grid = nullptr;
    upbp_hash_grid* g = new (std::nothrow) upbp_hash_grid();
    if (!g) return fail(ctx, UPBP_ENOMEM, "host allocation failed");
    upbp_status st = guarded(ctx, [&] {
        UPBP_REQUIRE(type == UPBP_PP3D || type == UPBP_SURF, UPBP_EINVAL, "grid type must be UPBP_PP3D or UPBP_SURF");
        UPBP_REQUIRE(count >= 0 && count < ((int64_t)1 << 31), UPBP_EINVAL, "light vertex count out of range");
        UPBP_REQUIRE(count == 0 || (positions && vertices), UPBP_EINVAL, "positions and vertices are required");
        UPBP_REQUIRE(std::isfinite(radius) && radius > 0.f, UPBP_EINVAL, "merging radius must be finite and > 0");
        UPBP_REQUIRE(cell_count >= 1 && cell_count <= ((int64_t)1 << 30), UPBP_EINVAL,
                     "cell_count must be in [1, 2^30]");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        g->ctx = ctx;
        const float* pos = positions;
        const uint32_t* flags = nullptr;
        int64_t stride = 1;
        DevBuf pos_d, flags_d;
        if (count > 0) {
            if (memory == UPBP_MEM_HOST) {
                pos_d.ensure((size_t)count * 12);
                UPBP_HIP(hipMemcpyAsync(pos_d.p, positions, (size_t)count * 12, hipMemcpyHostToDevice, ctx->stream));
                std::vector<uint32_t> fl((size_t)count);
                for (int64_t i = 0; i < count; ++i) fl[(size_t)i] = vertices[i].flags;
                flags_d.ensure((size_t)count * 4);
                UPBP_HIP(hipMemcpyAsync(flags_d.p, fl.data(), (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
                UPBP_HIP(hipStreamSynchronize(ctx->stream));  // fl leaves scope
                pos = pos_d.as<float>();
                flags = flags_d.as<uint32_t>();
            } else {
                flags = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(vertices) +
                                                          offsetof(upbp_light_vertex, flags));
                stride = (int64_t)(sizeof(upbp_light_vertex) / sizeof(uint32_t));
            }
        }
        hash_grid_build(*ctx, pos, flags, stride, count, radius, cell_count, type, *g);
    });
    if (st != UPBP_OK) {
        delete g;
        return st;
    }
    ctx_retain(ctx);
    *out_grid = g;
    return UPBP_OK;
}

upbp_status upbp_hash_grid_destroy(upbp_hash_grid* grid) {
    if (!grid) return UPBP_OK;
    Ctx* ctx = grid->ctx;
    (void)hipSetDevice(ctx->device);
    delete grid;
    ctx_release(ctx);
    return UPBP_OK;
}

int64_t upbp_hash_grid_count(const upbp_hash_grid* grid) { return grid ? grid->n : 0; }

}  // extern "C"

namespace {

// Shared driver of upbp_pp3d_eval and upbp_surf_eval.
template <class Query, class Launch>
void run_vm_eval(Ctx* ctx, const HashGrid* grid, uint32_t type, const Query* queries, int64_t count,
                 const upbp_mis_inputs* mis, uint32_t flags, float* out, uint32_t* sig, upbp_memory memory,
                 upbp_eval_stats* stats, Launch&& launch) {
    UPBP_REQUIRE(grid != nullptr, UPBP_EINVAL, "grid is NULL");
    UPBP_REQUIRE(grid->ctx == ctx, UPBP_EINVAL, "grid belongs to another context");
    UPBP_REQUIRE(grid->type == type, UPBP_EINVAL,
                 type == UPBP_PP3D ? "PP3D evaluation needs a grid built with type UPBP_PP3D"
                                   : "SURF evaluation needs a grid built with type UPBP_SURF");
    UPBP_REQUIRE(count >= 0 && count < ((int64_t)1 << 32), UPBP_EINVAL, "query count out of range");
    UPBP_REQUIRE(count == 0 || queries, UPBP_EINVAL, "queries is NULL");
    UPBP_REQUIRE(out || sig, UPBP_EINVAL, "out and sig are both NULL");
    UPBP_REQUIRE((flags & ~UPBP_VM_NO_MIS) == 0, UPBP_EINVAL, "unknown vertex-merging flags");
    UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
    UPBP_REQUIRE(type != UPBP_PP3D || ctx->n_media > 0, UPBP_EINVAL,
                 "upbp_set_media must be called before PP3D evaluation");
    UPBP_REQUIRE(mis != nullptr, UPBP_EINVAL, "mis inputs are required (light vertices)");
    UPBP_REQUIRE(mis->light_vertex_count >= grid->n_src && (mis->light_vertex_count == 0 || mis->light_vertices),
                 UPBP_EINVAL, "mis->light_vertices must hold the light vertices the grid was built from");
    const bool use_mis = !(flags & UPBP_VM_NO_MIS);
    UPBP_REQUIRE(!use_mis || (mis->camera_count > 0 && mis->camera), UPBP_EINVAL, "mis->camera is required");
    UPBP_REQUIRE(mis->path_count >= 0 && (mis->path_count == 0 || mis->path_ends), UPBP_EINVAL,
                 "mis->path_ends is required");
    const bool timing = stats != nullptr;
    if (timing) UPBP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    MisDev d = {};
    d.params = mis->params;
    DevBuf sc, sl, sp, sq;
    auto stage = [&](DevBuf& buf, const void* src, size_t bytes, upbp_memory mem) -> const void* {
        if (!src || mem == UPBP_MEM_DEVICE) return src;
        buf.ensure(std::max<size_t>(bytes, 16));
        UPBP_HIP(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return buf.p;
    };
    d.camera = use_mis ? (const upbp_mis_data*)stage(sc, mis->camera,
                                                     (size_t)mis->camera_count * sizeof(upbp_mis_data), mis->memory)
                       : nullptr;
    d.camera_count = use_mis ? mis->camera_count : 0;
    d.lv = (const upbp_light_vertex*)stage(sl, mis->light_vertices,
                                           (size_t)mis->light_vertex_count * sizeof(upbp_light_vertex), mis->memory);
    d.lv_count = mis->light_vertex_count;
    d.path_ends = (const int32_t*)stage(sp, mis->path_ends, (size_t)mis->path_count * 4, mis->memory);
    d.path_count = mis->path_count;
    const Query* dq = (const Query*)stage(sq, queries, (size_t)count * sizeof(Query), memory);
    float* dout = out;
    uint32_t* dsig = sig;
    if (memory == UPBP_MEM_HOST && count > 0) {
        if (out) {
            ctx->staging[4].ensure(count * 12);
            dout = ctx->staging[4].as<float>();
        }
        if (sig) {
            ctx->staging[6].ensure(count * 8);
            dsig = ctx->staging[6].as<uint32_t>();
        }
    }
    launch_clear_u64(*ctx, ctx->stats.as<uint64_t>(), kStatSlots);
    if (timing) UPBP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    if (use_mis && count > 0) launch_light_tails(*ctx, d);
    launch(dq, d, dout, dsig);
    if (timing) UPBP_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    if (memory == UPBP_MEM_HOST && count > 0) {
        if (out) UPBP_HIP(hipMemcpyAsync(out, dout, count * 12, hipMemcpyDeviceToHost, ctx->stream));
        if (sig) UPBP_HIP(hipMemcpyAsync(sig, dsig, count * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    // the staged inputs are freed on return, so the call always completes here
    uint64_t st[kStatSlots];
    UPBP_HIP(hipMemcpyAsync(st, ctx->stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    UPBP_HIP(hipStreamSynchronize(ctx->stream));
    float ms_other = 0.f, ms_k = 0.f;
    if (timing) {
        UPBP_HIP(hipEventElapsedTime(&ms_other, ctx->ev[0], ctx->ev[1]));
        UPBP_HIP(hipEventElapsedTime(&ms_k, ctx->ev[1], ctx->ev[2]));
    }
    fill_stats(stats, st, ms_k, 0.f, ms_other, (out ? 1u : 0u) + (sig ? 1u : 0u));
    const uint64_t e = st[kStatError];
    UPBP_REQUIRE(!(e & kErrNonFinite), UPBP_ENONFINITE, "non-finite camera vertex position, direction or frame");
    UPBP_REQUIRE(!(e & kErrBadMedium), UPBP_EINVAL, "camera vertex medium outside the media table");
    UPBP_REQUIRE(!(e & kBadLightIndex), UPBP_EINVAL,
                 "vertex merging inputs: camera range, light vertex or path index out of range");
}

}  // namespace

extern "C" {

upbp_status upbp_pp3d_eval(upbp_ctx* ctx, const upbp_hash_grid* grid, const upbp_pp3d_query* queries, int64_t count,
                           const upbp_mis_inputs* mis, uint32_t flags, float* out, uint32_t* sig,
                           upbp_memory memory, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        run_vm_eval(ctx, grid, UPBP_PP3D, queries, count, mis, flags, out, sig, memory, stats,
                    [&](const upbp_pp3d_query* dq, const MisDev& d, float* o, uint32_t* sg) {
                        pp3d_eval(*ctx, *grid, dq, count, d, flags, o, sg);
                    });
    });
}

upbp_status upbp_surf_eval(upbp_ctx* ctx, const upbp_hash_grid* grid, const upbp_surf_query* queries, int64_t count,
                           const upbp_mis_inputs* mis, uint32_t flags, float* out, uint32_t* sig,
                           upbp_memory memory, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        run_vm_eval(ctx, grid, UPBP_SURF, queries, count, mis, flags, out, sig, memory, stats,
                    [&](const upbp_surf_query* dq, const MisDev& d, float* o, uint32_t* sg) {
                        surf_eval(*ctx, *grid, dq, count, d, flags, o, sg);
                    });
    });
}

upbp_status upbp_trace_light_paths(upbp_ctx* ctx, const upbp_region* regions, int32_t region_count,
                                   const upbp_point_light* lights, int32_t light_count,
                                   const upbp_light_trace_params* params, int64_t path_count, upbp_light_paths* out,
                                   upbp_memory memory) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        UPBP_REQUIRE(params && out, UPBP_EINVAL, "params and out are required");
        UPBP_REQUIRE(ctx->n_media > 0, UPBP_EINVAL, "upbp_set_media must be called before light tracing");
        UPBP_REQUIRE(region_count >= 0 && region_count <= UPBP_MAX_REGIONS && (region_count == 0 || regions),
                     UPBP_EINVAL, "region_count must be 0..UPBP_MAX_REGIONS with regions given");
        UPBP_REQUIRE(light_count >= 1 && lights, UPBP_EINVAL, "at least one point light is required");
        UPBP_REQUIRE(path_count >= 0 && path_count < ((int64_t)1 << 31), UPBP_EINVAL, "path count out of range");
        UPBP_REQUIRE(memory == UPBP_MEM_HOST || memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "bad memory kind");
        const upbp_light_trace_params& P = *params;
        UPBP_REQUIRE(P.photon_beam_type == UPBP_SHORT_BEAM || P.photon_beam_type == UPBP_LONG_BEAM, UPBP_EINVAL,
                     "photon_beam_type must be UPBP_SHORT_BEAM or UPBP_LONG_BEAM");
        UPBP_REQUIRE(P.max_path_length >= 1 && P.max_path_length <= 1024, UPBP_EINVAL,
                     "max_path_length must be 1..1024");
        UPBP_REQUIRE(P.global_medium >= 0 && P.global_medium < ctx->n_media, UPBP_EINVAL,
                     "global_medium outside the media table");
        for (int32_t i = 0; i < region_count; ++i) {
            const upbp_region& g = regions[i];
            UPBP_REQUIRE(g.shape == UPBP_REGION_SPHERE || g.shape == UPBP_REGION_BOX, UPBP_EINVAL, "bad region shape");
            UPBP_REQUIRE(g.medium_id >= 0 && g.medium_id < ctx->n_media, UPBP_EINVAL,
                         "region medium outside the media table");
            bool fin = true;
            for (int a = 0; a < 3; ++a) fin = fin && std::isfinite(g.p0[a]) && std::isfinite(g.p1[a]);
            UPBP_REQUIRE(fin, UPBP_ENONFINITE, "non-finite region parameter");
            UPBP_REQUIRE(g.shape != UPBP_REGION_SPHERE || g.p1[0] > 0.f, UPBP_EINVAL, "sphere radius must be > 0");
        }
        for (int32_t i = 0; i < light_count; ++i) {
            bool fin = true;
            for (int a = 0; a < 3; ++a) fin = fin && std::isfinite(lights[i].position[a]) &&
                                             std::isfinite(lights[i].intensity[a]);
            UPBP_REQUIRE(fin, UPBP_ENONFINITE, "non-finite point light");
        }
        UPBP_REQUIRE(path_count == 0 || out->path_ends, UPBP_EINVAL, "out->path_ends is required");
        UPBP_REQUIRE(out->vertex_capacity >= 0 && out->beam_capacity >= 0, UPBP_EINVAL, "negative capacity");
        UPBP_REQUIRE(out->vertex_capacity == 0 || (out->vertices && out->positions), UPBP_EINVAL,
                     "out->vertices and out->positions are required");
        UPBP_REQUIRE(out->beam_capacity == 0 || (out->beam_origin && out->beam_direction && out->beam_length &&
                                                  out->beam_throughput && out->beam_flags && out->beam_medium &&
                                                  out->beam_mis),
                     UPBP_EINVAL, "every out->beam_* array is required");
        if (memory == UPBP_MEM_DEVICE) {
            trace_light_paths(*ctx, regions, region_count, lights, light_count, P, path_count, *out);
            return;
        }
        // host outputs: trace into device buffers of the same capacities, then copy what was made
        const int64_t cv = std::max<int64_t>(out->vertex_capacity, 1), cb = std::max<int64_t>(out->beam_capacity, 1);
        DevBuf v, x, pe, bo, bd, bl, bt, bf, bm, bs;
        v.ensure(cv * sizeof(upbp_light_vertex));
        x.ensure(cv * 12);
        pe.ensure(std::max<int64_t>(path_count, 1) * 4);
        bo.ensure(cb * 12);
        bd.ensure(cb * 12);
        bl.ensure(cb * 4);
        bt.ensure(cb * 12);
        bf.ensure(cb * 4);
        bm.ensure(cb * 4);
        bs.ensure(cb * sizeof(upbp_beam_mis));
        upbp_light_paths d = *out;
        d.vertices = v.as<upbp_light_vertex>();
        d.positions = x.as<float>();
        d.path_ends = pe.as<int32_t>();
        d.beam_origin = bo.as<float>();
        d.beam_direction = bd.as<float>();
        d.beam_length = bl.as<float>();
        d.beam_throughput = bt.as<float>();
        d.beam_flags = bf.as<uint32_t>();
        d.beam_medium = bm.as<int32_t>();
        d.beam_mis = bs.as<upbp_beam_mis>();
        struct CountsBack {
            upbp_light_paths& o;
            const upbp_light_paths& d;
            ~CountsBack() {
                o.vertex_count = d.vertex_count;
                o.beam_count = d.beam_count;
            }
        } counts_back{*out, d};
        trace_light_paths(*ctx, regions, region_count, lights, light_count, P, path_count, d);
        const int64_t V = d.vertex_count, B = d.beam_count;
        auto back = [&](void* dst, const DevBuf& src, size_t bytes) {
            if (bytes) UPBP_HIP(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
        };
        back(out->vertices, v, V * sizeof(upbp_light_vertex));
        back(out->positions, x, V * 12);
        back(out->path_ends, pe, path_count * 4);
        back(out->beam_origin, bo, B * 12);
        back(out->beam_direction, bd, B * 12);
        back(out->beam_length, bl, B * 4);
        back(out->beam_throughput, bt, B * 12);
        back(out->beam_flags, bf, B * 4);
        back(out->beam_medium, bm, B * 4);
        back(out->beam_mis, bs, B * sizeof(upbp_beam_mis));
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

upbp_status upbp_hit_signatures(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                                const upbp_query_batch* queries, uint32_t* photon_sig, uint32_t* beam_sig,
                                upbp_memory out_memory) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        RunOut o;
        o.sig_ph = bre ? photon_sig : nullptr;
        o.sig_bm = beams ? beam_sig : nullptr;
        run_eval(*ctx, photon_sig ? bre : nullptr, beam_sig ? beams : nullptr, queries, Mode::Signature, o,
                 out_memory, nullptr, nullptr, nullptr, true);
    });
}

upbp_status upbp_adjoint(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                         const upbp_query_batch* queries, const float* grad_out_photons, const float* grad_out_beams,
                         float* grad_scattering, float* grad_absorption, float* grad_photon_power,
                         float* grad_beam_power, upbp_memory memory, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        check_batch(*ctx, queries);
        const int64_t Q = queries->count;
        const int M = ctx->n_media;
        const bool do_ph = bre && grad_out_photons && bre->n > 0;
        const bool do_bm = beams && grad_out_beams && beams->pieces > 0;
        // upstream gradients on the device
        const float* gph = do_ph ? stage_q(*ctx, 4, grad_out_photons, 3 * Q, memory) : nullptr;
        const float* gbm = do_bm ? stage_q(*ctx, 5, grad_out_beams, 3 * Q, memory) : nullptr;
        DevBuf gmedia, items_ph, items_bm, gs_d, ga_d;
        gmedia.ensure((size_t)M * 6 * 4);
        launch_clear_f32(*ctx, gmedia.as<float>(), (int64_t)M * 6);
        float* dpp = grad_photon_power;
        float* dbp = grad_beam_power;
        if (memory == UPBP_MEM_HOST) {
            if (do_ph && grad_photon_power) {
                items_ph.ensure(bre->n_in * 12);
                launch_clear_f32(*ctx, items_ph.as<float>(), 3 * bre->n_in);
                dpp = items_ph.as<float>();
            }
            if (do_bm && grad_beam_power) {
                items_bm.ensure(beams->n * 12);
                launch_clear_f32(*ctx, items_bm.as<float>(), 3 * beams->n);
                dbp = items_bm.as<float>();
            }
        }
        // item gradients must have somewhere to go
        DevBuf dummy_ph, dummy_bm;
        if (do_ph && !dpp) {
            dummy_ph.ensure(bre->n_in * 12);
            dpp = dummy_ph.as<float>();
        }
        if (do_bm && !dbp) {
            dummy_bm.ensure(beams->n * 12);
            dbp = dummy_bm.as<float>();
        }
        AdjointOut aph{gph, gmedia.as<float>(), dpp};
        AdjointOut abm{gbm, gmedia.as<float>(), dbp};
        RunOut o;
        run_eval(*ctx, do_ph ? bre : nullptr, do_bm ? beams : nullptr, queries, Mode::Adjoint, o, UPBP_MEM_DEVICE,
                 stats, do_ph ? &aph : nullptr, do_bm ? &abm : nullptr, true);
        // combine d/d sigma_t into d/d sigma_s and d/d sigma_a, accumulate into the outputs; the
        // media use |sigma|, so each entry carries the sign of the caller's coefficient
        if (memory == UPBP_MEM_DEVICE) {
            launch_combine_media_grads(*ctx, gmedia.as<float>(), ctx->media_sign.as<float>(), M, grad_scattering,
                                       grad_absorption);
            UPBP_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            std::vector<float> gm((size_t)M * 6);
            UPBP_HIP(hipMemcpy(gm.data(), gmedia.p, gm.size() * 4, hipMemcpyDeviceToHost));
            const std::vector<float>& sg = ctx->media_sign_host;
            for (int m = 0; m < M; ++m)
                for (int c = 0; c < 3; ++c) {
                    if (grad_scattering)
                        grad_scattering[3 * m + c] += sg[6 * m + 3 + c] * (gm[6 * m + c] + gm[6 * m + 3 + c]);
                    if (grad_absorption) grad_absorption[3 * m + c] += sg[6 * m + c] * gm[6 * m + 3 + c];
                }
            if (do_ph && grad_photon_power) {
                std::vector<float> h((size_t)bre->n_in * 3);
                UPBP_HIP(hipMemcpy(h.data(), dpp, h.size() * 4, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < h.size(); ++i) grad_photon_power[i] += h[i];
            }
            if (do_bm && grad_beam_power) {
                std::vector<float> h((size_t)beams->n * 3);
                UPBP_HIP(hipMemcpy(h.data(), dbp, h.size() * 4, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < h.size(); ++i) grad_beam_power[i] += h[i];
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Copies every instance's photon and beam structures into the plan's merged arrays
// (device-to-device) and rebases their child references; records each instance's roots.
static void merge_structures(Ctx& ctx, upbp_batch_plan* plan, const upbp_instance* instances) {
    int64_t pn = 0, pi = 0, bn = 0, bi = 0;  // photon nodes / items, beam nodes / items
    for (int i = 0; i < plan->n; ++i) {
        const Bre* b = instances[i].bre;
        const Beams* m = instances[i].beams;
        // every instance's nodes start at the same offset as its items (n_nodes <= items), so
        // the item range that an inner node's own index gives stays right after rebasing
        if (plan->host[i].has_photons) {
            pn += std::max<int64_t>(b->bvh.n_nodes, b->n);
            pi += b->n;
        }
        if (plan->host[i].has_beams) {
            bn += std::max<int64_t>(m->bvh.n_nodes, m->pieces);
            bi += m->pieces;
        }
    }
    UPBP_REQUIRE(pn < 0x7FFFFFFF && pi < 0x7FFFFFFF && bn < 0x7FFFFFFF && bi < 0x7FFFFFFF, UPBP_EINVAL,
                 "batched structures too large");
    plan->m_ph_nodes.ensure(std::max<int64_t>(pn, 1) * sizeof(NodeDev));
    plan->m_ph_pos_r2.ensure(std::max<int64_t>(pi, 1) * 16);
    plan->m_ph_power.ensure(std::max<int64_t>(pi, 1) * 16);
    plan->m_ph_inc.ensure(std::max<int64_t>(pi, 1) * 16);
    plan->m_bm_nodes.ensure(std::max<int64_t>(bn, 1) * sizeof(NodeDev));
    plan->m_bm_pieces.ensure(std::max<int64_t>(bi, 1) * sizeof(PieceDev));
    plan->m_bm_flux.ensure(std::max<int64_t>(bi, 1) * 16);
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    pn = pi = bn = bi = 0;
    for (int i = 0; i < plan->n; ++i) {
        InstanceDev& d = plan->host[i];
        const Bre* b = instances[i].bre;
        const Beams* m = instances[i].beams;
        if (d.has_photons) {
            NodeDev* dst = plan->m_ph_nodes.as<NodeDev>() + pn;
            UPBP_HIP(hipMemcpyAsync(dst, b->bvh.nodes.p, b->bvh.n_nodes * sizeof(NodeDev), dd, ctx.stream));
            UPBP_HIP(hipMemcpyAsync(plan->m_ph_pos_r2.as<float4>() + pi, b->pos_r2.p, b->n * 16, dd, ctx.stream));
            UPBP_HIP(hipMemcpyAsync(plan->m_ph_power.as<float4>() + pi, b->power.p, b->n * 16, dd, ctx.stream));
            UPBP_HIP(hipMemcpyAsync(plan->m_ph_inc.as<float4>() + pi, b->inc.p, b->n * 16, dd, ctx.stream));
            launch_rebase_nodes(ctx, dst, b->bvh.n_nodes, (uint32_t)pn, (uint32_t)pi);
            d.ph_root = (uint32_t)pn;
            pn += std::max<int64_t>(b->bvh.n_nodes, b->n);
            pi += b->n;
        }
        if (d.has_beams) {
            NodeDev* dst = plan->m_bm_nodes.as<NodeDev>() + bn;
            UPBP_HIP(hipMemcpyAsync(dst, m->bvh.nodes.p, m->bvh.n_nodes * sizeof(NodeDev), dd, ctx.stream));
            UPBP_HIP(hipMemcpyAsync(plan->m_bm_pieces.as<PieceDev>() + bi, m->piece.p, m->pieces * sizeof(PieceDev), dd,
                                    ctx.stream));
            UPBP_HIP(hipMemcpyAsync(plan->m_bm_flux.as<float4>() + bi, m->piece_flux.p, m->pieces * 16, dd, ctx.stream));
            launch_rebase_nodes(ctx, dst, m->bvh.n_nodes, (uint32_t)bn, (uint32_t)bi);
            d.bm_root = (uint32_t)bn;
            bn += std::max<int64_t>(m->bvh.n_nodes, m->pieces);
            bi += m->pieces;
        }
    }
    plan->merged.ph_nodes = plan->m_ph_nodes.as<NodeDev>();
    plan->merged.ph_pos_r2 = plan->m_ph_pos_r2.as<float4>();
    plan->merged.ph_power = plan->m_ph_power.as<float4>();
    plan->merged.ph_inc = plan->m_ph_inc.as<float4>();
    plan->merged.bm_nodes = plan->m_bm_nodes.as<NodeDev>();
    plan->merged.bm_pieces = plan->m_bm_pieces.as<PieceDev>();
    plan->merged.bm_flux = plan->m_bm_flux.as<float4>();
}

upbp_status upbp_batch_plan_create(upbp_ctx* ctx, const upbp_instance* instances, int32_t count,
                                   upbp_batch_plan** out_plan) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    if (!out_plan) return fail(ctx, UPBP_EINVAL, "out_plan is NULL");
    *out_plan = nullptr;
    upbp_batch_plan* plan = new (std::nothrow) upbp_batch_plan();
    if (!plan) return fail(ctx, UPBP_ENOMEM, "host allocation failed");
    upbp_status st = guarded(ctx, [&] {
        UPBP_REQUIRE(count > 0 && instances, UPBP_EINVAL, "instances must be a non-empty array");
        plan->ctx = ctx;
        plan->n = count;
        plan->host.resize(count);
        plan->media_gen = ctx->media_gen;
        plan->media_base.resize(count);
        int64_t tiles = 0;
        for (int i = 0; i < count; ++i) {
            const upbp_instance& in = instances[i];
            check_batch(*ctx, &in.queries);
            UPBP_REQUIRE(in.queries.memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "batched queries must be device memory");
            UPBP_REQUIRE(!in.bre || in.out_photons, UPBP_EINVAL, "instance out_photons is NULL");
            UPBP_REQUIRE(!in.beams || in.out_beams, UPBP_EINVAL, "instance out_beams is NULL");
            UPBP_REQUIRE(!in.beams || in.beams->red.max == 0, UPBP_EUNSUPPORTED,
                         "per-cell beam reduction is supported by upbp_eval and upbp_hit_signatures only");
            UPBP_REQUIRE(in.media_base >= 0 && in.media_base < ctx->n_media, UPBP_EINVAL, "media_base out of range");
            plan->media_base[i] = in.media_base;
            const Bre* b = in.bre;
            const Beams* m = in.beams;
            InstanceDev& d = plan->host[i];
            std::memset(&d, 0, sizeof(d));
            d.has_photons = b && b->n > 0;
            d.has_beams = m && m->pieces > 0;
            if (d.has_photons) {
                d.ph_nodes = b->bvh.nodes.as<NodeDev>();
                d.ph_pos_r2 = b->pos_r2.as<float4>();
                d.ph_power = b->power.as<float4>();
                d.ph_inc = b->inc.as<float4>();
            }
            if (d.has_beams) {
                d.bm_nodes = m->bvh.nodes.as<NodeDev>();
                d.bm_pieces = m->piece.as<PieceDev>();
                d.bm_flux = m->piece_flux.as<float4>();
            }
            d.origin = in.queries.origin;
            d.direction = in.queries.direction;
            d.offsets = in.queries.segment_offsets;
            d.segments = in.queries.segments;
            d.count = in.queries.count;
            d.beam_type = in.queries.beam_type;
            d.lite = in.queries.lite_segments;
            d.media_base = in.media_base;
            d.out_photons = in.out_photons;
            d.out_beams = in.out_beams;
            d.tile_begin = tiles;
            if (in.queries.order == UPBP_ORDER_SORTED && in.queries.count > 64) {
                float lo[3], hi[3];
                scene_bounds(b, m, lo, hi);
                QueryView v;
                v.count = in.queries.count;
                v.origin = d.origin;
                v.direction = d.direction;
                v.offsets = d.offsets;
                v.segments = d.segments;
                DevBuf* perm = new DevBuf();
                plan->perms.push_back(perm);
                perm->ensure(v.count * 4);
                ctx->qkeys.ensure(v.count * 4);
                ctx->qkeys2.ensure(v.count * 4);
                ctx->qperm.ensure(v.count * 4);
                launch_query_keys(*ctx, v, lo, hi, ctx->qkeys.as<uint32_t>(), ctx->qperm.as<uint32_t>());
                sort_pairs_u32(*ctx, ctx->qkeys.as<uint32_t>(), ctx->qkeys2.as<uint32_t>(), ctx->qperm.as<uint32_t>(),
                               perm->as<uint32_t>(), v.count, 30);
                d.perm = perm->as<uint32_t>();
            }
            tiles += (in.queries.count + 63) / 64;
        }
        plan->total_tiles = tiles;
        // the 32-bit work counter also absorbs one extra increment per exiting wave
        UPBP_REQUIRE(tiles < (1ll << 31), UPBP_EINVAL, "too many query tiles");
        merge_structures(*ctx, plan, instances);
        plan->table.ensure(count * sizeof(InstanceDev));
        UPBP_HIP(hipMemcpyAsync(plan->table.p, plan->host.data(), count * sizeof(InstanceDev), hipMemcpyHostToDevice,
                                ctx->stream));
        plan->counters.ensure(16);
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != UPBP_OK) {
        delete plan;
        return st;
    }
    plan->ctx = ctx;
    ctx_retain(ctx);
    *out_plan = plan;
    return UPBP_OK;
}

upbp_status upbp_batch_plan_destroy(upbp_batch_plan* plan) {
    if (!plan) return UPBP_OK;
    Ctx* ctx = plan->ctx;
    (void)hipSetDevice(ctx->device);
    delete plan;
    ctx_release(ctx);
    return UPBP_OK;
}

// The photon kernel runs on the side stream (fork/join by events, which a graph
// capture records as edges), so the beam kernel's blocks fill the CUs the photon
// kernel frees as its queue drains.
static void enqueue_batch(Ctx& ctx, upbp_batch_plan* plan, int blocks) {
    launch_clear_u64(ctx, ctx.stats.as<uint64_t>(), kStatSlots);
    launch_fill_u32(ctx, plan->counters.as<uint32_t>(), 0u, 2);
    hipStream_t main_stream = ctx.stream;
    UPBP_HIP(hipEventRecord(ctx.fork, main_stream));
    UPBP_HIP(hipStreamWaitEvent(ctx.side, ctx.fork, 0));
    ctx.stream = ctx.side;
    try {
        launch_batched(ctx, plan->table.as<InstanceDev>(), plan->n, plan->total_tiles, plan->counters.as<uint32_t>(),
                       plan->merged, 0, blocks, ctx.tuning.waves_per_block);
        UPBP_HIP(hipEventRecord(ctx.join, ctx.side));
    } catch (...) {
        ctx.stream = main_stream;
        throw;
    }
    ctx.stream = main_stream;
    launch_batched(ctx, plan->table.as<InstanceDev>(), plan->n, plan->total_tiles, plan->counters.as<uint32_t>(),
                   plan->merged, 1, blocks, ctx.tuning.waves_per_block);
    UPBP_HIP(hipStreamWaitEvent(main_stream, ctx.join, 0));
}

upbp_status upbp_batch_run(upbp_ctx* ctx, upbp_batch_plan* plan, int use_graph, upbp_eval_stats* stats) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    if (!plan || plan->ctx != ctx) return fail(ctx, UPBP_EINVAL, "plan is NULL or belongs to another context");
    return guarded(ctx, [&] {
        // upbp_set_media since the last run: the instances' media ranges are checked again, and a
        // captured graph (which froze the old table's address and size) is dropped
        if (plan->media_gen != ctx->media_gen) {
            UPBP_REQUIRE(ctx->n_media > 0, UPBP_EINVAL, "upbp_set_media must be called before evaluation");
            for (int i = 0; i < plan->n; ++i)
                UPBP_REQUIRE(plan->media_base[i] < ctx->n_media, UPBP_EINVAL,
                             "an instance's media_base is outside the context's current media table");
            if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
            if (plan->graph) (void)hipGraphDestroy(plan->graph);
            plan->exec = nullptr;
            plan->graph = nullptr;
            plan->media_gen = ctx->media_gen;
        }
        // blocks_per_cu 0: each persistent kernel fills the GPU to its own occupancy limit
        const int per_cu = std::max(ctx->tuning.blocks_per_cu, 0);
        const int blocks = ctx->num_cus * per_cu;
        if (per_cu == 0) prepare_batched(*ctx, ctx->tuning.waves_per_block);
        if (stats) UPBP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
        if (use_graph) {
            UPBP_REQUIRE(ctx->stream != nullptr, UPBP_EINVAL, "graph capture needs a non-default stream");
            if (!plan->exec || plan->graph_blocks != blocks || plan->graph_wpb != ctx->tuning.waves_per_block) {
                if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
                if (plan->graph) (void)hipGraphDestroy(plan->graph);
                plan->exec = nullptr;
                plan->graph = nullptr;
                UPBP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_batch(*ctx, plan, blocks);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(ctx->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                UPBP_HIP(hipStreamEndCapture(ctx->stream, &plan->graph));
                UPBP_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
                plan->graph_blocks = blocks;
                plan->graph_wpb = ctx->tuning.waves_per_block;
            }
            UPBP_HIP(hipGraphLaunch(plan->exec, ctx->stream));
        } else {
            enqueue_batch(*ctx, plan, blocks);
        }
        if (stats) UPBP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
        if (!ctx->sync && !stats) return;
        uint64_t st[kStatSlots];
        UPBP_HIP(hipMemcpyAsync(st, ctx->stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
        UPBP_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        if (stats) UPBP_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        fill_stats(stats, st, ms, 0.f, 0.f, 2);
        check_stats_errors(*ctx, st);
    });
}

upbp_status upbp_set_tuning(upbp_ctx* ctx, const upbp_tuning* tuning) {
    if (!ctx || !tuning) return fail(ctx, UPBP_EINVAL, "ctx and tuning are required");
    if (tuning->waves_per_block < 1 || tuning->waves_per_block > 4 || tuning->blocks_per_cu < 0 ||
        tuning->blocks_per_cu > 64 || tuning->photon_waves_per_block < 0 || tuning->photon_waves_per_block > 4 ||
        tuning->frustum_leaf < 0 || tuning->frustum_leaf > 64)
        return fail(ctx, UPBP_EINVAL,
                    "waves_per_block must be 1..4, photon_waves_per_block 0..4, blocks_per_cu 0..64 and "
                    "frustum_leaf 0..64");
    ctx->tuning.waves_per_block = tuning->waves_per_block;
    ctx->tuning.blocks_per_cu = tuning->blocks_per_cu;
    ctx->tuning.photon_waves_per_block = tuning->photon_waves_per_block;
    ctx->tuning.frustum_leaf = tuning->frustum_leaf > 0 ? tuning->frustum_leaf : Tuning{}.frustum_leaf;
    return UPBP_OK;
}

upbp_status upbp_get_tuning(const upbp_ctx* ctx, upbp_tuning* tuning) {
    if (!ctx || !tuning) return UPBP_EINVAL;
    tuning->waves_per_block = ctx->tuning.waves_per_block;
    tuning->blocks_per_cu = ctx->tuning.blocks_per_cu;
    tuning->photon_waves_per_block = ctx->tuning.photon_waves_per_block;
    tuning->frustum_leaf = ctx->tuning.frustum_leaf;
    return UPBP_OK;
}

upbp_status upbp_autotune(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams, const upbp_query_batch* queries,
                          upbp_tuning* best, float* best_ms) {
    if (!ctx) return fail(nullptr, UPBP_EINVAL, "ctx is NULL");
    return guarded(ctx, [&] {
        check_batch(*ctx, queries);
        UPBP_REQUIRE(queries->memory == UPBP_MEM_DEVICE, UPBP_EINVAL, "autotune needs device-resident queries");
        const int64_t Q = queries->count;
        DevBuf o1, o2;
        o1.ensure(std::max<int64_t>(Q, 1) * 12);
        o2.ensure(std::max<int64_t>(Q, 1) * 12);
        Tuning saved = ctx->tuning;
        float best_t = INFINITY;
        Tuning best_cfg = saved;
        const int cand[3] = {1, 2, 4};
        for (int wpb : cand) {
            ctx->tuning.waves_per_block = wpb;
            RunOut o;
            o.out_ph = bre ? o1.as<float>() : nullptr;
            o.out_bm = beams ? o2.as<float>() : nullptr;
            upbp_eval_stats s;
            run_eval(*ctx, bre, beams, queries, Mode::Forward, o, UPBP_MEM_DEVICE, &s, nullptr, nullptr, true);  // warm
            float tsum = 0.f;
            for (int rep = 0; rep < 3; ++rep) {
                run_eval(*ctx, bre, beams, queries, Mode::Forward, o, UPBP_MEM_DEVICE, &s, nullptr, nullptr, true);
                tsum += s.kernel_ms_photons + s.kernel_ms_beams;
            }
            if (tsum < best_t) {
                best_t = tsum;
                best_cfg = ctx->tuning;
            }
        }
        ctx->tuning = best_cfg;
        if (best) {
            best->waves_per_block = best_cfg.waves_per_block;
            best->blocks_per_cu = best_cfg.blocks_per_cu;
            best->photon_waves_per_block = best_cfg.photon_waves_per_block;
            best->frustum_leaf = best_cfg.frustum_leaf;
        }
        if (best_ms) *best_ms = best_t / 3.f;
    });
}

}  // extern "C"
