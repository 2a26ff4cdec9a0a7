// upbp_internal.hpp -- host-side objects behind the opaque handles of include/upbp.h.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/upbp.h"
#include "upbp_device.hpp"
#include "beamgrid.hpp"

namespace upbp {

// Error carried from a failing HIP call up to the C ABI.
struct Error {
    upbp_status status;
    std::string message;
};

#define UPBP_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            throw ::upbp::Error{e_ == hipErrorOutOfMemory ? UPBP_ENOMEM : UPBP_EDEVICE,      \
                                std::string(#call) + ": " + hipGetErrorString(e_)};         \
        }                                                                                    \
    } while (0)

#define UPBP_REQUIRE(cond, status, msg)                  \
    do {                                                 \
        if (!(cond)) throw ::upbp::Error{(status), (msg)}; \
    } while (0)

// Owning device buffer (hipMalloc / hipFree), grows on demand.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        release();
        size_t want = n < 256 ? 256 : n;
        UPBP_HIP(hipMalloc(&p, want));
        bytes = want;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Build temporaries: grow-only device buffers that a context keeps and lends out for one scope
// at a time, last borrowed first returned (every call on a context runs on its one stream, and
// a context is used by one host thread at a time). A renderer that rebuilds its structures every
// iteration then allocates and frees no temporary device memory once the sizes are reached; the
// buffers are freed with the context.
struct TmpPool {
    static constexpr int kSlots = 48;
    DevBuf slot[kSlots];
    int top = 0;
};

class TmpBuf {
  public:
    explicit TmpBuf(TmpPool& pool) : pool_(pool), b_(take(pool)), p(b_.p) {}
    ~TmpBuf() { --pool_.top; }
    TmpBuf(const TmpBuf&) = delete;
    TmpBuf& operator=(const TmpBuf&) = delete;
    void ensure(size_t n) { b_.ensure(n); }
    template <class T>
    T* as() const { return b_.as<T>(); }
    operator DevBuf&() { return b_; }

  private:
    static DevBuf& take(TmpPool& pool) {
        UPBP_REQUIRE(pool.top < TmpPool::kSlots, UPBP_EDEVICE, "build temporaries exhausted");
        return pool.slot[pool.top++];
    }
    TmpPool& pool_;
    DevBuf& b_;

  public:
    void*& p;  // the lent buffer's pointer (follows ensure)
};

// Counters written by the traversal kernels (device, uint64 each).
enum StatSlot : int {
    kStatPhotonHits = 0,
    kStatBeamHits = 1,
    kStatPhotonCands = 2,
    kStatBeamCands = 3,
    kStatNodes = 4,
    kStatError = 5,     // bit mask of input errors found by kernels
    kStatOverflow = 6,  // traversal stack overflow count
    kStatSlots = 8
};
constexpr uint32_t kErrBadOffsets = 1u, kErrBadMedium = 2u, kErrBadSegment = 4u, kErrNonFinite = 8u;

struct Tuning {
    int waves_per_block = 4;
    int blocks_per_cu = 0;
    int photon_waves_per_block = 0;  // 0: waves_per_block
    int frustum_leaf = 64;           // largest subtree a frustum wave tests as one leaf
};

struct Ctx {
    // References: the caller's (dropped by upbp_ctx_destroy) plus one per photon
    // structure, beam structure and batch plan built on this context.
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool sync = true;
    std::string last_error;
    int n_media = 0;
    std::vector<MediumDev> media_host;
    DevBuf media;
    // sign of every input coefficient (+1 / -1) as [M][6] floats (absorption RGB, scattering RGB):
    // the media use |sigma| (Media.hxx:136-140), so gradients with respect to the caller's
    // values carry these factors
    DevBuf media_sign;
    std::vector<float> media_sign_host;
    // bumped by every upbp_set_media: a batched plan whose captured graph froze an older table
    // re-validates its instances and captures again
    uint64_t media_gen = 0;
    DevBuf stats;       // kStatSlots x uint64
    DevBuf sort_tmp;    // hipcub temporary storage
    TmpPool tmp;        // build temporaries
    DevBuf qkeys, qkeys2, qperm, qperm2;
    DevBuf light_tails;  // float4 per light vertex of the current UPBP call
    DevBuf beam_recs;    // 64-byte MIS record per photon beam of the current UPBP call
    DevBuf photon_recs;  // 64-byte MIS record per light vertex of the current UPBP call
    DevBuf vm_rec;       // merge records (4 float4 per grid slot) of the current PP3D / SURF call
    DevBuf lt_slots[8];  // light tracing: per-path vertex, position and beam slots, counts, offsets, lights
    DevBuf staging[8];  // host-input staging copies
    hipEvent_t ev[6] = {};
    // second stream: the photon and beam kernels of one eval call run concurrently
    // (fork from `stream`, join back), so one kernel's tail overlaps the other's work
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // resident blocks per CU of the two persistent batched kernels at batched_wpb waves per block
    int batched_wpb = 0;
    int batched_per_cu[2] = {0, 0};
    int num_cus = 256;
    Tuning tuning;
    // this device's copy of the allocation-free entry point's fixed region (null: lookup failed)
    void* scratch = nullptr;
    std::string scratch_error;
};

// A built LBVH: node array plus the item order it implies.
struct Lbvh {
    DevBuf nodes;
    int64_t n_nodes = 0;
    int64_t n_items = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of all item boxes
};

struct Bre {
    Ctx* ctx = nullptr;
    int64_t n_in = 0;  // photons given to build
    int64_t n = 0;     // kept photons
    Lbvh bvh;
    DevBuf pos_r2;  // float4 [n]
    DevBuf power;   // float4 [n]
    DevBuf inc;     // float4 [n]
    DevBuf orig;    // uint32 [n] original input index
    DevBuf radius_in;  // float [n_in] radius the build gave each input photon (0 if filtered out)
};

struct Beams {
    Ctx* ctx = nullptr;
    int64_t n = 0;        // beams
    int64_t pieces = 0;   // beam pieces in the LBVH
    Lbvh bvh;
    DevBuf piece;         // PieceDev [pieces]
    DevBuf piece_flux;    // float4 [pieces] throughput (rgb, 0)
    DevBuf piece_beam;    // uint32 [pieces] original beam index
    DevBuf r_start, r_end;  // float [n] start / end radius the build gave each beam
    // per-cell beam reduction (max_beams_in_cell > 0): the reference grid's cell lists
    RedDev red;            // device view; red.max == 0 without reduction
    int64_t red_cells = 0;
    DevBuf red_off, red_ids, red_rank;
    // the grid's box top and cell size (with red_cells > 0), and the resolution the build was
    // given (0 = 256): what a walk of the grid needs besides RedDev
    float grid_hi[3] = {0.f, 0.f, 0.f}, grid_cell[3] = {0.f, 0.f, 0.f};
    int32_t grid_resolution = 0;
};

// Device memory one evaluation works in besides its inputs and outputs: the kernels' counters,
// the Morton keys and permutations of the query sort and the sort's temporary storage. The
// context's own growable buffers by default; the allocation-free entry point passes a fixed
// region instead.
struct Workspace {
    unsigned long long* stats = nullptr;  // kStatSlots counters
    uint32_t *keys = nullptr, *keys2 = nullptr, *perm = nullptr, *perm2 = nullptr;  // [max_sorted] each
    int64_t max_sorted = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
};

// Device-resident view of a query batch.
struct QueryView {
    int64_t count = 0;
    const float* origin = nullptr;
    const float* direction = nullptr;
    const int64_t* offsets = nullptr;
    const upbp_volume_segment* segments = nullptr;
    uint32_t beam_type = UPBP_LONG_BEAM;
    uint32_t lite = 0;
    const uint32_t* perm = nullptr;  // lane slot -> query index, or null (identity)
};

// ---- lbvh.hip -------------------------------------------------------------
// Builds an LBVH over n item boxes (lo/hi as float4 .xyz). On return
// out_order[k] is the input index of the k-th item in leaf order.
void build_lbvh(Ctx& ctx, int64_t n, const float4* box_lo, const float4* box_hi, int leaf_size,
                Lbvh& out, DevBuf& out_order);
void sort_pairs_u64(Ctx& ctx, uint64_t* keys_in, uint64_t* keys_out, uint32_t* vals_in,
                    uint32_t* vals_out, int64_t n, int end_bit);
void sort_pairs_u32(Ctx& ctx, uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_in,
                    uint32_t* vals_out, int64_t n, int end_bit);
void exclusive_scan_u32(Ctx& ctx, const uint32_t* in, uint32_t* out, int64_t n);
// The same sort in caller-provided temporary storage (UPBP_EDEVICE if tmp_bytes is too small);
// sort_pairs_u32_temp_bytes is what it needs for n keys (no device work, no allocation).
size_t sort_pairs_u32_temp_bytes(int64_t n, int end_bit);
void sort_pairs_u32_in(Ctx& ctx, void* tmp, size_t tmp_bytes, uint32_t* keys_in, uint32_t* keys_out,
                       uint32_t* vals_in, uint32_t* vals_out, int64_t n, int end_bit);

// ---- knn.hip --------------------------------------------------------------
// Point LBVH answering "squared distance to the k-th nearest point" queries.
struct KnnIndex {
    int64_t n = 0;
    Lbvh bvh;
    DevBuf points;  // float4 [n] in leaf order
    // points: device xyz [*][3]; keep: optional device index list of the points to use
    void build(Ctx& ctx, int64_t n_points, const float* pts, const uint32_t* keep);
    // out[i] (device) = k-th smallest squared distance from q[i] (device xyz) to the points;
    // with n <= k the largest of the n.
    void kth_d2(Ctx& ctx, int64_t nq, const float* q, int k, float* out);
};
// KNN_RADIUS of EmbreeBre::build: out_radius[i] (device, [n_in]) = radius * 2 sqrt(d2_k)
// for the kept photons (keep: kept -> input index, or null for all).
void photon_knn_radius(Ctx& ctx, int64_t n_kept, const float* pos, const uint32_t* keep, float radius, int knn,
                       DevBuf& out_radius, int64_t n_in);
// KNN_RADIUS of PhotonBeamsEvaluator::build: start / end radius per beam (device).
void beam_knn_radius(Ctx& ctx, int64_t n, const float* origin, const float* dir, const float* length, float radius,
                     int knn, DevBuf& r0, DevBuf& r1);

// ---- build.hip ------------------------------------------------------------
void bre_build(Ctx& ctx, const upbp_photon_input& in, const upbp_build_params& params, Bre& out);
void bre_set_power(Bre& bre, const float* power_dev);
void beams_build(Ctx& ctx, const upbp_beam_input& in, const upbp_build_params& params, Beams& out);

// ---- beamgrid.hip ---------------------------------------------------------
// Registers the beams (device arrays, input order) in the reference's uniform grid and
// keeps each cell's beam list and reduction ranks in out.red*; a no-op without reduction.
// With `always`, the grid is built even without reduction (red.max stays 0).
void beam_reduction_build(Ctx& ctx, int64_t n, const float* origin, const float* dir, const float* length,
                          const float* r_start, const float* r_end, const upbp_build_params& params, Beams& out,
                          bool always = false);
// Per-query GridStats of the reference grid's walk (beams tested, cells crossed, overfull
// cells crossed) into out [q.count][3] on the device (query order as given). Beams built
// without reduction get a grid of their build's resolution for the call.
void beam_grid_stats(Ctx& ctx, const Beams& b, const QueryView& q, uint32_t* out);
// Selection pdf of the cell holding pos (1 without reduction).
float beam_selection_pdf(const Beams& b, const float pos[3]);
// The same for n positions (xyz [n][3]) into out [n], both in `mem`.
void beam_selection_pdfs(const Beams& b, const float* pos, int64_t n, float* out, upbp_memory mem);

// ---- traverse.hip ---------------------------------------------------------
// EmbreeBre::evalBreBruteForce per query (device arrays; out [count][3]).
void bre_eval_brute_force(Ctx& ctx, const Bre& b, const float* origin, const float* direction, const float* min_t,
                          const float* max_t, const int32_t* medium, int64_t count, float* out);

enum class Mode : int { Forward = 0, Signature = 1, Adjoint = 2, Upbp = 3 };

struct PhotonArgs {
    const NodeDev* nodes;
    const float4* pos_r2;
    const float4* power;
    const float4* inc;
    const uint32_t* orig;
    int64_t n;
    // UPBP mode: the reduction grid of the photon-beam evaluator of the same call (sel.max == 0:
    // none), whose cell pdf at a PB2D hit is the camera vertex's BB1D beam selection pdf
    RedDev sel;
};
struct BeamArgs {
    const NodeDev* nodes;
    const PieceDev* pieces;
    const float4* piece_flux;
    const uint32_t* piece_beam;
    int64_t pieces_n;
    RedDev red;
};
struct AdjointOut {
    const float* grad_out;  // [Q][3]
    float* grad_media;      // [M][6]: dsigma_s direct (0..2), dsigma_t (3..5)
    float* grad_items;      // photon [n_in][3] or beam [n][3]
};

// UPBP mode inputs on the device (upbp_mis_inputs after staging).
struct MisDev {
    const upbp_query_mis* queries;
    const upbp_mis_data* camera;
    int64_t camera_count;
    const upbp_light_vertex* lv;
    int64_t lv_count;
    const int32_t* path_ends;
    int64_t path_count;
    const upbp_beam_mis* beams;
    upbp_mis_params params;
    const float4* tails = nullptr;      // per light vertex, launch_light_tails
    const void* beam_recs = nullptr;    // per photon beam, launch_beam_records
    const void* photon_recs = nullptr;  // per light vertex, launch_photon_records
};

// stats: the kStatSlots counters the kernels add to (null: the context's own)
void launch_photons(Ctx& ctx, Mode mode, const PhotonArgs& ph, const QueryView& q, float* out,
                    uint32_t* sig, const AdjointOut* adj, const MisDev* mis = nullptr,
                    unsigned long long* stats = nullptr);
void launch_beams(Ctx& ctx, Mode mode, const BeamArgs& bm, const QueryView& q, float* out,
                  uint32_t* sig, const AdjointOut* adj, const MisDev* mis = nullptr,
                  unsigned long long* stats = nullptr);
void launch_query_keys(Ctx& ctx, const QueryView& q, const float* lo, const float* hi, uint32_t* keys,
                       uint32_t* idx);
// Folds each light vertex's light walk below it into a float4 (the O(1) light walk of the
// UPBP weights) into ctx.light_tails and points mis.tails at it; nothing for no vertices.
void launch_light_tails(Ctx& ctx, MisDev& mis);
// Gathers each photon beam's MIS inputs (its record, origin light vertex and that vertex's tail)
// into ctx.beam_recs and points mis.beam_recs at it; after launch_light_tails.
void launch_beam_records(Ctx& ctx, MisDev& mis, int64_t n_beams);
// The same per light vertex for the PB2D hits into ctx.photon_recs (mis.photon_recs).
void launch_photon_records(Ctx& ctx, MisDev& mis);
void launch_clear_u64(Ctx& ctx, uint64_t* p, int64_t n);
void launch_clear_f32(Ctx& ctx, float* p, int64_t n);
// gs += sign_s * (direct + d/d sigma_t), ga += sign_a * d/d sigma_t per medium and channel
// (sign: [M][6] of ctx.media_sign).
void launch_combine_media_grads(Ctx& ctx, const float* gm, const float* sign, int n_media, float* gs, float* ga);

// ---- hashgrid.hip: vertex-merging hashed grid (PP3D, SURF) --------------------
struct HashGrid {
    Ctx* ctx = nullptr;
    uint32_t type = 0;  // UPBP_PP3D (in-medium vertices) or UPBP_SURF (surface vertices)
    int64_t n_src = 0;  // light vertices given to the build
    int64_t n = 0;      // vertices in the grid
    int64_t cells = 0;  // hash buckets
    float radius = 0.f, r2 = 0.f, inv_cell = 0.f;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    DevBuf slots;  // float4 [n]: position, light-vertex index (bits); bucket order
    DevBuf ends;   // uint32 [cells]: one past each bucket's last slot
};
// flags[i * stride] holds the UPBP_LV_* word of light vertex i (device pointers).
void hash_grid_build(Ctx& ctx, const float* pos, const uint32_t* flags, int64_t stride, int64_t n, float radius,
                     int64_t cells, uint32_t type, HashGrid& out);
void pp3d_eval(Ctx& ctx, const HashGrid& grid, const upbp_pp3d_query* q, int64_t count, const MisDev& mis,
               uint32_t flags, float* out, uint32_t* sig);
void surf_eval(Ctx& ctx, const HashGrid& grid, const upbp_surf_query* q, int64_t count, const MisDev& mis,
               uint32_t flags, float* out, uint32_t* sig);

// ---- batched persistent kernels --------------------------------------------
struct InstanceDev {
    const NodeDev* ph_nodes;
    const float4* ph_pos_r2;
    const float4* ph_power;
    const float4* ph_inc;
    const NodeDev* bm_nodes;
    const PieceDev* bm_pieces;
    const float4* bm_flux;
    const float* origin;
    const float* direction;
    const int64_t* offsets;
    const upbp_volume_segment* segments;
    const uint32_t* perm;
    float* out_photons;
    float* out_beams;
    int64_t count;
    int64_t tile_begin;  // first global tile of this instance
    int32_t media_base;
    uint32_t beam_type;
    uint32_t lite;
    int32_t has_photons;
    int32_t has_beams;
    uint32_t ph_root;  // root of this instance's photon tree in the plan's merged node array
    uint32_t bm_root;  // root of its beam tree in the merged beam node array
    int32_t pad;
};

// A batched plan's copy of every instance's structures, one array each, so that the
// persistent kernel reads them through kernel-argument pointers (scalar loads).
// Child references are rebased: inner refs by the instance's node offset, leaf refs
// by its item offset.
struct MergedDev {
    const NodeDev* ph_nodes;
    const float4* ph_pos_r2;
    const float4* ph_power;
    const float4* ph_inc;
    const NodeDev* bm_nodes;
    const PieceDev* bm_pieces;
    const float4* bm_flux;
};
void launch_rebase_nodes(Ctx& ctx, NodeDev* nodes, int64_t n, uint32_t node_off, uint32_t item_off);

// Queries (outside any graph capture) how many blocks of each persistent kernel fit on a CU.
void prepare_batched(Ctx& ctx, int waves_per_block);
void launch_batched(Ctx& ctx, const InstanceDev* inst, int n_inst, int64_t total_tiles, uint32_t* counters,
                    const MergedDev& merged, int which, int grid_blocks, int waves_per_block);
void launch_fill_u32(Ctx& ctx, uint32_t* p, uint32_t v, int64_t n);

// ---- lighttrace.hip: light sub-path tracing in media ---------------------------
// Fills out (device memory) and its counts; UPBP_EINVAL when a capacity is exceeded.
void trace_light_paths(Ctx& ctx, const upbp_region* regions, int32_t n_regions, const upbp_point_light* lights,
                       int32_t n_lights, const upbp_light_trace_params& P, int64_t path_count,
                       upbp_light_paths& out);

}  // namespace upbp
