// upbp.hpp -- header-only C++ facade over the C ABI (upbp.h) with the method
// shapes of the reference estimator classes, so renderer code can switch with
// few edits:
//   upbp::EmbreeBre            <-> EmbreeBre            (SmallUPBP/src/Bre/Bre.hxx:40-184)
//   upbp::PhotonBeamsEvaluator <-> PhotonBeamsEvaluator (SmallUPBP/src/Beams/PhBeams.hxx:46-160)
//   upbp::HashGrid             <-> HashGrid + RangeQuery for PP3D and SURF merging
//                                  (SmallUPBP/src/Misc/HashGrid.hxx:38-230, SmallUPBP/src/Renderers/UPBP.hxx:63-133)
//
// Two ways to evaluate:
//   evalBre(beamType, origin, dir, segments, n)  one query, synchronous (the reference call shape;
//                                                correct but one GPU round trip per call);
//   QueryBatch + evalBre(batch, ...)             many queries per launch (the throughput path):
//                                                add() every camera ray of an iteration, then
//                                                evaluate once. Radiance is linear, so the
//                                                renderer's per-ray weighting (e.g. UPBP.hxx:890-891)
//                                                is applied after the batch returns.
// Errors throw upbp::Error (status + message); nothing calls exit().
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "upbp.h"

namespace upbp {

struct Error : std::runtime_error {
    upbp_status status;
    Error(upbp_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(upbp_status s, const upbp_ctx* ctx) {
    if (s != UPBP_OK) throw Error(s, upbp_ctx_last_error(ctx));
}

// One GPU context (one per renderer thread, like one EmbreeBre per OpenMP thread).
class Context {
public:
    explicit Context(int device = 0) { check(upbp_ctx_create(device, &h_), nullptr); }
    ~Context() { upbp_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    // Replaces scene.mMedia lookups (Bre.cxx:836, PhBeams.cxx:201): media indexed by medium id.
    void setMedia(const std::vector<upbp_medium>& media) {
        check(upbp_set_media(h_, media.data(), (int32_t)media.size()), h_);
    }
    upbp_ctx* get() const { return h_; }

private:
    upbp_ctx* h_ = nullptr;
};

// Host-side accumulation of query rays and their volume segments (CSR).
class QueryBatch {
public:
    // Adds one query (a camera ray and its VolumeSegments); returns its index in the batch.
    size_t add(const float origin[3], const float direction[3], const upbp_volume_segment* segments, size_t n) {
        for (int a = 0; a < 3; ++a) {
            o_.push_back(origin[a]);
            d_.push_back(direction[a]);
        }
        seg_.insert(seg_.end(), segments, segments + n);
        off_.push_back((int64_t)seg_.size());
        return off_.size() - 2;
    }
    size_t size() const { return off_.size() - 1; }
    void clear() {
        o_.clear();
        d_.clear();
        seg_.clear();
        off_.assign(1, 0);
    }
    upbp_query_batch view(uint32_t beamType, bool lite = false, uint32_t order = UPBP_ORDER_SORTED) const {
        upbp_query_batch q{};
        q.count = (int64_t)size();
        q.origin = o_.data();
        q.direction = d_.data();
        q.segment_offsets = off_.data();
        q.segments = seg_.data();
        q.beam_type = beamType;
        q.lite_segments = lite ? 1u : 0u;
        q.order = order;
        q.memory = UPBP_MEM_HOST;
        return q;
    }

private:
    std::vector<float> o_, d_;
    std::vector<int64_t> off_{0};
    std::vector<upbp_volume_segment> seg_;
};

class EmbreeBre {
public:
    explicit EmbreeBre(Context& ctx) : ctx_(ctx) {}
    ~EmbreeBre() { destroy(); }
    EmbreeBre(const EmbreeBre&) = delete;
    EmbreeBre& operator=(const EmbreeBre&) = delete;

    // EmbreeBre::build (Bre.cxx:526-632): photons are light vertices with in_medium != 0
    // (mIsInMedium); position = mHitpoint, incDir = mBSDF.WorldDirFix(), power = mThroughput.
    // Returns the number of photons kept.
    int build(const float* position, const float* incDir, const float* power, const uint8_t* inMedium,
              int numVertices, int radiusCalculation, float photonRadius, int knn, bool verbose = false) {
        (void)verbose;
        destroy();
        upbp_photon_input in{};
        in.count = numVertices;
        in.position = position;
        in.inc_dir = incDir;
        in.power = power;
        in.in_medium = inMedium;
        in.memory = UPBP_MEM_HOST;
        upbp_build_params p{};
        p.radius_calculation = radiusCalculation;
        p.radius = photonRadius;
        p.knn = knn;
        int64_t kept = 0;
        check(upbp_bre_build(ctx_.get(), &in, &p, &h_, &kept), ctx_.get());
        return (int)kept;
    }

    // EmbreeBre::evalBre for one query (Bre.cxx:812-879; lite = VolumeSegmentLite overload, :895-964).
    std::array<float, 3> evalBre(uint32_t beamType, const float origin[3], const float direction[3],
                                 const upbp_volume_segment* segments, size_t n, bool lite = false) {
        QueryBatch b;
        b.add(origin, direction, segments, n);
        std::array<float, 3> out{0.f, 0.f, 0.f};
        evalBre(b, beamType, out.data(), lite);
        return out;
    }

    // Batched evaluation: outRgb receives [batch.size()][3] floats (host memory).
    void evalBre(const QueryBatch& batch, uint32_t beamType, float* outRgb, bool lite = false,
                 upbp_eval_stats* stats = nullptr) {
        upbp_query_batch q = batch.view(beamType, lite);
        check(upbp_eval(ctx_.get(), h_, nullptr, &q, outRgb, nullptr, UPBP_MEM_HOST, stats), ctx_.get());
    }

    // EmbreeBre::evalBreBruteForce (Bre.cxx:980-1029): every photon against [minT, maxT] of one
    // ray in medium mediumId, with the reference's debug kernel. beamType is unused there too.
    std::array<float, 3> evalBreBruteForce(uint32_t beamType, const float origin[3], const float direction[3],
                                           float minT, float maxT, int32_t mediumId) {
        (void)beamType;
        std::array<float, 3> out{0.f, 0.f, 0.f};
        check(upbp_bre_eval_brute_force(ctx_.get(), h_, origin, direction, &minT, &maxT, &mediumId, 1, out.data(),
                                        UPBP_MEM_HOST),
              ctx_.get());
        return out;
    }

    // EmbreeBre::destroy (Bre.cxx:762-777); also run by the destructor.
    void destroy() {
        if (h_) upbp_bre_destroy(h_);
        h_ = nullptr;
    }
    upbp_bre* get() const { return h_; }

private:
    Context& ctx_;
    upbp_bre* h_ = nullptr;
};

class PhotonBeamsEvaluator {
public:
    explicit PhotonBeamsEvaluator(Context& ctx) : ctx_(ctx) {}
    ~PhotonBeamsEvaluator() { destroy(); }
    PhotonBeamsEvaluator(const PhotonBeamsEvaluator&) = delete;
    PhotonBeamsEvaluator& operator=(const PhotonBeamsEvaluator&) = delete;

    // PhotonBeamsEvaluator::build (PhBeams.cxx:66-148) over beams in structure-of-arrays form:
    // mRay.origin, mRay.direction, mLength, mThroughputAtOrigin, mFlags and the medium index.
    // maxBeamsInCell / reductionType / gridSize / seed: the static sMaxBeamsInCell,
    // sReductionType and the grid settings of the reference (PhBeams.cxx:51-52,126-136).
    void build(const upbp_beam_input& beams, int radiusCalculation, float beamRadius, int knn,
               bool verbose = false, uint32_t maxBeamsInCell = 0, uint32_t reductionType = 0, int gridSize = 256,
               int seed = 1234) {
        (void)verbose;
        destroy();
        upbp_build_params p{};
        p.radius_calculation = radiusCalculation;
        p.radius = beamRadius;
        p.knn = knn;
        p.max_beams_in_cell = maxBeamsInCell;
        p.reduction_type = reductionType;
        p.grid_resolution = gridSize;
        p.reduction_seed = seed;
        check(upbp_beams_build(ctx_.get(), &beams, &p, &h_), ctx_.get());
    }

    std::array<float, 3> evalBeamBeamEstimate(uint32_t beamType, const float origin[3], const float direction[3],
                                              const upbp_volume_segment* segments, size_t n, bool lite = false) {
        QueryBatch b;
        b.add(origin, direction, segments, n);
        std::array<float, 3> out{0.f, 0.f, 0.f};
        evalBeamBeamEstimate(b, beamType, out.data(), lite);
        return out;
    }

    void evalBeamBeamEstimate(const QueryBatch& batch, uint32_t beamType, float* outRgb, bool lite = false,
                              upbp_eval_stats* stats = nullptr) {
        upbp_query_batch q = batch.view(beamType, lite);
        check(upbp_eval(ctx_.get(), nullptr, h_, &q, nullptr, outRgb, UPBP_MEM_HOST, stats), ctx_.get());
    }

    // GridStats per query of the reference grid's walk (GridStats.hxx:37-54, Grid.hxx:318-399,
    // PhBeams.cxx:177-230): outStats [n][3] = intersectedBeams, intersectedCells, overfullCells.
    void gridStats(const QueryBatch& batch, uint32_t beamType, uint32_t* outStats, bool lite = false) {
        upbp_query_batch q = batch.view(beamType, lite);
        check(upbp_beams_grid_stats(ctx_.get(), h_, &q, outStats, UPBP_MEM_HOST), ctx_.get());
    }

    // PhotonBeamsEvaluator::getBeamSelectionPdf (1 without per-cell beam reduction).
    float getBeamSelectionPdf(const float position[3]) const { return upbp_beams_selection_pdf(h_, position); }
    // The same for n positions ([n][3] -> [n]) in one launch.
    void getBeamSelectionPdfs(const float* positions, int64_t n, float* out, upbp_memory mem = UPBP_MEM_HOST) const {
        check(upbp_beams_selection_pdfs(h_, positions, n, out, mem), ctx_.get());
    }

    void destroy() {
        if (h_) upbp_beams_destroy(h_);
        h_ = nullptr;
    }
    upbp_beams* get() const { return h_; }

private:
    Context& ctx_;
    upbp_beams* h_ = nullptr;
};

// Vertex merging: mPP3DHashGrid / mSurfHashGrid (UPBP.hxx:689-703) and the RangeQuery
// gathers (UPBP.hxx:1097-1120) for a batch of camera vertices.
class HashGrid {
public:
    explicit HashGrid(Context& ctx) : ctx_(ctx) {}
    ~HashGrid() { destroy(); }
    HashGrid(const HashGrid&) = delete;
    HashGrid& operator=(const HashGrid&) = delete;

    // HashGrid::Reserve (HashGrid.hxx:41-44): number of hash buckets.
    void Reserve(int64_t cellCount) { cells_ = cellCount; }
    // HashGrid::Build(mLightVertices, radius, type) (HashGrid.hxx:46-128), type UPBP_PP3D or
    // UPBP_SURF: positions [n][3] (mHitpoint) and the light-vertex records (their flags select
    // the grid's vertices as MatchesType does).
    void Build(const float* positions, const upbp_light_vertex* vertices, int64_t n, float radius, uint32_t type,
               upbp_memory memory = UPBP_MEM_HOST) {
        destroy();
        check(upbp_hash_grid_build(ctx_.get(), positions, vertices, n, radius, cells_, type, memory, &h_),
              ctx_.get());
    }
    // RangeQuery::GetContrib() per camera vertex, before the caller's cameraState.mThroughput *
    // mPP3DNormalization (UPBP.hxx:1117) or mSurfNormalization (:1104); ppm = mAlgorithm == kPPM.
    void Process(const upbp_pp3d_query* queries, int64_t n, const upbp_mis_inputs& mis, bool ppm, float* outRgb,
                 upbp_memory memory = UPBP_MEM_HOST, upbp_eval_stats* stats = nullptr) {
        check(upbp_pp3d_eval(ctx_.get(), h_, queries, n, &mis, ppm ? UPBP_VM_NO_MIS : 0u, outRgb, nullptr, memory,
                             stats),
              ctx_.get());
    }
    void Process(const upbp_surf_query* queries, int64_t n, const upbp_mis_inputs& mis, bool ppm, float* outRgb,
                 upbp_memory memory = UPBP_MEM_HOST, upbp_eval_stats* stats = nullptr) {
        check(upbp_surf_eval(ctx_.get(), h_, queries, n, &mis, ppm ? UPBP_VM_NO_MIS : 0u, outRgb, nullptr, memory,
                             stats),
              ctx_.get());
    }
    int64_t count() const { return upbp_hash_grid_count(h_); }
    void destroy() {
        if (h_) upbp_hash_grid_destroy(h_);
        h_ = nullptr;
    }
    upbp_hash_grid* get() const { return h_; }

private:
    Context& ctx_;
    upbp_hash_grid* h_ = nullptr;
    int64_t cells_ = 1;
};

}  // namespace upbp
