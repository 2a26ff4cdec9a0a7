/*
 * upbp.h -- C ABI of the MI355X volumetric merging estimators.
 *
 * Two estimators of SmallUPBP are served through this boundary:
 *   - the Beam Radiance Estimate (PB2D): a query beam gathers photon POINTS
 *     with the 2D Epanechnikov kernel (reference: SmallUPBP/src/Bre/Bre.hxx:40-184,
 *     SmallUPBP/src/Bre/Bre.cxx:199-277,812-964);
 *   - the photon-beam lookup (BB1D): a query beam gathers photon BEAMS with the
 *     1D kernel (reference: SmallUPBP/src/Beams/PhBeams.hxx:46-160,
 *     SmallUPBP/src/Beams/PhotonBeam.hxx:238-367, SmallUPBP/src/Beams/PhBeams.cxx:177-324).
 *
 * The reference interface is the pair of C++ classes EmbreeBre and
 * PhotonBeamsEvaluator, called once per camera ray. Here the same contract is
 * exposed as plain C over opaque handles: structures are built once per
 * iteration (build), evaluated for a BATCH of query rays (eval), and released
 * (destroy). Every array argument is a plain pointer plus a count; the
 * `memory` field of each input struct says whether its pointers are host or
 * device (HIP) pointers. No function exits the process: every failure returns
 * an upbp_status and leaves a message in upbp_ctx_last_error().
 *
 * Threading: a upbp_ctx is externally synchronised (one per host thread or a
 * caller-held mutex), like one EmbreeBre per OpenMP thread in the reference
 * (SmallUPBP/src/SmallUPBP.cxx:66-79). Handles built on a context belong to it.
 *
 * The library runs only on an AMD GPU (gfx950). There is no CPU fallback:
 * upbp_ctx_create fails with UPBP_EDEVICE when no device is usable.
 */
#ifndef UPBP_H_
#define UPBP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UPBP_ABI_VERSION 2

/* ---------------------------------------------------------------------------
 * Status codes. The reference has none: its assertions print and exit(2)
 * (SmallUPBP/src/Misc/Defs.hxx:96-97). */
typedef enum upbp_status {
    UPBP_OK = 0,
    UPBP_EINVAL = 1,        /* bad argument: null pointer, negative count, r <= 0, tmin > tmax, bad medium id */
    UPBP_ENOMEM = 2,        /* device or host allocation failed */
    UPBP_EDEVICE = 3,       /* HIP runtime error, no device, or a kernel reported an internal overflow */
    UPBP_ENONFINITE = 4,    /* NaN/Inf in an input position, direction, radius, power or segment */
    UPBP_EUNSUPPORTED = 5,  /* valid request this build does not serve (e.g. a reduced beam grid in a batched plan) */
    UPBP_ENOTBUILT = 6      /* handle is empty (build found no photons/beams) where data is required */
} upbp_status;

/* ---------------------------------------------------------------------------
 * Constants, numerically identical to the reference enums. */
/* BeamType, SmallUPBP/src/Misc/Defs.hxx:146-150 */
#define UPBP_SHORT_BEAM 1u
#define UPBP_LONG_BEAM 2u
/* EstimatorTechnique and OtherSettings, SmallUPBP/src/Misc/Defs.hxx:175-194 */
#define UPBP_BPT 4u
#define UPBP_SURF 8u
#define UPBP_PP3D 16u
#define UPBP_PB2D 32u
#define UPBP_BB1D 64u
#define UPBP_NO_SINE_IN_WEIGHTS 128u
#define UPBP_PREVIOUS 256u
#define UPBP_COMPATIBLE 512u
#define UPBP_BB1D_PREVIOUS 1024u
#define UPBP_SPECULAR_ONLY 2048u
/* AbstractMedium::RaySamplingFlags, SmallUPBP/src/Scene/Media.hxx:45-49 */
#define UPBP_ORIGIN_IN_MEDIUM 1u
#define UPBP_END_IN_MEDIUM 2u
/* RadiusCalculation, SmallUPBP/src/Misc/Defs.hxx:155-159 */
#define UPBP_CONSTANT_RADIUS 0
#define UPBP_KNN_RADIUS 1

typedef enum upbp_memory {
    UPBP_MEM_HOST = 0,   /* pointers are ordinary host memory; the library copies over PCIe */
    UPBP_MEM_DEVICE = 1  /* pointers are device memory on the context's GPU */
} upbp_memory;

/* Query ordering inside one eval call (the result order never changes). */
#define UPBP_ORDER_SORTED 0   /* default: queries are regrouped by a Morton key of their entry point so
                                 each 64-lane wave carries a spatially coherent bundle */
#define UPBP_ORDER_AS_GIVEN 1 /* lanes take queries in input order (caller already tiled them) */

typedef struct upbp_ctx upbp_ctx;
typedef struct upbp_bre upbp_bre;       /* photon-point structure: replaces EmbreeBre's Embree BVH4 */
typedef struct upbp_beams upbp_beams;   /* photon-beam structure: replaces PhotonBeamsEvaluator's grid */

/* ---------------------------------------------------------------------------
 * Homogeneous medium. Replaces HomogeneousMedium's constructor
 * (SmallUPBP/src/Scene/Media.hxx:129-159): the library takes absolute values,
 * forms sigma_t = sigma_a + sigma_s, picks the minimum positive sigma_t
 * component (first index on ties), and sets the has-scattering flag.
 * Emission is not used by either estimator and is not carried. */
typedef struct upbp_medium {
    float absorption[3];     /* sigma_a, RGB */
    float scattering[3];     /* sigma_s, RGB */
    float mean_cosine;       /* Henyey-Greenstein g; |g| < 1e-3 is isotropic (Media.hxx:111) */
    float continuation_prob; /* used by the UPBP MIS weights only (Media.hxx:117) */
} upbp_medium;

/* One medium interval of a query ray. Same fields as VolumeSegment
 * (SmallUPBP/src/Path/Ray.hxx:131-140) minus mEmission, which the estimators
 * never read. With lite segments (VolumeSegmentLite, Ray.hxx:222-227) only
 * dist_min, dist_max and medium_id are read. 32 bytes. */
typedef struct upbp_volume_segment {
    float dist_min;
    float dist_max;
    float ray_sample_pdf;
    float ray_sample_rev_pdf;
    float attenuation[3];
    int32_t medium_id;
} upbp_volume_segment;

/* ---------------------------------------------------------------------------
 * Build inputs. */
typedef struct upbp_build_params {
    int32_t radius_calculation; /* UPBP_CONSTANT_RADIUS or UPBP_KNN_RADIUS (see below) */
    float radius;               /* photon radius / beam radius, > 0, used when no per-item radius array;
                                   with UPBP_KNN_RADIUS the multiplier of 2 sqrt(d2_knn) */
    int32_t knn;                /* UPBP_KNN_RADIUS: which nearest neighbour (>= 1, the item itself
                                   counts as the first) sets the radius */
    int32_t leaf_size;          /* max items per LBVH leaf, 1..64; 0 = library default (8) */
    float beam_piece_factor;    /* beams only: target piece length in units of the max beam radius; 0 = default (8) */
    int32_t max_pieces_per_beam;/* beams only: cap on pieces per beam, 1..256; 0 = default */
    /* beams only: per-cell beam reduction of the reference's uniform grid (-gridmax / -gridred,
     * Grid.hxx:167-279,409-631; upbp_eval, upbp_hit_signatures and upbp_eval_upbp). A hit whose
     * query point lies in a cell holding n > max_beams_in_cell beams is kept as the reduction
     * type decides and weighted by 1 / (max_beams_in_cell / n). */
    uint32_t max_beams_in_cell; /* sMaxBeamsInCell; 0 = no reduction (the reference default) */
    uint32_t reduction_type;    /* sReductionType: UPBP_REDUCE_* */
    int32_t grid_resolution;    /* cells on the grid's longest axis (sGridSize); 0 = 256 */
    int32_t reduction_seed;     /* mSeed of the reduction's random numbers */
} upbp_build_params;

/* BeamReduction, SmallUPBP/src/Misc/Defs.hxx:164-170 */
#define UPBP_REDUCE_PRESAMPLE 0u      /* a fixed random subset of max beams per cell */
#define UPBP_REDUCE_OFFSET 1u         /* max consecutive beams of a shuffled cell from a random offset per query */
#define UPBP_REDUCE_RESAMPLE_FIXED 2u /* max beams drawn with replacement per query */
#define UPBP_REDUCE_RESAMPLE 3u       /* every beam kept with probability max / n per query */

/* UPBP_KNN_RADIUS (photons: Bre.cxx:562-602; beams: PhBeams.cxx:77-117):
 * d2_knn is the squared distance to the knn-th nearest item of the kd-tree the
 * reference builds (in-medium photon positions; beam origins), the query's own
 * item included, computed in float as ((dx*dx) + dy*dy) + dz*dz.
 *   photon radius      = radius * (2 sqrt(d2_knn at the photon))
 *   beam start radius  = max(2 sqrt(d2_knn at the origin) * radius, 0.001)
 *   beam end radius    = max(2 sqrt(d2_knn at origin + direction * length) * radius, 0.001)
 * With n <= knn items the largest of the n distances is used (the reference's
 * kd-tree then returns whichever item it visited first). Per-item radius arrays
 * must be NULL. A photon whose radius comes out 0 (knn coincident photons) makes
 * the build fail with UPBP_EINVAL. */

/* Photons: light vertices in media. Replaces the inputs of EmbreeBre::build
 * (SmallUPBP/src/Bre/Bre.cxx:526-632): position = mHitpoint,
 * inc_dir = mBSDF.WorldDirFix(), power = mThroughput. */
typedef struct upbp_photon_input {
    int64_t count;
    const float* position;    /* [count][3] */
    const float* inc_dir;     /* [count][3]; may be NULL when every medium is isotropic */
    const float* power;       /* [count][3] */
    const float* radius;      /* [count] per-photon radius, or NULL for params->radius */
    const uint8_t* in_medium; /* [count] filter (mIsInMedium, Bre.cxx:538-547), or NULL = all */
    upbp_memory memory;
} upbp_photon_input;

/* Photon beams. Replaces PhotonBeam's MC-filled fields
 * (SmallUPBP/src/Beams/PhotonBeam.hxx:44-54) and PhotonBeamsEvaluator::build's
 * radius assignment (SmallUPBP/src/Beams/PhBeams.cxx:92-117). */
typedef struct upbp_beam_input {
    int64_t count;
    const float* origin;       /* [count][3] mRay.origin */
    const float* direction;    /* [count][3] mRay.direction, unit */
    const float* length;       /* [count] mLength, > 0 */
    const float* throughput;   /* [count][3] mThroughputAtOrigin */
    const uint32_t* flags;     /* [count] UPBP_SHORT_BEAM or UPBP_LONG_BEAM (mFlags) */
    const int32_t* medium_id;  /* [count] mMedium as an index into the context's media */
    const float* start_radius; /* [count] or NULL for params->radius */
    const float* end_radius;   /* [count] or NULL for start radius (constant cone) */
    upbp_memory memory;
} upbp_beam_input;

/* ---------------------------------------------------------------------------
 * A batch of query rays. Each query is one call of EmbreeBre::evalBre /
 * PhotonBeamsEvaluator::evalBeamBeamEstimate with (beamType, ray, segments). */
typedef struct upbp_query_batch {
    int64_t count;
    const float* origin;                  /* [count][3] */
    const float* direction;               /* [count][3], unit */
    const int64_t* segment_offsets;       /* [count+1] CSR offsets into segments */
    const upbp_volume_segment* segments;  /* [segment_offsets[count]] */
    uint32_t beam_type;                   /* UPBP_SHORT_BEAM or UPBP_LONG_BEAM for every query */
    uint32_t lite_segments;               /* 1 = VolumeSegmentLite semantics (LONG only, Bre.cxx:895-964) */
    uint32_t order;                       /* UPBP_ORDER_SORTED or UPBP_ORDER_AS_GIVEN */
    upbp_memory memory;
} upbp_query_batch;

/* Counters of one eval/adjoint call. Hits are accepted (query, photon) and
 * (query, beam) pairs; candidates are items tested against a lane's ray. */
typedef struct upbp_eval_stats {
    uint64_t photon_hits;
    uint64_t beam_hits;
    uint64_t photon_candidates;
    uint64_t beam_candidates;
    uint64_t nodes_visited;       /* LBVH nodes fetched, summed over waves */
    float kernel_ms_photons;      /* HIP-event time of the photon traversal kernel(s) on the context stream */
    float kernel_ms_beams;        /* HIP-event time of the beam traversal kernel(s) */
    float kernel_ms_other;        /* validation, query sort, clears */
    uint32_t launches;            /* traversal kernel launches in the call */
} upbp_eval_stats;

/* ---------------------------------------------------------------------------
 * Context. */
upbp_status upbp_ctx_create(int device, upbp_ctx** out_ctx);
/* Drops the caller's reference. Photon and beam structures and batch plans
 * built on the context hold references of their own, so they may be destroyed
 * before or after it, in any order; the context is freed with the last one. */
upbp_status upbp_ctx_destroy(upbp_ctx* ctx);
const char* upbp_ctx_last_error(const upbp_ctx* ctx);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL restores the
 * context's own stream. Every kernel of the context runs on this stream. */
upbp_status upbp_ctx_set_stream(upbp_ctx* ctx, void* hip_stream);
void* upbp_ctx_get_stream(const upbp_ctx* ctx);
/* sync = 1 (default): calls return after the GPU work finished. sync = 0:
 * calls that write device outputs return after enqueueing (stats still wait). */
upbp_status upbp_ctx_set_sync(upbp_ctx* ctx, int sync);
/* Media table, indexed by medium_id. Replaces Scene::mMedia lookups
 * (Bre.cxx:836, PhBeams.cxx:201). Copied to the device. */
upbp_status upbp_set_media(upbp_ctx* ctx, const upbp_medium* media, int32_t count);

/* ---------------------------------------------------------------------------
 * Photon points (PB2D). */
/* Replaces EmbreeBre::build (SmallUPBP/src/Bre/Bre.cxx:526-632). Builds a GPU
 * LBVH over photon cubes pos +- r (Bre.cxx:174-177). *out_num_photons receives
 * the number of photons kept after the in_medium filter (the reference return
 * value). A build with zero photons succeeds with an empty handle whose
 * evaluation returns zeros (Bre.cxx:825-826). */
upbp_status upbp_bre_build(upbp_ctx* ctx, const upbp_photon_input* photons,
                           const upbp_build_params* params, upbp_bre** out_bre,
                           int64_t* out_num_photons);
/* Replaces EmbreeBre::destroy (Bre.cxx:762-777). NULL is accepted. */
upbp_status upbp_bre_destroy(upbp_bre* bre);
/* Replace the photon powers ([count][3], original input order of the kept
 * photons) without rebuilding the tree. Used by differentiable callers. */
upbp_status upbp_bre_set_power(upbp_bre* bre, const float* power, upbp_memory memory);
int64_t upbp_bre_count(const upbp_bre* bre);
/* Radius the build gave each input photon, in input order (0 for photons the
 * in-medium filter dropped): the `radius` of EmbreePhoton::set (Bre.cxx:596),
 * e.g. to read back UPBP_KNN_RADIUS radii. `radius` holds photons->count floats. */
upbp_status upbp_bre_radii(const upbp_bre* bre, float* radius, upbp_memory memory);

/* ---------------------------------------------------------------------------
 * Photon beams (BB1D). */
/* Replaces PhotonBeamsEvaluator::build (SmallUPBP/src/Beams/PhBeams.cxx:66-148):
 * assigns radii like the reference (constant radius: start = end = r,
 * mRadiusChange = 0; UPBP_KNN_RADIUS as above), then cuts every beam into pieces and builds a GPU LBVH
 * over the pieces' cone bounding boxes (PhotonBeam.hxx:186-198). Each
 * (query, beam) pair is counted once: a piece accepts only closest-approach
 * parameters t2 in its own half-open range. */
upbp_status upbp_beams_build(upbp_ctx* ctx, const upbp_beam_input* beams,
                             const upbp_build_params* params, upbp_beams** out_beams);
/* Replaces PhotonBeamsEvaluator::destroy (PhBeams.cxx:155-159). */
upbp_status upbp_beams_destroy(upbp_beams* beams);
int64_t upbp_beams_count(const upbp_beams* beams);
int64_t upbp_beams_piece_count(const upbp_beams* beams);
/* mStartRadius / mEndRadius the build gave each beam (PhBeams.cxx:92-117), in
 * input order; each array holds beams->count floats. */
upbp_status upbp_beams_radii(const upbp_beams* beams, float* start_radius, float* end_radius,
                             upbp_memory memory);
/* Replaces PhotonBeamsEvaluator::getBeamSelectionPdf (PhBeams.hxx, Grid.hxx:409-415): the
 * selection pdf of the reduction grid's cell containing the position (1 without reduction,
 * sMaxBeamsInCell = 0, the reference default). NaN when the device read fails; the reason is
 * then in upbp_ctx_last_error of the structure's context. */
float upbp_beams_selection_pdf(const upbp_beams* beams, const float position[3]);
/* getBeamSelectionPdf for many positions at once (xyz [count][3] -> out_pdf [count], both in
 * `memory`): the loop that sets mBB1DBeamSelectionPdf of every in-medium light vertex after the
 * beam build (SmallUPBP/src/Renderers/UPBP.hxx:723-728) in one launch. */
upbp_status upbp_beams_selection_pdfs(const upbp_beams* beams, const float* positions, int64_t count, float* out_pdf,
                                      upbp_memory memory);

/* Per-query statistics of the reference's uniform beam grid: struct GridStats
 * (SmallUPBP/src/Beams/GridStats.hxx:37-54) as Grid::intersect gathers it
 * (SmallUPBP/src/Beams/Grid.hxx:318-399, counts at :365,378-382) over every segment of a
 * query in a scattering medium, the walk of PhotonBeamsEvaluator::evalBeamBeamEstimate
 * (SmallUPBP/src/Beams/PhBeams.cxx:177-230; it ends once a channel of the attenuation product
 * is not above 1e-20). out_stats [count][3], in out_memory, query order as given:
 *   [0] intersectedBeams: the beams registered in each cell the query crosses, or
 *       max_beams_in_cell for a cell holding more;
 *   [1] intersectedCells: the cells crossed;
 *   [2] overfullCells: crossed cells holding more than max_beams_in_cell beams (0 without
 *       reduction).
 * The reference feeds these counts to its beam-density images (BeamDensity.hxx:110-140). The
 * grid is the one the structure's reduction uses; a structure built without reduction gets a
 * grid of its build's grid_resolution (0 = 256) for the call. The walk is the reference's FP32
 * arithmetic, so the counts are those of the reference's own grid. */
upbp_status upbp_beams_grid_stats(upbp_ctx* ctx, const upbp_beams* beams, const upbp_query_batch* queries,
                                  uint32_t* out_stats, upbp_memory out_memory);

/* Replaces EmbreeBre::evalBreBruteForce (SmallUPBP/src/Bre/Bre.cxx:980-1029), the reference's
 * debug estimate. Per query i: the ray origin[i], direction[i] (xyz each), the interval
 * (min_t[i], max_t[i]) and the medium medium_id[i]; every kept photon of the structure is
 * tested with the PB2D disc test (Bre.cxx:199-223) and a hit adds
 * flux exp(-sigma_t t) PhaseFunction::Evaluate(direction, incDir) / r^2; out[i] (RGB) is the
 * sum times sigma_s / pi. Its kernel is the reference's debug one, not evalBre's: constant over
 * the disc, attenuation from the ray origin, the stored incoming direction as given, photons
 * of every medium. All arrays in `memory`; one wave per query walks all photons, so the call is
 * meant for tests and debugging. UPBP_EINVAL for a medium outside the table. */
upbp_status upbp_bre_eval_brute_force(upbp_ctx* ctx, const upbp_bre* bre, const float* origin, const float* direction,
                                      const float* min_t, const float* max_t, const int32_t* medium_id, int64_t count,
                                      float* out, upbp_memory memory);

/* ---------------------------------------------------------------------------
 * Evaluation. out_photons receives, per query, EmbreeBre::evalBre
 * (Bre.cxx:812-879 full segments, :895-964 lite); out_beams receives
 * PhotonBeamsEvaluator::evalBeamBeamEstimate (PhBeams.cxx:177-241, :257-324),
 * both with techniques = PB2D / BB1D and no MIS data (the plain VolLightTracer
 * path, SmallUPBP/src/Renderers/VolLightTracer.hxx:378,384). Normalisation by
 * the light-path count stays with the caller, as in the reference.
 * Either structure may be NULL, then its output is not written.
 * Outputs are [count][3] floats in the query order of the batch. */
upbp_status upbp_eval(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                      const upbp_query_batch* queries, float* out_photons, float* out_beams,
                      upbp_memory out_memory, upbp_eval_stats* stats);

/* Size of the fixed per-device region upbp_eval_noalloc works in, and the largest
 * UPBP_ORDER_SORTED batch it accepts (1920x1080 fits; UPBP_ORDER_AS_GIVEN has no such limit). */
#define UPBP_NOALLOC_SCRATCH_BYTES (64u << 20)
#define UPBP_NOALLOC_MAX_SORTED_QUERIES (1 << 21)

/* No counterpart in the reference. upbp_eval for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same kernels and results as upbp_eval with out_memory = UPBP_MEM_DEVICE, but the counters,
 * query keys, permutations and sort storage live in a fixed region of
 * UPBP_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on until the
 * process ends, so the call makes no hipMalloc, hipFree or other allocating call. The queries
 * and both outputs must be device memory (UPBP_EINVAL otherwise), and a UPBP_ORDER_SORTED
 * batch may hold at most UPBP_NOALLOC_MAX_SORTED_QUERIES queries (UPBP_EINVAL otherwise).
 * Every call on one GPU shares that GPU's region, whichever context makes it: calls that
 * overlap in time on one GPU (from two host threads, or on two streams) race and must be
 * serialised by the caller. */
upbp_status upbp_eval_noalloc(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                              const upbp_query_batch* queries, float* out_photons, float* out_beams,
                              upbp_eval_stats* stats);

/* Hit-set signatures per query: sig[q] = {number of hits, sum over hits of
 * hash32(original item index)} (wrapping uint32). Identical hit sets give
 * identical signatures; used to compare against a CPU evaluation. Either
 * output may be NULL. Outputs are [count][2] uint32 in batch order. */
upbp_status upbp_hit_signatures(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                                const upbp_query_batch* queries, uint32_t* photon_sig,
                                uint32_t* beam_sig, upbp_memory out_memory);

/* ---------------------------------------------------------------------------
 * Adjoint of upbp_eval for a scalar loss L(out_photons, out_beams). Inputs are
 * the upstream gradients dL/dout ([count][3], NULL = zero). Outputs (each may
 * be NULL) are ACCUMULATED into (+=), so clear them first:
 *   grad_scattering, grad_absorption: [media][3]  dL/dsigma_s, dL/dsigma_a
 *   grad_photon_power: [photons][3] dL/dpower, original order of the kept photons
 *   grad_beam_power:   [beams][3]   dL/dthroughput
 * Segment attenuations are taken as exp(-sigma_t * (dist_max - dist_min)) of the
 * segment's medium (what Scene::Intersect stores, Scene.hxx:542), so their
 * derivative is -(dist_max - dist_min) * attenuation. Transmittance is
 * recomputed, never stored. The media use |sigma| (Media.hxx:136-140), so
 * the medium gradients are taken with respect to the values given to
 * upbp_set_media: an entry of a negative coefficient carries its sign.
 * SHORT query beams: the in-segment transmittance is normalised by its
 * minimum-sigma_t channel k (Bre.cxx:257-263, PhotonBeam.hxx:336-345), and an
 * earlier segment's factor attenuation / ray_sample_pdf is taken with the
 * pass-through pdf exp(-sigma_t_k * (dist_max - dist_min)) of a medium that
 * samples distances (RaySamplePdf, Media.hxx:294-323); k is held fixed (the
 * derivative at a tie of the minimum channels is one-sided). A beam structure
 * with per-cell reduction weights each beam hit as upbp_eval does (its count
 * over the cell's pdf, neither of which depends on the media). */
upbp_status upbp_adjoint(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                         const upbp_query_batch* queries, const float* grad_out_photons,
                         const float* grad_out_beams, float* grad_scattering,
                         float* grad_absorption, float* grad_photon_power,
                         float* grad_beam_power, upbp_memory memory, upbp_eval_stats* stats);

/* ---------------------------------------------------------------------------
 * UPBP mode: the same two estimators weighted by the UPBP multiple-importance-
 * sampling weights, i.e. the callbacks breIntersectFuncHomogeneous2
 * (SmallUPBP/src/Bre/Bre.cxx:293-464) and PhotonBeam::accumulate2
 * (SmallUPBP/src/Beams/PhotonBeam.hxx:386-607) with the path-weight walks
 * AccumulateCameraPathWeight / AccumulateLightPathWeight
 * (SmallUPBP/src/Path/PathWeight.hxx:33-270). The inputs mirror
 * AdditionalRayDataForMis (embree-2.0/embree/common/ray.h:31-99), MisData and
 * UPBPLightVertex (SmallUPBP/src/Path/UPBPLightVertex.hxx:32-78).
 * Nothing the caller passes is modified (the reference writes per-hit MIS data
 * into the caller's camera and light-vertex arrays, Bre.cxx:395-422,
 * PhotonBeam.hxx:499-595; here those values live in registers). */

/* MisData flags (the four bools of MisData) */
#define UPBP_MIS_DELTA 1u
#define UPBP_MIS_ON_LIGHT_SOURCE 2u
#define UPBP_MIS_SPECULAR 4u
#define UPBP_MIS_IN_MEDIUM_WITH_BEAMS 8u

/* MisData, UPBPLightVertex.hxx:32-52. 64 bytes. */
typedef struct upbp_mis_data {
    float pdf_a_inv;                 /* mPdfAInv */
    float rev_pdf_a;                 /* mRevPdfA */
    float rev_pdf_a_without_bsdf;    /* mRevPdfAWithoutBsdf */
    float ray_sample_pdf_inv;        /* mRaySamplePdfInv */
    float ray_sample_rev_pdf_inv;    /* mRaySampleRevPdfInv */
    float ray_sample_pdfs_ratio;     /* mRaySamplePdfsRatio */
    float ray_sample_rev_pdfs_ratio; /* mRaySampleRevPdfsRatio */
    float sin_theta;                 /* mSinTheta */
    float cos_theta_out;             /* mCosThetaOut */
    float surf_factor;               /* mSurfMisWeightFactor */
    float pp3d_factor;               /* mPP3DMisWeightFactor */
    float pb2d_factor;               /* mPB2DMisWeightFactor */
    float bb1d_factor;               /* mBB1DMisWeightFactor */
    float bb1d_beam_selection_pdf;   /* mBB1DBeamSelectionPdf */
    uint32_t flags;                  /* UPBP_MIS_* */
    uint32_t pad;
} upbp_mis_data;

/* The UPBPLightVertex fields the estimators read (UPBPLightVertex.hxx:55-67);
 * the position is the photon position given to upbp_bre_build. 112 bytes. */
typedef struct upbp_light_vertex {
    float throughput[3];        /* mThroughput */
    float dir_fix[3];           /* mBSDF.WorldDirFix() */
    float continuation_prob;    /* mBSDF.ContinuationProb() */
    int32_t path_idx;           /* mPathIdx */
    uint32_t path_length;       /* mPathLength */
    uint32_t is_finite;         /* mIsFinite */
    uint32_t flags;             /* UPBP_LV_*: read by PP3D / SURF merging only (0 for PB2D / BB1D use) */
    uint32_t pad;
    upbp_mis_data mis;          /* mMisData */
} upbp_light_vertex;

/* upbp_light_vertex.flags (UPBPLightVertex.hxx:64-66) */
#define UPBP_LV_CONNECTABLE 1u  /* mConnectable: the vertex can be merged with */
#define UPBP_LV_IN_MEDIUM 2u    /* mInMedium: the vertex enters a PP3D grid, else a SURF grid (MatchesType) */

/* PhotonBeam MIS fields (PhotonBeam.hxx:48-54), one per beam. 32 bytes. */
typedef struct upbp_beam_mis {
    int32_t light_vertex;         /* mLightVertex as an index into the light vertices */
    uint32_t ray_sampling_flags;  /* mRaySamplingFlags */
    float ray_sample_pdf;         /* mRaySamplePdf */
    float ray_sample_rev_pdf;     /* mRaySampleRevPdf */
    float last_pdf_w_inv;         /* mLastPdfWInv */
    uint32_t pad[3];
} upbp_beam_mis;

/* Per-query part of AdditionalRayDataForMis plus the camera sub-path's MisData
 * (mCameraVerticesMisData): entries camera_offset .. camera_offset + camera_path_length
 * of the camera array; the last entry is recomputed per hit (Bre.cxx:396-419). 24 bytes. */
typedef struct upbp_query_mis {
    int64_t camera_offset;
    uint32_t camera_path_length;  /* mCameraPathLength, >= 1 */
    uint32_t ray_sampling_flags;  /* raySamplingFlags of evalBre: 0 or UPBP_ORIGIN_IN_MEDIUM */
    float last_pdf_w_inv;         /* mLastPdfWInv */
    uint32_t pad;
} upbp_query_mis;

/* Per-call constants of AdditionalRayDataForMis and the renderer (UPBP.hxx:423-460,780-907). */
typedef struct upbp_mis_params {
    uint32_t techniques;          /* estimatorTechniques | OtherSettings flags (PREVIOUS, BB1D_PREVIOUS, NO_SINE...) */
    uint32_t query_beam_type;     /* mQueryBeamType */
    uint32_t photon_beam_type;    /* mPhotonBeamType */
    uint32_t min_path_length;     /* mMinPathLength */
    uint32_t max_path_length;     /* mMaxPathLength */
    float surf_factor;            /* mSurfMisWeightFactor */
    float pp3d_factor;            /* mPP3DMisWeightFactor */
    float pb2d_factor;            /* mPB2DMisWeightFactor */
    float bb1d_factor;            /* mBB1DMisWeightFactor */
    float bb1d_min_mfp;           /* mBB1DMinMFP */
    float min_dist_to_med;        /* mMinDistToMed */
    int32_t has_bb1d_beams;       /* mBB1DPhotonBeams != NULL (PB2D sets the beam selection pdf from it) */
} upbp_mis_params;

typedef struct upbp_mis_inputs {
    const upbp_query_mis* queries;            /* [queries.count] */
    const upbp_mis_data* camera;              /* [camera_count] */
    int64_t camera_count;
    const upbp_light_vertex* light_vertices;  /* mLightVertices: every vertex of every light sub-path */
    int64_t light_vertex_count;
    const int32_t* path_ends;                 /* mPathEnds */
    int64_t path_count;
    const upbp_beam_mis* beams;               /* [beams count], required with a beam structure */
    upbp_mis_params params;
    upbp_memory memory;                       /* of every pointer above */
} upbp_mis_inputs;

/* UPBP evaluation. The photon structure must have been built from the light
 * vertices themselves (photon i of the build input = light vertex i, filtered by
 * in_medium), as EmbreeBre::build does for UPBPLightVertex (Bre.cxx:649-755).
 * A beam structure built with per-cell reduction (max_beams_in_cell > 0, below 2^23)
 * weights each beam hit as the reference's reduced grid does: the hit is tested as often
 * as its query point's cell decides (upbp_eval's rule), and each test divides by the
 * cell's selection pdf and passes it to the MIS weights as beamSelectionPdf
 * (PhotonBeam.hxx:487-564). A PB2D hit's camera vertex then takes the pdf of the cell
 * holding the hit point as its BB1D selection pdf (Bre.cxx:407-415). With out_beams NULL
 * the beam structure is not evaluated and serves only as that evaluator. */
upbp_status upbp_eval_upbp(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                           const upbp_query_batch* queries, const upbp_mis_inputs* mis,
                           float* out_photons, float* out_beams, upbp_memory out_memory,
                           upbp_eval_stats* stats);

/* ---------------------------------------------------------------------------
 * Vertex merging through the reference's hashed uniform grid (HashGrid,
 * SmallUPBP/src/Misc/HashGrid.hxx:38-230) with the RangeQuery::Process weight
 * (SmallUPBP/src/Renderers/UPBP.hxx:63-133):
 *   PP3D: point-point 3D merging of camera vertices in media (UPBP.hxx:1110-1120);
 *   SURF: surface photon mapping at camera vertices on surfaces (UPBP.hxx:1097-1107).
 * A camera vertex gathers the light vertices within a fixed radius. The grid is
 * reproduced exactly, including two of its properties: a camera vertex outside the
 * bounding box of the grid's vertices finds nothing (HashGrid.hxx:137-143), and when two
 * of the eight neighbour cells hash to the same bucket that bucket is visited twice
 * (HashGrid.hxx:161-185). */
typedef struct upbp_hash_grid upbp_hash_grid;

/* Replaces HashGrid::Reserve(cell_count) + HashGrid::Build(mLightVertices, radius, type)
 * (UPBP.hxx:689-703, HashGrid.hxx:41-128). type = UPBP_PP3D keeps the vertices with
 * UPBP_LV_IN_MEDIUM, type = UPBP_SURF the others (UPBPLightVertex::MatchesType,
 * UPBPLightVertex.hxx:74-77). positions: [count][3] (mHitpoint of each light vertex);
 * vertices: [count], only `flags` is read. The grid keeps light-vertex indices, so the
 * evaluation takes the same light-vertex array (mis->light_vertices). cell_count >= 1 is
 * the number of hash buckets (the reference reserves one per light path). */
upbp_status upbp_hash_grid_build(upbp_ctx* ctx, const float* positions, const upbp_light_vertex* vertices,
                                 int64_t count, float radius, int64_t cell_count, uint32_t type,
                                 upbp_memory memory, upbp_hash_grid** out_grid);
upbp_status upbp_hash_grid_destroy(upbp_hash_grid* grid);
int64_t upbp_hash_grid_count(const upbp_hash_grid* grid);  /* vertices in the grid */

/* One camera vertex in a medium (PP3D): RangeQuery's position, camera BSDF (a medium
 * BSDF: scattering coefficient, mean cosine and continuation probability of `medium`,
 * Bsdf.hxx:144-156) and SubPathState (path length; the camera MisData of the path is
 * entries camera_offset .. camera_offset + camera_path_length of mis->camera, the
 * last being this vertex's, as mCameraVerticesMisData holds it). 48 bytes. */
typedef struct upbp_pp3d_query {
    float position[3];            /* hitPoint */
    int32_t medium;               /* medium id of the vertex */
    float dir_fix[3];             /* bsdf.WorldDirFix() = -ray.direction */
    uint32_t camera_path_length;  /* cameraState.mPathLength, >= 1 */
    int64_t camera_offset;
    int64_t pad;
} upbp_pp3d_query;

/* One camera vertex on a surface (SURF): position, path length and camera MisData as
 * for PP3D, and the camera vertex's surface BSDF as the renderer set it up
 * (Bsdf.hxx:102-140, GetComponentProbabilities :672-706): its shading frame, the fixed
 * direction in that frame, the lobe probabilities, the continuation probability and
 * the material's diffuse and Phong parameters. Only non-delta BSDFs merge
 * (UPBP.hxx:1098). 128 bytes. */
typedef struct upbp_surf_query {
    float position[3];            /* hitPoint */
    uint32_t camera_path_length;  /* cameraState.mPathLength, >= 1 */
    float frame_x[3];             /* mFrame.mX (Frame.hxx:56-62) */
    float continuation_prob;      /* mContinuationProb */
    float frame_y[3];             /* mFrame.mY */
    float diff_prob;              /* mDiffProb */
    float frame_z[3];             /* mFrame.mZ, the shading normal */
    float phong_prob;             /* mPhongProb */
    float local_dir_fix[3];       /* mLocalDirFix = mFrame.ToLocal(-ray.direction) */
    float phong_exponent;         /* Material::mPhongExponent */
    float diffuse[3];             /* Material::mDiffuseReflectance */
    float pad0;
    float phong[3];               /* Material::mPhongReflectance */
    float pad1;
    int64_t camera_offset;
    int64_t pad2;
} upbp_surf_query;

/* upbp_pp3d_eval / upbp_surf_eval flags */
#define UPBP_VM_NO_MIS 1u  /* mAlgorithm == kPPM: every MIS weight is 1 (UPBP.hxx:121) */

/* RangeQuery::GetContrib() per camera vertex (the caller multiplies by
 * cameraState.mThroughput * mPP3DNormalization, UPBP.hxx:1117, or mSurfNormalization,
 * UPBP.hxx:1104). The grid must have been built with type UPBP_PP3D / UPBP_SURF.
 * out: [count][3] or NULL; sig: [count][2] or NULL, {number of Process calls, sum of
 * hash32(light vertex index)} over the vertices the grid finds within the radius
 * (duplicates included), as upbp_hit_signatures. mis->camera, light_vertices, path_ends
 * and params are read; mis->queries and mis->beams are not. `memory` is that of
 * queries, out and sig. */
upbp_status upbp_pp3d_eval(upbp_ctx* ctx, const upbp_hash_grid* grid, const upbp_pp3d_query* queries,
                           int64_t count, const upbp_mis_inputs* mis, uint32_t flags, float* out,
                           uint32_t* sig, upbp_memory memory, upbp_eval_stats* stats);
upbp_status upbp_surf_eval(upbp_ctx* ctx, const upbp_hash_grid* grid, const upbp_surf_query* queries,
                           int64_t count, const upbp_mis_inputs* mis, uint32_t flags, float* out,
                           uint32_t* sig, upbp_memory memory, upbp_eval_stats* stats);

/* ---------------------------------------------------------------------------
 * Light sub-path tracing in media: the producer of the light vertices, MisData
 * and photon beams the estimators above consume. Replaces the light-tracing
 * loop of UPBPRenderer::RunIteration (SmallUPBP/src/Renderers/UPBP.hxx:501-674)
 * with GenerateLightSample (:1592-1662), AddBeams (:1783-1870) and
 * SampleScattering (:1908-1969), for scenes of homogeneous media bounded by
 * imaginary (index-matched) spheres and boxes and lit by point lights
 * (PointLight::Emit, Scene/Lights.hxx:363-384). Scene::Intersect's media part
 * (Scene/Scene.hxx:393-531) becomes: the ray is cut at every boundary crossing,
 * each piece takes the medium of the innermost region containing its midpoint,
 * and distance sampling follows HomogeneousMedium::SampleRay / RaySamplePdf
 * (Scene/Media.hxx:240-323). The random numbers come from a counter-based
 * generator (path index, draw index) instead of the reference's sequential
 * mt19937, so one path's numbers do not depend on the others. Real surfaces,
 * area and environment lights, camera connections (ConnectToCamera) and
 * heterogeneous media are not part of this producer. */
#define UPBP_MAX_REGIONS 16    /* regions per scene */
#define UPBP_REGION_SPHERE 0u /* p0 = centre, p1[0] = radius */
#define UPBP_REGION_BOX 1u    /* p0 = minimum corner, p1 = maximum corner */
typedef struct upbp_region {
    uint32_t shape;     /* UPBP_REGION_* */
    int32_t medium_id;  /* medium inside the region (index into the context's media) */
    float p0[3];
    float p1[3];
} upbp_region;

typedef struct upbp_point_light {
    float position[3];
    float intensity[3]; /* mIntensity */
} upbp_point_light;

typedef struct upbp_light_trace_params {
    uint64_t seed;
    uint32_t photon_beam_type;  /* mPhotonBeamType: UPBP_SHORT_BEAM or UPBP_LONG_BEAM */
    uint32_t max_path_length;   /* mMaxPathLength (>= 1) */
    int32_t global_medium;      /* medium outside every region (mScene global medium) */
    int32_t beam_path_count;    /* mBB1DUsedLightSubPathCount: paths below it store beams; 0 = none */
    int32_t merge_pb2d;         /* mMergeWithLightVerticesPB2D: beams and IN_MEDIUM_WITH_BEAMS need mfp > bb1d_min_mfp */
    float surf_factor, pp3d_factor, pb2d_factor, bb1d_factor; /* m*MisWeightFactor */
    float bb1d_min_mfp;         /* mBB1DMinMFP */
    uint32_t pad;
} upbp_light_trace_params;

/* Outputs, in caller memory of one kind. Light vertices are grouped per path in
 * path order (vertex 0 of a path is its light-source vertex); path_ends[p] is
 * the end of path p, as mPathEnds. Beams follow in path order, each naming its
 * origin light vertex by global index. counts are written even when a capacity
 * is exceeded (the call then fails with UPBP_EINVAL and writes nothing else). */
typedef struct upbp_light_paths {
    upbp_light_vertex* vertices;  /* [vertex_capacity] */
    float* positions;             /* [vertex_capacity][3] mHitpoint */
    int64_t vertex_capacity;
    int32_t* path_ends;           /* [path_count] */
    float* beam_origin;           /* [beam_capacity][3] */
    float* beam_direction;        /* [beam_capacity][3] */
    float* beam_length;           /* [beam_capacity] */
    float* beam_throughput;       /* [beam_capacity][3] mThroughputAtOrigin */
    uint32_t* beam_flags;         /* [beam_capacity] UPBP_SHORT_BEAM / UPBP_LONG_BEAM */
    int32_t* beam_medium;         /* [beam_capacity] */
    upbp_beam_mis* beam_mis;      /* [beam_capacity] */
    int64_t beam_capacity;
    int64_t vertex_count;         /* out */
    int64_t beam_count;           /* out */
} upbp_light_paths;

upbp_status upbp_trace_light_paths(upbp_ctx* ctx, const upbp_region* regions, int32_t region_count,
                                   const upbp_point_light* lights, int32_t light_count,
                                   const upbp_light_trace_params* params, int64_t path_count,
                                   upbp_light_paths* out, upbp_memory memory);

/* ---------------------------------------------------------------------------
 * Batched evaluation of many independent instances in one persistent launch.
 * Each instance names its structures, a device-resident query batch, its
 * media (a range of the context's media table: medium_id + media_base) and
 * device outputs. Waves take (instance, query tile) items from a device work
 * queue until it drains. An instance's media are the table entries from
 * media_base to the end: a segment medium_id outside [0, count - media_base)
 * fails the run with UPBP_EINVAL. upbp_set_media between runs is allowed: the
 * next run checks every media_base against the new table and captures its
 * graph again. */
typedef struct upbp_instance {
    const upbp_bre* bre;          /* may be NULL */
    const upbp_beams* beams;      /* may be NULL */
    upbp_query_batch queries;     /* memory must be UPBP_MEM_DEVICE */
    int32_t media_base;           /* added to every medium id of this instance */
    float* out_photons;           /* device [count][3] or NULL */
    float* out_beams;             /* device [count][3] or NULL */
} upbp_instance;

typedef struct upbp_batch_plan upbp_batch_plan;
/* Prepare a reusable plan: validates the instances, uploads the descriptor
 * table, sorts each instance's queries once, and COPIES every instance's photon
 * and beam structures into merged arrays owned by the plan (one node array and
 * one item array per estimator, child references rebased), so the persistent
 * kernel reads them through kernel arguments like a single evaluation. The
 * structures may be destroyed after this call; rebuilding or re-powering them
 * (upbp_bre_set_power) does not reach an existing plan. Query arrays and
 * output buffers are referenced, not copied: they must outlive the plan and
 * may be refilled between runs. */
upbp_status upbp_batch_plan_create(upbp_ctx* ctx, const upbp_instance* instances,
                                   int32_t count, upbp_batch_plan** out_plan);
upbp_status upbp_batch_plan_destroy(upbp_batch_plan* plan);
/* Run the plan: one fill kernel that resets the queue counters, then one
 * persistent photon kernel and one persistent beam kernel. With
 * use_graph = 1 the three launches are captured once into a hipGraph and
 * replayed on later calls. */
upbp_status upbp_batch_run(upbp_ctx* ctx, upbp_batch_plan* plan, int use_graph,
                           upbp_eval_stats* stats);

/* ---------------------------------------------------------------------------
 * Launch tuning. The autotuner times the traversal kernels over a grid of
 * wave-per-block and waves-per-CU settings and keeps the fastest per
 * (structure size, query count) bucket. */
typedef struct upbp_tuning {
    int32_t waves_per_block;        /* 1..4: waves per workgroup of the traversal kernels */
    int32_t blocks_per_cu;          /* persistent kernels: resident blocks per CU; 0 = occupancy-derived */
    int32_t photon_waves_per_block; /* 1..4: the photon kernel's own count; 0 = waves_per_block */
    int32_t frustum_leaf;           /* 1..64: a wave whose rays share their origin tests an inner
                                       node's subtree of at most this many items as one leaf
                                       (1 = only the built leaves); 0 = library default (64) */
} upbp_tuning;
upbp_status upbp_set_tuning(upbp_ctx* ctx, const upbp_tuning* tuning);
upbp_status upbp_get_tuning(const upbp_ctx* ctx, upbp_tuning* tuning);
upbp_status upbp_autotune(upbp_ctx* ctx, const upbp_bre* bre, const upbp_beams* beams,
                          const upbp_query_batch* queries, upbp_tuning* best, float* best_ms);

#ifdef __cplusplus
}
#endif

#endif /* UPBP_H_ */
