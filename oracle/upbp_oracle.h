/*
 * upbp_oracle.h -- CPU restatement of SmallUPBP's volumetric merging estimators.
 *
 * TEST INFRASTRUCTURE. This library is the checker of the GPU library and the
 * timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. The product (smallupbp_amd, libupbp_hip.so) never
 * links, loads or calls it.
 *
 * It is a clean-room restatement of the reference algorithm in the reference's
 * floating-point operation order (FP32, no FMA contraction), with an FP64
 * re-evaluation of the same mathematics. Every function cites the reference
 * file:line it follows. Accelerations:
 *   photons: brute force over all photons, or a binned-SAH binary BVH over the
 *            photon cubes (the role of Embree's BVH4, SmallUPBP/src/Bre/Bre.cxx:492-507);
 *   beams:   brute force over all beams (SmallUPBP/src/Beams/PhBrute.hxx:80-88),
 *            or the reference's uniform grid with 3D-DDA (SmallUPBP/src/Beams/Grid.hxx:167-399).
 * Culling never changes a hit decision: the predicate that decides a hit is
 * the same function on every path.
 */
#ifndef UPBP_ORACLE_H_
#define UPBP_ORACLE_H_

#include <stdint.h>
#include "../include/upbp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Derived medium constants (SmallUPBP/src/Scene/Media.hxx:129-159). */
typedef struct or_medium {
    float sigma_a[3], sigma_s[3], sigma_t[3];
    float mean_cosine, continuation_prob;
    float min_comp;   /* minimum positive sigma_t component (Media.hxx:142-145) */
    int32_t min_idx;  /* its index; 0 when no component is positive */
    int32_t has_scattering, has_attenuation;
    float mfp, max_beam_length;
} or_medium;

void or_prepare_media(const upbp_medium* in, int32_t n, or_medium* out);

/* Cephes/Pommier single-precision exp (the method behind exp_ps,
 * SmallUPBP/src/PrecompiledLibs/FastFloat/sse_mathfun.h:250-300). */
float or_exp_cephes(float x);

/* Phase function pdf (SmallUPBP/src/Path/PhaseFunction.hxx:52-80). */
float or_phase_pdf(const float fix[3], const float gen[3], float g, float* sin_theta);

/* Reference predicates, exposed for closed-form tests. */
int or_test_bre(const float o[3], const float d[3], float min_t, float max_t, const float p[3],
                float r2, float* t, float* s);
int or_test_beam_beam(const float o1[3], const float d1[3], float min_t1, float max_t1,
                      const float o2[3], const float d2[3], float min_t2, float max_t2,
                      float max_dist_sqr, float* dist, float* sin_theta, float* t1, float* t2);

/* The path-weight walks themselves in FP64 (AccumulateCameraPathWeight /
 * AccumulateLightPathWeight, PathWeight.hxx:33-270), for first-principles checks. The
 * camera walk reads cam[0 .. len] as given; the light walk reads the light path of
 * path_idx through path_ends. */
double or_camera_walk(const upbp_mis_data* cam, int len, double last_rev, double last_sin, double last_ray_rev_inv,
                      double last_ray_rev_ratio, double next_to_last_partial, uint32_t qbt, uint32_t pbt, uint32_t tech);
double or_light_walk(const upbp_light_vertex* lv, int64_t lv_count, const int32_t* path_ends, int64_t path_count,
                     int path_idx, int path_length, double last_rev, double next_to_last_partial, uint32_t evaluated,
                     uint32_t qbt, uint32_t pbt, uint32_t tech, int* err);

/* 32-bit hash of an item index used by hit signatures (same as the GPU library). */
uint32_t or_hash32(uint32_t i);

/* ---- photon points (PB2D) ------------------------------------------------ */
typedef struct or_photons or_photons;
/* EmbreeBre::build semantics (Bre.cxx:526-632): keep photons with in_medium != 0
 * (all when NULL), radius from the array or const_radius, r2 = r * r. Builds a BVH
 * (leaf_size items per leaf) for accel = 1. */
or_photons* or_photons_create(int64_t n, const float* pos, const float* inc_dir, const float* power,
                              const float* radius, float const_radius, const uint8_t* in_medium,
                              int32_t leaf_size);
void or_photons_free(or_photons* ph);
int64_t or_photons_count(const or_photons* ph);
/* Copy of the kept photons' original indices (int64 [count]). */
void or_photons_orig_index(const or_photons* ph, int64_t* out);

/* EmbreeBre::evalBre (Bre.cxx:812-879 full, :895-964 lite) with the plain
 * PB2D callback breIntersectFuncHomogeneous (Bre.cxx:239-277), per query.
 * accel: 0 brute force, 1 BVH. precision: 32 (reference order, Cephes exp) or
 * 64 (double). Writes out32 ([Q][3]) for 32, out64 for 64. sig ([Q][2]) and
 * counters ([2]: hits, candidates) are optional. Returns 0 or an upbp_status. */
int or_bre_eval(const or_photons* ph, const upbp_medium* media, int32_t n_media,
                const upbp_query_batch* q, int accel, int precision, int threads,
                float* out32, double* out64, uint32_t* sig, uint64_t* counters);

/* ---- photon beams (BB1D) ------------------------------------------------- */
typedef struct or_beams or_beams;
/* PhotonBeamsEvaluator::build semantics (PhBeams.cxx:92-117): constant radius
 * gives start = end = r, mMaxRadiusSqr = r*r, mRadiusChange = 0; per-beam radii
 * give mRadiusChange = (end - start) / length. grid_res > 0 builds the
 * reference uniform grid (Grid.hxx:223-279) with that many cells on the
 * longest axis. */
or_beams* or_beams_create(int64_t n, const float* origin, const float* dir, const float* length,
                          const float* throughput, const uint32_t* flags, const int32_t* medium,
                          const float* start_radius, const float* end_radius, float const_radius,
                          int32_t grid_res);
void or_beams_free(or_beams* bm);
int64_t or_beams_count(const or_beams* bm);
/* Per-cell beam reduction of the grid (Grid::reduceBeams, Grid.hxx:558-631; needs
 * grid_res > 0). or_beams_eval then applies it: accel 1 as the reference does, per crowded
 * cell of the DDA with the cell's ray interval; accel 0 by the cell of each hit's query
 * point (the GPU's formulation). Random numbers are hashes of (seed, query ray, cell, draw). */
int or_beams_set_reduction(or_beams* bm, uint32_t max_beams_in_cell, uint32_t reduction_type, int32_t seed);
/* Grid::pdf (Grid.hxx:409-415): selection pdf of the cell holding pos. */
float or_beams_selection_pdf(const or_beams* bm, const float* pos);
/* Grid statistics: res[3], cells, pointers (registered beam references). */
void or_beams_grid_info(const or_beams* bm, int64_t* info6);
/* The grid's box lo[3], hi[3] and cell size[3]; the beams registered in each of its
 * res[0] res[1] res[2] cells (x fastest; n must be that product). */
void or_beams_grid_geometry(const or_beams* bm, float* out9);
int or_beams_grid_cell_counts(const or_beams* bm, uint32_t* out, int64_t n);

/* PhotonBeamsEvaluator::evalBeamBeamEstimate (PhBeams.cxx:177-241, :257-324)
 * with PhotonBeam::accumulate (PhotonBeam.hxx:315-367).
 * accel: 0 brute force (PhBrute.hxx:80-88), 1 grid DDA (Grid.hxx:318-399). */
int or_beams_eval(const or_beams* bm, const upbp_medium* media, int32_t n_media,
                  const upbp_query_batch* q, int accel, int precision, int threads,
                  float* out32, double* out64, uint32_t* sig, uint64_t* counters);

/* EmbreeBre::evalBreBruteForce (Bre.cxx:980-1029): per query [count] (origin, direction, [min_t,
 * max_t], medium) -> out [count][3], every kept photon tested; its kernel is the reference's
 * debug one (constant disc kernel 1 / (pi r^2), attenuation from the ray origin, phase of the
 * stored incoming direction as given, no photon medium check). */
int or_bre_eval_brute_force(const or_photons* ph, const upbp_medium* media, int32_t n_media, const float* origin,
                            const float* direction, const float* min_t, const float* max_t, const int32_t* medium,
                            int64_t count, int threads, float* out);

/* GridStats of Grid::intersect per query over its scattering segments (GridStats.hxx:37-54,
 * Grid.hxx:318-399, PhBeams.cxx:177-230): out[3 q] beams tested (a reduced cell counts
 * max_beams_in_cell), out[3 q + 1] cells crossed, out[3 q + 2] reduced (overfull) cells
 * crossed. Needs grid_res > 0. */
int or_beam_grid_stats(const or_beams* bm, const upbp_medium* media, int32_t n_media, const upbp_query_batch* q,
                       int threads, uint32_t* out);

/* ---- UPBP mode ----------------------------------------------------------- */
/* evalBre / evalBeamBeamEstimate with MIS data: breIntersectFuncHomogeneous2
 * (Bre.cxx:293-464), PhotonBeam::accumulate2 (PhotonBeam.hxx:386-607),
 * AccumulateCameraPathWeight / AccumulateLightPathWeight (PathWeight.hxx:33-270),
 * exact path walks. which: 0 photons (photon i = light vertex orig[i]), 1 beams.
 * accel 0: brute force over items; accel 1: photons through the BVH, beams through the
 * reference's uniform grid (the timed CPU comparison). counters[0] = hits (beam hits
 * counted once per test). When bm reduces its cells (or_beams_set_reduction), a beam hit
 * passes its cell's pdf as beamSelectionPdf and is counted as often as the cell tests it
 * (accel 0: the cell of the hit's query point, as or_beams_eval), and a photon hit's
 * camera vertex takes the pdf of the cell holding the hit point (Bre.cxx:407-415);
 * bm is read for that even when which = 0. */
int or_upbp_eval(const or_photons* ph, const or_beams* bm, const upbp_medium* media, int32_t n_media,
                 const upbp_query_batch* q, const upbp_mis_inputs* mis, int which, int accel, int precision,
                 int threads, float* out32, double* out64, uint64_t* counters);

/* ---- adjoint (FP64 analytic gradient of the plain estimators) ----------- */
/* Gradient of L = sum_q <w_photons[q], out_photons[q]> + <w_beams[q], out_beams[q]>
 * (w are the upstream gradients). Brute-force traversal, double arithmetic. Segment
 * attenuations are exp(-sigma_t len); for SHORT queries the earlier segments' pdfs are the
 * pass-through RaySamplePdf exp(-sigma_t_k len) of media that sample distances, and the
 * minimum-sigma_t channel k is held fixed. Reduced beam grids weight each hit by its times
 * counted over the cell's pdf (held fixed). fp32_hits: hit decisions and hit parameters in
 * the reference's FP32 (the hit sets of the FP32 forward and the GPU), the gradient arithmetic
 * in double; else everything in double (the FP64 forward's hit sets). Outputs are overwritten
 * (not accumulated). */
int or_adjoint(const or_photons* ph, const or_beams* bm, const upbp_medium* media, int32_t n_media,
               const upbp_query_batch* q, const double* w_photons, const double* w_beams,
               int threads, double* grad_sigma_s, double* grad_sigma_a,
               double* grad_photon_power, double* grad_beam_power, int fp32_hits);

/* ---- k-nearest-neighbour radii (KdTmpl.hxx:600-792 result semantics) ----- */
/* out[i] = k-th smallest squared distance from q[i] to pts (brute force, bounded
 * max-heap); with n <= k the largest of the n distances. */
void or_knn_d2(const float* pts, int64_t n, const float* q, int64_t nq, int32_t k, int threads, float* out);
/* Bre.cxx:562-602: out[i] = radius * 2 sqrt(d2_k) over in-medium photons (in_medium NULL = all). */
void or_photon_knn_radius(int64_t n, const float* pos, const uint8_t* in_medium, float radius, int32_t knn,
                          int threads, float* out);
/* PhBeams.cxx:77-117: start / end radius of each beam from the k-th nearest beam origin. */
void or_beam_knn_radius(int64_t n, const float* origin, const float* dir, const float* length, float radius,
                        int32_t knn, int threads, float* r0, float* r1);

/* ---- vertex merging through the hashed grid (HashGrid.hxx:38-230, UPBP.hxx:63-133) ---- */
typedef struct or_hgrid or_hgrid;
/* HashGrid::Reserve(cell_count) + Build(vertices, radius, type): type UPBP_PP3D keeps the
 * vertices whose flags hold UPBP_LV_IN_MEDIUM, UPBP_SURF the others; FP32 bounding box,
 * cell size 2r, hashed buckets. */
or_hgrid* or_hgrid_create(int64_t n, const float* pos, const upbp_light_vertex* lv, float radius, int64_t cell_count,
                        uint32_t type);
void or_hgrid_free(or_hgrid* g);
int64_t or_hgrid_count(const or_hgrid* g);
/* bb6 = (min.xyz, max.xyz); indices [count] (mIndices); cell_ends [cell_count] (mCellEnds). */
void or_hgrid_info(const or_hgrid* g, float* bb6, int32_t* indices, int32_t* cell_ends);
/* RangeQuery over HashGrid::Process per camera vertex (PP3D: in a medium; SURF: on a
 * surface). accel 1: the grid's own walk (HashGrid.hxx:130-186); accel 0: every grid
 * vertex within the radius, once per neighbour bucket equal to its own bucket (the same
 * multiset, stated without the grid's bucket ranges). out32/out64 [nq][3] by precision,
 * sig [nq][2] optional (Process calls, sum of or_hash32(vertex)); counters[0] = Process
 * calls, counters[1] = distance tests. Returns 0 or an upbp_status. */
int or_pp3d_eval(const or_hgrid* g, const upbp_medium* media, int32_t n_media, const upbp_pp3d_query* q, int64_t nq,
                 const upbp_mis_inputs* mis, uint32_t flags, int accel, int precision, int threads, float* out32,
                 double* out64, uint32_t* sig, uint64_t* counters);
int or_surf_eval(const or_hgrid* g, const upbp_surf_query* q, int64_t nq, const upbp_mis_inputs* mis, uint32_t flags,
                 int accel, int precision, int threads, float* out32, double* out64, uint32_t* sig,
                 uint64_t* counters);

/* ---- light sub-path tracing in media (lighttrace.c) ---------------------- */
#define LT_MAX_REGIONS UPBP_MAX_REGIONS
/* upbp_trace_light_paths restated (UPBP.hxx:501-674,1592-1662,1783-1870,1908-1969): the
 * same outputs, the same counter-based random numbers, FP32. Returns 0 or an upbp_status
 * (UPBP_EINVAL with out->vertex_count / beam_count set when a capacity is exceeded). */
int or_trace_light_paths(const upbp_medium* media, int32_t n_media, const upbp_region* regions, int32_t n_regions,
                         const upbp_point_light* lights, int32_t n_lights, const upbp_light_trace_params* params,
                         int64_t path_count, int threads, upbp_light_paths* out);
/* The same, and per path the closest call of its branch decisions (distance sample against
 * the segment end, roulette draw against the continuation probability, sphere and box
 * crossings), as the relative distance of the two compared values (INFINITY: no decision). */
int or_trace_light_paths_margins(const upbp_medium* media, int32_t n_media, const upbp_region* regions,
                                 int32_t n_regions, const upbp_point_light* lights, int32_t n_lights,
                                 const upbp_light_trace_params* P, int64_t path_count, int threads,
                                 upbp_light_paths* out, float* margins);

#ifdef __cplusplus
}
#endif

#endif /* UPBP_ORACLE_H_ */
