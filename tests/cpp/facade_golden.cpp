// facade_golden.cpp -- a C++ caller of the drop-in, as a SmallUPBP renderer would be one.
//
// It includes only include/upbp.hpp, links libupbp_hip.so, and drives the estimators through
// the facade classes with the reference's method shapes: EmbreeBre::build / evalBre / destroy
// (SmallUPBP/src/Bre/Bre.hxx:40-184) and PhotonBeamsEvaluator::build / evalBeamBeamEstimate /
// getBeamSelectionPdf / destroy (SmallUPBP/src/Beams/PhBeams.hxx:46-160), both the batched
// QueryBatch path and the one-query synchronous calls of the reference (VolLightTracer.hxx:378,384).
//
// Usage: facade_golden IN OUT
//   IN  (little endian): int64 M, upbp_medium[M];
//       int64 N (photons, 0 = none), float position[N][3], inc_dir[N][3], power[N][3], float radius;
//       int64 B (beams, 0 = none), float origin[B][3], direction[B][3], length[B], throughput[B][3],
//       uint32 flags[B], int32 medium[B], float radius;
//       int64 Q, uint32 beam_type, float origin[Q][3], direction[Q][3], int64 offsets[Q+1],
//       upbp_volume_segment segments[offsets[Q]]
//   OUT: int64 Q, int64 K (single-query calls), int64 photons kept, uint64 photon hits, uint64 beam hits,
//        float batched_photons[Q][3], batched_beams[Q][3], single_photons[K][3], single_beams[K][3],
//        float selection_pdf
// Exit status: 0 ok, 2 bad input file, 3 a library error (message on stderr).
#include <cstdio>
#include <cstring>
#include <vector>

#include "upbp.hpp"

namespace {

struct Reader {
    FILE* f;
    bool ok = true;
    template <class T>
    T one() {
        T v{};
        ok = ok && std::fread(&v, sizeof(T), 1, f) == 1;
        return v;
    }
    template <class T>
    std::vector<T> many(int64_t n) {
        std::vector<T> v((size_t)(n > 0 ? n : 0));
        if (n > 0) ok = ok && std::fread(v.data(), sizeof(T), (size_t)n, f) == (size_t)n;
        return v;
    }
};

template <class T>
void put(FILE* f, const T* p, size_t n) {
    if (n) std::fwrite(p, sizeof(T), n, f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    Reader r{in};
    const int64_t M = r.one<int64_t>();
    std::vector<upbp_medium> media = r.many<upbp_medium>(M);
    const int64_t N = r.one<int64_t>();
    std::vector<float> ph_pos = r.many<float>(3 * N), ph_inc = r.many<float>(3 * N), ph_pow = r.many<float>(3 * N);
    const float ph_r = N ? r.one<float>() : 0.f;
    const int64_t B = r.one<int64_t>();
    std::vector<float> bo = r.many<float>(3 * B), bd = r.many<float>(3 * B), bl = r.many<float>(B),
                       bt = r.many<float>(3 * B);
    std::vector<uint32_t> bf = r.many<uint32_t>(B);
    std::vector<int32_t> bmed = r.many<int32_t>(B);
    const float bm_r = B ? r.one<float>() : 0.f;
    const int64_t Q = r.one<int64_t>();
    const uint32_t beam_type = r.one<uint32_t>();
    std::vector<float> qo = r.many<float>(3 * Q), qd = r.many<float>(3 * Q);
    std::vector<int64_t> off = r.many<int64_t>(Q + 1);
    std::vector<upbp_volume_segment> seg = r.ok && Q >= 0 ? r.many<upbp_volume_segment>(off[(size_t)Q]) :
                                                            std::vector<upbp_volume_segment>();
    std::fclose(in);
    if (!r.ok) {
        std::fprintf(stderr, "truncated input\n");
        return 2;
    }
    const int64_t K = Q < 16 ? Q : 16;
    std::vector<float> out_ph((size_t)Q * 3, 0.f), out_bm((size_t)Q * 3, 0.f);
    std::vector<float> one_ph((size_t)K * 3, 0.f), one_bm((size_t)K * 3, 0.f);
    int64_t kept = 0;
    upbp_eval_stats sp{}, sb{};
    float pdf = 0.f;
    try {
        upbp::Context ctx(0);
        ctx.setMedia(media);
        // the renderer adds every camera ray of an iteration, then evaluates once
        upbp::QueryBatch batch;
        for (int64_t q = 0; q < Q; ++q)
            batch.add(&qo[3 * q], &qd[3 * q], seg.data() + off[q], (size_t)(off[q + 1] - off[q]));
        if (N > 0) {
            upbp::EmbreeBre bre(ctx);
            kept = bre.build(ph_pos.data(), ph_inc.data(), ph_pow.data(), nullptr, (int)N, UPBP_CONSTANT_RADIUS, ph_r,
                             0);
            bre.evalBre(batch, beam_type, out_ph.data(), false, &sp);
            for (int64_t q = 0; q < K; ++q) {
                const auto rgb = bre.evalBre(beam_type, &qo[3 * q], &qd[3 * q], seg.data() + off[q],
                                             (size_t)(off[q + 1] - off[q]));
                std::memcpy(&one_ph[3 * q], rgb.data(), 12);
            }
            bre.destroy();
        }
        if (B > 0) {
            upbp_beam_input bi{};
            bi.count = B;
            bi.origin = bo.data();
            bi.direction = bd.data();
            bi.length = bl.data();
            bi.throughput = bt.data();
            bi.flags = bf.data();
            bi.medium_id = bmed.data();
            bi.memory = UPBP_MEM_HOST;
            upbp::PhotonBeamsEvaluator beams(ctx);
            beams.build(bi, UPBP_CONSTANT_RADIUS, bm_r, 0);
            beams.evalBeamBeamEstimate(batch, beam_type, out_bm.data(), false, &sb);
            for (int64_t q = 0; q < K; ++q) {
                const auto rgb = beams.evalBeamBeamEstimate(beam_type, &qo[3 * q], &qd[3 * q], seg.data() + off[q],
                                                            (size_t)(off[q + 1] - off[q]));
                std::memcpy(&one_bm[3 * q], rgb.data(), 12);
            }
            pdf = beams.getBeamSelectionPdf(&bo[0]);
            beams.destroy();
        }
    } catch (const upbp::Error& e) {
        std::fprintf(stderr, "upbp error %d: %s\n", (int)e.status, e.what());
        return 3;
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    const uint64_t hp = sp.photon_hits, hb = sb.beam_hits;
    put(out, &Q, 1);
    put(out, &K, 1);
    put(out, &kept, 1);
    put(out, &hp, 1);
    put(out, &hb, 1);
    put(out, out_ph.data(), out_ph.size());
    put(out, out_bm.data(), out_bm.size());
    put(out, one_ph.data(), one_ph.size());
    put(out, one_bm.data(), one_bm.size());
    put(out, &pdf, 1);
    std::fclose(out);
    std::printf("facade_golden: %lld queries, %lld photons kept, %llu photon hits, %llu beam hits\n", (long long)Q,
                (long long)kept, (unsigned long long)hp, (unsigned long long)hb);
    return 0;
}
