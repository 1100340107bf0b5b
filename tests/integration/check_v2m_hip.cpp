// check_v2m_hip.cpp -- the v2mesh-side binding (trussnet_amd/integration/v2m_hip.cpp)
// compiled against the reference's own headers and linked with the reference's CPU
// path (oracle/_ref/libv2m_ref.so) and libtrussnet_amd.
//
// Without a device every *_hip call must throw std::runtime_error, and the caller's
// fallback (restore the seeds, run the CPU path) must give relax_cpu's result, as
// run_pipeline does (src/v2m_pipeline.cpp:497-526). With a device the HIP results
// must equal the CPU path's bit for bit: build_grid_hip (prob_fields included, on a
// tissue-probability input) against build_grid_cpu, seed_hip against seed_cpu,
// relax_hip and relax_hip_batch against relax_cpu, thin_hip against thin_nodes on the
// seeds, protect_hip against protect_features (a grid with feature curves).
//
// Exit status 0 and a last line "ok <mode>" when every check passed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "v2m_hip.h"

using namespace tn;

static int failures = 0;

#define CHECK(cond, what)                                          \
    do {                                                           \
        if (!(cond)) {                                             \
            std::printf("FAIL %s (line %d)\n", what, __LINE__);    \
            ++failures;                                            \
        }                                                          \
    } while (0)

template <class T>
static bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

static bool same_nodes(const Nodes& a, const Nodes& b) {
    return same(a.P, b.P) && same(a.lab, b.lab) && same(a.typ, b.typ) && same(a.part, b.part);
}

// two nested spheres in an n^3 volume (labels 1 and 2), optionally as probabilities
static LabelVolume spheres(int n, bool tpm) {
    LabelVolume lv;
    lv.nx = lv.ny = lv.nz = n;
    lv.data.resize((size_t)n * n * n);
    const float c = 0.5f * (n - 1), r1 = 0.38f * n, r2 = 0.2f * n;
    if (tpm) {
        lv.nprob = 3;
        lv.prob.resize(3 * lv.data.size());
    }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const size_t v = ((size_t)z * n + y) * n + x;
                const float d = std::sqrt((x - c) * (x - c) + (y - c) * (y - c) + (z - c) * (z - c));
                lv.data[v] = d < r2 ? 2 : d < r1 ? 1 : 0;
                if (tpm) {   // softmax of three smooth logits; data = their argmax
                    const float l[3] = { 1.5f * (d - r1), 1.5f * std::fabs(d - 0.5f * (r1 + r2)) * -1.0f + 2.0f,
                                         1.5f * (r2 - d) };
                    const float m = std::max(l[0], std::max(l[1], l[2]));
                    float e[3], s = 0.0f;
                    for (int k = 0; k < 3; ++k) {
                        e[k] = std::exp(l[k] - m);
                        s += e[k];
                    }
                    int best = 0;
                    for (int k = 0; k < 3; ++k) {
                        lv.prob[(size_t)k * lv.data.size() + v] = e[k] / s;
                        best = e[k] > e[best] ? k : best;
                    }
                    lv.data[v] = (uint16_t)best;
                }
            }
    lv.maxlabel = 2;
    return lv;
}

static bool run_case(const char* name, bool tpm, bool device) {
    const LabelVolume lv = spheres(40, tpm);
    GridParams gp;
    Grid g;
    build_grid_cpu(lv, gp, g);
    CHECK(g.prob_fields == tpm, "build_grid_cpu prob_fields");
    RelaxParams prm;
    prm.max_iters = 20;
    prm.dptol = 0.0f;
    Nodes seeds;
    seed_cpu(g, prm, seeds);
    Nodes want = seeds;
    RelaxStats sw;
    relax_cpu(g, prm, want, sw);

    // the grid
    Grid gh;
    bool threw = false;
    try {
        build_grid_hip(lv, gp, gh, -1);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    if (!device) {
        CHECK(threw, "build_grid_hip without a device throws");
    } else {
        CHECK(!threw, "build_grid_hip");
        CHECK(gh.prob_fields == g.prob_fields, "build_grid_hip prob_fields");
        CHECK(same(gh.bl_cnt, g.bl_cnt) && same(gh.bl_lab, g.bl_lab) && same(gh.bl_slot, g.bl_slot), "bricks");
        CHECK(same(gh.slot_brick, g.slot_brick) && same(gh.slot_label, g.slot_label), "slots");
        CHECK(same(gh.phi, g.phi) && same(gh.h, g.h) && same(gh.grade, g.grade), "fields");
        CHECK(gh.hmin == g.hmin && gh.hmax == g.hmax && gh.hbase == g.hbase && gh.R == g.R, "scalars");
    }

    // seeding
    Nodes sh;
    threw = false;
    try {
        seed_hip(g, prm, sh, -1);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    if (!device) {
        CHECK(threw, "seed_hip without a device throws");
    } else {
        CHECK(!threw && same_nodes(sh, seeds) && same(sh.part3, seeds.part3), "seed_hip == seed_cpu");
    }

    // the relaxation with the pipeline's fallback
    Nodes nd = seeds;
    RelaxStats st;
    bool used_gpu = false;
    {
        const Nodes keep = nd;   // restored if the device fails part-way
        try {
            relax_hip(g, prm, nd, st, -1);
            used_gpu = true;
        } catch (const std::exception& e) {
            nd = keep;
        }
        if (!used_gpu) {
            relax_cpu(g, prm, nd, st);
        }
    }
    CHECK(used_gpu == device, "relax_hip ran where a device is");
    CHECK(same_nodes(nd, want), "relaxed nodes == relax_cpu");
    CHECK(st.iters == sw.iters && st.rebuilds == sw.rebuilds && st.fire_resets == sw.fire_resets, "stats");
    CHECK(st.n_interior == sw.n_interior && st.n_interface == sw.n_interface && st.n_junction == sw.n_junction, "types");

    // thinning of the seeds, as run_pipeline does before the relaxation
    for (float B : { 0.6f, 1.1f }) {
        RelaxParams tp = prm;
        tp.thin = B;
        Nodes tw = seeds, th = seeds;
        const size_t rw = thin_nodes(g, tp, tw);
        threw = false;
        size_t rh = 0;
        try {
            rh = thin_hip(g, tp, th, -1);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!device) {
            CHECK(threw, "thin_hip without a device throws");
        } else {
            CHECK(!threw && rh == rw && same_nodes(th, tw) && same(th.part3, tw.part3), "thin_hip == thin_nodes");
        }
    }
    // feature protection: pinned nodes along a segment through the volume, with
    // interface nodes between them (the grid's own feature curves stand in for a
    // shape input's)
    if (device) {
        Grid gf = g;
        const float a[3] = { 6.0f, 20.0f, 20.0f }, b[3] = { 34.0f, 20.0f, 20.0f };
        gf.feat = { 2.0f, a[0], a[1], a[2], b[0], b[1], b[2] };
        Nodes fn = want;
        for (int k = 0; k <= 14; ++k) {   // pins every 2 mm along the segment
            fn.P.insert(fn.P.end(), { a[0] + 2.0f * k, a[1], a[2] });
            fn.lab.push_back(1);
            fn.typ.push_back(3);
            fn.part.insert(fn.part.end(), { 2, 0 });
            fn.part3.push_back((uint16_t)0xFFFF);
        }
        for (int k = 0; k < 14; ++k) {   // and a node between each pair
            fn.P.insert(fn.P.end(), { a[0] + 2.0f * k + 1.0f, a[1] + 0.3f, a[2] });
            fn.lab.push_back(1);
            fn.typ.push_back(1);
            fn.part.insert(fn.part.end(), { 2, 0 });
            fn.part3.push_back((uint16_t)0xFFFF);
        }
        if (fn.part3.size() != fn.typ.size()) {
            fn.part3.resize(fn.typ.size(), (uint16_t)0xFFFF);
        }
        Nodes fw = fn, fh = fn;
        const size_t pw = protect_features(gf, fw);
        const size_t ph = protect_hip(gf, fh, -1);
        CHECK(pw > 0 && ph == pw && same_nodes(fh, fw) && same(fh.part3, fw.part3), "protect_hip == protect_features");
    }

    // the batch: the same volume twice
    if (device) {
        Nodes a = seeds, b = seeds;
        std::vector<const Grid*> gs{ &g, &g };
        std::vector<Nodes*> ns{ &a, &b };
        std::vector<RelaxStats> ss;
        relax_hip_batch(gs, prm, ns, ss, -1);
        CHECK(same_nodes(a, want) && same_nodes(b, want) && ss.size() == 2, "relax_hip_batch == relax_cpu");
    }
    std::printf("%s: %zu seeds, %d iterations, %s\n", name, seeds.size(), st.iters, used_gpu ? "device" : "fallback");
    return failures == 0;
}

int main(int argc, char** argv) {
    const bool device = argc > 1 && std::string(argv[1]) == "device";
    run_case("spheres40", false, device);
    run_case("spheres40-tpm", true, device);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ok %s\n", device ? "device" : "fallback");
    return 0;
}
