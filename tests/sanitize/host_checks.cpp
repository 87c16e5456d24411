// Host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer: the binned-SAH
// BVH builder, the procedural scene generator, the environment-map importance builder, the camera
// jitter patterns and the .spd table reader, driven over ordinary, degenerate and malformed inputs.
// Built and run by tests/test_host_sanitizers.py with clang++ -fsanitize=address,undefined; any report
// fails the run. Each check also asserts the result's structure, so a silently wrong answer fails.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "env_map.hpp"
#include "plt_internal.hpp"
#include "sample_pattern.hpp"
#include "scene_host.hpp"

// The library defines these in translation units built for the GPU (context.hip); the checks here
// never reach the profile builder, and errors are only counted.
static int g_errors = 0;
namespace plt {
int set_error(int code, const std::string&) {
    ++g_errors;
    return code;
}
}  // namespace plt
extern "C" int plt_spectral_profile_build(uint32_t, const float*, const float*, int, plt_spectral_profile*) {
    return PLT_ENOTIMPL;
}

#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

// Every triangle in exactly one leaf, leaves and inner references in range, levels within bound.
static void check_tree(const plt::HostBvh& b, size_t nt) {
    if (nt == 0) {
        CHECK(b.nodes.empty() && b.tri_order.empty());
        return;
    }
    CHECK(b.tri_order.size() == nt);
    std::vector<int> seen(nt, 0), visits(b.nodes.size(), 0);
    for (uint32_t t : b.tri_order) {
        CHECK(t < nt);
        ++seen[t];
    }
    for (int s : seen) CHECK(s == 1);
    std::vector<int> placed(nt, 0);
    std::vector<std::pair<int32_t, int>> stack{{0, 1}};
    int levels = 0;
    while (!stack.empty()) {
        auto [n, lvl] = stack.back();
        stack.pop_back();
        CHECK(n >= 0 && size_t(n) < b.nodes.size());
        ++visits[n];
        levels = lvl > levels ? lvl : levels;
        const plt::BvhNode& node = b.nodes[n];
        for (int c = 0; c < 4; ++c) {
            if (std::isinf(node.lox[c])) {
                CHECK(node.child[c] == 0);
                continue;
            }
            const int32_t r = node.child[c];
            if (r < 0) {
                const uint32_t leaf = uint32_t(~r), first = leaf >> 3, cnt = (leaf & 7u) + 1u;
                CHECK(first + cnt <= nt);
                for (uint32_t k = first; k < first + cnt; ++k) ++placed[k];
            } else {
                stack.push_back({r, lvl + 1});
            }
        }
    }
    for (int v : visits) CHECK(v == 1);
    for (int p : placed) CHECK(p == 1);
    CHECK(levels == b.max_depth);
    CHECK(levels <= 32);
}

static void soup(size_t n, uint32_t seed, float spread, std::vector<float>& pos, std::vector<uint32_t>& idx) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    pos.resize(9 * n);
    idx.resize(3 * n);
    for (size_t t = 0; t < n; ++t) {
        const float c[3] = {spread * u(rng), spread * u(rng), spread * u(rng)};
        for (int v = 0; v < 3; ++v) {
            for (int a = 0; a < 3; ++a) pos[9 * t + 3 * v + a] = c[a] + 0.3f * u(rng);
            idx[3 * t + v] = uint32_t(3 * t + v);
        }
    }
}

static void bvh_checks() {
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    for (size_t n : {size_t(0), size_t(1), size_t(3), size_t(9), size_t(1000), size_t(40000)}) {
        soup(n, uint32_t(n) + 1u, 10.f, pos, idx);
        check_tree(plt::build_bvh(pos, idx, 4), n);
    }
    // every triangle the same (no centroid extent), and all collapsed to one point
    soup(2000, 7, 0.f, pos, idx);
    for (size_t i = 9; i < pos.size(); ++i) pos[i] = pos[i % 9];
    check_tree(plt::build_bvh(pos, idx, 4), 2000);
    std::fill(pos.begin(), pos.end(), 1.5f);
    check_tree(plt::build_bvh(pos, idx, 4), 2000);
    // sizes and positions growing geometrically: the depth bound's median splits
    const size_t n = 300;
    pos.assign(9 * n, 0.f);
    idx.resize(3 * n);
    for (size_t t = 0; t < n; ++t) {
        const float s = std::pow(1.1f, float(t));
        const float v[9] = {s, 0, 0, s, 0.5f * s, 0, s, 0, 0.5f * s};
        for (int k = 0; k < 9; ++k) pos[9 * t + k] = v[k];
        for (int k = 0; k < 3; ++k) idx[3 * t + k] = uint32_t(3 * t + k);
    }
    check_tree(plt::build_bvh(pos, idx, 4), n);
}

static void scene_checks() {
    for (uint32_t spheres : {0u, 1u, 5u}) {
        plt_procedural_scene_desc d{};
        d.seed = spheres + 3;
        d.sphere_count = spheres;
        d.subdivision = 2;
        d.material_kinds = 8;
        d.ground_size = 20.f;
        d.aspect = 1.5f;
        plt::HostScene s;
        CHECK(plt::build_procedural_scene(d, s) == PLT_OK);
        const size_t nt = s.indices.size() / 3;
        CHECK(nt == 4 + size_t(spheres) * 20 * 16);
        for (uint32_t i : s.indices) CHECK(i < s.positions.size() / 3);
        check_tree(s.bvh, nt);
    }
}

static void env_checks() {
    for (uint32_t w : {1u, 7u, 256u}) {
        plt::HostEnvMap e;
        e.width = w;
        e.height = w > 1 ? w / 2 + 1 : 1;
        e.rgba.assign(size_t(e.width) * e.height * 4, 0.f);
        std::mt19937 rng(w);
        std::exponential_distribution<float> x(1.f);
        for (float& v : e.rgba) v = x(rng);
        plt::build_env_importance(e);
        CHECK(!e.importance.empty() && !e.mip_offset.empty());
        for (float v : e.importance) CHECK(std::isfinite(v) && v >= 0.f);
    }
}

static void pattern_checks() {
    for (int type : {0, 1, 2, 3}) {
        for (uint32_t count : {1u, 2u, 8u, 16u, 17u}) {
            plt::SamplePattern p;
            if (p.configure(type, count) != PLT_OK) continue;
            for (int i = 0; i < 64; ++i) {
                float xy[2];
                p.next(xy);
                CHECK(xy[0] >= -0.5f && xy[0] < 0.5f && xy[1] >= -0.5f && xy[1] < 0.5f);
            }
        }
    }
    std::vector<float> xy(2 * 40);
    CHECK(plt_sample_pattern_generate(2, 16, 40, xy.data()) == PLT_OK);
}

static void spd_checks(const std::string& tables) {
    std::vector<float> wl(4096), v(4096);
    uint32_t count = 0;
    for (const char* f : {"/ior/Cu.eta.spd", "/ior/Au.k.spd", "/emission/FL4.spd"}) {
        CHECK(plt_spectrum_load_spd((tables + f).c_str(), 4096, wl.data(), v.data(), &count) == PLT_OK);
        CHECK(count >= 2);
        for (uint32_t i = 1; i < count; ++i) CHECK(wl[i] > wl[i - 1]);
    }
    // a capacity smaller than the table, a missing file and malformed text
    CHECK(plt_spectrum_load_spd((tables + "/ior/Cu.eta.spd").c_str(), 2, wl.data(), v.data(), &count) < 0);
    CHECK(plt_spectrum_load_spd((tables + "/no/such.spd").c_str(), 4096, wl.data(), v.data(), &count) < 0);
    const std::string bad = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") + "/plt_sanitize_bad.spd";
    for (const char* text : {"", "#only a comment\n", "400 1.0\n410", "400 1.0\nabc def\n500 2\n", "   \n\t\n400\t1e39\n",
                             "400 nan\n410 -inf\n"}) {
        FILE* f = std::fopen(bad.c_str(), "wb");
        CHECK(f);
        std::fputs(text, f);
        std::fclose(f);
        (void)plt_spectrum_load_spd(bad.c_str(), 4096, wl.data(), v.data(), &count);
    }
    std::remove(bad.c_str());
}

int main(int argc, char** argv) {
    CHECK(argc == 2);
    bvh_checks();
    scene_checks();
    env_checks();
    pattern_checks();
    spd_checks(argv[1]);
    std::printf("host checks passed (%d errors reported by the library, as expected)\n", g_errors);
    return 0;
}
