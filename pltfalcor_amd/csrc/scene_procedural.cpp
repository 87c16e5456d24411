// Procedural workload scenes: a ground quad, icospheres on it, one quad area light, a pinhole
// camera, and PLT materials whose binary16 parameter blocks are filled the way the reference's
// material setters fill BasicMaterialData (Rendering/Materials/PLT/*Material.cpp).
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <utility>
#include "scene_host.hpp"

namespace plt {

uint16_t f32_to_h(float f) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(f)); }
float h_to_f32(uint16_t h) { return static_cast<float>(__builtin_bit_cast(_Float16, h)); }

namespace {

// UniformSampleGenerator twin for host-side scene layout draws (SplitMix64 -> xoshiro128**).
struct HostRng {
    uint32_t s[4];
    static uint32_t spread(uint32_t x) {
        x &= 0xffffu;
        x = (x | (x << 8)) & 0x00FF00FFu; x = (x | (x << 4)) & 0x0F0F0F0Fu;
        x = (x | (x << 2)) & 0x33333333u; x = (x | (x << 1)) & 0x55555555u;
        return x;
    }
    HostRng(uint32_t px, uint32_t py, uint32_t sn) {
        uint64_t st = (uint64_t(sn) << 32) | uint64_t(spread(px) | (spread(py) << 1));
        uint64_t v[2];
        for (int i = 0; i < 2; ++i) {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            v[i] = z ^ (z >> 31);
        }
        s[0] = uint32_t(v[0]); s[1] = uint32_t(v[0] >> 32); s[2] = uint32_t(v[1]); s[3] = uint32_t(v[1] >> 32);
    }
    float next1D() {
        const uint32_t m = s[0] * 5u;
        const uint32_t r = ((m << 7) | (m >> 25)) * 9u;
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = (s[3] << 11) | (s[3] >> 21);
        return float(r >> 8) * 0x1p-24f;
    }
};

struct Ico {
    std::vector<double> v;       // 3 per vertex (unit sphere)
    std::vector<uint32_t> f;     // 3 per face, outward CCW
};

Ico make_icosphere(uint32_t subdiv) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    Ico m;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : base) {
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        m.v.insert(m.v.end(), {p[0] / l, p[1] / l, p[2] / l});
    }
    m.f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
           3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double p[3];
            for (int k = 0; k < 3; ++k) p[k] = 0.5 * (m.v[3 * a + k] + m.v[3 * b + k]);
            const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            const uint32_t id = uint32_t(m.v.size() / 3);
            m.v.insert(m.v.end(), {p[0] / l, p[1] / l, p[2] / l});
            mid.emplace(key, id);
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(m.f.size() * 4);
        for (size_t i = 0; i < m.f.size(); i += 3) {
            const uint32_t a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        m.f.swap(nf);
    }
    return m;
}

plt_material blank(uint32_t type) {
    plt_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = type;
    for (int i = 0; i < 4; ++i) m.base_color[i] = f32_to_h(1.f);  // BasicMaterialData defaults
    for (int i = 0; i < 3; ++i) m.transmission[i] = f32_to_h(1.f);
    m.ior = f32_to_h(1.5f);
    m.emission_profile = m.ior_n_profile = m.ior_k_profile = PLT_NO_PROFILE;
    return m;
}

// PLTDielectricMaterial: constructor (ext IoR 1, Abbe 100 at the default IoR 1.5), then IoR and
// roughness setters (PLTDielectricMaterial.cpp:46-101).
plt_material dielectric(float ior, float roughness) {
    plt_material m = blank(PLT_MAT_DIELECTRIC);
    m.specular[3] = f32_to_h(1.f);
    m.specular[0] = f32_to_h(.52f * (h_to_f32(m.ior) - 1.f) / 100.f);
    m.ior = f32_to_h(ior);
    m.specular[1] = f32_to_h(roughness);
    return m;
}
// PLTConductorMaterial (PLTConductorMaterial.cpp:45-85): ext IoR 1, gamma 3, n,k profiles.
plt_material conductor(float roughness, uint32_t pn, uint32_t pk) {
    plt_material m = blank(PLT_MAT_CONDUCTOR);
    m.ior = f32_to_h(1.f);
    for (int i = 0; i < 4; ++i) m.specular[i] = f32_to_h(1.f);
    m.specular[1] = f32_to_h(roughness);
    m.specular[0] = f32_to_h(3.f);
    m.ior_n_profile = pn;
    m.ior_k_profile = pk;
    return m;
}
plt_material diffuse(float r, float g, float b, uint32_t emission) {
    plt_material m = blank(PLT_MAT_DIFFUSE);
    m.base_color[0] = f32_to_h(r); m.base_color[1] = f32_to_h(g); m.base_color[2] = f32_to_h(b);
    m.emission_profile = emission;
    return m;
}
// PLTCoatedOpaqueDielectricMaterial (PLTCoatedOpaqueDielectricMaterial.cpp:45-105).
plt_material coated_opaque(float r, float g, float b, float roughness, float coat_ior, float thickness) {
    plt_material m = blank(PLT_MAT_COATED_OPAQUE_DIELECTRIC);
    m.base_color[0] = f32_to_h(r); m.base_color[1] = f32_to_h(g); m.base_color[2] = f32_to_h(b);
    m.ior = f32_to_h(1.f);
    m.specular[1] = f32_to_h(roughness);
    m.data1[3] = f32_to_h(coat_ior);
    m.data1[0] = f32_to_h(thickness);
    m.specular[2] = f32_to_h(1.f);
    return m;
}
// PLTDiffractionGratedConductorMaterial (PLTDiffractionGratedConductorMaterial.cpp:45-147):
// sinusoidal 1-D grating, 3 lobes, height 1 x scale 0.5 um, power 1.
plt_material grated_conductor(float pitch, float direction, uint32_t pn, uint32_t pk) {
    plt_material m = blank(PLT_MAT_GRATED_CONDUCTOR);
    m.ior = f32_to_h(1.f);
    m.data1[0] = f32_to_h(pitch);
    m.data1[1] = f32_to_h(0.f);
    m.data1[2] = f32_to_h(direction);
    m.data2[0] = f32_to_h(1.f);
    m.specular[0] = f32_to_h(.5f);
    m.specular[1] = f32_to_h(1.f);
    m.data1[3] = uint16_t((0u << 8) + 3u);  // asfloat16(type << 8 | lobes)
    m.data2[1] = f32_to_h(0.f);
    m.ior_n_profile = pn;
    m.ior_k_profile = pk;
    return m;
}

// PLTCoatedConductorMaterial (PLTCoatedConductorMaterial.cpp:45-110): specular colour 1, ext IoR 1,
// coat thickness scale 1, then roughness, coat IoR and coat thickness (um) setters.
plt_material coated_conductor(float roughness, float coat_ior, float thickness, uint32_t pn, uint32_t pk) {
    plt_material m = blank(PLT_MAT_COATED_CONDUCTOR);
    for (int i = 0; i < 4; ++i) m.base_color[i] = f32_to_h(1.f);
    m.ior = f32_to_h(1.f);
    m.specular[1] = f32_to_h(roughness);
    m.specular[2] = f32_to_h(1.f);
    m.data1[3] = f32_to_h(coat_ior);
    m.data1[0] = f32_to_h(thickness);
    m.ior_n_profile = pn;
    m.ior_k_profile = pk;
    return m;
}
// PLTThinDielectricMaterial (PLTThinDielectricMaterial.cpp:45-90): IoR, thickness (mm) in
// specular.w, birefringence in baseColor.x with the scale specular.x set to 1, optic axis
// (normalised, z >= 0) xy in baseColor.yz, baseColor.a = 1 (opacity 0.5, not masked).
plt_material thin_dielectric(float ior, float thickness_mm, float birefringence, float ax, float ay, float az) {
    plt_material m = blank(PLT_MAT_THIN_DIELECTRIC);
    m.ior = f32_to_h(ior);
    const float l = std::sqrt(ax * ax + ay * ay + az * az);
    const float s = az < 0.f ? -1.f / l : 1.f / l;
    m.base_color[0] = f32_to_h(birefringence);
    m.base_color[1] = f32_to_h(ax * s);
    m.base_color[2] = f32_to_h(ay * s);
    m.base_color[3] = f32_to_h(1.f);
    m.specular[0] = f32_to_h(1.f);
    m.specular[3] = f32_to_h(thickness_mm);
    return m;
}
// PLTMultiLayeredStackMaterial (PLTMultiLayeredStackMaterial.cpp:45-115): base colour 1, ext IoR 1,
// layer IoRs in data2.xy, thicknesses (um) in data1.xy, layer count as raw bits in data1.w.
plt_material multilayer(float ior1, float ior2, float t1, float t2, uint32_t layers) {
    plt_material m = blank(PLT_MAT_MULTILAYER_STACK);
    for (int i = 0; i < 4; ++i) m.base_color[i] = f32_to_h(1.f);
    m.ior = f32_to_h(1.f);
    m.data2[0] = f32_to_h(ior1);
    m.data2[1] = f32_to_h(ior2);
    m.data1[0] = f32_to_h(t1);
    m.data1[1] = f32_to_h(t2);
    m.data1[3] = uint16_t(layers);
    return m;
}
// PLTOpaqueDielectricMaterial (PLTOpaqueDielectricMaterial.cpp): albedo and roughness.
plt_material opaque_dielectric(float r, float g, float b, float roughness) {
    plt_material m = blank(PLT_MAT_OPAQUE_DIELECTRIC);
    m.base_color[0] = f32_to_h(r); m.base_color[1] = f32_to_h(g); m.base_color[2] = f32_to_h(b);
    m.base_color[3] = f32_to_h(1.f);
    m.ior = f32_to_h(1.f);
    m.specular[1] = f32_to_h(roughness);
    return m;
}

void add_quad(HostScene& s, const float c[4][3], float nrm_y, uint32_t material) {
    const uint32_t mesh = uint32_t(s.mesh_first_tri.size());
    const uint32_t v0 = uint32_t(s.positions.size() / 3);
    for (int i = 0; i < 4; ++i) {
        s.positions.insert(s.positions.end(), {c[i][0], c[i][1], c[i][2]});
        s.normals.insert(s.normals.end(), {0.f, h_to_f32(f32_to_h(nrm_y)), 0.f});
    }
    s.mesh_first_tri.push_back(uint32_t(s.indices.size() / 3));
    s.mesh_material.push_back(material);
    s.indices.insert(s.indices.end(), {v0 + 0, v0 + 1, v0 + 2, v0 + 0, v0 + 2, v0 + 3});
    s.tri_mesh.insert(s.tri_mesh.end(), {mesh, mesh});
}
}  // namespace

int build_procedural_scene(const plt_procedural_scene_desc& d, HostScene& s) {
    if (d.subdivision > 7 || d.sphere_count > 65535 || d.material_kinds < 1 || d.material_kinds > 8 ||
        !(d.ground_size > 0.f) || !(d.aspect > 0.f))
        return PLT_EINVAL;

    // Spectral profiles: 0/1 = synthetic Drude metal n,k (64 samples over 380-780 nm; silver-like
    // plasma 9 eV, damping 0.05 eV), 2 = 6500 K blackbody emission (5 nm table, x1e-4).
    {
        const int np = 64;
        std::vector<float> wl(np), n(np), k(np);
        for (int i = 0; i < np; ++i) {
            const double l = 380.0 + 400.0 * double(i) / double(np - 1);
            const double w = 1239.841984 / l;
            const std::complex<double> eps = 1.0 - 81.0 / (w * w + std::complex<double>(0.0, 0.05 * w));
            const std::complex<double> nk = std::sqrt(eps);
            wl[i] = float(l);
            n[i] = float(nk.real());
            k[i] = float(std::fabs(nk.imag()));
        }
        plt_spectral_profile p;
        plt_spectral_profile_build(np, wl.data(), n.data(), 0, &p);
        s.profiles.push_back(p);
        plt_spectral_profile_build(np, wl.data(), k.data(), 0, &p);
        s.profiles.push_back(p);
        std::vector<float> bw, bv;
        const double c = 299792458., h = 6.62606957e-34, kb = 1.3806488e-23, T = 6500.0;
        for (double l = 380.0; l <= 780.0 + 1e-9; l += 5.0) {
            const double lm = l * 1e-9;
            const double le = (2. * h * c * c) / ((lm * lm * lm * lm * lm) * (std::exp((h * c) / (lm * kb * T)) - 1.) * 1e9);
            bw.push_back(float(l));
            bv.push_back(float(le * 1e-4));
        }
        plt_spectral_profile_build(uint32_t(bw.size()), bw.data(), bv.data(), 1, &p);
        s.profiles.push_back(p);
    }

    // Materials: 0 ground, 1 light, 2 + i sphere i.
    s.materials.push_back(diffuse(.5f, .5f, .5f, PLT_NO_PROFILE));
    s.materials.push_back(diffuse(.8f, .8f, .8f, 2));

    const float g = 0.5f * d.ground_size;
    const float ground[4][3] = {{-g, 0.f, -g}, {g, 0.f, -g}, {g, 0.f, g}, {-g, 0.f, g}};
    {
        // Upward winding: (0,3,2), (0,2,1)
        const uint32_t mesh = 0, v0 = 0;
        for (int i = 0; i < 4; ++i) {
            s.positions.insert(s.positions.end(), {ground[i][0], ground[i][1], ground[i][2]});
            s.normals.insert(s.normals.end(), {0.f, 1.f, 0.f});
        }
        s.mesh_first_tri.push_back(0);
        s.mesh_material.push_back(0);
        s.indices.insert(s.indices.end(), {v0 + 0, v0 + 3, v0 + 2, v0 + 0, v0 + 2, v0 + 1});
        s.tri_mesh.insert(s.tri_mesh.end(), {mesh, mesh});
    }

    const Ico ico = make_icosphere(d.subdivision);
    const uint32_t nv = uint32_t(ico.v.size() / 3), nf = uint32_t(ico.f.size() / 3);
    const float lim = 0.9f * g;
    for (uint32_t i = 0; i < d.sphere_count; ++i) {
        HostRng rng(i, 1u, d.seed);
        const float ux = rng.next1D(), uz = rng.next1D(), ur = rng.next1D(), uk = rng.next1D();
        const float p0 = rng.next1D(), p1 = rng.next1D(), p2 = rng.next1D(), p3 = rng.next1D(), p4 = rng.next1D();
        const float cx = -lim + 2.f * lim * ux, cz = -lim + 2.f * lim * uz;
        const float r = 0.2f + 0.8f * ur;
        const float cy = r;
        const uint32_t kind = std::min(uint32_t(uk * float(d.material_kinds)), d.material_kinds - 1);
        switch (kind) {
            case 0: s.materials.push_back(dielectric(1.3f + 0.5f * p0, 0.1f * p1)); break;
            case 1: s.materials.push_back(conductor(0.3f * p0, 0, 1)); break;
            case 2: s.materials.push_back(coated_opaque(.2f + .7f * p0, .2f + .7f * p1, .2f + .7f * p2, .05f + .25f * p3, 1.5f,
                                                        1.f + 4.f * p4)); break;
            case 3: s.materials.push_back(grated_conductor(0.5f + 1.5f * p0, 3.14159265f * p1, 0, 1)); break;
            // thickness U[0.2, 2] um (in mm), birefringence U[0.01, 0.2], IoR U[1.3, 1.8]
            case 4: s.materials.push_back(thin_dielectric(1.3f + .5f * p3, 2e-4f + 1.8e-3f * p0, .01f + .19f * p1,
                                                          .3f * (2.f * p2 - 1.f), .3f * (2.f * p4 - 1.f), 1.f)); break;
            // 20 layers of 1.38 / 2.3, quarter-wave at 550 nm
            case 5: s.materials.push_back(multilayer(1.38f, 2.3f, .55f / (4.f * 1.38f), .55f / (4.f * 2.3f), 20)); break;
            case 6: s.materials.push_back(coated_conductor(.05f + .25f * p0, 1.5f, 1.f + 4.f * p1, 0, 1)); break;
            default: s.materials.push_back(opaque_dielectric(.2f + .7f * p0, .2f + .7f * p1, .2f + .7f * p2, .05f + .3f * p3)); break;
        }
        const uint32_t mesh = uint32_t(s.mesh_first_tri.size());
        const uint32_t vbase = uint32_t(s.positions.size() / 3);
        s.mesh_first_tri.push_back(uint32_t(s.indices.size() / 3));
        s.mesh_material.push_back(2 + i);
        for (uint32_t v = 0; v < nv; ++v) {
            const double* u = &ico.v[3 * size_t(v)];
            s.positions.insert(s.positions.end(), {float(cx + r * u[0]), float(cy + r * u[1]), float(cz + r * u[2])});
            s.normals.insert(s.normals.end(), {h_to_f32(f32_to_h(float(u[0]))), h_to_f32(f32_to_h(float(u[1]))),
                                               h_to_f32(f32_to_h(float(u[2])))});
        }
        for (uint32_t f = 0; f < nf; ++f) {
            s.indices.insert(s.indices.end(), {vbase + ico.f[3 * f], vbase + ico.f[3 * f + 1], vbase + ico.f[3 * f + 2]});
            s.tri_mesh.push_back(mesh);
        }
    }

    // Light: 3 x 3 quad at y = 7 facing down, winding (0,1,2), (0,2,3).
    const float light[4][3] = {{-1.5f, 7.f, -1.5f}, {1.5f, 7.f, -1.5f}, {1.5f, 7.f, 1.5f}, {-1.5f, 7.f, 1.5f}};
    add_quad(s, light, -1.f, 1);

    // Camera
    plt_camera& c = s.camera;
    const float pos[3] = {0.f, 5.f, 16.f}, tgt[3] = {0.f, 0.6f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    std::memcpy(c.pos, pos, sizeof(pos));
    std::memcpy(c.target, tgt, sizeof(tgt));
    std::memcpy(c.up, up, sizeof(up));
    c.fov_y = 45.f * 3.14159265358979f / 180.f;
    c.aspect = d.aspect;
    c.focal_distance = 1.f;
    c.jitter[0] = c.jitter[1] = 0.f;

    s.bvh = build_bvh(s.positions, s.indices, 4);
    return PLT_OK;
}

}  // namespace plt
