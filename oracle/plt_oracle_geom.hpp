// PLT oracle, geometry part: ray/triangle intersection, closest- and any-hit queries (brute force
// and over the oracle's own median-split BVH), ray-origin offsetting and pinhole camera rays.
//
// TEST INFRASTRUCTURE ONLY (see plt_oracle.hpp).
//
// The reference's ray casting is DXR hardware traversal (Scene/Scene.cpp:389-407), which has no
// CPU form. Hit ids are pinned against brute force with the published watertight test of Woop,
// Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013 (FP32, double
// fallback on zero edge functions), accepting tmin <= t <= tmax; equal distances resolve to the
// smaller triangle id. The result does not depend on any acceleration structure.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace pltoracle {

struct ORay { float o[3], d[3], tmin, tmax; };
struct OHit { uint32_t tri = 0xFFFFFFFFu; float b1 = 0, b2 = 0, t = 0; };

struct WoopPrep {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
inline WoopPrep woop_prep(const ORay& r) {
    WoopPrep p;
    const float ax = std::fabs(r.d[0]), ay = std::fabs(r.d[1]), az = std::fabs(r.d[2]);
    p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz == 2 ? 0 : p.kz + 1;
    p.ky = p.kx == 2 ? 0 : p.kx + 1;
    if (r.d[p.kz] < 0.f) std::swap(p.kx, p.ky);
    p.Sx = r.d[p.kx] / r.d[p.kz];
    p.Sy = r.d[p.ky] / r.d[p.kz];
    p.Sz = 1.f / r.d[p.kz];
    return p;
}
inline bool woop_intersect(const ORay& r, const WoopPrep& p, const float* v0, const float* v1, const float* v2, float tmin,
                           float tmax, float& t_out, float& b1, float& b2) {
    const float Ak = v0[p.kx] - r.o[p.kx], Al = v0[p.ky] - r.o[p.ky], Az0 = v0[p.kz] - r.o[p.kz];
    const float Bk = v1[p.kx] - r.o[p.kx], Bl = v1[p.ky] - r.o[p.ky], Bz0 = v1[p.kz] - r.o[p.kz];
    const float Ck = v2[p.kx] - r.o[p.kx], Cl = v2[p.ky] - r.o[p.ky], Cz0 = v2[p.kz] - r.o[p.kz];
    const float Ax = Ak - p.Sx * Az0, Ay = Al - p.Sy * Az0;
    const float Bx = Bk - p.Sx * Bz0, By = Bl - p.Sy * Bz0;
    const float Cx = Ck - p.Sx * Cz0, Cy = Cl - p.Sy * Cz0;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = p.Sz * Az0, Bz = p.Sz * Bz0, Cz = p.Sz * Cz0;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    b1 = V / det;
    b2 = W / det;
    return true;
}

struct OMesh {
    const float* pos = nullptr;       // 3 per vertex
    const uint32_t* idx = nullptr;    // 3 per triangle
    uint32_t ntri = 0;
    const float* vtx(uint32_t t, int k) const { return pos + 3 * size_t(idx[3 * size_t(t) + k]); }
};

// Closest hit by testing every triangle.
inline OHit closest_brute(const OMesh& m, const ORay& r) {
    OHit h;
    h.t = r.tmax;
    const WoopPrep p = woop_prep(r);
    for (uint32_t t = 0; t < m.ntri; ++t) {
        float tt, b1, b2;
        if (woop_intersect(r, p, m.vtx(t, 0), m.vtx(t, 1), m.vtx(t, 2), r.tmin, h.t, tt, b1, b2)) {
            if (tt < h.t || t < h.tri) { h.tri = t; h.t = tt; h.b1 = b1; h.b2 = b2; }
        }
    }
    return h;
}

// Median-split BVH (leaf <= 4), boxes widened so the double-precision slab test never rejects a box
// that contains a hit. Only speed depends on it.
struct OBvh {
    struct Node { double lo[3], hi[3]; int32_t left, right; uint32_t start, count; };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    void build(const OMesh& m) {
        order.resize(m.ntri);
        for (uint32_t i = 0; i < m.ntri; ++i) order[i] = i;
        nodes.clear();
        nodes.reserve(2 * size_t(m.ntri) / 3 + 2);
        if (m.ntri) rec(m, 0, m.ntri);
    }
    int32_t rec(const OMesh& m, uint32_t s, uint32_t c) {
        const int32_t me = int32_t(nodes.size());
        nodes.push_back(Node{});
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t i = s; i < s + c; ++i)
            for (int a = 0; a < 3; ++a) {
                double mn = 1e300, mx = -1e300;
                for (int k = 0; k < 3; ++k) { mn = std::min(mn, double(m.vtx(order[i], k)[a])); mx = std::max(mx, double(m.vtx(order[i], k)[a])); }
                lo[a] = std::min(lo[a], mn); hi[a] = std::max(hi[a], mx);
                clo[a] = std::min(clo[a], 0.5 * (mn + mx)); chi[a] = std::max(chi[a], 0.5 * (mn + mx));
            }
        for (int a = 0; a < 3; ++a) {
            const double pad = 1e-5 * (hi[a] - lo[a]) + 1e-5 * std::max(std::fabs(lo[a]), std::fabs(hi[a])) + 1e-30;
            nodes[me].lo[a] = lo[a] - pad; nodes[me].hi[a] = hi[a] + pad;
        }
        if (c <= 4) { nodes[me].left = nodes[me].right = -1; nodes[me].start = s; nodes[me].count = c; return me; }
        int ax = 0;
        for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
        const uint32_t mid = s + c / 2;
        std::nth_element(order.begin() + s, order.begin() + mid, order.begin() + s + c, [&](uint32_t x, uint32_t y) {
            double cx = 0, cy = 0;
            for (int k = 0; k < 3; ++k) { cx += m.vtx(x, k)[ax]; cy += m.vtx(y, k)[ax]; }
            return cx < cy || (cx == cy && x < y);
        });
        nodes[me].count = 0;
        const int32_t l = rec(m, s, mid - s);
        const int32_t rr = rec(m, mid, s + c - mid);
        nodes[me].left = l; nodes[me].right = rr;
        return me;
    }
    static bool slab(const Node& n, const ORay& r, double tmax) {
        double t0 = r.tmin, t1 = tmax;
        for (int a = 0; a < 3; ++a) {
            const double d = r.d[a], o = r.o[a];
            if (d == 0.0) { if (o < n.lo[a] || o > n.hi[a]) return false; continue; }
            double ta = (n.lo[a] - o) / d, tb = (n.hi[a] - o) / d;
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta); t1 = std::min(t1, tb);
            if (t0 > t1) return false;
        }
        return true;
    }
    OHit closest(const OMesh& m, const ORay& r, bool any_hit) const {
        OHit h;
        h.t = r.tmax;
        if (nodes.empty()) return h;
        const WoopPrep p = woop_prep(r);
        int32_t st[128];
        int sp = 0;
        st[sp++] = 0;
        while (sp) {
            const Node& n = nodes[size_t(st[--sp])];
            if (!slab(n, r, h.t)) continue;
            if (n.left < 0) {
                for (uint32_t i = n.start; i < n.start + n.count; ++i) {
                    const uint32_t t = order[i];
                    float tt, b1, b2;
                    if (woop_intersect(r, p, m.vtx(t, 0), m.vtx(t, 1), m.vtx(t, 2), r.tmin, h.t, tt, b1, b2)) {
                        if (tt < h.t || t < h.tri) { h.tri = t; h.t = tt; h.b1 = b1; h.b2 = b2; }
                        if (any_hit) return h;
                    }
                }
            } else {
                st[sp++] = n.left;
                st[sp++] = n.right;
            }
        }
        return h;
    }
};

// Wachter-Binder ray origin offset (Utils/Geometry/GeometryHelpers.slang:47-61).
inline void compute_ray_origin(const float pos[3], const float n[3], float out[3]) {
    const float origin = 1.f / 16.f, fScale = 3.f / 65536.f, iScale = 3 * 256.f;
    for (int a = 0; a < 3; ++a) {
        const int32_t iOff = int32_t(n[a] * iScale);
        int32_t bits;
        std::memcpy(&bits, &pos[a], 4);
        bits += pos[a] < 0.f ? -iOff : iOff;
        float iPos;
        std::memcpy(&iPos, &bits, 4);
        const float fOff = n[a] * fScale;
        out[a] = std::fabs(pos[a]) < origin ? pos[a] + fOff : iPos;
    }
}

}  // namespace pltoracle
