// Software ray traversal for CDNA4: BVH4 with the four child boxes in the parent as SoA rows (one
// 128 B node = eight independent 16 B loads tests four boxes, so a level costs one memory latency),
// triangles re-ordered into leaf order as 48 B records (v0 + triangle id, v1, v2) fetched one
// triangle ahead of the test, and a per-lane traversal stack: its first kTraversalStackDepth
// entries in LDS laid out [depth][lane] (a wave's pushes/pops hit 64 distinct banks), the rest in
// private memory (scratch), touched only by rays deep in very deep trees.
//
// Hit results are independent of the tree: the watertight test (Woop, Benthin, Wald 2013,
// "Watertight Ray/Triangle Intersection", JCGT 2(1)) runs with FP contraction off and a double
// fallback on zero edge functions, box tests are conservative (far distance widened by
// 1 + 2*gamma(3), Ize 2013 "Robust BVH Ray Traversal"), and equal distances resolve to the
// smallest triangle id. The CPU oracle uses the same arithmetic, so hit ids are bit-exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plt {

struct Ray {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
};
struct Hit {
    uint32_t tri;  // 0xFFFFFFFF = miss
    float b1, b2, t;
};

struct TraceScene {
    const float4* __restrict__ nodes;  // 8 per node (BvhNode)
    const float4* __restrict__ tris;   // 3 per triangle in leaf order
};

constexpr uint32_t kNoHit = 0xFFFFFFFFu;

// Watertight ray/triangle state: axis permutation and shear constants.
struct RayPrep {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float o[3];
};

__device__ __forceinline__ float pick3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ RayPrep prep_ray(const Ray& r) {
#pragma clang fp contract(off)
    RayPrep p;
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz == 2 ? 0 : p.kz + 1;
    p.ky = p.kx == 2 ? 0 : p.kx + 1;
    const float dkz = pick3(r.dx, r.dy, r.dz, p.kz);
    if (dkz < 0.f) { const int t = p.kx; p.kx = p.ky; p.ky = t; }
    p.Sx = pick3(r.dx, r.dy, r.dz, p.kx) / dkz;
    p.Sy = pick3(r.dx, r.dy, r.dz, p.ky) / dkz;
    p.Sz = 1.f / dkz;
    p.o[0] = r.ox; p.o[1] = r.oy; p.o[2] = r.oz;
    return p;
}

// Returns true and (t, b1, b2) when the ray hits the triangle with tmin <= t <= tmax.
__device__ __forceinline__ bool intersect_tri(const RayPrep& p, float4 v0, float4 v1, float4 v2, float tmin, float tmax,
                                              float& t_out, float& b1, float& b2) {
#pragma clang fp contract(off)
    const float Ak = pick3(v0.x, v0.y, v0.z, p.kx) - pick3(p.o[0], p.o[1], p.o[2], p.kx);
    const float Al = pick3(v0.x, v0.y, v0.z, p.ky) - pick3(p.o[0], p.o[1], p.o[2], p.ky);
    const float Az0 = pick3(v0.x, v0.y, v0.z, p.kz) - pick3(p.o[0], p.o[1], p.o[2], p.kz);
    const float Bk = pick3(v1.x, v1.y, v1.z, p.kx) - pick3(p.o[0], p.o[1], p.o[2], p.kx);
    const float Bl = pick3(v1.x, v1.y, v1.z, p.ky) - pick3(p.o[0], p.o[1], p.o[2], p.ky);
    const float Bz0 = pick3(v1.x, v1.y, v1.z, p.kz) - pick3(p.o[0], p.o[1], p.o[2], p.kz);
    const float Ck = pick3(v2.x, v2.y, v2.z, p.kx) - pick3(p.o[0], p.o[1], p.o[2], p.kx);
    const float Cl = pick3(v2.x, v2.y, v2.z, p.ky) - pick3(p.o[0], p.o[1], p.o[2], p.ky);
    const float Cz0 = pick3(v2.x, v2.y, v2.z, p.kz) - pick3(p.o[0], p.o[1], p.o[2], p.kz);
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

// Conservative slab test of one child box; returns the entry distance or +inf on a miss.
__device__ __forceinline__ float box_entry(float lox, float hix, float loy, float hiy, float loz, float hiz, float ix,
                                           float iy, float iz, float oix, float oiy, float oiz, float tmin, float tmax) {
    const float tx0 = fmaf(lox, ix, -oix), tx1 = fmaf(hix, ix, -oix);
    const float ty0 = fmaf(loy, iy, -oiy), ty1 = fmaf(hiy, iy, -oiy);
    const float tz0 = fmaf(loz, iz, -oiz), tz1 = fmaf(hiz, iz, -oiz);
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tn <= tf * 1.0000005f ? tn : __builtin_inff();
}

// Entries of the traversal stack beyond the LDS part, in private memory. A BVH4 of L levels needs
// at most 3 (L - 1) entries (three siblings pushed per level).
constexpr int kTraversalSpill = 64;

// Sorts four keys ascending (five compare-exchanges).
__device__ __forceinline__ void sort4(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    uint32_t t;
    t = min(a, b); b = max(a, b); a = t;
    t = min(c, d); d = max(c, d); c = t;
    t = min(a, c); c = max(a, c); a = t;
    t = min(b, d); d = max(b, d); b = t;
    t = min(b, c); c = max(b, c); b = t;
}

// Traverses the BVH. anyHit: stop at the first hit in [tmin, tmax] (visibility rays); a per-lane
// value lets one call serve a wave whose lanes cast both kinds (whether a ray hits anything in
// [tmin, tmax] does not depend on it). `stack` points at this lane's first LDS slot; consecutive
// slots are `stride` ints apart. Children are visited nearest entry first (ties and order only
// change the speed: the closest hit is the smallest (t, triangle id) over every triangle tested,
// and boxes are conservative).
template <int kDepth>
__device__ __forceinline__ Hit traverse_rays(const TraceScene& sc, const Ray& r, int* stack, int stride, bool anyHit) {
    Hit h{kNoHit, 0.f, 0.f, r.tmax};
    const RayPrep p = prep_ray(r);
    const float dx = fabsf(r.dx) > 1e-30f ? r.dx : copysignf(1e-30f, r.dx);
    const float dy = fabsf(r.dy) > 1e-30f ? r.dy : copysignf(1e-30f, r.dy);
    const float dz = fabsf(r.dz) > 1e-30f ? r.dz : copysignf(1e-30f, r.dz);
    const float ix = 1.f / dx, iy = 1.f / dy, iz = 1.f / dz;
    const float oix = r.ox * ix, oiy = r.oy * iy, oiz = r.oz * iz;
    int spill[kTraversalSpill];
    int sp = 0;
    int node = 0;
    float best = r.tmax;
    auto push = [&](int x) {
        if (sp < kDepth) stack[sp * stride] = x;
        else if (sp < kDepth + kTraversalSpill) spill[sp - kDepth] = x;
        ++sp;
    };
    while (true) {
        if (node >= 0) {
            const float4* q = sc.nodes + 8 * node;
            const float4 lox = q[0], hix = q[1], loy = q[2], hiy = q[3], loz = q[4], hiz = q[5], cf = q[6];
            const float e0 = box_entry(lox.x, hix.x, loy.x, hiy.x, loz.x, hiz.x, ix, iy, iz, oix, oiy, oiz, r.tmin, best);
            const float e1 = box_entry(lox.y, hix.y, loy.y, hiy.y, loz.y, hiz.y, ix, iy, iz, oix, oiy, oiz, r.tmin, best);
            const float e2 = box_entry(lox.z, hix.z, loy.z, hiy.z, loz.z, hiz.z, ix, iy, iz, oix, oiy, oiz, r.tmin, best);
            const float e3 = box_entry(lox.w, hix.w, loy.w, hiy.w, loz.w, hiz.w, ix, iy, iz, oix, oiy, oiz, r.tmin, best);
            // sort keys: the entry distance's bits (non-negative floats order as integers) with the
            // slot in the low two bits; a missed box sorts last
            constexpr float kInf = __builtin_inff();
            uint32_t k0 = e0 != kInf ? (__float_as_uint(e0) & ~3u) : 0xFFFFFFFFu;
            uint32_t k1 = e1 != kInf ? ((__float_as_uint(e1) & ~3u) | 1u) : 0xFFFFFFFFu;
            uint32_t k2 = e2 != kInf ? ((__float_as_uint(e2) & ~3u) | 2u) : 0xFFFFFFFFu;
            uint32_t k3 = e3 != kInf ? ((__float_as_uint(e3) & ~3u) | 3u) : 0xFFFFFFFFu;
            sort4(k0, k1, k2, k3);
            // The child references are used only when a box is hit; left to itself the compiler
            // issues their load there, after the box tests, which costs every node visit a second
            // dependent memory round trip. Requiring them in registers here keeps the load with
            // the boxes' (the empty asm emits nothing).
            asm volatile("" : : "v"(cf.x), "v"(cf.y), "v"(cf.z), "v"(cf.w));
            if (k0 != 0xFFFFFFFFu) {
                const int c0 = __float_as_int(cf.x), c1 = __float_as_int(cf.y), c2 = __float_as_int(cf.z), c3 = __float_as_int(cf.w);
                auto child = [&](uint32_t k) {  // selects, no branches
                    const int lo = (k & 1u) ? c1 : c0, hi = (k & 1u) ? c3 : c2;
                    return (k & 2u) ? hi : lo;
                };
                // the farther hit children on the stack, farthest first; the nearest next
                if (k3 != 0xFFFFFFFFu) push(child(k3));
                if (k2 != 0xFFFFFFFFu) push(child(k2));
                if (k1 != 0xFFFFFFFFu) push(child(k1));
                node = child(k0);
                continue;
            }
        } else {
            const uint32_t leaf = uint32_t(~node);
            const uint32_t start = leaf >> 3, cnt = (leaf & 7u) + 1u;
            const float4* t = sc.tris + 3 * size_t(start);
            float4 v0 = t[0], v1 = t[1], v2 = t[2];
            for (uint32_t i = 0; i < cnt; ++i) {
                float4 w0 = v0, w1 = v1, w2 = v2;
                if (i + 1 < cnt) {  // the next triangle's fetch overlaps this test
                    w0 = t[3 * (i + 1) + 0];
                    w1 = t[3 * (i + 1) + 1];
                    w2 = t[3 * (i + 1) + 2];
                }
                float tt, b1, b2;
                if (intersect_tri(p, v0, v1, v2, r.tmin, best, tt, b1, b2)) {
                    const uint32_t id = __float_as_uint(v0.w);
                    if (tt < best || id < h.tri) {
                        best = tt;
                        h.tri = id; h.b1 = b1; h.b2 = b2; h.t = tt;
                        if (anyHit) return h;
                    }
                }
                v0 = w0; v1 = w1; v2 = w2;
            }
        }
        if (sp == 0) break;
        --sp;
        node = sp < kDepth ? stack[sp * stride] : spill[sp - kDepth];
    }
    return h;
}
template <bool kAnyHit, int kDepth>
__device__ __forceinline__ Hit traverse(const TraceScene& sc, const Ray& r, int* stack, int stride) {
    return traverse_rays<kDepth>(sc, r, stack, stride, kAnyHit);
}

}  // namespace plt
