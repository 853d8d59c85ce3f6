// rr_device.h — device-side arithmetic shared by the build and trace kernels.
//
// Every floating-point expression here is written out operation by
// operation and the library is compiled with -ffp-contract=off, so results
// are a fixed function of the inputs (IEEE-754 binary32, round to nearest),
// reproducible bit for bit by a CPU that evaluates the same sequence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rr_layout.h"

namespace rr {

#define RR_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// Order-preserving float <-> uint mapping used for atomic min/max of box
// coordinates (vlk/kernels/common.h:68-90): positive floats get the sign bit
// set, negative floats are bit-inverted, so unsigned order == float order.
RR_DEV uint32_t FloatToOrdered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
RR_DEV float OrderedToFloat(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// ---------------------------------------------------------------------------
// 30-bit Morton code (vlk/kernels/common.h:251-268): each coordinate in
// [0,1] is scaled by 1024, clamped to [0,1023], truncated, and its 10 bits
// interleaved as x2 y1 z0.
RR_DEV uint32_t SpreadBits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
RR_DEV uint32_t QuantizeUnit(float p) {
    float q = p * 1024.0f;
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}
RR_DEV uint32_t Morton30(float px, float py, float pz) {
    return (SpreadBits10(QuantizeUnit(px)) << 2) | (SpreadBits10(QuantizeUnit(py)) << 1) |
           SpreadBits10(QuantizeUnit(pz));
}
// Position of centroid c inside [lo, lo+extent]; a zero-extent axis maps to
// 0 instead of the reference's 0/0 (lbvh_calc_morton_codes_mesh.comp:119-123).
RR_DEV float UnitCoord(float c, float lo, float extent) { return extent > 0.0f ? (c - lo) / extent : 0.0f; }

// ---------------------------------------------------------------------------
// Ray set-up shared by the slab and triangle tests.
struct RayState {
    float ox, oy, oz;     // origin
    float dx, dy, dz;     // direction
    float tmin;           // lower bound of the accepted interval
    float ix, iy, iz;     // reciprocal direction for slab tests
    float sx, sy, sz;     // watertight shear constants
    uint32_t k;           // watertight axis permutation: kx | ky << 2 | kz << 4
};

// Conservative widening of slab intervals: slab distances (b - o) * (1/d)
// carry at most ~3 ulp of relative error and their signs are exact, so a box
// is kept when entry <= exit + 2^-20 |exit| (entry and exit clipped to the
// ray's interval first).
constexpr float kSlabWiden = 9.5367431640625e-07f;  // 2^-20
// Relative slack on the current closest distance when culling boxes, so a
// candidate at exactly the closest distance (tie) or within the triangle
// test's rounding of it is never culled.
constexpr float kLimitWiden = 1.52587890625e-05f;  // 2^-16

RR_DEV float Perm(int k, float x, float y, float z) { return k == 0 ? x : (k == 1 ? y : z); }

RR_DEV void SetupRay(RayState& r, float ox, float oy, float oz, float tmin, float dx, float dy, float dz) {
    r.ox = ox; r.oy = oy; r.oz = oz; r.tmin = tmin;
    r.dx = dx; r.dy = dy; r.dz = dz;
    // Exact reciprocals: a zero component gives +-inf (sign of the zero), so
    // an axis-parallel ray has an unbounded slab on that axis.
    // One division at a time (a scheduling barrier between them): ray set-up
    // runs while a lane's whole traversal state is live, and three
    // interleaved divisions would set the kernel's register peak.
    r.ix = 1.0f / dx;
    __builtin_amdgcn_sched_barrier(0);
    r.iy = 1.0f / dy;
    __builtin_amdgcn_sched_barrier(0);
    r.iz = 1.0f / dz;
    // Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013:
    // kz = dimension of largest |d| (first on ties), kx/ky the next two in
    // cyclic order, swapped when d[kz] < 0 to keep the winding.
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const float dkz = Perm(kz, dx, dy, dz);
    // (kx, ky) = (1, 2), (2, 0), (0, 1) for kz = 0, 1, 2, swapped when d[kz] < 0.
    constexpr uint32_t kFwd0 = 1u | 2u << 2, kFwd1 = 2u | 0u << 2 | 1u << 4, kFwd2 = 0u | 1u << 2 | 2u << 4;
    constexpr uint32_t kRev0 = 2u | 1u << 2, kRev1 = 0u | 2u << 2 | 1u << 4, kRev2 = 1u | 0u << 2 | 2u << 4;
    r.k = dkz < 0.0f ? (kz == 0 ? kRev0 : (kz == 1 ? kRev1 : kRev2)) : (kz == 0 ? kFwd0 : (kz == 1 ? kFwd1 : kFwd2));
    const int kx = (int)(r.k & 3u), ky = (int)((r.k >> 2) & 3u);
    // Shear constants from the reciprocal already computed for the slab test.
    r.sz = Perm(kz, r.ix, r.iy, r.iz);
    r.sx = Perm(kx, dx, dy, dz) * r.sz;
    r.sy = Perm(ky, dx, dy, dz) * r.sz;
}

// Conservative slab test of one box.  Near and far planes are chosen by the
// sign of the reciprocal direction.  An origin lying exactly on a plane of
// an axis the ray is parallel to gives 0 * inf = NaN for that plane; the
// IEEE max/min below ignore NaN operands, which makes that plane unbounded,
// as it is for the exact ray.  Returns the entry distance in t0 and whether
// the box is kept; the exit is clipped to `limit`.
RR_DEV bool SlabTest(const RayState& r, float limit, float bx0, float by0, float bz0, float bx1, float by1, float bz1,
                     float& t0) {
    const float lx = (bx0 - r.ox) * r.ix, hx = (bx1 - r.ox) * r.ix;
    const float ly = (by0 - r.oy) * r.iy, hy = (by1 - r.oy) * r.iy;
    const float lz = (bz0 - r.oz) * r.iz, hz = (bz1 - r.oz) * r.iz;
    const bool sx = r.ix < 0.0f, sy = r.iy < 0.0f, sz = r.iz < 0.0f;
    const float nx = sx ? hx : lx, fx = sx ? lx : hx;
    const float ny = sy ? hy : ly, fy = sy ? ly : hy;
    const float nz = sz ? hz : lz, fz = sz ? lz : hz;
    const float a = fmaxf(fmaxf(fmaxf(nx, ny), nz), r.tmin);
    const float b = fminf(fminf(fminf(fx, fy), fz), limit);
    t0 = a;
    return a <= fmaf(kSlabWiden, fabsf(b), b);
}

RR_DEV float WidenLimit(float t) { return fmaf(kLimitWiden, fabsf(t), t); }

// Leaf test of an AABB-list primitive: the ray's entry distance into the
// box (exact slab test, no widening), t = max(entry, tmin), a hit when
// t <= min(exit, tmax).  Same plane selection and NaN handling as SlabTest.
RR_DEV bool BoxLeafTest(const RayState& r, float tmax, float bx0, float by0, float bz0, float bx1, float by1, float bz1,
                        float& t) {
    const float lx = (bx0 - r.ox) * r.ix, hx = (bx1 - r.ox) * r.ix;
    const float ly = (by0 - r.oy) * r.iy, hy = (by1 - r.oy) * r.iy;
    const float lz = (bz0 - r.oz) * r.iz, hz = (bz1 - r.oz) * r.iz;
    const bool sx = r.ix < 0.0f, sy = r.iy < 0.0f, sz = r.iz < 0.0f;
    const float nx = sx ? hx : lx, fx = sx ? lx : hx;
    const float ny = sy ? hy : ly, fy = sy ? ly : hy;
    const float nz = sz ? hz : lz, fz = sz ? lz : hz;
    const float a = fmaxf(fmaxf(fmaxf(nx, ny), nz), r.tmin);
    const float b = fminf(fminf(fminf(fx, fy), fz), tmax);
    t = a;
    return a <= b;
}


// Watertight ray/triangle test (Woop et al. 2013, with the paper's double
// precision fallback for zero edge functions, no backface culling).  On a
// hit returns true with the barycentric numerators and denominator: the hit
// is u = V/det (weight of v1), v = W/det (weight of v2), t = T/det.
struct TriHit {
    float V, W, det, t;
};

RR_DEV bool WatertightTest(const RayState& r, float v0x, float v0y, float v0z, float v1x, float v1y, float v1z,
                           float v2x, float v2y, float v2z, TriHit& h) {
    const float Ax0 = v0x - r.ox, Ay0 = v0y - r.oy, Az0 = v0z - r.oz;
    const float Ax1 = v1x - r.ox, Ay1 = v1y - r.oy, Az1 = v1z - r.oz;
    const float Ax2 = v2x - r.ox, Ay2 = v2y - r.oy, Az2 = v2z - r.oz;
    const int kx = (int)(r.k & 3u), ky = (int)((r.k >> 2) & 3u), kz = (int)(r.k >> 4);
    const float akx = Perm(kx, Ax0, Ay0, Az0), aky = Perm(ky, Ax0, Ay0, Az0), akz = Perm(kz, Ax0, Ay0, Az0);
    const float bkx = Perm(kx, Ax1, Ay1, Az1), bky = Perm(ky, Ax1, Ay1, Az1), bkz = Perm(kz, Ax1, Ay1, Az1);
    const float ckx = Perm(kx, Ax2, Ay2, Az2), cky = Perm(ky, Ax2, Ay2, Az2), ckz = Perm(kz, Ax2, Ay2, Az2);
    const float Ax = akx - r.sx * akz, Ay = aky - r.sy * akz;
    const float Bx = bkx - r.sx * bkz, By = bky - r.sy * bkz;
    const float Cx = ckx - r.sx * ckz, Cy = cky - r.sy * ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.sz * akz, Bz = r.sz * bkz, Cz = r.sz * ckz;
    const float T = U * Az + V * Bz + W * Cz;
    h.V = V; h.W = W; h.det = det;
    h.t = T / det;
    return true;
}

// Leaf test by primitive kind: a triangle (watertight) or, for an AABB
// list, the leaf's box (stored as its min corner in v0 and max corner in
// v1), reported with barycentrics 0.
RR_DEV bool LeafTest(bool boxes, const RayState& r, float tmax, float v0x, float v0y, float v0z, float v1x, float v1y,
                     float v1z, float v2x, float v2y, float v2z, TriHit& h) {
    if (boxes) {
        h.V = 0.0f;
        h.W = 0.0f;
        h.det = 1.0f;
        return BoxLeafTest(r, tmax, v0x, v0y, v0z, v1x, v1y, v1z, h.t);
    }
    return WatertightTest(r, v0x, v0y, v0z, v1x, v1y, v1z, v2x, v2y, v2z, h);
}

// Closest-hit ordering: (t, inst, prim) lexicographic, so equal distances
// resolve to the lowest ids whatever order the BVH visits them in.
RR_DEV bool CloserHit(float t, uint32_t inst, uint32_t prim, float bt, uint32_t binst, uint32_t bprim) {
    return t < bt || (t == bt && (inst < binst || (inst == binst && prim < bprim)));
}

// ---------------------------------------------------------------------------
// 64-bit wave helpers (wave64 on CDNA).
// The gfx950 dispatcher hands workgroups to its 8 XCDs round-robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  Kernels whose
// neighbouring workgroups share cache lines (the refits: a mesh's triangles
// and nodes) renumber their workgroups with this bijection of [0, grid) so
// that each XCD takes one contiguous range of logical workgroups (XCD x the
// x-th eighth, the first grid % 8 ranges one longer).  Only placement
// depends on it, never results.
constexpr uint32_t kXcds = 8;
RR_DEV uint32_t XcdContiguousBlock(uint32_t block, uint32_t grid) {
    const uint32_t x = block % kXcds, k = block / kXcds;
    const uint32_t q = grid / kXcds, r = grid % kXcds;
    return x * q + (x < r ? x : r) + k;
}

RR_DEV uint64_t LaneMaskLt() {
    const uint32_t lane = __lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Number of set bits of m below this lane (v_mbcnt: no lane mask held in registers).
RR_DEV uint32_t RankIn(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
RR_DEV T WaveMin(T v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
template <typename T>
RR_DEV T WaveMax(T v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Loads and stores through the global address space.  A pointer read from
// memory (a BLAS address in a TLAS's table, a batched request's buffers) is
// generic to the compiler, which then emits flat instructions; those count
// against both the vector-memory and the LDS counters, so each wait for one
// also waits for the other (the LDS stack's accesses among them).  Every
// buffer here is device memory, so the cast to address space 1 is exact.
typedef float RrFloat4V __attribute__((ext_vector_type(4)));
template <typename T>
using GlobalPtr = const T __attribute__((address_space(1)))*;
template <typename T>
using GlobalPtrW = T __attribute__((address_space(1)))*;
template <typename T>
RR_DEV T LdG(const T* p) { return *(GlobalPtr<T>)p; }
RR_DEV float4 LdG4(const void* p) {
    const RrFloat4V v = *(GlobalPtr<RrFloat4V>)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
template <typename T>
RR_DEV void StG(T* p, T v) { *(GlobalPtrW<T>)p = v; }
RR_DEV void StG4(void* p, float4 v) {
    RrFloat4V w;
    w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
    *(GlobalPtrW<RrFloat4V>)p = w;
}

// A device address read from memory (a table of meshes), re-expressed as
// the same address reached from a kernel-argument pointer: the compiler then
// knows it points to global memory and emits global rather than flat
// instructions for every access through it, also in helpers that take plain
// pointers.  Null stays null.
template <typename T>
RR_DEV T* RebaseOn(const void* kernel_arg, uint64_t addr) {
    if (addr == 0) return nullptr;
    char* b = const_cast<char*>(static_cast<const char*>(kernel_arg));
    return reinterpret_cast<T*>(b + (int64_t)(addr - reinterpret_cast<uint64_t>(b)));
}

// Node loads as four 16-byte vector loads.
RR_DEV void LoadNode(const BvhNode* __restrict__ nodes, uint32_t i, float4& q0, float4& q1, float4& q2, float4& q3) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    q0 = LdG4(p); q1 = LdG4(p + 1); q2 = LdG4(p + 2); q3 = LdG4(p + 3);
}

}  // namespace rr
