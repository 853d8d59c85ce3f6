// oracle.cpp — CPU oracle for the build-and-intersect path.
//
// TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this as the checker; the library never links it.
//
// What it restates (reference files cited per function):
//   * the CPU traversal and validators of bvh_analyzer (bvh.h:129-328,
//     aabb.h:72-81, transform.h:56-125), with its two bugs fixed: the
//     reciprocal uses float fabs/copysign (float3.h:157-169 calls int abs), and
//     a triangle is accepted against the running closest distance
//     (triangle.h:59 bounds by ray.max_t); children are visited near-first as
//     the reference GPU kernel does (isect.comp:134-166);
//   * the LBVH build of the Vulkan kernels (lbvh_*.comp, restructure_bvh.comp).
// Numerics follow the library's contract: binary32 operations in a fixed
// order (built with -ffp-contract=off), a watertight triangle test (Woop,
// Benthin, Wald, JCGT 2(1) 2013) and (t, inst, prim) lexicographic ordering of
// closest hits.  The reference holds no golden vectors for this path; every
// function here is pinned instead by closed forms, brute force, invariants,
// FP64 and the reference's own bvh_analyzer compiled from its sources (see
// tests/test_oracle.py).
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <queue>
#include <set>
#include <stack>
#include <utility>
#include <vector>

namespace {

constexpr uint32_t kInv = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFEu;

struct Node {
    float a0min[3];
    uint32_t c0;
    float a0max[3];
    uint32_t c1;
    float a1min[3];
    uint32_t parent;
    float a1max[3];
    uint32_t update;
};
static_assert(sizeof(Node) == 64, "node size");

struct Header {
    uint32_t magic, kind, prim_count, node_count;
    uint64_t nodes_offset, transforms_offset, blas_table_offset;
    uint32_t flags, build_serial;
    uint32_t reserved[20];
};
static_assert(sizeof(Header) == 128, "header size");

constexpr uint32_t kMagic = 0x53415252u;
constexpr uint32_t kBlas = 1, kTlas = 2;
constexpr uint32_t kFlagAabbs = 2;  // leaves are boxes (AABB list)

size_t Align(size_t v, size_t a) { return (v + a - 1) / a * a; }
uint32_t NodeCount(uint32_t n) { return n ? 2 * n - 1 : 0; }

const Header* Hdr(const void* a) { return static_cast<const Header*>(a); }
const Node* NodesOf(const void* a) {
    return reinterpret_cast<const Node*>(static_cast<const uint8_t*>(a) + Hdr(a)->nodes_offset);
}
Node* NodesOf(void* a) { return reinterpret_cast<Node*>(static_cast<uint8_t*>(a) + Hdr(a)->nodes_offset); }

// IEEE minNum/maxNum (the non-NaN operand wins), as the GPU's min/max.
inline float Min(float a, float b) { return (a < b || b != b) ? a : b; }
inline float Max(float a, float b) { return (a > b || b != b) ? a : b; }

// ---------------------------------------------------------------------------
// Build restatement.

uint32_t ToOrdered(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
float FromOrdered(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Bit dilation and quantisation of vlk/kernels/common.h:251-268.
uint32_t Spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
uint32_t Quantize(float p) {
    float q = p * 1024.0f;
    q = Min(Max(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}
float Unit(float c, float lo, float ext) { return ext > 0.0f ? (c - lo) / ext : 0.0f; }

void Tri(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t t, float v[9]) {
    for (int k = 0; k < 3; ++k) {
        const float* p = vtx + (size_t)idx[3 * (size_t)t + k] * stride;
        v[3 * k] = p[0];
        v[3 * k + 1] = p[1];
        v[3 * k + 2] = p[2];
    }
}

// Min and max of ordered uints are exact, so the thread count and the order
// of the reduction do not change the result.
void Bounds(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t tris, float lo[3], float hi[3]) {
    uint32_t mn[3] = {kInv, kInv, kInv}, mx[3] = {0, 0, 0};
#pragma omp parallel
    {
        uint32_t tmn[3] = {kInv, kInv, kInv}, tmx[3] = {0, 0, 0};
#pragma omp for schedule(static) nowait
        for (int64_t t = 0; t < (int64_t)tris; ++t) {
            float v[9];
            Tri(vtx, stride, idx, (uint32_t)t, v);
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    tmn[a] = std::min(tmn[a], ToOrdered(v[3 * k + a]));
                    tmx[a] = std::max(tmx[a], ToOrdered(v[3 * k + a]));
                }
        }
#pragma omp critical
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], tmn[a]);
            mx[a] = std::max(mx[a], tmx[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = FromOrdered(mn[a]);
        hi[a] = FromOrdered(mx[a]);
    }
}

// delta of lbvh_emit_hierarchy_mesh.comp:85-103.
int Delta(const uint32_t* c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    if (c[i] != c[j]) return __builtin_clz(c[i] ^ c[j]);
    return 32 + __builtin_clz((uint32_t)(i ^ j));
}

// Every node is written by one iteration only (its own, and the parent
// link by its one parent's), so the loops run in parallel.
void Emit(const uint32_t* c, const uint32_t* prims, uint32_t n, Node* nodes) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Node& l = nodes[n - 1 + i];
        l.c0 = kInv;
        l.c1 = prims[i];
        l.update = 0;
    }
    if (n == 1) {
        nodes[0].parent = kInv;
        return;
    }
    const int N = (int)n;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < N - 1; ++I) {
        // FindSpan (lbvh_emit_hierarchy_mesh.comp:105-140).
        const int d = (Delta(c, N, I, I + 1) - Delta(c, N, I, I - 1)) > 0 ? 1 : -1;
        const int dmin = Delta(c, N, I, I - d);
        int lmax = 2;
        while (Delta(c, N, I, I + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (Delta(c, N, I, I + (l + t) * d) > dmin) l += t;
        const int j = I + l * d;
        const int first = std::min(I, j), last = std::max(I, j);
        // FindSplit (lbvh_emit_hierarchy_mesh.comp:142-168).
        const int common = Delta(c, N, first, last);
        int lo = first, hi = last;
        while (hi > lo + 1) {
            const int mid = (lo + hi) / 2;
            if (Delta(c, N, first, mid) > common) lo = mid;
            else hi = mid;
        }
        const uint32_t left = lo == first ? n - 1 + (uint32_t)lo : (uint32_t)lo;
        const uint32_t right = lo + 1 == last ? n - 1 + (uint32_t)lo + 1 : (uint32_t)lo + 1;
        nodes[I].c0 = left;
        nodes[I].c1 = right;
        nodes[I].update = 0;
        nodes[left].parent = (uint32_t)I;
        nodes[right].parent = (uint32_t)I;
    }
    nodes[0].parent = kInv;
}

struct Box {
    float lo[3], hi[3];
};

// Refit (lbvh_fit_aabb_mesh.comp:94-205): leaves get their triangle, internal
// nodes their children's boxes; post-order with an explicit stack.
void Refit(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t n, Node* nodes) {
    if (n == 0) return;
    const uint32_t nint = n - 1;
    std::vector<Box> box(NodeCount(n));
    std::vector<std::pair<uint32_t, bool>> st;
    st.push_back({0, false});
    while (!st.empty()) {
        auto [x, expanded] = st.back();
        st.pop_back();
        if (x >= nint) {
            Node& l = nodes[x];
            float v[9];
            Tri(vtx, stride, idx, l.c1, v);
            for (int a = 0; a < 3; ++a) {
                l.a0min[a] = v[a];
                l.a0max[a] = v[3 + a];
                l.a1min[a] = v[6 + a];
                l.a1max[a] = 0.0f;
                box[x].lo[a] = Min(Min(v[a], v[3 + a]), v[6 + a]);
                box[x].hi[a] = Max(Max(v[a], v[3 + a]), v[6 + a]);
            }
            l.update = 0;
            continue;
        }
        if (!expanded) {
            st.push_back({x, true});
            st.push_back({nodes[x].c0, false});
            st.push_back({nodes[x].c1, false});
            continue;
        }
        Node& nd = nodes[x];
        const Box& b0 = box[nd.c0];
        const Box& b1 = box[nd.c1];
        for (int a = 0; a < 3; ++a) {
            nd.a0min[a] = b0.lo[a];
            nd.a0max[a] = b0.hi[a];
            nd.a1min[a] = b1.lo[a];
            nd.a1max[a] = b1.hi[a];
            box[x].lo[a] = Min(b0.lo[a], b1.lo[a]);
            box[x].hi[a] = Max(b0.hi[a], b1.hi[a]);
        }
        nd.update = 0;  // the GPU refit's second arrival clears the arrival counter
    }
}

// --- treelet restructuring (restructure_bvh.comp:149-542) -----------------

constexpr int kLeaves = 7;  // the reference's treelet size; 5 and 9 are tuning options
constexpr int kMasks = 1 << kLeaves;
constexpr float kCInt = 1.2f;

float Area(const float lo[3], const float hi[3]) {
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float s = ex * ez;
    s = s + ey * ex;
    s = s + ez * ey;
    return 2.0f * s;
}

Box NodeBox(const Node* nodes, uint32_t x, uint32_t nint) {
    const Node& nd = nodes[x];
    Box b;
    for (int a = 0; a < 3; ++a) {
        float l = Min(Min(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
        float h = Max(Max(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
        if (x < nint) {
            l = Min(l, nd.a1max[a]);
            h = Max(h, nd.a1max[a]);
        }
        b.lo[a] = l;
        b.hi[a] = h;
    }
    return b;
}

// Optimal partition of every subset of the L treelet leaves by increasing
// size (restructure_bvh.comp:224-414, L = 7 there): candidates of subset s
// are the non-empty subsets p of s without its lowest leaf, in increasing
// order, the first cheapest wins; cost = C_INT * area + best split.
// Orientation as restructure_bvh.comp:395-397 (sizes 2-5 flipped for even
// sizes; larger sizes unflipped).
template <int L>
void TreeletDp(const Box lb[L], float sa[1 << L], float cost[1 << L], uint32_t part[1 << L]) {
    constexpr int kLeaves = L;
    constexpr int kMasks = 1 << L;
    auto union_of = [&](uint32_t m) {
        Box b;
        for (int a = 0; a < 3; ++a) { b.lo[a] = FLT_MAX; b.hi[a] = -FLT_MAX; }
        for (int k = 0; k < kLeaves; ++k)
            if (m & (1u << k))
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = Min(b.lo[a], lb[k].lo[a]);
                    b.hi[a] = Max(b.hi[a], lb[k].hi[a]);
                }
        return b;
    };
    for (uint32_t m = 0; m < kMasks; ++m) {
        Box b = union_of(m);
        sa[m] = m ? Area(b.lo, b.hi) : 0.0f;
        cost[m] = __builtin_popcount(m) == 1 ? kCInt * sa[m] : 0.0f;
        part[m] = 0;
    }
    for (int bits = 2; bits <= kLeaves; ++bits) {
        for (uint32_t s = 0; s < kMasks; ++s) {
            if (__builtin_popcount(s) != bits) continue;
            const uint32_t delta = (s - 1) & s;
            uint32_t p = (0u - delta) & s;
            float best_cost = FLT_MAX;
            uint32_t best = 0;
            do {
                const float c = cost[p] + cost[s ^ p];
                if (best == 0 || c < best_cost) {
                    best_cost = c;
                    best = p;
                }
                p = (p - delta) & s;
            } while (p != 0);
            cost[s] = kCInt * sa[s] + best_cost;
            part[s] = (bits <= 5 && (bits & 1) == 0) ? (s ^ best) : best;
        }
    }
}

template <int L>
void RestructureTreelet(Node* nodes, uint32_t nint, uint32_t root) {
    constexpr int kLeaves = L;
    constexpr int kMasks = 1 << L;
    uint32_t leaf[kLeaves], inner[kLeaves - 1];
    float area[kLeaves];
    inner[0] = root;
    leaf[0] = nodes[root].c0;
    leaf[1] = nodes[root].c1;
    for (int k = 0; k < 2; ++k) {
        Box b = NodeBox(nodes, leaf[k], nint);
        area[k] = Area(b.lo, b.hi);
    }
    for (int size = 2; size < kLeaves; ++size) {
        float largest = 0.0f;
        int sel = -1;
        for (int k = 0; k < size; ++k)
            if (leaf[k] < nint && (largest == 0.0f || area[k] > largest)) {
                largest = area[k];
                sel = k;
            }
        if (sel < 0) return;  // fewer than L leaves below: leave as is
        const uint32_t opened = leaf[sel];
        inner[size - 1] = opened;
        const uint32_t c0 = nodes[opened].c0, c1 = nodes[opened].c1;
        leaf[sel] = c0;
        leaf[size] = c1;
        Box b = NodeBox(nodes, c0, nint);
        area[sel] = Area(b.lo, b.hi);
        b = NodeBox(nodes, c1, nint);
        area[size] = Area(b.lo, b.hi);
    }
    Box lb[kLeaves];
    for (int k = 0; k < kLeaves; ++k) lb[k] = NodeBox(nodes, leaf[k], nint);
    float sa[kMasks], cost[kMasks];
    uint32_t part[kMasks];
    auto union_of = [&](uint32_t m) {
        Box b;
        for (int a = 0; a < 3; ++a) { b.lo[a] = FLT_MAX; b.hi[a] = -FLT_MAX; }
        for (int k = 0; k < kLeaves; ++k)
            if (m & (1u << k))
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = Min(b.lo[a], lb[k].lo[a]);
                    b.hi[a] = Max(b.hi[a], lb[k].hi[a]);
                }
        return b;
    };
    TreeletDp<L>(lb, sa, cost, part);
    // Rewire top-down.
    uint32_t stack_mask[kLeaves], stack_node[kLeaves];
    int sp = 0, allocated = 1;
    stack_mask[sp] = kMasks - 1;
    stack_node[sp] = inner[0];
    ++sp;
    while (sp > 0) {
        --sp;
        const uint32_t m = stack_mask[sp], node = stack_node[sp];
        const uint32_t side[2] = {part[m], m ^ part[m]};
        uint32_t child[2];
        for (int c = 0; c < 2; ++c) {
            if (__builtin_popcount(side[c]) > 1) {
                child[c] = inner[allocated++];
                stack_mask[sp] = side[c];
                stack_node[sp] = child[c];
                ++sp;
            } else {
                child[c] = leaf[__builtin_ctz(side[c])];
            }
        }
        const Box b0 = union_of(side[0]), b1 = union_of(side[1]);
        Node& nd = nodes[node];
        for (int a = 0; a < 3; ++a) {
            nd.a0min[a] = b0.lo[a];
            nd.a0max[a] = b0.hi[a];
            nd.a1min[a] = b1.lo[a];
            nd.a1max[a] = b1.hi[a];
        }
        nd.c0 = child[0];
        nd.c1 = child[1];
        nodes[child[0]].parent = node;
        nodes[child[1]].parent = node;
    }
}

void RefitFromChildren(Node* nodes, uint32_t x, uint32_t nint) {
    const Box b0 = NodeBox(nodes, nodes[x].c0, nint);
    const Box b1 = NodeBox(nodes, nodes[x].c1, nint);
    for (int a = 0; a < 3; ++a) {
        nodes[x].a0min[a] = b0.lo[a];
        nodes[x].a0max[a] = b0.hi[a];
        nodes[x].a1min[a] = b1.lo[a];
        nodes[x].a1max[a] = b1.hi[a];
    }
}

// One round: treelet roots are the nodes a find_treelet_roots.comp climb
// stops at (both children complete, >= min_prims primitives); every root
// and every ancestor of a root is processed after its subtree, ancestors
// being refit from their children first (restructure_bvh.comp:497-540).
template <int L>
void RestructureRound(Node* nodes, uint32_t n, uint32_t min_prims) {
    const uint32_t nint = n - 1, total = NodeCount(n);
    std::vector<uint32_t> count(total, 0);
    std::vector<uint8_t> complete(total, 0), root(total, 0), marked(total, 0);
    // Post-order.
    std::vector<uint32_t> order;
    order.reserve(total);
    {
        std::vector<std::pair<uint32_t, bool>> st{{0, false}};
        while (!st.empty()) {
            auto [x, e] = st.back();
            st.pop_back();
            if (x >= nint || e) {
                order.push_back(x);
                continue;
            }
            st.push_back({x, true});
            st.push_back({nodes[x].c1, false});
            st.push_back({nodes[x].c0, false});
        }
    }
    for (uint32_t x : order) {
        if (x >= nint) {
            count[x] = 1;
            complete[x] = 1;
            continue;
        }
        const uint32_t a = nodes[x].c0, b = nodes[x].c1;
        count[x] = count[a] + count[b];
        const bool both = complete[a] && complete[b];
        if (both && count[x] >= min_prims) root[x] = 1;
        complete[x] = both && count[x] < min_prims;
        marked[x] = root[x] || marked[a] || marked[b];
    }
    for (uint32_t x : order) {
        if (x >= nint || !marked[x]) continue;
        if (!root[x]) RefitFromChildren(nodes, x, nint);
        RestructureTreelet<L>(nodes, nint, x);
    }
}

// ---------------------------------------------------------------------------
// Traversal restatement.

struct RayState {
    float o[3], d[3], tmin, id[3], S[3];
    int k[3];
};

constexpr float kSlabWiden = 9.5367431640625e-07f;  // 2^-20
constexpr float kLimitWiden = 1.52587890625e-05f;  // 2^-16

void Setup(RayState& r, const float o[3], float tmin, const float d[3]) {
    for (int a = 0; a < 3; ++a) {
        r.o[a] = o[a];
        r.d[a] = d[a];
        r.id[a] = 1.0f / d[a];  // +-inf for a zero component (the reference clamps |d| to 1e-5, common.h:172-183)
    }
    r.tmin = tmin;
    const float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.k[0] = kx;
    r.k[1] = ky;
    r.k[2] = kz;
    r.S[2] = r.id[kz];  // 1/d[kz], shared with the slab test
    r.S[0] = d[kx] * r.S[2];
    r.S[1] = d[ky] * r.S[2];
}

// Slab test with near/far planes chosen by the direction's sign; NaN plane
// distances (origin on the plane of a parallel axis) drop out of the IEEE
// max/min.  The box is kept when entry <= exit + 2^-20 |exit|.
bool Slab(const RayState& r, float limit, const float lo[3], const float hi[3], float& t0) {
    float n3[3], f3[3];
    for (int a = 0; a < 3; ++a) {
        const float l = (lo[a] - r.o[a]) * r.id[a];
        const float h = (hi[a] - r.o[a]) * r.id[a];
        const bool neg = r.id[a] < 0.0f;
        n3[a] = neg ? h : l;
        f3[a] = neg ? l : h;
    }
    const float a = Max(Max(Max(n3[0], n3[1]), n3[2]), r.tmin);
    const float b = Min(Min(Min(f3[0], f3[1]), f3[2]), limit);
    t0 = a;
    return a <= std::fma(kSlabWiden, std::fabs(b), b);
}

float Widen(float t) { return std::fma(kLimitWiden, std::fabs(t), t); }

// Leaf of an AABB-list BLAS: entry distance into the box (exact slab, no
// widening), t = max(entry, tmin), a hit when t <= min(exit, tmax); the
// leaf stores the box's min corner in a0min and its max corner in a0max.
bool BoxLeaf(const RayState& r, float tmax, const float lo[3], const float hi[3], float& t) {
    float n3[3], f3[3];
    for (int a = 0; a < 3; ++a) {
        const float l = (lo[a] - r.o[a]) * r.id[a];
        const float h = (hi[a] - r.o[a]) * r.id[a];
        const bool neg = r.id[a] < 0.0f;
        n3[a] = neg ? h : l;
        f3[a] = neg ? l : h;
    }
    const float a = Max(Max(Max(n3[0], n3[1]), n3[2]), r.tmin);
    const float b = Min(Min(Min(f3[0], f3[1]), f3[2]), tmax);
    t = a;
    return a <= b;
}

struct Tri3 {
    float V, W, det, t;
};

// Woop et al. 2013 with the double-precision fallback for zero edge values.
bool Watertight(const RayState& r, const float v0[3], const float v1[3], const float v2[3], Tri3& h) {
    float A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = v0[a] - r.o[a];
        B[a] = v1[a] - r.o[a];
        C[a] = v2[a] - r.o[a];
    }
    const int kx = r.k[0], ky = r.k[1], kz = r.k[2];
    const float Ax = A[kx] - r.S[0] * A[kz], Ay = A[ky] - r.S[1] * A[kz];
    const float Bx = B[kx] - r.S[0] * B[kz], By = B[ky] - r.S[1] * B[kz];
    const float Cx = C[kx] - r.S[0] * C[kz], Cy = C[ky] - r.S[1] * C[kz];
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
    const float Az = r.S[2] * A[kz], Bz = r.S[2] * B[kz], Cz = r.S[2] * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    h.V = V;
    h.W = W;
    h.det = det;
    h.t = T / det;
    return true;
}

// Leaf test by the BLAS's primitive kind.
bool Leaf(bool boxes, const RayState& r, float tmax, const Node& l, Tri3& h) {
    if (boxes) {
        h.V = 0.0f;
        h.W = 0.0f;
        h.det = 1.0f;
        return BoxLeaf(r, tmax, l.a0min, l.a0max, h.t);
    }
    return Watertight(r, l.a0min, l.a0max, l.a1min, h);
}

bool Closer(float t, uint32_t inst, uint32_t prim, float bt, uint32_t binst, uint32_t bprim) {
    return t < bt || (t == bt && (inst < binst || (inst == binst && prim < bprim)));
}

struct Xf {
    float m[3][4];
};

void TransformRay(const Xf& x, const float o[3], const float d[3], float oo[3], float dd[3]) {
    for (int a = 0; a < 3; ++a) {
        oo[a] = ((x.m[a][0] * o[0] + x.m[a][1] * o[1]) + x.m[a][2] * o[2]) + x.m[a][3];
        dd[a] = (x.m[a][0] * d[0] + x.m[a][1] * d[1]) + x.m[a][2] * d[2];
    }
}

const Xf* XformsOf(const void* tlas) {
    return reinterpret_cast<const Xf*>(static_cast<const uint8_t*>(tlas) + Hdr(tlas)->transforms_offset);
}

OrcHit MissHit(float tmax) {
    OrcHit h{};
    h.inst_id = kInv;
    h.prim_id = kInv;
    h.t = tmax;
    return h;
}

// Depth-first traversal (bvh_analyzer/bvh.h:254-316 with isect.comp's
// near-first child order and closest-distance culling).
OrcHit TraceOne(const void* accel, const void* const* blas_host, const float* ray, bool any, OrcStats* st) {
    const Header* h = Hdr(accel);
    const bool two = h->kind == kTlas;
    const float tmax = ray[7];
    OrcHit out = MissHit(tmax);
    OrcStats s{0, 0, 0, 0};
    if (h->node_count == 0) {
        if (st) *st = s;
        return out;
    }
    const Node* top = NodesOf(accel);
    const Xf* xf = two ? XformsOf(accel) : nullptr;
    RayState r;
    Setup(r, ray, ray[3], ray + 4);
    float best_t = tmax, bV = 0, bW = 0, bdet = 1;
    uint32_t best_prim = kInv, best_inst = kInv;
    float limit = Widen(tmax);
    std::vector<uint32_t> stack;
    stack.reserve(128);
    const Node* nodes = top;
    uint32_t inst = kInv, node = 0;
    bool boxes = !two && (h->flags & kFlagAabbs) != 0;
    bool alive = true;
    while (alive) {
        const Node& nd = nodes[node];
        if (nd.c0 != kInv) {
            s.internal_nodes++;
            s.box_tests += 2;
            float ta, tb;
            const bool h0 = Slab(r, limit, nd.a0min, nd.a0max, ta);
            const bool h1 = Slab(r, limit, nd.a1min, nd.a1max, tb);
            if (h0 || h1) {
                const bool c1first = h1 && (!h0 || ta > tb);
                if (h0 && h1) stack.push_back(c1first ? nd.c0 : nd.c1);
                node = c1first ? nd.c1 : nd.c0;
                continue;
            }
        } else if (two && inst == kInv) {
            s.leaves++;
            const uint32_t id = nd.c1;
            const void* blas = blas_host[id];
            if (Hdr(blas)->node_count != 0) {
                float oo[3], dd[3];
                TransformRay(xf[2 * (size_t)id], ray, ray + 4, oo, dd);
                Setup(r, oo, ray[3], dd);
                inst = id;
                nodes = NodesOf(blas);
                boxes = (Hdr(blas)->flags & kFlagAabbs) != 0;
                stack.push_back(kSentinel);
                node = 0;
                continue;
            }
        } else {
            s.leaves++;
            s.tri_tests++;
            Tri3 th;
            if (Leaf(boxes, r, best_t, nd, th) && th.t >= r.tmin && th.t <= best_t) {
                const uint32_t hi = two ? inst : 0u;
                if (any) {
                    best_t = th.t; best_prim = nd.c1; best_inst = hi;
                    bV = th.V; bW = th.W; bdet = th.det;
                    break;
                }
                if (Closer(th.t, hi, nd.c1, best_t, best_inst, best_prim)) {
                    best_t = th.t; best_prim = nd.c1; best_inst = hi;
                    bV = th.V; bW = th.W; bdet = th.det;
                    limit = Widen(best_t);
                }
            }
        }
        for (;;) {
            if (stack.empty()) { alive = false; break; }
            node = stack.back();
            stack.pop_back();
            if (two && node == kSentinel) {
                inst = kInv;
                nodes = top;
                boxes = false;
                Setup(r, ray, ray[3], ray + 4);
                continue;
            }
            break;
        }
    }
    if (best_prim != kInv) {
        out.u = bV / bdet;
        out.v = bW / bdet;
        out.inst_id = best_inst;
        out.prim_id = best_prim;
        out.t = best_t;
    }
    if (st) *st = s;
    return out;
}

// Closest hit over every triangle of one BLAS for an already set-up ray.
void BruteBlas(const void* blas, const RayState& r, uint32_t inst, float& best_t, uint32_t& best_inst,
               uint32_t& best_prim, float& bV, float& bW, float& bdet) {
    const Header* h = Hdr(blas);
    if (h->node_count == 0) return;
    const Node* nodes = NodesOf(blas);
    const uint32_t n = h->prim_count;
    const uint32_t first = n - 1;
    const bool boxes = (h->flags & kFlagAabbs) != 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Node& l = nodes[first + i];
        Tri3 th;
        if (Leaf(boxes, r, best_t, l, th) && th.t >= r.tmin && th.t <= best_t &&
            Closer(th.t, inst, l.c1, best_t, best_inst, best_prim)) {
            best_t = th.t; best_inst = inst; best_prim = l.c1;
            bV = th.V; bW = th.W; bdet = th.det;
        }
    }
}

int Threads(int t) { return t > 0 ? t : omp_get_max_threads(); }

}  // namespace

extern "C" {

void orc_mesh_bounds(const float* vertices, uint32_t stride, const uint32_t* indices, uint32_t tris, float lo[3],
                     float hi[3]) {
    Bounds(vertices, stride, indices, tris, lo, hi);
}

uint32_t orc_morton3(float x, float y, float z) {
    return (Spread10(Quantize(x)) << 2) | (Spread10(Quantize(y)) << 1) | Spread10(Quantize(z));
}

void orc_morton_codes(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t tris, uint32_t* codes) {
    float lo[3], hi[3];
    Bounds(vtx, stride, idx, tris, lo, hi);
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = hi[a] - lo[a];
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)tris; ++t) {
        float v[9], c[3];
        Tri(vtx, stride, idx, (uint32_t)t, v);
        for (int a = 0; a < 3; ++a) {
            const float l = Min(Min(v[a], v[3 + a]), v[6 + a]);
            const float h = Max(Max(v[a], v[3 + a]), v[6 + a]);
            c[a] = (l + h) * 0.5f;
        }
        codes[t] = orc_morton3(Unit(c[0], lo[0], ext[0]), Unit(c[1], lo[1], ext[1]), Unit(c[2], lo[2], ext[2]));
    }
}

// Stable sort of (key, value) pairs by key: LSD radix sort, 8-bit digits,
// four passes.  Each thread counts and then scatters one contiguous chunk;
// a digit's output run takes the chunks in order, which keeps equal keys in
// input order (the property std::stable_sort gives, algos_test.h:378).
void orc_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n) {
    if (n < 2) return;
    std::vector<uint32_t> k2(n), v2(n);
    uint32_t *ki = keys, *vi = vals, *ko = k2.data(), *vo = v2.data();
    const int threads = std::max(1, std::min(omp_get_max_threads(), (int)(n / 4096) + 1));
    std::vector<uint64_t> count((size_t)threads * 256);
    for (int shift = 0; shift < 32; shift += 8) {
#pragma omp parallel num_threads(threads)
        {
            // The runtime may start fewer threads than asked for: split the
            // keys over the threads that run.
            const int nt = omp_get_num_threads();
            const int t = omp_get_thread_num();
            const uint64_t b = (uint64_t)n * t / nt, e = (uint64_t)n * (t + 1) / nt;
            uint64_t* c = count.data() + (size_t)t * 256;
            std::fill(c, c + 256, 0);
            for (uint64_t i = b; i < e; ++i) ++c[(ki[i] >> shift) & 255u];
#pragma omp barrier
#pragma omp single
            {
                uint64_t run = 0;
                for (int d = 0; d < 256; ++d)
                    for (int u = 0; u < nt; ++u) {
                        const uint64_t x = count[(size_t)u * 256 + d];
                        count[(size_t)u * 256 + d] = run;
                        run += x;
                    }
            }
            for (uint64_t i = b; i < e; ++i) {
                const uint64_t o = c[(ki[i] >> shift) & 255u]++;
                ko[o] = ki[i];
                vo[o] = vi[i];
            }
        }
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    // Four passes: the sorted pairs are back in keys / vals.
}

void orc_set_threads(int threads) {
    if (threads > 0) omp_set_num_threads(threads);
}

void orc_emit_hierarchy(const uint32_t* codes, const uint32_t* prims, uint32_t n, void* nodes) {
    if (n) Emit(codes, prims, n, static_cast<Node*>(nodes));
}

void orc_refit(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t tris, void* nodes) {
    Refit(vtx, stride, idx, tris, static_cast<Node*>(nodes));
}

// Three rounds with at least 64, 128, 256 primitives per treelet root
// (restructure_hlbvh.cpp:34-35,151,276), treelets of `leaves` leaves
// (5, 7 or 9; 7 in the reference).
void orc_restructure_leaves(uint32_t tris, void* nodes, int leaves) {
    if (tris < 64) return;
    uint32_t m = 64;
    for (int round = 0; round < 3; ++round, m *= 2) {
        if (leaves == 5) RestructureRound<5>(static_cast<Node*>(nodes), tris, m);
        else if (leaves == 9) RestructureRound<9>(static_cast<Node*>(nodes), tris, m);
        else RestructureRound<7>(static_cast<Node*>(nodes), tris, m);
    }
}

void orc_restructure(uint32_t tris, void* nodes) { orc_restructure_leaves(tris, nodes, kLeaves); }

size_t orc_blas_bytes(uint32_t tris) { return sizeof(Header) + Align((size_t)NodeCount(tris) * 64, 256); }

void orc_build_blas_leaves(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t n, int restructure,
                           int leaves, void* buffer) {
    std::memset(buffer, 0, orc_blas_bytes(n));
    Header* h = static_cast<Header*>(buffer);
    h->magic = kMagic;
    h->kind = kBlas;
    h->prim_count = n;
    h->node_count = NodeCount(n);
    h->nodes_offset = sizeof(Header);
    h->flags = (restructure && n >= 64) ? 1u : 0u;
    h->build_serial = 1;
    if (!n) return;
    std::vector<uint32_t> codes(n), prims(n);
    orc_morton_codes(vtx, stride, idx, n, codes.data());
    for (uint32_t i = 0; i < n; ++i) prims[i] = i;
    orc_sort_pairs(codes.data(), prims.data(), n);
    Node* nodes = NodesOf(buffer);
    Emit(codes.data(), prims.data(), n, nodes);
    Refit(vtx, stride, idx, n, nodes);
    if (restructure) orc_restructure_leaves(n, nodes, leaves);
}

void orc_build_blas(const float* vtx, uint32_t stride, const uint32_t* idx, uint32_t n, int restructure, void* buffer) {
    orc_build_blas_leaves(vtx, stride, idx, n, restructure, kLeaves, buffer);
}

size_t orc_tlas_bytes(uint32_t n) {
    return sizeof(Header) + Align((size_t)NodeCount(n) * 64, 256) + Align((size_t)n * 96, 256) + Align((size_t)n * 8, 256);
}

void orc_build_tlas(const float* xforms, const uint64_t* blas_addr, const void* const* blas_host, uint32_t n,
                    void* buffer) {
    std::memset(buffer, 0, orc_tlas_bytes(n));
    Header* h = static_cast<Header*>(buffer);
    h->magic = kMagic;
    h->kind = kTlas;
    h->prim_count = n;
    h->node_count = NodeCount(n);
    h->nodes_offset = sizeof(Header);
    h->transforms_offset = h->nodes_offset + Align((size_t)NodeCount(n) * 64, 256);
    h->blas_table_offset = h->transforms_offset + Align((size_t)n * 96, 256);
    h->build_serial = 1;
    if (!n) return;
    Xf* xf = reinterpret_cast<Xf*>(static_cast<uint8_t*>(buffer) + h->transforms_offset);
    uint64_t* table = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(buffer) + h->blas_table_offset);
    std::vector<Box> wb(n);
    constexpr float kWiden = 1.52587890625e-05f;  // 2^-16
    for (uint32_t i = 0; i < n; ++i) {
        Xf m;
        std::memcpy(m.m, xforms + 12 * (size_t)i, 48);
        // Affine inverse: adjugate / determinant, translation mapped back.
        const float(*a)[4] = m.m;
        Xf r;
        const float c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        const float c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
        const float c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        const float det = (a[0][0] * c00 + a[0][1] * c01) + a[0][2] * c02;
        const float inv = 1.0f / det;
        r.m[0][0] = c00 * inv;
        r.m[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv;
        r.m[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv;
        r.m[1][0] = c01 * inv;
        r.m[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv;
        r.m[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv;
        r.m[2][0] = c02 * inv;
        r.m[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv;
        r.m[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * inv;
        for (int k = 0; k < 3; ++k) r.m[k][3] = -((r.m[k][0] * a[0][3] + r.m[k][1] * a[1][3]) + r.m[k][2] * a[2][3]);
        xf[2 * (size_t)i] = r;
        xf[2 * (size_t)i + 1] = m;
        table[i] = blas_addr[i];
        Box& w = wb[i];
        for (int k = 0; k < 3; ++k) { w.lo[k] = FLT_MAX; w.hi[k] = -FLT_MAX; }
        const Header* bh = Hdr(blas_host[i]);
        if (bh->node_count) {
            const Node& root = NodesOf(blas_host[i])[0];
            float blo[3], bhi[3];
            for (int k = 0; k < 3; ++k) {
                if (bh->node_count == 1) {
                    blo[k] = Min(Min(root.a0min[k], root.a0max[k]), root.a1min[k]);
                    bhi[k] = Max(Max(root.a0min[k], root.a0max[k]), root.a1min[k]);
                } else {
                    blo[k] = Min(root.a0min[k], root.a1min[k]);
                    bhi[k] = Max(root.a0max[k], root.a1max[k]);
                }
            }
            float mag[3] = {0, 0, 0};
            for (int c = 0; c < 8; ++c) {
                const float x = (c & 4) ? bhi[0] : blo[0];
                const float y = (c & 2) ? bhi[1] : blo[1];
                const float z = (c & 1) ? bhi[2] : blo[2];
                for (int k = 0; k < 3; ++k) {
                    const float* mm = m.m[k];
                    const float p = ((mm[0] * x + mm[1] * y) + mm[2] * z) + mm[3];
                    const float g = ((std::fabs(mm[0] * x) + std::fabs(mm[1] * y)) + std::fabs(mm[2] * z)) + std::fabs(mm[3]);
                    w.lo[k] = Min(w.lo[k], p);
                    w.hi[k] = Max(w.hi[k], p);
                    mag[k] = Max(mag[k], g);
                }
            }
            for (int k = 0; k < 3; ++k) {
                const float e = mag[k] * kWiden;
                w.lo[k] = w.lo[k] - e;
                w.hi[k] = w.hi[k] + e;
            }
        }
    }
    // Scene bounds over instance boxes, Morton codes of their centroids.
    uint32_t mn[3] = {kInv, kInv, kInv}, mx[3] = {0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        if (Hdr(blas_host[i])->node_count == 0) continue;
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], ToOrdered(wb[i].lo[k]));
            mx[k] = std::max(mx[k], ToOrdered(wb[i].hi[k]));
        }
    }
    float lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = FromOrdered(mn[k]);
        ext[k] = FromOrdered(mx[k]) - lo[k];
    }
    std::vector<uint32_t> codes(n), ids(n);
    for (uint32_t i = 0; i < n; ++i) {
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = (wb[i].lo[k] + wb[i].hi[k]) * 0.5f;
        codes[i] = orc_morton3(Unit(c[0], lo[0], ext[0]), Unit(c[1], lo[1], ext[1]), Unit(c[2], lo[2], ext[2]));
        ids[i] = i;
    }
    orc_sort_pairs(codes.data(), ids.data(), n);
    Node* nodes = NodesOf(buffer);
    Emit(codes.data(), ids.data(), n, nodes);
    // Refit over instance boxes (lbvh_fit_aabb_scene.comp:98-167).
    const uint32_t nint = n - 1;
    std::vector<Box> box(NodeCount(n));
    std::vector<std::pair<uint32_t, bool>> st{{0, false}};
    while (!st.empty()) {
        auto [x, e] = st.back();
        st.pop_back();
        if (x >= nint) {
            Node& l = nodes[x];
            const Box& b = wb[l.c1];
            for (int k = 0; k < 3; ++k) {
                l.a0min[k] = b.lo[k]; l.a0max[k] = b.hi[k];
                l.a1min[k] = b.lo[k]; l.a1max[k] = b.hi[k];
            }
            box[x] = b;
            continue;
        }
        if (!e) {
            st.push_back({x, true});
            st.push_back({nodes[x].c0, false});
            st.push_back({nodes[x].c1, false});
            continue;
        }
        Node& nd = nodes[x];
        for (int k = 0; k < 3; ++k) {
            nd.a0min[k] = box[nd.c0].lo[k]; nd.a0max[k] = box[nd.c0].hi[k];
            nd.a1min[k] = box[nd.c1].lo[k]; nd.a1max[k] = box[nd.c1].hi[k];
            box[x].lo[k] = Min(box[nd.c0].lo[k], box[nd.c1].lo[k]);
            box[x].hi[k] = Max(box[nd.c0].hi[k], box[nd.c1].hi[k]);
        }
        nd.update = 0;
    }
}

void orc_trace(const void* accel, const void* const* blas_host, const float* rays, uint32_t ray_count, int any,
               OrcHit* hits, OrcStats* stats, int threads) {
#pragma omp parallel for schedule(dynamic, 1024) num_threads(Threads(threads))
    for (int64_t i = 0; i < (int64_t)ray_count; ++i)
        hits[i] = TraceOne(accel, blas_host, rays + 8 * i, any != 0, stats ? stats + i : nullptr);
}

void orc_brute_force(const void* accel, const void* const* blas_host, const float* rays, uint32_t ray_count,
                     OrcHit* hits, int threads) {
    const Header* h = Hdr(accel);
    const bool two = h->kind == kTlas;
#pragma omp parallel for schedule(dynamic, 16) num_threads(Threads(threads))
    for (int64_t i = 0; i < (int64_t)ray_count; ++i) {
        const float* ray = rays + 8 * i;
        float best_t = ray[7], bV = 0, bW = 0, bdet = 1;
        uint32_t best_inst = kInv, best_prim = kInv;
        RayState r;
        if (!two) {
            Setup(r, ray, ray[3], ray + 4);
            BruteBlas(accel, r, 0u, best_t, best_inst, best_prim, bV, bW, bdet);
        } else {
            const Xf* xf = XformsOf(accel);
            for (uint32_t k = 0; k < h->prim_count; ++k) {
                float oo[3], dd[3];
                TransformRay(xf[2 * (size_t)k], ray, ray + 4, oo, dd);
                Setup(r, oo, ray[3], dd);
                BruteBlas(blas_host[k], r, k, best_t, best_inst, best_prim, bV, bW, bdet);
            }
        }
        OrcHit o = MissHit(ray[7]);
        if (best_prim != kInv) {
            o.u = bV / bdet;
            o.v = bW / bdet;
            o.inst_id = best_inst;
            o.prim_id = best_prim;
            o.t = best_t;
        }
        hits[i] = o;
    }
}

int orc_triangle_fp64(const float* ray, const float* v0, const float* v1, const float* v2, double eps, double* t,
                      double* u, double* v) {
    const double o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
    double e1[3], e2[3], s[3];
    for (int a = 0; a < 3; ++a) {
        e1[a] = (double)v1[a] - v0[a];
        e2[a] = (double)v2[a] - v0[a];
        s[a] = o[a] - v0[a];
    }
    const double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return 0;
    const double inv = 1.0 / det;
    const double uu = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    const double vv = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    *t = tt;
    *u = uu;
    *v = vv;
    return (uu >= -eps && vv >= -eps && uu + vv <= 1.0 + eps) ? 1 : 0;
}

void orc_treelet_dp(const float* boxes, float* cost, uint32_t* part) {
    Box lb[kLeaves];
    for (int k = 0; k < kLeaves; ++k)
        for (int a = 0; a < 3; ++a) {
            lb[k].lo[a] = boxes[6 * k + a];
            lb[k].hi[a] = boxes[6 * k + 3 + a];
        }
    float sa[kMasks];
    TreeletDp<kLeaves>(lb, sa, cost, part);
}

int orc_is_valid(const void* accel) {
    // bvh_analyzer/bvh.h:129-184: BFS; parent links agree; each internal
    // node's child boxes lie in the box its parent holds for it; every node
    // reached exactly once.
    const Header* h = Hdr(accel);
    const uint32_t total = h->node_count;
    if (total == 0) return 1;
    const Node* nodes = NodesOf(accel);
    const uint32_t nint = h->prim_count - 1;
    std::vector<uint8_t> seen(total, 0);
    std::queue<std::pair<uint32_t, uint32_t>> q;
    q.push({0, kInv});
    seen[0] = 1;
    size_t visited = 1;
    auto includes = [](const float lo[3], const float hi[3], const float clo[3], const float chi[3]) {
        for (int a = 0; a < 3; ++a)
            if (clo[a] - lo[a] < -1e-8f || hi[a] - chi[a] < -1e-8f) return false;
        return true;
    };
    while (!q.empty()) {
        auto [x, parent] = q.front();
        q.pop();
        const Node& nd = nodes[x];
        if (nd.parent != parent) return 0;
        if (x >= nint && total > 1) continue;
        if (total == 1) break;
        const float *plo = nullptr, *phi = nullptr;
        if (parent != kInv) {
            const Node& p = nodes[parent];
            if (p.c0 == x) { plo = p.a0min; phi = p.a0max; }
            else { plo = p.a1min; phi = p.a1max; }
        }
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = c ? nd.c1 : nd.c0;
            if (ch >= total || seen[ch]) return 0;
            seen[ch] = 1;
            ++visited;
            q.push({ch, x});
            if (plo && !includes(plo, phi, c ? nd.a1min : nd.a0min, c ? nd.a1max : nd.a0max)) return 0;
        }
    }
    return visited == total ? 1 : 0;
}

int orc_check_consistency(const void* accel) {
    // test/test_vk/hlbvh_test.h:69-100 (parent links, parent box contains
    // child boxes with 1e-8 slack), tightened: an internal node's slot holds
    // exactly its children's union and a leaf's slot exactly its triangle's
    // box (the refit invariant).
    const Header* h = Hdr(accel);
    const uint32_t total = h->node_count;
    if (total == 0) return 1;
    const Node* nodes = NodesOf(accel);
    const uint32_t nint = h->prim_count - 1;
    const bool tlas = h->kind == kTlas;
    std::queue<std::pair<uint32_t, uint32_t>> q;
    q.push({0, kInv});
    size_t visited = 0;
    while (!q.empty()) {
        auto [x, parent] = q.front();
        q.pop();
        ++visited;
        if (visited > total) return 0;
        const Node& nd = nodes[x];
        if (nd.parent != parent) return 0;
        const bool leaf = nd.c0 == kInv;
        if (leaf != (x >= nint || total == 1)) return 0;
        Box own;
        if (leaf) {
            for (int a = 0; a < 3; ++a) {
                own.lo[a] = tlas ? nd.a0min[a] : Min(Min(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
                own.hi[a] = tlas ? nd.a0max[a] : Max(Max(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                own.lo[a] = Min(nd.a0min[a], nd.a1min[a]);
                own.hi[a] = Max(nd.a0max[a], nd.a1max[a]);
            }
            q.push({nd.c0, x});
            q.push({nd.c1, x});
        }
        if (parent != kInv) {
            const Node& p = nodes[parent];
            const float* slo = p.c0 == x ? p.a0min : p.a1min;
            const float* shi = p.c0 == x ? p.a0max : p.a1max;
            for (int a = 0; a < 3; ++a)
                if (!(slo[a] == own.lo[a] && shi[a] == own.hi[a])) return 0;
        }
    }
    return visited == total ? 1 : 0;
}

float orc_sah(const void* accel) {
    // test/test_vk/common.h:242-267 and bvh_analyzer/bvh.h:186-223: sum over
    // every node (triangles included) of its box area over the root's, in
    // depth-first order (child1 popped first).
    const Header* h = Hdr(accel);
    if (h->node_count == 0) return 0.0f;
    const Node* nodes = NodesOf(accel);
    auto box_area = [&](uint32_t x) {
        const Node& nd = nodes[x];
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            if (nd.c0 != kInv) {
                lo[a] = std::min(nd.a0min[a], nd.a1min[a]);
                hi[a] = std::max(nd.a0max[a], nd.a1max[a]);
            } else {
                lo[a] = std::min(std::min(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
                hi[a] = std::max(std::max(nd.a0min[a], nd.a0max[a]), nd.a1min[a]);
            }
        }
        const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        return 2 * (ex * ey + ex * ez + ey * ez);
    };
    const float root = box_area(0);
    float total = 0.0f;
    std::stack<uint32_t> st;
    st.push(0);
    while (!st.empty()) {
        const uint32_t x = st.top();
        st.pop();
        total += box_area(x) / root;
        if (nodes[x].c0 != kInv) {
            st.push(nodes[x].c0);
            st.push(nodes[x].c1);
        }
    }
    return total;
}

uint32_t orc_height(const void* accel) {
    const Header* h = Hdr(accel);
    if (h->node_count == 0) return 0;
    const Node* nodes = NodesOf(accel);
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st{{0, 1}};
    while (!st.empty()) {
        auto [x, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        if (nodes[x].c0 != kInv) {
            st.push_back({nodes[x].c0, d + 1});
            st.push_back({nodes[x].c1, d + 1});
        }
    }
    return best;
}

int orc_write_analyzer_dump(const void* accel, const float* rays, uint32_t width, uint32_t height, const char* bvh_path,
                            const char* rays_path, const char* config_path) {
    const Header* h = Hdr(accel);
    FILE* f = std::fopen(bvh_path, "wb");
    if (!f) return 1;
    std::fwrite(NodesOf(accel), 64, h->node_count, f);
    std::fclose(f);
    f = std::fopen(rays_path, "wb");
    if (!f) return 1;
    std::fwrite(rays, 32, (size_t)width * height, f);
    std::fclose(f);
    f = std::fopen(config_path, "w");
    if (!f) return 1;
    // bvh_path, type, internal node count, triangle count, rays path, width, height.
    std::fprintf(f, "%s\nvkbvh2\n%u\n%u\n%s\n%u\n%u\n", bvh_path, h->prim_count - 1, h->prim_count, rays_path, width,
                 height);
    std::fclose(f);
    return 0;
}

}  // extern "C"
