// Accel: host build + gfx950 traversal kernels + the C ABI that replaces pine's
// BVH (src/impl/accel/bvh.h:72-85, bvh.cpp:612-709).
//
// Traversal: one ray per lane, ordered depth-first descent (nearest child
// first, as pine orders by slab distance, bvh.cpp:584-592), with the
// per-lane stack kept in LDS, laid out [depth][lane] so that the 64 lanes of a
// wave touch 64 consecutive dwords whatever their depths.  Node records are
// read as 16-byte vector loads.  Nodes store the boxes of their children, so
// one node fetch tests all children at once.
#include "bvh_build.h"
#include "pg_api.h"
#include "pg_isect.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace pg {

thread_local std::string g_last_error;

// sort.hip
hipError_t reorder_sort_temp_bytes(int64_t n, int end_bit, size_t& bytes);
hipError_t reorder_sort(void* temp, size_t bytes, const uint32_t* keys_in, uint32_t* keys_out,
                        uint32_t* perm, int64_t n, int end_bit, hipStream_t s);
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace pg

using namespace pg;

#define PG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return pg_fail(PG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct HostShape {
    int32_t kind;        // 0 mesh, 1 analytic
    int32_t first_prim;  // global id of its first prim
    int32_t count;
};

struct pg_accel {
    int device = 0;
    // Host copies in global prim id order.
    std::vector<int32_t> prim_kind;   // 0 triangle, 1 analytic
    std::vector<int32_t> prim_index;  // index into tri_v / ana_rec
    std::vector<float> tri_v;         // 9 floats per triangle: v0, v1, v2
    std::vector<float> ana_rec;       // 16 floats per analytic record (prim slot = 0)
    std::vector<float> ana_box;       // 6 floats per analytic record
    std::vector<HostShape> shapes;
    std::vector<int32_t> prim_shape;  // global prim -> shape index
    // Device copies.
    float4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_ana = nullptr;
    int width = 2;
    int max_depth = 0;
    int64_t n_nodes = 0;
    int64_t n_tri_slots = 0, n_ana_slots = 0;
    double build_ms = 0.0;
    bool built = false;
    int block = 128;
    int32_t launch_flags = PG_LAUNCH_DEFAULT;
    int stack_entries = 1;  // LDS stack entries per lane, from the built depth
    int64_t generation = 0;  // bumped by every build and launch change (graph reuse key)
    // Ray reordering (PG_LAUNCH_REORDER): the hierarchy's bounds for the key,
    // and the sort's device buffers.
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    uint32_t* d_keys = nullptr;  // 2 x sort_cap: keys, then the sorted keys
    int32_t* d_perm = nullptr;   // key-order slot -> caller's slot
    void* d_sort_tmp = nullptr;  // the radix sort's temporary storage
    size_t sort_tmp_bytes = 0;
    int64_t sort_cap = 0;
    // Stackless traversal (PG_LAUNCH_STACKLESS): child codes by path key.
    int2* d_keytab = nullptr;
    int keytab_bits = 0;
    // PG_LAUNCH_TIMED: events around the last launch's traversal kernel.
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    bool timed_pending = false;

    void free_sort() {
        if (d_keys)
            (void)hipFree(d_keys);
        if (d_perm)
            (void)hipFree(d_perm);
        if (d_sort_tmp)
            (void)hipFree(d_sort_tmp);
        d_keys = nullptr;
        d_perm = nullptr;
        d_sort_tmp = nullptr;
        sort_tmp_bytes = 0;
        sort_cap = 0;
    }
    void free_device() {
        if (d_nodes)
            (void)hipFree(d_nodes);
        if (d_tris)
            (void)hipFree(d_tris);
        if (d_ana)
            (void)hipFree(d_ana);
        if (d_keytab)
            (void)hipFree(d_keytab);
        d_nodes = nullptr;
        d_tris = nullptr;
        d_ana = nullptr;
        d_keytab = nullptr;
        keytab_bits = 0;
        free_sort();
    }
    void free_events() {
        if (ev_start)
            (void)hipEventDestroy(ev_start);
        if (ev_end)
            (void)hipEventDestroy(ev_end);
        ev_start = ev_end = nullptr;
        timed_pending = false;
    }
};

// ---------------------------------------------------------------------------
// Device side

namespace {

// LDS stack entries per lane: a binary node pushes at most one sibling per
// level, a 4-wide node at most three.  The kernels size the stack to the built
// tree (dynamic LDS), so shallow trees leave room for more waves per CU.
constexpr int kMaxStackEntries = 64;
inline int stack_entries(int width, int max_depth) {
    return std::max(1, (width - 1) * max_depth);
}

struct RayIn {
    f3 o, d;
    float tmin, tmax;
    f3 inv, oinv;
};

__device__ __forceinline__ RayIn load_ray(const pg_ray* rays, int64_t i) {
    const float4* r4 = reinterpret_cast<const float4*>(rays + i);
    float4 a = r4[0], b = r4[1];
    RayIn r;
    r.o = mk(a.x, a.y, a.z);
    r.tmin = a.w;
    r.d = mk(b.x, b.y, b.z);
    r.tmax = b.w;
    // pine's SafeRcp (vecmath.h:1095-1100)
    r.inv.x = r.d.x == 0.0f ? 1e20f : 1.0f / r.d.x;
    r.inv.y = r.d.y == 0.0f ? 1e20f : 1.0f / r.d.y;
    r.inv.z = r.d.z == 0.0f ? 1e20f : 1.0f / r.d.z;
    r.oinv = mk(-r.o.x * r.inv.x, -r.o.y * r.inv.y, -r.o.z * r.inv.z);
    return r;
}

// Slab test of one box; returns entry distance or +inf-like miss flag.
// The ray's tmin and the running limit reach the node loop from other basic
// blocks, so fmaxf/fminf on them get a canonicalising v_max_f32 (IEEE mode
// quiets signalling NaNs) at every node.  The two combinations that take them
// are written as the hardware min/max instructions instead: the same result
// for every input that is not a signalling NaN (rays and limits are plain
// floats), two instructions fewer per node visit.
__device__ __forceinline__ float max3_raw(float a, float b, float c) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float min3_raw(float a, float b, float c) {
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ bool slab(const RayIn& r, float lx, float ly, float lz, float hx,
                                     float hy, float hz, float tmax, float& tnear) {
    float x0 = __builtin_fmaf(lx, r.inv.x, r.oinv.x), x1 = __builtin_fmaf(hx, r.inv.x, r.oinv.x);
    float y0 = __builtin_fmaf(ly, r.inv.y, r.oinv.y), y1 = __builtin_fmaf(hy, r.inv.y, r.oinv.y);
    float z0 = __builtin_fmaf(lz, r.inv.z, r.oinv.z), z1 = __builtin_fmaf(hz, r.inv.z, r.oinv.z);
    const float tn = max3_raw(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1), r.tmin);
    const float tf = min3_raw(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1), tmax);
    tnear = tn;
    return tn <= tf;
}

// kMode: 0 closest hit, 1 any hit with pine's Hit tests, 2 occlusion with
// pine's Intersect tests (first accepted prim ends the ray).
enum TraceMode : int { kClosest = 0, kAnyHit = 1, kOccluded = 2 };

template <int kMode>
__device__ __forceinline__ bool leaf_test(int32_t code, const float4* __restrict__ tris,
                                          const float4* __restrict__ ana, const RayIn& r,
                                          float& best_t, int32_t& best_prim, float& bu, float& bv,
                                          int& tested) {
    const int count = ((code >> kLeafCountShift) & 15) + 1;
    tested += count;
    const int32_t start = code & kLeafStartMask;
    if (!(code & kLeafAnalytic)) {
        for (int k = 0; k < count; k++) {
            const float4* t = tris + 3 * (int64_t)(start + k);
            float4 a = t[0], b = t[1], c = t[2];
            f3 v0 = mk(a.x, a.y, a.z), e1 = mk(b.x, b.y, b.z), e2 = mk(c.x, c.y, c.z);
            if (kMode == kAnyHit) {
                if (tri_any(r.o, r.d, r.tmin, r.tmax, v0, e1, e2))
                    return true;
            } else {
                bool acc = tri_closest(r.o, r.d, r.tmin, v0, e1, e2, __float_as_int(a.w), best_t,
                                       best_prim, bu, bv);
                if (kMode == kOccluded && acc)
                    return true;
            }
        }
    } else {
        for (int k = 0; k < count; k++) {
            const float4* q = ana + 4 * (int64_t)(start + k);
            float rec[16];
            float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
            rec[0] = q0.x, rec[1] = q0.y, rec[2] = q0.z, rec[3] = q0.w;
            rec[4] = q1.x, rec[5] = q1.y, rec[6] = q1.z, rec[7] = q1.w;
            rec[8] = q2.x, rec[9] = q2.y, rec[10] = q2.z, rec[11] = q2.w;
            rec[12] = q3.x, rec[13] = q3.y, rec[14] = q3.z, rec[15] = q3.w;
            if (kMode == kAnyHit) {
                if (analytic_any(rec, r.o, r.d, r.tmin, r.tmax))
                    return true;
            } else {
                bool acc = analytic_closest(rec, r.o, r.d, r.tmin, best_t, best_prim, bu, bv);
                if (kMode == kOccluded && acc)
                    return true;
            }
        }
    }
    return false;
}

// MI355X dispatches workgroup b to XCD b % 8, and each XCD has its own L2.
// With the remap, XCD x works through one contiguous range of rays (its
// share of blocks, in order), so neighbouring rays - which touch the same
// nodes - meet in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t logical_block(bool remap) {
    const int64_t b = blockIdx.x;
    if (!remap)
        return b;
    const int64_t g = gridDim.x, q = g / kXcds, rem = g % kXcds;
    const int64_t x = b % kXcds, k = b / kXcds;
    return x * q + (x < rem ? x : rem) + k;
}

struct Stack {
    int32_t* base;
    int lane, stride, sp;
    __device__ __forceinline__ void push(int32_t v) {
        base[sp * stride + lane] = v;
        sp++;
    }
    // Next node from the stack; false when it is empty (traversal finished).
    __device__ __forceinline__ bool pop(int32_t& node) {
        if (sp == 0)
            return false;
        sp--;
        node = base[sp * stride + lane];
        return true;
    }
};

// Binary node, 64 B = {lo0.xyz, hi0.x}, {hi0.yz, lo1.xy}, {lo1.z, hi1.xyz}, {c0, c1, -, -}.
// On return `node` is the next node to visit; `live` turns false when the
// stack ran empty.
__device__ __forceinline__ void visit2(const float4* __restrict__ nodes, const RayIn& r,
                                       float lim, Stack& st, int32_t& node, bool& live) {
    const float4* nd = nodes + 4 * (int64_t)node;
    float4 a = nd[0], b = nd[1], c = nd[2], d = nd[3];
    const int32_t c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
    float t0, t1;
    bool h0 = slab(r, a.x, a.y, a.z, a.w, b.x, b.y, lim, t0) && c0 != kEmptyChild;
    bool h1 = slab(r, b.z, b.w, c.x, c.y, c.z, c.w, lim, t1) && c1 != kEmptyChild;
    if (h0 && h1) {
        st.push(t0 <= t1 ? c1 : c0);
        node = t0 <= t1 ? c0 : c1;
    } else if (h0 || h1) {
        node = h0 ? c0 : c1;
    } else {
        live = st.pop(node);
    }
}

// 4-wide node, 128 B = lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4], child[4], pad.
// Hit children are ordered by entry distance; all but the nearest are pushed,
// farthest first.
__device__ __forceinline__ void visit4(const float4* __restrict__ nodes, const RayIn& r,
                                       float lim, Stack& st, int32_t& node, bool& live) {
    const float4* nd = nodes + 8 * (int64_t)node;
    float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
    int4 ch = *reinterpret_cast<const int4*>(nd + 6);
    float t[4];
    int32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
    bool h[4];
    h[0] = slab(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, lim, t[0]) && c[0] != kEmptyChild;
    h[1] = slab(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, lim, t[1]) && c[1] != kEmptyChild;
    h[2] = slab(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, lim, t[2]) && c[2] != kEmptyChild;
    h[3] = slab(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, lim, t[3]) && c[3] != kEmptyChild;
    float key[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        key[k] = h[k] ? t[k] : __int_as_float(0x7f800000);
#define PG_CSWAP(a, b)          \
    if (key[b] < key[a]) {      \
        float tk = key[a];      \
        key[a] = key[b];        \
        key[b] = tk;            \
        int32_t tc = c[a];      \
        c[a] = c[b];            \
        c[b] = tc;              \
    }
    PG_CSWAP(0, 1)
    PG_CSWAP(2, 3)
    PG_CSWAP(0, 2)
    PG_CSWAP(1, 3)
    PG_CSWAP(1, 2)
#undef PG_CSWAP
    const int nhit = (int)h[0] + (int)h[1] + (int)h[2] + (int)h[3];
    if (nhit == 0) {
        live = st.pop(node);
        return;
    }
    for (int k = nhit - 1; k >= 1; k--)
        st.push(c[k]);
    node = c[0];
}

template <int W>
__device__ __forceinline__ void visit(const float4* __restrict__ nodes, const RayIn& r,
                                      float lim, Stack& st, int32_t& node, bool& live) {
    if (W == 2)
        visit2(nodes, r, lim, st, node, live);
    else
        visit4(nodes, r, lim, st, node, live);
}

// One ray per lane.  kPostpone selects the traversal loop:
//  - false: each iteration visits one node or tests one leaf, so a wave whose
//    lanes are at different kinds of node executes both paths;
//  - true: leaves are postponed (Aila and Laine's while-while loop with
//    speculative traversal): a lane that reaches a leaf records it and keeps
//    descending until every lane of the wave holds a leaf, then the wave tests
//    its leaves together.  Any valid visiting order returns the same closest
//    hit, because ties are broken by (t, prim id).
template <int W, int kBlock, int kMode, bool kStats, bool kPostpone>
__global__ __launch_bounds__(kBlock) void k_trace(const float4* __restrict__ nodes,
                                                  const float4* __restrict__ tris,
                                                  const float4* __restrict__ ana,
                                                  const pg_ray* __restrict__ rays, int64_t n,
                                                  pg_hit* __restrict__ hits,
                                                  pg_hit_uv* __restrict__ uvs,
                                                  int32_t* __restrict__ occluded,
                                                  int2* __restrict__ counts, int xcd_remap,
                                                  const int32_t* __restrict__ d_n,
                                                  const int32_t* __restrict__ perm) {
    extern __shared__ int32_t lds_stack[];
    const int lane = threadIdx.x;
    // d_n: the ray count in device memory (a launch for a queue's capacity, as
    // in a captured bounce loop): the grid is capped (launch_grid) and strides
    // over the count, so a bounce with few rays costs a small launch.  With a
    // host count the grid covers the rays and each lane traces one.
    const int64_t cnt = d_n ? (int64_t)*d_n : n, step = (int64_t)gridDim.x * kBlock;
    for (int64_t i = logical_block(xcd_remap != 0) * kBlock + lane; i < cnt; i += step) {
        // With reordering, slot i traces ray perm[i] (key order) and its result
        // goes to that ray's slot.
        const int64_t o = perm ? (int64_t)perm[i] : i;
        const RayIn r = load_ray(rays, o);
        float best_t = r.tmax, bu = 0.0f, bv = 0.0f;
        int32_t best_prim = -1;
        Stack st{lds_stack, lane, kBlock, 0};
        int32_t node = 0;
        bool live = true, found_any = false;
        int visited = 0, tested = 0;
        if (!kPostpone) {
            while (live) {
                if (node >= 0) {
                    visited++;
                    visit<W>(nodes, r, kMode == kAnyHit ? r.tmax : best_t, st, node, live);
                } else {
                    if (leaf_test<kMode>(node, tris, ana, r, best_t, best_prim, bu, bv, tested)) {
                        found_any = true;
                        break;
                    }
                    live = st.pop(node);
                }
            }
        } else {
            int32_t leaf = 0;  // postponed leaf code; leaf codes are negative, 0 = none
            while (live) {
                while (live && node >= 0) {
                    visited++;
                    visit<W>(nodes, r, kMode == kAnyHit ? r.tmax : best_t, st, node, live);
                    if (live && node < 0 && leaf == 0) {
                        leaf = node;
                        live = st.pop(node);
                    }
                    if (__all(leaf != 0))
                        break;
                }
                while (leaf != 0) {
                    if (leaf_test<kMode>(leaf, tris, ana, r, best_t, best_prim, bu, bv, tested)) {
                        found_any = true;
                        live = false;
                        break;
                    }
                    leaf = 0;
                    if (live && node < 0) {
                        leaf = node;
                        live = st.pop(node);
                    }
                }
            }
        }
        if (kStats)
            counts[o] = make_int2(visited, tested);
        if (kMode != kClosest) {
            occluded[o] = found_any ? 1 : 0;
        } else {
            pg_hit h;
            h.prim = best_prim;
            h.t = best_prim >= 0 ? best_t : r.tmax;
            hits[o] = h;
            if (uvs) {
                pg_hit_uv uv;
                uv.u = bu;
                uv.v = bv;
                uvs[o] = uv;
            }
        }
    }
}

// --- stackless traversal with wave-wide node fetches (binary nodes) ---------
// No stack.  Every child slot of the tree has a path key: the root's is 1 and
// a child's is 2 x its parent's key + its slot (0 or 1), so a key spells the
// child choices from the root.  A table built with the hierarchy maps each
// key to its child code (inner node or leaf).  Per ray the kernel keeps the
// current item's key and level and a 32-bit trail with bit L set while the far
// child chosen at level L is still to be visited.  Children are visited
// nearest first, by the same slab distances as the stack traversal, so the
// same nodes are visited.  Backtracking takes the deepest pending level k,
// forms the key of the far child (the ancestor of the current key at level
// k + 1, last bit flipped) and fetches its code from the table: one probe
// (about 1.1 slots at a table load of at most 1/2) instead of a stack pop,
// whatever the depth, or none when the far child is the one set aside last
// (Binder and Keller, "Efficient stackless hierarchy
// traversal on GPUs with backtracking in constant time", HPG 2016).  When
// every active lane of the wave is at the same node (the top levels, and
// longer for coherent rays) the node is fetched once by scalar loads.
// Wave-wide fetches are tried in the top kWaveFetchLevels levels only (as
// the wave's first lane sees them): deeper down the lanes rarely share a node
// and the test itself costs more than it saves (measured on the reordered
// 16M-ray benchmark: 4.23 ms with the test at every level, 4.10 ms without).
constexpr int kWaveFetchLevels = 4;

__device__ __forceinline__ void fetch2(const float4* __restrict__ nodes, int32_t P, int level, float4& a,
                                       float4& b, float4& c, float4& d) {
    const int32_t p0 = __builtin_amdgcn_readfirstlane(P);
    if (__builtin_amdgcn_readfirstlane(level) < kWaveFetchLevels && __all(P == p0)) {
        const float4* nd = nodes + 4 * (int64_t)p0;
        a = nd[0], b = nd[1], c = nd[2], d = nd[3];
    } else {
        const float4* nd = nodes + 4 * (int64_t)P;
        a = nd[0], b = nd[1], c = nd[2], d = nd[3];
    }
}

// Slot of `key` in a table of 2^bits slots (multiplicative hashing).
__host__ __device__ __forceinline__ uint32_t key_slot(uint32_t key, int bits) {
    return (key * 2654435761u) >> (32 - bits);
}

// The child code stored under `key` (always present: only far children that
// were hit are looked up).  Linear probing; the bound only guards a table
// that does not hold the key.
__device__ __forceinline__ int32_t key_code(const int2* __restrict__ tab, int bits, uint32_t key) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t h = key_slot(key, bits);
    for (uint32_t probe = 0; probe <= mask; probe++) {
        const int2 e = tab[h];
        if ((uint32_t)e.x == key)
            return e.y;
        h = (h + 1u) & mask;
    }
    return kEmptyChild;
}

// Backtracking of the stackless traversal: the next item is the far child of
// the deepest pending level (`trail` empty: the traversal is finished).
// `key`/`level` name the item just finished and become the next item's.
__device__ __forceinline__ bool backtrack(uint32_t& trail, uint32_t& key, int& level, int32_t& node,
                                          uint32_t& far_key, int32_t far_code, const int2* __restrict__ tab,
                                          int bits) {
    if (trail == 0u)
        return false;
    const int k = 31 - __clz((int)trail);
    trail &= ~(1u << k);
    key = (key >> (level - (k + 1))) ^ 1u;
    level = k + 1;
    if (key == far_key) {
        node = far_code;
        far_key = 0u;
    } else {
        node = key_code(tab, bits, key);
    }
    return true;
}

template <int kBlock, int kMode, bool kStats>
__global__ __launch_bounds__(kBlock) void k_trace_stackless(const float4* __restrict__ nodes,
                                                            const float4* __restrict__ tris,
                                                            const float4* __restrict__ ana,
                                                            const pg_ray* __restrict__ rays, int64_t n,
                                                            pg_hit* __restrict__ hits,
                                                            pg_hit_uv* __restrict__ uvs,
                                                            int32_t* __restrict__ occluded,
                                                            int2* __restrict__ counts, int xcd_remap,
                                                            const int32_t* __restrict__ d_n,
                                                            const int32_t* __restrict__ perm,
                                                            const int2* __restrict__ keytab, int keybits) {
    const int64_t cnt = d_n ? (int64_t)*d_n : n, step = (int64_t)gridDim.x * kBlock;
    for (int64_t i = logical_block(xcd_remap != 0) * kBlock + threadIdx.x; i < cnt; i += step) {
        // With reordering, slot i traces ray perm[i] (key order) and its result
        // goes to that ray's slot.
        const int64_t o = perm ? (int64_t)perm[i] : i;
        const RayIn r = load_ray(rays, o);
        float best_t = r.tmax, bu = 0.0f, bv = 0.0f;
        int32_t best_prim = -1;
        int visited = 0, tested = 0;
        bool found_any = false;
        int32_t node = 0;    // the current item: inner node index or leaf code
        uint32_t key = 1u;   // its path key
        int level = 0;       // its depth (root 0)
        uint32_t trail = 0;  // bit L: the far child chosen at level L is pending
        // The far child set aside last (key 0: none): the deepest pending one as
        // long as nothing was set aside after it, which saves the table probe on
        // about 60% of backtracks.
        uint32_t far_key = 0u;
        int32_t far_code = 0;
        bool live = true;
        while (live) {
            if (node >= 0) {
                visited++;
                float4 a, b, c, d;
                fetch2(nodes, node, level, a, b, c, d);
                const int32_t c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
                const float lim = kMode == kAnyHit ? r.tmax : best_t;
                float t0, t1;
                const bool h0 = slab(r, a.x, a.y, a.z, a.w, b.x, b.y, lim, t0) && c0 != kEmptyChild;
                const bool h1 = slab(r, b.z, b.w, c.x, c.y, c.z, c.w, lim, t1) && c1 != kEmptyChild;
                if (h0 || h1) {
                    const uint32_t pick = h0 && h1 ? (t0 <= t1 ? 0u : 1u) : (h0 ? 0u : 1u);
                    if (h0 && h1) {
                        trail |= 1u << level;
                        far_key = 2u * key + (1u - pick);
                        far_code = pick ? c0 : c1;
                    }
                    node = pick ? c1 : c0;
                    key = 2u * key + pick;
                    level++;
                    continue;
                }
            } else if (leaf_test<kMode>(node, tris, ana, r, best_t, best_prim, bu, bv, tested)) {
                found_any = true;
                break;
            }
            // Every far child that was hit is in the table; a missing key
            // (kEmptyChild) is skipped rather than followed.
            do {
                live = backtrack(trail, key, level, node, far_key, far_code, keytab, keybits);
            } while (live && node == kEmptyChild);
        }
        if (kStats)
            counts[o] = make_int2(visited, tested);
        if (kMode != kClosest) {
            occluded[o] = found_any ? 1 : 0;
        } else {
            pg_hit h;
            h.prim = best_prim;
            h.t = best_prim >= 0 ? best_t : r.tmax;
            hits[o] = h;
            if (uvs) {
                pg_hit_uv uv;
                uv.u = bu;
                uv.v = bv;
                uvs[o] = uv;
            }
        }
    }
}

template <int kMode, bool kStats>
void launch_stackless(const pg_accel* a, int64_t grid, const pg_ray* rays, int64_t n, pg_hit* hits,
                      pg_hit_uv* uvs, int32_t* occ, int2* counts, hipStream_t s, const int32_t* d_n,
                      const int32_t* perm) {
    const int remap = (a->launch_flags & PG_LAUNCH_XCD_REMAP) ? 1 : 0;
#define PG_LAUNCH(B)                                                                              \
    hipLaunchKernelGGL((k_trace_stackless<B, kMode, kStats>), dim3((unsigned)grid), dim3(B), 0, s, \
                       a->d_nodes, a->d_tris, a->d_ana, rays, n, hits, uvs, occ, counts, remap, d_n, \
                       perm, a->d_keytab, a->keytab_bits)
    if (a->block == 64)
        PG_LAUNCH(64);
    else if (a->block == 256)
        PG_LAUNCH(256);
    else
        PG_LAUNCH(128);
#undef PG_LAUNCH
}

template <int W, int kMode, bool kStats, bool kPostpone>
void launch_w(const pg_accel* a, int64_t grid, size_t lds, const pg_ray* rays, int64_t n,
              pg_hit* hits, pg_hit_uv* uvs, int32_t* occ, int2* counts, hipStream_t s,
              const int32_t* d_n, const int32_t* perm) {
    const int remap = (a->launch_flags & PG_LAUNCH_XCD_REMAP) ? 1 : 0;
#define PG_LAUNCH(B)                                                                           \
    hipLaunchKernelGGL((k_trace<W, B, kMode, kStats, kPostpone>), dim3((unsigned)grid), dim3(B), \
                       lds, s, a->d_nodes, a->d_tris, a->d_ana, rays, n, hits, uvs, occ,         \
                       counts, remap, d_n, perm)
    if (a->block == 64)
        PG_LAUNCH(64);
    else if (a->block == 256)
        PG_LAUNCH(256);
    else
        PG_LAUNCH(128);
#undef PG_LAUNCH
}

// --- ray reordering (PG_LAUNCH_REORDER) ---------------------------------------
// Rays are traced in the order of a key: the octant of the direction (the
// major 3 bits), then the Morton code of the origin quantised to 2^c cells per
// axis of the hierarchy's bounds.  Rays that start close together and head the
// same way then share a wave and the nodes it fetches (fewer divergent
// iterations, more L2 hits); every ray's result is the same in any order.  A
// key kernel writes one key per ray; a device radix sort (sort.hip) turns the
// keys into perm, the ray index of each key-order slot; the trace kernel of
// slot i reads ray perm[i] and writes that ray's result.  Host-count launches
// use 7 bits per axis (24-bit keys, three 8-bit digits); device-count
// launches, sized for a capacity, use 6 bits per axis and mark the slots past
// the count with bit 21 so that they sort last (22-bit keys).
constexpr int64_t kReorderMinRays = 65536;

struct ReorderFrame {
    float lo[3], scale[3];  // cell = (o - lo) * scale, clamped to [0, 2^bits - 1]
    int bits;
    uint32_t past_count;  // key of slots past the device count
};

__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t cell_of(float o, float lo, float scale, float top) {
    // fmaxf/fminf map NaN to the bounds, so the conversion is always defined.
    return (uint32_t)fminf(fmaxf((o - lo) * scale, 0.0f), top);
}

__global__ __launch_bounds__(256) void k_reorder_keys(const pg_ray* __restrict__ rays, int64_t n,
                                                      const int32_t* __restrict__ d_n, ReorderFrame f,
                                                      uint32_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    if (d_n && i >= (int64_t)*d_n) {
        keys[i] = f.past_count;
        return;
    }
    const float4* r4 = reinterpret_cast<const float4*>(rays + i);
    const float4 a = r4[0], b = r4[1];
    const uint32_t oct = ((b.x < 0.0f) ? 4u : 0u) | ((b.y < 0.0f) ? 2u : 0u) | ((b.z < 0.0f) ? 1u : 0u);
    const float top = (float)((1 << f.bits) - 1);
    const uint32_t m = (spread3(cell_of(a.x, f.lo[0], f.scale[0], top)) << 2) |
                       (spread3(cell_of(a.y, f.lo[1], f.scale[1], top)) << 1) |
                       spread3(cell_of(a.z, f.lo[2], f.scale[2], top));
    keys[i] = (oct << (3 * f.bits)) | m;
}

// Sizes the accel's sort buffers for `cap` rays (allocating only when
// `may_alloc`); `ready` tells whether they can serve the launch.
int ensure_reorder(pg_accel* a, int64_t cap, bool may_alloc, bool& ready) {
    ready = false;
    if (cap > (int64_t)INT32_MAX)
        return PG_OK;
    if (a->sort_cap < cap) {
        if (!may_alloc)
            return PG_OK;
        a->free_sort();
        size_t tmp = 0;
        PG_HIP(reorder_sort_temp_bytes(cap, 24, tmp));
        PG_HIP(hipMalloc(&a->d_keys, 2 * cap * sizeof(uint32_t)));
        PG_HIP(hipMalloc(&a->d_perm, cap * sizeof(int32_t)));
        PG_HIP(hipMalloc(&a->d_sort_tmp, tmp));
        a->sort_tmp_bytes = tmp;
        a->sort_cap = cap;
    }
    ready = true;
    return PG_OK;
}

// Enqueues the key kernel and the sort of rays[0, n) (or of the first *d_n
// of them) on stream s; afterwards a->d_perm holds the key order.
hipError_t enqueue_reorder(const pg_accel* a, const pg_ray* rays, int64_t n, const int32_t* d_n,
                           hipStream_t s) {
    ReorderFrame f;
    f.bits = d_n ? 6 : 7;
    f.past_count = 1u << (3 + 3 * f.bits);
    const int end_bit = d_n ? 4 + 3 * f.bits : 3 + 3 * f.bits;
    for (int k = 0; k < 3; k++) {
        const float ext = a->root_hi[k] - a->root_lo[k];
        f.lo[k] = a->root_lo[k];
        f.scale[k] = ext > 0.0f ? (float)(1 << f.bits) / ext : 0.0f;
    }
    hipLaunchKernelGGL(k_reorder_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rays, n, d_n, f,
                       a->d_keys);
    size_t need = 0;
    const hipError_t q = reorder_sort_temp_bytes(n, end_bit, need);
    if (q != hipSuccess)
        return q;
    if (need > a->sort_tmp_bytes)
        return hipErrorInvalidValue;  // not reached: storage was sized for 24-bit keys at capacity
    return reorder_sort(a->d_sort_tmp, a->sort_tmp_bytes, a->d_keys, a->d_keys + a->sort_cap,
                        reinterpret_cast<uint32_t*>(a->d_perm), n, end_bit, s);
}

// Workgroups of a traversal launch: enough for every ray of a host count; for
// a device-held count (at most n) at most kDevCountLanes lanes striding over
// the count.  A bounce's queue is sized for its capacity but often holds far
// fewer rays.  4M lanes is about eight per lane the GPU holds at once.
// Measured on the 512^2 x 64-spp spheres and 2^20-triangle mesh scenes
// (16M-path batches, M samples/s): 1M lanes 3,274 and 966, 4M 3,345 and
// 1,001, no cap 3,270 and 1,020.
constexpr int64_t kDevCountLanes = 1 << 22;

int64_t launch_grid(int64_t n, int block, bool device_count) {
    const int64_t g = (n + block - 1) / block;
    return device_count ? std::min<int64_t>(g, kDevCountLanes / block) : g;
}

// Launches the traversal over rays[0, n) (or the first *d_n).  With
// PG_LAUNCH_REORDER and at least kReorderMinRays rays it first sorts the rays
// by key; host-count launches may allocate the sort's buffers, device-count
// launches (which a caller may be recording into a graph) use them only when
// they already hold the capacity.
template <int kMode, bool kStats>
hipError_t launch_trace(pg_accel* a, const pg_ray* rays, int64_t n, pg_hit* hits,
                        pg_hit_uv* uvs, int32_t* occ, int2* counts, hipStream_t s,
                        const int32_t* d_n = nullptr) {
    const int64_t grid = launch_grid(n, a->block, d_n != nullptr);
    const size_t lds = (size_t)a->stack_entries * a->block * sizeof(int32_t);
    const bool postpone = (a->launch_flags & PG_LAUNCH_POSTPONE_LEAVES) != 0;
    const int32_t* perm = nullptr;
    if ((a->launch_flags & PG_LAUNCH_REORDER) && n >= kReorderMinRays) {
        bool ready = false;
        // A launch being recorded into a graph keeps the callers' order: the
        // sort's library code clears memory with memset nodes.
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        const hipError_t ce = hipStreamIsCapturing(s, &cap);
        if (ce != hipSuccess)
            return ce;
        if (cap == hipStreamCaptureStatusNone && ensure_reorder(a, n, d_n == nullptr, ready) != PG_OK)
            return hipErrorOutOfMemory;
        if (ready) {
            const hipError_t e = enqueue_reorder(a, rays, n, d_n, s);
            if (e != hipSuccess)
                return e;
            perm = a->d_perm;
        }
    }
    const bool timed = (a->launch_flags & PG_LAUNCH_TIMED) != 0;
    if (timed) {
        if (!a->ev_start) {
            const hipError_t e0 = hipEventCreate(&a->ev_start);
            if (e0 != hipSuccess)
                return e0;
            const hipError_t e1 = hipEventCreate(&a->ev_end);
            if (e1 != hipSuccess)
                return e1;
        }
        const hipError_t er = hipEventRecord(a->ev_start, s);
        if (er != hipSuccess)
            return er;
    }
    if (a->width == 2 && (a->launch_flags & PG_LAUNCH_STACKLESS)) {
        if (!a->d_keytab)
            return hipErrorInvalidValue;  // refused by pg_accel_set_launch / build before this point
        launch_stackless<kMode, kStats>(a, grid, rays, n, hits, uvs, occ, counts, s, d_n, perm);
    } else if (a->width == 2) {
        if (postpone)
            launch_w<2, kMode, kStats, true>(a, grid, lds, rays, n, hits, uvs, occ, counts, s, d_n, perm);
        else
            launch_w<2, kMode, kStats, false>(a, grid, lds, rays, n, hits, uvs, occ, counts, s, d_n, perm);
    } else {
        if (postpone)
            launch_w<4, kMode, kStats, true>(a, grid, lds, rays, n, hits, uvs, occ, counts, s, d_n, perm);
        else
            launch_w<4, kMode, kStats, false>(a, grid, lds, rays, n, hits, uvs, occ, counts, s, d_n, perm);
    }
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess && timed) {
        a->timed_pending = true;
        return hipEventRecord(a->ev_end, s);
    }
    return e;
}

}  // namespace

// Path keys fit 32 bits for hierarchies of up to 31 levels.
constexpr int kMaxKeyDepth = 31;

// The stackless traversal's table: every non-empty child slot of the binary
// hierarchy under its path key (root 1, child = 2 x parent + slot), open
// addressing with linear probing in 2^bits slots at a load of at most 1/2
// (empty slots hold key 0, which no child has).
static void build_key_table(const FlatBVH& flat, std::vector<int2>& tab, int& bits) {
    struct Item {
        int64_t at;
        uint32_t key;
    };
    std::vector<std::pair<uint32_t, int32_t>> entries;
    std::vector<Item> todo{{0, 1u}};
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        const float* node = &flat.nodes[(size_t)it.at * 16];
        for (int slot = 0; slot < 2; slot++) {
            int32_t code;
            memcpy(&code, node + 12 + slot, 4);
            if (code == kEmptyChild)
                continue;
            const uint32_t ck = 2u * it.key + (uint32_t)slot;
            entries.push_back({ck, code});
            if (code >= 0)
                todo.push_back({code, ck});
        }
    }
    bits = 4;
    while ((size_t(1) << bits) < 2 * entries.size())
        bits++;
    const uint32_t mask = (1u << bits) - 1u;
    tab.assign(size_t(1) << bits, make_int2(0, 0));
    for (const auto& e : entries) {
        uint32_t h = key_slot(e.first, bits);
        while (tab[h].x != 0)
            h = (h + 1u) & mask;
        tab[h] = make_int2((int32_t)e.first, e.second);
    }
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" const char* pg_last_error(void) { return pg::g_last_error.c_str(); }
extern "C" const char* pg_version(void) { return "pinehip 0.1 gfx950"; }

extern "C" int pg_accel_create(pg_accel** out) {
    if (!out)
        return pg_fail(PG_E_ARG, "pg_accel_create: out is null");
    pg_accel* a = new (std::nothrow) pg_accel();
    if (!a)
        return pg_fail(PG_E_NOMEM, "pg_accel_create: out of host memory");
    *out = a;
    return PG_OK;
}

static int check_open(pg_accel* a, const char* fn) {
    if (!a)
        return pg_fail(PG_E_ARG, std::string(fn) + ": null accel");
    if (a->built)
        return pg_fail(PG_E_STATE, std::string(fn) + ": accel already built");
    return PG_OK;
}

extern "C" int pg_accel_add_mesh(pg_accel* a, const float* v, int64_t nv, const uint32_t* idx,
                                 int64_t nt, int32_t* first_prim_out) {
    if (int rc = check_open(a, "pg_accel_add_mesh"))
        return rc;
    if (!v || !idx || nv <= 0 || nt <= 0)
        return pg_fail(PG_E_ARG, "pg_accel_add_mesh: empty mesh");
    // bvh.cpp:148 CHECK_NE(numPrimitives, 0) and geometry.h:390-394 index checks.
    const int64_t first = (int64_t)a->prim_kind.size();
    if (first + nt >= (int64_t)kLeafStartMask)
        return pg_fail(PG_E_ARG, "pg_accel_add_mesh: too many prims for one accel");
    for (int64_t k = 0; k < 3 * nt; k++)
        if ((int64_t)idx[k] >= nv)
            return pg_fail(PG_E_ARG, "pg_accel_add_mesh: index out of range");
    const int32_t shape = (int32_t)a->shapes.size();
    a->shapes.push_back({0, (int32_t)first, (int32_t)nt});
    const int64_t base = (int64_t)a->tri_v.size() / 9;
    a->tri_v.resize((base + nt) * 9);
    for (int64_t k = 0; k < nt; k++) {
        for (int c = 0; c < 3; c++) {
            const float* p = v + 3 * (int64_t)idx[3 * k + c];
            a->tri_v[(base + k) * 9 + 3 * c + 0] = p[0];
            a->tri_v[(base + k) * 9 + 3 * c + 1] = p[1];
            a->tri_v[(base + k) * 9 + 3 * c + 2] = p[2];
        }
        a->prim_kind.push_back(0);
        a->prim_index.push_back((int32_t)(base + k));
        a->prim_shape.push_back(shape);
    }
    if (first_prim_out)
        *first_prim_out = (int32_t)first;
    return PG_OK;
}

static int add_analytic(pg_accel* a, const float rec[16], const float box[6], int32_t* prim_out) {
    const int32_t prim = (int32_t)a->prim_kind.size();
    const int32_t shape = (int32_t)a->shapes.size();
    a->shapes.push_back({1, prim, 1});
    a->prim_kind.push_back(1);
    a->prim_index.push_back((int32_t)(a->ana_rec.size() / 16));
    a->prim_shape.push_back(shape);
    a->ana_rec.insert(a->ana_rec.end(), rec, rec + 16);
    a->ana_box.insert(a->ana_box.end(), box, box + 6);
    if (prim_out)
        *prim_out = prim;
    return PG_OK;
}

extern "C" int pg_accel_add_sphere(pg_accel* a, const float c[3], float r, int32_t* prim_out) {
    if (int rc = check_open(a, "pg_accel_add_sphere"))
        return rc;
    if (!c || !(r > 0.0f))
        return pg_fail(PG_E_ARG, "pg_accel_add_sphere: bad centre or radius");
    float rec[16] = {c[0], c[1], c[2], 0, r, 0, 0, u2f(kSphere), 0, 0, 0, 0, 0, 0, 0, 0};
    float box[6] = {c[0] - r, c[1] - r, c[2] - r, c[0] + r, c[1] + r, c[2] + r};
    return add_analytic(a, rec, box, prim_out);
}

extern "C" int pg_accel_add_rect(pg_accel* a, const float pos[3], const float ex[3],
                                 const float ey[3], int32_t* prim_out) {
    if (int rc = check_open(a, "pg_accel_add_rect"))
        return rc;
    if (!pos || !ex || !ey)
        return pg_fail(PG_E_ARG, "pg_accel_add_rect: null vector");
    // Rect ctor (geometry.h:337-343), with pine's Normalize/Length.
    f3 X = mk(ex[0], ex[1], ex[2]), Y = mk(ey[0], ey[1], ey[2]), P = mk(pos[0], pos[1], pos[2]);
    f3 exn = normalize(X), eyn = normalize(Y), n = normalize(cross(X, Y));
    float lx = length(X), ly = length(Y);
    float rec[16] = {P.x,   P.y,   P.z,   0,  exn.x, exn.y, exn.z, u2f(kRect),
                     eyn.x, eyn.y, eyn.z, lx, n.x,   n.y,   n.z,   ly};
    // Conservative box over the four corners.
    float box[6] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2) {
            f3 q = P + X * (0.5f * sx) + Y * (0.5f * sy);
            float qq[3] = {q.x, q.y, q.z};
            for (int k = 0; k < 3; k++) {
                box[k] = std::min(box[k], qq[k]);
                box[3 + k] = std::max(box[3 + k], qq[k]);
            }
        }
    return add_analytic(a, rec, box, prim_out);
}

extern "C" int pg_accel_add_plane(pg_accel* a, const float pos[3], const float normal[3],
                                  int32_t* prim_out) {
    if (int rc = check_open(a, "pg_accel_add_plane"))
        return rc;
    if (!pos || !normal)
        return pg_fail(PG_E_ARG, "pg_accel_add_plane: null vector");
    // Plane::Create normalises (geometry.cpp:340-342); ctor builds u, v (geometry.h:121-125).
    f3 n = normalize(mk(normal[0], normal[1], normal[2]));
    f3 u, v;
    coordinate_system(n, u, v);
    float rec[16] = {pos[0], pos[1], pos[2], 0, n.x, n.y, n.z, u2f(kPlane),
                     u.x,    u.y,    u.z,    0, v.x, v.y, v.z, 0};
    // Plane::GetAABB (geometry.cpp:110-112)
    float box[6] = {-1e6f, -1e6f, -1e6f, 1e6f, 1e6f, 1e6f};
    return add_analytic(a, rec, box, prim_out);
}

extern "C" int pg_accel_rebuild(pg_accel* a, const pg_build_options* opts) {
    if (!a)
        return pg_fail(PG_E_ARG, "pg_accel_rebuild: null accel");
    a->built = false;
    return pg_accel_build(a, opts);
}

extern "C" int pg_accel_build(pg_accel* a, const pg_build_options* opts) {
    if (int rc = check_open(a, "pg_accel_build"))
        return rc;
    BuildParams bp;
    if (opts) {
        if (opts->width != 0 && opts->width != 2 && opts->width != 4)
            return pg_fail(PG_E_ARG, "pg_accel_build: width must be 2 or 4");
        if (opts->leaf_max < 0 || opts->leaf_max > 8 || opts->bins < 0 || opts->bins > 256)
            return pg_fail(PG_E_ARG, "pg_accel_build: leaf_max 1..8, bins 4..256");
        if (opts->width)
            bp.width = opts->width;
        if (opts->leaf_max)
            bp.leaf_max = opts->leaf_max;
        if (opts->bins)
            bp.bins = opts->bins;
        bp.threads = opts->threads;
    }
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n = (int64_t)a->prim_kind.size();
    std::vector<BuildPrim> prims((size_t)n);
    float scene_mag = 0.0f;
    for (int64_t i = 0; i < n; i++) {
        BuildPrim& p = prims[i];
        p.id = (int32_t)i;
        p.kind = a->prim_kind[i];
        if (p.kind == 0) {
            const float* t = &a->tri_v[(size_t)a->prim_index[i] * 9];
            for (int k = 0; k < 3; k++) {
                p.lo[k] = std::min(t[k], std::min(t[3 + k], t[6 + k]));
                p.hi[k] = std::max(t[k], std::max(t[3 + k], t[6 + k]));
            }
        } else {
            const float* b = &a->ana_box[(size_t)a->prim_index[i] * 6];
            for (int k = 0; k < 3; k++) {
                p.lo[k] = b[k];
                p.hi[k] = b[3 + k];
            }
        }
        for (int k = 0; k < 3; k++)
            scene_mag = std::max(scene_mag, std::max(std::fabs(p.lo[k]), std::fabs(p.hi[k])));
    }
    // The hierarchy's bounds frame the reordering key's cells.
    float blo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, bhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (const BuildPrim& p : prims)
        for (int k = 0; k < 3; k++) {
            blo[k] = std::min(blo[k], p.lo[k]);
            bhi[k] = std::max(bhi[k], p.hi[k]);
        }
    // Grow child boxes by 2^-18 of the scene's magnitude: float slab tests then
    // never cull a box that holds the true closest hit.
    const float inflate = scene_mag * (1.0f / 262144.0f) + 1e-30f;
    FlatBVH flat = build_bvh(prims, bp, inflate);
    if (stack_entries(flat.width, flat.max_depth) > kMaxStackEntries)
        return pg_fail(PG_E_STATE, "pg_accel_build: hierarchy deeper than the traversal stack");

    // Leaf records in slot order.
    std::vector<float4> tris(flat.tri_order.size() * 3);
    for (size_t s = 0; s < flat.tri_order.size(); s++) {
        int32_t prim = flat.tri_order[s];
        const float* t = &a->tri_v[(size_t)a->prim_index[prim] * 9];
        f3 v0 = mk(t[0], t[1], t[2]), v1 = mk(t[3], t[4], t[5]), v2 = mk(t[6], t[7], t[8]);
        f3 e1 = v1 - v0, e2 = v2 - v0;  // geometry.h:272-273
        tris[3 * s + 0] = make_float4(v0.x, v0.y, v0.z, u2f((uint32_t)prim));
        tris[3 * s + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
        tris[3 * s + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    }
    std::vector<float4> ana(flat.ana_order.size() * 4);
    for (size_t s = 0; s < flat.ana_order.size(); s++) {
        int32_t prim = flat.ana_order[s];
        const float* q = &a->ana_rec[(size_t)a->prim_index[prim] * 16];
        ana[4 * s + 0] = make_float4(q[0], q[1], q[2], u2f((uint32_t)prim));
        ana[4 * s + 1] = make_float4(q[4], q[5], q[6], q[7]);
        ana[4 * s + 2] = make_float4(q[8], q[9], q[10], q[11]);
        ana[4 * s + 3] = make_float4(q[12], q[13], q[14], q[15]);
    }
    a->free_device();
    PG_HIP(hipGetDevice(&a->device));
    const size_t node_bytes = flat.nodes.size() * sizeof(float);
    PG_HIP(hipMalloc(&a->d_nodes, node_bytes));
    PG_HIP(hipMemcpy(a->d_nodes, flat.nodes.data(), node_bytes, hipMemcpyHostToDevice));
    if (!tris.empty()) {
        PG_HIP(hipMalloc(&a->d_tris, tris.size() * sizeof(float4)));
        PG_HIP(hipMemcpy(a->d_tris, tris.data(), tris.size() * sizeof(float4),
                         hipMemcpyHostToDevice));
    }
    if (!ana.empty()) {
        PG_HIP(hipMalloc(&a->d_ana, ana.size() * sizeof(float4)));
        PG_HIP(hipMemcpy(a->d_ana, ana.data(), ana.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (flat.width == 2 && flat.max_depth <= kMaxKeyDepth) {
        std::vector<int2> tab;
        int bits = 0;
        build_key_table(flat, tab, bits);
        PG_HIP(hipMalloc(&a->d_keytab, tab.size() * sizeof(int2)));
        PG_HIP(hipMemcpy(a->d_keytab, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice));
        a->keytab_bits = bits;
    }
    a->width = flat.width;
    a->max_depth = flat.max_depth;
    for (int k = 0; k < 3; k++) {
        a->root_lo[k] = n > 0 ? blo[k] : 0.0f;
        a->root_hi[k] = n > 0 ? bhi[k] : 0.0f;
    }
    a->stack_entries = stack_entries(flat.width, flat.max_depth);
    a->n_nodes = flat.n_nodes;
    a->n_tri_slots = (int64_t)flat.tri_order.size();
    a->n_ana_slots = (int64_t)flat.ana_order.size();
    a->built = true;
    a->generation++;
    a->build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PG_OK;
}

static int check_trace(pg_accel* a, const pg_ray* rays, int64_t n, const void* out,
                       const char* fn) {
    if (!a)
        return pg_fail(PG_E_ARG, std::string(fn) + ": null accel");
    if (!a->built)
        return pg_fail(PG_E_STATE, std::string(fn) + ": accel not built");
    if (n < 0 || (n > 0 && (!rays || !out)))
        return pg_fail(PG_E_ARG, std::string(fn) + ": bad ray/hit buffers");
    return PG_OK;
}

extern "C" int pg_trace_closest(pg_accel* a, const pg_ray* d_rays, int64_t n, pg_hit* d_hits,
                                pg_hit_uv* d_uv, void* stream) {
    if (int rc = check_trace(a, d_rays, n, d_hits, "pg_trace_closest"))
        return rc;
    if (n == 0)
        return PG_OK;
    PG_HIP((launch_trace<kClosest, false>(a, d_rays, n, d_hits, d_uv, nullptr, nullptr,
                                       (hipStream_t)stream)));
    return PG_OK;
}

extern "C" int pg_trace_any(pg_accel* a, const pg_ray* d_rays, int64_t n, int32_t* d_occluded,
                            void* stream) {
    if (int rc = check_trace(a, d_rays, n, d_occluded, "pg_trace_any"))
        return rc;
    if (n == 0)
        return PG_OK;
    PG_HIP((launch_trace<kAnyHit, false>(a, d_rays, n, nullptr, nullptr, d_occluded, nullptr,
                                      (hipStream_t)stream)));
    return PG_OK;
}

extern "C" int pg_trace_closest_dev(pg_accel* a, const pg_ray* d_rays, const int32_t* d_n,
                                    int64_t capacity, pg_hit* d_hits, pg_hit_uv* d_uv, void* stream) {
    if (int rc = check_trace(a, d_rays, capacity, d_hits, "pg_trace_closest_dev"))
        return rc;
    if (!d_n)
        return pg_fail(PG_E_ARG, "pg_trace_closest_dev: null count");
    if (capacity == 0)
        return PG_OK;
    PG_HIP((launch_trace<kClosest, false>(a, d_rays, capacity, d_hits, d_uv, nullptr, nullptr,
                                       (hipStream_t)stream, d_n)));
    return PG_OK;
}
extern "C" int pg_trace_occluded_dev(pg_accel* a, const pg_ray* d_rays, const int32_t* d_n,
                                     int64_t capacity, int32_t* d_occluded, void* stream) {
    if (int rc = check_trace(a, d_rays, capacity, d_occluded, "pg_trace_occluded_dev"))
        return rc;
    if (!d_n)
        return pg_fail(PG_E_ARG, "pg_trace_occluded_dev: null count");
    if (capacity == 0)
        return PG_OK;
    PG_HIP((launch_trace<kOccluded, false>(a, d_rays, capacity, nullptr, nullptr, d_occluded, nullptr,
                                          (hipStream_t)stream, d_n)));
    return PG_OK;
}
extern "C" int pg_trace_occluded(pg_accel* a, const pg_ray* d_rays, int64_t n,
                                  int32_t* d_occluded, void* stream) {
    if (int rc = check_trace(a, d_rays, n, d_occluded, "pg_trace_occluded"))
        return rc;
    if (n == 0)
        return PG_OK;
    PG_HIP((launch_trace<kOccluded, false>(a, d_rays, n, nullptr, nullptr, d_occluded, nullptr,
                                          (hipStream_t)stream)));
    return PG_OK;
}

extern "C" int pg_trace_closest_counts(pg_accel* a, const pg_ray* d_rays, int64_t n,
                                       pg_hit* d_hits, int32_t* d_counts, void* stream) {
    if (int rc = check_trace(a, d_rays, n, d_hits, "pg_trace_closest_counts"))
        return rc;
    if (n > 0 && !d_counts)
        return pg_fail(PG_E_ARG, "pg_trace_closest_counts: null counts buffer");
    if (n == 0)
        return PG_OK;
    PG_HIP((launch_trace<kClosest, true>(a, d_rays, n, d_hits, nullptr, nullptr,
                                      reinterpret_cast<int2*>(d_counts), (hipStream_t)stream)));
    return PG_OK;
}

extern "C" int pg_accel_get_stats(const pg_accel* a, pg_accel_stats* out) {
    if (!a || !out)
        return pg_fail(PG_E_ARG, "pg_accel_get_stats: null argument");
    out->n_prims = (int64_t)a->prim_kind.size();
    out->n_nodes = a->n_nodes;
    out->node_bytes = a->n_nodes * (a->width == 2 ? 64 : 128);
    out->prim_bytes = a->n_tri_slots * 48 + a->n_ana_slots * 64;
    out->width = a->width;
    out->max_depth = a->max_depth;
    out->build_ms = a->build_ms;
    out->block = a->block;
    out->launch_flags = a->launch_flags;
    out->last_trace_ms = -1.0;
    if (a->timed_pending) {
        float ms = 0.0f;
        PG_HIP(hipEventSynchronize(a->ev_end));
        PG_HIP(hipEventElapsedTime(&ms, a->ev_start, a->ev_end));
        out->last_trace_ms = ms;
    }
    return PG_OK;
}

extern "C" int pg_accel_prim_shape(const pg_accel* a, int32_t prim, int32_t* shape,
                                   int32_t* local) {
    if (!a || !shape || !local)
        return pg_fail(PG_E_ARG, "pg_accel_prim_shape: null argument");
    if (prim < 0 || prim >= (int32_t)a->prim_shape.size())
        return pg_fail(PG_E_ARG, "pg_accel_prim_shape: prim out of range");
    *shape = a->prim_shape[prim];
    *local = prim - a->shapes[*shape].first_prim;
    return PG_OK;
}

extern "C" int pg_accel_set_launch(pg_accel* a, int32_t block, int32_t flags) {
    if (!a)
        return pg_fail(PG_E_ARG, "pg_accel_set_launch: null accel");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return pg_fail(PG_E_ARG, "pg_accel_set_launch: block must be 64, 128 or 256");
    if (flags < -1 ||
        flags > (PG_LAUNCH_POSTPONE_LEAVES | PG_LAUNCH_XCD_REMAP | PG_LAUNCH_STACKLESS | PG_LAUNCH_REORDER |
                 PG_LAUNCH_TIMED))
        return pg_fail(PG_E_ARG, "pg_accel_set_launch: unknown flags");
    if (flags >= 0 && (flags & PG_LAUNCH_STACKLESS) && a->built && a->width == 2 && !a->d_keytab)
        return pg_fail(PG_E_ARG, "pg_accel_set_launch: stackless traversal needs a hierarchy of at most 31 levels");
    if (block)
        a->block = block;
    if (flags >= 0)
        a->launch_flags = flags;
    a->generation++;
    return PG_OK;
}

int64_t pg::accel_generation(const pg_accel* a) { return a ? a->generation : 0; }

extern "C" int pg_accel_release(pg_accel* a) {
    if (!a)
        return PG_OK;
    a->free_device();
    a->free_events();
    delete a;
    return PG_OK;
}
