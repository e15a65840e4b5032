// HIP kernels of one resolve2d step (Solver.process, src/core/lib.zig:189-299)
// for gfx950, over a batch of independent worlds.
//
// Pipeline (every kernel reads its element counts from device memory so the
// whole step is a fixed launch sequence that can be captured in a hipGraph):
//
//   k_prep          per body: cos/sin of the angle, cell range of the stale AABB
//   scan            cell counts -> entry offsets                (E entries)
//   k_entries       per entry: bucket = hash(cell) % table, bucket counts
//   scan            bucket counts -> bucket offsets
//   k_bucket_fill   per entry: scatter body into its bucket
//   k_bucket_sort   per bucket: descending body order, duplicates dropped
//   k_ent_cnt       per entry: bodies in the bucket with a higher slot
//   scan            -> raw candidate offsets                    (R candidates)
//   k_cand_emit     per entry: (a, b) raw candidates in the reference's query order
//   k_narrow        per candidate: first occurrence, filters, SAT
//   scan            collides flags -> manifold index             (M manifolds)
//   k_man_build     per manifold: contact points (identifyCollisionPoints)
//   k_world_ranges  per world: manifold / item ranges
//   k_items_*       solver items and their ordering resources
//   scan, k_list_*  per-body ordered item lists -> ranks
//   k_solve         one workgroup per world: sub-steps, forces, integration,
//                   pre-step and the sequential-impulse iterations, run as a
//                   dataflow that reproduces the reference's sequential order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "r2d_device.h"
#include "r2d_kernels.h"

namespace r2d {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t grid_stride() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a < b ? b : a; }

// number of valid elements given a device count clamped to capacity
__device__ __forceinline__ uint32_t dev_count(const uint32_t* c, uint32_t cap)
{
    uint32_t v = *c;
    return v < cap ? v : cap;
}

// largest i in [0, n) with off[i] <= x (off ascending)
__device__ __forceinline__ uint32_t upper_idx(const uint32_t* off, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive block scan; blockDim multiple of 64, <= 1024
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* sm, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = wave_incl_scan(v);
    if (lane == 63) sm[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t t = lane < nw ? sm[lane] : 0u;
        t = wave_incl_scan(t);
        if (lane < nw) sm[lane] = t;
    }
    __syncthreads();
    uint32_t base = wid ? sm[wid - 1] : 0u;
    *total = sm[nw - 1];
    __syncthreads();
    return base + x - v;
}

// ------------------------------------------------------------- scan (u32)
// Exclusive scan of n = min(*n_dev, cap) (or n_const) values into out[0..n],
// out[n] = total; three fixed-grid kernels.
constexpr int SCAN_BLOCKS = 256;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;

__device__ __forceinline__ uint32_t scan_n(const uint32_t* n_dev, uint32_t n_const, uint32_t cap)
{
    return n_dev ? dev_count(n_dev, cap) : n_const;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* in, const uint32_t* n_dev,
                                                               uint32_t n_const, uint32_t cap, uint32_t* partials)
{
    __shared__ uint32_t sm[16];
    uint32_t n = scan_n(n_dev, n_const, cap);
    uint32_t chunk = (n + SCAN_BLOCKS - 1) / SCAN_BLOCKS;
    uint32_t b0 = blockIdx.x * chunk, b1 = umin(n, b0 + chunk);
    uint32_t s = 0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) s += in[i];
    uint32_t tot;
    block_excl_scan(s, sm, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(uint32_t* partials, uint32_t* out,
                                                                 const uint32_t* n_dev, uint32_t n_const,
                                                                 uint32_t cap, uint32_t* total_dev)
{
    __shared__ uint32_t sm[16];
    uint32_t n = scan_n(n_dev, n_const, cap);
    uint32_t v = threadIdx.x < SCAN_BLOCKS ? partials[threadIdx.x] : 0u;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, sm, &tot);
    if (threadIdx.x < SCAN_BLOCKS) partials[threadIdx.x] = ex;
    if (threadIdx.x == 0) {
        out[n] = tot;
        if (total_dev) *total_dev = tot;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t* in, uint32_t* out,
                                                             const uint32_t* n_dev, uint32_t n_const,
                                                             uint32_t cap, const uint32_t* partials)
{
    __shared__ uint32_t sm[16];
    uint32_t n = scan_n(n_dev, n_const, cap);
    uint32_t chunk = (n + SCAN_BLOCKS - 1) / SCAN_BLOCKS;
    uint32_t b0 = blockIdx.x * chunk, b1 = umin(n, b0 + chunk);
    uint32_t base = partials[blockIdx.x];
    const uint32_t tile = SCAN_THREADS * SCAN_ITEMS;
    for (uint32_t t0 = b0; t0 < b1; t0 += tile) {
        uint32_t v[SCAN_ITEMS];
        uint32_t s = 0;
        uint32_t i0 = t0 + threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = (i0 + j < b1) ? in[i0 + j] : 0u;
            s += v[j];
        }
        uint32_t tot;
        uint32_t ex = block_excl_scan(s, sm, &tot);
        uint32_t run = base + ex;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            if (i0 + j < b1) out[i0 + j] = run;
            run += v[j];
        }
        base += tot;
    }
}

void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_const, uint32_t cap,
              uint32_t* total_dev, uint32_t* partials, hipStream_t st)
{
    hipLaunchKernelGGL(k_scan_reduce, dim3(SCAN_BLOCKS), dim3(SCAN_THREADS), 0, st, in, n_dev, n_const, cap, partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_THREADS), 0, st, partials, out, n_dev, n_const, cap,
                       total_dev);
    hipLaunchKernelGGL(k_scan_down, dim3(SCAN_BLOCKS), dim3(SCAN_THREADS), 0, st, in, out, n_dev, n_const, cap,
                       partials);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = grid_tid(); i < n; i += grid_stride()) p[i] = 0u;
}

// ---------------------------------------------------------- body geometry
// The parts of a body that SAT and contact generation read.  World vertices
// are add2(rotate2(local_vertex, angle), pos) with the angle's cos/sin taken
// once per step (identical bits to recomputing them per call).
struct Geo {
    v2 pos;
    float r;          // disc radius
    bool rect;
    uint32_t id;
    v2 lv[4];         // Rectangle.local_vertices (Rectangle.zig:40)
    v2 wv[4];         // world vertices
    float c, s;       // cos / sin of the angle
};

// B: a type with the body arrays of WorldDev (px, py, flags, rad, id, hw, hh)
template <class B>
__device__ __forceinline__ Geo load_geo(const B& D, uint32_t slot, float c, float s)
{
    Geo g;
    g.pos = mk(D.px[slot], D.py[slot]);
    uint32_t f = D.flags[slot];
    g.rect = (f & BF_RECT) != 0;
    g.r = D.rad[slot];
    g.id = D.id[slot];
    g.c = c; g.s = s;
    float hw = D.hw[slot], hh = D.hh[slot];
    g.lv[0] = mk(-hw, -hh); g.lv[1] = mk(-hw, hh); g.lv[2] = mk(hw, hh); g.lv[3] = mk(hw, -hh);
#pragma unroll
    for (int i = 0; i < 4; ++i) g.wv[i] = add2(rotcs(g.lv[i], c, s), g.pos);
    return g;
}

// Disc.closestPoint (Disc.zig:80-84); Rectangle.closestPoint (Rectangle.zig:114-130)
__device__ __forceinline__ v2 closest_point(const Geo& b, v2 p)
{
    if (!b.rect) {
        v2 n = normalize2(sub2(p, b.pos));
        return addmult2(b.pos, n, b.r);
    }
    float best = __builtin_inff();
    v2 bp = p;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float d2 = length2sq(sub2(b.wv[i], p));
        if (d2 < best) { best = d2; bp = b.wv[i]; }
    }
    return bp;
}

struct Edge {
    v2 dir, a, b;
};

// a[i] for a run-time i with the array kept in registers (selects, not a
// dynamically indexed private array, which would live in scratch memory)
__device__ __forceinline__ v2 pick4(const v2 (&a)[4], int i)
{
    v2 r = a[0];
    r = i == 1 ? a[1] : r;
    r = i == 2 ? a[2] : r;
    r = i == 3 ? a[3] : r;
    return r;
}

// Disc.getNormal (Disc.zig:86-90); Rectangle.getNormal (Rectangle.zig:132-151)
__device__ __forceinline__ Edge get_normal(const Geo& self, const Geo& other, int iter)
{
    Edge e;
    if (!self.rect) {
        v2 closest = closest_point(other, self.pos);
        e.dir = normalize2(sub2(closest, self.pos));
        e.a = e.b = mk(0.f, 0.f);
        return e;
    }
    int nx = (iter == 3) ? 0 : iter + 1;
    v2 a1 = pick4(self.wv, iter), a2 = pick4(self.wv, nx);
    e.dir = rot90ccw(normalize2(sub2(a2, a1)));
    e.a = a1; e.b = a2;
    return e;
}

// projectAlongAxis: Disc.zig:92-97, Rectangle.zig:153-169
__device__ __forceinline__ void project(const Geo& b, v2 n, float& lo, float& hi)
{
    if (!b.rect) {
        float mid = dot2(b.pos, n);
        lo = mid - b.r; hi = mid + b.r;
        return;
    }
    lo = __builtin_inff(); hi = -__builtin_inff();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float d = dot2(b.wv[i], n);
        if (d < lo) lo = d;
        if (d > hi) hi = d;
    }
}

__device__ __forceinline__ bool should_flip(v2 n, const Geo& ref, const Geo& inc)
{
    return dot2(n, sub2(inc.pos, ref.pos)) < 0.0f;   // collision.zig:221-226
}

struct Sat {
    v2 normal;
    float pen;
    int nid;
    int ref;          // 0: first argument of performNarrowSAT's pair order, 1: second
    bool set;
};

// overlapSAT (collision.zig:228-289); `which` tags the reference body
__device__ __forceinline__ bool overlap_sat(Sat& ret, const Geo& ref, const Geo& inc, int which)
{
    const float EPS = SAT_OVERLAP_THRESHOLD;
    int nn = ref.rect ? 4 : 1;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        if (it >= nn) break;
        Edge e = get_normal(ref, inc, it);
        v2 normal = e.dir;
        bool flipped = false;
        if (should_flip(normal, ref, inc)) { normal = negself(normal); flipped = true; }
        float p1lo, p1hi, p2lo, p2hi;
        project(ref, normal, p1lo, p1hi);
        project(inc, normal, p2lo, p2hi);
        float d1 = p1hi - p2lo;
        float d2 = p2hi - p1lo;
        float d = minn(d1, d2);
        if (d <= -COLLISION_MARGIN) return false;
        if (!flipped && approx2(normal, ret.normal, EPS)) {
            float tref = dot2(ref.pos, normal);
            float tinc = dot2(inc.pos, normal);
            if (tref < tinc - EPS) { ret.pen = d; ret.normal = normal; ret.nid = it; ret.ref = which; ret.set = true; }
        }
        if (!flipped && approx2(normal, negate2(ret.normal), EPS)) {
            v2 diff = sub2(inc.pos, ref.pos);
            if (dot2(diff, normal) > dot2(diff, ret.normal) + EPS) {
                ret.pen = d; ret.normal = normal; ret.nid = it; ret.ref = which; ret.set = true;
            }
        }
        if (d + EPS < ret.pen) { ret.pen = d; ret.normal = normal; ret.nid = it; ret.ref = which; ret.set = true; }
    }
    return true;
}

// aabb.intersects (aabb.zig:12-19)
template <class B>
__device__ __forceinline__ bool aabb_hit(const B& D, uint32_t a, uint32_t b)
{
    float dx = absf(D.ax[b] - D.ax[a]);
    float dy = absf(D.ay[b] - D.ay[a]);
    return (dx <= (D.ahw[a] + D.ahw[b]) + AABB_EPS_OVERLAP) && (dy <= (D.ahh[a] + D.ahh[b]) + AABB_EPS_OVERLAP);
}

__device__ __forceinline__ bool excluded(const WorldDev& D, uint32_t ida, uint32_t idb)
{
    uint64_t key = ida < idb ? (((uint64_t)ida << 32) | idb) : (((uint64_t)idb << 32) | ida);
    int lo = 0, hi = D.nexcl;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        uint64_t v = D.excl[mid];
        if (v == key) return true;
        if (v < key) lo = mid + 1; else hi = mid;
    }
    return false;
}

// ------------------------------------------------------------ broadphase
__global__ void k_fill_bworld(StepDev S)
{
    // the descriptor's fields read once: the stores below could alias them,
    // so reading them in the loop would reload them every iteration
    const uint32_t n = (uint32_t)S.worlds[blockIdx.x].nbodies, base = S.worlds[blockIdx.x].body_base;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) S.bworld[base + i] = blockIdx.x;
}

// One body's broadphase preparation: cos / sin of the angle, the scheduling
// flag (a static body orders contacts only if a contact's "+= 0" could change
// it: a -0 component; joint-touched bodies are added later) and the cell
// range of SpatialHash.iterateAABBHashes (SpatialHash.zig:83-96) from the
// stale AABB.  err: 3 cell range too large to enumerate, 4 non-finite AABB
// (@intFromFloat is undefined in the reference), 0 otherwise.
struct Prep {
    float c, s;
    uint32_t res;
    int64_t x0, y0;
    uint32_t nx, ny, ncell;
    int err;
};

__device__ __forceinline__ Prep prep_body(const WorldDev& D, uint32_t sl)
{
    Prep o;
    float a = D.ang[sl];
    o.c = cosf_r(a);
    o.s = sinf_r(a);
    uint32_t f = D.flags[sl];
    o.res = 1u;
    if (f & BF_STATIC) {
        float m0 = D.mx[sl], m1 = D.my[sl], m2 = D.L[sl];
        bool negz = (m0 == 0.0f && __builtin_signbit(m0)) || (m1 == 0.0f && __builtin_signbit(m1)) ||
                    (m2 == 0.0f && __builtin_signbit(m2));
        o.res = negz ? 1u : 0u;
    }
    o.err = 0;
    o.ncell = 0;
    o.x0 = o.y0 = 0;
    o.nx = o.ny = 0;
    float cell = D.cell;
    float ax = D.ax[sl], ay = D.ay[sl], hw = D.ahw[sl], hh = D.ahh[sl];
    float v0 = floorf((ax - hw) / cell), v1 = floorf((ay - hh) / cell);
    float v2f = floorf((ax + hw) / cell), v3 = floorf((ay + hh) / cell);
    const float LIM = 4.0e18f;
    bool ok = (v0 > -LIM && v0 < LIM) && (v1 > -LIM && v1 < LIM) && (v2f > -LIM && v2f < LIM) &&
              (v3 > -LIM && v3 < LIM);
    if (ok) {
        int64_t x0 = (int64_t)v0, y0 = (int64_t)v1, x1 = (int64_t)v2f, y1 = (int64_t)v3;
        uint64_t nx = (x1 >= x0) ? (uint64_t)(x1 - x0 + 1) : 0u;
        uint64_t ny = (y1 >= y0) ? (uint64_t)(y1 - y0 + 1) : 0u;
        uint64_t tot = nx * ny;
        if (nx > (1u << 30) || ny > (1u << 30) || tot > (1u << 30)) {
            o.err = 3;
        } else {
            o.ncell = (uint32_t)tot;
            o.x0 = x0; o.y0 = y0; o.nx = (uint32_t)nx; o.ny = (uint32_t)ny;
        }
    } else {
        o.err = 4;
    }
    return o;
}

__global__ void k_prep(StepDev S)
{
    for (uint32_t i = grid_tid(); i < S.nb; i += grid_stride()) {
        uint32_t w = S.bworld[i];
        const WorldDev& D = S.worlds[w];
        const Prep o = prep_body(D, i - D.body_base);
        S.cc[i] = o.c;
        S.cs[i] = o.s;
        S.sflag[i] = o.res;
        if (o.err) S.w_status[w] = o.err;
        if (!o.err && o.ncell) { S.cx0[i] = o.x0; S.cy0[i] = o.y0; S.cnx[i] = o.nx; S.cny[i] = o.ny; }
        S.ncell[i] = o.ncell;
    }
}

// per cell entry: its body, its bucket; bucket occupancy counts
__global__ void k_entries(StepDev S)
{
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    if (S.counters[C_ENTRIES] > S.cap_e) { if (grid_tid() == 0) atomicOr(&S.counters[C_OVERFLOW], 1u); return; }
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t i = upper_idx(S.cell_off, S.nb, e);   // body whose cell range holds entry e
        uint32_t k = e - S.cell_off[i];
        const WorldDev& D = S.worlds[S.bworld[i]];
        uint32_t nx = S.cnx[i];
        int64_t yi = S.cy0[i] + (int64_t)(k / nx);
        int64_t xi = S.cx0[i] + (int64_t)(k % nx);
        uint32_t gb = D.bucket_base + (uint32_t)cell_hash((uint64_t)D.table, xi, yi);
        S.ent_body[e] = i;
        S.ent_bkt[e] = gb;
        atomicAdd(&S.bkt_cnt[gb], 1u);
    }
}

__global__ void k_bucket_fill(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t gb = S.ent_bkt[e];
        uint32_t pos = atomicAdd(&S.bkt_fill[gb], 1u);
        S.bkt_items[S.bkt_off[gb] + pos] = S.ent_body[e];
    }
}

// How many of l[0, c) are below v: four entries per iteration, their loads
// in flight together
__device__ __forceinline__ uint32_t count_below(const uint32_t* l, uint32_t c, uint32_t v)
{
    uint32_t r = 0, i = 0;
    for (; i + 4 <= c; i += 4) {
        const uint32_t a0 = l[i], a1 = l[i + 1], a2 = l[i + 2], a3 = l[i + 3];
        r += (a0 < v) + (a1 < v) + (a2 < v) + (a3 < v);
    }
    for (; i < c; ++i) r += l[i] < v;
    return r;
}

// A cell entry's rank among its bucket's members it[0, c): how many are
// greater than its body a and how many equal it.  Four members per
// iteration, so their loads are in flight together.
__device__ __forceinline__ void bucket_rank_of(const uint32_t* it, uint32_t c, uint32_t a, uint32_t& greater,
                                               uint32_t& same)
{
    uint32_t g = 0, e = 0, j = 0;
    for (; j + 4 <= c; j += 4) {
        const uint32_t v0 = it[j], v1 = it[j + 1], v2 = it[j + 2], v3 = it[j + 3];
        g += (v0 > a) + (v1 > a) + (v2 > a) + (v3 > a);
        e += (v0 == a) + (v1 == a) + (v2 == a) + (v3 == a);
    }
    for (; j < c; ++j) {
        const uint32_t v = it[j];
        g += v > a;
        e += v == a;
    }
    greater = g;
    same = e;
}

// Within a bucket the reference stores entries in reverse insertion order
// (pre-decrement scatter, SpatialHash.zig:36-40, 59-68): descending body slot.
// Each entry's place is the number of bucket members with a higher slot plus
// its rank among its own body's earlier entries in the same bucket (a body
// listed twice in one bucket -- two of its cells hash alike -- sits in
// adjacent places; readers skip the repeat, which cannot change which pair
// appears first).
__global__ void k_bucket_rank(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t a = S.ent_body[e], gb = S.ent_bkt[e];
        uint32_t o = S.bkt_off[gb], n = S.bkt_cnt[gb];
        const uint32_t* it = S.bkt_items + o;
        uint32_t greater, same;
        bucket_rank_of(it, n, a, greater, same);
        uint32_t dup = 0;
        if (same > 1)
            for (uint32_t x = S.cell_off[a]; x < e; ++x) dup += S.ent_bkt[x] == gb;
        S.bkt_sorted[o + greater + dup] = a;
    }
}

// raw candidates per entry: distinct bucket members with a higher slot than
// the entry's body
__global__ void k_ent_cnt(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t a = S.ent_body[e], gb = S.ent_bkt[e];
        const uint32_t* it = S.bkt_sorted + S.bkt_off[gb];
        uint32_t n = S.bkt_cnt[gb], c = 0;
        for (uint32_t j = 0; j < n && it[j] > a; ++j) c += (j == 0 || it[j] != it[j - 1]);
        S.ent_cnt[e] = c;
    }
}

__global__ void k_cand_emit(StepDev S)
{
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    if (S.counters[C_OVERFLOW]) return;
    if (S.counters[C_RAW] > S.cap_r) { if (grid_tid() == 0) atomicOr(&S.counters[C_OVERFLOW], 2u); return; }
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t a = S.ent_body[e], gb = S.ent_bkt[e];
        const uint32_t* it = S.bkt_sorted + S.bkt_off[gb];
        uint32_t o = S.ent_off[e], c = S.ent_cnt[e];
        uint32_t k = e - S.cell_off[a];
        for (uint32_t j = 0, q = 0; q < c; ++j) {
            if (j > 0 && it[j] == it[j - 1]) continue;
            S.cand[o + q] = make_uint4(a, it[j], k, 0u);   // one 16-byte store per candidate
            ++q;
        }
    }
}

// Solver.updateManifolds filters + performNarrowSAT (lib.zig:272-285,
// collision.zig:299-320) for one raw candidate: bodies sa < sb (world slots),
// found at cell k of sa.  ba: bucket of each cell of sa; bb: of each of the nb
// cells of sb.  A candidate counts only at the first position of its partner
// in the lower body's query list; later occurrences are the reference's
// "already a manifold / already rejected".  Returns whether it collides, with
// the reference face id | ref-is-sa << 2 in info and the normal.
// PREFILTERED: the both-static and AABB filters were applied when the
// candidate was emitted (the filters are pure predicates of the pair, so
// their order does not change which pairs pass)
template <bool PREFILTERED = false, class B>
__device__ __forceinline__ bool narrow_pair(const WorldDev& D, const B& bodies, uint32_t sa, uint32_t sb, uint32_t k,
                                            const uint32_t* ba, const uint32_t* bb, uint32_t nb, float ca, float sa_sin,
                                            float cb, float sb_sin, uint32_t& info, float& nx, float& ny)
{
    // the pair filters of updateManifolds (lib.zig:273-282) depend on the
    // pair only, so they run first: the first-occurrence test below costs
    // (cells of a before k) x (cells of b), large for wall-sized bodies
    bool ok = true;
    if (!PREFILTERED) {
        uint32_t fa = bodies.flags[sa], fb = bodies.flags[sb];
        ok = !((fa & BF_STATIC) && (fb & BF_STATIC));
    }
    if (ok && D.nexcl) ok = !excluded(D, bodies.id[sa], bodies.id[sb]);
    if (!PREFILTERED && ok) ok = aabb_hit(bodies, sa, sb);
    // b first appears in a's query list at cell k unless an earlier cell of a
    // hashes to a bucket b also occupies (buckets are deduplicated, so
    // within one bucket b appears once)
    if (ok) {
        for (uint32_t j = 0; j < nb && ok; ++j) {
            uint32_t h = bb[j];
            for (uint32_t kk = 0; kk < k; ++kk)
                if (ba[kk] == h) { ok = false; break; }
        }
    }
    if (!ok) return false;
    Geo ga = load_geo(bodies, sa, ca, sa_sin);
    Geo gb = load_geo(bodies, sb, cb, sb_sin);
    // performNarrowSAT: lower id is the first reference body
    bool a_first = ga.id < gb.id;
    const Geo o1 = a_first ? ga : gb;   // by value: selecting references would put both in scratch
    const Geo o2 = a_first ? gb : ga;
    Sat sat;
    const float undef = -3.0316488252093987e-13f;   // Zig Debug fill of `undefined` (0xAA)
    sat.normal = mk(undef, undef);
    sat.pen = __builtin_inff();
    sat.nid = 0; sat.ref = 0; sat.set = false;
    if (overlap_sat(sat, o1, o2, 0) && overlap_sat(sat, o2, o1, 1)) {
        // ref body expressed relative to (a, b)
        bool ref_is_a = (sat.ref == 0) == a_first;
        info = (uint32_t)sat.nid | (ref_is_a ? 4u : 0u);
        nx = sat.normal.x; ny = sat.normal.y;
        return true;
    }
    return false;
}

__global__ void k_narrow(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t R = dev_count(&S.counters[C_RAW], S.cap_r);
    for (uint32_t q = grid_tid(); q < R; q += grid_stride()) {
        const uint4 cq = S.cand[q];
        const uint32_t a = cq.x, b = cq.y, k = cq.z;
        uint32_t info = 0;
        float nx = 0.f, ny = 0.f;
        const WorldDev& D = S.worlds[S.bworld[a]];
        bool hit = narrow_pair(D, D, a - D.body_base, b - D.body_base, k, S.ent_bkt + S.cell_off[a],
                               S.ent_bkt + S.cell_off[b], S.ncell[b], S.cc[a], S.cs[a], S.cc[b], S.cs[b], info, nx, ny);
        S.cand_flag[q] = hit ? 1u : 0u;
        S.sat_info[q] = hit ? info : 0u;
        S.sat_nx[q] = hit ? nx : 0.f;
        S.sat_ny[q] = hit ? ny : 0.f;
    }
}

// CollisionPoint.init (collision.zig:38-51)
__device__ __forceinline__ void write_point(const StepDev& S, uint32_t m, int p, v2 pos, float depth,
                                            const Geo& ref, const Geo& inc, v2 normal)
{
    v2 middle = sub2(pos, scale2(normal, depth / 2));
    v2 r1 = sub2(middle, ref.pos), r2 = sub2(middle, inc.pos);
    float* P = S.m_pt;
    size_t cm = S.cap_m;
    P[(MP_ACC_PN * 2 + p) * cm + m] = 0.f;
    P[(MP_ACC_PT * 2 + p) * cm + m] = 0.f;
    P[(MP_R1X * 2 + p) * cm + m] = r1.x;
    P[(MP_R1Y * 2 + p) * cm + m] = r1.y;
    P[(MP_R2X * 2 + p) * cm + m] = r2.x;
    P[(MP_R2Y * 2 + p) * cm + m] = r2.y;
    P[(MP_POSX * 2 + p) * cm + m] = middle.x;
    P[(MP_POSY * 2 + p) * cm + m] = middle.y;
    P[(MP_DEPTH * 2 + p) * cm + m] = depth;
    P[(MP_ODEPTH * 2 + p) * cm + m] = depth;
    P[(MP_MASS_N * 2 + p) * cm + m] = 0.f;
    P[(MP_MASS_T * 2 + p) * cm + m] = 0.f;
}

// clipLineToLine (collision.zig:330-363): b clipped onto a
__device__ __forceinline__ void clip_line(v2 aa, v2 ab, v2 ba, v2 bb, v2& o1, v2& o2)
{
    v2 delta_a = sub2(ab, aa);
    float a_len = length2(delta_a);
    v2 tang = scale2(delta_a, 1.0f / a_len);
    v2 p1 = ba, p2 = bb;
    {
        float sc = dot2(sub2(ba, aa), tang);
        if (!(sc > 0.f && sc < a_len)) {
            v2 clos = (sc < 0.5f * a_len) ? aa : ab;
            v2 dp = sub2(ba, bb);
            float t = -dot2(tang, sub2(bb, clos)) / dot2(tang, dp);
            p1 = addmult2(bb, dp, t);
        }
    }
    {
        float sc = dot2(sub2(bb, aa), tang);
        if (!(sc > 0.f && sc < a_len)) {
            v2 clos = (sc < 0.5f * a_len) ? aa : ab;
            v2 dp = sub2(bb, ba);
            float t = -dot2(tang, sub2(ba, clos)) / dot2(tang, dp);
            p2 = addmult2(ba, dp, t);
        }
    }
    o1 = p1; o2 = p2;
}

// identifyCollisionPoints: Disc (Disc.zig:99-123), Rectangle (Rectangle.zig:171-211),
// with clipAgainstEdge (Rectangle.zig:213-238, Disc.zig:125-130), for the
// collision of world slots sa, sb (narrow_pair's info and normal) into
// manifold m; base: the world's first batch body index
template <class B>
__device__ __forceinline__ void man_build_one(const StepDev& S, const WorldDev& D, const B& bodies, uint32_t m,
                                              uint32_t base, uint32_t sa, uint32_t sb, uint32_t info, v2 normal,
                                              float ca, float sa_sin, float cb, float sb_sin)
{
    int nid = (int)(info & 3u);
    bool ref_is_a = (info & 4u) != 0;
    uint32_t sr = ref_is_a ? sa : sb, si = ref_is_a ? sb : sa;
    Geo gr = ref_is_a ? load_geo(bodies, sr, ca, sa_sin) : load_geo(bodies, sr, cb, sb_sin);
    Geo gi = ref_is_a ? load_geo(bodies, si, cb, sb_sin) : load_geo(bodies, si, ca, sa_sin);
    S.m_ref[m] = base + sr;
    S.m_inc[m] = base + si;
    S.m_nx[m] = normal.x;
    S.m_ny[m] = normal.y;
    S.m_fric[m] = 0.f;
    S.m_prev1[m] = D.ang[sr];
    S.m_prev2[m] = D.ang[si];
    uint32_t valid = 0;
    if (!gr.rect) {
        v2 pos = closest_point(gi, gr.pos);
        Edge e = get_normal(gr, gi, nid);
        v2 n = e.dir;
        if (should_flip(n, gr, gi)) n = negself(n);
        float d = dot2(n, sub2(pos, gr.pos));
        write_point(S, m, 0, pos, d - gr.r, gr, gi, n);
        valid = 1;
    } else {
        Edge e = get_normal(gr, gi, nid);
        if (!gi.rect) {
            v2 pos = addmult2(gi.pos, e.dir, -gi.r);
            float d = dot2(sub2(pos, e.a), e.dir);
            write_point(S, m, 0, pos, d, gr, gi, e.dir);
            valid = 1;
        } else {
            // incident edge: first i minimising dot(normal, rot90ccw(normalize(v[i+1]-v[i])))
            v2 ba = gi.wv[0], bb = gi.wv[1];
            float best = __builtin_inff();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int nx = (i == 3) ? 0 : i + 1;
                v2 tn = rot90ccw(normalize2(sub2(gi.wv[nx], gi.wv[i])));
                float d = dot2(e.dir, tn);
                if (d < best) { best = d; ba = gi.wv[i]; bb = gi.wv[nx]; }
            }
            v2 ca2, cb2;
            clip_line(e.a, e.b, ba, bb, ca2, cb2);
            int i = 0;
            float d = dot2(sub2(ca2, e.a), e.dir);
            if (d < COLLISION_MARGIN) { write_point(S, m, i, ca2, d, gr, gi, e.dir); valid |= 1u << i; i += 1; }
            d = dot2(sub2(cb2, e.b), e.dir);
            if (d < COLLISION_MARGIN) { write_point(S, m, i, cb2, d, gr, gi, e.dir); valid |= 1u << i; }
        }
    }
    S.m_valid[m] = valid;
}

__global__ void k_man_build(StepDev S)
{
    uint32_t R = dev_count(&S.counters[C_RAW], S.cap_r);
    if (S.counters[C_OVERFLOW]) return;
    if (S.counters[C_MANIFOLDS] > S.cap_m) { if (grid_tid() == 0) atomicOr(&S.counters[C_OVERFLOW], 4u); return; }
    for (uint32_t q = grid_tid(); q < R; q += grid_stride()) {
        if (!S.cand_flag[q]) continue;
        const uint4 cq = S.cand[q];
        const uint32_t a = cq.x, b = cq.y;
        const WorldDev& D = S.worlds[S.bworld[a]];
        man_build_one(S, D, D, S.cand_idx[q], D.body_base, a - D.body_base, b - D.body_base, S.sat_info[q],
                      mk(S.sat_nx[q], S.sat_ny[q]), S.cc[a], S.cs[a], S.cc[b], S.cs[b]);
    }
}

__global__ void k_world_ranges(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t M = dev_count(&S.counters[C_MANIFOLDS], S.cap_m);
    uint32_t R = dev_count(&S.counters[C_RAW], S.cap_r);
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    for (uint32_t w = grid_tid(); w <= S.nworlds; w += grid_stride()) {
        uint32_t mb;
        if (w == S.nworlds) {
            mb = M;
        } else {
            const WorldDev& D = S.worlds[w];
            uint32_t fb = D.body_base;
            uint32_t e = (fb < S.nb) ? S.cell_off[fb] : E;
            if (e > E) e = E;
            uint32_t r = S.ent_off[e];
            if (r > R) r = R;
            mb = S.cand_idx[r];
            if (mb > M) mb = M;
        }
        S.w_man_begin[w] = mb;
        uint32_t jb = (w == S.nworlds) ? S.njoints : S.worlds[w].joint_base;
        S.w_item_base[w] = jb + mb;
        if (w == S.nworlds) S.counters[C_ITEMS] = jb + mb;
    }
}

// Order in which the solve kernels' workgroups take worlds: most items first
// (longest-processing-time first shortens the tail of the launch).  Worlds
// are binned by log2 of their item count; any order gives the same results.
__global__ __launch_bounds__(1024) void k_world_order(StepDev S)
{
    __shared__ uint32_t hist[33];
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t nw = S.nworlds;
    if (threadIdx.x < 33) hist[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) {
        uint32_t items = S.w_item_base[w + 1] - S.w_item_base[w];
        atomicAdd(&hist[(uint32_t)__builtin_clz(items + 1u)], 1u);   // fewer leading zeros = more items
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int b = 0; b < 33; ++b) { uint32_t c = hist[b]; hist[b] = acc; acc += c; }
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) {
        uint32_t items = S.w_item_base[w + 1] - S.w_item_base[w];
        uint32_t pos = atomicAdd(&hist[(uint32_t)__builtin_clz(items + 1u)], 1u);
        S.w_order[pos] = w;
    }
}

// distance / offset joints write both bodies even when static
// (DistanceJoint.zig:72-74, OffsetDistanceJoint.zig:100-104): such a static
// body orders every item that touches it.
__global__ void k_joint_mark(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t j = grid_tid(); j < S.njoints; j += grid_stride()) {
        // world of batch joint j
        uint32_t lo = 0, hi = S.nworlds;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (S.worlds[mid].joint_base <= j) lo = mid; else hi = mid;
        }
        const WorldDev& D = S.worlds[lo];
        const Joint& J = D.joints[j - D.joint_base];
        if (J.type <= 1) {
            if (J.s1 >= 0) atomicOr(&S.sflag[D.body_base + J.s1], SF_RESOURCE);
            if (J.s2 >= 0) atomicOr(&S.sflag[D.body_base + J.s2], SF_RESOURCE);
        }
    }
}

__global__ void k_items_joints(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    if (S.counters[C_ITEMS] > S.cap_i) { if (grid_tid() == 0) atomicOr(&S.counters[C_OVERFLOW], 8u); return; }
    for (uint32_t j = grid_tid(); j < S.njoints; j += grid_stride()) {
        uint32_t lo = 0, hi = S.nworlds;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (S.worlds[mid].joint_base <= j) lo = mid; else hi = mid;
        }
        const WorldDev& D = S.worlds[lo];
        uint32_t jl = j - D.joint_base;
        const Joint& J = D.joints[jl];
        uint32_t item = S.w_item_base[lo] + jl;
        int32_t r0 = J.s1 >= 0 ? (int32_t)(D.body_base + J.s1) : -1;
        int32_t r1 = -1;
        if (J.type <= 1 && J.s2 >= 0 && J.s2 != J.s1) r1 = (int32_t)(D.body_base + J.s2);
        S.it_r0[item] = r0;
        S.it_r1[item] = r1;
        if (r0 >= 0) atomicAdd(&S.deg[r0], 1u);
        if (r1 >= 0) atomicAdd(&S.deg[r1], 1u);
    }
}

__global__ void k_items_mans(StepDev S)
{
    if (S.counters[C_OVERFLOW] || S.counters[C_ITEMS] > S.cap_i) return;
    uint32_t M = dev_count(&S.counters[C_MANIFOLDS], S.cap_m);
    for (uint32_t m = grid_tid(); m < M; m += grid_stride()) {
        uint32_t ref = S.m_ref[m], inc = S.m_inc[m];
        uint32_t w = S.bworld[ref];
        uint32_t item = S.w_item_base[w] + S.worlds[w].njoints + (m - S.w_man_begin[w]);
        int32_t r0 = (S.sflag[ref] & SF_RESOURCE) ? (int32_t)ref : -1;
        int32_t r1 = (S.sflag[inc] & SF_RESOURCE) ? (int32_t)inc : -1;
        S.it_r0[item] = r0;
        S.it_r1[item] = r1;
        if (r0 >= 0) atomicAdd(&S.deg[r0], 1u);
        if (r1 >= 0) atomicAdd(&S.deg[r1], 1u);
    }
}

__global__ void k_list_fill(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t I = dev_count(&S.counters[C_ITEMS], S.cap_i);
    for (uint32_t it = grid_tid(); it < I; it += grid_stride()) {
        int32_t r0 = S.it_r0[it], r1 = S.it_r1[it];
        if (r0 >= 0) { uint32_t p = atomicAdd(&S.deg_fill[r0], 1u); S.lists[S.deg_off[r0] + p] = it * 2u; }
        if (r1 >= 0) { uint32_t p = atomicAdd(&S.deg_fill[r1], 1u); S.lists[S.deg_off[r1] + p] = it * 2u + 1u; }
    }
}

__device__ __forceinline__ uint32_t ld_acq(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st_rel(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// polls of one item before its world is stopped as stalled: ~2 s of waiting
// with the poll's sleep; a wave can wait behind other worlds' waves on its
// SIMD for a tenth of that and still be making progress
constexpr uint32_t SPIN_LIMIT = 1u << 26;
// the colouring polls global memory (~1 us a poll): its own, smaller limit
constexpr uint32_t COLOR_SPIN_LIMIT = 1u << 22;

// Coloured schedule (a world with r2d_set_schedule(w, 1)): greedy colouring
// of the world's items in item order -- an item takes the smallest colour
// unused by earlier items sharing one of its ordering bodies; past 127
// colours an item takes 128 + its index, a colour of its own -- then the
// execution order (colour, item).  One workgroup per world.  The colouring
// runs as a dataflow over per-body counters in item order (an item waits
// until every earlier item on its bodies has taken its colour), the pattern
// of the solver; the order is a stable block-wide counting sort.  Worlds in
// the reference order get colour 0 and the identity order.
// Reference order: the execution positions of a world's items (it_perm) by
// level -- an item's level is 1 + the larger level of its predecessors, the
// items just before it in its ordering bodies' lists.  Items of one level
// share no ordering body, so their order among themselves is free.  The solve's lanes
// take positions t, t + T, ... of each pass in turn; with positions in item
// order a lane's next item often waits behind a long chain while items at
// later positions could run, and the lane holds them back.  Any order that
// lists every item after its predecessors keeps the dataflow free of
// deadlock (the earliest unfinished item in (pass, position) order is always
// its lane's current item, and its predecessors have run), and the counters
// still give each body its items in the reference order, so the results are
// the same bits; only which lane runs what, and when, changes.  Work arrays
// in LDS (lds_words): levels, packed predecessors, level counts, 3 nit + 2
// words; a world with more items keeps the item order.  Levels are found by
// sweeps in which each thread walks a contiguous chunk of items in item
// order (a predecessor in the same chunk is seen in the same sweep), until a
// sweep changes nothing.  The world's item arrays start at its first item
// (r0s .. perm); a body r's item list is lists[off[r - base] ..], whose
// entries are (item + sub) * 2 + side.  Not inlined: it runs once per step,
// and kept out of the caller it leaves the caller's register allocation as
// it is (the fused step kernel is at its register budget).
__device__ __attribute__((noinline)) void level_order(const int32_t* r0s, const int32_t* r1s, const uint32_t* k0s,
                                                      const uint32_t* k1s, uint32_t* perm, const uint32_t* lists,
                                                      const uint32_t* off, uint32_t base, uint32_t sub, uint32_t nit,
                                                      uint32_t* lds, uint32_t lds_words, uint32_t* flags)
{
    const uint32_t T = blockDim.x, t = threadIdx.x;
    if (nit < 2u || nit >= 0xffffu || 3u * nit + 2u > lds_words) {
        for (uint32_t p = t; p < nit; p += T) perm[p] = p;
        return;
    }
    uint32_t* lev = lds;
    uint32_t* pred = lds + nit;
    uint32_t* hist = lds + 2u * nit;
    for (uint32_t p = t; p < nit; p += T) {
        const int32_t r0 = r0s[p], r1 = r1s[p];
        uint32_t q0 = 0xffffu, q1 = 0xffffu;
        if (r0 >= 0 && k0s[p] > 0u) q0 = (lists[off[(uint32_t)r0 - base] + k0s[p] - 1u] >> 1) - sub;
        if (r1 >= 0 && k1s[p] > 0u) q1 = (lists[off[(uint32_t)r1 - base] + k1s[p] - 1u] >> 1) - sub;
        pred[p] = q0 | (q1 << 16);
        lev[p] = 1u;
    }
    if (t == 0) { flags[0] = 0u; flags[1] = 0u; flags[2] = 0u; }
    __syncthreads();
    const uint32_t chunk = (nit + T - 1) / T;
    const uint32_t lo = umin(t * chunk, nit), hi = umin(lo + chunk, nit);
    // sweep p reports a change in flags[p % 3] and clears the flag of sweep
    // p + 1, which every thread read before the previous barrier
    for (uint32_t sweep = 0;; ++sweep) {
        bool changed = false;
        for (uint32_t p = lo; p < hi; ++p) {
            const uint32_t pw = pred[p];
            const uint32_t q0 = pw & 0xffffu, q1 = pw >> 16;
            uint32_t l = 1u;
            if (q0 != 0xffffu)
                l = umax(l, 1u + __hip_atomic_load(&lev[q0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (q1 != 0xffffu)
                l = umax(l, 1u + __hip_atomic_load(&lev[q1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (l != __hip_atomic_load(&lev[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                __hip_atomic_store(&lev[p], l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                changed = true;
            }
        }
        if (changed) flags[sweep % 3u] = 1u;
        if (t == 0) flags[(sweep + 1u) % 3u] = 0u;
        __syncthreads();
        if (!flags[sweep % 3u]) break;
    }
    // counting sort by level: counts, exclusive offsets, then each item takes
    // the next free position of its level
    if (t == 0) flags[0] = 0u;
    __syncthreads();
    for (uint32_t p = lo; p < hi; ++p) atomicMax(&flags[0], lev[p]);
    __syncthreads();
    const uint32_t top = flags[0];
    for (uint32_t l = t; l <= top + 1u; l += T) hist[l] = 0u;
    __syncthreads();
    for (uint32_t p = t; p < nit; p += T) atomicAdd(&hist[lev[p]], 1u);
    __syncthreads();
    if (t < 64u) {
        const uint32_t len = top + 1u, c = (len + 63u) / 64u;
        const uint32_t a = umin(t * c, len), b = umin(a + c, len);
        uint32_t sum = 0u;
        for (uint32_t l = a; l < b; ++l) sum += hist[l];
        uint32_t run = wave_incl_scan(sum) - sum;
        for (uint32_t l = a; l < b; ++l) {
            const uint32_t x = hist[l];
            hist[l] = run;
            run += x;
        }
    }
    __syncthreads();
    for (uint32_t p = t; p < nit; p += T) perm[atomicAdd(&hist[lev[p]], 1u)] = p;
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_color(StepDev S, uint32_t lds_words)
{
    extern __shared__ __align__(16) uint32_t dyn[];
    __shared__ uint32_t s_scan[256];
    __shared__ uint32_t s_maxc, s_abort, s_base;
    __shared__ uint32_t s_flags[3];
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = blockIdx.x;
    const WorldDev& D = S.worlds[w];
    const uint32_t i0 = S.w_item_base[w], nit = S.w_item_base[w + 1] - i0;
    const uint32_t T = blockDim.x, t = threadIdx.x;
    if (D.sched == 0) {
        for (uint32_t p = t; p < nit; p += T) S.it_color[i0 + p] = 0u;
        level_order(S.it_r0 + i0, S.it_r1 + i0, S.it_k0 + i0, S.it_k1 + i0, S.it_perm + i0, S.lists, S.deg_off, 0u,
                    i0, nit, dyn, lds_words, s_flags);
        return;
    }
    const uint32_t base = D.body_base;
    for (int b = t; b < D.nbodies; b += T) {
        S.cmask[2 * (base + b)] = 0ull;
        S.cmask[2 * (base + b) + 1] = 0ull;
        S.g_cnt[base + b] = 0u;
    }
    if (t == 0) { s_maxc = 0u; s_abort = 0u; s_base = 0u; }
    __syncthreads();
    uint32_t spins = 0;
    for (uint32_t p = t; p < nit;) {
        if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
        const uint32_t it = i0 + p;
        const int32_t r0 = S.it_r0[it], r1 = S.it_r1[it];
        bool ready = r0 < 0 || ld_acq(&S.g_cnt[r0]) == S.it_k0[it];
        if (ready && r1 >= 0) ready = ld_acq(&S.g_cnt[r1]) == S.it_k1[it];
        if (ready) {
            uint64_t m0 = 0, m1 = 0;
            if (r0 >= 0) { m0 |= S.cmask[2 * r0]; m1 |= S.cmask[2 * r0 + 1]; }
            if (r1 >= 0) { m0 |= S.cmask[2 * r1]; m1 |= S.cmask[2 * r1 + 1]; }
            uint32_t c;
            if (~m0) c = (uint32_t)__builtin_ctzll(~m0);
            else if (~m1) c = 64u + (uint32_t)__builtin_ctzll(~m1);
            else c = 128u + p;
            S.it_color[it] = c;
            if (c < 128) {
                const uint32_t hi = c >> 6;
                const uint64_t bit = 1ull << (c & 63u);
                if (r0 >= 0) S.cmask[2 * r0 + hi] |= bit;
                if (r1 >= 0) S.cmask[2 * r1 + hi] |= bit;
            }
            if (r0 >= 0) st_rel(&S.g_cnt[r0], S.it_k0[it] + 1u);
            if (r1 >= 0) st_rel(&S.g_cnt[r1], S.it_k1[it] + 1u);
            atomicMax(&s_maxc, umin(c, 128u));
            p += T;
            spins = 0;
        } else {
            // every item's predecessors come first, so some item is always ready
            if (++spins > COLOR_SPIN_LIMIT) {
                __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                S.w_status[w] = 5;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    if (s_abort) return;
    // execution order: stable counting sort by min(colour, 128); each thread
    // owns a contiguous chunk of items, bins are placed one after another
    const uint32_t chunk = (nit + T - 1) / T;
    const uint32_t lo = umin(t * chunk, nit), hi = umin(lo + chunk, nit);
    const uint32_t maxc = s_maxc;
    for (uint32_t bin = 0; bin <= maxc; ++bin) {
        uint32_t cnt = 0;
        for (uint32_t p = lo; p < hi; ++p) cnt += umin(S.it_color[i0 + p], 128u) == bin;
        s_scan[t] = cnt;
        __syncthreads();
        for (uint32_t d = 1; d < T; d <<= 1) {
            uint32_t v = t >= d ? s_scan[t - d] : 0u;
            __syncthreads();
            s_scan[t] += v;
            __syncthreads();
        }
        uint32_t off = s_base + s_scan[t] - cnt;
        for (uint32_t p = lo; p < hi; ++p)
            if (umin(S.it_color[i0 + p], 128u) == bin) S.it_perm[i0 + off++] = p;
        __syncthreads();
        if (t == T - 1) s_base += s_scan[T - 1];
        __syncthreads();
    }
}

// rank of each item among the items sharing one of its resources: in item
// order (the reference's sequential order: constraints, then manifolds in
// discovery order); with BY_COLOR, for bodies of coloured worlds, in
// (colour, item) order
template <bool BY_COLOR>
__global__ void k_list_rank(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t x = grid_tid(); x < S.nb; x += grid_stride()) {
        if (BY_COLOR && S.worlds[S.bworld[x]].sched == 0) continue;
        uint32_t o = S.deg_off[x], n = S.deg_off[x + 1] - o;
        uint32_t* l = S.lists + o;
        auto key = [&](uint32_t v) -> uint64_t {
            return BY_COLOR ? (((uint64_t)S.it_color[v >> 1] << 32) | v) : (uint64_t)v;
        };
        for (uint32_t i = 1; i < n; ++i) {
            uint32_t v = l[i];
            uint64_t kv = key(v);
            uint32_t j = i;
            while (j > 0 && key(l[j - 1]) > kv) { l[j] = l[j - 1]; --j; }
            l[j] = v;
        }
        for (uint32_t p = 0; p < n; ++p) {
            uint32_t v = l[p];
            if (v & 1u) S.it_k1[v >> 1] = p; else S.it_k0[v >> 1] = p;
        }
    }
}

// ---------------------------------------------------------------- solver
template <bool LDS>
struct SolverState {
    float *mx, *my, *L, *im, *ii;
    uint32_t *cnt, *deg;
};



// ------------------------------------------------------- solver item records
// Each solver item of a step gets one 128-byte record at its execution
// position, built in parallel at the first sub-step: the ordering bodies and
// ranks, and for a manifold its normal, friction, both points' lever arms,
// depths, effective masses (identical every sub-step unless updateTGSDepth
// moves the point) and accumulated impulses.  A worker holds a record before
// the item's dependencies are met (item_chain reads its next record while it
// still waits on the current one; item_chain_lanes copies it into the lane's
// LDS slot), so a link of a dependency chain touches only LDS body state.
enum : uint32_t {
    RF_JOINT = 1u, RF_P0 = 2u, RF_P1 = 4u, RF_W1 = 8u, RF_W2 = 16u, RF_S1 = 32u, RF_S2 = 64u,
    RF_TGS = 128u,           // a point's depth is not its original depth: updateTGSDepth is live
    RF_WARM = 256u           // warm started: each sub-step adds MP_WARM_N / MP_WARM_T to the impulses
};
constexpr uint32_t NONE = 0xffffffffu;

// words: 0 flags, 1 body a (manifold ref / joint r0), 2 body b (inc / r1),
// 3 index (manifold in the step arrays / world-local joint), 4-5 ranks at the
// ordering bodies, 6 the ordering bodies' item counts, 7 the compact chain
// indices (worlds beyond LDS); then the six rows a chain link reads: 8-10
// normal + friction, 11-17 point 0 (r1x r1y r2x r2y depth mass_n mass_t),
// 18-24 point 1, 25-26 the points' Baumgarte biases, 28-31 impulses (pn0 pt0
// pn1 pt1)
struct Rec {
    uint32_t flags, a, b, idx, rk0, rk1;
    float nx, ny, fric;
    float p[2][7];
    float acc[4];
    uint32_t degs;       // items at r0 | items at r1 << 16; DEGS_WIDE when either needs more than 16 bits
    float bias[2];       // Baumgarte bias of the sub-step per point, -1 when depth >= 0
    uint32_t cpair;      // worlds beyond LDS: compact chain indices of (a, b) / the joint's (s1, s2)
};
constexpr uint32_t DEGS_WIDE = 0xffffffffu;

constexpr uint32_t REC_LINK = 2;    // first of the six rows a chain link reads
constexpr uint32_t REC_ACC = 7;     // the impulses' row

__device__ __forceinline__ void rec_load(Rec& r, const uint4* q)
{
    const uint4 v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3], v4 = q[4], v5 = q[5], v6 = q[6], v7 = q[7];
    r.flags = v0.x; r.a = v0.y; r.b = v0.z; r.idx = v0.w;
    r.rk0 = v1.x; r.rk1 = v1.y; r.degs = v1.z; r.cpair = v1.w;
    r.nx = __uint_as_float(v2.x); r.ny = __uint_as_float(v2.y); r.fric = __uint_as_float(v2.z);
    r.p[0][0] = __uint_as_float(v2.w);
    r.p[0][1] = __uint_as_float(v3.x); r.p[0][2] = __uint_as_float(v3.y); r.p[0][3] = __uint_as_float(v3.z);
    r.p[0][4] = __uint_as_float(v3.w);
    r.p[0][5] = __uint_as_float(v4.x); r.p[0][6] = __uint_as_float(v4.y); r.p[1][0] = __uint_as_float(v4.z);
    r.p[1][1] = __uint_as_float(v4.w);
    r.p[1][2] = __uint_as_float(v5.x); r.p[1][3] = __uint_as_float(v5.y); r.p[1][4] = __uint_as_float(v5.z);
    r.p[1][5] = __uint_as_float(v5.w);
    r.p[1][6] = __uint_as_float(v6.x); r.bias[0] = __uint_as_float(v6.y); r.bias[1] = __uint_as_float(v6.z);
    r.acc[0] = __uint_as_float(v7.x); r.acc[1] = __uint_as_float(v7.y);
    r.acc[2] = __uint_as_float(v7.z); r.acc[3] = __uint_as_float(v7.w);
}

// a record read by the compact chain: its bodies as compact indices
__device__ __forceinline__ void rec_to_compact(Rec& r)
{
    const uint32_t c0 = r.cpair & 0xffffu, c1 = r.cpair >> 16;
    r.a = c0;
    if (r.flags & RF_JOINT) r.b = (r.b == NONE) ? NONE : c1;
    else r.b = c1;
}

__device__ __forceinline__ void rec_store(const Rec& r, uint4* q)
{
    q[0] = make_uint4(r.flags, r.a, r.b, r.idx);
    q[1] = make_uint4(r.rk0, r.rk1, r.degs, r.cpair);
    q[2] = make_uint4(__float_as_uint(r.nx), __float_as_uint(r.ny), __float_as_uint(r.fric),
                      __float_as_uint(r.p[0][0]));
    q[3] = make_uint4(__float_as_uint(r.p[0][1]), __float_as_uint(r.p[0][2]), __float_as_uint(r.p[0][3]),
                      __float_as_uint(r.p[0][4]));
    q[4] = make_uint4(__float_as_uint(r.p[0][5]), __float_as_uint(r.p[0][6]), __float_as_uint(r.p[1][0]),
                      __float_as_uint(r.p[1][1]));
    q[5] = make_uint4(__float_as_uint(r.p[1][2]), __float_as_uint(r.p[1][3]), __float_as_uint(r.p[1][4]),
                      __float_as_uint(r.p[1][5]));
    q[6] = make_uint4(__float_as_uint(r.p[1][6]), __float_as_uint(r.bias[0]), __float_as_uint(r.bias[1]), 0u);
    q[7] = make_uint4(__float_as_uint(r.acc[0]), __float_as_uint(r.acc[1]), __float_as_uint(r.acc[2]),
                      __float_as_uint(r.acc[3]));
}

__device__ __forceinline__ int32_t rec_r0(const Rec& r)
{
    if (r.flags & RF_JOINT) return r.a == NONE ? -1 : (int32_t)r.a;
    return (r.flags & RF_W1) ? (int32_t)r.a : -1;
}
__device__ __forceinline__ int32_t rec_r1(const Rec& r)
{
    if (r.flags & RF_JOINT) return r.b == NONE ? -1 : (int32_t)r.b;
    return (r.flags & RF_W2) ? (int32_t)r.b : -1;
}

// the ordering bodies' item counts, packed when both fit 16 bits
template <class ST>
__device__ __forceinline__ uint32_t rec_degs(const Rec& r, const ST& st)
{
    int32_t q0 = rec_r0(r), q1 = rec_r1(r);
    uint32_t d0 = q0 >= 0 ? st.deg[q0] : 0u, d1 = q1 >= 0 ? st.deg[q1] : 0u;
    return (d0 > 0xffffu || d1 > 0xffffu) ? DEGS_WIDE : (d0 | (d1 << 16));
}

// Baumgarte bias of each point for a sub-step of length dt (collision.zig:172),
// -1 for a point whose depth is >= 0 (calculateImpulses resets and skips it)
__device__ __forceinline__ void rec_biases(Rec& r, float dt)
{
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        float depth = r.p[p][4];
        r.bias[p] = depth >= 0.f ? -1.0f : BAUMGARTE * maxn(0.0f, (-depth) - BAUMGARTE_SLOP) / dt;
    }
}

// CollisionManifold.preStep effective masses (collision.zig:117-132) for one point
template <class ST>
__device__ __forceinline__ void rec_masses(Rec& r, const ST& st, int p)
{
    float im1 = st.im[r.a], im2 = st.im[r.b], ii1 = st.ii[r.a], ii2 = st.ii[r.b];
    float inv_mass = im1 + im2;
    v2 n = mk(r.nx, r.ny), t = rot90cw(n);
    v2 r1 = mk(r.p[p][0], r.p[p][1]), r2 = mk(r.p[p][2], r.p[p][3]);
    float r1n = cross2(r1, n), r2n = cross2(r2, n);
    float kn = inv_mass + ii1 * (r1n * r1n) + ii2 * (r2n * r2n);
    float r1t = cross2(r1, t), r2t = cross2(r2, t);
    float kt = inv_mass + ii1 * (r1t * r1t) + ii2 * (r2t * r2t);
    r.p[p][5] = kn > 0.0f ? 1.0f / kn : 0.0f;
    r.p[p][6] = kt > 0.0f ? 1.0f / kt : 0.0f;
}

// record of the item at execution position `pos` (first sub-step)
template <bool WARM = false, class ST>
__device__ __forceinline__ void rec_build(Rec& r, const StepDev& S, const WorldDev& D, const ST& st, uint32_t pos,
                                          uint32_t ibase, uint32_t nj, uint32_t mbeg, uint32_t base)
{
    const size_t cm = S.cap_m;
    const float* P = S.m_pt;
    uint32_t j = S.it_perm[ibase + pos];
    uint32_t item = ibase + j;
    r.rk0 = S.it_k0[item];
    r.rk1 = S.it_k1[item];
    r.nx = r.ny = r.fric = 0.0f;
    r.degs = 0u;
    r.bias[0] = r.bias[1] = -1.0f;
    r.cpair = 0u;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int f = 0; f < 7; ++f) r.p[p][f] = 0.0f;
#pragma unroll
    for (int f = 0; f < 4; ++f) r.acc[f] = 0.0f;
    if (j < nj) {
        int32_t r0 = S.it_r0[item], r1 = S.it_r1[item];
        r.flags = RF_JOINT;
        r.a = r0 >= 0 ? (uint32_t)(r0 - (int32_t)base) : NONE;
        r.b = r1 >= 0 ? (uint32_t)(r1 - (int32_t)base) : NONE;
        r.idx = j;
        return;
    }
    uint32_t m = mbeg + (j - nj);
    r.idx = m;
    r.a = S.m_ref[m] - base;
    r.b = S.m_inc[m] - base;
    r.flags = 0u;
    if (S.sflag[base + r.a] & SF_RESOURCE) r.flags |= RF_W1;
    if (S.sflag[base + r.b] & SF_RESOURCE) r.flags |= RF_W2;
    if (D.flags[r.a] & BF_STATIC) r.flags |= RF_S1;
    if (D.flags[r.b] & BF_STATIC) r.flags |= RF_S2;
    r.nx = S.m_nx[m];
    r.ny = S.m_ny[m];
    r.fric = sqrt_rn(D.mu[r.a] * D.mu[r.b]);   // preStep (collision.zig:114)
    uint32_t valid = S.m_valid[m];
    if constexpr (WARM)
        if (valid & MV_WARM) r.flags |= RF_WARM;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(valid & (1u << p))) continue;
        r.flags |= (RF_P0 << p);
        r.p[p][0] = P[(MP_R1X * 2 + p) * cm + m];
        r.p[p][1] = P[(MP_R1Y * 2 + p) * cm + m];
        r.p[p][2] = P[(MP_R2X * 2 + p) * cm + m];
        r.p[p][3] = P[(MP_R2Y * 2 + p) * cm + m];
        r.p[p][4] = P[(MP_DEPTH * 2 + p) * cm + m];
        if (!approx(r.p[p][4], P[(MP_ODEPTH * 2 + p) * cm + m], 1e-4f)) r.flags |= RF_TGS;
        rec_masses(r, st, p);
    }
}

// updateTGSDepth (collision.zig:73-100) for a record whose points may move:
// rotate lever arms by the bodies' rotation since the previous sub-step,
// recompute depth and (preStep) the effective masses
// DV: WorldDev or PassArrays (the body arrays it reads: ang, px, py)
template <class ST, class DV>
__device__ __forceinline__ void rec_tgs(Rec& r, const StepDev& S, const DV& D, const ST& st, bool = false)
{
    const uint32_t m = r.idx;
    const size_t cm = S.cap_m;
    const float* P = S.m_pt;
    // everything the update may read, loaded before any of it is tested (one
    // round trip instead of three dependent ones; the indices are valid
    // whether or not a point is present or moved)
    const float a1 = D.ang[r.a], a2 = D.ang[r.b];
    const float prev1 = S.m_prev1[m], prev2 = S.m_prev2[m];
    const float ods[2] = {P[(MP_ODEPTH * 2 + 0) * cm + m], P[(MP_ODEPTH * 2 + 1) * cm + m]};
    const v2 pa = mk(D.px[r.a], D.py[r.a]), pb = mk(D.px[r.b], D.py[r.b]);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(r.flags & (RF_P0 << p))) continue;
        const float od = ods[p];
        if (approx(r.p[p][4], od, 1e-4f)) continue;
        v2 rr1 = rotate2(mk(r.p[p][0], r.p[p][1]), a1 - prev1);
        v2 rr2 = rotate2(mk(r.p[p][2], r.p[p][3]), a2 - prev2);
        v2 q1 = add2(rr1, pa);
        v2 q2 = add2(rr2, pb);
        float dd = dot2(mk(r.nx, r.ny), sub2(q2, q1));
        r.p[p][4] = dd + od;
        r.p[p][0] = rr1.x; r.p[p][1] = rr1.y; r.p[p][2] = rr2.x; r.p[p][3] = rr2.y;
        rec_masses(r, st, p);
    }
    S.m_prev1[m] = a1;
    S.m_prev2[m] = a2;
}

// What a joint link reads besides the bodies' momentum and inverse masses:
// parameters and the pose of its bodies, constant within a sub-step.
struct JointIn {
    uint32_t type;
    bool stat1;                       // body 1 is static (fixed / motor joints skip)
    float pmax, pmin, beta, target;
    float r1x, r1y, r2x, r2y, tx, ty;
    v2 pa, pb;                        // positions
    float ca, sa, cb, sb;             // cos / sin of the angles (offset joints)
    float I1, tq1;                    // inertia and torque of body 1 (motor joints)
};

// Constraint.solve (Constraints/*.zig) with momentum / inverse masses at st
// indices a and b
template <class ST>
__device__ __forceinline__ void joint_apply(const JointIn& j, const ST& st, uint32_t a, uint32_t b, float dt)
{
    const float pmax = j.pmax, pmin = j.pmin, beta = j.beta, target = j.target;
    if (j.type == 0) {
        // DistanceJoint.solve (DistanceJoint.zig:40-75)
        float w1 = st.im[a], w2 = st.im[b];
        v2 v1 = scale2(mk(st.mx[a], st.my[a]), w1);
        v2 v2v = scale2(mk(st.mx[b], st.my[b]), w2);
        v2 rv = sub2(v2v, v1);
        v2 pa = j.pa, pb = j.pb;
        v2 n = normalize2(sub2(pb, pa));
        float dist_err = length2(sub2(pb, pa)) - target;
        float pc = beta * dist_err / maxn(target, 0.001f);
        float Jv = -(dot2(rv, n) + pc) / (w1 + w2);
        float ndv = dot2(n, rv);
        float den = 1.0f / maxn(absf(ndv), GRAD_DIV_LIMIT);
        Jv = clampn(Jv, pmin * den, pmax * den);
        v2 dp = scale2(n, Jv);
        st.mx[a] -= dp.x; st.my[a] -= dp.y;
        st.mx[b] += dp.x; st.my[b] += dp.y;
    } else if (j.type == 1) {
        // OffsetDistanceJoint.solve (OffsetDistanceJoint.zig:44-105)
        v2 r1 = mk(j.r1x, j.r1y), r2 = mk(j.r2x, j.r2y);
        v2 rr1 = rotcs(r1, j.ca, j.sa), rr2 = rotcs(r2, j.cb, j.sb);
        v2 a1 = add2(rr1, j.pa);
        v2 a2 = add2(rr2, j.pb);
        v2 n = normalize2(sub2(a2, a1));
        float im1 = st.im[a], im2 = st.im[b], ii1 = st.ii[a], ii2 = st.ii[b];
        v2 vl1 = scale2(mk(st.mx[a], st.my[a]), im1);
        float w1 = st.L[a] * ii1;
        v2 vl2 = scale2(mk(st.mx[b], st.my[b]), im2);
        float w2 = st.L[b] * ii2;
        v2 v1 = add2(vl1, mk(-rr1.y * w1, rr1.x * w1));
        v2 v2v = add2(vl2, mk(-rr2.y * w2, rr2.x * w2));
        v2 dv = sub2(v2v, v1);
        float dist_err = length2(sub2(a2, a1)) - target;
        float pc = beta * dist_err / maxn(target, GRAD_DIV_LIMIT);
        float num = dot2(n, dv) + pc;
        float r1xn = cross2(rr1, n), r2xn = cross2(rr2, n);
        float den = im1 + im2 + r1xn * r1xn * ii1 + r2xn * r2xn * ii2;
        float Jv = -num / den;
        float ndv = dot2(n, dv);
        float den2 = 1.0f / maxn(absf(ndv), GRAD_DIV_LIMIT);
        Jv = clampn(Jv, pmin * den2, pmax * den2);
        v2 dp = scale2(n, Jv);
        st.mx[a] -= dp.x; st.my[a] -= dp.y;
        st.L[a] -= Jv * r1xn;
        st.mx[b] += dp.x; st.my[b] += dp.y;
        st.L[b] += Jv * r2xn;
    } else if (j.type == 2) {
        // FixedPositionJoint.solve (FixedPositionJoint.zig:38-72)
        if (j.stat1) return;
        float w = st.im[a];
        v2 v = scale2(mk(st.mx[a], st.my[a]), w);
        v2 delta = sub2(mk(j.tx, j.ty), j.pa);
        float dist = length2(delta);
        if (dist < ALLOWED_CONSTRAINT_VALUE) return;
        v2 n = scale2(delta, 1.0f / dist);
        float bias = beta * dist;
        float rv = dot2(v, n);
        float Jv = -(rv - bias) / w;
        float den = 1.0f / maxn(absf(rv), ALLOWED_CONSTRAINT_VALUE);
        Jv = clampn(Jv, pmin * den, pmax * den);
        v2 dp = scale2(n, Jv);
        st.mx[a] += dp.x; st.my[a] += dp.y;
    } else {
        // MotorJoint.solve (MotorJoint.zig:38-64)
        if (j.stat1) return;
        float I = j.I1;
        float omega = st.L[a] / I;
        float C = omega - target;
        if (absf(C) < ALLOWED_CONSTRAINT_VALUE) return;
        float bias = beta * C;
        float tq = j.tq1;
        float Jv = tq - bias;
        float rel = tq / I;
        float den = 1.0f / maxn(absf(rel), ALLOWED_CONSTRAINT_VALUE);
        Jv = clampn(Jv, pmin * den, pmax * den);
        st.L[a] += Jv * dt;
    }
}

// Joint records keep what the link reads in the fields a contact uses for its
// points: p[0] = pmax pmin beta r1x r1y r2x r2y, p[1][0..2] = target tx ty,
// p[1][3..6] = positions of the two bodies, (nx, ny, fric, acc[0]) = cos / sin
// of both angles, acc[1] / acc[2] = inertia / torque of body 1; the type in
// flags bits 24-30 and RF_S1 when body 1 is static.
__device__ __forceinline__ void rec_joint_params(Rec& r, const Joint& J, const WorldDev& D)
{
    r.flags |= (uint32_t)J.type << 24;
    if (D.flags[J.s1] & BF_STATIC) r.flags |= RF_S1;
    r.p[0][0] = J.pmax; r.p[0][1] = J.pmin; r.p[0][2] = J.beta;
    r.p[0][3] = J.r1x; r.p[0][4] = J.r1y; r.p[0][5] = J.r2x; r.p[0][6] = J.r2y;
    r.p[1][0] = J.target; r.p[1][1] = J.tx; r.p[1][2] = J.ty;
}

// the sub-step's pose of a joint record's bodies (after the previous
// sub-step's position integration; torque as set for this sub-step);
// DV: WorldDev or PassArrays
template <class DV>
__device__ __forceinline__ void rec_joint_pose(Rec& r, const Joint& J, const DV& D)
{
    const int32_t a = J.s1, b = J.s2;
    r.p[1][3] = D.px[a]; r.p[1][4] = D.py[a];
    r.acc[1] = D.inertia[a];
    r.acc[2] = D.tq[a];
    if (J.type == 1) {
        float angle_a = D.ang[a], angle_b = D.ang[b];
        r.nx = cosf_r(angle_a); r.ny = sinf_r(angle_a);
        r.fric = cosf_r(angle_b); r.acc[0] = sinf_r(angle_b);
    }
    if (J.type <= 1) { r.p[1][5] = D.px[b]; r.p[1][6] = D.py[b]; }
}

template <class ST>
__device__ __forceinline__ void rec_solve_joint(const Rec& r, const ST& st, uint32_t ua, uint32_t ub, float dt)
{
    JointIn j;
    j.type = (r.flags >> 24) & 0x7fu;
    j.stat1 = (r.flags & RF_S1) != 0;
    j.pmax = r.p[0][0]; j.pmin = r.p[0][1]; j.beta = r.p[0][2];
    j.r1x = r.p[0][3]; j.r1y = r.p[0][4]; j.r2x = r.p[0][5]; j.r2y = r.p[0][6];
    j.target = r.p[1][0]; j.tx = r.p[1][1]; j.ty = r.p[1][2];
    j.pa = mk(r.p[1][3], r.p[1][4]); j.pb = mk(r.p[1][5], r.p[1][6]);
    j.ca = r.nx; j.sa = r.ny; j.cb = r.fric; j.sb = r.acc[0];
    j.I1 = r.acc[1]; j.tq1 = r.acc[2];
    joint_apply(j, st, ua, ub, dt);
}

// one body's solver state as read by a chain link
struct BodyState {
    float mx, my, L, im, ii;
};

__device__ __forceinline__ float lds_relaxed(const float* p)
{
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP));
}

// a body's state read without ordering: only for LDS, where one workgroup's
// operations are served in order (see the dataflow loops)
template <class ST>
__device__ __forceinline__ BodyState body_state(const ST& st, uint32_t i)
{
    return BodyState{lds_relaxed(&st.mx[i]), lds_relaxed(&st.my[i]), lds_relaxed(&st.L[i]), lds_relaxed(&st.im[i]),
                     lds_relaxed(&st.ii[i])};
}

// The light compact chain's state (worlds beyond LDS whose touched bodies
// exceed the full compact state but whose momentum and counters fit LDS):
// momentum and counter of each touched body in LDS under its compact index;
// a link's inverse masses and inertias come with its item (a sixth row of the
// lane slot, from S.rmass), and the pair proxies answer st.im[i] / st.ii[i]
// for the item's two bodies (the first body's index, else the second's); an
// item count per body (only items whose counts need more than 16 bits read
// it) is looked up in global memory.
struct PairValue {
    uint32_t a;
    float va, vb;
    __device__ __forceinline__ float operator[](uint32_t i) const { return i == a ? va : vb; }
};
struct DegLookup {
    const uint32_t* deg;
    const uint32_t* cinv;
    uint32_t base;
    __device__ __forceinline__ uint32_t operator[](uint32_t c) const { return deg[base + cinv[base + c]]; }
};
struct LiteState {
    float *mx, *my, *L;
    uint32_t* cnt;
    PairValue im, ii;
    DegLookup deg;
    const uint4* rmass;      // the world's item masses, by execution position
};
template <class ST>
struct IsLite {
    static constexpr bool value = false;
};
template <>
struct IsLite<LiteState> {
    static constexpr bool value = true;
};

__device__ __forceinline__ BodyState body_state(const LiteState& st, uint32_t i)
{
    return BodyState{lds_relaxed(&st.mx[i]), lds_relaxed(&st.my[i]), lds_relaxed(&st.L[i]), st.im[i], st.ii[i]};
}

// the same with ordinary loads (after the item's counters were acquired)
template <class ST>
__device__ __forceinline__ BodyState body_state_ordered(const ST& st, uint32_t i)
{
    return BodyState{st.mx[i], st.my[i], st.L[i], st.im[i], st.ii[i]};
}

// CollisionManifold.calculateImpulses (collision.zig:135-218) on a slot or a
// record, with both bodies' state already read (the link's only reads of
// them) and the Baumgarte bias of the sub-step precomputed
// ((0.02 max(0, -depth - 0.005)) / dt, collision.zig:172; -1 marks depth >= 0)
template <bool WARM = false, class Item, class ST>
__device__ __forceinline__ void contact_impulses(Item& z, uint32_t l1, uint32_t l2, const BodyState& b1,
                                                 const BodyState& b2, const ST& st)
{
    v2 vl1 = scale2(mk(b1.mx, b1.my), b1.im);
    float om1 = b1.L * b1.ii;
    v2 vl2 = scale2(mk(b2.mx, b2.my), b2.im);
    float om2 = b2.L * b2.ii;
    v2 n = mk(z.nx, z.ny);
    v2 t = rot90cw(n);
    v2 app1 = mk(0.f, 0.f), app2 = mk(0.f, 0.f);
    float rot1 = 0.f, rot2 = 0.f;
    // straight-line code with selects: a point that is absent, separated
    // (bias < 0: its impulses reset) or below the minimum impulse changes
    // nothing else, exactly as skipping it would.  A warm-started manifold
    // has no minimum: its accumulated impulse may shrink (clamped at 0), so a
    // carried impulse that is too large can be taken back.
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        // the second point only when a lane running this link has one (a
        // wave whose links are all one-point manifolds skips its half of the
        // instructions; for an absent point the selects below change nothing)
        if (p == 1 && __ballot((z.flags & RF_P1) != 0u) == 0ull) break;
        const bool on = (z.flags & (RF_P0 << p)) != 0u;
        const bool sep = z.bias[p] < 0.f;
        v2 r1 = mk(z.p[p][0], z.p[p][1]), r2 = mk(z.p[p][2], z.p[p][3]);
        v2 v1 = add2(vl1, mk(-r1.y * om1, r1.x * om1));
        v2 v2v = add2(vl2, mk(-r2.y * om2, r2.x * om2));
        v2 dv = sub2(v1, v2v);
        float num = dot2(dv, n) + z.bias[p];
        float pn = num * z.p[p][5];
        const bool go = on && !sep && (!(pn < MIN_MANIFOLD_IMPULSE) || (WARM && (z.flags & RF_WARM) != 0u));
        num = dot2(dv, t);
        float pt = num * z.p[p][6];
        float acc_pn = z.acc[2 * p], acc_pt = z.acc[2 * p + 1];
        float new_pn = maxn(0.0f, acc_pn + pn);
        float applied_pn = new_pn - acc_pn;
        float max_pt = z.fric * absf(new_pn);
        float new_pt = clampn(acc_pt + pt, -max_pt, max_pt);
        float applied_pt = new_pt - acc_pt;
        const bool clr = on && sep;
        z.acc[2 * p] = clr ? 0.f : (go ? new_pn : acc_pn);
        z.acc[2 * p + 1] = clr ? 0.f : (go ? new_pt : acc_pt);
        v2 dp = add2(scale2(n, applied_pn), scale2(t, applied_pt));
        const bool g1 = go && !(z.flags & RF_S1), g2 = go && !(z.flags & RF_S2);
        const float a1x = app1.x - dp.x, a1y = app1.y - dp.y, q1 = rot1 - cross2(r1, dp);
        const float a2x = app2.x + dp.x, a2y = app2.y + dp.y, q2 = rot2 + cross2(r2, dp);
        app1.x = g1 ? a1x : app1.x; app1.y = g1 ? a1y : app1.y; rot1 = g1 ? q1 : rot1;
        app2.x = g2 ? a2x : app2.x; app2.y = g2 ? a2y : app2.y; rot2 = g2 ? q2 : rot2;
    }
    if (z.flags & RF_W1) { st.mx[l1] = b1.mx + app1.x; st.my[l1] = b1.my + app1.y; st.L[l1] = b1.L + rot1; }
    if (z.flags & RF_W2) { st.mx[l2] = b2.mx + app2.x; st.my[l2] = b2.my + app2.y; st.L[l2] = b2.L + rot2; }
}

// record -> the step's manifold arrays (what r2d_get_manifolds reports)
__device__ __forceinline__ void rec_writeback(const Rec& r, const StepDev& S)
{
    if (r.flags & RF_JOINT) return;
    const size_t cm = S.cap_m;
    float* P = S.m_pt;
    const uint32_t m = r.idx;
    S.m_fric[m] = r.fric;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(r.flags & (RF_P0 << p))) continue;
        P[(MP_ACC_PN * 2 + p) * cm + m] = r.acc[2 * p];
        P[(MP_ACC_PT * 2 + p) * cm + m] = r.acc[2 * p + 1];
        P[(MP_MASS_N * 2 + p) * cm + m] = r.p[p][5];
        P[(MP_MASS_T * 2 + p) * cm + m] = r.p[p][6];
        P[(MP_R1X * 2 + p) * cm + m] = r.p[p][0];
        P[(MP_R1Y * 2 + p) * cm + m] = r.p[p][1];
        P[(MP_R2X * 2 + p) * cm + m] = r.p[p][2];
        P[(MP_R2Y * 2 + p) * cm + m] = r.p[p][3];
        P[(MP_DEPTH * 2 + p) * cm + m] = r.p[p][4];
    }
}

// LDS of the register-slot kernel: 13 words of body state per body
__host__ __device__ inline size_t slot_lds_bytes(int nbodies) { return (size_t)nbodies * 13u * 4u; }

// Which solve kernel steps a world: 0 the register-slot kernel (at most
// slot_nmax bodies and slot_cap items), 1 the record kernel with LDS body
// state (at most lds_nmax bodies), 2 the record kernel with body state in
// global memory.  Every kernel evaluates the same rule, so each world is
// stepped by exactly one of them.
__device__ __forceinline__ int solve_route(uint32_t n, uint32_t nitems, uint32_t slot_cap, uint32_t slot_nmax,
                                           uint32_t lds_nmax, bool warm)
{
    if (slot_cap && !warm && n <= slot_nmax && nitems <= slot_cap) return 0;
    if (n <= lds_nmax) return 1;
    return 2;
}

// ---------------------------------------------- solve building blocks
// Per-body and per-item parts of Solver.process's sub-step loop
// (lib.zig:199-250), shared by the one-workgroup-per-world kernels and by the
// grid-wide kernels used for worlds whose body state does not fit LDS.

// inverse mass / inertia (static: 0) and item count of body i (degs: the
// world's item counts per body)
template <class ST>
__device__ __forceinline__ void body_setup(const WorldDev& D, const ST& st, int i, bool copy_state,
                                           const uint32_t* degs)
{
    bool stat = (D.flags[i] & BF_STATIC) != 0;
    if (copy_state) { st.mx[i] = D.mx[i]; st.my[i] = D.my[i]; st.L[i] = D.L[i]; }
    st.im[i] = stat ? 0.0f : 1.0f / D.mass[i];
    st.ii[i] = stat ? 0.0f : 1.0f / D.inertia[i];
    st.deg[i] = degs[i];
}

// The body arrays that the per-body passes of a sub-step read and write,
// copied out of the world's descriptor once per pass and typed by address
// space: global for the body arrays (their loads and stores then wait on the
// memory counter alone; a generic pointer's wait also covers LDS operations),
// constant for the force generators, which no kernel writes (scalar loads,
// one per generator for the whole wave).  Through a reference to the
// descriptor in global memory the compiler reloads every pointer for every
// body, since it cannot tell that the pass's own stores leave it unchanged.
typedef __attribute__((address_space(1))) float* GlbF;
typedef const __attribute__((address_space(1))) float* GlbCF;
typedef const __attribute__((address_space(1))) uint32_t* GlbCU;
typedef const __attribute__((address_space(4))) Gen* ConstGen;
struct BodyArrays {
    GlbF px, py, ang, fx, fy, tq, ax, ay, ahw, ahh;
    GlbCF mass, inertia, rad, hw, hh;
    GlbCU flags;
    GlbCF wsum;                      // the world's float4 warm sums, as floats
    ConstGen gens;
    int32_t ngens;
    float warm;
};
__device__ __forceinline__ BodyArrays body_arrays(const WorldDev& D)
{
    return BodyArrays{(GlbF)D.px, (GlbF)D.py, (GlbF)D.ang, (GlbF)D.fx, (GlbF)D.fy, (GlbF)D.tq, (GlbF)D.ax,
                      (GlbF)D.ay, (GlbF)D.ahw, (GlbF)D.ahh, (GlbCF)D.mass, (GlbCF)D.inertia, (GlbCF)D.rad,
                      (GlbCF)D.hw, (GlbCF)D.hh, (GlbCU)D.flags, (GlbCF)(const float*)D.wsum, (ConstGen)D.gens,
                      D.ngens, D.warm};
}

// body i's warm impulse per sub-step (momentum x, y, angular: the first
// three floats of its float4)
__device__ __forceinline__ const float* warm_sum(const WorldDev& D, int i)
{
    return reinterpret_cast<const float*>(&D.wsum[i]);
}
__device__ __forceinline__ GlbCF warm_sum(const BodyArrays& D, int i) { return D.wsum + 4 * i; }

// what body_forces reads of body i, loaded before any of it is used (one
// round trip; every index is valid whatever the flags)
struct ForceIn {
    uint32_t f;
    float fx, fy, mass, px, py, tq, ang, hw, hh, rad;
};
template <class DV>
__device__ __forceinline__ ForceIn force_in(const DV& D, int i)
{
    return ForceIn{D.flags[i], D.fx[i], D.fy[i], D.mass[i], D.px[i], D.py[i], D.tq[i], D.ang[i], D.hw[i], D.hh[i],
                   D.rad[i]};
}

// forces (lib.zig:200-205), AABB refresh + velocity integration
// (lib.zig:207-216) for body i from its loaded inputs; clears its solver
// counter.  DV: WorldDev or BodyArrays.
template <bool WARM = false, class ST, class DV>
__device__ __forceinline__ void body_forces_in(const DV& D, const ST& st, int i, const ForceIn& in, bool last,
                                               float sub_dt)
{
    const uint32_t f = in.f;
    const float fx0 = in.fx, fy0 = in.fy, mass = in.mass, px = in.px, py = in.py, tq = in.tq;
    const float ang = in.ang, bhw = in.hw, bhh = in.hh, rad = in.rad;
    const bool stat = (f & BF_STATIC) != 0;
    float fx = fx0, fy = fy0;
    if (!stat) {
        for (int g = 0; g < D.ngens; ++g) {
            const auto& G = D.gens[g];   // fields read one by one (keeps G's address space)
            if (G.type == 0) {
                fx += G.gx * mass;
                fy += G.gy * mass;
            } else if (G.slot == i) {
                float inv_m = 1.0f / mass;
                v2 v = scale2(mk(st.mx[i], st.my[i]), inv_m);
                v2 d = sub2(mk(px, py), mk(G.ax, G.ay));
                fx += (-G.k) * d.x - G.c * v.x;
                fy += (-G.k) * d.y - G.c * v.y;
            }
        }
    }
    if (last) {
        // Disc.updateAABB / Rectangle.updateAABB; only the last refresh is observable
        float hw, hh;
        if (f & BF_RECT) {
            float c = cosf_r(ang), sn = sinf_r(ang);
            float x = bhw, y = bhh;
            v2 lv[4] = {mk(-x, -y), mk(-x, y), mk(x, y), mk(x, -y)};
            hw = 0.f; hh = 0.f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v2 r = rotcs(lv[k], c, sn);
                if (r.x > hw) hw = r.x;
                if (r.y > hh) hh = r.y;
            }
        } else {
            hw = rad; hh = rad;
        }
        D.ax[i] = px; D.ay[i] = py; D.ahw[i] = hw; D.ahh[i] = hh;
    }
    if (!stat) {
        st.mx[i] += fx * sub_dt;
        st.my[i] += fy * sub_dt;
        st.L[i] += tq * sub_dt;
        if constexpr (WARM) {
            if (D.warm > 0.0f) {
                // warm starting: the body's share of the matched manifolds' carried impulses
                const auto ws = warm_sum(D, i);
                st.mx[i] += ws[0];
                st.my[i] += ws[1];
                st.L[i] += ws[2];
            }
        }
    }
    D.fx[i] = fx; D.fy[i] = fy;
    st.cnt[i] = 0u;
}

template <bool WARM = false, class ST, class DV>
__device__ __forceinline__ void body_forces(const DV& D, const ST& st, int i, bool last, float sub_dt)
{
    body_forces_in<WARM>(D, st, i, force_in(D, i), last, sub_dt);
}

// what body_positions reads of body i, loaded before any of it is used
struct PosIn {
    uint32_t f;
    float mass, inertia, px, py, ang;
};
template <class DV>
__device__ __forceinline__ PosIn pos_in(const DV& D, int i)
{
    return PosIn{D.flags[i], D.mass[i], D.inertia[i], D.px[i], D.py[i], D.ang[i]};
}

// position integration (lib.zig:238-249) for body i from its loaded inputs;
// DV: WorldDev or BodyArrays
template <class ST, class DV>
__device__ __forceinline__ void body_positions_in(const DV& D, const ST& st, int i, const PosIn& in, float sub_dt)
{
    const uint32_t f = in.f;
    const float mass = in.mass, inertia = in.inertia, px = in.px, py = in.py, ang = in.ang;
    if (f & BF_STATIC) return;
    float qd = sub_dt / mass;
    D.px[i] = px + st.mx[i] * qd;
    D.py[i] = py + st.my[i] * qd;
    D.ang[i] = ang + st.L[i] * sub_dt / inertia;
    D.fx[i] = 0.0f; D.fy[i] = 0.0f;
    D.tq[i] = 0.0f;
}

template <class ST, class DV>
__device__ __forceinline__ void body_positions(const DV& D, const ST& st, int i, float sub_dt)
{
    body_positions_in(D, st, i, pos_in(D, i), sub_dt);
}

// A per-body pass over [0, n) in rounds of K bodies per thread: each round
// loads the inputs of all its bodies before it applies the update to any, so
// their loads are in flight together (bodies are independent in these
// passes, so the order changes no result).
template <int K, class In, class Load, class Apply>
__device__ __forceinline__ void body_pass(int n, Load load, Apply apply)
{
    const int T = blockDim.x;
    for (int i0 = threadIdx.x; i0 < n; i0 += K * T) {
        In in[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (i0 + k * T < n) in[k] = load(i0 + k * T);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (i0 + k * T < n) apply(i0 + k * T, in[k]);
    }
}

// ------------------------------------------------------------ warm starting
// Opt-in, not in the reference (its roadmap item "persist manifolds across
// frames by keying with stable body IDs", README.md:60-61).  After a step
// every manifold is kept under the lower slot of its two bodies (WARM_K
// entries per body, discovery order); at the next step a manifold with the
// same (ref id, inc id), the same points, a normal within WARM_DOT of the old
// one and no live updateTGSDepth point carries warm * acc / sub_steps of each
// point's impulses into every sub-step: added to its accumulators before the
// iterations, and applied to its bodies with the forces.

// manifold m (step index): match against its lower body's table; the points'
// impulses per sub-step to MP_WARM_N / MP_WARM_T, MV_WARM on a match
__device__ __forceinline__ void warm_match(const StepDev& S, const WorldDev& D, uint32_t m, uint32_t base,
                                           uint32_t sub_steps)
{
    const size_t cm = S.cap_m;
    float* P = S.m_pt;
    const uint32_t valid = S.m_valid[m] & 3u;
    const uint32_t a = S.m_ref[m] - base, b = S.m_inc[m] - base;
    const uint32_t lo = umin(a, b);
    const v2 n = mk(S.m_nx[m], S.m_ny[m]);
    bool live = false;
#pragma unroll
    for (int p = 0; p < 2; ++p)
        if (valid & (1u << p)) live |= !approx(P[(MP_DEPTH * 2 + p) * cm + m], P[(MP_ODEPTH * 2 + p) * cm + m], 1e-4f);
    float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool hit = false;
    if (!live) {
        const uint32_t ida = D.id[a], idb = D.id[b];
        const uint32_t c = umin(D.wcnt[lo], WARM_K);
        const uint4* e = D.wtab + (size_t)lo * WARM_K * 3;
        for (uint32_t k = 0; k < c; ++k, e += 3) {
            const uint4 h = e[0];
            if (h.x != ida || h.y != idb) continue;
            const uint4 q = e[1], r = e[2];
            const float d = dot2(n, mk(__uint_as_float(h.w), __uint_as_float(q.x)));
            if (h.z == valid && d >= WARM_DOT) {
                hit = true;
                const float acc[4] = {__uint_as_float(q.y), __uint_as_float(q.z), __uint_as_float(q.w),
                                      __uint_as_float(r.x)};
                const float ss = (float)sub_steps;
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    if (!(valid & (1u << p)) || !(P[(MP_DEPTH * 2 + p) * cm + m] < 0.0f)) continue;
                    g[2 * p] = (D.warm * acc[2 * p]) / ss;
                    g[2 * p + 1] = (D.warm * acc[2 * p + 1]) / ss;
                }
            }
            break;
        }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        P[(MP_WARM_N * 2 + p) * cm + m] = g[2 * p];
        P[(MP_WARM_T * 2 + p) * cm + m] = g[2 * p + 1];
    }
    if (hit) S.m_valid[m] = valid | MV_WARM;
}

// body x (world slot) from its item list (entries item * 2 + side, items
// counted from ibase, in item order): the sum, in that order, of its matched
// manifolds' carried impulses, as calculateImpulses applies them
// (collision.zig:200-217): the reference body gets -dp, -r1 x dp, the
// incident body +dp, +r2 x dp
__device__ __forceinline__ void warm_body_sum(const StepDev& S, const WorldDev& D, const uint32_t* list, uint32_t cnt,
                                              uint32_t x, uint32_t ibase, uint32_t nj, uint32_t mbeg)
{
    const size_t cm = S.cap_m;
    const float* P = S.m_pt;
    float sx = 0.0f, sy = 0.0f, sl = 0.0f;
    for (uint32_t i = 0; i < cnt; ++i) {
        const uint32_t v = list[i], j = (v >> 1) - ibase;
        if (j < nj) continue;
        const uint32_t m = mbeg + (j - nj);
        const uint32_t valid = S.m_valid[m];
        if (!(valid & MV_WARM)) continue;
        const v2 n = mk(S.m_nx[m], S.m_ny[m]), t = rot90cw(n);
        for (int p = 0; p < 2; ++p) {
            if (!(valid & (1u << p))) continue;
            const v2 dp = add2(scale2(n, P[(MP_WARM_N * 2 + p) * cm + m]), scale2(t, P[(MP_WARM_T * 2 + p) * cm + m]));
            if (v & 1u) {
                const v2 r2 = mk(P[(MP_R2X * 2 + p) * cm + m], P[(MP_R2Y * 2 + p) * cm + m]);
                sx += dp.x; sy += dp.y; sl += cross2(r2, dp);
            } else {
                const v2 r1 = mk(P[(MP_R1X * 2 + p) * cm + m], P[(MP_R1Y * 2 + p) * cm + m]);
                sx -= dp.x; sy -= dp.y; sl -= cross2(r1, dp);
            }
        }
    }
    D.wsum[x] = make_float4(sx, sy, sl, 0.0f);
}

// after the step: manifold m of the world's [mbeg, mend) into its lower
// body's table at its rank among that body's manifolds (contiguous in
// discovery order: a pair is found from its lower slot); the last of them
// sets the count.  The world's counts are zero beforehand.
__device__ __forceinline__ void warm_save(const StepDev& S, const WorldDev& D, uint32_t m, uint32_t mbeg,
                                          uint32_t mend, uint32_t base)
{
    auto lo_of = [&](uint32_t k) { return umin(S.m_ref[k], S.m_inc[k]) - base; };
    const uint32_t lo = lo_of(m);
    uint32_t k = 0;
    while (k < WARM_K && m - k > mbeg && lo_of(m - k - 1) == lo) ++k;
    if (m + 1 == mend || lo_of(m + 1) != lo) D.wcnt[lo] = umin(k + 1, WARM_K);
    if (k >= WARM_K) return;
    const size_t cm = S.cap_m;
    const float* P = S.m_pt;
    uint4* e = D.wtab + ((size_t)lo * WARM_K + k) * 3;
    const uint32_t a = S.m_ref[m] - base, b = S.m_inc[m] - base;
    e[0] = make_uint4(D.id[a], D.id[b], S.m_valid[m] & 3u, __float_as_uint(S.m_nx[m]));
    e[1] = make_uint4(__float_as_uint(S.m_ny[m]), __float_as_uint(P[(MP_ACC_PN * 2 + 0) * cm + m]),
                      __float_as_uint(P[(MP_ACC_PT * 2 + 0) * cm + m]), __float_as_uint(P[(MP_ACC_PN * 2 + 1) * cm + m]));
    e[2] = make_uint4(__float_as_uint(P[(MP_ACC_PT * 2 + 1) * cm + m]), 0u, 0u, 0u);
}

// the record of execution position pos (first sub-step's preStep)
template <bool WARM = false, class ST>
__device__ __forceinline__ void item_build(const StepDev& S, const WorldDev& D, const ST& st, uint4* recs, uint32_t pos,
                                           uint32_t ibase, uint32_t mbeg, float sub_dt)
{
    Rec r;
    rec_build<WARM>(r, S, D, st, pos, ibase, (uint32_t)D.njoints, mbeg, D.body_base);
    r.degs = rec_degs(r, st);
    if (r.flags & RF_JOINT) {
        const Joint& J = D.joints[r.idx];
        rec_joint_params(r, J, D);
        rec_joint_pose(r, J, D);
    } else {
        rec_biases(r, sub_dt);
    }
    rec_store(r, recs + (size_t)pos * 8);
}

// The body arrays that the per-sub-step item pass reads, copied out of the
// world's descriptor once per pass: through a reference to the descriptor
// (global memory) the compiler reloads each pointer for every item, since it
// cannot tell that the pass's own stores leave the descriptor unchanged.
struct PassArrays {
    const float *px, *py, *ang, *inertia, *tq;
    const Joint* joints;
};
__device__ __forceinline__ PassArrays pass_arrays(const WorldDev& D)
{
    return PassArrays{D.px, D.py, D.ang, D.inertia, D.tq, D.joints};
}

// updateTGSDepth (collision.zig:73-100): live only for records flagged at
// build time; every manifold's previous angles end as the last sub-step's.
// DV: WorldDev or PassArrays.
// h: the record's header row (recs[pos * 8])
template <bool WARM = false, class ST, class DV>
__device__ __forceinline__ void item_tgs_h(const StepDev& S, const DV& D, const ST& st, uint4* recs, uint32_t pos,
                                           uint4 h, bool last, float sub_dt)
{
    uint4* q = recs + (size_t)pos * 8;
    if (h.x & RF_JOINT) {
        Rec r;
        rec_load(r, q);
        rec_joint_pose(r, D.joints[h.w], D);
        rec_store(r, q);
        return;
    }
    if (WARM && (h.x & RF_WARM)) {
        // warm starting: the carried impulse per sub-step joins the accumulators
        const size_t cm = S.cap_m;
        const float* P = S.m_pt;
        uint4 a = q[REC_ACC];
        if (h.x & RF_P0) {
            a.x = __float_as_uint(__uint_as_float(a.x) + P[(MP_WARM_N * 2 + 0) * cm + h.w]);
            a.y = __float_as_uint(__uint_as_float(a.y) + P[(MP_WARM_T * 2 + 0) * cm + h.w]);
        }
        if (h.x & RF_P1) {
            a.z = __float_as_uint(__uint_as_float(a.z) + P[(MP_WARM_N * 2 + 1) * cm + h.w]);
            a.w = __float_as_uint(__uint_as_float(a.w) + P[(MP_WARM_T * 2 + 1) * cm + h.w]);
        }
        q[REC_ACC] = a;
    }
    if (h.x & RF_TGS) {
        Rec r;
        rec_load(r, q);
        rec_tgs(r, S, D, st);
        rec_biases(r, sub_dt);
        rec_store(r, q);
    } else if (last) {
        S.m_prev1[h.w] = D.ang[h.y];
        S.m_prev2[h.w] = D.ang[h.z];
    }
}

template <bool WARM = false, class ST, class DV>
__device__ __forceinline__ void item_tgs(const StepDev& S, const DV& D, const ST& st, uint4* recs, uint32_t pos,
                                         bool last, float sub_dt)
{
    item_tgs_h<WARM>(S, D, st, recs, pos, recs[(size_t)pos * 8], last, sub_dt);
}

// K iterations of one sub-step's (constraints, manifolds) for the world of
// this workgroup, as an ordered dataflow
template <bool LDS, bool COMPACT, bool WARM = false, class ST>
__device__ __forceinline__ void item_chain(const StepDev& S, const WorldDev& D, const ST& st, uint4* recs,
                                           uint32_t nitems, uint32_t iters, float sub_dt, uint32_t w,
                                           uint32_t& s_abort)
{
    // K iterations of (constraints, manifolds) as one ordered dataflow.  With
    // LDS body state a link's counters and its bodies' state are read
    // together (LDS serves one workgroup's operations in order, so state read
    // after a counter that shows the item ready is what its predecessor
    // released; values read for an item that is not ready are discarded);
    // with body state in global memory the state is read after an acquire.
    {
        const uint32_t p_begin = threadIdx.x, p_step = blockDim.x;
        uint32_t pcur = p_begin, kcur = (p_begin < nitems) ? 0u : iters;
        Rec cur, nxt;
        uint32_t cur_pos = 0, cur_k = 0, nxt_pos = 0, nxt_k = 0;
        bool have = false, has_nxt = false;
        if (kcur < iters) {
            rec_load(nxt, recs + (size_t)pcur * 8);
            if (COMPACT) rec_to_compact(nxt);
            nxt_pos = pcur; nxt_k = kcur; has_nxt = true;
            pcur += p_step;
            if (pcur >= nitems) { pcur = p_begin; ++kcur; }
        }
        int32_t r0 = -1, r1 = -1;
        uint32_t t0 = 0, t1 = 0;
        uint32_t spins = 0;
        while (true) {
            if (!have) {
                if (!has_nxt) break;
                cur = nxt;
                cur_pos = nxt_pos;
                cur_k = nxt_k;
                has_nxt = false;
                have = true;
                // prefetch the following position; the same position in the
                // next iteration reads impulses this record is about to write,
                // so that case is forwarded from registers instead
                if (kcur < iters) {
                    if (pcur != cur_pos) {
                        rec_load(nxt, recs + (size_t)pcur * 8);
                        if (COMPACT) rec_to_compact(nxt);
                    }
                    nxt_pos = pcur; nxt_k = kcur; has_nxt = true;
                    pcur += p_step;
                    if (pcur >= nitems) { pcur = p_begin; ++kcur; }
                }
                r0 = rec_r0(cur);
                r1 = rec_r1(cur);
                if (cur.degs != DEGS_WIDE) {
                    t0 = cur_k * (cur.degs & 0xffffu) + cur.rk0;
                    t1 = cur_k * (cur.degs >> 16) + cur.rk1;
                } else {
                    if (r0 >= 0) t0 = cur_k * st.deg[r0] + cur.rk0;
                    if (r1 >= 0) t1 = cur_k * st.deg[r1] + cur.rk1;
                }
            }
            const bool contact = !(cur.flags & RF_JOINT);
            BodyState b1{}, b2{};
            bool ready;
            if (LDS) {
                uint32_t c0 = r0 >= 0 ? __hip_atomic_load(&st.cnt[r0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                      : t0;
                uint32_t c1 = r1 >= 0 ? __hip_atomic_load(&st.cnt[r1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                      : t1;
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (contact) {
                    b1 = body_state(st, cur.a);
                    b2 = body_state(st, cur.b);
                }
                ready = c0 == t0 && c1 == t1;
            } else {
                ready = true;
                if (r0 >= 0) ready = ld_acq(&st.cnt[r0]) == t0;
                if (ready && r1 >= 0) ready = ld_acq(&st.cnt[r1]) == t1;
                if (ready && contact) {
                    b1 = body_state_ordered(st, cur.a);
                    b2 = body_state_ordered(st, cur.b);
                }
            }
            if (ready) {
                // a wave with a ready link issues ahead of the waves that poll
                __builtin_amdgcn_s_setprio(2);
                if (!contact) {
                    const uint32_t ua = COMPACT ? (cur.cpair & 0xffffu) : cur.a;
                    const uint32_t ub = COMPACT ? (cur.cpair >> 16) : (cur.b == NONE ? cur.a : cur.b);
                    rec_solve_joint(cur, st, ua, ub, sub_dt);
                } else {
                    contact_impulses<WARM>(cur, cur.a, cur.b, b1, b2, st);
                    recs[(size_t)cur_pos * 8 + REC_ACC] = make_uint4(__float_as_uint(cur.acc[0]),
                                                                     __float_as_uint(cur.acc[1]),
                                                                     __float_as_uint(cur.acc[2]),
                                                                     __float_as_uint(cur.acc[3]));
                }
                if (LDS) {
                    // the state writes above reach LDS before the counters (in order)
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (r0 >= 0) __hip_atomic_store(&st.cnt[r0], t0 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (r1 >= 0) __hip_atomic_store(&st.cnt[r1], t1 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    if (r0 >= 0) st_rel(&st.cnt[r0], t0 + 1u);
                    if (r1 >= 0) st_rel(&st.cnt[r1], t1 + 1u);
                }
                __builtin_amdgcn_s_setprio(0);
                if (has_nxt && nxt_pos == cur_pos) nxt = cur;
                have = false;
                spins = 0;
            } else {
                // every item's predecessors precede it, so some item is always
                // ready; a stall this long means corrupt ordering data: stop
                if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                if (++spins > SPIN_LIMIT) {
                    __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    S.w_status[w] = 5;
                    break;
                }
            }
            // a wave with no lane ready yields its SIMD
            if (__ballot(ready) == 0ull) __builtin_amdgcn_s_sleep(1);
        }
    }
}

// Keeps the compiler from moving LDS accesses across this point, without a
// wait: one workgroup's LDS operations are served in the order they are
// issued, so program order is all the dataflow needs.
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// A lane's slot: the five rows of its current record that a chain link reads
// besides the header and the impulses (rows REC_LINK.., copied as they are):
// (nx ny fric p00) (p01 p02 p03 p04) (p05 p06 p10 p11) (p12 p13 p14 p15)
// (p16 bias0 bias1 -).
constexpr uint32_t LANE_ROWS = REC_ACC - REC_LINK;
constexpr uint32_t LANE_WORDS = LANE_ROWS * 4;
static_assert(REC_LINK == 2 && REC_ACC == 7 && LANE_ROWS == 5, "lane_load reads rows REC_LINK..REC_ACC-1 in this order");
// the record kernel takes the lane slots up to this workgroup size (larger
// ones keep the chain's records in registers)
constexpr uint32_t LANES_MAXT = 512;
// the record kernel's workgroup when a batch holds few worlds with LDS body state
constexpr int REC_WIDE_THREADS = 512;
__host__ __device__ constexpr size_t rec_lane_bytes(uint32_t threads)
{
    return threads <= LANES_MAXT ? (size_t)threads * LANE_WORDS * 4u : 0u;
}

// a contact's impulses back to its record
__device__ __forceinline__ void rec_put_acc(uint4* recs, uint32_t pos, const float* acc)
{
    recs[(size_t)pos * 8 + REC_ACC] = make_uint4(__float_as_uint(acc[0]), __float_as_uint(acc[1]),
                                                 __float_as_uint(acc[2]), __float_as_uint(acc[3]));
}

// a lane's slot rows are blockDim.x apart (row r of lane t at lanes[r * T + t]):
// a wave's 16-byte row reads and writes then cover consecutive LDS words,
// free of bank conflicts (rows contiguous per lane, 80 bytes apart, were not)
__device__ __forceinline__ void lane_load(Rec& r, const uint4* q, uint32_t stride)
{
    const uint4 a = q[0], b = q[stride], c = q[2 * stride], d = q[3 * stride], e = q[4 * stride];
    r.nx = __uint_as_float(a.x); r.ny = __uint_as_float(a.y); r.fric = __uint_as_float(a.z);
    r.p[0][0] = __uint_as_float(a.w);
    r.p[0][1] = __uint_as_float(b.x); r.p[0][2] = __uint_as_float(b.y); r.p[0][3] = __uint_as_float(b.z);
    r.p[0][4] = __uint_as_float(b.w);
    r.p[0][5] = __uint_as_float(c.x); r.p[0][6] = __uint_as_float(c.y); r.p[1][0] = __uint_as_float(c.z);
    r.p[1][1] = __uint_as_float(c.w);
    r.p[1][2] = __uint_as_float(d.x); r.p[1][3] = __uint_as_float(d.y); r.p[1][4] = __uint_as_float(d.z);
    r.p[1][5] = __uint_as_float(d.w);
    r.p[1][6] = __uint_as_float(e.x); r.bias[0] = __uint_as_float(e.y); r.bias[1] = __uint_as_float(e.z);
}

// The same ordered dataflow as item_chain (body state in LDS), with a
// lane's current record kept in its LDS slot (lanes: blockDim.x slots of
// LANE_WORDS, 16-byte aligned) instead of registers: the loop carries only
// the header (flags, bodies, ranks, targets) and the impulses, so an
// iteration moves no record between registers, and a poll reads just the two
// counters.  The slot and the bodies' state are read in one batch once the
// item is ready.  A link issues no global memory operation: a contact's
// impulses go back to its record when the lane moves to another position
// and when the chain ends (a wave's memory counter is waited on as a whole,
// so a store per link would make every later wait of the wave wait for it);
// a lane whose next position is its current one (one item per iteration)
// keeps slot and impulses.  COMPACT: body state under the compact chain
// indices of worlds beyond LDS (record row 1 .w), as in rec_to_compact.
// GLOBAL: body state and counters in global memory (worlds whose touched
// bodies do not fit LDS): counters are polled with acquire loads and
// released with release stores, and the state is read after the acquire.
// HYBRID (with GLOBAL): the state stays in global memory but the counters
// are in LDS (lcnt) under the compact chain indices, so a poll is an LDS
// read; acquire / release at workgroup scope order the state around them.
template <bool COMPACT = false, bool GLOBAL = false, bool WARM = false, bool HYBRID = false, class ST>
__device__ __forceinline__ void item_chain_lanes(int32_t* w_status, const ST& st, uint4* recs, uint4* lanes,
                                                 uint32_t nitems, uint32_t iters, float sub_dt, uint32_t w,
                                                 uint32_t& s_abort, uint32_t* lcnt = nullptr)
{
    static_assert(!HYBRID || (GLOBAL && !COMPACT), "the LDS counters go with global-memory state");
    constexpr bool LITE = IsLite<ST>::value;
    static_assert(!LITE || (COMPACT && !GLOBAL), "the light state is a compact LDS state");
    const uint32_t p_begin = threadIdx.x, p_step = blockDim.x;
    uint4* slot = lanes + threadIdx.x;          // row r at slot[r * p_step]
    uint32_t pos = p_begin, k = (p_begin < nitems) ? 0u : iters;
    uint32_t held = NONE;                    // position whose record the slot holds
    uint32_t flags = 0, a = 0, b = 0, rk0 = 0, rk1 = 0, degs = 0;
    int32_t r0 = -1, r1 = -1;
    uint32_t t0 = 0, t1 = 0, n0 = 0, n1 = 0;  // the counters' targets and released values
    uint32_t* const cbase = HYBRID ? lcnt : st.cnt;
    uint32_t *p0 = cbase, *p1 = cbase;       // the counters polled (body 0's when none)
    uint32_t cpair = 0;                      // HYBRID: the compact indices of a and b
    float acc[4] = {0.f, 0.f, 0.f, 0.f};     // the held record's impulses
    bool have = false;
    uint32_t spins = 0;
    while (true) {
        if (!have) {
            if (k >= iters) break;
            if (pos != held) {
                if (held != NONE && !(flags & RF_JOINT)) rec_put_acc(recs, held, acc);
                const uint4* q = recs + (size_t)pos * 8;
                const uint4 h0 = q[0], h1 = q[1], ac = q[REC_ACC];
                // every row's load is issued before any is written to the
                // slot (one round trip for the record): without the barrier
                // the scheduler may put each row's load after the previous
                // row's LDS write, one round trip per row
                const uint4 w0 = q[REC_LINK], w1 = q[REC_LINK + 1], w2 = q[REC_LINK + 2], w3 = q[REC_LINK + 3],
                            w4 = q[REC_LINK + 4];
                __builtin_amdgcn_sched_barrier(0);
                slot[0] = w0;
                slot[p_step] = w1;
                slot[2 * p_step] = w2;
                slot[3 * p_step] = w3;
                slot[4 * p_step] = w4;
                if constexpr (LITE) slot[LANE_ROWS * p_step] = st.rmass[pos];
                flags = h0.x; a = h0.y; b = h0.z; rk0 = h1.x; rk1 = h1.y; degs = h1.z;
                if (HYBRID) cpair = h1.w;
                if (COMPACT) {
                    const uint32_t c0 = h1.w & 0xffffu, c1 = h1.w >> 16;
                    a = c0;
                    b = ((flags & RF_JOINT) && b == NONE) ? NONE : c1;
                }
                acc[0] = __uint_as_float(ac.x); acc[1] = __uint_as_float(ac.y);
                acc[2] = __uint_as_float(ac.z); acc[3] = __uint_as_float(ac.w);
                held = pos;
            }
            if (flags & RF_JOINT) {
                r0 = a == NONE ? -1 : (int32_t)a;
                r1 = b == NONE ? -1 : (int32_t)b;
            } else {
                r0 = (flags & RF_W1) ? (int32_t)a : -1;
                r1 = (flags & RF_W2) ? (int32_t)b : -1;
            }
            if (degs != DEGS_WIDE) {
                t0 = k * (degs & 0xffffu) + rk0;
                t1 = k * (degs >> 16) + rk1;
            } else {
                t0 = r0 >= 0 ? k * st.deg[r0] + rk0 : 0u;
                t1 = r1 >= 0 ? k * st.deg[r1] + rk1 : 0u;
            }
            n0 = t0 + 1u;
            n1 = t1 + 1u;
            if (HYBRID) {
                p0 = cbase + (r0 >= 0 ? (cpair & 0xffffu) : 0u);
                p1 = cbase + (r1 >= 0 ? (cpair >> 16) : 0u);
            } else {
                p0 = cbase + (r0 >= 0 ? r0 : 0);
                p1 = cbase + (r1 >= 0 ? r1 : 0);
            }
            have = true;
        }
        // a poll reads the two counters only (a lane without an ordering
        // body reads body 0's and ignores it); the bodies' state is read once
        // the item is ready, in the same batch as the slot (LDS serves one
        // workgroup's operations in order, so it is what the predecessors
        // released)
        const bool contact = !(flags & RF_JOINT);
        bool ready;
        if (HYBRID) {
            ready = r0 < 0 || __hip_atomic_load(p0, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == t0;
            if (ready && r1 >= 0) ready = __hip_atomic_load(p1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == t1;
        } else if (GLOBAL) {
            ready = r0 < 0 || ld_acq(p0) == t0;
            if (ready && r1 >= 0) ready = ld_acq(p1) == t1;
        } else {
            const uint32_t c0 = __hip_atomic_load(p0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t c1 = __hip_atomic_load(p1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ready = (r0 < 0 || c0 == t0) && (r1 < 0 || c1 == t1);
        }
        if (ready) {
            // a wave with a ready link issues ahead of the waves that poll
            __builtin_amdgcn_s_setprio(2);
            lds_order();
            Rec cur;
            cur.flags = flags; cur.a = a; cur.b = b;
            lane_load(cur, slot, p_step);
#pragma unroll
            for (int f = 0; f < 4; ++f) cur.acc[f] = acc[f];
            // the light state's masses of this item's bodies, from its slot
            ST sl = st;
            if constexpr (LITE) {
                const uint4 m = slot[LANE_ROWS * p_step];
                sl.im = PairValue{a, __uint_as_float(m.x), __uint_as_float(m.z)};
                sl.ii = PairValue{a, __uint_as_float(m.y), __uint_as_float(m.w)};
            }
            const BodyState b1 = GLOBAL ? body_state_ordered(sl, contact ? a : 0u) : body_state(sl, contact ? a : 0u);
            const BodyState b2 = GLOBAL ? body_state_ordered(sl, contact ? b : 0u) : body_state(sl, contact ? b : 0u);
            if (!contact) {
                rec_solve_joint(cur, sl, a, b == NONE ? a : b, sub_dt);
            } else {
                contact_impulses<WARM>(cur, a, b, b1, b2, sl);
#pragma unroll
                for (int f = 0; f < 4; ++f) acc[f] = cur.acc[f];
            }
            // the state writes above reach LDS before the counters (in order)
            lds_order();
            if (HYBRID) {
                if (r0 >= 0) __hip_atomic_store(p0, n0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (r1 >= 0) __hip_atomic_store(p1, n1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else if (GLOBAL) {
                if (r0 >= 0) st_rel(p0, n0);
                if (r1 >= 0) st_rel(p1, n1);
            } else {
                if (r0 >= 0) __hip_atomic_store(p0, n0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (r1 >= 0) __hip_atomic_store(p1, n1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            __builtin_amdgcn_s_setprio(0);
            pos += p_step;
            if (pos >= nitems) { pos = p_begin; ++k; }
            have = false;
            spins = 0;
        } else {
            // every item's predecessors precede it, so some item is always
            // ready; a stall this long means corrupt ordering data: stop
            ++spins;
            if ((spins & 255u) == 0u) {
                if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                if (spins > SPIN_LIMIT) {
                    __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    w_status[w] = 5;
                    break;
                }
            }
        }
        // no sleep between polls: a sleeping wave notices a released
        // counter later, and the ready waves of other worlds on the SIMD
        // issue ahead of it anyway
    }
    if (held != NONE && !(flags & RF_JOINT)) rec_put_acc(recs, held, acc);
}

// The LDS-state chain of a sub-step as an out-of-line call.  Its register
// allocation is then its own: inlined, any change of register pressure
// elsewhere in the step kernel reshaped the loop (at worst a record pickup
// became one global round trip per row instead of one for all rows).  It
// takes pointers typed by address space (LDS: the body state laid out as in
// rec_world, the lane slots and the abort flag; global: the records and the
// world status) and values only, so the callee keeps LDS and global
// instructions and no argument block is copied.
typedef __attribute__((address_space(3))) uint32_t* LdsU32;
typedef __attribute__((address_space(3))) uint4* LdsU4;
typedef __attribute__((address_space(1))) int32_t* GlbI32;
typedef __attribute__((address_space(1))) uint4* GlbU4;
template <bool WARM>
__device__ __attribute__((noinline)) void chain_lds_call(GlbI32 w_status, LdsU32 lds, uint32_t n, GlbU4 recs,
                                                        LdsU4 lanes, uint32_t nitems, uint32_t iters, float sub_dt,
                                                        uint32_t w, LdsU32 s_abort)
{
    SolverState<true> st;
    st.mx = (float*)(uint32_t*)lds;
    st.my = st.mx + n;
    st.L = st.my + n;
    st.im = st.L + n;
    st.ii = st.im + n;
    st.cnt = (uint32_t*)(st.ii + n);
    st.deg = st.cnt + n;
    item_chain_lanes<false, false, WARM>((int32_t*)w_status, st, (uint4*)recs, (uint4*)lanes, nitems, iters, sub_dt, w,
                                         *(uint32_t*)s_abort);
}

// the records' impulses, effective masses, lever arms and depths -> the
// step's manifold arrays (what r2d_get_manifolds reports)
__device__ __forceinline__ void item_writeback(const StepDev& S, uint4* recs, uint32_t pos)
{
    Rec r;
    rec_load(r, recs + (size_t)pos * 8);
    rec_writeback(r, S);
}

// One workgroup steps one world: S sub-steps of forces, integration,
// pre-step and K solver iterations (lib.zig:199-250).  Solver items run as a
// dataflow: an item may run once every earlier item (in the schedule's order)
// that touches one of its ordering bodies has run, which gives the same bits
// as the sequential loop.  Workers own execution positions tid, tid+blockDim,
// ... and take them in (iteration, position) order, the order the ranks use,
// so the earliest unfinished item always has a worker.  Body state lives in
// LDS (lds: solve_lds_bytes(n) bytes); records at the world's item range
// [ibase, ibase + nitems) of S.recs; degs: the world's item counts per body.
// LANES: the chain keeps each lane's current record in its slot of `lanes`
// (LDS, blockDim.x * LANE_WORDS words, 16-byte aligned, disjoint from lds).
template <bool LANES, bool WARM = false>
__device__ __forceinline__ void rec_world(const StepDev& S, const WorldDev& D, uint32_t w, uint32_t ibase,
                                          uint32_t nitems, uint32_t mbeg, const uint32_t* degs, float dt,
                                          uint32_t sub_steps, uint32_t iters, uint32_t* lds, uint4* lanes,
                                          uint32_t& s_abort)
{
    if (threadIdx.x == 0) s_abort = 0u;
    const int n = D.nbodies;
    const float sub_dt = dt / (float)sub_steps;
    uint4* recs = S.recs + (size_t)ibase * 8;

    SolverState<true> st;
    st.mx = (float*)lds;
    st.my = st.mx + n;
    st.L = st.my + n;
    st.im = st.L + n;
    st.ii = st.im + n;
    st.cnt = (uint32_t*)(st.ii + n);
    st.deg = st.cnt + n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) body_setup(D, st, i, true, degs);
    __syncthreads();
    for (uint32_t pos = threadIdx.x; pos < nitems; pos += blockDim.x)
        item_build<WARM>(S, D, st, recs, pos, ibase, mbeg, sub_dt);
    for (uint32_t s = 0; s < sub_steps; ++s) {
        const bool last = s + 1 == sub_steps;
        const BodyArrays bf = body_arrays(D);
        body_pass<3, ForceIn>(
            n, [&](int i) { return force_in(bf, i); },
            [&](int i, const ForceIn& in) { body_forces_in<WARM>(bf, st, i, in, last, sub_dt); });
        __syncthreads();
        {
            // the next record's header is loaded while this one is updated
            const PassArrays pa = pass_arrays(D);
            uint32_t pos = threadIdx.x;
            uint4 hn = pos < nitems ? recs[(size_t)pos * 8] : make_uint4(0u, 0u, 0u, 0u);
            for (; pos < nitems; pos += blockDim.x) {
                const uint4 h = hn;
                if (pos + blockDim.x < nitems) hn = recs[(size_t)(pos + blockDim.x) * 8];
                item_tgs_h<WARM>(S, pa, st, recs, pos, h, last, sub_dt);
            }
        }
        __syncthreads();
        if constexpr (LANES)
            chain_lds_call<WARM>((GlbI32)S.w_status, (LdsU32)lds, (uint32_t)n, (GlbU4)recs, (LdsU4)lanes, nitems, iters,
                                 sub_dt, w, (LdsU32)&s_abort);
        else item_chain<true, false, WARM>(S, D, st, recs, nitems, iters, sub_dt, w, s_abort);
        __syncthreads();
        const BodyArrays bp = body_arrays(D);
        body_pass<3, PosIn>(
            n, [&](int i) { return pos_in(bp, i); },
            [&](int i, const PosIn& in) { body_positions_in(bp, st, i, in, sub_dt); });
        __syncthreads();
    }
    for (uint32_t pos = threadIdx.x; pos < nitems; pos += blockDim.x) item_writeback(S, recs, pos);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        D.mx[i] = st.mx[i]; D.my[i] = st.my[i]; D.L[i] = st.L[i];
    }
}

// the record kernel: the worlds solve_route gives it, one workgroup each
template <int MAXT, int WAVES = 1, bool WARM = false>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WAVES)))
void k_solve(StepDev S, float dt, uint32_t sub_steps, uint32_t iters, uint32_t slot_cap, uint32_t slot_nmax,
             uint32_t lds_nmax)
{
    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ uint32_t s_abort;
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = S.w_order[blockIdx.x];
    if (S.w_status[w] != 0) return;
    const uint32_t ibase = S.w_item_base[w];
    const uint32_t nitems = S.w_item_base[w + 1] - ibase;
    const WorldDev& D = S.worlds[w];
    if (solve_route((uint32_t)D.nbodies, nitems, slot_cap, slot_nmax, lds_nmax, D.warm > 0.0f) != 1) return;
    // LDS: the chain's lane slots (workgroups of at most LANES_MAXT threads),
    // then the body state
    constexpr bool LANES = MAXT <= LANES_MAXT;
    uint4* lanes = LANES ? reinterpret_cast<uint4*>(smem) : nullptr;
    uint32_t* body = smem + (LANES ? MAXT * LANE_WORDS : 0u);
    rec_world<LANES, WARM>(S, D, w, ibase, nitems, S.w_man_begin[w], S.deg + D.body_base, dt, sub_steps, iters, body,
                     lanes, s_abort);
}

// ---------------------------------- warm starting in the phase kernels
// grid-wide over the step's manifolds / bodies; worlds without warm starting
// (and stopped worlds) are skipped

// after the per-body item lists: the matches, then each body's sum
__global__ void k_warm_match(StepDev S, uint32_t sub_steps)
{
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t M = dev_count(&S.counters[C_MANIFOLDS], S.cap_m);
    for (uint32_t m = grid_tid(); m < M; m += grid_stride()) {
        const uint32_t w = S.bworld[S.m_ref[m]];
        const WorldDev& D = S.worlds[w];
        if (D.warm > 0.0f && S.w_status[w] == 0) warm_match(S, D, m, D.body_base, sub_steps);
    }
}

__global__ void k_warm_sums(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t x = grid_tid(); x < S.nb; x += grid_stride()) {
        const uint32_t w = S.bworld[x];
        const WorldDev& D = S.worlds[w];
        if (!(D.warm > 0.0f) || S.w_status[w] != 0) continue;
        const uint32_t o = S.deg_off[x];
        warm_body_sum(S, D, S.lists + o, S.deg_off[x + 1] - o, x - D.body_base, S.w_item_base[w],
                      (uint32_t)D.njoints, S.w_man_begin[w]);
    }
}

// after the solve: clear the tables' counts, then store the manifolds
__global__ void k_warm_clear(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t x = grid_tid(); x < S.nb; x += grid_stride()) {
        const uint32_t w = S.bworld[x];
        const WorldDev& D = S.worlds[w];
        if (D.warm > 0.0f && S.w_status[w] == 0) D.wcnt[x - D.body_base] = 0u;
    }
}

__global__ void k_warm_save(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t M = dev_count(&S.counters[C_MANIFOLDS], S.cap_m);
    for (uint32_t m = grid_tid(); m < M; m += grid_stride()) {
        const uint32_t w = S.bworld[S.m_ref[m]];
        const WorldDev& D = S.worlds[w];
        if (D.warm > 0.0f && S.w_status[w] == 0)
            warm_save(S, D, m, S.w_man_begin[w], S.w_man_begin[w + 1], D.body_base);
    }
}

// ------------------------------------- worlds beyond LDS: split sub-steps
// For a world whose body state does not fit LDS (more than lds_nmax bodies)
// the per-body and per-item parts run as grid-wide kernels over all such
// worlds' bodies, and only the dependency chains run one workgroup per world;
// body state, counters and records stay in global memory between launches.
__device__ __forceinline__ bool big_world(const WorldDev& D, uint32_t lds_nmax)
{
    return (uint32_t)D.nbodies > lds_nmax;
}

__device__ __forceinline__ SolverState<false> global_state(const StepDev& S, const WorldDev& D)
{
    SolverState<false> st;
    st.mx = D.mx; st.my = D.my; st.L = D.L;
    st.im = S.g_im + D.body_base; st.ii = S.g_ii + D.body_base;
    st.cnt = S.g_cnt + D.body_base; st.deg = S.deg + D.body_base;
    return st;
}

// phase 0: body setup; 1: forces + velocities (sub-step s); 2: positions
__global__ void k_big_bodies(StepDev S, uint32_t lds_nmax, int phase, uint32_t s, uint32_t sub_steps, float sub_dt)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t x = grid_tid(); x < S.nb; x += grid_stride()) {
        const uint32_t w = S.bworld[x];
        if (S.w_status[w] != 0) continue;
        const WorldDev& D = S.worlds[w];
        if (!big_world(D, lds_nmax)) continue;
        const SolverState<false> st = global_state(S, D);
        const int i = (int)(x - D.body_base);
        if (phase == 0) body_setup(D, st, i, false, S.deg + D.body_base);
        else if (phase == 1) body_forces<true>(D, st, i, s + 1 == sub_steps, sub_dt);
        else body_positions(D, st, i, sub_dt);
    }
}

// the per-item parts, gridDim.y workgroups per world: phase 0 builds the
// records, 1 is updateTGSDepth (sub-step s), 2 writes back
constexpr unsigned BIG_THREADS = 256;
constexpr unsigned BIG_SPLIT = 32;
__global__ __launch_bounds__(BIG_THREADS) void k_big_item_pass(StepDev S, uint32_t lds_nmax, int phase, uint32_t s,
                                                            uint32_t sub_steps, float sub_dt)
{
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = blockIdx.x;
    if (S.w_status[w] != 0) return;
    const WorldDev& D = S.worlds[w];
    if (!big_world(D, lds_nmax)) return;
    const SolverState<false> st = global_state(S, D);
    const uint32_t ibase = S.w_item_base[w], nitems = S.w_item_base[w + 1] - ibase, mbeg = S.w_man_begin[w];
    uint4* recs = S.recs + (size_t)ibase * 8;
    for (uint32_t pos = blockIdx.y * blockDim.x + threadIdx.x; pos < nitems; pos += gridDim.y * blockDim.x) {
        if (phase == 0) item_build<true>(S, D, st, recs, pos, ibase, mbeg, sub_dt);
        else if (phase == 1) item_tgs<true>(S, D, st, recs, pos, s + 1 == sub_steps, sub_dt);
        else item_writeback(S, recs, pos);
    }
}

// after the records: the compact chain indices, one workgroup per world
__global__ __launch_bounds__(BIG_THREADS) void k_big_items(StepDev S, uint32_t lds_nmax)
{
    __shared__ uint32_t s_scan[BIG_THREADS];
    __shared__ uint32_t s_total;
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = blockIdx.x;
    if (S.w_status[w] != 0) return;
    const WorldDev& D = S.worlds[w];
    if (!big_world(D, lds_nmax)) return;
    const uint32_t ibase = S.w_item_base[w], nitems = S.w_item_base[w + 1] - ibase;
    const uint32_t base = D.body_base, T = blockDim.x, t = threadIdx.x;
    const int n = D.nbodies;
    uint4* recs = S.recs + (size_t)ibase * 8;
    // compact chain indices: the bodies the items touch, numbered in slot order
    const uint32_t UNUSED = 0xffffffffu;
    for (int i = t; i < n; i += T) S.cmap[base + i] = UNUSED;
    __syncthreads();
    for (uint32_t pos = t; pos < nitems; pos += T) {
        const uint4 h = recs[(size_t)pos * 8];
        if (h.x & RF_JOINT) {
            const Joint& J = D.joints[h.w];
            S.cmap[base + J.s1] = 0u;
            if (J.s2 >= 0) S.cmap[base + J.s2] = 0u;
        } else {
            S.cmap[base + h.y] = 0u;
            S.cmap[base + h.z] = 0u;
        }
    }
    __syncthreads();
    // each wave numbers a contiguous range of slots, 64 at a time (coalesced
    // loads, ranks from the wave's ballot): count, the waves' offsets, assign
    constexpr uint32_t NW = BIG_THREADS / 64;
    const uint32_t lane = t & 63u, wv = t >> 6;
    const uint32_t span = (((uint32_t)n + NW - 1) / NW + 63u) & ~63u;
    const uint32_t lo = umin(wv * span, (uint32_t)n), hi = umin(lo + span, (uint32_t)n);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t cnt = 0;
    for (uint32_t i0 = lo; i0 < hi; i0 += 64u) {
        const uint32_t i = i0 + lane;
        cnt += (uint32_t)__popcll(__ballot(i < hi && S.cmap[base + i] == 0u));
    }
    if (lane == 0) s_scan[wv] = cnt;
    __syncthreads();
    uint32_t c = 0, total = 0;
    for (uint32_t k = 0; k < NW; ++k) {
        c += k < wv ? s_scan[k] : 0u;
        total += s_scan[k];
    }
    for (uint32_t i0 = lo; i0 < hi; i0 += 64u) {
        const uint32_t i = i0 + lane;
        const bool hit = i < hi && S.cmap[base + i] == 0u;
        const uint64_t bal = __ballot(hit);
        if (hit) {
            const uint32_t ci = c + (uint32_t)__popcll(bal & below);
            S.cmap[base + i] = ci;
            S.cinv[base + ci] = i;
        }
        c += (uint32_t)__popcll(bal);
    }
    if (t == 0) s_total = total;
    __syncthreads();
    const uint32_t nc = s_total;
    if (t == 0) S.w_nc[w] = nc;
    if (nc >= 0xffffu) return;   // the compact chain is not used (LDS holds far fewer)
    for (uint32_t pos = t; pos < nitems; pos += T) {
        uint4* q = recs + (size_t)pos * 8;
        const uint4 h = q[0];
        uint32_t c0, c1, s0, s1;
        if (h.x & RF_JOINT) {
            const Joint& J = D.joints[h.w];
            c0 = S.cmap[base + J.s1];
            c1 = J.s2 >= 0 ? S.cmap[base + J.s2] : 0xffffu;
            s0 = (uint32_t)J.s1;
            s1 = J.s2 >= 0 ? (uint32_t)J.s2 : s0;
        } else {
            c0 = S.cmap[base + h.y];
            c1 = S.cmap[base + h.z];
            s0 = h.y;
            s1 = h.z;
        }
        q[1].w = c0 | (c1 << 16);
        // the light compact chain's masses of the item's bodies
        S.rmass[ibase + pos] = make_uint4(__float_as_uint(S.g_im[base + s0]), __float_as_uint(S.g_ii[base + s0]),
                                          __float_as_uint(S.g_im[base + s1]), __float_as_uint(S.g_ii[base + s1]));
    }
}

// LDS bytes of the compact chain per body (momentum, inverse masses, counter, item count)
constexpr size_t COMPACT_BYTES = 7 * 4;
constexpr size_t LITE_BYTES = 4 * 4;   // the light compact state: momentum and counter
// the light compact chain runs in the 256-lane compact launch, with lane
// slots of six rows; both chain launches decide with this one size, so
// exactly one of them takes each world
constexpr uint32_t LITE_LANES = 256;
constexpr size_t LITE_SLOT_BYTES = (size_t)LITE_LANES * (LANE_ROWS + 1) * 16u;
constexpr size_t COMPACT_LDS = 160 * 1024 - 1024;
// lanes of the global-memory chain's workgroup
constexpr int BIG_GLOBAL_THREADS = 1024;

// one sub-step's K iterations; one workgroup per world.  When the bodies the
// items touch fit LDS, their state is gathered there under compact indices,
// the chain runs on LDS, and the state is scattered back; otherwise the
// chain runs on global memory.  PART 0 takes only the worlds of the first
// kind, 1 only those of the second (launched with more lanes: a lane owns
// fewer positions, so fewer of its items wait behind one whose operands are
// not ready, and the chain's loads and polls, all to L2, overlap more).
template <int MAXT, bool WARM = false, int PART = 0>
__global__ __launch_bounds__(MAXT) void k_big_chain(StepDev S, uint32_t lds_nmax, uint32_t iters, float sub_dt)
{
    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ uint32_t s_abort;
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = S.w_order[blockIdx.x];
    if (S.w_status[w] != 0) return;
    const WorldDev& D = S.worlds[w];
    if (!big_world(D, lds_nmax)) return;
    if (threadIdx.x == 0) s_abort = 0u;
    __syncthreads();
    const uint32_t ibase = S.w_item_base[w], nitems = S.w_item_base[w + 1] - ibase;
    const uint32_t base = D.body_base;
    uint4* recs = S.recs + (size_t)ibase * 8;
    const uint32_t nc = S.w_nc[w];
    const bool compact = (size_t)nc * COMPACT_BYTES <= COMPACT_LDS;
    // the light compact state (momentum and counters, 16 B per touched body,
    // beside lane slots of six rows) for the worlds the full one does not fit
    const bool lite = !compact && nc < 0xffffu && (size_t)nc * LITE_BYTES + LITE_SLOT_BYTES <= COMPACT_LDS;
    static_assert(PART != 0 || MAXT == LITE_LANES, "the light chain's slots are sized for the compact launch");
    if ((compact || lite) != (PART == 0)) return;   // uniform: the other launch takes this world
    // the lane-slot chain when the slots fit beside the compact state
    constexpr size_t LANE_BYTES = (size_t)MAXT * LANE_WORDS * 4u;
    const bool lanes = (size_t)nc * COMPACT_BYTES + LANE_BYTES <= COMPACT_LDS;
    if constexpr (PART == 0) {
        if (lite) {
            LiteState st;
            st.mx = (float*)(smem + LITE_SLOT_BYTES / 4);
            st.my = st.mx + nc;
            st.L = st.my + nc;
            st.cnt = (uint32_t*)(st.L + nc);
            st.im = PairValue{0u, 0.f, 0.f};
            st.ii = PairValue{0u, 0.f, 0.f};
            st.deg = DegLookup{S.deg, S.cinv, base};
            st.rmass = S.rmass + ibase;
            for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
                const uint32_t i = S.cinv[base + c];
                st.mx[c] = D.mx[i]; st.my[c] = D.my[i]; st.L[c] = D.L[i];
                st.cnt[c] = 0u;
            }
            __syncthreads();
            item_chain_lanes<true, false, WARM>(S.w_status, st, recs, reinterpret_cast<uint4*>(smem), nitems, iters, sub_dt, w,
                                                s_abort);
            __syncthreads();
            for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
                const uint32_t i = S.cinv[base + c];
                D.mx[i] = st.mx[c]; D.my[i] = st.my[c]; D.L[i] = st.L[c];
            }
            return;
        }
        SolverState<true> st;
        st.mx = (float*)(smem + (lanes ? MAXT * LANE_WORDS : 0u));
        st.my = st.mx + nc;
        st.L = st.my + nc;
        st.im = st.L + nc;
        st.ii = st.im + nc;
        st.cnt = (uint32_t*)(st.ii + nc);
        st.deg = st.cnt + nc;
        for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
            const uint32_t i = S.cinv[base + c];
            st.mx[c] = D.mx[i]; st.my[c] = D.my[i]; st.L[c] = D.L[i];
            st.im[c] = S.g_im[base + i]; st.ii[c] = S.g_ii[base + i];
            st.cnt[c] = 0u;
            st.deg[c] = S.deg[base + i];
        }
        __syncthreads();
        if (lanes)
            item_chain_lanes<true, false, WARM>(S.w_status, st, recs, reinterpret_cast<uint4*>(smem), nitems, iters, sub_dt, w,
                                                s_abort);
        else item_chain<true, true, WARM>(S, D, st, recs, nitems, iters, sub_dt, w, s_abort);
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
            const uint32_t i = S.cinv[base + c];
            D.mx[i] = st.mx[c]; D.my[i] = st.my[c]; D.L[i] = st.L[c];
        }
    } else {
        (void)lanes;
        // body state in global memory; the lane slots take the start of LDS,
        // then the counters under the compact indices when they fit (the
        // compact indices exist below 65,535 touched bodies), else the
        // counters stay in global memory too
        const SolverState<false> st = global_state(S, D);
        uint4* slots = reinterpret_cast<uint4*>(smem);
        if (nc < 0xffffu && (size_t)nc * 4u + LANE_BYTES <= COMPACT_LDS) {
            uint32_t* lcnt = smem + (size_t)MAXT * LANE_WORDS;
            for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) lcnt[c] = 0u;
            __syncthreads();
            item_chain_lanes<false, true, WARM, true>(S.w_status, st, recs, slots, nitems, iters, sub_dt, w, s_abort, lcnt);
        } else {
            item_chain_lanes<false, true, WARM>(S.w_status, st, recs, slots, nitems, iters, sub_dt, w, s_abort);
        }
    }
}

// ------------------------------------------------------ register-slot solver
// For worlds with at most NSLOT * MAXT solver items (and < 65,535 bodies) each
// thread keeps all of its items -- execution positions tid, tid + MAXT, ... --
// in registers for the whole step, and the body state a chain link reads
// (momentum, inverse masses, pose, cos / sin of the angle, inertia, torque)
// sits in LDS, so the dependency-chain loop issues no global memory operation:
// a wave's vector-memory wait is wave-wide, and any load inside the loop would
// put HBM latency on every chain link.
struct Slot {
    uint32_t flags;          // RF_* bits | joint type << 24 | SLOT_USED; 0 = unused
    uint32_t pair;           // manifold: ref | inc << 16; joint: r0 | r1 << 16 (0xffff = none)
    uint32_t ranks;          // rank at r0 | rank at r1 << 16
    uint32_t idx;            // manifold index / joint body slots (s1 | s2 << 16)
    float nx, ny, fric;
    float p[2][7];           // manifold: r1x r1y r2x r2y depth mass_n mass_t per point
                             // joint: pmax pmin beta r1x r1y r2x r2y | target tx ty
    float acc[4];            // pn0 pt0 pn1 pt1
    uint32_t degs;           // items at r0 | items at r1 << 16 (ordering bodies)
    float bias[2];           // per point: Baumgarte bias of this sub-step, -1 when depth >= 0
};
constexpr uint32_t SLOT_USED = 0x80000000u;

struct SlotState {
    float *mx, *my, *L, *im, *ii;
    uint32_t *cnt, *deg;
    float *px, *py, *c, *s, *I, *tq;
};

__device__ __forceinline__ void slot_from_rec(Slot& z, const Rec& r)
{
    z.flags = r.flags | SLOT_USED;
    z.pair = (r.a & 0xffffu) | (r.b << 16);
    z.ranks = (r.rk0 & 0xffffu) | (r.rk1 << 16);
    z.idx = r.idx;
    z.nx = r.nx; z.ny = r.ny; z.fric = r.fric;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int f = 0; f < 7; ++f) z.p[p][f] = r.p[p][f];
#pragma unroll
    for (int f = 0; f < 4; ++f) z.acc[f] = r.acc[f];
}

__device__ __forceinline__ void rec_from_slot(Rec& r, const Slot& z)
{
    r.flags = z.flags & ~SLOT_USED;
    uint32_t a = z.pair & 0xffffu, b = z.pair >> 16;
    r.a = (a == 0xffffu) ? NONE : a;
    r.b = (b == 0xffffu) ? NONE : b;
    r.rk0 = z.ranks & 0xffffu; r.rk1 = z.ranks >> 16;
    r.idx = z.idx;
    r.nx = z.nx; r.ny = z.ny; r.fric = z.fric;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int f = 0; f < 7; ++f) r.p[p][f] = z.p[p][f];
#pragma unroll
    for (int f = 0; f < 4; ++f) r.acc[f] = z.acc[f];
}

// joint slot: parameters from the world's joint record
__device__ __forceinline__ void slot_joint(Slot& z, const Joint& J, const WorldDev& D)
{
    z.flags |= (uint32_t)J.type << 24;
    if (D.flags[J.s1] & BF_STATIC) z.flags |= RF_S1;
    z.idx = ((uint32_t)J.s1 & 0xffffu) | ((uint32_t)(J.s2 < 0 ? 0xffff : J.s2) << 16);
    z.p[0][0] = J.pmax; z.p[0][1] = J.pmin; z.p[0][2] = J.beta;
    z.p[0][3] = J.r1x; z.p[0][4] = J.r1y; z.p[0][5] = J.r2x; z.p[0][6] = J.r2y;
    z.p[1][0] = J.target; z.p[1][1] = J.tx; z.p[1][2] = J.ty;
}

// Constraint.solve for a joint slot: pose, cos / sin, inertia and torque from LDS
__device__ __forceinline__ void slot_solve_joint(const Slot& z, const SlotState& st, float dt)
{
    const uint32_t a = z.idx & 0xffffu, b = z.idx >> 16;
    JointIn j;
    j.type = (z.flags >> 24) & 0x7fu;
    j.stat1 = (z.flags & RF_S1) != 0;
    j.pmax = z.p[0][0]; j.pmin = z.p[0][1]; j.beta = z.p[0][2];
    j.r1x = z.p[0][3]; j.r1y = z.p[0][4]; j.r2x = z.p[0][5]; j.r2y = z.p[0][6];
    j.target = z.p[1][0]; j.tx = z.p[1][1]; j.ty = z.p[1][2];
    j.pa = mk(st.px[a], st.py[a]);
    j.ca = st.c[a]; j.sa = st.s[a];
    j.I1 = st.I[a]; j.tq1 = st.tq[a];
    if (j.type <= 1) {
        j.pb = mk(st.px[b], st.py[b]);
        j.cb = st.c[b]; j.sb = st.s[b];
    } else {
        j.pb = mk(0.f, 0.f);
        j.cb = j.sb = 0.f;
    }
    joint_apply(j, st, a, b, dt);
}

// ordering bodies of a slot (-1: none) and the Baumgarte biases of a sub-step
__device__ __forceinline__ void slot_ordering(const Slot& z, int32_t& r0, int32_t& r1)
{
    uint32_t a = z.pair & 0xffffu, b = z.pair >> 16;
    if (z.flags & RF_JOINT) {
        r0 = a == 0xffffu ? -1 : (int32_t)a;
        r1 = b == 0xffffu ? -1 : (int32_t)b;
    } else {
        r0 = (z.flags & RF_W1) ? (int32_t)a : -1;
        r1 = (z.flags & RF_W2) ? (int32_t)b : -1;
    }
}

__device__ __forceinline__ void slot_biases(Slot& z, float dt)
{
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        float depth = z.p[p][4];
        z.bias[p] = depth >= 0.f ? -1.0f : BAUMGARTE * maxn(0.0f, (-depth) - BAUMGARTE_SLOP) / dt;
    }
}

// LOCKSTEP: a wave takes its lanes' slot z together, z = 0, 1, ... in turn
// (each slot's registers addressed directly, no per-lane selection of the
// current slot), instead of every lane walking its own slots independently.
// Both orders are deadlock-free: a wave works on positions in increasing
// (iteration, slot, lane) order, the order of the dependencies.
// One world's step on the register-slot path, one workgroup of MAXT threads
// (nitems <= NSLOT * MAXT; lds: slot_lds_bytes(n) bytes; degs: the world's
// item counts per body).
template <int NSLOT, int MAXT, bool LOCKSTEP>
__device__ __forceinline__ void slots_world(const StepDev& S, const WorldDev& D, uint32_t w, uint32_t ibase,
                                            uint32_t nitems, uint32_t mbeg, const uint32_t* degs, float dt,
                                            uint32_t sub_steps, uint32_t iters, uint32_t* lds, uint32_t& s_abort)
{
    const int n = D.nbodies;
    if (threadIdx.x == 0) s_abort = 0u;
    const uint32_t base = D.body_base;
    const uint32_t nj = D.njoints;
    const float fsub = (float)sub_steps;
    const float sub_dt = dt / fsub;

    SlotState st;
    st.mx = (float*)lds;
    st.my = st.mx + n;
    st.L = st.my + n;
    st.im = st.L + n;
    st.ii = st.im + n;
    st.cnt = (uint32_t*)(st.ii + n);
    st.deg = st.cnt + n;
    st.px = (float*)(st.deg + n);
    st.py = st.px + n;
    st.c = st.py + n;
    st.s = st.c + n;
    st.I = st.s + n;
    st.tq = st.I + n;
    for (int i = threadIdx.x; i < n; i += MAXT) {
        bool stat = (D.flags[i] & BF_STATIC) != 0;
        st.mx[i] = D.mx[i]; st.my[i] = D.my[i]; st.L[i] = D.L[i];
        st.im[i] = stat ? 0.0f : 1.0f / D.mass[i];
        st.ii[i] = stat ? 0.0f : 1.0f / D.inertia[i];
        st.deg[i] = degs[i];
        st.I[i] = D.inertia[i];
    }
    __syncthreads();
    Slot sl[NSLOT];
    uint32_t nslot = 0;
#pragma unroll
    for (int z = 0; z < NSLOT; ++z) {
        sl[z].flags = 0u;
        uint32_t pos = threadIdx.x + (uint32_t)z * MAXT;
        if (pos < nitems) {
            Rec r;
            rec_build(r, S, D, st, pos, ibase, nj, mbeg, base);
            slot_from_rec(sl[z], r);
            if (r.flags & RF_JOINT) slot_joint(sl[z], D.joints[r.idx], D);
            int32_t q0, q1;
            slot_ordering(sl[z], q0, q1);
            sl[z].degs = (q0 >= 0 ? st.deg[q0] : 0u) | ((q1 >= 0 ? st.deg[q1] : 0u) << 16);
            nslot = (uint32_t)z + 1u;
        }
    }

    for (uint32_t s = 0; s < sub_steps; ++s) {
        // forces (lib.zig:200-205), AABB refresh + velocity integration (lib.zig:207-216)
        for (int i = threadIdx.x; i < n; i += MAXT) {
            uint32_t f = D.flags[i];
            bool stat = (f & BF_STATIC) != 0;
            float fx = D.fx[i], fy = D.fy[i];
            float px = D.px[i], py = D.py[i], a = D.ang[i];
            for (int g = 0; g < D.ngens; ++g) {
                const Gen& G = D.gens[g];
                if (stat) continue;
                if (G.type == 0) {
                    fx += G.gx * D.mass[i];
                    fy += G.gy * D.mass[i];
                } else if (G.slot == i) {
                    float inv_m = 1.0f / D.mass[i];
                    v2 v = scale2(mk(st.mx[i], st.my[i]), inv_m);
                    v2 d = sub2(mk(px, py), mk(G.ax, G.ay));
                    fx += (-G.k) * d.x - G.c * v.x;
                    fy += (-G.k) * d.y - G.c * v.y;
                }
            }
            // cos / sin of the angle: offset joints read them in the solve
            float c = 0.0f, sn = 0.0f;
            if (nj || (s + 1 == sub_steps && (f & BF_RECT))) { c = cosf_r(a); sn = sinf_r(a); }
            if (s + 1 == sub_steps) {
                float hw, hh;
                if (f & BF_RECT) {
                    float x = D.hw[i], y = D.hh[i];
                    v2 lv[4] = {mk(-x, -y), mk(-x, y), mk(x, y), mk(x, -y)};
                    hw = 0.f; hh = 0.f;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v2 r = rotcs(lv[k], c, sn);
                        if (r.x > hw) hw = r.x;
                        if (r.y > hh) hh = r.y;
                    }
                } else {
                    hw = D.rad[i]; hh = D.rad[i];
                }
                D.ax[i] = px; D.ay[i] = py; D.ahw[i] = hw; D.ahh[i] = hh;
            }
            if (!stat) {
                st.mx[i] += fx * sub_dt;
                st.my[i] += fy * sub_dt;
                st.L[i] += D.tq[i] * sub_dt;
            }
            D.fx[i] = fx; D.fy[i] = fy;
            st.px[i] = px; st.py[i] = py; st.c[i] = c; st.s[i] = sn;
            st.tq[i] = D.tq[i];
            st.cnt[i] = 0u;
        }
        __syncthreads();
        // updateTGSDepth (collision.zig:73-100)
#pragma unroll
        for (int z = 0; z < NSLOT; ++z) {
            uint32_t fl = sl[z].flags;
            if (!fl || (fl & RF_JOINT)) continue;
            if (fl & RF_TGS) {
                Rec r;
                rec_from_slot(r, sl[z]);
                rec_tgs(r, S, D, st, true);
                slot_from_rec(sl[z], r);
            } else if (s + 1 == sub_steps) {
                S.m_prev1[sl[z].idx] = D.ang[sl[z].pair & 0xffffu];
                S.m_prev2[sl[z].idx] = D.ang[sl[z].pair >> 16];
            }
            slot_biases(sl[z], sub_dt);
        }
        if constexpr (LOCKSTEP) {
            // lockstep over slots: see the kernel's comment; reads as in the
            // independent walk below
            for (uint32_t k = 0; k < iters; ++k) {
#pragma unroll
                for (int z = 0; z < NSLOT; ++z) {
                    Slot& it = sl[z];
                    bool done = !it.flags;
                    int32_t r0 = -1, r1 = -1;
                    slot_ordering(it, r0, r1);
                    const uint32_t t0 = k * (it.degs & 0xffffu) + (it.ranks & 0xffffu);
                    const uint32_t t1 = k * (it.degs >> 16) + (it.ranks >> 16);
                    const bool contact = !(it.flags & RF_JOINT);
                    uint32_t spins = 0;
                    while (!done) {
                        uint32_t c0 = r0 >= 0 ? __hip_atomic_load(&st.cnt[r0], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WORKGROUP)
                                              : t0;
                        uint32_t c1 = r1 >= 0 ? __hip_atomic_load(&st.cnt[r1], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WORKGROUP)
                                              : t1;
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        BodyState b1{}, b2{};
                        if (contact) {
                            b1 = body_state(st, it.pair & 0xffffu);
                            b2 = body_state(st, it.pair >> 16);
                        }
                        const bool ready = c0 == t0 && c1 == t1;
                        if (ready) {
                            if (!contact) slot_solve_joint(it, st, sub_dt);
                            else contact_impulses(it, it.pair & 0xffffu, it.pair >> 16, b1, b2, st);
                            __atomic_signal_fence(__ATOMIC_SEQ_CST);
                            if (r0 >= 0)
                                __hip_atomic_store(&st.cnt[r0], t0 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (r1 >= 0)
                                __hip_atomic_store(&st.cnt[r1], t1 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            done = true;
                        } else {
                            if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                                done = true;
                            } else if (++spins > SPIN_LIMIT) {
                                __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                S.w_status[w] = 5;
                                done = true;
                            }
                        }
                        if (__ballot(ready) == 0ull) __builtin_amdgcn_s_sleep(1);
                    }
                }
            }
        } else
        // K iterations as an ordered dataflow over register-resident items.  A
        // link's counters and its bodies' state are read together: LDS serves
        // one workgroup's operations in order, so state read after a counter
        // that shows the item ready is the state its predecessor released
        // (values read for an item that is not ready are discarded)
        {
            uint32_t k = nslot ? 0u : iters, cz = 0, spins = 0;
            bool have = false;
            Slot cur;
            int32_t r0 = -1, r1 = -1;
            uint32_t t0 = 0, t1 = 0;
            while (k < iters) {
                if (!have) {
#pragma unroll
                    for (int z = 0; z < NSLOT; ++z)
                        if ((uint32_t)z == cz) cur = sl[z];
                    have = true;
                    slot_ordering(cur, r0, r1);
                    t0 = k * (cur.degs & 0xffffu) + (cur.ranks & 0xffffu);
                    t1 = k * (cur.degs >> 16) + (cur.ranks >> 16);
                }
                uint32_t c0 = r0 >= 0 ? __hip_atomic_load(&st.cnt[r0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                      : t0;
                uint32_t c1 = r1 >= 0 ? __hip_atomic_load(&st.cnt[r1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                      : t1;
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                const bool contact = !(cur.flags & RF_JOINT);
                BodyState b1{}, b2{};
                if (contact) {
                    b1 = body_state(st, cur.pair & 0xffffu);
                    b2 = body_state(st, cur.pair >> 16);
                }
                const bool ready = c0 == t0 && c1 == t1;
                if (ready) {
                    // a wave with a ready link issues ahead of the waves that poll
                    __builtin_amdgcn_s_setprio(2);
                    if (!contact) {
                        slot_solve_joint(cur, st, sub_dt);
                    } else {
                        contact_impulses(cur, cur.pair & 0xffffu, cur.pair >> 16, b1, b2, st);
#pragma unroll
                        for (int z = 0; z < NSLOT; ++z)
                            if ((uint32_t)z == cz)
#pragma unroll
                                for (int f = 0; f < 4; ++f) sl[z].acc[f] = cur.acc[f];
                    }
                    // the state writes above reach LDS before the counters (in order)
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (r0 >= 0) __hip_atomic_store(&st.cnt[r0], t0 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (r1 >= 0) __hip_atomic_store(&st.cnt[r1], t1 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __builtin_amdgcn_s_setprio(0);
                    have = false;
                    spins = 0;
                    if (++cz == nslot) { cz = 0; ++k; }
                } else {
                    // every item's predecessors precede it, so some item is always
                    // ready; a stall this long means corrupt ordering data: stop
                    if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                    if (++spins > SPIN_LIMIT) {
                        __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        S.w_status[w] = 5;
                        break;
                    }
                }
                // a wave with no lane ready yields its SIMD
                if (__ballot(ready) == 0ull) __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        // position integration (lib.zig:238-249)
        for (int i = threadIdx.x; i < n; i += MAXT) {
            if (D.flags[i] & BF_STATIC) continue;
            float qd = sub_dt / D.mass[i];
            D.px[i] += st.mx[i] * qd;
            D.py[i] += st.my[i] * qd;
            D.ang[i] += st.L[i] * sub_dt / D.inertia[i];
            D.fx[i] = 0.0f; D.fy[i] = 0.0f;
            D.tq[i] = 0.0f;
        }
        __syncthreads();
    }
#pragma unroll
    for (int z = 0; z < NSLOT; ++z) {
        if (!sl[z].flags || (sl[z].flags & RF_JOINT)) continue;
        Rec r;
        rec_from_slot(r, sl[z]);
        rec_writeback(r, S);
    }
    for (int i = threadIdx.x; i < n; i += MAXT) {
        D.mx[i] = st.mx[i]; D.my[i] = st.my[i]; D.L[i] = st.L[i];
    }
}

// the register-slot kernel: the worlds solve_route gives it, one workgroup each
template <int NSLOT, int MAXT, int WAVES, bool LOCKSTEP = false>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WAVES)))
void k_solve_slots(StepDev S, float dt, uint32_t sub_steps, uint32_t iters, uint32_t slot_nmax)
{
    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ uint32_t s_abort;
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = S.w_order[blockIdx.x];
    if (S.w_status[w] != 0) return;
    const uint32_t ibase = S.w_item_base[w];
    const uint32_t nitems = S.w_item_base[w + 1] - ibase;
    const WorldDev& D = S.worlds[w];
    if (solve_route((uint32_t)D.nbodies, nitems, (uint32_t)NSLOT * MAXT, slot_nmax, 0u, D.warm > 0.0f) != 0) return;
    slots_world<NSLOT, MAXT, LOCKSTEP>(S, D, w, ibase, nitems, S.w_man_begin[w], S.deg + D.body_base, dt, sub_steps,
                                       iters, smem, s_abort);
}

// ------------------------------------------------------ fused world steps
// Worlds of at most FUSED_NMAX bodies in the reference order can take whole
// steps in one workgroup: broadphase, narrowphase, schedule and solve of the
// world with workgroup barriers between the phases instead of kernel
// boundaries, and all the steps of a call in one launch (worlds are
// independent, so nothing crosses workgroups).  Each phase computes for its
// world exactly what the grid-wide kernels above compute: the same
// per-element functions (prep_body, narrow_pair, man_build_one, the solve
// paths) over the world's own scratch slices (WorldScratch), with the
// per-body and per-bucket arrays in LDS.

// The stale AABBs (aabb.zig:6-8) and flags of a world's bodies in LDS,
// for the pair filters of the fused broadphase
struct PairCache {
    float *ax, *ay, *ahw, *ahh;
    uint32_t* flags;
};
constexpr uint32_t PAIR_CACHE_WORDS = 5;   // per body

__device__ __forceinline__ PairCache pair_cache(uint32_t* p, uint32_t n)
{
    return PairCache{(float*)p, (float*)p + n, (float*)p + 2 * n, (float*)p + 3 * n, p + 4 * n};
}

// updateManifolds' pair filters that need no more than the cache: both
// bodies static (lib.zig:273), padded AABB overlap (lib.zig:282)
__device__ __forceinline__ uint32_t pair_passes(const PairCache& P, uint32_t a, uint32_t b)
{
    if ((P.flags[a] & BF_STATIC) && (P.flags[b] & BF_STATIC)) return 0u;
    return aabb_hit(P, a, b) ? 1u : 0u;
}

// The raw candidates of one cell entry of body a: the distinct members of
// its bucket above a (s[0, c), members in descending order, a body's repeats
// adjacent) that pass pair_passes, as the serial walk
//   for (j = 0; j < c && s[j] > a; ++j)
//       if (j == 0 || s[j] != s[j - 1]) { ok = pair_passes(P, a, s[j]); ... }
// counts them: their number, and the passing positions below 32 as a mask.
// Four members at a time, so the filters' LDS loads of four bodies are in
// flight together instead of one body's at a time.
__device__ __forceinline__ void entry_candidates(const PairCache& P, uint32_t a, const uint32_t* s, uint32_t c,
                                                 uint32_t& q_out, uint32_t& mask_out)
{
    uint32_t q = 0, mask = 0, prev = 0;
    for (uint32_t j = 0; j < c; j += 4) {
        uint32_t m[4];
        bool live[4];
        bool stop = false;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) m[k] = j + k < c ? s[j + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const bool above = j + k < c && m[k] > a;
            const uint32_t before = k ? m[k - 1] : prev;
            live[k] = !stop && above && (j + k == 0 || m[k] != before);
            stop = stop || !above;
        }
        uint32_t ok[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) ok[k] = pair_passes(P, a, m[k]);
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const bool pass = live[k] && ok[k] != 0u;
            q += pass;
            if (pass && j + k < 32u) mask |= 1u << (j + k);
        }
        if (stop) break;
        prev = m[3];
    }
    q_out = q;
    mask_out = mask;
}

// LDS of the fused kernel for a world of n bodies and tb hash buckets
struct FusedLds {
    uint32_t* coff;    // [n+1] cells per body -> entry offsets
    float *cc, *cs;    // [n] cos / sin of the angle at step start
    uint32_t* boff;    // [tb+1] bucket occupancy -> bucket offsets
    uint32_t* bfill;   // [tb]
    uint32_t* dcnt;    // [n] solver items per ordering body
    uint32_t* doff;    // [n+1] -> list offsets
    uint32_t* dfill;   // [n]
    uint32_t* solve;   // the solve path's body state
    uint4* lanes;      // the record chain's lane slots (over the front part)
};

// front part of a world's fused LDS, in words
__host__ __device__ constexpr uint32_t fused_front_words(uint32_t n, uint32_t tb) { return 6u * n + 2u * tb + 3u; }

// The front part is dead during the solve's chains (the item counts per body
// are read before them), so the record chain's lane slots (lane_words, at
// offset 0) overlay it; the solve's part starts after the larger of the two.
__device__ __forceinline__ FusedLds fused_carve(uint32_t* lds, uint32_t n, uint32_t tb, uint32_t lane_words)
{
    FusedLds L;
    uint32_t* p = lds;
    L.lanes = reinterpret_cast<uint4*>(lds);
    L.coff = p; p += n + 1;
    L.cc = (float*)p; p += n;
    L.cs = (float*)p; p += n;
    L.boff = p; p += tb + 1;
    L.bfill = p; p += tb;
    L.dcnt = p; p += n;
    L.doff = p; p += n + 1;
    L.dfill = p; p += n;
    L.solve = lds + umax(fused_front_words(n, tb), lane_words);
    return L;
}

// words per body after the front part: the solve's LDS body state (register
// slots 13, records 7), which also holds the pair cache and, when they fit,
// the bucket arrays until the solve
__host__ __device__ constexpr uint32_t fused_solve_words(bool slots) { return slots ? 13u : 7u; }

size_t fused_lds_bytes(uint32_t n, uint32_t tb, bool slots, uint32_t threads)
{
    const size_t front = std::max<size_t>(fused_front_words(n, tb), (size_t)threads * LANE_WORDS);
    return (front + (size_t)fused_solve_words(slots) * n) * 4u;
}

// The fused kernel's narrowphase test of one candidate, as a call: out of
// line, its registers are allocated apart from the step kernel's, which
// spills less (45 -> 31 VGPRs for k_fused<0, 192, 3>) and ran the 240-step
// launch 2 % faster.  It takes only pointers and values (no reference to the
// kernel's argument block, which would be copied to scratch).
__device__ __attribute__((noinline)) bool narrow_pair_call(const WorldDev& D, uint32_t sa, uint32_t sb, uint32_t k,
                                                         const uint32_t* ba, const uint32_t* bb, uint32_t nb,
                                                         float ca, float sa_sin, float cb, float sb_sin,
                                                         uint32_t& info, float& nx, float& ny)
{
    return narrow_pair<true>(D, D, sa, sb, k, ba, bb, nb, ca, sa_sin, cb, sb_sin, info, nx, ny);
}

// exclusive scan of a[0, len) in place (LDS or global), a[len] = the total,
// which is returned; called by the whole workgroup after a barrier, ends
// with one
__device__ __forceinline__ uint32_t group_scan(uint32_t* a, uint32_t len, uint32_t* sm)
{
    const uint32_t T = blockDim.x, t = threadIdx.x;
    const uint32_t chunk = (len + T - 1) / T;
    const uint32_t lo = umin(t * chunk, len), hi = umin(lo + chunk, len);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += a[i];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, sm, &total);
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
    if (t == 0) a[len] = total;
    __syncthreads();
    return total;
}

// A world's slice of family k (0 cell entries, 1 raw candidates, 2
// manifolds, 3 solver items) is too small for `need` elements: the world
// moves to a larger slice taken from the free end of the arrays (S.fpool).
// False when the arrays are full; the world then stops (fused_overflow).
// Called by the whole workgroup; the slice table entry is updated for the host.
// Inlined: as a call it takes the kernel's StepDev by reference, and the
// kernel then copies its whole argument block to scratch at entry.
__device__ __forceinline__ bool fused_grow(const StepDev& S, uint32_t w, WorldScratch& X, int k, uint32_t need,
                                           uint32_t& s_slot)
{
    const uint32_t cap = need + need / 2 + 64;
    const uint32_t take = k < 2 ? cap + 1 : cap;   // entry and candidate slices end with a scan total
    if (threadIdx.x == 0) {
        // claim [base, base + take) only if it fits: a refused request leaves
        // the free end for smaller ones
        const uint32_t lim = S.fpool[4 + k];
        uint32_t base = __hip_atomic_load(&S.fpool[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;) {
            if (base > lim || take > lim - base) { base = 0xffffffffu; break; }
            const uint32_t seen = atomicCAS(&S.fpool[k], base, base + take);
            if (seen == base) break;
            base = seen;
        }
        s_slot = base;
    }
    __syncthreads();
    const uint32_t base = s_slot;
    __syncthreads();
    if (base == 0xffffffffu) return false;
    if (k == 0) { X.e0 = base; X.ecap = cap; }
    else if (k == 1) { X.r0 = base; X.rcap = cap; }
    else if (k == 2) { X.m0 = base; X.mcap = cap; }
    else { X.i0 = base; X.icap = cap; }
    if (threadIdx.x == 0) S.wscr[w] = X;
    return true;
}

// A world's scratch slice is too small for this step and the arrays are
// full: record what it needs and stop the world before its state changes
// (the host grows the arrays and runs the remaining steps)
__device__ __forceinline__ void fused_overflow(const StepDev& S, uint32_t w, uint32_t e, uint32_t r, uint32_t m,
                                               uint32_t it)
{
    if (threadIdx.x == 0) {
        uint32_t* c = S.w_cnt + 8 * (size_t)w;
        c[4] = e; c[5] = r; c[6] = m; c[7] = it;
        S.w_status[w] = WS_OVERFLOW;
    }
}

// The solve of a fused step: the register-slot path when the step's items
// fit the slots, else the record path.  Inlined: a call would save and
// restore the step's registers around it every step (scratch traffic of the
// order of the step's own); WAVES sets the register budget of the shape.
template <int NSLOT, int MAXT, int WAVES, bool LOCKSTEP, bool WARM>
__device__ __forceinline__ void fused_solve(const StepDev& S, const WorldDev& D, uint32_t w, uint32_t ibase,
                                         uint32_t nitems, uint32_t mbeg, const uint32_t* degs, float dt,
                                         uint32_t sub_steps, uint32_t iters, uint32_t* lds, uint4* lanes,
                                         uint32_t* s_abort)
{
    if constexpr (NSLOT > 0) {
        if (nitems <= (uint32_t)NSLOT * MAXT && !(WARM && D.warm > 0.0f)) {
            slots_world<NSLOT, MAXT, LOCKSTEP>(S, D, w, ibase, nitems, mbeg, degs, dt, sub_steps, iters, lds,
                                               *s_abort);
            return;
        }
    }
    rec_world<true, WARM>(S, D, w, ibase, nitems, mbeg, degs, dt, sub_steps, iters, lds, lanes, *s_abort);
}

// One workgroup per world (S.w_order), up to nsteps steps per world counted
// in S.w_done, so a relaunch after an overflow resumes each world where it
// stopped.  The solve takes the register-slot path when the step's items fit
// NSLOT * MAXT slots, else the record path; both keep body state in LDS.
template <int NSLOT, int MAXT, int WAVES, bool LOCKSTEP, bool WARM = false>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WAVES)))
void k_fused(StepDev S, float dt, uint32_t sub_steps, uint32_t iters, uint32_t nsteps)
{
    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ uint32_t s_sm[16];
    __shared__ uint32_t s_err, s_abort, s_slot;
    __shared__ uint32_t s_lvl[3];
    const uint32_t w = S.w_order[blockIdx.x];
    const WorldDev& D = S.worlds[w];
    WorldScratch X = S.wscr[w];
    const uint32_t n = (uint32_t)D.nbodies, tb = D.table, base = D.body_base, nj = (uint32_t)D.njoints;
    const uint32_t t = threadIdx.x;
    constexpr uint32_t T = MAXT;
    const FusedLds L = fused_carve(smem, n, tb, T * LANE_WORDS);
    uint32_t* cnt = S.w_cnt + 8 * (size_t)w;
    uint64_t* ts = S.w_tstamp ? S.w_tstamp + (size_t)w * (S.tcap + 1) : nullptr;
    uint32_t step = S.w_done[w];
    if (ts && t == 0 && step < nsteps && step <= S.tcap) ts[step] = wall_clock64();
    // phase times (profiling): thread 0 adds the wall clock between marks
    uint64_t* prof = (S.w_prof && t == 0) ? S.w_prof + (size_t)w * FP_N : nullptr;
    uint64_t tmark = prof ? wall_clock64() : 0;
    auto mark = [&](int ph) {
        if (prof) {
            const uint64_t now = wall_clock64();
            prof[ph] += now - tmark;
            tmark = now;
        }
    };
    for (; step < nsteps; ++step) {
        if (S.w_status[w] != WS_OK) break;      // uniform: read before the barrier below
        // ---- broadphase: per body (k_prep)
        if (t == 0) s_err = 0u;
        __syncthreads();
        // stale AABBs and flags in LDS for the pair filters below (the solve's
        // words are free until the solve)
        const PairCache P = pair_cache(L.solve, n);
        for (uint32_t i = t; i < n; i += T) {
            const Prep o = prep_body(D, i);
            P.ax[i] = D.ax[i]; P.ay[i] = D.ay[i]; P.ahw[i] = D.ahw[i]; P.ahh[i] = D.ahh[i];
            P.flags[i] = D.flags[i];
            L.cc[i] = o.c;
            L.cs[i] = o.s;
            S.sflag[base + i] = o.res;
            if (o.err) atomicOr(&s_err, o.err == WS_NONFINITE ? 2u : 1u);
            if (!o.err && o.ncell) { S.cx0[base + i] = o.x0; S.cy0[base + i] = o.y0; S.cnx[base + i] = o.nx; }
            L.coff[i] = o.ncell;
        }
        __syncthreads();
        if (s_err) {
            if (t == 0) S.w_status[w] = (s_err & 2u) ? WS_NONFINITE : WS_CELLS;
            break;
        }
        const uint32_t E = group_scan(L.coff, n, s_sm);
        mark(FP_PREP);
        if (E > X.ecap && !fused_grow(S, w, X, 0, E, s_slot)) { fused_overflow(S, w, E, 0, 0, 0); break; }
        uint32_t* ent_body = S.ent_body + X.e0;
        uint32_t* ecnt = S.ent_cnt + X.e0;      // raw candidates per entry -> offsets
        // the entries' buckets and the bucket members are read in serial
        // loops: in LDS after the pair cache when they fit the solve's words
        // (3 E <= 8 n; E is about 1.5 n for bodies up to a cell wide)
        const bool e_lds = 3 * (size_t)E <= (size_t)(fused_solve_words(NSLOT > 0) - PAIR_CACHE_WORDS) * n;
        uint32_t* e_free = L.solve + PAIR_CACHE_WORDS * n;
        uint32_t* ent_bkt = e_lds ? e_free : S.ent_bkt + X.e0;
        uint32_t* items = e_lds ? e_free + E : S.bkt_items + X.e0;
        uint32_t* sorted = e_lds ? e_free + 2 * E : S.bkt_sorted + X.e0;
        for (uint32_t b = t; b <= tb; b += T) {
            L.boff[b] = 0u;
            if (b < tb) L.bfill[b] = 0u;
        }
        __syncthreads();
        // cell entries and bucket occupancy (k_entries)
        for (uint32_t e = t; e < E; e += T) {
            uint32_t i = upper_idx(L.coff, n, e);
            uint32_t k = e - L.coff[i];
            uint32_t nx = S.cnx[base + i];
            int64_t yi = S.cy0[base + i] + (int64_t)(k / nx);
            int64_t xi = S.cx0[base + i] + (int64_t)(k % nx);
            uint32_t b = (uint32_t)cell_hash((uint64_t)tb, xi, yi);
            ent_body[e] = i;
            ent_bkt[e] = b;
            atomicAdd(&L.boff[b], 1u);
        }
        __syncthreads();
        group_scan(L.boff, tb, s_sm);
        // bucket members (k_bucket_fill), then in descending slot with a
        // body's repeats adjacent (k_bucket_rank)
        for (uint32_t e = t; e < E; e += T) {
            uint32_t b = ent_bkt[e];
            uint32_t pos = atomicAdd(&L.bfill[b], 1u);
            items[L.boff[b] + pos] = ent_body[e];
        }
        __syncthreads();
        for (uint32_t e = t; e < E; e += T) {
            uint32_t a = ent_body[e], b = ent_bkt[e];
            uint32_t o = L.boff[b], c = L.boff[b + 1] - o;
            uint32_t greater, same;
            bucket_rank_of(items + o, c, a, greater, same);
            uint32_t dup = 0;
            if (same > 1)
                for (uint32_t x = L.coff[a]; x < e; ++x) dup += ent_bkt[x] == b;
            sorted[o + greater + dup] = a;
        }
        __syncthreads();
        mark(FP_BUCKETS);
        // raw candidates per entry: distinct bucket members above the entry's
        // body (k_ent_cnt) that pass the both-static and AABB filters; the
        // passing members among the first 32 as a mask (in `items`, whose
        // bucket lists are no longer read)
        for (uint32_t e = t; e < E; e += T) {
            const uint32_t a = ent_body[e], b = ent_bkt[e];
            const uint32_t o = L.boff[b], c = L.boff[b + 1] - o;
            uint32_t q, mask;
            entry_candidates(P, a, sorted + o, c, q, mask);
            ecnt[e] = q;
            items[e] = mask;
        }
        __syncthreads();
        const uint32_t R = group_scan(ecnt, E, s_sm);
        if (R > X.rcap && !fused_grow(S, w, X, 1, R, s_slot)) { fused_overflow(S, w, E, R, 0, 0); break; }
        uint4* cand = S.cand + X.r0;
        uint32_t* cflag = S.cand_flag + X.r0;   // collides per candidate -> manifold offsets
        // raw candidates in query order (k_cand_emit)
        for (uint32_t e = t; e < E; e += T) {
            uint32_t a = ent_body[e], b = ent_bkt[e];
            uint32_t o = L.boff[b];
            uint32_t ro = ecnt[e], c = ecnt[e + 1] - ro;
            uint32_t k = e - L.coff[a];
            uint32_t q = 0;
            for (uint32_t mask = items[e]; mask; mask &= mask - 1u, ++q)
                cand[ro + q] = make_uint4(a, sorted[o + (uint32_t)__builtin_ctz(mask)], k, 0u);
            for (uint32_t j = 32; q < c; ++j) {
                const uint32_t bj = sorted[o + j];
                if (bj == sorted[o + j - 1] || !pair_passes(P, a, bj)) continue;
                cand[ro + q] = make_uint4(a, bj, k, 0u);
                ++q;
            }
        }
        __syncthreads();
        mark(FP_CANDS);
        // ---- narrowphase (k_narrow): the collision overwrites its candidate
        // (partner, face id | ref-is-a, normal)
        for (uint32_t q = t; q < R; q += T) {
            const uint4 cq = cand[q];
            const uint32_t a = cq.x, b = cq.y;
            uint32_t info = 0;
            float nx = 0.f, ny = 0.f;
            bool hit = narrow_pair_call(D, a, b, cq.z, ent_bkt + L.coff[a], ent_bkt + L.coff[b],
                                         L.coff[b + 1] - L.coff[b], L.cc[a], L.cs[a], L.cc[b], L.cs[b], info, nx, ny);
            cflag[q] = hit ? 1u : 0u;
            if (hit) cand[q] = make_uint4(a | (b << 16), info, __float_as_uint(nx), __float_as_uint(ny));
        }
        __syncthreads();
        const uint32_t M = group_scan(cflag, R, s_sm);
        mark(FP_NARROW);
        if ((M > X.mcap && !fused_grow(S, w, X, 2, M, s_slot)) ||
            (nj + M > X.icap && !fused_grow(S, w, X, 3, nj + M, s_slot))) {
            fused_overflow(S, w, E, R, M, nj + M);
            break;
        }
        uint32_t* lists = S.lists + 2 * (size_t)X.i0;
        // contact points (k_man_build), manifolds in discovery order
        for (uint32_t q = t; q < R; q += T) {
            const uint32_t mi = cflag[q];
            if (cflag[q + 1] == mi) continue;
            const uint4 x = cand[q];
            const uint32_t a = x.x & 0xffffu, b = x.x >> 16;
            man_build_one(S, D, D, X.m0 + mi, base, a, b, x.y, mk(__uint_as_float(x.z), __uint_as_float(x.w)), L.cc[a],
                          L.cs[a], L.cc[b], L.cs[b]);
        }
        // ---- schedule: distance / offset joints make both bodies ordering (k_joint_mark)
        for (uint32_t j = t; j < nj; j += T) {
            const Joint& J = D.joints[j];
            if (J.type <= 1) {
                if (J.s1 >= 0) atomicOr(&S.sflag[base + J.s1], SF_RESOURCE);
                if (J.s2 >= 0) atomicOr(&S.sflag[base + J.s2], SF_RESOURCE);
            }
        }
        for (uint32_t i = t; i < n; i += T) { L.dcnt[i] = 0u; L.dfill[i] = 0u; }
        __syncthreads();
        mark(FP_CONTACTS);
        const uint32_t nit = nj + M;
        // items: joints, then manifolds, with their ordering bodies (k_items_*)
        for (uint32_t j = t; j < nit; j += T) {
            int32_t r0 = -1, r1 = -1;
            if (j < nj) {
                const Joint& J = D.joints[j];
                r0 = J.s1 >= 0 ? (int32_t)(base + J.s1) : -1;
                if (J.type <= 1 && J.s2 >= 0 && J.s2 != J.s1) r1 = (int32_t)(base + J.s2);
            } else {
                const uint32_t m = X.m0 + (j - nj);
                const uint32_t ref = S.m_ref[m], inc = S.m_inc[m];
                r0 = (S.sflag[ref] & SF_RESOURCE) ? (int32_t)ref : -1;
                r1 = (S.sflag[inc] & SF_RESOURCE) ? (int32_t)inc : -1;
            }
            S.it_r0[X.i0 + j] = r0;
            S.it_r1[X.i0 + j] = r1;
            S.it_perm[X.i0 + j] = j;
            if (r0 >= 0) atomicAdd(&L.dcnt[r0 - (int32_t)base], 1u);
            if (r1 >= 0) atomicAdd(&L.dcnt[r1 - (int32_t)base], 1u);
        }
        __syncthreads();
        for (uint32_t i = t; i < n; i += T) L.doff[i] = L.dcnt[i];
        __syncthreads();
        group_scan(L.doff, n, s_sm);
        // per-body item lists (k_list_fill), ranks in item order (k_list_rank).
        // When the lists fit the front part of LDS (its cell, trig and bucket
        // arrays are dead after the contacts), they are filled there in any
        // order and each entry then finds its rank by counting the smaller
        // entries of its body's list (independent LDS reads) and goes to that
        // place of the global lists; else they are filled in global memory and
        // sorted per body.
        uint32_t* const ul = L.coff;
        const bool lds_lists = L.doff[n] <= 3u * n + 2u * tb + 2u;
        uint32_t* const fill = lds_lists ? ul : lists;
        for (uint32_t j = t; j < nit; j += T) {
            const int32_t r0 = S.it_r0[X.i0 + j], r1 = S.it_r1[X.i0 + j];
            if (r0 >= 0) {
                const uint32_t x = (uint32_t)r0 - base;
                fill[L.doff[x] + atomicAdd(&L.dfill[x], 1u)] = j * 2u;
            }
            if (r1 >= 0) {
                const uint32_t x = (uint32_t)r1 - base;
                fill[L.doff[x] + atomicAdd(&L.dfill[x], 1u)] = j * 2u + 1u;
            }
        }
        __syncthreads();
        if (lds_lists) {
            for (uint32_t j = t; j < nit; j += T) {
#pragma unroll
                for (uint32_t side = 0; side < 2; ++side) {
                    const int32_t r = side ? S.it_r1[X.i0 + j] : S.it_r0[X.i0 + j];
                    if (r < 0) continue;
                    const uint32_t x = (uint32_t)r - base, o = L.doff[x];
                    const uint32_t v = j * 2u + side;
                    const uint32_t rank = count_below(ul + o, L.dcnt[x], v);
                    lists[o + rank] = v;
                    if (side) S.it_k1[X.i0 + j] = rank; else S.it_k0[X.i0 + j] = rank;
                }
            }
        } else {
            for (uint32_t x = t; x < n; x += T) {
                uint32_t* l = lists + L.doff[x];
                const uint32_t c = L.dcnt[x];
                for (uint32_t i = 1; i < c; ++i) {
                    uint32_t v = l[i];
                    uint32_t j = i;
                    while (j > 0 && l[j - 1] > v) { l[j] = l[j - 1]; --j; }
                    l[j] = v;
                }
                for (uint32_t p = 0; p < c; ++p) {
                    uint32_t v = l[p];
                    if (v & 1u) S.it_k1[X.i0 + (v >> 1)] = p; else S.it_k0[X.i0 + (v >> 1)] = p;
                }
            }
        }
        __syncthreads();
        // execution positions by dependency level when a lane holds more than
        // one item (with a lane per item the order of positions is moot); its
        // work arrays in the solve's part of LDS, free until the solve
        if (nit > T)
            level_order(S.it_r0 + X.i0, S.it_r1 + X.i0, S.it_k0 + X.i0, S.it_k1 + X.i0, S.it_perm + X.i0, lists, L.doff,
                        base, 0u, nit, L.solve, fused_solve_words(NSLOT > 0) * n, s_lvl);
        if (WARM && D.warm > 0.0f) {
            // warm starting: match last step's manifolds, then each body's carried impulse
            for (uint32_t q = t; q < M; q += T) warm_match(S, D, X.m0 + q, base, sub_steps);
            __syncthreads();
            for (uint32_t x = t; x < n; x += T) warm_body_sum(S, D, lists + L.doff[x], L.dcnt[x], x, 0u, nj, X.m0);
            __syncthreads();
        }
        mark(FP_SCHED);
        // ---- solve
        fused_solve<NSLOT, MAXT, WAVES, LOCKSTEP, WARM>(S, D, w, X.i0, nit, X.m0, L.dcnt, dt, sub_steps, iters, L.solve,
                                                  L.lanes,
                                                  &s_abort);
        __syncthreads();
        mark(FP_SOLVE);
        if (s_abort) break;                     // the solve stalled: status set
        if (WARM && D.warm > 0.0f) {
            // warm starting: this step's manifolds become the next step's tables
            for (uint32_t x = t; x < n; x += T) D.wcnt[x] = 0u;
            __syncthreads();
            for (uint32_t q = t; q < M; q += T) warm_save(S, D, X.m0 + q, X.m0, X.m0 + M, base);
            __syncthreads();
        }
        if (t == 0) {
            cnt[0] = E; cnt[1] = R; cnt[2] = M; cnt[3] = nit;
            unsigned long long* sums = reinterpret_cast<unsigned long long*>(S.counters + C_SUMS);
            atomicAdd(&sums[0], (unsigned long long)n);
            atomicAdd(&sums[1], (unsigned long long)E);
            atomicAdd(&sums[2], (unsigned long long)R);
            atomicAdd(&sums[3], (unsigned long long)M);
            atomicAdd(&sums[4], (unsigned long long)nit);
            S.w_man_begin[w] = X.m0;
            S.w_man_end[w] = X.m0 + M;
            S.w_done[w] = step + 1;
            if (ts && step < S.tcap) ts[step + 1] = wall_clock64();
        }
    }
    // the counts of the world's last step into the call's totals
    __syncthreads();
    if (t < 4) atomicAdd(&S.counters[t], cnt[t]);
}

// step bookkeeping: counters of one step start at zero; the overflow flag is
// sticky (it is cleared by the host), so a replayed sequence of steps stops
// committing at the first step whose scratch did not fit.
__global__ void k_begin_step(uint32_t* counters)
{
    if (threadIdx.x < C_OVERFLOW) counters[threadIdx.x] = 0u;
}

__global__ void k_commit_step(uint32_t* counters)
{
    if (threadIdx.x == 0 && counters[C_OVERFLOW] == 0u) counters[C_STEPS] += 1u;
}

// ------------------------------------------------------------ host side
static inline unsigned blocks_for(uint64_t n, unsigned bs, unsigned cap = 4096)
{
    uint64_t b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

size_t solve_lds_bytes(int nbodies) { return (size_t)nbodies * 7u * 4u; }

typedef void (*SlotFn)(StepDev, float, uint32_t, uint32_t, uint32_t);
typedef void (*RecFn)(StepDev, float, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t);
typedef void (*FusedFn)(StepDev, float, uint32_t, uint32_t, uint32_t);
struct SolveVariant {
    SlotFn slot_fn;      // register-slot kernel (nullptr: none)
    unsigned slot_threads;
    int slots;
    RecFn rec_fn;        // record kernel (LDS body state) for the worlds the slot kernel leaves
    unsigned threads;
    FusedFn fused_fn;    // whole steps per workgroup with the slot kernel's shape (nullptr: none)
    unsigned fused_t;    // the fused kernel's workgroup size when it is neither slot_threads nor threads
};
// solve variants: the register-slot kernel's slots x workgroup size, and the
// record kernel's workgroup size for worlds with more items than slots
// (fewer waves per world keep a dependency chain's wave from sharing its
// SIMD with the world's other waves: 256 threads beat 1,024 by about 3x on a
// 4,096-body world)
static const SolveVariant kSolveVariants[SOLVE_TABLE] = {
    {nullptr, 0, 0, k_solve<1024>, 1024, nullptr},
    {k_solve_slots<4, 256, 3>, 256, 4, k_solve<256, 2>, 256, k_fused<4, 256, 3, false>},
    {k_solve_slots<2, 512, 3>, 512, 2, k_solve<256, 2>, 256, k_fused<2, 512, 3, false>},
    {k_solve_slots<3, 256, 3>, 256, 3, k_solve<256, 2>, 256, k_fused<3, 256, 3, false>},
    {k_solve_slots<4, 256, 2>, 256, 4, k_solve<256, 2>, 256, k_fused<4, 256, 2, false>},
    {k_solve_slots<8, 128, 2>, 128, 8, k_solve<256, 2>, 256, k_fused<8, 128, 2, false>},
    {nullptr, 0, 0, k_solve<256, 2>, 256, nullptr},
    {k_solve_slots<4, 256, 3>, 256, 4, k_solve<128, 2>, 128, k_fused<4, 256, 3, false>},
    // the lockstep slot walk takes the phase kernels only: its fused kernel
    // hung once in a test with no cause found, so it is not built
    {k_solve_slots<3, 256, 3, true>, 256, 3, k_solve<256, 2>, 256, nullptr},
    // records only (as variant 6), fused with 3 or 4 waves per SIMD: fewer
    // registers than the slots, so more worlds resident per CU
    {nullptr, 0, 0, k_solve<256, 2>, 256, k_fused<0, 256, 3, false>},
    {nullptr, 0, 0, k_solve<256, 2>, 256, k_fused<0, 256, 4, false>},
    // records only, 192-thread workgroups: with the records' smaller LDS
    // state, four worlds of 500 bodies are resident per CU
    {nullptr, 0, 0, k_solve<192, 3>, 192, k_fused<0, 192, 3, false>},
    // records only, 128-thread workgroups at 2 waves per SIMD: the same four
    // worlds per CU with the full register file (no spills)
    {nullptr, 0, 0, k_solve<128, 2>, 128, k_fused<0, 128, 2, false>},
    // records only, 512-thread workgroups at 2 waves per SIMD (the full
    // register file): for a few worlds, each alone on a CU, whose chains are
    // latency-bound; up to 512 items no lane ever changes record
    {nullptr, 0, 0, k_solve<REC_WIDE_THREADS, 1>, REC_WIDE_THREADS, k_fused<0, 512, 2, false>},
    // WIDE_VARIANT: 768 lanes at 3 waves per SIMD (the register budget of
    // variant 11), up to 768 items without a pickup.  A lane that holds two
    // items serialises them, which can triple a chain (the 20-row pyramid's
    // 520-590 items: 148 links per sub-step with a lane per item, 403-437
    // with 512 lanes), so the widest shape that does not spill wins
    {nullptr, 0, 0, k_solve<REC_WIDE_THREADS, 1>, REC_WIDE_THREADS, k_fused<0, 768, 3, false>, 768},
    // WARM_VARIANT: batches with a warm-started world (the shape of variant
    // 11 with the warm-start passes compiled in; the kernels above carry none)
    {nullptr, 0, 0, k_solve<192, 3, true>, 192, k_fused<0, 192, 3, false, true>},
};

bool fused_slots(int v)
{
    return v >= 0 && v < SOLVE_TABLE && kSolveVariants[v].slot_fn != nullptr;
}

uint32_t fused_threads(int v)
{
    const SolveVariant& sv = kSolveVariants[v];
    if (sv.fused_t) return sv.fused_t;
    return sv.slot_fn ? (uint32_t)sv.slot_threads : (uint32_t)sv.threads;
}

// a fused kernel runs only at the workgroup size it was compiled for (its
// loops stride by MAXT): set_solve_smem_limit checks each against the width
// it would be launched with, and a mismatch leaves the variant without one
static bool g_fused_width_ok[SOLVE_TABLE];

bool fused_available(int v)
{
    return v >= 0 && v < SOLVE_TABLE && kSolveVariants[v].fused_fn != nullptr && g_fused_width_ok[v];
}

hipError_t fused_attributes(int v, hipFuncAttributes* out)
{
    return hipFuncGetAttributes(out, (const void*)kSolveVariants[v].fused_fn);
}

void launch_fused(const StepDev& S, int v, uint32_t nworlds, float dt, uint32_t sub_steps, uint32_t iters,
                  uint32_t nsteps, size_t lds, hipStream_t st)
{
    const SolveVariant& sv = kSolveVariants[v];
    const unsigned threads = fused_threads(v);
    hipLaunchKernelGGL(sv.fused_fn, dim3(nworlds), dim3(threads), lds, st, S, dt, sub_steps, iters, nsteps);
}

unsigned solve_variant_threads(int v) { return kSolveVariants[v < SOLVE_TABLE && v >= 0 ? v : 0].threads; }
int solve_variant_slots(int v) { return kSolveVariants[v < SOLVE_TABLE && v >= 0 ? v : 0].slots; }

void KernelTrace::mark(const char* name, hipStream_t st)
{
    if (marks.size() == pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        pool.push_back(e);
    }
    hipEvent_t e = pool[marks.size()];
    if (hipEventRecord(e, st) != hipSuccess) return;
    marks.push_back({name, e, st});
}

void KernelTrace::flush()
{
    if (marks.empty()) return;
    (void)hipDeviceSynchronize();
    // a main-stream kernel's interval starts at the previous main-stream mark;
    // a side-stream kernel's at the last main-stream mark before it (the fork)
    const Mark* prev_main = nullptr;
    for (const Mark& m : marks) {
        if (prev_main && m.name[0]) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, prev_main->ev, m.ev) == hipSuccess) {
                auto& t = totals[m.name];
                t.first += ms;
                t.second += 1;
            }
        }
        if (m.st == main) prev_main = &m;
    }
    marks.clear();
}

void KernelTrace::report(FILE* f)
{
    flush();
    std::vector<std::pair<double, std::string>> v;
    double all = 0;
    for (auto& kv : totals) { v.push_back({kv.second.first, kv.first}); all += kv.second.first; }
    std::sort(v.rbegin(), v.rend());
    std::fprintf(f, "[r2d] kernel times (eager launches, one event after each):\n");
    for (auto& e : v) {
        const auto& t = totals[e.second];
        std::fprintf(f, "[r2d] %-28s calls %8llu  total %10.3f ms  avg %9.2f us  %5.1f%%\n", e.second.c_str(),
                     (unsigned long long)t.second, t.first, 1e3 * t.first / (double)std::max<uint64_t>(t.second, 1),
                     100.0 * t.first / std::max(all, 1e-9));
    }
}

KernelTrace::~KernelTrace()
{
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

static inline void tmark(const Streams& ss, const char* name, hipStream_t st)
{
    if (ss.trace) ss.trace->mark(name, st);
}

SolvePlan plan_solve(int variant, const uint32_t* nbodies, size_t nw, uint32_t lds_limit, bool any_colored,
                     bool any_warm, bool allow_wide)
{
    SolvePlan p;
    p.any_colored = any_colored ? 1u : 0u;
    p.any_warm = any_warm ? 1u : 0u;
    p.variant = (variant >= 0 && variant < SOLVE_TABLE) ? variant : 0;
    const SolveVariant& v = kSolveVariants[p.variant];
    p.slot_cap = v.slot_fn ? (uint32_t)v.slots * v.slot_threads : 0u;
    p.slot_nmax = v.slot_fn ? SLOT_NMAX : 0u;
    p.lds_nmax = (uint32_t)((160u * 1024u - 1024u - rec_lane_bytes(v.threads)) / solve_lds_bytes(1));
    if (lds_limit) p.lds_nmax = std::min(p.lds_nmax, lds_limit);
    for (size_t i = 0; i < nw; ++i) {
        uint32_t n = nbodies[i];
        if (p.slot_cap && n <= p.slot_nmax) { p.any_slot = 1; p.slot_maxn = std::max(p.slot_maxn, n); }
        if (n <= p.lds_nmax) { p.any_lds = 1; p.lds_maxn = std::max(p.lds_maxn, n); p.n_lds += 1; }
        else { p.any_big = 1; p.n_big += 1; }
    }
    // A few worlds, each alone on a CU, are latency-bound on their chains:
    // more lanes hold fewer positions each and wait less (4,096-body world:
    // 0.71 -> 0.58 ms per step with 512 lanes).  Many worlds share CUs and
    // are issue-bound: the variant's shape.  The wide kernel's lane slots
    // leave LDS for fewer bodies.
    const uint32_t wide_nmax = (uint32_t)((160u * 1024u - 1024u - rec_lane_bytes(REC_WIDE_THREADS)) /
                                          solve_lds_bytes(1));
    p.rec_wide = (allow_wide && p.any_lds && !any_warm && p.n_lds <= 64u && p.lds_maxn <= wide_nmax &&
                  kSolveVariants[p.variant].threads < REC_WIDE_THREADS) ? 1u : 0u;
    return p;
}

void launch_phase(const StepDev& S, const Streams& ss, int phase, float dt, uint32_t sub_steps, uint32_t iters,
                  const SolvePlan& sp)
{
    const unsigned BS = 256;
    const unsigned GRID = 1024;   // grid-stride kernels over device-sized ranges
    hipStream_t st = ss.main;
    switch (phase) {
    case PH_BROAD:
        hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, st, S.counters);
        tmark(ss, "k_begin_step", st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nworlds + 1, BS)), dim3(BS), 0, st,
                           reinterpret_cast<uint32_t*>(S.w_status), S.nworlds + 1);
        tmark(ss, "k_zero_u32", st);
        hipLaunchKernelGGL(k_fill_bworld, dim3(S.nworlds), dim3(256), 0, st, S);
        tmark(ss, "k_fill_bworld", st);
        hipLaunchKernelGGL(k_prep, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
        tmark(ss, "k_prep", st);
        scan_u32(S.ncell, S.cell_off, nullptr, S.nb, S.nb, &S.counters[C_ENTRIES], S.partials, st);
        tmark(ss, "scan ncell", st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nbuckets + 1, BS)), dim3(BS), 0, st, S.bkt_cnt,
                           S.nbuckets + 1);
        tmark(ss, "k_zero_u32", st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nbuckets + 1, BS)), dim3(BS), 0, st, S.bkt_fill,
                           S.nbuckets + 1);
        tmark(ss, "k_zero_u32", st);
        hipLaunchKernelGGL(k_entries, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_entries", st);
        scan_u32(S.bkt_cnt, S.bkt_off, nullptr, S.nbuckets, S.nbuckets, nullptr, S.partials, st);
        tmark(ss, "scan bkt_cnt", st);
        hipLaunchKernelGGL(k_bucket_fill, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_bucket_fill", st);
        hipLaunchKernelGGL(k_bucket_rank, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_bucket_rank", st);
        hipLaunchKernelGGL(k_ent_cnt, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_ent_cnt", st);
        scan_u32(S.ent_cnt, S.ent_off, &S.counters[C_ENTRIES], 0, S.cap_e, &S.counters[C_RAW], S.partials, st);
        tmark(ss, "scan ent_cnt", st);
        hipLaunchKernelGGL(k_cand_emit, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_cand_emit", st);
        break;
    case PH_NARROW:
        hipLaunchKernelGGL(k_narrow, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_narrow", st);
        scan_u32(S.cand_flag, S.cand_idx, &S.counters[C_RAW], 0, S.cap_r, &S.counters[C_MANIFOLDS], S.partials, st);
        tmark(ss, "scan cand_flag", st);
        hipLaunchKernelGGL(k_man_build, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_man_build", st);
        break;
    case PH_SCHED:
        hipLaunchKernelGGL(k_world_ranges, dim3(blocks_for(S.nworlds + 1, BS)), dim3(BS), 0, st, S);
        tmark(ss, "k_world_ranges", st);
        hipLaunchKernelGGL(k_joint_mark, dim3(blocks_for(S.njoints, BS)), dim3(BS), 0, st, S);
        tmark(ss, "k_joint_mark", st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nb + 1, BS)), dim3(BS), 0, st, S.deg, S.nb + 1);
        tmark(ss, "k_zero_u32", st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nb + 1, BS)), dim3(BS), 0, st, S.deg_fill, S.nb + 1);
        tmark(ss, "k_zero_u32", st);
        hipLaunchKernelGGL(k_items_joints, dim3(blocks_for(S.njoints, BS)), dim3(BS), 0, st, S);
        tmark(ss, "k_items_joints", st);
        hipLaunchKernelGGL(k_items_mans, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_items_mans", st);
        scan_u32(S.deg, S.deg_off, nullptr, S.nb, S.nb, nullptr, S.partials, st);
        tmark(ss, "scan deg", st);
        hipLaunchKernelGGL(k_list_fill, dim3(GRID), dim3(BS), 0, st, S);
        tmark(ss, "k_list_fill", st);
        hipLaunchKernelGGL(k_list_rank<false>, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
        tmark(ss, "k_list_rank<false>", st);
        {
            // the level order's work arrays: 96 KB for batches of up to 256
            // worlds (items per world up to ~8,000), 32 KB for larger ones
            const size_t lds = S.nworlds <= 256u ? 96u * 1024u : 32u * 1024u;
            hipLaunchKernelGGL(k_color, dim3(S.nworlds), dim3(256), lds, st, S, (uint32_t)(lds / 4u));
        }
        tmark(ss, "k_color", st);
        hipLaunchKernelGGL(k_world_order, dim3(1), dim3(1024), 0, st, S);
        tmark(ss, "k_world_order", st);
        if (sp.any_colored) {
            hipLaunchKernelGGL(k_list_rank<true>, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
            tmark(ss, "k_list_rank<true>", st);
        }
        if (sp.any_warm) {
            hipLaunchKernelGGL(k_warm_match, dim3(GRID), dim3(BS), 0, st, S, sub_steps);
            tmark(ss, "k_warm_match", st);
            hipLaunchKernelGGL(k_warm_sums, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
            tmark(ss, "k_warm_sums", st);
        }
        break;
    default: {
        // the three solve kernels take disjoint sets of worlds (solve_route);
        // the record kernels run on side streams beside the slot kernel
        const SolveVariant& v = kSolveVariants[sp.variant];
        const bool run_slot = v.slot_fn && sp.any_slot, run_lds = sp.any_lds != 0, run_big = sp.any_big != 0;
        if (run_lds || run_big) {
            (void)hipEventRecord(ss.fork, st);
            if (run_big) (void)hipStreamWaitEvent(ss.aux[1], ss.fork, 0);
            if (run_lds) (void)hipStreamWaitEvent(ss.aux[0], ss.fork, 0);
        }
        if (run_big) {
            // worlds beyond LDS: grid-wide body and item phases around a
            // one-workgroup-per-world chain kernel, all on the second side stream
            const hipStream_t sb = ss.aux[1];
            const float sub_dt = dt / (float)sub_steps;
            const unsigned gb = blocks_for(S.nb, BS);
            // the per-item passes: up to BIG_SPLIT workgroups per world beyond
            // LDS, fewer when there are many of them or many other worlds
            // (every world gets the same count and the others exit at once)
            const uint32_t split = std::max<uint32_t>(
                1u, std::min<uint32_t>({BIG_SPLIT, 256u / std::max<uint32_t>(sp.n_big, 1u),
                                        8192u / std::max<uint32_t>(S.nworlds, 1u)}));
            const dim3 gi(S.nworlds, split);
            hipLaunchKernelGGL(k_big_bodies, dim3(gb), dim3(BS), 0, sb, S, sp.lds_nmax, 0, 0u, sub_steps, sub_dt);
            hipLaunchKernelGGL(k_big_item_pass, gi, dim3(BIG_THREADS), 0, sb, S, sp.lds_nmax, 0, 0u, sub_steps, sub_dt);
            hipLaunchKernelGGL(k_big_items, dim3(S.nworlds), dim3(BIG_THREADS), 0, sb, S, sp.lds_nmax);
            for (uint32_t s = 0; s < sub_steps; ++s) {
                hipLaunchKernelGGL(k_big_bodies, dim3(gb), dim3(BS), 0, sb, S, sp.lds_nmax, 1, s, sub_steps, sub_dt);
                hipLaunchKernelGGL(k_big_item_pass, gi, dim3(BIG_THREADS), 0, sb, S, sp.lds_nmax, 1, s, sub_steps,
                                   sub_dt);
                // worlds whose touched bodies fit LDS, then those that do not
                // the second launch's LDS: its lane slots, then the counters
                constexpr size_t GLOBAL_LDS = COMPACT_LDS;
                if (sp.any_warm) {
                    hipLaunchKernelGGL((k_big_chain<256, true, 0>), dim3(S.nworlds), dim3(256), COMPACT_LDS, sb, S,
                                       sp.lds_nmax, iters, sub_dt);
                    hipLaunchKernelGGL((k_big_chain<BIG_GLOBAL_THREADS, true, 1>), dim3(S.nworlds),
                                       dim3(BIG_GLOBAL_THREADS), GLOBAL_LDS, sb, S, sp.lds_nmax, iters, sub_dt);
                } else {
                    hipLaunchKernelGGL((k_big_chain<256, false, 0>), dim3(S.nworlds), dim3(256), COMPACT_LDS, sb, S,
                                       sp.lds_nmax, iters, sub_dt);
                    hipLaunchKernelGGL((k_big_chain<BIG_GLOBAL_THREADS, false, 1>), dim3(S.nworlds),
                                       dim3(BIG_GLOBAL_THREADS), GLOBAL_LDS, sb, S, sp.lds_nmax, iters, sub_dt);
                }
                hipLaunchKernelGGL(k_big_bodies, dim3(gb), dim3(BS), 0, sb, S, sp.lds_nmax, 2, s, sub_steps, sub_dt);
            }
            hipLaunchKernelGGL(k_big_item_pass, gi, dim3(BIG_THREADS), 0, sb, S, sp.lds_nmax, 2, 0u, sub_steps, sub_dt);
            tmark(ss, "k_big (global state)", sb);
        }
        if (run_lds && sp.rec_wide) {
            hipLaunchKernelGGL((k_solve<REC_WIDE_THREADS, 1>), dim3(S.nworlds), dim3(REC_WIDE_THREADS),
                               solve_lds_bytes((int)sp.lds_maxn) + rec_lane_bytes(REC_WIDE_THREADS), ss.aux[0], S, dt,
                               sub_steps, iters, sp.slot_cap, sp.slot_nmax, sp.lds_nmax);
            tmark(ss, "k_solve (record, wide)", ss.aux[0]);
        } else if (run_lds) {
            hipLaunchKernelGGL(v.rec_fn, dim3(S.nworlds), dim3(v.threads),
                               solve_lds_bytes((int)sp.lds_maxn) + rec_lane_bytes(v.threads),
                               ss.aux[0], S, dt, sub_steps, iters, sp.slot_cap, sp.slot_nmax, sp.lds_nmax);
            tmark(ss, "k_solve (record)", ss.aux[0]);
        }
        if (run_slot) {
            hipLaunchKernelGGL(v.slot_fn, dim3(S.nworlds), dim3(v.slot_threads), slot_lds_bytes((int)sp.slot_maxn),
                               st, S, dt, sub_steps, iters, sp.slot_nmax);
            tmark(ss, "k_solve_slots", st);
        }
        if (run_lds) {
            (void)hipEventRecord(ss.join[0], ss.aux[0]);
            (void)hipStreamWaitEvent(st, ss.join[0], 0);
        }
        if (run_big) {
            (void)hipEventRecord(ss.join[1], ss.aux[1]);
            (void)hipStreamWaitEvent(st, ss.join[1], 0);
        }
        if (sp.any_warm) {
            hipLaunchKernelGGL(k_warm_clear, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
            tmark(ss, "k_warm_clear", st);
            hipLaunchKernelGGL(k_warm_save, dim3(GRID), dim3(BS), 0, st, S);
            tmark(ss, "k_warm_save", st);
        }
        hipLaunchKernelGGL(k_commit_step, dim3(1), dim3(64), 0, st, S.counters);
        tmark(ss, "k_commit_step", st);
        break;
    }
    }
}

void set_solve_smem_limit()
{
    (void)hipFuncSetAttribute((const void*)k_color, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_big_chain<256, false, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_solve<REC_WIDE_THREADS, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_big_chain<256, true, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_big_chain<BIG_GLOBAL_THREADS, false, 1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void*)k_big_chain<BIG_GLOBAL_THREADS, true, 1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    for (int v = 0; v < SOLVE_TABLE; ++v) {
        if (kSolveVariants[v].slot_fn)
            (void)hipFuncSetAttribute((const void*)kSolveVariants[v].slot_fn,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)kSolveVariants[v].rec_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        if (kSolveVariants[v].fused_fn) {
            (void)hipFuncSetAttribute((const void*)kSolveVariants[v].fused_fn,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipFuncAttributes fa{};
            g_fused_width_ok[v] = hipFuncGetAttributes(&fa, (const void*)kSolveVariants[v].fused_fn) == hipSuccess &&
                                  fa.maxThreadsPerBlock == (int)fused_threads(v);
        }
    }
    (void)hipGetLastError();   // a refused attribute is not an error of a later call
}

}  // namespace r2d
