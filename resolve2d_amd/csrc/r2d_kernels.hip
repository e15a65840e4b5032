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

#include "r2d_device.h"
#include "r2d_kernels.h"

namespace r2d {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t grid_stride() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

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

__device__ __forceinline__ Geo load_geo(const WorldDev& D, uint32_t slot, float c, float s)
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
    v2 a1 = self.wv[iter], a2 = self.wv[nx];
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
__device__ bool overlap_sat(Sat& ret, const Geo& ref, const Geo& inc, int which)
{
    const float EPS = SAT_OVERLAP_THRESHOLD;
    int nn = ref.rect ? 4 : 1;
    for (int it = 0; it < nn; ++it) {
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
__device__ __forceinline__ bool aabb_hit(const WorldDev& D, uint32_t a, uint32_t b)
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
    const WorldDev& D = S.worlds[blockIdx.x];
    for (int i = threadIdx.x; i < D.nbodies; i += blockDim.x) S.bworld[D.body_base + i] = blockIdx.x;
}

// SpatialHash.iterateAABBHashes cell range (SpatialHash.zig:83-96) from the
// stale AABB, plus per-step trig and scheduling flags.
__global__ void k_prep(StepDev S)
{
    for (uint32_t i = grid_tid(); i < S.nb; i += grid_stride()) {
        uint32_t w = S.bworld[i];
        const WorldDev& D = S.worlds[w];
        uint32_t sl = i - D.body_base;
        float a = D.ang[sl];
        S.cc[i] = cosf_r(a);
        S.cs[i] = sinf_r(a);
        uint32_t f = D.flags[sl];
        uint32_t res = 1u;
        if (f & BF_STATIC) {
            // a static body orders contacts only if a contact's "+= 0" could
            // change it: a -0 component (joint-touched bodies are added later)
            float m0 = D.mx[sl], m1 = D.my[sl], m2 = D.L[sl];
            bool negz = (m0 == 0.0f && __builtin_signbit(m0)) || (m1 == 0.0f && __builtin_signbit(m1)) ||
                        (m2 == 0.0f && __builtin_signbit(m2));
            res = negz ? 1u : 0u;
        }
        S.sflag[i] = res;
        float cell = D.cell;
        float ax = D.ax[sl], ay = D.ay[sl], hw = D.ahw[sl], hh = D.ahh[sl];
        float v0 = floorf((ax - hw) / cell), v1 = floorf((ay - hh) / cell);
        float v2f = floorf((ax + hw) / cell), v3 = floorf((ay + hh) / cell);
        const float LIM = 4.0e18f;
        bool ok = (v0 > -LIM && v0 < LIM) && (v1 > -LIM && v1 < LIM) && (v2f > -LIM && v2f < LIM) &&
                  (v3 > -LIM && v3 < LIM);
        uint32_t nc = 0;
        if (ok) {
            int64_t x0 = (int64_t)v0, y0 = (int64_t)v1, x1 = (int64_t)v2f, y1 = (int64_t)v3;
            uint64_t nx = (x1 >= x0) ? (uint64_t)(x1 - x0 + 1) : 0u;
            uint64_t ny = (y1 >= y0) ? (uint64_t)(y1 - y0 + 1) : 0u;
            uint64_t tot = nx * ny;
            if (nx > (1u << 30) || ny > (1u << 30) || tot > (1u << 30)) {
                S.w_status[w] = 3;   // cell range too large to enumerate
            } else {
                nc = (uint32_t)tot;
                S.cx0[i] = x0; S.cy0[i] = y0; S.cnx[i] = (uint32_t)nx; S.cny[i] = (uint32_t)ny;
            }
        } else {
            S.w_status[w] = 4;       // non-finite AABB: @intFromFloat is undefined in the reference
        }
        S.ncell[i] = nc;
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

// Within a bucket the reference stores entries in reverse insertion order
// (pre-decrement scatter, SpatialHash.zig:36-40, 59-68): descending body slot.
// A body listed twice in one bucket cannot change which pair appears first.
__global__ void k_bucket_sort(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t gb = grid_tid(); gb < S.nbuckets; gb += grid_stride()) {
        uint32_t o = S.bkt_off[gb], n = S.bkt_cnt[gb];
        uint32_t* it = S.bkt_items + o;
        for (uint32_t i = 1; i < n; ++i) {
            uint32_t v = it[i];
            uint32_t j = i;
            while (j > 0 && it[j - 1] < v) { it[j] = it[j - 1]; --j; }
            it[j] = v;
        }
        uint32_t u = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (u == 0 || it[u - 1] != it[i]) it[u++] = it[i];
        S.bkt_len[gb] = u;
    }
}

// raw candidates per entry: bucket members with a higher slot than the entry's body
__global__ void k_ent_cnt(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t E = dev_count(&S.counters[C_ENTRIES], S.cap_e);
    for (uint32_t e = grid_tid(); e < E; e += grid_stride()) {
        uint32_t a = S.ent_body[e], gb = S.ent_bkt[e];
        const uint32_t* it = S.bkt_items + S.bkt_off[gb];
        uint32_t n = S.bkt_len[gb], c = 0;
        while (c < n && it[c] > a) ++c;
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
        const uint32_t* it = S.bkt_items + S.bkt_off[gb];
        uint32_t o = S.ent_off[e], c = S.ent_cnt[e];
        for (uint32_t j = 0; j < c; ++j) { S.cand_a[o + j] = a; S.cand_b[o + j] = it[j]; }
    }
}

// Solver.updateManifolds filters + performNarrowSAT (lib.zig:272-285,
// collision.zig:299-320) for each raw candidate.  A candidate counts only at
// the first position of its partner in the lower body's query list; later
// occurrences are the reference's "already a manifold / already rejected".
__global__ void k_narrow(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    uint32_t R = dev_count(&S.counters[C_RAW], S.cap_r);
    for (uint32_t q = grid_tid(); q < R; q += grid_stride()) {
        uint32_t a = S.cand_a[q], b = S.cand_b[q];
        uint32_t seg = S.ent_off[S.cell_off[a]];
        uint32_t flag = 0, info = 0;
        float nx = 0.f, ny = 0.f;
        bool first = true;
        for (uint32_t p = seg; p < q; ++p)
            if (S.cand_b[p] == b) { first = false; break; }
        if (first) {
            uint32_t w = S.bworld[a];
            const WorldDev& D = S.worlds[w];
            uint32_t sa = a - D.body_base, sb = b - D.body_base;
            uint32_t fa = D.flags[sa], fb = D.flags[sb];
            bool ok = !((fa & BF_STATIC) && (fb & BF_STATIC));
            if (ok && D.nexcl) ok = !excluded(D, D.id[sa], D.id[sb]);
            if (ok) ok = aabb_hit(D, sa, sb);
            if (ok) {
                Geo ga = load_geo(D, sa, S.cc[a], S.cs[a]);
                Geo gb = load_geo(D, sb, S.cc[b], S.cs[b]);
                // performNarrowSAT: lower id is the first reference body
                bool a_first = ga.id < gb.id;
                const Geo& o1 = a_first ? ga : gb;
                const Geo& o2 = a_first ? gb : ga;
                Sat sat;
                const float undef = -3.0316488252093987e-13f;   // Zig Debug fill of `undefined` (0xAA)
                sat.normal = mk(undef, undef);
                sat.pen = __builtin_inff();
                sat.nid = 0; sat.ref = 0; sat.set = false;
                if (overlap_sat(sat, o1, o2, 0) && overlap_sat(sat, o2, o1, 1)) {
                    flag = 1;
                    // ref body expressed relative to (a, b)
                    bool ref_is_a = (sat.ref == 0) == a_first;
                    info = (uint32_t)sat.nid | (ref_is_a ? 4u : 0u);
                    nx = sat.normal.x; ny = sat.normal.y;
                }
            }
        }
        S.cand_flag[q] = flag;
        S.sat_info[q] = info;
        S.sat_nx[q] = nx;
        S.sat_ny[q] = ny;
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
// with clipAgainstEdge (Rectangle.zig:213-238, Disc.zig:125-130)
__global__ void k_man_build(StepDev S)
{
    uint32_t R = dev_count(&S.counters[C_RAW], S.cap_r);
    if (S.counters[C_OVERFLOW]) return;
    if (S.counters[C_MANIFOLDS] > S.cap_m) { if (grid_tid() == 0) atomicOr(&S.counters[C_OVERFLOW], 4u); return; }
    for (uint32_t q = grid_tid(); q < R; q += grid_stride()) {
        if (!S.cand_flag[q]) continue;
        uint32_t m = S.cand_idx[q];
        uint32_t a = S.cand_a[q], b = S.cand_b[q];
        uint32_t info = S.sat_info[q];
        int nid = (int)(info & 3u);
        bool ref_is_a = (info & 4u) != 0;
        uint32_t ref = ref_is_a ? a : b, inc = ref_is_a ? b : a;
        const WorldDev& D = S.worlds[S.bworld[a]];
        uint32_t sr = ref - D.body_base, si = inc - D.body_base;
        Geo gr = load_geo(D, sr, S.cc[ref], S.cs[ref]);
        Geo gi = load_geo(D, si, S.cc[inc], S.cs[inc]);
        v2 normal = mk(S.sat_nx[q], S.sat_ny[q]);
        S.m_ref[m] = ref;
        S.m_inc[m] = inc;
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
                v2 ca, cb;
                clip_line(e.a, e.b, ba, bb, ca, cb);
                int i = 0;
                float d = dot2(sub2(ca, e.a), e.dir);
                if (d < COLLISION_MARGIN) { write_point(S, m, i, ca, d, gr, gi, e.dir); valid |= 1u << i; i += 1; }
                d = dot2(sub2(cb, e.b), e.dir);
                if (d < COLLISION_MARGIN) { write_point(S, m, i, cb, d, gr, gi, e.dir); valid |= 1u << i; }
            }
        }
        S.m_valid[m] = valid;
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

// Coloured schedule: greedy colouring of each world's items in item order;
// an item takes the smallest colour unused by earlier items sharing one of
// its ordering bodies.  Worlds in reference order keep colour 0 everywhere.
__global__ void k_color(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t w = grid_tid(); w < S.nworlds; w += grid_stride()) {
        const WorldDev& D = S.worlds[w];
        uint32_t i0 = S.w_item_base[w], i1 = S.w_item_base[w + 1];
        if (D.sched == 0) {
            for (uint32_t it = i0; it < i1; ++it) { S.it_color[it] = 0u; S.it_perm[it] = it - i0; }
            continue;
        }
        for (int b = 0; b < D.nbodies; ++b) { S.cmask[2 * (D.body_base + b)] = 0; S.cmask[2 * (D.body_base + b) + 1] = 0; }
        uint32_t overflow_color = 128u;
        for (uint32_t it = i0; it < i1; ++it) {
            int32_t r0 = S.it_r0[it], r1 = S.it_r1[it];
            uint64_t m0 = 0, m1 = 0;
            if (r0 >= 0) { m0 |= S.cmask[2 * r0]; m1 |= S.cmask[2 * r0 + 1]; }
            if (r1 >= 0) { m0 |= S.cmask[2 * r1]; m1 |= S.cmask[2 * r1 + 1]; }
            uint32_t c;
            if (~m0) c = (uint32_t)__builtin_ctzll(~m0);
            else if (~m1) c = 64u + (uint32_t)__builtin_ctzll(~m1);
            else c = overflow_color++;   // a colour of its own is always independent
            S.it_color[it] = c;
            if (c < 64) {
                uint64_t bit = 1ull << c;
                if (r0 >= 0) S.cmask[2 * r0] |= bit;
                if (r1 >= 0) S.cmask[2 * r1] |= bit;
            } else if (c < 128) {
                uint64_t bit = 1ull << (c - 64);
                if (r0 >= 0) S.cmask[2 * r0 + 1] |= bit;
                if (r1 >= 0) S.cmask[2 * r1 + 1] |= bit;
            }
        }
        // execution order = (colour, item): a stable counting sort; colours past
        // 127 are unique and increase with the item index, so they share the
        // last bin in item order.  Workers walk this order, the same order the
        // per-body ranks use, so the earliest unfinished item always has a worker.
        uint32_t* h = S.chist + 130 * w;
        for (int c = 0; c < 130; ++c) h[c] = 0u;
        for (uint32_t it = i0; it < i1; ++it) h[umin(S.it_color[it], 128u) + 1] += 1u;
        for (int c = 1; c < 130; ++c) h[c] += h[c - 1];
        for (uint32_t it = i0; it < i1; ++it) {
            uint32_t b = umin(S.it_color[it], 128u);
            S.it_perm[i0 + h[b]] = it - i0;
            h[b] += 1u;
        }
    }
}

// rank of each item among the items sharing one of its resources: in the
// reference's sequential order (constraints, then manifolds in discovery
// order), or colour by colour for the coloured schedule
__global__ void k_list_rank(StepDev S)
{
    if (S.counters[C_OVERFLOW]) return;
    for (uint32_t x = grid_tid(); x < S.nb; x += grid_stride()) {
        uint32_t o = S.deg_off[x], n = S.deg_off[x + 1] - o;
        uint32_t* l = S.lists + o;
        for (uint32_t i = 1; i < n; ++i) {
            uint32_t v = l[i];
            uint64_t kv = ((uint64_t)S.it_color[v >> 1] << 32) | v;
            uint32_t j = i;
            while (j > 0 && (((uint64_t)S.it_color[l[j - 1] >> 1] << 32) | l[j - 1]) > kv) { l[j] = l[j - 1]; --j; }
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

__device__ __forceinline__ uint32_t ld_acq(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st_rel(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// CollisionManifold.calculateImpulses (collision.zig:135-218)
template <class ST>
__device__ __forceinline__ void solve_manifold(const StepDev& S, const WorldDev& D, const ST& st, uint32_t m,
                                               uint32_t l1, uint32_t l2, bool w1, bool w2, float dt)
{
    float* P = S.m_pt;
    const size_t cm = S.cap_m;
    float im1 = st.im[l1], ii1 = st.ii[l1], im2 = st.im[l2], ii2 = st.ii[l2];
    v2 vl1 = scale2(mk(st.mx[l1], st.my[l1]), im1);
    float om1 = st.L[l1] * ii1;
    v2 vl2 = scale2(mk(st.mx[l2], st.my[l2]), im2);
    float om2 = st.L[l2] * ii2;
    v2 n = mk(S.m_nx[m], S.m_ny[m]);
    v2 t = rot90cw(n);
    float fric = S.m_fric[m];
    uint32_t valid = S.m_valid[m];
    v2 app1 = mk(0.f, 0.f), app2 = mk(0.f, 0.f);
    float rot1 = 0.f, rot2 = 0.f;
    const uint32_t st1 = D.flags[l1] & BF_STATIC;
    const uint32_t st2 = D.flags[l2] & BF_STATIC;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(valid & (1u << p))) continue;
        float depth = P[(MP_DEPTH * 2 + p) * cm + m];
        if (depth >= 0.f) {
            P[(MP_ACC_PN * 2 + p) * cm + m] = 0.f;
            P[(MP_ACC_PT * 2 + p) * cm + m] = 0.f;
            continue;
        }
        v2 r1 = mk(P[(MP_R1X * 2 + p) * cm + m], P[(MP_R1Y * 2 + p) * cm + m]);
        v2 r2 = mk(P[(MP_R2X * 2 + p) * cm + m], P[(MP_R2Y * 2 + p) * cm + m]);
        v2 v1 = add2(vl1, mk(-r1.y * om1, r1.x * om1));
        v2 v2v = add2(vl2, mk(-r2.y * om2, r2.x * om2));
        v2 dv = sub2(v1, v2v);
        float bias = BAUMGARTE * maxn(0.0f, (-depth) - BAUMGARTE_SLOP) / dt;
        float num = dot2(dv, n) + bias;
        float pn = num * P[(MP_MASS_N * 2 + p) * cm + m];
        if (pn < MIN_MANIFOLD_IMPULSE) continue;
        num = dot2(dv, t);
        float pt = num * P[(MP_MASS_T * 2 + p) * cm + m];
        float acc_pn = P[(MP_ACC_PN * 2 + p) * cm + m];
        float acc_pt = P[(MP_ACC_PT * 2 + p) * cm + m];
        float new_pn = maxn(0.0f, acc_pn + pn);
        float applied_pn = new_pn - acc_pn;
        acc_pn = new_pn;
        float max_pt = fric * absf(acc_pn);
        float new_pt = clampn(acc_pt + pt, -max_pt, max_pt);
        float applied_pt = new_pt - acc_pt;
        acc_pt = new_pt;
        P[(MP_ACC_PN * 2 + p) * cm + m] = acc_pn;
        P[(MP_ACC_PT * 2 + p) * cm + m] = acc_pt;
        v2 dp = add2(scale2(n, applied_pn), scale2(t, applied_pt));
        if (!st1) { app1.x -= dp.x; app1.y -= dp.y; rot1 -= cross2(r1, dp); }
        if (!st2) { app2.x += dp.x; app2.y += dp.y; rot2 += cross2(r2, dp); }
    }
    if (w1) { st.mx[l1] += app1.x; st.my[l1] += app1.y; st.L[l1] += rot1; }
    if (w2) { st.mx[l2] += app2.x; st.my[l2] += app2.y; st.L[l2] += rot2; }
}

template <class ST>
__device__ __forceinline__ void solve_joint(const WorldDev& D, const ST& st, const Joint& J, float dt)
{
    if (J.type == 0) {
        // DistanceJoint.solve (DistanceJoint.zig:40-75)
        uint32_t a = J.s1, b = J.s2;
        float w1 = st.im[a], w2 = st.im[b];
        v2 v1 = scale2(mk(st.mx[a], st.my[a]), w1);
        v2 v2v = scale2(mk(st.mx[b], st.my[b]), w2);
        v2 rv = sub2(v2v, v1);
        v2 pa = mk(D.px[a], D.py[a]), pb = mk(D.px[b], D.py[b]);
        v2 n = normalize2(sub2(pb, pa));
        float dist_err = length2(sub2(pb, pa)) - J.target;
        float pc = J.beta * dist_err / maxn(J.target, 0.001f);
        float Jv = -(dot2(rv, n) + pc) / (w1 + w2);
        float ndv = dot2(n, rv);
        float den = 1.0f / maxn(absf(ndv), GRAD_DIV_LIMIT);
        Jv = clampn(Jv, J.pmin * den, J.pmax * den);
        v2 dp = scale2(n, Jv);
        st.mx[a] -= dp.x; st.my[a] -= dp.y;
        st.mx[b] += dp.x; st.my[b] += dp.y;
    } else if (J.type == 1) {
        // OffsetDistanceJoint.solve (OffsetDistanceJoint.zig:44-105)
        uint32_t a = J.s1, b = J.s2;
        float ca = cosf_r(D.ang[a]), sa = sinf_r(D.ang[a]);
        float cb = cosf_r(D.ang[b]), sb = sinf_r(D.ang[b]);
        v2 r1 = mk(J.r1x, J.r1y), r2 = mk(J.r2x, J.r2y);
        v2 rr1 = rotcs(r1, ca, sa), rr2 = rotcs(r2, cb, sb);
        v2 a1 = add2(rr1, mk(D.px[a], D.py[a]));
        v2 a2 = add2(rr2, mk(D.px[b], D.py[b]));
        v2 n = normalize2(sub2(a2, a1));
        float im1 = st.im[a], im2 = st.im[b], ii1 = st.ii[a], ii2 = st.ii[b];
        v2 vl1 = scale2(mk(st.mx[a], st.my[a]), im1);
        float w1 = st.L[a] * ii1;
        v2 vl2 = scale2(mk(st.mx[b], st.my[b]), im2);
        float w2 = st.L[b] * ii2;
        v2 v1 = add2(vl1, mk(-rr1.y * w1, rr1.x * w1));
        v2 v2v = add2(vl2, mk(-rr2.y * w2, rr2.x * w2));
        v2 dv = sub2(v2v, v1);
        float dist_err = length2(sub2(a2, a1)) - J.target;
        float pc = J.beta * dist_err / maxn(J.target, GRAD_DIV_LIMIT);
        float num = dot2(n, dv) + pc;
        float r1xn = cross2(rr1, n), r2xn = cross2(rr2, n);
        float den = im1 + im2 + r1xn * r1xn * ii1 + r2xn * r2xn * ii2;
        float Jv = -num / den;
        float ndv = dot2(n, dv);
        float den2 = 1.0f / maxn(absf(ndv), GRAD_DIV_LIMIT);
        Jv = clampn(Jv, J.pmin * den2, J.pmax * den2);
        v2 dp = scale2(n, Jv);
        st.mx[a] -= dp.x; st.my[a] -= dp.y;
        st.L[a] -= Jv * r1xn;
        st.mx[b] += dp.x; st.my[b] += dp.y;
        st.L[b] += Jv * r2xn;
    } else if (J.type == 2) {
        // FixedPositionJoint.solve (FixedPositionJoint.zig:38-72)
        uint32_t a = J.s1;
        if (D.flags[a] & BF_STATIC) return;
        float w = st.im[a];
        v2 v = scale2(mk(st.mx[a], st.my[a]), w);
        v2 delta = sub2(mk(J.tx, J.ty), mk(D.px[a], D.py[a]));
        float dist = length2(delta);
        if (dist < ALLOWED_CONSTRAINT_VALUE) return;
        v2 n = scale2(delta, 1.0f / dist);
        float bias = J.beta * dist;
        float rv = dot2(v, n);
        float Jv = -(rv - bias) / w;
        float den = 1.0f / maxn(absf(rv), ALLOWED_CONSTRAINT_VALUE);
        Jv = clampn(Jv, J.pmin * den, J.pmax * den);
        v2 dp = scale2(n, Jv);
        st.mx[a] += dp.x; st.my[a] += dp.y;
    } else {
        // MotorJoint.solve (MotorJoint.zig:38-64)
        uint32_t a = J.s1;
        if (D.flags[a] & BF_STATIC) return;
        float I = D.inertia[a];
        float omega = st.L[a] / I;
        float C = omega - J.target;
        if (absf(C) < ALLOWED_CONSTRAINT_VALUE) return;
        float bias = J.beta * C;
        float tq = D.tq[a];
        float Jv = tq - bias;
        float rel = tq / I;
        float den = 1.0f / maxn(absf(rel), ALLOWED_CONSTRAINT_VALUE);
        Jv = clampn(Jv, J.pmin * den, J.pmax * den);
        st.L[a] += Jv * dt;
    }
}

// One workgroup steps one world: S sub-steps of forces, integration,
// pre-step and K solver iterations (lib.zig:199-250).  Solver items are run
// as a dataflow: an item may run once every earlier item (in the reference's
// order) that touches one of its ordering bodies has run, which gives the
// same bits as the sequential loop.

// --------------------------------------------------- register-resident items
// One thread owns the items at execution positions tid, tid + blockDim, ...
// (at most NSLOT of them) and keeps each item's contact geometry, effective
// masses and accumulated impulses in registers for the whole step.
struct Slot {
    uint32_t flags;          // SL_* bits
    uint32_t j;              // world-local item index
    int32_t r0, r1;          // ordering bodies (world-local) or -1
    uint32_t rk0, rk1;       // ranks at r0 / r1
    uint32_t dg0, dg1;       // item counts of r0 / r1 (targets advance by these per iteration)
    uint32_t l1, l2;         // manifold bodies (ref, inc), world-local
    uint32_t m;              // manifold index in the step's arrays
    float nx, ny, fric, prev1, prev2;
    float r1x[2], r1y[2], r2x[2], r2y[2], depth[2], odepth[2], mn[2], mt[2], apn[2], apt[2];
};
enum : uint32_t { SL_USED = 1u, SL_JOINT = 2u, SL_P0 = 4u, SL_P1 = 8u, SL_W1 = 16u, SL_W2 = 32u, SL_S1 = 64u, SL_S2 = 128u };

// CollisionManifold.preStep effective masses (collision.zig:117-132) for one point
template <class ST>
__device__ __forceinline__ void slot_masses(Slot& z, const ST& st, int p)
{
    float im1 = st.im[z.l1], im2 = st.im[z.l2], ii1 = st.ii[z.l1], ii2 = st.ii[z.l2];
    float inv_mass = im1 + im2;
    v2 n = mk(z.nx, z.ny), t = rot90cw(n);
    v2 r1 = mk(z.r1x[p], z.r1y[p]), r2 = mk(z.r2x[p], z.r2y[p]);
    float r1n = cross2(r1, n), r2n = cross2(r2, n);
    float kn = inv_mass + ii1 * (r1n * r1n) + ii2 * (r2n * r2n);
    float r1t = cross2(r1, t), r2t = cross2(r2, t);
    float kt = inv_mass + ii1 * (r1t * r1t) + ii2 * (r2t * r2t);
    z.mn[p] = kn > 0.0f ? 1.0f / kn : 0.0f;
    z.mt[p] = kt > 0.0f ? 1.0f / kt : 0.0f;
}

template <class ST>
__device__ __forceinline__ void slot_load(Slot& z, const StepDev& S, const WorldDev& D, const ST& st, uint32_t pos,
                                          uint32_t ibase, uint32_t nj, uint32_t mbeg, uint32_t base)
{
    const size_t cm = S.cap_m;
    const float* P = S.m_pt;
    z.j = S.it_perm[ibase + pos];
    uint32_t item = ibase + z.j;
    z.r0 = S.it_r0[item];
    z.r1 = S.it_r1[item];
    z.flags = SL_USED;
    if (z.r0 >= 0) { z.r0 -= (int32_t)base; z.rk0 = S.it_k0[item]; z.dg0 = st.deg[z.r0]; }
    if (z.r1 >= 0) { z.r1 -= (int32_t)base; z.rk1 = S.it_k1[item]; z.dg1 = st.deg[z.r1]; }
    if (z.j < nj) { z.flags |= SL_JOINT; return; }
    uint32_t m = mbeg + (z.j - nj);
    z.m = m;
    z.l1 = S.m_ref[m] - base;
    z.l2 = S.m_inc[m] - base;
    if (S.sflag[base + z.l1] & SF_RESOURCE) z.flags |= SL_W1;
    if (S.sflag[base + z.l2] & SF_RESOURCE) z.flags |= SL_W2;
    if (D.flags[z.l1] & BF_STATIC) z.flags |= SL_S1;
    if (D.flags[z.l2] & BF_STATIC) z.flags |= SL_S2;
    z.nx = S.m_nx[m];
    z.ny = S.m_ny[m];
    z.fric = sqrt_rn(D.mu[z.l1] * D.mu[z.l2]);
    z.prev1 = S.m_prev1[m];
    z.prev2 = S.m_prev2[m];
    uint32_t valid = S.m_valid[m];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (valid & (1u << p)) z.flags |= (SL_P0 << p);
        z.r1x[p] = P[(MP_R1X * 2 + p) * cm + m];
        z.r1y[p] = P[(MP_R1Y * 2 + p) * cm + m];
        z.r2x[p] = P[(MP_R2X * 2 + p) * cm + m];
        z.r2y[p] = P[(MP_R2Y * 2 + p) * cm + m];
        z.depth[p] = P[(MP_DEPTH * 2 + p) * cm + m];
        z.odepth[p] = P[(MP_ODEPTH * 2 + p) * cm + m];
        z.apn[p] = 0.0f;
        z.apt[p] = 0.0f;
        slot_masses(z, st, p);
    }
}

// updateTGSDepth (collision.zig:73-100): points whose depth left its original
// value are moved with the bodies' rotation since the last sub-step; the
// effective masses then follow the new lever arms (preStep runs every sub-step)
template <class ST>
__device__ __forceinline__ void slot_tgs(Slot& z, const WorldDev& D, const ST& st)
{
    float a1 = D.ang[z.l1], a2 = D.ang[z.l2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(z.flags & (SL_P0 << p))) continue;
        if (approx(z.depth[p], z.odepth[p], 1e-4f)) continue;
        v2 rr1 = rotate2(mk(z.r1x[p], z.r1y[p]), a1 - z.prev1);
        v2 rr2 = rotate2(mk(z.r2x[p], z.r2y[p]), a2 - z.prev2);
        v2 q1 = add2(rr1, mk(D.px[z.l1], D.py[z.l1]));
        v2 q2 = add2(rr2, mk(D.px[z.l2], D.py[z.l2]));
        float dd = dot2(mk(z.nx, z.ny), sub2(q2, q1));
        z.depth[p] = dd + z.odepth[p];
        z.r1x[p] = rr1.x; z.r1y[p] = rr1.y; z.r2x[p] = rr2.x; z.r2y[p] = rr2.y;
        slot_masses(z, st, p);
    }
    z.prev1 = a1;
    z.prev2 = a2;
}

// CollisionManifold.calculateImpulses (collision.zig:135-218) on a register slot
template <class ST>
__device__ __forceinline__ void slot_impulses(Slot& z, const ST& st, float dt)
{
    const uint32_t l1 = z.l1, l2 = z.l2;
    float im1 = st.im[l1], ii1 = st.ii[l1], im2 = st.im[l2], ii2 = st.ii[l2];
    v2 vl1 = scale2(mk(st.mx[l1], st.my[l1]), im1);
    float om1 = st.L[l1] * ii1;
    v2 vl2 = scale2(mk(st.mx[l2], st.my[l2]), im2);
    float om2 = st.L[l2] * ii2;
    v2 n = mk(z.nx, z.ny);
    v2 t = rot90cw(n);
    v2 app1 = mk(0.f, 0.f), app2 = mk(0.f, 0.f);
    float rot1 = 0.f, rot2 = 0.f;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(z.flags & (SL_P0 << p))) continue;
        float depth = z.depth[p];
        if (depth >= 0.f) { z.apn[p] = 0.f; z.apt[p] = 0.f; continue; }
        v2 r1 = mk(z.r1x[p], z.r1y[p]), r2 = mk(z.r2x[p], z.r2y[p]);
        v2 v1 = add2(vl1, mk(-r1.y * om1, r1.x * om1));
        v2 v2v = add2(vl2, mk(-r2.y * om2, r2.x * om2));
        v2 dv = sub2(v1, v2v);
        float bias = BAUMGARTE * maxn(0.0f, (-depth) - BAUMGARTE_SLOP) / dt;
        float num = dot2(dv, n) + bias;
        float pn = num * z.mn[p];
        if (pn < MIN_MANIFOLD_IMPULSE) continue;
        num = dot2(dv, t);
        float pt = num * z.mt[p];
        float acc_pn = z.apn[p], acc_pt = z.apt[p];
        float new_pn = maxn(0.0f, acc_pn + pn);
        float applied_pn = new_pn - acc_pn;
        acc_pn = new_pn;
        float max_pt = z.fric * absf(acc_pn);
        float new_pt = clampn(acc_pt + pt, -max_pt, max_pt);
        float applied_pt = new_pt - acc_pt;
        z.apn[p] = acc_pn;
        z.apt[p] = new_pt;
        v2 dp = add2(scale2(n, applied_pn), scale2(t, applied_pt));
        if (!(z.flags & SL_S1)) { app1.x -= dp.x; app1.y -= dp.y; rot1 -= cross2(r1, dp); }
        if (!(z.flags & SL_S2)) { app2.x += dp.x; app2.y += dp.y; rot2 += cross2(r2, dp); }
    }
    if (z.flags & SL_W1) { st.mx[l1] += app1.x; st.my[l1] += app1.y; st.L[l1] += rot1; }
    if (z.flags & SL_W2) { st.mx[l2] += app2.x; st.my[l2] += app2.y; st.L[l2] += rot2; }
}

__device__ __forceinline__ void slot_store(const Slot& z, const StepDev& S)
{
    if (!(z.flags & SL_USED) || (z.flags & SL_JOINT)) return;
    const size_t cm = S.cap_m;
    float* P = S.m_pt;
    const uint32_t m = z.m;
    S.m_fric[m] = z.fric;
    S.m_prev1[m] = z.prev1;
    S.m_prev2[m] = z.prev2;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if (!(z.flags & (SL_P0 << p))) continue;
        P[(MP_ACC_PN * 2 + p) * cm + m] = z.apn[p];
        P[(MP_ACC_PT * 2 + p) * cm + m] = z.apt[p];
        P[(MP_R1X * 2 + p) * cm + m] = z.r1x[p];
        P[(MP_R1Y * 2 + p) * cm + m] = z.r1y[p];
        P[(MP_R2X * 2 + p) * cm + m] = z.r2x[p];
        P[(MP_R2Y * 2 + p) * cm + m] = z.r2y[p];
        P[(MP_DEPTH * 2 + p) * cm + m] = z.depth[p];
        P[(MP_MASS_N * 2 + p) * cm + m] = z.mn[p];
        P[(MP_MASS_T * 2 + p) * cm + m] = z.mt[p];
    }
}

constexpr uint32_t SPIN_LIMIT = 1u << 22;

template <bool LDS, int NSLOT, int MAXT>
__global__ __launch_bounds__(MAXT) void k_solve(StepDev S, float dt, uint32_t sub_steps, uint32_t iters)
{
    extern __shared__ uint32_t smem[];
    __shared__ uint32_t s_abort;
    if (S.counters[C_OVERFLOW]) return;
    const uint32_t w = blockIdx.x;
    if (S.w_status[w] != 0) return;
    if (threadIdx.x == 0) s_abort = 0u;
    const WorldDev& D = S.worlds[w];
    const int n = D.nbodies;
    const uint32_t base = D.body_base;
    const uint32_t nj = D.njoints;
    const uint32_t mbeg = S.w_man_begin[w];
    const uint32_t ibase = S.w_item_base[w];
    const uint32_t nitems = S.w_item_base[w + 1] - ibase;
    const float fsub = (float)sub_steps;
    const float sub_dt = dt / fsub;

    SolverState<LDS> st;
    if (LDS) {
        st.mx = (float*)smem;
        st.my = st.mx + n;
        st.L = st.my + n;
        st.im = st.L + n;
        st.ii = st.im + n;
        st.cnt = (uint32_t*)(st.ii + n);
        st.deg = st.cnt + n;
    } else {
        st.mx = D.mx; st.my = D.my; st.L = D.L;
        st.im = S.g_im + base; st.ii = S.g_ii + base;
        st.cnt = S.g_cnt + base; st.deg = S.deg + base;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        bool stat = (D.flags[i] & BF_STATIC) != 0;
        if (LDS) { st.mx[i] = D.mx[i]; st.my[i] = D.my[i]; st.L[i] = D.L[i]; }
        st.im[i] = stat ? 0.0f : 1.0f / D.mass[i];
        st.ii[i] = stat ? 0.0f : 1.0f / D.inertia[i];
        st.deg[i] = S.deg_off[base + i + 1] - S.deg_off[base + i];
    }
    __syncthreads();

    // register path when every item of this world has an owner slot
    const bool use_reg = NSLOT > 0 && nitems <= (uint32_t)NSLOT * blockDim.x;
    Slot sl[NSLOT > 0 ? NSLOT : 1];
    uint32_t nslot = 0;
    if (use_reg) {
#pragma unroll
        for (int z = 0; z < (NSLOT > 0 ? NSLOT : 1); ++z) {
            sl[z].flags = 0u;
            uint32_t pos = threadIdx.x + (uint32_t)z * blockDim.x;
            if (pos < nitems) {
                slot_load(sl[z], S, D, st, pos, ibase, nj, mbeg, base);
                nslot = (uint32_t)z + 1u;
            }
        }
    }

    const uint32_t M = S.w_man_begin[w + 1] - mbeg;
    float* P = S.m_pt;
    const size_t cm = S.cap_m;
    for (uint32_t s = 0; s < sub_steps; ++s) {
        // forces (lib.zig:200-205), AABB refresh + velocity integration (lib.zig:207-216)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            uint32_t f = D.flags[i];
            bool stat = (f & BF_STATIC) != 0;
            float fx = D.fx[i], fy = D.fy[i];
            for (int g = 0; g < D.ngens; ++g) {
                const Gen& G = D.gens[g];
                if (stat) continue;
                if (G.type == 0) {
                    fx += G.gx * D.mass[i];
                    fy += G.gy * D.mass[i];
                } else if (G.slot == i) {
                    float inv_m = 1.0f / D.mass[i];
                    v2 v = scale2(mk(st.mx[i], st.my[i]), inv_m);
                    v2 d = sub2(mk(D.px[i], D.py[i]), mk(G.ax, G.ay));
                    fx += (-G.k) * d.x - G.c * v.x;
                    fy += (-G.k) * d.y - G.c * v.y;
                }
            }
            if (s + 1 == sub_steps) {
                // Disc.updateAABB / Rectangle.updateAABB; only the last refresh is observable
                float a = D.ang[i];
                float hw, hh;
                if (f & BF_RECT) {
                    float c = cosf_r(a), sn = sinf_r(a);
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
                D.ax[i] = D.px[i]; D.ay[i] = D.py[i]; D.ahw[i] = hw; D.ahh[i] = hh;
            }
            if (!stat) {
                st.mx[i] += fx * sub_dt;
                st.my[i] += fy * sub_dt;
                st.L[i] += D.tq[i] * sub_dt;
            }
            D.fx[i] = fx; D.fy[i] = fy;
            st.cnt[i] = 0u;
        }
        __syncthreads();
        // updateTGSDepth + preStep (collision.zig:73-133)
        if (use_reg) {
#pragma unroll
            for (int z = 0; z < (NSLOT > 0 ? NSLOT : 1); ++z)
                if ((sl[z].flags & SL_USED) && !(sl[z].flags & SL_JOINT)) slot_tgs(sl[z], D, st);
        }
        for (uint32_t mm = threadIdx.x; !use_reg && mm < M; mm += blockDim.x) {
            uint32_t m = mbeg + mm;
            uint32_t l1 = S.m_ref[m] - base, l2 = S.m_inc[m] - base;
            v2 n = mk(S.m_nx[m], S.m_ny[m]);
            uint32_t valid = S.m_valid[m];
            float a1 = D.ang[l1], a2 = D.ang[l2];
            float prev1 = S.m_prev1[m], prev2 = S.m_prev2[m];
            for (int p = 0; p < 2; ++p) {
                if (!(valid & (1u << p))) continue;
                float depth = P[(MP_DEPTH * 2 + p) * cm + m], od = P[(MP_ODEPTH * 2 + p) * cm + m];
                if (approx(depth, od, 1e-4f)) continue;
                v2 r1 = mk(P[(MP_R1X * 2 + p) * cm + m], P[(MP_R1Y * 2 + p) * cm + m]);
                v2 r2 = mk(P[(MP_R2X * 2 + p) * cm + m], P[(MP_R2Y * 2 + p) * cm + m]);
                v2 rr1 = rotate2(r1, a1 - prev1), rr2 = rotate2(r2, a2 - prev2);
                v2 q1 = add2(rr1, mk(D.px[l1], D.py[l1]));
                v2 q2 = add2(rr2, mk(D.px[l2], D.py[l2]));
                float dd = dot2(n, sub2(q2, q1));
                P[(MP_DEPTH * 2 + p) * cm + m] = dd + od;
                P[(MP_R1X * 2 + p) * cm + m] = rr1.x; P[(MP_R1Y * 2 + p) * cm + m] = rr1.y;
                P[(MP_R2X * 2 + p) * cm + m] = rr2.x; P[(MP_R2Y * 2 + p) * cm + m] = rr2.y;
            }
            S.m_prev1[m] = a1; S.m_prev2[m] = a2;
            float im1 = st.im[l1], im2 = st.im[l2], ii1 = st.ii[l1], ii2 = st.ii[l2];
            float inv_mass = im1 + im2;
            v2 t = rot90cw(n);
            S.m_fric[m] = sqrt_rn(D.mu[l1] * D.mu[l2]);
            for (int p = 0; p < 2; ++p) {
                if (!(valid & (1u << p))) continue;
                v2 r1 = mk(P[(MP_R1X * 2 + p) * cm + m], P[(MP_R1Y * 2 + p) * cm + m]);
                v2 r2 = mk(P[(MP_R2X * 2 + p) * cm + m], P[(MP_R2Y * 2 + p) * cm + m]);
                float r1n = cross2(r1, n), r2n = cross2(r2, n);
                float kn = inv_mass + ii1 * (r1n * r1n) + ii2 * (r2n * r2n);
                float r1t = cross2(r1, t), r2t = cross2(r2, t);
                float kt = inv_mass + ii1 * (r1t * r1t) + ii2 * (r2t * r2t);
                P[(MP_MASS_N * 2 + p) * cm + m] = kn > 0.0f ? 1.0f / kn : 0.0f;
                P[(MP_MASS_T * 2 + p) * cm + m] = kt > 0.0f ? 1.0f / kt : 0.0f;
            }
        }
        __syncthreads();
        // K iterations of (constraints, manifolds) as one ordered dataflow
        if (use_reg) {
            uint32_t k = 0, cz = 0, spins = 0;
            if (nslot == 0) k = iters;
            while (k < iters) {
                if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                bool ready = false;
#pragma unroll
                for (int z = 0; z < (NSLOT > 0 ? NSLOT : 1); ++z) {
                    if ((uint32_t)z == cz) {
                        ready = true;
                        if (sl[z].r0 >= 0) ready = ld_acq(&st.cnt[sl[z].r0]) == k * sl[z].dg0 + sl[z].rk0;
                        if (ready && sl[z].r1 >= 0) ready = ld_acq(&st.cnt[sl[z].r1]) == k * sl[z].dg1 + sl[z].rk1;
                    }
                }
                if (ready) {
#pragma unroll
                    for (int z = 0; z < (NSLOT > 0 ? NSLOT : 1); ++z) {
                        if ((uint32_t)z == cz) {
                            if (sl[z].flags & SL_JOINT) solve_joint(D, st, D.joints[sl[z].j], sub_dt);
                            else slot_impulses(sl[z], st, sub_dt);
                            if (sl[z].r0 >= 0) st_rel(&st.cnt[sl[z].r0], k * sl[z].dg0 + sl[z].rk0 + 1u);
                            if (sl[z].r1 >= 0) st_rel(&st.cnt[sl[z].r1], k * sl[z].dg1 + sl[z].rk1 + 1u);
                        }
                    }
                    spins = 0;
                    if (++cz == nslot) { cz = 0; ++k; }
                } else {
                    if (++spins > SPIN_LIMIT) {
                        __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        S.w_status[w] = 5;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        } else {
            const uint32_t total = iters * nitems;
            uint32_t q = threadIdx.x;
            uint32_t k = 0, j = 0;
            int32_t r0 = -1, r1 = -1;
            uint32_t t0 = 0, t1 = 0;
            bool have = false;
            uint32_t spins = 0;
            while (true) {
                if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                if (!have) {
                    if (q >= total) break;
                    k = q / nitems;
                    j = S.it_perm[ibase + (q - k * nitems)];
                    uint32_t item = ibase + j;
                    r0 = S.it_r0[item]; r1 = S.it_r1[item];
                    if (r0 >= 0) { r0 -= (int32_t)base; t0 = k * st.deg[r0] + S.it_k0[item]; }
                    if (r1 >= 0) { r1 -= (int32_t)base; t1 = k * st.deg[r1] + S.it_k1[item]; }
                    have = true;
                }
                bool ready = true;
                if (r0 >= 0) ready = ld_acq(&st.cnt[r0]) == t0;
                if (ready && r1 >= 0) ready = ld_acq(&st.cnt[r1]) == t1;
                if (ready) {
                    if (j < nj) {
                        solve_joint(D, st, D.joints[j], sub_dt);
                    } else {
                        uint32_t m = mbeg + (j - nj);
                        uint32_t l1 = S.m_ref[m] - base, l2 = S.m_inc[m] - base;
                        bool wr1 = (S.sflag[base + l1] & SF_RESOURCE) != 0;
                        bool wr2 = (S.sflag[base + l2] & SF_RESOURCE) != 0;
                        solve_manifold(S, D, st, m, l1, l2, wr1, wr2, sub_dt);
                    }
                    if (r0 >= 0) st_rel(&st.cnt[r0], t0 + 1u);
                    if (r1 >= 0) st_rel(&st.cnt[r1], t1 + 1u);
                    have = false;
                    spins = 0;
                    q += blockDim.x;
                } else {
                    // every item's predecessors precede it, so some item is always
                    // ready; a stall this long means corrupt ordering data: stop
                    if (++spins > SPIN_LIMIT) {
                        __hip_atomic_store(&s_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        S.w_status[w] = 5;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();
        // position integration (lib.zig:238-249)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
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
    if (use_reg) {
#pragma unroll
        for (int z = 0; z < (NSLOT > 0 ? NSLOT : 1); ++z) slot_store(sl[z], S);
    }
    if (LDS) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            D.mx[i] = st.mx[i]; D.my[i] = st.my[i]; D.L[i] = st.L[i];
        }
    }
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

typedef void (*SolveFn)(StepDev, float, uint32_t, uint32_t);
struct SolveVariant {
    SolveFn lds_fn;
    unsigned threads;
    int slots;
};
// solve kernel variants: register slots per thread x workgroup size
static const SolveVariant kSolveVariants[SOLVE_VARIANTS] = {
    {k_solve<true, 0, 1024>, 1024, 0},
    {k_solve<true, 4, 256>, 256, 4},
    {k_solve<true, 2, 512>, 512, 2},
    {k_solve<true, 2, 256>, 256, 2},
    {k_solve<true, 0, 256>, 256, 0},
};

unsigned solve_variant_threads(int v) { return kSolveVariants[v < SOLVE_VARIANTS && v >= 0 ? v : 0].threads; }
int solve_variant_slots(int v) { return kSolveVariants[v < SOLVE_VARIANTS && v >= 0 ? v : 0].slots; }

void launch_phase(const StepDev& S, hipStream_t st, int phase, float dt, uint32_t sub_steps, uint32_t iters,
                  uint32_t max_bodies, bool lds, int solve_variant)
{
    const unsigned BS = 256;
    const unsigned GRID = 1024;   // grid-stride kernels over device-sized ranges
    switch (phase) {
    case PH_BROAD:
        hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, st, S.counters);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nworlds + 1, BS)), dim3(BS), 0, st,
                           reinterpret_cast<uint32_t*>(S.w_status), S.nworlds + 1);
        hipLaunchKernelGGL(k_fill_bworld, dim3(S.nworlds), dim3(256), 0, st, S);
        hipLaunchKernelGGL(k_prep, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
        scan_u32(S.ncell, S.cell_off, nullptr, S.nb, S.nb, &S.counters[C_ENTRIES], S.partials, st);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nbuckets + 1, BS)), dim3(BS), 0, st, S.bkt_cnt,
                           S.nbuckets + 1);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nbuckets + 1, BS)), dim3(BS), 0, st, S.bkt_fill,
                           S.nbuckets + 1);
        hipLaunchKernelGGL(k_entries, dim3(GRID), dim3(BS), 0, st, S);
        scan_u32(S.bkt_cnt, S.bkt_off, nullptr, S.nbuckets, S.nbuckets, nullptr, S.partials, st);
        hipLaunchKernelGGL(k_bucket_fill, dim3(GRID), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_bucket_sort, dim3(blocks_for(S.nbuckets, BS)), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_ent_cnt, dim3(GRID), dim3(BS), 0, st, S);
        scan_u32(S.ent_cnt, S.ent_off, &S.counters[C_ENTRIES], 0, S.cap_e, &S.counters[C_RAW], S.partials, st);
        hipLaunchKernelGGL(k_cand_emit, dim3(GRID), dim3(BS), 0, st, S);
        break;
    case PH_NARROW:
        hipLaunchKernelGGL(k_narrow, dim3(GRID), dim3(BS), 0, st, S);
        scan_u32(S.cand_flag, S.cand_idx, &S.counters[C_RAW], 0, S.cap_r, &S.counters[C_MANIFOLDS], S.partials, st);
        hipLaunchKernelGGL(k_man_build, dim3(GRID), dim3(BS), 0, st, S);
        break;
    case PH_SCHED:
        hipLaunchKernelGGL(k_world_ranges, dim3(blocks_for(S.nworlds + 1, BS)), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_joint_mark, dim3(blocks_for(S.njoints, BS)), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nb + 1, BS)), dim3(BS), 0, st, S.deg, S.nb + 1);
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(S.nb + 1, BS)), dim3(BS), 0, st, S.deg_fill, S.nb + 1);
        hipLaunchKernelGGL(k_items_joints, dim3(blocks_for(S.njoints, BS)), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_items_mans, dim3(GRID), dim3(BS), 0, st, S);
        scan_u32(S.deg, S.deg_off, nullptr, S.nb, S.nb, nullptr, S.partials, st);
        hipLaunchKernelGGL(k_color, dim3(blocks_for(S.nworlds, 64)), dim3(64), 0, st, S);
        hipLaunchKernelGGL(k_list_fill, dim3(GRID), dim3(BS), 0, st, S);
        hipLaunchKernelGGL(k_list_rank, dim3(blocks_for(S.nb, BS)), dim3(BS), 0, st, S);
        break;
    default: {
        const SolveVariant& v = kSolveVariants[solve_variant < SOLVE_VARIANTS ? solve_variant : 0];
        if (lds) {
            size_t sm = solve_lds_bytes((int)max_bodies);
            hipLaunchKernelGGL(v.lds_fn, dim3(S.nworlds), dim3(v.threads), sm, st, S, dt, sub_steps, iters);
        } else {
            hipLaunchKernelGGL((k_solve<false, 0, 1024>), dim3(S.nworlds), dim3(1024), 0, st, S, dt, sub_steps, iters);
        }
        hipLaunchKernelGGL(k_commit_step, dim3(1), dim3(64), 0, st, S.counters);
        break;
    }
    }
}

void set_solve_smem_limit()
{
    for (int v = 0; v < SOLVE_VARIANTS; ++v)
        (void)hipFuncSetAttribute((const void*)kSolveVariants[v].lds_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
}

}  // namespace r2d
