// Host runtime of the C ABI in include/r2d.h: contexts, worlds (a host mirror
// plus resident device state), and the batched step.
//
// Scene construction restates EntityFactory (src/core/lib.zig:13-130) and the
// shape constructors (Bodies/Disc.zig:29-55, Bodies/Rectangle.zig:33-64) in
// f32 with the reference's operation order, so the initial state has the same
// bits as the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <deque>
#include <mutex>
#include <unordered_map>
#include <new>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/r2d.h"
#include "r2d_device.h"
#include "r2d_kernels.h"

using namespace r2d;

#define R2D_EXPORT extern "C" __attribute__((visibility("default")))

static thread_local char g_last_error[512] = "";
static std::atomic<uint64_t> g_epoch{0};

static void note_error(const char* what, hipError_t e, int line)
{
    std::snprintf(g_last_error, sizeof g_last_error, "%s failed: %s (%d) at r2d_api.hip:%d", what,
                  hipGetErrorString(e), (int)e, line);
    if (std::getenv("R2D_DEBUG")) std::fprintf(stderr, "[r2d] %s\n", g_last_error);
}

#define HIPCHK(x)                                        \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) {                          \
            note_error(#x, e_, __LINE__);                \
            return R2D_E_DEVICE;                         \
        }                                                \
    } while (0)

// ------------------------------------------ step scratch of r2d_process_static
// A file-scope device array holds the step scratch of r2d_process_static, the
// entry point for callers that cannot allocate (code under graph capture,
// serving loops that forbid allocation): every array of the fused step for
// one world of up to 1,024 bodies.  Each GPU holds its own copy for the life
// of the process; host code uses only the address hipGetSymbolAddress gave
// for the current device (taken at the entry point's first call on a
// device), never the array's name.  The array has external linkage: the HIP
// runtime finds a device variable by its global symbol in the code object,
// and a `static` one is a local symbol there (its lookup aborted the
// process); hidden visibility keeps the name out of the library's exports.
constexpr size_t kStaticScratchBytes = (size_t)8 << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_setup;
void* g_scratch_ptr[kMaxDevices] = {};     // device address of `scratch`, per device
std::mutex g_scratch_use[kMaxDevices];     // one r2d_process_static at a time per device
uint64_t g_scratch_call[kMaxDevices] = {}; // number of the last call that wrote it, per device
std::atomic<uint64_t> g_scratch_calls{0};

// capacities of the static layout: bodies, buckets, cell entries, raw
// candidate pairs, manifolds, joints, solver items
constexpr uint32_t ST_N = FUSED_NMAX;
constexpr uint32_t ST_B = 4 * ST_N;
constexpr uint32_t ST_E = 16 * ST_N;
constexpr uint32_t ST_R = 64 * ST_N;
constexpr uint32_t ST_M = 4 * ST_N;
constexpr uint32_t ST_J = 2 * ST_N;
constexpr uint32_t ST_I = ST_M + ST_J;

// Carves the step scratch of one world (every StepDev array, sized for the
// capacities above) and its descriptor out of the static array at `base`.
// Returns the bytes used; a null base only measures.
size_t static_layout(char* base, StepDev* S, WorldDev** desc)
{
    const size_t N = ST_N, B = ST_B, E = ST_E, R = ST_R, M = ST_M, I = ST_I, W = 1;
    size_t o = 0;
    auto take = [&](auto** p, size_t bytes) {
        using T = std::remove_reference_t<decltype(*p)>;
        *p = base ? reinterpret_cast<T>(base + o) : nullptr;
        o += (bytes + 255) & ~(size_t)255;
    };
    StepDev& F = *S;
    F = StepDev{};
    take(desc, sizeof(WorldDev));
    take(&F.bworld, N * 4);
    take(&F.cx0, N * 8); take(&F.cy0, N * 8);
    take(&F.cnx, N * 4); take(&F.cny, N * 4);
    take(&F.ncell, (N + 1) * 4); take(&F.cell_off, (N + 1) * 4);
    take(&F.cc, N * 4); take(&F.cs, N * 4); take(&F.sflag, N * 4);
    take(&F.bkt_cnt, (B + 1) * 4); take(&F.bkt_off, (B + 1) * 4); take(&F.bkt_fill, (B + 1) * 4);
    take(&F.bkt_items, (E + 1) * 4); take(&F.bkt_sorted, (E + 1) * 4);
    take(&F.ent_body, (E + 1) * 4); take(&F.ent_bkt, (E + 1) * 4);
    take(&F.ent_cnt, (E + 1) * 4); take(&F.ent_off, (E + 1) * 4);
    take(&F.cand, (R + 1) * 16);
    take(&F.cand_flag, (R + 1) * 4); take(&F.cand_idx, (R + 1) * 4);
    take(&F.sat_nx, (R + 1) * 4); take(&F.sat_ny, (R + 1) * 4); take(&F.sat_info, (R + 1) * 4);
    take(&F.m_ref, M * 4); take(&F.m_inc, M * 4);
    take(&F.m_nx, M * 4); take(&F.m_ny, M * 4); take(&F.m_fric, M * 4);
    take(&F.m_prev1, M * 4); take(&F.m_prev2, M * 4); take(&F.m_valid, M * 4);
    take(&F.m_pt, (size_t)MP_NFIELDS * 2 * M * 4);
    take(&F.w_man_begin, (W + 1) * 4); take(&F.w_item_base, (W + 1) * 4);
    take(&F.w_status, (W + 1) * 4); take(&F.w_order, (W + 1) * 4);
    take(&F.it_r0, I * 4); take(&F.it_r1, I * 4); take(&F.it_k0, I * 4); take(&F.it_k1, I * 4);
    take(&F.deg, (N + 1) * 4); take(&F.deg_off, (N + 1) * 4); take(&F.deg_fill, (N + 1) * 4);
    take(&F.lists, 2 * I * 4);
    take(&F.it_color, I * 4); take(&F.it_perm, I * 4);
    take(&F.recs, I * 128);
    take(&F.cmask, 2 * N * 8);
    take(&F.g_im, N * 4); take(&F.g_ii, N * 4); take(&F.g_cnt, N * 4);
    take(&F.cmap, N * 4); take(&F.cinv, N * 4); take(&F.w_nc, (W + 1) * 4);
    take(&F.counters, 64 * 4);
    take(&F.partials, 1024 * 4);
    take(&F.wscr, W * sizeof(WorldScratch));
    take(&F.fpool, 8 * 4);
    take(&F.w_done, W * 4);
    take(&F.w_cnt, W * 32);
    take(&F.w_man_end, W * 4);
    F.cap_e = ST_E; F.cap_r = ST_R; F.cap_m = ST_M; F.cap_i = ST_I;
    F.nb = 0; F.nworlds = 1; F.nbuckets = 0; F.njoints = 0;
    F.worlds = base ? *desc : nullptr;
    F.w_tstamp = nullptr; F.tcap = 0; F.w_prof = nullptr;
    return o;
}

// the address of `scratch` on the current device, taken once per device
int static_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return R2D_E_CAPACITY;
    std::lock_guard<std::mutex> g(g_scratch_setup);
    if (g_scratch_ptr[device]) return R2D_OK;
    StepDev S;
    WorldDev* d = nullptr;
    if (static_layout(nullptr, &S, &d) > kStaticScratchBytes) return R2D_E_CAPACITY;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();   // not an error of a later call
        std::snprintf(g_last_error, sizeof g_last_error, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                      hipGetErrorString(e));
        return R2D_E_DEVICE;
    }
    g_scratch_ptr[device] = p;
    return R2D_OK;
}

}  // namespace

namespace {

// body arrays of the packed host/device block, each `cap` long
enum BodyField {
    F_PX = 0, F_PY, F_ANG, F_MX, F_MY, F_L, F_FX, F_FY, F_TQ, F_MASS, F_INERTIA, F_MU, F_RAD, F_HW, F_HH,
    F_AX, F_AY, F_AHW, F_AHH, F_ID, F_FLAGS, F_NFIELDS
};

struct HJoint {
    int32_t type;
    uint32_t id1, id2;
    float pmax, pmin, beta, r1x, r1y, r2x, r2y, target, tx, ty;
};

struct HGen {
    int32_t type;
    float gx, gy;
    uint32_t id;
    float k, c, ax, ay;
};

}  // namespace

struct r2d_world {
    r2d_ctx* ctx = nullptr;
    r2d_ctx* owner = nullptr;   // the context that created it (ctx moves to a serving lane while served)
    uint32_t current_body_id = 0;
    int32_t n = 0;
    size_t cap = 0;
    std::vector<uint32_t> block;       // F_NFIELDS * cap words (floats stored by bit pattern)
    std::vector<float> width, height;  // rectangle shape (host only)
    std::vector<int32_t> id2slot;
    std::vector<HJoint> joints;
    std::vector<HGen> gens;
    std::vector<uint64_t> excl;        // canonical (min << 32 | max), sorted, unique
    float cell_arg = 4.0f;
    size_t mult_arg = 2;
    int hash_mode = 0;
    float hash_cell = 4.0f;
    size_t hash_mult = 2;
    int sched = 0;
    float warm = 0.0f;        // warm starting: fraction of the previous step's impulses carried (0: off)
    bool warm_clear = true;   // the warm tables are to be emptied before the next step
    // device
    uint32_t* dblock = nullptr;
    size_t dcap = 0;
    Joint* djoints = nullptr;
    size_t djcap = 0;
    Gen* dgens = nullptr;
    size_t dgcap = 0;
    uint64_t* dexcl = nullptr;
    size_t dxcap = 0;
    uint4* dwtab = nullptr;   // warm tables (WARM_K entries of 3 uint4 per body), counts, per-body sums
    uint32_t* dwcnt = nullptr;
    float4* dwsum = nullptr;
    size_t dwcap = 0;
    bool host_valid = true;   // host mirror is current
    bool dev_valid = false;   // device body block is current
    bool topo_dirty = true;   // joints / generators / exclusions need upload
    bool in_service = false;  // held by a server until its request completes
    uint64_t version = 1;     // bumped by every edit and device reallocation
    uint64_t desc_version = 0, checked_version = 0;
    WorldDev desc{};          // descriptor of `desc_version` (batch offsets aside)
    // manifolds of the last step this world took part in
    uint64_t epoch = 0;
    uint32_t batch_index = 0;
    bool fused_src = false;   // ... computed by the fused kernel (their range: w_man_begin / w_man_end)
    bool prestepped = true;   // ... by a step with sub-steps (a step without runs no preStep)
    uint64_t sepoch = 0;      // ... or by r2d_process_static call number `sepoch` (its scratch holds them)
    // capacities of the world's slices of the fused step scratch (cell
    // entries, raw candidates, manifolds, solver items); 0: initial estimate
    uint32_t fz[4] = {0, 0, 0, 0};

    float* f(int field) { return reinterpret_cast<float*>(block.data() + (size_t)field * cap); }
    uint32_t* u(int field) { return block.data() + (size_t)field * cap; }
    int32_t slot_of(uint32_t id) const { return id < id2slot.size() ? id2slot[id] : -1; }
};

struct r2d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[2] = {nullptr, nullptr};      // side streams of the solve phase
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;   // brackets of the last launched batch
    hipEvent_t ev_kend = nullptr;                      // end of the last fused kernel
    r2d_server* server = nullptr;
    uint64_t live_worlds = 0;   // worlds created on this context and not yet destroyed
    // step scratch
    void* arena = nullptr;
    size_t arena_bytes = 0;
    StepDev S{};
    uint32_t cap_n = 0, cap_b = 0, cap_w = 0, cap_j = 0;
    uint32_t cap_e = 0, cap_r = 0, cap_m = 0, cap_i = 0;
    WorldDev* dworlds = nullptr;
    size_t dworlds_cap = 0;
    uint32_t* hcounters = nullptr;     // pinned
    int32_t* hstatus = nullptr;        // pinned
    size_t hstatus_cap = 0;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;              // phase events of timed steps not yet summed
    size_t ev_used = 0;
    double phase_ms[4] = {0, 0, 0, 0};
    std::vector<float> step_ms;                   // device time of each timed step
    uint64_t phase_steps = 0;
    uint64_t epoch = 0;
    uint64_t last_counts[4] = {0, 0, 0, 0};
    // captured step: replayed while the batch layout and launch settings stay the same
    bool use_graph = true;
    hipGraphExec_t graph_exec = nullptr;          // whole step
    hipGraphExec_t phase_exec[PH_N] = {};         // one per phase (timed steps)
    std::vector<uint8_t> graph_key;
    uint64_t graph_uses = 0;                      // launches with the current key
    std::vector<WorldDev> uploaded_descs;
    int solve_variant = -1;                       // -1: choose by world size
    uint32_t lds_nmax = 0;                        // 0: LDS body state up to the LDS capacity
    bool wide_record = true;                      // batches of <= 64 LDS worlds may take the 512-lane record kernel

    KernelTrace* trace = nullptr;                 // per-kernel timing (R2D_TRACE_KERNELS)

    // whole steps per workgroup (k_fused) for batches of small worlds
    bool fused_enabled = true;
    void* farena = nullptr;                       // the worlds' slices (entries, candidates, manifolds, items)
    size_t farena_bytes = 0;
    uint32_t fcap_e = 0, fcap_r = 0, fcap_m = 0, fcap_i = 0;   // slice totals the arena holds
    void* fworld = nullptr;                       // per-world arrays of the fused kernel
    size_t fworld_cap = 0;
    StepDev SF{};                                 // the fused kernel's view: S with the slice arrays in farena
    std::vector<WorldScratch> fscr;               // current slices (host copy)
    std::vector<WorldScratch> fscr_uploaded;
    std::vector<uint32_t> forder, forder_uploaded;
    int num_cus = 0;                              // compute units of the device
    uint32_t* hdone = nullptr;                    // pinned: steps done per world, last-step / needed counts
    uint32_t* hcnt = nullptr;
    WorldScratch* hscr = nullptr;                 // pinned: the slice table as the last launch left it
    uint32_t* dfsum = nullptr;                    // a fused launch's summary gathered on the device (fsum_words)
    uint32_t* hfsum = nullptr;                    // pinned: the same, copied back in one piece
    size_t hfused_cap = 0;
    uint64_t* tstamp = nullptr;                   // per-world step times of a profiled fused call
    size_t tstamp_words = 0;
    uint64_t* wprof = nullptr;                    // per-world phase ticks of a profiled fused call
    size_t wprof_worlds = 0;
    double fused_phase_ms[FP_N] = {};             // summed over the worlds' profiled steps
    std::vector<double> fused_step_ms;            // world-step latencies of profiled fused calls
    double fused_ms = 0.0;                        // device time of profiled fused launches
    uint64_t fused_launches = 0;
    uint64_t fused_steps = 0;                     // world-steps of profiled fused launches
    uint64_t fused_sums[5] = {};                  // their bodies, entries, raw candidates, manifolds, items
    void* ingest = nullptr;                       // r2d_make_bodies' inputs on the device
    size_t ingest_cap = 0;                        // bodies they hold
    // device blocks of this context's worlds, reused across worlds: hipFree
    // waits for the whole device, so a world destroyed while other work runs
    // (a served scene retired beside running lanes) would stall the host on
    // every destroy.  Blocks come in power-of-two size classes; free ones wait
    // in per-class lists until a world needs one or the context goes.
    std::unordered_map<size_t, std::vector<void*>> pool_free;
    std::unordered_map<void*, size_t> pool_class;  // every block handed out -> its class
    size_t pool_cached = 0;                        // bytes in the free lists

    Streams streams() const { return Streams{stream, {aux[0], aux[1]}, fork, {join[0], join[1]}, trace}; }
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// words of a fused launch's summary for nw worlds (k_fused_gather): the 64
// counters, then per world status, steps done, 8 counts, 8-word slice entry
constexpr size_t FSUM_WORLD_WORDS = 2 + 8 + sizeof(WorldScratch) / 4;
static_assert(sizeof(WorldScratch) == 32, "the summary packs 8-word slice table entries");
size_t fsum_words(size_t nw) { return 64 + FSUM_WORLD_WORDS * nw; }

// the free lists hold at most this many bytes; a block released beyond it is
// freed (with hipFree's wait)
constexpr size_t kPoolCacheMax = (size_t)8 << 30;

size_t pool_class_of(size_t bytes)
{
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

// a device block of at least `bytes` for a world of context c
int pool_alloc(r2d_ctx* c, void** out, size_t bytes)
{
    const size_t cls = pool_class_of(bytes);
    std::vector<void*>& fl = c->pool_free[cls];
    void* p = nullptr;
    if (!fl.empty()) {
        p = fl.back();
        fl.pop_back();
        c->pool_cached -= cls;
    } else if (hipMalloc(&p, cls) != hipSuccess) {
        // the cached blocks of other classes back to the device, then once more
        (void)hipGetLastError();
        for (auto& kv : c->pool_free)
            for (void* q : kv.second) (void)hipFree(q);
        c->pool_free.clear();
        c->pool_cached = 0;
        p = nullptr;
        if (hipMalloc(&p, cls) != hipSuccess) {
            (void)hipGetLastError();
            return R2D_E_OOM;
        }
    }
    c->pool_class[p] = cls;
    *out = p;
    return R2D_OK;
}

// a block from pool_alloc back to its free list; no device work may use it
// any more (the caller has waited for the streams that did)
void pool_release(r2d_ctx* c, void* p)
{
    if (!p) return;
    auto it = c->pool_class.find(p);
    if (it == c->pool_class.end()) {   // not from the pool
        (void)hipFree(p);
        return;
    }
    const size_t cls = it->second;
    c->pool_class.erase(it);
    if (c->pool_cached + cls > kPoolCacheMax) {
        (void)hipFree(p);
        return;
    }
    c->pool_free[cls].push_back(p);
    c->pool_cached += cls;
}

void pool_destroy(r2d_ctx* c)
{
    for (auto& kv : c->pool_free)
        for (void* q : kv.second) (void)hipFree(q);
    c->pool_free.clear();
    c->pool_cached = 0;
}

void destroy_graphs(r2d_ctx* c);
int server_collect(r2d_server* sv);

int ensure_world_device(r2d_world* w)
{
    if (w->dcap < w->cap || !w->dblock) {
        if (w->dblock) {
            HIPCHK(hipStreamSynchronize(w->ctx->stream));
            pool_release(w->owner, w->dblock);
        }
        w->dblock = nullptr;
        size_t c = std::max<size_t>(w->cap, 16);
        if (pool_alloc(w->owner, (void**)&w->dblock, (size_t)F_NFIELDS * c * 4)) return R2D_E_OOM;
        w->dcap = c;
        w->dev_valid = false;
        w->version += 1;
    }
    return R2D_OK;
}

// warm-start tables of a warm-started world, as large as its device body
// block; a larger table keeps the old one's entries
int ensure_warm(r2d_world* w)
{
    if (!(w->warm > 0.0f)) return R2D_OK;
    hipStream_t st = w->ctx->stream;
    if (w->dwcap < w->dcap) {
        const size_t c = w->dcap;
        uint4* tab = nullptr;
        uint32_t* cnt = nullptr;
        float4* sum = nullptr;
        r2d_ctx* o = w->owner;
        if (pool_alloc(o, (void**)&tab, c * WARM_K * 3 * sizeof(uint4))) return R2D_E_OOM;
        if (pool_alloc(o, (void**)&cnt, c * 4)) { pool_release(o, tab); return R2D_E_OOM; }
        if (pool_alloc(o, (void**)&sum, c * sizeof(float4))) { pool_release(o, tab); pool_release(o, cnt); return R2D_E_OOM; }
        HIPCHK(hipMemsetAsync(cnt, 0, c * 4, st));
        if (w->dwcap && !w->warm_clear) {
            HIPCHK(hipMemcpyAsync(tab, w->dwtab, w->dwcap * WARM_K * 3 * sizeof(uint4), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(cnt, w->dwcnt, w->dwcap * 4, hipMemcpyDeviceToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        pool_release(o, w->dwtab);
        pool_release(o, w->dwcnt);
        pool_release(o, w->dwsum);
        w->dwtab = tab; w->dwcnt = cnt; w->dwsum = sum;
        w->dwcap = c;
        w->version += 1;
    } else if (w->warm_clear) {
        HIPCHK(hipMemsetAsync(w->dwcnt, 0, w->dwcap * 4, st));
    }
    w->warm_clear = false;
    return R2D_OK;
}

// host mirror <-> device block; the layouts match when dcap == cap
int upload_bodies(r2d_world* w)
{
    int rc = ensure_world_device(w);
    if (rc) return rc;
    if (w->dcap != w->cap) {
        // grow host block to the device capacity so one copy moves everything
        std::vector<uint32_t> nb((size_t)F_NFIELDS * w->dcap, 0u);
        for (int fld = 0; fld < F_NFIELDS; ++fld)
            std::memcpy(nb.data() + (size_t)fld * w->dcap, w->block.data() + (size_t)fld * w->cap,
                        (size_t)w->n * 4);
        w->block.swap(nb);
        w->cap = w->dcap;
        w->width.resize(w->cap);
        w->height.resize(w->cap);
    }
    HIPCHK(hipMemcpyAsync(w->dblock, w->block.data(), (size_t)F_NFIELDS * w->cap * 4, hipMemcpyHostToDevice,
                          w->ctx->stream));
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    w->dev_valid = true;
    return R2D_OK;
}

int download_bodies(r2d_world* w)
{
    if (w->host_valid) return R2D_OK;
    HIPCHK(hipMemcpyAsync(w->block.data(), w->dblock, (size_t)F_NFIELDS * w->cap * 4, hipMemcpyDeviceToHost,
                          w->ctx->stream));
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    w->host_valid = true;
    return R2D_OK;
}

// host edits need the current state; afterwards the device copy is stale
int begin_edit(r2d_world* w)
{
    if (w->in_service) return R2D_E_BUSY;
    int rc = download_bodies(w);
    if (rc) return rc;
    w->dev_valid = false;
    w->version += 1;
    return R2D_OK;
}

void grow_host(r2d_world* w, size_t need)
{
    if (need <= w->cap) return;
    size_t nc = std::max<size_t>(16, w->cap * 2);
    while (nc < need) nc *= 2;
    std::vector<uint32_t> nb((size_t)F_NFIELDS * nc, 0u);
    for (int fld = 0; fld < F_NFIELDS; ++fld)
        std::memcpy(nb.data() + (size_t)fld * nc, w->block.data() + (size_t)fld * w->cap, (size_t)w->n * 4);
    w->block.swap(nb);
    w->cap = nc;
    w->width.resize(nc);
    w->height.resize(nc);
}

// appendBody (lib.zig:66-71)
uint32_t append_body(r2d_world* w)
{
    grow_host(w, (size_t)w->n + 1);
    uint32_t id = w->current_body_id;
    if (id >= w->id2slot.size()) w->id2slot.resize(std::max<size_t>(64, (size_t)id * 2 + 2), -1);
    w->id2slot[id] = w->n;
    w->u(F_ID)[w->n] = id;
    w->n += 1;
    w->current_body_id += 1;
    return id;
}

int upload_topology(r2d_world* w)
{
    if (!w->topo_dirty) return R2D_OK;
    hipStream_t st = w->ctx->stream;
    // joints: ids resolved to current slots
    std::vector<Joint> js(w->joints.size());
    for (size_t i = 0; i < w->joints.size(); ++i) {
        const HJoint& h = w->joints[i];
        Joint& d = js[i];
        std::memset(&d, 0, sizeof d);
        d.type = h.type;
        d.s1 = w->slot_of(h.id1);
        d.s2 = (h.type <= 1) ? w->slot_of(h.id2) : -1;
        d.pmax = h.pmax; d.pmin = h.pmin; d.beta = h.beta;
        d.r1x = h.r1x; d.r1y = h.r1y; d.r2x = h.r2x; d.r2y = h.r2y;
        d.target = h.target; d.tx = h.tx; d.ty = h.ty;
    }
    const bool grow = js.size() > w->djcap || w->gens.size() > w->dgcap || w->excl.size() > w->dxcap;
    if (grow) HIPCHK(hipStreamSynchronize(st));   // the blocks replaced below are idle
    if (js.size() > w->djcap) {
        pool_release(w->owner, w->djoints);
        w->djoints = nullptr;
        w->djcap = std::max<size_t>(16, js.size() * 2);
        if (pool_alloc(w->owner, (void**)&w->djoints, w->djcap * sizeof(Joint))) { w->djcap = 0; return R2D_E_OOM; }
    }
    if (!js.empty()) HIPCHK(hipMemcpyAsync(w->djoints, js.data(), js.size() * sizeof(Joint), hipMemcpyHostToDevice, st));
    std::vector<Gen> gs(w->gens.size());
    for (size_t i = 0; i < w->gens.size(); ++i) {
        const HGen& h = w->gens[i];
        Gen& d = gs[i];
        d.type = h.type; d.gx = h.gx; d.gy = h.gy;
        d.slot = h.type == 1 ? w->slot_of(h.id) : -1;
        d.k = h.k; d.c = h.c; d.ax = h.ax; d.ay = h.ay;
    }
    if (gs.size() > w->dgcap) {
        pool_release(w->owner, w->dgens);
        w->dgens = nullptr;
        w->dgcap = std::max<size_t>(16, gs.size() * 2);
        if (pool_alloc(w->owner, (void**)&w->dgens, w->dgcap * sizeof(Gen))) { w->dgcap = 0; return R2D_E_OOM; }
    }
    if (!gs.empty()) HIPCHK(hipMemcpyAsync(w->dgens, gs.data(), gs.size() * sizeof(Gen), hipMemcpyHostToDevice, st));
    if (w->excl.size() > w->dxcap) {
        pool_release(w->owner, w->dexcl);
        w->dexcl = nullptr;
        w->dxcap = std::max<size_t>(16, w->excl.size() * 2);
        if (pool_alloc(w->owner, (void**)&w->dexcl, w->dxcap * sizeof(uint64_t))) { w->dxcap = 0; return R2D_E_OOM; }
    }
    if (!w->excl.empty())
        HIPCHK(hipMemcpyAsync(w->dexcl, w->excl.data(), w->excl.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    w->topo_dirty = false;
    w->version += 1;   // device topology pointers may have moved
    return R2D_OK;
}

uint32_t table_size(const r2d_world* w)
{
    size_t mult = w->hash_mode == 0 ? 2 : w->hash_mult;
    return (uint32_t)(mult * (size_t)w->n);
}

float hash_cell(const r2d_world* w) { return w->hash_mode == 0 ? 4.0f : w->hash_cell; }

WorldDev make_desc(r2d_world* w)
{
    WorldDev d;
    std::memset(&d, 0, sizeof d);
    float* b = reinterpret_cast<float*>(w->dblock);
    size_t c = w->dcap;
    d.px = b + F_PX * c; d.py = b + F_PY * c; d.ang = b + F_ANG * c;
    d.mx = b + F_MX * c; d.my = b + F_MY * c; d.L = b + F_L * c;
    d.fx = b + F_FX * c; d.fy = b + F_FY * c; d.tq = b + F_TQ * c;
    d.mass = b + F_MASS * c; d.inertia = b + F_INERTIA * c; d.mu = b + F_MU * c;
    d.rad = b + F_RAD * c; d.hw = b + F_HW * c; d.hh = b + F_HH * c;
    d.ax = b + F_AX * c; d.ay = b + F_AY * c; d.ahw = b + F_AHW * c; d.ahh = b + F_AHH * c;
    d.id = w->dblock + F_ID * c; d.flags = w->dblock + F_FLAGS * c;
    d.joints = w->djoints; d.gens = w->dgens; d.excl = w->dexcl;
    d.nbodies = w->n;
    d.njoints = (int32_t)w->joints.size();
    d.ngens = (int32_t)w->gens.size();
    d.nexcl = (int32_t)w->excl.size();
    d.cell = hash_cell(w);
    d.table = table_size(w);
    d.sched = w->sched;
    if (w->warm > 0.0f) {
        d.warm = w->warm;
        d.wtab = w->dwtab; d.wcnt = w->dwcnt; d.wsum = w->dwsum;
    }
    return d;
}

// carve the step scratch for the current capacities
int ensure_scratch(r2d_ctx* c, uint32_t n, uint32_t b, uint32_t wn, uint32_t j, uint32_t e, uint32_t r, uint32_t m,
                   uint32_t it)
{
    if (n <= c->cap_n && b <= c->cap_b && wn <= c->cap_w && j <= c->cap_j && e <= c->cap_e && r <= c->cap_r &&
        m <= c->cap_m && it <= c->cap_i && c->arena)
        return R2D_OK;
    auto grow = [](uint32_t have, uint32_t need) {
        if (need <= have) return have;
        uint64_t v = (uint64_t)need + need / 2 + 256;
        return (uint32_t)std::min<uint64_t>(v, 0xffffff00ull);
    };
    c->cap_n = grow(c->cap_n, n); c->cap_b = grow(c->cap_b, b); c->cap_w = grow(c->cap_w, wn);
    c->cap_j = grow(c->cap_j, j); c->cap_e = grow(c->cap_e, e); c->cap_r = grow(c->cap_r, r);
    c->cap_m = grow(c->cap_m, m); c->cap_i = grow(c->cap_i, std::max(it, c->cap_m + c->cap_j));
    size_t N = c->cap_n, B = c->cap_b, W = c->cap_w, E = c->cap_e, R = c->cap_r, M = c->cap_m, I = c->cap_i;
    size_t off = 0;
    std::vector<std::pair<void**, size_t>> parts;
    StepDev& S = c->S;
    auto add = [&](void** p, size_t bytes) { parts.push_back({p, bytes}); };
    add((void**)&S.bworld, N * 4);
    add((void**)&S.cx0, N * 8); add((void**)&S.cy0, N * 8);
    add((void**)&S.cnx, N * 4); add((void**)&S.cny, N * 4);
    add((void**)&S.ncell, (N + 1) * 4); add((void**)&S.cell_off, (N + 1) * 4);
    add((void**)&S.cc, N * 4); add((void**)&S.cs, N * 4); add((void**)&S.sflag, N * 4);
    add((void**)&S.bkt_cnt, (B + 1) * 4); add((void**)&S.bkt_off, (B + 1) * 4);
    add((void**)&S.bkt_fill, (B + 1) * 4);
    add((void**)&S.bkt_items, E * 4); add((void**)&S.bkt_sorted, E * 4);
    add((void**)&S.ent_body, E * 4); add((void**)&S.ent_bkt, E * 4);
    add((void**)&S.ent_cnt, (E + 1) * 4); add((void**)&S.ent_off, (E + 1) * 4);
    add((void**)&S.cand, R * 16);
    add((void**)&S.cand_flag, (R + 1) * 4); add((void**)&S.cand_idx, (R + 1) * 4);
    add((void**)&S.sat_nx, R * 4); add((void**)&S.sat_ny, R * 4); add((void**)&S.sat_info, R * 4);
    add((void**)&S.m_ref, M * 4); add((void**)&S.m_inc, M * 4);
    add((void**)&S.m_nx, M * 4); add((void**)&S.m_ny, M * 4); add((void**)&S.m_fric, M * 4);
    add((void**)&S.m_prev1, M * 4); add((void**)&S.m_prev2, M * 4); add((void**)&S.m_valid, M * 4);
    add((void**)&S.m_pt, (size_t)MP_NFIELDS * 2 * M * 4);
    add((void**)&S.w_man_begin, (W + 1) * 4); add((void**)&S.w_item_base, (W + 1) * 4);
    add((void**)&S.w_status, (W + 1) * 4); add((void**)&S.w_order, (W + 1) * 4);
    add((void**)&S.it_r0, I * 4); add((void**)&S.it_r1, I * 4);
    add((void**)&S.it_k0, I * 4); add((void**)&S.it_k1, I * 4);
    add((void**)&S.deg, (N + 1) * 4); add((void**)&S.deg_off, (N + 1) * 4); add((void**)&S.deg_fill, (N + 1) * 4);
    add((void**)&S.lists, 2 * I * 4);
    add((void**)&S.it_color, I * 4);
    add((void**)&S.it_perm, I * 4);
    add((void**)&S.recs, I * 128);
    add((void**)&S.rmass, I * 16);
    add((void**)&S.cmask, 2 * N * 8);
    add((void**)&S.g_im, N * 4); add((void**)&S.g_ii, N * 4); add((void**)&S.g_cnt, N * 4);
    add((void**)&S.cmap, N * 4); add((void**)&S.cinv, N * 4); add((void**)&S.w_nc, (W + 1) * 4);
    add((void**)&S.counters, 64 * 4);
    add((void**)&S.partials, 1024 * 4);
    for (auto& p : parts) off += align256(p.second);
    if (c->arena) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->arena));
        c->arena = nullptr;
    }
    if (hipMalloc(&c->arena, off) != hipSuccess) { c->arena = nullptr; c->arena_bytes = 0; return R2D_E_OOM; }
    c->arena_bytes = off;
    char* base = (char*)c->arena;
    size_t o = 0;
    for (auto& p : parts) { *p.first = base + o; o += align256(p.second); }
    HIPCHK(hipMemsetAsync(c->arena, 0, off, c->stream));
    S.cap_e = c->cap_e; S.cap_r = c->cap_r; S.cap_m = c->cap_m; S.cap_i = c->cap_i;
    if (c->hstatus_cap < W + 1) {
        if (c->hstatus) HIPCHK(hipHostFree(c->hstatus));
        if (hipHostMalloc((void**)&c->hstatus, (W + 1) * 4) != hipSuccess) return R2D_E_OOM;
        c->hstatus_cap = W + 1;
    }
    c->uploaded_descs.clear();
    destroy_graphs(c);
    if (c->dworlds_cap < W) {
        if (c->dworlds) HIPCHK(hipFree(c->dworlds));
        if (hipMalloc(&c->dworlds, W * sizeof(WorldDev)) != hipSuccess) return R2D_E_OOM;
        c->dworlds_cap = W;
    }
    return R2D_OK;
}

void destroy_graphs(r2d_ctx* c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    for (int ph = 0; ph < PH_N; ++ph) {
        if (c->phase_exec[ph]) (void)hipGraphExecDestroy(c->phase_exec[ph]);
        c->phase_exec[ph] = nullptr;
    }
    c->graph_key.clear();
}

// The launch shape of a batch the caller has not tuned: a few worlds (each
// alone on a CU, latency-bound on their chains) take the 512-lane record
// kernel; more take the register-slot kernel for small worlds and the
// 256-thread record kernel for the others.
int default_variant(size_t nworlds) { return nworlds <= WIDE_MAX_WORLDS ? WIDE_VARIANT : 1; }

template <class Fn>
bool capture_graph(r2d_ctx* c, hipGraphExec_t* out, Fn&& enqueue)
{
    hipGraph_t g = nullptr;
    *out = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    enqueue();
    bool ok = hipStreamEndCapture(c->stream, &g) == hipSuccess && g;
    if (ok) ok = hipGraphInstantiate(out, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    if (!ok) *out = nullptr;
    (void)hipGetLastError();
    return ok;
}

// One batched call in flight: prepared (uploads, scratch, plan), launched
// (the steps enqueued plus asynchronous copies of the counters), then waited
// for (overflowed scratch is grown and the uncommitted steps relaunched).
struct Batch {
    std::vector<r2d_world*> worlds;
    std::vector<WorldDev> descs;
    uint64_t N = 0, B = 0, J = 0;
    SolvePlan plan;
    float dt = 0.f;
    size_t sub = 0, iters = 0;
    size_t steps = 0, done = 0;
    int attempts = 0;
    int first_err = R2D_OK;
    std::vector<int32_t> world_status;   // device status per world of the last wait
    bool fused = false;                  // whole steps per workgroup (k_fused)
    size_t fused_lds = 0;
    int variant = 0;
};

// ------------------------------------------------ fused steps: scratch slices
constexpr size_t kFusedLdsMax = 160 * 1024 - 1024;

uint32_t grow_cap(uint32_t have, uint64_t need)
{
    if (need <= have) return have;
    uint64_t v = need + need / 2 + 256;
    return (uint32_t)std::min<uint64_t>(v, 0xffffff00ull);
}

// The fused arena: every world's slices of the entry, candidate, manifold and
// item arrays.  With `preserve` the manifold arrays keep their contents (the
// slices of worlds that finished before an overflow relaunch stay readable).
int ensure_fused_arena(r2d_ctx* c, uint64_t te, uint64_t tr, uint64_t tm, uint64_t ti, bool preserve)
{
    if (te <= c->fcap_e && tr <= c->fcap_r && tm <= c->fcap_m && ti <= c->fcap_i && c->farena) return R2D_OK;
    if (std::max(std::max(te, tr), std::max(tm, ti)) >= 0x7fffff00ull) return R2D_E_CAPACITY;
    const uint32_t E = grow_cap(c->fcap_e, te), R = grow_cap(c->fcap_r, tr), M = grow_cap(c->fcap_m, tm),
                   I = grow_cap(c->fcap_i, ti);
    StepDev F{};
    std::vector<std::pair<void**, size_t>> parts;
    auto add = [&](void** p, size_t bytes) { parts.push_back({p, bytes}); };
    add((void**)&F.bkt_items, (size_t)E * 4); add((void**)&F.bkt_sorted, (size_t)E * 4);
    add((void**)&F.ent_body, (size_t)E * 4); add((void**)&F.ent_bkt, (size_t)E * 4);
    add((void**)&F.ent_cnt, (size_t)E * 4);
    add((void**)&F.cand, (size_t)R * 16); add((void**)&F.cand_flag, (size_t)R * 4);
    add((void**)&F.sat_nx, (size_t)R * 4); add((void**)&F.sat_ny, (size_t)R * 4); add((void**)&F.sat_info, (size_t)R * 4);
    add((void**)&F.m_ref, (size_t)M * 4); add((void**)&F.m_inc, (size_t)M * 4);
    add((void**)&F.m_nx, (size_t)M * 4); add((void**)&F.m_ny, (size_t)M * 4); add((void**)&F.m_fric, (size_t)M * 4);
    add((void**)&F.m_prev1, (size_t)M * 4); add((void**)&F.m_prev2, (size_t)M * 4); add((void**)&F.m_valid, (size_t)M * 4);
    add((void**)&F.m_pt, (size_t)MP_NFIELDS * 2 * M * 4);
    add((void**)&F.it_r0, (size_t)I * 4); add((void**)&F.it_r1, (size_t)I * 4);
    add((void**)&F.it_k0, (size_t)I * 4); add((void**)&F.it_k1, (size_t)I * 4);
    add((void**)&F.it_perm, (size_t)I * 4);
    add((void**)&F.lists, 2 * (size_t)I * 4);
    add((void**)&F.recs, (size_t)I * 128);
    size_t bytes = 0;
    for (auto& p : parts) bytes += align256(p.second);
    void* arena = nullptr;
    if (hipMalloc(&arena, bytes) != hipSuccess) return R2D_E_OOM;
    size_t o = 0;
    for (auto& p : parts) { *p.first = (char*)arena + o; o += align256(p.second); }
    StepDev& O = c->SF;
    if (c->farena && preserve) {
        const size_t om = c->fcap_m;
        void* const dst[8] = {F.m_ref, F.m_inc, F.m_nx, F.m_ny, F.m_fric, F.m_prev1, F.m_prev2, F.m_valid};
        const void* const src[8] = {O.m_ref, O.m_inc, O.m_nx, O.m_ny, O.m_fric, O.m_prev1, O.m_prev2, O.m_valid};
        for (int k = 0; k < 8; ++k)
            HIPCHK(hipMemcpyAsync(dst[k], src[k], om * 4, hipMemcpyDeviceToDevice, c->stream));
        for (int r = 0; r < MP_NFIELDS * 2; ++r)
            HIPCHK(hipMemcpyAsync(F.m_pt + (size_t)r * M, O.m_pt + (size_t)r * om, om * 4, hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    if (c->farena) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->farena));
    }
    c->farena = arena;
    c->farena_bytes = bytes;
    c->fcap_e = E; c->fcap_r = R; c->fcap_m = M; c->fcap_i = I;
    O.bkt_items = F.bkt_items; O.bkt_sorted = F.bkt_sorted; O.ent_body = F.ent_body; O.ent_bkt = F.ent_bkt;
    O.ent_cnt = F.ent_cnt; O.cand = F.cand; O.cand_flag = F.cand_flag; O.sat_nx = F.sat_nx; O.sat_ny = F.sat_ny;
    O.sat_info = F.sat_info; O.m_ref = F.m_ref; O.m_inc = F.m_inc; O.m_nx = F.m_nx; O.m_ny = F.m_ny;
    O.m_fric = F.m_fric; O.m_prev1 = F.m_prev1; O.m_prev2 = F.m_prev2; O.m_valid = F.m_valid; O.m_pt = F.m_pt;
    O.it_r0 = F.it_r0; O.it_r1 = F.it_r1; O.it_k0 = F.it_k0; O.it_k1 = F.it_k1; O.it_perm = F.it_perm;
    O.lists = F.lists; O.recs = F.recs;
    O.cap_e = E; O.cap_r = R; O.cap_m = M; O.cap_i = I;
    return R2D_OK;
}

// per-world arrays of the fused kernel (device) and their pinned host copies
int ensure_fused_world(r2d_ctx* c, size_t W)
{
    if (W <= c->fworld_cap && c->fworld) return R2D_OK;
    const size_t cap = std::max<size_t>(W, 2 * c->fworld_cap);
    const size_t bytes = align256(cap * sizeof(WorldScratch)) + 4 * align256(cap * 4) + align256(cap * 32) + 256;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return R2D_E_OOM;
    if (c->fworld) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->fworld));
    }
    c->fworld = p;
    c->fworld_cap = cap;
    char* q = (char*)p;
    StepDev& O = c->SF;
    O.wscr = (WorldScratch*)q; q += align256(cap * sizeof(WorldScratch));
    O.w_done = (uint32_t*)q; q += align256(cap * 4);
    O.w_man_begin = (uint32_t*)q; q += align256(cap * 4);
    O.w_man_end = (uint32_t*)q; q += align256(cap * 4);
    O.w_order = (uint32_t*)q; q += align256(cap * 4);
    O.w_cnt = (uint32_t*)q; q += align256(cap * 32);
    O.fpool = (uint32_t*)q;
    c->fscr_uploaded.clear();
    c->forder_uploaded.clear();
    if (c->hfused_cap < cap) {
        if (c->hdone) HIPCHK(hipHostFree(c->hdone));
        if (c->hcnt) HIPCHK(hipHostFree(c->hcnt));
        if (c->hscr) HIPCHK(hipHostFree(c->hscr));
        c->hdone = nullptr;
        c->hcnt = nullptr;
        c->hscr = nullptr;
        c->hfused_cap = 0;
        if (c->dfsum) HIPCHK(hipFree(c->dfsum));
        if (c->hfsum) HIPCHK(hipHostFree(c->hfsum));
        c->dfsum = nullptr;
        c->hfsum = nullptr;
        if (hipHostMalloc((void**)&c->hdone, cap * 4) != hipSuccess) return R2D_E_OOM;
        if (hipHostMalloc((void**)&c->hcnt, cap * 32) != hipSuccess) return R2D_E_OOM;
        if (hipHostMalloc((void**)&c->hscr, cap * sizeof(WorldScratch)) != hipSuccess) return R2D_E_OOM;
        if (hipMalloc((void**)&c->dfsum, fsum_words(cap) * 4) != hipSuccess) return R2D_E_OOM;
        if (hipHostMalloc((void**)&c->hfsum, fsum_words(cap) * 4) != hipSuccess) return R2D_E_OOM;
        c->hfused_cap = cap;
    }
    return R2D_OK;
}

// slice capacities of a world: what it used before, else estimates (per
// body 4 cell entries, 32 raw candidates -- a settled pile has about 24 --
// and 3 manifolds)
void fused_caps(const r2d_world* w, uint32_t cap[4])
{
    const uint32_t n = (uint32_t)w->n, nj = (uint32_t)w->joints.size();
    cap[0] = w->fz[0] ? w->fz[0] : 4 * n + 64;
    cap[1] = w->fz[1] ? w->fz[1] : 32 * n + 256;
    cap[2] = w->fz[2] ? w->fz[2] : 3 * n + 64;
    cap[3] = std::max(w->fz[3] ? w->fz[3] : nj + cap[2], nj + 1);
}

// end of the used part of each array (entries, candidates, manifolds, items)
void fused_used(const r2d_ctx* c, uint64_t end[4])
{
    end[0] = end[1] = end[2] = end[3] = 0;
    for (const WorldScratch& x : c->fscr) {
        end[0] = std::max<uint64_t>(end[0], (uint64_t)x.e0 + x.ecap + 1);
        end[1] = std::max<uint64_t>(end[1], (uint64_t)x.r0 + x.rcap + 1);
        end[2] = std::max<uint64_t>(end[2], (uint64_t)x.m0 + x.mcap);
        end[3] = std::max<uint64_t>(end[3], (uint64_t)x.i0 + x.icap);
    }
}

// upload the slice table and the workgroup order when they changed
int fused_upload_tables(r2d_ctx* c)
{
    const size_t nw = c->fscr.size();
    if (c->fscr_uploaded.size() != nw ||
        std::memcmp(c->fscr_uploaded.data(), c->fscr.data(), nw * sizeof(WorldScratch)) != 0) {
        HIPCHK(hipMemcpyAsync((void*)c->SF.wscr, c->fscr.data(), nw * sizeof(WorldScratch), hipMemcpyHostToDevice,
                              c->stream));
        c->fscr_uploaded = c->fscr;
    }
    if (c->forder_uploaded != c->forder) {
        HIPCHK(hipMemcpyAsync(c->SF.w_order, c->forder.data(), nw * 4, hipMemcpyHostToDevice, c->stream));
        c->forder_uploaded = c->forder;
    }
    // the host copies are read by the asynchronous uploads above
    HIPCHK(hipStreamSynchronize(c->stream));
    return R2D_OK;
}

// slices of every world of a new batch, one after another
int fused_layout(r2d_ctx* c, Batch& b)
{
    const size_t nw = b.worlds.size();
    int rc = ensure_fused_world(c, nw);
    if (rc) return rc;
    c->fscr.resize(nw);
    uint64_t te = 0, tr = 0, tm = 0, ti = 0;
    for (size_t i = 0; i < nw; ++i) {
        uint32_t cap[4];
        fused_caps(b.worlds[i], cap);
        c->fscr[i] = WorldScratch{(uint32_t)te, cap[0], (uint32_t)tr, cap[1], (uint32_t)tm, cap[2], (uint32_t)ti,
                                  cap[3]};
        te += (uint64_t)cap[0] + 1; tr += (uint64_t)cap[1] + 1; tm += cap[2]; ti += cap[3];
    }
    // as much again free for worlds that outgrow their slices during the call
    if ((rc = ensure_fused_arena(c, 2 * te, 2 * tr, 2 * tm, 2 * ti, false))) return rc;
    // workgroups take the largest worlds first (the longest steps start first)
    c->forder.resize(nw);
    for (size_t i = 0; i < nw; ++i) c->forder[i] = (uint32_t)i;
    std::stable_sort(c->forder.begin(), c->forder.end(),
                     [&](uint32_t x, uint32_t y) { return b.worlds[x]->n > b.worlds[y]->n; });
    // Workgroups b, b + P, b + 2P, ... (P = compute units) land on one CU when
    // the dispatcher hands them out round-robin and they are resident at once,
    // so with plain largest-first order one CU holds the largest world of every
    // run of P.  Every other run is reversed: each CU mixes large and small
    // worlds, and the call, which lasts as long as its slowest CU, ends sooner.
    const size_t P = c->num_cus > 0 ? (size_t)c->num_cus : 0;
    for (size_t q = 1; P && (q + 1) * P <= nw; q += 2)
        std::reverse(c->forder.begin() + q * P, c->forder.begin() + (q + 1) * P);
    return fused_upload_tables(c);
}

// after an overflow: the worlds that stopped get larger slices appended at
// the end (the other slices, and the manifolds they hold, stay where they are)
int fused_regrow(r2d_ctx* c, Batch& b)
{
    const size_t nw = b.worlds.size();
    uint64_t end[4];
    fused_used(c, end);
    uint64_t te = end[0], tr = end[1], tm = end[2], ti = end[3];
    for (size_t i = 0; i < nw; ++i) {
        if (c->hstatus[i] != WS_OVERFLOW) continue;
        r2d_world* w = b.worlds[i];
        const uint32_t* need = c->hcnt + 8 * i + 4;
        uint32_t cap[4];
        fused_caps(w, cap);
        for (int k = 0; k < 4; ++k)
            if (need[k] > cap[k]) cap[k] = (uint32_t)std::min<uint64_t>((uint64_t)need[k] * 2 + 64, 0x7ffffff0u);
        cap[3] = std::max(cap[3], (uint32_t)w->joints.size() + cap[2]);
        for (int k = 0; k < 4; ++k) w->fz[k] = cap[k];
        c->fscr[i] = WorldScratch{(uint32_t)te, cap[0], (uint32_t)tr, cap[1], (uint32_t)tm, cap[2], (uint32_t)ti,
                                  cap[3]};
        te += (uint64_t)cap[0] + 1; tr += (uint64_t)cap[1] + 1; tm += cap[2]; ti += cap[3];
        c->hstatus[i] = WS_OK;
    }
    int rc = ensure_fused_arena(c, te + te / 2, tr + tr / 2, tm + tm / 2, ti + ti / 2, true);
    if (rc) return rc;
    if ((rc = fused_upload_tables(c))) return rc;
    HIPCHK(hipMemcpyAsync(c->S.w_status, c->hstatus, nw * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return R2D_OK;
}

// the fused kernel's view: the step scratch's per-body / per-world arrays
// and the fused arena's slices
// A fused launch's bookkeeping in two small kernels instead of four fills, a
// copy in and five copies out (each an enqueue of its own, ~10 us of a
// one-world call): k_fused_prep clears the counters (and, on a call's first
// launch, the worlds' steps done, status and counts) and writes the arrays'
// free ends; k_fused_gather packs what the host reads back into one block,
// copied in one piece: the 64 counters, then per world status, steps done,
// the 8 counts and the 8-word slice table entry.

struct PoolArgs {
    uint32_t v[8];
};

__global__ void k_fused_prep(uint32_t* counters, int32_t* w_status, uint32_t* w_done, uint32_t* w_cnt, uint32_t* fpool,
                             PoolArgs pool, uint32_t nw, uint32_t first)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64u) counters[i] = 0u;
    if (i < 8u) fpool[i] = pool.v[i];
    if (first) {
        if (i < nw) { w_status[i] = 0; w_done[i] = 0u; }
        if (i < 8u * nw) w_cnt[i] = 0u;
    }
}

__global__ void k_fused_gather(const uint32_t* counters, const int32_t* w_status, const uint32_t* w_done,
                               const uint32_t* w_cnt, const WorldScratch* wscr, uint32_t* out, uint32_t nw)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64u) out[i] = counters[i];
    uint32_t* o = out + 64;
    if (i < nw) { o[i] = (uint32_t)w_status[i]; o[nw + i] = w_done[i]; }
    if (i < 8u * nw) {
        o[2u * nw + i] = w_cnt[i];
        o[10u * nw + i] = reinterpret_cast<const uint32_t*>(wscr)[i];
    }
}

unsigned fsum_blocks(size_t nw) { return (unsigned)((std::max<size_t>(64, 8 * nw) + 255) / 256); }

// the gathered summary into the host arrays the rest of the code reads
void fsum_unpack(r2d_ctx* c, size_t nw)
{
    const uint32_t* h = c->hfsum;
    std::memcpy(c->hcounters, h, 64 * 4);
    h += 64;
    std::memcpy(c->hstatus, h, nw * 4);
    std::memcpy(c->hdone, h + nw, nw * 4);
    std::memcpy(c->hcnt, h + 2 * nw, nw * 32);
    std::memcpy(c->hscr, h + 10 * nw, nw * sizeof(WorldScratch));
}

StepDev fused_view(const r2d_ctx* c)
{
    StepDev F = c->S;
    const StepDev& O = c->SF;
    F.bkt_items = O.bkt_items; F.bkt_sorted = O.bkt_sorted; F.ent_body = O.ent_body; F.ent_bkt = O.ent_bkt;
    F.ent_cnt = O.ent_cnt; F.cand = O.cand; F.cand_flag = O.cand_flag; F.sat_nx = O.sat_nx; F.sat_ny = O.sat_ny;
    F.sat_info = O.sat_info; F.m_ref = O.m_ref; F.m_inc = O.m_inc; F.m_nx = O.m_nx; F.m_ny = O.m_ny;
    F.m_fric = O.m_fric; F.m_prev1 = O.m_prev1; F.m_prev2 = O.m_prev2; F.m_valid = O.m_valid; F.m_pt = O.m_pt;
    F.it_r0 = O.it_r0; F.it_r1 = O.it_r1; F.it_k0 = O.it_k0; F.it_k1 = O.it_k1; F.it_perm = O.it_perm;
    F.lists = O.lists; F.recs = O.recs;
    F.cap_e = O.cap_e; F.cap_r = O.cap_r; F.cap_m = O.cap_m; F.cap_i = O.cap_i;
    F.wscr = O.wscr; F.fpool = O.fpool; F.w_done = O.w_done; F.w_cnt = O.w_cnt; F.w_man_begin = O.w_man_begin;
    F.w_man_end = O.w_man_end; F.w_order = O.w_order;
    F.w_tstamp = nullptr; F.tcap = 0; F.w_prof = nullptr;
    return F;
}


int batch_prepare(r2d_ctx* c, r2d_world* const* worlds, size_t nw, float dt, size_t sub_steps, size_t iters,
                  size_t steps, Batch& b)
{
    // the reference raises InvalidRigidBodyId from the first constraint solve
    bool solves = sub_steps > 0 && iters > 0;
    for (size_t i = 0; i < nw; ++i) {
        r2d_world* w = worlds[i];
        if (!w || w->ctx != c) return R2D_E_INVALID_ARG;
        if (solves && w->checked_version != w->version) {
            for (const HJoint& h : w->joints)
                if (w->slot_of(h.id1) < 0 || (h.type <= 1 && w->slot_of(h.id2) < 0)) return R2D_E_INVALID_ID;
            w->checked_version = w->version;
        }
    }
    b = Batch();
    b.worlds.assign(worlds, worlds + nw);
    b.descs.resize(nw);
    b.dt = dt; b.sub = sub_steps; b.iters = iters; b.steps = steps;
    std::vector<uint32_t> sizes(nw);
    for (size_t i = 0; i < nw; ++i) {
        r2d_world* w = worlds[i];
        int rc;
        if (!w->dev_valid && (rc = upload_bodies(w))) return rc;
        if ((rc = upload_topology(w))) return rc;
        if ((rc = ensure_warm(w))) return rc;
        if (w->desc_version != w->version) {
            w->desc = make_desc(w);
            w->desc_version = w->version;
        }
        WorldDev d = w->desc;
        d.body_base = (uint32_t)b.N;
        d.bucket_base = (uint32_t)b.B;
        d.joint_base = (uint32_t)b.J;
        b.descs[i] = d;
        b.N += (uint64_t)w->n;
        b.B += d.table;
        b.J += w->joints.size();
        sizes[i] = (uint32_t)w->n;
    }
    if (b.N >= 0x7fffffffull || b.B >= 0x7fffffffull) return R2D_E_CAPACITY;

    // solve variant: context override (autotuner), environment, then the
    // default (default_variant)
    int variant = default_variant(nw);
    if (c->solve_variant >= 0) variant = c->solve_variant;
    if (const char* e = std::getenv("R2D_SOLVE_VARIANT")) {
        int v = std::atoi(e);
        if (v >= 0 && v < SOLVE_VARIANTS) variant = v;
    }
    bool colored = false, warm = false;
    for (size_t i = 0; i < nw; ++i) {
        colored = colored || worlds[i]->sched != 0;
        warm = warm || worlds[i]->warm > 0.0f;
    }
    // warm starting is compiled into one variant's kernels only
    if (warm) variant = WARM_VARIANT;
    b.plan = plan_solve(variant, sizes.data(), nw, c->lds_nmax, colored, warm, c->wide_record);
    b.variant = variant;
    // whole steps per workgroup: worlds of at most FUSED_NMAX bodies in the
    // reference order whose broadphase and solve arrays fit LDS
    bool fused = c->fused_enabled && fused_available(variant) && !colored;
    uint32_t maxn = 0, maxtb = 0;
    for (size_t i = 0; i < nw && fused; ++i) {
        fused = sizes[i] <= FUSED_NMAX;
        maxn = std::max(maxn, sizes[i]);
        maxtb = std::max(maxtb, b.descs[i].table);
    }
    if (fused) {
        b.fused_lds = fused_lds_bytes(maxn, maxtb, fused_slots(variant), fused_threads(variant));
        fused = b.fused_lds <= kFusedLdsMax;
    }
    b.fused = fused;
    int rc;
    if (fused) {
        // the step scratch holds the per-body and per-world arrays; the
        // worlds' slices live in the fused arena
        if ((rc = ensure_scratch(c, (uint32_t)b.N, (uint32_t)b.B, (uint32_t)nw, (uint32_t)b.J, 1, 1, 1, 1)))
            return rc;
        return fused_layout(c, b);
    }
    uint32_t est_e = std::max<uint32_t>(c->cap_e, (uint32_t)(4 * b.N + 1024));
    uint32_t est_r = std::max<uint32_t>(c->cap_r, (uint32_t)(8 * b.N + 1024));
    uint32_t est_m = std::max<uint32_t>(c->cap_m, (uint32_t)(4 * b.N + 1024));
    return ensure_scratch(c, (uint32_t)b.N, (uint32_t)b.B, (uint32_t)nw, (uint32_t)b.J, est_e, est_r, est_m,
                          (uint32_t)b.J + est_m);
}

// sum the phase times of the timed steps recorded so far
int timer_flush(r2d_ctx* c)
{
    if (!c->ev_used) return R2D_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k + KT_N <= c->ev_used; k += KT_N) {
        float total = 0.f;
        for (int p = 0; p < 4; ++p) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[k + p], c->ev_pool[k + p + 1]));
            c->phase_ms[p] += ms;
            total += ms;
        }
        c->step_ms.push_back(total);
        c->phase_steps += 1;
    }
    c->ev_used = 0;
    return R2D_OK;
}

int fused_launch(r2d_ctx* c, Batch& b);

int batch_launch(r2d_ctx* c, Batch& b)
{
    const size_t nw = b.worlds.size();
    if (c->uploaded_descs.size() != nw ||
        std::memcmp(c->uploaded_descs.data(), b.descs.data(), nw * sizeof(WorldDev)) != 0) {
        HIPCHK(hipMemcpyAsync(c->dworlds, b.descs.data(), nw * sizeof(WorldDev), hipMemcpyHostToDevice, c->stream));
        c->uploaded_descs = b.descs;
    }
    if (b.fused) return fused_launch(c, b);
    StepDev S = c->S;
    S.nb = (uint32_t)b.N; S.nworlds = (uint32_t)nw; S.nbuckets = (uint32_t)b.B; S.njoints = (uint32_t)b.J;
    S.worlds = c->dworlds;
    const Streams ss = c->streams();
    const uint32_t sub = (uint32_t)b.sub, iters = (uint32_t)b.iters;
    auto eager_phase = [&](int ph) { launch_phase(S, ss, ph, b.dt, sub, iters, b.plan); };
    const bool tm = c->profiling;
    hipGraphExec_t* want = nullptr;   // graphs this launch replays
    if (c->use_graph) {
        // key: everything the captured launches bake in (the world descriptors
        // are device data, uploaded above, and are not part of it)
        std::vector<uint8_t> key;
        auto put = [&](const void* p, size_t n) { key.insert(key.end(), (const uint8_t*)p, (const uint8_t*)p + n); };
        put(&S, sizeof S);
        put(&b.plan, sizeof b.plan);
        float kdt = b.dt; uint32_t ks = sub, ki = iters;
        put(&kdt, 4); put(&ks, 4); put(&ki, 4);
        if (key != c->graph_key) { destroy_graphs(c); c->graph_key.swap(key); c->graph_uses = 0; }
        c->graph_uses += 1;
        // capturing costs more than launching once: capture a shape that repeats
        // (a second launch with the same key, or several steps in this one)
        const bool repeat = c->graph_uses >= 2 || b.steps - b.done >= 2;
        bool ok = repeat;
        if (ok && !tm && !c->graph_exec)
            ok = capture_graph(c, &c->graph_exec, [&]() { for (int ph = 0; ph < PH_N; ++ph) eager_phase(ph); });
        for (int ph = 0; ok && tm && ph < PH_N; ++ph)
            if (!c->phase_exec[ph]) ok = capture_graph(c, &c->phase_exec[ph], [&]() { eager_phase(ph); });
        if (repeat && !ok) { destroy_graphs(c); c->use_graph = false; }   // capture unsupported: launch eagerly
        else if (ok) want = tm ? c->phase_exec : &c->graph_exec;
    }
    HIPCHK(hipMemsetAsync(S.counters, 0, 64 * 4, c->stream));
    HIPCHK(hipEventRecord(c->ev_begin, c->stream));
    const size_t todo = b.steps - b.done;
    for (size_t s = 0; s < todo; ++s) {
        if (!tm) {
            if (want) HIPCHK(hipGraphLaunch(*want, c->stream));
            else for (int ph = 0; ph < PH_N; ++ph) eager_phase(ph);
            continue;
        }
        // timed step: events on the stream between the phases, summed later
        // (timer_flush), so timing adds no synchronisation
        if (c->ev_used + KT_N > c->ev_pool.size()) {
            if (c->ev_pool.size() >= 4000) {
                int rc = timer_flush(c);
                if (rc) return rc;
            } else {
                for (int i = 0; i < 64 * KT_N; ++i) {
                    hipEvent_t e;
                    HIPCHK(hipEventCreate(&e));
                    c->ev_pool.push_back(e);
                }
            }
        }
        hipEvent_t* ev = c->ev_pool.data() + c->ev_used;
        for (int ph = 0; ph < PH_N; ++ph) {
            HIPCHK(hipEventRecord(ev[ph], c->stream));
            if (want) HIPCHK(hipGraphLaunch(want[ph], c->stream));
            else eager_phase(ph);
        }
        HIPCHK(hipEventRecord(ev[KT_END], c->stream));
        c->ev_used += KT_N;
    }
    HIPCHK(hipMemcpyAsync(c->hcounters, S.counters, 64 * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->hstatus, S.w_status, nw * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->ev_end, c->stream));
    b.attempts += 1;
    return R2D_OK;
}

// the steps of a fused launch: one kernel for the whole call
int fused_launch(r2d_ctx* c, Batch& b)
{
    const size_t nw = b.worlds.size();
    StepDev F = fused_view(c);
    F.nb = (uint32_t)b.N; F.nworlds = (uint32_t)nw; F.nbuckets = (uint32_t)b.B; F.njoints = (uint32_t)b.J;
    F.worlds = c->dworlds;
    if (c->profiling) {
        // The timing buffers grow rarely and by a lot: a launch that follows a
        // hipMalloc closely can stall all of its waves for several ms at once
        // (measured: one ~7 ms world-step in 889 of 1,024 worlds of the first
        // profiled 20-step call, none once the buffers existed before it).
        const size_t words = nw * (b.steps + 1);
        if (c->tstamp_words < words) {
            const size_t want = std::max<size_t>({words, 2 * c->tstamp_words, (size_t)1 << 18});
            if (c->tstamp) HIPCHK(hipFree(c->tstamp));
            c->tstamp = nullptr;
            c->tstamp_words = 0;
            if (hipMalloc((void**)&c->tstamp, want * 8) != hipSuccess) return R2D_E_OOM;
            c->tstamp_words = want;
        }
        if (b.attempts == 0) HIPCHK(hipMemsetAsync(c->tstamp, 0, words * 8, c->stream));
        F.w_tstamp = c->tstamp;
        F.tcap = (uint32_t)b.steps;
        if (c->wprof_worlds < nw) {
            if (c->wprof) HIPCHK(hipFree(c->wprof));
            c->wprof = nullptr;
            const size_t want = std::max<size_t>({nw, 2 * c->wprof_worlds, 4096});
            c->wprof_worlds = 0;
            if (hipMalloc((void**)&c->wprof, want * FP_N * 8) != hipSuccess) return R2D_E_OOM;
            c->wprof_worlds = want;
        }
        if (b.attempts == 0) HIPCHK(hipMemsetAsync(c->wprof, 0, nw * FP_N * 8, c->stream));
        F.w_prof = c->wprof;
    }
    // free ends of the arrays: where outgrown worlds take new slices
    uint64_t end[4];
    fused_used(c, end);
    const uint32_t caps[4] = {c->fcap_e, c->fcap_r, c->fcap_m, c->fcap_i};
    PoolArgs pool;
    for (int k = 0; k < 4; ++k) {
        pool.v[k] = (uint32_t)std::min<uint64_t>(end[k], caps[k]);
        pool.v[4 + k] = caps[k];
    }
    hipLaunchKernelGGL(k_fused_prep, dim3(fsum_blocks(nw)), dim3(256), 0, c->stream, F.counters, F.w_status,
                       F.w_done, F.w_cnt, F.fpool, pool, (uint32_t)nw, b.attempts == 0 ? 1u : 0u);
    HIPCHK(hipEventRecord(c->ev_begin, c->stream));
    (void)hipGetLastError();   // the launch check below sees this launch only
    launch_fused(F, b.variant, (uint32_t)nw, b.dt, (uint32_t)b.sub, (uint32_t)b.iters, (uint32_t)b.steps,
                 b.fused_lds, c->stream);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        hipFuncAttributes fa{};
        (void)fused_attributes(b.variant, &fa);
        std::snprintf(g_last_error, sizeof g_last_error,
                      "fused kernel launch failed: %s (%d); variant %d, %zu worlds, %zu B dynamic LDS; kernel: %d "
                      "regs, %zu B static LDS, %zu B max dynamic LDS, %zu B scratch per lane, %d max threads",
                      hipGetErrorString(le), (int)le, b.variant, nw, b.fused_lds, fa.numRegs, fa.sharedSizeBytes,
                      (size_t)fa.maxDynamicSharedSizeBytes, fa.localSizeBytes, fa.maxThreadsPerBlock);
        return R2D_E_DEVICE;
    }
    HIPCHK(hipEventRecord(c->ev_kend, c->stream));
    hipLaunchKernelGGL(k_fused_gather, dim3(fsum_blocks(nw)), dim3(256), 0, c->stream, F.counters, F.w_status,
                       F.w_done, F.w_cnt, F.wscr, c->dfsum, (uint32_t)nw);
    HIPCHK(hipMemcpyAsync(c->hfsum, c->dfsum, fsum_words(nw) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->ev_end, c->stream));
    b.attempts += 1;
    return R2D_OK;
}

// per-world step latencies of a profiled fused launch (wall clock at 100 MHz)
int fused_collect_times(r2d_ctx* c, const Batch& b)
{
    const size_t nw = b.worlds.size(), stride = b.steps + 1;
    std::vector<uint64_t> ts(nw * stride);
    HIPCHK(hipMemcpy(ts.data(), c->tstamp, ts.size() * 8, hipMemcpyDeviceToHost));
    int hz = 0;
    if (hipDeviceGetAttribute(&hz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || hz <= 0) hz = 100000;
    const double ms_per_tick = 1.0 / (double)hz;   // rate in kHz
    for (size_t i = 0; i < nw; ++i)
        for (size_t s = 0; s < b.steps; ++s) {
            const uint64_t a = ts[i * stride + s], e = ts[i * stride + s + 1];
            if (a && e && e >= a) c->step_ms.push_back((float)((double)(e - a) * ms_per_tick));
        }
    std::vector<uint64_t> ph(nw * FP_N);
    HIPCHK(hipMemcpy(ph.data(), c->wprof, ph.size() * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nw; ++i)
        for (int k = 0; k < FP_N; ++k) c->fused_phase_ms[k] += (double)ph[i * FP_N + k] * ms_per_tick;
    return R2D_OK;
}

int fused_wait(r2d_ctx* c, Batch& b)
{
    const size_t nw = b.worlds.size();
    for (;;) {
        HIPCHK(hipEventSynchronize(c->ev_end));
        fsum_unpack(c, nw);
        if (c->profiling) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev_begin, c->ev_kend));
            c->fused_ms += ms;
            c->fused_launches += 1;
            const uint64_t* sums = reinterpret_cast<const uint64_t*>(c->hcounters + C_SUMS);
            for (int k = 0; k < 5; ++k) c->fused_sums[k] += sums[k];
        }
        b.world_status.assign(c->hstatus, c->hstatus + nw);
        // slices the worlds moved to during the launch; later layouts start from them
        c->fscr.assign(c->hscr, c->hscr + nw);
        c->fscr_uploaded = c->fscr;
        for (size_t i = 0; i < nw; ++i) {
            const WorldScratch& x = c->fscr[i];
            uint32_t* fz = b.worlds[i]->fz;
            fz[0] = std::max(fz[0], x.ecap); fz[1] = std::max(fz[1], x.rcap);
            fz[2] = std::max(fz[2], x.mcap); fz[3] = std::max(fz[3], x.icap);
        }
        bool overflow = false;
        for (size_t i = 0; i < nw; ++i) {
            const int32_t st = c->hstatus[i];
            if (st == WS_OVERFLOW) { overflow = true; continue; }
            if (st == WS_STALL) std::snprintf(g_last_error, sizeof g_last_error, "solver schedule stalled in world %zu", i);
            if (st != WS_OK && b.first_err == R2D_OK)
                b.first_err = st == WS_NONFINITE ? R2D_E_NONFINITE : st == WS_STALL ? R2D_E_DEVICE : R2D_E_RANGE;
        }
        for (int k = 0; k < 4; ++k) c->last_counts[k] = c->hcounters[k];
        if (!overflow) break;
        if (b.attempts > 16) return R2D_E_CAPACITY;
        int rc = fused_regrow(c, b);
        if (rc) return rc;
        if ((rc = fused_launch(c, b))) return rc;
    }
    if (c->profiling) {
        int rc = fused_collect_times(c, b);
        if (rc) return rc;
        for (size_t i = 0; i < nw; ++i) c->fused_steps += c->hdone[i];
    }
    b.done = b.steps;
    return R2D_OK;
}

// the worlds of a completed batch: their manifolds are this batch's
int batch_finish(r2d_ctx* c, Batch& b)
{
    c->epoch = g_epoch.fetch_add(1) + 1;   // unique over contexts: a world's epoch names one batch
    for (size_t i = 0; i < b.worlds.size(); ++i) {
        b.worlds[i]->host_valid = false;
        b.worlds[i]->epoch = c->epoch;
        b.worlds[i]->batch_index = (uint32_t)i;
        b.worlds[i]->fused_src = b.fused;
        b.worlds[i]->prestepped = b.sub > 0;
        b.worlds[i]->sepoch = 0;
    }
    return b.first_err;
}

int batch_wait(r2d_ctx* c, Batch& b)
{
    if (b.fused) {
        int rc = fused_wait(c, b);
        if (rc) return rc;
        return batch_finish(c, b);
    }
    const size_t nw = b.worlds.size();
    for (;;) {
        HIPCHK(hipEventSynchronize(c->ev_end));
        if (c->trace && c->trace->marks.size() > 4096) c->trace->flush();
        uint32_t committed = c->hcounters[C_STEPS];
        b.done += committed;
        b.world_status.assign(c->hstatus, c->hstatus + nw);
        for (size_t i = 0; i < nw; ++i) {
            if (c->hstatus[i] == 5)
                std::snprintf(g_last_error, sizeof g_last_error, "solver schedule stalled in world %zu", i);
            if (c->hstatus[i] != 0 && b.first_err == R2D_OK)
                b.first_err = c->hstatus[i] == 4 ? R2D_E_NONFINITE : c->hstatus[i] == 5 ? R2D_E_DEVICE : R2D_E_RANGE;
        }
        for (int k = 0; k < 4; ++k) c->last_counts[k] = c->hcounters[k];
        if (c->hcounters[C_OVERFLOW]) {
            uint32_t e = std::max(c->hcounters[C_ENTRIES], c->cap_e);
            uint32_t r = std::max(c->hcounters[C_RAW], c->cap_r);
            uint32_t m = std::max(c->hcounters[C_MANIFOLDS], c->cap_m);
            uint32_t it = std::max(c->hcounters[C_ITEMS], c->cap_i);
            // a count is only known once the counts before it fit; grow the first short buffer
            if (c->hcounters[C_OVERFLOW] & 1u) e = c->hcounters[C_ENTRIES] * 2;
            if (c->hcounters[C_OVERFLOW] & 2u) r = c->hcounters[C_RAW] * 2;
            if (c->hcounters[C_OVERFLOW] & 4u) m = c->hcounters[C_MANIFOLDS] * 2;
            if (c->hcounters[C_OVERFLOW] & 8u) it = c->hcounters[C_ITEMS] * 2;
            int rc = ensure_scratch(c, (uint32_t)b.N, (uint32_t)b.B, (uint32_t)nw, (uint32_t)b.J, e + 1, r + 1,
                                    m + 1, it + 1);
            if (rc) return rc;
        }
        if (b.done >= b.steps) break;
        if (b.attempts > 16) return R2D_E_CAPACITY;
        int rc = batch_launch(c, b);
        if (rc) return rc;
    }
    return batch_finish(c, b);
}

// device time of the last launched batch (its final launch when it was relaunched)
float batch_gpu_ms(r2d_ctx* c)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_begin, c->ev_end) != hipSuccess) ms = 0.f;
    return ms;
}

int step_batch(r2d_ctx* c, r2d_world* const* worlds, size_t nw, float dt, size_t sub_steps, size_t iters,
               size_t steps)
{
    if (nw == 0 || steps == 0) return R2D_OK;
    for (size_t i = 0; i < nw; ++i)
        if (worlds[i] && worlds[i]->in_service) return R2D_E_BUSY;
    Batch b;
    int rc = batch_prepare(c, worlds, nw, dt, sub_steps, iters, steps, b);
    if (rc) return rc;
    if ((rc = batch_launch(c, b))) return rc;
    return batch_wait(c, b);
}

void fill_state(r2d_world* w, int32_t s, r2d_body_state* o)
{
    std::memset(o, 0, sizeof *o);
    o->id = w->u(F_ID)[s];
    uint32_t fl = w->u(F_FLAGS)[s];
    o->type = (fl & BF_RECT) ? 1u : 0u;
    o->is_static = (fl & BF_STATIC) ? 1u : 0u;
    o->num_normals = o->type ? 4u : 1u;
    o->pos_x = w->f(F_PX)[s]; o->pos_y = w->f(F_PY)[s];
    o->mom_x = w->f(F_MX)[s]; o->mom_y = w->f(F_MY)[s];
    o->force_x = w->f(F_FX)[s]; o->force_y = w->f(F_FY)[s];
    o->mass = w->f(F_MASS)[s]; o->angle = w->f(F_ANG)[s]; o->ang_momentum = w->f(F_L)[s];
    o->torque = w->f(F_TQ)[s]; o->inertia = w->f(F_INERTIA)[s]; o->mu = w->f(F_MU)[s];
    o->aabb_x = w->f(F_AX)[s]; o->aabb_y = w->f(F_AY)[s];
    o->aabb_half_width = w->f(F_AHW)[s]; o->aabb_half_height = w->f(F_AHH)[s];
    o->radius = w->f(F_RAD)[s];
    o->width = w->width[s]; o->height = w->height[s];
}

// ---------------------------------------------------------------- serving
// Admission estimate of one resident world: its device block and topology
// plus its share of the step scratch at the initial capacity estimates: the
// phase kernels' (per body: cell ranges, 4 cell entries, 8 raw candidates, 4
// manifolds and 4 solver items) or the fused kernel's slices and their free
// reserve, whichever is larger (per body 2 x (4 entries x 20 B + 32
// candidates x 32 B + 3 manifolds x 128 B + 3 items x 156 B)); per hash
// bucket 12 B; per joint one solver item.
constexpr size_t kScratchClassic = 96 + 4 * 24 + 8 * 36 + 4 * 128 + 4 * 160;
constexpr size_t kScratchFused = 96 + 2 * (4 * 20 + 32 * 32 + 3 * 128 + 3 * 156);
constexpr size_t kScratchPerBody = kScratchClassic > kScratchFused ? kScratchClassic : kScratchFused;
constexpr size_t kScratchPerBucket = 12;
constexpr size_t kScratchPerJoint = 160;
constexpr size_t kScratchPerWorld = 20 + sizeof(WorldDev);

size_t world_footprint(const r2d_world* w)
{
    size_t n = (size_t)w->n;
    size_t dev = (size_t)F_NFIELDS * std::max<size_t>(std::max(w->cap, w->dcap), 16) * 4 +
                 std::max<size_t>(16, w->joints.size()) * sizeof(Joint) +
                 std::max<size_t>(16, w->gens.size()) * sizeof(Gen) + std::max<size_t>(16, w->excl.size()) * 8;
    return dev + n * kScratchPerBody + (size_t)table_size(w) * kScratchPerBucket +
           w->joints.size() * kScratchPerJoint + kScratchPerWorld;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// A continuous-batching server: requests (a world and a number of steps) are
// admitted while the sum of admitted footprints fits the HBM budget and a
// batch slot is free.  Admitted worlds are stepped in lanes by size class,
// each lane with its own stream and step scratch (an internal context), so a
// lane of small worlds is not held to the step time of the largest world:
// every tick collects the lanes whose batch step has finished, retires
// finished requests, admits waiting ones (arrival order within a lane), and
// launches one step of every running world of each idle lane as one batch.
// Worlds stay resident on the device from admission to completion.
struct r2d_server {
    struct Req {
        r2d_world* w = nullptr;   // held while queued or running
        uint32_t n_bodies = 0;
        uint32_t lane = 0;
        uint32_t steps = 0, done = 0;
        size_t bytes = 0;
        int32_t state = R2D_REQ_QUEUED;
        int32_t status = R2D_OK;
        double t_submit = 0, t_admit = 0, t_done = 0;
        double max_step_ms = 0;
    };
    struct Lane {
        r2d_ctx* ctx = nullptr;               // own stream and step scratch
        std::deque<uint64_t> queue;           // waiting for admission, in arrival order
        std::vector<uint64_t> running;        // admitted, in order of admission
        bool inflight = false;
        Batch batch;                          // the launched batch step(s)
        std::vector<uint64_t> batch_tickets;
        uint32_t batch_steps = 1;             // steps of the launched batch
        double t_launch = 0;
    };
    r2d_ctx* home = nullptr;
    r2d_serve_options opt{};
    uint32_t nlanes = 1;
    Lane lanes[3];
    std::vector<Req> reqs;                    // ticket t is reqs[t - 1]
    size_t resident = 0, peak_resident = 0;
    size_t nrunning = 0;
    uint64_t rejected = 0, completed = 0, batch_steps = 0, scene_steps = 0, body_steps = 0;
    double gpu_ms = 0, wall_ms = 0;
    std::vector<float> step_ms;               // device time of each batch step
    std::vector<uint32_t> step_worlds;        // scenes in it
    uint32_t last_completed = 0;
    std::deque<uint64_t> done_tickets;        // finished since the last r2d_serve_completed
    double last_gpu_ms = 0, last_wall_ms = 0;
};

namespace {

uint32_t lane_of(const r2d_server* sv, uint32_t n)
{
    for (uint32_t l = 0; l + 1 < sv->nlanes; ++l)
        if (n <= sv->opt.split_bodies[l]) return l;
    return sv->nlanes - 1;
}

void retire(r2d_server* sv, uint64_t t, int32_t state, int32_t status)
{
    r2d_server::Req& q = sv->reqs[t - 1];
    if (q.state == R2D_REQ_RUNNING) sv->nrunning -= 1;
    q.state = state;
    q.status = status;
    q.t_done = now_ms();
    q.n_bodies = (uint32_t)q.w->n;
    q.w->in_service = false;
    q.w->ctx = sv->home;   // back to the caller's context
    q.w = nullptr;         // released: the caller may destroy the world
    sv->resident -= q.bytes;
    sv->completed += 1;
    sv->last_completed += 1;
    sv->done_tickets.push_back(t);
}

// wait for a lane's launched batch step and account for it
int lane_collect(r2d_server* sv, r2d_server::Lane& L)
{
    if (!L.inflight) return R2D_OK;
    L.inflight = false;
    int rc = batch_wait(L.ctx, L.batch);
    const uint32_t k = std::max<uint32_t>(L.batch_steps, 1);
    double wall = (now_ms() - L.t_launch) / k;   // per step of the launch
    float gms = batch_gpu_ms(L.ctx) / (float)k;
    const size_t nw = L.batch_tickets.size();
    // per-world failures arrive as world statuses; anything else fails the batch
    const bool per_world = rc == R2D_OK || L.batch.world_status.size() == nw;
    for (size_t i = 0; i < nw; ++i) {
        uint64_t t = L.batch_tickets[i];
        r2d_server::Req& q = sv->reqs[t - 1];
        if (!per_world) { retire(sv, t, R2D_REQ_FAILED, rc); continue; }
        int32_t ws = L.batch.world_status[i];
        if (ws != 0) {
            retire(sv, t, R2D_REQ_FAILED, ws == 4 ? R2D_E_NONFINITE : ws == 5 ? R2D_E_DEVICE : R2D_E_RANGE);
            continue;
        }
        q.done += k;
        q.max_step_ms = std::max(q.max_step_ms, (double)gms);
        sv->body_steps += (uint64_t)q.w->n * k;
        if (q.done >= q.steps) retire(sv, t, R2D_REQ_DONE, R2D_OK);
    }
    std::vector<uint64_t> still;
    for (uint64_t t : L.running)
        if (sv->reqs[t - 1].state == R2D_REQ_RUNNING) still.push_back(t);
    L.running.swap(still);
    sv->batch_steps += k;
    sv->scene_steps += nw * k;
    sv->gpu_ms += (double)gms * k;
    sv->wall_ms += wall * k;
    sv->step_ms.push_back(gms);
    sv->step_worlds.push_back((uint32_t)(nw * k));
    sv->last_gpu_ms = std::max(sv->last_gpu_ms, (double)gms);
    sv->last_wall_ms = std::max(sv->last_wall_ms, wall);
    return per_world ? R2D_OK : rc;
}

int server_collect(r2d_server* sv)
{
    int first = R2D_OK;
    for (uint32_t l = 0; l < sv->nlanes; ++l) {
        int rc = lane_collect(sv, sv->lanes[l]);
        if (rc && first == R2D_OK) first = rc;
    }
    return first;
}

uint32_t server_admit(r2d_server* sv)
{
    uint32_t n = 0;
    for (uint32_t l = 0; l < sv->nlanes; ++l) {
        r2d_server::Lane& L = sv->lanes[l];
        while (!L.queue.empty() && sv->nrunning < sv->opt.max_worlds) {
            uint64_t t = L.queue.front();
            r2d_server::Req& q = sv->reqs[t - 1];
            if (sv->resident + q.bytes > sv->opt.hbm_budget_bytes) break;   // no overtaking within a lane
            L.queue.pop_front();
            q.state = R2D_REQ_RUNNING;
            q.t_admit = now_ms();
            q.w->ctx = L.ctx;     // stepped (and read back) on the lane's stream until released
            sv->resident += q.bytes;
            sv->peak_resident = std::max(sv->peak_resident, sv->resident);
            sv->nrunning += 1;
            L.running.push_back(t);
            n += 1;
        }
    }
    return n;
}

int lane_launch(r2d_server* sv, r2d_server::Lane& L)
{
    if (L.running.empty() || L.inflight) return R2D_OK;
    std::vector<r2d_world*> ws;
    ws.reserve(L.running.size());
    for (uint64_t t : L.running) ws.push_back(sv->reqs[t - 1].w);
    L.ctx->solve_variant = sv->home->solve_variant;
    L.ctx->fused_enabled = sv->home->fused_enabled;
    // as many steps as every running world still needs, up to steps_per_tick
    uint32_t k = std::max<uint32_t>(sv->opt.steps_per_tick, 1);
    for (uint64_t t : L.running) {
        const r2d_server::Req& q = sv->reqs[t - 1];
        k = std::min<uint32_t>(k, q.steps > q.done ? q.steps - q.done : 1u);
    }
    int rc = batch_prepare(L.ctx, ws.data(), ws.size(), sv->opt.dt, sv->opt.sub_steps, sv->opt.collision_iters, k,
                           L.batch);
    if (!rc) rc = batch_launch(L.ctx, L.batch);
    if (rc) {
        for (uint64_t t : L.running) retire(sv, t, R2D_REQ_FAILED, rc);
        L.running.clear();
        return rc;
    }
    L.batch_tickets = L.running;
    L.batch_steps = k;
    L.t_launch = now_ms();
    L.inflight = true;
    return R2D_OK;
}

// percentile of the batch-step times, each weighted by its scene count
double weighted_pct(const std::vector<float>& v, const std::vector<uint32_t>& wts, double p)
{
    if (v.empty()) return 0.0;
    std::vector<std::pair<float, uint32_t>> a(v.size());
    uint64_t tot = 0;
    for (size_t i = 0; i < v.size(); ++i) { a[i] = {v[i], wts[i]}; tot += wts[i]; }
    std::sort(a.begin(), a.end());
    double target = p * (double)tot;
    uint64_t acc = 0;
    for (auto& e : a) {
        acc += e.second;
        if ((double)acc >= target) return e.first;
    }
    return a.back().first;
}

}  // namespace

// =================================================================== C ABI
R2D_EXPORT int r2d_ctx_create(int device, r2d_ctx** out)
{
    if (!out) return R2D_E_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return R2D_E_NO_DEVICE;
    if (device < 0 || device >= ndev) return R2D_E_INVALID_ARG;
    HIPCHK(hipSetDevice(device));
    r2d_ctx* c = new (std::nothrow) r2d_ctx();
    if (!c) return R2D_E_OOM;
    c->device = device;
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        c->num_cus = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return R2D_E_DEVICE; }
    for (int i = 0; i < 2; ++i) HIPCHK(hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) HIPCHK(hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming));
    HIPCHK(hipEventCreate(&c->ev_begin));
    HIPCHK(hipEventCreate(&c->ev_end));
    HIPCHK(hipEventCreate(&c->ev_kend));
    if (hipHostMalloc((void**)&c->hcounters, 64 * 4) != hipSuccess) { delete c; return R2D_E_OOM; }
    if (const char* e = std::getenv("R2D_NO_FUSED")) c->fused_enabled = !(e[0] == '1');
    set_solve_smem_limit();
    if (const char* e = std::getenv("R2D_NO_GRAPH")) c->use_graph = !(e[0] == '1');
    if (const char* e = std::getenv("R2D_TRACE_KERNELS")) {
        if (e[0] == '1') {
            c->trace = new (std::nothrow) KernelTrace();
            if (c->trace) c->trace->main = c->stream;
            c->use_graph = false;   // an event after every launch: eager launches
        }
    }
    *out = c;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_destroy(r2d_ctx* c)
{
    if (!c) return R2D_E_INVALID_ARG;
    // a server's lanes and the worlds of this context refer to it
    if (c->server || c->live_worlds) return R2D_E_BUSY;
    (void)hipStreamSynchronize(c->stream);
    if (c->trace) {
        c->trace->report(stderr);
        delete c->trace;
    }
    destroy_graphs(c);
    if (c->arena) (void)hipFree(c->arena);
    if (c->dworlds) (void)hipFree(c->dworlds);
    if (c->hcounters) (void)hipHostFree(c->hcounters);
    if (c->hstatus) (void)hipHostFree(c->hstatus);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c->fork);
    for (int i = 0; i < 2; ++i) (void)hipEventDestroy(c->join[i]);
    (void)hipEventDestroy(c->ev_begin);
    (void)hipEventDestroy(c->ev_end);
    (void)hipEventDestroy(c->ev_kend);
    if (c->farena) (void)hipFree(c->farena);
    if (c->fworld) (void)hipFree(c->fworld);
    if (c->tstamp) (void)hipFree(c->tstamp);
    if (c->wprof) (void)hipFree(c->wprof);
    if (c->hdone) (void)hipHostFree(c->hdone);
    if (c->hcnt) (void)hipHostFree(c->hcnt);
    if (c->hscr) (void)hipHostFree(c->hscr);
    if (c->dfsum) (void)hipFree(c->dfsum);
    if (c->hfsum) (void)hipHostFree(c->hfsum);
    if (c->ingest) (void)hipFree(c->ingest);
    pool_destroy(c);
    for (int i = 0; i < 2; ++i) (void)hipStreamDestroy(c->aux[i]);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_sync(r2d_ctx* c)
{
    if (!c) return R2D_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    return R2D_OK;
}

R2D_EXPORT int r2d_world_create(r2d_ctx* c, float cell, size_t mult, r2d_world** out)
{
    if (!c || !out) return R2D_E_INVALID_ARG;
    r2d_world* w = new (std::nothrow) r2d_world();
    if (!w) return R2D_E_OOM;
    w->ctx = c;
    w->owner = c;
    c->live_worlds += 1;
    w->cell_arg = cell;
    w->mult_arg = mult;
    *out = w;
    return R2D_OK;
}

R2D_EXPORT int r2d_world_destroy(r2d_world* w)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    // the world's last launches ran on its context's stream (a served world
    // is out of service only once its lane's launches are collected)
    (void)hipStreamSynchronize(w->ctx->stream);
    r2d_ctx* o = w->owner;
    pool_release(o, w->dblock);
    pool_release(o, w->djoints);
    pool_release(o, w->dgens);
    pool_release(o, w->dexcl);
    pool_release(o, w->dwtab);
    pool_release(o, w->dwcnt);
    pool_release(o, w->dwsum);
    o->live_worlds -= 1;
    delete w;
    return R2D_OK;
}

R2D_EXPORT int r2d_world_clear(r2d_world* w)
{
    if (!w) return R2D_E_INVALID_ARG;
    int rc = begin_edit(w);
    if (rc) return rc;
    std::fill(w->id2slot.begin(), w->id2slot.end(), -1);
    w->n = 0;
    w->warm_clear = true;
    w->joints.clear();
    w->gens.clear();
    w->excl.clear();
    w->topo_dirty = true;
    w->version += 1;
    w->epoch = 0;
    return R2D_OK;
}

static int make_body_common(r2d_world* w, const r2d_body_options* bo, bool rect, float a, float b, uint32_t* id_out)
{
    if (!w || !bo) return R2D_E_INVALID_ARG;
    int rc = begin_edit(w);
    if (rc) return rc;
    float mass;
    float inertia;
    if (!rect) {
        // makeDiscBody (lib.zig:73-82), Disc.init (Disc.zig:29-55)
        float r = a;
        mass = bo->mass_kind ? bo->mass_value : (float)M_PI * r * r * bo->mass_value;
        inertia = 0.5f * mass * r * r;
    } else {
        // makeRectangleBody (lib.zig:84-93), Rectangle.init (Rectangle.zig:33-64)
        mass = bo->mass_kind ? bo->mass_value : a * b * bo->mass_value;
        inertia = mass * (a * a + b * b) / 12;
    }
    uint32_t id = append_body(w);
    int32_t s = w->n - 1;
    w->f(F_PX)[s] = bo->pos_x; w->f(F_PY)[s] = bo->pos_y; w->f(F_ANG)[s] = bo->angle;
    w->f(F_FX)[s] = 0.f; w->f(F_FY)[s] = 0.f; w->f(F_TQ)[s] = 0.f;
    w->f(F_MASS)[s] = mass; w->f(F_INERTIA)[s] = inertia; w->f(F_MU)[s] = bo->mu;
    w->f(F_MX)[s] = bo->vel_x * mass; w->f(F_MY)[s] = bo->vel_y * mass;
    w->f(F_L)[s] = bo->omega * inertia;
    w->u(F_FLAGS)[s] = rect ? BF_RECT : 0u;
    if (!rect) {
        w->f(F_RAD)[s] = a; w->f(F_HW)[s] = 0.f; w->f(F_HH)[s] = 0.f;
        w->width[s] = 0.f; w->height[s] = 0.f;
        w->f(F_AX)[s] = bo->pos_x; w->f(F_AY)[s] = bo->pos_y; w->f(F_AHW)[s] = a; w->f(F_AHH)[s] = a;
    } else {
        float hw = a / 2, hh = b / 2;
        w->f(F_RAD)[s] = 0.f; w->f(F_HW)[s] = hw; w->f(F_HH)[s] = hh;
        w->width[s] = a; w->height[s] = b;
        // Rectangle.updateAABB (Rectangle.zig:71-86)
        v2 lv[4] = {mk(-hw, -hh), mk(-hw, hh), mk(hw, hh), mk(hw, -hh)};
        float c = cosf_r(bo->angle), sn = sinf_r(bo->angle);
        float ew = 0.f, eh = 0.f;
        for (int k = 0; k < 4; ++k) {
            v2 r = rotcs(lv[k], c, sn);
            if (r.x > ew) ew = r.x;
            if (r.y > eh) eh = r.y;
        }
        w->f(F_AX)[s] = bo->pos_x; w->f(F_AY)[s] = bo->pos_y; w->f(F_AHW)[s] = ew; w->f(F_AHH)[s] = eh;
    }
    if (id_out) *id_out = id;
    return R2D_OK;
}

R2D_EXPORT int r2d_make_disc(r2d_world* w, const r2d_body_options* bo, float radius, uint32_t* id_out)
{
    return make_body_common(w, bo, false, radius, 0.f, id_out);
}

R2D_EXPORT int r2d_make_rect(r2d_world* w, const r2d_body_options* bo, float width, float height, uint32_t* id_out)
{
    return make_body_common(w, bo, true, width, height, id_out);
}

R2D_EXPORT int r2d_set_static(r2d_world* w, uint32_t id, int is_static)
{
    if (!w) return R2D_E_INVALID_ARG;
    int32_t s = w->slot_of(id);
    if (s < 0) return R2D_E_INVALID_ID;
    int rc = begin_edit(w);
    if (rc) return rc;
    uint32_t& fl = w->u(F_FLAGS)[s];
    fl = is_static ? (fl | BF_STATIC) : (fl & ~BF_STATIC);
    return R2D_OK;
}

R2D_EXPORT int r2d_remove_body(r2d_world* w, uint32_t id)
{
    if (!w) return R2D_E_INVALID_ARG;
    int32_t s = w->slot_of(id);
    if (s < 0) return R2D_E_INVALID_ID;
    int rc = begin_edit(w);
    if (rc) return rc;
    int32_t last = w->n - 1;
    if (s != last) {
        for (int fld = 0; fld < F_NFIELDS; ++fld) w->block[(size_t)fld * w->cap + s] = w->block[(size_t)fld * w->cap + last];
        w->width[s] = w->width[last];
        w->height[s] = w->height[last];
        w->id2slot[w->u(F_ID)[s]] = s;
    }
    w->id2slot[id] = -1;
    w->n -= 1;
    w->warm_clear = true;   // the warm tables are kept under slots
    w->topo_dirty = true;   // joint / spring slots move
    w->version += 1;
    w->epoch = 0;
    return R2D_OK;
}

R2D_EXPORT int r2d_make_gravity(r2d_world* w, float g)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    HGen h{};
    h.type = 0; h.gx = 0.0f; h.gy = -g;
    w->gens.push_back(h);
    w->topo_dirty = true;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_make_spring(r2d_world* w, uint32_t id, float k, float c, float ax, float ay)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    if (w->slot_of(id) < 0) return R2D_E_INVALID_ID;
    HGen h{};
    h.type = 1; h.id = id; h.k = k; h.c = c; h.ax = ax; h.ay = ay;
    w->gens.push_back(h);
    w->topo_dirty = true;
    w->version += 1;
    return R2D_OK;
}

static const r2d_constraint_params kDefaultParams = {INFINITY, -INFINITY, 10.0f};

static int add_joint(r2d_world* w, const r2d_constraint_params* p, HJoint h, size_t* handle)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    if (!p) p = &kDefaultParams;
    h.pmax = p->power_max; h.pmin = p->power_min; h.beta = p->beta;
    w->joints.push_back(h);
    w->topo_dirty = true;
    w->version += 1;
    if (handle) *handle = w->joints.size() - 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_make_distance_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id1, uint32_t id2,
                                       float target, size_t* handle)
{
    HJoint h{};
    h.type = 0; h.id1 = id1; h.id2 = id2; h.target = target;
    return add_joint(w, p, h, handle);
}

R2D_EXPORT int r2d_make_offset_distance_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id1,
                                              uint32_t id2, float r1x, float r1y, float r2x, float r2y,
                                              float target, size_t* handle)
{
    HJoint h{};
    h.type = 1; h.id1 = id1; h.id2 = id2; h.r1x = r1x; h.r1y = r1y; h.r2x = r2x; h.r2y = r2y; h.target = target;
    return add_joint(w, p, h, handle);
}

R2D_EXPORT int r2d_make_fixed_position_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id, float tx,
                                             float ty, size_t* handle)
{
    HJoint h{};
    h.type = 2; h.id1 = id; h.tx = tx; h.ty = ty;
    return add_joint(w, p, h, handle);
}

R2D_EXPORT int r2d_make_motor_joint(r2d_world* w, const r2d_constraint_params* p, uint32_t id, float omega,
                                    size_t* handle)
{
    HJoint h{};
    h.type = 3; h.id1 = id; h.target = omega;
    return add_joint(w, p, h, handle);
}

R2D_EXPORT int r2d_exclude_collision_pair(r2d_world* w, uint32_t a, uint32_t b)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    uint64_t k = a < b ? (((uint64_t)a << 32) | b) : (((uint64_t)b << 32) | a);
    auto it = std::lower_bound(w->excl.begin(), w->excl.end(), k);
    if (it == w->excl.end() || *it != k) w->excl.insert(it, k);
    w->topo_dirty = true;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_process(r2d_world* w, float dt, size_t sub_steps, size_t iters)
{
    if (!w) return R2D_E_INVALID_ARG;
    return step_batch(w->ctx, &w, 1, dt, sub_steps, iters, 1);
}

R2D_EXPORT int r2d_process_batch(r2d_ctx* c, r2d_world* const* worlds, size_t n, float dt, size_t sub_steps,
                                 size_t iters, size_t steps)
{
    if (!c || (!worlds && n)) return R2D_E_INVALID_ARG;
    return step_batch(c, worlds, n, dt, sub_steps, iters, steps);
}

R2D_EXPORT int r2d_process_static(r2d_world* w, float dt, size_t sub_steps, size_t iters)
{
    if (!w) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    r2d_ctx* c = w->ctx;
    // the world must be resident: its device block and topology uploaded by
    // an earlier step and no bodies added since (state edits are fine)
    if (w->n > 0 && (!w->dblock || w->dcap != w->cap || w->topo_dirty)) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "r2d_process_static: the world is not resident on the device (step it once with r2d_process "
                      "after adding bodies, joints or force generators)");
        return R2D_E_INVALID_ARG;
    }
    if (w->sched != 0 || w->warm > 0.0f) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "r2d_process_static: only the reference order without warm starting");
        return R2D_E_INVALID_ARG;
    }
    const uint32_t n = (uint32_t)w->n, tb = table_size(w), nj = (uint32_t)w->joints.size();
    if (n > ST_N || tb > ST_B || nj > ST_J) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "r2d_process_static: %u bodies, %u buckets, %u joints; at most %u, %u, %u", n, tb, nj, ST_N,
                      ST_B, ST_J);
        return R2D_E_CAPACITY;
    }
    // the reference raises InvalidRigidBodyId from the first constraint solve
    if (sub_steps > 0 && iters > 0)
        for (const HJoint& h : w->joints)
            if (w->slot_of(h.id1) < 0 || (h.type <= 1 && w->slot_of(h.id2) < 0)) return R2D_E_INVALID_ID;
    int variant = c->solve_variant >= 0 ? c->solve_variant : default_variant(1);
    if (!fused_available(variant)) variant = 11;
    size_t lds = fused_lds_bytes(n, tb, fused_slots(variant), fused_threads(variant));
    if (lds > kFusedLdsMax) {
        variant = 11;
        lds = fused_lds_bytes(n, tb, fused_slots(variant), fused_threads(variant));
        if (lds > kFusedLdsMax) return R2D_E_CAPACITY;
    }
    if (n == 0) {   // nothing moves and no manifold exists
        w->epoch = 0;
        w->sepoch = 0;
        return R2D_OK;
    }
    const int dev = c->device;
    if (dev < 0 || dev >= kMaxDevices) return R2D_E_CAPACITY;
    if (!g_scratch_ptr[dev]) {
        // first call on this device: the static scratch's address here
        HIPCHK(hipSetDevice(dev));
        const int rc = static_scratch_setup(dev);
        if (rc) return rc;
    }
    char* base = (char*)g_scratch_ptr[dev];
    std::lock_guard<std::mutex> guard(g_scratch_use[dev]);
    hipStream_t st = c->stream;
    if (!w->dev_valid) {
        HIPCHK(hipMemcpyAsync(w->dblock, w->block.data(), (size_t)F_NFIELDS * w->cap * 4, hipMemcpyHostToDevice, st));
        w->dev_valid = true;
    }
    if (w->desc_version != w->version) {
        w->desc = make_desc(w);
        w->desc_version = w->version;
    }
    StepDev F;
    WorldDev* ddesc = nullptr;
    static_layout(base, &F, &ddesc);
    F.nb = n; F.nbuckets = tb; F.njoints = nj;
    // host values the asynchronous uploads read; the call waits for the
    // stream before they go out of scope
    WorldDev d = w->desc;
    d.body_base = 0; d.bucket_base = 0; d.joint_base = 0;
    const WorldScratch x{0u, ST_E, 0u, ST_R, 0u, ST_M, 0u, ST_I};
    // free ends at the capacities: a world that outgrows a slice stops
    // before its step (WS_OVERFLOW) instead of taking more
    const uint32_t pool[8] = {ST_E + 1, ST_R + 1, ST_M, ST_I, ST_E + 1, ST_R + 1, ST_M, ST_I};
    const uint32_t order = 0u;
    HIPCHK(hipMemcpyAsync(ddesc, &d, sizeof d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(F.wscr, &x, sizeof x, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(F.fpool, pool, sizeof pool, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(F.w_order, &order, 4, hipMemcpyHostToDevice, st));
    // what the fused kernel expects to start at zero; the array keeps the
    // previous call's contents otherwise
    HIPCHK(hipMemsetAsync(F.counters, 0, 64 * 4, st));
    HIPCHK(hipMemsetAsync(F.w_done, 0, 4, st));
    HIPCHK(hipMemsetAsync(F.w_status, 0, 4, st));
    HIPCHK(hipMemsetAsync(F.w_cnt, 0, 32, st));
    (void)hipGetLastError();
    launch_fused(F, variant, 1u, dt, (uint32_t)sub_steps, (uint32_t)iters, 1u, lds, st);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        note_error("launch_fused (r2d_process_static)", le, __LINE__);
        return R2D_E_DEVICE;
    }
    int32_t status = 0;
    uint32_t cnt[8] = {};
    HIPCHK(hipMemcpyAsync(&status, F.w_status, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cnt, F.w_cnt, 32, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    g_scratch_call[dev] = g_scratch_calls.fetch_add(1) + 1;
    w->epoch = 0;
    w->sepoch = 0;
    if (status == WS_OVERFLOW) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "r2d_process_static: the step needs %u cell entries, %u candidate pairs, %u manifolds, %u "
                      "solver items; the static scratch holds %u, %u, %u, %u (the world is unchanged)",
                      cnt[4], cnt[5], cnt[6], cnt[7], ST_E, ST_R, ST_M, ST_I);
        return R2D_E_CAPACITY;
    }
    w->host_valid = false;
    w->sepoch = g_scratch_call[dev];
    w->prestepped = sub_steps > 0;
    if (status == WS_NONFINITE) return R2D_E_NONFINITE;
    if (status == WS_STALL) {
        std::snprintf(g_last_error, sizeof g_last_error, "solver schedule stalled (r2d_process_static)");
        return R2D_E_DEVICE;
    }
    if (status != WS_OK) return R2D_E_RANGE;
    return R2D_OK;
}

R2D_EXPORT int r2d_num_bodies(r2d_world* w, size_t* n)
{
    if (!w || !n) return R2D_E_INVALID_ARG;
    *n = (size_t)w->n;
    return R2D_OK;
}

R2D_EXPORT int r2d_body_id_at(r2d_world* w, size_t i, uint32_t* id)
{
    if (!w || !id) return R2D_E_INVALID_ARG;
    if (i >= (size_t)w->n) return R2D_E_RANGE;
    *id = w->u(F_ID)[i];
    return R2D_OK;
}

R2D_EXPORT int r2d_get_body(r2d_world* w, uint32_t id, r2d_body_state* out)
{
    if (!w || !out) return R2D_E_INVALID_ARG;
    int32_t s = w->slot_of(id);
    if (s < 0) return R2D_E_INVALID_ID;
    int rc = download_bodies(w);
    if (rc) return rc;
    fill_state(w, s, out);
    return R2D_OK;
}

R2D_EXPORT int r2d_set_body(r2d_world* w, uint32_t id, const r2d_body_state* in)
{
    if (!w || !in) return R2D_E_INVALID_ARG;
    int32_t s = w->slot_of(id);
    if (s < 0) return R2D_E_INVALID_ID;
    int rc = begin_edit(w);
    if (rc) return rc;
    w->f(F_PX)[s] = in->pos_x; w->f(F_PY)[s] = in->pos_y;
    w->f(F_MX)[s] = in->mom_x; w->f(F_MY)[s] = in->mom_y;
    w->f(F_FX)[s] = in->force_x; w->f(F_FY)[s] = in->force_y;
    w->f(F_ANG)[s] = in->angle; w->f(F_L)[s] = in->ang_momentum; w->f(F_TQ)[s] = in->torque;
    return R2D_OK;
}

R2D_EXPORT int r2d_get_bodies(r2d_world* w, r2d_body_state* out, size_t cap, size_t* n)
{
    if (!w) return R2D_E_INVALID_ARG;
    int rc = download_bodies(w);
    if (rc) return rc;
    if (n) *n = (size_t)w->n;
    if (!out) return R2D_OK;
    size_t m = std::min(cap, (size_t)w->n);
    for (size_t i = 0; i < m; ++i) fill_state(w, (int32_t)i, &out[i]);
    return R2D_OK;
}

R2D_EXPORT int r2d_get_manifolds(r2d_world* w, r2d_manifold_state* out, size_t cap, size_t* n)
{
    if (!w) return R2D_E_INVALID_ARG;
    r2d_ctx* c = w->ctx;
    // a launched server step is rewriting the step scratch: finish it first
    if (c->server) {
        int rc = server_collect(c->server);
        if (rc) return rc;
    }
    // manifolds of r2d_process_static: in the static scratch until the next call on the device
    std::unique_lock<std::mutex> guard;
    StepDev S;
    const bool from_static = w->epoch == 0 && w->sepoch != 0;
    if (from_static) {
        guard = std::unique_lock<std::mutex>(g_scratch_use[c->device]);
        if (g_scratch_call[c->device] != w->sepoch) {
            if (n) *n = 0;
            std::snprintf(g_last_error, sizeof g_last_error,
                          "the manifolds of this world's r2d_process_static step were overwritten by a later call");
            return R2D_E_RANGE;
        }
        WorldDev* dd = nullptr;
        static_layout((char*)g_scratch_ptr[c->device], &S, &dd);
    } else if (w->epoch == 0 || w->epoch != c->epoch) {
        if (n) *n = 0;
        return w->epoch == 0 ? R2D_OK : R2D_E_RANGE;
    } else {
        // manifolds of the fused kernel are in the world's slice of the fused arena
        S = w->fused_src ? fused_view(c) : c->S;
    }
    const uint32_t bidx = from_static ? 0u : w->batch_index;
    uint32_t b[2];
    if (from_static || w->fused_src) {
        HIPCHK(hipMemcpy(b, S.w_man_begin + bidx, 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(b + 1, S.w_man_end + bidx, 4, hipMemcpyDeviceToHost));
    } else {
        HIPCHK(hipMemcpy(b, S.w_man_begin + bidx, 8, hipMemcpyDeviceToHost));
    }
    size_t cnt = b[1] - b[0];
    if (n) *n = cnt;
    if (!out) return R2D_OK;
    size_t m = std::min(cap, cnt);
    if (m == 0) return R2D_OK;
    std::vector<uint32_t> ref(m), inc(m), valid(m);
    std::vector<float> nx(m), ny(m), fr(m), pt((size_t)MP_NFIELDS * 2 * m);
    HIPCHK(hipMemcpy(ref.data(), S.m_ref + b[0], m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(inc.data(), S.m_inc + b[0], m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(valid.data(), S.m_valid + b[0], m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nx.data(), S.m_nx + b[0], m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ny.data(), S.m_ny + b[0], m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fr.data(), S.m_fric + b[0], m * 4, hipMemcpyDeviceToHost));
    for (int f = 0; f < MP_NFIELDS * 2; ++f)
        HIPCHK(hipMemcpy(pt.data() + (size_t)f * m, S.m_pt + (size_t)f * S.cap_m + b[0], m * 4, hipMemcpyDeviceToHost));
    int rc = download_bodies(w);
    if (rc) return rc;
    uint32_t wb[1] = {0u};
    // batch body base of this world = first body's batch index
    if (!from_static) HIPCHK(hipMemcpy(wb, &c->dworlds[w->batch_index].body_base, 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < m; ++i) {
        r2d_manifold_state& o = out[i];
        std::memset(&o, 0, sizeof o);
        o.ref_id = w->u(F_ID)[ref[i] - wb[0]];
        o.inc_id = w->u(F_ID)[inc[i] - wb[0]];
        o.normal_x = nx[i]; o.normal_y = ny[i];
        // preStep (collision.zig:102-133) sets the tangent, the friction and
        // the points' effective masses; after a step without sub-steps they
        // keep their initial zeros (the kernels precompute them at the
        // record build regardless)
        const bool pre = w->prestepped;
        o.tangent_x = pre ? ny[i] : 0.0f; o.tangent_y = pre ? -nx[i] : 0.0f;
        o.friction = pre ? fr[i] : 0.0f;
        for (int p = 0; p < 2; ++p) {
            r2d_manifold_point& q = o.points[p];
            q.valid = (valid[i] >> p) & 1u;
            if (!q.valid) continue;
            auto F = [&](int fld) { return pt[((size_t)fld * 2 + p) * m + i]; };
            q.acc_pn = F(MP_ACC_PN); q.acc_pt = F(MP_ACC_PT);
            q.ref_r_x = F(MP_R1X); q.ref_r_y = F(MP_R1Y); q.inc_r_x = F(MP_R2X); q.inc_r_y = F(MP_R2Y);
            q.pos_x = F(MP_POSX); q.pos_y = F(MP_POSY);
            q.depth = F(MP_DEPTH); q.original_depth = F(MP_ODEPTH);
            q.mass_n = pre ? F(MP_MASS_N) : 0.0f; q.mass_t = pre ? F(MP_MASS_T) : 0.0f;
        }
    }
    return R2D_OK;
}

R2D_EXPORT int r2d_set_hash_params(r2d_world* w, int mode, float cell, size_t mult)
{
    if (!w || (mode != 0 && mode != 1)) return R2D_E_INVALID_ARG;
    if (mode == 1 && (!(cell > 0.0f) || mult == 0)) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    w->hash_mode = mode;
    w->hash_cell = cell;
    w->hash_mult = mult;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_set_schedule(r2d_world* w, int mode)
{
    if (!w || (mode != 0 && mode != 1)) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    if (mode == 1 && w->warm > 0.0f) return R2D_E_INVALID_ARG;   // warm starting runs in the reference order
    w->sched = mode;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_set_warm_start(r2d_world* w, float factor)
{
    if (!w || !(factor >= 0.0f && factor <= 1.0f)) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    if (factor > 0.0f && w->sched != 0) return R2D_E_INVALID_ARG;
    w->warm = factor;
    w->warm_clear = true;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_set_profiling(r2d_ctx* c, int on)
{
    if (!c) return R2D_E_INVALID_ARG;
    c->profiling = on != 0;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_phase_times(r2d_ctx* c, double* out_ms, uint64_t* steps)
{
    if (!c) return R2D_E_INVALID_ARG;
    int rc = timer_flush(c);
    if (rc) return rc;
    if (out_ms) for (int i = 0; i < 4; ++i) out_ms[i] = c->phase_ms[i];
    if (steps) *steps = c->phase_steps;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_step_times(r2d_ctx* c, float* out, size_t cap, size_t* n)
{
    if (!c) return R2D_E_INVALID_ARG;
    int rc = timer_flush(c);
    if (rc) return rc;
    if (n) *n = c->step_ms.size();
    if (out) std::memcpy(out, c->step_ms.data(), std::min(cap, c->step_ms.size()) * sizeof(float));
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_reset_phase_times(r2d_ctx* c)
{
    if (!c) return R2D_E_INVALID_ARG;
    int rc = timer_flush(c);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) c->phase_ms[i] = 0;
    c->phase_steps = 0;
    c->step_ms.clear();
    c->fused_ms = 0.0;
    c->fused_launches = 0;
    c->fused_steps = 0;
    for (int k = 0; k < FP_N; ++k) c->fused_phase_ms[k] = 0.0;
    for (int k = 0; k < 5; ++k) c->fused_sums[k] = 0;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_fused_counts(r2d_ctx* c, uint64_t* out)
{
    if (!c || !out) return R2D_E_INVALID_ARG;
    for (int k = 0; k < 5; ++k) out[k] = c->fused_sums[k];
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_fused_phase_times(r2d_ctx* c, double* out_ms)
{
    if (!c || !out_ms) return R2D_E_INVALID_ARG;
    for (int k = 0; k < FP_N; ++k) out_ms[k] = c->fused_phase_ms[k];
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_fused_times(r2d_ctx* c, double* kernel_ms, uint64_t* launches, uint64_t* world_steps)
{
    if (!c) return R2D_E_INVALID_ARG;
    if (kernel_ms) *kernel_ms = c->fused_ms;
    if (launches) *launches = c->fused_launches;
    if (world_steps) *world_steps = c->fused_steps;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_set_fused(r2d_ctx* c, int on)
{
    if (!c) return R2D_E_INVALID_ARG;
    c->fused_enabled = on != 0;
    return R2D_OK;
}

__global__ void k_debug_eval(int op, const float* in, float* out, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float x = in[i];
        float r;
        switch (op) {
        case 0: r = sinf_r(x); break;
        case 1: r = cosf_r(x); break;
        case 2: r = sqrt_rn(x); break;
        case 3: r = 1.0f / x; break;
        default: r = in[2 * i] / in[2 * i + 1]; break;
        }
        out[i] = r;
    }
}

R2D_EXPORT int r2d_debug_eval(r2d_ctx* c, int op, const float* in, float* out, size_t n)
{
    if (!c || !in || !out || op < 0 || op > 4) return R2D_E_INVALID_ARG;
    if (n == 0) return R2D_OK;
    size_t nin = op == 4 ? 2 * n : n;
    float *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, nin * 4) != hipSuccess) return R2D_E_OOM;
    if (hipMalloc(&dout, n * 4) != hipSuccess) { (void)hipFree(din); return R2D_E_OOM; }
    HIPCHK(hipMemcpyAsync(din, in, nin * 4, hipMemcpyHostToDevice, c->stream));
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_debug_eval, dim3(blocks), dim3(256), 0, c->stream, op, din, dout, (uint32_t)n);
    HIPCHK(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(din);
    (void)hipFree(dout);
    return R2D_OK;
}

// ---- measured peaks: HBM read and copy streams, LDS reads
// Grid-stride stream over n16 16-byte elements, four independent loads in
// flight per thread (n16 is a multiple of 4 x the grid's threads, so every
// load is in bounds); the XOR of what was read goes to *sink only when it
// equals a value it cannot take in practice, which keeps the loads.
__global__ __launch_bounds__(256) void k_peak_read(const uint4* __restrict__ a, size_t n16, uint32_t* sink)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t acc = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += 4 * stride) {
        const uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        acc ^= v0.x ^ v0.y ^ v0.z ^ v0.w ^ v1.x ^ v1.y ^ v1.z ^ v1.w ^ v2.x ^ v2.y ^ v2.z ^ v2.w ^ v3.x ^ v3.y ^
               v3.z ^ v3.w;
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}

__global__ __launch_bounds__(256) void k_peak_copy(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += 4 * stride) {
        const uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        b[i] = v0; b[i + stride] = v1; b[i + 2 * stride] = v2; b[i + 3 * stride] = v3;
    }
}

// LDS reads: 16 KB per workgroup, each wave reading 64 consecutive 16-byte
// rows per load (conflict-free), eight loads per round; the rows move with
// the round so no load is loop-invariant
constexpr uint32_t PEAK_LDS_ROWS = 1024;
__global__ __launch_bounds__(256) void k_peak_lds(uint32_t rounds, uint32_t* sink)
{
    __shared__ uint4 buf[PEAK_LDS_ROWS];
    for (uint32_t i = threadIdx.x; i < PEAK_LDS_ROWS; i += blockDim.x)
        buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
    __syncthreads();
    uint32_t acc = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            const uint4 v = buf[(threadIdx.x + k * 128u + r * 16u) & (PEAK_LDS_ROWS - 1u)];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}

R2D_EXPORT int r2d_debug_peaks(r2d_ctx* c, double* out)
{
    if (!c || !out) return R2D_E_INVALID_ARG;
    HIPCHK(hipSetDevice(c->device));
    // 1 GiB per buffer: a multiple of 4 x every grid's threads swept below
    const unsigned threads = 256;
    const size_t n16 = ((size_t)1 << 30) / 16;
    uint4 *a = nullptr, *b = nullptr;
    uint32_t* sink = nullptr;
    if (hipMalloc((void**)&a, n16 * 16) != hipSuccess) return R2D_E_OOM;
    if (hipMalloc((void**)&b, n16 * 16) != hipSuccess) { (void)hipFree(a); return R2D_E_OOM; }
    if (hipMalloc((void**)&sink, 4) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return R2D_E_OOM; }
    // every HIP status after the allocations is kept, so the buffers are
    // freed on any failure
    hipError_t err = hipSuccess;
    auto chk = [&](hipError_t e) { if (err == hipSuccess && e != hipSuccess) err = e; };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    chk(hipEventCreate(&e0));
    chk(hipEventCreate(&e1));
    chk(hipMemsetAsync(a, 1, n16 * 16, c->stream));
    chk(hipMemsetAsync(b, 0, n16 * 16, c->stream));
    auto timed = [&](auto&& launch, int reps) -> double {
        if (err != hipSuccess) return 0.0;
        launch();                                  // warm-up
        chk(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r) launch();
        chk(hipEventRecord(e1, c->stream));
        chk(hipEventSynchronize(e1));
        float ms = 0.f;
        chk(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / 1e3 / reps;
    };
    // the streams at several grid sizes (1,024-16,384 workgroups): the fastest counts
    double t_read = 0.0, t_copy = 0.0;
    for (unsigned blocks = 1024; blocks <= 16384; blocks *= 2) {
        const double tr = timed([&] {
            hipLaunchKernelGGL(k_peak_read, dim3(blocks), dim3(threads), 0, c->stream, a, n16, sink);
        }, 5);
        const double tc = timed([&] {
            hipLaunchKernelGGL(k_peak_copy, dim3(blocks), dim3(threads), 0, c->stream, a, b, n16);
        }, 5);
        if (tr > 0.0 && (t_read == 0.0 || tr < t_read)) t_read = tr;
        if (tc > 0.0 && (t_copy == 0.0 || tc < t_copy)) t_copy = tc;
    }
    const uint32_t rounds = 4096, lds_blocks = 256 * 8;
    const double t_lds = timed([&] {
        hipLaunchKernelGGL(k_peak_lds, dim3(lds_blocks), dim3(threads), 0, c->stream, rounds, sink);
    }, 5);
    chk(hipGetLastError());
    chk(hipStreamSynchronize(c->stream));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(sink);
    if (err != hipSuccess || t_read <= 0.0 || t_copy <= 0.0 || t_lds <= 0.0) {
        note_error("r2d_debug_peaks", err, __LINE__);
        return R2D_E_DEVICE;
    }
    out[0] = (double)(n16 * 16) / t_read / 1e9;
    out[1] = (double)(2 * n16 * 16) / t_copy / 1e9;
    out[2] = (double)lds_blocks * threads * 16.0 * 8.0 * rounds / t_lds / 1e9;
    return R2D_OK;
}

R2D_EXPORT int r2d_last_error(char* buf, size_t n)
{
    if (!buf || n == 0) return R2D_E_INVALID_ARG;
    std::snprintf(buf, n, "%s", g_last_error);
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_set_solve_variant(r2d_ctx* c, int variant)
{
    if (!c || variant < -1 || variant >= SOLVE_VARIANTS) return R2D_E_INVALID_ARG;
    c->solve_variant = variant;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_set_lds_limit(r2d_ctx* c, uint32_t max_bodies)
{
    if (!c) return R2D_E_INVALID_ARG;
    c->lds_nmax = max_bodies;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_set_wide_record(r2d_ctx* c, int on)
{
    if (!c) return R2D_E_INVALID_ARG;
    c->wide_record = on != 0;
    return R2D_OK;
}

R2D_EXPORT int r2d_ctx_last_counts(r2d_ctx* c, uint64_t* out)
{
    if (!c || !out) return R2D_E_INVALID_ARG;
    for (int i = 0; i < 4; ++i) out[i] = c->last_counts[i];
    return R2D_OK;
}

// EntityFactory.makeDiscBody / makeRectangleBody (lib.zig:73-93) and the
// shape constructors (Disc.zig:29-55, Rectangle.zig:33-64) for n bodies on
// the device: the arithmetic of make_body_common, in the same order, written
// into the world's device block at slots n0 .. n0 + n - 1 with ids id0 ..
__global__ void k_make_bodies(uint32_t* blk, size_t cap, uint32_t n0, uint32_t n, const r2d_body_options* bo,
                              const int32_t* kind, const float* dim_a, const float* dim_b, uint32_t id0)
{
    float* f = reinterpret_cast<float*>(blk);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const r2d_body_options o = bo[i];
        const bool rect = kind[i] == 1;
        const float a = dim_a[i], b = rect ? dim_b[i] : 0.f;
        float mass, inertia;
        if (!rect) {
            mass = o.mass_kind ? o.mass_value : (float)M_PI * a * a * o.mass_value;
            inertia = 0.5f * mass * a * a;
        } else {
            mass = o.mass_kind ? o.mass_value : a * b * o.mass_value;
            inertia = mass * (a * a + b * b) / 12;
        }
        const size_t s = (size_t)n0 + i;
        auto F = [&](int fld) -> float& { return f[(size_t)fld * cap + s]; };
        F(F_PX) = o.pos_x; F(F_PY) = o.pos_y; F(F_ANG) = o.angle;
        F(F_FX) = 0.f; F(F_FY) = 0.f; F(F_TQ) = 0.f;
        F(F_MASS) = mass; F(F_INERTIA) = inertia; F(F_MU) = o.mu;
        F(F_MX) = o.vel_x * mass; F(F_MY) = o.vel_y * mass;
        F(F_L) = o.omega * inertia;
        blk[(size_t)F_FLAGS * cap + s] = rect ? BF_RECT : 0u;
        blk[(size_t)F_ID * cap + s] = id0 + i;
        F(F_AX) = o.pos_x; F(F_AY) = o.pos_y;
        if (!rect) {
            F(F_RAD) = a; F(F_HW) = 0.f; F(F_HH) = 0.f;
            F(F_AHW) = a; F(F_AHH) = a;
        } else {
            const float hw = a / 2, hh = b / 2;
            F(F_RAD) = 0.f; F(F_HW) = hw; F(F_HH) = hh;
            // Rectangle.updateAABB (Rectangle.zig:71-86)
            v2 lv[4] = {mk(-hw, -hh), mk(-hw, hh), mk(hw, hh), mk(hw, -hh)};
            const float c = cosf_r(o.angle), sn = sinf_r(o.angle);
            float ew = 0.f, eh = 0.f;
            for (int k = 0; k < 4; ++k) {
                v2 r = rotcs(lv[k], c, sn);
                if (r.x > ew) ew = r.x;
                if (r.y > eh) eh = r.y;
            }
            F(F_AHW) = ew; F(F_AHH) = eh;
        }
    }
}

namespace {
// bodies per r2d_make_bodies call from which the device builds them
constexpr size_t kDeviceIngestMin = 32;

// the per-context staging area of r2d_make_bodies' inputs
int ensure_ingest(r2d_ctx* c, size_t n)
{
    if (n <= c->ingest_cap && c->ingest) return R2D_OK;
    const size_t cap = std::max<size_t>(n, 2 * c->ingest_cap);
    if (c->ingest) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->ingest));
    }
    c->ingest = nullptr;
    c->ingest_cap = 0;
    if (hipMalloc(&c->ingest, cap * (sizeof(r2d_body_options) + 12) + 1024) != hipSuccess) return R2D_E_OOM;
    c->ingest_cap = cap;
    return R2D_OK;
}
}  // namespace

R2D_EXPORT int r2d_make_bodies(r2d_world* w, size_t n, const r2d_body_options* bo, const int32_t* kind,
                               const float* dim_a, const float* dim_b, uint32_t* ids_out)
{
    if (!w || (n && (!bo || !kind || !dim_a || !dim_b))) return R2D_E_INVALID_ARG;
    for (size_t i = 0; i < n; ++i)
        if (kind[i] != 0 && kind[i] != 1) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    if (n < kDeviceIngestMin) {
        int rc = begin_edit(w);
        if (rc) return rc;
        grow_host(w, (size_t)w->n + n);
        for (size_t i = 0; i < n; ++i) {
            uint32_t id = 0;
            rc = make_body_common(w, &bo[i], kind[i] == 1, dim_a[i], kind[i] == 1 ? dim_b[i] : 0.f, &id);
            if (rc) return rc;
            if (ids_out) ids_out[i] = id;
        }
        return R2D_OK;
    }
    // On the device (appendBody order, lib.zig:66-71): the world's device
    // block becomes the current copy with the new bodies at the end; the host
    // mirror is refreshed when the host next reads it (ids, shapes and the
    // id -> slot map are kept on the host at once).
    r2d_ctx* c = w->ctx;
    hipStream_t st = c->stream;
    const size_t n0 = (size_t)w->n, need = n0 + n;
    if (need > 0x7fffffffull || (uint64_t)w->current_body_id + n > 0xffffffffull) return R2D_E_CAPACITY;
    int rc;
    if (need > w->cap || need > w->dcap || !w->dblock) {
        // grow: the current state through the host mirror into a larger block
        if ((rc = download_bodies(w))) return rc;
        grow_host(w, need);
        if ((rc = ensure_world_device(w))) return rc;
        if ((rc = upload_bodies(w))) return rc;
    } else if (!w->dev_valid) {
        if ((rc = upload_bodies(w))) return rc;
    }
    if ((rc = ensure_ingest(c, n))) return rc;
    char* q = (char*)c->ingest;
    r2d_body_options* dbo = (r2d_body_options*)q;
    int32_t* dkind = (int32_t*)(q + align256(n * sizeof(r2d_body_options)));
    float* da = (float*)((char*)dkind + align256(n * 4));
    float* db = (float*)((char*)da + align256(n * 4));
    HIPCHK(hipMemcpyAsync(dbo, bo, n * sizeof(r2d_body_options), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dkind, kind, n * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(da, dim_a, n * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(db, dim_b, n * 4, hipMemcpyHostToDevice, st));
    const uint32_t id0 = w->current_body_id;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_make_bodies, dim3(blocks), dim3(256), 0, st, w->dblock, w->dcap, (uint32_t)n0,
                       (uint32_t)n, dbo, dkind, da, db, id0);
    HIPCHK(hipGetLastError());
    // host side: ids, the id -> slot map, the shapes the host keeps
    if ((size_t)id0 + n > w->id2slot.size())
        w->id2slot.resize(std::max<size_t>(64, ((size_t)id0 + n) * 2), -1);
    uint32_t* hid = w->u(F_ID);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t id = id0 + (uint32_t)i;
        const size_t slot = n0 + i;
        w->id2slot[id] = (int32_t)slot;
        hid[slot] = id;
        const bool rect = kind[i] == 1;
        w->width[slot] = rect ? dim_a[i] : 0.f;
        w->height[slot] = rect ? dim_b[i] : 0.f;
        if (ids_out) ids_out[i] = id;
    }
    w->n = (int32_t)need;
    w->current_body_id = id0 + (uint32_t)n;
    // the caller's arrays are read by the copies above
    HIPCHK(hipStreamSynchronize(st));
    w->host_valid = false;
    w->dev_valid = true;
    w->version += 1;
    return R2D_OK;
}

R2D_EXPORT int r2d_world_footprint(r2d_world* w, size_t* bytes)
{
    if (!w || !bytes) return R2D_E_INVALID_ARG;
    *bytes = world_footprint(w);
    return R2D_OK;
}

R2D_EXPORT int r2d_serve_create(r2d_ctx* c, const r2d_serve_options* opt, r2d_server** out)
{
    if (!c || !opt || !out || opt->max_worlds == 0 || opt->hbm_budget_bytes == 0) return R2D_E_INVALID_ARG;
    if (c->server) return R2D_E_BUSY;
    r2d_server* sv = new (std::nothrow) r2d_server();
    if (!sv) return R2D_E_OOM;
    sv->home = c;
    sv->opt = *opt;
    // lanes: one, plus one per increasing split
    sv->nlanes = 1;
    for (int l = 0; l < 2 && opt->split_bodies[l] > (l ? opt->split_bodies[l - 1] : 0u); ++l) sv->nlanes += 1;
    for (uint32_t l = 0; l < sv->nlanes; ++l) {
        int rc = r2d_ctx_create(c->device, &sv->lanes[l].ctx);
        if (rc) {
            for (uint32_t k = 0; k < l; ++k) r2d_ctx_destroy(sv->lanes[k].ctx);
            delete sv;
            return rc;
        }
        sv->lanes[l].ctx->server = sv;
    }
    c->server = sv;
    *out = sv;
    return R2D_OK;
}

R2D_EXPORT int r2d_serve_destroy(r2d_server* sv)
{
    if (!sv) return R2D_E_INVALID_ARG;
    int rc = server_collect(sv);
    for (auto& q : sv->reqs)   // unfinished requests release their worlds
        if (q.w) {
            q.w->in_service = false;
            q.w->ctx = sv->home;
            q.w->epoch = 0;
        }
    for (uint32_t l = 0; l < sv->nlanes; ++l) {
        sv->lanes[l].ctx->server = nullptr;
        r2d_ctx_destroy(sv->lanes[l].ctx);
    }
    sv->home->server = nullptr;
    delete sv;
    return rc;
}

R2D_EXPORT int r2d_serve_submit(r2d_server* sv, r2d_world* w, uint32_t steps, uint64_t* ticket)
{
    if (!sv || !w || !ticket || w->ctx != sv->home) return R2D_E_INVALID_ARG;
    if (w->in_service) return R2D_E_BUSY;
    // the reference raises InvalidRigidBodyId from the first constraint solve
    if (sv->opt.sub_steps > 0 && sv->opt.collision_iters > 0)
        for (const HJoint& h : w->joints)
            if (w->slot_of(h.id1) < 0 || (h.type <= 1 && w->slot_of(h.id2) < 0)) return R2D_E_INVALID_ID;
    size_t bytes = world_footprint(w);
    if (bytes > sv->opt.hbm_budget_bytes) { sv->rejected += 1; return R2D_E_CAPACITY; }
    r2d_server::Req q;
    q.w = w;
    q.steps = steps;
    q.bytes = bytes;
    q.lane = lane_of(sv, (uint32_t)w->n);
    q.t_submit = now_ms();
    sv->reqs.push_back(q);
    *ticket = sv->reqs.size();
    w->in_service = true;
    if (steps == 0) {
        sv->resident += bytes;   // retire() gives it back
        retire(sv, *ticket, R2D_REQ_DONE, R2D_OK);
        return R2D_OK;
    }
    sv->lanes[q.lane].queue.push_back(*ticket);
    return R2D_OK;
}

R2D_EXPORT int r2d_serve_tick(r2d_server* sv, r2d_serve_tick_info* info)
{
    if (!sv) return R2D_E_INVALID_ARG;
    sv->last_completed = 0;
    sv->last_gpu_ms = sv->last_wall_ms = 0;
    int first = R2D_OK;
    // collect the lanes whose step has finished; with none finished, wait for the first
    bool any = false, waiting = false;
    for (;;) {
        for (uint32_t l = 0; l < sv->nlanes; ++l) {
            r2d_server::Lane& L = sv->lanes[l];
            if (!L.inflight) continue;
            waiting = true;
            if (hipEventQuery(L.ctx->ev_end) == hipErrorNotReady) continue;
            int rc = lane_collect(sv, L);
            if (rc && first == R2D_OK) first = rc;
            any = true;
        }
        if (any || !waiting) break;
        std::this_thread::yield();
    }
    uint32_t admitted = server_admit(sv);
    uint32_t launched = 0;
    uint64_t bodies = 0;
    for (uint32_t l = 0; l < sv->nlanes; ++l) {
        r2d_server::Lane& L = sv->lanes[l];
        if (L.inflight) continue;
        int rc = lane_launch(sv, L);
        if (rc && first == R2D_OK) first = rc;
        if (L.inflight) { launched += (uint32_t)L.batch_tickets.size(); bodies += L.batch.N; }
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->admitted = admitted;
        info->completed = sv->last_completed;
        size_t queued = 0;
        for (uint32_t l = 0; l < sv->nlanes; ++l) queued += sv->lanes[l].queue.size();
        info->queued = (uint32_t)queued;
        info->running = (uint32_t)sv->nrunning;
        info->launched_worlds = launched;
        info->launched_bodies = bodies;
        info->resident_bytes = sv->resident;
        info->step_gpu_ms = sv->last_gpu_ms;
        info->step_wall_ms = sv->last_wall_ms;
    }
    return first;
}

R2D_EXPORT int r2d_serve_drain(r2d_server* sv)
{
    if (!sv) return R2D_E_INVALID_ARG;
    int first = R2D_OK;
    for (;;) {
        bool busy = sv->nrunning > 0;
        for (uint32_t l = 0; l < sv->nlanes; ++l)
            busy = busy || sv->lanes[l].inflight || !sv->lanes[l].queue.empty();
        if (!busy) break;
        int rc = r2d_serve_tick(sv, nullptr);
        if (rc && first == R2D_OK) first = rc;
    }
    return first;
}

R2D_EXPORT int r2d_serve_completed(r2d_server* sv, uint64_t* tickets, uint32_t cap, uint32_t* n)
{
    if (!sv || !n || (cap && !tickets)) return R2D_E_INVALID_ARG;
    uint32_t k = 0;
    while (k < cap && !sv->done_tickets.empty()) {
        tickets[k++] = sv->done_tickets.front();
        sv->done_tickets.pop_front();
    }
    *n = k;
    return R2D_OK;
}

R2D_EXPORT int r2d_serve_poll(r2d_server* sv, uint64_t ticket, r2d_request_status* out)
{
    if (!sv || !out || ticket == 0 || ticket > sv->reqs.size()) return R2D_E_INVALID_ARG;
    const r2d_server::Req& q = sv->reqs[ticket - 1];
    std::memset(out, 0, sizeof *out);
    out->state = q.state;
    out->status = q.status;
    out->steps_done = q.done;
    out->steps_requested = q.steps;
    out->footprint_bytes = q.bytes;
    out->lane = q.lane;
    double now = now_ms();
    out->wait_ms = (q.state == R2D_REQ_QUEUED ? now : q.t_admit) - q.t_submit;
    out->total_ms = (q.state >= R2D_REQ_DONE ? q.t_done : now) - q.t_submit;
    out->max_step_ms = q.max_step_ms;
    return R2D_OK;
}

R2D_EXPORT int r2d_serve_stats(r2d_server* sv, r2d_serve_summary* out)
{
    if (!sv || !out) return R2D_E_INVALID_ARG;
    std::memset(out, 0, sizeof *out);
    out->submitted = sv->reqs.size();
    out->completed = sv->completed;
    out->rejected = sv->rejected;
    size_t queued = 0;
    for (uint32_t l = 0; l < sv->nlanes; ++l) queued += sv->lanes[l].queue.size();
    out->queued = queued;
    out->running = sv->nrunning;
    out->batch_steps = sv->batch_steps;
    out->scene_steps = sv->scene_steps;
    out->body_steps = sv->body_steps;
    out->gpu_ms = sv->gpu_ms;
    out->wall_ms = sv->wall_ms;
    out->p50_step_ms = weighted_pct(sv->step_ms, sv->step_worlds, 0.50);
    out->p99_step_ms = weighted_pct(sv->step_ms, sv->step_worlds, 0.99);
    std::vector<uint32_t> ones(sv->step_ms.size(), 1u);
    out->p50_batch_ms = weighted_pct(sv->step_ms, ones, 0.50);
    out->p99_batch_ms = weighted_pct(sv->step_ms, ones, 0.99);
    out->resident_bytes = sv->resident;
    out->peak_resident_bytes = sv->peak_resident;
    out->lanes = sv->nlanes;
    return R2D_OK;
}
