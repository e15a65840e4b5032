// Host-side entry points of the step kernels (r2d_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "r2d_device.h"

namespace r2d {

// phase markers recorded by the optional per-step timer
enum { KT_BROAD_BEGIN = 0, KT_NARROW_BEGIN, KT_SCHED_BEGIN, KT_SOLVE_BEGIN, KT_END, KT_N };


// bytes of LDS the solve kernel needs for a world of n bodies
size_t solve_lds_bytes(int nbodies);

// the four phases of one step, enqueued in this order on one stream
enum { PH_BROAD = 0, PH_NARROW, PH_SCHED, PH_SOLVE, PH_N };

// solve-kernel variants (register-slot kernel shape + record-kernel workgroup
// size; table in r2d_kernels.hip)
constexpr int SOLVE_VARIANTS = 15;
// the default for batches of at most WIDE_MAX_WORLDS worlds (each alone on a
// CU): records only, 768-thread workgroups, so a world of up to 768 solver
// items holds one item per lane and its chain never picks up a record
constexpr int WIDE_VARIANT = 14;
constexpr size_t WIDE_MAX_WORLDS = 64;
// the variant every batch with a warm-started world takes (not tunable)
constexpr int WARM_VARIANT = SOLVE_VARIANTS;
constexpr int SOLVE_TABLE = SOLVE_VARIANTS + 1;
unsigned solve_variant_threads(int v);
int solve_variant_slots(int v);

// worlds of at most this many bodies may take the register-slot kernel
constexpr uint32_t SLOT_NMAX = 1024;

// how one batch's solve phase is split over the solve kernels
struct SolvePlan {
    int variant = 0;
    uint32_t slot_cap = 0;    // items a world may have to take the slot kernel (0: no slot kernel)
    uint32_t slot_nmax = 0;   // bodies a world may have to take the slot kernel
    uint32_t lds_nmax = 0;    // bodies a world may have to keep its body state in LDS
    uint32_t slot_maxn = 0;   // largest world each launch may take: sizes its LDS
    uint32_t lds_maxn = 0;
    uint32_t any_slot = 0, any_lds = 0, any_big = 0;
    uint32_t any_colored = 0;  // some world uses the coloured schedule
    uint32_t any_warm = 0;     // some world is warm started
    uint32_t n_big = 0;        // worlds beyond LDS
    uint32_t n_lds = 0;        // worlds of the record kernel with LDS body state
    uint32_t rec_wide = 0;     // they take the 512-lane record kernel (few worlds, each alone on a CU)
};
// lds_limit: largest world (bodies) that keeps its body state in LDS (0: as
// many as LDS holds); larger worlds use the global-memory record kernel
SolvePlan plan_solve(int variant, const uint32_t* nbodies, size_t nw, uint32_t lds_limit, bool any_colored,
                     bool any_warm, bool allow_wide = true);

// Optional per-kernel timing (a context created with R2D_TRACE_KERNELS set):
// an event after every launch, summed per kernel name; the steps run eagerly.
struct KernelTrace {
    struct Mark {
        const char* name;
        hipEvent_t ev;
        hipStream_t st;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, uint64_t>> totals;
    hipStream_t main = nullptr;
    void mark(const char* name, hipStream_t st);
    void flush();                 // synchronise and add the recorded intervals
    void report(FILE* f);
    ~KernelTrace();
};

// the context's main stream plus two side streams for the solve kernels
struct Streams {
    hipStream_t main;
    hipStream_t aux[2];
    hipEvent_t fork;
    hipEvent_t join[2];
    KernelTrace* trace;
};

// enqueue one phase of a batched step (phases in PH_* order on ss.main)
void launch_phase(const StepDev& S, const Streams& ss, int phase, float dt, uint32_t sub_steps, uint32_t iters,
                  const SolvePlan& sp);

void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_const, uint32_t cap,
              uint32_t* total_dev, uint32_t* partials, hipStream_t st);

void set_solve_smem_limit();

// Whole steps per workgroup (k_fused) for worlds of at most FUSED_NMAX bodies
// in the reference order; the variant picks the solve's slot shape.
constexpr uint32_t FUSED_NMAX = 1024;
bool fused_available(int variant);
// dynamic LDS of the fused kernel for a world of n bodies and tb hash buckets
// (slots: the variant solves with register slots, else with records;
// threads: its workgroup size, which sizes the record chain's lane slots)
size_t fused_lds_bytes(uint32_t n, uint32_t tb, bool slots, uint32_t threads);
bool fused_slots(int variant);
uint32_t fused_threads(int variant);
// one workgroup per world (S.w_order), up to nsteps steps per world (S.w_done)
// the fused kernel of a variant as compiled (registers, LDS, scratch): for error reports
hipError_t fused_attributes(int variant, hipFuncAttributes* out);
void launch_fused(const StepDev& S, int variant, uint32_t nworlds, float dt, uint32_t sub_steps, uint32_t iters,
                  uint32_t nsteps, size_t lds, hipStream_t st);

}  // namespace r2d
