// Host-side entry points of the step kernels (r2d_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "r2d_device.h"

namespace r2d {

// phase markers recorded by the optional per-step timer
enum { KT_BROAD_BEGIN = 0, KT_NARROW_BEGIN, KT_SCHED_BEGIN, KT_SOLVE_BEGIN, KT_END, KT_N };

struct KernelTimer {
    hipEvent_t ev[KT_N];
    void mark(int id, hipStream_t st) { (void)hipEventRecord(ev[id], st); }
};

// bytes of LDS the solve kernel needs for a world of n bodies
size_t solve_lds_bytes(int nbodies);

// the four phases of one step, enqueued in this order on one stream
enum { PH_BROAD = 0, PH_NARROW, PH_SCHED, PH_SOLVE, PH_N };

// solve-kernel variants (workgroup size x waves per SIMD; table in r2d_kernels.hip)
constexpr int SOLVE_VARIANTS = 7;
unsigned solve_variant_threads(int v);
int solve_variant_slots(int v);

// enqueue one phase of a batched step on stream st
void launch_phase(const StepDev& S, hipStream_t st, int phase, float dt, uint32_t sub_steps, uint32_t iters,
                  uint32_t max_bodies, bool lds, int solve_variant);

void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_dev, uint32_t n_const, uint32_t cap,
              uint32_t* total_dev, uint32_t* partials, hipStream_t st);

void set_solve_smem_limit();

}  // namespace r2d
