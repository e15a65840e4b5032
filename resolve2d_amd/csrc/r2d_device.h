// Device-side layout of resolve2d worlds and of one batched step.
//
// A world keeps its bodies as structure-of-arrays in HBM, indexed by the
// reference's iteration order (the AutoArrayHashMap order of lib.zig:135,
// including swapRemove moves).  A batched step views several worlds as one
// flat "batch body index" space: batch index = world.body_base + slot.
#pragma once

#include <stdint.h>

#include "r2d_math.h"

namespace r2d {

enum : uint32_t {
    BF_STATIC = 1u,    // RigidBody.static (RigidBody.zig:85)
    BF_RECT = 2u,      // RigidBody.Type.rectangle
};

// per-body scheduling flag: the body orders the solver's items
enum : uint32_t { SF_RESOURCE = 1u };

// Constraint record (Constraint.zig:27-31 parameters + per-type fields)
struct Joint {
    int32_t type;      // 0 distance, 1 offset distance, 2 fixed position, 3 motor
    int32_t s1, s2;    // world-local body slots (s2 unused for 2/3)
    float pmax, pmin, beta;
    float r1x, r1y, r2x, r2y;
    float target;      // target distance / target omega
    float tx, ty;      // fixed-position target
    float pad[3];
};

// ForceGen record (ForceGen.zig:12-19); type 1 = anchored spring extension
struct Gen {
    int32_t type;
    int32_t slot;      // spring body slot, -1 when its body is gone
    float gx, gy;      // DownwardsGravity.g_vec (DownwardsGravity.zig:22)
    float k, c, ax, ay;
};

struct WorldDev {
    // bodies (SoA, capacity >= nbodies)
    float *px, *py, *ang, *mx, *my, *L, *fx, *fy, *tq;
    float *mass, *inertia, *mu, *rad, *hw, *hh;
    float *ax, *ay, *ahw, *ahh;      // stale AABB (aabb.zig:6-8)
    uint32_t *id, *flags;
    const Joint* joints;
    const Gen* gens;
    const uint64_t* excl;            // sorted (min_id << 32 | max_id)
    int32_t nbodies, njoints, ngens, nexcl;
    float cell;                      // spatial-hash cell width
    uint32_t table;                  // spatial-hash table size (buckets)
    uint32_t body_base, bucket_base, joint_base;
    int32_t sched;                   // 0 reference order, 1 coloured
    // warm starting (opt-in; warm == 0: the reference's cold manifolds)
    float warm;                      // fraction of last step's impulses carried over, (0, 1]
    uint4* wtab;                     // [WARM_K x 3 per body] last step's manifolds under their lower slot
    uint32_t* wcnt;                  // [per body] entries of wtab in use
    float4* wsum;                    // [per body] warm impulse per sub-step: momentum x, y, angular
};

// Warm starting keeps, after each step, every manifold's ids, point mask,
// normal and accumulated impulses under the lower slot of its two bodies (at
// most WARM_K per body, in discovery order).  A manifold of the next step
// with the same (ref id, inc id), the same points and a normal within
// WARM_DOT of the old one starts each sub-step with warm * acc / sub_steps.
constexpr uint32_t WARM_K = 8;
constexpr float WARM_DOT = 0.99f;
// m_valid bit: the manifold matched last step's (warm starting)
enum : uint32_t { MV_WARM = 4u };

// device counters
enum : int {
    C_ENTRIES = 0,     // spatial-hash cell entries E
    C_RAW = 1,         // raw candidates (a, b, cell) R
    C_MANIFOLDS = 2,   // manifolds M
    C_ITEMS = 3,       // solver items (joints + manifolds)
    C_OVERFLOW = 4,    // capacity overflow flag (bit per buffer)
    C_STEPS = 5,       // committed steps
    C_COUNT = 8,
    C_SUMS = 16        // k_fused: 64-bit sums over the launch's world-steps of bodies, E, R, M, items
};

// A world's slices of the step scratch when whole steps of the world run in
// one workgroup (k_fused): cell entries [e0, e0 + ecap] (bkt_items,
// bkt_sorted, ent_body, ent_bkt, ent_cnt), raw candidates [r0, r0 + rcap]
// (cand, cand_flag, sat_*), manifolds [m0, m0 + mcap), solver items
// [i0, i0 + icap) (it_*, recs) and their lists [2 i0, 2 (i0 + icap)).
struct WorldScratch {
    uint32_t e0, ecap, r0, rcap, m0, mcap, i0, icap;
};

// world status codes (w_status)
enum : int32_t {
    WS_OK = 0,
    WS_CELLS = 3,        // cell range too large to enumerate
    WS_NONFINITE = 4,    // non-finite AABB
    WS_STALL = 5,        // solver dataflow stalled (corrupt ordering data)
    WS_OVERFLOW = 6      // k_fused: the world's scratch slice is too small (w_cnt[4..7] = counts needed)
};

// per-step scratch, sized by capacities chosen on the host
struct StepDev {
    uint32_t nb;             // batch bodies N
    uint32_t nworlds;        // W
    uint32_t nbuckets;       // B
    uint32_t njoints;        // batch joints
    uint32_t cap_e, cap_r, cap_m, cap_i;
    const WorldDev* worlds;
    uint32_t* bworld;        // [N] world index of each batch body
    // broadphase, per body
    int64_t *cx0, *cy0;      // first cell
    uint32_t *cnx, *cny;     // cells per row / rows
    uint32_t* ncell;         // [N+1] cells per body -> exclusive offsets
    uint32_t* cell_off;      // [N+1]
    float *cc, *cs;          // cos / sin of the angle at step start
    uint32_t* sflag;         // scheduling flags
    // spatial hash buckets
    uint32_t *bkt_cnt, *bkt_off, *bkt_fill;  // [B+1]
    uint32_t* bkt_items;     // [cap_e] bucket members as filled (any order)
    uint32_t* bkt_sorted;    // [cap_e] bucket members in descending slot
    // entries -> raw candidates
    uint32_t *ent_body, *ent_bkt;  // [cap_e]
    uint32_t *ent_cnt, *ent_off;   // [cap_e+1]
    uint4* cand;             // [cap_r] raw candidates: lower body, partner, cell index in the lower body, 0
    uint32_t *cand_flag, *cand_idx;      // [cap_r+1]
    float *sat_nx, *sat_ny;
    uint32_t* sat_info;      // normal id | ref-is-lower-slot << 2 | collides << 3
    // manifolds (discovery order)
    uint32_t *m_ref, *m_inc;           // batch body indices
    float *m_nx, *m_ny, *m_fric, *m_prev1, *m_prev2;
    uint32_t* m_valid;                 // bit p: point p present
    float* m_pt;                       // [MP_NFIELDS fields][2 points][cap_m]
    // per world
    uint32_t *w_man_begin, *w_item_base;   // [W+1]
    int32_t* w_status;
    uint32_t* w_order;       // [W] worlds in the order the solve workgroups take them
    // solver items (joints then manifolds, per world)
    int32_t *it_r0, *it_r1;
    uint32_t *it_k0, *it_k1;
    uint32_t *deg, *deg_off, *deg_fill;    // [N+1]
    uint32_t* lists;                       // [2*cap_i]
    uint32_t* it_color;                    // [cap_i] colour (0 in reference order)
    uint32_t* it_perm;                     // [cap_i] world-local items in execution order
    uint64_t* cmask;                       // [2*N] colours used per body (coloured schedule)
    uint4* recs;                           // [8*cap_i] solver item records (128 B each)
    uint4* rmass;                          // [cap_i] worlds beyond LDS: inverse mass and inertia of each item's
                                           // two bodies (im a, ii a, im b, ii b) for the light compact chain
    // global-memory solver state (worlds too large for LDS)
    float *g_im, *g_ii;
    uint32_t* g_cnt;
    uint32_t* cmap;          // [N] worlds beyond LDS: compact chain index of each body (or ~0)
    uint32_t* cinv;          // [N] compact index -> world slot, at the world's body base
    uint32_t* w_nc;          // [W] bodies in the compact chain
    uint32_t* counters;
    uint32_t* partials;                    // scan scratch
    // whole steps of each world in one workgroup (k_fused)
    WorldScratch* wscr;        // [W] the world's slices of the arrays above (a world may move to a larger one)
    uint32_t* fpool;           // [8] next free index and capacity of the entry, candidate, manifold and item arrays
    uint32_t* w_done;          // [W] steps the world has completed in the current call
    uint32_t* w_cnt;           // [8W] E, R, M, items of the world's last step; then the counts an overflow needed
    uint32_t* w_man_end;       // [W] end of the world's manifolds (from w_man_begin)
    uint64_t* w_tstamp;        // [W (tcap + 1)] wall clock at the call's first step and after each step, or null
    uint32_t tcap;
    uint64_t* w_prof;          // [W FP_N] wall-clock ticks per phase of the fused kernel, or null
};

// phases of a fused step (profiled with w_prof)
enum : int { FP_PREP = 0, FP_BUCKETS, FP_CANDS, FP_NARROW, FP_CONTACTS, FP_SCHED, FP_SOLVE, FP_N = 8 };

// manifold point fields: m_pt[(field * 2 + point) * cap_m + m]
enum : int {
    MP_ACC_PN = 0, MP_ACC_PT, MP_R1X, MP_R1Y, MP_R2X, MP_R2Y, MP_POSX, MP_POSY,
    MP_DEPTH, MP_ODEPTH, MP_MASS_N, MP_MASS_T,
    MP_WARM_N, MP_WARM_T,      // warm starting: impulse per sub-step carried from the previous step
    MP_NFIELDS
};

}  // namespace r2d
