/*
 * ORACLE (test infrastructure only).  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg load this library, and only as the checker
 * or the timed CPU baseline.  The product path never links it.
 *
 * Instantiates the restatement in r2d_oracle_impl.h twice:
 *   or32_*  -- f32, operation-for-operation the reference's arithmetic
 *   or64_*  -- f64, the same method in double precision (error reference)
 * and exposes a flat C API (doubles in, doubles out; f32 values round-trip
 * exactly through double) for ctypes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "r2d_trig_ref.h"

#define R2D_REAL float
#define R2D_P o32i_
#define R2D_SQRT sqrtf
#define R2D_FABS fabsf
#define R2D_FLOOR floorf
#define R2D_SIN r2dref_sinf
#define R2D_COS r2dref_cosf
#define R2D_PI ((float)3.14159265358979323846)
#include "r2d_oracle_impl.h"
#undef R2D_REAL
#undef R2D_P
#undef R2D_SQRT
#undef R2D_FABS
#undef R2D_FLOOR
#undef R2D_SIN
#undef R2D_COS
#undef R2D_PI

#define R2D_REAL double
#define R2D_P o64i_
#define R2D_SQRT sqrt
#define R2D_FABS fabs
#define R2D_FLOOR floor
#define R2D_SIN sin
#define R2D_COS cos
#define R2D_PI 3.14159265358979323846
#include "r2d_oracle_impl.h"

#define EXPORT __attribute__((visibility("default")))

EXPORT float or_sinf(float x) { return r2dref_sinf(x); }

/* f32 nmath probes (nmath.zig): op 0 add2, 1 sub2, 2 multelem2, 3 divelem2,
 * 4 dot2, 5 cross2, 6 length2sq, 7 length2, 8 normalize2, 9 rotate2 (angle = bx),
 * 10 rotate90clockwise, 11 rotate90counterclockwise.  out[0..1]. */
EXPORT void or32_nmath(int op, double ax, double ay, double bx, double by, double *out)
{
    o32i_vec2 a = o32i_v2((float)ax, (float)ay), b = o32i_v2((float)bx, (float)by), r = o32i_v2(0, 0);
    float s = 0;
    switch (op) {
    case 0: r = o32i_add2(a, b); break;
    case 1: r = o32i_sub2(a, b); break;
    case 2: r = o32i_v2(a.x * b.x, a.y * b.y); break;
    case 3: r = o32i_v2(a.x / b.x, a.y / b.y); break;
    case 4: s = o32i_dot2(a, b); break;
    case 5: s = o32i_cross2(a, b); break;
    case 6: s = o32i_length2sq(a); break;
    case 7: s = o32i_length2(a); break;
    case 8: r = o32i_normalize2(a); break;
    case 9: r = o32i_rotate2(a, (float)bx); break;
    case 10: r = o32i_rot90cw(a); break;
    default: r = o32i_rot90ccw(a); break;
    }
    if (op >= 4 && op <= 7) { out[0] = s; out[1] = 0; } else { out[0] = r.x; out[1] = r.y; }
}

/* AABB.intersects (aabb.zig:12-19) on two (x, y, half_width, half_height) boxes */
EXPORT int or32_aabb_intersects(const double *a, const double *b)
{
    o32i_body A, B;
    memset(&A, 0, sizeof A); memset(&B, 0, sizeof B);
    A.apos = o32i_v2((float)a[0], (float)a[1]); A.ahw = (float)a[2]; A.ahh = (float)a[3];
    B.apos = o32i_v2((float)b[0], (float)b[1]); B.ahw = (float)b[2]; B.ahh = (float)b[3];
    return o32i_aabb_intersects(&A, &B);
}

/* SpatialHash.hash (SpatialHash.zig:78-81) */
EXPORT uint64_t or_hash(uint64_t table, int64_t xi, int64_t yi) { return o32i_hash(table, xi, yi); }
EXPORT float or_cosf(float x) { return r2dref_cosf(x); }

#define API(EP, IP, R)                                                                                       \
    EXPORT void *EP##world_new(double cell_width, uint64_t table_mult)                                   \
    {                                                                                                   \
        return IP##world_new((R)cell_width, table_mult);                                                 \
    }                                                                                                   \
    EXPORT void EP##world_free_(void *w) { IP##world_free((IP##world *)w); }                               \
    EXPORT void EP##set_hash(void *w, double cell, uint64_t mult)                                        \
    {                                                                                                   \
        ((IP##world *)w)->hash_cell = (R)cell;                                                           \
        ((IP##world *)w)->hash_mult = mult;                                                              \
    }                                                                                                   \
    EXPORT void EP##clear(void *vw)                                                                      \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        for (uint32_t i = 0; i < w->capid; ++i) w->id2slot[i] = -1;                                     \
        w->nbodies = 0; w->ncons = 0; w->ngens = 0; w->nmans = 0; w->nexcl = 0;                         \
        if (w->wcnt) memset(w->wcnt, 0, sizeof(int32_t) * (size_t)w->wcap);                             \
    }                                                                                                   \
    EXPORT uint32_t EP##make_disc_(void *w, double px, double py, double vx, double vy, double ang,     \
                                  double om, double mu, int mk, double mv, double r)                    \
    {                                                                                                   \
        return IP##make_disc((IP##world *)w, (R)px, (R)py, (R)vx, (R)vy, (R)ang, (R)om, (R)mu, mk, (R)mv, \
                            (R)r);                                                                      \
    }                                                                                                   \
    EXPORT uint32_t EP##make_rect_(void *w, double px, double py, double vx, double vy, double ang,     \
                                  double om, double mu, int mk, double mv, double wd, double ht)        \
    {                                                                                                   \
        return IP##make_rect((IP##world *)w, (R)px, (R)py, (R)vx, (R)vy, (R)ang, (R)om, (R)mu, mk, (R)mv, \
                            (R)wd, (R)ht);                                                              \
    }                                                                                                   \
    EXPORT int EP##set_static(void *vw, uint32_t id, int st)                                             \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        int32_t s = IP##slot_of(w, id);                                                                  \
        if (s < 0) return -1;                                                                           \
        w->bodies[s].is_static = st;                                                                    \
        return 0;                                                                                       \
    }                                                                                                   \
    EXPORT int EP##remove_body_(void *w, uint32_t id) { return IP##remove_body((IP##world *)w, id); }      \
    EXPORT void EP##make_gravity(void *w, double g)                                                      \
    {                                                                                                   \
        IP##forcegen f;                                                                                  \
        memset(&f, 0, sizeof f);                                                                        \
        f.type = 0; f.g = (R)g; f.g_vec = IP##v2(0, -(R)g);                                              \
        IP##add_gen((IP##world *)w, &f);                                                                  \
    }                                                                                                   \
    EXPORT void EP##make_spring(void *w, uint32_t id, double k, double c, double ax, double ay)          \
    {                                                                                                   \
        IP##forcegen f;                                                                                  \
        memset(&f, 0, sizeof f);                                                                        \
        f.type = 1; f.id = id; f.k = (R)k; f.c = (R)c; f.anchor = IP##v2((R)ax, (R)ay);                  \
        IP##add_gen((IP##world *)w, &f);                                                                  \
    }                                                                                                   \
    EXPORT int64_t EP##make_joint(void *vw, int type, double pmax, double pmin, double beta, uint32_t id1, \
                                 uint32_t id2, double r1x, double r1y, double r2x, double r2y,          \
                                 double target, double tx, double ty)                                   \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        IP##constraint c;                                                                                \
        memset(&c, 0, sizeof c);                                                                        \
        c.type = type; c.pmax = (R)pmax; c.pmin = (R)pmin; c.beta = (R)beta;                            \
        c.id1 = id1; c.id2 = id2; c.r1 = IP##v2((R)r1x, (R)r1y); c.r2 = IP##v2((R)r2x, (R)r2y);           \
        c.target = (R)target; c.tpos = IP##v2((R)tx, (R)ty);                                             \
        IP##add_constraint(w, &c);                                                                       \
        return w->ncons - 1;                                                                            \
    }                                                                                                   \
    EXPORT void EP##exclude_(void *w, uint32_t a, uint32_t b) { IP##exclude((IP##world *)w, a, b); }       \
    EXPORT int EP##process_(void *w, double dt, uint64_t s, uint64_t k)                                  \
    {                                                                                                   \
        return IP##process((IP##world *)w, (R)dt, s, k);                                                  \
    }                                                                                                   \
    EXPORT int EP##num_bodies(void *w) { return ((IP##world *)w)->nbodies; }                              \
    EXPORT int EP##num_manifolds(void *w) { return ((IP##world *)w)->nmans; }                             \
    EXPORT uint32_t EP##current_id(void *w) { return ((IP##world *)w)->current_body_id; }                 \
    /* 19 doubles per body, iteration order */                                                          \
    EXPORT void EP##get_bodies(void *vw, double *o)                                                      \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        for (int32_t i = 0; i < w->nbodies; ++i) {                                                      \
            IP##body *b = &w->bodies[i];                                                                 \
            double *r = o + 19 * i;                                                                     \
            r[0] = b->id; r[1] = b->type; r[2] = b->is_static;                                          \
            r[3] = b->pos.x; r[4] = b->pos.y; r[5] = b->mom.x; r[6] = b->mom.y;                         \
            r[7] = b->force.x; r[8] = b->force.y; r[9] = b->mass; r[10] = b->angle; r[11] = b->L;       \
            r[12] = b->torque; r[13] = b->inertia; r[14] = b->mu;                                       \
            r[15] = b->apos.x; r[16] = b->apos.y; r[17] = b->ahw; r[18] = b->ahh;                       \
        }                                                                                               \
    }                                                                                                   \
    /* setters (wasm_root.zig:156-224) for one body by id */                                            \
    EXPORT int EP##set_body(void *vw, uint32_t id, double px, double py, double mx, double my, double fx, \
                           double fy, double ang, double L, double tq)                                  \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        int32_t s = IP##slot_of(w, id);                                                                  \
        if (s < 0) return -1;                                                                           \
        IP##body *b = &w->bodies[s];                                                                     \
        b->pos = IP##v2((R)px, (R)py); b->mom = IP##v2((R)mx, (R)my); b->force = IP##v2((R)fx, (R)fy);     \
        b->angle = (R)ang; b->L = (R)L; b->torque = (R)tq;                                              \
        return 0;                                                                                       \
    }                                                                                                   \
    /* 33 doubles per manifold, discovery order */                                                      \
    EXPORT void EP##get_manifolds(void *vw, double *o)                                                   \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        for (int32_t i = 0; i < w->nmans; ++i) {                                                        \
            IP##manifold *m = &w->mans[i];                                                               \
            double *r = o + 33 * i;                                                                     \
            r[0] = w->bodies[m->ref].id; r[1] = w->bodies[m->inc].id;                                   \
            r[2] = m->normal.x; r[3] = m->normal.y; r[4] = m->tangent.x; r[5] = m->tangent.y;           \
            r[6] = m->friction;                                                                         \
            for (int p = 0; p < 2; ++p) {                                                               \
                IP##point *q = &m->pts[p];                                                               \
                double *z = r + 7 + 13 * p;                                                             \
                z[0] = q->valid;                                                                        \
                if (!q->valid) { for (int j = 1; j < 13; ++j) z[j] = 0; continue; }                     \
                z[1] = q->acc_pn; z[2] = q->acc_pt; z[3] = q->ref_r.x; z[4] = q->ref_r.y;               \
                z[5] = q->inc_r.x; z[6] = q->inc_r.y; z[7] = q->pos.x; z[8] = q->pos.y;                 \
                z[9] = q->depth; z[10] = q->orig_depth; z[11] = q->mass_n; z[12] = q->mass_t;           \
            }                                                                                           \
        }                                                                                               \
    }                                                                                                   \
    EXPORT void EP##set_trace(void *w, int on) { ((IP##world *)w)->trace_on = on; }                       \
    EXPORT void EP##set_schedule(void *w, int mode) { ((IP##world *)w)->sched = mode; }                   \
    /* warm starting (extension): factor, kept manifolds emptied */                                     \
    EXPORT void EP##set_warm(void *vw, double f)                                                         \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        w->warm = (R)f;                                                                                 \
        if (w->wcnt) memset(w->wcnt, 0, sizeof(int32_t) * (size_t)w->wcap);                             \
    }                                                                                                   \
    EXPORT int EP##num_warm(void *w) { return ((IP##world *)w)->nwarm; }                                   \
    /* tests: a new factor with the kept manifolds left in place */                                     \
    EXPORT void EP##set_warm_factor(void *w, double f) { ((IP##world *)w)->warm = (R)f; }                 \
    EXPORT int64_t EP##num_trace(void *w) { return ((IP##world *)w)->ntrace / 2; }                        \
    /* (id_a, id_b) pairs that reached SAT, in visiting order */                                        \
    EXPORT void EP##get_trace(void *vw, uint32_t *o)                                                     \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        for (int64_t i = 0; i < w->ntrace; ++i) o[i] = w->bodies[w->trace[i]].id;                       \
    }                                                                                                   \
    /* narrowphase of one pair as updateManifolds would build it (lib.zig:284-292):                     \
     * out[0] collides, then the 33-double manifold record */                                           \
    EXPORT int EP##narrow_pair(void *vw, uint32_t id1, uint32_t id2, double *o)                          \
    {                                                                                                   \
        IP##world *w = (IP##world *)vw;                                                                   \
        int32_t a = IP##slot_of(w, id1), b = IP##slot_of(w, id2);                                         \
        if (a < 0 || b < 0) return -1;                                                                  \
        IP##sat sat = IP##narrow_sat(w->bodies, a, b);                                                    \
        o[0] = sat.collides; o[1] = sat.penetration; o[2] = sat.normal_id;                              \
        if (!sat.collides) return 0;                                                                    \
        IP##manifold m;                                                                                  \
        memset(&m, 0, sizeof m);                                                                        \
        m.ref = sat.key_ref; m.inc = sat.key_inc; m.normal = sat.normal;                                \
        IP##identify_points(&w->bodies[m.ref], &w->bodies[m.inc], sat.normal_id, m.pts);                 \
        int32_t save = w->nmans;                                                                        \
        IP##manifold *keep = w->mans; int32_t keepcap = w->capmans;                                      \
        w->mans = &m; w->nmans = 1;                                                                     \
        EP##get_manifolds(w, o + 3);                                                                     \
        w->mans = keep; w->nmans = save; w->capmans = keepcap;                                          \
        return 1;                                                                                       \
    }

API(or32_, o32i_, float)
API(or64_, o64i_, double)
