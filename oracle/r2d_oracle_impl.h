/*
 * ORACLE (test infrastructure only; never linked into the product path).
 *
 * A plain-C restatement of resolve2d's step, `Solver.process`
 * (src/core/lib.zig:189-299), and everything it calls.  This file is included
 * twice by r2d_oracle.c: once with R2D_REAL = float (the reference's own f32
 * arithmetic, operation by operation) and once with R2D_REAL = double (the
 * "FP64 run of the reference method" used to report error growth).
 *
 * Conventions kept from the reference, with the file:line each follows:
 *   - operation order of every f32 expression is the Zig source order;
 *     the library is compiled with -ffp-contract=off (no FMA contraction);
 *   - @max/@min lower to LLVM maxnum/minnum: NaN-suppressing, and for equal
 *     operands (incl. +0/-0) the FIRST operand is returned (x86 MAXSS/MINSS
 *     lowering and compiler_rt fmaxf agree on this);
 *   - std.math.clamp(v, lo, hi) = @max(lo, @min(v, hi));
 *   - sin/cos: the musl/msun method in r2d_trig_ref.h (f32); libm in f64.
 *
 * Every function cites the reference file:line it restates.
 *
 * R2D_PLANT (default 0) compiles a deliberately wrong variant of one formula.
 * It exists only so tests/test_oracle_spring_offset.py can show that its
 * checks fail when the restatement is wrong; the library built by the
 * Makefile always has R2D_PLANT = 0.
 *   1  OffsetDistanceJoint: body 1's angular impulse with the wrong sign
 *   2  OffsetDistanceJoint: lever arms not rotated into the world frame
 *   3  OffsetDistanceJoint: anchors placed without the body's rotation
 *   4  anchored spring: damping term dropped
 *   5  anchored spring: damping on momentum instead of velocity
 */
#ifndef R2D_PLANT
#define R2D_PLANT 0
#endif

#ifndef R2D_REAL
#error "define R2D_REAL and R2D_P before including"
#endif

#define P_(a, b) a##b
#define P2(a, b) P_(a, b)
#define F(name) P2(R2D_P, name)

typedef R2D_REAL F(real);

/* ---------------------------------------------------------------- nmath.zig */
typedef struct { R2D_REAL x, y; } F(vec2);

static inline F(vec2) F(v2)(R2D_REAL x, R2D_REAL y) { F(vec2) r; r.x = x; r.y = y; return r; }
/* nmath.zig:58-68 */
static inline F(vec2) F(add2)(F(vec2) a, F(vec2) b) { return F(v2)(a.x + b.x, a.y + b.y); }
static inline F(vec2) F(sub2)(F(vec2) a, F(vec2) b) { return F(v2)(a.x - b.x, a.y - b.y); }
static inline F(vec2) F(scale2)(F(vec2) a, R2D_REAL s) { return F(v2)(a.x * s, a.y * s); }
/* nmath.zig:79-93 */
static inline R2D_REAL F(dot2)(F(vec2) a, F(vec2) b) { return a.x * b.x + a.y * b.y; }
static inline R2D_REAL F(cross2)(F(vec2) a, F(vec2) b) { return a.x * b.y - a.y * b.x; }
static inline R2D_REAL F(length2sq)(F(vec2) a) { return F(dot2)(a, a); }
static inline R2D_REAL F(length2)(F(vec2) a) { return R2D_SQRT(F(length2sq)(a)); }
/* nmath.zig:95-103: below 1e-3 returns zero, else scale by the reciprocal */
static inline F(vec2) F(normalize2)(F(vec2) a)
{
    R2D_REAL len = F(length2)(a);
    if (len < (R2D_REAL)1e-3f) return F(v2)(0, 0);
    return F(scale2)(a, (R2D_REAL)1 / len);
}
/* nmath.zig:105-115 */
static inline F(vec2) F(negate2)(F(vec2) a) { return F(v2)(-a.x, -a.y); }
static inline F(vec2) F(addmult2)(F(vec2) a, F(vec2) b, R2D_REAL s) { return F(add2)(a, F(scale2)(b, s)); }
/* Vector2.negate (nmath.zig:48-51): multiply by -1 */
static inline F(vec2) F(negself)(F(vec2) a) { return F(v2)(a.x * (R2D_REAL)-1, a.y * (R2D_REAL)-1); }
/* nmath.zig:5-7, 121-123 */
static inline int F(approx)(R2D_REAL a, R2D_REAL b, R2D_REAL eps) { return a > b - eps && a < b + eps; }
static inline int F(approx2)(F(vec2) a, F(vec2) b, R2D_REAL eps) { return F(approx)(a.x, b.x, eps) && F(approx)(a.y, b.y, eps); }
/* nmath.zig:129-136 */
static inline F(vec2) F(rotate2)(F(vec2) a, R2D_REAL angle)
{
    R2D_REAL c = R2D_COS(angle);
    R2D_REAL s = R2D_SIN(angle);
    return F(v2)(a.x * c - a.y * s, a.x * s + a.y * c);
}
/* nmath.zig:142-148 */
static inline F(vec2) F(rot90cw)(F(vec2) a) { return F(v2)(a.y, -a.x); }
static inline F(vec2) F(rot90ccw)(F(vec2) a) { return F(v2)(-a.y, a.x); }

/* @max/@min (LLVM maxnum/minnum lowering: equal -> first operand) */
static inline int F(isnan)(R2D_REAL a) { return a != a; }
static inline R2D_REAL F(maxn)(R2D_REAL a, R2D_REAL b)
{
    if (F(isnan)(a)) return b;
    if (F(isnan)(b)) return a;
    return (b > a) ? b : a;
}
static inline R2D_REAL F(minn)(R2D_REAL a, R2D_REAL b)
{
    if (F(isnan)(a)) return b;
    if (F(isnan)(b)) return a;
    return (b < a) ? b : a;
}
static inline R2D_REAL F(clampn)(R2D_REAL v, R2D_REAL lo, R2D_REAL hi) { return F(maxn)(lo, F(minn)(v, hi)); }
static inline R2D_REAL F(absr)(R2D_REAL a) { return R2D_FABS(a); }

/* ------------------------------------------------- simulation_constants.zig */
#define K_MIN_MANIFOLD_IMPULSE ((R2D_REAL)1e-4f)
#define K_BAUMGARTE ((R2D_REAL)0.02f)
#define K_BAUMGARTE_SLOP ((R2D_REAL)0.005f)
#define K_SAT_OVERLAP_THRESHOLD ((R2D_REAL)1e-4f)
#define K_COLLISION_MARGIN ((R2D_REAL)0.01f)
#define K_AABB_EPS_OVERLAP ((R2D_REAL)0.01f)
#define K_GRAD_DIV_LIMIT ((R2D_REAL)1e-4f)
#define K_ALLOWED_CONSTRAINT_VALUE ((R2D_REAL)1e-6f)

/* ------------------------------------------------------------------ types */
/* Bodies/RigidBody.zig:69-91 (Props + id/aabb/static/type), Disc.zig, Rectangle.zig */
typedef struct {
    uint32_t id;
    int32_t type;            /* 0 disc, 1 rectangle (RigidBody.Type, RigidBody.zig:28-31) */
    int32_t is_static;
    R2D_REAL radius;         /* disc */
    R2D_REAL width, height;  /* rectangle */
    F(vec2) lv[4];           /* Rectangle.local_vertices (Rectangle.zig:38-40) */
    /* props */
    F(vec2) pos, mom, force;
    R2D_REAL mass, angle, L, torque, inertia, mu;
    /* aabb (aabb.zig:6-8) */
    F(vec2) apos;
    R2D_REAL ahw, ahh;
} F(body);

/* Constraints/Constraint.zig:27-31 + per-type fields */
typedef struct {
    int32_t type;            /* 0 distance, 1 offset distance, 2 fixed position, 3 motor */
    R2D_REAL pmax, pmin, beta;
    uint32_t id1, id2;
    F(vec2) r1, r2;          /* offset joint anchors (local) */
    R2D_REAL target;         /* distance / target omega */
    F(vec2) tpos;            /* fixed position target */
} F(constraint);

/* Forces/ForceGen.zig:12-19; type 1 is the anchored spring extension */
typedef struct {
    int32_t type;            /* 0 downwards gravity, 1 anchored spring (extension) */
    R2D_REAL g;
    F(vec2) g_vec;
    uint32_t id;
    R2D_REAL k, c;
    F(vec2) anchor;
} F(forcegen);

/* collision.zig:22-52 */
typedef struct {
    int32_t valid;
    R2D_REAL acc_pn, acc_pt;
    F(vec2) ref_r, inc_r, pos;
    R2D_REAL depth, orig_depth, mass_n, mass_t;
} F(point);

/* collision.zig:54-71 */
typedef struct {
    int32_t ref, inc;        /* body slots (the reference keys by pointer) */
    F(vec2) normal, tangent;
    F(point) pts[2];
    R2D_REAL prev_angle_1, prev_angle_2;
    F(vec2) app_lin1, app_lin2;
    R2D_REAL app_rot1, app_rot2;
    R2D_REAL friction;
    /* warm starting (an extension of this build): matched last step's manifold,
     * impulse per sub-step carried per point (pn0, pt0, pn1, pt1) */
    int32_t warm;
    R2D_REAL wg[4];
} F(manifold);

/* warm starting: a kept manifold (body ids, point mask, normal, impulses) */
#define WARM_K_ 8
typedef struct {
    uint32_t ida, idb;
    int32_t mask;
    F(vec2) n;
    R2D_REAL acc[4];
} F(wentry);

typedef struct {
    uint32_t current_body_id;
    F(body) *bodies; int32_t nbodies, capbodies;
    int32_t *id2slot; uint32_t capid;          /* id -> slot, -1 removed */
    F(constraint) *cons; int32_t ncons, capcons;
    F(forcegen) *gens; int32_t ngens, capgens;
    uint64_t *excl; int32_t nexcl, capexcl;     /* (id1<<32|id2), both orders */
    F(manifold) *mans; int32_t nmans, capmans;
    R2D_REAL cell_width_arg; uint64_t table_mult_arg; /* stored, unused (lib.zig:141-142) */
    /* hash parameters actually used (lib.zig:255 hard-codes 4.0 and 2*N) */
    R2D_REAL hash_cell; uint64_t hash_mult;
    /* solver schedule: 0 reference order; 1 graph-coloured (an extension of
     * this build, not of the reference: items run colour by colour) */
    int32_t sched;
    /* warm starting (an extension of this build, off when warm == 0): kept
     * manifolds under the lower slot of their bodies, WARM_K_ per slot */
    R2D_REAL warm;
    F(wentry) *wtab; int32_t *wcnt; int32_t wcap;
    int32_t nwarm;                              /* manifolds matched by the last step */
    /* broadphase trace for tests: candidate pairs visited in discovery order */
    int32_t trace_on;
    int32_t *trace; int64_t ntrace, captrace;
} F(world);

#define GROW(ptr, n, cap, T)                                                     \
    do {                                                                         \
        if ((n) >= (cap)) {                                                      \
            int64_t nc_ = (cap) ? (int64_t)(cap) * 2 : 16;                       \
            while (nc_ <= (int64_t)(n)) nc_ *= 2;                                \
            (ptr) = (T *)realloc((ptr), sizeof(T) * (size_t)nc_);                \
            (cap) = (__typeof__(cap))nc_;                                        \
        }                                                                        \
    } while (0)

/* ------------------------------------------------------ shape vtables */
/* Disc.updateAABB (Disc.zig:63-66), Rectangle.updateAABB (Rectangle.zig:71-86) */
static void F(update_aabb)(F(body) *b)
{
    if (b->type == 0) {
        b->apos = b->pos; b->ahw = b->radius; b->ahh = b->radius;
        return;
    }
    R2D_REAL w = 0, h = 0;
    for (int i = 0; i < 4; ++i) {
        F(vec2) r = F(rotate2)(b->lv[i], b->angle);
        if (r.x > w) w = r.x;
        if (r.y > h) h = r.y;
    }
    b->apos = b->pos; b->ahw = w; b->ahh = h;
}

/* Disc.closestPoint (Disc.zig:80-84); Rectangle.closestPoint (Rectangle.zig:114-130) */
static F(vec2) F(closest_point)(const F(body) *b, F(vec2) p)
{
    if (b->type == 0) {
        F(vec2) n = F(normalize2)(F(sub2)(p, b->pos));
        return F(addmult2)(b->pos, n, b->radius);
    }
    R2D_REAL best = (R2D_REAL)INFINITY;
    F(vec2) bp = p;
    for (int i = 0; i < 4; ++i) {
        F(vec2) v = F(add2)(F(rotate2)(b->lv[i], b->angle), b->pos);  /* getWorldVertices, Rectangle.zig:88-98 */
        R2D_REAL d2 = F(length2sq)(F(sub2)(v, p));
        if (d2 < best) { best = d2; bp = v; }
    }
    return bp;
}

typedef struct { F(vec2) dir; F(vec2) a, b; int has_edge; } F(edge);

/* Disc.getNormal (Disc.zig:86-90); Rectangle.getNormal (Rectangle.zig:132-151) */
static F(edge) F(get_normal)(const F(body) *self, const F(body) *other, int iter)
{
    F(edge) e;
    if (self->type == 0) {
        F(vec2) closest = F(closest_point)(other, self->pos);
        e.dir = F(normalize2)(F(sub2)(closest, self->pos));
        e.a = e.b = F(v2)(0, 0); e.has_edge = 0;
        return e;
    }
    int nx = (iter == 3) ? 0 : iter + 1;
    F(vec2) a1 = F(add2)(F(rotate2)(self->lv[iter], self->angle), self->pos);
    F(vec2) a2 = F(add2)(F(rotate2)(self->lv[nx], self->angle), self->pos);
    F(vec2) dir = F(normalize2)(F(sub2)(a2, a1));
    e.dir = F(rot90ccw)(dir);
    e.a = a1; e.b = a2; e.has_edge = 1;
    return e;
}

/* Disc.projectAlongAxis (Disc.zig:92-97); Rectangle.projectAlongAxis (Rectangle.zig:153-169) */
static void F(project)(const F(body) *b, F(vec2) n, R2D_REAL out[2])
{
    if (b->type == 0) {
        R2D_REAL mid = F(dot2)(b->pos, n);
        out[0] = mid - b->radius; out[1] = mid + b->radius;
        return;
    }
    R2D_REAL lo = (R2D_REAL)INFINITY, hi = -(R2D_REAL)INFINITY;
    for (int i = 0; i < 4; ++i) {
        F(vec2) w = F(add2)(F(rotate2)(b->lv[i], b->angle), b->pos);
        R2D_REAL d = F(dot2)(w, n);
        if (d < lo) lo = d;
        if (d > hi) hi = d;
    }
    out[0] = lo; out[1] = hi;
}

/* CollisionPoint.init (collision.zig:38-51) */
static F(point) F(point_init)(F(vec2) pos, R2D_REAL depth, const F(body) *ref, const F(body) *inc, F(vec2) normal)
{
    F(point) p;
    F(vec2) middle = F(sub2)(pos, F(scale2)(normal, depth / 2));
    p.valid = 1;
    p.acc_pn = 0; p.acc_pt = 0;
    p.ref_r = F(sub2)(middle, ref->pos);
    p.inc_r = F(sub2)(middle, inc->pos);
    p.pos = middle;
    p.depth = depth; p.orig_depth = depth;
    p.mass_n = 0; p.mass_t = 0;
    return p;
}

/* clipLineToLine (collision.zig:330-363): b is clipped onto a */
static void F(clip_line)(F(vec2) aa, F(vec2) ab, F(vec2) ba, F(vec2) bb, F(vec2) *o1, F(vec2) *o2)
{
    F(vec2) delta_a = F(sub2)(ab, aa);
    R2D_REAL a_len = F(length2)(delta_a);
    F(vec2) tang = F(scale2)(delta_a, (R2D_REAL)1 / a_len);
    F(vec2) p1 = ba, p2 = bb;
    {
        R2D_REAL sc = F(dot2)(F(sub2)(ba, aa), tang);
        if (!(sc > 0 && sc < a_len)) {
            F(vec2) clos = (sc < (R2D_REAL)0.5 * a_len) ? aa : ab;
            F(vec2) dp = F(sub2)(ba, bb);
            R2D_REAL t = -F(dot2)(tang, F(sub2)(bb, clos)) / F(dot2)(tang, dp);
            p1 = F(addmult2)(bb, dp, t);
        }
    }
    {
        R2D_REAL sc = F(dot2)(F(sub2)(bb, aa), tang);
        if (!(sc > 0 && sc < a_len)) {
            F(vec2) clos = (sc < (R2D_REAL)0.5 * a_len) ? aa : ab;
            F(vec2) dp = F(sub2)(bb, ba);
            R2D_REAL t = -F(dot2)(tang, F(sub2)(ba, clos)) / F(dot2)(tang, dp);
            p2 = F(addmult2)(ba, dp, t);
        }
    }
    *o1 = p1; *o2 = p2;
}

/* clipAgainstEdge: Rectangle (Rectangle.zig:213-238), Disc (Disc.zig:125-130).
 * Returns 1 for an edge (a,b), 0 for a point (a). */
static int F(clip_against_edge)(const F(body) *self, F(vec2) ea, F(vec2) eb, F(vec2) normal, F(vec2) *oa, F(vec2) *ob)
{
    if (self->type == 0) {
        *oa = F(addmult2)(self->pos, normal, -self->radius);
        return 0;
    }
    /* best_edge starts undefined; with NaN dots nothing is chosen: use the
     * first edge (documented divergence for non-finite input) */
    F(vec2) best_a = F(v2)(0, 0), best_b = F(v2)(0, 0);
    int chosen = 0;
    R2D_REAL best_dot = (R2D_REAL)INFINITY;
    F(vec2) curr = F(add2)(F(rotate2)(self->lv[0], self->angle), self->pos);
    F(vec2) first_a = curr, first_b = curr;
    for (int i = 0; i < 4; ++i) {
        int nx = (i == 3) ? 0 : i + 1;
        F(vec2) next = F(add2)(F(rotate2)(self->lv[nx], self->angle), self->pos);
        if (i == 0) first_b = next;
        F(vec2) tangent = F(normalize2)(F(sub2)(next, curr));
        F(vec2) tn = F(rot90ccw)(tangent);
        R2D_REAL d = F(dot2)(normal, tn);
        if (d < best_dot) { best_a = curr; best_b = next; best_dot = d; chosen = 1; }
        curr = next;
    }
    if (!chosen) { best_a = first_a; best_b = first_b; }
    F(clip_line)(ea, eb, best_a, best_b, oa, ob);
    return 1;
}

/* normalShouldFlipSAT (collision.zig:221-226) */
static int F(should_flip)(F(vec2) n, const F(body) *ref, const F(body) *inc)
{
    return F(dot2)(n, F(sub2)(inc->pos, ref->pos)) < (R2D_REAL)0.0;
}

/* identifyCollisionPoints: Disc (Disc.zig:99-123), Rectangle (Rectangle.zig:171-211) */
static void F(identify_points)(const F(body) *ref, const F(body) *inc, int normal_id, F(point) out[2])
{
    out[0].valid = 0; out[1].valid = 0;
    if (ref->type == 0) {
        F(vec2) pos = F(closest_point)(inc, ref->pos);
        F(edge) e = F(get_normal)(ref, inc, normal_id);
        F(vec2) n = e.dir;
        if (F(should_flip)(n, ref, inc)) n = F(negself)(n);
        R2D_REAL d = F(dot2)(n, F(sub2)(pos, ref->pos));
        R2D_REAL depth = d - ref->radius;
        out[0] = F(point_init)(pos, depth, ref, inc, n);
        return;
    }
    F(edge) e = F(get_normal)(ref, inc, normal_id);
    F(vec2) ia, ib;
    int is_edge = F(clip_against_edge)(inc, e.a, e.b, e.dir, &ia, &ib);
    if (is_edge) {
        int i = 0;
        R2D_REAL d = F(dot2)(F(sub2)(ia, e.a), e.dir);
        if (d < K_COLLISION_MARGIN) { out[i] = F(point_init)(ia, d, ref, inc, e.dir); i += 1; }
        d = F(dot2)(F(sub2)(ib, e.b), e.dir);
        if (d < K_COLLISION_MARGIN) { out[i] = F(point_init)(ib, d, ref, inc, e.dir); }
    } else {
        R2D_REAL d = F(dot2)(F(sub2)(ia, e.a), e.dir);
        out[0] = F(point_init)(ia, d, ref, inc, e.dir);
    }
}

/* SATResult (collision.zig:291-297) */
typedef struct {
    int collides;
    F(vec2) normal;
    R2D_REAL penetration;
    int normal_id;
    int key_ref, key_inc;    /* slots */
    int key_set;
} F(sat);

/* overlapSAT (collision.zig:228-289) */
static int F(overlap_sat)(F(sat) *ret, const F(body) *bodies, int ri, int ii)
{
    const F(body) *ref = &bodies[ri], *inc = &bodies[ii];
    const R2D_REAL EPS = K_SAT_OVERLAP_THRESHOLD;
    int nn = (ref->type == 0) ? 1 : 4;     /* num_normals: Disc.zig:47, Rectangle.zig:56 */
    for (int it = 0; it < nn; ++it) {
        F(edge) e = F(get_normal)(ref, inc, it);
        F(vec2) normal = e.dir;
        int flipped = 0;
        if (F(should_flip)(normal, ref, inc)) { normal = F(negself)(normal); flipped = 1; }
        R2D_REAL p1[2], p2[2];
        F(project)(ref, normal, p1);
        F(project)(inc, normal, p2);
        R2D_REAL d1 = p1[1] - p2[0];
        R2D_REAL d2 = p2[1] - p1[0];
        R2D_REAL d = F(minn)(d1, d2);
        if (d <= -K_COLLISION_MARGIN) return 0;
        if (!flipped && F(approx2)(normal, ret->normal, EPS)) {
            R2D_REAL tref = F(dot2)(ref->pos, normal);
            R2D_REAL tinc = F(dot2)(inc->pos, normal);
            if (tref < tinc - EPS) {
                ret->penetration = d; ret->normal = normal; ret->normal_id = it;
                ret->key_ref = ri; ret->key_inc = ii; ret->key_set = 1;
            }
        }
        if (!flipped && F(approx2)(normal, F(negate2)(ret->normal), EPS)) {
            F(vec2) diff = F(sub2)(inc->pos, ref->pos);
            if (F(dot2)(diff, normal) > F(dot2)(diff, ret->normal) + EPS) {
                ret->penetration = d; ret->normal = normal; ret->normal_id = it;
                ret->key_ref = ri; ret->key_inc = ii; ret->key_set = 1;
            }
        }
        if (d + EPS < ret->penetration) {
            ret->penetration = d; ret->normal = normal; ret->normal_id = it;
            ret->key_ref = ri; ret->key_inc = ii; ret->key_set = 1;
        }
    }
    return 1;
}

/* performNarrowSAT (collision.zig:299-320).  `ret.normal` starts undefined in
 * the reference; Debug builds fill undefined memory with 0xAA bytes, which is
 * used here (it only matters when every projection is NaN). */
static F(sat) F(narrow_sat)(const F(body) *bodies, int s1, int s2)
{
    F(sat) ret;
    const float undef = -3.0316488252093987e-13f;  /* 0xAAAAAAAA */
    ret.collides = 0;
    ret.penetration = (R2D_REAL)INFINITY;
    ret.normal = F(v2)((R2D_REAL)undef, (R2D_REAL)undef);
    ret.normal_id = 0; ret.key_ref = s1; ret.key_inc = s2; ret.key_set = 0;
    int o1 = (bodies[s1].id < bodies[s2].id) ? s1 : s2;
    int o2 = (o1 == s1) ? s2 : s1;
    if (!F(overlap_sat)(&ret, bodies, o1, o2)) return ret;
    if (!F(overlap_sat)(&ret, bodies, o2, o1)) return ret;
    ret.collides = 1;
    if (!ret.key_set) { ret.key_ref = o1; ret.key_inc = o2; }
    return ret;
}

/* aabb.intersects (aabb.zig:12-19) */
static int F(aabb_intersects)(const F(body) *a, const F(body) *b)
{
    R2D_REAL dx = F(absr)(b->apos.x - a->apos.x);
    R2D_REAL dy = F(absr)(b->apos.y - a->apos.y);
    return (dx <= (a->ahw + b->ahw) + K_AABB_EPS_OVERLAP) && (dy <= (a->ahh + b->ahh) + K_AABB_EPS_OVERLAP);
}

/* ------------------------------------------------------- SpatialHash.zig */
/* SpatialHash.hash (SpatialHash.zig:78-81): wrapping i64 arithmetic, u64 modulo */
static inline uint64_t F(hash)(uint64_t table_size, int64_t xi, int64_t yi)
{
    uint64_t h = ((uint64_t)xi * 92837111ull) ^ ((uint64_t)yi * 689287499ull);
    return h % table_size;
}

/* cell range of an AABB (SpatialHash.zig:83-96, aabb.zig:25-32) */
static int F(cell_range)(const F(body) *b, R2D_REAL cell, int64_t r[4])
{
    R2D_REAL minx = b->apos.x - b->ahw, miny = b->apos.y - b->ahh;
    R2D_REAL maxx = b->apos.x + b->ahw, maxy = b->apos.y + b->ahh;
    R2D_REAL v[4] = { R2D_FLOOR(minx / cell), R2D_FLOOR(miny / cell), R2D_FLOOR(maxx / cell), R2D_FLOOR(maxy / cell) };
    for (int i = 0; i < 4; ++i) {
        if (!(v[i] > (R2D_REAL)-9.0e18 && v[i] < (R2D_REAL)9.0e18)) return 0;  /* @intFromFloat UB in the reference */
        r[i] = (int64_t)v[i];
    }
    return 1;
}

/* ------------------------------------------------------------- world API */
static F(world) *F(world_new)(R2D_REAL cell_width, uint64_t table_mult)
{
    F(world) *w = (F(world) *)calloc(1, sizeof(F(world)));
    w->cell_width_arg = cell_width;
    w->table_mult_arg = table_mult;
    w->hash_cell = (R2D_REAL)4.0;
    w->hash_mult = 2;
    return w;
}

static void F(world_free)(F(world) *w)
{
    if (!w) return;
    free(w->bodies); free(w->id2slot); free(w->cons); free(w->gens); free(w->excl); free(w->mans); free(w->trace);
    free(w->wtab); free(w->wcnt);
    free(w);
}

static int32_t F(slot_of)(const F(world) *w, uint32_t id)
{
    if (id >= w->capid) return -1;
    return w->id2slot[id];
}

/* EntityFactory.appendBody (lib.zig:66-71) */
static uint32_t F(append_body)(F(world) *w, F(body) *b)
{
    uint32_t id = w->current_body_id;
    b->id = id;
    GROW(w->bodies, w->nbodies, w->capbodies, F(body));
    if (id >= w->capid) {
        uint32_t nc = w->capid ? w->capid : 64;
        while (nc <= id) nc *= 2;
        w->id2slot = (int32_t *)realloc(w->id2slot, sizeof(int32_t) * nc);
        for (uint32_t i = w->capid; i < nc; ++i) w->id2slot[i] = -1;
        w->capid = nc;
    }
    w->bodies[w->nbodies] = *b;
    w->id2slot[id] = w->nbodies;
    w->nbodies += 1;
    w->current_body_id += 1;
    return id;
}

/* makeDiscBody (lib.zig:73-82) + Disc.init (Disc.zig:29-55).
 * mass_kind: 0 density, 1 mass */
static uint32_t F(make_disc)(F(world) *w, R2D_REAL px, R2D_REAL py, R2D_REAL vx, R2D_REAL vy, R2D_REAL angle,
                             R2D_REAL omega, R2D_REAL mu, int mass_kind, R2D_REAL mass_val, R2D_REAL radius)
{
    F(body) b;
    memset(&b, 0, sizeof b);
    R2D_REAL mass = mass_kind ? mass_val : (R2D_PI * radius) * radius * mass_val;
    b.type = 0; b.is_static = 0; b.radius = radius;
    b.pos = F(v2)(px, py); b.mom = F(v2)(0, 0); b.force = F(v2)(0, 0);
    b.angle = angle; b.L = 0; b.torque = 0; b.mass = mass;
    b.inertia = (R2D_REAL)0.5 * mass * radius * radius;
    b.mu = mu;
    F(update_aabb)(&b);
    b.mom = F(scale2)(F(v2)(vx, vy), mass);
    b.L = omega * b.inertia;
    return F(append_body)(w, &b);
}

/* makeRectangleBody (lib.zig:84-93) + Rectangle.init (Rectangle.zig:33-64) */
static uint32_t F(make_rect)(F(world) *w, R2D_REAL px, R2D_REAL py, R2D_REAL vx, R2D_REAL vy, R2D_REAL angle,
                             R2D_REAL omega, R2D_REAL mu, int mass_kind, R2D_REAL mass_val, R2D_REAL width, R2D_REAL height)
{
    F(body) b;
    memset(&b, 0, sizeof b);
    R2D_REAL mass = mass_kind ? mass_val : width * height * mass_val;
    R2D_REAL hw = width / 2, hh = height / 2;
    b.type = 1; b.is_static = 0; b.width = width; b.height = height;
    b.lv[0] = F(v2)(-hw, -hh); b.lv[1] = F(v2)(-hw, hh); b.lv[2] = F(v2)(hw, hh); b.lv[3] = F(v2)(hw, -hh);
    b.pos = F(v2)(px, py); b.mom = F(v2)(0, 0); b.force = F(v2)(0, 0);
    b.angle = angle; b.L = 0; b.torque = 0; b.mass = mass;
    b.inertia = mass * (width * width + height * height) / 12;
    b.mu = mu;
    F(update_aabb)(&b);
    b.mom = F(scale2)(F(v2)(vx, vy), mass);
    b.L = omega * b.inertia;
    return F(append_body)(w, &b);
}

/* Solver.removeRigidBody (lib.zig:308-310): AutoArrayHashMap.swapRemove */
static int F(remove_body)(F(world) *w, uint32_t id)
{
    int32_t s = F(slot_of)(w, id);
    if (s < 0) return -1;
    int32_t last = w->nbodies - 1;
    if (s != last) {
        w->bodies[s] = w->bodies[last];
        w->id2slot[w->bodies[s].id] = s;
    }
    w->id2slot[id] = -1;
    w->nbodies -= 1;
    if (w->wcnt) memset(w->wcnt, 0, sizeof(int32_t) * (size_t)w->wcap);   /* tables are kept under slots */
    return 0;
}

static void F(add_constraint)(F(world) *w, const F(constraint) *c)
{
    GROW(w->cons, w->ncons, w->capcons, F(constraint));
    w->cons[w->ncons++] = *c;
}

static void F(add_gen)(F(world) *w, const F(forcegen) *g)
{
    GROW(w->gens, w->ngens, w->capgens, F(forcegen));
    w->gens[w->ngens++] = *g;
}

/* excludeCollisionPair (lib.zig:124-129): both orders */
static void F(exclude)(F(world) *w, uint32_t a, uint32_t b)
{
    GROW(w->excl, w->nexcl + 1, w->capexcl, uint64_t);
    w->excl[w->nexcl++] = ((uint64_t)a << 32) | b;
    w->excl[w->nexcl++] = ((uint64_t)b << 32) | a;
}

static int F(is_excluded)(const F(world) *w, uint32_t a, uint32_t b)
{
    uint64_t k = ((uint64_t)a << 32) | b;
    for (int32_t i = 0; i < w->nexcl; ++i)
        if (w->excl[i] == k) return 1;
    return 0;
}

/* --------------------------------------------------- manifold key set */
typedef struct { uint64_t *keys; int32_t *vals; uint64_t cap; } F(kset);
static void F(kset_init)(F(kset) *s, uint64_t n)
{
    uint64_t cap = 16;
    while (cap < 2 * n + 16) cap *= 2;
    s->cap = cap;
    s->keys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    s->vals = (int32_t *)malloc(sizeof(int32_t) * cap);
    for (uint64_t i = 0; i < cap; ++i) s->keys[i] = ~0ull;
}
static void F(kset_free)(F(kset) *s) { free(s->keys); free(s->vals); }
static uint64_t F(kmix)(uint64_t k) { k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; return k; }
static int F(kset_has)(const F(kset) *s, uint64_t k)
{
    uint64_t i = F(kmix)(k) & (s->cap - 1);
    while (s->keys[i] != ~0ull) { if (s->keys[i] == k) return 1; i = (i + 1) & (s->cap - 1); }
    return 0;
}
static void F(kset_put)(F(kset) *s, uint64_t k, int32_t v)
{
    uint64_t i = F(kmix)(k) & (s->cap - 1);
    while (s->keys[i] != ~0ull) { if (s->keys[i] == k) { s->vals[i] = v; return; } i = (i + 1) & (s->cap - 1); }
    s->keys[i] = k; s->vals[i] = v;
}

/* Solver.updateManifolds (lib.zig:253-299) with SpatialHash.init/query
 * (SpatialHash.zig:19-71, 108-137) restated step by step: counting pass,
 * inclusive prefix, pre-decrement scatter, raster-order queries. */
static void F(update_manifolds)(F(world) *w)
{
    int32_t N = w->nbodies;
    uint64_t T = w->hash_mult * (uint64_t)N;
    R2D_REAL cell = w->hash_cell;
    w->nmans = 0;
    w->ntrace = 0;
    if (N == 0) return;
    uint64_t *table = (uint64_t *)calloc(T + 1, sizeof(uint64_t));
    int64_t r[4];
    for (int32_t s = 0; s < N; ++s) {
        if (!F(cell_range)(&w->bodies[s], cell, r)) continue;
        for (int64_t yi = r[1]; yi <= r[3]; ++yi)
            for (int64_t xi = r[0]; xi <= r[2]; ++xi) table[F(hash)(T, xi, yi)] += 1;
    }
    uint64_t start = 0;
    for (uint64_t i = 0; i < T; ++i) { start += table[i]; table[i] = start; }
    table[T] = start;
    int32_t *indices = (int32_t *)malloc(sizeof(int32_t) * (start ? start : 1));
    for (int32_t s = 0; s < N; ++s) {
        if (!F(cell_range)(&w->bodies[s], cell, r)) continue;
        for (int64_t yi = r[1]; yi <= r[3]; ++yi)
            for (int64_t xi = r[0]; xi <= r[2]; ++xi) {
                uint64_t h = F(hash)(T, xi, yi);
                table[h] -= 1;
                indices[table[h]] = s;
            }
    }
    F(kset) ks;
    F(kset_init)(&ks, (uint64_t)N * 8);
    int32_t *q = NULL; int64_t nq = 0, capq = 0;
    for (int32_t a = 0; a < N; ++a) {
        F(body) *b1 = &w->bodies[a];
        nq = 0;
        if (F(cell_range)(b1, cell, r)) {
            for (int64_t yi = r[1]; yi <= r[3]; ++yi)
                for (int64_t xi = r[0]; xi <= r[2]; ++xi) {
                    uint64_t h = F(hash)(T, xi, yi);
                    for (uint64_t i = table[h]; i < table[h + 1]; ++i) {
                        GROW(q, nq, capq, int32_t);
                        q[nq++] = indices[i];
                    }
                }
        }
        for (int64_t qi = 0; qi < nq; ++qi) {
            int32_t bslot = q[qi];
            F(body) *b2 = &w->bodies[bslot];
            if (b1->is_static && b2->is_static) continue;
            if (a == bslot) continue;
            if (F(is_excluded)(w, b1->id, b2->id)) continue;
            if (F(is_excluded)(w, b2->id, b1->id)) continue;
            uint64_t k1 = ((uint64_t)(uint32_t)a << 32) | (uint32_t)bslot;
            uint64_t k2 = ((uint64_t)(uint32_t)bslot << 32) | (uint32_t)a;
            if (F(kset_has)(&ks, k1) || F(kset_has)(&ks, k2)) continue;
            if (!F(aabb_intersects)(b1, b2)) continue;
            if (w->trace_on) {
                GROW(w->trace, w->ntrace + 1, w->captrace, int32_t);
                w->trace[w->ntrace++] = a; w->trace[w->ntrace++] = bslot;
            }
            F(sat) sat = F(narrow_sat)(w->bodies, a, bslot);
            if (!sat.collides) continue;
            F(manifold) m;
            memset(&m, 0, sizeof m);
            m.ref = sat.key_ref; m.inc = sat.key_inc;
            m.normal = sat.normal;
            F(identify_points)(&w->bodies[m.ref], &w->bodies[m.inc], sat.normal_id, m.pts);
            m.prev_angle_1 = w->bodies[m.ref].angle;
            m.prev_angle_2 = w->bodies[m.inc].angle;
            GROW(w->mans, w->nmans, w->capmans, F(manifold));
            w->mans[w->nmans] = m;
            F(kset_put)(&ks, ((uint64_t)(uint32_t)m.ref << 32) | (uint32_t)m.inc, w->nmans);
            w->nmans += 1;
        }
    }
    free(q);
    F(kset_free)(&ks);
    free(indices);
    free(table);
}

/* CollisionManifold.updateTGSDepth (collision.zig:73-100) */
static void F(update_tgs_depth)(F(manifold) *m, const F(body) *b1, const F(body) *b2)
{
    for (int i = 0; i < 2; ++i) {
        F(point) *p = &m->pts[i];
        if (!p->valid) continue;
        if (F(approx)(p->depth, p->orig_depth, (R2D_REAL)1e-4f)) continue;
        F(vec2) r1 = p->ref_r, r2 = p->inc_r;
        F(vec2) rr1 = F(rotate2)(r1, b1->angle - m->prev_angle_1);
        F(vec2) rr2 = F(rotate2)(r2, b2->angle - m->prev_angle_2);
        F(vec2) a1 = F(add2)(rr1, b1->pos);
        F(vec2) a2 = F(add2)(rr2, b2->pos);
        R2D_REAL depth = F(dot2)(m->normal, F(sub2)(a2, a1));
        p->depth = depth + p->orig_depth;
        p->ref_r = rr1; p->inc_r = rr2;
    }
    m->prev_angle_1 = b1->angle;
    m->prev_angle_2 = b2->angle;
}

/* CollisionManifold.preStep (collision.zig:102-133) */
static void F(pre_step)(F(manifold) *m, const F(body) *b1, const F(body) *b2)
{
    R2D_REAL inv_m1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->mass);
    R2D_REAL inv_m2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->mass);
    R2D_REAL inv_mass = inv_m1 + inv_m2;
    R2D_REAL inv_i1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->inertia);
    R2D_REAL inv_i2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->inertia);
    m->tangent = F(rot90cw)(m->normal);
    m->friction = R2D_SQRT(b1->mu * b2->mu);
    for (int i = 0; i < 2; ++i) {
        F(point) *p = &m->pts[i];
        if (!p->valid) continue;
        F(vec2) r1 = p->ref_r, r2 = p->inc_r;
        R2D_REAL r1n = F(cross2)(r1, m->normal);
        R2D_REAL r2n = F(cross2)(r2, m->normal);
        R2D_REAL kn = inv_mass + inv_i1 * (r1n * r1n) + inv_i2 * (r2n * r2n);
        R2D_REAL r1t = F(cross2)(r1, m->tangent);
        R2D_REAL r2t = F(cross2)(r2, m->tangent);
        R2D_REAL kt = inv_mass + inv_i1 * (r1t * r1t) + inv_i2 * (r2t * r2t);
        p->mass_n = (kn > (R2D_REAL)0.0) ? ((R2D_REAL)1 / kn) : (R2D_REAL)0.0;
        p->mass_t = (kt > (R2D_REAL)0.0) ? ((R2D_REAL)1 / kt) : (R2D_REAL)0.0;
    }
}

/* CollisionManifold.calculateImpulses (collision.zig:135-218) */
static void F(calc_impulses)(F(manifold) *m, F(body) *b1, F(body) *b2, R2D_REAL dt)
{
    R2D_REAL inv_m1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->mass);
    R2D_REAL inv_m2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->mass);
    R2D_REAL inv_i1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->inertia);
    R2D_REAL inv_i2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->inertia);
    F(vec2) vl1 = F(scale2)(b1->mom, inv_m1);
    R2D_REAL w1 = b1->L * inv_i1;
    F(vec2) vl2 = F(scale2)(b2->mom, inv_m2);
    R2D_REAL w2 = b2->L * inv_i2;
    for (int i = 0; i < 2; ++i) {
        F(point) *p = &m->pts[i];
        if (!p->valid) continue;
        if (p->depth >= 0) { p->acc_pn = 0; p->acc_pt = 0; continue; }
        F(vec2) r1 = p->ref_r, r2 = p->inc_r;
        F(vec2) vr1 = F(v2)(-r1.y * w1, r1.x * w1);
        F(vec2) v1 = F(add2)(vl1, vr1);
        F(vec2) vr2 = F(v2)(-r2.y * w2, r2.x * w2);
        F(vec2) v2v = F(add2)(vl2, vr2);
        F(vec2) dv = F(sub2)(v1, v2v);
        R2D_REAL bias = K_BAUMGARTE * F(maxn)(0, (-p->depth) - K_BAUMGARTE_SLOP) / dt;
        R2D_REAL num = F(dot2)(dv, m->normal) + bias;
        R2D_REAL pn = num * p->mass_n;
        if (pn < K_MIN_MANIFOLD_IMPULSE && !m->warm) continue;
        num = F(dot2)(dv, m->tangent);
        R2D_REAL pt = num * p->mass_t;
        R2D_REAL new_pn = F(maxn)(0, p->acc_pn + pn);
        R2D_REAL applied_pn = new_pn - p->acc_pn;
        p->acc_pn = new_pn;
        R2D_REAL max_pt = m->friction * F(absr)(p->acc_pn);
        R2D_REAL new_pt = F(clampn)(p->acc_pt + pt, -max_pt, max_pt);
        R2D_REAL applied_pt = new_pt - p->acc_pt;
        p->acc_pt = new_pt;
        F(vec2) pn_vec = F(scale2)(m->normal, applied_pn);
        F(vec2) pt_vec = F(scale2)(m->tangent, applied_pt);
        F(vec2) dp = F(add2)(pn_vec, pt_vec);
        if (!b1->is_static) {
            m->app_lin1.x -= dp.x; m->app_lin1.y -= dp.y;
            m->app_rot1 -= F(cross2)(r1, dp);
        }
        if (!b2->is_static) {
            m->app_lin2.x += dp.x; m->app_lin2.y += dp.y;
            m->app_rot2 += F(cross2)(r2, dp);
        }
    }
    b1->mom.x += m->app_lin1.x; b1->mom.y += m->app_lin1.y;
    b1->L += m->app_rot1;
    b2->mom.x += m->app_lin2.x; b2->mom.y += m->app_lin2.y;
    b2->L += m->app_rot2;
    m->app_lin1 = F(v2)(0, 0); m->app_rot1 = 0;
    m->app_lin2 = F(v2)(0, 0); m->app_rot2 = 0;
}

/* -------------------------------------------------- Constraints (joints) */
/* DistanceJoint.solve (DistanceJoint.zig:40-75) */
static void F(solve_distance)(const F(constraint) *c, F(body) *b1, F(body) *b2)
{
    R2D_REAL eps = K_GRAD_DIV_LIMIT;
    R2D_REAL w1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->mass);
    R2D_REAL w2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->mass);
    F(vec2) v1 = F(scale2)(b1->mom, w1);
    F(vec2) v2v = F(scale2)(b2->mom, w2);
    F(vec2) rv = F(sub2)(v2v, v1);
    F(vec2) n = F(normalize2)(F(sub2)(b2->pos, b1->pos));
    R2D_REAL dist_err = F(length2)(F(sub2)(b2->pos, b1->pos)) - c->target;
    R2D_REAL pc = c->beta * dist_err / F(maxn)(c->target, (R2D_REAL)0.001f);
    R2D_REAL J = -(F(dot2)(rv, n) + pc) / (w1 + w2);
    R2D_REAL ndv = F(dot2)(n, rv);
    R2D_REAL den = (R2D_REAL)1 / F(maxn)(F(absr)(ndv), eps);
    R2D_REAL jmax = c->pmax * den, jmin = c->pmin * den;
    J = F(clampn)(J, jmin, jmax);
    F(vec2) dp = F(scale2)(n, J);
    b1->mom.x -= dp.x; b1->mom.y -= dp.y;
    b2->mom.x += dp.x; b2->mom.y += dp.y;
}

/* OffsetDistanceJoint.solve (OffsetDistanceJoint.zig:44-105) */
static void F(solve_offset)(const F(constraint) *c, F(body) *b1, F(body) *b2)
{
    R2D_REAL eps = K_GRAD_DIV_LIMIT;
#if R2D_PLANT == 3
    F(vec2) a1 = F(add2)(c->r1, b1->pos);
    F(vec2) a2 = F(add2)(c->r2, b2->pos);
#else
    F(vec2) a1 = F(add2)(F(rotate2)(c->r1, b1->angle), b1->pos);   /* localToWorld, RigidBody.zig:113-117 */
    F(vec2) a2 = F(add2)(F(rotate2)(c->r2, b2->angle), b2->pos);
#endif
    F(vec2) n = F(normalize2)(F(sub2)(a2, a1));
    R2D_REAL im1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->mass);
    R2D_REAL im2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->mass);
    R2D_REAL ii1 = b1->is_static ? 0 : ((R2D_REAL)1 / b1->inertia);
    R2D_REAL ii2 = b2->is_static ? 0 : ((R2D_REAL)1 / b2->inertia);
    F(vec2) vl1 = F(scale2)(b1->mom, im1);
    R2D_REAL w1 = b1->L * ii1;
    F(vec2) vl2 = F(scale2)(b2->mom, im2);
    R2D_REAL w2 = b2->L * ii2;
#if R2D_PLANT == 2
    F(vec2) rr1 = c->r1;
#else
    F(vec2) rr1 = F(rotate2)(c->r1, b1->angle);
#endif
    F(vec2) va1 = F(v2)(-rr1.y * w1, rr1.x * w1);
    F(vec2) v1 = F(add2)(vl1, va1);
#if R2D_PLANT == 2
    F(vec2) rr2 = c->r2;
#else
    F(vec2) rr2 = F(rotate2)(c->r2, b2->angle);
#endif
    F(vec2) va2 = F(v2)(-rr2.y * w2, rr2.x * w2);
    F(vec2) v2v = F(add2)(vl2, va2);
    F(vec2) dv = F(sub2)(v2v, v1);
    R2D_REAL dist_err = F(length2)(F(sub2)(a2, a1)) - c->target;
    R2D_REAL pc = c->beta * dist_err / F(maxn)(c->target, eps);
    R2D_REAL num = F(dot2)(n, dv) + pc;
    R2D_REAL r1xn = F(cross2)(rr1, n);
    R2D_REAL r2xn = F(cross2)(rr2, n);
    R2D_REAL den = im1 + im2 + r1xn * r1xn * ii1 + r2xn * r2xn * ii2;
    R2D_REAL J = -num / den;
    R2D_REAL ndv = F(dot2)(n, dv);
    R2D_REAL den2 = (R2D_REAL)1 / F(maxn)(F(absr)(ndv), eps);
    R2D_REAL jmax = c->pmax * den2, jmin = c->pmin * den2;
    J = F(clampn)(J, jmin, jmax);
    F(vec2) dp = F(scale2)(n, J);
    b1->mom.x -= dp.x; b1->mom.y -= dp.y;
#if R2D_PLANT == 1
    b1->L += J * r1xn;
#else
    b1->L -= J * r1xn;
#endif
    b2->mom.x += dp.x; b2->mom.y += dp.y;
    b2->L += J * r2xn;
}

/* FixedPositionJoint.solve (FixedPositionJoint.zig:38-72) */
static void F(solve_fixed)(const F(constraint) *c, F(body) *b)
{
    R2D_REAL eps = K_ALLOWED_CONSTRAINT_VALUE;
    if (b->is_static) return;
    R2D_REAL w = (R2D_REAL)1 / b->mass;
    F(vec2) v = F(scale2)(b->mom, w);
    F(vec2) delta = F(sub2)(c->tpos, b->pos);
    R2D_REAL dist = F(length2)(delta);
    if (dist < K_ALLOWED_CONSTRAINT_VALUE) return;
    F(vec2) n = F(scale2)(delta, (R2D_REAL)1 / dist);
    R2D_REAL bias = c->beta * dist;
    R2D_REAL rv = F(dot2)(v, n);
    R2D_REAL J = -(rv - bias) / w;
    R2D_REAL den = (R2D_REAL)1 / F(maxn)(F(absr)(rv), eps);
    R2D_REAL jmax = c->pmax * den, jmin = c->pmin * den;
    J = F(clampn)(J, jmin, jmax);
    F(vec2) dp = F(scale2)(n, J);
    b->mom.x += dp.x; b->mom.y += dp.y;
}

/* MotorJoint.solve (MotorJoint.zig:38-64); dt is the sub-step */
static void F(solve_motor)(const F(constraint) *c, F(body) *b, R2D_REAL dt)
{
    R2D_REAL eps = K_ALLOWED_CONSTRAINT_VALUE;
    if (b->is_static) return;
    R2D_REAL omega = b->L / b->inertia;
    R2D_REAL C = omega - c->target;
    if (F(absr)(C) < K_ALLOWED_CONSTRAINT_VALUE) return;
    R2D_REAL bias = c->beta * C;
    R2D_REAL J = b->torque - bias;
    R2D_REAL rel = b->torque / b->inertia;
    R2D_REAL den = (R2D_REAL)1 / F(maxn)(F(absr)(rel), eps);
    R2D_REAL jmax = c->pmax * den, jmin = c->pmin * den;
    J = F(clampn)(J, jmin, jmax);
    b->L += J * dt;
}

/* Constraint.solve dispatch; InvalidRigidBodyId when an id is gone */
static int F(solve_constraint)(F(world) *w, const F(constraint) *c, R2D_REAL sub_dt)
{
    int32_t s1 = F(slot_of)(w, c->id1);
    if (s1 < 0) return -2;
    if (c->type == 0 || c->type == 1) {
        int32_t s2 = F(slot_of)(w, c->id2);
        if (s2 < 0) return -2;
        if (c->type == 0) F(solve_distance)(c, &w->bodies[s1], &w->bodies[s2]);
        else F(solve_offset)(c, &w->bodies[s1], &w->bodies[s2]);
    } else if (c->type == 2) {
        F(solve_fixed)(c, &w->bodies[s1]);
    } else {
        F(solve_motor)(c, &w->bodies[s1], sub_dt);
    }
    return 0;
}

/* ForceGen.apply: DownwardsGravity.apply (DownwardsGravity.zig:35-39).
 * Spring (extension, no reference): F += -k (x - anchor) - c v, v = p / m,
 * applied to its one body when dynamic. */
static void F(apply_gen)(const F(forcegen) *g, F(body) *b)
{
    if (g->type == 0) {
        if (b->is_static) return;
        b->force.x += g->g_vec.x * b->mass;
        b->force.y += g->g_vec.y * b->mass;
        return;
    }
    if (b->id != g->id || b->is_static) return;
    R2D_REAL inv_m = (R2D_REAL)1 / b->mass;
    F(vec2) v = F(scale2)(b->mom, inv_m);
    F(vec2) d = F(sub2)(b->pos, g->anchor);
#if R2D_PLANT == 4
    b->force.x += (-g->k) * d.x;
    b->force.y += (-g->k) * d.y;
#elif R2D_PLANT == 5
    b->force.x += (-g->k) * d.x - g->c * b->mom.x;
    b->force.y += (-g->k) * d.y - g->c * b->mom.y;
#else
    b->force.x += (-g->k) * d.x - g->c * v.x;
    b->force.y += (-g->k) * d.y - g->c * v.y;
#endif
}

/* Coloured schedule (extension, no reference counterpart).  Items are the
 * constraints followed by the manifolds.  An item's ordering bodies are both
 * bodies of a distance / offset joint (static or not), the body of a fixed /
 * motor joint, and those bodies of a manifold that are dynamic, touched by a
 * distance / offset joint, or hold a -0 momentum component (a contact's
 * "+= 0" on any other static body changes nothing).  Greedy colouring in item
 * order: smallest colour unused at the item's ordering bodies; past 128
 * colours every item gets a colour of its own.  Writes the item permutation
 * sorted by (colour, item). */
static void F(color_items)(F(world) *w, int32_t *perm)
{
    int32_t N = w->nbodies, C = w->ncons, M = w->nmans, I = C + M;
    uint8_t *jt = (uint8_t *)calloc((size_t)N + 1, 1);
    for (int32_t c = 0; c < C; ++c) {
        if (w->cons[c].type > 1) continue;
        int32_t s1 = F(slot_of)(w, w->cons[c].id1), s2 = F(slot_of)(w, w->cons[c].id2);
        if (s1 >= 0) jt[s1] = 1;
        if (s2 >= 0) jt[s2] = 1;
    }
    uint64_t *mask = (uint64_t *)calloc(2 * (size_t)N + 2, sizeof(uint64_t));
    uint32_t *col = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(I + 1));
    for (int32_t it = 0; it < I; ++it) {
        int32_t r[2] = { -1, -1 };
        if (it < C) {
            const F(constraint) *c = &w->cons[it];
            r[0] = F(slot_of)(w, c->id1);
            if (c->type <= 1) { int32_t s2 = F(slot_of)(w, c->id2); if (s2 != r[0]) r[1] = s2; }
        } else {
            const F(manifold) *m = &w->mans[it - C];
            int32_t e[2] = { m->ref, m->inc };
            for (int k = 0; k < 2; ++k) {
                const F(body) *b = &w->bodies[e[k]];
                int negz = (b->mom.x == 0 && signbit(b->mom.x)) || (b->mom.y == 0 && signbit(b->mom.y)) ||
                           (b->L == 0 && signbit(b->L));
                if (!b->is_static || jt[e[k]] || negz) r[k] = e[k];
            }
        }
        uint64_t m0 = 0, m1 = 0;
        for (int k = 0; k < 2; ++k) if (r[k] >= 0) { m0 |= mask[2 * r[k]]; m1 |= mask[2 * r[k] + 1]; }
        uint32_t cc;
        if (~m0) cc = (uint32_t)__builtin_ctzll(~m0);
        else if (~m1) cc = 64u + (uint32_t)__builtin_ctzll(~m1);
        else cc = 128u + (uint32_t)it;   /* a colour of its own, increasing with the item */
        col[it] = cc;
        for (int k = 0; k < 2; ++k) {
            if (r[k] < 0) continue;
            if (cc < 64) mask[2 * r[k]] |= 1ull << cc;
            else if (cc < 128) mask[2 * r[k] + 1] |= 1ull << (cc - 64);
        }
    }
    /* stable counting order by colour */
    for (int32_t it = 0; it < I; ++it) perm[it] = it;
    for (int32_t i = 1; i < I; ++i) {
        int32_t v = perm[i], j = i;
        while (j > 0 && col[perm[j - 1]] > col[v]) { perm[j] = perm[j - 1]; --j; }
        perm[j] = v;
    }
    free(col); free(mask); free(jt);
}

/* Solver.process (lib.zig:189-251).  Returns 0, or -2 for
 * error.InvalidRigidBodyId (state is left as the reference leaves it). */
/* ------------------------------------------------------ warm starting
 * Not in the reference (its roadmap: "persist manifolds across frames by
 * keying with stable body IDs", README.md:60-61).  The rule, step by step:
 * match (after updateManifolds), per-body sums in manifold order, and at
 * every sub-step the sums with the forces and the carried impulses into
 * the accumulators before the iterations; after the step the manifolds are
 * kept for the next one. */
static void F(warm_grow)(F(world) *w)
{
    if (w->nbodies <= w->wcap) return;
    int32_t nc = w->wcap ? w->wcap : 16;
    while (nc < w->nbodies) nc *= 2;
    w->wtab = (F(wentry) *)realloc(w->wtab, sizeof(F(wentry)) * (size_t)nc * WARM_K_);
    w->wcnt = (int32_t *)realloc(w->wcnt, sizeof(int32_t) * (size_t)nc);
    for (int32_t i = w->wcap; i < nc; ++i) w->wcnt[i] = 0;
    w->wcap = nc;
}

static int32_t F(point_mask)(const F(manifold) *M) { return (M->pts[0].valid ? 1 : 0) | (M->pts[1].valid ? 2 : 0); }

static void F(warm_match)(F(world) *w, F(manifold) *M, R2D_REAL f_sub)
{
    M->warm = 0;
    for (int i = 0; i < 4; ++i) M->wg[i] = 0;
    int live = 0;
    for (int p = 0; p < 2; ++p)
        if (M->pts[p].valid) live |= !F(approx)(M->pts[p].depth, M->pts[p].orig_depth, (R2D_REAL)1e-4f);
    if (live) return;
    int32_t lo = M->ref < M->inc ? M->ref : M->inc;
    int32_t c = w->wcnt[lo] < WARM_K_ ? w->wcnt[lo] : WARM_K_;
    uint32_t ida = w->bodies[M->ref].id, idb = w->bodies[M->inc].id;
    for (int32_t k = 0; k < c; ++k) {
        const F(wentry) *e = &w->wtab[(size_t)lo * WARM_K_ + k];
        if (e->ida != ida || e->idb != idb) continue;
        R2D_REAL d = F(dot2)(M->normal, e->n);
        if (e->mask == F(point_mask)(M) && d >= (R2D_REAL)0.99f) {
            M->warm = 1;
            for (int p = 0; p < 2; ++p) {
                if (!M->pts[p].valid || !(M->pts[p].depth < 0)) continue;
                M->wg[2 * p] = (w->warm * e->acc[2 * p]) / f_sub;
                M->wg[2 * p + 1] = (w->warm * e->acc[2 * p + 1]) / f_sub;
            }
        }
        break;
    }
}

static void F(warm_save)(F(world) *w)
{
    for (int32_t b = 0; b < w->nbodies; ++b) w->wcnt[b] = 0;
    for (int32_t m = 0; m < w->nmans; ++m) {
        const F(manifold) *M = &w->mans[m];
        int32_t lo = M->ref < M->inc ? M->ref : M->inc;
        if (w->wcnt[lo] >= WARM_K_) continue;
        F(wentry) *e = &w->wtab[(size_t)lo * WARM_K_ + w->wcnt[lo]++];
        e->ida = w->bodies[M->ref].id;
        e->idb = w->bodies[M->inc].id;
        e->mask = F(point_mask)(M);
        e->n = M->normal;
        for (int p = 0; p < 2; ++p) { e->acc[2 * p] = M->pts[p].acc_pn; e->acc[2 * p + 1] = M->pts[p].acc_pt; }
    }
}

static int F(process)(F(world) *w, R2D_REAL dt, uint64_t sub_steps, uint64_t iters)
{
    R2D_REAL f_sub = (R2D_REAL)(float)sub_steps;
    R2D_REAL sub_dt = dt / f_sub;
    int32_t *perm = NULL;
    if (w->sched == 1 && sub_steps > 0 && iters > 0) {
        /* the coloured order is not the constraint order: validate ids up front */
        for (int32_t c = 0; c < w->ncons; ++c) {
            if (F(slot_of)(w, w->cons[c].id1) < 0) return -2;
            if (w->cons[c].type <= 1 && F(slot_of)(w, w->cons[c].id2) < 0) return -2;
        }
    }
    F(update_manifolds)(w);
    if (w->sched == 1) {
        perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w->ncons + w->nmans + 1));
        F(color_items)(w, perm);
    }
    /* warm starting: matches, then each body's carried impulse per sub-step,
     * summed in manifold order as calculateImpulses applies impulses */
    R2D_REAL *wsum = NULL;
    const int warm = w->warm > 0;
    if (warm) {
        F(warm_grow)(w);
        wsum = (R2D_REAL *)calloc((size_t)w->nbodies * 3 + 1, sizeof(R2D_REAL));
        w->nwarm = 0;
        for (int32_t m = 0; m < w->nmans; ++m) {
            F(manifold) *M = &w->mans[m];
            F(warm_match)(w, M, f_sub);
            if (!M->warm) continue;
            w->nwarm += 1;
            F(vec2) t = F(rot90cw)(M->normal);
            int s1 = w->bodies[M->ref].is_static, s2 = w->bodies[M->inc].is_static;
            for (int p = 0; p < 2; ++p) {
                if (!M->pts[p].valid) continue;
                F(vec2) dp = F(add2)(F(scale2)(M->normal, M->wg[2 * p]), F(scale2)(t, M->wg[2 * p + 1]));
                R2D_REAL *a = wsum + 3 * (size_t)M->ref, *b = wsum + 3 * (size_t)M->inc;
                if (!s1) { a[0] -= dp.x; a[1] -= dp.y; a[2] -= F(cross2)(M->pts[p].ref_r, dp); }
                if (!s2) { b[0] += dp.x; b[1] += dp.y; b[2] += F(cross2)(M->pts[p].inc_r, dp); }
            }
        }
    }
    for (uint64_t s = 0; s < sub_steps; ++s) {
        for (int32_t gi = 0; gi < w->ngens; ++gi)
            for (int32_t b = 0; b < w->nbodies; ++b) F(apply_gen)(&w->gens[gi], &w->bodies[b]);
        for (int32_t b = 0; b < w->nbodies; ++b) {
            F(body) *B = &w->bodies[b];
            F(update_aabb)(B);
            if (B->is_static) continue;
            B->mom.x += B->force.x * sub_dt;
            B->mom.y += B->force.y * sub_dt;
            B->L += B->torque * sub_dt;
            if (warm) {
                B->mom.x += wsum[3 * (size_t)b];
                B->mom.y += wsum[3 * (size_t)b + 1];
                B->L += wsum[3 * (size_t)b + 2];
            }
        }
        for (int32_t m = 0; m < w->nmans; ++m) {
            F(manifold) *M = &w->mans[m];
            F(update_tgs_depth)(M, &w->bodies[M->ref], &w->bodies[M->inc]);
            F(pre_step)(M, &w->bodies[M->ref], &w->bodies[M->inc]);
            if (M->warm)
                for (int p = 0; p < 2; ++p) {
                    if (!M->pts[p].valid) continue;
                    M->pts[p].acc_pn += M->wg[2 * p];
                    M->pts[p].acc_pt += M->wg[2 * p + 1];
                }
        }
        for (uint64_t k = 0; k < iters; ++k) {
            if (perm) {
                for (int32_t q = 0; q < w->ncons + w->nmans; ++q) {
                    int32_t it = perm[q];
                    if (it < w->ncons) {
                        F(solve_constraint)(w, &w->cons[it], sub_dt);
                    } else {
                        F(manifold) *M = &w->mans[it - w->ncons];
                        F(calc_impulses)(M, &w->bodies[M->ref], &w->bodies[M->inc], sub_dt);
                    }
                }
                continue;
            }
            for (int32_t c = 0; c < w->ncons; ++c) {
                int rc = F(solve_constraint)(w, &w->cons[c], sub_dt);
                if (rc) { free(wsum); return rc; }
            }
            for (int32_t m = 0; m < w->nmans; ++m) {
                F(manifold) *M = &w->mans[m];
                F(calc_impulses)(M, &w->bodies[M->ref], &w->bodies[M->inc], sub_dt);
            }
        }
        for (int32_t b = 0; b < w->nbodies; ++b) {
            F(body) *B = &w->bodies[b];
            if (B->is_static) continue;
            R2D_REAL q = sub_dt / B->mass;
            B->pos.x += B->mom.x * q;
            B->pos.y += B->mom.y * q;
            B->angle += B->L * sub_dt / B->inertia;
            B->force = F(v2)(0, 0);
            B->torque = 0;
        }
    }
    if (warm) F(warm_save)(w);
    free(wsum);
    free(perm);
    return 0;
}

#undef P_
#undef P2
#undef F
#undef GROW
#undef K_MIN_MANIFOLD_IMPULSE
#undef K_BAUMGARTE
#undef K_BAUMGARTE_SLOP
#undef K_SAT_OVERLAP_THRESHOLD
#undef K_COLLISION_MARGIN
#undef K_AABB_EPS_OVERLAP
#undef K_GRAD_DIV_LIMIT
#undef K_ALLOWED_CONSTRAINT_VALUE
