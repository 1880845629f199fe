/*
 * am_oracle.c - CPU restatement of the AdaptiveMerging 3D contact path.
 * TEST INFRASTRUCTURE ONLY (see am_oracle.h).  Citations are to
 * src/mergingBodies3D/ of the reference.  Build with -ffp-contract=off.
 */
#include "am_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Java Math.max / Math.min semantics (NaN propagates; -0 < +0).        */
/* ------------------------------------------------------------------ */
static double jmax(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && signbit(a)) return b;
    return (a >= b) ? a : b;
}
static double jmin(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && signbit(b)) return b;
    return (a <= b) ? a : b;
}

/* ------------------------------------------------------------------ */
/* javax.vecmath operations, in their documented evaluation order.     */
/* ------------------------------------------------------------------ */
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(double* o, const double* a, const double* b) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = b[0] * a[2] - b[2] * a[0];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
static double len3(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void normalize3(double* a) {
    double n = 1.0 / sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] *= n; a[1] *= n; a[2] *= n;
}
/* Matrix3d.transform(v, out): out = M v */
static void mat_vec(double* o, const double* M, const double* v) {
    double x = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    double y = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    double z = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
    o[0] = x; o[1] = y; o[2] = z;
}
/* R^T v as written in RigidTransform3D.inverseTransform (RigidTransform3D.java:160-164) */
static void mat_tvec(double* o, const double* M, const double* v) {
    double x = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
    double y = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
    double z = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
    o[0] = x; o[1] = y; o[2] = z;
}
/* RigidTransform3D.transform(Point3d): R p + t */
static void xform_point(double* o, const double* R, const double* t, const double* p) {
    double q[3];
    mat_vec(q, R, p);
    o[0] = q[0] + t[0]; o[1] = q[1] + t[1]; o[2] = q[2] + t[2];
}
/* RigidTransform3D.inverseTransform(Point3d): R^T (p - t) (RigidTransform3D.java:144-151) */
static void inv_xform_point(double* o, const double* R, const double* t, const double* p) {
    double d[3] = {p[0] - t[0], p[1] - t[1], p[2] - t[2]};
    mat_tvec(o, R, d);
}
/* Matrix3d.mul(A, B) */
static void mat_mul(double* o, const double* A, const double* B) {
    double r[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r[i * 3 + j] = A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j] + A[i * 3 + 2] * B[2 * 3 + j];
    memcpy(o, r, sizeof r);
}
/* Matrix3d.mulTransposeRight(A, B) = A B^T */
static void mat_mul_bt(double* o, const double* A, const double* B) {
    double r[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r[i * 3 + j] = A[i * 3 + 0] * B[j * 3 + 0] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
    memcpy(o, r, sizeof r);
}
/* RigidTransform3D.computeRM0RT: M2 = R M1 R^T (RigidTransform3D.java:224-227) */
static void rm0rt(double* M2, const double* R, const double* M1) {
    double t[9];
    mat_mul(t, R, M1);
    mat_mul_bt(M2, t, R);
}

/* ------------------------------------------------------------------ */
/* World                                                                */
/* ------------------------------------------------------------------ */
typedef struct orc_contact {
    int32_t b1, b2, bv1, bv2, info;
    double pB1[3], nB1[3], t1B1[3], t2B1[3];
    double viol, prev_viol;
    double J[36];
    double lam[3], lamw[3];
    double b[3], D[3];
    int32_t state, is_new;
} orc_contact;

#define ORC_MAX_SCENES 4096
struct orc_world {
    int32_t n;                /* entities */
    am_body_desc* desc;
    am_sphere_node* nodes;
    int32_t n_nodes;
    double* cw;               /* per-node world centre (BVSphere.cW) */
    double *x, *R, *v, *w, *f, *t, *jinv, *dv;
    int32_t* parent;
    int32_t* top; int32_t n_top;
    int32_t *mem_first, *mem_count, *members;
    orc_contact* c; int32_t nc, capc;
    orc_contact* old; int32_t nold, capold;
    /* sorted lookup over old contacts */
    uint64_t* okey; int32_t* oidx; int32_t nokey;
    am_spring* springs; int32_t n_springs;
    /* sub-bodies of composites (RigidBodyGeomComposite.bodies) and their
     * transforms into the composite's frame (RigidBody.transformB2C) */
    am_body_desc* subs; double* sub_b2c; int32_t n_subs;
    am_params p;
    /* per scene of the last main solve that stopped scenes one by one */
    int32_t n_scene_used;
    int32_t scene_used[ORC_MAX_SCENES];
    double scene_last[ORC_MAX_SCENES];
};

static void push_contact_slot(orc_world* w) {
    if (w->nc == w->capc) {
        w->capc = w->capc ? 2 * w->capc : 1024;
        w->c = (orc_contact*)realloc(w->c, (size_t)w->capc * sizeof(orc_contact));
    }
}

void orc_set_params(orc_world* w, const am_params* p) { w->p = *p; }

static void default_params(am_params* p) {
    memset(p, 0, sizeof *p);
    p->iterations = 30; p->iterations_in_collection = 1; p->tolerance = 1e-5; p->omega = 1.0;
    p->feedback_stiffness = 0.5; p->enable_post_stabilization = 0; p->enable_compliance = 1;
    p->compliance = 1e-3; p->restitution = 0.5; p->friction = 0.1; p->sliding_threshold = 0.01;
    p->warm_start = 1; p->use_gravity = 1; p->gravity_amount = 1.0; p->gravity_angle_deg = 90.0;
    p->viscous_linear_decay = 1.0; p->viscous_angular_decay = 1.0;
    p->spring_stiffness_mod = 1.0; p->spring_damping_mod = 1.0;
    p->organize_contacts = 1;
}

orc_world* orc_create(const am_body_desc* bodies, int32_t n, const am_sphere_node* nodes, int32_t n_nodes) {
    return orc_create_ex(bodies, n, nodes, n_nodes, NULL, NULL, 0);
}

orc_world* orc_create_ex(const am_body_desc* bodies, int32_t n, const am_sphere_node* nodes, int32_t n_nodes,
                         const am_body_desc* subs, const double* sub_b2c, int32_t n_subs) {
    orc_world* w = (orc_world*)calloc(1, sizeof *w);
    w->n_subs = n_subs;
    w->subs = (am_body_desc*)malloc((size_t)(n_subs ? n_subs : 1) * sizeof(am_body_desc));
    w->sub_b2c = (double*)malloc((size_t)(n_subs ? n_subs : 1) * 12 * sizeof(double));
    if (n_subs) {
        memcpy(w->subs, subs, (size_t)n_subs * sizeof(am_body_desc));
        memcpy(w->sub_b2c, sub_b2c, (size_t)n_subs * 12 * sizeof(double));
    }
    w->n = n;
    w->desc = (am_body_desc*)malloc((size_t)(n ? n : 1) * sizeof(am_body_desc));
    memcpy(w->desc, bodies, (size_t)n * sizeof(am_body_desc));
    w->n_nodes = n_nodes;
    w->nodes = (am_sphere_node*)malloc((size_t)(n_nodes ? n_nodes : 1) * sizeof(am_sphere_node));
    if (n_nodes) memcpy(w->nodes, nodes, (size_t)n_nodes * sizeof(am_sphere_node));
    w->cw = (double*)calloc((size_t)(n_nodes ? n_nodes : 1) * 3, sizeof(double));
    size_t nn = (size_t)(n ? n : 1);
    w->x = (double*)calloc(nn * 3, 8); w->R = (double*)calloc(nn * 9, 8);
    w->v = (double*)calloc(nn * 3, 8); w->w = (double*)calloc(nn * 3, 8);
    w->f = (double*)calloc(nn * 3, 8); w->t = (double*)calloc(nn * 3, 8);
    w->jinv = (double*)calloc(nn * 9, 8); w->dv = (double*)calloc(nn * 6, 8);
    w->parent = (int32_t*)malloc(nn * sizeof(int32_t));
    w->top = (int32_t*)malloc(nn * sizeof(int32_t));
    w->mem_first = (int32_t*)calloc(nn, sizeof(int32_t));
    w->mem_count = (int32_t*)calloc(nn, sizeof(int32_t));
    w->members = (int32_t*)calloc(nn, sizeof(int32_t));
    for (int32_t i = 0; i < n; i++) {
        w->parent[i] = -1;
        w->top[i] = i;
        w->R[i * 9 + 0] = w->R[i * 9 + 4] = w->R[i * 9 + 8] = 1.0;
        memcpy(&w->jinv[i * 9], w->desc[i].jinv0, 9 * sizeof(double));
    }
    w->n_top = n;
    default_params(&w->p);
    return w;
}

void orc_destroy(orc_world* w) {
    if (!w) return;
    free(w->desc); free(w->nodes); free(w->cw);
    free(w->x); free(w->R); free(w->v); free(w->w); free(w->f); free(w->t); free(w->jinv); free(w->dv);
    free(w->parent); free(w->top); free(w->mem_first); free(w->mem_count); free(w->members);
    free(w->c); free(w->old); free(w->okey); free(w->oidx); free(w->springs); free(w->subs); free(w->sub_b2c);
    free(w);
}

void orc_set_structure(orc_world* w, const int32_t* parent, const int32_t* top, int32_t n_top,
                       const int32_t* mf, const int32_t* mc, const int32_t* members) {
    int32_t nm = 0;
    for (int32_t i = 0; i < w->n; i++) {
        w->parent[i] = parent[i];
        w->mem_first[i] = mf ? mf[i] : 0;
        w->mem_count[i] = mc ? mc[i] : 0;
        if (w->mem_first[i] + w->mem_count[i] > nm) nm = w->mem_first[i] + w->mem_count[i];
    }
    free(w->members);
    w->members = (int32_t*)calloc((size_t)(nm ? nm : 1), sizeof(int32_t));
    if (members && nm) memcpy(w->members, members, (size_t)nm * sizeof(int32_t));
    free(w->top);
    w->top = (int32_t*)malloc((size_t)(n_top ? n_top : 1) * sizeof(int32_t));
    memcpy(w->top, top, (size_t)n_top * sizeof(int32_t));
    w->n_top = n_top;
}

static int pinned(const orc_world* w, int32_t i) { return (w->desc[i].flags & AM_BODY_PINNED) != 0; }
static int sleeping(const orc_world* w, int32_t i) { return (w->desc[i].flags & AM_BODY_SLEEPING) != 0; }

/* RigidBody.updateRotationalInertiaFromTransformation (RigidBody.java:311-321) */
static void update_jinv(orc_world* w, int32_t i) {
    if (pinned(w, i)) return;
    rm0rt(&w->jinv[i * 9], &w->R[i * 9], w->desc[i].jinv0);
}

/* Grows the entity arrays to n (new entities: unused collection slots,
 * pinned, identity rotation).  Collections of a merging run live there. */
void orc_resize(orc_world* w, int32_t n) {
    if (n <= w->n) return;
    size_t o = (size_t)w->n, m = (size_t)n;
#define GROW(ptr, T, k)                                             \
    ptr = (T*)realloc(ptr, m * (k) * sizeof(T));                    \
    memset(ptr + o * (k), 0, (m - o) * (k) * sizeof(T));
    GROW(w->desc, am_body_desc, 1)
    GROW(w->x, double, 3) GROW(w->R, double, 9) GROW(w->v, double, 3) GROW(w->w, double, 3)
    GROW(w->f, double, 3) GROW(w->t, double, 3) GROW(w->jinv, double, 9) GROW(w->dv, double, 6)
    GROW(w->parent, int32_t, 1) GROW(w->mem_first, int32_t, 1) GROW(w->mem_count, int32_t, 1)
#undef GROW
    for (size_t i = o; i < m; i++) {
        w->desc[i].geom = AM_GEOM_COLLECTION;
        w->desc[i].flags = AM_BODY_PINNED;
        w->R[i * 9 + 0] = w->R[i * 9 + 4] = w->R[i * 9 + 8] = 1.0;
        w->parent[i] = -1;
    }
    w->n = n;
}
void orc_set_entity(orc_world* w, int32_t e, const am_body_desc* d) { w->desc[e] = *d; }
void orc_get_entity(const orc_world* w, int32_t e, am_body_desc* d) { *d = w->desc[e]; }
int32_t orc_entities(const orc_world* w) { return w->n; }

void orc_set_state(orc_world* w, const double* x, const double* R, const double* v, const double* om) {
    for (int32_t i = 0; i < w->n; i++) {
        if (x) memcpy(&w->x[i * 3], &x[i * 3], 24);
        if (R) memcpy(&w->R[i * 9], &R[i * 9], 72);
        if (v) memcpy(&w->v[i * 3], &v[i * 3], 24);
        if (om) memcpy(&w->w[i * 3], &om[i * 3], 24);
        if (pinned(w, i)) memset(&w->jinv[i * 9], 0, 72);
        else update_jinv(w, i);
    }
}
void orc_get_state(const orc_world* w, double* x, double* R, double* v, double* om) {
    size_t n = (size_t)w->n;
    if (x) memcpy(x, w->x, n * 24);
    if (R) memcpy(R, w->R, n * 72);
    if (v) memcpy(v, w->v, n * 24);
    if (om) memcpy(om, w->w, n * 24);
}
void orc_set_jinv(orc_world* w, int32_t e, const double* jinv) { memcpy(&w->jinv[e * 9], jinv, 72); }
void orc_get_jinv(const orc_world* w, double* jinv) { memcpy(jinv, w->jinv, (size_t)w->n * 72); }
void orc_set_forces(orc_world* w, const double* f, const double* t) {
    if (f) memcpy(w->f, f, (size_t)w->n * 24);
    if (t) memcpy(w->t, t, (size_t)w->n * 24);
}
void orc_get_forces(const orc_world* w, double* f, double* t) {
    if (f) memcpy(f, w->f, (size_t)w->n * 24);
    if (t) memcpy(t, w->t, (size_t)w->n * 24);
}
void orc_get_deltav(const orc_world* w, double* dv) { memcpy(dv, w->dv, (size_t)w->n * 48); }
void orc_set_deltav(orc_world* w, const double* dv) { memcpy(w->dv, dv, (size_t)w->n * 48); }

/* RigidBody.clear (RigidBody.java:226-230) for every entity */
void orc_clear_bodies(orc_world* w) {
    memset(w->f, 0, (size_t)w->n * 24);
    memset(w->t, 0, (size_t)w->n * 24);
    memset(w->dv, 0, (size_t)w->n * 48);
}

/* RigidBodySystem.applyGravityForce (RigidBodySystem.java:276-290): the
 * force of every top-level entity and of every collection member. */
void orc_apply_gravity(orc_world* w) {
    if (!w->p.use_gravity) return;
    double theta = w->p.gravity_angle_deg / 180.0 * M_PI;
    double g = w->p.gravity_amount;
    double dir[3] = {g * cos(theta), g * sin(theta), 0.0};
    for (int32_t k = 0; k < w->n_top; k++) {
        int32_t i = w->top[k];
        double m = -w->desc[i].mass;
        for (int a = 0; a < 3; a++) w->f[i * 3 + a] = w->f[i * 3 + a] + m * dir[a];
        if (w->desc[i].geom == AM_GEOM_COLLECTION) {
            for (int32_t q = 0; q < w->mem_count[i]; q++) {
                int32_t b = w->members[w->mem_first[i] + q];
                double mb = -w->desc[b].mass;
                for (int a = 0; a < 3; a++) w->f[b * 3 + a] = w->f[b * 3 + a] + mb * dir[a];
            }
        }
    }
}

/* RigidBodySystem.springs (filled by XMLParser.setCommonAttributes :596-619) */
void orc_set_springs(orc_world* w, const am_spring* springs, int32_t n) {
    free(w->springs);
    w->springs = (am_spring*)malloc((size_t)(n ? n : 1) * sizeof(am_spring));
    if (n) memcpy(w->springs, springs, (size_t)n * sizeof(am_spring));
    w->n_springs = n;
}

/* RigidBody.applyForceW (RigidBody.java:335-340) */
static void apply_force_w(orc_world* w, int32_t e, const double* p, const double* F) {
    double r[3], tq[3];
    for (int a = 0; a < 3; a++) w->f[e * 3 + a] = w->f[e * 3 + a] + F[a];
    for (int a = 0; a < 3; a++) r[a] = p[a] - w->x[e * 3 + a];
    cross3(tq, r, F);
    for (int a = 0; a < 3; a++) w->t[e * 3 + a] = w->t[e * 3 + a] + tq[a];
}
/* RigidBody.getSpatialVelocity (RigidBody.java:458-462) */
static void point_velocity(const orc_world* w, int32_t e, const double* p, double* out) {
    double r[3];
    for (int a = 0; a < 3; a++) r[a] = p[a] - w->x[e * 3 + a];
    cross3(out, &w->w[e * 3], r);
    for (int a = 0; a < 3; a++) out[a] = out[a] + w->v[e * 3 + a];
}
/* body point to world: RigidTransform3D.transform (RigidTransform3D.java:107-110) */
static void body_to_world(const orc_world* w, int32_t e, const double* pb, double* out) {
    mat_vec(out, &w->R[e * 9], pb);
    for (int a = 0; a < 3; a++) out[a] = out[a] + w->x[e * 3 + a];
}

/* RigidBodySystem.applySpringForces (RigidBodySystem.java:309-322) with
 * Spring.applyOneBody / applyTwoBodies (Spring.java:165-214), springs in order */
void orc_apply_springs(orc_world* w) {
    const double ks = w->p.spring_stiffness_mod, ds = w->p.spring_damping_mod;
    for (int32_t i = 0; i < w->n_springs; i++) {
        const am_spring* sp = &w->springs[i];
        const int32_t b1 = sp->body1;
        double p1w[3], disp[3], F[3];
        body_to_world(w, b1, sp->pb1, p1w);
        if (sp->type != AM_SPRING_BODYBODY) {
            for (int a = 0; a < 3; a++) disp[a] = sp->pw[a] - p1w[a];
            double len = len3(disp);
            if (len < 1e-3) continue;
            double vel[3];
            point_velocity(w, b1, p1w, vel);
            double scl = -(sp->k * ks * (len - sp->l0 * sp->ls) - sp->d * ds * (dot3(vel, disp) / len)) / len;
            for (int a = 0; a < 3; a++) F[a] = -scl * disp[a];
            apply_force_w(w, b1, p1w, F);
            if (w->parent[b1] >= 0) apply_force_w(w, w->parent[b1], p1w, F);
        } else {
            const int32_t b2 = sp->body2;
            double p2w[3], v1[3], v2[3], rel[3];
            body_to_world(w, b2, sp->pb2, p2w);
            for (int a = 0; a < 3; a++) disp[a] = p2w[a] - p1w[a];
            double dist = len3(disp);
            if (dist < 1e-3) continue;
            double inv = 1. / dist;
            for (int a = 0; a < 3; a++) disp[a] = inv * disp[a];
            point_velocity(w, b2, p2w, v2);
            point_velocity(w, b1, p1w, v1);
            for (int a = 0; a < 3; a++) rel[a] = v2[a] - v1[a];
            double fi = sp->k * ks * (dist - sp->l0 * sp->ls) + sp->d * ds * dot3(rel, disp);
            for (int a = 0; a < 3; a++) F[a] = fi * disp[a];
            apply_force_w(w, b1, p1w, F);
            if (w->parent[b1] >= 0) apply_force_w(w, w->parent[b1], p1w, F);
            for (int a = 0; a < 3; a++) F[a] = -fi * disp[a];
            apply_force_w(w, b2, p2w, F);
            if (w->parent[b2] >= 0) apply_force_w(w, w->parent[b2], p2w, F);
        }
    }
}

/* solve body of a contact body: parent if in a collection and not computing
 * in collection (PGS.java:108-109) */
static int32_t solve_body(const orc_world* w, int32_t b, int in_collection) {
    return (w->parent[b] >= 0 && !in_collection) ? w->parent[b] : b;
}

/* Contact.computeJacobian(boolean, contactW, normalW, t1W, t2W) (Contact.java:235-255) */
static void compute_jacobian_w(const orc_world* w, orc_contact* c, int in_coll, const double* pW,
                               const double* nW, const double* t1, const double* t2) {
    int32_t a = solve_body(w, c->b1, in_coll), b = solve_body(w, c->b2, in_coll);
    double r1[3] = {pW[0] - w->x[a * 3 + 0], pW[1] - w->x[a * 3 + 1], pW[2] - w->x[a * 3 + 2]};
    double r2[3] = {pW[0] - w->x[b * 3 + 0], pW[1] - w->x[b * 3 + 1], pW[2] - w->x[b * 3 + 2]};
    const double* dirs[3] = {nW, t1, t2};
    for (int row = 0; row < 3; row++) {
        const double* d = dirs[row];
        double* J = &c->J[row * 12];
        J[0] = -1.0 * d[0]; J[1] = -1.0 * d[1]; J[2] = -1.0 * d[2];
        cross3(&J[3], d, r1);
        J[6] = d[0]; J[7] = d[1]; J[8] = d[2];
        cross3(&J[9], r2, d);
    }
}

/* Contact.computeJacobian(boolean) (Contact.java:265-271): rebuild the world
 * frame from the body-1 frame. */
static void compute_jacobian(const orc_world* w, orc_contact* c, int in_coll) {
    const double* R = &w->R[c->b1 * 9];
    const double* x = &w->x[c->b1 * 3];
    double pW[3], nW[3], t1[3], t2[3];
    xform_point(pW, R, x, c->pB1);
    mat_vec(nW, R, c->nB1);
    mat_vec(t1, R, c->t1B1);
    mat_vec(t2, R, c->t2B1);
    compute_jacobian_w(w, c, in_coll, pW, nW, t1, t2);
}

/* Contact.set (Contact.java:159-233) */
static void contact_set(orc_world* w, orc_contact* c, int32_t b1, int32_t b2, const double* pW,
                        const double* nW, int32_t bv1, int32_t bv2, int32_t info, double viol) {
    c->state = AM_STATE_CLEAR;
    c->b1 = b1; c->b2 = b2; c->info = info; c->bv1 = bv1; c->bv2 = bv2;
    c->viol = viol; c->prev_viol = 0.0;
    c->lam[0] = c->lam[1] = c->lam[2] = 0.0;
    c->lamw[0] = c->lamw[1] = c->lamw[2] = 0.0;
    double p[3] = {pW[0], pW[1], pW[2]};
    double n[3] = {nW[0], nW[1], nW[2]};
    double t1[3], t2[3];
    double anx = fabs(n[0]), any = fabs(n[1]), anz = fabs(n[2]);
    if (anx < any && anx < anz) { t1[0] = 1; t1[1] = 0; t1[2] = 0; }
    else if (any < anz) { t1[0] = 0; t1[1] = 1; t1[2] = 0; }
    else { t1[0] = 0; t1[1] = 0; t1[2] = 1; }
    cross3(t2, n, t1);
    normalize3(t2);
    cross3(t1, t2, n);
    compute_jacobian_w(w, c, 0, p, n, t1, t2);
    const double* R = &w->R[b1 * 9];
    const double* x = &w->x[b1 * 3];
    inv_xform_point(c->pB1, R, x, p);
    mat_tvec(c->nB1, R, n);
    mat_tvec(c->t1B1, R, t1);
    mat_tvec(c->t2B1, R, t2);
    c->b[0] = c->b[1] = c->b[2] = 0.0;
    c->D[0] = c->D[1] = c->D[2] = 0.0;
    c->is_new = 1;
}

static void emit(orc_world* w, int32_t b1, int32_t b2, const double* pW, const double* nW,
                 int32_t bv1, int32_t bv2, int32_t info, double viol) {
    push_contact_slot(w);
    contact_set(w, &w->c[w->nc], b1, b2, pW, nW, bv1, bv2, info, viol);
    w->nc++;
}

/* What a primitive collides: a body's geometry at its pose, or a sub-body of
 * a composite at the pose RigidBodyGeomComposite.updateBodyPositionsFromParent
 * gives it (RigidTransform3D.mult, RigidTransform3D.java:193-202).  Contacts
 * go to the physical body with the sub-body in the bv id (Contact.set keeps it
 * as csb, Contact.java:166-177). */
typedef struct {
    const am_body_desc* d;
    const double* x;  /* the body's state, or the caller's buffer for a sub-body */
    const double* R;
    int32_t body, sub;
} oshape;
static oshape body_shape(const orc_world* w, int32_t b) {
    oshape s;
    s.d = &w->desc[b];
    s.x = &w->x[b * 3];
    s.R = &w->R[b * 9];
    s.body = b; s.sub = -1;
    return s;
}
static void sub_pose(const orc_world* w, int32_t b, int32_t q, double* x, double* R) {
    const double* T = &w->sub_b2c[(size_t)(w->desc[b].tree_first + q) * 12];
    xform_point(x, &w->R[b * 9], &w->x[b * 3], T + 9);
    mat_mul(R, &w->R[b * 9], T);
}
/* buf: 12 doubles that hold the sub-body's pose while the shape is used */
static oshape sub_shape(const orc_world* w, int32_t b, int32_t q, double* buf) {
    oshape s;
    s.d = &w->subs[w->desc[b].tree_first + q];
    sub_pose(w, b, q, buf, buf + 3);
    s.x = buf;
    s.R = buf + 3;
    s.body = b; s.sub = q;
    return s;
}
static int32_t enc_bv(const oshape* s, int32_t bv) { return s->sub < 0 ? bv : AM_BV_SUB(s->sub, bv); }

/* ------------------------------------------------------------------ */
/* Box-plane: BoxPlane.dBoxPlane (collision/BoxPlane.java:27-51)        */
/* ------------------------------------------------------------------ */
static int32_t box_plane_s(orc_world* w, const oshape* box, const oshape* plane) {
    const am_body_desc* bd = box->d;
    const am_body_desc* pd = plane->d;
    const double* n = pd->plane_n;
    const double* x = box->x;
    const double* R = box->R;
    if (dot3(x, n) + pd->plane_d > bd->radius) return 0;
    double h[3] = {0.5 * bd->size[0], 0.5 * bd->size[1], 0.5 * bd->size[2]};
    int32_t cnt = 0;
    for (int k = 0; k < 8; k++) {
        double q[3] = {(k & 4) ? -h[0] : h[0], (k & 2) ? -h[1] : h[1], (k & 1) ? -h[2] : h[2]};
        double qw[3];
        xform_point(qw, R, x, q);
        double s = dot3(qw, n) + pd->plane_d;
        if (s < 0) {
            emit(w, plane->body, box->body, qw, n, enc_bv(plane, AM_BV_NONE), enc_bv(box, AM_BV_NONE), k, s);
            cnt++;
        }
    }
    return cnt;
}
int32_t orc_box_plane(orc_world* w, int32_t box, int32_t plane) {
    oshape a = body_shape(w, box), b = body_shape(w, plane);
    return box_plane_s(w, &a, &b);
}

/* ------------------------------------------------------------------ */
/* Box-box: BoxBox.dBoxBox (collision/BoxBox.java:433-859)              */
/* ------------------------------------------------------------------ */
static const double BB_FUDGE = 1.05, BB_TINY = 1e-5, BB_LCA_EPS = 0.0001;

/* BoxBox.intersectRectQuad (BoxBox.java:140-209) */
static int intersect_rect_quad(const double h[2], double p[8], double ret[16]) {
    int nq = 4, nr = 0;
    double buffer[16];
    double* q = p;
    double* r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            double* pq = q;
            double* pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if ((double)sign * pq[dir] < h[dir]) {
                    pr[0] = pq[0]; pr[1] = pq[1]; pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                double* nextq = (i > 1) ? pq + 2 : q;
                if (((double)sign * pq[dir] < h[dir]) ^ ((double)sign * nextq[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] + (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) *
                                                    ((double)sign * h[dir] - pq[dir]);
                    pr[dir] = (double)sign * h[dir];
                    pr += 2; nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                pq += 2;
            }
            q = r;
            r = (q == ret) ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret) memcpy(ret, q, (size_t)nr * 2 * sizeof(double));
    return nr;
}

/* Rij = column i of A dotted with column j of B (BoxBox.dDOT44) */
static double col_dot(const double* A, int i, const double* B, int j) {
    return A[0 * 3 + i] * B[0 * 3 + j] + A[1 * 3 + i] * B[1 * 3 + j] + A[2 * 3 + i] * B[2 * 3 + j];
}
/* v dotted with column j of M (BoxBox.dDOT14 / dDOT41) */
static double vec_col(const double* v, const double* M, int j) {
    return v[0] * M[0 * 3 + j] + v[1] * M[1 * 3 + j] + v[2] * M[2 * 3 + j];
}

typedef struct { double s; int code; int invert; const double* nR; int ncol; double nC[3]; } bb_tst;

static int bb_tst1(double e1, double e2, const double* M, int col, int cc, bb_tst* t) {
    double s2 = fabs(e1) - e2;
    if (s2 > 0) return 0;
    if (s2 > t->s + BB_TINY) { t->s = s2; t->nR = M; t->ncol = col; t->invert = (e1 < 0); t->code = cc; }
    return 1;
}
static int bb_tst2(double e1, double e2, double n1, double n2, double n3, int cc, bb_tst* t) {
    double s2 = fabs(e1) - e2;
    if (s2 > 0) return 0;
    double l = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    if (l > 0) {
        s2 /= l;
        if (s2 * BB_FUDGE > t->s) {
            t->s = s2; t->nR = NULL; t->ncol = 0;
            t->nC[0] = n1 / l; t->nC[1] = n2 / l; t->nC[2] = n3 / l;
            t->invert = (e1 < 0); t->code = cc;
        }
    }
    return 1;
}

static int32_t box_box_s(orc_world* w, const oshape* S1, const oshape* S2) {
    const double* p1 = S1->x;
    const double* R1 = S1->R;
    const double* p2 = S2->x;
    const double* R2 = S2->R;
    const double* side1 = S1->d->size;
    const double* side2 = S2->d->size;
    const int32_t b1 = S1->body, b2 = S2->body;
    const int32_t bv1 = enc_bv(S1, AM_BV_NONE), bv2 = enc_bv(S2, AM_BV_NONE);
    int info = 0;
    double p[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    if (len3(p) > S1->d->radius + S2->d->radius) return 0;
    double pp[3];
    mat_tvec(pp, R1, p);
    double A[3] = {0.5 * side1[0], 0.5 * side1[1], 0.5 * side1[2]};
    double B[3] = {0.5 * side2[0], 0.5 * side2[1], 0.5 * side2[2]};
    double R11 = col_dot(R1, 0, R2, 0), R12 = col_dot(R1, 0, R2, 1), R13 = col_dot(R1, 0, R2, 2);
    double R21 = col_dot(R1, 1, R2, 0), R22 = col_dot(R1, 1, R2, 1), R23 = col_dot(R1, 1, R2, 2);
    double R31 = col_dot(R1, 2, R2, 0), R32 = col_dot(R1, 2, R2, 1), R33 = col_dot(R1, 2, R2, 2);
    double Q11 = fabs(R11), Q12 = fabs(R12), Q13 = fabs(R13);
    double Q21 = fabs(R21), Q22 = fabs(R22), Q23 = fabs(R23);
    double Q31 = fabs(R31), Q32 = fabs(R32), Q33 = fabs(R33);
    bb_tst t;
    t.s = -INFINITY; t.invert = 0; t.code = 0; t.nR = NULL; t.ncol = 0;
    t.nC[0] = t.nC[1] = t.nC[2] = 0.0;
    if (!bb_tst1(pp[0], (A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13), R1, 0, 1, &t)) return 0;
    if (!bb_tst1(pp[1], (A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23), R1, 1, 2, &t)) return 0;
    if (!bb_tst1(pp[2], (A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33), R1, 2, 3, &t)) return 0;
    if (!bb_tst1(vec_col(p, R2, 0), (A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0]), R2, 0, 4, &t)) return 0;
    if (!bb_tst1(vec_col(p, R2, 1), (A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1]), R2, 1, 5, &t)) return 0;
    if (!bb_tst1(vec_col(p, R2, 2), (A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2]), R2, 2, 6, &t)) return 0;
    if (!bb_tst2(pp[2] * R21 - pp[1] * R31, (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0, -R31, R21, 7, &t)) return 0;
    if (!bb_tst2(pp[2] * R22 - pp[1] * R32, (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0, -R32, R22, 8, &t)) return 0;
    if (!bb_tst2(pp[2] * R23 - pp[1] * R33, (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0, -R33, R23, 9, &t)) return 0;
    if (!bb_tst2(pp[0] * R31 - pp[2] * R11, (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), R31, 0, -R11, 10, &t)) return 0;
    if (!bb_tst2(pp[0] * R32 - pp[2] * R12, (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), R32, 0, -R12, 11, &t)) return 0;
    if (!bb_tst2(pp[0] * R33 - pp[2] * R13, (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), R33, 0, -R13, 12, &t)) return 0;
    if (!bb_tst2(pp[1] * R11 - pp[0] * R21, (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -R21, R11, 0, 13, &t)) return 0;
    if (!bb_tst2(pp[1] * R12 - pp[0] * R22, (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -R22, R12, 0, 14, &t)) return 0;
    if (!bb_tst2(pp[1] * R13 - pp[0] * R23, (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -R23, R13, 0, 15, &t)) return 0;
    if (t.code == 0) return 0;

    double normal[3];
    if (t.nR != NULL) {
        normal[0] = t.nR[0 * 3 + t.ncol]; normal[1] = t.nR[1 * 3 + t.ncol]; normal[2] = t.nR[2 * 3 + t.ncol];
    } else {
        mat_vec(normal, R1, t.nC);
    }
    if (t.invert) { normal[0] *= -1.0; normal[1] *= -1.0; normal[2] *= -1.0; }
    double depth = -t.s;

    if (t.code > 6) {
        /* edge-edge (BoxBox.java:600-651) */
        double pa[3] = {p1[0], p1[1], p1[2]};
        for (int j = 0; j < 3; j++) {
            double sign = (vec_col(normal, R1, j) > 0) ? 1.0 : -1.0;
            pa[0] += sign * A[j] * R1[0 * 3 + j];
            pa[1] += sign * A[j] * R1[1 * 3 + j];
            pa[2] += sign * A[j] * R1[2 * 3 + j];
        }
        double pb[3] = {p2[0], p2[1], p2[2]};
        for (int j = 0; j < 3; j++) {
            double sign = (vec_col(normal, R2, j) > 0) ? -1.0 : 1.0;
            pb[0] += sign * B[j] * R2[0 * 3 + j];
            pb[1] += sign * B[j] * R2[1 * 3 + j];
            pb[2] += sign * B[j] * R2[2 * 3 + j];
        }
        int ca = (t.code - 7) / 3, cb = (t.code - 7) % 3;
        double ua[3] = {R1[0 * 3 + ca], R1[1 * 3 + ca], R1[2 * 3 + ca]};
        double ub[3] = {R2[0 * 3 + cb], R2[1 * 3 + cb], R2[2 * 3 + cb]};
        /* dLineClosestApproach (BoxBox.java:366-383) */
        double pl[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        double uaub = dot3(ua, ub);
        double q1 = dot3(ua, pl);
        double q2 = -dot3(ub, pl);
        double d = 1 - uaub * uaub;
        double alpha, beta;
        if (d <= BB_LCA_EPS) { alpha = 0; beta = 0; }
        else { d = 1.0 / d; alpha = (q1 + uaub * q2) * d; beta = (uaub * q1 + q2) * d; }
        for (int i = 0; i < 3; i++) pa[i] = alpha * ua[i] + pa[i];
        for (int i = 0; i < 3; i++) pb[i] = beta * ub[i] + pb[i];
        double pos[3];
        for (int i = 0; i < 3; i++) pos[i] = (pa[i] + pb[i]) * 0.5;
        emit(w, b1, b2, pos, normal, bv1, bv2, info++, -depth);
        return 1;
    }

    /* face-something (BoxBox.java:656-858) */
    const double *Ra, *Rb, *pa, *pb, *Sa, *Sb;
    if (t.code <= 3) { Ra = R1; Rb = R2; pa = p1; pb = p2; Sa = A; Sb = B; }
    else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }
    double normal2[3], nr[3], anr[3];
    if (t.code <= 3) { normal2[0] = normal[0]; normal2[1] = normal[1]; normal2[2] = normal[2]; }
    else { normal2[0] = -1.0 * normal[0]; normal2[1] = -1.0 * normal[1]; normal2[2] = -1.0 * normal[2]; }
    mat_tvec(nr, Rb, normal2);
    anr[0] = fabs(nr[0]); anr[1] = fabs(nr[1]); anr[2] = fabs(nr[2]);
    int lanr, a1, a2;
    if (anr[1] > anr[0]) {
        if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
        if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    }
    double center[3];
    if (nr[lanr] < 0) {
        for (int i = 0; i < 3; i++) center[i] = pb[i] - pa[i] + Sb[lanr] * Rb[i * 3 + lanr];
    } else {
        for (int i = 0; i < 3; i++) center[i] = pb[i] - pa[i] - Sb[lanr] * Rb[i * 3 + lanr];
    }
    int codeN, code1, code2;
    if (t.code <= 3) codeN = t.code - 1; else codeN = t.code - 4;
    if (codeN == 0) { code1 = 1; code2 = 2; }
    else if (codeN == 1) { code1 = 0; code2 = 2; }
    else { code1 = 0; code2 = 1; }
    double c1 = vec_col(center, Ra, code1);
    double c2 = vec_col(center, Ra, code2);
    double m11 = col_dot(Ra, code1, Rb, a1);
    double m12 = col_dot(Ra, code1, Rb, a2);
    double m21 = col_dot(Ra, code2, Rb, a1);
    double m22 = col_dot(Ra, code2, Rb, a2);
    double quad[8], ret[16], rect[2];
    {
        double k1 = m11 * Sb[a1];
        double k2 = m21 * Sb[a1];
        double k3 = m12 * Sb[a2];
        double k4 = m22 * Sb[a2];
        quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
        quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
        quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
        quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    rect[0] = Sa[code1];
    rect[1] = Sa[code2];
    int n = intersect_rect_quad(rect, quad, ret);
    if (n < 1) return 0;
    double det1 = 1.0 / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    int cnum = 0;
    double point[24], dep[8];
    for (int j = 0; j < n; j++) {
        double k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        double k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        for (int i = 0; i < 3; i++) point[cnum * 3 + i] = center[i] + k1 * Rb[i * 3 + a1] + k2 * Rb[i * 3 + a2];
        dep[cnum] = Sa[codeN] - (normal2[0] * point[cnum * 3 + 0] + normal2[1] * point[cnum * 3 + 1] +
                                 normal2[2] * point[cnum * 3 + 2]);
        if (dep[cnum] >= 0) {
            ret[cnum * 2] = ret[j * 2];
            ret[cnum * 2 + 1] = ret[j * 2 + 1];
            cnum++;
        }
    }
    if (cnum < 1) return 0;
    for (int j = 0; j < cnum; j++) {
        double pos[3];
        for (int i = 0; i < 3; i++) pos[i] = point[j * 3 + i] + pa[i];
        emit(w, b1, b2, pos, normal, bv1, bv2, info++, -dep[j]);
    }
    return cnum;
}
int32_t orc_box_box(orc_world* w, int32_t b1, int32_t b2) {
    oshape a = body_shape(w, b1), b = body_shape(w, b2);
    return box_box_s(w, &a, &b);
}

/* ------------------------------------------------------------------ */
/* Box-sphere: BoxSphere.java                                           */
/* ------------------------------------------------------------------ */
typedef struct { double pos[3], normal[3], depth; } bs_res;

static double dist3(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* BoxSphere.test1 (BoxSphere.java:242-253) */
static void bs_test1(const double* q, const double* cB, double r, bs_res* t) {
    double s = dist3(cB, q) - r;
    if (s > 0) return;
    if (s < t->depth) {
        t->pos[0] = q[0]; t->pos[1] = q[1]; t->pos[2] = q[2];
        t->normal[0] = cB[0] - q[0]; t->normal[1] = cB[1] - q[1]; t->normal[2] = cB[2] - q[2];
        normalize3(t->normal);
        t->depth = s;
    }
}
/* BoxSphere.test2 (BoxSphere.java:266-275) */
static void bs_test2(double depth, double px, double py, double pz, double nx, double ny, double nz, bs_res* t) {
    if (depth > 0) return;
    if (depth < t->depth) {
        t->pos[0] = px; t->pos[1] = py; t->pos[2] = pz;
        t->normal[0] = nx; t->normal[1] = ny; t->normal[2] = nz;
        t->depth = depth;
    }
}

/* BoxSphere.dBoxSphere (BoxSphere.java:58-144) */
static int32_t box_sphere_s(orc_world* w, const oshape* box, const oshape* other, const double* c, double r) {
    const am_body_desc* bd = box->d;
    const double* R = box->R;
    const double* x = box->x;
    double p[3] = {0.5 * bd->size[0], 0.5 * bd->size[1], 0.5 * bd->size[2]};
    double cB[3];
    inv_xform_point(cB, R, x, c);
    bs_res t;
    t.depth = 1;
    t.pos[0] = t.pos[1] = t.pos[2] = 0; t.normal[0] = t.normal[1] = t.normal[2] = 0;
    for (int k = 0; k < 8; k++) {
        double q[3] = {(k & 4) ? -p[0] : p[0], (k & 2) ? -p[1] : p[1], (k & 1) ? -p[2] : p[2]};
        bs_test1(q, cB, r, &t);
    }
    double s, v[3];
#define EDGE(vx, vy, vz, px_, py_, pz_)                       \
    {                                                         \
        v[0] = (vx); v[1] = (vy); v[2] = (vz);                \
        s = len3(v); s -= r;                                  \
        bs_test2(s, px_, py_, pz_, v[0], v[1], v[2], &t);     \
    }
    if (-p[0] <= cB[0] && cB[0] <= p[0]) {
        if (cB[1] >= p[1] && cB[2] >= p[2]) EDGE(0, cB[1] - p[1], cB[2] - p[2], cB[0], p[1], p[2])
        else if (cB[1] <= -p[1] && cB[2] >= p[2]) EDGE(0, cB[1] + p[1], cB[2] - p[2], cB[0], -p[1], p[2])
        else if (cB[1] >= p[1] && cB[2] <= -p[2]) EDGE(0, cB[1] - p[1], cB[2] + p[2], cB[0], p[1], -p[2])
        else if (cB[1] <= -p[1] && cB[2] <= -p[2]) EDGE(0, cB[1] + p[1], cB[2] + p[2], cB[0], -p[1], -p[2])
    }
    if (-p[1] <= cB[1] && cB[1] <= p[1]) {
        if (cB[0] >= p[0] && cB[2] >= p[2]) EDGE(cB[0] - p[0], 0, cB[2] - p[2], p[0], cB[1], p[2])
        else if (cB[0] <= -p[0] && cB[2] >= p[2]) EDGE(cB[0] + p[0], 0, cB[2] - p[2], -p[0], cB[1], p[2])
        else if (cB[0] >= p[0] && cB[2] <= -p[2]) EDGE(cB[0] - p[0], 0, cB[2] + p[2], p[0], cB[1], -p[2])
        else if (cB[0] <= -p[0] && cB[2] <= -p[2]) EDGE(cB[0] + p[0], 0, cB[2] + p[2], -p[0], cB[1], -p[2])
    }
    if (-p[2] <= cB[2] && cB[2] <= p[2]) {
        if (cB[0] >= p[0] && cB[1] >= p[1]) EDGE(cB[0] - p[0], cB[1] - p[1], 0, p[0], p[1], cB[2])
        else if (cB[0] <= -p[0] && cB[1] >= p[1]) EDGE(cB[0] + p[0], cB[1] - p[1], 0, -p[0], p[1], cB[2])
        else if (cB[0] >= p[0] && cB[1] <= -p[1]) EDGE(cB[0] - p[0], cB[1] + p[1], 0, p[0], -p[1], cB[2])
        else if (cB[0] <= -p[0] && cB[1] <= -p[1]) EDGE(cB[0] + p[0], cB[1] + p[1], 0, -p[0], -p[1], cB[2])
    }
#undef EDGE
    if (-p[0] <= cB[0] && cB[0] <= p[0] && -p[1] <= cB[1] && cB[1] <= p[1]) {
        if (cB[2] > 0) { s = cB[2] - p[2] - r; bs_test2(s, cB[0], cB[1], p[2], 0, 0, 1, &t); }
        else { s = -p[2] - cB[2] - r; bs_test2(s, cB[0], cB[1], -p[2], 0, 0, -1, &t); }
    }
    if (-p[0] <= cB[0] && cB[0] <= p[0] && -p[2] <= cB[2] && cB[2] <= p[2]) {
        if (cB[1] > 0) { s = cB[1] - p[1] - r; bs_test2(s, cB[0], p[1], cB[2], 0, 1, 0, &t); }
        else { s = -p[1] - cB[1] - r; bs_test2(s, cB[0], -p[1], cB[2], 0, -1, 0, &t); }
    }
    if (-p[1] <= cB[1] && cB[1] <= p[1] && -p[2] <= cB[2] && cB[2] <= p[2]) {
        if (cB[0] > 0) { s = cB[0] - p[0] - r; bs_test2(s, p[0], cB[1], cB[2], 1, 0, 0, &t); }
        else { s = -p[0] - cB[0] - r; bs_test2(s, -p[0], cB[1], cB[2], -1, 0, 0, &t); }
    }
    if (t.depth != 1) {
        double pw[3], nw[3];
        xform_point(pw, R, x, t.pos);
        mat_vec(nw, R, t.normal);
        normalize3(nw);
        emit(w, box->body, other->body, pw, nw, enc_bv(box, AM_BV_NONE), enc_bv(other, AM_BV_NONE), 0, t.depth);
        return 1;
    }
    return 0;
}
int32_t orc_box_sphere(orc_world* w, int32_t box, int32_t other, const double* c, double r) {
    oshape a = body_shape(w, box), b = body_shape(w, other);
    return box_sphere_s(w, &a, &b, c, r);
}

/* BoxSphere.dBoxSphereTest (BoxSphere.java:156-233) */
static int32_t box_sphere_test_s(const oshape* box, const double* c, double r) {
    const am_body_desc* bd = box->d;
    const double* R = box->R;
    const double* x = box->x;
    if (dist3(c, x) > bd->radius + r) return 0;
    double p[3] = {0.5 * bd->size[0], 0.5 * bd->size[1], 0.5 * bd->size[2]};
    double cB[3];
    inv_xform_point(cB, R, x, c);
    if (cB[0] - p[0] > r) return 0;
    if (-p[0] - cB[0] > r) return 0;
    if (cB[1] - p[1] > r) return 0;
    if (-p[1] - cB[1] > r) return 0;
    if (cB[2] - p[2] > r) return 0;
    if (-p[2] - cB[2] > r) return 0;
    for (int k = 0; k < 8; k++) {
        double q[3] = {(k & 4) ? -p[0] : p[0], (k & 2) ? -p[1] : p[1], (k & 1) ? -p[2] : p[2]};
        if (dist3(cB, q) - r < 0) return 1;
    }
    double v[3];
#define ETEST(vx, vy, vz) { v[0] = (vx); v[1] = (vy); v[2] = (vz); if (len3(v) - r < 0) return 1; }
    if (-p[0] <= cB[0] && cB[0] <= p[0]) {
        if (cB[1] >= p[1] && cB[2] >= p[2]) ETEST(0, cB[1] - p[1], cB[2] - p[2])
        else if (cB[1] <= -p[1] && cB[2] >= p[2]) ETEST(0, cB[1] + p[1], cB[2] - p[2])
        else if (cB[1] >= p[1] && cB[2] <= -p[2]) ETEST(0, cB[1] - p[1], cB[2] + p[2])
        else if (cB[1] <= -p[1] && cB[2] <= -p[2]) ETEST(0, cB[1] + p[1], cB[2] + p[2])
    }
    if (-p[1] <= cB[1] && cB[1] <= p[1]) {
        if (cB[0] >= p[0] && cB[2] >= p[2]) ETEST(cB[0] - p[0], 0, cB[2] - p[2])
        else if (cB[0] <= -p[0] && cB[2] >= p[2]) ETEST(cB[0] + p[0], 0, cB[2] - p[2])
        else if (cB[0] >= p[0] && cB[2] <= -p[2]) ETEST(cB[0] - p[0], 0, cB[2] + p[2])
        else if (cB[0] <= -p[0] && cB[2] <= -p[2]) ETEST(cB[0] + p[0], 0, cB[2] + p[2])
    }
    if (-p[2] < cB[2] && cB[2] < p[2]) {
        if (cB[0] >= p[0] && cB[1] >= p[1]) ETEST(cB[0] - p[0], cB[1] - p[1], 0)
        else if (cB[0] <= -p[0] && cB[1] >= p[1]) ETEST(cB[0] + p[0], cB[1] - p[1], 0)
        else if (cB[0] >= p[0] && cB[1] <= -p[1]) ETEST(cB[0] - p[0], cB[1] + p[1], 0)
        else if (cB[0] <= -p[0] && cB[1] <= -p[1]) ETEST(cB[0] + p[0], cB[1] + p[1], 0)
    }
#undef ETEST
    if (-p[0] <= cB[0] && cB[0] <= p[0] && -p[1] <= cB[1] && cB[1] <= p[1]) {
        if (cB[2] > 0) { if (cB[2] - p[2] - r < 0) return 1; }
        else { if (-p[2] - cB[2] - r < 0) return 1; }
    }
    if (-p[0] <= cB[0] && cB[0] <= p[0] && -p[2] <= cB[2] && cB[2] <= p[2]) {
        if (cB[1] > 0) { if (cB[1] - p[1] - r < 0) return 1; }
        else { if (-p[1] - cB[1] - r < 0) return 1; }
    }
    if (-p[1] <= cB[1] && cB[1] <= p[1] && -p[2] <= cB[2] && cB[2] <= p[2]) {
        if (cB[0] > 0) { if (cB[0] - p[0] - r < 0) return 1; }
        else { if (-p[0] - cB[0] - r < 0) return 1; }
    }
    return 0;
}
int32_t orc_box_sphere_test(const orc_world* w, int32_t box, const double* c, double r) {
    oshape a = body_shape(w, box);
    return box_sphere_test_s(&a, c, r);
}

/* ------------------------------------------------------------------ */
/* Sphere trees (CollisionProcessor.java:793-1056, BVSphere.java:174-186)*/
/* ------------------------------------------------------------------ */
/* BVSphere.updatecW: every node centre of the body, once per detection; the
 * nodes of a composite's sphere sub-bodies with the sub-body's transform */
static void update_tree_cw(orc_world* w, int32_t b) {
    const am_body_desc* d = &w->desc[b];
    if (d->geom == AM_GEOM_COMPOSITE) {
        for (int32_t q = 0; q < d->tree_count; q++) {
            const am_body_desc* sd = &w->subs[d->tree_first + q];
            if (sd->geom != AM_GEOM_TREE) continue;
            double x[3], R[9];
            sub_pose(w, b, q, x, R);
            for (int32_t k = 0; k < sd->tree_count; k++) {
                int32_t g = sd->tree_first + k;
                xform_point(&w->cw[g * 3], R, x, w->nodes[g].c);
            }
        }
        return;
    }
    for (int32_t k = 0; k < d->tree_count; k++) {
        int32_t g = d->tree_first + k;
        xform_point(&w->cw[g * 3], &w->R[b * 9], &w->x[b * 3], w->nodes[g].c);
    }
}

static double dist2(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

/* collideSphereTreeAndPlane (CollisionProcessor.java:793-826) */
static void tree_plane(orc_world* w, const oshape* body, int32_t node, const oshape* plane) {
    const am_body_desc* bd = body->d;
    const am_body_desc* pd = plane->d;
    int32_t g = bd->tree_first + node;
    const double* c = &w->cw[g * 3];
    const double* n = pd->plane_n;
    double d = n[0] * c[0] + n[1] * c[1] + n[2] * c[2] + pd->plane_d - w->nodes[g].r;
    if (d < 0) {
        if (w->nodes[g].n_children == 0) {
            double nm[3] = {c[0] - pd->plane_p[0], c[1] - pd->plane_p[1], c[2] - pd->plane_p[2]};
            double val = dot3(nm, n);
            nm[0] = val * n[0]; nm[1] = val * n[1]; nm[2] = val * n[2];
            double cw[3] = {c[0] - nm[0], c[1] - nm[1], c[2] - nm[2]};
            double nn[3] = {-1.0 * n[0], -1.0 * n[1], -1.0 * n[2]};
            emit(w, body->body, plane->body, cw, nn, enc_bv(body, node), AM_BV_PLANE, 0, d);
        } else {
            for (int32_t k = 0; k < w->nodes[g].n_children; k++)
                tree_plane(w, body, w->nodes[g].first_child + k, plane);
        }
    }
}

/* collideBoxAndSphereTree (CollisionProcessor.java:837-850) */
static void box_tree(orc_world* w, const oshape* box, const oshape* body2, int32_t node) {
    int32_t g = body2->d->tree_first + node;
    if (!box_sphere_test_s(box, &w->cw[g * 3], w->nodes[g].r)) return;
    if (w->nodes[g].n_children != 0) {
        for (int32_t k = 0; k < w->nodes[g].n_children; k++)
            box_tree(w, box, body2, w->nodes[g].first_child + k);
    } else {
        box_sphere_s(w, box, body2, &w->cw[g * 3], w->nodes[g].r);
    }
}

/* processCollision (CollisionProcessor.java:1036-1056) */
static void leaf_leaf(orc_world* w, const oshape* S1, int32_t n1, const oshape* S2, int32_t n2) {
    int32_t g1 = S1->d->tree_first + n1, g2 = S2->d->tree_first + n2;
    const double* c1 = &w->cw[g1 * 3];
    const double* c2 = &w->cw[g2 * 3];
    double r1 = w->nodes[g1].r, r2 = w->nodes[g2].r;
    double distance = dist3(c1, c2);
    double sum = r2 + r1;
    if (distance < sum) {
        double alpha = (r1 - r2 + distance) / (2 * distance);
        double p[3], n[3];
        for (int i = 0; i < 3; i++) p[i] = (1 - alpha) * c1[i] + alpha * c2[i];
        for (int i = 0; i < 3; i++) n[i] = c2[i] - c1[i];
        normalize3(n);
        emit(w, S1->body, S2->body, p, n, enc_bv(S1, n1), enc_bv(S2, n2), 0, distance - sum);
    }
}

/* collideSphereTrees (CollisionProcessor.java:975-1009) */
static void tree_tree(orc_world* w, const oshape* S1, int32_t n1, const oshape* S2, int32_t n2) {
    int32_t g1 = S1->d->tree_first + n1, g2 = S2->d->tree_first + n2;
    const am_sphere_node* s1 = &w->nodes[g1];
    const am_sphere_node* s2 = &w->nodes[g2];
    if (dist2(&w->cw[g1 * 3], &w->cw[g2 * 3]) < (s1->r + s2->r) * (s1->r + s2->r)) {
        int l1 = s1->n_children == 0, l2 = s2->n_children == 0;
        if (l1 && l2) leaf_leaf(w, S1, n1, S2, n2);
        else if (l1) { for (int32_t k = 0; k < s2->n_children; k++) tree_tree(w, S1, n1, S2, s2->first_child + k); }
        else if (l2) { for (int32_t k = 0; k < s1->n_children; k++) tree_tree(w, S1, s1->first_child + k, S2, n2); }
        else if (s1->r <= s2->r) { for (int32_t k = 0; k < s2->n_children; k++) tree_tree(w, S1, n1, S2, s2->first_child + k); }
        else { for (int32_t k = 0; k < s1->n_children; k++) tree_tree(w, S1, s1->first_child + k, S2, n2); }
    }
}

/* narrowPhase on two primitive shapes (CollisionProcessor.java:723-750) */
static void shape_phase(orc_world* w, const oshape* s1, const oshape* s2) {
    const int32_t g1 = s1->d->geom, g2 = s2->d->geom;
    if (g1 == AM_GEOM_PLANE) {
        if (g2 == AM_GEOM_PLANE) return; /* plane-plane: reported impossible, no contact */
        if (g2 == AM_GEOM_BOX) box_plane_s(w, s2, s1);
        else tree_plane(w, s2, 0, s1);
    } else if (g1 == AM_GEOM_BOX) {
        if (g2 == AM_GEOM_PLANE) box_plane_s(w, s1, s2);
        else if (g2 == AM_GEOM_BOX) box_box_s(w, s1, s2);
        else box_tree(w, s1, s2, 0);
    } else {
        if (g2 == AM_GEOM_PLANE) tree_plane(w, s1, 0, s2);
        else if (g2 == AM_GEOM_BOX) box_tree(w, s2, s1, 0);
        else tree_tree(w, s1, 0, s2, 0);
    }
}

/* narrowPhase type dispatch (CollisionProcessor.java:707-757, 768-791 with
 * collectionCD = 0): collections first, then composites, each sub-body in
 * turn (:711-722) */
static void narrow_phase(orc_world* w, int32_t b1, int32_t b2) {
    const am_body_desc* d1 = &w->desc[b1];
    const am_body_desc* d2 = &w->desc[b2];
    if (d1->geom == AM_GEOM_COLLECTION) {
        for (int32_t q = 0; q < w->mem_count[b1]; q++) narrow_phase(w, w->members[w->mem_first[b1] + q], b2);
        return;
    }
    if (d2->geom == AM_GEOM_COLLECTION) {
        for (int32_t q = 0; q < w->mem_count[b2]; q++) narrow_phase(w, b1, w->members[w->mem_first[b2] + q]);
        return;
    }
    double buf1[12], buf2[12];
    if (d1->geom == AM_GEOM_COMPOSITE) {
        oshape o = body_shape(w, b2);
        for (int32_t q = 0; q < d1->tree_count; q++) {
            oshape s = sub_shape(w, b1, q, buf1);
            if (d2->geom == AM_GEOM_COMPOSITE) {
                for (int32_t q2 = 0; q2 < d2->tree_count; q2++) {
                    oshape s2 = sub_shape(w, b2, q2, buf2);
                    shape_phase(w, &s, &s2);
                }
            } else {
                shape_phase(w, &s, &o);
            }
        }
        return;
    }
    oshape a = body_shape(w, b1);
    if (d2->geom == AM_GEOM_COMPOSITE) {
        for (int32_t q = 0; q < d2->tree_count; q++) {
            oshape s2 = sub_shape(w, b2, q, buf2);
            shape_phase(w, &a, &s2);
        }
        return;
    }
    oshape b = body_shape(w, b2);
    shape_phase(w, &a, &b);
}

/* collisionDetection + broadPhase (CollisionProcessor.java:91-102, 683-697) */
int32_t orc_detect(orc_world* w) {
    w->nc = 0;
    for (int32_t i = 0; i < w->n; i++)
        if (w->desc[i].geom == AM_GEOM_TREE || w->desc[i].geom == AM_GEOM_COMPOSITE) update_tree_cw(w, i);
    int32_t N = w->n_top;
    for (int32_t i = 0; i < N - 1; i++) {
        int32_t b1 = w->top[i];
        for (int32_t j = i + 1; j < N; j++) {
            int32_t b2 = w->top[j];
            if (pinned(w, b1) && pinned(w, b2)) continue;
            if ((pinned(w, b1) && sleeping(w, b2)) || (pinned(w, b2) && sleeping(w, b1))) continue;
            /* bodies of independent scenes (batched worlds) never collide */
            if (w->desc[b1].scene != w->desc[b2].scene) continue;
            narrow_phase(w, b1, b2);
        }
    }
    if (w->nc == 0) w->nold = 0;
    return w->nc;
}

void orc_contacts_set(orc_world* w, int32_t n, const int32_t* body1, const int32_t* body2,
                      const double* pw, const double* nw, const double* viol, const int32_t* info,
                      const int32_t* bv1, const int32_t* bv2) {
    w->nc = 0;
    for (int32_t k = 0; k < n; k++)
        emit(w, body1[k], body2[k], &pw[k * 3], &nw[k * 3], bv1 ? bv1[k] : AM_BV_NONE,
             bv2 ? bv2[k] : AM_BV_NONE, info ? info[k] : 0, viol[k]);
}

/* Installs contacts field by field (the internal contacts of collections
 * joining the list of the single sweep, or contacts coming back from it). */
void orc_contacts_set_raw(orc_world* w, int32_t n, const am_contact_arrays* in) {
    w->nc = 0;
    for (int32_t k = 0; k < n; k++) {
        push_contact_slot(w);
        orc_contact* c = &w->c[w->nc++];
        c->b1 = in->body1[k];
        c->b2 = in->body2[k];
        c->bv1 = in->bv1[k];
        c->bv2 = in->bv2[k];
        c->info = in->info[k];
        memcpy(c->pB1, &in->point_b1[k * 3], 24);
        memcpy(c->nB1, &in->normal_b1[k * 3], 24);
        memcpy(c->t1B1, &in->tangent1_b1[k * 3], 24);
        memcpy(c->t2B1, &in->tangent2_b1[k * 3], 24);
        c->viol = in->violation[k];
        c->prev_viol = in->prev_violation[k];
        memcpy(c->J, &in->jacobian[k * 36], 36 * 8);
        memcpy(c->lam, &in->lambda[k * 3], 24);
        memcpy(c->lamw, &in->lambda_warm[k * 3], 24);
        memcpy(c->b, &in->b[k * 3], 24);
        memcpy(c->D, &in->diag[k * 3], 24);
        c->state = in->state[k];
        c->is_new = in->new_this_step[k];
    }
}

int32_t orc_contacts_get(const orc_world* w, am_contact_arrays* o, int32_t cap) {
    int32_t n = w->nc < cap ? w->nc : cap;
    for (int32_t k = 0; k < n; k++) {
        const orc_contact* c = &w->c[k];
        if (o->body1) o->body1[k] = c->b1;
        if (o->body2) o->body2[k] = c->b2;
        if (o->bv1) o->bv1[k] = c->bv1;
        if (o->bv2) o->bv2[k] = c->bv2;
        if (o->info) o->info[k] = c->info;
        if (o->point_b1) memcpy(&o->point_b1[k * 3], c->pB1, 24);
        if (o->normal_b1) memcpy(&o->normal_b1[k * 3], c->nB1, 24);
        if (o->tangent1_b1) memcpy(&o->tangent1_b1[k * 3], c->t1B1, 24);
        if (o->tangent2_b1) memcpy(&o->tangent2_b1[k * 3], c->t2B1, 24);
        if (o->violation) o->violation[k] = c->viol;
        if (o->prev_violation) o->prev_violation[k] = c->prev_viol;
        if (o->jacobian) memcpy(&o->jacobian[k * 36], c->J, 36 * 8);
        if (o->lambda) memcpy(&o->lambda[k * 3], c->lam, 24);
        if (o->lambda_warm) memcpy(&o->lambda_warm[k * 3], c->lamw, 24);
        if (o->b) memcpy(&o->b[k * 3], c->b, 24);
        if (o->diag) memcpy(&o->diag[k * 3], c->D, 24);
        if (o->state) o->state[k] = c->state;
        if (o->new_this_step) o->new_this_step[k] = c->is_new;
    }
    return w->nc;
}

void orc_lambdas_set(orc_world* w, const double* lam, int32_t n) {
    for (int32_t k = 0; k < n && k < w->nc; k++) {
        memcpy(w->c[k].lam, &lam[k * 3], 24);
        memcpy(w->c[k].lamw, &lam[k * 3], 24);
    }
}

/* ------------------------------------------------------------------ */
/* Warm start (CollisionProcessor.java:443-665)                         */
/* ------------------------------------------------------------------ */
/* Contact identity (Contact.hashCode/equals, Contact.java:543-572): the
 * unordered body pair with each side's bounding volume, plus info. */
static uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
typedef struct { int32_t lo, hi, bvlo, bvhi, info; } ckey;
static ckey make_key(int32_t b1, int32_t b2, int32_t bv1, int32_t bv2, int32_t info) {
    ckey k;
    if (b1 < b2) { k.lo = b1; k.hi = b2; k.bvlo = bv1; k.bvhi = bv2; }
    else { k.lo = b2; k.hi = b1; k.bvlo = bv2; k.bvhi = bv1; }
    k.info = info;
    return k;
}
static uint64_t key_hash(ckey k) {
    uint64_t h = mix64(((uint64_t)(uint32_t)k.lo << 32) | (uint32_t)k.hi);
    h = mix64(h ^ (((uint64_t)(uint32_t)k.bvlo << 32) | (uint32_t)k.bvhi));
    return mix64(h ^ (uint64_t)(uint32_t)k.info);
}
static int key_eq(ckey a, const orc_contact* c) {
    ckey b = make_key(c->b1, c->b2, c->bv1, c->bv2, c->info);
    return a.lo == b.lo && a.hi == b.hi && a.bvlo == b.bvlo && a.bvhi == b.bvhi && a.info == b.info;
}

/* updateContactsMap (CollisionProcessor.java:443-448): the old list plus an
 * open-addressing table; for equal keys the later contact wins, as a
 * HashMap.put of an equal key replaces the value. */
void orc_update_contacts_map(orc_world* w) {
    if (w->capold < w->nc) {
        w->capold = w->nc;
        w->old = (orc_contact*)realloc(w->old, (size_t)w->capold * sizeof(orc_contact));
    }
    if (w->nc) memcpy(w->old, w->c, (size_t)w->nc * sizeof(orc_contact));
    w->nold = w->nc;
    int32_t cap = 16;
    while (cap < 2 * w->nold + 16) cap *= 2;
    free(w->okey); free(w->oidx);
    w->okey = (uint64_t*)calloc((size_t)cap, sizeof(uint64_t));
    w->oidx = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
    for (int32_t i = 0; i < cap; i++) w->oidx[i] = -1;
    w->nokey = cap;
    for (int32_t i = 0; i < w->nold; i++) {
        const orc_contact* c = &w->old[i];
        ckey k = make_key(c->b1, c->b2, c->bv1, c->bv2, c->info);
        uint64_t h = key_hash(k);
        int32_t s = (int32_t)(h & (uint64_t)(cap - 1));
        for (;;) {
            if (w->oidx[s] < 0) { w->oidx[s] = i; break; }
            if (key_eq(k, &w->old[w->oidx[s]])) { w->oidx[s] = i; break; }
            s = (s + 1) & (cap - 1);
        }
    }
}

static orc_contact* old_lookup(orc_world* w, ckey k) {
    if (!w->nold || !w->nokey) return NULL;
    uint64_t h = key_hash(k);
    int32_t cap = w->nokey;
    int32_t s = (int32_t)(h & (uint64_t)(cap - 1));
    for (;;) {
        int32_t i = w->oidx[s];
        if (i < 0) return NULL;
        if (key_eq(k, &w->old[i])) return &w->old[i];
        s = (s + 1) & (cap - 1);
    }
}

static void take_old(orc_contact* c, orc_contact* o, int post, int zero_old) {
    c->is_new = 0;
    memcpy(c->lam, o->lam, 24);
    memcpy(c->lamw, o->lam, 24);
    if (zero_old) o->lam[0] = o->lam[1] = o->lam[2] = 0.0;
    c->prev_viol = post ? o->prev_viol : o->viol;
}

/* vanillaWarmStart (CollisionProcessor.java:646-665) */
static void vanilla_warm_start(orc_world* w, orc_contact* c, int post) {
    orc_contact* o = old_lookup(w, make_key(c->b1, c->b2, c->bv1, c->bv2, c->info));
    if (o) take_old(c, o, post, 0);
    else c->is_new = 1;
}

static void contact_world_point(const orc_world* w, const orc_contact* c, double* out) {
    xform_point(out, &w->R[c->b1 * 9], &w->x[c->b1 * 3], c->pB1);
}

/* warmStart (CollisionProcessor.java:481-640).  Body pair contacts are the
 * runs of contacts sharing an unordered body pair; within a pair contacts are
 * visited in list order, which fixes the "take and zero" outcome. */
/* Whether side `bv` of a contact on `body` is a composite part that is not a
 * box (Contact.csb1 / csb2 non-null and not RigidBodyGeomBox). */
static int non_box_part(const orc_world* w, int32_t body, int32_t bv) {
    if (w->desc[body].geom != AM_GEOM_COMPOSITE || bv > -3) return 0;
    int32_t s = (-3 - bv) >> 16;
    return w->subs[w->desc[body].tree_first + s].geom != AM_GEOM_BOX;
}

void orc_warm_start(orc_world* w, int32_t post, int32_t* bad_out, int32_t* rep_out) {
    int32_t bad = 0, rep = 0;
    for (int32_t k = 0; k < w->nc; k++) {
        orc_contact* c = &w->c[k];
        /* box or composite on both sides takes the box-box path (:489-493) */
        int g1 = w->desc[c->b1].geom, g2 = w->desc[c->b2].geom;
        int s1 = g1 == AM_GEOM_BOX || g1 == AM_GEOM_COMPOSITE, s2 = g2 == AM_GEOM_BOX || g2 == AM_GEOM_COMPOSITE;
        if (!(s1 && s2)) { vanilla_warm_start(w, c, post); continue; }
        /* a part of body 1 that is not a box: vanilla, next contact (:497-500);
         * of body 2: vanilla, then on through the box-box path (:501-503) */
        if (non_box_part(w, c->b1, c->bv1)) { vanilla_warm_start(w, c, post); continue; }
        if (non_box_part(w, c->b2, c->bv2)) vanilla_warm_start(w, c, post);
        double pNew[3], pOld[3];
        contact_world_point(w, c, pNew);
        orc_contact* o = old_lookup(w, make_key(c->b1, c->b2, c->bv1, c->bv2, c->info));
        if (o) {
            contact_world_point(w, o, pOld);
            double dist = dist3(pNew, pOld);
            if (dist > 0.05) {
                bad++;
                int myInfo = c->info;
                double best = dist;
                int bestInfo = myInfo;
                orc_contact* bestOld = o;
                for (int info = 0; info < 9; info++) {
                    if (info == myInfo) continue;
                    o = old_lookup(w, make_key(c->b1, c->b2, c->bv1, c->bv2, info));
                    if (!o) break;
                    contact_world_point(w, o, pOld);
                    dist = dist3(pNew, pOld);
                    if (dist < best) { best = dist; bestInfo = info; bestOld = o; }
                }
                if (bestInfo != myInfo) rep++;
                o = bestOld;
                dist = best;
            }
            if (dist < 0.05) take_old(c, o, post, 1);
            else c->is_new = 1;
        } else {
            double best = DBL_MAX;
            int bestInfo = -1;
            orc_contact* bestOld = NULL;
            for (int info = 0; info < 9; info++) {
                o = old_lookup(w, make_key(c->b1, c->b2, c->bv1, c->bv2, info));
                if (!o) break;
                contact_world_point(w, o, pOld);
                double dist = dist3(pNew, pOld);
                if (dist < best) { best = dist; bestInfo = info; bestOld = o; }
            }
            if (bestInfo != -1) {
                rep++;
                if (best < 0.05) take_old(c, bestOld, post, 1);
            } else {
                c->is_new = 1;
            }
        }
    }
    if (bad_out) *bad_out = bad;
    if (rep_out) *rep_out = rep;
}

/* redoWarmStart / noWarmStart (CollisionProcessor.java:453-470) */
void orc_redo_warm_start(orc_world* w) {
    for (int32_t k = 0; k < w->nc; k++) memcpy(w->c[k].lam, w->c[k].lamw, 24);
}
void orc_no_warm_start(orc_world* w) {
    for (int32_t k = 0; k < w->nc; k++) w->c[k].lam[0] = w->c[k].lam[1] = w->c[k].lam[2] = 0.0;
}

/* ------------------------------------------------------------------ */
/* PGS (PGS.java:73-256, Contact.java:279-398)                          */
/* ------------------------------------------------------------------ */
/* Contact.computeB (Contact.java:279-326) */
static void compute_b(const orc_world* w, orc_contact* c, double dt, double fb, int in_coll,
                      int rest_ovr, double rest_val) {
    int32_t a = solve_body(w, c->b1, in_coll), b = solve_body(w, c->b2, in_coll);
    double restitution = (w->desc[c->b1].restitution + w->desc[c->b2].restitution) / 2.;
    if (rest_ovr) restitution = rest_val;
    double bn = 0, bt1 = 0, bt2 = 0, tmp[3];
    const int32_t ids[2] = {a, b};
    for (int s = 0; s < 2; s++) {
        int32_t e = ids[s];
        const double* Jn = &c->J[0 + s * 6];
        const double* Jt1 = &c->J[12 + s * 6];
        const double* Jt2 = &c->J[24 + s * 6];
        double sc = w->desc[e].minv * dt;
        for (int i = 0; i < 3; i++) tmp[i] = sc * w->f[e * 3 + i] + w->v[e * 3 + i];
        bn += dot3(tmp, Jn);
        bt1 += dot3(tmp, Jt1);
        bt2 += dot3(tmp, Jt2);
        mat_vec(tmp, &w->jinv[e * 9], &w->t[e * 3]);
        for (int i = 0; i < 3; i++) tmp[i] *= dt;
        for (int i = 0; i < 3; i++) tmp[i] = tmp[i] + w->w[e * 3 + i];
        bn += dot3(tmp, Jn + 3);
        bt1 += dot3(tmp, Jt1 + 3);
        bt2 += dot3(tmp, Jt2 + 3);
        double bounce = dot3(&w->v[e * 3], Jn) + dot3(&w->w[e * 3], Jn + 3);
        bounce *= restitution;
        bn += bounce;
    }
    bn += fb * c->viol;
    c->b[0] = bn; c->b[1] = bt1; c->b[2] = bt2;
}

/* Contact.computeJMinvJt (Contact.java:340-354) */
static void compute_d(const orc_world* w, orc_contact* c, int in_coll) {
    int32_t a = solve_body(w, c->b1, in_coll), b = solve_body(w, c->b2, in_coll);
    for (int row = 0; row < 3; row++) {
        const double* ja = &c->J[row * 12];
        const double* jb = &c->J[row * 12 + 6];
        double t1[3], t2[3];
        mat_vec(t1, &w->jinv[a * 9], ja + 3);
        mat_vec(t2, &w->jinv[b * 9], jb + 3);
        c->D[row] = w->desc[a].minv * dot3(ja, ja) + dot3(ja + 3, t1) + w->desc[b].minv * dot3(jb, jb) + dot3(jb + 3, t2);
    }
}

/* Contact.getJdv (Contact.java:362-378) */
static double get_jdv(const orc_world* w, const orc_contact* c, int in_coll, int row) {
    int32_t a = solve_body(w, c->b1, in_coll), b = solve_body(w, c->b2, in_coll);
    const double* ja = &c->J[row * 12];
    const double* jb = &c->J[row * 12 + 6];
    const double* d1 = &w->dv[a * 6];
    const double* d2 = &w->dv[b * 6];
    return (dot3(ja, d1) + dot3(ja + 3, d1 + 3)) + (dot3(jb, d2) + dot3(jb + 3, d2 + 3));
}

/* PGS.updateDeltaVwithLambdai (PGS.java:221-245) */
static void update_dv(orc_world* w, const orc_contact* c, double lambda, int row, int in_coll) {
    int32_t a = solve_body(w, c->b1, in_coll), b = solve_body(w, c->b2, in_coll);
    const double* ja = &c->J[row * 12];
    const double* jb = &c->J[row * 12 + 6];
    double tmp[3];
    double* d1 = &w->dv[a * 6];
    double* d2 = &w->dv[b * 6];
    double s1 = w->desc[a].minv * lambda;
    for (int i = 0; i < 3; i++) d1[i] = s1 * ja[i] + d1[i];
    mat_vec(tmp, &w->jinv[a * 9], ja + 3);
    for (int i = 0; i < 3; i++) d1[3 + i] = lambda * tmp[i] + d1[3 + i];
    double s2 = w->desc[b].minv * lambda;
    for (int i = 0; i < 3; i++) d2[i] = s2 * jb[i] + d2[i];
    mat_vec(tmp, &w->jinv[b * 9], jb + 3);
    for (int i = 0; i < 3; i++) d2[3 + i] = lambda * tmp[i] + d2[3 + i];
}

static int clamps(const orc_world* w, const orc_contact* c) {
    return !(w->desc[c->b1].flags & AM_BODY_MAGNET_ACTIVE) && !(w->desc[c->b2].flags & AM_BODY_MAGNET_ACTIVE);
}

/* friction coefficient rule (PGS.java:128-142) */
static double contact_mu(const orc_world* w, const orc_contact* c) {
    if (w->p.friction_override) return w->p.friction;
    double f1 = w->desc[c->b1].friction, f2 = w->desc[c->b2].friction;
    if (f1 < 0.2 || f2 < 0.2) return jmin(f1, f2);
    if (f1 > 1. || f2 > 1.) return jmax(f1, f2);
    return (f1 + f2) / 2.;
}

/* Contact.updateContactState (Contact.java:385-398) */
static void update_state(const orc_world* w, orc_contact* c, int in_coll) {
    double w1 = c->b[1] + get_jdv(w, c, in_coll, 1);
    double w2 = c->b[2] + get_jdv(w, c, in_coll, 2);
    if (fabs(c->lam[0]) <= 1e-14) c->state = AM_STATE_BROKEN;
    else if (fabs(w1) > w->p.sliding_threshold) c->state = AM_STATE_ONEDGE;
    else if (fabs(w2) > w->p.sliding_threshold) c->state = AM_STATE_ONEDGE;
    else c->state = AM_STATE_CLEAR;
}

/* CollisionProcessor.solveLCP / updateInCollections configuration of PGS,
 * then PGS.solve (PGS.java:73-194) over contacts in the given order. */
int32_t orc_pgs_solve(orc_world* w, const int32_t* order, int32_t n_order, double dt, int32_t main_solve,
                      int32_t post, double* changes, double* last_change) {
    const am_params* P = &w->p;
    int in_coll = !main_solve;
    int iterations = main_solve ? P->iterations : P->iterations_in_collection;
    double fb;
    if (main_solve) fb = (!P->enable_post_stabilization || post) ? P->feedback_stiffness : 0.0;
    else fb = P->feedback_stiffness;
    double compliance = P->enable_compliance ? P->compliance : 0.0;
    double omega = P->omega;
    int post_mode = main_solve ? post : 0;
    if (n_order == 0) { if (last_change) *last_change = 0; return 0; }
    /* updateJacobiansThatNeedUpdating (CollisionProcessor.java:326-334) */
    for (int32_t k = 0; k < n_order; k++) {
        orc_contact* c = &w->c[order[k]];
        if (w->parent[c->b1] >= 0 || w->parent[c->b2] >= 0) compute_jacobian(w, c, in_coll);
    }
    /* confidentWarmStart (PGS.java:250-256); solver.warmStart is true in both uses */
    for (int32_t k = 0; k < n_order; k++) {
        const orc_contact* c = &w->c[order[k]];
        update_dv(w, c, c->lam[0], 0, in_coll);
        update_dv(w, c, c->lam[1], 1, in_coll);
        update_dv(w, c, c->lam[2], 2, in_coll);
    }
    for (int32_t k = 0; k < n_order; k++) {
        orc_contact* c = &w->c[order[k]];
        if (post_mode) { c->b[0] = fb * c->viol; c->b[1] = 0.; c->b[2] = 0.; }
        else compute_b(w, c, dt, fb, in_coll, P->restitution_override, P->restitution);
        compute_d(w, c, in_coll);
    }
    /* Independent scenes in one list (batched worlds) stop one by one: each
     * keeps its own max |dl| and leaves the sweeps once it is under the
     * tolerance (PGS.java:190 per scene), which is the same as sweeping each
     * scene alone in this order restricted to its contacts. */
    int32_t n_sc = 1;
    for (int32_t i = 0; i < w->n; i++)
        if (w->desc[i].scene + 1 > n_sc) n_sc = w->desc[i].scene + 1;
    int per_scene = main_solve && n_sc > 1 && P->tolerance > 0;
    double* sc_max = NULL;
    int32_t* sc_used = NULL;
    if (per_scene) {
        sc_max = calloc((size_t)n_sc, sizeof(double));
        sc_used = calloc((size_t)n_sc, sizeof(int32_t));
    }
    w->n_scene_used = 0;
    int iter = iterations, done = 0;
    double maxabs = 0;
    while (iter > 0) {
        maxabs = 0;
        if (per_scene)
            for (int32_t s = 0; s < n_sc; s++) sc_max[s] = 0;
        for (int32_t k = 0; k < n_order; k++) {
            orc_contact* c = &w->c[order[k]];
            int32_t sc = per_scene ? w->desc[c->b1].scene : 0;
            if (per_scene && sc_used[sc]) continue;
            double cmax = 0;
            int cl = clamps(w, c);
            double Jdv = get_jdv(w, c, in_coll, 0);
            double prev = c->lam[0];
            c->lam[0] = (c->D[0] * c->lam[0] - omega * (c->b[0] + Jdv)) / (c->D[0] + compliance);
            if (cl) c->lam[0] = jmax(0, c->lam[0]);
            double diff = c->lam[0] - prev;
            update_dv(w, c, diff, 0, in_coll);
            cmax = jmax(cmax, fabs(diff));
            double mu = contact_mu(w, c);
            for (int row = 1; row < 3; row++) {
                Jdv = get_jdv(w, c, in_coll, row);
                prev = c->lam[row];
                c->lam[row] = (c->D[row] * c->lam[row] - omega * (c->b[row] + Jdv)) / (c->D[row] + compliance);
                if (cl) {
                    double limit = mu * c->lam[0];
                    c->lam[row] = jmax(c->lam[row], -limit);
                    c->lam[row] = jmin(c->lam[row], limit);
                }
                diff = c->lam[row] - prev;
                update_dv(w, c, diff, row, in_coll);
                cmax = jmax(cmax, fabs(diff));
            }
            maxabs = jmax(maxabs, cmax);
            if (per_scene) sc_max[sc] = jmax(sc_max[sc], cmax);
            if (iter == 1) update_state(w, c, in_coll);
        }
        iter--;
        if (main_solve && changes) changes[done] = log(maxabs);
        done++;
        if (per_scene) {
            int all = 1;
            for (int32_t s = 0; s < n_sc; s++) {
                if (sc_used[s]) continue;
                if (sc_max[s] < P->tolerance || iter == 0) {
                    sc_used[s] = done;
                    if (s < ORC_MAX_SCENES) w->scene_last[s] = sc_max[s];
                } else {
                    all = 0;
                }
            }
            if (all) break;
        } else if (main_solve && maxabs < P->tolerance) {
            break;
        }
    }
    if (per_scene) {
        w->n_scene_used = n_sc < ORC_MAX_SCENES ? n_sc : ORC_MAX_SCENES;
        for (int32_t s = 0; s < w->n_scene_used; s++) w->scene_used[s] = sc_used[s];
        free(sc_max);
        free(sc_used);
    }
    if (last_change) *last_change = maxabs;
    return done;
}

int32_t orc_scene_iterations(const orc_world* w, int32_t* used, double* last, int32_t capacity) {
    int32_t n = w->n_scene_used < capacity ? w->n_scene_used : capacity;
    for (int32_t s = 0; s < n; s++) {
        if (used) used[s] = w->scene_used[s];
        if (last) last[s] = w->scene_last[s];
    }
    return w->n_scene_used;
}

/* ------------------------------------------------------------------ */
/* Integration (RigidBody.java:382-441)                                 */
/* ------------------------------------------------------------------ */
/* RigidBody.advanceVelocities (RigidBody.java:409-417) */
static void advance_velocities_one(orc_world* w, int32_t i, double dt) {
    double s = dt * w->desc[i].minv, dom[3];
    for (int a = 0; a < 3; a++) w->v[i * 3 + a] = s * w->f[i * 3 + a] + w->v[i * 3 + a];
    for (int a = 0; a < 3; a++) w->v[i * 3 + a] = w->v[i * 3 + a] + w->dv[i * 6 + a];
    mat_vec(dom, &w->jinv[i * 9], &w->t[i * 3]);
    for (int a = 0; a < 3; a++) dom[a] *= dt;
    for (int a = 0; a < 3; a++) w->w[i * 3 + a] = w->w[i * 3 + a] + dom[a];
    for (int a = 0; a < 3; a++) w->w[i * 3 + a] = w->w[i * 3 + a] + w->dv[i * 6 + 3 + a];
}

/* Sine and cosine for expRodrigues, which calls Math.cos / Math.sin.  The
 * algorithm is fdlibm 5.3's (Sun Microsystems, 1993): s_sin.c / s_cos.c,
 * the kernels k_sin.c / k_cos.c (degree-13 and degree-14 minimax polynomials
 * on [-pi/4, pi/4]) and the Cody-Waite reduction of e_rem_pio2.c for
 * |x| <= 2^19 pi/2; java.lang.StrictMath.sin/cos are defined as fdlibm's
 * results.  Arguments beyond 2^19 pi/2 (fdlibm's Payne-Hanek range, which a
 * rotation per step never reaches) use the same three-step reduction with a
 * floating-point quotient. */
static int32_t fd_hi(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    return (int32_t)(b >> 32);
}
static double fd_with_hi(int32_t hi) {
    uint64_t b = (uint64_t)(uint32_t)hi << 32;
    double x;
    memcpy(&x, &b, 8);
    return x;
}
static const double fd_S1 = -1.66666666666666324348e-01, fd_S2 = 8.33333333332248946124e-03,
                    fd_S3 = -1.98412698298579493134e-04, fd_S4 = 2.75573137070700676789e-06,
                    fd_S5 = -2.50507602534068634195e-08, fd_S6 = 1.58969099521155010221e-10;
static const double fd_C1 = 4.16666666666666019037e-02, fd_C2 = -1.38888888888741095749e-03,
                    fd_C3 = 2.48015872894767294178e-05, fd_C4 = -2.75573143513906633035e-07,
                    fd_C5 = 2.08757232129817482790e-09, fd_C6 = -1.13596475577881948265e-11;
static const double fd_invpio2 = 6.36619772367581382433e-01, fd_pio2_1 = 1.57079632673412561417e+00,
                    fd_pio2_1t = 6.07710050650619224932e-11, fd_pio2_2 = 6.07710050630396597660e-11,
                    fd_pio2_2t = 2.02226624879595063154e-21, fd_pio2_3 = 2.02226624871116645580e-21,
                    fd_pio2_3t = 8.47842766036889956997e-32;
/* high words of the doubles nearest n pi/2, n = 1..32 */
static const int32_t fd_npio2_hw[32] = {
    0x3ff921fb, 0x400921fb, 0x4012d97c, 0x401921fb, 0x401f6a7a, 0x4022d97c, 0x4025fdbb, 0x402921fb,
    0x402c463a, 0x402f6a7a, 0x4031475c, 0x4032d97c, 0x40346b9c, 0x4035fdbb, 0x40378fdb, 0x403921fb,
    0x403ab41b, 0x403c463a, 0x403dd85a, 0x403f6a7a, 0x40407e4c, 0x4041475c, 0x4042106c, 0x4042d97c,
    0x4043a28c, 0x40446b9c, 0x404534ac, 0x4045fdbb, 0x4046c6cb, 0x40478fdb, 0x404858eb, 0x404921fb};

static double fd_kernel_sin(double x, double y, int iy) {
    int32_t ix = fd_hi(x) & 0x7fffffff;
    if (ix < 0x3e400000 && (int)x == 0) return x;
    double z = x * x, v = z * x;
    double r = fd_S2 + z * (fd_S3 + z * (fd_S4 + z * (fd_S5 + z * fd_S6)));
    if (iy == 0) return x + v * (fd_S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * fd_S1);
}
static double fd_kernel_cos(double x, double y) {
    int32_t ix = fd_hi(x) & 0x7fffffff;
    if (ix < 0x3e400000 && (int)x == 0) return 1.0;
    double z = x * x;
    double r = z * (fd_C1 + z * (fd_C2 + z * (fd_C3 + z * (fd_C4 + z * (fd_C5 + z * fd_C6)))));
    if (ix < 0x3fd33333) return 1.0 - (0.5 * z - (z * r - x * y));
    double qx = ix > 0x3fe90000 ? 0.28125 : fd_with_hi(ix - 0x00200000);
    double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}
/* e_rem_pio2.c for |x| up to 2^19 pi/2: x = n pi/2 + y[0] + y[1] */
static int fd_rem_pio2(double x, double* y) {
    int32_t hx = fd_hi(x), ix = hx & 0x7fffffff;
    if (ix < 0x4002d97c) { /* |x| < 3pi/4: n = +-1 */
        double z;
        if (hx > 0) {
            z = x - fd_pio2_1;
            if (ix != 0x3ff921fb) {
                y[0] = z - fd_pio2_1t;
                y[1] = (z - y[0]) - fd_pio2_1t;
            } else {
                z -= fd_pio2_2;
                y[0] = z - fd_pio2_2t;
                y[1] = (z - y[0]) - fd_pio2_2t;
            }
            return 1;
        }
        z = x + fd_pio2_1;
        if (ix != 0x3ff921fb) {
            y[0] = z + fd_pio2_1t;
            y[1] = (z - y[0]) + fd_pio2_1t;
        } else {
            z += fd_pio2_2;
            y[0] = z + fd_pio2_2t;
            y[1] = (z - y[0]) + fd_pio2_2t;
        }
        return -1;
    }
    double t = fabs(x);
    const int medium = ix <= 0x413921fb;
    double fn;
    int n;
    if (medium) {
        n = (int)(t * fd_invpio2 + 0.5);
        fn = (double)n;
    } else {
        fn = floor(t * fd_invpio2 + 0.5);
        n = (int)(fn - 4.0 * floor(fn * 0.25));
    }
    double r = t - fn * fd_pio2_1;
    double w = fn * fd_pio2_1t;
    if (medium && n < 32 && ix != fd_npio2_hw[n - 1]) {
        y[0] = r - w;
    } else {
        int32_t j = ix >> 20;
        y[0] = r - w;
        int32_t i = j - ((fd_hi(y[0]) >> 20) & 0x7ff);
        if (i > 16 || !medium) {
            t = r;
            w = fn * fd_pio2_2;
            r = t - w;
            w = fn * fd_pio2_2t - ((t - r) - w);
            y[0] = r - w;
            i = j - ((fd_hi(y[0]) >> 20) & 0x7ff);
            if (i > 49 || !medium) {
                t = r;
                w = fn * fd_pio2_3;
                r = t - w;
                w = fn * fd_pio2_3t - ((t - r) - w);
                y[0] = r - w;
            }
        }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) {
        y[0] = -y[0];
        y[1] = -y[1];
        return -n;
    }
    return n;
}
double orc_sin(double x) {
    int32_t ix = fd_hi(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) return fd_kernel_sin(x, 0.0, 0);
    if (ix >= 0x7ff00000) return x - x;
    double y[2];
    switch (fd_rem_pio2(x, y) & 3) {
        case 0: return fd_kernel_sin(y[0], y[1], 1);
        case 1: return fd_kernel_cos(y[0], y[1]);
        case 2: return -fd_kernel_sin(y[0], y[1], 1);
        default: return -fd_kernel_cos(y[0], y[1]);
    }
}
double orc_cos(double x) {
    int32_t ix = fd_hi(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) return fd_kernel_cos(x, 0.0);
    if (ix >= 0x7ff00000) return x - x;
    double y[2];
    switch (fd_rem_pio2(x, y) & 3) {
        case 0: return fd_kernel_cos(y[0], y[1]);
        case 1: return -fd_kernel_sin(y[0], y[1], 1);
        case 2: return -fd_kernel_cos(y[0], y[1]);
        default: return fd_kernel_sin(y[0], y[1], 1);
    }
}

/* The pose update of RigidBody.advancePositions (RigidBody.java:427-437):
 * x += dt v (Tuple3d.scaleAdd), then for a turn angle |omega| dt above 1e-8
 * R = normalizeCP(expRodrigues(omega / |omega|, |omega| dt) R)
 * (RigidBody.java:382-401, Matrix3d.normalizeCP) */
static void advance_pose(double* x, double* R, const double* v, const double* om, double dt) {
    for (int a = 0; a < 3; a++) x[a] = dt * v[a] + x[a];
    double t = len3(om) * dt;
    double d[3] = {om[0], om[1], om[2]};
    normalize3(d);
    if (t > 1e-8) {
        double c = orc_cos(t), s = orc_sin(t), c1 = 1 - c, dR[9];
        double wX = d[0], wY = d[1], wZ = d[2];
        dR[0] = c + wX * wX * c1;
        dR[3] = wZ * s + wX * wY * c1;
        dR[6] = -wY * s + wX * wZ * c1;
        dR[1] = -wZ * s + wX * wY * c1;
        dR[4] = c + wY * wY * c1;
        dR[7] = wX * s + wY * wZ * c1;
        dR[2] = wY * s + wX * wZ * c1;
        dR[5] = -wX * s + wY * wZ * c1;
        dR[8] = c + wZ * wZ * c1;
        mat_mul(dR, dR, R);
        double mag = 1.0 / sqrt(dR[0] * dR[0] + dR[3] * dR[3] + dR[6] * dR[6]);
        R[0] = dR[0] * mag; R[3] = dR[3] * mag; R[6] = dR[6] * mag;
        mag = 1.0 / sqrt(dR[1] * dR[1] + dR[4] * dR[4] + dR[7] * dR[7]);
        R[1] = dR[1] * mag; R[4] = dR[4] * mag; R[7] = dR[7] * mag;
        R[2] = R[3] * R[7] - R[4] * R[6];
        R[5] = R[1] * R[6] - R[0] * R[7];
        R[8] = R[0] * R[4] - R[1] * R[3];
    }
}

/* RigidBody.advancePositions (RigidBody.java:427-441) */
static void advance_positions_one(orc_world* w, int32_t i, double dt) {
    advance_pose(&w->x[i * 3], &w->R[i * 9], &w->v[i * 3], &w->w[i * 3], dt);
    update_jinv(w, i);
}

/* RigidBodySystem.advanceBodiesVelocities/Positions (RigidBodySystem.java:342-352):
 * top-level entities that are neither pinned nor sleeping. */
void orc_advance_velocities(orc_world* w, double dt) {
    for (int32_t k = 0; k < w->n_top; k++) {
        int32_t i = w->top[k];
        if (!pinned(w, i) && !sleeping(w, i)) advance_velocities_one(w, i, dt);
    }
}
void orc_advance_positions(orc_world* w, double dt) {
    for (int32_t k = 0; k < w->n_top; k++) {
        int32_t i = w->top[k];
        if (!pinned(w, i) && !sleeping(w, i)) advance_positions_one(w, i, dt);
    }
}
/* RigidBody.advancePositionsPostStabilization (RigidBody.java:423-425): the
 * pose update with deltaV's linear and angular parts as the velocities, over
 * the loop of RigidBodySystem.postStabilization (:374-377) */
void orc_advance_positions_post(orc_world* w, double dt) {
    for (int32_t k = 0; k < w->n_top; k++) {
        int32_t i = w->top[k];
        if (pinned(w, i) || sleeping(w, i)) continue;
        advance_pose(&w->x[i * 3], &w->R[i * 9], &w->dv[i * 6], &w->dv[i * 6 + 3], dt);
        update_jinv(w, i);
    }
}

/* ------------------------------------------------------------------ */
/* Relative motion (MotionMetricProcessor.java:39-73)                   */
/* ------------------------------------------------------------------ */
/* RigidBody.getSpatialVelocity (RigidBody.java:458-462) */
static void spatial_velocity(const orc_world* w, int32_t b, const double* pW, double* out) {
    double tmp[3] = {pW[0] - w->x[b * 3 + 0], pW[1] - w->x[b * 3 + 1], pW[2] - w->x[b * 3 + 2]};
    cross3(out, &w->w[b * 3], tmp);
    for (int a = 0; a < 3; a++) out[a] = out[a] + w->v[b * 3 + a];
}

double orc_motion_metric(const orc_world* w, int32_t a, int32_t b) {
    double largest = 0, pW[3], v1[3], v2[3];
    if (b < 0) {
        for (int k = 0; k < 8; k++) {
            xform_point(pW, &w->R[a * 9], &w->x[a * 3], &w->desc[a].bbox[k * 3]);
            spatial_velocity(w, a, pW, v1);
            largest = jmax(len3(v1), largest);
        }
        return largest;
    }
    for (int s = 0; s < 2; s++) {
        int32_t body = s == 0 ? a : b;
        for (int k = 0; k < 8; k++) {
            xform_point(pW, &w->R[body * 9], &w->x[body * 3], &w->desc[body].bbox[k * 3]);
            spatial_velocity(w, a, pW, v1);
            spatial_velocity(w, b, pW, v2);
            for (int q = 0; q < 3; q++) v1[q] = v1[q] - v2[q];
            largest = jmax(len3(v1), largest);
        }
    }
    return largest;
}

/* MotionMetricProcessor.getLargestVelocityNorm(body1, body2, dt)
 * (MotionMetricProcessor.java:75-116): copies of both bodies advanced by dt
 * (RigidBody.set + advancePositions); each bounding-box corner of one body
 * goes through the other body's current and advanced poses and back through
 * its own advanced pose; the largest displacement over dt. */
double orc_motion_metric_pos(const orc_world* w, int32_t a, int32_t b, double dt) {
    double xa[2][3], Ra[2][9];
    const int32_t ids[2] = {a, b};
    for (int s = 0; s < 2; s++) {
        memcpy(xa[s], &w->x[ids[s] * 3], sizeof xa[s]);
        memcpy(Ra[s], &w->R[ids[s] * 9], sizeof Ra[s]);
        advance_pose(xa[s], Ra[s], &w->v[ids[s] * 3], &w->w[ids[s] * 3], dt);
    }
    double largest = 0;
    for (int s = 0; s < 2; s++) {
        const int32_t self = ids[s], other = ids[1 - s];
        for (int k = 0; k < 8; k++) {
            const double* p = &w->desc[self].bbox[k * 3];
            double pW[3], pB[3], v[3];
            xform_point(pW, &w->R[self * 9], &w->x[self * 3], p);
            inv_xform_point(pB, &w->R[other * 9], &w->x[other * 3], pW);
            xform_point(pW, Ra[1 - s], xa[1 - s], pB);
            inv_xform_point(pB, Ra[s], xa[s], pW);
            for (int q = 0; q < 3; q++) v[q] = p[q] - pB[q];
            const double inv_dt = 1. / dt;
            for (int q = 0; q < 3; q++) v[q] = v[q] * inv_dt;
            largest = jmax(len3(v), largest);
        }
    }
    return largest;
}

/* ------------------------------------------------------------------ */
/* One step without merging (RigidBodySystem.advanceTime, :102-185)     */
/* ------------------------------------------------------------------ */
int32_t orc_step(orc_world* w, double dt, int32_t* iters, double* last) {
    orc_clear_bodies(w);
    orc_apply_gravity(w);
    orc_apply_springs(w);
    orc_update_contacts_map(w);
    orc_detect(w);
    orc_warm_start(w, 0, NULL, NULL);
    if (w->p.warm_start) orc_redo_warm_start(w);
    else orc_no_warm_start(w);
    int32_t* order = (int32_t*)malloc((size_t)(w->nc ? w->nc : 1) * sizeof(int32_t));
    for (int32_t k = 0; k < w->nc; k++) order[k] = k;
    int32_t it = orc_pgs_solve(w, order, w->nc, dt, 1, 0, NULL, last);
    free(order);
    if (iters) *iters = it;
    orc_advance_velocities(w, dt);
    orc_advance_positions(w, dt);
    return w->nc;
}
