/*
 * am_oracle.h - CPU restatement of AdaptiveMerging's 3D per-step contact path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for libam
 * (the HIP implementation).  It may be linked or called only by the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark; the
 * product path never calls it.
 *
 * Every function follows a reference routine (paths under
 * src/mergingBodies3D/ of EulalieCoevoet/AdaptiveMerging), cited at its
 * definition.  Arithmetic is IEEE double, evaluated in the reference's
 * operation order (javax.vecmath semantics: Matrix3d.transform is
 * m00*x+m01*y+m02*z, Vector3d.normalize multiplies by 1/sqrt, ...), built
 * with -ffp-contract=off so no fused multiply-add changes a rounding.
 *
 * Parity pinning (what checks each part beyond retyping it):
 *   box-plane, sphere-sphere, sphere-plane leaves: closed forms (tests)
 *   box-box: closed forms for stacked/offset faces and crossing edges,
 *            swap symmetry, invariants (tests)
 *   box-sphere: closed forms for face/edge/corner regions + SDF sampling
 *   sphere trees: equal to brute-force leaf pairs on conservative trees
 *   contact frame / Jacobian: orthonormality and closed form
 *   PGS: one contact on a plane has a closed-form lambda; projection
 *        invariants; convergence of the KKT residual
 *   motion metric: rigid co-motion gives 0, pure translation gives |dv|
 *   integration: closed-form free flight and constant-spin rotation
 * Parity unpinned: javax.vecmath rounding beyond its documented formulas
 * (the jar is compiled and not vendored), and Java Math.sin/cos in the
 * Rodrigues update (the C library's sin/cos are used).
 */
#ifndef AM_ORACLE_H
#define AM_ORACLE_H

#include <stdint.h>
#include "../include/am.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

orc_world* orc_create(const am_body_desc* bodies, int32_t n_entities,
                      const am_sphere_node* nodes, int32_t n_nodes);
/* orc_create with composite bodies (AM_GEOM_COMPOSITE): their sub-bodies and
 * transforms, as am_scene_upload_ex takes them. */
orc_world* orc_create_ex(const am_body_desc* bodies, int32_t n_entities, const am_sphere_node* nodes,
                         int32_t n_nodes, const am_body_desc* subs, const double* sub_b2c, int32_t n_subs);
void orc_destroy(orc_world* w);
void orc_set_params(orc_world* w, const am_params* p);

/* Entities = physical bodies followed by collection entities.  parent[i] is
 * the collection entity of body i or -1.  top[] is the broad-phase list
 * (RigidBodySystem.bodies order); members[] lists each collection's bodies
 * (member_first/member_count index it per entity). */
void orc_set_structure(orc_world* w, const int32_t* parent, const int32_t* top, int32_t n_top,
                       const int32_t* member_first, const int32_t* member_count,
                       const int32_t* members);

/* Grows the entity count (new entities are unused collection slots). */
void orc_resize(orc_world* w, int32_t n_entities);
int32_t orc_entities(const orc_world* w);
void orc_set_entity(orc_world* w, int32_t entity, const am_body_desc* desc);
void orc_get_entity(const orc_world* w, int32_t entity, am_body_desc* desc);

void orc_set_state(orc_world* w, const double* x, const double* R, const double* v,
                   const double* omega);
void orc_get_state(const orc_world* w, double* x, double* R, double* v, double* omega);
void orc_set_jinv(orc_world* w, int32_t entity, const double* jinv);
void orc_get_jinv(const orc_world* w, double* jinv);
void orc_set_forces(orc_world* w, const double* force, const double* torque);
void orc_get_forces(const orc_world* w, double* force, double* torque);
void orc_get_deltav(const orc_world* w, double* dv);
void orc_set_deltav(orc_world* w, const double* dv);
void orc_clear_bodies(orc_world* w);
void orc_apply_gravity(orc_world* w);
void orc_set_springs(orc_world* w, const am_spring* springs, int32_t n);
void orc_apply_springs(orc_world* w);

/* Detection. */
void orc_update_contacts_map(orc_world* w);
int32_t orc_detect(orc_world* w);
void orc_contacts_set(orc_world* w, int32_t n, const int32_t* body1, const int32_t* body2,
                      const double* point_w, const double* normal_w, const double* violation,
                      const int32_t* info, const int32_t* bv1, const int32_t* bv2);
int32_t orc_contacts_get(const orc_world* w, am_contact_arrays* out, int32_t capacity);
void orc_contacts_set_raw(orc_world* w, int32_t n, const am_contact_arrays* in);
void orc_lambdas_set(orc_world* w, const double* lambda, int32_t n);

/* Primitive narrow-phase entry points on a body pair (results appended to the
 * current contact list); return the number of contacts appended. */
int32_t orc_box_box(orc_world* w, int32_t b1, int32_t b2);
int32_t orc_box_plane(orc_world* w, int32_t box, int32_t plane);
int32_t orc_box_sphere(orc_world* w, int32_t box, int32_t other, const double* cW, double r);
int32_t orc_box_sphere_test(const orc_world* w, int32_t box, const double* cW, double r);

/* Warm start and solve. */
void orc_warm_start(orc_world* w, int32_t post_stabilization, int32_t* bad, int32_t* repaired);
void orc_redo_warm_start(orc_world* w);
void orc_no_warm_start(orc_world* w);
/* PGS.solve over contacts in the given order (a sequential sweep).
 * main_solve != 0 configures as CollisionProcessor.solveLCP, otherwise as the
 * single sweep of updateInCollections (computeInCollection).  Returns the
 * number of sweeps done; changes[] (may be NULL) receives log(max|dl|). */
int32_t orc_pgs_solve(orc_world* w, const int32_t* order, int32_t n_order, double dt,
                      int32_t main_solve, int32_t post_stabilization, double* changes,
                      double* last_change);

/* Sweeps used and the max |dl| of the last sweep per scene of the last main
 * solve whose contacts spanned several scenes (returns the scene count). */
int32_t orc_scene_iterations(const orc_world* w, int32_t* used, double* last, int32_t capacity);

void orc_advance_velocities(orc_world* w, double dt);
void orc_advance_positions(orc_world* w, double dt);
void orc_advance_positions_post(orc_world* w, double dt);

/* Math.sin / Math.cos as expRodrigues uses them, by fdlibm 5.3's algorithm
 * (java.lang.StrictMath's definition). */
double orc_sin(double x);
double orc_cos(double x);

/* MotionMetricProcessor.getLargestVelocityNorm; b < 0 -> one-body variant. */
double orc_motion_metric(const orc_world* w, int32_t a, int32_t b);
/* the position-level variant (MotionMetricProcessor.java:75-116), both bodies
 * unpinned entities */
double orc_motion_metric_pos(const orc_world* w, int32_t a, int32_t b, double dt);

/* Full step without merging (RigidBodySystem.advanceTime with merging,
 * sleeping and post-stabilization off), sweeping contacts in list order
 * unless order_cb supplies an order. Returns number of contacts. */
int32_t orc_step(orc_world* w, double dt, int32_t* iterations_used, double* last_change);

#ifdef __cplusplus
}
#endif
#endif
