/*
 * am.h - C ABI of libam, the MI355X (gfx950) implementation of the per-timestep
 * 3D contact path of AdaptiveMerging (broad phase, narrow phase, PGS friction-box
 * solve, warm start, relative-motion metric and merge/unmerge checks).
 *
 * Conventions
 *  - Every entry point returns int: AM_OK (0) or a negative AM_ERR_* code.
 *    am_last_error(ctx) returns a description of the last failure.  No C++
 *    exception and no abort crosses this boundary.
 *  - Host buffers are caller-owned and copied; device memory is owned by the
 *    context.  One HIP stream per context.  A context is not thread-safe (the
 *    reference is single-threaded too); independent contexts are.
 *  - All floating point is IEEE double.  Matrices are row-major 3x3
 *    (m00 m01 m02 m10 ... m22), the javax.vecmath Matrix3d field order.
 *  - Body indices are positions in the am_body_desc array passed to
 *    am_scene_upload ("physical" bodies).  Collections (merged bodies) are
 *    extra entities after them, created by am_merge / am_unmerge (see
 *    am_structure_pull).
 *
 * Reference interfaces replaced (paths under src/mergingBodies3D/):
 *   CollisionProcessor.collisionDetection   CollisionProcessor.java:91-102
 *   CollisionProcessor.updateContactsMap    CollisionProcessor.java:443-448
 *   CollisionProcessor.warmStart            CollisionProcessor.java:481-640
 *   CollisionProcessor.redoWarmStart        CollisionProcessor.java:453-459
 *   CollisionProcessor.noWarmStart          CollisionProcessor.java:464-470
 *   CollisionProcessor.solveLCP / PGS.solve CollisionProcessor.java:108-137,
 *                                           PGS.java:73-194
 *   CollisionProcessor.updateInCollections  CollisionProcessor.java:232-303
 *   MotionMetricProcessor.getMotionMetric   MotionMetricProcessor.java:15-116
 *   BodyPairContact.accumulateFor*          BodyPairContact.java:83-145
 *   BodyPairContact merge/unmerge tests     BodyPairContact.java:151-266
 *   Merging.merge / Merging.unmerge         Merging.java:73-393
 *   RigidCollection (mass properties,       RigidCollection.java:55-937
 *     velocities, member transforms)
 *   RigidBody.advanceVelocities/Positions   RigidBody.java:409-441
 *   RigidBodySystem.advanceTime             RigidBodySystem.java:102-185
 */
#ifndef AM_H
#define AM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AM_OK 0
#define AM_ERR_INVALID (-1)   /* bad argument or handle */
#define AM_ERR_HIP (-2)       /* HIP runtime failure */
#define AM_ERR_CAPACITY (-3)  /* a buffer would overflow; grown, call again */
#define AM_ERR_STATE (-4)     /* call out of order (e.g. solve before detect) */
#define AM_ERR_NUMERIC (-5)   /* NaN/Inf in body state */
#define AM_ERR_MISMATCH (-6)  /* two launch shapes of one kernel gave different bits */

enum am_geom {
    AM_GEOM_BOX = 0,        /* RigidBodyGeomBox */
    AM_GEOM_TREE = 1,       /* sphere tree: RigidBodyGeomSphere (1 leaf) or mesh (.sph) */
    AM_GEOM_PLANE = 2,      /* PlaneRigidBody, always pinned */
    AM_GEOM_COLLECTION = 3, /* RigidCollection entity (no own geometry) */
    AM_GEOM_COMPOSITE = 4   /* RigidBodyGeomComposite: tree_first/tree_count select its sub-bodies
                               in the sub-body table of am_scene_upload_ex */
};

enum am_body_flags {
    AM_BODY_PINNED = 1,        /* RigidBody.pinned */
    AM_BODY_MAGNET_ACTIVE = 2, /* magnetic && activateMagnet: no clamping in PGS */
    AM_BODY_SLEEPING = 4,      /* RigidBody.sleeping (set and cleared by am_sleep / am_wake) */
    AM_BODY_KINEMATIC = 8      /* pinned body whose pose and velocity follow a prescribed motion
                                  (am_kinematic_set); no reference counterpart.  Such a body never
                                  merges and never sleeps */
};

enum am_contact_state { AM_STATE_BROKEN = 0, AM_STATE_ONEDGE = 1, AM_STATE_CLEAR = 2 };

/* bv ids carried by contacts: leaf node index (>=0) within the body's tree,
 * AM_BV_NONE for primitives with no bounding volume, AM_BV_PLANE for the
 * plane's dummy bounding volume (CollisionProcessor.java:817). */
#define AM_BV_NONE (-1)
#define AM_BV_PLANE (-2)
/* A contact on sub-body s of a composite (Contact.csb1/csb2) carries that
 * sub-body in its bv id: AM_BV_SUB(s, bv) with bv the sub-body's own id
 * (AM_BV_NONE for a box, its leaf node for a sphere); decode with
 * s = (-3 - id) >> 16, bv = ((-3 - id) & 0xffff) - 1. */
#define AM_BV_SUB(s, bv) (-3 - ((s) << 16) - ((bv) + 1))

/* Immutable description of one physical body (RigidBody fields that do not
 * change per step; XMLParser.createBox/createSphere/createPlane). */
typedef struct am_body_desc {
    int32_t geom;        /* am_geom */
    int32_t flags;       /* am_body_flags */
    int32_t tree_first;  /* TREE: index of this body's root in the node array */
    int32_t tree_count;  /* TREE: number of nodes */
    int32_t scene;       /* independent scene id (batched worlds); 0 for one scene */
    int32_t reserved;
    double size[3];      /* BOX: full edge lengths */
    double radius;       /* RigidBody.radius: box half-diagonal, 0 otherwise */
    double mass;         /* massLinear (0 when pinned) */
    double minv;         /* 1/massLinear, 0 when pinned */
    double jinv0[9];     /* rest-pose inverse rotational inertia */
    double friction;     /* RigidBody.friction (0.8 default) */
    double restitution;  /* RigidBody.restitution (0 default) */
    double plane_n[3];   /* PLANE: unit normal */
    double plane_p[3];   /* PLANE: point */
    double plane_d;      /* PLANE: -(p.n) */
    double bbox[24];     /* boundingBoxB: 8 body-frame points, reference order */
    double inertia0[9];  /* massAngular0: rest-pose rotational inertia (0 when pinned) */
} am_body_desc;

/* One BVSphere node.  Children of a node are stored contiguously; indices are
 * relative to the owning body's tree_first. Node 0 of a body is the root. */
typedef struct am_sphere_node {
    double c[3];         /* centre in body coordinates (BVSphere.cB) */
    double r;            /* radius */
    int32_t first_child; /* -1 for a leaf */
    int32_t n_children;  /* 0 for a leaf */
} am_sphere_node;

/* Solver and merge parameters; defaults (am_params_default) are the
 * reference defaults, CollisionProcessor.java:1058-1083, Merging.java:31-48,
 * Contact.java:518, RigidBodySystem.java:545-551. */
typedef struct am_params {
    int32_t iterations;               /* 30 */
    int32_t iterations_in_collection; /* 1 */
    double tolerance;                 /* 1e-5 */
    double omega;                     /* 1 */
    double feedback_stiffness;        /* 0.5 */
    int32_t enable_post_stabilization;/* 0 */
    int32_t enable_compliance;        /* 1 */
    double compliance;                /* 1e-3 */
    int32_t restitution_override;     /* 0 */
    int32_t friction_override;        /* 0 */
    double restitution;               /* 0.5 (used only when overridden) */
    double friction;                  /* 0.1 (used only when overridden) */
    double sliding_threshold;         /* 0.01 */
    int32_t warm_start;               /* 1 */
    int32_t use_gravity;              /* 1 */
    double gravity_amount;            /* 1 */
    double gravity_angle_deg;         /* 90 */
    double viscous_linear_decay;      /* 1 */
    double viscous_angular_decay;     /* 1 */
    double spring_stiffness_mod;      /* 1 (RigidBodySystem.springStiffnessMod) */
    double spring_damping_mod;        /* 1 (RigidBodySystem.springDampingMod) */
    /* sleeping (Sleeping.SleepParameters, Sleeping.java:25-30) */
    int32_t enable_sleeping;          /* 1 */
    int32_t sleep_step_accum;         /* 10 (0..200): length of each body's metric history */
    double sleep_threshold;           /* 1e-5 */
    /* merging (Merging.MergeParameters) */
    int32_t enable_merging;           /* 1 */
    int32_t enable_merge_pinned;      /* 1 */
    int32_t enable_merge_stable_contact; /* 1 */
    int32_t enable_merge_let_it_breathe; /* 1 */
    int32_t enable_unmerging;         /* 1 */
    int32_t enable_unmerge_friction;  /* 1 */
    int32_t enable_unmerge_normal;    /* 1 */
    int32_t enable_unmerge_relative_motion; /* 1 */
    int32_t update_contacts_in_collections; /* 1 */
    int32_t step_accum_merging;       /* 3 */
    int32_t step_accum_unmerging;     /* 3 */
    int32_t steps_between_merge_events; /* 10 */
    int32_t metric_position_level;    /* 0: the merge and unmerge metrics compare the bodies' poses advanced
                                         by dt instead of their velocities (Merging.metricPositionLevel) */
    double threshold_merge;           /* 1e-2 */
    double threshold_unmerge;         /* 2e-2 */
    double threshold_breath;          /* 1e-5 */
    /* implementation knobs (no reference counterpart) */
    int32_t pgs_block;                /* threads per workgroup in the PGS kernel (0 = auto) */
    int32_t pgs_grid;                 /* workgroups of the PGS kernel (0 = auto) */
    int32_t pgs_plain_launch;         /* 1: launch the PGS kernel with an ordinary (not cooperative) launch,
                                         grid clamped to the co-resident limit; for profilers that do not
                                         support cooperative launches (0 = cooperative) */
    int32_t pgs_schedule;             /* 0: dataflow (each body's deltaV handed from update to update, no grid barriers);
                                         1: colour by colour with a grid barrier between colours.
                                         Both give bit-identical results (the exported sweep order) */
    /* learned warm start (no reference counterpart): 0 plain previous-frame
     * start; 1 the MLP's predicted start, falling back to the plain start
     * whenever its start residual is not smaller */
    int32_t learned_warm_start;
    int32_t collect_training;         /* 1: each main solve that reaches the tolerance appends
                                         (features, converged lambda) to the training set */
    int32_t organize_contacts;        /* 1 (Merging.organizeContacts): the single sweep inside collections
                                         visits the body pairs breadth first from those holding a new
                                         contact (getOrganizedContacts); 0: colour order only */
    int32_t host_threads;             /* host threads a merge or unmerge event may split its per-collection
                                         work over (0 = up to 16; 1 = the calling thread only); the
                                         results do not depend on it (no reference counterpart) */
} am_params;

/* A spring (Spring.java:10-104).  ZERO pulls body1's point pb1 towards the
 * world point it started at (rest length l0 = 0.5 unless set), WORLD towards
 * the world point pw, BODYBODY joins pb1 on body1 and pb2 on body2.  pb1/pb2
 * are in the bodies' frames; pw and l0 are what the constructors computed
 * from the initial poses.  k stiffness, d damping, ls rest-length scale. */
enum {
    AM_SPRING_ZERO = 0,
    AM_SPRING_WORLD = 1,
    AM_SPRING_BODYBODY = 2,
};
typedef struct am_spring {
    int32_t type;
    int32_t body1;
    int32_t body2;        /* BODYBODY only, else -1 */
    int32_t pad;
    double pb1[3];
    double pb2[3];
    double pw[3];
    double l0;
    double k;
    double d;
    double ls;
} am_spring;

/* Caller-owned destination arrays for am_contacts_pull; any may be NULL.
 * Vectors are 3 doubles per contact, the Jacobian 36 per contact in the
 * order jna.v jna.w jnb.v jnb.w jt1a.v jt1a.w jt1b.v jt1b.w jt2a.v jt2a.w
 * jt2b.v jt2b.w (Contact.java:61-66). */
typedef struct am_contact_arrays {
    int32_t* body1;
    int32_t* body2;
    int32_t* bv1;
    int32_t* bv2;
    int32_t* info;
    double* point_b1;      /* contactB1 */
    double* normal_b1;     /* normalB1 */
    double* tangent1_b1;
    double* tangent2_b1;
    double* violation;     /* constraintViolation */
    double* prev_violation;
    double* jacobian;
    double* lambda;        /* lambda0..2 */
    double* lambda_warm;   /* lambda0warm..2 */
    double* b;             /* bn bt1 bt2 */
    double* diag;          /* D00 D11 D22 */
    int32_t* state;        /* am_contact_state */
    int32_t* new_this_step;/* newThisTimeStep */
} am_contact_arrays;

/* Per-step timings (ms, from HIP events) and counters of am_step. */
typedef struct am_step_stats {
    int32_t n_contacts;
    int32_t n_pairs;          /* broad-phase candidate pairs */
    int32_t n_groups;         /* solve groups (body pairs) */
    int32_t n_colours;
    int32_t iterations_used;  /* PGS sweeps done */
    int32_t bad_warm_starts;
    int32_t bad_warm_starts_repaired;
    int32_t reserved;
    double last_change;       /* max |delta lambda| of the last sweep */
    double ms_detect;
    double ms_warm_start;
    double ms_solve;
    double ms_integrate;
    double ms_total;
    double ms_pgs_kernel;     /* duration of the persistent PGS kernel alone */
    double ms_setup_kernel;   /* b, D and record packing in sweep order */
    double ms_learned;        /* features + MLP inference + start selection */
    double start_residual_plain;   /* start residual of the plain warm start */
    double start_residual_learned; /* start residual of the predicted start (0 if unused) */
    int32_t used_learned;     /* 1 if the solve started from the predicted lambdas */
    int32_t n_collections;    /* collections after the step */
    int32_t n_merged;         /* body pairs merged by this step (Merging.merge) */
    int32_t n_unmerged;       /* collections cut by this step (Merging.unmerge) */
    int32_t n_internal_contacts; /* contacts inside collections */
    int32_t n_sleeping;       /* top-level entities asleep after the step */
    double ms_merging;        /* single sweep, unmerge, pair histories and merge */
} am_step_stats;

/* One body pair (BodyPairContact) as am_body_pairs_pull reports it. */
typedef struct am_body_pair {
    int32_t body1, body2;     /* physical bodies */
    int32_t in_collection;    /* BodyPairContact.inCollection */
    int32_t n_contacts;       /* active contacts (external) or internal contacts */
    int32_t hist_n;           /* motionMetricHist size */
    int32_t state_n;          /* contactStateHist size (external pairs) */
    int32_t flag;             /* merge (external) or unmerge (internal) condition */
    int32_t collection;       /* owning collection entity (internal pairs), else -1 */
    double hist[8];           /* motionMetricHist, oldest first (external pairs) */
    int32_t states[8];        /* contactStateHist, oldest first (external pairs) */
} am_body_pair;

typedef struct am_ctx am_ctx;

/* ---- context ---------------------------------------------------------- */
const char* am_version(void);
int am_ctx_create(int32_t device, am_ctx** out);
int am_ctx_destroy(am_ctx* ctx);
const char* am_last_error(const am_ctx* ctx);
int am_params_default(am_params* out);
int am_params_set(am_ctx* ctx, const am_params* params);

/* ---- scene and state (RigidBody fields, XMLParser output) ------------- */
/* Uploads bodies and sphere-tree nodes; resets contacts and warm-start data.
 * Replaces XMLParser.parse -> system.bodies (XMLParser.java:40-67). */
int am_scene_upload(am_ctx* ctx, const am_body_desc* bodies, int32_t n_bodies,
                    const am_sphere_node* nodes, int32_t n_nodes);
/* am_scene_upload with composite bodies: subs[n_subs] are the sub-bodies
 * (boxes, or spheres as one-node trees in the node array) of every
 * AM_GEOM_COMPOSITE body, sub_b2c[12 * n_subs] their transforms into the
 * composite's frame (RigidBody.transformB2C, XMLParser.java:336-343): R
 * row-major, then t. */
int am_scene_upload_ex(am_ctx* ctx, const am_body_desc* bodies, int32_t n_bodies, const am_sphere_node* nodes,
                       int32_t n_nodes, const am_body_desc* subs, const double* sub_b2c, int32_t n_subs);
/* Pose and velocity of every entity (n = bodies + collections).  jinv is
 * recomputed as R jinv0 R^T (RigidBody.updateRotationalInertiaFromTransformation,
 * RigidBody.java:311-321).  R may be NULL to keep the current rotation. */
int am_state_push(am_ctx* ctx, const double* x, const double* R, const double* v,
                  const double* w, int32_t n);
int am_state_pull(am_ctx* ctx, double* x, double* R, double* v, double* w, int32_t n);
/* Merged-body structure.  Entities are the physical bodies followed by
 * collection slots; n_entities grows when merging needs more slots.
 * parent[e] is the collection of entity e or -1; top[0..n_top) is the
 * broad-phase list (RigidBodySystem.bodies order).  Either array may be NULL. */
int am_structure_pull(am_ctx* ctx, int32_t* parent, int32_t cap_entities, int32_t* top, int32_t cap_top,
                      int32_t* n_entities, int32_t* n_top);
/* Descriptions of entities [0, n): for collections the mass properties that
 * RigidCollection computed (mass, minv, jinv0, inertia0, bbox, pinned). */
int am_entities_pull(am_ctx* ctx, am_body_desc* out, int32_t n);
int am_jinv_pull(am_ctx* ctx, double* jinv, int32_t n);
/* Force/torque accumulators (RigidBody.force/torque). */
int am_forces_push(am_ctx* ctx, const double* force, const double* torque, int32_t n);
int am_forces_pull(am_ctx* ctx, double* force, double* torque, int32_t n);
/* RigidBody.clear for all entities: zero force, torque and deltaV (RigidBody.java:226-230). */
int am_clear_bodies(am_ctx* ctx);
/* RigidBodySystem.applyGravityForce (RigidBodySystem.java:276-290). */
int am_apply_gravity(am_ctx* ctx);
/* Replaces RigidBodySystem.springs (RigidBodySystem.java:34; filled by
 * XMLParser.setCommonAttributes :596-619).  Bodies are physical body indices. */
int am_springs_set(am_ctx* ctx, const am_spring* springs, int32_t n);
/* RigidBodySystem.applySpringForces (RigidBodySystem.java:309-322) with
 * Spring.apply (Spring.java:143-214): in list order, each spring adds its force
 * and torque to its bodies (and to their collections).  am_step calls it right
 * after the gravity, as applyExternalForces does (:234-257). */
int am_apply_spring_forces(am_ctx* ctx);
int am_deltav_pull(am_ctx* ctx, double* dv, int32_t n);   /* 6 per entity: v then w */
/* A force at a world point, held until the next external-force pass (am_step
 * runs one after the gravity and the springs), which adds it to the body's
 * top-level entity as RigidBody.applyForceW does and then drops it: the
 * impulse tool's stored force, RigidBodySystem.applyImpulse
 * (RigidBodySystem.java:262-267, called from applyExternalForces :234-257).
 * Forces added before one pass are applied in the order added. */
int am_impulse_add(am_ctx* ctx, int32_t body, const double* point_w, const double* force);
/* RigidBodySystem.applyImpulse (:262-267) on its own: applies the held forces
 * and drops them (am_step does this inside its external-force pass). */
int am_apply_impulses(am_ctx* ctx);
/* RigidBodySystem.applyViscousDecay (:428-435) on its own (am_step does it last). */
int am_apply_viscous_decay(am_ctx* ctx);

/* ---- independent scenes (no reference counterpart) ----------------------
 * Bodies carry a scene id (am_body_desc.scene, 0 .. 2^20 - 1); bodies of
 * different scenes never collide, so one context steps many worlds at once.
 * With more than one scene and a tolerance > 0, the main solve stops each
 * scene on its own max |delta lambda| (PGS.java:190): a scene under the
 * tolerance skips the remaining sweeps, exactly as if it were solved alone
 * in the exported sweep order restricted to its contacts.  This entry point
 * reports, per scene of the last such solve, the sweeps used and the max
 * |delta lambda| of its last sweep (either array may be NULL; capacity >=
 * n_scenes). */
int am_scene_iterations_pull(am_ctx* ctx, int32_t* used, double* last, int32_t capacity, int32_t* n_scenes);

/* ---- prescribed motion (no reference counterpart) ----------------------
 * A kinematic body is a pinned body (AM_BODY_PINNED | AM_BODY_KINEMATIC in
 * its am_body_desc) that moves on a sinusoid: at time t its position is
 * x0 + (amplitude sin(omega t + phase)) dir and its velocity
 * (amplitude omega cos(omega t + phase)) dir, its rotation stays and its
 * angular velocity is 0.  Pinned, it has no mass in the solve, but its
 * velocity enters b = J (v + dt M^-1 f) as every body's does
 * (Contact.computeB, Contact.java:279-330), so bodies resting on it are
 * carried along.  The reference's nearest scenes hang platforms on springs
 * (scenes3D/dominosPlatforms.xml:141-152); its animate tool sets a body's
 * velocity at the same point of the step (Animation.apply, called from
 * RigidBodySystem.advanceTime :109). */
typedef struct am_kinematic {
    int32_t body;         /* physical body, flagged AM_BODY_PINNED | AM_BODY_KINEMATIC */
    int32_t pad;
    double x0[3];
    double dir[3];
    double amplitude;
    double omega;         /* rad/s */
    double phase;         /* rad */
} am_kinematic;
/* Replaces the table of prescribed motions. */
int am_kinematic_set(am_ctx* ctx, const am_kinematic* k, int32_t n);
/* Sets every kinematic body's pose and velocity at time t.  am_step calls it
 * right after clearing the bodies, with t = the simulation time. */
int am_kinematic_update(am_ctx* ctx, double t);
/* RigidBodySystem.simulationTime: 0 after am_scene_upload, dt added at the end
 * of every am_step (RigidBodySystem.java:176). */
int am_sim_time_get(am_ctx* ctx, double* t);
int am_sim_time_set(am_ctx* ctx, double t);

/* ---- detection (CollisionProcessor) ----------------------------------- */
int am_update_contacts_map(am_ctx* ctx);
int am_detect(am_ctx* ctx, double dt, int32_t* n_contacts);
/* Installs a contact list directly (as Contact.set would build it), for
 * callers that bring their own contacts (PGS used on its own). */
int am_contacts_push(am_ctx* ctx, int32_t n, const int32_t* body1, const int32_t* body2,
                     const double* point_w, const double* normal_w, const double* violation,
                     const int32_t* info, const int32_t* bv1, const int32_t* bv2);
int am_contacts_pull(am_ctx* ctx, am_contact_arrays* out, int32_t capacity, int32_t* n);
/* Sets lambda (and lambda_warm) of the current contacts, 3 per contact. */
int am_lambdas_push(am_ctx* ctx, const double* lambda, int32_t n);

/* ---- warm start and solve -------------------------------------------- */
int am_warm_start(am_ctx* ctx, int32_t post_stabilization, int32_t* bad, int32_t* repaired);
int am_redo_warm_start(am_ctx* ctx);
int am_no_warm_start(am_ctx* ctx);
/* solveLCP: coloured PGS.  The sweep order is the colour-sorted order
 * returned by am_solve_order_pull; it is the order a sequential Gauss-Seidel
 * sweep must use to reproduce the result bit for bit. */
int am_solve(am_ctx* ctx, double dt, int32_t post_stabilization, int32_t* iterations_used,
             double* last_change);
int am_solve_order_pull(am_ctx* ctx, int32_t* order, int32_t capacity, int32_t* n);
/* Colour of every contact in the current layout (64 = sequential tail). */
int am_colours_pull(am_ctx* ctx, int32_t* colour_of_contact, int32_t capacity, int32_t* n_colours);
/* CollisionProcessor.updateInCollections (CollisionProcessor.java:232-303):
 * one PGS sweep over the external and internal contacts with sub-body
 * Jacobians (computeInCollection), in colour order over physical body pairs
 * (am_solve_order_pull gives it right after the call); then the members of
 * collections advance their velocities and top-level deltaV is cleared. */
int am_update_in_collections(am_ctx* ctx, double dt, int32_t* iterations_used);
/* PGS.changes: log(max |delta lambda|) per sweep of the last am_solve. */
int am_changes_pull(am_ctx* ctx, double* changes, int32_t capacity, int32_t* n);

/* ---- integration (RigidBody.advanceVelocities / advancePositions) ----- */
int am_advance_velocities(am_ctx* ctx, double dt);
int am_advance_positions(am_ctx* ctx, double dt);
/* RigidBody.advancePositionsPostStabilization (RigidBody.java:423-425) over
 * the top-level entities that are neither pinned nor sleeping
 * (RigidBodySystem.postStabilization, :374-377): deltaV's parts as the
 * velocities.  am_step runs the whole post-stabilization pass when
 * enable_post_stabilization is set, with merging and unmerging off. */
int am_advance_positions_post(am_ctx* ctx, double dt);
/* One RigidBodySystem.advanceTime (RigidBodySystem.java:102-185): clear,
 * gravity, contacts map, detection, warm start, the single sweep inside
 * collections, unmerging, the main solve, body-pair histories, velocity and
 * position update, every steps_between_merge_events steps merging, and
 * sleeping (am_wake after the warm start, am_sleep last).  The merging phases
 * run only when enable_merging / enable_unmerging are set or collections
 * exist; the sleeping phases only when enable_sleeping is set.  Kinematic
 * bodies take their pose at the simulation time right after the bodies are
 * cleared; the step ends with RigidBodySystem.applyViscousDecay (:173,
 * :428-435): the velocities of the top-level entities scaled by
 * viscous_linear_decay and viscous_angular_decay. */
int am_step(am_ctx* ctx, double dt, am_step_stats* stats);

/* ---- adaptive merging (Merging, BodyPairContact, RigidCollection) ------
 * The phases of RigidBodySystem.advanceTime (RigidBodySystem.java:102-185)
 * that am_step runs when merging is on; callers that drive the phases one by
 * one call them in this order. */
/* RigidBodySystem.accumulateForUnmerging (:318-329); dt is used by the
 * position-level metric (must be > 0 when metric_position_level is set) */
int am_accumulate_for_unmerging(am_ctx* ctx, double dt);
/* Merging.unmerge (Merging.java:215-273); n_cut = collections cut */
int am_unmerge(am_ctx* ctx, double dt, int32_t* n_cut);
/* RigidBody.clear of the top-level entities and the external forces again,
 * as advanceTime does after a merging event (RigidBodySystem.java:138-142) */
int am_clear_top_level(am_ctx* ctx);
/* CollisionProcessor.clearBodyPairContacts (:213-226) with the body-pair
 * bookkeeping of updateBodyPairContacts (:145-206) */
int am_clear_body_pair_contacts(am_ctx* ctx);
/* RigidBodySystem.accumulateForMerging (:331-334); dt as above */
int am_accumulate_for_merging(am_ctx* ctx, double dt);
/* Merging.merge (Merging.java:73-163); n_merged = body pairs merged */
int am_merge(am_ctx* ctx, int32_t* n_merged);
/* ---- sleeping (Sleeping.java) ------------------------------------------
 * am_step calls am_wake right after the warm start and am_sleep at the end
 * of the step (RigidBodySystem.java:128,170) when enable_sleeping is set.
 * A sleeping top-level body is not integrated, keeps no contacts with pinned
 * bodies (CollisionProcessor.java:693), and a sleeping collection is left out
 * of the single sweep and of unmerging. */
/* Sleeping.wake (Sleeping.java:107-139): a sleeping body touching a body that
 * is not pinned wakes, with that body; so do the two ends of a body-body
 * spring when one end sleeps and the other does not.  n_woken = entities
 * woken. */
int am_wake(am_ctx* ctx, int32_t* n_woken);
/* Sleeping.sleep (Sleeping.java:49-98): every awake top-level body without an
 * active contact to a body that is not pinned appends its motion metric to
 * its history; it falls asleep once the last sleep_step_accum values are all
 * at most sleep_threshold and none exceeds its predecessor by more than 5e-5.
 * n_asleep = top-level entities asleep afterwards. */
int am_sleep(am_ctx* ctx, int32_t* n_asleep);
/* Sleeping.wakeAll (Sleeping.java:144-150) */
int am_wake_all(am_ctx* ctx);
/* Per entity: the sleeping flag, the history length and the history, oldest
 * first, AM_SLEEP_HIST_STRIDE doubles per entity (entries past the length are
 * stale).  Any of the three arrays may be NULL. */
#define AM_SLEEP_HIST_STRIDE 201
int am_sleep_state_pull(am_ctx* ctx, uint8_t* sleeping, int32_t* hist_n, double* hist, int32_t n);

/* Merging.unmergeAll (Merging.java:168-206), what the "unmerge all" request
 * does at the start of a step: every awake collection is dissolved, its
 * bodies return to the top-level list with the collection's velocity, its
 * body pairs restart their motion history and its internal contacts are
 * dropped.  n_dissolved = collections dissolved. */
int am_unmerge_all(am_ctx* ctx, int32_t* n_dissolved);
/* External pairs of the last am_clear_body_pair_contacts, then internal pairs. */
int am_body_pairs_pull(am_ctx* ctx, am_body_pair* out, int32_t capacity, int32_t* n_external, int32_t* n_internal);
/* RigidCollection.internalContacts of every collection, pair by pair. */
int am_internal_contacts_pull(am_ctx* ctx, am_contact_arrays* out, int32_t capacity, int32_t* n);

/* ---- learned warm start (no reference counterpart) --------------------
 * A per-contact MLP 24 -> 64 -> 64 -> 64 -> 3 (ReLU), BF16 weights on the
 * matrix cores with FP32 accumulation, predicts lambda / S from 24 contact
 * features; parameters are a flat FP32 vector W1 (24x64, row-major [in][out]),
 * b1 (64), W2 (64x64), b2, W3 (64x64), b3, W4 (64x3), b4 (3). */
#define AM_MLP_PARAMS 10115
#define AM_MLP_FEATURES 24
int am_mlp_init(am_ctx* ctx, uint64_t seed);
int am_mlp_set_params(am_ctx* ctx, const float* params, int32_t n);
int am_mlp_get_params(am_ctx* ctx, float* params, int32_t n);
int am_mlp_forward(am_ctx* ctx, const float* x, int32_t n, float* y);
/* d(mean squared error)/d(params) over n rows, same flat layout */
int am_mlp_gradients(am_ctx* ctx, const float* x, const float* t, int32_t n, float* grads, float* loss);
/* features and lambda scale S of the current contacts (list order) */
int am_features_pull(am_ctx* ctx, float* x, double* scale, int32_t capacity, int32_t* n);
int am_train_size(am_ctx* ctx, int32_t* n);
int am_train_clear(am_ctx* ctx);
/* Adam over the collected (features, lambda / S) pairs */
int am_train_epochs(am_ctx* ctx, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss);
/* lambda each contact started the last main solve from (list order) */
int am_start_lambdas_pull(am_ctx* ctx, double* lambda, int32_t capacity, int32_t* n);

/* ---- relative motion (MotionMetricProcessor.getLargestVelocityNorm) --- */
/* For each pair (a,b) of entity ids: the largest relative velocity over both
 * bodies' bounding-box corners; b < 0 means the one-body variant. */
int am_motion_metric(am_ctx* ctx, const int32_t* pairs, int32_t n_pairs, double* out);
/* am_motion_metric without any device allocation; no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, hosts that forbid hipMalloc): the pairs
 * and results are staged in a fixed device array of AM_METRIC_NOALLOC_MAX
 * pairs (1 MiB per GPU, held until the process ends).  More pairs fail with
 * AM_ERR_CAPACITY.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (from several threads or contexts) race. */
#define AM_METRIC_NOALLOC_MAX 65536
int am_motion_metric_noalloc(am_ctx* ctx, const int32_t* pairs, int32_t n_pairs, double* out);

/* ---- launch-shape autotuner (no reference counterpart) ----------------
 * Times the PGS kernel on the contact set currently held (after am_detect
 * and a warm start, or after an am_step) under each candidate launch shape:
 * the dataflow schedule at several wave counts and the coloured schedule at
 * several workgroup sizes and grids.  Every trial runs `sweeps` sweeps with
 * no early exit from the same lambdas and deltaV; the lambdas and deltaV of
 * every trial must equal the first trial's bit for bit, else the call fails
 * with AM_ERR_MISMATCH.  The contacts, lambdas, deltaV and the
 * last-solve statistics are restored afterwards.  *best is the fastest
 * trial, except that the first trial (the automatic dataflow shape) is kept
 * unless another is at least 2 % faster.  With apply != 0 that shape is
 * written into the context's pgs_schedule / pgs_grid / pgs_block;
 * am_params_set overrides it later like any other parameter. */
typedef struct am_tune_trial {
    int32_t pgs_schedule;    /* 0 dataflow, 1 coloured */
    int32_t pgs_grid;        /* as requested (0 = the automatic choice) */
    int32_t pgs_block;       /* as requested (0 = the automatic choice) */
    int32_t grid_used;       /* workgroups launched */
    int32_t block_used;      /* threads per workgroup launched */
    int32_t reserved;
    double ms_per_sweep;     /* best of `reps` kernel times / sweeps */
} am_tune_trial;
int am_autotune(am_ctx* ctx, double dt, int32_t sweeps, int32_t reps, int32_t apply, am_tune_trial* trials,
                int32_t capacity, int32_t* n_trials, int32_t* best);
/* Workgroups and threads per workgroup of the last main solve's PGS kernel
 * (0, 0 before the first solve with contacts). */
int am_launch_shape_pull(am_ctx* ctx, int32_t* grid, int32_t* block);

#ifdef __cplusplus
}
#endif
#endif /* AM_H */
