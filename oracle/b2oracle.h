/*
 * b2oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a scalar C++ restatement of dbox's
 * (the D port of Box2D 2.3.1) per-step hot path, b2World.Step, written to keep
 * the reference's floating-point operation order so that its results can be
 * compared with the GPU path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * it and never falls back to it.
 *
 * Every struct below is a plain C layout shared byte-for-byte with the scene
 * descriptions that the tests hand to the GPU library (dbx_body_def /
 * dbx_fixture_def in include/dbox_hip.h), so one numpy structured array feeds
 * both sides.
 */
#ifndef DBOX_ORACLE_H
#define DBOX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Body definition: mirrors b2BodyDef (src/dbox/dynamics/b2body.d:51-104). */
typedef struct {
    int32_t type; /* 0 static, 1 kinematic, 2 dynamic (b2BodyType) */
    float px, py, angle;
    float vx, vy, w;
    float linear_damping, angular_damping, gravity_scale;
    int32_t allow_sleep, awake, fixed_rotation, bullet;
} orc_body_def; /* 56 bytes */

/* Fixture definition: mirrors b2FixtureDef (src/dbox/dynamics/b2fixture.d:49-73)
 * plus the shape constructor that the scene used.
 *   shape 0 circle : radius, p[0..1] = local centre
 *   shape 1 edge   : p[0..3] = v1, v2; p[4..7] = v0, v3; flags bit1/bit2 = has v0 / has v3
 *   shape 2 polygon: count points in p[0..2*count-1], hulled by b2PolygonShape.Set
 *   shape 3 box    : p[0..1] = hx, hy (b2PolygonShape.SetAsBox);
 *                    flags bit3 = oriented box with centre p[2..3] and angle p[4]
 *   shape 4 chain  : count (2..8) points in p[0..2*count-1]; flags bit4 = loop
 *                    (b2ChainShape.CreateLoop), else CreateChain with flags bit1 / bit2 =
 *                    previous / next ghost vertex in p[16..17] / p[18..19]
 */
typedef struct {
    int32_t body;
    int32_t shape;
    int32_t count;
    int32_t flags; /* bit0 sensor */
    float density, friction, restitution;
    uint16_t category, mask;
    int16_t group, pad;
    float radius;
    float p[20];
} orc_fixture_def; /* 120 bytes */

/* Body state read back after a step (indexed by body creation order). */
typedef struct {
    float px, py, qs, qc;      /* m_xf */
    float cx, cy, a, sleep_time; /* m_sweep.c, m_sweep.a */
    float vx, vy, w;
    int32_t flags;             /* b2Body.m_flags (awake 0x2, ...) */
    float c0x, c0y, a0, inv_mass;
} orc_body_state; /* 64 bytes */

/* A contact read back in world contact-list order (m_contactList head first). */
typedef struct {
    int32_t fixture_a, fixture_b; /* fixture creation indices after the type-order swap */
    int32_t flags;                /* b2Contact.m_flags: touching 0x2, enabled 0x4 */
    int32_t type, point_count;    /* b2Manifold.Type, pointCount */
    float local_normal[2], local_point[2];
    float point_local[4];         /* points[0].localPoint, points[1].localPoint */
    float normal_impulse[2], tangent_impulse[2];
    uint32_t id[2];               /* b2ContactID.key */
    float friction, restitution;
} orc_contact; /* 84 bytes */

/* One b2ContactListener callback recorded during the last orc_step: type 1
 * BeginContact, 2 EndContact (b2contact.d:342-350, b2contactmanager.d:190-193),
 * 3 PostSolve with b2ContactImpulse (b2island.d:438-461). */
typedef struct {
    int32_t type;
    int32_t fixture_a, fixture_b;
    int32_t point_count;
    float normal_impulse[2], tangent_impulse[2];
} orc_event; /* 32 bytes */

/* Manifold produced by one collide call (b2Manifold, b2collision.d:96-114)
 * together with its b2WorldManifold (b2collision.d:123-191). */
typedef struct {
    int32_t type, point_count, swapped, pad;
    float local_normal[2], local_point[2];
    float point_local[4];
    uint32_t id[2];
    float normal[2], points[4], separations[2];
} orc_manifold;

/* Joint definition: b2JointDef and the concrete joint definitions
 * (src/dbox/dynamics/joints/b2*joint.d), one record for every kind; the byte layout
 * of dbx_joint_def in include/dbox_hip.h.  type is b2JointType (1 revolute,
 * 2 prismatic, 3 distance, 4 pulley, 5 mouse, 7 wheel, 8 weld, 9 friction, 10 rope,
 * 11 motor; the gear joint, 6, is not supported). */
typedef struct {
    int32_t type, body_a, body_b, collide_connected;
    int32_t enable_limit, enable_motor;
    int32_t joint1, joint2;   /* gear: the two joints it couples, by creation index in the world */
    float local_anchor_a[2], local_anchor_b[2];
    float local_axis_a[2];
    float reference_angle;
    float length;                      /* distance: length; rope: maxLength */
    float frequency_hz, damping_ratio;
    float lower, upper;                /* revolute: angles; prismatic: translations */
    float motor_speed, max_motor;      /* maxMotorTorque (revolute, wheel) / maxMotorForce (prismatic) */
    float max_force, max_torque;       /* friction, motor; mouse: max_force */
    float linear_offset[2], angular_offset, correction_factor; /* motor */
    float ground_anchor_a[2], ground_anchor_b[2], length_a, length_b, ratio; /* pulley */
    float target[2];                   /* mouse */
} orc_joint_def; /* 148 bytes */

void *orc_world_create(float gx, float gy);
void orc_world_destroy(void *w);
/* b2World.SetWarmStarting / SetAllowSleeping / SetContinuousPhysics
 * (continuous = 1 runs b2World.SolveTOI after the discrete solve). */
int orc_world_set_flags(void *w, int warm_starting, int allow_sleep, int continuous);
/* Test hook (not reference behaviour): 1 = solve each island's contacts in greedy
 * colour order, to measure what reordering Gauss-Seidel alone changes. */
int orc_world_set_solve_order(void *w, int mode);
/* Test hook: the contacts the last step solved, in solve order, as (island,
 * bodyA, bodyB, dynamic mask) rows; returns the row count. */
int orc_solve_log(void *w, int32_t *out, int cap);
int orc_create_body(void *w, const orc_body_def *def);
int orc_create_fixture(void *w, const orc_fixture_def *def);
/* Builds a whole scene: bodies in order, then fixtures in order. */
int orc_build(void *w, const orc_body_def *bodies, int nb, const orc_fixture_def *fixtures, int nf);
void orc_step(void *w, float dt, int velocity_iterations, int position_iterations);
int orc_body_count(void *w);
int orc_fixture_count(void *w);
int orc_contact_count(void *w);
void orc_get_bodies(void *w, orc_body_state *out);
/* Contacts in m_contactList order; returns the number written (<= cap). */
int orc_get_contacts(void *w, orc_contact *out, int cap);
/* A body's contact edges (b2Body.m_contactList), head first, as proxy pairs; returns
 * the edge count (pairs beyond cap are not written), -1 for a bad body index. */
int orc_body_contact_pairs(void *w, int body, int *out_pairs, int cap);
/* Listener callbacks of the last orc_step in call order; returns the count. */
int orc_events(void *w, orc_event *out, int cap);
/* World manifold of contact-list entry i at the bodies' current transforms. */
int orc_contact_world_manifold(void *w, int list_index, orc_manifold *out);
/* Fat AABB of each proxy, 4 floats each (lower.x, lower.y, upper.x, upper.y).
 * Proxies are the fixtures' children (one per fixture, one per chain segment)
 * counted in creation order; contacts, events and pairs name fixtures by this
 * proxy index, which is the fixture index when there are no chains. */
void orc_get_fat_aabbs(void *w, float *out);
int orc_proxy_count(void *w);
/* The unique pairs that the most recent UpdatePairs handed to AddPair, as
 * (fixture index, fixture index) with the smaller proxy id first, in the
 * sorted order of b2PairLessThan. Returns the count (<= cap pairs written). */
int orc_last_pairs(void *w, int32_t *out, int cap);
/* b2Profile of the last step: step, collide, solve, solveInit, solveVelocity,
 * solvePosition, broadphase, solveTOI (ms). */
void orc_profile(void *w, float *out8);
/* Number of dynamic bodies (for body-step counting). */
int orc_dynamic_body_count(void *w);
/* Maximum penetration over touching contacts: -min world-manifold separation. */
float orc_max_penetration(void *w);

/* Stand-alone collide call on two shapes built from fixture defs (bodies at
 * (px, py, angle) transforms).  The pair is put in the order b2Contact.Create
 * uses; out->swapped says whether A and B were exchanged.  Returns 0, or -1
 * when the shape pair has no contact type (edge-edge). */
int orc_collide(const orc_fixture_def *a, const float *xfa3, const orc_fixture_def *b,
                const float *xfb3, orc_manifold *out);

/* The chain-and-circle / chain-and-polygon collide of one chain segment
 * (b2ChainShape.GetChildEdge then the edge collide, b2chainandpolygoncontact.d,
 * b2chainandcirclecontact.d); the chain is shape A.  Returns -1 for a bad child
 * index or a shape B that is not a circle or polygon. */
int orc_collide_child(const orc_fixture_def *chain, int child, const float *xfa3, const orc_fixture_def *b,
                      const float *xfb3, orc_manifold *out);
/* Shape geometry exactly as the reference builds it (hull, normals, centroid),
 * for comparing the GPU library's own builder. out: count, radius, centroid(2),
 * vertices(16), normals(16) = 36 floats; mass_out: mass, cx, cy, I. */
int orc_shape_geometry(const orc_fixture_def *def, float *out36, float *mass_out4);

/* b2BroadPhase.UpdatePairs on a b2DynamicTree holding the given fat AABBs,
 * querying the proxies whose moved flag is set.  Returns the pair count; pairs
 * are written as (proxy, proxy) with the smaller index first, sorted. */
int orc_tree_pairs(const float *fat_aabbs, const uint8_t *moved, int n, int32_t *out, int cap);

/* b2Distance (b2distance.d:185-345) between two shapes at transforms (px, py,
 * angle), with an empty simplex cache.  out6: pointA, pointB, distance, iterations. */
int orc_shape_distance(const orc_fixture_def *a, const float *xfa3, const orc_fixture_def *b, const float *xfb3,
                       int use_radii, float *out6);
/* b2TimeOfImpact (b2timeofimpact.d:67-295) of two shapes on sweeps given as 9
 * floats (localCenter.xy, c0.xy, c.xy, a0, a, alpha0).  Returns the output state
 * (0 unknown, 1 failed, 2 overlapped, 3 touching, 4 separated); *t_out = time. */
int orc_time_of_impact(const orc_fixture_def *a, const float *sweep_a9, const orc_fixture_def *b, const float *sweep_b9,
                       float t_max, float *t_out);
/* Test hook: TOI events (sub-step island solves) of the last orc_step. */
int orc_toi_events(void *w);
/* Test hook standing in for b2ContactListener.PreSolve (b2contact.d:352-355):
 * mode 1 records each call as an event of type 4 (fixtures, and the old
 * manifold's point count and impulses); mode 2 also applies a fixed policy by
 * k = (fixtureA + fixtureB) % 7: 0 disable, 1 friction 0.05, 2 restitution 0.6,
 * 3 tangent speed 0.8. */
int orc_world_set_presolve(void *w, int mode);
/* b2World.CreateJoint (b2world.d:196-262) after the bodies it names exist; returns
 * the joint index (creation order) or -1 for an unsupported type or bad bodies. */
int orc_create_joint(void *w, const orc_joint_def *def);
int orc_joint_count(void *w);
/* Per joint in creation order, 8 floats: the impulses the joint keeps between steps
 * (revolute / prismatic / weld: m_impulse x, y, z; distance, rope, pulley, wheel:
 * m_impulse in [0]; friction, motor, mouse: linear impulse in [0..1], angular in [2]),
 * motor impulse, spring impulse (wheel), limit state, and 2 zeros. */
void orc_get_joints(void *w, float *out8);
/* Test switch for b2Rot.Set (not reference behaviour).  Mode 0 (default) calls
 * glibc sinf/cosf as dbox does; mode 1 evaluates sin and cos in double and rounds
 * once (Cody-Waite reduction, Taylor series in Horner form with fused
 * multiply-adds), the device library's evaluation, so that GPU and oracle results
 * can be compared bit for bit.  orc_set_rot_mode sets the process-wide mode (for
 * the stand-alone calls) and returns the previous one; orc_world_set_rot_mode
 * makes orc_step of one world run in mode 0 or 1 (-1: the process-wide mode). */
int orc_set_rot_mode(int mode);
int orc_world_set_rot_mode(void *w, int mode);
/* Test hook: b2ContactSolver.SolveVelocityConstraints (b2contactsolver.d:492-772)
 * on one constructed constraint, its masses computed as InitializeVelocityConstraints
 * does (b2contactsolver.d:384-447).  in30: vA(3) vB(3) invMassA invIA invMassB invIB
 * normal(2) rA1(2) rB1(2) rA2(2) rB2(2) normalImpulse1 tangentImpulse1 normalImpulse2
 * tangentImpulse2 velocityBias1 velocityBias2 friction tangentSpeed pointCount
 * iterations.  out11: vA(3) vB(3) the four impulses, pointCount after the block
 * solver's condition check. */
void orc_solve_one_constraint(const float *in30, float *out11);
/* Test hook: one b2ContactSolver.SolveTOIPositionConstraints pass
 * (b2contactsolver.d:152-242) over contact-list entry list_index, with body toi_body
 * as the TOI body.  out4: its sweep c (2) and a after the pass, minimum separation.
 * Returns -1 when the entry or the body does not exist. */
int orc_toi_position_once(void *w, int list_index, int toi_body, float *out4);
#ifdef __cplusplus
}
#endif
#endif
