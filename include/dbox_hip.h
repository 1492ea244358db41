/*
 * dbox_hip.h -- C ABI of the MI355X b2World.Step hot path.
 *
 * The reference (dbox, the D port of Box2D 2.3.1) has no FFI: its step is the
 * D member function b2World.Step (src/dbox/dynamics/b2world.d:367-434).  This
 * library replaces the body of that call for every world bound to a handle:
 *   broadphase   b2BroadPhase.UpdatePairs + b2ContactManager.AddPair
 *                (src/dbox/collision/b2broadphase.d:139-195, dynamics/b2contactmanager.d:52-176)
 *   narrowphase  b2ContactManager.Collide -> b2Contact.Update -> b2Collide*
 *                (dynamics/b2contactmanager.d:251-317, dynamics/contacts/b2contact.d:270-356)
 *   solver       b2World.Solve -> b2Island.Solve -> b2ContactSolver
 *                (dynamics/b2world.d:930-1124, dynamics/b2island.d:75-280,
 *                 dynamics/contacts/b2contactsolver.d:73-787)
 * The D side keeps b2Body, b2Fixture, b2Shape and b2ContactListener; a hook at
 * the top of b2World.Step calls dbx_step and copies results back.
 *
 * All entry points are plain C: pointers and sizes, no C++ or torch types.
 * Every pointer argument is caller-owned and is read (or written) before the
 * call returns.  Status codes are negative on error; dbx_last_error gives the
 * detail.  A handle is used by one host thread at a time.
 */
#ifndef DBOX_HIP_H
#define DBOX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dbx_status;
enum {
    DBX_OK = 0,
    DBX_E_INVALID_ARG = -1,
    DBX_E_CAPACITY = -2,
    DBX_E_UNSUPPORTED = -3, /* a feature this library does not implement */
    DBX_E_HIP = -4,
    DBX_E_LOCKED = -5,      /* world mutation after dbx_finalize */
    DBX_E_STATE = -6
};

/* dbx_step flags (b2World.SetWarmStarting / SetAllowSleeping / SetContinuousPhysics,
 * b2world.d:616-753, plus the solver schedule). */
enum {
    DBX_STEP_WARM = 1u << 0,    /* warm starting (default on in dbox) */
    DBX_STEP_SLEEP = 1u << 1,   /* allow sleeping (default on in dbox) */
    DBX_STEP_CCD = 1u << 2,     /* continuous physics: b2World.SolveTOI after the discrete solve (default on in dbox) */
    DBX_STEP_ORDERED = 1u << 3, /* solve in the reference's exact island order (dependency levels) */
    DBX_STEP_LEARNED = 1u << 4, /* MLP warm start for new contact points */
    DBX_STEP_RESIDUAL = 1u << 5, /* iterate velocities to a residual instead of a fixed count */
    DBX_STEP_LOG = 1u << 6,      /* append new points' features and solved impulses to the training log */
    DBX_STEP_PRESOLVE = 1u << 7  /* stop after b2ContactManager.Collide: the host runs its PreSolve
                                    callbacks, may change contacts with dbx_write_contact_settings,
                                    and dbx_step_finish runs the rest of the step */
};

/* Body definition, same fields as b2BodyDef (src/dbox/dynamics/b2body.d:51-104). */
typedef struct {
    int32_t type; /* 0 static, 1 kinematic, 2 dynamic */
    float px, py, angle;
    float vx, vy, w;
    float linear_damping, angular_damping, gravity_scale;
    int32_t allow_sleep, awake, fixed_rotation, bullet;
} dbx_body_def;

/* Fixture definition: b2FixtureDef (src/dbox/dynamics/b2fixture.d:49-73) plus the
 * shape constructor it was made with.
 *   shape 0 circle : radius; p[0..1] = local centre            (b2CircleShape)
 *   shape 1 edge   : p[0..3] = v1, v2; p[4..7] = v0, v3;
 *                    flags bit1 / bit2 = has v0 / has v3        (b2EdgeShape.Set)
 *   shape 2 polygon: count points in p[0 .. 2*count-1]          (b2PolygonShape.Set, hulled)
 *   shape 3 box    : p[0..1] = hx, hy; flags bit3 = centre p[2..3], angle p[4]
 *                                                               (b2PolygonShape.SetAsBox)
 *   shape 4 chain  : count (2..8) points in p[0 .. 2*count-1]; flags bit4 = loop
 *                    (b2ChainShape.CreateLoop, count >= 3), else CreateChain with
 *                    flags bit1 / bit2 = previous / next ghost vertex in p[16..17] /
 *                    p[18..19] (SetPrevVertex / SetNextVertex).  A longer chain is
 *                    several chain fixtures joined by ghost vertices.
 * flags bit0 = sensor (b2FixtureDef.isSensor): its contacts touch while the
 * shapes overlap (b2TestOverlap), carry no manifold, wake no body, give
 * BEGIN / END but no PRE_SOLVE / POST_SOLVE and are never solved.
 *
 * Proxy indices.  Every fixture child has a broadphase proxy (b2Fixture.m_proxies):
 * one per fixture, one per segment of a chain.  The read-backs below name the
 * fixtures of pairs, contacts and events by proxy index, counting the children of
 * all fixtures in creation order; without chains that is the fixture index.
 * dbx_read_proxies maps proxy indices to (fixture, child). */
typedef struct {
    int32_t body;
    int32_t shape;
    int32_t count;
    int32_t flags;
    float density, friction, restitution;
    uint16_t category, mask;
    int16_t group, pad;
    float radius;
    float p[20];
} dbx_fixture_def;

/* Joint definition: b2JointDef and the concrete joint definitions
 * (src/dbox/dynamics/joints/b2*joint.d), one record for every kind.  type is
 * b2JointType (b2joint.d:28-41).  A gear joint (b2GearJointDef) names the two
 * revolute or prismatic joints it couples in joint1 / joint2 and its ratio in ratio;
 * its bodies become those joints' bodies B, as in b2GearJoint's constructor, while
 * body_a / body_b must still be valid, distinct bodies.  Fields a kind does not use
 * are ignored. */
enum {
    DBX_JOINT_REVOLUTE = 1, DBX_JOINT_PRISMATIC = 2, DBX_JOINT_DISTANCE = 3, DBX_JOINT_PULLEY = 4,
    DBX_JOINT_MOUSE = 5, DBX_JOINT_GEAR = 6, DBX_JOINT_WHEEL = 7, DBX_JOINT_WELD = 8, DBX_JOINT_FRICTION = 9, DBX_JOINT_ROPE = 10,
    DBX_JOINT_MOTOR = 11
};
typedef struct {
    int32_t type;
    int32_t body_a, body_b;          /* body creation indices */
    int32_t collide_connected;       /* b2JointDef.collideConnected */
    int32_t enable_limit, enable_motor; /* revolute, prismatic (limit, motor); wheel (motor) */
    int32_t joint1, joint2;          /* gear: the revolute or prismatic joints it couples (creation
                                      * indices of this world's joints, both created before it) */
    float local_anchor_a[2], local_anchor_b[2]; /* all but motor and mouse */
    float local_axis_a[2];           /* prismatic (normalised on creation), wheel (as given) */
    float reference_angle;           /* revolute, prismatic, weld */
    float length;                    /* distance: length; rope: maxLength */
    float frequency_hz, damping_ratio; /* distance, weld, wheel, mouse */
    float lower, upper;              /* revolute: lower / upper angle; prismatic: translation */
    float motor_speed, max_motor;    /* revolute, wheel: maxMotorTorque; prismatic: maxMotorForce */
    float max_force, max_torque;     /* friction, motor; mouse: maxForce */
    float linear_offset[2], angular_offset, correction_factor; /* motor */
    float ground_anchor_a[2], ground_anchor_b[2], length_a, length_b, ratio; /* pulley */
    float target[2];                 /* mouse: the target, and the anchor on body B at creation */
} dbx_joint_def; /* 148 bytes */

/* Body state read back after a step: b2Body.m_xf, m_sweep, velocities, m_flags. */
typedef struct {
    float px, py, qs, qc;
    float cx, cy, a, sleep_time;
    float vx, vy, w;
    int32_t flags; /* b2Body flags: awake 0x2, autoSleep 0x4, ... */
    float c0x, c0y, a0, inv_mass;
} dbx_body_state;

/* A contact in b2World.GetContactList order (most recently created first). */
typedef struct {
    int32_t fixture_a, fixture_b; /* world-local fixture creation indices after the type-order swap */
    int32_t flags;                /* touching 0x2, enabled 0x4 */
    int32_t type, point_count;    /* b2Manifold.Type, pointCount */
    float local_normal[2], local_point[2];
    float point_local[4];
    float normal_impulse[2], tangent_impulse[2];
    uint32_t id[2];               /* b2ContactID.key */
    float friction, restitution;
} dbx_contact;

/* One b2ContactListener callback of the last step.  BEGIN and END come from
 * b2Contact.Update (src/dbox/dynamics/contacts/b2contact.d:342-350) and, for END,
 * from b2ContactManager.Destroy of a touching contact (src/dbox/dynamics/
 * b2contactmanager.d:190-193); POST_SOLVE from b2Island.Report
 * (src/dbox/dynamics/b2island.d:438-461), with b2ContactImpulse's count and
 * impulses. */
enum { DBX_EVENT_BEGIN = 1, DBX_EVENT_END = 2, DBX_EVENT_POST_SOLVE = 3, DBX_EVENT_PRE_SOLVE = 4 };
typedef struct {
    int32_t type;                 /* DBX_EVENT_* */
    int32_t fixture_a, fixture_b; /* as in dbx_contact */
    int32_t point_count;          /* POST_SOLVE: b2ContactImpulse.count (1 for a downgraded two-point manifold);
                                     PRE_SOLVE: the old manifold's pointCount */
    float normal_impulse[2], tangent_impulse[2]; /* PRE_SOLVE: the old manifold's impulses */
} dbx_contact_event;

/* What a PreSolve callback may change on a contact (b2Contact.SetEnabled,
 * SetFriction, SetRestitution, SetTangentSpeed: b2contact.d:108-198), for the
 * contact between two fixtures. */
enum { DBX_SET_ENABLED = 1, DBX_SET_FRICTION = 2, DBX_SET_RESTITUTION = 4, DBX_SET_TANGENT_SPEED = 8 };
typedef struct {
    int32_t fixture_a, fixture_b;
    int32_t mask;     /* DBX_SET_*: the fields below that are applied */
    int32_t enabled;
    float friction, restitution, tangent_speed;
} dbx_contact_setting; /* 28 bytes */

/* b2Profile (src/dbox/dynamics/b2timestep.d:37-46), milliseconds of the last step,
 * with b2World.Step's meaning of each field (b2world.d:367-434, 930-1124):
 *   step       the whole step (HIP events on the handle's stream)
 *   collide    b2ContactManager.Collide
 *   solve      all of b2World.Solve: island search, island solves, SynchronizeFixtures
 *              and FindNewContacts (so it includes broadphase)
 *   solveInit / solveVelocity / solvePosition
 *              the phases of b2Island.Solve (b2island.d:118-148: constraint set-up and
 *              warm start; 151-165: velocity iterations and StoreImpulses; 203-237:
 *              position iterations and the copy back), timed inside the solve kernel
 *              of the largest world with the device's constant-rate wall clock; dbox
 *              sums them over islands (b2world.d:1079-1081), here the islands of the
 *              largest world run in one workgroup while the other worlds run beside it
 *   broadphase SynchronizeFixtures + FindNewContacts (b2world.d:1102-1123)
 *   solveTOI   b2World.SolveTOI
 * A step replayed from a hipGraph records step and the three island-solve phases only;
 * split (DBX_STEP_PRESOLVE) steps are not timed.  The island-solve phases of the last
 * 4,096 steps before a dbx_sync are kept; an earlier step of a longer run without a
 * sync reports them as 0. */
typedef struct {
    float step, collide, solve, solveInit, solveVelocity, solvePosition, broadphase, solveTOI;
} dbx_profile;

typedef struct {
    int32_t device;                 /* HIP device ordinal */
    int32_t contacts_per_fixture;   /* contact capacity per fixture (0 = 8) */
    int32_t block_size;             /* threads per world workgroup (0 = autotuned default) */
    int32_t use_graph;              /* replay each step as a captured hipGraph (one per distinct
                                       step parameters and buffer parity) */
} dbx_config;

typedef struct dbx_handle dbx_handle;

/* Create a handle bound to one GPU.  Replaces constructing b2World (b2world.d:865-895). */
dbx_status dbx_create(const dbx_config *config, dbx_handle **out);
void dbx_destroy(dbx_handle *h);
const char *dbx_last_error(const dbx_handle *h);

/* Add one world, built from body and fixture definitions in creation order, as
 * b2World.CreateBody (b2world.d:75-99) and b2Body.CreateFixture (b2body.d:116-155)
 * would build it: hulls, mass data and proxies are computed here on the host. */
dbx_status dbx_add_world(dbx_handle *h, const dbx_body_def *bodies, int32_t body_count,
                         const dbx_fixture_def *fixtures, int32_t fixture_count, float gravity_x,
                         float gravity_y, int32_t *world_id);
/* b2World.CreateJoint (b2world.d:196-262) for n joints of one world, in creation
 * order, between dbx_add_world and dbx_finalize (DBX_E_LOCKED after it).  Islands,
 * sleeping and b2Body.ShouldCollide (collideConnected) then follow the joints as in
 * the reference, and each island's joints are solved with its contacts in
 * b2Island.Solve's sequence (b2island.d:137-224).  DBX_E_INVALID_ARG (and none of
 * the n joints added) for a bad body index, equal bodies, a value that is not
 * finite, or a gear whose joint1 / joint2 is not a revolute or prismatic joint made
 * before it. */
dbx_status dbx_add_joints(dbx_handle *h, int32_t world, const dbx_joint_def *joints, int32_t n);
/* Per joint of one world in creation order, 8 floats: the impulses the joint keeps
 * between steps (revolute, prismatic, weld: m_impulse x, y, z; distance, rope,
 * pulley, wheel, gear: m_impulse in [0]; friction, motor, mouse: linear impulse in [0..1],
 * angular impulse in [2]), the motor impulse, the wheel's spring impulse, the limit
 * state (b2LimitState), 0, 0.  b2Joint.GetReactionForce / GetReactionTorque are
 * these times 1 / dt (b2*joint.d). */
dbx_status dbx_read_joints(dbx_handle *h, int32_t world, float *out, int32_t n);
/* Upload every added world to the device.  After this the world set is fixed. */
dbx_status dbx_finalize(dbx_handle *h);

/* b2World.Step(dt, velocityIterations, positionIterations) (b2world.d:367-434) for
 * every world of the handle, asynchronously on the handle's stream. */
dbx_status dbx_step(dbx_handle *h, float dt, int32_t velocity_iterations, int32_t position_iterations,
                    uint32_t flags);
dbx_status dbx_sync(dbx_handle *h);

dbx_status dbx_world_count(const dbx_handle *h, int32_t *count);
/* body count, fixture count, contact count, touching contacts */
dbx_status dbx_world_info(dbx_handle *h, int32_t world, int32_t out[4]);
dbx_status dbx_read_bodies(dbx_handle *h, int32_t world, dbx_body_state *out, int32_t n);
/* Host-side mutations between steps (b2Body.SetTransform, SetLinearVelocity,
 * SetAngularVelocity, SetAwake, ApplyForce/ApplyTorque: b2body.d:430-454,
 * 570-640, 827-846): overwrites every body's pose, sweep, velocities, sleep
 * time and awake bit (flags & 0x2) from `in` (n = body count, in creation
 * order), and, when force3 is not null, its force and torque (3 floats per body,
 * the accumulators the next step integrates and then clears).  Fixtures whose
 * fat AABB no longer contains the new pose are re-fattened and moved, as
 * b2Fixture.Synchronize(xf, xf) does after SetTransform. */
dbx_status dbx_write_bodies(dbx_handle *h, int32_t world, const dbx_body_state *in, const float *force3, int32_t n);
/* For each proxy of one world (see "Proxy indices" above): its fixture's creation
 * index and its child index.  *count = proxies; nothing is written when cap is smaller. */
dbx_status dbx_read_proxies(dbx_handle *h, int32_t world, int32_t *fixture, int32_t *child, int32_t cap,
                            int32_t *count);
/* Contacts of one world in b2World contact-list order. */
dbx_status dbx_read_contacts(dbx_handle *h, int32_t world, dbx_contact *out, int32_t cap, int32_t *count);
/* The contact-listener callbacks dbox would have made during the last dbx_step of
 * one world, in its order: BEGIN / END as b2ContactManager.Collide walks the
 * contact list (b2contactmanager.d:251-317; a contact destroyed while touching
 * gives END at its place in the list), then POST_SOLVE for every solved contact
 * in island order (b2world.d:1085, b2island.d:239).  With more than one step per
 * dbx_step call only the last step's callbacks are kept.  After a DBX_STEP_PRESOLVE
 * step each touching contact Collide updated also gives PRE_SOLVE right after its
 * BEGIN / END (b2contact.d:342-355); while that step waits for dbx_step_finish only
 * the Collide callbacks exist.  With DBX_STEP_CCD the continuous pass's
 * callbacks follow, in its order (b2world.d:1127-1452): BEGIN / END (and, after a
 * split step, PRE_SOLVE) from each b2Contact.Update of a TOI event, then POST_SOLVE
 * for every contact of the event's sub-step island with the sub-step solver's
 * impulses (b2island.d:415).  PreSolve changes cannot reach the continuous pass:
 * its PRE_SOLVE events are reports only.  Each world's pass logs up to 128 + 4 x
 * its body count of records per step (its callbacks, and one per contact it
 * updates, which keeps that contact's discrete POST_SOLVE data); DBX_E_CAPACITY
 * when its last pass needed more.
 * *count = number of events; DBX_E_CAPACITY when cap is smaller (nothing else
 * written). */
dbx_status dbx_read_events(dbx_handle *h, int32_t world, dbx_contact_event *out, int32_t cap, int32_t *count);
/* Between dbx_step(..., DBX_STEP_PRESOLVE) and dbx_step_finish: apply what the
 * PreSolve callbacks changed, for n contacts of one world named by their fixture
 * pair (either order).  Friction, restitution and tangent speed stay with the
 * contact for later steps, as in b2Contact; enabled lasts for this step
 * (b2Contact.Update re-enables every contact it updates).  DBX_E_STATE outside
 * that window, DBX_E_INVALID_ARG for a pair without a contact. */
dbx_status dbx_write_contact_settings(dbx_handle *h, int32_t world, const dbx_contact_setting *s, int32_t n);
/* The rest of a DBX_STEP_PRESOLVE step (island search, solve, broadphase,
 * continuous pass) with its parameters; asynchronous like dbx_step.  Split steps
 * run each tier in turn on the handle's stream and are not timed in dbx_profile. */
dbx_status dbx_step_finish(dbx_handle *h);
/* Fat AABBs of each fixture's proxy (lower x, lower y, upper x, upper y). */
dbx_status dbx_read_fat_aabbs(dbx_handle *h, int32_t world, float *out, int32_t n);
/* Pairs the last broadphase pass handed to AddPair, as (fixture, fixture) in the
 * order b2PairLessThan sorts them (b2broadphase.d:312-325). */
dbx_status dbx_read_pairs(dbx_handle *h, int32_t world, int32_t *out, int32_t cap, int32_t *count);
dbx_status dbx_profile_read(dbx_handle *h, dbx_profile *out);
/* Sums of the per-phase times of every step resolved by dbx_sync since the last
 * reset, and the number of steps summed (the b2Profile accumulation the demo's
 * Test.Step does, examples/demo/framework/test.d:558-578). */
dbx_status dbx_profile_totals(dbx_handle *h, dbx_profile *sum, int32_t *steps, int32_t reset);
/* Launch-level timing (no counterpart in the reference): milliseconds of the lead
 * tier's kernels, from HIP events on its stream, summed over the eager steps resolved
 * by dbx_sync since the last reset, in the order Collide, island search, island
 * solve, broadphase (SynchronizeFixtures + FindNewContacts), continuous pass; *steps
 * = steps summed.  The other tiers run beside the lead tier. */
dbx_status dbx_phase_totals(dbx_handle *h, float *ms5, int32_t *steps, int32_t reset);
/* Totals over all worlds: dynamic bodies, contacts, touching contacts, awake bodies. */
dbx_status dbx_totals(dbx_handle *h, int64_t out[4]);
/* Per-world velocity iterations used by the last step (RESIDUAL mode) and max penetration. */
dbx_status dbx_read_solver_stats(dbx_handle *h, int32_t *iterations, float *max_penetration, int32_t nworlds);

/* Per-world velocity residual after the last velocity iteration of the last
 * step in RESIDUAL mode (-1 otherwise).  With a negative tolerance
 * (dbx_set_residual(h, -1, n)) a step runs exactly n iterations, so this is the
 * residual after n iterations. */
dbx_status dbx_read_residuals(dbx_handle *h, float *residual, int32_t nworlds);

/* Per-world number of TOI events the last step's continuous pass solved (the
 * sub-step islands of b2World.SolveTOI, b2world.d:1127-1452); 0 without
 * DBX_STEP_CCD.  The reference counts them nowhere; dbox's b2Profile.solveTOI
 * times the pass (dbx_profile.solveTOI here). */
dbx_status dbx_read_toi_events(dbx_handle *h, int32_t *events, int32_t nworlds);

/* Launch tiers (no counterpart in the reference; worlds are independent there):
 * worlds are launched largest first, in tiers of similar size that each run
 * their step chain on a stream of their own.  Writes (first launch slot, world
 * count, LDS bytes, workgroup size) per tier into tiers[4 * i .. 4 * i + 3] for
 * up to cap tiers, the launch order (world index per slot) into order when it
 * is not null (nworlds entries), and the tier count into count. */
dbx_status dbx_read_tiers(dbx_handle *h, int32_t *tiers, int32_t cap, int32_t *order, int32_t *count);

/* Launch-shape autotuner (no counterpart in the reference).  The next
 * 5 * steps_per_candidate calls of dbx_step run every tier's collide,
 * broadphase and coloured-solve kernels with workgroups of 64, 128, 256, 512
 * and 1024 threads in turn, timing each tier's step chain on its own stream;
 * after the last of them each tier keeps its fastest size (dbx_read_tiers shows
 * it).  Results never depend on the workgroup size, so tuning steps are
 * ordinary steps.  Those steps are not captured as graphs. */
dbx_status dbx_autotune(dbx_handle *h, int32_t steps_per_candidate);
/* Turns the hipGraph replay of dbx_config.use_graph on or off for later steps.
 * Replayed steps record only the step's total time in dbx_profile. */
dbx_status dbx_set_graph(dbx_handle *h, int32_t enabled);
/* Sets the workgroup size of each tier (n = tier count), e.g. from a cache of
 * earlier autotuning. */
dbx_status dbx_set_tier_blocks(dbx_handle *h, const int32_t *blocks, int32_t n);

/* The island solve order of the last step for one world: for each constraint in
 * the order b2World.Solve's island search adds contacts (b2world.d:963-1076), its
 * island, body A, body B and solve batch (4 int32 per row).  Returns the row count. */
dbx_status dbx_read_schedule(dbx_handle *h, int32_t world, int32_t *out, int32_t cap, int32_t *count);

/* Stand-alone broadphase (b2BroadPhase.UpdatePairs semantics) on given fat
 * AABBs and moved flags of n proxies; pairs are written sorted as (i, j), i < j. */
dbx_status dbx_broadphase_pairs(dbx_handle *h, const float *fat_aabbs, const uint8_t *moved, int32_t n,
                                int32_t *out, int32_t cap, int32_t *count);
/* Stand-alone broadphase that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The pairs of
 * dbx_broadphase_pairs, with every buffer in device memory: fat_aabbs (4 floats per
 * proxy, 16-byte aligned) and moved (one int32 per proxy, nonzero = moved) are read;
 * pairs (cap pairs of 2 int32, (i, j) with i < j, in b2PairLessThan order) is
 * written only when the pair count fits in cap; *count (device int32) receives the
 * pair count.  Asynchronous on the handle's stream (dbx_sync waits).  The working
 * memory is a fixed array of the library on each GPU, sized for at most
 * DBX_NOALLOC_MAX_PROXIES proxies (DBX_E_CAPACITY above that); it holds the last
 * call's data, and calls that overlap on one GPU (from other streams, handles or
 * threads) race on it. */
#define DBX_NOALLOC_MAX_PROXIES 16384
dbx_status dbx_broadphase_pairs_noalloc(dbx_handle *h, const float *fat_aabbs, const int32_t *moved, int32_t n,
                                        int32_t *pairs, int32_t cap, int32_t *count);
/* Stand-alone narrowphase: for each of n pairs, shape a[i] at transform xa[i]
 * (x, y, angle) against shape b[i] at xb[i]; the pair is put in b2Contact.Create's
 * type order first.  out holds 32 floats per pair: type, pointCount, swapped, 0,
 * localNormal(2), localPoint(2), points' local (4), ids (2, as uint32 bits),
 * world normal (2), world points (4), separations (2), 8 unused. */
dbx_status dbx_collide_pairs(dbx_handle *h, const dbx_fixture_def *a, const float *xa, const dbx_fixture_def *b,
                             const float *xb, int32_t n, float *out);

/* Residual mode (DBX_STEP_RESIDUAL): velocity iterations stop once the residual
 * (the largest over points of the normal complementarity error |vn - bias| for a
 * positive impulse, else max(0, bias - vn), and of |vt| where friction is not
 * saturated) is at most tolerance, or after max_iterations.  Defaults 1e-3, 100.
 * A negative tolerance never stops early (a residual is never negative): every
 * step runs max_iterations iterations and reports the residual after them.
 * No counterpart in the reference, whose iteration count is fixed. */
dbx_status dbx_set_residual(dbx_handle *h, float tolerance, int32_t max_iterations);

/* ---- Learned warm start (no counterpart in the reference) ----
 * A 14 -> 64 -> 64 -> 2 ReLU network, evaluated and trained on the GPU's bf16
 * matrix cores with fp32 accumulation, predicts the normal and tangent impulse
 * of a manifold point that b2Contact.Update (b2contact.d:306-324) found no match
 * for, in place of the zero such a point starts with; matched points keep the
 * previous step's impulses.  DBX_STEP_LEARNED needs DBX_STEP_WARM and a network.
 * Parameters: DBX_MLP_PARAMS floats, W1[64][14], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2].  Normalisation: DBX_MLP_NORM floats, input mean[14], input
 * scale[14] (the network sees (x - mean) * scale), output mean[2], output
 * scale[2] (impulse = y * scale + mean).  The DBX_MLP_FEATURES features of a
 * point, in order: relative normal velocity, relative tangent velocity, normal
 * x, normal y, separation, invMassA, invMassB, invIA, invIB, the sum of the
 * matched (warm-start) normal impulses on body A's contacts, the same for body
 * B, the number of those matched points, friction, the point's normal mass.  Predicted impulses are clamped to the solver's bounds
 * (normal >= 0, |tangent| <= friction * normal). */
#define DBX_MLP_PARAMS 5250
#define DBX_MLP_NORM 32
#define DBX_MLP_FEATURES 14
/* He-uniform random weights from seed, zero biases, identity normalisation. */
dbx_status dbx_mlp_init(dbx_handle *h, uint64_t seed);
/* Either pointer may be null to keep the current values. */
dbx_status dbx_mlp_set(dbx_handle *h, const float *params, const float *norm);
dbx_status dbx_mlp_get(dbx_handle *h, float *params, float *norm);
/* Predicted impulses y[2 i], y[2 i + 1] for feature rows x[14 i .. 14 i + 13]. */
dbx_status dbx_mlp_forward(dbx_handle *h, const float *x, int32_t n, float *y);
/* Gradient (DBX_MLP_PARAMS floats) and value of the loss
 * 1/n sum_i sum_o (y_o(i) - (t[2 i + o] - out_mean_o) / out_scale_o)^2. */
dbx_status dbx_mlp_gradients(dbx_handle *h, const float *x, const float *t, int32_t n, float *grad, float *loss);
/* Training log of up to capacity rows, filled by steps with DBX_STEP_LOG: the
 * features of each new point and its accumulated impulses after the velocity
 * iterations (the labels). */
dbx_status dbx_mlp_log_enable(dbx_handle *h, int64_t capacity);
dbx_status dbx_mlp_log_count(dbx_handle *h, int64_t *count);
dbx_status dbx_mlp_log_read(dbx_handle *h, int64_t first, int64_t n, float *x, float *t);
dbx_status dbx_mlp_log_clear(dbx_handle *h);
/* Sets the normalisation from the mean and standard deviation of the log. */
dbx_status dbx_mlp_fit_normalization(dbx_handle *h);
/* steps Adam updates (learning rate lr) on mini-batches of batch log rows drawn
 * with replacement; loss (steps floats, may be null) receives each step's loss. */
dbx_status dbx_mlp_train(dbx_handle *h, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss);

#ifdef __cplusplus
}
#endif
#endif
