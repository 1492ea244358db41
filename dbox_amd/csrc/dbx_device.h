// dbx_device.h -- device-side data layout shared by the kernels and the host runtime.
//
// All worlds of a handle live in flat arrays; world w owns the slices
//   bodies   [bodyBase,    bodyBase + bodyCount)
//   fixtures [fixBase,     fixBase + fixCount)        (one proxy per fixture)
//   contacts [contactBase, contactBase + contactCap)  (capacity; live count in WorldState)
//   pairs    [pairBase,    pairBase + pairCap)
// Indices stored inside a world (fixture -> body, contact -> fixture, adjacency)
// are world-local, so a world's results never depend on which other worlds share
// the launch.  Contacts are kept in creation order (oldest first); the reference's
// contact list (b2ContactManager.m_contactList, newest first) is that order reversed.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dbx_geom.h"
#include "dbx_mlp_dev.h"

namespace dbx {

// b2Body.m_flags (b2body.d:1118-1127) in the low 16 bits, body type in bits 16-17.
enum : uint32_t {
    kBodyIsland = 0x1,
    kBodyAwake = 0x2,
    kBodyAutoSleep = 0x4,
    kBodyBullet = 0x8,
    kBodyFixedRotation = 0x10,
    kBodyActive = 0x20,
};
enum : int { kStatic = 0, kKinematic = 1, kDynamic = 2 };
__device__ __host__ __forceinline__ int body_type(uint32_t f) { return (int)((f >> 16) & 3u); }

// b2Contact.m_flags (b2contact.d:242-261) plus transient bits of one step.
enum : uint32_t {
    kContactIsland = 0x1,
    kContactTouching = 0x2,
    kContactEnabled = 0x4,
    kContactDestroyed = 0x100, // fat AABBs stopped overlapping this step
    kContactDormant = 0x200,   // both bodies asleep or static when Collide began
    kContactBegan = 0x400,     // became touching this step (BeginContact)
    kContactEnded = 0x800,     // stopped touching this step (EndContact)
    kContactNewPt0 = 0x1000,   // manifold point 0 has no match in the previous manifold
    kContactNewPt1 = 0x2000,
    kContactSolvedOne = 0x4000, // a two-point manifold solved as one point (block-solver downgrade)
    kContactSensor = 0x8000,    // a fixture is a sensor: overlap only, never solved (b2contact.d:289-298)
};

struct WorldDesc {
    int bodyBase, bodyCount;
    int fixBase, fixCount;
    int contactBase, contactCap;
    int pairBase, pairCap;
    float gx, gy;
    int islandBase, pad;      // pad bit 0: a body of this world has several fixtures;
                              // bit 1: pairs come from the device-wide broadphase pass
    long long islScratchBase; // byte offset of the world's island-search scratch
    long long bpScratchBase;  // byte offset of the world's broadphase scratch
    int jointBase, jointCount; // the world's joints [jointBase, jointBase + jointCount)
};

struct WorldState {
    int contactCount;
    int pairCount;    // unique pairs of the last broadphase pass
    int islandCount;
    int batchCount;
    int error;        // bit0 contact capacity, bit1 pair capacity
    int touching;
    int iterations;   // velocity iterations used by the last solve
    int constraints;  // constraints solved by the last step
    float residual;   // velocity residual after the last iteration (RESIDUAL mode), else -1
    int collideCount; // contacts in the list the last Collide walked (its event flags stay there)
    int widePairs;    // pairs found by the device-wide broadphase pass (worlds too big for one workgroup)
    int toiEvents;    // TOI events (sub-step solves) of the last step's continuous pass
    int adjContacts;  // contacts the last Collide kept, which its adjacency lists cover
    int toiCallbacks; // listener callbacks the last continuous pass made (may exceed its log's capacity)
    int joints;       // joints the last island search added to islands
    int toiCand;      // contacts the continuous pass's first search of a big world gave a TOI (its candidate list)
};

// One listener callback of the continuous pass, laid out as dbx_contact_event:
// BEGIN / END / PRE_SOLVE from b2Contact.Update, POST_SOLVE from the sub-step
// island's b2Island.Report.
struct ToiCallback {
    int type, fa, fb, pointCount;
    float n0, n1, t0, t1;
};
// Log record that is no callback: the discrete solve's PostSolve count and
// impulses of contact slot fa, saved before the pass first updates it.
constexpr int kLogSnapshot = 0x100;

// A contact with its manifold, 80 bytes (b2Contact + b2Manifold, b2contact.d:437-465).
struct alignas(16) Contact {
    int fa, fb;            // world-local fixture indices, registry order
    uint32_t flags;
    int meta;              // manifold type | pointCount << 4
    float friction, restitution;
    uint32_t id0, id1;     // b2ContactID keys
    float4 man;            // localNormal.xy, localPoint.xy
    float4 pt0, pt1;       // point localPoint.xy, normalImpulse, tangentImpulse
};
static_assert(sizeof(Contact) == 80, "contact layout");

// b2ContactVelocityConstraint (b2contactsolver.d:43-58), 160 bytes.
struct alignas(16) VelCon {
    float2 rA0, rB0;
    float2 rA1, rB1;
    float nImp0, tImp0, nImp1, tImp1;
    float nMass0, tMass0, bias0, nMass1;
    float tMass1, bias1, friction, tangentSpeed;
    float2 normal;
    float invMassA, invMassB;
    float invIA, invIB;
    int indexA, indexB;     // world-local bodies
    float4 K;               // ex.x, ex.y, ey.x, ey.y
    float4 normalMass;      // ex.x, ex.y, ey.x, ey.y
    int pointCount, contact, island, dynMask; // dynMask bit0/1: body A/B is dynamic
};
static_assert(sizeof(VelCon) == 160, "velocity constraint layout");
static_assert(offsetof(VelCon, nImp0) == 32, "impulses form one float4");

// b2ContactPositionConstraint (b2contactsolver.d:801-814) plus body indices, 96 bytes.
struct alignas(16) PosCon {
    float2 lp0, lp1;
    float2 localNormal, localPoint;
    float2 lcA, lcB;
    float invMassA, invMassB, invIA, invIB;
    float radiusA, radiusB;
    int type, pointCount;
    int indexA, indexB, dynMask, island;
};
static_assert(sizeof(PosCon) == 96, "position constraint layout");

struct StepParams {
    float dt, inv_dt, dtRatio;
    int velocityIterations, positionIterations;
    uint32_t flags;
    float residualTol;
    int maxIterations;
};
// StepParams.flags bit: the islands of a big world searched by island (exact order,
// no joints) are solved by several workgroups at once (k_solve_grp); not set in the
// residual, learned and logging modes, which treat the world's constraints as one set.
constexpr uint32_t kStepGrouped = 1u << 8;
// StepParams.flags bit: in the exact-order solve of a world on the lead tier's
// two-wave workgroups, the warm start and the velocity iterations may overlap,
// one sweep per wave (velocity_piped), where the world's schedule lets them; not
// set in the residual mode, whose sweeps each end with a test of the whole sweep.
constexpr uint32_t kStepPiped = 1u << 9;
// With kStepPiped: overlap the sweeps even where the schedule says it gains little.
constexpr uint32_t kStepPipedAlways = 1u << 10;

struct JointRec; // dbx_joints.h

// Flat device arrays of all worlds.
struct Dev {
    int nworlds;
    const int *order;       // launch order of worlds (largest first)
    const WorldDesc *W;
    WorldState *S;
    // bodies
    uint32_t *bflags;
    float4 *bxf;            // p.x, p.y, q.s, q.c
    float4 *bpos;           // sweep c.x, c.y, a, sleepTime
    float4 *bpos0;          // sweep c0.x, c0.y, a0, 0
    float4 *bvel;           // v.x, v.y, w, 0
    float4 *bforce;         // force.x, force.y, torque, 0
    const float4 *bmass;    // invMass, invI, localCenter.x, localCenter.y
    const float4 *bparam;   // linearDamping, angularDamping, gravityScale, mass
    // fixtures / proxies
    const int *fbody;
    const Shape *fshape;
    const float4 *fmat;     // friction, restitution, density, shape radius
    const int2 *ffilter;    // category | mask << 16, group
    float4 *ffat;           // fat AABB
    int *fmoved;
    int *fcount;            // per fixture scratch (pair counts)
    // contacts, double buffered (collide compacts one into the other)
    Contact *cont[2];
    float *ctan[2];         // b2Contact.m_tangentSpeed of each contact slot, moving with the contacts
    int useTan;             // some contact has had a tangent speed set (until then every slot holds 0)
    // per body scratch
    int *adjOff, *adjCnt;   // contact adjacency of each body (CSR)
    int *bisl;              // island of a body this step, -1 when not simulated
    // island search of worlds too big for one workgroup (allocated when there are any)
    int *iwParent;          // union-find parent, then the component's root
    int4 *iwAgg;            // per root: seed (largest awake member), constraints, non-static bodies, island
    int4 *iwIsl;            // per island: seed, first constraint, first stack slot, constraints
    int *iwCur;             // per island: bodies placed in its list (its stack slots hold its bodies after the search)
    int *bscratch;          // DFS stack / schedule scratch
    int *wake;              // list position of the first wake in Collide
    unsigned long long *bmask; // colour masks
    // per contact-slot scratch
    int2 *adjList;          // 2 entries per contact slot: (contact * 2 + side, other body)
    int2 *pairs;
    int *ordC;              // island-ordered contact indices
    int *cbatch;            // batch of each ordered constraint
    int *cisl;              // island of each ordered constraint
    int *sched;             // ordered constraints grouped by batch
    int *batchOff;          // per world: contactCap + 65 batch offsets at contactBase + 65 * world
    int *batchCur;          // scatter cursors, same layout as batchOff
    char *islScratch;       // island-search scratch of worlds that do not fit in LDS
    char *bpScratch;        // broadphase scratch of worlds that do not fit in LDS
    VelCon *vc;
    PosCon *pc;
    // per island scratch (island index < body count)
    int *islMin;            // order-preserving int image of a float minimum
    int *islSolved;
    // learned warm start and its training log (null unless enabled)
    MlpDev mlp;
    float *mlpFeat;         // staged feature rows: 2 per contact slot from 2 * contactBase
    int2 *mlpList;          // per staged row: (constraint * 2 + point, log row or -1)
    int *mlpNew;            // per world: rows staged in the current step
    longlong2 *mlpBody;     // per body: sum (fixed point, 2^-32 N s) and count of matched warm-start normal
                            // impulses; integer sums do not depend on the order of the atomic adds
    float *logX;            // log rows of kMlpFeatStride features
    float *logT;            // 2 labels (normal, tangent impulse) per log row
    unsigned long long *logCount;
    long long logCap;
    long long *ktime;       // optional: phase timestamps of the largest world (s_memtime)
    // b2Profile phases of the largest world's island solve (b2island.d:118-237): a ring of
    // kProfSlots slots of kProfStamps wall-clock stamps (wall_clock64) after a counter in
    // prof[0..kProfStamps); slot s starts at prof[kProfStamps * (s + 1)]
    unsigned long long *prof;
    // continuous pass (b2World.SolveTOI) state, valid during that pass only
    int4 *toiC;             // per contact slot: (cached TOI as float bits, toiCount | flags, next edge of
                            // body A's and of body B's list of contacts made since the last Collide)
    int4 *toiB;             // per body: (sweep alpha0 as float bits, island flag, head edge of that list, 0)
    int *toiCand;           // per world from contactBase: the slots that got a TOI in this pass, each once
    const int *bfixOff, *bfixCnt; // per body: its fixtures' range in bfixList
    const int *bfixList;    // per world from fixBase: fixture indices grouped by body
    // joints (b2Joint) of worlds that have them, from each world's jointBase
    JointRec *joints;       // definitions, kept impulses and solver temporaries
    int *ordJ;              // joints in island order (from jointBase)
    int *islCoff, *islJoff; // per world from bodyBase + world: each island's first constraint and
                            // first joint in island order, one past the last island at the end
    const int *jadjOff, *jadjCnt; // per body: its joint edges' range in jadjList (from 2 * jointBase)
    const int2 *jadjList;   // (2 * joint + side, other body) in b2Body.m_jointList order, newest first
    ToiCallback *toiLog;    // per world from toiLogOff[w]: the continuous pass's callbacks in order
    const int *toiLogOff;   // nworlds + 1 offsets into toiLog
};

constexpr int kProfSlots = 4096; // steps between two syncs whose phases are kept (256 KiB)
constexpr int kProfStamps = 8;
// Stamp indices of one slot: solveInit from kProfInit0 to kProfVel0, solveVelocity to
// kProfVel1, solvePosition from kProfPos0 to kProfPos1.
enum : int { kProfInit0 = 0, kProfVel0 = 1, kProfVel1 = 2, kProfPos0 = 3, kProfPos1 = 4 };

// Phase stamps of the island solve of the largest world (the lead tier's first
// workgroup): the first stamp takes a slot, the others write into it.
struct ProfStamp {
    unsigned long long *slot = nullptr;
    __device__ __forceinline__ void begin(const Dev &d, bool lead) {
        if (lead && d.prof) {
            const unsigned long long k = atomicAdd(d.prof, 1ull) % (unsigned long long)kProfSlots;
            slot = d.prof + kProfStamps * (k + 1);
            slot[kProfInit0] = wall_clock64();
        }
    }
    __device__ __forceinline__ void at(int i) const {
        if (slot) slot[i] = wall_clock64();
    }
};

// Order-preserving int image of a float (for atomicMin on floats).
__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff)); }

} // namespace dbx
