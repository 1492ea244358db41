// dbx_runtime.hip -- host runtime behind include/dbox_hip.h.
//
// World construction mirrors what b2World.CreateBody and b2Body.CreateFixture do
// in dbox (hull, normals, centroid, mass data, initial fat AABBs), so that a
// world added here starts from the same state as the reference world built from
// the same definitions.  Those host computations keep the reference's float
// operation order and are compiled without FMA contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/dbox_hip.h"
#include "dbx_device.h"
#include "dbx_joints.h"
#include "dbx_kernels.h"
#include "dbx_mlp.h"

using namespace dbx;

// ---------------------------------------------------------------------------
// Scratch of dbx_broadphase_pairs_noalloc: a fixed device array (each GPU holds
// its own copy until the process ends), big enough for the sort-and-sweep of
// DBX_NOALLOC_MAX_PROXIES proxies (keys 8 B per slot of the power of two above
// the proxy count, fat AABB copy 16 B, three int arrays 12 B per proxy).  Host
// code never passes `scratch` itself to a kernel: its device address on the
// current GPU comes from hipGetSymbolAddress, looked up once per device when a
// handle is created there.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 36u * (size_t)DBX_NOALLOC_MAX_PROXIES;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchDev[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
hipError_t noalloc_scratch(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratchDev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchDev[device] = p;
    }
    *out = g_scratchDev[device];
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// host geometry (creation time only)
// ---------------------------------------------------------------------------
struct HV {
    float x, y;
};
inline HV hv(float x, float y) { return HV{x, y}; }
inline HV operator+(HV a, HV b) { return hv(a.x + b.x, a.y + b.y); }
inline HV operator-(HV a, HV b) { return hv(a.x - b.x, a.y - b.y); }
inline HV operator*(float s, HV a) { return hv(s * a.x, s * a.y); }
inline float hdot(HV a, HV b) { return a.x * b.x + a.y * b.y; }
inline float hcross(HV a, HV b) { return a.x * b.y - a.y * b.x; }
inline HV hmin(HV a, HV b) { return hv(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y); }
inline HV hmax(HV a, HV b) { return hv(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y); }
inline float hnormalize(HV &v) {
    float len = sqrtf(v.x * v.x + v.y * v.y);
    if (len < kEps) return 0.0f;
    float inv = 1.0f / len;
    v.x *= inv;
    v.y *= inv;
    return len;
}
struct HXf {
    HV p;
    float s, c;
};
inline HXf hxf(HV p, float angle) { return HXf{p, sinf(angle), cosf(angle)}; }
inline HV hrot(const HXf &x, HV v) { return hv(x.c * v.x - x.s * v.y, x.s * v.x + x.c * v.y); }
inline HV hmul(const HXf &x, HV v) { return hv((x.c * v.x - x.s * v.y) + x.p.x, (x.s * v.x + x.c * v.y) + x.p.y); }

void set_box(Shape &s, float hx, float hy) { // b2PolygonShape.SetAsBox (b2polygonshape.d:220-232)
    s.type = kPolygon;
    s.radius = kPolygonRadius;
    s.count = 4;
    const float vx[4] = {-hx, hx, hx, -hx}, vy[4] = {-hy, -hy, hy, hy};
    const float nx[4] = {0.0f, 1.0f, 0.0f, -1.0f}, ny[4] = {-1.0f, 0.0f, 1.0f, 0.0f};
    for (int i = 0; i < 4; ++i) {
        s.v[i] = make_float2(vx[i], vy[i]);
        s.n[i] = make_float2(nx[i], ny[i]);
    }
    s.c = make_float2(0.0f, 0.0f);
}

HV centroid_of(const HV *vs, int n) { // b2PolygonShape.ComputeCentroid (b2polygonshape.d:505-549)
    HV c = hv(0.0f, 0.0f);
    float area = 0.0f;
    const float inv3 = 1.0f / 3.0f;
    const HV o = hv(0.0f, 0.0f);
    for (int i = 0; i < n; ++i) {
        HV p2 = vs[i], p3 = i + 1 < n ? vs[i + 1] : vs[0];
        HV e1 = p2 - o, e2 = p3 - o;
        float tri = 0.5f * hcross(e1, e2);
        area += tri;
        c = c + (tri * inv3) * (o + p2 + p3);
    }
    float inv = 1.0f / area;
    return hv(c.x * inv, c.y * inv);
}

// b2PolygonShape.Set: weld, gift-wrap hull, normals, centroid (b2polygonshape.d:74-209).
bool set_polygon(Shape &s, const float *pts, int count) {
    s.type = kPolygon;
    s.radius = kPolygonRadius;
    if (count < 3 || count > kMaxPolygonVertices) return false;
    HV ps[kMaxPolygonVertices];
    int n = 0;
    for (int i = 0; i < count; ++i) {
        HV v = hv(pts[2 * i], pts[2 * i + 1]);
        bool unique = true;
        for (int j = 0; j < n && unique; ++j) {
            HV d = v - ps[j];
            if (hdot(d, d) < 0.5f * kLinearSlop) unique = false;
        }
        if (unique) ps[n++] = v;
    }
    if (n < 3) return false;
    int i0 = 0;
    float x0 = ps[0].x;
    for (int i = 1; i < n; ++i) {
        float x = ps[i].x;
        if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) {
            i0 = i;
            x0 = x;
        }
    }
    int hull[kMaxPolygonVertices];
    int m = 0, ih = i0;
    while (true) {
        hull[m] = ih;
        int ie = 0;
        for (int j = 1; j < n; ++j) {
            if (ie == ih) {
                ie = j;
                continue;
            }
            HV r = ps[ie] - ps[hull[m]];
            HV v = ps[j] - ps[hull[m]];
            float c = hcross(r, v);
            if (c < 0.0f) ie = j;
            if (c == 0.0f && hdot(v, v) > hdot(r, r)) ie = j;
        }
        ++m;
        ih = ie;
        if (ie == i0 || m >= kMaxPolygonVertices) break;
    }
    if (m < 3) return false;
    HV vs[kMaxPolygonVertices];
    for (int i = 0; i < m; ++i) vs[i] = ps[hull[i]];
    s.count = m;
    for (int i = 0; i < m; ++i) {
        HV e = vs[i + 1 < m ? i + 1 : 0] - vs[i];
        HV nrm = hv(1.0f * e.y, -1.0f * e.x);
        hnormalize(nrm);
        s.v[i] = make_float2(vs[i].x, vs[i].y);
        s.n[i] = make_float2(nrm.x, nrm.y);
    }
    HV c = centroid_of(vs, m);
    s.c = make_float2(c.x, c.y);
    return true;
}

struct HMass {
    float mass;
    HV center;
    float I;
};
// b2PolygonShape.ComputeMass / b2CircleShape.ComputeMass / b2EdgeShape.ComputeMass.
HMass shape_mass(const Shape &s, float density) {
    HMass md{0.0f, hv(0.0f, 0.0f), 0.0f};
    if (s.type == kCircle) {
        HV p = hv(s.c.x, s.c.y);
        md.mass = density * kPi * s.radius * s.radius;
        md.center = p;
        md.I = md.mass * (0.5f * s.radius * s.radius + hdot(p, p));
        return md;
    }
    if (s.type == kEdge) {
        md.center = 0.5f * (hv(s.v[0].x, s.v[0].y) + hv(s.v[1].x, s.v[1].y));
        return md;
    }
    int n = s.count;
    HV ref = hv(0.0f, 0.0f);
    for (int i = 0; i < n; ++i) ref = ref + hv(s.v[i].x, s.v[i].y);
    float invn = 1.0f / (float)n;
    ref = hv(ref.x * invn, ref.y * invn);
    HV center = hv(0.0f, 0.0f);
    float area = 0.0f, I = 0.0f;
    const float k_inv3 = 1.0f / 3.0f;
    for (int i = 0; i < n; ++i) {
        int j = i + 1 < n ? i + 1 : 0;
        HV e1 = hv(s.v[i].x, s.v[i].y) - ref;
        HV e2 = hv(s.v[j].x, s.v[j].y) - ref;
        float D = hcross(e1, e2);
        float tri = 0.5f * D;
        area += tri;
        center = center + (tri * k_inv3) * (e1 + e2);
        float intx2 = e1.x * e1.x + e2.x * e1.x + e2.x * e2.x;
        float inty2 = e1.y * e1.y + e2.y * e1.y + e2.y * e2.y;
        I += (0.25f * k_inv3 * D) * (intx2 + inty2);
    }
    md.mass = density * area;
    float inv = 1.0f / area;
    center = hv(center.x * inv, center.y * inv);
    md.center = center + ref;
    md.I = density * I;
    md.I += md.mass * (hdot(md.center, md.center) - hdot(center, center));
    return md;
}

float4 shape_aabb_host(const Shape &s, const HXf &xf) { // ComputeAABB at creation
    if (s.type == kPolygon) {
        HV lo = hmul(xf, hv(s.v[0].x, s.v[0].y)), hi = lo;
        for (int i = 1; i < s.count; ++i) {
            HV w = hmul(xf, hv(s.v[i].x, s.v[i].y));
            lo = hmin(lo, w);
            hi = hmax(hi, w);
        }
        return make_float4(lo.x - s.radius, lo.y - s.radius, hi.x + s.radius, hi.y + s.radius);
    }
    if (s.type == kCircle) {
        HV p = xf.p + hrot(xf, hv(s.c.x, s.c.y));
        return make_float4(p.x - s.radius, p.y - s.radius, p.x + s.radius, p.y + s.radius);
    }
    HV a = hmul(xf, hv(s.v[0].x, s.v[0].y)), b = hmul(xf, hv(s.v[1].x, s.v[1].y));
    HV lo = hmin(a, b), hi = hmax(a, b);
    const float r = (s.flags & 4) ? 0.0f : s.radius; // a chain's child: no radius (b2chainshape.d:214-226)
    return make_float4(lo.x - r, lo.y - r, hi.x + r, hi.y + r);
}

// b2ChainShape (b2chainshape.d:24-263) as its child edges (GetChildEdge,
// b2chainshape.d:134-170), one proxy each.  Definition: count vertices in
// p[0 .. 2 count), flags bit1 / bit2 a previous / next ghost vertex in p[16..17] /
// p[18..19] (SetPrevVertex / SetNextVertex), bit4 a loop (CreateLoop).
bool chain_children(const dbx_fixture_def &d, std::vector<Shape> &out, std::string &err) {
    const bool loop = (d.flags & 16) != 0;
    const int cnt = d.count;
    if (cnt < (loop ? 3 : 2) || cnt > 8) {
        err = "a chain needs 2 to 8 vertices (3 to 8 for a loop)";
        return false;
    }
    HV v[9];
    for (int i = 0; i < cnt; ++i) {
        if (!finite(d.p[2 * i]) || !finite(d.p[2 * i + 1])) {
            err = "chain vertex is not finite";
            return false;
        }
        v[i] = hv(d.p[2 * i], d.p[2 * i + 1]);
        if (i > 0) { // CreateChain / CreateLoop assert b2DistanceSquared > b2_linearSlop^2
            const HV e = v[i] - v[i - 1];
            if (!(hdot(e, e) > kLinearSlop * kLinearSlop)) {
                err = "chain vertices are too close";
                return false;
            }
        }
    }
    int m = cnt;
    HV prev = hv(0.0f, 0.0f), next = hv(0.0f, 0.0f);
    bool hasPrev = false, hasNext = false;
    if (loop) {
        m = cnt + 1;
        v[cnt] = v[0];
        prev = v[m - 2];
        next = v[1];
        hasPrev = hasNext = true;
    } else {
        if (d.flags & 2) {
            prev = hv(d.p[16], d.p[17]);
            hasPrev = true;
        }
        if (d.flags & 4) {
            next = hv(d.p[18], d.p[19]);
            hasNext = true;
        }
    }
    for (int i = 0; i + 1 < m; ++i) {
        Shape s;
        memset(&s, 0, sizeof(s));
        s.type = kEdge;
        s.radius = kPolygonRadius;
        s.v[0] = make_float2(v[i].x, v[i].y);
        s.v[1] = make_float2(v[i + 1].x, v[i + 1].y);
        const bool has0 = i > 0 ? true : hasPrev, has3 = i < m - 2 ? true : hasNext;
        const HV v0 = i > 0 ? v[i - 1] : prev, v3 = i < m - 2 ? v[i + 2] : next;
        s.v[2] = make_float2(v0.x, v0.y);
        s.v[3] = make_float2(v3.x, v3.y);
        s.flags = (has0 ? 1 : 0) | (has3 ? 2 : 0) | 4;
        out.push_back(s);
    }
    return true;
}

bool build_shape(const dbx_fixture_def &d, Shape &s, std::string &err) {
    memset(&s, 0, sizeof(s));
    switch (d.shape) {
    case 0:
        s.type = kCircle;
        s.radius = d.radius;
        s.c = make_float2(d.p[0], d.p[1]);
        if (!(d.radius > 0.0f)) { err = "circle radius must be positive"; return false; }
        return true;
    case 1:
        s.type = kEdge;
        s.radius = kPolygonRadius;
        s.v[0] = make_float2(d.p[0], d.p[1]);
        s.v[1] = make_float2(d.p[2], d.p[3]);
        s.flags = ((d.flags & 2) ? 1 : 0) | ((d.flags & 4) ? 2 : 0);
        s.v[2] = (d.flags & 2) ? make_float2(d.p[4], d.p[5]) : make_float2(0.0f, 0.0f);
        s.v[3] = (d.flags & 4) ? make_float2(d.p[6], d.p[7]) : make_float2(0.0f, 0.0f);
        return true;
    case 2:
        if (!set_polygon(s, d.p, d.count)) { err = "degenerate polygon"; return false; }
        return true;
    case 3:
        set_box(s, d.p[0], d.p[1]);
        if (d.flags & 8) { // SetAsBox(hx, hy, center, angle) (b2polygonshape.d:235-262)
            HXf x = hxf(hv(d.p[2], d.p[3]), d.p[4]);
            for (int i = 0; i < 4; ++i) {
                HV v = hmul(x, hv(s.v[i].x, s.v[i].y));
                HV n = hrot(x, hv(s.n[i].x, s.n[i].y));
                s.v[i] = make_float2(v.x, v.y);
                s.n[i] = make_float2(n.x, n.y);
            }
            s.c = make_float2(d.p[2], d.p[3]);
        }
        return true;
    default:
        err = "unknown shape kind";
        return false;
    }
}

bool finite(float x) { return std::isfinite(x); }

struct WorldHost {
    std::vector<uint32_t> bflags;
    std::vector<float4> bxf, bpos, bpos0, bvel, bforce, bmass, bparam;
    std::vector<int> fbody;
    std::vector<int> pfix, pchild; // per proxy: its fixture (definition index) and child
    std::vector<Shape> fshape;
    std::vector<float4> fmat, ffat;
    std::vector<int2> ffilter;
    int dynamicCount = 0;
    float gx = 0.0f, gy = -10.0f;
    std::vector<JointRec> joints; // b2World.CreateJoint order
};

// b2World.CreateJoint and the constructor of each b2Joint subclass: the members each
// sets from its definition (b2*joint.d), with body B's transform for the mouse
// joint's anchor.
// b2GearJoint's constructor (b2gearjoint.d:83-166): bodies A and B become the two
// joints' bodies B (the joint's edges follow them), C and D their bodies A; the
// constant is the two joint coordinates at the bodies' transforms when it is made.
static bool build_gear(const dbx_joint_def &d, const WorldHost &W, JointRec &j, std::string &err) {
    const int nj = (int)W.joints.size();
    if (d.joint1 < 0 || d.joint1 >= nj || d.joint2 < 0 || d.joint2 >= nj) {
        err = "gear: joint1 and joint2 must be joints of this world made before it";
        return false;
    }
    const JointRec &j1 = W.joints[d.joint1], &j2 = W.joints[d.joint2];
    auto hinge = [](int t) { return t == DBX_JOINT_REVOLUTE || t == DBX_JOINT_PRISMATIC; };
    if (!hinge(j1.type) || !hinge(j2.type)) {
        err = "gear: joint1 and joint2 must be revolute or prismatic joints";
        return false;
    }
    if (!finite(d.ratio)) {
        err = "joint definition is not finite";
        return false;
    }
    j = JointRec{};
    j.type = DBX_JOINT_GEAR;
    j.flags = d.collide_connected ? kJointCollide : 0;
    j.typeA = j1.type;
    j.typeB = j2.type;
    // one joint's coordinate: the relative angle (revolute) or the translation along
    // the axis (prismatic) of body `b` against body `g`
    auto coordinate = [&](const JointRec &jt, int g, int b, float refAngle, float2 anchorG, float2 anchorB, float2 axis) {
        if (jt.type == DBX_JOINT_REVOLUTE) return W.bpos[b].z - W.bpos[g].z - refAngle;
        const float4 xg = W.bxf[g], xb = W.bxf[b];
        const float rx = xb.w * anchorB.x - xb.z * anchorB.y, ry = xb.z * anchorB.x + xb.w * anchorB.y;
        const float vx = rx + (xb.x - xg.x), vy = ry + (xb.y - xg.y);
        const float px = xg.w * vx + xg.z * vy, py = -xg.z * vx + xg.w * vy; // b2MulT(xfG.q, v)
        return (px - anchorG.x) * axis.x + (py - anchorG.y) * axis.y;
    };
    j.bodyC = j1.bodyA;
    j.bodyA = j1.bodyB;
    j.localAnchorC = j1.localAnchorA;
    j.localAnchorA = j1.localAnchorB;
    j.referenceAngleA = j1.referenceAngle;
    j.localAxisC = j1.type == DBX_JOINT_REVOLUTE ? make_float2(0.0f, 0.0f) : j1.localXAxisA;
    const float coordinateA = coordinate(j1, j.bodyC, j.bodyA, j.referenceAngleA, j.localAnchorC, j.localAnchorA,
                                         j.localAxisC);
    j.bodyD = j2.bodyA;
    j.bodyB = j2.bodyB;
    j.localAnchorD = j2.localAnchorA;
    j.localAnchorB = j2.localAnchorB;
    j.referenceAngleB = j2.referenceAngle;
    j.localAxisD = j2.type == DBX_JOINT_REVOLUTE ? make_float2(0.0f, 0.0f) : j2.localXAxisA;
    const float coordinateB = coordinate(j2, j.bodyD, j.bodyB, j.referenceAngleB, j.localAnchorD, j.localAnchorB,
                                         j.localAxisD);
    j.ratio = d.ratio;
    j.constant = coordinateA + j.ratio * coordinateB;
    j.limitState = kLimitInactive;
    return true;
}

bool build_joint(const dbx_joint_def &d, const WorldHost &W, JointRec &j, std::string &err) {
    const int nb = (int)W.bflags.size();
    if (d.type == DBX_JOINT_GEAR) {
        if (d.body_a < 0 || d.body_a >= nb || d.body_b < 0 || d.body_b >= nb || d.body_a == d.body_b) {
            err = "bad joint bodies";
            return false;
        }
        return build_gear(d, W, j, err);
    }
    if (d.type < DBX_JOINT_REVOLUTE || d.type > DBX_JOINT_MOTOR) {
        err = "bad joint type";
        return false;
    }
    if (d.body_a < 0 || d.body_a >= nb || d.body_b < 0 || d.body_b >= nb || d.body_a == d.body_b) {
        err = "bad joint bodies";
        return false;
    }
    const float *fl = &d.local_anchor_a[0];
    for (int i = 0; i < 29; ++i)
        if (!finite(fl[i])) {
            err = "joint definition is not finite";
            return false;
        }
    if (d.type == DBX_JOINT_PULLEY && d.ratio == 0.0f) {
        err = "pulley ratio is zero";
        return false;
    }
    j = JointRec{};
    j.type = d.type;
    j.bodyA = d.body_a;
    j.bodyB = d.body_b;
    j.flags = (d.collide_connected ? kJointCollide : 0) | (d.enable_limit ? kJointLimit : 0) |
              (d.enable_motor ? kJointMotorOn : 0);
    j.localAnchorA = make_float2(d.local_anchor_a[0], d.local_anchor_a[1]);
    j.localAnchorB = make_float2(d.local_anchor_b[0], d.local_anchor_b[1]);
    j.referenceAngle = d.reference_angle;
    j.length = d.length;
    j.frequencyHz = d.frequency_hz;
    j.dampingRatio = d.damping_ratio;
    j.lower = d.lower;
    j.upper = d.upper;
    j.motorSpeed = d.motor_speed;
    j.maxMotor = d.max_motor;
    j.maxForce = d.max_force;
    j.maxTorque = d.max_torque;
    if (d.type == DBX_JOINT_PRISMATIC || d.type == DBX_JOINT_WHEEL) {
        HV x = hv(d.local_axis_a[0], d.local_axis_a[1]);
        if (d.type == DBX_JOINT_PRISMATIC) hnormalize(x); // b2prismaticjoint.d:174; the wheel keeps it as given
        j.localXAxisA = make_float2(x.x, x.y);
        j.localYAxisA = make_float2(-1.0f * x.y, 1.0f * x.x); // b2Cross(1.0f, axis)
    }
    if (d.type == DBX_JOINT_MOTOR) {
        j.linearOffset = make_float2(d.linear_offset[0], d.linear_offset[1]);
        j.angularOffset = d.angular_offset;
        j.correctionFactor = d.correction_factor;
    }
    if (d.type == DBX_JOINT_MOUSE) { // m_localAnchorB = b2MulT(bodyB.GetTransform(), target) (b2mousejoint.d:124)
        j.targetA = make_float2(d.target[0], d.target[1]);
        const float4 x = W.bxf[d.body_b];
        const float px = d.target[0] - x.x, py = d.target[1] - x.y;
        j.localAnchorB = make_float2(x.w * px + x.z * py, -x.z * px + x.w * py);
    }
    if (d.type == DBX_JOINT_PULLEY) {
        j.groundAnchorA = make_float2(d.ground_anchor_a[0], d.ground_anchor_a[1]);
        j.groundAnchorB = make_float2(d.ground_anchor_b[0], d.ground_anchor_b[1]);
        j.lengthA = d.length_a;
        j.lengthB = d.length_b;
        j.ratio = d.ratio;
        j.constant = d.length_a + j.ratio * d.length_b; // b2pulleyjoint.d:151
    }
    j.limitState = kLimitInactive;
    return true;
}

// b2World.CreateBody + b2Body.CreateFixture for a whole definition list.
bool build_world(const dbx_body_def *bodies, int nb, const dbx_fixture_def *fx, int nf, WorldHost &W, std::string &err) {
    W.bflags.resize(nb);
    W.bxf.resize(nb);
    W.bpos.resize(nb);
    W.bpos0.resize(nb);
    W.bvel.resize(nb);
    W.bforce.assign(nb, make_float4(0, 0, 0, 0));
    W.bmass.resize(nb);
    W.bparam.resize(nb);
    std::vector<HXf> xf(nb);
    std::vector<HV> lc(nb, hv(0.0f, 0.0f)), c(nb);
    std::vector<float> mass(nb), invMass(nb), I(nb), invI(nb);
    std::vector<std::vector<int>> fixturesOf(nb);
    for (int i = 0; i < nb; ++i) {
        const dbx_body_def &d = bodies[i];
        if (d.type < 0 || d.type > 2) { err = "bad body type"; return false; }
        if (!finite(d.px) || !finite(d.py) || !finite(d.angle) || !finite(d.vx) || !finite(d.vy) || !finite(d.w) ||
            !finite(d.linear_damping) || d.linear_damping < 0.0f || !finite(d.angular_damping) || d.angular_damping < 0.0f ||
            !finite(d.gravity_scale)) {
            err = "body definition is not finite";
            return false;
        }
        uint32_t f = kBodyActive;
        if (d.bullet) f |= kBodyBullet;
        if (d.fixed_rotation) f |= kBodyFixedRotation;
        if (d.allow_sleep) f |= kBodyAutoSleep;
        if (d.awake) f |= kBodyAwake;
        f |= (uint32_t)d.type << 16;
        W.bflags[i] = f;
        xf[i] = hxf(hv(d.px, d.py), d.angle);
        c[i] = xf[i].p;
        mass[i] = invMass[i] = d.type == kDynamic ? 1.0f : 0.0f;
        I[i] = invI[i] = 0.0f;
        W.bvel[i] = make_float4(d.vx, d.vy, d.w, 0.0f);
        W.bparam[i] = make_float4(d.linear_damping, d.angular_damping, d.gravity_scale, 0.0f);
        if (d.type == kDynamic) W.dynamicCount++;
    }
    // One proxy per fixture child (b2Fixture.CreateProxies, b2fixture.d:414-431), in
    // creation order: a chain contributes one edge proxy per segment, every other
    // shape one; the device's "fixture" arrays are these proxies.
    std::vector<int> firstProxy(nf);
    for (int i = 0; i < nf; ++i) {
        const dbx_fixture_def &d = fx[i];
        if (d.body < 0 || d.body >= nb) { err = "fixture body out of range"; return false; }
        if (!finite(d.density) || d.density < 0.0f || !finite(d.friction) || !finite(d.restitution)) {
            err = "fixture definition is not finite";
            return false;
        }
        std::vector<Shape> kids;
        if (d.shape == 4) {
            if (!chain_children(d, kids, err)) return false;
        } else {
            Shape s;
            if (!build_shape(d, s, err)) return false;
            kids.push_back(s);
        }
        if (d.flags & 1) // b2FixtureDef.isSensor
            for (Shape &s : kids) s.flags |= kShapeSensor;
        firstProxy[i] = (int)W.fshape.size();
        for (const Shape &s : kids) {
            W.fbody.push_back(d.body);
            W.fshape.push_back(s);
            W.fmat.push_back(make_float4(d.friction, d.restitution, d.density, s.radius));
            W.ffilter.push_back(make_int2((int)d.category | ((int)d.mask << 16), (int)d.group));
            // b2Fixture.CreateProxies -> b2DynamicTree.CreateProxy: fatten by the extension.
            float4 a = shape_aabb_host(s, xf[d.body]);
            W.ffat.push_back(
                make_float4(a.x - kAabbExtension, a.y - kAabbExtension, a.z + kAabbExtension, a.w + kAabbExtension));
            W.pfix.push_back(i);
            W.pchild.push_back((int)(&s - &kids[0]));
        }
        fixturesOf[d.body].insert(fixturesOf[d.body].begin(), i); // m_fixtureList: newest first
        // b2Body.ResetMassData (b2body.d:555-625) after each fixture with density
        if (d.density > 0.0f) {
            int b = d.body;
            int type = bodies[b].type;
            mass[b] = invMass[b] = I[b] = invI[b] = 0.0f;
            lc[b] = hv(0.0f, 0.0f);
            if (type != kDynamic) {
                c[b] = xf[b].p;
                continue;
            }
            HV center = hv(0.0f, 0.0f);
            for (int fi : fixturesOf[b]) {
                if (fx[fi].density == 0.0f) continue;
                // a chain has no mass (b2chainshape.d:228-234), as its first child edge
                HMass md = shape_mass(W.fshape[firstProxy[fi]], fx[fi].density);
                mass[b] += md.mass;
                center = center + md.mass * md.center;
                I[b] += md.I;
            }
            if (mass[b] > 0.0f) {
                invMass[b] = 1.0f / mass[b];
                center = hv(center.x * invMass[b], center.y * invMass[b]);
            } else {
                mass[b] = 1.0f;
                invMass[b] = 1.0f;
            }
            if (I[b] > 0.0f && (W.bflags[b] & kBodyFixedRotation) == 0) {
                I[b] -= mass[b] * hdot(center, center);
                invI[b] = 1.0f / I[b];
            } else {
                I[b] = 0.0f;
                invI[b] = 0.0f;
            }
            HV old = c[b];
            lc[b] = center;
            c[b] = hmul(xf[b], lc[b]);
            float w = W.bvel[b].z;
            HV dv = hv(-w * (c[b].y - old.y), w * (c[b].x - old.x));
            W.bvel[b].x += dv.x;
            W.bvel[b].y += dv.y;
        }
    }
    for (int i = 0; i < nb; ++i) {
        W.bxf[i] = make_float4(xf[i].p.x, xf[i].p.y, xf[i].s, xf[i].c);
        W.bpos[i] = make_float4(c[i].x, c[i].y, bodies[i].angle, 0.0f);
        W.bpos0[i] = W.bpos[i];
        W.bpos0[i].w = 0.0f;
        W.bmass[i] = make_float4(invMass[i], invI[i], lc[i].x, lc[i].y);
        W.bparam[i].w = mass[i];
    }
    return true;
}

} // namespace

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
// Worlds are launched in tiers of similar size, largest first.  A tier's island
// search, solver and broadphase workgroups get `lds` bytes of LDS (a world whose
// state does not fit there runs from global memory instead), and each tier
// runs its whole step chain on its own stream, so that the many small worlds
// proceed while the few largest ones, which set the step time, are solved.
struct Tier {
    int first = 0, count = 0; // range in the launch order
    int lds = 0;
    int block = 256;          // threads of its collide, broadphase and coloured-solve workgroups
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr; // chain timing while autotuning
    bool own = false;         // stream and event created for this tier
    int wideFirst = 0, wideCount = 0; // its worlds whose broadphase runs device-wide (range in the handle's list)
};

// A captured step: the launch sequence of one step for one combination of
// device state, step parameters, buffer parity and launch shapes.
struct GraphEntry {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
};
constexpr int kMaxGraphs = 8;
constexpr int kTuneBlocks[] = {64, 128, 256, 512, 1024};
constexpr int kTuneCandidates = 5;
constexpr int kStepEvents = 7;

struct dbx_handle {
    dbx_config cfg{};
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<WorldHost> pending;
    bool finalized = false;
    int nworlds = 0;
    std::vector<WorldDesc> W;
    std::vector<int> dynamicCount;
    Dev d{};
    int cur = 0;
    float inv_dt0 = 0.0f;
    bool newFixture = true;
    int block = 256;
    int ldsMax = 65536;           // dynamic LDS a workgroup may use on this device
    std::vector<Tier> tiers;      // tier 0 = the largest worlds (on `stream` unless leadCus)
    std::vector<WideBp> wide;     // worlds whose broadphase scratch does not fit in LDS, by tier
    std::vector<int> proxyFix, proxyChild; // per proxy (from each world's fixBase): fixture and child
    // DBX_STEP_PRESOLVE: a step stopped after Collide, finished by dbx_step_finish
    bool splitWaiting = false;    // the first half has run, the second has not
    bool lastSplit = false;       // the last (or pending) step was split: PreSolve events are kept
    bool lastCcd = false;         // the last step ran the continuous pass (its callbacks are in toiLog)
    std::vector<int> toiLogOff;   // host copy of Dev::toiLogOff
    StepParams pendSp{};
    uint32_t pendFlags = 0;
    Contact *contPre = nullptr;   // the list Collide walked, as it was before (old manifolds)
    std::vector<void *> allocs;
    hipEvent_t ev[6] = {};
    std::vector<hipEvent_t> ring;   // kStepEvents events per step since the last sync
    std::vector<char> ringGraph;    // per ring entry: the step ran as a graph (phase events unrecorded)
    std::vector<GraphEntry> graphs;
    hipEvent_t fork = nullptr;      // fork event inside captures
    int tuneSteps = 0, tuneDone = 0, tunePer = 0; // autotuner progress
    std::vector<double> tuneMs;     // [tier][candidate] summed chain milliseconds
    std::vector<double> tierMs;     // DBX_TIER_TIMES: summed chain milliseconds per tier
    int tierSteps = 0;
    int ringUsed = 0;
    dbx_profile prof{};             // last step
    void *noallocScratch = nullptr; // device address of `scratch` on this handle's GPU
    std::vector<char> ringSolved;   // per ring entry: the step ran the island solve (dt > 0)
    unsigned long long profSeen = 0; // phase-stamp slots of the solve kernels already read
    double wallKHz = 100000.0;      // rate of the kernels' wall_clock64 stamps
    int cuCount = 0;                // compute units of the device
    bool leadCus = false;           // the lead tier runs on CUs of its own (CU-masked streams)
    dbx_profile sum{};              // accumulated since the last dbx_profile_totals reset
    double phaseSum[5] = {0, 0, 0, 0, 0}; // lead tier: collide, islands, solve, broadphase, TOI (eager steps)
    int phaseSteps = 0;
    int sumSteps = 0;
    long long totalBodies = 0, totalFix = 0, totalContacts = 0, totalPairs = 0;
    float *dPen = nullptr;
    MlpState mlp{};
    float residualTol = 1e-3f;
    int maxIterations = 100;
    bool staged = false;            // learned / log staging buffers allocated
    float *logX = nullptr, *logT = nullptr;
    unsigned long long *logCount = nullptr;
    long long logCap = 0;
};

namespace {

dbx_status fail(dbx_handle *h, dbx_status s, const std::string &msg) {
    if (h) h->err = msg;
    return s;
}
dbx_status hip_fail(dbx_handle *h, hipError_t e, const char *what) {
    return fail(h, DBX_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define DBX_HIP(h, expr)                                                                                              \
    do {                                                                                                              \
        hipError_t _e = (expr);                                                                                       \
        if (_e != hipSuccess) return hip_fail(h, _e, #expr);                                                          \
    } while (0)

template <class T>
hipError_t dalloc(dbx_handle *h, T **p, size_t n) {
    size_t bytes = std::max<size_t>(1, n) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return e;
    h->allocs.push_back(*p);
    return hipMemsetAsync(*p, 0, bytes, h->stream);
}
template <class T>
hipError_t upload(dbx_handle *h, T *dst, const std::vector<T> &src) {
    if (src.empty()) return hipSuccess;
    return hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, h->stream);
}

} // namespace

template <class T>
static void key_append(std::vector<unsigned char> &k, const T &v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

extern "C" {

dbx_status dbx_create(const dbx_config *config, dbx_handle **out) {
    if (!out) return DBX_E_INVALID_ARG;
    *out = nullptr;
    dbx_handle *h = new dbx_handle();
    if (config) h->cfg = *config;
    if (h->cfg.contacts_per_fixture <= 0) h->cfg.contacts_per_fixture = 8;
    if (h->cfg.block_size > 0) h->block = h->cfg.block_size;
    if (h->block < 64 || h->block > kMaxBlock || (h->block & 63)) {
        delete h;
        return DBX_E_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || h->cfg.device < 0 || h->cfg.device >= ndev) {
        delete h;
        return DBX_E_HIP;
    }
    if (hipSetDevice(h->cfg.device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
        delete h;
        return DBX_E_HIP;
    }
    for (auto &e : h->ev) (void)hipEventCreate(&e);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->cfg.device) == hipSuccess) {
        size_t lim = prop.sharedMemPerBlock;
        if (lim > 160 * 1024) lim = 160 * 1024;
        if (lim >= 32768) h->ldsMax = (int)lim;
        h->cuCount = prop.multiProcessorCount;
    }
    if (configure_kernels(h->ldsMax) != hipSuccess) {
        h->ldsMax = 65536;
        (void)configure_kernels(h->ldsMax);
    }
    int wallKHz = 0;
    if (hipDeviceGetAttribute(&wallKHz, hipDeviceAttributeWallClockRate, h->cfg.device) == hipSuccess && wallKHz > 0)
        h->wallKHz = wallKHz;
    if (noalloc_scratch(h->cfg.device, &h->noallocScratch) != hipSuccess) h->noallocScratch = nullptr;
    (void)hipGetLastError(); // clear any error state left by the attribute calls
    if (getenv("DBX_VERBOSE"))
        fprintf(stderr, "dbx: device %d, %d CUs, LDS per workgroup %d bytes\n", h->cfg.device,
                prop.multiProcessorCount, h->ldsMax);
    *out = h;
    return DBX_OK;
}

void dbx_destroy(dbx_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->cfg.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    mlp_release(h->mlp);
    for (void *p : h->allocs) (void)hipFree(p);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->ring) (void)hipEventDestroy(e);
    for (GraphEntry &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    if (h->fork) (void)hipEventDestroy(h->fork);
    for (Tier &t : h->tiers) {
        if (t.t0) (void)hipEventDestroy(t.t0);
        if (t.t1) (void)hipEventDestroy(t.t1);
        if (!t.own) continue;
        (void)hipStreamSynchronize(t.stream);
        (void)hipStreamDestroy(t.stream);
        (void)hipEventDestroy(t.done);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *dbx_last_error(const dbx_handle *h) { return h ? h->err.c_str() : "null handle"; }

dbx_status dbx_add_world(dbx_handle *h, const dbx_body_def *bodies, int32_t nb, const dbx_fixture_def *fixtures, int32_t nf,
                         float gx, float gy, int32_t *world_id) {
    if (!h) return DBX_E_INVALID_ARG;
    if (h->finalized) return fail(h, DBX_E_LOCKED, "worlds cannot be added after dbx_finalize");
    if (nb <= 0 || nf < 0 || !bodies || (nf > 0 && !fixtures)) return fail(h, DBX_E_INVALID_ARG, "empty world");
    if (!std::isfinite(gx) || !std::isfinite(gy)) return fail(h, DBX_E_INVALID_ARG, "gravity is not finite");
    WorldHost W;
    std::string err;
    if (!build_world(bodies, nb, fixtures, nf, W, err)) {
        bool unsupported = err.find("not supported") != std::string::npos;
        return fail(h, unsupported ? DBX_E_UNSUPPORTED : DBX_E_INVALID_ARG, err);
    }
    W.gx = gx;
    W.gy = gy;
    h->pending.push_back(std::move(W));
    if (world_id) *world_id = (int32_t)h->pending.size() - 1;
    return DBX_OK;
}

dbx_status dbx_add_joints(dbx_handle *h, int32_t world, const dbx_joint_def *joints, int32_t n) {
    if (!h) return DBX_E_INVALID_ARG;
    if (h->finalized) return fail(h, DBX_E_LOCKED, "joints are added before dbx_finalize");
    if (world < 0 || world >= (int)h->pending.size()) return fail(h, DBX_E_INVALID_ARG, "bad world index");
    if (n < 0 || (n > 0 && !joints)) return fail(h, DBX_E_INVALID_ARG, "bad joint array");
    WorldHost &W = h->pending[world];
    const size_t had = W.joints.size();
    std::string err;
    for (int i = 0; i < n; ++i) { // one at a time: a gear may couple joints made earlier in the same call
        JointRec j;
        if (!build_joint(joints[i], W, j, err)) {
            W.joints.resize(had);
            return fail(h, DBX_E_INVALID_ARG, err);
        }
        W.joints.push_back(j);
    }
    return DBX_OK;
}

// The lead tier (the few largest worlds) sets the step time, and its kernels need
// whole CUs (up to all 160 KB of a CU's LDS).  Sharing the device, its workgroups
// wait at every dispatch until the thousands of small worlds' workgroups leave a
// CU free.  With more than one tier, the lead tier therefore gets a stream whose
// CU mask holds eight CUs of its own, one on each of the eight XCDs, and the other
// tiers' streams the remaining CUs.  Mask bits 33 k (k = 0..7) name one CU per XCD
// whether the runtime deals mask bits out to the XCDs in turn or in blocks of 32,
// so every XCD keeps CUs for both masks.  Needs 256 CUs and at most eight lead
// worlds with no device-wide pass; DBX_SHARED_CUS turns it off.
constexpr int kLeadCus = 8;
static hipError_t create_tier_streams(dbx_handle *h, std::vector<Tier> &tiers) {
    const bool reserve = tiers.size() > 1 && h->cuCount == 256 && tiers[0].count <= kLeadCus &&
                         tiers[0].wideCount == 0 && !getenv("DBX_SHARED_CUS");
    uint32_t lead[8] = {}, rest[8];
    for (uint32_t &m : rest) m = ~0u;
    for (int k = 0; k < kLeadCus; ++k) {
        const int bit = 33 * k;
        lead[bit >> 5] |= 1u << (bit & 31);
        rest[bit >> 5] &= ~(1u << (bit & 31));
    }
    h->leadCus = reserve;
    for (size_t t = 0; t < tiers.size(); ++t) {
        Tier &T = tiers[t];
        hipError_t e = hipSuccess;
        if (reserve) e = hipExtStreamCreateWithCUMask(&T.stream, 8, t == 0 ? lead : rest);
        else if (t == 0) T.stream = h->stream;
        else e = hipStreamCreate(&T.stream);
        if (e != hipSuccess) return e;
        if (T.stream == h->stream) continue;
        T.own = true;
        e = hipEventCreateWithFlags(&T.done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    if (reserve && getenv("DBX_VERBOSE")) {
        uint32_t m[8] = {};
        if (hipExtStreamGetCUMask(tiers[0].stream, 8, m) == hipSuccess)
            fprintf(stderr, "dbx: lead tier CU mask %08x %08x %08x %08x %08x %08x %08x %08x\n", m[0], m[1], m[2], m[3],
                    m[4], m[5], m[6], m[7]);
    }
    return hipSuccess;
}

dbx_status dbx_finalize(dbx_handle *h) {
    if (!h) return DBX_E_INVALID_ARG;
    if (h->finalized) return fail(h, DBX_E_STATE, "already finalized");
    if (h->pending.empty()) return fail(h, DBX_E_STATE, "no worlds");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    const int nw = (int)h->pending.size();
    h->nworlds = nw;
    h->W.resize(nw);
    long long nb = 0, nf = 0, nc = 0, np = 0, islBytes = 0, bpBytes = 0;
    int cpf = h->cfg.contacts_per_fixture;
    for (int w = 0; w < nw; ++w) {
        const WorldHost &H = h->pending[w];
        WorldDesc &D = h->W[w];
        D.bodyBase = (int)nb;
        D.bodyCount = (int)H.bflags.size();
        D.fixBase = (int)nf;
        D.fixCount = (int)H.fbody.size();
        D.contactBase = (int)nc;
        D.contactCap = std::max(64, cpf * D.fixCount);
        D.pairBase = (int)np;
        D.pairCap = std::max(64, cpf * D.fixCount);
        D.gx = H.gx;
        D.gy = H.gy;
        D.islandBase = D.bodyBase;
        // bit 0: some non-static body has more than one proxy, so two contacts of
        // one body can lead to the same other non-static body (the island search
        // must de-duplicate pushes; static bodies are never pushed)
        D.pad = 0;
        {
            std::vector<int> nfix(D.bodyCount, 0);
            for (int b : H.fbody)
                if (body_type(H.bflags[b]) != kStatic && ++nfix[b] > 1) D.pad |= 1;
        }
        D.jointBase = w == 0 ? 0 : h->W[w - 1].jointBase + h->W[w - 1].jointCount;
        D.jointCount = (int)H.joints.size();
        D.islScratchBase = islBytes;
        islBytes += (long long)islands_scratch_bytes(D.bodyCount, D.contactCap);
        D.bpScratchBase = bpBytes;
        bpBytes += (long long)broadphase_scratch_bytes(D.fixCount);
        nb += D.bodyCount;
        nf += D.fixCount;
        nc += D.contactCap;
        np += D.pairCap;
        h->dynamicCount.push_back(H.dynamicCount);
    }
    if (nc + nw > INT32_MAX / 2 || nb > INT32_MAX / 2) return fail(h, DBX_E_CAPACITY, "too many contacts for int32 indexing");
    h->totalBodies = nb;
    h->totalFix = nf;
    h->totalContacts = nc;
    h->totalPairs = np;
    // Launch order: largest LDS need first (island-search state at three contacts
    // per fixture, which also bounds the solver's), then split into tiers.
    std::vector<size_t> need(nw);
    for (int w = 0; w < nw; ++w)
        need[w] = islands_scratch_bytes(h->W[w].bodyCount, 3 * h->W[w].fixCount + 16);
    std::vector<int> order(nw);
    for (int w = 0; w < nw; ++w) order[w] = w;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return need[a] > need[b]; });
    {
        std::vector<int> sizes = {h->ldsMax};
        for (int l : {81920, 40960, 20480})
            if (l < h->ldsMax) sizes.push_back(l);
        std::vector<Tier> tiers;
        for (int i = 0; i < nw; ++i) {
            size_t n = need[order[i]];
            int t = 0;
            while (t + 1 < (int)sizes.size() && (size_t)sizes[t + 1] >= n) ++t;
            if (tiers.empty() || tiers.back().lds != sizes[t]) {
                Tier T;
                T.first = i;
                T.lds = sizes[t];
                T.block = h->block;
                tiers.push_back(T);
            }
            tiers.back().count++;
        }
        // A world whose broadphase scratch does not fit in its tier's LDS has its
        // pairs found by the device-wide pass.
        for (Tier &T : tiers) {
            const int lds1 = std::min(T.lds, h->ldsMax - 1024);
            T.wideFirst = (int)h->wide.size();
            for (int i = T.first; i < T.first + T.count; ++i) {
                WorldDesc &D = h->W[order[i]];
                if (broadphase_scratch_bytes(D.fixCount) <= (size_t)lds1) continue;
                D.pad |= 2;
                WideBp b{};
                b.world = order[i];
                b.tierFirst = i == T.first ? 1 : 0;
                b.n = D.fixCount;
                b.cap = D.pairCap;
                h->wide.push_back(b);
            }
            T.wideCount = (int)h->wide.size() - T.wideFirst;
        }
        DBX_HIP(h, create_tier_streams(h, tiers));
        h->tiers = tiers;
        if (getenv("DBX_VERBOSE"))
            for (const Tier &T : h->tiers)
                fprintf(stderr, "dbx: tier of %d worlds from %d, LDS %d bytes, block %d\n", T.count, T.first, T.lds,
                        T.block);
    }

    Dev &d = h->d;
    d.nworlds = nw;
    int *dOrder;
    WorldDesc *dW;
    DBX_HIP(h, dalloc(h, &dOrder, nw));
    DBX_HIP(h, dalloc(h, &dW, nw));
    DBX_HIP(h, dalloc(h, &d.S, nw));
    DBX_HIP(h, hipMemcpyAsync(dOrder, order.data(), nw * sizeof(int), hipMemcpyHostToDevice, h->stream));
    DBX_HIP(h, hipMemcpyAsync(dW, h->W.data(), nw * sizeof(WorldDesc), hipMemcpyHostToDevice, h->stream));
    d.order = dOrder;
    d.W = dW;
    // bodies
    std::vector<uint32_t> bflags;
    std::vector<float4> bxf, bpos, bpos0, bvel, bforce, bmass, bparam, fmat, ffat;
    std::vector<int> fbody;
    std::vector<Shape> fshape;
    std::vector<int2> ffilter;
    bflags.reserve(nb);
    for (const WorldHost &H : h->pending) {
        bflags.insert(bflags.end(), H.bflags.begin(), H.bflags.end());
        bxf.insert(bxf.end(), H.bxf.begin(), H.bxf.end());
        bpos.insert(bpos.end(), H.bpos.begin(), H.bpos.end());
        bpos0.insert(bpos0.end(), H.bpos0.begin(), H.bpos0.end());
        bvel.insert(bvel.end(), H.bvel.begin(), H.bvel.end());
        bforce.insert(bforce.end(), H.bforce.begin(), H.bforce.end());
        bmass.insert(bmass.end(), H.bmass.begin(), H.bmass.end());
        bparam.insert(bparam.end(), H.bparam.begin(), H.bparam.end());
        fbody.insert(fbody.end(), H.fbody.begin(), H.fbody.end());
        h->proxyFix.insert(h->proxyFix.end(), H.pfix.begin(), H.pfix.end());
        h->proxyChild.insert(h->proxyChild.end(), H.pchild.begin(), H.pchild.end());
        fshape.insert(fshape.end(), H.fshape.begin(), H.fshape.end());
        fmat.insert(fmat.end(), H.fmat.begin(), H.fmat.end());
        ffat.insert(ffat.end(), H.ffat.begin(), H.ffat.end());
        ffilter.insert(ffilter.end(), H.ffilter.begin(), H.ffilter.end());
    }
    float4 *p_bmass, *p_bparam, *p_fmat;
    int *p_fbody;
    Shape *p_fshape;
    int2 *p_ffilter;
    DBX_HIP(h, dalloc(h, &d.bflags, nb));
    DBX_HIP(h, dalloc(h, &d.bxf, nb));
    DBX_HIP(h, dalloc(h, &d.bpos, nb));
    DBX_HIP(h, dalloc(h, &d.bpos0, nb));
    DBX_HIP(h, dalloc(h, &d.bvel, nb));
    DBX_HIP(h, dalloc(h, &d.bforce, nb));
    DBX_HIP(h, dalloc(h, &p_bmass, nb));
    DBX_HIP(h, dalloc(h, &p_bparam, nb));
    DBX_HIP(h, dalloc(h, &p_fbody, nf));
    DBX_HIP(h, dalloc(h, &p_fshape, nf));
    DBX_HIP(h, dalloc(h, &p_fmat, nf));
    DBX_HIP(h, dalloc(h, &p_ffilter, nf));
    DBX_HIP(h, dalloc(h, &d.ffat, nf));
    DBX_HIP(h, dalloc(h, &d.fmoved, nf));
    DBX_HIP(h, dalloc(h, &d.fcount, nf));
    DBX_HIP(h, upload(h, d.bflags, bflags));
    DBX_HIP(h, upload(h, d.bxf, bxf));
    DBX_HIP(h, upload(h, d.bpos, bpos));
    DBX_HIP(h, upload(h, d.bpos0, bpos0));
    DBX_HIP(h, upload(h, d.bvel, bvel));
    DBX_HIP(h, upload(h, d.bforce, bforce));
    DBX_HIP(h, upload(h, p_bmass, bmass));
    DBX_HIP(h, upload(h, p_bparam, bparam));
    DBX_HIP(h, upload(h, p_fbody, fbody));
    DBX_HIP(h, upload(h, p_fshape, fshape));
    DBX_HIP(h, upload(h, p_fmat, fmat));
    DBX_HIP(h, upload(h, p_ffilter, ffilter));
    DBX_HIP(h, upload(h, d.ffat, ffat));
    std::vector<int> ones(nf, 1); // every new proxy is in the move buffer (b2BroadPhase.CreateProxy)
    DBX_HIP(h, upload(h, d.fmoved, ones));
    d.bmass = p_bmass;
    d.bparam = p_bparam;
    d.fbody = p_fbody;
    d.fshape = p_fshape;
    d.fmat = p_fmat;
    d.ffilter = p_ffilter;
    // contacts and scratch
    DBX_HIP(h, dalloc(h, &d.cont[0], nc));
    DBX_HIP(h, dalloc(h, &d.cont[1], nc));
    DBX_HIP(h, dalloc(h, &d.ctan[0], nc));
    DBX_HIP(h, dalloc(h, &d.ctan[1], nc));
    DBX_HIP(h, dalloc(h, &d.adjOff, nb));
    DBX_HIP(h, dalloc(h, &d.adjCnt, nb));
    DBX_HIP(h, dalloc(h, &d.bisl, nb));
    if (!h->wide.empty()) {
        DBX_HIP(h, dalloc(h, &d.iwParent, nb));
        DBX_HIP(h, dalloc(h, &d.iwAgg, nb));
        DBX_HIP(h, dalloc(h, &d.iwIsl, nb));
        DBX_HIP(h, dalloc(h, &d.iwCur, nb));
    }
    DBX_HIP(h, dalloc(h, &d.bscratch, nb));
    DBX_HIP(h, dalloc(h, &d.wake, nb));
    DBX_HIP(h, dalloc(h, &d.bmask, nb));
    DBX_HIP(h, dalloc(h, &d.adjList, 2 * nc));
    DBX_HIP(h, dalloc(h, &d.pairs, np));
    DBX_HIP(h, dalloc(h, &d.ordC, nc));
    DBX_HIP(h, dalloc(h, &d.cbatch, nc));
    DBX_HIP(h, dalloc(h, &d.cisl, nc));
    DBX_HIP(h, dalloc(h, &d.sched, nc));
    DBX_HIP(h, dalloc(h, &d.batchOff, nc + 65LL * nw));
    DBX_HIP(h, dalloc(h, &d.batchCur, nc + 65LL * nw));
    DBX_HIP(h, dalloc(h, &d.islScratch, islBytes));
    DBX_HIP(h, dalloc(h, &d.bpScratch, bpBytes));
    {
        size_t wideBytes = 0;
        for (WideBp &b : h->wide) {
            b.desc = h->W[b.world];
            wideBytes += wide_scratch_bytes(b);
        }
        char *wp = nullptr;
        if (wideBytes) DBX_HIP(h, dalloc(h, &wp, wideBytes));
        for (WideBp &b : h->wide) {
            wide_carve(b, wp);
            wp += wide_scratch_bytes(b);
        }
    }
    DBX_HIP(h, dalloc(h, &d.vc, nc));
    DBX_HIP(h, dalloc(h, &d.pc, nc));
    DBX_HIP(h, dalloc(h, &d.islMin, nb));
    DBX_HIP(h, dalloc(h, &d.toiC, nc));
    DBX_HIP(h, dalloc(h, &d.toiB, nb));
    DBX_HIP(h, dalloc(h, &d.toiCand, nc));
    // each body's fixtures (b2Body.m_fixtureList as a range), for the continuous pass;
    // the host arrays live until the stream synchronisation below
    std::vector<int> off(nb), cnt(nb), list(nf);
    {
        for (int w = 0; w < nw; ++w) {
            const WorldDesc &D = h->W[w];
            for (int f = 0; f < D.fixCount; ++f) cnt[D.bodyBase + fbody[D.fixBase + f]]++;
            int run = 0;
            for (int b = 0; b < D.bodyCount; ++b) {
                off[D.bodyBase + b] = run;
                run += cnt[D.bodyBase + b];
                cnt[D.bodyBase + b] = 0;
            }
            for (int f = 0; f < D.fixCount; ++f) {
                const int b = fbody[D.fixBase + f];
                list[D.fixBase + off[D.bodyBase + b] + cnt[D.bodyBase + b]++] = f;
            }
        }
        int *p_off, *p_cnt, *p_list;
        DBX_HIP(h, dalloc(h, &p_off, nb));
        DBX_HIP(h, dalloc(h, &p_cnt, nb));
        DBX_HIP(h, dalloc(h, &p_list, nf));
        DBX_HIP(h, upload(h, p_off, off));
        DBX_HIP(h, upload(h, p_cnt, cnt));
        DBX_HIP(h, upload(h, p_list, list));
        d.bfixOff = p_off;
        d.bfixCnt = p_cnt;
        d.bfixList = p_list;
    }
    // the continuous pass's listener callbacks: 128 + 4 per body for each world
    std::vector<int> logOff(nw + 1, 0);
    {
        for (int w = 0; w < nw; ++w) logOff[w + 1] = logOff[w] + 128 + 4 * h->W[w].bodyCount;
        int *p_off;
        DBX_HIP(h, dalloc(h, &p_off, nw + 1));
        DBX_HIP(h, upload(h, p_off, logOff));
        DBX_HIP(h, dalloc(h, &d.toiLog, logOff[nw]));
        d.toiLogOff = p_off;
        h->toiLogOff = logOff;
    }
    // Joints: records (definitions and kept state) from each world's jointBase, and
    // each body's joint edges in b2Body.m_jointList order (CreateJoint prepends, so the
    // newest joint comes first), offsets relative to the world's 2 * jointBase.
    {
        const long long nj = (long long)h->W[nw - 1].jointBase + h->W[nw - 1].jointCount;
        std::vector<JointRec> jr;
        jr.reserve((size_t)nj);
        std::vector<int> joff((size_t)nb, 0), jcnt((size_t)nb, 0);
        std::vector<int2> jlist((size_t)std::max<long long>(2 * nj, 1));
        for (int w = 0; w < nw; ++w) {
            const WorldDesc &D = h->W[w];
            const std::vector<JointRec> &J = h->pending[w].joints;
            jr.insert(jr.end(), J.begin(), J.end());
            for (const JointRec &j : J) {
                jcnt[D.bodyBase + j.bodyA]++;
                jcnt[D.bodyBase + j.bodyB]++;
            }
            int run = 0;
            for (int b = 0; b < D.bodyCount; ++b) {
                joff[D.bodyBase + b] = run;
                run += jcnt[D.bodyBase + b];
            }
            std::vector<int> fill(D.bodyCount, 0);
            for (int k = (int)J.size() - 1; k >= 0; --k) {
                const int a = J[k].bodyA, b = J[k].bodyB;
                jlist[2 * (size_t)D.jointBase + joff[D.bodyBase + a] + fill[a]++] = make_int2(2 * k, b);
                jlist[2 * (size_t)D.jointBase + joff[D.bodyBase + b] + fill[b]++] = make_int2(2 * k + 1, a);
            }
        }
        int *p_joff, *p_jcnt;
        int2 *p_jlist;
        DBX_HIP(h, dalloc(h, &d.joints, (size_t)std::max<long long>(nj, 1)));
        if (nj) DBX_HIP(h, upload(h, d.joints, jr));
        DBX_HIP(h, dalloc(h, &d.ordJ, (size_t)std::max<long long>(nj, 1)));
        DBX_HIP(h, dalloc(h, &p_joff, nb));
        DBX_HIP(h, dalloc(h, &p_jcnt, nb));
        DBX_HIP(h, dalloc(h, &p_jlist, jlist.size()));
        DBX_HIP(h, upload(h, p_joff, joff));
        DBX_HIP(h, upload(h, p_jcnt, jcnt));
        DBX_HIP(h, upload(h, p_jlist, jlist));
        d.jadjOff = p_joff;
        d.jadjCnt = p_jcnt;
        d.jadjList = p_jlist;
        DBX_HIP(h, dalloc(h, &d.islCoff, nb + nw));
        DBX_HIP(h, dalloc(h, &d.islJoff, nb + nw));
    }
    DBX_HIP(h, dalloc(h, &d.islSolved, nb));
    DBX_HIP(h, dalloc(h, &h->dPen, nw));
    if (getenv("DBX_KTIME")) DBX_HIP(h, dalloc(h, &d.ktime, 48));
    DBX_HIP(h, dalloc(h, &d.prof, (size_t)kProfStamps * (kProfSlots + 1)));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    h->pending.clear();
    h->pending.shrink_to_fit();
    h->finalized = true;
    return DBX_OK;
}

// Staging buffers of the learned warm start / training log, sized by the worlds.
static dbx_status stage_mlp(dbx_handle *h) {
    Dev &d = h->d;
    DBX_HIP(h, dalloc(h, &d.mlpFeat, 2 * h->totalContacts * kMlpFeatStride));
    DBX_HIP(h, dalloc(h, &d.mlpList, 2 * h->totalContacts));
    DBX_HIP(h, dalloc(h, &d.mlpNew, h->nworlds));
    DBX_HIP(h, dalloc(h, &d.mlpBody, h->totalBodies));
    h->staged = true;
    return DBX_OK;
}

// The step chain of one tier (b2World.Step, b2world.d:367-434, for its worlds).
// part bits: 1 = FindNewContacts (when fixtures were added), 2 = Collide, 4 = the
// island search and everything after it.
static dbx_status step_tier(dbx_handle *h, int t, const StepParams &sp, uint32_t flags, bool findNew,
                            const hipEvent_t *ev, int part = 7) {
    const Tier &T = h->tiers[t];
    bool lead = t == 0;
    const int cin = h->cur, cout = cin ^ 1;
    const int lds1 = std::min(T.lds, h->ldsMax - 1024); // kernels that also hold static LDS
    hipStream_t s = T.stream;
    const Dev &d = h->d;
    const WideBp *wide = T.wideCount ? &h->wide[T.wideFirst] : nullptr;
    if ((part & 1) && findNew) // FindNewContacts when fixtures were added (b2world.d:372-376)
        DBX_HIP(h, launch_broadphase(d, cin, T.first, T.count, lds1, T.block, lead, s, wide, T.wideCount));
    if (part & 2) DBX_HIP(h, launch_collide(d, cin, T.first, T.count, lds1, T.block, lead, s, wide, T.wideCount));
    if (!(part & 4)) return DBX_OK;
    const bool timed = lead && ev; // phase events: lead tier, eager steps only
    if (timed) (void)hipEventRecord(ev[1], s);
    if (sp.dt > 0.0f) {
        const int ordered = (flags & DBX_STEP_ORDERED) ? 1 : 0;
        DBX_HIP(h, launch_islands(d, cout, ordered, T.first, T.count, lds1, lead, s));
        if (ordered && T.wideCount && d.iwParent) DBX_HIP(h, launch_islands_wide(d, cout, wide, T.wideCount, s));
        if (timed) (void)hipEventRecord(ev[2], s);
        // Exact order: one wave per world with the state in LDS, except for a tier of
        // worlds too big for LDS (all with the device-wide broadphase): their levels
        // (hundreds of constraints wide in a 20,000-body container) are swept from
        // global memory by a 256-thread workgroup, level batches separated by barriers.
        if (ordered && T.wideCount < T.count) DBX_HIP(h, launch_solve_wave(d, cout, sp, T.first, T.count, lds1, lead, s));
        else DBX_HIP(h, launch_solve(d, cout, sp, T.first, T.count, ordered ? 256 : T.block, ordered ? lds1 : 0, lead, s));
        if (ordered && T.wideCount && d.iwParent && (sp.flags & kStepGrouped))
            DBX_HIP(h, launch_solve_groups(d, cout, sp, wide, T.wideCount, lds1, lead, s));
        if (timed) (void)hipEventRecord(ev[3], s);
        DBX_HIP(h, launch_broadphase(d, cout, T.first, T.count, lds1, T.block, lead, s, wide, T.wideCount));
        if (timed) (void)hipEventRecord(ev[4], s);
        // b2World.SolveTOI when continuous physics is on (b2world.d:414-419)
        if (flags & DBX_STEP_CCD) DBX_HIP(h, launch_toi(d, cout, sp, T.first, T.count, lead, s, wide, T.wideCount));
    } else if (timed) {
        (void)hipEventRecord(ev[2], s);
        (void)hipEventRecord(ev[3], s);
        (void)hipEventRecord(ev[4], s);
    }
    if (timed) (void)hipEventRecord(ev[5], s);
    return DBX_OK;
}

// Launches one step: fork the tier streams from the handle's stream, run each
// tier's chain, join.  ev (timing events of this step) is null while capturing.
static dbx_status issue_step(dbx_handle *h, const StepParams &sp, uint32_t flags, bool findNew, const hipEvent_t *ev,
                             bool timeTiers) {
    hipStream_t s = h->stream;
    hipEvent_t fork = ev ? ev[0] : h->fork;
    DBX_HIP(h, hipEventRecord(fork, s));
    for (Tier &T : h->tiers)
        if (T.own) DBX_HIP(h, hipStreamWaitEvent(T.stream, fork, 0));
    for (size_t t = 0; t < h->tiers.size(); ++t) {
        Tier &T = h->tiers[t];
        if (timeTiers) DBX_HIP(h, hipEventRecord(T.t0, T.stream));
        dbx_status st = step_tier(h, (int)t, sp, flags, findNew, ev);
        if (st != DBX_OK) return st;
        if (timeTiers) DBX_HIP(h, hipEventRecord(T.t1, T.stream));
    }
    for (Tier &T : h->tiers) {
        if (!T.own) continue;
        DBX_HIP(h, hipEventRecord(T.done, T.stream));
        DBX_HIP(h, hipStreamWaitEvent(s, T.done, 0));
    }
    if (ev) DBX_HIP(h, hipEventRecord(ev[6], s));
    return DBX_OK;
}

// The step as hipGraphs, one per tier: each tier's chain is captured once per
// distinct (device pointers, step parameters, buffer parity, launch shape) and
// replayed on the tier's own stream, so the tiers still run concurrently (one
// graph holding all tiers executes its branches one after another).  Kernel
// arguments are baked into a graph, so anything that changes them selects
// another graph.
static dbx_status tier_graph(dbx_handle *h, int t, const StepParams &sp, uint32_t flags, bool findNew,
                             hipGraphExec_t *out) {
    std::vector<unsigned char> key;
    key_append(key, t);
    key_append(key, h->d);
    key_append(key, sp);
    key_append(key, flags);
    key_append(key, h->cur);
    key_append(key, findNew);
    key_append(key, h->tiers[t].block);
    for (GraphEntry &g : h->graphs)
        if (g.key == key) {
            *out = g.exec;
            return DBX_OK;
        }
    hipStream_t s = h->tiers[t].stream;
    DBX_HIP(h, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    dbx_status st = step_tier(h, t, sp, flags, findNew, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != DBX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
    if ((int)h->graphs.size() >= kMaxGraphs * (int)h->tiers.size()) {
        DBX_HIP(h, hipDeviceSynchronize());
        (void)hipGraphExecDestroy(h->graphs.front().exec);
        h->graphs.erase(h->graphs.begin());
    }
    h->graphs.push_back(GraphEntry{key, exec});
    *out = exec;
    return DBX_OK;
}

static dbx_status step_graph(dbx_handle *h, const StepParams &sp, uint32_t flags, bool findNew, const hipEvent_t *ev) {
    std::vector<hipGraphExec_t> execs(h->tiers.size());
    for (size_t t = 0; t < h->tiers.size(); ++t) {
        dbx_status st = tier_graph(h, (int)t, sp, flags, findNew, &execs[t]);
        if (st != DBX_OK) return st;
    }
    hipStream_t s = h->stream;
    DBX_HIP(h, hipEventRecord(ev[0], s));
    for (Tier &T : h->tiers)
        if (T.own) DBX_HIP(h, hipStreamWaitEvent(T.stream, ev[0], 0));
    for (size_t t = 0; t < h->tiers.size(); ++t) DBX_HIP(h, hipGraphLaunch(execs[t], h->tiers[t].stream));
    for (Tier &T : h->tiers) {
        if (!T.own) continue;
        DBX_HIP(h, hipEventRecord(T.done, T.stream));
        DBX_HIP(h, hipStreamWaitEvent(s, T.done, 0));
    }
    DBX_HIP(h, hipEventRecord(ev[6], s));
    return DBX_OK;
}

// Autotuner bookkeeping after a tuning step: add each tier's chain time to its
// candidate, and after the last candidate keep each tier's fastest size.
static dbx_status tune_account(dbx_handle *h) {
    const int c = h->tuneDone / h->tunePer;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    for (size_t t = 0; t < h->tiers.size(); ++t) {
        float ms = 0.0f;
        DBX_HIP(h, hipEventElapsedTime(&ms, h->tiers[t].t0, h->tiers[t].t1));
        h->tuneMs[t * kTuneCandidates + c] += ms;
    }
    if (++h->tuneDone < h->tuneSteps) return DBX_OK;
    for (size_t t = 0; t < h->tiers.size(); ++t) {
        int best = 0;
        for (int k = 1; k < kTuneCandidates; ++k)
            if (h->tuneMs[t * kTuneCandidates + k] < h->tuneMs[t * kTuneCandidates + best]) best = k;
        h->tiers[t].block = kTuneBlocks[best];
    }
    if (getenv("DBX_VERBOSE"))
        for (size_t t = 0; t < h->tiers.size(); ++t) {
            fprintf(stderr, "dbx: autotune tier %zu:", t);
            for (int k = 0; k < kTuneCandidates; ++k)
                fprintf(stderr, " %d:%.3fms", kTuneBlocks[k], h->tuneMs[t * kTuneCandidates + k] / h->tunePer);
            fprintf(stderr, " -> %d\n", h->tiers[t].block);
        }
    return DBX_OK;
}

dbx_status dbx_step(dbx_handle *h, float dt, int32_t vel, int32_t pos, uint32_t flags) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (h->splitWaiting) return fail(h, DBX_E_STATE, "a PRESOLVE step is waiting for dbx_step_finish");
    if (!std::isfinite(dt) || vel < 0 || pos < 0) return fail(h, DBX_E_INVALID_ARG, "bad step parameters");
    if ((flags & DBX_STEP_LEARNED) && !h->mlp.ready) return fail(h, DBX_E_STATE, "no warm-start network loaded");
    if ((flags & DBX_STEP_LEARNED) && !(flags & DBX_STEP_WARM))
        return fail(h, DBX_E_INVALID_ARG, "the learned warm start needs warm starting");
    if ((flags & DBX_STEP_LOG) && !h->logX) return fail(h, DBX_E_STATE, "the training log is not enabled");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    StepParams sp{};
    sp.dt = dt;
    sp.inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
    sp.dtRatio = h->inv_dt0 * dt;
    sp.velocityIterations = vel;
    sp.positionIterations = pos;
    sp.flags = ((flags & DBX_STEP_WARM) ? 1u : 0u) | ((flags & DBX_STEP_SLEEP) ? 2u : 0u) |
               ((flags & DBX_STEP_RESIDUAL) ? (1u << 5) : 0u) | ((flags & DBX_STEP_LEARNED) ? (1u << 4) : 0u) |
               ((flags & DBX_STEP_LOG) ? (1u << 6) : 0u);
    if ((flags & DBX_STEP_ORDERED) && !(flags & (DBX_STEP_RESIDUAL | DBX_STEP_LEARNED | DBX_STEP_LOG)) &&
        !getenv("DBX_SOLVE_UNGROUPED"))
        sp.flags |= kStepGrouped;
    if ((flags & DBX_STEP_ORDERED) && !(flags & DBX_STEP_RESIDUAL) && !getenv("DBX_SWEEPS_SEQUENTIAL"))
        sp.flags |= kStepPiped | (getenv("DBX_SWEEPS_OVERLAP_ALWAYS") ? kStepPipedAlways : 0u);
    if ((flags & (DBX_STEP_LEARNED | DBX_STEP_LOG)) && !h->staged) {
        dbx_status st = stage_mlp(h);
        if (st != DBX_OK) return st;
    }
    sp.residualTol = h->residualTol;
    sp.maxIterations = h->maxIterations;
    if ((int)h->ring.size() < kStepEvents * (h->ringUsed + 1)) {
        for (int i = 0; i < kStepEvents; ++i) {
            hipEvent_t e;
            DBX_HIP(h, hipEventCreate(&e));
            h->ring.push_back(e);
        }
        h->ringGraph.push_back(0);
        h->ringSolved.push_back(0);
    }
    hipEvent_t *ev = &h->ring[kStepEvents * h->ringUsed];
    h->d.mlp = mlp_dev(h->mlp);
    h->d.logX = h->logX;
    h->d.logT = h->logT;
    h->d.logCount = h->logCount;
    h->d.logCap = h->logCap;
    const bool findNew = h->newFixture;
    h->lastCcd = (flags & DBX_STEP_CCD) != 0;
    if (flags & DBX_STEP_PRESOLVE) {
        // b2World.Step up to the end of Collide, every tier in turn on the handle's
        // stream: FindNewContacts, a copy of the list Collide is about to walk (the
        // old manifolds that PreSolve receives), Collide.  dbx_step_finish runs the rest.
        if (!h->contPre) DBX_HIP(h, dalloc(h, &h->contPre, h->totalContacts));
        for (size_t t = 0; t < h->tiers.size(); ++t) {
            hipStream_t ts = h->tiers[t].stream;
            h->tiers[t].stream = h->stream;
            dbx_status st = step_tier(h, (int)t, sp, flags, findNew, nullptr, 1);
            h->tiers[t].stream = ts;
            if (st != DBX_OK) return st;
        }
        DBX_HIP(h, hipMemcpyAsync(h->contPre, h->d.cont[h->cur], h->totalContacts * sizeof(Contact),
                                  hipMemcpyDeviceToDevice, h->stream));
        for (size_t t = 0; t < h->tiers.size(); ++t) {
            hipStream_t ts = h->tiers[t].stream;
            h->tiers[t].stream = h->stream;
            dbx_status st = step_tier(h, (int)t, sp, flags, findNew, nullptr, 2);
            h->tiers[t].stream = ts;
            if (st != DBX_OK) return st;
        }
        h->newFixture = false;
        h->splitWaiting = true;
        h->lastSplit = true;
        h->pendSp = sp;
        h->pendFlags = flags;
        return DBX_OK;
    }
    h->lastSplit = false;
    const bool tuning = h->tuneDone < h->tuneSteps;
    if (tuning) {
        const int c = h->tuneDone / h->tunePer;
        for (Tier &T : h->tiers) T.block = kTuneBlocks[c];
    }
    // DBX_TIER_TIMES (diagnostics): time every tier's chain; each step then waits
    // for its own end, so the steps no longer overlap the host
    static const bool tierTimes = getenv("DBX_TIER_TIMES") != nullptr;
    const bool timeTiers = tuning || (tierTimes && !h->cfg.use_graph);
    if (timeTiers)
        for (Tier &T : h->tiers) {
            if (!T.t0) DBX_HIP(h, hipEventCreate(&T.t0));
            if (!T.t1) DBX_HIP(h, hipEventCreate(&T.t1));
        }
    dbx_status st = DBX_OK;
    if (h->cfg.use_graph && !tuning) {
        st = step_graph(h, sp, flags, findNew, ev);
        h->ringGraph[h->ringUsed] = 1;
    } else {
        st = issue_step(h, sp, flags, findNew, ev, timeTiers);
        h->ringGraph[h->ringUsed] = 0;
    }
    if (st != DBX_OK) return st;
    h->ringSolved[h->ringUsed] = dt > 0.0f;
    if (timeTiers && !tuning) {
        DBX_HIP(h, hipStreamSynchronize(h->stream));
        h->tierMs.resize(h->tiers.size(), 0.0);
        for (size_t t = 0; t < h->tiers.size(); ++t) {
            float ms = 0.0f;
            DBX_HIP(h, hipEventElapsedTime(&ms, h->tiers[t].t0, h->tiers[t].t1));
            h->tierMs[t] += ms;
        }
        h->tierSteps++;
    }
    h->ringUsed++;
    h->newFixture = false;
    h->cur ^= 1;
    if (dt > 0.0f) h->inv_dt0 = sp.inv_dt;
    if (tuning) return tune_account(h);
    return DBX_OK;
}

static dbx_status check_world(dbx_handle *h, int32_t world);

dbx_status dbx_step_finish(dbx_handle *h) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->splitWaiting) return fail(h, DBX_E_STATE, "no PRESOLVE step is waiting");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    for (size_t t = 0; t < h->tiers.size(); ++t) {
        hipStream_t ts = h->tiers[t].stream;
        h->tiers[t].stream = h->stream;
        dbx_status st = step_tier(h, (int)t, h->pendSp, h->pendFlags, false, nullptr, 4);
        h->tiers[t].stream = ts;
        if (st != DBX_OK) return st;
    }
    h->splitWaiting = false;
    h->cur ^= 1;
    if (h->pendSp.dt > 0.0f) h->inv_dt0 = h->pendSp.inv_dt;
    return DBX_OK;
}

dbx_status dbx_write_contact_settings(dbx_handle *h, int32_t world, const dbx_contact_setting *s, int32_t n) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    if (!h->splitWaiting) return fail(h, DBX_E_STATE, "contact settings are written between dbx_step(PRESOLVE) and dbx_step_finish");
    if (n < 0 || (n > 0 && !s)) return fail(h, DBX_E_INVALID_ARG, "bad settings array");
    const WorldDesc &D = h->W[world];
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    const int m = S.contactCount, buf = h->cur ^ 1; // the list Collide left
    std::vector<Contact> c(m);
    std::vector<float> tan(m);
    if (m) {
        DBX_HIP(h, hipMemcpy(c.data(), h->d.cont[buf] + D.contactBase, m * sizeof(Contact), hipMemcpyDeviceToHost));
        DBX_HIP(h, hipMemcpy(tan.data(), h->d.ctan[buf] + D.contactBase, m * sizeof(float), hipMemcpyDeviceToHost));
    }
    std::unordered_map<long long, int> at;
    at.reserve(2 * (size_t)m);
    for (int i = 0; i < m; ++i) {
        at[((long long)c[i].fa << 32) | (unsigned)c[i].fb] = i;
        at[((long long)c[i].fb << 32) | (unsigned)c[i].fa] = i;
    }
    for (int k = 0; k < n; ++k) {
        auto it = at.find(((long long)s[k].fixture_a << 32) | (unsigned)s[k].fixture_b);
        if (it == at.end()) return fail(h, DBX_E_INVALID_ARG, "no contact between these fixtures");
        if (!std::isfinite(s[k].friction) || !std::isfinite(s[k].restitution) || !std::isfinite(s[k].tangent_speed))
            return fail(h, DBX_E_INVALID_ARG, "contact setting is not finite");
        Contact &x = c[it->second];
        const int mk = s[k].mask;
        // b2Contact.SetEnabled / SetFriction / SetRestitution / SetTangentSpeed (b2contact.d:108-198)
        if (mk & DBX_SET_ENABLED) x.flags = s[k].enabled ? (x.flags | kContactEnabled) : (x.flags & ~kContactEnabled);
        if (mk & DBX_SET_FRICTION) x.friction = s[k].friction;
        if (mk & DBX_SET_RESTITUTION) x.restitution = s[k].restitution;
        if (mk & DBX_SET_TANGENT_SPEED) {
            tan[it->second] = s[k].tangent_speed;
            if (s[k].tangent_speed != 0.0f) h->d.useTan = 1; // from now on the slots' speeds move with the contacts
        }
    }
    if (m) {
        DBX_HIP(h, hipMemcpy(h->d.cont[buf] + D.contactBase, c.data(), m * sizeof(Contact), hipMemcpyHostToDevice));
        DBX_HIP(h, hipMemcpy(h->d.ctan[buf] + D.contactBase, tan.data(), m * sizeof(float), hipMemcpyHostToDevice));
    }
    return DBX_OK;
}

dbx_status dbx_sync(dbx_handle *h) {
    if (!h) return DBX_E_INVALID_ARG;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    if (!h->finalized) return DBX_OK;
    std::vector<WorldState> S(h->nworlds);
    DBX_HIP(h, hipMemcpy(S.data(), h->d.S, h->nworlds * sizeof(WorldState), hipMemcpyDeviceToHost));
    for (int w = 0; w < h->nworlds; ++w) {
        if (S[w].error) {
            // bit 1: contact slots; bit 2: the step's broadphase pair list; bit 4: the
            // continuous pass's per-event lists (pairs of the moved proxies, moved proxies)
            char buf[160];
            const int e = S[w].error;
            snprintf(buf, sizeof buf, "world %d exceeded its %s capacity", w,
                     (e & 1) ? "contact" : (e & 2) ? "broadphase pair" : "continuous-pass pair/moved-proxy list");
            return fail(h, DBX_E_CAPACITY, buf);
        }
    }
    // b2Profile's island-solve phases (b2island.d:118-237) from the stamps the largest
    // world's solve left: one slot per step that ran the solve, in step order.
    std::vector<unsigned long long> stamps((size_t)kProfStamps * (kProfSlots + 1));
    DBX_HIP(h, hipMemcpy(stamps.data(), h->d.prof, stamps.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    const unsigned long long profCount = stamps[0];
    int solvedSteps = 0;
    for (int k = 0; k < h->ringUsed; ++k) solvedSteps += h->ringSolved[k] ? 1 : 0;
    const bool stampsMatch = profCount - h->profSeen == (unsigned long long)solvedSteps;
    unsigned long long nextSlot = h->profSeen;
    h->profSeen = profCount;
    for (int k = 0; k < h->ringUsed; ++k) {
        hipEvent_t *ev = &h->ring[kStepEvents * k];
        float ms[5] = {0, 0, 0, 0, 0}, total = 0.0f;
        float phase[3] = {0, 0, 0};
        if (h->ringSolved[k]) {
            const unsigned long long slot = nextSlot++;
            if (stampsMatch && profCount - slot <= (unsigned long long)kProfSlots) {
                const unsigned long long *t = &stamps[(size_t)kProfStamps * (slot % kProfSlots + 1)];
                const double perMs = h->wallKHz;
                phase[0] = (float)((double)(t[kProfVel0] - t[kProfInit0]) / perMs);
                phase[1] = (float)((double)(t[kProfVel1] - t[kProfVel0]) / perMs);
                phase[2] = (float)((double)(t[kProfPos1] - t[kProfPos0]) / perMs);
            }
        }
        if (!h->ringGraph[k]) // a replayed graph records only the step's ends
            for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
        (void)hipEventElapsedTime(&total, ev[0], ev[6]);
        // b2World.Step's fields (b2world.d:397-431): collide; solve = all of b2World.Solve,
        // i.e. the island search, the island solves, SynchronizeFixtures and FindNewContacts
        // (whose share is also broadphase, b2world.d:1102-1123); the island-solve phases
        // accumulate over islands (b2world.d:1079-1081), here the phases of the largest
        // world, which the other worlds run beside.
        dbx_profile p{};
        p.collide = ms[0];
        p.solve = ms[1] + ms[2] + ms[3];
        p.solveInit = phase[0];
        p.solveVelocity = phase[1];
        p.solvePosition = phase[2];
        p.broadphase = ms[3];
        p.solveTOI = ms[4];
        p.step = total;
        h->prof = p;
        h->sum.step += p.step;
        h->sum.collide += p.collide;
        h->sum.solve += p.solve;
        h->sum.solveInit += p.solveInit;
        h->sum.solveVelocity += p.solveVelocity;
        h->sum.solvePosition += p.solvePosition;
        h->sum.broadphase += p.broadphase;
        h->sum.solveTOI += p.solveTOI;
        h->sumSteps++;
        if (!h->ringGraph[k]) {
            for (int i = 0; i < 5; ++i) h->phaseSum[i] += ms[i];
            h->phaseSteps++;
        }
    }
    h->ringUsed = 0;
    if (h->tierSteps > 0) {
        fprintf(stderr, "dbx tier chain ms/step over %d steps:", h->tierSteps);
        for (size_t t = 0; t < h->tierMs.size(); ++t) fprintf(stderr, " [%zu] %.3f", t, h->tierMs[t] / h->tierSteps);
        fprintf(stderr, "\n");
        h->tierMs.assign(h->tierMs.size(), 0.0);
        h->tierSteps = 0;
    }
    if (h->d.ktime && getenv("DBX_KTIME")) {
        long long t[48];
        if (hipMemcpy(t, h->d.ktime, sizeof t, hipMemcpyDeviceToHost) == hipSuccess) {
            fprintf(stderr, "dbx ktime (shader cycles) islands: stage %lld dfs %lld batch %lld [nc %lld nbatch %lld] | "
                            "solve: stage+init %lld warm %lld vel %lld pos %lld tail %lld\n",
                    t[1] - t[0], t[2] - t[1], t[3] - t[2], t[15], t[14], t[5] - t[4], t[6] - t[5], t[7] - t[6],
                    t[8] - t[7], t[9] - t[8]);
            fprintf(stderr, "dbx ktime collide: narrowphase %lld dormant+wake %lld compact %lld adj-build %lld "
                            "adj-sort %lld | broadphase: pairs %lld addpair %lld\n",
                    t[17] - t[16], t[18] - t[17], t[19] - t[18], t[20] - t[19], t[21] - t[20], t[23] - t[22],
                    t[24] - t[23]);
            fprintf(stderr, "dbx ktime toi: init+first search %lld, then %lld events: searches %lld events %lld "
                            "(total %lld, %lld contacts)\n",
                    t[26] - t[25], t[28], t[29], t[30], t[27] - t[25], t[31]);
            fprintf(stderr, "dbx ktime toi kernel entry to exit: %.1f us wall clock, %lld shader cycles\n",
                    (double)(t[34] - t[32]) * 1000.0 / h->wallKHz, t[35] - t[33]);
            fprintf(stderr, "dbx ktime solve waves' HW_ID (overlapped sweeps): %llx %llx; solve %lld shader cycles in "
                            "%.1f us wall clock\n",
                    (unsigned long long)t[10], (unsigned long long)t[11], t[9] - t[4],
                    (double)(t[13] - t[12]) * 1000.0 / h->wallKHz);
            fprintf(stderr, "dbx ktime big-world LDS solve: count %lld number %lld init %lld warm %lld velocity %lld "
                            "store %lld position %lld [bodies %lld constraints %lld levels %lld]\n",
                    t[37] - t[36], t[38] - t[37], t[39] - t[38], t[40] - t[39], t[41] - t[40], t[42] - t[41],
                    t[44] - t[43], t[45], t[46], t[47]);
        }
    }
    return DBX_OK;
}

dbx_status dbx_read_tiers(dbx_handle *h, int32_t *tiers, int32_t cap, int32_t *order, int32_t *count) {
    if (!h || !count || (cap > 0 && !tiers)) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    *count = (int32_t)h->tiers.size();
    for (int i = 0; i < cap && i < (int)h->tiers.size(); ++i) {
        tiers[4 * i + 0] = h->tiers[i].first;
        tiers[4 * i + 1] = h->tiers[i].count;
        tiers[4 * i + 2] = h->tiers[i].lds;
        tiers[4 * i + 3] = h->tiers[i].block;
    }
    if (order) DBX_HIP(h, hipMemcpy(order, h->d.order, h->nworlds * sizeof(int), hipMemcpyDeviceToHost));
    return DBX_OK;
}

dbx_status dbx_world_count(const dbx_handle *h, int32_t *count) {
    if (!h || !count) return DBX_E_INVALID_ARG;
    *count = h->finalized ? h->nworlds : (int32_t)h->pending.size();
    return DBX_OK;
}

static dbx_status check_world(dbx_handle *h, int32_t world) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (world < 0 || world >= h->nworlds) return fail(h, DBX_E_INVALID_ARG, "world index out of range");
    return DBX_OK;
}

dbx_status dbx_world_info(dbx_handle *h, int32_t world, int32_t out[4]) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    out[0] = h->W[world].bodyCount;
    out[1] = h->W[world].fixCount;
    out[2] = S.contactCount;
    out[3] = S.touching;
    return DBX_OK;
}

dbx_status dbx_read_bodies(dbx_handle *h, int32_t world, dbx_body_state *out, int32_t n) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    if (!out || n < D.bodyCount) return fail(h, DBX_E_INVALID_ARG, "output too small");
    int nb = D.bodyCount, b0 = D.bodyBase;
    std::vector<uint32_t> fl(nb);
    std::vector<float4> xf(nb), pos(nb), pos0(nb), vel(nb), ms(nb);
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(fl.data(), h->d.bflags + b0, nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(xf.data(), h->d.bxf + b0, nb * sizeof(float4), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(pos.data(), h->d.bpos + b0, nb * sizeof(float4), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(pos0.data(), h->d.bpos0 + b0, nb * sizeof(float4), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(vel.data(), h->d.bvel + b0, nb * sizeof(float4), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(ms.data(), h->d.bmass + b0, nb * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; ++i) {
        dbx_body_state &s = out[i];
        s.px = xf[i].x;
        s.py = xf[i].y;
        s.qs = xf[i].z;
        s.qc = xf[i].w;
        s.cx = pos[i].x;
        s.cy = pos[i].y;
        s.a = pos[i].z;
        s.sleep_time = pos[i].w;
        s.vx = vel[i].x;
        s.vy = vel[i].y;
        s.w = vel[i].z;
        s.flags = (int32_t)(fl[i] & 0xffffu);
        s.c0x = pos0[i].x;
        s.c0y = pos0[i].y;
        s.a0 = pos0[i].z;
        s.inv_mass = ms[i].x;
    }
    return DBX_OK;
}

dbx_status dbx_write_bodies(dbx_handle *h, int32_t world, const dbx_body_state *in, const float *force3, int32_t n) {
    dbx_status st = check_world(h, world);
    if (h->splitWaiting) return fail(h, DBX_E_STATE, "a PRESOLVE step is waiting for dbx_step_finish");
    if (st) return st;
    const WorldDesc &D = h->W[world];
    if (!in || n != D.bodyCount) return fail(h, DBX_E_INVALID_ARG, "one state per body is required");
    const int nb = D.bodyCount, b0 = D.bodyBase;
    std::vector<uint32_t> fl(nb);
    std::vector<float4> xf(nb), pos(nb), pos0(nb), vel(nb), frc;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(fl.data(), h->d.bflags + b0, nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; ++i) {
        const dbx_body_state &s = in[i];
        const float v[] = {s.px, s.py, s.qs, s.qc, s.cx, s.cy, s.a, s.sleep_time, s.vx, s.vy, s.w, s.c0x, s.c0y, s.a0};
        for (float x : v)
            if (!std::isfinite(x)) return fail(h, DBX_E_INVALID_ARG, "body state is not finite");
        xf[i] = make_float4(s.px, s.py, s.qs, s.qc);
        pos[i] = make_float4(s.cx, s.cy, s.a, s.sleep_time);
        pos0[i] = make_float4(s.c0x, s.c0y, s.a0, 0.0f);
        vel[i] = make_float4(s.vx, s.vy, s.w, 0.0f);
        fl[i] = (fl[i] & ~(uint32_t)kBodyAwake) | ((uint32_t)s.flags & kBodyAwake);
    }
    if (force3) {
        frc.resize(nb);
        for (int i = 0; i < nb; ++i) {
            if (!std::isfinite(force3[3 * i]) || !std::isfinite(force3[3 * i + 1]) || !std::isfinite(force3[3 * i + 2]))
                return fail(h, DBX_E_INVALID_ARG, "force is not finite");
            frc[i] = make_float4(force3[3 * i], force3[3 * i + 1], force3[3 * i + 2], 0.0f);
        }
    }
    DBX_HIP(h, hipMemcpy(h->d.bflags + b0, fl.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice));
    DBX_HIP(h, hipMemcpy(h->d.bxf + b0, xf.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    DBX_HIP(h, hipMemcpy(h->d.bpos + b0, pos.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    DBX_HIP(h, hipMemcpy(h->d.bpos0 + b0, pos0.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    DBX_HIP(h, hipMemcpy(h->d.bvel + b0, vel.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    if (force3) DBX_HIP(h, hipMemcpy(h->d.bforce + b0, frc.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    DBX_HIP(h, launch_refit(h->d, world, h->stream));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    return DBX_OK;
}

dbx_status dbx_read_events(dbx_handle *h, int32_t world, dbx_contact_event *out, int32_t cap, int32_t *count) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    if (!count) return fail(h, DBX_E_INVALID_ARG, "count is required");
    const WorldDesc &D = h->W[world];
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    // The last step's Collide walked the list held in the other buffer, where its
    // flags stay; the solved contacts are in the current list, in island order.
    // While a PRESOLVE step waits, Collide's walked list is still the current one
    // and nothing has been solved yet.
    const int walkedBuf = h->splitWaiting ? h->cur : h->cur ^ 1, nowBuf = walkedBuf ^ 1;
    const int n0 = S.collideCount, m = S.contactCount, nc = h->splitWaiting ? 0 : S.constraints;
    std::vector<Contact> walked(n0), now(m), pre;
    std::vector<int32_t> ord(nc);
    if (n0) DBX_HIP(h, hipMemcpy(walked.data(), h->d.cont[walkedBuf] + D.contactBase, n0 * sizeof(Contact),
                                 hipMemcpyDeviceToHost));
    if (m) DBX_HIP(h, hipMemcpy(now.data(), h->d.cont[nowBuf] + D.contactBase, m * sizeof(Contact), hipMemcpyDeviceToHost));
    if (nc) DBX_HIP(h, hipMemcpy(ord.data(), h->d.ordC + D.contactBase, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    const bool preSolve = h->lastSplit && h->contPre;
    if (preSolve && n0) {
        pre.resize(n0);
        DBX_HIP(h, hipMemcpy(pre.data(), h->contPre + D.contactBase, n0 * sizeof(Contact), hipMemcpyDeviceToHost));
    }
    // the continuous pass's log: its callbacks, and the discrete PostSolve data of
    // the contacts it updated (kLogSnapshot), which their slots no longer hold
    std::vector<ToiCallback> log;
    if (h->lastCcd && !h->splitWaiting) {
        const int cap = h->toiLogOff[world + 1] - h->toiLogOff[world];
        if (S.toiCallbacks > cap)
            return fail(h, DBX_E_CAPACITY, "the continuous pass made more listener callbacks than its log holds");
        log.resize(S.toiCallbacks);
        if (!log.empty())
            DBX_HIP(h, hipMemcpy(log.data(), h->d.toiLog + h->toiLogOff[world], log.size() * sizeof(ToiCallback),
                                 hipMemcpyDeviceToHost));
    }
    std::unordered_map<int, const ToiCallback *> solvedBefore;
    for (const ToiCallback &r : log)
        if (r.type == kLogSnapshot) solvedBefore.emplace(r.fa, &r);
    std::vector<dbx_contact_event> ev;
    ev.reserve(nc + 16);
    for (int i = n0 - 1; i >= 0; --i) { // list order: newest first
        const Contact &c = walked[i];
        dbx_contact_event e{};
        e.fixture_a = c.fa;
        e.fixture_b = c.fb;
        if (c.flags & kContactBegan) {
            e.type = DBX_EVENT_BEGIN;
            ev.push_back(e);
        } else if (c.flags & kContactEnded) {
            e.type = DBX_EVENT_END;
            ev.push_back(e);
        }
        // PreSolve after the update of a contact that touches (b2contact.d:352-355),
        // with the manifold as it was before the update
        if (preSolve && (c.flags & kContactTouching) && !(c.flags & (kContactDormant | kContactDestroyed | kContactSensor))) {
            const Contact &o = pre[i];
            dbx_contact_event p{};
            p.type = DBX_EVENT_PRE_SOLVE;
            p.fixture_a = c.fa;
            p.fixture_b = c.fb;
            p.point_count = (o.meta >> 4) & 0xf;
            if (p.point_count > 0) {
                p.normal_impulse[0] = o.pt0.z;
                p.tangent_impulse[0] = o.pt0.w;
            }
            if (p.point_count > 1) {
                p.normal_impulse[1] = o.pt1.z;
                p.tangent_impulse[1] = o.pt1.w;
            }
            ev.push_back(p);
        }
    }
    for (int k = 0; k < nc; ++k) {
        if (ord[k] < 0 || ord[k] >= m) return fail(h, DBX_E_STATE, "solve order out of range");
        const Contact &c = now[ord[k]];
        dbx_contact_event e{};
        e.type = DBX_EVENT_POST_SOLVE;
        e.fixture_a = c.fa;
        e.fixture_b = c.fb;
        auto snap = solvedBefore.find(ord[k]);
        if (snap != solvedBefore.end()) {
            const ToiCallback &r = *snap->second;
            e.point_count = r.pointCount;
            e.normal_impulse[0] = r.n0;
            e.tangent_impulse[0] = r.t0;
            e.normal_impulse[1] = r.n1;
            e.tangent_impulse[1] = r.t1;
            ev.push_back(e);
            continue;
        }
        e.point_count = (c.flags & kContactSolvedOne) ? 1 : ((c.meta >> 4) & 0xf);
        e.normal_impulse[0] = c.pt0.z;
        e.tangent_impulse[0] = c.pt0.w;
        if (e.point_count > 1) {
            e.normal_impulse[1] = c.pt1.z;
            e.tangent_impulse[1] = c.pt1.w;
        }
        ev.push_back(e);
    }
    // the continuous pass's callbacks, which come after the discrete solve's
    // (b2world.d:425-428); PRE_SOLVE only for a split step, as above
    static_assert(sizeof(ToiCallback) == sizeof(dbx_contact_event), "callback records share one layout");
    for (const ToiCallback &r : log) {
        if (r.type == kLogSnapshot || (r.type == DBX_EVENT_PRE_SOLVE && !h->lastSplit)) continue;
        dbx_contact_event e;
        std::memcpy(&e, &r, sizeof e);
        ev.push_back(e);
    }
    *count = (int32_t)ev.size();
    if (!out) return DBX_OK;
    if (cap < (int32_t)ev.size()) return fail(h, DBX_E_CAPACITY, "event buffer too small");
    std::copy(ev.begin(), ev.end(), out);
    return DBX_OK;
}

dbx_status dbx_read_proxies(dbx_handle *h, int32_t world, int32_t *fixture, int32_t *child, int32_t cap,
                            int32_t *count) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    if (!count) return fail(h, DBX_E_INVALID_ARG, "count is required");
    const WorldDesc &D = h->W[world];
    *count = D.fixCount;
    if (cap < D.fixCount || !fixture || !child) return DBX_OK;
    for (int i = 0; i < D.fixCount; ++i) {
        fixture[i] = h->proxyFix[D.fixBase + i];
        child[i] = h->proxyChild[D.fixBase + i];
    }
    return DBX_OK;
}

dbx_status dbx_read_contacts(dbx_handle *h, int32_t world, dbx_contact *out, int32_t cap, int32_t *count) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    int n = S.contactCount;
    if (count) *count = n;
    if (!out || cap <= 0) return DBX_OK;
    std::vector<Contact> c(n);
    const int buf = h->splitWaiting ? h->cur ^ 1 : h->cur; // while a PRESOLVE step waits: the list Collide left
    if (n) DBX_HIP(h, hipMemcpy(c.data(), h->d.cont[buf] + D.contactBase, n * sizeof(Contact), hipMemcpyDeviceToHost));
    int m = std::min(n, cap);
    for (int k = 0; k < m; ++k) {
        const Contact &s = c[n - 1 - k]; // list order: newest first
        dbx_contact &o = out[k];
        o.fixture_a = s.fa;
        o.fixture_b = s.fb;
        o.flags = (int32_t)(s.flags & (kContactTouching | kContactEnabled));
        o.type = s.meta & 0xf;
        o.point_count = (s.meta >> 4) & 0xf;
        o.local_normal[0] = s.man.x;
        o.local_normal[1] = s.man.y;
        o.local_point[0] = s.man.z;
        o.local_point[1] = s.man.w;
        o.point_local[0] = s.pt0.x;
        o.point_local[1] = s.pt0.y;
        o.point_local[2] = s.pt1.x;
        o.point_local[3] = s.pt1.y;
        o.normal_impulse[0] = s.pt0.z;
        o.normal_impulse[1] = s.pt1.z;
        o.tangent_impulse[0] = s.pt0.w;
        o.tangent_impulse[1] = s.pt1.w;
        o.id[0] = s.id0;
        o.id[1] = s.id1;
        o.friction = s.friction;
        o.restitution = s.restitution;
    }
    return DBX_OK;
}

dbx_status dbx_read_joints(dbx_handle *h, int32_t world, float *out, int32_t n) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    if (n != D.jointCount || (n > 0 && !out)) return fail(h, DBX_E_INVALID_ARG, "n must be the world's joint count");
    if (n == 0) return DBX_OK;
    std::vector<JointRec> J(n);
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(J.data(), h->d.joints + D.jointBase, n * sizeof(JointRec), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        const JointRec &j = J[i];
        float *o = out + 8 * i;
        const bool lin = j.type == DBX_JOINT_FRICTION || j.type == DBX_JOINT_MOTOR || j.type == DBX_JOINT_MOUSE;
        o[0] = lin ? j.linearImpulse.x : j.ix;
        o[1] = lin ? j.linearImpulse.y : j.iy;
        o[2] = lin ? j.angularImpulse : j.iz;
        o[3] = j.motorImpulse;
        o[4] = j.springImpulse;
        o[5] = (float)j.limitState;
        o[6] = 0.0f;
        o[7] = 0.0f;
    }
    return DBX_OK;
}

dbx_status dbx_read_fat_aabbs(dbx_handle *h, int32_t world, float *out, int32_t n) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    if (!out || n < D.fixCount) return fail(h, DBX_E_INVALID_ARG, "output too small");
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(out, h->d.ffat + D.fixBase, D.fixCount * sizeof(float4), hipMemcpyDeviceToHost));
    return DBX_OK;
}

dbx_status dbx_read_pairs(dbx_handle *h, int32_t world, int32_t *out, int32_t cap, int32_t *count) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    if (count) *count = S.pairCount;
    int m = std::min(S.pairCount, cap);
    if (out && m > 0) DBX_HIP(h, hipMemcpy(out, h->d.pairs + D.pairBase, m * sizeof(int2), hipMemcpyDeviceToHost));
    return DBX_OK;
}

dbx_status dbx_profile_read(dbx_handle *h, dbx_profile *out) {
    if (!h || !out) return DBX_E_INVALID_ARG;
    *out = h->prof;
    return DBX_OK;
}

dbx_status dbx_profile_totals(dbx_handle *h, dbx_profile *sum, int32_t *steps, int32_t reset) {
    if (!h) return DBX_E_INVALID_ARG;
    if (sum) *sum = h->sum;
    if (steps) *steps = h->sumSteps;
    if (reset) {
        h->sum = dbx_profile{};
        h->sumSteps = 0;
    }
    return DBX_OK;
}

dbx_status dbx_phase_totals(dbx_handle *h, float *ms5, int32_t *steps, int32_t reset) {
    if (!h || !ms5 || !steps) return DBX_E_INVALID_ARG;
    for (int i = 0; i < 5; ++i) ms5[i] = (float)h->phaseSum[i];
    *steps = h->phaseSteps;
    if (reset) {
        for (double &x : h->phaseSum) x = 0.0;
        h->phaseSteps = 0;
    }
    return DBX_OK;
}

dbx_status dbx_totals(dbx_handle *h, int64_t out[4]) {
    if (!h || !out) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    std::vector<WorldState> S(h->nworlds);
    DBX_HIP(h, hipMemcpy(S.data(), h->d.S, h->nworlds * sizeof(WorldState), hipMemcpyDeviceToHost));
    std::vector<uint32_t> fl(h->totalBodies);
    DBX_HIP(h, hipMemcpy(fl.data(), h->d.bflags, fl.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t dyn = 0, nc = 0, touch = 0, awake = 0;
    for (int w = 0; w < h->nworlds; ++w) {
        dyn += h->dynamicCount[w];
        nc += S[w].contactCount;
        touch += S[w].touching;
    }
    for (uint32_t f : fl) awake += (body_type(f) == kDynamic) && (f & kBodyAwake);
    out[0] = dyn;
    out[1] = nc;
    out[2] = touch;
    out[3] = awake;
    return DBX_OK;
}

dbx_status dbx_read_solver_stats(dbx_handle *h, int32_t *iterations, float *max_penetration, int32_t nworlds) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (nworlds < h->nworlds) return fail(h, DBX_E_INVALID_ARG, "output too small");
    std::vector<WorldState> S(h->nworlds);
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(S.data(), h->d.S, h->nworlds * sizeof(WorldState), hipMemcpyDeviceToHost));
    if (iterations)
        for (int w = 0; w < h->nworlds; ++w) iterations[w] = S[w].iterations;
    if (max_penetration) {
        DBX_HIP(h, launch_max_penetration(h->d, h->cur, h->dPen, h->stream));
        DBX_HIP(h, hipMemcpyAsync(max_penetration, h->dPen, h->nworlds * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        DBX_HIP(h, hipStreamSynchronize(h->stream));
    }
    return DBX_OK;
}

dbx_status dbx_read_toi_events(dbx_handle *h, int32_t *events, int32_t nworlds) {
    if (!h || !events) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (nworlds < h->nworlds) return fail(h, DBX_E_INVALID_ARG, "output too small");
    std::vector<WorldState> S(h->nworlds);
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(S.data(), h->d.S, h->nworlds * sizeof(WorldState), hipMemcpyDeviceToHost));
    for (int w = 0; w < h->nworlds; ++w) events[w] = S[w].toiEvents;
    return DBX_OK;
}

dbx_status dbx_read_residuals(dbx_handle *h, float *residual, int32_t nworlds) {
    if (!h || !residual) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (nworlds < h->nworlds) return fail(h, DBX_E_INVALID_ARG, "output too small");
    std::vector<WorldState> S(h->nworlds);
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(S.data(), h->d.S, h->nworlds * sizeof(WorldState), hipMemcpyDeviceToHost));
    for (int w = 0; w < h->nworlds; ++w) residual[w] = S[w].residual;
    return DBX_OK;
}

dbx_status dbx_read_schedule(dbx_handle *h, int32_t world, int32_t *out, int32_t cap, int32_t *count) {
    dbx_status st = check_world(h, world);
    if (st) return st;
    const WorldDesc &D = h->W[world];
    WorldState S;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&S, h->d.S + world, sizeof S, hipMemcpyDeviceToHost));
    int n = S.constraints;
    if (count) *count = n;
    if (!out || cap <= 0 || n == 0) return DBX_OK;
    std::vector<int> ord(n), isl(n), bat(n), fb(D.fixCount);
    std::vector<Contact> c(S.contactCount);
    DBX_HIP(h, hipMemcpy(ord.data(), h->d.ordC + D.contactBase, n * sizeof(int), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(isl.data(), h->d.cisl + D.contactBase, n * sizeof(int), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(bat.data(), h->d.cbatch + D.contactBase, n * sizeof(int), hipMemcpyDeviceToHost));
    DBX_HIP(h, hipMemcpy(fb.data(), h->d.fbody + D.fixBase, D.fixCount * sizeof(int), hipMemcpyDeviceToHost));
    if (S.contactCount)
        DBX_HIP(h, hipMemcpy(c.data(), h->d.cont[h->cur] + D.contactBase, S.contactCount * sizeof(Contact),
                             hipMemcpyDeviceToHost));
    for (int i = 0; i < n && i < cap; ++i) {
        const Contact &k = c[ord[i]];
        out[4 * i] = isl[i];
        out[4 * i + 1] = fb[k.fa];
        out[4 * i + 2] = fb[k.fb];
        out[4 * i + 3] = bat[i];
    }
    return DBX_OK;
}

dbx_status dbx_broadphase_pairs(dbx_handle *h, const float *fat_aabbs, const uint8_t *moved, int32_t n, int32_t *out,
                                int32_t cap, int32_t *count) {
    if (!h || n < 0 || (n > 0 && (!fat_aabbs || !moved)) || cap < 0) return DBX_E_INVALID_ARG;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    std::vector<int> mv(n);
    for (int i = 0; i < n; ++i) mv[i] = moved[i] ? 1 : 0;
    float4 *dfat = nullptr;
    int *dmv = nullptr, *dtot = nullptr;
    char *dcnt = nullptr;
    int2 *dout = nullptr;
    size_t m = (size_t)std::max(1, n);
    hipError_t e = hipMalloc(&dfat, m * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&dmv, m * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dcnt, broadphase_scratch_bytes(std::max(1, n)));
    if (e == hipSuccess) e = hipMalloc(&dtot, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dout, (size_t)std::max(1, cap) * sizeof(int2));
    if (e == hipSuccess && n) e = hipMemcpy(dfat, fat_aabbs, n * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(dmv, mv.data(), n * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_bp_single(dfat, dmv, n, dcnt, dout, cap, dtot, h->stream);
    int total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, dtot, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && total <= cap && total > 0 && out)
        e = hipMemcpy(out, dout, total * sizeof(int2), hipMemcpyDeviceToHost);
    (void)hipFree(dfat);
    (void)hipFree(dmv);
    (void)hipFree(dcnt);
    (void)hipFree(dtot);
    (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(h, e, "broadphase");
    if (count) *count = total;
    return total > cap ? fail(h, DBX_E_CAPACITY, "pair output too small") : DBX_OK;
}

dbx_status dbx_broadphase_pairs_noalloc(dbx_handle *h, const float *fat_aabbs, const int32_t *moved, int32_t n,
                                        int32_t *pairs, int32_t cap, int32_t *count) {
    if (!h) return DBX_E_INVALID_ARG;
    if (n < 0 || cap < 0 || !count || (n > 0 && (!fat_aabbs || !moved)) || (cap > 0 && !pairs))
        return fail(h, DBX_E_INVALID_ARG, "bad broadphase arguments");
    if (((uintptr_t)fat_aabbs & 15) || ((uintptr_t)pairs & 7) || ((uintptr_t)moved & 3) || ((uintptr_t)count & 3))
        return fail(h, DBX_E_INVALID_ARG, "misaligned device pointer (fat AABBs need 16 bytes, pairs 8, ints 4)");
    if (n > DBX_NOALLOC_MAX_PROXIES || broadphase_scratch_bytes(n) > kNoAllocScratchBytes)
        return fail(h, DBX_E_CAPACITY, "more proxies than the fixed scratch holds (DBX_NOALLOC_MAX_PROXIES)");
    if (!h->noallocScratch) return fail(h, DBX_E_HIP, "no device address for the fixed scratch on this GPU");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    const hipError_t e = launch_bp_single(reinterpret_cast<const float4 *>(fat_aabbs), moved, n,
                                          static_cast<char *>(h->noallocScratch), reinterpret_cast<int2 *>(pairs), cap,
                                          count, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "broadphase");
    return DBX_OK;
}

dbx_status dbx_collide_pairs(dbx_handle *h, const dbx_fixture_def *a, const float *xa, const dbx_fixture_def *b,
                             const float *xb, int32_t n, float *out) {
    if (!h || n < 0 || (n > 0 && (!a || !b || !xa || !xb || !out))) return DBX_E_INVALID_ARG;
    if (n == 0) return DBX_OK;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    std::vector<Shape> sa(n), sb(n);
    std::string err;
    for (int i = 0; i < n; ++i) {
        if (!build_shape(a[i], sa[i], err) || !build_shape(b[i], sb[i], err)) return fail(h, DBX_E_INVALID_ARG, err);
    }
    Shape *dsa = nullptr, *dsb = nullptr;
    float *dxa = nullptr, *dxb = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&dsa, n * sizeof(Shape));
    if (e == hipSuccess) e = hipMalloc(&dsb, n * sizeof(Shape));
    if (e == hipSuccess) e = hipMalloc(&dxa, n * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dxb, n * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dout, n * 32 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dsa, sa.data(), n * sizeof(Shape), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dsb, sb.data(), n * sizeof(Shape), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dxa, xa, n * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dxb, xb, n * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_collide_pairs(dsa, dxa, dsb, dxb, n, dout, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * 32 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dsa);
    (void)hipFree(dsb);
    (void)hipFree(dxa);
    (void)hipFree(dxb);
    (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(h, e, "collide");
    return DBX_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Residual mode and the learned warm start.
// ---------------------------------------------------------------------------
dbx_status dbx_set_residual(dbx_handle *h, float tolerance, int32_t max_iterations) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!std::isfinite(tolerance) || max_iterations < 1 || max_iterations > 100000)
        return fail(h, DBX_E_INVALID_ARG, "bad residual settings");
    h->residualTol = tolerance;
    h->maxIterations = max_iterations;
    return DBX_OK;
}

static uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

dbx_status dbx_mlp_init(dbx_handle *h, uint64_t seed) {
    if (!h) return DBX_E_INVALID_ARG;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    std::vector<float> P(kMlpParams, 0.0f), N(kMlpNorm, 0.0f);
    uint64_t st = seed;
    auto uni = [&](float lim) { return (float)(((splitmix(st) >> 11) * 0x1.0p-53) * 2.0 - 1.0) * lim; };
    const float l1 = std::sqrt(6.0f / kMlpIn), l2 = std::sqrt(6.0f / kMlpHidden), l3 = std::sqrt(1.0f / kMlpHidden);
    for (int i = 0; i < kMlpHidden * kMlpIn; ++i) P[kOffW1 + i] = uni(l1);
    for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) P[kOffW2 + i] = uni(l2);
    for (int i = 0; i < kMlpOut * kMlpHidden; ++i) P[kOffW3 + i] = uni(l3);
    for (int i = 0; i < kMlpIn; ++i) N[kNormInScale + i] = 1.0f;
    N[kNormOutScale] = N[kNormOutScale + 1] = 1.0f;
    DBX_HIP(h, mlp_upload(h->mlp, P.data(), N.data(), h->stream));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    return DBX_OK;
}

dbx_status dbx_mlp_set(dbx_handle *h, const float *params, const float *norm) {
    if (!h || (!params && !norm)) return DBX_E_INVALID_ARG;
    if (!h->mlp.ready && (!params || !norm)) return fail(h, DBX_E_STATE, "the first network needs parameters and normalisation");
    if (params)
        for (int i = 0; i < kMlpParams; ++i)
            if (!std::isfinite(params[i])) return fail(h, DBX_E_INVALID_ARG, "parameter is not finite");
    if (norm)
        for (int i = 0; i < kMlpNorm; ++i)
            if (!std::isfinite(norm[i])) return fail(h, DBX_E_INVALID_ARG, "normalisation is not finite");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, mlp_upload(h->mlp, params, norm, h->stream));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    return DBX_OK;
}

dbx_status dbx_mlp_get(dbx_handle *h, float *params, float *norm) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->mlp.ready) return fail(h, DBX_E_STATE, "no warm-start network loaded");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    if (params) DBX_HIP(h, hipMemcpy(params, h->mlp.params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) DBX_HIP(h, hipMemcpy(norm, h->mlp.norm, kMlpNorm * sizeof(float), hipMemcpyDeviceToHost));
    return DBX_OK;
}

// Device copy of n feature rows of DBX_MLP_FEATURES floats, padded to kMlpFeatStride.
static dbx_status upload_rows(dbx_handle *h, const float *x, int32_t n, float **dX) {
    std::vector<float> pad((size_t)n * kMlpFeatStride, 0.0f);
    for (int32_t i = 0; i < n; ++i)
        for (int j = 0; j < kMlpIn; ++j) pad[(size_t)i * kMlpFeatStride + j] = x[(size_t)i * kMlpIn + j];
    DBX_HIP(h, hipMalloc((void **)dX, pad.size() * sizeof(float)));
    hipError_t e = hipMemcpy(*dX, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*dX);
        return hip_fail(h, e, "hipMemcpy");
    }
    return DBX_OK;
}

dbx_status dbx_mlp_forward(dbx_handle *h, const float *x, int32_t n, float *y) {
    if (!h || n < 0 || (n > 0 && (!x || !y))) return DBX_E_INVALID_ARG;
    if (!h->mlp.ready) return fail(h, DBX_E_STATE, "no warm-start network loaded");
    if (n == 0) return DBX_OK;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    float *dX = nullptr, *dY = nullptr;
    dbx_status st = upload_rows(h, x, n, &dX);
    if (st != DBX_OK) return st;
    hipError_t e = hipMalloc((void **)&dY, 2 * (size_t)n * sizeof(float));
    if (e == hipSuccess) e = mlp_forward(h->mlp, dX, n, dY, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(y, dY, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    if (e != hipSuccess) return hip_fail(h, e, "mlp forward");
    return DBX_OK;
}

dbx_status dbx_mlp_gradients(dbx_handle *h, const float *x, const float *t, int32_t n, float *grad, float *loss) {
    if (!h || n <= 0 || !x || !t || !grad) return DBX_E_INVALID_ARG;
    if (!h->mlp.ready) return fail(h, DBX_E_STATE, "no warm-start network loaded");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    float *dX = nullptr, *dT = nullptr;
    dbx_status st = upload_rows(h, x, n, &dX);
    if (st != DBX_OK) return st;
    hipError_t e = hipMalloc((void **)&dT, 2 * (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dT, t, 2 * (size_t)n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_clear_grad(h->mlp, h->stream);
    if (e == hipSuccess) e = mlp_gradients(h->mlp, dX, dT, n, n, 0, nullptr, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(grad, h->mlp.grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, h->mlp.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = mlp_clear_grad(h->mlp, h->stream);
    (void)hipFree(dX);
    if (dT) (void)hipFree(dT);
    if (e != hipSuccess) return hip_fail(h, e, "mlp gradients");
    return DBX_OK;
}

dbx_status dbx_mlp_log_enable(dbx_handle *h, int64_t capacity) {
    if (!h || capacity <= 0 || capacity > INT32_MAX) return DBX_E_INVALID_ARG;
    if (h->logX) return fail(h, DBX_E_STATE, "the training log is already enabled");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    DBX_HIP(h, dalloc(h, &h->logX, capacity * kMlpFeatStride));
    DBX_HIP(h, dalloc(h, &h->logT, 2 * capacity));
    DBX_HIP(h, dalloc(h, &h->logCount, 1));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    h->logCap = capacity;
    return DBX_OK;
}

static dbx_status log_rows(dbx_handle *h, long long *count) {
    if (!h->logX) return fail(h, DBX_E_STATE, "the training log is not enabled");
    unsigned long long c = 0;
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    DBX_HIP(h, hipMemcpy(&c, h->logCount, sizeof c, hipMemcpyDeviceToHost));
    *count = (long long)std::min<unsigned long long>(c, (unsigned long long)h->logCap);
    return DBX_OK;
}

dbx_status dbx_mlp_log_count(dbx_handle *h, int64_t *count) {
    if (!h || !count) return DBX_E_INVALID_ARG;
    long long c = 0;
    dbx_status st = log_rows(h, &c);
    if (st != DBX_OK) return st;
    *count = c;
    return DBX_OK;
}

dbx_status dbx_mlp_log_read(dbx_handle *h, int64_t first, int64_t n, float *x, float *t) {
    if (!h || first < 0 || n < 0) return DBX_E_INVALID_ARG;
    long long c = 0;
    dbx_status st = log_rows(h, &c);
    if (st != DBX_OK) return st;
    if (first + n > c) return fail(h, DBX_E_INVALID_ARG, "log rows out of range");
    if (n == 0) return DBX_OK;
    std::vector<float> rows((size_t)n * kMlpFeatStride);
    DBX_HIP(h, hipMemcpy(rows.data(), h->logX + first * kMlpFeatStride, rows.size() * sizeof(float),
                         hipMemcpyDeviceToHost));
    if (x)
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < kMlpIn; ++j) x[i * kMlpIn + j] = rows[(size_t)i * kMlpFeatStride + j];
    if (t) DBX_HIP(h, hipMemcpy(t, h->logT + 2 * first, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return DBX_OK;
}

dbx_status dbx_mlp_log_clear(dbx_handle *h) {
    if (!h) return DBX_E_INVALID_ARG;
    if (!h->logX) return fail(h, DBX_E_STATE, "the training log is not enabled");
    DBX_HIP(h, hipMemsetAsync(h->logCount, 0, sizeof(unsigned long long), h->stream));
    DBX_HIP(h, hipStreamSynchronize(h->stream));
    return DBX_OK;
}

dbx_status dbx_mlp_fit_normalization(dbx_handle *h) {
    if (!h) return DBX_E_INVALID_ARG;
    long long c = 0;
    dbx_status st = log_rows(h, &c);
    if (st != DBX_OK) return st;
    if (c < 2) return fail(h, DBX_E_STATE, "the training log has fewer than two rows");
    std::vector<float> x((size_t)c * kMlpIn), t(2 * (size_t)c);
    st = dbx_mlp_log_read(h, 0, c, x.data(), t.data());
    if (st != DBX_OK) return st;
    std::vector<float> N(kMlpNorm, 0.0f);
    auto stats = [&](const float *v, int stride, int col, float &mean, float &sd) {
        double s = 0.0, s2 = 0.0;
        for (long long i = 0; i < c; ++i) s += v[i * stride + col];
        const double m = s / (double)c;
        for (long long i = 0; i < c; ++i) {
            const double dv = v[i * stride + col] - m;
            s2 += dv * dv;
        }
        mean = (float)m;
        sd = (float)std::sqrt(s2 / (double)c);
        if (!(sd > 1e-6f)) sd = 1.0f; // a constant feature carries no information
    };
    for (int j = 0; j < kMlpIn; ++j) {
        float m, sd;
        stats(x.data(), kMlpIn, j, m, sd);
        N[kNormInMean + j] = m;
        N[kNormInScale + j] = 1.0f / sd;
    }
    for (int o = 0; o < kMlpOut; ++o) {
        float m, sd;
        stats(t.data(), 2, o, m, sd);
        N[kNormOutMean + o] = m;
        N[kNormOutScale + o] = sd;
    }
    if (!h->mlp.ready) {
        st = dbx_mlp_init(h, 0x5EED);
        if (st != DBX_OK) return st;
    }
    return dbx_mlp_set(h, nullptr, N.data());
}

dbx_status dbx_mlp_train(dbx_handle *h, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss) {
    if (!h || steps < 0 || batch <= 0 || !(lr > 0.0f)) return DBX_E_INVALID_ARG;
    if (!h->mlp.ready) return fail(h, DBX_E_STATE, "no warm-start network loaded");
    long long c = 0;
    dbx_status st = log_rows(h, &c);
    if (st != DBX_OK) return st;
    if (c < 1) return fail(h, DBX_E_STATE, "the training log is empty");
    if (steps == 0) return DBX_OK;
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    float *dLoss = nullptr;
    DBX_HIP(h, hipMalloc((void **)&dLoss, (size_t)steps * sizeof(float)));
    hipError_t e = mlp_fill(dLoss, steps, 0.0f, h->stream);
    for (int32_t i = 0; i < steps && e == hipSuccess; ++i) {
        const uint64_t sd = (seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(i + 1))) | 1ull; // non-zero: sampled rows
        e = mlp_gradients(h->mlp, h->logX, h->logT, c, batch, sd, dLoss + i, h->stream);
        if (e == hipSuccess) e = mlp_adam_step(h->mlp, lr, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, dLoss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dLoss);
    if (e != hipSuccess) return hip_fail(h, e, "mlp training");
    return DBX_OK;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner.
// ---------------------------------------------------------------------------
dbx_status dbx_autotune(dbx_handle *h, int32_t steps_per_candidate) {
    if (!h || steps_per_candidate < 1 || steps_per_candidate > 1000) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    DBX_HIP(h, hipSetDevice(h->cfg.device));
    for (Tier &T : h->tiers) {
        if (!T.t0) DBX_HIP(h, hipEventCreate(&T.t0));
        if (!T.t1) DBX_HIP(h, hipEventCreate(&T.t1));
    }
    h->tunePer = steps_per_candidate;
    h->tuneSteps = steps_per_candidate * kTuneCandidates;
    h->tuneDone = 0;
    h->tuneMs.assign(h->tiers.size() * kTuneCandidates, 0.0);
    return DBX_OK;
}

dbx_status dbx_set_tier_blocks(dbx_handle *h, const int32_t *blocks, int32_t n) {
    if (!h || !blocks) return DBX_E_INVALID_ARG;
    if (!h->finalized) return fail(h, DBX_E_STATE, "dbx_finalize has not been called");
    if (n != (int32_t)h->tiers.size()) return fail(h, DBX_E_INVALID_ARG, "one size per tier is required");
    for (int32_t i = 0; i < n; ++i)
        if (blocks[i] < 64 || blocks[i] > kMaxBlock || (blocks[i] & 63))
            return fail(h, DBX_E_INVALID_ARG, "workgroup size must be a multiple of 64 in [64, 1024]");
    for (int32_t i = 0; i < n; ++i) h->tiers[i].block = blocks[i];
    return DBX_OK;
}

dbx_status dbx_set_graph(dbx_handle *h, int32_t enabled) {
    if (!h) return DBX_E_INVALID_ARG;
    h->cfg.use_graph = enabled ? 1 : 0;
    return DBX_OK;
}
