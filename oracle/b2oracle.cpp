/*
 * b2oracle.cpp -- CPU oracle for the b2World.Step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see b2oracle.h).  A scalar restatement of dbox
 * (D port of Box2D 2.3.1) that keeps the reference's float operation order.
 * Built with -O3 -ffp-contract=off -fno-fast-math against glibc libm, which is
 * what dbox reaches through core.stdc.math.  Citations are relative to the
 * reference's src/dbox/.
 *
 * Parity pins (what checks each part; see tests/test_oracle.py):
 *   math / shapes / mass       closed forms (box, circle, regular polygon mass and inertia)
 *   collide functions          textbook special cases (axis-aligned box stacks, circle gaps,
 *                              face/vertex regions) + A/B swap symmetry
 *   dynamic tree + UpdatePairs brute-force fat-AABB overlap (exact)
 *   solver / step              hello_world scene at rest at y = 1.01 m after 60 steps,
 *                              free-fall closed form, resting box impulse m*g*dt,
 *                              momentum conservation, sleep after 0.5 s
 *   multi-step stacking order  parity unpinned against the dbox binary (no D compiler):
 *                              pinned only to this line-by-line restatement.
 *   b2Distance                 brute-force polygon/segment/point distance, circle closed form
 *                              (tests/test_oracle_toi.py)
 *   b2TimeOfImpact             closed-form impact time of a translating box over an edge;
 *                              brute-force sampling of the swept distance (no earlier
 *                              crossing of the target, separation at the returned time)
 *   SolveTOI                   a fast box over a thin edge stops on it instead of tunnelling and
 *                              rests at the height the target separation fixes; the sub-step
 *                              island order and the b2_toiBaugarte factor are pinned only to
 *                              this restatement (parity unpinned).
 */
#include "b2oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

/* ---- b2settings.d:61-149 (float-typed constants evaluated in float) ---- */
constexpr float kPi = 3.14159265359f;
constexpr float kEps = FLT_EPSILON;
constexpr float kMaxFloat = FLT_MAX;
constexpr int kMaxManifoldPoints = 2;
constexpr int kMaxPolygonVertices = 8;
constexpr float kAabbExtension = 0.1f;
constexpr float kAabbMultiplier = 2.0f;
constexpr float kLinearSlop = 0.005f;
constexpr float kAngularSlop = (2.0f / 180.0f * kPi);
constexpr float kPolygonRadius = (2.0f * kLinearSlop);
constexpr float kVelocityThreshold = 1.0f;
constexpr float kMaxLinearCorrection = 0.2f;
constexpr float kMaxTranslation = 2.0f;
constexpr float kMaxTranslationSquared = (kMaxTranslation * kMaxTranslation);
constexpr float kMaxRotation = (0.5f * kPi);
constexpr float kMaxRotationSquared = (kMaxRotation * kMaxRotation);
constexpr float kBaumgarte = 0.2f;
constexpr float kTimeToSleep = 0.5f;
constexpr float kLinearSleepTolerance = 0.01f;
constexpr float kAngularSleepTolerance = (2.0f / 180.0f * kPi);

/* ---- b2math.d:60-766 ---- */
struct V2 {
    float x = 0, y = 0;
};
static inline V2 mk(float x, float y) { V2 v; v.x = x; v.y = y; return v; }
static inline V2 operator+(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }
static inline V2 operator-(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }
static inline V2 operator-(V2 a) { return mk(-a.x, -a.y); }
/* b2Vec2.opBinary(float) computes (s op x): b2math.d:113-117 */
static inline V2 operator*(float s, V2 a) { return mk(s * a.x, s * a.y); }
static inline void operator+=(V2 &a, V2 b) { a.x += b.x; a.y += b.y; }
static inline void operator-=(V2 &a, V2 b) { a.x -= b.x; a.y -= b.y; }
static inline void operator*=(V2 &a, float s) { a.x *= s; a.y *= s; }
static inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
static inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
static inline V2 cross(V2 a, float s) { return mk(s * a.y, -s * a.x); }
static inline V2 cross(float s, V2 a) { return mk(-s * a.y, s * a.x); }
static inline float length(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static inline float lengthSq(V2 a) { return a.x * a.x + a.y * a.y; }
static inline float distSq(V2 a, V2 b) { V2 c = a - b; return dot(c, c); }
/* b2Vec2.Normalize b2math.d:162-175 */
static inline float normalize(V2 &v) {
    float len = length(v);
    if (len < kEps) return 0.0f;
    float inv = 1.0f / len;
    v.x *= inv;
    v.y *= inv;
    return len;
}
static inline float fmin_(float a, float b) { return a < b ? a : b; }
static inline float fmax_(float a, float b) { return a > b ? a : b; }
static inline V2 vmin(V2 a, V2 b) { return mk(fmin_(a.x, b.x), fmin_(a.y, b.y)); }
static inline V2 vmax(V2 a, V2 b) { return mk(fmax_(a.x, b.x), fmax_(a.y, b.y)); }
static inline float fabs_(float a) { return a > 0.0f ? a : -a; }
static inline float clampf(float a, float lo, float hi) { return fmax_(lo, fmin_(a, hi)); }

/* Test switch (not reference behaviour): 0 = b2Rot.Set with glibc sinf/cosf as dbox
 * calls them; 1 = sin and cos evaluated in double and rounded once to float, by the
 * textbook method below (Cody-Waite reduction by pi/2 split into three doubles, Taylor
 * series of sin r and cos r up to r^15 / r^16 in Horner form with fused multiply-adds).
 * Mode 1 is the same evaluation, operation for operation, as the device library's, so
 * GPU results can be compared with this restatement bit for bit. */
static int g_rotMode = 0;
static void rotExact(float angle, float &s, float &c) {
    const double x = (double)angle;
    if (!(std::fabs(x) < 1.0e6)) {
        s = (float)std::sin(x);
        c = (float)std::cos(x);
        return;
    }
    const double P1 = 0x1.921fb54400000p+0, P2 = 0x1.0b4611a600000p-34, P3 = 0x1.3198a2e037073p-69;
    const double k = std::rint(x * 0.6366197723675814);
    double r = std::fma(-k, P1, x);
    r = std::fma(-k, P2, r);
    r = std::fma(-k, P3, r);
    const double z = r * r;
    /* 1/n! with alternating signs: sin r = r + r z (s6 + z (s5 + ...)), cos r = 1 + z (-1/2 + z (c6 + ...)) */
    static const double S[7] = {-7.647163731819816e-13, 1.6059043836821613e-10, -2.505210838544172e-08,
                                2.7557319223985893e-06, -0.0001984126984126984, 0.008333333333333333,
                                -0.16666666666666666};
    static const double Cc[7] = {4.779477332387385e-14, -1.1470745597729725e-11, 2.08767569878681e-09,
                                 -2.755731922398589e-07, 2.48015873015873e-05, -0.001388888888888889,
                                 0.041666666666666664};
    double ps = std::fma(S[0], z, S[1]);
    double pc = std::fma(Cc[0], z, Cc[1]);
    for (int i = 2; i < 7; ++i) {
        ps = std::fma(ps, z, S[i]);
        pc = std::fma(pc, z, Cc[i]);
    }
    pc = std::fma(pc, z, -0.5);
    pc = std::fma(pc, z, 1.0);
    ps = std::fma(ps * z, r, r);
    const unsigned q = (unsigned)(int)k;
    const float fs = (float)ps, fc = (float)pc;
    const bool odd = (q & 1u) != 0u;
    float so = odd ? fc : fs, co = odd ? fs : fc;
    if (q & 2u) so = -so;
    if ((q + 1u) & 2u) co = -co;
    s = so;
    c = co;
}

struct Rot {
    float s = 0, c = 0;
    void set(float angle) { /* b2math.d:483-488 */
        if (g_rotMode) rotExact(angle, s, c);
        else { s = sinf(angle); c = cosf(angle); }
    }
};
struct Xf {
    V2 p;
    Rot q;
};
struct Mat22 {
    V2 ex, ey;
    /* b2Mat22.GetInverse b2math.d:322-337 */
    Mat22 inverse() const {
        float a = ex.x, b = ey.x, c = ex.y, d = ey.y;
        Mat22 B;
        float det = a * d - b * c;
        if (det != 0.0f) det = 1.0f / det;
        B.ex.x = det * d;
        B.ey.x = -det * b;
        B.ex.y = -det * c;
        B.ey.y = det * a;
        return B;
    }
};
static inline V2 mul(const Mat22 &A, V2 v) { return mk(A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y); }
/* b2Mat22.Solve b2math.d:341-352 */
static inline V2 solve22m(const Mat22 &A, V2 b) {
    float a11 = A.ex.x, a12 = A.ey.x, a21 = A.ex.y, a22 = A.ey.y;
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return mk(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}
static inline V2 mul(Rot q, V2 v) { return mk(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
static inline V2 mulT(Rot q, V2 v) { return mk(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
static inline V2 mul(const Xf &T, V2 v) {
    float x = (T.q.c * v.x - T.q.s * v.y) + T.p.x;
    float y = (T.q.s * v.x + T.q.c * v.y) + T.p.y;
    return mk(x, y);
}
static inline V2 mulT(const Xf &T, V2 v) {
    float px = v.x - T.p.x;
    float py = v.y - T.p.y;
    float x = (T.q.c * px + T.q.s * py);
    float y = (-T.q.s * px + T.q.c * py);
    return mk(x, y);
}
static inline Rot mulT(Rot q, Rot r) {
    Rot qr;
    qr.s = q.c * r.s - q.s * r.c;
    qr.c = q.c * r.c + q.s * r.s;
    return qr;
}
static inline Xf mulT(const Xf &A, const Xf &B) {
    Xf C;
    C.q = mulT(A.q, B.q);
    C.p = mulT(A.q, B.p - A.p);
    return C;
}

/* ---- b2collision.d ---- */
struct AABB {
    V2 lo, hi;
    float perimeter() const {
        float wx = hi.x - lo.x;
        float wy = hi.y - lo.y;
        return 2.0f * (wx + wy);
    }
    void combine(const AABB &a, const AABB &b) { lo = vmin(a.lo, b.lo); hi = vmax(a.hi, b.hi); }
    bool contains(const AABB &a) const {
        bool r = true;
        r = r && lo.x <= a.lo.x;
        r = r && lo.y <= a.lo.y;
        r = r && a.hi.x <= hi.x;
        r = r && a.hi.y <= hi.y;
        return r;
    }
};
/* b2TestOverlap(b2AABB, b2AABB) b2collision.d:470-483 */
static inline bool testOverlap(const AABB &a, const AABB &b) {
    V2 d1 = b.lo - a.hi;
    V2 d2 = a.lo - b.hi;
    if (d1.x > 0.0f || d1.y > 0.0f) return false;
    if (d2.x > 0.0f || d2.y > 0.0f) return false;
    return true;
}

enum { e_vertex = 0, e_face = 1 };
/* b2ContactID: 4 uint8 packed little-endian into a uint32 key (b2collision.d:38-60) */
static inline uint32_t makeKey(int indexA, int indexB, int typeA, int typeB) {
    return (uint32_t)(uint8_t)indexA | ((uint32_t)(uint8_t)indexB << 8) | ((uint32_t)(uint8_t)typeA << 16) |
           ((uint32_t)(uint8_t)typeB << 24);
}
static inline int keyIndexA(uint32_t k) { return k & 0xff; }
static inline int keyIndexB(uint32_t k) { return (k >> 8) & 0xff; }
static inline int keyTypeA(uint32_t k) { return (k >> 16) & 0xff; }
static inline int keyTypeB(uint32_t k) { return (k >> 24) & 0xff; }

struct ManifoldPoint {
    V2 localPoint;
    float normalImpulse = 0, tangentImpulse = 0;
    uint32_t id = 0;
};
enum { m_circles = 0, m_faceA = 1, m_faceB = 2 };
struct Manifold {
    ManifoldPoint points[kMaxManifoldPoints];
    V2 localNormal, localPoint;
    int type = 0;
    int pointCount = 0;
};
struct WorldManifold {
    V2 normal;
    V2 points[2];
    float separations[2] = {0, 0};
    /* b2WorldManifold.Initialize b2collision.d:123-191 */
    void initialize(const Manifold &m, const Xf &xfA, float radiusA, const Xf &xfB, float radiusB) {
        if (m.pointCount == 0) return;
        switch (m.type) {
        case m_circles: {
            normal = mk(1.0f, 0.0f);
            V2 pointA = mul(xfA, m.localPoint);
            V2 pointB = mul(xfB, m.points[0].localPoint);
            if (distSq(pointA, pointB) > kEps * kEps) {
                normal = pointB - pointA;
                normalize(normal);
            }
            V2 cA = pointA + radiusA * normal;
            V2 cB = pointB - radiusB * normal;
            points[0] = 0.5f * (cA + cB);
            separations[0] = dot(cB - cA, normal);
        } break;
        case m_faceA: {
            normal = mul(xfA.q, m.localNormal);
            V2 planePoint = mul(xfA, m.localPoint);
            for (int i = 0; i < m.pointCount; ++i) {
                V2 clipPoint = mul(xfB, m.points[i].localPoint);
                V2 cA = clipPoint + (radiusA - dot(clipPoint - planePoint, normal)) * normal;
                V2 cB = clipPoint - radiusB * normal;
                points[i] = 0.5f * (cA + cB);
                separations[i] = dot(cB - cA, normal);
            }
        } break;
        case m_faceB: {
            normal = mul(xfB.q, m.localNormal);
            V2 planePoint = mul(xfB, m.localPoint);
            for (int i = 0; i < m.pointCount; ++i) {
                V2 clipPoint = mul(xfA, m.points[i].localPoint);
                V2 cB = clipPoint + (radiusB - dot(clipPoint - planePoint, normal)) * normal;
                V2 cA = clipPoint - radiusA * normal;
                points[i] = 0.5f * (cA + cB);
                separations[i] = dot(cA - cB, normal);
            }
            normal = -normal;
        } break;
        default:
            break;
        }
    }
};

struct ClipVertex {
    V2 v;
    uint32_t id = 0;
};
/* b2ClipSegmentToLine b2collision.d:413-446 */
static int clipSegmentToLine(ClipVertex vOut[2], const ClipVertex vIn[2], V2 normal, float offset, int vertexIndexA) {
    int numOut = 0;
    float distance0 = dot(normal, vIn[0].v) - offset;
    float distance1 = dot(normal, vIn[1].v) - offset;
    if (distance0 <= 0.0f) vOut[numOut++] = vIn[0];
    if (distance1 <= 0.0f) vOut[numOut++] = vIn[1];
    if (distance0 * distance1 < 0.0f) {
        float interp = distance0 / (distance0 - distance1);
        vOut[numOut].v = vIn[0].v + interp * (vIn[1].v - vIn[0].v);
        vOut[numOut].id = makeKey(vertexIndexA, keyIndexB(vIn[0].id), e_vertex, e_face);
        ++numOut;
    }
    return numOut;
}

/* ---- shapes (src/dbox/collision/shapes) ---- */
enum { s_circle = 0, s_edge = 1, s_polygon = 2, s_chain = 3 };
struct Shape {
    int type = s_circle;
    float radius = 0;
    V2 p;                        /* circle centre */
    V2 centroid;                 /* polygon */
    V2 v[kMaxPolygonVertices];   /* polygon vertices */
    V2 n[kMaxPolygonVertices];   /* polygon normals */
    int count = 0;
    V2 v0, v1, v2, v3;           /* edge */
    bool has0 = false, has3 = false;
    V2 cv[kMaxPolygonVertices + 1]; /* chain: m_vertices (a loop repeats its first vertex) */
    int ccount = 0;              /* chain: m_count */
    V2 cprev, cnext;             /* chain: m_prevVertex, m_nextVertex */
    bool chasPrev = false, chasNext = false;
};
/* b2ChainShape.GetChildCount / GetChildEdge (b2chainshape.d:129-170) */
static int childCount(const Shape &s) { return s.type == s_chain ? s.ccount - 1 : 1; }
static void childEdge(const Shape &chain, int index, Shape &edge) {
    edge = Shape();
    edge.type = s_edge;
    edge.radius = chain.radius;
    edge.v1 = chain.cv[index + 0];
    edge.v2 = chain.cv[index + 1];
    if (index > 0) {
        edge.v0 = chain.cv[index - 1];
        edge.has0 = true;
    } else {
        edge.v0 = chain.cprev;
        edge.has0 = chain.chasPrev;
    }
    if (index < chain.ccount - 2) {
        edge.v3 = chain.cv[index + 2];
        edge.has3 = true;
    } else {
        edge.v3 = chain.cnext;
        edge.has3 = chain.chasNext;
    }
}
struct MassData {
    float mass = 0;
    V2 center;
    float I = 0;
};

/* b2PolygonShape.SetAsBox b2polygonshape.d:220-232 */
static void polySetAsBox(Shape &s, float hx, float hy) {
    s.type = s_polygon;
    s.radius = kPolygonRadius;
    s.count = 4;
    s.v[0] = mk(-hx, -hy);
    s.v[1] = mk(hx, -hy);
    s.v[2] = mk(hx, hy);
    s.v[3] = mk(-hx, hy);
    s.n[0] = mk(0.0f, -1.0f);
    s.n[1] = mk(1.0f, 0.0f);
    s.n[2] = mk(0.0f, 1.0f);
    s.n[3] = mk(-1.0f, 0.0f);
    s.centroid = mk(0.0f, 0.0f);
}
/* b2PolygonShape.SetAsBox(hx, hy, center, angle) b2polygonshape.d:235-262 */
static void polySetAsBoxOriented(Shape &s, float hx, float hy, V2 center, float angle) {
    polySetAsBox(s, hx, hy);
    s.centroid = center;
    Xf xf;
    xf.p = center;
    xf.q.set(angle);
    for (int i = 0; i < s.count; ++i) {
        s.v[i] = mul(xf, s.v[i]);
        s.n[i] = mul(xf.q, s.n[i]);
    }
}
/* b2PolygonShape.ComputeCentroid b2polygonshape.d:505-549 (reference point at origin) */
static V2 computeCentroid(const V2 *vs, int count) {
    V2 c = mk(0.0f, 0.0f);
    float area = 0.0f;
    V2 pRef = mk(0.0f, 0.0f);
    const float inv3 = 1.0f / 3.0f;
    for (int i = 0; i < count; ++i) {
        V2 p1 = pRef;
        V2 p2 = vs[i];
        V2 p3 = i + 1 < count ? vs[i + 1] : vs[0];
        V2 e1 = p2 - p1;
        V2 e2 = p3 - p1;
        float D = cross(e1, e2);
        float triangleArea = 0.5f * D;
        area += triangleArea;
        c += (triangleArea * inv3) * (p1 + p2 + p3);
    }
    c *= 1.0f / area;
    return c;
}
/* b2PolygonShape.Set: weld + gift wrapping hull b2polygonshape.d:74-209 */
static bool polySet(Shape &s, const V2 *vertices, int count) {
    s.type = s_polygon;
    s.radius = kPolygonRadius;
    if (count < 3) { polySetAsBox(s, 1.0f, 1.0f); return false; }
    int n = count < kMaxPolygonVertices ? count : kMaxPolygonVertices;
    V2 ps[kMaxPolygonVertices];
    int tempCount = 0;
    for (int i = 0; i < n; ++i) {
        V2 v = vertices[i];
        bool unique = true;
        for (int j = 0; j < tempCount; ++j) {
            if (distSq(v, ps[j]) < 0.5f * kLinearSlop) { unique = false; break; }
        }
        if (unique) ps[tempCount++] = v;
    }
    n = tempCount;
    if (n < 3) { polySetAsBox(s, 1.0f, 1.0f); return false; }
    int i0 = 0;
    float x0 = ps[0].x;
    for (int i = 1; i < n; ++i) {
        float x = ps[i].x;
        if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) { i0 = i; x0 = x; }
    }
    int hull[kMaxPolygonVertices];
    int m = 0;
    int ih = i0;
    for (;;) {
        hull[m] = ih;
        int ie = 0;
        for (int j = 1; j < n; ++j) {
            if (ie == ih) { ie = j; continue; }
            V2 r = ps[ie] - ps[hull[m]];
            V2 v = ps[j] - ps[hull[m]];
            float c = cross(r, v);
            if (c < 0.0f) ie = j;
            if (c == 0.0f && lengthSq(v) > lengthSq(r)) ie = j;
        }
        ++m;
        ih = ie;
        if (ie == i0) break;
        if (m >= kMaxPolygonVertices) break; /* a convex input never gets here */
    }
    if (m < 3) { polySetAsBox(s, 1.0f, 1.0f); return false; }
    s.count = m;
    for (int i = 0; i < m; ++i) s.v[i] = ps[hull[i]];
    for (int i = 0; i < m; ++i) {
        int i1 = i;
        int i2 = i + 1 < m ? i + 1 : 0;
        V2 edge = s.v[i2] - s.v[i1];
        s.n[i] = cross(edge, 1.0f);
        normalize(s.n[i]);
    }
    s.centroid = computeCentroid(s.v, m);
    return true;
}
/* b2PolygonShape.ComputeMass b2polygonshape.d:376-459 */
static void polyMass(const Shape &sh, MassData &md, float density) {
    V2 center = mk(0.0f, 0.0f);
    float area = 0.0f;
    float I = 0.0f;
    V2 s = mk(0.0f, 0.0f);
    for (int i = 0; i < sh.count; ++i) s += sh.v[i];
    s *= 1.0f / (float)sh.count;
    const float k_inv3 = 1.0f / 3.0f;
    for (int i = 0; i < sh.count; ++i) {
        V2 e1 = sh.v[i] - s;
        V2 e2 = i + 1 < sh.count ? sh.v[i + 1] - s : sh.v[0] - s;
        float D = cross(e1, e2);
        float triangleArea = 0.5f * D;
        area += triangleArea;
        center += (triangleArea * k_inv3) * (e1 + e2);
        float ex1 = e1.x, ey1 = e1.y;
        float ex2 = e2.x, ey2 = e2.y;
        float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
        float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
        I += (0.25f * k_inv3 * D) * (intx2 + inty2);
    }
    md.mass = density * area;
    center *= 1.0f / area;
    md.center = center + s;
    md.I = density * I;
    md.I += md.mass * (dot(md.center, md.center) - dot(center, center));
}
/* b2CircleShape.ComputeMass b2circleshape.d:120-127 */
static void circleMass(const Shape &sh, MassData &md, float density) {
    md.mass = density * kPi * sh.radius * sh.radius;
    md.center = sh.p;
    md.I = md.mass * (0.5f * sh.radius * sh.radius + dot(sh.p, sh.p));
}
static void computeMass(const Shape &sh, MassData &md, float density) {
    if (sh.type == s_polygon) polyMass(sh, md, density);
    else if (sh.type == s_circle) circleMass(sh, md, density);
    else if (sh.type == s_chain) { md.mass = 0.0f; md.center = mk(0.0f, 0.0f); md.I = 0.0f; } /* b2chainshape.d:228-234 */
    else { md.mass = 0.0f; md.center = 0.5f * (sh.v1 + sh.v2); md.I = 0.0f; } /* b2edgeshape.d:179-186 */
}
/* ComputeAABB: polygon b2polygonshape.d:356-373, circle b2circleshape.d:110-117, edge b2edgeshape.d:163-176 */
static void computeAABB(const Shape &sh, AABB &aabb, const Xf &xf, int childIndex = 0) {
    if (sh.type == s_chain) { /* b2chainshape.d:214-226: no radius */
        int i1 = childIndex, i2 = childIndex + 1;
        if (i2 == sh.ccount) i2 = 0;
        V2 v1 = mul(xf, sh.cv[i1]);
        V2 v2 = mul(xf, sh.cv[i2]);
        aabb.lo = vmin(v1, v2);
        aabb.hi = vmax(v1, v2);
        return;
    }
    if (sh.type == s_polygon) {
        V2 lower = mul(xf, sh.v[0]);
        V2 upper = lower;
        for (int i = 1; i < sh.count; ++i) {
            V2 v = mul(xf, sh.v[i]);
            lower = vmin(lower, v);
            upper = vmax(upper, v);
        }
        V2 r = mk(sh.radius, sh.radius);
        aabb.lo = lower - r;
        aabb.hi = upper + r;
    } else if (sh.type == s_circle) {
        V2 p = xf.p + mul(xf.q, sh.p);
        aabb.lo = mk(p.x - sh.radius, p.y - sh.radius);
        aabb.hi = mk(p.x + sh.radius, p.y + sh.radius);
    } else {
        V2 v1 = mul(xf, sh.v1);
        V2 v2 = mul(xf, sh.v2);
        V2 lower = vmin(v1, v2);
        V2 upper = vmax(v1, v2);
        V2 r = mk(sh.radius, sh.radius);
        aabb.lo = lower - r;
        aabb.hi = upper + r;
    }
}

/* ---- narrowphase: b2collidepolygon.d ---- */
static float findMaxSeparation(int *edgeIndex, const Shape &poly1, const Xf &xf1, const Shape &poly2, const Xf &xf2) {
    int count1 = poly1.count;
    int count2 = poly2.count;
    Xf xf = mulT(xf2, xf1);
    int bestIndex = 0;
    float maxSeparation = -kMaxFloat;
    for (int i = 0; i < count1; ++i) {
        V2 n = mul(xf.q, poly1.n[i]);
        V2 v1 = mul(xf, poly1.v[i]);
        float si = kMaxFloat;
        for (int j = 0; j < count2; ++j) {
            float sij = dot(n, poly2.v[j] - v1);
            if (sij < si) si = sij;
        }
        if (si > maxSeparation) { maxSeparation = si; bestIndex = i; }
    }
    *edgeIndex = bestIndex;
    return maxSeparation;
}
static void findIncidentEdge(ClipVertex c[2], const Shape &poly1, const Xf &xf1, int edge1, const Shape &poly2, const Xf &xf2) {
    int count2 = poly2.count;
    V2 normal1 = mulT(xf2.q, mul(xf1.q, poly1.n[edge1]));
    int index = 0;
    float minDot = kMaxFloat;
    for (int i = 0; i < count2; ++i) {
        float d = dot(normal1, poly2.n[i]);
        if (d < minDot) { minDot = d; index = i; }
    }
    int i1 = index;
    int i2 = i1 + 1 < count2 ? i1 + 1 : 0;
    c[0].v = mul(xf2, poly2.v[i1]);
    c[0].id = makeKey(edge1, i1, e_face, e_vertex);
    c[1].v = mul(xf2, poly2.v[i2]);
    c[1].id = makeKey(edge1, i2, e_face, e_vertex);
}
/* b2CollidePolygons b2collidepolygon.d:127-254 */
static void collidePolygons(Manifold &manifold, const Shape &polyA, const Xf &xfA, const Shape &polyB, const Xf &xfB) {
    manifold.pointCount = 0;
    float totalRadius = polyA.radius + polyB.radius;
    int edgeA = 0;
    float separationA = findMaxSeparation(&edgeA, polyA, xfA, polyB, xfB);
    if (separationA > totalRadius) return;
    int edgeB = 0;
    float separationB = findMaxSeparation(&edgeB, polyB, xfB, polyA, xfA);
    if (separationB > totalRadius) return;
    const Shape *poly1, *poly2;
    Xf xf1, xf2;
    int edge1;
    int flip;
    const float k_tol = 0.1f * kLinearSlop;
    if (separationB > separationA + k_tol) {
        poly1 = &polyB; poly2 = &polyA; xf1 = xfB; xf2 = xfA; edge1 = edgeB;
        manifold.type = m_faceB;
        flip = 1;
    } else {
        poly1 = &polyA; poly2 = &polyB; xf1 = xfA; xf2 = xfB; edge1 = edgeA;
        manifold.type = m_faceA;
        flip = 0;
    }
    ClipVertex incidentEdge[2];
    findIncidentEdge(incidentEdge, *poly1, xf1, edge1, *poly2, xf2);
    int count1 = poly1->count;
    int iv1 = edge1;
    int iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
    V2 v11 = poly1->v[iv1];
    V2 v12 = poly1->v[iv2];
    V2 localTangent = v12 - v11;
    normalize(localTangent);
    V2 localNormal = cross(localTangent, 1.0f);
    V2 planePoint = 0.5f * (v11 + v12);
    V2 tangent = mul(xf1.q, localTangent);
    V2 normal = cross(tangent, 1.0f);
    v11 = mul(xf1, v11);
    v12 = mul(xf1, v12);
    float frontOffset = dot(normal, v11);
    float sideOffset1 = -dot(tangent, v11) + totalRadius;
    float sideOffset2 = dot(tangent, v12) + totalRadius;
    ClipVertex clipPoints1[2];
    ClipVertex clipPoints2[2];
    int np = clipSegmentToLine(clipPoints1, incidentEdge, -tangent, sideOffset1, iv1);
    if (np < 2) return;
    np = clipSegmentToLine(clipPoints2, clipPoints1, tangent, sideOffset2, iv2);
    if (np < 2) return;
    manifold.localNormal = localNormal;
    manifold.localPoint = planePoint;
    int pointCount = 0;
    for (int i = 0; i < kMaxManifoldPoints; ++i) {
        float separation = dot(normal, clipPoints2[i].v) - frontOffset;
        if (separation <= totalRadius) {
            ManifoldPoint &cp = manifold.points[pointCount];
            cp.localPoint = mulT(xf2, clipPoints2[i].v);
            cp.id = clipPoints2[i].id;
            if (flip) {
                uint32_t k = cp.id;
                cp.id = makeKey(keyIndexB(k), keyIndexA(k), keyTypeB(k), keyTypeA(k));
            }
            ++pointCount;
        }
    }
    manifold.pointCount = pointCount;
}

/* ---- b2collidecircle.d ---- */
static void collideCircles(Manifold &manifold, const Shape &circleA, const Xf &xfA, const Shape &circleB, const Xf &xfB) {
    manifold.pointCount = 0;
    V2 pA = mul(xfA, circleA.p);
    V2 pB = mul(xfB, circleB.p);
    V2 d = pB - pA;
    float distSqr = dot(d, d);
    float rA = circleA.radius, rB = circleB.radius;
    float radius = rA + rB;
    if (distSqr > radius * radius) return;
    manifold.type = m_circles;
    manifold.localPoint = circleA.p;
    manifold.localNormal = mk(0.0f, 0.0f);
    manifold.pointCount = 1;
    manifold.points[0].localPoint = circleB.p;
    manifold.points[0].id = 0;
}
static void collidePolygonAndCircle(Manifold &manifold, const Shape &polygonA, const Xf &xfA, const Shape &circleB, const Xf &xfB) {
    manifold.pointCount = 0;
    V2 c = mul(xfB, circleB.p);
    V2 cLocal = mulT(xfA, c);
    int normalIndex = 0;
    float separation = -kMaxFloat;
    float radius = polygonA.radius + circleB.radius;
    int vertexCount = polygonA.count;
    const V2 *vertices = polygonA.v;
    const V2 *normals = polygonA.n;
    for (int i = 0; i < vertexCount; ++i) {
        float s = dot(normals[i], cLocal - vertices[i]);
        if (s > radius) return;
        if (s > separation) { separation = s; normalIndex = i; }
    }
    int vertIndex1 = normalIndex;
    int vertIndex2 = vertIndex1 + 1 < vertexCount ? vertIndex1 + 1 : 0;
    V2 v1 = vertices[vertIndex1];
    V2 v2 = vertices[vertIndex2];
    if (separation < kEps) {
        manifold.pointCount = 1;
        manifold.type = m_faceA;
        manifold.localNormal = normals[normalIndex];
        manifold.localPoint = 0.5f * (v1 + v2);
        manifold.points[0].localPoint = circleB.p;
        manifold.points[0].id = 0;
        return;
    }
    float u1 = dot(cLocal - v1, v2 - v1);
    float u2 = dot(cLocal - v2, v1 - v2);
    if (u1 <= 0.0f) {
        if (distSq(cLocal, v1) > radius * radius) return;
        manifold.pointCount = 1;
        manifold.type = m_faceA;
        manifold.localNormal = cLocal - v1;
        normalize(manifold.localNormal);
        manifold.localPoint = v1;
        manifold.points[0].localPoint = circleB.p;
        manifold.points[0].id = 0;
    } else if (u2 <= 0.0f) {
        if (distSq(cLocal, v2) > radius * radius) return;
        manifold.pointCount = 1;
        manifold.type = m_faceA;
        manifold.localNormal = cLocal - v2;
        normalize(manifold.localNormal);
        manifold.localPoint = v2;
        manifold.points[0].localPoint = circleB.p;
        manifold.points[0].id = 0;
    } else {
        V2 faceCenter = 0.5f * (v1 + v2);
        float separation2 = dot(cLocal - faceCenter, normals[vertIndex1]);
        if (separation2 > radius) return;
        manifold.pointCount = 1;
        manifold.type = m_faceA;
        manifold.localNormal = normals[vertIndex1];
        manifold.localPoint = faceCenter;
        manifold.points[0].localPoint = circleB.p;
        manifold.points[0].id = 0;
    }
}

/* ---- b2collideedge.d ---- */
static void collideEdgeAndCircle(Manifold &manifold, const Shape &edgeA, const Xf &xfA, const Shape &circleB, const Xf &xfB) {
    manifold.pointCount = 0;
    V2 Q = mulT(xfA, mul(xfB, circleB.p));
    V2 A = edgeA.v1, B = edgeA.v2;
    V2 e = B - A;
    float u = dot(e, B - Q);
    float v = dot(e, Q - A);
    float radius = edgeA.radius + circleB.radius;
    /* cf.indexB = 0, cf.typeB = e_vertex */
    if (v <= 0.0f) {
        V2 P = A;
        V2 d = Q - P;
        float dd = dot(d, d);
        if (dd > radius * radius) return;
        if (edgeA.has0) {
            V2 A1 = edgeA.v0;
            V2 B1 = A;
            V2 e1 = B1 - A1;
            float u1 = dot(e1, B1 - Q);
            if (u1 > 0.0f) return;
        }
        manifold.pointCount = 1;
        manifold.type = m_circles;
        manifold.localNormal = mk(0.0f, 0.0f);
        manifold.localPoint = P;
        manifold.points[0].id = makeKey(0, 0, e_vertex, e_vertex);
        manifold.points[0].localPoint = circleB.p;
        return;
    }
    if (u <= 0.0f) {
        V2 P = B;
        V2 d = Q - P;
        float dd = dot(d, d);
        if (dd > radius * radius) return;
        if (edgeA.has3) {
            V2 B2 = edgeA.v3;
            V2 A2 = B;
            V2 e2 = B2 - A2;
            float v2 = dot(e2, Q - A2);
            if (v2 > 0.0f) return;
        }
        manifold.pointCount = 1;
        manifold.type = m_circles;
        manifold.localNormal = mk(0.0f, 0.0f);
        manifold.localPoint = P;
        manifold.points[0].id = makeKey(1, 0, e_vertex, e_vertex);
        manifold.points[0].localPoint = circleB.p;
        return;
    }
    float den = dot(e, e);
    V2 P = (1.0f / den) * (u * A + v * B);
    V2 d = Q - P;
    float dd = dot(d, d);
    if (dd > radius * radius) return;
    V2 n = mk(-e.y, e.x);
    if (dot(n, Q - A) < 0.0f) n = mk(-n.x, -n.y);
    normalize(n);
    manifold.pointCount = 1;
    manifold.type = m_faceA;
    manifold.localNormal = n;
    manifold.localPoint = A;
    manifold.points[0].id = makeKey(0, 0, e_face, e_vertex);
    manifold.points[0].localPoint = circleB.p;
}

struct EPAxis {
    enum { e_unknown = 0, e_edgeA = 1, e_edgeB = 2 };
    int type = e_unknown;
    int index = 0;
    float separation = 0;
};
/* b2EPCollider b2collideedge.d:206-716 */
struct EPCollider {
    V2 pv[kMaxPolygonVertices], pn[kMaxPolygonVertices];
    int pcount = 0;
    Xf m_xf;
    V2 m_centroidB, m_v0, m_v1, m_v2, m_v3, m_normal0, m_normal1, m_normal2, m_normal, m_lowerLimit, m_upperLimit;
    float m_radius = 0;
    bool m_front = false;

    void collide(Manifold &manifold, const Shape &edgeA, const Xf &xfA, const Shape &polygonB, const Xf &xfB) {
        m_xf = mulT(xfA, xfB);
        m_centroidB = mul(m_xf, polygonB.centroid);
        m_v0 = edgeA.v0; m_v1 = edgeA.v1; m_v2 = edgeA.v2; m_v3 = edgeA.v3;
        bool hasVertex0 = edgeA.has0;
        bool hasVertex3 = edgeA.has3;
        V2 edge1 = m_v2 - m_v1;
        normalize(edge1);
        m_normal1 = mk(edge1.y, -edge1.x);
        float offset1 = dot(m_normal1, m_centroidB - m_v1);
        float offset0 = 0.0f, offset2 = 0.0f;
        bool convex1 = false, convex2 = false;
        if (hasVertex0) {
            V2 edge0 = m_v1 - m_v0;
            normalize(edge0);
            m_normal0 = mk(edge0.y, -edge0.x);
            convex1 = cross(edge0, edge1) >= 0.0f;
            offset0 = dot(m_normal0, m_centroidB - m_v0);
        }
        if (hasVertex3) {
            V2 edge2 = m_v3 - m_v2;
            normalize(edge2);
            m_normal2 = mk(edge2.y, -edge2.x);
            convex2 = cross(edge1, edge2) > 0.0f;
            offset2 = dot(m_normal2, m_centroidB - m_v2);
        }
        if (hasVertex0 && hasVertex3) {
            if (convex1 && convex2) {
                m_front = offset0 >= 0.0f || offset1 >= 0.0f || offset2 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal0; m_upperLimit = m_normal2; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = -m_normal1; }
            } else if (convex1) {
                m_front = offset0 >= 0.0f || (offset1 >= 0.0f && offset2 >= 0.0f);
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal0; m_upperLimit = m_normal1; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal2; m_upperLimit = -m_normal1; }
            } else if (convex2) {
                m_front = offset2 >= 0.0f || (offset0 >= 0.0f && offset1 >= 0.0f);
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal1; m_upperLimit = m_normal2; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = -m_normal0; }
            } else {
                m_front = offset0 >= 0.0f && offset1 >= 0.0f && offset2 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal1; m_upperLimit = m_normal1; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal2; m_upperLimit = -m_normal0; }
            }
        } else if (hasVertex0) {
            if (convex1) {
                m_front = offset0 >= 0.0f || offset1 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal0; m_upperLimit = -m_normal1; }
                else { m_normal = -m_normal1; m_lowerLimit = m_normal1; m_upperLimit = -m_normal1; }
            } else {
                m_front = offset0 >= 0.0f && offset1 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = m_normal1; m_upperLimit = -m_normal1; }
                else { m_normal = -m_normal1; m_lowerLimit = m_normal1; m_upperLimit = -m_normal0; }
            }
        } else if (hasVertex3) {
            if (convex2) {
                m_front = offset1 >= 0.0f || offset2 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = m_normal2; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = m_normal1; }
            } else {
                m_front = offset1 >= 0.0f && offset2 >= 0.0f;
                if (m_front) { m_normal = m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = m_normal1; }
                else { m_normal = -m_normal1; m_lowerLimit = -m_normal2; m_upperLimit = m_normal1; }
            }
        } else {
            m_front = offset1 >= 0.0f;
            if (m_front) { m_normal = m_normal1; m_lowerLimit = -m_normal1; m_upperLimit = -m_normal1; }
            else { m_normal = -m_normal1; m_lowerLimit = m_normal1; m_upperLimit = m_normal1; }
        }
        pcount = polygonB.count;
        for (int i = 0; i < polygonB.count; ++i) {
            pv[i] = mul(m_xf, polygonB.v[i]);
            pn[i] = mul(m_xf.q, polygonB.n[i]);
        }
        m_radius = 2.0f * kPolygonRadius;
        manifold.pointCount = 0;
        EPAxis edgeAxis = computeEdgeSeparation();
        if (edgeAxis.type == EPAxis::e_unknown) return;
        if (edgeAxis.separation > m_radius) return;
        EPAxis polygonAxis = computePolygonSeparation();
        if (polygonAxis.type != EPAxis::e_unknown && polygonAxis.separation > m_radius) return;
        const float k_relativeTol = 0.98f;
        const float k_absoluteTol = 0.001f;
        EPAxis primaryAxis;
        if (polygonAxis.type == EPAxis::e_unknown) primaryAxis = edgeAxis;
        else if (polygonAxis.separation > k_relativeTol * edgeAxis.separation + k_absoluteTol) primaryAxis = polygonAxis;
        else primaryAxis = edgeAxis;
        ClipVertex ie[2];
        int rf_i1, rf_i2;
        V2 rf_v1, rf_v2, rf_normal;
        if (primaryAxis.type == EPAxis::e_edgeA) {
            manifold.type = m_faceA;
            int bestIndex = 0;
            float bestValue = dot(m_normal, pn[0]);
            for (int i = 1; i < pcount; ++i) {
                float value = dot(m_normal, pn[i]);
                if (value < bestValue) { bestValue = value; bestIndex = i; }
            }
            int i1 = bestIndex;
            int i2 = i1 + 1 < pcount ? i1 + 1 : 0;
            ie[0].v = pv[i1];
            ie[0].id = makeKey(0, i1, e_face, e_vertex);
            ie[1].v = pv[i2];
            ie[1].id = makeKey(0, i2, e_face, e_vertex);
            if (m_front) { rf_i1 = 0; rf_i2 = 1; rf_v1 = m_v1; rf_v2 = m_v2; rf_normal = m_normal1; }
            else { rf_i1 = 1; rf_i2 = 0; rf_v1 = m_v2; rf_v2 = m_v1; rf_normal = -m_normal1; }
        } else {
            manifold.type = m_faceB;
            ie[0].v = m_v1;
            ie[0].id = makeKey(0, primaryAxis.index, e_vertex, e_face);
            ie[1].v = m_v2;
            ie[1].id = makeKey(0, primaryAxis.index, e_vertex, e_face);
            rf_i1 = primaryAxis.index;
            rf_i2 = rf_i1 + 1 < pcount ? rf_i1 + 1 : 0;
            rf_v1 = pv[rf_i1];
            rf_v2 = pv[rf_i2];
            rf_normal = pn[rf_i1];
        }
        V2 sideNormal1 = mk(rf_normal.y, -rf_normal.x);
        V2 sideNormal2 = -sideNormal1;
        float sideOffset1 = dot(sideNormal1, rf_v1);
        float sideOffset2 = dot(sideNormal2, rf_v2);
        ClipVertex clipPoints1[2], clipPoints2[2];
        int np = clipSegmentToLine(clipPoints1, ie, sideNormal1, sideOffset1, rf_i1);
        if (np < kMaxManifoldPoints) return;
        np = clipSegmentToLine(clipPoints2, clipPoints1, sideNormal2, sideOffset2, rf_i2);
        if (np < kMaxManifoldPoints) return;
        if (primaryAxis.type == EPAxis::e_edgeA) {
            manifold.localNormal = rf_normal;
            manifold.localPoint = rf_v1;
        } else {
            manifold.localNormal = polygonB.n[rf_i1];
            manifold.localPoint = polygonB.v[rf_i1];
        }
        int pointCount = 0;
        for (int i = 0; i < kMaxManifoldPoints; ++i) {
            float separation = dot(rf_normal, clipPoints2[i].v - rf_v1);
            if (separation <= m_radius) {
                ManifoldPoint &cp = manifold.points[pointCount];
                if (primaryAxis.type == EPAxis::e_edgeA) {
                    cp.localPoint = mulT(m_xf, clipPoints2[i].v);
                    cp.id = clipPoints2[i].id;
                } else {
                    uint32_t k = clipPoints2[i].id;
                    cp.localPoint = clipPoints2[i].v;
                    cp.id = makeKey(keyIndexB(k), keyIndexA(k), keyTypeB(k), keyTypeA(k));
                }
                ++pointCount;
            }
        }
        manifold.pointCount = pointCount;
    }
    EPAxis computeEdgeSeparation() const {
        EPAxis axis;
        axis.type = EPAxis::e_edgeA;
        axis.index = m_front ? 0 : 1;
        axis.separation = FLT_MAX;
        for (int i = 0; i < pcount; ++i) {
            float s = dot(m_normal, pv[i] - m_v1);
            if (s < axis.separation) axis.separation = s;
        }
        return axis;
    }
    EPAxis computePolygonSeparation() const {
        EPAxis axis;
        axis.type = EPAxis::e_unknown;
        axis.index = -1;
        axis.separation = -FLT_MAX;
        V2 perp = mk(-m_normal.y, m_normal.x);
        for (int i = 0; i < pcount; ++i) {
            V2 n = -pn[i];
            float s1 = dot(n, pv[i] - m_v1);
            float s2 = dot(n, pv[i] - m_v2);
            float s = fmin_(s1, s2);
            if (s > m_radius) {
                axis.type = EPAxis::e_edgeB;
                axis.index = i;
                axis.separation = s;
                return axis;
            }
            if (dot(n, perp) >= 0.0f) {
                if (dot(n - m_upperLimit, m_normal) < -kAngularSlop) continue;
            } else {
                if (dot(n - m_lowerLimit, m_normal) < -kAngularSlop) continue;
            }
            if (s > axis.separation) {
                axis.type = EPAxis::e_edgeB;
                axis.index = i;
                axis.separation = s;
            }
        }
        return axis;
    }
};
static void collideEdgeAndPolygon(Manifold &m, const Shape &edgeA, const Xf &xfA, const Shape &polyB, const Xf &xfB) {
    EPCollider c;
    c.collide(m, edgeA, xfA, polyB, xfB);
}

/* Contact type registry b2contact.d:430-436: primary orders are
 * circle-circle, polygon-circle, polygon-polygon, edge-circle, edge-polygon. */
/* b2Contact.InitializeRegisters (b2contact.d:425-435): edge and chain pair only
 * with circles and polygons, and are then shape A */
static bool hasContactType(int ta, int tb) {
    const bool la = ta == s_edge || ta == s_chain, lb = tb == s_edge || tb == s_chain;
    return !(la && lb);
}
static bool isPrimary(int ta, int tb) {
    if (ta == tb) return true;
    if (ta == s_polygon && tb == s_circle) return true;
    if (ta == s_edge || ta == s_chain) return true;
    return false;
}
/* The Evaluate override of each contact class (b2polygoncontact.d:56 etc.); a
 * chain is evaluated as its child edge indexA (b2chainandpolygoncontact.d,
 * b2chainandcirclecontact.d: GetChildEdge then the edge collide). */
static void evaluate(Manifold &m, const Shape &a, const Xf &xfA, const Shape &b, const Xf &xfB, int indexA = 0) {
    if (a.type == s_chain) {
        Shape edge;
        childEdge(a, indexA, edge);
        if (b.type == s_circle) collideEdgeAndCircle(m, edge, xfA, b, xfB);
        else if (b.type == s_polygon) collideEdgeAndPolygon(m, edge, xfA, b, xfB);
        else m.pointCount = 0;
        return;
    }
    if (a.type == s_polygon && b.type == s_polygon) collidePolygons(m, a, xfA, b, xfB);
    else if (a.type == s_circle && b.type == s_circle) collideCircles(m, a, xfA, b, xfB);
    else if (a.type == s_polygon && b.type == s_circle) collidePolygonAndCircle(m, a, xfA, b, xfB);
    else if (a.type == s_edge && b.type == s_circle) collideEdgeAndCircle(m, a, xfA, b, xfB);
    else if (a.type == s_edge && b.type == s_polygon) collideEdgeAndPolygon(m, a, xfA, b, xfB);
    else m.pointCount = 0;
}

/* ---- b2DynamicTree b2dynamictree.d ---- */
constexpr int kNull = -1;
struct TreeNode {
    AABB aabb;
    int userData = -1;
    int parentOrNext = kNull;
    int child1 = kNull, child2 = kNull;
    int height = 0;
    bool isLeaf() const { return child1 == kNull; }
};
struct DynamicTree {
    int root = kNull;
    std::vector<TreeNode> nodes;
    int nodeCount = 0, nodeCapacity = 16, freeList = 0;
    DynamicTree() {
        nodes.resize(nodeCapacity);
        for (int i = 0; i < nodeCapacity - 1; ++i) { nodes[i].parentOrNext = i + 1; nodes[i].height = -1; }
        nodes[nodeCapacity - 1].parentOrNext = kNull;
        nodes[nodeCapacity - 1].height = -1;
    }
    int allocateNode() {
        if (freeList == kNull) {
            nodeCapacity *= 2;
            nodes.resize(nodeCapacity);
            for (int i = nodeCount; i < nodeCapacity - 1; ++i) { nodes[i].parentOrNext = i + 1; nodes[i].height = -1; }
            nodes[nodeCapacity - 1].parentOrNext = kNull;
            nodes[nodeCapacity - 1].height = -1;
            freeList = nodeCount;
        }
        int id = freeList;
        freeList = nodes[id].parentOrNext;
        nodes[id].parentOrNext = kNull;
        nodes[id].child1 = kNull;
        nodes[id].child2 = kNull;
        nodes[id].height = 0;
        nodes[id].userData = -1;
        ++nodeCount;
        return id;
    }
    void freeNode(int id) {
        nodes[id].parentOrNext = freeList;
        nodes[id].height = -1;
        freeList = id;
        --nodeCount;
    }
    int createProxy(const AABB &aabb, int userData) {
        int id = allocateNode();
        V2 r = mk(kAabbExtension, kAabbExtension);
        nodes[id].aabb.lo = aabb.lo - r;
        nodes[id].aabb.hi = aabb.hi + r;
        nodes[id].userData = userData;
        nodes[id].height = 0;
        insertLeaf(id);
        return id;
    }
    int createProxyFat(const AABB &fat, int userData) { /* test hook: leaf with a given fat AABB */
        int id = allocateNode();
        nodes[id].aabb = fat;
        nodes[id].userData = userData;
        nodes[id].height = 0;
        insertLeaf(id);
        return id;
    }
    /* b2DynamicTree.MoveProxy b2dynamictree.d:140-184 */
    bool moveProxy(int id, const AABB &aabb, V2 displacement) {
        if (nodes[id].aabb.contains(aabb)) return false;
        removeLeaf(id);
        AABB b = aabb;
        V2 r = mk(kAabbExtension, kAabbExtension);
        b.lo = b.lo - r;
        b.hi = b.hi + r;
        V2 d = kAabbMultiplier * displacement;
        if (d.x < 0.0f) b.lo.x += d.x; else b.hi.x += d.x;
        if (d.y < 0.0f) b.lo.y += d.y; else b.hi.y += d.y;
        nodes[id].aabb = b;
        insertLeaf(id);
        return true;
    }
    template <class F>
    void query(F &&callback, const AABB &aabb) const {
        int stackBuf[256];
        std::vector<int> big;
        int count = 0;
        auto push = [&](int v) {
            if (count < 256) stackBuf[count++] = v;
            else { if (big.empty()) big.assign(stackBuf, stackBuf + 256); big.push_back(v); ++count; }
        };
        auto pop = [&]() -> int {
            --count;
            if (!big.empty()) { int v = big.back(); big.pop_back(); if (big.size() == 256) big.clear(); return v; }
            return stackBuf[count];
        };
        push(root);
        while (count > 0) {
            int nodeId = pop();
            if (nodeId == kNull) continue;
            const TreeNode &node = nodes[nodeId];
            if (testOverlap(node.aabb, aabb)) {
                if (node.isLeaf()) {
                    if (!callback(nodeId)) return;
                } else {
                    push(node.child1);
                    push(node.child2);
                }
            }
        }
    }
    void insertLeaf(int leaf) {
        if (root == kNull) { root = leaf; nodes[root].parentOrNext = kNull; return; }
        AABB leafAABB = nodes[leaf].aabb;
        int index = root;
        while (!nodes[index].isLeaf()) {
            int child1 = nodes[index].child1;
            int child2 = nodes[index].child2;
            float area = nodes[index].aabb.perimeter();
            AABB combinedAABB;
            combinedAABB.combine(nodes[index].aabb, leafAABB);
            float combinedArea = combinedAABB.perimeter();
            float cost = 2.0f * combinedArea;
            float inheritanceCost = 2.0f * (combinedArea - area);
            float cost1;
            if (nodes[child1].isLeaf()) {
                AABB aabb; aabb.combine(leafAABB, nodes[child1].aabb);
                cost1 = aabb.perimeter() + inheritanceCost;
            } else {
                AABB aabb; aabb.combine(leafAABB, nodes[child1].aabb);
                float oldArea = nodes[child1].aabb.perimeter();
                float newArea = aabb.perimeter();
                cost1 = (newArea - oldArea) + inheritanceCost;
            }
            float cost2;
            if (nodes[child2].isLeaf()) {
                AABB aabb; aabb.combine(leafAABB, nodes[child2].aabb);
                cost2 = aabb.perimeter() + inheritanceCost;
            } else {
                AABB aabb; aabb.combine(leafAABB, nodes[child2].aabb);
                float oldArea = nodes[child2].aabb.perimeter();
                float newArea = aabb.perimeter();
                cost2 = newArea - oldArea + inheritanceCost;
            }
            if (cost < cost1 && cost < cost2) break;
            index = (cost1 < cost2) ? child1 : child2;
        }
        int sibling = index;
        int oldParent = nodes[sibling].parentOrNext;
        int newParent = allocateNode();
        nodes[newParent].parentOrNext = oldParent;
        nodes[newParent].userData = -1;
        nodes[newParent].aabb.combine(leafAABB, nodes[sibling].aabb);
        nodes[newParent].height = nodes[sibling].height + 1;
        if (oldParent != kNull) {
            if (nodes[oldParent].child1 == sibling) nodes[oldParent].child1 = newParent;
            else nodes[oldParent].child2 = newParent;
        } else {
            root = newParent;
        }
        nodes[newParent].child1 = sibling;
        nodes[newParent].child2 = leaf;
        nodes[sibling].parentOrNext = newParent;
        nodes[leaf].parentOrNext = newParent;
        index = nodes[leaf].parentOrNext;
        while (index != kNull) {
            index = balance(index);
            int c1 = nodes[index].child1, c2 = nodes[index].child2;
            nodes[index].height = 1 + std::max(nodes[c1].height, nodes[c2].height);
            nodes[index].aabb.combine(nodes[c1].aabb, nodes[c2].aabb);
            index = nodes[index].parentOrNext;
        }
    }
    void removeLeaf(int leaf) {
        if (leaf == root) { root = kNull; return; }
        int parent = nodes[leaf].parentOrNext;
        int grandParent = nodes[parent].parentOrNext;
        int sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;
        if (grandParent != kNull) {
            if (nodes[grandParent].child1 == parent) nodes[grandParent].child1 = sibling;
            else nodes[grandParent].child2 = sibling;
            nodes[sibling].parentOrNext = grandParent;
            freeNode(parent);
            int index = grandParent;
            while (index != kNull) {
                index = balance(index);
                int c1 = nodes[index].child1, c2 = nodes[index].child2;
                nodes[index].aabb.combine(nodes[c1].aabb, nodes[c2].aabb);
                nodes[index].height = 1 + std::max(nodes[c1].height, nodes[c2].height);
                index = nodes[index].parentOrNext;
            }
        } else {
            root = sibling;
            nodes[sibling].parentOrNext = kNull;
            freeNode(parent);
        }
    }
    /* b2DynamicTree.Balance b2dynamictree.d:773-915 */
    int balance(int iA) {
        TreeNode *A = &nodes[iA];
        if (A->isLeaf() || A->height < 2) return iA;
        int iB = A->child1, iC = A->child2;
        TreeNode *B = &nodes[iB], *C = &nodes[iC];
        int bal = C->height - B->height;
        if (bal > 1) {
            int iF = C->child1, iG = C->child2;
            TreeNode *F = &nodes[iF], *G = &nodes[iG];
            C->child1 = iA;
            C->parentOrNext = A->parentOrNext;
            A->parentOrNext = iC;
            if (C->parentOrNext != kNull) {
                if (nodes[C->parentOrNext].child1 == iA) nodes[C->parentOrNext].child1 = iC;
                else nodes[C->parentOrNext].child2 = iC;
            } else {
                root = iC;
            }
            if (F->height > G->height) {
                C->child2 = iF; A->child2 = iG; G->parentOrNext = iA;
                A->aabb.combine(B->aabb, G->aabb);
                C->aabb.combine(A->aabb, F->aabb);
                A->height = 1 + std::max(B->height, G->height);
                C->height = 1 + std::max(A->height, F->height);
            } else {
                C->child2 = iG; A->child2 = iF; F->parentOrNext = iA;
                A->aabb.combine(B->aabb, F->aabb);
                C->aabb.combine(A->aabb, G->aabb);
                A->height = 1 + std::max(B->height, F->height);
                C->height = 1 + std::max(A->height, G->height);
            }
            return iC;
        }
        if (bal < -1) {
            int iD = B->child1, iE = B->child2;
            TreeNode *D = &nodes[iD], *E = &nodes[iE];
            B->child1 = iA;
            B->parentOrNext = A->parentOrNext;
            A->parentOrNext = iB;
            if (B->parentOrNext != kNull) {
                if (nodes[B->parentOrNext].child1 == iA) nodes[B->parentOrNext].child1 = iB;
                else nodes[B->parentOrNext].child2 = iB;
            } else {
                root = iB;
            }
            if (D->height > E->height) {
                B->child2 = iD; A->child1 = iE; E->parentOrNext = iA;
                A->aabb.combine(C->aabb, E->aabb);
                B->aabb.combine(A->aabb, D->aabb);
                A->height = 1 + std::max(C->height, E->height);
                B->height = 1 + std::max(A->height, D->height);
            } else {
                B->child2 = iE; A->child1 = iD; D->parentOrNext = iA;
                A->aabb.combine(C->aabb, D->aabb);
                B->aabb.combine(A->aabb, E->aabb);
                A->height = 1 + std::max(C->height, D->height);
                B->height = 1 + std::max(A->height, E->height);
            }
            return iB;
        }
        return iA;
    }
};

/* ---- b2BroadPhase b2broadphase.d ---- */
struct Pair {
    int a, b;
};
struct BroadPhase {
    DynamicTree tree;
    std::vector<int> moveBuffer;
    std::vector<Pair> pairBuffer;
    int queryProxyId = kNull;
    int createProxy(const AABB &aabb, int userData) {
        int id = tree.createProxy(aabb, userData);
        moveBuffer.push_back(id);
        return id;
    }
    void moveProxy(int id, const AABB &aabb, V2 displacement) {
        if (tree.moveProxy(id, aabb, displacement)) moveBuffer.push_back(id);
    }
    bool testOverlapProxies(int a, int b) const { return testOverlap(tree.nodes[a].aabb, tree.nodes[b].aabb); }
    /* b2BroadPhase.UpdatePairs b2broadphase.d:139-195 */
    template <class F>
    void updatePairs(F &&addPair) {
        pairBuffer.clear();
        for (size_t i = 0; i < moveBuffer.size(); ++i) {
            queryProxyId = moveBuffer[i];
            if (queryProxyId == kNull) continue;
            AABB fat = tree.nodes[queryProxyId].aabb;
            tree.query(
                [&](int proxyId) {
                    if (proxyId == queryProxyId) return true;
                    Pair p;
                    p.a = std::min(proxyId, queryProxyId);
                    p.b = std::max(proxyId, queryProxyId);
                    pairBuffer.push_back(p);
                    return true;
                },
                fat);
        }
        moveBuffer.clear();
        std::sort(pairBuffer.begin(), pairBuffer.end(), [](const Pair &p1, const Pair &p2) {
            if (p1.a < p2.a) return true;
            if (p1.a == p2.a) return p1.b < p2.b;
            return false;
        });
        size_t i = 0;
        while (i < pairBuffer.size()) {
            Pair primary = pairBuffer[i];
            addPair(tree.nodes[primary.a].userData, tree.nodes[primary.b].userData, primary);
            ++i;
            while (i < pairBuffer.size()) {
                if (pairBuffer[i].a != primary.a || pairBuffer[i].b != primary.b) break;
                ++i;
            }
        }
    }
};

/* ---- bodies, fixtures, contacts ---- */
enum { b_static = 0, b_kinematic = 1, b_dynamic = 2 };
enum {
    bf_island = 0x0001, bf_awake = 0x0002, bf_autoSleep = 0x0004, bf_bullet = 0x0008,
    bf_fixedRotation = 0x0010, bf_active = 0x0020
};
enum { cf_island = 0x0001, cf_touching = 0x0002, cf_enabled = 0x0004, cf_filter = 0x0008, cf_toi = 0x0020 };
constexpr int kMaxSubSteps = 8;     /* b2settings.d:105 */
constexpr int kMaxTOIContacts = 32; /* b2settings.d:110 */
constexpr float kToiBaumgarte = 0.75f; /* b2settings.d:138 */

struct Sweep {
    V2 localCenter, c0, c;
    float a0 = 0, a = 0, alpha0 = 0;
    /* b2Sweep.GetTransform b2math.d:556-564 */
    void getTransform(Xf &xf, float beta) const {
        xf.p = (1.0f - beta) * c0 + beta * c;
        float angle = (1.0f - beta) * a0 + beta * a;
        xf.q.set(angle);
        xf.p -= mul(xf.q, localCenter);
    }
    /* b2Sweep.Advance b2math.d:568-575 */
    void advance(float alpha) {
        float beta = (alpha - alpha0) / (1.0f - alpha0);
        c0 += beta * (c - c0);
        a0 += beta * (a - a0);
        alpha0 = alpha;
    }
    /* b2Sweep.Normalize b2math.d:578-584 */
    void normalize() {
        float twoPi = 2.0f * kPi;
        float d = twoPi * floorf(a0 / twoPi);
        a0 -= d;
        a -= d;
    }
};

/* ---- b2Distance: GJK with a simplex cache (b2distance.d:30-707) ---- */
struct DistanceProxy {
    V2 buf[2];
    const V2 *vertices = nullptr;
    int count = 0;
    float radius = 0;
    DistanceProxy() = default;
    DistanceProxy(const DistanceProxy &) = delete; /* vertices may point into buf */
    /* b2DistanceProxy.Set b2distance.d:34-91 (no chain shapes) */
    void set(const Shape &s, int index = 0) {
        if (s.type == s_circle) {
            vertices = &s.p;
            count = 1;
        } else if (s.type == s_chain) { /* b2distance.d:56-74 */
            buf[0] = s.cv[index];
            buf[1] = index + 1 < s.ccount ? s.cv[index + 1] : s.cv[0];
            vertices = buf;
            count = 2;
        } else if (s.type == s_polygon) {
            vertices = s.v;
            count = s.count;
        } else {
            buf[0] = s.v1;
            buf[1] = s.v2;
            vertices = buf;
            count = 2;
        }
        radius = s.radius;
    }
    /* GetSupport b2distance.d:94-111 */
    int support(V2 d) const {
        int bestIndex = 0;
        float bestValue = dot(vertices[0], d);
        for (int i = 1; i < count; ++i) {
            float value = dot(vertices[i], d);
            if (value > bestValue) {
                bestIndex = i;
                bestValue = value;
            }
        }
        return bestIndex;
    }
};
struct SimplexCache {
    float metric = 0;
    int count = 0;
    int indexA[3] = {0, 0, 0}, indexB[3] = {0, 0, 0};
};
struct SimplexVertex {
    V2 wA, wB, w;
    float a = 0;
    int indexA = 0, indexB = 0;
};
struct Simplex {
    SimplexVertex v[3];
    int count = 0;
    /* ReadCache b2distance.d:574-624 */
    void readCache(const SimplexCache &cache, const DistanceProxy &pA, const Xf &xfA, const DistanceProxy &pB,
                   const Xf &xfB) {
        count = cache.count;
        for (int i = 0; i < count; ++i) {
            SimplexVertex &s = v[i];
            s.indexA = cache.indexA[i];
            s.indexB = cache.indexB[i];
            s.wA = mul(xfA, pA.vertices[s.indexA]);
            s.wB = mul(xfB, pB.vertices[s.indexB]);
            s.w = s.wB - s.wA;
            s.a = 0.0f;
        }
        if (count > 1) {
            float metric1 = cache.metric;
            float metric2 = metric();
            if (metric2 < 0.5f * metric1 || 2.0f * metric1 < metric2 || metric2 < kEps) count = 0;
        }
        if (count == 0) {
            SimplexVertex &s = v[0];
            s.indexA = 0;
            s.indexB = 0;
            s.wA = mul(xfA, pA.vertices[0]);
            s.wB = mul(xfB, pB.vertices[0]);
            s.w = s.wB - s.wA;
            s.a = 1.0f;
            count = 1;
        }
    }
    /* WriteCache b2distance.d:626-637 */
    void writeCache(SimplexCache &cache) const {
        cache.metric = metric();
        cache.count = count;
        for (int i = 0; i < count; ++i) {
            cache.indexA[i] = v[i].indexA;
            cache.indexB[i] = v[i].indexB;
        }
    }
    /* GetSearchDirection b2distance.d:639-662 */
    V2 searchDirection() const {
        if (count == 1) return -v[0].w;
        V2 e12 = v[1].w - v[0].w;
        float sgn = cross(e12, -v[0].w);
        return sgn > 0.0f ? cross(1.0f, e12) : cross(e12, 1.0f);
    }
    /* GetClosestPoint b2distance.d:664-684 */
    V2 closestPoint() const {
        if (count == 1) return v[0].w;
        if (count == 2) return v[0].a * v[0].w + v[1].a * v[1].w;
        return mk(0.0f, 0.0f);
    }
    /* GetWitnessPoints b2distance.d:686-708 */
    void witnessPoints(V2 &pA, V2 &pB) const {
        if (count == 1) {
            pA = v[0].wA;
            pB = v[0].wB;
        } else if (count == 2) {
            pA = v[0].a * v[0].wA + v[1].a * v[1].wA;
            pB = v[0].a * v[0].wB + v[1].a * v[1].wB;
        } else {
            pA = v[0].a * v[0].wA + v[1].a * v[1].wA + v[2].a * v[2].wA;
            pB = pA;
        }
    }
    /* GetMetric b2distance.d:710-728 */
    float metric() const {
        if (count == 1) return 0.0f;
        if (count == 2) return length(v[0].w - v[1].w);
        return cross(v[1].w - v[0].w, v[2].w - v[0].w);
    }
    /* Solve2 b2distance.d:373-406 */
    void solve2() {
        V2 w1 = v[0].w, w2 = v[1].w;
        V2 e12 = w2 - w1;
        float d12_2 = -dot(w1, e12);
        if (d12_2 <= 0.0f) {
            v[0].a = 1.0f;
            count = 1;
            return;
        }
        float d12_1 = dot(w2, e12);
        if (d12_1 <= 0.0f) {
            v[1].a = 1.0f;
            count = 1;
            v[0] = v[1];
            return;
        }
        float inv_d12 = 1.0f / (d12_1 + d12_2);
        v[0].a = d12_1 * inv_d12;
        v[1].a = d12_2 * inv_d12;
        count = 2;
    }
    /* Solve3 b2distance.d:413-571 */
    void solve3() {
        V2 w1 = v[0].w, w2 = v[1].w, w3 = v[2].w;
        V2 e12 = w2 - w1;
        float w1e12 = dot(w1, e12), w2e12 = dot(w2, e12);
        float d12_1 = w2e12, d12_2 = -w1e12;
        V2 e13 = w3 - w1;
        float w1e13 = dot(w1, e13), w3e13 = dot(w3, e13);
        float d13_1 = w3e13, d13_2 = -w1e13;
        V2 e23 = w3 - w2;
        float w2e23 = dot(w2, e23), w3e23 = dot(w3, e23);
        float d23_1 = w3e23, d23_2 = -w2e23;
        float n123 = cross(e12, e13);
        float d123_1 = n123 * cross(w2, w3);
        float d123_2 = n123 * cross(w3, w1);
        float d123_3 = n123 * cross(w1, w2);
        if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
            v[0].a = 1.0f;
            count = 1;
            return;
        }
        if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
            float inv_d12 = 1.0f / (d12_1 + d12_2);
            v[0].a = d12_1 * inv_d12;
            v[1].a = d12_2 * inv_d12;
            count = 2;
            return;
        }
        if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
            float inv_d13 = 1.0f / (d13_1 + d13_2);
            v[0].a = d13_1 * inv_d13;
            v[2].a = d13_2 * inv_d13;
            count = 2;
            v[1] = v[2];
            return;
        }
        if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
            v[1].a = 1.0f;
            count = 1;
            v[0] = v[1];
            return;
        }
        if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
            v[2].a = 1.0f;
            count = 1;
            v[0] = v[2];
            return;
        }
        if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
            float inv_d23 = 1.0f / (d23_1 + d23_2);
            v[1].a = d23_1 * inv_d23;
            v[2].a = d23_2 * inv_d23;
            count = 2;
            v[0] = v[2];
            return;
        }
        float inv_d123 = 1.0f / (d123_1 + d123_2 + d123_3);
        v[0].a = d123_1 * inv_d123;
        v[1].a = d123_2 * inv_d123;
        v[2].a = d123_3 * inv_d123;
        count = 3;
    }
};
struct DistanceOutput {
    V2 pointA, pointB;
    float distance = 0;
    int iterations = 0;
};
/* b2Distance b2distance.d:185-345 */
static void gjkDistance(DistanceOutput &output, SimplexCache &cache, const DistanceProxy &proxyA, const Xf &transformA,
                        const DistanceProxy &proxyB, const Xf &transformB, bool useRadii) {
    Simplex simplex;
    simplex.readCache(cache, proxyA, transformA, proxyB, transformB);
    const int k_maxIters = 20;
    int saveA[3], saveB[3];
    int iter = 0;
    while (iter < k_maxIters) {
        int saveCount = simplex.count;
        for (int i = 0; i < saveCount; ++i) {
            saveA[i] = simplex.v[i].indexA;
            saveB[i] = simplex.v[i].indexB;
        }
        if (simplex.count == 2) simplex.solve2();
        else if (simplex.count == 3) simplex.solve3();
        if (simplex.count == 3) break;
        V2 d = simplex.searchDirection();
        if (lengthSq(d) < kEps * kEps) break;
        SimplexVertex &vertex = simplex.v[simplex.count];
        vertex.indexA = proxyA.support(mulT(transformA.q, -d));
        vertex.wA = mul(transformA, proxyA.vertices[vertex.indexA]);
        vertex.indexB = proxyB.support(mulT(transformB.q, d));
        vertex.wB = mul(transformB, proxyB.vertices[vertex.indexB]);
        vertex.w = vertex.wB - vertex.wA;
        ++iter;
        bool duplicate = false;
        for (int i = 0; i < saveCount; ++i) {
            if (vertex.indexA == saveA[i] && vertex.indexB == saveB[i]) {
                duplicate = true;
                break;
            }
        }
        if (duplicate) break;
        ++simplex.count;
    }
    simplex.witnessPoints(output.pointA, output.pointB);
    output.distance = length(output.pointA - output.pointB);
    output.iterations = iter;
    simplex.writeCache(cache);
    if (useRadii) {
        float rA = proxyA.radius, rB = proxyB.radius;
        if (output.distance > rA + rB && output.distance > kEps) {
            output.distance -= rA + rB;
            V2 normal = output.pointB - output.pointA;
            normalize(normal);
            output.pointA += rA * normal;
            output.pointB -= rB * normal;
        } else {
            V2 p = 0.5f * (output.pointA + output.pointB);
            output.pointA = p;
            output.pointB = p;
            output.distance = 0.0f;
        }
    }
}

/* b2TestOverlap of two shapes b2collision.d:449-468: b2Distance with radii from an
 * empty cache, overlapping below 10 epsilon */
static bool testOverlap(const Shape &a, int indexA, const Xf &xfA, const Shape &b, int indexB, const Xf &xfB) {
    DistanceProxy pA, pB;
    pA.set(a, indexA);
    pB.set(b, indexB);
    SimplexCache cache;
    cache.count = 0;
    DistanceOutput out;
    gjkDistance(out, cache, pA, xfA, pB, xfB, true);
    return out.distance < 10.0f * kEps;
}

/* ---- b2TimeOfImpact: conservative advancement on separating axes (b2timeofimpact.d:67-530) ---- */
enum { toi_unknown = 0, toi_failed = 1, toi_overlapped = 2, toi_touching = 3, toi_separated = 4 };
struct SeparationFunction {
    enum { e_points, e_faceA, e_faceB };
    const DistanceProxy *proxyA = nullptr, *proxyB = nullptr;
    Sweep sweepA, sweepB;
    int type = e_points;
    V2 localPoint, axis;
    /* Initialize b2timeofimpact.d:331-410 */
    float initialize(const SimplexCache &cache, const DistanceProxy &pA, const Sweep &sA, const DistanceProxy &pB,
                     const Sweep &sB, float t1) {
        proxyA = &pA;
        proxyB = &pB;
        int count = cache.count;
        sweepA = sA;
        sweepB = sB;
        Xf xfA, xfB;
        sweepA.getTransform(xfA, t1);
        sweepB.getTransform(xfB, t1);
        if (count == 1) {
            type = e_points;
            V2 pointA = mul(xfA, pA.vertices[cache.indexA[0]]);
            V2 pointB = mul(xfB, pB.vertices[cache.indexB[0]]);
            axis = pointB - pointA;
            return normalize(axis);
        } else if (cache.indexA[0] == cache.indexA[1]) {
            type = e_faceB;
            V2 localPointB1 = pB.vertices[cache.indexB[0]];
            V2 localPointB2 = pB.vertices[cache.indexB[1]];
            axis = cross(localPointB2 - localPointB1, 1.0f);
            normalize(axis);
            V2 normal = mul(xfB.q, axis);
            localPoint = 0.5f * (localPointB1 + localPointB2);
            V2 pointB = mul(xfB, localPoint);
            V2 pointA = mul(xfA, pA.vertices[cache.indexA[0]]);
            float s = dot(pointA - pointB, normal);
            if (s < 0.0f) {
                axis = -axis;
                s = -s;
            }
            return s;
        } else {
            type = e_faceA;
            V2 localPointA1 = pA.vertices[cache.indexA[0]];
            V2 localPointA2 = pA.vertices[cache.indexA[1]];
            axis = cross(localPointA2 - localPointA1, 1.0f);
            normalize(axis);
            V2 normal = mul(xfA.q, axis);
            localPoint = 0.5f * (localPointA1 + localPointA2);
            V2 pointA = mul(xfA, localPoint);
            V2 pointB = mul(xfB, pB.vertices[cache.indexB[0]]);
            float s = dot(pointB - pointA, normal);
            if (s < 0.0f) {
                axis = -axis;
                s = -s;
            }
            return s;
        }
    }
    /* FindMinSeparation b2timeofimpact.d:413-470 */
    float findMinSeparation(int &indexA, int &indexB, float t) const {
        Xf xfA, xfB;
        sweepA.getTransform(xfA, t);
        sweepB.getTransform(xfB, t);
        if (type == e_points) {
            V2 axisA = mulT(xfA.q, axis);
            V2 axisB = mulT(xfB.q, -axis);
            indexA = proxyA->support(axisA);
            indexB = proxyB->support(axisB);
            V2 pointA = mul(xfA, proxyA->vertices[indexA]);
            V2 pointB = mul(xfB, proxyB->vertices[indexB]);
            return dot(pointB - pointA, axis);
        } else if (type == e_faceA) {
            V2 normal = mul(xfA.q, axis);
            V2 pointA = mul(xfA, localPoint);
            V2 axisB = mulT(xfB.q, -normal);
            indexA = -1;
            indexB = proxyB->support(axisB);
            V2 pointB = mul(xfB, proxyB->vertices[indexB]);
            return dot(pointB - pointA, normal);
        } else {
            V2 normal = mul(xfB.q, axis);
            V2 pointB = mul(xfB, localPoint);
            V2 axisA = mulT(xfA.q, -normal);
            indexB = -1;
            indexA = proxyA->support(axisA);
            V2 pointA = mul(xfA, proxyA->vertices[indexA]);
            return dot(pointA - pointB, normal);
        }
    }
    /* Evaluate b2timeofimpact.d:473-518 */
    float evaluate(int indexA, int indexB, float t) const {
        Xf xfA, xfB;
        sweepA.getTransform(xfA, t);
        sweepB.getTransform(xfB, t);
        if (type == e_points) {
            V2 pointA = mul(xfA, proxyA->vertices[indexA]);
            V2 pointB = mul(xfB, proxyB->vertices[indexB]);
            return dot(pointB - pointA, axis);
        } else if (type == e_faceA) {
            V2 normal = mul(xfA.q, axis);
            V2 pointA = mul(xfA, localPoint);
            V2 pointB = mul(xfB, proxyB->vertices[indexB]);
            return dot(pointB - pointA, normal);
        } else {
            V2 normal = mul(xfB.q, axis);
            V2 pointB = mul(xfB, localPoint);
            V2 pointA = mul(xfA, proxyA->vertices[indexA]);
            return dot(pointA - pointB, normal);
        }
    }
};
/* b2TimeOfImpact b2timeofimpact.d:67-295; returns the state, t in [0, tMax] */
static int timeOfImpact(float &tOut, const DistanceProxy &proxyA, const DistanceProxy &proxyB, Sweep sweepA, Sweep sweepB,
                        float tMax) {
    int state = toi_unknown;
    tOut = tMax;
    sweepA.normalize();
    sweepB.normalize();
    float totalRadius = proxyA.radius + proxyB.radius;
    float target = fmax_(kLinearSlop, totalRadius - 3.0f * kLinearSlop);
    float tolerance = 0.25f * kLinearSlop;
    float t1 = 0.0f;
    const int k_maxIterations = 20;
    int iter = 0;
    SimplexCache cache;
    cache.count = 0;
    for (;;) {
        Xf xfA, xfB;
        sweepA.getTransform(xfA, t1);
        sweepB.getTransform(xfB, t1);
        DistanceOutput distanceOutput;
        gjkDistance(distanceOutput, cache, proxyA, xfA, proxyB, xfB, false);
        if (distanceOutput.distance <= 0.0f) {
            state = toi_overlapped;
            tOut = 0.0f;
            break;
        }
        if (distanceOutput.distance < target + tolerance) {
            state = toi_touching;
            tOut = t1;
            break;
        }
        SeparationFunction fcn;
        fcn.initialize(cache, proxyA, sweepA, proxyB, sweepB, t1);
        bool done = false;
        float t2 = tMax;
        int pushBackIter = 0;
        for (;;) {
            int indexA, indexB;
            float s2 = fcn.findMinSeparation(indexA, indexB, t2);
            if (s2 > target + tolerance) {
                state = toi_separated;
                tOut = tMax;
                done = true;
                break;
            }
            if (s2 > target - tolerance) {
                t1 = t2;
                break;
            }
            float s1 = fcn.evaluate(indexA, indexB, t1);
            if (s1 < target - tolerance) {
                state = toi_failed;
                tOut = t1;
                done = true;
                break;
            }
            if (s1 <= target + tolerance) {
                state = toi_touching;
                tOut = t1;
                done = true;
                break;
            }
            int rootIterCount = 0;
            float a1 = t1, a2 = t2;
            for (;;) {
                float t;
                if (rootIterCount & 1) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
                else t = 0.5f * (a1 + a2);
                ++rootIterCount;
                float s = fcn.evaluate(indexA, indexB, t);
                if (fabs_(s - target) < tolerance) {
                    t2 = t;
                    break;
                }
                if (s > target) {
                    a1 = t;
                    s1 = s;
                } else {
                    a2 = t;
                    s2 = s;
                }
                if (rootIterCount == 50) break;
            }
            ++pushBackIter;
            if (pushBackIter == kMaxPolygonVertices) break;
        }
        ++iter;
        if (done) break;
        if (iter == k_maxIterations) {
            state = toi_failed;
            tOut = t1;
            break;
        }
    }
    return state;
}
struct Body;
struct Fixture {
    int index = 0;
    int body = 0;
    Shape shape;
    float density = 0, friction = 0, restitution = 0;
    uint16_t category = 1, mask = 0xFFFF;
    int16_t group = 0;
    bool sensor = false;
    std::vector<int> proxyIds;  /* m_proxies: one tree proxy per child */
    std::vector<AABB> proxyAabbs;
    int proxyBase = 0;          /* canonical index of child 0: children counted in creation order */
};
struct Body {
    int type = b_static;
    uint16_t flags = 0;
    int islandIndex = 0;
    Xf xf;
    Sweep sweep;
    V2 v;
    float w = 0;
    V2 force;
    float torque = 0;
    float mass = 0, invMass = 0, I = 0, invI = 0;
    float linearDamping = 0, angularDamping = 0, gravityScale = 1;
    float sleepTime = 0;
    std::vector<int> fixtures; /* m_fixtureList: most recently created first */
    int contactList = kNull;   /* head edge: contact index * 2 + (0 = nodeA, 1 = nodeB) */
    int jointList = kNull;     /* head joint edge: joint index * 2 + (0 = edgeA, 1 = edgeB) */
    int prev = kNull, next = kNull;
    bool awake() const { return (flags & bf_awake) != 0; }
    /* b2Body.SetAwake b2body.d:827-846 */
    void setAwake(bool flag) {
        if (flag) {
            if ((flags & bf_awake) == 0) { flags |= bf_awake; sleepTime = 0.0f; }
        } else {
            flags &= ~bf_awake;
            sleepTime = 0.0f;
            v = mk(0.0f, 0.0f);
            w = 0.0f;
            force = mk(0.0f, 0.0f);
            torque = 0.0f;
        }
    }
    void synchronizeTransform() {
        xf.q.set(sweep.a);
        xf.p = sweep.c - mul(xf.q, sweep.localCenter);
    }
    /* b2Body.Advance b2body.d:1172-1180 */
    void advance(float alpha) {
        sweep.advance(alpha);
        sweep.c = sweep.c0;
        sweep.a = sweep.a0;
        xf.q.set(sweep.a);
        xf.p = sweep.c - mul(xf.q, sweep.localCenter);
    }
};
struct ContactEdge {
    int other = kNull;
    int prev = kNull, next = kNull; /* encoded edges (contact * 2 + side) */
};
struct Contact {
    uint32_t flags = 0;
    int prev = kNull, next = kNull;
    ContactEdge node[2];
    int fixtureA = 0, fixtureB = 0;
    int indexA = 0, indexB = 0; /* child indices (chains) */
    Manifold manifold;
    float friction = 0, restitution = 0, tangentSpeed = 0;
    bool alive = false;
    int toiCount = 0;
    float toi = 0;
};

struct VelocityConstraintPoint {
    V2 rA, rB;
    float normalImpulse = 0, tangentImpulse = 0, normalMass = 0, tangentMass = 0, velocityBias = 0;
};
struct VelocityConstraint {
    VelocityConstraintPoint points[2];
    V2 normal;
    Mat22 normalMass, K;
    int indexA = 0, indexB = 0;
    float invMassA = 0, invMassB = 0, invIA = 0, invIB = 0;
    float friction = 0, restitution = 0, tangentSpeed = 0;
    int pointCount = 0, contactIndex = 0;
};
struct PositionConstraint {
    V2 localPoints[2];
    V2 localNormal, localPoint;
    int indexA = 0, indexB = 0;
    float invMassA = 0, invMassB = 0;
    V2 localCenterA, localCenterB;
    float invIA = 0, invIB = 0;
    int type = 0;
    float radiusA = 0, radiusB = 0;
    int pointCount = 0;
};
struct Position { V2 c; float a = 0; };
struct Velocity { V2 v; float w = 0; };

struct TimeStep {
    float dt = 0, inv_dt = 0, dtRatio = 0;
    int velocityIterations = 0, positionIterations = 0;
    bool warmStarting = true;
};

#include "b2oracle_joints.h"

using Clock = std::chrono::steady_clock;
static inline float msSince(Clock::time_point t0) {
    return std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

struct World {
    V2 gravity;
    std::vector<Body> bodies;
    std::vector<Fixture> fixtures;
    std::vector<Contact> contacts;
    std::vector<int> freeContacts;
    int bodyList = kNull;
    int contactList = kNull;
    int contactCount = 0;
    BroadPhase bp;
    bool warmStarting = true, continuousPhysics = false, allowSleep = true;
    bool newFixture = false;
    bool stepComplete = true;
    int toiEvents = 0; /* test hook: TOI events solved by the last step */
    /* test hook, standing in for a b2ContactListener.PreSolve: 0 none, 1 record the
     * calls, 2 record them and apply presolvePolicy */
    int presolveMode = 0;
    int rotMode = -1; /* test switch: -1 = leave g_rotMode as it is, else g_rotMode during step() */
    float inv_dt0 = 0.0f;
    int solveOrder = 0;
    int islandCount = 0; /* test hook: 1 = solve island contacts in greedy-colour order */
    float profile[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<Pair> lastPairs; /* fixture-index pairs handed to AddPair */
    std::vector<int> solveLog;   /* test hook: per solved contact (island, bodyA, bodyB, dynamic mask) */
    std::vector<int> canonFix, canonChild; /* per canonical proxy index: fixture, child */
    int proxyOf(int fixture, int child) const { return fixtures[fixture].proxyBase + child; }

    /* island scratch */
    std::vector<int> islandBodies, islandContacts, islandJoints, stack;
    std::vector<Joint> joints;
    int jointList = kNull;
    JointEdge &jedge(int e) { return joints[e >> 1].node[e & 1]; }

    /* b2World.CreateJoint b2world.d:196-262 (no contact exists to flag for filtering
     * before the first step) */
    int createJoint(const orc_joint_def &d) {
        const int nb = (int)bodies.size();
        if (d.body_a < 0 || d.body_a >= nb || d.body_b < 0 || d.body_b >= nb || d.body_a == d.body_b) return -1;
        if (d.type < j_revolute || d.type > j_motor) return -1;
        Joint j;
        if (d.type == j_gear) { /* b2GearJointDef: two revolute or prismatic joints made before it */
            const int nj = (int)joints.size();
            if (d.joint1 < 0 || d.joint1 >= nj || d.joint2 < 0 || d.joint2 >= nj) return -1;
            const Joint &j1 = joints[d.joint1], &j2 = joints[d.joint2];
            if ((j1.type != j_revolute && j1.type != j_prismatic) || (j2.type != j_revolute && j2.type != j_prismatic))
                return -1;
            gearFromDef(j, d, j1, j2, bodies.data());
        } else {
            jointFromDef(j, d, bodies[d.body_b].xf);
        }
        const int ji = (int)joints.size();
        j.prev = kNull;
        j.next = jointList;
        if (jointList != kNull) joints[jointList].prev = ji;
        jointList = ji;
        joints.push_back(j);
        const int ja = joints[ji].bodyA, jb = joints[ji].bodyB; /* a gear's are its joints' bodies B */
        for (int side = 0; side < 2; ++side) {
            Body &b = bodies[side == 0 ? ja : jb];
            JointEdge &e = joints[ji].node[side];
            e.other = side == 0 ? jb : ja;
            e.prev = kNull;
            e.next = b.jointList;
            if (b.jointList != kNull) jedge(b.jointList).prev = 2 * ji + side;
            b.jointList = 2 * ji + side;
        }
        return ji;
    }
    std::vector<Position> positions;
    std::vector<Velocity> velocities;
    std::vector<VelocityConstraint> vcs;
    std::vector<PositionConstraint> pcs;

    static int edgeContact(int e) { return e >> 1; }
    ContactEdge &edge(int e) { return contacts[e >> 1].node[e & 1]; }

    /* b2World.CreateBody b2world.d:75-99 + b2Body ctor b2body.d:1030-1115 */
    int createBody(const orc_body_def &d) {
        Body b;
        b.flags = 0;
        if (d.bullet) b.flags |= bf_bullet;
        if (d.fixed_rotation) b.flags |= bf_fixedRotation;
        if (d.allow_sleep) b.flags |= bf_autoSleep;
        if (d.awake) b.flags |= bf_awake;
        b.flags |= bf_active;
        b.xf.p = mk(d.px, d.py);
        b.xf.q.set(d.angle);
        b.sweep.localCenter = mk(0.0f, 0.0f);
        b.sweep.c0 = b.xf.p;
        b.sweep.c = b.xf.p;
        b.sweep.a0 = d.angle;
        b.sweep.a = d.angle;
        b.sweep.alpha0 = 0.0f;
        b.v = mk(d.vx, d.vy);
        b.w = d.w;
        b.linearDamping = d.linear_damping;
        b.angularDamping = d.angular_damping;
        b.gravityScale = d.gravity_scale;
        b.force = mk(0.0f, 0.0f);
        b.torque = 0.0f;
        b.sleepTime = 0.0f;
        b.type = d.type;
        if (b.type == b_dynamic) { b.mass = 1.0f; b.invMass = 1.0f; }
        else { b.mass = 0.0f; b.invMass = 0.0f; }
        b.I = 0.0f;
        b.invI = 0.0f;
        int id = (int)bodies.size();
        b.prev = kNull;
        b.next = bodyList;
        if (bodyList != kNull) bodies[bodyList].prev = id;
        bodies.push_back(b);
        bodyList = id;
        return id;
    }
    /* b2Body.ResetMassData b2body.d:555-625 */
    void resetMassData(Body &b) {
        b.mass = 0.0f;
        b.invMass = 0.0f;
        b.I = 0.0f;
        b.invI = 0.0f;
        b.sweep.localCenter = mk(0.0f, 0.0f);
        if (b.type == b_static || b.type == b_kinematic) {
            b.sweep.c0 = b.xf.p;
            b.sweep.c = b.xf.p;
            b.sweep.a0 = b.sweep.a;
            return;
        }
        V2 localCenter = mk(0.0f, 0.0f);
        for (int fi : b.fixtures) {
            const Fixture &f = fixtures[fi];
            if (f.density == 0.0f) continue;
            MassData md;
            computeMass(f.shape, md, f.density);
            b.mass += md.mass;
            localCenter += md.mass * md.center;
            b.I += md.I;
        }
        if (b.mass > 0.0f) {
            b.invMass = 1.0f / b.mass;
            localCenter *= b.invMass;
        } else {
            b.mass = 1.0f;
            b.invMass = 1.0f;
        }
        if (b.I > 0.0f && (b.flags & bf_fixedRotation) == 0) {
            b.I -= b.mass * dot(localCenter, localCenter);
            b.invI = 1.0f / b.I;
        } else {
            b.I = 0.0f;
            b.invI = 0.0f;
        }
        V2 oldCenter = b.sweep.c;
        b.sweep.localCenter = localCenter;
        b.sweep.c = mul(b.xf, b.sweep.localCenter);
        b.sweep.c0 = b.sweep.c;
        b.v += cross(b.w, b.sweep.c - oldCenter);
    }
    /* b2Body.CreateFixture b2body.d:116-155 + b2Fixture.Create/CreateProxies b2fixture.d:367-465 */
    int createFixture(const orc_fixture_def &d) {
        Fixture f;
        f.index = (int)fixtures.size();
        f.body = d.body;
        f.density = d.density;
        f.friction = d.friction;
        f.restitution = d.restitution;
        f.category = d.category;
        f.mask = d.mask;
        f.group = d.group;
        f.sensor = (d.flags & 1) != 0;
        buildShape(d, f.shape);
        Body &b = bodies[d.body];
        /* b2Fixture.CreateProxies (b2fixture.d:414-431); a proxy's user data is its
         * canonical index, which orders like the tree's ids (allocated in the same order) */
        f.proxyBase = (int)canonFix.size();
        const int kids = childCount(f.shape);
        for (int i = 0; i < kids; ++i) {
            AABB a;
            computeAABB(f.shape, a, b.xf, i);
            f.proxyAabbs.push_back(a);
            f.proxyIds.push_back(bp.createProxy(a, f.proxyBase + i));
            canonFix.push_back(f.index);
            canonChild.push_back(i);
        }
        fixtures.push_back(f);
        b.fixtures.insert(b.fixtures.begin(), f.index);
        if (f.density > 0.0f) resetMassData(b);
        newFixture = true;
        return f.index;
    }
    static void buildShape(const orc_fixture_def &d, Shape &s) {
        if (d.shape == 0) {
            s.type = s_circle;
            s.radius = d.radius;
            s.p = mk(d.p[0], d.p[1]);
        } else if (d.shape == 1) {
            s.type = s_edge;
            s.radius = kPolygonRadius;
            s.v1 = mk(d.p[0], d.p[1]);
            s.v2 = mk(d.p[2], d.p[3]);
            s.v0 = mk(d.p[4], d.p[5]);
            s.v3 = mk(d.p[6], d.p[7]);
            s.has0 = (d.flags & 2) != 0;
            s.has3 = (d.flags & 4) != 0;
            if (!s.has0) s.v0 = mk(0.0f, 0.0f);
            if (!s.has3) s.v3 = mk(0.0f, 0.0f);
        } else if (d.shape == 2) {
            V2 pts[kMaxPolygonVertices];
            int n = d.count < kMaxPolygonVertices ? d.count : kMaxPolygonVertices;
            for (int i = 0; i < n; ++i) pts[i] = mk(d.p[2 * i], d.p[2 * i + 1]);
            polySet(s, pts, n);
        } else if (d.shape == 4) { /* b2ChainShape.CreateLoop / CreateChain (b2chainshape.d:46-100) */
            s.type = s_chain;
            s.radius = kPolygonRadius;
            const int n = d.count;
            for (int i = 0; i < n; ++i) s.cv[i] = mk(d.p[2 * i], d.p[2 * i + 1]);
            if (d.flags & 16) {
                s.ccount = n + 1;
                s.cv[n] = s.cv[0];
                s.cprev = s.cv[s.ccount - 2];
                s.cnext = s.cv[1];
                s.chasPrev = s.chasNext = true;
            } else {
                s.ccount = n;
                if (d.flags & 2) { s.cprev = mk(d.p[16], d.p[17]); s.chasPrev = true; } /* SetPrevVertex */
                if (d.flags & 4) { s.cnext = mk(d.p[18], d.p[19]); s.chasNext = true; } /* SetNextVertex */
            }
        } else {
            if (d.flags & 8) polySetAsBoxOriented(s, d.p[0], d.p[1], mk(d.p[2], d.p[3]), d.p[4]);
            else polySetAsBox(s, d.p[0], d.p[1]);
        }
    }

    /* b2Body.ShouldCollide b2body.d:1149-1170 */
    bool bodyShouldCollide(int ai, int bi) {
        const Body &a = bodies[ai], &b = bodies[bi];
        if (a.type != b_dynamic && b.type != b_dynamic) return false;
        for (int e = a.jointList; e != kNull; e = jedge(e).next)
            if (jedge(e).other == bi && joints[e >> 1].collideConnected == false) return false;
        return true;
    }
    /* b2ContactFilter.ShouldCollide b2worldcallbacks.d:52-65 */
    static bool filterShouldCollide(const Fixture &fa, const Fixture &fb) {
        if (fa.group == fb.group && fa.group != 0) return fa.group > 0;
        return (fa.mask & fb.category) != 0 && (fa.category & fb.mask) != 0;
    }

    /* b2ContactManager.AddPair b2contactmanager.d:52-176 */
    /* b2ContactManager.AddPair (b2contactmanager.d:52-176) for two canonical proxies */
    void addPair(int proxyA, int proxyB) {
        int fixA = canonFix[proxyA], fixB = canonFix[proxyB];
        int indexA = canonChild[proxyA], indexB = canonChild[proxyB];
        Fixture *fixtureA = &fixtures[fixA];
        Fixture *fixtureB = &fixtures[fixB];
        int bodyA = fixtureA->body, bodyB = fixtureB->body;
        if (bodyA == bodyB) return;
        for (int e = bodies[bodyB].contactList; e != kNull; e = edge(e).next) {
            if (edge(e).other == bodyA) {
                const Contact &c = contacts[edgeContact(e)];
                if (c.fixtureA == fixA && c.fixtureB == fixB && c.indexA == indexA && c.indexB == indexB) return;
                if (c.fixtureA == fixB && c.fixtureB == fixA && c.indexA == indexB && c.indexB == indexA) return;
            }
        }
        if (!bodyShouldCollide(bodyB, bodyA)) return;
        if (!filterShouldCollide(*fixtureA, *fixtureB)) return;
        int ta = fixtureA->shape.type, tb = fixtureB->shape.type;
        if (!hasContactType(ta, tb)) return;
        if (!isPrimary(ta, tb)) {
            std::swap(fixA, fixB);
            std::swap(fixtureA, fixtureB);
            std::swap(indexA, indexB);
        }
        int ci;
        if (!freeContacts.empty()) { ci = freeContacts.back(); freeContacts.pop_back(); }
        else { ci = (int)contacts.size(); contacts.emplace_back(); }
        Contact &c = contacts[ci];
        c = Contact();
        c.alive = true;
        c.flags = cf_enabled;
        c.fixtureA = fixA;
        c.fixtureB = fixB;
        c.indexA = indexA;
        c.indexB = indexB;
        c.manifold.pointCount = 0;
        c.friction = sqrtf(fixtureA->friction * fixtureB->friction);   /* b2MixFriction */
        c.restitution = fixtureA->restitution > fixtureB->restitution ? fixtureA->restitution : fixtureB->restitution;
        c.tangentSpeed = 0.0f;
        bodyA = fixtureA->body;
        bodyB = fixtureB->body;
        c.prev = kNull;
        c.next = contactList;
        if (contactList != kNull) contacts[contactList].prev = ci;
        contactList = ci;
        int eA = ci * 2 + 0, eB = ci * 2 + 1;
        c.node[0].other = bodyB;
        c.node[0].prev = kNull;
        c.node[0].next = bodies[bodyA].contactList;
        if (bodies[bodyA].contactList != kNull) edge(bodies[bodyA].contactList).prev = eA;
        bodies[bodyA].contactList = eA;
        c.node[1].other = bodyA;
        c.node[1].prev = kNull;
        c.node[1].next = bodies[bodyB].contactList;
        if (bodies[bodyB].contactList != kNull) edge(bodies[bodyB].contactList).prev = eB;
        bodies[bodyB].contactList = eB;
        if (!fixtureA->sensor && !fixtureB->sensor) {
            bodies[bodyA].setAwake(true);
            bodies[bodyB].setAwake(true);
        }
        ++contactCount;
    }
    /* b2ContactManager.Destroy b2contactmanager.d:183-246 + b2Contact.Destroy b2contact.d:402-423 */
    void destroyContact(int ci) {
        Contact &c = contacts[ci];
        int bodyA = fixtures[c.fixtureA].body, bodyB = fixtures[c.fixtureB].body;
        if (c.flags & cf_touching) recordEvent(2, c, nullptr); /* b2contactmanager.d:190-193 */
        if (c.prev != kNull) contacts[c.prev].next = c.next;
        if (c.next != kNull) contacts[c.next].prev = c.prev;
        if (ci == contactList) contactList = c.next;
        for (int side = 0; side < 2; ++side) {
            int body = side == 0 ? bodyA : bodyB;
            ContactEdge &n = c.node[side];
            if (n.prev != kNull) edge(n.prev).next = n.next;
            if (n.next != kNull) edge(n.next).prev = n.prev;
            if (bodies[body].contactList == ci * 2 + side) bodies[body].contactList = n.next;
        }
        if (c.manifold.pointCount > 0 && !fixtures[c.fixtureA].sensor && !fixtures[c.fixtureB].sensor) {
            bodies[bodyA].setAwake(true);
            bodies[bodyB].setAwake(true);
        }
        c.alive = false;
        freeContacts.push_back(ci);
        --contactCount;
    }
    void findNewContacts() {
        lastPairs.clear();
        bp.updatePairs([&](int fa, int fb, const Pair &) {
            Pair p;
            p.a = fa;
            p.b = fb;
            lastPairs.push_back(p);
            addPair(fa, fb);
        });
    }
    /* b2Contact.Update b2contact.d:270-356 with the test listener */
    void updateContact(Contact &c) {
        Manifold oldManifold = c.manifold;
        c.flags |= cf_enabled;
        bool wasTouching = (c.flags & cf_touching) == cf_touching;
        Body &bodyA = bodies[fixtures[c.fixtureA].body];
        Body &bodyB = bodies[fixtures[c.fixtureB].body];
        if (fixtures[c.fixtureA].sensor || fixtures[c.fixtureB].sensor) {
            /* a sensor: shape overlap by b2TestOverlap, no manifold, no wake, no
             * PreSolve (b2contact.d:289-298,352) */
            const bool touching = testOverlap(fixtures[c.fixtureA].shape, c.indexA, bodyA.xf,
                                              fixtures[c.fixtureB].shape, c.indexB, bodyB.xf);
            c.manifold.pointCount = 0;
            if (touching) c.flags |= cf_touching;
            else c.flags &= ~cf_touching;
            if (!wasTouching && touching) recordEvent(1, c, nullptr);
            if (wasTouching && !touching) recordEvent(2, c, nullptr);
            return;
        }
        evaluate(c.manifold, fixtures[c.fixtureA].shape, bodyA.xf, fixtures[c.fixtureB].shape, bodyB.xf, c.indexA);
        bool touching = c.manifold.pointCount > 0;
        for (int i = 0; i < c.manifold.pointCount; ++i) {
            ManifoldPoint &mp2 = c.manifold.points[i];
            mp2.normalImpulse = 0.0f;
            mp2.tangentImpulse = 0.0f;
            for (int j = 0; j < oldManifold.pointCount; ++j) {
                if (oldManifold.points[j].id == mp2.id) {
                    mp2.normalImpulse = oldManifold.points[j].normalImpulse;
                    mp2.tangentImpulse = oldManifold.points[j].tangentImpulse;
                    break;
                }
            }
        }
        if (touching != wasTouching) {
            bodyA.setAwake(true);
            bodyB.setAwake(true);
        }
        if (touching) c.flags |= cf_touching;
        else c.flags &= ~cf_touching;
        /* listener calls b2contact.d:342-355 */
        if (!wasTouching && touching) recordEvent(1, c, nullptr);
        if (wasTouching && !touching) recordEvent(2, c, nullptr);
        if (touching && presolveMode) {
            orc_event e = {};
            e.type = 4;
            e.fixture_a = proxyOf(c.fixtureA, c.indexA);
            e.fixture_b = proxyOf(c.fixtureB, c.indexB);
            e.point_count = oldManifold.pointCount;
            for (int j = 0; j < oldManifold.pointCount; ++j) {
                e.normal_impulse[j] = oldManifold.points[j].normalImpulse;
                e.tangent_impulse[j] = oldManifold.points[j].tangentImpulse;
            }
            events.push_back(e);
            if (presolveMode == 2) presolvePolicy(c);
        }
    }
    /* The test listener's PreSolve: by k = fixtureA + fixtureB, k % 7 == 0 disables
     * the contact, 1 sets friction 0.05, 2 restitution 0.6, 3 tangent speed 0.8
     * (b2Contact.SetEnabled / SetFriction / SetRestitution / SetTangentSpeed,
     * b2contact.d:108-198). */
    static void presolvePolicy(Contact &c) {
        const int k = (c.fixtureA + c.fixtureB) % 7;
        if (k == 0) c.flags &= ~cf_enabled;
        else if (k == 1) c.friction = 0.05f;
        else if (k == 2) c.restitution = 0.6f;
        else if (k == 3) c.tangentSpeed = 0.8f;
    }
    /* b2ContactManager.Collide b2contactmanager.d:251-317 */
    void collide() {
        int ci = contactList;
        while (ci != kNull) {
            Contact &c = contacts[ci];
            Body &bodyA = bodies[fixtures[c.fixtureA].body];
            Body &bodyB = bodies[fixtures[c.fixtureB].body];
            bool activeA = bodyA.awake() && bodyA.type != b_static;
            bool activeB = bodyB.awake() && bodyB.type != b_static;
            if (!activeA && !activeB) { ci = c.next; continue; }
            bool overlap = bp.testOverlapProxies(fixtures[c.fixtureA].proxyIds[c.indexA],
                                                 fixtures[c.fixtureB].proxyIds[c.indexB]);
            if (!overlap) {
                int nuke = ci;
                ci = c.next;
                destroyContact(nuke);
                continue;
            }
            updateContact(c);
            ci = c.next;
        }
    }
    /* b2Body.SynchronizeFixtures b2body.d:1129-1141 + b2Fixture.Synchronize b2fixture.d:480-502 */
    void synchronizeFixtures(Body &b) {
        Xf xf1;
        xf1.q.set(b.sweep.a0);
        xf1.p = b.sweep.c0 - mul(xf1.q, b.sweep.localCenter);
        for (int fi : b.fixtures) {
            Fixture &f = fixtures[fi];
            for (size_t i = 0; i < f.proxyIds.size(); ++i) {
                AABB aabb1, aabb2;
                computeAABB(f.shape, aabb1, xf1, (int)i);
                computeAABB(f.shape, aabb2, b.xf, (int)i);
                f.proxyAabbs[i].combine(aabb1, aabb2);
                V2 displacement = b.xf.p - xf1.p;
                bp.moveProxy(f.proxyIds[i], f.proxyAabbs[i], displacement);
            }
        }
    }

    /* ---- b2ContactSolver b2contactsolver.d ---- */
    void solverCreate(const TimeStep &step) {
        int n = (int)islandContacts.size();
        vcs.assign(n, VelocityConstraint());
        pcs.assign(n, PositionConstraint());
        for (int i = 0; i < n; ++i) {
            Contact &contact = contacts[islandContacts[i]];
            const Fixture &fa = fixtures[contact.fixtureA];
            const Fixture &fb = fixtures[contact.fixtureB];
            const Body &bodyA = bodies[fa.body];
            const Body &bodyB = bodies[fb.body];
            const Manifold &manifold = contact.manifold;
            int pointCount = manifold.pointCount;
            VelocityConstraint &vc = vcs[i];
            vc.friction = contact.friction;
            vc.restitution = contact.restitution;
            vc.tangentSpeed = contact.tangentSpeed;
            vc.indexA = bodyA.islandIndex;
            vc.indexB = bodyB.islandIndex;
            vc.invMassA = bodyA.invMass;
            vc.invMassB = bodyB.invMass;
            vc.invIA = bodyA.invI;
            vc.invIB = bodyB.invI;
            vc.contactIndex = i;
            vc.pointCount = pointCount;
            PositionConstraint &pc = pcs[i];
            pc.indexA = bodyA.islandIndex;
            pc.indexB = bodyB.islandIndex;
            pc.invMassA = bodyA.invMass;
            pc.invMassB = bodyB.invMass;
            pc.localCenterA = bodyA.sweep.localCenter;
            pc.localCenterB = bodyB.sweep.localCenter;
            pc.invIA = bodyA.invI;
            pc.invIB = bodyB.invI;
            pc.localNormal = manifold.localNormal;
            pc.localPoint = manifold.localPoint;
            pc.pointCount = pointCount;
            pc.radiusA = fa.shape.radius;
            pc.radiusB = fb.shape.radius;
            pc.type = manifold.type;
            for (int j = 0; j < pointCount; ++j) {
                const ManifoldPoint &cp = manifold.points[j];
                VelocityConstraintPoint &vcp = vc.points[j];
                if (step.warmStarting) {
                    vcp.normalImpulse = step.dtRatio * cp.normalImpulse;
                    vcp.tangentImpulse = step.dtRatio * cp.tangentImpulse;
                } else {
                    vcp.normalImpulse = 0.0f;
                    vcp.tangentImpulse = 0.0f;
                }
                pc.localPoints[j] = cp.localPoint;
            }
        }
    }
    void initializeVelocityConstraints() {
        for (size_t i = 0; i < vcs.size(); ++i) {
            VelocityConstraint &vc = vcs[i];
            PositionConstraint &pc = pcs[i];
            float radiusA = pc.radiusA, radiusB = pc.radiusB;
            const Manifold &manifold = contacts[islandContacts[vc.contactIndex]].manifold;
            int indexA = vc.indexA, indexB = vc.indexB;
            V2 localCenterA = pc.localCenterA, localCenterB = pc.localCenterB;
            V2 cA = positions[indexA].c;
            float aA = positions[indexA].a;
            V2 vA = velocities[indexA].v;
            float wA = velocities[indexA].w;
            V2 cB = positions[indexB].c;
            float aB = positions[indexB].a;
            V2 vB = velocities[indexB].v;
            float wB = velocities[indexB].w;
            Xf xfA, xfB;
            xfA.q.set(aA);
            xfB.q.set(aB);
            xfA.p = cA - mul(xfA.q, localCenterA);
            xfB.p = cB - mul(xfB.q, localCenterB);
            WorldManifold wm;
            wm.initialize(manifold, xfA, radiusA, xfB, radiusB);
            vc.normal = wm.normal;
            int pointCount = vc.pointCount;
            for (int j = 0; j < pointCount; ++j) {
                VelocityConstraintPoint &vcp = vc.points[j];
                vcp.rA = wm.points[j] - cA;
                vcp.rB = wm.points[j] - cB;
                pointMasses(vc, vcp);
                vcp.velocityBias = 0.0f;
                float vRel = dot(vc.normal, vB + cross(wB, vcp.rB) - vA - cross(wA, vcp.rA));
                if (vRel < -kVelocityThreshold) vcp.velocityBias = -vc.restitution * vRel;
            }
            blockMasses(vc);
        }
    }
    /* normal and tangent mass of one point (b2contactsolver.d:384-402) */
    static void pointMasses(const VelocityConstraint &vc, VelocityConstraintPoint &vcp) {
        float mA = vc.invMassA, mB = vc.invMassB, iA = vc.invIA, iB = vc.invIB;
        float rnA = cross(vcp.rA, vc.normal);
        float rnB = cross(vcp.rB, vc.normal);
        float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
        vcp.normalMass = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
        V2 tangent = cross(vc.normal, 1.0f);
        float rtA = cross(vcp.rA, tangent);
        float rtB = cross(vcp.rB, tangent);
        float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
        vcp.tangentMass = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
    }
    /* the two-point block K, its inverse, or the downgrade to one point (b2contactsolver.d:418-447) */
    static void blockMasses(VelocityConstraint &vc) {
        if (vc.pointCount != 2) return; /* g_blockSolve == true (b2contactsolver.d:799) */
        float mA = vc.invMassA, mB = vc.invMassB, iA = vc.invIA, iB = vc.invIB;
        VelocityConstraintPoint &vcp1 = vc.points[0];
        VelocityConstraintPoint &vcp2 = vc.points[1];
        float rn1A = cross(vcp1.rA, vc.normal);
        float rn1B = cross(vcp1.rB, vc.normal);
        float rn2A = cross(vcp2.rA, vc.normal);
        float rn2B = cross(vcp2.rB, vc.normal);
        float k11 = mA + mB + iA * rn1A * rn1A + iB * rn1B * rn1B;
        float k22 = mA + mB + iA * rn2A * rn2A + iB * rn2B * rn2B;
        float k12 = mA + mB + iA * rn1A * rn2A + iB * rn1B * rn2B;
        const float k_maxConditionNumber = 1000.0f;
        if (k11 * k11 < k_maxConditionNumber * (k11 * k22 - k12 * k12)) {
            vc.K.ex = mk(k11, k12);
            vc.K.ey = mk(k12, k22);
            vc.normalMass = vc.K.inverse();
        } else {
            vc.pointCount = 1;
        }
    }
    void warmStart() {
        for (size_t i = 0; i < vcs.size(); ++i) {
            VelocityConstraint &vc = vcs[i];
            int indexA = vc.indexA, indexB = vc.indexB;
            float mA = vc.invMassA, iA = vc.invIA, mB = vc.invMassB, iB = vc.invIB;
            V2 vA = velocities[indexA].v;
            float wA = velocities[indexA].w;
            V2 vB = velocities[indexB].v;
            float wB = velocities[indexB].w;
            V2 normal = vc.normal;
            V2 tangent = cross(normal, 1.0f);
            for (int j = 0; j < vc.pointCount; ++j) {
                VelocityConstraintPoint &vcp = vc.points[j];
                V2 P = vcp.normalImpulse * normal + vcp.tangentImpulse * tangent;
                wA -= iA * cross(vcp.rA, P);
                vA -= mA * P;
                wB += iB * cross(vcp.rB, P);
                vB += mB * P;
            }
            velocities[indexA].v = vA;
            velocities[indexA].w = wA;
            velocities[indexB].v = vB;
            velocities[indexB].w = wB;
        }
    }
    void solveVelocityConstraints() {
        for (size_t i = 0; i < vcs.size(); ++i) {
            VelocityConstraint &vc = vcs[i];
            int indexA = vc.indexA, indexB = vc.indexB;
            float mA = vc.invMassA, iA = vc.invIA, mB = vc.invMassB, iB = vc.invIB;
            int pointCount = vc.pointCount;
            V2 vA = velocities[indexA].v;
            float wA = velocities[indexA].w;
            V2 vB = velocities[indexB].v;
            float wB = velocities[indexB].w;
            V2 normal = vc.normal;
            V2 tangent = cross(normal, 1.0f);
            float friction = vc.friction;
            for (int j = 0; j < pointCount; ++j) {
                VelocityConstraintPoint &vcp = vc.points[j];
                V2 dv = vB + cross(wB, vcp.rB) - vA - cross(wA, vcp.rA);
                float vt = dot(dv, tangent) - vc.tangentSpeed;
                float lambda = vcp.tangentMass * (-vt);
                float maxFriction = friction * vcp.normalImpulse;
                float newImpulse = clampf(vcp.tangentImpulse + lambda, -maxFriction, maxFriction);
                lambda = newImpulse - vcp.tangentImpulse;
                vcp.tangentImpulse = newImpulse;
                V2 P = lambda * tangent;
                vA -= mA * P;
                wA -= iA * cross(vcp.rA, P);
                vB += mB * P;
                wB += iB * cross(vcp.rB, P);
            }
            if (pointCount == 1) {
                VelocityConstraintPoint &vcp = vc.points[0];
                V2 dv = vB + cross(wB, vcp.rB) - vA - cross(wA, vcp.rA);
                float vn = dot(dv, normal);
                float lambda = -vcp.normalMass * (vn - vcp.velocityBias);
                float newImpulse = fmax_(vcp.normalImpulse + lambda, 0.0f);
                lambda = newImpulse - vcp.normalImpulse;
                vcp.normalImpulse = newImpulse;
                V2 P = lambda * normal;
                vA -= mA * P;
                wA -= iA * cross(vcp.rA, P);
                vB += mB * P;
                wB += iB * cross(vcp.rB, P);
            } else {
                VelocityConstraintPoint &cp1 = vc.points[0];
                VelocityConstraintPoint &cp2 = vc.points[1];
                V2 a = mk(cp1.normalImpulse, cp2.normalImpulse);
                V2 dv1 = vB + cross(wB, cp1.rB) - vA - cross(wA, cp1.rA);
                V2 dv2 = vB + cross(wB, cp2.rB) - vA - cross(wA, cp2.rA);
                float vn1 = dot(dv1, normal);
                float vn2 = dot(dv2, normal);
                V2 b;
                b.x = vn1 - cp1.velocityBias;
                b.y = vn2 - cp2.velocityBias;
                b -= mul(vc.K, a);
                for (;;) {
                    V2 x = -mul(vc.normalMass, b);
                    if (x.x >= 0.0f && x.y >= 0.0f) {
                        applyBlock(vc, x, a, normal, vA, wA, vB, wB, mA, iA, mB, iB);
                        break;
                    }
                    x.x = -cp1.normalMass * b.x;
                    x.y = 0.0f;
                    vn1 = 0.0f;
                    vn2 = vc.K.ex.y * x.x + b.y;
                    if (x.x >= 0.0f && vn2 >= 0.0f) {
                        applyBlock(vc, x, a, normal, vA, wA, vB, wB, mA, iA, mB, iB);
                        break;
                    }
                    x.x = 0.0f;
                    x.y = -cp2.normalMass * b.y;
                    vn1 = vc.K.ey.x * x.y + b.x;
                    vn2 = 0.0f;
                    if (x.y >= 0.0f && vn1 >= 0.0f) {
                        applyBlock(vc, x, a, normal, vA, wA, vB, wB, mA, iA, mB, iB);
                        break;
                    }
                    x.x = 0.0f;
                    x.y = 0.0f;
                    vn1 = b.x;
                    vn2 = b.y;
                    if (vn1 >= 0.0f && vn2 >= 0.0f) {
                        applyBlock(vc, x, a, normal, vA, wA, vB, wB, mA, iA, mB, iB);
                        break;
                    }
                    break;
                }
            }
            velocities[indexA].v = vA;
            velocities[indexA].w = wA;
            velocities[indexB].v = vB;
            velocities[indexB].w = wB;
        }
    }
    static void applyBlock(VelocityConstraint &vc, V2 x, V2 a, V2 normal, V2 &vA, float &wA, V2 &vB, float &wB, float mA,
                           float iA, float mB, float iB) {
        VelocityConstraintPoint &cp1 = vc.points[0];
        VelocityConstraintPoint &cp2 = vc.points[1];
        V2 d = x - a;
        V2 P1 = d.x * normal;
        V2 P2 = d.y * normal;
        vA -= mA * (P1 + P2);
        wA -= iA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
        vB += mB * (P1 + P2);
        wB += iB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
        cp1.normalImpulse = x.x;
        cp2.normalImpulse = x.y;
    }
    std::vector<orc_event> events; /* listener callbacks of the current step */
    void recordEvent(int type, const Contact &c, const VelocityConstraint *vc) {
        orc_event e = {};
        e.type = type;
        e.fixture_a = proxyOf(c.fixtureA, c.indexA);
        e.fixture_b = proxyOf(c.fixtureB, c.indexB);
        if (vc) {
            e.point_count = vc->pointCount;
            for (int j = 0; j < vc->pointCount; ++j) {
                e.normal_impulse[j] = vc->points[j].normalImpulse;
                e.tangent_impulse[j] = vc->points[j].tangentImpulse;
            }
        }
        events.push_back(e);
    }
    void storeImpulses() {
        for (size_t i = 0; i < vcs.size(); ++i) {
            const VelocityConstraint &vc = vcs[i];
            Manifold &manifold = contacts[islandContacts[vc.contactIndex]].manifold;
            for (int j = 0; j < vc.pointCount; ++j) {
                manifold.points[j].normalImpulse = vc.points[j].normalImpulse;
                manifold.points[j].tangentImpulse = vc.points[j].tangentImpulse;
            }
        }
    }
    bool solvePositionConstraints() {
        float minSeparation = 0.0f;
        for (size_t i = 0; i < pcs.size(); ++i) {
            const PositionConstraint &pc = pcs[i];
            int indexA = pc.indexA, indexB = pc.indexB;
            V2 localCenterA = pc.localCenterA;
            float mA = pc.invMassA, iA = pc.invIA;
            V2 localCenterB = pc.localCenterB;
            float mB = pc.invMassB, iB = pc.invIB;
            int pointCount = pc.pointCount;
            V2 cA = positions[indexA].c;
            float aA = positions[indexA].a;
            V2 cB = positions[indexB].c;
            float aB = positions[indexB].a;
            for (int j = 0; j < pointCount; ++j) {
                Xf xfA, xfB;
                xfA.q.set(aA);
                xfB.q.set(aB);
                xfA.p = cA - mul(xfA.q, localCenterA);
                xfB.p = cB - mul(xfB.q, localCenterB);
                V2 normal, point;
                float separation = 0.0f;
                positionManifold(pc, xfA, xfB, j, normal, point, separation);
                V2 rA = point - cA;
                V2 rB = point - cB;
                minSeparation = fmin_(minSeparation, separation);
                float C = clampf(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
                float rnA = cross(rA, normal);
                float rnB = cross(rB, normal);
                float K = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
                float impulse = K > 0.0f ? -C / K : 0.0f;
                V2 P = impulse * normal;
                cA -= mA * P;
                aA -= iA * cross(rA, P);
                cB += mB * P;
                aB += iB * cross(rB, P);
            }
            positions[indexA].c = cA;
            positions[indexA].a = aA;
            positions[indexB].c = cB;
            positions[indexB].a = aB;
        }
        return minSeparation >= -3.0f * kLinearSlop;
    }
    /* b2PositionSolverManifold.Initialize b2contactsolver.d:818-863 */
    static void positionManifold(const PositionConstraint &pc, const Xf &xfA, const Xf &xfB, int index, V2 &normal, V2 &point,
                                 float &separation) {
        switch (pc.type) {
        case m_circles: {
            V2 pointA = mul(xfA, pc.localPoint);
            V2 pointB = mul(xfB, pc.localPoints[0]);
            normal = pointB - pointA;
            normalize(normal);
            point = 0.5f * (pointA + pointB);
            separation = dot(pointB - pointA, normal) - pc.radiusA - pc.radiusB;
        } break;
        case m_faceA: {
            normal = mul(xfA.q, pc.localNormal);
            V2 planePoint = mul(xfA, pc.localPoint);
            V2 clipPoint = mul(xfB, pc.localPoints[index]);
            separation = dot(clipPoint - planePoint, normal) - pc.radiusA - pc.radiusB;
            point = clipPoint;
        } break;
        case m_faceB: {
            normal = mul(xfB.q, pc.localNormal);
            V2 planePoint = mul(xfB, pc.localPoint);
            V2 clipPoint = mul(xfA, pc.localPoints[index]);
            separation = dot(clipPoint - planePoint, normal) - pc.radiusA - pc.radiusB;
            point = clipPoint;
            normal = -normal;
        } break;
        default:
            break;
        }
    }

    /* b2Island.Solve b2island.d:75-280 */
    void solveIsland(const TimeStep &step) {
        auto t0 = Clock::now();
        float h = step.dt;
        int nb = (int)islandBodies.size();
        positions.resize(nb);
        velocities.resize(nb);
        for (int i = 0; i < nb; ++i) {
            Body &b = bodies[islandBodies[i]];
            V2 c = b.sweep.c;
            float a = b.sweep.a;
            V2 v = b.v;
            float w = b.w;
            b.sweep.c0 = b.sweep.c;
            b.sweep.a0 = b.sweep.a;
            if (b.type == b_dynamic) {
                v += h * (b.gravityScale * gravity + b.invMass * b.force);
                w += h * b.invI * b.torque;
                v *= 1.0f / (1.0f + h * b.linearDamping);
                w *= 1.0f / (1.0f + h * b.angularDamping);
            }
            positions[i].c = c;
            positions[i].a = a;
            velocities[i].v = v;
            velocities[i].w = w;
        }
        auto t1 = Clock::now();
        solverCreate(step);
        initializeVelocityConstraints();
        if (step.warmStarting) warmStart();
        SolverData data{step, positions.data(), velocities.data()};
        for (int ji : islandJoints) {
            Joint &jt = joints[ji];
            if (jt.type == j_gear) gearBodies(jt, bodies[jt.bodyC], bodies[jt.bodyD]);
            jointInitVelocity(jt, bodies[jt.bodyA], bodies[jt.bodyB], data);
        }
        profile[3] += msSince(t1);
        auto t2 = Clock::now();
        for (int i = 0; i < step.velocityIterations; ++i) {
            for (int ji : islandJoints) jointSolveVelocity(joints[ji], data);
            solveVelocityConstraints();
        }
        storeImpulses();
        profile[4] += msSince(t2);
        for (int i = 0; i < nb; ++i) {
            V2 c = positions[i].c;
            float a = positions[i].a;
            V2 v = velocities[i].v;
            float w = velocities[i].w;
            V2 translation = h * v;
            if (dot(translation, translation) > kMaxTranslationSquared) {
                float ratio = kMaxTranslation / length(translation);
                v *= ratio;
            }
            float rotation = h * w;
            if (rotation * rotation > kMaxRotationSquared) {
                float ratio = kMaxRotation / fabs_(rotation);
                w *= ratio;
            }
            c += h * v;
            a += h * w;
            positions[i].c = c;
            positions[i].a = a;
            velocities[i].v = v;
            velocities[i].w = w;
        }
        auto t3 = Clock::now();
        bool positionSolved = false;
        for (int i = 0; i < step.positionIterations; ++i) {
            bool contactsOkay = solvePositionConstraints();
            bool jointsOkay = true;
            for (int ji : islandJoints) {
                bool jointOkay = jointSolvePosition(joints[ji], data);
                jointsOkay = jointsOkay && jointOkay;
            }
            if (contactsOkay && jointsOkay) { positionSolved = true; break; }
        }
        for (int i = 0; i < nb; ++i) {
            Body &b = bodies[islandBodies[i]];
            b.sweep.c = positions[i].c;
            b.sweep.a = positions[i].a;
            b.v = velocities[i].v;
            b.w = velocities[i].w;
            b.synchronizeTransform();
        }
        profile[5] += msSince(t3);
        /* b2Island.Report b2island.d:239,438-461 */
        for (size_t i = 0; i < vcs.size(); ++i) recordEvent(3, contacts[islandContacts[vcs[i].contactIndex]], &vcs[i]);
        if (allowSleep) {
            float minSleepTime = kMaxFloat;
            const float linTolSqr = kLinearSleepTolerance * kLinearSleepTolerance;
            const float angTolSqr = kAngularSleepTolerance * kAngularSleepTolerance;
            for (int i = 0; i < nb; ++i) {
                Body &b = bodies[islandBodies[i]];
                if (b.type == b_static) continue;
                if ((b.flags & bf_autoSleep) == 0 || b.w * b.w > angTolSqr || dot(b.v, b.v) > linTolSqr) {
                    b.sleepTime = 0.0f;
                    minSleepTime = 0.0f;
                } else {
                    b.sleepTime += h;
                    minSleepTime = fmin_(minSleepTime, b.sleepTime);
                }
            }
            if (minSleepTime >= kTimeToSleep && positionSolved) {
                for (int i = 0; i < nb; ++i) bodies[islandBodies[i]].setAwake(false);
            }
        }
        (void)t0;
    }

    /* Test hook, not reference behaviour: reorder the island's contacts by greedy
     * colour (lowest colour not used by either dynamic body, in island order),
     * stable within a colour. */
    void colourOrder() {
        std::vector<unsigned long long> mask(bodies.size(), 0ull);
        std::vector<std::pair<int, int>> keyed;
        for (size_t k = 0; k < islandContacts.size(); ++k) {
            const Contact &c = contacts[islandContacts[k]];
            int a = fixtures[c.fixtureA].body, b = fixtures[c.fixtureB].body;
            bool da = bodies[a].type == b_dynamic, db = bodies[b].type == b_dynamic;
            unsigned long long used = (da ? mask[a] : 0ull) | (db ? mask[b] : 0ull);
            int col = 0;
            while (col < 64 && (used >> col & 1ull)) ++col;
            if (col < 64) {
                if (da) mask[a] |= 1ull << col;
                if (db) mask[b] |= 1ull << col;
            } else {
                col = 64 + (int)k;
            }
            keyed.push_back({col, (int)k});
        }
        std::stable_sort(keyed.begin(), keyed.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
            return x.first < y.first;
        });
        std::vector<int> reordered;
        for (auto &kv : keyed) reordered.push_back(islandContacts[kv.second]);
        islandContacts.swap(reordered);
    }

    /* b2World.Solve b2world.d:930-1124 */
    void solve(const TimeStep &step) {
        profile[3] = profile[4] = profile[5] = 0.0f;
        solveLog.clear();
        islandCount = 0;
        for (int b = bodyList; b != kNull; b = bodies[b].next) bodies[b].flags &= ~bf_island;
        for (int c = contactList; c != kNull; c = contacts[c].next) contacts[c].flags &= ~cf_island;
        for (int j = jointList; j != kNull; j = joints[j].next) joints[j].islandFlag = false;
        stack.resize(bodies.size());
        for (int seed = bodyList; seed != kNull; seed = bodies[seed].next) {
            Body &sb = bodies[seed];
            if (sb.flags & bf_island) continue;
            if (!sb.awake() || (sb.flags & bf_active) == 0) continue;
            if (sb.type == b_static) continue;
            islandBodies.clear();
            islandContacts.clear();
            islandJoints.clear();
            int stackCount = 0;
            stack[stackCount++] = seed;
            sb.flags |= bf_island;
            while (stackCount > 0) {
                int bi = stack[--stackCount];
                Body &b = bodies[bi];
                b.islandIndex = (int)islandBodies.size();
                islandBodies.push_back(bi);
                b.setAwake(true);
                if (b.type == b_static) continue;
                for (int e = b.contactList; e != kNull; e = edge(e).next) {
                    int ci = edgeContact(e);
                    Contact &contact = contacts[ci];
                    if (contact.flags & cf_island) continue;
                    if ((contact.flags & cf_enabled) == 0 || (contact.flags & cf_touching) == 0) continue;
                    if (fixtures[contact.fixtureA].sensor || fixtures[contact.fixtureB].sensor) continue;
                    islandContacts.push_back(ci);
                    contact.flags |= cf_island;
                    int other = edge(e).other;
                    if (bodies[other].flags & bf_island) continue;
                    stack[stackCount++] = other;
                    bodies[other].flags |= bf_island;
                }
                for (int e = b.jointList; e != kNull; e = jedge(e).next) { /* b2world.d:1048-1075 */
                    Joint &jt = joints[e >> 1];
                    if (jt.islandFlag) continue;
                    int other = jedge(e).other;
                    if ((bodies[other].flags & bf_active) == 0) continue;
                    islandJoints.push_back(e >> 1);
                    jt.islandFlag = true;
                    if (bodies[other].flags & bf_island) continue;
                    stack[stackCount++] = other;
                    bodies[other].flags |= bf_island;
                }
            }
            if (solveOrder == 1) colourOrder();
            for (int ci : islandContacts) {
                const Contact &c = contacts[ci];
                int a = fixtures[c.fixtureA].body, b = fixtures[c.fixtureB].body;
                solveLog.push_back(islandCount);
                solveLog.push_back(a);
                solveLog.push_back(b);
                solveLog.push_back((bodies[a].type == b_dynamic ? 1 : 0) | (bodies[b].type == b_dynamic ? 2 : 0));
            }
            ++islandCount;
            solveIsland(step);
            for (int bi : islandBodies) {
                if (bodies[bi].type == b_static) bodies[bi].flags &= ~bf_island;
            }
        }
        auto t0 = Clock::now();
        for (int b = bodyList; b != kNull; b = bodies[b].next) {
            Body &body = bodies[b];
            if ((body.flags & bf_island) == 0) continue;
            if (body.type == b_static) continue;
            synchronizeFixtures(body);
        }
        findNewContacts();
        profile[6] = msSince(t0);
    }

    /* b2ContactSolver.SolveTOIPositionConstraints b2contactsolver.d:152-242 */
    bool solveTOIPositionConstraints(int toiIndexA, int toiIndexB) {
        return solveTOIPositionConstraintsMin(toiIndexA, toiIndexB) >= -1.5f * kLinearSlop;
    }
    float solveTOIPositionConstraintsMin(int toiIndexA, int toiIndexB) {
        float minSeparation = 0.0f;
        for (size_t i = 0; i < pcs.size(); ++i) {
            const PositionConstraint &pc = pcs[i];
            int indexA = pc.indexA, indexB = pc.indexB;
            V2 localCenterA = pc.localCenterA, localCenterB = pc.localCenterB;
            int pointCount = pc.pointCount;
            float mA = 0.0f, iA = 0.0f;
            if (indexA == toiIndexA || indexA == toiIndexB) {
                mA = pc.invMassA;
                iA = pc.invIA;
            }
            float mB = 0.0f, iB = 0.0f;
            if (indexB == toiIndexA || indexB == toiIndexB) {
                mB = pc.invMassB;
                iB = pc.invIB;
            }
            V2 cA = positions[indexA].c;
            float aA = positions[indexA].a;
            V2 cB = positions[indexB].c;
            float aB = positions[indexB].a;
            for (int j = 0; j < pointCount; ++j) {
                Xf xfA, xfB;
                xfA.q.set(aA);
                xfB.q.set(aB);
                xfA.p = cA - mul(xfA.q, localCenterA);
                xfB.p = cB - mul(xfB.q, localCenterB);
                V2 normal, point;
                float separation = 0.0f;
                positionManifold(pc, xfA, xfB, j, normal, point, separation);
                V2 rA = point - cA;
                V2 rB = point - cB;
                minSeparation = fmin_(minSeparation, separation);
                float C = clampf(kToiBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
                float rnA = cross(rA, normal);
                float rnB = cross(rB, normal);
                float K = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
                float impulse = K > 0.0f ? -C / K : 0.0f;
                V2 P = impulse * normal;
                cA -= mA * P;
                aA -= iA * cross(rA, P);
                cB += mB * P;
                aB += iB * cross(rB, P);
            }
            positions[indexA].c = cA;
            positions[indexA].a = aA;
            positions[indexB].c = cB;
            positions[indexB].a = aB;
        }
        return minSeparation;
    }

    /* b2Island.SolveTOI b2island.d:282-416 */
    void solveIslandTOI(const TimeStep &subStep, int toiIndexA, int toiIndexB) {
        int nb = (int)islandBodies.size();
        positions.resize(nb);
        velocities.resize(nb);
        for (int i = 0; i < nb; ++i) {
            const Body &b = bodies[islandBodies[i]];
            positions[i].c = b.sweep.c;
            positions[i].a = b.sweep.a;
            velocities[i].v = b.v;
            velocities[i].w = b.w;
        }
        solverCreate(subStep);
        for (int i = 0; i < subStep.positionIterations; ++i)
            if (solveTOIPositionConstraints(toiIndexA, toiIndexB)) break;
        /* leap of faith to the new safe state */
        Body &ta = bodies[islandBodies[toiIndexA]];
        ta.sweep.c0 = positions[toiIndexA].c;
        ta.sweep.a0 = positions[toiIndexA].a;
        Body &tb = bodies[islandBodies[toiIndexB]];
        tb.sweep.c0 = positions[toiIndexB].c;
        tb.sweep.a0 = positions[toiIndexB].a;
        initializeVelocityConstraints();
        for (int i = 0; i < subStep.velocityIterations; ++i) solveVelocityConstraints();
        float h = subStep.dt;
        for (int i = 0; i < nb; ++i) {
            V2 c = positions[i].c;
            float a = positions[i].a;
            V2 v = velocities[i].v;
            float w = velocities[i].w;
            V2 translation = h * v;
            if (dot(translation, translation) > kMaxTranslationSquared) {
                float ratio = kMaxTranslation / length(translation);
                v *= ratio;
            }
            float rotation = h * w;
            if (rotation * rotation > kMaxRotationSquared) {
                float ratio = kMaxRotation / fabs_(rotation);
                w *= ratio;
            }
            c += h * v;
            a += h * w;
            positions[i].c = c;
            positions[i].a = a;
            velocities[i].v = v;
            velocities[i].w = w;
            Body &b = bodies[islandBodies[i]];
            b.sweep.c = c;
            b.sweep.a = a;
            b.v = v;
            b.w = w;
            b.synchronizeTransform();
        }
        /* b2Island.Report b2island.d:415,438-461 */
        for (size_t i = 0; i < vcs.size(); ++i) recordEvent(3, contacts[islandContacts[vcs[i].contactIndex]], &vcs[i]);
    }
    void addIslandBody(int bi) {
        bodies[bi].islandIndex = (int)islandBodies.size();
        islandBodies.push_back(bi);
    }

    /* b2World.SolveTOI b2world.d:1127-1452 (no sub-stepping, no sensors) */
    void solveTOI(const TimeStep &step) {
        if (stepComplete) {
            for (int b = bodyList; b != kNull; b = bodies[b].next) {
                bodies[b].flags &= ~bf_island;
                bodies[b].sweep.alpha0 = 0.0f;
            }
            for (int c = contactList; c != kNull; c = contacts[c].next) {
                contacts[c].flags &= ~(cf_toi | cf_island);
                contacts[c].toiCount = 0;
                contacts[c].toi = 1.0f;
            }
        }
        for (;;) {
            /* the first TOI over the contact list */
            int minContact = kNull;
            float minAlpha = 1.0f;
            for (int ci = contactList; ci != kNull; ci = contacts[ci].next) {
                Contact &c = contacts[ci];
                if ((c.flags & cf_enabled) == 0) continue;
                if (c.toiCount > kMaxSubSteps) continue;
                float alpha = 1.0f;
                if (c.flags & cf_toi) {
                    alpha = c.toi;
                } else {
                    const Fixture &fA = fixtures[c.fixtureA];
                    const Fixture &fB = fixtures[c.fixtureB];
                    if (fA.sensor || fB.sensor) continue;
                    Body &bA = bodies[fA.body];
                    Body &bB = bodies[fB.body];
                    int typeA = bA.type, typeB = bB.type;
                    bool activeA = bA.awake() && typeA != b_static;
                    bool activeB = bB.awake() && typeB != b_static;
                    if (!activeA && !activeB) continue;
                    bool collideA = (bA.flags & bf_bullet) != 0 || typeA != b_dynamic;
                    bool collideB = (bB.flags & bf_bullet) != 0 || typeB != b_dynamic;
                    if (!collideA && !collideB) continue;
                    /* put the sweeps onto the same time interval */
                    float alpha0 = bA.sweep.alpha0;
                    if (bA.sweep.alpha0 < bB.sweep.alpha0) {
                        alpha0 = bB.sweep.alpha0;
                        bA.sweep.advance(alpha0);
                    } else if (bB.sweep.alpha0 < bA.sweep.alpha0) {
                        alpha0 = bA.sweep.alpha0;
                        bB.sweep.advance(alpha0);
                    }
                    DistanceProxy pA, pB;
                    pA.set(fA.shape, c.indexA);
                    pB.set(fB.shape, c.indexB);
                    float beta = 0.0f;
                    int state = timeOfImpact(beta, pA, pB, bA.sweep, bB.sweep, 1.0f);
                    if (state == toi_touching) alpha = fmin_(alpha0 + (1.0f - alpha0) * beta, 1.0f);
                    else alpha = 1.0f;
                    c.toi = alpha;
                    c.flags |= cf_toi;
                }
                if (alpha < minAlpha) {
                    minContact = ci;
                    minAlpha = alpha;
                }
            }
            if (minContact == kNull || 1.0f - 10.0f * kEps < minAlpha) {
                stepComplete = true;
                break;
            }
            /* advance the bodies to the TOI */
            const int ia = fixtures[contacts[minContact].fixtureA].body;
            const int ib = fixtures[contacts[minContact].fixtureB].body;
            Sweep backup1 = bodies[ia].sweep, backup2 = bodies[ib].sweep;
            bodies[ia].advance(minAlpha);
            bodies[ib].advance(minAlpha);
            updateContact(contacts[minContact]);
            contacts[minContact].flags &= ~cf_toi;
            ++contacts[minContact].toiCount;
            if ((contacts[minContact].flags & cf_enabled) == 0 || (contacts[minContact].flags & cf_touching) == 0) {
                contacts[minContact].flags &= ~cf_enabled;
                bodies[ia].sweep = backup1;
                bodies[ib].sweep = backup2;
                bodies[ia].synchronizeTransform();
                bodies[ib].synchronizeTransform();
                continue;
            }
            bodies[ia].setAwake(true);
            bodies[ib].setAwake(true);
            /* the island: both bodies, the TOI contact, and their contacts with
             * static, kinematic or bullet bodies that touch at the TOI */
            islandBodies.clear();
            islandContacts.clear();
            addIslandBody(ia);
            addIslandBody(ib);
            islandContacts.push_back(minContact);
            bodies[ia].flags |= bf_island;
            bodies[ib].flags |= bf_island;
            contacts[minContact].flags |= cf_island;
            const int pair[2] = {ia, ib};
            for (int i = 0; i < 2; ++i) {
                Body &body = bodies[pair[i]];
                if (body.type != b_dynamic) continue;
                for (int e = body.contactList; e != kNull; e = edge(e).next) {
                    if ((int)islandBodies.size() == 2 * kMaxTOIContacts) break;
                    if ((int)islandContacts.size() == kMaxTOIContacts) break;
                    const int ci = edgeContact(e);
                    Contact &contact = contacts[ci];
                    if (contact.flags & cf_island) continue;
                    const int oi = edge(e).other;
                    Body &other = bodies[oi];
                    if (other.type == b_dynamic && (body.flags & bf_bullet) == 0 && (other.flags & bf_bullet) == 0) continue;
                    if (fixtures[contact.fixtureA].sensor || fixtures[contact.fixtureB].sensor) continue;
                    Sweep backup = other.sweep;
                    if ((other.flags & bf_island) == 0) other.advance(minAlpha);
                    updateContact(contact);
                    if ((contact.flags & cf_enabled) == 0 || (contact.flags & cf_touching) == 0) {
                        other.sweep = backup;
                        other.synchronizeTransform();
                        continue;
                    }
                    contact.flags |= cf_island;
                    islandContacts.push_back(ci);
                    if (other.flags & bf_island) continue;
                    other.flags |= bf_island;
                    if (other.type != b_static) other.setAwake(true);
                    addIslandBody(oi);
                }
            }
            TimeStep subStep;
            subStep.dt = (1.0f - minAlpha) * step.dt;
            subStep.inv_dt = 1.0f / subStep.dt;
            subStep.dtRatio = 1.0f;
            subStep.positionIterations = 20;
            subStep.velocityIterations = step.velocityIterations;
            subStep.warmStarting = false;
            solveIslandTOI(subStep, bodies[ia].islandIndex, bodies[ib].islandIndex);
            /* reset island flags and synchronize broad-phase proxies */
            for (int bi : islandBodies) {
                Body &body = bodies[bi];
                body.flags &= ~bf_island;
                if (body.type != b_dynamic) continue;
                synchronizeFixtures(body);
                for (int e = body.contactList; e != kNull; e = edge(e).next)
                    contacts[edgeContact(e)].flags &= ~(cf_toi | cf_island);
            }
            findNewContacts();
            ++toiEvents;
        }
    }

    /* b2World.Step b2world.d:367-434 */
    /* Test hook: b2ContactSolver.SolveVelocityConstraints (b2contactsolver.d:492-772) on one
     * constructed constraint between island bodies 0 and 1, after the masses of
     * InitializeVelocityConstraints (b2contactsolver.d:384-447). */
    void solveOneConstraint(const float *in, float *out) {
        velocities.assign(2, Velocity());
        velocities[0].v = mk(in[0], in[1]);
        velocities[0].w = in[2];
        velocities[1].v = mk(in[3], in[4]);
        velocities[1].w = in[5];
        VelocityConstraint vc;
        vc.indexA = 0;
        vc.indexB = 1;
        vc.invMassA = in[6];
        vc.invIA = in[7];
        vc.invMassB = in[8];
        vc.invIB = in[9];
        vc.normal = mk(in[10], in[11]);
        vc.pointCount = (int)in[28];
        for (int j = 0; j < vc.pointCount; ++j) {
            VelocityConstraintPoint &vcp = vc.points[j];
            vcp.rA = mk(in[12 + 4 * j], in[13 + 4 * j]);
            vcp.rB = mk(in[14 + 4 * j], in[15 + 4 * j]);
            vcp.normalImpulse = in[20 + 2 * j];
            vcp.tangentImpulse = in[21 + 2 * j];
            vcp.velocityBias = in[24 + j];
            pointMasses(vc, vcp);
        }
        blockMasses(vc);
        vc.friction = in[26];
        vc.tangentSpeed = in[27];
        vcs.assign(1, vc);
        const int iterations = (int)in[29];
        for (int i = 0; i < iterations; ++i) solveVelocityConstraints();
        const VelocityConstraint &r = vcs[0];
        out[0] = velocities[0].v.x;
        out[1] = velocities[0].v.y;
        out[2] = velocities[0].w;
        out[3] = velocities[1].v.x;
        out[4] = velocities[1].v.y;
        out[5] = velocities[1].w;
        out[6] = r.points[0].normalImpulse;
        out[7] = r.points[0].tangentImpulse;
        out[8] = r.points[1].normalImpulse;
        out[9] = r.points[1].tangentImpulse;
        out[10] = (float)r.pointCount;
        vcs.clear();
    }
    /* Test hook: one b2ContactSolver.SolveTOIPositionConstraints pass (b2contactsolver.d:152-242)
     * over the contact at list index listIndex, with the bodies of fixture toiFixture's body as
     * the TOI body (the other body keeps zero mass in the pass).  out: the TOI body's sweep c
     * and a after the pass, and the pass's minimum separation. */
    int toiPositionOnce(int listIndex, int toiBody, float *out) {
        int ci = contactList;
        for (int i = 0; i < listIndex && ci != kNull; ++i) ci = contacts[ci].next;
        if (ci == kNull) return -1;
        const Contact &c = contacts[ci];
        const int ba = fixtures[c.fixtureA].body, bb = fixtures[c.fixtureB].body;
        if (toiBody != ba && toiBody != bb) return -1;
        islandBodies = {ba, bb};
        bodies[ba].islandIndex = 0;
        bodies[bb].islandIndex = 1;
        islandContacts = {ci};
        positions.assign(2, Position());
        velocities.assign(2, Velocity());
        for (int i = 0; i < 2; ++i) {
            const Body &b = bodies[islandBodies[i]];
            positions[i].c = b.sweep.c;
            positions[i].a = b.sweep.a;
        }
        TimeStep sub;
        sub.dtRatio = 1.0f;
        sub.warmStarting = false;
        solverCreate(sub);
        const int t = toiBody == ba ? 0 : 1;
        out[3] = solveTOIPositionConstraintsMin(t, t);
        out[0] = positions[t].c.x;
        out[1] = positions[t].c.y;
        out[2] = positions[t].a;
        return 0;
    }
    void step(float dt, int velocityIterations, int positionIterations) {
        const int savedRotMode = g_rotMode;
        if (rotMode >= 0) g_rotMode = rotMode;
        stepImpl(dt, velocityIterations, positionIterations);
        g_rotMode = savedRotMode;
    }
    void stepImpl(float dt, int velocityIterations, int positionIterations) {
        auto t0 = Clock::now();
        events.clear();
        if (newFixture) {
            findNewContacts();
            newFixture = false;
        }
        TimeStep st;
        st.dt = dt;
        st.velocityIterations = velocityIterations;
        st.positionIterations = positionIterations;
        st.inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
        st.dtRatio = inv_dt0 * dt;
        st.warmStarting = warmStarting;
        auto t1 = Clock::now();
        collide();
        profile[1] = msSince(t1);
        if (stepComplete && st.dt > 0.0f) {
            auto t2 = Clock::now();
            solve(st);
            profile[2] = msSince(t2);
        }
        toiEvents = 0;
        profile[7] = 0.0f;
        if (continuousPhysics && st.dt > 0.0f) {
            auto t3 = Clock::now();
            solveTOI(st);
            profile[7] = msSince(t3);
        }
        if (st.dt > 0.0f) inv_dt0 = st.inv_dt;
        for (int b = bodyList; b != kNull; b = bodies[b].next) { /* ClearForces b2world.d:443-450 */
            bodies[b].force = mk(0.0f, 0.0f);
            bodies[b].torque = 0.0f;
        }
        profile[0] = msSince(t0);
    }
};

static void manifoldOut(const Manifold &m, orc_manifold *out) {
    out->type = m.type;
    out->point_count = m.pointCount;
    out->local_normal[0] = m.localNormal.x;
    out->local_normal[1] = m.localNormal.y;
    out->local_point[0] = m.localPoint.x;
    out->local_point[1] = m.localPoint.y;
    for (int i = 0; i < 2; ++i) {
        out->point_local[2 * i] = m.points[i].localPoint.x;
        out->point_local[2 * i + 1] = m.points[i].localPoint.y;
        out->id[i] = m.points[i].id;
    }
}
static void worldManifoldOut(const WorldManifold &wm, orc_manifold *out) {
    out->normal[0] = wm.normal.x;
    out->normal[1] = wm.normal.y;
    for (int i = 0; i < 2; ++i) {
        out->points[2 * i] = wm.points[i].x;
        out->points[2 * i + 1] = wm.points[i].y;
        out->separations[i] = wm.separations[i];
    }
}

} /* namespace orc */

using namespace orc;

extern "C" {

void *orc_world_create(float gx, float gy) {
    World *w = new World();
    w->gravity = mk(gx, gy);
    return w;
}
void orc_world_destroy(void *w) { delete (World *)w; }
int orc_world_set_flags(void *w, int warm, int sleep, int continuous) {
    World *W = (World *)w;
    W->warmStarting = warm != 0;
    W->allowSleep = sleep != 0;
    W->continuousPhysics = continuous != 0;
    return 0;
}
int orc_solve_log(void *w, int32_t *out, int cap) {
    World *W = (World *)w;
    int n = (int)W->solveLog.size() / 4;
    for (int i = 0; i < 4 * n && i < 4 * cap; ++i) out[i] = W->solveLog[i];
    return n;
}
int orc_world_set_solve_order(void *w, int mode) {
    ((World *)w)->solveOrder = mode;
    return 0;
}
int orc_create_body(void *w, const orc_body_def *def) { return ((World *)w)->createBody(*def); }
int orc_create_fixture(void *w, const orc_fixture_def *def) {
    World *W = (World *)w;
    if (def->body < 0 || def->body >= (int)W->bodies.size()) return -1;
    return W->createFixture(*def);
}
int orc_build(void *w, const orc_body_def *bodies, int nb, const orc_fixture_def *fixtures, int nf) {
    for (int i = 0; i < nb; ++i) orc_create_body(w, &bodies[i]);
    for (int i = 0; i < nf; ++i)
        if (orc_create_fixture(w, &fixtures[i]) < 0) return -1;
    return 0;
}
void orc_step(void *w, float dt, int vel, int pos) { ((World *)w)->step(dt, vel, pos); }
int orc_body_count(void *w) { return (int)((World *)w)->bodies.size(); }
int orc_fixture_count(void *w) { return (int)((World *)w)->fixtures.size(); }
int orc_contact_count(void *w) { return ((World *)w)->contactCount; }
int orc_events(void *w, orc_event *out, int cap) {
    const std::vector<orc_event> &ev = ((World *)w)->events;
    int n = (int)ev.size();
    for (int i = 0; i < n && i < cap; ++i) out[i] = ev[i];
    return n;
}
int orc_dynamic_body_count(void *w) {
    int n = 0;
    for (const Body &b : ((World *)w)->bodies) n += b.type == b_dynamic;
    return n;
}
void orc_get_bodies(void *w, orc_body_state *out) {
    World *W = (World *)w;
    for (size_t i = 0; i < W->bodies.size(); ++i) {
        const Body &b = W->bodies[i];
        orc_body_state &s = out[i];
        s.px = b.xf.p.x; s.py = b.xf.p.y; s.qs = b.xf.q.s; s.qc = b.xf.q.c;
        s.cx = b.sweep.c.x; s.cy = b.sweep.c.y; s.a = b.sweep.a; s.sleep_time = b.sleepTime;
        s.vx = b.v.x; s.vy = b.v.y; s.w = b.w; s.flags = b.flags;
        s.c0x = b.sweep.c0.x; s.c0y = b.sweep.c0.y; s.a0 = b.sweep.a0; s.inv_mass = b.invMass;
    }
}
/* The proxy pairs of a body's contact edges (b2Body.m_contactList, b2body.d, kept by
 * AddPair / Destroy at b2contactmanager.d:138-166 and 88-127), head first. */
int orc_body_contact_pairs(void *w, int body, int *out_pairs, int cap) {
    World *W = (World *)w;
    if (body < 0 || body >= (int)W->bodies.size()) return -1;
    int n = 0;
    for (int e = W->bodies[body].contactList; e != kNull; e = W->edge(e).next, ++n) {
        if (n >= cap) continue;
        const Contact &c = W->contacts[e >> 1];
        out_pairs[2 * n] = W->proxyOf(c.fixtureA, c.indexA);
        out_pairs[2 * n + 1] = W->proxyOf(c.fixtureB, c.indexB);
    }
    return n;
}

int orc_get_contacts(void *w, orc_contact *out, int cap) {
    World *W = (World *)w;
    int n = 0;
    for (int ci = W->contactList; ci != kNull && n < cap; ci = W->contacts[ci].next, ++n) {
        const Contact &c = W->contacts[ci];
        orc_contact &o = out[n];
        o.fixture_a = W->proxyOf(c.fixtureA, c.indexA);
        o.fixture_b = W->proxyOf(c.fixtureB, c.indexB);
        o.flags = (int)c.flags;
        o.type = c.manifold.type;
        o.point_count = c.manifold.pointCount;
        o.local_normal[0] = c.manifold.localNormal.x;
        o.local_normal[1] = c.manifold.localNormal.y;
        o.local_point[0] = c.manifold.localPoint.x;
        o.local_point[1] = c.manifold.localPoint.y;
        for (int i = 0; i < 2; ++i) {
            o.point_local[2 * i] = c.manifold.points[i].localPoint.x;
            o.point_local[2 * i + 1] = c.manifold.points[i].localPoint.y;
            o.normal_impulse[i] = c.manifold.points[i].normalImpulse;
            o.tangent_impulse[i] = c.manifold.points[i].tangentImpulse;
            o.id[i] = c.manifold.points[i].id;
        }
        o.friction = c.friction;
        o.restitution = c.restitution;
    }
    return n;
}
int orc_contact_world_manifold(void *w, int list_index, orc_manifold *out) {
    World *W = (World *)w;
    int ci = W->contactList;
    for (int i = 0; i < list_index && ci != kNull; ++i) ci = W->contacts[ci].next;
    if (ci == kNull) return -1;
    const Contact &c = W->contacts[ci];
    const Fixture &fa = W->fixtures[c.fixtureA];
    const Fixture &fb = W->fixtures[c.fixtureB];
    WorldManifold wm;
    wm.initialize(c.manifold, W->bodies[fa.body].xf, fa.shape.radius, W->bodies[fb.body].xf, fb.shape.radius);
    manifoldOut(c.manifold, out);
    worldManifoldOut(wm, out);
    out->swapped = 0;
    return 0;
}
float orc_max_penetration(void *w) {
    World *W = (World *)w;
    float minSep = 0.0f;
    for (int ci = W->contactList; ci != kNull; ci = W->contacts[ci].next) {
        const Contact &c = W->contacts[ci];
        if ((c.flags & cf_touching) == 0) continue;
        const Fixture &fa = W->fixtures[c.fixtureA];
        const Fixture &fb = W->fixtures[c.fixtureB];
        WorldManifold wm;
        wm.initialize(c.manifold, W->bodies[fa.body].xf, fa.shape.radius, W->bodies[fb.body].xf, fb.shape.radius);
        for (int i = 0; i < c.manifold.pointCount; ++i) minSep = fmin_(minSep, wm.separations[i]);
    }
    return -minSep;
}
int orc_proxy_count(void *w) { return (int)((World *)w)->canonFix.size(); }
void orc_get_fat_aabbs(void *w, float *out) {
    World *W = (World *)w;
    for (size_t i = 0; i < W->canonFix.size(); ++i) {
        const AABB &a = W->bp.tree.nodes[W->fixtures[W->canonFix[i]].proxyIds[W->canonChild[i]]].aabb;
        out[4 * i + 0] = a.lo.x;
        out[4 * i + 1] = a.lo.y;
        out[4 * i + 2] = a.hi.x;
        out[4 * i + 3] = a.hi.y;
    }
}
int orc_last_pairs(void *w, int32_t *out, int cap) {
    World *W = (World *)w;
    int n = (int)W->lastPairs.size();
    for (int i = 0; i < n && i < cap; ++i) {
        out[2 * i] = W->lastPairs[i].a;
        out[2 * i + 1] = W->lastPairs[i].b;
    }
    return n;
}
void orc_profile(void *w, float *out8) {
    World *W = (World *)w;
    for (int i = 0; i < 8; ++i) out8[i] = W->profile[i];
}
int orc_collide(const orc_fixture_def *a, const float *xfa3, const orc_fixture_def *b, const float *xfb3, orc_manifold *out) {
    Shape sa, sb;
    World::buildShape(*a, sa);
    World::buildShape(*b, sb);
    Xf xfa, xfb;
    xfa.p = mk(xfa3[0], xfa3[1]);
    xfa.q.set(xfa3[2]);
    xfb.p = mk(xfb3[0], xfb3[1]);
    xfb.q.set(xfb3[2]);
    memset(out, 0, sizeof(*out));
    if (!hasContactType(sa.type, sb.type)) return -1;
    const Shape *A = &sa, *B = &sb;
    const Xf *XA = &xfa, *XB = &xfb;
    if (!isPrimary(sa.type, sb.type)) { std::swap(A, B); std::swap(XA, XB); out->swapped = 1; }
    Manifold m;
    evaluate(m, *A, *XA, *B, *XB);
    WorldManifold wm;
    wm.initialize(m, *XA, A->radius, *XB, B->radius);
    int swapped = out->swapped;
    manifoldOut(m, out);
    if (m.pointCount > 0) worldManifoldOut(wm, out);
    out->swapped = swapped;
    return 0;
}
int orc_shape_distance(const orc_fixture_def *a, const float *xfa3, const orc_fixture_def *b, const float *xfb3,
                       int use_radii, float *out6) {
    Shape sa, sb;
    World::buildShape(*a, sa);
    World::buildShape(*b, sb);
    Xf xfa, xfb;
    xfa.p = mk(xfa3[0], xfa3[1]);
    xfa.q.set(xfa3[2]);
    xfb.p = mk(xfb3[0], xfb3[1]);
    xfb.q.set(xfb3[2]);
    DistanceProxy pa, pb;
    pa.set(sa);
    pb.set(sb);
    SimplexCache cache;
    DistanceOutput o;
    gjkDistance(o, cache, pa, xfa, pb, xfb, use_radii != 0);
    out6[0] = o.pointA.x;
    out6[1] = o.pointA.y;
    out6[2] = o.pointB.x;
    out6[3] = o.pointB.y;
    out6[4] = o.distance;
    out6[5] = (float)o.iterations;
    return 0;
}
static Sweep sweepIn(const float *s9) {
    Sweep sw;
    sw.localCenter = mk(s9[0], s9[1]);
    sw.c0 = mk(s9[2], s9[3]);
    sw.c = mk(s9[4], s9[5]);
    sw.a0 = s9[6];
    sw.a = s9[7];
    sw.alpha0 = s9[8];
    return sw;
}
int orc_time_of_impact(const orc_fixture_def *a, const float *sweep_a9, const orc_fixture_def *b, const float *sweep_b9,
                       float t_max, float *t_out) {
    Shape sa, sb;
    World::buildShape(*a, sa);
    World::buildShape(*b, sb);
    DistanceProxy pa, pb;
    pa.set(sa);
    pb.set(sb);
    return timeOfImpact(*t_out, pa, pb, sweepIn(sweep_a9), sweepIn(sweep_b9), t_max);
}
int orc_toi_events(void *w) { return ((World *)w)->toiEvents; }
int orc_create_joint(void *w, const orc_joint_def *def) { return ((World *)w)->createJoint(*def); }
int orc_joint_count(void *w) { return (int)((World *)w)->joints.size(); }
void orc_get_joints(void *w, float *out) {
    const World &W = *(const World *)w;
    for (size_t i = 0; i < W.joints.size(); ++i) {
        const Joint &j = W.joints[i];
        float *o = out + 8 * i;
        const bool lin = j.type == j_friction || j.type == j_motor || j.type == j_mouse;
        o[0] = lin ? j.linearImpulse.x : j.impulse.x;
        o[1] = lin ? j.linearImpulse.y : j.impulse.y;
        o[2] = lin ? j.angularImpulse : j.impulse.z;
        o[3] = j.motorImpulse;
        o[4] = j.springImpulse;
        o[5] = (float)j.limitState;
        o[6] = 0.0f;
        o[7] = 0.0f;
    }
}
int orc_set_rot_mode(int mode) {
    const int prev = g_rotMode;
    g_rotMode = mode ? 1 : 0;
    return prev;
}
int orc_world_set_rot_mode(void *w, int mode) {
    if (mode < -1 || mode > 1) return -1;
    ((World *)w)->rotMode = mode;
    return 0;
}
void orc_solve_one_constraint(const float *in30, float *out11) {
    World w;
    w.solveOneConstraint(in30, out11);
}
int orc_toi_position_once(void *w, int list_index, int toi_body, float *out4) {
    return ((World *)w)->toiPositionOnce(list_index, toi_body, out4);
}
int orc_world_set_presolve(void *w, int mode) {
    if (mode < 0 || mode > 2) return -1;
    ((World *)w)->presolveMode = mode;
    return 0;
}
int orc_collide_child(const orc_fixture_def *chain, int child, const float *xfa3, const orc_fixture_def *b,
                      const float *xfb3, orc_manifold *out) {
    Shape sa, sb;
    World::buildShape(*chain, sa);
    World::buildShape(*b, sb);
    if (sa.type != s_chain || child < 0 || child >= childCount(sa)) return -1;
    if (sb.type != s_circle && sb.type != s_polygon) return -1;
    Xf xfa, xfb;
    xfa.p = mk(xfa3[0], xfa3[1]);
    xfa.q.set(xfa3[2]);
    xfb.p = mk(xfb3[0], xfb3[1]);
    xfb.q.set(xfb3[2]);
    memset(out, 0, sizeof(*out));
    Manifold m;
    evaluate(m, sa, xfa, sb, xfb, child);
    WorldManifold wm;
    wm.initialize(m, xfa, sa.radius, xfb, sb.radius);
    manifoldOut(m, out);
    if (m.pointCount > 0) worldManifoldOut(wm, out);
    return 0;
}
int orc_shape_geometry(const orc_fixture_def *def, float *o, float *mass_out4) {
    Shape s;
    World::buildShape(*def, s);
    memset(o, 0, 36 * sizeof(float));
    o[0] = (float)s.count;
    o[1] = s.radius;
    o[2] = s.type == s_circle ? s.p.x : s.centroid.x;
    o[3] = s.type == s_circle ? s.p.y : s.centroid.y;
    for (int i = 0; i < s.count; ++i) {
        o[4 + 2 * i] = s.v[i].x;
        o[5 + 2 * i] = s.v[i].y;
        o[20 + 2 * i] = s.n[i].x;
        o[21 + 2 * i] = s.n[i].y;
    }
    MassData md;
    computeMass(s, md, def->density);
    mass_out4[0] = md.mass;
    mass_out4[1] = md.center.x;
    mass_out4[2] = md.center.y;
    mass_out4[3] = md.I;
    return 0;
}
int orc_tree_pairs(const float *fat, const uint8_t *moved, int n, int32_t *out, int cap) {
    BroadPhase bp;
    std::vector<int> id(n);
    for (int i = 0; i < n; ++i) {
        AABB a;
        a.lo = mk(fat[4 * i], fat[4 * i + 1]);
        a.hi = mk(fat[4 * i + 2], fat[4 * i + 3]);
        id[i] = bp.tree.createProxyFat(a, i);
    }
    for (int i = 0; i < n; ++i)
        if (moved[i]) bp.moveBuffer.push_back(id[i]);
    int count = 0;
    bp.updatePairs([&](int ua, int ub, const Pair &) {
        if (count < cap) {
            out[2 * count] = ua < ub ? ua : ub;
            out[2 * count + 1] = ua < ub ? ub : ua;
        }
        ++count;
    });
    return count;
}

} /* extern "C" */
