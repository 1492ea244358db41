// dbx_geom.h -- device geometry for the step kernels: 2-D vector algebra,
// shapes, the narrowphase collide functions and the world manifold.
//
// Every expression keeps the floating-point operation order of the reference
// (src/dbox/common/b2math.d, collision/*.d), and the library is compiled with
// -ffp-contract=off, so a kernel result differs from the reference's scalar
// result only where a libm call (sin, cos) rounds differently.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dbx {

// b2settings.d:61-149, evaluated in float.
constexpr float kPi = 3.14159265359f;
constexpr float kEps = 1.1920928955078125e-07f; // FLT_EPSILON
constexpr float kMaxFloat = 3.402823466e+38f;   // FLT_MAX
constexpr float kLinearSlop = 0.005f;
constexpr float kAngularSlop = (2.0f / 180.0f * kPi);
constexpr float kPolygonRadius = (2.0f * kLinearSlop);
constexpr float kAabbExtension = 0.1f;
constexpr float kAabbMultiplier = 2.0f;
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
constexpr int kMaxPolygonVertices = 8;

// Shape kinds (b2Shape.Type) and manifold kinds (b2Manifold.Type).
enum : int { kCircle = 0, kEdge = 1, kPolygon = 2 };
enum : int { kManCircles = 0, kManFaceA = 1, kManFaceB = 2 };
enum : int { kFeatVertex = 0, kFeatFace = 1 };

// One shape, 160 bytes, read-only on the device.
//  polygon: v[0..count) vertices, n[0..count) normals, c = centroid
//  circle : c = local centre
//  edge   : v[0] = vertex1, v[1] = vertex2, v[2] = vertex0, v[3] = vertex3, flags bit0/bit1 = has v0/v3;
//           flags bit2: a child edge of a chain (b2ChainShape.GetChildEdge), whose AABB has no radius
//  any    : flags bit3 (kShapeSensor): the fixture is a sensor (b2Fixture.m_isSensor)
constexpr int kShapeSensor = 8;
struct alignas(16) Shape {
    int type;
    int count;
    float radius;
    int flags;
    float2 c;
    float2 pad;
    float2 v[kMaxPolygonVertices];
    float2 n[kMaxPolygonVertices];
};
static_assert(sizeof(Shape) == 160, "shape layout");

struct Rot {
    float s, c;
};
struct Xf {
    float2 p;
    Rot q;
};

__device__ __forceinline__ float2 v2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float2 add(float2 a, float2 b) { return v2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return v2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 neg(float2 a) { return v2(-a.x, -a.y); }
__device__ __forceinline__ float2 scl(float s, float2 a) { return v2(s * a.x, s * a.y); }
__device__ __forceinline__ float dot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float crs(float2 a, float2 b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ float2 crs(float2 a, float s) { return v2(s * a.y, -s * a.x); }
__device__ __forceinline__ float2 crs(float s, float2 a) { return v2(-s * a.y, s * a.x); }
__device__ __forceinline__ float lensq(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float dist2(float2 a, float2 b) { float2 c = sub(a, b); return dot(c, c); }
__device__ __forceinline__ float fminr(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmaxr(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fabsr(float a) { return a > 0.0f ? a : -a; }
__device__ __forceinline__ float clampr(float a, float lo, float hi) { return fmaxr(lo, fminr(a, hi)); }
// Component-wise select; a conditional expression on float2 objects can be
// lowered to a select of their addresses, which keeps them in scratch memory.
__device__ __forceinline__ float2 sel(bool c, float2 a, float2 b) { return v2(c ? a.x : b.x, c ? a.y : b.y); }
__device__ __forceinline__ float2 vminr(float2 a, float2 b) { return v2(fminr(a.x, b.x), fminr(a.y, b.y)); }
__device__ __forceinline__ float2 vmaxr(float2 a, float2 b) { return v2(fmaxr(a.x, b.x), fmaxr(a.y, b.y)); }

// b2Vec2.Normalize (b2math.d:162-175); sqrtf and 1/x are correctly rounded in HIP.
__device__ __forceinline__ float normalize(float2 &v) {
    float len = sqrtf(v.x * v.x + v.y * v.y);
    if (len < kEps) return 0.0f;
    float inv = 1.0f / len;
    v.x *= inv;
    v.y *= inv;
    return len;
}

// b2Rot.Set (b2math.d:483-488) calls sinf/cosf.  Here both are evaluated in
// double precision and rounded once to float:
//   x = k * pi/2 + r, |r| <= pi/4, with pi/2 split into three doubles whose
//   first two have 33 significant bits, so k * P1 and k * P2 are exact for
//   |k| < 2^20 (Cody-Waite reduction);
//   sin r and cos r from their Taylor series up to r^15 / r^16, whose truncation
//   error on |r| <= pi/4 is below 5e-17 (about half a double ulp of the result);
//   the quadrant k mod 4 picks the signs and the swap of sin and cos.
// The float results equal the correctly rounded sin and cos except when the exact
// value lies within a few double ulps of a float rounding boundary.
//
// Both Horner chains are written out as one block of v_fma_f64 instructions, the
// two chains interleaved: written as fma(), the compiler turns each step into the
// two-operand v_fmac_f64 and copies the loop-invariant coefficient before it, and
// separate asm statements get a wait state between them.  Each instruction
// computes exactly what the corresponding fma() computes, in the same order.
__device__ __forceinline__ void sincos_poly(double z, double &ps, double &pc) {
    const double s0 = -7.647163731819816e-13, s1 = 1.6059043836821613e-10, s2 = -2.505210838544172e-08,
                 s3 = 2.7557319223985893e-06, s4 = -0.0001984126984126984, s5 = 0.008333333333333333,
                 s6 = -0.16666666666666666;
    const double c0 = 4.779477332387385e-14, c1 = -1.1470745597729725e-11, c2 = 2.08767569878681e-09,
                 c3 = -2.755731922398589e-07, c4 = 2.48015873015873e-05, c5 = -0.001388888888888889,
                 c6 = 0.041666666666666664;
    asm("v_fma_f64 %0, %3, %2, %4\n\t"
        "v_fma_f64 %1, %10, %2, %11\n\t"
        "v_fma_f64 %0, %0, %2, %5\n\t"
        "v_fma_f64 %1, %1, %2, %12\n\t"
        "v_fma_f64 %0, %0, %2, %6\n\t"
        "v_fma_f64 %1, %1, %2, %13\n\t"
        "v_fma_f64 %0, %0, %2, %7\n\t"
        "v_fma_f64 %1, %1, %2, %14\n\t"
        "v_fma_f64 %0, %0, %2, %8\n\t"
        "v_fma_f64 %1, %1, %2, %15\n\t"
        "v_fma_f64 %0, %0, %2, %9\n\t"
        "v_fma_f64 %1, %1, %2, %16\n\t"
        "v_fma_f64 %1, %1, %2, -0.5\n\t"
        "v_fma_f64 %1, %1, %2, 1.0"
        : "=&v"(ps), "=&v"(pc)
        : "v"(z), "v"(s0), "v"(s1), "v"(s2), "v"(s3), "v"(s4), "v"(s5), "v"(s6), "v"(c0), "v"(c1), "v"(c2),
          "v"(c3), "v"(c4), "v"(c5), "v"(c6));
}
__device__ __forceinline__ Rot rot(float angle) {
    const double x = (double)angle;
    const double P1 = 0x1.921fb54400000p+0, P2 = 0x1.0b4611a600000p-34, P3 = 0x1.3198a2e037073p-69;
    const double k = rint(x * 0.6366197723675814);
    double r = fma(-k, P1, x);
    r = fma(-k, P2, r);
    r = fma(-k, P3, r);
    const double z = r * r;
    double ps, pc;
    sincos_poly(z, ps, pc); // ps = s0..s6 in z; pc = c0..c6, -1/2, 1 in z
    ps = fma(ps * z, r, r);
    // quadrant q: sin = (ps, pc, -ps, -pc)[q], cos = (pc, -ps, -pc, ps)[q]; rounding to
    // float commutes with negation, so the swap and signs are applied to the floats,
    // the signs as a flip of the sign bit.  |k| < 2^20 wherever this result is kept,
    // so k converts to int32 exactly.
    const unsigned q = (unsigned)__double2int_rn(k);
    const float fs = (float)ps, fc = (float)pc;
    const bool odd = (q & 1u) != 0u;
    Rot qr;
    qr.s = __uint_as_float(__float_as_uint(odd ? fc : fs) ^ ((q & 2u) << 30));
    qr.c = __uint_as_float(__float_as_uint(odd ? fs : fc) ^ (((q + 1u) & 2u) << 30));
    // |angle| >= 1e6 (or not finite): libm's sincos.  The branch is taken by the
    // whole wave or not at all, so the common path carries no divergent branch.
    const bool big = !(fabs(x) < 1.0e6);
    if (__ballot(big) != 0ull) {
        double s, c;
        sincos(x, &s, &c);
        qr.s = big ? (float)s : qr.s;
        qr.c = big ? (float)c : qr.c;
    }
    return qr;
}
__device__ __forceinline__ float2 mul(Rot q, float2 v) { return v2(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
__device__ __forceinline__ float2 mulT(Rot q, float2 v) { return v2(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
__device__ __forceinline__ float2 mul(const Xf &T, float2 v) {
    return v2((T.q.c * v.x - T.q.s * v.y) + T.p.x, (T.q.s * v.x + T.q.c * v.y) + T.p.y);
}
__device__ __forceinline__ float2 mulT(const Xf &T, float2 v) {
    float px = v.x - T.p.x;
    float py = v.y - T.p.y;
    return v2(T.q.c * px + T.q.s * py, -T.q.s * px + T.q.c * py);
}
__device__ __forceinline__ Xf mulT(const Xf &A, const Xf &B) {
    Xf C;
    C.q.s = A.q.c * B.q.s - A.q.s * B.q.c;
    C.q.c = A.q.c * B.q.c + A.q.s * B.q.s;
    C.p = mulT(A.q, sub(B.p, A.p));
    return C;
}
// Transform of a body from its centre of mass c, angle a and local centre lc.
__device__ __forceinline__ Xf xf_from(float2 c, float a, float2 lc) {
    Xf x;
    x.q = rot(a);
    x.p = sub(c, mul(x.q, lc));
    return x;
}

// b2ContactID key: indexA | indexB << 8 | typeA << 16 | typeB << 24 (b2collision.d:38-60).
__device__ __forceinline__ uint32_t fkey(uint32_t ia, uint32_t ib, uint32_t ta, uint32_t tb) {
    return (ia & 0xffu) | ((ib & 0xffu) << 8) | ((ta & 0xffu) << 16) | ((tb & 0xffu) << 24);
}
__device__ __forceinline__ uint32_t fkey_swap(uint32_t k) {
    return ((k >> 8) & 0xffu) | ((k & 0xffu) << 8) | (((k >> 24) & 0xffu) << 16) | (((k >> 16) & 0xffu) << 24);
}

// Local manifold (b2Manifold, b2collision.d:96-114) in registers.
struct Manifold {
    int type;
    int count;
    float2 ln;    // localNormal
    float2 lp;    // localPoint
    float2 p[2];  // points' localPoint
    uint32_t id[2];
};

struct AABB {
    float2 lo, hi;
};
// b2TestOverlap(b2AABB, b2AABB) b2collision.d:470-483: touching counts as overlap.
__device__ __forceinline__ bool overlap(float4 a, float4 b) {
    float d1x = b.x - a.z, d1y = b.y - a.w;
    float d2x = a.x - b.z, d2y = a.y - b.w;
    if (d1x > 0.0f || d1y > 0.0f) return false;
    if (d2x > 0.0f || d2y > 0.0f) return false;
    return true;
}

// ComputeAABB of a shape at a transform (b2polygonshape.d:356-373,
// b2circleshape.d:110-117, b2edgeshape.d:163-176).
__device__ inline float4 shape_aabb(const Shape *__restrict__ s, const Xf &xf) {
    int t = s->type;
    if (t == kPolygon) {
        float2 lo = mul(xf, s->v[0]);
        float2 hi = lo;
        int n = s->count;
        for (int i = 1; i < n; ++i) {
            float2 w = mul(xf, s->v[i]);
            lo = vminr(lo, w);
            hi = vmaxr(hi, w);
        }
        float r = s->radius;
        return make_float4(lo.x - r, lo.y - r, hi.x + r, hi.y + r);
    }
    if (t == kCircle) {
        float2 p = add(xf.p, mul(xf.q, s->c));
        float r = s->radius;
        return make_float4(p.x - r, p.y - r, p.x + r, p.y + r);
    }
    float2 a = mul(xf, s->v[0]);
    float2 b = mul(xf, s->v[1]);
    float2 lo = vminr(a, b), hi = vmaxr(a, b);
    // b2EdgeShape.ComputeAABB adds the radius; b2ChainShape.ComputeAABB does not
    // (b2edgeshape.d:163-176, b2chainshape.d:214-226); x - 0 is x exactly
    float r = (s->flags & 4) ? 0.0f : s->radius;
    return make_float4(lo.x - r, lo.y - r, hi.x + r, hi.y + r);
}

struct ClipV {
    float2 v;
    uint32_t id;
};
// b2ClipSegmentToLine (b2collision.d:413-446).  The outputs are, in order, the
// input points behind the plane and then the intersection point; written with
// selects so that no local array is indexed at run time.
__device__ __forceinline__ int clip_segment(ClipV &o0, ClipV &o1, const ClipV &i0, const ClipV &i1, float2 normal,
                                            float offset, int vertexIndexA) {
    float d0 = dot(normal, i0.v) - offset;
    float d1 = dot(normal, i1.v) - offset;
    bool k0 = d0 <= 0.0f, k1 = d1 <= 0.0f, cut = d0 * d1 < 0.0f;
    ClipV x;
    if (cut) {
        float interp = d0 / (d0 - d1);
        x.v = add(i0.v, scl(interp, sub(i1.v, i0.v)));
        x.id = fkey((uint32_t)vertexIndexA, (i0.id >> 8) & 0xffu, kFeatVertex, kFeatFace);
    } else {
        x = i0;
    }
    o0 = k0 ? i0 : (k1 ? i1 : x);
    o1 = (k0 && k1) ? i1 : x;
    return (int)k0 + (int)k1 + (int)cut;
}

// Appends a manifold point (count < 2) without run-time indexing.
__device__ __forceinline__ void push_point(Manifold &m, int &pc, float2 p, uint32_t id) {
    if (pc == 0) {
        m.p[0] = p;
        m.id[0] = id;
    } else {
        m.p[1] = p;
        m.id[1] = id;
    }
    ++pc;
}

// b2FindMaxSeparation (b2collidepolygon.d:29-71): edge normals of p1 against p2 in p2's frame.
__device__ inline float max_separation(int &edge, const Shape *__restrict__ p1, const Xf &xf1, const Shape *__restrict__ p2,
                                       const Xf &xf2) {
    Xf x = mulT(xf2, xf1);
    int n1 = p1->count, n2 = p2->count;
    int best = 0;
    float maxSep = -kMaxFloat;
    for (int i = 0; i < n1; ++i) {
        float2 n = mul(x.q, p1->n[i]);
        float2 w = mul(x, p1->v[i]);
        float si = kMaxFloat;
        for (int j = 0; j < n2; ++j) {
            float sij = dot(n, sub(p2->v[j], w));
            si = sij < si ? sij : si;
        }
        if (si > maxSep) {
            maxSep = si;
            best = i;
        }
    }
    edge = best;
    return maxSep;
}

// b2CollidePolygons (b2collidepolygon.d:127-254) with b2FindIncidentEdge (73-118).
__device__ inline void collide_polygons(Manifold &m, const Shape *__restrict__ A, const Xf &xfA, const Shape *__restrict__ B,
                                        const Xf &xfB) {
    m.count = 0;
    float totalRadius = A->radius + B->radius;
    int edgeA = 0;
    float sepA = max_separation(edgeA, A, xfA, B, xfB);
    if (sepA > totalRadius) return;
    int edgeB = 0;
    float sepB = max_separation(edgeB, B, xfB, A, xfA);
    if (sepB > totalRadius) return;
    const float k_tol = 0.1f * kLinearSlop;
    bool flip = sepB > sepA + k_tol;
    const Shape *p1 = flip ? B : A;
    const Shape *p2 = flip ? A : B;
    Xf xf1 = flip ? xfB : xfA;
    Xf xf2 = flip ? xfA : xfB;
    int edge1 = flip ? edgeB : edgeA;
    m.type = flip ? kManFaceB : kManFaceA;

    // incident edge on p2
    float2 normal1 = mulT(xf2.q, mul(xf1.q, p1->n[edge1]));
    int n2 = p2->count;
    int index = 0;
    float minDot = kMaxFloat;
    for (int i = 0; i < n2; ++i) {
        float d = dot(normal1, p2->n[i]);
        if (d < minDot) {
            minDot = d;
            index = i;
        }
    }
    int i2 = index + 1 < n2 ? index + 1 : 0;
    ClipV inc0, inc1;
    inc0.v = mul(xf2, p2->v[index]);
    inc0.id = fkey(edge1, index, kFeatFace, kFeatVertex);
    inc1.v = mul(xf2, p2->v[i2]);
    inc1.id = fkey(edge1, i2, kFeatFace, kFeatVertex);

    int n1 = p1->count;
    int iv1 = edge1;
    int iv2 = edge1 + 1 < n1 ? edge1 + 1 : 0;
    float2 v11 = p1->v[iv1];
    float2 v12 = p1->v[iv2];
    float2 localTangent = sub(v12, v11);
    normalize(localTangent);
    float2 localNormal = crs(localTangent, 1.0f);
    float2 planePoint = scl(0.5f, add(v11, v12));
    float2 tangent = mul(xf1.q, localTangent);
    float2 normal = crs(tangent, 1.0f);
    v11 = mul(xf1, v11);
    v12 = mul(xf1, v12);
    float frontOffset = dot(normal, v11);
    float sideOffset1 = -dot(tangent, v11) + totalRadius;
    float sideOffset2 = dot(tangent, v12) + totalRadius;
    ClipV a0, a1, b0, b1;
    if (clip_segment(a0, a1, inc0, inc1, neg(tangent), sideOffset1, iv1) < 2) return;
    if (clip_segment(b0, b1, a0, a1, tangent, sideOffset2, iv2) < 2) return;
    m.ln = localNormal;
    m.lp = planePoint;
    int pc = 0;
    if (dot(normal, b0.v) - frontOffset <= totalRadius)
        push_point(m, pc, mulT(xf2, b0.v), flip ? fkey_swap(b0.id) : b0.id);
    if (dot(normal, b1.v) - frontOffset <= totalRadius)
        push_point(m, pc, mulT(xf2, b1.v), flip ? fkey_swap(b1.id) : b1.id);
    m.count = pc;
}

// b2CollideCircles (b2collidecircle.d:29-56).
__device__ inline void collide_circles(Manifold &m, const Shape *__restrict__ A, const Xf &xfA, const Shape *__restrict__ B,
                                       const Xf &xfB) {
    m.count = 0;
    float2 pA = mul(xfA, A->c);
    float2 pB = mul(xfB, B->c);
    float2 d = sub(pB, pA);
    float distSqr = dot(d, d);
    float radius = A->radius + B->radius;
    if (distSqr > radius * radius) return;
    m.type = kManCircles;
    m.lp = A->c;
    m.ln = v2(0.0f, 0.0f);
    m.count = 1;
    m.p[0] = B->c;
    m.id[0] = 0;
}

// b2CollidePolygonAndCircle (b2collidecircle.d:59-164).
__device__ inline void collide_polygon_circle(Manifold &m, const Shape *__restrict__ A, const Xf &xfA,
                                              const Shape *__restrict__ B, const Xf &xfB) {
    m.count = 0;
    float2 cLocal = mulT(xfA, mul(xfB, B->c));
    int normalIndex = 0;
    float separation = -kMaxFloat;
    float radius = A->radius + B->radius;
    int nv = A->count;
    for (int i = 0; i < nv; ++i) {
        float s = dot(A->n[i], sub(cLocal, A->v[i]));
        if (s > radius) return;
        if (s > separation) {
            separation = s;
            normalIndex = i;
        }
    }
    int vi1 = normalIndex;
    int vi2 = vi1 + 1 < nv ? vi1 + 1 : 0;
    float2 v1 = A->v[vi1], w2 = A->v[vi2];
    m.type = kManFaceA;
    m.p[0] = B->c;
    m.id[0] = 0;
    if (separation < kEps) {
        m.count = 1;
        m.ln = A->n[normalIndex];
        m.lp = scl(0.5f, add(v1, w2));
        return;
    }
    float u1 = dot(sub(cLocal, v1), sub(w2, v1));
    float u2 = dot(sub(cLocal, w2), sub(v1, w2));
    if (u1 <= 0.0f) {
        if (dist2(cLocal, v1) > radius * radius) return;
        m.count = 1;
        m.ln = sub(cLocal, v1);
        normalize(m.ln);
        m.lp = v1;
    } else if (u2 <= 0.0f) {
        if (dist2(cLocal, w2) > radius * radius) return;
        m.count = 1;
        m.ln = sub(cLocal, w2);
        normalize(m.ln);
        m.lp = w2;
    } else {
        float2 faceCenter = scl(0.5f, add(v1, w2));
        float s2 = dot(sub(cLocal, faceCenter), A->n[vi1]);
        if (s2 > radius) return;
        m.count = 1;
        m.ln = A->n[vi1];
        m.lp = faceCenter;
    }
}

// b2CollideEdgeAndCircle (b2collideedge.d:31-160).
__device__ inline void collide_edge_circle(Manifold &m, const Shape *__restrict__ E, const Xf &xfA, const Shape *__restrict__ B,
                                           const Xf &xfB) {
    m.count = 0;
    float2 Q = mulT(xfA, mul(xfB, B->c));
    float2 A = E->v[0], Bv = E->v[1];
    float2 e = sub(Bv, A);
    float u = dot(e, sub(Bv, Q));
    float v = dot(e, sub(Q, A));
    float radius = E->radius + B->radius;
    m.p[0] = B->c;
    if (v <= 0.0f) {
        float2 d = sub(Q, A);
        if (dot(d, d) > radius * radius) return;
        if (E->flags & 1) {
            float2 e1 = sub(A, E->v[2]);
            if (dot(e1, sub(A, Q)) > 0.0f) return;
        }
        m.count = 1;
        m.type = kManCircles;
        m.ln = v2(0.0f, 0.0f);
        m.lp = A;
        m.id[0] = fkey(0, 0, kFeatVertex, kFeatVertex);
        return;
    }
    if (u <= 0.0f) {
        float2 d = sub(Q, Bv);
        if (dot(d, d) > radius * radius) return;
        if (E->flags & 2) {
            float2 e2 = sub(E->v[3], Bv);
            if (dot(e2, sub(Q, Bv)) > 0.0f) return;
        }
        m.count = 1;
        m.type = kManCircles;
        m.ln = v2(0.0f, 0.0f);
        m.lp = Bv;
        m.id[0] = fkey(1, 0, kFeatVertex, kFeatVertex);
        return;
    }
    float den = dot(e, e);
    float2 P = scl(1.0f / den, add(scl(u, A), scl(v, Bv)));
    float2 d = sub(Q, P);
    if (dot(d, d) > radius * radius) return;
    float2 n = v2(-e.y, e.x);
    if (dot(n, sub(Q, A)) < 0.0f) n = v2(-n.x, -n.y);
    normalize(n);
    m.count = 1;
    m.type = kManFaceA;
    m.ln = n;
    m.lp = A;
    m.id[0] = fkey(0, 0, kFeatFace, kFeatVertex);
}

// b2EPCollider.Collide (b2collideedge.d:217-622): edge A with adjacency against polygon B.
__device__ inline void collide_edge_polygon(Manifold &m, const Shape *__restrict__ E, const Xf &xfA, const Shape *__restrict__ P,
                                            const Xf &xfB) {
    Xf xf = mulT(xfA, xfB);
    float2 centroidB = mul(xf, P->c);
    float2 ev1 = E->v[0], ev2 = E->v[1], ev0 = E->v[2], ev3 = E->v[3];
    bool has0 = (E->flags & 1) != 0, has3 = (E->flags & 2) != 0;
    float2 edge1 = sub(ev2, ev1);
    normalize(edge1);
    float2 normal1 = v2(edge1.y, -edge1.x);
    float2 normal0 = v2(0.0f, 0.0f), normal2 = v2(0.0f, 0.0f);
    float offset1 = dot(normal1, sub(centroidB, ev1));
    float offset0 = 0.0f, offset2 = 0.0f;
    bool convex1 = false, convex2 = false;
    if (has0) {
        float2 edge0 = sub(ev1, ev0);
        normalize(edge0);
        normal0 = v2(edge0.y, -edge0.x);
        convex1 = crs(edge0, edge1) >= 0.0f;
        offset0 = dot(normal0, sub(centroidB, ev0));
    }
    if (has3) {
        float2 edge2 = sub(ev3, ev2);
        normalize(edge2);
        normal2 = v2(edge2.y, -edge2.x);
        convex2 = crs(edge1, edge2) > 0.0f;
        offset2 = dot(normal2, sub(centroidB, ev2));
    }
    // Front/back classification and normal limits (b2collideedge.d:257-403).
    bool front;
    float2 normal, lower, upper;
    float2 nn1 = neg(normal1);
    if (has0 && has3) {
        if (convex1 && convex2) {
            front = offset0 >= 0.0f || offset1 >= 0.0f || offset2 >= 0.0f;
            normal = front ? normal1 : nn1; lower = front ? normal0 : nn1; upper = front ? normal2 : nn1;
        } else if (convex1) {
            front = offset0 >= 0.0f || (offset1 >= 0.0f && offset2 >= 0.0f);
            normal = front ? normal1 : nn1; lower = front ? normal0 : neg(normal2); upper = front ? normal1 : nn1;
        } else if (convex2) {
            front = offset2 >= 0.0f || (offset0 >= 0.0f && offset1 >= 0.0f);
            normal = front ? normal1 : nn1; lower = front ? normal1 : nn1; upper = front ? normal2 : neg(normal0);
        } else {
            front = offset0 >= 0.0f && offset1 >= 0.0f && offset2 >= 0.0f;
            normal = front ? normal1 : nn1; lower = front ? normal1 : neg(normal2); upper = front ? normal1 : neg(normal0);
        }
    } else if (has0) {
        if (convex1) {
            front = offset0 >= 0.0f || offset1 >= 0.0f;
            normal = front ? normal1 : nn1; lower = front ? normal0 : normal1; upper = nn1;
        } else {
            front = offset0 >= 0.0f && offset1 >= 0.0f;
            normal = front ? normal1 : nn1; lower = normal1; upper = front ? nn1 : neg(normal0);
        }
    } else if (has3) {
        if (convex2) {
            front = offset1 >= 0.0f || offset2 >= 0.0f;
            normal = front ? normal1 : nn1; lower = nn1; upper = front ? normal2 : normal1;
        } else {
            front = offset1 >= 0.0f && offset2 >= 0.0f;
            normal = front ? normal1 : nn1; lower = front ? nn1 : neg(normal2); upper = normal1;
        }
    } else {
        front = offset1 >= 0.0f;
        normal = front ? normal1 : nn1; lower = front ? nn1 : normal1; upper = front ? nn1 : normal1;
    }
    // Polygon B in frame A: vertices and normals are transformed where used.
    const int pn = P->count;
    const float radius = 2.0f * kPolygonRadius;
    m.count = 0;
    // ComputeEdgeSeparation (b2collideedge.d:624-642)
    float edgeSep = kMaxFloat;
    for (int i = 0; i < pn; ++i) {
        float s = dot(normal, sub(mul(xf, P->v[i]), ev1));
        if (s < edgeSep) edgeSep = s;
    }
    if (edgeSep > radius) return;
    // ComputePolygonSeparation (b2collideedge.d:644-695)
    int polyType = 0, polyIndex = -1;
    float polySep = -kMaxFloat;
    float2 perp = v2(-normal.y, normal.x);
    for (int i = 0; i < pn; ++i) {
        float2 n = neg(mul(xf.q, P->n[i]));
        float2 pvi = mul(xf, P->v[i]);
        float s1 = dot(n, sub(pvi, ev1));
        float s2 = dot(n, sub(pvi, ev2));
        float s = fminr(s1, s2);
        if (s > radius) {
            polyType = 2;
            polyIndex = i;
            polySep = s;
            break;
        }
        if (dot(n, perp) >= 0.0f) {
            if (dot(sub(n, upper), normal) < -kAngularSlop) continue;
        } else {
            if (dot(sub(n, lower), normal) < -kAngularSlop) continue;
        }
        if (s > polySep) {
            polyType = 2;
            polyIndex = i;
            polySep = s;
        }
    }
    if (polyType != 0 && polySep > radius) return;
    const float k_relativeTol = 0.98f;
    const float k_absoluteTol = 0.001f;
    bool useEdge = polyType == 0 || !(polySep > k_relativeTol * edgeSep + k_absoluteTol);
    ClipV ie0, ie1;
    int rf_i1, rf_i2;
    float2 rf_v1, rf_v2, rf_normal;
    if (useEdge) {
        m.type = kManFaceA;
        int best = 0;
        float bestValue = dot(normal, mul(xf.q, P->n[0]));
        for (int i = 1; i < pn; ++i) {
            float value = dot(normal, mul(xf.q, P->n[i]));
            if (value < bestValue) {
                bestValue = value;
                best = i;
            }
        }
        int j2 = best + 1 < pn ? best + 1 : 0;
        ie0.v = mul(xf, P->v[best]);
        ie0.id = fkey(0, best, kFeatFace, kFeatVertex);
        ie1.v = mul(xf, P->v[j2]);
        ie1.id = fkey(0, j2, kFeatFace, kFeatVertex);
        rf_i1 = front ? 0 : 1;
        rf_i2 = front ? 1 : 0;
        rf_v1 = front ? ev1 : ev2;
        rf_v2 = front ? ev2 : ev1;
        rf_normal = front ? normal1 : nn1;
    } else {
        m.type = kManFaceB;
        ie0.v = ev1;
        ie0.id = fkey(0, polyIndex, kFeatVertex, kFeatFace);
        ie1.v = ev2;
        ie1.id = fkey(0, polyIndex, kFeatVertex, kFeatFace);
        rf_i1 = polyIndex;
        rf_i2 = rf_i1 + 1 < pn ? rf_i1 + 1 : 0;
        rf_v1 = mul(xf, P->v[rf_i1]);
        rf_v2 = mul(xf, P->v[rf_i2]);
        rf_normal = mul(xf.q, P->n[rf_i1]);
    }
    float2 side1 = v2(rf_normal.y, -rf_normal.x);
    float2 side2 = neg(side1);
    float off1 = dot(side1, rf_v1);
    float off2 = dot(side2, rf_v2);
    ClipV a0, a1, b0, b1;
    if (clip_segment(a0, a1, ie0, ie1, side1, off1, rf_i1) < 2) return;
    if (clip_segment(b0, b1, a0, a1, side2, off2, rf_i2) < 2) return;
    if (useEdge) {
        m.ln = rf_normal;
        m.lp = rf_v1;
    } else {
        m.ln = P->n[rf_i1];
        m.lp = P->v[rf_i1];
    }
    int pc = 0;
    if (dot(rf_normal, sub(b0.v, rf_v1)) <= radius) {
        if (useEdge) push_point(m, pc, mulT(xf, b0.v), b0.id);
        else push_point(m, pc, b0.v, fkey_swap(b0.id));
    }
    if (dot(rf_normal, sub(b1.v, rf_v1)) <= radius) {
        if (useEdge) push_point(m, pc, mulT(xf, b1.v), b1.id);
        else push_point(m, pc, b1.v, fkey_swap(b1.id));
    }
    m.count = pc;
}

// b2Contact.Evaluate overrides: the pair is already in registry order
// (circle-circle, polygon-circle, polygon-polygon, edge-circle, edge-polygon).
__device__ inline void evaluate(Manifold &m, const Shape *__restrict__ A, const Xf &xfA, const Shape *__restrict__ B,
                                const Xf &xfB) {
    int ta = A->type, tb = B->type;
    if (ta == kPolygon) {
        if (tb == kPolygon) collide_polygons(m, A, xfA, B, xfB);
        else collide_polygon_circle(m, A, xfA, B, xfB);
    } else if (ta == kCircle) {
        collide_circles(m, A, xfA, B, xfB);
    } else {
        if (tb == kCircle) collide_edge_circle(m, A, xfA, B, xfB);
        else collide_edge_polygon(m, A, xfA, B, xfB);
    }
}

// Registry order of b2Contact.Create (b2contact.d:375-400, 430-436).
__device__ __forceinline__ bool pair_has_type(int ta, int tb) { return !(ta == kEdge && tb == kEdge); }
__device__ __forceinline__ bool pair_is_primary(int ta, int tb) {
    return ta == tb || (ta == kPolygon && tb == kCircle) || ta == kEdge;
}

// b2WorldManifold.Initialize (b2collision.d:123-191).
struct WorldManifold {
    float2 normal;
    float2 points[2];
    float sep[2];
};
__device__ __forceinline__ void world_manifold(WorldManifold &w, int type, int count, float2 ln, float2 lp, float2 p0,
                                               float2 p1, const Xf &xfA, float rA, const Xf &xfB, float rB) {
    if (count == 0) return;
    if (type == kManCircles) {
        w.normal = v2(1.0f, 0.0f);
        float2 pointA = mul(xfA, lp);
        float2 pointB = mul(xfB, p0);
        if (dist2(pointA, pointB) > kEps * kEps) {
            w.normal = sub(pointB, pointA);
            normalize(w.normal);
        }
        float2 cA = add(pointA, scl(rA, w.normal));
        float2 cB = sub(pointB, scl(rB, w.normal));
        w.points[0] = scl(0.5f, add(cA, cB));
        w.sep[0] = dot(sub(cB, cA), w.normal);
    } else if (type == kManFaceA) {
        w.normal = mul(xfA.q, ln);
        float2 planePoint = mul(xfA, lp);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (i >= count) break;
            float2 clip = mul(xfB, i == 0 ? p0 : p1);
            float2 cA = add(clip, scl(rA - dot(sub(clip, planePoint), w.normal), w.normal));
            float2 cB = sub(clip, scl(rB, w.normal));
            w.points[i] = scl(0.5f, add(cA, cB));
            w.sep[i] = dot(sub(cB, cA), w.normal);
        }
    } else {
        w.normal = mul(xfB.q, ln);
        float2 planePoint = mul(xfB, lp);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (i >= count) break;
            float2 clip = mul(xfA, i == 0 ? p0 : p1);
            float2 cB = add(clip, scl(rB - dot(sub(clip, planePoint), w.normal), w.normal));
            float2 cA = sub(clip, scl(rA, w.normal));
            w.points[i] = scl(0.5f, add(cA, cB));
            w.sep[i] = dot(sub(cA, cB), w.normal);
        }
        w.normal = neg(w.normal);
    }
}

} // namespace dbx
