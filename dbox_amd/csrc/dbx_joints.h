// dbx_joints.h -- the joints of an island solve on the device (b2Joint subclasses,
// src/dbox/dynamics/joints/b2*joint.d): InitVelocityConstraints (with the joint's
// warm start), SolveVelocityConstraints and SolvePositionConstraints of the
// distance, revolute, prismatic, weld, wheel, rope, friction, motor, mouse and
// pulley joints.  Every expression keeps the reference's float operation order
// (the library is built with -ffp-contract=off).
//
// A joint's record lives in HBM for the world's lifetime and holds what the
// reference keeps in the joint object between steps: the accumulated impulses,
// the limit state and the solver's temporaries (the wheel joint reads its spring
// axis and the position solver reads the velocity init's sAy/sBy, as the
// reference does).  Body state is read from and written to the world's SoA arrays:
// bpos (sweep c, a), bvel (v, w), bmass (invMass, invI, localCenter), bparam.w (mass).
#pragma once

#include "dbx_geom.h"

namespace dbx {

enum : int { kJointRevolute = 1, kJointPrismatic = 2, kJointDistance = 3, kJointPulley = 4, kJointMouse = 5,
             kJointGear = 6, kJointWheel = 7, kJointWeld = 8, kJointFriction = 9, kJointRope = 10, kJointMotor = 11 };
enum : int { kLimitInactive = 0, kLimitLower = 1, kLimitUpper = 2, kLimitEqual = 3 }; // b2LimitState
// JointRec.flags: b2JointDef.collideConnected, enableLimit, enableMotor, and b2Joint.m_islandFlag
enum : int { kJointCollide = 1, kJointLimit = 2, kJointMotorOn = 4, kJointInIsland = 8 };
constexpr float kMaxAngularCorrection = (8.0f / 180.0f * kPi); // b2settings.d:127

struct JointRec {
    int type, bodyA, bodyB, flags;
    // definition (set on the host when the joint is created)
    float2 localAnchorA, localAnchorB, localXAxisA, localYAxisA;
    float2 linearOffset, groundAnchorA, groundAnchorB, targetA;
    float referenceAngle, length, frequencyHz, dampingRatio;
    float lower, upper, motorSpeed, maxMotor;
    float maxForce, maxTorque, angularOffset, correctionFactor;
    float lengthA, lengthB, ratio, constant;
    // kept between steps: impulse (x, y, z), motor impulse; linear impulse, angular and spring impulse
    float ix, iy, iz, motorImpulse;
    float2 linearImpulse;
    float angularImpulse, springImpulse;
    int limitState, indexA, indexB, pad0; // indexA / indexB: world-local body indices
    // solver temporaries
    float2 localCenterA, localCenterB;
    float invMassA, invMassB, invIA, invIB;
    float2 rA, rB, u, uA, uB, axis, perp, ax, ay, linearError, Cm, pad3;
    float s1, s2, a1, a2, sAx, sBx, sAy, sBy;
    float mass, motorMass, springMass, angularMass, bias, gamma, beta, curLength;
    float angularError, pad4, pad5, pad6;
    float m3[9];   // b2Mat33 (revolute, weld: m_mass; prismatic: m_K): ex.xyz, ey.xyz, ez.xyz
    float pad7[3];
    float m2[4];   // b2Mat22 (friction, motor: m_linearMass; mouse: m_mass): ex.xy, ey.xy
    // gear (b2gearjoint.d:488-527): joint1 couples C (its body A) to A (its body B),
    // joint2 D to B; the impulse is ix, the effective mass `mass`
    int bodyC, bodyD, typeA, typeB;
    float2 localAnchorC, localAnchorD, localAxisC, localAxisD, lcC, lcD, JvAC, JvBD;
    float referenceAngleA, referenceAngleB, mC, mD, iC, iD, JwA, JwB, JwC, JwD;
};

struct J3 {
    float x, y, z;
};
__device__ __forceinline__ J3 j3(float x, float y, float z) { return J3{x, y, z}; }
__device__ __forceinline__ float dot3(J3 a, J3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ J3 cross3(J3 a, J3 b) {
    return j3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// b2Mat33 on 9 floats: ex = m[0..2], ey = m[3..5], ez = m[6..8] (b2math.d:365-470)
struct M33 {
    J3 ex, ey, ez;
};
__device__ __forceinline__ M33 m33_load(const float *m) {
    return M33{j3(m[0], m[1], m[2]), j3(m[3], m[4], m[5]), j3(m[6], m[7], m[8])};
}
__device__ __forceinline__ void m33_store(float *m, const M33 &A) {
    m[0] = A.ex.x; m[1] = A.ex.y; m[2] = A.ex.z;
    m[3] = A.ey.x; m[4] = A.ey.y; m[5] = A.ey.z;
    m[6] = A.ez.x; m[7] = A.ez.y; m[8] = A.ez.z;
}
__device__ __forceinline__ J3 solve33(const M33 &A, J3 b) {
    float det = dot3(A.ex, cross3(A.ey, A.ez));
    if (det != 0.0f) det = 1.0f / det;
    return j3(det * dot3(b, cross3(A.ey, A.ez)), det * dot3(A.ex, cross3(b, A.ez)), det * dot3(A.ex, cross3(A.ey, b)));
}
__device__ __forceinline__ float2 solve22(const M33 &A, float2 b) {
    const float a11 = A.ex.x, a12 = A.ey.x, a21 = A.ex.y, a22 = A.ey.y;
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return v2(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}
__device__ __forceinline__ M33 inverse22(const M33 &A) {
    const float a = A.ex.x, b = A.ey.x, c = A.ex.y, d = A.ey.y;
    float det = a * d - b * c;
    if (det != 0.0f) det = 1.0f / det;
    return M33{j3(det * d, -det * c, 0.0f), j3(-det * b, det * a, 0.0f), j3(0.0f, 0.0f, 0.0f)};
}
__device__ __forceinline__ M33 sym_inverse33(const M33 &A) {
    float det = dot3(A.ex, cross3(A.ey, A.ez));
    if (det != 0.0f) det = 1.0f / det;
    const float a11 = A.ex.x, a12 = A.ey.x, a13 = A.ez.x, a22 = A.ey.y, a23 = A.ez.y, a33 = A.ez.z;
    M33 M;
    M.ex.x = det * (a22 * a33 - a23 * a23);
    M.ex.y = det * (a13 * a23 - a12 * a33);
    M.ex.z = det * (a12 * a23 - a13 * a22);
    M.ey.x = M.ex.y;
    M.ey.y = det * (a11 * a33 - a13 * a13);
    M.ey.z = det * (a13 * a12 - a11 * a23);
    M.ez.x = M.ex.z;
    M.ez.y = M.ey.z;
    M.ez.z = det * (a11 * a22 - a12 * a12);
    return M;
}
// b2Mul(b2Mat33, b2Vec3) = v.x ex + v.y ey + v.z ez; b2Mul22 (b2math.d:680-688)
__device__ __forceinline__ J3 mul33(const M33 &A, J3 v) {
    return j3((v.x * A.ex.x + v.y * A.ey.x) + v.z * A.ez.x, (v.x * A.ex.y + v.y * A.ey.y) + v.z * A.ez.y,
              (v.x * A.ex.z + v.y * A.ey.z) + v.z * A.ez.z);
}
__device__ __forceinline__ float2 mul22(const M33 &A, float2 v) {
    return v2(A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y);
}
// b2Mat22 on 4 floats ex.xy, ey.xy: GetInverse, Solve, b2Mul (b2math.d:278-352, 626-629)
__device__ __forceinline__ void m22_inverse(float *out, float a, float c, float b, float d) {
    float det = a * d - b * c;
    if (det != 0.0f) det = 1.0f / det;
    out[0] = det * d;
    out[2] = -det * b;
    out[1] = -det * c;
    out[3] = det * a;
}
__device__ __forceinline__ float2 m22_mul(const float *m, float2 v) {
    return v2(m[0] * v.x + m[2] * v.y, m[1] * v.x + m[3] * v.y);
}
__device__ __forceinline__ float2 m22_solve(float a11, float a21, float a12, float a22, float2 b) {
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return v2(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}
__device__ __forceinline__ float2 operator-(float2 a) { return v2(-a.x, -a.y); }
__device__ __forceinline__ float clampf(float a, float lo, float hi) { return fmaxr(lo, fminr(a, hi)); }

// The step's fields the joints read (b2TimeStep).
struct JointStep {
    float dt, inv_dt, dtRatio;
    bool warm;
};

// b2Joint.InitVelocityConstraints for the bodies' current positions and velocities.
__device__ void joint_init(JointRec &j, float4 *bpos, float4 *bvel, const float4 *bmass, const float4 *bparam,
                           const JointStep &st) {
    j.indexA = j.bodyA;
    j.indexB = j.bodyB;
    const float4 mAs = bmass[j.bodyA], mBs = bmass[j.bodyB];
    j.localCenterA = v2(mAs.z, mAs.w);
    j.localCenterB = v2(mBs.z, mBs.w);
    j.invMassA = mAs.x;
    j.invMassB = mBs.x;
    j.invIA = mAs.y;
    j.invIB = mBs.y;
    const float4 pA = bpos[j.indexA], pB = bpos[j.indexB], wVA = bvel[j.indexA], wVB = bvel[j.indexB];
    const float2 cA = v2(pA.x, pA.y), cB = v2(pB.x, pB.y);
    const float aA = pA.z, aB = pB.z;
    float2 vA = v2(wVA.x, wVA.y), vB = v2(wVB.x, wVB.y);
    float wA = wVA.z, wB = wVB.z;
    const Rot qA = rot(aA), qB = rot(aB);
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    const bool enableLimit = (j.flags & kJointLimit) != 0, enableMotor = (j.flags & kJointMotorOn) != 0;
    bool write = true;
    switch (j.type) {
    case kJointGear: { // b2gearjoint.d:245-350: all four velocities read, then written A, B, C, D
        const float4 mCs = bmass[j.bodyC], mDs = bmass[j.bodyD];
        j.lcC = v2(mCs.z, mCs.w);
        j.lcD = v2(mDs.z, mDs.w);
        j.mC = mCs.x;
        j.mD = mDs.x;
        j.iC = mCs.y;
        j.iD = mDs.y;
        const float4 pC = bpos[j.bodyC], pD = bpos[j.bodyD], wVC = bvel[j.bodyC], wVD = bvel[j.bodyD];
        float2 vC = v2(wVC.x, wVC.y), vD = v2(wVD.x, wVD.y);
        float wC = wVC.z, wD = wVD.z;
        const Rot qC = rot(pC.z), qD = rot(pD.z);
        j.mass = 0.0f;
        if (j.typeA == kJointRevolute) {
            j.JvAC = v2(0.0f, 0.0f);
            j.JwA = 1.0f;
            j.JwC = 1.0f;
            j.mass += iA + j.iC;
        } else {
            const float2 u = mul(qC, j.localAxisC);
            const float2 rC = mul(qC, sub(j.localAnchorC, j.lcC));
            const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
            j.JvAC = u;
            j.JwC = crs(rC, u);
            j.JwA = crs(rA, u);
            j.mass += j.mC + mA + j.iC * j.JwC * j.JwC + iA * j.JwA * j.JwA;
        }
        if (j.typeB == kJointRevolute) {
            j.JvBD = v2(0.0f, 0.0f);
            j.JwB = j.ratio;
            j.JwD = j.ratio;
            j.mass += j.ratio * j.ratio * (iB + j.iD);
        } else {
            const float2 u = mul(qD, j.localAxisD);
            const float2 rD = mul(qD, sub(j.localAnchorD, j.lcD));
            const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
            j.JvBD = scl(j.ratio, u);
            j.JwD = j.ratio * crs(rD, u);
            j.JwB = j.ratio * crs(rB, u);
            j.mass += j.ratio * j.ratio * (j.mD + mB) + j.iD * j.JwD * j.JwD + iB * j.JwB * j.JwB;
        }
        j.mass = j.mass > 0.0f ? 1.0f / j.mass : 0.0f;
        if (st.warm) {
            const float P = j.ix;
            vA = add(vA, scl(mA * P, j.JvAC));
            wA += iA * P * j.JwA;
            vB = add(vB, scl(mB * P, j.JvBD));
            wB += iB * P * j.JwB;
            vC = sub(vC, scl(j.mC * P, j.JvAC));
            wC -= j.iC * P * j.JwC;
            vD = sub(vD, scl(j.mD * P, j.JvBD));
            wD -= j.iD * P * j.JwD;
        } else {
            j.ix = 0.0f;
        }
        bvel[j.indexA] = make_float4(vA.x, vA.y, wA, wVA.w);
        bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
        bvel[j.bodyC] = make_float4(vC.x, vC.y, wC, wVC.w);
        bvel[j.bodyD] = make_float4(vD.x, vD.y, wD, wVD.w);
        return;
    }
    case kJointDistance: { // b2distancejoint.d:235-310
        j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        j.u = sub(sub(add(cB, j.rB), cA), j.rA);
        const float length = sqrtf(j.u.x * j.u.x + j.u.y * j.u.y);
        if (length > kLinearSlop) j.u = scl(1.0f / length, j.u);
        else j.u = v2(0.0f, 0.0f);
        const float crAu = crs(j.rA, j.u), crBu = crs(j.rB, j.u);
        float invMass = j.invMassA + j.invIA * crAu * crAu + j.invMassB + j.invIB * crBu * crBu;
        j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (j.frequencyHz > 0.0f) {
            const float C = length - j.length;
            const float omega = 2.0f * kPi * j.frequencyHz;
            const float d = 2.0f * j.mass * j.dampingRatio * omega;
            const float k = j.mass * omega * omega;
            const float h = st.dt;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invMass += j.gamma;
            j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        } else {
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        if (st.warm) {
            j.ix *= st.dtRatio;
            const float2 P = scl(j.ix, j.u);
            vA = sub(vA, scl(j.invMassA, P));
            wA -= j.invIA * crs(j.rA, P);
            vB = add(vB, scl(j.invMassB, P));
            wB += j.invIB * crs(j.rB, P);
        } else {
            j.ix = 0.0f;
        }
        break;
    }
    case kJointRevolute: { // b2revolutejoint.d:315-410
        j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const bool fixedRotation = (iA + iB == 0.0f);
        M33 M;
        M.ex.x = mA + mB + j.rA.y * j.rA.y * iA + j.rB.y * j.rB.y * iB;
        M.ey.x = -j.rA.y * j.rA.x * iA - j.rB.y * j.rB.x * iB;
        M.ez.x = -j.rA.y * iA - j.rB.y * iB;
        M.ex.y = M.ey.x;
        M.ey.y = mA + mB + j.rA.x * j.rA.x * iA + j.rB.x * j.rB.x * iB;
        M.ez.y = j.rA.x * iA + j.rB.x * iB;
        M.ex.z = M.ez.x;
        M.ey.z = M.ez.y;
        M.ez.z = iA + iB;
        m33_store(j.m3, M);
        j.motorMass = iA + iB;
        if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        if (enableMotor == false || fixedRotation) j.motorImpulse = 0.0f;
        if (enableLimit && fixedRotation == false) {
            const float jointAngle = aB - aA - j.referenceAngle;
            if (fabsr(j.upper - j.lower) < 2.0f * kAngularSlop) {
                j.limitState = kLimitEqual;
            } else if (jointAngle <= j.lower) {
                if (j.limitState != kLimitLower) j.iz = 0.0f;
                j.limitState = kLimitLower;
            } else if (jointAngle >= j.upper) {
                if (j.limitState != kLimitUpper) j.iz = 0.0f;
                j.limitState = kLimitUpper;
            } else {
                j.limitState = kLimitInactive;
                j.iz = 0.0f;
            }
        } else {
            j.limitState = kLimitInactive;
        }
        if (st.warm) {
            j.ix *= st.dtRatio;
            j.iy *= st.dtRatio;
            j.iz *= st.dtRatio;
            j.motorImpulse *= st.dtRatio;
            const float2 P = v2(j.ix, j.iy);
            vA = sub(vA, scl(mA, P));
            wA -= iA * (crs(j.rA, P) + j.motorImpulse + j.iz);
            vB = add(vB, scl(mB, P));
            wB += iB * (crs(j.rB, P) + j.motorImpulse + j.iz);
        } else {
            j.ix = j.iy = j.iz = 0.0f;
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case kJointPrismatic: { // b2prismaticjoint.d:380-500
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const float2 d = sub(add(sub(cB, cA), rB), rA);
        j.axis = mul(qA, j.localXAxisA);
        j.a1 = crs(add(d, rA), j.axis);
        j.a2 = crs(rB, j.axis);
        j.motorMass = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
        if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        j.perp = mul(qA, j.localYAxisA);
        j.s1 = crs(add(d, rA), j.perp);
        j.s2 = crs(rB, j.perp);
        {
            const float k11 = mA + mB + iA * j.s1 * j.s1 + iB * j.s2 * j.s2;
            const float k12 = iA * j.s1 + iB * j.s2;
            const float k13 = iA * j.s1 * j.a1 + iB * j.s2 * j.a2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            const float k23 = iA * j.a1 + iB * j.a2;
            const float k33 = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
            m33_store(j.m3, M33{j3(k11, k12, k13), j3(k12, k22, k23), j3(k13, k23, k33)});
        }
        if (enableLimit) {
            const float jointTranslation = dot(j.axis, d);
            if (fabsr(j.upper - j.lower) < 2.0f * kLinearSlop) {
                j.limitState = kLimitEqual;
            } else if (jointTranslation <= j.lower) {
                if (j.limitState != kLimitLower) {
                    j.limitState = kLimitLower;
                    j.iz = 0.0f;
                }
            } else if (jointTranslation >= j.upper) {
                if (j.limitState != kLimitUpper) {
                    j.limitState = kLimitUpper;
                    j.iz = 0.0f;
                }
            } else {
                j.limitState = kLimitInactive;
                j.iz = 0.0f;
            }
        } else {
            j.limitState = kLimitInactive;
            j.iz = 0.0f;
        }
        if (enableMotor == false) j.motorImpulse = 0.0f;
        if (st.warm) {
            j.ix *= st.dtRatio;
            j.iy *= st.dtRatio;
            j.iz *= st.dtRatio;
            j.motorImpulse *= st.dtRatio;
            const float2 P = add(scl(j.ix, j.perp), scl(j.motorImpulse + j.iz, j.axis));
            const float LA = j.ix * j.s1 + j.iy + (j.motorImpulse + j.iz) * j.a1;
            const float LB = j.ix * j.s2 + j.iy + (j.motorImpulse + j.iz) * j.a2;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        } else {
            j.ix = j.iy = j.iz = 0.0f;
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case kJointWeld: { // b2weldjoint.d:206-290
        j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        M33 K;
        K.ex.x = mA + mB + j.rA.y * j.rA.y * iA + j.rB.y * j.rB.y * iB;
        K.ey.x = -j.rA.y * j.rA.x * iA - j.rB.y * j.rB.x * iB;
        K.ez.x = -j.rA.y * iA - j.rB.y * iB;
        K.ex.y = K.ey.x;
        K.ey.y = mA + mB + j.rA.x * j.rA.x * iA + j.rB.x * j.rB.x * iB;
        K.ez.y = j.rA.x * iA + j.rB.x * iB;
        K.ex.z = K.ez.x;
        K.ey.z = K.ez.y;
        K.ez.z = iA + iB;
        M33 M;
        if (j.frequencyHz > 0.0f) {
            M = inverse22(K);
            float invM = iA + iB;
            const float m = invM > 0.0f ? 1.0f / invM : 0.0f;
            const float C = aB - aA - j.referenceAngle;
            const float omega = 2.0f * kPi * j.frequencyHz;
            const float d = 2.0f * m * j.dampingRatio * omega;
            const float k = m * omega * omega;
            const float h = st.dt;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invM += j.gamma;
            M.ez.z = invM != 0.0f ? 1.0f / invM : 0.0f;
        } else if (K.ez.z == 0.0f) {
            M = inverse22(K);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        } else {
            M = sym_inverse33(K);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        m33_store(j.m3, M);
        if (st.warm) {
            j.ix *= st.dtRatio;
            j.iy *= st.dtRatio;
            j.iz *= st.dtRatio;
            const float2 P = v2(j.ix, j.iy);
            vA = sub(vA, scl(mA, P));
            wA -= iA * (crs(j.rA, P) + j.iz);
            vB = add(vB, scl(mB, P));
            wB += iB * (crs(j.rB, P) + j.iz);
        } else {
            j.ix = j.iy = j.iz = 0.0f;
        }
        break;
    }
    case kJointWheel: { // b2wheeljoint.d:310-420
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const float2 d = sub(sub(add(cB, rB), cA), rA);
        j.ay = mul(qA, j.localYAxisA);
        j.sAy = crs(add(d, rA), j.ay);
        j.sBy = crs(rB, j.ay);
        j.mass = mA + mB + iA * j.sAy * j.sAy + iB * j.sBy * j.sBy;
        if (j.mass > 0.0f) j.mass = 1.0f / j.mass;
        j.springMass = 0.0f;
        j.bias = 0.0f;
        j.gamma = 0.0f;
        if (j.frequencyHz > 0.0f) {
            j.ax = mul(qA, j.localXAxisA);
            j.sAx = crs(add(d, rA), j.ax);
            j.sBx = crs(rB, j.ax);
            const float invMass = mA + mB + iA * j.sAx * j.sAx + iB * j.sBx * j.sBx;
            if (invMass > 0.0f) {
                j.springMass = 1.0f / invMass;
                const float C = dot(d, j.ax);
                const float omega = 2.0f * kPi * j.frequencyHz;
                const float d2 = 2.0f * j.springMass * j.dampingRatio * omega;
                const float k = j.springMass * omega * omega;
                const float h = st.dt;
                j.gamma = h * (d2 + h * k);
                if (j.gamma > 0.0f) j.gamma = 1.0f / j.gamma;
                j.bias = C * h * k * j.gamma;
                j.springMass = invMass + j.gamma;
                if (j.springMass > 0.0f) j.springMass = 1.0f / j.springMass;
            }
        } else {
            j.springImpulse = 0.0f;
        }
        if (enableMotor) {
            j.motorMass = iA + iB;
            if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        } else {
            j.motorMass = 0.0f;
            j.motorImpulse = 0.0f;
        }
        if (st.warm) {
            j.ix *= st.dtRatio;
            j.springImpulse *= st.dtRatio;
            j.motorImpulse *= st.dtRatio;
            const float2 P = add(scl(j.ix, j.ay), scl(j.springImpulse, j.ax));
            const float LA = j.ix * j.sAy + j.springImpulse * j.sAx + j.motorImpulse;
            const float LB = j.ix * j.sBy + j.springImpulse * j.sBx + j.motorImpulse;
            vA = sub(vA, scl(j.invMassA, P));
            wA -= j.invIA * LA;
            vB = add(vB, scl(j.invMassB, P));
            wB += j.invIB * LB;
        } else {
            j.ix = 0.0f;
            j.springImpulse = 0.0f;
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case kJointRope: { // b2ropejoint.d:190-255
        j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        j.u = sub(sub(add(cB, j.rB), cA), j.rA);
        j.curLength = sqrtf(j.u.x * j.u.x + j.u.y * j.u.y);
        const float C = j.curLength - j.length;
        j.limitState = C > 0.0f ? kLimitUpper : kLimitInactive;
        if (j.curLength > kLinearSlop) {
            j.u = scl(1.0f / j.curLength, j.u);
        } else {
            j.u = v2(0.0f, 0.0f);
            j.mass = 0.0f;
            j.ix = 0.0f;
            write = false; // the reference returns before writing the velocities back
            break;
        }
        const float crA = crs(j.rA, j.u), crB = crs(j.rB, j.u);
        const float invMass = j.invMassA + j.invIA * crA * crA + j.invMassB + j.invIB * crB * crB;
        j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (st.warm) {
            j.ix *= st.dtRatio;
            const float2 P = scl(j.ix, j.u);
            vA = sub(vA, scl(j.invMassA, P));
            wA -= j.invIA * crs(j.rA, P);
            vB = add(vB, scl(j.invMassB, P));
            wB += j.invIB * crs(j.rB, P);
        } else {
            j.ix = 0.0f;
        }
        break;
    }
    case kJointFriction:
    case kJointMotor: { // b2frictionjoint.d:190-245, b2motorjoint.d:237-300
        const bool motor = j.type == kJointMotor;
        if (motor) {
            j.rA = mul(qA, -j.localCenterA);
            j.rB = mul(qB, -j.localCenterB);
        } else {
            j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
            j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        }
        const float kxx = mA + mB + iA * j.rA.y * j.rA.y + iB * j.rB.y * j.rB.y;
        const float kxy = -iA * j.rA.x * j.rA.y - iB * j.rB.x * j.rB.y;
        const float kyy = mA + mB + iA * j.rA.x * j.rA.x + iB * j.rB.x * j.rB.x;
        m22_inverse(j.m2, kxx, kxy, kxy, kyy);
        j.angularMass = iA + iB;
        if (j.angularMass > 0.0f) j.angularMass = 1.0f / j.angularMass;
        if (motor) {
            j.linearError = sub(sub(sub(add(cB, j.rB), cA), j.rA), mul(qA, j.linearOffset));
            j.angularError = aB - aA - j.angularOffset;
        }
        if (st.warm) {
            j.linearImpulse = scl(st.dtRatio, j.linearImpulse);
            j.angularImpulse *= st.dtRatio;
            const float2 P = j.linearImpulse;
            vA = sub(vA, scl(mA, P));
            wA -= iA * (crs(j.rA, P) + j.angularImpulse);
            vB = add(vB, scl(mB, P));
            wB += iB * (crs(j.rB, P) + j.angularImpulse);
        } else {
            j.linearImpulse = v2(0.0f, 0.0f);
            j.angularImpulse = 0.0f;
        }
        break;
    }
    case kJointMouse: { // b2mousejoint.d:160-215: only body B moves
        const float massB = bparam[j.bodyB].w;
        const float omega = 2.0f * kPi * j.frequencyHz;
        const float d = 2.0f * massB * j.dampingRatio * omega;
        const float k = massB * (omega * omega);
        const float h = st.dt;
        j.gamma = h * (d + h * k);
        if (j.gamma != 0.0f) j.gamma = 1.0f / j.gamma;
        j.beta = h * k * j.gamma;
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const float kxx = j.invMassB + j.invIB * j.rB.y * j.rB.y + j.gamma;
        const float kxy = -j.invIB * j.rB.x * j.rB.y;
        const float kyy = j.invMassB + j.invIB * j.rB.x * j.rB.x + j.gamma;
        m22_inverse(j.m2, kxx, kxy, kxy, kyy);
        j.Cm = sub(add(cB, j.rB), j.targetA);
        j.Cm = scl(j.beta, j.Cm);
        wB *= 0.98f;
        if (st.warm) {
            j.linearImpulse = scl(st.dtRatio, j.linearImpulse);
            vB = add(vB, scl(j.invMassB, j.linearImpulse));
            wB += j.invIB * crs(j.rB, j.linearImpulse);
        } else {
            j.linearImpulse = v2(0.0f, 0.0f);
        }
        bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
        return;
    }
    case kJointPulley: { // b2pulleyjoint.d:240-315
        j.rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        j.rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        j.uA = sub(add(cA, j.rA), j.groundAnchorA);
        j.uB = sub(add(cB, j.rB), j.groundAnchorB);
        const float lengthA = sqrtf(j.uA.x * j.uA.x + j.uA.y * j.uA.y);
        const float lengthB = sqrtf(j.uB.x * j.uB.x + j.uB.y * j.uB.y);
        if (lengthA > 10.0f * kLinearSlop) j.uA = scl(1.0f / lengthA, j.uA);
        else j.uA = v2(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) j.uB = scl(1.0f / lengthB, j.uB);
        else j.uB = v2(0.0f, 0.0f);
        const float ruA = crs(j.rA, j.uA), ruB = crs(j.rB, j.uB);
        const float pmA = j.invMassA + j.invIA * ruA * ruA;
        const float pmB = j.invMassB + j.invIB * ruB * ruB;
        j.mass = pmA + j.ratio * j.ratio * pmB;
        if (j.mass > 0.0f) j.mass = 1.0f / j.mass;
        if (st.warm) {
            j.ix *= st.dtRatio;
            const float2 PA = scl(-(j.ix), j.uA);
            const float2 PB = scl(-j.ratio * j.ix, j.uB);
            vA = add(vA, scl(j.invMassA, PA));
            wA += j.invIA * crs(j.rA, PA);
            vB = add(vB, scl(j.invMassB, PB));
            wB += j.invIB * crs(j.rB, PB);
        } else {
            j.ix = 0.0f;
        }
        break;
    }
    default:
        write = false;
        break;
    }
    if (write) {
        bvel[j.indexA] = make_float4(vA.x, vA.y, wA, wVA.w);
        bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
    }
}

// b2Joint.SolveVelocityConstraints.
__device__ void joint_solve_velocity(JointRec &j, float4 *bvel, const JointStep &st) {
    const float4 wVA = bvel[j.indexA], wVB = bvel[j.indexB];
    float2 vA = v2(wVA.x, wVA.y), vB = v2(wVB.x, wVB.y);
    float wA = wVA.z, wB = wVB.z;
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    const bool enableLimit = (j.flags & kJointLimit) != 0, enableMotor = (j.flags & kJointMotorOn) != 0;
    switch (j.type) {
    case kJointGear: { // b2gearjoint.d:352-386
        const float4 wVC = bvel[j.bodyC], wVD = bvel[j.bodyD];
        float2 vC = v2(wVC.x, wVC.y), vD = v2(wVD.x, wVD.y);
        float wC = wVC.z, wD = wVD.z;
        float Cdot = dot(j.JvAC, sub(vA, vC)) + dot(j.JvBD, sub(vB, vD));
        Cdot += (j.JwA * wA - j.JwC * wC) + (j.JwB * wB - j.JwD * wD);
        const float impulse = -j.mass * Cdot;
        j.ix += impulse;
        vA = add(vA, scl(mA * impulse, j.JvAC));
        wA += iA * impulse * j.JwA;
        vB = add(vB, scl(mB * impulse, j.JvBD));
        wB += iB * impulse * j.JwB;
        vC = sub(vC, scl(j.mC * impulse, j.JvAC));
        wC -= j.iC * impulse * j.JwC;
        vD = sub(vD, scl(j.mD * impulse, j.JvBD));
        wD -= j.iD * impulse * j.JwD;
        bvel[j.indexA] = make_float4(vA.x, vA.y, wA, wVA.w);
        bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
        bvel[j.bodyC] = make_float4(vC.x, vC.y, wC, wVC.w);
        bvel[j.bodyD] = make_float4(vD.x, vD.y, wD, wVD.w);
        return;
    }
    case kJointDistance: { // b2distancejoint.d:312-336
        const float2 vpA = add(vA, crs(wA, j.rA)), vpB = add(vB, crs(wB, j.rB));
        const float Cdot = dot(j.u, sub(vpB, vpA));
        const float impulse = -j.mass * (Cdot + j.bias + j.gamma * j.ix);
        j.ix += impulse;
        const float2 P = scl(impulse, j.u);
        vA = sub(vA, scl(j.invMassA, P));
        wA -= j.invIA * crs(j.rA, P);
        vB = add(vB, scl(j.invMassB, P));
        wB += j.invIB * crs(j.rB, P);
        break;
    }
    case kJointRevolute: { // b2revolutejoint.d:412-515
        const bool fixedRotation = (iA + iB == 0.0f);
        if (enableMotor && j.limitState != kLimitEqual && fixedRotation == false) {
            const float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = st.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        const M33 M = m33_load(j.m3);
        if (enableLimit && j.limitState != kLimitInactive && fixedRotation == false) {
            const float2 Cdot1 = sub(sub(add(vB, crs(wB, j.rB)), vA), crs(wA, j.rA));
            const float Cdot2 = wB - wA;
            J3 impulse = solve33(M, j3(Cdot1.x, Cdot1.y, Cdot2));
            impulse = j3(-impulse.x, -impulse.y, -impulse.z);
            if (j.limitState == kLimitEqual) {
                j.ix += impulse.x;
                j.iy += impulse.y;
                j.iz += impulse.z;
            } else {
                const float newImpulse = j.iz + impulse.z;
                if ((j.limitState == kLimitLower && newImpulse < 0.0f) ||
                    (j.limitState == kLimitUpper && newImpulse > 0.0f)) {
                    const float2 rhs = add(-Cdot1, scl(j.iz, v2(M.ez.x, M.ez.y)));
                    const float2 reduced = solve22(M, rhs);
                    impulse.x = reduced.x;
                    impulse.y = reduced.y;
                    impulse.z = -j.iz;
                    j.ix += reduced.x;
                    j.iy += reduced.y;
                    j.iz = 0.0f;
                } else {
                    j.ix += impulse.x;
                    j.iy += impulse.y;
                    j.iz += impulse.z;
                }
            }
            const float2 P = v2(impulse.x, impulse.y);
            vA = sub(vA, scl(mA, P));
            wA -= iA * (crs(j.rA, P) + impulse.z);
            vB = add(vB, scl(mB, P));
            wB += iB * (crs(j.rB, P) + impulse.z);
        } else {
            const float2 Cdot = sub(sub(add(vB, crs(wB, j.rB)), vA), crs(wA, j.rA));
            const float2 impulse = solve22(M, -Cdot);
            j.ix += impulse.x;
            j.iy += impulse.y;
            vA = sub(vA, scl(mA, impulse));
            wA -= iA * crs(j.rA, impulse);
            vB = add(vB, scl(mB, impulse));
            wB += iB * crs(j.rB, impulse);
        }
        break;
    }
    case kJointPrismatic: { // b2prismaticjoint.d:502-585
        if (enableMotor && j.limitState != kLimitEqual) {
            const float Cdot = dot(j.axis, sub(vB, vA)) + j.a2 * wB - j.a1 * wA;
            float impulse = j.motorMass * (j.motorSpeed - Cdot);
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = st.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            const float2 P = scl(impulse, j.axis);
            const float LA = impulse * j.a1, LB = impulse * j.a2;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        }
        float2 Cdot1;
        Cdot1.x = dot(j.perp, sub(vB, vA)) + j.s2 * wB - j.s1 * wA;
        Cdot1.y = wB - wA;
        const M33 K = m33_load(j.m3);
        if (enableLimit && j.limitState != kLimitInactive) {
            const float Cdot2 = dot(j.axis, sub(vB, vA)) + j.a2 * wB - j.a1 * wA;
            const J3 f1 = j3(j.ix, j.iy, j.iz);
            J3 df = solve33(K, j3(-Cdot1.x, -Cdot1.y, -Cdot2));
            j.ix += df.x;
            j.iy += df.y;
            j.iz += df.z;
            if (j.limitState == kLimitLower) j.iz = fmaxr(j.iz, 0.0f);
            else if (j.limitState == kLimitUpper) j.iz = fminr(j.iz, 0.0f);
            const float2 b = sub(-Cdot1, scl(j.iz - f1.z, v2(K.ez.x, K.ez.y)));
            const float2 f2r = add(solve22(K, b), v2(f1.x, f1.y));
            j.ix = f2r.x;
            j.iy = f2r.y;
            df = j3(j.ix - f1.x, j.iy - f1.y, j.iz - f1.z);
            const float2 P = add(scl(df.x, j.perp), scl(df.z, j.axis));
            const float LA = df.x * j.s1 + df.y + df.z * j.a1;
            const float LB = df.x * j.s2 + df.y + df.z * j.a2;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        } else {
            const float2 df = solve22(K, -Cdot1);
            j.ix += df.x;
            j.iy += df.y;
            const float2 P = scl(df.x, j.perp);
            const float LA = df.x * j.s1 + df.y, LB = df.x * j.s2 + df.y;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        }
        break;
    }
    case kJointWeld: { // b2weldjoint.d:292-345
        const M33 M = m33_load(j.m3);
        if (j.frequencyHz > 0.0f) {
            const float Cdot2 = wB - wA;
            const float impulse2 = -M.ez.z * (Cdot2 + j.bias + j.gamma * j.iz);
            j.iz += impulse2;
            wA -= iA * impulse2;
            wB += iB * impulse2;
            const float2 Cdot1 = sub(sub(add(vB, crs(wB, j.rB)), vA), crs(wA, j.rA));
            const float2 impulse1 = -mul22(M, Cdot1);
            j.ix += impulse1.x;
            j.iy += impulse1.y;
            const float2 P = impulse1;
            vA = sub(vA, scl(mA, P));
            wA -= iA * crs(j.rA, P);
            vB = add(vB, scl(mB, P));
            wB += iB * crs(j.rB, P);
        } else {
            const float2 Cdot1 = sub(sub(add(vB, crs(wB, j.rB)), vA), crs(wA, j.rA));
            const float Cdot2 = wB - wA;
            J3 impulse = mul33(M, j3(Cdot1.x, Cdot1.y, Cdot2));
            impulse = j3(-impulse.x, -impulse.y, -impulse.z);
            j.ix += impulse.x;
            j.iy += impulse.y;
            j.iz += impulse.z;
            const float2 P = v2(impulse.x, impulse.y);
            vA = sub(vA, scl(mA, P));
            wA -= iA * (crs(j.rA, P) + impulse.z);
            vB = add(vB, scl(mB, P));
            wB += iB * (crs(j.rB, P) + impulse.z);
        }
        break;
    }
    case kJointWheel: { // b2wheeljoint.d:422-480
        {
            const float Cdot = dot(j.ax, sub(vB, vA)) + j.sBx * wB - j.sAx * wA;
            const float impulse = -j.springMass * (Cdot + j.bias + j.gamma * j.springImpulse);
            j.springImpulse += impulse;
            const float2 P = scl(impulse, j.ax);
            const float LA = impulse * j.sAx, LB = impulse * j.sBx;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        }
        {
            const float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = st.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            const float Cdot = dot(j.ay, sub(vB, vA)) + j.sBy * wB - j.sAy * wA;
            const float impulse = -j.mass * Cdot;
            j.ix += impulse;
            const float2 P = scl(impulse, j.ay);
            const float LA = impulse * j.sAy, LB = impulse * j.sBy;
            vA = sub(vA, scl(mA, P));
            wA -= iA * LA;
            vB = add(vB, scl(mB, P));
            wB += iB * LB;
        }
        break;
    }
    case kJointRope: { // b2ropejoint.d:257-283
        const float2 vpA = add(vA, crs(wA, j.rA)), vpB = add(vB, crs(wB, j.rB));
        const float C = j.curLength - j.length;
        float Cdot = dot(j.u, sub(vpB, vpA));
        if (C < 0.0f) Cdot += st.inv_dt * C;
        float impulse = -j.mass * Cdot;
        const float oldImpulse = j.ix;
        j.ix = fminr(0.0f, j.ix + impulse);
        impulse = j.ix - oldImpulse;
        const float2 P = scl(impulse, j.u);
        vA = sub(vA, scl(j.invMassA, P));
        wA -= j.invIA * crs(j.rA, P);
        vB = add(vB, scl(j.invMassB, P));
        wB += j.invIB * crs(j.rB, P);
        break;
    }
    case kJointFriction:
    case kJointMotor: { // b2frictionjoint.d:247-300, b2motorjoint.d:302-360
        const bool motor = j.type == kJointMotor;
        const float h = st.dt, inv_h = st.inv_dt;
        {
            float Cdot = wB - wA;
            if (motor) Cdot = wB - wA + inv_h * j.correctionFactor * j.angularError;
            float impulse = -j.angularMass * Cdot;
            const float oldImpulse = j.angularImpulse;
            const float maxImpulse = h * j.maxTorque;
            j.angularImpulse = clampf(j.angularImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.angularImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            float2 Cdot = sub(sub(add(vB, crs(wB, j.rB)), vA), crs(wA, j.rA));
            if (motor) Cdot = add(Cdot, scl(inv_h * j.correctionFactor, j.linearError));
            float2 impulse = -m22_mul(j.m2, Cdot);
            const float2 oldImpulse = j.linearImpulse;
            j.linearImpulse = add(j.linearImpulse, impulse);
            const float maxImpulse = h * j.maxForce;
            if (lensq(j.linearImpulse) > maxImpulse * maxImpulse) {
                normalize(j.linearImpulse);
                j.linearImpulse = scl(maxImpulse, j.linearImpulse);
            }
            impulse = sub(j.linearImpulse, oldImpulse);
            vA = sub(vA, scl(mA, impulse));
            wA -= iA * crs(j.rA, impulse);
            vB = add(vB, scl(mB, impulse));
            wB += iB * crs(j.rB, impulse);
        }
        break;
    }
    case kJointMouse: { // b2mousejoint.d:217-245
        const float2 Cdot = add(vB, crs(wB, j.rB));
        float2 impulse = m22_mul(j.m2, -add(add(Cdot, j.Cm), scl(j.gamma, j.linearImpulse)));
        const float2 oldImpulse = j.linearImpulse;
        j.linearImpulse = add(j.linearImpulse, impulse);
        const float maxImpulse = st.dt * j.maxForce;
        if (lensq(j.linearImpulse) > maxImpulse * maxImpulse)
            j.linearImpulse = scl(maxImpulse / sqrtf(j.linearImpulse.x * j.linearImpulse.x +
                                                     j.linearImpulse.y * j.linearImpulse.y),
                                  j.linearImpulse);
        impulse = sub(j.linearImpulse, oldImpulse);
        vB = add(vB, scl(j.invMassB, impulse));
        wB += j.invIB * crs(j.rB, impulse);
        bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
        return;
    }
    case kJointPulley: { // b2pulleyjoint.d:317-345
        const float2 vpA = add(vA, crs(wA, j.rA)), vpB = add(vB, crs(wB, j.rB));
        const float Cdot = -dot(j.uA, vpA) - j.ratio * dot(j.uB, vpB);
        const float impulse = -j.mass * Cdot;
        j.ix += impulse;
        const float2 PA = scl(-impulse, j.uA);
        const float2 PB = scl(-j.ratio * impulse, j.uB);
        vA = add(vA, scl(j.invMassA, PA));
        wA += j.invIA * crs(j.rA, PA);
        vB = add(vB, scl(j.invMassB, PB));
        wB += j.invIB * crs(j.rB, PB);
        break;
    }
    default:
        return;
    }
    bvel[j.indexA] = make_float4(vA.x, vA.y, wA, wVA.w);
    bvel[j.indexB] = make_float4(vB.x, vB.y, wB, wVB.w);
}

// b2Joint.SolvePositionConstraints; returns whether the joint's position error is
// within the slop.
__device__ bool joint_solve_position(JointRec &j, float4 *bpos) {
    const float4 pA = bpos[j.indexA], pB = bpos[j.indexB];
    float2 cA = v2(pA.x, pA.y), cB = v2(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    const bool enableLimit = (j.flags & kJointLimit) != 0;
    bool ok = true;
    switch (j.type) {
    case kJointGear: { // b2gearjoint.d:388-486
        const float4 pC = bpos[j.bodyC], pD = bpos[j.bodyD];
        float2 cC = v2(pC.x, pC.y), cD = v2(pD.x, pD.y);
        float aC = pC.z, aD = pD.z;
        const Rot qA = rot(aA), qB = rot(aB), qC = rot(aC), qD = rot(aD);
        float coordinateA = 0.0f, coordinateB = 0.0f;
        float2 JvAC = v2(0.0f, 0.0f), JvBD = v2(0.0f, 0.0f);
        float JwA = 0.0f, JwB = 0.0f, JwC = 0.0f, JwD = 0.0f;
        float mass = 0.0f;
        if (j.typeA == kJointRevolute) {
            JwA = 1.0f;
            JwC = 1.0f;
            mass += iA + j.iC;
            coordinateA = aA - aC - j.referenceAngleA;
        } else {
            const float2 u = mul(qC, j.localAxisC);
            const float2 rC = mul(qC, sub(j.localAnchorC, j.lcC));
            const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
            JvAC = u;
            JwC = crs(rC, u);
            JwA = crs(rA, u);
            mass += j.mC + mA + j.iC * JwC * JwC + iA * JwA * JwA;
            const float2 pC_ = sub(j.localAnchorC, j.lcC);
            const float2 pA_ = mulT(qC, add(rA, sub(cA, cC)));
            coordinateA = dot(sub(pA_, pC_), j.localAxisC);
        }
        if (j.typeB == kJointRevolute) {
            JwB = j.ratio;
            JwD = j.ratio;
            mass += j.ratio * j.ratio * (iB + j.iD);
            coordinateB = aB - aD - j.referenceAngleB;
        } else {
            const float2 u = mul(qD, j.localAxisD);
            const float2 rD = mul(qD, sub(j.localAnchorD, j.lcD));
            const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
            JvBD = scl(j.ratio, u);
            JwD = j.ratio * crs(rD, u);
            JwB = j.ratio * crs(rB, u);
            mass += j.ratio * j.ratio * (j.mD + mB) + j.iD * JwD * JwD + iB * JwB * JwB;
            const float2 pD_ = sub(j.localAnchorD, j.lcD);
            const float2 pB_ = mulT(qD, add(rB, sub(cB, cD)));
            coordinateB = dot(sub(pB_, pD_), j.localAxisD);
        }
        const float C = (coordinateA + j.ratio * coordinateB) - j.constant;
        float impulse = 0.0f;
        if (mass > 0.0f) impulse = -C / mass;
        cA = add(cA, scl(mA * impulse, JvAC));
        aA += iA * impulse * JwA;
        cB = add(cB, scl(mB * impulse, JvBD));
        aB += iB * impulse * JwB;
        cC = sub(cC, scl(j.mC * impulse, JvAC));
        aC -= j.iC * impulse * JwC;
        cD = sub(cD, scl(j.mD * impulse, JvBD));
        aD -= j.iD * impulse * JwD;
        bpos[j.indexA] = make_float4(cA.x, cA.y, aA, pA.w);
        bpos[j.indexB] = make_float4(cB.x, cB.y, aB, pB.w);
        bpos[j.bodyC] = make_float4(cC.x, cC.y, aC, pC.w);
        bpos[j.bodyD] = make_float4(cD.x, cD.y, aD, pD.w);
        return true; // the linear error is never computed (b2gearjoint.d:484-485)
    }
    case kJointDistance: { // b2distancejoint.d:338-371
        if (j.frequencyHz > 0.0f) return true;
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        float2 u = sub(sub(add(cB, rB), cA), rA);
        const float length = normalize(u);
        float C = length - j.length;
        C = clampf(C, -kMaxLinearCorrection, kMaxLinearCorrection);
        const float impulse = -j.mass * C;
        const float2 P = scl(impulse, u);
        cA = sub(cA, scl(j.invMassA, P));
        aA -= j.invIA * crs(rA, P);
        cB = add(cB, scl(j.invMassB, P));
        aB += j.invIB * crs(rB, P);
        ok = fabsr(C) < kLinearSlop;
        break;
    }
    case kJointRevolute: { // b2revolutejoint.d:517-590
        float angularError = 0.0f, positionError = 0.0f;
        const bool fixedRotation = (j.invIA + j.invIB == 0.0f);
        if (enableLimit && j.limitState != kLimitInactive && fixedRotation == false) {
            const float angle = aB - aA - j.referenceAngle;
            float limitImpulse = 0.0f;
            if (j.limitState == kLimitEqual) {
                const float C = clampf(angle - j.lower, -kMaxAngularCorrection, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
                angularError = fabsr(C);
            } else if (j.limitState == kLimitLower) {
                float C = angle - j.lower;
                angularError = -C;
                C = clampf(C + kAngularSlop, -kMaxAngularCorrection, 0.0f);
                limitImpulse = -j.motorMass * C;
            } else if (j.limitState == kLimitUpper) {
                float C = angle - j.upper;
                angularError = C;
                C = clampf(C - kAngularSlop, 0.0f, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
            }
            aA -= j.invIA * limitImpulse;
            aB += j.invIB * limitImpulse;
        }
        {
            const Rot qA = rot(aA), qB = rot(aB);
            const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
            const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
            const float2 C = sub(sub(add(cB, rB), cA), rA);
            positionError = sqrtf(C.x * C.x + C.y * C.y);
            const float kxx = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
            const float kxy = -iA * rA.x * rA.y - iB * rB.x * rB.y;
            const float kyy = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
            const float2 impulse = -m22_solve(kxx, kxy, kxy, kyy, C);
            cA = sub(cA, scl(mA, impulse));
            aA -= iA * crs(rA, impulse);
            cB = add(cB, scl(mB, impulse));
            aB += iB * crs(rB, impulse);
        }
        ok = positionError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case kJointPrismatic: { // b2prismaticjoint.d:587-700
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const float2 d = sub(sub(add(cB, rB), cA), rA);
        const float2 axis = mul(qA, j.localXAxisA);
        const float a1 = crs(add(d, rA), axis), a2 = crs(rB, axis);
        const float2 perp = mul(qA, j.localYAxisA);
        const float s1 = crs(add(d, rA), perp), s2 = crs(rB, perp);
        J3 impulse;
        float2 C1;
        C1.x = dot(perp, d);
        C1.y = aB - aA - j.referenceAngle;
        float linearError = fabsr(C1.x);
        const float angularError = fabsr(C1.y);
        bool active = false;
        float C2 = 0.0f;
        if (enableLimit) {
            const float translation = dot(axis, d);
            if (fabsr(j.upper - j.lower) < 2.0f * kLinearSlop) {
                C2 = clampf(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
                linearError = fmaxr(linearError, fabsr(translation));
                active = true;
            } else if (translation <= j.lower) {
                C2 = clampf(translation - j.lower + kLinearSlop, -kMaxLinearCorrection, 0.0f);
                linearError = fmaxr(linearError, j.lower - translation);
                active = true;
            } else if (translation >= j.upper) {
                C2 = clampf(translation - j.upper - kLinearSlop, 0.0f, kMaxLinearCorrection);
                linearError = fmaxr(linearError, translation - j.upper);
                active = true;
            }
        }
        if (active) {
            const float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            const float k12 = iA * s1 + iB * s2;
            const float k13 = iA * s1 * a1 + iB * s2 * a2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            const float k23 = iA * a1 + iB * a2;
            const float k33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
            impulse = solve33(M33{j3(k11, k12, k13), j3(k12, k22, k23), j3(k13, k23, k33)}, j3(-C1.x, -C1.y, -C2));
        } else {
            const float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            const float k12 = iA * s1 + iB * s2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            const float2 impulse1 = m22_solve(k11, k12, k12, k22, -C1);
            impulse = j3(impulse1.x, impulse1.y, 0.0f);
        }
        const float2 P = add(scl(impulse.x, perp), scl(impulse.z, axis));
        const float LA = impulse.x * s1 + impulse.y + impulse.z * a1;
        const float LB = impulse.x * s2 + impulse.y + impulse.z * a2;
        cA = sub(cA, scl(mA, P));
        aA -= iA * LA;
        cB = add(cB, scl(mB, P));
        aB += iB * LB;
        ok = linearError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case kJointWeld: { // b2weldjoint.d:347-415
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        float positionError, angularError;
        M33 K;
        K.ex.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
        K.ey.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
        K.ez.x = -rA.y * iA - rB.y * iB;
        K.ex.y = K.ey.x;
        K.ey.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
        K.ez.y = rA.x * iA + rB.x * iB;
        K.ex.z = K.ez.x;
        K.ey.z = K.ez.y;
        K.ez.z = iA + iB;
        if (j.frequencyHz > 0.0f) {
            const float2 C1 = sub(sub(add(cB, rB), cA), rA);
            positionError = sqrtf(C1.x * C1.x + C1.y * C1.y);
            angularError = 0.0f;
            const float2 P = -solve22(K, C1);
            cA = sub(cA, scl(mA, P));
            aA -= iA * crs(rA, P);
            cB = add(cB, scl(mB, P));
            aB += iB * crs(rB, P);
        } else {
            const float2 C1 = sub(sub(add(cB, rB), cA), rA);
            const float C2 = aB - aA - j.referenceAngle;
            positionError = sqrtf(C1.x * C1.x + C1.y * C1.y);
            angularError = fabsr(C2);
            J3 impulse;
            if (K.ez.z > 0.0f) {
                impulse = solve33(K, j3(C1.x, C1.y, C2));
                impulse = j3(-impulse.x, -impulse.y, -impulse.z);
            } else {
                const float2 impulse2 = -solve22(K, C1);
                impulse = j3(impulse2.x, impulse2.y, 0.0f);
            }
            const float2 P = v2(impulse.x, impulse.y);
            cA = sub(cA, scl(mA, P));
            aA -= iA * (crs(rA, P) + impulse.z);
            cB = add(cB, scl(mB, P));
            aB += iB * (crs(rB, P) + impulse.z);
        }
        ok = positionError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case kJointWheel: { // b2wheeljoint.d:482-530: k from the velocity init's sAy, sBy
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        const float2 d = sub(add(sub(cB, cA), rB), rA);
        const float2 ay = mul(qA, j.localYAxisA);
        const float sAy = crs(add(d, rA), ay), sBy = crs(rB, ay);
        const float C = dot(d, ay);
        const float k = j.invMassA + j.invMassB + j.invIA * j.sAy * j.sAy + j.invIB * j.sBy * j.sBy;
        float impulse;
        if (k != 0.0f) impulse = -C / k;
        else impulse = 0.0f;
        const float2 P = scl(impulse, ay);
        const float LA = impulse * sAy, LB = impulse * sBy;
        cA = sub(cA, scl(j.invMassA, P));
        aA -= j.invIA * LA;
        cB = add(cB, scl(j.invMassB, P));
        aB += j.invIB * LB;
        ok = fabsr(C) <= kLinearSlop;
        break;
    }
    case kJointRope: { // b2ropejoint.d:285-318
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        float2 u = sub(sub(add(cB, rB), cA), rA);
        const float length = normalize(u);
        float C = length - j.length;
        C = clampf(C, 0.0f, kMaxLinearCorrection);
        const float impulse = -j.mass * C;
        const float2 P = scl(impulse, u);
        cA = sub(cA, scl(j.invMassA, P));
        aA -= j.invIA * crs(rA, P);
        cB = add(cB, scl(j.invMassB, P));
        aB += j.invIB * crs(rB, P);
        ok = length - j.length < kLinearSlop;
        break;
    }
    case kJointPulley: { // b2pulleyjoint.d:347-418
        const Rot qA = rot(aA), qB = rot(aB);
        const float2 rA = mul(qA, sub(j.localAnchorA, j.localCenterA));
        const float2 rB = mul(qB, sub(j.localAnchorB, j.localCenterB));
        float2 uA = sub(add(cA, rA), j.groundAnchorA);
        float2 uB = sub(add(cB, rB), j.groundAnchorB);
        const float lengthA = sqrtf(uA.x * uA.x + uA.y * uA.y);
        const float lengthB = sqrtf(uB.x * uB.x + uB.y * uB.y);
        if (lengthA > 10.0f * kLinearSlop) uA = scl(1.0f / lengthA, uA);
        else uA = v2(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) uB = scl(1.0f / lengthB, uB);
        else uB = v2(0.0f, 0.0f);
        const float ruA = crs(rA, uA), ruB = crs(rB, uB);
        const float pmA = j.invMassA + j.invIA * ruA * ruA;
        const float pmB = j.invMassB + j.invIB * ruB * ruB;
        float mass = pmA + j.ratio * j.ratio * pmB;
        if (mass > 0.0f) mass = 1.0f / mass;
        const float C = j.constant - lengthA - j.ratio * lengthB;
        const float linearError = fabsr(C);
        const float impulse = -mass * C;
        const float2 PA = scl(-impulse, uA);
        const float2 PB = scl(-j.ratio * impulse, uB);
        cA = add(cA, scl(j.invMassA, PA));
        aA += j.invIA * crs(rA, PA);
        cB = add(cB, scl(j.invMassB, PB));
        aB += j.invIB * crs(rB, PB);
        ok = linearError < kLinearSlop;
        break;
    }
    default: // friction, motor, mouse: no position constraint
        return true;
    }
    bpos[j.indexA] = make_float4(cA.x, cA.y, aA, pA.w);
    bpos[j.indexB] = make_float4(cB.x, cB.y, aB, pB.w);
    return ok;
}

} // namespace dbx
