/*
 * b2oracle_joints.h -- the oracle's joints (TEST INFRASTRUCTURE ONLY, see b2oracle.h).
 *
 * A scalar restatement of dbox's joint solvers, keeping the reference's float
 * operation order; included into b2oracle.cpp inside namespace orc.  Citations are
 * relative to the reference's src/dbox/.  One record type holds every joint kind's
 * fields; members the reference keeps in a joint between steps (impulses, limit
 * state, and the wheel joint's spring axis) persist here too.
 *
 *   b2DistanceJoint  dynamics/joints/b2distancejoint.d:235-371
 *   b2RevoluteJoint  dynamics/joints/b2revolutejoint.d:315-601
 *   b2PrismaticJoint dynamics/joints/b2prismaticjoint.d:380-722
 *   b2WeldJoint      dynamics/joints/b2weldjoint.d:206-431
 *   b2WheelJoint     dynamics/joints/b2wheeljoint.d:310-554
 *   b2RopeJoint      dynamics/joints/b2ropejoint.d:190-318
 *   b2FrictionJoint  dynamics/joints/b2frictionjoint.d:190-322
 *   b2MotorJoint     dynamics/joints/b2motorjoint.d:237-371
 *   b2MouseJoint     dynamics/joints/b2mousejoint.d:160-297
 *   b2PulleyJoint    dynamics/joints/b2pulleyjoint.d:240-418
 * (b2GearJoint, which couples four bodies through two other joints, is not restated.)
 */

enum { j_revolute = 1, j_prismatic = 2, j_distance = 3, j_pulley = 4, j_mouse = 5, j_gear = 6, j_wheel = 7,
       j_weld = 8, j_friction = 9, j_rope = 10, j_motor = 11 };
enum { l_inactive = 0, l_atLower = 1, l_atUpper = 2, l_equal = 3 }; /* b2LimitState (b2joint.d:46-52) */
constexpr float kMaxAngularCorrection = (8.0f / 180.0f * kPi); /* b2settings.d:127 */

struct V3 {
    float x = 0, y = 0, z = 0;
};
static inline V3 mk3(float x, float y, float z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
static inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross3(V3 a, V3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

/* b2Mat33 b2math.d:365-470 */
struct Mat33 {
    V3 ex, ey, ez;
    V3 solve33(V3 b) const {
        float det = dot3(ex, cross3(ey, ez));
        if (det != 0.0f) det = 1.0f / det;
        V3 x;
        x.x = det * dot3(b, cross3(ey, ez));
        x.y = det * dot3(ex, cross3(b, ez));
        x.z = det * dot3(ex, cross3(ey, b));
        return x;
    }
    V2 solve22(V2 b) const {
        float a11 = ex.x, a12 = ey.x, a21 = ex.y, a22 = ey.y;
        float det = a11 * a22 - a12 * a21;
        if (det != 0.0f) det = 1.0f / det;
        return mk(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
    }
    void inverse22(Mat33 &M) const {
        float a = ex.x, b = ey.x, c = ex.y, d = ey.y;
        float det = a * d - b * c;
        if (det != 0.0f) det = 1.0f / det;
        M.ex.x = det * d;
        M.ey.x = -det * b;
        M.ex.z = 0.0f;
        M.ex.y = -det * c;
        M.ey.y = det * a;
        M.ey.z = 0.0f;
        M.ez.x = 0.0f;
        M.ez.y = 0.0f;
        M.ez.z = 0.0f;
    }
    void symInverse33(Mat33 &M) const {
        float det = dot3(ex, cross3(ey, ez));
        if (det != 0.0f) det = 1.0f / det;
        float a11 = ex.x, a12 = ey.x, a13 = ez.x;
        float a22 = ey.y, a23 = ez.y;
        float a33 = ez.z;
        M.ex.x = det * (a22 * a33 - a23 * a23);
        M.ex.y = det * (a13 * a23 - a12 * a33);
        M.ex.z = det * (a12 * a23 - a13 * a22);
        M.ey.x = M.ex.y;
        M.ey.y = det * (a11 * a33 - a13 * a13);
        M.ey.z = det * (a13 * a12 - a11 * a23);
        M.ez.x = M.ex.z;
        M.ez.y = M.ey.z;
        M.ez.z = det * (a11 * a22 - a12 * a12);
    }
};
/* b2Mul(b2Mat33, b2Vec3) = v.x * ex + v.y * ey + v.z * ez (b2math.d:680-683) */
static inline V3 mul33(const Mat33 &A, V3 v) {
    return mk3((v.x * A.ex.x + v.y * A.ey.x) + v.z * A.ez.x, (v.x * A.ex.y + v.y * A.ey.y) + v.z * A.ez.y,
               (v.x * A.ex.z + v.y * A.ey.z) + v.z * A.ez.z);
}
/* b2Mul22 (b2math.d:685-688) */
static inline V2 mul22(const Mat33 &A, V2 v) { return mk(A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y); }

struct JointEdge {
    int other = kNull;
    int prev = kNull, next = kNull; /* encoded edges (joint * 2 + side) */
};

struct Joint {
    int type = 0;
    int bodyA = 0, bodyB = 0;
    bool collideConnected = false, islandFlag = false;
    int prev = kNull, next = kNull;
    JointEdge node[2];
    /* definition */
    V2 localAnchorA, localAnchorB, localXAxisA, localYAxisA, linearOffset, groundAnchorA, groundAnchorB, targetA;
    float referenceAngle = 0, length = 0, frequencyHz = 0, dampingRatio = 0, lower = 0, upper = 0;
    float motorSpeed = 0, maxMotor = 0, maxForce = 0, maxTorque = 0, angularOffset = 0, correctionFactor = 0;
    float lengthA = 0, lengthB = 0, ratio = 0, constant = 0;
    bool enableLimit = false, enableMotor = false;
    /* kept between steps */
    V3 impulse;                 /* revolute, prismatic, weld: x, y, z; distance, rope, pulley, wheel: x */
    V2 linearImpulse;           /* friction, motor, mouse */
    float angularImpulse = 0, motorImpulse = 0, springImpulse = 0;
    int limitState = l_inactive;
    /* solver temporaries (some read back from an earlier step, as in the reference) */
    int indexA = 0, indexB = 0;
    V2 localCenterA, localCenterB;
    float invMassA = 0, invMassB = 0, invIA = 0, invIB = 0;
    V2 rA, rB, u, uA, uB, axis, perp, ax, ay, linearError, Cm;
    float s1 = 0, s2 = 0, a1 = 0, a2 = 0, sAx = 0, sBx = 0, sAy = 0, sBy = 0;
    float mass = 0, motorMass = 0, springMass = 0, angularMass = 0, bias = 0, gamma = 0, beta = 0;
    float curLength = 0, angularError = 0;
    Mat33 m3;                   /* revolute, weld: m_mass; prismatic: m_K */
    Mat22 m2;                   /* friction, motor: m_linearMass; mouse: m_mass */
    /* gear (b2gearjoint.d:488-527): joint1 couples C (its body A) to A (its body B), joint2 D to B */
    int joint1 = -1, joint2 = -1, typeA = 0, typeB = 0, bodyC = 0, bodyD = 0, indexC = 0, indexD = 0;
    V2 localAnchorC, localAnchorD, localAxisC, localAxisD, lcC, lcD, JvAC, JvBD;
    float referenceAngleA = 0, referenceAngleB = 0, mC = 0, mD = 0, iC = 0, iD = 0;
    float JwA = 0, JwB = 0, JwC = 0, JwD = 0;
};

struct SolverData {
    TimeStep step;
    Position *positions;
    Velocity *velocities;
};

/* b2Joint subclass constructors: the members each sets from its definition */
static void jointFromDef(Joint &j, const orc_joint_def &d, const Xf &xfB) {
    j.type = d.type;
    j.bodyA = d.body_a;
    j.bodyB = d.body_b;
    j.collideConnected = d.collide_connected != 0;
    j.localAnchorA = mk(d.local_anchor_a[0], d.local_anchor_a[1]);
    j.localAnchorB = mk(d.local_anchor_b[0], d.local_anchor_b[1]);
    j.referenceAngle = d.reference_angle;
    j.enableLimit = d.enable_limit != 0;
    j.enableMotor = d.enable_motor != 0;
    j.lower = d.lower;
    j.upper = d.upper;
    j.motorSpeed = d.motor_speed;
    j.maxMotor = d.max_motor;
    j.frequencyHz = d.frequency_hz;
    j.dampingRatio = d.damping_ratio;
    j.maxForce = d.max_force;
    j.maxTorque = d.max_torque;
    switch (d.type) {
    case j_distance: j.length = d.length; break;
    case j_rope: j.length = d.length; break; /* m_maxLength */
    case j_prismatic: /* b2prismaticjoint.d:168-190 */
        j.localXAxisA = mk(d.local_axis_a[0], d.local_axis_a[1]);
        normalize(j.localXAxisA);
        j.localYAxisA = cross(1.0f, j.localXAxisA);
        break;
    case j_wheel: /* b2wheeljoint.d:150-175: the axis is not normalised */
        j.localXAxisA = mk(d.local_axis_a[0], d.local_axis_a[1]);
        j.localYAxisA = cross(1.0f, j.localXAxisA);
        break;
    case j_motor:
        j.linearOffset = mk(d.linear_offset[0], d.linear_offset[1]);
        j.angularOffset = d.angular_offset;
        j.correctionFactor = d.correction_factor;
        break;
    case j_mouse: /* b2mousejoint.d:117-135 */
        j.targetA = mk(d.target[0], d.target[1]);
        j.localAnchorB = mulT(xfB, j.targetA);
        break;
    case j_pulley: /* b2pulleyjoint.d:140-155 */
        j.groundAnchorA = mk(d.ground_anchor_a[0], d.ground_anchor_a[1]);
        j.groundAnchorB = mk(d.ground_anchor_b[0], d.ground_anchor_b[1]);
        j.lengthA = d.length_a;
        j.lengthB = d.length_b;
        j.ratio = d.ratio;
        j.constant = d.length_a + j.ratio * d.length_b;
        break;
    default: break;
    }
}

/* b2GearJoint's constructor (b2gearjoint.d:83-166): bodies A and B become the two
 * joints' bodies B (the joint's edges follow them), C and D their bodies A, and the
 * constant is the two joint coordinates at the bodies' current transforms. */
static void gearFromDef(Joint &j, const orc_joint_def &d, const Joint &j1, const Joint &j2, const Body *bodies) {
    j.type = j_gear;
    j.collideConnected = d.collide_connected != 0;
    j.joint1 = d.joint1;
    j.joint2 = d.joint2;
    j.typeA = j1.type;
    j.typeB = j2.type;
    float coordinateA = 0.0f, coordinateB = 0.0f;
    j.bodyC = j1.bodyA;
    j.bodyA = j1.bodyB;
    {
        const Body &A = bodies[j.bodyA], &C = bodies[j.bodyC];
        j.localAnchorC = j1.localAnchorA;
        j.localAnchorA = j1.localAnchorB;
        j.referenceAngleA = j1.referenceAngle;
        if (j.typeA == j_revolute) {
            j.localAxisC = mk(0.0f, 0.0f);
            coordinateA = A.sweep.a - C.sweep.a - j.referenceAngleA;
        } else {
            j.localAxisC = j1.localXAxisA;
            V2 pC = j.localAnchorC;
            V2 pA = mulT(C.xf.q, mul(A.xf.q, j.localAnchorA) + (A.xf.p - C.xf.p));
            coordinateA = dot(pA - pC, j.localAxisC);
        }
    }
    j.bodyD = j2.bodyA;
    j.bodyB = j2.bodyB;
    {
        const Body &B = bodies[j.bodyB], &D = bodies[j.bodyD];
        j.localAnchorD = j2.localAnchorA;
        j.localAnchorB = j2.localAnchorB;
        j.referenceAngleB = j2.referenceAngle;
        if (j.typeB == j_revolute) {
            j.localAxisD = mk(0.0f, 0.0f);
            coordinateB = B.sweep.a - D.sweep.a - j.referenceAngleB;
        } else {
            j.localAxisD = j2.localXAxisA;
            V2 pD = j.localAnchorD;
            V2 pB = mulT(D.xf.q, mul(B.xf.q, j.localAnchorB) + (B.xf.p - D.xf.p));
            coordinateB = dot(pB - pD, j.localAxisD);
        }
    }
    j.ratio = d.ratio;
    j.constant = coordinateA + j.ratio * coordinateB;
    j.impulse = V3{0.0f, 0.0f, 0.0f};
}

/* The gear's bodies C and D for the island solve (b2gearjoint.d:247-262) */
static void gearBodies(Joint &j, const Body &bC, const Body &bD) {
    j.indexC = bC.islandIndex;
    j.indexD = bD.islandIndex;
    j.lcC = bC.sweep.localCenter;
    j.lcD = bD.sweep.localCenter;
    j.mC = bC.invMass;
    j.mD = bD.invMass;
    j.iC = bC.invI;
    j.iD = bD.invI;
}

static void jointBodies(Joint &j, const Body &bA, const Body &bB) {
    j.indexA = bA.islandIndex;
    j.indexB = bB.islandIndex;
    j.localCenterA = bA.sweep.localCenter;
    j.localCenterB = bB.sweep.localCenter;
    j.invMassA = bA.invMass;
    j.invMassB = bB.invMass;
    j.invIA = bA.invI;
    j.invIB = bB.invI;
}

static void jointInitVelocity(Joint &j, const Body &bA, const Body &bB, SolverData &data) {
    jointBodies(j, bA, bB);
    const TimeStep &step = data.step;
    V2 cA = data.positions[j.indexA].c;
    float aA = data.positions[j.indexA].a;
    V2 vA = data.velocities[j.indexA].v;
    float wA = data.velocities[j.indexA].w;
    V2 cB = data.positions[j.indexB].c;
    float aB = data.positions[j.indexB].a;
    V2 vB = data.velocities[j.indexB].v;
    float wB = data.velocities[j.indexB].w;
    Rot qA, qB;
    qA.set(aA);
    qB.set(aB);
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    switch (j.type) {
    case j_gear: { /* b2gearjoint.d:245-350: all four velocities read, then written A, B, C, D */
        float aC = data.positions[j.indexC].a;
        V2 vC = data.velocities[j.indexC].v;
        float wC = data.velocities[j.indexC].w;
        float aD = data.positions[j.indexD].a;
        V2 vD = data.velocities[j.indexD].v;
        float wD = data.velocities[j.indexD].w;
        Rot qC, qD;
        qC.set(aC);
        qD.set(aD);
        j.mass = 0.0f;
        if (j.typeA == j_revolute) {
            j.JvAC = mk(0.0f, 0.0f);
            j.JwA = 1.0f;
            j.JwC = 1.0f;
            j.mass += iA + j.iC;
        } else {
            V2 u = mul(qC, j.localAxisC);
            V2 rC = mul(qC, j.localAnchorC - j.lcC);
            V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
            j.JvAC = u;
            j.JwC = cross(rC, u);
            j.JwA = cross(rA, u);
            j.mass += j.mC + mA + j.iC * j.JwC * j.JwC + iA * j.JwA * j.JwA;
        }
        if (j.typeB == j_revolute) {
            j.JvBD = mk(0.0f, 0.0f);
            j.JwB = j.ratio;
            j.JwD = j.ratio;
            j.mass += j.ratio * j.ratio * (iB + j.iD);
        } else {
            V2 u = mul(qD, j.localAxisD);
            V2 rD = mul(qD, j.localAnchorD - j.lcD);
            V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
            j.JvBD = j.ratio * u;
            j.JwD = j.ratio * cross(rD, u);
            j.JwB = j.ratio * cross(rB, u);
            j.mass += j.ratio * j.ratio * (j.mD + mB) + j.iD * j.JwD * j.JwD + iB * j.JwB * j.JwB;
        }
        j.mass = j.mass > 0.0f ? 1.0f / j.mass : 0.0f;
        if (step.warmStarting) {
            const float P = j.impulse.x;
            vA += (mA * P) * j.JvAC;
            wA += iA * P * j.JwA;
            vB += (mB * P) * j.JvBD;
            wB += iB * P * j.JwB;
            vC -= (j.mC * P) * j.JvAC;
            wC -= j.iC * P * j.JwC;
            vD -= (j.mD * P) * j.JvBD;
            wD -= j.iD * P * j.JwD;
        } else {
            j.impulse.x = 0.0f;
        }
        data.velocities[j.indexA].v = vA;
        data.velocities[j.indexA].w = wA;
        data.velocities[j.indexB].v = vB;
        data.velocities[j.indexB].w = wB;
        data.velocities[j.indexC].v = vC;
        data.velocities[j.indexC].w = wC;
        data.velocities[j.indexD].v = vD;
        data.velocities[j.indexD].w = wD;
        return;
    }
    case j_distance: { /* b2distancejoint.d:235-310 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.u = cB + j.rB - cA - j.rA;
        float length = sqrtf(j.u.x * j.u.x + j.u.y * j.u.y);
        if (length > kLinearSlop) j.u *= 1.0f / length;
        else j.u = mk(0.0f, 0.0f);
        float crAu = cross(j.rA, j.u);
        float crBu = cross(j.rB, j.u);
        float invMass = j.invMassA + j.invIA * crAu * crAu + j.invMassB + j.invIB * crBu * crBu;
        j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (j.frequencyHz > 0.0f) {
            float C = length - j.length;
            float omega = 2.0f * kPi * j.frequencyHz;
            float d = 2.0f * j.mass * j.dampingRatio * omega;
            float k = j.mass * omega * omega;
            float h = step.dt;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invMass += j.gamma;
            j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        } else {
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            V2 P = j.impulse.x * j.u;
            vA -= j.invMassA * P;
            wA -= j.invIA * cross(j.rA, P);
            vB += j.invMassB * P;
            wB += j.invIB * cross(j.rB, P);
        } else {
            j.impulse.x = 0.0f;
        }
        break;
    }
    case j_revolute: { /* b2revolutejoint.d:315-410 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        bool fixedRotation = (iA + iB == 0.0f);
        j.m3.ex.x = mA + mB + j.rA.y * j.rA.y * iA + j.rB.y * j.rB.y * iB;
        j.m3.ey.x = -j.rA.y * j.rA.x * iA - j.rB.y * j.rB.x * iB;
        j.m3.ez.x = -j.rA.y * iA - j.rB.y * iB;
        j.m3.ex.y = j.m3.ey.x;
        j.m3.ey.y = mA + mB + j.rA.x * j.rA.x * iA + j.rB.x * j.rB.x * iB;
        j.m3.ez.y = j.rA.x * iA + j.rB.x * iB;
        j.m3.ex.z = j.m3.ez.x;
        j.m3.ey.z = j.m3.ez.y;
        j.m3.ez.z = iA + iB;
        j.motorMass = iA + iB;
        if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        if (j.enableMotor == false || fixedRotation) j.motorImpulse = 0.0f;
        if (j.enableLimit && fixedRotation == false) {
            float jointAngle = aB - aA - j.referenceAngle;
            if (fabs_(j.upper - j.lower) < 2.0f * kAngularSlop) {
                j.limitState = l_equal;
            } else if (jointAngle <= j.lower) {
                if (j.limitState != l_atLower) j.impulse.z = 0.0f;
                j.limitState = l_atLower;
            } else if (jointAngle >= j.upper) {
                if (j.limitState != l_atUpper) j.impulse.z = 0.0f;
                j.limitState = l_atUpper;
            } else {
                j.limitState = l_inactive;
                j.impulse.z = 0.0f;
            }
        } else {
            j.limitState = l_inactive;
        }
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            j.impulse.y *= step.dtRatio;
            j.impulse.z *= step.dtRatio;
            j.motorImpulse *= step.dtRatio;
            V2 P = mk(j.impulse.x, j.impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.motorImpulse + j.impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.motorImpulse + j.impulse.z);
        } else {
            j.impulse = V3();
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case j_prismatic: { /* b2prismaticjoint.d:380-500 */
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 d = (cB - cA) + rB - rA;
        {
            j.axis = mul(qA, j.localXAxisA);
            j.a1 = cross(d + rA, j.axis);
            j.a2 = cross(rB, j.axis);
            j.motorMass = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
            if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        }
        {
            j.perp = mul(qA, j.localYAxisA);
            j.s1 = cross(d + rA, j.perp);
            j.s2 = cross(rB, j.perp);
            float k11 = mA + mB + iA * j.s1 * j.s1 + iB * j.s2 * j.s2;
            float k12 = iA * j.s1 + iB * j.s2;
            float k13 = iA * j.s1 * j.a1 + iB * j.s2 * j.a2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            float k23 = iA * j.a1 + iB * j.a2;
            float k33 = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
            j.m3.ex = mk3(k11, k12, k13);
            j.m3.ey = mk3(k12, k22, k23);
            j.m3.ez = mk3(k13, k23, k33);
        }
        if (j.enableLimit) {
            float jointTranslation = dot(j.axis, d);
            if (fabs_(j.upper - j.lower) < 2.0f * kLinearSlop) {
                j.limitState = l_equal;
            } else if (jointTranslation <= j.lower) {
                if (j.limitState != l_atLower) {
                    j.limitState = l_atLower;
                    j.impulse.z = 0.0f;
                }
            } else if (jointTranslation >= j.upper) {
                if (j.limitState != l_atUpper) {
                    j.limitState = l_atUpper;
                    j.impulse.z = 0.0f;
                }
            } else {
                j.limitState = l_inactive;
                j.impulse.z = 0.0f;
            }
        } else {
            j.limitState = l_inactive;
            j.impulse.z = 0.0f;
        }
        if (j.enableMotor == false) j.motorImpulse = 0.0f;
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            j.impulse.y *= step.dtRatio;
            j.impulse.z *= step.dtRatio;
            j.motorImpulse *= step.dtRatio;
            V2 P = j.impulse.x * j.perp + (j.motorImpulse + j.impulse.z) * j.axis;
            float LA = j.impulse.x * j.s1 + j.impulse.y + (j.motorImpulse + j.impulse.z) * j.a1;
            float LB = j.impulse.x * j.s2 + j.impulse.y + (j.motorImpulse + j.impulse.z) * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        } else {
            j.impulse = V3();
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case j_weld: { /* b2weldjoint.d:206-290 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        Mat33 K;
        K.ex.x = mA + mB + j.rA.y * j.rA.y * iA + j.rB.y * j.rB.y * iB;
        K.ey.x = -j.rA.y * j.rA.x * iA - j.rB.y * j.rB.x * iB;
        K.ez.x = -j.rA.y * iA - j.rB.y * iB;
        K.ex.y = K.ey.x;
        K.ey.y = mA + mB + j.rA.x * j.rA.x * iA + j.rB.x * j.rB.x * iB;
        K.ez.y = j.rA.x * iA + j.rB.x * iB;
        K.ex.z = K.ez.x;
        K.ey.z = K.ez.y;
        K.ez.z = iA + iB;
        if (j.frequencyHz > 0.0f) {
            K.inverse22(j.m3);
            float invM = iA + iB;
            float m = invM > 0.0f ? 1.0f / invM : 0.0f;
            float C = aB - aA - j.referenceAngle;
            float omega = 2.0f * kPi * j.frequencyHz;
            float d = 2.0f * m * j.dampingRatio * omega;
            float k = m * omega * omega;
            float h = step.dt;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invM += j.gamma;
            j.m3.ez.z = invM != 0.0f ? 1.0f / invM : 0.0f;
        } else if (K.ez.z == 0.0f) {
            K.inverse22(j.m3);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        } else {
            K.symInverse33(j.m3);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            j.impulse.y *= step.dtRatio;
            j.impulse.z *= step.dtRatio;
            V2 P = mk(j.impulse.x, j.impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.impulse.z);
        } else {
            j.impulse = V3();
        }
        break;
    }
    case j_wheel: { /* b2wheeljoint.d:310-420 */
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 d = cB + rB - cA - rA;
        {
            j.ay = mul(qA, j.localYAxisA);
            j.sAy = cross(d + rA, j.ay);
            j.sBy = cross(rB, j.ay);
            j.mass = mA + mB + iA * j.sAy * j.sAy + iB * j.sBy * j.sBy;
            if (j.mass > 0.0f) j.mass = 1.0f / j.mass;
        }
        j.springMass = 0.0f;
        j.bias = 0.0f;
        j.gamma = 0.0f;
        if (j.frequencyHz > 0.0f) {
            j.ax = mul(qA, j.localXAxisA);
            j.sAx = cross(d + rA, j.ax);
            j.sBx = cross(rB, j.ax);
            float invMass = mA + mB + iA * j.sAx * j.sAx + iB * j.sBx * j.sBx;
            if (invMass > 0.0f) {
                j.springMass = 1.0f / invMass;
                float C = dot(d, j.ax);
                float omega = 2.0f * kPi * j.frequencyHz;
                float d2 = 2.0f * j.springMass * j.dampingRatio * omega;
                float k = j.springMass * omega * omega;
                float h = step.dt;
                j.gamma = h * (d2 + h * k);
                if (j.gamma > 0.0f) j.gamma = 1.0f / j.gamma;
                j.bias = C * h * k * j.gamma;
                j.springMass = invMass + j.gamma;
                if (j.springMass > 0.0f) j.springMass = 1.0f / j.springMass;
            }
        } else {
            j.springImpulse = 0.0f;
        }
        if (j.enableMotor) {
            j.motorMass = iA + iB;
            if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        } else {
            j.motorMass = 0.0f;
            j.motorImpulse = 0.0f;
        }
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            j.springImpulse *= step.dtRatio;
            j.motorImpulse *= step.dtRatio;
            V2 P = j.impulse.x * j.ay + j.springImpulse * j.ax;
            float LA = j.impulse.x * j.sAy + j.springImpulse * j.sAx + j.motorImpulse;
            float LB = j.impulse.x * j.sBy + j.springImpulse * j.sBx + j.motorImpulse;
            vA -= j.invMassA * P;
            wA -= j.invIA * LA;
            vB += j.invMassB * P;
            wB += j.invIB * LB;
        } else {
            j.impulse.x = 0.0f;
            j.springImpulse = 0.0f;
            j.motorImpulse = 0.0f;
        }
        break;
    }
    case j_rope: { /* b2ropejoint.d:190-255 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.u = cB + j.rB - cA - j.rA;
        j.curLength = sqrtf(j.u.x * j.u.x + j.u.y * j.u.y);
        float C = j.curLength - j.length;
        j.limitState = C > 0.0f ? l_atUpper : l_inactive;
        if (j.curLength > kLinearSlop) {
            j.u *= 1.0f / j.curLength;
        } else {
            j.u = mk(0.0f, 0.0f);
            j.mass = 0.0f;
            j.impulse.x = 0.0f;
            return; /* the velocities are left as they are */
        }
        float crA = cross(j.rA, j.u);
        float crB = cross(j.rB, j.u);
        float invMass = j.invMassA + j.invIA * crA * crA + j.invMassB + j.invIB * crB * crB;
        j.mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            V2 P = j.impulse.x * j.u;
            vA -= j.invMassA * P;
            wA -= j.invIA * cross(j.rA, P);
            vB += j.invMassB * P;
            wB += j.invIB * cross(j.rB, P);
        } else {
            j.impulse.x = 0.0f;
        }
        break;
    }
    case j_friction: { /* b2frictionjoint.d:190-245 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        Mat22 K;
        K.ex.x = mA + mB + iA * j.rA.y * j.rA.y + iB * j.rB.y * j.rB.y;
        K.ex.y = -iA * j.rA.x * j.rA.y - iB * j.rB.x * j.rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = mA + mB + iA * j.rA.x * j.rA.x + iB * j.rB.x * j.rB.x;
        j.m2 = K.inverse();
        j.angularMass = iA + iB;
        if (j.angularMass > 0.0f) j.angularMass = 1.0f / j.angularMass;
        if (step.warmStarting) {
            j.linearImpulse *= step.dtRatio;
            j.angularImpulse *= step.dtRatio;
            V2 P = mk(j.linearImpulse.x, j.linearImpulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.angularImpulse);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.angularImpulse);
        } else {
            j.linearImpulse = mk(0.0f, 0.0f);
            j.angularImpulse = 0.0f;
        }
        break;
    }
    case j_motor: { /* b2motorjoint.d:237-300 */
        j.rA = mul(qA, -j.localCenterA);
        j.rB = mul(qB, -j.localCenterB);
        Mat22 K;
        K.ex.x = mA + mB + iA * j.rA.y * j.rA.y + iB * j.rB.y * j.rB.y;
        K.ex.y = -iA * j.rA.x * j.rA.y - iB * j.rB.x * j.rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = mA + mB + iA * j.rA.x * j.rA.x + iB * j.rB.x * j.rB.x;
        j.m2 = K.inverse();
        j.angularMass = iA + iB;
        if (j.angularMass > 0.0f) j.angularMass = 1.0f / j.angularMass;
        j.linearError = cB + j.rB - cA - j.rA - mul(qA, j.linearOffset);
        j.angularError = aB - aA - j.angularOffset;
        if (step.warmStarting) {
            j.linearImpulse *= step.dtRatio;
            j.angularImpulse *= step.dtRatio;
            V2 P = mk(j.linearImpulse.x, j.linearImpulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.angularImpulse);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.angularImpulse);
        } else {
            j.linearImpulse = mk(0.0f, 0.0f);
            j.angularImpulse = 0.0f;
        }
        break;
    }
    case j_mouse: { /* b2mousejoint.d:160-215: only body B moves */
        const float massB = bB.mass;
        float omega = 2.0f * kPi * j.frequencyHz;
        float d = 2.0f * massB * j.dampingRatio * omega;
        float k = massB * (omega * omega);
        float h = step.dt;
        j.gamma = h * (d + h * k);
        if (j.gamma != 0.0f) j.gamma = 1.0f / j.gamma;
        j.beta = h * k * j.gamma;
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        Mat22 K;
        K.ex.x = j.invMassB + j.invIB * j.rB.y * j.rB.y + j.gamma;
        K.ex.y = -j.invIB * j.rB.x * j.rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = j.invMassB + j.invIB * j.rB.x * j.rB.x + j.gamma;
        j.m2 = K.inverse();
        j.Cm = cB + j.rB - j.targetA;
        j.Cm *= j.beta;
        wB *= 0.98f;
        if (step.warmStarting) {
            j.linearImpulse *= step.dtRatio;
            vB += j.invMassB * j.linearImpulse;
            wB += j.invIB * cross(j.rB, j.linearImpulse);
        } else {
            j.linearImpulse = mk(0.0f, 0.0f);
        }
        data.velocities[j.indexB].v = vB;
        data.velocities[j.indexB].w = wB;
        return;
    }
    case j_pulley: { /* b2pulleyjoint.d:240-315 */
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.uA = cA + j.rA - j.groundAnchorA;
        j.uB = cB + j.rB - j.groundAnchorB;
        float lengthA = sqrtf(j.uA.x * j.uA.x + j.uA.y * j.uA.y);
        float lengthB = sqrtf(j.uB.x * j.uB.x + j.uB.y * j.uB.y);
        if (lengthA > 10.0f * kLinearSlop) j.uA *= 1.0f / lengthA;
        else j.uA = mk(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) j.uB *= 1.0f / lengthB;
        else j.uB = mk(0.0f, 0.0f);
        float ruA = cross(j.rA, j.uA);
        float ruB = cross(j.rB, j.uB);
        float pmA = j.invMassA + j.invIA * ruA * ruA;
        float pmB = j.invMassB + j.invIB * ruB * ruB;
        j.mass = pmA + j.ratio * j.ratio * pmB;
        if (j.mass > 0.0f) j.mass = 1.0f / j.mass;
        if (step.warmStarting) {
            j.impulse.x *= step.dtRatio;
            V2 PA = -(j.impulse.x) * j.uA;
            V2 PB = (-j.ratio * j.impulse.x) * j.uB;
            vA += j.invMassA * PA;
            wA += j.invIA * cross(j.rA, PA);
            vB += j.invMassB * PB;
            wB += j.invIB * cross(j.rB, PB);
        } else {
            j.impulse.x = 0.0f;
        }
        break;
    }
    default: break;
    }
    data.velocities[j.indexA].v = vA;
    data.velocities[j.indexA].w = wA;
    data.velocities[j.indexB].v = vB;
    data.velocities[j.indexB].w = wB;
}

static void jointSolveVelocity(Joint &j, SolverData &data) {
    const TimeStep &step = data.step;
    V2 vA = data.velocities[j.indexA].v;
    float wA = data.velocities[j.indexA].w;
    V2 vB = data.velocities[j.indexB].v;
    float wB = data.velocities[j.indexB].w;
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    switch (j.type) {
    case j_gear: { /* b2gearjoint.d:352-386 */
        V2 vC = data.velocities[j.indexC].v;
        float wC = data.velocities[j.indexC].w;
        V2 vD = data.velocities[j.indexD].v;
        float wD = data.velocities[j.indexD].w;
        float Cdot = dot(j.JvAC, vA - vC) + dot(j.JvBD, vB - vD);
        Cdot += (j.JwA * wA - j.JwC * wC) + (j.JwB * wB - j.JwD * wD);
        float impulse = -j.mass * Cdot;
        j.impulse.x += impulse;
        vA += (mA * impulse) * j.JvAC;
        wA += iA * impulse * j.JwA;
        vB += (mB * impulse) * j.JvBD;
        wB += iB * impulse * j.JwB;
        vC -= (j.mC * impulse) * j.JvAC;
        wC -= j.iC * impulse * j.JwC;
        vD -= (j.mD * impulse) * j.JvBD;
        wD -= j.iD * impulse * j.JwD;
        data.velocities[j.indexA].v = vA;
        data.velocities[j.indexA].w = wA;
        data.velocities[j.indexB].v = vB;
        data.velocities[j.indexB].w = wB;
        data.velocities[j.indexC].v = vC;
        data.velocities[j.indexC].w = wC;
        data.velocities[j.indexD].v = vD;
        data.velocities[j.indexD].w = wD;
        return;
    }
    case j_distance: { /* b2distancejoint.d:312-336 */
        V2 vpA = vA + cross(wA, j.rA);
        V2 vpB = vB + cross(wB, j.rB);
        float Cdot = dot(j.u, vpB - vpA);
        float impulse = -j.mass * (Cdot + j.bias + j.gamma * j.impulse.x);
        j.impulse.x += impulse;
        V2 P = impulse * j.u;
        vA -= j.invMassA * P;
        wA -= j.invIA * cross(j.rA, P);
        vB += j.invMassB * P;
        wB += j.invIB * cross(j.rB, P);
        break;
    }
    case j_revolute: { /* b2revolutejoint.d:412-515 */
        bool fixedRotation = (iA + iB == 0.0f);
        if (j.enableMotor && j.limitState != l_equal && fixedRotation == false) {
            float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            float oldImpulse = j.motorImpulse;
            float maxImpulse = step.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        if (j.enableLimit && j.limitState != l_inactive && fixedRotation == false) {
            V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            float Cdot2 = wB - wA;
            V3 Cdot = mk3(Cdot1.x, Cdot1.y, Cdot2);
            V3 impulse = j.m3.solve33(Cdot);
            impulse = mk3(-impulse.x, -impulse.y, -impulse.z);
            if (j.limitState == l_equal) {
                j.impulse = mk3(j.impulse.x + impulse.x, j.impulse.y + impulse.y, j.impulse.z + impulse.z);
            } else if (j.limitState == l_atLower || j.limitState == l_atUpper) {
                float newImpulse = j.impulse.z + impulse.z;
                if ((j.limitState == l_atLower && newImpulse < 0.0f) || (j.limitState == l_atUpper && newImpulse > 0.0f)) {
                    V2 rhs = -Cdot1 + j.impulse.z * mk(j.m3.ez.x, j.m3.ez.y);
                    V2 reduced = j.m3.solve22(rhs);
                    impulse.x = reduced.x;
                    impulse.y = reduced.y;
                    impulse.z = -j.impulse.z;
                    j.impulse.x += reduced.x;
                    j.impulse.y += reduced.y;
                    j.impulse.z = 0.0f;
                } else {
                    j.impulse = mk3(j.impulse.x + impulse.x, j.impulse.y + impulse.y, j.impulse.z + impulse.z);
                }
            }
            V2 P = mk(impulse.x, impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + impulse.z);
        } else {
            V2 Cdot = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            V2 impulse = j.m3.solve22(-Cdot);
            j.impulse.x += impulse.x;
            j.impulse.y += impulse.y;
            vA -= mA * impulse;
            wA -= iA * cross(j.rA, impulse);
            vB += mB * impulse;
            wB += iB * cross(j.rB, impulse);
        }
        break;
    }
    case j_prismatic: { /* b2prismaticjoint.d:502-585 */
        if (j.enableMotor && j.limitState != l_equal) {
            float Cdot = dot(j.axis, vB - vA) + j.a2 * wB - j.a1 * wA;
            float impulse = j.motorMass * (j.motorSpeed - Cdot);
            float oldImpulse = j.motorImpulse;
            float maxImpulse = step.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            V2 P = impulse * j.axis;
            float LA = impulse * j.a1;
            float LB = impulse * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        V2 Cdot1;
        Cdot1.x = dot(j.perp, vB - vA) + j.s2 * wB - j.s1 * wA;
        Cdot1.y = wB - wA;
        if (j.enableLimit && j.limitState != l_inactive) {
            float Cdot2 = dot(j.axis, vB - vA) + j.a2 * wB - j.a1 * wA;
            V3 Cdot = mk3(Cdot1.x, Cdot1.y, Cdot2);
            V3 f1 = j.impulse;
            V3 df = j.m3.solve33(mk3(-Cdot.x, -Cdot.y, -Cdot.z));
            j.impulse = mk3(j.impulse.x + df.x, j.impulse.y + df.y, j.impulse.z + df.z);
            if (j.limitState == l_atLower) j.impulse.z = fmax_(j.impulse.z, 0.0f);
            else if (j.limitState == l_atUpper) j.impulse.z = fmin_(j.impulse.z, 0.0f);
            V2 b = -Cdot1 - (j.impulse.z - f1.z) * mk(j.m3.ez.x, j.m3.ez.y);
            V2 f2r = j.m3.solve22(b) + mk(f1.x, f1.y);
            j.impulse.x = f2r.x;
            j.impulse.y = f2r.y;
            df = mk3(j.impulse.x - f1.x, j.impulse.y - f1.y, j.impulse.z - f1.z);
            V2 P = df.x * j.perp + df.z * j.axis;
            float LA = df.x * j.s1 + df.y + df.z * j.a1;
            float LB = df.x * j.s2 + df.y + df.z * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        } else {
            V2 df = j.m3.solve22(-Cdot1);
            j.impulse.x += df.x;
            j.impulse.y += df.y;
            V2 P = df.x * j.perp;
            float LA = df.x * j.s1 + df.y;
            float LB = df.x * j.s2 + df.y;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        break;
    }
    case j_weld: { /* b2weldjoint.d:292-345 */
        if (j.frequencyHz > 0.0f) {
            float Cdot2 = wB - wA;
            float impulse2 = -j.m3.ez.z * (Cdot2 + j.bias + j.gamma * j.impulse.z);
            j.impulse.z += impulse2;
            wA -= iA * impulse2;
            wB += iB * impulse2;
            V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            V2 impulse1 = -mul22(j.m3, Cdot1);
            j.impulse.x += impulse1.x;
            j.impulse.y += impulse1.y;
            V2 P = impulse1;
            vA -= mA * P;
            wA -= iA * cross(j.rA, P);
            vB += mB * P;
            wB += iB * cross(j.rB, P);
        } else {
            V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            float Cdot2 = wB - wA;
            V3 Cdot = mk3(Cdot1.x, Cdot1.y, Cdot2);
            V3 impulse = mul33(j.m3, Cdot);
            impulse = mk3(-impulse.x, -impulse.y, -impulse.z);
            j.impulse = mk3(j.impulse.x + impulse.x, j.impulse.y + impulse.y, j.impulse.z + impulse.z);
            V2 P = mk(impulse.x, impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + impulse.z);
        }
        break;
    }
    case j_wheel: { /* b2wheeljoint.d:422-480 */
        {
            float Cdot = dot(j.ax, vB - vA) + j.sBx * wB - j.sAx * wA;
            float impulse = -j.springMass * (Cdot + j.bias + j.gamma * j.springImpulse);
            j.springImpulse += impulse;
            V2 P = impulse * j.ax;
            float LA = impulse * j.sAx;
            float LB = impulse * j.sBx;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        {
            float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            float oldImpulse = j.motorImpulse;
            float maxImpulse = step.dt * j.maxMotor;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            float Cdot = dot(j.ay, vB - vA) + j.sBy * wB - j.sAy * wA;
            float impulse = -j.mass * Cdot;
            j.impulse.x += impulse;
            V2 P = impulse * j.ay;
            float LA = impulse * j.sAy;
            float LB = impulse * j.sBy;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        break;
    }
    case j_rope: { /* b2ropejoint.d:257-283 */
        V2 vpA = vA + cross(wA, j.rA);
        V2 vpB = vB + cross(wB, j.rB);
        float C = j.curLength - j.length;
        float Cdot = dot(j.u, vpB - vpA);
        if (C < 0.0f) Cdot += step.inv_dt * C;
        float impulse = -j.mass * Cdot;
        float oldImpulse = j.impulse.x;
        j.impulse.x = fmin_(0.0f, j.impulse.x + impulse);
        impulse = j.impulse.x - oldImpulse;
        V2 P = impulse * j.u;
        vA -= j.invMassA * P;
        wA -= j.invIA * cross(j.rA, P);
        vB += j.invMassB * P;
        wB += j.invIB * cross(j.rB, P);
        break;
    }
    case j_friction:
    case j_motor: { /* b2frictionjoint.d:247-300, b2motorjoint.d:302-360 */
        const bool motor = j.type == j_motor;
        float h = step.dt;
        float inv_h = step.inv_dt;
        {
            float Cdot = wB - wA;
            if (motor) Cdot = wB - wA + inv_h * j.correctionFactor * j.angularError;
            float impulse = -j.angularMass * Cdot;
            float oldImpulse = j.angularImpulse;
            float maxImpulse = h * j.maxTorque;
            j.angularImpulse = clampf(j.angularImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.angularImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            V2 Cdot = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            if (motor) Cdot = Cdot + (inv_h * j.correctionFactor) * j.linearError;
            V2 impulse = -mul(j.m2, Cdot);
            V2 oldImpulse = j.linearImpulse;
            j.linearImpulse += impulse;
            float maxImpulse = h * j.maxForce;
            if (lengthSq(j.linearImpulse) > maxImpulse * maxImpulse) {
                normalize(j.linearImpulse);
                j.linearImpulse *= maxImpulse;
            }
            impulse = j.linearImpulse - oldImpulse;
            vA -= mA * impulse;
            wA -= iA * cross(j.rA, impulse);
            vB += mB * impulse;
            wB += iB * cross(j.rB, impulse);
        }
        break;
    }
    case j_mouse: { /* b2mousejoint.d:217-245 */
        V2 Cdot = vB + cross(wB, j.rB);
        V2 impulse = mul(j.m2, -(Cdot + j.Cm + j.gamma * j.linearImpulse));
        V2 oldImpulse = j.linearImpulse;
        j.linearImpulse += impulse;
        float maxImpulse = step.dt * j.maxForce;
        if (lengthSq(j.linearImpulse) > maxImpulse * maxImpulse)
            j.linearImpulse *= maxImpulse / length(j.linearImpulse);
        impulse = j.linearImpulse - oldImpulse;
        vB += j.invMassB * impulse;
        wB += j.invIB * cross(j.rB, impulse);
        data.velocities[j.indexB].v = vB;
        data.velocities[j.indexB].w = wB;
        return;
    }
    case j_pulley: { /* b2pulleyjoint.d:317-345 */
        V2 vpA = vA + cross(wA, j.rA);
        V2 vpB = vB + cross(wB, j.rB);
        float Cdot = -dot(j.uA, vpA) - j.ratio * dot(j.uB, vpB);
        float impulse = -j.mass * Cdot;
        j.impulse.x += impulse;
        V2 PA = -impulse * j.uA;
        V2 PB = -j.ratio * impulse * j.uB;
        vA += j.invMassA * PA;
        wA += j.invIA * cross(j.rA, PA);
        vB += j.invMassB * PB;
        wB += j.invIB * cross(j.rB, PB);
        break;
    }
    default: break;
    }
    data.velocities[j.indexA].v = vA;
    data.velocities[j.indexA].w = wA;
    data.velocities[j.indexB].v = vB;
    data.velocities[j.indexB].w = wB;
}

static bool jointSolvePosition(Joint &j, SolverData &data) {
    V2 cA = data.positions[j.indexA].c;
    float aA = data.positions[j.indexA].a;
    V2 cB = data.positions[j.indexB].c;
    float aB = data.positions[j.indexB].a;
    const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
    bool ok = true;
    switch (j.type) {
    case j_gear: { /* b2gearjoint.d:388-486 */
        V2 cC = data.positions[j.indexC].c;
        float aC = data.positions[j.indexC].a;
        V2 cD = data.positions[j.indexD].c;
        float aD = data.positions[j.indexD].a;
        Rot qA, qB, qC, qD;
        qA.set(aA);
        qB.set(aB);
        qC.set(aC);
        qD.set(aD);
        float coordinateA = 0.0f, coordinateB = 0.0f;
        V2 JvAC = mk(0.0f, 0.0f), JvBD = mk(0.0f, 0.0f);
        float JwA = 0.0f, JwB = 0.0f, JwC = 0.0f, JwD = 0.0f;
        float mass = 0.0f;
        if (j.typeA == j_revolute) {
            JwA = 1.0f;
            JwC = 1.0f;
            mass += iA + j.iC;
            coordinateA = aA - aC - j.referenceAngleA;
        } else {
            V2 u = mul(qC, j.localAxisC);
            V2 rC = mul(qC, j.localAnchorC - j.lcC);
            V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
            JvAC = u;
            JwC = cross(rC, u);
            JwA = cross(rA, u);
            mass += j.mC + mA + j.iC * JwC * JwC + iA * JwA * JwA;
            V2 pC = j.localAnchorC - j.lcC;
            V2 pA = mulT(qC, rA + (cA - cC));
            coordinateA = dot(pA - pC, j.localAxisC);
        }
        if (j.typeB == j_revolute) {
            JwB = j.ratio;
            JwD = j.ratio;
            mass += j.ratio * j.ratio * (iB + j.iD);
            coordinateB = aB - aD - j.referenceAngleB;
        } else {
            V2 u = mul(qD, j.localAxisD);
            V2 rD = mul(qD, j.localAnchorD - j.lcD);
            V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
            JvBD = j.ratio * u;
            JwD = j.ratio * cross(rD, u);
            JwB = j.ratio * cross(rB, u);
            mass += j.ratio * j.ratio * (j.mD + mB) + j.iD * JwD * JwD + iB * JwB * JwB;
            V2 pD = j.localAnchorD - j.lcD;
            V2 pB = mulT(qD, rB + (cB - cD));
            coordinateB = dot(pB - pD, j.localAxisD);
        }
        float C = (coordinateA + j.ratio * coordinateB) - j.constant;
        float impulse = 0.0f;
        if (mass > 0.0f) impulse = -C / mass;
        cA += mA * impulse * JvAC;
        aA += iA * impulse * JwA;
        cB += mB * impulse * JvBD;
        aB += iB * impulse * JwB;
        cC -= j.mC * impulse * JvAC;
        aC -= j.iC * impulse * JwC;
        cD -= j.mD * impulse * JvBD;
        aD -= j.iD * impulse * JwD;
        data.positions[j.indexA].c = cA;
        data.positions[j.indexA].a = aA;
        data.positions[j.indexB].c = cB;
        data.positions[j.indexB].a = aB;
        data.positions[j.indexC].c = cC;
        data.positions[j.indexC].a = aC;
        data.positions[j.indexD].c = cD;
        data.positions[j.indexD].a = aD;
        return true; /* the linear error is never computed (b2gearjoint.d:484-485) */
    }
    case j_distance: { /* b2distancejoint.d:338-371 */
        if (j.frequencyHz > 0.0f) return true;
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 u = cB + rB - cA - rA;
        float length = normalize(u);
        float C = length - j.length;
        C = clampf(C, -kMaxLinearCorrection, kMaxLinearCorrection);
        float impulse = -j.mass * C;
        V2 P = impulse * u;
        cA -= j.invMassA * P;
        aA -= j.invIA * cross(rA, P);
        cB += j.invMassB * P;
        aB += j.invIB * cross(rB, P);
        ok = fabs_(C) < kLinearSlop;
        break;
    }
    case j_revolute: { /* b2revolutejoint.d:517-590 */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        float angularError = 0.0f, positionError = 0.0f;
        bool fixedRotation = (j.invIA + j.invIB == 0.0f);
        if (j.enableLimit && j.limitState != l_inactive && fixedRotation == false) {
            float angle = aB - aA - j.referenceAngle;
            float limitImpulse = 0.0f;
            if (j.limitState == l_equal) {
                float C = clampf(angle - j.lower, -kMaxAngularCorrection, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
                angularError = fabs_(C);
            } else if (j.limitState == l_atLower) {
                float C = angle - j.lower;
                angularError = -C;
                C = clampf(C + kAngularSlop, -kMaxAngularCorrection, 0.0f);
                limitImpulse = -j.motorMass * C;
            } else if (j.limitState == l_atUpper) {
                float C = angle - j.upper;
                angularError = C;
                C = clampf(C - kAngularSlop, 0.0f, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
            }
            aA -= j.invIA * limitImpulse;
            aB += j.invIB * limitImpulse;
        }
        {
            qA.set(aA);
            qB.set(aB);
            V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
            V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
            V2 C = cB + rB - cA - rA;
            positionError = length(C);
            Mat22 K;
            K.ex.x = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
            K.ex.y = -iA * rA.x * rA.y - iB * rB.x * rB.y;
            K.ey.x = K.ex.y;
            K.ey.y = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
            V2 impulse = -solve22m(K, C);
            cA -= mA * impulse;
            aA -= iA * cross(rA, impulse);
            cB += mB * impulse;
            aB += iB * cross(rB, impulse);
        }
        ok = positionError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case j_prismatic: { /* b2prismaticjoint.d:587-700 */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 d = cB + rB - cA - rA;
        V2 axis = mul(qA, j.localXAxisA);
        float a1 = cross(d + rA, axis);
        float a2 = cross(rB, axis);
        V2 perp = mul(qA, j.localYAxisA);
        float s1 = cross(d + rA, perp);
        float s2 = cross(rB, perp);
        V3 impulse;
        V2 C1;
        C1.x = dot(perp, d);
        C1.y = aB - aA - j.referenceAngle;
        float linearError = fabs_(C1.x);
        float angularError = fabs_(C1.y);
        bool active = false;
        float C2 = 0.0f;
        if (j.enableLimit) {
            float translation = dot(axis, d);
            if (fabs_(j.upper - j.lower) < 2.0f * kLinearSlop) {
                C2 = clampf(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
                linearError = fmax_(linearError, fabs_(translation));
                active = true;
            } else if (translation <= j.lower) {
                C2 = clampf(translation - j.lower + kLinearSlop, -kMaxLinearCorrection, 0.0f);
                linearError = fmax_(linearError, j.lower - translation);
                active = true;
            } else if (translation >= j.upper) {
                C2 = clampf(translation - j.upper - kLinearSlop, 0.0f, kMaxLinearCorrection);
                linearError = fmax_(linearError, translation - j.upper);
                active = true;
            }
        }
        if (active) {
            float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            float k12 = iA * s1 + iB * s2;
            float k13 = iA * s1 * a1 + iB * s2 * a2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            float k23 = iA * a1 + iB * a2;
            float k33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
            Mat33 K;
            K.ex = mk3(k11, k12, k13);
            K.ey = mk3(k12, k22, k23);
            K.ez = mk3(k13, k23, k33);
            impulse = K.solve33(mk3(-C1.x, -C1.y, -C2));
        } else {
            float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            float k12 = iA * s1 + iB * s2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            Mat22 K;
            K.ex = mk(k11, k12);
            K.ey = mk(k12, k22);
            V2 impulse1 = solve22m(K, -C1);
            impulse = mk3(impulse1.x, impulse1.y, 0.0f);
        }
        V2 P = impulse.x * perp + impulse.z * axis;
        float LA = impulse.x * s1 + impulse.y + impulse.z * a1;
        float LB = impulse.x * s2 + impulse.y + impulse.z * a2;
        cA -= mA * P;
        aA -= iA * LA;
        cB += mB * P;
        aB += iB * LB;
        ok = linearError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case j_weld: { /* b2weldjoint.d:347-415 */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        float positionError, angularError;
        Mat33 K;
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
            V2 C1 = cB + rB - cA - rA;
            positionError = length(C1);
            angularError = 0.0f;
            V2 P = -K.solve22(C1);
            cA -= mA * P;
            aA -= iA * cross(rA, P);
            cB += mB * P;
            aB += iB * cross(rB, P);
        } else {
            V2 C1 = cB + rB - cA - rA;
            float C2 = aB - aA - j.referenceAngle;
            positionError = length(C1);
            angularError = fabs_(C2);
            V3 C = mk3(C1.x, C1.y, C2);
            V3 impulse;
            if (K.ez.z > 0.0f) {
                impulse = K.solve33(C);
                impulse = mk3(-impulse.x, -impulse.y, -impulse.z);
            } else {
                V2 impulse2 = -K.solve22(C1);
                impulse = mk3(impulse2.x, impulse2.y, 0.0f);
            }
            V2 P = mk(impulse.x, impulse.y);
            cA -= mA * P;
            aA -= iA * (cross(rA, P) + impulse.z);
            cB += mB * P;
            aB += iB * (cross(rB, P) + impulse.z);
        }
        ok = positionError <= kLinearSlop && angularError <= kAngularSlop;
        break;
    }
    case j_wheel: { /* b2wheeljoint.d:482-530: K uses the members sAy, sBy set at velocity init */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 d = (cB - cA) + rB - rA;
        V2 ay = mul(qA, j.localYAxisA);
        float sAy = cross(d + rA, ay);
        float sBy = cross(rB, ay);
        float C = dot(d, ay);
        float k = j.invMassA + j.invMassB + j.invIA * j.sAy * j.sAy + j.invIB * j.sBy * j.sBy;
        float impulse;
        if (k != 0.0f) impulse = -C / k;
        else impulse = 0.0f;
        V2 P = impulse * ay;
        float LA = impulse * sAy;
        float LB = impulse * sBy;
        cA -= j.invMassA * P;
        aA -= j.invIA * LA;
        cB += j.invMassB * P;
        aB += j.invIB * LB;
        ok = fabs_(C) <= kLinearSlop;
        break;
    }
    case j_rope: { /* b2ropejoint.d:285-318 */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 u = cB + rB - cA - rA;
        float length = normalize(u);
        float C = length - j.length;
        C = clampf(C, 0.0f, kMaxLinearCorrection);
        float impulse = -j.mass * C;
        V2 P = impulse * u;
        cA -= j.invMassA * P;
        aA -= j.invIA * cross(rA, P);
        cB += j.invMassB * P;
        aB += j.invIB * cross(rB, P);
        ok = length - j.length < kLinearSlop;
        break;
    }
    case j_pulley: { /* b2pulleyjoint.d:347-418 */
        Rot qA, qB;
        qA.set(aA);
        qB.set(aB);
        V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 uA = cA + rA - j.groundAnchorA;
        V2 uB = cB + rB - j.groundAnchorB;
        float lengthA = length(uA);
        float lengthB = length(uB);
        if (lengthA > 10.0f * kLinearSlop) uA *= 1.0f / lengthA;
        else uA = mk(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) uB *= 1.0f / lengthB;
        else uB = mk(0.0f, 0.0f);
        float ruA = cross(rA, uA);
        float ruB = cross(rB, uB);
        float pmA = j.invMassA + j.invIA * ruA * ruA;
        float pmB = j.invMassB + j.invIB * ruB * ruB;
        float mass = pmA + j.ratio * j.ratio * pmB;
        if (mass > 0.0f) mass = 1.0f / mass;
        float C = j.constant - lengthA - j.ratio * lengthB;
        float linearError = fabs_(C);
        float impulse = -mass * C;
        V2 PA = -impulse * uA;
        V2 PB = -j.ratio * impulse * uB;
        cA += j.invMassA * PA;
        aA += j.invIA * cross(rA, PA);
        cB += j.invMassB * PB;
        aB += j.invIB * cross(rB, PB);
        ok = linearError < kLinearSlop;
        break;
    }
    default: /* friction, motor, mouse: no position constraint */
        return true;
    }
    data.positions[j.indexA].c = cA;
    data.positions[j.indexA].a = aA;
    data.positions[j.indexB].c = cB;
    data.positions[j.indexB].a = aB;
    return ok;
}
