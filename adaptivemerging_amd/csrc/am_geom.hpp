// am_geom.hpp - narrow-phase primitives for gfx950, one pair per lane, FP64.
//
// Each routine reports contacts through a Sink (count-only or write) so the
// same code runs in the counting pass and in the emitting pass of the
// two-pass narrow phase.  Arithmetic follows the reference routines cited at
// each function, in their evaluation order.
#pragma once

#include "am_internal.hpp"

namespace am {

struct BodyPose {
    d3 x;
    m33 R;
};

AMD_INLINE BodyPose pose(const BodiesView& B, int b) { return BodyPose{ld3(&B.x[b * 3]), ldm(&B.R[b * 9])}; }

// What a primitive routine collides: a body's geometry at its pose, or one
// sub-body of a composite (RigidBodyGeomComposite) at the pose
// updateBodyPositionsFromParent gives it.  Contacts are reported on the
// physical body, with the sub-body folded into the bounding-volume id
// (Contact.set, Contact.java:166-177, keeps it as csb for the warm start).
struct Shape {
    const am_body_desc* d;
    BodyPose P;
    int body;  // the contact body
    int sub;   // sub-body inside the composite, -1 for a whole body
};
AMD_INLINE Shape body_shape(const BodiesView& B, int b) { return Shape{&B.desc[b], pose(B, b), b, -1}; }
// RigidTransform3D.mult(A, B) (RigidTransform3D.java:193-202): t = R_A t_B + t_A, R = R_A R_B
AMD_INLINE BodyPose sub_pose(const BodiesView& B, int b, int s) {
    const int g = B.desc[b].tree_first + s;
    const BodyPose Pb = pose(B, b);
    return BodyPose{xform(Pb.R, Pb.x, ld3(&B.sub_b2c[(size_t)g * 12 + 9])), mmul(Pb.R, ldm(&B.sub_b2c[(size_t)g * 12]))};
}
AMD_INLINE Shape sub_shape(const BodiesView& B, int b, int s) {
    return Shape{&B.sub_desc[B.desc[b].tree_first + s], sub_pose(B, b, s), b, s};
}
AMD_INLINE int enc_bv(const Shape& s, int bv) { return s.sub < 0 ? bv : AM_BV_SUB(s.sub, bv); }

// ---------------------------------------------------------------------------
// Box-plane: BoxPlane.dBoxPlane (collision/BoxPlane.java:27-51).  Contacts have
// body1 = plane, normal = plane normal, info = corner index.
template <class Sink>
AMD_INLINE void box_plane(const Shape& box, const Shape& plane, Sink& sink) {
    const am_body_desc& bd = *box.d;
    const am_body_desc& pd = *plane.d;
    d3 n = ld3(pd.plane_n);
    const BodyPose& P = box.P;
    if (dot(P.x, n) + pd.plane_d > bd.radius) return;
    d3 h = mk(0.5 * bd.size[0], 0.5 * bd.size[1], 0.5 * bd.size[2]);
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        d3 q = mk((k & 4) ? -h.x : h.x, (k & 2) ? -h.y : h.y, (k & 1) ? -h.z : h.z);
        d3 qw = xform(P.R, P.x, q);
        double s = dot(qw, n) + pd.plane_d;
        if (s < 0) sink.emit(plane.body, box.body, qw, n, enc_bv(plane, AM_BV_NONE), enc_bv(box, AM_BV_NONE), k, s);
    }
}

// ---------------------------------------------------------------------------
// Box-box: BoxBox.dBoxBox (collision/BoxBox.java:433-859), the ODE separating
// axis test with the reference's tinyOffset (1e-5) tie-break on face axes and
// fudge factor 1.05 on edge axes.
constexpr double kBBFudge = 1.05;
constexpr double kBBTiny = 1e-5;
constexpr double kBBLineEps = 0.0001;

// BoxBox.intersectRectQuad (BoxBox.java:140-209); buffers are caller storage.
AMD_INLINE int intersect_rect_quad(const double h[2], double p[8], double ret[16], double buffer[16]) {
    int nq = 4, nr = 0;
    double* q = p;
    double* r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            double* pq = q;
            double* pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if ((double)sign * pq[dir] < h[dir]) {
                    pr[0] = pq[0];
                    pr[1] = pq[1];
                    pr += 2;
                    nr++;
                    if (nr & 8) {
                        q = r;
                        goto done;
                    }
                }
                double* nextq = (i > 1) ? pq + 2 : q;
                if (((double)sign * pq[dir] < h[dir]) ^ ((double)sign * nextq[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] +
                                  (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) * ((double)sign * h[dir] - pq[dir]);
                    pr[dir] = (double)sign * h[dir];
                    pr += 2;
                    nr++;
                    if (nr & 8) {
                        q = r;
                        goto done;
                    }
                }
                pq += 2;
            }
            q = r;
            r = (q == ret) ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret)
        for (int i = 0; i < nr * 2; i++) ret[i] = q[i];
    return nr;
}

struct BBTest {
    double s;
    int code;
    bool invert;
    int nr_mat;  // 0: none (use nC), 1: column of R1, 2: column of R2
    int nr_col;
    d3 nC;
};

AMD_INLINE bool bb_tst1(double e1, double e2, int mat, int col_, int cc, BBTest& t) {
    double s2 = fabs(e1) - e2;
    if (s2 > 0) return false;
    if (s2 > t.s + kBBTiny) {
        t.s = s2;
        t.nr_mat = mat;
        t.nr_col = col_;
        t.invert = (e1 < 0);
        t.code = cc;
    }
    return true;
}
AMD_INLINE bool bb_tst2(double e1, double e2, double n1, double n2, double n3, int cc, BBTest& t) {
    double s2 = fabs(e1) - e2;
    if (s2 > 0) return false;
    double l = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    if (l > 0) {
        s2 /= l;
        if (s2 * kBBFudge > t.s) {
            t.s = s2;
            t.nr_mat = 0;
            t.nr_col = 0;
            t.nC = mk(n1 / l, n2 / l, n3 / l);
            t.invert = (e1 < 0);
            t.code = cc;
        }
    }
    return true;
}

AMD_INLINE double col_dot(const m33& A, int i, const m33& B, int j) {
    return A.m[i] * B.m[j] + A.m[3 + i] * B.m[3 + j] + A.m[6 + i] * B.m[6 + j];
}
AMD_INLINE double vec_col(d3 v, const m33& M, int j) { return v.x * M.m[j] + v.y * M.m[3 + j] + v.z * M.m[6 + j]; }

template <class Sink>
AMD_INLINE void box_box(const Shape& S1, const Shape& S2, Sink& sink) {
    const am_body_desc& D1 = *S1.d;
    const am_body_desc& D2 = *S2.d;
    const BodyPose& P1 = S1.P;
    const BodyPose& P2 = S2.P;
    const int b1 = S1.body, b2 = S2.body;
    const int bv1 = enc_bv(S1, AM_BV_NONE), bv2 = enc_bv(S2, AM_BV_NONE);
    const m33& R1 = P1.R;
    const m33& R2 = P2.R;
    int info = 0;
    d3 p = sub(P2.x, P1.x);
    if (length(p) > D1.radius + D2.radius) return;
    d3 pp = mul_t(R1, p);
    d3 A = mk(0.5 * D1.size[0], 0.5 * D1.size[1], 0.5 * D1.size[2]);
    d3 Bh = mk(0.5 * D2.size[0], 0.5 * D2.size[1], 0.5 * D2.size[2]);
    double R11 = col_dot(R1, 0, R2, 0), R12 = col_dot(R1, 0, R2, 1), R13 = col_dot(R1, 0, R2, 2);
    double R21 = col_dot(R1, 1, R2, 0), R22 = col_dot(R1, 1, R2, 1), R23 = col_dot(R1, 1, R2, 2);
    double R31 = col_dot(R1, 2, R2, 0), R32 = col_dot(R1, 2, R2, 1), R33 = col_dot(R1, 2, R2, 2);
    double Q11 = fabs(R11), Q12 = fabs(R12), Q13 = fabs(R13);
    double Q21 = fabs(R21), Q22 = fabs(R22), Q23 = fabs(R23);
    double Q31 = fabs(R31), Q32 = fabs(R32), Q33 = fabs(R33);
    BBTest t;
    t.s = -INFINITY;
    t.invert = false;
    t.code = 0;
    t.nr_mat = 0;
    t.nr_col = 0;
    t.nC = mk(0, 0, 0);
    if (!bb_tst1(pp.x, (A.x + Bh.x * Q11 + Bh.y * Q12 + Bh.z * Q13), 1, 0, 1, t)) return;
    if (!bb_tst1(pp.y, (A.y + Bh.x * Q21 + Bh.y * Q22 + Bh.z * Q23), 1, 1, 2, t)) return;
    if (!bb_tst1(pp.z, (A.z + Bh.x * Q31 + Bh.y * Q32 + Bh.z * Q33), 1, 2, 3, t)) return;
    if (!bb_tst1(vec_col(p, R2, 0), (A.x * Q11 + A.y * Q21 + A.z * Q31 + Bh.x), 2, 0, 4, t)) return;
    if (!bb_tst1(vec_col(p, R2, 1), (A.x * Q12 + A.y * Q22 + A.z * Q32 + Bh.y), 2, 1, 5, t)) return;
    if (!bb_tst1(vec_col(p, R2, 2), (A.x * Q13 + A.y * Q23 + A.z * Q33 + Bh.z), 2, 2, 6, t)) return;
    if (!bb_tst2(pp.z * R21 - pp.y * R31, (A.y * Q31 + A.z * Q21 + Bh.y * Q13 + Bh.z * Q12), 0, -R31, R21, 7, t)) return;
    if (!bb_tst2(pp.z * R22 - pp.y * R32, (A.y * Q32 + A.z * Q22 + Bh.x * Q13 + Bh.z * Q11), 0, -R32, R22, 8, t)) return;
    if (!bb_tst2(pp.z * R23 - pp.y * R33, (A.y * Q33 + A.z * Q23 + Bh.x * Q12 + Bh.y * Q11), 0, -R33, R23, 9, t)) return;
    if (!bb_tst2(pp.x * R31 - pp.z * R11, (A.x * Q31 + A.z * Q11 + Bh.y * Q23 + Bh.z * Q22), R31, 0, -R11, 10, t)) return;
    if (!bb_tst2(pp.x * R32 - pp.z * R12, (A.x * Q32 + A.z * Q12 + Bh.x * Q23 + Bh.z * Q21), R32, 0, -R12, 11, t)) return;
    if (!bb_tst2(pp.x * R33 - pp.z * R13, (A.x * Q33 + A.z * Q13 + Bh.x * Q22 + Bh.y * Q21), R33, 0, -R13, 12, t)) return;
    if (!bb_tst2(pp.y * R11 - pp.x * R21, (A.x * Q21 + A.y * Q11 + Bh.y * Q33 + Bh.z * Q32), -R21, R11, 0, 13, t)) return;
    if (!bb_tst2(pp.y * R12 - pp.x * R22, (A.x * Q22 + A.y * Q12 + Bh.x * Q33 + Bh.z * Q31), -R22, R12, 0, 14, t)) return;
    if (!bb_tst2(pp.y * R13 - pp.x * R23, (A.x * Q23 + A.y * Q13 + Bh.x * Q32 + Bh.y * Q31), -R23, R13, 0, 15, t)) return;
    if (t.code == 0) return;

    d3 normal;
    if (t.nr_mat == 1) normal = col(R1, t.nr_col);
    else if (t.nr_mat == 2) normal = col(R2, t.nr_col);
    else normal = mul(R1, t.nC);
    if (t.invert) normal = mk(normal.x * -1.0, normal.y * -1.0, normal.z * -1.0);
    double depth = -t.s;

    if (t.code > 6) {
        // edge-edge (BoxBox.java:600-651)
        d3 pa = P1.x;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double sign = (vec_col(normal, R1, j) > 0) ? 1.0 : -1.0;
            double aj = comp(A, j);
            pa.x += sign * aj * R1.m[0 * 3 + j];
            pa.y += sign * aj * R1.m[1 * 3 + j];
            pa.z += sign * aj * R1.m[2 * 3 + j];
        }
        d3 pb = P2.x;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double sign = (vec_col(normal, R2, j) > 0) ? -1.0 : 1.0;
            double bj = comp(Bh, j);
            pb.x += sign * bj * R2.m[0 * 3 + j];
            pb.y += sign * bj * R2.m[1 * 3 + j];
            pb.z += sign * bj * R2.m[2 * 3 + j];
        }
        d3 ua = col(R1, (t.code - 7) / 3);
        d3 ub = col(R2, (t.code - 7) % 3);
        // dLineClosestApproach (BoxBox.java:366-383)
        d3 pl = sub(pb, pa);
        double uaub = dot(ua, ub);
        double q1 = dot(ua, pl);
        double q2 = -dot(ub, pl);
        double d = 1 - uaub * uaub;
        double alpha, beta;
        if (d <= kBBLineEps) {
            alpha = 0;
            beta = 0;
        } else {
            d = 1.0 / d;
            alpha = (q1 + uaub * q2) * d;
            beta = (uaub * q1 + q2) * d;
        }
        pa = scale_add(alpha, ua, pa);
        pb = scale_add(beta, ub, pb);
        d3 pos = mk((pa.x + pb.x) * 0.5, (pa.y + pb.y) * 0.5, (pa.z + pb.z) * 0.5);
        sink.emit(b1, b2, pos, normal, bv1, bv2, info++, -depth);
        return;
    }

    // face-something (BoxBox.java:656-858)
    const bool ref1 = t.code <= 3;
    const m33& Ra = ref1 ? R1 : R2;
    const m33& Rb = ref1 ? R2 : R1;
    const d3 pa = ref1 ? P1.x : P2.x;
    const d3 pb = ref1 ? P2.x : P1.x;
    const d3 Sa = ref1 ? A : Bh;
    const d3 Sb = ref1 ? Bh : A;
    d3 normal2 = ref1 ? normal : mk(-1.0 * normal.x, -1.0 * normal.y, -1.0 * normal.z);
    d3 nr = mul_t(Rb, normal2);
    d3 anr = mk(fabs(nr.x), fabs(nr.y), fabs(nr.z));
    int lanr, a1, a2;
    if (anr.y > anr.x) {
        if (anr.y > anr.z) {
            a1 = 0; lanr = 1; a2 = 2;
        } else {
            a1 = 0; a2 = 1; lanr = 2;
        }
    } else {
        if (anr.x > anr.z) {
            lanr = 0; a1 = 1; a2 = 2;
        } else {
            a1 = 0; a2 = 1; lanr = 2;
        }
    }
    d3 center;
    double sbl = comp(Sb, lanr);
    if (comp(nr, lanr) < 0) {
        center = mk(pb.x - pa.x + sbl * Rb.m[0 * 3 + lanr], pb.y - pa.y + sbl * Rb.m[1 * 3 + lanr],
                    pb.z - pa.z + sbl * Rb.m[2 * 3 + lanr]);
    } else {
        center = mk(pb.x - pa.x - sbl * Rb.m[0 * 3 + lanr], pb.y - pa.y - sbl * Rb.m[1 * 3 + lanr],
                    pb.z - pa.z - sbl * Rb.m[2 * 3 + lanr]);
    }
    int codeN, code1, code2;
    codeN = ref1 ? t.code - 1 : t.code - 4;
    if (codeN == 0) {
        code1 = 1; code2 = 2;
    } else if (codeN == 1) {
        code1 = 0; code2 = 2;
    } else {
        code1 = 0; code2 = 1;
    }
    double c1 = vec_col(center, Ra, code1);
    double c2 = vec_col(center, Ra, code2);
    double m11 = col_dot(Ra, code1, Rb, a1);
    double m12 = col_dot(Ra, code1, Rb, a2);
    double m21 = col_dot(Ra, code2, Rb, a1);
    double m22 = col_dot(Ra, code2, Rb, a2);
    double quad[8], ret[16], buffer[16], rect[2];
    {
        double sba1 = comp(Sb, a1), sba2 = comp(Sb, a2);
        double k1 = m11 * sba1;
        double k2 = m21 * sba1;
        double k3 = m12 * sba2;
        double k4 = m22 * sba2;
        quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
        quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
        quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
        quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    rect[0] = comp(Sa, code1);
    rect[1] = comp(Sa, code2);
    int n = intersect_rect_quad(rect, quad, ret, buffer);
    if (n < 1) return;
    double det1 = 1.0 / (m11 * m22 - m12 * m21);
    m11 *= det1;
    m12 *= det1;
    m21 *= det1;
    m22 *= det1;
    double san = comp(Sa, codeN);
    d3 ra1 = col(Rb, a1), ra2 = col(Rb, a2);
    for (int j = 0; j < n; j++) {
        double k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        double k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        d3 point = mk(center.x + k1 * ra1.x + k2 * ra2.x, center.y + k1 * ra1.y + k2 * ra2.y,
                      center.z + k1 * ra1.z + k2 * ra2.z);
        double dep = san - (normal2.x * point.x + normal2.y * point.y + normal2.z * point.z);
        if (dep >= 0) {
            d3 pos = mk(point.x + pa.x, point.y + pa.y, point.z + pa.z);
            sink.emit(b1, b2, pos, normal, bv1, bv2, info++, -dep);
        }
    }
}

// ---------------------------------------------------------------------------
// Box-sphere: BoxSphere.dBoxSphere / dBoxSphereTest (collision/BoxSphere.java).
struct BSResult {
    d3 pos, normal;
    double depth;
};
AMD_INLINE void bs_test1(d3 q, d3 cB, double r, BSResult& t) {
    double s = distance(cB, q) - r;
    if (s > 0) return;
    if (s < t.depth) {
        t.pos = q;
        t.normal = normalized(sub(cB, q));
        t.depth = s;
    }
}
AMD_INLINE void bs_test2(double depth, double px, double py, double pz, double nx, double ny, double nz, BSResult& t) {
    if (depth > 0) return;
    if (depth < t.depth) {
        t.pos = mk(px, py, pz);
        t.normal = mk(nx, ny, nz);
        t.depth = depth;
    }
}

template <class Sink>
AMD_INLINE void box_sphere(const Shape& box, const Shape& other, d3 c, double r, int bv_other, Sink& sink) {
    const am_body_desc& bd = *box.d;
    const BodyPose& P = box.P;
    d3 p = mk(0.5 * bd.size[0], 0.5 * bd.size[1], 0.5 * bd.size[2]);
    d3 cB = inv_xform(P.R, P.x, c);
    BSResult t;
    t.depth = 1;
    t.pos = mk(0, 0, 0);
    t.normal = mk(0, 0, 0);
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        d3 q = mk((k & 4) ? -p.x : p.x, (k & 2) ? -p.y : p.y, (k & 1) ? -p.z : p.z);
        bs_test1(q, cB, r, t);
    }
    double s;
    d3 v;
#define AM_BS_EDGE(vx, vy, vz, px_, py_, pz_)            \
    {                                                    \
        v = mk((vx), (vy), (vz));                        \
        s = length(v);                                   \
        s -= r;                                          \
        bs_test2(s, px_, py_, pz_, v.x, v.y, v.z, t);    \
    }
    if (-p.x <= cB.x && cB.x <= p.x) {
        if (cB.y >= p.y && cB.z >= p.z) AM_BS_EDGE(0, cB.y - p.y, cB.z - p.z, cB.x, p.y, p.z)
        else if (cB.y <= -p.y && cB.z >= p.z) AM_BS_EDGE(0, cB.y + p.y, cB.z - p.z, cB.x, -p.y, p.z)
        else if (cB.y >= p.y && cB.z <= -p.z) AM_BS_EDGE(0, cB.y - p.y, cB.z + p.z, cB.x, p.y, -p.z)
        else if (cB.y <= -p.y && cB.z <= -p.z) AM_BS_EDGE(0, cB.y + p.y, cB.z + p.z, cB.x, -p.y, -p.z)
    }
    if (-p.y <= cB.y && cB.y <= p.y) {
        if (cB.x >= p.x && cB.z >= p.z) AM_BS_EDGE(cB.x - p.x, 0, cB.z - p.z, p.x, cB.y, p.z)
        else if (cB.x <= -p.x && cB.z >= p.z) AM_BS_EDGE(cB.x + p.x, 0, cB.z - p.z, -p.x, cB.y, p.z)
        else if (cB.x >= p.x && cB.z <= -p.z) AM_BS_EDGE(cB.x - p.x, 0, cB.z + p.z, p.x, cB.y, -p.z)
        else if (cB.x <= -p.x && cB.z <= -p.z) AM_BS_EDGE(cB.x + p.x, 0, cB.z + p.z, -p.x, cB.y, -p.z)
    }
    if (-p.z <= cB.z && cB.z <= p.z) {
        if (cB.x >= p.x && cB.y >= p.y) AM_BS_EDGE(cB.x - p.x, cB.y - p.y, 0, p.x, p.y, cB.z)
        else if (cB.x <= -p.x && cB.y >= p.y) AM_BS_EDGE(cB.x + p.x, cB.y - p.y, 0, -p.x, p.y, cB.z)
        else if (cB.x >= p.x && cB.y <= -p.y) AM_BS_EDGE(cB.x - p.x, cB.y + p.y, 0, p.x, -p.y, cB.z)
        else if (cB.x <= -p.x && cB.y <= -p.y) AM_BS_EDGE(cB.x + p.x, cB.y + p.y, 0, -p.x, -p.y, cB.z)
    }
#undef AM_BS_EDGE
    if (-p.x <= cB.x && cB.x <= p.x && -p.y <= cB.y && cB.y <= p.y) {
        if (cB.z > 0) {
            s = cB.z - p.z - r;
            bs_test2(s, cB.x, cB.y, p.z, 0, 0, 1, t);
        } else {
            s = -p.z - cB.z - r;
            bs_test2(s, cB.x, cB.y, -p.z, 0, 0, -1, t);
        }
    }
    if (-p.x <= cB.x && cB.x <= p.x && -p.z <= cB.z && cB.z <= p.z) {
        if (cB.y > 0) {
            s = cB.y - p.y - r;
            bs_test2(s, cB.x, p.y, cB.z, 0, 1, 0, t);
        } else {
            s = -p.y - cB.y - r;
            bs_test2(s, cB.x, -p.y, cB.z, 0, -1, 0, t);
        }
    }
    if (-p.y <= cB.y && cB.y <= p.y && -p.z <= cB.z && cB.z <= p.z) {
        if (cB.x > 0) {
            s = cB.x - p.x - r;
            bs_test2(s, p.x, cB.y, cB.z, 1, 0, 0, t);
        } else {
            s = -p.x - cB.x - r;
            bs_test2(s, -p.x, cB.y, cB.z, -1, 0, 0, t);
        }
    }
    if (t.depth != 1) {
        d3 pw = xform(P.R, P.x, t.pos);
        d3 nw = normalized(mul(P.R, t.normal));
        (void)bv_other;
        sink.emit(box.body, other.body, pw, nw, enc_bv(box, AM_BV_NONE), enc_bv(other, AM_BV_NONE), 0, t.depth);
    }
}

AMD_INLINE bool box_sphere_test(const Shape& box, d3 c, double r) {
    const am_body_desc& bd = *box.d;
    const BodyPose& P = box.P;
    if (distance(c, P.x) > bd.radius + r) return false;
    d3 p = mk(0.5 * bd.size[0], 0.5 * bd.size[1], 0.5 * bd.size[2]);
    d3 cB = inv_xform(P.R, P.x, c);
    if (cB.x - p.x > r) return false;
    if (-p.x - cB.x > r) return false;
    if (cB.y - p.y > r) return false;
    if (-p.y - cB.y > r) return false;
    if (cB.z - p.z > r) return false;
    if (-p.z - cB.z > r) return false;
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        d3 q = mk((k & 4) ? -p.x : p.x, (k & 2) ? -p.y : p.y, (k & 1) ? -p.z : p.z);
        if (distance(cB, q) - r < 0) return true;
    }
    d3 v;
#define AM_BS_ET(vx, vy, vz)              \
    {                                     \
        v = mk((vx), (vy), (vz));         \
        if (length(v) - r < 0) return true; \
    }
    if (-p.x <= cB.x && cB.x <= p.x) {
        if (cB.y >= p.y && cB.z >= p.z) AM_BS_ET(0, cB.y - p.y, cB.z - p.z)
        else if (cB.y <= -p.y && cB.z >= p.z) AM_BS_ET(0, cB.y + p.y, cB.z - p.z)
        else if (cB.y >= p.y && cB.z <= -p.z) AM_BS_ET(0, cB.y - p.y, cB.z + p.z)
        else if (cB.y <= -p.y && cB.z <= -p.z) AM_BS_ET(0, cB.y + p.y, cB.z + p.z)
    }
    if (-p.y <= cB.y && cB.y <= p.y) {
        if (cB.x >= p.x && cB.z >= p.z) AM_BS_ET(cB.x - p.x, 0, cB.z - p.z)
        else if (cB.x <= -p.x && cB.z >= p.z) AM_BS_ET(cB.x + p.x, 0, cB.z - p.z)
        else if (cB.x >= p.x && cB.z <= -p.z) AM_BS_ET(cB.x - p.x, 0, cB.z + p.z)
        else if (cB.x <= -p.x && cB.z <= -p.z) AM_BS_ET(cB.x + p.x, 0, cB.z + p.z)
    }
    if (-p.z < cB.z && cB.z < p.z) {
        if (cB.x >= p.x && cB.y >= p.y) AM_BS_ET(cB.x - p.x, cB.y - p.y, 0)
        else if (cB.x <= -p.x && cB.y >= p.y) AM_BS_ET(cB.x + p.x, cB.y - p.y, 0)
        else if (cB.x >= p.x && cB.y <= -p.y) AM_BS_ET(cB.x - p.x, cB.y + p.y, 0)
        else if (cB.x <= -p.x && cB.y <= -p.y) AM_BS_ET(cB.x + p.x, cB.y + p.y, 0)
    }
#undef AM_BS_ET
    if (-p.x <= cB.x && cB.x <= p.x && -p.y <= cB.y && cB.y <= p.y) {
        if (cB.z > 0) {
            if (cB.z - p.z - r < 0) return true;
        } else {
            if (-p.z - cB.z - r < 0) return true;
        }
    }
    if (-p.x <= cB.x && cB.x <= p.x && -p.z <= cB.z && cB.z <= p.z) {
        if (cB.y > 0) {
            if (cB.y - p.y - r < 0) return true;
        } else {
            if (-p.y - cB.y - r < 0) return true;
        }
    }
    if (-p.y <= cB.y && cB.y <= p.y && -p.z <= cB.z && cB.z <= p.z) {
        if (cB.x > 0) {
            if (cB.x - p.x - r < 0) return true;
        } else {
            if (-p.x - cB.x - r < 0) return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------
// Sphere trees: depth-first traversals with an explicit stack that visits
// children in stored order, i.e. the reference recursion order
// (CollisionProcessor.java:793-850, 975-1056).
struct NodeRef {
    d3 c;
    double r;
    int first, count;
};
AMD_INLINE NodeRef node_ref(const TreeView& T, int g) {
    const am_sphere_node& s = T.nodes[g];
    return NodeRef{ld3(&T.cw[g * 3]), s.r, s.first_child, s.n_children};
}

// collideSphereTreeAndPlane: contacts body1 = tree body, body2 = plane,
// bv2 = the plane's dummy volume, violation = d.
template <class Sink>
AMD_INLINE void tree_plane(const TreeView& T, const Shape& body, const Shape& plane, Sink& sink, int* overflow) {
    const am_body_desc& bd = *body.d;
    const am_body_desc& pd = *plane.d;
    d3 n = ld3(pd.plane_n);
    d3 pp = ld3(pd.plane_p);
    int stack[kTreeStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int node = stack[--sp];
        NodeRef s = node_ref(T, bd.tree_first + node);
        double d = n.x * s.c.x + n.y * s.c.y + n.z * s.c.z + pd.plane_d - s.r;
        if (d < 0) {
            if (s.count == 0) {
                d3 nm = sub(s.c, pp);
                double val = dot(nm, n);
                nm = scale(val, n);
                d3 cw = sub(s.c, nm);
                d3 nn = mk(-1.0 * n.x, -1.0 * n.y, -1.0 * n.z);
                sink.emit(body.body, plane.body, cw, nn, enc_bv(body, node), AM_BV_PLANE, 0, d);
            } else {
                if (sp + s.count > kTreeStack) {
                    atomicOr(overflow, 2);
                    return;
                }
                for (int k = s.count - 1; k >= 0; k--) stack[sp++] = s.first + k;
            }
        }
    }
}

// collideBoxAndSphereTree: contacts body1 = box, body2 = tree body.
template <class Sink>
AMD_INLINE void box_tree(const TreeView& T, const Shape& box, const Shape& body2, Sink& sink, int* overflow) {
    const am_body_desc& bd = *body2.d;
    int stack[kTreeStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int node = stack[--sp];
        NodeRef s = node_ref(T, bd.tree_first + node);
        if (!box_sphere_test(box, s.c, s.r)) continue;
        if (s.count != 0) {
            if (sp + s.count > kTreeStack) {
                atomicOr(overflow, 2);
                return;
            }
            for (int k = s.count - 1; k >= 0; k--) stack[sp++] = s.first + k;
        } else {
            box_sphere(box, body2, s.c, s.r, node, sink);
        }
    }
}

// collideSphereTrees + processCollision.
template <class Sink>
AMD_INLINE void tree_tree(const TreeView& T, const Shape& S1, const Shape& S2, Sink& sink, int* overflow) {
    const int f1 = S1.d->tree_first, f2 = S2.d->tree_first;
    // stack entries pack node1 (low 16 bits) and node2 (high 16 bits)
    int stack[kTreeStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int e = stack[--sp];
        int n1 = e & 0xffff, n2 = (e >> 16) & 0xffff;
        NodeRef s1 = node_ref(T, f1 + n1);
        NodeRef s2 = node_ref(T, f2 + n2);
        if (!(distance_sq(s1.c, s2.c) < (s1.r + s2.r) * (s1.r + s2.r))) continue;
        bool l1 = s1.count == 0, l2 = s2.count == 0;
        if (l1 && l2) {
            double dist = distance(s1.c, s2.c);
            double sum = s2.r + s1.r;
            if (dist < sum) {
                double alpha = (s1.r - s2.r + dist) / (2 * dist);
                d3 p = mk((1 - alpha) * s1.c.x + alpha * s2.c.x, (1 - alpha) * s1.c.y + alpha * s2.c.y,
                          (1 - alpha) * s1.c.z + alpha * s2.c.z);
                d3 nrm = normalized(sub(s2.c, s1.c));
                sink.emit(S1.body, S2.body, p, nrm, enc_bv(S1, n1), enc_bv(S2, n2), 0, dist - sum);
            }
            continue;
        }
        bool down2 = l1 || (!l2 && s1.r <= s2.r);
        int cnt = down2 ? s2.count : s1.count;
        if (sp + cnt > kTreeStack) {
            atomicOr(overflow, 2);
            return;
        }
        if (down2) {
            for (int k = cnt - 1; k >= 0; k--) stack[sp++] = n1 | ((s2.first + k) << 16);
        } else {
            for (int k = cnt - 1; k >= 0; k--) stack[sp++] = (s1.first + k) | (n2 << 16);
        }
    }
}

// narrowPhase on two primitive shapes (CollisionProcessor.java:723-750).
template <class Sink>
AMD_INLINE void shape_dispatch(const TreeView& T, const Shape& s1, const Shape& s2, Sink& sink, int* overflow) {
    const int g1 = s1.d->geom, g2 = s2.d->geom;
    if (g1 == AM_GEOM_PLANE) {
        if (g2 == AM_GEOM_PLANE) return;
        if (g2 == AM_GEOM_BOX) box_plane(s2, s1, sink);
        else tree_plane(T, s2, s1, sink, overflow);
    } else if (g1 == AM_GEOM_BOX) {
        if (g2 == AM_GEOM_PLANE) box_plane(s1, s2, sink);
        else if (g2 == AM_GEOM_BOX) box_box(s1, s2, sink);
        else box_tree(T, s1, s2, sink, overflow);
    } else {
        if (g2 == AM_GEOM_PLANE) tree_plane(T, s1, s2, sink, overflow);
        else if (g2 == AM_GEOM_BOX) box_tree(T, s2, s1, sink, overflow);
        else tree_tree(T, s1, s2, sink, overflow);
    }
}

// narrowPhase dispatch on two physical bodies (CollisionProcessor.java:707-757):
// a composite on either side collides each of its sub-bodies in turn, the
// first body's sub-bodies outermost (:711-722).
template <class Sink>
AMD_INLINE void narrow_dispatch(const BodiesView& B, const TreeView& T, int b1, int b2, Sink& sink, int* overflow) {
    const bool c1 = B.desc[b1].geom == AM_GEOM_COMPOSITE, c2 = B.desc[b2].geom == AM_GEOM_COMPOSITE;
    if (!c1 && !c2) {
        shape_dispatch(T, body_shape(B, b1), body_shape(B, b2), sink, overflow);
        return;
    }
    const int n1 = c1 ? B.desc[b1].tree_count : 1, n2 = c2 ? B.desc[b2].tree_count : 1;
    for (int i = 0; i < n1; i++) {
        const Shape s1 = c1 ? sub_shape(B, b1, i) : body_shape(B, b1);
        for (int j = 0; j < n2; j++) shape_dispatch(T, s1, c2 ? sub_shape(B, b2, j) : body_shape(B, b2), sink, overflow);
    }
}

}  // namespace am
