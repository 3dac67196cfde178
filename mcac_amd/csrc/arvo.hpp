// ARVO (a12): the part of a union of balls' volume and area carried by one
// ball (Busa et al., Comput. Phys. Commun. 165 (2005) 59-96; the method of the
// reference's arvo.cpp).  The ball's neighbours cut circles out of its
// surface; their stereographic images (from the north pole onto the plane
// tangent at the south pole) bound the exposed region, whose area and
// z-field volume (divergence theorem with F = (0, 0, z)) are sums of closed-form
// integrals over the boundary arcs.  The parts of all balls sum to the
// union's volume and area.
//
// One thread computes one ball from its neighbours (at most kOverlapSlots),
// in nanometres like the reference so that its absolute tolerances keep their
// scale, with the reference's fixes: positions relative to the aggregate's
// sphere 0, SI results, a ball inside a neighbour carries nothing and a
// neighbour inside the ball, or overlapping it by no more than the contact
// tolerance (1e-10 of the radii sum), is no neighbour.
#pragma once
#include "device_math.hpp"

namespace mcac_dev {

constexpr int kArvoBalls = 16;                    // neighbours of one ball
constexpr double kArvoEpsNorthPole = 0.0001;      // nm (arvo.cpp:38)
constexpr double kArvoNorthPoleReduce = 0.9999;   // arvo.cpp:39
constexpr double kArvoEpsAngle = 1e-12;           // arvo.cpp:41
constexpr double kArvoEpsTwoPi = 1e-12;           // arvo.cpp:42
constexpr int kArvoShrinkCap = 64;                // north-pole shrink steps per neighbour (1 or 2 in practice)

struct ArvoCircle {
    double t, s, r, o;  // image centre, radius, orientation (+1: north pole inside the neighbour)
};

// Is (t, s) exposed with respect to circle m (inside a positive circle or
// outside a negative one)?
__host__ __device__ inline bool arvo_exposed(double t, double s, const ArvoCircle &m) {
    const double d = sqrt((t - m.t) * (t - m.t) + (s - m.s) * (s - m.s));
    return d < m.r ? m.o > 0 : m.o < 0;
}

// Volume and area parts of one arc of circle c from `start` over the oriented
// angle `delta`, for a ball of radius r1 whose centre sits at height z1.
__host__ __device__ inline void arvo_arc(const ArvoCircle &c, double start, double delta, double r1, double z1, double &dv,
                                double &da) {
    const double A = (4.0 * r1 * r1 + c.t * c.t + c.s * c.s + c.r * c.r) / 2.0;
    const double B = c.t * c.r, C = c.s * c.r;
    const double S = sqrt(A * A - B * B - C * C);
    const double rr = c.r * c.r - A, S2 = S * S;
    double I1, I2, I3, J1;
    if (fabs(fabs(delta) - 2.0 * kPi) < kArvoEpsTwoPi) {  // whole circle
        I1 = 2.0 * kPi / S;
        I2 = 2.0 * kPi * A / (S2 * S);
        I3 = kPi * (2.0 * A * A + B * B + C * C) / (S2 * S2 * S);
        J1 = kPi + rr / 2.0 * I1;
    } else {
        const double al = delta < 0 ? start + delta : start;
        const double be = delta < 0 ? start : start + delta;
        const double sa = sin(al), ca = cos(al), sb = sin(be), cb = cos(be);
        const double da_ = A + B * ca + C * sa, db = A + B * cb + C * sb;
        const double fa1 = (-B * sa + C * ca) / da_, fb1 = (-B * sb + C * cb) / db;
        const double fa2 = fa1 / da_, fb2 = fb1 / db;
        const double h = (be - al) / 2.0, m = (al + be) / 2.0;
        I1 = 2.0 * (kPi / 2.0 - atan((A * cos(h) + B * cos(m) + C * sin(m)) / (S * sin(h)))) / S;
        I2 = (fb1 - fa1 + A * I1) / S2;
        I3 = (fb2 - fa2 + (fb1 - fa1) / A + (2.0 * A * A + B * B + C * C) * I2 / A) / (2.0 * S2);
        J1 = ((be - al) + rr * I1) / 2.0;
    }
    const double J2 = (I1 + rr * I2) / 4.0, J3 = (I2 + rr * I3) / 8.0;
    const double r2 = r1 * r1, r3 = r2 * r1, r4 = r2 * r2;
    dv = (128.0 * J3 * (r4 * r3) + 8.0 * J2 * (r3 * r2) + 2.0 * J1 * r3) / 3.0 - 8.0 * r4 * J2 * (z1 + r1);
    da = 2.0 * J1 * r2;
    if (delta < 0) {
        dv = -dv;
        da = -da;
    }
}

// Part of ball (x[0], y[0], z[0], r[0]) with neighbours 1..k (nm): *vol, *area.
__host__ __device__ inline void arvo_ball(double *x, double *y, double *z, double *r, int k, double *vol, double *area) {
    const double r1 = r[0];
    const double VF = 4 * kPi / 3, SF = 4 * kPi;
    if (k == 0) {
        *vol = VF * r1 * r1 * r1;
        *area = SF * r1 * r1;
        return;
    }
    // neighbours whose surface passes at the north pole are shrunk (arvo.cpp:690-707)
    const double nx = x[0], ny = y[0], nz = z[0] + r1;
    for (int j = 1; j <= k; ++j) {
        for (int it = 0; it < kArvoShrinkCap; ++it) {
            const double d = fabs(sqrt((nx - x[j]) * (nx - x[j]) + (ny - y[j]) * (ny - y[j]) + (nz - z[j]) * (nz - z[j])) -
                                  r[j]);
            if (d >= kArvoEpsNorthPole) break;
            r[j] = r[j] * kArvoNorthPoleReduce;
        }
    }
    ArvoCircle c[kArvoBalls];
    bool positive = false;
    for (int j = 0; j < k; ++j) {  // image circles (arvo.cpp:651-678)
        const double dx = x[0] - x[j + 1], dy = y[0] - y[j + 1];
        const double zn = z[0] + r1 - z[j + 1], zs = z[0] - r1 - z[j + 1], rj = r[j + 1];
        const double a = dx * dx + dy * dy + zn * zn - rj * rj;
        const double b = 8.0 * r1 * r1 * dx, cc = 8.0 * r1 * r1 * dy;
        const double d = 4.0 * r1 * r1 * (dx * dx + dy * dy + zs * zs - rj * rj);
        c[j] = {-b / (2.0 * a), -cc / (2.0 * a), sqrt((b * b + cc * cc - 4.0 * a * d) / (4.0 * a * a)), a > 0 ? -1.0 : 1.0};
        positive = positive || c[j].o > 0;
    }
    double pv = 0.0, pa = 0.0, dv, da;
    double ang[2 * kArvoBalls];
    for (int i = 0; i < k; ++i) {
        int na = 0;
        for (int m = 0; m < k; ++m) {  // crossings of circle i with the others, as angles on circle i
            if (m == i) continue;
            const double ex = c[m].t - c[i].t, ey = c[m].s - c[i].s;
            const double d = sqrt(ex * ex + ey * ey);
            if (!(d < c[m].r + c[i].r && fabs(c[m].r - c[i].r) < d)) continue;
            const double along = (c[i].r * c[i].r - c[m].r * c[m].r + d * d) / (2.0 * d);
            const double h = sqrt(smax(c[i].r * c[i].r - along * along, 0.0));
            const double base = atan2(ey, ex), w = atan2(h, along);
            double p = base + w, q = base - w;
            for (int it = 0; it < 4 && p < 0; ++it) p += 2.0 * kPi;
            for (int it = 0; it < 4 && q < 0; ++it) q += 2.0 * kPi;
            for (int it = 0; it < 4 && p >= 2.0 * kPi; ++it) p -= 2.0 * kPi;
            for (int it = 0; it < 4 && q >= 2.0 * kPi; ++it) q -= 2.0 * kPi;
            ang[na++] = p;
            ang[na++] = q;
        }
        if (na == 0) {  // the whole circle if it lies in every other exposed region
            bool all = true;
            for (int m = 0; m < k && all; ++m) {
                if (m == i) continue;
                double pt = c[i].t + c[i].r;
                const double ps = c[i].s;
                if (sqrt((pt - c[m].t) * (pt - c[m].t) + (ps - c[m].s) * (ps - c[m].s)) == c[m].r) pt = c[i].t;
                all = arvo_exposed(pt, ps, c[m]);
            }
            if (all) {
                arvo_arc(c[i], 0.0, 2.0 * kPi * c[i].o, r1, z[0], dv, da);
                pv = pv + dv;
                pa = pa + da;
            }
            continue;
        }
        // ascending for a positive circle, descending for a negative one; merge near-equal angles
        for (int p = 1; p < na; ++p) {
            const double v = ang[p];
            int q = p - 1;
            while (q >= 0 && (c[i].o > 0 ? ang[q] > v : ang[q] < v)) {
                ang[q + 1] = ang[q];
                --q;
            }
            ang[q + 1] = v;
        }
        int nu = 1;
        for (int p = 1; p < na; ++p)
            if (fabs(ang[p] - ang[nu - 1]) > kArvoEpsAngle) ang[nu++] = ang[p];
        for (int p = 0; p < nu; ++p) {  // arc from ang[p] to the next (the last one wraps round)
            const bool wrap = p == nu - 1;
            const double mid = wrap ? (ang[0] + 2.0 * kPi + ang[p]) / 2.0 : (ang[p] + ang[p + 1]) / 2.0;
            const double pt = c[i].t + c[i].r * cos(mid), ps = c[i].s + c[i].r * sin(mid);
            bool ok = true;
            for (int m = 0; m < k && ok; ++m)
                if (m != i) ok = arvo_exposed(pt, ps, c[m]);
            if (!ok) continue;
            const double delta = wrap ? ang[0] + c[i].o * 2.0 * kPi - ang[p] : ang[p + 1] - ang[p];
            arvo_arc(c[i], ang[p], delta, r1, z[0], dv, da);
            pv = pv + dv;
            pa = pa + da;
        }
    }
    *vol = pv;
    *area = pa;
    if (!positive) {  // every circle negative: the arcs bound the covered part
        *vol = *vol + VF * r1 * r1 * r1;
        *area = *area + SF * r1 * r1;
    }
}

}  // namespace mcac_dev
