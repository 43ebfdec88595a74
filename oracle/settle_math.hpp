// SETTLE's closed-form rigid-triangle solve, shared by the FP64 oracle
// (oracle.cpp, T = double) and the FP32 OpenMP CPU baseline (cpu_md.cpp,
// T = float). TEST AND MEASUREMENT INFRASTRUCTURE ONLY: the product
// (sponge_amd/) never includes it.
//
// Get_Rabcde_From_SSS (constrain/SETTLE.cu:8-47) and SETTLE_DO_TRIANGLE
// (constrain/SETTLE.cu:100-188) restated: Miyamoto & Kollman's SETTLE without
// mass weights in the torque condition, as the reference solves it.
#pragma once

#include <cmath>

namespace orc {

template <class T>
inline void v_sub(const T* a, const T* b, T* o) {
    for (int d = 0; d < 3; d++) o[d] = a[d] - b[d];
}
template <class T>
inline T v_dot(const T* a, const T* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class T>
inline void v_cross(const T* a, const T* b, T* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T>
inline void v_unit(T* a) {
    T s = T(1) / std::sqrt(v_dot(a, a));
    for (int d = 0; d < 3; d++) a[d] *= s;
}
// Get_Periodic_Displacement of float coordinates (common.cu:99-107)
template <class T>
inline void v_pbc(const T* a, const T* b, const float* L, T* o) {
    for (int d = 0; d < 3; d++) {
        T x = a[d] - b[d];
        o[d] = x - std::floor(x / T(L[d]) + T(0.5)) * T(L[d]);
    }
}

// A rigid triangle with sides |AB|, |AC|, |BC| placed with its centre of mass
// at the origin, A at (0, ra), B at (rc, rb), C at (rd, re); g = (ra, rb, rc,
// rd, re).
template <class T>
inline void settle_geometry(T rAB, T rAC, T rBC, T mA, T mB, T mC, T* g) {
    T mT = mA + mB + mC;
    T Ax = 0, Ay = 0, Bx = -rAB, By = 0;
    T c = (rBC * rBC - rAC * rAC - rAB * rAB) / (2 * rAC * rAB);
    T Cx = rAC * c, Cy = rAC * std::sqrt(T(1) - c * c);
    T Ox = (Bx * mB + Cx * mC) / mT, Oy = Cy * mC / mT;
    Ax -= Ox; Ay -= Oy; Bx -= Ox; By -= Oy; Cx -= Ox; Cy -= Oy;
    c = T(1) / std::sqrt(T(1) + Ax * Ax / Ay / Ay);
    T s = c * Ax / Ay;
    T ra = Ax * s + Ay * c;
    T rc = Bx * c - By * s, rb = Bx * s + By * c;
    T rd = Cx * c - Cy * s, re = Cx * s + Cy * c;
    if (ra < 0) { ra = -ra; rb = -rb; re = -re; }
    g[0] = ra; g[1] = rb; g[2] = rc; g[3] = rd; g[4] = re;
}

// rB0, rC0: old B-A, C-A; rA1..rC1: new positions relative to the new centre
// of mass, replaced by the constrained ones; v: the three velocities,
// corrected. Returns the triangle's virial F_B . rB0 + F_C . rC0.
template <class T>
inline T settle_do_triangle(const T* rB0, const T* rC0, T* rA1, T* rB1, T* rC1, const T* g, T mB, T mC, T dt,
                            T vfac, T xfac, T* vA, T* vB, T* vC) {
    T ra = g[0], rb = g[1], rc = g[2], rd = g[3], re = g[4];
    T ez[3], ex[3], ey[3];
    v_cross(rB0, rC0, ez);
    v_cross(rA1, ez, ex);
    v_cross(ez, ex, ey);
    v_unit(ex); v_unit(ey); v_unit(ez);
    T B0x = v_dot(ex, rB0), B0y = v_dot(ey, rB0), C0x = v_dot(ex, rC0), C0y = v_dot(ey, rC0);
    T A1z = v_dot(ez, rA1);
    T B1x = v_dot(ex, rB1), B1y = v_dot(ey, rB1), B1z = v_dot(ez, rB1);
    T C1x = v_dot(ex, rC1), C1y = v_dot(ey, rC1), C1z = v_dot(ez, rC1);
    T sinphi = A1z / ra, cosphi = std::sqrt(T(1) - sinphi * sinphi);
    T sinpsi = (B1z - C1z - (rb - re) * sinphi) / ((rd - rc) * cosphi);
    T cospsi = std::sqrt(T(1) - sinpsi * sinpsi);
    T A2y = ra * cosphi;
    T B2x = rc * cospsi, B2y = rb * cosphi + rc * sinpsi * sinphi;
    T C2x = rd * cospsi, C2y = re * cosphi + rd * sinpsi * sinphi;
    T alpha = B2x * B0x + C2x * C0x + B2y * B0y + C2y * C0y;
    T beta = -B2x * B0y - C2x * C0y + B2y * B0x + C2y * C0x;
    T gamma = B1y * B0x - B1x * B0y + C1y * C0x - C1x * C0y;
    T t = alpha * alpha + beta * beta;
    T sinth = (alpha * gamma - beta * std::sqrt(t - gamma * gamma)) / t;
    T costh = std::sqrt(T(1) - sinth * sinth);
    T A3[3] = {-A2y * sinth, A2y * costh, A1z};
    T B3[3] = {B2x * costh - B2y * sinth, B2x * sinth + B2y * costh, B1z};
    T C3[3] = {C2x * costh - C2y * sinth, C2x * sinth + C2y * costh, C1z};
    T a3[3], b3[3], c3[3];
    for (int d = 0; d < 3; d++) {
        a3[d] = A3[0] * ex[d] + A3[1] * ey[d] + A3[2] * ez[d];
        b3[d] = B3[0] * ex[d] + B3[1] * ey[d] + B3[2] * ez[d];
        c3[d] = C3[0] * ex[d] + C3[1] * ey[d] + C3[2] * ez[d];
    }
    T kv = vfac / dt / xfac, kf = T(1) / dt / dt / xfac, vir = 0;
    for (int d = 0; d < 3; d++) {
        vA[d] += kv * (a3[d] - rA1[d]);
        vB[d] += kv * (b3[d] - rB1[d]);
        vC[d] += kv * (c3[d] - rC1[d]);
        vir += kf * mB * (b3[d] - rB1[d]) * rB0[d] + kf * mC * (c3[d] - rC1[d]) * rC0[d];
        rA1[d] = a3[d]; rB1[d] = b3[d]; rC1[d] = c3[d];
    }
    return vir;
}

}  // namespace orc
