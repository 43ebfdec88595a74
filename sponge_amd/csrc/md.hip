// NVE leap-frog integrator and SETTLE rigid-body constraints for gfx950.
//
// Reference behaviour: MD_Iteration_Leap_Frog(_With_Max_Velocity)
// (MD_core/MD_core.cu:13-45, driver :1535-1547) and the SETTLE module
// (constrain/SETTLE.cu: canonical triangle geometry :8-47, group discovery
// :189-380, remember :385-410, triangle :412-441 with the solver :100-188,
// pair :442-483, driver :494-525). SETTLE itself is the analytic solution of
// the SHAKE equations of a rigid triangle (Miyamoto & Kollman, J. Comput.
// Chem. 13, 952 (1992)); the code below derives it in its own terms.
//
// Layout: one thread per rigid group. Rigid triangles first, rigid pairs
// next, then (fused iteration only) the atoms that belong to no group. The
// fused iteration does the reference's Remember_Last_Coordinates, Leap_Frog
// and Do_SETTLE in one pass over each group: the three atoms of a water are
// read once and written once, instead of five kernel passes over the arrays.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sponge_md.h"
#include "common.hpp"

using namespace snb;

namespace {

// Canonical shape of a rigid triangle in its own plane, centre of mass at the
// origin: A = (0, ra), B = (rc, rb), C = (rd, re), with ra > 0 and the
// triangle A -> B -> C counter-clockwise (the old-frame normal below is
// (B - A) x (C - A), so the canonical shape must have the same sense).
struct SettleTriangle {
    int a, b, c;
    float ra, rb, rc, rd, re;
};
struct SettlePair {
    int a, b;
    float r0;
    float inv_msum;  // 1 / (m_a + m_b)
};

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 unit(F3 a) { return mul(rsqrtf(dot(a, a)), a); }
__device__ __forceinline__ F3 ld3(const float* p, int i) { return f3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(float* p, int i, F3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// Minimum-image displacement a - b of fixed-point coordinates (wrap-around
// 32-bit difference times L / 2^32), as the pair kernels form it.
__device__ __forceinline__ F3 uint_disp(const unsigned* u, int a, int b, F3 scaler) {
    return f3((float)(int)(u[3 * a] - u[3 * b]) * scaler.x, (float)(int)(u[3 * a + 1] - u[3 * b + 1]) * scaler.y,
              (float)(int)(u[3 * a + 2] - u[3 * b + 2]) * scaler.z);
}
// Minimum-image displacement a - b of float coordinates in a box of sides L.
__device__ __forceinline__ float wrap1(float d, float L) { return d - floorf(d / L + 0.5f) * L; }
__device__ __forceinline__ F3 box_disp(F3 a, F3 b, F3 L) {
    F3 d = sub(a, b);
    return f3(wrap1(d.x, L.x), wrap1(d.y, L.y), wrap1(d.z, L.z));
}

// acc = f / m, v += dt acc (capped to |v| <= vmax when vmax > 0), x += dt v.
__device__ __forceinline__ void leap_frog_one(float* crd, float* vel, const float* frc, float* acc, const float* im,
                                              int i, float dt, float vmax, F3& x, F3& v) {
    const float w = im[i];
    F3 a = mul(w, ld3(frc, i));
    v = add(ld3(vel, i), mul(dt, a));
    if (vmax > 0.f) {
        v = mul(fminf(1.f, vmax * rsqrtf(dot(v, v))), v);
    } else if (acc) {
        st3(acc, i, a);
    }
    x = add(ld3(crd, i), mul(dt, v));
}

struct ConstraintFactors {
    float dt, vfac, xfac;
};

// One rigid triangle. In: b0 = B - A and c0 = C - A before the step, the
// atoms' unconstrained positions xa, xb, xc after it. Out: constrained
// positions, velocities moved by vfac / (dt xfac) times each position
// correction, and the virial sum_k F_k . r_k,old of the constraint forces
// F_k = m_k (x_k - x_k,1) / (dt^2 xfac) (A is the origin of the old vectors).
__device__ __forceinline__ float settle_triangle(const SettleTriangle& t, F3 b0, F3 c0, float ma, float mb, float mc,
                                                 F3 L, ConstraintFactors cf, F3& xa, F3& xb, F3& xc, F3& va, F3& vb,
                                                 F3& vc) {
    // positions relative to the group's centre of mass; the constraint
    // forces are internal, so the centre of mass is where it was
    const F3 db = box_disp(xb, xa, L), dc = box_disp(xc, xa, L);
    const float inv_m = 1.f / (ma + mb + mc);
    const F3 com = add(xa, mul(inv_m, add(mul(mb, db), mul(mc, dc))));
    const F3 a1 = sub(xa, com);
    const F3 b1 = add(a1, db), c1 = add(a1, dc);

    // Frame: n normal to the old triangle; ex normal to n and to the new
    // centre-of-mass -> A vector, ey completing the triad. Constraint
    // corrections run along old bond vectors, i.e. inside the old plane, so
    // every atom keeps its n-component.
    const F3 n = unit(cross(b0, c0));
    const F3 ex = unit(cross(a1, n));
    const F3 ey = unit(cross(n, ex));
    const float xb0 = dot(ex, b0), yb0 = dot(ey, b0);
    const float xc0 = dot(ex, c0), yc0 = dot(ey, c0);
    const float za = dot(n, a1);
    const float xb1 = dot(ex, b1), yb1 = dot(ey, b1), zb1 = dot(n, b1);
    const float xc1 = dot(ex, c1), yc1 = dot(ey, c1), zc1 = dot(n, c1);

    // Tilt the canonical shape out of the old plane: A rises to za (angle
    // phi about ex), then B and C to their heights (angle psi).
    const float sphi = za / t.ra;
    const float cphi = sqrtf(fmaxf(0.f, 1.f - sphi * sphi));
    const float spsi = (zb1 - zc1 - (t.rb - t.re) * sphi) / ((t.rd - t.rc) * cphi);
    const float cpsi = sqrtf(fmaxf(0.f, 1.f - spsi * spsi));
    const float ya2 = t.ra * cphi;
    const float xb2 = t.rc * cpsi, yb2 = t.rb * cphi + t.rc * spsi * sphi;
    const float xc2 = t.rd * cpsi, yc2 = t.re * cphi + t.rd * spsi * sphi;

    // In-plane rotation theta: the corrections exert no torque about the old
    // A, sum_k r_k,old x (x_k - x_k,1) = 0 over B and C, i.e.
    //   p sin(theta) + q cos(theta) = g.
    const float p = xb2 * xb0 + yb2 * yb0 + xc2 * xc0 + yc2 * yc0;
    const float q = xb0 * yb2 - yb0 * xb2 + xc0 * yc2 - yc0 * xc2;
    const float g = xb0 * yb1 - yb0 * xb1 + xc0 * yc1 - yc0 * xc1;
    const float h = p * p + q * q;
    const float sth = (p * g - q * sqrtf(fmaxf(0.f, h - g * g))) / h;
    const float cth = sqrtf(fmaxf(0.f, 1.f - sth * sth));

    // constrained positions in the frame, then back in the lab
    const float xa3 = -ya2 * sth, ya3 = ya2 * cth;
    const float xb3 = xb2 * cth - yb2 * sth, yb3 = xb2 * sth + yb2 * cth;
    const float xc3 = xc2 * cth - yc2 * sth, yc3 = xc2 * sth + yc2 * cth;
    const F3 a3 = add(add(mul(xa3, ex), mul(ya3, ey)), mul(za, n));
    const F3 b3 = add(add(mul(xb3, ex), mul(yb3, ey)), mul(zb1, n));
    const F3 c3 = add(add(mul(xc3, ex), mul(yc3, ey)), mul(zc1, n));

    const F3 da = sub(a3, a1), dbb = sub(b3, b1), dcc = sub(c3, c1);
    const float kv = cf.vfac / cf.dt / cf.xfac;
    va = add(va, mul(kv, da));
    vb = add(vb, mul(kv, dbb));
    vc = add(vc, mul(kv, dcc));
    const float kf = 1.f / cf.dt / cf.dt / cf.xfac;
    const F3 fb = mul(kf * mb, dbb), fc = mul(kf * mc, dcc);
    xa = add(a3, com);
    xb = add(b3, com);
    xc = add(c3, com);
    return fb.x * b0.x + fc.x * c0.x + fb.y * b0.y + fc.y * c0.y + fb.z * b0.z + fc.z * c0.z;
}

// One rigid pair: the correction k r_old with |r_new + k r_old| = r0 (the
// root nearer zero), split in inverse proportion to the masses.
__device__ __forceinline__ float settle_pair(const SettlePair& pr, F3 r2, float ma, float mb, F3 L,
                                             ConstraintFactors cf, F3& xa, F3& xb, F3& va, F3& vb) {
    const F3 r1 = box_disp(xb, xa, L);
    const float s11 = dot(r1, r1), s12 = dot(r1, r2), s22 = dot(r2, r2);
    const float k = (sqrtf(fmaxf(0.f, s12 * s12 - s11 * s22 + s22 * pr.r0 * pr.r0)) - s12) / s22;
    const F3 kr = mul(k * pr.inv_msum, r2);
    const F3 da = mul(-mb, kr), db = mul(ma, kr);
    xa = add(xa, da);
    xb = add(xb, db);
    const float kv = cf.vfac / cf.dt / cf.xfac;
    va = add(va, mul(kv, da));
    vb = add(vb, mul(kv, db));
    const F3 fb = mul(mb / cf.dt / cf.dt / cf.xfac, db);
    return fb.x * r2.x + fb.y * r2.y + fb.z * r2.z;
}

__device__ __forceinline__ void add_pressure(float vir, int need_pressure, float pfac, float* pressure) {
    if (!need_pressure) return;
    vir = wave_sum(vir);
    if ((threadIdx.x & 63) == 0 && vir != 0.f) atomicAdd(pressure, vir * pfac);
}

// MD_INFORMATION::system_information::Get_Pressure (MD_core.cu:1777-1795):
// MD_Atom_Ek (:241-250), Sum_Of_List of the kinetic energies, Add_Sum_List
// of the per-atom virials into *virial, Calculate_Pressure_Cuda (:183-186),
// one block, FP64 sums. With `parts` (the fused step's FP64 energy / virial
// parts) the virial is taken from them instead:
// W = sum(-f_LJ r^2) + 6 E_lrc + E_direct + E_correction + E_self + E_reciprocal.
__global__ __launch_bounds__(1024) void pressure_kernel(int n, const float* __restrict__ vel,
                                                        const float* __restrict__ mass,
                                                        const float* __restrict__ atom_virial, float* virial,
                                                        const double* __restrict__ parts, double inv_v,
                                                        float* __restrict__ pressure, float* __restrict__ ek_out) {
    __shared__ double red[16];
    double ek = 0.0, av = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        ek += 0.5 * (vx * vx + vy * vy + vz * vz) * (double)mass[i];
        if (atom_virial) av += atom_virial[i];
    }
    ek = block_sum<1024>(ek, red);
    av = block_sum<1024>(av, red);
    if (threadIdx.x == 0) {
        double w;
        if (parts) {
            w = parts[SPONGE_PART_LJ_VIRIAL] + 6.0 * parts[SPONGE_PART_LRC] + parts[SPONGE_PART_DIRECT] +
                parts[SPONGE_PART_CORRECTION] + parts[SPONGE_PART_SELF] + parts[SPONGE_PART_RECIPROCAL];
        } else {
            w = (double)virial[0] + av;
            virial[0] = (float)w;
        }
        pressure[0] = (float)((2.0 * ek + w) * (1.0 / 3.0) * inv_v);
        if (ek_out) ek_out[0] = (float)ek;
    }
}

// Scale_List (common.cu): x *= factor.
__global__ void scale_list_kernel(int count, float* __restrict__ x, float factor) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) x[i] *= factor;
}

__global__ void leap_frog_kernel(int n, float* vel, float* crd, const float* frc, float* acc, const float* im,
                                 float dt, float vmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 x, v;
    leap_frog_one(crd, vel, frc, acc, im, i, dt, vmax, x, v);
    st3(vel, i, v);
    st3(crd, i, x);
}

// last: 6 floats per triangle (B - A, C - A), then 3 per pair (j - i)
__global__ void settle_remember_kernel(int ntri, int npair, const SettleTriangle* tri, const SettlePair* prs,
                                       const unsigned* u, F3 scaler, float* last) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ntri) {
        const SettleTriangle t = tri[g];
        const F3 b0 = uint_disp(u, t.b, t.a, scaler), c0 = uint_disp(u, t.c, t.a, scaler);
        st3(last, 2 * g, b0);
        st3(last, 2 * g + 1, c0);
    } else if (g < ntri + npair) {
        const SettlePair p = prs[g - ntri];
        st3(last, ntri + g, uint_disp(u, p.b, p.a, scaler));
    }
}

__global__ void settle_do_kernel(int ntri, int npair, const SettleTriangle* tri, const SettlePair* prs,
                                 const float* last, const float* mass, float* crd, float* vel, F3 L,
                                 ConstraintFactors cf, int need_pressure, float pfac, float* pressure) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    float vir = 0.f;
    if (g < ntri) {
        const SettleTriangle t = tri[g];
        F3 xa = ld3(crd, t.a), xb = ld3(crd, t.b), xc = ld3(crd, t.c);
        F3 va = ld3(vel, t.a), vb = ld3(vel, t.b), vc = ld3(vel, t.c);
        vir = settle_triangle(t, ld3(last, 2 * g), ld3(last, 2 * g + 1), mass[t.a], mass[t.b], mass[t.c], L, cf, xa,
                              xb, xc, va, vb, vc);
        st3(crd, t.a, xa);
        st3(crd, t.b, xb);
        st3(crd, t.c, xc);
        st3(vel, t.a, va);
        st3(vel, t.b, vb);
        st3(vel, t.c, vc);
    } else if (g < ntri + npair) {
        const SettlePair p = prs[g - ntri];
        F3 xa = ld3(crd, p.a), xb = ld3(crd, p.b);
        F3 va = ld3(vel, p.a), vb = ld3(vel, p.b);
        vir = settle_pair(p, ld3(last, ntri + g), mass[p.a], mass[p.b], L, cf, xa, xb, va, vb);
        st3(crd, p.a, xa);
        st3(crd, p.b, xb);
        st3(vel, p.a, va);
        st3(vel, p.b, vb);
    }
    add_pressure(vir, need_pressure, pfac, pressure);
}

// Remember + leap-frog + SETTLE in one pass for group g of [triangles |
// pairs | unconstrained atoms]; returns the group's constraint virial.
__device__ __forceinline__ float md_group(int g, int ntri, int npair, int nfree, const SettleTriangle* tri,
                                          const SettlePair* prs, const int* free_atoms, const unsigned* u, F3 scaler,
                                          const float* im, const float* mass, float* crd, float* vel,
                                          const float* frc, float* acc, F3 L, float dt, float vmax,
                                          ConstraintFactors cf) {
    float vir = 0.f;
    if (g < ntri) {
        const SettleTriangle t = tri[g];
        const F3 b0 = uint_disp(u, t.b, t.a, scaler), c0 = uint_disp(u, t.c, t.a, scaler);
        F3 xa, xb, xc, va, vb, vc;
        leap_frog_one(crd, vel, frc, acc, im, t.a, dt, vmax, xa, va);
        leap_frog_one(crd, vel, frc, acc, im, t.b, dt, vmax, xb, vb);
        leap_frog_one(crd, vel, frc, acc, im, t.c, dt, vmax, xc, vc);
        vir = settle_triangle(t, b0, c0, mass[t.a], mass[t.b], mass[t.c], L, cf, xa, xb, xc, va, vb, vc);
        st3(crd, t.a, xa);
        st3(crd, t.b, xb);
        st3(crd, t.c, xc);
        st3(vel, t.a, va);
        st3(vel, t.b, vb);
        st3(vel, t.c, vc);
    } else if (g < ntri + npair) {
        const SettlePair p = prs[g - ntri];
        const F3 r2 = uint_disp(u, p.b, p.a, scaler);
        F3 xa, xb, va, vb;
        leap_frog_one(crd, vel, frc, acc, im, p.a, dt, vmax, xa, va);
        leap_frog_one(crd, vel, frc, acc, im, p.b, dt, vmax, xb, vb);
        vir = settle_pair(p, r2, mass[p.a], mass[p.b], L, cf, xa, xb, va, vb);
        st3(crd, p.a, xa);
        st3(crd, p.b, xb);
        st3(vel, p.a, va);
        st3(vel, p.b, vb);
    } else if (g < ntri + npair + nfree) {
        const int i = free_atoms[g - ntri - npair];
        F3 x, v;
        leap_frog_one(crd, vel, frc, acc, im, i, dt, vmax, x, v);
        st3(vel, i, v);
        st3(crd, i, x);
    }
    return vir;
}

__global__ void md_iteration_kernel(int ntri, int npair, int nfree, const SettleTriangle* tri, const SettlePair* prs,
                                    const int* free_atoms, const unsigned* u, F3 scaler, const float* im,
                                    const float* mass, float* crd, float* vel, const float* frc, float* acc, F3 L,
                                    float dt, float vmax, ConstraintFactors cf, int need_pressure, float pfac,
                                    float* pressure) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const float vir = md_group(g, ntri, npair, nfree, tri, prs, free_atoms, u, scaler, im, mass, crd, vel, frc, acc,
                               L, dt, vmax, cf);
    add_pressure(vir, need_pressure, pfac, pressure);
}

// One system's integrator record in a batch: its rigid groups and arrays
// (the coordinates, forces, fixed-point coordinates and box come from the
// batch's own system record).
struct MdSys {
    const SettleTriangle* tri;
    const SettlePair* prs;
    const int* free_atoms;
    int ntri, npair, nfree;
    float* vel;
    float* acc;
    const float* im;
    const float* mass;
    float dt, vmax;
    ConstraintFactors cf;
};

// md_iteration_kernel over every system of a batch: block -> {system, block
// within the system} from the map, the system's box from its PME record.
__global__ __launch_bounds__(256) void batch_md_iteration_kernel(const int2* __restrict__ map,
                                                                 const BatchSys* __restrict__ bs,
                                                                 const MdSys* __restrict__ ms) {
    const int2 e = map[blockIdx.x];
    if (e.x < 0) return;
    const MdSys& m = ms[e.x];
    const BatchSys& b = bs[e.x];
    const F3 L{b.pme.box[0], b.pme.box[1], b.pme.box[2]};
    const F3 scaler{L.x / 4294967296.0f, L.y / 4294967296.0f, L.z / 4294967296.0f};
    (void)md_group(e.y * blockDim.x + threadIdx.x, m.ntri, m.npair, m.nfree, m.tri, m.prs, m.free_atoms,
                   b.io.uint_crd, scaler, m.im, m.mass, b.io.crd, m.vel, b.io.frc, m.acc, L, m.dt, m.vmax, m.cf);
}

constexpr int kMdThreads = 256;

// Canonical coordinates of a rigid triangle with sides |AB|, |AC|, |BC| and
// the given masses (FP64): centre of mass at the origin, A on the +y axis,
// A -> B -> C counter-clockwise.
void canonical_triangle(double rab, double rac, double rbc, double ma, double mb, double mc, float out[5]) {
    // A at the origin, B on the -x axis, C above it (law of cosines)
    const double bx = -rab, by = 0.0;
    const double cx = (rbc * rbc - rab * rab - rac * rac) / (2.0 * rab);
    const double cy = std::sqrt(std::max(0.0, rac * rac - cx * cx));
    const double m = ma + mb + mc;
    const double ox = (mb * bx + mc * cx) / m, oy = (mb * by + mc * cy) / m;
    // unit vector from the centre of mass to A, and a perpendicular one
    const double ra = std::sqrt(ox * ox + oy * oy);
    const double uy0 = -ox / ra, uy1 = -oy / ra;
    double ux0 = uy1, ux1 = -uy0;
    auto X = [&](double px, double py) { return (px - ox) * ux0 + (py - oy) * ux1; };
    auto Y = [&](double px, double py) { return (px - ox) * uy0 + (py - oy) * uy1; };
    double rc = X(bx, by), rb = Y(bx, by), rd = X(cx, cy), re = Y(cx, cy);
    // orientation of A -> B -> C in the (x, y) frame; mirror x if clockwise
    const double sense = rc * (re - ra) - (rb - ra) * rd;
    if (sense < 0) {
        rc = -rc;
        rd = -rd;
    }
    out[0] = (float)ra;
    out[1] = (float)rb;
    out[2] = (float)rc;
    out[3] = (float)rd;
    out[4] = (float)re;
}

int launch_blocks(int n) { return ceil_div(n, kMdThreads); }

}  // namespace

struct sponge_settle {
    int n = 0, ntri = 0, npair = 0, nfree = 0, nrest = 0;
    float dt = 0.f, vfac = 1.f, xfac = 1.f;
    std::vector<SettleTriangle> h_tri;
    std::vector<SettlePair> h_pair;
    SettleTriangle* d_tri = nullptr;
    SettlePair* d_pair = nullptr;
    int* d_free = nullptr;
    float* d_last = nullptr;
};

static void settle_free(sponge_settle* s) {
    if (!s) return;
    (void)hipFree(s->d_tri);
    (void)hipFree(s->d_pair);
    (void)hipFree(s->d_free);
    (void)hipFree(s->d_last);
    delete s;
}

extern "C" int sponge_md_leap_frog(int atom_numbers, float* d_vel, float* d_crd, const float* d_frc, float* d_acc,
                                   const float* d_inverse_mass, float dt, float max_velocity, void* stream) {
    SNB_ARG(atom_numbers >= 0 && (atom_numbers == 0 || (d_vel && d_crd && d_frc && d_inverse_mass)), "leap_frog");
    if (atom_numbers == 0) return SPONGE_NB_OK;
    leap_frog_kernel<<<launch_blocks(atom_numbers), kMdThreads, 0, as_stream(stream)>>>(
        atom_numbers, d_vel, d_crd, d_frc, d_acc, d_inverse_mass, dt, max_velocity);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_md_get_pressure(int atom_numbers, const float* d_vel, const float* d_mass,
                                      const float* d_atom_virial, float* d_virial, float volume, float* d_pressure,
                                      float* d_ek, void* stream) {
    SNB_ARG(atom_numbers > 0 && d_vel && d_mass && d_virial && d_pressure && volume > 0.f, "get_pressure");
    pressure_kernel<<<1, 1024, 0, as_stream(stream)>>>(atom_numbers, d_vel, d_mass, d_atom_virial, d_virial, nullptr,
                                                       1.0 / (double)volume, d_pressure, d_ek);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_md_pressure_from_parts(int atom_numbers, const float* d_vel, const float* d_mass,
                                             const double* d_parts, float volume, float* d_pressure, float* d_ek,
                                             void* stream) {
    SNB_ARG(atom_numbers > 0 && d_vel && d_mass && d_parts && d_pressure && volume > 0.f, "pressure_from_parts");
    pressure_kernel<<<1, 1024, 0, as_stream(stream)>>>(atom_numbers, d_vel, d_mass, nullptr, nullptr, d_parts,
                                                       1.0 / (double)volume, d_pressure, d_ek);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_md_scale_list(int count, float* d_list, float factor, void* stream) {
    SNB_ARG(count >= 0 && (count == 0 || d_list), "scale_list");
    if (count == 0) return SPONGE_NB_OK;
    scale_list_kernel<<<ceil_div(count, 256), 256, 0, as_stream(stream)>>>(count, d_list, factor);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

// Group discovery of SETTLE::Initial (constrain/SETTLE.cu:209-331): each
// atom keeps up to two constraint partners; an atom that meets a third is
// marked unusable together with its partner. Three atoms that are each
// other's two partners form a triangle whose A is the lowest-indexed of them
// (B, C = its partners in the order their constraints came); two atoms that
// are each other's only partner form a pair (i = the lower index).
// Constraints whose first atom ends up in no group are left over.
extern "C" int sponge_settle_create(int atom_numbers, int constrain_pair_numbers, const int* atom_i,
                                    const int* atom_j, const float* constant_r, const float* h_mass, float dt,
                                    float exp_gamma, int is_minimization, sponge_settle** out) {
    SNB_ARG(out && atom_numbers > 0 && constrain_pair_numbers >= 0 && h_mass && dt > 0.f, "settle_create");
    SNB_ARG(constrain_pair_numbers == 0 || (atom_i && atom_j && constant_r), "settle_create pairs");
    *out = nullptr;
    const int n = atom_numbers;
    for (int k = 0; k < constrain_pair_numbers; k++)
        SNB_ARG(atom_i[k] >= 0 && atom_i[k] < n && atom_j[k] >= 0 && atom_j[k] < n && atom_i[k] != atom_j[k],
                "settle_create: constraint atom index");
    std::vector<int> deg(n, 0), mate(2 * (size_t)n, -1);
    std::vector<float> len(3 * (size_t)n, 0.f);
    for (int k = 0; k < constrain_pair_numbers; k++) {
        const int i = atom_i[k], j = atom_j[k];
        if (deg[i] < 2 && deg[j] < 2) {
            mate[2 * i + deg[i]] = j;
            mate[2 * j + deg[j]] = i;
            len[3 * i + deg[i]] = constant_r[k];
            len[3 * j + deg[j]] = constant_r[k];
            deg[i]++;
            deg[j]++;
        } else {
            deg[i] = 3;
            deg[j] = 3;
        }
    }
    // classify: deg 2 at a triangle's A, -2 at its B and C; 1 / -1 for a pair
    for (int i = 0; i < n; i++) {
        if (deg[i] == 2) {
            const int b = mate[2 * i], c = mate[2 * i + 1];
            const bool closed = deg[b] == 2 && deg[c] == 2 &&
                                ((mate[2 * b] == i && mate[2 * b + 1] == c) || (mate[2 * b + 1] == i && mate[2 * b] == c));
            if (closed) {
                deg[b] = -2;
                deg[c] = -2;
                len[3 * i + 2] = mate[2 * b + 1] == c ? len[3 * b + 1] : len[3 * b];  // |BC|
            } else {
                deg[i] = deg[b] = deg[c] = 3;
            }
        } else if (deg[i] == 1) {
            const int j = mate[2 * i];
            if (deg[j] == 1) deg[j] = -1;
            else deg[i] = deg[j] = 3;
        }
    }
    auto s = new sponge_settle();
    s->n = n;
    s->dt = dt;
    s->vfac = is_minimization ? 0.f : exp_gamma;
    s->xfac = 0.5f * (1.f + exp_gamma);
    std::vector<char> grouped(n, 0);
    for (int i = 0; i < n; i++) {
        if (deg[i] == 2) {
            deg[i] = -2;
            SettleTriangle t;
            t.a = i;
            t.b = mate[2 * i];
            t.c = mate[2 * i + 1];
            float g[5];
            canonical_triangle(len[3 * i], len[3 * i + 1], len[3 * i + 2], h_mass[t.a], h_mass[t.b], h_mass[t.c], g);
            t.ra = g[0];
            t.rb = g[1];
            t.rc = g[2];
            t.rd = g[3];
            t.re = g[4];
            s->h_tri.push_back(t);
            grouped[t.a] = grouped[t.b] = grouped[t.c] = 1;
        } else if (deg[i] == 1) {
            deg[i] = -1;
            SettlePair p;
            p.a = i;
            p.b = mate[2 * i];
            p.r0 = len[3 * i];
            p.inv_msum = 1.f / (h_mass[p.a] + h_mass[p.b]);
            s->h_pair.push_back(p);
            grouped[p.a] = grouped[p.b] = 1;
        }
    }
    for (int k = 0; k < constrain_pair_numbers; k++)
        if (deg[atom_i[k]] > 0) s->nrest++;
    std::vector<int> fr;
    for (int i = 0; i < n; i++)
        if (!grouped[i]) fr.push_back(i);
    s->ntri = (int)s->h_tri.size();
    s->npair = (int)s->h_pair.size();
    s->nfree = (int)fr.size();
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return SPONGE_NB_OK;
        SNB_HIP(hipMalloc(dst, bytes));
        SNB_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return SPONGE_NB_OK;
    };
    int rc = upload((void**)&s->d_tri, s->h_tri.data(), sizeof(SettleTriangle) * s->ntri);
    if (!rc) rc = upload((void**)&s->d_pair, s->h_pair.data(), sizeof(SettlePair) * s->npair);
    if (!rc) rc = upload((void**)&s->d_free, fr.data(), sizeof(int) * s->nfree);
    if (!rc && s->ntri + s->npair > 0) {
        const size_t bytes = sizeof(float) * (6 * (size_t)s->ntri + 3 * (size_t)s->npair);
        hipError_t e = hipMalloc((void**)&s->d_last, bytes);
        if (e == hipSuccess) e = hipMemset(s->d_last, 0, bytes);
        if (e != hipSuccess) {
            set_error("settle_create: %s", hipGetErrorString(e));
            rc = SPONGE_NB_EHIP;
        }
    }
    if (rc) {
        settle_free(s);
        return rc;
    }
    *out = s;
    return SPONGE_NB_OK;
}

extern "C" void sponge_settle_destroy(sponge_settle* s) { settle_free(s); }

extern "C" int sponge_settle_counts(const sponge_settle* s, int* triangle_numbers, int* pair_numbers,
                                    int* remaining) {
    SNB_ARG(s, "settle_counts");
    if (triangle_numbers) *triangle_numbers = s->ntri;
    if (pair_numbers) *pair_numbers = s->npair;
    if (remaining) *remaining = s->nrest;
    return SPONGE_NB_OK;
}

extern "C" int sponge_settle_groups(const sponge_settle* s, int* triangles, float* geometry, int* pairs,
                                    float* pair_r) {
    SNB_ARG(s, "settle_groups");
    for (int k = 0; k < s->ntri; k++) {
        const SettleTriangle& t = s->h_tri[k];
        if (triangles) {
            triangles[3 * k] = t.a;
            triangles[3 * k + 1] = t.b;
            triangles[3 * k + 2] = t.c;
        }
        if (geometry) {
            const float g[5] = {t.ra, t.rb, t.rc, t.rd, t.re};
            std::memcpy(geometry + 5 * k, g, sizeof(g));
        }
    }
    for (int k = 0; k < s->npair; k++) {
        if (pairs) {
            pairs[2 * k] = s->h_pair[k].a;
            pairs[2 * k + 1] = s->h_pair[k].b;
        }
        if (pair_r) pair_r[k] = s->h_pair[k].r0;
    }
    return SPONGE_NB_OK;
}

extern "C" int sponge_settle_set_factors(sponge_settle* s, float v_factor, float x_factor) {
    SNB_ARG(s && x_factor != 0.f, "settle_set_factors");
    s->vfac = v_factor;
    s->xfac = x_factor;
    return SPONGE_NB_OK;
}

extern "C" int sponge_settle_remember_last_coordinates(sponge_settle* s, const unsigned int* d_uint_crd,
                                                       const float scaler[3], void* stream) {
    SNB_ARG(s && scaler, "settle_remember");
    const int groups = s->ntri + s->npair;
    if (groups == 0) return SPONGE_NB_OK;
    SNB_ARG(d_uint_crd, "settle_remember uint_crd");
    settle_remember_kernel<<<launch_blocks(groups), kMdThreads, 0, as_stream(stream)>>>(
        s->ntri, s->npair, s->d_tri, s->d_pair, d_uint_crd, F3{scaler[0], scaler[1], scaler[2]}, s->d_last);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

static float pressure_factor(const float L[3]) { return 0.33333f / L[0] / L[1] / L[2]; }

extern "C" int sponge_settle_do(sponge_settle* s, const float* d_mass, float* d_crd, const float box_length[3],
                                float* d_vel, int need_pressure, float* d_pressure, void* stream) {
    SNB_ARG(s && box_length, "settle_do");
    const int groups = s->ntri + s->npair;
    if (groups == 0) return SPONGE_NB_OK;
    SNB_ARG(d_mass && d_crd && d_vel && (!need_pressure || d_pressure), "settle_do arrays");
    settle_do_kernel<<<launch_blocks(groups), kMdThreads, 0, as_stream(stream)>>>(
        s->ntri, s->npair, s->d_tri, s->d_pair, s->d_last, d_mass, d_crd, d_vel,
        F3{box_length[0], box_length[1], box_length[2]}, ConstraintFactors{s->dt, s->vfac, s->xfac}, need_pressure,
        pressure_factor(box_length), d_pressure);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_md_iteration(sponge_settle* s, int atom_numbers, float* d_vel, float* d_crd,
                                   const float* d_frc, float* d_acc, const float* d_inverse_mass,
                                   const float* d_mass, const unsigned int* d_uint_crd, const float box_length[3],
                                   float dt, float max_velocity, int need_pressure, float* d_pressure,
                                   void* stream) {
    SNB_ARG(atom_numbers >= 0 && box_length && dt > 0.f, "md_iteration");
    if (!s) return sponge_md_leap_frog(atom_numbers, d_vel, d_crd, d_frc, d_acc, d_inverse_mass, dt, max_velocity,
                                       stream);
    SNB_ARG(s->n == atom_numbers, "md_iteration: atom_numbers differs from the SETTLE handle's");
    SNB_ARG(d_vel && d_crd && d_frc && d_inverse_mass && d_mass && (!need_pressure || d_pressure),
            "md_iteration arrays");
    SNB_ARG(s->ntri + s->npair == 0 || d_uint_crd, "md_iteration uint_crd");
    const int groups = s->ntri + s->npair + s->nfree;
    if (groups == 0) return SPONGE_NB_OK;
    const F3 scaler{box_length[0] / 4294967296.0f, box_length[1] / 4294967296.0f, box_length[2] / 4294967296.0f};
    md_iteration_kernel<<<launch_blocks(groups), kMdThreads, 0, as_stream(stream)>>>(
        s->ntri, s->npair, s->nfree, s->d_tri, s->d_pair, s->d_free, d_uint_crd, scaler, d_inverse_mass, d_mass,
        d_crd, d_vel, d_frc, d_acc, F3{box_length[0], box_length[1], box_length[2]}, dt, max_velocity,
        ConstraintFactors{s->dt, s->vfac, s->xfac}, need_pressure, pressure_factor(box_length), d_pressure);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

// ---------------------------------------------------------------------------
// Batched integrator: the systems' MdSys records follow their SETTLE handles
// and arrays; the launch map gives each system ceil(groups / 256) blocks.

int snb_batch_md_sync(sponge_batch* b) {
    std::vector<MdSys> h(b->nsys);
    for (int s = 0; s < b->nsys; s++) {
        const sponge_settle* st = b->settle[s];
        const sponge_batch_md_io& io = b->md_io[s];
        MdSys& m = h[s];
        if (st) {
            m.tri = st->d_tri;
            m.prs = st->d_pair;
            m.free_atoms = st->d_free;
            m.ntri = st->ntri;
            m.npair = st->npair;
            m.nfree = st->nfree;
            m.cf = ConstraintFactors{st->dt, st->vfac, st->xfac};
        } else {
            m.tri = nullptr;
            m.prs = nullptr;
            m.free_atoms = b->d_ident;
            m.ntri = m.npair = 0;
            m.nfree = b->nl[s]->n;
            m.cf = ConstraintFactors{io.dt, 1.f, 1.f};
        }
        m.vel = io.vel;
        m.acc = io.acc;
        m.im = io.inverse_mass;
        m.mass = io.mass;
        m.dt = io.dt;
        m.vmax = io.max_velocity;
    }
    SNB_HIP(hipMemcpy(b->d_md, h.data(), h.size() * sizeof(MdSys), hipMemcpyHostToDevice));
    return SPONGE_NB_OK;
}

extern "C" int sponge_batch_set_integrator(sponge_batch* b, sponge_settle* const* settle,
                                           const sponge_batch_md_io* io) {
    SNB_ARG(b && io, "batch_set_integrator");
    std::vector<int2> map;
    int plain_n = 0;
    for (int s = 0; s < b->nsys; s++) {
        const sponge_settle* st = settle ? settle[s] : nullptr;
        const int n = b->nl[s]->n;
        SNB_ARG(io[s].vel && io[s].inverse_mass && io[s].dt > 0.f && io[s].max_velocity >= 0.f,
                "batch_set_integrator: system arrays or step");
        SNB_ARG(!st || st->n == n, "batch_set_integrator: a SETTLE handle's atom count differs from its system's");
        SNB_ARG(!st || st->ntri + st->npair == 0 || io[s].mass, "batch_set_integrator: SETTLE needs the masses");
        const int groups = st ? st->ntri + st->npair + st->nfree : n;
        if (!st) plain_n = std::max(plain_n, n);
        for (int j = 0; j < launch_blocks(groups); j++) map.push_back(make_int2(s, j));
    }
    // new buffers first: on an error the batch keeps its previous setting
    int2* d_map = nullptr;
    void* d_md = nullptr;
    int* d_ident = nullptr;
    bool ok = hipMalloc((void**)&d_md, b->nsys * sizeof(MdSys)) == hipSuccess;
    if (ok && !map.empty())
        ok = hipMalloc((void**)&d_map, map.size() * sizeof(int2)) == hipSuccess &&
             hipMemcpy(d_map, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && plain_n > b->ident_n) {
        std::vector<int> iota(plain_n);
        for (int i = 0; i < plain_n; i++) iota[i] = i;
        ok = hipMalloc((void**)&d_ident, plain_n * sizeof(int)) == hipSuccess &&
             hipMemcpy(d_ident, iota.data(), plain_n * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        (void)hipFree(d_map);
        (void)hipFree(d_md);
        (void)hipFree(d_ident);
        set_error("batch_set_integrator: device allocation or upload failed");
        return SPONGE_NB_EHIP;
    }
    (void)hipFree(b->d_md);
    (void)hipFree(b->md_map);
    b->d_md = d_md;
    b->md_map = d_map;
    b->md_blocks = (int)map.size();
    if (d_ident) {
        (void)hipFree(b->d_ident);
        b->d_ident = d_ident;
        b->ident_n = plain_n;
    }
    b->settle.assign(b->nsys, nullptr);
    if (settle) b->settle.assign(settle, settle + b->nsys);
    b->md_io.assign(io, io + b->nsys);
    return snb_batch_md_sync(b);
}

extern "C" int sponge_batch_md_iteration(sponge_batch* b, void* stream) {
    SNB_ARG(b, "batch_md_iteration");
    SNB_ARG(b->d_md, "batch_md_iteration: no integrator (sponge_batch_set_integrator)");
    if (b->md_blocks == 0) return SPONGE_NB_OK;
    batch_md_iteration_kernel<<<b->md_blocks, kMdThreads, 0, as_stream(stream)>>>(b->md_map, b->d,
                                                                                 (const MdSys*)b->d_md);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}
