// OpenMP CPU baseline of SPONGE's NVE step on the host cores: the same
// formulas as the GPU path in FP32 production arithmetic, for bench.py's
// cpu_baseline leg.
//
// MEASUREMENT AND TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline
// leg and by tests/; the product (sponge_amd/) never links or calls it.
//
// One step is main.cu's loop body for NVE with rigid waters:
//   * neighbor list every `refresh` steps (Neighbor_List_Update interval mode,
//     neighbor_list.cu:510-556): Crd_Periodic_Map (common.cu:299-352), uint
//     coordinates (common.cu:264-287), the reference's cell grid of cells at
//     least (rc + skin) / 2 wide with the 5 x 5 x 5 stencil (neighbor_list.cu:
//     12-25, 57-142), a half list (j > i, FP32 dr2 < (rc + skin)^2,
//     neighbor_list.cu:223-283) minus the exclusions (:353-403);
//   * LJ + PME direct force over the half list (Lennard_Jones_force.cu:
//     424-562), FP32, erfc(x) = exp(-x^2) x erfcx(x) with the rational form
//     the GPU kernels use (refit of Abramowitz & Stegun 7.1.26);
//   * excluded-pair correction (PME_force.cu:620-695), FP32;
//   * SPME reciprocal force (PME_force.cu:367-773): order-4 B-spline charge
//     spread, a 3D real-to-complex FFT written here (Stockham mixed-radix
//     passes, two real z-lines per complex transform), BC convolution,
//     complex-to-real inverse, force gather;
//   * leap-frog (MD_core.cu:13-45) and SETTLE (constrain/SETTLE.cu:100-188,
//     settle_math.hpp) per rigid water.
// Threads: pair and correction forces go to per-thread force arrays (the
// half list writes both atoms), reduced once per step; the spread gives each
// thread whole x planes (every atom filed under planes x..x+3 contributes to
// plane x), so no grid point has two writers; FFT lines, the gather and the
// waters are independent.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "settle_math.hpp"

namespace {

typedef std::complex<float> cf;

inline int num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
inline int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

// erfcx(x) ~ t P7(t), t = 1 / (1 + 0.3195 x): the coefficients the GPU pair
// kernel evaluates (relative error of erfc below 5.2e-7 on [0, 4.5] in FP32).
inline float erfcx_poly(float x) {
    const float t = 1.0f / (1.0f + 0.3195f * x);
    float p = -0.181163783781557f;
    p = p * t + 0.7263733451877996f;
    p = p * t - 0.8054935211563752f;
    p = p * t + 0.8311706379959746f;
    p = p * t - 0.18734797244699758f;
    p = p * t + 0.27142876592060167f;
    p = p * t + 0.16357488112143515f;
    p = p * t + 0.18145758631968023f;
    return t * p;
}

// ---- FFT -------------------------------------------------------------------
// One length: radix passes (4 first, then 2, 3, 5, 7) and per-pass twiddles
// w[k * R + q] = exp(sign 2 pi i q k / (Ns R)) for the Stockham step
//   a_q = x[t + q N / R] w(q, k), k = t mod Ns; DFT_R(a);
//   y[(t / Ns) Ns R + k + q Ns] = a_q;   Ns *= R.
struct Fft1 {
    int n = 0;
    std::vector<int> radix;
    std::vector<std::vector<cf>> tw[2];  // [0] forward (sign -1), [1] inverse

    void plan(int len) {
        n = len;
        radix.clear();
        int m = len;
        for (int r : {4, 2, 3, 5, 7})
            while (m % r == 0) {
                radix.push_back(r);
                m /= r;
            }
        for (int r = 11; m > 1; r++)
            while (m % r == 0) {
                radix.push_back(r);
                m /= r;
            }
        for (int s = 0; s < 2; s++) {
            const double sign = s == 0 ? -1.0 : 1.0;
            tw[s].clear();
            int Ns = 1;
            for (int R : radix) {
                std::vector<cf> w((size_t)Ns * R);
                for (int k = 0; k < Ns; k++)
                    for (int q = 0; q < R; q++) {
                        const double a = sign * 2.0 * M_PI * (double)q * k / ((double)Ns * R);
                        w[(size_t)k * R + q] = cf((float)std::cos(a), (float)std::sin(a));
                    }
                tw[s].push_back(std::move(w));
                Ns *= R;
            }
        }
    }

    // In-register DFT of size R (direct for the primes above 5).
    static void dft(cf* a, int R, int inv) {
        const float sg = inv ? 1.0f : -1.0f;
        if (R == 2) {
            const cf x = a[0], y = a[1];
            a[0] = x + y;
            a[1] = x - y;
        } else if (R == 4) {
            const cf s02 = a[0] + a[2], d02 = a[0] - a[2], s13 = a[1] + a[3];
            const cf d = a[1] - a[3];
            const cf d13(-sg * d.imag(), sg * d.real());  // sign * i * d
            a[0] = s02 + s13;
            a[2] = s02 - s13;
            a[1] = d02 + d13;
            a[3] = d02 - d13;
        } else if (R == 3) {
            const float s = sg * 0.86602540378443865f;
            const cf t = a[1] + a[2], m = a[0] - 0.5f * t, d = a[1] - a[2];
            const cf jd(-s * d.imag(), s * d.real());
            a[0] = a[0] + t;
            a[1] = m + jd;
            a[2] = m - jd;
        } else {
            cf o[16];
            for (int k = 0; k < R; k++) {
                cf acc(0.f, 0.f);
                for (int q = 0; q < R; q++) {
                    const double ang = sg * 2.0 * M_PI * (double)((q * k) % R) / R;
                    acc += a[q] * cf((float)std::cos(ang), (float)std::sin(ang));
                }
                o[k] = acc;
            }
            for (int k = 0; k < R; k++) a[k] = o[k];
        }
    }

    // Unnormalised transform of x (length n) in place; y is scratch of n.
    void run(cf* x, cf* y, int inv) const {
        int Ns = 1;
        cf* src = x;
        cf* dst = y;
        for (size_t p = 0; p < radix.size(); p++) {
            const int R = radix[p];
            const int stride = n / R;
            const cf* w = tw[inv][p].data();
            for (int t = 0; t < stride; t++) {
                const int k = t % Ns;
                cf a[16];
                for (int q = 0; q < R; q++) a[q] = src[t + q * stride] * w[(size_t)k * R + q];
                dft(a, R, inv);
                const int o = (t / Ns) * Ns * R + k;
                for (int q = 0; q < R; q++) dst[o + q * Ns] = a[q];
            }
            std::swap(src, dst);
            Ns *= R;
        }
        if (src != x) std::memcpy(x, src, sizeof(cf) * n);
    }
};

struct CpuMd {
    int n = 0, ntypes = 0, refresh = 20, steps_done = 0;
    float box[3], rc = 10.f, skin = 2.f, beta = 0.f, dt = 0.f;
    std::vector<float> crd, vel, frc, mass, inv_mass, q;
    std::vector<int> type;
    std::vector<float> A, B;       // ntypes^2 square tables (x12 / x6 stored)
    std::vector<int> es, en, el;   // exclusion CSR
    // neighbor list
    int nc[3] = {1, 1, 1};
    std::vector<uint32_t> u;       // uint coordinates
    std::vector<int> cell_start, cell_atoms, nb_n, nb;
    int nb_cap = 1024;
    bool overflow = false;
    // per-thread force arrays
    int nth = 1;
    std::vector<float> fth;
    // PME
    int K[3] = {0, 0, 0}, H = 0;
    std::vector<float> bc;         // [x][y][kz], kz < H
    std::vector<float> grid;       // [x][y][z] real grid, convolved in place
    std::vector<cf> spec;          // [x][y][kz] half spectrum
    std::vector<int> base;         // 3 per atom
    std::vector<float> frac;       // 3 per atom
    std::vector<int> bin_start, bin_atoms;
    Fft1 fx, fy, fz;
    // SETTLE (waters: atoms 3w, 3w + 1, 3w + 2 = O, H, H)
    int nw = 0;
    float geo[5];
    std::vector<float> last;       // old B - A, C - A per water
};

void nl_build(CpuMd& s) {
    const int n = s.n;
    // Crd_Periodic_Map and Crd_To_Uint_Crd with cof = 0.25 * 2^32 / L
    for (int d = 0; d < 3; d++) {
        const float L = s.box[d], cof = 0.25f * (4294967296.0f / L);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            float x = s.crd[3 * i + d];
            if (x >= 0.0f) {
                if (!(x < L)) x -= L;
            } else {
                x += L;
            }
            s.crd[3 * i + d] = x;
            s.u[3 * i + d] = ((uint32_t)(int)(x * cof)) << 2;
        }
    }
    const int ncell = s.nc[0] * s.nc[1] * s.nc[2];
    std::vector<int> cell_of(n), cnt(ncell + 1, 0);
    for (int i = 0; i < n; i++) {
        int c[3];
        for (int d = 0; d < 3; d++) c[d] = (int)(((uint64_t)s.u[3 * i + d] * (uint32_t)s.nc[d]) >> 32);
        cell_of[i] = (c[0] * s.nc[1] + c[1]) * s.nc[2] + c[2];
        cnt[cell_of[i] + 1]++;
    }
    for (int c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    s.cell_start = cnt;
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; i++) s.cell_atoms[pos[cell_of[i]]++] = i;
    // stencil offsets: -2..2 per axis, each distinct cell once on small grids
    std::vector<int> off[3];
    for (int d = 0; d < 3; d++) {
        if (s.nc[d] >= 5) off[d] = {-2, -1, 0, 1, 2};
        else for (int k = 0; k < s.nc[d]; k++) off[d].push_back(k);
    }
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * s.box[d];
    const float rl2 = (s.rc + s.skin) * (s.rc + s.skin);
    bool over = false;
#pragma omp parallel for schedule(dynamic, 64) reduction(|| : over)
    for (int i = 0; i < n; i++) {
        const int ci = cell_of[i];
        const int cz = ci % s.nc[2], cy = (ci / s.nc[2]) % s.nc[1], cx = ci / (s.nc[2] * s.nc[1]);
        int* out = &s.nb[(size_t)i * s.nb_cap];
        int m = 0;
        const uint32_t* ui = &s.u[3 * i];
        for (int ox : off[0]) {
            const int x = s.nc[0] >= 5 ? (cx + ox + s.nc[0]) % s.nc[0] : ox;
            for (int oy : off[1]) {
                const int y = s.nc[1] >= 5 ? (cy + oy + s.nc[1]) % s.nc[1] : oy;
                for (int oz : off[2]) {
                    const int z = s.nc[2] >= 5 ? (cz + oz + s.nc[2]) % s.nc[2] : oz;
                    const int c = (x * s.nc[1] + y) * s.nc[2] + z;
                    for (int k = s.cell_start[c]; k < s.cell_start[c + 1]; k++) {
                        const int j = s.cell_atoms[k];
                        if (j <= i) continue;
                        const uint32_t* uj = &s.u[3 * j];
                        const float dx = (float)(int)(uj[0] - ui[0]) * cof[0];
                        const float dy = (float)(int)(uj[1] - ui[1]) * cof[1];
                        const float dz = (float)(int)(uj[2] - ui[2]) * cof[2];
                        if (dx * dx + dy * dy + dz * dz < rl2) {
                            if (m < s.nb_cap) out[m] = j;
                            m++;
                        }
                    }
                }
            }
        }
        if (m > s.nb_cap) {
            over = true;
            m = s.nb_cap;
        }
        // delete the excluded partners (exclusion rows hold j > i)
        for (int e = s.es[i]; e < s.es[i] + s.en[i]; e++) {
            const int j = s.el[e];
            for (int k = 0; k < m; k++)
                if (out[k] == j) {
                    out[k] = out[--m];
                    break;
                }
        }
        s.nb_n[i] = m;
    }
    s.overflow = s.overflow || over;
}

// LJ + PME direct force and the excluded-pair correction into s.frc.
void direct_force(CpuMd& s) {
    const int n = s.n, T = s.nth;
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * s.box[d];
    const float rc2 = s.rc * s.rc, beta = s.beta, tsp = 1.1283791670955126f;
#pragma omp parallel
    {
        const int t = thread_id();
        float* F = &s.fth[(size_t)t * 3 * n];
        std::fill(F, F + 3 * (size_t)n, 0.f);
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            const uint32_t* ui = &s.u[3 * i];
            const float qi = s.q[i];
            const float* Ai = &s.A[(size_t)s.type[i] * s.ntypes];
            const float* Bi = &s.B[(size_t)s.type[i] * s.ntypes];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            const int* lst = &s.nb[(size_t)i * s.nb_cap];
            for (int k = 0; k < s.nb_n[i]; k++) {
                const int j = lst[k];
                const uint32_t* uj = &s.u[3 * j];
                const float dx = (float)(int)(uj[0] - ui[0]) * cof[0];
                const float dy = (float)(int)(uj[1] - ui[1]) * cof[1];
                const float dz = (float)(int)(uj[2] - ui[2]) * cof[2];
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (!(r2 < rc2)) continue;
                const float ir2 = 1.0f / r2, ir = std::sqrt(ir2), ir6 = ir2 * ir2 * ir2;
                const int tj = s.type[j];
                const float flj = (Bi[tj] - Ai[tj] * ir6) * ir6 * ir2;
                const float x = beta * r2 * ir;
                const float ex = std::exp(-x * x);
                const float fc = qi * s.q[j] * ex * (tsp * x + erfcx_poly(x)) * ir * ir2;
                const float f = flj - fc;
                fx += f * dx;
                fy += f * dy;
                fz += f * dz;
                F[3 * j] -= f * dx;
                F[3 * j + 1] -= f * dy;
                F[3 * j + 2] -= f * dz;
            }
            for (int e = s.es[i]; e < s.es[i] + s.en[i]; e++) {
                const int j = s.el[e];
                const uint32_t* uj = &s.u[3 * j];
                const float dx = (float)(int)(uj[0] - ui[0]) * cof[0];
                const float dy = (float)(int)(uj[1] - ui[1]) * cof[1];
                const float dz = (float)(int)(uj[2] - ui[2]) * cof[2];
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float r = std::sqrt(r2), br = beta * r;
                const float fa = -qi * s.q[j] * (tsp * br * std::exp(-br * br) + std::erfc(br) - 1.0f) / r2 / r;
                fx += fa * dx;
                fy += fa * dy;
                fz += fa * dz;
                F[3 * j] -= fa * dx;
                F[3 * j + 1] -= fa * dy;
                F[3 * j + 2] -= fa * dz;
            }
            F[3 * i] += fx;
            F[3 * i + 1] += fy;
            F[3 * i + 2] += fz;
        }
        // reduce the per-thread arrays (each thread sums its share of atoms)
#pragma omp for schedule(static)
        for (long k = 0; k < 3L * n; k++) {
            float acc = 0.f;
            for (int th = 0; th < T; th++) acc += s.fth[(size_t)th * 3 * n + k];
            s.frc[k] = acc;
        }
    }
}

inline void bspline(float f, float* w, float* dw) {
    const float f2 = f * f, f3 = f2 * f, g = 1.0f - f;
    w[0] = f3 * (1.0f / 6.0f);
    w[1] = (-3.0f * f3 + 3.0f * f2 + 3.0f * f + 1.0f) * (1.0f / 6.0f);
    w[2] = (3.0f * f3 - 6.0f * f2 + 4.0f) * (1.0f / 6.0f);
    w[3] = g * g * g * (1.0f / 6.0f);
    if (dw) {
        dw[0] = 0.5f * f2;
        dw[1] = -1.5f * f2 + f + 0.5f;
        dw[2] = 1.5f * f2 - 2.0f * f;
        dw[3] = -0.5f * g * g;
    }
}

// SPME reciprocal force added to s.frc.
void reciprocal_force(CpuMd& s) {
    const int n = s.n, Kx = s.K[0], Ky = s.K[1], Kz = s.K[2], H = s.H;
    // base and fraction from the uint coordinate, t = u K / 2^32
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            const uint64_t t = (uint64_t)s.u[3 * i + d] * (uint32_t)s.K[d];
            s.base[3 * i + d] = (int)(t >> 32);
            s.frac[3 * i + d] = (float)(uint32_t)t * 2.3283064365386963e-10f;
        }
    // atoms filed under their x base plane
    std::fill(s.bin_start.begin(), s.bin_start.end(), 0);
    for (int i = 0; i < n; i++) s.bin_start[s.base[3 * i] + 1]++;
    for (int x = 0; x < Kx; x++) s.bin_start[x + 1] += s.bin_start[x];
    {
        std::vector<int> pos(s.bin_start.begin(), s.bin_start.end() - 1);
        for (int i = 0; i < n; i++) s.bin_atoms[pos[s.base[3 * i]]++] = i;
    }
    // spread: plane gx takes the atoms filed under gx .. gx + 3
#pragma omp parallel for schedule(dynamic, 1)
    for (int gx = 0; gx < Kx; gx++) {
        float* plane = &s.grid[(size_t)gx * Ky * Kz];
        std::fill(plane, plane + (size_t)Ky * Kz, 0.f);
        for (int a = 0; a < 4; a++) {
            const int bx = (gx + a) % Kx;
            for (int k = s.bin_start[bx]; k < s.bin_start[bx + 1]; k++) {
                const int i = s.bin_atoms[k];
                float wx[4], wy[4], wz[4];
                bspline(s.frac[3 * i], wx, nullptr);
                bspline(s.frac[3 * i + 1], wy, nullptr);
                bspline(s.frac[3 * i + 2], wz, nullptr);
                const float qa = s.q[i] * wx[a];
                for (int b = 0; b < 4; b++) {
                    int gy = s.base[3 * i + 1] - b;
                    gy += gy < 0 ? Ky : 0;
                    float* row = plane + (size_t)gy * Kz;
                    const float qb = qa * wy[b];
                    for (int c = 0; c < 4; c++) {
                        int gz = s.base[3 * i + 2] - c;
                        gz += gz < 0 ? Kz : 0;
                        row[gz] += qb * wz[c];
                    }
                }
            }
        }
    }
    // forward z: two real lines per complex transform, unpacked into the
    // half spectrum
    const int nlines = Kx * Ky;
#pragma omp parallel
    {
        std::vector<cf> buf(std::max({Kx, Ky, Kz})), tmp(buf.size());
#pragma omp for schedule(static)
        for (int l = 0; l < nlines; l += 2) {
            const bool two = l + 1 < nlines;
            const float* a = &s.grid[(size_t)l * Kz];
            for (int z = 0; z < Kz; z++) buf[z] = cf(a[z], two ? a[Kz + z] : 0.f);
            s.fz.run(buf.data(), tmp.data(), 0);
            cf* A = &s.spec[(size_t)l * H];
            for (int k = 0; k < H; k++) {
                const cf zk = buf[k], zc = std::conj(buf[(Kz - k) % Kz]);
                A[k] = 0.5f * (zk + zc);
                if (two) {
                    const cf d = 0.5f * (zk - zc);
                    A[H + k] = cf(d.imag(), -d.real());  // d / i
                }
            }
        }
        // forward y, then x, convolution, inverse x, inverse y
#pragma omp for schedule(static)
        for (int l = 0; l < Kx * H; l++) {
            const int x = l / H, k = l % H;
            cf* p = &s.spec[(size_t)x * Ky * H + k];
            for (int y = 0; y < Ky; y++) buf[y] = p[(size_t)y * H];
            s.fy.run(buf.data(), tmp.data(), 0);
            for (int y = 0; y < Ky; y++) p[(size_t)y * H] = buf[y];
        }
#pragma omp for schedule(static)
        for (int l = 0; l < Ky * H; l++) {
            const int y = l / H, k = l % H;
            cf* p = &s.spec[(size_t)y * H + k];
            const float* g = &s.bc[(size_t)y * H + k];
            const size_t st = (size_t)Ky * H;
            for (int x = 0; x < Kx; x++) buf[x] = p[x * st];
            s.fx.run(buf.data(), tmp.data(), 0);
            for (int x = 0; x < Kx; x++) buf[x] *= g[x * st];
            s.fx.run(buf.data(), tmp.data(), 1);
            for (int x = 0; x < Kx; x++) p[x * st] = buf[x];
        }
#pragma omp for schedule(static)
        for (int l = 0; l < Kx * H; l++) {
            const int x = l / H, k = l % H;
            cf* p = &s.spec[(size_t)x * Ky * H + k];
            for (int y = 0; y < Ky; y++) buf[y] = p[(size_t)y * H];
            s.fy.run(buf.data(), tmp.data(), 1);
            for (int y = 0; y < Ky; y++) p[(size_t)y * H] = buf[y];
        }
        // inverse z (complex to real): two Hermitian half lines A, B as one
        // complex line A + i B, whose inverse gives the two real lines
#pragma omp for schedule(static)
        for (int l = 0; l < nlines; l += 2) {
            const bool two = l + 1 < nlines;
            const cf* A = &s.spec[(size_t)l * H];
            const cf* B = two ? A + H : nullptr;
            for (int k = 0; k < Kz; k++) {
                const bool lo = k < H;
                const cf a = lo ? A[k] : std::conj(A[Kz - k]);
                const cf b = two ? (lo ? B[k] : std::conj(B[Kz - k])) : cf(0.f, 0.f);
                buf[k] = a + cf(-b.imag(), b.real());  // a + i b
            }
            s.fz.run(buf.data(), tmp.data(), 1);
            float* g = &s.grid[(size_t)l * Kz];
            for (int z = 0; z < Kz; z++) {
                g[z] = buf[z].real();
                if (two) g[Kz + z] = buf[z].imag();
            }
        }
    }
    // gather: F = -q sum R dtheta K / L
    const float sx = (float)Kx / s.box[0], sy = (float)Ky / s.box[1], sz = (float)Kz / s.box[2];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
        bspline(s.frac[3 * i], wx, dx);
        bspline(s.frac[3 * i + 1], wy, dy);
        bspline(s.frac[3 * i + 2], wz, dz);
        float F[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < 4; a++) {
            int gx = s.base[3 * i] - a;
            gx += gx < 0 ? Kx : 0;
            for (int b = 0; b < 4; b++) {
                int gy = s.base[3 * i + 1] - b;
                gy += gy < 0 ? Ky : 0;
                const float* row = &s.grid[((size_t)gx * Ky + gy) * Kz];
                for (int c = 0; c < 4; c++) {
                    int gz = s.base[3 * i + 2] - c;
                    gz += gz < 0 ? Kz : 0;
                    const float phi = row[gz];
                    F[0] += phi * dx[a] * wy[b] * wz[c];
                    F[1] += phi * wx[a] * dy[b] * wz[c];
                    F[2] += phi * wx[a] * wy[b] * dz[c];
                }
            }
        }
        s.frc[3 * i] -= s.q[i] * F[0] * sx;
        s.frc[3 * i + 1] -= s.q[i] * F[1] * sy;
        s.frc[3 * i + 2] -= s.q[i] * F[2] * sz;
    }
}

void compute_forces(CpuMd& s) {
    if (s.refresh > 0 ? s.steps_done % s.refresh == 0 : s.steps_done == 0) nl_build(s);
    else {
        // Crd_To_Uint_Crd of the step (MD_core.cu:1528-1532)
        const float cq[3] = {0.25f * (4294967296.0f / s.box[0]), 0.25f * (4294967296.0f / s.box[1]),
                             0.25f * (4294967296.0f / s.box[2])};
#pragma omp parallel for schedule(static)
        for (int i = 0; i < s.n; i++)
            for (int d = 0; d < 3; d++) s.u[3 * i + d] = ((uint32_t)(int)(s.crd[3 * i + d] * cq[d])) << 2;
    }
    direct_force(s);
    reciprocal_force(s);
}

// Remember_Last_Coordinates (SETTLE.cu:385-410), leap-frog (MD_core.cu:
// 13-45), SETTLE per water.
void integrate(CpuMd& s) {
    const float dt = s.dt;
    const float sc[3] = {s.box[0] / 4294967296.0f, s.box[1] / 4294967296.0f, s.box[2] / 4294967296.0f};
#pragma omp parallel for schedule(static)
    for (int w = 0; w < s.nw; w++) {
        const int a = 3 * w;
        for (int k = 1; k <= 2; k++)
            for (int d = 0; d < 3; d++)
                s.last[6 * w + 3 * (k - 1) + d] = (float)(int32_t)(s.u[3 * (a + k) + d] - s.u[3 * a + d]) * sc[d];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; i++)
        for (int d = 0; d < 3; d++) {
            const float v = s.vel[3 * i + d] + dt * s.inv_mass[i] * s.frc[3 * i + d];
            s.vel[3 * i + d] = v;
            s.crd[3 * i + d] += dt * v;
        }
#pragma omp parallel for schedule(static)
    for (int w = 0; w < s.nw; w++) {
        const int a = 3 * w, b = a + 1, c = a + 2;
        float rA[3], rB[3], rC[3], rO[3];
        for (int d = 0; d < 3; d++) rA[d] = s.crd[3 * a + d];
        orc::v_pbc(&s.crd[3 * b], rA, s.box, rB);
        orc::v_pbc(&s.crd[3 * c], rA, s.box, rC);
        const float mA = s.mass[a], mB = s.mass[b], mC = s.mass[c];
        for (int d = 0; d < 3; d++) {
            rO[d] = (mB * rB[d] + mC * rC[d]) / (mA + mB + mC) + rA[d];
            rA[d] -= rO[d];
            rB[d] += rA[d];
            rC[d] += rA[d];
        }
        orc::settle_do_triangle<float>(&s.last[6 * w], &s.last[6 * w + 3], rA, rB, rC, s.geo, mB, mC, dt, 1.0f,
                                       1.0f, &s.vel[3 * a], &s.vel[3 * b], &s.vel[3 * c]);
        for (int d = 0; d < 3; d++) {
            s.crd[3 * a + d] = rA[d] + rO[d];
            s.crd[3 * b + d] = rB[d] + rO[d];
            s.crd[3 * c + d] = rC[d] + rO[d];
        }
    }
}

}  // namespace

extern "C" {

// Handle over host copies of the system. A / B: lower-triangular pair tables
// (x12 / x6); exclusions: the reference's upper-triangular CSR; bc: the BC
// table over the full spectrum [x][y][z] (FP64, as orc_pme_bc_full fills it);
// the first 3 * n_water atoms are rigid waters (O, H, H) with sides r_oh,
// r_oh, r_hh.
void* cpu_md_create(int n, const float* box, const float* crd, const float* vel, const float* mass,
                    const float* charge, const int* type, int ntypes, const float* A_tri, const float* B_tri,
                    const int* es, const int* en, const int* el, float rc, float skin, int refresh, int Kx, int Ky,
                    int Kz, float beta, const double* bc_full, float dt, int n_water, float r_oh, float r_hh) {
    CpuMd* s = new CpuMd();
    s->n = n;
    for (int d = 0; d < 3; d++) s->box[d] = box[d];
    s->crd.assign(crd, crd + 3 * (size_t)n);
    s->vel.assign(vel, vel + 3 * (size_t)n);
    s->frc.assign(3 * (size_t)n, 0.f);
    s->mass.assign(mass, mass + n);
    s->inv_mass.resize(n);
    for (int i = 0; i < n; i++) s->inv_mass[i] = mass[i] > 0 ? 1.0f / mass[i] : 0.f;
    s->q.assign(charge, charge + n);
    s->type.assign(type, type + n);
    s->ntypes = ntypes;
    s->A.resize((size_t)ntypes * ntypes);
    s->B.resize((size_t)ntypes * ntypes);
    for (int x = 0; x < ntypes; x++)
        for (int y = 0; y < ntypes; y++) {
            const int hi = std::max(x, y), lo = std::min(x, y), k = hi * (hi + 1) / 2 + lo;
            s->A[(size_t)x * ntypes + y] = A_tri[k];
            s->B[(size_t)x * ntypes + y] = B_tri[k];
        }
    s->es.assign(es, es + n);
    s->en.assign(en, en + n);
    long ne = 0;
    for (int i = 0; i < n; i++) ne = std::max(ne, (long)es[i] + en[i]);
    s->el.assign(el, el + ne);
    s->rc = rc;
    s->skin = skin;
    s->refresh = refresh;
    for (int d = 0; d < 3; d++) s->nc[d] = std::max(1, (int)std::floor(box[d] / (0.5f * (rc + skin))));
    s->u.assign(3 * (size_t)n, 0u);
    s->cell_atoms.resize(n);
    s->nb_n.assign(n, 0);
    s->nb.resize((size_t)n * s->nb_cap);
    s->nth = num_threads();
    s->fth.resize((size_t)s->nth * 3 * n);
    s->K[0] = Kx;
    s->K[1] = Ky;
    s->K[2] = Kz;
    s->H = Kz / 2 + 1;
    s->beta = beta;
    s->bc.resize((size_t)Kx * Ky * s->H);
    for (int x = 0; x < Kx; x++)
        for (int y = 0; y < Ky; y++)
            for (int z = 0; z < s->H; z++)
                s->bc[((size_t)x * Ky + y) * s->H + z] = (float)bc_full[((size_t)x * Ky + y) * Kz + z];
    s->grid.resize((size_t)Kx * Ky * Kz);
    s->spec.resize((size_t)Kx * Ky * s->H);
    s->base.resize(3 * (size_t)n);
    s->frac.resize(3 * (size_t)n);
    s->bin_start.resize(Kx + 1);
    s->bin_atoms.resize(n);
    s->fx.plan(Kx);
    s->fy.plan(Ky);
    s->fz.plan(Kz);
    s->dt = dt;
    s->nw = n_water;
    const float mO = n_water ? mass[0] : 1.f, mH = n_water ? mass[1] : 1.f;
    orc::settle_geometry<float>(r_oh, r_oh, r_hh, mO, mH, mH, s->geo);
    s->last.resize(6 * (size_t)n_water);
    return s;
}

void cpu_md_destroy(void* h) { delete static_cast<CpuMd*>(h); }

// Forces of the current coordinates (a list rebuild first); frc_out (3n).
void cpu_md_forces(void* h, float* frc_out) {
    CpuMd& s = *static_cast<CpuMd*>(h);
    nl_build(s);
    direct_force(s);
    reciprocal_force(s);
    std::memcpy(frc_out, s.frc.data(), sizeof(float) * 3 * (size_t)s.n);
}

// `steps` NVE steps; returns 1 if a neighbor list overflowed its capacity.
int cpu_md_run(void* h, int steps) {
    CpuMd& s = *static_cast<CpuMd*>(h);
    for (int k = 0; k < steps; k++) {
        compute_forces(s);
        integrate(s);
        s.steps_done++;
    }
    return s.overflow ? 1 : 0;
}

void cpu_md_state(void* h, float* crd, float* vel) {
    CpuMd& s = *static_cast<CpuMd*>(h);
    if (crd) std::memcpy(crd, s.crd.data(), sizeof(float) * 3 * (size_t)s.n);
    if (vel) std::memcpy(vel, s.vel.data(), sizeof(float) * 3 * (size_t)s.n);
}

int cpu_md_threads(void) { return num_threads(); }

}  // extern "C"
