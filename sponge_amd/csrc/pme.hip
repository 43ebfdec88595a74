// Smooth PME reciprocal part for gfx950 (Particle_Mesh_Ewald,
// PME_force/PME_force.cu), with a hand-written 3D real-to-complex FFT in
// place of cuFFT.
//
// Pipeline per call (real grid [x][y][z], z fastest, as the reference's
// PME_Q; the half spectrum and the BC table are stored [kz][x][y], kz <= Kz/2,
// so that one kz plane is contiguous):
//   pme_prep      per atom: grid base index and spline fraction from the uint
//                 coordinate with one 32x32->64-bit multiply (t = u K / 2^32)
//   spread_z      one block per (x plane, 16 y rows): the stencil rows of every
//                 atom filed under planes x..x+3 that fall in the block are
//                 spread in LDS as exact fixed-point sums (no global atomics
//                 and no charge grid in HBM), then the real z-rows are paired
//                 into complex z-FFTs and unpacked (R2C)
//   conv2d        one block per kz: y- and x-FFT of the plane, energy
//                 1/2 sum m BC |F|^2 (FP64), multiply by BC, inverse x- and
//                 y-FFT
//   inv_z         one block per (x plane, 16 y rows): Hermitian repack and
//                 paired complex z-IFFT (C2R), write the convolved grid
//   pme_gather    16 lanes per atom: force -q sum grid * dtheta (K/L)
// All FFT passes are mixed-radix (8,4,2,3,5,7) Stockham steps over lines held
// in LDS, each line owned by one wave for all of its passes.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "pme_prep.hpp"
#include "tune_table.hpp"

using namespace snb;

namespace {

// y rows per block of the spread / z-transform kernels (even: rows pair into
// complex z-lines)
constexpr int kRowBlock = 16;
constexpr int kZThreads = 256;
constexpr int kSpreadPerThread = 8;
// Fixed-point unit of the spread grid: one contribution q*Mx*My*Mz must stay
// below 2^31 units, i.e. |q| < 128 / (2/3)^3 = 432 (18.2223 x 23.7 e).
constexpr float kSpreadScale = 16777216.0f;
// one block of the convolution kernel holds a whole Kx x Ky plane; four
// waves, one per SIMD, so that a block finds room on CUs the concurrent pair
// kernel already fills
constexpr int kConvThreads = 256;
// Issue priority of the reciprocal pipeline's waves (s_setprio): the pipeline
// is the fused step's critical path and shares the CUs with the pair kernel's
// waves, which have slack
constexpr int kPmeWavePriority = 3;
// global loads in flight per thread in the convolution kernels (a 64 x 64
// plane in two round trips), and spectrum points per thread of the inverse z
// kernel (16 rows x (Kz/2 + 1), Kz <= 512)
constexpr int kConvBatch = 8;
constexpr int kZPer = 17;
constexpr int kGatherThreads = 256;
// The reciprocal chain's kernels run beside the pair kernel, whose waves hold
// 64 of a SIMD's 512 VGPRs each: a SIMD with six pair waves has 128 left, so
// the chain kernels are held to 128 VGPRs (four waves per SIMD) and one of
// their waves still fits there (radix-8 transforms otherwise take ~132, and
// wait for pair waves to retire). Held to 64, they spill.
constexpr int kPmeMinWaves = 4;
// (radix-8 sets only: the others need more registers and would spill)
#define SNB_PME_CHAIN __attribute__((amdgpu_waves_per_eu(RM == kRadix8 ? kPmeMinWaves : 1)))

__host__ __device__ __forceinline__ int row_blocks(int Ky) { return (Ky + kRowBlock - 1) / kRowBlock; }
// LDS layout of transform lines. Stockham passes write each line at strides
// of the radix (positions R j + q), which on a plain layout puts a wave's
// stores on a few banks; one pad slot after every 16 positions, pd(p), and
// line pitches of 1 mod 16 (in 8-byte elements) spread them over all banks.
// Both indices of a 2D plane are padded, so rows and columns stay lines.
__host__ __device__ __forceinline__ int pd(int p) { return p + (p >> 4); }
__host__ __device__ __forceinline__ int padded_len(int K) { return pd(K - 1) + 1; }
__host__ __device__ __forceinline__ int pitch16(int n) { return ((n + 14) / 16) * 16 + 1; }
// pitch of the packed z-lines and of the y direction of the convolution plane
__host__ __device__ __forceinline__ int line_pitch(int K) { return pitch16(padded_len(K)); }
// LDS offset of point idx = x Ky + y of a padded plane (row pitch `pitch`);
// rky = 1 / Ky. 24-bit multiplies (full rate; every index here < 2^24).
__device__ __forceinline__ int plane_lds(int idx, float rky, int Ky, int pitch) {
    const int xx = (int)(((float)idx + 0.5f) * rky);
    const int y = idx - (int)__umul24((unsigned)xx, (unsigned)Ky);
    return (int)__umul24((unsigned)pd(xx), (unsigned)pitch) + pd(y);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by (SIGN * i)
template <int SIGN>
__device__ __forceinline__ float2 mul_i(float2 a) {
    return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// In-register DFT of size R with kernel exp(SIGN 2 pi i q s / R).
template <int R, int SIGN>
__device__ __forceinline__ void dft(float2* v) {
    if (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = mul_i<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if (R == 8) {
        // two radix-4 transforms of the even and odd elements, then the
        // eighth-root twiddles exp(SIGN i pi k / 4)
        const float c = 0.70710678118654752f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft<4, SIGN>(e);
        dft<4, SIGN>(o);
        const float2 o1 = cmul(o[1], make_float2(c, SIGN * c));
        const float2 o2 = mul_i<SIGN>(o[2]);
        const float2 o3 = cmul(o[3], make_float2(-c, SIGN * c));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if (R == 3) {
        const float c = -0.5f, s = 0.86602540378443865f;
        float2 t = cadd(v[1], v[2]);
        float2 m = make_float2(v[0].x + c * t.x, v[0].y + c * t.y);
        float2 d = csub(v[1], v[2]);
        float2 rd = mul_i<SIGN>(make_float2(s * d.x, s * d.y));
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, rd);
        v[2] = csub(m, rd);
    } else {
        // generic R (5, 7): y_s = sum_q v_q exp(SIGN 2 pi i q s / R)
        float2 out[R];
#pragma unroll
        for (int s_ = 0; s_ < R; s_++) {
            float2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; q++) {
                int m = (q * s_) % R;
                float ang = (float)(2.0 * M_PI / R) * m;
                float cs = __builtin_cosf(ang), sn = SIGN * __builtin_sinf(ang);
                acc.x += v[q].x * cs - v[q].y * sn;
                acc.y += v[q].x * sn + v[q].y * cs;
            }
            out[s_] = acc;
        }
#pragma unroll
        for (int s_ = 0; s_ < R; s_++) v[s_] = out[s_];
    }
}

// n / d for 0 <= n < 2^21 from rd = 1.0f / d (d >= 1): (n + 1/2) / d lies at
// least 1/(2d) away from an integer and the float product is within
// 2^-22 (n + 1/2) / d of it, so the truncation is exact. A few VALU
// operations instead of an integer division.
__device__ __forceinline__ int div_small(int n, float rd) { return (int)(((float)n + 0.5f) * rd); }

// One Stockham radix-R step over nlines lines of length K held in LDS;
// element (line l, position p) lives at buf[l*ls + p*ks]; tw is the LDS
// twiddle table exp(-2 pi i j / K); Ns = product of the radices of the
// earlier passes, kons = K / Ns. Line l is owned by wave l % nw in every
// pass (a wave works on G of its lines at a time), and a wave's LDS
// operations execute in program order, so a line's reads all precede its
// writes and the next pass sees them without any workgroup barrier.
template <int R, int SIGN>
__device__ __forceinline__ void wave_pass(float2* buf, int nlines, int ls, int ks, int K, int Ns, int kons,
                                          const float2* tw, int wid, int nw, int lane) {
    const int per = K / R;
    const int tws = kons / R;
    const float rns = 1.0f / (float)Ns;
    if (per <= 64) {
        const float rper = 1.0f / (float)per;
        const int G = div_small(64, rper);
        const int sub = div_small(lane, rper);
        const int j = lane - sub * per;
        const int jn = div_small(j, rns);
        const int k = j - jn * Ns;
        const int d = jn * Ns * R + k;
        // Two groups of G lines per round: every LDS read of both groups
        // (elements and twiddles) is issued before the first butterfly, so
        // one read latency is exposed per two groups. A lane without a line
        // in a group reads line 0 and discards it.
        float2 t[R > 1 ? R - 1 : 1];
#pragma unroll
        for (int q = 1; q < R; q++) {
            t[q - 1] = tw[q * k * tws];
            if (SIGN > 0) t[q - 1].y = -t[q - 1].y;
        }
        // the lane's element offsets within a line, the same for every line
        // of the pass (24-bit multiplies: full rate, every index < 2^24)
        int oin[R], oout[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            oin[q] = (int)__umul24((unsigned)pd(j + (int)__umul24(q, per)), (unsigned)ks);
            oout[q] = (int)__umul24((unsigned)pd(d + (int)__umul24(q, Ns)), (unsigned)ks);
        }
        for (int i0 = 0; wid + nw * i0 < nlines; i0 += 2 * G) {
            const int la = wid + nw * (i0 + sub), lb = la + nw * G;
            const bool oka = sub < G && la < nlines, okb = sub < G && lb < nlines;
            const int ra = (int)__umul24((unsigned)pd(oka ? la : 0), (unsigned)ls);
            const int rb = (int)__umul24((unsigned)pd(okb ? lb : 0), (unsigned)ls);
            float2 va[R], vb[R];
            if (wid + nw * (i0 + G) < nlines) {
                // some lane of the wave has a line in the second group
#pragma unroll
                for (int q = 0; q < R; q++) {
                    va[q] = buf[ra + oin[q]];
                    vb[q] = buf[rb + oin[q]];
                }
#pragma unroll
                for (int q = 1; q < R; q++) {
                    va[q] = cmul(va[q], t[q - 1]);
                    vb[q] = cmul(vb[q], t[q - 1]);
                }
                dft<R, SIGN>(va);
                dft<R, SIGN>(vb);
                if (oka) {
#pragma unroll
                    for (int q = 0; q < R; q++) buf[ra + oout[q]] = va[q];
                }
                if (okb) {
#pragma unroll
                    for (int q = 0; q < R; q++) buf[rb + oout[q]] = vb[q];
                }
            } else {
                // first group only (few lines: the z lines of a row block,
                // or 8 waves on a 64-line plane)
#pragma unroll
                for (int q = 0; q < R; q++) va[q] = buf[ra + oin[q]];
#pragma unroll
                for (int q = 1; q < R; q++) va[q] = cmul(va[q], t[q - 1]);
                dft<R, SIGN>(va);
                if (oka) {
#pragma unroll
                    for (int q = 0; q < R; q++) buf[ra + oout[q]] = va[q];
                }
            }
        }
    } else {
        // long lines (per > 64, at most 128 here): one line per wave at a time
        for (int l = wid; l < nlines; l += nw) {
            float2 v[2][R];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int j = lane + 64 * c;
                if (j < per) {
                    const int k = j - div_small(j, rns) * Ns;
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        float2 x = buf[pd(l) * ls + pd(j + q * per) * ks];
                        if (q > 0) {
                            float2 t = tw[q * k * tws];
                            if (SIGN > 0) t.y = -t.y;
                            x = cmul(x, t);
                        }
                        v[c][q] = x;
                    }
                    dft<R, SIGN>(v[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int j = lane + 64 * c;
                if (j < per) {
                    const int jn = div_small(j, rns);
                    const int d = jn * Ns * R + (j - jn * Ns);
#pragma unroll
                    for (int q = 0; q < R; q++) buf[pd(l) * ls + pd(d + q * Ns) * ks] = v[c][q];
                }
            }
        }
    }
}

// Full multi-pass transform of the wave's lines; callers put a workgroup
// barrier before (data written by other waves) and after (data read by other
// waves).
// Radix sets of the transform kernels: bit r set = radix-r passes compiled
// in. Kernels are instantiated for a few sets (kRadix8: 64- and 512-point
// axes; kRadix34: e.g. 108 = 4 x 3 x 3 x 3; kRadix248: other powers of two)
// besides all six radices: without the code of the radices an axis does not
// use, the kernels need ~120 instead of ~210 VGPRs, so twice the waves fit a
// SIMD. The host picks the smallest instantiated set that holds every pass.
constexpr int kRadixAll = (1 << 2) | (1 << 3) | (1 << 4) | (1 << 5) | (1 << 7) | (1 << 8);
constexpr int kRadix8 = 1 << 8;
constexpr int kRadix34 = (1 << 3) | (1 << 4);
constexpr int kRadix248 = (1 << 2) | (1 << 4) | (1 << 8);

template <int SIGN, int NT, int RM = kRadixAll>
__device__ __forceinline__ void lds_fft(float2* buf, int nlines, int ls, int ks, const FftPlan& p, const float2* tw) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = NT / 64;
    int Ns = 1, kons = p.K;
    for (int ps = 0; ps < p.npass; ps++) {
        const int R = (int)((p.code >> (4 * ps)) & 15ull);
        if ((RM & (1 << 8)) && R == 8) {
            wave_pass<8, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        } else if ((RM & (1 << 4)) && R == 4) {
            wave_pass<4, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        } else if ((RM & (1 << 2)) && R == 2) {
            wave_pass<2, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        } else if ((RM & (1 << 3)) && R == 3) {
            wave_pass<3, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        } else if ((RM & (1 << 5)) && R == 5) {
            wave_pass<5, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        } else if (RM & (1 << 7)) {
            wave_pass<7, SIGN>(buf, nlines, ls, ks, p.K, Ns, kons, tw, wid, nw, lane);
        }
        kons /= R;
        Ns *= R;
    }
}

template <int NT>
__device__ __forceinline__ void load_twiddles(float2* dst, const float2* __restrict__ src, int K) {
    for (int j = threadIdx.x; j < K; j += NT) dst[j] = src[j];
}

// ---------------------------------------------------------------------------
// PME_Atom_Near (PME_force.cu:367-421): per-atom base, fraction and x-plane
// bin (pme_prep_block).
__global__ __launch_bounds__(256) void pme_prep_kernel(int n, const unsigned int* __restrict__ uc,
                                                       const float* __restrict__ q, PmePrep P, int energy) {
    // pme_prep_block holds workgroup barriers: every thread reaches the one
    // call site, atoms past n pass i = -1.
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < n;
    pme_prep_block(ok ? i : -1, ok ? uc[3 * i] : 0u, ok ? uc[3 * i + 1] : 0u, ok ? uc[3 * i + 2] : 0u,
                   ok ? q[i] : 0.f, P, energy);
}

// Order-4 cardinal B-spline weights M4(f + k), k = 0..3, and derivatives
// (coefficient tables PME_force.cu:8-14).
__device__ __forceinline__ float bsp_w(float f, int k) {
    float f2 = f * f, f3 = f2 * f;
    switch (k) {
        case 0: return f3 * (1.0f / 6.0f);
        case 1: return (-3.0f * f3 + 3.0f * f2 + 3.0f * f + 1.0f) * (1.0f / 6.0f);
        case 2: return (3.0f * f3 - 6.0f * f2 + 4.0f) * (1.0f / 6.0f);
        default: {
            float g = 1.0f - f;
            return g * g * g * (1.0f / 6.0f);
        }
    }
}
__device__ __forceinline__ float bsp_dw(float f, int k) {
    float f2 = f * f;
    switch (k) {
        case 0: return 0.5f * f2;
        case 1: return -1.5f * f2 + f + 0.5f;
        case 2: return 1.5f * f2 - 2.0f * f;
        default: {
            float g = 1.0f - f;
            return -0.5f * g * g;
        }
    }
}

// PME_Q_Spread (PME_force.cu:428-471) fused with the forward z transform of
// one block of kRowBlock y rows of one x plane: grid point (base - k) gets
// q M4(fx+kx) M4(fy+ky) M4(fz+kz); plane x receives the atoms filed under
// planes x .. x+3 (kx = their offset), of which the block keeps the stencil
// rows that fall inside it. The rows are summed as fixed-point integers in
// LDS (LDS integer atomics are an order of magnitude faster than
// LDS float atomics on this chip, and integer sums do not depend on the order
// in which atoms arrive), converted, paired into complex z-lines (row r is
// the real (r even) or imaginary (r odd) part of line r/2), transformed, and
// unpacked (R2C) straight into the spectrum, stored [kz][x][y].
//
// MODE 1 spreads, instead of the charge, the adjoint source of the
// force-matching gradient, q_i w_i . grad M_i(g): wv[i] = (w_x Kx/Lx,
// w_y Ky/Ly, w_z Kz/Lz), with the fixed-point unit read from scale_ptr (set
// on the device so that no contribution reaches 2^31).
template <int MODE = 0, int RM = kRadixAll>
__device__ __forceinline__ void spread_z_body(
    int x, int yb, const unsigned* __restrict__ rec_base, const float4* __restrict__ rec_frac,
    const int* __restrict__ bin_count, const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx,
    int Ky, int Kz, const FftPlan& pz, const float2* __restrict__ twz, double* __restrict__ acc,
    const float4* __restrict__ wv = nullptr, const float* __restrict__ scale_ptr = nullptr) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    // the four bins' counts are read first: their latency overlaps the
    // twiddle loads and the clearing of the accumulators
    int b0 = x, b1 = x + 1, b2 = x + 2, b3 = x + 3;
    b1 -= b1 >= Kx ? Kx : 0;
    b2 -= b2 >= Kx ? Kx : 0;
    b3 -= b3 >= Kx ? Kx : 0;
    const int c0 = bin_count[b0], c1 = bin_count[b1], c2 = bin_count[b2], c3 = bin_count[b3];
    const float unit = MODE ? scale_ptr[0] : kSpreadScale;
    const int H = Kz / 2 + 1, zp = line_pitch(Kz);
    const int y0 = yb * kRowBlock, nr = min(kRowBlock, Ky - y0), nrp = (nr + 1) / 2;
    // Charge spread: each 32-bit contribution c is summed as two 32-bit
    // words, hi = c >> 16 (signed) and lo = c & 0xffff, in two arrays; with
    // fewer than 2^16 contributions per point neither word overflows and
    // hi 2^16 + lo is the exact 64-bit sum. A wave's 32-bit atomics touch
    // one bank per lane, not the two of a 64-bit atomic. The adjoint source
    // (MODE 1) keeps 64-bit atomics.
    unsigned long long* qacc = reinterpret_cast<unsigned long long*>(sm);  // MODE 1: nr x Kz
    int* qhi = reinterpret_cast<int*>(sm);                                  // MODE 0: nr x Kz
    unsigned* qlo = reinterpret_cast<unsigned*>(sm) + kRowBlock * Kz;       //         nr x Kz
    float2* pk = sm + kRowBlock * Kz;                                       // nrp lines at pitch zp
    float2* tz = pk + (kRowBlock / 2) * zp;
    load_twiddles<kZThreads>(tz, twz, Kz);
    if (acc && x == 0 && yb == 0 && threadIdx.x == 0) acc[SPONGE_PART_RECIPROCAL] = 0.0;
    const int npt = nr * Kz;
    const float rkz = 1.0f / (float)Kz, rnr = 1.0f / (float)nr;
    for (int idx = threadIdx.x; idx < npt; idx += kZThreads) {
        if (MODE == 0) {
            qhi[idx] = 0;
            qlo[idx] = 0u;
        } else {
            qacc[idx] = 0ull;
        }
    }
    __syncthreads();
    // The four bins are walked as one flat range of records (base, fraction
    // and charge filed by the prep in bin order); each thread issues the
    // record loads of up to kSpreadPerThread atoms before any LDS update, so
    // their latencies overlap.
    const int o1 = min(c0, bin_cap);
    const int o2 = o1 + min(c1, bin_cap);
    const int o3 = o2 + min(c2, bin_cap);
    const int total = o3 + min(c3, bin_cap);
    for (int t0 = 0; t0 < total; t0 += kZThreads * kSpreadPerThread) {
        long sl[kSpreadPerThread];
        int akx[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            const int t = t0 + r * kZThreads + (int)threadIdx.x;
            sl[r] = -1;
            akx[r] = 0;
            if (t < total) {
                const int k = (t >= o1) + (t >= o2) + (t >= o3);
                const int b = k == 0 ? b0 : k == 1 ? b1 : k == 2 ? b2 : b3;
                const int o = k == 0 ? 0 : k == 1 ? o1 : k == 2 ? o2 : o3;
                sl[r] = (long)b * bin_cap + (t - o);
                akx[r] = k;
            }
        }
        // every candidate's record is loaded together (one memory round
        // trip), the rows are tested afterwards
        unsigned bb[kSpreadPerThread];
        float4 ff[kSpreadPerThread];
        int ai[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++)
            if (sl[r] >= 0) {
                bb[r] = rec_base[sl[r]];
                ff[r] = rec_frac[sl[r]];
                if (MODE) ai[r] = bin_atoms[sl[r]];
            }
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            if (sl[r] < 0) continue;
            // stencil rows by - ky, ky = 0..3, relative to the block
            int d = (int)((bb[r] >> 10) & 1023u) - y0;
            d += d < 0 ? Ky : 0;
            bool inside = false;
#pragma unroll
            for (int ky = 0; ky < 4; ky++) {
                int rr = d - ky;
                rr += rr < 0 ? Ky : 0;
                inside |= rr < nr;
            }
            if (!inside) continue;
            const float4 f = ff[r];
            const int bz = (int)(bb[r] >> 20);
            float wz[4], dz[4];
            int gz[4];
#pragma unroll
            for (int kz = 0; kz < 4; kz++) {
                wz[kz] = bsp_w(f.z, kz);
                if (MODE) dz[kz] = bsp_dw(f.z, kz);
                int tt = bz - kz;
                gz[kz] = tt + (tt < 0 ? Kz : 0);
            }
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            if (MODE) w = wv[ai[r]];
            const float qs = f.w * unit;
            const float ax = bsp_w(f.x, akx[r]);
            const float dax = MODE ? bsp_dw(f.x, akx[r]) : 0.f;
#pragma unroll
            for (int ky = 0; ky < 4; ky++) {
                int rr = d - ky;
                rr += rr < 0 ? Ky : 0;
                if (rr >= nr) continue;
                const float ay = bsp_w(f.y, ky);
                unsigned long long* row = qacc + rr * Kz;
                if (MODE == 0) {
                    const float wy = qs * ax * ay;
                    int* rhi = qhi + rr * Kz;
                    unsigned* rlo = qlo + rr * Kz;
#pragma unroll
                    for (int kz = 0; kz < 4; kz++) {
                        const int c = __float2int_rn(wy * wz[kz]);
                        atomicAdd(rhi + gz[kz], c >> 16);
                        atomicAdd(rlo + gz[kz], (unsigned)c & 0xffffu);
                    }
                } else {
                    const float day = bsp_dw(f.y, ky);
                    const float c1 = qs * (w.x * dax * ay + w.y * ax * day);
                    const float c2 = qs * w.z * ax * ay;
#pragma unroll
                    for (int kz = 0; kz < 4; kz++)
                        atomicAdd(row + gz[kz],
                                  (unsigned long long)(long long)__float2int_rn(fmaf(c1, wz[kz], c2 * dz[kz])));
                }
            }
        }
    }
    __syncthreads();
    float* pkf = reinterpret_cast<float*>(pk);
    for (int idx = threadIdx.x; idx < npt; idx += kZThreads) {
        const int r = div_small(idx, rkz), z = idx - r * Kz;
        const long long q =
            MODE == 0 ? (long long)qhi[idx] * 65536ll + (long long)qlo[idx] : (long long)qacc[idx];
        pkf[2 * ((r >> 1) * zp + pd(z)) + (r & 1)] = (float)((double)q / (double)unit);
    }
    if (nr & 1)
        for (int z = threadIdx.x; z < Kz; z += kZThreads) pkf[2 * ((nr >> 1) * zp + pd(z)) + 1] = 0.f;
    __syncthreads();
    lds_fft<-1, kZThreads, RM>(pk, nrp, zp, 1, pz, tz);
    __syncthreads();
    // Unpack the paired transforms Z = FFT(a + i b): A_k = (Z_k + conj Z_-k)/2,
    // B_k = (Z_k - conj Z_-k)/(2i); consecutive threads take consecutive rows
    // so the spectrum is written in runs of nr y values.
    for (int idx = threadIdx.x; idx < nr * H; idx += kZThreads) {
        const int k = div_small(idx, rnr), r = idx - k * nr;
        const int p = r >> 1;
        const float2 zk = pk[p * zp + pd(k)];
        const float2 zm = pk[p * zp + pd(k == 0 ? 0 : Kz - k)];
        const float2 v = (r & 1) == 0 ? make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y))
                                      : make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        FQ[((size_t)k * Kx + x) * Ky + y0 + r] = v;
    }
}

template <int RM>
__global__ __launch_bounds__(kZThreads) SNB_PME_CHAIN void spread_z_kernel(
    const unsigned* __restrict__ rec_base, const float4* __restrict__ rec_frac, const int* __restrict__ bin_count,
    const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz, int nyb,
    FftPlan pz, const float2* __restrict__ twz, double* __restrict__ acc) {
    const int x = blockIdx.x / nyb;
    spread_z_body<0, RM>(x, blockIdx.x - x * nyb, rec_base, rec_frac, bin_count, bin_atoms, bin_cap, FQ, Kx, Ky, Kz,
                         pz, twz, acc);
}

__global__ __launch_bounds__(kZThreads) void spread_z_adjoint_kernel(
    const unsigned* __restrict__ rec_base, const float4* __restrict__ rec_frac, const int* __restrict__ bin_count,
    const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz, int nyb,
    FftPlan pz, const float2* __restrict__ twz, const float4* __restrict__ wv, const float* __restrict__ scale_ptr) {
    const int x = blockIdx.x / nyb;
    spread_z_body<1>(x, blockIdx.x - x * nyb, rec_base, rec_frac, bin_count, bin_atoms, bin_cap, FQ, Kx, Ky, Kz, pz,
                     twz, nullptr, wv, scale_ptr);
}

// Batched: local block l = x * (row blocks) + row block.
__global__ __launch_bounds__(kZThreads) void spread_z_batch(const BatchSys* __restrict__ bs,
                                                            const int2* __restrict__ map, int energy) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    const int nyb = row_blocks(P.K[1]);
    const int x = m.y / nyb;
    spread_z_body(x, m.y - x * nyb, P.bin_base, P.bin_frac, P.bin_count, P.bin_atoms, P.bin_cap, P.FQ, P.K[0], P.K[1],
                  P.K[2], b.pz, P.d_tw[2], energy ? P.acc : nullptr);
}

// One block per kz: the Kx x Ky plane of the spectrum in LDS (padded rows
// and columns), FFT along y then x, energy 1/2 sum m BC |F|^2 (FP64; m = 2 for
// the kz that stand for their conjugates too), multiply by BC, inverse FFT
// along x then y.
template <int RM = kRadixAll, int NT = kConvThreads>
__device__ __forceinline__ void conv2d_body(int kz, float2* __restrict__ FQ, const float* __restrict__ bc, int Kx,
                                            int Ky, int Kz, const FftPlan& px, const FftPlan& py,
                                            const float2* __restrict__ twx, const float2* __restrict__ twy, int energy,
                                            double* __restrict__ acc, int* __restrict__ bin_count) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    __shared__ double red[NT / 64];
    const int pitch = line_pitch(Ky);
    float2* tx = sm + padded_len(Kx) * pitch;
    float2* ty = tx + Kx;
    // the spread has consumed the x-plane bins: clear them for the next prep
    if (kz == 0)
        for (int b = threadIdx.x; b < Kx; b += NT) bin_count[b] = 0;
    float2* plane = FQ + (size_t)kz * Kx * Ky;
    const float* bcp = bc + (size_t)kz * Kx * Ky;
    const int np = Kx * Ky;
    const float rky = 1.0f / (float)Ky;
    // The plane and its BC values are loaded in one memory round trip when
    // they fit CB points per thread (4,096 points: a 64 x 64 plane with 256
    // or 512 threads; the BC values wait in registers through the forward
    // transforms): with one block per CU nothing else hides the memory
    // latency.
    constexpr int CB = kConvBatch * (512 / NT);
    float bv[CB];
    const bool onepass = np <= CB * NT;
    for (int s0 = threadIdx.x; s0 < np; s0 += CB * NT) {
        float2 v[CB];
#pragma unroll
        for (int s = 0; s < CB; s++) {
            const int idx = s0 + s * NT;
            if (idx < np) {
                v[s] = plane[idx];
                if (onepass) bv[s] = bcp[idx];
            }
        }
        if (s0 == (int)threadIdx.x) {
            // the twiddle loads join the first round trip
            load_twiddles<NT>(tx, twx, Kx);
            load_twiddles<NT>(ty, twy, Ky);
        }
#pragma unroll
        for (int s = 0; s < CB; s++) {
            const int idx = s0 + s * NT;
            if (idx < np) {
                sm[plane_lds(idx, rky, Ky, pitch)] = v[s];
            }
        }
    }
    __syncthreads();
    lds_fft<-1, NT, RM>(sm, Kx, pitch, 1, py, ty);
    __syncthreads();
    lds_fft<-1, NT, RM>(sm, Ky, 1, pitch, px, tx);
    __syncthreads();
    double e = 0.0;
    for (int s0 = threadIdx.x; s0 < np; s0 += CB * NT) {
        if (!onepass) {
#pragma unroll
            for (int s = 0; s < CB; s++) {
                const int idx = s0 + s * NT;
                if (idx < np) bv[s] = bcp[idx];
            }
        }
#pragma unroll
        for (int s = 0; s < CB; s++) {
            const int idx = s0 + s * NT;
            if (idx < np) {
                const int o = plane_lds(idx, rky, Ky, pitch);
                const float b = bv[s];
                const float2 w = sm[o];
                if (energy) e += (double)b * ((double)w.x * w.x + (double)w.y * w.y);
                sm[o] = make_float2(w.x * b, w.y * b);
            }
        }
    }
    __syncthreads();
    lds_fft<1, NT, RM>(sm, Ky, 1, pitch, px, tx);
    __syncthreads();
    lds_fft<1, NT, RM>(sm, Kx, pitch, 1, py, ty);
    __syncthreads();
    for (int idx = threadIdx.x; idx < np; idx += NT) {
        plane[idx] = sm[plane_lds(idx, rky, Ky, pitch)];
    }
    if (energy) {
        const double t = block_sum<NT>(e, red);
        const double m = (kz == 0 || 2 * kz == Kz) ? 1.0 : 2.0;
        if (threadIdx.x == 0) atomicAdd(acc + SPONGE_PART_RECIPROCAL, 0.5 * m * t);
    }
}

template <int RM>
__global__ __launch_bounds__(kConvThreads) SNB_PME_CHAIN void conv2d_kernel(float2* __restrict__ FQ, const float* __restrict__ bc,
                                                              int Kx, int Ky, int Kz, FftPlan px, FftPlan py,
                                                              const float2* __restrict__ twx,
                                                              const float2* __restrict__ twy, int energy,
                                                              double* __restrict__ acc, int* __restrict__ bin_count) {
    conv2d_body<RM>(blockIdx.x, FQ, bc, Kx, Ky, Kz, px, py, twx, twy, energy, acc, bin_count);
}

// The same with 512 threads (8 waves) per plane: more waves to hide the LDS
// latency of the transform passes, where the CUs are free to hold them.
template <int RM>
__global__ __launch_bounds__(2 * kConvThreads) SNB_PME_CHAIN void conv2d_wide_kernel(
    float2* __restrict__ FQ, const float* __restrict__ bc, int Kx, int Ky, int Kz, FftPlan px, FftPlan py,
    const float2* __restrict__ twx, const float2* __restrict__ twy, int energy, double* __restrict__ acc,
    int* __restrict__ bin_count) {
    conv2d_body<RM, 2 * kConvThreads>(blockIdx.x, FQ, bc, Kx, Ky, Kz, px, py, twx, twy, energy, acc, bin_count);
}

__global__ __launch_bounds__(kConvThreads) void conv2d_batch(const BatchSys* __restrict__ bs,
                                                             const int2* __restrict__ map, int energy) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    conv2d_body(m.y, P.FQ, P.d_bc, P.K[0], P.K[1], P.K[2], b.px, b.py, P.d_tw[0], P.d_tw[1], energy, P.acc,
                P.bin_count);
}

// Split convolution for large planes: one workgroup per plane leaves most of
// the chip idle (55 planes of 108 x 108 at a 108^3 grid) and cannot hold
// planes past ~128 x 128 in LDS, so the same steps run as three launches of
// kLines lines per workgroup: y-lines forward, x-lines forward + energy + BC +
// inverse, y-lines inverse. Each line is read and written once per launch.
constexpr int kLines = 16;
constexpr int kLineThreads = 256;
// largest Kx * Ky that runs as one workgroup per plane by default
constexpr int kConvPlaneMax = 64 * 64;

// kLines consecutive x rows of one kz plane (contiguous in FQ), transformed
// along y in place.
template <int SIGN, int RM>
__device__ __forceinline__ void conv_y_body(int bx, int kz, float2* __restrict__ FQ, int Kx, int Ky, const FftPlan& py,
                                            const float2* __restrict__ twy) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    const int pitch = line_pitch(Ky);
    float2* ty = sm + kLines * pitch;
    const int x0 = bx * kLines, nl = min(kLines, Kx - x0);
    float2* rows = FQ + ((size_t)kz * Kx + x0) * Ky;
    const int np = nl * Ky;
    const float rky = 1.0f / (float)Ky;
    // kConvBatch loads per thread in flight at a time; the twiddle loads
    // join the first round
    for (int s0 = threadIdx.x; s0 < np; s0 += kConvBatch * kLineThreads) {
        float2 v[kConvBatch];
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * kLineThreads;
            if (idx < np) v[s] = rows[idx];
        }
        if (s0 == (int)threadIdx.x) load_twiddles<kLineThreads>(ty, twy, Ky);
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * kLineThreads;
            if (idx < np) {
                const int r = div_small(idx, rky), y = idx - r * Ky;
                sm[r * pitch + pd(y)] = v[s];
            }
        }
    }
    __syncthreads();
    lds_fft<SIGN, kLineThreads, RM>(sm, nl, pitch, 1, py, ty);
    __syncthreads();
    for (int idx = threadIdx.x; idx < np; idx += kLineThreads) {
        const int r = div_small(idx, rky), y = idx - r * Ky;
        rows[idx] = sm[r * pitch + pd(y)];
    }
}

template <int RM>
__global__ __launch_bounds__(kLineThreads) SNB_PME_CHAIN void conv_y_forward_kernel(float2* __restrict__ FQ, int Kx, int Ky,
                                                                      FftPlan py, const float2* __restrict__ twy) {
    conv_y_body<-1, RM>(blockIdx.x, blockIdx.y, FQ, Kx, Ky, py, twy);
}

template <int RM>
__global__ __launch_bounds__(kLineThreads) SNB_PME_CHAIN void conv_y_inverse_kernel(float2* __restrict__ FQ, int Kx, int Ky,
                                                                      FftPlan py, const float2* __restrict__ twy) {
    conv_y_body<1, RM>(blockIdx.x, blockIdx.y, FQ, Kx, Ky, py, twy);
}

// kLines consecutive y columns of one kz plane: forward x transform, energy
// 1/2 sum m BC |F|^2 (FP64, as conv2d_body), multiply by BC, inverse x
// transform. Column c of the block is LDS line c (x contiguous).
template <int RM>
__device__ __forceinline__ void conv_x_body(int by, int kz, float2* __restrict__ FQ, const float* __restrict__ bc,
                                            int Kx, int Ky, int Kz, const FftPlan& px, const float2* __restrict__ twx,
                                            int energy, double* __restrict__ acc, int* __restrict__ bin_count) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    __shared__ double red[kLineThreads / 64];
    const int pitch = line_pitch(Kx);
    float2* tx = sm + kLines * pitch;
    const int y0 = by * kLines, nc = min(kLines, Ky - y0);
    // the spread has consumed the x-plane bins: clear them for the next prep
    if (kz == 0 && by == 0)
        for (int b = threadIdx.x; b < Kx; b += kLineThreads) bin_count[b] = 0;
    float2* cols = FQ + (size_t)kz * Kx * Ky + y0;
    const float* bcp = bc + (size_t)kz * Kx * Ky + y0;
    const int np = Kx * kLines;
    // BC values of up to kConvBatch points per thread wait in registers
    // through the forward transform
    float bv[kConvBatch];
    const bool onepass = np <= kConvBatch * kLineThreads;
    for (int s0 = threadIdx.x; s0 < np; s0 += kConvBatch * kLineThreads) {
        float2 v[kConvBatch];
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * kLineThreads, x = idx / kLines, c = idx % kLines;
            if (idx < np && c < nc) {
                v[s] = cols[(size_t)x * Ky + c];
                if (onepass) bv[s] = bcp[(size_t)x * Ky + c];
            }
        }
        if (s0 == (int)threadIdx.x) load_twiddles<kLineThreads>(tx, twx, Kx);  // same round trip
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * kLineThreads, x = idx / kLines, c = idx % kLines;
            if (idx < np && c < nc) sm[c * pitch + pd(x)] = v[s];
        }
    }
    __syncthreads();
    lds_fft<-1, kLineThreads, RM>(sm, nc, pitch, 1, px, tx);
    __syncthreads();
    double e = 0.0;
    for (int s0 = threadIdx.x; s0 < np; s0 += kConvBatch * kLineThreads) {
        if (!onepass) {
#pragma unroll
            for (int s = 0; s < kConvBatch; s++) {
                const int idx = s0 + s * kLineThreads, x = idx / kLines, c = idx % kLines;
                if (idx < np && c < nc) bv[s] = bcp[(size_t)x * Ky + c];
            }
        }
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * kLineThreads, x = idx / kLines, c = idx % kLines;
            if (idx < np && c < nc) {
                const float b = bv[s];
                const float2 w = sm[c * pitch + pd(x)];
                if (energy) e += (double)b * ((double)w.x * w.x + (double)w.y * w.y);
                sm[c * pitch + pd(x)] = make_float2(w.x * b, w.y * b);
            }
        }
    }
    __syncthreads();
    lds_fft<1, kLineThreads, RM>(sm, nc, pitch, 1, px, tx);
    __syncthreads();
    for (int idx = threadIdx.x; idx < np; idx += kLineThreads) {
        const int x = idx / kLines, c = idx % kLines;
        if (c < nc) cols[(size_t)x * Ky + c] = sm[c * pitch + pd(x)];
    }
    if (energy) {
        const double t = block_sum<kLineThreads>(e, red);
        const double m = (kz == 0 || 2 * kz == Kz) ? 1.0 : 2.0;
        if (threadIdx.x == 0) atomicAdd(acc + SPONGE_PART_RECIPROCAL, 0.5 * m * t);
    }
}

template <int RM>
__global__ __launch_bounds__(kLineThreads) SNB_PME_CHAIN void conv_x_kernel(float2* __restrict__ FQ, const float* __restrict__ bc,
                                                              int Kx, int Ky, int Kz, FftPlan px,
                                                              const float2* __restrict__ twx, int energy,
                                                              double* __restrict__ acc, int* __restrict__ bin_count) {
    conv_x_body<RM>(blockIdx.x, blockIdx.y, FQ, bc, Kx, Ky, Kz, px, twx, energy, acc, bin_count);
}

// Inverse z transform (C2R) of one block of kRowBlock y rows of one x plane:
// rows (2p, 2p+1) are repacked into one complex line Z = A + i B over all kz,
// each row extended by Hermitian symmetry (self-conjugate bins keep only their
// real part, the C2R convention), transformed, and written to the real grid
// R [x][y][z].
template <int RM = kRadixAll>
__device__ __forceinline__ void inv_z_body(int x, int yb, const float2* __restrict__ FQ, float* __restrict__ R, int Kx,
                                           int Ky, int Kz, const FftPlan& pz, const float2* __restrict__ twz) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    const int H = Kz / 2 + 1, zp = line_pitch(Kz);
    const int y0 = yb * kRowBlock, nr = min(kRowBlock, Ky - y0), nrp = (nr + 1) / 2;
    float2* rows = sm;                      // nr x H
    float2* pk = sm + kRowBlock * H;        // nrp lines at pitch zp
    float2* tz = pk + (kRowBlock / 2) * zp;
    const float rkz = 1.0f / (float)Kz, rnr = 1.0f / (float)nr;
    float2 v[kZPer];
#pragma unroll
    for (int s = 0; s < kZPer; s++) {
        const int idx = threadIdx.x + s * kZThreads;
        if (idx < nr * H) {
            const int k = div_small(idx, rnr), r = idx - k * nr;
            v[s] = FQ[((size_t)k * Kx + x) * Ky + y0 + r];
        }
    }
    // the twiddle loads join the same round trip
    load_twiddles<kZThreads>(tz, twz, Kz);
#pragma unroll
    for (int s = 0; s < kZPer; s++) {
        const int idx = threadIdx.x + s * kZThreads;
        if (idx < nr * H) {
            const int k = div_small(idx, rnr), r = idx - k * nr;
            rows[r * H + k] = v[s];
        }
    }
    __syncthreads();
    for (int o = threadIdx.x; o < nrp * Kz; o += kZThreads) {
        const int p = div_small(o, rkz), k = o - p * Kz;
        const bool hi = k > Kz / 2;
        const int kk = hi ? Kz - k : k;
        const bool selfc = (kk == 0) || (2 * kk == Kz);
        float2 A = rows[(2 * p) * H + kk];
        float2 B = (2 * p + 1 < nr) ? rows[(2 * p + 1) * H + kk] : make_float2(0.f, 0.f);
        if (hi) {
            A.y = -A.y;
            B.y = -B.y;
        }
        if (selfc) {
            A.y = 0.f;
            B.y = 0.f;
        }
        pk[p * zp + pd(k)] = make_float2(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    lds_fft<1, kZThreads, RM>(pk, nrp, zp, 1, pz, tz);
    __syncthreads();
    float* Rp = R + ((size_t)x * Ky + y0) * Kz;
    for (int idx = threadIdx.x; idx < nrp * Kz; idx += kZThreads) {
        const int p = div_small(idx, rkz), z = idx - p * Kz;
        const float2 v = pk[p * zp + pd(z)];
        Rp[(2 * p) * Kz + z] = v.x;
        if (2 * p + 1 < nr) Rp[(2 * p + 1) * Kz + z] = v.y;
    }
}

template <int RM>
__global__ __launch_bounds__(kZThreads) SNB_PME_CHAIN void inv_z_kernel(const float2* __restrict__ FQ, float* __restrict__ R, int Kx,
                                                          int Ky, int Kz, int nyb, FftPlan pz,
                                                          const float2* __restrict__ twz) {
    const int x = blockIdx.x / nyb;
    inv_z_body<RM>(x, blockIdx.x - x * nyb, FQ, R, Kx, Ky, Kz, pz, twz);
}

__global__ __launch_bounds__(kZThreads) void inv_z_batch(const BatchSys* __restrict__ bs,
                                                         const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    const int nyb = row_blocks(P.K[1]);
    const int x = m.y / nyb;
    inv_z_body(x, m.y - x * nyb, P.FQ, P.R, P.K[0], P.K[1], P.K[2], b.pz, P.d_tw[2]);
}

// ---------------------------------------------------------------------------
// Slab form (convolution form 2): one workgroup per x plane takes the z and
// y transforms of its Ky x Kz slab, so that the convolution is left with the
// x lines (conv_x) and the reciprocal chain is three launches:
//   spread_yz  spread of every atom filed under planes x..x+3 (all its
//              stencil rows fall in the slab) into the exact fixed-point sums
//              of spread_z, paired real rows -> complex z-lines, z-FFT,
//              R2C unpack into a [kz][y] tile, y-FFT, store FQ[kz][x][y]
//   conv_x     x-FFT, energy, BC, inverse x-FFT (as in the split form)
//   inv_yz     load the tile, inverse y-FFT, Hermitian repack into paired
//              complex z-lines, inverse z-FFT (C2R), store R[x][y][z]
// Same arithmetic as spread_z + conv_y_forward and conv_y_inverse + inv_z,
// with the y lines in LDS instead of a round trip through FQ.
constexpr int kYZThreads = 512;

// LDS (float2 units): max(accumulators, y tile) | z-lines | twiddles z, y.
// Lines of a transform sit at padded line positions (line l at pd(l) x
// pitch, as lds_fft addresses them).
__host__ __device__ __forceinline__ int yz_tile(int Ky, int Kz) { return padded_len(Kz / 2 + 1) * line_pitch(Ky); }
__host__ __device__ __forceinline__ int yz_lines(int Ky, int Kz) { return padded_len((Ky + 1) / 2) * line_pitch(Kz); }
__host__ __device__ __forceinline__ int yz_acc(int Ky, int Kz) {
    const int a = Ky * Kz, t = yz_tile(Ky, Kz);
    return a > t ? a : t;
}

template <int RM = kRadixAll>
__device__ __forceinline__ void spread_yz_body(int x, const unsigned* __restrict__ rec_base,
                                               const float4* __restrict__ rec_frac,
                                               const int* __restrict__ bin_count, int bin_cap,
                                               float2* __restrict__ FQ, int Kx, int Ky, int Kz, const FftPlan& pz,
                                               const FftPlan& py, const float2* __restrict__ twz,
                                               const float2* __restrict__ twy, double* __restrict__ acc) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    constexpr int NT = kYZThreads;
    int b0 = x, b1 = x + 1, b2 = x + 2, b3 = x + 3;
    b1 -= b1 >= Kx ? Kx : 0;
    b2 -= b2 >= Kx ? Kx : 0;
    b3 -= b3 >= Kx ? Kx : 0;
    const int c0 = bin_count[b0], c1 = bin_count[b1], c2 = bin_count[b2], c3 = bin_count[b3];
    const int H = Kz / 2 + 1, zp = line_pitch(Kz), yp = line_pitch(Ky), nrp = (Ky + 1) / 2;
    int* qhi = reinterpret_cast<int*>(sm);                        // Ky x Kz
    unsigned* qlo = reinterpret_cast<unsigned*>(sm) + Ky * Kz;    // Ky x Kz
    float2* T = sm;                                               // H x yp, after the accumulators
    float2* pk = sm + yz_acc(Ky, Kz);                             // nrp lines at pitch zp
    float2* tz = pk + yz_lines(Ky, Kz);
    float2* ty = tz + Kz;
    load_twiddles<NT>(tz, twz, Kz);
    load_twiddles<NT>(ty, twy, Ky);
    if (acc && x == 0 && threadIdx.x == 0) acc[SPONGE_PART_RECIPROCAL] = 0.0;
    const int npt = Ky * Kz;
    for (int idx = threadIdx.x; idx < npt; idx += NT) {
        qhi[idx] = 0;
        qlo[idx] = 0u;
    }
    __syncthreads();
    const int o1 = min(c0, bin_cap);
    const int o2 = o1 + min(c1, bin_cap);
    const int o3 = o2 + min(c2, bin_cap);
    const int total = o3 + min(c3, bin_cap);
    const float unit = kSpreadScale;
    for (int t0 = 0; t0 < total; t0 += NT * kSpreadPerThread) {
        long sl[kSpreadPerThread];
        int akx[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            const int t = t0 + r * NT + (int)threadIdx.x;
            sl[r] = -1;
            akx[r] = 0;
            if (t < total) {
                const int k = (t >= o1) + (t >= o2) + (t >= o3);
                const int b = k == 0 ? b0 : k == 1 ? b1 : k == 2 ? b2 : b3;
                const int o = k == 0 ? 0 : k == 1 ? o1 : k == 2 ? o2 : o3;
                sl[r] = (long)b * bin_cap + (t - o);
                akx[r] = k;
            }
        }
        unsigned bb[kSpreadPerThread];
        float4 ff[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++)
            if (sl[r] >= 0) {
                bb[r] = rec_base[sl[r]];
                ff[r] = rec_frac[sl[r]];
            }
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            if (sl[r] < 0) continue;
            const float4 f = ff[r];
            const int by = (int)((bb[r] >> 10) & 1023u), bz = (int)(bb[r] >> 20);
            float wz[4];
            int gz[4];
#pragma unroll
            for (int kz = 0; kz < 4; kz++) {
                wz[kz] = bsp_w(f.z, kz);
                int tt = bz - kz;
                gz[kz] = tt + (tt < 0 ? Kz : 0);
            }
            const float qs = f.w * unit;
            const float ax = bsp_w(f.x, akx[r]);
#pragma unroll
            for (int ky = 0; ky < 4; ky++) {
                int rr = by - ky;
                rr += rr < 0 ? Ky : 0;
                const float wy = qs * ax * bsp_w(f.y, ky);
                int* rhi = qhi + rr * Kz;
                unsigned* rlo = qlo + rr * Kz;
#pragma unroll
                for (int kz = 0; kz < 4; kz++) {
                    const int cq = __float2int_rn(wy * wz[kz]);
                    atomicAdd(rhi + gz[kz], cq >> 16);
                    atomicAdd(rlo + gz[kz], (unsigned)cq & 0xffffu);
                }
            }
        }
    }
    __syncthreads();
    // rows (2p, 2p + 1) -> the real and imaginary parts of z-line p
    float* pkf = reinterpret_cast<float*>(pk);
    const float rkz = 1.0f / (float)Kz, rky = 1.0f / (float)Ky;
    for (int idx = threadIdx.x; idx < npt; idx += NT) {
        const int r = div_small(idx, rkz), z = idx - r * Kz;
        const long long q = (long long)qhi[idx] * 65536ll + (long long)qlo[idx];
        pkf[2 * (pd(r >> 1) * zp + pd(z)) + (r & 1)] = (float)((double)q / (double)unit);
    }
    if (Ky & 1)
        for (int z = threadIdx.x; z < Kz; z += NT) pkf[2 * (pd(Ky >> 1) * zp + pd(z)) + 1] = 0.f;
    __syncthreads();
    lds_fft<-1, NT, RM>(pk, nrp, zp, 1, pz, tz);
    __syncthreads();
    // R2C unpack of the paired transforms into the [kz][y] tile (the
    // accumulators are dead: the tile takes their place)
    for (int idx = threadIdx.x; idx < Ky * H; idx += NT) {
        const int k = div_small(idx, rky), r = idx - k * Ky;
        const int pl = pd(r >> 1) * zp;
        const float2 zk = pk[pl + pd(k)];
        const float2 zm = pk[pl + pd(k == 0 ? 0 : Kz - k)];
        T[pd(k) * yp + pd(r)] = (r & 1) == 0 ? make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y))
                                             : make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
    __syncthreads();
    lds_fft<-1, NT, RM>(T, H, yp, 1, py, ty);
    __syncthreads();
    for (int idx = threadIdx.x; idx < Ky * H; idx += NT) {
        const int k = div_small(idx, rky), y = idx - k * Ky;
        FQ[((size_t)k * Kx + x) * Ky + y] = T[pd(k) * yp + pd(y)];
    }
}

template <int RM>
__global__ __launch_bounds__(kYZThreads) SNB_PME_CHAIN void spread_yz_kernel(
    const unsigned* __restrict__ rec_base, const float4* __restrict__ rec_frac, const int* __restrict__ bin_count,
    int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz, FftPlan pz, FftPlan py,
    const float2* __restrict__ twz, const float2* __restrict__ twy, double* __restrict__ acc) {
    spread_yz_body<RM>(blockIdx.x, rec_base, rec_frac, bin_count, bin_cap, FQ, Kx, Ky, Kz, pz, py, twz, twy, acc);
}

template <int RM = kRadixAll>
__device__ __forceinline__ void inv_yz_body(int x, const float2* __restrict__ FQ, float* __restrict__ R, int Kx,
                                            int Ky, int Kz, const FftPlan& pz, const FftPlan& py,
                                            const float2* __restrict__ twz, const float2* __restrict__ twy) {
    __builtin_amdgcn_s_setprio(kPmeWavePriority);
    extern __shared__ float2 sm[];
    constexpr int NT = kYZThreads;
    const int H = Kz / 2 + 1, zp = line_pitch(Kz), yp = line_pitch(Ky), nrp = (Ky + 1) / 2;
    float2* T = sm;                    // H x yp
    float2* pk = sm + yz_tile(Ky, Kz); // nrp lines at pitch zp
    float2* tz = pk + yz_lines(Ky, Kz);
    float2* ty = tz + Kz;
    const float rky = 1.0f / (float)Ky, rkz = 1.0f / (float)Kz;
    const int np = Ky * H;
    // kConvBatch loads per thread in flight at a time; the twiddle loads
    // join the first round
    for (int s0 = threadIdx.x; s0 < np; s0 += kConvBatch * NT) {
        float2 v[kConvBatch];
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * NT;
            if (idx < np) {
                const int k = div_small(idx, rky), y = idx - k * Ky;
                v[s] = FQ[((size_t)k * Kx + x) * Ky + y];
            }
        }
        if (s0 == (int)threadIdx.x) {
            load_twiddles<NT>(tz, twz, Kz);
            load_twiddles<NT>(ty, twy, Ky);
        }
#pragma unroll
        for (int s = 0; s < kConvBatch; s++) {
            const int idx = s0 + s * NT;
            if (idx < np) {
                const int k = div_small(idx, rky), y = idx - k * Ky;
                T[pd(k) * yp + pd(y)] = v[s];
            }
        }
    }
    __syncthreads();
    lds_fft<1, NT, RM>(T, H, yp, 1, py, ty);
    __syncthreads();
    // rows (2p, 2p + 1) repacked into one complex z-line over all kz, each
    // extended by Hermitian symmetry (self-conjugate bins keep their real
    // part), as inv_z_body does
    for (int o = threadIdx.x; o < nrp * Kz; o += NT) {
        const int p = div_small(o, rkz), k = o - p * Kz;
        const bool hi = k > Kz / 2;
        const int kk = hi ? Kz - k : k;
        const bool selfc = (kk == 0) || (2 * kk == Kz);
        const int tl = pd(kk) * yp;
        float2 A = T[tl + pd(2 * p)];
        float2 B = (2 * p + 1 < Ky) ? T[tl + pd(2 * p + 1)] : make_float2(0.f, 0.f);
        if (hi) {
            A.y = -A.y;
            B.y = -B.y;
        }
        if (selfc) {
            A.y = 0.f;
            B.y = 0.f;
        }
        pk[pd(p) * zp + pd(k)] = make_float2(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    lds_fft<1, NT, RM>(pk, nrp, zp, 1, pz, tz);
    __syncthreads();
    float* Rp = R + (size_t)x * Ky * Kz;
    for (int idx = threadIdx.x; idx < nrp * Kz; idx += NT) {
        const int p = div_small(idx, rkz), z = idx - p * Kz;
        const float2 v = pk[pd(p) * zp + pd(z)];
        Rp[(2 * p) * Kz + z] = v.x;
        if (2 * p + 1 < Ky) Rp[(2 * p + 1) * Kz + z] = v.y;
    }
}

template <int RM>
__global__ __launch_bounds__(kYZThreads) SNB_PME_CHAIN void inv_yz_kernel(const float2* __restrict__ FQ,
                                                                          float* __restrict__ R, int Kx, int Ky,
                                                                          int Kz, FftPlan pz, FftPlan py,
                                                                          const float2* __restrict__ twz,
                                                                          const float2* __restrict__ twy) {
    inv_yz_body<RM>(blockIdx.x, FQ, R, Kx, Ky, Kz, pz, py, twz, twy);
}

// PME_Final (PME_force.cu:485-550): 16 lanes per atom (lane = kx, ky), each
// walking the 4 kz points; force += -q sum R dtheta (K/L).
__device__ __forceinline__ void pme_gather_body(int blk, int n, const unsigned* __restrict__ base,
                                                const float4* __restrict__ frac, const float* __restrict__ R, int Kx,
                                                int Ky, int Kz, float ibx, float iby, float ibz,
                                                float* __restrict__ frc) {
    const int gid = blk * kGatherThreads + threadIdx.x;
    const int i = gid >> 4, sub = gid & 15;
    const int kx = sub >> 2, ky = sub & 3;
    float sx = 0.f, sy = 0.f, sz = 0.f, qi = 0.f;
    // the force being added to is read with the atom's base and fraction
    // (one memory round trip before the grid reads)
    float f0x = 0.f, f0y = 0.f, f0z = 0.f;
    if (i < n && sub == 0) {
        f0x = frc[3 * i];
        f0y = frc[3 * i + 1];
        f0z = frc[3 * i + 2];
    }
    if (i < n) {
        const unsigned bb = base[i];
        const int4 b = make_int4((int)(bb & 1023u), (int)((bb >> 10) & 1023u), (int)(bb >> 20), 0);
        float4 f = frac[i];
        qi = f.w;
        float wx = bsp_w(f.x, kx), dwx = bsp_dw(f.x, kx);
        float wy = bsp_w(f.y, ky), dwy = bsp_dw(f.y, ky);
        int gx = b.x - kx, gy = b.y - ky;
        gx += gx < 0 ? Kx : 0;
        gy += gy < 0 ? Ky : 0;
        const float* row = R + ((size_t)gx * Ky + gy) * Kz;
#pragma unroll
        for (int kz = 0; kz < 4; kz++) {
            int gz = b.z - kz;
            gz += gz < 0 ? Kz : 0;
            float wz = bsp_w(f.z, kz), dwz = bsp_dw(f.z, kz);
            float phi = row[gz];
            sx += phi * dwx * wy * wz;
            sy += phi * wx * dwy * wz;
            sz += phi * wx * wy * dwz;
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        sx += __shfl_xor(sx, o);
        sy += __shfl_xor(sy, o);
        sz += __shfl_xor(sz, o);
    }
    if (i < n && sub == 0) {
        frc[3 * i] = f0x + -qi * sx * ibx;
        frc[3 * i + 1] = f0y + -qi * sy * iby;
        frc[3 * i + 2] = f0z + -qi * sz * ibz;
    }
}

// Adjoint gathers, 16 lanes per atom as pme_gather_body:
// out[i] += sign * sum_g R[g] m_i(g), m = M (MODE 0) or wv_i . grad M (MODE 1,
// wv carrying K/L per axis like the force).
template <int MODE>
__global__ __launch_bounds__(kGatherThreads) void pme_adj_gather_kernel(
    int n, const unsigned* __restrict__ base, const float4* __restrict__ frac, const float* __restrict__ R, int Kx,
    int Ky, int Kz, const float4* __restrict__ wv, float sign, float* __restrict__ out) {
    const int gid = blockIdx.x * kGatherThreads + threadIdx.x;
    const int i = gid >> 4, sub = gid & 15;
    const int kx = sub >> 2, ky = sub & 3;
    float acc = 0.f;
    if (i < n) {
        const unsigned bb = base[i];
        const int bx = (int)(bb & 1023u), by = (int)((bb >> 10) & 1023u), bz = (int)(bb >> 20);
        const float4 f = frac[i];
        const float wx = bsp_w(f.x, kx), wy = bsp_w(f.y, ky);
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        float dwx = 0.f, dwy = 0.f;
        if (MODE) {
            w = wv[i];
            dwx = bsp_dw(f.x, kx);
            dwy = bsp_dw(f.y, ky);
        }
        int gx = bx - kx, gy = by - ky;
        gx += gx < 0 ? Kx : 0;
        gy += gy < 0 ? Ky : 0;
        const float* row = R + ((size_t)gx * Ky + gy) * Kz;
#pragma unroll
        for (int kz = 0; kz < 4; kz++) {
            int gz = bz - kz;
            gz += gz < 0 ? Kz : 0;
            const float phi = row[gz];
            const float wz = bsp_w(f.z, kz);
            if (MODE == 0) {
                acc += phi * wx * wy * wz;
            } else {
                const float dwz = bsp_dw(f.z, kz);
                acc += phi * (w.x * dwx * wy * wz + w.y * wx * dwy * wz + w.z * wx * wy * dwz);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) acc += __shfl_xor(acc, o);
    if (i < n && sub == 0) out[i] += sign * acc;
}

// step_counter (the fused step's count of completed steps since the last
// list rebuild, or NULL) is advanced here: the gather is the step's last
// kernel, so the count stays fixed while the prep and pair kernels read it.
__global__ __launch_bounds__(kGatherThreads) void pme_gather_kernel(int n, const unsigned* __restrict__ base,
                                                                    const float4* __restrict__ frac,
                                                                    const float* __restrict__ R, int Kx, int Ky,
                                                                    int Kz, float ibx, float iby, float ibz,
                                                                    float* __restrict__ frc, int* step_counter) {
    if (step_counter && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step_counter, 1);
    pme_gather_body(blockIdx.x, n, base, frac, R, Kx, Ky, Kz, ibx, iby, ibz, frc);
}

__global__ __launch_bounds__(kGatherThreads) void pme_gather_batch(const BatchSys* __restrict__ bs,
                                                                   const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    pme_gather_body(m.y, P.n, P.base, P.frac, P.R, P.K[0], P.K[1], P.K[2], (float)P.K[0] / P.box[0],
                    (float)P.K[1] / P.box[1], (float)P.K[2] / P.box[2], b.io.frc);
}

// Energy / virial bookkeeping of PME_Reciprocal_Force_With_Energy_And_Virial
// (PME_force.cu:753-771) with FP64 sums: potential += corr + self + rec,
// virial += direct + corr + self + rec.
__global__ void pme_finalize_kernel(int n, const float* __restrict__ dir_atom, const float* __restrict__ corr_atom,
                                    double beta, double nf, int need_energy, int need_virial,
                                    float* __restrict__ d_potential, float* __restrict__ d_virial,
                                    double* __restrict__ acc) {
    __shared__ double red[16];
    double sd = 0.0, sc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (dir_atom) sd += dir_atom[i];
        sc += corr_atom[i];
    }
    double tsd = block_sum<1024>(sd, red);
    double tsc = block_sum<1024>(sc, red);
    if (threadIdx.x == 0) {
        double q2 = acc[kAccQ2], q1 = acc[kAccQ1];
        double eself = -beta / sqrt(M_PI) * q2 + nf * q1 * q1;
        double erec = acc[SPONGE_PART_RECIPROCAL];
        acc[SPONGE_PART_SELF] = eself;
        acc[SPONGE_PART_CORRECTION] = tsc;
        acc[SPONGE_PART_DIRECT] = tsd;
        acc[kAccQ2] = 0.0;
        acc[kAccQ1] = 0.0;
        if (need_energy && d_potential) d_potential[0] += (float)(tsc + eself + erec);
        if (need_virial && d_virial) d_virial[0] += (float)(tsd + tsc + eself + erec);
    }
}

// Batched energy-step bookkeeping: per system the finalize sums above, then
// the parts array of the fused step.
__global__ __launch_bounds__(1024) void pme_finalize_batch(const BatchSys* __restrict__ bs,
                                                           const int2* __restrict__ map) {
    __shared__ double red[16];
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    double sc = 0.0;
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) sc += P.corr_atom[i];
    const double tsc = block_sum<1024>(sc, red);
    if (threadIdx.x == 0) {
        double* acc = P.acc;
        const double* la = b.lj.ps.acc;
        const double q2 = acc[kAccQ2], q1 = acc[kAccQ1];
        const double eself = -(double)P.beta / sqrt(M_PI) * q2 + P.neutralizing_factor * q1 * q1;
        acc[SPONGE_PART_SELF] = eself;
        acc[SPONGE_PART_CORRECTION] = tsc;
        acc[kAccQ2] = 0.0;
        acc[kAccQ1] = 0.0;
        double* parts = b.io.parts;
        if (parts) {
            parts[SPONGE_PART_LJ] = la[SPONGE_PART_LJ];
            parts[SPONGE_PART_DIRECT] = la[SPONGE_PART_DIRECT];
            parts[SPONGE_PART_LJ_VIRIAL] = la[SPONGE_PART_LJ_VIRIAL];
            parts[SPONGE_PART_CORRECTION] = tsc;
            parts[SPONGE_PART_SELF] = eself;
            parts[SPONGE_PART_RECIPROCAL] = acc[SPONGE_PART_RECIPROCAL];
            parts[SPONGE_PART_LRC] = b.lrc;
            parts[SPONGE_PART_PAIRS_IN] = la[SPONGE_PART_PAIRS_IN];
        }
    }
}

int plan_axis(int K, FftPlan* p) {
    p->K = K;
    p->npass = 0;
    p->code = 0;
    int m = K;
    auto push = [&](int r) {
        p->code |= (unsigned long long)r << (4 * p->npass);
        p->npass++;
        m /= r;
    };
    while (m % 8 == 0 && p->npass < 16) push(8);
    while (m % 4 == 0 && p->npass < 16) push(4);
    for (int r : {2, 3, 5, 7})
        while (m % r == 0 && p->npass < 16) push(r);
    if (m != 1) return -1;
    // every pass's butterflies (K / R of them) must fit the two lanes per
    // element of the long-line path
    for (int ps = 0; ps < p->npass; ps++)
        if (K / (int)((p->code >> (4 * ps)) & 15ull) > 128) return -1;
    return 0;
}

// Dynamic LDS of the three transform kernels (bytes).
size_t spread_z_lds(int Kz) {
    return (size_t)(kRowBlock * Kz + (kRowBlock / 2) * line_pitch(Kz) + Kz) * sizeof(float2);
}
size_t inv_z_lds(int Kz) {
    return (size_t)(kRowBlock * (Kz / 2 + 1) + (kRowBlock / 2) * line_pitch(Kz) + Kz) * sizeof(float2);
}
size_t conv2d_lds(int Kx, int Ky) { return (size_t)(padded_len(Kx) * line_pitch(Ky) + Kx + Ky) * sizeof(float2); }
size_t lines_lds(int K) { return (size_t)(kLines * line_pitch(K) + K) * sizeof(float2); }
size_t spread_yz_lds(int Ky, int Kz) {
    return (size_t)(yz_acc(Ky, Kz) + yz_lines(Ky, Kz) + Kz + Ky) * sizeof(float2);
}
size_t inv_yz_lds(int Ky, int Kz) { return (size_t)(yz_tile(Ky, Kz) + yz_lines(Ky, Kz) + Kz + Ky) * sizeof(float2); }
// LDS of a convolution form's launches (2: the slab form's spread and
// inverse, which replace spread_z and inv_z, and conv_x)
size_t form_lds(int form, int Kx, int Ky, int Kz) {
    if (form == 2) return std::max({spread_yz_lds(Ky, Kz), inv_yz_lds(Ky, Kz), lines_lds(Kx)});
    return form == 1 ? std::max(lines_lds(Kx), lines_lds(Ky)) : conv2d_lds(Kx, Ky);
}

// Lets every transform kernel use up to `bytes` of dynamic LDS (once per
// size; the attribute is per function).
int raise_lds_limits(size_t bytes) {
    static size_t done = 64 * 1024;
    if (bytes <= done) return SPONGE_NB_OK;
    const void* fns[] = {(const void*)spread_z_kernel<kRadixAll>, (const void*)spread_z_kernel<kRadix8>,
                         (const void*)spread_z_kernel<kRadix34>,  (const void*)spread_z_kernel<kRadix248>,
                         (const void*)spread_z_adjoint_kernel,    (const void*)spread_z_batch,
                         (const void*)conv2d_kernel<kRadixAll>,   (const void*)conv2d_kernel<kRadix8>,
                         (const void*)conv2d_kernel<kRadix34>,    (const void*)conv2d_kernel<kRadix248>,
                         (const void*)conv2d_wide_kernel<kRadixAll>, (const void*)conv2d_wide_kernel<kRadix8>,
                         (const void*)conv2d_wide_kernel<kRadix34>,  (const void*)conv2d_wide_kernel<kRadix248>,
                         (const void*)conv2d_batch,               (const void*)inv_z_kernel<kRadixAll>,
                         (const void*)inv_z_kernel<kRadix8>,      (const void*)inv_z_kernel<kRadix34>,
                         (const void*)inv_z_kernel<kRadix248>,    (const void*)inv_z_batch,
                         (const void*)conv_x_kernel<kRadixAll>,   (const void*)conv_x_kernel<kRadix8>,
                         (const void*)conv_x_kernel<kRadix34>,    (const void*)conv_x_kernel<kRadix248>,
                         (const void*)conv_y_forward_kernel<kRadixAll>, (const void*)conv_y_forward_kernel<kRadix8>,
                         (const void*)conv_y_forward_kernel<kRadix34>,  (const void*)conv_y_forward_kernel<kRadix248>,
                         (const void*)conv_y_inverse_kernel<kRadixAll>, (const void*)conv_y_inverse_kernel<kRadix8>,
                         (const void*)conv_y_inverse_kernel<kRadix34>,  (const void*)conv_y_inverse_kernel<kRadix248>,
                         (const void*)spread_yz_kernel<kRadixAll>,      (const void*)spread_yz_kernel<kRadix8>,
                         (const void*)spread_yz_kernel<kRadix34>,       (const void*)spread_yz_kernel<kRadix248>,
                         (const void*)inv_yz_kernel<kRadixAll>,         (const void*)inv_yz_kernel<kRadix8>,
                         (const void*)inv_yz_kernel<kRadix34>,          (const void*)inv_yz_kernel<kRadix248>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
            set_error("cannot raise the LDS limit of the PME kernels to %zu B", bytes);
            return SPONGE_NB_EHIP;
        }
    done = bytes;
    return SPONGE_NB_OK;
}

}  // namespace

// Get_Beta (PME_force.cu:71-100), FP64 bisection.
static double pme_get_beta(double rc, double tol) {
    double high = 1.0;
    int ihigh = 1;
    while (std::erfc(high * rc) / rc > tol) {
        high *= 2;
        ihigh++;
    }
    ihigh += 50;
    double low = 0.0, beta = 0.0;
    for (int ilow = 1; ilow < ihigh; ilow++) {
        beta = (low + high) / 2;
        if (std::erfc(beta * rc) / rc >= tol) low = beta;
        else high = beta;
    }
    return beta;
}

// Check_2357_Factor / Get_Fft_Patameter (common.cu:467-534).
static bool smooth2357(int n) {
    if (n <= 0) return false;
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}
static int pme_fft_size(float length) {
    int t = ((int)std::ceil(length + 3)) >> 2 << 2;
    const int snaps[5][3] = {{60, 68, 64}, {120, 136, 128}, {240, 272, 256}, {480, 544, 512}, {960, 1088, 1024}};
    for (auto& s : snaps)
        if (t >= s[0] && t <= s[1]) {
            t = s[2];
            break;
        }
    while (!smooth2357(t)) t += 4;
    return t;
}

// |b(k)|^2 Euler spline moduli, order 4 (getb, PME_force.cu:44-67).
static double bspline_mod2(int k, int K) {
    const double M4[3] = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};  // M4(1), M4(2), M4(3)
    std::complex<double> den(0, 0);
    for (int kk = 0; kk < 3; kk++) den += M4[kk] * std::exp(std::complex<double>(0, 2 * M_PI * k * kk / K));
    return 1.0 / std::norm(den);
}

static void pme_fill_bc(sponge_pme* pme, std::vector<float>& h) {
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    const double Lx = pme->box[0], Ly = pme->box[1], Lz = pme->box[2];
    const double V = Lx * Ly * Lz, b = pme->beta;
    std::vector<double> B1(Kx), B2(Ky), B3(Kz);
    for (int k = 0; k < Kx; k++) B1[k] = bspline_mod2(k, Kx);
    for (int k = 0; k < Ky; k++) B2[k] = bspline_mod2(k, Ky);
    for (int k = 0; k < Kz; k++) B3[k] = bspline_mod2(k, Kz);
    h.assign((size_t)Kx * Ky * H, 0.f);
    for (int x = 0; x < Kx; x++) {
        int xr = x > Kx / 2 ? Kx - x : x;
        for (int y = 0; y < Ky; y++) {
            int yr = y > Ky / 2 ? Ky - y : y;
            for (int z = 0; z < H; z++) {
                double msq = xr * xr / (Lx * Lx) + yr * yr / (Ly * Ly) + (double)z * z / (Lz * Lz);
                size_t idx = ((size_t)z * Kx + x) * Ky + y;  // spectrum layout [kz][x][y]
                if (x == 0 && y == 0 && z == 0) h[idx] = 0.f;
                else h[idx] = (float)(std::exp(-M_PI * M_PI * msq / (b * b)) / (M_PI * msq * V) * B1[x] * B2[y] * B3[z]);
            }
        }
    }
    pme->neutralizing_factor = -M_PI / (2.0 * b * b * V);
}

extern "C" int sponge_pme_create(int atom_numbers, const float box_length[3], float cutoff, float direct_tolerance,
                                 int fftx, int ffty, int fftz, sponge_pme** out) {
    SNB_ARG(out && box_length, "null pointer");
    SNB_ARG(atom_numbers > 0 && cutoff > 0, "atom_numbers / cutoff");
    for (int d = 0; d < 3; d++) SNB_ARG(box_length[d] > 0, "box_length");
    *out = nullptr;
    sponge_pme* pme = new sponge_pme();
    memset(pme, 0, sizeof(*pme));
    pme->n = atom_numbers;
    for (int d = 0; d < 3; d++) pme->box[d] = box_length[d];
    pme->cutoff = cutoff;
    pme->tolerance = direct_tolerance > 0 ? direct_tolerance : 1e-5f;
    pme->beta = (float)pme_get_beta(cutoff, pme->tolerance);
    int req[3] = {fftx, ffty, fftz};
    for (int d = 0; d < 3; d++) pme->K[d] = req[d] > 0 ? req[d] : pme_fft_size(box_length[d]);
    FftPlan p;
    for (int d = 0; d < 3; d++) {
        if (pme->K[d] < 4 || plan_axis(pme->K[d], &p) != 0) {
            set_error("PME grid size %d unsupported: it must be >= 4 with prime factors 2, 3, 5, 7 only, and "
                      "K / r <= 128 for each of its radix-r passes (8 and 4 first)",
                      pme->K[d]);
            delete pme;
            return SPONGE_NB_EARG;
        }
    }
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        max_lds = 64 * 1024;
    // planes beyond 64 x 64 (or beyond the LDS) run the split convolution
    // unless the tuned table says otherwise; SPONGE_PME_SPLIT_CONV=0/1 forces
    // either form where it fits
    const TuneRow& tune = tune_row(atom_numbers);
    const bool plane_fits = conv2d_lds(Kx, Ky) <= (size_t)max_lds;
    pme->conv_split = (Kx * Ky > kConvPlaneMax || !plane_fits) ? 1 : 0;
    if (tune.conv_split == 0 && plane_fits) pme->conv_split = 0;
    if (tune.conv_split == 1) pme->conv_split = 1;
    if (tune.conv_split == 2 && form_lds(2, Kx, Ky, Kz) <= (size_t)max_lds) pme->conv_split = 2;
    pme->conv_threads = tune.conv_threads == 2 * kConvThreads ? 2 * kConvThreads : kConvThreads;
    pme->step_form = tune.step_form == 1 ? 1 : 0;
    if (const char* env = getenv("SPONGE_PME_SPLIT_CONV")) {
        if (env[0] == '1') pme->conv_split = 1;
        if (env[0] == '0' && conv2d_lds(Kx, Ky) <= (size_t)max_lds) pme->conv_split = 0;
    }
    const size_t conv_lds = form_lds(pme->conv_split, Kx, Ky, Kz);
    const size_t lds_need = std::max({spread_z_lds(Kz), inv_z_lds(Kz), conv_lds});
    if (lds_need > (size_t)max_lds || Kx > kMaxGridAxis || Ky > kMaxGridAxis || Kz > kMaxGridAxis) {
        set_error("PME grid %dx%dx%d too large for the in-LDS transforms (%zu B of LDS, %d available)", Kx, Ky, Kz,
                  lds_need, max_lds);
        delete pme;
        return SPONGE_NB_EARG;
    }
    {
        const int rc = raise_lds_limits(lds_need);
        if (rc) {
            delete pme;
            return rc;
        }
    }
    pme->nfft = Kx * Ky * H;
    std::vector<float> hbc;
    pme_fill_bc(pme, hbc);
    size_t NG = (size_t)Kx * Ky * Kz;
    hipError_t e = hipSuccess;
    auto A = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, bytes);
    };
    A((void**)&pme->d_bc, hbc.size() * sizeof(float));
    A((void**)&pme->FQ, (size_t)pme->nfft * sizeof(float2));
    A((void**)&pme->R, NG * sizeof(float));
    A((void**)&pme->base, (size_t)atom_numbers * sizeof(unsigned));
    A((void**)&pme->frac, (size_t)atom_numbers * sizeof(float4));
    A((void**)&pme->corr_atom, (size_t)atom_numbers * sizeof(float));
    pme->bin_cap = atom_numbers;  // a plane can never hold more than every atom
    A((void**)&pme->bin_count, (size_t)Kx * sizeof(int));
    A((void**)&pme->bin_atoms, (size_t)Kx * pme->bin_cap * sizeof(int));
    A((void**)&pme->bin_base, (size_t)Kx * pme->bin_cap * sizeof(unsigned));
    A((void**)&pme->bin_frac, (size_t)Kx * pme->bin_cap * sizeof(float4));
    A((void**)&pme->acc, SPONGE_NB_NPARTS * sizeof(double));
    for (int d = 0; d < 3; d++) {
        std::vector<float2> tw(pme->K[d]);
        for (int j = 0; j < pme->K[d]; j++) {
            double ang = 2.0 * M_PI * j / pme->K[d];
            tw[j] = make_float2((float)std::cos(ang), (float)-std::sin(ang));
        }
        A((void**)&pme->d_tw[d], tw.size() * sizeof(float2));
        if (e == hipSuccess) e = hipMemcpy(pme->d_tw[d], tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(pme->d_bc, hbc.data(), hbc.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(pme->corr_atom, 0, (size_t)atom_numbers * sizeof(float));
    if (e == hipSuccess) e = hipMemset(pme->bin_count, 0, (size_t)Kx * sizeof(int));
    if (e == hipSuccess) e = hipMemset(pme->acc, 0, SPONGE_NB_NPARTS * sizeof(double));
    // an atom that an overflowing neighbor list leaves without a slot keeps
    // these (the fused step's prep writes slotted atoms only): grid point 0
    // rather than an undefined index for the gather
    if (e == hipSuccess) e = hipMemset(pme->base, 0, (size_t)atom_numbers * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(pme->frac, 0, (size_t)atom_numbers * sizeof(float4));
    // default-priority side stream and fork / join events of the fused step
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&pme->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pme->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pme->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("sponge_pme_create: %s", hipGetErrorString(e));
        sponge_pme_destroy(pme);
        return SPONGE_NB_EHIP;
    }
    *out = pme;
    return SPONGE_NB_OK;
}

extern "C" void sponge_pme_destroy(sponge_pme* pme) {
    if (!pme) return;
    if (pme->side) {
        (void)hipStreamSynchronize(pme->side);
        (void)hipStreamDestroy(pme->side);
    }
    if (pme->fork) (void)hipEventDestroy(pme->fork);
    if (pme->join) (void)hipEventDestroy(pme->join);
    void* ptrs[] = {pme->d_bc, pme->FQ,       pme->R,        pme->base,
                    pme->frac, pme->corr_atom, pme->acc,      pme->d_tw[0],  pme->d_tw[1],
                    pme->d_tw[2]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (pme->bin_count) (void)hipFree(pme->bin_count);
    if (pme->bin_atoms) (void)hipFree(pme->bin_atoms);
    if (pme->bin_base) (void)hipFree(pme->bin_base);
    if (pme->bin_frac) (void)hipFree(pme->bin_frac);
    if (pme->zero_type) (void)hipFree(pme->zero_type);
    if (pme->zero_ab) (void)hipFree(pme->zero_ab);
    snb_fep_free(pme->fep);
    snb_scratch_free(&pme->ps);
    delete pme;
}

extern "C" int sponge_pme_info(const sponge_pme* pme, float* beta, int fft[3]) {
    SNB_ARG(pme, "pme_info");
    if (beta) *beta = pme->beta;
    if (fft)
        for (int d = 0; d < 3; d++) fft[d] = pme->K[d];
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_update_box_length(sponge_pme* pme, const float box_length[3]) {
    SNB_ARG(pme && box_length, "pme_update_box_length");
    for (int d = 0; d < 3; d++) {
        SNB_ARG(box_length[d] > 0, "box_length");
        pme->box[d] = box_length[d];
    }
    std::vector<float> hbc;
    pme_fill_bc(pme, hbc);
    SNB_HIP(hipMemcpy(pme->d_bc, hbc.data(), hbc.size() * sizeof(float), hipMemcpyHostToDevice));
    return SPONGE_NB_OK;
}

// Radix set of a plan (bit r per radix-r pass) and the smallest
// instantiated set holding it.
static int radix_mask(const FftPlan& p) {
    int m = 0;
    for (int k = 0; k < p.npass; k++) m |= 1 << (int)((p.code >> (4 * k)) & 15ull);
    return m;
}
static int kernel_set(int m) {
    for (int c : {kRadix8, kRadix34, kRadix248})
        if ((m & ~c) == 0) return c;
    return kRadixAll;
}

#define SNB_RADIX_DISPATCH(SET, KERNEL, GRID, BLOCK, LDS, STREAM, ...)                                  \
    do {                                                                                                \
        switch (SET) {                                                                                  \
            case kRadix8: KERNEL<kRadix8><<<GRID, BLOCK, LDS, STREAM>>>(__VA_ARGS__); break;             \
            case kRadix34: KERNEL<kRadix34><<<GRID, BLOCK, LDS, STREAM>>>(__VA_ARGS__); break;           \
            case kRadix248: KERNEL<kRadix248><<<GRID, BLOCK, LDS, STREAM>>>(__VA_ARGS__); break;         \
            default: KERNEL<kRadixAll><<<GRID, BLOCK, LDS, STREAM>>>(__VA_ARGS__); break;                \
        }                                                                                               \
    } while (0)

// The same dispatch through hipExtLaunchKernelGGL, whose launch records the
// two events at the kernel's own start and end (measurement hook).
#define SNB_RADIX_DISPATCH_EV(SET, KERNEL, GRID, BLOCK, LDS, STREAM, E0, E1, ...)                          \
    do {                                                                                                   \
        switch (SET) {                                                                                     \
            case kRadix8:                                                                                  \
                hipExtLaunchKernelGGL(KERNEL<kRadix8>, dim3(GRID), dim3(BLOCK), (std::uint32_t)(LDS), STREAM, \
                                      E0, E1, 0u, __VA_ARGS__);                                            \
                break;                                                                                     \
            case kRadix34:                                                                                 \
                hipExtLaunchKernelGGL(KERNEL<kRadix34>, dim3(GRID), dim3(BLOCK), (std::uint32_t)(LDS),      \
                                      STREAM, E0, E1, 0u, __VA_ARGS__);                                    \
                break;                                                                                     \
            case kRadix248:                                                                                \
                hipExtLaunchKernelGGL(KERNEL<kRadix248>, dim3(GRID), dim3(BLOCK), (std::uint32_t)(LDS),     \
                                      STREAM, E0, E1, 0u, __VA_ARGS__);                                    \
                break;                                                                                     \
            default:                                                                                       \
                hipExtLaunchKernelGGL(KERNEL<kRadixAll>, dim3(GRID), dim3(BLOCK), (std::uint32_t)(LDS),     \
                                      STREAM, E0, E1, 0u, __VA_ARGS__);                                    \
                break;                                                                                     \
        }                                                                                                  \
    } while (0)

// Stage 1 of the pipeline: spread + z transforms of the binned atoms.
int snb_pme_spread(sponge_pme* pme, int energy, hipStream_t s) {
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2];
    FftPlan pz;
    plan_axis(Kz, &pz);
    const int nyb = row_blocks(Ky);
    double* acc = energy ? pme->acc : nullptr;
    if (pme->conv_split == 2) {
        FftPlan py;
        plan_axis(Ky, &py);
        const int set = kernel_set(radix_mask(pz) | radix_mask(py));
        if (pme->ev[0])
            SNB_RADIX_DISPATCH_EV(set, spread_yz_kernel, Kx, kYZThreads, spread_yz_lds(Ky, Kz), s, pme->ev[0],
                                  pme->ev[1], pme->bin_base, pme->bin_frac, pme->bin_count, pme->bin_cap, pme->FQ, Kx,
                                  Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc);
        else
            SNB_RADIX_DISPATCH(set, spread_yz_kernel, Kx, kYZThreads, spread_yz_lds(Ky, Kz), s, pme->bin_base,
                               pme->bin_frac, pme->bin_count, pme->bin_cap, pme->FQ, Kx, Ky, Kz, pz, py,
                               pme->d_tw[2], pme->d_tw[1], acc);
        SNB_CHECK_LAUNCH();
        return SPONGE_NB_OK;
    }
    if (pme->ev[0]) {
        SNB_RADIX_DISPATCH_EV(kernel_set(radix_mask(pz)), spread_z_kernel, Kx * nyb, kZThreads, spread_z_lds(Kz), s,
                              pme->ev[0], pme->ev[1], pme->bin_base, pme->bin_frac, pme->bin_count, pme->bin_atoms,
                              pme->bin_cap, pme->FQ, Kx, Ky, Kz, nyb, pz, pme->d_tw[2], acc);
    } else {
        SNB_RADIX_DISPATCH(kernel_set(radix_mask(pz)), spread_z_kernel, Kx * nyb, kZThreads, spread_z_lds(Kz), s,
                           pme->bin_base, pme->bin_frac, pme->bin_count, pme->bin_atoms, pme->bin_cap, pme->FQ, Kx, Ky,
                           Kz,
                           nyb, pz, pme->d_tw[2], acc);
    }
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

// The convolution of a spectrum that holds the z transforms only (forms 0
// and 1: the plane or the three line launches, then inv_z) or the z and y
// transforms (form 2: conv_x, then inv_yz), each kernel in the smallest radix
// set of its axes.
static int convolve_launch(sponge_pme* pme, int energy, hipStream_t s, int form) {
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    FftPlan px, py, pz;
    plan_axis(Kx, &px);
    plan_axis(Ky, &py);
    plan_axis(Kz, &pz);
    const int nyb = row_blocks(Ky);
    if (form == 2) {
        const dim3 gx(ceil_div(Ky, kLines), H);
        SNB_RADIX_DISPATCH(kernel_set(radix_mask(px)), conv_x_kernel, gx, kLineThreads, lines_lds(Kx), s, pme->FQ,
                           pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc, pme->bin_count);
        SNB_CHECK_LAUNCH();
        SNB_RADIX_DISPATCH(kernel_set(radix_mask(pz) | radix_mask(py)), inv_yz_kernel, Kx, kYZThreads,
                           inv_yz_lds(Ky, Kz), s, pme->FQ, pme->R, Kx, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]);
        SNB_CHECK_LAUNCH();
        return SPONGE_NB_OK;
    }
    if (form == 1) {
        const dim3 gy(ceil_div(Kx, kLines), H), gx(ceil_div(Ky, kLines), H);
        const int sy = kernel_set(radix_mask(py)), sx = kernel_set(radix_mask(px));
        SNB_RADIX_DISPATCH(sy, conv_y_forward_kernel, gy, kLineThreads, lines_lds(Ky), s, pme->FQ, Kx, Ky, py,
                           pme->d_tw[1]);
        SNB_CHECK_LAUNCH();
        SNB_RADIX_DISPATCH(sx, conv_x_kernel, gx, kLineThreads, lines_lds(Kx), s, pme->FQ, pme->d_bc, Kx, Ky, Kz, px,
                           pme->d_tw[0], energy, pme->acc, pme->bin_count);
        SNB_CHECK_LAUNCH();
        SNB_RADIX_DISPATCH(sy, conv_y_inverse_kernel, gy, kLineThreads, lines_lds(Ky), s, pme->FQ, Kx, Ky, py,
                           pme->d_tw[1]);
        SNB_CHECK_LAUNCH();
    } else if (pme->conv_threads == 2 * kConvThreads) {
        SNB_RADIX_DISPATCH(kernel_set(radix_mask(px) | radix_mask(py)), conv2d_wide_kernel, H, 2 * kConvThreads,
                           conv2d_lds(Kx, Ky), s, pme->FQ, pme->d_bc, Kx, Ky, Kz, px, py, pme->d_tw[0],
                           pme->d_tw[1], energy, pme->acc, pme->bin_count);
        SNB_CHECK_LAUNCH();
    } else {
        SNB_RADIX_DISPATCH(kernel_set(radix_mask(px) | radix_mask(py)), conv2d_kernel, H, kConvThreads,
                           conv2d_lds(Kx, Ky), s, pme->FQ, pme->d_bc, Kx, Ky, Kz, px, py, pme->d_tw[0],
                           pme->d_tw[1], energy, pme->acc, pme->bin_count);
        SNB_CHECK_LAUNCH();
    }
    SNB_RADIX_DISPATCH(kernel_set(radix_mask(pz)), inv_z_kernel, Kx * nyb, kZThreads, inv_z_lds(Kz), s, pme->FQ,
                       pme->R, Kx, Ky, Kz, nyb, pz, pme->d_tw[2]);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_pme_convolve(sponge_pme* pme, int energy, hipStream_t s) {
    return convolve_launch(pme, energy, s, pme->conv_split);
}

// The reciprocal pipeline; forces added to d_frc (original atom order).
int snb_pme_reciprocal(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge, float* d_frc,
                       int energy, int prep_done, hipStream_t s) {
    if (!prep_done) {
        pme_prep_kernel<<<ceil_div(pme->n, 256), 256, 0, s>>>(pme->n, d_uint_crd, d_charge, pme_prep_args(pme),
                                                               energy);
        SNB_CHECK_LAUNCH();
    }
    int rc = snb_pme_spread(pme, energy, s);
    if (rc) return rc;
    rc = snb_pme_convolve(pme, energy, s);
    if (rc) return rc;
    return d_frc ? snb_pme_gather(pme, d_frc, s) : SPONGE_NB_OK;
}

// Reciprocal forces from the convolved grid R, added to d_frc.
int snb_pme_gather(sponge_pme* pme, float* d_frc, hipStream_t s, int* step_counter) {
    const int n = pme->n, Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2];
    const int grid = ceil_div((long)n * 16, kGatherThreads);
    if (pme->ev[2])
        hipExtLaunchKernelGGL(pme_gather_kernel, dim3(grid), dim3(kGatherThreads), 0u, s, pme->ev[2], pme->ev[3], 0u,
                              n, pme->base, pme->frac, pme->R, Kx, Ky, Kz, (float)Kx / pme->box[0],
                              (float)Ky / pme->box[1], (float)Kz / pme->box[2], d_frc, step_counter);
    else
        pme_gather_kernel<<<grid, kGatherThreads, 0, s>>>(n, pme->base, pme->frac, pme->R, Kx, Ky, Kz,
                                                           (float)Kx / pme->box[0], (float)Ky / pme->box[1],
                                                           (float)Kz / pme->box[2], d_frc, step_counter);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_set_tuning(sponge_pme* pme, int step_form, int conv_split) {
    SNB_ARG(pme, "pme_set_tuning");
    SNB_ARG(step_form >= -1 && step_form <= 1, "step_form must be -1, 0 or 1");
    if (step_form >= 0) pme->step_form = step_form;
    if (conv_split >= 0 && conv_split <= 2) {
        int dev = 0, max_lds = 64 * 1024;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            max_lds = 64 * 1024;
        const size_t need = form_lds(conv_split, pme->K[0], pme->K[1], pme->K[2]);
        SNB_ARG(need <= (size_t)max_lds, "this grid's planes or slabs do not fit one workgroup's LDS");
        const int rc = raise_lds_limits(need);
        if (rc) return rc;
        pme->conv_split = conv_split;
    }
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_set_events(sponge_pme* pme, void* spread_start, void* spread_stop, void* gather_start,
                                     void* gather_stop) {
    SNB_ARG(pme, "pme_set_events");
    SNB_ARG((spread_start == nullptr) == (spread_stop == nullptr) &&
                (gather_start == nullptr) == (gather_stop == nullptr),
            "both events of a kernel or neither");
    pme->ev[0] = reinterpret_cast<hipEvent_t>(spread_start);
    pme->ev[1] = reinterpret_cast<hipEvent_t>(spread_stop);
    pme->ev[2] = reinterpret_cast<hipEvent_t>(gather_start);
    pme->ev[3] = reinterpret_cast<hipEvent_t>(gather_stop);
    return SPONGE_NB_OK;
}

int snb_pme_finalize(sponge_pme* pme, const float* d_direct_atom_energy, int need_energy, int need_virial,
                     float* d_potential, float* d_virial, hipStream_t s) {
    pme_finalize_kernel<<<1, 1024, 0, s>>>(pme->n, need_virial ? d_direct_atom_energy : nullptr, pme->corr_atom,
                                           (double)pme->beta, pme->neutralizing_factor, need_energy, need_virial,
                                           d_potential, d_virial, pme->acc);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_reciprocal_force(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge,
                                           float* d_frc, int need_virial, int need_energy, float* d_virial,
                                           float* d_potential, const float* d_direct_atom_energy, void* stream) {
    SNB_ARG(pme && d_uint_crd && d_charge && d_frc, "pme_reciprocal_force");
    hipStream_t s = as_stream(stream);
    int energy = (need_virial > 0 || need_energy > 0) ? 1 : 0;
    int rc = snb_pme_reciprocal(pme, d_uint_crd, d_charge, d_frc, energy, 0, s);
    if (rc || !energy) return rc;
    return snb_pme_finalize(pme, d_direct_atom_energy, need_energy > 0, need_virial > 0, d_potential, d_virial, s);
}

// Host-side read of the PME FP64 parts (after a synchronising copy).
int snb_pme_parts(sponge_pme* pme, double* parts, hipStream_t s) {
    SNB_HIP(hipMemcpyAsync(parts, pme->acc, SPONGE_NB_NPARTS * sizeof(double), hipMemcpyDeviceToHost, s));
    SNB_HIP(hipStreamSynchronize(s));
    return SPONGE_NB_OK;
}

// ---------------------------------------------------------------------------
// Batched reciprocal pipeline: one launch per stage over every system of a
// batch (block maps built by sponge_batch_create).
void snb_pme_plans(const sponge_pme* pme, FftPlan* px, FftPlan* py, FftPlan* pz) {
    plan_axis(pme->K[0], px);
    plan_axis(pme->K[1], py);
    plan_axis(pme->K[2], pz);
}

int snb_pme_row_blocks(const sponge_pme* pme) { return row_blocks(pme->K[1]); }

size_t snb_pme_lds_fwd(const sponge_pme* pme) { return spread_z_lds(pme->K[2]); }

size_t snb_pme_lds_conv(const sponge_pme* pme) { return conv2d_lds(pme->K[0], pme->K[1]); }

// The batched kernels hold a whole plane per workgroup also where the
// system's own pipeline runs the split convolution.
int snb_pme_raise_lds(size_t bytes) {
    int dev = 0, max_lds = 0;
    SNB_HIP(hipGetDevice(&dev));
    SNB_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (bytes > (size_t)max_lds) {
        set_error("batched PME transforms need %zu B of LDS per workgroup, %d available", bytes, max_lds);
        return SPONGE_NB_EARG;
    }
    return raise_lds_limits(bytes);
}

size_t snb_pme_lds_inv(const sponge_pme* pme) { return inv_z_lds(pme->K[2]); }

int snb_batch_reciprocal(sponge_batch* b, int energy, hipStream_t s) {
    spread_z_batch<<<b->nblk[kStFwd], kZThreads, b->lds_fwd, s>>>(b->d, b->map[kStFwd], energy);
    SNB_CHECK_LAUNCH();
    conv2d_batch<<<b->nblk[kStConv], kConvThreads, b->lds_conv, s>>>(b->d, b->map[kStConv], energy);
    SNB_CHECK_LAUNCH();
    inv_z_batch<<<b->nblk[kStInv], kZThreads, b->lds_inv, s>>>(b->d, b->map[kStInv]);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_batch_gather(sponge_batch* b, hipStream_t s) {
    pme_gather_batch<<<b->nblk[kStGather], kGatherThreads, 0, s>>>(b->d, b->map[kStGather]);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_batch_finalize(sponge_batch* b, hipStream_t s) {
    pme_finalize_batch<<<b->nblk[kStFin], 1024, 0, s>>>(b->d, b->map[kStFin]);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

// ---------------------------------------------------------------------------
// Pieces of the adjoint (gradient) path.
int snb_pme_prep(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge, hipStream_t s) {
    pme_prep_kernel<<<ceil_div(pme->n, 256), 256, 0, s>>>(pme->n, d_uint_crd, d_charge, pme_prep_args(pme), 0);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

// R = theta * D for the adjoint source D[g] = sum_i q_i wv_i . grad M_i(g);
// the atom bins must have been filled (snb_pme_prep) since the last pipeline.
int snb_pme_adjoint_grid(sponge_pme* pme, const float4* wv, const float* scale, hipStream_t s) {
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    FftPlan px, py, pz;
    snb_pme_plans(pme, &px, &py, &pz);
    const int nyb = row_blocks(Ky);
    spread_z_adjoint_kernel<<<Kx * nyb, kZThreads, spread_z_lds(Kz), s>>>(
        pme->bin_base, pme->bin_frac, pme->bin_count, pme->bin_atoms, pme->bin_cap, pme->FQ, Kx, Ky, Kz, nyb, pz,
        pme->d_tw[2], wv, scale);
    SNB_CHECK_LAUNCH();
    (void)H;
    (void)px;
    (void)py;
    // the adjoint source was spread with z transforms only: the line form
    // stands in for the slab form
    return convolve_launch(pme, 0, s, pme->conv_split == 2 ? 1 : pme->conv_split);
}

int snb_pme_adj_gather(sponge_pme* pme, int mode, const float4* wv, float sign, float* out, hipStream_t s) {
    const int n = pme->n;
    const int grid = ceil_div((long)n * 16, kGatherThreads);
    if (mode == 0)
        pme_adj_gather_kernel<0><<<grid, kGatherThreads, 0, s>>>(n, pme->base, pme->frac, pme->R, pme->K[0],
                                                                 pme->K[1], pme->K[2], wv, sign, out);
    else
        pme_adj_gather_kernel<1><<<grid, kGatherThreads, 0, s>>>(n, pme->base, pme->frac, pme->R, pme->K[0],
                                                                 pme->K[1], pme->K[2], wv, sign, out);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}
