// Smooth PME reciprocal part for gfx950 (Particle_Mesh_Ewald,
// PME_force/PME_force.cu), with a hand-written 3D real-to-complex FFT in
// place of cuFFT.
//
// Pipeline per call (grid layout [x][y][z], z fastest, half spectrum
// [kx][ky][kz <= Kz/2] exactly as the reference's PME_Q / PME_FQ / PME_BC):
//   pme_prep      per atom: grid base index and spline fraction from the uint
//                 coordinate with one 32x32->64-bit multiply (t = u K / 2^32)
//   fwd_plane     one block per x plane: every atom whose order-4 stencil
//                 touches the plane is spread into the plane in LDS
//                 (ds_add_f32, no global atomics and no charge grid in HBM),
//                 then real z-rows are packed two at a time into complex
//                 z-FFTs and unpacked (R2C), then y-FFTs
//   xconv         one block per (y, kz chunk): x-FFT, multiply by BC, energy
//                 1/2 sum BC |F|^2 (FP64), inverse x-FFT
//   inv_plane     one block per x plane: inverse y-FFTs, Hermitian repack and
//                 paired complex z-IFFT (C2R), write the convolved grid
//   pme_gather    16 lanes per atom: force -q sum grid * dtheta (K/L)
// All FFT passes are mixed-radix (2,3,4,5,7) Stockham steps over a whole
// plane held in LDS, each step staged through registers so one LDS buffer
// suffices.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "pme_prep.hpp"

using namespace snb;

namespace {

constexpr int kPlaneThreads = 512;
constexpr int kSpreadPerThread = 4;
// Fixed-point unit of the spread grid: one contribution q*Mx*My*Mz must stay
// below 2^31 units, i.e. |q| < 128 / (2/3)^3 = 432 (18.2223 x 23.7 e).
constexpr float kSpreadScale = 16777216.0f;
constexpr int kConvThreads = 256;
constexpr int kConvChunk = 16;
// LDS pitch of one x position in the convolution kernel: odd, so that the
// x-FFT's element stride (the pitch) spreads a wave's accesses over the banks
constexpr int kConvPitch = kConvChunk + 1;
constexpr int kGatherThreads = 256;

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

// q = n / d for n * d < 2^32 via one multiply-high with m = ceil(2^32 / d);
// d == 1 (m would be 2^32) is passed through.
__device__ __forceinline__ unsigned div_magic(unsigned d) { return d > 1u ? 0xFFFFFFFFu / d + 1u : 0u; }
__device__ __forceinline__ unsigned fast_div(unsigned n, unsigned d, unsigned m) { return d > 1u ? __umulhi(n, m) : n; }

// One Stockham radix-R step over nlines lines of length K held in LDS;
// element (line l, position p) lives at buf[l*ls + p*ks]; tw is the LDS
// twiddle table exp(-2 pi i j / K). Line l is owned by wave l % nw in every
// pass (a wave works on G of its lines at a time), and a wave's LDS
// operations execute in program order, so a line's reads all precede its
// writes and the next pass sees them without any workgroup barrier.
template <int R, int SIGN>
__device__ void wave_pass(float2* buf, int nlines, int ls, int ks, int K, int Ns, const float2* tw, int wid,
                          int nw, int lane) {
    const int per = K / R;
    const int tws = K / (Ns * R);
    const unsigned mns = div_magic((unsigned)Ns);
    if (per <= 64) {
        const int G = 64 / per;
        const int sub = (int)fast_div((unsigned)lane, (unsigned)per, div_magic((unsigned)per));
        const int j = lane - sub * per;
        const int jn = (int)fast_div((unsigned)j, Ns, mns);
        const int k = j - jn * Ns;
        const int d = jn * Ns * R + k;
        for (int i0 = 0; wid + nw * i0 < nlines; i0 += G) {
            const int l = wid + nw * (i0 + sub);
            if (sub < G && l < nlines) {
                float2 v[R];
#pragma unroll
                for (int q = 0; q < R; q++) {
                    float2 x = buf[l * ls + (j + q * per) * ks];
                    if (q > 0) {
                        float2 t = tw[q * k * tws];
                        if (SIGN > 0) t.y = -t.y;
                        x = cmul(x, t);
                    }
                    v[q] = x;
                }
                dft<R, SIGN>(v);
#pragma unroll
                for (int q = 0; q < R; q++) buf[l * ls + (d + q * Ns) * ks] = v[q];
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
                    const int k = j - (int)fast_div((unsigned)j, Ns, mns) * Ns;
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        float2 x = buf[l * ls + (j + q * per) * ks];
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
                    const int jn = (int)fast_div((unsigned)j, Ns, mns);
                    const int d = jn * Ns * R + (j - jn * Ns);
#pragma unroll
                    for (int q = 0; q < R; q++) buf[l * ls + (d + q * Ns) * ks] = v[c][q];
                }
            }
        }
    }
}

// Full multi-pass transform of the wave's lines; callers put a workgroup
// barrier before (data written by other waves) and after (data read by other
// waves).
template <int SIGN, int NT>
__device__ void lds_fft(float2* buf, int nlines, int ls, int ks, const FftPlan& p, const float2* tw) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = NT / 64;
    int Ns = 1;
    for (int ps = 0; ps < p.npass; ps++) {
        int R = p.radix[ps];
        switch (R) {
            case 4: wave_pass<4, SIGN>(buf, nlines, ls, ks, p.K, Ns, tw, wid, nw, lane); break;
            case 2: wave_pass<2, SIGN>(buf, nlines, ls, ks, p.K, Ns, tw, wid, nw, lane); break;
            case 3: wave_pass<3, SIGN>(buf, nlines, ls, ks, p.K, Ns, tw, wid, nw, lane); break;
            case 5: wave_pass<5, SIGN>(buf, nlines, ls, ks, p.K, Ns, tw, wid, nw, lane); break;
            default: wave_pass<7, SIGN>(buf, nlines, ls, ks, p.K, Ns, tw, wid, nw, lane); break;
        }
        Ns *= R;
    }
}

template <int NT>
__device__ void load_twiddles(float2* dst, const float2* __restrict__ src, int K) {
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

// PME_Q_Spread (PME_force.cu:428-471) fused with the forward transform of
// one x plane: grid point (base - k) gets q M4(fx+kx) M4(fy+ky) M4(fz+kz);
// plane x receives the atoms filed under planes x .. x+3 (with kx = their
// offset). The real plane is held directly in the packed complex layout of the
// z-FFT: row y is the real (y even) or imaginary (y odd) part of line y/2.
//
// MODE 1 spreads, instead of the charge, the adjoint source of the
// force-matching gradient, q_i w_i . grad M_i(g): wv[i] = (w_x Kx/Lx,
// w_y Ky/Ly, w_z Kz/Lz), with the fixed-point unit read from scale_ptr (set
// on the device so that no contribution reaches 2^31).
template <int MAXB, int MODE = 0>
__device__ __forceinline__ void fwd_plane_body(
    int x, const unsigned* __restrict__ base, const float4* __restrict__ frac, const int* __restrict__ bin_count,
    const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz,
    const FftPlan& pz, const FftPlan& py, const float2* __restrict__ twz, const float2* __restrict__ twy,
    double* __restrict__ acc, const float4* __restrict__ wv = nullptr, const float* __restrict__ scale_ptr = nullptr) {
    extern __shared__ float2 sm[];
    const float unit = MODE ? scale_ptr[0] : kSpreadScale;
    const int H = Kz / 2 + 1, nrp = (Ky + 1) / 2;
    // packed z-lines sit at an odd pitch, so that the lines a wave transforms
    // together fall on different LDS banks
    const int zp = Kz + 1;
    const int plane = max(Ky * H, nrp * zp);
    float2* tz = sm + max(plane, Ky * Kz);  // past the fixed-point grid too
    float2* ty = tz + Kz;
    load_twiddles<kPlaneThreads>(tz, twz, Kz);
    load_twiddles<kPlaneThreads>(ty, twy, Ky);
    if (x == 0 && threadIdx.x == 0 && acc) acc[SPONGE_PART_RECIPROCAL] = 0.0;
    // Spread into a row-major Ky x Kz grid of 64-bit fixed-point sums
    // (kSpreadScale units) that aliases the front of the plane buffer. LDS
    // integer atomics are an order of magnitude faster than LDS float atomics
    // on this chip, and integer sums make the grid independent of the order
    // in which atoms arrive.
    unsigned long long* qacc = reinterpret_cast<unsigned long long*>(sm);
    const int npt = Ky * Kz;
    for (int idx = threadIdx.x; idx < npt; idx += kPlaneThreads) qacc[idx] = 0ull;
    __syncthreads();
    // The four bins are walked as one flat range; each thread issues the
    // index, base and fraction loads of up to kSpreadPerThread atoms before
    // any LDS update, so their latencies overlap.
    int c[4], off[5];
    off[0] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int b = x + k;
        b -= b >= Kx ? Kx : 0;
        c[k] = min(bin_count[b], bin_cap);
        off[k + 1] = off[k] + c[k];
    }
    const int total = off[4];
    for (int t0 = 0; t0 < total; t0 += kPlaneThreads * kSpreadPerThread) {
        int ai[kSpreadPerThread], akx[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            const int t = t0 + r * kPlaneThreads + (int)threadIdx.x;
            ai[r] = -1;
            akx[r] = 0;
            if (t < total) {
                const int k = (t >= off[1]) + (t >= off[2]) + (t >= off[3]);
                int b = x + k;
                b -= b >= Kx ? Kx : 0;
                ai[r] = bin_atoms[(size_t)b * bin_cap + (t - off[k])];
                akx[r] = k;
            }
        }
        unsigned bb[kSpreadPerThread];
        float4 ff[kSpreadPerThread], ww[kSpreadPerThread];
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            if (ai[r] >= 0) {
                bb[r] = base[ai[r]];
                ff[r] = frac[ai[r]];
                if (MODE) ww[r] = wv[ai[r]];
            }
        }
#pragma unroll
        for (int r = 0; r < kSpreadPerThread; r++) {
            if (ai[r] < 0) continue;
            const int by = (int)((bb[r] >> 10) & 1023u), bz = (int)(bb[r] >> 20);
            const float4 f = ff[r];
            float wz[4], dz[4];
            int gz[4];
#pragma unroll
            for (int kz = 0; kz < 4; kz++) {
                wz[kz] = bsp_w(f.z, kz);
                if (MODE) dz[kz] = bsp_dw(f.z, kz);
                int tt = bz - kz;
                gz[kz] = tt + (tt < 0 ? Kz : 0);
            }
            if (MODE == 0) {
                const float wx = f.w * unit * bsp_w(f.x, akx[r]);
#pragma unroll
                for (int ky = 0; ky < 4; ky++) {
                    int gy = by - ky;
                    gy += gy < 0 ? Ky : 0;
                    const float wy = wx * bsp_w(f.y, ky);
                    unsigned long long* row = qacc + gy * Kz;
#pragma unroll
                    for (int kz = 0; kz < 4; kz++)
                        atomicAdd(row + gz[kz], (unsigned long long)(long long)__float2int_rn(wy * wz[kz]));
                }
            } else {
                const float4 w = ww[r];
                const float qs = f.w * unit;
                const float ax = bsp_w(f.x, akx[r]), dax = bsp_dw(f.x, akx[r]);
#pragma unroll
                for (int ky = 0; ky < 4; ky++) {
                    int gy = by - ky;
                    gy += gy < 0 ? Ky : 0;
                    const float ay = bsp_w(f.y, ky), day = bsp_dw(f.y, ky);
                    const float c1 = qs * (w.x * dax * ay + w.y * ax * day);
                    const float c2 = qs * w.z * ax * ay;
                    unsigned long long* row = qacc + gy * Kz;
#pragma unroll
                    for (int kz = 0; kz < 4; kz++)
                        atomicAdd(row + gz[kz],
                                  (unsigned long long)(long long)__float2int_rn(fmaf(c1, wz[kz], c2 * dz[kz])));
                }
            }
        }
    }
    __syncthreads();
    // Fixed point -> float, into the packed complex layout of the z-FFT: row y
    // is the real (y even) or imaginary (y odd) part of line y/2; with Ky odd
    // the imaginary part of the last line is zero.
    float qv[2 * MAXB];
#pragma unroll
    for (int s = 0; s < 2 * MAXB; s++) {
        const int o = threadIdx.x + s * kPlaneThreads;
        if (o < npt) qv[s] = (float)((double)(long long)qacc[o] / (double)unit);
    }
    __syncthreads();
    float* smf = reinterpret_cast<float*>(sm);
#pragma unroll
    for (int s = 0; s < 2 * MAXB; s++) {
        const int o = threadIdx.x + s * kPlaneThreads;
        if (o < npt) {
            const int y = o / Kz, z = o - y * Kz;
            smf[2 * ((y >> 1) * zp + z) + (y & 1)] = qv[s];
        }
    }
    if (Ky & 1)
        for (int z = threadIdx.x; z < Kz; z += kPlaneThreads) smf[2 * ((Ky >> 1) * zp + z) + 1] = 0.f;
    __syncthreads();
    lds_fft<-1, kPlaneThreads>(sm, nrp, zp, 1, pz, tz);
    __syncthreads();
    // Unpack the paired transforms Z = FFT(a + i b): A_k = (Z_k + conj Z_-k)/2,
    // B_k = (Z_k - conj Z_-k)/(2i), into rows y of the half spectrum.
    float2 st[MAXB];
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int o = threadIdx.x + s * kPlaneThreads;
        if (o < Ky * H) {
            int y = o / H, k = o - y * H;
            int p = y >> 1;
            float2 zk = sm[p * zp + k];
            float2 zm = sm[p * zp + (k == 0 ? 0 : Kz - k)];
            if ((y & 1) == 0) st[s] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            else st[s] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int o = threadIdx.x + s * kPlaneThreads;
        if (o < Ky * H) sm[o] = st[s];
    }
    __syncthreads();
    lds_fft<-1, kPlaneThreads>(sm, H, 1, H, py, ty);
    __syncthreads();
    float2* out = FQ + (size_t)x * Ky * H;
    for (int idx = threadIdx.x; idx < Ky * H; idx += kPlaneThreads) out[idx] = sm[idx];
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void fwd_plane_kernel(
    const unsigned* __restrict__ base, const float4* __restrict__ frac, const int* __restrict__ bin_count,
    const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz, FftPlan pz,
    FftPlan py, const float2* __restrict__ twz, const float2* __restrict__ twy, double* __restrict__ acc) {
    fwd_plane_body<MAXB>(blockIdx.x, base, frac, bin_count, bin_atoms, bin_cap, FQ, Kx, Ky, Kz, pz, py, twz, twy,
                         acc);
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void fwd_plane_adjoint_kernel(
    const unsigned* __restrict__ base, const float4* __restrict__ frac, const int* __restrict__ bin_count,
    const int* __restrict__ bin_atoms, int bin_cap, float2* __restrict__ FQ, int Kx, int Ky, int Kz, FftPlan pz,
    FftPlan py, const float2* __restrict__ twz, const float2* __restrict__ twy, const float4* __restrict__ wv,
    const float* __restrict__ scale_ptr) {
    fwd_plane_body<MAXB, 1>(blockIdx.x, base, frac, bin_count, bin_atoms, bin_cap, FQ, Kx, Ky, Kz, pz, py, twz,
                            twy, nullptr, wv, scale_ptr);
}

// Batched: block -> (system, x plane) from the block map.
template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void fwd_plane_batch(const BatchSys* __restrict__ bs,
                                                                 const int2* __restrict__ map, int energy) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    fwd_plane_body<MAXB>(m.y, P.base, P.frac, P.bin_count, P.bin_atoms, P.bin_cap, P.FQ, P.K[0], P.K[1], P.K[2],
                         b.pz, b.py, P.d_tw[2], P.d_tw[1], energy ? P.acc : nullptr);
}

// One block per (y, chunk of kConvChunk kz): FFT along x, energy, BC
// multiply, inverse FFT along x.
__device__ __forceinline__ void xconv_body(int y, int chunk, float2* __restrict__ FQ, const float* __restrict__ bc,
                                           int Kx, int Ky, int Kz, const FftPlan& px,
                                           const float2* __restrict__ twx, int energy, double* __restrict__ acc,
                                           int* __restrict__ bin_count) {
    __shared__ float2 sm[256 * kConvPitch];
    __shared__ float2 tx[256];
    __shared__ double red[kConvThreads / 64];
    const int H = Kz / 2 + 1;
    const int kz0 = chunk * kConvChunk;
    const int cw = min(kConvChunk, H - kz0);
    // the forward plane kernels have consumed the x-plane bins: clear them
    if (y == 0 && chunk == 0)
        for (int b = threadIdx.x; b < Kx; b += kConvThreads) bin_count[b] = 0;
    load_twiddles<kConvThreads>(tx, twx, Kx);
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        sm[xx * kConvPitch + c] = FQ[((size_t)xx * Ky + y) * H + kz0 + c];
    }
    __syncthreads();
    lds_fft<-1, kConvThreads>(sm, cw, 1, kConvPitch, px, tx);
    __syncthreads();
    double e = 0.0;
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        int kz = kz0 + c;
        float b = bc[((size_t)xx * Ky + y) * H + kz];
        float2 v = sm[xx * kConvPitch + c];
        if (energy) {
            double m = (kz == 0 || (2 * kz == Kz)) ? 1.0 : 2.0;
            e += m * (double)b * ((double)v.x * v.x + (double)v.y * v.y);
        }
        sm[xx * kConvPitch + c] = make_float2(v.x * b, v.y * b);
    }
    __syncthreads();
    lds_fft<1, kConvThreads>(sm, cw, 1, kConvPitch, px, tx);
    __syncthreads();
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        FQ[((size_t)xx * Ky + y) * H + kz0 + c] = sm[xx * kConvPitch + c];
    }
    if (energy) {
        double t = block_sum<kConvThreads>(e, red);
        if (threadIdx.x == 0) atomicAdd(acc + SPONGE_PART_RECIPROCAL, 0.5 * t);
    }
}

__global__ __launch_bounds__(kConvThreads) void xconv_kernel(float2* __restrict__ FQ, const float* __restrict__ bc,
                                                             int Kx, int Ky, int Kz, FftPlan px,
                                                             const float2* __restrict__ twx, int energy,
                                                             double* __restrict__ acc, int* __restrict__ bin_count) {
    xconv_body(blockIdx.x, blockIdx.y, FQ, bc, Kx, Ky, Kz, px, twx, energy, acc, bin_count);
}

// Batched: the local block l is (y = l mod Ky, chunk = l / Ky).
__global__ __launch_bounds__(kConvThreads) void xconv_batch(const BatchSys* __restrict__ bs,
                                                            const int2* __restrict__ map, int energy) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    const int Ky = P.K[1];
    const int chunk = m.y / Ky;
    xconv_body(m.y - chunk * Ky, chunk, P.FQ, P.d_bc, P.K[0], Ky, P.K[2], b.px, P.d_tw[0], energy, P.acc,
               P.bin_count);
}

template <int MAXB>
__device__ __forceinline__ void inv_plane_body(int x, const float2* __restrict__ FQ, float* __restrict__ R, int Ky,
                                               int Kz, const FftPlan& pz, const FftPlan& py,
                                               const float2* __restrict__ twz, const float2* __restrict__ twy) {
    extern __shared__ float2 sm[];
    const int H = Kz / 2 + 1, nrp = (Ky + 1) / 2;
    const int zp = Kz + 1;  // odd pitch of the packed z-lines
    const int plane = max(Ky * H, nrp * zp);
    float2* tz = sm + plane;
    float2* ty = tz + Kz;
    load_twiddles<kPlaneThreads>(tz, twz, Kz);
    load_twiddles<kPlaneThreads>(ty, twy, Ky);
    const float2* in = FQ + (size_t)x * Ky * H;
    for (int idx = threadIdx.x; idx < Ky * H; idx += kPlaneThreads) sm[idx] = in[idx];
    __syncthreads();
    lds_fft<1, kPlaneThreads>(sm, H, 1, H, py, ty);
    __syncthreads();
    // Repack rows (2p, 2p+1) into one complex line Z = A + i B over all kz,
    // extending each row by Hermitian symmetry; self-conjugate bins keep only
    // their real part (the C2R convention).
    float2 st[MAXB];
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int o = threadIdx.x + s * kPlaneThreads;
        if (o < nrp * Kz) {
            int p = o / Kz, k = o - p * Kz;
            bool hi = k > Kz / 2;
            int kk = hi ? Kz - k : k;
            bool selfc = (kk == 0) || (2 * kk == Kz);
            float2 A = sm[(2 * p) * H + kk];
            float2 B = (2 * p + 1 < Ky) ? sm[(2 * p + 1) * H + kk] : make_float2(0.f, 0.f);
            if (hi) {
                A.y = -A.y;
                B.y = -B.y;
            }
            if (selfc) {
                A.y = 0.f;
                B.y = 0.f;
            }
            st[s] = make_float2(A.x - B.y, A.y + B.x);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int o = threadIdx.x + s * kPlaneThreads;
        if (o < nrp * Kz) {
            const int p = o / Kz;
            sm[o + p] = st[s];  // line p at pitch zp = Kz + 1
        }
    }
    __syncthreads();
    lds_fft<1, kPlaneThreads>(sm, nrp, zp, 1, pz, tz);
    __syncthreads();
    float* Rp = R + (size_t)x * Ky * Kz;
    for (int idx = threadIdx.x; idx < nrp * Kz; idx += kPlaneThreads) {
        int p = idx / Kz, z = idx - p * Kz;
        float2 v = sm[p * zp + z];
        Rp[(2 * p) * Kz + z] = v.x;
        if (2 * p + 1 < Ky) Rp[(2 * p + 1) * Kz + z] = v.y;
    }
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void inv_plane_kernel(const float2* __restrict__ FQ,
                                                                  float* __restrict__ R, int Ky, int Kz, FftPlan pz,
                                                                  FftPlan py, const float2* __restrict__ twz,
                                                                  const float2* __restrict__ twy) {
    inv_plane_body<MAXB>(blockIdx.x, FQ, R, Ky, Kz, pz, py, twz, twy);
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void inv_plane_batch(const BatchSys* __restrict__ bs,
                                                                 const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_pme& P = b.pme;
    inv_plane_body<MAXB>(m.y, P.FQ, P.R, P.K[1], P.K[2], b.pz, b.py, P.d_tw[2], P.d_tw[1]);
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
        frc[3 * i] += -qi * sx * ibx;
        frc[3 * i + 1] += -qi * sy * iby;
        frc[3 * i + 2] += -qi * sz * ibz;
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

__global__ __launch_bounds__(kGatherThreads) void pme_gather_kernel(int n, const unsigned* __restrict__ base,
                                                                    const float4* __restrict__ frac,
                                                                    const float* __restrict__ R, int Kx, int Ky,
                                                                    int Kz, float ibx, float iby, float ibz,
                                                                    float* __restrict__ frc) {
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
    int m = K;
    while (m % 4 == 0 && p->npass < 12) {
        p->radix[p->npass++] = 4;
        m /= 4;
    }
    for (int r : {2, 3, 5, 7})
        while (m % r == 0 && p->npass < 12) {
            p->radix[p->npass++] = r;
            m /= r;
        }
    return m == 1 ? 0 : -1;
}

// Dynamic LDS of the forward plane kernel: the fixed-point spread grid
// (8 B per point) or the complex plane, whichever is larger, then the z and
// y twiddles.
size_t fwd_plane_lds(int Ky, int Kz) {
    const long H = Kz / 2 + 1;
    const long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * (Kz + 1));
    return (size_t)(std::max(plane, (long)Ky * Kz) + Ky + Kz) * sizeof(float2);
}

int pick_maxb(long need) {
    for (int b : {1, 2, 4, 8, 16})
        if (need <= b) return b;
    return -1;
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
                size_t idx = ((size_t)x * Ky + y) * H + z;
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
            set_error("PME grid size %d must be >= 4 with prime factors 2,3,5,7 only", pme->K[d]);
            delete pme;
            return SPONGE_NB_EARG;
        }
    }
    const int Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * (Kz + 1));
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        max_lds = 64 * 1024;
    const size_t lds_fwd = fwd_plane_lds(Ky, Kz);
    const int mb = pick_maxb((plane + kPlaneThreads - 1) / kPlaneThreads);
    if (mb < 0 || lds_fwd > (size_t)max_lds || Kx > 256 ||
        pick_maxb(((long)Kx * kConvChunk / 2 + kConvThreads - 1) / kConvThreads) < 0) {
        set_error("PME grid %dx%dx%d too large for the in-LDS plane FFT (%zu B of LDS, %d available)", Kx, Ky, Kz,
                  lds_fwd, max_lds);
        delete pme;
        return SPONGE_NB_EARG;
    }
    if (lds_fwd > 64 * 1024) {
        const void* fn = mb == 1 ? (const void*)fwd_plane_kernel<1>
                         : mb == 2 ? (const void*)fwd_plane_kernel<2>
                         : mb == 4 ? (const void*)fwd_plane_kernel<4>
                         : mb == 8 ? (const void*)fwd_plane_kernel<8>
                                   : (const void*)fwd_plane_kernel<16>;
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fwd) != hipSuccess) {
            set_error("cannot raise the LDS limit of the plane kernel to %zu B", lds_fwd);
            delete pme;
            return SPONGE_NB_EHIP;
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
    if (pme->zero_type) (void)hipFree(pme->zero_type);
    if (pme->zero_ab) (void)hipFree(pme->zero_ab);
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

// The reciprocal pipeline; forces added to d_frc (original atom order).
int snb_pme_reciprocal(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge, float* d_frc,
                       int energy, int prep_done, hipStream_t s) {
    const int n = pme->n, Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    FftPlan px, py, pz;
    plan_axis(Kx, &px);
    plan_axis(Ky, &py);
    plan_axis(Kz, &pz);
    if (!prep_done) {
        pme_prep_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_uint_crd, d_charge, pme_prep_args(pme), energy);
        SNB_CHECK_LAUNCH();
    }
    long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * (Kz + 1));
    int mb = pick_maxb((plane + kPlaneThreads - 1) / kPlaneThreads);
    size_t lds = (plane + Ky + Kz) * sizeof(float2);
    size_t lds_fwd = fwd_plane_lds(Ky, Kz);
    double* acc = energy ? pme->acc : nullptr;
#define SNB_FWD(MB) fwd_plane_kernel<MB><<<Kx, kPlaneThreads, lds_fwd, s>>>(pme->base, pme->frac, pme->bin_count, pme->bin_atoms, pme->bin_cap, pme->FQ, Kx, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc)
    switch (mb) {
        case 1: SNB_FWD(1); break;
        case 2: SNB_FWD(2); break;
        case 4: SNB_FWD(4); break;
        case 8: SNB_FWD(8); break;
        default: SNB_FWD(16); break;
    }
#undef SNB_FWD
    SNB_CHECK_LAUNCH();
    dim3 cg(Ky, ceil_div(H, kConvChunk));
    xconv_kernel<<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc,
                                             pme->bin_count);
    SNB_CHECK_LAUNCH();
    switch (mb) {
        case 1: inv_plane_kernel<1><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 2: inv_plane_kernel<2><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 4: inv_plane_kernel<4><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 8: inv_plane_kernel<8><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        default: inv_plane_kernel<16><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
    }
    SNB_CHECK_LAUNCH();
    (void)n;
    return d_frc ? snb_pme_gather(pme, d_frc, s) : SPONGE_NB_OK;
}

// Reciprocal forces from the convolved grid R, added to d_frc.
int snb_pme_gather(sponge_pme* pme, float* d_frc, hipStream_t s) {
    const int n = pme->n, Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2];
    pme_gather_kernel<<<ceil_div((long)n * 16, kGatherThreads), kGatherThreads, 0, s>>>(
        n, pme->base, pme->frac, pme->R, Kx, Ky, Kz, (float)Kx / pme->box[0], (float)Ky / pme->box[1],
        (float)Kz / pme->box[2], d_frc);
    SNB_CHECK_LAUNCH();
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

int snb_pme_plane_maxb(const sponge_pme* pme) {
    const int Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    const long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * (Kz + 1));
    return pick_maxb((plane + kPlaneThreads - 1) / kPlaneThreads);
}

size_t snb_pme_lds_fwd(const sponge_pme* pme) { return fwd_plane_lds(pme->K[1], pme->K[2]); }

size_t snb_pme_lds_inv(const sponge_pme* pme) {
    const int Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    const long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * (Kz + 1));
    return (size_t)(plane + Ky + Kz) * sizeof(float2);
}

int snb_batch_reciprocal(sponge_batch* b, int energy, hipStream_t s) {
#define SNB_FWDB(MB)                                                                                         \
    fwd_plane_batch<MB><<<b->nblk[kStFwd], kPlaneThreads, b->lds_fwd, s>>>(b->d, b->map[kStFwd], energy)
#define SNB_INVB(MB) inv_plane_batch<MB><<<b->nblk[kStInv], kPlaneThreads, b->lds_inv, s>>>(b->d, b->map[kStInv])
    switch (b->mb_plane) {
        case 1: SNB_FWDB(1); break;
        case 2: SNB_FWDB(2); break;
        case 4: SNB_FWDB(4); break;
        case 8: SNB_FWDB(8); break;
        default: SNB_FWDB(16); break;
    }
    SNB_CHECK_LAUNCH();
    xconv_batch<<<b->nblk[kStConv], kConvThreads, 0, s>>>(b->d, b->map[kStConv], energy);
    SNB_CHECK_LAUNCH();
    switch (b->mb_plane) {
        case 1: SNB_INVB(1); break;
        case 2: SNB_INVB(2); break;
        case 4: SNB_INVB(4); break;
        case 8: SNB_INVB(8); break;
        default: SNB_INVB(16); break;
    }
    SNB_CHECK_LAUNCH();
#undef SNB_FWDB
#undef SNB_INVB
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

// Raises the dynamic-LDS limit of the batched forward plane kernel when the
// largest system's plane needs more than 64 KB.
int snb_batch_pme_attributes(int mb, size_t lds_fwd) {
    if (lds_fwd <= 64 * 1024) return SPONGE_NB_OK;
    const void* fn = mb == 1 ? (const void*)fwd_plane_batch<1>
                     : mb == 2 ? (const void*)fwd_plane_batch<2>
                     : mb == 4 ? (const void*)fwd_plane_batch<4>
                     : mb == 8 ? (const void*)fwd_plane_batch<8>
                               : (const void*)fwd_plane_batch<16>;
    SNB_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fwd));
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
    const int mb = snb_pme_plane_maxb(pme);
    const size_t lds_fwd = fwd_plane_lds(Ky, Kz), lds = snb_pme_lds_inv(pme);
    if (lds_fwd > 64 * 1024) {
        const void* fn = mb == 1 ? (const void*)fwd_plane_adjoint_kernel<1>
                         : mb == 2 ? (const void*)fwd_plane_adjoint_kernel<2>
                         : mb == 4 ? (const void*)fwd_plane_adjoint_kernel<4>
                         : mb == 8 ? (const void*)fwd_plane_adjoint_kernel<8>
                                   : (const void*)fwd_plane_adjoint_kernel<16>;
        SNB_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fwd));
    }
#define SNB_FWDA(MB)                                                                                           \
    fwd_plane_adjoint_kernel<MB><<<Kx, kPlaneThreads, lds_fwd, s>>>(pme->base, pme->frac, pme->bin_count,      \
                                                                    pme->bin_atoms, pme->bin_cap, pme->FQ, Kx, \
                                                                    Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], \
                                                                    wv, scale)
    switch (mb) {
        case 1: SNB_FWDA(1); break;
        case 2: SNB_FWDA(2); break;
        case 4: SNB_FWDA(4); break;
        case 8: SNB_FWDA(8); break;
        default: SNB_FWDA(16); break;
    }
#undef SNB_FWDA
    SNB_CHECK_LAUNCH();
    dim3 cg(Ky, ceil_div(H, kConvChunk));
    xconv_kernel<<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], 0, pme->acc,
                                             pme->bin_count);
    SNB_CHECK_LAUNCH();
    switch (mb) {
        case 1: inv_plane_kernel<1><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 2: inv_plane_kernel<2><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 4: inv_plane_kernel<4><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 8: inv_plane_kernel<8><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        default: inv_plane_kernel<16><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
    }
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
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
