// Smooth PME reciprocal part for gfx950 (Particle_Mesh_Ewald,
// PME_force/PME_force.cu), with a hand-written 3D real-to-complex FFT in
// place of cuFFT.
//
// Pipeline per call (grid layout [x][y][z], z fastest, half spectrum
// [kx][ky][kz <= Kz/2] exactly as the reference's PME_Q / PME_FQ / PME_BC):
//   pme_prep      per atom: grid base index and spline fraction from the uint
//                 coordinate with one 32x32->64-bit multiply (t = u K / 2^32)
//   pme_spread    one wave per atom, lane = one of the 64 order-4 stencil
//                 points, FP32 atomics into Q
//   fwd_plane     one block per x plane: real z-rows packed two at a time
//                 into complex z-FFTs and unpacked (R2C), then y-FFTs; the plane
//                 of Q is zeroed for the next call
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

using namespace snb;

namespace {

struct FftPlan {
    int K;
    int npass;
    int radix[12];
};

constexpr int kPlaneThreads = 1024;
constexpr int kConvThreads = 256;
constexpr int kConvChunk = 16;
constexpr int kGatherThreads = 256;
constexpr int kSpreadAtomsPerWave = 4;
// PME acc slots 6 and 7 hold sum q^2 and sum q between prep and finalize.
constexpr int kAccQ2 = 6;
constexpr int kAccQ1 = 7;

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

// One Stockham radix-R step over nlines lines of length K held in LDS; element
// (line l, position p) lives at buf[l*ls + p*ks]. Every thread first loads and
// transforms all of its butterflies into registers, then all threads write.
template <int R, int SIGN, int NT, int MAXB>
__device__ void fft_pass(float2* buf, int nlines, int ls, int ks, int K, int Ns, const float2* __restrict__ tw) {
    const int per = K / R;
    const int nb = nlines * per;
    const int tws = K / (Ns * R);
    float2 v[MAXB][R];
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int b = threadIdx.x + s * NT;
        if (b < nb) {
            int l = b / per, j = b - l * per;
            int k = j % Ns;
#pragma unroll
            for (int q = 0; q < R; q++) {
                float2 x = buf[l * ls + (j + q * per) * ks];
                if (q > 0) {
                    float2 t = tw[q * k * tws];
                    if (SIGN > 0) t.y = -t.y;
                    x = cmul(x, t);
                }
                v[s][q] = x;
            }
            dft<R, SIGN>(v[s]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int b = threadIdx.x + s * NT;
        if (b < nb) {
            int l = b / per, j = b - l * per;
            int k = j % Ns;
            int d = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) buf[l * ls + (d + q * Ns) * ks] = v[s][q];
        }
    }
    __syncthreads();
}

template <int SIGN, int NT, int MAXB>
__device__ void lds_fft(float2* buf, int nlines, int ls, int ks, const FftPlan& p, const float2* __restrict__ tw) {
    int Ns = 1;
    for (int ps = 0; ps < p.npass; ps++) {
        int R = p.radix[ps];
        switch (R) {
            case 4: fft_pass<4, SIGN, NT, MAXB>(buf, nlines, ls, ks, p.K, Ns, tw); break;
            case 2: fft_pass<2, SIGN, NT, MAXB>(buf, nlines, ls, ks, p.K, Ns, tw); break;
            case 3: fft_pass<3, SIGN, NT, MAXB>(buf, nlines, ls, ks, p.K, Ns, tw); break;
            case 5: fft_pass<5, SIGN, NT, MAXB>(buf, nlines, ls, ks, p.K, Ns, tw); break;
            default: fft_pass<7, SIGN, NT, MAXB>(buf, nlines, ls, ks, p.K, Ns, tw); break;
        }
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------
// PME_Atom_Near (PME_force.cu:367-421) without the 64-index cache: base and
// fraction of t = u K / 2^32 from the exact 64-bit product.
__global__ void pme_prep_kernel(int n, const unsigned int* __restrict__ uc, const float* __restrict__ q, int Kx,
                                int Ky, int Kz, int4* __restrict__ base, float4* __restrict__ frac, int energy,
                                double* __restrict__ acc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double q2 = 0.0, q1 = 0.0;
    if (i < n) {
        const float s = 2.3283064365386963e-10f;  // 2^-32
        unsigned long long tx = (unsigned long long)uc[3 * i] * (unsigned)Kx;
        unsigned long long ty = (unsigned long long)uc[3 * i + 1] * (unsigned)Ky;
        unsigned long long tz = (unsigned long long)uc[3 * i + 2] * (unsigned)Kz;
        base[i] = make_int4((int)(tx >> 32), (int)(ty >> 32), (int)(tz >> 32), 0);
        frac[i] = make_float4((float)(unsigned)tx * s, (float)(unsigned)ty * s, (float)(unsigned)tz * s, q[i]);
        if (energy) {
            q1 = q[i];
            q2 = q1 * q1;
        }
    }
    if (energy) {
        q2 = wave_sum(q2);
        q1 = wave_sum(q1);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(acc + kAccQ2, q2);
            atomicAdd(acc + kAccQ1, q1);
        }
    }
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

// PME_Q_Spread (PME_force.cu:428-471): grid point (base - k) gets
// q * M4(fx+kx) M4(fy+ky) M4(fz+kz).
__global__ void pme_spread_kernel(int n, const int4* __restrict__ base, const float4* __restrict__ frac, int Kx,
                                  int Ky, int Kz, float* __restrict__ Q) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int kx = lane >> 4, ky = (lane >> 2) & 3, kz = lane & 3;
    for (int t = 0; t < kSpreadAtomsPerWave; t++) {
        int i = wave * kSpreadAtomsPerWave + t;
        if (i >= n) return;
        int4 b = base[i];
        float4 f = frac[i];
        float w = f.w * bsp_w(f.x, kx) * bsp_w(f.y, ky) * bsp_w(f.z, kz);
        int gx = b.x - kx, gy = b.y - ky, gz = b.z - kz;
        gx += gx < 0 ? Kx : 0;
        gy += gy < 0 ? Ky : 0;
        gz += gz < 0 ? Kz : 0;
        atomicAdd(Q + ((size_t)gx * Ky + gy) * Kz + gz, w);
    }
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void fwd_plane_kernel(float* __restrict__ Q, float2* __restrict__ FQ,
                                                                  int Ky, int Kz, FftPlan pz, FftPlan py,
                                                                  const float2* __restrict__ twz,
                                                                  const float2* __restrict__ twy,
                                                                  double* __restrict__ acc) {
    extern __shared__ float2 sm[];
    const int x = blockIdx.x;
    const int H = Kz / 2 + 1, nrp = (Ky + 1) / 2;
    float* Qp = Q + (size_t)x * Ky * Kz;
    if (x == 0 && threadIdx.x == 0 && acc) acc[SPONGE_PART_RECIPROCAL] = 0.0;
    for (int idx = threadIdx.x; idx < nrp * Kz; idx += kPlaneThreads) {
        int p = idx / Kz, z = idx - p * Kz;
        float a = Qp[(2 * p) * Kz + z];
        float b = (2 * p + 1 < Ky) ? Qp[(2 * p + 1) * Kz + z] : 0.f;
        sm[idx] = make_float2(a, b);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < Ky * Kz; idx += kPlaneThreads) Qp[idx] = 0.f;
    lds_fft<-1, kPlaneThreads, MAXB>(sm, nrp, Kz, 1, pz, twz);
    // Unpack the paired transforms Z = FFT(a + i b): A_k = (Z_k + conj Z_-k)/2,
    // B_k = (Z_k - conj Z_-k)/(2i), into rows y of the half spectrum.
    float2 st[MAXB];
#pragma unroll
    for (int s = 0; s < MAXB; s++) {
        int o = threadIdx.x + s * kPlaneThreads;
        if (o < Ky * H) {
            int y = o / H, k = o - y * H;
            int p = y >> 1;
            float2 zk = sm[p * Kz + k];
            float2 zm = sm[p * Kz + (k == 0 ? 0 : Kz - k)];
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
    lds_fft<-1, kPlaneThreads, MAXB>(sm, H, 1, H, py, twy);
    float2* out = FQ + (size_t)x * Ky * H;
    for (int idx = threadIdx.x; idx < Ky * H; idx += kPlaneThreads) out[idx] = sm[idx];
}

template <int MAXB>
__global__ __launch_bounds__(kConvThreads) void xconv_kernel(float2* __restrict__ FQ, const float* __restrict__ bc,
                                                             int Kx, int Ky, int Kz, FftPlan px,
                                                             const float2* __restrict__ twx, int energy,
                                                             double* __restrict__ acc) {
    __shared__ float2 sm[256 * kConvChunk];
    __shared__ double red[kConvThreads / 64];
    const int H = Kz / 2 + 1;
    const int y = blockIdx.x, kz0 = blockIdx.y * kConvChunk;
    const int cw = min(kConvChunk, H - kz0);
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        sm[xx * kConvChunk + c] = FQ[((size_t)xx * Ky + y) * H + kz0 + c];
    }
    __syncthreads();
    lds_fft<-1, kConvThreads, MAXB>(sm, cw, 1, kConvChunk, px, twx);
    double e = 0.0;
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        int kz = kz0 + c;
        float b = bc[((size_t)xx * Ky + y) * H + kz];
        float2 v = sm[xx * kConvChunk + c];
        if (energy) {
            double m = (kz == 0 || (2 * kz == Kz)) ? 1.0 : 2.0;
            e += m * (double)b * ((double)v.x * v.x + (double)v.y * v.y);
        }
        sm[xx * kConvChunk + c] = make_float2(v.x * b, v.y * b);
    }
    __syncthreads();
    lds_fft<1, kConvThreads, MAXB>(sm, cw, 1, kConvChunk, px, twx);
    for (int idx = threadIdx.x; idx < Kx * cw; idx += kConvThreads) {
        int xx = idx / cw, c = idx - xx * cw;
        FQ[((size_t)xx * Ky + y) * H + kz0 + c] = sm[xx * kConvChunk + c];
    }
    if (energy) {
        double t = block_sum<kConvThreads>(e, red);
        if (threadIdx.x == 0) atomicAdd(acc + SPONGE_PART_RECIPROCAL, 0.5 * t);
    }
}

template <int MAXB>
__global__ __launch_bounds__(kPlaneThreads) void inv_plane_kernel(const float2* __restrict__ FQ,
                                                                  float* __restrict__ R, int Ky, int Kz, FftPlan pz,
                                                                  FftPlan py, const float2* __restrict__ twz,
                                                                  const float2* __restrict__ twy) {
    extern __shared__ float2 sm[];
    const int x = blockIdx.x;
    const int H = Kz / 2 + 1, nrp = (Ky + 1) / 2;
    const float2* in = FQ + (size_t)x * Ky * H;
    for (int idx = threadIdx.x; idx < Ky * H; idx += kPlaneThreads) sm[idx] = in[idx];
    __syncthreads();
    lds_fft<1, kPlaneThreads, MAXB>(sm, H, 1, H, py, twy);
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
        if (o < nrp * Kz) sm[o] = st[s];
    }
    __syncthreads();
    lds_fft<1, kPlaneThreads, MAXB>(sm, nrp, Kz, 1, pz, twz);
    float* Rp = R + (size_t)x * Ky * Kz;
    for (int idx = threadIdx.x; idx < nrp * Kz; idx += kPlaneThreads) {
        int p = idx / Kz, z = idx - p * Kz;
        float2 v = sm[idx];
        Rp[(2 * p) * Kz + z] = v.x;
        if (2 * p + 1 < Ky) Rp[(2 * p + 1) * Kz + z] = v.y;
    }
}

// PME_Final (PME_force.cu:485-550): 16 lanes per atom (lane = kx, ky), each
// walking the 4 kz points; force += -q sum R dtheta (K/L).
__global__ __launch_bounds__(kGatherThreads) void pme_gather_kernel(int n, const int4* __restrict__ base,
                                                                    const float4* __restrict__ frac,
                                                                    const float* __restrict__ R, int Kx, int Ky,
                                                                    int Kz, float ibx, float iby, float ibz,
                                                                    float* __restrict__ frc) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gid >> 4, sub = gid & 15;
    const int kx = sub >> 2, ky = sub & 3;
    float sx = 0.f, sy = 0.f, sz = 0.f, qi = 0.f;
    if (i < n) {
        int4 b = base[i];
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

int pick_maxb(long need) {
    for (int b : {1, 2, 4, 8})
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
    long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * Kz);
    if (pick_maxb((plane + kPlaneThreads - 1) / kPlaneThreads) < 0 || plane * 8 > 64 * 1024 ||
        Kx > 256 || pick_maxb(((long)Kx * kConvChunk / 2 + kConvThreads - 1) / kConvThreads) < 0) {
        set_error("PME grid %dx%dx%d too large for the in-LDS plane FFT", Kx, Ky, Kz);
        delete pme;
        return SPONGE_NB_EARG;
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
    A((void**)&pme->Q, NG * sizeof(float));
    A((void**)&pme->FQ, (size_t)pme->nfft * sizeof(float2));
    A((void**)&pme->R, NG * sizeof(float));
    A((void**)&pme->base, (size_t)atom_numbers * sizeof(int4));
    A((void**)&pme->frac, (size_t)atom_numbers * sizeof(float4));
    A((void**)&pme->corr_atom, (size_t)atom_numbers * sizeof(float));
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
    if (e == hipSuccess) e = hipMemset(pme->Q, 0, NG * sizeof(float));
    if (e == hipSuccess) e = hipMemset(pme->corr_atom, 0, (size_t)atom_numbers * sizeof(float));
    if (e == hipSuccess) e = hipMemset(pme->acc, 0, SPONGE_NB_NPARTS * sizeof(double));
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
    void* ptrs[] = {pme->d_bc, pme->Q,         pme->FQ,       pme->R,        pme->base,
                    pme->frac, pme->corr_atom, pme->acc,      pme->d_tw[0],  pme->d_tw[1],
                    pme->d_tw[2]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
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
                       int energy, hipStream_t s) {
    const int n = pme->n, Kx = pme->K[0], Ky = pme->K[1], Kz = pme->K[2], H = Kz / 2 + 1;
    FftPlan px, py, pz;
    plan_axis(Kx, &px);
    plan_axis(Ky, &py);
    plan_axis(Kz, &pz);
    pme_prep_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_uint_crd, d_charge, Kx, Ky, Kz, pme->base, pme->frac,
                                                     energy, pme->acc);
    SNB_CHECK_LAUNCH();
    int nw = ceil_div(n, kSpreadAtomsPerWave);
    pme_spread_kernel<<<ceil_div(nw, 4), 256, 0, s>>>(n, pme->base, pme->frac, Kx, Ky, Kz, pme->Q);
    SNB_CHECK_LAUNCH();
    long plane = std::max((long)Ky * H, (long)((Ky + 1) / 2) * Kz);
    int mb = pick_maxb((plane + kPlaneThreads - 1) / kPlaneThreads);
    size_t lds = plane * sizeof(float2);
    double* acc = energy ? pme->acc : nullptr;
    switch (mb) {
        case 1: fwd_plane_kernel<1><<<Kx, kPlaneThreads, lds, s>>>(pme->Q, pme->FQ, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc); break;
        case 2: fwd_plane_kernel<2><<<Kx, kPlaneThreads, lds, s>>>(pme->Q, pme->FQ, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc); break;
        case 4: fwd_plane_kernel<4><<<Kx, kPlaneThreads, lds, s>>>(pme->Q, pme->FQ, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc); break;
        default: fwd_plane_kernel<8><<<Kx, kPlaneThreads, lds, s>>>(pme->Q, pme->FQ, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1], acc); break;
    }
    SNB_CHECK_LAUNCH();
    dim3 cg(Ky, ceil_div(H, kConvChunk));
    int mbc = pick_maxb(((long)Kx * kConvChunk / 2 + kConvThreads - 1) / kConvThreads);
    switch (mbc) {
        case 1: xconv_kernel<1><<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc); break;
        case 2: xconv_kernel<2><<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc); break;
        case 4: xconv_kernel<4><<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc); break;
        default: xconv_kernel<8><<<cg, kConvThreads, 0, s>>>(pme->FQ, pme->d_bc, Kx, Ky, Kz, px, pme->d_tw[0], energy, pme->acc); break;
    }
    SNB_CHECK_LAUNCH();
    switch (mb) {
        case 1: inv_plane_kernel<1><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 2: inv_plane_kernel<2><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        case 4: inv_plane_kernel<4><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
        default: inv_plane_kernel<8><<<Kx, kPlaneThreads, lds, s>>>(pme->FQ, pme->R, Ky, Kz, pz, py, pme->d_tw[2], pme->d_tw[1]); break;
    }
    SNB_CHECK_LAUNCH();
    if (d_frc) {
        pme_gather_kernel<<<ceil_div((long)n * 16, kGatherThreads), kGatherThreads, 0, s>>>(
            n, pme->base, pme->frac, pme->R, Kx, Ky, Kz, (float)Kx / pme->box[0], (float)Ky / pme->box[1],
            (float)Kz / pme->box[2], d_frc);
        SNB_CHECK_LAUNCH();
    }
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
    int rc = snb_pme_reciprocal(pme, d_uint_crd, d_charge, d_frc, energy, s);
    if (rc || !energy) return rc;
    return snb_pme_finalize(pme, d_direct_atom_energy, need_energy > 0, need_virial > 0, d_potential, d_virial, s);
}

// Host-side read of the PME FP64 parts (after a synchronising copy).
int snb_pme_parts(sponge_pme* pme, double* parts, hipStream_t s) {
    SNB_HIP(hipMemcpyAsync(parts, pme->acc, SPONGE_NB_NPARTS * sizeof(double), hipMemcpyDeviceToHost, s));
    SNB_HIP(hipStreamSynchronize(s));
    return SPONGE_NB_OK;
}
