// FEP / TI frame evaluators for gfx950: the soft-core energy split and
// dH/dlambda of LJ_SOFT_CORE (Lennard_Jones_force/LJ_soft_core.cu:970-1418,
// :1809-1955), the charge-perturbation terms of TI_CORE (TI_core/TI_core.cu:
// 13-125, :670-739) and FEP_CORE's direct-energy split (FEP_core/FEP_core.cu:
// 12-66, :545-558).
//
// One walk over the cluster list (one wave per 4-atom i-cluster, lanes as in
// the pair kernel, each unordered pair taken once from its lower-index atom)
// splits the pairs in two:
//  * ordinary pairs (same mask, neither end state appearing from zero LJ):
//    their energy is linear in lambda for LJ and quadratic for Coulomb
//    through the charges q + t dq, so 13 lambda-free FP64 sums carry every
//    window (vA, vB per mask class; sum q q g, sum (q dq + dq q) g,
//    sum dq dq g with g = erfc(b r)/r; the reference-form derivative sums
//    with erfc(b r); the pair count);
//  * soft-core pairs are appended to a compact list (wave-aggregated, one
//    atomic per wave), and a second launch evaluates every (pair, window)
//    with one thread each, so the handful of waves that meet the perturbed
//    atoms do not serialise the windows.
// The host combines the sums per window. Every sum is FP64 from the exact
// integer displacement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "pair_math.hpp"
#include "../../include/sponge_fep.h"

using namespace snb;

namespace {

constexpr int kFepWaves = 4;
constexpr int kSoftThreads = 256;
constexpr int kWinParams = 8;   // lambda, 1 - lambda, a lambda^p, a (1-lambda)^p, p a lambda^(p-1), p a (1-lambda)^(p-1), t, 0
constexpr int kSoftFields = 6;  // LJ inter, LJ intra, direct inter, direct intra, dH/dl LJ, dH/dl direct
constexpr int kInitialSoftCap = 1 << 15;

// lambda-free sums of ordinary pairs
enum HardSum {
    H_VA_INTRA, H_VB_INTRA, H_VA_INTER, H_VB_INTER,
    H_Q0_INTRA, H_Q1_INTRA, H_Q2_INTRA, H_Q0_INTER, H_Q1_INTER, H_Q2_INTER,
    H_R1, H_R2, H_PAIRS, kHard
};
// reduction slots of the per-atom cross kernel
enum CrossSum { X_RECIP, X_QDQ, X_Q, X_DQ, X_CORR, kCross };

constexpr double kOneOver3SqrtPi = 0.18806319451591877;  // 1 / (3 sqrt(pi))

__global__ void fep_zero_kernel(double* __restrict__ acc, int nacc, int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nacc) acc[i] = 0.0;
    if (i == 0) count[0] = 0;
}

// Walk of the cluster list. TABLE: soft-core handle records
// fe[pt] = (AA, AB, BA, BB), fe[nt^2 + pt] = (soft, same mask, 0, 0)
// for the pair type pt = t_i nt + t_j of combined types; without a table all
// pairs are ordinary with zero LJ and the mask class comes from `mask`
// (original order; NULL: one class).
template <bool TABLE>
__global__ __launch_bounds__(kFepWaves * 64) void fep_list_kernel(
    const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list, int list_cap,
    const uint4* __restrict__ pk, const int* __restrict__ s2o, const float4* __restrict__ fe, int nt, float cx,
    float cy, float cz, float rc2, double beta, const float* __restrict__ dq, const int* __restrict__ mask,
    double* __restrict__ acc, int4* __restrict__ soft, int* __restrict__ count, int soft_cap) {
    __shared__ double red[kFepWaves][kHard];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ic = blockIdx.x * kFepWaves + w;
    double h[kHard];
#pragma unroll
    for (int k = 0; k < kHard; k++) h[k] = 0.0;
    if (ic < flags[4]) {
        const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
        const int islot = ic * kClusterSize + ii;
        const uint4 a = pk[islot];
        const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
        const int ti = unpack_type(a);
        const int ai = s2o[islot];
        const double qi = (double)__uint_as_float(a.w);
        const double dqi = (dq && ai >= 0) ? (double)dq[ai] : 0.0;
        const int mi = (mask && ai >= 0) ? mask[ai] : 0;
        const int n = list_n[ic];
        const uint2* lst = list + (size_t)ic * list_cap;
        for (int base = 0; base < n; base += 4) {
            const int e = base + g;
            const uint2 ent = e < n ? lst[e] : make_uint2((unsigned)ic, 0u);
            const int jc = (int)ent.x;
            const int jslot = jc * kClusterSize + jj;
            const uint4 b = pk[jslot];
            const int ix = (int)((b.x & ~3u) - ax), iy = (int)((b.y & ~3u) - ay), iz = (int)((b.z & ~3u) - az);
            const float r2f = dr2_exact((float)ix * cx, (float)iy * cy, (float)iz * cz);
            const int aj = s2o[jslot];
            const bool in = ((ent.y >> bit) & 1u) && r2f < rc2 && ai >= 0 && ai < aj;
            const int tj = unpack_type(b);
            const int pt = ti * nt + tj;
            bool sp = false, intra = true;
            float4 f0 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (in) {
                if (TABLE) {
                    const float4 f1 = fe[nt * nt + pt];
                    sp = f1.x != 0.f;
                    intra = f1.y != 0.f;
                    f0 = fe[pt];
                } else {
                    intra = !mask || mi == mask[aj];
                }
            }
            // soft-core pairs: one slot per lane, one atomic per wave
            const unsigned long long sb = __ballot(in && sp);
            if (sb) {
                int basei = 0;
                const int leader = __builtin_ctzll(sb);
                if (lane == leader) basei = atomicAdd(count, __popcll(sb));
                basei = __shfl(basei, leader);
                if (in && sp) {
                    const int k = basei + __popcll(sb & ((1ull << lane) - 1ull));
                    if (k < soft_cap) {
                        soft[2 * k] = make_int4(ai, aj, pt, 0);
                        soft[2 * k + 1] = make_int4(ix, iy, iz, 0);
                    }
                }
            }
            if (in && !sp) {
                const double dx = (double)ix * (double)cx, dy = (double)iy * (double)cy, dz = (double)iz * (double)cz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = sqrt(r2);
                const double ec = erfc(beta * r);
                const double gr = ec / r;
                const double qj = (double)__uint_as_float(b.w);
                const double dqj = dq ? (double)dq[aj] : 0.0;
                const double q0 = qi * qj, q1 = qi * dqj + dqi * qj, q2 = dqi * dqj;
                double vA = 0.0, vB = 0.0;
                if (TABLE) {
                    const double ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2;
                    vA = ((double)f0.x * (1.0 / 12.0) * ir6 - (double)f0.y * (1.0 / 6.0)) * ir6;
                    vB = ((double)f0.z * (1.0 / 12.0) * ir6 - (double)f0.w * (1.0 / 6.0)) * ir6;
                }
                if (intra) {
                    h[H_VA_INTRA] += vA;
                    h[H_VB_INTRA] += vB;
                    h[H_Q0_INTRA] += q0 * gr;
                    h[H_Q1_INTRA] += q1 * gr;
                    h[H_Q2_INTRA] += q2 * gr;
                } else {
                    h[H_VA_INTER] += vA;
                    h[H_VB_INTER] += vB;
                    h[H_Q0_INTER] += q0 * gr;
                    h[H_Q1_INTER] += q1 * gr;
                    h[H_Q2_INTER] += q2 * gr;
                }
                h[H_R1] += q1 * ec;
                h[H_R2] += q2 * ec;
            }
            if (in) h[H_PAIRS] += 1.0;
        }
    }
#pragma unroll
    for (int k = 0; k < kHard; k++) {
        const double v = wave_sum(h[k]);
        if (lane == 0) red[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kHard) {
        double s = 0.0;
        for (int k = 0; k < kFepWaves; k++) s += red[k][threadIdx.x];
        if (s != 0.0) atomicAdd(acc + threadIdx.x, s);
    }
}

// Every recorded soft-core pair at window blockIdx.y (LJ_soft_core.cu:
// 1050-1110 energies, :1230-1270 dH/dlambda). sigma^6 of each end state is
// derived per pair from the tables in FP64 (:1058-1065): soft_core_sigma^6
// when A or B is below 1e-6, else max(A / (2 B), soft_core_sigma_min^6).
// Energies keep an end state's terms only while its soft-core radius is
// below the cutoff (:1073, :1087); the derivative terms do not test it (as
// the reference's dH/dlambda kernels).
__global__ __launch_bounds__(kSoftThreads) void fep_soft_kernel(
    const int4* __restrict__ soft, const int* __restrict__ count, int soft_cap, const float4* __restrict__ fe, int nt,
    double cx, double cy, double cz, double rc, double beta, double sig6, double sig6min, const float* __restrict__ q,
    const float* __restrict__ dq, const double* __restrict__ win, double* __restrict__ out) {
    __shared__ double red[kSoftThreads / 64];
    const int wi = blockIdx.y;
    const double lam = win[wi * kWinParams + 0], l_ = win[wi * kWinParams + 1];
    const double aA = win[wi * kWinParams + 2], aB = win[wi * kWinParams + 3];
    const double dA = win[wi * kWinParams + 4], dB = win[wi * kWinParams + 5];
    const double t = win[wi * kWinParams + 6];
    const int ns = min(count[0], soft_cap);
    double s[kSoftFields];
#pragma unroll
    for (int k = 0; k < kSoftFields; k++) s[k] = 0.0;
    for (int k = blockIdx.x * kSoftThreads + threadIdx.x; k < ns; k += gridDim.x * kSoftThreads) {
        const int4 id = soft[2 * k], d = soft[2 * k + 1];
        const float4 f0 = fe[id.z], f1 = fe[nt * nt + id.z];
        const double AA = f0.x, AB = f0.y, BA = f0.z, BB = f0.w;
        const double sA6 = (AA < 1e-6 || AB < 1e-6) ? sig6 : fmax(0.5 * AA / AB, sig6min);
        const double sB6 = (BA < 1e-6 || BB < 1e-6) ? sig6 : fmax(0.5 * BA / BB, sig6min);
        const double dx = (double)d.x * cx, dy = (double)d.y * cy, dz = (double)d.z * cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r6 = r2 * r2 * r2;
        const double uA = r6 + aA * sA6, uB = r6 + aB * sB6;
        const double iA = 1.0 / uA, iB = 1.0 / uB;
        const double sA = cbrt(sqrt(uA)), sB = cbrt(sqrt(uB));
        const double VA = (AA * (1.0 / 12.0) * iA - AB * (1.0 / 6.0)) * iA;
        const double VB = (BA * (1.0 / 12.0) * iB - BB * (1.0 / 6.0)) * iB;
        const double fA = erfc(beta * sA) / sA, fB = erfc(beta * sB) / sB;
        const double xA = exp(-beta * beta * sA * sA), xB = exp(-beta * beta * sB * sB);
        const double qi = (double)q[id.x] + (dq ? t * (double)dq[id.x] : 0.0);
        const double qj = (double)q[id.y] + (dq ? t * (double)dq[id.y] : 0.0);
        const double qq = qi * qj;
        const double dqq = dq ? qi * (double)dq[id.y] + (double)dq[id.x] * qj : 0.0;
        const bool inA = sA < rc, inB = sB < rc;
        const double elj = (inA ? l_ * VA : 0.0) + (inB ? lam * VB : 0.0);
        const double ec = qq * ((inA ? l_ * fA : 0.0) + (inB ? lam * fB : 0.0));
        const double dlj = VB - VA + l_ * dA * sA6 * iA * iA * (AB * (1.0 / 6.0) - AA * (1.0 / 6.0) * iA) +
                           lam * dB * sB6 * iB * iB * (BA * (1.0 / 6.0) * iB - BB * (1.0 / 6.0));
        const double dcf = qq * (fB - fA - l_ * dA * sA6 * iA * (beta * kOneOver3SqrtPi * xA + fA * (1.0 / 6.0)) +
                                 lam * dB * sB6 * iB * (beta * kOneOver3SqrtPi * xB + fB * (1.0 / 6.0))) +
                           dqq * (l_ * fA + lam * fB);
        if (f1.y != 0.f) {
            s[1] += elj;
            s[3] += ec;
        } else {
            s[0] += elj;
            s[2] += ec;
        }
        s[4] += dlj;
        s[5] += dcf;
    }
#pragma unroll
    for (int k = 0; k < kSoftFields; k++) {
        const double v = block_sum<kSoftThreads>(s[k], red);
        if (threadIdx.x == 0 && v != 0.0) atomicAdd(out + kHard + wi * kSoftFields + k, v);
    }
}

// Per-atom cross terms of a charge perturbation dq (TI_core.cu:676-718):
// sum dq phi (phi: convolved grid at the atom), sum q dq, sum q, sum dq, and
// the excluded-pair cross correction -sum (q_i dq_j + dq_i q_j) erf(b r)/r
// over atom i's exclusion list (FP64).
__global__ __launch_bounds__(256) void fep_cross_kernel(int n, const float* __restrict__ q,
                                                        const float* __restrict__ dq, const float* __restrict__ phi,
                                                        const unsigned* __restrict__ uc, double sx, double sy,
                                                        double sz, double beta, const int* __restrict__ es,
                                                        const int* __restrict__ el, const int* __restrict__ en,
                                                        double* __restrict__ acc) {
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[kCross] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const double qi = q[i], dqi = dq[i];
        v[X_RECIP] = dqi * (double)phi[i];
        v[X_QDQ] = qi * dqi;
        v[X_Q] = qi;
        v[X_DQ] = dqi;
        if (es) {
            const unsigned ux = uc[3 * i], uy = uc[3 * i + 1], uz = uc[3 * i + 2];
            for (int k = es[i]; k < es[i] + en[i]; k++) {
                const int j = el[k];
                const double dx = (double)(int)(uc[3 * j] - ux) * sx;
                const double dy = (double)(int)(uc[3 * j + 1] - uy) * sy;
                const double dz = (double)(int)(uc[3 * j + 2] - uz) * sz;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                v[X_CORR] -= (qi * (double)dq[j] + dqi * (double)q[j]) * erf(beta * r) / r;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kCross; k++) {
        const double t = block_sum<256>(v[k], red);
        if (threadIdx.x == 0 && t != 0.0) atomicAdd(acc + k, t);
    }
}

__global__ void fep_fill_kernel(float* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace

// Evaluator scratch of one LJ or PME handle (allocated on first use).
struct FepScratch {
    double* acc = nullptr;      // kHard + SPONGE_FEP_MAX_WINDOWS * kSoftFields (and the cross sums)
    int* count = nullptr;       // soft pairs recorded
    int4* soft = nullptr;       // 2 int4 per soft pair: (atom i, atom j, pair type), (dx, dy, dz) in uint units
    int soft_cap = 0;
    double* win = nullptr;      // SPONGE_FEP_MAX_WINDOWS * kWinParams
    float* phi = nullptr;       // n potentials (PME)
    int* zero_type = nullptr;   // n zeros (tableless walks)
    int n = 0;
    PairScratch ps{};           // cluster-ordered records of the walk
};

void snb_fep_free(void* p) {
    FepScratch* f = static_cast<FepScratch*>(p);
    if (!f) return;
    void* ptrs[] = {f->acc, f->count, f->soft, f->win, f->phi, f->zero_type};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    snb_scratch_free(&f->ps);
    delete f;
}

namespace {

int fep_scratch_alloc(FepScratch* f, int n) {
    f->n = n;
    const int nacc = kHard + SPONGE_FEP_MAX_WINDOWS * kSoftFields;
    SNB_HIP(hipMalloc((void**)&f->acc, nacc * sizeof(double)));
    SNB_HIP(hipMalloc((void**)&f->count, sizeof(int)));
    SNB_HIP(hipMalloc((void**)&f->win, SPONGE_FEP_MAX_WINDOWS * kWinParams * sizeof(double)));
    SNB_HIP(hipMalloc((void**)&f->zero_type, n * sizeof(int)));
    SNB_HIP(hipMemset(f->zero_type, 0, n * sizeof(int)));
    SNB_HIP(hipMalloc((void**)&f->soft, 2 * (size_t)kInitialSoftCap * sizeof(int4)));
    f->soft_cap = kInitialSoftCap;
    return SPONGE_NB_OK;
}

// The scratch is stored in *slot only once every buffer is allocated, so a
// failed allocation leaves the handle as it was and the next call retries.
int fep_scratch(void** slot, int n, FepScratch** out) {
    FepScratch* f = static_cast<FepScratch*>(*slot);
    if (!f) {
        f = new FepScratch();
        const int rc = fep_scratch_alloc(f, n);
        if (rc) {
            snb_fep_free(f);
            return rc;
        }
        *slot = f;
    }
    *out = f;
    return SPONGE_NB_OK;
}

struct WalkArgs {
    const sponge_nl* nl;
    const unsigned* uc;
    const float* q;
    const float* dq;
    const int* type;          // combined types (table) or NULL (zero types)
    const float4* fe;         // soft-core records or NULL
    int nt;
    const int* mask;
    float cutoff;
    float cof[3];
    double beta;
    double sig6, sig6min;
};

// Walk + soft-pair windows; h_sums[kHard] and h_soft[nwin * kSoftFields] on
// the host. Re-runs once with a larger soft-pair list when it overflowed.
int fep_walk(FepScratch* f, const WalkArgs& A, int nwin, const double* h_win, double* h_sums, double* h_soft,
             hipStream_t s) {
    const sponge_nl* nl = A.nl;
    for (int attempt = 0; attempt < 2; attempt++) {
        const int nacc = kHard + nwin * kSoftFields;
        fep_zero_kernel<<<ceil_div(nacc, 256), 256, 0, s>>>(f->acc, nacc, f->count);
        SNB_CHECK_LAUNCH();
        int rc = snb_gather(nl, A.uc, A.q, A.type ? A.type : f->zero_type, &f->ps, s);
        if (rc) return rc;
        const int nblk = ceil_div(nl->ncl_cap, kFepWaves);
        const float rc2 = A.cutoff * A.cutoff;
        if (A.fe)
            fep_list_kernel<true><<<nblk, kFepWaves * 64, 0, s>>>(
                nl->flags, nl->list_n, nl->list, nl->list_cap, f->ps.pk, nl->s2o, A.fe, A.nt, A.cof[0], A.cof[1],
                A.cof[2], rc2, A.beta, A.dq, A.mask, f->acc, f->soft, f->count, f->soft_cap);
        else
            fep_list_kernel<false><<<nblk, kFepWaves * 64, 0, s>>>(
                nl->flags, nl->list_n, nl->list, nl->list_cap, f->ps.pk, nl->s2o, nullptr, 1, A.cof[0], A.cof[1],
                A.cof[2], rc2, A.beta, A.dq, A.mask, f->acc, f->soft, f->count, f->soft_cap);
        SNB_CHECK_LAUNCH();
        if (A.fe && nwin > 0) {
            SNB_HIP(hipMemcpyAsync(f->win, h_win, nwin * kWinParams * sizeof(double), hipMemcpyHostToDevice, s));
            // enough blocks for a few thousand soft pairs per window; larger
            // lists are strided
            const dim3 grid(64, nwin);
            fep_soft_kernel<<<grid, kSoftThreads, 0, s>>>(
                f->soft, f->count, f->soft_cap, A.fe, A.nt, (double)A.cof[0], (double)A.cof[1], (double)A.cof[2],
                (double)A.cutoff, A.beta, A.sig6, A.sig6min, A.q, A.dq, f->win, f->acc);
            SNB_CHECK_LAUNCH();
        }
        std::vector<double> h(nacc);
        int cnt = 0;
        SNB_HIP(hipMemcpyAsync(h.data(), f->acc, nacc * sizeof(double), hipMemcpyDeviceToHost, s));
        SNB_HIP(hipMemcpyAsync(&cnt, f->count, sizeof(int), hipMemcpyDeviceToHost, s));
        SNB_HIP(hipStreamSynchronize(s));
        if (cnt <= f->soft_cap) {
            memcpy(h_sums, h.data(), kHard * sizeof(double));
            if (h_soft) memcpy(h_soft, h.data() + kHard, (size_t)nwin * kSoftFields * sizeof(double));
            return SPONGE_NB_OK;
        }
        // the capacity stays 0 until the larger list is allocated, so a
        // failed allocation leaves no pointer the kernels would write through
        SNB_HIP(hipFree(f->soft));
        f->soft = nullptr;
        f->soft_cap = 0;
        const int cap = cnt + cnt / 4;
        SNB_HIP(hipMalloc((void**)&f->soft, 2 * (size_t)cap * sizeof(int4)));
        f->soft_cap = cap;
    }
    set_error("fep: soft-core pair list still overflowing");
    return SPONGE_NB_ECAPACITY;
}

// Window parameters of the soft-core terms (LJ_soft_core.cu:1563-1568) in FP64.
void window_params(const sponge_lj* lj, double lam, double* w) {
    const double l_ = 1.0 - lam, a = lj->fep_alpha, p = lj->fep_p;
    w[0] = lam;
    w[1] = l_;
    w[2] = a * std::pow(lam, p);
    w[3] = a * std::pow(l_, p);
    w[4] = p * a * std::pow(lam, p - 1.0);
    w[5] = p * a * std::pow(l_, p - 1.0);
    w[6] = lam - lj->fep_lambda;
    w[7] = 0.0;
}

int soft_windows(sponge_lj* lj, const sponge_nl* nl, const unsigned* uc, const float* q, const float* dq,
                 float beta, int nwin, const double* lams, double* out, hipStream_t s) {
    SNB_ARG(lj && nl && uc && q && out && lams, "lj_soft windows");
    SNB_ARG(lj->d_fe, "not a soft-core handle (sponge_lj_soft_create)");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    SNB_ARG(nwin >= 1 && nwin <= SPONGE_FEP_MAX_WINDOWS, "1 <= nwin <= SPONGE_FEP_MAX_WINDOWS");
    FepScratch* f = nullptr;
    int rc = fep_scratch(&lj->fep, lj->n, &f);
    if (rc) return rc;
    std::vector<double> w((size_t)nwin * kWinParams);
    for (int k = 0; k < nwin; k++) {
        SNB_ARG(lams[k] >= 0.0 && lams[k] <= 1.0, "window lambda must be in [0, 1]");
        window_params(lj, lams[k], &w[(size_t)k * kWinParams]);
    }
    WalkArgs A{nl, uc, q, dq, lj->d_type, lj->d_fe, lj->ntypes, nullptr, lj->cutoff,
               {lj->cof_dr[0], lj->cof_dr[1], lj->cof_dr[2]}, (double)beta, lj->fep_sig6, lj->fep_sig6min};
    double h[kHard];
    std::vector<double> so((size_t)nwin * kSoftFields);
    rc = fep_walk(f, A, nwin, w.data(), h, so.data(), s);
    if (rc) return rc;
    const double V = lj->volume;
    for (int k = 0; k < nwin; k++) {
        const double lam = lams[k], l_ = 1.0 - lam, t = lam - lj->fep_lambda;
        const double* S = &so[(size_t)k * kSoftFields];
        double* o = out + (size_t)k * SPONGE_FEP_NFIELDS;
        o[SPONGE_FEP_VDW_INTERSYS] = l_ * h[H_VA_INTER] + lam * h[H_VB_INTER] + S[0];
        o[SPONGE_FEP_VDW_INTRASYS] = l_ * h[H_VA_INTRA] + lam * h[H_VB_INTRA] + S[1];
        o[SPONGE_FEP_COUL_DIRECT_INTERSYS] = h[H_Q0_INTER] + t * h[H_Q1_INTER] + t * t * h[H_Q2_INTER] + S[2];
        o[SPONGE_FEP_COUL_DIRECT_INTRASYS] = h[H_Q0_INTRA] + t * h[H_Q1_INTRA] + t * t * h[H_Q2_INTRA] + S[3];
        o[SPONGE_FEP_DHDL_LJ] = (h[H_VB_INTRA] - h[H_VA_INTRA]) + (h[H_VB_INTER] - h[H_VA_INTER]) + S[4];
        o[SPONGE_FEP_DHDL_DIRECT] = h[H_R1] + 2.0 * t * h[H_R2] + S[5];
        o[SPONGE_FEP_DHDL_DIRECT_EXACT] =
            h[H_Q1_INTRA] + h[H_Q1_INTER] + 2.0 * t * (h[H_Q2_INTRA] + h[H_Q2_INTER]) + S[5];
        o[SPONGE_FEP_VDW_LRC] = (l_ * lj->lrc_A + lam * lj->lrc_B) / V;
        o[SPONGE_FEP_DHDL_LRC] = (lj->lrc_B - lj->lrc_A) / V;
    }
    return SPONGE_NB_OK;
}

// Cross sums of sponge_pme_cross_dhdl: the reciprocal pipeline on q (R =
// theta * Q), the interpolated potential at every atom, then the per-atom
// FP64 sums. Returns {recip, self, corr} in x[0..2].
int pme_cross(sponge_pme* pme, FepScratch* f, const unsigned* uc, const float* q, const float* dq,
              const float* scaler, const int* es, const int* el, const int* en, double* x, hipStream_t s) {
    const int n = pme->n;
    if (!f->phi) SNB_HIP(hipMalloc((void**)&f->phi, n * sizeof(float)));
    int rc = snb_pme_reciprocal(pme, uc, q, nullptr, 0, 0, s);
    if (rc) return rc;
    fep_fill_kernel<<<ceil_div(n, 256), 256, 0, s>>>(f->phi, n);
    SNB_CHECK_LAUNCH();
    rc = snb_pme_adj_gather(pme, 0, nullptr, 1.0f, f->phi, s);
    if (rc) return rc;
    fep_zero_kernel<<<1, 64, 0, s>>>(f->acc, kCross, f->count);
    SNB_CHECK_LAUNCH();
    fep_cross_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, q, dq, f->phi, uc, (double)scaler[0], (double)scaler[1],
                                                       (double)scaler[2], (double)pme->beta, es, el, en, f->acc);
    SNB_CHECK_LAUNCH();
    double h[kCross];
    SNB_HIP(hipMemcpyAsync(h, f->acc, sizeof(h), hipMemcpyDeviceToHost, s));
    SNB_HIP(hipStreamSynchronize(s));
    x[0] = h[X_RECIP];
    x[1] = -2.0 * (double)pme->beta / std::sqrt(M_PI) * h[X_QDQ] + 2.0 * pme->neutralizing_factor * h[X_Q] * h[X_DQ];
    x[2] = h[X_CORR];
    return SPONGE_NB_OK;
}

// {reciprocal, self, correction} energies of one charge set.
int pme_parts(sponge_pme* pme, const unsigned* uc, const float* q, const float* scaler, const int* es,
              const int* el, const int* en, double* e, hipStream_t s) {
    int rc = snb_pme_reciprocal(pme, uc, q, nullptr, 1, 0, s);
    if (rc) return rc;
    rc = snb_excluded(pme, uc, scaler, q, es, el, en, nullptr, nullptr, s);
    if (rc) return rc;
    rc = snb_pme_finalize(pme, nullptr, 0, 0, nullptr, nullptr, s);
    if (rc) return rc;
    double acc[SPONGE_NB_NPARTS];
    SNB_HIP(hipMemcpyAsync(acc, pme->acc, sizeof(acc), hipMemcpyDeviceToHost, s));
    SNB_HIP(hipStreamSynchronize(s));
    e[0] = acc[SPONGE_PART_RECIPROCAL];
    e[1] = acc[SPONGE_PART_SELF];
    e[2] = acc[SPONGE_PART_CORRECTION];
    return SPONGE_NB_OK;
}

}  // namespace

extern "C" int sponge_lj_soft_energy_with_coulomb_direct(sponge_lj* lj, const sponge_nl* nl,
                                                         const unsigned int* d_uint_crd, const float* d_charge,
                                                         float pme_beta, double h_energy[4], void* stream) {
    SNB_ARG(h_energy, "h_energy");
    double o[SPONGE_FEP_NFIELDS];
    const double lam = lj ? lj->fep_lambda : 0.0;
    int rc = soft_windows(lj, nl, d_uint_crd, d_charge, nullptr, pme_beta, 1, &lam, o, as_stream(stream));
    if (rc) return rc;
    for (int k = 0; k < 4; k++) h_energy[k] = o[k];
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_soft_dhdl_with_coulomb_direct(sponge_lj* lj, const sponge_nl* nl,
                                                       const unsigned int* d_uint_crd, const float* d_charge,
                                                       const float* d_charge_B_A, int charge_perturbed,
                                                       float pme_beta, double h_dhdl[2], void* stream) {
    SNB_ARG(h_dhdl, "h_dhdl");
    SNB_ARG(!charge_perturbed || d_charge_B_A, "charge_perturbed needs d_charge_B_A");
    double o[SPONGE_FEP_NFIELDS];
    const double lam = lj ? lj->fep_lambda : 0.0;
    int rc = soft_windows(lj, nl, d_uint_crd, d_charge, charge_perturbed ? d_charge_B_A : nullptr, pme_beta, 1,
                          &lam, o, as_stream(stream));
    if (rc) return rc;
    h_dhdl[0] = o[SPONGE_FEP_DHDL_LJ];
    h_dhdl[1] = o[SPONGE_FEP_DHDL_DIRECT];
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_soft_long_range_correction(const sponge_lj* lj, double* h_lrc, double* h_dhdl_lrc) {
    SNB_ARG(lj && lj->d_fe, "not a soft-core handle (sponge_lj_soft_create)");
    const double lam = lj->fep_lambda;
    if (h_lrc) *h_lrc = ((1.0 - lam) * lj->lrc_A + lam * lj->lrc_B) / lj->volume;
    if (h_dhdl_lrc) *h_dhdl_lrc = (lj->lrc_B - lj->lrc_A) / lj->volume;
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_soft_lambda_windows(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                             const float* d_charge, const float* d_charge_B_A, float pme_beta,
                                             int nwin, const double* h_lambda, double* h_out, void* stream) {
    return soft_windows(lj, nl, d_uint_crd, d_charge, d_charge_B_A, pme_beta, nwin, h_lambda, h_out,
                        as_stream(stream));
}

extern "C" int sponge_pme_cross_dhdl(sponge_pme* pme, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                     const float* d_charge, const float* d_charge_B_A, const float scaler[3],
                                     const int* d_excluded_list_start, const int* d_excluded_list,
                                     const int* d_excluded_numbers, int lj_perturbed, double h_out[4],
                                     void* stream) {
    SNB_ARG(pme && d_uint_crd && d_charge && d_charge_B_A && scaler && h_out, "pme_cross_dhdl");
    SNB_ARG(d_excluded_list_start && d_excluded_list && d_excluded_numbers, "exclusion CSR");
    SNB_ARG(lj_perturbed || (nl && nl->n == pme->n), "neighbor list");
    hipStream_t s = as_stream(stream);
    FepScratch* f = nullptr;
    int rc = fep_scratch(&pme->fep, pme->n, &f);
    if (rc) return rc;
    double x[3];
    rc = pme_cross(pme, f, d_uint_crd, d_charge, d_charge_B_A, scaler, d_excluded_list_start, d_excluded_list,
                   d_excluded_numbers, x, s);
    if (rc) return rc;
    double direct = 0.0;
    if (!lj_perturbed) {
        float cof[3];
        for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * pme->box[d];
        WalkArgs A{nl, d_uint_crd, d_charge, d_charge_B_A, nullptr, nullptr, 1, nullptr, pme->cutoff,
                   {cof[0], cof[1], cof[2]}, (double)pme->beta, 0.0, 0.0};
        double h[kHard];
        rc = fep_walk(f, A, 0, nullptr, h, nullptr, s);
        if (rc) return rc;
        direct = h[H_Q1_INTRA] + h[H_Q1_INTER];
    }
    h_out[0] = x[0];
    h_out[1] = x[1];
    h_out[2] = x[2];
    h_out[3] = direct;
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_lambda_windows(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge,
                                         const float* d_charge_B_A, const float scaler[3],
                                         const int* d_excluded_list_start, const int* d_excluded_list,
                                         const int* d_excluded_numbers, float lambda0, int nwin,
                                         const double* h_lambda, double* h_out, void* stream) {
    SNB_ARG(pme && d_uint_crd && d_charge && scaler && h_lambda && h_out, "pme_lambda_windows");
    SNB_ARG(d_excluded_list_start && d_excluded_list && d_excluded_numbers, "exclusion CSR");
    SNB_ARG(nwin >= 1 && nwin <= SPONGE_FEP_MAX_WINDOWS, "1 <= nwin <= SPONGE_FEP_MAX_WINDOWS");
    hipStream_t s = as_stream(stream);
    FepScratch* f = nullptr;
    int rc = fep_scratch(&pme->fep, pme->n, &f);
    if (rc) return rc;
    double c0[3], c1[3] = {0.0, 0.0, 0.0}, c2[3] = {0.0, 0.0, 0.0};
    if (d_charge_B_A) {
        rc = pme_cross(pme, f, d_uint_crd, d_charge, d_charge_B_A, scaler, d_excluded_list_start, d_excluded_list,
                       d_excluded_numbers, c1, s);
        if (rc) return rc;
        rc = pme_parts(pme, d_uint_crd, d_charge_B_A, scaler, d_excluded_list_start, d_excluded_list,
                       d_excluded_numbers, c2, s);
        if (rc) return rc;
    }
    rc = pme_parts(pme, d_uint_crd, d_charge, scaler, d_excluded_list_start, d_excluded_list, d_excluded_numbers, c0,
                   s);
    if (rc) return rc;
    for (int k = 0; k < nwin; k++) {
        const double t = h_lambda[k] - (double)lambda0;
        for (int p = 0; p < 3; p++) h_out[(size_t)k * 3 + p] = c0[p] + t * c1[p] + t * t * c2[p];
    }
    return SPONGE_NB_OK;
}

extern "C" int sponge_fep_separate_direct_energy(sponge_pme* pme, const sponge_nl* nl,
                                                 const unsigned int* d_uint_crd, const float* d_charge,
                                                 const int* d_subsys_division, double h_out[2], void* stream) {
    SNB_ARG(pme && nl && d_uint_crd && d_charge && h_out, "fep_separate_direct_energy");
    SNB_ARG(nl->n == pme->n, "atom count of PME and neighbor list differ");
    hipStream_t s = as_stream(stream);
    FepScratch* f = nullptr;
    int rc = fep_scratch(&pme->fep, pme->n, &f);
    if (rc) return rc;
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * pme->box[d];
    WalkArgs A{nl, d_uint_crd, d_charge, nullptr, nullptr, nullptr, 1, d_subsys_division, pme->cutoff,
               {cof[0], cof[1], cof[2]}, (double)pme->beta, 0.0, 0.0};
    double h[kHard];
    rc = fep_walk(f, A, 0, nullptr, h, nullptr, s);
    if (rc) return rc;
    h_out[0] = h[H_Q0_INTER];
    h_out[1] = h[H_Q0_INTRA];
    return SPONGE_NB_OK;
}
