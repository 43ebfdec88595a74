// Adjoint (gradient) of the nonbonded path with respect to per-atom charges
// and the LJ pair tables, for the total energy and for a force-matching loss
// L = sum_i |F_i - T_i|^2. The reference has no adjoint; this follows the
// analytic derivatives of the same formulas (LJ + erfc direct pairs over the
// cluster list with the forward cutoff predicate, the excluded-pair
// correction, smooth PME with the same spread / FFT / convolution kernels,
// self and neutralising terms, long-range dispersion). Pair terms are
// recomputed, nothing is stored between forward and backward.
//
// With w_i = 2 (F_i - T_i) and d = r_j - r_i:
//   dL/dq_i = sum_j -q_j g(r) d.(w_i - w_j) + (exclusions with g - 1/r^3)
//             - sum_g R[g] w_i.grad M_i(g) - sum_g (theta*D)[g] M_i(g),
//   D[g] = sum_i q_i w_i.grad M_i(g),  g(r) = (2b/sqrt(pi) r e^{-b^2r^2} + erfc(b r))/r^3,
//   dL/dA_p = sum_{ordered pairs of type p} -r^-14 d.w_i,  dL/dB_p = sum r^-8 d.w_i.
// The list is full (both directions), so each i-cluster wave owns the whole
// charge gradient of its atoms: no atomics on the per-atom gradient.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "pair_math.hpp"

using namespace snb;

namespace {

constexpr int kAdjWaves = 4;
constexpr int kAdjMaxTypes = 8;  // per-lane accumulators of the table gradients

struct AdjScratch {
    int n, slot_cap, ntypes;
    float* gq;          // n, per-atom charge gradient (FP32 accumulation)
    float* w3;          // 3n, w = 2 (F - T)
    float4* wv;         // n, w scaled by K/L per axis (spread / gather of the reciprocal adjoint)
    float4* ws;         // slot_cap, w in slot order (pair adjoint)
    double* red;        // [0] loss, [1] sum q
    unsigned* bound;    // max_i |q_i| sum_a |wv_ia| as FP32 bits
    float* scale;       // fixed-point unit of the adjoint spread
    double* lrc_count;  // P: ordered atom pairs per pair type
};

__global__ void adj_clear_kernel(double* red, unsigned* bound, double* gA, double* gB, int P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2) red[t] = 0.0;
    if (t == 0) bound[0] = 0u;
    if (t < P) {
        gA[t] = 0.0;
        gB[t] = 0.0;
    }
}

// w = 2 (F - T), its K/L-scaled copy for the reciprocal adjoint, the loss,
// the bound of one adjoint spread contribution and sum q.
__global__ __launch_bounds__(256) void adj_residual_kernel(int n, const float* __restrict__ F,
                                                           const float* __restrict__ T, const float* __restrict__ q,
                                                           float kx, float ky, float kz, float* __restrict__ w3,
                                                           float4* __restrict__ wv, double* __restrict__ red,
                                                           unsigned* __restrict__ bound) {
    __shared__ double sm[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    float bnd = 0.f;
    if (i < n) {
        float e[3];
        for (int d = 0; d < 3; d++) {
            e[d] = F[3 * i + d] - T[3 * i + d];
            l += (double)e[d] * e[d];
            w3[3 * i + d] = 2.0f * e[d];
        }
        const float4 v = make_float4(2.0f * e[0] * kx, 2.0f * e[1] * ky, 2.0f * e[2] * kz, 0.f);
        wv[i] = v;
        bnd = fabsf(q[i]) * (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
    }
    const double t = block_sum<256>(l, sm);
    if (threadIdx.x == 0) atomicAdd(red, t);
    // wave max, then one atomic per wave (positive floats order like their bits)
    for (int o = 32; o > 0; o >>= 1) bnd = fmaxf(bnd, __shfl_xor(bnd, o));
    if ((threadIdx.x & 63) == 0) atomicMax(bound, __float_as_uint(bnd));
}

// One adjoint-spread contribution is at most bound * (2/3)^3 (every B-spline
// weight and derivative of order 4 is at most 2/3): keep it below 2^30 units.
__global__ void adj_scale_kernel(const unsigned* bound, float* scale) {
    const float b = __uint_as_float(bound[0]) * (8.0f / 27.0f);
    scale[0] = b > 0.f ? 1073741824.0f / b : 1.0f;
}

__global__ __launch_bounds__(256) void adj_sumq_kernel(int n, const float* __restrict__ q, double* __restrict__ red) {
    __shared__ double sm[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double t = block_sum<256>(i < n ? (double)q[i] : 0.0, sm);
    if (threadIdx.x == 0) atomicAdd(red + 1, t);
}

// Slot-order records (as the forward gather: type in the low uint bits) and
// w in slot order.
__global__ void adj_gather_kernel(const int* __restrict__ flags, const int* __restrict__ s2o,
                                  const unsigned int* __restrict__ uc, const float* __restrict__ q,
                                  const int* __restrict__ type, uint4* __restrict__ pk, const float* __restrict__ w3,
                                  float4* __restrict__ ws) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= flags[3]) return;
    const int a = s2o[p];
    if (a >= 0) {
        const unsigned t = (unsigned)type[a];
        pk[p] = make_uint4((uc[3 * a] & ~3u) | (t & 3u), (uc[3 * a + 1] & ~3u) | ((t >> 2) & 3u),
                           (uc[3 * a + 2] & ~3u) | ((t >> 4) & 3u), __float_as_uint(q[a]));
        if (ws) ws[p] = make_float4(w3[3 * a], w3[3 * a + 1], w3[3 * a + 2], 0.f);
    } else {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
        if (ws) ws[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ int pidx(int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// One wave per i-cluster over the full cluster-pair list (the forward
// kernel's layout and cutoff predicate). LOSS = false: energy gradient
// (potential of atom i; half of every ordered pair's table term, since the
// full list holds each unordered pair twice). LOSS = true: force-matching
// gradient. gq[atom] is assigned; the table gradients are reduced per
// (type i, type j) over the wave and added to gA / gB in FP64.
template <bool LOSS>
__global__ __launch_bounds__(kAdjWaves * 64) void adj_pair_kernel(
    const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list, int list_cap,
    const uint4* __restrict__ pk, const float4* __restrict__ ws, int nt, float cx, float cy, float cz, float rc2,
    float beta, const int* __restrict__ s2o, float* __restrict__ gq, double* __restrict__ gA,
    double* __restrict__ gB) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lb = xcd_block(gridDim.x);
    const int ic = lb * kAdjWaves + w;
    if (lb < 0 || ic >= flags[4]) return;
    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const int ti = unpack_type(a);
    float4 wi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (LOSS) wi = ws[islot];
    const float mlog2e_b2 = -1.4426950408889634f * (beta * beta);
    // per-lane sums in FP64: single terms are FP32, but the sums over a
    // lane's pairs cancel (d.w changes sign) and near pairs dominate
    double sq = 0.0;
    double ta[kAdjMaxTypes], tb[kAdjMaxTypes];
#pragma unroll
    for (int t = 0; t < kAdjMaxTypes; t++) ta[t] = tb[t] = 0.0;
    const int n = list_n[ic];
    const uint2* lst = list + (size_t)ic * list_cap;
    for (int base = 0; base < n; base += 64) {
        const uint2 my = base + lane < n ? lst[base + lane] : make_uint2((unsigned)ic, 0u);
        const int cnt = min(64, n - base);
        for (int it = 0; it < cnt; it += 4) {
            const int src = min(it + g, 63);
            const int jc = __shfl((int)my.x, src);
            const unsigned m = (unsigned)__shfl((int)my.y, src);
            const int jslot = jc * kClusterSize + jj;
            const uint4 b = pk[jslot];
            const float dx = (float)(int)((b.x & ~3u) - ax) * cx;
            const float dy = (float)(int)((b.y & ~3u) - ay) * cy;
            const float dz = (float)(int)((b.z & ~3u) - az) * cz;
            const float r2 = dr2_exact(dx, dy, dz);
            const bool in = (it + g < cnt) && ((m >> bit) & 1u) && r2 < rc2;
            const float r2s = in ? r2 : 1.0f;
            const float rinv = __builtin_amdgcn_rsqf(r2s);
            const float rinv2 = rinv * rinv;
            const float x = beta * (r2s * rinv);
            const float ex = __builtin_amdgcn_exp2f(mlog2e_b2 * r2s);
            const float ecx = erfcx_poly(x);
            const float qj = __uint_as_float(b.w);
            const int tj = unpack_type(b);
            // table terms in FP64 from the exact integer displacement: r^-14
            // amplifies the FP32 rounding of r^2 seven-fold
            double cA = 0.0, cB = 0.0;
            if (in) {
                const double ddx = (double)(int)((b.x & ~3u) - ax) * (double)cx;
                const double ddy = (double)(int)((b.y & ~3u) - ay) * (double)cy;
                const double ddz = (double)(int)((b.z & ~3u) - az) * (double)cz;
                const double ir2 = 1.0 / (ddx * ddx + ddy * ddy + ddz * ddz);
                const double ir6 = ir2 * ir2 * ir2;
                if (!LOSS) {
                    sq += (double)(qj * ex * ecx * rinv);
                    cA = ir6 * ir6 * (0.5 / 12.0);
                    cB = ir6 * (-0.5 / 6.0);
                } else {
                    const float4 wj = ws[jslot];
                    const float sdi = dx * wi.x + dy * wi.y + dz * wi.z;
                    const float sdj = dx * wj.x + dy * wj.y + dz * wj.z;
                    const float gr = ex * fmaf(kTwoOverSqrtPi, x, ecx) * rinv * rinv2;
                    sq += (double)(-qj * gr * (sdi - sdj));
                    const double sdid = ddx * (double)wi.x + ddy * (double)wi.y + ddz * (double)wi.z;
                    cA = -ir6 * ir6 * ir2 * sdid;
                    cB = ir6 * ir2 * sdid;
                }
            }
#pragma unroll
            for (int t = 0; t < kAdjMaxTypes; t++) {
                ta[t] += tj == t ? cA : 0.0;
                tb[t] += tj == t ? cB : 0.0;
            }
        }
    }
    // sum the 16 lanes sharing ii
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
        sq += __shfl_xor(sq, o);
#pragma unroll
        for (int t = 0; t < kAdjMaxTypes; t++) {
            ta[t] += __shfl_xor(ta[t], o);
            tb[t] += __shfl_xor(tb[t], o);
        }
    }
    if (lane < 4) {
        const int at = s2o[islot];
        if (at >= 0) {
            gq[at] = (float)sq;
#pragma unroll
            for (int t = 0; t < kAdjMaxTypes; t++) {
                if (t < nt && (ta[t] != 0.0 || tb[t] != 0.0)) {
                    atomicAdd(gA + pidx(ti, t), ta[t]);
                    atomicAdd(gB + pidx(ti, t), tb[t]);
                }
            }
        }
    }
}

// Excluded pairs (PME_force.cu:620-695 restated for the gradient): one
// thread per atom over its CSR row, atomics on both ends.
template <bool LOSS>
__global__ __launch_bounds__(256) void adj_excl_kernel(int n, const unsigned int* __restrict__ uc, float sx, float sy,
                                                       float sz, const float* __restrict__ q, float beta,
                                                       const int* __restrict__ es, const int* __restrict__ el,
                                                       const int* __restrict__ en, const float* __restrict__ w3,
                                                       float* __restrict__ gq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ne = en[i];
    if (ne <= 0) return;
    const int s = es[i];
    const float qi = q[i];
    const unsigned ux = uc[3 * i], uy = uc[3 * i + 1], uz = uc[3 * i + 2];
    float gi = 0.f;
    for (int k = s; k < s + ne; k++) {
        const int j = el[k];
        const float dx = (float)(int)(uc[3 * j] - ux) * sx;
        const float dy = (float)(int)(uc[3 * j + 1] - uy) * sy;
        const float dz = (float)(int)(uc[3 * j + 2] - uz) * sz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float r = sqrtf(r2);
        const float br = beta * r;
        const float qj = q[j];
        if (!LOSS) {
            const float pot = erff(br) / r;
            gi -= qj * pot;
            atomicAdd(gq + j, -qi * pot);
        } else {
            const float h = (kTwoOverSqrtPi * br * expf(-br * br) + erfcf(br) - 1.0f) / r2 / r;
            const float sd = dx * (w3[3 * i] - w3[3 * j]) + dy * (w3[3 * i + 1] - w3[3 * j + 1]) +
                             dz * (w3[3 * i + 2] - w3[3 * j + 2]);
            gi += -qj * h * sd;
            atomicAdd(gq + j, -qi * h * sd);
        }
    }
    atomicAdd(gq + i, gi);
}

// Final per-atom gradient in FP64: energy mode adds the self and
// neutralising derivatives -2b/sqrt(pi) q_i + 2 nf sum q.
__global__ void adj_finish_kernel(int n, const float* __restrict__ gq, const float* __restrict__ q, int energy,
                                  double beta, double nf, const double* __restrict__ red, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = gq[i];
    if (energy) v += -2.0 * beta / sqrt(M_PI) * (double)q[i] + 2.0 * nf * red[1];
    out[i] = v;
}

__global__ void adj_lrc_kernel(int P, const double* __restrict__ count, double c, double* __restrict__ gB,
                               const double* __restrict__ red, double* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (count && p < P) gB[p] += c * count[p];
    if (loss && p == 0) loss[0] = red[0];
}

AdjScratch* adj_get(sponge_lj* lj, const sponge_nl* nl) {
    AdjScratch* A = static_cast<AdjScratch*>(lj->adj);
    if (A && A->slot_cap >= nl->slot_cap) return A;
    if (A) snb_adj_free(lj);
    A = new AdjScratch();
    memset(A, 0, sizeof(*A));
    A->n = lj->n;
    A->slot_cap = nl->slot_cap;
    A->ntypes = lj->ntypes;
    const int T = lj->ntypes, P = T * (T + 1) / 2;
    std::vector<double> cnt(T, 0.0), pc(P, 0.0);
    for (int i = 0; i < lj->n; i++) cnt[lj->h_type[i]] += 1;
    for (int x = 0; x < T; x++)
        for (int y = 0; y < T; y++) {
            const int hi = x > y ? x : y, lo = x > y ? y : x;
            pc[hi * (hi + 1) / 2 + lo] += cnt[x] * cnt[y];
        }
    hipError_t e = hipSuccess;
    auto M = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    M((void**)&A->gq, lj->n * sizeof(float));
    M((void**)&A->w3, 3 * (size_t)lj->n * sizeof(float));
    M((void**)&A->wv, lj->n * sizeof(float4));
    M((void**)&A->ws, nl->slot_cap * sizeof(float4));
    M((void**)&A->red, 2 * sizeof(double));
    M((void**)&A->bound, sizeof(unsigned));
    M((void**)&A->scale, sizeof(float));
    M((void**)&A->lrc_count, P * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(A->lrc_count, pc.data(), P * sizeof(double), hipMemcpyHostToDevice);
    lj->adj = A;
    if (e != hipSuccess) {
        set_error("adjoint scratch: %s", hipGetErrorString(e));
        snb_adj_free(lj);
        return nullptr;
    }
    return A;
}

int adj_common(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const unsigned int* uc, const float* q,
               const int* es, const int* el, const int* en, int loss, AdjScratch* A, double* gA, double* gB,
               hipStream_t s) {
    const int n = lj->n;
    const int T = lj->ntypes, P = T * (T + 1) / 2;
    int rc = snb_lj_scratch(lj, nl);
    if (rc) return rc;
    adj_gather_kernel<<<ceil_div(nl->slot_cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, uc, q, lj->d_type, lj->ps.pk,
                                                                   A->w3, loss ? A->ws : nullptr);
    SNB_CHECK_LAUNCH();
    const int nblk = xcd_grid(ceil_div(nl->ncl_cap, kAdjWaves));
    const float rc2 = lj->cutoff * lj->cutoff;
#define SNB_ADJ(L)                                                                                               \
    adj_pair_kernel<L><<<nblk, kAdjWaves * 64, 0, s>>>(nl->flags, nl->list_n, nl->list, nl->list_cap, lj->ps.pk, \
                                                       A->ws, T, lj->cof_dr[0], lj->cof_dr[1], lj->cof_dr[2], rc2, \
                                                       pme->beta, nl->s2o, A->gq, gA, gB)
    if (loss) SNB_ADJ(true);
    else SNB_ADJ(false);
#undef SNB_ADJ
    SNB_CHECK_LAUNCH();
    if (loss)
        adj_excl_kernel<true><<<ceil_div(n, 256), 256, 0, s>>>(n, uc, lj->cof_dr[0], lj->cof_dr[1], lj->cof_dr[2], q,
                                                               pme->beta, es, el, en, A->w3, A->gq);
    else
        adj_excl_kernel<false><<<ceil_div(n, 256), 256, 0, s>>>(n, uc, lj->cof_dr[0], lj->cof_dr[1], lj->cof_dr[2],
                                                                q, pme->beta, es, el, en, A->w3, A->gq);
    SNB_CHECK_LAUNCH();
    (void)P;
    // reciprocal: R = theta * Q (prep, spread, FFTs, convolution; no gather)
    return snb_pme_reciprocal(pme, uc, q, nullptr, 0, 0, s);
}

int adj_check(const sponge_nl* nl, const sponge_lj* lj, const sponge_pme* pme) {
    SNB_ARG(nl && lj && pme, "gradient: null handle");
    SNB_ARG(nl->n == lj->n && nl->n == pme->n, "gradient: atom counts of the modules differ");
    SNB_ARG(!lj->d_sc, "gradient: soft-core handles are not supported");
    SNB_ARG(lj->ntypes <= kAdjMaxTypes, "gradient: at most 8 LJ types");
    return SPONGE_NB_OK;
}

}  // namespace

void snb_adj_free(sponge_lj* lj) {
    AdjScratch* A = static_cast<AdjScratch*>(lj->adj);
    if (!A) return;
    void* ptrs[] = {A->gq, A->w3, A->wv, A->ws, A->red, A->bound, A->scale, A->lrc_count};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete A;
    lj->adj = nullptr;
}

extern "C" int sponge_nb_energy_gradient(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const unsigned int* d_uint_crd,
                                         const float* d_charge, const int* d_excluded_list_start,
                                         const int* d_excluded_list, const int* d_excluded_numbers,
                                         double* d_grad_charge, double* d_grad_A, double* d_grad_B, void* stream) {
    int rc = adj_check(nl, lj, pme);
    if (rc) return rc;
    SNB_ARG(d_uint_crd && d_charge && d_excluded_list_start && d_excluded_list && d_excluded_numbers &&
                d_grad_charge && d_grad_A && d_grad_B,
            "energy_gradient: null pointer");
    hipStream_t s = as_stream(stream);
    AdjScratch* A = adj_get(lj, nl);
    if (!A) return SPONGE_NB_EHIP;
    const int n = lj->n, T = lj->ntypes, P = T * (T + 1) / 2;
    adj_clear_kernel<<<ceil_div(std::max(P, 2), 256), 256, 0, s>>>(A->red, A->bound, d_grad_A, d_grad_B, P);
    SNB_CHECK_LAUNCH();
    rc = adj_common(nl, lj, pme, d_uint_crd, d_charge, d_excluded_list_start, d_excluded_list, d_excluded_numbers, 0,
                    A, d_grad_A, d_grad_B, s);
    if (rc) return rc;
    rc = snb_pme_adj_gather(pme, 0, nullptr, 1.0f, A->gq, s);
    if (rc) return rc;
    adj_sumq_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_charge, A->red);
    SNB_CHECK_LAUNCH();
    adj_finish_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, A->gq, d_charge, 1, (double)pme->beta,
                                                       pme->neutralizing_factor, A->red, d_grad_charge);
    SNB_CHECK_LAUNCH();
    const double c = (-2.0 / 3.0) * M_PI / ((double)lj->cutoff * lj->cutoff * lj->cutoff) / 6.0 / lj->volume;
    adj_lrc_kernel<<<ceil_div(P, 256), 256, 0, s>>>(P, A->lrc_count, c, d_grad_B, A->red, nullptr);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_nb_force_matching_gradient(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme,
                                                 const unsigned int* d_uint_crd, const float* d_charge,
                                                 const int* d_excluded_list_start, const int* d_excluded_list,
                                                 const int* d_excluded_numbers, const float* d_frc,
                                                 const float* d_target_frc, double* d_loss, double* d_grad_charge,
                                                 double* d_grad_A, double* d_grad_B, void* stream) {
    int rc = adj_check(nl, lj, pme);
    if (rc) return rc;
    SNB_ARG(d_uint_crd && d_charge && d_excluded_list_start && d_excluded_list && d_excluded_numbers && d_frc &&
                d_target_frc && d_grad_charge && d_grad_A && d_grad_B,
            "force_matching_gradient: null pointer");
    hipStream_t s = as_stream(stream);
    AdjScratch* A = adj_get(lj, nl);
    if (!A) return SPONGE_NB_EHIP;
    const int n = lj->n, T = lj->ntypes, P = T * (T + 1) / 2;
    adj_clear_kernel<<<ceil_div(std::max(P, 2), 256), 256, 0, s>>>(A->red, A->bound, d_grad_A, d_grad_B, P);
    SNB_CHECK_LAUNCH();
    adj_residual_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_frc, d_target_frc, d_charge,
                                                         (float)pme->K[0] / pme->box[0], (float)pme->K[1] / pme->box[1],
                                                         (float)pme->K[2] / pme->box[2], A->w3, A->wv, A->red,
                                                         A->bound);
    SNB_CHECK_LAUNCH();
    adj_scale_kernel<<<1, 1, 0, s>>>(A->bound, A->scale);
    SNB_CHECK_LAUNCH();
    rc = adj_common(nl, lj, pme, d_uint_crd, d_charge, d_excluded_list_start, d_excluded_list, d_excluded_numbers, 1,
                    A, d_grad_A, d_grad_B, s);
    if (rc) return rc;
    // (a) explicit charge in the adjoint source: -sum_g R[g] w_i.grad M_i(g)
    rc = snb_pme_adj_gather(pme, 1, A->wv, -1.0f, A->gq, s);
    if (rc) return rc;
    // (b) through the charge grid: -sum_g (theta * D)[g] M_i(g)
    rc = snb_pme_prep(pme, d_uint_crd, d_charge, s);
    if (rc) return rc;
    rc = snb_pme_adjoint_grid(pme, A->wv, A->scale, s);
    if (rc) return rc;
    rc = snb_pme_adj_gather(pme, 0, nullptr, -1.0f, A->gq, s);
    if (rc) return rc;
    adj_finish_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, A->gq, d_charge, 0, 0.0, 0.0, A->red, d_grad_charge);
    SNB_CHECK_LAUNCH();
    adj_lrc_kernel<<<1, 64, 0, s>>>(P, nullptr, 0.0, d_grad_B, A->red, d_loss);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}
