// LJ + PME direct-space pair force for gfx950 (LENNARD_JONES_INFORMATION,
// Lennard_Jones_force/Lennard_Jones_force.cu) and the excluded-pair
// correction (Particle_Mesh_Ewald::PME_Excluded_Force_With_Atom_Energy,
// PME_force/PME_force.cu:620-710).
//
// Pair kernel layout: one wave64 per 4-atom i-cluster; each iteration takes
// four j-clusters from the cluster-pair list, lane = (g:2, jj:2, ii:2) holds
// the pair (ii, jj) of j-cluster g. Coordinates, charges and LJ types are
// gathered once per step into cluster order as one 16-byte record per atom.
// The list names every neighbouring cluster from both sides, so each wave
// only accumulates the force on its own four atoms (in registers, reduced
// with cross-lane adds at the end) and adds it to the caller's force array
// with plain read-modify-writes: no atomics anywhere in the pair path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "pme_prep.hpp"

using namespace snb;

namespace {

constexpr int kPairWaves = 4;
constexpr float kTwoOverSqrtPi = 1.1283791670955126f;

// Per-step gather into cluster order (replaces Copy_Crd_And_Charge_To_New_Crd,
// Lennard_Jones_force.cu:24-34). The LJ type (< 64) rides in the low two bits
// of each uint coordinate, which Crd_To_Uint_Crd always leaves zero (`<< 2`,
// common.cu:283-285): one 16-byte load per j atom carries position, charge
// and type.
__global__ void gather_kernel(const int* __restrict__ flags, const int* __restrict__ s2o,
                              const unsigned int* __restrict__ uc, const float* __restrict__ q,
                              const int* __restrict__ type, uint4* __restrict__ pk, int energy,
                              double* __restrict__ acc) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (energy && p < 3) acc[p] = 0.0;  // SPONGE_PART_LJ, _DIRECT, _LJ_VIRIAL
    if (energy && p == 3) acc[SPONGE_PART_PAIRS_IN] = 0.0;
    if (p >= flags[3]) return;
    int a = s2o[p];
    if (a >= 0) {
        unsigned t = (unsigned)type[a];
        pk[p] = make_uint4((uc[3 * a] & ~3u) | (t & 3u), (uc[3 * a + 1] & ~3u) | ((t >> 2) & 3u),
                           (uc[3 * a + 2] & ~3u) | ((t >> 4) & 3u), __float_as_uint(q[a]));
    } else {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Fused per-step prep of the whole force path, one thread per cluster slot:
// optional MD_Information_Crd_To_Uint_Crd (MD_core.cu:1528-1532, written back
// to the caller's uint array), the pair-kernel record, and the PME base,
// fraction and x-plane bin of the slot's atom.
__global__ __launch_bounds__(256) void step_prep_kernel(const int* __restrict__ flags, const int* __restrict__ s2o,
                                                        const float* __restrict__ crd, float cqx, float cqy, float cqz,
                                                        unsigned int* __restrict__ uc, const float* __restrict__ q,
                                                        const int* __restrict__ type, uint4* __restrict__ pk,
                                                        PmePrep P, int energy, double* __restrict__ acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (energy && p < 3) acc[p] = 0.0;  // SPONGE_PART_LJ, _DIRECT, _LJ_VIRIAL
    if (energy && p == 3) acc[SPONGE_PART_PAIRS_IN] = 0.0;
    const int nslot = flags[3];
    int a = p < nslot ? s2o[p] : -1;
    unsigned u0 = 0u, u1 = 0u, u2 = 0u;
    float qa = 0.f;
    if (a >= 0) {
        if (crd) {
            u0 = ((unsigned)(int)__fmul_rn(crd[3 * a], cqx)) << 2;
            u1 = ((unsigned)(int)__fmul_rn(crd[3 * a + 1], cqy)) << 2;
            u2 = ((unsigned)(int)__fmul_rn(crd[3 * a + 2], cqz)) << 2;
            uc[3 * a] = u0;
            uc[3 * a + 1] = u1;
            uc[3 * a + 2] = u2;
        } else {
            u0 = uc[3 * a];
            u1 = uc[3 * a + 1];
            u2 = uc[3 * a + 2];
        }
        qa = q[a];
        const unsigned t = (unsigned)type[a];
        pk[p] = make_uint4((u0 & ~3u) | (t & 3u), (u1 & ~3u) | ((t >> 2) & 3u), (u2 & ~3u) | ((t >> 4) & 3u),
                           __float_as_uint(qa));
    } else if (p < nslot) {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
    }
    pme_prep_block(a, u0, u1, u2, qa, P, energy);
}

__device__ __forceinline__ int unpack_type(uint4 v) { return (int)((v.x & 3u) | ((v.y & 3u) << 2) | ((v.z & 3u) << 4)); }

// erfc(x) = exp(-x^2) * erfcx(x) with erfcx(x) ~ t P7(t), t = 1/(1 + p x):
// the form of Abramowitz & Stegun 7.1.26 with eight coefficients refitted
// (iteratively reweighted least squares on [0, 5]) for a relative error of
// 6.3e-8 in erfcx; evaluated in FP32 the relative error of erfc and of the
// force factor 2x/sqrt(pi) e^{-x^2} + erfc(x) stays below 5.2e-7 on [0, 4.5].
__device__ __forceinline__ float erfcx_poly(float x) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3195f, x, 1.0f));
    float p = -0.181163783781557f;
    p = fmaf(p, t, 0.7263733451877996f);
    p = fmaf(p, t, -0.8054935211563752f);
    p = fmaf(p, t, 0.8311706379959746f);
    p = fmaf(p, t, -0.18734797244699758f);
    p = fmaf(p, t, 0.27142876592060167f);
    p = fmaf(p, t, 0.16357488112143515f);
    p = fmaf(p, t, 0.18145758631968023f);
    return t * p;
}

// Cutoff test shared with the oracle: FP32 dr2 = (dx*dx + dy*dy) + dz*dz with
// every operation rounded (the neighbor list's expression) against rc*rc in
// FP32, so both sides agree on every pair, including those within one ulp of
// the cutoff.
template <bool ENERGY>
__global__ __launch_bounds__(kPairWaves * 64) void pair_kernel(
    const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list, int list_cap,
    const uint4* __restrict__ pk, const float2* __restrict__ ab, int nt, float cx, float cy, float cz, float rc2,
    float beta, const int* __restrict__ s2o, float* __restrict__ frc, float* __restrict__ atom_e,
    float* __restrict__ atom_v, float* __restrict__ atom_d, double* __restrict__ acc, int overwrite) {
    extern __shared__ float2 sab[];
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) sab[k] = ab[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lb = xcd_block(gridDim.x);
    const int ic = lb * kPairWaves + w;
    if (lb < 0 || ic >= flags[4]) return;
    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const float qi = __uint_as_float(a.w);
    const float2* srow = sab + unpack_type(a) * nt;
    const float mlog2e_b2 = -1.4426950408889634f * (beta * beta);
    int ai = 0;
    if (ENERGY) ai = s2o[islot];
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    float eie = 0.f, eiv = 0.f, eid = 0.f;
    double tlj = 0.0, tdir = 0.0, tvir = 0.0, tcnt = 0.0;
    const int n = list_n[ic];
    const uint2* lst = list + (size_t)ic * list_cap;
    for (int base = 0; base < n; base += 64) {
        // 64 list entries (16 iterations of 4 j-clusters) held one per lane.
        const uint2 my = base + lane < n ? lst[base + lane] : make_uint2((unsigned)ic, 0u);
        const int cnt = min(64, n - base);
        int jc = __shfl((int)my.x, g);
        unsigned m = (unsigned)__shfl((int)my.y, g);
        uint4 b = pk[jc * kClusterSize + jj];
        for (int it = 0; it < cnt; it += 4) {
            // prefetch the next iteration's j record
            const int src = min(it + 4 + g, 63);
            const int jcn = __shfl((int)my.x, src);
            const unsigned mn = (unsigned)__shfl((int)my.y, src);
            const uint4 bn = pk[jcn * kClusterSize + jj];
            const float dx = (float)(int)((b.x & ~3u) - ax) * cx;
            const float dy = (float)(int)((b.y & ~3u) - ay) * cy;
            const float dz = (float)(int)((b.z & ~3u) - az) * cz;
            const float r2 = dr2_exact(dx, dy, dz);
            const bool in = ((m >> bit) & 1u) && r2 < rc2;
            const float r2s = in ? r2 : 1.0f;
            const float rinv = __builtin_amdgcn_rsqf(r2s);
            const float rinv2 = rinv * rinv;
            const float rinv6 = rinv2 * rinv2 * rinv2;
            const float2 AB = srow[unpack_type(b)];
            const float flj = (AB.y - AB.x * rinv6) * rinv6 * rinv2;
            const float x = beta * (r2s * rinv);
            const float ex = __builtin_amdgcn_exp2f(mlog2e_b2 * r2s);
            const float ecx = erfcx_poly(x);
            const float qq = qi * __uint_as_float(b.w);
            const float fc = qq * ex * fmaf(kTwoOverSqrtPi, x, ecx) * rinv * rinv2;
            const float f = in ? flj - fc : 0.f;
            fix += f * dx;
            fiy += f * dy;
            fiz += f * dz;
            if (ENERGY) {
                // Energy steps evaluate the pair energies in FP64 from the exact
                // integer displacement: the totals are nets of large cancelling
                // terms, which FP32 terms cannot hold to 1e-6. Each unordered
                // pair is credited once, to its lower-index atom, as the
                // reference's half list does.
                const int aj = s2o[jc * kClusterSize + jj];
                if (in && ai < aj) {
                    const double ddx = (double)(int)((b.x & ~3u) - ax) * (double)cx;
                    const double ddy = (double)(int)((b.y & ~3u) - ay) * (double)cy;
                    const double ddz = (double)(int)((b.z & ~3u) - az) * (double)cz;
                    const double dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    const double dr = sqrt(dr2);
                    const double ir2 = 1.0 / dr2, ir6 = ir2 * ir2 * ir2;
                    const double elj = ((double)AB.x * (1.0 / 12.0) * ir6 - (double)AB.y * (1.0 / 6.0)) * ir6;
                    const double edir = (double)qi * (double)__uint_as_float(b.w) * erfc((double)beta * dr) / dr;
                    const double vlj = ((double)AB.x * ir6 - (double)AB.y) * ir6;
                    tlj += elj;
                    tdir += edir;
                    tvir += vlj;
                    tcnt += 1.0;
                    eie += (float)(elj + edir);
                    eiv += (float)vlj;
                    eid += (float)edir;
                }
            }
            b = bn;
            m = mn;
            jc = jcn;
        }
    }
    // i force: sum the 16 lanes sharing ii (lanes ii, ii+4, ..., ii+60).
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
        fix += __shfl_xor(fix, o);
        fiy += __shfl_xor(fiy, o);
        fiz += __shfl_xor(fiz, o);
        if (ENERGY) {
            eie += __shfl_xor(eie, o);
            eiv += __shfl_xor(eiv, o);
            eid += __shfl_xor(eid, o);
        }
    }
    if (lane < 4) {
        const int at = s2o[islot];
        if (at >= 0) {
            if (frc) {
                // every atom has exactly one slot, so this is the only writer
                if (overwrite) {
                    frc[3 * at] = fix;
                    frc[3 * at + 1] = fiy;
                    frc[3 * at + 2] = fiz;
                } else {
                    frc[3 * at] += fix;
                    frc[3 * at + 1] += fiy;
                    frc[3 * at + 2] += fiz;
                }
            }
            if (ENERGY) {
                if (atom_e) atom_e[at] += eie;
                if (atom_v) atom_v[at] += eiv;
                if (atom_d) atom_d[at] = eid;
            }
        }
    }
    if (ENERGY) {
        tlj = wave_sum(tlj);
        tdir = wave_sum(tdir);
        tvir = wave_sum(tvir);
        tcnt = wave_sum(tcnt);
        if (lane == 0) {
            atomicAdd(acc + SPONGE_PART_LJ, tlj);
            atomicAdd(acc + SPONGE_PART_DIRECT, tdir);
            atomicAdd(acc + SPONGE_PART_LJ_VIRIAL, tvir);
            atomicAdd(acc + SPONGE_PART_PAIRS_IN, tcnt);
        }
    }
}

__global__ void add_scalar_kernel(float* a, float va, float* b, float vb) {
    if (a) a[0] += va;
    if (b) b[0] += vb;
}

// PME_Excluded_Force_With_Atom_Energy_Correction (PME_force.cu:620-695):
// one thread per atom over its exclusion CSR row, FP32; energy
// -q_i q_j erf(b r)/r credited to atom i; FP64 total in acc.
__global__ void excluded_kernel(int n, const unsigned int* __restrict__ uc, float sx, float sy, float sz,
                                const float* __restrict__ q, float beta, const int* __restrict__ es,
                                const int* __restrict__ el, const int* __restrict__ en, float* __restrict__ frc,
                                float* __restrict__ atom_e, float* __restrict__ corr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        int ne = en[i];
        float e = 0.f;
        if (ne > 0) {
            int s = es[i];
            float qi = q[i];
            unsigned ux = uc[3 * i], uy = uc[3 * i + 1], uz = uc[3 * i + 2];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            for (int k = s; k < s + ne; k++) {
                int j = el[k];
                float dx = (float)(int)(uc[3 * j] - ux) * sx;
                float dy = (float)(int)(uc[3 * j + 1] - uy) * sy;
                float dz = (float)(int)(uc[3 * j + 2] - uz) * sz;
                float r2 = dx * dx + dy * dy + dz * dz;
                float r = sqrtf(r2);
                float br = beta * r;
                float qq = qi * q[j];
                float fa = (kTwoOverSqrtPi * br * expf(-br * br) + erfcf(br) - 1.0f) / r2 / r;
                fa = -qq * fa;
                float lx = fa * dx, ly = fa * dy, lz = fa * dz;
                fx += lx;
                fy += ly;
                fz += lz;
                if (frc) {
                    atomicAdd(frc + 3 * j, -lx);
                    atomicAdd(frc + 3 * j + 1, -ly);
                    atomicAdd(frc + 3 * j + 2, -lz);
                }
                e -= qq * erff(br) / r;
            }
            if (frc) {
                atomicAdd(frc + 3 * i, fx);
                atomicAdd(frc + 3 * i + 1, fy);
                atomicAdd(frc + 3 * i + 2, fz);
            }
            if (atom_e) atomicAdd(atom_e + i, e);
        }
        if (corr) corr[i] = e;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int sponge_lj_create(int atom_numbers, int atom_type_numbers, const float* h_LJ_A, const float* h_LJ_B,
                                const int* h_atom_LJ_type, float cutoff, const float box_length[3], sponge_lj** out) {
    SNB_ARG(out && h_LJ_A && h_LJ_B && h_atom_LJ_type && box_length, "null pointer");
    SNB_ARG(atom_numbers > 0, "atom_numbers");
    SNB_ARG(atom_type_numbers > 0 && atom_type_numbers <= kMaxLjTypes, "atom_type_numbers (1..64)");
    *out = nullptr;
    const int T = atom_type_numbers, P = T * (T + 1) / 2;
    for (int i = 0; i < atom_numbers; i++)
        SNB_ARG(h_atom_LJ_type[i] >= 0 && h_atom_LJ_type[i] < T, "atom LJ type out of range");
    sponge_lj* lj = new sponge_lj();
    memset(lj, 0, sizeof(*lj));
    lj->n = atom_numbers;
    lj->ntypes = T;
    lj->cutoff = cutoff;
    lj->h_A = new float[P];
    lj->h_B = new float[P];
    lj->h_type = new int[atom_numbers];
    memcpy(lj->h_A, h_LJ_A, P * sizeof(float));
    memcpy(lj->h_B, h_LJ_B, P * sizeof(float));
    memcpy(lj->h_type, h_atom_LJ_type, atom_numbers * sizeof(int));
    sponge_lj_update_volume(lj, box_length);
    // Long-range dispersion factor (Total_C6_Get, Lennard_Jones_force.cu:1042-1065,
    // :1127) from type counts in FP64: sum over all ordered atom pairs.
    std::vector<double> cnt(T, 0.0);
    for (int i = 0; i < atom_numbers; i++) cnt[h_atom_LJ_type[i]] += 1;
    double s = 0;
    std::vector<float2> sq((size_t)T * T);
    for (int x = 0; x < T; x++)
        for (int y = 0; y < T; y++) {
            int hi = x > y ? x : y, lo = x > y ? y : x;
            int k = hi * (hi + 1) / 2 + lo;
            s += cnt[x] * cnt[y] * (double)h_LJ_B[k];
            sq[(size_t)x * T + y] = make_float2(h_LJ_A[k], h_LJ_B[k]);
        }
    lj->lrc_factor = s * (-2.0 / 3.0) * M_PI / ((double)cutoff * cutoff * cutoff) / 6.0;
    hipError_t e = hipMalloc((void**)&lj->d_ab, sq.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(lj->d_ab, sq.data(), sq.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&lj->d_type, atom_numbers * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(lj->d_type, h_atom_LJ_type, atom_numbers * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("sponge_lj_create: %s", hipGetErrorString(e));
        sponge_lj_destroy(lj);
        return SPONGE_NB_EHIP;
    }
    *out = lj;
    return SPONGE_NB_OK;
}

extern "C" void sponge_lj_destroy(sponge_lj* lj) {
    if (!lj) return;
    void* ptrs[] = {lj->d_ab, lj->d_type};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    snb_scratch_free(&lj->ps);
    delete[] lj->h_A;
    delete[] lj->h_B;
    delete[] lj->h_type;
    delete lj;
}

extern "C" int sponge_lj_update_volume(sponge_lj* lj, const float box_length[3]) {
    SNB_ARG(lj && box_length, "lj_update_volume");
    for (int d = 0; d < 3; d++) {
        SNB_ARG(box_length[d] > 0, "box_length");
        lj->box[d] = box_length[d];
        lj->cof_dr[d] = (1.0f / 4294967296.0f) * box_length[d];
    }
    lj->volume = box_length[0] * box_length[1] * box_length[2];
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_host_tables(const sponge_lj* lj, const float** A, const float** B, const int** t) {
    SNB_ARG(lj, "lj_host_tables");
    if (A) *A = lj->h_A;
    if (B) *B = lj->h_B;
    if (t) *t = lj->h_type;
    return SPONGE_NB_OK;
}

static int scratch_ensure(PairScratch* ps, const sponge_nl* nl) {
    if (ps->cap >= nl->slot_cap && ps->acc) return SPONGE_NB_OK;
    // Sized once from the list's slot capacity, on the first call (before any
    // graph capture of the step can start).
    snb_scratch_free(ps);
    int cap = nl->slot_cap;
    SNB_HIP(hipMalloc((void**)&ps->pk, cap * sizeof(uint4)));
    SNB_HIP(hipMalloc((void**)&ps->acc, SPONGE_NB_NPARTS * sizeof(double)));
    SNB_HIP(hipMemset(ps->acc, 0, SPONGE_NB_NPARTS * sizeof(double)));
    ps->cap = cap;
    return SPONGE_NB_OK;
}

void snb_scratch_free(PairScratch* ps) {
    void* ptrs[] = {ps->pk, ps->acc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    hipEvent_t e0 = ps->ev_start, e1 = ps->ev_stop;
    memset(ps, 0, sizeof(*ps));
    ps->ev_start = e0;  // events belong to the caller
    ps->ev_stop = e1;
}

// Gather + pair kernel: forces added to d_frc (original order, may be NULL),
// per-atom energies to the given arrays, FP64 totals to ps->acc.
int snb_pairs(const sponge_nl* nl, const float2* ab, int nt, const int* d_type, float cutoff, const float* cof,
              const unsigned int* d_uint_crd, const float* d_charge, float beta, int energy, PairScratch* ps,
              float* d_frc, float* atom_e, float* atom_v, float* atom_d, hipStream_t s, int gathered,
              int overwrite) {
    int rc = scratch_ensure(ps, nl);
    if (rc) return rc;
    const int cap = ps->cap;
    if (!gathered) {
        gather_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, d_uint_crd, d_charge, d_type, ps->pk,
                                                         energy, ps->acc);
        SNB_CHECK_LAUNCH();
    }
    const int nblk = xcd_grid(ceil_div(nl->ncl_cap, kPairWaves));
    const float rc2 = cutoff * cutoff;
    const size_t lds = (size_t)nt * nt * sizeof(float2);
    if (ps->ev_start) SNB_HIP(hipEventRecord(ps->ev_start, s));
    if (energy) {
        pair_kernel<true><<<nblk, kPairWaves * 64, lds, s>>>(nl->flags, nl->list_n, nl->list, nl->list_cap, ps->pk,
                                                             ab, nt, cof[0], cof[1], cof[2], rc2, beta, nl->s2o, d_frc,
                                                             atom_e, atom_v, atom_d, ps->acc, overwrite);
    } else {
        pair_kernel<false><<<nblk, kPairWaves * 64, lds, s>>>(nl->flags, nl->list_n, nl->list, nl->list_cap, ps->pk,
                                                              ab, nt, cof[0], cof[1], cof[2], rc2, beta, nl->s2o, d_frc,
                                                              nullptr, nullptr, nullptr, ps->acc, overwrite);
    }
    SNB_CHECK_LAUNCH();
    if (ps->ev_stop) SNB_HIP(hipEventRecord(ps->ev_stop, s));
    return SPONGE_NB_OK;
}

int snb_step_prep(const sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const float* d_crd, unsigned int* d_uint_crd,
                  const float* d_charge, int energy, hipStream_t s) {
    int rc = scratch_ensure(&lj->ps, nl);
    if (rc) return rc;
    float cq[3];
    for (int d = 0; d < 3; d++) cq[d] = 0.25f * (4294967296.0f / lj->box[d]);
    const int cap = lj->ps.cap;
    step_prep_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, d_crd, cq[0], cq[1], cq[2], d_uint_crd,
                                                        d_charge, lj->d_type, lj->ps.pk, pme_prep_args(pme), energy,
                                                        lj->ps.acc);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_set_pair_events(sponge_lj* lj, void* start_event, void* stop_event) {
    SNB_ARG(lj, "lj_set_pair_events");
    lj->ps.ev_start = reinterpret_cast<hipEvent_t>(start_event);
    lj->ps.ev_stop = reinterpret_cast<hipEvent_t>(stop_event);
    return SPONGE_NB_OK;
}


extern "C" int sponge_lj_pme_direct_force(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                          const float* d_charge, float* d_frc, float pme_beta, int need_atom_energy,
                                          float* d_atom_energy, int need_virial, float* d_atom_lj_virial,
                                          float* d_atom_direct_pme_energy, void* stream) {
    SNB_ARG(lj && nl && d_uint_crd && d_charge && d_frc, "lj_pme_direct_force");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    hipStream_t s = as_stream(stream);
    const int energy = (need_atom_energy > 0 || need_virial > 0) ? 1 : 0;
    return snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, d_charge, pme_beta,
                     energy, &lj->ps, d_frc, need_atom_energy > 0 ? d_atom_energy : nullptr,
                     need_virial > 0 ? d_atom_lj_virial : nullptr, energy ? d_atom_direct_pme_energy : nullptr, s);
}

extern "C" int sponge_lj_long_range_correction(sponge_lj* lj, int need_pressure, float* d_virial, int need_potential,
                                               float* d_potential, void* stream) {
    SNB_ARG(lj, "lj_long_range_correction");
    float f = (float)lj->lrc_factor;
    float* a = need_pressure > 0 ? d_virial : nullptr;
    float* b = need_potential > 0 ? d_potential : nullptr;
    if (!a && !b) return SPONGE_NB_OK;
    add_scalar_kernel<<<1, 1, 0, as_stream(stream)>>>(a, f * 6.0f / lj->volume, b, f / lj->volume);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_get_energy(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                    double* h_energy, void* stream) {
    SNB_ARG(lj && nl && d_uint_crd && h_energy, "lj_get_energy");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    hipStream_t s = as_stream(stream);
    // LJ_Energy_CUDA takes no charges; run the pair kernel with zero charges
    // (the scratch charge array is the zero-filled energy buffer of PME-less
    // use: a dedicated zeroed array of n floats).
    float* zq = nullptr;
    SNB_HIP(hipMalloc((void**)&zq, lj->n * sizeof(float)));
    SNB_HIP(hipMemset(zq, 0, lj->n * sizeof(float)));
    int rc = snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, zq, 0.0f, 1,
                       &lj->ps, nullptr, nullptr, nullptr, nullptr, s);
    double acc[SPONGE_NB_NPARTS];
    if (rc == 0) {
        SNB_HIP(hipMemcpyAsync(acc, lj->ps.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
        SNB_HIP(hipStreamSynchronize(s));
        *h_energy = acc[SPONGE_PART_LJ] + lj->lrc_factor / lj->volume;
    }
    (void)hipFree(zq);
    return rc;
}

// ---------------------------------------------------------------------------
int snb_excluded(sponge_pme* pme, const unsigned int* d_uint_crd, const float scaler[3], const float* d_charge,
                 const int* es, const int* el, const int* en, float* d_frc, float* d_atom_energy, hipStream_t s) {
    const int n = pme->n;
    excluded_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_uint_crd, scaler[0], scaler[1], scaler[2], d_charge,
                                                     pme->beta, es, el, en, d_frc, d_atom_energy, pme->corr_atom);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_excluded_force(sponge_pme* pme, const unsigned int* d_uint_crd, const float scaler[3],
                                         const float* d_charge, const int* d_excluded_list_start,
                                         const int* d_excluded_list, const int* d_excluded_atom_numbers,
                                         float* d_frc, float* d_atom_energy, void* stream) {
    SNB_ARG(pme && d_uint_crd && scaler && d_charge && d_excluded_list_start && d_excluded_list &&
                d_excluded_atom_numbers && d_frc,
            "pme_excluded_force");
    hipStream_t s = as_stream(stream);
    return snb_excluded(pme, d_uint_crd, scaler, d_charge, d_excluded_list_start, d_excluded_list,
                        d_excluded_atom_numbers, d_frc, d_atom_energy, s);
}
