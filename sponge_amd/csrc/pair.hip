// LJ + PME direct-space pair force for gfx950 (LENNARD_JONES_INFORMATION,
// Lennard_Jones_force/Lennard_Jones_force.cu) and the excluded-pair
// correction (Particle_Mesh_Ewald::PME_Excluded_Force_With_Atom_Energy,
// PME_force/PME_force.cu:620-710).
//
// Pair kernel layout: one wave64 per 4-atom i-cluster; each iteration takes
// four j-clusters from the cluster-pair list, lane = (g:2, jj:2, ii:2) holds
// the pair (ii, jj) of j-cluster g. Coordinates and charges are gathered once
// per step into cluster order as one 16-byte record per atom; the i force
// stays in registers for the whole list, the j force is reduced over the four
// i lanes with DPP quad permutes and leaves with one atomic per component and
// j atom.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace snb;

namespace {

constexpr int kPairWaves = 4;
constexpr float kTwoOverSqrtPi = 1.1283791670955126f;

// Per-step gather into cluster order (replaces Copy_Crd_And_Charge_To_New_Crd,
// Lennard_Jones_force.cu:24-34) and reset of the sorted accumulators.
__global__ void gather_kernel(int cap, const int* __restrict__ flags, const int* __restrict__ s2o,
                              const unsigned int* __restrict__ uc, const float* __restrict__ q,
                              const int* __restrict__ type, uint4* __restrict__ pk, int* __restrict__ pty,
                              float* __restrict__ fs, float* __restrict__ es, int energy,
                              double* __restrict__ acc) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (energy && p < 3) acc[p] = 0.0;  // SPONGE_PART_LJ, _DIRECT, _LJ_VIRIAL
    if (energy && p == 3) acc[SPONGE_PART_PAIRS_IN] = 0.0;
    if (p >= flags[3]) return;
    int a = s2o[p];
    if (a >= 0) {
        pk[p] = make_uint4(uc[3 * a], uc[3 * a + 1], uc[3 * a + 2], __float_as_uint(q[a]));
        pty[p] = type[a];
    } else {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
        pty[p] = 0;
    }
    fs[p] = 0.f;
    fs[cap + p] = 0.f;
    fs[2 * cap + p] = 0.f;
    if (energy) {
        es[p] = 0.f;
        es[cap + p] = 0.f;
        es[2 * cap + p] = 0.f;
    }
}

// Adds the cluster-ordered accumulators into the caller's original-order
// arrays (frc AoS float3, per-atom energy arrays).
__global__ void scatter_kernel(int cap, const int* __restrict__ flags, const int* __restrict__ s2o,
                               const float* __restrict__ fs, const float* __restrict__ es,
                               float* __restrict__ frc, float* __restrict__ atom_e, float* __restrict__ atom_v,
                               float* __restrict__ atom_d) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= flags[3]) return;
    int a = s2o[p];
    if (a < 0) return;
    if (frc) {
        frc[3 * a] += fs[p];
        frc[3 * a + 1] += fs[cap + p];
        frc[3 * a + 2] += fs[2 * cap + p];
    }
    if (atom_e) atom_e[a] += es[p];
    if (atom_v) atom_v[a] += es[cap + p];
    if (atom_d) atom_d[a] = es[2 * cap + p];
}

template <bool ENERGY>
__global__ __launch_bounds__(kPairWaves * 64) void pair_kernel(
    const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list, int list_cap,
    const uint4* __restrict__ pk, const int* __restrict__ pty, const float2* __restrict__ ab, int nt, float cx,
    float cy, float cz, float rc2, float beta, float* __restrict__ fs, int cap, const int* __restrict__ s2o,
    float* __restrict__ es, double* __restrict__ acc) {
    __shared__ float2 sab[kMaxLjTypes * kMaxLjTypes];
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) sab[k] = ab[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lb = xcd_block(gridDim.x);
    const int ic = lb * kPairWaves + w;
    if (lb < 0 || ic >= flags[4]) return;
    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const float qi = __uint_as_float(a.w);
    const float2* srow = sab + pty[islot] * nt;
    int ai = 0;
    if (ENERGY) ai = s2o[islot];
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    float eie = 0.f, eiv = 0.f, eid = 0.f;
    double tlj = 0.0, tdir = 0.0, tvir = 0.0, tcnt = 0.0;
    const int n = list_n[ic];
    const uint2* lst = list + (size_t)ic * list_cap;
    for (int e0 = 0; e0 < n; e0 += 4) {
        const int e = e0 + g;
        uint2 ent = e < n ? lst[e] : make_uint2((unsigned)ic, 0u);
        const int jslot = (int)ent.x * kClusterSize + jj;
        const uint4 b = pk[jslot];
        const float dx = (float)(int)(b.x - a.x) * cx;
        const float dy = (float)(int)(b.y - a.y) * cy;
        const float dz = (float)(int)(b.z - a.z) * cz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool in = ((ent.y >> bit) & 1u) && r2 < rc2;
        const float r2s = in ? r2 : 1.0f;
        const float rinv = rsqrtf(r2s);
        const float rinv2 = rinv * rinv;
        const float rinv6 = rinv2 * rinv2 * rinv2;
        const float2 AB = srow[pty[jslot]];
        const float flj = (AB.y - AB.x * rinv6) * rinv6 * rinv2;
        const float r = r2s * rinv;
        const float br = beta * r;
        const float qq = qi * __uint_as_float(b.w);
        const float ex = expf(-br * br);
        const float ec = erfcf(br);
        const float fc = qq * (kTwoOverSqrtPi * br * ex + ec) * rinv * rinv2;
        const float f = in ? flj - fc : 0.f;
        const float fx = f * dx, fy = f * dy, fz = f * dz;
        fix += fx;
        fiy += fy;
        fiz += fz;
        // j force: reduce over the four i lanes of each (g, jj) quad.
        float jx = fx + dpp_quad_xor1(fx), jy = fy + dpp_quad_xor1(fy), jz = fz + dpp_quad_xor1(fz);
        jx += dpp_quad_xor2(jx);
        jy += dpp_quad_xor2(jy);
        jz += dpp_quad_xor2(jz);
        const bool jlead = ii == 0 && e < n;
        if (jlead) {
            atomicAdd(fs + jslot, -jx);
            atomicAdd(fs + cap + jslot, -jy);
            atomicAdd(fs + 2 * cap + jslot, -jz);
        }
        if (ENERGY) {
            // Energy steps evaluate the pair energies in FP64 from the exact
            // integer displacement: the totals are nets of large cancelling
            // terms, which FP32 terms cannot hold to 1e-6.
            double elj = 0.0, edir = 0.0, vlj = 0.0;
            if (in) {
                const double ddx = (double)(int)(b.x - a.x) * (double)cx;
                const double ddy = (double)(int)(b.y - a.y) * (double)cy;
                const double ddz = (double)(int)(b.z - a.z) * (double)cz;
                const double dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
                const double dr = sqrt(dr2);
                const double ir2 = 1.0 / dr2, ir6 = ir2 * ir2 * ir2;
                elj = ((double)AB.x * (1.0 / 12.0) * ir6 - (double)AB.y * (1.0 / 6.0)) * ir6;
                edir = (double)qi * (double)__uint_as_float(b.w) * erfc((double)beta * dr) / dr;
                vlj = ((double)AB.x * ir6 - (double)AB.y) * ir6;
            }
            tlj += elj;
            tdir += edir;
            tvir += vlj;
            tcnt += in ? 1.0 : 0.0;
            const int aj = s2o[jslot];
            const bool to_i = ai < aj;  // the reference credits the lower-index atom
            const float fe = (float)(elj + edir), fv = (float)vlj, fd = (float)edir;
            float ce = to_i ? 0.f : fe, cv = to_i ? 0.f : fv, cd = to_i ? 0.f : fd;
            if (to_i) {
                eie += fe;
                eiv += fv;
                eid += fd;
            }
            ce += dpp_quad_xor1(ce);
            cv += dpp_quad_xor1(cv);
            cd += dpp_quad_xor1(cd);
            ce += dpp_quad_xor2(ce);
            cv += dpp_quad_xor2(cv);
            cd += dpp_quad_xor2(cd);
            if (jlead) {
                atomicAdd(es + jslot, ce);
                atomicAdd(es + cap + jslot, cv);
                atomicAdd(es + 2 * cap + jslot, cd);
            }
        }
    }
    // i force: sum the 16 lanes sharing ii (lanes ii, ii+4, ..., ii+60).
    fix += __shfl_xor(fix, 4);
    fiy += __shfl_xor(fiy, 4);
    fiz += __shfl_xor(fiz, 4);
    fix += __shfl_xor(fix, 8);
    fiy += __shfl_xor(fiy, 8);
    fiz += __shfl_xor(fiz, 8);
    fix += __shfl_xor(fix, 16);
    fiy += __shfl_xor(fiy, 16);
    fiz += __shfl_xor(fiz, 16);
    fix += __shfl_xor(fix, 32);
    fiy += __shfl_xor(fiy, 32);
    fiz += __shfl_xor(fiz, 32);
    if (lane < 4) {
        atomicAdd(fs + islot, fix);
        atomicAdd(fs + cap + islot, fiy);
        atomicAdd(fs + 2 * cap + islot, fiz);
    }
    if (ENERGY) {
        for (int o = 4; o < 64; o <<= 1) {
            eie += __shfl_xor(eie, o);
            eiv += __shfl_xor(eiv, o);
            eid += __shfl_xor(eid, o);
        }
        if (lane < 4) {
            atomicAdd(es + islot, eie);
            atomicAdd(es + cap + islot, eiv);
            atomicAdd(es + 2 * cap + islot, eid);
        }
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
    SNB_HIP(hipMalloc((void**)&ps->pty, cap * sizeof(int)));
    SNB_HIP(hipMalloc((void**)&ps->fs, 3 * (size_t)cap * sizeof(float)));
    SNB_HIP(hipMalloc((void**)&ps->es, 3 * (size_t)cap * sizeof(float)));
    SNB_HIP(hipMalloc((void**)&ps->acc, SPONGE_NB_NPARTS * sizeof(double)));
    SNB_HIP(hipMemset(ps->acc, 0, SPONGE_NB_NPARTS * sizeof(double)));
    ps->cap = cap;
    return SPONGE_NB_OK;
}

void snb_scratch_free(PairScratch* ps) {
    void* ptrs[] = {ps->pk, ps->pty, ps->fs, ps->es, ps->acc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    hipEvent_t e0 = ps->ev_start, e1 = ps->ev_stop;
    memset(ps, 0, sizeof(*ps));
    ps->ev_start = e0;  // events belong to the caller
    ps->ev_stop = e1;
}

// Gather + pair kernel; results stay cluster-ordered in ps->fs / ps->es and
// the FP64 totals in ps->acc.
int snb_pairs(const sponge_nl* nl, const float2* ab, int nt, const int* d_type, float cutoff, const float* cof,
              const unsigned int* d_uint_crd, const float* d_charge, float beta, int energy, PairScratch* ps,
              hipStream_t s) {
    int rc = scratch_ensure(ps, nl);
    if (rc) return rc;
    const int cap = ps->cap;
    gather_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(cap, nl->flags, nl->s2o, d_uint_crd, d_charge, d_type, ps->pk,
                                                     ps->pty, ps->fs, ps->es, energy, ps->acc);
    SNB_CHECK_LAUNCH();
    const int nblk = xcd_grid(ceil_div(nl->ncl_cap, kPairWaves));
    const float rc2 = cutoff * cutoff;
    if (ps->ev_start) SNB_HIP(hipEventRecord(ps->ev_start, s));
    if (energy) {
        pair_kernel<true><<<nblk, kPairWaves * 64, 0, s>>>(nl->flags, nl->list_n, nl->list, nl->list_cap, ps->pk,
                                                           ps->pty, ab, nt, cof[0], cof[1], cof[2], rc2, beta, ps->fs,
                                                           cap, nl->s2o, ps->es, ps->acc);
    } else {
        pair_kernel<false><<<nblk, kPairWaves * 64, 0, s>>>(nl->flags, nl->list_n, nl->list, nl->list_cap, ps->pk,
                                                            ps->pty, ab, nt, cof[0], cof[1], cof[2], rc2, beta, ps->fs,
                                                            cap, nl->s2o, ps->es, ps->acc);
    }
    SNB_CHECK_LAUNCH();
    if (ps->ev_stop) SNB_HIP(hipEventRecord(ps->ev_stop, s));
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_set_pair_events(sponge_lj* lj, void* start_event, void* stop_event) {
    SNB_ARG(lj, "lj_set_pair_events");
    lj->ps.ev_start = reinterpret_cast<hipEvent_t>(start_event);
    lj->ps.ev_stop = reinterpret_cast<hipEvent_t>(stop_event);
    return SPONGE_NB_OK;
}

int snb_pairs_scatter(const sponge_nl* nl, const PairScratch* ps, float* d_frc, float* atom_e, float* atom_v,
                      float* atom_d, hipStream_t s) {
    const int cap = ps->cap;
    scatter_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(cap, nl->flags, nl->s2o, ps->fs, ps->es, d_frc, atom_e, atom_v,
                                                      atom_d);
    SNB_CHECK_LAUNCH();
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
    int rc = snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, d_charge, pme_beta,
                       energy, &lj->ps, s);
    if (rc) return rc;
    return snb_pairs_scatter(nl, &lj->ps, d_frc, need_atom_energy > 0 ? d_atom_energy : nullptr,
                             need_virial > 0 ? d_atom_lj_virial : nullptr,
                             energy ? d_atom_direct_pme_energy : nullptr, s);
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
                       &lj->ps, s);
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
