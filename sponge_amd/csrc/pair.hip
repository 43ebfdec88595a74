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

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <hip/hip_ext.h>

#include "common.hpp"
#include "pair_math.hpp"
#include "pme_prep.hpp"
#include "pair_body.hpp"
#include "tune_table.hpp"

using namespace snb;

namespace {

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

// Slot records from the caller's UINT_VECTOR_LJ_TYPE array (uint x, y, z, LJ
// type, charge: 5 words per atom, Lennard_Jones_force.cuh:26-33).
__global__ void gather_records_kernel(const int* __restrict__ flags, const int* __restrict__ s2o,
                                      const unsigned int* __restrict__ rec, uint4* __restrict__ pk) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= flags[3]) return;
    const int a = s2o[p];
    if (a >= 0) {
        const unsigned* r = rec + 5 * (size_t)a;
        const unsigned t = r[3];
        pk[p] = make_uint4((r[0] & ~3u) | (t & 3u), (r[1] & ~3u) | ((t >> 2) & 3u), (r[2] & ~3u) | ((t >> 4) & 3u),
                           r[4]);
    } else {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Fused per-step prep of the whole force path, one thread per cluster slot:
// optional MD_Information_Crd_To_Uint_Crd (MD_core.cu:1528-1532, written back
// to the caller's uint array), the pair-kernel record, and the PME base,
// fraction and x-plane bin of the slot's atom. The threads past the slots
// take the atoms a column overflow left out of the list (flags[5] of them;
// slots + those <= slot capacity): fresh uint coordinates, zero charge on
// the PME grid (so no reciprocal force either), zero force when the step
// assigns forces (zero_frc) -- the overflow is reported by the list status.
__device__ __forceinline__ void step_prep_body(int blk, const int* __restrict__ flags, const int* __restrict__ s2o,
                                               const float* __restrict__ crd, float cqx, float cqy, float cqz,
                                               unsigned int* __restrict__ uc, const float* __restrict__ q,
                                               const int* __restrict__ type, uint4* __restrict__ pk,
                                               const PmePrep& P, int energy, double* __restrict__ acc,
                                               int* __restrict__ step_counter, const int* __restrict__ dropped,
                                               float* __restrict__ zero_frc, const uint4* __restrict__ prune_ref = nullptr,
                                               float vote_r2 = 0.f, float cx = 0.f, float cy = 0.f, float cz = 0.f,
                                               int* __restrict__ vote = nullptr, int halo = 0) {
    const int p = blk * blockDim.x + threadIdx.x;
    if (energy && p < 3) acc[p] = 0.0;  // SPONGE_PART_LJ, _DIRECT, _LJ_VIRIAL
    if (energy && p == 3) acc[SPONGE_PART_PAIRS_IN] = 0.0;
    const int nslot = flags[3];
    int a = p < nslot ? s2o[p] : -1;
    bool left_out = false;
    if (p >= nslot && p < nslot + flags[5]) {
        a = dropped[p - nslot];
        left_out = true;
    }
    unsigned u0 = 0u, u1 = 0u, u2 = 0u;
    float qa = 0.f;
    if (left_out) {
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
        if (zero_frc) {
            zero_frc[3 * a] = 0.f;
            zero_frc[3 * a + 1] = 0.f;
            zero_frc[3 * a + 2] = 0.f;
        }
    } else if (a >= 0) {
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
        const uint4 rec = make_uint4((u0 & ~3u) | (t & 3u), (u1 & ~3u) | ((t >> 2) & 3u),
                                     (u2 & ~3u) | ((t >> 4) & 3u), __float_as_uint(qa));
        pk[p] = rec;
        // periodic images for the 16-bit partner offsets: needed only when
        // some offset wraps, i.e. beyond 32767 slots
        if (halo && nslot > 32767) {
            if (p < kPkHalo) pk[nslot + p] = rec;
            if (p >= nslot - kPkHalo) pk[p - nslot] = rec;
        }
    } else if (p < nslot) {
        pk[p] = make_uint4(0u, 0u, 0u, 0u);
        if (halo && nslot > 32767) {
            if (p < kPkHalo) pk[nslot + p] = make_uint4(0u, 0u, 0u, 0u);
            if (p >= nslot - kPkHalo) pk[p - nslot] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
    if (vote) {
        // displacement of the slot's atom since the inner list's last prune
        bool moved = false;
        if (a >= 0 && !left_out) {
            const uint4 r = prune_ref[p];
            const float dx = (float)(int)(u0 - (r.x & ~3u)) * cx;
            const float dy = (float)(int)(u1 - (r.y & ~3u)) * cy;
            const float dz = (float)(int)(u2 - (r.z & ~3u)) * cz;
            moved = dx * dx + dy * dy + dz * dz > vote_r2;
        }
        // this step's number since the last rebuild (the count of completed
        // steps is advanced by the step's last kernel, so it is fixed here)
        if (__ballot(moved) && (threadIdx.x & 63) == 0) atomicMax(vote, *step_counter + 1);
    }
    pme_prep_block(a, u0, u1, u2, qa, P, energy);
}

__global__ __launch_bounds__(256) void step_prep_kernel(const int* __restrict__ flags, const int* __restrict__ s2o,
                                                        const float* __restrict__ crd, float cqx, float cqy, float cqz,
                                                        unsigned int* __restrict__ uc, const float* __restrict__ q,
                                                        const int* __restrict__ type, uint4* __restrict__ pk,
                                                        PmePrep P, int energy, double* __restrict__ acc,
                                                        int* __restrict__ step_counter, const int* __restrict__ dropped,
                                                        float* __restrict__ zero_frc, const uint4* __restrict__ prune_ref,
                                                        float vote_r2, float cx, float cy, float cz,
                                                        int* __restrict__ vote, int halo) {
    step_prep_body(blockIdx.x, flags, s2o, crd, cqx, cqy, cqz, uc, q, type, pk, P, energy, acc, step_counter, dropped,
                   zero_frc, prune_ref, vote_r2, cx, cy, cz, vote, halo);
}

__global__ __launch_bounds__(256) void step_prep_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map,
                                                       int energy, int from_crd, int overwrite) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const PmePrep P = pme_prep_args(&b.pme);
    step_prep_body(m.y, b.nl.flags, b.nl.s2o, from_crd ? b.io.crd : nullptr, b.cq[0], b.cq[1], b.cq[2],
                   b.io.uint_crd, b.io.charge, b.lj.d_type, b.lj.ps.pk, P, energy, b.lj.ps.acc, nullptr, b.nl.dropped,
                   overwrite ? b.io.frc : nullptr);
}



// W: i-clusters (waves) per block; the force-only plain-LJ launch takes the
// handle's tuned value (2, 4 or 8), every other launch 4. The force-only
// kernels are held to 64 VGPRs (8 waves per SIMD): the per-slot force loops
// wait on gathered slot records, which the other waves cover; the filtering
// loop of the two-level lists would take 66 (two registers are spilled
// around it, outside its loop).
template <bool ENERGY, bool SOFT, int TB, int W = kPairWaves>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(!ENERGY && !SOFT ? 8 : 1))) void pair_kernel(
    const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list, int list_cap,
    const uint4* __restrict__ pk, const float2* __restrict__ ab, int nt, float cx, float cy, float cz, float rc2,
    float beta, const int* __restrict__ s2o, float* __restrict__ frc, float* __restrict__ atom_e,
    float* __restrict__ atom_v, float* __restrict__ atom_d, double* __restrict__ acc, int overwrite,
    const float4* __restrict__ sc, PruneArgs pa, ExclArgs ex) {
    pair_body<ENERGY, SOFT, TB, W>(xcd_block(gridDim.x), flags, list_n, list, list_cap, pk, ab, nt, cx, cy, cz, rc2,
                                   beta, s2o, frc, atom_e, atom_v, atom_d, acc, overwrite, sc, pa, ex);
}

// Batched: logical blocks are spread over the XCDs as in the single-system
// launch, then mapped to (system, block of 4 i-clusters).
template <bool ENERGY, bool SOFT, int TB>
__global__ __launch_bounds__(kPairWaves * 64) void pair_batch(const BatchSys* __restrict__ bs,
                                                              const int2* __restrict__ map, int nmap,
                                                              int overwrite) {
    const int g = xcd_block(nmap);
    if (g < 0) return;
    const int2 m = map[g];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const float rc = b.lj.cutoff;
    pair_body<ENERGY, SOFT, TB>(m.y, b.nl.flags, b.nl.list_n, b.nl.list, b.nl.list_cap, b.lj.ps.pk, b.lj.d_ab,
                            b.lj.ntypes, b.lj.cof_dr[0], b.lj.cof_dr[1], b.lj.cof_dr[2], rc * rc, b.pme.beta,
                            b.nl.s2o, b.io.frc, ENERGY ? b.io.atom_energy : nullptr,
                            ENERGY ? b.io.atom_virial : nullptr, nullptr, b.lj.ps.acc, overwrite, b.lj.d_sc);
}

__global__ void add_scalar_kernel(float* a, float va, float* b, float vb) {
    if (a) a[0] += va;
    if (b) b[0] += vb;
}

// PME_Excluded_Force_With_Atom_Energy_Correction (PME_force.cu:620-695):
// one thread per atom over its exclusion CSR row, FP32; energy
// -q_i q_j erf(b r)/r credited to atom i; FP64 total in acc.
__device__ __forceinline__ void excluded_body(int blk, int n, const unsigned int* __restrict__ uc, float sx, float sy,
                                              float sz, const float* __restrict__ q, float beta,
                                              const int* __restrict__ es, const int* __restrict__ el,
                                              const int* __restrict__ en, float* __restrict__ frc,
                                              float* __restrict__ atom_e, float* __restrict__ corr) {
    int i = blk * blockDim.x + threadIdx.x;
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

__global__ void excluded_kernel(int n, const unsigned int* __restrict__ uc, float sx, float sy, float sz,
                                const float* __restrict__ q, float beta, const int* __restrict__ es,
                                const int* __restrict__ el, const int* __restrict__ en, float* __restrict__ frc,
                                float* __restrict__ atom_e, float* __restrict__ corr) {
    excluded_body(blockIdx.x, n, uc, sx, sy, sz, q, beta, es, el, en, frc, atom_e, corr);
}

__global__ __launch_bounds__(256) void excluded_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    excluded_body(m.y, b.pme.n, b.io.uint_crd, b.lj.cof_dr[0], b.lj.cof_dr[1], b.lj.cof_dr[2], b.io.charge,
                  b.pme.beta, b.io.excluded_list_start, b.io.excluded_list, b.io.excluded_numbers, b.io.frc, nullptr,
                  b.pme.corr_atom);
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free direct force: the cluster-order records and FP64 totals of
// sponge_lj_pme_direct_force live in this file-scope device array instead of
// the module's scratch (allocated on a handle's first call). One copy per
// GPU, held until the process ends; every call on a GPU shares it.
constexpr int kNoAllocSlots = 131072;  // list slots (atoms + column padding)
constexpr size_t kNoAllocPkBytes = (size_t)kNoAllocSlots * sizeof(uint4);
constexpr size_t kNoAllocBytes = kNoAllocPkBytes + SPONGE_NB_NPARTS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the array's host-side name is not a device address).
constexpr int kNoAllocMaxDevices = 64;
static void* g_scratch_ptr[kNoAllocMaxDevices];
static hipError_t g_scratch_err[kNoAllocMaxDevices];
static std::once_flag g_scratch_once[kNoAllocMaxDevices];

static int noalloc_scratch(void** out) {
    int dev = 0;
    SNB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("allocation-free entry points support devices 0..%d (current %d)", kNoAllocMaxDevices - 1, dev);
        return SPONGE_NB_ENODEV;
    }
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = p;
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev]) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(g_scratch_err[dev]));
        return SPONGE_NB_EHIP;
    }
    *out = g_scratch_ptr[dev];
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_pme_direct_force_noalloc(sponge_lj* lj, const sponge_nl* nl,
                                                  const unsigned int* d_uint_crd, const float* d_charge, float* d_frc,
                                                  float pme_beta, int need_atom_energy, float* d_atom_energy,
                                                  int need_virial, float* d_atom_lj_virial,
                                                  float* d_atom_direct_pme_energy, void* stream) {
    SNB_ARG(lj && nl && d_uint_crd && d_charge && d_frc, "lj_pme_direct_force_noalloc");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    if (nl->slot_cap > kNoAllocSlots) {
        set_error("lj_pme_direct_force_noalloc: the list has %d slots; the static scratch holds %d "
                  "(about 125,000 atoms); use sponge_lj_pme_direct_force",
                  nl->slot_cap, kNoAllocSlots);
        return SPONGE_NB_ECAPACITY;
    }
    void* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    // The gather kernel rewrites every record the pair kernel reads and, with
    // energy, zeroes the totals first: nothing of an earlier call is read.
    PairScratch ps{};
    ps.cap = kNoAllocSlots;
    ps.pk = reinterpret_cast<uint4*>(base);
    ps.acc = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(base) + kNoAllocPkBytes);
    hipStream_t s = as_stream(stream);
    const int energy = (need_atom_energy > 0 || need_virial > 0) ? 1 : 0;
    return snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, d_charge, pme_beta,
                     energy, &ps, d_frc, need_atom_energy > 0 ? d_atom_energy : nullptr,
                     need_virial > 0 ? d_atom_lj_virial : nullptr, energy ? d_atom_direct_pme_energy : nullptr, s, 0,
                     0, lj->d_sc);
}

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
    lj->ps.waves = tune_row(atom_numbers).pair_waves;
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

// LJ_SOFT_CORE::Initial (LJ_soft_core.cu:1420-1610) after its file reads.
// Atoms are grouped into combined types c = (type A, type B, mask); every
// (c_i, c_j) pair type is decided once on the host: ordinary pair types get
// the lambda-mixed (A, B) of the plain kernel, soft-core ones (mask differs,
// or exactly one end state has A below 1e-6, LJ_soft_core.cu:251) get the
// records soft_pair reads. The long-range factor sums (1 - lambda) AB + lambda BB
// over all ordered atom pairs in FP64 (Total_C6_Get, :88-118, :1588-1600).
extern "C" int sponge_lj_soft_create(int atom_numbers, int atom_type_numbers_A, int atom_type_numbers_B,
                                     const float* h_LJ_AA, const float* h_LJ_AB, const float* h_LJ_BA,
                                     const float* h_LJ_BB, const int* h_atom_LJ_type_A,
                                     const int* h_atom_LJ_type_B, const int* h_subsys_division, float lambda,
                                     float alpha, float p, float sigma, float sigma_min, float cutoff,
                                     const float box_length[3], sponge_lj** out) {
    SNB_ARG(out && h_LJ_AA && h_LJ_AB && h_LJ_BA && h_LJ_BB && h_atom_LJ_type_A && h_atom_LJ_type_B && box_length,
            "null pointer");
    SNB_ARG(atom_numbers > 0, "atom_numbers");
    SNB_ARG(atom_type_numbers_A > 0 && atom_type_numbers_B > 0, "atom_type_numbers");
    SNB_ARG(lambda >= 0.f && lambda <= 1.f, "lambda_lj must be in [0, 1]");
    *out = nullptr;
    const int TA = atom_type_numbers_A, TB = atom_type_numbers_B;
    std::map<std::tuple<int, int, int>, int> combo;
    std::vector<int> ctype(atom_numbers);
    std::vector<std::tuple<int, int, int>> cdesc;
    for (int i = 0; i < atom_numbers; i++) {
        const int a = h_atom_LJ_type_A[i], b = h_atom_LJ_type_B[i], m = h_subsys_division ? h_subsys_division[i] : 0;
        SNB_ARG(a >= 0 && a < TA && b >= 0 && b < TB, "atom LJ type out of range");
        auto key = std::make_tuple(a, b, m);
        auto it = combo.find(key);
        if (it == combo.end()) {
            SNB_ARG((int)cdesc.size() < kMaxLjTypes, "more than 64 distinct (type A, type B, mask) combinations");
            it = combo.emplace(key, (int)cdesc.size()).first;
            cdesc.push_back(key);
        }
        ctype[i] = it->second;
    }
    const int C = (int)cdesc.size();
    auto pidx = [](int x, int y) {
        const int hi = x > y ? x : y, lo = x > y ? y : x;
        return hi * (hi + 1) / 2 + lo;
    };
    const double lam = lambda, l_ = 1.0 - lam;
    const double alp = alpha * std::pow(lam, (double)p), alp_ = alpha * std::pow(l_, (double)p);
    const double sig6 = std::pow((double)sigma, 6.0), sig6min = std::pow((double)sigma_min, 6.0);
    std::vector<float2> sq((size_t)C * C);
    std::vector<float4> sc(2 * (size_t)C * C, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float4> fe(2 * (size_t)C * C);
    std::vector<double> cnt(C, 0.0);
    for (int i = 0; i < atom_numbers; i++) cnt[ctype[i]] += 1;
    double lrc = 0, lrc_a = 0, lrc_b = 0;
    for (int x = 0; x < C; x++)
        for (int y = 0; y < C; y++) {
            const int ax = std::get<0>(cdesc[x]), bx = std::get<1>(cdesc[x]), mx = std::get<2>(cdesc[x]);
            const int ay = std::get<0>(cdesc[y]), by = std::get<1>(cdesc[y]), my = std::get<2>(cdesc[y]);
            const int pa = pidx(ax, ay), pb = pidx(bx, by);
            const double aa = h_LJ_AA[pa], ab = h_LJ_AB[pa], ba = h_LJ_BA[pb], bb = h_LJ_BB[pb];
            lrc += cnt[x] * cnt[y] * (l_ * ab + lam * bb);
            lrc_a += cnt[x] * cnt[y] * ab;
            lrc_b += cnt[x] * cnt[y] * bb;
            const bool soft = mx != my || (ba > 1e-6 && aa < 1e-6) || (ba < 1e-6 && aa > 1e-6);
            const size_t k = (size_t)x * C + y;
            const double sa6 = (aa < 1e-6 || ab < 1e-6) ? sig6 : std::max(0.5 * aa / ab, sig6min);
            const double sb6 = (ba < 1e-6 || bb < 1e-6) ? sig6 : std::max(0.5 * ba / bb, sig6min);
            fe[k] = make_float4((float)aa, (float)ab, (float)ba, (float)bb);
            fe[(size_t)C * C + k] = make_float4(soft ? 1.f : 0.f, mx == my ? 1.f : 0.f, 0.f, 0.f);
            if (!soft) {
                sq[k] = make_float2((float)(l_ * aa + lam * ba), (float)(l_ * ab + lam * bb));
            } else {
                sq[k] = make_float2(-1.0f, 0.0f);
                sc[k] = make_float4((float)(l_ * aa), (float)(l_ * ab), (float)(lam * ba), (float)(lam * bb));
                sc[(size_t)C * C + k] = make_float4((float)(alp * sa6), (float)(alp_ * sb6), (float)l_, (float)lam);
            }
        }
    sponge_lj* lj = new sponge_lj();
    memset(lj, 0, sizeof(*lj));
    lj->n = atom_numbers;
    lj->ntypes = C;
    lj->cutoff = cutoff;
    const int P = TA * (TA + 1) / 2;
    lj->h_A = new float[P];
    lj->h_B = new float[P];
    lj->h_type = new int[atom_numbers];
    memcpy(lj->h_A, h_LJ_AA, P * sizeof(float));
    memcpy(lj->h_B, h_LJ_AB, P * sizeof(float));
    memcpy(lj->h_type, h_atom_LJ_type_A, atom_numbers * sizeof(int));
    sponge_lj_update_volume(lj, box_length);
    const double kf = (-2.0 / 3.0) * M_PI / ((double)cutoff * cutoff * cutoff) / 6.0;
    lj->lrc_factor = lrc * kf;
    lj->lrc_A = lrc_a * kf;
    lj->lrc_B = lrc_b * kf;
    lj->fep_lambda = lam;
    lj->fep_alpha = alpha;
    lj->fep_p = p;
    lj->fep_sig6 = sig6;
    lj->fep_sig6min = sig6min;
    hipError_t e = hipMalloc((void**)&lj->d_ab, sq.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(lj->d_ab, sq.data(), sq.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&lj->d_fe, fe.size() * sizeof(float4));
    if (e == hipSuccess) e = hipMemcpy(lj->d_fe, fe.data(), fe.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&lj->d_sc, sc.size() * sizeof(float4));
    if (e == hipSuccess) e = hipMemcpy(lj->d_sc, sc.data(), sc.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&lj->d_type, atom_numbers * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(lj->d_type, ctype.data(), atom_numbers * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("sponge_lj_soft_create: %s", hipGetErrorString(e));
        sponge_lj_destroy(lj);
        return SPONGE_NB_EHIP;
    }
    *out = lj;
    return SPONGE_NB_OK;
}

// New A / B tables for a plain handle (x12 / x6 triangular, as created): the
// square device table, the host copies and the long-range factor. Synchronous.
extern "C" int sponge_lj_set_tables(sponge_lj* lj, const float* h_LJ_A, const float* h_LJ_B) {
    SNB_ARG(lj && h_LJ_A && h_LJ_B, "lj_set_tables");
    SNB_ARG(!lj->d_sc, "lj_set_tables: soft-core handles keep their tables");
    const int T = lj->ntypes, P = T * (T + 1) / 2;
    memcpy(lj->h_A, h_LJ_A, P * sizeof(float));
    memcpy(lj->h_B, h_LJ_B, P * sizeof(float));
    std::vector<double> cnt(T, 0.0);
    for (int i = 0; i < lj->n; i++) cnt[lj->h_type[i]] += 1;
    std::vector<float2> sq((size_t)T * T);
    double sum = 0;
    for (int x = 0; x < T; x++)
        for (int y = 0; y < T; y++) {
            const int hi = x > y ? x : y, lo = x > y ? y : x, k = hi * (hi + 1) / 2 + lo;
            sum += cnt[x] * cnt[y] * (double)h_LJ_B[k];
            sq[(size_t)x * T + y] = make_float2(h_LJ_A[k], h_LJ_B[k]);
        }
    lj->lrc_factor = sum * (-2.0 / 3.0) * M_PI / ((double)lj->cutoff * lj->cutoff * lj->cutoff) / 6.0;
    SNB_HIP(hipMemcpy(lj->d_ab, sq.data(), sq.size() * sizeof(float2), hipMemcpyHostToDevice));
    return SPONGE_NB_OK;
}

extern "C" void sponge_lj_destroy(sponge_lj* lj) {
    if (!lj) return;
    snb_adj_free(lj);
    snb_fep_free(lj->fep);
    void* ptrs[] = {lj->d_ab, lj->d_sc, lj->d_fe, lj->d_type};
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
    // the records with kPkHalo slots on either side (16-bit partner offsets
    // of the per-slot inner lists reach them); zeroed so that no lane ever
    // reads an undefined record
    const size_t nrec = (size_t)cap + 2 * (size_t)kPkHalo;
    SNB_HIP(hipMalloc((void**)&ps->pk_alloc, nrec * sizeof(uint4)));
    SNB_HIP(hipMemset(ps->pk_alloc, 0, nrec * sizeof(uint4)));
    ps->pk = ps->pk_alloc + kPkHalo;
    SNB_HIP(hipMalloc((void**)&ps->acc, SPONGE_NB_NPARTS * sizeof(double)));
    SNB_HIP(hipMemset(ps->acc, 0, SPONGE_NB_NPARTS * sizeof(double)));
    ps->cap = cap;
    return SPONGE_NB_OK;
}

// Cubic Hermite pieces of the direct Coulomb force factor
// G(r) = erfc(beta r) / r^3 + 2 beta / sqrt(pi) exp(-beta^2 r^2) / r^2 (force
// on atom i from j: -q_i q_j G(r) r_ij, the PME direct term of
// Lennard_Jones_force.cu:495-510) on kCoulTab intervals of [0, rc (1 + 1e-4)]:
// piece k holds (G, h G', 3 (G1 - G0) - 2 h G0' - h G1', 2 (G0 - G1) + h G0' +
// h G1') at its ends, so G(k h + u h) = c0 + u (c1 + u (c2 + u c3)). FP64 on
// the host; the first piece (r below one interval, never a listed pair) is a
// copy of the second. Built on first use and when beta or the cutoff changes
// (a synchronous copy: not during graph capture).
static int ensure_coulomb_table(PairScratch* ps, float beta, float rc) {
    if (ps->ctab && ps->ctab_beta == beta && ps->ctab_rc == rc) return SPONGE_NB_OK;
    if (!ps->ctab) SNB_HIP(hipMalloc((void**)&ps->ctab, kCoulTab * sizeof(float4)));
    const double b = beta, rmax = (double)rc * 1.0001, h = rmax / kCoulTab, tsp = 2.0 * b / std::sqrt(M_PI);
    auto G = [&](double r) { return std::erfc(b * r) / (r * r * r) + tsp * std::exp(-b * b * r * r) / (r * r); };
    auto dG = [&](double r) {
        return -3.0 * std::erfc(b * r) / (r * r * r * r) - tsp * std::exp(-b * b * r * r) * (3.0 / (r * r * r) + 2.0 * b * b / r);
    };
    std::vector<float4> t(kCoulTab);
    for (int k = 1; k < kCoulTab; k++) {
        const double r0 = k * h, r1 = r0 + h;
        const double g0 = G(r0), g1 = G(r1), d0 = h * dG(r0), d1 = h * dG(r1);
        t[k] = make_float4((float)g0, (float)d0, (float)(3.0 * (g1 - g0) - 2.0 * d0 - d1),
                           (float)(2.0 * (g0 - g1) + d0 + d1));
    }
    t[0] = t[1];
    SNB_HIP(hipMemcpy(ps->ctab, t.data(), kCoulTab * sizeof(float4), hipMemcpyHostToDevice));
    ps->ctab_beta = beta;
    ps->ctab_rc = rc;
    return SPONGE_NB_OK;
}

void snb_scratch_free(PairScratch* ps) {
    void* ptrs[] = {ps->pk_alloc, ps->acc, ps->ctab};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    hipEvent_t e0 = ps->ev_start, e1 = ps->ev_stop;
    const int waves = ps->waves;
    memset(ps, 0, sizeof(*ps));
    ps->ev_start = e0;  // events belong to the caller
    ps->ev_stop = e1;
    ps->waves = waves;  // a setting, not storage
}

// Gather + pair kernel: forces added to d_frc (original order, may be NULL),
// per-atom energies to the given arrays, FP64 totals to ps->acc.
int snb_pairs(const sponge_nl* nl, const float2* ab, int nt, const int* d_type, float cutoff, const float* cof,
              const unsigned int* d_uint_crd, const float* d_charge, float beta, int energy, PairScratch* ps,
              float* d_frc, float* atom_e, float* atom_v, float* atom_d, hipStream_t s, int gathered,
              int overwrite, const float4* sc, int inner, ExclArgs ex) {
    // the per-slot inner lists address records through 16-bit offsets,
    // which needs the record halo of an owned scratch
    const bool fused = inner != 0;
    inner = inner && nl->prune_interval > 0 && ps->pk_alloc != nullptr;
    PruneArgs pa{};
    if (inner) {
        pa.interval = nl->prune_interval;
        pa.r2 = nl->prune_r2;
        pa.in_n = nl->list_in_n;
        pa.in_list = nl->list_in;
        pa.atom_n = nl->atom_in_n;
        pa.atom_list = nl->atom_in;
        pa.atom_cap = nl->atom_in_cap;
        pa.diag = nl->flags + 11;
        pa.ref = nl->prune_ref;
        if (!sc && snb_two_level(nl)) {
            pa.atom_on = nl->atom_out_n;
            pa.atom_olist = nl->atom_out;
        }
    }
    const int* list_n = nl->list_n;
    const uint2* list = nl->list;
    int rc = scratch_ensure(ps, nl);
    if (rc) return rc;
    const int cap = ps->cap;
    pa.pk_records = cap + 2 * kPkHalo;
    if (!gathered) {
        gather_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, d_uint_crd, d_charge, d_type, ps->pk,
                                                         energy, ps->acc);
        SNB_CHECK_LAUNCH();
    }
    const int W = (!energy && !sc && (ps->waves == 2 || ps->waves == 8)) ? ps->waves : kPairWaves;
    const int nblk = xcd_grid(ceil_div(nl->ncl_cap, W));
    const float rc2 = cutoff * cutoff;
    size_t lds = (size_t)nt * nt * sizeof(float2);
    // the per-slot force loop of force-only plain steps reads the Coulomb
    // force factor from a table in LDS; every fused step makes sure it
    // exists (built by the first one, so that a prune schedule turned on
    // later allocates nothing inside a graph capture)
    if (fused && !energy && !sc && ps->pk_alloc) {
        rc = ensure_coulomb_table(ps, beta, cutoff);
        if (rc) return rc;
    }
    if (inner && !energy && !sc) {
        pa.ctab = ps->ctab;
        pa.ctab_scale = (float)kCoulTab / (cutoff * 1.0001f);
        lds = ((lds + 15) & ~(size_t)15) + kCoulTab * sizeof(float4);
    }
    // With timing events set, the launch itself records them (hipExtLaunchKernel
    // takes the kernel's own start and end times; events recorded around a
    // launch add the marker packets' latency, ~8 us per launch here)
    const bool timed = ps->ev_start && ps->ev_stop;
#define SNB_PAIR(E, S, W)                      \
    do {                                       \
        if (nt <= 4) SNB_PAIR_TB(E, S, 2, W);  \
        else SNB_PAIR_TB(E, S, 6, W);          \
    } while (0)
#define SNB_PAIR_TB(E, S, TB, W)                                                                                     \
    do {                                                                                                            \
        if (timed)                                                                                                  \
            hipExtLaunchKernelGGL(pair_kernel<E, S, TB, W>, dim3(nblk), dim3(W * 64), (std::uint32_t)lds, s,         \
                                  ps->ev_start, ps->ev_stop, 0u, nl->flags, list_n, list, nl->list_cap,              \
                                  (const uint4*)ps->pk, ab, nt, cof[0], cof[1], cof[2], rc2, beta, nl->s2o, d_frc,   \
                                  E ? atom_e : nullptr, E ? atom_v : nullptr, E ? atom_d : nullptr, ps->acc,         \
                                  overwrite, sc, pa, ex);                                                            \
        else                                                                                                        \
            pair_kernel<E, S, TB, W><<<nblk, W * 64, lds, s>>>(                                                     \
                nl->flags, list_n, list, nl->list_cap, ps->pk, ab, nt, cof[0], cof[1], cof[2], rc2, beta, nl->s2o,  \
                d_frc, E ? atom_e : nullptr, E ? atom_v : nullptr, E ? atom_d : nullptr, ps->acc, overwrite, sc,    \
                pa, ex);                                                                                            \
    } while (0)
    if (energy) {
        if (sc) SNB_PAIR(true, true, kPairWaves);
        else SNB_PAIR(true, false, kPairWaves);
    } else if (sc) {
        SNB_PAIR(false, true, kPairWaves);
    } else if (W == 2) {
        SNB_PAIR(false, false, 2);
    } else if (W == 8) {
        SNB_PAIR(false, false, 8);
    } else {
        SNB_PAIR(false, false, kPairWaves);
    }
#undef SNB_PAIR
#undef SNB_PAIR_TB
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_gather(const sponge_nl* nl, const unsigned int* d_uint_crd, const float* d_charge, const int* d_type,
               PairScratch* ps, hipStream_t s) {
    int rc = scratch_ensure(ps, nl);
    if (rc) return rc;
    gather_kernel<<<ceil_div(ps->cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, d_uint_crd, d_charge, d_type, ps->pk, 0,
                                                         ps->acc);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_step_prep(const sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const float* d_crd, unsigned int* d_uint_crd,
                  const float* d_charge, int energy, float* d_frc_assigned, hipStream_t s) {
    int rc = scratch_ensure(&lj->ps, nl);
    if (rc) return rc;
    float cq[3];
    for (int d = 0; d < 3; d++) cq[d] = 0.25f * (4294967296.0f / lj->box[d]);
    const int cap = lj->ps.cap;
    const bool prune = nl->prune_interval > 0;
    // displacement votes: the per-slot lists of plain LJ prune by distance
    // only in the two-level form (else they hold every pair of the full
    // list); soft-core LJ prunes its cluster-pair inner list by distance
    const bool vote = prune && (lj->d_sc != nullptr || snb_two_level(nl));
    step_prep_kernel<<<ceil_div(cap, 256), 256, 0, s>>>(nl->flags, nl->s2o, d_crd, cq[0], cq[1], cq[2], d_uint_crd,
                                                        d_charge, lj->d_type, lj->ps.pk, pme_prep_args(pme), energy,
                                                        lj->ps.acc, prune ? nl->flags + 10 : nullptr, nl->dropped,
                                                        d_frc_assigned, nl->prune_ref, nl->prune_vote_r2,
                                                        lj->cof_dr[0], lj->cof_dr[1], lj->cof_dr[2],
                                                        vote ? nl->flags + 13 : nullptr,
                                                        (prune && lj->ps.pk_alloc) ? 1 : 0);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_lj_set_pair_events(sponge_lj* lj, void* start_event, void* stop_event) {
    SNB_ARG(lj, "lj_set_pair_events");
    SNB_ARG((start_event == nullptr) == (stop_event == nullptr), "both timing events or neither");
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
                     need_virial > 0 ? d_atom_lj_virial : nullptr, energy ? d_atom_direct_pme_energy : nullptr, s, 0,
                     0, lj->d_sc);
}

extern "C" int sponge_lj_force_with_pme_direct_force(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                                    const float* d_charge, float* d_frc, float pme_beta, void* stream) {
    SNB_ARG(lj && nl && d_uint_crd && d_charge && d_frc, "lj_force_with_pme_direct_force");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    return snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, d_charge, pme_beta, 0,
                     &lj->ps, d_frc, nullptr, nullptr, nullptr, as_stream(stream), 0, 0, lj->d_sc);
}

extern "C" int sponge_lj_force_with_pme_direct_force_records(sponge_lj* lj, const sponge_nl* nl, int atom_numbers,
                                                            const void* d_uint_crd_with_LJ, const float scaler[3],
                                                            float* d_frc, float cutoff, float pme_beta,
                                                            void* stream) {
    SNB_ARG(lj && nl && d_uint_crd_with_LJ && scaler && d_frc, "lj_force_with_pme_direct_force_records");
    SNB_ARG(nl->n == lj->n, "atom count of LJ and neighbor list differ");
    SNB_ARG(atom_numbers == lj->n,
            "atom_numbers must be the module's atom count (partial ranges are used by SITS, not provided here)");
    SNB_ARG(!lj->d_sc, "records carry plain LJ types; soft-core handles take sponge_lj_pme_direct_force");
    SNB_ARG(cutoff > 0, "cutoff");
    hipStream_t s = as_stream(stream);
    int rc = scratch_ensure(&lj->ps, nl);
    if (rc) return rc;
    gather_records_kernel<<<ceil_div(lj->ps.cap, 256), 256, 0, s>>>(
        nl->flags, nl->s2o, reinterpret_cast<const unsigned int*>(d_uint_crd_with_LJ), lj->ps.pk);
    SNB_CHECK_LAUNCH();
    return snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, cutoff, scaler, nullptr, nullptr, pme_beta, 0, &lj->ps,
                     d_frc, nullptr, nullptr, nullptr, s, 1, 0, nullptr);
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
                       &lj->ps, nullptr, nullptr, nullptr, nullptr, s, 0, 0, lj->d_sc);
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

// ---------------------------------------------------------------------------
// Batched launches (one per stage over every system of a batch).
int snb_batch_prep(sponge_batch* b, int energy, int from_crd, int overwrite, hipStream_t s) {
    step_prep_batch<<<b->nblk[kStPrep], 256, 0, s>>>(b->d, b->map[kStPrep], energy, from_crd, overwrite);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_batch_pairs(sponge_batch* b, int energy, int overwrite, hipStream_t s) {
    const int nmap = b->nblk[kStPair];
    const int grid = xcd_grid(nmap);
#define SNB_PAIRB(E, S)                                                                                          \
    do {                                                                                                         \
        if (b->max_nt <= 4)                                                                                      \
            pair_batch<E, S, 2><<<grid, kPairWaves * 64, b->lds_pair, s>>>(b->d, b->map[kStPair], nmap,          \
                                                                          overwrite);                           \
        else                                                                                                     \
            pair_batch<E, S, 6><<<grid, kPairWaves * 64, b->lds_pair, s>>>(b->d, b->map[kStPair], nmap,          \
                                                                          overwrite);                           \
    } while (0)
    if (energy) {
        if (b->soft) SNB_PAIRB(true, true);
        else SNB_PAIRB(true, false);
    } else {
        if (b->soft) SNB_PAIRB(false, true);
        else SNB_PAIRB(false, false);
    }
#undef SNB_PAIRB
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_batch_excluded(sponge_batch* b, hipStream_t s) {
    excluded_batch<<<b->nblk[kStExcl], 256, 0, s>>>(b->d, b->map[kStExcl]);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

int snb_lj_scratch(sponge_lj* lj, const sponge_nl* nl) { return scratch_ensure(&lj->ps, nl); }

