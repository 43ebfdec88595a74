// Internal definitions shared by the HIP translation units of libsponge_nb.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include <cstdint>
#include <cstdio>

#include "../../include/sponge_nb.h"
#include "../../include/sponge_md.h"

namespace snb {

void set_error(const char* fmt, ...);

#define SNB_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::snb::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                  \
            return SPONGE_NB_EHIP;                                                 \
        }                                                                          \
    } while (0)

#define SNB_CHECK_LAUNCH() SNB_HIP(hipGetLastError())

#define SNB_ARG(cond, msg)                   \
    do {                                         \
        if (!(cond)) {                           \
            ::snb::set_error("bad argument: %s", msg); \
            return SPONGE_NB_EARG;               \
        }                                        \
    } while (0)

constexpr int kWave = 64;
constexpr int kClusterSize = 4;          // atoms per cluster (4x4 cluster pairs, 4 per wave64)
// Slot records copied on either side of the record array (periodic images of
// the first and last kPkHalo slots): a 16-bit partner offset from any slot
// then reads a record inside the allocation.
constexpr int kPkHalo = 32768;

// 16-bit offset of slot j from slot i in the per-slot lists: j - i, or the
// offset of j's periodic image (modulo the slot count nsl) when j - i does
// not fit; false when neither fits. i + offset always indexes the slot
// records or their copies kPkHalo to either side.
__device__ __forceinline__ bool slot_offset(int i, int j, int nsl, int& off) {
    int d = j - i;
    if (d > 32767) d -= nsl;
    else if (d < -32768) d += nsl;
    off = d;
    return d >= -32768 && d <= 32767;
}
// Intervals of the tabulated direct Coulomb force factor
// G(r) = erfc(beta r) / r^3 + 2 beta / sqrt(pi) exp(-beta^2 r^2) / r^2 over
// [0, rc]: one float4 of cubic Hermite coefficients each (16 KB, held in LDS
// by the per-slot force loop).
constexpr int kCoulTab = 1024;
constexpr int kMaxLjTypes = 64;          // LDS pair table holds kMaxLjTypes^2 (A,B)
constexpr int kNumXcd = 8;               // MI355X: 8 XCDs, workgroups dealt round-robin
// Largest PME grid axis: sponge_pme_create rejects larger axes, and the prep
// kernels' per-x-plane LDS histograms hold this many planes
constexpr int kMaxGridAxis = 512;

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// blockIdx -> logical block so that each XCD gets one contiguous range of
// logical blocks (workgroup b is dispatched to XCD b % 8). Launch with
// gridDim.x a multiple of 8; returns -1 for padding blocks.
__device__ __forceinline__ int xcd_block(int nlogical) {
    int per = gridDim.x / kNumXcd;
    int l = (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
    return l < nlogical ? l : -1;
}
static inline int xcd_grid(int nlogical) { return ceil_div(nlogical, kNumXcd) * kNumXcd; }

// ---- wave64 lane exchange ------------------------------------------------
// DPP quad permutations (within groups of 4 lanes).
__device__ __forceinline__ float dpp_quad_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_quad_xor2(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
}
// Rotate right within each row of 16 lanes.
template <int N>
__device__ __forceinline__ float dpp_row_ror(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + N, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// FP32 squared distance (dx*dx + dy*dy) + dz*dz with every operation rounded
// separately, in this order. This is the neighbor list's distance test (and
// the pair kernel's cutoff test); the oracle evaluates the identical
// expression, so it must never be contracted into FMAs.
__device__ __forceinline__ float dr2_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    float s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// Block-wide FP64 sum; every thread must call it. Result valid in thread 0.
template <int NT>
__device__ double block_sum(double v, double* smem /* NT/64 */) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) smem[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < NT / 64; k++) r += smem[k];
    __syncthreads();
    return r;
}

}  // namespace snb

// ---- module objects --------------------------------------------------------
// Cluster-ordered per-step buffers of one pair-kernel user.
struct PairScratch {
    int cap;                     // slots (the list's slot capacity)
    uint4* pk;                   // packed (ux|t, uy|t, uz|t, q): LJ type in the low bits
    uint4* pk_alloc;             // allocation holding pk with kPkHalo records on either side
                                 // (NULL: pk is a caller's buffer without them)
    double* acc;                 // SPONGE_NB_NPARTS FP64 totals
    hipEvent_t ev_start, ev_stop;  // optional, recorded around the pair kernel
    int waves;                   // i-clusters (waves) per block of the force-only pair kernel: 2, 4 or 8
                                 // (0: 4)
    float4* ctab;                // kCoulTab cubic pieces of the direct Coulomb force factor (per-slot
                                 // force loop), built for (ctab_beta, ctab_rc)
    float ctab_beta, ctab_rc;
};

// Mixed-radix pass list of one FFT axis (radix 8 first, then 4, 2, 3, 5, 7),
// radix of pass p in bits 4p..4p+3 of `code`: a scalar, so that kernels
// never index an array with the pass number (which would put the plan in
// scratch memory).
struct FftPlan {
    int K;
    int npass;
    unsigned long long code;
};

struct sponge_nl {
    int n;
    float box[3];
    float cutoff, skin;
    int refresh_interval, refresh_count;
    int max_atom_in_grid_numbers, max_neighbor_numbers;
    float skin_permit;          // internal value (already doubled like neighbor_list.cu:610)
    float rl2;                  // (cutoff+skin)^2 in FP32
    float half_skin_square;     // 0.25 skin^2
    float cof_q[3];             // 0.25f * 2^32 / L
    float cof_dr[3];            // (1/2^32) * L
    // column grid used to form clusters of 4 z-sorted atoms
    int ncx, ncy, ncol;
    int col_cap;                // max atoms per column
    int slot_cap;               // max padded slots (multiple of 4)
    int ncl_cap;                // slot_cap / 4
    int list_cap;               // max cluster-pair entries per i-cluster
    int cand_cap;               // max candidate j-clusters per i-cluster
    // device buffers
    float* old_crd;             // 3n
    unsigned int* nl_uint;      // 3n, uint coordinates at the last rebuild
    int* col_count;             // ncol
    int* col_start;             // ncol+1 (slots)
    int* col_atoms;             // ncol * col_cap
    int* s2o;                   // slot_cap: slot -> atom (-1 = padding)
    int* o2s;                   // n: atom -> slot (-1: left out by a column overflow)
    int* dropped;               // n: atoms left out by a column overflow, flags[5] of them
    float4* cl_lo;              // ncl_cap bounding boxes (Angstrom)
    float4* cl_hi;
    int* list_n;                // ncl_cap
    uint2* list;                // ncl_cap * list_cap : (j cluster, 16-bit pair mask)
    uint4* pack;                // slot_cap sorted (ux, uy, uz, atom) at the last rebuild
    int* flags;                 // [0] rebuild flag, [1] overflow bits, [2] rebuild count,
                                // [3] slot count, [4] cluster count, [5] atoms left out
                                // (column overflow), [7] skin-check vote,
                                // [8..9] listed pair count (64-bit), [10] fused steps
                                // completed since the last rebuild
    int device;
    // inner list of the fused step: the entries and masks of `list` whose
    // pairs lie within prune_r2, recomputed every prune_interval fused steps
    // (0 = off) from the step's coordinates
    int prune_interval;
    float prune_r2;
    // displacement vote: the fused step's prep kernel writes its step number
    // (flags[10] + 1) into flags[13] when an atom moved more than half the
    // prune margin, (prune radius - cutoff) / 2, since the last prune; the
    // pair kernel of that step then prunes at once (a pair left out at the
    // prune was at least the prune radius apart, so it cannot be inside the
    // cutoff before some atom moved that far). flags[15] counts the prunes
    // the vote caused.
    float prune_vote_r2;
    uint4* prune_ref;           // slot_cap: slot records at the last prune
    int* list_in_n;             // ncl_cap
    uint2* list_in;             // ncl_cap * list_cap
    // the inner pairs per i slot (j slots within prune_r2) for plain LJ,
    // which force-only fused steps read (list_in is soft-core LJ's form),
    // each as the 16-bit offset j - i, taken modulo the slot count when that
    // is the shorter way (the slot records carry kPkHalo copies on either
    // side, so i + offset is always a readable record). flags[11] = 1 once a
    // slot held more than atom_in_cap partners, or a partner out of 16-bit
    // reach, at a prune since the last rebuild (its cluster then reads the
    // full list); flags[12] counts the launches that read the per-slot form
    int atom_in_cap;
    int* atom_in_n;             // slot_cap
    unsigned short* atom_in;    // slot_cap * atom_in_cap
    // outer per-slot lists (same form and cap), used when the prune radius
    // is below the list radius: every pair of the full list (within
    // rc + skin at the rebuild), written by the rebuild or by the first
    // fused step after a new prune setting (flags[14] holds that step's
    // number); the inner lists are then pruned from these
    int* atom_out_n;            // slot_cap
    unsigned short* atom_out;   // slot_cap * atom_in_cap
    // symmetric exclusion lists in original atom order (partners of atom a:
    // sym_list[sym_start[a] .. sym_start[a+1])), built on the host from the
    // exclusion CSR the first time an update passes those arrays
    const int* ex_es;
    const int* ex_el;
    const int* ex_en;
    int* sym_start;             // n + 1
    int* sym_list;              // 2 x excluded pairs
};

struct sponge_lj {
    int n, ntypes;
    float cutoff;
    float box[3];
    float cof_dr[3];
    float volume;
    double lrc_factor;           // -(2/3) pi sum_ij B_ij/6 / rc^3
    float* h_A;                  // triangular, x12
    float* h_B;                  // triangular, x6
    int* h_type;
    float2* d_ab;                // ntypes^2 (A, B) square table; A < 0 marks a soft-core pair type
    float4* d_sc;                // soft-core handles: 2 ntypes^2 per-pair-type records, else NULL
    int* d_type;                 // n (soft-core handles: combined (type A, type B, mask) index)
    PairScratch ps;              // sized from the neighbor list on first use
    void* adj;                   // adjoint scratch, allocated by the first gradient call
    // soft-core handles: raw end-state records of every combined pair type,
    // 2 ntypes^2 float4 (AA, AB, BA, BB), (soft, same mask, 0, 0), and the
    // parameters the lambda-window evaluators need
    float4* d_fe;
    double fep_lambda, fep_alpha, fep_p, fep_sig6, fep_sig6min;
    double lrc_A, lrc_B;         // long-range factors of state A (sum AB) and B (sum BB)
    void* fep;                   // frame-evaluator scratch, allocated by the first call
    int conv_split;              // 2D transforms as line launches (large planes)
};

struct sponge_pme {
    int n;
    float box[3];
    float cutoff, tolerance, beta;
    int K[3];
    int nfft;                    // K0*K1*(K2/2+1)
    float* d_bc;                 // nfft, stored [kz][x][y] like FQ
    float2* d_tw[3];             // twiddles exp(-2 pi i j / K) per axis
    float2* FQ;                  // nfft half spectrum, stored [kz][x][y]
    float* R;                    // K^3 convolved grid
    unsigned* base;              // n grid base indices, x | y << 10 | z << 20
    int* bin_count;              // K0 atoms filed under each x base plane
    int* bin_atoms;              // K0 * bin_cap atom indices
    unsigned* bin_base;          // K0 * bin_cap: the filed atoms' base and
    float4* bin_frac;            //   fraction records, in bin order
    int bin_cap;
    float4* frac;                // n fractional parts
    float* corr_atom;            // n correction energies
    double* acc;                 // SPONGE_NB_NPARTS
    double neutralizing_factor;  // -pi / (2 beta^2 V)
    PairScratch ps;              // direct-energy query (Get_Energy DIRECT part)
    int* zero_type;              // n zeros: every atom of LJ type 0
    float2* zero_ab;             // one (A, B) = (0, 0) entry
    hipStream_t side;            // reciprocal pipeline of the fused step runs here
    hipEvent_t fork, join;       // concurrently with the pair kernel
    void* fep;                   // frame-evaluator scratch, allocated by the first call
    int conv_split;              // 2D transforms as line launches (large planes)
    hipEvent_t ev[4];            // optional: recorded by the spread and gather launches
    int step_form;               // fused step: 0 pair kernel on a second stream, 1 every kernel
                                 // on its own
    int conv_threads;            // threads of the one-workgroup-per-plane convolution: 256 or 512
};

// ---- batched systems ------------------------------------------------------
// Caller-owned device arrays of one system of a batch (MD_INFORMATION's).
using BatchIo = sponge_batch_io;

// Everything a batched kernel needs of one system: the module structs by
// value (their device pointers and scalars) plus derived launch constants.
struct BatchSys {
    sponge_nl nl;
    sponge_lj lj;
    sponge_pme pme;
    BatchIo io;
    FftPlan px, py, pz;
    float rlm;                  // list radius with the bounding-box margin
    float thresh;               // skin-check threshold (squared)
    float cq[3];                // MD_core's 0.25 * 2^32 / L
    double lrc;                 // lrc_factor / volume
};

enum BatchStage {
    kStPrep, kStPair, kStExcl, kStFwd, kStConv, kStInv, kStGather, kStFin,
    kStNlCheck, kStNlBegin, kStNlBin, kStNlScan, kStNlSort, kStNlPairs, kNumStages
};

struct sponge_batch {
    int nsys;
    std::vector<sponge_nl*> nl;
    std::vector<sponge_lj*> lj;
    std::vector<sponge_pme*> pme;
    std::vector<BatchSys> h;    // host copy of d
    BatchSys* d;
    int2* map[kNumStages];      // block -> (system, block within the system); x = -1 past the end
    int nblk[kNumStages];
    int max_nt, soft;
    size_t lds_pair, lds_fwd, lds_conv, lds_inv;
    int refresh_interval, refresh_count;
    // integrator (sponge_batch_set_integrator): handles and arrays per
    // system, their device records, and the block map of its launch
    std::vector<sponge_settle*> settle;
    std::vector<sponge_batch_md_io> md_io;
    void* d_md = nullptr;
    int2* md_map = nullptr;
    int md_blocks = 0;
    int* d_ident = nullptr;     // 0 .. n-1: the atom list of systems without rigid groups
    int ident_n = 0;
    // the force step's reciprocal stages run here, beside the pair stage
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// Inner-list arguments of the pair kernel (interval 0: no inner list).
struct PruneArgs {
    int interval = 0;
    float r2 = 0.f;
    int* in_n = nullptr;
    uint2* in_list = nullptr;
    int* atom_n = nullptr;      // per-slot inner lists (NULL: cluster form only)
    unsigned short* atom_list = nullptr;  // 16-bit slot offsets
    int atom_cap = 0;
    int* diag = nullptr;        // flags + 11: [0] per-slot overflow, [1] per-slot launches,
                                // [2] step of the last displacement vote, [3] step whose
                                // per-slot lists are rewritten (a new prune setting),
                                // [4] voted prunes
    uint4* ref = nullptr;       // slot records at the last prune
    // two-level per-slot lists (atom_olist non-NULL): the outer lists hold
    // every pair of the full list, and a prune of the inner lists filters
    // them instead of the cluster-pair list
    int* atom_on = nullptr;
    unsigned short* atom_olist = nullptr;
    const float4* ctab = nullptr;  // the Coulomb table (per-slot force loop), NULL: the formula
    float ctab_scale = 0.f;        // intervals per Angstrom
    int pk_records = 0;            // the slot records with their copies on either side
};

// Excluded-pair correction folded into the pair kernel (start NULL: off):
// the symmetric exclusion lists and the atom -> slot map of the list.
struct ExclArgs {
    const int* start;
    const int* list;
    const int* o2s;
};

// Internal entry points shared between translation units.
int snb_pairs(const sponge_nl* nl, const float2* ab, int nt, const int* d_type, float cutoff, const float* cof,
              const unsigned int* d_uint_crd, const float* d_charge, float beta, int energy, PairScratch* ps,
              float* d_frc, float* atom_e, float* atom_v, float* atom_d, hipStream_t s, int gathered = 0,
              int overwrite = 0, const float4* sc = nullptr, int inner = 0,
              ExclArgs ex = ExclArgs{nullptr, nullptr, nullptr});
void snb_scratch_free(PairScratch* ps);
// Cluster-order records (uint coordinates, charge, type) into ps->pk.
int snb_gather(const sponge_nl* nl, const unsigned int* d_uint_crd, const float* d_charge, const int* d_type,
               PairScratch* ps, hipStream_t s);
void snb_fep_free(void* fep_scratch);
int snb_excluded(sponge_pme* pme, const unsigned int* d_uint_crd, const float scaler[3], const float* d_charge,
                 const int* es, const int* el, const int* en, float* d_frc, float* d_atom_energy, hipStream_t s);
int snb_pme_reciprocal(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge, float* d_frc,
                       int energy, int prep_done, hipStream_t s);
int snb_pme_gather(sponge_pme* pme, float* d_frc, hipStream_t s, int* step_counter = nullptr);
int snb_pme_spread(sponge_pme* pme, int energy, hipStream_t s);
int snb_pme_convolve(sponge_pme* pme, int energy, hipStream_t s);
int snb_step_prep(const sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const float* d_crd, unsigned int* d_uint_crd,
                  const float* d_charge, int energy, float* d_frc_assigned, hipStream_t s);
int snb_pme_finalize(sponge_pme* pme, const float* d_direct_atom_energy, int need_energy, int need_virial,
                     float* d_potential, float* d_virial, hipStream_t s);
int snb_fill_zero_f64(double* p, int n, hipStream_t s);
void snb_pme_plans(const sponge_pme* pme, FftPlan* px, FftPlan* py, FftPlan* pz);
int snb_pme_row_blocks(const sponge_pme* pme);
size_t snb_pme_lds_fwd(const sponge_pme* pme);
size_t snb_pme_lds_conv(const sponge_pme* pme);
int snb_pme_raise_lds(size_t bytes);
size_t snb_pme_lds_inv(const sponge_pme* pme);
float snb_nl_rlm(const sponge_nl* nl);
float snb_prune_vote_r2(float cutoff, float prune_radius);
// Two-level per-slot lists: an inner prune radius below the list radius.
inline bool snb_two_level(const sponge_nl* nl) { return nl->prune_interval > 0 && nl->prune_r2 < nl->rl2; }
// The per-slot lists of every listed pair, which each rebuild writes after
// the list (the outer ones of the two-level form); NULL while the inner list
// is off.
inline unsigned short* snb_nl_slot_lists(const sponge_nl* nl) {
    return nl->prune_interval > 0 ? (snb_two_level(nl) ? nl->atom_out : nl->atom_in) : nullptr;
}
inline int* snb_nl_slot_counts(const sponge_nl* nl) {
    return nl->prune_interval > 0 ? (snb_two_level(nl) ? nl->atom_out_n : nl->atom_in_n) : nullptr;
}
int snb_batch_prep(sponge_batch* b, int energy, int from_crd, int overwrite, hipStream_t s);
int snb_batch_pairs(sponge_batch* b, int energy, int overwrite, hipStream_t s);
int snb_batch_excluded(sponge_batch* b, hipStream_t s);
int snb_lj_scratch(sponge_lj* lj, const sponge_nl* nl);
void snb_adj_free(sponge_lj* lj);
int snb_batch_reciprocal(sponge_batch* b, int energy, hipStream_t s);
int snb_batch_gather(sponge_batch* b, hipStream_t s);
int snb_batch_finalize(sponge_batch* b, hipStream_t s);
int snb_pme_prep(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge, hipStream_t s);

int snb_pme_adjoint_grid(sponge_pme* pme, const float4* wv, const float* scale, hipStream_t s);
int snb_pme_adj_gather(sponge_pme* pme, int mode, const float4* wv, float sign, float* out, hipStream_t s);
int snb_batch_nl(sponge_batch* b, int forced, hipStream_t s);
int snb_batch_md_sync(sponge_batch* b);
