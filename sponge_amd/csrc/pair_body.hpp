// Device side of the LJ + PME direct pair kernel (LENNARD_JONES_INFORMATION,
// Lennard_Jones_force/Lennard_Jones_force.cu:424-562; soft-core terms of
// LJ_soft_core.cu:254-300; the folded exclusion correction of
// PME_force.cu:620-695): one i-cluster per wave (pair_wave), the kernels'
// blocks (pair_body).
//
// Layout: one wave64 per 4-atom i-cluster; each iteration takes four
// j-clusters from the cluster-pair list, lane = (g:2, jj:2, ii:2) holds the
// pair (ii, jj) of j-cluster g. Coordinates, charges and LJ types are read as
// one 16-byte record per slot. The list names every neighbouring cluster
// from both sides, so each wave only accumulates the force on its own four
// atoms and writes it with plain read-modify-writes: no atomics.
#pragma once
#include "common.hpp"
#include "pair_math.hpp"

namespace snb {

constexpr int kPairWaves = 4;

// Soft-core pair terms (LJ_soft_core.cu:254-300): with s^6 = r^6 + c per end
// state, F = (f_LJ - f_C) dr where
//   f_LJ = r^4 [(l_AB - l_AA s_A^-6) s_A^-12 + (lB_B - lB_A s_B^-6) s_B^-12]
//   f_C  = q_i q_j r^4 [l_ (2b/sqrt(pi) e^{-b^2 s_A^2} + erfc(b s_A)/s_A) s_A^-6 + (B alike)]
// s0 = {l_ AA, l_ AB, lam BA, lam BB} (l_ = 1 - lambda), s1 = {c_A, c_B, l_, lam}.
// s = (s^6)^(1/6) is taken as exp2(log2(s^6) / 6).
__device__ __forceinline__ void soft_pair(float r2, float qq, float beta, float mlog2e_b2, float4 s0, float4 s1,
                                          float& flj, float& fc) {
    const float r4 = r2 * r2, r6 = r4 * r2;
    const float SA6 = r6 + s1.x, SB6 = r6 + s1.y;
    const float iA6 = __builtin_amdgcn_rcpf(SA6), iB6 = __builtin_amdgcn_rcpf(SB6);
    flj = r4 * ((s0.y - s0.x * iA6) * iA6 * iA6 + (s0.w - s0.z * iB6) * iB6 * iB6);
    const float sA = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(SA6) * (1.0f / 6.0f));
    const float sB = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(SB6) * (1.0f / 6.0f));
    const float isA = __builtin_amdgcn_rcpf(sA), isB = __builtin_amdgcn_rcpf(sB);
    const float exA = __builtin_amdgcn_exp2f(mlog2e_b2 * (sA * sA));
    const float exB = __builtin_amdgcn_exp2f(mlog2e_b2 * (sB * sB));
    const float tA = fmaf(exA * erfcx_poly(beta * sA), isA, kTwoOverSqrtPi * beta * exA);
    const float tB = fmaf(exB * erfcx_poly(beta * sB), isB, kTwoOverSqrtPi * beta * exB);
    fc = qq * r4 * (s1.z * tA * iA6 + s1.w * tB * iB6);
}

// FP64 energy-step terms of one soft-core pair from the exact displacement:
// E_LJ, E_C and the LJ virial -f_LJ r^2 (LJ_soft_core.cu:434-473, 598-641).
__device__ __forceinline__ void soft_pair_energy(double r2, double qq, double beta, float4 s0, float4 s1,
                                                 double& elj, double& edir, double& vlj) {
    const double r4 = r2 * r2, r6 = r4 * r2;
    const double SA6 = r6 + (double)s1.x, SB6 = r6 + (double)s1.y;
    const double iA6 = 1.0 / SA6, iB6 = 1.0 / SB6;
    const double sA = cbrt(sqrt(SA6)), sB = cbrt(sqrt(SB6));
    elj = ((double)s0.x * (1.0 / 12.0) * iA6 - (double)s0.y * (1.0 / 6.0)) * iA6 +
          ((double)s0.z * (1.0 / 12.0) * iB6 - (double)s0.w * (1.0 / 6.0)) * iB6;
    edir = qq * ((double)s1.z * erfc(beta * sA) / sA + (double)s1.w * erfc(beta * sB) / sB);
    const double flj = r4 * (((double)s0.y - (double)s0.x * iA6) * iA6 * iA6 +
                             ((double)s0.w - (double)s0.z * iB6) * iB6 * iB6);
    vlj = -flj * r2;
}

// PME_Excluded_Force_With_Atom_Energy's force (PME_force.cu:620-695) on one
// atom (slot record a) from its exclusion partners k0, k0 + kstep, ...: each
// atom collects its own side of every excluded pair, so no other atom's force
// is written.
__device__ __forceinline__ void excluded_partners(const ExclArgs& ex, const uint4* __restrict__ pk, int at, int k0,
                                                  int kstep, unsigned ax, unsigned ay, unsigned az, float qi,
                                                  float cx, float cy, float cz, float beta, float& fix, float& fiy,
                                                  float& fiz) {
    const int r1 = ex.start[at + 1];
    for (int k = ex.start[at] + k0; k < r1; k += kstep) {
        // a partner left out of an overflowing list has no slot
        const int js = ex.o2s[ex.list[k]];
        if (js < 0) continue;
        const uint4 b = pk[js];
        const float dx = (float)(int)((b.x & ~3u) - ax) * cx;
        const float dy = (float)(int)((b.y & ~3u) - ay) * cy;
        const float dz = (float)(int)((b.z & ~3u) - az) * cz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float r = sqrtf(r2);
        const float br = beta * r;
        const float qq = qi * __uint_as_float(b.w);
        const float fa = -qq * ((kTwoOverSqrtPi * br * expf(-br * br) + erfcf(br) - 1.0f) / r2 / r);
        fix += fa * dx;
        fiy += fa * dy;
        fiz += fa * dz;
    }
}

// Per-slot lists of cluster ic: every pair the full list holds (its masks:
// the pairs within rc + skin at the rebuild, exclusions left out), per i slot
// in list order (lanes ii, ii + 4, ... of a round in lane order), into
// `lists` / `counts`; no slot record is read. Returns the partner count of
// slot lane >> 4 (the lane mapping of pair_atoms); a count above atom_cap
// (partners past the cap dropped, or a partner beyond 16-bit reach) sends the
// cluster to the full list until the lists are written again. The list
// stores are complete and visible to the whole wave on return.
__device__ __forceinline__ int slot_lists(int ic, int nsl, const int* __restrict__ list_n,
                                          const uint2* __restrict__ list, int list_cap, const PruneArgs& pa,
                                          unsigned short* __restrict__ lists, int* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const int n = list_n[ic];
    const uint2* lst = list + (size_t)ic * list_cap;
    unsigned short* out = lists + (size_t)islot * pa.atom_cap;
    const unsigned long long sel = 0x1111111111111111ull << ii, below = (1ull << lane) - 1ull;
    int count = 0;
    bool far = false;
    for (int base = 0; base < n; base += 64) {
        const uint2 my = base + lane < n ? lst[base + lane] : make_uint2((unsigned)ic, 0u);
        const int cnt = min(64, n - base);
        // four entries per round, lane (g, jj, ii) holding pair (ii, jj) of
        // entry g; the lane exchanges of four rounds are issued together
        for (int it = 0; it < cnt; it += 16) {
            int jcs[4];
            unsigned ms[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = min(it + 4 * u + g, 63);
                jcs[u] = __shfl((int)my.x, e);
                ms[u] = (unsigned)__shfl((int)my.y, e);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int jc = jcs[u];
                const bool keep = it + 4 * u + g < cnt && ((ms[u] >> bit) & 1u);
                const unsigned long long mi = __ballot(keep) & sel;
                const int pos = count + __popcll(mi & below);
                // below 32,768 slots every offset fits without the periodic
                // image (a wave-uniform branch)
                int off = jc * kClusterSize + jj - islot;
                bool reach = true;
                if (nsl > 32767) reach = slot_offset(islot, jc * kClusterSize + jj, nsl, off);
                far = far || (keep && !reach);
                if (keep && pos < pa.atom_cap) out[pos] = (unsigned short)off;
                count += __popcll(mi);
            }
        }
    }
    if (__ballot(far) & sel) count = pa.atom_cap + 1;
    if (lane < kClusterSize) {
        counts[islot] = count;
        if (count > pa.atom_cap) atomicOr(pa.diag, 1);
    }
    // other lanes of the wave read these lists next: the wave's stores must
    // have completed (one CU, so the workgroup scope suffices; the lists of
    // different slots share no cache line)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    return __shfl(count, lane >> 4);
}

// Inner lists of the two-level form from the outer ones, without forces
// (energy steps, whose forces come from the full list): the partners of each
// slot of cluster ic within pa.r2 at this step's positions, in outer-list
// order, with their counts, as pair_atoms_filter writes them. Lane (i, k), i = lane >> 4, takes
// outer partners 2k and 2k + 1 of every 32 of atom i (one dword of offsets;
// two dwords in flight per lane); n = the outer count of slot i, nmax the
// longest of the four (at most atom_cap, so every offset read lies in the
// slot's list). A partner 2k + h of a round lands after those of the lanes
// below k, h = 0 before h = 1. Returns the inner count of slot i; the stores
// are complete and visible to the whole wave on return.
__device__ __forceinline__ int filter_slots(int ic, int n, int nmax, const uint4* __restrict__ pk, float cx,
                                            float cy, float cz, const PruneArgs& pa) {
    const int lane = threadIdx.x & 63, ii = lane >> 4, k = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const unsigned* __restrict__ lw =
        reinterpret_cast<const unsigned*>(pa.atom_olist + (size_t)islot * pa.atom_cap);
    unsigned short* __restrict__ out = pa.atom_list + (size_t)islot * pa.atom_cap;
    const unsigned below = (1u << k) - 1u;
    const int nwords = (nmax + 1) >> 1;
    int count = 0;
    for (int t = 0; t < nmax; t += 64) {
        unsigned w[2];
        uint4 b[4];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int kw = (t >> 1) + 16 * u + k;
            w[u] = kw < nwords ? lw[kw] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            b[2 * u] = pk[islot + (int)(short)(w[u] & 0xffffu)];
            b[2 * u + 1] = pk[islot + ((int)w[u] >> 16)];
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            bool keep[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint4 r = b[2 * u + h];
                const float dx = (float)(int)((r.x & ~3u) - ax) * cx;
                const float dy = (float)(int)((r.y & ~3u) - ay) * cy;
                const float dz = (float)(int)((r.z & ~3u) - az) * cz;
                keep[h] = t + 32 * u + 2 * k + h < n && dr2_exact(dx, dy, dz) < pa.r2;
            }
            const unsigned m0 = (unsigned)(__ballot(keep[0]) >> (16 * ii)) & 0xffffu;
            const unsigned m1 = (unsigned)(__ballot(keep[1]) >> (16 * ii)) & 0xffffu;
            const int pos = count + __popc(m0 & below) + __popc(m1 & below);
            if (keep[0]) out[pos] = (unsigned short)(w[u] & 0xffffu);
            if (keep[1]) out[pos + (keep[0] ? 1 : 0)] = (unsigned short)(w[u] >> 16);
            count += __popc(m0) + __popc(m1);
        }
    }
    if (k == 0) pa.atom_n[islot] = count;
    // other lanes read these lists next (as in slot_lists)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    return count;
}

// Slot records of the per-slot force loops through a structured buffer
// resource (16-byte records) over the records and their copies on either
// side: a partner's record is one indexed buffer load at the slot's own
// index plus the sign-extended 16-bit offset (one SDWA add to decode; no
// 64-bit address per pair), and an index outside the allocation reads zeros.
typedef unsigned SlotRecordBits __attribute__((ext_vector_type(4)));
__device__ SlotRecordBits snb_struct_buffer_load_b128(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                                      int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");

struct SlotRecords {
    __amdgpu_buffer_rsrc_t rsrc;
    int self;  // index of the slot's own record
};

__device__ __forceinline__ SlotRecords slot_records(const uint4* pk, int islot, const PruneArgs& pa) {
    SlotRecords r;
    r.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(pk - kPkHalo), (short)sizeof(uint4), pa.pk_records,
                                               0x00020000);
    r.self = islot + kPkHalo;
    return r;
}

// The record at a signed 16-bit slot offset (bits 0-15 or 16-31 of w).
template <bool HIGH>
__device__ __forceinline__ uint4 slot_record(const SlotRecords& r, unsigned w) {
    const int off = HIGH ? (int)w >> 16 : (int)(short)(w & 0xffffu);
    const SlotRecordBits v = snb_struct_buffer_load_b128(r.rsrc, r.self + off, 0, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

// One listed pair of the per-slot force loop (TABLE: the Coulomb force
// factor from its cubic piece in LDS, else from the fitted erfc): the
// displacement and r^2 to the partner record b, and the factor f of the
// pair force f (dx, dy, dz) on atom i; 0 unless listed and r < rc. The pair
// terms are pair_body's, in FP32.
template <int TB, bool TABLE>
__device__ __forceinline__ float slot_pair(const uint4 b, unsigned ax, unsigned ay, unsigned az, float qi,
                                           const float2* srow, const float4* stab, float cx, float cy, float cz,
                                           float rc2, float beta, float mlog2e_b2, float ctab_scale, bool listed,
                                           float& dx, float& dy, float& dz, float& r2) {
    // with at most 4 types the y and z words carry no type bits
    dx = (float)(int)((b.x & ~3u) - ax) * cx;
    dy = (float)(int)((TB == 2 ? b.y : b.y & ~3u) - ay) * cy;
    dz = (float)(int)((TB == 2 ? b.z : b.z & ~3u) - az) * cz;
    r2 = dr2_exact(dx, dy, dz);
    const bool in = listed && r2 < rc2;
    // masked lanes may give inf / NaN below; the final select discards them
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    const float rinv6 = rinv2 * rinv2 * rinv2;
    const int tj = TB == 2 ? (int)(b.x & 3u) : unpack_type(b);
    const float2 AB = srow[tj];
    const float flj = (AB.y - AB.x * rinv6) * rinv6 * rinv2;
    const float qq = qi * __uint_as_float(b.w);
    float fc;
    if (TABLE) {
        // the force factor from its cubic piece on [k h, (k + 1) h]
        const float t = (r2 * rinv) * ctab_scale;
        const unsigned kk = min((unsigned)t, (unsigned)(kCoulTab - 1));
        // t - kk for t below the table's end (beyond it the pair is out)
        const float u = __builtin_amdgcn_fractf(t);
        const float4 c = stab[kk];
        fc = qq * fmaf(fmaf(fmaf(c.w, u, c.z), u, c.y), u, c.x);
    } else {
        const float x = beta * (r2 * rinv);
        const float ex_ = __builtin_amdgcn_exp2f(mlog2e_b2 * r2);
        const float ecx = erfcx_poly(x);
        fc = qq * ex_ * fmaf(kTwoOverSqrtPi, x, ecx) * rinv * rinv2;
    }
    return in ? flj - fc : 0.f;
}

// End of the per-slot force loop: the folded exclusion corrections of atom i
// (lane k takes partners k, k + 16, ...), the 16-lane sum and the store.
__device__ __forceinline__ void slot_finish(int islot, int k, const uint4* __restrict__ pk, unsigned ax, unsigned ay,
                                            unsigned az, float qi, float cx, float cy, float cz, float beta,
                                            const int* __restrict__ s2o, float* __restrict__ frc, int overwrite,
                                            const ExclArgs& ex, float fix, float fiy, float fiz) {
    const int at = s2o[islot];
    if (ex.start && at >= 0) excluded_partners(ex, pk, at, k, 16, ax, ay, az, qi, cx, cy, cz, beta, fix, fiy, fiz);
    // i force: sum the 16 lanes of atom i
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        fix += __shfl_xor(fix, o);
        fiy += __shfl_xor(fiy, o);
        fiz += __shfl_xor(fiz, o);
    }
    if (k == 0 && at >= 0 && frc) {
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
}

// Force-only steps of plain LJ from the per-slot lists: lane (i, k),
// i = lane >> 4, takes partners 2k and 2k + 1 of every 32 of atom i of the
// wave's cluster (n of them; nmax the longest of the four lists) from one
// dword of two 16-bit slot offsets, so every lane holds listed pairs (the
// cluster form evaluates all 16 pairs of each listed cluster pair, in cutoff
// or not). A dword of offsets is fetched two rounds ahead and the first
// record of a round one round ahead.
// Lanes past the end of a list read a record
// too (offsets left from earlier prunes, or zero: always inside the records
// and their copies) and are masked out.
template <int TB, bool TABLE>
__device__ __forceinline__ void pair_atoms(int ic, int n, int nmax, const uint4* __restrict__ pk, const float2* sab,
                                           const float4* stab, int nt, float cx, float cy, float cz, float rc2,
                                           float beta, const int* __restrict__ s2o, float* __restrict__ frc,
                                           int overwrite, const PruneArgs& pa, const ExclArgs& ex) {
    const int lane = threadIdx.x & 63, ii = lane >> 4, k = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const float qi = __uint_as_float(a.w);
    const int ti = TB == 2 ? (int)(a.x & 3u) : unpack_type(a);
    const float2* srow = sab + ti * nt;
    const float mlog2e_b2 = -1.4426950408889634f * (beta * beta);
    const unsigned* __restrict__ lw =
        reinterpret_cast<const unsigned*>(pa.atom_list + (size_t)islot * pa.atom_cap);
    const int nwords = (nmax + 1) >> 1;  // atom_cap is even: the dword stays inside the slot's list
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    const SlotRecords rec = slot_records(pk, islot, pa);
    const int nk = n - 2 * k;  // partner t + 2k + h is listed when t + h < nk
    // One round of 32 partners per atom: pairs (b0, b1) from offset dword
    // w; the next round's first record bn and the offsets two rounds ahead
    // are fetched first. Rounds go in pairs so that the registers rotate
    // once per two rounds.
    auto round = [&](int t, unsigned w_, unsigned wn_, const uint4& b0_, unsigned& w2_, uint4& bn_) {
        const int k2 = (t >> 1) + 32 + k;
        w2_ = k2 < nwords ? lw[k2] : 0u;
        const uint4 b1 = slot_record<true>(rec, w_);
        bn_ = slot_record<false>(rec, wn_);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            float dx, dy, dz, r2;
            const float f = slot_pair<TB, TABLE>(h ? b1 : b0_, ax, ay, az, qi, srow, stab, cx, cy, cz, rc2, beta,
                                                 mlog2e_b2, pa.ctab_scale, t + h < nk, dx, dy, dz, r2);
            fix += f * dx;
            fiy += f * dy;
            fiz += f * dz;
        }
    };
    // offsets two rounds ahead, the first record of a round one round ahead
    unsigned w = k < nwords ? lw[k] : 0u;
    unsigned wn = 16 + k < nwords ? lw[16 + k] : 0u;
    uint4 b0 = slot_record<false>(rec, w);
    for (int t = 0; t < nmax; t += 64) {
        unsigned w2, w3;
        uint4 bn, bnn;
        round(t, w, wn, b0, w2, bn);
        if (t + 32 >= nmax) break;
        round(t + 32, wn, w2, bn, w3, bnn);
        w = w2;
        wn = w3;
        b0 = bnn;
    }
    slot_finish(islot, k, pk, ax, ay, az, qi, cx, cy, cz, beta, s2o, frc, overwrite, ex, fix, fiy, fiz);
}

// The same force loop over the outer lists of the two-level form that also
// writes the inner lists: the partners within pa.r2 at this step's
// positions, with their counts, in the order the loop meets them (per round
// of 32 outer partners, the even ones 2k in k order, then the odd ones
// 2k + 1). Nothing is
// fetched ahead across rounds (the kernel keeps its 64 VGPRs; the other waves
// of the SIMD cover the latency).
template <int TB, bool TABLE>
__device__ __forceinline__ void pair_atoms_filter(int ic, int n, int nmax, const uint4* __restrict__ pk,
                                                  const float2* sab, const float4* stab, int nt, float cx, float cy,
                                                  float cz, float rc2, float beta, const int* __restrict__ s2o,
                                                  float* __restrict__ frc, int overwrite, const PruneArgs& pa,
                                                  const ExclArgs& ex) {
    const int lane = threadIdx.x & 63, ii = lane >> 4, k = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const float qi = __uint_as_float(a.w);
    const int ti = TB == 2 ? (int)(a.x & 3u) : unpack_type(a);
    const float2* srow = sab + ti * nt;
    const float mlog2e_b2 = -1.4426950408889634f * (beta * beta);
    const unsigned lo = (unsigned)islot * (unsigned)pa.atom_cap;  // below 2^32: the lists' allocation
    const unsigned* __restrict__ lw = reinterpret_cast<const unsigned*>(pa.atom_olist) + (lo >> 1);
    const int nwords = (nmax + 1) >> 1;
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    const SlotRecords rec = slot_records(pk, islot, pa);
    int count = 0;
    for (int t = 0; t < nmax; t += 32) {
        const int kw = (t >> 1) + k;
        const unsigned w = kw < nwords ? lw[kw] : 0u;
        const uint4 b0 = slot_record<false>(rec, w);
        const uint4 b1 = slot_record<true>(rec, w);
        const unsigned below = (1u << k) - 1u;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            float dx, dy, dz, r2;
            const bool listed = t + 2 * k + h < n;
            const float f = slot_pair<TB, TABLE>(h ? b1 : b0, ax, ay, az, qi, srow, stab, cx, cy, cz, rc2, beta,
                                                 mlog2e_b2, pa.ctab_scale, listed, dx, dy, dz, r2);
            fix += f * dx;
            fiy += f * dy;
            fiz += f * dz;
            // atom ii's kept partners 2k + h of this round (a 16-bit mask
            // over k), after those of h = 0
            const bool keep = listed && r2 < pa.r2;
            const unsigned m = (unsigned)(__ballot(keep) >> (16 * ii)) & 0xffffu;
            if (keep) pa.atom_list[lo + (unsigned)(count + __popc(m & below))] =
                (unsigned short)(h ? w >> 16 : w & 0xffffu);
            count += __popc(m);
        }
    }
    if (k == 0) pa.atom_n[islot] = count;
    slot_finish(islot, k, pk, ax, ay, az, qi, cx, cy, cz, beta, s2o, frc, overwrite, ex, fix, fiy, fiz);
}

// Cutoff test shared with the oracle: FP32 dr2 = (dx*dx + dy*dy) + dz*dz with
// every operation rounded (the neighbor list's expression) against rc*rc in
// FP32, so both sides agree on every pair, including those within one ulp of
// the cutoff.
// SOFT: the type index is the combined (type A, type B, mask) index of a
// soft-core handle; `ab` holds the lambda-mixed (A, B) of ordinary pair
// types and A < 0 for soft-core pair types, whose records are sc[t_i nt + t_j]
// (s0) and sc[nt^2 + t_i nt + t_j] (s1), read from global memory only by
// waves that meet such a pair.
// TB = 2: every type index is below 4 and sits in the x coordinate's low bits
// alone (two instructions fewer per pair); TB = 6: the general 6-bit index.
template <bool ENERGY, bool SOFT, int TB>
__device__ __forceinline__ void pair_wave(
    int ic, const float2* sab, const float4* stab, const int* __restrict__ flags, const int* __restrict__ list_n,
    const uint2* __restrict__ list, int list_cap, const uint4* __restrict__ pk, int nt, float cx, float cy, float cz,
    float rc2, float beta, const int* __restrict__ s2o, float* __restrict__ frc, float* __restrict__ atom_e,
    float* __restrict__ atom_v, float* __restrict__ atom_d, double* __restrict__ acc, int overwrite,
    const float4* __restrict__ sc, const PruneArgs& pa, const ExclArgs& ex) {
    const int lane = threadIdx.x & 63;
    // Inner lists (pa.interval > 0). c = this step's number since the
    // rebuild (the count of completed steps is advanced by the step's
    // gather, after this kernel); a step is scheduled every pa.interval-th
    // step and on the first after a rebuild, voted when its prep kernel
    // found an atom that moved more than half the prune margin since the
    // last prune, and renewing when it is the first after a new prune
    // setting.
    // Plain LJ keeps per-slot lists (pa.atom_list; force-only steps read
    // them, energy steps the full list). One level (a prune radius at or
    // beyond rc + skin): the lists hold every pair of the full list, written
    // by each rebuild (nl_slot_lists_kernel) and on renewing steps; no other
    // step prunes and nothing is voted (the pairs are those of the list's
    // own radius at the rebuild). Two levels
    // (pa.atom_olist set): those lists are the outer ones, and the inner
    // lists, the outer pairs within pa.r2, are written on scheduled, voted
    // and renewing steps. Soft-core LJ prunes the cluster-pair list to pa.r2
    // on scheduled and voted steps and reads the pruned one in between. A
    // step that prunes by distance keeps the slot records it pruned at.
    const bool two = pa.atom_olist != nullptr;
    const bool slots = !SOFT && pa.atom_list != nullptr && pa.interval > 0;
    bool use_inner = false, write_inner = false, write_outer = false;
    if (pa.interval > 0) {
        const int c = flags[10] + 1;
        // the rebuild writes the lists of a new period; a new prune setting
        // has the next step write them
        const bool renew = pa.diag[3] == c;
        const bool scheduled = c <= 1 || (c - 1) % pa.interval == 0;
        const bool voted = pa.diag[2] == c;
        if (slots) {
            write_outer = two && renew;
            write_inner = two ? renew || scheduled || voted : renew;
        } else {
            write_inner = scheduled || voted;
        }
        use_inner = !write_inner;
        if (write_inner && lane < kClusterSize) pa.ref[ic * kClusterSize + lane] = pk[ic * kClusterSize + lane];
        if (voted && !scheduled && (!slots || two) && ic == 0 && lane == 0) atomicAdd(pa.diag + 4, 1);
    }
    if (slots && !ENERGY) {
        if (ic == 0 && lane == 0) atomicAdd(pa.diag + 1, 1);
        if (two && write_inner) {
            const int n = write_outer ? slot_lists(ic, flags[3], list_n, list, list_cap, pa, pa.atom_olist, pa.atom_on)
                                      : pa.atom_on[ic * kClusterSize + (lane >> 4)];
            const int nmax = max(max(__shfl(n, 0), __shfl(n, 16)), max(__shfl(n, 32), __shfl(n, 48)));
            if (nmax <= pa.atom_cap) {
                if (stab)
                    pair_atoms_filter<TB, true>(ic, n, nmax, pk, sab, stab, nt, cx, cy, cz, rc2, beta, s2o, frc,
                                                overwrite, pa, ex);
                else
                    pair_atoms_filter<TB, false>(ic, n, nmax, pk, sab, stab, nt, cx, cy, cz, rc2, beta, s2o, frc,
                                                 overwrite, pa, ex);
                return;
            }
            // an outer list overflowed: the inner lists follow it to the
            // full list until the outer ones are written again
            if (lane < kClusterSize) pa.atom_n[ic * kClusterSize + lane] = pa.atom_cap + 1;
        } else {
            const int n = write_inner ? slot_lists(ic, flags[3], list_n, list, list_cap, pa, pa.atom_list, pa.atom_n)
                                      : pa.atom_n[ic * kClusterSize + (lane >> 4)];
            const int nmax = max(max(__shfl(n, 0), __shfl(n, 16)), max(__shfl(n, 32), __shfl(n, 48)));
            if (nmax <= pa.atom_cap) {
                if (stab)
                    pair_atoms<TB, true>(ic, n, nmax, pk, sab, stab, nt, cx, cy, cz, rc2, beta, s2o, frc, overwrite,
                                         pa, ex);
                else
                    pair_atoms<TB, false>(ic, n, nmax, pk, sab, stab, nt, cx, cy, cz, rc2, beta, s2o, frc, overwrite,
                                          pa, ex);
                return;
            }
        }
        write_inner = false;  // a slot overflowed: this cluster's forces from the full list
        write_outer = false;
    }
    // energy steps that write the per-slot lists take every listed pair from
    // the full list (the outer ones in the two-level form, which then writes
    // the inner ones from them below)
    const bool write_atoms = slots && (two ? write_outer : write_inner);
    const bool write_cluster = !slots && write_inner;
    unsigned short* const wl = two ? pa.atom_olist : pa.atom_list;
    int* const wn = two ? pa.atom_on : pa.atom_n;
    const float wr2 = slots ? __builtin_huge_valf() : pa.r2;
    if (slots) use_inner = false;
    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4, bit = lane & 15;
    const int islot = ic * kClusterSize + ii;
    const uint4 a = pk[islot];
    const unsigned ax = a.x & ~3u, ay = a.y & ~3u, az = a.z & ~3u;
    const float qi = __uint_as_float(a.w);
    const int ti = TB == 2 ? (int)(a.x & 3u) : unpack_type(a);
    const float2* srow = sab + ti * nt;
    const float mlog2e_b2 = -1.4426950408889634f * (beta * beta);
    int ai = 0;
    if (ENERGY) ai = s2o[islot];
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    float eie = 0.f, eiv = 0.f, eid = 0.f;
    double tlj = 0.0, tdir = 0.0, tvir = 0.0, tcnt = 0.0;
    const int n = use_inner ? pa.in_n[ic] : list_n[ic];
    const uint2* lst = (use_inner ? pa.in_list : list) + (size_t)ic * list_cap;
    uint2* inner_out = write_cluster ? pa.in_list + (size_t)ic * list_cap : nullptr;
    int inner_count = 0, atom_count = 0;
    bool atom_far = false;
    const int nsl = flags[3];
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
            if (write_inner) {
                const bool keep = ((m >> bit) & 1u) && r2 < wr2;
                const unsigned long long kb = __ballot(keep);
                if (write_cluster) {
                    const unsigned mk = (unsigned)(kb >> (g * 16)) & 0xffffu;
                    const bool leader = (lane & 15) == 0 && mk != 0u;
                    const unsigned long long lb2 = __ballot(leader);
                    if (leader)
                        inner_out[inner_count + __popcll(lb2 & ((1ull << lane) - 1ull))] =
                            make_uint2((unsigned)jc, mk);
                    inner_count += __popcll(lb2);
                } else {
                    // lanes ii, ii + 4, ... append to slot ii's list in lane order
                    const unsigned long long mi = kb & (0x1111111111111111ull << ii);
                    const int pos = atom_count + __popcll(mi & ((1ull << lane) - 1ull));
                    int off = 0;
                    const bool reach = slot_offset(islot, jc * kClusterSize + jj, nsl, off);
                    atom_far = atom_far || (keep && !reach);
                    if (keep && pos < pa.atom_cap) wl[(size_t)islot * pa.atom_cap + pos] = (unsigned short)off;
                    atom_count += __popcll(mi);
                }
            }
            // masked lanes (padding, the atom itself) may give inf / NaN
            // below; the final select discards them
            const float r2s = r2;
            const float rinv = __builtin_amdgcn_rsqf(r2s);
            const float rinv2 = rinv * rinv;
            const float rinv6 = rinv2 * rinv2 * rinv2;
            const int tj = TB == 2 ? (int)(b.x & 3u) : unpack_type(b);
            const float2 AB = srow[tj];
            float flj = (AB.y - AB.x * rinv6) * rinv6 * rinv2;
            const float x = beta * (r2s * rinv);
            const float ex = __builtin_amdgcn_exp2f(mlog2e_b2 * r2s);
            const float ecx = erfcx_poly(x);
            const float qq = qi * __uint_as_float(b.w);
            float fc = qq * ex * fmaf(kTwoOverSqrtPi, x, ecx) * rinv * rinv2;
            bool sp = false;
            float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
            if (SOFT) {
                sp = in && AB.x < 0.f;
                if (sp) {
                    s0 = sc[ti * nt + tj];
                    s1 = sc[nt * nt + ti * nt + tj];
                    soft_pair(r2s, qq, beta, mlog2e_b2, s0, s1, flj, fc);
                }
            }
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
                    double elj, edir, vlj;
                    if (SOFT && sp) {
                        soft_pair_energy(dr2, (double)qi * (double)__uint_as_float(b.w), (double)beta, s0, s1, elj,
                                         edir, vlj);
                    } else {
                        const double dr = sqrt(dr2);
                        const double ir2 = 1.0 / dr2, ir6 = ir2 * ir2 * ir2;
                        elj = ((double)AB.x * (1.0 / 12.0) * ir6 - (double)AB.y * (1.0 / 6.0)) * ir6;
                        edir = (double)qi * (double)__uint_as_float(b.w) * erfc((double)beta * dr) / dr;
                        vlj = ((double)AB.x * ir6 - (double)AB.y) * ir6;
                    }
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
    if (write_cluster && lane == 0) pa.in_n[ic] = inner_count;
    if (write_atoms && (__ballot(atom_far) & (0x1111111111111111ull << ii))) atom_count = pa.atom_cap + 1;
    if (write_atoms && lane < kClusterSize) {
        wn[islot] = atom_count;
        if (atom_count > pa.atom_cap) atomicOr(pa.diag, 1);
    }
    if (ENERGY && slots && two && write_inner) {
        // the inner lists from the outer ones (no force: this step's forces
        // come from the full list)
        if (write_atoms) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const int no = pa.atom_on[ic * kClusterSize + (lane >> 4)];
        const int nomax = max(max(__shfl(no, 0), __shfl(no, 16)), max(__shfl(no, 32), __shfl(no, 48)));
        if (nomax <= pa.atom_cap)
            filter_slots(ic, no, nomax, pk, cx, cy, cz, pa);
        else if (lane < kClusterSize)
            pa.atom_n[ic * kClusterSize + lane] = pa.atom_cap + 1;
    }
    if (ex.start) {
        // this cluster's atoms: lane (ii, k) takes the k-th exclusion partner
        // of atom ii, 16 per pass
        const int at = s2o[islot];
        if (at >= 0) excluded_partners(ex, pk, at, lane >> 2, 16, ax, ay, az, qi, cx, cy, cz, beta, fix, fiy, fiz);
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


// One block of the pair kernel: the LJ (A, B) table into LDS, then one
// i-cluster per wave (logical block lb covers clusters W lb .. W lb + W - 1).
template <bool ENERGY, bool SOFT, int TB, int W = kPairWaves>
__device__ __forceinline__ void pair_body(
    int lb, const int* __restrict__ flags, const int* __restrict__ list_n, const uint2* __restrict__ list,
    int list_cap, const uint4* __restrict__ pk, const float2* __restrict__ ab, int nt, float cx, float cy, float cz,
    float rc2, float beta, const int* __restrict__ s2o, float* __restrict__ frc, float* __restrict__ atom_e,
    float* __restrict__ atom_v, float* __restrict__ atom_d, double* __restrict__ acc, int overwrite,
    const float4* __restrict__ sc, PruneArgs pa = PruneArgs{},
    ExclArgs ex = ExclArgs{nullptr, nullptr, nullptr}) {
    extern __shared__ float2 sab[];
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) sab[k] = ab[k];
    // the Coulomb table after the LJ table (16-byte aligned), when given
    float4* stab = nullptr;
    if (!ENERGY && !SOFT && pa.ctab) {
        stab = reinterpret_cast<float4*>(reinterpret_cast<char*>(sab) + ((nt * nt * 8 + 15) & ~15));
        for (int k = threadIdx.x; k < kCoulTab; k += blockDim.x) stab[k] = pa.ctab[k];
    }
    __syncthreads();
    const int ic = lb * W + (int)(threadIdx.x >> 6);
    if (lb < 0 || ic >= flags[4]) return;
    pair_wave<ENERGY, SOFT, TB>(ic, sab, stab, flags, list_n, list, list_cap, pk, nt, cx, cy, cz, rc2, beta, s2o,
                                frc, atom_e, atom_v, atom_d, acc, overwrite, sc, pa, ex);
}

}  // namespace snb
