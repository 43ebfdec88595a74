// Per-atom PME preparation shared by the PME module's own prep kernel and the
// fused per-step prep kernel (PME_Atom_Near, PME_force.cu:367-421, without
// the 64-index cache).
#pragma once
#include "common.hpp"

namespace snb {

// PME acc slots 6 and 7 hold sum q^2 and sum q between prep and finalize.
constexpr int kAccQ2 = 6;
constexpr int kAccQ1 = 7;

struct PmePrep {
    int Kx, Ky, Kz;
    unsigned* base;     // x | y << 10 | z << 20
    float4* frac;       // (fx, fy, fz, q)
    int* bin_count;     // atoms filed under each x base plane
    int* bin_atoms;
    unsigned* bin_base;  // the filed atom's base and fraction, in bin order
    float4* bin_frac;
    int bin_cap;
    double* acc;
};

__host__ __device__ static inline PmePrep pme_prep_args(const sponge_pme* pme) {
    PmePrep p;
    p.Kx = pme->K[0];
    p.Ky = pme->K[1];
    p.Kz = pme->K[2];
    p.base = pme->base;
    p.frac = pme->frac;
    p.bin_count = pme->bin_count;
    p.bin_atoms = pme->bin_atoms;
    p.bin_base = pme->bin_base;
    p.bin_frac = pme->bin_frac;
    p.bin_cap = pme->bin_cap;
    p.acc = pme->acc;
    return p;
}

// Block-cooperative: it contains workgroup barriers, so every thread of the
// block must reach the same single call site (no call under divergent control
// flow, no early return before it); i < 0 = no atom. t = u K / 2^32 from one
// exact 32x32->64-bit product per axis; base index and fraction stored; the
// atom filed under its x base plane with its base and fraction (LDS
// histogram, one global atomic per block and plane); with energy, sum q^2 and
// sum q added to acc in FP64.
__device__ __forceinline__ void pme_prep_block(int i, unsigned ux, unsigned uy, unsigned uz, float q,
                                               const PmePrep& P, int energy) {
    __shared__ int hist[kMaxGridAxis];
    __shared__ int boff[kMaxGridAxis];
    for (int b = threadIdx.x; b < P.Kx; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    double q2 = 0.0, q1 = 0.0;
    int bx = -1, rank = 0;
    unsigned bb = 0u;
    float4 ff = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i >= 0) {
        const float s = 2.3283064365386963e-10f;  // 2^-32
        unsigned long long tx = (unsigned long long)ux * (unsigned)P.Kx;
        unsigned long long ty = (unsigned long long)uy * (unsigned)P.Ky;
        unsigned long long tz = (unsigned long long)uz * (unsigned)P.Kz;
        bx = (int)(tx >> 32);
        bb = (unsigned)bx | ((unsigned)(ty >> 32) << 10) | ((unsigned)(tz >> 32) << 20);
        ff = make_float4((float)(unsigned)tx * s, (float)(unsigned)ty * s, (float)(unsigned)tz * s, q);
        P.base[i] = bb;
        P.frac[i] = ff;
        rank = atomicAdd(hist + bx, 1);
        if (energy) {
            q1 = q;
            q2 = q1 * q1;
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < P.Kx; b += blockDim.x)
        if (hist[b] > 0) boff[b] = atomicAdd(P.bin_count + b, hist[b]);
    __syncthreads();
    if (bx >= 0) {
        const int slot = boff[bx] + rank;
        if (slot < P.bin_cap) {
            // the record is filed with the index, so the spread reads it in
            // bin order without a dependent load through the index
            const size_t o = (size_t)bx * P.bin_cap + slot;
            P.bin_atoms[o] = i;
            P.bin_base[o] = bb;
            P.bin_frac[o] = ff;
        }
    }
    if (energy) {
        q2 = wave_sum(q2);
        q1 = wave_sum(q1);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(P.acc + kAccQ2, q2);
            atomicAdd(P.acc + kAccQ1, q1);
        }
    }
}

}  // namespace snb
