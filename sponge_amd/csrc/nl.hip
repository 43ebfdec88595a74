// Neighbor-list rebuild for gfx950: NEIGHBOR_LIST (neighbor_list/neighbor_list.cu)
// re-designed around 4-atom clusters.
//
// The reference bins atoms into cells of edge >= (rc+skin)/2 and appends, per
// atom i, every j>i in the 125 surrounding cells with FP32 dr2 < (rc+skin)^2
// (neighbor_list.cu:223-283), then deletes excluded partners (:353-403).
// Here atoms are counting-sorted into xy columns, z-sorted inside each column
// and cut into clusters of 4; the list holds, per i-cluster, every j-cluster
// (both directions, so each unordered cluster pair appears twice and the pair
// kernel needs no atomics) whose bounding box comes within rc+skin, each with
// a 16-bit mask whose bit (jj*4+ii) is set exactly when the reference would
// list that atom pair: the same FP32 dr2 expression, the same strict
// comparison, minus exclusions and padding. The set of listed unordered atom
// pairs is therefore identical to the reference's half list; only the storage
// differs.
//
// Rebuild decision stays on the device (no host sync, no device-side
// launch): every rebuild kernel reads flags[0] and returns at once when it
// is 0, so the skin-triggered mode can sit inside a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "tune_table.hpp"

using namespace snb;

namespace {

constexpr int kSortThreads = 256;
constexpr int kMaxColumn = 1024;     // bitonic sort capacity per column (LDS)
constexpr int kPairWaves = 4;
constexpr int kCandCap = 1024;       // candidate j-clusters per i-cluster (LDS)
constexpr int kColCap = 128;         // cell columns of one search window held in LDS

// Is_need_refresh_neighbor_list_cuda (neighbor_list.cu:336-351): FP32
// periodic displacement against the coordinates of the last rebuild.
__device__ __forceinline__ void nl_check_body(int blk, int n, const float* __restrict__ crd,
                                              const float* __restrict__ old_crd, float bx, float by, float bz,
                                              float thresh, int* flags) {
#pragma clang fp contract(off)
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float dx = crd[3 * i] - old_crd[3 * i];
    float dy = crd[3 * i + 1] - old_crd[3 * i + 1];
    float dz = crd[3 * i + 2] - old_crd[3 * i + 2];
    dx = dx - floorf(dx / bx + 0.5f) * bx;
    dy = dy - floorf(dy / by + 0.5f) * by;
    dz = dz - floorf(dz / bz + 0.5f) * bz;
    float r2 = dr2_exact(dx, dy, dz);
    if (r2 > thresh) atomicOr(flags + 7, 1);
}

__global__ void nl_check_kernel(int n, const float* __restrict__ crd, const float* __restrict__ old_crd,
                                float bx, float by, float bz, float thresh, int* flags) {
    nl_check_body(blockIdx.x, n, crd, old_crd, bx, by, bz, thresh, flags);
}

// Latches the rebuild decision into flags[0] and clears per-rebuild counters.
__device__ __forceinline__ void nl_begin_body(int forced, int* flags, int* col_count, int ncol) {
    __shared__ int go;
    if (threadIdx.x == 0) {
        int f = forced ? 1 : flags[7];
        flags[7] = 0;
        flags[0] = f;
        go = f;
        if (f) {
            flags[1] = 0;
            flags[5] = 0;
            flags[8] = 0;
            flags[9] = 0;
            flags[2] += 1;
            flags[10] = 0;  // the next fused step prunes the new list
            flags[13] = 0;  // no displacement vote on the new list yet
            flags[14] = 0;
            flags[11] = 0;  // and rewrites the per-slot inner lists
        }
    }
    __syncthreads();
    if (!go) return;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) col_count[c] = 0;
}

__global__ void nl_begin_kernel(int forced, int* flags, int* col_count, int ncol) {
    nl_begin_body(forced, flags, col_count, ncol);
}

// Crd_Periodic_Map (common.cu:299-352), old_crd copy (neighbor_list.cu:474)
// and Crd_To_Uint_Crd with the list's own cof (common.cu:264-287,
// neighbor_list.cu:622), then the column of each atom from its uint x, y.
__device__ __forceinline__ void nl_bin_body(int blk, int n, float* __restrict__ crd, float* __restrict__ old_crd,
                                            unsigned int* __restrict__ nl_uint, float bx, float by, float bz,
                                            float cqx, float cqy, float cqz, int ncx, int ncy, int col_cap,
                                            int* __restrict__ col_count, int* __restrict__ col_atoms, int* flags,
                                            int* __restrict__ o2s, int* __restrict__ dropped) {
    if (!flags[0]) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float b[3] = {bx, by, bz};
    float cq[3] = {cqx, cqy, cqz};
    unsigned int u[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float x = crd[3 * i + d];
        if (x >= 0.0f) {
            if (!(x < b[d])) x = x - b[d];
        } else {
            x = x + b[d];
        }
        crd[3 * i + d] = x;
        old_crd[3 * i + d] = x;
        int t = (int)__fmul_rn(x, cq[d]);
        u[d] = ((unsigned int)t) << 2;
        nl_uint[3 * i + d] = u[d];
    }
    int cx = (int)(((unsigned long long)u[0] * (unsigned)ncx) >> 32);
    int cy = (int)(((unsigned long long)u[1] * (unsigned)ncy) >> 32);
    int c = cx * ncy + cy;
    int s = atomicAdd(col_count + c, 1);
    if (s < col_cap) {
        col_atoms[(size_t)c * col_cap + s] = i;
    } else {
        // left out of this list (reported): no slot, which the slot map's
        // readers test (the sort writes the slots of the others); the fused
        // step's prep finds the atom in `dropped` and gives it a defined
        // state (fresh uint coordinates, zero charge on the PME grid, zero
        // force when forces are assigned)
        atomicOr(flags + 1, SPONGE_NL_OVERFLOW_COLUMN);
        o2s[i] = -1;
        dropped[atomicAdd(flags + 5, 1)] = i;
    }
}

__global__ void nl_bin_kernel(int n, float* __restrict__ crd, float* __restrict__ old_crd,
                              unsigned int* __restrict__ nl_uint, float bx, float by, float bz,
                              float cqx, float cqy, float cqz, int ncx, int ncy, int col_cap,
                              int* __restrict__ col_count, int* __restrict__ col_atoms, int* flags,
                              int* __restrict__ o2s, int* __restrict__ dropped) {
    nl_bin_body(blockIdx.x, n, crd, old_crd, nl_uint, bx, by, bz, cqx, cqy, cqz, ncx, ncy, col_cap, col_count,
                col_atoms, flags, o2s, dropped);
}

// nl_bin_body with the column slots taken through a block histogram in LDS:
// consecutive atoms (the atoms of a molecule, molecules placed next to each
// other) fall into the same few columns, and one global atomic per atom on
// the same counters serialises; here a block adds one count per column it
// touches. The atom order within a column changes, which the sort (by uint z,
// then atom index) removes. LDS: 2 ncol ints.
__global__ __launch_bounds__(256) void nl_bin_hist_kernel(
    int n, float* __restrict__ crd, float* __restrict__ old_crd, unsigned int* __restrict__ nl_uint, float bx,
    float by, float bz, float cqx, float cqy, float cqz, int ncx, int ncy, int col_cap, int* __restrict__ col_count,
    int* __restrict__ col_atoms, int* flags, int* __restrict__ o2s, int* __restrict__ dropped) {
    if (!flags[0]) return;  // uniform: every thread of the grid returns
    extern __shared__ int hbin[];
    const int ncol = ncx * ncy;
    int* hist = hbin;
    int* hbase = hbin + ncol;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1, rank = 0;
    if (i < n) {
        const float b[3] = {bx, by, bz};
        const float cq[3] = {cqx, cqy, cqz};
        unsigned int u[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float x = crd[3 * i + d];
            if (x >= 0.0f) {
                if (!(x < b[d])) x = x - b[d];
            } else {
                x = x + b[d];
            }
            crd[3 * i + d] = x;
            old_crd[3 * i + d] = x;
            const int t = (int)__fmul_rn(x, cq[d]);
            u[d] = ((unsigned int)t) << 2;
            nl_uint[3 * i + d] = u[d];
        }
        const int cx = (int)(((unsigned long long)u[0] * (unsigned)ncx) >> 32);
        const int cy = (int)(((unsigned long long)u[1] * (unsigned)ncy) >> 32);
        c = cx * ncy + cy;
        rank = atomicAdd(hist + c, 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ncol; k += blockDim.x)
        if (hist[k] > 0) hbase[k] = atomicAdd(col_count + k, hist[k]);
    __syncthreads();
    if (c >= 0) {
        const int s = hbase[c] + rank;
        if (s < col_cap) {
            col_atoms[(size_t)c * col_cap + s] = i;
        } else {
            // left out of this list, as nl_bin_body does
            atomicOr(flags + 1, SPONGE_NL_OVERFLOW_COLUMN);
            o2s[i] = -1;
            dropped[atomicAdd(flags + 5, 1)] = i;
        }
    }
}

// Largest column count binned through the LDS histogram (64 KB of LDS).
constexpr int kBinHistMax = 8192;

// Exclusive scan of the padded (multiple of 4) column sizes; one block.
__device__ __forceinline__ void nl_scan_body(int ncol, int col_cap, const int* __restrict__ col_count,
                                             int* __restrict__ col_start, int* flags) {
    if (!flags[0]) return;
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int base = 0; base < ncol; base += blockDim.x) {
        int c = base + threadIdx.x;
        int v = 0;
        if (c < ncol) {
            int k = min(col_count[c], col_cap);
            v = (k + 3) & ~3;
        }
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int off = carry;
        for (int k = 0; k < w; k++) off += wsum[k];
        if (c < ncol) col_start[c] = off + x - v;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int k = 0; k < nw; k++) t += wsum[k];
            carry += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        col_start[ncol] = carry;
        flags[3] = carry;
        flags[4] = carry / kClusterSize;
    }
}

__global__ void nl_scan_kernel(int ncol, int col_cap, const int* __restrict__ col_count,
                               int* __restrict__ col_start, int* flags) {
    nl_scan_body(ncol, col_cap, col_count, col_start, flags);
}

// One block per column: sort its atoms by (uint z, atom index) -- a
// deterministic order -- write the slot maps, the sorted uint coordinates and
// the bounding box of every 4-atom cluster.
__device__ __forceinline__ void nl_sort_body(
    int c, int col_cap, const int* __restrict__ col_count, const int* __restrict__ col_start,
    const int* __restrict__ col_atoms, const unsigned int* __restrict__ nl_uint,
    float crx, float cry, float crz, int* __restrict__ s2o, int* __restrict__ o2s,
    uint4* __restrict__ pack, float4* __restrict__ cl_lo, float4* __restrict__ cl_hi, const int* flags) {
    if (!flags[0]) return;
    __shared__ unsigned long long key[kMaxColumn];
    int cnt = min(col_count[c], col_cap);
    int start = col_start[c];
    int pad = (cnt + 3) & ~3;
    int P = 1;
    while (P < cnt) P <<= 1;
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
        if (k < cnt) {
            int a = col_atoms[(size_t)c * col_cap + k];
            key[k] = ((unsigned long long)nl_uint[3 * a + 2] << 32) | (unsigned)a;
        } else {
            key[k] = ~0ull;
        }
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < P; k += blockDim.x) {
                int partner = k ^ stride;
                if (partner > k) {
                    bool up = (k & size) == 0;
                    unsigned long long a = key[k], b = key[partner];
                    if ((a > b) == up) {
                        key[k] = b;
                        key[partner] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < pad; k += blockDim.x) {
        int slot = start + k;
        if (k < cnt) {
            int a = (int)(key[k] & 0xffffffffu);
            s2o[slot] = a;
            o2s[a] = slot;
            pack[slot] = make_uint4(nl_uint[3 * a], nl_uint[3 * a + 1], nl_uint[3 * a + 2], (unsigned)a);
        } else {
            s2o[slot] = -1;
            pack[slot] = make_uint4(0u, 0u, 0u, 0xffffffffu);
        }
    }
    __syncthreads();
    int ncl = pad / kClusterSize;
    for (int k = threadIdx.x; k < ncl; k += blockDim.x) {
        float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        for (int m = 0; m < kClusterSize; m++) {
            int kk = k * kClusterSize + m;
            if (kk >= cnt) break;
            int a = (int)(key[kk] & 0xffffffffu);
            float x[3] = {(float)nl_uint[3 * a] * crx, (float)nl_uint[3 * a + 1] * cry,
                          (float)nl_uint[3 * a + 2] * crz};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = fminf(lo[d], x[d]);
                hi[d] = fmaxf(hi[d], x[d]);
            }
        }
        int cid = start / kClusterSize + k;
        cl_lo[cid] = make_float4(lo[0], lo[1], lo[2], 0.f);
        cl_hi[cid] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
}

__global__ __launch_bounds__(kSortThreads) void nl_sort_kernel(
    int col_cap, const int* __restrict__ col_count, const int* __restrict__ col_start,
    const int* __restrict__ col_atoms, const unsigned int* __restrict__ nl_uint,
    float crx, float cry, float crz, int* __restrict__ s2o, int* __restrict__ o2s,
    uint4* __restrict__ pack, float4* __restrict__ cl_lo, float4* __restrict__ cl_hi, const int* flags) {
    nl_sort_body(blockIdx.x, col_cap, col_count, col_start, col_atoms, nl_uint, crx, cry, crz, s2o, o2s, pack, cl_lo,
                 cl_hi, flags);
}

// FP32 neighbor-list distance, the reference's expression with every
// operation rounded in a fixed order (no contraction): must agree bit for
// bit with the oracle.
__device__ __forceinline__ float nl_dr2(uint4 a, uint4 b, float cx, float cy, float cz) {
    float dx = (float)(int)(b.x - a.x) * cx;
    float dy = (float)(int)(b.y - a.y) * cy;
    float dz = (float)(int)(b.z - a.z) * cz;
    return dr2_exact(dx, dy, dz);
}

__device__ __forceinline__ bool excluded_pair(int a, int b, const int* __restrict__ es,
                                              const int* __restrict__ en, const int* __restrict__ el) {
    int lo = min(a, b), hi = max(a, b);
    int s = es[lo], e = s + en[lo];
    for (int k = s; k < e; k++)
        if (el[k] == hi) return true;
    return false;
}

// One wave per i-cluster: gather candidate j-clusters (index >= i) from the
// columns within rc+skin, test bounding boxes, then compute the exact 16-bit
// pair masks four cluster pairs at a time and append the non-empty ones.
__device__ __forceinline__ void nl_pairs_body(
    int lb, int ncx, int ncy, float bx, float by, float bz, float crx, float cry, float crz, float rl2,
    float rlm, const int* __restrict__ col_start, const float4* __restrict__ cl_lo,
    const float4* __restrict__ cl_hi, const uint4* __restrict__ pack, const int* __restrict__ es,
    const int* __restrict__ en, const int* __restrict__ el, int list_cap, int* __restrict__ list_n,
    uint2* __restrict__ list, int* flags, const int* __restrict__ sym_start = nullptr,
    const int* __restrict__ sym_list = nullptr) {
    if (!flags[0]) return;
    __shared__ int cand[kPairWaves][kCandCap];
    __shared__ int colj0[kPairWaves][kColCap];
    __shared__ int colpre[kPairWaves][kColCap];
    __shared__ int colzc[kPairWaves][kColCap];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ic = lb * kPairWaves + w;
    const int ncl = flags[4];
    if (lb < 0 || ic >= ncl) return;
    const float4 li = cl_lo[ic], hi = cl_hi[ic];
    const float cix = 0.5f * (li.x + hi.x), ciy = 0.5f * (li.y + hi.y), ciz = 0.5f * (li.z + hi.z);
    const float hix = 0.5f * (hi.x - li.x), hiy = 0.5f * (hi.y - li.y), hiz = 0.5f * (hi.z - li.z);
    const float cwx = bx / ncx, cwy = by / ncy;
    int ax0 = (int)floorf((li.x - rlm) / cwx), ax1 = (int)floorf((hi.x + rlm) / cwx);
    int ay0 = (int)floorf((li.y - rlm) / cwy), ay1 = (int)floorf((hi.y + rlm) / cwy);
    int nxr = ax1 - ax0 + 1, nyr = ay1 - ay0 + 1;
    if (nxr >= ncx) { ax0 = 0; nxr = ncx; }
    if (nyr >= ncy) { ay0 = 0; nyr = ncy; }
    const float rlm2 = rlm * rlm;
    int ncand = 0;
    bool over = false;
    // bounding-box test of candidate j-cluster jc; appends it to cand[w]
    auto test_box = [&](int jc, bool live) {
        bool ok = false;
        if (live) {
            const float4 lj = cl_lo[jc], hj = cl_hi[jc];
            float dx = 0.5f * (lj.x + hj.x) - cix, dy = 0.5f * (lj.y + hj.y) - ciy, dz = 0.5f * (lj.z + hj.z) - ciz;
            dx -= bx * rintf(dx / bx);
            dy -= by * rintf(dy / by);
            dz -= bz * rintf(dz / bz);
            const float gx = fmaxf(fabsf(dx) - hix - 0.5f * (hj.x - lj.x), 0.f);
            const float gy = fmaxf(fabsf(dy) - hiy - 0.5f * (hj.y - lj.y), 0.f);
            const float gz = fmaxf(fabsf(dz) - hiz - 0.5f * (hj.z - lj.z), 0.f);
            ok = gx * gx + gy * gy + gz * gz < rlm2;
        }
        const unsigned long long bal = __ballot(ok);
        const int pos = ncand + __popcll(bal & ((1ull << lane) - 1ull));
        if (ok) {
            if (pos < kCandCap) cand[w][pos] = jc;
            else over = true;
        }
        ncand += __popcll(bal);
    };
    const int ncols = nxr * nyr;
    if (ncols <= kColCap) {
        // The window's columns as one flat range of j-clusters: a table of
        // each column's first cluster and running cluster count, then 64
        // clusters per step whatever their column (instead of one column per
        // step, most of whose lanes would idle).
        // Within a column the clusters are sorted by z (their atoms are), so
        // only those whose z extent meets [lo.z - rlm, hi.z + rlm] of the
        // i-cluster (periodically: a circular run of the column's clusters,
        // found by two binary searches) can pass the box test; a window at
        // least a box length long takes the whole column.
        const float zpad = rlm + 1.0e-2f;
        const float zlo = li.z - zpad, zhi = hi.z + zpad;
        const bool zwin = zhi - zlo < bz;
        const float za = zlo < 0.f ? zlo + bz : zlo;   // window start in [0, bz)
        const float zb = zhi >= bz ? zhi - bz : zhi;   // window end in [0, bz)
        const bool zwrap = zlo < 0.f || zhi >= bz;
        int total = 0;
        for (int base = 0; base < ncols; base += 64) {
            const int k = base + lane;
            int j0 = 0, cnt = 0, zs = 0, zl = 0;
            if (k < ncols) {
                const int ox = k / nyr, oy = k - ox * nyr;
                const int cx = ((ax0 + ox) % ncx + ncx) % ncx, cy = ((ay0 + oy) % ncy + ncy) % ncy;
                const int c = cx * ncy + cy;
                j0 = col_start[c] / kClusterSize;
                cnt = col_start[c + 1] / kClusterSize - j0;
                zl = cnt;
                if (zwin && cnt > 0) {
                    // s: the first cluster whose top reaches za; e: the first
                    // whose bottom lies above zb
                    int lo = 0, up = cnt;
                    while (lo < up) {
                        const int mid = (lo + up) >> 1;
                        if (cl_hi[j0 + mid].z < za) lo = mid + 1;
                        else up = mid;
                    }
                    const int s = lo;
                    lo = 0;
                    up = cnt;
                    while (lo < up) {
                        const int mid = (lo + up) >> 1;
                        if (cl_lo[j0 + mid].z <= zb) lo = mid + 1;
                        else up = mid;
                    }
                    const int e = lo;
                    zs = s == cnt ? 0 : s;
                    zl = zwrap ? min(cnt - s + e, cnt) : max(e - s, 0);
                }
            }
            int incl = zl;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            if (k < ncols) {
                colj0[w][k] = j0;
                colpre[w][k] = total + incl - zl;
                colzc[w][k] = zs | (cnt << 16);
            }
            total += __shfl(incl, 63);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < total; t0 += 64) {
            const int t = t0 + lane;
            int jc = 0;
            if (t < total) {
                // last column whose run starts at or before t (empty columns
                // share their successor's start and are skipped)
                int lo = 0, hi = ncols - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (colpre[w][mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int zc = colzc[w][lo], zs = zc & 0xffff, cnt = zc >> 16;
                const int u = zs + (t - colpre[w][lo]);
                jc = colj0[w][lo] + (u >= cnt ? u - cnt : u);
            }
            test_box(jc, t < total);
        }
    } else {
    for (int ox = 0; ox < nxr; ox++) {
        int cx = ((ax0 + ox) % ncx + ncx) % ncx;
        for (int oy = 0; oy < nyr; oy++) {
            int cy = ((ay0 + oy) % ncy + ncy) % ncy;
            int c = cx * ncy + cy;
            int j0 = col_start[c] / kClusterSize, j1 = col_start[c + 1] / kClusterSize;
            for (int jb = j0; jb < j1; jb += 64) test_box(jb + lane, jb + lane < j1);
        }
    }
    }
    if (ncand > kCandCap) {
        ncand = kCandCap;
        over = true;
    }
    if (over && lane == 0) atomicOr(flags + 1, SPONGE_NL_OVERFLOW_CANDIDATES);
    // cand[w] is private to this wave: order its LDS writes before the reads.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();

    const int ii = lane & 3, jj = (lane >> 2) & 3, g = lane >> 4;
    const uint4 pi = pack[ic * kClusterSize + ii];
    const int ai = (int)pi.w;
    // The lane's i-atom exclusion partners, the first four in registers, from
    // the symmetric lists when there are some (else the CSR is searched).
    int e0 = -1, e1 = -1, e2 = -1, e3 = -1, r0 = 0, r1 = 0;
    if (sym_start && ai >= 0) {
        r0 = sym_start[ai];
        r1 = sym_start[ai + 1];
        const int len = r1 - r0;
        if (len > 0) e0 = sym_list[r0];
        if (len > 1) e1 = sym_list[r0 + 1];
        if (len > 2) e2 = sym_list[r0 + 2];
        if (len > 3) e3 = sym_list[r0 + 3];
    }
    auto is_excluded = [&](int aj) -> bool {
        if (!sym_start) return excluded_pair(ai, aj, es, en, el);
        if (aj == e0 || aj == e1 || aj == e2 || aj == e3) return true;
        for (int k = r0 + 4; k < r1; k++)
            if (sym_list[k] == aj) return true;
        return false;
    };
    int count = 0;
    long npairs = 0;
    // four groups of 4 candidates per pass: their j records are loaded
    // together, then tested and appended in candidate order (the loop is
    // latency bound; the list stores of a pass no longer stall the next
    // group's loads)
    for (int cb = 0; cb < ncand; cb += 16) {
        int jcs[4];
        uint4 pjs[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int c = cb + 4 * u + g;
            jcs[u] = c < ncand ? cand[w][c] : -1;
            pjs[u] = jcs[u] >= 0 ? pack[jcs[u] * kClusterSize + jj] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int jc = jcs[u];
            const uint4 pj = pjs[u];
            bool bit = false;
            if (jc >= 0) {
                int aj = (int)pj.w;
                bool valid = ai >= 0 && aj >= 0 && (jc != ic || jj != ii);
                if (valid && nl_dr2(pi, pj, crx, cry, crz) < rl2) bit = !is_excluded(aj);
            }
            unsigned long long bal = __ballot(bit);
            unsigned m = (unsigned)(bal >> (g * 16)) & 0xffffu;
            bool leader = (lane & 15) == 0 && m != 0;
            unsigned long long lb2 = __ballot(leader);
            int slot = count + __popcll(lb2 & ((1ull << lane) - 1ull));
            if (leader) {
                if (slot < list_cap) list[(size_t)ic * list_cap + slot] = make_uint2((unsigned)jc, m);
            }
            count += __popcll(lb2);
            npairs += __popcll(bal);
        }
    }
    if (lane == 0) {
        if (count > list_cap) {
            atomicOr(flags + 1, SPONGE_NL_OVERFLOW_LIST);
            count = list_cap;
        }
        list_n[ic] = count;
        // every unordered pair is listed from both of its atoms
        atomicAdd((unsigned long long*)(flags + 8), (unsigned long long)npairs);
    }
}

__global__ __launch_bounds__(kPairWaves * 64) void nl_pairs_kernel(
    int ncx, int ncy, float bx, float by, float bz, float crx, float cry, float crz, float rl2,
    float rlm, const int* __restrict__ col_start, const float4* __restrict__ cl_lo,
    const float4* __restrict__ cl_hi, const uint4* __restrict__ pack, const int* __restrict__ es,
    const int* __restrict__ en, const int* __restrict__ el, int list_cap, int* __restrict__ list_n,
    uint2* __restrict__ list, int* flags, const int* __restrict__ sym_start, const int* __restrict__ sym_list) {
    nl_pairs_body(xcd_block(gridDim.x), ncx, ncy, bx, by, bz, crx, cry, crz, rl2, rlm, col_start, cl_lo, cl_hi, pack,
                  es, en, el, list_cap, list_n, list, flags, sym_start, sym_list);
}

// Per-slot lists of the fused step from the stored list (plain LJ, inner list
// on): for each slot ii of cluster ic, the partners jc*4 + jj of every entry
// whose mask holds bit jj*4 + ii, in list order and jj order within an entry
// (the order in which the pair kernel's own walk writes them), as 16-bit slot
// offsets. One wave per cluster, one entry per lane: each lane's place in a
// slot's list is the sum of the bit counts of the lanes below, so a wave
// writes 64 entries' partners at once. A slot with more partners than cap,
// or one beyond 16-bit reach, gets the count cap + 1 (its cluster then reads
// the full list) and raises flags[11].
__global__ __launch_bounds__(256) void nl_slot_lists_kernel(const int* __restrict__ flags,
                                                            const int* __restrict__ list_n,
                                                            const uint2* __restrict__ list, int list_cap,
                                                            unsigned short* __restrict__ slot_list,
                                                            int* __restrict__ slot_n, int cap, int* diag) {
    if (!flags[0]) return;
    const int lane = threadIdx.x & 63;
    const int lb = xcd_block(gridDim.x);
    const int ic = lb * 4 + (int)(threadIdx.x >> 6);
    if (lb < 0 || ic >= flags[4]) return;
    const int nsl = flags[3];
    const int n = min(list_n[ic], list_cap);
    const uint2* lst = list + (size_t)ic * list_cap;
    const unsigned long long below = (1ull << lane) - 1ull;
    int base[kClusterSize] = {0, 0, 0, 0};
    unsigned farmask = 0u;  // slots of this lane's entries with a partner beyond 16-bit reach
    // the entries of four batches of 64 are loaded together; each slot's
    // place for a lane is the count of its bits in the entries of the lanes
    // below (three ballots of the count's bits, counted below the lane)
    for (int b0 = 0; b0 < n; b0 += 256) {
        uint2 ens[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = b0 + 64 * u + lane;
            ens[u] = e < n ? lst[e] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (b0 + 64 * u >= n) break;
            const uint2 en = ens[u];
            const unsigned m = en.y;
#pragma unroll
            for (int ii = 0; ii < kClusterSize; ii++) {
                const unsigned c = (unsigned)__popc(m & (0x1111u << ii));
                const unsigned long long q0 = __ballot(c & 1u), q1 = __ballot(c & 2u), q2 = __ballot(c & 4u);
                int pos = base[ii] + __popcll(q0 & below) + 2 * __popcll(q1 & below) + 4 * __popcll(q2 & below);
                const int islot = ic * kClusterSize + ii;
                unsigned short* out = slot_list + (size_t)islot * cap;
#pragma unroll
                for (int jj = 0; jj < kClusterSize; jj++) {
                    if ((m >> (jj * 4 + ii)) & 1u) {
                        const int j = (int)en.x * kClusterSize + jj;
                        int off = j - islot;
                        bool reach = true;
                        if (nsl > 32767) reach = slot_offset(islot, j, nsl, off);
                        if (!reach) farmask |= 1u << ii;
                        if (pos < cap) out[pos] = (unsigned short)off;
                        pos++;
                    }
                }
                base[ii] += __popcll(q0) + 2 * __popcll(q1) + 4 * __popcll(q2);
            }
        }
    }
#pragma unroll
    for (int ii = 0; ii < kClusterSize; ii++) {
        const bool slot_far = __ballot((farmask >> ii) & 1u) != 0ull;
        int count = base[ii];
        if (slot_far) count = cap + 1;
        if (lane == ii) {
            slot_n[ic * kClusterSize + ii] = count;
            if (count > cap) atomicOr(diag, 1);
        }
    }
}

// ---- batched rebuild: one launch per stage over every system --------------
__global__ void nl_check_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    nl_check_body(m.y, b.nl.n, b.io.crd, b.nl.old_crd, b.nl.box[0], b.nl.box[1], b.nl.box[2], b.thresh, b.nl.flags);
}

__global__ void nl_begin_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map, int forced) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    nl_begin_body(forced, b.nl.flags, b.nl.col_count, b.nl.ncol);
}

__global__ void nl_bin_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_nl& L = b.nl;
    nl_bin_body(m.y, L.n, b.io.crd, L.old_crd, L.nl_uint, L.box[0], L.box[1], L.box[2], L.cof_q[0], L.cof_q[1],
                L.cof_q[2], L.ncx, L.ncy, L.col_cap, L.col_count, L.col_atoms, L.flags, L.o2s, L.dropped);
}

__global__ void nl_scan_batch(const BatchSys* __restrict__ bs, const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const sponge_nl& L = bs[m.x].nl;
    nl_scan_body(L.ncol, L.col_cap, L.col_count, L.col_start, L.flags);
}

__global__ __launch_bounds__(kSortThreads) void nl_sort_batch(const BatchSys* __restrict__ bs,
                                                              const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const sponge_nl& L = bs[m.x].nl;
    nl_sort_body(m.y, L.col_cap, L.col_count, L.col_start, L.col_atoms, L.nl_uint, L.cof_dr[0], L.cof_dr[1],
                 L.cof_dr[2], L.s2o, L.o2s, L.pack, L.cl_lo, L.cl_hi, L.flags);
}

__global__ __launch_bounds__(kPairWaves * 64) void nl_pairs_batch(const BatchSys* __restrict__ bs,
                                                                  const int2* __restrict__ map, int nmap) {
    const int g = xcd_block(nmap);
    if (g < 0) return;
    const int2 m = map[g];
    if (m.x < 0) return;
    const BatchSys& b = bs[m.x];
    const sponge_nl& L = b.nl;
    nl_pairs_body(m.y, L.ncx, L.ncy, L.box[0], L.box[1], L.box[2], L.cof_dr[0], L.cof_dr[1], L.cof_dr[2], L.rl2,
                  b.rlm, L.col_start, L.cl_lo, L.cl_hi, L.pack, b.io.excluded_list_start, b.io.excluded_numbers, b.io.excluded_list,
                  L.list_cap, L.list_n, L.list, L.flags);
}

__global__ void nl_uint_kernel(int n, const float* __restrict__ crd, float cqx, float cqy, float cqz,
                               unsigned int* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float cq[3] = {cqx, cqy, cqz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int t = (int)__fmul_rn(crd[3 * i + d], cq[d]);
        out[3 * i + d] = ((unsigned int)t) << 2;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
static void nl_set_box(sponge_nl* nl, const float* box) {
    for (int d = 0; d < 3; d++) {
        nl->box[d] = box[d];
        nl->cof_q[d] = (0.25f * 4294967296.0f) / box[d];
        nl->cof_dr[d] = (1.0f / 4294967296.0f) * box[d];
    }
}

extern "C" int sponge_crd_to_uint(int atom_numbers, const float box_length[3], const float* d_crd,
                                  unsigned int* d_uint_crd, void* stream) {
    SNB_ARG(atom_numbers >= 0 && box_length && d_crd && d_uint_crd, "crd_to_uint");
    if (atom_numbers == 0) return SPONGE_NB_OK;
    // MD_core's pbc cof (MD_core.cu:485-490): 0.25 * (2^32 / L), equal to the
    // list's 0.25*2^32 / L bit for bit (scaling by 0.25 is exact).
    float cq[3];
    for (int d = 0; d < 3; d++) cq[d] = 0.25f * (4294967296.0f / box_length[d]);
    nl_uint_kernel<<<ceil_div(atom_numbers, 256), 256, 0, as_stream(stream)>>>(atom_numbers, d_crd, cq[0], cq[1],
                                                                              cq[2], d_uint_crd);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_create(int atom_numbers, const float box_length[3], float cutoff, float skin,
                                int refresh_interval, int max_atom_in_grid_numbers,
                                int max_neighbor_numbers, float skin_permit, sponge_nl** out) {
    SNB_ARG(out && box_length, "null pointer");
    SNB_ARG(atom_numbers > 0, "atom_numbers must be positive");
    SNB_ARG(cutoff > 0 && skin >= 0, "cutoff/skin");
    for (int d = 0; d < 3; d++) SNB_ARG(box_length[d] > 0, "box_length");
    *out = nullptr;
    sponge_nl* nl = new sponge_nl();
    memset(nl, 0, sizeof(*nl));
    nl->n = atom_numbers;
    nl->cutoff = cutoff;
    nl->skin = skin;
    nl->refresh_interval = refresh_interval;
    nl->max_atom_in_grid_numbers = max_atom_in_grid_numbers > 0 ? max_atom_in_grid_numbers : 64;
    nl->max_neighbor_numbers = max_neighbor_numbers > 0 ? max_neighbor_numbers : 800;
    nl->skin_permit = skin_permit > 0 ? 2.0f * skin_permit : 1.0f;
    float cws = cutoff + skin;
    nl->rl2 = cws * cws;
    nl->half_skin_square = (float)(0.25 * skin * skin);
    nl_set_box(nl, box_length);
    (void)hipGetDevice(&nl->device);

    double V = (double)box_length[0] * box_length[1] * box_length[2];
    double dens = atom_numbers / V;
    double w = std::cbrt((double)kClusterSize / dens);
    nl->ncx = std::max(1, (int)std::floor(box_length[0] / w));
    nl->ncy = std::max(1, (int)std::floor(box_length[1] / w));
    nl->ncol = nl->ncx * nl->ncy;
    double mean = (double)atom_numbers / nl->ncol;
    int cap = (int)std::ceil(2.5 * mean + 32);
    cap = (cap + 63) & ~63;
    nl->col_cap = std::min(std::max(cap, 64), kMaxColumn);
    nl->slot_cap = atom_numbers + 3 * nl->ncol;
    nl->slot_cap = (nl->slot_cap + 3) & ~3;
    nl->ncl_cap = nl->slot_cap / kClusterSize;
    double rl = cws;
    double est = (4.0 / 3.0) * M_PI * std::pow(rl + 2.0 * w, 3) * dens / kClusterSize;
    int lc = (int)std::ceil(2.0 * est) + 64;
    nl->list_cap = std::min(std::max(lc, 64), std::min(nl->ncl_cap, kCandCap));
    nl->cand_cap = kCandCap;

    size_t n = atom_numbers;
    hipError_t e = hipSuccess;
    auto A = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes > 0 ? bytes : 4);
    };
    A((void**)&nl->old_crd, 3 * n * sizeof(float));
    A((void**)&nl->nl_uint, 3 * n * sizeof(unsigned));
    A((void**)&nl->col_count, nl->ncol * sizeof(int));
    A((void**)&nl->col_start, (nl->ncol + 1) * sizeof(int));
    A((void**)&nl->col_atoms, (size_t)nl->ncol * nl->col_cap * sizeof(int));
    A((void**)&nl->s2o, nl->slot_cap * sizeof(int));
    A((void**)&nl->o2s, n * sizeof(int));
    A((void**)&nl->dropped, n * sizeof(int));
    A((void**)&nl->cl_lo, nl->ncl_cap * sizeof(float4));
    A((void**)&nl->cl_hi, nl->ncl_cap * sizeof(float4));
    A((void**)&nl->list_n, nl->ncl_cap * sizeof(int));
    A((void**)&nl->list, (size_t)nl->ncl_cap * nl->list_cap * sizeof(uint2));
    A((void**)&nl->list_in, (size_t)nl->ncl_cap * nl->list_cap * sizeof(uint2));
    A((void**)&nl->list_in_n, nl->ncl_cap * sizeof(int));
    {
        // partners of one slot within the list radius: the mean count at the
        // system's density with a quarter and 64 to spare (a slot in a denser
        // region sends its cluster to the full list); each slot's list starts
        // on a 128-byte line of its own
        const double rl3 = std::pow((double)cutoff + (double)skin, 3);
        const int est = (int)std::ceil(1.25 * (4.0 / 3.0) * M_PI * rl3 * dens) + 64;
        nl->atom_in_cap = (std::min(est, nl->slot_cap) + 63) & ~63;
    }
    A((void**)&nl->atom_in_n, nl->slot_cap * sizeof(int));
    A((void**)&nl->atom_in, (size_t)nl->slot_cap * nl->atom_in_cap * sizeof(unsigned short));
    A((void**)&nl->atom_out_n, nl->slot_cap * sizeof(int));
    A((void**)&nl->atom_out, (size_t)nl->slot_cap * nl->atom_in_cap * sizeof(unsigned short));
    A((void**)&nl->pack, nl->slot_cap * sizeof(uint4));
    A((void**)&nl->flags, 16 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->flags, 0, 16 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->list_n, 0, nl->ncl_cap * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->list_in_n, 0, nl->ncl_cap * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->atom_in_n, 0, nl->slot_cap * sizeof(int));
    if (e == hipSuccess)
        e = hipMemset(nl->atom_in, 0, (size_t)nl->slot_cap * nl->atom_in_cap * sizeof(unsigned short));
    if (e == hipSuccess) e = hipMemset(nl->atom_out_n, 0, nl->slot_cap * sizeof(int));
    if (e == hipSuccess)
        e = hipMemset(nl->atom_out, 0, (size_t)nl->slot_cap * nl->atom_in_cap * sizeof(unsigned short));
    // Inner list of the fused step (interval mode): pruned every
    // P = refresh_interval / prune_div fused steps to rc + margin * skin, and
    // at once on a step where an atom has moved more than half the margin
    // since the last prune; P and the margin are the measured defaults of
    // tune_table.hpp for this many atoms (sponge_nl_set_prune changes them).
    if (refresh_interval > 0) {
        const TuneRow& t = tune_row(atom_numbers);
        if (t.prune_div > 0) {
            nl->prune_interval = std::max(1, refresh_interval / t.prune_div);
            const float rp = cutoff + t.prune_margin * skin;
            nl->prune_r2 = rp * rp;
            nl->prune_vote_r2 = snb_prune_vote_r2(cutoff, rp);
        }
    }
    A((void**)&nl->prune_ref, nl->slot_cap * sizeof(uint4));
    if (e == hipSuccess) e = hipMemset(nl->prune_ref, 0, nl->slot_cap * sizeof(uint4));
    if (e == hipSuccess) e = hipMemset(nl->s2o, 0xff, nl->slot_cap * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->col_start, 0, (nl->ncol + 1) * sizeof(int));
    if (e == hipSuccess) e = hipMemset(nl->old_crd, 0, 3 * n * sizeof(float));
    if (e != hipSuccess) {
        set_error("sponge_nl_create: %s", hipGetErrorString(e));
        sponge_nl_destroy(nl);
        return SPONGE_NB_EHIP;
    }
    *out = nl;
    return SPONGE_NB_OK;
}

extern "C" void sponge_nl_destroy(sponge_nl* nl) {
    if (!nl) return;
    void* ptrs[] = {nl->old_crd, nl->nl_uint, nl->col_count, nl->col_start, nl->col_atoms, nl->s2o,
                    nl->o2s,     nl->cl_lo,   nl->cl_hi,     nl->list_n,    nl->list,      nl->pack,
                    nl->flags,   nl->list_in, nl->list_in_n, nl->sym_start, nl->sym_list, nl->atom_in_n,
                    nl->atom_in, nl->dropped, nl->prune_ref, nl->atom_out_n, nl->atom_out};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete nl;
}

// Squared displacement past which the fused step re-prunes its inner list:
// half the prune margin, less 1 % for the FP32 rounding of the distances
// compared at the prune.
float snb_prune_vote_r2(float cutoff, float prune_radius) {
    const float h = 0.5f * (prune_radius - cutoff);
    return 0.98f * h * h;
}

// Bounding-box search radius: rc + skin with a small margin for the FP32
// rounding of the cluster boxes.
float snb_nl_rlm(const sponge_nl* nl) { return std::sqrt(nl->rl2) + 1.0e-3f * (1.0f + std::sqrt(nl->rl2)); }

// Launches the rebuild pipeline; `forced` latches the decision on the host,
// otherwise the device flag written by nl_check_kernel decides.
static int nl_rebuild(sponge_nl* nl, float* d_crd, const int* es, const int* el, const int* en, int forced,
                      hipStream_t s) {
    const int n = nl->n;
    nl_begin_kernel<<<1, 256, 0, s>>>(forced, nl->flags, nl->col_count, nl->ncol);
    SNB_CHECK_LAUNCH();
    if (nl->ncol <= kBinHistMax)
        nl_bin_hist_kernel<<<ceil_div(n, 256), 256, (size_t)2 * nl->ncol * sizeof(int), s>>>(
            n, d_crd, nl->old_crd, nl->nl_uint, nl->box[0], nl->box[1], nl->box[2], nl->cof_q[0], nl->cof_q[1],
            nl->cof_q[2], nl->ncx, nl->ncy, nl->col_cap, nl->col_count, nl->col_atoms, nl->flags, nl->o2s,
            nl->dropped);
    else
        nl_bin_kernel<<<ceil_div(n, 256), 256, 0, s>>>(n, d_crd, nl->old_crd, nl->nl_uint, nl->box[0], nl->box[1],
                                                       nl->box[2], nl->cof_q[0], nl->cof_q[1], nl->cof_q[2], nl->ncx,
                                                       nl->ncy, nl->col_cap, nl->col_count, nl->col_atoms, nl->flags,
                                                       nl->o2s, nl->dropped);
    SNB_CHECK_LAUNCH();
    nl_scan_kernel<<<1, 1024, 0, s>>>(nl->ncol, nl->col_cap, nl->col_count, nl->col_start, nl->flags);
    SNB_CHECK_LAUNCH();
    nl_sort_kernel<<<nl->ncol, kSortThreads, 0, s>>>(nl->col_cap, nl->col_count, nl->col_start, nl->col_atoms,
                                                      nl->nl_uint, nl->cof_dr[0], nl->cof_dr[1], nl->cof_dr[2],
                                                      nl->s2o, nl->o2s, nl->pack, nl->cl_lo, nl->cl_hi, nl->flags);
    SNB_CHECK_LAUNCH();
    const float rlm = snb_nl_rlm(nl);
    int nblk = xcd_grid(ceil_div(nl->ncl_cap, kPairWaves));
    nl_pairs_kernel<<<nblk, kPairWaves * 64, 0, s>>>(
        nl->ncx, nl->ncy, nl->box[0], nl->box[1], nl->box[2], nl->cof_dr[0], nl->cof_dr[1], nl->cof_dr[2], nl->rl2,
        rlm, nl->col_start, nl->cl_lo, nl->cl_hi, nl->pack, es, en, el, nl->list_cap, nl->list_n, nl->list,
        nl->flags, nl->sym_start, nl->sym_list);
    SNB_CHECK_LAUNCH();
    if (unsigned short* sl = snb_nl_slot_lists(nl)) {
        nl_slot_lists_kernel<<<xcd_grid(ceil_div(nl->ncl_cap, 4)), 256, 0, s>>>(
            nl->flags, nl->list_n, nl->list, nl->list_cap, sl, snb_nl_slot_counts(nl), nl->atom_in_cap, nl->flags + 11);
        SNB_CHECK_LAUNCH();
    }
    return SPONGE_NB_OK;
}

// Symmetric exclusion lists from the reference's CSR (each excluded pair
// listed once, in the row of one of its atoms; MD_core/MD_core.cu:284-483):
// built on the host the first time these CSR arrays are passed (one
// synchronising copy), reused while the same arrays come back (the cache is
// keyed on the three device pointers: a caller that rewrites the CSR in
// place calls sponge_nl_set_exclusions). A first pass inside a stream
// capture is refused, since the copy would synchronise the capture.
static int nl_sym_exclusions(sponge_nl* nl, const int* es, const int* el, const int* en, hipStream_t s) {
    if (es == nl->ex_es && el == nl->ex_el && en == nl->ex_en && nl->sym_start) return SPONGE_NB_OK;
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            set_error("Neighbor_List_Update: new exclusion arrays inside a stream capture; pass them once "
                      "(or call sponge_nl_set_exclusions) before capturing");
            return SPONGE_NB_EARG;
        }
    }
    const int n = nl->n;
    std::vector<int> hs(n), hn(n);
    SNB_HIP(hipDeviceSynchronize());
    SNB_HIP(hipMemcpy(hs.data(), es, n * sizeof(int), hipMemcpyDeviceToHost));
    SNB_HIP(hipMemcpy(hn.data(), en, n * sizeof(int), hipMemcpyDeviceToHost));
    long end = 0;
    for (int i = 0; i < n; i++) {
        SNB_ARG(hn[i] >= 0 && hs[i] >= 0, "exclusion CSR: negative start or count");
        end = std::max(end, (long)hs[i] + hn[i]);
    }
    std::vector<int> hl(std::max(end, 1L));
    if (end > 0) SNB_HIP(hipMemcpy(hl.data(), el, end * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> start(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = hs[i]; k < hs[i] + hn[i]; k++) {
            SNB_ARG(hl[k] >= 0 && hl[k] < n, "exclusion CSR: partner index out of range");
            start[i + 1]++;
            start[hl[k] + 1]++;
        }
    for (int i = 0; i < n; i++) start[i + 1] += start[i];
    std::vector<int> pos(start.begin(), start.end() - 1), sym(std::max(start[n], 1));
    for (int i = 0; i < n; i++)
        for (int k = hs[i]; k < hs[i] + hn[i]; k++) {
            sym[pos[i]++] = hl[k];
            sym[pos[hl[k]]++] = i;
        }
    (void)hipFree(nl->sym_start);
    (void)hipFree(nl->sym_list);
    nl->sym_start = nullptr;
    nl->sym_list = nullptr;
    nl->ex_es = nl->ex_el = nl->ex_en = nullptr;
    SNB_HIP(hipMalloc((void**)&nl->sym_start, (n + 1) * sizeof(int)));
    SNB_HIP(hipMalloc((void**)&nl->sym_list, sym.size() * sizeof(int)));
    SNB_HIP(hipMemcpy(nl->sym_start, start.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    SNB_HIP(hipMemcpy(nl->sym_list, sym.data(), sym.size() * sizeof(int), hipMemcpyHostToDevice));
    nl->ex_es = es;
    nl->ex_el = el;
    nl->ex_en = en;
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_update(sponge_nl* nl, float* d_crd, const int* d_excluded_list_start,
                                const int* d_excluded_list, const int* d_excluded_numbers, int forced_update,
                                int forced_check, void* stream) {
    SNB_ARG(nl && d_crd && d_excluded_list_start && d_excluded_list && d_excluded_numbers, "nl_update");
    hipStream_t s = as_stream(stream);
    {
        const int rc = nl_sym_exclusions(nl, d_excluded_list_start, d_excluded_list, d_excluded_numbers, s);
        if (rc) return rc;
    }
    if (forced_update) {
        return nl_rebuild(nl, d_crd, d_excluded_list_start, d_excluded_list, d_excluded_numbers, 1, s);
    }
    if (nl->refresh_interval > 0 && !forced_check) {
        int rc = SPONGE_NB_OK;
        if (nl->refresh_count % nl->refresh_interval == 0)
            rc = nl_rebuild(nl, d_crd, d_excluded_list_start, d_excluded_list, d_excluded_numbers, 1, s);
        nl->refresh_count += 1;
        return rc;
    }
    float thresh = nl->skin_permit * nl->skin_permit * nl->half_skin_square;
    nl_check_kernel<<<ceil_div(nl->n, 256), 256, 0, s>>>(nl->n, d_crd, nl->old_crd, nl->box[0], nl->box[1],
                                                         nl->box[2], thresh, nl->flags);
    SNB_CHECK_LAUNCH();
    return nl_rebuild(nl, d_crd, d_excluded_list_start, d_excluded_list, d_excluded_numbers, 0, s);
}

extern "C" int sponge_nl_set_exclusions(sponge_nl* nl, const int* d_excluded_list_start, const int* d_excluded_list,
                                        const int* d_excluded_numbers) {
    SNB_ARG(nl && d_excluded_list_start && d_excluded_list && d_excluded_numbers, "nl_set_exclusions");
    nl->ex_es = nl->ex_el = nl->ex_en = nullptr;  // rebuild even for the same pointers
    return nl_sym_exclusions(nl, d_excluded_list_start, d_excluded_list, d_excluded_numbers, nullptr);
}

// The next fused step re-prunes: it sees its own step number in the vote
// word (flags[13]), as after a displacement vote, and rewrites the per-slot
// lists (flags[14]).
__global__ void nl_prune_now_kernel(int* flags) {
    if (threadIdx.x == 0) {
        flags[13] = flags[10] + 1;
        flags[14] = flags[10] + 1;
    }
}

extern "C" int sponge_nl_set_prune(sponge_nl* nl, int prune_interval, float prune_radius) {
    SNB_ARG(nl && prune_interval >= 0, "nl_set_prune");
    SNB_ARG(prune_interval == 0 || prune_radius >= nl->cutoff, "prune radius below the cutoff");
    nl->prune_interval = prune_interval;
    nl->prune_r2 = prune_radius * prune_radius;
    nl->prune_vote_r2 = snb_prune_vote_r2(nl->cutoff, prune_radius);
    // the inner lists were pruned to the old radius: rebuild them on the
    // next step (ordered on the null stream before any later launch)
    nl_prune_now_kernel<<<1, 64, 0, nullptr>>>(nl->flags);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_update_volume(sponge_nl* nl, const float box_length[3]) {
    SNB_ARG(nl && box_length, "nl_update_volume");
    for (int d = 0; d < 3; d++) SNB_ARG(box_length[d] > 0, "box_length");
    nl_set_box(nl, box_length);
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_status(sponge_nl* nl, long* pair_count, int* overflow_flags, int* rebuilds) {
    SNB_ARG(nl, "nl_status");
    int f[16];
    SNB_HIP(hipDeviceSynchronize());
    SNB_HIP(hipMemcpy(f, nl->flags, sizeof(f), hipMemcpyDeviceToHost));
    unsigned long long pc;
    memcpy(&pc, f + 8, sizeof(pc));
    if (pair_count) *pair_count = (long)(pc / 2);
    if (overflow_flags) *overflow_flags = f[1];
    if (rebuilds) *rebuilds = f[2];
    return f[1] ? SPONGE_NB_ECAPACITY : SPONGE_NB_OK;
}

extern "C" int sponge_nl_inner_status(sponge_nl* nl, int* slot_overflow, int* slot_launches) {
    SNB_ARG(nl, "nl_inner_status");
    int f[16];
    SNB_HIP(hipDeviceSynchronize());
    SNB_HIP(hipMemcpy(f, nl->flags, sizeof(f), hipMemcpyDeviceToHost));
    if (slot_overflow) *slot_overflow = f[11];
    if (slot_launches) *slot_launches = f[12];
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_prune_votes(sponge_nl* nl, int* voted_prunes) {
    SNB_ARG(nl && voted_prunes, "nl_prune_votes");
    int f[16];
    SNB_HIP(hipDeviceSynchronize());
    SNB_HIP(hipMemcpy(f, nl->flags, sizeof(f), hipMemcpyDeviceToHost));
    *voted_prunes = f[15];
    return SPONGE_NB_OK;
}

extern "C" int sponge_nl_get_pairs(sponge_nl* nl, int* h_pairs, long capacity, long* pair_count) {
    SNB_ARG(nl, "nl_get_pairs");
    int f[16];
    SNB_HIP(hipDeviceSynchronize());
    SNB_HIP(hipMemcpy(f, nl->flags, sizeof(f), hipMemcpyDeviceToHost));
    int nslot = f[3], ncl = f[4];
    std::vector<int> ln(ncl), s2o(nslot);
    std::vector<uint2> lst((size_t)ncl * nl->list_cap);
    if (ncl > 0) {
        SNB_HIP(hipMemcpy(ln.data(), nl->list_n, ncl * sizeof(int), hipMemcpyDeviceToHost));
        SNB_HIP(hipMemcpy(s2o.data(), nl->s2o, nslot * sizeof(int), hipMemcpyDeviceToHost));
        SNB_HIP(hipMemcpy(lst.data(), nl->list, lst.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    }
    // Every unordered pair must be listed from both of its atoms (the pair
    // kernel accumulates only its own atoms' forces): the pairs seen from
    // the lower atom and from the higher one are collected separately and
    // must be the same set.
    std::vector<unsigned long long> lo_side, hi_side;
    for (int ic = 0; ic < ncl; ic++)
        for (int e = 0; e < ln[ic]; e++) {
            uint2 ent = lst[(size_t)ic * nl->list_cap + e];
            for (int b = 0; b < 16; b++)
                if (ent.y >> b & 1u) {
                    int ii = b & 3, jj = b >> 2;
                    int a = s2o[ic * kClusterSize + ii], c = s2o[ent.x * kClusterSize + jj];
                    const unsigned long long key =
                        ((unsigned long long)(unsigned)std::min(a, c) << 32) | (unsigned)std::max(a, c);
                    (a < c ? lo_side : hi_side).push_back(key);
                }
        }
    std::sort(lo_side.begin(), lo_side.end());
    std::sort(hi_side.begin(), hi_side.end());
    const long m = (long)lo_side.size();
    if (h_pairs)
        for (long k = 0; k < std::min(m, capacity); k++) {
            h_pairs[2 * k] = (int)(lo_side[k] >> 32);
            h_pairs[2 * k + 1] = (int)(lo_side[k] & 0xffffffffu);
        }
    if (pair_count) *pair_count = m;
    if (f[1]) return SPONGE_NB_ECAPACITY;
    if (lo_side != hi_side) {
        set_error("neighbor list is not symmetric: %zu pairs listed from the lower atom, %zu from the higher",
                  lo_side.size(), hi_side.size());
        return SPONGE_NB_EINTERNAL;
    }
    return SPONGE_NB_OK;
}

// Batched Neighbor_List_Update: forced = rebuild every system; otherwise the
// skin check of every system votes into its own flag (skin-triggered mode).
int snb_batch_nl(sponge_batch* b, int forced, hipStream_t s) {
    if (!forced) {
        nl_check_batch<<<b->nblk[kStNlCheck], 256, 0, s>>>(b->d, b->map[kStNlCheck]);
        SNB_CHECK_LAUNCH();
    }
    nl_begin_batch<<<b->nblk[kStNlBegin], 256, 0, s>>>(b->d, b->map[kStNlBegin], forced);
    SNB_CHECK_LAUNCH();
    nl_bin_batch<<<b->nblk[kStNlBin], 256, 0, s>>>(b->d, b->map[kStNlBin]);
    SNB_CHECK_LAUNCH();
    nl_scan_batch<<<b->nblk[kStNlScan], 1024, 0, s>>>(b->d, b->map[kStNlScan]);
    SNB_CHECK_LAUNCH();
    nl_sort_batch<<<b->nblk[kStNlSort], kSortThreads, 0, s>>>(b->d, b->map[kStNlSort]);
    SNB_CHECK_LAUNCH();
    const int nmap = b->nblk[kStNlPairs];
    nl_pairs_batch<<<xcd_grid(nmap), kPairWaves * 64, 0, s>>>(b->d, b->map[kStNlPairs], nmap);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}
