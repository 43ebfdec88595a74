// FP64 CPU restatement of SPONGE's nonbonded hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
// product (sponge_amd/) never links or calls it.
//
// Every function cites the reference file:line whose formula it restates
// (paths relative to the reference's SPONGE/ directory). Arithmetic is FP64
// except where the reference's *definition* is an FP32 expression that the
// contract requires bit-for-bit (uint coordinates and the neighbor-list
// distance test), which are restated in FP32 with a fixed operation order
// (this file is compiled with -ffp-contract=off).
//
// Pins (see tests/test_oracle_*.py): closed forms (LJ dimer, Coulomb pair,
// excluded pair), brute force vs the reference cell-grid algorithm, numpy's
// pocketfft for the FFT, exact Ewald sums and the NaCl Madelung constant for
// the reciprocal part, finite differences for every force.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "settle_math.hpp"

typedef std::complex<double> cplx;
static const double PI_D = 3.14159265358979323846;

extern "C" {

// ---------------------------------------------------------------------------
// a2 Crd_To_Uint_Crd (common.cu:264-287) with the NEIGHBOR_LIST cof
// quarter_crd_to_uint_crd_cof = 0.25f * 2^32 / L (neighbor_list.cu:622).
// FP32 by definition: u = ((int)(x * cof)) << 2.
void orc_crd_to_uint(int n, const float* crd, const float* box, uint32_t* out) {
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (0.25f * 4294967296.0f) / box[d];
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            float t = crd[3 * i + d] * cof[d];
            int32_t ti = (int32_t)t;
            out[3 * i + d] = ((uint32_t)ti) << 2;
        }
}

// a4 Crd_Periodic_Map (common.cu:299-352): shift by at most one box length.
void orc_periodic_map(int n, float* crd, const float* box) {
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            float x = crd[3 * i + d];
            if (x >= 0) {
                if (!(x < box[d])) x = x - box[d];
            } else {
                x = x + box[d];
            }
            crd[3 * i + d] = x;
        }
}

// The neighbor-list distance test, FP32 as in Find_atom_neighbors_gridly
// (neighbor_list.cu:269-274) with cof = (1/2^32) * L (neighbor_list.cu:623).
// Fixed order: dr2 = (dx*dx + dy*dy) + dz*dz, every operation rounded.
static inline float nl_dr2(const uint32_t* a, const uint32_t* b, const float* cof) {
    float dx = (float)(int32_t)(b[0] - a[0]) * cof[0];
    float dy = (float)(int32_t)(b[1] - a[1]) * cof[1];
    float dz = (float)(int32_t)(b[2] - a[2]) * cof[2];
    float s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

static inline bool is_excluded(int a, int b, const int* es, const int* en, const int* el) {
    // a < b; exclusion lists are upper triangular and sorted (MD_core.cu:441).
    for (int k = es[a]; k < es[a] + en[a]; k++)
        if (el[k] == b) return true;
    return false;
}

// a5-a9 as a set: every unordered pair i<j whose FP32 dr2 < (rc+skin)^2,
// minus exclusions (Delete_Excluded_Atoms_Serial_In_Neighbor_List,
// neighbor_list.cu:353-403). Pairs written as (i, j), i<j, sorted. Returns
// the pair count; pairs is filled only when cap is large enough.
long orc_nl_brute(int n, const uint32_t* u, const float* box, float cutoff, float skin,
                  const int* es, const int* en, const int* el, int* pairs, long cap) {
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * box[d];
    float cws = cutoff + skin;
    float rl2 = cws * cws;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<int>> part(nt);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        for (int j = i + 1; j < n; j++) {
            if (nl_dr2(u + 3 * i, u + 3 * j, cof) < rl2 && !is_excluded(i, j, es, en, el)) {
                part[t].push_back(i);
                part[t].push_back(j);
            }
        }
    }
    std::vector<std::pair<int, int>> all;
    for (auto& p : part)
        for (size_t k = 0; k < p.size(); k += 2) all.push_back({p[k], p[k + 1]});
    std::sort(all.begin(), all.end());
    long m = (long)all.size();
    if (pairs && cap >= m)
        for (long k = 0; k < m; k++) { pairs[2 * k] = all[k].first; pairs[2 * k + 1] = all[k].second; }
    return m;
}

// The reference's own cell-grid search restated: grid sizing
// (neighbor_list.cu:12-25), 5x5x5 stencil with the small-box rules
// (:57-142), binning with the (n-N)>>31 mask (:165-180), the j>i filter and
// FP32 test (:260-281), then exclusion deletion. Returns the multiset of
// pairs (duplicates would show a stencil bug), sorted.
long orc_nl_grid(int n, const float* crd_in, const float* box, float cutoff, float skin,
                 const int* es, const int* en, const int* el, int* pairs, long cap, int* grid_dims) {
    std::vector<float> crd(crd_in, crd_in + 3 * n);
    orc_periodic_map(n, crd.data(), box);
    std::vector<uint32_t> u(3 * n);
    orc_crd_to_uint(n, crd.data(), box, u.data());
    float cws = cutoff + skin;
    float half = 0.5f * cws;
    int N[3];
    float glen_inv[3];
    for (int d = 0; d < 3; d++) {
        N[d] = (int)floorf(box[d] / half);
        float gl = (float)box[d] / N[d];
        glen_inv[d] = (float)(1. / gl);
    }
    if (grid_dims) { grid_dims[0] = N[0]; grid_dims[1] = N[1]; grid_dims[2] = N[2]; }
    int Nxy = N[0] * N[1], G = Nxy * N[2];
    std::vector<std::vector<int>> bucket(G + 1);
    for (int i = 0; i < n; i++) {
        int c[3];
        for (int d = 0; d < 3; d++) {
            int v = (int)(crd[3 * i + d] * glen_inv[d]);
            v = v & ((v - N[d]) >> 31);
            c[d] = v;
        }
        bucket[c[2] * Nxy + c[1] * N[0] + c[0]].push_back(i);
    }
    float cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * box[d];
    float rl2 = cws * cws;
    std::vector<std::pair<int, int>> all;
    for (int g = 0; g < G; g++) {
        int gz = g / Nxy, gy = (g - Nxy * gz) / N[0], gx = g - Nxy * gz - N[0] * gy;
        int gc[3] = {gx, gy, gz};
        std::vector<int> sten;
        for (int l = -2; l <= 2; l++)
            for (int m = -2; m <= 2; m++)
                for (int k = -2; k <= 2; k++) {
                    int off[3] = {l, m, k};
                    int cc[3];
                    bool out = false;
                    for (int d = 0; d < 3; d++) {
                        int v = gc[d] + off[d];
                        bool keep = N[d] >= 5 || (v >= 0 && v < N[d]) ||
                                    (N[d] == 4 && ((gc[d] == 0 && off[d] == -1) || (gc[d] == 3 && off[d] == 1)));
                        if (!keep) out = true;
                        if (!out) {
                            if (v < 0) v += N[d];
                            else if (v >= N[d]) v -= N[d];
                        }
                        cc[d] = v;
                    }
                    sten.push_back(out ? G : cc[2] * Nxy + cc[1] * N[0] + cc[0]);
                }
        std::sort(sten.begin(), sten.end());
        for (int s : sten)
            for (int a : bucket[g])
                for (int b : bucket[s])
                    if (b > a && nl_dr2(&u[3 * a], &u[3 * b], cof) < rl2) all.push_back({a, b});
    }
    std::vector<std::pair<int, int>> kept;
    for (auto& p : all)
        if (!is_excluded(p.first, p.second, es, en, el)) kept.push_back(p);
    std::sort(kept.begin(), kept.end());
    long m = (long)kept.size();
    if (pairs && cap >= m)
        for (long k = 0; k < m; k++) { pairs[2 * k] = kept[k].first; pairs[2 * k + 1] = kept[k].second; }
    return m;
}

// ---------------------------------------------------------------------------
// a13 LJ + PME direct pair terms (Lennard_Jones_force.cu:424-562), FP64, over
// a half list of pairs (i<j). The per-atom arrays follow the reference's
// half-list convention: every pair term is credited to atom i (the lower
// index). dr = r_j - r_i from the wrap-around uint difference (common.cu:89).
// The cutoff test is the FP32 expression of the neighbor list (nl_dr2 against
// rc*rc in FP32); the reference's norm3df(dr) < rc (:480-482) can differ only
// for pairs within one FP32 ulp of the cutoff, and a fixed FP32 predicate
// lets the GPU path and this oracle agree on every pair.
// Stored A, B are x12 / x6 (Lennard_Jones_force.cu:1092,1097).
//   F_i += (f_LJ - f_C) dr, F_j -= (f_LJ - f_C) dr
//   f_LJ = (-A r^-6 + B) r^-8,  f_C = q_i q_j (2/sqrt(pi) b r e^{-b^2r^2} + erfc(b r)) r^-3
//   E_LJ = A/12 r^-12 - B/6 r^-6, E_C = q_i q_j erfc(b r)/r, W_LJ = -f_LJ r^2
// totals[0]=E_LJ, [1]=E_direct, [2]=W_LJ, [3]=number of pairs inside rc.
static inline int pair_index(int a, int b) {
    int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

void orc_direct(int n, const uint32_t* u, const float* q, const int* type, const float* A,
                const float* B, const float* box, double rc, double beta, long npairs,
                const int* pairs, double* frc, double* atom_e, double* atom_vir,
                double* atom_dir, double* totals) {
    double cof[3];
    float coff[3];
    for (int d = 0; d < 3; d++) {
        cof[d] = (double)box[d] / 4294967296.0;
        coff[d] = (1.0f / 4294967296.0f) * box[d];
    }
    const float rcf = (float)rc;
    const float rc2f = rcf * rcf;
    const double two_over_sqrtpi = 2.0 / std::sqrt(PI_D);
    double e_lj = 0, e_c = 0, w_lj = 0, cnt = 0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> fpart(nt, std::vector<double>(3 * (size_t)n, 0.0));
    std::vector<std::vector<double>> epart(nt, std::vector<double>(3 * (size_t)n, 0.0));
#pragma omp parallel for reduction(+ : e_lj, e_c, w_lj, cnt) schedule(static)
    for (long k = 0; k < npairs; k++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int i = pairs[2 * k], j = pairs[2 * k + 1];
        double dr[3];
        for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        double r = std::sqrt(r2);
        if (!(nl_dr2(u + 3 * i, u + 3 * j, coff) < rc2f)) continue;
        int p = pair_index(type[i], type[j]);
        double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2, ri8 = ri6 * ri2;
        double flj = (-(double)A[p] * ri6 + (double)B[p]) * ri8;
        double br = beta * r;
        double qq = (double)q[i] * (double)q[j];
        double erfcv = std::erfc(br);
        double fc = qq * (two_over_sqrtpi * br * std::exp(-br * br) + erfcv) * ri2 / r;
        double f = flj - fc;
        for (int d = 0; d < 3; d++) {
            fpart[t][3 * (size_t)i + d] += f * dr[d];
            fpart[t][3 * (size_t)j + d] -= f * dr[d];
        }
        double elj = ((double)A[p] / 12.0 * ri6 - (double)B[p] / 6.0) * ri6;
        double ec = qq * erfcv / r;
        double wlj = -flj * r2;
        epart[t][3 * (size_t)i + 0] += elj + ec;
        epart[t][3 * (size_t)i + 1] += wlj;
        epart[t][3 * (size_t)i + 2] += ec;
        e_lj += elj;
        e_c += ec;
        w_lj += wlj;
        cnt += 1;
    }
    for (int t = 0; t < nt; t++)
        for (size_t a = 0; a < (size_t)n; a++) {
            for (int d = 0; d < 3; d++) frc[3 * a + d] += fpart[t][3 * a + d];
            if (atom_e) atom_e[a] += epart[t][3 * a + 0];
            if (atom_vir) atom_vir[a] += epart[t][3 * a + 1];
            if (atom_dir) atom_dir[a] += epart[t][3 * a + 2];
        }
    totals[0] = e_lj;
    totals[1] = e_c;
    totals[2] = w_lj;
    totals[3] = cnt;
}

// ---------------------------------------------------------------------------
// a27 soft-core LJ + PME direct pair terms (LJ_soft_core.cu:159-317 force,
// :320-489 energy, :491-662 virial), FP64, half list, same cutoff predicate
// and per-atom crediting as orc_direct. Per atom: LJ type in state A and in
// state B, and a subsystem mask. With lam = lambda, l_ = 1 - lambda, pA / pB
// the pair indices of the A / B types (tables stored x12 / x6):
//   soft = mask_i != mask_j or exactly one of AA[pA], BA[pB] is below 1e-6.
//   not soft: plain LJ with A = l_ AA + lam BA, B = l_ AB + lam BB, plain
//             Coulomb q_i q_j erfc(b r)/r.
//   soft:     sA^6 = r^6 + alpha lam^p sigA6, sB^6 = r^6 + alpha l_^p sigB6,
//             sig6 = soft_core_sigma^6 when A or B of the state is below 1e-6,
//             else max(A / (2B), sigma_min^6);
//             E_LJ = l_ (AA/12 sA^-12 - AB/6 sA^-6) + lam (BA/12 sB^-12 - BB/6 sB^-6)
//             E_C  = q_i q_j (l_ erfc(b sA)/sA + lam erfc(b sB)/sB)
//             f_LJ = r^4 (l_ (-AA sA^-6 + AB) sA^-12 + lam (-BA sB^-6 + BB) sB^-12)
//             f_C  = q_i q_j r^4 (l_ (2b/sqrt(pi) e^{-b^2 sA^2} + erfc(b sA)/sA) sA^-6 + (same for B))
//   F_i += (f_LJ - f_C) dr, W_LJ = -f_LJ r^2. totals as orc_direct.
void orc_direct_soft(int n, const uint32_t* u, const float* q, const int* typeA, const int* typeB,
                     const int* mask, const float* AA, const float* AB, const float* BA, const float* BB,
                     double lambda, double alpha, double p, double sigma, double sigma_min,
                     const float* box, double rc, double beta, long npairs, const int* pairs, double* frc,
                     double* atom_e, double* atom_vir, double* atom_dir, double* totals) {
    double cof[3];
    float coff[3];
    for (int d = 0; d < 3; d++) {
        cof[d] = (double)box[d] / 4294967296.0;
        coff[d] = (1.0f / 4294967296.0f) * box[d];
    }
    const float rcf = (float)rc;
    const float rc2f = rcf * rcf;
    const double two_over_sqrtpi = 2.0 / std::sqrt(PI_D);
    const double l_ = 1.0 - lambda;
    const double alp = alpha * std::pow(lambda, p), alp_ = alpha * std::pow(l_, p);
    const double sig6 = std::pow(sigma, 6.0), sig6min = std::pow(sigma_min, 6.0);
    double e_lj = 0, e_c = 0, w_lj = 0, cnt = 0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> fpart(nt, std::vector<double>(3 * (size_t)n, 0.0));
    std::vector<std::vector<double>> epart(nt, std::vector<double>(3 * (size_t)n, 0.0));
#pragma omp parallel for reduction(+ : e_lj, e_c, w_lj, cnt) schedule(static)
    for (long k = 0; k < npairs; k++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int i = pairs[2 * k], j = pairs[2 * k + 1];
        if (!(nl_dr2(u + 3 * i, u + 3 * j, coff) < rc2f)) continue;
        double dr[3];
        for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
        const double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        const double r = std::sqrt(r2);
        const int pa = pair_index(typeA[i], typeA[j]), pb = pair_index(typeB[i], typeB[j]);
        const double aa = AA[pa], ab = AB[pa], ba = BA[pb], bb = BB[pb];
        const double qq = (double)q[i] * (double)q[j];
        const bool soft = mask[i] != mask[j] || (ba > 1e-6 && aa < 1e-6) || (ba < 1e-6 && aa > 1e-6);
        double flj, fc, elj, ec;
        if (!soft) {
            const double A = l_ * aa + lambda * ba, B = l_ * ab + lambda * bb;
            const double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2, ri8 = ri6 * ri2;
            flj = (-A * ri6 + B) * ri8;
            const double br = beta * r, erfcv = std::erfc(br);
            fc = qq * (two_over_sqrtpi * br * std::exp(-br * br) + erfcv) * ri2 / r;
            elj = (A / 12.0 * ri6 - B / 6.0) * ri6;
            ec = qq * erfcv / r;
        } else {
            const double sa6 = (aa < 1e-6 || ab < 1e-6) ? sig6 : std::max(0.5 * aa / ab, sig6min);
            const double sb6 = (ba < 1e-6 || bb < 1e-6) ? sig6 : std::max(0.5 * ba / bb, sig6min);
            const double r4 = r2 * r2, r6 = r4 * r2;
            const double SA6 = r6 + alp * sa6, SB6 = r6 + alp_ * sb6;  // s^6
            const double iA6 = 1.0 / SA6, iB6 = 1.0 / SB6;
            const double sA = std::pow(SA6, 1.0 / 6.0), sB = std::pow(SB6, 1.0 / 6.0);
            flj = r4 * (l_ * (-aa * iA6 + ab) * iA6 * iA6 + lambda * (-ba * iB6 + bb) * iB6 * iB6);
            const double bA = beta * sA, bB = beta * sB;
            const double eA = std::erfc(bA), eB = std::erfc(bB);
            fc = qq * r4 *
                 (l_ * (two_over_sqrtpi * beta * std::exp(-bA * bA) + eA / sA) * iA6 +
                  lambda * (two_over_sqrtpi * beta * std::exp(-bB * bB) + eB / sB) * iB6);
            elj = l_ * (aa / 12.0 * iA6 - ab / 6.0) * iA6 + lambda * (ba / 12.0 * iB6 - bb / 6.0) * iB6;
            ec = qq * (l_ * eA / sA + lambda * eB / sB);
        }
        const double f = flj - fc;
        for (int d = 0; d < 3; d++) {
            fpart[t][3 * (size_t)i + d] += f * dr[d];
            fpart[t][3 * (size_t)j + d] -= f * dr[d];
        }
        const double wlj = -flj * r2;
        epart[t][3 * (size_t)i + 0] += elj + ec;
        epart[t][3 * (size_t)i + 1] += wlj;
        epart[t][3 * (size_t)i + 2] += ec;
        e_lj += elj;
        e_c += ec;
        w_lj += wlj;
        cnt += 1;
    }
    for (int t = 0; t < nt; t++)
        for (size_t a = 0; a < (size_t)n; a++) {
            for (int d = 0; d < 3; d++) frc[3 * a + d] += fpart[t][3 * a + d];
            if (atom_e) atom_e[a] += epart[t][3 * a + 0];
            if (atom_vir) atom_vir[a] += epart[t][3 * a + 1];
            if (atom_dir) atom_dir[a] += epart[t][3 * a + 2];
        }
    totals[0] = e_lj;
    totals[1] = e_c;
    totals[2] = w_lj;
    totals[3] = cnt;
}

// Long-range dispersion factor of the soft-core module (LJ_soft_core.cu:
// 88-118, 1588-1600): -2/3 pi / rc^3 / 6 * sum_ij (l_ AB[pA_ij] + lam BB[pB_ij]),
// every ordered pair (i, j) including i = j; divide by the volume for the energy.
double orc_lj_soft_lrc_factor(int n, const int* typeA, const int* typeB, int ntA, int ntB, const float* AB,
                              const float* BB, double lambda, double rc) {
    std::vector<double> ca(ntA, 0.0), cb(ntB, 0.0);
    std::vector<double> cab((size_t)ntA * ntB, 0.0);
    for (int i = 0; i < n; i++) cab[(size_t)typeA[i] * ntB + typeB[i]] += 1.0;
    double s = 0;
    for (int a1 = 0; a1 < ntA; a1++)
        for (int b1 = 0; b1 < ntB; b1++) {
            const double c1 = cab[(size_t)a1 * ntB + b1];
            if (c1 == 0) continue;
            for (int a2 = 0; a2 < ntA; a2++)
                for (int b2 = 0; b2 < ntB; b2++) {
                    const double c2 = cab[(size_t)a2 * ntB + b2];
                    if (c2 == 0) continue;
                    s += c1 * c2 * ((1.0 - lambda) * AB[pair_index(a1, a2)] + lambda * BB[pair_index(b1, b2)]);
                }
        }
    return s * (-2.0 / 3.0) * PI_D / (rc * rc * rc) / 6.0;
}

// FEP / TI pair terms of one frame at several lambda windows, each window a
// literal restatement of LJ_Soft_Core_Energy_With_Coulumb_Direct_CUDA
// (LJ_soft_core.cu:970-1135) and LJ_Soft_Core_With_Drect_Columb_dH_dlambda_CUDA
// (:1137-1284), FP64, half list with the same cutoff predicate as orc_direct.
// Window w: lam = lambdas[w], l_ = 1 - lam, charges q + (lam - lambda0) dq
// (dq may be NULL), alp = alpha lam^p, alp_ = alpha l_^p, alp1 = alpha
// lam^(p-1), alp1_ = alpha l_^(p-1), sig6 as orc_direct_soft.
//   ordinary pair: E_LJ = (A/12 r^-6 - B/6) r^-6 with A = l_ AA + lam BA,
//     B = l_ AB + lam BB; E_C = q_i q_j erfc(b r)/r;
//     dH_LJ = ((BA - AA)/12 r^-6 - (BB - AB)/6) r^-6;
//     dH_C (reference, :1271) = (dq_j q_i + q_j dq_i) erfc(b r);
//     dH_C (exact) = (dq_j q_i + q_j dq_i) erfc(b r)/r.
//   soft pair: uA = r^6 + alp sigA6, uB = r^6 + alp_ sigB6, s = u^(1/6);
//     energies keep state X only while s_X < rc (:1073, :1087):
//     E_LJ = l_ (AA/12 uA^-2 - AB/6 uA^-1) + lam (BA/12 uB^-2 - BB/6 uB^-1),
//     E_C = q_i q_j (l_ erfc(b sA)/sA + lam erfc(b sB)/sB);
//     dH_LJ = uB^-2 BA/12 - BB/6 uB^-1 - AA/12 uA^-2 + AB/6 uA^-1
//             + lam uB^-2 p sigB6 alp1_ (BA/6 uB^-1 - BB/6)
//             + l_ uA^-2 p sigA6 alp1 (-AA/6 uA^-1 + AB/6);
//     dH_C = q_i q_j [lam p sigB6 alp1_ uB^-1 (b/(3 sqrt pi) e^{-b^2 sB^2} + erfc(b sB)/(6 sB))
//                     - l_ p sigA6 alp1 uA^-1 (b/(3 sqrt pi) e^{-b^2 sA^2} + erfc(b sA)/(6 sA))
//                     - erfc(b sA)/sA + erfc(b sB)/sB]
//            + (dq_j q_i + q_j dq_i)(l_ erfc(b sA)/sA + lam erfc(b sB)/sB)   (both forms).
// truncate_soft = 0 drops the soft-core-radius test (the energy whose lambda
// derivative the dH/dlambda formulas are).
// out[w * 7 + k]: LJ intersys, LJ intrasys, direct intersys, direct intrasys,
// dH LJ, dH direct (reference form), dH direct (exact); intrasys = same mask.
void orc_fep_windows(int n, const uint32_t* u, const float* q, const float* dq, const int* typeA,
                     const int* typeB, const int* mask, const float* AA, const float* AB, const float* BA,
                     const float* BB, double lambda0, double alpha, double p, double sigma, double sigma_min,
                     const float* box, double rc, double beta, long npairs, const int* pairs, int nwin,
                     const double* lambdas, int truncate_soft, double* out) {
    double cof[3];
    float coff[3];
    for (int d = 0; d < 3; d++) {
        cof[d] = (double)box[d] / 4294967296.0;
        coff[d] = (1.0f / 4294967296.0f) * box[d];
    }
    const float rcf = (float)rc;
    const float rc2f = rcf * rcf;
    const double k3 = 1.0 / (3.0 * std::sqrt(PI_D));
    const double sig6 = std::pow(sigma, 6.0), sig6min = std::pow(sigma_min, 6.0);
    for (int w = 0; w < nwin; w++) {
        const double lam = lambdas[w], l_ = 1.0 - lam, t = lam - lambda0;
        const double alp = alpha * std::pow(lam, p), alp_ = alpha * std::pow(l_, p);
        const double alp1 = alpha * std::pow(lam, p - 1.0), alp1_ = alpha * std::pow(l_, p - 1.0);
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2, s3, s4, s5, s6) schedule(static)
        for (long k = 0; k < npairs; k++) {
            const int i = pairs[2 * k], j = pairs[2 * k + 1];
            if (!(nl_dr2(u + 3 * i, u + 3 * j, coff) < rc2f)) continue;
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
            const double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            const double r = std::sqrt(r2);
            const int pa = pair_index(typeA[i], typeA[j]), pb = pair_index(typeB[i], typeB[j]);
            const double aa = AA[pa], ab = AB[pa], ba = BA[pb], bb = BB[pb];
            const double dqi = dq ? (double)dq[i] : 0.0, dqj = dq ? (double)dq[j] : 0.0;
            const double qi = (double)q[i] + t * dqi, qj = (double)q[j] + t * dqj;
            const double cross = dqj * qi + qj * dqi;
            const bool soft = mask[i] != mask[j] || (ba > 1e-6 && aa < 1e-6) || (ba < 1e-6 && aa > 1e-6);
            double elj, ec, dlj, dref, dexact;
            if (!soft) {
                const double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2;
                const double A = l_ * aa + lam * ba, B = l_ * ab + lam * bb;
                const double erfcv = std::erfc(beta * r);
                elj = (A / 12.0 * ri6 - B / 6.0) * ri6;
                ec = qi * qj * erfcv / r;
                dlj = ((ba - aa) / 12.0 * ri6 - (bb - ab) / 6.0) * ri6;
                dref = erfcv * cross;
                dexact = erfcv / r * cross;
            } else {
                const double sa6 = (aa < 1e-6 || ab < 1e-6) ? sig6 : std::max(0.5 * aa / ab, sig6min);
                const double sb6 = (ba < 1e-6 || bb < 1e-6) ? sig6 : std::max(0.5 * ba / bb, sig6min);
                const double r6 = r2 * r2 * r2;
                const double uA = r6 + alp * sa6, uB = r6 + alp_ * sb6;
                const double iA = 1.0 / uA, iB = 1.0 / uB;
                const double sA = std::pow(uA, 1.0 / 6.0), sB = std::pow(uB, 1.0 / 6.0);
                const double gA = std::erfc(beta * sA) / sA, gB = std::erfc(beta * sB) / sB;
                const bool inA = !truncate_soft || sA < rc, inB = !truncate_soft || sB < rc;
                const double vA = inA ? (aa / 12.0 * iA - ab / 6.0) * iA : 0.0;
                const double vB = inB ? (ba / 12.0 * iB - bb / 6.0) * iB : 0.0;
                const double eA = inA ? gA : 0.0, eB = inB ? gB : 0.0;
                elj = l_ * vA + lam * vB;
                ec = qi * qj * (l_ * eA + lam * eB);
                dlj = iB * iB * ba / 12.0 - bb / 6.0 * iB - aa / 12.0 * iA * iA + ab / 6.0 * iA +
                      lam * iB * iB * p * sb6 * alp1_ * (ba / 6.0 * iB - bb / 6.0) +
                      l_ * iA * iA * p * sa6 * alp1 * (-aa / 6.0 * iA + ab / 6.0);
                double dc = lam * p * sb6 * alp1_ * iB * (beta * k3 * std::exp(-beta * beta * sB * sB) + gB / 6.0);
                dc -= l_ * p * sa6 * alp1 * iA * (beta * k3 * std::exp(-beta * beta * sA * sA) + gA / 6.0);
                dc += gB - gA;
                dc = dc * qi * qj + cross * (l_ * gA + lam * gB);
                dref = dc;
                dexact = dc;
            }
            if (mask[i] == mask[j]) {
                s1 += elj;
                s3 += ec;
            } else {
                s0 += elj;
                s2 += ec;
            }
            s4 += dlj;
            s5 += dref;
            s6 += dexact;
        }
        double* o = out + 7 * (size_t)w;
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
        o[3] = s3;
        o[4] = s4;
        o[5] = s5;
        o[6] = s6;
    }
}

// a24 exclusion correction (PME_force.cu:620-695): for every excluded pair,
// E -= q_i q_j erf(b r)/r, F_i += -q_i q_j (2/sqrt(pi) b r e^{-b^2r^2} + erfc(b r) - 1)/r^3 dr.
// Energy credited to atom i (the list owner). Returns the total.
double orc_excluded(int n, const uint32_t* u, const float* q, const float* box, double beta,
                    const int* es, const int* en, const int* el, double* frc, double* atom_corr) {
    double cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (double)box[d] / 4294967296.0;
    const double two_over_sqrtpi = 2.0 / std::sqrt(PI_D);
    double tot = 0;
    for (int i = 0; i < n; i++) {
        for (int k = es[i]; k < es[i] + en[i]; k++) {
            int j = el[k];
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
            double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            double r = std::sqrt(r2);
            double br = beta * r;
            double qq = (double)q[i] * (double)q[j];
            double f = -qq * (two_over_sqrtpi * br * std::exp(-br * br) + std::erfc(br) - 1.0) / r2 / r;
            for (int d = 0; d < 3; d++) {
                frc[3 * i + d] += f * dr[d];
                frc[3 * j + d] -= f * dr[d];
            }
            double e = -qq * std::erf(br) / r;
            tot += e;
            if (atom_corr) atom_corr[i] += e;
        }
    }
    return tot;
}

// PME_Cross_Excluded_Energy_Correction (TI_core.cu:69-125), FP64: the lambda
// derivative of the exclusion correction when the charges move along dq,
// -sum over excluded pairs (q_i dq_j + dq_i q_j) erf(b r)/r.
double orc_excluded_cross(int n, const uint32_t* u, const float* q, const float* dq, const float* box,
                          double beta, const int* es, const int* en, const int* el) {
    double cof[3];
    for (int d = 0; d < 3; d++) cof[d] = (double)box[d] / 4294967296.0;
    double tot = 0;
    for (int i = 0; i < n; i++) {
        for (int k = es[i]; k < es[i] + en[i]; k++) {
            const int j = el[k];
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
            const double r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
            tot -= ((double)q[i] * (double)dq[j] + (double)dq[i] * (double)q[j]) * std::erf(beta * r) / r;
        }
    }
    return tot;
}

// ---------------------------------------------------------------------------
// a16 PME setup. Get_Beta (PME_force.cu:71-100) restated in FP64: double the
// upper bound until erfc(b rc)/rc <= tol, then (iterations+50) bisection
// steps.
double orc_pme_beta(double rc, double tol) {
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

// Cardinal B-spline M_n(u) (PME_force.cu:21-31), recursive definition.
static double bspline_M(double uu, int n) {
    if (n == 2) return (uu < 0 || uu > 2) ? 0.0 : 1.0 - std::fabs(uu - 1.0);
    return uu / (n - 1) * bspline_M(uu, n - 1) + (n - uu) / (n - 1) * bspline_M(uu - 1, n - 1);
}

// Euler exponential spline modulus |b(k)|^2 (getb, PME_force.cu:44-67;
// Essmann et al. 1995 eq. 4.4), order 4.
double orc_bspline_mod2(int k, int K) {
    const int order = 4;
    cplx den(0, 0);
    for (int kk = 0; kk < order - 1; kk++)
        den += bspline_M(kk + 1, order) * std::exp(cplx(0, 2 * PI_D * k * kk / K));
    return 1.0 / std::norm(den);
}

// BC table over the full spectrum (kx, ky, kz all in [0,K)), FP64, folded
// |m| as in PME_force.cu:238-286:
//   BC(m) = exp(-pi^2 |m|^2 / b^2) / (pi |m|^2 V) * B1 B2 B3, BC(0) = 0.
void orc_pme_bc_full(int Kx, int Ky, int Kz, const float* box, double beta, double* bc) {
    double V = (double)box[0] * box[1] * box[2];
    std::vector<double> B1(Kx), B2(Ky), B3(Kz);
    for (int k = 0; k < Kx; k++) B1[k] = orc_bspline_mod2(k, Kx);
    for (int k = 0; k < Ky; k++) B2[k] = orc_bspline_mod2(k, Ky);
    for (int k = 0; k < Kz; k++) B3[k] = orc_bspline_mod2(k, Kz);
    for (int x = 0; x < Kx; x++) {
        int xr = x > Kx / 2 ? Kx - x : x;
        for (int y = 0; y < Ky; y++) {
            int yr = y > Ky / 2 ? Ky - y : y;
            for (int z = 0; z < Kz; z++) {
                int zr = z > Kz / 2 ? Kz - z : z;
                double msq = (double)xr * xr / ((double)box[0] * box[0]) +
                             (double)yr * yr / ((double)box[1] * box[1]) +
                             (double)zr * zr / ((double)box[2] * box[2]);
                size_t idx = ((size_t)x * Ky + y) * Kz + z;
                if (x == 0 && y == 0 && z == 0) bc[idx] = 0;
                else bc[idx] = std::exp(-PI_D * PI_D * msq / (beta * beta)) / (PI_D * msq * V) * B1[x] * B2[y] * B3[z];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 mixed-radix FFT (recursive decimation in time; any length, prime
// factors handled by a direct DFT of that radix). sign=-1 forward.
static void fft1d_rec(cplx* x, int n, int stride, cplx* out, int sign) {
    if (n == 1) { out[0] = x[0]; return; }
    int r = 0;
    for (int p : {4, 2, 3, 5, 7})
        if (n % p == 0) { r = p; break; }
    if (r == 0) {
        for (int p = 11; p <= n; p++)
            if (n % p == 0) { r = p; break; }
    }
    int m = n / r;
    std::vector<cplx> sub((size_t)n);
    for (int q = 0; q < r; q++) fft1d_rec(x + (size_t)q * stride, m, stride * r, sub.data() + (size_t)q * m, sign);
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < r; s++) {
            cplx acc(0, 0);
            for (int q = 0; q < r; q++) {
                double ang = sign * 2.0 * PI_D * (double)q * (k + (double)s * m) / n;
                acc += sub[(size_t)q * m + k] * cplx(std::cos(ang), std::sin(ang));
            }
            out[k + (size_t)s * m] = acc;
        }
    }
}

static void fft3d(std::vector<cplx>& a, int Kx, int Ky, int Kz, int sign) {
    int nmax = std::max(Kx, std::max(Ky, Kz));
#pragma omp parallel
    {
        std::vector<cplx> in(nmax), out(nmax);
#pragma omp for
        for (int l = 0; l < Kx * Ky; l++) {
            cplx* p = a.data() + (size_t)l * Kz;
            for (int z = 0; z < Kz; z++) in[z] = p[z];
            fft1d_rec(in.data(), Kz, 1, out.data(), sign);
            for (int z = 0; z < Kz; z++) p[z] = out[z];
        }
#pragma omp for
        for (int l = 0; l < Kx * Kz; l++) {
            int x = l / Kz, z = l % Kz;
            for (int y = 0; y < Ky; y++) in[y] = a[((size_t)x * Ky + y) * Kz + z];
            fft1d_rec(in.data(), Ky, 1, out.data(), sign);
            for (int y = 0; y < Ky; y++) a[((size_t)x * Ky + y) * Kz + z] = out[y];
        }
#pragma omp for
        for (int l = 0; l < Ky * Kz; l++) {
            int y = l / Kz, z = l % Kz;
            for (int x = 0; x < Kx; x++) in[x] = a[((size_t)x * Ky + y) * Kz + z];
            fft1d_rec(in.data(), Kx, 1, out.data(), sign);
            for (int x = 0; x < Kx; x++) a[((size_t)x * Ky + y) * Kz + z] = out[x];
        }
    }
}

// Forward 3D real-to-complex transform of a row-major real grid (z fastest),
// unnormalised, returning the half spectrum kz in [0, Kz/2] as interleaved
// (re, im) -- the cufftExecR2C layout the reference relies on
// (PME_force.cu:159, :741).
void orc_fft3_r2c(int Kx, int Ky, int Kz, const double* in, double* out) {
    std::vector<cplx> a((size_t)Kx * Ky * Kz);
    for (size_t k = 0; k < a.size(); k++) a[k] = cplx(in[k], 0);
    fft3d(a, Kx, Ky, Kz, -1);
    int hz = Kz / 2 + 1;
    for (int x = 0; x < Kx; x++)
        for (int y = 0; y < Ky; y++)
            for (int z = 0; z < hz; z++) {
                cplx v = a[((size_t)x * Ky + y) * Kz + z];
                size_t o = ((size_t)x * Ky + y) * hz + z;
                out[2 * o] = v.real();
                out[2 * o + 1] = v.imag();
            }
}

// Inverse (unnormalised) complex-to-real transform from the half spectrum,
// the cufftExecC2R convention (PME_force.cu:747).
void orc_fft3_c2r(int Kx, int Ky, int Kz, const double* in, double* out) {
    int hz = Kz / 2 + 1;
    std::vector<cplx> a((size_t)Kx * Ky * Kz);
    for (int x = 0; x < Kx; x++)
        for (int y = 0; y < Ky; y++)
            for (int z = 0; z < Kz; z++) {
                cplx v;
                if (z < hz) {
                    size_t o = ((size_t)x * Ky + y) * hz + z;
                    v = cplx(in[2 * o], in[2 * o + 1]);
                } else {
                    int xm = (Kx - x) % Kx, ym = (Ky - y) % Ky, zm = Kz - z;
                    size_t o = ((size_t)xm * Ky + ym) * hz + zm;
                    v = std::conj(cplx(in[2 * o], in[2 * o + 1]));
                }
                a[((size_t)x * Ky + y) * Kz + z] = v;
            }
    fft3d(a, Kx, Ky, Kz, +1);
    for (size_t k = 0; k < a.size(); k++) out[k] = a[k].real();
}

// ---------------------------------------------------------------------------
// a17-a23 SPME reciprocal part, FP64. Fractional grid coordinate
// t = u K / 2^32 (PME_Atom_Near, PME_force.cu:373-386, exact in FP64);
// grid point (base - k) mod K gets weight M4(frac + k) (PME_Q_Spread :428-471);
// convolution with BC over the full spectrum; E = 1/2 sum Q (theta*Q)
// (:753-755); F = -q sum (theta*Q) dM4 (K/L) (PME_Final :485-550).
// Returns E_rec; frc += reciprocal forces; optional outputs: the charge grid
// Q and the convolved grid (theta*Q).
double orc_pme_reciprocal(int n, const uint32_t* u, const float* q, const float* box, int Kx,
                          int Ky, int Kz, double beta, double* frc, double* Q_out,
                          double* conv_out) {
    int K[3] = {Kx, Ky, Kz};
    size_t NG = (size_t)Kx * Ky * Kz;
    std::vector<double> Q(NG, 0.0);
    std::vector<int> base(3 * (size_t)n);
    std::vector<double> frac(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double t = (double)u[3 * i + d] * (double)K[d] / 4294967296.0;
            double b = std::floor(t);
            base[3 * i + d] = (int)b;
            frac[3 * i + d] = t - b;
        }
    auto w4 = [](double f, int k) { return bspline_M(f + k, 4); };
    auto dw4 = [](double f, int k) {
        double uu = f + k;
        return bspline_M(uu, 3) - bspline_M(uu - 1, 3);   // d/du M_n = M_{n-1}(u) - M_{n-1}(u-1)
    };
    for (int i = 0; i < n; i++) {
        double wx[4], wy[4], wz[4];
        for (int k = 0; k < 4; k++) {
            wx[k] = w4(frac[3 * i], k);
            wy[k] = w4(frac[3 * i + 1], k);
            wz[k] = w4(frac[3 * i + 2], k);
        }
        for (int a = 0; a < 4; a++) {
            int gx = ((base[3 * i] - a) % Kx + Kx) % Kx;
            for (int b = 0; b < 4; b++) {
                int gy = ((base[3 * i + 1] - b) % Ky + Ky) % Ky;
                for (int c = 0; c < 4; c++) {
                    int gz = ((base[3 * i + 2] - c) % Kz + Kz) % Kz;
                    Q[((size_t)gx * Ky + gy) * Kz + gz] += (double)q[i] * wx[a] * wy[b] * wz[c];
                }
            }
        }
    }
    std::vector<double> bc(NG);
    orc_pme_bc_full(Kx, Ky, Kz, box, beta, bc.data());
    std::vector<cplx> a(NG);
    for (size_t k = 0; k < NG; k++) a[k] = cplx(Q[k], 0);
    fft3d(a, Kx, Ky, Kz, -1);
    for (size_t k = 0; k < NG; k++) a[k] *= bc[k];
    fft3d(a, Kx, Ky, Kz, +1);
    double E = 0;
    for (size_t k = 0; k < NG; k++) E += Q[k] * a[k].real();
    E *= 0.5;
    if (Q_out) for (size_t k = 0; k < NG; k++) Q_out[k] = Q[k];
    if (conv_out) for (size_t k = 0; k < NG; k++) conv_out[k] = a[k].real();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
        for (int k = 0; k < 4; k++) {
            wx[k] = w4(frac[3 * i], k); dx[k] = dw4(frac[3 * i], k);
            wy[k] = w4(frac[3 * i + 1], k); dy[k] = dw4(frac[3 * i + 1], k);
            wz[k] = w4(frac[3 * i + 2], k); dz[k] = dw4(frac[3 * i + 2], k);
        }
        double F[3] = {0, 0, 0};
        for (int aa = 0; aa < 4; aa++) {
            int gx = ((base[3 * i] - aa) % Kx + Kx) % Kx;
            for (int b = 0; b < 4; b++) {
                int gy = ((base[3 * i + 1] - b) % Ky + Ky) % Ky;
                for (int c = 0; c < 4; c++) {
                    int gz = ((base[3 * i + 2] - c) % Kz + Kz) % Kz;
                    double phi = a[((size_t)gx * Ky + gy) * Kz + gz].real();
                    F[0] += phi * dx[aa] * wy[b] * wz[c];
                    F[1] += phi * wx[aa] * dy[b] * wz[c];
                    F[2] += phi * wx[aa] * wy[b] * dz[c];
                }
            }
        }
        for (int d = 0; d < 3; d++) frc[3 * i + d] += -(double)q[i] * F[d] * (double)K[d] / (double)box[d];
    }
    return E;
}

// ---------------------------------------------------------------------------
// Gradients of the nonbonded energy and of a force-matching loss with respect
// to per-atom charges and the (stored, x12 / x6) LJ pair tables. The reference
// has no adjoint; these are the analytic derivatives of the formulas restated
// above (orc_direct, orc_excluded, orc_pme_reciprocal, orc_pme_self,
// orc_lj_lrc_factor), FP64, over the same half list and cutoff predicate.
//
// Energy: dE/dq_i = sum_j q_j erfc(b r)/r - sum_excl q_j erf(b r)/r
//                   + sum_g R[g] M_i(g) - 2b/sqrt(pi) q_i + 2 nf sum q,
//         dE/dA_p = sum_{pairs of type p} r^-12/12,
//         dE/dB_p = -sum r^-6/6 + d(lrc/V)/dB_p,
// with R = theta*Q the convolved grid and nf = -pi/(2 b^2 V).
// Loss L = sum_i |F_i - T_i|^2, w_i = 2 (F_i - T_i), d = r_j - r_i:
//   pair:      dL/dq_i += -q_j g(r) d.(w_i - w_j),  g = (2b/sqrt(pi) r e^{-b^2r^2} + erfc(b r))/r^3
//              dL/dA_p += -r^-14 d.(w_i - w_j),     dL/dB_p += r^-8 d.(w_i - w_j)
//   excluded:  dL/dq_i += -q_j h(r) d.(w_i - w_j),  h = g - 1/r^3 (and i <-> j)
//   reciprocal: dL/dq_k += -sum_g R[g] w_k.grad M_k(g) - sum_g (theta*D)[g] M_k(g),
//              D[g] = sum_i q_i w_i.grad M_i(g)  (grad M carries K/L per axis).
static void pme_frac_base(int n, const uint32_t* u, const int* K, std::vector<int>& base, std::vector<double>& frac) {
    base.resize(3 * (size_t)n);
    frac.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double t = (double)u[3 * i + d] * (double)K[d] / 4294967296.0;
            double b = std::floor(t);
            base[3 * i + d] = (int)b;
            frac[3 * i + d] = t - b;
        }
}

// sum over the 64 grid points of atom i of grid[g] * (weight chosen by mode):
// mode 0: M(g); mode 1: w.grad M(g) with wv = (w_x Kx/Lx, w_y Ky/Ly, w_z Kz/Lz).
static double pme_atom_gather(const double* grid, const int* base, const double* frac, const int* K,
                              const double* wv, int mode) {
    double wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
    for (int k = 0; k < 4; k++) {
        wx[k] = bspline_M(frac[0] + k, 4);
        wy[k] = bspline_M(frac[1] + k, 4);
        wz[k] = bspline_M(frac[2] + k, 4);
        dx[k] = bspline_M(frac[0] + k, 3) - bspline_M(frac[0] + k - 1, 3);
        dy[k] = bspline_M(frac[1] + k, 3) - bspline_M(frac[1] + k - 1, 3);
        dz[k] = bspline_M(frac[2] + k, 3) - bspline_M(frac[2] + k - 1, 3);
    }
    double s = 0;
    for (int a = 0; a < 4; a++) {
        int gx = ((base[0] - a) % K[0] + K[0]) % K[0];
        for (int b = 0; b < 4; b++) {
            int gy = ((base[1] - b) % K[1] + K[1]) % K[1];
            for (int c = 0; c < 4; c++) {
                int gz = ((base[2] - c) % K[2] + K[2]) % K[2];
                double m = mode == 0 ? wx[a] * wy[b] * wz[c]
                                     : wv[0] * dx[a] * wy[b] * wz[c] + wv[1] * wx[a] * dy[b] * wz[c] +
                                           wv[2] * wx[a] * wy[b] * dz[c];
                s += grid[((size_t)gx * K[1] + gy) * K[2] + gz] * m;
            }
        }
    }
    return s;
}

void orc_energy_grad(int n, const uint32_t* u, const float* q, const int* type, int ntypes, const float* A,
                     const float* B, const float* box, double rc, double beta, long npairs, const int* pairs,
                     const int* es, const int* en, const int* el, int Kx, int Ky, int Kz, double* gq, double* gA,
                     double* gB) {
    double cof[3];
    float coff[3];
    for (int d = 0; d < 3; d++) {
        cof[d] = (double)box[d] / 4294967296.0;
        coff[d] = (1.0f / 4294967296.0f) * box[d];
    }
    const float rcf = (float)rc;
    const float rc2f = rcf * rcf;
    const int P = ntypes * (ntypes + 1) / 2;
    for (int i = 0; i < n; i++) gq[i] = 0;
    for (int p = 0; p < P; p++) gA[p] = gB[p] = 0;
    for (long k = 0; k < npairs; k++) {
        int i = pairs[2 * k], j = pairs[2 * k + 1];
        if (!(nl_dr2(u + 3 * i, u + 3 * j, coff) < rc2f)) continue;
        double dr[3];
        for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2], r = std::sqrt(r2);
        double pot = std::erfc(beta * r) / r;
        gq[i] += (double)q[j] * pot;
        gq[j] += (double)q[i] * pot;
        double ri6 = 1.0 / (r2 * r2 * r2);
        int p = pair_index(type[i], type[j]);
        gA[p] += ri6 * ri6 / 12.0;
        gB[p] += -ri6 / 6.0;
    }
    for (int i = 0; i < n; i++)
        for (int k = es[i]; k < es[i] + en[i]; k++) {
            int j = el[k];
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
            double r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
            double pot = std::erf(beta * r) / r;
            gq[i] -= (double)q[j] * pot;
            gq[j] -= (double)q[i] * pot;
        }
    int K[3] = {Kx, Ky, Kz};
    size_t NG = (size_t)Kx * Ky * Kz;
    std::vector<double> R(NG), frc(3 * (size_t)n, 0.0);
    orc_pme_reciprocal(n, u, q, box, Kx, Ky, Kz, beta, frc.data(), nullptr, R.data());
    std::vector<int> base;
    std::vector<double> frac;
    pme_frac_base(n, u, K, base, frac);
    double sq = 0;
    for (int i = 0; i < n; i++) sq += q[i];
    const double V = (double)box[0] * box[1] * box[2];
    const double nf = -PI_D / (2.0 * beta * beta * V);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        gq[i] += pme_atom_gather(R.data(), &base[3 * (size_t)i], &frac[3 * (size_t)i], K, nullptr, 0) -
                 2.0 * beta / std::sqrt(PI_D) * q[i] + 2.0 * nf * sq;
    // long-range dispersion: lrc/V = c sum_{ordered ij} B[p_ij], c = -2/3 pi / rc^3 / 6 / V
    std::vector<double> cnt(ntypes, 0.0);
    for (int i = 0; i < n; i++) cnt[type[i]] += 1;
    const double c = (-2.0 / 3.0) * PI_D / (rc * rc * rc) / 6.0 / V;
    for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++) gB[pair_index(a, b)] += c * cnt[a] * cnt[b];
}

double orc_loss_grad(int n, const uint32_t* u, const float* q, const int* type, int ntypes, const float* A,
                     const float* B, const float* box, double rc, double beta, long npairs, const int* pairs,
                     const int* es, const int* en, const int* el, int Kx, int Ky, int Kz, const double* F,
                     const double* T, double* gq, double* gA, double* gB) {
    (void)A;
    (void)B;
    double cof[3];
    float coff[3];
    for (int d = 0; d < 3; d++) {
        cof[d] = (double)box[d] / 4294967296.0;
        coff[d] = (1.0f / 4294967296.0f) * box[d];
    }
    const float rcf = (float)rc;
    const float rc2f = rcf * rcf;
    const double tsp = 2.0 / std::sqrt(PI_D);
    const int P = ntypes * (ntypes + 1) / 2;
    std::vector<double> w(3 * (size_t)n);
    double L = 0;
    for (size_t a = 0; a < 3 * (size_t)n; a++) {
        double e = F[a] - T[a];
        L += e * e;
        w[a] = 2.0 * e;
    }
    for (int i = 0; i < n; i++) gq[i] = 0;
    for (int p = 0; p < P; p++) gA[p] = gB[p] = 0;
    for (long k = 0; k < npairs; k++) {
        int i = pairs[2 * k], j = pairs[2 * k + 1];
        if (!(nl_dr2(u + 3 * i, u + 3 * j, coff) < rc2f)) continue;
        double dr[3];
        for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2], r = std::sqrt(r2);
        double sd = 0;
        for (int d = 0; d < 3; d++) sd += dr[d] * (w[3 * (size_t)i + d] - w[3 * (size_t)j + d]);
        double br = beta * r;
        double g = (tsp * br * std::exp(-br * br) + std::erfc(br)) / (r2 * r);
        gq[i] += -(double)q[j] * g * sd;
        gq[j] += -(double)q[i] * g * sd;
        double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2, ri8 = ri6 * ri2;
        int p = pair_index(type[i], type[j]);
        gA[p] += -ri6 * ri8 * sd;
        gB[p] += ri8 * sd;
    }
    for (int i = 0; i < n; i++)
        for (int k = es[i]; k < es[i] + en[i]; k++) {
            int j = el[k];
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = (double)(int32_t)(u[3 * j + d] - u[3 * i + d]) * cof[d];
            double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2], r = std::sqrt(r2);
            double sd = 0;
            for (int d = 0; d < 3; d++) sd += dr[d] * (w[3 * (size_t)i + d] - w[3 * (size_t)j + d]);
            double br = beta * r;
            double h = (tsp * br * std::exp(-br * br) + std::erfc(br) - 1.0) / (r2 * r);
            gq[i] += -(double)q[j] * h * sd;
            gq[j] += -(double)q[i] * h * sd;
        }
    int K[3] = {Kx, Ky, Kz};
    size_t NG = (size_t)Kx * Ky * Kz;
    std::vector<double> R(NG), frc(3 * (size_t)n, 0.0);
    orc_pme_reciprocal(n, u, q, box, Kx, Ky, Kz, beta, frc.data(), nullptr, R.data());
    std::vector<int> base;
    std::vector<double> frac;
    pme_frac_base(n, u, K, base, frac);
    std::vector<double> wv(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) wv[3 * (size_t)i + d] = w[3 * (size_t)i + d] * K[d] / (double)box[d];
    // D grid: sum_i q_i w_i.grad M_i
    std::vector<cplx> a(NG, cplx(0, 0));
    for (int i = 0; i < n; i++) {
        const double* f = &frac[3 * (size_t)i];
        const int* b0 = &base[3 * (size_t)i];
        const double* v = &wv[3 * (size_t)i];
        double wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
        for (int k = 0; k < 4; k++) {
            wx[k] = bspline_M(f[0] + k, 4);
            wy[k] = bspline_M(f[1] + k, 4);
            wz[k] = bspline_M(f[2] + k, 4);
            dx[k] = bspline_M(f[0] + k, 3) - bspline_M(f[0] + k - 1, 3);
            dy[k] = bspline_M(f[1] + k, 3) - bspline_M(f[1] + k - 1, 3);
            dz[k] = bspline_M(f[2] + k, 3) - bspline_M(f[2] + k - 1, 3);
        }
        for (int aa = 0; aa < 4; aa++) {
            int gx = ((b0[0] - aa) % Kx + Kx) % Kx;
            for (int bb = 0; bb < 4; bb++) {
                int gy = ((b0[1] - bb) % Ky + Ky) % Ky;
                for (int cc = 0; cc < 4; cc++) {
                    int gz = ((b0[2] - cc) % Kz + Kz) % Kz;
                    double m = v[0] * dx[aa] * wy[bb] * wz[cc] + v[1] * wx[aa] * dy[bb] * wz[cc] +
                               v[2] * wx[aa] * wy[bb] * dz[cc];
                    a[((size_t)gx * Ky + gy) * Kz + gz] += (double)q[i] * m;
                }
            }
        }
    }
    std::vector<double> bc(NG);
    orc_pme_bc_full(Kx, Ky, Kz, box, beta, bc.data());
    fft3d(a, Kx, Ky, Kz, -1);
    for (size_t k = 0; k < NG; k++) a[k] *= bc[k];
    fft3d(a, Kx, Ky, Kz, +1);
    std::vector<double> RD(NG);
    for (size_t k = 0; k < NG; k++) RD[k] = a[k].real();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        gq[i] += -pme_atom_gather(R.data(), &base[3 * (size_t)i], &frac[3 * (size_t)i], K, &wv[3 * (size_t)i], 1) -
                 pme_atom_gather(RD.data(), &base[3 * (size_t)i], &frac[3 * (size_t)i], K, nullptr, 0);
    return L;
}

// Self and neutralising terms (PME_force.cu:757-762, neutralizing_factor :166-167):
// E_self = -b/sqrt(pi) sum q^2, E_neutral = -pi/(2 b^2 V) (sum q)^2.
void orc_pme_self(int n, const float* q, double beta, const float* box, double* out2) {
    double s2 = 0, s1 = 0;
    for (int i = 0; i < n; i++) { s2 += (double)q[i] * q[i]; s1 += q[i]; }
    double V = (double)box[0] * box[1] * box[2];
    out2[0] = -beta / std::sqrt(PI_D) * s2;
    out2[1] = -PI_D / (2.0 * beta * beta * V) * s1 * s1;
}

// a11/a14 long-range dispersion factor (Lennard_Jones_force.cu:1042-1065,
// :1127): factor = -(2/3) pi sum_i sum_j (B_ij/6) / rc^3 over all ordered
// pairs including i=j; E_lrc = factor/V, W_lrc = 6 factor/V (:1178-1191).
double orc_lj_lrc_factor(int n, const int* type, int ntypes, const float* B, double rc) {
    std::vector<double> count(ntypes, 0.0);
    for (int i = 0; i < n; i++) count[type[i]] += 1;
    double s = 0;
    for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++) s += count[a] * count[b] * (double)B[pair_index(a, b)];
    return s * (-2.0 / 3.0) * PI_D / (rc * rc * rc) / 6.0;
}

// Exact Ewald sum (FP64) for tiny periodic systems of point charges with no
// exclusions, used only to pin the SPME restatement: real space over all
// images within rreal with erfc screening, reciprocal space over all |k| <= kmax with the same 1/(pi V m^2) convention,
// plus self and neutralising terms. Returns the total; frc += forces.
double orc_ewald_exact(int n, const double* xyz, const double* q, const float* box, double beta,
                       double rreal, int kmax, double* frc) {
    double L[3] = {box[0], box[1], box[2]};
    double V = L[0] * L[1] * L[2];
    double E = 0;
    int ni[3];
    for (int d = 0; d < 3; d++) ni[d] = (int)std::ceil(rreal / L[d]) + 1;
    const double tsp = 2.0 / std::sqrt(PI_D);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int a = -ni[0]; a <= ni[0]; a++)
                for (int b = -ni[1]; b <= ni[1]; b++)
                    for (int c = -ni[2]; c <= ni[2]; c++) {
                        if (i == j && a == 0 && b == 0 && c == 0) continue;
                        double dr[3] = {xyz[3 * j] - xyz[3 * i] + a * L[0], xyz[3 * j + 1] - xyz[3 * i + 1] + b * L[1],
                                        xyz[3 * j + 2] - xyz[3 * i + 2] + c * L[2]};
                        double r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
                        if (r > rreal) continue;
                        double qq = q[i] * q[j];
                        E += 0.5 * qq * std::erfc(beta * r) / r;
                        double f = -qq * (tsp * beta * r * std::exp(-beta * beta * r * r) + std::erfc(beta * r)) / (r * r * r);
                        for (int d = 0; d < 3; d++) frc[3 * i + d] += f * dr[d];
                    }
    for (int mx = -kmax; mx <= kmax; mx++)
        for (int my = -kmax; my <= kmax; my++)
            for (int mz = -kmax; mz <= kmax; mz++) {
                if (mx == 0 && my == 0 && mz == 0) continue;
                if (mx * mx + my * my + mz * mz > kmax * kmax) continue;
                double m[3] = {mx / L[0], my / L[1], mz / L[2]};
                double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                double pref = std::exp(-PI_D * PI_D * m2 / (beta * beta)) / (PI_D * V * m2);
                double sr = 0, si = 0;
                for (int i = 0; i < n; i++) {
                    double ph = 2 * PI_D * (m[0] * xyz[3 * i] + m[1] * xyz[3 * i + 1] + m[2] * xyz[3 * i + 2]);
                    sr += q[i] * std::cos(ph);
                    si += q[i] * std::sin(ph);
                }
                E += 0.5 * pref * (sr * sr + si * si);
                for (int i = 0; i < n; i++) {
                    double ph = 2 * PI_D * (m[0] * xyz[3 * i] + m[1] * xyz[3 * i + 1] + m[2] * xyz[3 * i + 2]);
                    // dE/dr_i = pref * q_i * 2 pi m (-sin(ph) sr + cos(ph) si)
                    double g = pref * q[i] * 2 * PI_D * (-std::sin(ph) * sr + std::cos(ph) * si);
                    for (int d = 0; d < 3; d++) frc[3 * i + d] -= g * m[d];
                }
            }
    double s2 = 0, s1 = 0;
    for (int i = 0; i < n; i++) { s2 += q[i] * q[i]; s1 += q[i]; }
    E += -beta / std::sqrt(PI_D) * s2 - PI_D / (2 * beta * beta * V) * s1 * s1;
    return E;
}

// ---------------------------------------------------------------------------
// Integrator and SETTLE (next to the force path).

// Get_Rabcde_From_SSS (constrain/SETTLE.cu:8-47) in FP64 (settle_math.hpp).
void orc_settle_geometry(double rAB, double rAC, double rBC, double mA, double mB, double mC, double* g) {
    orc::settle_geometry<double>(rAB, rAC, rBC, mA, mB, mC, g);
}

// SETTLE::Initial group discovery (constrain/SETTLE.cu:209-331, pairs' k at
// :325, leftover count :339). counts = {triangles, pairs, leftover pairs};
// tri = (A, B, C) per triangle with geo = its 5 canonical coordinates;
// prs = (i, j) per pair with plen = its length.
void orc_settle_groups(int n, int np, const int* ai, const int* aj, const float* r, const float* mass, int* tri,
                       double* geo, int* prs, double* plen, int* counts) {
    std::vector<int> ln(n, 0), la(2 * (size_t)n, -1);
    std::vector<double> lr(3 * (size_t)n, 0.0);
    for (int k = 0; k < np; k++) {
        int i = ai[k], j = aj[k];
        if (ln[i] < 2 && ln[j] < 2) {
            la[2 * i + ln[i]] = j; la[2 * j + ln[j]] = i;
            lr[3 * i + ln[i]] = r[k]; lr[3 * j + ln[j]] = r[k];
            ln[i]++; ln[j]++;
        } else {
            ln[i] = 3; ln[j] = 3;
        }
    }
    for (int i = 0; i < n; i++) {
        if (ln[i] == 2) {
            int a = la[2 * i], b = la[2 * i + 1];
            if (ln[a] == 2 && ln[b] == 2 &&
                ((la[2 * a] == i && la[2 * a + 1] == b) || (la[2 * a + 1] == i && la[2 * a] == b))) {
                ln[a] = -2; ln[b] = -2;
                lr[3 * i + 2] = (la[2 * a + 1] == b) ? lr[3 * a + 1] : lr[3 * a];
            } else {
                ln[i] = 3; ln[a] = 3; ln[b] = 3;
            }
        } else if (ln[i] == 1) {
            int a = la[2 * i];
            if (ln[a] == 1) ln[a] = -1;
            else { ln[i] = 3; ln[a] = 3; }
        }
    }
    int nt = 0, npr = 0, rest = 0;
    for (int i = 0; i < n; i++) {
        if (ln[i] == 2) {
            ln[i] = -2;
            tri[3 * nt] = i; tri[3 * nt + 1] = la[2 * i]; tri[3 * nt + 2] = la[2 * i + 1];
            orc_settle_geometry(lr[3 * i], lr[3 * i + 1], lr[3 * i + 2], mass[i], mass[la[2 * i]],
                                mass[la[2 * i + 1]], geo + 5 * nt);
            nt++;
        } else if (ln[i] == 1) {
            ln[i] = -1;
            prs[2 * npr] = i; prs[2 * npr + 1] = la[2 * i]; plen[npr] = lr[3 * i];
            npr++;
        }
    }
    for (int k = 0; k < np; k++) if (ln[ai[k]] > 0) rest++;
    counts[0] = nt; counts[1] = npr; counts[2] = rest;
}

using orc::settle_do_triangle;
using orc::v_dot;
using orc::v_pbc;

// One NVE step of main.cu:282-320 in FP64: SETTLE::Remember_Last_Coordinates
// from the step's uint coordinates (SETTLE.cu:385-410, scaler L / 2^32),
// MD_Iteration_Leap_Frog(_With_Max_Velocity) (MD_core.cu:13-45) on every
// atom, then settle_pair / settle_triangle (SETTLE.cu:412-483) with
// Get_Periodic_Displacement of the float coordinates. crd / vel updated in
// place; returns the constraint virial sum (before the 1/(3V) factor).
double orc_md_iteration(int n, double* crd, double* vel, const double* frc, const double* inv_mass,
                        const double* mass, const uint32_t* u, const float* box, double dt, double vmax, double vfac,
                        double xfac, int ntri, const int* tri, const double* geo, int npair, const int* prs,
                        const double* plen) {
    double sc[3];
    for (int d = 0; d < 3; d++) sc[d] = (double)box[d] / 4294967296.0;
    auto udisp = [&](int a, int b, double* o) {
        for (int d = 0; d < 3; d++) o[d] = (double)(int32_t)(u[3 * a + d] - u[3 * b + d]) * sc[d];
    };
    std::vector<double> last(6 * (size_t)ntri + 3 * (size_t)npair);
    for (int k = 0; k < ntri; k++) {
        udisp(tri[3 * k + 1], tri[3 * k], &last[6 * k]);
        udisp(tri[3 * k + 2], tri[3 * k], &last[6 * k + 3]);
    }
    for (int k = 0; k < npair; k++) udisp(prs[2 * k + 1], prs[2 * k], &last[6 * (size_t)ntri + 3 * k]);
    for (int i = 0; i < n; i++) {
        double v[3];
        for (int d = 0; d < 3; d++) v[d] = vel[3 * i + d] + dt * inv_mass[i] * frc[3 * i + d];
        if (vmax > 0) {
            double s = std::min(1.0, vmax / std::sqrt(v_dot(v, v)));
            for (int d = 0; d < 3; d++) v[d] *= s;
        }
        for (int d = 0; d < 3; d++) { vel[3 * i + d] = v[d]; crd[3 * i + d] += dt * v[d]; }
    }
    double vir = 0;
    for (int k = 0; k < npair; k++) {
        int i = prs[2 * k], j = prs[2 * k + 1];
        double r1[3], *r2 = &last[6 * (size_t)ntri + 3 * k];
        v_pbc(&crd[3 * j], &crd[3 * i], box, r1);
        double mA = mass[i], mB = mass[j], r0r0 = plen[k] * plen[k];
        double r1r1 = v_dot(r1, r1), r1r2 = v_dot(r1, r2), r2r2 = v_dot(r2, r2);
        double kk = (std::sqrt(r1r2 * r1r2 - r1r1 * r2r2 + r2r2 * r0r0) - r1r2) / r2r2;
        double ck = 1.0 / (mA + mB);
        for (int d = 0; d < 3; d++) {
            double di = -mB * ck * kk * r2[d], dj = mA * ck * kk * r2[d];
            crd[3 * i + d] += di; crd[3 * j + d] += dj;
            vel[3 * i + d] += vfac / dt / xfac * di; vel[3 * j + d] += vfac / dt / xfac * dj;
            vir += mB / dt / dt / xfac * dj * r2[d];
        }
    }
    for (int k = 0; k < ntri; k++) {
        int a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
        double rA[3], rB[3], rC[3], rO[3];
        for (int d = 0; d < 3; d++) rA[d] = crd[3 * a + d];
        v_pbc(&crd[3 * b], rA, box, rB);
        v_pbc(&crd[3 * c], rA, box, rC);
        double mA = mass[a], mB = mass[b], mC = mass[c];
        for (int d = 0; d < 3; d++) {
            rO[d] = (mB * rB[d] + mC * rC[d]) / (mA + mB + mC) + rA[d];
            rA[d] -= rO[d];
            rB[d] += rA[d];
            rC[d] += rA[d];
        }
        vir += settle_do_triangle<double>(&last[6 * k], &last[6 * k + 3], rA, rB, rC, geo + 5 * k, mB, mC, dt, vfac, xfac,
                                  &vel[3 * a], &vel[3 * b], &vel[3 * c]);
        for (int d = 0; d < 3; d++) {
            crd[3 * a + d] = rA[d] + rO[d];
            crd[3 * b + d] = rB[d] + rO[d];
            crd[3 * c + d] = rC[d] + rO[d];
        }
    }
    return vir;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
