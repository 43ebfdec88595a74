// C ABI glue: error reporting, reductions, the fused per-step force path and
// the energy queries of Particle_Mesh_Ewald::Get_Energy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "tune_table.hpp"

using namespace snb;

namespace snb {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace snb

namespace {

__global__ void fill_zero_f64_kernel(double* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void sum_list_kernel(int n, const float* __restrict__ in, float* __restrict__ out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += in[i];
    double t = block_sum<1024>(s, red);
    if (threadIdx.x == 0) out[0] += (float)t;
}

__global__ void parts_kernel(const double* __restrict__ lj_acc, const double* __restrict__ pme_acc, double lrc,
                             double* __restrict__ parts) {
    if (threadIdx.x != 0) return;
    parts[SPONGE_PART_LJ] = lj_acc[SPONGE_PART_LJ];
    parts[SPONGE_PART_DIRECT] = lj_acc[SPONGE_PART_DIRECT];
    parts[SPONGE_PART_LJ_VIRIAL] = lj_acc[SPONGE_PART_LJ_VIRIAL];
    parts[SPONGE_PART_CORRECTION] = pme_acc[SPONGE_PART_CORRECTION];
    parts[SPONGE_PART_SELF] = pme_acc[SPONGE_PART_SELF];
    parts[SPONGE_PART_RECIPROCAL] = pme_acc[SPONGE_PART_RECIPROCAL];
    parts[SPONGE_PART_LRC] = lrc;
    parts[SPONGE_PART_PAIRS_IN] = lj_acc[SPONGE_PART_PAIRS_IN];
}

}  // namespace

int snb_fill_zero_f64(double* p, int n, hipStream_t s) {
    fill_zero_f64_kernel<<<ceil_div(n, 256), 256, 0, s>>>(p, n);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" const char* sponge_nb_last_error(void) { return snb::g_err; }

extern "C" int sponge_nb_device_info(int* count, char* name, int name_len) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c == 0) {
        if (count) *count = 0;
        set_error("no HIP device: %s", hipGetErrorString(e));
        return SPONGE_NB_ENODEV;
    }
    if (count) *count = c;
    int dev = 0;
    SNB_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SNB_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) snprintf(name, name_len, "%s", prop.gcnArchName);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %s is not gfx950", prop.gcnArchName);
        return SPONGE_NB_ENODEV;
    }
    return SPONGE_NB_OK;
}

extern "C" int sponge_sum_list(int n, const float* d_list, float* d_out, void* stream) {
    SNB_ARG(n >= 0 && d_list && d_out, "sum_list");
    sum_list_kernel<<<1, 1024, 0, as_stream(stream)>>>(n, d_list, d_out);
    SNB_CHECK_LAUNCH();
    return SPONGE_NB_OK;
}

extern "C" int sponge_nb_force_step(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const float* d_crd,
                                    unsigned int* d_uint_crd, const float* d_charge, const int* d_excluded_list_start,
                                    const int* d_excluded_list, const int* d_excluded_numbers, float* d_frc, int flags,
                                    int need_energy, int need_virial, float* d_atom_energy, float* d_atom_virial,
                                    double* d_parts, void* stream) {
    SNB_ARG(nl && lj && pme && d_uint_crd && d_charge && d_frc, "nb_force_step");
    SNB_ARG(d_excluded_list_start && d_excluded_list && d_excluded_numbers, "exclusion CSR");
    SNB_ARG(nl->n == lj->n && nl->n == pme->n, "atom counts of the modules differ");
    SNB_ARG(pme->box[0] == lj->box[0] && pme->box[1] == lj->box[1] && pme->box[2] == lj->box[2],
            "LJ and PME boxes differ");
    hipStream_t s = as_stream(stream);
    const int energy = (need_energy > 0 || need_virial > 0) ? 1 : 0;
    const bool overwrite = (flags & SPONGE_STEP_OVERWRITE_FORCE) != 0;
    int rc = snb_step_prep(nl, lj, pme, d_crd, d_uint_crd, d_charge, energy, overwrite ? d_frc : nullptr, s);
    if (rc) return rc;
    // Force-only steps fold the excluded-pair correction into the pair kernel
    // when the neighbor list holds the symmetric form of this exclusion CSR;
    // energy steps keep the separate kernel, which also fills the per-atom
    // correction energies.
    const bool fold_excl = !energy && nl->sym_start && nl->ex_es == d_excluded_list_start &&
                           nl->ex_el == d_excluded_list && nl->ex_en == d_excluded_numbers;
    const ExclArgs ex = fold_excl ? ExclArgs{nl->sym_start, nl->sym_list, nl->o2s} : ExclArgs{nullptr, nullptr, nullptr};
    // launch form: forced by the flags, else the PME handle's (tuned) form
    const bool serial = (flags & SPONGE_STEP_SERIAL) ||
                        (!(flags & SPONGE_STEP_TWO_STREAMS) && pme->step_form == 1);
    int* step_count = nl->prune_interval > 0 ? nl->flags + 10 : nullptr;
    // Fork: the reciprocal pipeline (spread, FFTs, convolution) stays on s
    // while the pair and exclusion kernels run on a stream of the PME handle;
    // both read only what step_prep wrote. Join before the gather, which
    // adds into d_frc. The pipeline is the longer branch: eager, it starts
    // without a cross-stream wait; captured, it is the first branch after the
    // fork, which a graph keeps on the queue of the node before it (the other
    // branch gets a second queue, whose start costs 10-40 us).
    hipStream_t ps = serial ? s : pme->side;
    if (!serial) SNB_HIP(hipEventRecord(pme->fork, s));
    rc = snb_pme_reciprocal(pme, d_uint_crd, d_charge, nullptr, energy, 1, s);
    if (rc) return rc;
    if (!serial) SNB_HIP(hipStreamWaitEvent(ps, pme->fork, 0));
    rc = snb_pairs(nl, lj->d_ab, lj->ntypes, lj->d_type, lj->cutoff, lj->cof_dr, d_uint_crd, d_charge, pme->beta,
                   energy, &lj->ps, d_frc, need_energy > 0 ? d_atom_energy : nullptr,
                   need_virial > 0 ? d_atom_virial : nullptr, nullptr, ps, 1,
                   overwrite ? 1 : 0, lj->d_sc, 1, ex);
    if (rc) return rc;
    if (!fold_excl) {
        rc = snb_excluded(pme, d_uint_crd, lj->cof_dr, d_charge, d_excluded_list_start, d_excluded_list,
                          d_excluded_numbers, d_frc, nullptr, ps);
        if (rc) return rc;
    }
    if (!serial) {
        SNB_HIP(hipEventRecord(pme->join, ps));
        SNB_HIP(hipStreamWaitEvent(s, pme->join, 0));
    }
    rc = snb_pme_gather(pme, d_frc, s, step_count);
    if (rc) return rc;
    if (energy) {
        rc = snb_pme_finalize(pme, nullptr, 0, 0, nullptr, nullptr, s);
        if (rc) return rc;
        if (d_parts) {
            parts_kernel<<<1, 64, 0, s>>>(lj->ps.acc, pme->acc, lj->lrc_factor / lj->volume, d_parts);
            SNB_CHECK_LAUNCH();
        }
    }
    return SPONGE_NB_OK;
}

extern "C" int sponge_pme_get_energy(sponge_pme* pme, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                     const float* d_charge, const float scaler[3], const int* d_excluded_list_start,
                                     const int* d_excluded_list, const int* d_excluded_atom_numbers, int which_part,
                                     double* h_energy, void* stream) {
    SNB_ARG(pme && d_uint_crd && d_charge && h_energy, "pme_get_energy");
    SNB_ARG(which_part >= 0 && which_part <= 4, "which_part must be 0..4 (PME_force.cuh:86-93)");
    hipStream_t s = as_stream(stream);
    const bool all = which_part == 0;
    double e_rec = 0, e_self = 0, e_dir = 0, e_corr = 0;
    if (all || which_part == 2 || which_part == 4) {
        int rc = snb_pme_reciprocal(pme, d_uint_crd, d_charge, nullptr, 1, 0, s);
        if (rc) return rc;
    }
    if (all || which_part == 3) {
        SNB_ARG(scaler && d_excluded_list_start && d_excluded_list && d_excluded_atom_numbers, "exclusion CSR");
        int rc = snb_excluded(pme, d_uint_crd, scaler, d_charge, d_excluded_list_start, d_excluded_list,
                              d_excluded_atom_numbers, nullptr, nullptr, s);
        if (rc) return rc;
    }
    if (all || which_part == 2 || which_part == 3 || which_part == 4) {
        int rc = snb_pme_finalize(pme, nullptr, 0, 0, nullptr, nullptr, s);
        if (rc) return rc;
        double acc[SPONGE_NB_NPARTS];
        SNB_HIP(hipMemcpyAsync(acc, pme->acc, sizeof(acc), hipMemcpyDeviceToHost, s));
        SNB_HIP(hipStreamSynchronize(s));
        e_rec = acc[SPONGE_PART_RECIPROCAL];
        e_self = acc[SPONGE_PART_SELF];
        e_corr = acc[SPONGE_PART_CORRECTION];
    }
    if (all || which_part == 1) {
        SNB_ARG(nl && nl->n == pme->n, "neighbor list");
        if (!pme->zero_type) {
            SNB_HIP(hipMalloc((void**)&pme->zero_type, pme->n * sizeof(int)));
            SNB_HIP(hipMemset(pme->zero_type, 0, pme->n * sizeof(int)));
            SNB_HIP(hipMalloc((void**)&pme->zero_ab, sizeof(float2)));
            SNB_HIP(hipMemset(pme->zero_ab, 0, sizeof(float2)));
        }
        float cof[3];
        for (int d = 0; d < 3; d++) cof[d] = (1.0f / 4294967296.0f) * pme->box[d];
        int rc = snb_pairs(nl, pme->zero_ab, 1, pme->zero_type, pme->cutoff, cof, d_uint_crd, d_charge, pme->beta, 1,
                           &pme->ps, nullptr, nullptr, nullptr, nullptr, s);
        if (rc) return rc;
        double acc[SPONGE_NB_NPARTS];
        SNB_HIP(hipMemcpyAsync(acc, pme->ps.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
        SNB_HIP(hipStreamSynchronize(s));
        e_dir = acc[SPONGE_PART_DIRECT];
    }
    switch (which_part) {
        case 0: *h_energy = e_rec + e_self + e_dir + e_corr; break;
        case 1: *h_energy = e_dir; break;
        case 2: *h_energy = e_rec; break;
        case 3: *h_energy = e_corr; break;
        default: *h_energy = e_self; break;
    }
    return SPONGE_NB_OK;
}

// ---------------------------------------------------------------------------
// Launch shapes (tune_table.hpp holds the measured defaults).

extern "C" int sponge_nb_default_shapes(int atom_numbers, int refresh_interval, float cutoff, float skin,
                                        sponge_shapes* out) {
    SNB_ARG(out && atom_numbers > 0 && cutoff > 0 && skin >= 0, "nb_default_shapes");
    const TuneRow& t = tune_row(atom_numbers);
    out->pair_waves = t.pair_waves;
    out->conv_split = t.conv_split < 0 ? -1 : t.conv_split;
    out->conv_threads = t.conv_threads;
    out->step_form = t.step_form;
    out->prune_interval = (refresh_interval > 0 && t.prune_div > 0) ? std::max(1, refresh_interval / t.prune_div) : 0;
    out->prune_radius = cutoff + t.prune_margin * skin;
    return SPONGE_NB_OK;
}

extern "C" int sponge_nb_get_shapes(const sponge_nl* nl, const sponge_lj* lj, const sponge_pme* pme,
                                    sponge_shapes* out) {
    SNB_ARG(nl && lj && pme && out, "nb_get_shapes");
    out->pair_waves = (lj->ps.waves == 2 || lj->ps.waves == 8) ? lj->ps.waves : 4;  // else the 4-wave block
    out->conv_split = pme->conv_split;
    out->conv_threads = pme->conv_threads > 0 ? pme->conv_threads : 256;
    out->step_form = pme->step_form;
    out->prune_interval = nl->prune_interval;
    out->prune_radius = nl->prune_interval > 0 ? std::sqrt(nl->prune_r2) : nl->cutoff;
    return SPONGE_NB_OK;
}

extern "C" int sponge_nb_set_shapes(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const sponge_shapes* sh) {
    SNB_ARG(nl && lj && pme && sh, "nb_set_shapes");
    SNB_ARG(sh->pair_waves == 2 || sh->pair_waves == 4 || sh->pair_waves == 8, "pair_waves must be 2, 4 or 8");
    SNB_ARG(sh->conv_threads == 256 || sh->conv_threads == 512, "conv_threads must be 256 or 512");
    SNB_ARG(sh->step_form == 0 || sh->step_form == 1, "step_form must be 0 or 1");
    SNB_ARG(sh->conv_split >= 0 && sh->conv_split <= 2, "conv_split must be 0, 1 or 2");
    SNB_ARG(sh->prune_interval >= 0, "prune_interval");
    SNB_ARG(sh->prune_interval == 0 || sh->prune_radius >= nl->cutoff, "prune radius below the cutoff");
    // the one check that needs the grid: the plane form must fit the LDS
    int rc = sponge_pme_set_tuning(pme, -1, sh->conv_split);
    if (rc) return rc;
    pme->step_form = sh->step_form;
    pme->conv_threads = sh->conv_threads;
    lj->ps.waves = sh->pair_waves;
    if (sh->prune_interval != nl->prune_interval ||
        (sh->prune_interval > 0 && sh->prune_radius * sh->prune_radius != nl->prune_r2))
        return sponge_nl_set_prune(nl, sh->prune_interval, sh->prune_radius);
    return SPONGE_NB_OK;
}
