// Batched systems: one launch per stage of the nonbonded step over many
// independent systems (FEP lambda windows, replicas of uneven size).
//
// Each batched kernel reads a block map entry {system, block within the
// system} and the system's BatchSys (its module structs by value plus the
// caller's arrays) and runs the same device body as the single-system
// kernel. Maps are built once here; grids are sized from capacities, so the
// launches are static and the whole step can be captured into one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace snb;

static int stage_blocks(int st, const sponge_nl* nl, const sponge_pme* pme) {
    const int n = pme->n, H = pme->K[2] / 2 + 1;
    switch (st) {
        case kStPrep: return ceil_div(nl->slot_cap, 256);
        case kStPair: return ceil_div(nl->ncl_cap, 4);
        case kStExcl: return ceil_div(n, 256);
        case kStFwd: return pme->K[0] * snb_pme_row_blocks(pme);
        case kStConv: return H;
        case kStInv: return pme->K[0] * snb_pme_row_blocks(pme);
        case kStGather: return ceil_div(16 * n, 256);
        case kStFin: return 1;
        case kStNlCheck: return ceil_div(n, 256);
        case kStNlBegin: return 1;
        case kStNlBin: return ceil_div(n, 256);
        case kStNlScan: return 1;
        case kStNlSort: return nl->ncol;
        default: return ceil_div(nl->ncl_cap, 4);  // kStNlPairs
    }
}

static void batch_free(sponge_batch* b) {
    if (!b) return;
    if (b->d) (void)hipFree(b->d);
    for (int st = 0; st < kNumStages; st++)
        if (b->map[st]) (void)hipFree(b->map[st]);
    if (b->d_md) (void)hipFree(b->d_md);
    if (b->md_map) (void)hipFree(b->md_map);
    if (b->d_ident) (void)hipFree(b->d_ident);
    if (b->side) {
        (void)hipStreamSynchronize(b->side);
        (void)hipStreamDestroy(b->side);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    delete b;
}

extern "C" int sponge_batch_sync(sponge_batch* b) {
    SNB_ARG(b, "batch_sync");
    for (int s = 0; s < b->nsys; s++) {
        BatchSys& v = b->h[s];
        const sponge_nl* nl = b->nl[s];
        const sponge_lj* lj = b->lj[s];
        const sponge_pme* pme = b->pme[s];
        SNB_ARG(nl->box[0] == pme->box[0] && nl->box[1] == pme->box[1] && nl->box[2] == pme->box[2] &&
                    lj->box[0] == pme->box[0] && lj->box[1] == pme->box[1] && lj->box[2] == pme->box[2],
                "the modules of a system disagree on the box");
        v.nl = *nl;
        v.lj = *lj;
        v.pme = *pme;
        v.rlm = snb_nl_rlm(nl);
        v.thresh = nl->skin_permit * nl->skin_permit * nl->half_skin_square;
        for (int d = 0; d < 3; d++) v.cq[d] = 0.25f * (4294967296.0f / pme->box[d]);
        v.lrc = lj->lrc_factor / lj->volume;
    }
    SNB_HIP(hipMemcpy(b->d, b->h.data(), b->nsys * sizeof(BatchSys), hipMemcpyHostToDevice));
    return b->d_md ? snb_batch_md_sync(b) : SPONGE_NB_OK;
}

extern "C" int sponge_batch_create(int nsys, sponge_nl* const* nl, sponge_lj* const* lj, sponge_pme* const* pme,
                                   const sponge_batch_io* io, sponge_batch** out) {
    SNB_ARG(out && nl && lj && pme && io && nsys > 0, "batch_create");
    *out = nullptr;
    for (int s = 0; s < nsys; s++) {
        SNB_ARG(nl[s] && lj[s] && pme[s], "null module handle");
        SNB_ARG(nl[s]->n == lj[s]->n && nl[s]->n == pme[s]->n, "atom counts of a system's modules differ");
        SNB_ARG(nl[s]->refresh_interval == nl[0]->refresh_interval, "refresh_interval differs between systems");
        const sponge_batch_io& o = io[s];
        SNB_ARG(o.crd && o.uint_crd && o.charge && o.excluded_list_start && o.excluded_list && o.excluded_numbers &&
                    o.frc,
                "null system array");
    }
    sponge_batch* b = new sponge_batch();
    b->nsys = nsys;
    b->nl.assign(nl, nl + nsys);
    b->lj.assign(lj, lj + nsys);
    b->pme.assign(pme, pme + nsys);
    b->h.resize(nsys);
    b->d = nullptr;
    for (int st = 0; st < kNumStages; st++) b->map[st] = nullptr;
    b->refresh_interval = nl[0]->refresh_interval;
    b->refresh_count = 0;
    b->max_nt = 0;
    b->soft = 0;
    b->lds_fwd = b->lds_conv = b->lds_inv = 0;
    int rc = SPONGE_NB_OK;
    for (int s = 0; s < nsys && rc == SPONGE_NB_OK; s++) {
        rc = snb_lj_scratch(lj[s], nl[s]);
        b->h[s].io = io[s];
        snb_pme_plans(pme[s], &b->h[s].px, &b->h[s].py, &b->h[s].pz);
        b->max_nt = std::max(b->max_nt, lj[s]->ntypes);
        b->soft |= lj[s]->d_sc != nullptr;
        b->lds_fwd = std::max(b->lds_fwd, snb_pme_lds_fwd(pme[s]));
        b->lds_conv = std::max(b->lds_conv, snb_pme_lds_conv(pme[s]));
        b->lds_inv = std::max(b->lds_inv, snb_pme_lds_inv(pme[s]));
    }
    b->lds_pair = (size_t)b->max_nt * b->max_nt * sizeof(float2);
    if (rc == SPONGE_NB_OK) rc = snb_pme_raise_lds(std::max({b->lds_fwd, b->lds_conv, b->lds_inv}));
    if (rc == SPONGE_NB_OK && hipMalloc((void**)&b->d, nsys * sizeof(BatchSys)) != hipSuccess) {
        set_error("batch_create: out of device memory");
        rc = SPONGE_NB_EHIP;
    }
    // default-priority side stream and the fork / join events of the force step
    if (rc == SPONGE_NB_OK && (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
                               hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
                               hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)) {
        set_error("batch_create: cannot create the side stream or its events");
        rc = SPONGE_NB_EHIP;
    }
    for (int st = 0; st < kNumStages && rc == SPONGE_NB_OK; st++) {
        std::vector<int2> m;
        for (int s = 0; s < nsys; s++) {
            const int k = stage_blocks(st, nl[s], pme[s]);
            for (int j = 0; j < k; j++) m.push_back(make_int2(s, j));
        }
        b->nblk[st] = (int)m.size();
        if (hipMalloc((void**)&b->map[st], m.size() * sizeof(int2)) != hipSuccess ||
            hipMemcpy(b->map[st], m.data(), m.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("batch_create: block map upload failed");
            rc = SPONGE_NB_EHIP;
        }
    }
    if (rc == SPONGE_NB_OK) rc = sponge_batch_sync(b);
    if (rc != SPONGE_NB_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return SPONGE_NB_OK;
}

extern "C" void sponge_batch_destroy(sponge_batch* b) { batch_free(b); }

extern "C" int sponge_batch_neighbor_list_update(sponge_batch* b, int forced_update, void* stream) {
    SNB_ARG(b, "batch_neighbor_list_update");
    hipStream_t s = as_stream(stream);
    if (forced_update) return snb_batch_nl(b, 1, s);
    if (b->refresh_interval > 0) {
        int rc = SPONGE_NB_OK;
        if (b->refresh_count % b->refresh_interval == 0) rc = snb_batch_nl(b, 1, s);
        b->refresh_count += 1;
        return rc;
    }
    return snb_batch_nl(b, 0, s);
}

extern "C" int sponge_batch_force_step(sponge_batch* b, int flags, int need_energy, int need_virial, void* stream) {
    SNB_ARG(b, "batch_force_step");
    hipStream_t s = as_stream(stream);
    const int energy = (need_energy > 0 || need_virial > 0) ? 1 : 0;
    int rc = snb_batch_prep(b, energy, (flags & SPONGE_STEP_FROM_CRD) ? 1 : 0,
                            (flags & SPONGE_STEP_OVERWRITE_FORCE) ? 1 : 0, s);
    if (rc) return rc;
    // Fork: the pair and exclusion stages (the longer branch: it stays on
    // the caller's queue, captured first) and the reciprocal stages on the
    // side stream, both after the prep; join before the gather, which adds
    // into the forces. SPONGE_STEP_SERIAL keeps every stage on `stream`.
    const bool serial = (flags & SPONGE_STEP_SERIAL) != 0;
    hipStream_t ps = serial ? s : b->side;
    if (!serial) {
        SNB_HIP(hipEventRecord(b->fork, s));
        SNB_HIP(hipStreamWaitEvent(ps, b->fork, 0));
    }
    rc = snb_batch_pairs(b, energy, (flags & SPONGE_STEP_OVERWRITE_FORCE) ? 1 : 0, s);
    if (!rc) rc = snb_batch_excluded(b, s);
    if (!rc) rc = snb_batch_reciprocal(b, energy, ps);
    if (rc) return rc;
    if (!serial) {
        SNB_HIP(hipEventRecord(b->join, ps));
        SNB_HIP(hipStreamWaitEvent(s, b->join, 0));
    }
    rc = snb_batch_gather(b, s);
    if (!rc && energy) rc = snb_batch_finalize(b, s);
    return rc;
}
