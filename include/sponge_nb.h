/*
 * sponge_nb.h -- C ABI of the MI355X (gfx950) nonbonded hot path of SPONGE.
 *
 * Three module handles mirror the reference's three C++ modules; every entry
 * point below replaces one reference method (cited as file:line relative to
 * the reference's SPONGE/ directory). All pointers named d_* are device
 * pointers (HIP), laid out exactly as the reference lays them out
 * (VECTOR = 3 floats AoS, UNSIGNED_INT_VECTOR = 3 uint32 AoS, original atom
 * order), so the reference's MD_INFORMATION buffers can be passed unchanged.
 * Outputs follow the reference's accumulate-into (+=) semantics.
 *
 * `stream` is a hipStream_t (NULL = the default stream). No entry point that
 * takes a stream allocates, frees or synchronises in steady state (scratch is
 * sized on the first call, and sponge_nl_update reads a new exclusion CSR
 * once), so a whole MD step can be captured into one hipGraph after one eager
 * step; functions that return host values synchronise and say so.
 *
 * Return value: 0 on success, a negative SPONGE_NB_E* code on failure;
 * sponge_nb_last_error() describes the last failure on the calling thread.
 */
#ifndef SPONGE_NB_H
#define SPONGE_NB_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPONGE_NB_OK 0
#define SPONGE_NB_EARG (-1)      /* invalid argument / shape */
#define SPONGE_NB_EHIP (-2)      /* HIP runtime error */
#define SPONGE_NB_ECAPACITY (-3) /* a fixed capacity was exceeded (see overflow flags) */
#define SPONGE_NB_ENODEV (-4)    /* no usable gfx950 device */
#define SPONGE_NB_EINTERNAL (-5) /* an internal consistency check failed */

/* Overflow flags reported by sponge_nl_status (bit set = overflowed). */
#define SPONGE_NL_OVERFLOW_COLUMN 1  /* too many atoms in one cell column */
#define SPONGE_NL_OVERFLOW_LIST 2    /* too many cluster pairs for one cluster */
#define SPONGE_NL_OVERFLOW_CANDIDATES 4

/* Energy / virial parts accumulated in FP64 on the device by the fused step
 * and by the energy queries (index into a double[SPONGE_NB_NPARTS]). */
#define SPONGE_PART_LJ 0          /* sum of LJ pair energies (r < rc)          */
#define SPONGE_PART_DIRECT 1      /* PME direct-space energy                   */
#define SPONGE_PART_LJ_VIRIAL 2   /* sum of -f_LJ r^2                          */
#define SPONGE_PART_CORRECTION 3  /* excluded-pair correction energy          */
#define SPONGE_PART_SELF 4        /* self + neutralising energy                */
#define SPONGE_PART_RECIPROCAL 5  /* reciprocal-space energy                   */
#define SPONGE_PART_LRC 6         /* long-range dispersion energy (factor / V) */
#define SPONGE_PART_PAIRS_IN 7    /* number of listed pairs inside the cutoff  */
#define SPONGE_NB_NPARTS 8

typedef struct sponge_nl sponge_nl;
typedef struct sponge_lj sponge_lj;
typedef struct sponge_pme sponge_pme;

const char* sponge_nb_last_error(void);

/* Number of usable devices and the name of the device the library runs on. */
int sponge_nb_device_info(int* count, char* name, int name_len);

/* ------------------------------------------------------------------ */
/* MD_INFORMATION::MD_Information_Crd_To_Uint_Crd (MD_core/MD_core.cu:1528-1532),
 * kernel Crd_To_Uint_Crd (common.cu:264-287): u = ((int)(x * 2^30 / L)) << 2. */
int sponge_crd_to_uint(int atom_numbers, const float box_length[3], const float* d_crd,
                       unsigned int* d_uint_crd, void* stream);

/* ------------------------------------------------------------------ */
/* NEIGHBOR_LIST (neighbor_list/neighbor_list.cuh:44-105).                */

/* NEIGHBOR_LIST::Initial (neighbor_list/neighbor_list.cu:580-642). The mdin
 * keys neighbor_list_refresh_interval / _max_atom_in_grid_numbers /
 * _max_neighbor_numbers / _skin_permit arrive already parsed (skin_permit as
 * written in mdin; the doubling of neighbor_list.cu:610 happens inside). */
int sponge_nl_create(int atom_numbers, const float box_length[3], float cutoff, float skin,
                     int refresh_interval, int max_atom_in_grid_numbers,
                     int max_neighbor_numbers, float skin_permit, sponge_nl** out);
/* NEIGHBOR_LIST::Clear (neighbor_list/neighbor_list.cu:659-711). */
void sponge_nl_destroy(sponge_nl* nl);

/* NEIGHBOR_LIST::Neighbor_List_Update (neighbor_list/neighbor_list.cu:510-556):
 * forced / every-refresh_interval / skin-triggered rebuild. d_crd (AoS
 * float3, device) is wrapped into the box in place on a rebuild, as the
 * reference's Crd_Periodic_Map does. The exclusion CSR is the reference's
 * (sorted, upper triangular). The skin-triggered path is decided on the
 * device (no host round trip, no device-side launch). */
int sponge_nl_update(sponge_nl* nl, float* d_crd, const int* d_excluded_list_start,
                     const int* d_excluded_list, const int* d_excluded_numbers,
                     int forced_update, int forced_check, void* stream);

/* No counterpart in the reference (its exclusion deletion reads the CSR on
 * every rebuild). Synchronises. The list keeps a symmetric copy of the
 * exclusion CSR (the partners of every atom, both directions) for its
 * kernels; sponge_nl_update builds it the first time it sees a set of CSR
 * pointers (synchronising once; refused inside a stream capture) and reuses
 * it while the same pointers come back. A caller that rewrites the CSR in
 * place, or reuses the addresses for new arrays, calls this to rebuild it. */
int sponge_nl_set_exclusions(sponge_nl* nl, const int* d_excluded_list_start, const int* d_excluded_list,
                             const int* d_excluded_numbers);

/* Inner list of the fused step (no counterpart in the reference);
 * prune_interval 0 turns it off (the cluster-pair list every step). Plain LJ
 * (force-only steps): each rebuild also writes the list's pairs per atom
 * (every pair within cutoff + skin at the rebuild, the reference's own
 * list) and the fused steps read them; with
 * prune_radius >= cutoff + skin nothing else is pruned. With a smaller
 * prune_radius (>= cutoff) those per-atom lists are the outer ones: every
 * prune_interval fused steps an inner list keeps their pairs within
 * prune_radius of the current coordinates. Soft-core LJ prunes its
 * cluster-pair list to prune_radius on the same schedule. Between
 * scheduled prunes of a distance-pruned list, a step at which some atom has
 * moved more than (prune_radius - cutoff) / 2 since the last prune prunes at
 * once (decided on the device), so the inner list never misses a pair
 * inside the cutoff that the list it is pruned from holds. Default for
 * refresh_interval > 0: the tuned table's values for the atom count
 * (sponge_nb_default_shapes); off otherwise. The next fused step after this
 * call writes the lists itself (a kernel on the null stream marks it: not
 * during a graph capture). The other entry points always use the full
 * list. */
int sponge_nl_set_prune(sponge_nl* nl, int prune_interval, float prune_radius);

/* Synchronises. Number of inner-list prunes that the displacement test above
 * (or a new prune setting) caused beyond the scheduled ones since the list
 * was created. */
int sponge_nl_prune_votes(sponge_nl* nl, int* voted_prunes);

/* NEIGHBOR_LIST::Update_Volume (neighbor_list/neighbor_list.cu:645-656). */
int sponge_nl_update_volume(sponge_nl* nl, const float box_length[3]);

/* Synchronises. Number of (i<j) atom pairs currently in the list, overflow
 * flags (SPONGE_NL_OVERFLOW_*), and how many rebuilds have run. */
int sponge_nl_status(sponge_nl* nl, long* pair_count, int* overflow_flags, int* rebuilds);

/* Synchronises. Diagnostics of the fused step's inner list (no counterpart in
 * the reference): with plain LJ, force-only fused steps read per-slot partner
 * lists pruned to the inner radius; `slot_overflow` = 1 when a slot had more
 * partners than those lists hold at a prune since the last rebuild (its
 * cluster then reads the full list), and `slot_launches` counts the
 * pair-kernel launches that read the per-slot form since the list was
 * created. */
int sponge_nl_inner_status(sponge_nl* nl, int* slot_overflow, int* slot_launches);

/* Synchronises. The list as canonical atom pairs (i<j, original indices),
 * sorted; at most `capacity` pairs are written to h_pairs (2 ints each).
 * Replaces reading d_nl[i].atom_serial on the host (neighbor_list.cu:490-503).
 * Also checks that every pair is stored from both of its atoms (the layout
 * the pair kernel relies on): SPONGE_NB_EINTERNAL if not. */
int sponge_nl_get_pairs(sponge_nl* nl, int* h_pairs, long capacity, long* pair_count);

/* ------------------------------------------------------------------ */
/* LENNARD_JONES_INFORMATION (Lennard_Jones_force/Lennard_Jones_force.cuh:37-149). */

/* LENNARD_JONES_INFORMATION::Initial (Lennard_Jones_force.cu:1067-1139) after
 * the LJ_in_file / amber_parm7 read: h_LJ_A / h_LJ_B already multiplied by
 * 12 / 6, lower-triangular pair storage, h_atom_LJ_type per atom. */
int sponge_lj_create(int atom_numbers, int atom_type_numbers, const float* h_LJ_A,
                     const float* h_LJ_B, const int* h_atom_LJ_type, float cutoff,
                     const float box_length[3], sponge_lj** out);
void sponge_lj_destroy(sponge_lj* lj);

/* LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial
 * (Lennard_Jones_force.cu:1318-1352 and kernels :47-562). Accumulates the
 * fused LJ + erfc direct-Coulomb force into d_frc; with need_atom_energy the
 * LJ + direct energy into d_atom_energy; with need_virial the LJ virial into
 * d_atom_lj_virial. d_atom_direct_pme_energy is reset and refilled whenever
 * energy or virial is requested (the reference resets it only on the virial
 * path). Per-atom terms are credited to the lower-index atom of each pair,
 * as the reference's half list does. */
int sponge_lj_pme_direct_force(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                               const float* d_charge, float* d_frc, float pme_beta,
                               int need_atom_energy, float* d_atom_energy, int need_virial,
                               float* d_atom_lj_virial, float* d_atom_direct_pme_energy,
                               void* stream);

/* LENNARD_JONES_INFORMATION::LJ_Force_With_PME_Direct_Force, the overload of
 * Lennard_Jones_force.cuh:124 (uint coordinates and charges, the module's
 * cutoff and box): the LJ + PME direct force alone, added to d_frc (what
 * LJ_PME_Direct_Force_With_Atom_Energy_And_Virial does without energy or
 * virial, Lennard_Jones_force.cu:1318-1328). */
int sponge_lj_force_with_pme_direct_force(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                          const float* d_charge, float* d_frc, float pme_beta, void* stream);

/* LENNARD_JONES_INFORMATION::LJ_Force_With_PME_Direct_Force, the overload of
 * Lennard_Jones_force.cuh:121 (.cu:1297-1306): the same force from the
 * caller's UINT_VECTOR_LJ_TYPE records (uint x, y, z, LJ type, charge; 20
 * bytes per atom, Lennard_Jones_force.cuh:26-33) with an explicit scaler
 * (uint to Angstrom per axis) and cutoff. atom_numbers must be the module's
 * atom count: the reference lets it be smaller to evaluate part of the
 * system for SITS, which is not provided here (SPONGE_NB_EARG). Plain LJ
 * handles only. */
int sponge_lj_force_with_pme_direct_force_records(sponge_lj* lj, const sponge_nl* nl, int atom_numbers,
                                                 const void* d_uint_crd_with_LJ, const float scaler[3], float* d_frc,
                                                 float cutoff, float pme_beta, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and kernels as
 * sponge_lj_pme_direct_force, but its scratch (cluster-order records and FP64
 * totals) is a static device array of the library (2 MiB per GPU, held until
 * the process ends) instead of the module's lazily allocated buffers: it
 * calls no hipMalloc / hipFree. Lists of more than 131,072 slots (about
 * 125,000 atoms) are refused with SPONGE_NB_ECAPACITY. Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller. */
int sponge_lj_pme_direct_force_noalloc(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                       const float* d_charge, float* d_frc, float pme_beta,
                                       int need_atom_energy, float* d_atom_energy, int need_virial,
                                       float* d_atom_lj_virial, float* d_atom_direct_pme_energy,
                                       void* stream);

/* LENNARD_JONES_INFORMATION::Long_Range_Correction (Lennard_Jones_force.cu:1178-1191). */
int sponge_lj_long_range_correction(sponge_lj* lj, int need_pressure, float* d_virial,
                                    int need_potential, float* d_potential, void* stream);

/* LENNARD_JONES_INFORMATION::Get_Energy (Lennard_Jones_force.cu:1407-1433):
 * LJ energy over the list (r < rc) plus the long-range correction, FP64.
 * Synchronises. */
int sponge_lj_get_energy(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                         double* h_energy, void* stream);

/* LENNARD_JONES_INFORMATION::Update_Volume (Lennard_Jones_force.cu:1450-1456). */
int sponge_lj_update_volume(sponge_lj* lj, const float box_length[3]);

/* Replaces the A / B tables of a plain handle (same layout as
 * sponge_lj_create; used to evaluate new parameters, e.g. by the gradient
 * path's callers). Synchronous; not for soft-core handles. No counterpart in
 * the reference, whose tables are fixed after Initial. */
int sponge_lj_set_tables(sponge_lj* lj, const float* h_LJ_A, const float* h_LJ_B);

/* Host copies the reference keeps public for nb14 (Lennard_Jones_force.cuh:52-58). */
int sponge_lj_host_tables(const sponge_lj* lj, const float** h_LJ_A, const float** h_LJ_B,
                          const int** h_atom_LJ_type);

/* LJ_SOFT_CORE (Lennard_Jones_force/LJ_soft_core.cuh:37-160): Initial
 * (LJ_soft_core.cu:1420-1610) after its reads of LJ_soft_core_in_file
 * (`N TA TB`, AA and AB of state A, BA and BB of state B, then per atom
 * `typeA typeB`; tables x12 / x6 as stored) and subsys_division_in_file
 * (h_subsys_division may be NULL: every mask 0). lambda_lj, soft_core_alpha
 * (0.5), soft_core_powfer (1), soft_core_sigma (3.0), soft_core_sigma_min (0).
 * The handle is a sponge_lj: LJ_Soft_Core_PME_Direct_Force_With_Atom_Energy_And_Virial
 * (:1740-1774) is sponge_lj_pme_direct_force on it, and Long_Range_Correction,
 * Get_Energy, Update_Volume, sponge_nb_force_step take it like a plain handle.
 * At most 64 distinct (type A, type B, mask) combinations. */
int sponge_lj_soft_create(int atom_numbers, int atom_type_numbers_A, int atom_type_numbers_B,
                          const float* h_LJ_AA, const float* h_LJ_AB, const float* h_LJ_BA,
                          const float* h_LJ_BB, const int* h_atom_LJ_type_A,
                          const int* h_atom_LJ_type_B, const int* h_subsys_division, float lambda,
                          float alpha, float p, float sigma, float sigma_min, float cutoff,
                          const float box_length[3], sponge_lj** out);

/* Measurement hook (no counterpart in the reference): when set, every later
 * pair-kernel launch of this module records the two hipEvent_t as its start
 * and end (hipExtLaunchKernel on the launch stream: the kernel's own times,
 * without the latency of marker packets around it), so a caller can time the
 * dominant kernel inside a timed region. Both events or neither; NULL events
 * switch the hook off. Not for graph capture. */
int sponge_lj_set_pair_events(sponge_lj* lj, void* start_event, void* stop_event);

/* ------------------------------------------------------------------ */
/* Particle_Mesh_Ewald (PME_force/PME_force.cuh:25-144).               */

/* Particle_Mesh_Ewald::Initial (PME_force.cu:108-306): beta from the
 * PME_Direct_Tolerance bisection, grid from PME_fftx/y/z (<=0 = automatic,
 * Get_Fft_Patameter), order-4 BC table. */
int sponge_pme_create(int atom_numbers, const float box_length[3], float cutoff,
                      float direct_tolerance, int fftx, int ffty, int fftz, sponge_pme** out);
void sponge_pme_destroy(sponge_pme* pme);

/* Measurement hook (no counterpart in the reference), like
 * sponge_lj_set_pair_events: when set, every later launch of the charge
 * spread (with its z transforms) records spread_start / spread_stop and every
 * force gather records gather_start / gather_stop (hipEvent_t) at the
 * kernel's own start and end. NULL pairs switch them off. Not for graph
 * capture. */
int sponge_pme_set_events(sponge_pme* pme, void* spread_start, void* spread_stop, void* gather_start,
                          void* gather_stop);

/* Launch-shape tuning (no counterpart in the reference; -1 keeps a
 * setting). step_form: how sponge_nb_force_step issues a step when its flags
 * do not say -- 0 the pair kernel on a second stream, concurrent with the
 * reciprocal pipeline (default), 1 every kernel on its own. conv_split 0 / 1
 * / 2 selects the one-workgroup-per-plane convolution, the three line
 * launches, or the slab form (the spread and the inverse transform take the
 * y transforms of one x plane each, the convolution the x lines); a form
 * whose workgroups do not fit the LDS is refused. */
int sponge_pme_set_tuning(sponge_pme* pme, int step_form, int conv_split);

/* Launch shapes and schedule of the fused step as one record (no counterpart
 * in the reference, whose launch shapes are constants:
 * Lennard_Jones_force.cu:1327-1350, neighbor_list.cu:435, 480). The handles'
 * creators take their defaults from a table measured per system size by the
 * shape sweep of sponge_amd/autotune.py; a caller may sweep its own system
 * and set the winner. */
typedef struct sponge_shapes {
    int pair_waves;     /* i-clusters (waves) per block of the force-only pair kernel: 2, 4 or 8 */
    int conv_split;     /* 0: one workgroup per kz plane; 1: three line launches; 2: slab form */
    int conv_threads;   /* threads of the per-plane workgroup: 256 or 512 */
    int step_form;      /* 0: pair kernel on a second stream; 1: every kernel on the caller's stream */
    int prune_interval; /* fused steps between scheduled prunes of the inner list; 0: no inner list */
    float prune_radius; /* A, >= cutoff */
} sponge_shapes;

/* The tuned table's shapes for a system of atom_numbers atoms with the given
 * list settings (planes beyond the LDS take conv_split 1 whatever the
 * table says). */
int sponge_nb_default_shapes(int atom_numbers, int refresh_interval, float cutoff, float skin, sponge_shapes* out);

/* The shapes the handles run with now. */
int sponge_nb_get_shapes(const sponge_nl* nl, const sponge_lj* lj, const sponge_pme* pme, sponge_shapes* out);

/* Sets every field (validated first: nothing changes on an error). Outside
 * graph capture: a new prune setting marks the next step for a prune. */
int sponge_nb_set_shapes(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const sponge_shapes* shapes);

/* Public members the reference's callers read: beta (PME_force.cuh:62) and
 * fftx/ffty/fftz (:30-32). */
int sponge_pme_info(const sponge_pme* pme, float* beta, int fft[3]);

/* Particle_Mesh_Ewald::PME_Excluded_Force_With_Atom_Energy
 * (PME_force.cu:697-710, kernel :620-695). Also keeps the per-atom
 * correction energy internally (d_correction_atom_energy). */
int sponge_pme_excluded_force(sponge_pme* pme, const unsigned int* d_uint_crd,
                              const float scaler[3], const float* d_charge,
                              const int* d_excluded_list_start, const int* d_excluded_list,
                              const int* d_excluded_atom_numbers, float* d_frc,
                              float* d_atom_energy, void* stream);

/* Particle_Mesh_Ewald::PME_Reciprocal_Force_With_Energy_And_Virial
 * (PME_force.cu:722-773): spread, 3D R2C FFT, BC convolution, C2R FFT,
 * gather into d_frc; with need_energy / need_virial adds
 * correction + self + reciprocal energy to *d_potential and
 * direct + correction + self + reciprocal to *d_virial (FP64 reductions).
 * d_direct_atom_energy is the array last filled by
 * sponge_lj_pme_direct_force (may be NULL when neither is needed). */
int sponge_pme_reciprocal_force(sponge_pme* pme, const unsigned int* d_uint_crd,
                                const float* d_charge, float* d_frc, int need_virial,
                                int need_energy, float* d_virial, float* d_potential,
                                const float* d_direct_atom_energy, void* stream);

/* Particle_Mesh_Ewald::Get_Energy (PME_force.cu:892-994). which_part:
 * 0 total, 1 direct, 2 reciprocal, 3 correction, 4 self (PME_force.cuh:86-93).
 * FP64, synchronises. */
int sponge_pme_get_energy(sponge_pme* pme, const sponge_nl* nl, const unsigned int* d_uint_crd,
                          const float* d_charge, const float scaler[3],
                          const int* d_excluded_list_start, const int* d_excluded_list,
                          const int* d_excluded_atom_numbers, int which_part, double* h_energy,
                          void* stream);

/* Particle_Mesh_Ewald::Update_Box_Length (PME_force.cu:1033-1042); the
 * reference's Update_Volume (:996-999) forwards the stale member box, this
 * entry point takes the new box for both. */
int sponge_pme_update_box_length(sponge_pme* pme, const float box_length[3]);

/* ------------------------------------------------------------------ */
/* Fused per-step force path: the body of Main_Calculate_Force's nonbonded
 * calls (main.cu:170-193) in the minimum number of launches. With d_crd
 * non-NULL the uint coordinates are first computed from it and written to
 * d_uint_crd (MD_Information_Crd_To_Uint_Crd, MD_core.cu:1528-1532, fused);
 * with d_crd NULL d_uint_crd is read. All nonbonded forces are added to d_frc
 * (or, with SPONGE_STEP_OVERWRITE_FORCE, d_frc is overwritten: the caller need
 * not zero it). When need_energy / need_virial is set, the FP64 parts
 * (SPONGE_PART_*) are written to d_parts (device, 8 doubles) and the per-atom
 * arrays are filled like the separate calls fill them.
 * The pair kernel runs on a stream of the PME handle, forked from and joined
 * back into `stream` with events (so the call stays capturable into a graph),
 * concurrently with the reciprocal spread / FFT / convolution on `stream`;
 * SPONGE_STEP_SERIAL keeps every launch on `stream`, SPONGE_STEP_TWO_STREAMS
 * forces the concurrent form; without either the form set by
 * sponge_pme_set_tuning applies. */
#define SPONGE_STEP_OVERWRITE_FORCE 1
#define SPONGE_STEP_SERIAL 2
#define SPONGE_STEP_TWO_STREAMS 8
int sponge_nb_force_step(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme, const float* d_crd,
                         unsigned int* d_uint_crd, const float* d_charge,
                         const int* d_excluded_list_start, const int* d_excluded_list,
                         const int* d_excluded_numbers, float* d_frc, int flags, int need_energy,
                         int need_virial, float* d_atom_energy, float* d_atom_virial,
                         double* d_parts, void* stream);

/* ------------------------------------------------------------------ */
/* Gradients (no counterpart in the reference, which has no adjoint). Both
 * recompute the pair terms over the current neighbor list, run the
 * reciprocal pipeline, and write FP64 device arrays: d_grad_charge (n, in
 * the units of the stored charges), d_grad_A / d_grad_B (T(T+1)/2, with
 * respect to the stored x12 / x6 pair tables). Plain LJ handles with at most
 * 8 types. The first call allocates the adjoint scratch.
 * Energy: E = E_LJ + E_lrc + E_direct + E_correction + E_reciprocal + E_self.
 * Force matching: L = sum_i |F_i - T_i|^2 from the caller's forces F (of the
 * same parameters, e.g. from sponge_nb_force_step) and targets T; *d_loss = L. */
int sponge_nb_energy_gradient(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme,
                              const unsigned int* d_uint_crd, const float* d_charge,
                              const int* d_excluded_list_start, const int* d_excluded_list,
                              const int* d_excluded_numbers, double* d_grad_charge,
                              double* d_grad_A, double* d_grad_B, void* stream);
int sponge_nb_force_matching_gradient(sponge_nl* nl, sponge_lj* lj, sponge_pme* pme,
                                      const unsigned int* d_uint_crd, const float* d_charge,
                                      const int* d_excluded_list_start, const int* d_excluded_list,
                                      const int* d_excluded_numbers, const float* d_frc,
                                      const float* d_target_frc, double* d_loss,
                                      double* d_grad_charge, double* d_grad_A, double* d_grad_B,
                                      void* stream);

/* ------------------------------------------------------------------ */
/* Batched systems (no counterpart in the reference, which runs one system
 * per process; the FEP windows of main_fep.cu / TI_core each get their own).
 * A batch groups already-created module handles of independent systems --
 * FEP lambda windows (soft-core handles) and replicas of any size -- and runs
 * each stage of the step as ONE launch over all of them (block maps with
 * per-system offsets), so small systems fill the GPU together. Per system
 * the caller names its device arrays once, like MD_INFORMATION owns them. */
typedef struct sponge_batch sponge_batch;
typedef struct {
    float* crd;                     /* wrapped in place by list rebuilds */
    unsigned int* uint_crd;
    const float* charge;
    const int* excluded_list_start;
    const int* excluded_list;
    const int* excluded_numbers;
    float* frc;
    float* atom_energy;             /* energy steps, may be NULL */
    float* atom_virial;             /* energy steps, may be NULL */
    double* parts;                  /* SPONGE_NB_NPARTS doubles, energy steps, may be NULL */
} sponge_batch_io;

/* Handles stay owned by the caller and must outlive the batch; every system's
 * list, LJ and PME handles must agree on atom count and box, and all lists on
 * refresh_interval. Call sponge_batch_sync after changing any member handle
 * (e.g. *_update_volume). */
int sponge_batch_create(int nsys, sponge_nl* const* nl, sponge_lj* const* lj, sponge_pme* const* pme,
                        const sponge_batch_io* io, sponge_batch** out);
void sponge_batch_destroy(sponge_batch* batch);
int sponge_batch_sync(sponge_batch* batch);
/* Neighbor_List_Update of every system (refresh_interval > 0: rebuild all on
 * the interval; 0: each system's own skin check decides on the device). */
int sponge_batch_neighbor_list_update(sponge_batch* batch, int forced_update, void* stream);
/* sponge_nb_force_step of every system; SPONGE_STEP_FROM_CRD computes each
 * system's uint coordinates from its crd first. The reciprocal stages run on
 * a stream of the batch beside the pair and exclusion stages (forked and
 * joined with events, so the step stays capturable); SPONGE_STEP_SERIAL keeps
 * every stage on `stream`. */
#define SPONGE_STEP_FROM_CRD 4
int sponge_batch_force_step(sponge_batch* batch, int flags, int need_energy, int need_virial,
                            void* stream);

/* Sum_Of_List / Add_Sum_List (common.cu:192-263, MD_core.cu:173-181) with an
 * FP64 accumulator: *d_out += sum(d_list[0..n)). */
int sponge_sum_list(int n, const float* d_list, float* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SPONGE_NB_H */
