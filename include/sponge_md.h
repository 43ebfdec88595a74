/*
 * sponge_md.h -- C ABI of the MI355X (gfx950) step on either side of the
 * nonbonded force path: the NVE leap-frog integrator and SETTLE rigid-body
 * constraints (rigid triangles such as 3-site waters, and rigid pairs).
 *
 * Conventions are those of sponge_nb.h: d_* are device pointers in the
 * reference's MD_INFORMATION layout (VECTOR = 3 floats AoS, original atom
 * order), `stream` is a hipStream_t, no stream entry point allocates, frees
 * or synchronises (so a whole MD step can be captured into one hipGraph),
 * and every function returns 0 or a negative SPONGE_NB_E* code with
 * sponge_nb_last_error() describing the failure. Citations are relative to
 * the reference's SPONGE/ directory. Units are SPONGE's: dt in internal time
 * units (ps x 20.455), masses in g/mol, lengths in Angstrom.
 */
#ifndef SPONGE_MD_H
#define SPONGE_MD_H

#include "sponge_nb.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sponge_settle sponge_settle;

/* MD_INFORMATION::NVE_iteration::Leap_Frog (MD_core/MD_core.cu:1535-1547),
 * kernels MD_Iteration_Leap_Frog / _With_Max_Velocity (MD_core.cu:13-45):
 * acc = frc / m, vel += dt acc (norm capped at max_velocity when > 0),
 * crd += dt vel. d_acc may be NULL (then it is not written). */
int sponge_md_leap_frog(int atom_numbers, float* d_vel, float* d_crd, const float* d_frc, float* d_acc,
                        const float* d_inverse_mass, float dt, float max_velocity, void* stream);

/* SETTLE::Initial (constrain/SETTLE.cu:189-380) over the constraint pairs
 * that CONSTRAIN::Initial_Constrain (constrain/constrain.cu:4-80) holds
 * (atom_i[k], atom_j[k], constant_r[k]; host arrays). Atoms linked by three
 * mutual constraints form a rigid triangle, atoms linked by one constraint
 * to each other only form a rigid pair, with the reference's discovery rule
 * and atom roles; h_mass is the host mass array. dt is the step, exp_gamma
 * the thermostat's velocity factor (1 for NVE): velocities are corrected by
 * exp_gamma / (dt * x_factor) times the position correction, with
 * x_factor = (1 + exp_gamma) / 2 (constrain.cu:16-17); is_minimization sets
 * the velocity factor to 0 (constrain.cu:19-22). */
int sponge_settle_create(int atom_numbers, int constrain_pair_numbers, const int* atom_i, const int* atom_j,
                         const float* constant_r, const float* h_mass, float dt, float exp_gamma,
                         int is_minimization, sponge_settle** out);
void sponge_settle_destroy(sponge_settle* s);

/* Number of rigid triangles, rigid pairs, and constraint pairs left over for
 * another constraint method (constrain/SETTLE.cu:339-360 keeps those for
 * SIMPLE_CONSTRAIN / SHAKE, which this library does not provide: a caller
 * with remaining > 0 must constrain them itself). Host only. */
int sponge_settle_counts(const sponge_settle* s, int* triangle_numbers, int* pair_numbers, int* remaining);

/* The rigid groups as found (host arrays of the sizes given by counts, any
 * may be NULL): triangles[3k..3k+2] = atoms A, B, C; geometry[5k..5k+4] = the
 * canonical positions (ra, rb, rc, rd, re) of Get_Rabcde_From_SSS
 * (SETTLE.cu:8-47); pairs[2k..2k+1] = atoms i, j; pair_r[k] = length. */
int sponge_settle_groups(const sponge_settle* s, int* triangles, float* geometry, int* pairs, float* pair_r);

/* Thermostat coupling: update exp_gamma (v_factor = exp_gamma,
 * x_factor = (1 + exp_gamma) / 2), as main.cu:303-312 does between steps. */
int sponge_settle_set_factors(sponge_settle* s, float v_factor, float x_factor);

/* SETTLE::Remember_Last_Coordinates (constrain/SETTLE.cu:494-505): stores the
 * periodic B-A, C-A (triangles) and j-i (pairs) displacements of the current
 * fixed-point coordinates; scaler = box_length / 2^32 per axis. */
int sponge_settle_remember_last_coordinates(sponge_settle* s, const unsigned int* d_uint_crd,
                                            const float scaler[3], void* stream);

/* SETTLE::Do_SETTLE (constrain/SETTLE.cu:507-525): moves each rigid group
 * back onto its constraints after an unconstrained step, corrects the
 * velocities, and when need_pressure adds the constraint virial
 * sum(F . r_old) / (3 V) to *d_pressure (one float, device). */
int sponge_settle_do(sponge_settle* s, const float* d_mass, float* d_crd, const float box_length[3],
                     float* d_vel, int need_pressure, float* d_pressure, void* stream);

/* One launch doing what main.cu:282-320 does for NVE with SETTLE:
 * Remember_Last_Coordinates (from d_uint_crd), Leap_Frog (dt, max_velocity)
 * and Do_SETTLE, fused over the rigid groups plus the unconstrained atoms.
 * Same arithmetic as the three calls in order (results are bit-identical);
 * s may be NULL (plain leap-frog of every atom). d_acc may be NULL. */
int sponge_md_iteration(sponge_settle* s, int atom_numbers, float* d_vel, float* d_crd, const float* d_frc,
                        float* d_acc, const float* d_inverse_mass, const float* d_mass,
                        const unsigned int* d_uint_crd, const float box_length[3], float dt,
                        float max_velocity, int need_pressure, float* d_pressure, void* stream);

/* ------------------------------------------------------------------ */
/* Batched integrator. No counterpart in the reference, which runs each FEP
 * window or replica as its own process (main.cu:282-320 per process): per
 * system it is sponge_md_iteration, for every system of a sponge_batch in
 * ONE launch, so a batched MD step is sponge_batch_force_step (FROM_CRD |
 * OVERWRITE_FORCE) + sponge_batch_md_iteration + sponge_batch_neighbor_
 * list_update, each one launch per stage, and capturable as one hipGraph.
 * The coordinates, forces, fixed-point coordinates and box of each system
 * are the batch's (sponge_batch_io, its handles' box). */
typedef struct {
    float* vel;
    float* acc;                     /* may be NULL */
    const float* inverse_mass;
    const float* mass;              /* may be NULL for a system without SETTLE */
    float dt;
    float max_velocity;
} sponge_batch_md_io;

/* Names each system's SETTLE handle (settle may be NULL, or settle[k] NULL:
 * plain leap-frog of every atom of system k) and integrator arrays; a SETTLE
 * handle must have the system's atom count and stay alive with the batch.
 * Replaces an earlier call's setting; on an error the batch keeps it.
 * sponge_batch_sync re-reads the SETTLE handles' factors
 * (sponge_settle_set_factors) and the boxes. Allocates; host-synchronous. */
int sponge_batch_set_integrator(sponge_batch* batch, sponge_settle* const* settle, const sponge_batch_md_io* io);
/* sponge_md_iteration (need_pressure 0) of every system, one launch. */
int sponge_batch_md_iteration(sponge_batch* batch, void* stream);

/* ------------------------------------------------------------------ */
/* Pressure and volume coupling. */

/* MD_INFORMATION::system_information::Get_Pressure (MD_core/MD_core.cu:
 * 1777-1795, kernels MD_Atom_Ek :241-250 and Calculate_Pressure_Cuda
 * :183-186), the body of Calculate_Pressure_And_Potential_If_Needed
 * (:1765-1775) when need_pressure: Ek = sum m v^2 / 2 of d_vel,
 * *d_virial += sum of d_atom_virial (may be NULL), then
 * *d_pressure = (2 Ek + *d_virial) / (3 volume) (assigned), in kcal/mol/A^3;
 * *d_ek = Ek when d_ek is not NULL. FP64 sums, one launch. The constraint
 * virial is added afterwards by sponge_settle_do / sponge_md_iteration. */
int sponge_md_get_pressure(int atom_numbers, const float* d_vel, const float* d_mass, const float* d_atom_virial,
                           float* d_virial, float volume, float* d_pressure, float* d_ek, void* stream);

/* The same pressure from the FP64 parts of sponge_nb_force_step (no
 * counterpart in the reference, whose force calls accumulate *d_virial):
 * W = parts[LJ_VIRIAL] + 6 parts[LRC] + parts[DIRECT] + parts[CORRECTION]
 *   + parts[SELF] + parts[RECIPROCAL]. */
int sponge_md_pressure_from_parts(int atom_numbers, const float* d_vel, const float* d_mass, const double* d_parts,
                                  float volume, float* d_pressure, float* d_ek, void* stream);

/* Scale_List (common.cu): d_list[0..count) *= factor; the barostats' scaling
 * of coordinates and velocities (main.cu:334-372). */
int sponge_md_scale_list(int count, float* d_list, float factor, void* stream);

#ifdef __cplusplus
}
#endif
#endif
