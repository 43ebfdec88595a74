/*
 * sponge_fep.h -- C ABI of the MI355X (gfx950) FEP / TI frame evaluators:
 * the nonbonded parts of SPONGE_FEP's partition energies (main_fep.cu:59-134)
 * and of SPONGE_TI's dH/dlambda (main_ti.cu:132-157), evaluated on one frame
 * of a trajectory over the cluster neighbor list of sponge_nb.h.
 *
 * Conventions are those of sponge_nb.h: d_* are device pointers in the
 * reference's layout (original atom order), `stream` is a hipStream_t, every
 * function returns 0 or a negative SPONGE_NB_E* code with
 * sponge_nb_last_error() describing the failure, and citations are relative
 * to the reference's SPONGE/ directory. Every entry point here returns host
 * values and synchronises `stream` (the reference's is_download = 1 path);
 * the first call on a handle allocates its evaluator scratch.
 *
 * Pair sums run in FP64 from the exact integer displacement, each unordered
 * pair of the list inside the cutoff once. The list must be current for the
 * frame (main_fep.cu:167-172 rebuilds it for every frame).
 */
#ifndef SPONGE_FEP_H
#define SPONGE_FEP_H

#include "sponge_nb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fields of one lambda window in sponge_lj_soft_lambda_windows' output
 * (index into double[SPONGE_FEP_NFIELDS] per window). */
#define SPONGE_FEP_VDW_INTERSYS 0          /* soft-core LJ, pairs of different masks      */
#define SPONGE_FEP_VDW_INTRASYS 1          /* soft-core LJ, pairs of one mask             */
#define SPONGE_FEP_COUL_DIRECT_INTERSYS 2  /* PME direct Coulomb, different masks         */
#define SPONGE_FEP_COUL_DIRECT_INTRASYS 3  /* PME direct Coulomb, one mask                */
#define SPONGE_FEP_DHDL_LJ 4               /* dH/dlambda of the soft-core LJ               */
#define SPONGE_FEP_DHDL_DIRECT 5           /* dH/dlambda of direct Coulomb, reference form */
#define SPONGE_FEP_DHDL_DIRECT_EXACT 6     /* dH/dlambda of direct Coulomb, exact          */
#define SPONGE_FEP_VDW_LRC 7               /* Long_Range_Correction() at the window        */
#define SPONGE_FEP_DHDL_LRC 8              /* its lambda derivative                        */
#define SPONGE_FEP_NFIELDS 9
#define SPONGE_FEP_MAX_WINDOWS 64

/* ------------------------------------------------------------------ */
/* LJ_SOFT_CORE (Lennard_Jones_force/LJ_soft_core.cuh:37-160) on a handle of
 * sponge_lj_soft_create. pme_beta is the module's Get_Beta_LJ(cutoff,
 * PME_Direct_Tolerance) (LJ_soft_core.cu:6-35, :1569-1574), the same
 * bisection as Particle_Mesh_Ewald's: pass sponge_pme_info's beta. */

/* LJ_SOFT_CORE::Get_Energy_With_Coulomb_Direct (LJ_soft_core.cu:1809-1852,
 * kernel :970-1135) at the handle's lambda_lj with the given charges:
 * h_energy = {LJ intersys, LJ intrasys, direct intersys, direct intrasys}
 * (h_LJ_energy_sum_intersys, _intrasys, h_direct_ene_sum_intersys,
 * _intrasys). Soft-core pairs keep an end state's terms only while that
 * state's soft-core radius is below the cutoff, as the reference does. */
int sponge_lj_soft_energy_with_coulomb_direct(sponge_lj* lj, const sponge_nl* nl,
                                              const unsigned int* d_uint_crd, const float* d_charge,
                                              float pme_beta, double h_energy[4], void* stream);

/* LJ_SOFT_CORE::Get_Partial_H_Partial_Lambda_With_Columb_Direct
 * (LJ_soft_core.cu:1854-1894, kernels :1137-1284 and :1286-1418):
 * h_dhdl = {sum dH/dlambda LJ (h_sigma_of_dH_dlambda_lj), sum dH/dlambda
 * direct (h_sigma_of_dH_dlambda_direct)}. charge_perturbed = 0 ignores
 * d_charge_B_A (may be NULL). With charges perturbed, the reference's
 * ordinary (not soft-core) pairs contribute (dq_i q_j + q_i dq_j) erfc(b r)
 * without the 1/r of the pair's Coulomb energy (LJ_soft_core.cu:1271); this
 * entry point keeps that form so that results match the reference, and
 * sponge_lj_soft_lambda_windows reports both forms. */
int sponge_lj_soft_dhdl_with_coulomb_direct(sponge_lj* lj, const sponge_nl* nl,
                                            const unsigned int* d_uint_crd, const float* d_charge,
                                            const float* d_charge_B_A, int charge_perturbed,
                                            float pme_beta, double h_dhdl[2], void* stream);

/* LJ_SOFT_CORE::Long_Range_Correction() (LJ_soft_core.cu:1927-1937) and
 * Partial_H_Partial_Lambda_Long_Range_Correction (:1939-1955): factor / V
 * with the factor of the handle's lambda, and (sum BB - sum AB) factor / V
 * (Total_C6_B_A_Get, :123-157). Host values, no device work. */
int sponge_lj_soft_long_range_correction(const sponge_lj* lj, double* h_lrc, double* h_dhdl_lrc);

/* All lambda windows of one frame in one pass (no counterpart in the
 * reference, which evaluates one lambda per SPONGE_FEP / SPONGE_TI run):
 * the list is walked once; pairs whose energy is polynomial in lambda
 * (ordinary pairs: LJ linear, Coulomb quadratic through the charges) are
 * reduced to lambda-free sums, soft-core pairs are recorded and evaluated
 * for every window in a second launch. Window w has LJ lambda h_lambda[w]
 * and charges d_charge + (h_lambda[w] - lambda_lj) d_charge_B_A (d_charge
 * are the charges of the handle's lambda_lj; d_charge_B_A may be NULL for
 * unperturbed charges), i.e. the inputs SPONGE_FEP / SPONGE_TI would read at
 * that lambda. h_out[w * SPONGE_FEP_NFIELDS + field] for
 * 1 <= nwin <= SPONGE_FEP_MAX_WINDOWS. At h_lambda[w] == lambda_lj the
 * fields equal the entry points above. */
int sponge_lj_soft_lambda_windows(sponge_lj* lj, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                  const float* d_charge, const float* d_charge_B_A, float pme_beta,
                                  int nwin, const double* h_lambda, double* h_out, void* stream);

/* ------------------------------------------------------------------ */
/* Charge-perturbed PME terms (TI_core/TI_core.cu, FEP_core/FEP_core.cu). */

/* TI_CORE::Get_Cross_PME_Partial_H_Partial_Lambda (TI_core.cu:670-739) with
 * kernels PME_Cross_Direct_Energy (:18-66) and
 * PME_Cross_Excluded_Energy_Correction (:69-125): the lambda derivatives of
 * the PME energy parts when the charges move along d_charge_B_A:
 * h_out = {reciprocal, self, correction, direct} (cross_reciprocal_ene,
 * cross_self_ene, cross_correction_ene, cross_direct_ene; the reference's
 * pme_dH_dlambda is the sum of the first three). direct is computed only
 * when lj_perturbed == 0 (else 0, as in the reference, whose soft-core
 * module supplies it). The self term's neutralising part is
 * 2 f (sum q)(sum dq) with f = -pi / (2 beta^2 V); the reference multiplies
 * f (sum q) by a sum of dq it never fills (TI_core.cu:663, :704), which
 * agrees when the system is neutral. The reciprocal term is
 * sum_i dq_i phi(r_i) with phi the convolved grid interpolated at the atom,
 * equal to the reference's sum over the grid of Q_dq . theta*Q. */
int sponge_pme_cross_dhdl(sponge_pme* pme, const sponge_nl* nl, const unsigned int* d_uint_crd,
                          const float* d_charge, const float* d_charge_B_A, const float scaler[3],
                          const int* d_excluded_list_start, const int* d_excluded_list,
                          const int* d_excluded_numbers, int lj_perturbed, double h_out[4],
                          void* stream);

/* PME reciprocal, self and correction energies at every lambda window (no
 * counterpart in the reference): the three parts are quadratic in the
 * charges, E(t) = E(q) + t cross + t^2 E(dq) for charges q + t dq with
 * t = h_lambda[w] - lambda0, from one pipeline on q, one on dq and the
 * cross terms of sponge_pme_cross_dhdl. h_out[w * 3 + {0, 1, 2}] =
 * {reciprocal, self, correction}; the reference's coul_long_range
 * (main_fep.cu:115-126) is their sum. */
int sponge_pme_lambda_windows(sponge_pme* pme, const unsigned int* d_uint_crd, const float* d_charge,
                              const float* d_charge_B_A, const float scaler[3],
                              const int* d_excluded_list_start, const int* d_excluded_list,
                              const int* d_excluded_numbers, float lambda0, int nwin,
                              const double* h_lambda, double* h_out, void* stream);

/* FEP_CORE::Seperate_Direct_Atom_Energy (FEP_core.cu:545-558, kernel :12-66):
 * PME direct energy split by subsystem mask when no soft-core module runs:
 * h_out = {intersys, intrasys} (coul_direct_intersys_ene, _intrasys).
 * Cutoff and beta are the PME handle's; d_subsys_division may be NULL
 * (every mask 0). */
int sponge_fep_separate_direct_energy(sponge_pme* pme, const sponge_nl* nl, const unsigned int* d_uint_crd,
                                      const float* d_charge, const int* d_subsys_division,
                                      double h_out[2], void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SPONGE_FEP_H */
