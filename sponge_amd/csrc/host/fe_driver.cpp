// sponge_fe_driver: SPONGE's free-energy post-processors over a trajectory,
// `sponge_fe_driver fep [-mdin mdin.txt]` (main_fep.cu:16-29, 59-134, 161-172)
// and `sponge_fe_driver ti [...]` (main_ti.cu:18-31, 65-157, 196-201), with
// the nonbonded terms on the GPU through the mirror classes. Bonded terms
// (bond, angle, dihedral, nb14, bond_soft) are not part of this library and
// are reported as 0, as the reference does for modules without input.
//
// Inputs (FEP_core.cu:278-465, TI_core.cu:127-567): atom_numbers,
// frame_numbers (1000), crd (float32 frames, as mdcrd.dat), box (one line
// `Lx Ly Lz a b c` per frame), charge_pertubated (0), charge_in_file,
// chargeA_in_file / chargeB_in_file (TI with perturbed charges),
// subsys_division_in_file, exclude_in_file, cutoff (10), skin (2),
// target_pressure (1 bar), target_temperature (300), LJ_soft_core_in_file
// and its keys, PME_*. Outputs: one mdout line per frame; FEP appends the
// partition_energy_data records (13 float32 per frame, FEP_core.cuh:23-40) to
// float32ene (float32ene.dat) after the last frame (FEP_core.cu:539-543); TI
// prints the ensemble average of dH/dlambda and appends it to TI (TI.txt)
// (TI_core.cu:652-657).
//
// Extension of this driver (no counterpart in the reference): with
// fep_lambda_windows = N (2..64) the FEP run also evaluates every frame at
// the N windows lambda = k / (N - 1) in one pass per frame (charges move from
// chargeA to chargeB when both files are given) and writes float64 records
// [frame][window][12] to fep_windows (fep_windows.dat): the 9 fields of
// sponge_lj_soft_lambda_windows followed by PME reciprocal, self, correction.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sponge_host.h"

static void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "Error: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

static FILE* open_or_die(const char* path, const char* mode) {
    FILE* fp = std::fopen(path, mode);
    if (!fp) {
        std::fprintf(stderr, "Error: cannot open %s\n", path);
        std::exit(1);
    }
    return fp;
}

// `N`, then N values (charge_in_file / chargeA / chargeB / subsys_division).
template <class T>
static std::vector<T> read_list(const char* path, int n, const char* fmt) {
    FILE* fp = open_or_die(path, "r");
    int m = 0;
    if (std::fscanf(fp, "%d", &m) != 1 || m != n) {
        std::fprintf(stderr, "Error: atom_numbers is not equal in %s: %d %d\n", path, n, m);
        std::exit(1);
    }
    std::vector<T> v(n);
    for (int i = 0; i < n; i++)
        if (std::fscanf(fp, fmt, &v[i]) != 1) {
            std::fprintf(stderr, "Error: %s ends early\n", path);
            std::exit(1);
        }
    std::fclose(fp);
    return v;
}

template <class T>
static T* to_device(const std::vector<T>& h) {
    T* d = nullptr;
    size_t bytes = (h.empty() ? 1 : h.size()) * sizeof(T);
    hip_ok(hipMalloc((void**)&d, bytes), "hipMalloc");
    if (!h.empty()) hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

// FEP_CORE::partition_energy_data (FEP_core/FEP_core.cuh:23-40)
struct partition_energy_data {
    float bond_ene, angle_ene, dihedral_ene, nb14_LJ_ene, nb14_EE_ene;
    float bond_soft_ene, vdw_intersys_ene, vdw_intrasys_ene, coul_direct_intersys_ene, coul_direct_intrasys_ene;
    float vdw_long_range_correction, coul_long_range;
    float pV;
};

int main(int argc, char** argv) {
    bool ti = false;
    if (argc >= 2 && (std::strcmp(argv[1], "ti") == 0 || std::strcmp(argv[1], "TI") == 0)) ti = true;
    else if (!(argc >= 2 && (std::strcmp(argv[1], "fep") == 0 || std::strcmp(argv[1], "FEP") == 0))) {
        std::fprintf(stderr, "usage: %s fep|ti [-mdin mdin.txt] [-key value ...]\n", argv[0]);
        return 1;
    }
    CONTROLLER controller;
    if (!controller.Initial(argc, argv)) return 1;
    controller.printf("%s\n", controller.md_name.c_str());
    if (!controller.Command_Exist("atom_numbers")) {
        std::fprintf(stderr, "Error: missing value of atom numbers.\n");
        return 1;
    }
    const int n = std::atoi(controller.Command("atom_numbers"));
    int frame_numbers = 1000;
    if (controller.Command_Exist("frame_numbers")) frame_numbers = std::atoi(controller.Command("frame_numbers"));
    int charge_pertubated = 0;
    if (controller.Command_Exist("charge_pertubated"))
        charge_pertubated = std::atoi(controller.Command("charge_pertubated"));
    float cutoff = 10.0f, skin = 2.0f;
    if (controller.Command_Exist("cutoff")) cutoff = (float)std::atof(controller.Command("cutoff"));
    if (controller.Command_Exist("skin")) skin = (float)std::atof(controller.Command("skin"));
    if (!controller.Command_Exist("crd") || !controller.Command_Exist("box")) {
        std::fprintf(stderr, "Error: missing trajectory file (crd) or box trajectory file (box).\n");
        return 1;
    }
    FILE* crd_traj = open_or_die(controller.Command("crd"), "rb");
    FILE* box_traj = open_or_die(controller.Command("box"), "r");

    std::vector<float> charge(n, 0.f), charge_B_A(n, 0.f);
    if (controller.Command_Exist("charge_in_file")) charge = read_list<float>(controller.Command("charge_in_file"), n, "%f");
    const bool have_AB = controller.Command_Exist("chargeA_in_file") && controller.Command_Exist("chargeB_in_file");
    if (ti && charge_pertubated > 0 && !have_AB) {
        std::fprintf(stderr, "Error: missing value of charge A and charge B, These value must be given in TI mode "
                             "if charge is pertubated.\n");
        return 1;
    }
    std::vector<float> qA, qB;
    if (have_AB) {
        qA = read_list<float>(controller.Command("chargeA_in_file"), n, "%f");
        qB = read_list<float>(controller.Command("chargeB_in_file"), n, "%f");
        for (int i = 0; i < n; i++) charge_B_A[i] = qB[i] - qA[i];  // TI_core.cu:459-462
    }
    std::vector<int> ex_start(n, 0), ex_numbers(n, 0), ex_list;
    if (controller.Command_Exist("exclude_in_file")) {
        FILE* fp = open_or_die(controller.Command("exclude_in_file"), "r");
        int m = 0, total = 0;
        if (std::fscanf(fp, "%d %d", &m, &total) != 2 || m != n) {
            std::fprintf(stderr, "Error: atom_numbers is not equal in exclude_in_file\n");
            return 1;
        }
        for (int i = 0; i < n; i++) {
            int k = 0;
            if (std::fscanf(fp, "%d", &k) != 1) k = -1;
            ex_start[i] = (int)ex_list.size();
            ex_numbers[i] = k;
            for (int j = 0; j < k; j++) {
                int a = -1;
                if (std::fscanf(fp, "%d", &a) != 1 || a < 0 || a >= n) k = -1;
                ex_list.push_back(a);
            }
            if (k < 0) {
                std::fprintf(stderr, "Error: malformed exclude_in_file at atom %d\n", i);
                return 1;
            }
        }
        std::fclose(fp);
    }
    // data.pressure in kcal/mol/A^3 (FEP_core.cu:455-461, the bar conversion of common.cuh:26-27)
    float pressure = 1.0f;
    if (controller.Command_Exist("target_pressure")) pressure = (float)std::atof(controller.Command("target_pressure"));
    pressure *= 0.00001439506089041446f;
    int nwin = 0;
    if (!ti && controller.Command_Exist("fep_lambda_windows")) {
        nwin = std::atoi(controller.Command("fep_lambda_windows"));
        if (nwin < 2 || nwin > SPONGE_FEP_MAX_WINDOWS) {
            std::fprintf(stderr, "Error: fep_lambda_windows must be in 2..%d\n", SPONGE_FEP_MAX_WINDOWS);
            return 1;
        }
    }

    // Read_Next_Frame (FEP_core.cu:473-485 / TI_core.cu:575-592)
    std::vector<VECTOR> coordinate(n);
    VECTOR box{1.f, 1.f, 1.f};
    auto read_frame = [&]() {
        if (std::fread(coordinate.data(), sizeof(VECTOR), n, crd_traj) != (size_t)n) {
            std::fprintf(stderr, "Error: the coordinate trajectory ends early\n");
            std::exit(1);
        }
        float a[3];
        if (std::fscanf(box_traj, "%f %f %f %f %f %f", &box.x, &box.y, &box.z, &a[0], &a[1], &a[2]) != 6) {
            std::fprintf(stderr, "Error: the box trajectory ends early\n");
            std::exit(1);
        }
    };
    read_frame();

    VECTOR* crd = to_device(coordinate);
    float* d_charge = to_device(charge);
    float* d_charge_B_A = to_device(charge_B_A);
    int* d_es = to_device(ex_start);
    int* d_en = to_device(ex_numbers);
    int* d_el = to_device(ex_list);
    UNSIGNED_INT_VECTOR* uint_crd = nullptr;
    hip_ok(hipMalloc((void**)&uint_crd, n * sizeof(UNSIGNED_INT_VECTOR)), "hipMalloc");
    auto crd_to_uint = [&]() {
        const float b[3] = {box.x, box.y, box.z};
        sponge_host_check(sponge_crd_to_uint(n, b, reinterpret_cast<const float*>(crd),
                                             reinterpret_cast<unsigned*>(uint_crd), nullptr),
                          "Crd_To_Uint_Crd");
    };
    crd_to_uint();

    LJ_SOFT_CORE lj_soft;
    Particle_Mesh_Ewald pme;
    NEIGHBOR_LIST neighbor_list;
    lj_soft.Initial(&controller, cutoff, box);
    if (charge_pertubated) pme.Initial(&controller, n, box, cutoff);
    const bool need_nl = charge_pertubated || lj_soft.is_initialized;
    if (need_nl) {
        neighbor_list.Initial(&controller, n, box, cutoff, skin);
        neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en, NEIGHBOR_LIST::FORCED_UPDATE);
    }
    if (nwin && !lj_soft.is_initialized) {
        std::fprintf(stderr, "Error: fep_lambda_windows needs the soft-core module (LJ_soft_core_in_file)\n");
        return 1;
    }
    int* d_mask = lj_soft.is_initialized ? lj_soft.d_subsys_division : nullptr;
    std::vector<int> mask_h;
    if (!lj_soft.is_initialized && controller.Command_Exist("subsys_division_in_file")) {
        mask_h = read_list<int>(controller.Command("subsys_division_in_file"), n, "%d");
        d_mask = to_device(mask_h);
    }
    // the window evaluation needs the charges of lambda_lj and their direction
    float* d_window_charge = d_charge;
    float* d_window_dq = nullptr;
    std::vector<float> q_lambda;
    if (nwin && have_AB) {
        q_lambda.resize(n);
        const double lam = lj_soft.lambda;
        for (int i = 0; i < n; i++) q_lambda[i] = (float)((1.0 - lam) * qA[i] + lam * qB[i]);
        d_window_charge = to_device(q_lambda);
        d_window_dq = d_charge_B_A;
    }
    FILE* win_out = nullptr;
    if (nwin) win_out = open_or_die(controller.Command_Exist("fep_windows") ? controller.Command("fep_windows")
                                                                            : "fep_windows.dat",
                                    "wb");

    std::vector<partition_energy_data> parts(frame_numbers);
    double total_dhdl = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    for (int frame = 1; frame <= frame_numbers; frame++) {
        const VECTOR scaler{box.x / 4294967296.0f, box.y / 4294967296.0f, box.z / 4294967296.0f};
        const float sc3[3] = {scaler.x, scaler.y, scaler.z};
        if (!ti) {
            // main_fep.cu:59-134 (nonbonded part)
            partition_energy_data& P = parts[frame - 1];
            std::memset(&P, 0, sizeof(P));
            float lj_soft_ene = 0.f;
            if (lj_soft.is_initialized) {
                lj_soft_ene = lj_soft.Get_Energy_With_Coulomb_Direct(uint_crd, neighbor_list.d_nl, d_charge);
                P.vdw_intersys_ene = lj_soft.h_LJ_energy_sum_intersys;
                P.vdw_intrasys_ene = lj_soft.h_LJ_energy_sum_intrasys;
                P.coul_direct_intersys_ene = lj_soft.h_direct_ene_sum_intersys;
                P.coul_direct_intrasys_ene = lj_soft.h_direct_ene_sum_intrasys;
                P.vdw_long_range_correction = lj_soft.Long_Range_Correction();
            }
            if (charge_pertubated) {
                P.coul_long_range = pme.Get_Energy(uint_crd, d_charge, neighbor_list.d_nl, scaler, d_es, d_el, d_en,
                                                   Particle_Mesh_Ewald::CORRECTION);
                P.coul_long_range += pme.Get_Energy(uint_crd, d_charge, neighbor_list.d_nl, scaler, d_es, d_el, d_en,
                                                    Particle_Mesh_Ewald::SELF);
                P.coul_long_range += pme.Get_Energy(uint_crd, d_charge, neighbor_list.d_nl, scaler, d_es, d_el, d_en,
                                                    Particle_Mesh_Ewald::RECIPROCAL);
                if (!lj_soft.is_initialized) {
                    double sep[2];
                    sponge_host_check(sponge_fep_separate_direct_energy(pme.handle, neighbor_list.d_nl,
                                                                        reinterpret_cast<unsigned*>(uint_crd),
                                                                        d_charge, d_mask, sep, nullptr),
                                      "Seperate_Direct_Atom_Energy");
                    P.coul_direct_intersys_ene = (float)sep[0];
                    P.coul_direct_intrasys_ene = (float)sep[1];
                }
            }
            P.pV = pressure * box.x * box.y * box.z;  // main_fep.cu:128
            const double ene = (double)P.vdw_intersys_ene + P.vdw_intrasys_ene + P.coul_direct_intersys_ene +
                               P.coul_direct_intrasys_ene + P.vdw_long_range_correction + P.coul_long_range + P.pV;
            controller.printf("frame %d ene %.2f LJ(sc.) %.2f Coul(direct.) %.2f LR_corr(sc.) %.2f PME(reci.) %.2f "
                              "PME(self.) %.2f PME(corr.) %.2f pV %.2f\n",
                              frame, ene, lj_soft_ene, P.coul_direct_intersys_ene + P.coul_direct_intrasys_ene,
                              P.vdw_long_range_correction, pme.reciprocal_ene, pme.self_ene, pme.correction_ene, P.pV);
            if (nwin) {
                std::vector<double> lam(nwin), pw((size_t)nwin * SPONGE_FEP_NFIELDS), qe((size_t)nwin * 3, 0.0);
                for (int k = 0; k < nwin; k++) lam[k] = (double)k / (nwin - 1);
                sponge_host_check(sponge_lj_soft_lambda_windows(lj_soft.handle, neighbor_list.d_nl,
                                                                reinterpret_cast<unsigned*>(uint_crd),
                                                                d_window_charge, d_window_dq, lj_soft.pme_beta, nwin,
                                                                lam.data(), pw.data(), nullptr),
                                  "lambda windows");
                if (charge_pertubated)
                    sponge_host_check(sponge_pme_lambda_windows(pme.handle, reinterpret_cast<unsigned*>(uint_crd),
                                                                d_window_charge, d_window_dq, sc3, d_es, d_el, d_en,
                                                                lj_soft.lambda, nwin, lam.data(), qe.data(), nullptr),
                                      "PME lambda windows");
                for (int k = 0; k < nwin; k++) {
                    double rec[SPONGE_FEP_NFIELDS + 3];
                    std::memcpy(rec, &pw[(size_t)k * SPONGE_FEP_NFIELDS], SPONGE_FEP_NFIELDS * sizeof(double));
                    std::memcpy(rec + SPONGE_FEP_NFIELDS, &qe[(size_t)k * 3], 3 * sizeof(double));
                    std::fwrite(rec, sizeof(double), SPONGE_FEP_NFIELDS + 3, win_out);
                }
            }
        } else {
            // main_ti.cu:65-157 (nonbonded part)
            float lj_dhdl = 0.f, coul_direct_dhdl = 0.f, lrc_dhdl = 0.f, pme_dhdl = 0.f;
            double cross[4] = {0, 0, 0, 0};
            if (lj_soft.is_initialized) {
                lj_dhdl = lj_soft.Get_Partial_H_Partial_Lambda_With_Columb_Direct(uint_crd, d_charge, neighbor_list.d_nl,
                                                                                  d_charge_B_A, charge_pertubated);
                coul_direct_dhdl = lj_soft.h_sigma_of_dH_dlambda_direct;
                lrc_dhdl = lj_soft.Partial_H_Partial_Lambda_Long_Range_Correction();
            }
            if (charge_pertubated) {
                sponge_host_check(sponge_pme_cross_dhdl(pme.handle, neighbor_list.d_nl,
                                                        reinterpret_cast<unsigned*>(uint_crd), d_charge, d_charge_B_A,
                                                        sc3, d_es, d_el, d_en, lj_soft.is_initialized, cross, nullptr),
                                  "Get_Cross_PME_Partial_H_Partial_Lambda");
                pme_dhdl = (float)(cross[0] + cross[1] + cross[2]);
                coul_direct_dhdl += (float)cross[3];
            }
            const float dhdl = lj_dhdl + coul_direct_dhdl + lrc_dhdl + pme_dhdl;  // Sum_One_Frame, TI_core.cu:640-645
            total_dhdl += dhdl;
            controller.printf("frame %d dH_dlambda %.6f LJ(sc.) %.6f Coul(direct.) %.6f LR_corr(sc.) %.6f "
                              "PME(reci.) %.6f PME(self.) %.6f PME(corr.) %.6f\n",
                              frame, dhdl, lj_dhdl, coul_direct_dhdl, lrc_dhdl, cross[0], cross[1], cross[2]);
        }
        if (frame != frame_numbers) {
            // Main_Iteration: next frame, Main_Volume_Update, forced list update
            read_frame();
            hip_ok(hipMemcpy(crd, coordinate.data(), n * sizeof(VECTOR), hipMemcpyHostToDevice), "hipMemcpy");
            crd_to_uint();
            lj_soft.Update_Volume(box);
            if (charge_pertubated) pme.Update_Volume(box);
            if (need_nl) {
                neighbor_list.Update_Volume(box);
                neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en, NEIGHBOR_LIST::FORCED_UPDATE);
            }
        }
    }
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ti) {
        FILE* fp = open_or_die(controller.Command_Exist("float32ene") ? controller.Command("float32ene")
                                                                       : "float32ene.dat",
                               "ab");
        std::fwrite(parts.data(), sizeof(partition_energy_data), frame_numbers, fp);
        std::fclose(fp);
    } else {
        const double avg = total_dhdl / frame_numbers;
        std::fprintf(stdout, "Ensemble Average <dH/dlambda>: %.6f\n", avg);
        FILE* fp = open_or_die(controller.Command_Exist("TI") ? controller.Command("TI") : "TI.txt", "a");
        std::fprintf(fp, "%.6f\n", avg);
        std::fclose(fp);
    }
    if (win_out) std::fclose(win_out);
    controller.printf("Core Run Wall Time: %f second(s), %f frames/s\n", secs, frame_numbers / secs);
    std::fclose(crd_traj);
    std::fclose(box_traj);
    lj_soft.Clear();
    pme.Clear();
    neighbor_list.Clear();
    return 0;
}
