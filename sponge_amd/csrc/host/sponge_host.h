// C++ host mirror of SPONGE's nonbonded modules over the C ABI of
// libsponge_nb (include/sponge_nb.h).
//
// The three structs keep the reference's names, public members and method
// signatures (neighbor_list/neighbor_list.cuh:44-105,
// Lennard_Jones_force/Lennard_Jones_force.cuh:37-149, PME_force/PME_force.cuh:25-144)
// so that an MD step loop written against SPONGE compiles against them. The
// one type change: NEIGHBOR_LIST::d_nl is the opaque list handle instead of
// an ATOM_GROUP array (both of its consumers are in this library). CONTROLLER
// restates the mdin reader of control.cu for the keys this path uses.
#pragma once
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../../include/sponge_fep.h"
#include "../../../include/sponge_md.h"
#include "../../../include/sponge_nb.h"

struct VECTOR {
    float x, y, z;
};
struct UNSIGNED_INT_VECTOR {
    unsigned int uint_x, uint_y, uint_z;
};
// Lennard_Jones_force.cuh:26-33
struct UINT_VECTOR_LJ_TYPE {
    unsigned int uint_x, uint_y, uint_z;
    int LJ_type;
    float charge;
};

// control.cu:52-164, 176-407 (mdin file: first line = task name; `key =
// value` entries separated by newlines or commas; `prefix { ... }` blocks
// name keys prefix_key (`main` and `comments` blocks are special); `##prefix
// { ... }` blocks and `#` lines are comments; a value is its first
// whitespace-separated token; `-key value` on the command line, applied
// before the file; a key set twice is an error; `*_in_file` keys fall back
// to default_in_file_prefix).
class CONTROLLER {
  public:
    std::map<std::string, std::string> commands;
    std::map<std::string, std::string> original_commands;
    std::string md_name;

    // Returns false (and prints) on a malformed input instead of exiting.
    bool Initial(int argc, char** argv);
    bool Commands_From_In_File(const char* path);
    bool Set_Command(const char* flag, const char* value, const char* prefix = nullptr);
    bool Command_Exist(const char* key);
    bool Command_Exist(const char* prefix, const char* key);
    const char* Command(const char* key) const;
    const char* Command(const char* prefix, const char* key) const;
    bool Command_Choice(const char* key, const char* value) const;
    void printf(const char* fmt, ...) const;

    // mdout columns (control.cu:490-573): a module registers a column with
    // its printf format once (Step_Print_Initial); Step_Print fills it on a
    // print step; Print_To_Screen_And_Mdout writes the row (`%12s ` per
    // column to the mdout file, `%12s = %12s, ` three a line to the screen)
    // and resets every column to `****`. The mdout file (key `mdout`,
    // default mdout.txt) opens with the header line, so a run that only
    // parses its input writes no file.
    std::vector<std::string> outputs_key;
    std::map<std::string, std::string> outputs_format, outputs_content;
    FILE* mdout = nullptr;
    void Step_Print_Initial(const char* head, const char* format);
    void Step_Print(const char* head, float value);
    void Step_Print(const char* head, double value);
    void Step_Print(const char* head, int value);
    void Print_First_Line_To_Mdout(FILE* fp = nullptr);
    void Print_To_Screen_And_Mdout(FILE* fp = nullptr);
    void Clear();
};

struct NEIGHBOR_LIST {
    enum NEIGHBOR_LIST_UPDATE_PARAMETER { CONDITIONAL_UPDATE = 0, FORCED_UPDATE = 1 };
    enum NEIGHBOR_LIST_CHECK_PARAMETER { CONDITIONAL_CHECK = 0, FORCED_CHECK = 1 };
    char module_name[256];
    int is_initialized = 0;
    int atom_numbers = 0;
    int refresh_interval = 0;
    float skin = 2.0f;
    float cutoff = 10.0f;
    float skin_permit = 1.0f;
    int max_neighbor_numbers = 800;
    int max_atom_in_grid_numbers = 64;
    VECTOR box_length{0, 0, 0};
    sponge_nl* d_nl = nullptr;

    // neighbor_list.cu:580-642; mdin keys <module>_refresh_interval,
    // _max_atom_in_grid_numbers, _max_neighbor_numbers, _skin_permit.
    void Initial(CONTROLLER* controller, int md_atom_numbers, VECTOR box_length, float cut, float skin,
                 const char* module_name = nullptr);
    void Neighbor_List_Update(VECTOR* crd, int* d_excluded_list_start, int* d_excluded_list, int* d_excluded_numbers,
                              int forced_update = 0, int forced_check = 0);
    void Update_Volume(VECTOR box_length);
    void Clear();
};

struct LENNARD_JONES_INFORMATION {
    char module_name[256];
    int is_initialized = 0;
    int atom_numbers = 0;
    int atom_type_numbers = 0;
    int pair_type_numbers = 0;
    int* h_atom_LJ_type = nullptr;  // public for nb14 (main.cu:81)
    float* h_LJ_A = nullptr;        // x12
    float* h_LJ_B = nullptr;        // x6
    float cutoff = 10.0f;
    float long_range_factor = 0.0f;
    float h_LJ_energy_sum = 0.0f;
    sponge_lj* handle = nullptr;
    int is_controller_printf_initialized = 0;  // its mdout columns are registered

    // Lennard_Jones_force.cu:1067-1139: LJ_in_file (`N ntypes`, A x T(T+1)/2,
    // B x T(T+1)/2, N types).
    void Initial(CONTROLLER* controller, float cutoff, VECTOR box_length, const char* module_name = nullptr);
    void LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(const int atom_numbers, const UNSIGNED_INT_VECTOR* uint_crd,
                                                         const float* charge, VECTOR* frc, const sponge_nl* nl,
                                                         const float pme_beta, const int need_atom_energy,
                                                         float* atom_energy, const int need_virial,
                                                         float* atom_lj_virial, float* atom_direct_pme_energy);
    // Lennard_Jones_force.cuh:121-124, both overloads (force only)
    void LJ_Force_With_PME_Direct_Force(const int atom_numbers, const UINT_VECTOR_LJ_TYPE* uint_crd,
                                        const VECTOR scaler, VECTOR* frc, const sponge_nl* nl, const float cutoff,
                                        const float pme_beta);
    void LJ_Force_With_PME_Direct_Force(const UNSIGNED_INT_VECTOR* uint_crd, const float* charge, VECTOR* frc,
                                        const sponge_nl* nl, const float pme_beta);
    void Long_Range_Correction(int need_pressure, float* d_virial, int need_potential, float* d_potential);
    float Get_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const sponge_nl* nl, int is_download = 1);
    void Update_Volume(VECTOR box_length);
    void Clear();
};

struct Particle_Mesh_Ewald {
    enum PME_ENERGY_PART { TOTAL = 0, DIRECT = 1, RECIPROCAL = 2, CORRECTION = 3, SELF = 4 };
    char module_name[256];
    int is_initialized = 0;
    int atom_numbers = 0;
    int fftx = -1, ffty = -1, fftz = -1;
    float beta = 0.0f;
    float cutoff = 10.0f;
    float tolerance = 1e-5f;
    VECTOR boxlength{0, 0, 0};
    float* d_direct_atom_energy = nullptr;      // device, filled by the LJ direct call
    float* d_correction_atom_energy = nullptr;  // device
    float reciprocal_ene = 0, self_ene = 0, direct_ene = 0, correction_ene = 0, ee_ene = 0;
    sponge_pme* handle = nullptr;
    int is_controller_printf_initialized = 0;  // its mdout columns are registered

    // PME_force.cu:108-306; mdin keys <module>_Direct_Tolerance, _fftx/y/z.
    void Initial(CONTROLLER* controller, int atom_numbers, VECTOR boxlength, float cutoff,
                 const char* module_name = nullptr);
    void PME_Excluded_Force_With_Atom_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const VECTOR sacler,
                                             const float* charge, const int* excluded_list_start,
                                             const int* excluded_list, const int* excluded_atom_numbers,
                                             VECTOR* frc, float* atom_energy);
    void PME_Reciprocal_Force_With_Energy_And_Virial(const UNSIGNED_INT_VECTOR* uint_crd, const float* charge,
                                                     VECTOR* force, int need_virial, int need_energy,
                                                     float* d_virial, float* d_potential);
    float Get_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const float* charge, const sponge_nl* nl,
                     const VECTOR scaler, const int* excluded_list_start, const int* excluded_list,
                     const int* excluded_atom_numbers, int which_part = 0, int is_download = 1);
    void Update_Volume(VECTOR box_length);
    void Update_Box_Length(VECTOR boxlength);
    void Clear();
};

// LJ_SOFT_CORE (Lennard_Jones_force/LJ_soft_core.cuh:37-160): Initial reads
// <module>_in_file (`N TA TB`, AA and AB of the TA(TA+1)/2 A pairs, BA and BB
// of the B pairs before the x12 / x6 scaling, then per atom `typeA typeB`),
// subsys_division_in_file, lambda_lj (required), soft_core_alpha (0.5),
// soft_core_powfer (1), soft_core_sigma (3), soft_core_sigma_min (0) and
// PME_Direct_Tolerance for its own beta (LJ_soft_core.cu:1420-1610).
struct LJ_SOFT_CORE {
    char module_name[256];
    int is_initialized = 0;
    int atom_numbers = 0;
    int atom_type_numbers_A = 0, atom_type_numbers_B = 0;
    float lambda = 0.f, alpha = 0.5f, p = 1.f, sigma = 3.f, sigma_min = 0.f;
    float cutoff = 10.0f;
    float pme_beta = 0.f;
    float h_LJ_energy_sum = 0.f;
    float h_LJ_energy_sum_intersys = 0.f, h_LJ_energy_sum_intrasys = 0.f;
    float h_direct_ene_sum_intersys = 0.f, h_direct_ene_sum_intrasys = 0.f;
    float h_sigma_of_dH_dlambda_lj = 0.f, h_sigma_of_dH_dlambda_direct = 0.f;
    std::vector<int> h_subsys_division;
    int* d_subsys_division = nullptr;
    sponge_lj* handle = nullptr;

    void Initial(CONTROLLER* controller, float cutoff, VECTOR box_length, const char* module_name = nullptr);
    // LJ_soft_core.cu:1809-1852; returns the LJ total and fills the four split sums
    float Get_Energy_With_Coulomb_Direct(const UNSIGNED_INT_VECTOR* uint_crd, const sponge_nl* nl,
                                         const float* charge, int is_download = 1);
    // LJ_soft_core.cu:1854-1894; returns the LJ sum, the direct one is kept
    float Get_Partial_H_Partial_Lambda_With_Columb_Direct(const UNSIGNED_INT_VECTOR* uint_crd, const float* charge,
                                                          const sponge_nl* nl, const float* charge_B_A,
                                                          const int charge_pertubated, int is_download = 1);
    float Long_Range_Correction();                                 // :1927-1937
    float Partial_H_Partial_Lambda_Long_Range_Correction();        // :1939-1955
    void Update_Volume(VECTOR box_length);                         // :1896-1902
    void Clear();
};

// Failure of a library call: prints sponge_nb_last_error() and exits with
// status 1, the reference's behaviour for fatal input errors without the
// interactive getchar() pause.
void sponge_host_check(int rc, const char* what);

// ---- the step on either side of the force path --------------------------

// constrain/constrain.cuh: one distance constraint, the reference's layout.
struct CONSTRAIN_PAIR {
    int atom_i_serial;
    int atom_j_serial;
    float constant_r;
    float constrain_k;
};

// CONSTRAIN (constrain/constrain.cuh) for a host without SPONGE's own: the
// pairs of <module>_in_file (a count, then `i j r` lines) and the thermostat
// factors (constrain/constrain.cu:4-80). A SPONGE build keeps its own
// CONSTRAIN, which SETTLE::Initial accepts as well.
struct CONSTRAIN {
    char module_name[256] = "constrain";
    int atom_numbers = 0;
    float dt = 0.f, dt_inverse = 0.f;
    float v_factor = 1.f, x_factor = 1.f;
    int constrain_pair_numbers = 0;
    std::vector<CONSTRAIN_PAIR> pairs;
    CONSTRAIN_PAIR* h_constrain_pair = nullptr;
    void Initial_Constrain(CONTROLLER* controller, int atom_numbers, float dt, VECTOR box_length, float exp_gamma,
                           int is_Minimization, const float* atom_mass, int* system_freedom);
};

// SETTLE (constrain/SETTLE.cuh:37-67) over sponge_settle_*: rigid triangles
// and pairs found among the constraint pairs; the pairs it takes over are
// removed from the constraint's host list, as SETTLE.cu:338-360 does (a
// build that also runs SIMPLE_CONSTRAIN re-uploads that list).
struct SETTLE {
    char module_name[256] = "settle";
    int is_initialized = 0;
    int triangle_numbers = 0;
    int pair_numbers = 0;
    sponge_settle* handle = nullptr;

    template <class CONSTRAIN_T>
    void Initial(CONTROLLER* controller, CONSTRAIN_T* constrain, float* h_mass, const char* module_name = nullptr) {
        if (module_name) std::snprintf(this->module_name, sizeof(this->module_name), "%s", module_name);
        const int np = constrain->constrain_pair_numbers;
        if (np <= 0) {
            controller->printf("SETTLE IS NOT INITIALIZED\n\n");
            return;
        }
        controller->printf("START INITIALIZING SETTLE:\n");
        std::vector<int> ai(np), aj(np);
        std::vector<float> r(np);
        for (int k = 0; k < np; k++) {
            ai[k] = constrain->h_constrain_pair[k].atom_i_serial;
            aj[k] = constrain->h_constrain_pair[k].atom_j_serial;
            r[k] = constrain->h_constrain_pair[k].constant_r;
        }
        sponge_host_check(sponge_settle_create(constrain->atom_numbers, np, ai.data(), aj.data(), r.data(), h_mass,
                                               constrain->dt, constrain->v_factor, 0, &handle),
                          "SETTLE::Initial");
        sponge_host_check(sponge_settle_set_factors(handle, constrain->v_factor, constrain->x_factor),
                          "SETTLE factors");
        int rest = 0;
        sponge_host_check(sponge_settle_counts(handle, &triangle_numbers, &pair_numbers, &rest), "SETTLE counts");
        controller->printf("    rigid triangle numbers is %d\n", triangle_numbers);
        controller->printf("    rigid pair numbers is %d\n", pair_numbers);
        if (triangle_numbers + pair_numbers == 0) {
            controller->printf("SETTLE IS NOT INITIALIZED\n\n");
            return;
        }
        // keep the constraints whose first atom is in no rigid group
        std::vector<char> grouped(constrain->atom_numbers, 0);
        std::vector<int> tri(3 * (size_t)triangle_numbers), prs(2 * (size_t)pair_numbers);
        sponge_host_check(sponge_settle_groups(handle, tri.data(), nullptr, prs.data(), nullptr), "SETTLE groups");
        for (int a : tri) grouped[a] = 1;
        for (int a : prs) grouped[a] = 1;
        int kept = 0;
        for (int k = 0; k < np; k++)
            if (!grouped[constrain->h_constrain_pair[k].atom_i_serial])
                constrain->h_constrain_pair[kept++] = constrain->h_constrain_pair[k];
        constrain->constrain_pair_numbers = kept;
        controller->printf("    remaining simple constrain pair numbers is %d\n", kept);
        is_initialized = 1;
        controller->printf("END INITIALIZING SETTLE\n\n");
    }
    // SETTLE.cu:494-505; scaler = box_length / 2^32
    void Remember_Last_Coordinates(UNSIGNED_INT_VECTOR* uint_crd, VECTOR scaler);
    // SETTLE.cu:507-525
    void Do_SETTLE(const float* d_mass, VECTOR* crd, VECTOR box_length, VECTOR* vel, int need_pressure,
                   float* d_pressure);
    void Clear();
};

// MD_INFORMATION::NVE_iteration::Leap_Frog (MD_core/MD_core.cu:1535-1547)
// on MD_INFORMATION's arrays.
void Leap_Frog(int atom_numbers, VECTOR* vel, VECTOR* crd, VECTOR* frc, VECTOR* acc, const float* d_mass_inverse,
               float dt, float max_velocity);

// MD_INFORMATION::system_information::Get_Pressure (MD_core/MD_core.cu:
// 1777-1795) on MD_INFORMATION's arrays: *d_virial += sum of d_atom_virial,
// *d_pressure = (2 Ek + *d_virial) / (3 V); returns the pressure when
// is_download (kcal/mol/A^3).
float Get_Pressure(int atom_numbers, const VECTOR* vel, const float* d_mass, const float* d_atom_virial,
                   float* d_virial, VECTOR box_length, float* d_pressure, int is_download = 1);

// BERENDSEN_BAROSTAT_INFORMATION (barostats/Berendsen_barostat.cuh): Initial
// reads dt (ps), berendsen_barostat_tau (1 ps), _compressibility (4.5e-5
// bar^-1), _stochastic_term (0), _update_interval (10, or 1 with the
// stochastic term), _random_seed (Berendsen_barostat.cu:5-77).
// Scale_Factor is the step of main.cu:336-344 for the pressure p_now;
// vel_factor is 1 / mu with the stochastic term, else 0 (no velocity scaling).
struct BERENDSEN_BAROSTAT_INFORMATION {
    char module_name[256];
    int is_initialized = 0;
    float dt = 1e-3f, taup = 1.0f, compressibility = 4.5e-5f;
    int stochastic_term = 0, update_interval = 10;
    float V0 = 0.f, newV = 0.f;
    float target_temperature = 300.f;
    unsigned long long rng_state = 0;
    int is_controller_printf_initialized = 0;  // its mdout columns are registered
    void Initial(CONTROLLER* controller, float target_pressure, VECTOR box_length, const char* module_name = nullptr);
    void Ask_For_Calculate_Pressure(int steps, int* need_pressure);
    float Scale_Factor(float p_now, float target_pressure, float* vel_factor);
};

// ANDERSEN_BAROSTAT_INFORMATION (barostats/andersen_barostat.cuh): Initial
// reads andersen_barostat_tau (1 ps), _compressibility (4.5e-5 bar^-1) and
// _dV/dt (0) (andersen_barostat.cu:5-57); Scale_Factor is main.cu:358-362.
struct ANDERSEN_BAROSTAT_INFORMATION {
    char module_name[256];
    int is_initialized = 0;
    float V0 = 0.f, new_V = 0.f, h_mass_inverse = 0.f, dV_dt = 0.f;
    int is_controller_printf_initialized = 0;  // its mdout columns are registered
    void Initial(CONTROLLER* controller, float target_pressure, VECTOR box_length, const char* module_name = nullptr);
    float Scale_Factor(float p_now, float target_pressure, float volume, float dt);
};
