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

#include "../../../include/sponge_nb.h"

struct VECTOR {
    float x, y, z;
};
struct UNSIGNED_INT_VECTOR {
    unsigned int uint_x, uint_y, uint_z;
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

    // Lennard_Jones_force.cu:1067-1139: LJ_in_file (`N ntypes`, A x T(T+1)/2,
    // B x T(T+1)/2, N types).
    void Initial(CONTROLLER* controller, float cutoff, VECTOR box_length, const char* module_name = nullptr);
    void LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(const int atom_numbers, const UNSIGNED_INT_VECTOR* uint_crd,
                                                         const float* charge, VECTOR* frc, const sponge_nl* nl,
                                                         const float pme_beta, const int need_atom_energy,
                                                         float* atom_energy, const int need_virial,
                                                         float* atom_lj_virial, float* atom_direct_pme_energy);
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

// Failure of a library call: prints sponge_nb_last_error() and exits with
// status 1, the reference's behaviour for fatal input errors without the
// interactive getchar() pause.
void sponge_host_check(int rc, const char* what);
