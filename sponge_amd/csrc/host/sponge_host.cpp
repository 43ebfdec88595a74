// Implementation of the C++ host mirror (see sponge_host.h).
#include "sponge_host.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

void sponge_host_check(int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, sponge_nb_last_error());
        std::exit(1);
    }
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "Error: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

// ---------------------------------------------------------------- CONTROLLER
static std::string first_token(const std::string& v) {
    std::istringstream is(v);
    std::string t;
    is >> t;
    return t;
}

bool CONTROLLER::Set_Command(const char* flag, const char* value, const char* prefix) {
    if (prefix && std::strcmp(prefix, "comments") == 0) return true;
    std::string key;
    if (prefix && prefix[0] && std::strcmp(prefix, "main") != 0) key = std::string(prefix) + "_";
    key += flag;
    if (commands.count(key)) {
        std::fprintf(stderr, "Error: %s is set more than once.\n", key.c_str());
        return false;
    }
    original_commands[key] = value;
    commands[key] = first_token(value);
    return true;
}

// Statements end at any of "\n , { }"; the run of enders and blanks after a
// statement is its ender set (control.cu:229-278). `name {` opens a block
// whose keys are stored as name_key, `}` closes it, `#` starts a comment that
// runs to the end of the line and `##name {` opens a commented-out block
// (control.cu:294-341, Get_Command :205-225).
bool CONTROLLER::Commands_From_In_File(const char* path) {
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "Error: cannot open %s\n", path);
        return false;
    }
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t p = text.find('\n');
    md_name = text.substr(0, p);
    if (p == std::string::npos) return true;
    auto is_ender = [](char c) { return c == '\n' || c == ',' || c == '{' || c == '}'; };
    auto is_blank = [](char c) { return c == ' ' || c == '\t'; };
    std::string prefix;
    size_t i = p + 1;
    const size_t n = text.size();
    while (i < n) {
        std::string line, ender;
        while (i < n && line.empty() && is_blank(text[i])) i++;
        while (i < n && !is_ender(text[i])) line.push_back(text[i++]);
        if (i < n) ender.push_back(text[i++]);
        while (i < n && (is_blank(text[i]) || is_ender(text[i])))
            if (is_ender(text[i++])) ender.push_back(text[i - 1]);
        const bool opens = ender.find('{') != std::string::npos;
        const bool closes = ender.find('}') != std::string::npos;
        if (!line.empty() && line[0] == '#') {
            if (line.size() > 1 && line[1] == '#') {
                if (opens) prefix = first_token(line);
                if (closes) prefix.clear();
            }
            if (ender.find('\n') == std::string::npos) {  // the comment runs to the end of the line
                while (i < n && text[i] != '\n') i++;
            }
            continue;
        }
        if (opens) {
            prefix = first_token(line);
        } else if (!(prefix.size() > 1 && prefix[0] == '#' && prefix[1] == '#')) {
            size_t eq = line.find('=');
            if (eq != std::string::npos) {
                size_t eq2 = line.find('=', eq + 1);
                std::string flag = first_token(line.substr(0, eq));
                std::string value = line.substr(eq + 1, eq2 == std::string::npos ? std::string::npos : eq2 - eq - 1);
                if (!flag.empty() && !first_token(value).empty() &&
                    !Set_Command(flag.c_str(), value.c_str(), prefix.c_str()))
                    return false;
            }
        }
        if (closes) prefix.clear();
    }
    return true;
}

bool CONTROLLER::Initial(int argc, char** argv) {
    // control.cu:164-204: `-flag v1 v2 ...`; a flag is `-` followed by a non-digit
    auto is_flag = [](const char* a) { return std::strlen(a) > 1 && a[0] == '-' && !(a[1] >= '0' && a[1] <= '9'); };
    for (int i = 1; i < argc; i++) {
        if (!is_flag(argv[i])) continue;
        std::string v;
        for (int j = i + 1; j < argc && !is_flag(argv[j]); j++) {
            v += " ";
            v += argv[j];
        }
        if (!Set_Command(argv[i] + 1, v.c_str())) return false;
    }
    if (Command_Exist("mdin")) return Commands_From_In_File(Command("mdin"));
    std::ifstream probe("mdin.txt");
    if (probe) return Commands_From_In_File("mdin.txt");
    md_name = "Default SPONGE MD Task Name";
    return true;
}

// A missing `<name>_in_file` falls back to `<default_in_file_prefix>_<name>.txt`
// when that file exists (control.cu:52-87).
bool CONTROLLER::Command_Exist(const char* key) {
    if (commands.count(key)) return true;
    const std::string k(key), tail("_in_file");
    if (k.size() > tail.size() && k.compare(k.size() - tail.size(), tail.size(), tail) == 0 &&
        commands.count("default_in_file_prefix")) {
        std::string path = commands["default_in_file_prefix"] + "_" + k.substr(0, k.size() - tail.size()) + ".txt";
        if (FILE* f = std::fopen(path.c_str(), "r")) {
            std::fclose(f);
            commands[k] = path;
            return true;
        }
    }
    return false;
}
bool CONTROLLER::Command_Exist(const char* prefix, const char* key) {
    return Command_Exist((std::string(prefix) + "_" + key).c_str());
}
const char* CONTROLLER::Command(const char* key) const {
    auto it = commands.find(key);
    return it == commands.end() ? "" : it->second.c_str();
}
const char* CONTROLLER::Command(const char* prefix, const char* key) const {
    return Command((std::string(prefix) + "_" + key).c_str());
}
bool CONTROLLER::Command_Choice(const char* key, const char* value) const {
    auto it = commands.find(key);
    if (it == commands.end()) return false;
    const std::string& a = it->second;
    if (a.size() != std::strlen(value)) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)value[i])) return false;
    return true;
}
void CONTROLLER::printf(const char* fmt, ...) const {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stdout, fmt, ap);
    va_end(ap);
}

void CONTROLLER::Step_Print_Initial(const char* head, const char* format) {
    if (outputs_format.count(head)) return;
    outputs_format[head] = format;
    outputs_content[head] = "****";
    outputs_key.push_back(head);
}
template <class T>
static void step_print(CONTROLLER& c, const char* head, T value) {
    auto it = c.outputs_format.find(head);
    if (it == c.outputs_format.end()) return;  // a column nobody registered is ignored
    char buf[256];
    std::snprintf(buf, sizeof(buf), it->second.c_str(), value);
    c.outputs_content[head] = buf;
}
void CONTROLLER::Step_Print(const char* head, float value) { step_print(*this, head, (double)value); }
void CONTROLLER::Step_Print(const char* head, double value) { step_print(*this, head, value); }
void CONTROLLER::Step_Print(const char* head, int value) { step_print(*this, head, value); }
void CONTROLLER::Print_First_Line_To_Mdout(FILE* fp) {
    if (!fp) {
        if (!mdout) {
            const char* path = Command_Exist("mdout") ? Command("mdout") : "mdout.txt";
            mdout = std::fopen(path, "w");
            if (!mdout) {
                std::fprintf(stderr, "Error: cannot open %s\n", path);
                std::exit(1);
            }
        }
        fp = mdout;
    }
    for (const std::string& k : outputs_key) std::fprintf(fp, "%12s ", k.c_str());
    std::fprintf(fp, "\n");
    std::printf("---------------------------------------------------------------------------------------\n");
}
void CONTROLLER::Print_To_Screen_And_Mdout(FILE* fp) {
    if (!fp) fp = mdout;
    int n = 0;
    for (const std::string& k : outputs_key) {
        std::string& v = outputs_content[k];
        std::printf("%12s = %12s, ", k.c_str(), v.c_str());
        if (fp) std::fprintf(fp, "%12s ", v.c_str());
        v = "****";
        if (++n % 3 == 0) std::printf("\n");
    }
    if (n % 3 != 0) std::printf("\n");
    std::printf("---------------------------------------------------------------------------------------\n");
    if (fp) {
        std::fprintf(fp, "\n");
        std::fflush(fp);
    }
}
void CONTROLLER::Clear() {
    if (mdout) std::fclose(mdout);
    mdout = nullptr;
    outputs_key.clear();
    outputs_format.clear();
    outputs_content.clear();
}

// ------------------------------------------------------------- NEIGHBOR_LIST
void NEIGHBOR_LIST::Initial(CONTROLLER* controller, int md_atom_numbers, VECTOR box, float cut, float sk,
                            const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "neighbor_list");
    atom_numbers = md_atom_numbers;
    float permit_in_mdin = 0.5f;  // mdin value 0.5 = internal 1.0 (neighbor_list.cu:608-611)
    if (controller->Command_Exist(module_name, "refresh_interval"))
        refresh_interval = std::atoi(controller->Command(module_name, "refresh_interval"));
    if (controller->Command_Exist(module_name, "max_atom_in_grid_numbers"))
        max_atom_in_grid_numbers = std::atoi(controller->Command(module_name, "max_atom_in_grid_numbers"));
    if (controller->Command_Exist(module_name, "max_neighbor_numbers"))
        max_neighbor_numbers = std::atoi(controller->Command(module_name, "max_neighbor_numbers"));
    if (controller->Command_Exist(module_name, "skin_permit"))
        permit_in_mdin = (float)std::atof(controller->Command(module_name, "skin_permit"));
    skin_permit = 2.0f * permit_in_mdin;
    skin = sk;
    cutoff = cut;
    box_length = box;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_nl_create(atom_numbers, b, cutoff, skin, refresh_interval, max_atom_in_grid_numbers,
                                       max_neighbor_numbers, permit_in_mdin, &d_nl),
                      "NEIGHBOR_LIST::Initial");
    is_initialized = 1;
}

void NEIGHBOR_LIST::Neighbor_List_Update(VECTOR* crd, int* es, int* el, int* en, int forced_update, int forced_check) {
    if (!is_initialized) return;
    sponge_host_check(sponge_nl_update(d_nl, reinterpret_cast<float*>(crd), es, el, en, forced_update, forced_check,
                                       nullptr),
                      "Neighbor_List_Update");
}

void NEIGHBOR_LIST::Update_Volume(VECTOR box) {
    if (!is_initialized) return;
    box_length = box;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_nl_update_volume(d_nl, b), "NEIGHBOR_LIST::Update_Volume");
}

void NEIGHBOR_LIST::Clear() {
    if (is_initialized) {
        sponge_nl_destroy(d_nl);
        d_nl = nullptr;
        is_initialized = 0;
    }
}

// ------------------------------------------------- LENNARD_JONES_INFORMATION
static bool read_parm7_lj(const char* path, int* n, int* ntypes, std::vector<float>& A, std::vector<float>& B,
                          std::vector<int>& types) {
    // Lennard_Jones_force.cu:1201-1276: POINTERS (atoms, types),
    // ATOM_TYPE_INDEX (1-based), LENNARD_JONES_ACOEF / _BCOEF (x12 / x6).
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    auto read_block = [&](int count, auto&& sink) {
        std::getline(in, line);  // %FORMAT line
        int got = 0;
        while (got < count && std::getline(in, line)) {
            std::istringstream is(line);
            std::string tok;
            while (got < count && is >> tok) sink(got++, tok);
        }
    };
    while (std::getline(in, line)) {
        std::istringstream is(line);
        std::string a, b;
        if (!(is >> a >> b) || a != "%FLAG") continue;
        if (b == "POINTERS") {
            read_block(2, [&](int k, const std::string& t) {
                if (k == 0) *n = std::atoi(t.c_str());
                else *ntypes = std::atoi(t.c_str());
            });
            int P = *ntypes * (*ntypes + 1) / 2;
            A.assign(P, 0.f);
            B.assign(P, 0.f);
            types.assign(*n, 0);
        } else if (b == "ATOM_TYPE_INDEX") {
            read_block(*n, [&](int k, const std::string& t) { types[k] = std::atoi(t.c_str()) - 1; });
        } else if (b == "LENNARD_JONES_ACOEF") {
            read_block((int)A.size(), [&](int k, const std::string& t) { A[k] = (float)(12.0 * std::atof(t.c_str())); });
        } else if (b == "LENNARD_JONES_BCOEF") {
            read_block((int)B.size(), [&](int k, const std::string& t) { B[k] = (float)(6.0 * std::atof(t.c_str())); });
        }
    }
    return *n > 0 && *ntypes > 0;
}

void LENNARD_JONES_INFORMATION::Initial(CONTROLLER* controller, float cut, VECTOR box, const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "LJ");
    std::vector<float> A, B;
    std::vector<int> t;
    if (controller->Command_Exist(module_name, "in_file")) {
        FILE* fp = std::fopen(controller->Command(module_name, "in_file"), "r");
        if (!fp) {
            std::fprintf(stderr, "Error: cannot open %s\n", controller->Command(module_name, "in_file"));
            std::exit(1);
        }
        if (std::fscanf(fp, "%d %d", &atom_numbers, &atom_type_numbers) != 2) std::exit(1);
        pair_type_numbers = atom_type_numbers * (atom_type_numbers + 1) / 2;
        A.resize(pair_type_numbers);
        B.resize(pair_type_numbers);
        t.resize(atom_numbers);
        for (int i = 0; i < pair_type_numbers; i++) {
            if (std::fscanf(fp, "%f", &A[i]) != 1) std::exit(1);
            A[i] *= 12.0f;
        }
        for (int i = 0; i < pair_type_numbers; i++) {
            if (std::fscanf(fp, "%f", &B[i]) != 1) std::exit(1);
            B[i] *= 6.0f;
        }
        for (int i = 0; i < atom_numbers; i++)
            if (std::fscanf(fp, "%d", &t[i]) != 1) std::exit(1);
        std::fclose(fp);
    } else if (controller->Command_Exist("amber_parm7")) {
        if (!read_parm7_lj(controller->Command("amber_parm7"), &atom_numbers, &atom_type_numbers, A, B, t)) {
            std::fprintf(stderr, "Error: LJ information not found in %s\n", controller->Command("amber_parm7"));
            std::exit(1);
        }
        pair_type_numbers = atom_type_numbers * (atom_type_numbers + 1) / 2;
    } else {
        return;  // not initialised: every method is a no-op, as in the reference
    }
    cutoff = cut;
    h_LJ_A = new float[pair_type_numbers];
    h_LJ_B = new float[pair_type_numbers];
    h_atom_LJ_type = new int[atom_numbers];
    std::memcpy(h_LJ_A, A.data(), pair_type_numbers * sizeof(float));
    std::memcpy(h_LJ_B, B.data(), pair_type_numbers * sizeof(float));
    std::memcpy(h_atom_LJ_type, t.data(), atom_numbers * sizeof(int));
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_lj_create(atom_numbers, atom_type_numbers, h_LJ_A, h_LJ_B, h_atom_LJ_type, cutoff, b,
                                       &handle),
                      "LENNARD_JONES_INFORMATION::Initial");
    // long_range_factor (Lennard_Jones_force.cu:1115-1127), for callers that read it
    std::vector<double> cnt(atom_type_numbers, 0.0);
    for (int i = 0; i < atom_numbers; i++) cnt[t[i]] += 1;
    double s = 0;
    for (int x = 0; x < atom_type_numbers; x++)
        for (int y = 0; y < atom_type_numbers; y++) {
            int hi = x > y ? x : y, lo = x > y ? y : x;
            s += cnt[x] * cnt[y] * B[hi * (hi + 1) / 2 + lo];
        }
    long_range_factor = (float)(s * (-2.0 / 3.0) * M_PI / ((double)cut * cut * cut) / 6.0);
    is_initialized = 1;
    if (!is_controller_printf_initialized) {  // Lennard_Jones_force.cu:1132-1136
        controller->Step_Print_Initial(module_name, "%.2f");
        is_controller_printf_initialized = 1;
    }
}

void LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
    const int n, const UNSIGNED_INT_VECTOR* uint_crd, const float* charge, VECTOR* frc, const sponge_nl* nl,
    const float pme_beta, const int need_atom_energy, float* atom_energy, const int need_virial, float* atom_lj_virial,
    float* atom_direct_pme_energy) {
    if (!is_initialized) return;
    (void)n;
    sponge_host_check(sponge_lj_pme_direct_force(handle, nl, reinterpret_cast<const unsigned*>(uint_crd), charge,
                                                 reinterpret_cast<float*>(frc), pme_beta, need_atom_energy,
                                                 atom_energy, need_virial, atom_lj_virial, atom_direct_pme_energy,
                                                 nullptr),
                      "LJ_PME_Direct_Force_With_Atom_Energy_And_Virial");
}

void LENNARD_JONES_INFORMATION::LJ_Force_With_PME_Direct_Force(const int n, const UINT_VECTOR_LJ_TYPE* uint_crd,
                                                               const VECTOR scaler, VECTOR* frc, const sponge_nl* nl,
                                                               const float cutoff_, const float pme_beta) {
    if (!is_initialized) return;
    const float sc[3] = {scaler.x, scaler.y, scaler.z};
    sponge_host_check(sponge_lj_force_with_pme_direct_force_records(handle, nl, n, uint_crd, sc,
                                                                    reinterpret_cast<float*>(frc), cutoff_, pme_beta,
                                                                    nullptr),
                      "LJ_Force_With_PME_Direct_Force");
}

void LENNARD_JONES_INFORMATION::LJ_Force_With_PME_Direct_Force(const UNSIGNED_INT_VECTOR* uint_crd,
                                                               const float* charge, VECTOR* frc, const sponge_nl* nl,
                                                               const float pme_beta) {
    if (!is_initialized) return;
    sponge_host_check(sponge_lj_force_with_pme_direct_force(handle, nl, reinterpret_cast<const unsigned*>(uint_crd),
                                                            charge, reinterpret_cast<float*>(frc), pme_beta, nullptr),
                      "LJ_Force_With_PME_Direct_Force");
}

void LENNARD_JONES_INFORMATION::Long_Range_Correction(int need_pressure, float* d_virial, int need_potential,
                                                      float* d_potential) {
    if (!is_initialized) return;
    sponge_host_check(sponge_lj_long_range_correction(handle, need_pressure, d_virial, need_potential, d_potential,
                                                      nullptr),
                      "Long_Range_Correction");
}

float LENNARD_JONES_INFORMATION::Get_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const sponge_nl* nl, int) {
    if (!is_initialized) return NAN;
    double e = 0;
    sponge_host_check(sponge_lj_get_energy(handle, nl, reinterpret_cast<const unsigned*>(uint_crd), &e, nullptr),
                      "LJ Get_Energy");
    h_LJ_energy_sum = (float)e;
    return h_LJ_energy_sum;
}

void LENNARD_JONES_INFORMATION::Update_Volume(VECTOR box) {
    if (!is_initialized) return;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_lj_update_volume(handle, b), "LJ Update_Volume");
}

void LENNARD_JONES_INFORMATION::Clear() {
    if (!is_initialized) return;
    sponge_lj_destroy(handle);
    handle = nullptr;
    delete[] h_LJ_A;
    delete[] h_LJ_B;
    delete[] h_atom_LJ_type;
    h_LJ_A = h_LJ_B = nullptr;
    h_atom_LJ_type = nullptr;
    is_initialized = 0;
}

// ------------------------------------------------------------ LJ_SOFT_CORE
// Get_Beta_LJ (LJ_soft_core.cu:6-35): double the upper bound until
// erfc(b rc)/rc <= tol, then (doublings + 50) bisection steps; FP64 here, as
// the PME module's beta, so that both modules hold the same value.
static float soft_core_beta(double rc, double tol) {
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
    return (float)beta;
}

void LJ_SOFT_CORE::Initial(CONTROLLER* controller, float cut, VECTOR box, const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "LJ_soft_core");
    if (!controller->Command_Exist(module_name, "in_file")) return;  // not initialised, as the reference
    if (!controller->Command_Exist("lambda_lj")) {
        std::fprintf(stderr, "Error: FEP lambda of LJ must be given for the calculation of SOFT CORE.\n");
        std::exit(1);
    }
    lambda = (float)std::atof(controller->Command("lambda_lj"));
    if (controller->Command_Exist("soft_core_alpha")) alpha = (float)std::atof(controller->Command("soft_core_alpha"));
    if (controller->Command_Exist("soft_core_powfer")) p = (float)std::atof(controller->Command("soft_core_powfer"));
    if (controller->Command_Exist("soft_core_sigma")) sigma = (float)std::atof(controller->Command("soft_core_sigma"));
    if (controller->Command_Exist("soft_core_sigma_min"))
        sigma_min = (float)std::atof(controller->Command("soft_core_sigma_min"));
    FILE* fp = std::fopen(controller->Command(module_name, "in_file"), "r");
    if (!fp) {
        std::fprintf(stderr, "Error: cannot open %s\n", controller->Command(module_name, "in_file"));
        std::exit(1);
    }
    auto bad = [&](const char* what) {
        std::fprintf(stderr, "Error: %s ends early (%s)\n", controller->Command(module_name, "in_file"), what);
        std::exit(1);
    };
    if (std::fscanf(fp, "%d %d %d", &atom_numbers, &atom_type_numbers_A, &atom_type_numbers_B) != 3) bad("header");
    const int PA = atom_type_numbers_A * (atom_type_numbers_A + 1) / 2;
    const int PB = atom_type_numbers_B * (atom_type_numbers_B + 1) / 2;
    std::vector<float> AA(PA), AB(PA), BA(PB), BB(PB);
    std::vector<int> tA(atom_numbers), tB(atom_numbers);
    for (int i = 0; i < PA; i++) {
        if (std::fscanf(fp, "%f", &AA[i]) != 1) bad("AA");
        AA[i] *= 12.0f;
    }
    for (int i = 0; i < PA; i++) {
        if (std::fscanf(fp, "%f", &AB[i]) != 1) bad("AB");
        AB[i] *= 6.0f;
    }
    for (int i = 0; i < PB; i++) {
        if (std::fscanf(fp, "%f", &BA[i]) != 1) bad("BA");
        BA[i] *= 12.0f;
    }
    for (int i = 0; i < PB; i++) {
        if (std::fscanf(fp, "%f", &BB[i]) != 1) bad("BB");
        BB[i] *= 6.0f;
    }
    for (int i = 0; i < atom_numbers; i++)
        if (std::fscanf(fp, "%d %d", &tA[i], &tB[i]) != 2) bad("types");
    std::fclose(fp);
    h_subsys_division.assign(atom_numbers, 0);
    if (controller->Command_Exist("subsys_division_in_file")) {
        FILE* fm = std::fopen(controller->Command("subsys_division_in_file"), "r");
        int n = 0;
        if (!fm || std::fscanf(fm, "%d", &n) != 1 || n != atom_numbers) {
            std::fprintf(stderr, "Error: subsys_division_in_file missing or atom_numbers differ\n");
            std::exit(1);
        }
        for (int i = 0; i < n; i++)
            if (std::fscanf(fm, "%d", &h_subsys_division[i]) != 1) {
                std::fprintf(stderr, "Error: subsys_division_in_file ends early\n");
                std::exit(1);
            }
        std::fclose(fm);
    }
    float tol = 1e-5f;
    if (controller->Command_Exist("PME_Direct_Tolerance"))
        tol = (float)std::atof(controller->Command("PME_Direct_Tolerance"));
    pme_beta = soft_core_beta(cut, tol);
    cutoff = cut;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_lj_soft_create(atom_numbers, atom_type_numbers_A, atom_type_numbers_B, AA.data(),
                                            AB.data(), BA.data(), BB.data(), tA.data(), tB.data(),
                                            h_subsys_division.data(), lambda, alpha, p, sigma, sigma_min, cut, b,
                                            &handle),
                      "LJ_SOFT_CORE::Initial");
    hip_check(hipMalloc((void**)&d_subsys_division, atom_numbers * sizeof(int)), "hipMalloc");
    hip_check(hipMemcpy(d_subsys_division, h_subsys_division.data(), atom_numbers * sizeof(int),
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    controller->printf("    FEP lj lambda %f, soft core alpha %f powfer %f sigma %f sigma_min %f\n", lambda, alpha,
                       p, sigma, sigma_min);
    is_initialized = 1;
}

float LJ_SOFT_CORE::Get_Energy_With_Coulomb_Direct(const UNSIGNED_INT_VECTOR* uint_crd, const sponge_nl* nl,
                                                   const float* charge, int is_download) {
    if (!is_initialized) return NAN;
    double e[4];
    sponge_host_check(sponge_lj_soft_energy_with_coulomb_direct(handle, nl, reinterpret_cast<const unsigned*>(uint_crd),
                                                                charge, pme_beta, e, nullptr),
                      "Get_Energy_With_Coulomb_Direct");
    h_LJ_energy_sum_intersys = (float)e[0];
    h_LJ_energy_sum_intrasys = (float)e[1];
    h_direct_ene_sum_intersys = (float)e[2];
    h_direct_ene_sum_intrasys = (float)e[3];
    h_LJ_energy_sum = (float)(e[0] + e[1]);
    return is_download ? h_LJ_energy_sum : 0.f;
}

float LJ_SOFT_CORE::Get_Partial_H_Partial_Lambda_With_Columb_Direct(const UNSIGNED_INT_VECTOR* uint_crd,
                                                                    const float* charge, const sponge_nl* nl,
                                                                    const float* charge_B_A,
                                                                    const int charge_pertubated, int is_download) {
    if (!is_initialized) return NAN;
    double d[2];
    sponge_host_check(sponge_lj_soft_dhdl_with_coulomb_direct(handle, nl, reinterpret_cast<const unsigned*>(uint_crd),
                                                              charge, charge_B_A, charge_pertubated > 0, pme_beta, d,
                                                              nullptr),
                      "Get_Partial_H_Partial_Lambda_With_Columb_Direct");
    h_sigma_of_dH_dlambda_lj = (float)d[0];
    h_sigma_of_dH_dlambda_direct = (float)d[1];
    return is_download ? h_sigma_of_dH_dlambda_lj : 0.f;
}

float LJ_SOFT_CORE::Long_Range_Correction() {
    if (!is_initialized) return NAN;
    double v = 0;
    sponge_host_check(sponge_lj_soft_long_range_correction(handle, &v, nullptr), "Long_Range_Correction");
    return (float)v;
}

float LJ_SOFT_CORE::Partial_H_Partial_Lambda_Long_Range_Correction() {
    if (!is_initialized) return NAN;
    double v = 0;
    sponge_host_check(sponge_lj_soft_long_range_correction(handle, nullptr, &v),
                      "Partial_H_Partial_Lambda_Long_Range_Correction");
    return (float)v;
}

void LJ_SOFT_CORE::Update_Volume(VECTOR box) {
    if (!is_initialized) return;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_lj_update_volume(handle, b), "LJ_SOFT_CORE Update_Volume");
}

void LJ_SOFT_CORE::Clear() {
    if (!is_initialized) return;
    sponge_lj_destroy(handle);
    handle = nullptr;
    (void)hipFree(d_subsys_division);
    d_subsys_division = nullptr;
    is_initialized = 0;
}

// ------------------------------------------------------ Particle_Mesh_Ewald
void Particle_Mesh_Ewald::Initial(CONTROLLER* controller, int n, VECTOR box, float cut, const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "PME");
    atom_numbers = n;
    cutoff = cut;
    boxlength = box;
    tolerance = 1e-5f;
    if (controller->Command_Exist(module_name, "Direct_Tolerance"))
        tolerance = (float)std::atof(controller->Command(module_name, "Direct_Tolerance"));
    fftx = ffty = fftz = -1;
    if (controller->Command_Exist(module_name, "fftx")) fftx = std::atoi(controller->Command(module_name, "fftx"));
    if (controller->Command_Exist(module_name, "ffty")) ffty = std::atoi(controller->Command(module_name, "ffty"));
    if (controller->Command_Exist(module_name, "fftz")) fftz = std::atoi(controller->Command(module_name, "fftz"));
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_pme_create(n, b, cut, tolerance, fftx, ffty, fftz, &handle), "Particle_Mesh_Ewald::Initial");
    int k[3];
    sponge_host_check(sponge_pme_info(handle, &beta, k), "PME info");
    fftx = k[0];
    ffty = k[1];
    fftz = k[2];
    hip_check(hipMalloc((void**)&d_direct_atom_energy, n * sizeof(float)), "hipMalloc");
    hip_check(hipMalloc((void**)&d_correction_atom_energy, n * sizeof(float)), "hipMalloc");
    hip_check(hipMemset(d_direct_atom_energy, 0, n * sizeof(float)), "hipMemset");
    hip_check(hipMemset(d_correction_atom_energy, 0, n * sizeof(float)), "hipMemset");
    controller->printf("    PME beta %f, fft %d %d %d\n", beta, fftx, ffty, fftz);
    is_initialized = 1;
    if (!is_controller_printf_initialized) {  // PME_force.cu:298-303
        controller->Step_Print_Initial(module_name, "%.2f");
        is_controller_printf_initialized = 1;
    }
}

void Particle_Mesh_Ewald::PME_Excluded_Force_With_Atom_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const VECTOR sacler,
                                                              const float* charge, const int* es, const int* el,
                                                              const int* en, VECTOR* frc, float* atom_energy) {
    if (!is_initialized) return;
    float s[3] = {sacler.x, sacler.y, sacler.z};
    // The reference resets d_correction_atom_energy before its kernel adds
    // into atom_energy (usually that same array); the library's kernel assigns
    // each atom's correction energy, so no reset is needed.
    sponge_host_check(sponge_pme_excluded_force(handle, reinterpret_cast<const unsigned*>(uint_crd), s, charge, es, el,
                                                en, reinterpret_cast<float*>(frc), atom_energy, nullptr),
                      "PME_Excluded_Force_With_Atom_Energy");
}

void Particle_Mesh_Ewald::PME_Reciprocal_Force_With_Energy_And_Virial(const UNSIGNED_INT_VECTOR* uint_crd,
                                                                      const float* charge, VECTOR* force,
                                                                      int need_virial, int need_energy,
                                                                      float* d_virial, float* d_potential) {
    if (!is_initialized) return;
    sponge_host_check(sponge_pme_reciprocal_force(handle, reinterpret_cast<const unsigned*>(uint_crd), charge,
                                                  reinterpret_cast<float*>(force), need_virial, need_energy, d_virial,
                                                  d_potential, d_direct_atom_energy, nullptr),
                      "PME_Reciprocal_Force_With_Energy_And_Virial");
}

float Particle_Mesh_Ewald::Get_Energy(const UNSIGNED_INT_VECTOR* uint_crd, const float* charge, const sponge_nl* nl,
                                      const VECTOR scaler, const int* es, const int* el, const int* en,
                                      int which_part, int) {
    if (!is_initialized) return NAN;
    float s[3] = {scaler.x, scaler.y, scaler.z};
    auto part = [&](int w) {
        double e = 0;
        sponge_host_check(sponge_pme_get_energy(handle, nl, reinterpret_cast<const unsigned*>(uint_crd), charge, s, es,
                                                el, en, w, &e, nullptr),
                          "PME Get_Energy");
        return (float)e;
    };
    if (which_part == TOTAL || which_part == RECIPROCAL) reciprocal_ene = part(RECIPROCAL);
    if (which_part == TOTAL || which_part == SELF) self_ene = part(SELF);
    if (which_part == TOTAL || which_part == DIRECT) direct_ene = part(DIRECT);
    if (which_part == TOTAL || which_part == CORRECTION) correction_ene = part(CORRECTION);
    switch (which_part) {
        case TOTAL: return ee_ene = reciprocal_ene + self_ene + direct_ene + correction_ene;
        case RECIPROCAL: return reciprocal_ene;
        case SELF: return self_ene;
        case DIRECT: return direct_ene;
        default: return correction_ene;
    }
}

void Particle_Mesh_Ewald::Update_Box_Length(VECTOR box) {
    if (!is_initialized) return;
    boxlength = box;
    float b[3] = {box.x, box.y, box.z};
    sponge_host_check(sponge_pme_update_box_length(handle, b), "PME Update_Box_Length");
}

// The reference's Update_Volume forwards its stale member box
// (PME_force.cu:996-999); the new box is used here.
void Particle_Mesh_Ewald::Update_Volume(VECTOR box) { Update_Box_Length(box); }

void Particle_Mesh_Ewald::Clear() {
    if (!is_initialized) return;
    sponge_pme_destroy(handle);
    handle = nullptr;
    (void)hipFree(d_direct_atom_energy);
    (void)hipFree(d_correction_atom_energy);
    d_direct_atom_energy = d_correction_atom_energy = nullptr;
    is_initialized = 0;
}

// ---- the step on either side of the force path --------------------------

void CONSTRAIN::Initial_Constrain(CONTROLLER* controller, int n, float step, VECTOR box_length, float exp_gamma,
                                  int is_Minimization, const float* atom_mass, int* system_freedom) {
    (void)box_length;
    atom_numbers = n;
    dt = step;
    dt_inverse = 1.0f / step;
    v_factor = exp_gamma;
    x_factor = 0.5f * (1.0f + exp_gamma);
    if (is_Minimization) v_factor = 0.0f;
    pairs.clear();
    if (controller->Command_Exist(module_name, "in_file")) {
        const char* path = controller->Command(module_name, "in_file");
        FILE* fp = std::fopen(path, "r");
        if (!fp) {
            std::fprintf(stderr, "Error: cannot open %s\n", path);
            std::exit(1);
        }
        int count = 0;
        if (std::fscanf(fp, "%d", &count) != 1 || count < 0) {
            std::fprintf(stderr, "Error: %s: constraint count not found\n", path);
            std::exit(1);
        }
        for (int k = 0; k < count; k++) {
            CONSTRAIN_PAIR p{};
            if (std::fscanf(fp, "%d %d %f", &p.atom_i_serial, &p.atom_j_serial, &p.constant_r) != 3 ||
                p.atom_i_serial < 0 || p.atom_i_serial >= n || p.atom_j_serial < 0 || p.atom_j_serial >= n) {
                std::fprintf(stderr, "Error: %s: malformed constraint %d\n", path, k);
                std::exit(1);
            }
            const float mi = atom_mass[p.atom_i_serial], mj = atom_mass[p.atom_j_serial];
            p.constrain_k = mi * mj / (mi + mj) / x_factor;
            pairs.push_back(p);
        }
        std::fclose(fp);
    }
    constrain_pair_numbers = (int)pairs.size();
    h_constrain_pair = pairs.data();
    if (system_freedom) *system_freedom -= constrain_pair_numbers;
    controller->printf("    constrain pair number is %d\n", constrain_pair_numbers);
}

void SETTLE::Remember_Last_Coordinates(UNSIGNED_INT_VECTOR* uint_crd, VECTOR scaler) {
    if (!is_initialized) return;
    const float s[3] = {scaler.x, scaler.y, scaler.z};
    sponge_host_check(sponge_settle_remember_last_coordinates(handle, reinterpret_cast<const unsigned*>(uint_crd), s,
                                                              nullptr),
                      "SETTLE::Remember_Last_Coordinates");
}

void SETTLE::Do_SETTLE(const float* d_mass, VECTOR* crd, VECTOR box_length, VECTOR* vel, int need_pressure,
                       float* d_pressure) {
    if (!is_initialized) return;
    const float b[3] = {box_length.x, box_length.y, box_length.z};
    sponge_host_check(sponge_settle_do(handle, d_mass, reinterpret_cast<float*>(crd), b,
                                       reinterpret_cast<float*>(vel), need_pressure, d_pressure, nullptr),
                      "SETTLE::Do_SETTLE");
}

void SETTLE::Clear() {
    if (handle) sponge_settle_destroy(handle);
    handle = nullptr;
    is_initialized = 0;
}

void Leap_Frog(int atom_numbers, VECTOR* vel, VECTOR* crd, VECTOR* frc, VECTOR* acc, const float* d_mass_inverse,
               float dt, float max_velocity) {
    sponge_host_check(sponge_md_leap_frog(atom_numbers, reinterpret_cast<float*>(vel), reinterpret_cast<float*>(crd),
                                          reinterpret_cast<const float*>(frc), reinterpret_cast<float*>(acc),
                                          d_mass_inverse, dt, max_velocity, nullptr),
                      "Leap_Frog");
}

float Get_Pressure(int atom_numbers, const VECTOR* vel, const float* d_mass, const float* d_atom_virial,
                   float* d_virial, VECTOR box, float* d_pressure, int is_download) {
    const float volume = box.x * box.y * box.z;
    sponge_host_check(sponge_md_get_pressure(atom_numbers, reinterpret_cast<const float*>(vel), d_mass, d_atom_virial,
                                             d_virial, volume, d_pressure, nullptr, nullptr),
                      "Get_Pressure");
    if (!is_download) return 0.f;
    float p = 0.f;
    hip_check(hipMemcpy(&p, d_pressure, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    return p;
}

// ---------------------------------------------------------------- barostats
static const float kPresConvertion = 6.946827162543585e4f;  // bar per kcal/mol/A^3 (common.cuh:24-25)

void BERENDSEN_BAROSTAT_INFORMATION::Initial(CONTROLLER* controller, float target_pressure, VECTOR box,
                                             const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "berendsen_barostat");
    V0 = box.x * box.y * box.z;
    newV = V0;
    if (controller->Command_Exist("dt")) dt = (float)std::atof(controller->Command("dt"));
    if (controller->Command_Exist(module_name, "tau")) taup = (float)std::atof(controller->Command(module_name, "tau"));
    if (controller->Command_Exist(module_name, "compressibility"))
        compressibility = (float)std::atof(controller->Command(module_name, "compressibility"));
    compressibility *= kPresConvertion;
    if (controller->Command_Exist(module_name, "stochastic_term"))
        stochastic_term = (int)std::atof(controller->Command(module_name, "stochastic_term"));
    update_interval = stochastic_term == 1 ? 1 : 10;
    if (controller->Command_Exist(module_name, "update_interval"))
        update_interval = std::atoi(controller->Command(module_name, "update_interval"));
    if (stochastic_term && update_interval != 1) {
        std::fprintf(stderr, "Error: the update interval should be 1 when using stochastic term\n");
        std::exit(1);
    }
    if (controller->Command_Exist("target_temperature"))
        target_temperature = (float)std::atof(controller->Command("target_temperature"));
    rng_state = 1;
    if (controller->Command_Exist(module_name, "random_seed"))
        rng_state = (unsigned long long)std::atoll(controller->Command(module_name, "random_seed"));
    controller->printf("    Berendsen barostat: target %.2f bar, tau %f ps, compressibility %e, interval %d\n",
                       target_pressure * kPresConvertion, taup, compressibility / kPresConvertion, update_interval);
    is_initialized = 1;
    if (!is_controller_printf_initialized) {  // Berendsen_barostat.cu:70-76
        controller->Step_Print_Initial("density", "%.4f");
        controller->Step_Print_Initial("pressure", "%.2f");
        is_controller_printf_initialized = 1;
    }
}

void BERENDSEN_BAROSTAT_INFORMATION::Ask_For_Calculate_Pressure(int steps, int* need_pressure) {
    if (is_initialized && steps % update_interval == 0) *need_pressure += 1;
}

// Standard normal deviate from a 64-bit linear congruential generator and the
// Box-Muller transform (the reference draws from std::normal_distribution).
static double lcg_normal(unsigned long long* s) {
    auto next = [s]() {
        *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((*s >> 11) + 0.5) / 9007199254740992.0;
    };
    const double u1 = next(), u2 = next();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

float BERENDSEN_BAROSTAT_INFORMATION::Scale_Factor(float p_now, float target_pressure, float* vel_factor) {
    float mu = 1 - update_interval * compressibility * dt / taup / 3 * (target_pressure - p_now);
    *vel_factor = 0.f;
    if (stochastic_term) {
        const float sd = update_interval * dt * 20.455f;
        mu += std::sqrt(2 * 0.00198716f * target_temperature * compressibility / taup / newV) / 3 *
              (float)(sd * lcg_normal(&rng_state));
        *vel_factor = 1.0f / mu;
    }
    return mu;
}

void ANDERSEN_BAROSTAT_INFORMATION::Initial(CONTROLLER* controller, float target_pressure, VECTOR box,
                                            const char* name) {
    std::snprintf(module_name, sizeof(module_name), "%s", name ? name : "andersen_barostat");
    V0 = box.x * box.y * box.z;
    new_V = V0;
    float taup = 1.0f, compressibility = 4.5e-5f;
    if (controller->Command_Exist(module_name, "tau")) taup = (float)std::atof(controller->Command(module_name, "tau"));
    if (controller->Command_Exist(module_name, "compressibility"))
        compressibility = (float)std::atof(controller->Command(module_name, "compressibility"));
    taup *= 20.455f;
    compressibility *= kPresConvertion;
    h_mass_inverse = V0 * compressibility / taup / taup;
    dV_dt = 0.f;
    if (controller->Command_Exist(module_name, "dV/dt")) dV_dt = (float)std::atof(controller->Command(module_name, "dV/dt"));
    controller->printf("    Andersen barostat: target %.2f bar, tau %f ps, compressibility %e\n",
                       target_pressure * kPresConvertion, taup / 20.455f, compressibility / kPresConvertion);
    is_initialized = 1;
    if (!is_controller_printf_initialized) {  // andersen_barostat.cu:46-53
        controller->Step_Print_Initial("density", "%.4f");
        controller->Step_Print_Initial("pressure", "%.2f");
        controller->Step_Print_Initial("dV/dt", "%.2f");
        is_controller_printf_initialized = 1;
    }
}

float ANDERSEN_BAROSTAT_INFORMATION::Scale_Factor(float p_now, float target_pressure, float volume, float dt) {
    dV_dt += (p_now - target_pressure) * h_mass_inverse;
    new_V = volume + dV_dt * dt;
    return (float)std::cbrt((double)new_V / (double)volume);
}
