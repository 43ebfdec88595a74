// sponge_nb_driver: SPONGE's step loop (main.cu:33-46, 48-79, 162-194,
// 282-320, 376-434) over the C++ mirror classes, so the call order a SPONGE
// build keeps is exercised end to end. Each step converts the coordinates to
// the uint grid, clears the force, runs the LJ + PME direct pair force, the
// long-range correction, the excluded-pair correction and the reciprocal
// force; with a molecular-dynamics mode it then remembers the constrained
// bond vectors, takes the leap-frog step and applies SETTLE (NVE); finally it
// updates the neighbor list as Main_Iteration does and writes the outputs of
// Main_Print. Without `mode` or `mass_in_file` the coordinates stay fixed.
//
// Inputs (mdin keys, as MD_core.cu:257-298, 308-345, 494-551, 736-826,
// 860-1049, 1053-1080, 1212-1347 read them): coordinate_in_file and
// velocity_in_file, or amber_rst7 with amber_irest; mass_in_file,
// charge_in_file, exclude_in_file, residue_in_file, LJ_in_file,
// constrain_in_file (count, then `i j r`; constrain.cu:27-64), mode (NVE),
// dt (ps), cutoff, skin, step_limit, write_mdout_interval,
// write_information_interval, write_restart_file_interval, mdout, crd, box,
// rst, neighbor_list_*, PME_*. Outputs: the mdout columns (control.cu:490-573:
// step, time, temperature, potential, LJ, PME, and density, pressure, dV/dt
// with a barostat), mdcrd.dat (float32 xyz frames, MD_core.cu:1349-1360),
// mdbox.txt (MD_core.cu:1402-1412) and the restart: an AMBER rst7 named
// <rst> after an rst7 input, else <rst>_coordinate.txt / <rst>_velocity.txt
// (MD_core.cu:1414-1487). Extra keys of this driver: frc_out_file (forces
// after the first step: a line `N`, then N lines `fx fy fz`); check_input
// (parse, print what was read, write the restart of the state as read when
// `rst` is given, and stop before any device call).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

struct MdInputs {
    int atom_numbers = 0;
    VECTOR box{0, 0, 0};
    double start_time = 0.0;
    int amber_irest = -1;  // >= 0: restarts are AMBER rst7 files (MD_core.cu:1425)
    std::vector<VECTOR> crd, vel;
    std::vector<float> charge, mass;
    std::vector<int> ex_start, ex_numbers, ex_list;
    std::vector<int> res_start, res_end;
};

static void check_count(MdInputs& in, int n, const char* what) {
    if (n <= 0) {
        std::fprintf(stderr, "Error: atom_numbers not found in %s\n", what);
        std::exit(1);
    }
    if (in.atom_numbers > 0 && in.atom_numbers != n) {
        std::fprintf(stderr, "Error: atom_numbers is not equal: %d %d\n", in.atom_numbers, n);
        std::exit(1);
    }
    in.atom_numbers = n;
}

// The wrap of Crd_Periodic_Map (common.cu:299-345): a coordinate below 0
// gains one box length, one at or past the box loses one.
static float periodic_map(float x, float L) {
    if (x < 0.f) return x + L;
    return x < L ? x : x - L;
}

// MD_INFORMATION::Read_Rst7 (MD_core.cu:1258-1347): a title line; `natom`
// or `natom time` (a time means the file holds velocities); natom
// coordinates; natom velocities if it holds them; the box lengths. The
// velocities are zero when the file holds none or amber_irest is 0. Ten
// passes of the periodic map bring badly wrapped coordinates into the box.
static void read_rst7(const char* path, int irest, MdInputs& in) {
    FILE* fp = open_or_die(path, "r");
    char line[1024];
    int n = 0, fields = 0;
    if (std::fgets(line, sizeof(line), fp) && std::fgets(line, sizeof(line), fp))
        fields = std::sscanf(line, "%d %lf", &n, &in.start_time);
    if (fields < 1) n = 0;
    check_count(in, n, "amber_rst7");
    const bool has_vel = fields == 2;
    if (!has_vel) in.start_time = 0.0;
    std::printf("    Start reading AMBER rst7:\n        atom_numbers is %d\n        system start time is %lf\n", n,
                in.start_time);
    if (!has_vel || irest == 0) std::printf("        All velocity will be set to 0\n");
    in.crd.resize(n);
    in.vel.assign(n, VECTOR{0.f, 0.f, 0.f});
    auto read3 = [&](VECTOR& v, const char* what) {
        if (std::fscanf(fp, "%f %f %f", &v.x, &v.y, &v.z) != 3) {
            std::fprintf(stderr, "Error: amber_rst7 ends early in the %s\n", what);
            std::exit(1);
        }
    };
    for (int i = 0; i < n; i++) read3(in.crd[i], "coordinates");
    if (has_vel)
        for (int i = 0; i < n; i++) read3(in.vel[i], "velocities");
    if (!has_vel || irest == 0) in.vel.assign(n, VECTOR{0.f, 0.f, 0.f});
    read3(in.box, "box");
    std::fclose(fp);
    std::printf("        system size is %f %f %f\n    End reading AMBER rst7\n\n", in.box.x, in.box.y, in.box.z);
    for (int pass = 0; pass < 10; pass++)
        for (VECTOR& v : in.crd) {
            v.x = periodic_map(v.x, in.box.x);
            v.y = periodic_map(v.y, in.box.y);
            v.z = periodic_map(v.z, in.box.z);
        }
}

// MD_INFORMATION::trajectory_output::Export_Restart_File with amber_irest
// >= 0 (MD_core.cu:1425-1463): title, `%8d %.3lf` atoms and time, then the
// coordinates and the velocities as `%12.7f` triples two atoms a line, then
// the box lengths and three 90 degree angles.
static void write_rst7(const char* path, const std::string& title, double now, const std::vector<VECTOR>& crd,
                       const std::vector<VECTOR>& vel, VECTOR box) {
    FILE* fp = open_or_die(path, "w");
    std::fprintf(fp, "%s\n", title.c_str());
    std::fprintf(fp, "%8d %.3lf\n", (int)crd.size(), now);
    for (const std::vector<VECTOR>* list : {&crd, &vel}) {
        int s = 0;
        for (const VECTOR& v : *list) {
            std::fprintf(fp, "%12.7f%12.7f%12.7f", v.x, v.y, v.z);
            if (++s == 2) {
                s = 0;
                std::fprintf(fp, "\n");
            }
        }
        if (s == 1) std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "%12.7f%12.7f%12.7f", box.x, box.y, box.z);
    std::fprintf(fp, "%12.7f%12.7f%12.7f", 90.0f, 90.0f, 90.0f);
    std::fclose(fp);
}

// MD_INFORMATION::residue_information::Initial (MD_core.cu:736-826):
// residue_in_file holds `natom nres` and the atom count of each residue in
// order; without it every atom is its own residue.
static void read_residues(CONTROLLER& c, MdInputs& in) {
    const int n = in.atom_numbers;
    in.res_start.clear();
    in.res_end.clear();
    if (!c.Command_Exist("residue_in_file")) {
        for (int i = 0; i < n; i++) {
            in.res_start.push_back(i);
            in.res_end.push_back(i + 1);
        }
        return;
    }
    FILE* fp = open_or_die(c.Command("residue_in_file"), "r");
    int atoms = 0, nres = 0;
    if (std::fscanf(fp, "%d %d", &atoms, &nres) != 2) atoms = 0;
    check_count(in, atoms, "residue_in_file");
    int count = 0;
    for (int r = 0; r < nres; r++) {
        int k = -1;
        if (std::fscanf(fp, "%d", &k) != 1 || k < 0 || count + k > n) {
            std::fprintf(stderr, "Error: malformed residue_in_file at residue %d\n", r);
            std::exit(1);
        }
        in.res_start.push_back(count);
        count += k;
        in.res_end.push_back(count);
    }
    std::fclose(fp);
}

static void read_inputs(CONTROLLER& c, MdInputs& in) {
    // Read_Coordinate_And_Velocity (MD_core.cu:860-920): coordinate_in_file
    // (with velocity_in_file) first, else amber_rst7
    if (!c.Command_Exist("coordinate_in_file")) {
        if (!c.Command_Exist("amber_rst7")) {
            std::fprintf(stderr, "Error: coordinate_in_file or amber_rst7 is required\n");
            std::exit(1);
        }
        in.amber_irest = 1;
        if (c.Command_Exist("amber_irest")) in.amber_irest = std::atoi(c.Command("amber_irest"));
        read_rst7(c.Command("amber_rst7"), in.amber_irest, in);
    } else {
        FILE* fp = open_or_die(c.Command("coordinate_in_file"), "r");
        char line[1024];
        int n = 0;
        if (!std::fgets(line, sizeof(line), fp) || std::sscanf(line, "%d %lf", &n, &in.start_time) < 1) n = 0;
        check_count(in, n, "coordinate_in_file");
        in.crd.resize(n);
        for (int i = 0; i < n; i++)
            if (std::fscanf(fp, "%f %f %f", &in.crd[i].x, &in.crd[i].y, &in.crd[i].z) != 3) {
                std::fprintf(stderr, "Error: coordinate_in_file ends early\n");
                std::exit(1);
            }
        if (std::fscanf(fp, "%f %f %f", &in.box.x, &in.box.y, &in.box.z) != 3) {
            std::fprintf(stderr, "Error: box_length not found in coordinate_in_file\n");
            std::exit(1);
        }
        std::fclose(fp);
        in.vel.assign(in.atom_numbers, VECTOR{0.f, 0.f, 0.f});
    }
    if (c.Command_Exist("coordinate_in_file") && c.Command_Exist("velocity_in_file")) {
        FILE* fp = open_or_die(c.Command("velocity_in_file"), "r");
        char line[1024];
        int n = 0;
        if (!std::fgets(line, sizeof(line), fp) || std::sscanf(line, "%d", &n) != 1) n = 0;
        check_count(in, n, "velocity_in_file");
        for (int i = 0; i < n; i++)
            if (std::fscanf(fp, "%f %f %f", &in.vel[i].x, &in.vel[i].y, &in.vel[i].z) != 3) {
                std::fprintf(stderr, "Error: velocity_in_file ends early\n");
                std::exit(1);
            }
        std::fclose(fp);
    }
    in.mass.assign(in.atom_numbers, 20.f);  // MD_core.cu:1025-1037 default
    if (c.Command_Exist("mass_in_file")) {
        FILE* fp = open_or_die(c.Command("mass_in_file"), "r");
        int n = 0;
        if (std::fscanf(fp, "%d", &n) != 1) n = 0;
        check_count(in, n, "mass_in_file");
        for (int i = 0; i < n; i++)
            if (std::fscanf(fp, "%f", &in.mass[i]) != 1 || in.mass[i] < 0.f) {
                std::fprintf(stderr, "Error: mass_in_file ends early or holds a negative mass\n");
                std::exit(1);
            }
        std::fclose(fp);
    }
    in.charge.assign(in.atom_numbers, 0.f);
    if (c.Command_Exist("charge_in_file")) {
        FILE* fp = open_or_die(c.Command("charge_in_file"), "r");
        int n = 0;
        if (std::fscanf(fp, "%d", &n) != 1) n = 0;
        check_count(in, n, "charge_in_file");
        for (int i = 0; i < n; i++)
            if (std::fscanf(fp, "%f", &in.charge[i]) != 1) {
                std::fprintf(stderr, "Error: charge_in_file ends early\n");
                std::exit(1);
            }
        std::fclose(fp);
    }
    in.ex_start.assign(in.atom_numbers, 0);
    in.ex_numbers.assign(in.atom_numbers, 0);
    if (c.Command_Exist("exclude_in_file")) {
        FILE* fp = open_or_die(c.Command("exclude_in_file"), "r");
        int n = 0, total = 0;
        if (std::fscanf(fp, "%d %d", &n, &total) != 2) n = 0;
        check_count(in, n, "exclude_in_file");
        in.ex_list.reserve(total);
        for (int i = 0; i < n; i++) {
            int k = 0;
            if (std::fscanf(fp, "%d", &k) != 1) k = -1;
            in.ex_start[i] = (int)in.ex_list.size();
            in.ex_numbers[i] = k;
            for (int j = 0; j < k; j++) {
                int a = -1;
                if (std::fscanf(fp, "%d", &a) != 1 || a < 0 || a >= n) k = -1;
                in.ex_list.push_back(a);
            }
            if (k < 0 || (int)in.ex_list.size() > total) {
                std::fprintf(stderr, "Error: malformed exclude_in_file at atom %d\n", i);
                std::exit(1);
            }
        }
        std::fclose(fp);
    }
    read_residues(c, in);
}

template <class T>
static T* to_device(const std::vector<T>& h) {
    T* d = nullptr;
    size_t bytes = (h.empty() ? 1 : h.size()) * sizeof(T);
    hip_ok(hipMalloc((void**)&d, bytes), "hipMalloc");
    if (!h.empty()) hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

// Export_Restart_File (MD_core.cu:1414-1487): an AMBER rst7 named by `rst`
// (default `restart`) when the input was one (amber_irest >= 0), else
// <rst>_coordinate.txt and <rst>_velocity.txt.
static void export_restart(CONTROLLER& c, const MdInputs& in, const std::vector<VECTOR>& crd,
                           const std::vector<VECTOR>& vel, VECTOR box, double now) {
    const char* rst = c.Command_Exist("rst") ? c.Command("rst") : "restart";
    if (in.amber_irest >= 0) {
        write_rst7(rst, c.md_name, now, crd, vel, box);
        return;
    }
    const int n = (int)crd.size();
    char path[1100];
    std::snprintf(path, sizeof(path), "%s_coordinate.txt", rst);
    FILE* fc = open_or_die(path, "w");
    std::snprintf(path, sizeof(path), "%s_velocity.txt", rst);
    FILE* fv = open_or_die(path, "w");
    std::fprintf(fc, "%d %.3lf\n", n, now);
    std::fprintf(fv, "%d %.3lf\n", n, now);
    for (int i = 0; i < n; i++) {
        std::fprintf(fc, "%12.7f %12.7f %12.7f\n", crd[i].x, crd[i].y, crd[i].z);
        std::fprintf(fv, "%12.7f %12.7f %12.7f\n", vel[i].x, vel[i].y, vel[i].z);
    }
    std::fprintf(fc, "%12.7f %12.7f %12.7f %12.7f %12.7f %12.7f", box.x, box.y, box.z, 90.0f, 90.0f, 90.0f);
    std::fclose(fc);
    std::fclose(fv);
}

// Get_Residue_Temperature (MD_core.cu:1721-1742, kernel :216-239): the
// kinetic energy of each residue's centre-of-mass motion, T = 2 Ek / (3 kB
// nres) with kB = 0.00198716 kcal/mol/K.
static float residue_temperature(const MdInputs& in, const std::vector<VECTOR>& vel) {
    double ek = 0.0;
    for (size_t r = 0; r < in.res_start.size(); r++) {
        float px = 0.f, py = 0.f, pz = 0.f, m = 0.f;
        for (int i = in.res_start[r]; i < in.res_end[r]; i++) {
            px += in.mass[i] * vel[i].x;
            py += in.mass[i] * vel[i].y;
            pz += in.mass[i] * vel[i].z;
            m += in.mass[i];
        }
        ek += 0.5f * (px * px + py * py + pz * pz) / m;
    }
    return (float)(ek * 2.0 / 0.00198716f / (double)in.res_start.size() / 3.0);
}

int main(int argc, char** argv) {
    CONTROLLER controller;
    if (!controller.Initial(argc, argv)) return 1;
    controller.printf("%s\n", controller.md_name.c_str());

    MdInputs in;
    read_inputs(controller, in);
    const int n = in.atom_numbers;
    float cutoff = 10.0f, skin = 2.0f;
    int step_limit = 1, mdout_interval = 1000;
    if (controller.Command_Exist("cutoff")) cutoff = (float)std::atof(controller.Command("cutoff"));
    if (controller.Command_Exist("skin")) skin = (float)std::atof(controller.Command("skin"));
    if (controller.Command_Exist("step_limit")) step_limit = std::atoi(controller.Command("step_limit"));
    if (controller.Command_Exist("write_mdout_interval"))
        mdout_interval = std::atoi(controller.Command("write_mdout_interval"));
    if (mdout_interval <= 0) mdout_interval = 1;
    // MD_INFORMATION::Read_Mode / Read_dt (MD_core.cu:494-551): this build
    // integrates NVE only
    const bool md = controller.Command_Exist("mode") || controller.Command_Exist("mass_in_file");
    const bool npt = controller.Command_Exist("mode") && std::strcmp(controller.Command("mode"), "NPT") == 0;
    if (controller.Command_Exist("mode") && std::strcmp(controller.Command("mode"), "NVE") != 0 && !npt) {
        std::fprintf(stderr, "Error: mode %s is not provided here (NVE, or NPT with a barostat)\n",
                     controller.Command("mode"));
        return 1;
    }
    // NPT: a barostat (main.cu:115-126) over the NVE leap-frog; thermostats
    // are not provided here
    if (npt && controller.Command_Exist("thermostat")) {
        std::fprintf(stderr, "Error: thermostat %s is not provided here (NPT runs the NVE leap-frog with the "
                             "barostat)\n",
                     controller.Command("thermostat"));
        return 1;
    }
    const bool use_bd = npt && controller.Command_Choice("barostat", "berendsen_barostat");
    const bool use_ad = npt && controller.Command_Choice("barostat", "andersen_barostat");
    if (npt && !use_bd && !use_ad) {
        std::fprintf(stderr, "Error: mode NPT needs barostat = berendsen_barostat or andersen_barostat\n");
        return 1;
    }
    // MD_INFORMATION::system_information::Initial (MD_core.cu:268-271): bar -> kcal/mol/A^3
    float target_pressure = 1.0f;
    if (npt && controller.Command_Exist("target_pressure"))
        target_pressure = (float)std::atof(controller.Command("target_pressure"));
    target_pressure *= 0.00001439506089041446f;
    double dt_ps = 0.001;
    if (controller.Command_Exist("dt")) dt_ps = std::atof(controller.Command("dt"));
    const float dt = (float)dt_ps * 20.455f;
    int info_interval = 1000;
    if (controller.Command_Exist("write_information_interval"))
        info_interval = std::atoi(controller.Command("write_information_interval"));
    int restart_interval = info_interval;
    if (controller.Command_Exist("write_restart_file_interval"))
        restart_interval = std::atoi(controller.Command("write_restart_file_interval"));
    if (controller.Command_Exist("check_input")) {  // parse and report only; no device call
        for (auto& kv : controller.commands) controller.printf("command %s = %s\n", kv.first.c_str(), kv.second.c_str());
        double qsum = 0;
        for (float q : in.charge) qsum += q;
        controller.printf("atoms %d box %g %g %g excluded %zu charge_sum %.6f cutoff %g skin %g steps %d\n", n,
                          in.box.x, in.box.y, in.box.z, in.ex_list.size(), qsum, cutoff, skin, step_limit);
        controller.printf("residues %zu start_time %.3f amber_irest %d\n", in.res_start.size(), in.start_time,
                          in.amber_irest);
        // with `rst`, the restart file of the state as read (time start_time)
        if (controller.Command_Exist("rst")) export_restart(controller, in, in.crd, in.vel, in.box, in.start_time);
        return 0;
    }

    VECTOR* crd = to_device(in.crd);
    VECTOR* vel = to_device(in.vel);
    std::vector<float> inv_mass(n);
    for (int i = 0; i < n; i++) inv_mass[i] = in.mass[i] == 0.f ? 0.f : 1.0f / in.mass[i];
    float* d_mass = to_device(in.mass);
    float* d_mass_inverse = to_device(inv_mass);
    VECTOR* acc = to_device(std::vector<VECTOR>(n, VECTOR{0.f, 0.f, 0.f}));
    float* d_charge = to_device(in.charge);
    int* d_es = to_device(in.ex_start);
    int* d_en = to_device(in.ex_numbers);
    int* d_el = to_device(in.ex_list);
    UNSIGNED_INT_VECTOR* uint_crd = nullptr;
    VECTOR* frc = nullptr;
    float *d_atom_energy = nullptr, *d_atom_virial = nullptr, *d_scalars = nullptr, *d_pressure = nullptr;
    hip_ok(hipMalloc((void**)&uint_crd, n * sizeof(UNSIGNED_INT_VECTOR)), "hipMalloc");
    hip_ok(hipMalloc((void**)&frc, n * sizeof(VECTOR)), "hipMalloc");
    hip_ok(hipMalloc((void**)&d_atom_energy, n * sizeof(float)), "hipMalloc");
    hip_ok(hipMalloc((void**)&d_atom_virial, n * sizeof(float)), "hipMalloc");
    hip_ok(hipMalloc((void**)&d_scalars, 2 * sizeof(float)), "hipMalloc");  // d_virial, d_potential
    hip_ok(hipMalloc((void**)&d_pressure, sizeof(float)), "hipMalloc");
    hip_ok(hipMemset(d_pressure, 0, sizeof(float)), "hipMemset");
    VECTOR box = in.box;
    VECTOR scaler{box.x / 4294967296.0f, box.y / 4294967296.0f, box.z / 4294967296.0f};
    float box3[3] = {box.x, box.y, box.z};

    // MD_INFORMATION's mdout columns (MD_core.cu:273-276), ahead of the modules'
    controller.Step_Print_Initial("step", "%d");
    controller.Step_Print_Initial("time", "%.3lf");
    controller.Step_Print_Initial("temperature", "%.2f");
    controller.Step_Print_Initial("potential", "%.2f");
    double total_mass = 0.0;
    for (float m : in.mass) total_mass += m;
    NEIGHBOR_LIST neighbor_list;
    LENNARD_JONES_INFORMATION lj;
    Particle_Mesh_Ewald pme;
    neighbor_list.Initial(&controller, n, in.box, cutoff, skin);
    neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en, NEIGHBOR_LIST::FORCED_UPDATE);
    lj.Initial(&controller, cutoff, in.box);
    pme.Initial(&controller, n, in.box, cutoff);
    CONSTRAIN constrain;
    SETTLE settle;
    BERENDSEN_BAROSTAT_INFORMATION bd_baro;
    ANDERSEN_BAROSTAT_INFORMATION ad_baro;
    if (use_bd) bd_baro.Initial(&controller, target_pressure, box);
    if (use_ad) ad_baro.Initial(&controller, target_pressure, box);
    float h_pressure = 0.f;
    // Main_Volume_Change (main.cu:438-448) after the coordinates were scaled,
    // and Main_Volume_Change_Largely (:463-472) past a third of the volume
    auto volume_change = [&](float factor, float* V0) {
        box = VECTOR{factor * box.x, factor * box.y, factor * box.z};
        scaler = VECTOR{box.x / 4294967296.0f, box.y / 4294967296.0f, box.z / 4294967296.0f};
        box3[0] = box.x;
        box3[1] = box.y;
        box3[2] = box.z;
        neighbor_list.Update_Volume(box);
        neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en, NEIGHBOR_LIST::CONDITIONAL_UPDATE,
                                           NEIGHBOR_LIST::FORCED_CHECK);
        lj.Update_Volume(box);
        pme.Update_Volume(box);
        const float V = box.x * box.y * box.z;
        if (V > 1.331f * *V0 || V < 0.729f * *V0) {
            controller.printf("Some modules are based on the meshing methods, and it is more precise to "
                              "re-initialize these modules now for a long time or a large volume change.\n");
            neighbor_list.Clear();
            pme.Clear();
            neighbor_list.Initial(&controller, n, box, cutoff, skin);
            neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en, NEIGHBOR_LIST::FORCED_UPDATE);
            pme.Initial(&controller, n, box, cutoff);
            *V0 = V;
        }
    };
    FILE* crd_traj = nullptr;
    FILE* box_traj = nullptr;
    if (md) {
        int freedom = 3 * n;
        constrain.Initial_Constrain(&controller, n, dt, in.box, 1.0f, 0, in.mass.data(), &freedom);
        settle.Initial(&controller, &constrain, in.mass.data());
        if (constrain.constrain_pair_numbers > 0) {
            std::fprintf(stderr, "Error: %d constraints are not rigid triangles or pairs (SIMPLE_CONSTRAIN and "
                                 "SHAKE are not provided here)\n",
                         constrain.constrain_pair_numbers);
            return 1;
        }
        if (info_interval > 0) {
            crd_traj = open_or_die(controller.Command_Exist("crd") ? controller.Command("crd") : "mdcrd.dat", "wb");
            box_traj = open_or_die(controller.Command_Exist("box") ? controller.Command("box") : "mdbox.txt", "w");
        }
    }
    auto download = [&](VECTOR* d, std::vector<VECTOR>& h) {
        h.resize(n);
        hip_ok(hipMemcpy(h.data(), d, n * sizeof(VECTOR), hipMemcpyDeviceToHost), "hipMemcpy");
    };
    controller.Print_First_Line_To_Mdout();  // main.cu:132
    std::vector<float> h_atom_energy(n);

    auto t0 = std::chrono::steady_clock::now();
    for (int step = 1; step <= step_limit; step++) {
        const int need_potential = step % mdout_interval == 0;
        int need_pressure = 0;
        bd_baro.Ask_For_Calculate_Pressure(step, &need_pressure);  // main.cu:177-178
        if (ad_baro.is_initialized) need_pressure = 1;
        sponge_host_check(sponge_crd_to_uint(n, box3, reinterpret_cast<const float*>(crd),
                                             reinterpret_cast<unsigned*>(uint_crd), nullptr),
                          "Crd_To_Uint_Crd");
        hip_ok(hipMemsetAsync(frc, 0, n * sizeof(VECTOR), nullptr), "reset force");
        hip_ok(hipMemsetAsync(d_atom_energy, 0, n * sizeof(float), nullptr), "reset atom energy");
        hip_ok(hipMemsetAsync(d_atom_virial, 0, n * sizeof(float), nullptr), "reset atom virial");
        hip_ok(hipMemsetAsync(d_scalars, 0, 2 * sizeof(float), nullptr), "reset virial");
        lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(n, uint_crd, d_charge, frc, neighbor_list.d_nl, pme.beta,
                                                           need_potential, d_atom_energy, need_pressure,
                                                           d_atom_virial, pme.d_direct_atom_energy);
        lj.Long_Range_Correction(need_pressure, d_scalars, need_potential, d_scalars + 1);
        pme.PME_Excluded_Force_With_Atom_Energy(uint_crd, scaler, d_charge, d_es, d_el, d_en, frc,
                                                pme.d_correction_atom_energy);
        pme.PME_Reciprocal_Force_With_Energy_And_Virial(uint_crd, d_charge, frc, need_pressure, need_potential,
                                                        d_scalars, d_scalars + 1);
        if (need_pressure)  // Calculate_Pressure_And_Potential_If_Needed (MD_core.cu:1765-1775)
            Get_Pressure(n, vel, d_mass, d_atom_virial, d_scalars, box, d_pressure, 0);
        if (step == 1 && controller.Command_Exist("frc_out_file")) {
            std::vector<VECTOR> h(n);
            hip_ok(hipMemcpy(h.data(), frc, n * sizeof(VECTOR), hipMemcpyDeviceToHost), "hipMemcpy");
            FILE* fp = open_or_die(controller.Command("frc_out_file"), "w");
            std::fprintf(fp, "%d\n", n);
            for (int i = 0; i < n; i++) std::fprintf(fp, "%.9g %.9g %.9g\n", h[i].x, h[i].y, h[i].z);
            std::fclose(fp);
        }
        float e_pme = 0.f, e_lj = 0.f, potential = 0.f;
        if (need_potential) {
            // Get_Potential (MD_core.cu:1800-1812): the atom energies (LJ +
            // direct Coulomb) plus the scalar terms (long-range correction,
            // PME correction + self + reciprocal)
            float scalar_potential = 0.f;
            hip_ok(hipMemcpy(h_atom_energy.data(), d_atom_energy, n * sizeof(float), hipMemcpyDeviceToHost),
                   "hipMemcpy");
            hip_ok(hipMemcpy(&scalar_potential, d_scalars + 1, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
            double sum = scalar_potential;
            for (float e : h_atom_energy) sum += e;
            potential = (float)sum;
            e_pme = pme.Get_Energy(uint_crd, d_charge, neighbor_list.d_nl, scaler, d_es, d_el, d_en);
            e_lj = lj.Get_Energy(uint_crd, neighbor_list.d_nl);
            controller.printf("step %d LJ %.6f PME %.6f reciprocal %.6f self %.6f direct %.6f correction %.6f\n",
                              step, e_lj, e_pme, pme.reciprocal_ene, pme.self_ene, pme.direct_ene,
                              pme.correction_ene);
        }
        if (md) {
            // Main_Iteration (main.cu:282-320), NVE
            settle.Remember_Last_Coordinates(uint_crd, scaler);
            Leap_Frog(n, vel, crd, frc, acc, d_mass_inverse, dt, -1.0f);
            settle.Do_SETTLE(d_mass, crd, box, vel, need_pressure, d_pressure);
            if (bd_baro.is_initialized && step % bd_baro.update_interval == 0) {  // main.cu:334-354
                hip_ok(hipMemcpy(&h_pressure, d_pressure, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
                float vel_factor = 0.f;
                const float mu = bd_baro.Scale_Factor(h_pressure, target_pressure, &vel_factor);
                if (vel_factor != 0.f)
                    sponge_host_check(sponge_md_scale_list(3 * n, reinterpret_cast<float*>(vel), vel_factor, nullptr),
                                      "Scale_List");
                sponge_host_check(sponge_md_scale_list(3 * n, reinterpret_cast<float*>(crd), mu, nullptr),
                                  "Scale_List");
                volume_change(mu, &bd_baro.V0);
                bd_baro.newV = box.x * box.y * box.z;
            }
            if (ad_baro.is_initialized) {  // main.cu:356-372
                hip_ok(hipMemcpy(&h_pressure, d_pressure, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
                const float mu = ad_baro.Scale_Factor(h_pressure, target_pressure, box.x * box.y * box.z, dt);
                sponge_host_check(sponge_md_scale_list(3 * n, reinterpret_cast<float*>(crd), mu, nullptr),
                                  "Scale_List");
                sponge_host_check(sponge_md_scale_list(3 * n, reinterpret_cast<float*>(vel), 1.0f / mu, nullptr),
                                  "Scale_List");
                volume_change(mu, &ad_baro.V0);
            }
            if (need_pressure && need_potential)
                controller.printf("step %d pressure %.4f bar box %.6f %.6f %.6f\n", step, h_pressure * 6.946827162543585e4f,
                                  box.x, box.y, box.z);
        }
        neighbor_list.Neighbor_List_Update(crd, d_es, d_el, d_en);
        if (need_potential) {
            // print steps synchronise anyway: stop on a list that left atoms
            // or pairs out (the reference writes past its buckets silently)
            int ovf = 0;
            if (sponge_nl_status(neighbor_list.d_nl, nullptr, &ovf, nullptr) != SPONGE_NB_OK) {
                std::fprintf(stderr, "step %d: neighbor list capacity overflow (flags 0x%x); stopping\n", step, ovf);
                return 2;
            }
        }
        if (md && info_interval > 0 && step % info_interval == 0) {
            // Main_Print's trajectory frames (MD_core.cu:1349-1360, 1402-1412)
            std::vector<VECTOR> h;
            download(crd, h);
            std::fwrite(&h[0].x, sizeof(VECTOR), n, crd_traj);
            std::fprintf(box_traj, "%f %f %f %.0f %.0f %.0f\n", box.x, box.y, box.z, 90.0f, 90.0f, 90.0f);
        }
        const double now = in.start_time + dt_ps * step;
        if (need_potential) {  // Main_Print's mdout row (main.cu:382-409)
            std::vector<VECTOR> hv;
            download(vel, hv);
            controller.Step_Print("step", step);
            controller.Step_Print("time", now);
            controller.Step_Print("temperature", residue_temperature(in, hv));
            controller.Step_Print("potential", potential);
            controller.Step_Print("PME", e_pme);
            controller.Step_Print("LJ", e_lj);
            const float V = box.x * box.y * box.z;  // Get_Density (MD_core.cu:1510-1514)
            controller.Step_Print("density", (float)total_mass * 1e24f / 6.023e23f / V);
            controller.Step_Print("pressure", h_pressure * 6.946827162543585e4f);
            controller.Step_Print("dV/dt", ad_baro.dV_dt);
            controller.Print_To_Screen_And_Mdout();
        }
        if (md && restart_interval > 0 && step % restart_interval == 0) {
            std::vector<VECTOR> hc, hv;
            download(crd, hc);
            download(vel, hv);
            export_restart(controller, in, hc, hv, box, now);
        }
    }
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (crd_traj) std::fclose(crd_traj);
    if (box_traj) std::fclose(box_traj);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    long pairs = 0;
    int overflow = 0, rebuilds = 0;
    sponge_host_check(sponge_nl_status(neighbor_list.d_nl, &pairs, &overflow, &rebuilds), "sponge_nl_status");
    controller.printf("neighbor pairs %ld overflow %d rebuilds %d\n", pairs, overflow, rebuilds);
    controller.printf("Core Run Wall Time: %f second(s), %f steps/s\n", secs, step_limit / secs);

    controller.Clear();
    settle.Clear();
    pme.Clear();
    lj.Clear();
    neighbor_list.Clear();
    for (void* p : {(void*)vel, (void*)d_mass, (void*)d_mass_inverse, (void*)acc, (void*)crd, (void*)d_charge, (void*)d_es, (void*)d_en, (void*)d_el, (void*)uint_crd, (void*)frc,
                    (void*)d_atom_energy, (void*)d_atom_virial, (void*)d_scalars, (void*)d_pressure})
        (void)hipFree(p);
    return overflow ? 2 : 0;
}
