// ms_run: the host driver of a MorphoShell run on one MI355X.
//
//   ms_run <settings file> <reference VTK> [<ansatz VTK>]
//
// The command line, the settings file, the input VTK files, the output
// directory (output_using_<settings file name>, the previous one kept as
// prev_...), the log and the stepcount_<n>_output.vtk files follow the
// reference's executable (main.cpp, preliminary_setup.cpp,
// initialize_directories.cpp, write_output.cpp).  The work between the
// calls is the device's: set-up through the C-ABI host routines, then the
// whole dialling / waiting / equilibrium state machine of main.cpp:269-507
// runs on the GPU (ms_relax_setup / ms_relax_run), pausing at every
// equilibrium so this program can write its output, and stopping at a
// regular output step to write that one.  The output's device data are
// copied off synchronously; formatting and writing the file happen on a
// writer thread while the device steps on.
//
// Differences, each deliberate: the run stops at the first equilibrium at
// dial factor 1 (the reference's loop has no end unless glass slides are
// on, main.cpp:493); glass slides are refused.

#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <cstdio>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "morphoshell_amd.h"
#include "run_settings.h"
#include "vtk_io.h"

namespace fs = std::filesystem;
using namespace msrun;

namespace {

// get_real_time.cpp:31-40: day/month/year hour:minute.
std::string real_time() {
    const std::time_t t = std::time(nullptr);
    char buf[32];
    std::strftime(buf, sizeof buf, "%d/%m/%y %H:%M", std::localtime(&t));
    return buf;
}

std::string now_string() {
    const std::time_t t = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&t));
    return buf;
}

// Everything logged goes to the terminal and to <output dir>/log.txt.
class Log {
  public:
    void open(const std::string& path) { file_.open(path, std::ios::app); }
    template <class T>
    Log& operator<<(const T& v) {
        std::cout << v;
        if (file_) file_ << v;
        return *this;
    }
    Log& operator<<(std::ostream& (*m)(std::ostream&)) {
        std::cout << m;
        if (file_) file_ << m;
        return *this;
    }
    void precision(int p, bool sci) {
        for (std::ostream* o : {static_cast<std::ostream*>(&std::cout), static_cast<std::ostream*>(&file_)}) {
            if (sci)
                *o << std::scientific;
            else
                *o << std::defaultfloat;
            *o << std::setprecision(p);
        }
    }

  private:
    std::ofstream file_;
};

void check(int code, const char* what) {
    if (code != MS_OK) throw std::runtime_error(std::string(what) + ": " + ms_last_error());
}

std::string base_name(const std::string& p) { return fs::path(p).filename().string(); }

// initialize_directories.cpp:43-110: keep the previous output directory as
// prev_..., make a fresh one and copy the inputs into input_files_used/.
std::string make_output_dir(const std::string& settings, const std::string& ref, const std::string& ansatz,
                            std::string& note) {
    const fs::path cwd = fs::current_path();
    note += "Current working directory: " + cwd.string() + "\n";
    const std::string name = "output_using_" + base_name(settings);
    const fs::path dir = cwd / name, prev = cwd / ("prev_" + name);
    if (fs::exists(dir)) {
        if (fs::exists(prev)) {
            fs::remove_all(prev);
            note += "Deleted existing output directory from run-before-last: prev_" + name + "\n";
        }
        fs::rename(dir, prev);
        note += "Found a previous output directory for these settings; it was saved as prev_" + name + ".\n";
    } else {
        note += "No previous output directory was found that needed saving.\n";
    }
    fs::create_directory(dir);
    const fs::path inputs = dir / "input_files_used";
    fs::create_directory(inputs);
    fs::copy_file(settings, inputs / base_name(settings));
    fs::copy_file(ref, inputs / base_name(ref));
    if (!ansatz.empty()) fs::copy_file(ansatz, inputs / base_name(ansatz));
    return dir.string();
}

// One output's data, copied off the device; formatted and written by the
// writer thread while the device keeps stepping.
struct OutputJob {
    int64_t stepcount;
    double dial_factor;
    int status;
    std::vector<double> dofs, gauss, mean, strain, dens, stresses, deficit;
};

// Single background thread writing stepcount_<n>_output.vtk files (and the
// aux_output.txt rows) in submission order.
class OutputWriter {
  public:
    using Fn = std::function<void(const OutputJob&)>;
    explicit OutputWriter(Fn fn) : fn_(std::move(fn)), th_([this] { loop(); }) {}
    ~OutputWriter() { finish(); }
    void submit(OutputJob job) {
        std::unique_lock<std::mutex> lk(mu_);
        if (!error_.empty()) throw std::runtime_error(error_);
        q_.push_back(std::move(job));
        cv_.notify_all();
    }
    // Wait for every submitted job; rethrows the first write error.
    void finish() {
        {
            std::unique_lock<std::mutex> lk(mu_);
            done_ = true;
            cv_.notify_all();
        }
        if (th_.joinable()) th_.join();
        if (!error_.empty()) {
            const std::string e = error_;
            error_.clear();
            throw std::runtime_error(e);
        }
    }

  private:
    void loop() {
        for (;;) {
            OutputJob job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return done_ || !q_.empty(); });
                if (q_.empty()) return;
                job = std::move(q_.front());
                q_.pop_front();
            }
            try {
                fn_(job);
            } catch (const std::exception& e) {
                std::unique_lock<std::mutex> lk(mu_);
                if (error_.empty()) error_ = e.what();
            }
        }
    }
    Fn fn_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<OutputJob> q_;
    bool done_ = false;
    std::string error_;
    std::thread th_;
};

struct Run {
    RunSettings st;
    RefMesh mesh;
    std::string out_dir, aux_path;
    Log log;
    ms_scales sc{};
    double eta2 = 0.0;
    std::vector<double> areas, cents, masses, G, H;
    std::vector<int32_t> pid;
    std::vector<uint8_t> boundary;
    ms_mesh_info_out info{};
    double upper_slide_z = 0.0, lower_slide_z = 0.0;
    ms_ctx* ctx = nullptr;
    std::unique_ptr<OutputWriter> writer;
    ~Run() {
        if (writer) try {
                writer->finish();
            } catch (const std::exception& e) {
                std::cerr << e.what() << std::endl;
            }
        ms_destroy(ctx);
    }
};

ms_scales scales_for(const Run& r, double damping_prefactor, bool gradient_descent) {
    ms_scales s{};
    s.poisson_ratio = r.st.def_poisson_ratio;
    s.ref_density = r.st.ref_density;
    s.timestep_prefactor = r.st.timestep_prefactor;
    s.damping_prefactor = damping_prefactor;
    s.dial_phase_time_prefactor = r.st.dial_phase_time_prefactor;
    s.time_between_equil_checks_prefactor = r.st.time_between_equil_checks_prefactor;
    s.outputs_per_char_long_time = r.st.outputs_per_char_long_time;
    s.gradient_descent = gradient_descent ? 1 : 0;
    check(ms_calc_scales(r.mesh.num_tris, r.mesh.thickness.data(), r.mesh.shear_modulus.data(), r.areas.data(),
                         r.info.perimeter, r.info.approx_min_tri_size, &s),
          "ms_calc_scales");
    if (r.st.timesteps_per_dial_phase > 0) s.timesteps_per_dial_phase = r.st.timesteps_per_dial_phase;
    return s;
}

ms_shell_params params_for(const Run& r, const ms_scales& s, bool gradient_descent) {
    ms_shell_params p{};
    p.poisson_ratio = r.st.def_poisson_ratio;
    p.bend_stiffening_scale_factor = r.st.bend_stiffening_scale_factor;
    p.ref_density = r.st.ref_density;
    p.min_ref_thickness = s.min_ref_thickness;
    p.timestep = s.timestep;
    p.damping_factor = s.damping_factor;
    p.char_force_scale = s.char_force_scale;
    p.char_speed_scale = s.char_speed_scale;
    p.gradient_descent = gradient_descent ? 1 : 0;
    p.dial_protocol = r.st.dial_protocol == 3 ? MS_DIAL_LCE : MS_DIAL_BLEND;
    p.dial_from_ansatz = r.st.dialling_from_ansatz_rather_than_ref_state ? 1 : 0;
    return p;
}

// Formats and writes one output (writer thread): the VTK file and the
// aux_output.txt row (main.cpp:409-437, write_output.cpp).
void write_job(const Run& r, const OutputJob& j) {
    const int T = r.mesh.num_tris;
    OutputFields f{j.dofs.data(),   j.gauss.data(),    j.mean.data(),    j.dens.data(), j.dens.data() + T,
                   j.strain.data(), j.stresses.data(), j.deficit.data()};
    write_output_vtk(r.out_dir, j.stepcount, j.dial_factor, r.mesh, f);
    std::ofstream aux(r.aux_path, std::ios::app);
    if (j.stepcount == 0)
        aux << "stepcount  dial_factor  stuff.status  upper_slide_z  upper_slide_force  lower_slide_z  "
               "lower_slide_force  \n";
    aux << j.stepcount << "  " << j.dial_factor << " " << j.status << "  " << r.upper_slide_z << " " << 0.0 << " "
        << r.lower_slide_z << " " << 0.0 << " \n";
}

// The device part of an output (state, curvatures, strain, energy
// densities, Cauchy stress eigenpairs, angle deficits), copied to the host
// and handed to the writer thread.
void write_output(Run& r, int64_t stepcount, double dial_factor, int status, double time_plus_dt) {
    const int N = r.mesh.num_nodes, T = r.mesh.num_tris;
    OutputJob j{stepcount,
                dial_factor,
                status,
                std::vector<double>(3 * (size_t)N),
                std::vector<double>(T),
                std::vector<double>(T),
                std::vector<double>(T),
                std::vector<double>(2 * (size_t)T),
                std::vector<double>(8 * (size_t)T),
                std::vector<double>(N)};
    check(ms_get_state(r.ctx, 0, j.dofs.data(), nullptr), "ms_get_state");
    check(ms_calc_diagnostics(r.ctx, 0, j.gauss.data(), j.mean.data(), j.strain.data()), "ms_calc_diagnostics");
    check(ms_calc_energy_densities(r.ctx, 0, j.dens.data()), "ms_calc_energy_densities");
    check(ms_calc_stresses(r.ctx, 0, j.stresses.data()), "ms_calc_stresses");
    check(ms_calc_angle_deficits(r.ctx, 0, r.boundary.data(), j.deficit.data()), "ms_calc_angle_deficits");
    const double scale = r.sc.char_stretch_energy_density_scale;
    for (auto& d : j.dens) d /= scale;
    // write_output.cpp:139-152: eigenvalues over the reference shear modulus
    // times the reference thickness of each triangle
    for (int t = 0; t < T; ++t) {
        const double mh = r.mesh.shear_modulus[t] * r.mesh.thickness[t];
        j.stresses[t] /= mh;
        j.stresses[4 * (size_t)T + t] /= mh;
    }
    r.writer->submit(std::move(j));
    // main.cpp:437 prints the simulation time plus one timestep
    r.log << "Wrote VTK output at " << real_time() << ", stepcount = " << stepcount
          << ", simulation time = " << time_plus_dt << ", current dial factor = " << dial_factor << std::endl;
}

int run(int argc, char** argv) {
    if (argc != 3 && argc != 4) {
        std::cerr << "Error: Please provide at least one settings file path followed by one data file path as "
                     "command line arguments.\nA path to a file with ansatz degrees of freedom can be given as a "
                     "third argument if desired."
                  << std::endl;
        return -1;
    }
    const std::string settings_path = argv[1], ref_path = argv[2], ansatz_path = argc == 4 ? argv[3] : "";
    Run r;
    std::string note = "Simulation run began at: " + now_string() + "\nThe command that was run was\n";
    for (int i = 0; i < argc; ++i) note += std::string(argv[i]) + " ";
    note += "\n";
    try {
        r.out_dir = make_output_dir(settings_path, ref_path, ansatz_path, note);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\nError: initialize_directories failed." << std::endl;
        return -1;
    }
    r.log.open(r.out_dir + "/log.txt");
    r.aux_path = r.out_dir + "/aux_output.txt";
    r.log << note << std::endl;
    r.log << "Device library " << ms_version() << std::endl;

    try {
        r.st = read_run_settings(settings_path, !ansatz_path.empty());
    } catch (const std::exception& e) {
        r.log << e.what() << std::endl;
        return -1;
    }
    double dial_start = 0.0;
    std::vector<double> ansatz;
    try {
        bool cached = false;
        const auto tr = std::chrono::steady_clock::now();
        r.mesh = read_ref_vtk_cached(ref_path, r.st.ref_thickness_if_uniform, r.st.ref_shear_modulus_if_uniform,
                                     cached);
        r.log << (cached ? "Read the reference data from its binary sidecar " : "Read the reference data file ")
              << "in " << std::chrono::duration<double>(std::chrono::steady_clock::now() - tr).count() << " s"
              << std::endl;
        if (!ansatz_path.empty()) read_ansatz_vtk(ansatz_path, r.mesh.num_nodes, ansatz, dial_start);
    } catch (const std::exception& e) {
        r.log << e.what() << "\nAn error while reading data likely implies a problem with a data file, for example "
                             "a mismatch between the number of nodes or triangles stated and the number actually "
                             "present; or other similar mismatches in numbers of data values; or a format problem. "
                             "Remember the input file format is stricter than standard vtk."
              << std::endl;
        return -1;
    }
    if (r.st.dialling_from_ansatz_rather_than_ref_state) dial_start = 0.0;
    const int N = r.mesh.num_nodes, T = r.mesh.num_tris;
    r.log << "Dial factor will start from " << dial_start << std::endl;
    r.log << "Number of nodes = " << N << std::endl;
    r.log << "Number of triangles = " << T << std::endl;
    if (T < 50) r.log << "Your mesh has a small number of triangles." << std::endl;

    try {
        // calc_mesh_info, set_up_patch_fitting + calc_deriv_mats, calc_dof_masses
        r.areas.resize(T);
        r.cents.resize(3 * (size_t)T);
        r.boundary.resize(N);
        check(ms_mesh_info(N, T, r.mesh.pos.data(), r.mesh.tris.data(), r.areas.data(), r.cents.data(),
                           r.boundary.data(), &r.info),
              "calc_mesh_info");
        r.pid.resize(6 * (size_t)T);
        r.G.resize(6 * (size_t)T);
        r.H.resize(18 * (size_t)T);
        int32_t retried = 0;
        check(ms_patch_fit(N, T, r.mesh.pos.data(), r.mesh.tris.data(), r.st.patch_mat_dimless_conditioning_thresh,
                           r.pid.data(), r.G.data(), r.H.data(), &retried),
              "set_up_patch_fitting");
        r.log << "Patch fitting: " << retried << " triangles needed a second patch choice." << std::endl;
        r.masses.resize(N);
        check(ms_dof_masses(N, T, r.mesh.tris.data(), r.areas.data(), r.mesh.thickness.data(), r.st.ref_density,
                            r.masses.data()),
              "calc_dof_masses");
    } catch (const std::exception& e) {
        r.log << e.what() << std::endl;
        return -1;
    }
    bool gd = r.st.num_steps_gradient_descent > 0;
    r.sc = scales_for(r, r.st.damping_prefactor_1, gd);
    r.eta2 = scales_for(r, r.st.damping_prefactor_2, gd).damping_factor;
    r.log << "Timestep = " << r.sc.timestep << ", damping factor = " << r.sc.damping_factor
          << ", timesteps per dial phase = " << r.sc.timesteps_per_dial_phase
          << ", timesteps between equilibrium checks = " << r.sc.timesteps_between_equil_checks << std::endl;

    // perturb_node_positions_with_noise.cpp:45-61, then the ansatz if given
    std::vector<double> dofs = r.mesh.pos;
    {
        std::minstd_rand eng(123456);
        const double amp = r.info.approx_min_tri_size * 0.001;
        std::uniform_real_distribution<double> distr(-amp, amp);
        for (int n = 0; n < N; ++n)
            for (int c = 0; c < 3; ++c) dofs[(size_t)c * N + n] += distr(eng);
    }
    if (!ansatz.empty()) dofs = ansatz;
    r.upper_slide_z = *std::max_element(dofs.begin() + 2 * (size_t)N, dofs.end());
    r.lower_slide_z = *std::min_element(dofs.begin() + 2 * (size_t)N, dofs.end());

    ms_shell_desc d{};
    d.num_nodes = N;
    d.num_tris = T;
    d.patch_ids = r.pid.data();
    d.G = r.G.data();
    d.H = r.H.data();
    d.ref_pos = r.mesh.pos.data();
    d.ref_areas = r.areas.data();
    d.thickness = r.mesh.thickness.data();
    d.shear_modulus = r.mesh.shear_modulus.data();
    d.abar_info = r.mesh.abar_info.data();
    d.abar_info_cols = r.mesh.abar_cols;
    d.bbar_info = r.mesh.bbar_info.data();
    d.bbar_info_cols = r.mesh.bbar_cols;
    d.node_masses = r.masses.data();
    d.constrained = r.mesh.constrained.data();
    d.params = params_for(r, r.sc, gd);
    const int precision = r.st.precision == "mixed" ? MS_MIXED : MS_FP64;
    if (ms_create(&d, 1, precision, &r.ctx) != MS_OK) {
        r.log << "Error: " << ms_last_error() << std::endl;
        return -1;
    }
    // Launch shape by mesh size, from measured step times on discs of 4.5k to
    // 7.3M triangles.  Every shape is the fused step and gives the same bits.
    // Small meshes are launch-bound: many small tiles, one workgroup each, keep
    // most CUs busy (4.5k triangles: 16-node tiles in 128-thread workgroups,
    // 11.1 / 9.8 us per step in FP64 / mixed; 18k: 48-node tiles in 256-thread
    // workgroups, 9.6 / 8.8 us; 72k: 48-node tiles, 13.2 / 10.7 us; against
    // 12.4 / 11.9, 12.0 / 11.3 and 17.2 / 13.8 us for a triangle and a node
    // launch of 64-wide blocks).  From ~100k triangles: tiles of up to 512
    // nodes (two 256-node chunks in FP64, one chunk in mixed mode) on a
    // persistent grid of two workgroups per CU, cut smaller while they are
    // fewer than the grid's workgroups (the 290k-triangle disc: 512 tiles of
    // 285 nodes, 1.35x faster than 284 of 512).
    ms_tuning tune{};
    tune.tri_grid_per_cu = -1;
    tune.node_grid_per_cu = -1;
    tune.fused = 1;
    tune.tile_chunks = 1;
    if (T < 10000) {
        tune.node_block = 128;
        tune.tile_nodes = 16;
        tune.node_grid_per_cu = 0;
    } else if (T < 100000) {
        tune.node_block = 256;
        tune.tile_nodes = 48;
        tune.node_grid_per_cu = 0;
    } else {
        tune.tile_nodes = -2;
        tune.node_block = precision == MS_MIXED ? 512 : 256;
        tune.tile_chunks = precision == MS_MIXED ? 1 : 2;
        tune.node_grid_per_cu = 2;
        tune.fused = 1;
    }
    if (ms_set_tuning(r.ctx, &tune) != MS_OK)
        r.log << "Launch shape unavailable: " << ms_ctx_last_error(r.ctx) << std::endl;
    check(ms_set_state(r.ctx, 0, dofs.data(), nullptr), "ms_set_state");
    if (r.st.dialling_from_ansatz_rather_than_ref_state) {
        // main.cpp:321-329: a and b of the start state become the zero-dial geometry
        std::vector<double> a0(3 * (size_t)T), b0(3 * (size_t)T);
        check(ms_calc_geometry(r.ctx, 0, a0.data(), b0.data(), nullptr), "calc_a_comps_and_b_comps_and_normals");
        check(ms_set_zero_dial_geometry(r.ctx, 0, a0.data(), b0.data()), "ms_set_zero_dial_geometry");
    }
    ms_relax_params rp{};
    rp.tol_f = r.st.force_to_char_force_ratio_equil_threshold;
    rp.tol_v = r.st.speed_to_char_speed_ratio_equil_threshold;
    rp.equil_interval = r.sc.timesteps_between_equil_checks;
    rp.steps_per_dial_phase = r.sc.timesteps_per_dial_phase;
    rp.dial_spacing = r.st.dial_factors_to_hold_at_spacing;
    rp.dial_start = dial_start;
    rp.eta1 = r.sc.damping_factor;
    rp.eta2 = r.eta2;
    rp.pause_on_eq = 1;
    check(ms_relax_setup(r.ctx, 0, &rp), "ms_relax_setup");

    // every protocol check and phase end of the device run is logged as the
    // reference's loop logs it (main.cpp:303, 396-407)
    constexpr int32_t kLogCapacity = 16384;
    check(ms_set_check_log(r.ctx, kLogCapacity), "ms_set_check_log");
    const int64_t steps_per_call = (int64_t)kLogCapacity * std::max<int64_t>(1, rp.equil_interval + 1);
    // simulation time of a step: the timestep changes when gradient descent ends
    const double dt_gd = r.sc.timestep;
    const double dt_newton = gd ? scales_for(r, r.st.damping_prefactor_1, false).timestep : r.sc.timestep;
    const int64_t gd_steps = gd ? (int64_t)r.st.num_steps_gradient_descent + 1 : 0;
    auto sim_time = [&](int64_t k) {
        return k <= gd_steps ? (double)k * dt_gd : (double)gd_steps * dt_gd + (double)(k - gd_steps) * dt_newton;
    };
    auto time_plus_dt = [&](int64_t k) { return sim_time(k) + (k < gd_steps ? dt_gd : dt_newton); };
    std::vector<ms_check_record> events(kLogCapacity);
    auto drain_log = [&]() {
        int32_t n = 0;
        int64_t dropped = 0;
        check(ms_read_check_log(r.ctx, 0, events.data(), kLogCapacity, &n, &dropped), "ms_read_check_log");
        if (dropped) r.log << "(" << dropped << " protocol events not logged)" << std::endl;
        for (int32_t i = 0; i < n; ++i) {
            const ms_check_record& e = events[i];
            if (e.kind == 1) {
                r.log << "Reached dial_factor = " << e.dial_factor << ", now waiting for equilibrium." << std::endl;
                continue;
            }
            r.log << "Checking for equilibrium at " << real_time() << ", stepcount = " << e.step
                  << ", simulation time = " << sim_time(e.step) << ", dial_factor = " << e.dial_factor << std::endl;
            r.log << "Max ratio of non-damping force to characteristic force = " << e.max_force_ratio << " (node "
                  << e.max_force_node << ")" << "\n";
            r.log << "Max ratio of node speed to characteristic speed = " << e.max_speed_ratio << " (node "
                  << e.max_speed_node << ")" << "\n";
            if (e.at_equilibrium)
                r.log << "Equilibrium reached. Writing VTK output at " << real_time() << ", stepcount = " << e.step
                      << ", simulation time = " << sim_time(e.step) << ", dial_factor = " << e.dial_factor
                      << std::endl;
        }
    };

    r.writer = std::make_unique<OutputWriter>([&r](const OutputJob& j) { write_job(r, j); });
    r.log.precision(8, true);
    r.log << "Beginning dynamics loop." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    int64_t every = r.sc.steps_between_outputs;  // <= 0: outputs at equilibria only
    const int64_t gd_until = gd ? (int64_t)r.st.num_steps_gradient_descent + 1 : -1;
    const int64_t limit = r.st.max_steps > 0 ? r.st.max_steps : INT64_MAX;
    int64_t steps = 0;
    if (every > 0) write_output(r, 0, dial_start, 0, time_plus_dt(0));
    ms_relax_report rep{};
    while (true) {
        int64_t target = limit;
        if (every > 0) target = std::min(target, (steps / every + 1) * every);
        if (gd_until > steps) target = std::min(target, gd_until);
        const int64_t n = std::min(target - steps, steps_per_call);  // the log ring holds one call's events
        if (n <= 0) break;
        check(ms_relax_run(r.ctx, n, std::min<int64_t>(n, 50000), &rep), "ms_relax_run");
        drain_log();
        steps = rep.steps;
        if (rep.failed) {
            // main.cpp:440-470: output of the failing step's input, then the error
            write_output(r, steps, rep.dial_factor, rep.status, time_plus_dt(steps));
            r.writer->finish();
            r.log << "At stepcount = " << steps << ", dial_factor = " << rep.dial_factor
                  << " there was a problem: implausibly_high_force" << std::endl;
            return -1;
        }
        if (rep.halted) {
            // an equilibrium (logged with its check): the checked state is the current one
            write_output(r, rep.eq_step, rep.dial_factor, 2, time_plus_dt(rep.eq_step));
            if (rep.status == 2) break;  // dial factor 1: the run is complete
            check(ms_resume_shell(r.ctx, 0), "ms_resume_shell");
            steps = rep.eq_step + 1;  // resuming applies the checked step's update
            continue;
        }
        if (steps == gd_until) {
            // main.cpp:275-282: gradient descent ends, Newtonian dynamics from here
            r.log << "Finished doing gradient descent, now switching to Newtonian dynamics." << std::endl;
            gd = false;
            const ms_scales s2 = scales_for(r, r.st.damping_prefactor_1, false);
            ms_shell_params p2 = params_for(r, s2, false);
            p2.damping_factor = rep.status == 0 ? rp.eta1 : rp.eta2;  // the current phase's damping
            check(ms_set_shell_params(r.ctx, 0, &p2), "ms_set_shell_params");
            every = s2.steps_between_outputs;
        }
        if (every > 0 && steps % every == 0)
            write_output(r, steps, rep.dial_factor, rep.status, time_plus_dt(steps));
        if (steps >= limit) {
            r.log << "Stopped after max_steps = " << limit << " steps." << std::endl;
            break;
        }
    }
    r.writer->finish();  // every output written
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double E[3] = {0, 0, 0};
    check(ms_calc_energies(r.ctx, 0, E), "calc_energies");
    r.log << "Simulation completed at time = " << sim_time(rep.steps) << ", stepcount = " << rep.steps << std::endl;
    r.log << "(" << secs << " s of dynamics on the device; energies " << E[0] << " " << E[1] << " " << E[2] << ")"
          << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return -1;
    }
}
