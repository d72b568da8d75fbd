// Sanitizer harness of the device context's host bookkeeping: builds small
// shells with the library's own set-up calls, then drives the C-ABI through
// the launch shapes, graph captures, running-shell tables, tile rebuilds and
// allocation failures the product uses, on the host-memory HIP stand-in
// (mock_runtime.cpp, mock_kernels.cpp).  Built with g++
// -fsanitize=address,undefined; a failed check aborts with a message naming
// the launch, and AddressSanitizer reports any host-side overflow or
// use-after-free.  Prints "HOSTCHECK OK" at the end.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "hip/hip_runtime.h"
#include "morphoshell_amd.h"

extern "C" void mock_schedule_halt(int s, long long step);
extern "C" void mock_clear_halts();

namespace {

#define REQUIRE(cond, ...)                                   \
    do {                                                     \
        if (!(cond)) {                                       \
            std::fprintf(stderr, "HOSTCHECK FAILED: ");      \
            std::fprintf(stderr, __VA_ARGS__);               \
            std::fprintf(stderr, " (%s:%d)\n", __FILE__, __LINE__); \
            std::exit(1);                                    \
        }                                                    \
    } while (0)

#define OK(call) REQUIRE((call) == MS_OK, "%s -> %s", #call, ms_last_error())

struct Shell {
    int N = 0, T = 0;
    std::vector<double> pos;  // 3N SoA
    std::vector<int32_t> tri;
    std::vector<int32_t> pid;
    std::vector<double> G, H, area, cen, h, mu, ainfo, binfo, mass;
    std::vector<uint8_t> clamp;
    ms_shell_params params{};
    ms_shell_desc desc() const {
        ms_shell_desc d{};
        d.num_nodes = N;
        d.num_tris = T;
        d.patch_ids = pid.data();
        d.G = G.data();
        d.H = H.data();
        d.ref_pos = pos.data();
        d.ref_areas = area.data();
        d.thickness = h.data();
        d.shear_modulus = mu.data();
        d.abar_info = ainfo.data();
        d.abar_info_cols = 3;
        d.bbar_info = binfo.data();
        d.bbar_info_cols = 3;
        d.node_masses = mass.data();
        d.constrained = clamp.data();
        d.params = params;
        return d;
    }
};

// An n x n jittered grid on the unit square, each cell split along a
// diagonal, set up with the library's own host calls.
Shell make_shell(int n, unsigned seed, bool lce = false) {
    Shell s;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(-0.15, 0.15);
    s.N = n * n;
    s.pos.assign(3 * (size_t)s.N, 0.0);
    const double d = 1.0 / (n - 1);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int k = j * n + i;
            const bool edge = i == 0 || j == 0 || i == n - 1 || j == n - 1;
            s.pos[k] = i * d + (edge ? 0.0 : u(rng) * d);
            s.pos[s.N + k] = j * d + (edge ? 0.0 : u(rng) * d);
        }
    for (int j = 0; j + 1 < n; ++j)
        for (int i = 0; i + 1 < n; ++i) {
            const int a = j * n + i, b = a + 1, c = a + n, e = c + 1;
            s.tri.insert(s.tri.end(), {a, b, e, a, e, c});
        }
    s.T = (int)s.tri.size() / 3;
    const int T = s.T;
    s.area.resize(T);
    s.cen.resize(3 * (size_t)T);
    ms_mesh_info_out mi{};
    OK(ms_mesh_info(s.N, T, s.pos.data(), s.tri.data(), s.area.data(), s.cen.data(), nullptr, &mi));
    s.pid.resize(6 * (size_t)T);
    s.G.resize(6 * (size_t)T);
    s.H.resize(18 * (size_t)T);
    OK(ms_patch_fit(s.N, T, s.pos.data(), s.tri.data(), 1000.0, s.pid.data(), s.G.data(), s.H.data(), nullptr));
    s.h.assign(T, 0.01);
    s.mu.assign(T, 1.0);
    s.ainfo.resize(3 * (size_t)T);
    s.binfo.assign(3 * (size_t)T, 0.0);
    for (int t = 0; t < T; ++t) {
        if (lce) {
            s.ainfo[3 * t] = std::atan2(s.cen[T + t] - 0.5, s.cen[t] - 0.5) + M_PI / 2;
            s.ainfo[3 * t + 1] = 0.9;
            s.ainfo[3 * t + 2] = 0.5;
        } else {
            const double l = 1.0 + 0.05 * std::sin(6.0 * s.cen[t]) * std::cos(5.0 * s.cen[T + t]);
            s.ainfo[3 * t] = l * l;
            s.ainfo[3 * t + 1] = 0.0;
            s.ainfo[3 * t + 2] = l * l;
        }
    }
    s.mass.resize(s.N);
    OK(ms_dof_masses(s.N, T, s.tri.data(), s.area.data(), s.h.data(), 1.0, s.mass.data()));
    s.clamp.assign(s.N, 0);
    for (int i = 0; i < n; ++i) s.clamp[i] = 1;  // one clamped edge
    ms_scales sc{};
    sc.poisson_ratio = 0.5;
    sc.ref_density = 1.0;
    sc.timestep_prefactor = 1.0;
    sc.damping_prefactor = 1.0;
    sc.dial_phase_time_prefactor = 1.0;
    sc.time_between_equil_checks_prefactor = 0.1;
    sc.outputs_per_char_long_time = 1.0;
    OK(ms_calc_scales(T, s.h.data(), s.mu.data(), s.area.data(), mi.perimeter, mi.approx_min_tri_size, &sc));
    ms_shell_params& p = s.params;
    p.poisson_ratio = 0.5;
    p.bend_stiffening_scale_factor = 1e17;
    p.ref_density = 1.0;
    p.min_ref_thickness = sc.min_ref_thickness;
    p.timestep = sc.timestep;
    p.damping_factor = sc.damping_factor;
    p.char_force_scale = sc.char_force_scale;
    p.char_speed_scale = sc.char_speed_scale;
    p.dial_protocol = lce ? MS_DIAL_LCE : MS_DIAL_BLEND;
    return s;
}

ms_tuning tuning(int tb, int nb, int tg, int ng, int gs, int fused, int chunks, int tile_nodes = 0) {
    ms_tuning t{};
    t.tile_nodes = tile_nodes;
    t.tri_block = tb;
    t.node_block = nb;
    t.tri_grid_per_cu = tg;
    t.node_grid_per_cu = ng;
    t.graph_steps = gs;
    t.fused = fused;
    t.tile_chunks = chunks;
    return t;
}

void set_states(ms_ctx* c, const std::vector<Shell>& shells) {
    for (int s = 0; s < (int)shells.size(); ++s) OK(ms_set_state(c, s, shells[s].pos.data(), nullptr));
}

void exercise_ops(ms_ctx* c, const std::vector<Shell>& shells) {
    std::vector<ms_step_report> rep(shells.size());
    OK(ms_calc_deformation_forces(c));
    OK(ms_calc_non_deformation_forces(c));
    OK(ms_impose_constraints(c));
    REQUIRE(ms_advance_dynamics(c, rep.data()) == MS_OK, "advance");
    for (int s = 0; s < (int)shells.size(); ++s) {
        const Shell& sh = shells[s];
        double e[3];
        OK(ms_calc_energies(c, s, e));
        ms_step_report r{};
        OK(ms_is_at_equilibrium(c, s, 1e-6, 1e-5, &r));
        std::vector<double> a(3 * (size_t)sh.T), b(3 * (size_t)sh.T), n(3 * (size_t)sh.T), f(3 * (size_t)sh.N);
        OK(ms_calc_geometry(c, s, a.data(), b.data(), n.data()));
        OK(ms_calc_diagnostics(c, s, a.data(), b.data(), n.data()));
        std::vector<double> st(8 * (size_t)sh.T), ad(sh.N);
        OK(ms_calc_stresses(c, s, st.data()));
        std::vector<uint8_t> bnd(sh.N, 0);
        OK(ms_calc_angle_deficits(c, s, bnd.data(), ad.data()));
        OK(ms_calc_energy_densities(c, s, a.data()));
        OK(ms_get_forces(c, s, f.data()));
        OK(ms_set_forces(c, s, f.data()));
        std::vector<double> x(3 * (size_t)sh.N), v(3 * (size_t)sh.N);
        OK(ms_get_state(c, s, x.data(), v.data()));
    }
}

// The launch shape of the profiled small-disc run: 64-wide blocks, persistent
// grids of 8 per CU, 16-step graphs, two-launch and fused forms, FP64 and
// mixed; then the driver's own shapes.
void scenario_small_shapes() {
    std::vector<Shell> sh{make_shell(48, 1)};  // 4418 triangles
    for (int prec : {MS_FP64, MS_MIXED}) {
        ms_ctx* c = nullptr;
        std::vector<ms_shell_desc> d{sh[0].desc()};
        OK(ms_create(d.data(), 1, prec, &c));
        set_states(c, sh);
        for (int fused : {0, 1}) {
            ms_tuning t = tuning(64, 64, 8, 8, 16, fused, 1);
            OK(ms_set_tuning(c, &t));
            std::vector<ms_step_report> rep(1);
            OK(ms_step(c, 100, 0, 0.0, 0.0, 0, rep.data()));
            OK(ms_step(c, 37, 10, 1e-6, 1e-5, 0, rep.data()));
            OK(ms_enable_timing(c, 1));
            OK(ms_step(c, 5, 0, 0.0, 0.0, 0, rep.data()));
            OK(ms_enable_timing(c, 0));
        }
        for (int nb : {128, 256, 512})
            for (int tn : {-1, -2, 100, -1}) {
                // full tiles, balanced over the grid, 100-node groups of
                // partial chunks, and back
                ms_tuning t = tuning(256, nb, 0, 2, 16, 1, nb == 512 ? 1 : 2, tn);
                OK(ms_set_tuning(c, &t));
                OK(ms_step(c, 40, 0, 0.0, 0.0, 0, nullptr));
            }
        OK(ms_set_precision(c, prec == MS_FP64 ? MS_MIXED : MS_FP64));
        OK(ms_step(c, 20, 0, 0.0, 0.0, 0, nullptr));
        exercise_ops(c, sh);
        ms_destroy(c);
    }
    std::printf("small shapes ok\n");
}

// A ragged batch whose shells halt one after another (emulated), over
// persistent and one-tile-per-workgroup grids and one- and two-chunk tiles:
// the running-shell tables are rebuilt while graphs come and go.
void scenario_batch_halts() {
    std::vector<Shell> sh{make_shell(20, 2), make_shell(33, 3), make_shell(12, 4), make_shell(27, 5, true)};
    std::vector<ms_shell_desc> d;
    for (auto& s : sh) d.push_back(s.desc());
    for (int fused : {0, 1})
        for (int grid : {0, 2, 8})
            for (int chunks : {1, 2}) {
                ms_ctx* c = nullptr;
                OK(ms_create(d.data(), (int)d.size(), MS_MIXED, &c));
                set_states(c, sh);
                ms_tuning t = tuning(64, 128, grid, grid, 8, fused, chunks, chunks == 2 ? -2 : -1);
                OK(ms_set_tuning(c, &t));
                mock_clear_halts();
                mock_schedule_halt(1, 30);
                mock_schedule_halt(3, 55);
                mock_schedule_halt(0, 80);
                std::vector<ms_step_report> rep(sh.size());
                for (int k = 0; k < 12; ++k) OK(ms_step(c, 11, 0, 0.0, 0.0, 0, rep.data()));
                REQUIRE(rep[0].halted && rep[1].halted && rep[3].halted && !rep[2].halted, "halts not applied");
                OK(ms_resume_shell(c, 1));
                OK(ms_step(c, 9, 0, 0.0, 0.0, 0, rep.data()));
                mock_clear_halts();
                for (int s = 0; s < (int)sh.size(); ++s) OK(ms_resume_shell(c, s));
                OK(ms_step(c, 9, 0, 0.0, 0.0, 0, rep.data()));
                exercise_ops(c, sh);
                ms_destroy(c);
            }
    std::printf("batch halts ok\n");
}

// Balanced tiles (tile_nodes -2) on a device of 1 CU: the running shells'
// tiles are re-cut smaller as shells halt and full again when all run.
void scenario_balanced_retile() {
    setenv("MOCK_NUM_CUS", "1", 1);
    std::vector<Shell> sh{make_shell(30, 11), make_shell(24, 12), make_shell(12, 13), make_shell(26, 14)};
    std::vector<ms_shell_desc> d;
    for (auto& s : sh) d.push_back(s.desc());
    for (int prec : {MS_FP64, MS_MIXED}) {
        ms_ctx* c = nullptr;
        OK(ms_create(d.data(), (int)d.size(), prec, &c));
        set_states(c, sh);
        ms_tuning t = tuning(64, 128, 0, 0, 8, 1, 2, -2);
        OK(ms_set_tuning(c, &t));
        ms_tuning got{};
        OK(ms_get_tuning(c, &got));
        REQUIRE(got.tile_nodes == 256, "balanced tiles on 1 CU: %d nodes, expected full 256", got.tile_nodes);
        mock_clear_halts();
        mock_schedule_halt(0, 20);
        mock_schedule_halt(1, 20);
        mock_schedule_halt(3, 20);
        std::vector<ms_step_report> rep(sh.size());
        for (int k = 0; k < 6; ++k) OK(ms_step(c, 9, 0, 0.0, 0.0, 0, rep.data()));
        OK(ms_step(c, 9, 0, 0.0, 0.0, 0, rep.data()));  // the halts are seen: re-cut
        OK(ms_get_tuning(c, &got));
        REQUIRE(got.tile_nodes < 256 && got.fused == 1, "running shell's tiles not re-cut: %d nodes", got.tile_nodes);
        mock_clear_halts();
        for (int s = 0; s < (int)sh.size(); ++s) OK(ms_resume_shell(c, s));
        OK(ms_step(c, 9, 0, 0.0, 0.0, 0, rep.data()));
        OK(ms_get_tuning(c, &got));
        REQUIRE(got.tile_nodes == 256, "tiles not full again with every shell running: %d", got.tile_nodes);
        exercise_ops(c, sh);
        ms_destroy(c);
    }
    unsetenv("MOCK_NUM_CUS");
    std::printf("balanced retile ok\n");
}

// Dial-from-ansatz geometry set while the fused step's graph is live, and the
// device-run relaxation protocol.
void scenario_zero_dial_and_relax() {
    std::vector<Shell> sh{make_shell(30, 6), make_shell(25, 7)};
    std::vector<ms_shell_desc> d{sh[0].desc(), sh[1].desc()};
    ms_ctx* c = nullptr;
    OK(ms_create(d.data(), 2, MS_FP64, &c));
    set_states(c, sh);
    ms_tuning t = tuning(128, 128, 0, 0, 16, 1, 1);
    OK(ms_set_tuning(c, &t));
    OK(ms_step(c, 40, 0, 0.0, 0.0, 0, nullptr));
    std::vector<double> a0(3 * (size_t)sh[1].T, 1.0), b0(3 * (size_t)sh[1].T, 0.0);
    OK(ms_set_zero_dial_geometry(c, 1, a0.data(), b0.data()));
    OK(ms_step(c, 40, 0, 0.0, 0.0, 0, nullptr));
    ms_relax_params rp{};
    rp.steps_per_dial_phase = 50;
    rp.dial_spacing = 0.5;
    rp.dial_start = 0.0;
    rp.equil_interval = 10;
    rp.eta1 = sh[0].params.damping_factor;
    rp.eta2 = sh[0].params.damping_factor;
    rp.tol_f = 1e-6;
    rp.tol_v = 1e-5;
    for (int s = 0; s < 2; ++s) OK(ms_relax_setup(c, s, &rp));
    std::vector<ms_relax_report> rr(2);
    OK(ms_relax_run(c, 300, 64, rr.data()));
    ms_destroy(c);
    std::printf("zero-dial and relax ok\n");
}

// The allocation-free geometry call: no hipMalloc on its first call or
// later ones, its launch writes inside the scratch array, a second call with
// a smaller shell follows a larger one, and a shell over the limit is refused
// without touching the device.
void scenario_noalloc() {
    std::vector<Shell> sh{make_shell(40, 9), make_shell(14, 10)};
    std::vector<ms_shell_desc> d{sh[0].desc(), sh[1].desc()};
    ms_ctx* c = nullptr;
    OK(ms_create(d.data(), 2, MS_FP64, &c));
    set_states(c, sh);
    for (int s : {0, 1, 0}) {
        std::vector<double> a(3 * (size_t)sh[s].T), b(a.size()), n(a.size());
        const long long before = mock::malloc_calls();
        OK(ms_calc_geometry_noalloc(c, s, a.data(), b.data(), n.data()));
        REQUIRE(mock::malloc_calls() == before, "ms_calc_geometry_noalloc called hipMalloc");
        OK(ms_calc_geometry_noalloc(c, s, nullptr, b.data(), nullptr));
        REQUIRE(mock::malloc_calls() == before, "ms_calc_geometry_noalloc called hipMalloc");
    }
    ms_destroy(c);
    std::printf("allocation-free geometry ok\n");
}

// Every hipMalloc that a tuning change or a dial-geometry update makes is
// failed in turn; the call must report the error and leave a context whose
// later steps touch only live buffers (the replayed graph re-checks them).
void scenario_alloc_failures() {
    std::vector<Shell> sh{make_shell(26, 8)};
    std::vector<ms_shell_desc> d{sh[0].desc()};
    int failures = 0;
    for (int k = 0; k < 40; ++k) {
        ms_ctx* c = nullptr;
        OK(ms_create(d.data(), 1, MS_FP64, &c));
        set_states(c, sh);
        ms_tuning t0 = tuning(128, 128, 0, 0, 8, 1, 1);
        OK(ms_set_tuning(c, &t0));
        OK(ms_step(c, 20, 0, 0.0, 0.0, 0, nullptr));  // a live graph over the tiles
        const long long before = mock::live_allocations();
        mock::fail_malloc_after(k);
        std::vector<double> a0(3 * (size_t)sh[0].T, 1.0), b0(3 * (size_t)sh[0].T, 0.0);
        int r1 = ms_set_zero_dial_geometry(c, 0, a0.data(), b0.data());
        mock::fail_malloc_after(k);
        ms_tuning t1 = tuning(64, 256, 0, 2, 8, 1, 2);
        int r2 = ms_set_tuning(c, &t1);
        mock::fail_malloc_after(-1);
        failures += (r1 != MS_OK) + (r2 != MS_OK);
        // whatever failed, the context keeps working (or reports an error)
        const int r3 = ms_step(c, 20, 0, 0.0, 0.0, 0, nullptr);
        REQUIRE(r3 == MS_OK || r3 == MS_ERR_HIP, "step after a failed allocation: %d", r3);
        (void)before;
        ms_destroy(c);
    }
    // creation itself under every allocation failure
    for (int k = 0; k < 40; ++k) {
        const long long live = mock::live_allocations();
        ms_ctx* c = nullptr;
        mock::fail_malloc_after(k);
        const int r = ms_create(d.data(), 1, MS_FP64, &c);
        mock::fail_malloc_after(-1);
        if (r == MS_OK) ms_destroy(c);
        REQUIRE(mock::live_allocations() == live, "ms_create leaked %lld allocations after failure %d",
                mock::live_allocations() - live, k);
    }
    REQUIRE(failures > 0, "no allocation failure was injected");
    std::printf("allocation failures ok (%d calls failed)\n", failures);
}

}  // namespace

int main() {
    scenario_small_shapes();
    scenario_batch_halts();
    scenario_balanced_retile();
    scenario_zero_dial_and_relax();
    scenario_alloc_failures();
    scenario_noalloc();
    REQUIRE(mock::live_allocations() == 0, "%lld device allocations leaked", mock::live_allocations());
    std::printf("HOSTCHECK OK\n");
    return 0;
}
