"""The host driver ms_run (morphoshell_amd/driver): the reference's command
line, settings file, VTK input and output directory, with the relaxation on
the device.

CPU: usage and settings errors carry the reference's messages; the inputs are
read and the set-up runs before the device is needed.
GPU: a full run stops at the equilibrium at dial factor 1 with the same
state, bit for bit, as the Python context running the same protocol, and the
written output holds that state and the diagnostics.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import grf_disc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "morphoshell_amd", "_lib", "ms_run")

SETTINGS = """// settings of a small test run
outputs_per_char_long_time = -1.0;
ref_thickness_if_uniform = -1.0;
timestep_prefactor = 0.6;
dialling_from_ansatz_rather_than_ref_state = false;
dial_factors_to_hold_at_spacing = {spacing};
dial_phase_time_prefactor = 0.25;
time_between_equil_checks_prefactor = 0.01;
damping_prefactor_1 = 1.0;
damping_prefactor_2 = 1.0;
num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics = 0;
force_to_char_force_ratio_equil_threshold = 0.000001;
speed_to_char_speed_ratio_equil_threshold = 0.00001;
def_poisson_ratio = 0.5; // comment after a setting
ref_shear_modulus_if_uniform = -1.0;
ref_density = 1.0;
bend_stiffening_scale_factor = 100000000000000000.0;
patch_mat_dimless_conditioning_thresh = 10.0;
slide_stiffness_prefactor = 0.0;
slide_friction_coefficent = 0.0;
slide_speed_prefactor = 0.0;
using_initial_glass_slide_z_coords_from_settings_file = false;
initial_upper_slide_z_coord = 12345.0;
initial_lower_slide_z_coord = 12345.0;
timesteps_per_dial_phase = 1500;
"""


def _mesh():
    m = grf_disc(0.09, seed=4, h=0.03, kappa=0.6)
    return m


def _ansatz(path, m, seed=3):
    rng = np.random.default_rng(seed)
    P = m.points.copy()
    P[:, 2] = 1e-3 * rng.standard_normal(len(P))
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 4.2\ndial_factor = 0.00000000000000e+00, stepcount = 0\nASCII\n"
                f"DATASET POLYDATA\nPOINTS {len(P)} double\n")
        np.savetxt(f, P, fmt="%.17e")
    return np.ascontiguousarray(P.T.reshape(-1))


def _inputs(tmp_path, spacing=1.0, settings=None):
    from morphoshell_amd.meshing import write_vtk

    m = _mesh()
    ref = tmp_path / "disc.vtk"
    write_vtk(str(ref), m)
    st = tmp_path / "settings_test.txt"
    st.write_text(settings if settings is not None else SETTINGS.format(spacing=spacing))
    ans = tmp_path / "ansatz.vtk"
    x0 = _ansatz(str(ans), m)
    return m, ref, st, ans, x0


def _run(tmp_path, *args, timeout=600):
    return subprocess.run([EXE, *map(str, args)], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)


def _read_output(path, N):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "# vtk DataFile Version 4.2"
    assert lines[4] == f"POINTS {N} double"
    P = np.array([[float(v) for v in ln.split()] for ln in lines[5:5 + N]])
    return lines, np.ascontiguousarray(P.T.reshape(-1))


def test_driver_usage_error(tmp_path):
    r = _run(tmp_path)
    assert r.returncode != 0
    assert "Please provide at least one settings file path" in r.stderr


def test_driver_rejects_a_missing_semicolon(tmp_path):
    bad = SETTINGS.format(spacing=1.0).replace("timestep_prefactor = 0.6;", "timestep_prefactor = 0.6")
    _, ref, st, _, _ = _inputs(tmp_path, settings=bad)
    r = _run(tmp_path, st, ref)
    assert r.returncode != 0
    assert "did not have equal numbers of = and ;" in r.stdout
    log = (tmp_path / "output_using_settings_test.txt" / "log.txt").read_text()
    assert "did not have equal numbers of = and ;" in log


def test_driver_rejects_out_of_order_keys(tmp_path):
    s = SETTINGS.format(spacing=1.0)
    a, b = "damping_prefactor_1 = 1.0;", "damping_prefactor_2 = 1.0;"
    bad = s.replace(a, "@@").replace(b, a).replace("@@", b)
    _, ref, st, _, _ = _inputs(tmp_path, settings=bad)
    r = _run(tmp_path, st, ref)
    assert r.returncode != 0
    assert "something went wrong when reading the damping_prefactor_1 setting" in r.stdout


def test_driver_reads_inputs_and_keeps_the_previous_output(tmp_path):
    m, ref, st, ans, _ = _inputs(tmp_path)
    for _ in range(2):
        r = _run(tmp_path, st, ref, ans)
    out = tmp_path / "output_using_settings_test.txt"
    log = (out / "log.txt").read_text()
    assert f"Number of nodes = {m.num_nodes}" in log and f"Number of triangles = {m.num_tris}" in log
    assert "Timestep = " in log
    for name in ("settings_test.txt", "disc.vtk", "ansatz.vtk"):
        assert (out / "input_files_used" / name).exists()
    assert (tmp_path / "prev_output_using_settings_test.txt").is_dir()
    # without a device the run stops at the context, with the device it completes
    assert r.returncode == 0 or "no HIP device" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("spacing", [1.0, 0.5])
def test_driver_run_matches_context_relax(tmp_path, spacing):
    from morphoshell_amd.shell import Context, prepare_shell, read_settings

    m, ref, st, ans, x0 = _inputs(tmp_path, spacing=spacing)
    r = _run(tmp_path, st, ref, ans)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = tmp_path / "output_using_settings_test.txt"
    files = sorted(out.glob("stepcount_*_output.vtk"), key=lambda p: int(p.name.split("_")[1]))
    n_eq = 1 if spacing == 1.0 else 2
    assert len(files) == n_eq
    last_step = int(files[-1].name.split("_")[1])
    lines, x_drv = _read_output(files[-1], m.num_nodes)
    assert lines[1].startswith("dial_factor = 1.00000000000000e+00")
    assert "SCALARS gauss_curv double 1" in lines and "SCALARS angleDeficit double 1" in lines
    # write_output.cpp:136-156: the Cauchy stress block between the cell
    # scalars and the point data
    T, N = m.num_tris, m.num_nodes
    i = lines.index("FIELD cauchy_stress_info 4")
    assert lines[i - T - 2] == "SCALARS strain_measure double 1"
    blocks = {}
    j = i + 1
    for name, nc in (("dimless_cauchy_stress_eigval_1", 1), ("cauchy_stress_eigvec_1", 3),
                     ("dimless_cauchy_stress_eigval_2", 1), ("cauchy_stress_eigvec_2", 3)):
        assert lines[j] == f"{name} {nc} {T} double"
        blocks[name] = np.array([[float(v) for v in ln.split()] for ln in lines[j + 1:j + 1 + T]])
        j += T + 1
    assert lines[j] == f"POINT_DATA {N}"
    deficit = np.array([float(v) for v in lines[j + 3:j + 3 + N]])

    # the same protocol through the Python context, from the same start
    sh = prepare_shell(m, read_settings(st.read_text()))
    with Context(sh, precision="fp64") as ctx:
        ctx.set_state(0, x0)
        ctx.relax_setup(0, pause_on_eq=True)
        for _ in range(n_eq):
            rep = ctx.relax_run(2_000_000, poll_every=50000)[0]
            assert rep["halted"] and not rep["failed"]
            eq_step = rep["eq_step"]
            x_ctx, _ = ctx.get_state(0)
            if rep["status"] != 2:
                ctx.resume(0)
        assert rep["status"] == 2
        # the output's stress eigenpairs and angle deficits are the context's at that state
        S = ctx.calc_stresses(0)
        d_ctx = ctx.calc_angle_deficits(0)
    assert eq_step == last_step
    # the output prints 15 significant digits
    assert np.allclose(x_drv, x_ctx, rtol=0, atol=1e-13 * np.abs(x_ctx).max() + 1e-300)
    mh = sh.shear_modulus * sh.thickness
    for r, (ev, vec) in enumerate((("dimless_cauchy_stress_eigval_1", "cauchy_stress_eigvec_1"),
                                   ("dimless_cauchy_stress_eigval_2", "cauchy_stress_eigvec_2"))):
        lam = S[4 * r] / mh
        assert np.allclose(blocks[ev][:, 0], lam, rtol=1e-13, atol=1e-13 * np.abs(lam).max())
        assert np.allclose(blocks[vec], S[4 * r + 1:4 * r + 4].T, rtol=1e-13, atol=1e-13)
    assert np.allclose(deficit, d_ctx, rtol=1e-13, atol=1e-13)
    # log.txt carries the reference's per-check lines (main.cpp:396-407,
    # is_at_equilibrium.cpp:89-90) and phase ends (main.cpp:303)
    log = (out / "log.txt").read_text().split("\n")
    checks = [i for i, ln in enumerate(log) if ln.startswith("Checking for equilibrium at ")]
    assert len(checks) >= n_eq
    for i in checks:
        assert log[i + 1].startswith("Max ratio of non-damping force to characteristic force = ")
        assert log[i + 2].startswith("Max ratio of node speed to characteristic speed = ")
    reached = [ln for ln in log if ln.startswith("Equilibrium reached. Writing VTK output at ")]
    assert len(reached) == n_eq and f"stepcount = {last_step}," in reached[-1]
    assert sum(ln.endswith(", now waiting for equilibrium.") for ln in log) == n_eq
    assert sum(ln.startswith("Wrote VTK output at ") for ln in log) == n_eq
    assert any(ln.startswith(f"Simulation completed at time = ") and ln.endswith(f"stepcount = {last_step}")
               for ln in log)


@pytest.mark.gpu
def test_driver_regular_outputs(tmp_path):
    """outputs_per_char_long_time > 0: an output at step 0 and every
    steps_between_outputs steps (written by the writer thread while the
    device steps on), plus the equilibrium output."""
    settings = SETTINGS.format(spacing=1.0).replace("outputs_per_char_long_time = -1.0;",
                                                    "outputs_per_char_long_time = 20.0;")
    m, ref, st, ans, _ = _inputs(tmp_path, settings=settings)
    r = _run(tmp_path, st, ref, ans)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = tmp_path / "output_using_settings_test.txt"
    steps = sorted(int(p.name.split("_")[1]) for p in out.glob("stepcount_*_output.vtk"))
    assert steps[0] == 0 and len(steps) >= 3
    every = steps[1]
    regular, last = steps[:-1], steps[-1]
    assert regular == [k * every for k in range(len(regular))]
    assert last > regular[-1]
    aux = (out / "aux_output.txt").read_text().strip().split("\n")
    assert aux[0].startswith("stepcount") and len(aux) == len(steps) + 1
    assert [int(ln.split()[0]) for ln in aux[1:]] == steps
    for s in steps:
        lines, x = _read_output(out / f"stepcount_{s}_output.vtk", m.num_nodes)
        assert np.isfinite(x).all()
        assert lines[-1] == "" and "POINT_DATA" in " ".join(lines[-m.num_nodes - 4:])


def test_driver_binary_sidecar(tmp_path):
    """The first run writes <mesh>.msbin; the next one reads it; a changed
    text file is parsed again."""
    import time

    m, ref, st, ans, _ = _inputs(tmp_path)
    logs = []
    for k in range(3):
        if k == 2:
            time.sleep(0.01)
            ref.write_text(ref.read_text())  # same content, new modification time
        _run(tmp_path, st, ref, ans)
        logs.append((tmp_path / "output_using_settings_test.txt" / "log.txt").read_text())
    assert (tmp_path / "disc.vtk.msbin").exists()
    assert "Read the reference data file in" in logs[0]
    assert "Read the reference data from its binary sidecar" in logs[1]
    assert "Read the reference data file in" in logs[2]
    for log in logs:
        assert f"Number of nodes = {m.num_nodes}" in log and f"Number of triangles = {m.num_tris}" in log
