"""GPU parity of the per-step path against the FP64 oracle, through the C-ABI.

Tolerances (stated per test):
  * FP64 forces: max |F_gpu - F_orc| <= 1e-10 * max |F_orc|;
  * FP64 energies: 1e-11 relative; geometry 1e-12 (a) / 1e-9 (b, from H ~ 1/s^2);
  * elementwise dynamics given the same forces: bit-identical;
  * mixed mode: force error <= 1e-6 * max |F| (measured 1e-7 .. 4e-7; FP32
    coefficients, normal, stresses and contributions, with the tangent
    vectors, second derivatives and strains formed in FP64), printed for the
    record.
"""
import numpy as np
import pytest

from conftest import deformed_state, grf_disc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _shell(spacing=0.04, seed=0, h=0.01, kappa=0.5, gamma=40.0, clamp_every=0, protocol=1, **kw):
    from morphoshell_amd.meshing import lce_defect_abar_info
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(spacing, seed=seed, h=h, kappa=kappa)
    if protocol == 3:
        m.abar_info = lce_defect_abar_info(m.centroids(), 0.9, 0.5)
    if clamp_every:
        cl = np.zeros(m.num_nodes, np.uint8)
        cl[::clamp_every] = 1
        m.constrained = cl
    st = Settings(bend_stiffening_scale_factor=gamma, dial_protocol=protocol, **kw)
    return prepare_shell(m, st)


@pytest.fixture(scope="module")
def sh():
    return _shell(clamp_every=11)


def _ctx(shells, precision="fp64"):
    from morphoshell_amd.shell import Context

    return Context(shells, precision=precision)


def test_forces_fp64_match_oracle(sh):
    x, v = deformed_state(sh, seed=3)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        ctx.calc_deformation_forces()
        Fg = ctx.get_forces(0)
    Fo = orc.System.from_shell(sh).deformation_forces(x)
    err = np.abs(Fg - Fo).max() / np.abs(Fo).max()
    print(f"FP64 force rel err {err:.3e}")
    assert err <= 1e-10


def test_full_force_damping_clamp_match_oracle(sh):
    x, v = deformed_state(sh, seed=4)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        ctx.calc_deformation_forces()
        ctx.calc_non_deformation_forces()
        ctx.impose_constraints()
        Fg = ctx.get_forces(0)
    Fo = orc.System.from_shell(sh).full_forces(x, v)
    assert np.abs(Fg - Fo).max() <= 1e-10 * np.abs(Fo).max()
    cl = sh.mesh.constrained.astype(bool)
    assert np.all(Fg.reshape(3, -1)[:, cl] == 0)


def test_damping_clamp_and_advance_bitwise(sh):
    """Given identical forces, damping, clamp and update round exactly as the CPU path."""
    x, v = deformed_state(sh, seed=5)
    rng = np.random.default_rng(0)
    f = rng.normal(size=3 * sh.N) * sh.scales["char_force_scale"]
    sysm = orc.System.from_shell(sh)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        ctx.set_forces(0, f)
        ctx.calc_non_deformation_forces()
        ctx.impose_constraints()
        fg = ctx.get_forces(0)
        fo = f.copy()
        sysm.damping(v, fo)
        sysm.constraints(fo)
        assert np.array_equal(fg, fo)
        ctx.advance_dynamics()
        xg, vg = ctx.get_state(0)
    xo, vo = x.copy(), v.copy()
    d, _ = sysm.advance(xo, vo, fo)
    assert d == -1
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo)


def test_energies_and_densities_match_oracle(sh):
    x, v = deformed_state(sh, seed=6)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        Eg = ctx.calc_energies(0)
        Dg = ctx.calc_energy_densities(0)
    sysm = orc.System.from_shell(sh)
    Eo = sysm.energies(x, v)
    Do = sysm.energy_densities(x)
    print("energies gpu", Eg, "oracle", Eo)
    assert np.allclose(Eg, Eo, rtol=1e-11, atol=0)
    assert np.abs(Dg - Do).max() <= 1e-10 * np.abs(Do).max()


def test_geometry_matches_oracle(sh):
    x, _ = deformed_state(sh, seed=7)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x)
        a, b, n = ctx.calc_a_comps_and_b_comps_and_normals(0)
    ao, bo, no = orc.System.from_shell(sh).geometry(x)
    assert np.abs(a - ao).max() <= 1e-12
    assert np.abs(n - no).max() <= 1e-12
    assert np.abs(b - bo).max() <= 1e-9 * max(1.0, np.abs(bo).max())


def test_equilibrium_report_matches_oracle(sh):
    x, v = deformed_state(sh, seed=8)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        ctx.calc_deformation_forces()
        ctx.calc_non_deformation_forces()
        ctx.impose_constraints()
        Fg = ctx.get_forces(0)
        rep = ctx.is_at_equilibrium(0)
    fr, fi, vr, vi = orc.System.from_shell(sh).equilibrium(Fg, v)
    assert rep["max_force_node"] == fi and rep["max_speed_node"] == vi
    assert rep["max_force_ratio"] == fr and rep["max_speed_ratio"] == vr


def test_fused_step_equals_ops_sequence_bitwise(sh):
    x, v = deformed_state(sh, seed=9)
    with _ctx(sh) as a, _ctx(sh) as b:
        a.set_state(0, x, v)
        b.set_state(0, x, v)
        for _ in range(3):
            a.calc_deformation_forces()
            a.calc_non_deformation_forces()
            a.impose_constraints()
            a.advance_dynamics()
        b.step(3)
        xa, va = a.get_state(0)
        xb, vb = b.get_state(0)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_trajectory_matches_oracle_and_is_deterministic(sh):
    x, v = deformed_state(sh, seed=10)
    n = 200
    runs = []
    for _ in range(2):
        with _ctx(sh) as ctx:
            ctx.set_state(0, x, v)
            ctx.step(n)
            runs.append(ctx.get_state(0))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    sysm = orc.System.from_shell(sh)
    xo, vo = x.copy(), v.copy()
    assert sysm.run_steps(xo, vo, n) == -1
    dx = np.abs(runs[0][0] - xo).max()
    print(f"200-step position deviation {dx:.3e}")
    assert dx <= 1e-11


def test_graph_replay_equals_direct_launches(sh):
    x, v = deformed_state(sh, seed=11)
    with _ctx(sh) as a, _ctx(sh) as b:
        a.set_state(0, x, v)
        b.set_state(0, x, v)
        a.set_tuning(graph_steps=8)
        a.step(37)          # 4 graph replays + 5 direct steps
        b.enable_timing(True)
        b.step(37)          # all direct, with events
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])
        t = b.last_step_timing()
        assert t[0] > 0 and t[1] > 0


@pytest.mark.parametrize("tri_block,node_block,grid", [(64, 128, 0), (128, 64, 2), (256, 256, 1)])
def test_launch_shapes_do_not_change_results(sh, tri_block, node_block, grid):
    x, v = deformed_state(sh, seed=12)
    with _ctx(sh) as a, _ctx(sh) as b:
        a.set_state(0, x, v)
        b.set_state(0, x, v)
        b.set_tuning(tri_block=tri_block, node_block=node_block, tri_grid_per_cu=grid, node_grid_per_cu=grid)
        a.step(5)
        b.step(5)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])


def test_blowup_reports_first_offending_dof(sh):
    from morphoshell_amd.shell import HighForceError

    x, v = deformed_state(sh, seed=13)
    N = sh.N
    free = np.where(sh.mesh.constrained == 0)[0]
    k1, k2 = free[40], free[300]
    x[k1] += 0.5   # displaced far beyond an edge length
    x[2 * N + k2] += 0.5
    sysm = orc.System.from_shell(sh)
    fo = sysm.full_forces(x, v)
    d_o, bad_o = sysm.advance(x.copy(), v.copy(), fo)
    assert d_o >= 0
    for fused in (False, True):
        with _ctx(sh) as ctx:
            ctx.set_state(0, x, v)
            with pytest.raises(HighForceError) as ei:
                if fused:
                    ctx.step(4)
                else:
                    ctx.calc_deformation_forces()
                    ctx.calc_non_deformation_forces()
                    ctx.impose_constraints()
                    ctx.advance_dynamics()
            assert ei.value.dof == d_o
            assert str(ei.value) == "implausibly_high_force"
            assert abs(ei.value.force - bad_o) <= 1e-10 * abs(bad_o)
            # nothing advanced: the retained state is the failing step's input
            xs, _ = ctx.get_state(0)
            assert np.array_equal(xs, x)


def test_batch_of_shells_equals_single_shells():
    shells = [_shell(0.08, seed=1), _shell(0.05, seed=2, kappa=-1.0), _shell(0.11, seed=3, h=0.02)]
    states = [deformed_state(s, seed=20 + i) for i, s in enumerate(shells)]
    with _ctx(shells) as ctx:
        for i, (x, v) in enumerate(states):
            ctx.set_state(i, x, v)
        ctx.step(25)
        batch = [ctx.get_state(i) for i in range(3)]
        E = [ctx.calc_energies(i) for i in range(3)]
    for i, s in enumerate(shells):
        with _ctx(s) as c1:
            c1.set_state(0, *states[i])
            c1.step(25)
            xs, vs = c1.get_state(0)
            assert np.array_equal(xs, batch[i][0]) and np.array_equal(vs, batch[i][1])
            assert np.allclose(c1.calc_energies(0), E[i], rtol=1e-14, atol=0)


def test_device_dialling_matches_oracle():
    for protocol in (1, 3):
        s = _shell(0.06, seed=4, protocol=protocol)
        x, v = deformed_state(s, seed=30)
        f = 0.37
        with _ctx(s) as ctx:
            ctx.set_state(0, x, v)
            K = 1000
            ctx.set_dial_clock(0, 0.0, 1.0, K, int(f * K), True)
            assert ctx.dial_factor(0) == pytest.approx(f)
            ctx.calc_deformation_forces()
            Fg = ctx.get_forces(0)
        ab, bb = orc.dial(s.T, protocol, False, f, s.mesh.abar_info, s.mesh.bbar_info if protocol == 1 else None)
        Fo = orc.System.from_shell(s, abar=ab, bbar=bb).deformation_forces(x)
        assert np.abs(Fg - Fo).max() <= 1e-10 * np.abs(Fo).max(), protocol


def test_dial_clock_advances_per_step():
    s = _shell(0.1, seed=5)
    x, v = deformed_state(s, seed=31)
    with _ctx(s) as ctx:
        ctx.set_state(0, x, v)
        ctx.set_dial_clock(0, 0.2, 0.5, 100, 10, True)
        ctx.step(30)
        assert ctx.dial_factor(0) == pytest.approx(0.2 + 0.5 * 40 / 100)
        ctx.set_dial_clock(0, 0.2, 0.5, 100, 40, False)
        ctx.step(5)
        assert ctx.dial_factor(0) == pytest.approx(0.2 + 0.5 * 40 / 100)


def test_mixed_precision_force_error(sh):
    x, v = deformed_state(sh, seed=14)
    with _ctx(sh, "mixed") as ctx:
        ctx.set_state(0, x, v)
        ctx.calc_deformation_forces()
        Fm = ctx.get_forces(0)
        ctx.set_precision("fp64")
        ctx.calc_deformation_forces()
        Fd = ctx.get_forces(0)
    Fo = orc.System.from_shell(sh).deformation_forces(x)
    em = np.abs(Fm - Fo).max() / np.abs(Fo).max()
    ed = np.abs(Fd - Fo).max() / np.abs(Fo).max()
    print(f"mixed force rel err {em:.3e}, fp64 {ed:.3e}")
    assert ed <= 1e-10
    assert em <= 1e-6


def test_fused_equilibrium_check_and_halt():
    s = _shell(0.1, seed=6, kappa=0.0)
    # the flat reference state with abar = I, bbar = 0 is an equilibrium
    s.mesh.abar_info[:] = [1.0, 0.0, 1.0]
    with _ctx(s) as ctx, _ctx(s) as plain:
        ctx.set_state(0, s.ref_pos)
        plain.set_state(0, s.ref_pos)
        # checks run on clock steps 4, 9, ...: the shell halts at step 4 and
        # keeps the state that was checked (the input of step 4)
        reps = ctx.step(10, equil_every=5, force_tol=1e-6, speed_tol=1e-5, halt_on_equilibrium=True)
        r = reps[0]
        assert r["checked"] and r["at_equilibrium"] and r["halted"]
        assert r["max_force_ratio"] < 1e-10
        plain.step(4)
        x, v = ctx.get_state(0)
        xp, vp = plain.get_state(0)
        assert np.array_equal(x, xp) and np.array_equal(v, vp)
        # a halted shell is skipped by further steps
        ctx.step(7)
        assert np.array_equal(ctx.get_state(0)[0], xp)
        ctx.resume(0)
        ctx.step(1)
        plain.step(1)
        assert np.array_equal(ctx.get_state(0)[0], plain.get_state(0)[0])


# ------------------------------------------------- fused owner-computes step
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("node_block,grid,tile_chunks", [(64, 0, 1), (128, 2, 1), (256, 0, 1), (64, 0, 2),
                                                         (256, 3, 2), (512, 0, 1), (512, 2, 1)])
def test_fused_tile_step_is_bit_identical(sh, precision, node_block, grid, tile_chunks):
    """The fused tile launch sums the same contributions in the same
    patch-incidence order as the two-launch path: states agree bit for bit,
    including the equilibrium maxima of checked steps (per-chunk partials,
    also when a tile spans two chunks)."""
    x, v = deformed_state(sh, seed=40)
    with _ctx(sh, precision) as a, _ctx(sh, precision) as b:
        a.set_state(0, x, v)
        b.set_state(0, x, v)
        a.set_tuning(node_block=node_block, node_grid_per_cu=grid)
        b.set_tuning(node_block=node_block, node_grid_per_cu=grid, fused=1, tile_chunks=tile_chunks)
        assert b.get_tuning()["fused"] == 1 and b.get_tuning()["tile_chunks"] == tile_chunks
        ra = a.step(41, equil_every=10, force_tol=1e-30, speed_tol=1e-30)
        rb = b.step(41, equil_every=10, force_tol=1e-30, speed_tol=1e-30)
        xa, va = a.get_state(0)
        xb, vb = b.get_state(0)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        for k in ("max_force_ratio", "max_speed_ratio", "max_force_node", "max_speed_node", "checked"):
            assert ra[0][k] == rb[0][k], k
        b.enable_timing(True)
        b.step(3)
        a.step(3)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])


@pytest.mark.parametrize("scale", [1e-5, 1e8])
def test_forces_match_oracle_at_other_length_scales(scale):
    """The FP64 mechanics use refined hardware reciprocals and reciprocal
    square roots where the oracle divides: on a disc scaled by 1e-5 and 1e8
    (thickness and target curvature with it), the forces still match the
    oracle to 1e-10 of max |F|, and a fused step equals a two-launch step bit
    for bit.  (The patch fit's rank rule, relative to the largest pivot of a
    matrix whose rows scale as 1, L and L^2, as in the reference's
    FullPivLU, refuses this mesh at scales of 1e-6 and 1e9.)"""
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.04, seed=3, h=0.01 * scale, kappa=0.5 / scale)
    m.points = m.points * scale
    s = prepare_shell(m, Settings(bend_stiffening_scale_factor=40.0, dial_protocol=1))
    x, v = deformed_state(s, seed=21, amp=0.05 * scale)
    x[2 * s.N:] = 0.05 * scale * np.sin(2.5 * x[:s.N] / scale) * np.cos(1.7 * x[s.N:2 * s.N] / scale)
    v = v * scale
    Fo = orc.System.from_shell(s).deformation_forces(x)
    with _ctx(s) as a, _ctx(s) as b:
        a.set_state(0, x, v)
        a.calc_deformation_forces()
        Fg = a.get_forces(0)
        b.set_state(0, x, v)
        b.set_tuning(node_block=128, fused=1)
        a.step(1)
        b.step(1)
        xa, va = a.get_state(0)
        xb, vb = b.get_state(0)
    err = np.abs(Fg - Fo).max() / np.abs(Fo).max()
    print(f"scale {scale:g}: max |F| {np.abs(Fo).max():.3e}, rel err {err:.3e}")
    assert np.isfinite(Fg).all() and err <= 1e-10
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_fused_step_forms_areas_only_where_they_match(sh):
    """The FP64 fused step forms each reference area as 0.5 |x1 y2 - x2 y1|
    from the reference positions of patch nodes 0-2 when that equals every
    given area bit for bit (so the fused runs of the other tests, on planar
    reference states, take that path), and reads the area stream otherwise:
    with one area scaled by 1 + 1e-9 the fused step still equals the
    two-launch form, which reads the stream, and differs from the run on the
    unperturbed areas."""
    import copy

    p = np.asarray(sh.patch_ids)
    X, Y = sh.ref_pos[: sh.N], sh.ref_pos[sh.N: 2 * sh.N]
    x1, y1 = X[p[:, 1]] - X[p[:, 0]], Y[p[:, 1]] - Y[p[:, 0]]
    x2, y2 = X[p[:, 2]] - X[p[:, 0]], Y[p[:, 2]] - Y[p[:, 0]]
    assert np.array_equal(0.5 * np.abs(x1 * y2 - x2 * y1), sh.areas)
    s2 = copy.copy(sh)
    s2.areas = sh.areas.copy()
    s2.areas[sh.T // 3] *= 1.0 + 1e-9
    x, v = deformed_state(sh, seed=41)
    with _ctx(s2) as a, _ctx(s2) as b, _ctx(sh) as c:
        for ctx in (a, b, c):
            ctx.set_state(0, x, v)
        b.set_tuning(node_block=256, fused=1)
        c.set_tuning(node_block=256, fused=1)
        for ctx in (a, b, c):
            ctx.step(20)
        xa, xb, xc = (ctx.get_state(0)[0] for ctx in (a, b, c))
    assert np.array_equal(xa, xb)
    assert not np.array_equal(xb, xc)


@pytest.mark.parametrize("tile_chunks", [1, 2])
def test_fused_tile_step_batch_matches_oracle(tile_chunks):
    shells = [_shell(0.08, seed=1), _shell(0.05, seed=2, kappa=-1.0, clamp_every=7), _shell(0.11, seed=3, h=0.02)]
    states = [deformed_state(s, seed=50 + i) for i, s in enumerate(shells)]
    with _ctx(shells) as ctx, _ctx(shells) as ref:
        ctx.set_tuning(node_block=128, fused=1, tile_chunks=tile_chunks)
        for i, st in enumerate(states):
            ctx.set_state(i, *st)
            ref.set_state(i, *st)
        ctx.step(30)
        ref.step(30)
        for i, s in enumerate(shells):
            xg, vg = ctx.get_state(i)
            assert np.array_equal(xg, ref.get_state(i)[0]) and np.array_equal(vg, ref.get_state(i)[1])
            xo, vo = states[i][0].copy(), states[i][1].copy()
            assert orc.System.from_shell(s).run_steps(xo, vo, 30) == -1
            assert np.abs(xg - xo).max() <= 1e-11


def test_fused_tile_step_reports_blowup(sh):
    from morphoshell_amd.shell import HighForceError

    x, v = deformed_state(sh, seed=13)
    N = sh.N
    free = np.where(sh.mesh.constrained == 0)[0]
    x[free[40]] += 0.5
    x[2 * N + free[300]] += 0.5
    sysm = orc.System.from_shell(sh)
    d_o, bad_o = sysm.advance(x.copy(), v.copy(), sysm.full_forces(x, v))
    with _ctx(sh) as ctx:
        ctx.set_tuning(fused=1)
        ctx.set_state(0, x, v)
        with pytest.raises(HighForceError) as ei:
            ctx.step(4)
        assert ei.value.dof == d_o
        assert np.array_equal(ctx.get_state(0)[0], x)


def test_dial_from_ansatz_matches_oracle_in_both_step_forms():
    """Protocol-1 blend from the ansatz geometry (do_dialling.cpp:194-228):
    forces against the oracle, and the fused form (tiles built before the
    zero-dial geometry arrives, so their record copies are refreshed) against
    the two-launch form."""
    s = _shell(0.06, seed=7, dialling_from_ansatz_rather_than_ref_state=True)
    rng = np.random.default_rng(3)
    T = s.T
    a0 = np.concatenate([1 + 0.02 * rng.random(T), 0.01 * rng.standard_normal(T), 1 + 0.02 * rng.random(T)])
    b0 = 0.2 * rng.standard_normal(3 * T)
    x, v = deformed_state(s, seed=33)
    f, K = 0.4, 1000
    with _ctx(s) as a, _ctx(s) as b:
        b.set_tuning(node_block=128, fused=1)
        for c in (a, b):
            c.set_zero_dial_geometry(0, a0, b0)
            c.set_state(0, x, v)
            c.set_dial_clock(0, 0.0, 1.0, K, int(f * K), True)
        a.calc_deformation_forces()
        Fg = a.get_forces(0)
        a.set_state(0, x, v)
        a.step(12)
        b.step(12)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])
    ab, bb = orc.dial(T, 1, True, f, s.mesh.abar_info, s.mesh.bbar_info, a0, b0)
    Fo = orc.System.from_shell(s, abar=ab, bbar=bb).deformation_forces(x)
    err = np.abs(Fg - Fo).max() / np.abs(Fo).max()
    print(f"ansatz-dial force rel err {err:.3e}")
    assert err <= 1e-10


def _reference_form_diagnostics(F, b, a, abar):
    """calc_curvatures.cpp:52-66 and calc_strains.cpp:52-71 as written there:
    QR of the 3x2 deformation gradient for the symmetric shape operator, and
    the pseudo-inverse strain tensor (numpy, per triangle)."""
    T = F.shape[0]
    gauss, mean, strain = np.empty(T), np.empty(T), np.empty(T)
    for t in range(T):
        R = np.linalg.qr(F[t], mode="r")
        Si = np.linalg.inv(R[:2, :2])
        B = np.array([[b[0, t], b[1, t]], [b[1, t], b[2, t]]])
        S = Si.T @ B @ Si
        gauss[t] = np.linalg.det(S)
        mean[t] = 0.5 * np.trace(S)
        A = np.array([[a[0, t], a[1, t]], [a[1, t], a[2, t]]])
        Ab = np.array([[abar[0, t], abar[1, t]], [abar[1, t], abar[2, t]]])
        Fp = np.linalg.inv(F[t].T @ F[t]) @ F[t].T
        E = 0.5 * (Fp.T @ (A - Ab) @ Fp)
        strain[t] = np.sqrt(0.5 * np.trace(E.T @ E))
    return gauss, mean, strain


def test_output_diagnostics_match_reference_forms():
    """Device curvatures and strain against the reference's QR and
    pseudo-inverse forms, evaluated in numpy on the oracle's geometry."""
    s = _shell(0.08, seed=9, kappa=0.7)
    x, v = deformed_state(s, seed=60)
    sysm = orc.System.from_shell(s)
    with _ctx(s) as ctx:
        ctx.set_state(0, x, v)
        ctx.set_dial_clock(0, 1.0, 1.0, 1, 0, False)
        d = ctx.calc_diagnostics(0)
        a, b, _ = ctx.calc_a_comps_and_b_comps_and_normals(0)
    q = sysm.surface_derivs(x).reshape(-1, 3, 5)   # per triangle: X, Y, Z rows of (Rx, Ry, Rxx, Rxy, Ryy)
    F = q[:, :, 0:2]
    abar = s.mesh.abar_info.T
    g_ref, m_ref, s_ref = _reference_form_diagnostics(F, b, a, abar)
    for name, got, ref in (("gauss", d["gauss_curv"], g_ref), ("mean", d["mean_curv"], m_ref),
                           ("strain", d["strain"], s_ref)):
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
        print(f"{name}: rel err {err:.2e}")
        assert err <= 1e-10, name


def test_create_and_call_argument_errors(sh):
    """Bad inputs are refused with MS_ERR_ARG and a message, nothing runs."""
    import copy

    from morphoshell_amd.native import MS_ERR_ARG, NativeError

    bad = copy.copy(sh)
    bad.patch_ids = sh.patch_ids.copy()
    bad.patch_ids[5, 4] = sh.N + 3
    with pytest.raises(NativeError, match="patch id out of range") as ei:
        _ctx(bad)
    assert ei.value.code == MS_ERR_ARG
    bad = copy.copy(sh)
    bad.mesh = copy.copy(sh.mesh)
    bad.mesh.abar_info = sh.mesh.abar_info[:, :2].copy()
    with pytest.raises(NativeError, match="3-column"):
        _ctx(bad)
    with _ctx(sh) as ctx:
        # the library's own check, below the wrapper's (which raises IndexError)
        from morphoshell_amd import native as nat

        x0 = np.ascontiguousarray(sh.ref_pos, dtype=np.float64)
        assert ctx.L.ms_set_state(ctx.h, 1, nat.dptr(x0), None) == MS_ERR_ARG
        assert b"out of range" in ctx.L.ms_ctx_last_error(ctx.h)
        with pytest.raises(NativeError, match="fused must be"):
            ctx.set_tuning(fused=3)
        with pytest.raises(NativeError, match="node_block"):
            ctx.set_tuning(node_block=96)
        with pytest.raises(NativeError, match="tile_chunks"):
            ctx.set_tuning(tile_chunks=3)
        assert ctx.get_tuning()["tile_chunks"] == 1
        with pytest.raises(NativeError, match="nsteps"):
            ctx.step(-1)
        # the context still works after refused calls
        ctx.set_state(0, sh.ref_pos)
        ctx.step(2)


def test_autotuner_picks_a_shape_without_changing_results(tmp_path, monkeypatch):
    """autotune.tune times the candidates, applies the fastest, caches it by
    workload signature; results are the same bits under any choice."""
    from morphoshell_amd import autotune

    monkeypatch.setenv("MS_TUNE_CACHE", str(tmp_path / "tune.json"))
    shells = [_shell(0.08, seed=1), _shell(0.05, seed=2, kappa=-1.0)]
    states = [deformed_state(s, seed=70 + i) for i, s in enumerate(shells)]
    space = {"tri_block": (64, 256), "node_block": (128,), "grid_per_cu": (0, 4), "graph_steps": (8,),
             "fused_node_block": (64, 256)}
    with _ctx(shells) as a, _ctx(shells) as b:
        best = autotune.tune(a, steps=10, space=space)
        assert not best["cached"] and best["candidates"] == 8
        assert autotune._device() in best["signature"]  # cached per device
        t = a.get_tuning()
        assert t["fused"] == best["fused"] and t["node_block"] == best["node_block"]
        again = autotune.tune(b, steps=10, space=space)
        assert again["cached"] and again["fused"] == best["fused"]
        for c in (a, b):
            for i, st in enumerate(states):
                c.set_state(i, *st)
                c.resume(i)
        a.step(20)
        b.set_tuning(tri_block=256, node_block=256, tri_grid_per_cu=0, node_grid_per_cu=0, fused=0)
        b.step(20)
        for i in range(2):
            assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0])


def _annulus(spacing=0.06, r_in=0.35, seed=5):
    """A disc with its centre removed (inner boundary, one hole), nodes
    relabelled to the kept ones."""
    from morphoshell_amd.meshing import Mesh, grf, isotropic_abar_info, uniform_bbar_info
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(spacing, seed=seed, h=0.02, kappa=0.4)
    c = m.centroids()
    keep = np.hypot(c[:, 0], c[:, 1]) > r_in
    tris = m.tris[keep]
    used = np.unique(tris)
    remap = np.full(m.num_nodes, -1, np.int64)
    remap[used] = np.arange(len(used))
    a = Mesh(points=m.points[used], tris=remap[tris].astype(np.int32))
    ca = a.centroids()
    a.abar_info = isotropic_abar_info(1.0 + 0.08 * grf(ca[:, :2], 0.3, seed=seed + 1))
    a.bbar_info = uniform_bbar_info(a.num_tris, 0.4)
    a.fill_defaults(h=0.02)
    cl = np.zeros(a.num_nodes, np.uint8)
    cl[np.hypot(a.points[:, 0], a.points[:, 1]) > 0.97] = 1  # clamp the outer rim
    a.constrained = cl
    return prepare_shell(a, Settings())


def test_annulus_matches_oracle_in_both_step_forms():
    s = _annulus()
    assert s.info["num_holes"] == 1
    x, v = deformed_state(s, seed=80)
    with _ctx(s) as a, _ctx(s) as b:
        b.set_tuning(node_block=64, fused=1)
        for c in (a, b):
            c.set_state(0, x, v)
            c.step(25)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])
        xg = a.get_state(0)[0]
    xo, vo = x.copy(), v.copy()
    assert orc.System.from_shell(s).run_steps(xo, vo, 25) == -1
    assert np.abs(xg - xo).max() <= 1e-11


@pytest.mark.parametrize("precision,tile_chunks", [("fp64", 1), ("mixed", 1), ("fp64", 2), ("mixed", 2)])
def test_many_tiny_shells_fused_equals_two_launch(precision, tile_chunks):
    """A ragged batch of 40 shells of 30-400 triangles (most smaller than one
    tile): the fused step equals the two-launch step bit for bit."""
    shells = [_shell(0.16 + 0.25 * (i % 5) / 4, seed=i, h=0.02 + 0.005 * (i % 3), kappa=0.3 * (i % 4) - 0.4)
              for i in range(40)]
    states = [deformed_state(sh, seed=100 + i) for i, sh in enumerate(shells)]
    with _ctx(shells, precision) as a, _ctx(shells, precision) as b:
        b.set_tuning(node_block=64, fused=1, tile_chunks=tile_chunks)
        for i, st in enumerate(states):
            a.set_state(i, *st)
            b.set_state(i, *st)
        ra = a.step(30, equil_every=7, force_tol=1e-30, speed_tol=1e-30)
        rb = b.step(30, equil_every=7, force_tol=1e-30, speed_tol=1e-30)
        for i in range(len(shells)):
            assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0]), i
            assert ra[i]["max_force_ratio"] == rb[i]["max_force_ratio"], i


_GRF_CHECKS = {}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grf_disc_checked_at_steps_0_100_5000(precision):
    """The large smooth-random-field disc (isotropic abar = lambda^2 I,
    lambda = 1 + 0.1 g, 64-mode field of correlation length 0.2, bbar = 0,
    h = 0.005) on the same generator at spacing 0.02, run with the fused
    step from a flat start with z-noise and checked at steps 0, 100, 5000.
    At every check, against the oracle at the device's own state: forces
    within 1e-10 of max |F| (FP64) or 1e-6 (mixed), stretch and bend energy
    within 1e-11 relative (the energy sums are FP64 in both modes).
    Against the oracle's own trajectory: the FP64 positions agree to 1e-9
    and the energies to 1e-9 at step 100 (mixed: 1e-4).  By step 5000 the
    sheet is buckling and the trajectories have parted through the
    instability (max |dx| ~0.1), so only the total elastic energy is held to
    10%.  The errors are printed for the record."""
    from morphoshell_amd import workloads
    from morphoshell_amd.shell import Context

    sh = workloads.grf_disc(0.02, seed=3)
    x0, v0 = workloads.flat_noisy_state(sh, seed=5)
    sysm = orc.System.from_shell(sh)
    if "oracle" not in _GRF_CHECKS:
        xo, vo, done, states = x0.copy(), v0.copy(), 0, {}
        for k in (0, 100, 5000):
            assert sysm.run_steps(xo, vo, k - done) == -1
            done = k
            states[k] = (xo.copy(), sysm.energies(xo, vo))
        _GRF_CHECKS["oracle"] = states
    states = _GRF_CHECKS["oracle"]
    f_tol = 1e-10 if precision == "fp64" else 1e-6
    with Context(sh, precision=precision) as ctx:
        ctx.set_tuning(node_block=256, fused=1)
        ctx.set_state(0, x0, v0)
        done = 0
        for k in (0, 100, 5000):
            ctx.step(k - done)
            done = k
            x, v = ctx.get_state(0)
            ctx.calc_deformation_forces()
            F = ctx.get_forces(0)
            Fo = sysm.deformation_forces(x)
            ef = np.abs(F - Fo).max() / np.abs(Fo).max()
            E = ctx.calc_energies(0)
            Es = sysm.energies(x, v)
            es = np.abs(E[:2] - Es[:2]) / np.abs(Es[:2])
            xo, Eo = states[k]
            et = np.abs(E[:2] - Eo[:2]) / np.abs(Eo[:2])
            dx = np.abs(x - xo).max()
            print(f"{precision} step {k}: force rel err {ef:.2e}, energy rel err at this state {es.max():.2e}; "
                  f"against the oracle trajectory: stretch/bend energy {et[0]:.2e}/{et[1]:.2e}, max |dx| {dx:.2e}")
            assert ef <= f_tol
            assert (es <= 1e-11).all()
            if k <= 100:
                assert (et <= (1e-9 if precision == "fp64" else 1e-4)).all()
                if precision == "fp64":
                    assert dx <= 1e-9
            else:
                assert abs((E[0] + E[1]) - (Eo[0] + Eo[1])) <= 0.1 * (Eo[0] + Eo[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_per_triangle_thickness_modulus_and_bbar(precision):
    """A shell with per-triangle thickness, shear modulus and bbar next to one
    whose values are uniform: the kernels that stream these arrays run (a
    context takes them from the shell record only when every shell is
    uniform), both step forms agree bit for bit, and FP64 forces match the
    oracle to 1e-10 of max |F|; the uniform shell alone (record values)
    gives the same bits as in the mixed context."""
    from morphoshell_amd.meshing import uniform_bbar_info
    from morphoshell_amd.shell import Settings, prepare_shell

    rng = np.random.default_rng(21)
    m = grf_disc(0.06, seed=21, h=0.02, kappa=0.0)
    m.thickness = 0.02 * (1.0 + 0.3 * rng.uniform(size=m.num_tris))
    m.shear_modulus = 1.0 + 0.5 * rng.uniform(size=m.num_tris)
    m.bbar_info = uniform_bbar_info(m.num_tris, 0.4) + 0.2 * rng.standard_normal((m.num_tris, 3))
    varied = prepare_shell(m, Settings())
    uniform = _shell(0.07, seed=22, h=0.02, kappa=0.6)
    xs = [deformed_state(varied, seed=31), deformed_state(uniform, seed=32)]
    ends = {}
    for fused in (0, 1):
        with _ctx([varied, uniform], precision) as ctx:
            ctx.set_tuning(node_block=128, fused=fused)
            for i, st in enumerate(xs):
                ctx.set_state(i, *st)
            if fused == 0 and precision == "fp64":
                ctx.calc_deformation_forces()
                F = ctx.get_forces(0)
                Fo = orc.System.from_shell(varied).deformation_forces(xs[0][0])
                assert np.abs(F - Fo).max() <= 1e-10 * np.abs(Fo).max()
            ctx.step(20)
            ends[fused] = [ctx.get_state(i)[0] for i in range(2)]
    for i in range(2):
        assert np.array_equal(ends[0][i], ends[1][i]), i
    with _ctx(uniform, precision) as ctx:
        ctx.set_tuning(node_block=128, fused=1)
        ctx.set_state(0, *xs[1])
        ctx.step(20)
        assert np.array_equal(ctx.get_state(0)[0], ends[1][1])
    if precision == "fp64":
        xo, vo = xs[0][0].copy(), xs[0][1].copy()
        assert orc.System.from_shell(varied).run_steps(xo, vo, 20) == -1
        assert np.abs(ends[1][0] - xo).max() <= 1e-11


def test_fused_tiles_survive_a_precision_switch(sh):
    """Tiles are built with the pass length of the context's precision at
    the time (2B triangles in FP64, B in mixed mode); switching precision on
    a live fused context keeps them, and the states equal those of a context
    built in the new precision, bit for bit, both ways round."""
    x, v = deformed_state(sh, seed=91)
    for first, second in (("fp64", "mixed"), ("mixed", "fp64")):
        with _ctx(sh, first) as a, _ctx(sh, second) as b:
            a.set_tuning(node_block=64, fused=1)
            b.set_tuning(node_block=64, fused=1)
            a.set_precision(second)
            for c in (a, b):
                c.set_state(0, x, v)
            ra = a.step(23, equil_every=5, force_tol=1e-30, speed_tol=1e-30)
            rb = b.step(23, equil_every=5, force_tol=1e-30, speed_tol=1e-30)
            assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0]), (first, second)
            assert ra[0]["max_force_ratio"] == rb[0]["max_force_ratio"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lce_shell_fused_equals_two_launch_while_dialling(precision):
    """An LCE-protocol shell (protocol 3 targets, the kernel variant with the
    pow() of the opto-thermal law) stepped while its dial factor advances:
    the fused step equals the two-launch step bit for bit in both
    precisions."""
    s = _shell(0.06, seed=8, protocol=3)
    x, v = deformed_state(s, seed=81)
    K = 400
    with _ctx(s, precision) as a, _ctx(s, precision) as b:
        b.set_tuning(node_block=128, fused=1)
        for c in (a, b):
            c.set_state(0, x, v)
            c.set_dial_clock(0, 0.0, 1.0, K, 100, True)
        ra = a.step(37, equil_every=9, force_tol=1e-30, speed_tol=1e-30)
        rb = b.step(37, equil_every=9, force_tol=1e-30, speed_tol=1e-30)
        assert a.dial_factor(0) == b.dial_factor(0) == pytest.approx((100 + 37) / K)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])
        assert ra[0]["max_force_ratio"] == rb[0]["max_force_ratio"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lce_metric_cache_equals_recomputing(precision):
    """The fused step keeps an LCE shell's dialled metric in its tile records
    and reuses it while the dial factor stays put.  Against the two-launch
    form, which evaluates the opto-thermal law at every step, the states
    stay bit-identical: while dialling (the cache is rewritten every step),
    at the dial's end (f = 1 for many steps: reused), after the host moves
    the dial clock, after the tiles are rebuilt (a new record generation),
    across a precision switch and under the device-run protocol.  A blend
    shell in the same context keeps its bbar-info record slots."""
    shells = [_shell(0.06, seed=8, protocol=3), _shell(0.08, seed=9, kappa=0.6)]
    states = [deformed_state(s, seed=91 + i) for i, s in enumerate(shells)]
    other = "mixed" if precision == "fp64" else "fp64"
    with _ctx(shells, precision) as a, _ctx(shells, precision) as b:
        a.set_tuning(node_block=128, fused=1)
        b.set_tuning(node_block=128, fused=0)

        def same():
            for i in range(len(shells)):
                assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0]), i
                assert np.array_equal(a.get_state(i)[1], b.get_state(i)[1]), i

        for c in (a, b):
            for i, st in enumerate(states):
                c.set_state(i, *st)
                c.set_dial_clock(i, 0.0, 1.0, 40, 0, True)
        for c in (a, b):
            c.step(90, equil_every=25)  # 40 dialled steps, then 50 at f = 1
        same()
        for c in (a, b):
            c.set_dial_clock(0, 0.5, 0.0, 40, 0, False)  # held at f = 0.5
            c.step(30)
        same()
        a.set_tuning(node_block=64, fused=1)  # new tiles, new record generation
        for c in (a, b):
            c.step(30)
        same()
        for c in (a, b):
            c.set_precision(other)
            c.step(20)
        same()
        for c in (a, b):
            c.set_precision(precision)
            for i in range(len(shells)):
                c.relax_setup(i, steps_per_dial_phase=30, dial_spacing=0.5, dial_start=0.0, equil_interval=10,
                              tol_f=1e-30, tol_v=1e-30)
            c.relax_run(150, poll_every=50)
        same()
        assert a.dial_factor(0) == b.dial_factor(0)


@pytest.mark.parametrize("precision,fused", [("fp64", 0), ("mixed", 0), ("fp64", 1), ("mixed", 1)])
def test_halted_shells_drop_out_of_the_step_launches(precision, fused):
    """Once shells halt, the step launches cover the running shells' chunks
    and tiles only.  The running shells' states and checked residuals equal
    those of a context where nothing halts, bit for bit, the halted shells'
    states stay as they halted, and a re-armed shell runs again."""
    shells = [_shell(0.12 + 0.04 * (i % 3), seed=i, h=0.02, kappa=0.4 - 0.2 * i) for i in range(6)]
    states = [deformed_state(sh, seed=200 + i) for i, sh in enumerate(shells)]
    halting = (1, 4)
    with _ctx(shells, precision) as a, _ctx(shells, precision) as b:
        for c in (a, b):
            c.set_tuning(node_block=64, tri_block=64, fused=fused)
            for i, st in enumerate(states):
                c.set_state(i, *st)
        for i in halting:  # fully dialled, any residual counts as equilibrium: halts at its first check
            a.relax_setup(i, steps_per_dial_phase=1, dial_spacing=1.0, dial_start=1.0, equil_interval=2,
                          tol_f=1e30, tol_v=1e30)
        ra = a.step(5)
        assert all(ra[i]["halted"] for i in halting)
        held = {i: a.get_state(i) for i in halting}
        b.step(5)
        ra = a.step(40, equil_every=6, force_tol=1e-30, speed_tol=1e-30)
        rb = b.step(40, equil_every=6, force_tol=1e-30, speed_tol=1e-30)
        for i in range(len(shells)):
            xa, va = a.get_state(i)
            if i in halting:
                assert np.array_equal(xa, held[i][0]) and np.array_equal(va, held[i][1]), i
                continue
            xb, vb = b.get_state(i)
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), i
            assert ra[i]["max_force_ratio"] == rb[i]["max_force_ratio"], i
            assert ra[i]["max_force_node"] == rb[i]["max_force_node"], i
        # re-armed: the shell is in the launches again
        a.relax_setup(1, steps_per_dial_phase=1, dial_spacing=1.0, dial_start=1.0, equil_interval=2,
                      tol_f=1e-30, tol_v=1e-30)
        a.step(10)
        b.step(10)
        assert not np.array_equal(a.get_state(1)[0], held[1][0])
        for i in (0, 2, 3, 5):
            assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0]), i


def test_device_info_reports_the_device():
    import ctypes as C

    from morphoshell_amd import native

    name = C.create_string_buffer(64)
    cus, hbm, clk, bus = C.c_int32(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
    native.check(native.lib().ms_device_info(name, 64, C.byref(cus), C.byref(hbm), C.byref(clk), C.byref(bus)))
    assert name.value and len(name.value) < 64
    assert cus.value > 0 and hbm.value > 0 and clk.value > 0 and bus.value > 0
    # a short buffer is truncated and terminated; null outputs are allowed
    short = C.create_string_buffer(b"x" * 4, 4)
    native.check(native.lib().ms_device_info(short, 4, None, None, None, None))
    assert short.raw[3:4] == b"\0" and short.value == name.value[:3]


def _relabelled_shell(seed=5):
    """A 9.2k-node disc whose node labels are randomly permuted: a tile (a
    run of consecutive labels) is scattered over the whole disc, so its
    triangles reach most of the nodes (1024-node tiles: 8289 of 9214, whose
    positions alone need 199 KB of LDS)."""
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.02, seed=seed, h=0.02, kappa=0.2)
    perm = np.random.default_rng(seed).permutation(m.num_nodes)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    m.points = m.points[perm]
    m.tris = inv[m.tris]
    m.constrained = m.constrained[perm]
    return prepare_shell(m, Settings())


def test_refused_tile_shape_leaves_a_working_two_launch_context():
    """A fused shape whose tiles do not fit the LDS (1024-node tiles of a
    mesh with scattered node labels) is refused with an error; the context
    falls back to the two-launch form at the new block size and steps exactly
    as a context set to it."""
    from morphoshell_amd.native import NativeError

    sh = _relabelled_shell()
    x, v = deformed_state(sh, seed=41)
    with _ctx(sh) as a, _ctx(sh) as b:
        a.set_tuning(node_block=128, fused=1)
        assert a.get_tuning()["fused"] == 1
        with pytest.raises(NativeError, match="LDS"):
            a.set_tuning(node_block=512, fused=1, tile_chunks=2)
        t = a.get_tuning()
        assert t["fused"] == 0 and t["node_block"] == 512
        b.set_tuning(node_block=512, fused=0)
        for c in (a, b):
            c.set_state(0, x, v)
            c.step(20)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])


@pytest.mark.parametrize("node_block,tile_chunks", [(512, 2), (256, 2)])
def test_two_chunk_fp64_tiles_take_short_passes_and_match_two_launch(sh, node_block, tile_chunks):
    """FP64 tiles of two chunks run passes of B triangles (2B passes would
    not leave the LDS for the workgroups per CU the registers allow; at 512
    threads they would not fit at all) and step bit-identically to the
    two-launch form, the equilibrium check included."""
    x, v = deformed_state(sh, seed=43)
    with _ctx(sh) as a, _ctx(sh) as b:
        a.set_tuning(node_block=node_block, tile_chunks=tile_chunks, fused=1)
        assert a.get_tuning()["fused"] == 1
        b.set_tuning(node_block=node_block, fused=0)
        for c in (a, b):
            c.set_state(0, x, v)
        ra = a.step(24, equil_every=8)
        rb = b.step(24, equil_every=8)
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])
        assert np.array_equal(a.get_state(0)[1], b.get_state(0)[1])
        assert ra[0]["max_force_ratio"] == rb[0]["max_force_ratio"]
        assert ra[0]["max_speed_ratio"] == rb[0]["max_speed_ratio"]


@pytest.mark.parametrize("tile_nodes", [-2, 333, 100])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_tiles_of_any_size_step_like_two_launch(precision, tile_nodes):
    """Tiles smaller than their chunks: the size balanced over a persistent
    grid (here the 64-node floor: two small shells give fewer full tiles than
    the grid has workgroups) and explicit sizes that cut a partial second
    chunk (333 = 256 + 77) or one partial chunk (100).  Node groups restart
    at each shell.  The fused step stays bit-identical to the two-launch
    form, the equilibrium check included, and get_tuning reports the size in
    force."""
    shells = [_shell(0.02, seed=5), _shell(0.03, seed=6, kappa=-0.4)]
    states = [deformed_state(s, seed=60 + i) for i, s in enumerate(shells)]
    with _ctx(shells, precision) as a, _ctx(shells, precision) as b:
        a.set_tuning(node_block=256, tile_chunks=2, fused=1, node_grid_per_cu=2, tile_nodes=tile_nodes)
        t = a.get_tuning()
        assert t["fused"] == 1
        assert t["tile_nodes"] == (64 if tile_nodes == -2 else tile_nodes)
        b.set_tuning(node_block=256, fused=0)
        for c in (a, b):
            for i, st in enumerate(states):
                c.set_state(i, *st)
        ra = a.step(24, equil_every=8)
        rb = b.step(24, equil_every=8)
        for i in range(len(shells)):
            assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0])
            assert np.array_equal(a.get_state(i)[1], b.get_state(i)[1])
            assert ra[i]["max_force_ratio"] == rb[i]["max_force_ratio"]
            assert ra[i]["max_speed_ratio"] == rb[i]["max_speed_ratio"]


def test_wrapper_checks_array_sizes_and_shell_indices(sh):
    """The C-ABI reads as many values as the shell has; the Python wrapper
    rejects arrays of another size and shell indices out of range before
    any call."""
    x, v = deformed_state(sh, seed=42)
    with _ctx(sh) as ctx:
        with pytest.raises(ValueError, match="expected"):
            ctx.set_state(0, x[:-1], v)
        with pytest.raises(ValueError, match="expected"):
            ctx.set_state(0, x, v[:-3])
        with pytest.raises(ValueError, match="expected"):
            ctx.set_forces(0, np.zeros(3 * sh.N + 3))
        with pytest.raises(ValueError, match="expected"):
            ctx.set_zero_dial_geometry(0, np.ones(3 * sh.T - 1), None)
        with pytest.raises(IndexError):
            ctx.get_state(1)
        with pytest.raises(IndexError):
            ctx.set_state(-1, x, v)
        with pytest.raises(ValueError, match="precision"):
            ctx.set_precision("fp32")
        ctx.set_state(0, x, v)  # the right sizes still work
        assert np.array_equal(ctx.get_state(0)[0], x)


@pytest.mark.parametrize("tile_nodes", [-1, -2])
@pytest.mark.parametrize("tile_chunks", [1, 2])
@pytest.mark.parametrize("grid_per_cu", [0, 2, 8])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_halted_shells_fused_active_tiles(precision, grid_per_cu, tile_chunks, tile_nodes):
    """With the fused form, after some shells halt the running shells' tiles
    run from the active-tile list, with a grid of one workgroup per tile
    (grid_per_cu 0) or a persistent grid of 2 or 8 workgroups per CU that
    strides over the list when it holds more tiles than the grid has
    workgroups, and with tiles of one or two chunks.  With balanced tiles
    (tile_nodes -2) the tiles are re-cut as shells halt (two-chunk tiles on
    the grid of 8 per CU: 109 nodes while all 24 shells run, 64 once 12
    do).  The running shells' states equal, bit for bit, those of a context
    where nothing halts and that runs one workgroup per tile of one chunk."""
    # 24 shells of 9.2k nodes: 144 tiles each at 64-node tiles of one chunk
    # (72 at two chunks), so 3456 / 1728 tiles, 2880 / 1440 running after the
    # first halts and 1728 / 864 after the second; a grid of 2 per CU (512
    # workgroups on 256 CUs) strides over the list throughout, one of 8 per
    # CU (2048) strides before the second halts at one chunk
    shells = [_shell(0.02, seed=i, h=0.02, kappa=0.3 - 0.02 * i) for i in range(24)]
    states = [deformed_state(s, seed=300 + i) for i, s in enumerate(shells)]
    with _ctx(shells, precision) as a, _ctx(shells, precision) as b:
        a.set_tuning(node_block=64, tri_block=64, fused=1, node_grid_per_cu=grid_per_cu, tile_chunks=tile_chunks,
                     tile_nodes=tile_nodes)
        b.set_tuning(node_block=64, tri_block=64, fused=1, node_grid_per_cu=0, tile_chunks=1)
        L0 = a.get_tuning()["tile_nodes"]
        assert a.get_tuning()["node_grid_per_cu"] == grid_per_cu
        assert a.get_tuning()["tile_chunks"] == tile_chunks
        for c in (a, b):
            for i, st in enumerate(states):
                c.set_state(i, *st)
        running = set(range(len(shells)))
        for halting in (range(0, 4), range(4, 12)):
            for i in halting:
                a.relax_setup(i, steps_per_dial_phase=1, dial_spacing=1.0, dial_start=1.0, equil_interval=2,
                              tol_f=1e30, tol_v=1e30)
            ra = a.step(6)
            b.step(6)
            assert all(ra[i]["halted"] for i in halting)
            running -= set(halting)
            a.step(20)
            b.step(20)
            for i in running:
                assert np.array_equal(a.get_state(i)[0], b.get_state(i)[0]), i
                assert np.array_equal(a.get_state(i)[1], b.get_state(i)[1]), i
        if tile_nodes == -2 and tile_chunks == 2 and grid_per_cu == 8:
            assert (L0, a.get_tuning()["tile_nodes"]) == (109, 64)
            assert a.get_tuning()["fused"] == 1


def test_cauchy_stress_matches_oracle(sh):
    """calc_stresses (calc_stresses.cpp:45-179) at a bent and stretched state
    with bend stiffening active: the device's kept eigenvalues match the
    oracle's (its own gradient, LAPACK eigensolver) to 1e-10 of the largest,
    and the eigenvectors agree to 1e-9 (and in sign) wherever the pair is
    well separated from each other and from the dropped eigenvalue."""
    x, v = deformed_state(sh, seed=31)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        Sd = ctx.calc_stresses(0)
    So = orc.System.from_shell(sh).stresses(x)
    lam_o = So[[0, 4]]
    scale = np.abs(lam_o).max()
    err = np.abs(Sd[[0, 4]] - lam_o).max() / scale
    print(f"stress eigenvalue err {err:.2e} of the largest ({scale:.3e})")
    assert err <= 1e-10
    gap = np.minimum(np.abs(lam_o[0] - lam_o[1]), np.minimum(np.abs(lam_o[0]), np.abs(lam_o[1])))
    ok = gap > 1e-6 * scale
    assert ok.mean() > 0.9
    worst = 0.0
    for r in range(2):
        vd, vo = Sd[4 * r + 1:4 * r + 4, ok], So[4 * r + 1:4 * r + 4, ok]
        worst = max(worst, float(np.abs(vd - vo).max()))
    print(f"eigenvector err {worst:.2e} over {ok.sum()} of {sh.T} triangles")
    assert worst <= 1e-9


def test_cauchy_stress_closed_form_on_the_device():
    """A flat sheet with abar = I, bbar = 0, stretched by (1.02, 0.97) and
    rotated: eigenvalues 2 A c_s g0 sx/sy and 2 A c_s g2 sy/sx along the
    rotated stretch axes (c_s = h mu/(1 - nu), g0 = 2 nu (ex + ey) +
    2 (1 - nu) ex, ex = sx^2 - 1), the normal's zero dropped."""
    import math

    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.08, seed=32, h=0.02, kappa=0.0)
    m.abar_info[:] = [1.0, 0.0, 1.0]
    s = prepare_shell(m, Settings())
    N = s.N
    sx, sy = 1.02, 0.97
    x = s.ref_pos.copy()
    x[:N] *= sx
    x[N:2 * N] *= sy
    x[2 * N:] = 0.0
    c, si = math.cos(0.5), math.sin(0.5)
    Q = np.array([[c, 0.0, si], [0.0, 1.0, 0.0], [-si, 0.0, c]])
    x = (Q @ x.reshape(3, N)).reshape(-1)
    with _ctx(s) as ctx:
        ctx.set_state(0, x)
        S = ctx.calc_stresses(0)
    nu = s.settings.def_poisson_ratio
    cs = s.thickness * s.shear_modulus / (1.0 - nu)
    ex, ey = sx * sx - 1.0, sy * sy - 1.0
    l_x = 2.0 * s.areas * cs * (2.0 * nu * (ex + ey) + 2.0 * (1.0 - nu) * ex) * sx / sy
    l_y = 2.0 * s.areas * cs * (2.0 * nu * (ex + ey) + 2.0 * (1.0 - nu) * ey) * sy / sx
    # ascending eigenvalues are (l_y, 0, l_x) here (l_y < 0 < l_x): the zero is
    # dropped and the pair kept in cyclic order is (l_x, l_y)
    assert (l_y < 0).all() and (l_x > 0).all()
    assert np.abs(S[0] - l_x).max() <= 1e-12 * np.abs(l_x).max()
    assert np.abs(S[4] - l_y).max() <= 1e-12 * np.abs(l_x).max()
    assert np.abs(np.abs(Q[:, 0] @ S[1:4]) - 1.0).max() <= 1e-12
    assert np.abs(np.abs(Q[:, 1] @ S[5:8]) - 1.0).max() <= 1e-12


def test_angle_deficits_match_host_and_gauss_bonnet(sh):
    """calc_angle_deficits (calc_angle_deficits.cpp:40-78): against a numpy
    restatement summing the corner angles in triangle order (to 1e-13,
    acos may differ by an ulp between the device and the host libraries),
    and discrete Gauss-Bonnet: over a disc the deficits sum to 2 pi."""
    x, v = deformed_state(sh, seed=33)
    with _ctx(sh) as ctx:
        ctx.set_state(0, x, v)
        d = ctx.calc_angle_deficits(0)
    N = sh.N
    P = x.reshape(3, N).T
    tri = sh.mesh.tris
    ref = np.where(sh.info["boundary_nodes"] != 0, np.pi, 2.0 * np.pi)
    s0 = P[tri[:, 1]] - P[tri[:, 0]]
    s1 = P[tri[:, 2]] - P[tri[:, 1]]
    s2 = -s1 - s0
    sides = [s0, s1, s2]
    ang = []
    for k in range(3):
        a, b = sides[(k + 2) % 3], sides[k]
        ang.append(np.arccos(-(a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))))
    corners = np.stack(ang, 1)  # T x 3, in the order of the triangle loop
    np.subtract.at(ref, tri.reshape(-1), corners.reshape(-1))
    err = np.abs(d - ref).max()
    print(f"angle deficit err {err:.2e}; sum {d.sum():.15f} vs 2 pi {2 * np.pi:.15f}")
    assert err <= 1e-13
    assert d.sum() == pytest.approx(2.0 * np.pi, abs=1e-9)
