"""The device-run relaxation protocol (main.cpp:269-507 state machine) against
the oracle's restatement of the same loop, through the C-ABI.

Shell: 711-triangle disc, GRF metric (amplitude 0.05), bbar = 0.3 I, h = 0.05,
damping prefactors 20, two dialling phases of 300 steps (spacing 0.5),
equilibrium checks every 25 steps with both tolerances 1e-2.  The oracle
reaches equilibria at steps 3316 and 5125; the device must reach them at the
same steps and end in the same state (positions within 1e-9 after ~5000
FP64 steps).
"""
import numpy as np
import pytest

from conftest import grf_disc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

K, SPACING, INTERVAL, TOL = 300, 0.5, 25, 1e-2


def _setup(seed=2, h=0.05, kappa=0.3):
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.1, seed=seed, h=h, kappa=kappa, amp=0.05)
    sh = prepare_shell(m, Settings(damping_prefactor_1=20.0, damping_prefactor_2=20.0))
    rng = np.random.default_rng(seed)
    x = sh.ref_pos.copy()
    x[2 * sh.N:] += 1e-4 * rng.standard_normal(sh.N)
    return sh, x


def _oracle(sh, x):
    from morphoshell_amd.shell import scales_for

    sysm = orc.System.from_shell(sh)
    eta = scales_for(sh, 20.0)["damping_factor"]
    xx, vv = x.copy(), np.zeros_like(x)
    r = orc.run_protocol(sysm, lambda f: orc.dial(sh.T, 1, False, f, sh.mesh.abar_info, sh.mesh.bbar_info),
                         xx, vv, K, SPACING, 0.0, INTERVAL, eta, eta, TOL, TOL, 30000)
    return r, xx, vv


def _arm(ctx, i, pause=False):
    ctx.relax_setup(i, steps_per_dial_phase=K, dial_spacing=SPACING, dial_start=0.0, equil_interval=INTERVAL,
                    tol_f=TOL, tol_v=TOL, pause_on_eq=pause)


def test_device_protocol_matches_oracle():
    from morphoshell_amd.shell import Context

    sh, x = _setup()
    r, xo, vo = _oracle(sh, x)
    assert r["status"] == 2 and len(r["eq_steps"]) == 2
    with Context(sh) as ctx:
        ctx.set_state(0, x)
        _arm(ctx, 0)
        rep = ctx.relax_run(30000, poll_every=1000)[0]
        xg, vg = ctx.get_state(0)
    print("oracle", r, "device", rep)
    assert rep["status"] == 2 and rep["halted"] and not rep["failed"]
    assert rep["num_eq"] == 2 and rep["eq_step"] == r["eq_steps"][-1]
    assert rep["dial_factor"] == 1.0
    assert np.abs(xg - xo).max() <= 1e-9
    assert np.abs(vg - vo).max() <= 1e-9


def test_check_log_matches_the_oracle_loop():
    """The device's protocol log (ms_set_check_log / ms_read_check_log)
    against the events of the oracle's restatement of the loop: every
    dialling-phase end and every equilibrium check, in order, at the same
    steps with the same dial factors, first-argmax nodes and verdicts, and
    the force and speed ratios to 1e-9 relative (the FP64 states agree to
    ~1e-11).  Read in pieces over the run, with a ring too small to hold it
    all at once."""
    from morphoshell_amd.shell import Context

    sh, x = _setup()
    r, _, _ = _oracle(sh, x)
    ev_o = r["events"]
    checks = sum(e[0] == "check" for e in ev_o)
    assert checks > 50 and sum(e[0] == "phase_end" for e in ev_o) == 2
    got = []
    with Context(sh) as ctx:
        ctx.set_tuning(node_block=128, fused=1)
        ctx.set_check_log(64)
        ctx.set_state(0, x)
        _arm(ctx, 0)
        while True:
            rep = ctx.relax_run(1000, poll_every=500)[0]  # at most 1000 / (INTERVAL + 1) < 64 events
            evs, dropped = ctx.read_check_log(0)
            assert dropped == 0
            got += evs
            if rep["halted"]:
                break
    assert len(got) == len(ev_o)
    for g, o in zip(got, ev_o):
        if o[0] == "phase_end":
            assert g["kind"] == 1 and g["step"] == o[1] and g["dial_factor"] == o[2], (g, o)
        else:
            _, step, f, fr, fi, vr, vi, ok = o
            assert g["kind"] == 0 and g["step"] == step and g["dial_factor"] == f, (g, o)
            assert (g["max_force_node"], g["max_speed_node"], bool(g["at_equilibrium"])) == (fi, vi, ok), (g, o)
            assert g["max_force_ratio"] == pytest.approx(fr, rel=1e-9)
            assert g["max_speed_ratio"] == pytest.approx(vr, rel=1e-9)


def test_pause_and_resume_equals_uninterrupted():
    from morphoshell_amd.shell import Context

    sh, x = _setup()
    with Context(sh) as a, Context(sh) as b:
        a.set_state(0, x)
        b.set_state(0, x)
        _arm(a, 0)
        _arm(b, 0, pause=True)
        ra = a.relax_run(30000, poll_every=1000)[0]
        rb = b.relax_run(30000, poll_every=1000)[0]
        assert rb["halted"] and rb["status"] == 1 and rb["num_eq"] == 1  # paused at the first equilibrium
        b.resume(0)
        rb = b.relax_run(30000, poll_every=1000)[0]
        assert rb["halted"] and rb["status"] == 2 and rb["num_eq"] == 2
        # the second halt is the final equilibrium's check: same step, same state
        assert rb["eq_step"] == ra["eq_step"]
        assert np.array_equal(a.get_state(0)[0], b.get_state(0)[0])


def test_ragged_batch_protocol_equals_single_shells():
    from morphoshell_amd.shell import Context

    setups = [_setup(seed=2), _setup(seed=5, h=0.04, kappa=-0.5), _setup(seed=7, h=0.06, kappa=0.0)]
    shells = [s for s, _ in setups]
    with Context(shells) as ctx:
        for i, (_, x) in enumerate(setups):
            ctx.set_state(i, x)
            _arm(ctx, i)
        reps = ctx.relax_run(40000, poll_every=997)
        batch = [ctx.get_state(i)[0] for i in range(3)]
    for i, (sh, x) in enumerate(setups):
        with Context(sh) as c1:
            c1.set_state(0, x)
            _arm(c1, 0)
            r1 = c1.relax_run(40000, poll_every=1000)[0]
            assert reps[i]["status"] == r1["status"] == 2
            assert reps[i]["eq_step"] == r1["eq_step"]
            assert np.array_equal(c1.get_state(0)[0], batch[i])
