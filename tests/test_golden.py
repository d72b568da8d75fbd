"""Golden fixtures of the per-step path (tests/golden/*.npz, written by
tests/golden/make_golden.py with the FP64 oracle).

The reference ships no test vectors for this path, so these fixtures are the
oracle's own, kept as regression anchors:
  * CPU: the manifest checksums, the product's mesh generator and host set-up
    (patch ids exact, G and H to 1e-12, areas/masses/scales to 1e-14) and the
    oracle itself (forces, energies and the 100-step trajectory, exact);
  * GPU: the device path in both step forms against the fixture (forces
    1e-10 of max |F|, energies 1e-11, 100-step positions 1e-11 absolute).
"""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

CASES = ["disc_c1", "hand_200"]


def _load(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))


def _shell(name):
    """The shell of a fixture, rebuilt by the product's own generators."""
    from morphoshell_amd import workloads
    from morphoshell_amd.shell import Settings, prepare_shell

    if name == "disc_c1":
        return workloads.small_defect_disc()
    from make_golden import hand_mesh

    return prepare_shell(hand_mesh(), Settings(def_poisson_ratio=0.4))


def _oracle_system(d):
    from oracle import oracle as orc

    T = len(d["tris"])
    protocol = int(d["protocol"])
    ab, bb = orc.dial(T, protocol, False, float(d["dial_factor"]), d["abar_info"],
                      None if protocol == 3 else d["bbar_info"])
    return orc.System(len(d["points"]), T, d["patch_ids"], d["G"], d["H"], d["areas"], d["thickness"],
                      d["shear_modulus"], ab, bb, d["masses"], d["constrained"], float(d["poisson_ratio"]),
                      float(d["gamma"]), float(d["ref_density"]), float(d["min_ref_thickness"]), float(d["timestep"]),
                      float(d["damping_factor"]), float(d["char_force_scale"]), float(d["char_speed_scale"]), False)


def _state(d, k):
    return (d["x_start"], d["v_start"]) if k == 0 else (d[f"x_{k}"], d[f"v_{k}"])


def test_manifest_checksums():
    with open(os.path.join(GOLDEN, "MANIFEST.sha256")) as fh:
        for line in fh:
            digest, name = line.split()
            with open(os.path.join(GOLDEN, name), "rb") as f2:
                assert hashlib.sha256(f2.read()).hexdigest() == digest, name


@pytest.mark.parametrize("name", CASES)
def test_host_setup_reproduces_fixture(name):
    d = _load(name)
    sh = _shell(name)
    assert np.array_equal(sh.mesh.points, d["points"]) and np.array_equal(sh.mesh.tris, d["tris"])
    assert np.array_equal(sh.patch_ids, d["patch_ids"])
    for k in ("G", "H"):
        ref = d[k].reshape(len(d["tris"]), -1)
        got = np.asarray(getattr(sh, k)).reshape(ref.shape)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), k
    assert np.allclose(sh.areas, d["areas"], rtol=1e-14, atol=0)
    assert np.allclose(sh.masses, d["masses"], rtol=1e-14, atol=0)
    for k in ("timestep", "damping_factor", "char_force_scale", "char_speed_scale", "min_ref_thickness"):
        assert sh.scales[k] == pytest.approx(float(d[k]), rel=1e-14), k


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    from oracle import oracle as orc

    from morphoshell_amd.shell import Settings

    d = _load(name)
    thresh = Settings().patch_mat_dimless_conditioning_thresh
    pid, G, H, failed = orc.patch_fit(np.ascontiguousarray(d["points"].T.reshape(-1)), d["tris"], thresh)
    assert failed < 0
    assert np.array_equal(pid, d["patch_ids"]) and np.array_equal(G, d["G"]) and np.array_equal(H, d["H"])
    sysm = _oracle_system(d)
    for k in (0, 1, 100):
        x, v = _state(d, k)
        assert np.array_equal(sysm.full_forces(x, v), d[f"force_{k}"]), k
        assert np.array_equal(sysm.energies(x, v), d[f"energies_{k}"]), k
    x, v = d["x_start"].copy(), d["v_start"].copy()
    assert sysm.run_steps(x, v, 100) == -1
    assert np.array_equal(x, d["x_100"]) and np.array_equal(v, d["v_100"])


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("name", CASES)
def test_device_matches_fixture(name, fused):
    from morphoshell_amd.shell import Context

    d = _load(name)
    sh = _shell(name)
    f = float(d["dial_factor"])
    with Context(sh, precision="fp64") as ctx:
        ctx.set_tuning(node_block=64, fused=fused)
        ctx.set_dial_clock(0, f, 1.0, 1, 0, False)
        for k in (0, 1, 100):
            x, v = _state(d, k)
            ctx.set_state(0, x, v)
            ctx.calc_deformation_forces()
            ctx.calc_non_deformation_forces()
            ctx.impose_constraints()
            F = ctx.get_forces(0)
            Fo = d[f"force_{k}"]
            err = np.abs(F - Fo).max() / np.abs(Fo).max()
            assert err <= 1e-10, (k, err)
            E, Eo = ctx.calc_energies(0), d[f"energies_{k}"]
            assert np.all(np.abs(E - Eo) <= 1e-11 * np.abs(Eo).max()), (k, E, Eo)
        ctx.set_state(0, d["x_start"], d["v_start"])
        ctx.step(100)
        x, _ = ctx.get_state(0)
        dx = np.abs(x - d["x_100"]).max()
        print(f"{name} fused={fused}: 100-step position deviation {dx:.3e}")
        assert dx <= 1e-11
