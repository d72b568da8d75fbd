"""Pins of the FP64 oracle (and of the product's host set-up) on the CPU.

The reference ships no tests, fixtures or golden vectors for this path, so
every pin here comes from the mathematics: exactness of the patch operators
on affine/quadratic maps, closed-form energy densities, central finite
differences of the total energy, symmetry invariants, textbook compensated
summation, the discrete invariant of symplectic Euler, and scalar formulas.
Each is chosen so that a dropped term, a wrong sign or a swapped index fails.
"""
import math

import numpy as np
import pytest

from conftest import deformed_state, grf_disc
from oracle import oracle as orc


@pytest.fixture(scope="module")
def shell_c1():
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.04, seed=0, h=0.01, kappa=0.5)
    return prepare_shell(m, Settings(bend_stiffening_scale_factor=40.0))


@pytest.fixture(scope="module")
def shell_small():
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.12, seed=3, h=0.05, kappa=0.8)
    return prepare_shell(m, Settings(bend_stiffening_scale_factor=30.0))


# ------------------------------------------------------------ patch operators

def test_patch_choice_product_equals_oracle(shell_c1):
    sh = shell_c1
    pid, G, H, failed = orc.patch_fit(sh.ref_pos, sh.mesh.tris, sh.settings.patch_mat_dimless_conditioning_thresh)
    assert failed == -1
    assert np.array_equal(pid, sh.patch_ids)
    assert np.abs(G - sh.G).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(H - sh.H).max() <= 1e-12 * np.abs(H).max()
    # slots 0-2 are the triangle's own vertices in file order; 3-5 are distinct others
    assert np.array_equal(pid[:, :3], sh.mesh.tris)
    for row in pid[:50]:
        assert len(set(row.tolist())) == 6


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_G_exact_on_affine_and_H_exact_on_quadratic(shell_c1, which):
    sh = shell_c1
    if which == "product":
        pid, G, H = sh.patch_ids, sh.G, sh.H
    else:
        pid, G, H, _ = orc.patch_fit(sh.ref_pos, sh.mesh.tris, 10.0)
    N = sh.N
    x, y = sh.ref_pos[:N], sh.ref_pos[N:2 * N]
    # affine map f = 3 + 2x - 5y: R_x = 2, R_y = -5 exactly; H kills it
    f = 3.0 + 2.0 * x - 5.0 * y
    fv = f[pid]  # (T, 6)
    rx = (G[:, 0, :] * fv[:, :3]).sum(1)
    ry = (G[:, 1, :] * fv[:, :3]).sum(1)
    assert np.allclose(rx, 2.0, rtol=0, atol=1e-11)
    assert np.allclose(ry, -5.0, rtol=0, atol=1e-11)
    r2 = np.einsum("tnk,tn->tk", H, fv)
    assert np.abs(r2).max() < 1e-8
    # quadratic f = 0.7 x^2 - 1.3 x y + 0.4 y^2 + x: second derivatives exact
    q = 0.7 * x * x - 1.3 * x * y + 0.4 * y * y + x
    r2 = np.einsum("tnk,tn->tk", H, q[pid])
    assert np.allclose(r2[:, 0], 1.4, atol=1e-7)
    assert np.allclose(r2[:, 1], -1.3, atol=1e-7)
    assert np.allclose(r2[:, 2], 0.8, atol=1e-7)
    # moment identities of G: sum_v G = 0, sum_v G x_v = e_k
    assert np.abs(G.sum(2)).max() < 1e-9 * np.abs(G).max()


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_patch_fit_follows_from_reference_positions_and_non_vertex_rows(shell_c1, which):
    """The FP64 step streams only H's rows of the three non-vertex patch
    nodes.  In edge-vector form G is the inverse of the 2x2 matrix E of
    reference edge vectors from patch node 0, and H's vertex rows follow from
    its columns annihilating 1, X and Y: (H_1, H_2) = -E^-T sum_{n>=3} H_n
    (xi_n, eta_n), H_0 = -(H_1 + H_2 + sum_{n>=3} H_n).  Both patch fits
    (the product's LU, the oracle's Gauss-Jordan) agree with that to
    rounding."""
    sh = shell_c1
    if which == "product":
        pid, G, H = sh.patch_ids, sh.G, sh.H
    else:
        pid, G, H, _ = orc.patch_fit(sh.ref_pos, sh.mesh.tris, sh.settings.patch_mat_dimless_conditioning_thresh)
    N = sh.N
    X, Y = sh.ref_pos[:N], sh.ref_pos[N:2 * N]
    xi = X[pid[:, 1:]] - X[pid[:, [0]]]
    eta = Y[pid[:, 1:]] - Y[pid[:, [0]]]
    idet = 1.0 / (xi[:, 0] * eta[:, 1] - xi[:, 1] * eta[:, 0])
    g = np.stack([eta[:, 1] * idet, -eta[:, 0] * idet, -xi[:, 1] * idet, xi[:, 0] * idet], 1)
    Gs = np.stack([G[:, 0, 1], G[:, 0, 2], G[:, 1, 1], G[:, 1, 2]], 1)
    assert np.abs(g - Gs).max() <= 1e-14 * np.abs(Gs).max()
    assert np.abs(-(g[:, 0] + g[:, 1]) - G[:, 0, 0]).max() <= 1e-14 * np.abs(Gs).max()
    assert np.abs(-(g[:, 2] + g[:, 3]) - G[:, 1, 0]).max() <= 1e-14 * np.abs(Gs).max()
    Hn = H[:, 3:, :]
    t1 = (Hn * xi[:, 2:, None]).sum(1)
    t2 = (Hn * eta[:, 2:, None]).sum(1)
    H1 = -(g[:, [0]] * t1 + g[:, [2]] * t2)
    H2 = -(g[:, [1]] * t1 + g[:, [3]] * t2)
    H0 = -(Hn.sum(1) + H1 + H2)
    err = np.abs(np.stack([H0, H1, H2], 1) - H[:, :3, :]).max() / np.abs(H).max()
    assert err <= 1e-14, err
    # a wrong sign or a swapped edge in the rebuild is far off
    bad = -(g[:, [2]] * t1 + g[:, [0]] * t2)
    assert np.abs(bad - H[:, 1, :]).max() > 1e-3 * np.abs(H).max()


def test_patch_search_exhaustion_raises():
    from morphoshell_amd.native import NativeError
    from morphoshell_amd.shell import Settings, prepare_shell

    m = grf_disc(0.2, seed=0)
    with pytest.raises(NativeError, match="exhausted"):
        prepare_shell(m, Settings(patch_mat_dimless_conditioning_thresh=1e-6))
    pid, G, H, failed = orc.patch_fit(np.ascontiguousarray(m.points.T.reshape(-1)), m.tris, 1e-6)
    assert failed >= 0


# ------------------------------------------------------------------ geometry

def _q_from(Rx, Ry, Rxx, Rxy, Ryy):
    q = np.zeros(15)
    for c in range(3):
        q[5 * c:5 * c + 5] = [Rx[c], Ry[c], Rxx[c], Rxy[c], Ryy[c]]
    return q


def test_energy_closed_forms():
    h, mu, nu = 0.01, 1.3, 0.35
    # isotropic abar = lam^2 I, a = s^2 I: W_s = 2 h mu (1+nu)/(1-nu) (s^2-lam^2)^2 / lam^2
    lam, s = 1.1, 0.97
    q = _q_from([s, 0, 0], [0, s, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0])
    W = orc.tri_energy_q(q, [lam ** 2, 0, lam ** 2], [0, 0, 0], h, mu, nu, 1e300)
    ws = 2 * h * mu * (1 + nu) / (1 - nu) * (s * s - lam * lam) ** 2 / lam ** 2
    assert W[0] == pytest.approx(ws, rel=1e-13)
    assert W[1] == pytest.approx(0.0, abs=1e-300)
    # b = kappa I (R_xx = R_yy = -kappa n with n = +z), bbar = kbar I:
    # W_b0 = mu h^3 (1+nu)(kappa-kbar)^2 / (6 (1-nu) lam^2)
    kap, kb = 2.0, 0.5
    q = _q_from([lam, 0, 0], [0, lam, 0], [0, 0, -kap], [0, 0, 0], [0, 0, -kap])
    gam = 7.0
    W = orc.tri_energy_q(q, [lam ** 2, 0, lam ** 2], [kb, 0, kb], h, mu, nu, gam)
    wb0 = mu * h ** 3 * (1 + nu) * (kap - kb) ** 2 / (6 * (1 - nu) * lam ** 2)
    assert W[0] == pytest.approx(0.0, abs=1e-300)
    assert W[1] == pytest.approx(wb0 * (1 + wb0 / (gam * mu * h)), rel=1e-13)
    # anisotropic: pure shear a = diag(s1^2, s2^2), abar = I:
    # Phi = nu (d1+d2)^2 + (1-nu)(d1^2+d2^2)
    s1, s2 = 1.05, 0.9
    q = _q_from([s1, 0, 0], [0, s2, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0])
    W = orc.tri_energy_q(q, [1, 0, 1], [0, 0, 0], h, mu, nu, 1e300)
    d1, d2 = s1 * s1 - 1, s2 * s2 - 1
    assert W[0] == pytest.approx(h * mu / (1 - nu) * (nu * (d1 + d2) ** 2 + (1 - nu) * (d1 * d1 + d2 * d2)),
                                 rel=1e-13)
    # off-diagonal strain enters twice: a = [[1, e],[e, 1]], abar = I
    e = 0.03
    Rx = np.array([1.0, 0, 0])
    Ry = np.array([e, math.sqrt(1 - e * e), 0])
    W = orc.tri_energy_q(_q_from(Rx, Ry, [0] * 3, [0] * 3, [0] * 3), [1, 0, 1], [0, 0, 0], h, mu, nu, 1e300)
    assert W[0] == pytest.approx(h * mu / (1 - nu) * (1 - nu) * 2 * e * e, rel=1e-12)


def test_dE_dq_matches_finite_differences_per_triangle():
    rng = np.random.default_rng(7)
    for trial in range(5):
        q = rng.normal(size=15) * 0.3
        q[0], q[6] = 1.0 + 0.1 * rng.normal(), 1.0 + 0.1 * rng.normal()  # well-conditioned R_x, R_y
        ab = np.array([1.2, 0.1, 0.9])
        bb = rng.normal(size=3)
        h, mu, nu, gam = 0.05, 1.1, 0.4, 3.0
        _, g = orc.tri_energy_q(q, ab, bb, h, mu, nu, gam, want_grad=True)
        for j in range(15):
            eps = 1e-6
            qp, qm = q.copy(), q.copy()
            qp[j] += eps
            qm[j] -= eps
            Wp = orc.tri_energy_q(qp, ab, bb, h, mu, nu, gam).sum()
            Wm = orc.tri_energy_q(qm, ab, bb, h, mu, nu, gam).sum()
            fd = (Wp - Wm) / (2 * eps)
            assert abs(fd - g[j]) <= 1e-7 * max(1.0, np.abs(g).max()), (trial, j, fd, g[j])


def test_geometry_flat_stretch_and_quadratic_graph(shell_c1):
    sh = shell_c1
    sysm = orc.System.from_shell(sh)
    N = sh.N
    a, b, n = sysm.geometry(sh.ref_pos)
    assert np.allclose(a[0], 1) and np.allclose(a[1], 0, atol=1e-12) and np.allclose(a[2], 1)
    assert np.abs(b).max() < 1e-9
    assert np.allclose(n[2], 1.0)
    x = sh.ref_pos.copy()
    x[:N] *= 1.3
    x[N:2 * N] *= 0.8
    a, b, n = sysm.geometry(x)
    assert np.allclose(a[0], 1.69) and np.allclose(a[1], 0, atol=1e-12) and np.allclose(a[2], 0.64)
    # z = 1/2 (k1 x^2 + k2 y^2): R_xx = k1 e_z exactly, so b = -n.(k1, 0, k2) e_z
    k1, k2 = 0.6, -0.4
    X, Y = sh.ref_pos[:N], sh.ref_pos[N:2 * N]
    x = sh.ref_pos.copy()
    x[2 * N:] = 0.5 * (k1 * X * X + k2 * Y * Y)
    a, b, n = sysm.geometry(x)
    assert np.allclose(b[0], -n[2] * k1, atol=1e-8)
    assert np.allclose(b[1], 0.0, atol=1e-8)
    assert np.allclose(b[2], -n[2] * k2, atol=1e-8)


# -------------------------------------------------------------------- forces

def test_forces_match_finite_differences(shell_small):
    sh = shell_small
    sysm = orc.System.from_shell(sh)
    x, _ = deformed_state(sh, seed=1)
    F = sysm.deformation_forces(x)
    rng = np.random.default_rng(0)
    scale = np.abs(F).max()
    for d in rng.choice(3 * sh.N, 30, replace=False):
        eps = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[d] += eps
        xm[d] -= eps
        fd = -(sysm.total_energy(xp) - sysm.total_energy(xm)) / (2 * eps)
        assert abs(fd - F[d]) <= 2e-8 * scale, (d, fd, F[d])


def test_force_invariants(shell_c1):
    sh = shell_c1
    sysm = orc.System.from_shell(sh)
    x, _ = deformed_state(sh, seed=2)
    F = sysm.deformation_forces(x).reshape(3, -1)
    scale = np.abs(F).max()
    assert np.abs(F.sum(1)).max() < 1e-11 * scale * np.sqrt(sh.N)
    P = x.reshape(3, -1)
    torque = np.cross(P.T, F.T).sum(0)
    assert np.abs(torque).max() < 1e-11 * scale * np.sqrt(sh.N)


def test_zero_force_at_programmed_state(shell_c1):
    sh = shell_c1
    # flat mesh with abar = I, bbar = 0: every force vanishes
    sysm = orc.System.from_shell(sh, abar=np.tile([1.0, 0, 1.0], (sh.T, 1)).T.reshape(-1).copy(),
                                 bbar=np.zeros(3 * sh.T))
    F = sysm.deformation_forces(sh.ref_pos)
    assert np.abs(F).max() < 1e-14
    # homogeneous stretch with abar = s^2 I: interior forces vanish
    s = 1.07
    sysm = orc.System.from_shell(sh, abar=np.tile([s * s, 0, s * s], (sh.T, 1)).T.reshape(-1).copy(),
                                 bbar=np.zeros(3 * sh.T))
    F = sysm.deformation_forces(sh.ref_pos * s)
    assert np.abs(F).max() < 1e-13


# --------------------------------------------------------- sums and dynamics

def _reference_pick(lam3):
    """calc_stresses.cpp:165-176 on ascending eigenvalues: drop the first of
    least magnitude, keep the next two cyclically (indices)."""
    m = int(np.argmin(np.abs(lam3)))
    return (m + 1) % 3, (m + 2) % 3


@pytest.mark.parametrize("sx,sy", [(1.03, 1.0), (0.97, 1.0), (1.02, 0.99), (0.98, 1.04)])
def test_stress_closed_forms_stretched_flat_sheet(shell_c1, sx, sy):
    """A flat sheet with abar = I, bbar = 0 stretched by (sx, sy), then
    rotated rigidly by Q: the tensor (A/J) sum_alpha dW/dR_alpha (x) R_alpha
    is A c_s (2 g0 sx/sy Qe1 Qe1^T + 2 g2 sy/sx Qe2 Qe2^T) with c_s =
    h mu/(1 - nu), g0 = 2 nu (ex + ey) + 2 (1 - nu) ex, g2 the same with ex,
    ey swapped (ex = sx^2 - 1), and the normal's eigenvalue 0, so the two
    kept eigenpairs are these (in the reference's order), eigenvectors Qe1
    and Qe2 with their largest component positive."""
    sh = shell_c1
    T, N = sh.T, sh.N
    abar = np.concatenate([np.ones(T), np.zeros(T), np.ones(T)])
    sysm = orc.System.from_shell(sh, abar=abar, bbar=np.zeros(3 * T))
    x = sh.ref_pos.copy()
    x[:N] *= sx
    x[N:2 * N] *= sy
    x[2 * N:] = 0.0
    ang = 0.7
    Q = np.array([[math.cos(ang), 0.0, math.sin(ang)], [0.0, 1.0, 0.0], [-math.sin(ang), 0.0, math.cos(ang)]])
    Q = Q @ np.array([[math.cos(0.4), -math.sin(0.4), 0.0], [math.sin(0.4), math.cos(0.4), 0.0], [0.0, 0.0, 1.0]])
    x = (Q @ x.reshape(3, N)).reshape(-1)
    S = sysm.stresses(x)
    nu = sh.settings.def_poisson_ratio
    cs = sh.thickness * sh.shear_modulus / (1.0 - nu)
    ex, ey = sx * sx - 1.0, sy * sy - 1.0
    g0 = 2.0 * nu * (ex + ey) + 2.0 * (1.0 - nu) * ex
    g2 = 2.0 * nu * (ex + ey) + 2.0 * (1.0 - nu) * ey
    l1 = 2.0 * sh.areas * cs * g0 * sx / sy
    l2 = 2.0 * sh.areas * cs * g2 * sy / sx
    e1, e2 = Q[:, 0], Q[:, 1]
    scale = np.maximum(np.abs(l1), np.abs(l2))
    for t in range(0, T, 37):
        vals = [(0.0, None), (l1[t], e1), (l2[t], e2)]
        vals.sort(key=lambda p: p[0])
        i1, i2 = _reference_pick(np.array([v[0] for v in vals]))
        for r, k in enumerate((i1, i2)):
            lam, vec = vals[k]
            assert S[4 * r, t] == pytest.approx(lam, abs=1e-12 * scale[t])
            v = S[4 * r + 1:4 * r + 4, t]
            assert abs(abs(v @ vec) - 1.0) <= 1e-12
            assert v[np.argmax(np.abs(v))] > 0.0


def test_kahan_against_fsum():
    rng = np.random.default_rng(0)
    v = np.concatenate([[1.0], np.full(10000, 1e-16), rng.normal(size=1000) * 1e-3])
    assert orc.kahan_sum(v) == pytest.approx(math.fsum(v.tolist()), rel=0, abs=4e-16)
    naive = 0.0
    for t in v:
        naive += t
    assert abs(naive - math.fsum(v.tolist())) > 1e-13  # the naive sum is visibly wrong


def test_symplectic_euler_invariant(shell_small):
    """A one-dof harmonic oscillator driven through orc_advance conserves the
    discrete quadratic invariant of symplectic Euler exactly (to round-off)."""
    sh = shell_small
    N = sh.N
    k, m = 3.0, sh.masses[0]
    sysm = orc.System.from_shell(sh, eta=0.0)
    dt = sh.scales["timestep"]
    k = 0.25 * m / dt ** 2  # omega dt = 0.5
    x, v = np.zeros(3 * N), np.zeros(3 * N)
    x[0], v[0] = 0.01, 0.0
    inv0 = None
    for _ in range(2000):
        f = np.zeros(3 * N)
        f[0] = -k * x[0]
        d, _ = sysm.advance(x, v, f)
        assert d == -1
        # v' = v - dt k/m x ; x' = x + dt v'  conserves  k x^2 + m v^2 - dt k x v
        # evaluated as (x_n, v_n) with v_n the post-update velocity of step n
        inv = k * x[0] ** 2 + m * v[0] ** 2 - dt * k * x[0] * v[0]
        inv0 = inv if inv0 is None else inv0
        assert inv == pytest.approx(inv0, rel=1e-11)


def test_damping_clamp_advance_elementwise(shell_small):
    sh = shell_small
    N = sh.N
    sysm = orc.System.from_shell(sh)
    rng = np.random.default_rng(1)
    f = rng.normal(size=3 * N) * 1e-3
    v = rng.normal(size=3 * N) * 1e-2
    f0 = f.copy()
    sysm.damping(v, f)
    eta, rho, hmin = sh.scales["damping_factor"], sh.settings.ref_density, sh.scales["min_ref_thickness"]
    m3 = np.tile(sh.masses, 3)
    assert np.allclose(f, f0 - eta * m3 * v / (rho * hmin * hmin), rtol=1e-15, atol=0)
    # clamp
    cl = np.zeros(N, np.uint8)
    cl[::7] = 1
    sh2 = sh
    sysc = orc.System(N, sh.T, sh.patch_ids, sh.G, sh.H, sh.areas, sh.thickness, sh.shear_modulus,
                      np.ascontiguousarray(sh.mesh.abar_info.T.reshape(-1)),
                      np.ascontiguousarray(sh.mesh.bbar_info.T.reshape(-1)), sh.masses, cl, 0.5, 1e17, 1.0, hmin,
                      sh2.scales["timestep"], eta, sh.scales["char_force_scale"], sh.scales["char_speed_scale"])
    g = f.copy()
    sysc.constraints(g)
    for c in range(3):
        assert np.all(g[c * N:(c + 1) * N][cl == 1] == 0)
        assert np.array_equal(g[c * N:(c + 1) * N][cl == 0], f[c * N:(c + 1) * N][cl == 0])
    # blow-up: first offending dof, nothing updated
    x = np.zeros(3 * N)
    vv = np.zeros(3 * N)
    g = np.zeros(3 * N)
    lim = 5e4 * sh.scales["char_force_scale"]
    g[N + 5] = lim * 2
    g[2 * N + 1] = np.nan
    d, bad = sysm.advance(x, vv, g)
    assert d == N + 5 and bad == lim * 2
    assert not x.any() and not vv.any()


def test_equilibrium_first_argmax(shell_small):
    sh = shell_small
    N = sh.N
    sysm = orc.System.from_shell(sh)
    f = np.zeros(3 * N)
    v = np.zeros(3 * N)
    f[3] = 2.0
    f[2 * N + 9] = -2.0  # same norm, later node
    v[N + 4] = 1.0
    fr, fi, vr, vi = sysm.equilibrium(f, v)
    assert fi == 3 and vi == 4
    assert fr == pytest.approx(2.0 / sh.scales["char_force_scale"], rel=1e-15)
    assert vr == pytest.approx(1.0 / sh.scales["char_speed_scale"], rel=1e-15)


# ---------------------------------------------------------------- dialling

def test_dial_protocols():
    T = 5
    rng = np.random.default_rng(0)
    info = rng.uniform(0.8, 1.2, size=(T, 3))
    binfo = rng.normal(size=(T, 3))
    ab, bb = orc.dial(T, 1, False, 0.0, info, binfo)
    assert np.allclose(ab.reshape(3, T).T, [1, 0, 1]) and not bb.any()
    ab, bb = orc.dial(T, 1, False, 1.0, info, binfo)
    assert np.allclose(ab.reshape(3, T).T, info) and np.allclose(bb.reshape(3, T).T, binfo)
    ab, _ = orc.dial(T, 1, False, 0.25, info, binfo)
    assert np.allclose(ab.reshape(3, T).T, 0.75 * np.array([1, 0, 1]) + 0.25 * info)
    # LCE: director along x, lambda 0.9, nu_ot 0.5: abar = diag(lam^2, lam^-1)
    lce = np.tile([0.0, 0.9, 0.5], (T, 1))
    ab, bb = orc.dial(T, 3, False, 1.0, lce, None)
    assert np.allclose(ab.reshape(3, T).T, [0.81, 0, 1 / 0.9]) and not bb.any()
    # director at 90 degrees swaps the axes
    lce[:, 0] = np.pi / 2
    ab, _ = orc.dial(T, 3, False, 1.0, lce, None)
    assert np.allclose(ab.reshape(3, T).T, [1 / 0.9, 0, 0.81], atol=1e-12)
    # from ansatz: (1-f) a0 + f * final
    a0 = rng.uniform(0.9, 1.1, size=3 * T)
    ab, _ = orc.dial(T, 1, True, 0.5, info, binfo, a0=a0)
    assert np.allclose(ab, 0.5 * a0 + 0.5 * info.T.reshape(-1))


# --------------------------------------------------- host set-up (product)

def test_mesh_info_and_masses_equilateral():
    from morphoshell_amd.meshing import Mesh
    from morphoshell_amd.shell import Settings, prepare_shell

    # a perfect hexagon of 6 equilateral triangles, side 1
    ang = np.arange(6) * np.pi / 3
    pts = np.zeros((7, 3))
    pts[1:, 0], pts[1:, 1] = np.cos(ang), np.sin(ang)
    tris = np.array([[0, 1 + i, 1 + (i + 1) % 6] for i in range(6)], np.int32)
    m = Mesh(points=pts, tris=tris).fill_defaults(h=0.1, mu=2.0)
    st = Settings(patch_mat_dimless_conditioning_thresh=1e9, def_poisson_ratio=0.3, ref_density=1.5)
    sh = prepare_shell(m, st)
    A = math.sqrt(3) / 4
    assert np.allclose(sh.areas, A)
    assert sh.info["perimeter"] == pytest.approx(6.0)
    assert sh.info["approx_min_tri_size"] == pytest.approx(math.sqrt(3) / 2)
    assert sh.info["num_edges"] == 12 and sh.info["num_boundary_edges"] == 6 and sh.info["num_holes"] == 0
    assert sh.masses.sum() == pytest.approx(1.5 * 6 * A * 0.1)
    assert sh.masses[0] == pytest.approx(6 * 1.5 * A * 0.1 / 3)
    sc = sh.scales
    D = 2.0 * 0.1 ** 3 / (6 * 0.7)
    assert sc["char_flexural_rigidity"] == pytest.approx(D)
    assert sc["char_force_scale"] == pytest.approx(2.0 * math.sqrt(3) / 2 * 0.1)
    k = 2 * math.pi / 6.0
    assert sc["damping_factor"] == pytest.approx(1.0 * 2 * math.sqrt(D * 1.5 * 0.1 ** 3) * k * k)
    st_ = 0.185 * 0.6 * math.sqrt(1.5 / 2.0) * math.sqrt(3) / 2
    bt_ = 0.6 * math.sqrt(1.5 * 0.1 / D) * 0.75 / (2 * math.pi)
    assert sc["timestep"] == pytest.approx(min(st_, bt_))
    assert sc["char_long_time"] == pytest.approx(max(2 * math.pi * math.sqrt(1.5 / 2.0) / k,
                                                     2 * math.pi * math.sqrt(1.5 * 0.1 / D) / k ** 2))


def test_mesh_errors():
    from morphoshell_amd.meshing import Mesh
    from morphoshell_amd.native import NativeError
    from morphoshell_amd.shell import prepare_shell

    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    with pytest.raises(NativeError, match="start from 0"):
        prepare_shell(Mesh(points=pts, tris=np.array([[1, 2, 3]], np.int32)))
    with pytest.raises(NativeError, match="outside"):
        prepare_shell(Mesh(points=pts, tris=np.array([[0, 1, 7]], np.int32)))
    with pytest.raises(NativeError, match="touches no triangle"):
        prepare_shell(Mesh(points=pts, tris=np.array([[0, 1, 2]], np.int32)))


def test_settings_parser_strict_order():
    from morphoshell_amd.shell import REFERENCE_KEYS, read_settings

    vals = {
        "outputs_per_char_long_time": "400.0", "ref_thickness_if_uniform": "0.01", "timestep_prefactor": "0.6",
        "dialling_from_ansatz_rather_than_ref_state": "false", "dial_factors_to_hold_at_spacing": "1.0",
        "dial_phase_time_prefactor": "0.25", "time_between_equil_checks_prefactor": "0.01",
        "damping_prefactor_1": "35.0", "damping_prefactor_2": "35.0",
        "num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics": "0",
        "force_to_char_force_ratio_equil_threshold": "0.000001",
        "speed_to_char_speed_ratio_equil_threshold": "0.00001", "def_poisson_ratio": "0.5",
        "ref_shear_modulus_if_uniform": "1.0", "ref_density": "1.0",
        "bend_stiffening_scale_factor": "100000000000000000.0", "patch_mat_dimless_conditioning_thresh": "10.0",
        "slide_stiffness_prefactor": "0.0", "slide_friction_coefficent": "0.0", "slide_speed_prefactor": "0.0",
        "using_initial_glass_slide_z_coords_from_settings_file": "false", "initial_upper_slide_z_coord": "12345.0",
        "initial_lower_slide_z_coord": "12345.0",
    }
    text = "// comment line\n" + "\n".join(f"{k} = {vals[k]}; // trailing" for k in REFERENCE_KEYS)
    st = read_settings(text + "\nprecision = mixed;\ndial_protocol = 3;\n")
    assert st.damping_prefactor_1 == 35.0 and st.bend_stiffening_scale_factor == 1e17
    assert st.precision == "mixed" and st.dial_protocol == 3 and not st.dialling_from_ansatz_rather_than_ref_state
    swapped = list(REFERENCE_KEYS)
    swapped[1], swapped[2] = swapped[2], swapped[1]
    with pytest.raises(ValueError, match="ref_thickness_if_uniform"):
        read_settings("\n".join(f"{k} = {vals[k]};" for k in swapped))
    with pytest.raises(ValueError, match="semicolon"):
        read_settings(text.replace("0.6;", "0.6"))


def test_c_abi_library_exports_header_symbols():
    import ctypes
    import os
    import re

    from morphoshell_amd import native

    hdr = open(os.path.join(os.path.dirname(native.__file__), "..", "include", "morphoshell_amd.h")).read()
    names = sorted(set(re.findall(r"\b(ms_[a-z0-9_]+)\s*\(", hdr)))
    assert set(names) == set(native.EXPORTED)
    L = native.lib()
    for n in names:
        assert hasattr(L, n), n
    assert native.lib().ms_version().decode().count(".") == 2
    # argument errors are reported before any device is touched
    h = ctypes.c_void_p()
    assert L.ms_create(None, 0, 0, ctypes.byref(h)) == native.MS_ERR_ARG
    assert b"bad arguments" in L.ms_last_error()


def test_context_rejects_an_unknown_precision_before_any_device_call():
    import pytest

    from morphoshell_amd.meshing import disc_mesh
    from morphoshell_amd.shell import Context, Settings, prepare_shell

    sh = prepare_shell(disc_mesh(0.2), Settings())
    with pytest.raises(ValueError, match="precision"):
        Context(sh, precision="fp32")
