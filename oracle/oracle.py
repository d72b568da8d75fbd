"""ctypes front-end of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product path never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of oracle.c (the CPU baseline's -march=native copy)
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)
_lib = None


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    L = C.CDLL(LIB)
    P = C.c_void_p
    L.orc_kahan_sum.restype = C.c_double
    L.orc_kahan_sum.argtypes = [_dp, C.c_int64]
    L.orc_patch_fit.restype = C.c_int64
    L.orc_patch_fit.argtypes = [C.c_int32, C.c_int32, _dp, _ip, C.c_double, _ip, _dp, _dp]
    L.orc_create.restype = P
    L.orc_create.argtypes = [C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _u8,
                             C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_int]
    L.orc_destroy.argtypes = [P]
    L.orc_set_targets.argtypes = [P, _dp, _dp, _dp, _dp]
    L.orc_surface_derivs.argtypes = [P, _dp, _dp]
    L.orc_dE_dq.argtypes = [P, _dp, _dp]
    L.orc_deformation_forces.argtypes = [P, _dp, _dp]
    L.orc_geometry.argtypes = [P, _dp, _dp, _dp, _dp]
    L.orc_energy_densities.argtypes = [P, _dp, _dp]
    L.orc_energies.argtypes = [P, _dp, _dp, _dp]
    L.orc_damping.argtypes = [P, _dp, _dp]
    L.orc_constraints.argtypes = [P, _dp]
    L.orc_advance.restype = C.c_int64
    L.orc_advance.argtypes = [P, _dp, _dp, _dp, _dp]
    L.orc_equilibrium.argtypes = [P, _dp, _dp, _dp, _ip, _dp, _ip]
    L.orc_run_steps.restype = C.c_int64
    L.orc_run_steps.argtypes = [P, _dp, _dp, _dp, C.c_int64]
    L.orc_dial.argtypes = [C.c_int32, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp,
                           _dp]
    L.orc_tri_energy_q.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]
    L.orc_set_eta.argtypes = [P, C.c_double]
    L.orc_num_threads.restype = C.c_int
    L.orc_set_num_threads.argtypes = [C.c_int]
    _lib = L
    return L


def kahan_sum(v: np.ndarray) -> float:
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_kahan_sum(_d(v), len(v))


def patch_fit(ref_pos: np.ndarray, tris: np.ndarray, thresh: float):
    """Oracle restatement of the patch choice, G and H. Returns (pid, G, H, failed_tri)."""
    N = len(ref_pos) // 3
    T = len(tris)
    pid = np.zeros((T, 6), np.int32)
    G = np.zeros((T, 2, 3))
    H = np.zeros((T, 6, 3))
    f = lib().orc_patch_fit(N, T, _d(np.ascontiguousarray(ref_pos)), _i(np.ascontiguousarray(tris, np.int32)),
                            thresh, _i(pid), _d(G), _d(H))
    return pid, G, H, int(f)


def tri_energy_q(q, abar, bbar, h, mu, nu, gamma, want_grad=False):
    """W_s, W_b (and dW/dq for unit area) of one triangle from its 15 derivatives."""
    out = np.zeros(2)
    g = np.zeros(15) if want_grad else None
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (q, abar, bbar)]
    lib().orc_tri_energy_q(_d(arr[0]), _d(arr[1]), _d(arr[2]), h, mu, nu, gamma, _d(out), _d(g))
    return (out, g) if want_grad else out


def dial(T, protocol, from_ansatz, f, abar_info, bbar_info, a0=None, b0=None):
    ab = np.zeros(3 * T)
    bb = np.zeros(3 * T)
    ai = np.ascontiguousarray(abar_info, dtype=np.float64)
    bi = None if bbar_info is None else np.ascontiguousarray(bbar_info, dtype=np.float64)
    lib().orc_dial(T, protocol, 1 if from_ansatz else 0, f, _d(ai), ai.shape[1], _d(bi),
                   0 if bi is None else bi.shape[1], _d(a0), _d(b0), _d(ab), _d(bb))
    return ab, bb


class System:
    """One shell in the oracle: the reference's per-step dataflow in FP64."""

    def __init__(self, N, T, pid, G, H, area, h, mu, abar, bbar, mass, constrained, nu, gamma, rho, hmin, dt, eta,
                 fchar, cchar, gradient_descent=False):
        self.N, self.T = N, T
        self._keep = [np.ascontiguousarray(x) for x in (pid, G, H, area, h, mu, abar, bbar, mass)]
        cl = np.ascontiguousarray(constrained if constrained is not None else np.zeros(N), dtype=np.uint8)
        pid, G, H, area, h, mu, abar, bbar, mass = self._keep
        self.h = lib().orc_create(N, T, _i(pid.astype(np.int32)), _d(G.astype(np.float64)), _d(H), _d(area), _d(h),
                                  _d(mu), _d(abar), _d(bbar), _d(mass), cl.ctypes.data_as(_u8), nu, gamma, rho,
                                  hmin, dt, eta, fchar, cchar, 1 if gradient_descent else 0)

    @classmethod
    def from_shell(cls, sh, abar=None, bbar=None, eta=None, gradient_descent=None):
        """Build from a morphoshell_amd.shell.ShellHost's set-up products."""
        T = sh.T
        st = sh.settings
        if abar is None:
            abar = np.ascontiguousarray(sh.mesh.abar_info.T.reshape(-1))
        if bbar is None:
            bbar = np.ascontiguousarray(sh.mesh.bbar_info.T.reshape(-1))
        sc = sh.scales
        return cls(sh.N, T, sh.patch_ids.astype(np.int32), sh.G, sh.H, sh.areas, sh.thickness, sh.shear_modulus,
                   np.ascontiguousarray(abar, dtype=np.float64), np.ascontiguousarray(bbar, dtype=np.float64),
                   sh.masses, sh.mesh.constrained, st.def_poisson_ratio, st.bend_stiffening_scale_factor,
                   st.ref_density, sc["min_ref_thickness"], sc["timestep"],
                   sc["damping_factor"] if eta is None else eta, sc["char_force_scale"], sc["char_speed_scale"],
                   st.gradient_descent if gradient_descent is None else gradient_descent)

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def set_targets(self, abar=None, bbar=None, h=None, mu=None):
        lib().orc_set_targets(self.h, _d(abar), _d(bbar), _d(h), _d(mu))

    def surface_derivs(self, dofs):
        q = np.zeros(15 * self.T)
        lib().orc_surface_derivs(self.h, _d(np.ascontiguousarray(dofs)), _d(q))
        return q

    def dE_dq(self, q):
        g = np.zeros(15 * self.T)
        lib().orc_dE_dq(self.h, _d(np.ascontiguousarray(q)), _d(g))
        return g

    def deformation_forces(self, dofs):
        f = np.zeros(3 * self.N)
        lib().orc_deformation_forces(self.h, _d(np.ascontiguousarray(dofs)), _d(f))
        return f

    def geometry(self, dofs):
        a, b, n = np.zeros(3 * self.T), np.zeros(3 * self.T), np.zeros(3 * self.T)
        lib().orc_geometry(self.h, _d(np.ascontiguousarray(dofs)), _d(a), _d(b), _d(n))
        return a.reshape(3, -1), b.reshape(3, -1), n.reshape(3, -1)

    def stresses(self, dofs):
        """calc_stresses (calc_stresses.cpp:45-179) restated from the
        oracle's own gradient g = d(A (W_s + W_b))/dq (dE_dq, pinned by
        finite differences): per triangle sigma_ij = (1/J) sum_alpha
        g[R^i_alpha] R^j_alpha with J = |R_x x R_y|; the symmetric matrix of
        its lower triangle (what Eigen's SelfAdjointEigenSolver reads,
        calc_stresses.cpp:162) diagonalised by LAPACK (numpy eigh, lower
        triangle, ascending); the first eigenvalue of least magnitude is
        dropped and the next two kept in cyclic order (:165-176).  Returns
        (8, T): eigval_1, eigvec_1 (3), eigval_2, eigvec_2 (3), each
        eigenvector with its largest-magnitude component positive."""
        q = self.surface_derivs(dofs).reshape(self.T, 15)
        g = self.dE_dq(q.reshape(-1)).reshape(self.T, 15)
        Rx, Ry = q[:, [0, 5, 10]], q[:, [1, 6, 11]]
        gx, gy = g[:, [0, 5, 10]], g[:, [1, 6, 11]]
        J = np.linalg.norm(np.cross(Rx, Ry), axis=1)
        sig = (gx[:, :, None] * Rx[:, None, :] + gy[:, :, None] * Ry[:, None, :]) / J[:, None, None]
        lam, V = np.linalg.eigh(sig, UPLO="L")
        m = np.argmin(np.abs(lam), axis=1)  # first of least magnitude
        out = np.empty((8, self.T))
        rows = np.arange(self.T)
        for r in range(2):
            k = (m + 1 + r) % 3
            vec = V[rows, :, k]
            big = np.argmax(np.abs(vec), axis=1)
            sgn = np.where(vec[rows, big] < 0.0, -1.0, 1.0)
            out[4 * r] = lam[rows, k]
            out[4 * r + 1:4 * r + 4] = (vec * sgn[:, None]).T
        return out

    def energy_densities(self, dofs):
        d = np.zeros(2 * self.T)
        lib().orc_energy_densities(self.h, _d(np.ascontiguousarray(dofs)), _d(d))
        return d.reshape(2, -1)

    def energies(self, dofs, vel):
        out = np.zeros(3)
        lib().orc_energies(self.h, _d(np.ascontiguousarray(dofs)), _d(np.ascontiguousarray(vel)), _d(out))
        return out

    def total_energy(self, dofs):
        e = self.energies(dofs, np.zeros(3 * self.N))
        return e[0] + e[1]

    def damping(self, vel, forces):
        lib().orc_damping(self.h, _d(vel), _d(forces))

    def constraints(self, forces):
        lib().orc_constraints(self.h, _d(forces))

    def advance(self, dofs, vel, forces):
        bad = C.c_double()
        d = lib().orc_advance(self.h, _d(dofs), _d(vel), _d(forces), C.byref(bad))
        return int(d), bad.value

    def equilibrium(self, forces, vel):
        fr, vr = C.c_double(), C.c_double()
        fi, vi = C.c_int32(), C.c_int32()
        lib().orc_equilibrium(self.h, _d(forces), _d(vel), C.byref(fr), C.byref(fi), C.byref(vr), C.byref(vi))
        return fr.value, fi.value, vr.value, vi.value

    def full_forces(self, dofs, vel):
        """Elastic forces + damping + clamp: the force advance_dynamics sees."""
        f = self.deformation_forces(dofs)
        self.damping(np.ascontiguousarray(vel), f)
        self.constraints(f)
        return f

    def run_steps(self, dofs, vel, nsteps):
        f = np.zeros(3 * self.N)
        r = lib().orc_run_steps(self.h, _d(dofs), _d(vel), _d(f), nsteps)
        return int(r)


def run_protocol(sysm: System, dial, x, v, K, spacing, start, interval, eta1, eta2, tol_f, tol_v, max_steps):
    """The reference's dynamics-loop state machine (main.cpp:269-507) over the
    oracle's per-step calls, for one shell, stopping at the first equilibrium
    whose dial factor is 1 (the reference's own loop never stops there,
    main.cpp:493).  ``dial(f)`` returns (abar, bbar) 3T SoA at dial factor f.

    Returns dict(steps, eq_steps, status, dial_factor, events) and updates x, v
    in place (x, v end as the checked state of the final equilibrium, or
    after max_steps).  events lists what the reference's loop logs, in order:
    ("phase_end", step, dial factor) when a dialling phase ends
    (main.cpp:293-304) and ("check", step, dial factor, force ratio, its node,
    speed ratio, its node, verdict) at every equilibrium check
    (main.cpp:396-407, is_at_equilibrium.cpp:89-90).
    """
    import math

    DIAL, WAIT, EQ = 0, 1, 2
    status = DIAL
    step = 0
    phase = int(math.fmod(start, spacing) * K)
    last_eq = start
    f = start
    since_check = -123456
    eq_steps = []
    events = []
    lib().orc_set_eta(sysm.h, eta1)
    while step < max_steps:
        if f >= 1.0:
            f = 1.0
        if status == DIAL and phase >= K:
            status = WAIT
            since_check = 0
            lib().orc_set_eta(sysm.h, eta2)
            events.append(("phase_end", step, f))
        if status == DIAL or step == 0:
            ab, bb = dial(f)
            sysm.set_targets(ab, bb)
        F = sysm.full_forces(x, v)
        if status == WAIT and since_check > interval:
            fr, fi, vr, vi = sysm.equilibrium(F, v)
            events.append(("check", step, f, fr, fi, vr, vi, fr < tol_f and vr < tol_v))
            if fr < tol_f and vr < tol_v:
                status = EQ
                last_eq = f
                eq_steps.append(step)
            since_check = 0
        if status == EQ and f >= 1.0:
            return {"steps": step, "eq_steps": eq_steps, "status": 2, "dial_factor": f, "events": events}
        d, _ = sysm.advance(x, v, F)
        if d >= 0:
            return {"steps": step, "eq_steps": eq_steps, "status": -1, "dial_factor": f, "bad_dof": d,
                    "events": events}
        step += 1
        if status == DIAL:
            phase += 1
        f = last_eq + spacing * (phase / K)
        if status == WAIT:
            since_check += 1
        if status == EQ:
            status = DIAL
            phase = 0
            lib().orc_set_eta(sysm.h, eta1)
    return {"steps": step, "eq_steps": eq_steps, "status": status, "dial_factor": f, "events": events}
