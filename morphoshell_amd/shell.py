"""Host-side mirror of the reference's per-step interface.

``Settings`` is the reference's settings file (read_settings.cpp:137-238: 22
keys in a fixed order) plus this project's appended keys.  ``prepare_shell``
runs the once-per-run set-up (calc_mesh_info, set_up_patch_fitting,
calc_deriv_mats, calc_dof_masses and the scale calculators) through the
native library.  ``Context`` owns a device context for one or many shells and
exposes the reference's per-step calls by their reference names
(calc_deformation_forces, calc_non_deformation_forces, impose_constraints,
advance_dynamics, is_at_equilibrium, calc_energies,
calc_a_comps_and_b_comps_and_normals) plus the fused ``step``.

Errors keep the reference's behaviour: a force blow-up raises
``HighForceError`` whose message is "implausibly_high_force"
(advance_dynamics.cpp:58-73) and which names the offending dof, node and force.
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field, fields

import numpy as np

from . import native as nat
from .meshing import Mesh

__all__ = ["Settings", "read_settings", "ShellHost", "prepare_shell", "Context", "HighForceError"]


class HighForceError(RuntimeError):
    """advance_dynamics.cpp:58-73: raised as std::runtime_error("implausibly_high_force")."""

    def __init__(self, shell: int, dof: int, node: int, force: float, step: int):
        super().__init__("implausibly_high_force")
        self.shell, self.dof, self.node, self.force, self.step = shell, dof, node, force, step


# ------------------------------------------------------------------ settings

REFERENCE_KEYS = [
    "outputs_per_char_long_time",
    "ref_thickness_if_uniform",
    "timestep_prefactor",
    "dialling_from_ansatz_rather_than_ref_state",
    "dial_factors_to_hold_at_spacing",
    "dial_phase_time_prefactor",
    "time_between_equil_checks_prefactor",
    "damping_prefactor_1",
    "damping_prefactor_2",
    "num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics",
    "force_to_char_force_ratio_equil_threshold",
    "speed_to_char_speed_ratio_equil_threshold",
    "def_poisson_ratio",
    "ref_shear_modulus_if_uniform",
    "ref_density",
    "bend_stiffening_scale_factor",
    "patch_mat_dimless_conditioning_thresh",
    "slide_stiffness_prefactor",
    "slide_friction_coefficent",
    "slide_speed_prefactor",
    "using_initial_glass_slide_z_coords_from_settings_file",
    "initial_upper_slide_z_coord",
    "initial_lower_slide_z_coord",
]


@dataclass
class Settings:
    # the reference's keys, in file order (read_settings.cpp:137-238)
    outputs_per_char_long_time: float = -1.0
    ref_thickness_if_uniform: float = -1.0
    timestep_prefactor: float = 0.6
    dialling_from_ansatz_rather_than_ref_state: bool = False
    dial_factors_to_hold_at_spacing: float = 1.0
    dial_phase_time_prefactor: float = 0.25
    time_between_equil_checks_prefactor: float = 0.01
    damping_prefactor_1: float = 1.0
    damping_prefactor_2: float = 1.0
    num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics: int = 0
    force_to_char_force_ratio_equil_threshold: float = 1e-6
    speed_to_char_speed_ratio_equil_threshold: float = 1e-5
    def_poisson_ratio: float = 0.5
    ref_shear_modulus_if_uniform: float = -1.0
    ref_density: float = 1.0
    bend_stiffening_scale_factor: float = 1e17
    patch_mat_dimless_conditioning_thresh: float = 10.0
    slide_stiffness_prefactor: float = 0.0
    slide_friction_coefficent: float = 0.0
    slide_speed_prefactor: float = 0.0
    using_initial_glass_slide_z_coords_from_settings_file: bool = False
    initial_upper_slide_z_coord: float = 12345.0
    initial_lower_slide_z_coord: float = 12345.0
    # appended keys (accepted after the reference's 23rd)
    precision: str = "fp64"          # fp64 | mixed
    dial_protocol: int = 1           # 1 blend (do_dialling.cpp:83-84) | 3 LCE (104-124)
    timesteps_per_dial_phase: int = 0  # > 0 overrides the derived dial-phase length
    extras: dict = field(default_factory=dict)

    @property
    def gradient_descent(self) -> bool:
        return self.num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics > 0


def _boolstr(v: str) -> bool:
    # read_settings.cpp:290-300 accepts exactly "true" and "false"
    if v == "true":
        return True
    if v == "false":
        return False
    raise ValueError("Error: you passed a std::string to boolstring_to_bool that read neither true nor false.")


def read_settings(text: str) -> Settings:
    """Parse a settings file as read_settings.cpp:60-238 does.

    ``//`` comments are removed to the end of the line, all whitespace is
    removed, the rest must be ``key=value;`` pairs whose first 23 keys are the
    reference's, in its order.  Appended keys of this project may follow.
    """
    s = text.replace("\r", "\n") + "\n"
    s = re.sub(r"//[^\n]*\n", "", s)
    s = re.sub(r"\s+", "", s)
    eq = [m.start() for m in re.finditer("=", s)]
    sc = [m.start() for m in re.finditer(";", s)]
    if len(eq) != len(sc):
        raise ValueError("settings file: unequal numbers of = and ; (the most common mistake is a missing semicolon)")
    seq = [p for pair in zip(eq, sc) for p in pair]
    if seq != sorted(seq):
        raise ValueError("settings file: = and ; are not in the correct order")
    pairs, cur = [], 0
    for e, c in zip(eq, sc):
        pairs.append((s[cur:e], s[e + 1:c]))
        cur = c + 1
    st = Settings()
    types = {f.name: f.type for f in fields(Settings)}
    for i, key in enumerate(REFERENCE_KEYS):
        if i >= len(pairs) or pairs[i][0] != key:
            raise ValueError(f"Error: something went wrong when reading the {key} setting.")
        val = pairs[i][1]
        t = types[key]
        if t in ("bool", bool):
            setattr(st, key, _boolstr(val))
        elif t in ("int", int):
            setattr(st, key, int(val))
        else:
            setattr(st, key, float(val))
    for key, val in pairs[len(REFERENCE_KEYS):]:
        if key == "precision":
            if val not in ("fp64", "mixed"):
                raise ValueError("precision must be fp64 or mixed")
            st.precision = val
        elif key in ("dial_protocol", "timesteps_per_dial_phase"):
            setattr(st, key, int(val))
        else:
            st.extras[key] = val
    if st.dial_factors_to_hold_at_spacing > 1.0:
        raise ValueError("Error: The largest permitted value of dial_factors_to_hold_at_spacing is 1.0")
    return st


# ------------------------------------------------------------------- set-up

@dataclass
class ShellHost:
    """Host-side set-up products of one shell (the kernels' input streams)."""

    mesh: Mesh
    settings: Settings
    ref_pos: np.ndarray          # 3N dofs SoA
    areas: np.ndarray            # T
    centroids: np.ndarray        # 3T SoA
    info: dict
    patch_ids: np.ndarray        # (T, 6) int32
    G: np.ndarray                # (T, 2, 3)
    H: np.ndarray                # (T, 6, 3)
    thickness: np.ndarray        # T  (ref = def, dialling protocol 1 of thickness)
    shear_modulus: np.ndarray    # T
    masses: np.ndarray           # N
    scales: dict
    num_retried: int = 0

    @property
    def N(self) -> int:
        return self.mesh.num_nodes

    @property
    def T(self) -> int:
        return self.mesh.num_tris

    def params(self, damping_prefactor: float | None = None, gradient_descent: bool | None = None) -> nat.ShellParams:
        st = self.settings
        sc = self.scales
        gd = st.gradient_descent if gradient_descent is None else gradient_descent
        eta = sc["damping_factor"] if damping_prefactor is None else \
            scales_for(self, damping_prefactor, gd)["damping_factor"]
        p = nat.ShellParams()
        p.poisson_ratio = st.def_poisson_ratio
        p.bend_stiffening_scale_factor = st.bend_stiffening_scale_factor
        p.ref_density = st.ref_density
        p.min_ref_thickness = sc["min_ref_thickness"]
        p.timestep = sc["timestep"]
        p.damping_factor = eta
        p.char_force_scale = sc["char_force_scale"]
        p.char_speed_scale = sc["char_speed_scale"]
        p.gradient_descent = 1 if gd else 0
        p.dial_protocol = st.dial_protocol
        p.dial_from_ansatz = 1 if st.dialling_from_ansatz_rather_than_ref_state else 0
        return p

    def dofs_from_points(self, pts: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(pts.T.reshape(-1))


def scales_for(sh: ShellHost, damping_prefactor: float, gradient_descent: bool | None = None) -> dict:
    st = sh.settings
    s = nat.Scales()
    s.poisson_ratio = st.def_poisson_ratio
    s.ref_density = st.ref_density
    s.timestep_prefactor = st.timestep_prefactor
    s.damping_prefactor = damping_prefactor
    s.dial_phase_time_prefactor = st.dial_phase_time_prefactor
    s.time_between_equil_checks_prefactor = st.time_between_equil_checks_prefactor
    s.outputs_per_char_long_time = st.outputs_per_char_long_time
    s.gradient_descent = int(st.gradient_descent if gradient_descent is None else gradient_descent)
    L = nat.lib()
    nat.check(L.ms_calc_scales(sh.T, nat.dptr(sh.thickness), nat.dptr(sh.shear_modulus), nat.dptr(sh.areas),
                               sh.info["perimeter"], sh.info["approx_min_tri_size"], C.byref(s)))
    d = {f: getattr(s, f) for f, _ in s._fields_}
    if st.timesteps_per_dial_phase > 0:
        d["timesteps_per_dial_phase"] = st.timesteps_per_dial_phase
    return d


def prepare_shell(mesh: Mesh, settings: Settings | None = None) -> ShellHost:
    """Run the once-per-run set-up for one shell (host, OpenMP)."""
    st = settings or Settings()
    mesh.fill_defaults()
    L = nat.lib()
    N, T = mesh.num_nodes, mesh.num_tris
    ref_pos = np.ascontiguousarray(mesh.points.T.reshape(-1), dtype=np.float64)
    tris = np.ascontiguousarray(mesh.tris, dtype=np.int32)
    areas = np.empty(T)
    cents = np.empty(3 * T)
    bnodes = np.empty(N, dtype=np.uint8)
    info = nat.MeshInfoOut()
    nat.check(L.ms_mesh_info(N, T, nat.dptr(ref_pos), nat.iptr(tris), nat.dptr(areas), nat.dptr(cents),
                             nat.u8ptr(bnodes), C.byref(info)))
    pid = np.empty((T, 6), dtype=np.int32)
    G = np.empty((T, 2, 3))
    H = np.empty((T, 6, 3))
    nret = C.c_int32(0)
    nat.check(L.ms_patch_fit(N, T, nat.dptr(ref_pos), nat.iptr(tris), st.patch_mat_dimless_conditioning_thresh,
                             nat.iptr(pid), nat.dptr(G), nat.dptr(H), C.byref(nret)))
    h = np.ascontiguousarray(mesh.thickness, dtype=np.float64)
    if st.ref_thickness_if_uniform > 0:
        h = np.full(T, st.ref_thickness_if_uniform)
    mu = np.ascontiguousarray(mesh.shear_modulus, dtype=np.float64)
    if st.ref_shear_modulus_if_uniform > 0:
        mu = np.full(T, st.ref_shear_modulus_if_uniform)
    masses = np.empty(N)
    nat.check(L.ms_dof_masses(N, T, nat.iptr(tris), nat.dptr(areas), nat.dptr(h), st.ref_density,
                              nat.dptr(masses)))
    infod = {f: getattr(info, f) for f, _ in info._fields_}
    infod["boundary_nodes"] = bnodes
    sh = ShellHost(mesh=mesh, settings=st, ref_pos=ref_pos, areas=areas, centroids=cents, info=infod,
                   patch_ids=pid, G=G, H=H, thickness=h, shear_modulus=mu, masses=masses, scales={},
                   num_retried=int(nret.value))
    sh.scales = scales_for(sh, st.damping_prefactor_1)
    return sh


# ------------------------------------------------------------------ context

class Context:
    """A device context over one or many shells that step together."""

    def __init__(self, shells: list[ShellHost] | ShellHost, precision: str | None = None,
                 a0: list | None = None, b0: list | None = None):
        if isinstance(shells, ShellHost):
            shells = [shells]
        self.shells = list(shells)
        prec = precision or self.shells[0].settings.precision
        if prec not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {prec!r}")
        self.precision = nat.MS_MIXED if prec == "mixed" else nat.MS_FP64
        L = nat.lib()
        descs = (nat.ShellDesc * len(self.shells))()
        self._keep = []
        for i, sh in enumerate(self.shells):
            d = descs[i]
            st = sh.settings
            m = sh.mesh
            keep = {}
            keep["pid"] = np.ascontiguousarray(sh.patch_ids, dtype=np.int32)
            keep["G"] = np.ascontiguousarray(sh.G, dtype=np.float64)
            keep["H"] = np.ascontiguousarray(sh.H, dtype=np.float64)
            keep["ai"] = np.ascontiguousarray(m.abar_info, dtype=np.float64)
            keep["bi"] = np.ascontiguousarray(m.bbar_info, dtype=np.float64)
            keep["cl"] = np.ascontiguousarray(m.constrained, dtype=np.uint8)
            if a0 is not None and a0[i] is not None:
                keep["a0"] = np.ascontiguousarray(a0[i], dtype=np.float64)
            if b0 is not None and b0[i] is not None:
                keep["b0"] = np.ascontiguousarray(b0[i], dtype=np.float64)
            d.num_nodes = sh.N
            d.num_tris = sh.T
            d.patch_ids = nat.iptr(keep["pid"])
            d.G = nat.dptr(keep["G"])
            d.H = nat.dptr(keep["H"])
            keep["rp"] = np.ascontiguousarray(sh.ref_pos, dtype=np.float64)
            d.ref_pos = nat.dptr(keep["rp"])
            d.ref_areas = nat.dptr(sh.areas)
            d.thickness = nat.dptr(sh.thickness)
            d.shear_modulus = nat.dptr(sh.shear_modulus)
            d.abar_info = nat.dptr(keep["ai"])
            d.abar_info_cols = keep["ai"].shape[1]
            d.bbar_info = nat.dptr(keep["bi"])
            d.bbar_info_cols = keep["bi"].shape[1]
            d.a0 = nat.dptr(keep.get("a0"))
            d.b0 = nat.dptr(keep.get("b0"))
            d.node_masses = nat.dptr(sh.masses)
            d.constrained = nat.u8ptr(keep["cl"])
            d.params = sh.params()
            self._keep.append(keep)
        h = C.c_void_p()
        nat.check(L.ms_create(descs, len(self.shells), self.precision, C.byref(h)))
        self.h = h
        self.L = L
        self._keep = None  # arrays were copied

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.ms_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, code: int) -> None:
        nat.check(code, self.h)

    def _vec(self, a, n: int, what: str) -> np.ndarray:
        """a as a contiguous float64 array of n values (the C-ABI reads n)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != n:
            raise ValueError(f"{what}: expected {n} values, got {a.size}")
        return a

    def _shell(self, shell: int):
        if not 0 <= shell < len(self.shells):
            raise IndexError(f"shell {shell} out of range (0..{len(self.shells) - 1})")
        return self.shells[shell]

    # -- state
    def set_state(self, shell: int, dofs: np.ndarray, vel: np.ndarray | None = None) -> None:
        n = 3 * self._shell(shell).N
        dofs = self._vec(dofs, n, "set_state dofs")
        vel = None if vel is None else self._vec(vel, n, "set_state velocities")
        self._ck(self.L.ms_set_state(self.h, shell, nat.dptr(dofs), nat.dptr(vel)))

    def get_state(self, shell: int = 0) -> tuple[np.ndarray, np.ndarray]:
        n = self._shell(shell).N
        x, v = np.empty(3 * n), np.empty(3 * n)
        self._ck(self.L.ms_get_state(self.h, shell, nat.dptr(x), nat.dptr(v)))
        return x, v

    def get_forces(self, shell: int = 0) -> np.ndarray:
        f = np.empty(3 * self._shell(shell).N)
        self._ck(self.L.ms_get_forces(self.h, shell, nat.dptr(f)))
        return f

    def set_forces(self, shell: int, f: np.ndarray) -> None:
        f = self._vec(f, 3 * self._shell(shell).N, "set_forces")
        self._ck(self.L.ms_set_forces(self.h, shell, nat.dptr(f)))

    def set_dial_clock(self, shell: int, last_eq: float, spacing: float, steps_per_phase: int, phase_steps: int,
                       dialling: bool) -> None:
        self._ck(self.L.ms_set_dial_clock(self.h, shell, last_eq, spacing, int(steps_per_phase), int(phase_steps),
                                          1 if dialling else 0))

    def dial_factor(self, shell: int = 0) -> float:
        f = C.c_double()
        ph = C.c_int64()
        self._ck(self.L.ms_get_dial_factor(self.h, shell, C.byref(f), C.byref(ph)))
        return f.value

    def set_shell_params(self, shell: int, p: nat.ShellParams) -> None:
        self._ck(self.L.ms_set_shell_params(self.h, shell, C.byref(p)))

    def set_zero_dial_geometry(self, shell: int, a0: np.ndarray | None, b0: np.ndarray | None) -> None:
        n = 3 * self._shell(shell).T
        a0 = None if a0 is None else self._vec(a0, n, "set_zero_dial_geometry a0")
        b0 = None if b0 is None else self._vec(b0, n, "set_zero_dial_geometry b0")
        self._ck(self.L.ms_set_zero_dial_geometry(self.h, shell, nat.dptr(a0), nat.dptr(b0)))

    def set_precision(self, precision: str) -> None:
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.precision = nat.MS_MIXED if precision == "mixed" else nat.MS_FP64
        self._ck(self.L.ms_set_precision(self.h, self.precision))

    # -- the reference's per-step calls
    def calc_deformation_forces(self) -> None:
        self._ck(self.L.ms_calc_deformation_forces(self.h))

    def calc_non_deformation_forces(self) -> None:
        self._ck(self.L.ms_calc_non_deformation_forces(self.h))

    def impose_constraints(self) -> None:
        self._ck(self.L.ms_impose_constraints(self.h))

    def advance_dynamics(self) -> list[dict]:
        reps = (nat.StepReport * len(self.shells))()
        code = self.L.ms_advance_dynamics(self.h, reps)
        out = [r.as_dict() for r in reps]
        if code == nat.MS_ERR_HIGH_FORCE:
            self._raise_high_force(out)
        self._ck(code)
        return out

    def _raise_high_force(self, reps: list[dict]) -> None:
        for s, r in enumerate(reps):
            if r["status"] == nat.MS_ERR_HIGH_FORCE and r["bad_dof"] >= 0:
                n = self.shells[s].N
                raise HighForceError(s, r["bad_dof"], r["bad_dof"] % n, r["bad_force"], r["fail_step"])
        raise HighForceError(-1, -1, -1, float("nan"), -1)

    def is_at_equilibrium(self, shell: int = 0, force_tol: float | None = None,
                          speed_tol: float | None = None) -> dict:
        st = self._shell(shell).settings
        ft = st.force_to_char_force_ratio_equil_threshold if force_tol is None else force_tol
        vt = st.speed_to_char_speed_ratio_equil_threshold if speed_tol is None else speed_tol
        r = nat.StepReport()
        self._ck(self.L.ms_is_at_equilibrium(self.h, shell, ft, vt, C.byref(r)))
        return r.as_dict()

    def calc_energies(self, shell: int = 0) -> np.ndarray:
        out = np.empty(3)
        self._ck(self.L.ms_calc_energies(self.h, shell, nat.dptr(out)))
        return out

    def calc_a_comps_and_b_comps_and_normals(self, shell: int = 0):
        T = self._shell(shell).T
        a, b, n = np.empty(3 * T), np.empty(3 * T), np.empty(3 * T)
        self._ck(self.L.ms_calc_geometry(self.h, shell, nat.dptr(a), nat.dptr(b), nat.dptr(n)))
        return a.reshape(3, T), b.reshape(3, T), n.reshape(3, T)

    def calc_energy_densities(self, shell: int = 0) -> np.ndarray:
        T = self._shell(shell).T
        d = np.empty(2 * T)
        self._ck(self.L.ms_calc_energy_densities(self.h, shell, nat.dptr(d)))
        return d.reshape(2, T)

    def calc_diagnostics(self, shell: int = 0) -> dict:
        """Gauss and mean curvature and the strain measure per triangle
        (calc_curvatures.cpp:35-94, calc_strains.cpp:35-78)."""
        T = self._shell(shell).T
        g, m, s = np.empty(T), np.empty(T), np.empty(T)
        self._ck(self.L.ms_calc_diagnostics(self.h, shell, nat.dptr(g), nat.dptr(m), nat.dptr(s)))
        return {"gauss_curv": g, "mean_curv": m, "strain": s}

    def calc_stresses(self, shell: int = 0) -> np.ndarray:
        """calc_stresses (calc_stresses.cpp:45-179): (8, T) per triangle --
        eigenvalue 1, eigenvector 1 (3), eigenvalue 2, eigenvector 2 (3) of
        the reference-area-weighted Cauchy stress; write_output divides the
        eigenvalues by the reference shear modulus times thickness."""
        T = self._shell(shell).T
        out = np.empty(8 * T)
        self._ck(self.L.ms_calc_stresses(self.h, shell, nat.dptr(out)))
        return out.reshape(8, T)

    def calc_angle_deficits(self, shell: int = 0) -> np.ndarray:
        """calc_angle_deficits (calc_angle_deficits.cpp:40-78): per node 2 pi
        (pi on the boundary) minus its triangles' corner angles."""
        sh = self._shell(shell)
        bnd = np.ascontiguousarray(sh.info["boundary_nodes"], dtype=np.uint8)
        out = np.empty(sh.N)
        self._ck(self.L.ms_calc_angle_deficits(self.h, shell, bnd.ctypes.data, nat.dptr(out)))
        return out

    # -- fused path
    def step(self, nsteps: int, equil_every: int = 0, force_tol: float = 1e-6, speed_tol: float = 1e-5,
             halt_on_equilibrium: bool = False, raise_on_blowup: bool = True) -> list[dict]:
        reps = (nat.StepReport * len(self.shells))()
        code = self.L.ms_step(self.h, int(nsteps), int(equil_every), force_tol, speed_tol,
                              1 if halt_on_equilibrium else 0, reps)
        out = [r.as_dict() for r in reps]
        if code == nat.MS_ERR_HIGH_FORCE:
            if raise_on_blowup:
                self._raise_high_force(out)
            return out
        self._ck(code)
        return out

    def resume(self, shell: int) -> None:
        self._ck(self.L.ms_resume_shell(self.h, shell))

    # -- device-run relaxation protocol (main.cpp:269-507)
    def relax_setup(self, shell: int, steps_per_dial_phase: int | None = None, dial_spacing: float | None = None,
                    dial_start: float = 0.0, equil_interval: int | None = None, tol_f: float | None = None,
                    tol_v: float | None = None, pause_on_eq: bool = False) -> None:
        sh = self._shell(shell)
        st, sc = sh.settings, sh.scales
        p = nat.RelaxParams()
        p.tol_f = st.force_to_char_force_ratio_equil_threshold if tol_f is None else tol_f
        p.tol_v = st.speed_to_char_speed_ratio_equil_threshold if tol_v is None else tol_v
        p.equil_interval = sc["timesteps_between_equil_checks"] if equil_interval is None else equil_interval
        p.steps_per_dial_phase = sc["timesteps_per_dial_phase"] if steps_per_dial_phase is None else \
            steps_per_dial_phase
        p.dial_spacing = st.dial_factors_to_hold_at_spacing if dial_spacing is None else dial_spacing
        p.dial_start = dial_start
        p.eta1 = scales_for(sh, st.damping_prefactor_1)["damping_factor"]
        p.eta2 = scales_for(sh, st.damping_prefactor_2)["damping_factor"]
        p.pause_on_eq = 1 if pause_on_eq else 0
        self._ck(self.L.ms_relax_setup(self.h, shell, C.byref(p)))

    def relax_run(self, max_steps: int, poll_every: int = 2000) -> list[dict]:
        reps = (nat.RelaxReport * len(self.shells))()
        self._ck(self.L.ms_relax_run(self.h, int(max_steps), int(poll_every), reps))
        return [r.as_dict() for r in reps]

    def set_check_log(self, capacity: int) -> None:
        """Record every protocol equilibrium check and dialling-phase end of
        each shell in a device ring of `capacity` records (0 turns it off)."""
        self._ck(self.L.ms_set_check_log(self.h, int(capacity)))

    def read_check_log(self, shell: int, max_records: int = 1 << 20) -> tuple[list[dict], int]:
        """Protocol events of the shell since the previous read, oldest first,
        and the number overwritten before they were read."""
        buf = (nat.CheckRecord * max_records)()
        n, dropped = C.c_int32(0), C.c_int64(0)
        self._ck(self.L.ms_read_check_log(self.h, shell, buf, max_records, C.byref(n), C.byref(dropped)))
        return [buf[i].as_dict() for i in range(n.value)], int(dropped.value)

    @property
    def steps_taken(self) -> int:
        return int(self.L.ms_steps_taken(self.h))

    def set_tuning(self, tri_block=0, node_block=0, tri_grid_per_cu=-1, node_grid_per_cu=-1, graph_steps=0,
                   fused=-1, tile_chunks=0, tile_nodes=0):
        t = nat.Tuning()
        t.tri_block, t.node_block = tri_block, node_block
        t.tri_grid_per_cu, t.node_grid_per_cu, t.graph_steps = tri_grid_per_cu, node_grid_per_cu, graph_steps
        t.fused = fused
        t.tile_chunks = tile_chunks
        t.tile_nodes = tile_nodes
        self._ck(self.L.ms_set_tuning(self.h, C.byref(t)))

    def get_tuning(self) -> dict:
        t = nat.Tuning()
        self._ck(self.L.ms_get_tuning(self.h, C.byref(t)))
        return {"tri_block": t.tri_block, "node_block": t.node_block, "tri_grid_per_cu": t.tri_grid_per_cu,
                "node_grid_per_cu": t.node_grid_per_cu, "graph_steps": t.graph_steps, "fused": t.fused,
                "tile_chunks": t.tile_chunks, "tile_nodes": t.tile_nodes}

    def enable_timing(self, on: bool = True) -> None:
        self._ck(self.L.ms_enable_timing(self.h, 1 if on else 0))

    def last_step_timing(self) -> np.ndarray:
        out = np.empty(3)
        self._ck(self.L.ms_last_step_timing(self.h, nat.dptr(out)))
        return out

    @property
    def device_bytes(self) -> int:
        return int(self.L.ms_device_bytes(self.h))

    def synchronize(self) -> None:
        self._ck(self.L.ms_synchronize(self.h))
