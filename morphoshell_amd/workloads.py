"""The measured workloads, as concrete seeded inputs.

* ``small_defect_disc``: unit disc, spacing 0.04, seed 0 (~4.5k triangles),
  h = 0.01, nu = 0.5, +1 azimuthal-director LCE metric (director angle
  phi + pi/2) dialled to lambda = 0.9 with opto-thermal nu = 0.5 over 20k
  steps (do_dialling.cpp:104-124 protocol), bbar = 0, Gaussian z-noise
  sigma = 1e-4 (seed 0), damping prefactors 1.
* ``grf_disc``: the large single shell (spacing 0.002 -> ~1.81M triangles),
  lambda = 1 + 0.1 g, g a 64-mode Gaussian random field of correlation length
  0.2, abar = lambda^2 I, bbar = 0, h = 0.005.
* ``ragged_batch``: discs and squares with triangle counts log-uniform in
  [2k, 100k], per-shell GRF lambda (amplitude U(0.05, 0.2)), bbar = kappa I
  with kappa ~ U(-2, 2), h ~ U(0.002, 0.02).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .meshing import (Mesh, disc_mesh, grf, isotropic_abar_info, lce_defect_abar_info, square_mesh,
                      uniform_bbar_info)
from .shell import Settings, prepare_shell

__all__ = ["small_defect_disc", "grf_disc", "ragged_batch_meshes", "ragged_batch", "flat_noisy_state"]


def small_defect_disc(spacing: float = 0.04, seed: int = 0):
    m = disc_mesh(spacing, seed=seed)
    m.abar_info = lce_defect_abar_info(m.centroids(), 0.9, 0.5)
    m.bbar_info = np.zeros((m.num_tris, 3))
    m.fill_defaults(h=0.01, mu=1.0)
    st = Settings(def_poisson_ratio=0.5, dial_protocol=3, damping_prefactor_1=1.0, damping_prefactor_2=1.0,
                  timesteps_per_dial_phase=20000)
    return prepare_shell(m, st)


def grf_disc(spacing: float = 0.002, seed: int = 0, h: float = 0.005, amp: float = 0.1, corr: float = 0.2):
    m = disc_mesh(spacing, seed=seed)
    c = m.centroids()
    m.abar_info = isotropic_abar_info(1.0 + amp * grf(c[:, :2], corr, n_modes=64, seed=seed + 1))
    m.bbar_info = np.zeros((m.num_tris, 3))
    m.fill_defaults(h=h, mu=1.0)
    return prepare_shell(m, Settings(def_poisson_ratio=0.5))


def flat_noisy_state(sh, sigma: float = 1e-4, seed: int = 0):
    rng = np.random.default_rng(seed)
    x = sh.ref_pos.copy()
    x[2 * sh.N:] += sigma * rng.standard_normal(sh.N)
    return x, np.zeros_like(x)


def _batch_mesh(args) -> Mesh:
    i, seed, tmin, tmax = args
    rng = np.random.default_rng([seed, i])
    T = math.exp(rng.uniform(math.log(tmin), math.log(tmax)))
    kind = "disc" if rng.uniform() < 0.5 else "square"
    if kind == "disc":
        m = disc_mesh(math.sqrt(7.26 / T), seed=seed * 100003 + i)
    else:
        m = square_mesh(math.sqrt(9.24 / T), seed=seed * 100003 + i)
    amp = rng.uniform(0.05, 0.2)
    kappa = rng.uniform(-2.0, 2.0)
    h = rng.uniform(0.002, 0.02)
    c = m.centroids()
    # the field is clipped at 3 standard deviations: an unclipped 64-mode
    # sample can reach |g| ~ 5, and with amplitude 0.2 that programs
    # lambda ~ 0 (abar ~ 0), which no explicit step survives
    g = np.clip(grf(c[:, :2], 0.2, n_modes=64, seed=seed * 7919 + i), -3.0, 3.0)
    m.abar_info = isotropic_abar_info(1.0 + amp * g)
    m.bbar_info = uniform_bbar_info(m.num_tris, kappa)
    m.fill_defaults(h=h, mu=1.0)
    m.meta.update({"amp": amp, "kappa": kappa, "h": h, "kind": kind})
    return m


def ragged_batch_meshes(num_shells: int = 512, seed: int = 0, tmin: int = 2000, tmax: int = 100000,
                        workers: int | None = None) -> list[Mesh]:
    workers = workers or min(16, os.cpu_count() or 1)
    jobs = [(i, seed, tmin, tmax) for i in range(num_shells)]
    if workers <= 1:
        return [_batch_mesh(j) for j in jobs]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_batch_mesh, jobs, chunksize=4))


def ragged_batch(num_shells: int = 512, seed: int = 0, tmin: int = 2000, tmax: int = 100000, precision="mixed"):
    meshes = ragged_batch_meshes(num_shells, seed, tmin, tmax)
    st = Settings(def_poisson_ratio=0.5, precision=precision)
    return [prepare_shell(m, st) for m in meshes]
