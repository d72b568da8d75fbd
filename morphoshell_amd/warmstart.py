"""Learned warm start: a mesh message-passing network that predicts
near-equilibrium 3-D node positions from the flat reference mesh and the
programmed metric (abar, bbar, h).

It replaces the ansatz start of the reference
(move_nodes_to_ansatz_positions.cpp:26-39 with positions read by
read_vtk_data.cpp:380-445): the predicted positions become the start state,
dialled by default from their own geometry (the reference's
dialling_from_ansatz_rather_than_ref_state, do_dialling.cpp:194-228: the
start's a and b are the targets at dial factor 0, so it starts force-free)
or, with start="direct", started at dial factor 1.  The plain damped
dynamics finish the relaxation and judge it by residual force, and a flat
start with noise is used instead whenever the prediction has a higher
elastic energy than that flat start.

The network itself (forward, backward, Adam) runs on the GPU in the native
library (``ms_nn_*``, BF16 matrix-core GEMMs with FP32 accumulation).  This
module builds graphs and features, generates training equilibria with the
batched solver, trains, and applies the model.

Features per node (10): centred reference position / L (2), boundary flag,
mean over incident triangles of (abar_xx - 1, abar_xy, abar_yy - 1) and of
bbar * L / 4 (3), and h / l; per directed edge (3): (dx, dy, |d|) / l, with
L the largest distance from the centroid and l the mean edge length.
Output per node: 10 * (displacement / L) in the frame that aligns the
equilibrium with the reference (centroid and Kabsch rotation removed); the
up/down reflection is left to the loss (min over z -> -z).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys
import time

import numpy as np

from . import native as nat

NODE_FEATURES = 10
EDGE_FEATURES = 3
WIDTH = 64
LAYERS = 4
OUT_SCALE = 10.0


def log(msg: str) -> None:
    print(f"[warmstart] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ graphs

def final_metric(sh) -> tuple[np.ndarray, np.ndarray]:
    """abar, bbar (T x 3) the shell is programmed to reach at dial factor 1."""
    m = sh.mesh
    if sh.settings.dial_protocol == 3:
        ang, lam, nu = m.abar_info[:, 0], m.abar_info[:, 1], m.abar_info[:, 2]
        c, s = np.cos(ang), np.sin(ang)
        l2, lp = lam * lam, lam ** (-2.0 * nu)
        ab = np.stack([l2 * c * c + lp * s * s, (l2 - lp) * c * s, lp * c * c + l2 * s * s], axis=1)
        return ab, np.zeros_like(ab)
    return np.asarray(m.abar_info[:, :3], float), np.asarray(m.bbar_info[:, :3], float)


def shell_graph(sh) -> dict:
    m = sh.mesh
    N, T = sh.N, sh.T
    tris = m.tris.astype(np.int64)
    pairs = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]], axis=0)
    pairs = np.concatenate([pairs, pairs[:, ::-1]], axis=0)
    key = np.unique(pairs[:, 1] * N + pairs[:, 0])  # sorted by (dst, src)
    dst = (key // N).astype(np.int32)
    src = (key % N).astype(np.int32)
    xy = m.points[:, :2]
    cen = xy.mean(axis=0)
    rel = xy - cen
    L = float(np.sqrt((rel ** 2).sum(1)).max())
    d = xy[dst] - xy[src]
    ell = float(np.sqrt((d ** 2).sum(1)).mean())
    ab, bb = final_metric(sh)
    # per-node means over incident triangles
    cnt = np.bincount(tris.ravel(), minlength=N).astype(float)
    def node_mean(v):
        acc = np.zeros(N)
        for k in range(3):
            np.add.at(acc, tris[:, k], v)
        return acc / cnt
    h = sh.thickness
    feat = np.zeros((N, NODE_FEATURES), np.float32)
    feat[:, 0:2] = rel / L
    feat[:, 2] = sh.info["boundary_nodes"].astype(np.float32)
    feat[:, 3] = node_mean(ab[:, 0] - 1.0)
    feat[:, 4] = node_mean(ab[:, 1])
    feat[:, 5] = node_mean(ab[:, 2] - 1.0)
    for k in range(3):
        feat[:, 6 + k] = node_mean(bb[:, k] * L / 4.0)
    feat[:, 9] = node_mean(h / ell)
    ef = np.zeros((len(src), EDGE_FEATURES), np.float32)
    ef[:, 0:2] = d / ell
    ef[:, 2] = np.sqrt((d ** 2).sum(1)) / ell
    return {"N": N, "src": src, "dst": dst, "node_feat": feat, "edge_feat": ef, "L": L, "ell": ell,
            "center": cen}


def batch_graphs(graphs: list[dict]) -> dict:
    offs = np.zeros(len(graphs) + 1, np.int32)
    srcs, dsts, nfs, efs = [], [], [], []
    for i, g in enumerate(graphs):
        offs[i + 1] = offs[i] + g["N"]
        srcs.append(g["src"] + offs[i])
        dsts.append(g["dst"] + offs[i])
        nfs.append(g["node_feat"])
        efs.append(g["edge_feat"])
    return {"offsets": offs, "src": np.concatenate(srcs).astype(np.int32), "dst": np.concatenate(dsts).astype(np.int32),
            "node_feat": np.ascontiguousarray(np.concatenate(nfs), np.float32),
            "edge_feat": np.ascontiguousarray(np.concatenate(efs), np.float32)}


def kabsch_target(sh, graph: dict, dofs: np.ndarray) -> np.ndarray:
    """Network target of an equilibrium: 10 * displacement / L after removing
    the rigid motion that best maps the equilibrium onto the flat reference."""
    N = sh.N
    P = dofs.reshape(3, N).T.copy()
    R = sh.mesh.points.copy()
    Pc, Rc = P.mean(0), R.mean(0)
    H = (P - Pc).T @ (R - Rc)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    Rot = Vt.T @ D @ U.T
    Pa = (P - Pc) @ Rot.T + Rc
    disp = Pa - R
    return (OUT_SCALE * disp / graph["L"]).astype(np.float32)


def positions_from_output(sh, graph: dict, out: np.ndarray) -> np.ndarray:
    """dofs (3N, SoA) of the reference positions moved by the network output."""
    disp = out.astype(np.float64) * graph["L"] / OUT_SCALE
    P = sh.mesh.points + disp
    return np.ascontiguousarray(P.T.reshape(-1))


# --------------------------------------------------------------- the model

class WarmStartNet:
    def __init__(self, seed: int = 0, width: int = WIDTH, layers: int = LAYERS):
        a = nat.NNArch()
        a.node_features, a.edge_features, a.width, a.layers, a.out_dim, a.seed = (
            NODE_FEATURES, EDGE_FEATURES, width, layers, 3, seed)
        self.arch = (width, layers)
        self.L = nat.lib()
        h = C.c_void_p()
        nat.check(self.L.ms_nn_create(C.byref(a), C.byref(h)))
        self.h = h
        self.n = int(self.L.ms_nn_num_params(h))
        self._graph = None

    def close(self):
        if getattr(self, "h", None):
            self.L.ms_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, code):
        if code != nat.MS_OK:
            raise nat.NativeError(code, (self.L.ms_nn_last_error(self.h) or b"").decode())

    def params(self) -> np.ndarray:
        p = np.empty(self.n, np.float32)
        self._ck(self.L.ms_nn_get_params(self.h, nat.fptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        self._ck(self.L.ms_nn_set_params(self.h, nat.fptr(p)))

    def save(self, path: str) -> None:
        np.savez_compressed(path, params=self.params(), width=self.arch[0], layers=self.arch[1])

    @classmethod
    def load(cls, path: str) -> "WarmStartNet":
        z = np.load(path)
        net = cls(width=int(z["width"]), layers=int(z["layers"]))
        net.set_params(z["params"])
        return net

    @staticmethod
    def _desc(b: dict):
        d = nat.NNGraphDesc()
        d.num_graphs = len(b["offsets"]) - 1
        d.num_nodes = int(b["offsets"][-1])
        d.num_edges = len(b["src"])
        d.node_offsets = nat.iptr(b["offsets"])
        d.edge_src = nat.iptr(b["src"])
        d.edge_dst = nat.iptr(b["dst"])
        d.node_feat = nat.fptr(b["node_feat"])
        d.edge_feat = nat.fptr(b["edge_feat"])
        return d

    def set_graph(self, b: dict) -> None:
        self._graph = b  # keep the host arrays alive for the call
        self._ck(self.L.ms_nn_set_graph(self.h, C.byref(self._desc(b))))

    def predict(self, b: dict) -> np.ndarray:
        """The network's output for the batched graph b without the training
        state (ms_nn_predict: fused kernels, no activation buffers); equal
        bit for bit to set_graph(b) + forward()."""
        out = np.empty((int(b["offsets"][-1]), 3), np.float32)
        self._ck(self.L.ms_nn_predict(self.h, C.byref(self._desc(b)), nat.fptr(out)))
        return out

    def forward(self) -> np.ndarray:
        n = int(self._graph["offsets"][-1])
        out = np.empty((n, 3), np.float32)
        self._ck(self.L.ms_nn_forward(self.h, nat.fptr(out)))
        return out

    def loss_backward(self, target: np.ndarray, want_grads: bool = False):
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        g = np.empty(self.n, np.float32) if want_grads else None
        self._ck(self.L.ms_nn_loss_backward(self.h, nat.fptr(t), C.byref(loss), nat.fptr(g)))
        return (loss.value, g) if want_grads else loss.value

    def train_step(self, target: np.ndarray, lr: float, wd: float = 0.0) -> float:
        self._ck(self.L.ms_nn_forward(self.h, None))
        loss = self.loss_backward(target)
        self._ck(self.L.ms_nn_adam_step(self.h, lr, 0.9, 0.999, 1e-8, wd))
        return loss

    def predict_dofs(self, sh, graph: dict | None = None) -> np.ndarray:
        g = graph or shell_graph(sh)
        if self.arch[0] in (32, 64):
            return positions_from_output(sh, g, self.predict(batch_graphs([g])))
        self.set_graph(batch_graphs([g]))
        return positions_from_output(sh, g, self.forward())


# ------------------------------------------------------------ data + train

FAMILIES = ("defect", "grf")


def training_family(num: int, seed: int, tmin: int = 1000, tmax: int = 6000, family: str = "defect"):
    """Small thick shells that relax in ~1e5 steps, h in [0.02, 0.05].

    ``defect``: discs with a +1 azimuthal LCE defect metric (final lambda
    0.85-0.95 modulated by a clipped GRF of amplitude 0.03, opto-thermal nu
    0.5) and bbar = kappa I, kappa in [0, 1]: the equilibrium is a cone-like
    shape fixed by the defect, unique up to the z -> -z reflection.
    ``grf``: discs/squares with an isotropic GRF lambda (amplitude 0.05-0.15)
    and bbar = kappa I, |kappa| <= 1.5: thin discs then bend cylindrically
    about an axis that only the noise picks, so the equilibrium is not a
    function of the programmed metric."""
    from .meshing import (defect_metric_abar_info, disc_mesh, grf, isotropic_abar_info, square_mesh,
                          uniform_bbar_info)
    from .shell import Settings, prepare_shell

    if family not in FAMILIES:
        raise ValueError(f"unknown family {family!r}")
    out = []
    for i in range(num):
        rng = np.random.default_rng([seed, i, 7 if family == "grf" else 11])
        T = math.exp(rng.uniform(math.log(tmin), math.log(tmax)))
        if family == "defect" or rng.uniform() < 0.5:
            m = disc_mesh(math.sqrt(7.26 / T), seed=seed * 1009 + i)
        else:
            m = square_mesh(math.sqrt(9.24 / T), seed=seed * 1009 + i)
        c = m.centroids()
        g = np.clip(grf(c[:, :2], 0.3, 64, seed * 7907 + i), -3, 3)
        if family == "defect":
            lam = rng.uniform(0.85, 0.95) * (1.0 + 0.03 * g)
            m.abar_info = defect_metric_abar_info(c, lam, 0.5)
            m.bbar_info = uniform_bbar_info(m.num_tris, rng.uniform(0.0, 1.0))
        else:
            m.abar_info = isotropic_abar_info(1.0 + rng.uniform(0.05, 0.15) * g)
            m.bbar_info = uniform_bbar_info(m.num_tris, rng.uniform(-1.5, 1.5))
        m.fill_defaults(h=rng.uniform(0.02, 0.05))
        out.append(prepare_shell(m, Settings(def_poisson_ratio=0.5)))
    return out


def relax_batch(shells, tol: float, max_steps: int, plans=None, dial_steps: int = 2000, precision: str = "fp64",
                poll: int = 5000, equil_interval: int = 100):
    """Relax a batch with the device-run protocol.  plans[i] = (start dofs,
    dial_start, steps per dial phase[, from_ansatz]) or None for the flat
    start with noise dialled from 0 over dial_steps.  With from_ansatz the
    shell is dialled from the start shape's own geometry (the reference's
    dialling_from_ansatz_rather_than_ref_state, do_dialling.cpp:194-228): its
    a and b are the targets at dial factor 0, so the start carries no elastic
    force, and the programmed ones at 1.  Returns (reports, end dofs, seconds)."""
    import copy
    import dataclasses

    from .shell import Context
    from .workloads import flat_noisy_state

    def ansatz(i):
        return plans is not None and plans[i] is not None and len(plans[i]) > 3 and plans[i][3]

    run = []
    for i, sh in enumerate(shells):
        if ansatz(i):
            sh = copy.copy(sh)
            sh.settings = dataclasses.replace(sh.settings, dialling_from_ansatz_rather_than_ref_state=True)
        run.append(sh)
    t0 = time.perf_counter()
    with Context(run, precision=precision) as ctx:
        starts = []
        for i, sh in enumerate(shells):
            plan = None if plans is None else plans[i]
            if plan is None:
                x, v = flat_noisy_state(sh, seed=i)
                f0, K = 0.0, dial_steps
            else:
                x, v = plan[0], None
                f0, K = plan[1], plan[2]
            ctx.set_state(i, x, v)
            if ansatz(i):
                a, b, _ = ctx.calc_a_comps_and_b_comps_and_normals(i)
                ctx.set_zero_dial_geometry(i, a.reshape(-1), b.reshape(-1))
            starts.append((f0, K))
        ctx.set_tuning(node_block=128, fused=1, tile_nodes=-2)  # small shells: small tiles, re-cut as they halt
        for i, (f0, K) in enumerate(starts):
            ctx.relax_setup(i, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=f0,
                            equil_interval=equil_interval, tol_f=tol, tol_v=tol)
        # in parts of at most 200 000 steps, with a progress line after each
        done, reps = 0, []
        while done < max_steps:
            n = min(200_000, max_steps - done)
            reps = ctx.relax_run(n, poll_every=poll)
            done += n
            running = sum(1 for r in reps if not r["halted"])
            log(f"relax: {done} steps, {len(shells) - running} of {len(shells)} shells halted, "
                f"{time.perf_counter() - t0:.1f}s")
            if running == 0:
                break
        ends = [ctx.get_state(i)[0] for i in range(len(shells))]
    return reps, ends, time.perf_counter() - t0


def full_dial_energies(shells, states) -> list:
    """Elastic energy E_s + E_b at dial factor 1 of each shell's state; states[i]
    may be one dofs vector or a list of them (a list of energies back)."""
    from .shell import Context

    out = []
    with Context(shells, precision="fp64") as ctx:
        for i, xs in enumerate(states):
            ctx.set_dial_clock(i, 1.0, 1.0, 1, 0, False)
            one = isinstance(xs, np.ndarray)
            es = []
            for x in ([xs] if one else xs):
                ctx.set_state(i, x)
                e = ctx.calc_energies(i)
                es.append(float(e[0] + e[1]))
            out.append(es[0] if one else es)
    return out


def generate(num: int, seed: int, batch: int = 128, dial_steps: int = 2000, tol: float = 1e-5,
             max_steps: int = 600000, family: str = "defect") -> list[tuple]:
    """Equilibria of the training family produced by the batched solver:
    (shell, end dofs, (family seed, family size, index)) per converged shell."""
    data = []
    for b0 in range(0, num, batch):
        nb = min(batch, num - b0)
        shells = training_family(nb, seed + b0, family=family)
        reps, ends, secs = relax_batch(shells, tol, max_steps, dial_steps=dial_steps)
        ok = [i for i, r in enumerate(reps) if r["status"] == 2]
        st = np.bincount([r["status"] for r in reps], minlength=3)
        fails = sum(r["failed"] for r in reps)
        fr = np.array([r["max_force_ratio"] for r in reps])
        log(f"batch {b0 // batch}: {len(ok)}/{len(shells)} equilibria in {secs:.1f}s (status counts {st.tolist()}, "
            f"{fails} blow-ups, force ratio median {np.median(fr):.2e} max {fr.max():.2e})")
        data += [(shells[i], ends[i], (seed + b0, nb, i)) for i in ok]
    return data


def save_equilibria(path: str, data: list[tuple], family: str) -> None:
    keys = np.array([d[2] for d in data], np.int64).reshape(-1, 3)
    sizes = np.array([len(d[1]) for d in data], np.int64)
    np.savez_compressed(path, keys=keys, split=np.cumsum(sizes)[:-1], family=family,
                        dofs=(np.concatenate([d[1] for d in data]) if data else np.zeros(0)).astype(np.float32))


def load_equilibria(path: str) -> list[tuple]:
    """Inverse of save_equilibria; the shells are rebuilt from their seeds."""
    z = np.load(path)
    family = str(z["family"]) if "family" in z else "grf"
    fams = {}
    out = []
    for (sd, nb, i), x in zip(z["keys"], np.split(z["dofs"], z["split"])):
        key = (int(sd), int(nb))
        if key not in fams:
            fams[key] = training_family(key[1], key[0], family=family)
        out.append((fams[key][int(i)], np.asarray(x, np.float64), (key[0], key[1], int(i))))
    return out


def train(net: WarmStartNet, data: list[tuple], epochs: int = 60, graphs_per_step: int = 8, lr: float = 2e-3,
          seed: int = 0) -> list[float]:
    graphs = [shell_graph(d[0]) for d in data]
    targets = [kabsch_target(d[0], g, d[1]) for d, g in zip(data, graphs)]
    zero = float(np.mean(np.concatenate([(t.astype(np.float64) ** 2).ravel() for t in targets])))
    log(f"{len(data)} equilibria; loss of a zero prediction {zero:.5f}")
    rng = np.random.default_rng(seed)
    hist = []
    steps = epochs * max(1, len(data) // graphs_per_step)
    k = 0
    for ep in range(epochs):
        order = rng.permutation(len(data))
        losses = []
        for j in range(0, len(order) - graphs_per_step + 1, graphs_per_step):
            idx = order[j:j + graphs_per_step]
            net.set_graph(batch_graphs([graphs[i] for i in idx]))
            lr_k = lr * 0.5 * (1.0 + math.cos(math.pi * k / max(1, steps)))  # cosine decay
            losses.append(net.train_step(np.concatenate([targets[i] for i in idx]), lr_k))
            k += 1
        hist.append(float(np.mean(losses)) if losses else float("nan"))
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"epoch {ep}: loss {hist[-1]:.5f}")
    return hist


# scales of the predicted displacement tried before the energy test: the
# network's errors cost stretch energy, so a shorter step along the
# prediction often beats both the flat start and the full prediction
ALPHAS = (1.0, 0.75, 0.5, 0.35, 0.25, 0.15)
# the network is trained up to z -> -z (its loss takes the better
# reflection), but a programmed bbar prefers one side: both are tried
SIGNS = (1.0, -1.0)


def warm_start_plans(net: WarmStartNet, shells, dial_steps: int = 2000, start: str = "ansatz") -> tuple[list, list]:
    """Start plans with the energy fallback.  For each shell the predicted
    displacement, in either z orientation, is scaled by the alpha of least
    elastic energy (at dial factor 1); the scaled prediction is used when its
    energy is below that of the flat start with noise, otherwise the flat
    start is dialled as usual.  start = "direct": the prediction starts at
    dial factor 1; "ansatz": it is dialled from its own geometry over
    dial_steps (no elastic force at the start).  Returns (plans, the chosen
    signed alpha or None)."""
    from .workloads import flat_noisy_state

    flats = [flat_noisy_state(sh, seed=i)[0] for i, sh in enumerate(shells)]
    cands, labels = [], [(s, a) for s in SIGNS for a in ALPHAS]
    for sh in shells:
        d = net.predict_dofs(sh) - sh.ref_pos
        N = sh.N
        one = []
        for s, a in labels:
            dd = a * d
            dd[2 * N:] *= s
            one.append(sh.ref_pos + dd)
        cands.append(one)
    e_pred = full_dial_energies(shells, cands)
    e_flat = full_dial_energies(shells, flats)
    plans, used = [], []
    for i in range(len(shells)):
        e = np.asarray(e_pred[i])
        k = int(np.nanargmin(e)) if np.isfinite(e).any() else -1
        ok = k >= 0 and e[k] < e_flat[i]
        if not ok:
            plans.append(None)
        elif start == "ansatz":
            plans.append((cands[i][k], 0.0, dial_steps, True))
        else:
            plans.append((cands[i][k], 1.0, 1))
        used.append(labels[k][0] * labels[k][1] if ok else None)
    return plans, used


def evaluate(net: WarmStartNet, num: int = 64, seed: int = 10_000, dial_steps: int = 2000, tol: float = 1e-5,
             max_steps: int = 600000, family: str = "defect", start: str = "ansatz", tmin: int = 1000,
             tmax: int = 6000) -> dict:
    """Steps to equilibrium from the learned warm start (started at dial
    factor 1, or dialled from its own geometry: start, see warm_start_plans;
    with the fallback) against the flat start with noise dialled as usual, on
    unseen seeds of the family with triangle counts log-uniform in [tmin,
    tmax]; final energies of both at dial factor 1."""
    shells = training_family(num, seed, tmin=tmin, tmax=tmax, family=family)
    log(f"evaluation: {num} shells of {tmin}-{tmax} triangles set up")
    t0 = time.perf_counter()
    plans, used = warm_start_plans(net, shells, dial_steps, start=start)
    t_pred = time.perf_counter() - t0
    log(f"predictions: {sum(u is not None for u in used)} of {num} kept, {t_pred:.1f}s")
    base, base_end, t_base = relax_batch(shells, tol, max_steps, dial_steps=dial_steps)
    warm, warm_end, t_warm = relax_batch(shells, tol, max_steps, plans=plans, dial_steps=dial_steps)
    e_base = full_dial_energies(shells, base_end)
    e_warm = full_dial_energies(shells, warm_end)
    both = [i for i in range(num) if base[i]["status"] == 2 and warm[i]["status"] == 2]
    pred = [i for i in both if used[i] is not None]
    sb = np.array([base[i]["eq_step"] + 1 for i in both], float)
    sw = np.array([warm[i]["eq_step"] + 1 for i in both], float)
    sbp = np.array([base[i]["eq_step"] + 1 for i in pred], float)
    swp = np.array([warm[i]["eq_step"] + 1 for i in pred], float)
    de = np.array([(e_warm[i] - e_base[i]) / max(abs(e_base[i]), 1e-300) for i in both])
    f = lambda a: float(a) if np.size(a) else None  # noqa: E731
    return {
        "shells": num, "prediction_used": sum(u is not None for u in used),
        "alphas_chosen": [u for u in used if u is not None], "both_converged": len(both),
        "flat_converged": sum(1 for r in base if r["status"] == 2),
        "warm_converged": sum(1 for r in warm if r["status"] == 2),
        "steps_flat_mean": f(sb.mean()) if len(sb) else None,
        "steps_warm_mean": f(sw.mean()) if len(sw) else None,
        "steps_ratio_median": f(np.median(sb / sw)) if len(sb) else None,
        "steps_ratio_median_predicted": f(np.median(sbp / swp)) if len(sbp) else None,
        "seconds_flat": t_base, "seconds_warm": t_warm, "seconds_predict": t_pred,
        "final_energy_rel_diff_median": f(np.median(de)) if len(de) else None,
        "final_energy_lower_or_equal_fraction": f(np.mean(de <= 1e-6)) if len(de) else None,
        "tolerance": tol,
        "start": start,
        "triangles": [int(tmin), int(tmax)],
    }


# Parameters trained on 5789 equilibria of the defect family (2048 shells
# from each of the family seeds 1, 200000 and 300000, as generate() makes them
# in batches of 256, 600 000 steps at most, tolerance 1e-5; 60 000 Adam steps
# of 8 graphs at lr 2e-3 with cosine decay; loss 0.251 -> 0.189), written by
# WarmStartNet.save.
PRETRAINED_DEFECT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_defect.npz")


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(description="generate equilibria, train the warm-start network, evaluate")
    ap.add_argument("--num-train", type=int, default=512)
    ap.add_argument("--num-eval", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=0, help="0: from --train-steps")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="warmstart_params.npz")
    ap.add_argument("--max-steps", type=int, default=600000)
    ap.add_argument("--data", default="", help="npz of generated equilibria: loaded if present, else written")
    ap.add_argument("--family", default="defect", choices=FAMILIES)
    ap.add_argument("--params", default="", help="trained parameters to evaluate (skips generation and training)")
    ap.add_argument("--start", default="ansatz", choices=("direct", "ansatz"),
                    help="warm start at dial factor 1, or dialled from the prediction's own geometry")
    ap.add_argument("--eval-tmin", type=int, default=1000, help="evaluation shells: fewest triangles")
    ap.add_argument("--eval-tmax", type=int, default=6000, help="evaluation shells: most triangles")
    args = ap.parse_args(argv)
    if args.params:
        net = WarmStartNet.load(PRETRAINED_DEFECT if args.params == "pretrained" else args.params)
        ev = evaluate(net, args.num_eval, max_steps=args.max_steps, family=args.family, start=args.start,
                      tmin=args.eval_tmin, tmax=args.eval_tmax)
        print(json.dumps({"params": args.params, "eval": ev}), flush=True)
        return 0
    t0 = time.time()
    if args.data and os.path.exists(args.data):
        data = load_equilibria(args.data)
        log(f"loaded {len(data)} equilibria from {args.data}")
    else:
        data = generate(args.num_train, seed=1, max_steps=args.max_steps, family=args.family)
        if args.data:
            save_equilibria(args.data, data, args.family)
    t_gen = time.time() - t0
    net = WarmStartNet(seed=0)
    t0 = time.time()
    epochs = args.epochs or max(1, math.ceil(args.train_steps * 8 / max(1, len(data))))
    hist = train(net, data, epochs=epochs, lr=args.lr)
    t_train = time.time() - t0
    net.save(args.out)
    ev = evaluate(net, args.num_eval, max_steps=args.max_steps, family=args.family, start=args.start,
                  tmin=args.eval_tmin, tmax=args.eval_tmax)
    print(json.dumps({"equilibria": len(data), "generate_s": t_gen, "train_s": t_train,
                      "loss_first": hist[0], "loss_last": hist[-1], "eval": ev}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
