"""CPU pins of the warm-start network's reference and of the host-side
graph/target construction."""
import numpy as np
import pytest

from oracle import mpnn_ref


def _tiny_graph(rng, N=9):
    # a triangulated 3x3 patch, both edge directions, sorted by (dst, src)
    tris = [(0, 1, 4), (0, 4, 3), (1, 2, 5), (1, 5, 4), (3, 4, 7), (3, 7, 6), (4, 5, 8), (4, 8, 7)]
    pairs = set()
    for a, b, c in tris:
        for u, v in ((a, b), (b, c), (c, a)):
            pairs.add((u, v))
            pairs.add((v, u))
    pairs = sorted(pairs, key=lambda p: (p[1], p[0]))
    src = np.array([p[0] for p in pairs])
    dst = np.array([p[1] for p in pairs])
    return src, dst


def test_reference_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    Fn, Fe, W, L = 4, 3, 8, 2
    net = mpnn_ref.Net(Fn, Fe, W, L, bf16_ops=False)
    src, dst = _tiny_graph(rng)
    N, E = 9, len(src)
    p = rng.normal(size=net.lay["_n"]) * 0.4
    xn = rng.normal(size=(N, Fn))
    xe = rng.normal(size=(E, Fe))
    tgt = rng.normal(size=(N, 3))
    goff = np.array([0, 5, 9])

    def loss(pp):
        return net.loss(net.forward(pp, xn, xe, src, dst), tgt, goff)[0]

    c = {}
    out = net.forward(p, xn, xe, src, dst, c)
    _, dout = net.loss(out, tgt, goff)
    g = net.backward(p, xn, xe, src, dst, c, dout)
    for j in rng.choice(len(p), 60, replace=False):
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[j] += e
        pm[j] -= e
        fd = (loss(pp) - loss(pm)) / (2 * e)
        assert abs(fd - g[j]) <= 1e-7 * max(1.0, np.abs(g).max()), (j, fd, g[j])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, np.float32(1.0 + 2 ** -9)], np.float32)
    r = mpnn_ref.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # tie handled by evenness below
    assert r[2] == 1.0 + 2 ** -7               # above the tie rounds up
    assert r[3] == -2.5
    assert r[5] == 1.0                          # exact tie rounds to even (mantissa ...0)


def test_graph_and_targets():
    from morphoshell_amd.meshing import disc_mesh
    from morphoshell_amd.shell import Settings, prepare_shell
    from morphoshell_amd.warmstart import EDGE_FEATURES, NODE_FEATURES, kabsch_target, shell_graph

    m = disc_mesh(0.12, seed=1).fill_defaults(h=0.03)
    sh = prepare_shell(m, Settings())
    g = shell_graph(sh)
    assert g["node_feat"].shape == (sh.N, NODE_FEATURES) and g["edge_feat"].shape[1] == EDGE_FEATURES
    assert np.all(np.diff(g["dst"]) >= 0)
    assert len(g["src"]) == 2 * sh.info["num_edges"]
    assert np.isfinite(g["node_feat"]).all() and np.isfinite(g["edge_feat"]).all()
    # a rigid motion of an equilibrium does not change its target
    rng = np.random.default_rng(2)
    P = sh.mesh.points.copy()
    P[:, 2] = 0.05 * np.sin(3 * P[:, 0])
    t0 = kabsch_target(sh, g, np.ascontiguousarray(P.T.reshape(-1)))
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Q = P @ Rz.T + np.array([0.3, -1.2, 2.0])
    t1 = kabsch_target(sh, g, np.ascontiguousarray(Q.T.reshape(-1)))
    assert np.abs(t0 - t1).max() < 1e-5
