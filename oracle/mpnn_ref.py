"""CPU reference of the warm-start message-passing network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker of the GPU network's forward pass and
gradients.  Same architecture and flat parameter layout as the device model
(``morphoshell_amd/csrc/nn_model.cpp``); with ``bf16=True`` every GEMM
operand is rounded to bfloat16 (round-to-nearest-even) as the matrix cores
see it, and accumulation is FP64.  The backward pass is written by hand and
pinned against central finite differences in tests/test_warmstart_cpu.py.
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    out[nan] = np.nan
    return out


def layout(Fn: int, Fe: int, W: int, L: int, out_dim: int = 3) -> dict:
    """Offsets (w, b, in, out) of every linear map in the flat parameter vector."""
    off = 0
    lay = {}

    def lin(name, i, o):
        nonlocal off
        lay[name] = (off, off + i * o, i, o)
        off += i * o + o

    lin("enc_n", Fn, W)
    lin("enc_e", Fe, W)
    for l in range(L):
        lin(f"e1_{l}", 3 * W, W)
        lin(f"e2_{l}", W, W)
        lin(f"n1_{l}", 2 * W, W)
        lin(f"n2_{l}", W, W)
    lin("dec1", W, W)
    lin("dec2", W, out_dim)
    lay["_n"] = off
    return lay


def _wb(p, lay, name):
    w, b, i, o = lay[name]
    return p[w:w + i * o].reshape(i, o), p[b:b + o]


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def _silu(x):
    return x * _sig(x)


def _silu_d(x):
    s = _sig(x)
    return s + x * s * (1.0 - s)


class Net:
    def __init__(self, Fn, Fe, W, L, bf16_ops=True):
        self.Fn, self.Fe, self.W, self.L = Fn, Fe, W, L
        self.lay = layout(Fn, Fe, W, L)
        self.q = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))

    def _lin(self, p, name, X):
        Wm, b = _wb(p, self.lay, name)
        return self.q(X) @ self.q(Wm) + b

    def forward(self, p, xn, xe, src, dst, cache=None):
        W, L = self.W, self.L
        N = xn.shape[0]
        c = {} if cache is None else cache
        c["pre_n"] = self._lin(p, "enc_n", xn)
        h = _silu(c["pre_n"])
        c["pre_e"] = self._lin(p, "enc_e", xe)
        e = _silu(c["pre_e"])
        c["h"], c["e"] = [h], [e]
        for l in range(L):
            cat_e = np.concatenate([h[dst], h[src], e], axis=1)
            z1 = self._lin(p, f"e1_{l}", cat_e)
            a1 = _silu(z1)
            m = self._lin(p, f"e2_{l}", a1)
            agg = np.zeros((N, W))
            np.add.at(agg, dst, m)
            cat_n = np.concatenate([h, agg], axis=1)
            y1 = self._lin(p, f"n1_{l}", cat_n)
            s1 = _silu(y1)
            h = h + self._lin(p, f"n2_{l}", s1)
            e = e + m
            c.setdefault("layers", []).append(dict(cat_e=cat_e, z1=z1, a1=a1, m=m, cat_n=cat_n, y1=y1, s1=s1))
            c["h"].append(h)
            c["e"].append(e)
        c["d1pre"] = self._lin(p, "dec1", h)
        c["d1"] = _silu(c["d1pre"])
        out = self._lin(p, "dec2", c["d1"])
        c["out"] = out
        return out

    @staticmethod
    def loss(out, tgt, goff):
        """sum_g min(err(t), err(t with z -> -z)) / (3 N); returns (loss, dloss/dout)."""
        N = out.shape[0]
        scale = 1.0 / (3.0 * N)
        d = np.zeros_like(out)
        total = 0.0
        for g in range(len(goff) - 1):
            a, b = goff[g], goff[g + 1]
            tp = tgt[a:b]
            tm = tp * np.array([1.0, 1.0, -1.0])
            ep = ((out[a:b] - tp) ** 2).sum()
            em = ((out[a:b] - tm) ** 2).sum()
            t = tm if em < ep else tp
            total += min(ep, em)
            d[a:b] = 2.0 * scale * (out[a:b] - t)
        return total * scale, d

    def backward(self, p, xn, xe, src, dst, c, dout):
        """Gradient of the loss w.r.t. every parameter (same flat layout)."""
        W, L, lay, q = self.W, self.L, self.lay, self.q
        g = np.zeros_like(p)
        N, E = xn.shape[0], xe.shape[0]

        def lin_bwd(name, X, dY, want_dx=True):
            w, b, i, o = lay[name]
            Wm = p[w:w + i * o].reshape(i, o)
            g[w:w + i * o] += (q(X).T @ q(dY)).reshape(-1)
            g[b:b + o] += dY.sum(0)
            return q(dY) @ q(Wm).T if want_dx else None

        dd1 = lin_bwd("dec2", c["d1"], dout)
        dd1pre = dd1 * _silu_d(c["d1pre"])
        dH = lin_bwd("dec1", c["h"][L], dd1pre)
        dE = np.zeros((E, W))
        for l in range(L - 1, -1, -1):
            C = c["layers"][l]
            ds1 = lin_bwd(f"n2_{l}", C["s1"], dH)
            dy1 = ds1 * _silu_d(C["y1"])
            dcat_n = lin_bwd(f"n1_{l}", C["cat_n"], dy1)
            dH = dH + dcat_n[:, :W]
            dagg = dcat_n[:, W:]
            dm = dagg[dst] + dE
            da1 = lin_bwd(f"e2_{l}", C["a1"], dm)
            dz1 = da1 * _silu_d(C["z1"])
            dcat_e = lin_bwd(f"e1_{l}", C["cat_e"], dz1)
            np.add.at(dH, dst, dcat_e[:, :W])
            np.add.at(dH, src, dcat_e[:, W:2 * W])
            dE = dE + dcat_e[:, 2 * W:]
        dpre_n = dH * _silu_d(c["pre_n"])
        lin_bwd("enc_n", xn, dpre_n, want_dx=False)
        dpre_e = dE * _silu_d(c["pre_e"])
        lin_bwd("enc_e", xe, dpre_e, want_dx=False)
        return g
