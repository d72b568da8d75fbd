"""The allocation-free geometry entry point (ms_calc_geometry_noalloc)
against ms_calc_geometry, through the C-ABI.

Two shells of one context are evaluated in a row, the larger first and then
the smaller, so the second call's outputs come from a scratch array that
still holds the first call's larger result; each must equal
ms_calc_geometry's a, b and n bit for bit (the same kernel arithmetic).
A shell above MS_NOALLOC_MAX_TRIS is refused with MS_ERR_ARG and the
context keeps working.
"""
import numpy as np
import pytest

from conftest import deformed_state, grf_disc

pytestmark = pytest.mark.gpu

MAX_TRIS = 524288  # MS_NOALLOC_MAX_TRIS


def _noalloc(ctx, s, T):
    from morphoshell_amd import native as nat

    a, b, n = np.full(3 * T, np.nan), np.full(3 * T, np.nan), np.full(3 * T, np.nan)
    code = nat.lib().ms_calc_geometry_noalloc(ctx.h, s, nat.dptr(a), nat.dptr(b), nat.dptr(n))
    return code, a.reshape(3, -1), b.reshape(3, -1), n.reshape(3, -1)


def test_noalloc_geometry_equals_geometry_on_two_calls_in_a_row():
    from morphoshell_amd import native as nat
    from morphoshell_amd.shell import Context, Settings, prepare_shell

    big = prepare_shell(grf_disc(0.012, seed=41, h=0.01, kappa=0.3), Settings())
    small = prepare_shell(grf_disc(0.05, seed=42, h=0.02, kappa=-0.5), Settings())
    assert big.T > 4 * small.T
    with Context([big, small]) as ctx:
        for i, sh in enumerate((big, small)):
            ctx.set_state(i, *deformed_state(sh, seed=43 + i))
        ctx.step(7)
        for s, sh in ((0, big), (1, small)):  # the second call has the smaller input
            code, a, b, n = _noalloc(ctx, s, sh.T)
            assert code == nat.MS_OK, nat.lib().ms_ctx_last_error(ctx.h)
            ar, br, nr = ctx.calc_a_comps_and_b_comps_and_normals(s)
            assert np.array_equal(a, ar) and np.array_equal(b, br) and np.array_equal(n, nr), s
        # only the requested outputs are written
        bb = np.full(3 * small.T, np.nan)
        assert nat.lib().ms_calc_geometry_noalloc(ctx.h, 1, None, nat.dptr(bb), None) == nat.MS_OK
        assert np.array_equal(bb.reshape(3, -1), ctx.calc_a_comps_and_b_comps_and_normals(1)[1])


def test_noalloc_geometry_refuses_a_shell_above_the_limit():
    from morphoshell_amd import native as nat
    from morphoshell_amd.shell import Context, Settings, prepare_shell

    sh = prepare_shell(grf_disc(0.0035, seed=44, h=0.01), Settings())
    assert sh.T > MAX_TRIS
    with Context(sh) as ctx:
        ctx.set_state(0, *deformed_state(sh, seed=45))
        code, a, _, _ = _noalloc(ctx, 0, sh.T)
        assert code == nat.MS_ERR_ARG
        assert b"limit" in nat.lib().ms_ctx_last_error(ctx.h)
        assert np.isnan(a).all()
        ar, _, _ = ctx.calc_a_comps_and_b_comps_and_normals(0)  # the context keeps working
        assert np.isfinite(ar).all()
