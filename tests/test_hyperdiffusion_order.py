"""Higher-order hyperdiffusion (HyDiffOrder >= 6) in the oracle (CPU).

THOR runs HyDiffOrder / 2 - 1 Diffusion_Op passes before the last one
(thor_driver.cu:435-470); a pass after the first takes the previous pass's
output unweighted (thor_diff.h:218-221). The oracle's intermediate pass
(diffusion, stepnum 1, not last) must therefore be the first pass's
Laplacian, which is pinned bit for bit to THOR's kernel, applied to each of
the six fields of its input. The first pass applies the Laplacian to the
density unweighted (variable 0), so feeding field v of a random diff array in
as the density must give, bit for bit, variable v of the intermediate pass.
A dropped or extra density weighting, a wrong variable slot or a transposed
layout fails this. The chain of passes is checked by its sign: at orders 4
and 6 the density tendency of a random density perturbation opposes the
perturbation (THOR flips the sign of the coefficient with (-1)^(order/2) so
that every even order damps).
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid


@pytest.fixture(scope="module")
def setup(built):
    from oracle.oracle import Oracle
    capi = built
    g = grid(capi, 4, 8)
    p = earth_params(capi, 4, dt=600.0, HyDiffOrder=6)
    st = capi.ic_rest(g, p, seed=1021, temperature_noise=1.0, wind_noise=2.0)
    o = Oracle(g, p)
    o.set_state(st)
    return capi, g, p, o


def test_intermediate_pass_is_first_pass_laplacian(setup):
    capi, g, p, o = setup
    P, nv = g.P, g.nv
    rng = np.random.default_rng(5)
    D = rng.standard_normal((P, nv, 6)) * np.array([1e-3, 2.0, 2.0, 2.0, 0.1, 300.0])
    o.set("diff", D.reshape(-1))
    o.kernel("diffusion", 1, 0)
    mid = o.get("diff").reshape(P, nv, 6)
    assert np.abs(mid).max() > 0
    rho0 = o.get("Rhok").copy()
    for v in range(6):
        o.set("Rhok", np.ascontiguousarray(D[:, :, v]).reshape(-1))
        o.set("diff", np.zeros(P * nv * 6))
        o.kernel("diffusion", 0, 0)
        first = o.get("diff").reshape(P, nv, 6)[:, :, 0]
        np.testing.assert_array_equal(mid[:, :, v], first, err_msg=f"variable {v}")
    o.set("Rhok", rho0)


def test_order_six_damps_the_grid_scale_harder_than_order_four(built):
    """One intermediate pass is one more Laplacian: on a random density
    perturbation (dominated by grid-scale modes) the order-6 tendency is the
    order-4 one times about the Laplacian's grid-scale eigenvalue times
    dbar^2 (of order one), and both oppose the perturbation."""
    from oracle.oracle import Oracle
    capi = built
    g = grid(capi, 4, 8)
    out = {}
    for order in (4, 6):
        p = earth_params(capi, 4, dt=600.0, HyDiffOrder=order)
        st = capi.ic_rest(g, p, seed=1022, temperature_noise=1.0, wind_noise=0.0)
        o = Oracle(g, p)
        o.set_state(st)
        rng = np.random.default_rng(9)
        rho = o.get("Rho")
        o.set("Rhok", rho * (1.0 + 1e-3 * rng.standard_normal(rho.size)))
        o.set("diff", np.zeros(g.P * g.nv * 6))
        stepnum = 0
        for _ in range(order // 2 - 1):
            o.kernel("diffusion", stepnum, 0)
            stepnum += 1
        o.kernel("diffusion", stepnum, 1)
        out[order] = o.get("diffrh").reshape(g.P, g.nv)
    drho = o.get("Rhok").reshape(g.P, g.nv) - o.get("Rho").reshape(g.P, g.nv)
    for order in (4, 6):
        # damping: the tendency opposes the perturbation
        corr = float((out[order] * drho).sum())
        print(f"order {order}: <diffrh, drho> = {corr:.3e}")
        assert corr < 0, order
    ratio = np.abs(out[6]).max() / np.abs(out[4]).max()
    print(f"max |diffrh| order 6 / order 4 = {ratio:.3f}")
    assert 0.05 < ratio < 20
