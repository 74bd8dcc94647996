"""The oracle's vertical solve (THOR's Vertical_Eq, thor_vertical_int.h:52-563)
pinned by the residual of its own tridiagonal systems (CPU).

The reference checks its solve the same way (thor_vertical_int.h:488-525).
After one full step from a perturbed rest state, the oracle forms, per column,
the system a_l x_{l-1} + b_l x_l + c_l x_{l+1} = d_l for the vertical-momentum
deviation at interfaces 1..nv-1 (x_0 = x_nv = 0) and solves it by the modified
Thomas algorithm. Checked: the residual of the solution is at round-off, and
the solution equals a dense LU solve (numpy) of the same system. A mistake in
the elimination or the back substitution (an index shift, a sign, a missing
pivot) fails both; the coefficient formulas themselves are pinned only by the
invariants of the full step (rest preserved, mass conserved).
"""
from __future__ import annotations

import numpy as np
import pytest

from common import earth_params, grid


@pytest.mark.parametrize("deep,nonhydro", [(1, 1), (0, 0)])
def test_vertical_solve_residual(built, deep, nonhydro):
    from oracle.oracle import Oracle
    capi = built
    g = grid(capi, 4, 20)
    p = earth_params(capi, 4, dt=600.0, ns=6, DeepModel=deep, NonHydro=nonhydro)
    st = capi.ic_rest(g, p, seed=1015, temperature_noise=1.0, wind_noise=2.0)
    o = Oracle(g, p)
    o.set_state(st)
    o.steps(1)  # realistic sources, deviations and interface fields
    o.kernel("vertical_eq", int(round(1e3 * p.timestep / p.n_substeps)), 1)
    P, nv = g.P, g.nv
    x = o.get("Whs").reshape(P, nv + 1)
    a, b, c, d = (o.get(n).reshape(P, nv + 1)[:, 1:nv] for n in ("tri_a", "tri_b", "tri_c", "tri_d"))
    assert np.all(x[:, 0] == 0) and np.all(x[:, nv] == 0)
    xm, x0, xp = x[:, 0:nv - 1], x[:, 1:nv], x[:, 2:nv + 1]
    r = a * xm + b * x0 + c * xp - d
    scale = np.abs(a * xm) + np.abs(b * x0) + np.abs(c * xp) + np.abs(d)
    res = float((np.abs(r) / np.where(scale > 0, scale, 1.0)).max())
    # dense LU solve of the same systems on a sample of columns
    worst = 0.0
    for col in np.linspace(0, P - 1, 64).astype(int):
        n = nv - 1
        A = np.diag(b[col]) + np.diag(a[col, 1:], -1) + np.diag(c[col, :-1], 1)
        ref = np.linalg.solve(A, d[col])
        worst = max(worst, float(np.abs(ref - x0[col]).max() / max(np.abs(ref).max(), 1e-300)))
        assert n == len(ref)
    print(f"DeepModel {deep} NonHydro {nonhydro}: max scaled residual {res:.2e}, "
          f"max rel diff vs dense solve {worst:.2e}, max |Whs| {np.abs(x).max():.3e}")
    assert np.abs(x).max() > 0
    assert res < 1e-13
    assert worst < 1e-10
