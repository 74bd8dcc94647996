"""The C-ABI library loads without a GPU and exports every symbol that
include/*.h declares; host-only entry points work on the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.common import ROOT


def _declared(headers):
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in headers:
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(thor_\w+)\s*\(", src))
    return names


def test_exports_every_declared_symbol(built):
    """thor_dyn.h and thor_mlp.h: both builds of libthor_amd; thor_io.h: libthor_io."""
    headers = sorted(fn for fn in os.listdir(os.path.join(ROOT, "include")) if fn.endswith(".h"))
    assert headers == ["thor_dyn.h", "thor_io.h", "thor_mlp.h"], headers
    names = _declared(["thor_dyn.h", "thor_mlp.h"])
    assert len(names) > 15
    for variant in ("", "pm"):
        L = built.lib(variant)
        missing = [n for n in names if not hasattr(L, n)]
        assert not missing, f"{variant or 'default'} build lacks {missing}"
    from thor_amd import io
    names = _declared(["thor_io.h"])
    missing = [n for n in names if not hasattr(io.lib(), n)]
    assert len(names) > 10 and not missing, f"libthor_io lacks {missing}"


def test_no_device_is_reported_not_crashed(built):
    g = built.Grid(4, 4, 6.371e6, 36000.0)
    if os.environ.get("HIP_VISIBLE_DEVICES", "") == "" and not os.path.exists("/dev/kfd"):
        try:
            built.Dyn(g, built.default_params())
        except built.ThorError as e:
            assert e.code == built.THOR_ERR_NO_DEVICE


def test_bad_arguments_are_rejected(built):
    try:
        built.Grid(3, 4, 6.371e6, 36000.0)
    except built.ThorError as e:
        assert e.code == built.THOR_ERR_BAD_ARG
    else:
        raise AssertionError("glevel 3 accepted")
    msg = built.lib().thor_last_error().decode()
    assert "glevel" in msg and "4..8" in msg, msg


@pytest.mark.parametrize("order", [2, 5])
def test_unsupported_hyperdiffusion_order_is_rejected(built, order):
    """Orders below 4 or odd orders are refused before any device call."""
    from tests.common import earth_params, grid
    g = grid(built, 4, 8)
    p = earth_params(built, 4, HyDiffOrder=order)
    with pytest.raises(built.ThorError) as e:
        built.Dyn(g, p)
    assert e.value.code == built.THOR_ERR_UNSUPPORTED
    assert "HyDiffOrder" in str(e.value)


@pytest.mark.parametrize("thermo,precision,code", [(1, 1, "THOR_ERR_UNSUPPORTED"), (2, 0, "THOR_ERR_BAD_ARG")])
def test_thermo_equation_checks(built, thermo, precision, code):
    """The energy equation exists in the FP64 form only; thermo_equation takes
    0 (entropy) or 1 (energy). Both are refused before any device call."""
    from tests.common import earth_params, grid
    g = grid(built, 4, 8)
    p = earth_params(built, 4, thermo_equation=thermo, precision=precision)
    with pytest.raises(built.ThorError) as e:
        built.Dyn(g, p)
    assert e.value.code == getattr(built, code)
    assert ("energy" if thermo == 1 else "thermo_equation") in str(e.value)
