"""THOR's file formats on the CPU: the .thr parser (config_file.cpp:71-131,
keys of esp.cu:252-475) and the HDF5 layout of the output files
(storage.h; esp_output.cu:136-486).

The planet file this build writes for the Earth Held-Suarez configuration
is compared, dataset by dataset, with the layout of the planet file THOR
ships (ifile/esp_initial_planet.h5, recorded by
tests/golden/make_planet_layout.py): the same dataset names, integer or
float type, sizes, "Variable" and "units" texts and values. phy_module is
the one exception: its "Variable" names the physics module a build links
("multi" in THOR's file, "none" here, where Held-Suarez is a built-in
forcing).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def io(built):
    from thor_amd import io
    io.lib()
    return io


class RunConfig(C.Structure):
    pass


def _run_config(built):
    """ctypes mirror of thor_run_config (include/thor_io.h)."""
    capi = built
    if not hasattr(RunConfig, "_fields_"):
        RunConfig._fields_ = [
            ("params", capi.Params), ("top_altitude", C.c_double), ("vlevel", C.c_int),
            ("spring_dynamics", C.c_int), ("spring_beta", C.c_double), ("num_steps", C.c_int), ("n_out", C.c_int),
            ("globdiag", C.c_int), ("rest", C.c_int), ("gcm_off", C.c_int), ("forcing", C.c_int),
            ("simulation_ID", C.c_char * 160), ("results_path", C.c_char * 512), ("initial", C.c_char * 512),
            ("GPU_ID_N", C.c_int), ("RayleighSponge", C.c_int), ("sponge", capi.SpongeParams),
            ("rt", capi.RTParams), ("Tstar", C.c_double), ("radius_star", C.c_double),
            ("planet_star_dist", C.c_double), ("boundary_layer", C.c_int), ("bl", capi.BLParams)]
    return RunConfig


def test_config_values_defaults_and_unknown_keys(io, built, tmp_path):
    path = os.path.join(GOLDEN, "earth_hs_g4.thr")
    c = io.Config(path)
    assert c.get("glevel") == "4" and c.get("HyDiff") == "true"  # trailing comment removed
    assert c.get("core_benchmark") == "HeldSuarez" and c.get("does_not_exist") is None
    assert c.unknown() == []
    rc = _run_config(built)()
    io._check(io.lib().thor_config_run(c._h, C.byref(rc)))
    p = rc.params
    assert (p.A, p.Omega, p.Gravit, p.Rd, p.Cp, p.P_Ref) == (6371000.0, 7.292e-5, 9.8, 287.0, 1005.0, 1e5)
    assert (p.glevel, rc.vlevel, rc.top_altitude, p.timestep, rc.num_steps, rc.n_out) == (4, 32, 36000.0, 1000, 4, 2)
    assert p.HyDiffOrder == 4 and p.n_substeps == 6 and p.precision == 0  # defaults (define.h, thor_driver.cu)
    assert p.thermo_equation == built.THERMO_ENTROPY
    assert rc.forcing == built.FORCING_HELD_SUAREZ and rc.globdiag == 1 and rc.rest == 1
    assert rc.simulation_ID == b"Earth"
    # unknown keys are reported, not fatal; malformed lines and unsupported values are
    bad = tmp_path / "x.thr"
    bad.write_text("glevel = 5\nnot_a_key = 3\n")
    assert io.Config(str(bad)).unknown() == ["not_a_key"]
    bad.write_text("glevel = 5\nthis line has no equals sign\n")
    with pytest.raises(built.ThorError, match="line 2"):
        io.Config(str(bad))
    # thermo_equation (esp.cu:696-710): entropy, energy or an error
    bad.write_text("thermo_equation = energy\n")
    ce = io.Config(str(bad))
    io._check(io.lib().thor_config_run(ce._h, C.byref(rc)))
    assert rc.params.thermo_equation == built.THERMO_ENERGY
    # boundary layer (boundary_layer.cu:339-362): RayleighHS with THOR's defaults
    bad.write_text("boundary_layer = true\nbl_type = RayleighHS\nsurf_drag = 2e-5\n")
    cl = io.Config(str(bad))
    assert cl.unknown() == []
    io._check(io.lib().thor_config_run(cl._h, C.byref(rc)))
    assert rc.boundary_layer == 1 and rc.bl.surf_drag == 2e-5 and rc.bl.bl_sigma == 0.7
    for text, word in (("boundary_layer = true\nbl_type = LocalMixL\n", "bl_type"),
                       ("thermo_equation = enthalpy\n", "thermo_equation"), ("core_benchmark = JetSteady\n",
                                                                             "JetSteady"),
                       ("vert_refined = true\n", "vert_refined"), ("glevel = four\n", "glevel")):
        bad.write_text(text)
        cb = io.Config(str(bad))
        with pytest.raises(built.ThorError, match=word):
            io._check(io.lib().thor_config_run(cb._h, C.byref(rc)))


def test_radiative_transfer_incident_flux(io, built, tmp_path):
    """radiative_transfer = true: the double-gray forcing with
    incflx = (R* / a)^2 sigma T*^4, R* in solar radii and a in au
    (radiative_transfer.cu:1431-1432)."""
    f = tmp_path / "rt.thr"
    f.write_text("core_benchmark = NoBenchmark\nradiative_transfer = true\nTstar = 4500\nradius_star = 0.7\n"
                 "planet_star_dist = 0.015\nP_ref = 1e7\n")
    rc = _run_config(built)()
    c = io.Config(str(f))
    io._check(io.lib().thor_config_run(c._h, C.byref(rc)))
    want = (0.7 * 695700000.0 / (0.015 * 149597870700.0)) ** 2 * 5.6703744191844314e-08 * 4500.0 ** 4
    assert rc.forcing == built.FORCING_DOUBLE_GRAY
    assert abs(rc.rt.incflx / want - 1) < 1e-14 and rc.rt.ps0 == 1e7


def test_planet_file_layout_matches_thor(io, built, tmp_path):
    c = io.Config(os.path.join(GOLDEN, "earth_hs_g4.thr"))
    rc = _run_config(built)()
    io._check(io.lib().thor_config_run(c._h, C.byref(rc)))
    L = io.lib()
    L.thor_output_planet.argtypes = [C.c_char_p, C.c_void_p]
    io._check(L.thor_output_planet(str(tmp_path).encode(), C.byref(rc)))
    mine = io.h5_layout(str(tmp_path / "esp_output_planet_Earth.h5"))
    with open(os.path.join(GOLDEN, "planet_layout_earth_g4.json")) as fh:
        ref = json.load(fh)
    assert set(mine) == set(ref), (set(mine) ^ set(ref))
    for k, r in ref.items():
        m = mine[k]
        assert (m["int"], m["size"], m["units"]) == (r["int"], r["size"], r["units"]), k
        if k != "phy_module":
            assert m["Variable"] == r["Variable"], k
        assert m["value"] == r["value"], k


def test_h5_round_trip_and_attributes(io, built, tmp_path):
    path = str(tmp_path / "t.h5").encode()
    a = np.random.default_rng(1).standard_normal(1000)
    L = io.lib()
    dp = C.POINTER(C.c_double)
    io._check(L.thor_h5_write(path, 1, b"/Rho", a.ctypes.data_as(dp), a.size, b"kg/m^3", b"Density"))
    io._check(L.thor_h5_write(path, 0, b"Pressure", a.ctypes.data_as(dp), 10, b"Pa", b"Pressure"))
    assert io.h5_names(path.decode()) == ["Pressure", "Rho"]
    assert np.array_equal(io.h5_read(path.decode(), "Rho"), a)
    assert np.array_equal(io.h5_read(path.decode(), "/Pressure"), a[:10])
    assert io.h5_attr(path.decode(), "Rho", "units") == "kg/m^3"
    assert io.h5_attr(path.decode(), "Rho", "Variable") == "Density"
    assert not io.h5_is_int(path.decode(), "Rho")
    assert L.thor_h5_size(path, b"/missing") == -1
