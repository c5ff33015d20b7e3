"""Host builders (a1-a2), scenario generator and the C ABI surface (no GPU)."""
import ctypes
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import Case, oracle_config
from oracle import model as omodel

ROOT = Path(__file__).resolve().parent.parent


def test_product_static_matches_oracle_builder():
    from icedem.ice_state import build_static_data
    for name in ("hexagon", "c0", "c1small"):
        case = Case.make(name)
        sc = case.scene
        a = build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, case.cfg)
        b = omodel.build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids,
                                     oracle_config(case.cfg))
        for f in b._fields:
            x, y = getattr(a, f), getattr(b, f)
            if isinstance(y, float):
                assert x == y, f
            else:
                np.testing.assert_array_equal(np.asarray(x), np.asarray(y), err_msg=f)


def test_static_closed_forms():
    case = Case.make("hexagon")
    st = case.oracle(np.float64).st64
    A = 3 * math.sqrt(3) / 2 * 9.0
    np.testing.assert_allclose(st.areas, A, rtol=1e-12)
    np.testing.assert_allclose(st.radii, math.sqrt(A / math.pi), rtol=1e-12)
    np.testing.assert_allclose(st.masses, A * 1.0 * 920.0, rtol=1e-12)
    np.testing.assert_allclose(st.inertia, 0.5 * st.masses * st.radii ** 2, rtol=1e-12)
    E, nu = 5e9, 0.3
    assert abs(st.k_n - E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))) < 1e-3
    assert abs(st.k_s - st.k_n * (1 - 2 * nu) / (2 * (1 - nu))) < 1e-3
    np.testing.assert_allclose(st.rest_lengths, 10.0)
    w = min(st.radii[0], 4.0)
    np.testing.assert_allclose(st.bond_strengths, 0.3e6 * 1.0 * w * 1.0, rtol=1e-12)
    # sigma_bar floored at 1 MPa (F/(w h) = 0.3 MPa here)
    np.testing.assert_allclose(st.sigma_bar_bond, 1e6)
    np.testing.assert_array_equal(st.local_polys_padded[0], case.scene.polygons[0] - case.scene.centroids[0])


def test_hull_polygon_matches_reference_shape():
    from icedem.ice_state import ship_hull_polygon
    h = ship_hull_polygon(60.0, 12.0)
    assert h.shape == (62, 2)
    np.testing.assert_array_equal(h, omodel.hull_polygon(60.0, 12.0))
    assert h[0, 0] == 30.0 and abs(h[0, 1] - 6.0 * 0.1) < 1e-12     # taper at +x, 1-0.9 = 0.1
    assert ship_hull_polygon(60.0, 12.0, 5).shape == (12, 2)


def test_scenario_determinism_and_bonds():
    from icedem import scenario
    a = scenario.scene_c0(3); b = scenario.scene_c0(3)
    assert len(a.polygons) == len(b.polygons)
    for p, q in zip(a.polygons, b.polygons):
        np.testing.assert_array_equal(p, q)
    assert a.bonds == b.bonds
    # bonds: exactly the pairs of the reference criterion (bonds.py:6-16)
    c = a.centroids
    r = np.sqrt(np.array([scenario.polygon_area(p) for p in a.polygons]) / np.pi)
    ref = [(i, j) for i in range(len(c)) for j in range(i + 1, len(c))
           if np.linalg.norm(c[i] - c[j]) < (r[i] + r[j]) * 1.5]
    assert a.bonds == ref
    # tessellation conserves each floe's area
    for f in range(a.n_floes):
        cells = [p for p, k in zip(a.polygons, a.parent_ids) if k == f]
        assert 15 <= len(cells) <= 25


def test_header_symbols_exported_and_abi_layout(tmp_path):
    from icedem import _lib
    lib = _lib.load()
    hdr = (ROOT / "include" / "icedem.h").read_text()
    names = re.findall(r"^(?:const char\*|int)\s+(icedem_\w+)\s*\(", hdr, flags=re.M)
    assert len(names) >= 14
    sur = (ROOT / "include" / "icesur.h").read_text()
    sur_names = re.findall(r"^int\s+(icesur_\w+)\s*\(", sur, flags=re.M)
    assert len(sur_names) >= 11
    names += sur_names
    for nm in names:
        assert hasattr(lib, nm), nm
        assert nm in _lib.SIGNATURES, nm
    # struct layout: ctypes mirrors == C sizeof/offsetof
    src = tmp_path / "abi.c"
    structs = {"IceDemStatic": _lib.IceDemStatic, "IceDemParams": _lib.IceDemParams,
               "IceDemState": _lib.IceDemState, "IceDemInputs": _lib.IceDemInputs,
               "IceDemHistory": _lib.IceDemHistory, "IceDemDiag": _lib.IceDemDiag}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/icedem.h"', 'int main(void){']
    for s, cls in structs.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{s}.{fname} %zu\\n", offsetof({s}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for s, cls in structs.items():
        assert int(got[s]) == ctypes.sizeof(cls), s
        for fname, _ in cls._fields_:
            assert int(got[f"{s}.{fname}"]) == getattr(cls, fname).offset, f"{s}.{fname}"


def test_product_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from icedem.engine import DemEngine
    from icedem import _lib
    case = Case.make("hexagon")
    with pytest.raises(_lib.IceDemError):
        case.engine()
