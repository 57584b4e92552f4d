"""CPU checks of the C ABI: the library loads without a GPU, exports every
symbol include/fslbm.h declares, and its host-side builders and error
paths behave (no compute call needs a device here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from flowcontrast_amd import _lib, inputs

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(fslbm_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_binding():
    decl = declared_functions()
    assert decl, "no declarations parsed"
    assert sorted(_lib.EXPORTS) == decl


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_errors():
    L = _lib.lib()
    assert L.fslbm_abi_version() == 1
    assert L.fslbm_error_string(-1) == b"invalid argument"


def test_circle_mask_counts():
    m, nf = inputs.circle_mask(32, 32, 4, 14.0)
    assert m.shape == (4, 32, 32)
    assert nf == int((m[0] == 0).sum())
    # centred at 15.5: fluid never touches the faces, frame of 2 cells
    fl = np.argwhere(m[0] == 0)
    assert fl.min() == 2 and fl.max() == 29
    # closed-form count of lattice points in the disc centred at (15.5, 15.5)
    xs = np.arange(32) - 15.5
    assert nf == int(((xs[None] ** 2 + xs[:, None] ** 2) <= 196).sum())


def test_poiseuille_state_is_seeded_and_masked():
    m, _ = inputs.circle_mask(20, 20, 6, 8.0)
    a = inputs.poiseuille_state(m, 8.0, 0.05, 1e-3, 0.01, 7)
    b = inputs.poiseuille_state(m, 8.0, 0.05, 1e-3, 0.01, 7)
    c = inputs.poiseuille_state(m, 8.0, 0.05, 1e-3, 0.01, 8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    wall = m != 0
    for f in a:
        assert (f[wall] == 0).all()
    # ramp p = pin (nz - z)/nz on fluid cells, noise amplitude 1% of 2*ubar
    assert np.allclose(a[0][:, 10, 10], 1e-3 * (6 - np.arange(6)) / 6)
    assert np.abs(a[1]).max() <= 0.01 * 0.1
    assert a[3][0, 10, 10] < -0.09  # stored uz = -physical


def test_taylor_green_state():
    p, u, v, w = inputs.taylor_green_state(16, 0.05, 1.0)
    assert np.abs(w).max() == 0
    # divergence-free in the continuum; spectral check of the mode content
    assert abs(u.sum()) < 1e-12 and abs(v.sum()) < 1e-12
    assert abs(p.mean() - (1.0 + 3 * 0.05 ** 2 / 16 * 0.0)) < 1e-12
    assert abs(np.abs(u).max() - 0.05) < 1e-3


def test_create_rejects_bad_mask_without_gpu():
    # validation runs before any device query
    L = _lib.lib()
    prm = _lib.Params(1.0, 0.0, 0.01, 0.0, 2, 1)
    h = C.c_void_p()
    bad = np.zeros(8 * 8 * 8, dtype=np.uint8)  # fluid on the faces
    rc = L.fslbm_create(C.byref(prm), bad.ctypes.data_as(C.c_void_p), 8, 8, 8, 1, C.byref(h))
    assert rc == -1
    assert b"wall frame" in L.fslbm_last_error(None)
    odd = np.full(8 * 8 * 8, 7, dtype=np.uint8)
    rc = L.fslbm_create(C.byref(prm), odd.ctypes.data_as(C.c_void_p), 8, 8, 8, 1, C.byref(h))
    assert rc == -1 and b"mask codes" in L.fslbm_last_error(None)
    prm.d1 = 0
    rc = L.fslbm_create(C.byref(prm), None, 8, 8, 8, 1, C.byref(h))
    assert rc == -1 and b"d1" in L.fslbm_last_error(None)


def test_laminar_and_blasius_pin():
    # laminar: ubar = pin R^2 / (24 nu L)
    pin = inputs.laminar_pin(0.01, 64, 0.0714, 14.0)
    assert abs(pin * 196 / (24 * 0.01 * 64) - 0.0714) < 1e-12
    assert inputs.blasius_pin(0.01, 1024, 5300, 120) > 0


def test_server_create_validates_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.fslbm_server_create(3, 1 << 30, 100, C.byref(h)) == -1
    assert b"d1" in L.fslbm_server_last_error(None)
    assert L.fslbm_server_create(2, 1 << 30, 1, C.byref(h)) == -1
    assert b"quantum" in L.fslbm_server_last_error(None)
    assert L.fslbm_server_create(2, 0, 100, C.byref(h)) == -1
    # a valid configuration needs a device: none here
    rc = L.fslbm_server_create(2, 1 << 30, 100, C.byref(h))
    assert rc in (0, -4)
    if rc == 0:
        L.fslbm_server_destroy(h)
