"""GPU parity on general geometry: masks that vary along z.

The reference takes an arbitrary per-cell mask IO (FSLBM.cu:137-138,152,164:
every neighbour of the 27-point gather is tested for `IO < 125`), and its
paper validates rough-walled pipes (MethodDoc_withFSLBMpaper.pdf p.6).  The
folded layout keeps that: any 0/125 mask with a wall frame on the x/y faces.

Cases (all z periodic with the +-pin jump):
  rough pipe   radius varying with angle and z plus seeded wall bumps
  orifice      a straight pipe with a thin orifice plate (a constriction of
               a few planes), so neighbouring z planes differ sharply
  sphere       a sphere obstacle in a square duct

Kernels: FP64 runs the plain kernel (one thread per cell, the mask byte of
every neighbour); FP32 and the 16-bit storage modes run the pipelined tiled
kernel's variant for such masks, which reads a per-cell wall code (the
in-plane wall weights of the cell's 3 x 3 neighbourhood and its own wall
bit) and combines the codes of three planes into the wall rule's weights.

Tolerances (as for the straight pipe):
  FP64   max |gpu - oracle| / max |oracle| <= 1e-12 per field
  FP32   relative L2 <= 1e-5 against the FP32 oracle, p and velocity
  16-bit relative L2 of the velocity against the FP64 oracle after 7 steps
         <= 3 ulp sqrt(7/3) (one storage ulp per stochastically rounded
         store, independent over 7 stores); the pressure's rms error within
         the same fraction of the velocity rms (it is stored as a deviation
         from a mean line, scaled like the velocity)
"""
import numpy as np
import pytest

from flowcontrast_amd import KERNEL_AUTO, KERNEL_PLAIN, KERNEL_TILED, RunParams, Solver
from flowcontrast_amd._lib import FslbmError

pytestmark = pytest.mark.gpu


def rough_pipe(nx=36, ny=34, nz=40, R=13.0, amp=2.0, nbumps=40, seed=11):
    """Pipe whose wall radius varies with angle and z, plus random bumps."""
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float64)
    cx, cy = (nx - 1) / 2, (ny - 1) / 2
    r = np.hypot(x - cx, y - cy)
    th = np.arctan2(y - cy, x - cx)
    rw = R + amp * np.sin(3 * th + 2 * np.pi * z / nz) * np.cos(2 * np.pi * 2 * z / nz)
    wall = r > rw
    for _ in range(nbumps):  # hemispherical bumps on the wall
        t = rng.uniform(0, 2 * np.pi)
        zc = rng.uniform(0, nz)
        rb = rng.uniform(1.5, 3.0)
        bx, by = cx + R * np.cos(t), cy + R * np.sin(t)
        dz = np.minimum(np.abs(z - zc), nz - np.abs(z - zc))  # periodic in z
        wall |= (x - bx) ** 2 + (y - by) ** 2 + dz ** 2 <= rb * rb
    wall[:, :, :2] = wall[:, :, -2:] = True
    wall[:, :2, :] = wall[:, -2:, :] = True
    return np.where(wall, 125, 0).astype(np.uint8)


def orifice_pipe(nx=30, ny=30, nz=32, R=12.0, r_hole=6.0, z0=10, thick=3):
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float64)
    r = np.hypot(x - (nx - 1) / 2, y - (ny - 1) / 2)
    wall = r > R
    wall |= (z >= z0) & (z < z0 + thick) & (r > r_hole)
    return np.where(wall, 125, 0).astype(np.uint8)


def sphere_duct(nx=28, ny=26, nz=36, rs=6.0):
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float64)
    wall = (x < 2) | (x > nx - 3) | (y < 2) | (y > ny - 3)
    wall |= (x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 3) ** 2 <= rs * rs
    return np.where(wall, 125, 0).astype(np.uint8)


GEOMETRIES = {"rough": rough_pipe, "orifice": orifice_pipe, "sphere": sphere_duct}


def flow_state(mask, ubar=0.03, pin=2e-3, noise=0.05, seed=7):
    """Pressure ramp, a plug-like axial flow and seeded noise on fluid cells."""
    nz, ny, nx = mask.shape
    rng = np.random.default_rng(seed)
    fl = mask == 0
    z = np.arange(nz)[:, None, None]
    p = np.where(fl, pin * (nz - z) / nz, 0.0)
    out = [p]
    for base in (0.0, 0.0, -ubar):
        u = base + noise * ubar * rng.uniform(-1, 1, mask.shape)
        out.append(np.where(fl, u, 0.0))
    return out


def max_rel(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def l2_pu(got, ref):
    ep = rel_l2(got[0], ref[0])
    num = sum(np.sum((np.asarray(got[f], dtype=np.float64) - ref[f]) ** 2) for f in (1, 2, 3))
    den = sum(np.sum(np.asarray(ref[f], dtype=np.float64) ** 2) for f in (1, 2, 3))
    return ep, float(np.sqrt(num / den))


@pytest.mark.parametrize("geom", sorted(GEOMETRIES))
@pytest.mark.parametrize("blend, d1", [(0.0, 2), (1.0, 2), (0.5, 1)])
def test_z_varying_mask_f64(orc, geom, blend, d1):
    mask = GEOMETRIES[geom]()
    assert any(not np.array_equal(mask[0], mask[k]) for k in range(1, mask.shape[0]))
    st = flow_state(mask)
    prm = RunParams(tblend=0.9, pin=2e-3, nu=0.02, d1=d1, delta=1, blend=blend)
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    ref = orc.step_folded(st, op, mask=mask, gauge=4e-4, nsteps=20)
    with Solver(mask.shape, prm, mask=mask, precision="f64", kernel=KERNEL_AUTO) as s:
        assert s.kernel == KERNEL_PLAIN  # the one kernel that takes this mask
        s.set_state(*st)
        s.set_gauge(4e-4)
        s.step(20)
        got = s.get_state()
    wall = mask != 0
    for f in range(4):
        assert not np.any(np.asarray(got[f])[wall]), f
        assert max_rel(got[f], ref[f]) <= 1e-12, f


@pytest.mark.parametrize("variant", [KERNEL_TILED, KERNEL_PLAIN])
@pytest.mark.parametrize("geom", sorted(GEOMETRIES))
def test_z_varying_mask_f32(orc, geom, variant):
    mask = GEOMETRIES[geom]()
    st = flow_state(mask)
    prm = RunParams(tblend=1.0, pin=2e-3, nu=0.015, d1=2, delta=1, blend=0.0)
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    ref32 = orc.step_folded(st, op, mask=mask, gauge=4e-4, nsteps=50, dtype=np.float32)
    ref64 = orc.step_folded(st, op, mask=mask, gauge=4e-4, nsteps=50)
    with Solver(mask.shape, prm, mask=mask, precision="f32") as s:
        assert s.kernel == KERNEL_TILED  # the default for FP32
        s.set_kernel(variant)
        s.set_state(*[a.astype(np.float32) for a in st])
        s.set_gauge(4e-4)
        s.step(50)
        got = s.get_state()
    assert max(l2_pu(got, ref32)) <= 1e-5
    assert max(l2_pu(got, ref64)) <= 1e-5


TOL16 = {"f16s": 3 * 2.0 ** -10 * (7 / 3) ** 0.5, "bf16s": 3 * 2.0 ** -7 * (7 / 3) ** 0.5}


@pytest.mark.parametrize("precision", ["f32", "f16s", "bf16s"])
@pytest.mark.parametrize("blend, d1", [(0.0, 2), (1.0, 2), (0.5, 1), (0.0, 1)])
@pytest.mark.parametrize("geom", sorted(GEOMETRIES))
def test_z_varying_mask_tiled(orc, geom, blend, d1, precision):
    """The tiled kernel on z-varying masks: d1 1 and 2, blend 0 (order-2
    corrector), 1 (order 4) and between, every 32-bit storage mode, with a
    gauge; wall cells stay exactly zero."""
    mask = GEOMETRIES[geom]()
    st = flow_state(mask)
    prm = RunParams(tblend=0.9, pin=2e-3, nu=0.02, d1=d1, delta=1, blend=blend)
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    with Solver(mask.shape, prm, mask=mask, precision=precision) as s:
        s.set_kernel(KERNEL_TILED)
        s.set_state(*([a.astype(np.float32) for a in st] if precision == "f32" else st))
        s.set_gauge(1e-3)
        s.step(7)
        got = s.get_state()
    wall = mask != 0
    for f in range(4):
        assert not np.any(np.asarray(got[f])[wall]), f
    if precision == "f32":
        ref32 = orc.step_folded([a.astype(np.float32) for a in st], op, mask=mask,
                                gauge=1e-3, nsteps=7, dtype=np.float32)
        assert max(l2_pu(got, ref32)) <= 1e-5
        return
    ref = orc.step_folded(st, op, mask=mask, gauge=1e-3, nsteps=7)
    _, eu = l2_pu(got, ref)
    assert eu <= TOL16[precision], (precision, eu)
    urms = np.sqrt(np.mean(sum(np.asarray(ref[f]) ** 2 for f in (1, 2, 3))))
    prms = np.sqrt(np.mean((np.asarray(got[0]) - ref[0]) ** 2))
    assert prms <= TOL16[precision] * urms, (precision, prms / urms)


@pytest.mark.parametrize("chunk", [5, 13, 40])
def test_z_varying_mask_tiled_chunks(orc, chunk):
    """Work items of 5 and 13 planes (chunk ends inside the bumps and at the
    periodic wrap) and one item per tile column: the wall codes of the
    planes around each chunk's first and last plane."""
    mask = rough_pipe()
    st = flow_state(mask, seed=3)
    prm = RunParams(tblend=1.0, pin=2e-3, nu=0.015, d1=2, delta=1, blend=1.0)
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    ref32 = orc.step_folded([a.astype(np.float32) for a in st], op, mask=mask, nsteps=9,
                            dtype=np.float32)
    with Solver(mask.shape, prm, mask=mask, precision="f32") as s:
        s.set_tile_chunk(chunk)
        s.set_state(*[a.astype(np.float32) for a in st])
        s.step(9)
        got = s.get_state()
        assert s.tile_chunk == min(chunk, mask.shape[0])
    assert max(l2_pu(got, ref32)) <= 1e-5


def test_z_varying_wide_rough_pipe(orc):
    """A rough pipe three tile columns wide and over two tile rows tall:
    the x exchange between tiles (halo columns) and the halo rows carry
    each plane's own wall flags."""
    mask = rough_pipe(nx=150, ny=70, nz=24, R=32.0, amp=3.0, nbumps=60, seed=4)
    assert mask.shape == (24, 70, 150)
    st = flow_state(mask, seed=9)
    prm = RunParams(tblend=0.9, pin=2e-3, nu=0.02, d1=2, delta=1, blend=0.5)
    op = orc.params(prm.tblend, prm.pin, prm.nu, prm.blend, prm.d1)
    ref32 = orc.step_folded([a.astype(np.float32) for a in st], op, mask=mask, gauge=2e-4,
                            nsteps=6, dtype=np.float32)
    with Solver(mask.shape, prm, mask=mask, precision="f32") as s:
        s.set_state(*[a.astype(np.float32) for a in st])
        s.set_gauge(2e-4)
        s.step(6)
        got = s.get_state()
    assert max(l2_pu(got, ref32)) <= 1e-5


def test_fp64_tiled_kernel_refuses_z_varying_mask():
    """FP64 (the parity mode) runs such masks on the plain kernel only."""
    mask = rough_pipe()
    prm = RunParams(tblend=1.0, pin=2e-3, nu=0.02, d1=2, delta=1, blend=0.0)
    with Solver(mask.shape, prm, mask=mask, precision="f64") as s:
        with pytest.raises(FslbmError):
            s.set_kernel(KERNEL_TILED)
        assert s.kernel == KERNEL_PLAIN
