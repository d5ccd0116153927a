"""End to end through the drop-in: parameters1.txt (RK33 with LDG diffusion, 4 reference blocks)
-> C++ driver -> HDF5 files in the reference layout, checked against the oracle on the same run.
Both driver paths: the fused stage kernels (FUSED = 1, the default) and Grid::RK33 through the
Block operators (FUSED = 0)."""
import os
import subprocess

import numpy as np
import pytest

import h5read
from helpers import problem
from oracle_bind import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "3ddg_amd", "dg3d_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.gpu


def _params(tmp_path, extra):
    text = open(os.path.join(GOLDEN, "parameters1.txt")).read()
    text = text.replace("NUM_TIME_SUBSTEPS = 50", "NUM_TIME_SUBSTEPS = 3").replace("NUM_OUTSTEPS = 10", "NUM_OUTSTEPS = 2")
    text += extra
    pf = tmp_path / "p.txt"
    pf.write_text(text)
    return pf


def _oracle_reference_run(steps, diffusion=True, euler=False):
    """The same run on the oracle: 16x16x16 elements (2x2x1 blocks of 8x8x16), N=3, RK33 (or
    Grid::forwardEuler), DT=1e-6, B/A=5, D=5, linear sponge sigma_max 15000 width 1, Gaussian IC
    at the origin."""
    nE = (16, 16, 16)
    p = problem._base("p1", 3, nE, (-1.0, -1.0, -1.0), problem._uniform_widths(nE, (0.125,) * 3), 1500.0, 1000.0, 5.0)
    p.diffusion = np.full(p.E, 5.0)
    import importlib
    dg3d = importlib.import_module("3ddg_amd").dg3d
    p.sigma = dg3d.sponge_sigma(nE, (1, 1, 1), (0, 0, 0), (0.125,) * 3, (1,) * 6, 15000.0, 0)
    p.ic = dict(gauss=(1, 1, 1), amp_rho=1.0, amp_vel=(0.0, 0.0, 0.0), sigma=40.0, center=(0.0, 0.0, 0.0))
    o = Oracle(p, num_registers=4)
    o.fill_ic(0)
    dt = float(np.float32(0.000001))
    for _ in range(steps):
        if euler:
            o.forward_euler(dt, 3, 0)
        else:
            o.rk33(dt, 3, 0, 1, 2, adv=1, diff=1 if diffusion else 0, sponge=1)
        o.copy(0, 3)
    return p, o.get_state(0)


def _gather(p, out_dir, step):
    """Reassemble the 4 per-block files into the global field-major state."""
    n = p.n
    glob = np.zeros((4,) + tuple(p.num_elements) + (n, n, n))
    for rank in range(4):
        bi, bj = rank // 2, rank % 2
        rho = h5read.read(out_dir / "raw" / f"dRho_{rank}_{step}.h5", "rho")
        vel = h5read.read(out_dir / "raw" / f"vel_{rank}_{step}.h5", "velocity")
        assert rho.shape == (8 * n, 8 * n, 16 * n) and vel.shape == (3, 8 * n, 8 * n, 16 * n)
        t = h5read.read(out_dir / "raw" / f"dRho_{rank}_{step}.h5", "time")
        assert t.shape == (1,)
        # the reference's clock: float t += float DT once per step (driver.cpp:231,276)
        tf = np.float32(0.0)
        for _ in range(step):
            tf = np.float32(tf + np.float32(0.000001))
        assert t[0] == float(tf), (t[0], float(tf))
        for f, arr in enumerate([rho, vel[0], vel[1], vel[2]]):
            a = arr.reshape(8, n, 8, n, 16, n).transpose(0, 2, 4, 1, 3, 5)
            glob[f, bi * 8:(bi + 1) * 8, bj * 8:(bj + 1) * 8, :] = a
    return glob.reshape(-1)


@pytest.mark.skipif(not h5read.available(), reason="HDF5 library not available")
@pytest.mark.parametrize("case", ["fused", "operators", "fused_no_diffusion", "euler"])
def test_driver_parameters1_matches_oracle(tmp_path, case):
    """FUSED = 1 (fused stages) and FUSED = 0 (Grid::RK33 through the Block operators) with
    diffusion; EVALUATE_DIFFUSION = 0 on the fused path with DIFF_CONST still 5 (the stages must
    leave diffusion out); SCHEME = EULER (Grid::forwardEuler, which always runs the x-axis LDG
    pair)."""
    print(f"driver_parameters1_matches_oracle case={case}", flush=True)
    extra = {"fused": "FUSED = 1\n", "operators": "FUSED = 0\n", "euler": "SCHEME = EULER\n",
             "fused_no_diffusion": "FUSED = 1\nEVALUATE_DIFFUSION = 0\n"}[case]
    pf = _params(tmp_path, extra)
    out = tmp_path / "out"
    r = subprocess.run([DRIVER, str(pf), str(out)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Max dRho=" in r.stdout
    for rank in range(4):
        for name in ("coords", "ba", "diff", "rho0", "c0"):
            assert (out / "raw" / f"{name}_{rank}_0.h5").exists()
    c = h5read.read(out / "raw" / "coords_0_0.h5", "coords")
    assert c.shape == (3, 32, 32, 64)
    assert abs(c[0, 0, 0, 0] - (-1.0)) < 1e-12 and abs(c[2, 0, 0, -1] - 1.0) < 1e-12
    p, ref = _oracle_reference_run(6, diffusion=case != "fused_no_diffusion", euler=case == "euler")
    got = _gather(p, out, 6)
    a = got.reshape(4, -1)
    b = ref.reshape(4, -1)
    for f in range(4):
        err = np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f])
        assert err < 1e-12, (f, err)
