"""The C++ host drop-in (reference-style SolverInterface / GridInfoInterface over the C ABI)."""
import numpy as np
import pytest

from oracle.zlbm_oracle import Oracle
from tests import host_driver, scenarios as sc


def test_driver_reports_device_errors(tmp_path):
    """Without a usable device (or with an invalid setup) the driver exits non-zero with the
    library's message instead of aborting."""
    grid, kw = sc.cavity(n=8)
    kw = dict(kw)
    kw["bc"] = {1: kw["bc"][1]}            # missing boundary conditions: zlh_create rejects
    proc, out = host_driver.run(grid, kw, 1, str(tmp_path))
    assert proc.returncode == 1 and out is None
    assert "zlh_create failed" in proc.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["cavity", "amr2d"])
def test_driver_matches_oracle(tmp_path, case):
    if case == "cavity":
        grid, kw = sc.cavity(n=24)
        coll = "mrt"
    else:
        grid, kw = sc.amr2d(levels=3)
        coll = "srt"
    proc, out = host_driver.run(grid, kw, 10, str(tmp_path), collision=coll)
    assert proc.returncode == 0, proc.stderr
    o = Oracle(grid, collision=(1 if coll == "mrt" else 0,), **kw)
    o.time_marching(10)
    for l in range(grid.max_level + 1):
        ref = o.f(l)
        assert np.max(np.abs(out[l] - ref)) < 1e-11 * np.max(np.abs(ref))
