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


@pytest.mark.gpu
def test_driver_time_interpolation_matches_oracle(tmp_path):
    """SetTimeInterpolation(true) on the host: InformationFromGridOfDifferentLevel then also
    refills the ghost layers after odd substeps (TransferInfoFromCoarseGrid over the C ABI), like
    the oracle with the same option."""
    grid, kw = sc.amr2d(levels=3)
    proc, out = host_driver.run(grid, kw, 10, str(tmp_path), collision="mrt", time_interp=True)
    assert proc.returncode == 0, proc.stderr
    o = Oracle(grid, collision=(1,), time_interp=True, **kw)
    o.time_marching(10)
    for l in range(grid.max_level + 1):
        ref = o.f(l)
        assert np.max(np.abs(out[l] - ref)) < 1e-11 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_driver_graph_replay_matches_per_substep(tmp_path):
    """SetGraphReplay: the first AdvancingAtCurrentTime of each background step replays the whole
    step (zlh_time_marching; graphs of one step, for both buffer parities of level 0) and the
    finer levels' calls only check the order.  Its populations equal those of per-substep
    launches bit for bit, from an initial field given in the init file, and the oracle's from
    the same field; the driver reports the timed steps' wall time."""
    grid, kw = sc.amr2d(levels=3)
    init = [sc.level_fields(grid, l, 0.05) for l in range(grid.max_level + 1)]
    outs = []
    for replay in (True, False):
        d = tmp_path / ("replay" if replay else "substeps")
        d.mkdir()
        proc, out = host_driver.run(grid, kw, 5, str(d), collision="mrt", graph_replay=replay, init=init, warmup=1)
        assert proc.returncode == 0, proc.stderr
        assert host_driver.ms_per_step(proc) > 0
        outs.append(out)
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    o = Oracle(grid, collision=(1,), **kw)
    for l, (rho, u) in enumerate(init):
        o.init_macro(l, rho, u)
    o.time_marching(5)
    for l in range(grid.max_level + 1):
        ref = o.f(l)
        assert np.max(np.abs(outs[0][l] - ref)) < 1e-11 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_driver_immersed_boundary_matches_oracle(tmp_path):
    """SetImmersedBoundary through the driver's marker file: the D2Q9 channel with an immersed
    cylinder (101 markers, stencil 2) after 20 background steps, against the oracle with the same
    markers, under graph replay."""
    grid, kw, (xy, area) = sc.cylinder_channel()
    kw = dict(kw)
    kw.pop("has_force", None)
    proc, out = host_driver.run(grid, kw, 20, str(tmp_path), collision="srt",
                                ib=[(0, xy, area, np.zeros_like(xy), 2, 2.0)])
    assert proc.returncode == 0, proc.stderr
    o = Oracle(grid, collision=(0,), has_force=True, **kw)
    o.set_ib(0, xy, area, np.zeros_like(xy), stencil_dis=2)
    o.time_marching(20)
    ref = o.f(0)
    assert np.max(np.abs(out[0] - ref)) < 1e-10 * np.max(np.abs(ref))


def _cavity3d(n=10, lid=0.05):
    from zllbm_amd import grid as g
    grid = g.uniform_grid(3, (n, n, n))
    bc = sc.all_bc(3, sc.BB)
    bc[g.ZMAX] = (sc.BB, (lid, 0.0, 0.0))
    return grid, bc


@pytest.mark.gpu
def test_driver_learned_closure_matches_oracle(tmp_path):
    """The learned closure selected through the host's SetLesModel("learned") and
    SetLearnedClosure, on a 3D lid-driven cavity with FP64 populations, against the oracle (the
    device network runs in FP32, the oracle's in FP64)."""
    from tests.test_gpu_mlp import _random_model
    grid, bc = _cavity3d()
    nu = 0.002
    kw = dict(nu0=nu, bc=bc, les=2, cs=0.1)
    m = _random_model(3 * nu + 0.5)
    proc, out = host_driver.run(grid, kw, 20, str(tmp_path), collision="mrt", model=m)
    assert proc.returncode == 0, proc.stderr
    o = Oracle(grid, collision=(1,), **kw)
    o.set_mlp(m.params.astype(np.float64), m.mean, m.std, m.fmin, m.fmax, m.nu_scale)
    o.time_marching(20)
    ref = o.f(0)
    assert np.max(np.abs(ref - np.tile(ref.mean(0), (len(ref), 1)))) > 1e-4     # the lid has driven a flow
    assert np.max(np.abs(out[0] - ref)) < 2e-7 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_driver_bf16_precision(tmp_path):
    """precision = bf16 through the host setters: within BF16 accuracy of the oracle."""
    grid, bc = _cavity3d()
    kw = dict(nu0=0.01, bc=bc)
    proc, out = host_driver.run(grid, kw, 20, str(tmp_path), collision="mrt", precision="bf16")
    assert proc.returncode == 0, proc.stderr
    o = Oracle(grid, collision=(1,), **kw)
    o.time_marching(20)
    from tests.test_gpu_parity import _macros
    rho_o, u_o = _macros(o.f(0), 3)
    rho_d, u_d = _macros(out[0], 3)
    l2_u = np.sqrt(np.sum((u_d - u_o) ** 2) / np.sum(u_o ** 2))
    print("host driver, BF16 vs oracle: L2(u) %.3e, max |rho - rho_ref| %.3e" % (l2_u, np.max(np.abs(rho_d - rho_o))))
    assert l2_u < 1e-2 and np.max(np.abs(rho_d - rho_o)) < 1e-3



@pytest.mark.gpu
@pytest.mark.parametrize("case", ["amr2d", "amr3d"])
def test_driver_checkpoint_format_and_exact_restart(tmp_path, case):
    """The host writes the reference's grid-node checkpoint records (header, record order, flag
    and post-stream populations), and a run restarted from them continues as the uninterrupted
    run, bit for bit."""
    grid, kw = sc.amr2d(levels=3) if case == "amr2d" else sc.amr3d()
    q = 9 if grid.dims == 2 else 19
    ck = str(tmp_path / "grid_nodes")
    dirs = []
    for name in ("first", "restart", "straight"):
        (tmp_path / name).mkdir()
        dirs.append(str(tmp_path / name))
    proc, out10 = host_driver.run(grid, kw, 10, dirs[0], collision="mrt", checkpoint_out=ck)
    assert proc.returncode == 0, proc.stderr
    count, payload, data, rest = host_driver.read_checkpoint(ck, q)
    assert count == sum(lev.n for lev in grid.levels) and payload == 4 + 8 * q and rest == 0
    assert [(int(l), int(k)) for l, k in zip(data["level"], data["key"])] == host_driver.reference_record_order(grid)
    # payload: the grid's flags and the populations the driver reports after 10 steps (both in
    # key order; the grid's keys are sorted)
    for l, lev in enumerate(grid.levels):
        sel = np.flatnonzero(data["level"] == l)
        sel = sel[np.argsort(data["key"][sel])]
        np.testing.assert_array_equal(data["key"][sel], lev.keys)
        np.testing.assert_array_equal(data["flag"][sel], lev.flags)
        np.testing.assert_array_equal(data["f"][sel], out10[l])
    proc, restarted = host_driver.run(grid, kw, 10, dirs[1], collision="mrt", checkpoint_in=ck)
    assert proc.returncode == 0, proc.stderr
    proc, straight = host_driver.run(grid, kw, 20, dirs[2], collision="mrt")
    assert proc.returncode == 0, proc.stderr
    for l in range(grid.max_level + 1):
        np.testing.assert_array_equal(restarted[l], straight[l])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["amr2d", "amr3d"])
def test_driver_checkpoint_folder_and_exact_restart(tmp_path, case):
    """The checkpoint folder as the reference writes it (loadings, grid_nodes, interface_nodes):
    loadings and interface_nodes byte for byte as the reference's own loops produce them from this
    grid, grid_nodes as the single file; a run restarted from the folder continues bit for bit;
    a folder whose loadings describe another grid is refused."""
    grid, kw = sc.amr2d(levels=3) if case == "amr2d" else sc.amr3d()
    q = 9 if grid.dims == 2 else 19
    dirs = []
    for name in ("first", "restart", "straight", "bad"):
        (tmp_path / name).mkdir()
        dirs.append(str(tmp_path / name))
    ck = str(tmp_path / "ck")
    proc, _ = host_driver.run(grid, kw, 10, dirs[0], collision="mrt", checkpoint_dir_out=ck)
    assert proc.returncode == 0, proc.stderr
    with open(ck + "/loadings", "rb") as fh:
        assert fh.read() == host_driver.reference_loadings(grid)
    with open(ck + "/interface_nodes", "rb") as fh:
        got = fh.read()
    want = host_driver.reference_interface_nodes(grid)
    assert len(got) == len(want) and got == want
    count, payload, data, rest = host_driver.read_checkpoint(ck + "/grid_nodes", q)
    assert count == sum(lev.n for lev in grid.levels) and rest == 0
    proc, restarted = host_driver.run(grid, kw, 10, dirs[1], collision="mrt", checkpoint_dir_in=ck)
    assert proc.returncode == 0, proc.stderr
    proc, straight = host_driver.run(grid, kw, 20, dirs[2], collision="mrt")
    assert proc.returncode == 0, proc.stderr
    for l in range(grid.max_level + 1):
        np.testing.assert_array_equal(restarted[l], straight[l])
    raw = bytearray(open(ck + "/loadings", "rb").read())
    raw[-4] ^= 1                                            # the last cell's load
    open(ck + "/loadings", "wb").write(bytes(raw))
    proc, _ = host_driver.run(grid, kw, 1, dirs[3], collision="mrt", checkpoint_dir_in=ck)
    assert proc.returncode != 0 and "loadings do not describe" in proc.stderr


@pytest.mark.parametrize("case", ["amr2d", "amr3d"])
def test_checkpoint_order_rule_matches_reference_loops(case):
    """The host's record order (sort by background cell, level, then z, y, x inside the cell)
    equals the order of the reference writer's nested loops."""
    from zllbm_amd import grid as g
    grid, _ = sc.amr2d(levels=3) if case == "amr2d" else sc.amr3d()
    dims = grid.dims
    recs = []
    for l, lev in enumerate(grid.levels):
        c = g.morton_decode(dims, lev.keys).astype(np.int64)
        m = (1 << l) - 1
        for k, cc in zip(lev.keys, c):
            recs.append((int(k) >> (dims * l), l, int(cc[2]) & m if dims == 3 else 0, int(cc[1]) & m,
                         int(cc[0]) & m, int(k)))
    recs.sort()
    assert [(r[1], r[5]) for r in recs] == host_driver.reference_record_order(grid)


def _vtu_arrays(path):
    import xml.etree.ElementTree as ET
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
    arrays = {}
    for da in piece.iter("DataArray"):
        name = da.get("Name") or "points"
        dtype = {"Int32": np.int32, "Int64": np.int64, "UInt8": np.uint8, "Float32": np.float32}[da.get("type")]
        arrays[name] = np.array(da.text.split(), dtype=np.float64).astype(dtype)
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["amr2d", "amr3d"])
def test_driver_vtu_output(tmp_path, case):
    """VTU files of every level: flags, FP32 density and velocity of the final populations, node
    coordinates, and one pixel / voxel per node whose cell's corners are all nodes of the level."""
    from zllbm_amd import grid as g
    from tests.test_gpu_parity import _macros
    grid, kw = sc.amr2d(levels=3) if case == "amr2d" else sc.amr3d()
    dims = grid.dims
    prefix = str(tmp_path / "flow")
    proc, out = host_driver.run(grid, kw, 5, str(tmp_path), collision="mrt", vtu_out=prefix)
    assert proc.returncode == 0, proc.stderr
    pvtu = open(prefix + ".pvtu").read()
    for l, lev in enumerate(grid.levels):
        assert 'Source="flow_level%d.vtu"' % l in pvtu
        npts, ncell, a = _vtu_arrays(prefix + "_level%d.vtu" % l)
        assert npts == lev.n
        np.testing.assert_array_equal(a["flag"], lev.flags)
        rho, u = _macros(out[l], dims)
        np.testing.assert_allclose(a["rho"], rho.astype(np.float32), rtol=2e-7)
        np.testing.assert_allclose(a["velocity"].reshape(-1, dims), u.astype(np.float32), rtol=1e-5, atol=1e-9)
        ijk = g.morton_decode(dims, lev.keys).astype(np.int64)
        pts = a["points"].reshape(-1, 3)[:, :dims]
        np.testing.assert_array_equal(pts, (ijk / float(1 << l)).astype(np.float32))
        # cells: every node whose +1 neighbours along all axes (and diagonals) exist
        pos = {tuple(c): i for i, c in enumerate(ijk)}
        corners = [np.array([(v >> a) & 1 for a in range(dims)]) for v in range(1 << dims)]
        expect = [[pos[tuple(c + d)] for d in corners] for c in ijk if all(tuple(c + d) in pos for d in corners)]
        assert ncell == len(expect)
        np.testing.assert_array_equal(a["connectivity"].reshape(-1, 1 << dims), np.array(expect).reshape(-1, 1 << dims))
        assert np.all(a["types"] == (11 if dims == 3 else 8))
        np.testing.assert_array_equal(a["offsets"], (1 << dims) * np.arange(1, ncell + 1))


def _write_checkpoint(path, grid, q, records=None, payload=None):
    """A grid-node checkpoint of the grid at rest (f = 1/q per direction; values do not matter
    here), with optional record list override and header payload size."""
    rec = np.dtype([("key", "<u8"), ("level", "<i4"), ("flag", "<i4"), ("f", "<f8", (q,))])
    if records is None:
        records = [(int(k), l, int(fl)) for l, lev in enumerate(grid.levels) for k, fl in zip(lev.keys, lev.flags)]
    data = np.zeros(len(records), rec)
    for i, (k, l, fl) in enumerate(records):
        data[i] = (k, l, fl, np.full(q, 1.0 / q))
    with open(path, "wb") as fh:
        np.array([len(records)], np.uint32).tofile(fh)
        np.array([4 + 8 * q if payload is None else payload], np.int32).tofile(fh)
        data.tofile(fh)


@pytest.mark.parametrize("defect", ["payload", "duplicate", "foreign", "truncated"])
def test_driver_rejects_bad_checkpoint(tmp_path, defect):
    """A checkpoint that does not fit the grid stops the driver with a message before any device
    work: a wrong payload size, a node recorded twice (and another not at all), a key the grid
    does not have, a truncated file."""
    grid, kw = sc.amr2d(levels=2)
    q = 9
    recs = [(int(k), l, int(fl)) for l, lev in enumerate(grid.levels) for k, fl in zip(lev.keys, lev.flags)]
    path = str(tmp_path / "ck")
    if defect == "payload":
        _write_checkpoint(path, grid, q, payload=4 + 8 * q + 8)
        msg = "payload size"
    elif defect == "duplicate":
        recs[1] = recs[0]
        _write_checkpoint(path, grid, q, recs)
        msg = "two records"
    elif defect == "foreign":
        recs[0] = (int(recs[0][0]) | (1 << 62), recs[0][1], recs[0][2])
        _write_checkpoint(path, grid, q, recs)
        msg = "not in the grid"
    else:
        _write_checkpoint(path, grid, q)
        with open(path, "r+b") as fh:
            fh.truncate(8 + 50 * (16 + 8 * q))
        msg = "truncated"
    proc, out = host_driver.run(grid, kw, 1, str(tmp_path), checkpoint_in=path)
    assert proc.returncode == 1 and out is None
    assert msg in proc.stderr, proc.stderr
