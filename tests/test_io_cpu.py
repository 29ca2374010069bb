"""HDF5 snapshot output (periodicflow.io, after src/periodicflow/io.py:15-196)
on the CPU: the run-time layout (shenfun's ``<var>/<d>D/<step>`` and
``<var>/mesh/x<i>``), ``reconfigure_dataset``'s flat layout (root x, y[, z],
t; per-variable datasets named by the zero-padded step; periodic end
planes), read back through the writer's own libhdf5 binding and, where the
image has it, independently through h5py (a separate interpreter) and
h5dump."""
import json
import os
import subprocess

import numpy as np
import pytest

from periodicflow import io as pio
from periodicflow.basis import FourierSpace

H5PY_PYTHON = "/opt/conda/bin/python3.9"
H5DUMP = "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def h5():
    try:
        return pio.hdf5()
    except OSError as e:
        pytest.skip(f"no libhdf5: {e}")


def _fields(rs, shape, names):
    return {nm: [rs.standard_normal(shape)] for nm in names}


def _run(space, names, steps, tmp_path, periodic, end_time=1.0, fname="snap"):
    """Write one snapshot per step (fresh values each time, written in place
    into the arrays the writer holds) and return the writer and the values."""
    rs = np.random.RandomState(5)
    data = _fields(rs, space.N, names)
    w = pio.HDF5Writer(None, 0, 1, str(tmp_path / fname), solution={"space": space, "data": data},
                       time_points=[0.0, end_time], end_points=[b for _, b in space.domain], periodic=periodic)
    written = {}
    for st in steps:
        for nm in names:
            data[nm][0][...] = rs.standard_normal(space.N)
        written[st] = {nm: data[nm][0].copy() for nm in names}
        w.write(st)
    return w, written


def _h5py_dump(path):
    """Structure and contents of a file as h5py sees it (names in h5py's
    iteration order, shapes, dtypes, values)."""
    code = (
        "import h5py, json, sys\n"
        "out = {}\n"
        "def visit(name, obj):\n"
        "    if isinstance(obj, h5py.Dataset):\n"
        "        out[name] = {'shape': list(obj.shape), 'dtype': str(obj.dtype), 'data': obj[()].ravel().tolist()}\n"
        "with h5py.File(sys.argv[1], 'r') as f:\n"
        "    out['__root__'] = list(f)\n"
        "    f.visititems(visit)\n"
        "print(json.dumps(out))\n"
    )
    r = subprocess.run([H5PY_PYTHON, "-c", code, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_periodic_bc_matches_wrap_padding():
    a = np.arange(24.0).reshape(2, 3, 4)
    assert np.array_equal(pio.periodic_bc(a), np.pad(a, [(0, 1)] * 3, mode="wrap"))
    b = np.arange(6.0).reshape(2, 3)
    assert np.array_equal(pio.periodic_bc(b), np.pad(b, [(0, 1)] * 2, mode="wrap"))


def test_h5_path():
    assert str(pio.h5_path("TG")) == "TG.h5"
    assert str(pio.h5_path("out/TG.h5")) == "out/TG.h5"


def test_run_layout_3d(h5, tmp_path):
    space = FourierSpace(N=[8, 6, 4], domain=[(0.0, 2 * np.pi), (0.0, 1.0), (-1.0, 1.0)])
    steps = [0, 5, 10, 100]
    w, written = _run(space, ["u", "v", "w"], steps, tmp_path, periodic=False)
    w.close()
    fid = h5.open(w.path)
    try:
        assert h5.members(fid) == ["u", "v", "w"]
        assert h5.members(fid, "u") == ["3D", "mesh"]
        assert sorted(h5.members(fid, "u/3D"), key=int) == ["0", "5", "10", "100"]
        assert h5.members(fid, "v/mesh") == ["x0", "x1", "x2"]
        # x_j = a + j (b - a) / N on every axis
        for i, (a, b) in enumerate(space.domain):
            n = space.N[i]
            assert np.allclose(h5.read(fid, f"w/mesh/x{i}"), a + np.arange(n) * (b - a) / n, rtol=0, atol=1e-15)
        for st in steps:
            for nm in ("u", "v", "w"):
                assert np.array_equal(h5.read(fid, f"{nm}/3D/{st}"), written[st][nm])
    finally:
        h5.close(fid)
    # after close (the end of a solve) a later write reopens the file and appends
    w.write(200)
    w.close()
    fid = h5.open(w.path)
    try:
        assert sorted(h5.members(fid, "u/3D"), key=int) == ["0", "5", "10", "100", "200"]
        assert np.array_equal(h5.read(fid, "u/3D/100"), written[100]["u"])
    finally:
        h5.close(fid)


@pytest.mark.parametrize("periodic", [False, True])
def test_reconfigure_3d(h5, tmp_path, periodic):
    space = FourierSpace(N=[8, 6, 4], domain=[(0.0, 2 * np.pi), (0.0, 1.0), (-1.0, 1.0)])
    steps = [0, 5, 10, 100]
    w, written = _run(space, ["u", "v", "w"], steps, tmp_path, periodic, end_time=2.5)
    w.reconfigure_dataset()
    assert w.path.exists() and not (tmp_path / "snap_temp.h5").exists()
    shape = tuple(n + (1 if periodic else 0) for n in space.N)
    fid = h5.open(w.path)
    try:
        assert h5.members(fid) == ["t", "u", "v", "w", "x", "y", "z"]
        assert h5.members(fid, "u") == ["000", "005", "010", "100"]
        assert np.allclose(h5.read(fid, "t"), [0.0, 2.5 / 3, 5.0 / 3, 2.5])
        for i, lab in enumerate("xyz"):
            a, b = space.domain[i]
            n = space.N[i]
            want = a + np.arange(n) * (b - a) / n
            if periodic:
                want = np.append(want, b)
            assert np.allclose(h5.read(fid, lab), want, rtol=0, atol=1e-15)
        for st in steps:
            for nm in ("u", "v", "w"):
                got = h5.read(fid, f"{nm}/{st:03d}")
                assert got.shape == shape
                want = written[st][nm]
                if periodic:
                    want = np.pad(want, [(0, 1)] * 3, mode="wrap")
                assert np.array_equal(got, want)
    finally:
        h5.close(fid)


def test_reconfigure_2d_read_by_h5py(h5, tmp_path):
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no second interpreter with h5py")
    space = FourierSpace(N=[6, 4], domain=2 * np.pi)
    steps = [0, 2, 4, 6, 8, 10, 12]
    w, written = _run(space, ["u", "v"], steps, tmp_path, periodic=True, end_time=1.2, fname="thermal")
    w.reconfigure_dataset()
    out = _h5py_dump(w.path)
    assert out["__root__"] == ["t", "u", "v", "x", "y"]
    names = ["00", "02", "04", "06", "08", "10", "12"]
    for st, nm in zip(steps, names):
        for var in ("u", "v"):
            d = out[f"{var}/{nm}"]
            assert d["shape"] == [7, 5] and d["dtype"] == "float64"
            assert np.array_equal(np.array(d["data"]).reshape(7, 5), np.pad(written[st][var], [(0, 1)] * 2, mode="wrap"))
    assert np.allclose(out["t"]["data"], np.linspace(0.0, 1.2, 7))
    assert np.allclose(out["x"]["data"], np.append(np.arange(6) * 2 * np.pi / 6, 2 * np.pi))
    assert out["y"]["shape"] == [5]


def test_run_file_read_by_h5py_and_h5dump(h5, tmp_path):
    if not os.path.exists(H5PY_PYTHON) or not os.path.exists(H5DUMP):
        pytest.skip("no h5py interpreter or h5dump")
    space = FourierSpace(N=[4, 4, 4], domain=2 * np.pi)
    w, written = _run(space, ["u", "v", "w"], [0, 3], tmp_path, periodic=False)
    w.close()
    out = _h5py_dump(w.path)
    assert out["__root__"] == ["u", "v", "w"]
    assert np.array_equal(np.array(out["v/3D/3"]["data"]).reshape(4, 4, 4), written[3]["v"])
    assert out["u/mesh/x2"]["shape"] == [4]
    r = subprocess.run([H5DUMP, "-H", str(w.path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and 'DATASET "0"' in r.stdout and "H5T_IEEE_F64LE" in r.stdout


def test_reconfigure_missing_file(h5, tmp_path):
    space = FourierSpace(N=[4, 4], domain=2 * np.pi)
    w, _ = _run(space, ["u", "v"], [0], tmp_path, periodic=False)
    w.close()
    w.path.unlink()
    with pytest.raises(FileNotFoundError):
        w.reconfigure_dataset()
