"""The FileHFM wire format (hfm_io_read_files / hfm_io_write_files) and the FileHFM-compatible
executables, driven the way JMM_CPPLibs/FileIO/FileIO.py WriteCallRead drives the reference's."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from hamiltonfastmarching_amd import _lib
from hamiltonfastmarching_amd import synthetic as syn

LIBDIR = os.path.dirname(_lib.LIB_PATH)
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _get_array(L, io, key):
    nd = ctypes.c_int(0)
    shape = (ctypes.c_int64 * 8)()
    assert L.hfm_io_get_array_shape(io, key.encode(), ctypes.byref(nd), shape, 8) == 0
    dims = [shape[i] for i in range(nd.value)]
    out = np.zeros(int(np.prod(dims)) if dims else 1)
    assert L.hfm_io_get_array(io, key.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size) == 0
    return out.reshape(dims) if dims else out[0]


def test_read_files_roundtrip(tmp_path):
    L = _lib.load()
    p = syn.riemann2(16, seed=3)
    prefix = str(tmp_path / "input")
    oracle.write_rules(p, prefix)
    io = L.hfm_io_create()
    try:
        assert L.hfm_io_read_files(io, prefix.encode()) == 0, L.hfm_io_last_error(io)
        buf = ctypes.create_string_buffer(64)
        L.hfm_io_get_string(io, b"model", buf, 64)
        assert buf.value == b"Riemann2"
        assert np.array_equal(_get_array(L, io, "metric"), p["metric"])
        assert np.array_equal(_get_array(L, io, "seeds"), p["seeds"])
        assert _get_array(L, io, "gridScale") == p["gridScale"]
    finally:
        L.hfm_io_destroy(io)


def test_read_files_errors(tmp_path):
    L = _lib.load()
    io = L.hfm_io_create()
    try:
        assert L.hfm_io_read_files(io, str(tmp_path / "missing").encode()) == 1
        assert b"Input format file could not be opened" in L.hfm_io_last_error(io)
        prefix = str(tmp_path / "bad")
        oracle.write_rules({"model": "Riemann2", "dims": np.array([4.0, 4.0])}, prefix)
        with open(prefix + "_Data.dat", "ab") as f:
            f.write(b"\0" * 8)
        assert L.hfm_io_read_files(io, prefix.encode()) == 1
        assert b"input data file has incorrect size" in L.hfm_io_last_error(io)
    finally:
        L.hfm_io_destroy(io)


def test_executable_model_check(tmp_path):
    """FileHFM_<Model> applies to its model only (DispatchAndRun.h:227-233); no GPU needed."""
    exe = os.path.join(LIBDIR, "FileHFM_Riemann3")
    assert os.access(exe, os.X_OK)
    p = syn.riemann2(8, seed=1)
    oracle.write_rules(p, str(tmp_path / "input"))
    r = subprocess.run([exe, "input", "output"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "Executable applies to Riemann3, not to Riemann2." in r.stderr
    assert os.path.exists(tmp_path / "output_Format.txt")


@pytest.mark.gpu
@pytest.mark.parametrize("model,name,make", [
    ("Riemann2", "riemann2_64", lambda: syn.riemann2(64, seed=1)),
    ("ReedsSheppForward2", "rsf2_32x32x16", lambda: syn.rsf2(32, 16, seed=3))])
def test_executable_vs_golden(tmp_path, model, name, make):
    """WriteCallRead-style call of FileHFM_<Model> (no 'model' key: the name selects it) against the
    reference's fixtures: values, activeNeighs and geodesics from tips."""
    print(f"file executable {model}", flush=True)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    geo = np.load(os.path.join(GOLD, "geodesics.npz"))
    key = "riemann2" if model == "Riemann2" else "rsf2"
    p = {k: v for k, v in make().items() if k not in ("model", "tips")}
    p.update(exportValues=1.0, exportActiveNeighs=1.0, tips=geo[key + "_tips"])
    oracle.write_rules(p, str(tmp_path / "input"))
    exe = os.path.join(LIBDIR, f"FileHFM_{model}")
    r = subprocess.run(f"{exe} input output > log.txt", shell=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, open(tmp_path / "log.txt").read()
    out = oracle.read_rules(str(tmp_path / "output"))
    ref = g["values"]
    fin = np.isfinite(ref) & (ref > 0)
    assert np.max(np.abs(out["values"][fin] - ref[fin]) / ref[fin]) <= 1e-12
    assert np.mean(out["activeNeighs"] == g["activeNeighs"]) >= 0.999
    assert np.array_equal(out["geodesicLengths"].astype(int), geo[key + "_lengths"].astype(int))
    assert np.max(np.abs(out["geodesicPoints"] - geo[key + "_points"])) <= 1e-9
    for k in ("FMCPUTime", "StencilCPUTime", "GeodesicCPUTime"):
        assert k in out


def test_walls_input_errors():
    """The walls field is validated on the host before any device work (GetIntegralField,
    HFMInterface.hxx:33-42,288-292): wrong extent or depth is the reference's error."""
    import hamiltonfastmarching_amd as hfm
    p = syn.riemann2(16, seed=3)
    with pytest.raises(hfm.PyExp, match="walls has inconsistent dimensions"):
        hfm.run(dict(p, walls=np.zeros((16, 15))))
    with pytest.raises(hfm.PyExp, match="walls has incorrect depth"):
        hfm.run(dict(p, walls=np.zeros((16, 16, 2))))
