"""The C++ drop-in layer without a GPU: .nlos round trip through
NLosData::loadBinaryFile, and the error path for a missing capture."""
import json
import subprocess

import numpy as np

from fasterbackprojection_amd import nlos_io
from fasterbackprojection_amd.build import CLI
from fasterbackprojection_amd.scene import small_confocal


def test_nlos_roundtrip_python_and_cpp(tmp_path):
    cap = small_confocal(seed=0, n_scan=8, T=64, res=8, n_points=20)
    rec = nlos_io.from_geometry(cap.geometry, cap.transient)
    p = tmp_path / "cap.nlos"
    nlos_io.write(p, rec)
    back = nlos_io.read(p)
    np.testing.assert_array_equal(back.data, cap.transient.ravel())
    np.testing.assert_array_equal(back.laser_grid_positions, cap.geometry.laser_targets)
    assert back.dims == [8, 8, 64] and back.temporal_resolution == 64
    out = subprocess.run([str(CLI), "info", str(p)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    info = json.loads(out.stdout)
    assert info["dims"] == [8, 8, 64] and info["t"] == 64 and info["confocal"] == 1
    assert info["values"] == cap.transient.size
    assert abs(info["sum"] - float(cap.transient.astype(np.float64).sum())) < 1e-6 * max(1.0, info["sum"])
    assert np.allclose(info["hmin"], cap.geometry.hidden_min, atol=1e-7)
    assert abs(info["deltaT"] - np.float32(cap.geometry.time_step)) < 1e-12


def test_missing_capture_raises(tmp_path):
    out = subprocess.run([str(CLI), "info", str(tmp_path / "nope.hdf5")], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1 and "no readable .nlos" in out.stderr
