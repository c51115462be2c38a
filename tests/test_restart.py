"""Instantaneous-field files in the reference's restart format
(src/io_restart.f90:55-130, lib/2decomp-fft/src/io_mpi.f90:855-919): names,
byte layout (native FP64, Fortran order, no header) and round trips."""
import os

import numpy as np
import pytest

from chapsim2_amd import restart


def test_file_names_follow_generate_file_name():
    # "domain"//int2str(idom)//'_'//keyword//'_'//int2str(iter)//".bin", int2str without padding
    assert restart.file_name("qx", 1000) == "domain1_qx_1000.bin"
    assert restart.file_name("pr", 0, idom=2) == "domain2_pr_0.bin"


def test_byte_layout_is_fortran_order_fp64(tmp_path):
    shape = (3, 4, 5)
    a = np.arange(np.prod(shape), dtype=np.float64).reshape(shape, order="F") * 0.5 - 7.0
    p = tmp_path / "f.bin"
    restart.write_array(p, a)
    raw = p.read_bytes()
    assert len(raw) == 8 * a.size
    # element (i, j, k) (0-based) sits at byte 8 * (i + 3 j + 12 k), x fastest, little endian
    v = np.frombuffer(raw, dtype="<f8")
    for i, j, k in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 3, 4)):
        assert v[i + 3 * j + 12 * k] == a[i, j, k]
    np.testing.assert_array_equal(restart.read_array(p, shape), a)


def test_size_mismatch_is_an_error(tmp_path):
    p = tmp_path / "f.bin"
    restart.write_array(p, np.zeros((4, 4, 4)))
    with pytest.raises(ValueError):
        restart.read_array(p, (4, 5, 4))


class _FakeDomain:
    """The state interface of Domain, on the host."""

    def __init__(self, nc):
        nx, ny, nz = nc
        self.shapes = {"qx": (nx, ny, nz), "qy": (nx, ny + 1, nz), "qz": (nx, ny, nz), "pr": (nx, ny, nz)}
        rng = np.random.default_rng(3)
        self.state = [np.asfortranarray(rng.standard_normal(self.shapes[k])) for k in restart.FLOW_FIELDS]

    def get_state(self):
        return [a.copy(order="F") for a in self.state]

    def set_state(self, *arrs):
        self.state = [np.asfortranarray(a) for a in arrs]


def test_flow_round_trip(tmp_path):
    a = _FakeDomain((6, 5, 4))
    paths = restart.write_flow(a, 120, directory=str(tmp_path))
    assert [os.path.basename(p) for p in paths] == [f"domain1_{k}_120.bin" for k in ("qx", "qy", "qz", "pr")]
    b = _FakeDomain((6, 5, 4))
    b.state = [np.zeros_like(x) for x in b.state]
    restart.read_flow(b, 120, directory=str(tmp_path))
    for x, y in zip(a.state, b.state):
        np.testing.assert_array_equal(x, y)
