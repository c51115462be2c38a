"""Instantaneous-field files in the reference's restart format.

The reference writes each field of the flow with ``write_instantanous_array``
(src/io_restart.f90:55-69) through ``decomp_2d_write_one``: one file per
field, the whole global array as native FP64 in Fortran order (MPI-IO
subarray view, lib/2decomp-fft/src/io_mpi.f90:855-919), no header.  Files
are named ``domain<idom>_<keyword>_<iteration>.bin`` in ``1_data``
(``generate_pathfile_name`` / ``generate_file_name``, src/io_tools.f90:44-80;
``dir_data``, src/modules.f90:766) and a restart reads qx, qy, qz, pr back
with ``read_instantanous_flow`` (src/io_restart.f90:112-130).  These helpers write
and read the same files from a :class:`chapsim2_amd.domain.Domain`, so a run
can resume from the reference's files and the reference from ours.
"""
import os

import numpy as np

FLOW_FIELDS = ("qx", "qy", "qz", "pr")
DATA_DIR = "1_data"


def file_name(keyword, iteration, idom=1):
    """``domain<idom>_<keyword>_<iteration>.bin`` (integers without padding)."""
    return f"domain{int(idom)}_{keyword}_{int(iteration)}.bin"


def write_array(path, a):
    """The global array ``a`` as native FP64 in Fortran order, no header."""
    np.asarray(a, dtype=np.float64).ravel(order="F").tofile(path)


def read_array(path, shape):
    """A field written by :func:`write_array` (or by the reference) with the
    given global extents; the file size must match them exactly."""
    n = int(np.prod(shape))
    size = os.path.getsize(path)
    if size != 8 * n:
        raise ValueError(f"{path}: {size} bytes, expected {8 * n} for FP64 extents {tuple(shape)}")
    return np.fromfile(path, dtype=np.float64, count=n).reshape(shape, order="F")


def write_flow(dm, iteration, directory=DATA_DIR, idom=1):
    """Write qx, qy, qz and pr of ``dm`` as the reference's
    ``write_instantanous_flow`` (src/io_restart.f90:72-89) does; returns the
    file paths."""
    os.makedirs(directory, exist_ok=True)
    paths = []
    for key, a in zip(FLOW_FIELDS, dm.get_state()):
        p = os.path.join(directory, file_name(key, iteration, idom))
        write_array(p, a)
        paths.append(p)
    return paths


def read_flow(dm, iteration, directory=DATA_DIR, idom=1):
    """Load qx, qy, qz and pr written at ``iteration`` into ``dm`` (extents
    checked against the domain's staggered shapes), as the reference's
    ``read_instantanous_flow``; returns the arrays."""
    arrs = [read_array(os.path.join(directory, file_name(k, iteration, idom)), dm.shapes[k]) for k in FLOW_FIELDS]
    dm.set_state(*arrs)
    return arrs
