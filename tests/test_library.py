"""The C-ABI library: it loads without a GPU, exports every entry point the
public headers declare, and its host-side set-up (exact projection
operators) matches the oracle.  No compute entry point is called here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            with open(os.path.join(inc, fn)) as f:
                txt = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(chs_\w+)\s*\(", txt, flags=re.M))
    return names


def test_exports_every_declared_symbol(chs):
    from chapsim2_amd import _lib

    names = _declared()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(chs, n), n
    assert names == set(_lib.EXPORTS), names ^ set(_lib.EXPORTS)
    assert chs.chs_abi_version() >= 1


def test_create_without_device_fails_cleanly(chs):
    """Without a GPU, chs_create returns an error code instead of crashing."""
    import torch

    if torch.cuda.is_available():
        return
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import make_desc

    d = make_desc((16, 16, 16), 1.0, -1.0, 1.0, 1.0)
    h = C.c_void_p()
    rc = chs.chs_create(C.byref(d), C.byref(h))
    assert rc == _lib.CHS_EHIP and not h.value


def test_exact_projection_operators_match_oracle(chs, po):
    from chapsim2_amd import _lib
    from chapsim2_amd.domain import make_desc

    for nc, acc in (((24, 20, 12), 4), ((16, 32, 10), 2), ((20, 16, 8), 3)):
        lx, lz = 2 * np.pi, np.pi
        d = make_desc(nc, lx, -1.0, 1.0, lz, istret=2, rstret=0.075, iacc=acc)
        ny = nc[1]
        ly = np.zeros(ny * ny)
        lam = np.zeros(ny)
        lxv = np.zeros(nc[0] // 2 + 1)
        lzv = np.zeros(nc[2])
        res = C.c_double()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        _lib.check(chs.chs_inspect_projection(C.byref(d), dp(ly), dp(lam), dp(lxv), dp(lzv), C.byref(res)))
        ly = ly.reshape(ny, ny)
        dom = po.Dom(list(nc), [1, 0, 1], lx, -1.0, 1.0, lz, 2, 0.075)
        G = np.zeros((ny + 1, ny))
        D = np.zeros((ny, ny + 1))
        for k in range(ny):
            e = np.zeros(ny)
            e[k] = 1
            G[:, k] = dom.op1d(1, 4, acc, [5, 5], e, ny + 1, np.zeros(4))
        for k in range(ny + 1):
            e = np.zeros(ny + 1)
            e[k] = 1
            D[:, k] = dom.op1d(1, 5, acc, [4, 4], e, ny)
        G[0] = G[-1] = 0.0
        Lo = D @ G
        assert np.abs(ly - Lo).max() < 1e-13 * np.abs(Lo).max()
        w = np.sort(np.linalg.eigvals(Lo).real)
        assert np.abs(w - lam).max() < 1e-11 * np.abs(w).max()
        assert res.value < 1e-12
        # Fourier blocks: minus the square of the staggered scheme's
        # modified wavenumber
        for n, L, vals in ((nc[0], lx, lxv), (nc[2], lz, lzv)):
            e0 = np.zeros(n)
            e0[0] = 1.0
            col = dom.op1d(0 if vals is lxv else 2, 5, acc, [1, 1],
                           dom.op1d(0 if vals is lxv else 2, 4, acc, [1, 1], e0, n), n)
            ref = np.fft.fft(col).real[: len(vals)]
            assert np.abs(vals - ref).max() < 1e-12 * np.abs(ref).max()


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
