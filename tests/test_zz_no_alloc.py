"""xbg_mass_balance_noalloc (partials in a file-scope device array, no
allocation) against xbg_mass_balance of a context holding the same fields:
two calls in a row, the second on a smaller grid, results bit for bit."""
import ctypes as C

import numpy as np
import pytest

from xbeach_gpu_amd import Domain, setups
from xbeach_gpu_amd._ffi import lib

pytestmark = pytest.mark.gpu


def noalloc(hh_dev, zb_dev, n, dx):
    L = lib()
    w, b = C.c_double(), C.c_double()
    rc = L.xbg_mass_balance_noalloc(C.c_void_p(hh_dev), C.c_void_p(zb_dev), C.c_size_t(n), C.c_double(dx), None,
                                    C.byref(w), C.byref(b))
    assert rc == 0, L.xbg_noalloc_last_error().decode()
    return w.value, b.value


def test_noalloc_mass_balance_matches_context_entry_point(need_gpu):
    import torch

    rng = np.random.default_rng(17)
    for nx, ny, dx in ((300, 140, 2.5), (70, 33, 4.0)):   # second call: smaller input
        p = setups.small_params(nx=nx, ny=ny, dx=dx)
        hh = rng.uniform(0.0, 9.0, (ny, nx)).astype(np.float32)
        zb = rng.uniform(-3.0, 12.0, (ny, nx)).astype(np.float32)
        d = Domain(p)
        d.set("hh", hh)
        d.set("zb", zb)
        want = d.mass_balance()
        d.close()
        th = torch.from_numpy(hh).cuda()
        tz = torch.from_numpy(zb).cuda()
        torch.cuda.synchronize()
        got = noalloc(th.data_ptr(), tz.data_ptr(), nx * ny, dx)
        assert got == want, (nx, ny, got, want)
        assert got[0] == pytest.approx(float(hh.astype(np.float64).sum()) * dx * dx, rel=1e-12)
    # sizes beyond the maximum are refused before any launch
    L = lib()
    w, b = C.c_double(), C.c_double()
    rc = L.xbg_mass_balance_noalloc(C.c_void_p(th.data_ptr()), C.c_void_p(tz.data_ptr()), C.c_size_t(1 << 31),
                                    C.c_double(1.0), None, C.byref(w), C.byref(b))
    assert rc == -1 and b"exceeds the maximum" in L.xbg_noalloc_last_error()
