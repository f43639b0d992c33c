"""photons_force_device_static: the force evaluation without allocation, through the
C ABI, against photons_force_device on the same inputs.

Every buffer of the static evaluation comes from the library's static device scratch;
the kernels are those of the ordinary evaluation, so both give the same result bit for
bit.  Two calls in a row (the second on fewer particles, the scratch keeping the first
call's contents), a size above the maximum refused with a message, and the context's own
last evaluation left as it was.
"""
import ctypes as C

import numpy as np
import pytest

import ic
import photons_gpu as pg

pytestmark = pytest.mark.gpu


def clustered_box(ns, seed):
    n = ns ** 3
    box = 500000.0 * (n / 256.0 ** 3) ** (1.0 / 3.0)
    pos, _ = ic.clustered(n, box, seed=seed)
    return pos, ic.force_params(n, box, ns, rc_factor=6.0, order=4)


@pytest.mark.parametrize("precision", [pg.FP64, pg.FP32_MIXED])
def test_static_equals_ordinary_twice(precision):
    import torch
    L = pg.lib()
    for ns, seed in ((32, 3), (20, 4)):  # 32768 particles (the maximum), then fewer
        pos, prm = clustered_box(ns, seed)
        print(f"static evaluation, {ns}^3 particles, precision {precision}", flush=True)
        d_pos = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
        d_ord = torch.zeros_like(d_pos)
        d_sta = torch.full_like(d_pos, np.nan)
        ctx = pg.Context(prm, precision)
        ctx.force_device(len(pos), d_pos.data_ptr(), d_ord.data_ptr())
        ctx.sync()
        st_before = ctx.stats()
        rc = L.photons_force_device_static(ctx.h, len(pos), C.c_void_p(d_pos.data_ptr()),
                                           C.c_void_p(d_sta.data_ptr()))
        assert rc == 0, L.photons_last_error().decode()
        torch.cuda.synchronize()
        a_ord, a_sta = d_ord.cpu().numpy(), d_sta.cpu().numpy()
        assert np.isfinite(a_sta).all()
        assert np.array_equal(a_sta, a_ord)
        # the context's own evaluation is untouched
        assert ctx.stats() == st_before
        ctx.close()


def test_static_refuses_more_than_its_maximum():
    import torch
    pos, prm = clustered_box(33, 5)  # 35937 particles
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    d_acc = torch.zeros_like(d_pos)
    ctx = pg.Context(prm, pg.FP32_MIXED)
    L = pg.lib()
    rc = L.photons_force_device_static(ctx.h, len(pos), C.c_void_p(d_pos.data_ptr()), C.c_void_p(d_acc.data_ptr()))
    assert rc != 0
    assert b"32768" in L.photons_last_error()
    assert float(d_acc.abs().max()) == 0.0
