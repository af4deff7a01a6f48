"""Argument errors of the C ABI (no GPU): every entry point validates its arguments on the host before it
enqueues anything, returns BZ_EINVAL with a message naming itself, and the Python layer raises ValueError
(the reference's own error type for bad arguments, fdtd.py:33-63). The pointers passed here are non-null
placeholders that the failing checks never dereference."""
import pytest

from beamz_amd import _native as N

FAKE = 256          # a non-null address, never read: every call below fails a check first


def einval(name, *args, match):
    with pytest.raises(ValueError, match=match):
        N.call(name, *args)


def test_3d_step_checks():
    a = N.ptrs6([FAKE * (q + 1) for q in range(6)])
    b = N.ptrs6([FAKE * (q + 11) for q in range(6)])
    geo = (1e-8, 1e-8, 1e-8, 1e-17, 1.256e-6, 8.85e-12)
    einval("bz_yee3d_step2", a, a, FAKE, FAKE, 10, 10, 10, 64, *geo, None, None, 0, None, None, 0, 2, 8, None,
           match="bz_yee3d_step2: bad field buffers")                          # src and dst alike
    einval("bz_yee3d_step2", a, b, FAKE, FAKE, 10, 10, 1, 64, *geo, None, None, 0, None, None, 0, 2, 8, None,
           match="bz_yee3d_step2: bad shape")
    einval("bz_yee3d_step2", a, b, FAKE, FAKE, 10, 10, 10, 8, *geo, None, None, 0, None, None, 0, 2, 8, None,
           match="bz_yee3d_step2: bad shape")                                  # row pitch below nx
    einval("bz_yee3d_step2", a, b, FAKE, FAKE, 10, 10, 10, 64, *geo, None, None, 1, None, None, 0, 2, 8, None,
           match="bz_yee3d_step2: bad source list")                            # a box count without boxes


def test_2d_and_adjoint_checks():
    einval("bz_yee2d_run", None, 0, None, 0, 2, None, 1, None, 0, None, None, 0, 0, None, 0, None, None, None,
           match="bz_yee2d_run: bad arguments")
    einval("bz_yee2d_run", FAKE, 1, FAKE, 1, 2, FAKE, 1, None, 5, None, None, 0, 0, None, 0, None, None, None,
           match="bz_yee2d_run: monitor points without a slot table")
    einval("bz_yee2d_run", FAKE, 1, FAKE, 1, 2, FAKE, 1, FAKE, 5, FAKE, FAKE, 0, 3, None, 0, None, None, None,
           match="bz_yee2d_run: point groups need nslot > 0 and nrow > 0")
    einval("bz_yee2d_run2", FAKE, 1, FAKE, 1, 8, FAKE, 2, 1, None, None, 0, None, None, 0, 0, None, 2, None, None,
           None, match="bz_yee2d_run2: bad arguments")                         # rd0 must be 0 or 1
    einval("bz_yee2d_runk", FAKE, 1, FAKE, 1, 8, 8, 4, FAKE, 0, 1, FAKE, FAKE, FAKE, 5, FAKE, FAKE, 10, 0, None, 0,
           None, None, 0, None, match="bz_yee2d_runk: point groups need nslot > 0 and nrow > 0")
    einval("bz_adj2d_run2", FAKE, 1, FAKE, 1, 4, None, 0, 2, 10, 3, 0, None,
           match="bz_adj2d_run2: an odd step count needs one-step tiles")
    einval("bz_adj2d_run2", FAKE, 1, FAKE, 1, 4, FAKE, 1, 2, 10, 2, 2, None, match="bz_adj2d_run2: bad arguments")
    einval("bz_adj2d_run2", FAKE, 1, FAKE, 1, 4, FAKE, 1, 2, 2, 4, 0, None,
           match="bz_adj2d_run2: bad arguments")                               # more steps than exist below n_hi
    einval("bz_adj2d_run", FAKE, 1, FAKE, 1, 16, 10, 2, None, match="bz_adj2d_run: seg must be 2, 4 or 8")


def test_surrogate_and_watchdog_checks():
    einval("bz_nn_conv3x3", FAKE, FAKE, None, FAKE, 1, 4, 4, 33, 32, 1, None, match="multiple of 32")
    einval("bz_nonfinite_count", FAKE, -1, FAKE, None, match="bz_nonfinite_count: bad arguments")
    L = N.lib()
    need = L.bz_nn_wgrad_work(1, 4, 4, 32, 32, 1)
    assert need > 32 * 9 * 32 and L.bz_nn_wgrad_work(0, 4, 4, 32, 32, 1) == -1
    einval("bz_nn_conv3x3_wgrad", FAKE, FAKE, FAKE, FAKE, 1, 4, 4, 32, 32, 1, FAKE, need - 1, None,
           match="workspace smaller than bz_nn_wgrad_work")
    first = L.bz_nn_first_work(1, 8, 8, 32)
    assert first >= 320 and L.bz_nn_first_work(1, 8, 8, 0) == -1
    einval("bz_nn_conv_first_bwd", FAKE, FAKE, FAKE, FAKE, FAKE, None, 1, 8, 8, 32, FAKE, first - 1, None,
           match="workspace smaller than bz_nn_first_work")


def test_gradient_pass_injection_table_filter_and_weight_copy_checks():
    import ctypes
    einval("bz_adj2d_grad", None, 1, 16, 0, 10, None, match="bz_adj2d_grad: bad arguments")
    einval("bz_adj2d_grad", FAKE, 1, 16, 10, 5, None, match="bz_adj2d_grad: bad arguments")   # n_hi below n_lo
    einval("bz_adj2d_injtab", FAKE, 0, 16, 1, 10, None, match="bz_adj2d_injtab: bad arguments")
    einval("bz_adj2d_injtab", FAKE, 1, -1, 1, 10, None, match="bz_adj2d_injtab: bad arguments")
    einval("bz_filter2d", FAKE, FAKE, 1, 8, 8, FAKE, 2, None, None, None,
           match="bz_filter2d: null or aliased pointer")                        # in and out alike
    einval("bz_filter2d", FAKE, FAKE * 2, 1, 8, 8, FAKE, 9, None, None, None, match="radius must be 0..8")
    einval("bz_filter2d", FAKE, FAKE * 2, 0, 8, 8, FAKE, 2, None, None, None, match="bz_filter2d: bad shape")
    P, I = ctypes.c_void_p * 2, ctypes.c_int * 2
    w, w16, wt16 = P(FAKE, FAKE * 2), P(FAKE * 3, FAKE * 4), P(None, None)
    einval("bz_nn_weights_n", ctypes.addressof(w), ctypes.addressof(w16), ctypes.addressof(wt16),
           ctypes.addressof(I(8, 8)), ctypes.addressof(I(8, 8)), 17, None, match="1..16 layers")
    einval("bz_nn_weights_n", ctypes.addressof(w), ctypes.addressof(w16), ctypes.addressof(wt16),
           ctypes.addressof(I(8, 0)), ctypes.addressof(I(8, 8)), 2, None, match="bz_nn_weights_n: bad layer")
