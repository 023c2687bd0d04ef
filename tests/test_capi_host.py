"""Host-side checks of the C ABI that run without a GPU: argument validation that
must happen before any device work."""
import ctypes as C
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    if not os.path.exists(sa.LIB_PATH):
        pytest.skip("libsa_hip.so not built")
    sa.lib()
    return sa


def test_batch_refuses_instances_beyond_32_bit_work_items(sa):
    """The batch kernels index an instance's work items (half-sweep items x systems)
    in 32 bits: a 2D L=11586 instance of 32 systems (67,119,698 x 32 >= 2^31) is
    refused before anything touches the device, and so is the served request."""
    L = sa.lib()
    T = np.full(32, 1.0)
    rid = C.c_int(-1)
    assert L.sa_batch_add(None, 1, 11586, 32, 1, 10, T, 0, 0.0, C.byref(rid)) == -2  # SA_ERR_UNSUPPORTED
    assert b"2^31" in L.sa_last_error()
    assert L.sa_batch_add_pt(None, 1, 11586, 32, T, 0.0, 1, 10, 1, C.byref(rid)) == -2
    r = sa.make_request(1, 11586, 1, 100, 7, nT=32, every=10)
    assert L.sa_request_bytes(C.byref(r)) == -2
    # one system fewer than the limit is accepted by the same check
    ok = sa.make_request(1, 11586, 1, 100, 7, nT=31, every=10)
    assert L.sa_request_bytes(C.byref(ok)) > 0
    # a valid geometry reaches the null-batch check
    assert L.sa_batch_add(None, 1, 64, 1, 1, 10, T, 0, 0.0, C.byref(rid)) == -1  # SA_ERR_ARG
