"""The C++ drop-in of the integration notes compiles and links against libsa_hip.so
as code written for the reference's gpu_sys would (SA_REFERENCE_NAMES, gpu_3d /
gpu_2d per temperature, parallel_tempering<gpu_sys>, a user class derived from
gpu_sys without get_T / clone, delete through the base pointer).  No GPU needed."""
import os
import tempfile

import pytest

import dropin_build


def test_dropin_program_compiles_and_links():
    if not os.path.exists(os.path.join(dropin_build.LIBDIR, "libsa_hip.so")):
        pytest.skip("libsa_hip.so not built")
    with tempfile.TemporaryDirectory() as d:
        exe = dropin_build.build(os.path.join(d, "pt_dropin"))
        assert os.access(exe, os.X_OK)
