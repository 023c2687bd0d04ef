"""The C-ABI library loads on a machine without a GPU and exports every function
include/sa_hip.h declares (no compute calls here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(sa_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_and_python_binding_agree():
    import simulated_annealing_amd as sa

    assert declared_functions() == set(sa.EXPORTS)


def test_library_exports_every_declared_symbol():
    import simulated_annealing_amd as sa

    if not os.path.exists(sa.LIB_PATH):
        pytest.skip("libsa_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", sa.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared_functions() - exported
    assert not missing, missing
    L = sa.lib()                      # loads without a GPU
    for name in declared_functions():
        assert hasattr(L, name)
    assert L.sa_version() == 1
    assert L.sa_status_string(-1) == b"invalid argument"


def test_product_does_not_touch_the_oracle():
    """The product package and its C++ sources never import or link the oracle."""
    pkg = os.path.join(ROOT, "simulated_annealing_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text and "oracle/" not in text, f
    for f in os.listdir(os.path.join(ROOT, "include", "sa")):
        text = open(os.path.join(ROOT, "include", "sa", f)).read()
        assert "oracle" not in text, f
