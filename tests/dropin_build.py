"""Builds tests/dropin/pt_dropin (the C++ drop-in of the integration notes) with g++
against the in-tree libsa_hip.so: host C++ only, no GPU needed to compile."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "dropin", "pt_dropin.cpp")
EXE = os.path.join(ROOT, "tests", "dropin", "pt_dropin")
LIBDIR = os.path.join(ROOT, "simulated_annealing_amd")


def build(exe: str = EXE) -> str:
    cxx = shutil.which("g++") or "c++"
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lsa_hip", f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe
