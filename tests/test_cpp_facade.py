"""The C++ façade (include/alamo_amd.hpp) compiles with the host compiler against the C ABI,
and on a GPU drives a multigrid solve and a Krylov solve whose results match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from alamo_amd import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_solve.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "build", "facade_solve")


def build_exe():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    libdir = os.path.join(ROOT, "alamo_amd", "lib")
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           SRC, "-o", EXE, "-L" + libdir, "-lalamo_amd", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + libdir,
           "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.check_call(cmd)


def test_facade_compiles():
    build_exe()
    assert os.path.exists(EXE)


@pytest.mark.gpu
def test_facade_solve_matches_oracle(cuda, ora, tmp_path):
    if not os.path.exists(EXE):
        build_exe()
    out, outk = tmp_path / "x.bin", tmp_path / "xk.bin"
    res = subprocess.run([EXE, str(out), str(outk)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and res.stdout.count("converged 1") == 2 and "net params" in res.stdout
    assert "fixed iterations 2" in res.stdout
    assert "max diff to linear 0.000e+00" in res.stdout and "ngrow 2 ncomp 3 singular 0" in res.stdout
    x = np.fromfile(out, dtype=np.float64)
    xk = np.fromfile(outk, dtype=np.float64)
    n = 16
    N = (n + 1) ** 3
    i = np.arange(N) % (n + 1)
    C = P.iso_field(1.0 + 0.5 * i / n, np.ones(N))
    bct, bcv = P.bc_constant(vals={"XHI": [0.01, 0, 0]})
    b = ora.bc_init((n, n, n), bct, bcv)
    xref, it, *_ = ora.mg_solve((n, n, n), (1.0 / n,) * 3, P.ISO, C, bct, b, tol_rel=1e-12)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) < 1e-9
    assert np.linalg.norm(xk - xref) / np.linalg.norm(xref) < 1e-9
