"""The 64x4 LDS-tile kernel, the path of isotropic levels beyond the row kernel's 32-bit byte
offsets (arrays of 4 GiB and more), stays parity-green: ALAMO_ISO_KERNEL=tiles routes every level
to it.  The selection is read when the library loads, so it runs in a child process, which checks
Fapply, the residual and one Fsmooth call against the oracle at 1e-12 / 1e-11 of the field scale,
on general and on Dirichlet x faces."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, %(root)r); sys.path.insert(0, %(root)r + "/tests"); sys.path.insert(0, %(root)r + "/oracle")
import numpy as np, torch
import oracle_ctypes as ora
from alamo_amd import problems as P
from alamo_amd.operator import Elastic, LAYOUT_AOS
from helpers import random_iso, random_bc, rng
dev = torch.device("cuda:0")
T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
H = lambda t: (torch.cuda.synchronize(), t.cpu().numpy())[1]
def close(a, b, tol):
    s = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * s, (np.abs(a - b).max(), s)
for ncell, xdir in (((128, 100, 80), False), ((130, 97, 83), True)):
    nodes = int(np.prod([n + 1 for n in ncell])); dx = tuple(1.0 / n for n in ncell)
    C = random_iso(nodes, 141)
    t, v = random_bc(142, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    if xdir:
        t[0, :] = P.DISPLACEMENT; t[3, :] = P.DISPLACEMENT
    op = Elastic(ncell, dx, sym=P.ISO, bc=(t, v)); op.SetModel(C, LAYOUT_AOS)
    u = rng(143).uniform(-1, 1, 3 * nodes); b = rng(144).uniform(-1, 1, 3 * nodes)
    out = T(np.zeros(3 * nodes))
    op.Fapply(0, 0, out, T(u)); close(H(out), ora.fapply(ncell, dx, P.ISO, C, t, u), 1e-12)
    op.correctionResidual(0, 0, out, T(u), T(b)); close(H(out), ora.residual(ncell, dx, P.ISO, C, t, u, b), 1e-12)
    D = ora.diagonal(ncell, dx, P.ISO, C, t)
    x = T(u); op.Fsmooth(0, 0, x, T(b))
    close(H(x), ora.jacobi(ncell, dx, P.ISO, C, t, D, 2.0 / 3.0, 2, u, b), 1e-11)
print("variant ok")
'''


@pytest.mark.parametrize("env", [{"ALAMO_ISO_KERNEL": "tiles"}])
def test_kernel_variant_parity(cuda, env):
    e = dict(os.environ)
    e.update(env)
    print("variant", env, flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD % {"root": ROOT}], env=e, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "variant ok" in r.stdout
