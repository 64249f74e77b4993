"""The C ABI from a compiled C++ caller (camellia2_amd/lib/dpg_capi_example,
built from camellia2_amd/csrc/capi_example.cpp with the library): its K, rhs
and Schur complement equal the Python binding's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from camellia2_amd import capi, dpg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EXE = os.path.join(os.path.dirname(capi.LIB_PATH), "dpg_capi_example")


def test_cpp_caller_matches_the_python_binding(tmp_path):
    assert os.path.exists(EXE), "build the library first (__graft_entry__.build())"
    nx, ny = 6, 5
    out = str(tmp_path / "k.bin")
    r = subprocess.run([EXE, str(nx), str(ny), out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 4)
    s = dpg.sizes(f, b)
    C = nx * ny
    raw = np.fromfile(out, dtype=np.float64)
    K = raw[:C * s.M * s.M].reshape(C, s.M, s.M)
    rhs = raw[C * s.M * s.M:C * s.M * (s.M + 1)].reshape(C, s.M)
    S = raw[C * s.M * (s.M + 1):].reshape(C, s.T, s.T)
    nodes, par = dpg.mesh_quad(nx, ny, 0.10, 0xC0FFEE + 1)
    load = dpg.load_sine_modes(2, 1, 8, 6, 0xC0FFEE + 1 + 100)
    A = dpg.LocalAssembler(f, [b])
    o, nbad = A.local_stiffness_and_rhs(0, torch.from_numpy(nodes).cuda(), torch.from_numpy(par).cuda(), None, load,
                                        want_schur=True)
    assert nbad == 0
    assert np.array_equal(o["K"].cpu().numpy(), K)
    assert np.array_equal(o["rhs"].cpu().numpy(), rhs)
    assert np.array_equal(o["schur"].cpu().numpy(), S)
    A.close()
