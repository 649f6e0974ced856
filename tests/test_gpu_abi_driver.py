"""The C ABI driven from C++ (tests/abi/drive_tailor.cpp: create, load_case,
assemble, solve, get_eigenpair, destroy, as a reference-side caller binds
it), against the same case through the ctypes mirror.

Without a GPU: include/tailor.h compiles as C99 and the driver compiles and
links against libtailor.so.  On the GPU: the driver's eigenpairs equal the
ctypes path's, with unit-norm modes whose largest entry is real and
positive, and an invalid case fails with the reference's message and a
nonzero status.
"""
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, small_case

sys.path.insert(0, str(ROOT / "tests" / "abi"))
import build_driver  # noqa: E402

DRIVER = build_driver.DRIVER


def write_case(path, c, nev=8, maxit=1000, tol=1e-10):
    xr, er = c.xi_rule, c.eta_rule
    ph = c.physics
    with open(path, "wb") as f:
        np.array([c.ny, c.nz, xr.stencil_size, er.stencil_size, nev, maxit], np.int32).tofile(f)
        np.array([xr.period, er.period, c.spanwise_period, ph.reynolds, ph.mach, ph.prandtl,
                  ph.gamma, ph.gas_constant, ph.reference_mu, ph.reference_temperature,
                  ph.sutherland_constant, c.alpha.real, c.alpha.imag, c.omega_shift.real,
                  c.omega_shift.imag, tol], np.float64).tofile(f)
        np.asarray(xr.nodes, np.float64).tofile(f)
        np.asarray(er.nodes, np.float64).tofile(f)
        for r in (xr, er):
            np.ascontiguousarray(r.stencil_indices, np.int64).tofile(f)
            np.ascontiguousarray(r.stencil_offsets, np.int64).tofile(f)
        np.ascontiguousarray(xr.weights, np.float64).tofile(f)
        np.ascontiguousarray(er.weights, np.float64).tofile(f)
        np.ascontiguousarray(c.grid, np.complex128).view(np.float64).tofile(f)
        np.ascontiguousarray(c.baseflow, np.complex128).view(np.float64).tofile(f)


def test_header_compiles_as_c_and_the_driver_links():
    assert build_driver.build() == DRIVER
    assert DRIVER.is_file()


@pytest.mark.gpu
def test_cpp_driver_matches_the_ctypes_path(tmp_path):
    import tailor_amd as ta
    assert DRIVER.is_file(), "tests/abi/drive_tailor is not built: run __graft_entry__.build()"
    c = small_case(24, 12)
    case = tmp_path / "case.bin"
    write_case(case, c)
    print("drive_tailor <24x12 synthetic case>", flush=True)
    r = subprocess.run([str(DRIVER), str(case)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    pairs = [l.split() for l in lines if l.startswith("pair ")]
    assert len(pairs) >= 8
    lam = np.array([complex(float(p[2]), float(p[3])) for p in pairs])
    res = np.array([float(p[4]) for p in pairs])
    rel = np.array([float(p[5]) for p in pairs])
    norm = np.array([float(p[6]) for p in pairs])
    piv = np.array([complex(float(p[7]), float(p[8])) for p in pairs])
    assert np.all(res[:8] <= 1e-10) and np.all(rel[:8] <= 1e-10)
    assert np.all(np.abs(norm - 1.0) <= 1e-13)
    assert np.all(np.abs(piv.imag) <= 1e-13) and np.all(piv.real > 0)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        diag = ctx.assemble(diagnostics=True)
        ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10, maxit=1000)
        ref = np.array([p["lambda"] for p in ctx.eigenpairs()])
    asm = lines[0].split()
    assert abs(float(asm[6]) - diag["a_fro"]) <= 1e-12 * diag["a_fro"]
    for l in ref[:8]:
        assert np.min(np.abs(lam - l)) <= 1e-10 * abs(l)


@pytest.mark.gpu
def test_cpp_driver_reports_the_reference_message(tmp_path):
    c = small_case(24, 12)
    c.baseflow = c.baseflow.copy()
    c.baseflow[5, 4] = -1.0  # negative temperature
    case = tmp_path / "case.bin"
    write_case(case, c)
    print("drive_tailor <24x12 case with a negative temperature>", flush=True)
    r = subprocess.run([str(DRIVER), str(case)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "density and temperature must both be positive" in r.stderr
