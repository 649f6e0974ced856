"""CPU tests of the drop-in boundary: HDF5 files, YAML schema, launcher.

Mirrors the reference's own checks of these boundaries:
test_eigen_output.py (output schema), test_fdq_loader_failures.py (cache
corruptions and their messages), test_physics_config.py (YAML validation),
test_main.py (launcher commands and exit statuses).  The HDF5 module is
checked against the reference's shipped cache (read) and against libhdf5
through the conda interpreter's h5py when the image has it (both ways).
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
from pathlib import Path
from unittest import mock

import numpy as np
import pytest
import yaml

from conftest import small_case
from tailor_amd import cli, h5

GOLDEN = Path(__file__).resolve().parent / "golden"
REFERENCE_CACHE = Path("/root/reference/case/data/FD-q/fdq_sample_qy10_qz6.h5")
CONDA_PY = Path("/opt/conda/bin/python3.9")


def _has_libhdf5() -> bool:
    if not CONDA_PY.is_file():
        return False
    return subprocess.run([str(CONDA_PY), "-c", "import h5py"], capture_output=True).returncode == 0


def _tree() -> h5.Group:
    root = h5.Group("/", attrs={"schema_version": 1, "case": "Sample case", "alpha_real": 2.43,
                                "Ny": 201})
    root.members["grid"] = h5.complex_dataset("grid", np.arange(12).reshape(4, 3) * (1 + 2j))
    modes = h5.Group("modes")
    for k in range(21):                   # more than one symbol-table node
        modes.members[f"mode_{k:03d}"] = h5.complex_dataset(
            f"mode_{k:03d}", np.full((3, 5), k - 0.25j))
    root.members["modes"] = modes
    root.members["empty"] = h5.Group("empty", attrs={"topology": "periodic"})
    root.members["ints"] = h5.Dataset("ints", np.arange(-5, 6, dtype=np.int64).reshape(11, 1))
    return root


def test_h5_round_trip(tmp_path):
    h5.write(tmp_path / "t.h5", _tree())
    back = h5.read(tmp_path / "t.h5")
    assert back.attrs == {"schema_version": 1, "case": "Sample case", "alpha_real": 2.43, "Ny": 201}
    assert sorted(back["modes"].members) == [f"mode_{k:03d}" for k in range(21)]
    np.testing.assert_array_equal(h5.as_complex(back["modes/mode_017"]), np.full((3, 5), 17 - 0.25j))
    np.testing.assert_array_equal(h5.as_complex(back["grid"]), np.arange(12).reshape(4, 3) * (1 + 2j))
    assert back["modes/mode_003"].attrs == {"complex": 1}
    assert back["empty"].attrs == {"topology": "periodic"} and not back["empty"].members
    np.testing.assert_array_equal(back["ints"].data.ravel(), np.arange(-5, 6))


@pytest.mark.skipif(not _has_libhdf5(), reason="no libhdf5 binding in the image")
def test_h5_written_file_reads_in_libhdf5(tmp_path):
    h5.write(tmp_path / "t.h5", _tree())
    script = (
        "import h5py, json, numpy as np, sys\n"
        "f = h5py.File(sys.argv[1], 'r')\n"
        "m = f['modes/mode_017'][()]\n"
        "print(json.dumps({'attrs': {k: (v.decode() if isinstance(v, bytes) else v.item())"
        " for k, v in f.attrs.items()}, 'modes': sorted(f['modes']),"
        " 'm': [m[0, 0, 0], m[0, 0, 1]], 'complex': int(f['grid'].attrs['complex']),"
        " 'grid': f['grid'][()].tolist(), 'ints': f['ints'][()].ravel().tolist(),"
        " 'topology': f['empty'].attrs['topology'].decode()}))\n")
    out = subprocess.run([str(CONDA_PY), "-c", script, str(tmp_path / "t.h5")],
                         capture_output=True, text=True, check=True).stdout
    d = json.loads(out)
    assert d["attrs"] == {"schema_version": 1, "case": "Sample case", "alpha_real": 2.43, "Ny": 201}
    assert d["modes"] == [f"mode_{k:03d}" for k in range(21)]
    assert d["m"] == [17.0, -0.25] and d["complex"] == 1 and d["topology"] == "periodic"
    g = np.arange(12).reshape(4, 3) * (1 + 2j)
    np.testing.assert_array_equal(np.array(d["grid"]), np.stack([g.real, g.imag], -1))
    assert d["ints"] == list(range(-5, 6))


@pytest.mark.skipif(not REFERENCE_CACHE.is_file(), reason="reference case not present")
def test_h5_reads_the_reference_cache():
    """The shipped cache's rules equal the shipped rule files (fixtures)."""
    root = h5.read(REFERENCE_CACHE)
    assert root.attrs["schema_version"] == 2 and root.attrs["ordering"] == "k_j_dof"
    assert (root.attrs["Ny"], root.attrs["Nz"]) == (201, 50)
    for group, fixture in (("y", "fdq_N200_q10.npz"), ("z", "fdq_N50_q6p.npz")):
        ref = np.load(GOLDEN / fixture)
        for key in ("nodes", "stencil_indices", "stencil_offsets", "d0", "d1", "d2"):
            path = f"discretization/{group}/" + (f"weights/{key}" if key[0] == "d" else key)
            np.testing.assert_array_equal(root[path].data, ref[key])
    assert root["grid"].data.shape == (10050, 3, 2) and root["grid"].attrs["complex"] == 1
    assert root["baseflow"].data.shape == (10050, 5, 2)


@pytest.mark.skipif(not (REFERENCE_CACHE.is_file() and _has_libhdf5()),
                    reason="reference case or libhdf5 binding not present")
def test_h5_reference_cache_matches_libhdf5(tmp_path):
    script = ("import h5py, sys, numpy as np\nf = h5py.File(sys.argv[1], 'r')\n"
              "np.save(sys.argv[2], np.concatenate([f['grid'][()].ravel(),"
              " f['baseflow'][()].ravel()]))\n")
    subprocess.run([str(CONDA_PY), "-c", script, str(REFERENCE_CACHE), str(tmp_path / "r.npy")],
                   capture_output=True, check=True)
    ref = np.load(tmp_path / "r.npy")
    root = h5.read(REFERENCE_CACHE)
    mine = np.concatenate([root["grid"].data.ravel(), root["baseflow"].data.ravel()])
    np.testing.assert_array_equal(mine, ref)


# ---------------------------------------------------------------- YAML schema

def _case_dir(tmp_path, c=None):
    c = c or small_case()
    cfg = cli.write_config(tmp_path / "case" / "config.yaml", c)
    cli.write_cache(c, tmp_path / "case" / "data" / "FD-q" / "fdq_case_qy10_qz6.h5")
    return cfg, c


def test_config_schema_and_cache_path(tmp_path):
    cfg, c = _case_dir(tmp_path)
    r = cli.parse_config(cfg)
    assert r.case_title == c.title and (r.q_y, r.q_z) == (10, 6)
    assert r.input_file == tmp_path / "case" / "data" / "FD-q" / "fdq_case_qy10_qz6.h5"
    assert r.source_file == tmp_path / "case" / "data" / "case.h5"
    assert r.output_file == tmp_path / "case" / "results" / "eigenmodes.h5"
    assert r.alpha == c.alpha and r.nev == 8 and r.tol == 1e-10 and r.maxit == 1000
    assert r.search_center_lambda == complex(c.omega_shift.imag, -c.omega_shift.real)


def _edit(cfg: Path, path: tuple, value):
    doc = yaml.safe_load(cfg.read_text())
    node = doc
    for key in path[:-1]:
        node = node[key]
    if value is KeyError:
        del node[path[-1]]
    else:
        node[path[-1]] = value
    cfg.write_text(yaml.safe_dump(doc, sort_keys=False))


@pytest.mark.parametrize("path,value,message", [
    (("CaseTitle",), KeyError, "Missing YAML key: CaseTitle"),
    (("EigenSolver",), KeyError, "Missing or invalid YAML mapping: EigenSolver"),
    (("Stability", "Alpha"), KeyError, "Missing YAML key: Stability.Alpha"),
    (("Q-Value", "y"), 1, "Q-Value.y and Q-Value.z must both be at least 2"),
    (("Physics", "Targets", "ReynoldsNumber"), 0.0,
     "The physical parameters in config.yaml are outside their valid ranges"),
    (("Physics", "Gas", "RatioOfSpecificHeats"), 1.0,
     "The physical parameters in config.yaml are outside their valid ranges"),
    (("Physics", "Transport", "SutherlandConstant"), -1.0,
     "The physical parameters in config.yaml are outside their valid ranges"),
    (("Physics", "Targets", "MachNumber"), float("nan"), "MachNumber must be finite"),
    (("EigenSolver", "NumberOfEigenvalues"), 0,
     "EigenSolver counts, tolerance, and iteration limit must be positive"),
    (("EigenSolver", "Tolerance"), -1e-10,
     "EigenSolver counts, tolerance, and iteration limit must be positive"),
    (("File",), "other.h5", "FD-q input file does not exist"),
])
def test_config_validation_messages(tmp_path, path, value, message):
    cfg, _ = _case_dir(tmp_path)
    _edit(cfg, path, value)
    with pytest.raises(cli.InputError, match=message.replace(".", r"\.")):
        cli.parse_config(cfg)


def test_zero_lower_bound_sutherland_is_valid(tmp_path):
    cfg, _ = _case_dir(tmp_path)
    _edit(cfg, ("Physics", "Transport", "SutherlandConstant"), 0.0)
    assert cli.parse_config(cfg).physics.sutherland_constant == 0.0


# ---------------------------------------------------------------- FD-q cache

def test_cache_round_trip(tmp_path):
    cfg, c = _case_dir(tmp_path)
    got = cli.load_cache(cli.parse_config(cfg))
    np.testing.assert_array_equal(got.grid, c.grid)
    np.testing.assert_array_equal(got.baseflow, c.baseflow)
    for a, b in ((got.xi_rule, c.xi_rule), (got.eta_rule, c.eta_rule)):
        np.testing.assert_array_equal(a.nodes, b.nodes)
        np.testing.assert_array_equal(a.stencil_indices, b.stencil_indices)
        np.testing.assert_array_equal(a.stencil_offsets, b.stencil_offsets)
        np.testing.assert_array_equal(a.weights, b.weights)
        assert (a.q, a.topology, a.period) == (b.q, b.topology, b.period)
    assert got.spanwise_period == c.spanwise_period and got.physics == c.physics


def _mutate(cache: Path, fn):
    root = h5.read(cache)
    fn(root)
    h5.write(cache, root)


@pytest.mark.parametrize("name,fn,message", [
    ("missing_weight", lambda r: r["discretization/y/weights"].members.pop("d1"),
     "Missing HDF5 dataset"),
    ("wrong_q", lambda r: r.attrs.__setitem__("q_y", r.attrs["q_y"] - 1),
     "do not match the YAML configuration"),
    ("wrong_shape", lambda r: setattr(r["discretization/y/weights/d2"], "data",
                                      r["discretization/y/weights/d2"].data[:, :-1].copy()),
     "Unexpected shape for HDF5 dataset"),
    ("wrong_schema", lambda r: r.attrs.__setitem__("schema_version", 1),
     "Unsupported FD-q schema version 1"),
    ("wrong_ordering", lambda r: r.attrs.__setitem__("ordering", "dof_j_k"),
     "FD-q ordering must be k_j_dof"),
    ("missing_grid", lambda r: r.members.pop("grid"), "HDF5 dataset /grid is missing"),
    ("wrong_topology", lambda r: r["discretization/z"].attrs.__setitem__("topology", "bounded"),
     "Expected bounded y and periodic z FD-q topologies"),
])
def test_cache_corruptions_fail_with_the_reference_messages(tmp_path, name, fn, message):
    cfg, _ = _case_dir(tmp_path)
    recipe = cli.parse_config(cfg)
    _mutate(recipe.input_file, fn)
    with pytest.raises(cli.InputError, match=message) as err:
        cli.load_cache(recipe)
    assert err.value.status != 0, name


# ---------------------------------------------------------------- launcher

def test_prepare_only_stops_after_the_cache_check(tmp_path, capsys):
    cfg, _ = _case_dir(tmp_path)
    with mock.patch("subprocess.run") as run:
        assert cli.main(["-c", str(cfg), "--prepare-only"]) == 0
    run.assert_not_called()
    assert "Tailor FD-q input ready:" in capsys.readouterr().out


def test_launcher_runs_the_solver_with_passthrough(tmp_path):
    cfg, _ = _case_dir(tmp_path)
    with mock.patch("subprocess.run", return_value=subprocess.CompletedProcess([], 0)) as run:
        assert cli.main(["-c", str(cfg), "--", "-eps_nev", "4", "-tailor_assemble_only"]) == 0
    command = run.call_args.args[0]
    assert command[-5:] == ["-c", str(cfg), "-eps_nev", "4", "-tailor_assemble_only"]
    assert command[1:3] == ["-m", "tailor_amd.tailor"]


def test_launcher_returns_the_solver_status(tmp_path):
    cfg, _ = _case_dir(tmp_path)
    with mock.patch("subprocess.run", return_value=subprocess.CompletedProcess([], 9)):
        assert cli.main(["-c", str(cfg)]) == 9


def test_launcher_preprocessing_failure_does_not_start_the_solver(tmp_path):
    cfg, _ = _case_dir(tmp_path)
    os.remove(cli.parse_config(cfg).input_file)
    with mock.patch("subprocess.run") as run:
        assert cli.main(["-c", str(cfg)]) != 0
    run.assert_not_called()


def test_launcher_launch_errors_return_2(tmp_path):
    cfg, _ = _case_dir(tmp_path)
    with mock.patch("subprocess.run") as run:
        assert cli.main(["-c", str(cfg), "--solver", str(tmp_path / "missing")]) == 2
        assert cli.main(["-c", str(cfg), "-n", "2"]) == 2
    run.assert_not_called()
    with pytest.raises(SystemExit):
        cli.main(["-c", str(cfg), "-n", "0"])


def test_stale_cache_is_reported(tmp_path):
    cfg, c = _case_dir(tmp_path)
    recipe = cli.parse_config(cfg)
    recipe.source_file.write_bytes(b"source flow v2")
    cli.write_cache(c, recipe.input_file, source_sha256="0" * 64)
    assert cli.main(["-c", str(cfg), "--prepare-only"]) != 0
    import hashlib
    cli.write_cache(c, recipe.input_file,
                    source_sha256=hashlib.sha256(b"source flow v2").hexdigest())
    assert cli.main(["-c", str(cfg), "--prepare-only"]) == 0


def test_solver_requires_a_config(capsys):
    assert cli.solver_main([]) == cli.EXIT_USER_INPUT
    assert "Required option -c <file> was not provided" in capsys.readouterr().err


def test_solver_reports_a_missing_cache(tmp_path, capsys):
    cfg, _ = _case_dir(tmp_path)
    os.remove(cli.parse_config(cfg).input_file)
    assert cli.solver_main(["-c", str(cfg), "-tailor_assemble_only"]) != 0
    assert "FD-q input file does not exist" in capsys.readouterr().err


# ---------------------------------------------------------------- preprocessing

REFERENCE_SOURCE = Path("/root/reference/case/data/sample.h5")


def _write_source(path: Path, ny: int, nz: int, period: float):
    """A source flow the quintic tensor spline must reproduce exactly: every
    field is a polynomial of degree <= 5 in xi, constant in eta, except the
    grid's z = period (eta + 1) / 2, whose periodic part is constant."""
    xi = np.linspace(-1.0, 1.0, ny)
    eta = -1.0 + 2.0 * np.arange(nz) / nz
    X, E = np.meshgrid(xi, eta)                       # (nz, ny)
    y = 20.0 * (1.0 + X) + 0.3 * X ** 5 - 0.1 * X ** 3
    z = period * (E + 1.0) / 2.0
    grid = np.stack([np.zeros_like(X), y, z], axis=-1).astype(np.complex128)
    u = 0.5 * (1.0 + X) - 0.05 * X ** 4
    t = 1.5 - 0.2 * X ** 2 + 0.01 * X ** 5
    base = np.stack([1.0 / 1.5 + 0.1 * X ** 2, u, 0 * X, 0 * X, t], axis=-1).astype(np.complex128)
    base[..., 2] += 1e-3j * X ** 3                    # an imaginary part travels too
    root = h5.Group("/", attrs={"Ny": ny, "Nz": nz, "spanwise_periodic": 1,
                                "spanwise_period": period})
    root.members["grid"] = h5.complex_dataset("grid", grid)
    root.members["baseflow"] = h5.complex_dataset("baseflow", base)
    h5.write(path, root)
    return lambda xn, en: (np.meshgrid(xn, en))


def test_prepare_reproduces_quintic_fields_exactly(tmp_path):
    from tailor_amd import fdq, prepare as prep
    ny, nz, period = 24, 12, 5.0
    src = tmp_path / "data" / "flow.h5"
    src.parent.mkdir(parents=True)
    _write_source(src, ny, nz, period)
    nodes = -np.cos(np.pi * np.arange(ny) / (ny - 1))
    nodes[0], nodes[-1] = -1.0, 1.0
    with mock.patch.object(fdq, "load_bounded_nodes", lambda n, q: nodes):
        out = prep.prepare_case(src, tmp_path / "data" / "FD-q" / "fdq_flow_qy10_qz6.h5", 10, 6,
                                log=lambda *_: None)
    root = h5.read(out)
    X, E = np.meshgrid(nodes, -1.0 + 2.0 * np.arange(nz) / nz)
    grid = h5.as_complex(root["grid"]).reshape(nz, ny, 3)
    base = h5.as_complex(root["baseflow"]).reshape(nz, ny, 5)
    np.testing.assert_allclose(grid[..., 1].real, 20.0 * (1.0 + X) + 0.3 * X ** 5 - 0.1 * X ** 3,
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(grid[..., 2].real, period * (E + 1.0) / 2.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(base[..., 4].real, 1.5 - 0.2 * X ** 2 + 0.01 * X ** 5, atol=1e-13)
    np.testing.assert_allclose(base[..., 2].imag, 1e-3 * X ** 3, atol=1e-15)
    assert root.attrs["source_sha256"] == prep.sha256_file(src)
    assert root.attrs["interpolation_method"] == "bounded_y_periodic_z_quintic_bspline"
    assert root["discretization/y"].attrs["N"] == ny - 1


def test_prepare_rejects_bad_sources(tmp_path):
    from tailor_amd import prepare as prep
    src = tmp_path / "flow.h5"
    _write_source(src, 24, 12, 5.0)
    root = h5.read(src)
    root["baseflow"].data[0, 0, 4, 0] = -1.0
    h5.write(src, root)
    with pytest.raises(prep.PrepareError, match="sample temperature must be strictly positive"):
        prep.prepare_case(src, tmp_path / "out.h5", 10, 6)
    _write_source(src, 24, 12, 5.0)
    with pytest.raises(prep.PrepareError, match="periodic q_z must be even"):
        prep.prepare_case(src, tmp_path / "out.h5", 10, 5)
    with pytest.raises(FileNotFoundError, match="structured sample file does not exist"):
        prep.prepare_case(tmp_path / "missing.h5", tmp_path / "out.h5", 10, 6)


def test_launcher_builds_a_missing_cache_then_reuses_it(tmp_path, capsys):
    from tailor_amd import fdq
    c = small_case()
    cfg = cli.write_config(tmp_path / "case" / "config.yaml", c)
    src = tmp_path / "case" / "data" / "case.h5"
    src.parent.mkdir(parents=True)
    _write_source(src, 24, 12, 5.0)
    nodes = c.xi_rule.nodes
    with mock.patch.object(fdq, "load_bounded_nodes", lambda n, q: nodes):
        assert cli.main(["-c", str(cfg), "--prepare-only"]) == 0
        assert "FD-q cache generated" in capsys.readouterr().out
        assert cli.main(["-c", str(cfg), "--prepare-only"]) == 0
        assert "FD-q cache hit" in capsys.readouterr().out


@pytest.mark.skipif(not (REFERENCE_SOURCE.is_file() and REFERENCE_CACHE.is_file()),
                    reason="reference case not present")
def test_prepare_matches_the_shipped_cache(tmp_path):
    """The reference's own sample.h5 -> its shipped FD-q cache (the output of
    interpolate.py), with this package's N = 200 equal-error nodes."""
    from tailor_amd import prepare as prep
    out = prep.prepare_case(REFERENCE_SOURCE, tmp_path / "fdq.h5", 10, 6, log=lambda *_: None)
    mine, ref = h5.read(out), h5.read(REFERENCE_CACHE)
    for name in ("grid", "baseflow"):
        a, b = mine[name].data, ref[name].data
        assert a.shape == b.shape
        assert np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))) <= 1e-12, name
    for key in ("schema_version", "source_file", "source_sha256", "ordering", "Ny", "Nz",
                "q_y", "q_z", "spanwise_periodic", "spanwise_period", "interpolation_method",
                "interpolation_version"):
        assert mine.attrs[key] == ref.attrs[key], key
    np.testing.assert_allclose(mine["discretization/y/nodes"].data,
                               ref["discretization/y/nodes"].data, rtol=0, atol=1e-14)
