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
