"""Drop-in entry points: the ``Tailor.py`` launcher and the ``tailor`` solver.

* :func:`parse_config` reads the YAML case file with the schema and messages
  of Parser.cpp:36-145 and resolves the cache path
  ``<source dir>/FD-q/fdq_<stem>_qy<q_y>_qz<q_z>.h5`` (Parser.cpp:53-62).
* :func:`load_cache` reads the schema-v2 FD-q cache with the checks of
  Prepare.cpp:443-668 (rule contents are validated again by
  ``tailor_load_case`` with the messages of Prepare.cpp:328-437).
* :func:`solver_main` is the ``tailor`` executable (main.cpp, Process.cpp):
  ``-c <file>`` (required), ``-tailor_assemble_only``, and the PETSc/SLEPc
  options ``-eps_nev``, ``-eps_tol``, ``-eps_max_it``, ``-eps_ncv``; other
  ``-eps_*``/``-st_*``/``-ksp_*``/``-pc_*``/``-mat_*`` options name solver
  packages this build replaces and are accepted and ignored.  It prints the
  diagnostic lines of Process.cpp:62-176 and writes the schema-v1 eigenmode
  file of EigenOutput.cpp:85-304 (temporary file, validation, rename).
* :func:`main` is ``Tailor.py`` (main.py:35-332): same arguments, the FD-q
  preprocessing step (reuse a valid cache, else build it from the source
  flow with :mod:`tailor_amd.prepare`), then the solver as a child process;
  a preprocessing failure's status is propagated and a launch error is 2.

Everything runs the HIP library; there is no CPU path.
"""
from __future__ import annotations

import argparse
import hashlib
import math
import os
import subprocess
import sys
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from . import fdq, h5
from .case import Case, Physics

PACKAGE_ROOT = Path(__file__).resolve().parent
PROJECT_ROOT = PACKAGE_ROOT.parent

# exit statuses of the solver stage (the C ABI's status classes)
EXIT_USER_INPUT = 7
EXIT_FILE_UNEXPECTED = 1


class InputError(RuntimeError):
    """A configuration or cache error, with the reference's message text."""

    def __init__(self, message: str, status: int = EXIT_USER_INPUT):
        super().__init__(message)
        self.status = status


# --------------------------------------------------------------------------
# configuration (Parser.cpp)
# --------------------------------------------------------------------------

@dataclass
class Recipe:
    case_title: str
    q_y: int
    q_z: int
    source_file: Path
    input_file: Path
    alpha: complex
    search_center_omega: complex
    nev: int
    tol: float
    maxit: int
    output_file: Path
    physics: Physics
    config_path: Path

    @property
    def search_center_lambda(self) -> complex:
        """lambda_s = Im(omega_s) - i Re(omega_s) (EigenSolver.cpp:46-48)."""
        w = self.search_center_omega
        return complex(w.imag, -w.real)


def _required(node: Any, key: str, kind: type, label: str | None = None):
    if not isinstance(node, dict) or key not in node or node[key] is None:
        raise InputError(f"Missing YAML key: {label or key}")
    value = node[key]
    if kind is int:
        if isinstance(value, bool) or not isinstance(value, int):
            raise InputError(f"YAML key {label or key} must be an integer")
        return int(value)
    if kind is float:
        if isinstance(value, bool) or not isinstance(value, (int, float)):
            raise InputError(f"YAML key {label or key} must be a number")
        return float(value)
    if not isinstance(value, (str, int, float)) or isinstance(value, bool):
        raise InputError(f"YAML key {label or key} must be a string")
    return str(value)


def _mapping(node: Any, key: str, label: str) -> dict:
    value = node.get(key) if isinstance(node, dict) else None
    if not isinstance(value, dict):
        raise InputError(f"Missing or invalid YAML mapping: {label}")
    return value


def cache_path(config_path: Path, folder: str, file: str, q_y: int, q_z: int) -> Path:
    """``<source dir>/FD-q/fdq_<stem>_qy<q_y>_qz<q_z>.h5`` (Parser.cpp:53-62)."""
    source = Path(file)
    if not source.is_absolute():
        source = config_path.parent / folder / source
    source = Path(os.path.normpath(source))
    return source, source.parent / "FD-q" / f"fdq_{source.stem}_qy{q_y}_qz{q_z}.h5"


def parse_config(path: str | os.PathLike, require_cache: bool = True) -> Recipe:
    """Parser::parseOnRoot (Parser.cpp:36-145)."""
    import yaml

    config_path = Path(os.path.normpath(Path(path).expanduser().absolute()))
    if not config_path.is_file():
        raise InputError(f"bad file: {config_path}")
    with config_path.open("r", encoding="utf-8") as stream:
        try:
            root = yaml.safe_load(stream)
        except yaml.YAMLError as error:
            raise InputError(f"YAML parse error: {error}") from error
    if not isinstance(root, dict):
        raise InputError("the YAML document root must be a mapping")

    title = _required(root, "CaseTitle", str)
    q = _mapping(root, "Q-Value", "Q-Value")
    q_y, q_z = _required(q, "y", int), _required(q, "z", int)
    source, cache = cache_path(config_path, _required(root, "Folder", str),
                               _required(root, "File", str), q_y, q_z)

    stability = root.get("Stability")
    alpha_node = stability.get("Alpha") if isinstance(stability, dict) else None
    if alpha_node is None:
        raise InputError("Missing YAML key: Stability.Alpha")
    alpha = complex(_required(alpha_node, "Real", float), _required(alpha_node, "Imag", float))

    es = _mapping(root, "EigenSolver", "EigenSolver")
    sc = _mapping(es, "SearchCenterOmega", "EigenSolver.SearchCenterOmega")
    omega = complex(_required(sc, "Real", float), _required(sc, "Imag", float))
    nev = _required(es, "NumberOfEigenvalues", int)
    tol = _required(es, "Tolerance", float)
    maxit = _required(es, "MaximumIterations", int)

    out = _mapping(root, "Output", "Output")
    output = Path(_required(out, "File", str))
    if not output.is_absolute():
        output = config_path.parent / output
    output = Path(os.path.normpath(output))

    phys = root.get("Physics") if isinstance(root.get("Physics"), dict) else {}
    targets, gas, transport = (phys.get(k) for k in ("Targets", "Gas", "Transport"))
    physics = Physics(
        reynolds=_required(targets, "ReynoldsNumber", float),
        mach=_required(targets, "MachNumber", float),
        prandtl=_required(gas, "PrandtlNumber", float),
        gamma=_required(gas, "RatioOfSpecificHeats", float),
        gas_constant=_required(gas, "GasConstant", float),
        reference_mu=_required(transport, "ReferenceMu", float),
        reference_temperature=_required(transport, "ReferenceTemperature", float),
        sutherland_constant=_required(transport, "SutherlandConstant", float),
    )

    if q_y < 2 or q_z < 2:
        raise InputError("Q-Value.y and Q-Value.z must both be at least 2")
    if require_cache and not cache.is_file():
        raise InputError(f"FD-q input file does not exist: {cache}. "
                         f"Run Tailor.py -c {config_path} first")
    for value, name in ((alpha.real, "Stability.Alpha.Real"), (alpha.imag, "Stability.Alpha.Imag"),
                        (omega.real, "EigenSolver.SearchCenterOmega.Real"),
                        (omega.imag, "EigenSolver.SearchCenterOmega.Imag"),
                        (tol, "EigenSolver.Tolerance"), (physics.reynolds, "ReynoldsNumber"),
                        (physics.mach, "MachNumber"), (physics.prandtl, "PrandtlNumber"),
                        (physics.gas_constant, "GasConstant"),
                        (physics.gamma, "RatioOfSpecificHeats"),
                        (physics.reference_mu, "ReferenceMu"),
                        (physics.reference_temperature, "ReferenceTemperature"),
                        (physics.sutherland_constant, "SutherlandConstant")):
        if not math.isfinite(value):
            raise InputError(f"{name} must be finite")
    p = physics
    if (p.reynolds <= 0 or p.mach <= 0 or p.prandtl <= 0 or p.gas_constant <= 0
            or p.gamma <= 1 or p.reference_mu <= 0 or p.reference_temperature <= 0
            or p.sutherland_constant < 0):
        raise InputError("The physical parameters in config.yaml are outside their valid ranges")
    if nev <= 0 or tol <= 0 or maxit <= 0:
        raise InputError("EigenSolver counts, tolerance, and iteration limit must be positive")
    return Recipe(title, q_y, q_z, source, cache, alpha, omega, nev, tol, maxit, output,
                  physics, config_path)


# --------------------------------------------------------------------------
# FD-q cache (Prepare.cpp)
# --------------------------------------------------------------------------

def _dataset(root: h5.Group, path: str, shape: tuple[int, ...]) -> np.ndarray:
    try:
        node = root[path]
    except KeyError:
        raise InputError(f"Missing HDF5 dataset: {path}", EXIT_FILE_UNEXPECTED) from None
    if not isinstance(node, h5.Dataset):
        raise InputError(f"Missing HDF5 dataset: {path}", EXIT_FILE_UNEXPECTED)
    if node.data.ndim != len(shape):
        raise InputError(f"Unexpected rank for HDF5 dataset: {path}", EXIT_FILE_UNEXPECTED)
    if tuple(node.data.shape) != tuple(shape):
        raise InputError(f"Unexpected shape for HDF5 dataset: {path}", EXIT_FILE_UNEXPECTED)
    return node.data


def _attr(root: h5.Group, obj: str, name: str):
    node = root[obj] if obj != "/" else root
    if name not in node.attrs:
        raise InputError(f"Missing HDF5 attribute: {obj}/{name}", EXIT_FILE_UNEXPECTED)
    return node.attrs[name]


def _rule(root: h5.Group, group: str, n: int, q: int, topology: str,
          period: float) -> fdq.FDQRule:
    prefix = f"/discretization/{group}"
    st = (n, q + 1)
    nodes = _dataset(root, prefix + "/nodes", (n,)).astype(np.float64)
    idx = _dataset(root, prefix + "/stencil_indices", st).astype(np.int64)
    off = _dataset(root, prefix + "/stencil_offsets", st).astype(np.int64)
    w = np.stack([_dataset(root, f"{prefix}/weights/d{m}", st).astype(np.float64)
                  for m in range(3)])
    return fdq.FDQRule(nodes, q, idx, off, w, topology, period)


def load_cache(recipe: Recipe) -> Case:
    """Prepare::initialize (Prepare.cpp:443-687): the cache as a :class:`Case`."""
    try:
        root = h5.read(recipe.input_file)
    except (OSError, h5.H5Error) as error:
        raise InputError(f"Failed to read FD-q metadata: {error}", EXIT_FILE_UNEXPECTED) from error
    try:
        ny, nz = int(_attr(root, "/", "Ny")), int(_attr(root, "/", "Nz"))
        version = int(_attr(root, "/", "schema_version"))
        qy, qz = int(_attr(root, "/", "q_y")), int(_attr(root, "/", "q_z"))
        periodic = int(_attr(root, "/", "spanwise_periodic"))
        period = float(_attr(root, "/", "spanwise_period"))
        ordering = str(_attr(root, "/", "ordering"))
    except InputError as error:
        raise InputError(f"Failed to read FD-q metadata: {error}", EXIT_FILE_UNEXPECTED) from None
    fu = EXIT_FILE_UNEXPECTED
    if ny <= 0 or nz <= 0:
        raise InputError("HDF5 attributes Ny and Nz must be positive", fu)
    if version != 2:
        raise InputError(f"Unsupported FD-q schema version {version}", fu)
    if ordering != "k_j_dof":
        raise InputError("FD-q ordering must be k_j_dof", fu)
    if periodic != 1:
        raise InputError("FD-q input must declare spanwise_periodic=1", fu)
    if not (math.isfinite(period) and period > 0):
        raise InputError("FD-q spanwise_period must be finite and positive", fu)
    if qy != recipe.q_y or qz != recipe.q_z:
        raise InputError("FD-q q_y/q_z attributes do not match the YAML configuration", fu)

    n = ny * nz
    if "baseflow" not in root:
        raise InputError("HDF5 dataset /baseflow is missing", fu)
    if "grid" not in root:
        raise InputError("HDF5 dataset /grid is missing", fu)
    base, grid = root["baseflow"].data, root["grid"].data
    if base.size != 10 * n:
        raise InputError(f"Unexpected baseflow size: expected {5 * n}, got {base.size // 2}", fu)
    if grid.size != 6 * n:
        raise InputError(f"Unexpected grid size: expected {3 * n}, got {grid.size // 2}", fu)
    base = base.reshape(n, 5, 2).astype(np.float64)
    grid = grid.reshape(n, 3, 2).astype(np.float64)

    try:
        y_n, y_q = int(_attr(root, "/discretization/y", "node_count")), int(
            _attr(root, "/discretization/y", "q"))
        z_n, z_q = int(_attr(root, "/discretization/z", "node_count")), int(
            _attr(root, "/discretization/z", "q"))
        y_period = float(_attr(root, "/discretization/y", "period"))
        z_period = float(_attr(root, "/discretization/z", "period"))
    except KeyError as error:
        raise InputError(f"Missing HDF5 object: {error.args[0]}", fu) from None
    if y_n != ny or y_q != recipe.q_y:
        raise InputError("The /discretization/y metadata is inconsistent", fu)
    if z_n != nz or z_q != recipe.q_z:
        raise InputError("The /discretization/z metadata is inconsistent", fu)
    topo = (str(_attr(root, "/discretization/y", "topology")),
            str(_attr(root, "/discretization/z", "topology")))
    if topo != ("bounded", "periodic"):
        raise InputError("Expected bounded y and periodic z FD-q topologies", fu)
    try:
        xi = _rule(root, "y", ny, y_q, "bounded", y_period)
        eta = _rule(root, "z", nz, z_q, "periodic", z_period)
    except InputError as error:
        raise InputError(f"Failed to load FD-q discretization: {error}", fu) from None

    return Case(ny, nz, xi, eta, grid[..., 0] + 1j * grid[..., 1],
                base[..., 0] + 1j * base[..., 1], period, recipe.physics, recipe.alpha,
                recipe.search_center_omega, recipe.nev, recipe.tol, recipe.maxit,
                recipe.case_title)


def write_cache(case: Case, path: str | os.PathLike, source_file: str = "synthetic",
                source_sha256: str | None = None, extra_attrs: dict | None = None) -> Path:
    """Write ``case`` as a schema-v2 FD-q cache (the layout Prepare.cpp:443-668
    reads and interpolate.py:399-436 writes): root attributes, ``/grid`` and
    ``/baseflow`` in natural ``k_j`` order, ``/discretization/{y,z}``."""
    root = h5.Group("/", attrs={
        "Ny": case.ny, "Nz": case.nz, "schema_version": 2, "q_y": case.xi_rule.q,
        "q_z": case.eta_rule.q, "spanwise_periodic": 1,
        "spanwise_period": float(case.spanwise_period), "ordering": "k_j_dof",
        "source_file": source_file})
    if source_sha256 is not None:
        root.attrs["source_sha256"] = source_sha256
    root.attrs.update(extra_attrs or {})
    root.members["grid"] = h5.complex_dataset("grid", case.grid)
    root.members["baseflow"] = h5.complex_dataset("baseflow", case.baseflow)
    disc = h5.Group("discretization")
    for name, rule, n in (("y", case.xi_rule, case.ny), ("z", case.eta_rule, case.nz)):
        g = h5.Group(name, attrs={"N": n - 1 if rule.topology == "bounded" else n,
                                  "q": rule.q, "node_count": n, "topology": rule.topology,
                                  "period": float(rule.period)})
        g.members["nodes"] = h5.Dataset("nodes", np.asarray(rule.nodes, dtype=np.float64))
        g.members["stencil_indices"] = h5.Dataset(
            "stencil_indices", np.asarray(rule.stencil_indices, dtype=np.int64))
        g.members["stencil_offsets"] = h5.Dataset(
            "stencil_offsets", np.asarray(rule.stencil_offsets, dtype=np.int64))
        w = h5.Group("weights")
        for m in range(3):
            w.members[f"d{m}"] = h5.Dataset(f"d{m}", np.asarray(rule.weights[m], dtype=np.float64))
        g.members["weights"] = w
        disc.members[name] = g
    root.members["discretization"] = disc
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    h5.write(out, root)
    return out


def write_config(path: str | os.PathLike, case: Case, source: str = "case.h5",
                 folder: str = "data", output: str = "results/eigenmodes.h5") -> Path:
    """A YAML case file for ``case`` in the schema of Parser.cpp:41-108."""
    import yaml

    ph = case.physics
    doc = {
        "CaseTitle": case.title, "Folder": folder, "File": source,
        "Q-Value": {"y": int(case.xi_rule.q), "z": int(case.eta_rule.q)},
        "Stability": {"Alpha": {"Real": case.alpha.real, "Imag": case.alpha.imag}},
        "EigenSolver": {"SearchCenterOmega": {"Real": case.omega_shift.real,
                                              "Imag": case.omega_shift.imag},
                        "NumberOfEigenvalues": int(case.nev), "Tolerance": float(case.tol),
                        "MaximumIterations": int(case.maxit)},
        "Output": {"File": output},
        "Physics": {"Targets": {"ReynoldsNumber": ph.reynolds, "MachNumber": ph.mach},
                    "Gas": {"PrandtlNumber": ph.prandtl, "GasConstant": ph.gas_constant,
                            "RatioOfSpecificHeats": ph.gamma},
                    "Transport": {"ReferenceMu": ph.reference_mu,
                                  "ReferenceTemperature": ph.reference_temperature,
                                  "SutherlandConstant": ph.sutherland_constant}},
    }
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(yaml.safe_dump(doc, sort_keys=False), encoding="utf-8")
    return out


# --------------------------------------------------------------------------
# output (EigenOutput.cpp)
# --------------------------------------------------------------------------

def write_output(recipe: Recipe, case: Case, pairs: list[dict], requested: int) -> Path:
    """EigenOutput::write (EigenOutput.cpp:177-304): schema-v1 file written to
    ``<output>.tmp``, validated, then renamed over ``<output>``.  ``pairs``
    carry unit-norm, phase-fixed modes in natural ``k_j_dof`` order."""
    if not pairs:
        raise InputError("No converged modes are available for output", 6)
    out = Path(recipe.output_file)
    try:
        out.parent.mkdir(parents=True, exist_ok=True)
    except OSError:
        raise InputError(f"Could not create the output directory for {out}", 1) from None
    lam_s = recipe.search_center_lambda
    root = h5.Group("/", attrs={
        "schema_version": 1, "case": recipe.case_title, "ordering": "k_j_dof",
        "mode_dof": "[rho,u,v,w,T]",
        "lambda_convention": "lambda=-i*omega; A_bc*q=lambda*B_bc*q",
        "xi_boundary": "isothermal_no_slip_wall|characteristic_farfield",
        "eta_boundary": "periodic", "Ny": case.ny, "Nz": case.nz,
        "q_y": recipe.q_y, "q_z": recipe.q_z,
        "alpha_real": recipe.alpha.real, "alpha_imag": recipe.alpha.imag,
        "search_center_omega_real": recipe.search_center_omega.real,
        "search_center_omega_imag": recipe.search_center_omega.imag,
        "search_center_lambda_real": lam_s.real, "search_center_lambda_imag": lam_s.imag,
        "spanwise_period": float(case.spanwise_period),
        "requested_modes": int(requested), "converged_modes": len(pairs),
    })
    root.members["grid"] = h5.complex_dataset("grid", case.grid)
    root.members["baseflow"] = h5.complex_dataset("baseflow", case.baseflow)
    lam = np.array([p["lambda"] for p in pairs])
    spec = h5.Group("spectrum")
    spec.members["lambda"] = h5.complex_dataset("lambda", lam)
    spec.members["omega"] = h5.complex_dataset("omega", 1j * lam)
    spec.members["residual"] = h5.complex_dataset(
        "residual", np.array([p["relative_error"] for p in pairs], dtype=np.complex128))
    root.members["spectrum"] = spec
    modes = h5.Group("modes")
    for k, p in enumerate(pairs):
        name = f"mode_{k:03d}"
        modes.members[name] = h5.complex_dataset(name, np.asarray(p["mode"]).reshape(-1, 5))
    root.members["modes"] = modes

    tmp = Path(f"{out}.tmp")
    h5.write(tmp, root)
    ok = False
    try:
        back = h5.read(tmp)
        ok = all(p in back for p in ("/grid", "/baseflow", "/spectrum/lambda",
                                     "/spectrum/omega", "/spectrum/residual"))
        ok = ok and "/modes/mode_000" in back
        ok = ok and "lambda_convention" in back.attrs and "converged_modes" in back.attrs
        if ok:
            os.replace(tmp, out)
    except (OSError, h5.H5Error):
        ok = False
    if not ok:
        raise InputError(f"HDF5 validation or atomic replacement failed for {out}", 1)
    return out


# --------------------------------------------------------------------------
# the ``tailor`` executable (main.cpp + Process.cpp)
# --------------------------------------------------------------------------

_IGNORED_PREFIXES = ("-eps_", "-st_", "-ksp_", "-pc_", "-mat_", "-log_", "-options_",
                     "-malloc", "-info", "-help", "-bv_", "-ds_", "-rg_")


def _solver_options(argv: Sequence[str]) -> dict:
    opts: dict[str, str | None] = {}
    i = 0
    while i < len(argv):
        arg = argv[i]
        if not arg.startswith("-"):
            raise InputError(f"Unexpected argument {arg!r}")
        value = None
        if i + 1 < len(argv) and (not argv[i + 1].startswith("-") or _is_number(argv[i + 1])):
            value = argv[i + 1]
            i += 1
        opts[arg] = value
        i += 1
    return opts


def _is_number(s: str) -> bool:
    try:
        float(s)
        return True
    except ValueError:
        return False


def _print_error(message: str) -> None:
    print(f"[0]PETSC ERROR: {message}", file=sys.stderr, flush=True)


def solver_main(argv: Sequence[str] | None = None) -> int:
    """The ``tailor`` solver: Process::run (Process.cpp:34-176)."""
    argv = list(sys.argv[1:] if argv is None else argv)
    try:
        opts = _solver_options(argv)
        config = opts.get("-c")
        if not config:
            raise InputError("Required option -c <file> was not provided")
        recipe = parse_config(config)
        nev = int(opts.get("-eps_nev") or recipe.nev)
        tol = float(opts.get("-eps_tol") or recipe.tol)
        maxit = int(opts.get("-eps_max_it") or recipe.maxit)
        ncv = int(opts.get("-eps_ncv") or 0)
        for key in opts:
            if key not in ("-c", "-tailor_assemble_only", "-eps_nev", "-eps_tol",
                           "-eps_max_it", "-eps_ncv") and not key.startswith(_IGNORED_PREFIXES):
                print(f"WARNING! There are options you set that were not used!\n"
                      f"WARNING! could be spelling mistake, etc!\nOption left: name:{key}",
                      file=sys.stderr)
        case = load_cache(recipe)
    except InputError as error:
        _print_error(str(error))
        return error.status
    except (ValueError, TypeError) as error:
        _print_error(str(error))
        return EXIT_USER_INPUT

    import tailor_amd as ta

    try:
        with ta.Context(0) as ctx:
            ctx.load_case(case)
            diag = ctx.assemble(diagnostics=True)
            report = ta.format_assembly_report(case, diag).split("\n")
            report.insert(1, f"  input: {recipe.input_file}")
            print("\n".join(report), flush=True)
            if "-tailor_assemble_only" in opts:
                print("  assemble-only: boundary matrices validated; EPSSolve skipped", flush=True)
                return 0
            # a solve that ends with fewer than nev pairs raises here, before
            # the diagnostics and before any output (EigenSolver.cpp:84-88)
            sd = ctx.solve(nev=nev, omega=recipe.search_center_omega, tol=tol, maxit=maxit,
                           ncv=ncv)
            lam_s = recipe.search_center_lambda
            print(f"  eigensolver: requested={sd['requested']}, converged={sd['converged']}, "
                  f"iterations={sd['iterations']}\n"
                  f"  lambda target: {lam_s.real:.16g} {lam_s.imag:+.16g}i; "
                  f"max relative error: {sd['max_relative_error']:.16g}", flush=True)
            if sd["converged"] <= 0:
                raise InputError("No converged modes are available for output", 6)
            pairs = ctx.eigenpairs(vectors=True)
            out = write_output(recipe, case, pairs, sd["requested"])
            print(f"  output: {out} ({len(pairs)} modes)", flush=True)
            return 0
    except InputError as error:
        _print_error(str(error))
        return error.status
    except ta.TailorError as error:
        _print_error(error.message)
        return error.status


# --------------------------------------------------------------------------
# the ``Tailor.py`` launcher (main.py)
# --------------------------------------------------------------------------

def _parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(
        prog="Tailor.py",
        description="Check the FD-q case cache, then run the tailor solver on the GPU. "
                    "Arguments after '--' are passed to the solver.")
    p.add_argument("-c", "--config", required=True, type=Path, help="YAML case configuration file")
    p.add_argument("--solver", type=Path,
                   help="solver executable (default: TAILOR_EXECUTABLE, else this "
                        "package's solver module)")
    p.add_argument("-n", "--mpi-processes", type=int, default=1, metavar="N",
                   help="number of solver processes (this build runs one process per GPU "
                        "and accepts 1)")
    p.add_argument("--mpiexec", type=Path, help="accepted for compatibility; unused")
    p.add_argument("--prepare-only", action="store_true",
                   help="check the FD-q cache without running the solver")
    p.add_argument("solver_arguments", nargs=argparse.REMAINDER,
                   help="solver arguments following '--'")
    return p


def _sha256(path: Path) -> str:
    h = hashlib.sha256()
    with path.open("rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def prepare(config: Path) -> tuple[int, Recipe | None]:
    """The preprocessing stage (main.py:258-285 + interpolate.py:501-605).

    An existing cache is reused when it loads cleanly and, if the source
    file is present, was built from it (``source_sha256``); otherwise the
    cache is rebuilt from the source flow by :mod:`tailor_amd.prepare`.
    Without a source file an existing valid cache is used as is."""
    from . import prepare as prep

    try:
        recipe = parse_config(config, require_cache=False)
    except InputError as error:
        print(f"Tailor FD-q preprocessing error: {error}", file=sys.stderr)
        return 1, None
    print("Tailor FD-q preprocessing\n"
          f"  config: {recipe.config_path}\n  source: {recipe.source_file}\n"
          f"  output: {recipe.input_file}\n  q_y: {recipe.q_y}\n  q_z: {recipe.q_z}", flush=True)
    reason = None
    if recipe.input_file.is_file():
        try:
            load_cache(recipe)
            want = h5.read(recipe.input_file).attrs.get("source_sha256")
            if (recipe.source_file.is_file() and want is not None
                    and _sha256(recipe.source_file) != want):
                reason = "source_sha256 differs from the source file"
        except InputError as error:
            reason = str(error)
        if reason is None:
            print(f"FD-q cache hit: {recipe.input_file}", flush=True)
    else:
        reason = "no cache"
    if reason is not None:
        if recipe.input_file.is_file():
            print(f"FD-q cache invalid, rebuilding: {reason}", file=sys.stderr)
        try:
            recipe.input_file.parent.mkdir(parents=True, exist_ok=True)
            prep.prepare_case(recipe.source_file, recipe.input_file, recipe.q_y, recipe.q_z)
            load_cache(recipe)
        except (OSError, ValueError, InputError, h5.H5Error) as error:
            print(f"Tailor FD-q preprocessing error: {error}", file=sys.stderr)
            return 1, None
    print(f"Tailor FD-q input ready: {recipe.input_file}", flush=True)
    return 0, recipe


def solver_command(solver: Path | None, config: Path, extra: Sequence[str] = ()) -> list[str]:
    forwarded = list(extra)
    if forwarded[:1] == ["--"]:
        forwarded.pop(0)
    base = [str(solver)] if solver is not None else [sys.executable, "-m", "tailor_amd.tailor"]
    return [*base, "-c", str(config), *forwarded]


def main(argv: Sequence[str] | None = None) -> int:
    """Tailor.py (main.py:232-332)."""
    parser = _parser()
    args = parser.parse_args(argv)
    if args.mpi_processes < 1:
        parser.error("--mpi-processes must be at least 1")
    status, recipe = prepare(args.config)
    if status != 0:
        print(f"Tailor preprocessing subprocess exited with status {status}", file=sys.stderr)
        return status
    if args.prepare_only:
        return 0
    try:
        if args.mpi_processes > 1:
            raise ValueError("this build runs the solver as one process on one GPU; "
                             "use -n 1")
        solver = args.solver or (Path(os.environ["TAILOR_EXECUTABLE"])
                                 if os.environ.get("TAILOR_EXECUTABLE") else None)
        if solver is not None:
            solver = Path(solver).expanduser().resolve()
            if not solver.is_file():
                raise FileNotFoundError(f"Tailor solver does not exist: {solver}")
            if not os.access(solver, os.X_OK):
                raise PermissionError(f"Tailor solver is not executable: {solver}")
        command = solver_command(solver, recipe.config_path, args.solver_arguments)
    except (FileNotFoundError, OSError, PermissionError, ValueError) as error:
        print(f"Tailor C++ launch error: {error}", file=sys.stderr)
        return 2
    print(f"Starting Tailor solver with 1 process: {command[0]}", flush=True)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join(filter(None, [str(PROJECT_ROOT), env.get("PYTHONPATH")]))
    completed = subprocess.run(command, cwd=PROJECT_ROOT, check=False, env=env)
    if completed.returncode != 0:
        print(f"Tailor C++ solver exited with status {completed.returncode}", file=sys.stderr)
    return completed.returncode
