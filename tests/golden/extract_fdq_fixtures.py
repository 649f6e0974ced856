"""Copy the FD-q rule data that ship with the reference into small .npz fixtures.

The reference holds three rule caches as HDF5 data files (fdqNodes/N100_q10.h5,
N200_q10.h5 and N50_q6p.h5).  They are fixtures: nodes, stencil indices and
offsets, and the d0/d1/d2 weights.  This script reads them with h5py (which
executes nothing from the file) and writes plain numpy arrays, so that tests
run without h5py or the reference tree.

Run with an interpreter that has h5py, e.g.
    /opt/conda/bin/python3.9 tests/golden/extract_fdq_fixtures.py
"""
from pathlib import Path

import h5py
import numpy as np

SOURCE = Path("/root/reference/fdqNodes")
HERE = Path(__file__).resolve().parent

for name in ("N100_q10", "N200_q10", "N50_q6p"):
    with h5py.File(SOURCE / f"{name}.h5", "r") as handle:
        out = {
            "nodes": np.asarray(handle["nodes"], dtype=np.float64),
            "stencil_indices": np.asarray(handle["stencil_indices"], dtype=np.int64),
            "d0": np.asarray(handle["weights/d0"], dtype=np.float64),
            "d1": np.asarray(handle["weights/d1"], dtype=np.float64),
            "d2": np.asarray(handle["weights/d2"], dtype=np.float64),
            "log_error": np.asarray(handle["diagnostics/log_error"], dtype=np.float64),
            "q": np.int64(handle.attrs["q"]),
        }
        if "stencil_offsets" in handle:
            out["stencil_offsets"] = np.asarray(handle["stencil_offsets"], dtype=np.int64)
    np.savez(HERE / f"fdq_{name}.npz", **out)
    print(name, out["nodes"].shape)
