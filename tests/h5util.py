"""Reads datasets of the HDF5 files the dpa_hip CLI writes, through the
h5dump tool of the image's HDF5 (no h5py here)."""
import subprocess
import tempfile
from pathlib import Path

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"


def available() -> bool:
    return Path(H5DUMP).exists()


def read(path, dataset, dtype) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "x.bin"
        subprocess.run([H5DUMP, "-d", dataset, "-b", "LE", "-o", str(out), str(path)],
                       check=True, capture_output=True)
        return np.fromfile(out, dtype=dtype)


def datasets(path) -> list[str]:
    txt = subprocess.run([H5DUMP, "-n", str(path)], check=True, capture_output=True,
                         text=True).stdout
    return [l.split()[-1].lstrip("/") for l in txt.splitlines() if l.strip().startswith("dataset")]
