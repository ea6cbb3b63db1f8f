"""Write the fixtures taken from the reference's shipped demo output
(run_sim/expected_output.zip: the SmallSimulation input, 256 x 1 x 50 boxes,
BoxZAgar 32, two agar levels).

    python tests/golden/make_expected_golden.py [path/to/expected_output.zip]

Outputs (committed):
  * restart_5.npz, restart_7.npz, restart_8.npz: Restart/k.txt as t, dt and
    the raw 208-byte Cell records (129, 852 and 1952 cells);
  * expected_files.npz: the files AgarField/8.txt, WallField/8.txt,
    Environment/8.txt, Density/8.txt and lineage/0.txt, each stored
    lzma-compressed as bytes under its name, with its SHA-256, and the
    SHA-256 of each Restart/k.txt above.
Every fixture holds the reference's output data only.
"""
from __future__ import annotations

import hashlib
import lzma
import os
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ZIP = "/root/reference/AgentBasedModel-BacColonyGrowth/run_sim/expected_output.zip"
RESTARTS = (5, 7, 8)
FILES = ("AgarField/8.txt", "WallField/8.txt", "Environment/8.txt", "Density/8.txt", "lineage/0.txt")
CELL_BYTES = 208


def restart_arrays(raw: bytes):
    """(t, dt, cells as (N, 208) uint8) of a Restart file: 16-byte header,
    then whole Cell records."""
    if len(raw) < 16 or (len(raw) - 16) % CELL_BYTES:
        raise ValueError(f"not a restart file ({len(raw)} bytes)")
    t, dt = np.frombuffer(raw[:16], "<f8")
    cells = np.frombuffer(raw[16:], np.uint8).reshape(-1, CELL_BYTES).copy()
    return np.float64(t), np.float64(dt), cells


def key(name: str) -> str:
    return name.replace("/", "_").replace(".txt", "")


def main(zip_path: str) -> None:
    z = zipfile.ZipFile(zip_path)
    for k in RESTARTS:
        t, dt, cells = restart_arrays(z.read(f"expected_output/Restart/{k}.txt"))
        path = os.path.join(HERE, f"restart_{k}.npz")
        np.savez_compressed(path, t=t, dt=dt, cells=cells)
        print(path, len(cells), "cells")
    out = {f"Restart_{k}_sha256": np.array(hashlib.sha256(z.read(f"expected_output/Restart/{k}.txt")).hexdigest())
           for k in RESTARTS}
    for name in FILES:
        raw = z.read(f"expected_output/{name}")
        out[key(name)] = np.frombuffer(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME), np.uint8)
        out[key(name) + "_sha256"] = np.array(hashlib.sha256(raw).hexdigest())
    path = os.path.join(HERE, "expected_files.npz")
    np.savez(path, **out)
    print(path, os.path.getsize(path), "bytes")


def expected_file(g, name: str) -> bytes:
    """The original bytes of one stored file, checked against its digest."""
    raw = lzma.decompress(g[key(name)].tobytes())
    assert hashlib.sha256(raw).hexdigest() == str(g[key(name) + "_sha256"])
    return raw


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else ZIP)
