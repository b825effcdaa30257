"""Touchstone (version 1) export of S-parameters.

Replaces generate_touchstone (fem/physics/edm/touchstone.py:6-49) and
EMSimData.export_touchstone (fem/physics/edm/emdata.py:306-333).  The
reference writes through scikit-rf's Network.write_touchstone; this writes the
same file format directly: option line `# Hz S <RI|MA|DB> R 50`, one block per
frequency, 2-port data in the v1 column order S11 S21 S12 S22, other port
counts row by row (S11 S12 ... S1N, then S21 ...) with at most four pairs per
line, angles in degrees.
"""
from __future__ import annotations

import os
from typing import Literal

import numpy as np


def _pairs(S: np.ndarray, form: str):
    if form == "RI":
        return S.real, S.imag
    mag = np.abs(S)
    ang = np.degrees(np.angle(S))
    if form == "MA":
        return mag, ang
    with np.errstate(divide="ignore"):
        return 20 * np.log10(mag), ang


def generate_touchstone(filename: str, freq: np.ndarray, Smat: np.ndarray,
                        data_format: Literal["RI", "MA", "DB"], z0: float = 50.0) -> str:
    """Write Smat (M, N, N) at freq (M,) Hz; returns the file name written
    (extension .sNp added when `filename` has none)."""
    freq = np.asarray(freq, dtype=np.float64)
    Smat = np.asarray(Smat)
    if Smat.ndim != 3:
        raise ValueError(f"Smat must be 3-D with shape (M,N,N), got ndim={Smat.ndim}")
    M, N1, N2 = Smat.shape
    if N1 != N2:
        raise ValueError(f"Smat must be square in its last two dims, got {N1}x{N2}")
    if freq.ndim != 1 or freq.size != M:
        raise ValueError(f"freq must be 1-D of length {M}, got shape {freq.shape}")
    form = data_format.upper()
    if form not in ("RI", "MA", "DB"):
        raise ValueError(f"data_format must be RI, MA or DB, not {data_format!r}")
    base, ext = os.path.splitext(filename)
    out = base + (ext if ext else f".s{N1}p")
    A, B = _pairs(Smat, form)
    lines = ["! Touchstone file written by emerge_amd", f"# Hz S {form} R {z0:g}"]
    for m in range(M):
        if N1 == 2:
            order = [(0, 0), (1, 0), (0, 1), (1, 1)]
            vals = " ".join(f"{A[m, i, j]: .12e} {B[m, i, j]: .12e}" for i, j in order)
            lines.append(f"{freq[m]:.12e} {vals}")
            continue
        first = True
        for i in range(N1):
            for c0 in range(0, N1, 4):
                vals = " ".join(f"{A[m, i, j]: .12e} {B[m, i, j]: .12e}" for j in range(c0, min(N1, c0 + 4)))
                lines.append((f"{freq[m]:.12e} " if first else "  ") + vals)
                first = False
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return out


def read_touchstone(filename: str):
    """(freq Hz, S (M, N, N)) from a file written by generate_touchstone (any of RI/MA/DB)."""
    N = int(os.path.splitext(filename)[1][2:-1])
    form = "MA"
    scale = 1.0
    nums: list[float] = []
    with open(filename) as fh:
        for line in fh:
            line = line.split("!")[0].strip()
            if not line:
                continue
            if line.startswith("#"):
                tok = line[1:].upper().split()
                form = next(t for t in tok if t in ("RI", "MA", "DB"))
                scale = {"HZ": 1.0, "KHZ": 1e3, "MHZ": 1e6, "GHZ": 1e9}[tok[0]]
                continue
            nums.extend(float(t) for t in line.split())
    per = 1 + 2 * N * N
    data = np.array(nums).reshape(-1, per)
    freq = data[:, 0] * scale
    a, b = data[:, 1::2], data[:, 2::2]
    if form == "RI":
        v = a + 1j * b
    elif form == "MA":
        v = a * np.exp(1j * np.radians(b))
    else:
        v = 10 ** (a / 20) * np.exp(1j * np.radians(b))
    S = np.zeros((data.shape[0], N, N), dtype=np.complex128)
    if N == 2:
        for k, (i, j) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1)]):
            S[:, i, j] = v[:, k]
    else:
        S[:] = v.reshape(-1, N, N)
    return freq, S
