"""Trajectory analysis: DCD reading and the mean-square displacement (MSD)
with the semantics of the reference's script (py/Diffusion Coefficient/
MSD.py:15-82), computed on the GPU by selm_msd.

    time1, MSD = msd(0.005, 41, 1, 1, "1base.dcd")   # writes 1base.txt

The script plots as well; this module only computes and writes the table.
"""
import os
import struct

import numpy as np

from . import selm


def read_dcd(path):
    """Read a CHARMM/LAMMPS DCD with a unit cell in every frame.

    Returns a dict:
      icntrl   the 20 header control words;
      natoms;
      nsavc    steps between frames (the script's `freq`);
      delta    the timestep stored in the header;
      cells    [F][6] float64;
      frames   [F][3][natoms] float32, the x, y, z records of each frame.
    """
    b = open(path, "rb").read()
    if struct.unpack_from("<i", b, 0)[0] != 84 or b[4:8] != b"CORD":
        raise ValueError("%s: not a little-endian CORD DCD" % path)
    icntrl = list(struct.unpack_from("<20i", b, 8))
    delta = struct.unpack_from("<f", b, 8 + 36)[0]
    has_cell = icntrl[10] != 0
    off = 92
    n2 = struct.unpack_from("<i", b, off)[0]
    off += 4 + n2 + 4
    natoms = struct.unpack_from("<i", b, off + 4)[0]
    off += 12
    cell_len = 56 if has_cell else 0
    frame_len = cell_len + 3 * (8 + 4 * natoms)
    nfr = (len(b) - off) // frame_len
    cells = np.zeros((nfr, 6))
    frames = np.empty((nfr, 3, natoms), dtype=np.float32)
    for f in range(nfr):
        o = off + f * frame_len
        if has_cell:
            cells[f] = struct.unpack_from("<6d", b, o + 4)
            o += cell_len
        for d in range(3):
            frames[f, d] = np.frombuffer(b, dtype="<f4", count=natoms, offset=o + 4)
            o += 8 + 4 * natoms
    return {"icntrl": icntrl, "natoms": natoms, "nsavc": icntrl[2], "delta": delta, "cells": cells,
            "frames": frames}


def msd_frames(frames, box_x, nlags, mult=1):
    """MSD[dot] for dot = 0 .. nlags-1 (lag dot*mult) of frames [F][3][natoms]
    float32, with the script's conventions (origins from frame F//5, minimum
    image with half of box_x in every dimension), on the GPU."""
    frames = np.ascontiguousarray(frames, dtype=np.float32)
    nfr, three, natoms = frames.shape
    assert three == 3
    out = np.zeros(nlags)
    box = np.array([box_x, box_x, box_x], dtype=np.float64)
    L = selm.load()
    import ctypes as C
    rc = L.selm_msd(frames.ctypes.data_as(C.POINTER(C.c_float)), nfr, natoms,
                    box.ctypes.data_as(C.POINTER(C.c_double)), int(nlags), int(mult),
                    out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise selm.SelmError(rc, "selm_msd")
    return out


def msd(timestep, dotnum, mult, tau_t, dcd_filename, write_txt=True):
    """The script's msd(): returns (time1, MSD) for dot = 0 .. dotnum-1 with
    time1[dot] = dot*mult*timestep*freq/tau_t (freq = steps between frames)
    and writes "<base>.txt" with one "time MSD" pair per line."""
    d = read_dcd(dcd_filename)
    box_x = float(np.float32(d["cells"][0][0]))
    M = msd_frames(d["frames"], box_x, dotnum, mult)
    freq = d["nsavc"]
    time1 = [0] + [dot * mult * timestep * freq / tau_t for dot in range(1, dotnum)]
    MSD = [float(v) for v in M]
    MSD[0] = 0
    if write_txt:
        base = os.path.splitext(os.path.basename(dcd_filename))[0]
        with open(base + ".txt", "w") as f:
            for t, m in zip(time1, MSD):
                f.write(f"{t} {m}\n")
    return time1, MSD


def diffusivity(time1, MSD, t0, t1):
    """Self-diffusivity from the least-squares slope of MSD over t0 <= t <= t1
    (MSD = 6 D t in three dimensions)."""
    t = np.asarray(time1, dtype=np.float64)
    m = np.asarray(MSD, dtype=np.float64)
    sel = (t >= t0) & (t <= t1) & np.isfinite(m)
    slope = np.polyfit(t[sel], m[sel], 1)[0]
    return slope / 6.0
