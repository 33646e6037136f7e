"""Checkpoint directories in the reference's format v3 (Solver::save /
Solver::restore, crates/lbm-core/src/solver.rs:4536-4830): manifest.json plus
one rank file per part.  The rank file (populations, the previous step's
post-collision populations, moments, solid flags, force field, scalar
distributions, FNV-1a 64 payload hash) is written and read by the library
(lbm_checkpoint_write_part / lbm_checkpoint_read_part); this module builds
the manifest and the scenario and decomposition hashes (solver.rs:883-965)
and checks them on restore, before any state is installed."""
from __future__ import annotations

import ctypes as C
import json
import os
import struct

import numpy as np

from . import _lib as _L

FORMAT_VERSION = 3
RANK_FILE = "rank_0000_part_0000.bin"
_OFFSET = 0xCBF29CE484222325
_PRIME = 0x100000001B3


class CheckpointError(RuntimeError):
    """Structured failure; `code` is one of the reference's CKPT_* codes."""

    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code


def fnv1a64(data: bytes, h: int = _OFFSET) -> int:
    for b in data:
        h ^= b
        h = (h * _PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def _u64(v: int) -> bytes:
    return struct.pack("<Q", v)


def _f64(v: float) -> bytes:
    return struct.pack("<d", v)


def _f32(v: float) -> bytes:
    return struct.pack("<f", v)


_FACE_KIND_IDS = {"closed": 0, "velocity": 1, "pressure": 2, "outflow": 3, "convective": 4}


def _face_bc_bytes(kind: str, value) -> bytes:
    """hash_face_bc (solver.rs:860-880): kind id, then its f32 parameters."""
    data = _u64(_FACE_KIND_IDS[kind])
    v = [0.0, 0.0, 0.0] if value is None else list(np.atleast_1d(np.asarray(value, np.float64)))
    if kind == "velocity":
        v = (v + [0.0, 0.0, 0.0])[:3]
        for x in v:
            data += _f32(float(x))
    elif kind in ("pressure", "convective"):
        data += _f32(float(v[0]))
    return data


def spec_hash(Q: int, dims, nu: float, collision: str, magic: float, omega_shear, periodic, force,
              sources=(), faces=None, patches=()) -> int:
    """spec_hash (solver.rs:883-945), fp32 run (reals hashed as f32).
    `sources` are the volume sources in order, each (lo, hi, q_lu) for a
    mass flow or (lo, hi, q_lu, u) for a jet (SourceKind ids 1 and 2);
    `faces` the six (kind, value) face rules in Face::ALL order (None: all
    closed); `patches` (face, lo, hi, kind, value) in order."""
    data = _u64(3) + _u64(Q)
    for v in dims:
        data += _u64(int(v))
    data += _f64(float(nu))
    if collision == "bgk":
        data += _u64(1)
    elif collision == "trt":
        data += _u64(2) + _f64(float(magic))
    else:
        data += _u64(3) + _f64(float(omega_shear))
    for p in periodic:
        data += _u64(int(bool(p)))
    for kind, value in (faces or [("closed", None)] * 6):
        data += _face_bc_bytes(kind, value)
    for v in force:
        data += _f32(float(v))
    for src in sources:
        for v in list(src[0]) + list(src[1]):
            data += _u64(int(v))
        if len(src) > 3:
            data += _u64(2) + _f32(float(src[2]))
            for v in src[3]:
                data += _f32(float(v))
        else:
            data += _u64(1) + _f32(float(src[2]))
    for face, lo, hi, kind, value in patches:
        data += _u64(int(face))
        for v in list(lo) + list(hi):
            data += _u64(int(v))
        data += _face_bc_bytes(kind, value)
    return fnv1a64(data)


def decomp_hash(dims, periodic) -> int:
    """decomp_hash (solver.rs:947-965) of the monolithic part
    (Subdomain::monolithic, subdomain.rs:31-45)."""
    data = _u64(1)
    for p in periodic:
        data += _u64(int(bool(p)))
    for _ in range(3):
        data += _u64(0)  # origin
    for v in dims:
        data += _u64(int(v))
    for axis in range(3):  # faces XNeg, XPos, YNeg, YPos, ZNeg, ZPos
        for _ in range(2):
            data += _u64(1 if periodic[axis] else 0)  # Some(0) -> 0 + 1, None -> 0
    return fnv1a64(data)


def _hs(h: int) -> str:
    return f"fnv1a64:{h:016x}"


def _lattice(Q: int) -> str:
    return {19: "D3Q19", 27: "D3Q27"}[Q]


def _solver_spec_hash(s) -> int:
    return spec_hash(s.Q, s.dims, s.nu, s.collision, s.magic, s.params.omega_shear, s.periodic, s.force,
                     getattr(s, "sources", ()), getattr(s, "faces", None), getattr(s, "patches", ()))


def save(solver, directory: str) -> dict:
    """Write `directory`/manifest.json and the rank file of `solver`."""
    os.makedirs(directory, exist_ok=True)
    info = _L.CkptInfo()
    path = os.path.join(directory, RANK_FILE)
    solver._chk(solver.lib.lbm_checkpoint_write_part(solver.h, path.encode(), C.byref(info)))
    manifest = {
        "kind": "lbmflow-checkpoint",
        "format_version": FORMAT_VERSION,
        "step": int(info.step),
        "time": float(info.step),
        "dtype": "f32",
        "lattice": _lattice(solver.Q),
        "global": list(solver.dims),
        "scenario_hash": _hs(_solver_spec_hash(solver)),
        "decomp_hash": _hs(decomp_hash(solver.dims, solver.periodic)),
        "nranks": 1,
        "ranks": [{"rank": 0, "part": 0, "file": RANK_FILE, "origin": [0, 0, 0], "core": list(solver.dims),
                   "bytes": int(info.bytes), "payload_hash": _hs(info.payload_hash),
                   "mask_hash": _hs(info.mask_hash)}],
        "reserved": {"particles": False, "rng": False, "stats": False},
    }
    with open(os.path.join(directory, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=2)
    return manifest


def restore(solver, directory: str) -> int:
    """Install the checkpoint in `directory` into `solver`, which must be
    built with the saved run's scenario, masks and rotating body; returns the
    restored step.  Mismatches raise CheckpointError before state changes."""
    with open(os.path.join(directory, "manifest.json")) as fh:
        m = json.load(fh)
    if m.get("kind") != "lbmflow-checkpoint":
        raise CheckpointError("CKPT_BAD_MAGIC", f"manifest kind is {m.get('kind')}")
    if m.get("format_version") != FORMAT_VERSION:
        raise CheckpointError("CKPT_VERSION_MISMATCH", f"format_version {m.get('format_version')}")
    if m.get("dtype") != "f32":
        raise CheckpointError("CKPT_DTYPE_MISMATCH", f"checkpoint dtype {m.get('dtype')} differs from run dtype f32")
    if m.get("lattice") != _lattice(solver.Q):
        raise CheckpointError("CKPT_LATTICE_MISMATCH", f"checkpoint lattice {m.get('lattice')}")
    if list(m.get("global", [])) != list(solver.dims):
        raise CheckpointError("CKPT_GEOM_MISMATCH", f"checkpoint global {m.get('global')} differs from {solver.dims}")
    cur = _hs(_solver_spec_hash(solver))
    if m.get("scenario_hash") != cur:
        raise CheckpointError("CKPT_SCENARIO_MISMATCH",
                              f"scenario_hash differed: checkpoint={m.get('scenario_hash')} current={cur}")
    if m.get("decomp_hash") != _hs(decomp_hash(solver.dims, solver.periodic)) or m.get("nranks") != 1:
        raise CheckpointError("CKPT_DECOMP_MISMATCH", "checkpoint decomposition differs from this run")
    rank = m["ranks"][0]
    info = _L.CkptInfo()
    # the library checks the current masks against mask_hash before it reads
    # the file, and the file's payload against payload_hash, before any state
    # is installed
    info.mask_hash = _parse_hash(rank.get("mask_hash"), "mask_hash")
    info.payload_hash = _parse_hash(rank.get("payload_hash"), "payload_hash")
    rc = solver.lib.lbm_checkpoint_read_part(solver.h, os.path.join(directory, rank["file"]).encode(),
                                             int(m["step"]), C.byref(info))
    if rc != 0:
        msg = solver.lib.lbm_last_error(solver.ctx.h)
        msg = msg.decode() if msg else ""
        code = msg.split(":", 1)[0] if msg.startswith("CKPT_") else "CKPT_IO"
        raise CheckpointError(code, msg)
    return int(m["step"])


def _parse_hash(s, what: str) -> int:
    if not isinstance(s, str) or not s.startswith("fnv1a64:") or len(s) != 24:
        raise CheckpointError("CKPT_MANIFEST_INVALID", f"{what} is not an fnv1a64 hash string: {s!r}")
    try:
        return int(s[8:], 16)
    except ValueError:
        raise CheckpointError("CKPT_MANIFEST_INVALID", f"{what} is not an fnv1a64 hash string: {s!r}") from None
