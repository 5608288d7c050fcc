"""Restart archives of a solver context.

Follows the reference's CheckpointArchive
(Sources/BirdFlowMetal/CheckpointArchive.swift): a directory holding
manifest.json (sorted keys) and the numerical state as compressed
files, written into a temporary sibling directory that is then moved into
place (:43-79), and read back with the format, the byte counts and the
SHA-256 of every file's contents checked (:81-118).  save_checkpoint and
load_checkpoint mirror BirdFlowSimulation.saveCheckpoint
(BirdFlowSimulation.swift:1214-1298) and init(checkpointURL:) with
restoreCheckpointState (:661-680, :1920-1990).

Where this archive differs from the reference's, and why:

* The manifest names its own format ("format": "birdflow-hip", "schema"
  counting that format's versions).  Its fields are not those of the
  reference's schema-1 BirdFlowCheckpointManifest (populationBytes,
  populationSHA256, ...; CheckpointArchive.swift:23-37), so a reader of
  either format refuses the other's archive instead of misreading it.

* zlib (Python's standard library) compresses instead of LZFSE, an Apple
  library.
* The populations are the stored form itself (FP32 values, or the FP16 /
  BF16 bits of the offsets f_q - w_q) in direction-major order, so a restart
  is bit-exact for every storage; the reference copies its buffer bytes in
  its own layout.
* The geometry source travels in the archive (mesh arrays, sphere,
  analytic-bird or flapping-wing parameters, or the static wall field); the reference's
  archive holds the bird parameters and its prepared frame, which a restore
  here prepares again from the body pose, and a mesh surface is caller
  data.
* The completed step index and the mesh clock origin are restored exactly
  (bf_set_restart_state), so the continued run sees the same mesh times.

A restored context continues bit-identically to the saved one for
prescribed geometry (tests/test_gpu_checkpoint.py).  Load and body histories
restart empty, as in the reference.
"""
from __future__ import annotations

import ctypes
import dataclasses
import hashlib
import json
import os
import shutil
import uuid
import zlib
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import _abi
from .simulation import ArticulatedMesh, BirdFlowSimulation, BirdGeometry, BodyParameters, FlappingWing, SolverConfig

FORMAT_NAME = "birdflow-hip"
SCHEMA_VERSION = 1
MANIFEST_NAME = "manifest.json"
POPULATIONS_NAME = "populations.bin"
MASK_NAME = "solid-mask.bin"
COMPRESSION_SUFFIX = {"zlib": ".z", "none": ""}
_STORAGE_ELEMENTS = {_abi.STORAGE_FP32: "float32", _abi.STORAGE_FP16_OFFSET: "float16-offset-bits",
                     _abi.STORAGE_BF16_OFFSET: "bfloat16-offset-bits"}
_GEOMETRY_ARRAYS = {
    "mesh": (("points", np.float32), ("frame_times", np.float32), ("triangles", np.uint32),
             ("triangle_parts", np.uint8), ("translation", np.float32)),
    "articulated": (("rest", np.float32), ("vertex_joint", np.uint8), ("span_fraction", np.float32),
                    ("triangles", np.uint32), ("triangle_parts", np.uint8)),
    "static": (("wall", np.float32), ("mask_prev", np.uint8)),
    "sphere": (),
    "bird": (("keyframes", np.float32),),
    "wing": (),
}


class CheckpointError(RuntimeError):
    """BirdFlowCheckpointError (CheckpointArchive.swift:6-21): an invalid
    archive or a checksum that does not match."""


@dataclass
class Checkpoint:
    manifest: dict
    populations: np.ndarray          # stored form, direction-major
    solid_mask: np.ndarray           # uint8 per cell, x fastest
    geometry: Dict[str, np.ndarray]


def sha256(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()


def _compress(data: bytes, compression: str) -> bytes:
    if compression == "none":
        return data
    return zlib.compress(data, 1)


def _decompress(blob: bytes, compression: str, expected_size: int, name: str) -> bytes:
    if compression == "none":
        data = blob
    else:
        try:
            d = zlib.decompressobj()
            data = d.decompress(blob, expected_size + 1)
            if not d.eof:
                raise CheckpointError(f"invalid archive: {name} is truncated or longer than {expected_size} bytes")
        except zlib.error as exc:
            raise CheckpointError(f"invalid archive: {name} does not decompress ({exc})") from None
    if len(data) != expected_size:
        raise CheckpointError(f"invalid archive: {name} holds {len(data)} bytes, the manifest says {expected_size}")
    return data


def _plain(v):
    if isinstance(v, (tuple, list)):
        return [_plain(x) for x in v]
    if isinstance(v, np.generic):
        return v.item()
    return v


def write_archive(destination: str, manifest: dict, files: Dict[str, bytes], compression: str = "zlib") -> None:
    """CheckpointArchive.write (CheckpointArchive.swift:43-79): every file
    into a temporary sibling directory, then moved over the destination.
    `files` maps base names to raw contents; their byte counts and SHA-256
    go into the manifest's `files` table."""
    if compression not in COMPRESSION_SUFFIX:
        raise ValueError(f"compression must be one of {sorted(COMPRESSION_SUFFIX)}")
    destination = os.path.abspath(destination)
    parent, base = os.path.split(destination)
    os.makedirs(parent, exist_ok=True)
    temporary = os.path.join(parent, f".{base}-{uuid.uuid4()}")
    os.makedirs(temporary)
    try:
        table = {}
        for name, data in files.items():
            stored = name + COMPRESSION_SUFFIX[compression]
            with open(os.path.join(temporary, stored), "wb") as fh:
                fh.write(_compress(data, compression))
            table[name] = {"file": stored, "bytes": len(data), "sha256": sha256(data)}
        full = dict(manifest, format=FORMAT_NAME, schema=SCHEMA_VERSION, compression=compression, files=table)
        with open(os.path.join(temporary, MANIFEST_NAME), "w") as fh:
            json.dump(full, fh, indent=2, sort_keys=True)
            fh.write("\n")
        if os.path.isdir(destination):
            shutil.rmtree(destination)
        elif os.path.exists(destination):
            os.remove(destination)
        os.replace(temporary, destination)
    except BaseException:
        shutil.rmtree(temporary, ignore_errors=True)
        raise


def read_archive(source: str):
    """CheckpointArchive.read (CheckpointArchive.swift:81-118): the manifest
    and every listed file, with the format, schema, byte counts and checksums
    checked.  Returns (manifest, {name: bytes})."""
    try:
        with open(os.path.join(source, MANIFEST_NAME)) as fh:
            manifest = json.load(fh)
    except (OSError, ValueError) as exc:
        raise CheckpointError(f"invalid archive: manifest unreadable ({exc})") from None
    if not isinstance(manifest, dict) or manifest.get("format") != FORMAT_NAME:
        raise CheckpointError(f"invalid archive: not a {FORMAT_NAME} archive")
    if manifest.get("schema") != SCHEMA_VERSION:
        raise CheckpointError("invalid archive: unsupported schema")
    compression = manifest.get("compression")
    table = manifest.get("files")
    if compression not in COMPRESSION_SUFFIX or not isinstance(table, dict):
        raise CheckpointError("invalid archive: unknown compression or no file table")
    out = {}
    for name, entry in sorted(table.items()):
        try:
            stored, size, digest = entry["file"], int(entry["bytes"]), str(entry["sha256"])
        except (KeyError, TypeError, ValueError):
            raise CheckpointError(f"invalid archive: malformed entry for {name}") from None
        if os.path.basename(stored) != stored or stored != name + COMPRESSION_SUFFIX[compression]:
            raise CheckpointError(f"invalid archive: unexpected file name {stored!r}")
        try:
            with open(os.path.join(source, stored), "rb") as fh:
                blob = fh.read()
        except OSError as exc:
            raise CheckpointError(f"invalid archive: {stored} unreadable ({exc})") from None
        data = _decompress(blob, compression, size, stored)
        if sha256(data) != digest:
            raise CheckpointError(f"checksum failed for {stored}")
        out[name] = data
    return manifest, out


def save_checkpoint(sim: BirdFlowSimulation, destination: str, compression: str = "zlib") -> dict:
    """BirdFlowSimulation.saveCheckpoint (BirdFlowSimulation.swift:1214-1298)
    at a completed step: the stored populations, the solid mask the next step
    starts from, the body state and last load, the step index and mesh clock
    origin, the configuration and the geometry source.  Returns the
    manifest."""
    if sim.geometry_source is None:
        raise CheckpointError("a checkpoint needs a geometry source (mesh, articulated mesh, sphere, bird, wing or static geometry)")
    cfg = sim.config
    step, origin = sim.restart_state()
    populations = sim.read_population_storage()
    src = sim.geometry_source
    kind = src["kind"]
    mask = np.empty(sim.n, np.uint8)
    wall = np.empty((sim.n, 4), np.float32) if kind == "static" else None
    sim._check(sim._lib.bf_read_geometry(sim._ctx, mask.ctypes.data_as(ctypes.c_void_p),
                                         None if wall is None else wall.ctypes.data_as(ctypes.c_void_p)))
    geometry = {"kind": kind}
    arrays = {}
    if kind == "mesh":
        geometry["half_thickness"] = float(src["half_thickness"])
        for name, _ in _GEOMETRY_ARRAYS["mesh"]:
            arrays[name] = src[name]
        geometry["shapes"] = {k: list(v.shape) for k, v in arrays.items()}
    elif kind == "articulated":
        # rest pose, joints and the kinematics: every step's vertices follow
        geometry["kinematics"] = src["kinematics"]
        for name, _ in _GEOMETRY_ARRAYS["articulated"]:
            arrays[name] = src[name]
        geometry["shapes"] = {k: list(v.shape) for k, v in arrays.items()}
    elif kind == "sphere":
        for name in ("center_radius", "geometry_velocity", "wall_velocity"):
            geometry[name] = _plain(src[name])
    elif kind == "wing":
        # prescribed: the parameters fix every step's frame, mask and links
        geometry["wing"] = {k: _plain(v) for k, v in src["wing"].items()}
    elif kind == "bird":
        # the shape; a restore prepares the frame again from the body pose
        # and keeps the archive's mask as the next step's previous mask
        geometry["bird"] = {k: _plain(v) for k, v in src["bird"].items()}
        if src.get("keyframes") is not None:
            arrays["keyframes"] = src["keyframes"]
            geometry["shapes"] = {"keyframes": list(src["keyframes"].shape)}
    else:
        arrays["wall"] = wall
        if src["mask_prev"] is not None and step == src["set_at_step"]:
            arrays["mask_prev"] = src["mask_prev"]
        geometry["shapes"] = {k: list(v.shape) for k, v in arrays.items()}
    manifest = {
        "configuration": {k: _plain(v) for k, v in dataclasses.asdict(cfg).items()},
        "step": int(step),
        "mesh_time_origin": float(origin),
        "time_seconds": float(np.float32(origin) + np.float32(float(step) * cfg.time_step)),
        "storage_element": _STORAGE_ELEMENTS[cfg.storage],
        "body": [float(v) for v in sim.body_state()],
        "body_parameters": None if sim.body_parameters is None
        else {k: _plain(v) for k, v in dataclasses.asdict(sim.body_parameters).items()},
        "load": [float(v) for v in sim.read_loads()],
        "geometry": geometry,
    }
    files = {POPULATIONS_NAME: populations.tobytes(), MASK_NAME: mask.tobytes()}
    for name, a in arrays.items():
        files[f"geometry-{name}.bin"] = np.ascontiguousarray(a).tobytes()
    write_archive(destination, manifest, files, compression)
    return read_manifest(destination)


def read_manifest(source: str) -> dict:
    with open(os.path.join(source, MANIFEST_NAME)) as fh:
        return json.load(fh)


def read_checkpoint(source: str) -> Checkpoint:
    """The archive with every size and checksum checked, its arrays typed
    and shaped from the manifest."""
    manifest, files = read_archive(source)
    try:
        cfg = manifest["configuration"]
        cells = int(cfg["shape"][0]) * int(cfg["shape"][1]) * int(cfg["shape"][2])
        storage = int(cfg["storage"])
        geometry = manifest["geometry"]
        kind = geometry["kind"]
    except (KeyError, TypeError, ValueError, IndexError):
        raise CheckpointError("invalid archive: manifest lacks the configuration or geometry") from None
    if storage not in _STORAGE_ELEMENTS or manifest.get("storage_element") != _STORAGE_ELEMENTS[storage]:
        raise CheckpointError("invalid archive: storage form does not match the configuration")
    if kind not in _GEOMETRY_ARRAYS:
        raise CheckpointError(f"invalid archive: unknown geometry source {kind!r}")
    element = np.float32 if storage == _abi.STORAGE_FP32 else np.uint16
    need = (POPULATIONS_NAME, MASK_NAME)
    if any(n not in files for n in need):
        raise CheckpointError("invalid archive: populations or solid mask missing")
    pops, mask = files[POPULATIONS_NAME], files[MASK_NAME]
    if cells <= 0 or len(pops) != 19 * cells * np.dtype(element).itemsize or len(mask) != cells:
        raise CheckpointError("invalid archive: grid dimensions do not match the state byte counts")
    arrays = {}
    shapes = geometry.get("shapes", {})
    for name, dtype in _GEOMETRY_ARRAYS[kind]:
        key = f"geometry-{name}.bin"
        if key not in files:
            if (kind == "static" and name == "mask_prev") or (kind == "bird" and name == "keyframes"):
                continue
            raise CheckpointError(f"invalid archive: {key} missing")
        a = np.frombuffer(files[key], dtype)
        try:
            arrays[name] = a.reshape(shapes[name]).copy()
        except (KeyError, ValueError):
            raise CheckpointError(f"invalid archive: {key} does not match its shape") from None
    return Checkpoint(manifest, np.frombuffer(pops, element).copy(), np.frombuffer(mask, np.uint8).copy(), arrays)


def load_checkpoint(source: str) -> BirdFlowSimulation:
    """init(checkpointURL:) (BirdFlowSimulation.swift:661-680) with
    restoreCheckpointState (:1920-1990): a new context at the saved step with
    the saved geometry source, body, last load, solid mask and stored
    populations."""
    ck = read_checkpoint(source)
    m = ck.manifest
    cfg = SolverConfig(**m["configuration"])
    sim = BirdFlowSimulation(cfg)
    try:
        g = m["geometry"]
        if g["kind"] == "mesh":
            a = ck.geometry
            sim.upload_mesh(a["points"], a["frame_times"], a["triangles"], a["triangle_parts"], g["half_thickness"],
                            translation=a["translation"])
        elif g["kind"] == "articulated":
            sim.upload_articulated(ArticulatedMesh.from_arrays(ck.geometry, g["kinematics"]))
        elif g["kind"] == "sphere":
            sim.set_translating_sphere(g["center_radius"], g["geometry_velocity"], g["wall_velocity"])
        elif g["kind"] == "wing":
            sim.set_flapping_wing(FlappingWing(**g["wing"]))
        elif g["kind"] == "bird":
            sim.set_bird(BirdGeometry(**g["bird"]), keyframes=ck.geometry.get("keyframes"))
        else:
            sim.set_geometry(ck.solid_mask, ck.geometry["wall"], mask_prev=ck.geometry.get("mask_prev"))
        params = None if m["body_parameters"] is None else BodyParameters(**m["body_parameters"])
        sim.set_body(m["body"], params)
        sim.set_restart_state(m["step"], m["mesh_time_origin"], last_load=m["load"], solid_mask=ck.solid_mask)
        if g["kind"] == "static":
            sim.geometry_source["set_at_step"] = int(m["step"])
        sim.write_population_storage(ck.populations)
    except BaseException:
        sim.close()
        raise
    return sim
