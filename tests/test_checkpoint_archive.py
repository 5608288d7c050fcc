"""CPU checks of the restart archive format (no device calls): the write /
read round trip, deterministic bytes, atomic replacement, and the refusals of
CheckpointArchive.read (Sources/BirdFlowMetal/CheckpointArchive.swift:81-118):
unsupported schema, byte counts that do not match, corrupted or truncated
files and checksum mismatches."""
import hashlib
import json
import os

import numpy as np
import pytest

import birdflow_hip as bh
from birdflow_hip import checkpoint


def _files(directory):
    out = {}
    for name in sorted(os.listdir(directory)):
        with open(os.path.join(directory, name), "rb") as fh:
            out[name] = fh.read()
    return out


def _payload(seed=3):
    rng = np.random.default_rng(seed)
    # compressible and incompressible parts
    a = np.repeat(rng.integers(0, 255, 4096, dtype=np.uint8), 7).tobytes()
    b = rng.integers(0, 255, 20000, dtype=np.uint8).tobytes()
    return {"populations.bin": a + b, "solid-mask.bin": bytes(5000) + b"\x01" * 300}


@pytest.mark.parametrize("compression", ["zlib", "none"])
def test_archive_round_trip_is_exact_and_deterministic(tmp_path, compression):
    files = _payload()
    manifest = {"step": 12, "note": "x"}
    d1, d2 = str(tmp_path / "one"), str(tmp_path / "two")
    checkpoint.write_archive(d1, manifest, files, compression)
    checkpoint.write_archive(d2, manifest, files, compression)
    assert _files(d1) == _files(d2)
    m, got = checkpoint.read_archive(d1)
    assert got == files
    assert m["schema"] == checkpoint.SCHEMA_VERSION and m["compression"] == compression and m["step"] == 12
    for name, data in files.items():
        entry = m["files"][name]
        assert entry["bytes"] == len(data)
        assert entry["sha256"] == hashlib.sha256(data).hexdigest()
    if compression == "zlib":
        assert os.path.getsize(os.path.join(d1, "populations.bin.z")) < len(files["populations.bin"])


def test_archive_replaces_destination_and_leaves_no_temporaries(tmp_path):
    dest = str(tmp_path / "ck")
    checkpoint.write_archive(dest, {"step": 1}, {"populations.bin": b"old" * 10, "solid-mask.bin": b"\x00"})
    checkpoint.write_archive(dest, {"step": 2}, {"populations.bin": b"new" * 10, "solid-mask.bin": b"\x01"})
    m, got = checkpoint.read_archive(dest)
    assert m["step"] == 2 and got["populations.bin"] == b"new" * 10
    assert sorted(os.listdir(tmp_path)) == ["ck"]


def test_archive_write_failure_removes_the_temporary(tmp_path):
    # a manifest JSON cannot hold: the write fails after the data files
    with pytest.raises(TypeError):
        checkpoint.write_archive(str(tmp_path / "ck"), {"bad": object()}, {"populations.bin": b"1"})
    assert os.listdir(tmp_path) == []


def _corrupt(path, offset, mask=0x40):
    with open(path, "r+b") as fh:
        fh.seek(offset)
        b = fh.read(1)
        fh.seek(offset)
        fh.write(bytes([b[0] ^ mask]))


@pytest.mark.parametrize("compression", ["zlib", "none"])
def test_corrupted_file_is_refused(tmp_path, compression):
    d = str(tmp_path / "ck")
    checkpoint.write_archive(d, {}, _payload(), compression)
    name = "populations.bin" + checkpoint.COMPRESSION_SUFFIX[compression]
    _corrupt(os.path.join(d, name), os.path.getsize(os.path.join(d, name)) // 2)
    with pytest.raises(checkpoint.CheckpointError):
        checkpoint.read_archive(d)


def test_truncated_and_missing_files_are_refused(tmp_path):
    d = str(tmp_path / "ck")
    checkpoint.write_archive(d, {}, _payload())
    path = os.path.join(d, "solid-mask.bin.z")
    blob = open(path, "rb").read()
    with open(path, "wb") as fh:
        fh.write(blob[: len(blob) // 2])
    with pytest.raises(checkpoint.CheckpointError, match="solid-mask"):
        checkpoint.read_archive(d)
    os.remove(path)
    with pytest.raises(checkpoint.CheckpointError, match="unreadable"):
        checkpoint.read_archive(d)


def test_manifest_checks(tmp_path):
    d = str(tmp_path / "ck")
    checkpoint.write_archive(d, {}, _payload(), "none")
    mpath = os.path.join(d, checkpoint.MANIFEST_NAME)
    good = json.load(open(mpath))

    def rewrite(**changes):
        m = json.loads(json.dumps(good))
        for k, v in changes.items():
            m[k] = v
        json.dump(m, open(mpath, "w"))

    rewrite(schema=2)
    with pytest.raises(checkpoint.CheckpointError, match="schema"):
        checkpoint.read_archive(d)
    # the reference's own schema-1 manifest (no format field) is refused,
    # not misread
    reference_like = {k: v for k, v in good.items() if k != "format"}
    json.dump(dict(reference_like, populationBytes=0, populationSHA256="0" * 64), open(mpath, "w"))
    with pytest.raises(checkpoint.CheckpointError, match="not a birdflow-hip archive"):
        checkpoint.read_archive(d)
    rewrite(format="birdflow-metal")
    with pytest.raises(checkpoint.CheckpointError, match="not a birdflow-hip archive"):
        checkpoint.read_archive(d)
    files = json.loads(json.dumps(good["files"]))
    files["populations.bin"]["bytes"] += 1
    rewrite(files=files)
    with pytest.raises(checkpoint.CheckpointError, match="bytes"):
        checkpoint.read_archive(d)
    files = json.loads(json.dumps(good["files"]))
    files["populations.bin"]["sha256"] = "0" * 64
    rewrite(files=files)
    with pytest.raises(checkpoint.CheckpointError, match="checksum failed for populations.bin"):
        checkpoint.read_archive(d)
    files = json.loads(json.dumps(good["files"]))
    files["populations.bin"]["file"] = "../populations.bin"
    rewrite(files=files)
    with pytest.raises(checkpoint.CheckpointError, match="file name"):
        checkpoint.read_archive(d)
    with open(mpath, "w") as fh:
        fh.write("{not json")
    with pytest.raises(checkpoint.CheckpointError, match="manifest"):
        checkpoint.read_archive(d)


def _solver_manifest(shape, storage, geometry):
    cfg = bh.SolverConfig(shape=shape, omega_plus=1.6, omega_minus=1.2, storage=storage)
    import dataclasses
    return {"configuration": {k: checkpoint._plain(v) for k, v in dataclasses.asdict(cfg).items()},
            "storage_element": checkpoint._STORAGE_ELEMENTS[storage], "geometry": geometry, "step": 3,
            "mesh_time_origin": 0.0, "body": [0.0] * 13, "body_parameters": None, "load": [0.0] * 8}


def test_read_checkpoint_types_and_shapes_the_state(tmp_path):
    shape = (8, 4, 2)
    n = 64
    rng = np.random.default_rng(5)
    pops = rng.integers(0, 65535, 19 * n, dtype=np.uint16)
    mask = (rng.random(n) < 0.2).astype(np.uint8)
    pts = rng.random((3, 5, 3), dtype=np.float32)
    geometry = {"kind": "mesh", "half_thickness": 0.5,
                "shapes": {"points": [3, 5, 3], "frame_times": [3], "triangles": [2, 3], "triangle_parts": [2],
                           "translation": [3]}}
    files = {"populations.bin": pops.tobytes(), "solid-mask.bin": mask.tobytes(),
             "geometry-points.bin": pts.tobytes(), "geometry-frame_times.bin": np.arange(3, dtype=np.float32).tobytes(),
             "geometry-triangles.bin": np.array([[0, 1, 2], [2, 3, 4]], np.uint32).tobytes(),
             "geometry-triangle_parts.bin": np.ones(2, np.uint8).tobytes(),
             "geometry-translation.bin": np.zeros(3, np.float32).tobytes()}
    d = str(tmp_path / "ck")
    checkpoint.write_archive(d, _solver_manifest(shape, bh.STORAGE_FP16_OFFSET, geometry), files)
    ck = checkpoint.read_checkpoint(d)
    assert ck.populations.dtype == np.uint16 and np.array_equal(ck.populations, pops)
    assert np.array_equal(ck.solid_mask, mask)
    assert np.array_equal(ck.geometry["points"], pts)
    assert ck.geometry["triangles"].shape == (2, 3)
    # FP32 storage expects twice the population bytes
    checkpoint.write_archive(d, _solver_manifest(shape, bh.STORAGE_FP32, geometry), files)
    with pytest.raises(checkpoint.CheckpointError, match="grid dimensions"):
        checkpoint.read_checkpoint(d)
    # a storage label that does not match the configuration
    m = _solver_manifest(shape, bh.STORAGE_FP16_OFFSET, geometry)
    m["storage_element"] = "float32"
    checkpoint.write_archive(d, m, files)
    with pytest.raises(checkpoint.CheckpointError, match="storage"):
        checkpoint.read_checkpoint(d)
    # a geometry array that does not match its shape
    bad = dict(geometry, shapes=dict(geometry["shapes"], points=[4, 5, 3]))
    checkpoint.write_archive(d, _solver_manifest(shape, bh.STORAGE_FP16_OFFSET, bad), files)
    with pytest.raises(checkpoint.CheckpointError, match="points"):
        checkpoint.read_checkpoint(d)
