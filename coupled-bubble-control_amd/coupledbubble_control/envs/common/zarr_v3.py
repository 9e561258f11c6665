"""Zarr v3 directory stores without the zarr package: groups, attributes and
uncompressed arrays, plus a reader for the same subset.

The reference writes its trajectories with zarr-python 3 (``zarr.open`` /
``create_group`` / ``create_array``, envs/common/trajectory_buffer.py:158-226),
which stores the Zarr v3 format: every node is a directory with a
``zarr.json`` metadata document; an array's chunks are files under ``c/``
named by their grid coordinates.  This module writes that format as the Zarr
v3 core specification defines it -- regular chunk grid, default chunk-key
encoding with separator "/", the ``bytes`` codec (little endian) and no
compressor -- so zarr-python 3 opens the stores as ordinary groups and
arrays.  Arrays are written as one chunk.
"""
from __future__ import annotations

import json
import os
import shutil
from pathlib import Path
from typing import Any, Dict, List, Union

import numpy as np

PathLike = Union[str, Path]

_DTYPES = {"float32": "<f4", "float64": "<f8", "int32": "<i4", "int64": "<i8", "uint8": "|u1", "bool": "|b1"}


def _jsonable(v: Any) -> Any:
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    return v


def _write_json(path: Path, doc: Dict) -> None:
    with open(path, "w") as f:
        f.write(json.dumps(_jsonable(doc), separators=(",", ":")))


def is_node(path: PathLike) -> bool:
    return (Path(path) / "zarr.json").is_file()


def open_group(path: PathLike) -> Dict:
    """Create the group at ``path`` if it does not exist (zarr.open(mode="a"));
    returns its metadata document."""
    p = Path(path)
    meta = p / "zarr.json"
    if meta.is_file():
        doc = json.loads(meta.read_text())
        if doc.get("node_type") != "group":
            raise ValueError(f"{p} is a zarr array, not a group")
        return doc
    p.mkdir(parents=True, exist_ok=True)
    doc = {"zarr_format": 3, "node_type": "group", "attributes": {}}
    _write_json(meta, doc)
    return doc


def group_attrs(path: PathLike) -> Dict:
    return dict(open_group(path).get("attributes", {}))


def update_attrs(path: PathLike, attrs: Dict) -> None:
    """attrs.update for a group or an array node."""
    meta = Path(path) / "zarr.json"
    doc = json.loads(meta.read_text())
    doc.setdefault("attributes", {}).update(_jsonable(attrs))
    _write_json(meta, doc)


def create_group(parent: PathLike, name: str) -> Path:
    p = Path(parent) / name
    open_group(p)
    return p


def delete(path: PathLike) -> None:
    shutil.rmtree(path, ignore_errors=True)


def children(path: PathLike) -> List[str]:
    p = Path(path)
    return sorted(c.name for c in p.iterdir() if c.is_dir() and (c / "zarr.json").is_file())


def write_array(parent: PathLike, name: str, data, dtype: str = "float32") -> Path:
    """An array node ``parent/name`` holding ``data`` cast to ``dtype``, one chunk."""
    if dtype not in _DTYPES:
        raise ValueError(f"unsupported dtype {dtype}")
    a = np.ascontiguousarray(np.asarray(data).astype(np.dtype(_DTYPES[dtype]), copy=False))
    p = Path(parent) / name
    if p.exists():
        delete(p)
    shape = list(a.shape)
    key = p / "c"
    for _ in shape:
        key = key / "0"
    os.makedirs(key.parent if a.size else p, exist_ok=True)
    doc = {
        "zarr_format": 3,
        "node_type": "array",
        "shape": shape,
        "data_type": dtype,
        "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [max(1, d) for d in shape]}},
        "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
        "fill_value": False if dtype == "bool" else (0.0 if dtype.startswith("float") else 0),
        "codecs": [{"name": "bytes", "configuration": {"endian": "little"}}],
        "attributes": {},
        "storage_transformers": [],
    }
    _write_json(p / "zarr.json", doc)
    if a.size:
        with open(key, "wb") as f:
            f.write(a.tobytes(order="C"))
    return p


def _array_doc(shape, dtype: str) -> str:
    fill = "false" if dtype == "bool" else ("0.0" if dtype.startswith("float") else "0")
    chunk = ",".join(str(max(1, d)) for d in shape)
    return ('{"zarr_format":3,"node_type":"array","shape":[%s],"data_type":"%s",'
            '"chunk_grid":{"name":"regular","configuration":{"chunk_shape":[%s]}},'
            '"chunk_key_encoding":{"name":"default","configuration":{"separator":"/"}},'
            '"fill_value":%s,"codecs":[{"name":"bytes","configuration":{"endian":"little"}}],'
            '"attributes":{},"storage_transformers":[]}' % (",".join(str(d) for d in shape), dtype, chunk, fill))


def write_new_group(path: PathLike, arrays: Dict[str, np.ndarray], attrs: Dict, dtype: str = "float32") -> None:
    """A new group at ``path`` (which must not exist) holding ``arrays`` (each cast
    to ``dtype``, one chunk) and ``attrs``: the same nodes as open_group +
    write_array + update_attrs, with one directory and file per node part."""
    base = os.fspath(path)
    os.mkdir(base)
    npdt = np.dtype(_DTYPES[dtype])
    for name, data in arrays.items():
        a = np.ascontiguousarray(np.asarray(data).astype(npdt, copy=False))
        node = os.path.join(base, name)
        os.mkdir(node)
        with open(os.path.join(node, "zarr.json"), "w") as f:
            f.write(_array_doc(a.shape, dtype))
        if a.size:
            d = os.path.join(node, "c")
            os.mkdir(d)
            for _ in range(a.ndim - 1):
                d = os.path.join(d, "0")
                os.mkdir(d)
            with open(os.path.join(d, "0"), "wb") as f:
                f.write(a.tobytes(order="C"))
    with open(os.path.join(base, "zarr.json"), "w") as f:
        f.write(json.dumps({"zarr_format": 3, "node_type": "group", "attributes": _jsonable(attrs)},
                           separators=(",", ":")))


def read_array(path: PathLike) -> np.ndarray:
    """An array written in the subset above (one chunk or none, bytes codec)."""
    p = Path(path)
    doc = json.loads((p / "zarr.json").read_text())
    if doc.get("node_type") != "array":
        raise ValueError(f"{p} is not a zarr array")
    codecs = doc["codecs"]
    if len(codecs) != 1 or codecs[0]["name"] != "bytes":
        raise ValueError(f"{p}: only the bytes codec is supported here, got {codecs}")
    shape = tuple(doc["shape"])
    chunk = tuple(doc["chunk_grid"]["configuration"]["chunk_shape"])
    if any(c < s for c, s in zip(chunk, shape)):
        raise ValueError(f"{p}: more than one chunk")
    endian = "<" if codecs[0].get("configuration", {}).get("endian", "little") == "little" else ">"
    dt = np.dtype(_DTYPES[doc["data_type"]]).newbyteorder(endian) if doc["data_type"] != "bool" else np.dtype(bool)
    key = p / "c"
    for _ in shape:
        key = key / "0"
    if not key.is_file():
        return np.full(shape, doc["fill_value"], dtype=dt)
    return np.frombuffer(key.read_bytes(), dtype=dt).reshape(chunk)[tuple(slice(0, s) for s in shape)].copy()
