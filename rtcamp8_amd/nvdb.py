"""NanoVDB volumes: .nvdb files and grid buffers, through the library's codec.

Mirrors what the reference does with nanovdb::io (ext/nanovdb/nanovdb/util/IO.h): ``read_grid`` is
``nanovdb::io::readGrid(fileName, n)`` and ``load_grid`` adds the device upload and majorant of
``Scene::allocateVolumeGrid`` (renderer/renderer_scene.cpp:491-537, 2467-2470).  ``encode_grid`` and
``write_grids`` produce files the reference reads back unchanged.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _abi
from ._lib import call

CODECS = {"none": _abi.NVDB_CODEC_NONE, "zip": _abi.NVDB_CODEC_ZIP, "blosc": _abi.NVDB_CODEC_BLOSC}


def _bytes(data) -> bytes:
    if isinstance(data, (str, Path)):
        return Path(data).read_bytes()
    return bytes(data)


@dataclass
class GridInfo:
    name: str
    grid_type: int
    grid_class: int
    codec: int
    grid_size: int
    file_size: int
    voxel_count: int
    node_count: tuple
    tile_count: tuple
    world_bbox: tuple
    index_bbox: tuple
    voxel_size: tuple
    version: tuple
    segment: int


def list_grids(data) -> list[GridInfo]:
    """The grids of a .nvdb file image (path or bytes), in file order (readGridMetaData)."""
    buf = _bytes(data)
    count = C.c_uint32(0)
    call("rtc8x_nvdb_list", buf, len(buf), None, 0, C.byref(count))
    infos = (_abi.NvdbGridInfo * max(count.value, 1))()
    call("rtc8x_nvdb_list", buf, len(buf), infos, count.value, C.byref(count))
    out = []
    for i in range(count.value):
        g = infos[i]
        out.append(GridInfo(
            name=g.name.decode(errors="replace"), grid_type=g.grid_type, grid_class=g.grid_class, codec=g.codec,
            grid_size=g.grid_size, file_size=g.file_size, voxel_count=g.voxel_count,
            node_count=tuple(g.node_count), tile_count=tuple(g.tile_count), world_bbox=tuple(g.world_bbox),
            index_bbox=tuple(g.index_bbox), voxel_size=tuple(g.voxel_size),
            version=(g.version >> 21, (g.version >> 10) & 2047, g.version & 1023), segment=g.segment))
    return out


class HostGrid:
    """Grid tables in host memory owned by the library (released by ``close``)."""

    def __init__(self, grid: _abi.Grid):
        self.grid = grid

    def close(self) -> None:
        if self.grid.root_tiles:
            call("rtc8x_grid_free_host", C.byref(self.grid))

    def __enter__(self) -> "HostGrid":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - safety net
        try:
            self.close()
        except Exception:
            pass

    # views of the tables (valid until close)
    def _view(self, ptr: int, ctype, count: int) -> np.ndarray:
        if count == 0 or not ptr:
            return np.zeros(0, dtype=np.dtype(ctype))
        return np.ctypeslib.as_array((ctype * count).from_address(ptr))

    def leaf_values(self) -> np.ndarray:
        g = self.grid
        return self._view(g.leaf_values, C.c_float, g.num_leaves * _abi.LEAF_SIZE).reshape(-1, _abi.LEAF_SIZE)

    def leaf_origins(self) -> np.ndarray:
        g = self.grid
        return self._view(g.leaf_origins, C.c_int32, 3 * g.num_leaves).reshape(-1, 3)

    def root_tiles(self) -> np.ndarray:
        g = self.grid
        raw = self._view(g.root_tiles, C.c_uint64, 2 * g.num_root_tiles).reshape(-1, 2)
        return raw


def decode(buffer) -> HostGrid:
    """A FloatGrid buffer (what a GridHandle holds) as host tables."""
    buf = _bytes(buffer)
    g = _abi.Grid()
    call("rtc8x_nanovdb_decode", buf, len(buf), C.byref(g))
    return HostGrid(g)


def read_grid(data, n: int = 0) -> HostGrid:
    """Grid n of a .nvdb file image as host tables (nanovdb::io::readGrid)."""
    buf = _bytes(data)
    g = _abi.Grid()
    call("rtc8x_nvdb_read_grid", buf, len(buf), n, C.byref(g))
    return HostGrid(g)


def load_grid(data, n: int = 0) -> _abi.Grid:
    """Grid n of a .nvdb file image on the device, with the reference's majorant.  Release with
    rtc8x_grid_destroy."""
    buf = _bytes(data)
    g = _abi.Grid()
    call("rtc8x_grid_load_nvdb", buf, len(buf), n, C.byref(g))
    return g


def download(device_grid: _abi.Grid) -> HostGrid:
    """Host copy of a device grid's tables."""
    g = _abi.Grid()
    call("rtc8x_grid_download", C.byref(device_grid), C.byref(g))
    return HostGrid(g)


def encode_grid(host: _abi.Grid, name: str = "density") -> bytes:
    """Host tables as a NanoVDB FloatGrid buffer."""
    size = C.c_size_t(0)
    call("rtc8x_nanovdb_encode", C.byref(host), name.encode(), None, 0, C.byref(size))
    out = C.create_string_buffer(size.value)
    call("rtc8x_nanovdb_encode", C.byref(host), name.encode(), out, size.value, C.byref(size))
    return out.raw[: size.value]


def write_grids(buffers: list[bytes], codec: str = "none") -> bytes:
    """FloatGrid buffers as one .nvdb segment (nanovdb::io::writeGrids)."""
    n = len(buffers)
    keep = [C.create_string_buffer(b, len(b)) for b in buffers]
    ptrs = (C.c_void_p * n)(*[C.addressof(k) for k in keep])
    sizes = (C.c_size_t * n)(*[len(b) for b in buffers])
    size = C.c_size_t(0)
    call("rtc8x_nvdb_write", ptrs, sizes, n, CODECS[codec], None, 0, C.byref(size))
    out = C.create_string_buffer(size.value)
    call("rtc8x_nvdb_write", ptrs, sizes, n, CODECS[codec], out, size.value, C.byref(size))
    return out.raw[: size.value]


def save_device_grid(device_grid: _abi.Grid, path, name: str = "density", codec: str = "none") -> int:
    """Writes a device grid (e.g. the benchmark's procedural volume) as a one-grid .nvdb file; returns
    its size in bytes."""
    with download(device_grid) as h:
        data = write_grids([encode_grid(h.grid, name)], codec)
    Path(path).write_bytes(data)
    return len(data)
