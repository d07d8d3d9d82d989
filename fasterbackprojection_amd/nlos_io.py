"""Reader/writer of the reference's .nlos binary capture cache
(NLosData::saveBinaryFile / loadBinaryFile, FB/NLosData.cpp:320-405).

Layout (little endian): vec3 cameraPosition, vec3 laserPosition, uint32
temporalResolution, float deltaT, float t0, bool isConfocal (1 byte), AABB
(8-byte vptr slot, vec3 max, vec3 min), uint64 numDims + uint64 dims[],
uint64 n + vec3 cameraGridPositions[n] + vec3 cameraGridNormals[n],
vec2 cameraGridSize, the same three for the laser grid, uint64 numValues +
float data[numValues].
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np


@dataclass
class NlosFile:
    camera_position: np.ndarray
    laser_position: np.ndarray
    temporal_resolution: int
    delta_t: float
    t0: float
    is_confocal: bool
    hidden_max: np.ndarray
    hidden_min: np.ndarray
    dims: list
    camera_grid_positions: np.ndarray
    camera_grid_normals: np.ndarray
    camera_grid_size: np.ndarray
    laser_grid_positions: np.ndarray
    laser_grid_normals: np.ndarray
    laser_grid_size: np.ndarray
    data: np.ndarray


def write(path, f: NlosFile) -> None:
    v3 = lambda a: np.asarray(a, np.float32).reshape(3).tobytes()
    with open(path, "wb") as fh:
        fh.write(v3(f.camera_position))
        fh.write(v3(f.laser_position))
        fh.write(struct.pack("<Iff?", int(f.temporal_resolution), f.delta_t, f.t0, bool(f.is_confocal)))
        fh.write(b"\0" * 8)
        fh.write(v3(f.hidden_max))
        fh.write(v3(f.hidden_min))
        fh.write(struct.pack("<Q", len(f.dims)))
        fh.write(np.asarray(f.dims, np.uint64).tobytes())
        for pos, nrm, size in ((f.camera_grid_positions, f.camera_grid_normals, f.camera_grid_size),
                               (f.laser_grid_positions, f.laser_grid_normals, f.laser_grid_size)):
            pos = np.ascontiguousarray(pos, np.float32)
            fh.write(struct.pack("<Q", len(pos)))
            fh.write(pos.tobytes())
            fh.write(np.ascontiguousarray(nrm, np.float32).tobytes())
            fh.write(np.asarray(size, np.float32).reshape(2).tobytes())
        d = np.ascontiguousarray(f.data, np.float32).ravel()
        fh.write(struct.pack("<Q", d.size))
        fh.write(d.tobytes())


def read(path) -> NlosFile:
    b = open(path, "rb").read()
    o = 0

    def take(n):
        nonlocal o
        r = b[o:o + n]
        o += n
        return r
    f32 = lambda n: np.frombuffer(take(4 * n), np.float32).copy()
    cam = f32(3)
    las = f32(3)
    T, dt, t0, conf = struct.unpack("<Iff?", take(13))
    take(8)
    hmax = f32(3)
    hmin = f32(3)
    nd = struct.unpack("<Q", take(8))[0]
    dims = [int(x) for x in np.frombuffer(take(8 * nd), np.uint64)]
    grids = []
    for _ in range(2):
        n = struct.unpack("<Q", take(8))[0]
        grids.append((f32(3 * n).reshape(n, 3), f32(3 * n).reshape(n, 3), f32(2)))
    nv = struct.unpack("<Q", take(8))[0]
    data = f32(nv)
    return NlosFile(cam, las, T, dt, t0, conf, hmax, hmin, dims, grids[0][0], grids[0][1], grids[0][2],
                    grids[1][0], grids[1][1], grids[1][2], data)


def from_geometry(g, transient: np.ndarray) -> NlosFile:
    """A .nlos record for a synthetic capture (fasterbackprojection_amd.scene)."""
    L = len(g.laser_targets)
    S = len(g.sensor_targets)
    T = int(g.num_time_bins)
    if g.confocal:
        n = int(round(np.sqrt(L)))
        dims = [n, L // n, T] if n * n == L else [1, L, T]
    else:
        dims = [1, L, 1, S, T]
    up = np.tile(np.array([0, 1, 0], np.float32), (1, 1))
    return NlosFile(np.asarray(g.sensor_position, np.float32), np.asarray(g.laser_position, np.float32),
                    T, float(np.float32(g.time_step)), float(np.float32(g.time_offset)), bool(g.confocal),
                    np.asarray(g.hidden_max, np.float32), np.asarray(g.hidden_min, np.float32), dims,
                    np.asarray(g.sensor_targets, np.float32), np.repeat(up, S, 0),
                    np.asarray(g.wall_size, np.float32), np.asarray(g.laser_targets, np.float32),
                    np.repeat(up, L, 0), np.asarray(g.wall_size, np.float32),
                    np.asarray(transient, np.float32).ravel())
