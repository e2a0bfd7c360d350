"""Sample and snapshot files, and initial data read back from them.

azeban writes NetCDF (src/azeban/io/netcdf_sample_writer.cpp,
netcdf_snapshot_writer.cpp) and reads initial data from NetCDF files
(src/azeban/init/init_from_file.cpp).  netcdf-c is not part of this stack, so the
same variables, dimensions and group structure are kept in NumPy's own formats:

* a sample file is a directory; the group ``flow_field_<N>`` is a sub-directory
  holding ``time.npy`` (snapshot times), ``real_time.npy`` (member x time, seconds
  since the writer started), and one array per velocity component ``u.npy``,
  ``v.npy`` (, ``w.npy``) of shape (member, time, N, ..., N), float64, written in
  place through memory maps; with ``store_mean_var`` the sub-groups ``mean/`` and
  ``variance/`` hold (time, N, ..., N) arrays per component
  (netcdf_sample_writer.cpp:293-358);
* a snapshot is one ``.npz`` per written time with ``t`` and either the physical
  components ``u, v, w`` or ``u_hat_abs``/``u_hat_arg`` per component
  (netcdf_snapshot_writer.cpp:30-61, 255-271);
* initial data is read from ``<experiment>/sample_<i>_time_<t>.npz`` holding
  ``u, v(, w)`` in physical space (init_from_file.cpp:57-61, 134-155).

Physical fields come from the device path (``to_physical``: u_hat / N^d, then
the inverse transform; downsampled output truncates the spectrum first).  The
files themselves are plain host I/O.
"""
from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

import numpy as np

_COMPONENTS = ("u", "v", "w")


def _as_numpy(x) -> np.ndarray:
    if hasattr(x, "detach"):  # torch tensor, host or device
        x = x.detach()
        if x.is_cuda:
            import torch

            torch.cuda.synchronize()
        x = x.cpu().numpy()
    return np.asarray(x)


class Writer:
    """Snapshot-time bookkeeping shared by the writers (src/azeban/io/writer.cpp):
    next_timestep() is the next snapshot time or +inf; reset() moves to the next
    sample and back to the first snapshot."""

    def __init__(self, snapshot_times, sample_idx_start: int = 0):
        self.snapshot_times = [float(t) for t in snapshot_times]
        self.sample_idx = int(sample_idx_start)
        self.snapshot_idx = 0

    def next_timestep(self) -> float:
        if self.snapshot_idx >= len(self.snapshot_times):
            return math.inf
        return self.snapshot_times[self.snapshot_idx]

    def set_snapshot_idx(self, idx: int) -> None:
        self.snapshot_idx = int(idx)

    def reset(self) -> None:
        self.snapshot_idx = 0
        self.sample_idx += 1


class _Welford:
    """Running mean and population variance over samples, per snapshot
    (src/azeban/operations/statistics_recorder.cpp:35-58)."""

    def __init__(self, shape):
        self.count = 0
        self.mean = np.zeros(shape)
        self.m2 = np.zeros(shape)

    def update(self, x: np.ndarray) -> None:
        self.count += 1
        d1 = x - self.mean
        self.mean += d1 / self.count
        self.m2 += d1 * (x - self.mean)

    def variance(self) -> np.ndarray:
        return self.m2 / self.count if self.count else self.m2


class SampleWriter(Writer):
    """One group ``flow_field_<N>`` of a sample file (NetCDFSampleWriter).

    ``write(u, t)`` takes the physical velocity (dim, N, ..., N) of the current
    sample at snapshot ``t``; ``write_hat(u_hat, t)`` takes the spectral state of
    the solver's grid and writes its physical field on this writer's N (truncating
    the spectrum when N is smaller, netcdf_sample_writer.cpp:140-170)."""

    def __init__(self, path, dim: int, N: int, snapshot_times, num_samples: int,
                 store_mean_var: bool = False, sample_idx_start: int = 0):
        super().__init__(snapshot_times, sample_idx_start)
        self.dim = int(dim)
        self.N = int(N)
        self.num_samples = int(num_samples)
        self.store_mean_var = bool(store_mean_var)
        self.root = Path(path)
        self.group = self.root / f"flow_field_{self.N}"
        self.group.mkdir(parents=True, exist_ok=True)
        T = len(self.snapshot_times)
        np.save(self.group / "time.npy", np.asarray(self.snapshot_times, dtype=np.float64))
        self._real_time = np.lib.format.open_memmap(self.group / "real_time.npy", mode="w+",
                                                    dtype=np.float64, shape=(self.num_samples, T))
        shape = (self.num_samples, T) + (self.N,) * self.dim
        self._fields = [np.lib.format.open_memmap(self.group / f"{c}.npy", mode="w+", dtype=np.float64,
                                                  shape=shape)
                        for c in _COMPONENTS[: self.dim]]
        self._stats = [_Welford((self.dim,) + (self.N,) * self.dim) for _ in range(T)] \
            if self.store_mean_var else None
        meta = {"dim": self.dim, "N": self.N, "members": self.num_samples, "snapshots": T,
                "variables": list(_COMPONENTS[: self.dim]), "dtype": "float64",
                "layout": "member, time, x, y[, z]"}
        (self.group / "attributes.json").write_text(json.dumps(meta, indent=1))
        self._start = time.monotonic()

    def write(self, u, t: float) -> None:
        u = _as_numpy(u)
        want = (self.dim,) + (self.N,) * self.dim
        if u.shape != want:
            raise ValueError(f"expected a physical field of shape {want}, got {u.shape}")
        if self.snapshot_idx >= len(self.snapshot_times):
            raise ValueError("all snapshots of this sample are written")
        m = self.sample_idx
        if not 0 <= m < self.num_samples:
            raise ValueError(f"sample index {m} outside the file's {self.num_samples} members")
        for c in range(self.dim):
            self._fields[c][m, self.snapshot_idx] = u[c]
        if self._stats is not None:
            self._stats[self.snapshot_idx].update(u.astype(np.float64))
        self._real_time[m, self.snapshot_idx] = time.monotonic() - self._start
        self.snapshot_idx += 1

    def write_hat(self, u_hat, t: float) -> None:
        from .solver import to_physical

        self.write(to_physical(u_hat, self.N), t)

    def close(self) -> None:
        """Flush the fields and write the per-snapshot mean and variance."""
        for f in self._fields:
            f.flush()
        self._real_time.flush()
        if self._stats is not None:
            T = len(self.snapshot_times)
            for grp, get in (("mean", lambda s: s.mean), ("variance", lambda s: s.variance())):
                d = self.group / grp
                d.mkdir(exist_ok=True)
                for c, name in enumerate(_COMPONENTS[: self.dim]):
                    arr = np.stack([get(self._stats[i])[c] for i in range(T)])
                    np.save(d / f"{name}.npy", arr)


class SnapshotWriter(Writer):
    """One file per snapshot (NetCDFSnapshotWriter): ``<prefix>_<sample>_time_<idx>.npz``
    with ``t`` and the physical components, or with ``store_u_hat`` the modulus and
    argument of each spectral component (netcdf_snapshot_writer.cpp:255-271)."""

    def __init__(self, prefix, dim: int, snapshot_times, store_u_hat: bool = False,
                 sample_idx_start: int = 0):
        super().__init__(snapshot_times, sample_idx_start)
        self.prefix = str(prefix)
        self.dim = int(dim)
        self.store_u_hat = bool(store_u_hat)
        Path(self.prefix).parent.mkdir(parents=True, exist_ok=True)

    def _path(self) -> str:
        return f"{self.prefix}_{self.sample_idx}_time_{self.snapshot_idx}.npz"

    def write(self, u, t: float) -> str:
        u = _as_numpy(u)
        path = self._path()
        np.savez(path, t=np.float64(t), **{c: u[i] for i, c in enumerate(_COMPONENTS[: self.dim])})
        self.snapshot_idx += 1
        return path

    def write_hat(self, u_hat, t: float) -> str:
        if not self.store_u_hat:
            from .solver import to_physical

            return self.write(to_physical(u_hat), t)
        uh = _as_numpy(u_hat)
        path = self._path()
        arrays = {}
        for i, c in enumerate(_COMPONENTS[: self.dim]):
            arrays[f"{c}_hat_abs"] = np.abs(uh[i])
            arrays[f"{c}_hat_arg"] = np.angle(uh[i])
        np.savez(path, t=np.float64(t), **arrays)
        self.snapshot_idx += 1
        return path


class InitFromFile:
    """Initial data from ``<experiment>/sample_<i>_time_<time>.npz`` (InitFromFile):
    each call reads the next sample's physical ``u, v(, w)``; ``initialize`` returns
    the spectral state on the device (forward transform, then Leray, as
    Initializer::initialize does for spectral states, initializer.hpp:48-75)."""

    def __init__(self, experiment, time_: str, sample_idx_start: int = 0):
        self.experiment = str(experiment)
        self.time = str(time_)
        self.sample = int(sample_idx_start)

    def filename(self) -> str:
        return os.path.join(self.experiment, f"sample_{self.sample}_time_{self.time}.npz")

    def read_u(self, dim: int, N: int) -> np.ndarray:
        with np.load(self.filename(), allow_pickle=False) as f:
            names = _COMPONENTS[:dim]
            missing = [c for c in names if c not in f.files]
            if missing:
                raise ValueError(f"{self.filename()} lacks {missing}")
            out = np.stack([np.asarray(f[c], dtype=np.float64) for c in names])
        if out.shape != (dim,) + (N,) * dim:
            raise ValueError(f"initial data has shape {out.shape}, expected {(dim,) + (N,) * dim}")
        self.sample += 1
        return out

    def initialize(self, dim: int, N: int):
        from .solver import initialize

        return initialize(self.read_u(dim, N))


def read_samples(path, N: int) -> dict:
    """Load a sample file's group back: {'time', 'real_time', 'u', 'v'(, 'w'),
    'mean'?, 'variance'?} (memory-mapped field arrays)."""
    g = Path(path) / f"flow_field_{N}"
    meta = json.loads((g / "attributes.json").read_text())
    out = {"time": np.load(g / "time.npy"), "real_time": np.load(g / "real_time.npy")}
    for c in meta["variables"]:
        out[c] = np.load(g / f"{c}.npy", mmap_mode="r")
    for grp in ("mean", "variance"):
        if (g / grp).is_dir():
            out[grp] = {c: np.load(g / grp / f"{c}.npy") for c in meta["variables"]}
    return out
