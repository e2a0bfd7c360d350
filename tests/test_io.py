"""Sample/snapshot files and file initial data (host side; no GPU).

Mirrors the reference's writer contract (src/azeban/io/writer.cpp: next_timestep,
reset), the sample writer's layout and statistics
(src/azeban/io/netcdf_sample_writer.cpp:119-138, 293-358;
src/azeban/operations/statistics_recorder.cpp:35-58: population variance) and
InitFromFile's file naming (src/azeban/init/init_from_file.cpp:57-61)."""
import math

import numpy as np
import pytest

from azeban_amd import io as IO


def test_writer_timesteps_and_reset():
    w = IO.Writer([0.5, 1.0], sample_idx_start=3)
    assert w.next_timestep() == 0.5
    w.snapshot_idx += 1
    assert w.next_timestep() == 1.0
    w.snapshot_idx += 1
    assert w.next_timestep() == math.inf
    w.reset()
    assert (w.sample_idx, w.snapshot_idx, w.next_timestep()) == (4, 0, 0.5)


@pytest.mark.parametrize("dim", [2, 3])
def test_sample_writer_round_trip_and_statistics(tmp_path, dim):
    N, S, times = 8, 4, [0.25, 0.5, 1.0]
    rng = np.random.default_rng(1)
    data = rng.standard_normal((S, len(times), dim) + (N,) * dim)
    w = IO.SampleWriter(tmp_path / "samples", dim, N, times, S, store_mean_var=True)
    for s in range(S):
        ti = 0
        while (t := w.next_timestep()) != math.inf:
            assert t == times[ti]
            w.write(data[s, ti], t)
            ti += 1
        w.reset()
    w.close()
    back = IO.read_samples(tmp_path / "samples", N)
    assert np.array_equal(back["time"], np.array(times))
    for c, name in enumerate("uvw"[:dim]):
        assert np.array_equal(np.asarray(back[name]), data[:, :, c])
        assert np.allclose(back["mean"][name], data[:, :, c].mean(axis=0), rtol=0, atol=1e-14)
        assert np.allclose(back["variance"][name], data[:, :, c].var(axis=0), rtol=0, atol=1e-13)
    rt = back["real_time"]
    assert rt.shape == (S, len(times)) and np.all(np.diff(rt, axis=1) >= 0)


def test_sample_writer_rejects_bad_writes(tmp_path):
    w = IO.SampleWriter(tmp_path / "s", 2, 8, [1.0], 1)
    with pytest.raises(ValueError):
        w.write(np.zeros((2, 4, 4)), 1.0)
    w.write(np.zeros((2, 8, 8)), 1.0)
    with pytest.raises(ValueError):
        w.write(np.zeros((2, 8, 8)), 1.0)
    w.reset()
    with pytest.raises(ValueError):  # member index beyond the file
        w.write(np.zeros((2, 8, 8)), 1.0)


def test_snapshot_writer(tmp_path):
    rng = np.random.default_rng(2)
    u = rng.standard_normal((3, 4, 4, 4))
    w = IO.SnapshotWriter(tmp_path / "snap" / "run", 3, [0.1, 0.2])
    p = w.write(u, 0.1)
    with np.load(p) as f:
        assert float(f["t"]) == 0.1
        for i, c in enumerate("uvw"):
            assert np.array_equal(f[c], u[i])
    uh = rng.standard_normal((3, 4, 4, 3)) + 1j * rng.standard_normal((3, 4, 4, 3))
    wh = IO.SnapshotWriter(tmp_path / "snap" / "hat", 3, [0.1], store_u_hat=True)
    with np.load(wh.write_hat(uh, 0.1)) as f:
        for i, c in enumerate("uvw"):
            rec = f[f"{c}_hat_abs"] * np.exp(1j * f[f"{c}_hat_arg"])
            assert np.allclose(rec, uh[i], rtol=0, atol=1e-14)


def test_init_from_file_reads_samples_in_order(tmp_path):
    rng = np.random.default_rng(3)
    fields = [rng.standard_normal((2, 8, 8)) for _ in range(2)]
    for i, f in enumerate(fields):
        np.savez(tmp_path / f"sample_{i}_time_0.5.npz", u=f[0], v=f[1])
    init = IO.InitFromFile(tmp_path, "0.5")
    assert init.filename().endswith("sample_0_time_0.5.npz")
    assert np.array_equal(init.read_u(2, 8), fields[0])
    assert np.array_equal(init.read_u(2, 8), fields[1])
    with pytest.raises(FileNotFoundError):
        init.read_u(2, 8)
    np.savez(tmp_path / "sample_0_time_x.npz", u=fields[0][0])
    with pytest.raises(ValueError):
        IO.InitFromFile(tmp_path, "x").read_u(2, 8)
