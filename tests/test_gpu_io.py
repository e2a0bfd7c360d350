"""Device side of the output path and file initial data, against the oracle:
truncation (copy_from_padded), the 1/N^d-scaled physical field
(src/azeban/run_from_config.cpp:93-98), downsampled sample output
(src/azeban/io/netcdf_sample_writer.cpp:140-170) and InitFromFile + Leray
(include/azeban/init/initializer.hpp:67-73)."""
import numpy as np
import pytest
import torch

from oracle import azeban_oracle as O

import azeban_amd as A
from azeban_amd import init as I
from azeban_amd import io as IO

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("dim,M,N", [(2, 24, 16), (2, 128, 64), (3, 24, 16), (3, 96, 64), (3, 32, 32)])
def test_copy_from_padded_exact(cuda, dim, M, N):
    print(f"[case] copy_from_padded dim={dim} M={M} N={N}", flush=True)
    rng = np.random.default_rng(9)
    shape = (M,) * (dim - 1) + (M // 2 + 1,)
    x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    got = host(A.copy_from_padded(dev(x), N))
    assert np.array_equal(got, O.copy_from_padded(x, (N,) * (dim - 1) + (N // 2 + 1,)))


def test_scale(cuda):
    x = torch.arange(10, dtype=torch.float64, device="cuda")
    A.scale_(0.5, x)
    assert np.array_equal(host(x), np.arange(10) * 0.5)


@pytest.mark.parametrize("dim,N,n_out", [(2, 64, 64), (2, 64, 32), (3, 32, 32), (3, 32, 16)])
def test_to_physical(cuda, dim, N, n_out):
    print(f"[case] to_physical dim={dim} N={N} n_out={n_out}", flush=True)
    uh = O.initialize(I.fractional_brownian(dim, N, seed=4))
    got = host(A.to_physical(dev(uh), n_out))
    small = np.stack([O.copy_from_padded(uh[c], (n_out,) * (dim - 1) + (n_out // 2 + 1,))
                      for c in range(dim)])
    expect = O.fft_c2r(small / float(N) ** dim, n_out)
    assert rel(got, expect) <= 1e-14
    if n_out == N:  # the round trip of the initial data
        assert np.max(np.abs(got - O.fft_c2r(uh, N) / N**dim)) <= 1e-13 * np.max(np.abs(got))


def test_sample_writer_from_device_state(cuda, tmp_path):
    N, dim = 32, 2
    uh = O.initialize(I.fractional_brownian(dim, N, seed=6))
    w = IO.SampleWriter(tmp_path / "s", dim, 16, [1.0], 1)
    w.write_hat(dev(uh), 1.0)
    w.close()
    back = IO.read_samples(tmp_path / "s", 16)
    small = np.stack([O.copy_from_padded(uh[c], (16, 9)) for c in range(dim)])
    expect = O.fft_c2r(small / float(N) ** dim, 16)
    assert rel(np.stack([back["u"][0, 0], back["v"][0, 0]]), expect) <= 1e-14


def test_init_from_file_matches_oracle_initialize(cuda, tmp_path):
    u = I.fractional_brownian(3, 16, seed=8)
    np.savez(tmp_path / "sample_0_time_0.npz", u=u[0], v=u[1], w=u[2])
    got = host(IO.InitFromFile(tmp_path, "0").initialize(3, 16))
    assert rel(got, O.initialize(u)) <= 1e-14
