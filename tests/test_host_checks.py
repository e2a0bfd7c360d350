"""The Python seams reject tensors of the wrong dtype or shape before any library
call (the C ABI takes bare pointers and element counts, so a mismatch would read or
write outside the tensor on the device).  CPU only: the library loader is replaced
by one that fails the test if it is reached."""
import pytest
import torch

import azeban_amd as A
from azeban_amd import solver as S


@pytest.fixture
def no_library(monkeypatch):
    def boom(*a, **k):
        raise AssertionError("the library was called before the arguments were checked")

    monkeypatch.setattr(S, "load", boom)


def c128(*shape):
    return torch.zeros(shape, dtype=torch.complex128)


CASES = [
    ("axpby_ float64 x", lambda: A.axpby_(1.0, torch.zeros(8, dtype=torch.float64), 1.0, c128(8)), "dtype"),
    ("axpby_ float64 y", lambda: A.axpby_(1.0, c128(8), 1.0, torch.zeros(8, dtype=torch.float64)), "dtype"),
    ("axpby_ short y", lambda: A.axpby_(1.0, c128(8), 1.0, c128(4)), "elements"),
    ("max_norm complex", lambda: A.max_norm(c128(8)), "dtype"),
    ("max_norm float32", lambda: A.max_norm(torch.zeros(8, dtype=torch.float32)), "dtype"),
    ("leray_ complex64", lambda: A.leray_(torch.zeros((2, 8, 5), dtype=torch.complex64)), "dtype"),
    ("leray_ bad last axis", lambda: A.leray_(c128(2, 8, 4)), "shape"),
    ("leray_ components", lambda: A.leray_(c128(3, 8, 5)), "shape"),
    ("spectrum complex64", lambda: A.energy_spectrum(torch.zeros((3, 8, 8, 5), dtype=torch.complex64)), "dtype"),
    ("spectrum unequal axes", lambda: A.energy_spectrum(c128(3, 8, 16, 5)), "shape"),
    ("pad complex64", lambda: A.copy_to_padded(torch.zeros((8, 5), dtype=torch.complex64), 12), "dtype"),
    ("pad bad shape", lambda: A.copy_to_padded(c128(8, 8), 12), "shape"),
    ("pad M < N", lambda: A.copy_to_padded(c128(8, 5), 6), "M must"),
    ("unpad complex64", lambda: A.copy_from_padded(torch.zeros((12, 7), dtype=torch.complex64), 8), "dtype"),
    ("unpad N > M", lambda: A.copy_from_padded(c128(12, 7), 16), "N must"),
    ("r2c unequal axes", lambda: A.fft_r2c(torch.zeros((1, 8, 16), dtype=torch.float64)), "shape"),
    ("r2c complex", lambda: A.fft_r2c(c128(1, 8, 8)), "dtype"),
    ("c2r wrong n", lambda: A.fft_c2r(c128(1, 8, 5), 16), "shape"),
    ("c2r real", lambda: A.fft_c2r(torch.zeros((1, 8, 5), dtype=torch.float64), 8), "dtype"),
    ("scale_ complex64", lambda: A.scale_(2.0, torch.zeros(4, dtype=torch.complex64)), "float64"),
]


@pytest.mark.parametrize("name,call,needle", CASES, ids=[c[0] for c in CASES])
def test_seams_reject_bad_tensors_before_the_library(no_library, name, call, needle):
    with pytest.raises(ValueError, match=needle):
        call()
