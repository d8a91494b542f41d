"""Host side of frame ingestion (no GPU): the library's DCD reader against
files written to the format trajectory-formats/dcd.f90 reads (both byte
orders, with and without unit-cell records, truncated and foreign files),
and the PermIndex reader's rules (setconf.F90:866-905)."""
import numpy as np
import pytest

from ermod_amd import api as A
from ermod_amd import synth


def frames(nf=4, n=37, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-30, 60, (nf, n, 3)).astype(np.float32)
    cells = [np.diag(rng.uniform(20, 40, 3)) for _ in range(nf)]
    return x, cells


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("with_cell", [True, False])
def test_dcd_roundtrip(tmp_path, big, with_cell):
    x, cells = frames()
    p = tmp_path / "t.dcd"
    synth.write_dcd(p, x, cells if with_cell else None, big_endian=big)
    got, gc = A.read_dcd(p, periodic=with_cell)
    assert got.dtype == np.float32 and np.array_equal(got, x)
    if with_cell:
        np.testing.assert_array_equal(gc, np.array(cells))
    else:
        assert not gc.any()


def test_dcd_first_atoms_of_longer_records(tmp_path):
    x, cells = frames(n=50)
    p = tmp_path / "t.dcd"
    synth.write_dcd(p, x, cells)
    got, _ = A.read_dcd(p, natom=20)
    assert np.array_equal(got, x[:, :20])


def test_dcd_errors(tmp_path):
    x, cells = frames()
    p = tmp_path / "t.dcd"
    synth.write_dcd(p, x, cells)
    with pytest.raises(A.ErmError):          # the cell record is requested but absent
        q = tmp_path / "nocell.dcd"
        synth.write_dcd(q, x, None)
        A.read_dcd(q, periodic=True)
    raw = p.read_bytes()
    (tmp_path / "cut.dcd").write_bytes(raw[:-30])
    with pytest.raises(A.ErmError):          # the last frame is cut short
        A.read_dcd(tmp_path / "cut.dcd")
    (tmp_path / "bad.dcd").write_bytes(raw[:4] + b"XORD" + raw[8:])
    with pytest.raises(A.ErmError):          # no CORD magic
        A.read_dcd(tmp_path / "bad.dcd")


def test_permutation_rules(tmp_path):
    n = 6
    assert A.read_permutation(tmp_path / "missing", n) is None
    p = tmp_path / "PermIndex"
    p.write_text("1 3\n3 1\n\n5 6\n6 5\n")
    assert list(A.read_permutation(p, n)) == [2, 1, 0, 3, 5, 4]
    p.write_text("1 3\n")                     # atom 3 targeted twice (1 -> 3, 3 -> 3)
    with pytest.raises(A.ErmError):
        A.read_permutation(p, n)
    p.write_text("7 1\n")                     # out of range
    with pytest.raises(A.ErmError):
        A.read_permutation(p, n)
