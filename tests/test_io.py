"""Input files (no GPU): the reference's ASCII triple, parsed from a read-only
mapping, and the binary triple (read, mapped and written), round-tripped on
the BCSSTK18 fixture, random matrices and edge cases, plus malformed files."""
import os
import struct

import numpy as np
import pytest

from conftest import random_matrix


def _same(A, B):
    assert A.n == B.n and A.nnz == B.nnz
    assert np.array_equal(A.colbeg, B.colbeg) and np.array_equal(A.colend, B.colend)
    assert np.array_equal(A.csr, B.csr)


def _cases(bcsstk18):
    import gsofa_amd as g
    rng = np.random.default_rng(17)
    out = [bcsstk18, g.CsrMatrix.from_rows([[0]]), g.CsrMatrix.from_rows([[], [1, 0], [2, 1]]),
           g.CsrMatrix(0, [], [], [])]
    for _ in range(4):
        out.append(random_matrix(rng, int(rng.integers(1, 500)), 0.03)[0])
    return out


def test_ascii_round_trip(bcsstk18, tmp_path):
    import gsofa_amd as g
    for k, A in enumerate(_cases(bcsstk18)):
        p = [str(tmp_path / ("%d_%s.dat" % (k, s))) for s in ("beg", "end", "csr")]
        g.write_ascii(A, *p)
        _same(g.read_ascii(*p), A)


def test_ascii_reference_layout(tmp_path):
    """The reference's csr file: nnz, n, then the indices, any whitespace;
    colbeg/colend: n integers each."""
    import gsofa_amd as g
    (tmp_path / "c.dat").write_text("5 3\n0 1\t1\r\n2\n  0\n")
    (tmp_path / "b.dat").write_text("0 2 4")
    (tmp_path / "e.dat").write_text("2\n4\n5\n")
    A = g.read_ascii(str(tmp_path / "b.dat"), str(tmp_path / "e.dat"), str(tmp_path / "c.dat"))
    assert A.n == 3 and list(A.csr) == [0, 1, 1, 2, 0]
    assert list(A.colbeg) == [0, 2, 4] and list(A.colend) == [2, 4, 5]


@pytest.mark.parametrize("csr_text", ["5 3 0 1 1 2", "5 3 0 1 x 2 0", "5 3 0 1 1- 2 0", "3", "",
                                      "5 3 0 1 1 99999999999 0"])
def test_ascii_malformed(tmp_path, csr_text):
    import gsofa_amd as g
    (tmp_path / "c.dat").write_text(csr_text)
    (tmp_path / "b.dat").write_text("0 2 4")
    (tmp_path / "e.dat").write_text("2 4 5")
    with pytest.raises(g.GsofaError) as ei:
        g.read_ascii(str(tmp_path / "b.dat"), str(tmp_path / "e.dat"), str(tmp_path / "c.dat"))
    assert ei.value.code == -5
    with pytest.raises(g.GsofaError) as ei:
        g.read_ascii(str(tmp_path / "b.dat"), str(tmp_path / "missing.dat"), str(tmp_path / "c.dat"))
    assert ei.value.code == -5


def test_binary_round_trip_and_map(bcsstk18, tmp_path):
    import gsofa_amd as g
    for k, A in enumerate(_cases(bcsstk18)):
        path = str(tmp_path / ("m%d.bin" % k))
        g.write_binary(A, path)
        # layout: 64-byte header, arrays on 64-byte boundaries
        al = lambda x: (x + 63) & ~63
        assert os.path.getsize(path) == al(al(64 + 4 * A.n) + 4 * A.n) + 4 * A.nnz
        hdr = open(path, "rb").read(64)
        magic, ver, hb, n, nnz, ob, oe, oc, _ = struct.unpack("<8sIIqqQQQQ", hdr)
        assert magic == b"GSOFACS1" and ver == 1 and hb == 64 and n == A.n and nnz == A.nnz
        assert (ob, oe, oc) == (64, al(64 + 4 * A.n), al(al(64 + 4 * A.n) + 4 * A.n))
        _same(g.read_binary(path), A)
        with g.MappedMatrix(path) as mm:
            _same(mm.matrix, A)
            assert not mm.matrix.csr.flags.writeable or A.nnz == 0
    assert not os.path.exists(str(tmp_path / "m0.bin.tmp"))


def test_binary_malformed(bcsstk18, tmp_path):
    import gsofa_amd as g
    good = str(tmp_path / "good.bin")
    g.write_binary(bcsstk18, good)
    raw = open(good, "rb").read()
    bad = {
        "magic": b"XSOFACS1" + raw[8:],
        "version": raw[:8] + struct.pack("<I", 2) + raw[12:],
        "truncated": raw[:-4],
        "short": raw[:40],
        "offsets": raw[:40] + struct.pack("<Q", 72) + raw[48:],
        "negative n": raw[:16] + struct.pack("<q", -1) + raw[24:],
    }
    for name, data in bad.items():
        p = str(tmp_path / ("bad_%s.bin" % name.replace(" ", "_")))
        open(p, "wb").write(data)
        with pytest.raises(g.GsofaError) as ei:
            g.read_binary(p)
        assert ei.value.code == -5, name
        with pytest.raises(g.GsofaError) as ei:
            g.MappedMatrix(p)
        assert ei.value.code == -5, name
    with pytest.raises(g.GsofaError) as ei:
        g.read_binary(str(tmp_path / "missing.bin"))
    assert ei.value.code == -5
