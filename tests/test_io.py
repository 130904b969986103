"""The parallel .mmn/.amn readers of libw90gpu.so (w90gpu_read_mmn,
w90gpu_read_amn), host code only.

Each file is also read by a plain sequential parse written here, record by
record as overlap_read does it (src/overlap.F90:234-308 for .mmn, :323-366
for .amn), with Python's float() for the numbers: float() rounds correctly,
so the two must agree bit for bit.  Inputs: the reference's own regression
files (tests/golden/inputs, testw90_basic1/example02/example05) and small
files that hold the forms a list-directed read accepts (d/D exponents,
exponents without a letter, commas, tabs, blank lines, signed zeros, more
than 19 significant digits), duplicate and missing blocks, and errors.
"""
import gzip
import os

import numpy as np
import pytest

from wannier90_amd import capi, w90io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
CASES = [("testw90_basic1", "wannier"), ("testw90_example02", "lead"), ("testw90_example05", "diamond")]


def _fortran_float(tok):
    """float() of a Fortran real: d/D exponents, and an exponent sign
    without a letter after the mantissa."""
    t = tok.replace("d", "e").replace("D", "e").replace("E", "e")
    if "e" not in t:
        for i in range(1, len(t)):
            if t[i] in "+-" and (t[i - 1].isdigit() or t[i - 1] == "."):
                t = t[:i] + "e" + t[i:]
                break
    return float(t)


def _records(text):
    lines = text.split("\n")
    recs = [l.replace(",", " ").split() for l in lines[1:]]
    return [r for r in recs if r]


class _Mesh:
    def __init__(self, nnlist, nncell):
        self.nnlist = np.asfortranarray(nnlist, dtype=np.int32)
        self.nncell = np.asfortranarray(nncell, dtype=np.int32)
        self.nntot = self.nnlist.shape[1]


def _mesh_from_headers(text, order=None):
    """k-mesh tables with the b order of the file's own block headers
    (or permuted per k by order[k])."""
    recs = _records(text)
    nb, nk, nn = (int(x) for x in recs[0][:3])
    nnlist = np.zeros((nk, nn), dtype=np.int32)
    nncell = np.zeros((3, nk, nn), dtype=np.int32)
    seen = np.zeros(nk, dtype=int)
    pos = 1
    for _ in range(nk * nn):
        k, k2, g0, g1, g2 = (int(x) for x in recs[pos][:5])
        j = seen[k - 1]
        b = order[k - 1][j] if order is not None else j
        nnlist[k - 1, b] = k2
        nncell[:, k - 1, b] = (g0, g1, g2)
        seen[k - 1] += 1
        pos += 1 + nb * nb
    return nb, _Mesh(nnlist, nncell)


def _py_mmn(text, nb, mesh):
    """Sequential read as overlap_read: later blocks overwrite earlier ones."""
    recs = _records(text)
    nk, nn = mesh.nnlist.shape
    m = np.zeros((nb, nb, nn, nk), dtype=np.complex128, order="F")
    pos = 1
    for _ in range(nk * nn):
        k, k2, g0, g1, g2 = (int(x) for x in recs[pos][:5])
        pos += 1
        re = np.array([_fortran_float(recs[pos + e][0]) for e in range(nb * nb)])
        im = np.array([_fortran_float(recs[pos + e][1]) for e in range(nb * nb)])
        pos += nb * nb
        hits = [b for b in range(nn) if mesh.nnlist[k - 1, b] == k2 and
                tuple(mesh.nncell[:, k - 1, b]) == (g0, g1, g2)]
        assert len(hits) == 1
        blk = np.empty(nb * nb, dtype=np.complex128)
        blk.real = re
        blk.imag = im
        m[:, :, hits[0], k - 1] = blk.reshape((nb, nb), order="F")
    return m


def _py_amn(text):
    recs = _records(text)
    nb, nk, npj = (int(x) for x in recs[0][:3])
    a = np.zeros((nb, npj, nk), dtype=np.complex128, order="F")
    for r in recs[1:1 + nb * npj * nk]:
        m, n, k = (int(x) for x in r[:3])
        z = np.empty(1, dtype=np.complex128)
        z.real = _fortran_float(r[3])
        z.imag = _fortran_float(r[4])
        a[m - 1, n - 1, k - 1] = z[0]
    return a


def _same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _unzip(tmp_path, case, seed, ext):
    src = os.path.join(INPUTS, case, f"{seed}.{ext}.gz")
    text = gzip.open(src, "rt").read()
    dst = tmp_path / f"{seed}.{ext}"
    dst.write_text(text)
    return str(dst), text


@pytest.mark.parametrize("case,seed", CASES)
@pytest.mark.parametrize("nthreads", [1, 3, 16])
def test_mmn_reference_files(tmp_path, case, seed, nthreads):
    path, text = _unzip(tmp_path, case, seed, "mmn")
    nb, mesh = _mesh_from_headers(text)
    got = capi.read_mmn(path, nb, mesh, nthreads=nthreads)
    assert _same_bits(got, _py_mmn(text, nb, mesh))


@pytest.mark.parametrize("case,seed", CASES)
def test_mmn_blocks_go_to_their_neighbour(tmp_path, case, seed):
    """b order of the k-mesh differs from the file's: every block lands at
    the b whose k' and G match its header."""
    path, text = _unzip(tmp_path, case, seed, "mmn")
    recs = _records(text)
    nk, nn = int(recs[0][1]), int(recs[0][2])
    rng = np.random.default_rng(7)
    order = [rng.permutation(nn) for _ in range(nk)]
    nb, mesh = _mesh_from_headers(text, order)
    got = capi.read_mmn(path, nb, mesh, nthreads=5)
    ref = _py_mmn(text, nb, mesh)
    assert _same_bits(got, ref)
    nb0, mesh0 = _mesh_from_headers(text)
    plain = capi.read_mmn(path, nb0, mesh0, nthreads=1)
    for k in range(nk):
        for j in range(nn):
            assert _same_bits(got[:, :, order[k][j], k], plain[:, :, j, k])


@pytest.mark.parametrize("case,seed", CASES)
@pytest.mark.parametrize("nthreads", [1, 4])
def test_amn_reference_files(tmp_path, case, seed, nthreads):
    path, text = _unzip(tmp_path, case, seed, "amn")
    nb, nk, npj = (int(x) for x in _records(text)[0][:3])
    got = capi.read_amn(path, nb, nk, npj, nthreads=nthreads)
    assert _same_bits(got, _py_amn(text))
    # the w90io view (nk, nb, nproj)
    assert _same_bits(w90io.read_amn(path, nthreads=nthreads), np.ascontiguousarray(got.transpose(2, 0, 1)))


# numbers in the forms a list-directed read accepts; several leave the
# exact fast path (long mantissas, large exponents)
FORMS = [
    "0.920873394227", "-0.330458418073", "1.5D-3", "-2.5d+02", "1.5-003", "-7.25+1", "-0.000", "0.0",
    "+.5", "5.", "1e22", "1e23", "9007199254740993", "0.1234567890123456789012", "-1.7976931348623157e308",
    "4.9e-324", "2.2250738585072014E-308", "123456789012345678901234567890", "0.30000000000000004",
    "1.00000000000000011102230246251565404236316680908203125", "7.0e-10", "  3.25", "-0.5E-0",
]


def _edge_mmn(path, rng, dup=False):
    """nb=2, nk=3, nn=2 with the forms above, blank lines, commas, tabs and
    trailing text; with dup, block (k=1, b=0) appears twice and the block of
    (k=3, b=1) is missing."""
    nb, nk, nn = 2, 3, 2
    nnlist = np.array([[2, 3], [3, 1], [1, 2]], dtype=np.int32)
    nncell = np.zeros((3, nk, nn), dtype=np.int32)
    nncell[:, 0, 1] = (0, 0, 1)
    nncell[:, 2, 0] = (-1, 0, 0)
    blocks = [(k, b) for k in range(nk) for b in range(nn)]
    if dup:
        blocks[-1] = (0, 0)
    lines = ["header comment, with commas 1 2 3", "", f"   {nb}  {nk}  {nn}  trailing"]
    i = 0
    for k, b in blocks:
        g = nncell[:, k, b]
        lines.append(f"{k + 1} {nnlist[k, b]}\t{g[0]} {g[1]} {g[2]}")
        for e in range(nb * nb):
            x = FORMS[i % len(FORMS)]
            y = FORMS[(i * 7 + 3) % len(FORMS)]
            i += 1
            sep = [" ", ", ", "\t", " ,  "][e % 4]
            tail = "   extra 1 2" if e == 1 else ""
            if dup and (k, b) == (0, 0) and i <= nb * nb:
                tail += " " * 400 + "ignored"  # the first copy fills most of the file
            lines.append(f"  {x}{sep}{y}{tail}")
            if rng.random() < 0.2:
                lines.append("   ")
    path.write_text("\n".join(lines) + "\n")
    return nb, _Mesh(nnlist, nncell)


@pytest.mark.parametrize("nthreads", [1, 2, 7, 31])
def test_mmn_fortran_forms(tmp_path, nthreads):
    p = tmp_path / "edge.mmn"
    nb, mesh = _edge_mmn(p, np.random.default_rng(1))
    got = capi.read_mmn(str(p), nb, mesh, nthreads=nthreads)
    assert _same_bits(got, _py_mmn(p.read_text(), nb, mesh))


def test_mmn_later_duplicate_wins_and_missing_is_zero(tmp_path):
    """The first copy of the duplicated block fills most of the file, so for
    2-40 ranges its elements past the first range are parsed after the
    later copy (the elements at the start of a range whose header lies in
    the range before are parsed last); the later copy must still win."""
    p = tmp_path / "dup.mmn"
    nb, mesh = _edge_mmn(p, np.random.default_rng(2), dup=True)
    runs = [capi.read_mmn(str(p), nb, mesh, nthreads=nt) for nt in range(1, 41)]
    got = runs[0]
    for r in runs[1:]:
        assert _same_bits(r, got)
    # sequential semantics written out: the last (0, 0) block wins, (2, 1) stays zero
    recs = _records(p.read_text())
    ref = np.zeros_like(got)
    pos = 1
    for _ in range(6):
        k, k2, g0, g1, g2 = (int(x) for x in recs[pos][:5])
        b = [j for j in range(2) if mesh.nnlist[k - 1, j] == k2 and
             tuple(mesh.nncell[:, k - 1, j]) == (g0, g1, g2)][0]
        blk = np.empty(nb * nb, dtype=np.complex128)
        blk.real = [_fortran_float(recs[pos + 1 + e][0]) for e in range(nb * nb)]
        blk.imag = [_fortran_float(recs[pos + 1 + e][1]) for e in range(nb * nb)]
        ref[:, :, b, k - 1] = blk.reshape((nb, nb), order="F")
        pos += 1 + nb * nb
    assert _same_bits(got, ref)
    assert not np.any(got[:, :, 1, 2])


def test_mmn_many_ranges_synthetic(tmp_path):
    """A larger file (17-digit numbers, the slow conversion path) read with
    1 and 16 threads, against the sequential parse."""
    from wannier90_amd import synth
    rng = np.random.default_rng(3)
    nb = 5
    prob = synth.config(1)
    mesh = prob.kmesh
    nk, nn = mesh.nnlist.shape
    m = rng.standard_normal((nb, nb, nn, nk)) + 1j * rng.standard_normal((nb, nb, nn, nk))
    p = tmp_path / "syn.mmn"
    w90io.write_mmn(str(p), np.asfortranarray(m), mesh)
    a = capi.read_mmn(str(p), nb, mesh, nthreads=1)
    b = capi.read_mmn(str(p), nb, mesh, nthreads=16)
    assert _same_bits(a, b)
    assert _same_bits(a, np.asfortranarray(m))  # %.17e round-trips every double
    assert _same_bits(w90io.read_mmn(str(p), mesh), a)


def test_mmn_errors(tmp_path):
    p = tmp_path / "e.mmn"
    nb, mesh = _edge_mmn(p, np.random.default_rng(4))
    with pytest.raises(capi.W90Error, match="not the right number of bands"):
        capi.read_mmn(str(p), nb + 1, mesh)
    bad = _Mesh(mesh.nnlist[:2, :], mesh.nncell[:, :2, :])
    with pytest.raises(capi.W90Error, match="not the right number of k-points"):
        capi.read_mmn(str(p), nb, bad)
    with pytest.raises(capi.W90Error, match="Problem opening input file"):
        capi.read_mmn(str(tmp_path / "absent.mmn"), nb, mesh)
    # a neighbour the k-mesh does not have
    wrong = _Mesh(mesh.nnlist.copy(), mesh.nncell.copy())
    wrong.nncell[:, 1, 1] = (5, 5, 5)
    with pytest.raises(capi.W90Error, match="Neighbour not found"):
        capi.read_mmn(str(p), nb, wrong, nthreads=3)
    # two entries of the k-mesh match one block
    twice = _Mesh(mesh.nnlist.copy(), mesh.nncell.copy())
    twice.nnlist[1, 1] = twice.nnlist[1, 0]
    twice.nncell[:, 1, 1] = twice.nncell[:, 1, 0]
    with pytest.raises(capi.W90Error, match="More than one matching nearest neighbour"):
        capi.read_mmn(str(p), nb, twice, nthreads=2)
    # a truncated file and an unreadable number
    lines = p.read_text().split("\n")
    (tmp_path / "short.mmn").write_text("\n".join(lines[:-6]) + "\n")
    with pytest.raises(capi.W90Error, match="Problem reading input file"):
        capi.read_mmn(str(tmp_path / "short.mmn"), nb, mesh)
    txt = p.read_text().replace("0.920873394227", "0.92x", 1)
    (tmp_path / "garbled.mmn").write_text(txt)
    with pytest.raises(capi.W90Error, match="Problem reading input file"):
        capi.read_mmn(str(tmp_path / "garbled.mmn"), nb, mesh, nthreads=4)


def test_amn_order_duplicates_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    nb, npj, nk = 3, 2, 4
    recs = [(m, n, k) for k in range(1, nk + 1) for n in range(1, npj + 1) for m in range(1, nb + 1)]
    rng.shuffle(recs)
    recs[-1] = recs[0]  # the first record repeated last: it wins, one entry stays zero
    lines = ["amn comment", f"{nb} {nk} {npj}"]
    for i, (m, n, k) in enumerate(recs):
        lines.append(f"{m:5d}{n:5d}{k:5d} {FORMS[i % len(FORMS)]} {FORMS[(3 * i + 1) % len(FORMS)]}")
    p = tmp_path / "x.amn"
    p.write_text("\n".join(lines) + "\n")
    for nt in (1, 3, 8):
        got = capi.read_amn(str(p), nb, nk, npj, nthreads=nt)
        assert _same_bits(got, _py_amn(p.read_text()))
    assert np.count_nonzero(got == 0) >= 1
    with pytest.raises(capi.W90Error, match="not the right number of projections"):
        capi.read_amn(str(p), nb, nk, npj + 1)
    (tmp_path / "oob.amn").write_text(p.read_text().replace(f"{recs[2][0]:5d}{recs[2][1]:5d}{recs[2][2]:5d}",
                                                            f"{nb + 1:5d}{recs[2][1]:5d}{recs[2][2]:5d}", 1))
    with pytest.raises(capi.W90Error, match="Problem reading input file"):
        capi.read_amn(str(tmp_path / "oob.amn"), nb, nk, npj)
