"""Field files (th_field_write / th_field_read: the reference's field_out.d of Output_field and the
field_in.d that Init_cond reads with field_input_proc = 1; th_field_write_reduced / th_field_read_reduced:
field_out_reduced.d of Output_field_reduced and the reduced field_in.d of Init_cond_reduced).  The format is
pinned by a program that writes and reads through the C++ standard library's complex<double> stream
operators, as Write_data_MPI / Read_data_MPI do (array_basic.cc:294-335, 814-858, 997-1170): fixed notation
with 12 decimals for field_out.d (file_operations.cc:229-230), the stream's default 6 significant digits and
no separator for the reduced file.  Files must be equal byte for byte, and each side must read the other's
file to the printed digits.  The reduced layout is pinned by what the mathematics fixes: at Nr = N it is the
full array in order, and a reduced field read into a larger grid keeps every wavenumber, so its inverse
transform sampled on the coarse points is the coarse field's.  CPU only."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ref_io(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("refio") / "ref_field_io")
    subprocess.check_call([gxx, "-O1", "-std=c++11", "-o", exe, os.path.join(HERE, "golden", "ref_field_io.cpp")])
    return exe


def _fields(N, ncomp, seed):
    rng = np.random.default_rng(seed)
    shape = (N[0], N[1], N[2] // 2 + 1)
    out = []
    for c in range(ncomp):
        mag = 10.0 ** rng.uniform(-14, 6, size=shape)
        a = (rng.standard_normal(shape) * mag + 1j * rng.standard_normal(shape) * mag).astype(np.complex128)
        out.append(a)
    # values whose printed form has special cases: integers, negative zero, exact zero, large exponents
    out[0][0, 0, 0] = 2.0 + 0.0j
    out[0][0, 0, 1] = complex(-0.0, 0.0)
    out[-1][-1, -1, -1] = 1.2345678901234567e+200 - 1e-300j
    return out


def _printed(a):
    """The value the fixed 12-decimal text carries."""
    f = np.vectorize(lambda x: float(f"{x:.12f}"))
    return f(a.real) + 1j * f(a.imag)


def _printed6(a):
    """The value the default-format (6 significant digits) text carries."""
    f = np.vectorize(lambda x: float(f"{x:g}"))
    return f(a.real) + 1j * f(a.imag)


def _parse_values(text):
    """Every "(re,im)" of a field file body in order (the "%%" lines skipped)."""
    import re
    body = "\n".join(l for l in text.splitlines() if not l.startswith("%%"))
    return np.array([complex(float(a), float(b)) for a, b in re.findall(r"\(([^,()]+),([^,()]+)\)", body)])


def _kx(i, n):
    return i if i <= n // 2 else i - n


def test_reduced_at_full_size_is_the_whole_array_in_order(tmp_path):
    """Nr = N: the reduced file lists every mode of the array in its storage order (FOUR and SCFT)."""
    from tarang_amd import field_write_reduced
    N = (6, 4, 8)
    V = _fields(N, 2, 5)
    for basis in ("FOUR", "SCFT"):
        p = tmp_path / f"red_{basis}.d"
        field_write_reduced(p, N, N, V, time=0.5, basis=basis)
        txt = p.read_text()
        assert txt.startswith("%% Time = 0.5\n")
        assert ") (" not in txt and ")(" in txt          # no separator between values
        vals = _parse_values(txt)
        assert np.array_equal(vals, _printed6(np.concatenate([a.ravel() for a in V])))


def test_reduced_writer_matches_the_iostream_writer_byte_for_byte(tmp_path, ref_io):
    from tarang_amd import field_write_reduced
    N, Nr = (8, 6, 10), (4, 4, 6)
    V = _fields(N, 3, 6)
    ours = tmp_path / "ours.d"
    field_write_reduced(ours, N, Nr, V, time=0.03125)
    # file order of the reduced grid: rows kx = 0..Nr1/2, -Nr1/2+1..-1 and columns ky likewise, kz <= Nr3/2
    rows = [i for i in range(N[0]) if _kx(i, N[0]) in [_kx(r, Nr[0]) for r in range(Nr[0])] and
            not (_kx(i, N[0]) == -Nr[0] // 2)]
    rows = sorted(rows, key=lambda i: [_kx(r, Nr[0]) for r in range(Nr[0])].index(_kx(i, N[0])))
    cols = sorted([j for j in range(N[1]) if _kx(j, N[1]) in [_kx(c, Nr[1]) for c in range(Nr[1])] and
                   not (_kx(j, N[1]) == -Nr[1] // 2)],
                  key=lambda j: [_kx(c, Nr[1]) for c in range(Nr[1])].index(_kx(j, N[1])))
    sub = [a[np.ix_(rows, cols, range(Nr[2] // 2 + 1))] for a in V]
    raw = tmp_path / "v.bin"
    np.concatenate([a.ravel() for a in sub]).view(np.float64).tofile(raw)
    ref = tmp_path / "ref.d"
    subprocess.check_call([ref_io, "writered", str(raw), str(ref), str(sub[0].size), "3", "0.03125"])
    assert ours.read_bytes() == ref.read_bytes()


@pytest.mark.parametrize("basis", ["FOUR", "SCFT"])
def test_reduced_read_keeps_every_wavenumber(tmp_path, basis):
    """A coarse field written at its own size (th_field_write) and read with th_field_read_reduced into a finer
    grid keeps each mode at its wavenumber: the fine inverse transform sampled on the coarse points equals the
    coarse one (Nyquist planes zero); the coarse Nyquist planes land on both sides of the fine grid; every other
    mode is zero."""
    from tarang_amd import field_read_reduced, field_write
    Nr, N = (8, 6, 10), (16, 12, 20)
    rng = np.random.default_rng(11)
    c = rng.standard_normal((Nr[0], Nr[1], Nr[2] // 2 + 1)) + 1j * rng.standard_normal((Nr[0], Nr[1], Nr[2] // 2 + 1))
    p = tmp_path / "coarse.d"
    field_write(p, Nr, [c], time=2.0)
    (f,), t = field_read_reduced(p, N, Nr, 1, basis=basis)
    assert t == 2.0
    cp = _printed(c)
    hits = np.zeros(f.shape, bool)
    for i in range(Nr[0]):
        kx = i if basis == "SCFT" else _kx(i, Nr[0])
        fis = [kx if kx >= 0 else N[0] + kx]
        if basis == "FOUR" and kx == Nr[0] // 2:
            fis.append(N[0] - Nr[0] // 2)
        for j in range(Nr[1]):
            ky = _kx(j, Nr[1])
            fjs = [ky if ky >= 0 else N[1] + ky] + ([N[1] - Nr[1] // 2] if ky == Nr[1] // 2 else [])
            for fi in fis:
                for fj in fjs:
                    assert np.array_equal(f[fi, fj, :Nr[2] // 2 + 1], cp[i, j]), (i, j, fi, fj)
                    hits[fi, fj, :Nr[2] // 2 + 1] = True
    assert not np.any(f[~hits])
    if basis == "FOUR":
        z = cp.copy()
        z[Nr[0] // 2] = 0
        z[:, Nr[1] // 2] = 0
        z[:, :, Nr[2] // 2] = 0
        q = tmp_path / "coarse0.d"
        field_write(q, Nr, [z], time=2.0)
        (g,), _ = field_read_reduced(q, N, Nr, 1)
        fine = np.fft.irfftn(g, s=N, axes=(0, 1, 2)) * np.prod(N)
        coarse = np.fft.irfftn(z, s=Nr, axes=(0, 1, 2)) * np.prod(Nr)
        sx, sy, sz = N[0] // Nr[0], N[1] // Nr[1], N[2] // Nr[2]
        assert np.abs(fine[::sx, ::sy, ::sz] - coarse).max() <= 1e-12 * np.abs(coarse).max()


def test_reduced_round_trip_and_refusals(tmp_path):
    from tarang_amd import field_read_reduced, field_write_reduced
    N, Nr = (8, 8, 8), (4, 6, 4)
    V = _fields(N, 3, 9)
    p = tmp_path / "r.d"
    field_write_reduced(p, N, Nr, V, time=1.0)
    field_write_reduced(p, N, Nr, [2 * a for a in V], time=2.0, append=True)
    got, t = field_read_reduced(p, N, Nr, 3, snapshot=1)
    assert t == 2.0
    kept = [(i, j) for i in range(N[0]) for j in range(N[1])
            if -Nr[0] // 2 <= _kx(i, N[0]) <= Nr[0] // 2 and -Nr[1] // 2 <= _kx(j, N[1]) <= Nr[1] // 2]
    for a, b in zip(got, V):
        for i, j in kept:
            src_i = i if _kx(i, N[0]) != -Nr[0] // 2 else Nr[0] // 2
            src_j = j if _kx(j, N[1]) != -Nr[1] // 2 else Nr[1] // 2
            assert np.array_equal(a[i, j, :Nr[2] // 2 + 1], _printed6(2 * b[src_i, src_j, :Nr[2] // 2 + 1]))
    for bad in ((3, 6, 4), (4, 6, 10), (0, 6, 4)):
        with pytest.raises(RuntimeError, match="bad args"):
            field_write_reduced(p, N, bad, V)
        with pytest.raises(RuntimeError, match="bad args"):
            field_read_reduced(p, N, bad, 3)


def test_writer_matches_the_iostream_writer_byte_for_byte(tmp_path, ref_io):
    from tarang_amd import field_write
    N, ncomp = (4, 6, 10), 3
    V = _fields(N, ncomp, 1)
    ours = tmp_path / "ours.d"
    field_write(ours, N, V, time=0.125)
    raw = tmp_path / "v.bin"
    np.concatenate([a.ravel() for a in V]).view(np.float64).tofile(raw)
    ref = tmp_path / "ref.d"
    n = V[0].size
    subprocess.check_call([ref_io, "write", str(raw), str(ref), str(n), str(ncomp), "0.125"])
    assert ours.read_bytes() == ref.read_bytes()


def test_each_side_reads_the_other(tmp_path, ref_io):
    from tarang_amd import field_read, field_write
    N, ncomp = (6, 4, 8), 4
    V = _fields(N, ncomp, 2)
    n = V[0].size
    raw = tmp_path / "v.bin"
    np.concatenate([a.ravel() for a in V]).view(np.float64).tofile(raw)
    ref = tmp_path / "ref.d"
    subprocess.check_call([ref_io, "write", str(raw), str(ref), str(n), str(ncomp), "3.5"])
    got, t = field_read(ref, N, ncomp)
    assert t == 3.5
    for a, b in zip(got, V):
        assert np.array_equal(a, _printed(b))
    ours = tmp_path / "ours.d"
    field_write(ours, N, V, time=3.5)
    back = tmp_path / "back.bin"
    subprocess.check_call([ref_io, "read", str(ours), str(back), str(n), str(ncomp)])
    vals = np.fromfile(back, np.float64).view(np.complex128).reshape(ncomp, *V[0].shape)
    for c in range(ncomp):
        assert np.array_equal(vals[c], _printed(V[c]))


def test_snapshots_fp32_and_accepted_value_forms(tmp_path):
    from tarang_amd import field_read, field_write
    N = (2, 2, 4)
    A = [np.arange(12, dtype=np.float32).reshape(2, 2, 3).astype(np.complex64) * (1 + 0.5j)]
    B = [(a * -3).astype(np.complex64) for a in A]
    p = tmp_path / "field_out.d"
    field_write(p, N, A, time=1.0)
    field_write(p, N, B, time=2.0, append=True)
    (a,), ta = field_read(p, N, 1, snapshot=0, precision="fp32")
    (b,), tb = field_read(p, N, 1, snapshot=1, precision="fp32")
    assert (ta, tb) == (1.0, 2.0)
    assert np.array_equal(a, A[0]) and np.array_equal(b, B[0])
    # istream >> complex also takes "(re)" and a bare "re"; no "%%" line gives time NaN
    q = tmp_path / "field_in.d"
    q.write_text("(1,2) (3) 4\t(-5e-3 , 6E2)\n\n(0,0) 7 (8,9) (10)\n")
    (c,), tc = field_read(q, (2, 1, 2), 1)
    assert math.isnan(tc)
    assert np.array_equal(c.ravel(), np.array([1 + 2j, 3, 4, -5e-3 + 600j]))


def test_short_or_malformed_file_is_an_error(tmp_path):
    from tarang_amd import field_read
    p = tmp_path / "bad.d"
    p.write_text("%% Time = 1\n(1,2) (3,4) (5,6\n")
    with pytest.raises(RuntimeError, match="malformed"):
        field_read(p, (2, 1, 2), 1)
    p.write_text("%% Time = 1\n(1,2) (3,4)\n")
    with pytest.raises(RuntimeError, match="missing"):
        field_read(p, (2, 2, 2), 1)
    with pytest.raises(RuntimeError, match="cannot open"):
        field_read(tmp_path / "absent.d", (2, 2, 2), 1)
