"""Field files (th_field_write / th_field_read: the reference's field_out.d of Output_field and the
field_in.d that Init_cond reads with field_input_proc = 1).  The format is pinned by a program that
writes and reads through the C++ standard library's complex<double> stream operators at precision
12, as Write_data_MPI / Read_data_MPI do (array_basic.cc:294-335, 814-858): files must be equal
byte for byte, and each side must read the other's file to the printed 12 digits.  CPU only."""
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
    """The value the 12-digit text carries."""
    f = np.vectorize(lambda x: float(f"{x:.12g}"))
    return f(a.real) + 1j * f(a.imag)


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
