"""States to and from files in qpp's formats (include/qpp/input_output.hpp).

The expected bytes are written from the formats' definitions: text =
"rows cols" then one line per row of "(re,im)" entries separated by a space,
reals with max_digits10 = 17 significant digits (input_output.hpp:177-211,
internal/util.hpp:440-446); binary = "TYPE::Eigen::Matrix", rows and cols as
8-byte size_t, column-major complex<double> (input_output.hpp:318-345).
"""
import struct

import numpy as np
import pytest

from helpers import randket, randrho

pytestmark = pytest.mark.gpu


def qpp_text(A):
    """qpp::save's text for a matrix (a ket is D x 1), written independently
    from the definition above."""
    A = np.atleast_2d(np.asarray(A, dtype=complex).T).T if np.ndim(A) == 1 else np.asarray(A)
    rows, cols = A.shape
    out = [f"{rows} {cols}\n"]
    for i in range(rows):
        out.append(" ".join(f"({z.real:.17g},{z.imag:.17g})" for z in A[i]) + "\n")
    return "".join(out)


def test_text_closed_form(qa, tmp_path):
    s = np.sqrt(0.5)  # the double nearest 1/sqrt(2) (1/np.sqrt(2) is one ulp below)
    st =qa.State.from_array(np.array([s, 1j * s]), [2])
    f = tmp_path / "ket.txt"
    st.save(f)
    assert f.read_text() == "2 1\n(0.70710678118654757,0)\n(0,0.70710678118654757)\n"
    rho = np.array([[0.5, -0.25j], [0.25j, 0.5]])
    qa.State.from_array(rho, [2]).save(tmp_path / "rho.txt")
    assert (tmp_path / "rho.txt").read_text() == "2 2\n(0.5,0) (-0,-0.25)\n(0,0.25) (0.5,0)\n"


@pytest.mark.parametrize("kind", ["ket", "rho"])
def test_text_bytes_and_round_trip(qa, tmp_path, kind):
    rng = np.random.default_rng(21)
    A = randket(2 ** 6, rng) if kind == "ket" else randrho(2 ** 4, rng)
    n = 6 if kind == "ket" else 4
    st = qa.State.from_array(A, [2] * n)
    f = tmp_path / "a.txt"
    st.save(f, qa.TEXT)
    assert f.read_text() == qpp_text(A if kind == "rho" else A.reshape(-1, 1))
    back = qa.State.load(f, [2] * n, qa.KET if kind == "ket" else qa.RHO).download()
    assert np.array_equal(back, A)  # 17 significant digits round-trip exactly


@pytest.mark.parametrize("kind", ["ket", "rho"])
def test_binary_bytes_and_round_trip(qa, tmp_path, kind):
    rng = np.random.default_rng(22)
    n = 7 if kind == "ket" else 4
    A = randket(2 ** n, rng) if kind == "ket" else randrho(2 ** n, rng)
    f = tmp_path / "a.dat"
    qa.save(A, f, [2] * n, qa.BINARY)
    rows, cols = (A.shape[0], 1) if kind == "ket" else A.shape
    exp = b"TYPE::Eigen::Matrix" + struct.pack("<QQ", rows, cols) + \
        np.asarray(A, dtype=np.complex128).tobytes(order="F")
    assert f.read_bytes() == exp
    back = qa.load(f, [2] * n, qa.KET if kind == "ket" else qa.RHO, qa.BINARY)
    assert np.array_equal(back, A)


def test_binary_streams_in_chunks(qa, tmp_path):
    """A 23-qubit ket (2^23 amplitudes) crosses the 2^22-element transfer
    chunks on both save and load."""
    n = 23
    rng = np.random.default_rng(23)
    A = (rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)) / np.sqrt(2 ** (n + 1))
    st = qa.State.from_array(A, [2] * n)
    f = tmp_path / "big.dat"
    st.save(f, qa.BINARY)
    raw = f.read_bytes()
    assert raw[:19] == b"TYPE::Eigen::Matrix" and struct.unpack("<QQ", raw[19:35]) == (2 ** n, 1)
    assert np.array_equal(np.frombuffer(raw[35:], dtype=np.complex128), A)
    assert np.array_equal(qa.State.load(f, [2] * n, fmt=qa.BINARY).download(), A)


def test_complex64_states(qa, tmp_path):
    rng = np.random.default_rng(24)
    A = randket(2 ** 8, rng)
    st = qa.State.from_array(A, [2] * 8).astype(qa.C64)
    f = tmp_path / "c64.dat"
    st.save(f, qa.BINARY)
    got = np.frombuffer(f.read_bytes()[35:], dtype=np.complex128)
    assert np.array_equal(got, A.astype(np.complex64).astype(np.complex128))
    back = qa.State.load(f, [2] * 8, dtype=qa.C64, fmt=qa.BINARY)
    assert np.array_equal(back.download(), A.astype(np.complex64).astype(np.complex128))


def test_qutrits_and_whitespace(qa, tmp_path):
    """qpp::load reads with operator>> (whitespace-insensitive); dims of any
    subsystem sizes."""
    f = tmp_path / "q.txt"
    f.write_text("9 1\n(1,0)\n (0,0)\n(0,0) (0,0)\n(0,-0.5)\n(0,0)\n(0,0)\n(0,0)\n(0.25,0)\n")
    got = qa.load(f, [3, 3])
    exp = np.zeros(9, dtype=complex)
    exp[0], exp[4], exp[8] = 1, complex(0, -0.5), 0.25
    assert np.array_equal(got, exp)


def test_errors(qa, tmp_path):
    with pytest.raises(qa.RuntimeError_, match=r"qpp::load\(\): Error opening input stream!"):
        qa.load(tmp_path / "missing.txt", [2])
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"NOT::Eigen::Matrix!" + struct.pack("<QQ", 2, 1) + bytes(32))
    with pytest.raises(qa.RuntimeError_, match=r"qpp::obsolete::load\(\): Input stream is corrupted!"):
        qa.load(bad, [2], fmt=qa.BINARY)
    short = tmp_path / "short.dat"
    short.write_bytes(b"TYPE::Eigen::Matrix" + struct.pack("<QQ", 4, 1) + bytes(16))
    with pytest.raises(qa.RuntimeError_):
        qa.load(short, [2, 2], fmt=qa.BINARY)
    f = tmp_path / "k.txt"
    qa.save(np.array([1, 0, 0, 0], dtype=complex), f, [2, 2])
    with pytest.raises(qa.DimsMismatchCvector):
        qa.load(f, [2, 2, 2])
    with pytest.raises(qa.DimsMismatchMatrix):
        qa.load(f, [2, 2], qa.RHO)
