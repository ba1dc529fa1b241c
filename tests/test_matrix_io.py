"""BitMatrix text / BMMPYB1 binary formats and the canonical hash (host code,
no device): against the reference's own outputs (tests/golden/golden_io.npz,
written by the reference build), against hashlib's SHA-256 of the byte
stream the hash is defined over, and against the live reference where it is
built here."""
import hashlib
import os
import struct

import numpy as np
import pytest

import bmmpy as bmm
import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gio():
    return np.load(os.path.join(GOLD, "golden_io.npz"), allow_pickle=False)


def _matrix(words, cols):
    """A BitMatrix holding exactly these row words (stray tail bits kept)."""
    words = np.asarray(words, np.uint64)
    m = bmm.BitMatrix(words.shape[0], cols)
    padded = np.zeros((words.shape[0], m.stride_words), np.uint64)
    padded[:, :words.shape[1]] = words
    m.copy_from_words([int(v) for v in padded.reshape(-1)])
    return m


def _closed_form_hash(words, rows, cols):
    w = (cols + 63) // 64
    last = (1 << (cols % 64)) - 1 if cols % 64 else (1 << 64) - 1
    data = struct.pack("<QQ", rows, cols)
    for r in range(rows):
        for q in range(w):
            v = int(words[r, q])
            data += struct.pack("<Q", v & last if q == w - 1 else v)
    return hashlib.sha256(data).hexdigest()


def test_sha256_known_answers():
    # FIPS 180-4 examples through the matrix hash's byte stream: an empty
    # 0 x 0 matrix hashes the 16 zero bytes of its shape.
    assert bmm.BitMatrix(0, 0).hash() == hashlib.sha256(bytes(16)).hexdigest()
    # Multi-block messages (> 55 and > 64 bytes) against hashlib.
    for rows, cols in ((1, 64 * 5), (3, 64 * 7), (9, 1000), (40, 65)):
        m = oracle.oracle().bernoulli(rows * 7 + cols, rows, cols, 0.5)
        assert _matrix(m, cols).hash() == _closed_form_hash(m, rows, cols)


def test_golden_writers_and_hash(gio, tmp_path):
    for q in range(int(gio["io_cases"][0])):
        words, cols = gio[f"io{q}_in"], int(gio[f"io{q}_cols"][0])
        m = _matrix(words, cols)
        assert m.hash() == str(gio[f"io{q}_hash"][0]), q
        assert m.hash() == _closed_form_hash(words, words.shape[0], cols), q
        assert m.to_text() == str(gio[f"io{q}_text"][0]) == str(m), q
        path = tmp_path / f"m{q}.bmm"
        m.save_binary_file(path)
        assert path.read_bytes() == gio[f"io{q}_bin"].tobytes(), q
        back = bmm.BitMatrix.load_binary_file(str(path))
        assert back.shape == (words.shape[0], cols)
        assert np.array_equal(np.asarray(back.to_words(), np.uint64).reshape(words.shape), words)
        m.save_text_file(tmp_path / f"m{q}.txt")
        t = bmm.BitMatrix.load_text_file(tmp_path / f"m{q}.txt")
        assert t.hash() == m.hash() and t.to_rows() == m.to_rows()
        assert bmm.BitMatrix.from_text(m.to_text()).hash() == m.hash()


def test_hash_is_canonical_and_data_sensitive():
    # tests/cpp/bit_matrix_hash_tests.cpp:11-31.
    canonical = bmm.BitMatrix(1, 65)
    canonical[0, 0] = True
    canonical[0, 64] = True
    noisy = bmm.BitMatrix(1, 65)
    noisy.copy_from_words([0x1, 0x1 | (1 << 17), 0xAAAAAAAAAAAAAAAA, 0x5555555555555555])
    assert canonical.hash() == noisy.hash()
    changed = noisy.copy()
    changed.set(0, 1, True)
    assert changed.hash() != noisy.hash()


def test_malformed_inputs_match_reference_messages(gio, tmp_path):
    for q in range(int(gio["bad_cases"][0])):
        kind = int(gio[f"bad{q}_kind"][0])
        data = gio[f"bad{q}_data"].tobytes()
        msg = str(gio[f"bad{q}_msg"][0])
        path = tmp_path / f"bad{q}"
        path.write_bytes(data)
        load = bmm.BitMatrix.load_text_file if kind else bmm.BitMatrix.load_binary_file
        if msg:
            with pytest.raises(RuntimeError) as e:
                load(path)
            assert str(e.value) == msg, q
            if kind:
                with pytest.raises(RuntimeError, match=msg):
                    bmm.BitMatrix.from_text(data.decode())
        else:
            m = load(path)
            r, c = (int(v) for v in gio[f"bad{q}_shape"])
            assert m.shape == (r, c)
            w = gio[f"bad{q}_words"]
            assert np.array_equal(np.asarray(m.to_words(), np.uint64).reshape(w.shape), w)


def test_missing_files_and_pathlike(tmp_path):
    with pytest.raises(RuntimeError, match="failed to open matrix binary file for reading"):
        bmm.BitMatrix.load_binary_file(tmp_path / "none.bmm")
    with pytest.raises(RuntimeError, match="failed to open matrix text file for reading"):
        bmm.BitMatrix.load_text_file(str(tmp_path / "none.txt"))
    with pytest.raises(RuntimeError, match="failed to open matrix binary file for writing"):
        bmm.BitMatrix(1, 1).save_binary_file(tmp_path / "no" / "dir.bmm")


def test_live_reference_random(ref, tmp_path):
    ora = oracle.oracle()
    for rows, cols in ((1, 1), (2, 63), (3, 64), (17, 129), (33, 777)):
        m = ora.bernoulli(rows * 1000 + cols, rows, cols, 0.5)
        # stray bits past the last column, as the reference keeps them
        if cols % 64:
            m[:, -1] |= np.uint64(1 << 63)
        h, b, t = ref.matrix_io(m, cols)
        pm = _matrix(m, cols)
        assert pm.hash() == h and pm.to_text() == t
        pm.save_binary_file(tmp_path / "x.bmm")
        assert (tmp_path / "x.bmm").read_bytes() == b
        words, r, c = ref.matrix_load(b)
        back = bmm.BitMatrix.load_binary_file(tmp_path / "x.bmm")
        assert np.array_equal(np.asarray(back.to_words(), np.uint64), words)


def test_host_algebra_helpers():
    a = bmm.matrix_from_rows(["1100", "0110", "0011"])
    b = bmm.matrix_from_rows(["1010", "1010", "1111"])
    assert a.add(b).to_rows() == ["0110", "1100", "1100"]
    with pytest.raises(bmm.MatrixError, match="Matrix dimensions mismatch"):
        a.add(bmm.BitMatrix(2, 4))
    e = a.extract_rows_by_indices([2, 0])
    assert e.to_rows() == ["0011", "1100"]
    with pytest.raises(IndexError):
        a.extract_rows_by_indices([3])
    a.insert_rows_by_indices(bmm.matrix_from_rows(["1111"]), [1])
    assert a.to_rows() == ["1100", "1111", "0011"]
    with pytest.raises(bmm.MatrixError):
        a.insert_rows_by_indices(bmm.matrix_from_rows(["1111", "0000"]), [1])
    assert a.total_bytes == a.total_words * 8
    assert repr(a) == "BitMatrix(rows=3, cols=4, weight=8)"
    with pytest.raises(bmm.MatrixError):
        bmm.BitMatrix(2, 3).power(2)
    assert bmm.MatrixErr.SizeTooLarge != bmm.MatrixErr.DimensionMismatch
    with pytest.raises(ValueError, match="all rows must have equal width"):
        bmm.matrix_from_rows(["10", "1"])


def test_batch_files_round_trip(tmp_path):
    from bmmpy_amd import _lib, batch
    rows, cols, dens, words, offs = _lib.gen_batch(0x5EED0004, 24)
    paths = batch.save_batch(tmp_path, words, offs, rows, cols)
    w2, o2, r2, c2 = batch.load_batch(paths)
    assert np.array_equal(r2, rows) and np.array_equal(c2, cols)
    for i in range(rows.size):
        n = int(rows[i]) * ((int(cols[i]) + 63) // 64)
        assert np.array_equal(w2[o2[i]:o2[i] + n], words[offs[i]:offs[i] + n])
    assert batch.batch_hashes(w2, o2, r2, c2) == batch.batch_hashes(words, offs, rows, cols)
