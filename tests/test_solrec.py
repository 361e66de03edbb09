"""CPU: trajectory spill files in the reference's SolRec layout
(SolRec::wrtie_sol_to_disk dualmesh.C:87-142, read_sol_from_disk
dualmesh.C:234-297).  The writer is checked against an independent parse
(Python gzip + struct), the reader against files packed by hand with
struct, so a swapped count, a missing key or a transposed record fails one
of them."""
import gzip
import struct

import numpy as np
import pytest

from discrete_adjoint_amd import mesh


def _parse(path):
    raw = gzip.open(path, "rb").read()
    nz, nnz = struct.unpack_from("<ii", raw, 0)
    off = 8
    zero = [struct.unpack_from("<II", raw, off + 8 * i) for i in range(nz)]
    off += 8 * nz
    rec = [struct.unpack_from("<IIddd", raw, off + 32 * i) for i in range(nnz)]
    assert off + 32 * nnz == len(raw)
    return zero, rec


def _case(n=40, seed=5):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2**32, size=(n, 2), dtype=np.uint64).astype(np.uint32)
    h = rng.uniform(0.0, 2.0, n)
    h[::7] = 0.0
    h[3] = -0.0
    h[5] = -1e-3
    h[11] = np.nan
    hu, hv = rng.standard_normal(n), rng.standard_normal(n)
    return keys, (h, hu, hv)


def test_writer_layout(tmp_path):
    keys, U = _case()
    h = U[0]
    order = np.random.default_rng(1).permutation(len(h)).astype(np.int32)
    p = tmp_path / "solution_0000_00000007"
    mesh.write_solution(p, keys, U, order=order)
    zero, rec = _parse(p)
    # record_solution: h > 0 keeps the state, anything else is a zero entry
    wet = [e for e in order if h[e] > 0]
    dry = [e for e in order if not h[e] > 0]
    assert [tuple(keys[e]) for e in dry] == zero
    assert len(rec) == len(wet)
    for e, r in zip(wet, rec):
        assert r[:2] == tuple(keys[e])
        assert r[2:] == (h[e], U[1][e], U[2][e])  # exact FP64 bytes
    # order None is index order
    mesh.write_solution(p, keys, U)
    zero2, rec2 = _parse(p)
    assert zero2 == [tuple(keys[e]) for e in range(len(h)) if not h[e] > 0]
    assert [r[:2] for r in rec2] == [tuple(keys[e]) for e in range(len(h)) if h[e] > 0]


def test_reader_on_hand_packed_file(tmp_path):
    keys = np.array([[1, 2], [3, 4], [5, 6], [7, 8]], dtype=np.uint32)
    body = struct.pack("<ii", 1, 2)
    body += struct.pack("<II", 5, 6)                        # a zero entry
    body += struct.pack("<IIddd", 3, 4, 1.5, -0.25, 2.0)
    body += struct.pack("<IIddd", 9, 9, 7.0, 7.0, 7.0)     # not in the mesh
    p = tmp_path / "hand.gz"
    with gzip.open(p, "wb") as f:
        f.write(body)
    init = [np.full(4, 42.0) for _ in range(3)]
    (h, hu, hv), found, unknown = mesh.read_solution(p, keys, init)
    assert (found, unknown) == (2, 1)
    assert list(h) == [42.0, 1.5, 0.0, 42.0]
    assert list(hu) == [42.0, -0.25, 0.0, 42.0]
    assert list(hv) == [42.0, 2.0, 0.0, 42.0]


def test_round_trip_in_another_element_order(tmp_path):
    keys, U = _case(n=300, seed=9)
    p = tmp_path / "rt"
    mesh.write_solution(p, keys, U)
    perm = np.random.default_rng(2).permutation(len(keys))
    (h, hu, hv), found, unknown = mesh.read_solution(p, keys[perm])
    assert (found, unknown) == (len(keys), 0)
    wet = U[0][perm] > 0
    assert np.array_equal(h[wet], U[0][perm][wet])
    assert np.array_equal(hu[wet], U[1][perm][wet]) and np.array_equal(hv[wet], U[2][perm][wet])
    # zero entries come back as (0, 0, 0), whatever the state held
    assert not np.any(h[~wet]) and not np.any(hu[~wet]) and not np.any(hv[~wet])


def test_bad_files_are_errors(tmp_path):
    keys, U = _case(n=16)
    with pytest.raises(OSError):
        mesh.read_solution(tmp_path / "missing", keys)
    p = tmp_path / "t"
    mesh.write_solution(p, keys, U)
    raw = gzip.open(p, "rb").read()
    with gzip.open(tmp_path / "short", "wb") as f:
        f.write(raw[:-5])
    with pytest.raises(OSError):
        mesh.read_solution(tmp_path / "short", keys)
    with gzip.open(tmp_path / "long", "wb") as f:
        f.write(raw + b"\0")
    with pytest.raises(OSError):
        mesh.read_solution(tmp_path / "long", keys)
    with pytest.raises(OSError):
        mesh.write_solution(p, keys, U, order=np.full(len(keys), len(keys), np.int32))
