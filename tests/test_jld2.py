"""JLD2 result files (qdfn.jld2): the reference's `save(SaveResultFileName,
"History_V", ..., "History_NormalStress", ...)` (scripts/Functions_RSFDFN3DMain_D.jl:
286-291) and `load(...)` over the HDF5 subset described in the module.

Pins: Jenkins' lookup3 hashlittle against the published values of its test
driver (lookup3.c, driver5: "" -> deadbeef, "" seeded deadbeef -> bd5b7dde,
"Four score and seven years ago" -> 17770551 / seed 1 -> cd628161); the file
layout read back field by field as the HDF5 specification places it
(superblock version 2 at offset 512, end-of-file address, object-header
signatures and checksums, dataspace dims in reversed Julia order, IEEE
datatype fields); round trips of every supported element type and shape;
checksum failures on corrupted metadata.  No JLD2.jl or HDF5 library runs
here, so reading by those libraries is unpinned."""
import struct

import numpy as np
import pytest

from qdfn import jld2


def test_lookup3_published_values():
    assert jld2.lookup3(b"") == 0xDEADBEEF
    assert jld2.lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    assert jld2.lookup3(b"Four score and seven years ago", 0) == 0x17770551
    assert jld2.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def test_lookup3_block_boundaries_differ():
    # lengths around the 12-byte block size hash differently (no tail bytes lost)
    vals = {jld2.lookup3(b"x" * k) for k in range(0, 40)}
    assert len(vals) == 40


def test_result_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    rec, n = 37, 11
    hist = {"History_Time": rng.random(rec), "History_V": rng.random((rec, n)),
            "History_Disp": rng.random((rec, n)), "History_Theta": rng.random((rec, n)),
            "History_NormalStress": rng.random((rec, n))}
    p = tmp_path / "Result.jld2"
    jld2.save_result(p, hist)
    assert jld2.keys(p) == list(jld2.RESULT_KEYS)
    V, T = jld2.load(p, "History_V", "History_Time")
    np.testing.assert_array_equal(V, hist["History_V"])
    assert T.shape == (rec, 1)
    np.testing.assert_array_equal(T[:, 0], hist["History_Time"])
    allv = jld2.load(p)
    for k in jld2.RESULT_KEYS:
        assert allv[k].dtype == np.float64


@pytest.mark.parametrize("value", [
    np.arange(12.0).reshape(3, 4), np.arange(6, dtype=np.float32), np.arange(-5, 5, dtype=np.int64).reshape(5, 2),
    np.arange(7, dtype=np.int32), np.arange(8, dtype=np.uint8).reshape(2, 2, 2), np.zeros((0, 4)),
    np.float64(2.5e-3), np.int64(-17),
])
def test_types_and_shapes_round_trip(tmp_path, value):
    p = tmp_path / "a.jld2"
    jld2.save(p, "x", value, "LoadingFaultCount", np.int64(5))
    got = jld2.load(p, "x")
    assert np.asarray(got).shape == np.asarray(value).shape
    assert np.asarray(got).dtype == np.asarray(value).dtype
    np.testing.assert_array_equal(got, value)
    assert jld2.load(p, "LoadingFaultCount") == 5


def test_layout_follows_the_specification(tmp_path):
    p = tmp_path / "b.jld2"
    A = np.asfortranarray(np.arange(15.0).reshape(3, 5))  # a Julia 3 x 5 matrix
    jld2.save(p, "A", A)
    data = p.read_bytes()
    assert data.startswith(b"HDF5-based Julia Data Format, version 0.1.5")
    sb = data[512:]
    assert sb[:8] == b"\x89HDF\r\n\x1a\n" and sb[8] == 2 and sb[9] == 8 and sb[10] == 8
    base, ext, eof, root = struct.unpack_from("<QQQQ", sb, 12)
    assert base == 0 and ext == jld2.UNDEF and eof == len(data) - 512
    assert struct.unpack_from("<I", sb, 44)[0] == jld2.lookup3(sb[:44])
    assert sb[root:root + 4] == b"OHDR"
    # the root group's link names the dataset; its header has a reversed
    # dataspace (5, 3), an IEEE 64-bit little-endian datatype and a
    # contiguous layout pointing at the matrix's column-major bytes
    msgs = dict(jld2._parse_object_header(sb, root))
    link = [b for t, b in jld2._parse_object_header(sb, root) if t == 0x0006][0]
    assert link[3:4] == b"A"
    addr = struct.unpack_from("<Q", link, 4)[0]
    assert 0x0002 in msgs and 0x000A in msgs
    ds = dict(jld2._parse_object_header(sb, addr))
    space = ds[0x0001]
    assert space[:4] == bytes([2, 2, 0, 1]) and struct.unpack_from("<QQ", space, 4) == (5, 3)
    dtype = ds[0x0003]
    assert dtype[0] == 0x11 and dtype[2] == 63 and struct.unpack_from("<I", dtype, 4)[0] == 8
    assert struct.unpack_from("<HHBBBBI", dtype, 8) == (0, 64, 52, 11, 0, 52, 1023)
    lay = ds[0x0008]
    assert lay[:2] == bytes([3, 1])
    daddr, size = struct.unpack_from("<QQ", lay, 2)
    assert size == 15 * 8
    np.testing.assert_array_equal(np.frombuffer(sb[daddr:daddr + size], dtype="<f8"), A.ravel(order="F"))


def test_corrupted_metadata_is_refused(tmp_path):
    p = tmp_path / "c.jld2"
    jld2.save(p, "x", np.ones(4))
    data = bytearray(p.read_bytes())
    sb = data[512:]
    root = struct.unpack_from("<Q", sb, 36)[0]
    bad = bytearray(data)
    bad[512 + root + 14] ^= 0x01  # inside the root group's messages
    q = tmp_path / "d.jld2"
    q.write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        jld2.load(q, "x")
    bad = bytearray(data)
    bad[512 + root + 8] ^= 0x40  # the root header's chunk size
    q.write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="past the end|checksum"):
        jld2.load(q, "x")
    bad = bytearray(data)
    bad[512 + 20] ^= 0x01  # the superblock's base address
    q.write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        jld2.load(q)


def test_argument_errors(tmp_path):
    p = tmp_path / "e.jld2"
    with pytest.raises(ValueError):
        jld2.save(p, "x")
    with pytest.raises(ValueError):
        jld2.save(p, "x", 1.0, "x", 2.0)
    with pytest.raises(TypeError):
        jld2.save(p, "flag", np.array([True, False]))
    jld2.save(p, "x", 1.0)
    with pytest.raises(KeyError):
        jld2.load(p, "y")
