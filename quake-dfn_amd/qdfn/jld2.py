"""JLD2 files of numeric arrays: `save` / `load` with the FileIO call shape the
reference uses for its result file,

    save(SaveResultFileName, "History_V", History_V, "History_Disp", History_Disp, ...)
        (scripts/Functions_RSFDFN3DMain_D.jl:286-291)
    load(LoadFileName, "History_V", "History_Disp", ...)

JLD2 (JLD2.jl) files are HDF5 files behind a 512-byte header naming the
format; no HDF5 library is needed to write the subset used here, which follows
the HDF5 file format specification (version 3.0) directly:
  - the JLD2 header ("HDF5-based Julia Data Format, version 0.1.5", zero
    padded) at offset 0, the superblock (version 2) at offset 512, every other
    address relative to it (base address 0, the convention of HDF5 files with
    a user block, which JLD2 writes);
  - the root group as a version-2 object header with a Link Info, a Group Info
    and one hard Link message per dataset (compact link storage);
  - each dataset as a version-2 object header with Dataspace (version 2),
    Datatype (version 1: IEEE little-endian floating point or little-endian
    two's-complement integers), Fill Value (version 3) and Data Layout
    (version 3, contiguous) messages, and its raw data stored contiguously.
Metadata blocks carry Jenkins' lookup3 `hashlittle` checksums (seed 0), as
the specification prescribes for version-2 structures.  A Julia array of
size (m, n) is stored as the HDF5 dataspace (n, m) with its column-major
bytes, which is how JLD2 maps Julia arrays; numpy arrays given here are taken
with the Julia (column-major) meaning of their shape.

`load` reads files of this subset (it checks every checksum); files written
by JLD2.jl with Julia-specific types (the H-matrix block vectors of
Input_Discretized.jld2, LowRankApprox factorizations) are outside it.
"""
import struct

import numpy as np

JLD2_HEADER = b"HDF5-based Julia Data Format, version 0.1.5\x00 (Julia 1.10.0 64-bit LE)\x00"
HEADER_LENGTH = 512
SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF

_MSG_DATASPACE, _MSG_LINK_INFO, _MSG_DATATYPE, _MSG_FILL, _MSG_LINK, _MSG_LAYOUT, _MSG_GROUP_INFO = \
    0x0001, 0x0002, 0x0003, 0x0005, 0x0006, 0x0008, 0x000A


def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 hashlittle (public domain, 2006) of a byte string:
    the checksum of HDF5's version-2 metadata structures."""
    data = bytes(data)
    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & 0xFFFFFFFF
    if n == 0:
        return c
    pad = data + b"\x00" * ((-n) % 12)
    words = struct.unpack("<%dI" % (len(pad) // 4), pad)
    nblk = len(words) // 3
    M = 0xFFFFFFFF
    for blk in range(nblk - 1):
        a = (a + words[3 * blk]) & M
        b = (b + words[3 * blk + 1]) & M
        c = (c + words[3 * blk + 2]) & M
        # mix
        a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M
    # last block (1..12 bytes, zero padded), then final
    a = (a + words[-3]) & M
    b = (b + words[-2]) & M
    c = (c + words[-1]) & M
    c ^= b; c = (c - _rot(b, 14)) & M
    a ^= c; a = (a - _rot(c, 11)) & M
    b ^= a; b = (b - _rot(a, 25)) & M
    c ^= b; c = (c - _rot(b, 16)) & M
    a ^= c; a = (a - _rot(c, 4)) & M
    b ^= a; b = (b - _rot(a, 14)) & M
    c ^= b; c = (c - _rot(b, 24)) & M
    return c


# ---- datatype messages --------------------------------------------------

def _datatype(dt):
    """Datatype message body (version 1) of a numpy scalar type."""
    dt = np.dtype(dt)
    if dt.kind == "f" and dt.itemsize in (4, 8):
        if dt.itemsize == 8:
            props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        else:
            props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
        # class 1 (floating point), version 1; little-endian, implied-MSB
        # mantissa normalization (bits 4-5 = 2), sign bit location
        bits = bytes([0x20, dt.itemsize * 8 - 1, 0])
        return bytes([0x11]) + bits + struct.pack("<I", dt.itemsize) + props
    if dt.kind in "iu" and dt.itemsize in (1, 2, 4, 8):
        # class 0 (fixed point), version 1; little-endian, bit 3 = signed
        bits = bytes([0x08 if dt.kind == "i" else 0x00, 0, 0])
        return bytes([0x10]) + bits + struct.pack("<I", dt.itemsize) + struct.pack("<HH", 0, dt.itemsize * 8)
    raise TypeError(f"jld2: unsupported element type {dt}")


def _parse_datatype(body):
    cls_ver = body[0]
    cls = cls_ver & 0x0F
    size = struct.unpack_from("<I", body, 4)[0]
    if cls == 1:
        if body[1] & 0x01:
            raise ValueError("jld2: big-endian floating point not supported")
        return np.dtype("<f%d" % size)
    if cls == 0:
        if body[1] & 0x01:
            raise ValueError("jld2: big-endian integers not supported")
        return np.dtype(("<i%d" if body[1] & 0x08 else "<u%d") % size)
    raise ValueError(f"jld2: unsupported datatype class {cls}")


# ---- object headers -----------------------------------------------------

def _message(mtype, body, flags=0):
    return struct.pack("<BHB", mtype, len(body), flags) + body


def _object_header(messages):
    """Version-2 object header: 'OHDR', version 2, flags 2 (4-byte chunk size),
    chunk #0 of the messages, checksum."""
    chunk = b"".join(messages)
    head = b"OHDR" + bytes([2, 0x02]) + struct.pack("<I", len(chunk)) + chunk
    return head + struct.pack("<I", lookup3(head))


def _parse_object_header(buf, addr):
    if buf[addr:addr + 4] != b"OHDR":
        raise ValueError(f"jld2: no object header at relative address {addr}")
    version, flags = buf[addr + 4], buf[addr + 5]
    if version != 2:
        raise ValueError("jld2: object header version %d not supported" % version)
    p = addr + 6
    if flags & 0x20:
        p += 16  # access, modification, change, birth times
    if flags & 0x10:
        p += 4   # attribute phase change values
    width = 1 << (flags & 0x03)
    chunk = int.from_bytes(buf[p:p + width], "little")
    p += width
    end = p + chunk
    if end + 4 > len(buf):
        raise ValueError("jld2: object header extends past the end of the file")
    if lookup3(buf[addr:end]) != struct.unpack_from("<I", buf, end)[0]:
        raise ValueError("jld2: object header checksum mismatch")
    msgs = []
    while p + 4 <= end:
        mtype, size, mflags = struct.unpack_from("<BHB", buf, p)
        p += 4
        if flags & 0x04:
            p += 2  # creation order
        msgs.append((mtype, bytes(buf[p:p + size])))
        p += size
    return msgs


# ---- save / load ------------------------------------------------------

def _julia_array(v):
    """(numpy element type, Julia dims, column-major bytes) of a value."""
    a = np.asarray(v)
    if a.dtype == np.bool_:
        raise TypeError("jld2: Bool arrays are not in the supported subset")
    if a.dtype.kind not in "fiu":
        raise TypeError(f"jld2: unsupported value of type {a.dtype}")
    dt = a.dtype.newbyteorder("<")
    return dt, a.shape, np.asfortranarray(a, dtype=dt).tobytes(order="F")


def save(path, *args):
    """save(path, "name1", value1, "name2", value2, ...): numeric arrays and
    scalars (Float64/Float32/Int64/... in Julia terms) under the given names."""
    if len(args) % 2:
        raise ValueError("jld2.save: name / value pairs expected")
    items = [(str(args[i]), args[i + 1]) for i in range(0, len(args), 2)]
    names = [k for k, _ in items]
    if len(set(names)) != len(names):
        raise ValueError("jld2.save: duplicate names")
    for k in names:
        if not (0 < len(k.encode()) < 256):
            raise ValueError("jld2.save: names of 1 to 255 bytes")
    sb_len = 48
    pos = sb_len  # relative address: the superblock occupies [0, 48)
    headers, blobs = [], []
    # data first, then the dataset headers, then the root group
    data_addr = []
    for _, v in items:
        dt, dims, raw = _julia_array(v)
        data_addr.append((pos if raw else UNDEF, dt, dims, raw))
        blobs.append(raw)
        pos += len(raw)
    ds_addr = []
    for addr, dt, dims, raw in data_addr:
        if len(dims) == 0:
            space = bytes([2, 0, 0, 0])  # version 2, rank 0, flags 0, scalar
        else:
            space = bytes([2, len(dims), 0, 1]) + b"".join(struct.pack("<Q", d) for d in reversed(dims))
        fill = bytes([3, 0x05])  # version 3: allocation early, fill never written, no value
        layout = bytes([3, 1]) + struct.pack("<QQ", addr, len(raw))  # version 3, contiguous
        oh = _object_header([_message(_MSG_DATASPACE, space), _message(_MSG_DATATYPE, _datatype(dt), flags=1),
                             _message(_MSG_FILL, fill), _message(_MSG_LAYOUT, layout)])
        ds_addr.append(pos)
        headers.append(oh)
        pos += len(oh)
    link_info = bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)
    group_info = bytes([0, 0])
    links = [_message(_MSG_LINK, bytes([1, 0, len(k.encode())]) + k.encode() + struct.pack("<Q", a))
             for (k, _), a in zip(items, ds_addr)]
    root = _object_header([_message(_MSG_LINK_INFO, link_info), _message(_MSG_GROUP_INFO, group_info)] + links)
    root_addr = pos
    pos += len(root)
    sb = SIGNATURE + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", 0, UNDEF, pos, root_addr)
    sb += struct.pack("<I", lookup3(sb))
    assert len(sb) == sb_len
    with open(path, "wb") as f:
        f.write(JLD2_HEADER + b"\x00" * (HEADER_LENGTH - len(JLD2_HEADER)))
        f.write(sb)
        for raw in blobs:
            f.write(raw)
        for oh in headers:
            f.write(oh)
        f.write(root)


def _read(path):
    with open(path, "rb") as f:
        data = f.read()
    if not data.startswith(b"HDF5-based Julia Data Format, version "):
        raise ValueError("jld2: not a JLD2 file")
    buf = memoryview(data)[HEADER_LENGTH:]
    if bytes(buf[:8]) != SIGNATURE:
        raise ValueError("jld2: no HDF5 superblock at offset 512")
    version, so, sl, _ = buf[8], buf[9], buf[10], buf[11]
    if version not in (2, 3) or so != 8 or sl != 8:
        raise ValueError("jld2: superblock version 2/3 with 8-byte offsets and lengths expected")
    base, ext, eof, root = struct.unpack_from("<QQQQ", buf, 12)
    if lookup3(buf[:44]) != struct.unpack_from("<I", buf, 44)[0]:
        raise ValueError("jld2: superblock checksum mismatch")
    if eof > len(buf):
        raise ValueError("jld2: file truncated")
    out = {}
    for mtype, body in _parse_object_header(buf, root):
        if mtype != _MSG_LINK:
            continue
        lflags = body[1]
        p = 2
        if lflags & 0x08:
            if body[p] != 0:
                raise ValueError("jld2: only hard links are supported")
            p += 1
        if lflags & 0x04:
            p += 8
        if lflags & 0x10:
            p += 1
        nw = 1 << (lflags & 0x03)
        nlen = int.from_bytes(body[p:p + nw], "little")
        p += nw
        name = body[p:p + nlen].decode()
        p += nlen
        out[name] = struct.unpack_from("<Q", body, p)[0]
    return buf, out


def _dataset(buf, addr):
    dims, dt, layout = None, None, None
    for mtype, body in _parse_object_header(buf, addr):
        if mtype == _MSG_DATASPACE:
            if body[0] != 2:
                raise ValueError("jld2: dataspace version 2 expected")
            rank = body[1]
            if body[3] == 2:
                raise ValueError("jld2: null dataspace")
            dims = [struct.unpack_from("<Q", body, 4 + 8 * k)[0] for k in range(rank)]
        elif mtype == _MSG_DATATYPE:
            dt = _parse_datatype(body)
        elif mtype == _MSG_LAYOUT:
            if body[0] != 3 or body[1] != 1:
                raise ValueError("jld2: contiguous layout (version 3) expected")
            layout = struct.unpack_from("<QQ", body, 2)
    if dims is None or dt is None or layout is None:
        raise ValueError("jld2: incomplete dataset header")
    jdims = tuple(reversed(dims))
    count = int(np.prod(jdims)) if jdims else 1
    daddr, size = layout
    if size != count * dt.itemsize:
        raise ValueError("jld2: data size does not match the dataspace")
    raw = bytes(buf[daddr:daddr + size]) if size else b""
    a = np.frombuffer(raw, dtype=dt, count=count)
    if not jdims:
        return a[0]
    return a.reshape(jdims, order="F").astype(dt.newbyteorder("="))


def keys(path):
    """Names of the datasets in the file's root group, in file order."""
    return list(_read(path)[1].keys())


def load(path, *names):
    """load(path): dict of every dataset; load(path, "name"): its value;
    load(path, "a", "b", ...): a tuple of values (FileIO's load)."""
    buf, links = _read(path)
    want = names or tuple(links.keys())
    vals = []
    for k in want:
        if k not in links:
            raise KeyError(f"jld2: no dataset {k!r}")
        vals.append(_dataset(buf, links[k]))
    if not names:
        return dict(zip(want, vals))
    return vals[0] if len(vals) == 1 else tuple(vals)


RESULT_KEYS = ("History_V", "History_Disp", "History_Time", "History_Theta", "History_NormalStress")


def save_result(path, history):
    """The reference's periodic result file (Functions_RSFDFN3DMain_D.jl:286-291)
    from a run's histories (Simulation.history(): History_Time as a records x 1
    matrix, the others records x n)."""
    h = dict(history)
    h["History_Time"] = np.asarray(h["History_Time"], dtype=np.float64).reshape(-1, 1)
    args = []
    for k in RESULT_KEYS:
        args += [k, np.asarray(h[k], dtype=np.float64)]
    save(path, *args)
