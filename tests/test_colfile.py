"""Column files (include/flsgpu.h flsgpu_file_*), host side: every encoding
writes and reads back array for array, and a damaged file (bad magic, short
payload, a bit width, offset, count, position or table index that would make
a decoder read outside its arrays) is rejected before any kernel sees it.
Also the data-file reader of the end-to-end driver against the reference's
read_file_as rules (src/engine/data.cuh:184-233)."""
import os

import numpy as np
import pytest

from flsgpu import _capi as A
from flsgpu import alp as AL
from flsgpu import colfile as CF
from flsgpu import e2e as E2E

HDR = np.dtype([("magic", "S8"), ("version", "<u4"), ("encoding", "<i4"), ("n_values", "<u8"),
                ("n_vecs", "<u8"), ("payload", "<u8"), ("n_sections", "<u4"), ("rd", "<u4", 3),
                ("dict", "<u2", 8),
                ("sections", [("id", "<u4"), ("elem", "<u4"), ("count", "<u8"), ("offset", "<u8")],
                 16)])
SECTION = {"packed": 1, "bit_widths": 2, "vector_offsets": 3, "bases": 4, "factor_indices": 5,
           "fraction_indices": 6, "exceptions_offsets": 7, "exceptions": 8, "positions": 9,
           "counts": 10, "offsets_counts": 11}


def ffor_column(rng, width, n_vecs, with_bases=True):
    nl = 1024 // width
    bws = rng.integers(0, width + 1, n_vecs).astype(np.uint8)
    offs = np.zeros(n_vecs, np.uint64)
    if n_vecs:
        offs[1:] = np.cumsum(bws.astype(np.uint64) * np.uint64(nl))[:-1]
    ut = np.uint32 if width == 32 else np.uint64
    packed = np.frombuffer(rng.bytes(int(bws.astype(np.int64).sum()) * nl * width // 8), ut).copy()
    bases = np.frombuffer(rng.bytes(n_vecs * width // 8), ut).copy() if with_bases else None
    return CF.FforHostColumn(n_vecs * 1024, packed, bws, offs, bases)


def decimal_values(rng, width, n):
    if width == 64:
        return rng.integers(-10**6, 10**6, n) / 100.0
    return (rng.integers(-10**4, 10**4, n).astype(np.float32) / np.float32(10.0))


def columns(rng):
    out = []
    for w in (32, 64):
        out.append(ffor_column(rng, w, 7, True))
        out.append(ffor_column(rng, w, 5, False))
        dt = np.float32 if w == 32 else np.float64
        x = rng.integers(-10**4, 10**4, 12 * 1024).astype(dt)
        x[::97] += dt(0.37)  # exceptions; integer-valued data always samples as ALP
        col = AL.alp_encode(x)
        assert col is not None
        out.append(col)
        out.append(AL.galp_convert(col))
        out.append(AL.alprd_encode(rng.random(6 * 1024).astype(dt)))
    return out


def same(a, b):
    assert type(a) is type(b)
    for k, v in vars(a).items():
        if k.startswith("compressed_bytes"):
            continue
        u = getattr(b, k)
        if v is None or u is None:
            assert v is None and u is None, k
        elif isinstance(v, np.ndarray):
            assert v.dtype == u.dtype and np.array_equal(v, u), k
        else:
            assert v == u, k


def test_roundtrip_every_encoding(tmp_path):
    rng = np.random.default_rng(3)
    for i, col in enumerate(columns(rng)):
        p = str(tmp_path / f"c{i}.flsc")
        CF.write(p, col)
        info = CF.stat(p)
        assert info["encoding"] == CF.ENC_NAME[CF.encoding_of(col)]
        assert info["n_values"] == col.n_values and info["file_bytes"] == os.path.getsize(p)
        assert info["payload_bytes"] % 256 == 0
        back = CF.read_host(p)
        if isinstance(col, AL.AlpHostColumn) and col.offsets_counts is not None:
            back.counts = col.counts  # the G-ALP layout stores no counts
        same(col, back)


def test_empty_column(tmp_path):
    p = str(tmp_path / "e.flsc")
    col = CF.FforHostColumn(0, np.zeros(0, np.uint32), np.zeros(0, np.uint8),
                            np.zeros(0, np.uint64), np.zeros(0, np.uint32))
    CF.write(p, col)
    back = CF.read_host(p)
    assert back.n_values == 0 and len(back.packed) == 0


def test_write_rejects_bad_columns(tmp_path):
    rng = np.random.default_rng(4)
    col = ffor_column(rng, 32, 3)
    col.bit_widths = col.bit_widths.copy()
    col.bit_widths[1] = 33
    with pytest.raises(A.FlsgpuError):
        CF.write(str(tmp_path / "x.flsc"), col)


def header(path):
    with open(path, "rb") as f:
        return np.frombuffer(f.read(HDR.itemsize), HDR)[0].copy()


def patch_bytes(path, pos, data):
    with open(path, "r+b") as f:
        f.seek(pos)
        f.write(data)


def section_pos(path, name):
    h = header(path)
    for s in h["sections"][: h["n_sections"]]:
        if s["id"] == SECTION[name]:
            return 4096 + int(s["offset"]), int(s["elem"]), int(s["count"])
    raise KeyError(name)


def fresh(tmp_path, col, name="d.flsc"):
    p = str(tmp_path / name)
    CF.write(p, col)
    return p


def rejected(p):
    with pytest.raises(A.FlsgpuError):
        CF.read_host(p)


def test_damaged_files_are_rejected(tmp_path):
    rng = np.random.default_rng(5)
    ffor = ffor_column(rng, 64, 6)
    ffor.bit_widths[:] = 9
    ffor.vector_offsets = np.arange(6, dtype=np.uint64) * np.uint64(9 * 16)
    ffor.packed = np.frombuffer(rng.bytes(6 * 9 * 16 * 8), np.uint64).copy()
    x = decimal_values(rng, 64, 8 * 1024)
    x[::97] = rng.random(len(x[::97])) * 1e3  # exceptions in every vector
    alp = AL.alp_encode(x)
    assert alp is not None and len(alp.exceptions) > 0
    galp = AL.galp_convert(alp)

    p = fresh(tmp_path, ffor)
    patch_bytes(p, 0, b"NOTACOLF")
    rejected(p)

    p = fresh(tmp_path, ffor)
    with open(p, "r+b") as f:
        f.truncate(os.path.getsize(p) - 300)
    rejected(p)

    p = fresh(tmp_path, ffor)
    pos, _, _ = section_pos(p, "bit_widths")
    patch_bytes(p, pos + 2, bytes([65]))  # width above 64
    rejected(p)

    p = fresh(tmp_path, ffor)
    pos, _, _ = section_pos(p, "vector_offsets")
    patch_bytes(p, pos + 8 * 5, np.uint64(6 * 9 * 16 - 10).tobytes())  # last vector runs past the end
    rejected(p)

    p = fresh(tmp_path, alp)
    pos, _, _ = section_pos(p, "counts")
    patch_bytes(p, pos + 2 * 7, np.uint16(1000).tobytes())  # more exceptions than stored
    rejected(p)

    p = fresh(tmp_path, alp)
    pos, _, n = section_pos(p, "positions")
    patch_bytes(p, pos + 2 * (n - 1), np.uint16(1024).tobytes())
    rejected(p)

    p = fresh(tmp_path, alp)
    pos, _, _ = section_pos(p, "factor_indices")
    patch_bytes(p, pos, bytes([19]))  # past the 19-entry f64 factor table
    rejected(p)

    p = fresh(tmp_path, galp)
    pos, _, _ = section_pos(p, "offsets_counts")
    patch_bytes(p, pos, np.uint16((63 << 10) | 1000).tobytes())
    rejected(p)

    p = fresh(tmp_path, galp)
    with pytest.raises(A.FlsgpuError):  # the wrong encoding for this file
        st = A.ALPColumn()
        import ctypes as C
        blk = C.c_void_p()
        A.check(CF._lib().flsgpu_file_read_host(p.encode(), CF.ENC["alp_f64"], C.byref(st),
                                                C.byref(blk)))
    # the undamaged files still read
    for col in (ffor, alp, galp):
        CF.read_host(fresh(tmp_path, col, "ok.flsc"))


def test_read_file_as_rules(tmp_path):
    p = str(tmp_path / "v.bin")
    vals = np.arange(2500, dtype=np.float64)
    vals.tofile(p)
    got = E2E.read_file_as(p, np.float64, 0)  # whole vectors of the file
    assert np.array_equal(got, vals[:2048])
    got = E2E.read_file_as(p, np.float64, 5 * 1024)  # repeated to fill
    assert len(got) == 5120 and np.array_equal(got[:2500], vals) and np.array_equal(got[2500:5000], vals)
    assert np.array_equal(got[5000:], vals[:120])
    got = E2E.read_file_as(p, np.float64, 1500)  # rounded down to whole vectors
    assert np.array_equal(got, vals[:1024])
    with open(p, "ab") as f:
        f.write(b"\0\0\0")
    with pytest.raises(ValueError):
        E2E.read_file_as(p, np.float64, 0)


def test_query_value_bits():
    assert np.float32(E2E.query_value(np.float32)).view(np.uint32) == 0x4100F9DB
    assert np.float64(E2E.query_value(np.float64)).view(np.uint64) == 0x4100F9DB


def test_unsorted_or_duplicate_positions_are_rejected(tmp_path):
    """The decoders take exception positions as strictly increasing within a
    vector's list (within a lane's list for G-ALP): the writer, the upload
    helper's planner and the loader reject any other list."""
    rng = np.random.default_rng(6)
    x = decimal_values(rng, 64, 4 * 1024)
    x[::61] = rng.random(len(x[::61])) * 1e3  # ~17 exceptions per vector
    alp = AL.alp_encode(x)
    assert alp is not None and int(alp.counts[1]) >= 3
    galp = AL.galp_convert(alp)
    rd = AL.alprd_encode(rng.random(4 * 1024))
    e1 = int(alp.exceptions_offsets[1])

    # loader: duplicate the second position of vector 1 into the third
    p = fresh(tmp_path, alp)
    pos, _, _ = section_pos(p, "positions")
    patch_bytes(p, pos + 2 * (e1 + 2), alp.positions[e1 + 1].tobytes())
    rejected(p)
    # loader: swap two positions of vector 1 (out of order, still unique)
    p = fresh(tmp_path, alp)
    pos, _, _ = section_pos(p, "positions")
    patch_bytes(p, pos + 2 * (e1 + 1), alp.positions[e1 + 2].tobytes() + alp.positions[e1 + 1].tobytes())
    rejected(p)
    # loader, G-ALP: a lane's first entry moved to a position of another lane
    lane = int(np.nonzero(galp.offsets_counts[16:32] >> 10)[0][0])
    oc = int(galp.offsets_counts[16 + lane])
    first = int(galp.exceptions_offsets[1]) + (oc & 0x3FF)
    p = fresh(tmp_path, galp)
    pos, _, _ = section_pos(p, "positions")
    patch_bytes(p, pos + 2 * first, np.uint16((int(galp.positions[first]) + 1) % 1024).tobytes())
    rejected(p)
    # writer (the same planner runs for flsgpu_column_upload)
    bad = AL.AlpHostColumn(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                              for k, v in vars(alp).items()})
    bad.positions[e1 + 2] = bad.positions[e1 + 1]
    with pytest.raises(A.FlsgpuError):
        CF.write(str(tmp_path / "dup.flsc"), bad)
    if len(rd.positions) >= 2 and int(rd.counts[0]) >= 2:
        badrd = AL.AlprdHostColumn(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                                      for k, v in vars(rd).items()})
        badrd.positions[1] = badrd.positions[0]
        with pytest.raises(A.FlsgpuError):
            CF.write(str(tmp_path / "duprd.flsc"), badrd)
    # the undamaged columns still write and read
    for col in (alp, galp, rd):
        CF.read_host(fresh(tmp_path, col, "ok.flsc"))
