"""Column files loaded into device memory (flsgpu_file_load) decode bit for
bit like the same column uploaded array by array, for every encoding; and the
end-to-end driver (encode a data file, decode or query it on the GPU; the
reference's src/bench-compressors.cu path) on synthetic stand-ins for its
datasets, which are absent here."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import alp as AL  # noqa: E402
from flsgpu import colfile as CF  # noqa: E402
from flsgpu import e2e as E2E  # noqa: E402

import oracle_py as O  # noqa: E402
import test_colfile as TC  # noqa: E402

UT = {32: np.uint32, 64: np.uint64}


def expected(col):
    """Decoded values from the oracle (FFOR/BP) or the original data."""
    w = col.width
    nl = 1024 // w
    ocol = O.FFORColumn(col.n_values, col.packed if len(col.packed) else np.zeros(1, UT[w]),
                        col.bit_widths, col.vector_offsets,
                        col.bases if col.bases is not None else np.zeros(col.n_vecs, UT[w]))
    assert nl
    return O.ffor_column_decode(ocol)


def test_load_and_decode_every_encoding(tmp_path):
    rng = np.random.default_rng(11)
    for w in (32, 64):
        for with_bases in (True, False):
            col = TC.ffor_column(rng, w, 9, with_bases)
            p = str(tmp_path / f"f{w}{with_bases}.flsc")
            CF.write(p, col)
            got = CF.load(p).decompress().cpu().numpy()
            assert np.array_equal(got, expected(col)), (w, with_bases)
        dt = np.float32 if w == 32 else np.float64
        x = rng.integers(-10**4, 10**4, 20 * 1024).astype(dt)
        x[::131] += dt(0.37)  # one value in 131 is not an integer: exceptions
        alp = AL.alp_encode(x)
        assert alp is not None and len(alp.exceptions) > 0
        for host in (alp, AL.galp_convert(alp)):
            p = str(tmp_path / f"a{w}.flsc")
            CF.write(p, host)
            got = CF.load(p).decompress().cpu().numpy()
            assert np.array_equal(got.view(UT[w]), x.view(UT[w]))
            lc = CF.load(p)
            assert lc.query_any(float(x[777])).item() == 1
        y = rng.random(7 * 1024).astype(dt)
        rd = AL.alprd_encode(y)
        p = str(tmp_path / f"r{w}.flsc")
        CF.write(p, rd)
        got = CF.load(p).decompress().cpu().numpy()
        assert np.array_equal(got.view(UT[w]), y.view(UT[w]))


def test_load_larger_than_one_staging_chunk(tmp_path):
    """A payload of several 32 MiB staging chunks (double-buffered copy)."""
    rng = np.random.default_rng(12)
    n_vecs = 24 * 1024  # 24 Mi u64 values at width 64: 192 MiB packed
    nl = 16
    bws = np.full(n_vecs, 64, np.uint8)
    offs = np.arange(n_vecs, dtype=np.uint64) * np.uint64(64 * nl)
    packed = rng.integers(0, 2**63, n_vecs * 1024, dtype=np.uint64)
    col = CF.FforHostColumn(n_vecs * 1024, packed, bws, offs, None)
    p = str(tmp_path / "big.flsc")
    CF.write(p, col)
    assert CF.stat(p)["payload_bytes"] > 3 * (32 << 20)
    got = CF.load(p).decompress().cpu().numpy()
    assert np.array_equal(got, expected(col))
    os.remove(p)


def test_upload_every_encoding_with_padding(tmp_path):
    """flsgpu_column_upload (host::*Column::copy_to_device): decodes like the
    array-by-array upload; the reference's over-read padding after the packed
    words is there and zero."""
    import ctypes as C
    from flsgpu import _capi as A
    rng = np.random.default_rng(21)
    for col in TC.columns(rng):
        dev = CF.upload(col)
        got = dev.decompress().cpu().numpy()
        if isinstance(col, CF.FforHostColumn):
            assert np.array_equal(got, expected(col))
            words = len(col.packed)
            pad = np.ones(512, dtype=col.packed.dtype)
            bp = dev.struct.bp if col.bases is not None else dev.struct
            A.check(A.lib().flsgpu_download(pad.ctypes.data_as(C.c_void_p),
                                            C.c_void_p(bp.packed_array + words * col.packed.itemsize),
                                            pad.nbytes, None))
            assert not pad.any()
        else:
            w = 32 if got.dtype == np.float32 else 64
            ref = (AL.alprd_decompress(AL.DeviceAlprdColumn.from_host(col))
                   if isinstance(col, AL.AlprdHostColumn) else
                   AL.decompress(AL.DeviceAlpColumn.from_host(col))).cpu().numpy()
            assert np.array_equal(got.view(UT[w]), ref.view(UT[w]))


@pytest.mark.parametrize("compressor", ["ALP", "GALP", "ALPRD"])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_end_to_end_driver(tmp_path, compressor, dtype, capsys):
    rng = np.random.default_rng(13)
    w = 32 if dtype == "f32" else 64
    if compressor == "ALPRD":
        x = rng.random(10 * 1024 + 77).astype(np.float32 if w == 32 else np.float64)
    else:
        x = TC.decimal_values(rng, w, 10 * 1024 + 77)
    x[300] = E2E.query_value(x.dtype)  # the reference's search value is present
    data = str(tmp_path / f"stand_in_{dtype}.bin")
    x.tofile(data)
    flsc = str(tmp_path / "col.flsc")
    for via in (None, flsc):
        argv = [dtype, "decompression", compressor, data, "12", "--repeats", "2"]
        argv += ["--via-file", via] if via else []
        assert E2E.main(argv) == 0
        line = capsys.readouterr().out.strip().splitlines()[0].split(",")
        assert line[:4] == ["decompression", compressor, f"stand_in_{dtype}", "1"]
        assert int(line[4]) == 12 * 1024 * w // 8 and float(line[6]) > 1.0
    if compressor != "ALPRD":
        assert E2E.main([dtype, "decompression_query", compressor, data, "0", "--repeats", "2"]) == 0
        line = capsys.readouterr().out.strip().split(",")
        assert line[0] == "decompression_query" and line[3] == "1"
