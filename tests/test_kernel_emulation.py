"""The device encoder's kernels (csrc/encode.hip), the ALP decoders
(csrc/alp_kernels.hpp) and the batched decoder (csrc/batch.hip: plan, task
list, persistent kernel over every encoding) run on the CPU through a host emulation of the HIP
features they use (tests/emu/hip/hip_runtime.h: one OS thread per GPU
thread, block and wave barriers, group shuffles), built with
AddressSanitizer. Every output buffer has the exact size the library's
callers give it, so an out-of-bounds access aborts the driver. The encoder's
arrays must equal the host encoder's byte for byte; the decoders must
reproduce the values, for every variant shape and both the uniform-width and
mixed-width unpack paths. The range-scan kernels' per-lane arithmetic (the
fused filter + sum's popcount / bit-plane statistics, its range and equality
predicates, the ALP query's interval of matching offsets) is checked against
brute force at every bit width."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from flsgpu import _capi as A
from flsgpu import alp as AL

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UT = {32: np.uint32, 64: np.uint64}
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")


def _compiler():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    cxx = _compiler()
    if cxx is None:
        pytest.skip("no clang++ for the host emulation build")
    out = tmp_path_factory.mktemp("emu")
    builds = {"encoder_driver": ("encoder_driver", []), "batch_driver": ("batch_driver", []),
              "scan_math_driver": ("scan_math_driver", [])}
    for t in ("float", "double"):
        for g in (0, 1):
            builds[f"alp_decoder_{t}_{g}"] = ("alp_decoder_driver", [f"-DEMU_T={t}", f"-DEMU_GALP={g}"])
    procs = {}
    for key, (src, defs) in builds.items():  # the builds run in parallel
        cmd = [cxx, "-std=c++20", "-O1", "-g0", "-fsanitize=address", "-fno-omit-frame-pointer",
               "-ffp-contract=off", *defs, "-x", "c++", "-I", EMU, "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(ROOT, "msc-thesis-project_amd", "csrc"),
               os.path.join(EMU, src + ".cpp"), "-o", str(out / key), "-lpthread"]
        procs[key] = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    for key, p in procs.items():
        _, err = p.communicate()
        assert p.returncode == 0, err[-3000:]
    return {key: str(out / key) for key in builds}


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=ENV, timeout=600)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return r


def encode_emulated(drivers, tmp_path, x, e=None, f=None, galp=False):
    """Runs the three encoder kernels; returns (totals, arrays)."""
    w = x.dtype.itemsize * 8
    nv = len(x) // 1024
    ffor = x.dtype.kind == "u"
    src, dst = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(src, "wb") as fh:
        fh.write(struct.pack("<4I", w, nv, int(galp), int(ffor)))
        fh.write(x.tobytes())
        if not ffor:
            fh.write(e.tobytes())
            fh.write(f.tobytes())
    r = _run(drivers["encoder_driver"], src, dst)
    assert r.returncode == 0, r.stderr[-3000:]
    d = open(dst, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(d, dt, n, off)
        off += a.nbytes
        return a

    totals = take(np.uint64, 3)
    words, ne = int(totals[0]), int(totals[1])
    arrs = {"packed": take(UT[w], words), "bit_widths": take(np.uint8, nv),
            "vector_offsets": take(np.uint64, nv), "bases": take(UT[w], nv)}
    if not ffor:
        arrs.update({"factor_indices": take(np.uint8, nv), "fraction_indices": take(np.uint8, nv),
                     "exceptions_offsets": take(np.uint64, nv), "exceptions": take(x.dtype, ne),
                     "positions": take(np.uint16, ne), "counts": take(np.uint16, nv)})
        if galp:
            arrs["offsets_counts"] = take(np.uint16, nv * (1024 // w))
    return totals, arrs


def same_as_host(arrs, host):
    for k, v in arrs.items():
        if k == "counts" and host.counts is None:
            continue
        h = np.asarray(getattr(host, k))
        if k == "exceptions":
            assert np.array_equal(v.view(UT[v.itemsize * 8]), h.view(UT[h.itemsize * 8])), k
        else:
            assert h.shape == v.shape and np.array_equal(v, h.astype(v.dtype)), k


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["search", "random"])
@pytest.mark.parametrize("galp", [False, True])
def test_encoder_alp_matches_host(drivers, tmp_path, dtype, mode, galp):
    import test_encode_gpu as TE  # the GPU test's columns and parameters
    rng = np.random.default_rng(3)
    x = TE.columns(dtype, 11)
    if galp:
        v = x.reshape(-1, 1024)
        v[0] = np.round(v[6], 2)
        v[1] = np.round(v[7], 1)
        if dtype == np.float64:
            v[3] = np.round(v[8], 2)
    e, f = TE.params(x, dtype, rng, mode)
    host = AL.alp_encode_with_params(x, e, f)
    if galp:
        try:
            host = AL.galp_convert(host)
        except A.FlsgpuError:
            # random f64 parameters leave whole lanes as exceptions, which the
            # 16-bit G-ALP header cannot hold: the device encoder reports it
            assert mode == "random"
            totals, _ = encode_emulated(drivers, tmp_path, x, e, f, True)
            assert int(totals[2]) == 1
            return
    totals, arrs = encode_emulated(drivers, tmp_path, x, e, f, galp)
    assert int(totals[2]) == 0
    same_as_host(arrs, host)


def test_encoder_lane_overflow(drivers, tmp_path):
    x = np.round(np.random.default_rng(1).uniform(-100, 100, 2048), 2)
    x[1024 + 3::16] = np.pi  # every slot of lane 3 of vector 1 is an exception
    e, f = np.full(2, 2, np.uint8), np.zeros(2, np.uint8)
    totals, arrs = encode_emulated(drivers, tmp_path, x, e, f, False)
    same_as_host(arrs, AL.alp_encode_with_params(x, e, f))
    totals, _ = encode_emulated(drivers, tmp_path, x, e, f, True)
    assert int(totals[2]) == 1  # reported: the 16-bit G-ALP header cannot hold 64


@pytest.mark.parametrize("width", [32, 64])
def test_encoder_ffor_matches_host(drivers, tmp_path, width):
    rng = np.random.default_rng(width)
    ut = UT[width]
    nv = 40
    bws = rng.integers(0, width + 1, nv)
    bws[:3] = [0, width, 1]
    vals = np.empty((nv, 1024), ut)
    for v in range(nv):
        b = int(bws[v])
        span = (1 << b) - 1 if b < width else (1 << width) - 1
        off = np.frombuffer(rng.bytes(1024 * width // 8), dtype=ut) & ut(span)
        room = (1 << width) - span  # bases that keep base + offset below 2^width
        base = 0 if b == width else int.from_bytes(rng.bytes(8), "little") % room
        vals[v] = off + ut(base)
    vals = vals.reshape(-1)
    mn = vals.reshape(nv, 1024).min(axis=1)
    mx = vals.reshape(nv, 1024).max(axis=1)
    bw = np.array([(int(a) - int(b)).bit_length() for a, b in zip(mx, mn)], np.uint8)
    packed, offs = AL.ffor_pack(vals, bw, mn)
    totals, arrs = encode_emulated(drivers, tmp_path, vals)
    assert int(totals[0]) == len(packed)
    assert np.array_equal(arrs["packed"], packed) and np.array_equal(arrs["bit_widths"], bw)
    assert np.array_equal(arrs["vector_offsets"], offs) and np.array_equal(arrs["bases"], mn)


def dense_column(dtype, galp):
    """Mixed widths and exception densities: a lane full of exceptions
    (standard layout; 63 in G-ALP), a vector of exceptions only (standard),
    specials, sparse outliers."""
    rng = np.random.default_rng(17)
    w = np.dtype(dtype).itemsize * 8
    nl = 1024 // w
    nv = 7
    x = np.round(rng.uniform(-10 ** (1 + np.arange(nv) % 4)[:, None], 10 ** (1 + np.arange(nv) % 4)[:, None],
                             (nv, 1024)), 2).astype(dtype)
    x[0, 5::nl] = dtype(np.pi) if galp is False else x[0, 5::nl]
    if galp:
        x[0, 5 + nl::nl] = dtype(np.pi)  # 63 of the lane's slots
    else:
        x[1] = dtype(np.pi)              # every value an exception
    x[2, :4] = [np.nan, -0.0, np.inf, -np.inf]
    x[3, rng.choice(1024, 40, replace=False)] = rng.uniform(-1e6, 1e6, 40).astype(dtype)
    x = x.reshape(-1)
    col = AL.alp_encode(x)
    assert col is not None
    return x, (AL.galp_convert(col) if galp else col)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("galp", [False, True])
def test_alp_decoders_emulated(drivers, tmp_path, dtype, galp):
    x, col = dense_column(dtype, galp)
    w = x.dtype.itemsize * 8
    src = str(tmp_path / "col.bin")
    with open(src, "wb") as fh:
        fh.write(struct.pack("<6Q", w, col.n_vecs, len(col.packed), len(col.exceptions), int(galp), 0))
        for a, dt in ((col.packed, UT[w]), (col.bit_widths, np.uint8), (col.vector_offsets, np.uint64),
                      (col.bases, UT[w]), (col.factor_indices, np.uint8), (col.fraction_indices, np.uint8),
                      (col.exceptions_offsets, np.uint64), (col.exceptions, x.dtype),
                      (col.positions, np.uint16),
                      (col.offsets_counts if galp else col.counts, np.uint16)):
            fh.write(np.ascontiguousarray(a).astype(dt).tobytes())
        fh.write(x.tobytes())
    r = _run(drivers[f"alp_decoder_{'float' if w == 32 else 'double'}_{int(galp)}"], src)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 16 and all(l.endswith("differences=0") for l in lines), r.stdout


BATCH_ENC = {"bp_u32": 0, "ffor_u32": 1, "bp_u64": 2, "ffor_u64": 3, "alp_f32": 4, "alp_f64": 5,
             "alp_ext_f32": 6, "alp_ext_f64": 7, "alprd_f32": 8, "alprd_f64": 9}


def write_batch(path, specs):
    """The batch driver's input: each column's descriptor arrays as the
    device columns hold them (an empty array as one zero element)."""
    from flsgpu import _capi as A

    def arr(fh, a):
        a = np.ascontiguousarray(a)
        if a.size == 0:
            a = np.zeros(1, dtype=a.dtype)
        fh.write(struct.pack("<Q", a.nbytes))
        fh.write(a.tobytes())

    with open(path, "wb") as fh:
        fh.write(struct.pack("<I", len(specs)))
        for s in specs:
            enc = BATCH_ENC[s.kind]
            patcher = A.PATCHERS["prefetch-all" if enc >= 6 else "stateful"] if 4 <= enc <= 7 else 0
            fh.write(struct.pack("<iQi", enc, len(s.values) // 1024, patcher))
            c = s.host
            if enc <= 3:
                for a in c[:3] + ((c[3],) if enc in (1, 3) else ()):
                    arr(fh, a)
            elif enc <= 7:
                for a in (c.packed, c.bit_widths, c.vector_offsets, c.bases, c.factor_indices,
                          c.fraction_indices, c.exceptions_offsets, c.exceptions, c.positions,
                          c.offsets_counts if enc >= 6 else c.counts):
                    arr(fh, a)
            else:
                d = [int(v) for v in c.dictionary] + [0] * (8 - len(c.dictionary))
                fh.write(struct.pack("<III8H", c.right_bw, c.left_bw, len(c.dictionary), *d))
                for a in (c.right_packed, c.left_packed, c.exceptions_offsets, c.exceptions,
                          c.positions, c.counts):
                    arr(fh, a)
            arr(fh, s.values)


@pytest.mark.parametrize("ballot", [0, 1])
def test_batch_emulated(drivers, tmp_path, ballot):
    """The GPU batch test's columns (every encoding, mixed widths, small and
    ragged) plus a mixed batch: two runs of one plan, every output exact."""
    import test_batch_gpu as TB
    from flsgpu import workloads as WL
    specs = TB.every_encoding(np.random.default_rng(1)) + WL.mixed_batch(seed=2, n_columns=9,
                                                                         min_log2=10, max_log2=15)
    src = str(tmp_path / "batch.bin")
    write_batch(src, specs)
    r = subprocess.run([drivers["batch_driver"], src], capture_output=True, text=True, timeout=600,
                       env=dict(ENV, EMU_BALLOT=str(ballot)))
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    runs = [l for l in r.stdout.splitlines() if l.startswith("run ")]
    assert len(runs) == 2 and all(f"0 of {len(specs)} items differ" in l for l in runs), r.stdout


def test_scan_arithmetic_brute_force(drivers):
    """The fused filter + sum's per-lane statistics (every bit width of u32 and
    u64, every chunk shape, the offset, all-rows, range and equality
    predicates, bases that do and do not wrap) against a per-value
    restatement over the same packed words, and the ALP query's interval of
    matching offsets against the decode of every offset (csrc/fls_kernels.hpp
    span_filter_sum, csrc/alp_kernels.hpp alp_target_interval)."""
    r = _run(drivers["scan_math_driver"])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "mismatches: 0 of" in r.stdout, r.stdout[-2000:]
