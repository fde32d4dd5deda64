"""GPU parity of the batched persistent decoder: every encoding in one
launch, bit for bit equal to the CPU oracle's decode of the same encoded
buffers and to the original values; a plan re-runs correctly, plainly and
from hipGraph replays. The full-size case is the
benchmark's own 96-column mixed batch (2^12..2^24 values per column)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle_py as O  # noqa: E402
from flsgpu import alp as AL  # noqa: E402
from flsgpu import batch as B  # noqa: E402
from flsgpu import workloads as WL  # noqa: E402


def every_encoding(rng):
    specs = []
    for w in (32, 64):
        ut = np.uint32 if w == 32 else np.uint64
        for with_base in (False, True):
            n = 1024 * int(rng.integers(1, 40))
            vals = np.frombuffer(rng.bytes(n * w // 8), dtype=ut).copy() >> ut(w - 17)
            bws = rng.integers(0, 18, n // 1024).astype(np.uint8)
            bases = np.frombuffer(rng.bytes(n // 1024 * w // 8), dtype=ut).copy() if with_base else None
            vals = (vals & ((ut(1) << bws.astype(ut).repeat(1024)) - ut(1)))
            if bases is not None:
                vals = (vals.reshape(-1, 1024) + bases[:, None]).reshape(-1).astype(ut)
            packed, offs = AL.ffor_pack(vals, bws, bases)
            kind = ("ffor" if with_base else "bp") + f"_u{w}"
            specs.append(WL.BatchColumnSpec(kind, (packed, bws, offs, bases), vals))
        ft = np.float32 if w == 32 else np.float64
        for ext in (False, True):
            lim = 2000 if w == 32 else 8000  # float: stay below the ALP-RD estimate threshold
            x = (rng.integers(-lim, lim, 1024 * int(rng.integers(1, 30))) / 8).astype(ft)
            x[rng.choice(len(x), len(x) // 50, replace=False)] = rng.uniform(-1, 1, len(x) // 50).astype(ft)
            col = AL.alp_encode(x)
            assert col is not None, (w, ext)
            specs.append(WL.BatchColumnSpec(("alp_ext" if ext else "alp") + f"_f{w}",
                                            AL.galp_convert(col) if ext else col, x))
        x = rng.uniform(0, 1, 1024 * int(rng.integers(1, 30))).astype(ft)
        specs.append(WL.BatchColumnSpec(f"alprd_f{w}", AL.alprd_encode(x), x))
    return specs


def _nonempty(a, dtype):
    return a if len(a) else np.zeros(1, dtype)


def oracle_decode(spec):
    """The CPU oracle's decode of the spec's encoded buffers (the same
    buffers the batch decodes on the device)."""
    k = spec.kind
    if k.startswith("ffor") or k.startswith("bp"):
        p, b, o, s = spec.host
        return O.ffor_column_decode(O.FFORColumn(len(spec.values), _nonempty(p, p.dtype), b, o,
                                                 None if k.startswith("bp") else s))
    c = spec.host
    if k.startswith("alprd"):
        return O.alprd_column_decode(O.ALPRDColumn(
            c.n_values, c.right_bw, c.left_bw, c.dictionary, c.right_packed, c.left_packed,
            c.exceptions_offsets, c.exceptions, c.positions, c.counts))
    ext = c.offsets_counts is not None
    ocol = O.ALPColumn(c.n_values,
                       O.FFORColumn(c.n_values, _nonempty(c.packed, c.packed.dtype), c.bit_widths,
                                    c.vector_offsets, c.bases),
                       c.factor_indices, c.fraction_indices, c.exceptions_offsets,
                       _nonempty(c.exceptions, c.dtype), _nonempty(c.positions, np.uint16),
                       None if ext else c.counts, c.offsets_counts if ext else None)
    return O.alp_column_decode(ocol)


def bits(a):
    if isinstance(a, torch.Tensor):
        a = a.cpu().numpy()
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def check_all(specs, expected, outs, what):
    for i, (s, e, o) in enumerate(zip(specs, expected, outs)):
        got = bits(o)
        assert np.array_equal(got, bits(e)), (what, i, s.kind, len(s.values))
        assert np.array_equal(got, bits(s.values)), (what, i, s.kind, len(s.values))


def test_batch_every_encoding_matches_oracle():
    rng = np.random.default_rng(1)
    specs = every_encoding(rng) + WL.mixed_batch(seed=2, n_columns=9, min_log2=10, max_log2=15)
    expected = [oracle_decode(s) for s in specs]
    cols = [WL.to_device(s) for s in specs]
    outs = [B.output_for(c) for c in cols]
    plan = B.BatchPlan(cols, outs)
    assert plan.n_tasks > len(cols) // 2
    plan.run()
    torch.cuda.synchronize()
    check_all(specs, expected, outs, "run 1")
    # a plan is reusable: a second run decodes everything again
    for o in outs:
        o.zero_()
    plan.run()
    torch.cuda.synchronize()
    check_all(specs, expected, outs, "run 2")
    # and the plain per-column launches agree with both
    outs_p = [B.output_for(c) for c in cols]
    B.decode_each(cols, outs_p)
    torch.cuda.synchronize()
    check_all(specs, expected, outs_p, "plain")
    plan.close()


def test_batch_graph_replays():
    specs = WL.mixed_batch(seed=5, n_columns=12, min_log2=10, max_log2=16)
    expected = [oracle_decode(s) for s in specs]
    cols = [WL.to_device(s) for s in specs]
    outs = [B.output_for(c) for c in cols]
    plan = B.BatchPlan(cols, outs)
    for rep in range(4):
        for o in outs:
            o.fill_(0)                     # eager work between replays
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        check_all(specs, expected, outs, f"replay {rep}")
    plan.close()


@pytest.mark.timeout(900)
def test_c4_full_batch_matches_oracle():
    """The benchmark's C4 workload: 96 columns, log-uniform 2^12..2^24 values,
    a third each FFOR u32 (width 1..24), ALP double and ALP-RD double; one
    persistent launch, then two hipGraph replays, every column against the
    oracle's decode."""
    specs = WL.mixed_batch(seed=4, n_columns=96)
    assert len(specs) == 96
    n_values = sum(len(s.values) for s in specs)
    assert n_values > 10 ** 8
    kinds = {s.kind for s in specs}
    assert {"ffor_u32", "alprd_f64"} <= kinds and kinds & {"alp_f64", "alp_ext_f64"}
    expected = [oracle_decode(s) for s in specs]
    cols = [WL.to_device(s) for s in specs]
    outs = [B.output_for(c) for c in cols]
    plan = B.BatchPlan(cols, outs)
    plan.run()
    torch.cuda.synchronize()
    check_all(specs, expected, outs, "run")
    for rep in range(2):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        check_all(specs, expected, outs, f"graph {rep}")
    plan.close()
    del cols, outs
    torch.cuda.empty_cache()
