"""GPU parity of the batched persistent decoder: every encoding in one
launch, bitwise equal to one plain launch per column and to the original
values; hipGraph replays re-run the self-re-arming queue correctly."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

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


def bits(t):
    a = t.cpu().numpy()
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def test_batch_every_encoding_matches_plain_and_original():
    rng = np.random.default_rng(1)
    specs = every_encoding(rng) + WL.mixed_batch(seed=2, n_columns=9, min_log2=10, max_log2=15)
    cols = [WL.to_device(s) for s in specs]
    outs_b = [B.output_for(c) for c in cols]
    outs_p = [B.output_for(c) for c in cols]
    plan = B.BatchPlan(cols, outs_b)
    assert plan.n_tasks > len(cols) // 2
    plan.run()
    B.decode_each(cols, outs_p)
    torch.cuda.synchronize()
    for s, ob, op in zip(specs, outs_b, outs_p):
        assert np.array_equal(bits(ob), bits(op)), s.kind
        v = s.values
        assert np.array_equal(bits(ob), v.view({4: np.uint32, 8: np.uint64}[v.dtype.itemsize])), s.kind
    # the queue re-armed itself: a second run decodes everything again
    for o in outs_b:
        o.zero_()
    plan.run()
    torch.cuda.synchronize()
    for ob, op in zip(outs_b, outs_p):
        assert np.array_equal(bits(ob), bits(op))
    plan.close()


def test_batch_graph_replays():
    specs = WL.mixed_batch(seed=5, n_columns=12, min_log2=10, max_log2=16)
    cols = [WL.to_device(s) for s in specs]
    outs = [B.output_for(c) for c in cols]
    plan = B.BatchPlan(cols, outs)
    for rep in range(4):
        for o in outs:
            o.fill_(0)                     # eager work between replays
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for s, o in zip(specs, outs):
            v = s.values
            assert np.array_equal(bits(o), v.view({4: np.uint32, 8: np.uint64}[v.dtype.itemsize])), (rep, s.kind)
    plan.close()
