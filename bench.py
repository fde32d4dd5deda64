#!/usr/bin/env python3
"""Benchmark: decompressed GB/s and % of the HBM roofline on one MI355X.

Workload (one "step"): the FFOR bit-width sweep — for uint32 at every bit
width 0..32 and uint64 at every bit width 0..64, a column of 2^28 values
(per-vector random base plus uniform offsets in [0, 2^w)) is decoded to HBM by
the HIP kernels. All packed inputs (~87.5 GB) are resident in HBM before
timing; the output buffer (2 GiB) is reused. value = decompressed bytes of the
whole sweep / step time.

Extra, untimed-by-the-driver fields report the other shapes: the 2^25-value
uint32 width-13 column (decode and fused sum), the fused filter + sum sweep,
and per-width roofline fractions.

`roofline` is measured on the dominant kernel (the uint64 FFOR decode, the
autotuned variant at each width) with HIP events around each launch on the
launch stream. `cpu_baseline` times the
reference's own CPU decoder (compiled from its sources into oracle/_ref) on a
bounded sample of the same sweep, single-threaded.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "msc-thesis-project_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--values-log2", type=int, default=28, help="values per (dtype, width) column")
    ap.add_argument("--variant", type=int, default=0, help="0 = autotuned default")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the timed sweep")
    ap.add_argument("--extras", default="c1,filter,alp,batch,selector",
                    help="comma list of the extra workloads to run")
    ap.add_argument("--out", default=None, help="also write the JSON line here")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the decoder has no multi-GPU path); "
              f"got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)
    import torch
    from flsgpu import _capi as A
    from flsgpu import device as D
    from flsgpu import encode as E

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    A.check(A.lib().flsgpu_init(0))
    stream = torch.cuda.current_stream()
    nvals = 1 << args.values_log2
    n_vecs = nvals // 1024
    print(f"[bench] building sweep columns: 2^{args.values_log2} values per width",
          file=sys.stderr, flush=True)

    cols = []  # (width, bw, DeviceFFORColumn)
    for width in (32, 64):
        for bw in range(width + 1):
            p, b, o, s = E.synthetic_ffor_column_device(n_vecs, bw, width, seed=width * 100 + bw,
                                                        device=dev)
            cols.append((width, bw, D.DeviceFFORColumn(nvals, p, b, o, s, width)))
    out32 = torch.empty(nvals, dtype=torch.uint32, device=dev)
    out64 = torch.empty(nvals, dtype=torch.uint64, device=dev)
    torch.cuda.synchronize()
    print(f"[bench] resident: {torch.cuda.memory_allocated() / 1e9:.1f} GB", file=sys.stderr,
          flush=True)

    def algo_bytes(width, bw):
        nl = 1024 // width
        read = n_vecs * (bw * nl * width // 8 + 1 + 8 + width // 8)  # packed + bw + offset + base
        return read + nvals * width // 8

    from flsgpu import autotune as AT
    tuned = os.path.exists(AT.DEFAULT_TABLE) and args.variant == 0
    if tuned:
        AT.load_table()
    variants = [AT.variant_for(f"ffor_u{w}", bw) if tuned else args.variant for w, bw, _ in cols]

    def run_step(events=None, plain=False):
        for i, (width, bw, col) in enumerate(cols):
            out = out32 if width == 32 else out64
            if events is not None:
                events[i][0].record(stream)
            D.decompress(col, out, variant=A.VARIANT_LANE if plain else variants[i], stream=stream)
            if events is not None:
                events[i][1].record(stream)

    for _ in range(args.warmup):
        run_step()
    torch.cuda.synchronize()

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in cols]
    kernel_ms = [0.0] * len(cols)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_step(events)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(events):
            kernel_ms[i] += a.elapsed_time(b)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the per-step synchronize to read events is part of the timed region
    step_ms = (t1 - t0) * 1e3 / args.steps

    dec_bytes = sum(nvals * w // 8 for w, _, _ in cols)
    value = dec_bytes / (step_ms / 1e3) / 1e9

    # the same sweep with the plain one-thread-per-lane variant, for comparison
    run_step(plain=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_step(plain=True)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    per_width = []
    for i, (width, bw, _) in enumerate(cols):
        ms = kernel_ms[i] / args.steps
        ab = algo_bytes(width, bw)
        per_width.append({"dtype": f"u{width}", "bw": bw, "us": round(ms * 1e3, 2),
                          "decompressed_GBps": round(nvals * width / 8 / (ms / 1e3) / 1e9, 1),
                          "hbm_frac": round(ab / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)})
    u64 = [i for i, (w, _, _) in enumerate(cols) if w == 64]
    avg_bytes = sum(algo_bytes(64, cols[i][1]) for i in u64) / len(u64)
    avg_ms = sum(kernel_ms[i] for i in u64) / len(u64) / args.steps
    achieved = avg_bytes / (avg_ms / 1e3) / 1e9
    u32 = [i for i, (w, _, _) in enumerate(cols) if w == 32]
    avg_bytes32 = sum(algo_bytes(32, cols[i][1]) for i in u32) / len(u32)
    avg_ms32 = sum(kernel_ms[i] for i in u32) / len(u32) / args.steps

    result = {
        "metric": "decompressed GB/s + % of HBM roofline, plain vs batched/autotuned, 1 GPU",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+u64",
        "data": "synthetic",
        "config": {"workload": f"FFOR bit-width sweep: u32 w=0..32 and u64 w=0..64, "
                               f"2^{args.values_log2} values per width, decode to HBM",
                   "values_per_width": nvals, "variant": args.variant},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(),
                     "kernel": "FFOR u64 decode, autotuned variant per width (average over "
                               "w=0..64): " + variant_mix(A, [variants[i] for i in u64]),
                     "avg_launch_us": round(avg_ms * 1e3, 2),
                     "avg_algorithmic_bytes": int(avg_bytes)},
        "u32_sweep": {"avg_launch_us": round(avg_ms32 * 1e3, 2),
                      "hbm_frac": round(avg_bytes32 / (avg_ms32 / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "autotuned": tuned,
        "plain_lane_variant": {"value": round(dec_bytes / (plain_ms / 1e3) / 1e9, 1),
                               "ms_per_step": round(plain_ms, 3)},
        "hbm_copy_GBps": round(copy_bandwidth(torch, dev, stream), 1),
        "per_width": per_width,
    }

    host = {}  # host copies of the GPU run's buffers for the CPU leg
    if not args.no_cpu:
        host["sweep"] = [(w, bw, host_prefix(c, 1 << 10)) for w, bw, c in cols]
        host["c2"] = {(w, bw): host_prefix(c, 1 << 14) for w, bw, c in cols if (w, bw) in C2_CPU}
    del cols
    torch.cuda.empty_cache()
    if not args.no_extras:
        result["extras"] = extras(args, torch, A, D, E, dev, stream, out32,
                                  None if args.no_cpu else host)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(args, host)
    line = json.dumps(result)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


def variant_mix(A, vs):
    """'name xN, ...' of the decoder variants used, most used first."""
    names = {v: k for k, v in A.VARIANTS.items()}
    counts = {}
    for v in vs:
        counts[v] = counts.get(v, 0) + 1
    return ", ".join(f"{names.get(v, v)} x{c}" for v, c in
                     sorted(counts.items(), key=lambda kv: -kv[1]))


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "ffor_decompress_u64_sweep_2p28.md")


def measured_traffic():
    """HBM bytes per launch of the u64 sweep decoder (average over widths),
    from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in the profile,
    each scaled by the same counter's reading for a 2 GiB copy; None if the
    summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_bandwidth(torch, dev, stream):
    """Achievable HBM bandwidth for reference: a 2 GiB device-to-device copy
    (read + write bytes per second)."""
    a = torch.empty(1 << 28, dtype=torch.int64, device=dev)
    b = torch.empty_like(a)
    a.fill_(1)
    ms = time_launches(torch, stream, lambda: b.copy_(a), 10)
    del a, b
    return 2 * (1 << 31) / ms / 1e6


class CacheFlush:
    """Evicts the 256 MB Infinity Cache and the L2s around a timed launch by
    reading a 1 GiB scratch buffer, so a launch repeated on one column reads
    it from HBM, as a first launch would. A read leaves clean lines behind; a
    write would leave dirty ones whose write-back lands in the timed launch."""

    def __init__(self, torch, dev):
        self.torch = torch
        self.buf = torch.ones(1 << 27, dtype=torch.int64, device=dev)
        self.clean_ms = None

    def __call__(self):
        self.torch.sum(self.buf)


def _flushed_events_ms(torch, stream, body, reps, flush):
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for a, b in evs:
        flush()
        a.record(stream)
        body()
        b.record(stream)
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in evs) / reps


def cold_launch_ms(torch, stream, fn, reps, flush, writes=True):
    """Mean cost of `reps` launches of fn on `stream`, each after a cache
    flush. For work that writes (writes=True) the sustained cost: the launch
    plus a second flush inside its events, less what that flush takes on
    clean caches. A kernel with plain stores ends with up to 256 MB of dirty
    output lines in the Infinity Cache; their write-back is part of its HBM
    traffic but lands in whatever runs next, so the launch alone would not be
    charged for it (back-to-back launches, as in the timed sweep, are charged
    in sequence). Read-only work leaves nothing to write back and is timed
    alone between its events."""
    fn()
    torch.cuda.synchronize()
    if not writes:
        return _flushed_events_ms(torch, stream, fn, reps, flush)
    if flush.clean_ms is None:
        flush.clean_ms = _flushed_events_ms(torch, stream, flush, reps, flush)

    def body():
        fn()
        flush()

    return _flushed_events_ms(torch, stream, body, reps, flush) - flush.clean_ms


def time_launches(torch, stream, fn, reps):
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def extras(args, torch, A, D, E, dev, stream, out32, host=None):
    """Other configurations (not the headline): C1 u32 w=13 column, fused
    filter+sum sweep at 2^28 values."""
    import numpy as np
    res = {}
    which = set(args.extras.split(","))
    r = torch.empty(2, dtype=torch.uint64, device=dev)
    flush = CacheFlush(torch, dev)
    if "c1" in which:
        res.update(c1_extras(args, torch, A, D, E, dev, stream, r, flush, host))
    if "filter" in which:
        res["filter_sum_sweep"] = filter_sum_extras(args, torch, A, D, E, dev, stream, r, flush)
    torch.cuda.empty_cache()
    if "alp" in which:
        res.update(alp_extras(torch, dev, stream, flush, host))
    del flush
    if "batch" in which:
        res["c4_batch"] = batch_extras(torch, dev, stream)
    if "selector" in which:
        res["c5_selector"] = selector_extras(torch, dev)
    return res


def c1_extras(args, torch, A, D, E, dev, stream, r, flush, host=None):
    """2^25 uint32 values 1e6 + U[0, 2^13), FOR base = vector min, width 13:
    decode to HBM and fused sum; every timed launch starts with cold caches."""
    import numpy as np
    res = {}
    rng = np.random.default_rng(2025)
    vals = (1_000_000 + rng.integers(0, 1 << 13, 1 << 25)).astype(np.uint32)
    p, b, o, s = E.ffor_encode_uniform(vals, 13)
    col = D.DeviceFFORColumn.from_host(p, b, o, s, len(vals), dev)
    if host is not None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        host["c1"] = O.FFORColumn(len(vals), p, b.astype(np.uint8), o.astype(np.uint64), s)
    out = torch.empty(len(vals), dtype=torch.uint32, device=dev)
    ms = cold_launch_ms(torch, stream, lambda: D.decompress(col, out, variant=args.variant,
                                                            stream=stream), 20, flush)
    nbytes = len(vals) * 4
    algo = p.nbytes + b.nbytes + o.nbytes + s.nbytes + nbytes
    res["c1_u32_w13_2^25_decode"] = {"us": round(ms * 1e3, 2),
                                      "decompressed_GBps": round(nbytes / ms / 1e6, 1),
                                      "hbm_frac": round(algo / ms / 1e6 / HBM_PEAK_GBS, 4),
                                      "algorithmic_bytes": int(algo)}

    ms = cold_launch_ms(torch, stream, lambda: D.query_filter_sum(col, A.PRED_ALL, result=r,
                                                                  stream=stream), 20, flush,
                        writes=False)
    got_sum = int(r.cpu().numpy()[1])
    res["c1_u32_w13_2^25_sum"] = {"us": round(ms * 1e3, 2),
                                   "decompressed_equiv_GBps": round(nbytes / ms / 1e6, 1),
                                   "hbm_frac": round((algo - nbytes) / ms / 1e6 / HBM_PEAK_GBS, 4),
                                   "algorithmic_bytes": int(algo - nbytes),
                                   "sum_ok": got_sum == int(vals.astype(np.uint64).sum())}
    del col
    return res


def filter_sum_extras(args, torch, A, D, E, dev, stream, r, flush):
    """Fused filter (offset < 2^(w-1)) + sum over sweep widths (2^28 values)."""
    nvals = 1 << args.values_log2
    n_vecs = nvals // 1024
    fs = []
    for width, bws in ((32, (1, 8, 13, 16, 24, 32)), (64, (1, 16, 32, 48, 64))):
        for bw in bws:
            pk, bb, oo, ss = E.synthetic_ffor_column_device(n_vecs, bw, width, seed=7 + bw, device=dev)
            col = D.DeviceFFORColumn(nvals, pk, bb, oo, ss, width)
            ms = cold_launch_ms(torch, stream, lambda: D.query_filter_sum(
                col, A.PRED_OFFSET_LOWER_HALF, result=r, stream=stream), 5, flush, writes=False)
            read = pk.numel() * pk.element_size() + n_vecs * (1 + 8 + width // 8)
            fs.append({"dtype": f"u{width}", "bw": bw, "us": round(ms * 1e3, 1),
                       "decompressed_equiv_GBps": round(nvals * width / 8 / ms / 1e6, 1),
                       "hbm_frac": round(read / ms / 1e6 / HBM_PEAK_GBS, 4)})
            del col, pk, bb, oo, ss
    return fs


def selector_extras(torch, dev):
    """Learned ALP selector: 2^24 training vectors (exhaustive-search labels
    on the host), trained on the GPU with the library's kernels; evaluated on
    a held-out 2^24-value double column (2^14 vectors): compression ratio and
    selection time vs the exhaustive search, pure proposal and verified
    top-4 proposal."""
    from flsgpu import selector as SEL
    params, info = SEL.train_default(train_log2=24, epochs=4, seed=1, device=dev, log=None,
                                     batch=4096)
    res = SEL.evaluate(params, n_vecs=1 << 14, seed=99, device=dev, n_candidates=4)
    res.update({k: (round(v, 3) if isinstance(v, float) else v) for k, v in info.items()})
    return {k: (round(v, 6) if isinstance(v, float) else v) for k, v in res.items()}


def batch_extras(torch, dev, stream):
    """96 columns, log-uniform lengths 2^12..2^24, a third each FFOR u32
    (width 1..24), ALP double (G-ALP layout) and ALP-RD double: one plain
    launch per column vs one persistent batched launch vs its hipGraph replay."""
    import numpy as np
    from flsgpu import batch as B
    from flsgpu import workloads as WL
    t0 = time.perf_counter()
    specs = WL.mixed_batch(seed=4, n_columns=96)
    gen_s = time.perf_counter() - t0
    cols = [WL.to_device(s, dev) for s in specs]
    outs = [B.output_for(c, dev) for c in cols]
    plan = B.BatchPlan(cols, outs)
    dec_bytes = sum(s.values.nbytes for s in specs)
    reps = 10
    flush = CacheFlush(torch, dev)
    t_plain = cold_launch_ms(torch, stream, lambda: B.decode_each(cols, outs, stream=stream), reps,
                             flush)
    t_batch = cold_launch_ms(torch, stream, lambda: plan.run(stream), reps, flush)
    t_graph = cold_launch_ms(torch, stream, lambda: plan.graph_launch(stream), reps, flush)
    # the plain launches per kind, and the bytes a decode must move
    per_kind = {}
    for kind in sorted({s.kind for s in specs}):
        idx = [i for i, s in enumerate(specs) if s.kind == kind]
        kc, ko = [cols[i] for i in idx], [outs[i] for i in idx]
        t = cold_launch_ms(torch, stream, lambda: B.decode_each(kc, ko, stream=stream), 5, flush)
        per_kind[kind] = {"plain_us": round(t * 1e3, 1),
                          "values": int(sum(len(specs[i].values) for i in idx))}
    del flush
    read_bytes = sum(t.numel() * t.element_size() for c in cols for t in vars(c).values()
                     if isinstance(t, torch.Tensor))
    algo = read_bytes + dec_bytes
    ok = True
    for s, o in zip(specs, outs):
        v = s.values.view(np.int32 if s.values.itemsize == 4 else np.int64)
        ok &= bool(torch.equal(o.view(torch.int32 if s.values.itemsize == 4 else torch.int64),
                               torch.from_numpy(v).to(dev)))
    kinds = {}
    for s in specs:
        kinds[s.kind] = kinds.get(s.kind, 0) + 1
    res = {"columns": len(specs), "kinds": kinds, "values": int(sum(len(s.values) for s in specs)),
           "tasks": plan.n_tasks, "bit_exact": ok, "host_generate_encode_s": round(gen_s, 1),
           "plain_per_column_us": round(t_plain * 1e3, 1),
           "batched_us": round(t_batch * 1e3, 1), "graph_us": round(t_graph * 1e3, 1),
           "plain_GBps": round(dec_bytes / t_plain / 1e6, 1),
           "batched_GBps": round(dec_bytes / t_batch / 1e6, 1),
           "graph_GBps": round(dec_bytes / t_graph / 1e6, 1),
           "speedup_batched_vs_plain": round(t_plain / t_batch, 3),
           "speedup_graph_vs_plain": round(t_plain / t_graph, 3),
           "algorithmic_bytes": int(algo),
           "batched_hbm_frac": round(algo / t_batch / 1e6 / HBM_PEAK_GBS, 4),
           "plain_per_kind": per_kind}
    plan.close()
    del cols, outs
    torch.cuda.empty_cache()
    return res


def c3_values(np, width, rng):
    """2^26 doubles k/100, k ~ U[-10^6, 10^6] (2 decimals in [-10^4, 10^4]), or
    2^27 floats k/10, k ~ U[-10^4, 10^4]; in every 1024-vector 10 positions
    (1%) replaced by full-precision uniform values."""
    if width == 64:
        n = 1 << 26
        x = rng.integers(-1_000_000, 1_000_001, n) / 100.0
        repl = rng.uniform(-1e4, 1e4, (n // 1024, 10))
    else:
        n = 1 << 27
        x = rng.integers(-10_000, 10_001, n).astype(np.float32) / np.float32(10.0)
        repl = rng.uniform(-1e3, 1e3, (n // 1024, 10)).astype(np.float32)
    nv = n // 1024  # 10 distinct positions per vector: an odd stride is coprime with 1024
    pos = (rng.integers(0, 1024, (nv, 1)) + (2 * rng.integers(0, 512, (nv, 1)) + 1)
           * np.arange(10)[None, :]) % 1024
    v = x.reshape(-1, 1024)
    np.put_along_axis(v, pos, repl, axis=1)
    return x


def alp_extras(torch, dev, stream, flush, host=None):
    import numpy as np
    from flsgpu import alp as AL
    res = {}
    rng = np.random.default_rng(33)
    for width in (64, 32):
        x = c3_values(np, width, rng)
        t0 = time.perf_counter()
        col = AL.alp_encode(x)
        enc_s = time.perf_counter() - t0
        galp = AL.galp_convert(col)
        if host is not None:
            name = f"c3_alp_f{width}_2^{int(np.log2(len(x)))}"
            host.setdefault("c3", {})[name] = col
            host.setdefault("c3_values", {})[name] = x
        ref = torch.from_numpy(x.view(np.int64 if width == 64 else np.int32)).to(dev)
        nbytes = x.nbytes
        for name, c in (("alp", col), ("galp", galp)):
            d = AL.DeviceAlpColumn.from_host(c, dev)
            out = torch.empty(len(x), dtype=torch.float64 if width == 64 else torch.float32, device=dev)
            ms = cold_launch_ms(torch, stream, lambda: AL.decompress(d, out, stream=stream), 10,
                                flush)
            ok = bool(torch.equal(out.view(ref.dtype), ref))
            read = (c.packed.nbytes + c.bit_widths.nbytes + c.vector_offsets.nbytes + c.bases.nbytes
                    + c.factor_indices.nbytes * 2 + c.exceptions_offsets.nbytes + c.exceptions.nbytes
                    + c.positions.nbytes + (c.counts.nbytes if c.offsets_counts is None else c.offsets_counts.nbytes))
            res[f"c3_{name}_f{width}_2^{int(np.log2(len(x)))}"] = {
                "us": round(ms * 1e3, 1), "decompressed_GBps": round(nbytes / ms / 1e6, 1),
                "hbm_frac": round((read + nbytes) / ms / 1e6 / HBM_PEAK_GBS, 4), "bit_exact": ok,
                "algorithmic_bytes": int(read + nbytes),
                "bits_per_value": round(float(col.bit_widths.mean()), 2),
                "exceptions_per_vector": round(len(col.exceptions) / col.n_vecs, 2),
                "compression_ratio": round(c.compression_ratio(extended=name == "galp"), 3),
                "host_encode_s": round(enc_s, 2)}
            # "any value == v" over the column (the reference's decompression_query),
            # v absent so the whole column is scanned; read-only
            absent = 123456.789 if width == 64 else 12345.678
            found = torch.empty(1, dtype=torch.uint8, device=dev)
            qms = cold_launch_ms(torch, stream, lambda: AL.query_any(d, absent, found, stream=stream),
                                 10, flush, writes=False)
            res[f"c3_{name}_f{width}_query_any"] = {
                "us": round(qms * 1e3, 1), "hbm_frac": round(read / qms / 1e6 / HBM_PEAK_GBS, 4),
                "found": int(found.item()), "expected": int(np.any(x == x.dtype.type(absent)))}
            del d, out
        del ref
        torch.cuda.empty_cache()
    # ALP-RD (the C4 batch's third kind): 2^26 doubles uniform in [0, 1)
    x = rng.uniform(0.0, 1.0, 1 << 26)
    t0 = time.perf_counter()
    rd = AL.alprd_encode(x)
    enc_s = time.perf_counter() - t0
    d = AL.DeviceAlprdColumn.from_host(rd, dev)
    out = torch.empty(len(x), dtype=torch.float64, device=dev)
    ms = cold_launch_ms(torch, stream, lambda: AL.alprd_decompress(d, out, stream=stream), 10, flush)
    ok = bool(torch.equal(out.view(torch.int64), torch.from_numpy(x.view(np.int64)).to(dev)))
    read = sum(a.nbytes for a in (rd.right_packed, rd.left_packed, rd.exceptions_offsets,
                                  rd.exceptions, rd.positions, rd.counts))
    res["alprd_f64_2^26"] = {"us": round(ms * 1e3, 1), "decompressed_GBps": round(x.nbytes / ms / 1e6, 1),
                             "hbm_frac": round((read + x.nbytes) / ms / 1e6 / HBM_PEAK_GBS, 4),
                             "algorithmic_bytes": int(read + x.nbytes),
                             "bit_exact": ok, "right_bit_width": int(rd.right_bw),
                             "left_bit_width": int(rd.left_bw),
                             "exceptions_per_vector": round(len(rd.exceptions) / rd.n_vecs, 2),
                             "compression_ratio": round(x.nbytes / rd.compressed_bytes, 3),
                             "host_encode_s": round(enc_s, 2)}
    del d, out
    torch.cuda.empty_cache()
    return res


def host_prefix(col, n_vecs):
    """The first n_vecs vectors of a device FFOR column, copied to host as an
    oracle column (the CPU leg decodes the GPU run's own buffers)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    n_vecs = min(n_vecs, col.n_vecs)
    nl = 1024 // col.width
    bws = col.bit_widths[:n_vecs].cpu().numpy()
    offs = col.vector_offsets[:n_vecs].cpu().numpy()
    end = int(offs[-1]) + int(bws[-1]) * nl
    packed = col.packed[:max(end, 1)].cpu().numpy()
    bases = col.bases[:n_vecs].cpu().numpy()
    return O.FFORColumn(n_vecs * 1024, packed, bws, offs, bases)


# C2 columns the CPU leg also times on their own: (element bits, bit width)
C2_CPU = {(32, 13), (32, 16), (32, 32), (64, 13), (64, 16), (64, 32), (64, 64)}
CPU_RUNS = 11  # timed runs per figure after one warm-up; the median is reported


def _median_s(fn, runs=CPU_RUNS):
    import statistics
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def _ranges(n, k):
    return [(n * i // k, n * (i + 1) // k) for i in range(k) if n * i // k < n * (i + 1) // k]


def cpu_baseline(args, host):
    """The reference's own CPU decoders, compiled from its sources into
    oracle/_ref, on the GPU run's buffers, one thread and all host threads (at
    most 16, the CPU share of a one-GPU job; the ctypes calls release the GIL,
    so a thread pool over contiguous vector ranges runs them in parallel).
    Each figure is the median of CPU_RUNS runs after a warm-up.

    - headline: the bit-width sweep (u32 w=0..32, u64 w=0..64), the first
      2^20 values of every device column: fls::unffor per vector, the loop of
      src/engine/data.cuh:300-313;
    - C2 widths 13/16/32 (u32) and 13/16/32/64 (u64): the first 2^24 values
      of those device columns;
    - C1: the whole 2^25-value width-13 column;
    - C3: the whole ALP columns (2^26 doubles, 2^27 floats), standard layout:
      falp + patch_exceptions per vector (src/alp/alp-bindings.cu:208-228)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle_py as O
    from concurrent.futures import ThreadPoolExecutor
    R = O.ref()
    if R is None:
        return {"value": None, "unit": "GB/s", "cores": 1, "kind": "reference",
                "sample": "oracle/_ref is not built: no CPU baseline"}
    n_threads = max(1, min(16, len(os.sched_getaffinity(0))))
    pool = ThreadPoolExecutor(n_threads)

    def ffor_fn(col, out):
        isz = col.packed.itemsize
        f = getattr(R, f"ref_ffor_column_decode_u{isz * 8}")
        pk, bw = col.packed.ctypes.data, col.bit_widths.ctypes.data
        of, bs, od = col.vector_offsets.ctypes.data, col.bases.ctypes.data, out.ctypes.data

        def run(v0, v1, keep=(col, out)):  # keep: the arrays the addresses point into
            f(v1 - v0, pk, bw + v0, of + 8 * v0, bs + isz * v0, od + 1024 * isz * v0)
        return run

    def alp_fn(col, out):
        return lambda v0, v1: O.ref_alp_column_decode_range(col, v0, v1, out)

    def timed(items):
        """items: [(decode(v0, v1), n_vecs, decompressed bytes)] decoded in
        one run; returns (GB/s on 1 thread, GB/s on n_threads)."""
        nbytes = sum(b for _, _, b in items)
        t1 = _median_s(lambda: [fn(0, n) for fn, n, _ in items])

        # pieces of about an equal share of all vectors (many small columns
        # go whole to one thread; a large one is split)
        piece = max(1, sum(n for _, n, _ in items) // (4 * n_threads))

        def par():
            futs = [pool.submit(fn, a, b) for fn, n, _ in items
                    for a, b in _ranges(n, max(1, -(-n // piece)))]
            for f in futs:
                f.result()
        tn = _median_s(par)
        return round(nbytes / t1 / 1e9, 3), round(nbytes / tn / 1e9, 3)

    def ffor_items(cols):
        items = []
        for c in cols:
            out = np.empty(c.n_vecs * 1024, dtype=c.packed.dtype)
            items.append((ffor_fn(c, out), c.n_vecs, out.nbytes))
        return items

    t0 = time.perf_counter()
    sweep = host["sweep"]
    one, many = timed(ffor_items([c for _, _, c in sweep]))
    configs = {}
    for (width, bw), c in sorted(host.get("c2", {}).items()):
        a, b = timed(ffor_items([c]))
        configs[f"c2_u{width}_w{bw}_2^{int(np.log2(c.n_values))}"] = {
            "threads_1": a, f"threads_{n_threads}": b}
    if "c1" in host:
        a, b = timed(ffor_items([host["c1"]]))
        configs["c1_u32_w13_2^25"] = {"threads_1": a, f"threads_{n_threads}": b}
    for name, col in host.get("c3", {}).items():
        out = np.empty(col.n_values, dtype=col.exceptions.dtype)
        a, b = timed([(alp_fn(col, out), col.bit_widths.size, out.nbytes)])
        ok = bool(np.array_equal(out.view(np.uint8), host["c3_values"][name].view(np.uint8)))
        configs[name] = {"threads_1": a, f"threads_{n_threads}": b, "bit_exact": ok}
    pool.shutdown()
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v.strip()
    except OSError:
        pass
    model = (f"{info.get('model name', '?')} (family {info.get('cpu family', '?')}, model "
             f"{info.get('model', '?')}, stepping {info.get('stepping', '?')}, "
             f"{info.get('cpu MHz', '?')} MHz)")
    return {"value": one, "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"the first 2^20 values of every sweep column (u32 w=0..32, u64 w=0..64) "
                      f"copied from the device, fls::unffor per vector, median of {CPU_RUNS} "
                      f"runs, single thread",
            "all_cores": {"value": many, "unit": "GB/s", "cores": n_threads,
                          "sample": f"the same columns over {n_threads} threads, in pieces of "
                                    f"about 1/{4 * n_threads} of the vectors"},
            "configs": configs, "cpu_seconds": round(time.perf_counter() - t0, 1),
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


if __name__ == "__main__":
    main()
