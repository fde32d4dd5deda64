"""Autotuner: measures every decoder variant per (encoding, bit width) on the
current GPU and keeps the fastest, replacing the reference's offline
"fastest unpacker per bit width" SQL query over benchmark CSVs
(sql/select-best-unpacker-ffor.sql:1-7). The table is JSON; load_table()
pushes it into the library (flsgpu_set_tuning) so FLSGPU_VARIANT_AUTO picks it
up, and variant_for() serves hosts that know a column's bit width.

Run:  python -m flsgpu.autotune --out <table.json> [--values-log2 28]

Every launch is timed alone after a cache flush, at the column size given
(default 2^28 values, the size of the bit-width sweep): smaller columns fit in
the 256 MB Infinity Cache and rank the variants differently.
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np
import torch

from . import _capi as A

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
ENC = {"ffor_u32": 1, "ffor_u64": 3, "alp_f32": 4, "alp_f64": 5, "alp_ext_f32": 6,
       "alp_ext_f64": 7, "alprd_f32": 8, "alprd_f64": 9}
_table_cache: dict | None = None


class _Flush:
    """Reads a 1 GiB scratch buffer to evict the 256 MB Infinity Cache and
    the L2s, so every timed launch reads its column from HBM (a column that
    fits in the cache would otherwise favour latency-bound variants). A read
    leaves clean lines; a write would leave dirty lines whose write-back
    would be charged to the timed launch."""

    def __init__(self, device="cuda"):
        self.buf = torch.ones(1 << 27, dtype=torch.int64, device=device)

    def __call__(self):
        torch.sum(self.buf)


_flush = None


_flush_ms = None


def _events(body, reps):
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for a, b in evs:
        _flush()
        a.record()
        body()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in evs) / reps


def _time(fn, reps=10):
    """Mean cost of `reps` launches, each after a cache flush: the launch plus
    a second flush inside the events, less what that flush takes on clean
    caches. A variant with plain stores leaves up to the Infinity Cache's
    256 MB of dirty output lines, whose write-back lands in whatever runs
    next; timing the launch alone would not charge it. (Measured in the
    bit-width sweep: a non-temporal-store kernel after a plain-store one on
    the same output took 277 us instead of 205 us.)"""
    global _flush, _flush_ms
    if _flush is None:
        _flush = _Flush()
        _flush_ms = _events(_flush, reps)
    fn()
    torch.cuda.synchronize()

    def body():
        fn()
        _flush()

    return _events(body, reps) - _flush_ms


TIMES: dict = {}  # (encoding, bit width) -> {variant name: ms}, filled by the tuners


def tune_ffor(values_log2: int = 24, reps: int = 10, device="cuda"):
    from . import device as D
    from . import encode as E
    out = {}
    n = 1 << values_log2
    for width in (32, 64):
        key = f"ffor_u{width}"
        res = torch.empty(n, dtype=torch.uint32 if width == 32 else torch.uint64, device=device)
        table = {}
        for bw in range(width + 1):
            cols = [D.DeviceFFORColumn(n, *E.synthetic_ffor_column_device(n // 1024, bw, width, seed=s),
                                       width) for s in range(2)]
            best = None
            for name, v in A.VARIANTS.items():
                k = [0]

                def run():
                    D.decompress(cols[k[0] % 2], res, variant=v)
                    k[0] += 1
                ms = _time(run, reps)
                TIMES.setdefault(f"{key}/{bw}", {})[name] = round(ms, 4)
                if best is None or ms < best[1]:
                    best = (v, ms)
            table[str(bw)] = best[0]
        out[key] = table
    return out


def tune_alp(values_log2: int = 24, reps: int = 10, device="cuda"):
    from . import alp as AL
    from . import workloads as WL
    out = {}
    rng = np.random.default_rng(7)
    for width in (32, 64):
        x = WL.alp_decimal_values(width, 1 << values_log2, rng)
        col = AL.alp_encode(x)
        for name, c in ((f"alp_f{width}", col), (f"alp_ext_f{width}", AL.galp_convert(col))):
            d = AL.DeviceAlpColumn.from_host(c, device)
            res = torch.empty(len(x), dtype=torch.float32 if width == 32 else torch.float64, device=device)
            best = None
            for vname, v in A.VARIANTS.items():
                ms = _time(lambda: AL.decompress(d, res, variant=v), reps)
                TIMES.setdefault(f"{name}/255", {})[vname] = round(ms, 4)
                if best is None or ms < best[1]:
                    best = (v, ms)
            out[name] = {"255": best[0]}
    for width in (32, 64):  # ALP-RD: uniform values in [0, 1)
        dt = np.float32 if width == 32 else np.float64
        y = rng.uniform(0.0, 1.0, 1 << values_log2).astype(dt)
        d = AL.DeviceAlprdColumn.from_host(AL.alprd_encode(y), device)
        res = torch.empty(len(y), dtype=torch.float32 if width == 32 else torch.float64, device=device)
        best = None
        name = f"alprd_f{width}"
        for vname, v in A.VARIANTS.items():
            ms = _time(lambda: AL.alprd_decompress(d, res, variant=v), reps)
            TIMES.setdefault(f"{name}/255", {})[vname] = round(ms, 4)
            if best is None or ms < best[1]:
                best = (v, ms)
        out[name] = {"255": best[0]}
    return out


def load_table(path: str = DEFAULT_TABLE) -> dict:
    """Reads a tuning table and loads it into the library."""
    global _table_cache
    with open(path) as f:
        table = json.load(f)
    L = A.lib()
    for enc, entries in table.items():
        if enc not in ENC:
            continue
        for bw, v in entries.items():
            A.check(L.flsgpu_set_tuning(ENC[enc], int(bw), int(v)))
    _table_cache = table
    return table


def variant_for(encoding: str, bit_width: int) -> int:
    """Tuned variant for a column whose bit width the host knows."""
    t = _table_cache or {}
    e = t.get(encoding, {})
    if str(bit_width) in e:
        return int(e[str(bit_width)])
    return int(A.lib().flsgpu_get_tuning(ENC[encoding], 255))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--values-log2", type=int, default=28)
    ap.add_argument("--only", choices=["all", "ffor", "alp"], default="all",
                    help="re-tune these encodings; the others keep their entries in the "
                         "shipped table")
    args = ap.parse_args()
    A.check(A.lib().flsgpu_init(0))
    table = {}
    if args.only != "all":
        with open(DEFAULT_TABLE) as f:
            table = json.load(f)
    if args.only in ("all", "ffor"):
        table.update(tune_ffor(args.values_log2))
        # the column default is the variant that wins most widths
        for enc in ("ffor_u32", "ffor_u64"):
            vals = [v for k, v in table[enc].items() if k != "255"]
            table[enc]["255"] = max(set(vals), key=vals.count)
    if args.only in ("all", "alp"):
        # ALP columns: the C3 size (2^26 doubles / floats); encoding 2^28
        # values on the host would dominate the run
        table.update(tune_alp(min(args.values_log2, 26)))
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    with open(os.path.splitext(args.out)[0] + "_times.json", "w") as f:
        json.dump({"values_log2": args.values_log2, "ms": TIMES}, f, indent=1, sort_keys=True)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
